/root/repo/oracle/_ref/randomBound.x modelIn=/root/repo/tests/golden/sep/velin.H modelOut=/root/repo/tests/golden/sep/velout.H seed=9 pad=8 pctG=0.9 datapath=/root/repo/tests/golden/sep/ 

in=/root/repo/tests/golden/sep/velout.H@
data_format=native_float

n1=48

o1=0

d1=10

n2=40

o2=5

d2=12.5

n3=1

o3=0

d3=1
