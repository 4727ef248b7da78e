n1=48 o1=0 d1=10
n2=40 o2=5 d2=12.5
n3=1 o3=0 d3=1
in=/root/repo/tests/golden/sep/velin.bin
data_format=native_float
