// Host-side helpers of the VTI path (parameter maps, damping profiles,
// boundary shells, wavelet, random boundary).  All are cheap, one-off
// computations done on the CPU exactly as the reference does them, because
// their results must be reproduced bit for bit (the random boundary) or feed
// the bit-exact kernels (damping, taper).
#pragma once

#include <cstddef>
#include <cstdint>

namespace vti {

// init_abc(damping,nx,ny,nz,npad), lib/vti-fwi-3d/gpu/abc3d.cu:23-50
void init_damping3d(float *damping, int nx, int ny, int nz, int npad);

// init_abc(taper,nx,nz,npad), lib/vti-fwi-2d/init.cpp:70-89 (nx, nz unpadded)
void init_taper2d(float *taper, int nx, int nz, int npad);
// the 1D profile t0[npad] of init_abc: taper = t0(x) * t0(z) in the padding
void taper_profile2d(float *t0, int npad);

// VEpsDel2Cij / dVEpsDel2dCij / GradCij2GradVEpsDel, lib/vti-fwi-2d/conversions.cpp:42-81
void vepsdel_to_cij(float *c11, float *c13, float *c33, const float *v, const float *eps, const float *del, float v0,
                    float eps0, float del0, size_t n);
void dvepsdel_to_dcij(float *c11, float *c13, float *c33, float *dc11, float *dc13, float *dc33, const float *v,
                      const float *eps, const float *del, const float *dv, const float *deps, const float *ddel,
                      float v0, float eps0, float del0, size_t n);
void gradcij_to_gradvepsdel(float *gv, float *geps, float *gdel, const float *gc11, const float *gc13,
                            const float *gc33, const float *v, const float *eps, const float *del, float v0, float eps0,
                            float del0, size_t n);

// getBoundary / putBoundary / zeroBoundary, lib/vti-fwi-3d/gpu/boundary.cpp:10-117
// checkEpsDel (lib/vti-fwi-2d/check.cpp:22-40): where del0*del > eps0*eps,
// project (eps0*eps, del0*del) onto the line eps0*eps = del0*del; m marks
// equal points with 1 and projected ones with 2.  Returns the number projected.
long long check_eps_del(float *eps, float *del, float eps0, float del0, size_t n, float *m);

long long boundary_size(int nx, int ny, int nz, int npad);
void get_boundary(float *packed, const float *v, int nx, int ny, int nz, int npad);
void put_boundary(const float *packed, float *v, int nx, int ny, int nz, int npad);
void zero_boundary(float *v, int nx, int ny, int nz, int npad);

// ricker, lib/vti-fwi-2d/ricker.cpp:11-21
void ricker(float *wavelet, float freq, int nt, float dt, float ot, float tdelay, float scalefactor);

// glibc rand()/srand() sequence (TYPE_3 additive feedback generator), so the
// random boundary is reproducible without touching the C library's global
// generator.
class GlibcRand {
public:
    explicit GlibcRand(unsigned seed = 1) { seed_(seed); }
    void seed_(unsigned seed);
    int next();  // same value as rand() after srand(seed)

private:
    int32_t s_[31];
    int f_ = 3, r_ = 0;
};

// RandomMY, lib/random-boundary/randomBound.h:51-276: multiplies the
// boundary shell of `field` (dims n1 fastest, n3 = 1 for 2D) by per-cluster
// random factors.  Returns the number of clusters, or -1 on invalid input.
long long random_boundary(float *field, float bval, float eval, int n1, int n2, int n3, int t1, int t2, int t3,
                          float pctG, int seed);

}  // namespace vti
