// Checker-only shim (test infrastructure): exposes, with C linkage, the
// reference's own host functions compiled from their sources under the
// reference tree, so tests/golden/make_golden.py can record their outputs.
// Nothing in the product links this.
#include <cstddef>
#include <vector>

#include "boundary.h"
#include "check.h"
#include "conversions.h"
#include "init.h"

void ricker(float *wavelet, float freq, int nt, float dt, float ot, float tdelay, float scalefactor);

extern "C" {
void ref_VEpsDel2Cij(float *c11, float *c13, float *c33, const float *v, const float *eps, const float *del,
                     float v0, float eps0, float del0, size_t n)
{
    VEpsDel2Cij(c11, c13, c33, v, eps, del, v0, eps0, del0, n);
}
void ref_dVEpsDel2dCij(float *c11, float *c13, float *c33, float *dc11, float *dc13, float *dc33, const float *v,
                       const float *eps, const float *del, const float *dv, const float *deps, const float *ddel,
                       float v0, float eps0, float del0, size_t n)
{
    dVEpsDel2dCij(c11, c13, c33, dc11, dc13, dc33, v, eps, del, dv, deps, ddel, v0, eps0, del0, n);
}
void ref_GradCij2GradVEpsDel(float *gv, float *geps, float *gdel, const float *gc11, const float *gc13,
                             const float *gc33, const float *v, const float *eps, const float *del, float v0,
                             float eps0, float del0, size_t n)
{
    GradCij2GradVEpsDel(gv, geps, gdel, gc11, gc13, gc33, v, eps, del, v0, eps0, del0, n);
}
void ref_checkEpsDel(float *eps, float *del, float eps0, float del0, size_t n, float *m)
{
    checkEpsDel(eps, del, eps0, del0, n, m);
}
void ref_init_abc2d(float *taper, int nx, int nz, int npad) { init_abc(taper, nx, nz, npad); }
void ref_getBoundary(float *b, const float *v, int nx, int ny, int nz, int npad) { getBoundary(b, v, nx, ny, nz, npad); }
void ref_putBoundary(const float *b, float *v, int nx, int ny, int nz, int npad) { putBoundary(b, v, nx, ny, nz, npad); }
void ref_zeroBoundary(float *v, int nx, int ny, int nz, int npad) { zeroBoundary(v, nx, ny, nz, npad); }
void ref_ricker(float *w, float freq, int nt, float dt, float ot, float tdelay, float s)
{
    ricker(w, freq, nt, dt, ot, tdelay, s);
}
}
