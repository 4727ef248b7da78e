// The v/eps/delta parameterisation of the 2D drivers: objFuncGradientVEpsDel
// (lib/vti-fwi-2d/objFuncGradient.cu:28-54) and GNhessianVEpsDel
// (lib/vti-fwi-2d/hessians.cu:72-95).  Host code around the HIP drivers:
// checkEpsDel, the maps v/eps/delta -> c11/c13/c33 (and their Jacobian), the
// chain rule back to v/eps/delta, and the zeroed rows above the water bottom.
#include <algorithm>
#include <cstring>
#include <vector>

#include "vti_amd.h"
#include "vti_host.h"
#include "vti_internal.h"

namespace {

bool valid_grid(int nx, int nz, int npad) { return nx > 0 && nz > 0 && npad >= 0; }

}  // namespace

extern "C" {

int vti_check_eps_del(float *eps, float *del, float eps0, float del0, long long n, float *m)
{
    VTI_CHECK(eps && del && m && n >= 0, "vti_check_eps_del: bad argument");
    return (int)vti::check_eps_del(eps, del, eps0, del0, (size_t)n, m);
}

int vti2d_objfunc_gradient_vepsdel(vti_ctx *ctx, float *objfunc, float *gv, float *geps, float *gdel,
                                   const float *data, float *v, float *eps, float *del, float v0, float eps0,
                                   float del0, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                                   const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                                   float rate, float ot, float wbottom, float *m)
{
    VTI_CHECK(objfunc && gv && geps && gdel && v && eps && del && m && valid_grid(nx, nz, npad),
              "vti2d_objfunc_gradient_vepsdel: bad argument");
    const int nnx = nx + 2 * npad, nnz = nz + 2 * npad;
    const size_t nnxz = (size_t)nnx * nnz;
    vti::check_eps_del(eps, del, eps0, del0, nnxz, m);
    std::vector<float> c(3 * nnxz), gc(3 * nnxz);
    float *c11 = c.data(), *c13 = c11 + nnxz, *c33 = c13 + nnxz;
    float *gc11 = gc.data(), *gc13 = gc11 + nnxz, *gc33 = gc13 + nnxz;
    vti::vepsdel_to_cij(c11, c13, c33, v, eps, del, v0, eps0, del0, nnxz);
    const int rc = vti2d_objfunc_gradient(ctx, objfunc, gc11, gc13, gc33, data, c11, c13, c33, wavelet, sloc, ns, rloc,
                                          nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
    if (rc) return rc;
    vti::gradcij_to_gradvepsdel(gv, geps, gdel, gc11, gc13, gc33, v, eps, del, v0, eps0, del0, nnxz);
    // rows down to the water bottom (objFuncGradient.cu:46-51)
    const int nwbottom = (int)(wbottom / dz + 1);
    const size_t rows = (size_t)std::min(std::max(npad + nwbottom, 0), nnz);
    std::memset(gv, 0, rows * nnx * sizeof(float));
    std::memset(geps, 0, rows * nnx * sizeof(float));
    std::memset(gdel, 0, rows * nnx * sizeof(float));
    return vti::kOk;
}

int vti2d_gn_hessian_vepsdel(vti_ctx *ctx, float *gv, float *geps, float *gdel, const float *data, float *v,
                             float *eps, float *del, const float *dv, const float *deps, const float *ddel, float v0,
                             float eps0, float del0, const float *wavelet, const int *sloc, int ns, const int *rloc,
                             int nr, const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                             float rate, float ot, float wbottom, float *m)
{
    VTI_CHECK(gv && geps && gdel && v && eps && del && dv && deps && ddel && m && valid_grid(nx, nz, npad),
              "vti2d_gn_hessian_vepsdel: bad argument");
    const int nnx = nx + 2 * npad, nnz = nz + 2 * npad;
    const size_t nnxz = (size_t)nnx * nnz;
    vti::check_eps_del(eps, del, eps0, del0, nnxz, m);
    std::vector<float> c(6 * nnxz), gc(3 * nnxz);
    float *c11 = c.data(), *c13 = c11 + nnxz, *c33 = c13 + nnxz;
    float *dc11 = c33 + nnxz, *dc13 = dc11 + nnxz, *dc33 = dc13 + nnxz;
    float *gc11 = gc.data(), *gc13 = gc11 + nnxz, *gc33 = gc13 + nnxz;
    vti::dvepsdel_to_dcij(c11, c13, c33, dc11, dc13, dc33, v, eps, del, dv, deps, ddel, v0, eps0, del0, nnxz);
    const int rc = vti2d_gn_hessian(ctx, gc11, gc13, gc33, data, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, ns,
                                    rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
    if (rc) return rc;
    vti::gradcij_to_gradvepsdel(gv, geps, gdel, gc11, gc13, gc33, v, eps, del, v0, eps0, del0, nnxz);
    // hessians.cu:86: the row count is formed in float, then truncated
    const int n = (int)((float)npad + wbottom / dz + 1.f);
    const size_t rows = (size_t)std::min(std::max(n, 0), nnz);
    std::memset(gv, 0, rows * nnx * sizeof(float));
    std::memset(geps, 0, rows * nnx * sizeof(float));
    std::memset(gdel, 0, rows * nnx * sizeof(float));
    return vti::kOk;
}

}  // extern "C"
