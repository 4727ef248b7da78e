// extern "C" wrappers of the host-side helpers (declared in include/vti_amd.h).
#include "vti_host.h"
#include "vti_internal.h"

extern "C" {

int vti_init_damping3d(float *damping, int nx, int ny, int nz, int npad)
{
    VTI_CHECK(damping && nx > 0 && ny > 0 && nz > 0 && npad > 0, "vti_init_damping3d: bad argument");
    vti::init_damping3d(damping, nx, ny, nz, npad);
    return vti::kOk;
}

int vti_init_taper2d(float *taper, int nx, int nz, int npad)
{
    VTI_CHECK(taper && nx > 0 && nz > 0 && npad > 0, "vti_init_taper2d: bad argument");
    vti::init_taper2d(taper, nx, nz, npad);
    return vti::kOk;
}

int vti_vepsdel_to_cij(float *c11, float *c13, float *c33, const float *v, const float *eps, const float *del,
                       float v0, float eps0, float del0, long long n)
{
    VTI_CHECK(c11 && c13 && c33 && v && eps && del && n >= 0, "vti_vepsdel_to_cij: bad argument");
    vti::vepsdel_to_cij(c11, c13, c33, v, eps, del, v0, eps0, del0, (size_t)n);
    return vti::kOk;
}

int vti_dvepsdel_to_dcij(float *c11, float *c13, float *c33, float *dc11, float *dc13, float *dc33, const float *v,
                         const float *eps, const float *del, const float *dv, const float *deps, const float *ddel,
                         float v0, float eps0, float del0, long long n)
{
    VTI_CHECK(c11 && c13 && c33 && dc11 && dc13 && dc33 && v && eps && del && dv && deps && ddel && n >= 0,
              "vti_dvepsdel_to_dcij: bad argument");
    vti::dvepsdel_to_dcij(c11, c13, c33, dc11, dc13, dc33, v, eps, del, dv, deps, ddel, v0, eps0, del0, (size_t)n);
    return vti::kOk;
}

int vti_gradcij_to_gradvepsdel(float *gv, float *geps, float *gdel, const float *gc11, const float *gc13,
                               const float *gc33, const float *v, const float *eps, const float *del, float v0,
                               float eps0, float del0, long long n)
{
    VTI_CHECK(gv && geps && gdel && gc11 && gc13 && gc33 && v && eps && del && n >= 0,
              "vti_gradcij_to_gradvepsdel: bad argument");
    vti::gradcij_to_gradvepsdel(gv, geps, gdel, gc11, gc13, gc33, v, eps, del, v0, eps0, del0, (size_t)n);
    return vti::kOk;
}

long long vti_boundary_size(int nx, int ny, int nz, int npad) { return vti::boundary_size(nx, ny, nz, npad); }

int vti_get_boundary(float *packed, const float *v, int nx, int ny, int nz, int npad)
{
    VTI_CHECK(packed && v && 2 * npad <= nx && 2 * npad <= ny && 2 * npad <= nz, "vti_get_boundary: bad argument");
    vti::get_boundary(packed, v, nx, ny, nz, npad);
    return vti::kOk;
}

int vti_put_boundary(const float *packed, float *v, int nx, int ny, int nz, int npad)
{
    VTI_CHECK(packed && v && 2 * npad <= nx && 2 * npad <= ny && 2 * npad <= nz, "vti_put_boundary: bad argument");
    vti::put_boundary(packed, v, nx, ny, nz, npad);
    return vti::kOk;
}

int vti_zero_boundary(float *v, int nx, int ny, int nz, int npad)
{
    VTI_CHECK(v && 2 * npad <= nx && 2 * npad <= ny && 2 * npad <= nz, "vti_zero_boundary: bad argument");
    vti::zero_boundary(v, nx, ny, nz, npad);
    return vti::kOk;
}

int vti_ricker(float *wavelet, float freq, int nt, float dt, float ot, float tdelay, float scalefactor)
{
    VTI_CHECK(wavelet && nt >= 0, "vti_ricker: bad argument");
    vti::ricker(wavelet, freq, nt, dt, ot, tdelay, scalefactor);
    return vti::kOk;
}

int vti_glibc_rand(int seed, int n, int *out)
{
    VTI_CHECK(out && n >= 0, "vti_glibc_rand: bad argument");
    vti::GlibcRand g((unsigned)seed);
    for (int i = 0; i < n; ++i) out[i] = g.next();
    return vti::kOk;
}

long long vti_random_boundary(float *field, float bval, float eval, int n1, int n2, int n3, int t1, int t2, int t3,
                              float pctG, int seed)
{
    if (!field || n1 <= 2 * t1 || n2 <= 2 * t2 || (n3 > 1 && n3 <= 2 * t3) || n3 < 1) {
        vti::set_error("vti_random_boundary: bad dimensions");
        return vti::kErrArg;
    }
    return vti::random_boundary(field, bval, eval, n1, n2, n3, t1, t2, t3, pctG, seed);
}

}  // extern "C"
