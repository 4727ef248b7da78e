// vti_compat.hpp -- the reference's C++ driver signatures on top of the C ABI
// of libvtiamd.so (vti_amd.h).  Header-only; link with -lvtiamd.
//
// Lets code written against lib/vti-fwi-2d/wave.h and
// lib/vti-fwi-3d/gpu/wave3d.h (the L-BFGS drivers, HessianOp, the *-param
// mains) call this implementation unchanged, in place of the CUDA objects
// that define these functions: exactly one translation unit defines
// VTI_COMPAT_IMPLEMENTATION before including this header and so holds the
// (non-inline, external) definitions; every other one keeps including the
// reference's wave.h, or this header for the declarations.
//
// Behaviour kept from the reference:
//  - outputs are overwritten (zeroed by the callee), inputs are host memory;
//  - objFuncGradientCij3d_f leaves cij's boundary shell holding
//    randboundaryCij, as the reference's putBoundary does
//    (objFuncGradient3d_f.cu:786);
//  - on failure the reference prints to stderr and carries on; so does this
//    shim (the message comes from vti_last_error()).
//
// What the reference reads from its global argv inside the drivers -- the
// GPU list (gpu=), the observed data (data=), the tpow= gain and the
// objtype=/halfwidth= choice of the 3D objective -- is supplied here in one
// of two ways:
//  - define VTI_COMPAT_USE_MYIO and include the reference's lib/io/myio.h
//    first: the shim then calls get_array("gpu"), read("data") and
//    get_param("tpow") exactly where the reference does;
//  - otherwise call vti_compat::set_observed_data() / set_tpow() /
//    set_objective() / set_ngpu() before the 3D drivers.
// All work runs on one device (vti_compat::set_device, default 0); the GPU
// list only fixes the reference's pipeline length nround*NGPU*64
// (objFuncGradient3d_f.cu:539).
#ifndef VTI_COMPAT_HPP
#define VTI_COMPAT_HPP

#include <cmath>
#include <cstdio>
#include <map>
#include <vector>

#include "vti_amd.h"

namespace vti_compat {

struct State {
    vti_ctx *ctx = nullptr;
    int device = 0;
    int ngpu = 1;
    float tpow = 0.f;
    int objtype = 0, halfwidth = 20;
    const float *observed = nullptr;
};

inline State &state()
{
    static State s;
    return s;
}

inline void set_device(int device) { state().device = device; }
inline void set_ngpu(int ngpu) { state().ngpu = ngpu > 0 ? ngpu : 1; }
inline void set_tpow(float tpow) { state().tpow = tpow; }
// objtype= / halfwidth= of the 3D objective (0: least squares, 1: AGC)
inline void set_objective(int objtype, int halfwidth)
{
    state().objtype = objtype;
    state().halfwidth = halfwidth;
}
// trace-major observed data of every receiver (the reference's "data" file)
inline void set_observed_data(const float *data) { state().observed = data; }

inline vti_ctx *context()
{
    State &s = state();
    if (!s.ctx && vti_ctx_create(&s.ctx, s.device) != 0) {
        std::fprintf(stderr, "vti_compat: %s\n", vti_last_error());
        s.ctx = nullptr;
    }
    return s.ctx;
}

inline void report(int rc, const char *what)
{
    if (rc != 0) std::fprintf(stderr, "%s: %s\n", what, vti_last_error());
}

inline int ngpu()
{
#ifdef VTI_COMPAT_USE_MYIO
    std::vector<int> gpus;
    if (get_array("gpu", gpus) && !gpus.empty()) return (int)gpus.size();
#endif
    return state().ngpu;
}

inline float tpow()
{
#ifdef VTI_COMPAT_USE_MYIO
    float t = 0.f;
    get_param("tpow", t);
    return t;
#else
    return state().tpow;
#endif
}

inline int objtype()
{
#ifdef VTI_COMPAT_USE_MYIO
    int t = 0;
    get_param("objtype", t);
    return t;
#else
    return state().objtype;
#endif
}

inline int halfwidth()
{
#ifdef VTI_COMPAT_USE_MYIO
    int h = 20;
    get_param("halfwidth", h);
    return h;
#else
    return state().halfwidth;
#endif
}

// The receivers the listed shots use: one past the last receiver index.
inline long long receivers_used(const float *souloc, const std::vector<int> &shotid)
{
    long long n = 0;
    for (int is : shotid) {
        const long long end = (long long)souloc[5 * is + 4] + (long long)souloc[5 * is + 3];
        if (end > n) n = end;
    }
    return n;
}

// The reference's split of a run over GPU groups (objFuncGradient3d_f.cu:311-335,
// modelData3d_f.cu:396-420): group i gets GPUs [i*g, (i+1)*g), g =
// |GPUs|/ngpugroup, and the shots in order, the first |shotid| % ngpugroup
// groups one more than the others.  Here the groups run one after another on
// the one device; g only sets each group's pipeline length.
struct Group {
    std::vector<int> shots, gpus;
};
inline std::vector<Group> groups(const std::vector<int> &shotid, const std::vector<int> &GPUs, int ngpugroup)
{
    std::vector<Group> out;
    if (ngpugroup < 1) ngpugroup = 1;
    const size_t g = GPUs.size() / (size_t)ngpugroup;
    const size_t n1 = shotid.size() / (size_t)ngpugroup, r = shotid.size() % (size_t)ngpugroup;
    size_t s0 = 0;
    for (size_t i = 0; i < (size_t)ngpugroup; ++i) {
        Group gr;
        const size_t ns = n1 + (i < r ? 1 : 0);
        gr.shots.assign(shotid.begin() + (long)s0, shotid.begin() + (long)(s0 + ns));
        gr.gpus.assign(GPUs.begin() + (long)(i * g), GPUs.begin() + (long)((i + 1) * g));
        s0 += ns;
        out.push_back(gr);
    }
    return out;
}

// zeroBoundary on each of n fields, then the water layer: rows
// [npad, npad + (int)((wbottom - oz)/dz + 1) - npad) of every field
// (objFuncGradient3d_f.cu:276-282, rtm3d_f.cu:153-155)
inline void mask_shell_and_water(float **f, int n, int nx, int ny, int nz, int npad, float oz, float dz, float wbottom)
{
    const long long nxy = (long long)nx * ny;
    const int nw = (int)((wbottom - oz) / dz + 1) - npad;
    for (int k = 0; k < n; ++k) {
        report(vti_zero_boundary(f[k], nx, ny, nz, npad), "zeroBoundary");
        if (nw > 0)
            for (long long i = 0; i < (long long)nw * nxy && npad * nxy + i < nxy * nz; ++i) f[k][npad * nxy + i] = 0.f;
    }
}

inline int nnt3d(int nt, float dt, float samplingRate)
{
    return (nt - 1) / (int)std::round(samplingRate / dt) + 1;
}

// Observed data for the 3D objective: the reference's read("data") (or the
// registered buffer).  buf keeps what was read alive for the call.
inline const float *observed(std::vector<float> &buf, const float *souloc, const std::vector<int> &shotid, int nt,
                             float dt, float samplingRate)
{
#ifdef VTI_COMPAT_USE_MYIO
    buf.assign((size_t)nnt3d(nt, dt, samplingRate) * (size_t)receivers_used(souloc, shotid), 0.f);
    read("data", buf.data(), buf.size());
    return buf.data();
#else
    (void)buf;
    (void)souloc;
    (void)shotid;
    (void)nt;
    (void)dt;
    (void)samplingRate;
    return state().observed;
#endif
}

}  // namespace vti_compat

// ---- lib/vti-fwi-2d/wave.h ------------------------------------------------
#ifndef WAVE_H
void synthetic_f(float *d, const float *c11, const float *c13, const float *c33, const float *wavelet, const int *sloc,
                 int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx,
                 float dz, float dt, float rate, float ot = 0.);
float objFuncGradient_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11, const float *c13,
                        const float *c33, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                        const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate,
                        float ot = 0.);
void GNhessian_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11, const float *c13,
                 const float *c33, const float *dc11, const float *dc13, const float *dc33, const float *wavelet,
                 const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                 int npad, float dx, float dz, float dt, float rate, float ot);
void hessian_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11, const float *c13,
               const float *c33, const float *dc11, const float *dc13, const float *dc33, const float *wavelet,
               const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
               float dx, float dz, float dt, float rate, float ot);
void rtm_f(float *image, const float *data, const float *c11, const float *c13, const float *c33, const float *wavelet,
           const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
           float dx, float dz, float dt, float rate, float ot = 0.);
void odcig_f(float *image, const float *data, const float *c11, const float *c13, const float *c33,
             const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz,
             int nhx, int nt, int npad, float dx, float dz, float dt, float rate, float ot = 0.);
void modeling_f(float *d, const float *c11, const float *c13, const float *c33, const float *wavelet, int slocxz,
                const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate,
                float ot = 0.);
float objFuncGradientVEpsDel(float *gv, float *geps, float *gdel, const float *data, float *v, float *eps, float *del,
                             float v0, float eps0, float del0, const float *wavelet, const int *sloc, int ns,
                             const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx,
                             float dz, float dt, float rate, float ot, float wbottom, float *m);
void GNhessianVEpsDel(float *gv, float *geps, float *gdel, const float *d0, float *v, float *eps, float *del,
                      const float *dv, const float *deps, const float *ddel, float v0, float eps0, float del0,
                      const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                      int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot, float wbottom,
                      float *m);
#endif

// 3D declarations (wave3d.h itself needs CUDA's stream type, so it is not
// included here)
void modelDataCij3d_f(float *ddata, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                      const float *wavelet, float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt,
                      int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz, float dt,
                      float samplingRate, std::vector<int> &GPUs);
void objFuncCij3d_f(float *f, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                    float *cij, float *padboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                    float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                    std::vector<int> &GPUs);
void objFuncGradientCij3d_f(float *fgcij, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                            float *wavelet, float *cij, float *padboundaryCij, float *randboundaryCij, int nx, int ny,
                            int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy,
                            float dz, float dt, float samplingRate, std::vector<int> &GPUs);
void rtmCij3d_f(float *image, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate);
void rtmVEpsDel(float *image, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet, float *v,
                float *eps, float *del, float *randboundaryV, int nx, int ny, int nz, int nt, int npad, float ox,
                float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                float wbottom);
void objFuncGradientCij3d(float *fgcij, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                          float *wavelet, float *cij, float *padboundaryCij, float *randboundaryCij, int nx, int ny,
                          int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy,
                          float dz, float dt, float samplingRate, std::vector<int> &GPUs, int ngpugroup);
void objFuncCij3d(float *f, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                  float *cij, float *padboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                  float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                  std::vector<int> &GPUs, int ngpugroup);
double objFuncGradientVEpsDel(float *gv, float *geps, float *gdel, float *souloc, int ns, std::vector<int> &shotid,
                              float *recloc, float *wavelet, float *v, float *eps, float *del, float *padboundaryV,
                              float *randboundaryV, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                              float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                              float v0, float eps0, float wbottom, std::vector<int> &GPUs, int ngpugroup);
void modelData3d_f(float *ddata, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                   const float *wavelet, float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad,
                   float ox, float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                   std::vector<int> &GPUs, int ngpugroup);

// lib/lbfgs/lbfgs.h:32
#ifndef LBFGS_H
void lbfgs(size_t n, int &m, float *x, float &f, float *g, bool diagco, float *diag, float *w, int &iflag, int *isave,
           float *dsave);
#endif

#ifdef VTI_COMPAT_IMPLEMENTATION

// synthetic_f (wave.h:4, synthetic_f.cu:13)
void synthetic_f(float *d, const float *c11, const float *c13, const float *c33, const float *wavelet,
                        const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                        int npad, float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_synthetic(vti_compat::context(), d, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper, nx,
                                       nz, nt, npad, dx, dz, dt, rate, ot),
                       "synthetic_f");
}

// objFuncGradient_f (wave.h:12, objFuncGradient_f.cu:15)
float objFuncGradient_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                               const float *c13, const float *c33, const float *wavelet, const int *sloc, int ns,
                               const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx,
                               float dz, float dt, float rate, float ot)
{
    float f = 0.f;
    vti_compat::report(vti2d_objfunc_gradient(vti_compat::context(), &f, gc11, gc13, gc33, data, c11, c13, c33,
                                              wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate,
                                              ot),
                       "objFuncGradient_f");
    return f;
}

// GNhessian_f (wave.h:34, hessian_f.cu:575)
void GNhessian_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11, const float *c13,
                        const float *c33, const float *dc11, const float *dc13, const float *dc33,
                        const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper,
                        int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_gn_hessian(vti_compat::context(), gc11, gc13, gc33, data, c11, c13, c33, dc11, dc13, dc33,
                                        wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot),
                       "GNhessian_f");
}

// hessian_f (wave.h:32, hessian_f.cu:15)
void hessian_f(float *gc11, float *gc13, float *gc33, const float *data, const float *c11, const float *c13,
               const float *c33, const float *dc11, const float *dc13, const float *dc33, const float *wavelet,
               const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
               float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_hessian(vti_compat::context(), gc11, gc13, gc33, data, c11, c13, c33, dc11, dc13, dc33,
                                     wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot),
                       "hessian_f");
}

// rtm_f (wave.h:6, rtm_f.cu:13)
void rtm_f(float *image, const float *data, const float *c11, const float *c13, const float *c33, const float *wavelet,
           const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
           float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_rtm(vti_compat::context(), image, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper,
                                 nx, nz, nt, npad, dx, dz, dt, rate, ot),
                       "rtm_f");
}

// modeling_f (wave.h:14, modeling_f.cu:10)
void modeling_f(float *d, const float *c11, const float *c13, const float *c33, const float *wavelet, int slocxz,
                const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_modeling(vti_compat::context(), d, c11, c13, c33, wavelet, slocxz, taper, nx, nz, nt, npad,
                                      dx, dz, dt, rate, ot),
                       "modeling_f");
}

// odcig_f (wave.h:16, odcig_f.cu:12)
void odcig_f(float *image, const float *data, const float *c11, const float *c13, const float *c33,
             const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz,
             int nhx, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    vti_compat::report(vti2d_odcig(vti_compat::context(), image, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr,
                                   taper, nx, nz, nhx, nt, npad, dx, dz, dt, rate, ot),
                       "odcig_f");
}

// objFuncGradientVEpsDel (wave.h:10, objFuncGradient.cu:28)
float objFuncGradientVEpsDel(float *gv, float *geps, float *gdel, const float *data, float *v, float *eps, float *del,
                             float v0, float eps0, float del0, const float *wavelet, const int *sloc, int ns,
                             const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx,
                             float dz, float dt, float rate, float ot, float wbottom, float *m)
{
    float f = 0.f;
    vti_compat::report(vti2d_objfunc_gradient_vepsdel(vti_compat::context(), &f, gv, geps, gdel, data, v, eps, del, v0,
                                                      eps0, del0, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad,
                                                      dx, dz, dt, rate, ot, wbottom, m),
                       "objFuncGradientVEpsDel");
    return f;
}

// GNhessianVEpsDel (wave.h:42, hessians.cu:72)
void GNhessianVEpsDel(float *gv, float *geps, float *gdel, const float *d0, float *v, float *eps, float *del,
                      const float *dv, const float *deps, const float *ddel, float v0, float eps0, float del0,
                      const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                      int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot, float wbottom,
                      float *m)
{
    vti_compat::report(vti2d_gn_hessian_vepsdel(vti_compat::context(), gv, geps, gdel, d0, v, eps, del, dv, deps, ddel,
                                                v0, eps0, del0, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad,
                                                dx, dz, dt, rate, ot, wbottom, m),
                       "GNhessianVEpsDel");
}

// lbfgs (lbfgs.h:32, lbfgs_c.cpp:11) on the FP64 method of vti_lbfgs_*.  The
// caller's w/isave/dsave keep their sizes but not the state: that lives in an
// FP64 object keyed by w, made on iflag == 0 and dropped when the run ends.
// diag is not read (H0 is (y's)/(y'y)); diagco = true is refused.
void lbfgs(size_t n, int &m, float *x, float &f, float *g, bool diagco, float *diag, float *w, int &iflag, int *isave,
           float *dsave)
{
    (void)diag;
    (void)isave;
    (void)dsave;
    static std::map<float *, vti_lbfgs *> states;
    if (diagco) {
        std::fprintf(stderr, "lbfgs: a user-supplied diagonal (diagco) is not supported\n");
        iflag = -3;
        return;
    }
    vti_lbfgs *&st = states[w];
    if (iflag == 0) {
        if (st) vti_lbfgs_destroy(st);
        st = nullptr;
        if (vti_lbfgs_create(&st, (long long)n, m, 0) != 0) {
            vti_compat::report(-1, "lbfgs");
            states.erase(w);
            iflag = -3;
            return;
        }
    } else if (!st) {
        std::fprintf(stderr, "lbfgs: re-entry with iflag=%d but no run started on this workspace\n", iflag);
        states.erase(w);
        iflag = -3;
        return;
    }
    int it = 0, nf = 0;
    if (vti_lbfgs_step(st, x, (double)f, g, &iflag, &it, &nf) != 0) {
        vti_compat::report(-1, "lbfgs");
        iflag = -1;
    } else if (iflag < 0) {
        std::fprintf(stderr, "lbfgs: %s\n", vti_last_error());
    }
    if (iflag != 1) {
        vti_lbfgs_destroy(st);
        states.erase(w);
    }
}

// ---- lib/vti-fwi-3d/gpu/wave3d.h ------------------------------------------

// modelDataCij3d_f (wave3d.h:110, modelData3d_f.cu:427)
void modelDataCij3d_f(float *ddata, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                             const float *wavelet, float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt,
                             int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz, float dt,
                             float samplingRate, std::vector<int> &GPUs)
{
    vti_compat::report(vti3d_model_data(vti_compat::context(), ddata, souloc, ns, shotid.data(), (int)shotid.size(),
                                        recloc, wavelet, c11, c13, c33, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy,
                                        dz, dt, samplingRate, GPUs.empty() ? 1 : (int)GPUs.size()),
                       "modelDataCij3d_f");
}

// objFuncCij3d_f (wave3d.h:100, objFuncGradient3d_f.cu:1272)
void objFuncCij3d_f(float *f, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                           float *cij, float *padboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox,
                           float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                           std::vector<int> &GPUs)
{
    std::vector<float> buf;
    const float *obs = vti_compat::observed(buf, souloc, shotid, nt, dt, samplingRate);
    vti_compat::report(vti3d_objfunc(vti_compat::context(), f, souloc, ns, shotid.data(), (int)shotid.size(), recloc,
                                     wavelet, cij, padboundaryCij, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz,
                                     dt, samplingRate, GPUs.empty() ? 1 : (int)GPUs.size(), obs, vti_compat::tpow(),
                                     vti_compat::objtype(), vti_compat::halfwidth()),
                       "objFuncCij3d_f");
}

// objFuncGradientCij3d_f (wave3d.h:96, objFuncGradient3d_f.cu:361)
void objFuncGradientCij3d_f(float *fgcij, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                                   float *wavelet, float *cij, float *padboundaryCij, float *randboundaryCij, int nx,
                                   int ny, int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx,
                                   float dy, float dz, float dt, float samplingRate, std::vector<int> &GPUs)
{
    std::vector<float> buf;
    const float *obs = vti_compat::observed(buf, souloc, shotid, nt, dt, samplingRate);
    vti_compat::report(vti3d_objfunc_gradient(vti_compat::context(), fgcij, souloc, ns, shotid.data(),
                                              (int)shotid.size(), recloc, wavelet, cij, padboundaryCij,
                                              randboundaryCij, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt,
                                              samplingRate, GPUs.empty() ? 1 : (int)GPUs.size(), obs,
                                              vti_compat::tpow(), vti_compat::objtype(), vti_compat::halfwidth()),
                       "objFuncGradientCij3d_f");
}

// rtmCij3d_f (wave3d.h:88, rtm3d_f.cu:161); the reference reads gpu= itself
void rtmCij3d_f(float *image, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                       float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad, float ox,
                       float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate)
{
    std::vector<float> buf;
    const float *obs = vti_compat::observed(buf, souloc, shotid, nt, dt, samplingRate);
    vti_compat::report(vti3d_rtm(vti_compat::context(), image, souloc, ns, shotid.data(), (int)shotid.size(), recloc,
                                 wavelet, c11, c13, c33, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt,
                                 samplingRate, vti_compat::ngpu(), obs),
                       "rtmCij3d_f");
}

// rtmVEpsDel (wave3d.h:90, rtm3d_f.cu:139): v/eps/del to Cij with unit
// scales, rtmCij3d_f, then the boundary shell and the water layer of the
// image zeroed.  randboundaryV is not read (the reference's putBoundary of it
// is commented out, rtm3d_f.cu:148).
void rtmVEpsDel(float *image, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet, float *v,
                float *eps, float *del, float *randboundaryV, int nx, int ny, int nz, int nt, int npad, float ox,
                float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                float wbottom)
{
    (void)randboundaryV;
    const long long n = (long long)nx * ny * nz;
    std::vector<float> c(3 * (size_t)n);
    vti_compat::report(vti_vepsdel_to_cij(c.data(), c.data() + n, c.data() + 2 * n, v, eps, del, 1.f, 1.f, 1.f, n),
                       "rtmVEpsDel");
    rtmCij3d_f(image, souloc, ns, shotid, recloc, wavelet, c.data(), c.data() + n, c.data() + 2 * n, nx, ny, nz, nt,
               npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate);
    float *f[1] = {image};
    vti_compat::mask_shell_and_water(f, 1, nx, ny, nz, npad, oz, dz, wbottom);
}

// objFuncGradientCij3d (wave3d.h:94, objFuncGradient3d_f.cu:307): with one
// group, objFuncGradientCij3d_f itself (cij's shell left holding the random
// boundary); with more, each group on its own copy of cij (the caller's cij
// untouched) and the groups' {f, g} summed in FP32, last group first.
void objFuncGradientCij3d(float *fgcij, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                          float *wavelet, float *cij, float *padboundaryCij, float *randboundaryCij, int nx, int ny,
                          int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy,
                          float dz, float dt, float samplingRate, std::vector<int> &GPUs, int ngpugroup)
{
    if (ngpugroup <= 1) {
        objFuncGradientCij3d_f(fgcij, souloc, ns, shotid, recloc, wavelet, cij, padboundaryCij, randboundaryCij, nx,
                               ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, GPUs);
        return;
    }
    const size_t nn = 3 * (size_t)nx * ny * nz;
    std::vector<vti_compat::Group> gs = vti_compat::groups(shotid, GPUs, ngpugroup);
    std::vector<std::vector<float>> fg(gs.size());
    for (size_t i = 0; i < gs.size(); ++i) {
        std::vector<float> tc(cij, cij + nn);
        fg[i].assign(nn + 1, 0.f);
        objFuncGradientCij3d_f(fg[i].data(), souloc, ns, gs[i].shots, recloc, wavelet, tc.data(), padboundaryCij,
                               randboundaryCij, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate,
                               gs[i].gpus);
    }
    for (size_t j = 0; j < nn + 1; ++j) fgcij[j] = 0.f;
    for (size_t i = gs.size(); i-- > 0;)
        for (size_t j = 0; j < nn + 1; ++j) fgcij[j] += fg[i][j];
}

// objFuncCij3d (wave3d.h:98, objFuncGradient3d_f.cu:1233): *f gains each
// group's objective, first group first (*f is not cleared, as in the reference)
void objFuncCij3d(float *f, float *souloc, int ns, std::vector<int> &shotid, float *recloc, float *wavelet,
                  float *cij, float *padboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                  float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                  std::vector<int> &GPUs, int ngpugroup)
{
    std::vector<vti_compat::Group> gs = vti_compat::groups(shotid, GPUs, ngpugroup);
    std::vector<float> objf(gs.size(), 0.f);
    for (size_t i = 0; i < gs.size(); ++i)
        objFuncCij3d_f(&objf[i], souloc, ns, gs[i].shots, recloc, wavelet, cij, padboundaryCij, nx, ny, nz, nt, npad,
                       ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, gs[i].gpus);
    for (size_t i = 0; i < gs.size(); ++i) *f += objf[i];
}

// objFuncGradientVEpsDel, 3D (wave3d.h:102, objFuncGradient3d_f.cu:249): the
// model and both boundary shells to Cij (v0, eps0 and unit delta scale), the
// Cij gradient (grouped when ngpugroup > 1), its shell and water layer
// zeroed, the chain rule to (v, eps, del).  Returns the objective f of fgcij.
double objFuncGradientVEpsDel(float *gv, float *geps, float *gdel, float *souloc, int ns, std::vector<int> &shotid,
                              float *recloc, float *wavelet, float *v, float *eps, float *del, float *padboundaryV,
                              float *randboundaryV, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                              float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                              float v0, float eps0, float wbottom, std::vector<int> &GPUs, int ngpugroup)
{
    const long long nxyz = (long long)nx * ny * nz;
    const long long nb = vti_boundary_size(nx, ny, nz, npad);
    std::vector<float> cij(3 * (size_t)nxyz), fgcij(3 * (size_t)nxyz + 1, 0.f);
    std::vector<float> beps((size_t)nb), bdel((size_t)nb), padc(3 * (size_t)nb), randc(3 * (size_t)nb);
    vti_compat::report(vti_get_boundary(beps.data(), eps, nx, ny, nz, npad), "objFuncGradientVEpsDel");
    vti_compat::report(vti_get_boundary(bdel.data(), del, nx, ny, nz, npad), "objFuncGradientVEpsDel");
    float *c = cij.data();
    vti_compat::report(vti_vepsdel_to_cij(c, c + nxyz, c + 2 * nxyz, v, eps, del, v0, eps0, 1.f, nxyz),
                       "objFuncGradientVEpsDel");
    vti_compat::report(vti_vepsdel_to_cij(padc.data(), padc.data() + nb, padc.data() + 2 * nb, padboundaryV,
                                          beps.data(), bdel.data(), v0, eps0, 1.f, nb),
                       "objFuncGradientVEpsDel");
    vti_compat::report(vti_vepsdel_to_cij(randc.data(), randc.data() + nb, randc.data() + 2 * nb, randboundaryV,
                                          beps.data(), bdel.data(), v0, eps0, 1.f, nb),
                       "objFuncGradientVEpsDel");
    if (ngpugroup > 1)
        objFuncGradientCij3d(fgcij.data(), souloc, ns, shotid, recloc, wavelet, c, padc.data(), randc.data(), nx, ny,
                             nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, GPUs, ngpugroup);
    else
        objFuncGradientCij3d_f(fgcij.data(), souloc, ns, shotid, recloc, wavelet, c, padc.data(), randc.data(), nx,
                               ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, GPUs);
    const double objfunc = fgcij[0];
    float *g = fgcij.data() + 1;
    float *gf[3] = {g, g + nxyz, g + 2 * nxyz};
    vti_compat::mask_shell_and_water(gf, 3, nx, ny, nz, npad, oz, dz, wbottom);
    vti_compat::report(vti_gradcij_to_gradvepsdel(gv, geps, gdel, gf[0], gf[1], gf[2], v, eps, del, v0, eps0, 1.f,
                                                  nxyz),
                       "objFuncGradientVEpsDel");
    return objfunc;
}

// modelData3d_f, group form (wave3d.h:108, modelData3d_f.cu:395): each shot
// group's modelDataCij3d_f with that group's GPUs
void modelData3d_f(float *ddata, float *souloc, int ns, std::vector<int> &shotid, float *recloc,
                   const float *wavelet, float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad,
                   float ox, float oy, float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate,
                   std::vector<int> &GPUs, int ngpugroup)
{
    std::vector<vti_compat::Group> gs = vti_compat::groups(shotid, GPUs, ngpugroup);
    for (size_t i = gs.size(); i-- > 0;)
        modelDataCij3d_f(ddata, souloc, ns, gs[i].shots, recloc, wavelet, c11, c13, c33, nx, ny, nz, nt, npad, ox, oy,
                         oz, ot, dx, dy, dz, dt, samplingRate, gs[i].gpus);
}

#endif  // VTI_COMPAT_IMPLEMENTATION

#endif  // VTI_COMPAT_HPP
