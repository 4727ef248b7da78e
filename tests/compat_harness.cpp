// Test harness (test infrastructure): the reference's C++ driver signatures,
// as defined by include/vti_compat.hpp, behind extern "C" entry points that a
// ctypes test can call with flat arrays.  Each entry point builds the
// std::vector arguments the reference's callers pass (shot ids, the GPU
// list) and calls the shim exactly as fwi3d-lbfgsc.cu, rtm3d.cu,
// objFuncGradient3d-cij.cu and gradient-param.cu call the reference.
// Built into tests/libcompat_harness.so by the package Makefile.
#include <vector>

#define VTI_COMPAT_IMPLEMENTATION
#include "vti_compat.hpp"

namespace {
std::vector<int> ids(const int *p, int n) { return std::vector<int>(p, p + n); }
std::vector<int> gpus(int n)
{
    std::vector<int> g;
    for (int i = 0; i < n; ++i) g.push_back(i);
    return g;
}
}  // namespace

extern "C" {

void h_synthetic_f(float *d, const float *c11, const float *c13, const float *c33, const float *w, const int *sloc,
                   int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx,
                   float dz, float dt, float rate)
{
    synthetic_f(d, c11, c13, c33, w, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate);
}

void h_modeling_f(float *d, const float *c11, const float *c13, const float *c33, const float *w, int slocxz,
                  const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate)
{
    modeling_f(d, c11, c13, c33, w, slocxz, taper, nx, nz, nt, npad, dx, dz, dt, rate);
}

float h_objFuncGradient_f(float *g11, float *g13, float *g33, const float *data, const float *c11, const float *c13,
                          const float *c33, const float *w, const int *sloc, int ns, const int *rloc, int nr,
                          const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                          float rate)
{
    return objFuncGradient_f(g11, g13, g33, data, c11, c13, c33, w, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx,
                             dz, dt, rate);
}

void h_objFuncGradientCij3d_f(float *fgcij, float *souloc, int ns, const int *shot, int nshot, float *recloc,
                              float *w, float *cij, float *padb, float *randb, int nx, int ny, int nz, int nt, int npad,
                              float o, float d, float dt, float rate, int ngpu, const float *obs)
{
    vti_compat::set_observed_data(obs);
    std::vector<int> s = ids(shot, nshot), g = gpus(ngpu);
    objFuncGradientCij3d_f(fgcij, souloc, ns, s, recloc, w, cij, padb, randb, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d,
                           d, dt, rate, g);
}

void h_objFuncGradientCij3d(float *fgcij, float *souloc, int ns, const int *shot, int nshot, float *recloc, float *w,
                            float *cij, float *padb, float *randb, int nx, int ny, int nz, int nt, int npad, float o,
                            float d, float dt, float rate, int ngpu, int ngroup, const float *obs)
{
    vti_compat::set_observed_data(obs);
    std::vector<int> s = ids(shot, nshot), g = gpus(ngpu);
    objFuncGradientCij3d(fgcij, souloc, ns, s, recloc, w, cij, padb, randb, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d,
                         d, dt, rate, g, ngroup);
}

void h_objFuncCij3d(float *f, float *souloc, int ns, const int *shot, int nshot, float *recloc, float *w, float *cij,
                    float *padb, int nx, int ny, int nz, int nt, int npad, float o, float d, float dt, float rate,
                    int ngpu, int ngroup, const float *obs)
{
    vti_compat::set_observed_data(obs);
    std::vector<int> s = ids(shot, nshot), g = gpus(ngpu);
    objFuncCij3d(f, souloc, ns, s, recloc, w, cij, padb, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d, d, dt, rate, g,
                 ngroup);
}

double h_objFuncGradientVEpsDel3d(float *gv, float *geps, float *gdel, float *souloc, int ns, const int *shot,
                                  int nshot, float *recloc, float *w, float *v, float *eps, float *del, float *padv,
                                  float *randv, int nx, int ny, int nz, int nt, int npad, float o, float d, float dt,
                                  float rate, float v0, float eps0, float wbottom, int ngpu, int ngroup,
                                  const float *obs)
{
    vti_compat::set_observed_data(obs);
    std::vector<int> s = ids(shot, nshot), g = gpus(ngpu);
    return objFuncGradientVEpsDel(gv, geps, gdel, souloc, ns, s, recloc, w, v, eps, del, padv, randv, nx, ny, nz, nt,
                                  npad, o, o, o, 0.f, d, d, d, dt, rate, v0, eps0, wbottom, g, ngroup);
}

void h_rtmCij3d_f(float *image, float *souloc, int ns, const int *shot, int nshot, float *recloc3, float *w,
                  float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad, float o, float d,
                  float dt, float rate, int ngpu, const float *obs)
{
    vti_compat::set_observed_data(obs);
    vti_compat::set_ngpu(ngpu);
    std::vector<int> s = ids(shot, nshot);
    rtmCij3d_f(image, souloc, ns, s, recloc3, w, c11, c13, c33, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d, d, dt, rate);
}

void h_rtmVEpsDel(float *image, float *souloc, int ns, const int *shot, int nshot, float *recloc3, float *w, float *v,
                  float *eps, float *del, int nx, int ny, int nz, int nt, int npad, float o, float d, float dt,
                  float rate, float wbottom, int ngpu, const float *obs)
{
    vti_compat::set_observed_data(obs);
    vti_compat::set_ngpu(ngpu);
    std::vector<int> s = ids(shot, nshot);
    rtmVEpsDel(image, souloc, ns, s, recloc3, w, v, eps, del, nullptr, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d, d, dt,
               rate, wbottom);
}

void h_modelData3d_f(float *dd, float *souloc, int ns, const int *shot, int nshot, float *recloc, float *w,
                     float *c11, float *c13, float *c33, int nx, int ny, int nz, int nt, int npad, float o, float d,
                     float dt, float rate, int ngpu, int ngroup)
{
    std::vector<int> s = ids(shot, nshot), g = gpus(ngpu);
    modelData3d_f(dd, souloc, ns, s, recloc, w, c11, c13, c33, nx, ny, nz, nt, npad, o, o, o, 0.f, d, d, d, dt, rate,
                  g, ngroup);
}

}  // extern "C"
