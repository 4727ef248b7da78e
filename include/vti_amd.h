/*
 * vti_amd.h -- C ABI of libvtiamd.so, the MI355X-native implementation of
 * the pseudo-acoustic VTI wave-equation path of huyleq/thesis-source-code
 * (lib/vti-fwi-2d, lib/vti-fwi-3d).
 *
 * Conventions
 *  - Plain C: pointers, 32/64-bit integers and floats; no C++ types.
 *  - Every array argument is host memory owned by the caller, in the
 *    reference's layout (x fastest: 2D i = ix + iz*nnx, 3D
 *    j = ix + iy*nx + iz*nx*ny).  Device memory is the library's business.
 *  - Functions return 0 on success or a negative status; vti_last_error()
 *    gives a message for the calling thread's last failure.
 *  - A vti_ctx binds one device and one HIP stream.  Calls on different
 *    contexts may run concurrently; calls on one context are serialised by
 *    the caller.
 *
 * Each entry point names the reference interface it replaces (file:line,
 * paths relative to the reference repository).
 */
#ifndef VTI_AMD_H
#define VTI_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vti_ctx vti_ctx;
typedef struct vti3d_prop vti3d_prop;

/* ---- status and context ------------------------------------------------ */
const char *vti_last_error(void);
int vti_ctx_create(vti_ctx **ctx, int device);
int vti_ctx_destroy(vti_ctx *ctx);
int vti_ctx_synchronize(vti_ctx *ctx);

/* ---- host helpers (bit-exact restatements of the reference's host code) - */

/* init_abc(damping,nx,ny,nz,npad), lib/vti-fwi-3d/gpu/abc3d.cu:23-50.
 * damping has nx*ny + nz floats. */
int vti_init_damping3d(float *damping, int nx, int ny, int nz, int npad);

/* init_abc(taper,nx,nz,npad), lib/vti-fwi-2d/init.cpp:70-89.
 * nx, nz are the unpadded sizes; taper has (nx+2npad)*(nz+2npad) floats. */
int vti_init_taper2d(float *taper, int nx, int nz, int npad);

/* VEpsDel2Cij, lib/vti-fwi-2d/conversions.cpp:42-50 */
int vti_vepsdel_to_cij(float *c11, float *c13, float *c33, const float *v, const float *eps, const float *del,
                       float v0, float eps0, float del0, long long n);
/* dVEpsDel2dCij, lib/vti-fwi-2d/conversions.cpp:52-67 */
int vti_dvepsdel_to_dcij(float *c11, float *c13, float *c33, float *dc11, float *dc13, float *dc33, const float *v,
                         const float *eps, const float *del, const float *dv, const float *deps, const float *ddel,
                         float v0, float eps0, float del0, long long n);
/* GradCij2GradVEpsDel, lib/vti-fwi-2d/conversions.cpp:69-81 */
int vti_gradcij_to_gradvepsdel(float *gv, float *geps, float *gdel, const float *gc11, const float *gc13,
                               const float *gc33, const float *v, const float *eps, const float *del, float v0,
                               float eps0, float del0, long long n);

/* getBoundary / putBoundary / zeroBoundary, lib/vti-fwi-3d/gpu/boundary.cpp:10-117 */
long long vti_boundary_size(int nx, int ny, int nz, int npad);
int vti_get_boundary(float *packed, const float *v, int nx, int ny, int nz, int npad);
int vti_put_boundary(const float *packed, float *v, int nx, int ny, int nz, int npad);
int vti_zero_boundary(float *v, int nx, int ny, int nz, int npad);

/* ricker, lib/vti-fwi-2d/ricker.cpp:11-21 */
int vti_ricker(float *wavelet, float freq, int nt, float dt, float ot, float tdelay, float scalefactor);

/* The first n values rand() returns after srand(seed) with glibc's default
 * generator (the sequence lib/random-boundary/randomBound.h:52 draws from). */
int vti_glibc_rand(int seed, int n, int *out);

/* SEPMY::RandomMY(field,bval,eval,{n1,n2,n3},{t1,t2,t3},pctG,seed),
 * lib/random-boundary/randomBound.h:51-276; n3 = 1 for 2D.  Bit-exact.
 * Returns the number of clusters (>0) or a negative status. */
long long vti_random_boundary(float *field, float bval, float eval, int n1, int n2, int n3, int t1, int t2, int t3,
                              float pctG, int seed);

/* ---- 3D device-resident propagator ---------------------------------------
 * Holds c11/c13/c33, the damping profile and two time levels of (sigma_x,
 * sigma_z) in HBM.  Replaces the per-call host-streamed block pipeline of
 * lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:538-735 (forwardKernel*,
 * injectSource, abcXYZ/abcXY, recordData). */
int vti3d_prop_create(vti_ctx *ctx, vti3d_prop **prop, int nx, int ny, int nz, int npad, float dx, float dy,
                      float dz, float dt);
int vti3d_prop_destroy(vti3d_prop *prop);
size_t vti3d_prop_bytes(const vti3d_prop *prop);
int vti3d_prop_set_zchunk(vti3d_prop *prop, int zchunk);
int vti3d_prop_set_model(vti3d_prop *prop, const float *c11, const float *c13, const float *c33);
int vti3d_prop_reset(vti3d_prop *prop, long long sou_index, float value);
int vti3d_prop_set_state(vti3d_prop *prop, const float *cur_x, const float *cur_z, const float *prev_x,
                         const float *prev_z);
int vti3d_prop_get_state(vti3d_prop *prop, float *cur_x, float *cur_z, float *prev_x, float *prev_z);
int vti3d_prop_forward_steps(vti3d_prop *prop, int nsteps, int absorbing, float *ms_per_step);
int vti3d_prop_forward_shot(vti3d_prop *prop, int lmax, const float *wavelet, long long sou_index, int nr,
                            const long long *rec, int sts, int nnt, int absorbing, float *traces_tm);

#ifdef __cplusplus
}
#endif

#endif /* VTI_AMD_H */
