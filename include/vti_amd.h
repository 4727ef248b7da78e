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
typedef struct vti2d_batch_plan vti2d_batch_plan;

/* ---- status and context ------------------------------------------------ */
const char *vti_last_error(void);
int vti_ctx_create(vti_ctx **ctx, int device);
int vti_ctx_destroy(vti_ctx *ctx);
int vti_ctx_synchronize(vti_ctx *ctx);
/* Arithmetic of the 3D bulk stencil for propagators and drivers created on
 * ctx afterwards: 1 (default) evaluates the reference's expressions exactly
 * as lib/vti-fwi-3d/gpu/kernel3d.cu:54-69 does -- double weights, products
 * and sums over float storage; 0 evaluates them in FP32. */
int vti_ctx_set_arith(vti_ctx *ctx, int arith);
/* Snapshot storage of the 2D objective + gradient (vti2d_objfunc_gradient
 * and its v/eps/del wrapper) on ctx: 0 = FP32 (default, the reference's
 * storage, objFuncGradient_f.cu:188-202), 1 = FP16, 2 = BF16 with a
 * power-of-two scale per slice, 3 = S16 (16-bit integers with a power-of-two
 * scale per 64-point piece of a grid row).  A 16-bit format applies to calls whose
 * arguments the batched driver takes (the init_abc taper, ot = 0, rate/dt an
 * integer >= 2 dividing nt-1, sources and receivers in the interior); those
 * run through it, also for one shot, with half the snapshot memory.  Other
 * calls keep FP32.  The 3D drivers store no snapshots (they rebuild the
 * source wavefield from a random boundary), so the setting does not reach
 * them.  No counterpart in the reference. */
int vti_ctx_set_snapshots(vti_ctx *ctx, int snap_fmt);

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
int vti3d_prop_set_arith(vti3d_prop *prop, int arith);
/* Forward levels advanced per launch by vti3d_prop_forward_steps: 1 or 2
 * (two levels per launch, a temporally blocked kernel; the same bits). */
int vti3d_prop_set_levels(vti3d_prop *p, int levels);
int vti3d_prop_set_model(vti3d_prop *prop, const float *c11, const float *c13, const float *c33);
int vti3d_prop_reset(vti3d_prop *prop, long long sou_index, float value);
/* The two time levels as the propagator holds them.  In absorbing runs the
 * reference damps the current level in place after each block's step
 * (abc3d.cu:82-125), so the level it reads as prev carries the factor d =
 * dxy*dz twice; the propagator stores every level damped once and applies
 * the second factor when it reads prev.  So its prev equals the reference's
 * prev divided by d at the points the steps update (cur is the same). */
int vti3d_prop_set_state(vti3d_prop *prop, const float *cur_x, const float *cur_z, const float *prev_x,
                         const float *prev_z);
int vti3d_prop_get_state(vti3d_prop *prop, float *cur_x, float *cur_z, float *prev_x, float *prev_z);
int vti3d_prop_forward_steps(vti3d_prop *prop, int nsteps, int absorbing, float *ms_per_step);
int vti3d_prop_forward_shot(vti3d_prop *prop, int lmax, const float *wavelet, long long sou_index, int nr,
                            const long long *rec, int sts, int nnt, int absorbing, float *traces_tm);
/* Fused backward levels on the current state (mode 0: gradientKernel +
 * forwardKernel on lambda + imaging, objFuncGradient3d_f.cu:1008-1028;
 * mode 1: RTM, rtm3d_f.cu:585-610), no source/receivers; for timing. */
int vti3d_prop_backward_steps(vti3d_prop *prop, int mode, int nsteps, float *ms_per_step);
/* The adjoint wavefield (lambda_x, lambda_z at the current and previous
 * level) that the fused backward levels advance, and the FP32 accumulators
 * they add into: gc11, gc13, gc33 for mode 0, the image for mode 1 -- the
 * device-resident counterparts of curLambdaX/Z, prevLambdaX/Z and gc11/13/33
 * (objFuncGradient3d_f.cu:916-925) and image (rtm3d_f.cu:178).  set_adjoint
 * loads the state (reference layout; NULL leaves a field as is) and clears
 * the accumulators; get_adjoint reads them back (NULL skips a field). */
int vti3d_prop_set_adjoint(vti3d_prop *prop, int mode, const float *cur_x, const float *cur_z, const float *prev_x,
                           const float *prev_z);
int vti3d_prop_get_adjoint(vti3d_prop *prop, float *cur_x, float *cur_z, float *prev_x, float *prev_z, float *acc0,
                           float *acc1, float *acc2);

/* ---- 3D drivers with the reference's semantics ----------------------------
 * Geometry arrays as the reference takes them: souloc = 5 floats per shot
 * {x, y, z, nr, first receiver}, recloc = 4 floats per receiver
 * {x, y, z, live} (3 for RTM), positions in the units of o/d.  shotid lists
 * the shots to run.  ngpu is the number of GPUs the reference would have used:
 * it only fixes how many time levels its pipeline advances
 * (nround*ngpu*64, objFuncGradient3d_f.cu:539); all work runs on ctx's device.
 * obs stands in for the reference's read("data") inside the drivers:
 * trace-major, nnt = (nt-1)/round(samplingRate/dt)+1 samples per receiver. */

/* modelDataCij3d_f, lib/vti-fwi-3d/gpu/modelData3d_f.cu:427 (wave3d.h:110) */
int vti3d_model_data(vti_ctx *ctx, float *ddata, const float *souloc, int ns, const int *shotid, int nshot,
                     const float *recloc, const float *wavelet, const float *c11, const float *c13, const float *c33,
                     int nx, int ny, int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx,
                     float dy, float dz, float dt, float samplingRate, int ngpu);

/* objFuncCij3d_f, lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:1272 (wave3d.h:100).
 * f[0] = half the objective.  tpow: the tpow= gain (0 = none,
 * agc.cpp:103-115); objtype/halfwidth: the objtype= and halfwidth= parameters
 * (0 = least squares, 1 = AGC residual of agc.cpp:70-101 over a circular
 * window of 2*halfwidth+1 samples; the reference's default half width is 20). */
int vti3d_objfunc(vti_ctx *ctx, float *f, const float *souloc, int ns, const int *shotid, int nshot,
                  const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij, int nx, int ny,
                  int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
                  float dt, float samplingRate, int ngpu, const float *obs, float tpow, int objtype, int halfwidth);

/* objFuncGradientCij3d_f, lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:361 (wave3d.h:96).
 * fgcij = {f, gc11[nxyz], gc13[nxyz], gc33[nxyz]}; cij = {c11, c13, c33}
 * whose boundary shell is left holding randboundaryCij, as in the reference. */
int vti3d_objfunc_gradient(vti_ctx *ctx, float *fgcij, const float *souloc, int ns, const int *shotid, int nshot,
                           const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij,
                           const float *randboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                           float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate, int ngpu,
                           const float *obs, float tpow, int objtype, int halfwidth);

/* rtmCij3d_f, lib/vti-fwi-3d/gpu/rtm3d_f.cu:161 (wave3d.h:88).  recloc stride 3. */
int vti3d_rtm(vti_ctx *ctx, float *image, const float *souloc, int ns, const int *shotid, int nshot,
              const float *recloc, const float *wavelet, const float *c11, const float *c13, const float *c33, int nx,
              int ny, int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
              float dt, float samplingRate, int ngpu, const float *obs);

/* ---- 2D drivers with the reference's semantics ----------------------------
 * Arguments exactly as lib/vti-fwi-2d/wave.h declares them (padded grid
 * nnx = nx+2npad, nnz = nz+2npad; sloc = 4 ints per shot {ix, iz, nr_shot,
 * first receiver}; rloc = 2 ints per receiver {ix, iz}; taper from
 * vti_init_taper2d).  The reference reads the GPU list from the global argv
 * ("gpu="); all work here runs on ctx's device. */

/* synthetic_f, lib/vti-fwi-2d/synthetic_f.cu:13 (wave.h:4).  data: nr columns
 * (every shot's receivers side by side) x ((nt-ntNeg-1)/ratio+1) rows. */
int vti2d_synthetic(vti_ctx *ctx, float *data, const float *c11, const float *c13, const float *c33,
                    const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                    int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot);

/* modeling_f, lib/vti-fwi-2d/modeling_f.cu:10-86 (wave.h:14): one source at
 * the padded-grid index slocxz (inside the updated interior); wavefield holds
 * (nt - ntNeg - 1)/ratio + 1 slots of nnx*nnz floats (ratio = round(rate/dt),
 * ntNeg = round(|ot|/dt)); slot (it-ntNeg)/ratio receives the p field after
 * step it (the source adds w[it-1], with no dt^2, as there), slot 1 level 1
 * when ratio = 1 and ot = 0; slots never recorded keep their input values. */
int vti2d_modeling(vti_ctx *ctx, float *wavefield, const float *c11, const float *c13, const float *c33,
                   const float *wavelet, int slocxz, const float *taper, int nx, int nz, int nt, int npad, float dx,
                   float dz, float dt, float rate, float ot);

/* vti2d_synthetic with no allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same data as vti2d_synthetic (its one-shot schedule, bit
 * for bit); every device buffer lives in a fixed device array of about
 * 52 MiB per GPU, held until the process ends.  Limits (kErrArg beyond
 * them): nnx*nnz <= 2^20 grid points, <= 65536 receivers per shot,
 * receivers x recorded rows <= 2^22 per shot.  Calls on one GPU share the
 * array: calls that overlap on one GPU (two contexts, two threads) race.
 * The only host allocation is the process-wide divisor memo the first call
 * per distinct dx, dz fills. */
int vti2d_synthetic_noalloc(vti_ctx *ctx, float *data, const float *c11, const float *c13, const float *c33,
                            const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                            const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                            float rate, float ot);

/* objFuncGradient_f, lib/vti-fwi-2d/objFuncGradient_f.cu:15 (wave.h:12).
 * *objfunc receives the value the reference returns (0.5*sum r^2).
 * Requires rate/dt >= 2, (nt-1) % (rate/dt) == 0 and ot == 0, the conditions
 * under which the reference's snapshot buffers line up. */
int vti2d_objfunc_gradient(vti_ctx *ctx, float *objfunc, float *gc11, float *gc13, float *gc33, const float *data,
                           const float *c11, const float *c13, const float *c33, const float *wavelet, const int *sloc,
                           int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
                           float dx, float dz, float dt, float rate, float ot);

/* hessian_f, lib/vti-fwi-2d/hessian_f.cu:15 (wave.h:32): the full
 * Hessian-vector product H*dc -- the Gauss-Newton part plus the second-order
 * terms of the residual's adjoint (scatteringa), with the exact adjoint
 * backwardDC.  The reference fills its Born snapshot pipeline's first buffer
 * from the background snapshots (hessian_f.cu:381-382); this keeps that, so
 * the product is affine in dc.  Same schedule conditions as
 * vti2d_objfunc_gradient, and (nt-1)/(rate/dt) >= 3. */
int vti2d_hessian(vti_ctx *ctx, float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                  const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
                  const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                  int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot);

/* rtm_f, lib/vti-fwi-2d/rtm_f.cu:13 (wave.h:6): dipole source and receivers
 * (the pole one row below each point subtracted), the receiver side
 * propagated with the forward operator, image += 0.5 (p+q) x the source
 * wavefield slice interpolated in time.  image: nnx*nnz.  Needs
 * (nt-1) % (rate/dt) == 0 (otherwise the reference reads a slice past its
 * array), the source, every receiver and the points below them inside the
 * updated interior, and no receiver right below another. */
int vti2d_rtm(vti_ctx *ctx, float *image, const float *data, const float *c11, const float *c13, const float *c33,
              const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz,
              int nt, int npad, float dx, float dz, float dt, float rate, float ot);

/* odcig_f, lib/vti-fwi-2d/odcig_f.cu:12 (wave.h:16): rtm_f with the
 * subsurface-offset imaging condition, image[h][i] += 0.5 (p+q)(i-h) x
 * slice(i+h), h = -nhx..nhx; image: (2*nhx+1) planes of nnx*nnz.  Keeps the
 * reference's source scaling: dt2*wavelet[0] at the first level, wavelet[it-1]
 * (no dt2) after it (odcig_f.cu:111,123). */
int vti2d_odcig(vti_ctx *ctx, float *image, const float *data, const float *c11, const float *c13, const float *c33,
                const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                int nz, int nhx, int nt, int npad, float dx, float dz, float dt, float rate, float ot);

/* GNhessian_f, lib/vti-fwi-2d/hessian_f.cu:575 (wave.h:34). */
int vti2d_gn_hessian(vti_ctx *ctx, float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                     const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
                     const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                     int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot);

/* ---- 2D v/eps/delta parameterisation ---------------------------------------- */

/* checkEpsDel, lib/vti-fwi-2d/check.cpp:22-40: projects points with
 * del0*del > eps0*eps onto del0*del = eps0*eps, in place; m[i] = 1 where they
 * are equal, 2 where projected, 0 elsewhere.  Returns the number projected. */
int vti_check_eps_del(float *eps, float *del, float eps0, float del0, long long n, float *m);

/* objFuncGradientVEpsDel, lib/vti-fwi-2d/objFuncGradient.cu:28-54 (wave.h:10):
 * checkEpsDel on eps/del (in place, mask in m), VEpsDel2Cij, objFuncGradient_f,
 * the chain rule to gv/geps/gdel, and zero rows [0, npad + (int)(wbottom/dz+1)).
 * v/eps/del are divided by v0/eps0/del0 as the reference takes them. */
int vti2d_objfunc_gradient_vepsdel(vti_ctx *ctx, float *objfunc, float *gv, float *geps, float *gdel,
                                   const float *data, float *v, float *eps, float *del, float v0, float eps0,
                                   float del0, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                                   const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                                   float rate, float ot, float wbottom, float *m);

/* GNhessianVEpsDel, lib/vti-fwi-2d/hessians.cu:72-95 (wave.h:42): the
 * Gauss-Newton product for a perturbation (dv, deps, ddel). */
int vti2d_gn_hessian_vepsdel(vti_ctx *ctx, float *gv, float *geps, float *gdel, const float *data, float *v,
                             float *eps, float *del, const float *dv, const float *deps, const float *ddel, float v0,
                             float eps0, float del0, const float *wavelet, const int *sloc, int ns, const int *rloc,
                             int nr, const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt,
                             float rate, float ot, float wbottom, float *m);

/* ---- batched 2D gradient --------------------------------------------------------
 * Many shots, each on its own aperture sub-grid of one global model
 * (NNX = NX + 2*npad columns, nnz = nz + 2*npad rows, the reference layout),
 * advanced in lockstep: one launch per time level for all shots of a wave.
 * Two waves are in flight at once, each starting its forward pass when the
 * previous one's is done, so that a wave's forward pass overlaps the
 * previous wave's adjoint: waves of (max_wave+1)/2 shots when max_wave > 0
 * (about max_wave shots held at once), else sized to half the free device
 * memory.  Per shot the result is objFuncGradient_f
 * (lib/vti-fwi-2d/objFuncGradient_f.cu:15) on the model cropped to the shot's
 * columns [c0, c0 + nx + 2*npad), with the standard init_abc taper of that
 * sub-grid; the shots' gradients are summed in FP64 into gc11/gc13/gc33
 * (NNX*nnz doubles each), *objfunc = 0.5*sum r^2 over all shots.
 * shots: 6 ints per shot {c0, nx, ix_src, iz_src, nr_shot, first receiver};
 * rloc: 2 ints per receiver {ix, iz} in the shot's sub-grid coordinates;
 * data: the shots' observed traces one after the other, each time-major
 * ((nt-1)/(rate/dt)+1 rows of nr_shot).  Same schedule conditions as
 * vti2d_objfunc_gradient (rate/dt >= 2, (nt-1) % (rate/dt) == 0, ot = 0).
 * levels: time levels advanced per launch, 1 or 2 (0: 2); tile_rows: rows
 * of a two-level workgroup tile, 4, 8 or 16 (0: 16).  Neither changes a bit of
 * the result.  snap_fmt: 0 keeps the imaging snapshots in FP32 (the result
 * then equals the one-shot driver's bit for bit); 1 (FP16) and 2 (BF16)
 * store each snapshot slice scaled by a power of two, 3 (S16) 16-bit
 * integers with a power-of-two scale per 64-point row piece, halving their memory;
 * the gradient then differs at the level the format's rounding sets.  The
 * wave split changes no bit of the result.
 * No counterpart in the reference, which runs one shot per call. */
int vti2d_objfunc_gradient_batch(vti_ctx *ctx, double *objfunc, double *gc11, double *gc13, double *gc33,
                                 const float *data, const float *c11, const float *c13, const float *c33,
                                 const float *wavelet, const int *shots, int ns, const int *rloc, int nr, int NX, int nz,
                                 int nt, int npad, float dx, float dz, float dt, float rate, int max_wave, int levels,
                                 int tile_rows, int snap_fmt);

/* The batched gradient captured once and replayed: for the iterations of an
 * inversion, which evaluate the same shots (geometry, observed data, wavelet)
 * on a new model each time.  vti2d_batch_plan_create checks the arguments as
 * vti2d_objfunc_gradient_batch does, splits the shots into waves as the
 * direct call does (half of max_wave shots each when > 0, within the free
 * device memory, two in flight), allocates two sets of wavefields,
 * snapshots and trace buffers (one for the even waves, one for the odd
 * ones), uploads each wave's shot layout, receiver maps and observed data
 * (an error when the whole does not fit the free device memory), and
 * records each wave's launch sequence -- forward pass and residual, then
 * adjoint pass and FP64 accumulation, every buffer clearing done by a
 * kernel -- into two hipGraphs on the context's stream.
 * vti2d_batch_plan_run uploads c11/c13/c33 and replays the graphs, wave k
 * on the context's stream or (odd k) a second stream the plan holds, each
 * forward pass beside the previous wave's adjoint, and returns what
 * vti2d_objfunc_gradient_batch returns for the same arguments, bit for
 * bit.  The plan holds its device memory until
 * vti2d_batch_plan_destroy, which must come before vti_ctx_destroy.  Runs of
 * one plan, or of plans on one context, are serialised by the caller.
 * No counterpart in the reference, which sets every call up anew. */
int vti2d_batch_plan_create(vti_ctx *ctx, vti2d_batch_plan **plan, const float *data, const float *wavelet,
                            const int *shots, int ns, const int *rloc, int nr, int NX, int nz, int nt, int npad,
                            float dx, float dz, float dt, float rate, int max_wave, int levels, int tile_rows,
                            int snap_fmt);
int vti2d_batch_plan_run(vti2d_batch_plan *plan, double *objfunc, double *gc11, double *gc13, double *gc33,
                         const float *c11, const float *c13, const float *c33);
int vti2d_batch_plan_destroy(vti2d_batch_plan *plan);

/* Launch-shape autotuner for vti2d_objfunc_gradient_batch: times short runs
 * (24 snapshot intervals) of the same shots on a synthetic model over levels
 * per launch (1, 2) and tile rows (8, 16), and returns the fastest in
 * *levels, *tile_rows (and its trial time in *best_ms when not NULL).
 * *max_wave is modelled, not timed: the most shots whose adjoint working set
 * (16 FP32 planes per shot) fits 192 MB of the Infinity Cache, or 0 when all
 * shots do.  Cached per geometry for the life of the process.  No
 * counterpart in the reference. */
int vti2d_batch_tune(vti_ctx *ctx, const int *shots, int ns, const int *rloc, int nr, int NX, int nz, int nt,
                     int npad, float dx, float dz, float dt, float rate, int *levels, int *tile_rows, int *max_wave,
                     double *best_ms);

/* ---- L-BFGS (FP64) -----------------------------------------------------------
 * lbfgs, lib/lbfgs/lbfgs_c.cpp:11 (lbfgs.h:32) with mcsrch/mcstep,
 * lib/lbfgs/lsearch.cpp:9-606: reverse communication.  Start with *iflag = 0
 * and f, g at x; while the call returns *iflag = 1, evaluate f and g at the
 * new x and call again; *iflag = 0 on convergence (|g|/max(1,|x|) <= 1e-5),
 * -1 when the line search fails (vti_last_error() says why).  The state
 * (m corrections, line search) is FP64; the diagonal H0 is (y's)/(y'y).
 * unit_step = 0 keeps the reference's choice of starting every line search
 * after the first at the step the previous one accepted (lbfgs_c.cpp:433-436);
 * 1 restarts at stp = 1 as in Liu & Nocedal's method. */
typedef struct vti_lbfgs vti_lbfgs;
int vti_lbfgs_create(vti_lbfgs **state, long long n, int m, int unit_step);
/* The same method with its vectors resident in device memory (FP64: the m
 * stored steps and gradient differences, the direction, the base point, the
 * gradients), every vector operation a kernel on ctx's stream and the dot
 * products reduced there in a fixed order (lbfgs_c.cpp:226-240,350-428,
 * la.cpp:7-31); vti_lbfgs_step then takes x and g in host or device memory. */
int vti_lbfgs_create_device(vti_lbfgs **state, vti_ctx *ctx, long long n, int m, int unit_step);
int vti_lbfgs_destroy(vti_lbfgs *state);
int vti_lbfgs_step(vti_lbfgs *state, float *x, double f, const float *g, int *iflag, int *iter, int *nfun);

/* ---- request server ------------------------------------------------------------
 * No counterpart in the reference, whose drivers run one call at a time.
 * A server on one device runs 2D gradient and 3D modeling requests of any
 * grid size as they arrive: `slots` worker threads (1..16), each with its own
 * context and stream, take queued requests in arrival order as soon as a slot
 * is free and the request's device-memory estimate fits in what the budget
 * (bytes; <= 0: the free memory less 2 GiB) leaves after the resident models
 * and the requests in flight; a request that would not fit even alone is
 * refused when submitted.  Requests on different slots run concurrently on
 * the device.  Each request's result equals that of the same driver call
 * made alone.  Models are uploaded once (put_model: c11, c13, c33 of n
 * points, the layout the drivers take) and stay resident until dropped.
 * Every host buffer a request names must stay valid until vti_serve_wait
 * returns for it.  Latency = completion - submission; vti_serve_stats gives
 * the p50/p99 over completed requests (nearest rank), the completion rate
 * (completed / (last completion - first submission)) and the peak of resident
 * plus in-flight bytes. */
typedef struct vti_server vti_server;
int vti_serve_create(vti_server **out, int device, int slots, double budget_bytes);
int vti_serve_destroy(vti_server *s); /* drains the queue first */
int vti_serve_put_model(vti_server *s, const float *c11, const float *c13, const float *c33, long long n, int *handle);
int vti_serve_drop_model(vti_server *s, int handle);
/* vti2d_objfunc_gradient's arguments (ot = 0) on a resident model */
int vti_serve_grad2d(vti_server *s, long long *id, int model, float *objfunc, float *gc11, float *gc13, float *gc33,
                     const float *data, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                     const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate);
/* vti3d_model_data's arguments (ngpu = 1) on a resident model; nr = the
 * number of receivers in recloc */
int vti_serve_model3d(vti_server *s, long long *id, int model, float *ddata, const float *souloc, int ns,
                      const int *shotid, int nshot, const float *recloc, int nr, const float *wavelet, int nx, int ny,
                      int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
                      float dt, float samplingRate);
/* Kernel-level batching of 2D gradient requests: a worker merges up to
 * max_requests queued requests that share a schedule (nz, npad, nt, dx, dz,
 * dt, rate, wavelet) and use the standard taper into one batched run over
 * their models placed side by side; each still gets the bits
 * vti2d_objfunc_gradient returns for it alone.  1 turns merging off; the
 * default is 8.  vti_serve_batch_stats returns how many merged runs there
 * were and how many requests they held.  No counterpart in the reference. */
int vti_serve_set_batching(vti_server *s, int max_requests);
int vti_serve_batch_stats(vti_server *s, long long *merged_runs, long long *merged_requests);
/* Snapshot storage of the 2D gradient requests that workers start from now
 * on, as vti_ctx_set_snapshots sets it for a context (0 = FP32, the default;
 * 1 = FP16, 2 = BF16, 3 = S16); merged runs and single requests alike.  No
 * counterpart in the reference. */
int vti_serve_set_snapshots(vti_server *s, int snap_fmt);
/* on = 1 (default): a 3D modeling request starts only when nothing else
 * runs, and nothing starts beside it.  The 3D step is bound by HBM
 * bandwidth, so work beside it only stretches it; alone it keeps the
 * latency the one-request-at-a-time path gives it, while 2D requests still
 * overlap and merge with each other.  on = 0: 3D requests share the device
 * like the rest.  No counterpart in the reference. */
int vti_serve_set_exclusive3d(vti_server *s, int on);

/* blocks until request `id` is done; returns its status (vti_last_error()
 * then holds its message) and its latency and service time in seconds */
int vti_serve_wait(vti_server *s, long long id, double *latency_s, double *service_s);
int vti_serve_stats(vti_server *s, double *p50, double *p99, double *req_per_s, long long *completed,
                    double *peak_bytes);

#ifdef __cplusplus
}
#endif

#endif /* VTI_AMD_H */
