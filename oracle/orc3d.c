/*
 * Oracle (test infrastructure only) -- CPU restatement of the reference's 3D
 * VTI path in plain C.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker.
 *
 * Restates:
 *   lib/vti-fwi-3d/gpu/kernel3d.cu:3-468     forwardKernel, gradientKernel, imagingKernel
 *   lib/vti-fwi-3d/gpu/abc3d.cu:23-125       init_abc (3D), abcXYZ / abcXY
 *   lib/vti-fwi-3d/gpu/injectRecord3d.cu     injectSource, recordData, injectResidual,
 *                                            interpolateResidual
 *   lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:361-1231  objFuncGradientCij3d_f
 *   lib/vti-fwi-3d/gpu/rtm3d_f.cu:161-830    rtmCij3d_f
 *   lib/vti-fwi-3d/gpu/boundary.cpp:10-117   getBoundary / putBoundary / zeroBoundary
 *
 * Three arithmetic modes (suffixes _f32, _ref, _f64) share one body
 * (orc3d_body.h).  Parity status of each function is listed in orc_status().
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"

typedef struct {
    int nx, ny, nz, npad, nb;
    float dt2dx2, dt2dy2, dt2dz2, dx2, dy2, dz2, dt2;
} orc3d_geom;

typedef struct {
    int nr;
    long *rec_eff;  /* index the reference records / injects at (block-local index + recBlock) */
    long *rec_true; /* true receiver index (used once, for the first adjoint sample) */
    int rec_block;
    long sou_index;
    int sou_block;
} orc3d_shot;

/* ---- instantiate the body three times ----------------------------------- */
#define T float
#define SFX f32
#define ARITH_F32
#define CK0 C0F
#define CK1 C1F
#define CK2 C2F
#define CK3 C3F
#define CK4 C4F
#define KTWOTHIRD TWOTHIRD_F
#define KONETHIRD ONETHIRD_F
#include "orc3d_body.h"
#undef T
#undef SFX
#undef ARITH_F32
#undef CK0
#undef CK1
#undef CK2
#undef CK3
#undef CK4
#undef KTWOTHIRD
#undef KONETHIRD

#define T float
#define SFX ref
#define ARITH_REF
#include "orc3d_body.h"
#undef T
#undef SFX
#undef ARITH_REF

#define T double
#define SFX f64
#define ARITH_F64
#define CK0 C0D
#define CK1 C1D
#define CK2 C2D
#define CK3 C3D
#define CK4 C4D
#define KTWOTHIRD TWOTHIRD_D
#define KONETHIRD ONETHIRD_D
#include "orc3d_body.h"
#undef T
#undef SFX
#undef ARITH_F64
#undef CK0
#undef CK1
#undef CK2
#undef CK3
#undef CK4
#undef KTWOTHIRD
#undef KONETHIRD

/* ---- host helpers ------------------------------------------------------- */

/* init_abc(damping,nx,ny,nz,npad), abc3d.cu:23-50: damping[0..nxy) holds the
 * x*y product, damping[nxy..nxy+nz) the z factor. */
void orc3d_init_damping(float *damping, int nx, int ny, int nz, int npad)
{
    for (int iy = 0; iy < ny; iy++) {
        float dist;
        if (iy < npad) dist = npad - iy;
        else if (iy > ny - npad) dist = iy - ny + npad;
        else dist = 0.;
        float dampingY = DAMPER3D + (1. - DAMPER3D) * cos(PI3D * dist / npad);
        for (int ix = 0; ix < nx; ix++) {
            if (ix < npad) dist = npad - ix;
            else if (ix > nx - npad) dist = ix - nx + npad;
            else dist = 0.;
            float dampingX = DAMPER3D + (1. - DAMPER3D) * cos(PI3D * dist / npad);
            damping[ix + iy * nx] = dampingX * dampingY;
        }
    }
    long nxy = (long)nx * ny;
    for (int iz = 0; iz < nz; iz++) {
        float dist;
        if (iz < npad) dist = npad - iz;
        else if (iz > nz - npad) dist = iz - nz + npad;
        else dist = 0.;
        damping[nxy + iz] = DAMPER3D + (1. - DAMPER3D) * cos(PI3D * dist / npad);
    }
}

/* getBoundary / putBoundary / zeroBoundary, boundary.cpp:10-117.  op: 0 get
 * (volume -> packed), 1 put (packed -> volume), 2 zero the shell. */
void orc3d_boundary(int op, float *packed, float *v, int nx, int ny, int nz, int npad)
{
    long cb = 0, cv = 0;
#define SEG(n_)                                                                  \
    do {                                                                         \
        long n__ = (n_);                                                         \
        if (op == 0) memcpy(packed + cb, v + cv, n__ * sizeof(float));           \
        else if (op == 1) memcpy(v + cv, packed + cb, n__ * sizeof(float));      \
        else memset(v + cv, 0, n__ * sizeof(float));                             \
        cb += n__;                                                               \
        cv += n__;                                                               \
    } while (0)
    SEG((long)nx * ny * npad);
    for (int iz = 0; iz < nz - 2 * npad; iz++) {
        for (int iy = 0; iy < npad; iy++) SEG(nx);
        for (int iy = 0; iy < ny - 2 * npad; iy++) {
            SEG(npad);
            cv += nx - 2 * npad;
            SEG(npad);
        }
        for (int iy = 0; iy < npad; iy++) SEG(nx);
    }
    SEG((long)nx * ny * npad);
#undef SEG
}

long orc3d_nboundary(int nx, int ny, int nz, int npad)
{
    return (long)nx * ny * nz - (long)(nx - 2 * npad) * (ny - 2 * npad) * (nz - 2 * npad);
}

/* number of time levels the reference's pipeline actually advances:
 * nround*NGPU*NUPDATE steps, objFuncGradient3d_f.cu:539 */
int orc3d_lmax(int nt, int ngpu)
{
    int nround = (nt - 2) / (ngpu * 64);
    return nround * ngpu * 64 + 1;
}

/* shot geometry exactly as objFuncGradient3d_f.cu:545-576 computes it;
 * recstride is 4 for the gradient driver and 3 for RTM (rtm3d_f.cu:348). */
static void shot_setup(orc3d_shot *sh, const float *souloc, int is, const float *recloc, int recstride, int nx,
                       int ny, float ox, float oy, float oz, float dx, float dy, float dz)
{
    int nxy = nx * ny;
    sh->nr = (int)souloc[5 * is + 3];
    int irbegin = (int)souloc[5 * is + 4];
    sh->rec_eff = (long *)malloc(sizeof(long) * (sh->nr > 0 ? sh->nr : 1));
    sh->rec_true = (long *)malloc(sizeof(long) * (sh->nr > 0 ? sh->nr : 1));
    sh->rec_block = (int)((recloc[2] - oz) / dz / 4);
    for (int ir = 0; ir < sh->nr; ir++) {
        int ir1 = ir + irbegin;
        int ix = (int)((recloc[recstride * ir1] - ox) / dx);
        int iy = (int)((recloc[recstride * ir1 + 1] - oy) / dy);
        int iz = (int)((recloc[recstride * ir1 + 2] - oz) / dz);
        int ixy = ix + iy * nx;
        sh->rec_true[ir] = (long)ixy + (long)iz * nxy;
        sh->rec_eff[ir] = (long)ixy + (long)(iz % 4) * nxy + (long)4 * sh->rec_block * nxy;
    }
    int sx = (int)((souloc[5 * is] - ox) / dx);
    int sy = (int)((souloc[5 * is + 1] - oy) / dy);
    int sz = (int)((souloc[5 * is + 2] - oz) / dz);
    sh->sou_index = (long)sx + (long)sy * nx + (long)sz * nxy;
    sh->sou_block = sz / 4;
}

static void shot_free(orc3d_shot *sh)
{
    free(sh->rec_eff);
    free(sh->rec_true);
}

static void geom_setup(orc3d_geom *g, int nx, int ny, int nz, int npad, float dx, float dy, float dz, float dt)
{
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->npad = npad;
    g->nb = nz / 4;
    g->dx2 = dx * dx;
    g->dy2 = dy * dy;
    g->dz2 = dz * dz;
    g->dt2 = dt * dt;
    g->dt2dx2 = g->dt2 / g->dx2;
    g->dt2dy2 = g->dt2 / g->dy2;
    g->dt2dz2 = g->dt2 / g->dz2;
}

/* ---- drivers ------------------------------------------------------------- */

/* Forward modeling of one source: levels 2..lmax with (damping!=NULL) or
 * without absorbing damping, recording traces.  The initial state is the
 * reference's: level 1 = dt2*wavelet[0] at the source, level 0 = 0.
 * mode: 0 f32, 1 ref, 2 f64.  traces: trace-major (nnt per receiver), sample 0
 * left as the caller set it.  Final cur/prev fields are returned when the
 * pointers are non-NULL (float for f32/ref, double for f64). */
int orc3d_forward(int mode, const float *c11, const float *c13, const float *c33, int nx, int ny, int nz,
                  int npad, float dx, float dy, float dz, float dt, int lmax, const float *wavelet, long sou_index,
                  int nr, const long *rec_idx, int sts, int nnt, int absorbing, float *traces, double *traces64,
                  void *out_cur_x, void *out_cur_z, void *out_prev_x, void *out_prev_z)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2) return -1;
    long n = (long)nx * ny * nz;
    float *damping = NULL;
    if (absorbing) {
        damping = (float *)malloc(sizeof(float) * ((long)nx * ny + nz));
        orc3d_init_damping(damping, nx, ny, nz, npad);
    }
    orc3d_shot sh;
    sh.nr = nr;
    sh.rec_eff = (long *)rec_idx;
    sh.rec_true = (long *)rec_idx;
    sh.rec_block = 0;
    sh.sou_index = sou_index;
    sh.sou_block = (int)((sou_index / ((long)nx * ny)) / 4);
    float dt2 = g.dt2;
    if (mode == 2) {
        double *b[4];
        for (int k = 0; k < 4; k++) b[k] = (double *)calloc(n, sizeof(double));
        double *C[3];
        const float *src[3] = {c11, c13, c33};
        for (int k = 0; k < 3; k++) {
            C[k] = (double *)malloc(sizeof(double) * n);
            for (long i = 0; i < n; i++) C[k][i] = src[k][i];
        }
        double *cx = b[0], *px = b[1], *cz = b[2], *pz = b[3];
        float temp = dt2 * wavelet[0];
        cx[sou_index] = temp;
        cz[sou_index] = temp;
        forward_pass_f64(&g, &sh, C[0], C[1], C[2], damping, wavelet, lmax, &cx, &px, &cz, &pz, traces, traces64,
                         sts, nnt);
        if (out_cur_x) memcpy(out_cur_x, cx, sizeof(double) * n);
        if (out_cur_z) memcpy(out_cur_z, cz, sizeof(double) * n);
        if (out_prev_x) memcpy(out_prev_x, px, sizeof(double) * n);
        if (out_prev_z) memcpy(out_prev_z, pz, sizeof(double) * n);
        for (int k = 0; k < 4; k++) free(b[k]);
        for (int k = 0; k < 3; k++) free(C[k]);
    } else {
        float *b[4];
        for (int k = 0; k < 4; k++) b[k] = (float *)calloc(n, sizeof(float));
        float *cx = b[0], *px = b[1], *cz = b[2], *pz = b[3];
        float temp = dt2 * wavelet[0];
        cx[sou_index] = temp;
        cz[sou_index] = temp;
        if (mode == 0)
            forward_pass_f32(&g, &sh, c11, c13, c33, damping, wavelet, lmax, &cx, &px, &cz, &pz, traces, traces64,
                             sts, nnt);
        else
            forward_pass_ref(&g, &sh, c11, c13, c33, damping, wavelet, lmax, &cx, &px, &cz, &pz, traces, traces64,
                             sts, nnt);
        if (out_cur_x) memcpy(out_cur_x, cx, sizeof(float) * n);
        if (out_cur_z) memcpy(out_cur_z, cz, sizeof(float) * n);
        if (out_prev_x) memcpy(out_prev_x, px, sizeof(float) * n);
        if (out_prev_z) memcpy(out_prev_z, pz, sizeof(float) * n);
        for (int k = 0; k < 4; k++) free(b[k]);
    }
    free(damping);
    return 0;
}

/* Forward levels on caller-owned float state (time-reversibility checks and
 * the CPU baseline): `nsteps` levels with the block order and damping of the
 * forward pass, no source.  mode 0 FP32, 1 the reference's arithmetic.
 * State arrays are updated in place; after each level cur/prev swap roles,
 * so after an odd number of levels the caller's cur array holds the old prev. */
int orc3d_steps(int mode, float *cur_x, float *prev_x, float *cur_z, float *prev_z, const float *c11,
                const float *c13, const float *c33, int nx, int ny, int nz, int npad, float dx, float dy, float dz,
                float dt, int nsteps, int absorbing)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2 || (mode != 0 && mode != 1)) return -1;
    float *damping = NULL;
    if (absorbing) {
        damping = (float *)malloc(sizeof(float) * ((long)nx * ny + nz));
        orc3d_init_damping(damping, nx, ny, nz, npad);
    }
    for (int s = 0; s < nsteps; ++s) {
        for (int ib = 0; ib < g.nb; ++ib) {
            if (mode == 0) {
                fwd_block_f32(&g, ib, prev_x, cur_x, prev_x, prev_z, cur_z, prev_z, c11, c13, c33);
                if (damping) abc_block_f32(&g, ib, damping, prev_x, cur_x, prev_z, cur_z);
            } else {
                fwd_block_ref(&g, ib, prev_x, cur_x, prev_x, prev_z, cur_z, prev_z, c11, c13, c33);
                if (damping) abc_block_ref(&g, ib, damping, prev_x, cur_x, prev_z, cur_z);
            }
        }
        float *t = cur_x; cur_x = prev_x; prev_x = t;
        t = cur_z; cur_z = prev_z; prev_z = t;
    }
    free(damping);
    return 0;
}

/* FP64 counterpart of orc3d_steps on double state; `gradient_form` selects
 * the source-reconstruction formula of gradientKernel (no damping). */
int orc3d_steps_f64(double *cur_x, double *prev_x, double *cur_z, double *prev_z, const double *c11,
                    const double *c13, const double *c33, int nx, int ny, int nz, int npad, float dx, float dy,
                    float dz, float dt, int nsteps, int absorbing, int gradient_form)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2) return -1;
    long n = (long)nx * ny * nz;
    float *damping = NULL;
    if (absorbing) {
        damping = (float *)malloc(sizeof(float) * ((long)nx * ny + nz));
        orc3d_init_damping(damping, nx, ny, nz, npad);
    }
    double *dummy = gradient_form ? (double *)calloc(5 * n, sizeof(double)) : NULL;
    for (int s = 0; s < nsteps; ++s) {
        for (int ib = 0; ib < g.nb; ++ib) {
            if (gradient_form)
                grad_block_f64(&g, ib, dummy, dummy + n, dummy + 2 * n, prev_x, cur_x, prev_x, prev_z, cur_z, prev_z,
                               dummy + 3 * n, dummy + 4 * n, c11, c13, c33);
            else
                fwd_block_f64(&g, ib, prev_x, cur_x, prev_x, prev_z, cur_z, prev_z, c11, c13, c33);
            if (damping && !gradient_form) abc_block_f64(&g, ib, damping, prev_x, cur_x, prev_z, cur_z);
        }
        double *t = cur_x; cur_x = prev_x; prev_x = t;
        t = cur_z; cur_z = prev_z; prev_z = t;
    }
    free(dummy);
    free(damping);
    return 0;
}

/* Backward levels on caller-owned float state with no source and no
 * receivers (the fused device backward kernels' window tests): per level and
 * z-block, mode 0 runs gradientKernel on sigma (undamped reconstruction and
 * the imaging into acc0..2 = gc11, gc13, gc33) and forwardKernel + abc on
 * lambda (objFuncGradient3d_f.cu:1008-1028); mode 1 runs forwardKernel on
 * sigma and on lambda, abc on lambda and imagingKernel into acc0
 * (rtm3d_f.cu:585-610).  arith 0 FP32, 1 the reference's arithmetic.  The
 * level pairs swap roles after each level, as in orc3d_steps. */
int orc3d_bwd_steps(int mode, int arith, float *cx, float *px, float *cz, float *pz, float *lcx, float *lpx,
                    float *lcz, float *lpz, float *acc0, float *acc1, float *acc2, const float *c11, const float *c13,
                    const float *c33, int nx, int ny, int nz, int npad, float dx, float dy, float dz, float dt,
                    int nsteps)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2 || (mode != 0 && mode != 1) || (arith != 0 && arith != 1)) return -1;
    float *damping = (float *)malloc(sizeof(float) * ((long)nx * ny + nz));
    orc3d_init_damping(damping, nx, ny, nz, npad);
    for (int s = 0; s < nsteps; ++s) {
        for (int ib = 0; ib < g.nb; ++ib) {
            if (mode == 0) {
                if (arith) grad_block_ref(&g, ib, acc0, acc1, acc2, px, cx, px, pz, cz, pz, lcx, lcz, c11, c13, c33);
                else grad_block_f32(&g, ib, acc0, acc1, acc2, px, cx, px, pz, cz, pz, lcx, lcz, c11, c13, c33);
            } else {
                if (arith) fwd_block_ref(&g, ib, px, cx, px, pz, cz, pz, c11, c13, c33);
                else fwd_block_f32(&g, ib, px, cx, px, pz, cz, pz, c11, c13, c33);
            }
            if (arith) {
                fwd_block_ref(&g, ib, lpx, lcx, lpx, lpz, lcz, lpz, c11, c13, c33);
                abc_block_ref(&g, ib, damping, lpx, lcx, lpz, lcz);
                if (mode == 1) image_block_ref(&g, ib, acc0, cx, cz, lcx, lcz);
            } else {
                fwd_block_f32(&g, ib, lpx, lcx, lpx, lpz, lcz, lpz, c11, c13, c33);
                abc_block_f32(&g, ib, damping, lpx, lcx, lpz, lcz);
                if (mode == 1) image_block_f32(&g, ib, acc0, cx, cz, lcx, lcz);
            }
        }
        float *t;
        t = cx; cx = px; px = t;
        t = cz; cz = pz; pz = t;
        t = lcx; lcx = lpx; lpx = t;
        t = lcz; lcz = lpz; lpz = t;
    }
    free(damping);
    return 0;
}

/* ---- objective variants of the 3D drivers (lib/vti-fwi-3d/gpu/agc.cpp) ---- */
#define AGC_EPS 1e-6f

/* runsum, agc.cpp:14-26: circular running sum over [i-h, i+h], summed in
 * ascending j in the element type */
#define DEF_AGC(T, SFX, SQRT, MAX)                                                                       \
    static void runsum_##SFX(int nt, int h, const T *w, T *s)                                           \
    {                                                                                                   \
        for (int i = 0; i < nt; i++) {                                                                  \
            s[i] = 0;                                                                                   \
            for (int j = i - h; j <= i + h; j++) s[i] += w[j < 0 ? j + nt : (j >= nt ? j - nt : j)];    \
        }                                                                                               \
    }                                                                                                   \
    /* agc, agc.cpp:28-35 */                                                                            \
    static void agc_##SFX(int nt, int h, const T *d, T *hd, T *sd, T *d2)                               \
    {                                                                                                   \
        for (int i = 0; i < nt; i++) d2[i] = d[i] * d[i];                                               \
        runsum_##SFX(nt, h, d2, sd);                                                                    \
        for (int i = 0; i < nt; i++) hd[i] = d[i] / SQRT(MAX((T)AGC_EPS, sd[i]));                      \
    }                                                                                                   \
    /* residualAGC (in place), agc.cpp:70-101: d becomes the adjoint source */                          \
    static double residual_agc_##SFX(int nt, int ntr, int h, T *d, const float *d0f)                    \
    {                                                                                                   \
        T *hd = (T *)malloc(sizeof(T) * nt), *sd = (T *)malloc(sizeof(T) * nt);                         \
        T *hd0 = (T *)malloc(sizeof(T) * nt), *sd0 = (T *)malloc(sizeof(T) * nt);                       \
        T *d0 = (T *)malloc(sizeof(T) * nt), *d2 = (T *)malloc(sizeof(T) * nt);                         \
        double fi = 0.;                                                                                 \
        for (int j = 0; j < ntr; j++) {                                                                 \
            T *dj = d + (long)j * nt;                                                                   \
            for (int i = 0; i < nt; i++) d0[i] = (T)d0f[(long)j * nt + i];                             \
            agc_##SFX(nt, h, dj, hd, sd, d2);                                                           \
            agc_##SFX(nt, h, d0, hd0, sd0, d2);                                                         \
            for (int i = 0; i < nt; i++) hd[i] = hd[i] - hd0[i];                                        \
            for (int i = 0; i < nt; i++) fi += (double)(hd[i] * hd[i]);                                 \
            T *rs = sd, *drss = sd0;                                                                    \
            for (int i = 0; i < nt; i++) {                                                              \
                const T t1 = MAX((T)AGC_EPS, sd[i]);                                                    \
                const T t2 = hd[i] / SQRT(t1);                                                          \
                drss[i] = dj[i] * t2 / t1;                                                              \
                rs[i] = t2;                                                                             \
            }                                                                                           \
            runsum_##SFX(nt, h, drss, hd);                                                              \
            for (int i = 0; i < nt; i++) dj[i] = rs[i] - dj[i] * hd[i];                                 \
        }                                                                                               \
        free(hd); free(sd); free(hd0); free(sd0); free(d0); free(d2);                                   \
        return fi;                                                                                      \
    }

static inline float fmax_ref(float a, float b) { return a < b ? b : a; }
static inline double dmax_ref(double a, double b) { return a < b ? b : a; }
DEF_AGC(float, f32, sqrtf, fmax_ref)
DEF_AGC(double, f64, sqrt, dmax_ref)

/* tpower, agc.cpp:103-115: data *= (ot + it*dt)^p in float */
static void tpower_f32(float *data, int nt, float ot, float dt, int nr, float p)
{
    for (int ir = 0; ir < nr; ir++)
        for (int it = 0; it < nt; it++) {
            float t = ot + it * dt;
            t = powf(t, TPOW_EXP(p));
            data[it + (long)ir * nt] *= t;
        }
}

static void tpower_f64(double *data, int nt, float ot, float dt, int nr, float p)
{
    for (int ir = 0; ir < nr; ir++)
        for (int it = 0; it < nt; it++) {
            const double t = pow((double)(ot + it * dt), (double)TPOW_EXP(p));
            data[it + (long)ir * nt] *= t;
        }
}

/* AGC objective and adjoint source of traces (trace-major, F32 or F64):
 * test hook for the objective variants. */
double orc3d_agc_residual(int mode, void *d, const float *d0, int nt, int ntr, int halfwidth)
{
    if (mode == 2) return residual_agc_f64(nt, ntr, halfwidth, (double *)d, d0);
    return residual_agc_f32(nt, ntr, halfwidth, (float *)d, d0);
}

/* objFuncGradientCij3d_f (objFuncGradient3d_f.cu:361-1231) for every shot in
 * shotid.  obs is the observed data file contents (trace-major, nnt samples
 * per receiver, receivers in recloc order), which the reference reads from
 * disk inside the function (:760).  cij's boundary shell is left holding the
 * random boundary, as the reference leaves it.
 * mode 0/1 write fgcij (float, 1+3*nxyz); mode 2 writes fgcij64 (double). */
double orc3d_gradient(int mode, float *fgcij, double *fgcij64, const float *souloc, const int *shotid, int nshot,
                      const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij,
                      const float *randboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                      float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate, int ngpu,
                      const float *obs, float tpow, int objtype, int halfwidth)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2) return -1;
    long nxy = (long)nx * ny, nxyz = nxy * nz;
    long nbnd = orc3d_nboundary(nx, ny, nz, npad);
    int sts = (int)roundf(samplingRate / dt);
    int nnt = (nt - 1) / sts + 1;
    int lmax = orc3d_lmax(nt, ngpu);
    float *damping = (float *)malloc(sizeof(float) * (nxy + nz));
    orc3d_init_damping(damping, nx, ny, nz, npad);
    float *c11 = cij, *c13 = cij + nxyz, *c33 = cij + 2 * nxyz;
    double objfunc = 0.;

    void *bufs[8];
    size_t es = (mode == 2) ? sizeof(double) : sizeof(float);
    for (int k = 0; k < 8; k++) bufs[k] = malloc(es * nxyz);
    void *gcv = calloc(3 * nxyz, es);
    double *C64[3] = {NULL, NULL, NULL};
    if (mode == 2)
        for (int k = 0; k < 3; k++) C64[k] = (double *)malloc(sizeof(double) * nxyz);

    for (int is_ = 0; is_ < nshot; is_++) {
        int is = shotid[is_];
        orc3d_shot sh;
        shot_setup(&sh, souloc, is, recloc, 4, nx, ny, ox, oy, oz, dx, dy, dz);
        int irbegin = (int)souloc[5 * is + 4];
        float *data = (float *)calloc((size_t)nnt * (sh.nr > 0 ? sh.nr : 1), sizeof(float));
        double *data64 = (mode == 2) ? (double *)calloc((size_t)nnt * (sh.nr > 0 ? sh.nr : 1), sizeof(double)) : NULL;

        /* pass 1: absorbing boundary, record */
        orc3d_boundary(1, (float *)padboundaryCij, c11, nx, ny, nz, npad);
        orc3d_boundary(1, (float *)padboundaryCij + nbnd, c13, nx, ny, nz, npad);
        orc3d_boundary(1, (float *)padboundaryCij + 2 * nbnd, c33, nx, ny, nz, npad);
        for (int k = 0; k < 8; k++) memset(bufs[k], 0, es * nxyz);
        float temp = g.dt2 * wavelet[0];
        if (mode == 2) {
            for (int k = 0; k < 3; k++)
                for (long i = 0; i < nxyz; i++) C64[k][i] = cij[k * nxyz + i];
            double *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            forward_pass_f64(&g, &sh, C64[0], C64[1], C64[2], damping, wavelet, lmax, &cx, &px, &cz, &pz, NULL,
                             data64, sts, nnt);
        } else {
            float *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            if (mode == 0)
                forward_pass_f32(&g, &sh, c11, c13, c33, damping, wavelet, lmax, &cx, &px, &cz, &pz, data, NULL,
                                 sts, nnt);
            else
                forward_pass_ref(&g, &sh, c11, c13, c33, damping, wavelet, lmax, &cx, &px, &cz, &pz, data, NULL,
                                 sts, nnt);
        }
        /* mute dead traces, t^p gain, residual, objective (:736-779) */
        for (int ir = 0; ir < sh.nr; ir++) {
            int ir1 = ir + irbegin;
            if (recloc[4 * ir1 + 3] == 0.f) {
                memset(data + (long)ir * nnt, 0, sizeof(float) * nnt);
                if (data64) memset(data64 + (long)ir * nnt, 0, sizeof(double) * nnt);
            }
        }
        if (tpow != 0.f) {
            if (mode == 2) tpower_f64(data64, nnt, ot, samplingRate, sh.nr, tpow);
            else tpower_f32(data, nnt, ot, samplingRate, sh.nr, tpow);
        }
        const float *ob = obs + (long)nnt * irbegin;
        if (objtype != 0) {
            objfunc += mode == 2 ? residual_agc_f64(nnt, sh.nr, halfwidth, data64, ob)
                                 : residual_agc_f32(nnt, sh.nr, halfwidth, data, ob);
        } else {
            for (long i = 0; i < (long)nnt * sh.nr; i++) {
                if (mode == 2) {
                    data64[i] = RESID(data64[i], (double)ob[i]);
                    objfunc += data64[i] * data64[i];
                } else {
                    data[i] = RESID(data[i], ob[i]);
                    objfunc += (double)(data[i] * data[i]);
                }
            }
        }

        /* pass 2: random boundary, no damping, no recording */
        orc3d_boundary(1, (float *)randboundaryCij, c11, nx, ny, nz, npad);
        orc3d_boundary(1, (float *)randboundaryCij + nbnd, c13, nx, ny, nz, npad);
        orc3d_boundary(1, (float *)randboundaryCij + 2 * nbnd, c33, nx, ny, nz, npad);
        for (int k = 0; k < 8; k++) memset(bufs[k], 0, es * nxyz);
        if (mode == 2) {
            for (int k = 0; k < 3; k++)
                for (long i = 0; i < nxyz; i++) C64[k][i] = cij[k * nxyz + i];
            double *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            double *lcx = bufs[4], *lpx = bufs[5], *lcz = bufs[6], *lpz = bufs[7];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            forward_pass_f64(&g, &sh, C64[0], C64[1], C64[2], NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL,
                             sts, nnt);
            /* flip (:911-914) */
            double *t = cx; cx = px; px = t;
            t = cz; cz = pz; pz = t;
            for (int ir = 0; ir < sh.nr; ir++) {
                double tt = (double)g.dt2 * data64[(nnt - 1) + (long)ir * nnt];
                lcx[sh.rec_true[ir]] = LAM0_WX * tt;
                lcz[sh.rec_true[ir]] = LAM0_WZ * tt;
            }
            double *gc = (double *)gcv;
            backward_grad_pass_f64(&g, &sh, C64[0], C64[1], C64[2], damping, wavelet, lmax, nt, &cx, &px, &cz, &pz,
                                   &lcx, &lpx, &lcz, &lpz, gc, gc + nxyz, gc + 2 * nxyz, data64, sts, nnt);
        } else {
            float *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            float *lcx = bufs[4], *lpx = bufs[5], *lcz = bufs[6], *lpz = bufs[7];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            if (mode == 0)
                forward_pass_f32(&g, &sh, c11, c13, c33, NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL, sts,
                                 nnt);
            else
                forward_pass_ref(&g, &sh, c11, c13, c33, NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL, sts,
                                 nnt);
            float *t = cx; cx = px; px = t;
            t = cz; cz = pz; pz = t;
            for (int ir = 0; ir < sh.nr; ir++) {
                float tt = g.dt2 * data[(nnt - 1) + (long)ir * nnt];
                lcx[sh.rec_true[ir]] = (float)(LAM0_WX * tt);
                lcz[sh.rec_true[ir]] = (float)(LAM0_WZ * tt);
            }
            float *gc = (float *)gcv;
            if (mode == 0)
                backward_grad_pass_f32(&g, &sh, c11, c13, c33, damping, wavelet, lmax, nt, &cx, &px, &cz, &pz, &lcx,
                                       &lpx, &lcz, &lpz, gc, gc + nxyz, gc + 2 * nxyz, data, sts, nnt);
            else
                backward_grad_pass_ref(&g, &sh, c11, c13, c33, damping, wavelet, lmax, nt, &cx, &px, &cz, &pz, &lcx,
                                       &lpx, &lcz, &lpz, gc, gc + nxyz, gc + 2 * nxyz, data, sts, nnt);
        }
        free(data);
        free(data64);
        shot_free(&sh);
    }
    if (mode == 2) {
        fgcij64[0] = objfunc / 2;
        memcpy(fgcij64 + 1, gcv, sizeof(double) * 3 * nxyz);
    } else {
        fgcij[0] = (float)(objfunc / 2);
        memcpy(fgcij + 1, gcv, sizeof(float) * 3 * nxyz);
    }
    for (int k = 0; k < 8; k++) free(bufs[k]);
    for (int k = 0; k < 3; k++) free(C64[k]);
    free(gcv);
    free(damping);
    return objfunc / 2;
}

/* rtmCij3d_f (rtm3d_f.cu:161-830): random-boundary forward, then backward
 * with source reconstruction, receiver-side propagation of the observed data
 * (recloc stride 3) and the imaging condition.  image: float (mode 0/1) or
 * double (mode 2), zeroed here as the reference does (:178). */
int orc3d_rtm(int mode, void *image, const float *souloc, const int *shotid, int nshot, const float *recloc,
              const float *wavelet, const float *c11, const float *c13, const float *c33, int nx, int ny, int nz,
              int nt, int npad, float ox, float oy, float oz, float dx, float dy, float dz, float dt,
              float samplingRate, int ngpu, const float *obs)
{
    orc3d_geom g;
    geom_setup(&g, nx, ny, nz, npad, dx, dy, dz, dt);
    if (g.nb < 2) return -1;
    long nxy = (long)nx * ny, nxyz = nxy * nz;
    int sts = (int)roundf(samplingRate / dt);
    int nnt = (nt - 1) / sts + 1;
    int lmax = orc3d_lmax(nt, ngpu);
    float *damping = (float *)malloc(sizeof(float) * (nxy + nz));
    orc3d_init_damping(damping, nx, ny, nz, npad);
    size_t es = (mode == 2) ? sizeof(double) : sizeof(float);
    memset(image, 0, es * nxyz);
    void *bufs[8];
    for (int k = 0; k < 8; k++) bufs[k] = malloc(es * nxyz);
    double *C64[3] = {NULL, NULL, NULL};
    if (mode == 2) {
        const float *src[3] = {c11, c13, c33};
        for (int k = 0; k < 3; k++) {
            C64[k] = (double *)malloc(sizeof(double) * nxyz);
            for (long i = 0; i < nxyz; i++) C64[k][i] = src[k][i];
        }
    }
    for (int is_ = 0; is_ < nshot; is_++) {
        int is = shotid[is_];
        orc3d_shot sh;
        shot_setup(&sh, souloc, is, recloc, 3, nx, ny, ox, oy, oz, dx, dy, dz);
        int irbegin = (int)souloc[5 * is + 4];
        const float *ob = obs + (long)nnt * irbegin;
        for (int k = 0; k < 8; k++) memset(bufs[k], 0, es * nxyz);
        float temp = g.dt2 * wavelet[0];
        if (mode == 2) {
            double *ob64 = (double *)malloc(sizeof(double) * (size_t)nnt * (sh.nr > 0 ? sh.nr : 1));
            for (long i = 0; i < (long)nnt * sh.nr; i++) ob64[i] = ob[i];
            double *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            double *lcx = bufs[4], *lpx = bufs[5], *lcz = bufs[6], *lpz = bufs[7];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            forward_pass_f64(&g, &sh, C64[0], C64[1], C64[2], NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL,
                             sts, nnt);
            double *t = cx; cx = px; px = t;
            t = cz; cz = pz; pz = t;
            for (int ir = 0; ir < sh.nr; ir++) {
                double tt = (double)g.dt2 * ob64[(nnt - 1) + (long)ir * nnt];
                lcx[sh.rec_true[ir]] = LAM0_WX * tt;
                lcz[sh.rec_true[ir]] = LAM0_WZ * tt;
            }
            backward_rtm_pass_f64(&g, &sh, C64[0], C64[1], C64[2], damping, wavelet, lmax, nt, &cx, &px, &cz, &pz,
                                  &lcx, &lpx, &lcz, &lpz, (double *)image, ob64, sts, nnt);
            free(ob64);
        } else {
            float *cx = bufs[0], *px = bufs[1], *cz = bufs[2], *pz = bufs[3];
            float *lcx = bufs[4], *lpx = bufs[5], *lcz = bufs[6], *lpz = bufs[7];
            cx[sh.sou_index] = temp;
            cz[sh.sou_index] = temp;
            if (mode == 0)
                forward_pass_f32(&g, &sh, c11, c13, c33, NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL, sts,
                                 nnt);
            else
                forward_pass_ref(&g, &sh, c11, c13, c33, NULL, wavelet, lmax, &cx, &px, &cz, &pz, NULL, NULL, sts,
                                 nnt);
            float *t = cx; cx = px; px = t;
            t = cz; cz = pz; pz = t;
            for (int ir = 0; ir < sh.nr; ir++) {
                float tt = g.dt2 * ob[(nnt - 1) + (long)ir * nnt];
                lcx[sh.rec_true[ir]] = (float)(LAM0_WX * tt);
                lcz[sh.rec_true[ir]] = (float)(LAM0_WZ * tt);
            }
            if (mode == 0)
                backward_rtm_pass_f32(&g, &sh, c11, c13, c33, damping, wavelet, lmax, nt, &cx, &px, &cz, &pz, &lcx,
                                      &lpx, &lcz, &lpz, (float *)image, ob, sts, nnt);
            else
                backward_rtm_pass_ref(&g, &sh, c11, c13, c33, damping, wavelet, lmax, nt, &cx, &px, &cz, &pz, &lcx,
                                      &lpx, &lcz, &lpz, (float *)image, ob, sts, nnt);
        }
        shot_free(&sh);
    }
    for (int k = 0; k < 8; k++) free(bufs[k]);
    for (int k = 0; k < 3; k++) free(C64[k]);
    free(damping);
    return 0;
}
