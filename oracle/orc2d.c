/*
 * Oracle (test infrastructure only) -- CPU restatement of the reference's 2D
 * VTI path in plain C:
 *   lib/vti-fwi-2d/kernels.cu     D, forwardC, forwardCD, backwardDC, abc,
 *                                 scattering, gradientCrossCor, recordData,
 *                                 injectData, injectSource
 *   lib/vti-fwi-2d/synthetic_f.cu:13-152        synthetic_f
 *   lib/vti-fwi-2d/modeling_f.cu:10-86          modeling_f
 *   lib/vti-fwi-2d/objFuncGradient_f.cu:15-371  objFuncGradient_f
 *   lib/vti-fwi-2d/hessian_f.cu:575-1010        GNhessian_f
 *   lib/vti-fwi-2d/hessian_f.cu:15-573          hessian_f
 *   lib/vti-fwi-2d/rtm_f.cu:13-269              rtm_f
 *   lib/vti-fwi-2d/odcig_f.cu:12-247            odcig_f
 * mode 0: the reference's FP32 arithmetic; mode 2: FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"

#define T float
#define SFX f32
#define CK0 C0F
#define CK1 C1F
#define CK2 C2F
#define CK3 C3F
#define CK4 C4F
#include "orc2d_body.h"
#undef T
#undef SFX
#undef CK0
#undef CK1
#undef CK2
#undef CK3
#undef CK4

#define T double
#define SFX f64
#define CK0 C0D
#define CK1 C1D
#define CK2 C2D
#define CK3 C3D
#define CK4 C4D
#include "orc2d_body.h"
#undef T
#undef SFX
#undef CK0
#undef CK1
#undef CK2
#undef CK3
#undef CK4

static double *widen(const float *a, long n)
{
    if (!a) return NULL;
    double *d = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (long i = 0; i < n; ++i) d[i] = a[i];
    return d;
}

/* synthetic_f; data: float (mode 0) or double (mode 2), nr * nnt_data samples,
 * time-major with all shots' receivers side by side (synthetic_f.cu:110). */
int orc2d_synthetic(int mode, void *data, const float *c11, const float *c13, const float *c33, const float *wavelet,
                    const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                    int npad, float dx, float dz, float dt, float rate, float ot)
{
    const long n = (long)(nx + 2 * npad) * (nz + 2 * npad);
    if (mode == 0) {
        synthetic_f32((float *)data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt,
                      rate, ot);
    } else if (mode == 2) {
        double *a = widen(c11, n), *b = widen(c13, n), *c = widen(c33, n);
        synthetic_f64((double *)data, a, b, c, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt,
                      rate, ot);
        free(a);
        free(b);
        free(c);
    } else {
        return -1;
    }
    return 0;
}

/* modeling_f; wavefield: float (mode 0) or double (mode 2), nnxz samples per
 * slot, (nt - ntNeg - 1) / ratio + 1 slots. */
int orc2d_modeling(int mode, void *wavefield, const float *c11, const float *c13, const float *c33,
                   const float *wavelet, int slocxz, const float *taper, int nx, int nz, int nt, int npad, float dx,
                   float dz, float dt, float rate, float ot)
{
    const long n = (long)(nx + 2 * npad) * (nz + 2 * npad);
    if (mode == 0) {
        modeling_f32((float *)wavefield, c11, c13, c33, wavelet, slocxz, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
    } else if (mode == 2) {
        double *a = widen(c11, n), *b = widen(c13, n), *c = widen(c33, n);
        modeling_f64((double *)wavefield, a, b, c, wavelet, slocxz, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
        free(a);
        free(b);
        free(c);
    } else {
        return -1;
    }
    return 0;
}

/* objFuncGradient_f (born = 0) / GNhessian_f (born = 1) / gradient path with
 * the forwardCD adjoint (born = 2, a check only), as run on ngpu GPUs.  gc: 3 * nnxz of
 * float (mode 0) or double (mode 2): gc11, gc13, gc33.  Returns the objective
 * 0.5*sum r^2 (0 for the Hessian). */
double orc2d_fg(int mode, int born, int ngpu, void *gc, const float *data, const float *c11, const float *c13, const float *c33,
                const float *dc11, const float *dc13, const float *dc33, const float *wavelet, const int *sloc, int ns,
                const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad, float dx, float dz,
                float dt, float rate, float ot)
{
    const long n = (long)(nx + 2 * npad) * (nz + 2 * npad);
    double obj = 0;
    if (mode == 0) {
        float *g = (float *)gc;
        obj = fg_or_gn_f32(born, ngpu, g, g + n, g + 2 * n, data, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, ns, rloc,
                           nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
        /* objFuncGradient_f returns 0.5f*objfunc as float (objFuncGradient_f.cu:370) */
        return (double)(float)(0.5f * obj);
    }
    double *a = widen(c11, n), *b = widen(c13, n), *c = widen(c33, n);
    double *da = widen(dc11, n), *db = widen(dc13, n), *dc = widen(dc33, n);
    double *g = (double *)gc;
    obj = fg_or_gn_f64(born, ngpu, g, g + n, g + 2 * n, data, a, b, c, da, db, dc, wavelet, sloc, ns, rloc, nr, taper, nx, nz,
                       nt, npad, dx, dz, dt, rate, ot);
    free(a);
    free(b);
    free(c);
    free(da);
    free(db);
    free(dc);
    return 0.5 * obj;
}

/* One forwardCD step (float) and one backwardDC step on caller state, for the
 * adjoint (dot-product) test of the pair. */
void orc2d_forwardCD_f32(float *p0, float *q0, const float *p1, const float *q1, const float *c11, const float *c13,
                         const float *c33, float dx2, float dz2, float dt2, int nnx, int nnz)
{
    forwardCD_f32(p0, q0, p1, q1, c11, c13, c33, dx2, dz2, dt2, nnx, nnz);
}

void orc2d_forwardCD_f64(double *p0, double *q0, const double *p1, const double *q1, const double *c11,
                         const double *c13, const double *c33, double dx2, double dz2, double dt2, int nnx, int nnz)
{
    forwardCD_f64(p0, q0, p1, q1, c11, c13, c33, dx2, dz2, dt2, nnx, nnz);
}

void orc2d_backwardDC_f64(double *p1, double *q1, const double *p0, const double *q0, const double *c11,
                          const double *c13, const double *c33, double dx2, double dz2, double dt2, int nnx, int nnz)
{
    const long n = (long)nnx * nnz;
    double *sp = (double *)calloc(n, sizeof(double)), *sq = (double *)calloc(n, sizeof(double));
    backwardDC_f64(p1, q1, p0, q0, c11, c13, c33, dx2, dz2, dt2, nnx, nnz, sp, sq);
    free(sp);
    free(sq);
}

/* rtm_f (odcig = 0) / odcig_f (odcig = 1) as run on ngpu GPUs.  image:
 * nnxz (rtm) or (2*nhx+1)*nnxz (odcig) of float (mode 0) or double (mode 2). */
int orc2d_rtm(int mode, int odcig, int ngpu, void *image, const float *data, const float *c11, const float *c13,
              const float *c33, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
              const float *taper, int nx, int nz, int nhx, int nt, int npad, float dx, float dz, float dt, float rate,
              float ot)
{
    const long n = (long)(nx + 2 * npad) * (nz + 2 * npad);
    if (mode == 0) {
        rtm_ngpu_f32(odcig, ngpu, (float *)image, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nhx,
                     nt, npad, dx, dz, dt, rate, ot);
        return 0;
    }
    if (mode != 2) return -1;
    double *a = widen(c11, n), *b = widen(c13, n), *c = widen(c33, n);
    rtm_ngpu_f64(odcig, ngpu, (double *)image, data, a, b, c, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nhx, nt, npad,
                 dx, dz, dt, rate, ot);
    free(a);
    free(b);
    free(c);
    return 0;
}

/* hessian_f as run on ngpu GPUs; gc: 3 * nnxz of float (mode 0) or double
 * (mode 2); seed_bg = 1 as the reference, 0 with its Born snapshot buffer
 * seeded from the Born snapshots (a check) */
int orc2d_hessian(int mode, int seed_bg, int ngpu, void *gc, const float *data, const float *c11, const float *c13,
                  const float *c33, const float *dc11, const float *dc13, const float *dc33, const float *wavelet,
                  const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                  int npad, float dx, float dz, float dt, float rate, float ot)
{
    const long n = (long)(nx + 2 * npad) * (nz + 2 * npad);
    if (mode == 0) {
        float *g = (float *)gc;
        hess_ngpu_f32(seed_bg, ngpu, g, g + n, g + 2 * n, data, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, ns, rloc, nr,
                      taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
        return 0;
    }
    if (mode != 2) return -1;
    double *a = widen(c11, n), *b = widen(c13, n), *c = widen(c33, n);
    double *da = widen(dc11, n), *db = widen(dc13, n), *dc = widen(dc33, n);
    double *g = (double *)gc;
    hess_ngpu_f64(seed_bg, ngpu, g, g + n, g + 2 * n, data, a, b, c, da, db, dc, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt,
                  npad, dx, dz, dt, rate, ot);
    free(a);
    free(b);
    free(c);
    free(da);
    free(db);
    free(dc);
    return 0;
}
