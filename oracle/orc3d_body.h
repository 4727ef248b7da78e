/*
 * Body of the 3D oracle, instantiated three times by orc3d.c:
 *   T = float,  ARITH_F32 : FP32 arithmetic in the operation order the HIP kernels use
 *   T = float,  ARITH_REF : the reference's mixed arithmetic (double constants, float storage)
 *   T = double, ARITH_F64 : FP64 storage and arithmetic
 *
 * Test infrastructure only: nothing in the product links or calls this.
 *
 * Layout follows the reference: j = ix + iy*nx + iz*nx*ny (x fastest).
 * The reference streams the volume through the GPU in z-blocks of 4 planes
 * and, for each time level, runs the blocks in ascending z order with the
 * damping of each block applied right after that block's step
 * (lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:638-667).  This restatement runs
 * the same block order on the whole volume, so every block sees its
 * neighbours in exactly the state the reference's pipeline gives it.
 */

#define S_(a, b) a##_##b
#define S(a, b) S_(a, b)
#define FN(name) S(name, SFX)

/* ---- one z-block of forwardKernel (kernel3d.cu:3-220) ------------------- */
/* next may alias prev: each point reads its own prev before writing next. */
static void FN(fwd_block)(const orc3d_geom *g, int ib, T *next_x, const T *cur_x, const T *prev_x,
                          T *next_z, const T *cur_z, const T *prev_z, const T *c11, const T *c13,
                          const T *c33)
{
    const long nx = g->nx, ny = g->ny, nxy = nx * ny;
    const int zmax = 4 * g->nb;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 4 * ib; iz < 4 * ib + 4; ++iz)
        for (long iy = 4; iy < ny - 4; ++iy) {
            for (long ix = 4; ix < nx - 4; ++ix) {
                const long j = ix + iy * nx + iz * nxy;
                T zq[9];
                for (int k = -4; k <= 4; ++k) {
                    int zz = iz + k;
                    zq[k + 4] = (zz < 0 || zz >= zmax) ? (T)0 : cur_z[j + (long)k * nxy];
                }
                const T *s = cur_x + j;
#if defined(ARITH_REF)
                /* float pair sums, double products/accumulation, rounded to float */
                float tx = (float)((C0D * s[0] + C1D * (float)(s[-1] + s[1]) + C2D * (float)(s[-2] + s[2]) +
                                    C3D * (float)(s[-3] + s[3]) + C4D * (float)(s[-4] + s[4])) *
                                   (double)g->dt2dx2);
                float ty = (float)((C0D * s[0] + C1D * (float)(s[-nx] + s[nx]) +
                                    C2D * (float)(s[-2 * nx] + s[2 * nx]) + C3D * (float)(s[-3 * nx] + s[3 * nx]) +
                                    C4D * (float)(s[-4 * nx] + s[4 * nx])) *
                                   (double)g->dt2dy2);
                ty += tx;
                float tz = (float)((C0D * zq[4] + C1D * (float)(zq[3] + zq[5]) + C2D * (float)(zq[2] + zq[6]) +
                                    C3D * (float)(zq[1] + zq[7]) + C4D * (float)(zq[0] + zq[8])) *
                                   (double)g->dt2dz2);
                float ax = c11[j] * ty + c13[j] * tz;
                float az = c13[j] * ty + c33[j] * tz;
                next_x[j] = (float)((double)ax + 2. * (double)s[0] - (double)prev_x[j]);
                next_z[j] = (float)((double)az + 2. * (double)zq[4] - (double)prev_z[j]);
#else
                T tx = (CK0 * s[0] + CK1 * (s[-1] + s[1]) + CK2 * (s[-2] + s[2]) + CK3 * (s[-3] + s[3]) +
                        CK4 * (s[-4] + s[4])) *
                       (T)g->dt2dx2;
                T ty = (CK0 * s[0] + CK1 * (s[-nx] + s[nx]) + CK2 * (s[-2 * nx] + s[2 * nx]) +
                        CK3 * (s[-3 * nx] + s[3 * nx]) + CK4 * (s[-4 * nx] + s[4 * nx])) *
                       (T)g->dt2dy2;
                ty = ty + tx;
                T tz = (CK0 * zq[4] + CK1 * (zq[3] + zq[5]) + CK2 * (zq[2] + zq[6]) + CK3 * (zq[1] + zq[7]) +
                        CK4 * (zq[0] + zq[8])) *
                       (T)g->dt2dz2;
                T ax = c11[j] * ty + c13[j] * tz;
                T az = c13[j] * ty + c33[j] * tz;
                next_x[j] = (ax + (T)2 * s[0]) - prev_x[j];
                next_z[j] = (az + (T)2 * zq[4]) - prev_z[j];
#endif
            }
        }
}

/* ---- one z-block of gradientKernel (kernel3d.cu:222-454) ---------------- */
/* Source-field step with /d2 stencils and dt2 outside, plus the imaging
 * gc11 += lx*ty, gc33 += lz*tz, gc13 += lx*tz + lz*ty at the current level. */
static void FN(grad_block)(const orc3d_geom *g, int ib, T *gc11, T *gc13, T *gc33, T *next_x, const T *cur_x,
                           const T *prev_x, T *next_z, const T *cur_z, const T *prev_z, const T *lam_x,
                           const T *lam_z, const T *c11, const T *c13, const T *c33)
{
    const long nx = g->nx, ny = g->ny, nxy = nx * ny;
    const int zmax = 4 * g->nb;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 4 * ib; iz < 4 * ib + 4; ++iz)
        for (long iy = 4; iy < ny - 4; ++iy) {
            for (long ix = 4; ix < nx - 4; ++ix) {
                const long j = ix + iy * nx + iz * nxy;
                T zq[9];
                for (int k = -4; k <= 4; ++k) {
                    int zz = iz + k;
                    zq[k + 4] = (zz < 0 || zz >= zmax) ? (T)0 : cur_z[j + (long)k * nxy];
                }
                const T *s = cur_x + j;
#if defined(ARITH_REF)
                float tx = (float)((C0D * s[0] + C1D * (float)(s[-1] + s[1]) + C2D * (float)(s[-2] + s[2]) +
                                    C3D * (float)(s[-3] + s[3]) + C4D * (float)(s[-4] + s[4])) /
                                   (double)g->dx2);
                float ty = (float)((C0D * s[0] + C1D * (float)(s[-nx] + s[nx]) +
                                    C2D * (float)(s[-2 * nx] + s[2 * nx]) + C3D * (float)(s[-3 * nx] + s[3 * nx]) +
                                    C4D * (float)(s[-4 * nx] + s[4 * nx])) /
                                   (double)g->dy2);
                ty += tx;
                float tz = (float)((C0D * zq[4] + C1D * (float)(zq[3] + zq[5]) + C2D * (float)(zq[2] + zq[6]) +
                                    C3D * (float)(zq[1] + zq[7]) + C4D * (float)(zq[0] + zq[8])) /
                                   (double)g->dz2);
                float ax = c11[j] * ty + c13[j] * tz;
                float az = c13[j] * ty + c33[j] * tz;
                next_x[j] = (float)((double)(g->dt2 * ax) + 2. * (double)s[0] - (double)prev_x[j]);
                next_z[j] = (float)((double)(g->dt2 * az) + 2. * (double)zq[4] - (double)prev_z[j]);
#else
                T tx = (CK0 * s[0] + CK1 * (s[-1] + s[1]) + CK2 * (s[-2] + s[2]) + CK3 * (s[-3] + s[3]) +
                        CK4 * (s[-4] + s[4])) /
                       (T)g->dx2;
                T ty = (CK0 * s[0] + CK1 * (s[-nx] + s[nx]) + CK2 * (s[-2 * nx] + s[2 * nx]) +
                        CK3 * (s[-3 * nx] + s[3 * nx]) + CK4 * (s[-4 * nx] + s[4 * nx])) /
                       (T)g->dy2;
                ty = ty + tx;
                T tz = (CK0 * zq[4] + CK1 * (zq[3] + zq[5]) + CK2 * (zq[2] + zq[6]) + CK3 * (zq[1] + zq[7]) +
                        CK4 * (zq[0] + zq[8])) /
                       (T)g->dz2;
                T ax = c11[j] * ty + c13[j] * tz;
                T az = c13[j] * ty + c33[j] * tz;
                next_x[j] = ((T)g->dt2 * ax + (T)2 * s[0]) - prev_x[j];
                next_z[j] = ((T)g->dt2 * az + (T)2 * zq[4]) - prev_z[j];
#endif
                const T lx = lam_x[j], lz = lam_z[j];
                gc11[j] = gc11[j] + lx * ty;
                gc33[j] = gc33[j] + GC33_TERM(lz, ty, tz);
                gc13[j] = gc13[j] + GC13_TERM(lx, lz, ty, tz);
            }
        }
}

/* ---- damping of one block, abc3d.cu:82-125 ------------------------------- */
static void FN(abc_block)(const orc3d_geom *g, int ib, const float *damping, T *next_x, T *cur_x, T *next_z,
                          T *cur_z)
{
    const long nx = g->nx, ny = g->ny, nxy = nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 4 * ib; iz < 4 * ib + 4; ++iz)
        for (long iy = 4; iy < ny - 4; ++iy)
            for (long ix = 4; ix < nx - 4; ++ix) {
                const long j = ix + iy * nx + iz * nxy;
                /* abcXY and abcXYZ agree because damping[nxy+iz] is exactly 1 in the z interior */
                const float d = damping[ix + iy * nx] * damping[nxy + iz];
                next_x[j] *= d;
                cur_x[j] *= d;
                next_z[j] *= d;
                cur_z[j] *= d;
            }
}

/* ---- RTM imaging of one block, kernel3d.cu:456-468 ----------------------- */
static void FN(image_block)(const orc3d_geom *g, int ib, T *image, const T *sx, const T *sz, const T *lx,
                            const T *lz)
{
    const long nx = g->nx, ny = g->ny, nxy = nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 4 * ib; iz < 4 * ib + 4; ++iz)
        for (long iy = 4; iy < ny - 4; ++iy)
            for (long ix = 4; ix < nx - 4; ++ix) {
                const long j = ix + iy * nx + iz * nxy;
#if defined(ARITH_REF)
                image[j] = (float)((double)image[j] + (TWOTHIRD_D * sx[j] + ONETHIRD_D * sz[j]) *
                                                          (IMG_LWX * lx[j] + IMG_LWZ * lz[j]));
#else
                image[j] = image[j] + (KTWOTHIRD * sx[j] + KONETHIRD * sz[j]) *
                                          ((T)IMG_LWX * lx[j] + (T)IMG_LWZ * lz[j]);
#endif
            }
}

/* ---- point operations (injectRecord3d.cu) --------------------------------
 * These are tiny and use the reference's own mixed arithmetic in every mode
 * except F64 (which keeps double throughout). */
static void FN(record)(const orc3d_shot *sh, int itdata, int nnt, float *data, const T *sx, const T *sz)
{
    for (int ir = 0; ir < sh->nr; ++ir) {
        long i = sh->rec_eff[ir];
#if defined(ARITH_F64)
        data[itdata + (long)ir * nnt] = (float)(REC_WX * sx[i] + REC_WZ * sz[i]);
#else
        data[itdata + (long)ir * nnt] = (float)(REC_WX * (double)sx[i] + REC_WZ * (double)sz[i]);
#endif
    }
}

static void FN(record64)(const orc3d_shot *sh, int itdata, int nnt, double *data, const T *sx, const T *sz)
{
    for (int ir = 0; ir < sh->nr; ++ir) {
        long i = sh->rec_eff[ir];
        data[itdata + (long)ir * nnt] = REC_WX * (double)sx[i] + REC_WZ * (double)sz[i];
    }
}

/* injectResidual: s = dt2*res; x += (2/3)s; z += (1/3)s (double products) */
static void FN(inject_res)(const orc3d_shot *sh, const double *res, float dt2, T *lx, T *lz)
{
    for (int ir = 0; ir < sh->nr; ++ir) {
        long i = sh->rec_eff[ir];
#if defined(ARITH_F64)
        double s = (double)dt2 * res[ir];
        lx[i] += INJ_WX * s;
        lz[i] += INJ_WZ * s;
#else
        float s = dt2 * (float)res[ir];
        lx[i] = (float)((double)lx[i] + INJ_WX * (double)s);
        lz[i] = (float)((double)lz[i] + INJ_WZ * (double)s);
#endif
    }
}

/* interpolateResidual (injectRecord3d.cu:60-78): the float fine-time
 * residual for time index ti from the trace-major coarse array. */
static void FN(interp_res)(const orc3d_shot *sh, const void *coarse_v, int ti, int nnt, int sts, double *fine)
{
#if defined(ARITH_F64)
    const double *coarse = (const double *)coarse_v;
    double f = (double)ti / (double)sts;
    int i = (int)f;
    if (i >= nnt - 1) {
        for (int ir = 0; ir < sh->nr; ++ir) fine[ir] = coarse[(nnt - 1) + (long)ir * nnt];
    } else {
        f = f - i;
        for (int ir = 0; ir < sh->nr; ++ir)
            fine[ir] = (1. - f) * coarse[i + (long)ir * nnt] + f * coarse[(i + 1) + (long)ir * nnt];
    }
#else
    const float *coarse = (const float *)coarse_v;
    float f = (float)ti / (float)sts;
    int i = (int)f;
    if (i >= nnt - 1) {
        for (int ir = 0; ir < sh->nr; ++ir) fine[ir] = (double)coarse[(nnt - 1) + (long)ir * nnt];
    } else {
        f = f - i;
        for (int ir = 0; ir < sh->nr; ++ir)
            fine[ir] = (double)(float)((1. - f) * coarse[i + (long)ir * nnt] + f * coarse[(i + 1) + (long)ir * nnt]);
    }
#endif
}

/* ---- forward pass: levels 2..lmax (objFuncGradient3d_f.cu:594-735) -------
 * On entry cur = level 1, prev = level 0; on exit cur = level lmax,
 * prev = level lmax-1.  damping==NULL gives the random-boundary pass.
 * data (trace-major, nnt per receiver) receives samples it/sts when
 * it % sts == 0, from the damped field, as the reference records them. */
static void FN(forward_pass)(const orc3d_geom *g, const orc3d_shot *sh, const T *c11, const T *c13,
                             const T *c33, const float *damping, const float *wavelet, int lmax, T **cx,
                             T **px, T **cz, T **pz, float *data, double *data64, int sts, int nnt)
{
    for (int it = 2; it <= lmax; ++it) {
        for (int ib = 0; ib < g->nb; ++ib) {
            FN(fwd_block)(g, ib, *px, *cx, *px, *pz, *cz, *pz, c11, c13, c33);
            if (ib == sh->sou_block) {
                float source = g->dt2 * wavelet[it - 1];
                (*px)[sh->sou_index] += source;
                (*pz)[sh->sou_index] += source;
            }
            if (damping) FN(abc_block)(g, ib, damping, *px, *cx, *pz, *cz);
        }
        if ((data || data64) && it % sts == 0 && it / sts < nnt) {
            if (data) FN(record)(sh, it / sts, nnt, data, *px, *pz);
            if (data64) FN(record64)(sh, it / sts, nnt, data64, *px, *pz);
        }
        T *t = *cx; *cx = *px; *px = t;
        t = *cz; *cz = *pz; *pz = t;
    }
}

/* ---- gradient backward pass (objFuncGradient3d_f.cu:929-1078) ----------- */
static void FN(backward_grad_pass)(const orc3d_geom *g, const orc3d_shot *sh, const T *c11, const T *c13,
                                   const T *c33, const float *damping, const float *wavelet, int lmax, int nt,
                                   T **cx, T **px, T **cz, T **pz, T **lcx, T **lpx, T **lcz, T **lpz, T *gc11,
                                   T *gc13, T *gc33, const void *res, int sts, int nnt)
{
    double *fine = (double *)malloc(sizeof(double) * (sh->nr > 0 ? sh->nr : 1));
    for (int s = 2; s <= lmax; ++s) {
        int it = nt - 1 - s;
        FN(interp_res)(sh, res, RES_TI(it), nnt, sts, fine);
        for (int ib = 0; ib < g->nb; ++ib) {
            FN(grad_block)(g, ib, gc11, gc13, gc33, *px, *cx, *px, *pz, *cz, *pz, *lcx, *lcz, c11, c13, c33);
            FN(fwd_block)(g, ib, *lpx, *lcx, *lpx, *lpz, *lcz, *lpz, c11, c13, c33);
            if (ib == sh->sou_block) {
                float source = g->dt2 * wavelet[it + 1];
                (*px)[sh->sou_index] += source;
                (*pz)[sh->sou_index] += source;
            }
            if (ib == sh->rec_block) FN(inject_res)(sh, fine, g->dt2, *lpx, *lpz);
            FN(abc_block)(g, ib, damping, *lpx, *lcx, *lpz, *lcz);
        }
        T *t;
        t = *cx; *cx = *px; *px = t;
        t = *cz; *cz = *pz; *pz = t;
        t = *lcx; *lcx = *lpx; *lpx = t;
        t = *lcz; *lcz = *lpz; *lpz = t;
    }
    free(fine);
}

/* ---- RTM backward pass (rtm3d_f.cu:514-640) ------------------------------ */
static void FN(backward_rtm_pass)(const orc3d_geom *g, const orc3d_shot *sh, const T *c11, const T *c13,
                                  const T *c33, const float *damping, const float *wavelet, int lmax, int nt,
                                  T **cx, T **px, T **cz, T **pz, T **lcx, T **lpx, T **lcz, T **lpz, T *image,
                                  const void *obs, int sts, int nnt)
{
    double *fine = (double *)malloc(sizeof(double) * (sh->nr > 0 ? sh->nr : 1));
    for (int s = 2; s <= lmax; ++s) {
        int it = nt - 1 - s;
        FN(interp_res)(sh, obs, RES_TI(it), nnt, sts, fine);
        for (int ib = 0; ib < g->nb; ++ib) {
            FN(fwd_block)(g, ib, *px, *cx, *px, *pz, *cz, *pz, c11, c13, c33);
            FN(fwd_block)(g, ib, *lpx, *lcx, *lpx, *lpz, *lcz, *lpz, c11, c13, c33);
            if (ib == sh->sou_block) {
                float source = g->dt2 * wavelet[it + 1];
                (*px)[sh->sou_index] += source;
                (*pz)[sh->sou_index] += source;
            }
            if (ib == sh->rec_block) FN(inject_res)(sh, fine, g->dt2, *lpx, *lpz);
            FN(abc_block)(g, ib, damping, *lpx, *lcx, *lpz, *lcz);
            FN(image_block)(g, ib, image, *cx, *cz, *lcx, *lcz);
        }
        T *t;
        t = *cx; *cx = *px; *px = t;
        t = *cz; *cz = *pz; *pz = t;
        t = *lcx; *lcx = *lpx; *lpx = t;
        t = *lcz; *lcz = *lpz; *lpz = t;
    }
    free(fine);
}

#undef FN
#undef S
#undef S_
