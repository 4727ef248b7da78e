/*
 * The reference's CPU path for the 3D VTI step, restated in C with OpenMP.
 * Test infrastructure and the CPU baseline of bench.py only.
 *
 * lib/vti-fwi-3d/cpu/laplacian3d.ispc:3-24 (forward: c11, c13, c33 formed
 * per point from v, eps, del) and :113-131 (forwardCij), one x-row per call,
 * all in FP32 (ISPC floating literals are float); the modeling loop of
 * lib/vti-fwi-3d/cpu/vtifd3d8o-modeling.cpp:165-189: the update over the
 * interior [4, n-4) of every axis, swap, source at the new level, then abc
 * (:26-57) on both levels, with the 1D profile
 * damping[i] = 0.95 + 0.05*cos(pi*(npad-i)/npad), i < npad (:81).
 * The reference runs the x-rows through TBB over x/y/z blocks; here OpenMP
 * runs them over (z, y) -- the arithmetic per point is the same.
 *
 * The backward level of the CPU gradient (vtifd3d8o-objfuncGrad.cpp:322-367):
 * gradientCij (chapter2/Src/laplacian3d.ispc:202-223, c_ij formed per point
 * from v, eps, del; the source-side level reconstructed with /d2 stencils and
 * the imaging gc11 += lx (dxx+dyy), gc33 += lz dzz, gc13 += lx dzz + lz
 * (dxx+dyy)), forward on the adjoint, then abc on the adjoint's two levels.
 */
#include <math.h>
#include <stddef.h>

#define CC0 (-2.84722222222f)
#define CC1 (1.6f)
#define CC2 (-0.2f)
#define CC3 (0.02539682539f)
#define CC4 (-0.00178571428f)

static inline float d2(const float *a, long s)
{
    return CC0 * a[0] + CC1 * (a[s] + a[-s]) + CC2 * (a[2 * s] + a[-2 * s]) + CC3 * (a[3 * s] + a[-3 * s]) +
           CC4 * (a[4 * s] + a[-4 * s]);
}

/* forward / forwardCij on one x-row of n points starting at i0 */
static void row_forward(int mode, long n, long nx, long nxy, float dx2, float dy2, float dz2, float dt2,
                        const float *a, const float *b, const float *c, float *px, const float *cx, float *pz,
                        const float *cz)
{
    for (long i = 0; i < n; ++i) {
        float c11, c13, c33;
        if (mode == 0) { /* a, b, c = v, eps, del */
            c33 = a[i] * a[i];
            c11 = c33 * (1.f + 2.f * b[i]);
            c13 = c33 * sqrtf(1.f + 2.f * c[i]);
        } else { /* a, b, c = c11, c13, c33 */
            c11 = a[i];
            c13 = b[i];
            c33 = c[i];
        }
        const float dxx = d2(cx + i, 1) / dx2;
        const float dyy = d2(cx + i, nx) / dy2;
        const float dzz = d2(cz + i, nxy) / dz2;
        px[i] = 2.f * cx[i] + dt2 * (c11 * (dxx + dyy) + c13 * dzz) - px[i];
        pz[i] = 2.f * cz[i] + dt2 * (c13 * (dxx + dyy) + c33 * dzz) - pz[i];
    }
}

static void abc(float *s, const float *damping, int nx, int ny, int nz, int npad)
{
    const long nxy = (long)nx * ny;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < npad; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                s[ix + iy * (long)nx + iz * nxy] *= damping[iz];
                s[ix + iy * (long)nx + (nz - 1 - iz) * nxy] *= damping[iz];
            }
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < npad; iy++)
            for (int ix = 0; ix < nx; ix++) {
                s[ix + iy * (long)nx + iz * nxy] *= damping[iy];
                s[ix + (ny - 1 - iy) * (long)nx + iz * nxy] *= damping[iy];
            }
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < npad; ix++) {
                s[ix + iy * (long)nx + iz * nxy] *= damping[ix];
                s[nx - 1 - ix + iy * (long)nx + iz * nxy] *= damping[ix];
            }
}

/* gradientCij on one x-row */
static void row_gradient(long n, long nx, long nxy, float dx2, float dy2, float dz2, float dt2, const float *v,
                         const float *eps, const float *del, float *px, const float *cx, float *pz, const float *cz,
                         const float *lx, const float *lz, float *g11, float *g13, float *g33)
{
    for (long i = 0; i < n; ++i) {
        const float c33 = v[i] * v[i];
        const float c11 = c33 * (1.f + 2.f * eps[i]);
        const float c13 = c33 * sqrtf(1.f + 2.f * del[i]);
        const float dxx = d2(cx + i, 1) / dx2;
        const float dyy = d2(cx + i, nx) / dy2;
        const float dzz = d2(cz + i, nxy) / dz2;
        px[i] = 2.f * cx[i] + dt2 * (c11 * (dxx + dyy) + c13 * dzz) - px[i];
        pz[i] = 2.f * cz[i] + dt2 * (c13 * (dxx + dyy) + c33 * dzz) - pz[i];
        g11[i] += lx[i] * (dxx + dyy);
        g33[i] += lz[i] * dzz;
        g13[i] += lx[i] * dzz + lz[i] * (dxx + dyy);
    }
}

void cpu3d_damping(float *damping, int npad)
{
    for (int i = 0; i < npad; i++) damping[i] = (float)(0.95 + (1. - 0.95) * cos(3.14159265359 * (npad - i) / npad));
}

/* nsteps levels of the modeling loop on (prev, cur) of sigma_x and sigma_z;
 * mode 0: a, b, c = v, eps, del; mode 1: c11, c13, c33.  src_index < 0: no
 * source; otherwise wavelet[k] * dt2 is added at the new level of step k.
 * On return the newest level is in cur_*. */
int cpu3d_steps(int mode, float *prev_x, float *cur_x, float *prev_z, float *cur_z, const float *a, const float *b,
                const float *c, int nx, int ny, int nz, int npad, float dx, float dy, float dz, float dt, int nsteps,
                long src_index, const float *wavelet)
{
    const long nxy = (long)nx * ny, nxyz = nxy * nz;
    const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz, dt2 = dt * dt;
    float damping[4096];
    if (npad > 4096 || nx < 9 || ny < 9 || nz < 9) return -1;
    cpu3d_damping(damping, npad);
    float *px = prev_x, *cx = cur_x, *pz = prev_z, *cz = cur_z;
    for (int k = 0; k < nsteps; ++k) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int iz = 4; iz < nz - 4; ++iz)
            for (int iy = 4; iy < ny - 4; ++iy) {
                const long i = 4 + iy * (long)nx + iz * nxy;
                row_forward(mode, nx - 8, nx, nxy, dx2, dy2, dz2, dt2, a + i, b + i, c + i, px + i, cx + i, pz + i,
                            cz + i);
            }
        float *t = px;
        px = cx;
        cx = t;
        t = pz;
        pz = cz;
        cz = t;
        if (src_index >= 0 && src_index < nxyz && wavelet) {
            const float s = wavelet[k] * dt2;
            cx[src_index] += s;
            cz[src_index] += s;
        }
        if (npad > 0) {
            abc(px, damping, nx, ny, nz, npad);
            abc(pz, damping, nx, ny, nz, npad);
            abc(cx, damping, nx, ny, nz, npad);
            abc(cz, damping, nx, ny, nz, npad);
        }
    }
    if (cx != cur_x) { /* odd number of steps: the newest level sits in the prev arrays; swap contents */
        for (long i = 0; i < nxyz; ++i) {
            float u = cur_x[i];
            cur_x[i] = prev_x[i];
            prev_x[i] = u;
            u = cur_z[i];
            cur_z[i] = prev_z[i];
            prev_z[i] = u;
        }
    }
    return 0;
}

/* nsteps backward levels of the CPU gradient on caller-owned state (v, eps,
 * del model; sigma and lambda as (prev, cur) pairs; gc accumulated), no
 * source and no receivers: the loop body of vtifd3d8o-objfuncGrad.cpp:322-367
 * without its point injections.  The pairs swap after each level, so after an
 * odd count the newest levels sit in the prev arrays. */
int cpu3d_backward_steps(float *sprev_x, float *scur_x, float *sprev_z, float *scur_z, float *lprev_x, float *lcur_x,
                         float *lprev_z, float *lcur_z, float *g11, float *g13, float *g33, const float *v,
                         const float *eps, const float *del, int nx, int ny, int nz, int npad, float dx, float dy,
                         float dz, float dt, int nsteps)
{
    const long nxy = (long)nx * ny;
    const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz, dt2 = dt * dt;
    float damping[4096];
    if (npad > 4096 || nx < 9 || ny < 9 || nz < 9) return -1;
    cpu3d_damping(damping, npad);
    float *px = sprev_x, *cx = scur_x, *pz = sprev_z, *cz = scur_z;
    float *qx = lprev_x, *ax = lcur_x, *qz = lprev_z, *az = lcur_z;
    for (int k = 0; k < nsteps; ++k) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int iz = 4; iz < nz - 4; ++iz)
            for (int iy = 4; iy < ny - 4; ++iy) {
                const long i = 4 + iy * (long)nx + iz * nxy;
                row_gradient(nx - 8, nx, nxy, dx2, dy2, dz2, dt2, v + i, eps + i, del + i, px + i, cx + i, pz + i,
                             cz + i, ax + i, az + i, g11 + i, g13 + i, g33 + i);
                row_forward(0, nx - 8, nx, nxy, dx2, dy2, dz2, dt2, v + i, eps + i, del + i, qx + i, ax + i, qz + i,
                            az + i);
            }
        float *t;
        t = px; px = cx; cx = t;
        t = pz; pz = cz; cz = t;
        t = qx; qx = ax; ax = t;
        t = qz; qz = az; az = t;
        if (npad > 0) {
            abc(qx, damping, nx, ny, nz, npad);
            abc(qz, damping, nx, ny, nz, npad);
            abc(ax, damping, nx, ny, nz, npad);
            abc(az, damping, nx, ny, nz, npad);
        }
    }
    return 0;
}
