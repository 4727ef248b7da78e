// Batched 2D kernels: many shots, each on its own aperture sub-grid of one
// global model, advanced one time level per launch.
//
// A shot's fields (p, q at two levels, its gradient and its receiver map)
// live in its own nnz x nnx block of the batch buffers; its coefficients are
// the columns [c_off, c_off + nnx) of the global c11/c13/c33 (row pitch
// cpitch).  Workgroups are numbered over all shots' 64 x 8 tiles;
// tile_shot[] maps a workgroup to its shot.  The taper is the product of the
// 1D profile t0 of init_abc (lib/vti-fwi-2d/init.cpp:70-89) along x and z,
// formed in float as the reference forms its taper array.
//
// Per point the arithmetic is exactly that of vti2d_kernels.h (FP32, source
// order, contraction off), so a shot in a batch reproduces the single-shot
// driver on its cropped model bit for bit.
#pragma once

#include "vti2d_kernels.h"

#pragma clang fp contract(off)

namespace vti {

struct ShotB {
    int nnx;            // padded width of the shot's sub-grid
    int ntx;            // 64-wide tiles across it
    int tile0;          // first workgroup of this shot
    int nr;             // receivers
    long long f_off;    // offset of the shot's nnz x nnx block in the field buffers
    long long c_off;    // first column of the sub-grid in the global model
    long long snap_off; // offset of the shot's snapshots (nnt slots of 2 nnz x nnx planes)
    long long r_off;    // offset of the shot's traces (time-major, nr per row)
    int src;            // source point within the sub-grid (ix + iz * nnx)
    int pad_;
};

struct BatchGeo {
    int nnz, npad, cpitch;
    long long nnxz_total;
};

// taper of the point (ix, iz) of a sub-grid nnx wide: t0 along x, then along z
__device__ __forceinline__ float taper_at(const float *__restrict__ t0, int npad, int nnx, int nnz, int ix, int iz)
{
    float t = 1.f;
    if (ix < npad) t = t * t0[ix];
    else if (ix >= nnx - npad) t = t * t0[nnx - 1 - ix];
    if (iz < npad) t = t * t0[iz];
    else if (iz >= nnz - npad) t = t * t0[nnz - 1 - iz];
    return t;
}

// LDS tiles of a sub-grid whose coefficients sit at another pitch/offset
template <int MODE>
__device__ __forceinline__ void load_tiles_b(Tiles2 &t, const float *__restrict__ P, const float *__restrict__ Q,
                                             const float *__restrict__ c11, const float *__restrict__ c13,
                                             const float *__restrict__ c33, int nnx, int nnz, int cpitch, int x0,
                                             int z0)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto pv = [&](int x, int z) -> float {
        if (x < 0 || x >= nnx || z < 0 || z >= nnz) return 0.f;
        const long long i = x + (long long)z * nnx;
        if (MODE == 0) return P[i];
        const long long c = x + (long long)z * cpitch;
        return c11[c] * P[i] + c13[c] * Q[i];
    };
    auto qv = [&](int x, int z) -> float {
        if (x < 0 || x >= nnx || z < 0 || z >= nnz) return 0.f;
        const long long i = x + (long long)z * nnx;
        if (MODE == 0) return Q[i];
        const long long c = x + (long long)z * cpitch;
        return c13[c] * P[i] + c33[c] * Q[i];
    };
    const int ix = x0 + tx, iz = z0 + ty;
    t.p[ty][tx + 4] = pv(ix, iz);
    if (tx < 4) {
        t.p[ty][tx] = pv(x0 - 4 + tx, iz);
        t.p[ty][tx + 68] = pv(x0 + 64 + tx, iz);
    }
    t.q[ty + 4][tx] = qv(ix, iz);
    if (ty < 4) {
        t.q[ty][tx] = qv(ix, z0 - 4 + ty);
        t.q[ty + kBY2 + 4][tx] = qv(ix, z0 + kBY2 + ty);
    }
}

// forwardC step of every shot with injectSource, abc and (slot >= 0) the D
// snapshot of the level read (objFuncGradient_f.cu:160-210 per shot).
static __global__ __launch_bounds__(64 * kBY2) void fwd2db_kernel(
    float *__restrict__ P_n, float *__restrict__ Q_n, const float *__restrict__ P_c, const float *__restrict__ Q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ t0, const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
    float dx2, float dz2, float dt2, float src_val, float *__restrict__ snap, int slot)
{
    __shared__ Tiles2 t;
    const ShotB sh = shots[tile_shot[blockIdx.x]];
    const int tl = blockIdx.x - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    const float *p_c = P_c + sh.f_off, *q_c = Q_c + sh.f_off;
    load_tiles_b<0>(t, p_c, q_c, nullptr, nullptr, nullptr, nnx, nnz, 0, x0, z0);
    __syncthreads();
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    if (ix >= nnx - 4 || iz >= nnz - 4) return;
    const long long i = ix + (long long)iz * nnx;
    const long long c = sh.c_off + ix + (long long)iz * bg.cpitch;
    const float tp = dxx(t, dx2), tq = dzz(t, dz2);
    if (slot >= 0) {
        float *sp = snap + sh.snap_off + (long long)slot * 2 * nnx * nnz;
        sp[i] = tp;
        sp[i + (long long)nnx * nnz] = tq;
    }
    const float tap = taper_at(t0, bg.npad, nnx, nnz, ix, iz);
    float *p_n = P_n + sh.f_off, *q_n = Q_n + sh.f_off;
    const float a11 = c11[c], a13 = c13[c], a33 = c33[c];
    float pn = 2.f * p_c[i] + dt2 * (a11 * tp + a13 * tq) - p_n[i] * tap;
    float qn = 2.f * q_c[i] + dt2 * (a13 * tp + a33 * tq) - q_n[i] * tap;
    if (i == sh.src) {
        pn += src_val;
        qn += src_val;
    }
    p_n[i] = pn * tap;
    q_n[i] = qn * tap;
}

// D of the last forward level of every shot into its snapshot slot
static __global__ __launch_bounds__(64 * kBY2) void dsnap2db_kernel(const float *__restrict__ P, const float *__restrict__ Q,
                                                             const ShotB *__restrict__ shots,
                                                             const int *__restrict__ tile_shot, BatchGeo bg, float dx2,
                                                             float dz2, float *__restrict__ snap, int slot)
{
    __shared__ Tiles2 t;
    const ShotB sh = shots[tile_shot[blockIdx.x]];
    const int tl = blockIdx.x - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    load_tiles_b<0>(t, P + sh.f_off, Q + sh.f_off, nullptr, nullptr, nullptr, nnx, nnz, 0, x0, z0);
    __syncthreads();
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    if (ix >= nnx - 4 || iz >= nnz - 4) return;
    const long long i = ix + (long long)iz * nnx;
    float *sp = snap + sh.snap_off + (long long)slot * 2 * nnx * nnz;
    sp[i] = dxx(t, dx2);
    sp[i + (long long)nnx * nnz] = dzz(t, dz2);
}

// backwardDC step of every shot with its residual injection (receiver map),
// abc and gradientCrossCor against the interpolated snapshots
// (objFuncGradient_f.cu:249-300 per shot).
static __global__ __launch_bounds__(64 * kBY2) void bwd2db_kernel(
    float *__restrict__ P_n, float *__restrict__ Q_n, const float *__restrict__ P_c, const float *__restrict__ Q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ t0, const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
    const int *__restrict__ recmap, const float *__restrict__ traces, int row0, float fr,
    const float *__restrict__ snap, int slot0, float f, float *__restrict__ G11, float *__restrict__ G13,
    float *__restrict__ G33, float dx2, float dz2, float dt2)
{
    __shared__ Tiles2 t;
    const ShotB sh = shots[tile_shot[blockIdx.x]];
    const int tl = blockIdx.x - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    const float *p_c = P_c + sh.f_off, *q_c = Q_c + sh.f_off;
    const float *cc11 = c11 + sh.c_off, *cc13 = c13 + sh.c_off, *cc33 = c33 + sh.c_off;
    load_tiles_b<1>(t, p_c, q_c, cc11, cc13, cc33, nnx, nnz, bg.cpitch, x0, z0);
    __syncthreads();
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    if (ix >= nnx - 4 || iz >= nnz - 4) return;
    const long long i = ix + (long long)iz * nnx;
    const float tap = taper_at(t0, bg.npad, nnx, nnz, ix, iz);
    float *p_n = P_n + sh.f_off, *q_n = Q_n + sh.f_off;
    const float tp0 = dxx(t, dx2), tq0 = dzz(t, dz2);
    float pn = 2.f * p_c[i] + dt2 * tp0 - p_n[i] * tap;
    float qn = 2.f * q_c[i] + dt2 * tq0 - q_n[i] * tap;
    const int ir = recmap[sh.f_off + i];
    if (ir >= 0) {
        const float *r0 = traces + sh.r_off + (long long)row0 * sh.nr;
        const float s = 0.5f * ((1.f - fr) * r0[ir] + fr * r0[ir + sh.nr]);
        pn += dt2 * s;
        qn += dt2 * s;
    }
    pn = pn * tap;
    qn = qn * tap;
    p_n[i] = pn;
    q_n[i] = qn;
    const long long plane = (long long)nnx * nnz;
    const float *D0 = snap + sh.snap_off + (long long)slot0 * 2 * plane, *D1 = D0 + 2 * plane;
    const float Dp = (1.f - f) * D0[i] + f * D1[i];
    const float Dq = (1.f - f) * D0[i + plane] + f * D1[i + plane];
    G11[sh.f_off + i] += pn * Dp;
    G13[sh.f_off + i] += pn * Dq + qn * Dp;
    G33[sh.f_off + i] += qn * Dq;
}

// gradientCrossCor of the first adjoint level of every shot
static __global__ void xcor2db_kernel(const float *__restrict__ P, const float *__restrict__ Q,
                               const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
                               const float *__restrict__ snap, int slot0, float f, float *__restrict__ G11,
                               float *__restrict__ G13, float *__restrict__ G33)
{
    const ShotB sh = shots[tile_shot[blockIdx.x]];
    const int tl = blockIdx.x - sh.tile0;
    const int ix = 4 + (tl % sh.ntx) * 64 + threadIdx.x, iz = 4 + (tl / sh.ntx) * kBY2 + threadIdx.y;
    const int nnx = sh.nnx, nnz = bg.nnz;
    if (ix >= nnx - 4 || iz >= nnz - 4) return;
    const long long i = ix + (long long)iz * nnx, k = sh.f_off + i;
    const long long plane = (long long)nnx * nnz;
    const float *D0 = snap + sh.snap_off + (long long)slot0 * 2 * plane, *D1 = D0 + 2 * plane;
    const float Dp = (1.f - f) * D0[i] + f * D1[i];
    const float Dq = (1.f - f) * D0[i + plane] + f * D1[i + plane];
    G11[k] += P[k] * Dp;
    G13[k] += P[k] * Dq + Q[k] * Dp;
    G33[k] += Q[k] * Dq;
}

}  // namespace vti
