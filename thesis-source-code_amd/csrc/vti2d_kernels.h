// 2D VTI kernels for gfx950 (included by vti2d.hip).
//
// Grid: the reference's layout i = ix + iz*nnx (x fastest); every kernel
// updates the interior ix in [4, nnx-4), iz in [4, nnz-4) only
// (lib/vti-fwi-2d/kernels.cu).  p needs only its x derivative and q only its
// z derivative, so a workgroup of 64 x BY threads keeps two LDS tiles: p rows
// with a 4-point x halo and q columns with a 4-point z halo.
//
// Arithmetic is the reference's: FP32 with float weights, every expression
// evaluated in source order with contraction off, so results reproduce the
// oracle's FP32 restatement bit for bit.
//
// Each step kernel fuses what the reference runs as separate launches per
// time step: the stencil update, the source or receiver injection, the
// taper of the absorbing boundary (abc), the D snapshot of the current level
// and the imaging condition.  The taper the reference applies in place to the
// current level after each step is applied here when that level is read
// back as the previous level.
#pragma once

#include "vti_internal.h"

#pragma clang fp contract(off)

namespace vti {

struct Geo2 {
    int nnx, nnz;
    float rx2, rz2;  // RN(1/dx2), RN(1/dz2) when qdiv may use them, else 0
};

constexpr int kBY2 = 8;

// 9-point second difference of the reference (kernels.cu:486-493)
__device__ __forceinline__ float st9(float c, float a1, float b1, float a2, float b2, float a3, float b3, float a4,
                                     float b4)
{
    return VTI_C0 * c + VTI_C1 * (a1 + b1) + VTI_C2 * (a2 + b2) + VTI_C3 * (a3 + b3) + VTI_C4 * (a4 + b4);
}

// Origin of the workgroup's 64 x kBY2 tile in a (tiles across) x (tiles
// down) grid, in XCD-aware order.  Workgroups are dealt to the 8 XCDs in
// turn (linear id b to XCD b % 8) and each XCD has its own L2, so XCD x
// takes the contiguous run [x q + min(x, r), ...) of the nb = 8 q + r tiles
// in row-major order: tiles that read each other's halos share one L2.
__device__ __forceinline__ void tile_origin(int &x0, int &z0)
{
    const int nb = gridDim.x * gridDim.y, b = blockIdx.x + blockIdx.y * gridDim.x;
    const int x = b & 7, q = nb >> 3, r = nb & 7;
    const int t = x * q + min(x, r) + (b >> 3);
    x0 = 4 + (t % (int)gridDim.x) * 64;
    z0 = 4 + (t / (int)gridDim.x) * kBY2;
}

struct Tiles2 {
    float p[kBY2][72];      // x-halo tile
    float q[kBY2 + 8][64];  // z-halo tile
};

// Load the tiles.  MODE 0: p and q themselves.  MODE 1: c11*p + c13*q and
// c13*p + c33*q (the products the exact adjoint differentiates, kernels.cu:116-132).
template <int MODE>
__device__ __forceinline__ void load_tiles(Tiles2 &t, const float *__restrict__ P, const float *__restrict__ Q,
                                           const float *__restrict__ c11, const float *__restrict__ c13,
                                           const float *__restrict__ c33, const Geo2 &g, int x0, int z0)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto pv = [&](int x, int z) -> float {
        if (x < 0 || x >= g.nnx || z < 0 || z >= g.nnz) return 0.f;
        const long long i = x + (long long)z * g.nnx;
        if (MODE == 0) return P[i];
        return c11[i] * P[i] + c13[i] * Q[i];
    };
    auto qv = [&](int x, int z) -> float {
        if (x < 0 || x >= g.nnx || z < 0 || z >= g.nnz) return 0.f;
        const long long i = x + (long long)z * g.nnx;
        if (MODE == 0) return Q[i];
        return c13[i] * P[i] + c33[i] * Q[i];
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

// load_tiles<0> that also returns the raw p and q of the thread's own point
// (0 outside the grid), so the step need not read them again
__device__ __forceinline__ void load_tiles_c(Tiles2 &t, const float *__restrict__ P, const float *__restrict__ Q,
                                             const Geo2 &g, int x0, int z0, float &pc, float &qc)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto pv = [&](int x, int z) -> float {
        if (x < 0 || x >= g.nnx || z < 0 || z >= g.nnz) return 0.f;
        return P[x + (long long)z * g.nnx];
    };
    auto qv = [&](int x, int z) -> float {
        if (x < 0 || x >= g.nnx || z < 0 || z >= g.nnz) return 0.f;
        return Q[x + (long long)z * g.nnx];
    };
    const int ix = x0 + tx, iz = z0 + ty;
    pc = pv(ix, iz);
    qc = qv(ix, iz);
    t.p[ty][tx + 4] = pc;
    if (tx < 4) {
        t.p[ty][tx] = pv(x0 - 4 + tx, iz);
        t.p[ty][tx + 68] = pv(x0 + 64 + tx, iz);
    }
    t.q[ty + 4][tx] = qc;
    if (ty < 4) {
        t.q[ty][tx] = qv(ix, z0 - 4 + ty);
        t.q[ty + kBY2 + 4][tx] = qv(ix, z0 + kBY2 + ty);
    }
}

// x / d correctly rounded, for a divisor d fixed per launch.  y is RN(1/d)
// when the host has shown (vti::div_reciprocal) that Markstein's correction
// step -- q0 = RN(x y), r = x - q0 d (exact, one fma), q = RN(q0 + r y) --
// gives RN(x/d) for every x in [1,2).  For d in [2^-20, 2^20] every
// intermediate then scales exactly with x's exponent, which carries the
// result to all |x| in [2^-100, 2^100).  A zero x comes out as +0 and
// takes x's sign bit (a nonzero q has x's sign already).  Other magnitudes,
// and y == 0 (no proof for this d), take the division.
__device__ __forceinline__ float qdiv(float x, float d, float y)
{
    if (y == 0.f) return x / d;
    const float q0 = x * y;
    const float r = __builtin_fmaf(-q0, d, x);
    const unsigned xb = __float_as_uint(x), a = xb & 0x7fffffffu;
    float q = __uint_as_float((__float_as_uint(__builtin_fmaf(r, y, q0)) & 0x7fffffffu) | (xb & 0x80000000u));
    if (a - 0x0d800000u >= 0x64000000u && a != 0u) {
        // the empty volatile statement keeps the compiler from computing the
        // division for every lane and selecting it afterwards
        float xs = x;
        asm volatile("" : "+v"(xs));
        q = xs / d;
    }
    return q;
}

__device__ __forceinline__ float dxx(const Tiles2 &t, float dx2, float rx2)
{
    const int c = threadIdx.x + 4, r = threadIdx.y;
    const float *a = t.p[r];
    return qdiv(st9(a[c], a[c - 1], a[c + 1], a[c - 2], a[c + 2], a[c - 3], a[c + 3], a[c - 4], a[c + 4]), dx2, rx2);
}

__device__ __forceinline__ float dzz(const Tiles2 &t, float dz2, float rz2)
{
    const int c = threadIdx.x, r = threadIdx.y + 4;
    return qdiv(st9(t.q[r][c], t.q[r - 1][c], t.q[r + 1][c], t.q[r - 2][c], t.q[r + 2][c], t.q[r - 3][c],
                    t.q[r + 3][c], t.q[r - 4][c], t.q[r + 4][c]),
                dz2, rz2);
}

// Forward step: forwardCD / forwardC semantics (kernels.cu:452-498, 546-556)
// + injectSource (or, with src_neg >= 0, injectDipoleSource) + abc;
// optionally writes D of the current level (the
// snapshot the reference's D kernel takes, kernels.cu:345-387).
// BORN: also steps the Born field dp (GNhessian_f, hessian_f.cu:752-806) with
// the scattering source dt2*(dC * D p) (kernels.cu:282-291), and optionally
// writes D of the Born level read (bsnap).
template <bool BORN>
__global__ __launch_bounds__(64 * kBY2) void fwd2d_kernel(
    float *__restrict__ p_n, float *__restrict__ q_n, const float *__restrict__ p_c, const float *__restrict__ q_c,
    float *__restrict__ dp_n, float *__restrict__ dq_n, const float *__restrict__ dp_c,
    const float *__restrict__ dq_c, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ dc11, const float *__restrict__ dc13,
    const float *__restrict__ dc33, const float *__restrict__ taper, Geo2 g, float dx2, float dz2, float dt2,
    long long src, float src_val, float *__restrict__ snap_p, float *__restrict__ snap_q, long long src_neg,
    float *__restrict__ bsnap_p, float *__restrict__ bsnap_q)
{
    __shared__ Tiles2 t;
    __shared__ Tiles2 tb;
    int x0, z0;
    tile_origin(x0, z0);
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    const bool act = ix < g.nnx - 4 && iz < g.nnz - 4;
    const long long i = ix + (long long)iz * g.nnx;
    // the point's own operands, loaded with the tiles before the barrier
    float tap = 0.f, a11 = 0.f, a13 = 0.f, a33 = 0.f, pp = 0.f, qp = 0.f, pcur, qcur;
    if (act) {
        tap = taper[i];
        a11 = c11[i];
        a13 = c13[i];
        a33 = c33[i];
        pp = p_n[i];
        qp = q_n[i];
    }
    load_tiles_c(t, p_c, q_c, g, x0, z0, pcur, qcur);
    if (BORN) load_tiles<0>(tb, dp_c, dq_c, nullptr, nullptr, nullptr, g, x0, z0);
    __syncthreads();
    if (!act) return;
    const float tp = dxx(t, dx2, g.rx2), tq = dzz(t, dz2, g.rz2);
    if (snap_p) {
        snap_p[i] = tp;
        snap_q[i] = tq;
    }
    float pn = 2.f * pcur + dt2 * (a11 * tp + a13 * tq) - pp * tap;
    float qn = 2.f * qcur + dt2 * (a13 * tp + a33 * tq) - qp * tap;
    if (i == src) {
        pn += src_val;
        qn += src_val;
    }
    if (i == src_neg) {  // the negative pole of injectDipoleSource (kernels.cu:251-257)
        pn -= src_val;
        qn -= src_val;
    }
    p_n[i] = pn * tap;
    q_n[i] = qn * tap;
    if (BORN) {
        const float btp = dxx(tb, dx2, g.rx2), btq = dzz(tb, dz2, g.rz2);
        if (bsnap_p) {  // D of the Born level read (hessian_f.cu:236)
            bsnap_p[i] = btp;
            bsnap_q[i] = btq;
        }
        float dpn = 2.f * dp_c[i] + dt2 * (a11 * btp + a13 * btq) - dp_n[i] * tap;
        float dqn = 2.f * dq_c[i] + dt2 * (a13 * btp + a33 * btq) - dq_n[i] * tap;
        dpn += dt2 * (dc11[i] * tp + dc13[i] * tq);
        dqn += dt2 * (dc13[i] * tp + dc33[i] * tq);
        dp_n[i] = dpn * tap;
        dq_n[i] = dqn * tap;
    }
}

// D of a level into a snapshot slot (the last level of the forward pass).
static __global__ __attribute__((unused)) __launch_bounds__(64 * kBY2) void dsnap2d_kernel(const float *__restrict__ p, const float *__restrict__ q,
                                                           Geo2 g, float dx2, float dz2, float *__restrict__ sp,
                                                           float *__restrict__ sq)
{
    __shared__ Tiles2 t;
    int x0, z0;
    tile_origin(x0, z0);
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    load_tiles<0>(t, p, q, nullptr, nullptr, nullptr, g, x0, z0);
    __syncthreads();
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    sp[i] = dxx(t, dx2, g.rx2);
    sq[i] = dzz(t, dz2, g.rz2);
}

// Adjoint step fused with injectData, abc and gradientCrossCor
// (objFuncGradient_f.cu:249-300 / hessian_f.cu:853-900).
// ADJ 0: backwardDC, the exact transpose (kernels.cu:104-148).
// ADJ 1: forwardCD, the forward operator (GNhessian_f's adjoint, hessian_f.cu:854).
// recmap[i] is the receiver index injected at point i (or -1); the injected
// value is 0.5*((1-fr)*r0 + fr*r1) of the two trace rows bracketing the time.
// The image uses the snapshots D0, D1 of the levels bracketing this one.
template <int ADJ>
__global__ __launch_bounds__(64 * kBY2) void bwd2d_kernel(
    float *__restrict__ p_n, float *__restrict__ q_n, const float *__restrict__ p_c, const float *__restrict__ q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ taper, const int *__restrict__ recmap, const float *__restrict__ r0,
    const float *__restrict__ r1, float fr, const float *__restrict__ D0p, const float *__restrict__ D0q,
    const float *__restrict__ D1p, const float *__restrict__ D1q, float f, float *__restrict__ g11,
    float *__restrict__ g13, float *__restrict__ g33, Geo2 g, float dx2, float dz2, float dt2)
{
    __shared__ Tiles2 t;
    int x0, z0;
    tile_origin(x0, z0);
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    const bool act = ix < g.nnx - 4 && iz < g.nnz - 4;
    const long long i = ix + (long long)iz * g.nnx;
    // the point's own operands, loaded with the tiles before the barrier
    float tap = 0.f, pc = 0.f, qc = 0.f, pp = 0.f, qp = 0.f, a11 = 0.f, a13 = 0.f, a33 = 0.f, s = 0.f;
    float d0p = 0.f, d0q = 0.f, d1p = 0.f, d1q = 0.f, o11 = 0.f, o13 = 0.f, o33 = 0.f;
    int ir = -1;
    if (act) {
        tap = taper[i];
        pc = p_c[i];
        qc = q_c[i];
        pp = p_n[i];
        qp = q_n[i];
        if (ADJ == 1) {
            a11 = c11[i];
            a13 = c13[i];
            a33 = c33[i];
        }
        ir = recmap[i];
        if (ir >= 0) s = 0.5f * ((1.f - fr) * r0[ir] + fr * r1[ir]);
        d0p = D0p[i];
        d0q = D0q[i];
        d1p = D1p[i];
        d1q = D1q[i];
        o11 = g11[i];
        o13 = g13[i];
        o33 = g33[i];
    }
    load_tiles<ADJ == 0 ? 1 : 0>(t, p_c, q_c, c11, c13, c33, g, x0, z0);
    __syncthreads();
    if (!act) return;
    float pn, qn;
    if (ADJ == 0) {
        const float tp0 = dxx(t, dx2, g.rx2), tq0 = dzz(t, dz2, g.rz2);
        pn = 2.f * pc + dt2 * tp0 - pp * tap;
        qn = 2.f * qc + dt2 * tq0 - qp * tap;
    } else {
        const float tp = dxx(t, dx2, g.rx2), tq = dzz(t, dz2, g.rz2);
        pn = 2.f * pc + dt2 * (a11 * tp + a13 * tq) - pp * tap;
        qn = 2.f * qc + dt2 * (a13 * tp + a33 * tq) - qp * tap;
    }
    if (ir >= 0) {
        pn += dt2 * s;
        qn += dt2 * s;
    }
    pn = pn * tap;
    qn = qn * tap;
    p_n[i] = pn;
    q_n[i] = qn;
    const float Dp = (1.f - f) * d0p + f * d1p;
    const float Dq = (1.f - f) * d0q + f * d1q;
    g11[i] = o11 + pn * Dp;
    g13[i] = o13 + (pn * Dq + qn * Dp);
    g33[i] = o33 + qn * Dq;
}

// hessian_f's adjoint level (hessian_f.cu:387-443): the residual's adjoint
// lambda and its perturbation dlambda each take a backwardDC step
// (kernels.cu:104-148) with their injected traces (the residual, the Born
// data), dlambda the second-order source dt2 * D(dC lambda) (scatteringa,
// kernels.cu:150-195), abc, and the image of both: lambda against the Born
// snapshots, then dlambda against the background ones.
static __global__ __attribute__((unused)) __launch_bounds__(64 * kBY2) void hess_bwd2d_kernel(
    float *__restrict__ p_n, float *__restrict__ q_n, const float *__restrict__ p_c, const float *__restrict__ q_c,
    float *__restrict__ dp_n, float *__restrict__ dq_n, const float *__restrict__ dp_c,
    const float *__restrict__ dq_c, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ dc11, const float *__restrict__ dc13,
    const float *__restrict__ dc33, const float *__restrict__ taper, const int *__restrict__ recmap,
    const float *__restrict__ r0, const float *__restrict__ r1, const float *__restrict__ b0,
    const float *__restrict__ b1, float fr, const float *__restrict__ BD0p, const float *__restrict__ BD0q,
    const float *__restrict__ BD1p, const float *__restrict__ BD1q, const float *__restrict__ D0p,
    const float *__restrict__ D0q, const float *__restrict__ D1p, const float *__restrict__ D1q, float f,
    float *__restrict__ g11, float *__restrict__ g13, float *__restrict__ g33, Geo2 g, float dx2, float dz2,
    float dt2)
{
    __shared__ Tiles2 ta, tb, ts;
    int x0, z0;
    tile_origin(x0, z0);
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    load_tiles<1>(ta, p_c, q_c, c11, c13, c33, g, x0, z0);
    load_tiles<1>(tb, dp_c, dq_c, c11, c13, c33, g, x0, z0);
    load_tiles<1>(ts, p_c, q_c, dc11, dc13, dc33, g, x0, z0);
    __syncthreads();
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    const float tap = taper[i];
    float pn = 2.f * p_c[i] + dt2 * dxx(ta, dx2, g.rx2) - p_n[i] * tap;
    float qn = 2.f * q_c[i] + dt2 * dzz(ta, dz2, g.rz2) - q_n[i] * tap;
    float dpn = 2.f * dp_c[i] + dt2 * dxx(tb, dx2, g.rx2) - dp_n[i] * tap;
    float dqn = 2.f * dq_c[i] + dt2 * dzz(tb, dz2, g.rz2) - dq_n[i] * tap;
    const int ir = recmap[i];
    if (ir >= 0) {
        const float s = 0.5f * ((1.f - fr) * r0[ir] + fr * r1[ir]);
        pn += dt2 * s;
        qn += dt2 * s;
        const float sb = 0.5f * ((1.f - fr) * b0[ir] + fr * b1[ir]);
        dpn += dt2 * sb;
        dqn += dt2 * sb;
    }
    dpn += dt2 * dxx(ts, dx2, g.rx2);
    dqn += dt2 * dzz(ts, dz2, g.rz2);
    pn = pn * tap;
    qn = qn * tap;
    dpn = dpn * tap;
    dqn = dqn * tap;
    p_n[i] = pn;
    q_n[i] = qn;
    dp_n[i] = dpn;
    dq_n[i] = dqn;
    const float BDp = (1.f - f) * BD0p[i] + f * BD1p[i];
    const float BDq = (1.f - f) * BD0q[i] + f * BD1q[i];
    const float Dp = (1.f - f) * D0p[i] + f * D1p[i];
    const float Dq = (1.f - f) * D0q[i] + f * D1q[i];
    float a11 = g11[i], a13 = g13[i], a33 = g33[i];
    a11 += pn * BDp;
    a13 += pn * BDq + qn * BDp;
    a33 += qn * BDq;
    a11 += dpn * Dp;
    a13 += dpn * Dq + dqn * Dp;
    a33 += dqn * Dq;
    g11[i] = a11;
    g13[i] = a13;
    g33[i] = a33;
}

// gradientCrossCor on a given level (the first adjoint level, which the
// reference images before its backward loop, objFuncGradient_f.cu:247).
static __global__ __attribute__((unused)) void xcor2d_kernel(const float *__restrict__ ap, const float *__restrict__ aq,
                              const float *__restrict__ D0p, const float *__restrict__ D0q,
                              const float *__restrict__ D1p, const float *__restrict__ D1q, float f,
                              float *__restrict__ g11, float *__restrict__ g13, float *__restrict__ g33, Geo2 g)
{
    const int ix = 4 + blockIdx.x * blockDim.x + threadIdx.x, iz = 4 + blockIdx.y;
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    const float Dp = (1.f - f) * D0p[i] + f * D1p[i];
    const float Dq = (1.f - f) * D0q[i] + f * D1q[i];
    g11[i] += ap[i] * Dp;
    g13[i] += ap[i] * Dq + aq[i] * Dp;
    g33[i] += aq[i] * Dq;
}

}  // namespace vti
