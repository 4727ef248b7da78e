// Batched 2D kernels: many shots, each on its own aperture sub-grid of one
// global model, advanced one or two time levels per launch.
//
// A shot's fields (p, q, its gradient and its receiver map) live in its own
// nnz x nnx block of the batch buffers; its coefficients are the columns
// [c_off, c_off + nnx) of the global c11/c13/c33 (row pitch cpitch).
// Workgroups are numbered over all shots' tiles; tile_shot[] maps a
// workgroup to its shot.  The taper is the product of the 1D profile t0 of
// init_abc (lib/vti-fwi-2d/init.cpp:70-89) along x and z, formed in float as
// the reference forms its taper array.
//
// Per point the arithmetic is exactly that of vti2d_kernels.h (FP32, source
// order, contraction off), so a shot in a batch reproduces the single-shot
// driver on its cropped model bit for bit when the snapshots are kept in
// FP32.
//
// Snapshot formats (SF): 0 = FP32; 1 = FP16 and 2 = BF16, each slice (one
// shot, one slot, one of the two components) scaled by a power of two s so
// that its largest magnitude lies in [2^14, 2^15).  The forward kernels then
// write the FP32 slice into a staging area and fold |D| into a per-slice
// maximum (one atomic per block); pack_snap_kernel rounds the staged slice to
// 16 bits after the launch.  The adjoint reads h * s.
// 3 = S16: 16-bit signed integers with a power-of-two scale per segment of 64
// consecutive points of a slice (the scale puts the segment's largest
// magnitude in [2^14, 2^15)), so every value keeps 15 bits relative to its
// neighbourhood's largest instead of FP16's 11 bits relative to itself;
// pack_s16_kernel finds each segment's maximum itself (no atomics).
//
// Every step kernel issues the loads of its points' own operands together
// with the LDS tile loads, before the barrier, so that a workgroup waits for
// one round trip to memory rather than two.
#pragma once

#include <hip/hip_fp16.h>

#include "vti2d_kernels.h"

#pragma clang fp contract(off)

namespace vti {

struct ShotB {
    int nnx;              // padded width of the shot's sub-grid
    int ntx;              // 64-wide tiles across it
    int tile0;            // first workgroup of this shot
    int nr;               // receivers
    long long f_off;      // offset of the shot's nnz x nnx block in the field buffers
    long long c_off;      // first column of the sub-grid in the global model
    long long snap_off;   // offset of the shot's snapshots (nnt slots of 2 nnz x nnx planes)
    long long r_off;      // offset of the shot's traces (time-major, nr per row)
    long long stage_off;  // offset of the shot's staging planes (16-bit snapshots)
    int src;              // source point within the sub-grid (ix + iz * nnx)
    int rz0, rz1;         // rows holding receivers (rz0 > rz1 when there are none)
    long long scale_off;  // first of the shot's slice scales / maxima (2 per slot), or segment scales (S16)
};

// The tile a workgroup works on.  Workgroups are dealt to the 8 XCDs in
// turn (workgroup b to XCD b % 8) and each XCD has its own L2, so this gives
// XCD x the contiguous run of tiles [x q + min(x, r), ...) of the nb = 8 q + r
// in the launch, in order: neighbouring tiles, which read each other's
// halos, then run on one XCD close together in time and share its L2.  A
// bijection of [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile()
{
    const int nb = gridDim.x, b = blockIdx.x, x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + (b >> 3);
}

struct BatchGeo {
    int nnz, npad, cpitch;
    long long nnxz_total;
    float rx2, rz2;  // RN(1/dx2), RN(1/dz2) when qdiv may use them, else 0
};

// Where a forward kernel puts the D snapshot of the level it reads.
struct SnapW {
    float *d;        // FP32 slots (staged == false) or the staging planes
    unsigned *amax;  // per-slice maxima of |D| as float bits (FP16/BF16), or nullptr
    bool staged;     // FP16/BF16: D goes to the staging planes
    // S16: the kernels round D to 16 bits themselves (snap_put_s16)
    unsigned short *h = nullptr;
    float *scale = nullptr;
    bool s16 = false;
};

// Where the adjoint kernels read them.
struct SnapR {
    const float *f;           // FP32 slots
    const unsigned short *h;  // 16-bit slots
    const float *scale;       // per-slice (FP16/BF16) or per-segment (S16) scales of the 16-bit slots
};

// S16 segments: each row z of a component plane nnx wide is cut into the
// 64-point pieces [4 + 64 s, 68 + 64 s), the rows of the kernels' 64-wide
// tiles, so that one wavefront holds a whole segment.  s16_seg is the
// segment of point (x, z) (x >= 4) within its plane, s16_nseg the segments
// of a plane (every row counted; rows outside the interior are not used).
__host__ __device__ __forceinline__ long long s16_nseg(int nnx, int nnz)
{
    return (long long)nnz * ((nnx + 59) >> 6);
}
__device__ __forceinline__ long long s16_seg(int nnx, int x, int z)
{
    return (long long)z * ((nnx + 59) >> 6) + ((x - 4) >> 6);
}
__device__ __forceinline__ long long s16_scale_idx(const ShotB &sh, int nnz, int slot, int c, long long seg)
{
    return sh.scale_off + (2LL * slot + c) * s16_nseg(sh.nnx, nnz) + seg;
}

// S16 scale exponent of a segment whose largest magnitude is m: the segment
// / 2^e lies in [2^14, 2^15); 0 (scale 1) for an all-zero segment
__device__ __forceinline__ int s16_exp(float m)
{
    int e = 0;
    if (m > 0.f) {
        (void)frexpf(m, &e);  // m in [2^(e-1), 2^e)
        e = max(e - 15, -126);
    }
    return e;
}
// v / 2^e rounded to nearest even, clamped to +-32767 (m itself can round to 2^15)
__device__ __forceinline__ unsigned short s16_q(float v, int e)
{
    return (unsigned short)(short)fminf(fmaxf(rintf(ldexpf(v, -e)), -32767.f), 32767.f);
}


// taper of the point (ix, iz) of a sub-grid nnx wide: t0 along x, then along z
// The taper of point (ix, iz) as two loaded factors whose product it is:
// t0[d] for a distance d < npad from an edge, else t0[npad], which holds 1.
// The loads carry no branch and their product is formed where the taper is
// used, so that no load is waited for while the others are in flight.
struct TapF {
    float x, z;
};
__device__ __forceinline__ TapF taper_ld(const float *__restrict__ t0, int npad, int nnx, int nnz, int ix, int iz)
{
    const int dx = ix < npad ? ix : (ix >= nnx - npad ? nnx - 1 - ix : npad);
    const int dz = iz < npad ? iz : (iz >= nnz - npad ? nnz - 1 - iz : npad);
    return {t0[dx], t0[dz]};
}
__device__ __forceinline__ float taper_of(TapF f)
{
    return f.x * f.z;
}

// base[idx] through a 32-bit byte offset from a workgroup-uniform base, the
// form gfx950's global loads and stores take as SGPR base + VGPR offset with
// no 64-bit address arithmetic per lane (the host keeps the model, and so
// every shot plane, under 2^30 points)
template <class T>
__device__ __forceinline__ T ld32(const T *__restrict__ base, unsigned idx)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T *__restrict__ base, unsigned idx, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (unsigned)sizeof(T)) = v;
}
__device__ __forceinline__ TapF taper_ld32(const float *__restrict__ t0, int npad, int nnx, int nnz, int ix, int iz)
{
    const int dx = ix < npad ? ix : (ix >= nnx - npad ? nnx - 1 - ix : npad);
    const int dz = iz < npad ? iz : (iz >= nnz - npad ? nnz - 1 - iz : npad);
    return {ld32(t0, (unsigned)dx), ld32(t0, (unsigned)dz)};
}
// S16: D of the wavefront's 64 points of row z, x = x0 + lane (one
// segment), into slot `slot`: the segment's scales from the wave's largest
// magnitudes, then the 16-bit values of the interior points (on).  Every
// lane of the wave calls this, with tp = tq = 0 where !on; z is the same for
// the whole wave, and a row outside the interior writes nothing.
__device__ __forceinline__ void snap_put_s16(const SnapW &w, const ShotB &sh, int nnz, int slot, int x, int z,
                                             bool on, float tp, float tq)
{
    float mp = on ? fabsf(tp) : 0.f, mq = on ? fabsf(tq) : 0.f;
    for (int o = 32; o > 0; o >>= 1) {
        mp = fmaxf(mp, __shfl_xor(mp, o, 64));
        mq = fmaxf(mq, __shfl_xor(mq, o, 64));
    }
    if (z >= nnz - 4) return;
    const int ep = s16_exp(mp), eq = s16_exp(mq);
    const long long sg = s16_seg(sh.nnx, x, z);
    if ((threadIdx.x & 63) == 0) {
        w.scale[s16_scale_idx(sh, nnz, slot, 0, sg)] = ldexpf(1.f, ep);
        w.scale[s16_scale_idx(sh, nnz, slot, 1, sg)] = ldexpf(1.f, eq);
    }
    if (on) {
        const long long plane = (long long)sh.nnx * nnz;
        unsigned short *o = w.h + sh.snap_off + (long long)slot * 2 * plane;
        const unsigned i = (unsigned)(x + z * sh.nnx);
        st32(o, i, s16_q(tp, ep));
        st32(o + plane, i, s16_q(tq, eq));
    }
}

// the snapshot plane of component c at a slot, and its element i widened
template <int SF>
__device__ __forceinline__ float snap_raw32(const SnapR &r, const ShotB &sh, long long plane, int slot, int c,
                                            unsigned i)
{
    const long long o = sh.snap_off + (long long)slot * 2 * plane + c * plane;
    if (SF == 0) return ld32(r.f + o, i);
    const unsigned short h = ld32(r.h + o, i);
    if (SF == 1) return __half2float(__ushort_as_half(h));
    if (SF == 3) return (float)(short)h;
    return __uint_as_float((unsigned)h << 16);
}
__device__ __forceinline__ void snap_put32(const SnapW &w, const ShotB &sh, long long plane, int slot, unsigned i,
                                           float tp, float tq)
{
    float *sp = w.staged ? w.d + sh.stage_off : w.d + sh.snap_off + (long long)slot * 2 * plane;
    st32(sp, i, tp);
    st32(sp + plane, i, tq);
}

// write D of point i (shot-relative) into its slot or the staging planes
__device__ __forceinline__ void snap_put(const SnapW &w, const ShotB &sh, long long plane, int slot, long long i,
                                         float tp, float tq)
{
    float *sp = w.staged ? w.d + sh.stage_off : w.d + sh.snap_off + (long long)slot * 2 * plane;
    sp[i] = tp;
    sp[i + plane] = tq;
}

// fold the block's largest |D| of both components into the slice maxima (one
// atomic per block and component); every thread of the block must reach this
// call, and red must hold one float2 per wave
__device__ __forceinline__ void snap_amax(const SnapW &w, const ShotB &sh, int slot, float mp, float mq,
                                          float2 *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        mp = fmaxf(mp, __shfl_xor(mp, o, 64));
        mq = fmaxf(mq, __shfl_xor(mq, o, 64));
    }
    const int tid = threadIdx.x + blockDim.x * threadIdx.y, nw = (blockDim.x * blockDim.y) / 64;
    __syncthreads();
    if (tid % 64 == 0) red[tid / 64] = make_float2(mp, mq);
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < nw; ++k) {
            mp = fmaxf(mp, red[k].x);
            mq = fmaxf(mq, red[k].y);
        }
        unsigned *a = w.amax + sh.scale_off + 2 * slot;
        if (mp > 0.f) atomicMax(a, __float_as_uint(mp));
        if (mq > 0.f) atomicMax(a + 1, __float_as_uint(mq));
    }
}

// D component c of point i at a slot, in two parts: the stored value
// (widened to float) and the slot's scale, multiplied by snap_mul where the
// value is used, so the load is not waited for where it is issued
template <int SF>
__device__ __forceinline__ float snap_raw(const SnapR &r, const ShotB &sh, long long plane, int slot, int c,
                                          long long i)
{
    const long long k = sh.snap_off + (long long)slot * 2 * plane + c * plane + i;
    if (SF == 0) return r.f[k];
    if (SF == 1) return __half2float(__ushort_as_half(r.h[k]));
    if (SF == 3) return (float)(short)r.h[k];
    return __uint_as_float((unsigned)r.h[k] << 16);
}
// A scale read through the constant address space: the scales are not
// written while an adjoint kernel runs, so a wave-uniform index becomes one
// scalar load instead of a vector load per lane.
__device__ __forceinline__ float scale_ld(const float *scale, long long k)
{
    return ((const __attribute__((address_space(4))) float *)scale)[k];
}

// (sg: the point's S16 segment, s16_seg, uniform across the wave for a
// scalar load of its scale)
template <int SF>
__device__ __forceinline__ float snap_mul(const SnapR &r, const ShotB &sh, int nnz, int slot, int c, long long sg,
                                          float raw)
{
    if (SF == 0) return raw;
    if (SF == 3) return raw * scale_ld(r.scale, s16_scale_idx(sh, nnz, slot, c, sg));
    return raw * scale_ld(r.scale, sh.scale_off + 2 * slot + c);
}

// D component c of point i (S16 segment sg) at a slot
template <int SF>
__device__ __forceinline__ float snap_get(const SnapR &r, const ShotB &sh, int nnz, int slot, int c, long long i,
                                          long long sg)
{
    const long long plane = (long long)sh.nnx * nnz;
    const long long k = sh.snap_off + (long long)slot * 2 * plane + c * plane + i;
    if (SF == 0) return r.f[k];
    if (SF == 3) return (float)(short)r.h[k] * r.scale[s16_scale_idx(sh, nnz, slot, c, sg)];
    const float s = r.scale[sh.scale_off + 2 * slot + c];
    if (SF == 1) return __half2float(__ushort_as_half(r.h[k])) * s;
    return __uint_as_float((unsigned)r.h[k] << 16) * s;
}

// exponent e of the scale 2^e of a slice whose largest magnitude has float
// bits mbits: the slice / 2^e lies in [2^14, 2^15); 0 for an empty slice
__device__ __forceinline__ int slice_exp(unsigned mbits)
{
    const float m = __uint_as_float(mbits);
    if (!(m > 0.f)) return 0;
    int e;
    (void)frexpf(m, &e);  // m in [2^(e-1), 2^e)
    return max(e - 15, -110);
}

// Round the staged slices of slot `slot` of every shot (blockIdx.y) to 16
// bits (FP16 when bf16 == 0, else BF16; round to nearest even) and record
// their scales.
static __global__ void pack_snap_kernel(unsigned short *__restrict__ snap, float *__restrict__ scale,
                                        const float *__restrict__ stage, const unsigned *__restrict__ amax,
                                        const ShotB *__restrict__ shots, int nnz, int slot, int bf16)
{
    const ShotB sh = shots[blockIdx.y];
    const long long plane = (long long)sh.nnx * nnz;
    const int e0 = slice_exp(amax[sh.scale_off + 2 * slot]), e1 = slice_exp(amax[sh.scale_off + 2 * slot + 1]);
    if (blockIdx.x == 0 && threadIdx.x < 2)
        scale[sh.scale_off + 2 * slot + threadIdx.x] = ldexpf(1.f, threadIdx.x ? e1 : e0);
    const float inv0 = ldexpf(1.f, -e0), inv1 = ldexpf(1.f, -e1);
    unsigned short *out = snap + sh.snap_off + (long long)slot * 2 * plane;
    const float *in = stage + sh.stage_off;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < 2 * plane;
         k += (long long)gridDim.x * blockDim.x) {
        const float v = in[k] * (k < plane ? inv0 : inv1);
        if (bf16) {
            const unsigned u = __float_as_uint(v);
            out[k] = (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        } else {
            out[k] = __half_as_ushort(__float2half_rn(v));
        }
    }
}

// LDS tiles of a sub-grid whose coefficients sit at another pitch/offset
// (pc, qc: the raw p and q of the thread's own point, 0 outside the sub-grid)
template <int MODE>
__device__ __forceinline__ void load_tiles_b(Tiles2 &t, const float *__restrict__ P, const float *__restrict__ Q,
                                             const float *__restrict__ c11, const float *__restrict__ c13,
                                             const float *__restrict__ c33, int nnx, int nnz, int cpitch, int x0,
                                             int z0, float &pc, float &qc)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    pc = 0.f;
    qc = 0.f;
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
    if (ix < nnx && iz < nnz) {
        const long long i = ix + (long long)iz * nnx;
        pc = P[i];
        qc = Q[i];
        if (MODE == 0) {
            t.p[ty][tx + 4] = pc;
            t.q[ty + 4][tx] = qc;
        } else {
            const long long c = ix + (long long)iz * cpitch;
            const float a11 = c11[c], a13 = c13[c], a33 = c33[c];
            t.p[ty][tx + 4] = a11 * pc + a13 * qc;
            t.q[ty + 4][tx] = a13 * pc + a33 * qc;
        }
    } else {
        t.p[ty][tx + 4] = 0.f;
        t.q[ty + 4][tx] = 0.f;
    }
    if (tx < 4) {
        t.p[ty][tx] = pv(x0 - 4 + tx, iz);
        t.p[ty][tx + 68] = pv(x0 + 64 + tx, iz);
    }
    if (ty < 4) {
        t.q[ty][tx] = qv(ix, z0 - 4 + ty);
        t.q[ty + kBY2 + 4][tx] = qv(ix, z0 + kBY2 + ty);
    }
}

// forwardC step of every shot with injectSource, abc and (slot >= 0) the D
// snapshot of the level read (objFuncGradient_f.cu:160-210 per shot).  The
// point's own operands are loaded with the tiles, before the barrier.
static __global__ __launch_bounds__(64 * kBY2) void fwd2db_kernel(
    float *__restrict__ P_n, float *__restrict__ Q_n, const float *__restrict__ P_c, const float *__restrict__ Q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ t0, const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
    float dx2, float dz2, float dt2, float src_val, SnapW sw, int slot)
{
    __shared__ Tiles2 t;
    __shared__ float2 red[kBY2];
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    const bool act = ix < nnx - 4 && iz < nnz - 4;
    const long long i = ix + (long long)iz * nnx;
    float *p_n = P_n + sh.f_off, *q_n = Q_n + sh.f_off;
    float pp = 0.f, qp = 0.f, a11 = 0.f, a13 = 0.f, a33 = 0.f, pc, qc;
    TapF tw{0.f, 0.f};
    if (act) {
        const long long c = sh.c_off + ix + (long long)iz * bg.cpitch;
        pp = p_n[i];
        qp = q_n[i];
        a11 = c11[c];
        a13 = c13[c];
        a33 = c33[c];
        tw = taper_ld(t0, bg.npad, nnx, nnz, ix, iz);
    }
    load_tiles_b<0>(t, P_c + sh.f_off, Q_c + sh.f_off, nullptr, nullptr, nullptr, nnx, nnz, 0, x0, z0, pc, qc);
    __syncthreads();
    float mp = 0.f, mq = 0.f, dp = 0.f, dq = 0.f;
    if (act) {
        const float tap = taper_of(tw);
        const float tp = dxx(t, dx2, bg.rx2), tq = dzz(t, dz2, bg.rz2);
        dp = tp;
        dq = tq;
        if (slot >= 0 && !sw.s16) {
            snap_put(sw, sh, (long long)nnx * nnz, slot, i, tp, tq);
            mp = fabsf(tp);
            mq = fabsf(tq);
        }
        float pn = 2.f * pc + dt2 * (a11 * tp + a13 * tq) - pp * tap;
        float qn = 2.f * qc + dt2 * (a13 * tp + a33 * tq) - qp * tap;
        if (i == sh.src) {
            pn += src_val;
            qn += src_val;
        }
        p_n[i] = pn * tap;
        q_n[i] = qn * tap;
    }
    if (sw.s16 && slot >= 0) snap_put_s16(sw, sh, nnz, slot, ix, iz, act, dp, dq);
    if (sw.amax && slot >= 0) snap_amax(sw, sh, slot, mp, mq, red);
}

// D of the last forward level of every shot into its snapshot slot
static __global__ __launch_bounds__(64 * kBY2) void dsnap2db_kernel(const float *__restrict__ P,
                                                                    const float *__restrict__ Q,
                                                                    const ShotB *__restrict__ shots,
                                                                    const int *__restrict__ tile_shot, BatchGeo bg,
                                                                    float dx2, float dz2, SnapW sw, int slot)
{
    __shared__ Tiles2 t;
    __shared__ float2 red[kBY2];
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    float pc, qc;
    load_tiles_b<0>(t, P + sh.f_off, Q + sh.f_off, nullptr, nullptr, nullptr, nnx, nnz, 0, x0, z0, pc, qc);
    __syncthreads();
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    const bool act = ix < nnx - 4 && iz < nnz - 4;
    float mp = 0.f, mq = 0.f, dp = 0.f, dq = 0.f;
    if (act) {
        const long long i = ix + (long long)iz * nnx;
        dp = dxx(t, dx2, bg.rx2);
        dq = dzz(t, dz2, bg.rz2);
        if (!sw.s16) snap_put(sw, sh, (long long)nnx * nnz, slot, i, dp, dq);
        mp = fabsf(dp);
        mq = fabsf(dq);
    }
    if (sw.s16) snap_put_s16(sw, sh, nnz, slot, ix, iz, act, dp, dq);
    if (sw.amax) snap_amax(sw, sh, slot, mp, mq, red);
}

// backwardDC step of every shot with its residual injection (receiver map),
// abc and gradientCrossCor against the interpolated snapshots
// (objFuncGradient_f.cu:249-300 per shot).  The point's own operands (the
// previous level, the injected residual, the snapshots, the gradient) are
// loaded with the tiles, before the barrier.
template <int SF>
static __global__ __launch_bounds__(64 * kBY2) void bwd2db_kernel(
    float *__restrict__ P_n, float *__restrict__ Q_n, const float *__restrict__ P_c, const float *__restrict__ Q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ t0, const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
    const int *__restrict__ recmap, const float *__restrict__ traces, int row0, float fr, SnapR sr, int slot0,
    float f, float *__restrict__ G11, float *__restrict__ G13, float *__restrict__ G33, float dx2, float dz2,
    float dt2)
{
    __shared__ Tiles2 t;
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * kBY2;
    const int nnx = sh.nnx, nnz = bg.nnz;
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    const bool act = ix < nnx - 4 && iz < nnz - 4;
    const long long i = ix + (long long)iz * nnx, k = sh.f_off + i;
    const long long plane = (long long)nnx * nnz;
    float *p_n = P_n + sh.f_off, *q_n = Q_n + sh.f_off;
    float pp = 0.f, qp = 0.f, inj = 0.f, g11 = 0.f, g13 = 0.f, g33 = 0.f;
    float d0p = 0.f, d1p = 0.f, d0q = 0.f, d1q = 0.f, pc, qc;
    TapF tw{0.f, 0.f};
    int ir = -1;
    if (act) {
        pp = p_n[i];
        qp = q_n[i];
        tw = taper_ld(t0, bg.npad, nnx, nnz, ix, iz);
        if (iz >= sh.rz0 && iz <= sh.rz1) ir = recmap[k];
        d0p = snap_raw<SF>(sr, sh, plane, slot0, 0, i);
        d1p = snap_raw<SF>(sr, sh, plane, slot0 + 1, 0, i);
        d0q = snap_raw<SF>(sr, sh, plane, slot0, 1, i);
        d1q = snap_raw<SF>(sr, sh, plane, slot0 + 1, 1, i);
        g11 = G11[k];
        g13 = G13[k];
        g33 = G33[k];
    }
    load_tiles_b<1>(t, P_c + sh.f_off, Q_c + sh.f_off, c11 + sh.c_off, c13 + sh.c_off, c33 + sh.c_off, nnx, nnz,
                    bg.cpitch, x0, z0, pc, qc);
    if (ir >= 0) {  // depends on recmap: issued after the other loads
        const float *r0 = traces + sh.r_off + (long long)row0 * sh.nr;
        inj = 0.5f * ((1.f - fr) * r0[ir] + fr * r0[ir + sh.nr]);
    }
    __syncthreads();
    if (!act) return;
    const float tap = taper_of(tw);
    const long long sg = s16_seg(nnx, ix, iz);
    d0p = snap_mul<SF>(sr, sh, nnz, slot0, 0, sg, d0p);
    d1p = snap_mul<SF>(sr, sh, nnz, slot0 + 1, 0, sg, d1p);
    d0q = snap_mul<SF>(sr, sh, nnz, slot0, 1, sg, d0q);
    d1q = snap_mul<SF>(sr, sh, nnz, slot0 + 1, 1, sg, d1q);
    const float tp0 = dxx(t, dx2, bg.rx2), tq0 = dzz(t, dz2, bg.rz2);
    float pn = 2.f * pc + dt2 * tp0 - pp * tap;
    float qn = 2.f * qc + dt2 * tq0 - qp * tap;
    if (ir >= 0) {
        pn += dt2 * inj;
        qn += dt2 * inj;
    }
    pn = pn * tap;
    qn = qn * tap;
    p_n[i] = pn;
    q_n[i] = qn;
    const float Dp = (1.f - f) * d0p + f * d1p;
    const float Dq = (1.f - f) * d0q + f * d1q;
    G11[k] = g11 + pn * Dp;
    G13[k] = g13 + (pn * Dq + qn * Dp);
    G33[k] = g33 + qn * Dq;
}

// gradientCrossCor of the first adjoint level of every shot
template <int SF>
static __global__ void xcor2db_kernel(const float *__restrict__ P, const float *__restrict__ Q,
                                      const ShotB *__restrict__ shots, const int *__restrict__ tile_shot, BatchGeo bg,
                                      SnapR sr, int slot0, float f, float *__restrict__ G11, float *__restrict__ G13,
                                      float *__restrict__ G33)
{
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int ix = 4 + (tl % sh.ntx) * 64 + threadIdx.x, iz = 4 + (tl / sh.ntx) * kBY2 + threadIdx.y;
    const int nnx = sh.nnx, nnz = bg.nnz;
    if (ix >= nnx - 4 || iz >= nnz - 4) return;
    const long long i = ix + (long long)iz * nnx, k = sh.f_off + i, sg = s16_seg(nnx, ix, iz);
    const float Dp = (1.f - f) * snap_get<SF>(sr, sh, nnz, slot0, 0, i, sg) +
                     f * snap_get<SF>(sr, sh, nnz, slot0 + 1, 0, i, sg);
    const float Dq = (1.f - f) * snap_get<SF>(sr, sh, nnz, slot0, 1, i, sg) +
                     f * snap_get<SF>(sr, sh, nnz, slot0 + 1, 1, i, sg);
    G11[k] += P[k] * Dp;
    G13[k] += P[k] * Dq + Q[k] * Dp;
    G33[k] += Q[k] * Dq;
}

// ---- Two levels per launch (temporal blocking) -----------------------------
//
// A workgroup of 64 x BY threads owns the tile [x0, x0+64) x [z0, z0+BY) of
// one shot.  It first advances level L+1 on the rectangle
// R1 = [x0-4, x0+68) x [z0-4, z0+BY+4), which holds every point its own step
// to L+2 reads, then takes that step on the tile.  Points of R1 outside the
// shot's interior hold 0 at every level, as in the one-level kernels, whose
// per-point arithmetic this repeats operation for operation: one two-level
// launch gives the same bits as two one-level launches.  Neighbouring
// workgroups each compute the overlap of their rectangles for themselves.
//
// Level L-1 is read from (Pa, Qa) and level L from (Pb, Qb); L+1 goes to
// (Pc, Qc) and L+2 to (Pd, Qd) -- four distinct buffers per field, since
// neighbours still read the halos of L-1 and L during the launch.
//
// HBM per point and two levels: forward 7 reads (two levels, three
// coefficients) + 4 writes; adjoint the same + the snapshot slots + the
// gradient once (3 reads, 3 writes) instead of twice.

template <int BY, bool ADJ>
struct TB2 {
    float a[BY + 8][80];              // level L: p (forward) or c11 p + c13 q (adjoint); x0-8.., z0-4..
    float b[BY + 16][72];             // level L: q or c13 p + c33 q; x0-4.., z0-8..
    float p1[BY + 8][72];             // level L+1 on R1
    float q1[BY + 8][72];
    float pp1[ADJ ? BY + 8 : 1][72];  // adjoint: c11 p + c13 q of level L+1 on R1
    float qq1[ADJ ? BY + 8 : 1][72];  //          c13 p + c33 q
};

// trace rows and snapshot slots bracketing one adjoint level
struct LevelB {
    int row;   // first of the two trace rows
    float fr;  // weight of the second row
    int slot;  // first of the two snapshot slots
    float f;   // weight of the second slot
};

// The tiles' points outside R1: columns x0-8..x0-5 and x0+68..x0+71 of the
// x tile (NA points), rows z0-8..z0-5 and z0+BY+4..z0+BY+7 of the z tile
// (NB points).  Each thread loads at most NH of them into registers with the
// rest of its operands, before any LDS write, so that the workgroup waits for
// one memory latency rather than one per load loop.  The R1 part of the tiles
// is written from the operands the threads already hold.
template <int BY>
struct HaloT2 {
    // tile a's points take the first NAW slots (whole wavefronts), tile b's
    // the next NB, so that which tile a wavefront serves is uniform across it
    static constexpr int NT = 64 * BY, NA = 8 * (BY + 8), NAW = (NA + 63) / 64 * 64, NB = 8 * 72,
                         NH = (NAW + NB + NT - 1) / NT;
    // LDS position (tile a or b, row, column) and grid point of halo slot h;
    // false for a slot that holds no point (the outputs then name point 0)
    __device__ static __forceinline__ bool at(int h, int x0, int z0, bool &in_a, int &r, int &c, int &x, int &z)
    {
        const bool ok = h < NA || (h >= NAW && h < NAW + NB);
        h = ok ? (h < NA ? h : h - NAW + NA) : 0;
        if (h < NA) {
            in_a = true;
            r = h >> 3;
            c = h & 7;
            c = c < 4 ? c : c + 72;
            x = x0 - 8 + c;
            z = z0 - 4 + r;
        } else {
            h -= NA;
            in_a = false;
            r = h / 72;
            c = h - r * 72;
            r = r < 4 ? r : r + BY + 8;
            x = x0 - 4 + c;
            z = z0 - 8 + r;
        }
        return ok;
    }
};

// The loads behind one halo point, made at the point clamped into the grid:
// p (tile a) or q (tile b) in p; for the adjoint p, q and the two
// coefficients that weight them (c11, c13 in tile a; c13, c33 in tile b).
// The tile value is formed at the LDS write, 0 outside the grid.
template <bool ADJ>
struct HaloV {
    float p, q, u, w;
};

template <int BY, bool ADJ, bool IN>
__device__ __forceinline__ void halo_load_t2(HaloV<ADJ> (&hv)[HaloT2<BY>::NH], const float *__restrict__ P,
                                             const float *__restrict__ Q, const float *__restrict__ c11,
                                             const float *__restrict__ c13, const float *__restrict__ c33, int nnx,
                                             int nnz, int cpitch, int x0, int z0, int tid)
{
#pragma unroll
    for (int j = 0; j < HaloT2<BY>::NH; ++j) {
        bool in_a;
        int r, c, x, z;
        (void)HaloT2<BY>::at(tid + j * HaloT2<BY>::NT, x0, z0, in_a, r, c, x, z);
        if (!IN) {
            x = min(max(x, 0), nnx - 1);
            z = min(max(z, 0), nnz - 1);
        }
        const unsigned i = x + z * nnx;
        // the wavefront's tile, uniform across it by the slot layout
        const bool wa = __builtin_amdgcn_readfirstlane((tid + j * HaloT2<BY>::NT) < HaloT2<BY>::NAW);
        if (!ADJ) {
            hv[j].p = ld32(wa ? P : Q, i);
        } else {
            const unsigned cc = x + z * cpitch;
            hv[j].p = ld32(P, i);
            hv[j].q = ld32(Q, i);
            hv[j].u = ld32(wa ? c11 : c13, cc);
            hv[j].w = ld32(wa ? c13 : c33, cc);
        }
    }
}

template <int BY, bool ADJ, bool IN>
__device__ __forceinline__ void halo_store_t2(TB2<BY, ADJ> &t, const HaloV<ADJ> (&hv)[HaloT2<BY>::NH], int nnx,
                                              int nnz, int x0, int z0, int tid)
{
#pragma unroll
    for (int j = 0; j < HaloT2<BY>::NH; ++j) {
        bool in_a;
        int r, c, x, z;
        if (HaloT2<BY>::at(tid + j * HaloT2<BY>::NT, x0, z0, in_a, r, c, x, z)) {
            // x tile: p (adjoint c11 p + c13 q); z tile: q (adjoint c13 p + c33 q)
            float v = ADJ ? hv[j].u * hv[j].p + hv[j].w * hv[j].q : hv[j].p;
            v = (IN || (x >= 0 && x < nnx && z >= 0 && z < nnz)) ? v : 0.f;
            if (in_a)
                t.a[r][c] = v;
            else
                t.b[r][c] = v;
        }
    }
}

// the reference's 9-point second difference along an LDS row (x) and down an
// LDS column (z), operands in the order of dxx/dzz
__device__ __forceinline__ float st9_row(const float *a, int c)
{
    return st9(a[c], a[c - 1], a[c + 1], a[c - 2], a[c + 2], a[c - 3], a[c + 3], a[c - 4], a[c + 4]);
}
template <int W>
__device__ __forceinline__ float st9_col(const float (*a)[W], int r, int c)
{
    return st9(a[r][c], a[r - 1][c], a[r + 1][c], a[r - 2][c], a[r + 2][c], a[r - 3][c], a[r + 3][c], a[r - 4][c],
               a[r + 4][c]);
}

// Whether a tile's R1 (and so its tile and halo) lies where the taper is 1:
// every point then has the taper factor t0[npad] = 1 in both directions, is
// inside the grid and in the interior, so the clamps, the bounds tests and
// the taper loads and products drop out (a product by 1 is exact).
template <int BY>
__device__ __forceinline__ bool t2_inner(int x0, int z0, int nnx, int nnz, int npad)
{
    return x0 - 4 >= npad && x0 + 68 <= nnx - npad && z0 - 4 >= npad && z0 + BY + 4 <= nnz - npad;
}

// Forward: two forwardC steps with injectSource, abc and the D snapshots of
// the two levels read (slot1 for level L, slot2 for level L+1; < 0: none).
// IN: the tile is inner (t2_inner).
template <int BY, bool IN>
__device__ __forceinline__ void fwd2db_t2_body(
    TB2<BY, false> &t, float2 *red, const ShotB &sh, int x0, int z0, float *__restrict__ Pc,
    float *__restrict__ Qc, float *__restrict__ Pd, float *__restrict__ Qd, const float *__restrict__ Pa,
    const float *__restrict__ Qa, const float *__restrict__ Pb, const float *__restrict__ Qb,
    const float *__restrict__ cc11, const float *__restrict__ cc13, const float *__restrict__ cc33,
    const float *__restrict__ t0, const BatchGeo &bg, float dx2, float dz2, float dt2, float src1, float src2,
    const SnapW &sw, int slot1, int slot2)
{
    constexpr int NT = 64 * BY, N1 = (BY + 8) * 72, NK = (N1 + NT - 1) / NT;
    const int nnx = sh.nnx, nnz = bg.nnz, tid = threadIdx.x + 64 * threadIdx.y;
    const long long plane = (long long)nnx * nnz;
    // operands of this thread's points of R1 (levels L and L-1, coefficients,
    // taper) and of its own point of the tile.  Every load is made, at the
    // point clamped into the grid, without a branch: values of points outside
    // the grid (or the interior) are replaced or left unused afterwards.
    float pb[NK], qb[NK], pa[NK], qa[NK], k11[NK], k13[NK], k33[NK];
    TapF tk[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = min(tid + j * NT, N1 - 1);
        int x = x0 - 4 + k % 72, z = z0 - 4 + k / 72;
        if (!IN) {
            x = min(x, nnx - 1);
            z = min(z, nnz - 1);
        }
        const unsigned i = x + z * nnx, c = x + z * bg.cpitch;
        pb[j] = ld32(Pb, i);
        qb[j] = ld32(Qb, i);
        pa[j] = ld32(Pa, i);
        qa[j] = ld32(Qa, i);
        k11[j] = ld32(cc11, c);
        k13[j] = ld32(cc13, c);
        k33[j] = ld32(cc33, c);
        tk[j] = IN ? TapF{1.f, 1.f} : taper_ld32(t0, bg.npad, nnx, nnz, x, z);
    }
    HaloV<false> hv[HaloT2<BY>::NH];
    halo_load_t2<BY, false, IN>(hv, Pb, Qb, nullptr, nullptr, nullptr, nnx, nnz, 0, x0, z0, tid);
    const int x = x0 + threadIdx.x, z = z0 + threadIdx.y;
    const bool act = IN || (x < nnx - 4 && z < nnz - 4);
    float a11, a13, a33;
    TapF tw;
    {
        const int xo = IN ? x : min(x, nnx - 1), zo = IN ? z : min(z, nnz - 1);
        const unsigned c = xo + zo * bg.cpitch;
        a11 = ld32(cc11, c);
        a13 = ld32(cc13, c);
        a33 = ld32(cc33, c);
        tw = IN ? TapF{1.f, 1.f} : taper_ld32(t0, bg.npad, nnx, nnz, xo, zo);
    }
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = tid + j * NT;
        if (k < N1) {
            const int rx = k % 72, rz = k / 72;
            const bool in = IN || (x0 - 4 + rx < nnx && z0 - 4 + rz < nnz);
            t.a[rz][rx + 4] = in ? pb[j] : 0.f;
            t.b[rz + 4][rx] = in ? qb[j] : 0.f;
        }
    }
    halo_store_t2<BY, false, IN>(t, hv, nnx, nnz, x0, z0, tid);
    __syncthreads();
    float mp = 0.f, mq = 0.f;
    // level L+1 on R1
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = tid + j * NT;
        if (k < N1) {
            const int rx = k % 72, rz = k / 72;
            const int xx = x0 - 4 + rx, zz = z0 - 4 + rz;
            float p1 = 0.f, q1 = 0.f;
            if (IN || (xx >= 4 && xx < nnx - 4 && zz >= 4 && zz < nnz - 4)) {
                const int i = xx + zz * nnx;
                const float tp = qdiv(st9_row(t.a[rz], rx + 4), dx2, bg.rx2);
                const float tq = qdiv(st9_col<72>(t.b, rz + 4, rx), dz2, bg.rz2);
                if (slot1 >= 0 && !sw.s16 && rx >= 4 && rx < 68 && rz >= 4 && rz < BY + 4) {
                    snap_put32(sw, sh, plane, slot1, i, tp, tq);
                    mp = fmaxf(mp, fabsf(tp));
                    mq = fmaxf(mq, fabsf(tq));
                }
                const float tkj = taper_of(tk[j]);
                float pn = 2.f * pb[j] + dt2 * (k11[j] * tp + k13[j] * tq) - pa[j] * tkj;
                float qn = 2.f * qb[j] + dt2 * (k13[j] * tp + k33[j] * tq) - qa[j] * tkj;
                if (i == sh.src) {
                    pn += src1;
                    qn += src1;
                }
                p1 = pn * tkj;
                q1 = qn * tkj;
            }
            t.p1[rz][rx] = p1;
            t.q1[rz][rx] = q1;
        }
    }
    if (sw.amax && slot1 >= 0) snap_amax(sw, sh, slot1, mp, mq, red);
    lds_barrier();  // level L+1 reaches level L+2 through LDS only
    // level L+2 on the tile
    mp = 0.f;
    mq = 0.f;
    float d2p = 0.f, d2q = 0.f;
    if (act) {
        const float tap = taper_of(tw);
        const int rx = threadIdx.x + 4, rz = threadIdx.y + 4;
        const int i = x + z * nnx;
        const float tp = qdiv(st9_row(t.p1[rz], rx), dx2, bg.rx2);
        const float tq = qdiv(st9_col<72>(t.q1, rz, rx), dz2, bg.rz2);
        d2p = tp;
        d2q = tq;
        if (slot2 >= 0 && !sw.s16) {
            snap_put32(sw, sh, plane, slot2, i, tp, tq);
            mp = fabsf(tp);
            mq = fabsf(tq);
        }
        const float p1 = t.p1[rz][rx], q1 = t.q1[rz][rx];
        float pn = 2.f * p1 + dt2 * (a11 * tp + a13 * tq) - t.a[rz][rx + 4] * tap;
        float qn = 2.f * q1 + dt2 * (a13 * tp + a33 * tq) - t.b[rz + 4][rx] * tap;
        if (i == sh.src) {
            pn += src2;
            qn += src2;
        }
        st32(Pc, i, p1);
        st32(Qc, i, q1);
        st32(Pd, i, pn * tap);
        st32(Qd, i, qn * tap);
    }
    if (sw.s16) {
        // S16 by tile rows, one wavefront each: D of level L at the tile
        // point again from the level-L planes in LDS (the same operations as
        // on R1 above, so the same bits), and D of level L+1 from above
        if (slot1 >= 0) {
            float dp = 0.f, dq = 0.f;
            if (act) {
                dp = qdiv(st9_row(t.a[threadIdx.y + 4], threadIdx.x + 8), dx2, bg.rx2);
                dq = qdiv(st9_col<72>(t.b, threadIdx.y + 8, threadIdx.x + 4), dz2, bg.rz2);
            }
            snap_put_s16(sw, sh, nnz, slot1, x, z, act, dp, dq);
        }
        if (slot2 >= 0) snap_put_s16(sw, sh, nnz, slot2, x, z, act, d2p, d2q);
    }
    if (sw.amax && slot2 >= 0) snap_amax(sw, sh, slot2, mp, mq, red);
}

// A 16-row workgroup is 16 waves; two of them per CU need 8 waves per SIMD,
// which gfx950 grants only to kernels of at most 80 SGPRs.  The inner and
// edge bodies together would take about 93, so the occupancy is pinned and
// the compiler keeps a few uniform values in VGPR lanes instead.
template <int BY>
static __global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(BY == 16 ? 8 : 1, BY == 16 ? 8 : 10))) void fwd2db_t2_kernel(
    float *__restrict__ Pc, float *__restrict__ Qc, float *__restrict__ Pd, float *__restrict__ Qd,
    const float *__restrict__ Pa, const float *__restrict__ Qa, const float *__restrict__ Pb,
    const float *__restrict__ Qb, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ t0, const ShotB *__restrict__ shots,
    const int *__restrict__ tile_shot, BatchGeo bg, float dx2, float dz2, float dt2, float src1, float src2,
    SnapW sw, int slot1, int slot2)
{
    __shared__ TB2<BY, false> t;
    __shared__ float2 red[BY];
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * BY;
    const long long fo = sh.f_off, co = sh.c_off;
    if (t2_inner<BY>(x0, z0, sh.nnx, bg.nnz, bg.npad))
        fwd2db_t2_body<BY, true>(t, red, sh, x0, z0, Pc + fo, Qc + fo, Pd + fo, Qd + fo, Pa + fo, Qa + fo, Pb + fo,
                                 Qb + fo, c11 + co, c13 + co, c33 + co, t0, bg, dx2, dz2, dt2, src1, src2, sw, slot1,
                                 slot2);
    else
        fwd2db_t2_body<BY, false>(t, red, sh, x0, z0, Pc + fo, Qc + fo, Pd + fo, Qd + fo, Pa + fo, Qa + fo, Pb + fo,
                                  Qb + fo, c11 + co, c13 + co, c33 + co, t0, bg, dx2, dz2, dt2, src1, src2, sw, slot1,
                                  slot2);
}

// Adjoint: two backwardDC steps with the residual injection, abc and
// gradientCrossCor of both levels, the gradient read and written once.
// IN: the tile is inner (t2_inner).
template <int BY, int SF, bool IN>
__device__ __forceinline__ void bwd2db_t2_body(
    TB2<BY, true> &t, const ShotB &sh, int x0, int z0, float *__restrict__ Pc, float *__restrict__ Qc,
    float *__restrict__ Pd, float *__restrict__ Qd, const float *__restrict__ Pa, const float *__restrict__ Qa,
    const float *__restrict__ Pb, const float *__restrict__ Qb, const float *__restrict__ cc11,
    const float *__restrict__ cc13, const float *__restrict__ cc33, const float *__restrict__ t0, const BatchGeo &bg,
    const int *__restrict__ rmap, const float *__restrict__ traces, const LevelB &l1, const LevelB &l2,
    const SnapR &sr, float *__restrict__ g11p, float *__restrict__ g13p, float *__restrict__ g33p, float dx2,
    float dz2, float dt2)
{
    constexpr int NT = 64 * BY, N1 = (BY + 8) * 72, NK = (N1 + NT - 1) / NT;
    const int nnx = sh.nnx, nnz = bg.nnz, tid = threadIdx.x + 64 * threadIdx.y;
    const long long plane = (long long)nnx * nnz;
    // operands of this thread's points of R1 (levels L and L-1, coefficients,
    // taper, receiver index) and of its own point of the tile (level L, taper,
    // receiver index, the snapshots of both levels, the gradient).  Every load
    // is made, at the point clamped into the grid, without a branch: values of
    // points outside the grid (or the interior) are replaced or left unused
    // afterwards.  Receiver indices are read only by workgroups whose rows
    // meet the receiver rows.
    const bool rrows = z0 - 4 <= sh.rz1 && z0 + BY + 4 > sh.rz0;
    float pb[NK], qb[NK], pa[NK], qa[NK], k11[NK], k13[NK], k33[NK], inj[NK];
    TapF tk[NK];
    int irj[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = min(tid + j * NT, N1 - 1);
        int x = x0 - 4 + k % 72, z = z0 - 4 + k / 72;
        if (!IN) {
            x = min(x, nnx - 1);
            z = min(z, nnz - 1);
        }
        const unsigned i = x + z * nnx, c = x + z * bg.cpitch;
        pb[j] = ld32(Pb, i);
        qb[j] = ld32(Qb, i);
        pa[j] = ld32(Pa, i);
        qa[j] = ld32(Qa, i);
        k11[j] = ld32(cc11, c);
        k13[j] = ld32(cc13, c);
        k33[j] = ld32(cc33, c);
        tk[j] = IN ? TapF{1.f, 1.f} : taper_ld32(t0, bg.npad, nnx, nnz, x, z);
        irj[j] = -1;
        inj[j] = 0.f;
    }
    const int x = x0 + threadIdx.x, z = z0 + threadIdx.y;
    const bool act = IN || (x < nnx - 4 && z < nnz - 4);
    const int xo = IN ? x : min(x, nnx - 1), zo = IN ? z : min(z, nnz - 1);
    const unsigned i = xo + zo * nnx;
    const float pbc = ld32(Pb, i), qbc = ld32(Qb, i);
    const TapF tw = IN ? TapF{1.f, 1.f} : taper_ld32(t0, bg.npad, nnx, nnz, xo, zo);
    HaloV<true> hv[HaloT2<BY>::NH];
    halo_load_t2<BY, true, IN>(hv, Pb, Qb, cc11, cc13, cc33, nnx, nnz, bg.cpitch, x0, z0, tid);
    // the injected residual: recmap, then the trace reads that depend on it
    float inj2 = 0.f;
    int ir2 = -1;
    if (rrows) {
        int rraw[NK];
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            const int k = min(tid + j * NT, N1 - 1);
            const int xr = min(x0 - 4 + k % 72, nnx - 1), zr = min(z0 - 4 + k / 72, nnz - 1);
            rraw[j] = ld32(rmap, xr + zr * nnx);
        }
        const int r2raw = ld32(rmap, i);
        const float *tr1 = traces + sh.r_off + (long long)l1.row * sh.nr;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            const int k = tid + j * NT;
            const int xr = x0 - 4 + k % 72, zr = z0 - 4 + k / 72;
            const bool in = k < N1 && xr >= 4 && xr < nnx - 4 && zr >= 4 && zr < nnz - 4 && zr >= sh.rz0 && zr <= sh.rz1;
            irj[j] = in ? rraw[j] : -1;
            if (irj[j] >= 0)
                inj[j] = 0.5f * ((1.f - l1.fr) * ld32(tr1, irj[j]) + l1.fr * ld32(tr1, irj[j] + sh.nr));
        }
        ir2 = (act && z >= sh.rz0 && z <= sh.rz1) ? r2raw : -1;
        if (ir2 >= 0) {
            const float *tr2 = traces + sh.r_off + (long long)l2.row * sh.nr;
            inj2 = 0.5f * ((1.f - l2.fr) * ld32(tr2, ir2) + l2.fr * ld32(tr2, ir2 + sh.nr));
        }
    }
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = tid + j * NT;
        if (k < N1) {
            const int rx = k % 72, rz = k / 72;
            const bool in = IN || (x0 - 4 + rx < nnx && z0 - 4 + rz < nnz);
            t.a[rz][rx + 4] = in ? k11[j] * pb[j] + k13[j] * qb[j] : 0.f;
            t.b[rz + 4][rx] = in ? k13[j] * pb[j] + k33[j] * qb[j] : 0.f;
        }
    }
    halo_store_t2<BY, true, IN>(t, hv, nnx, nnz, x0, z0, tid);
    __syncthreads();
    // the snapshots of both levels and the gradient at this thread's point
    // (clamped into the grid), issued here so that they arrive while level
    // L+1 is computed on R1.  Both levels usually interpolate between the
    // same two slots; the second level's are then not read again.
    float a0p = snap_raw32<SF>(sr, sh, plane, l1.slot, 0, i);
    float a1p = snap_raw32<SF>(sr, sh, plane, l1.slot + 1, 0, i);
    float a0q = snap_raw32<SF>(sr, sh, plane, l1.slot, 1, i);
    float a1q = snap_raw32<SF>(sr, sh, plane, l1.slot + 1, 1, i);
    // (no copies here: a copy would wait for the loads)
    const bool same = l2.slot == l1.slot;
    float b0p = 0.f, b1p = 0.f, b0q = 0.f, b1q = 0.f;
    if (!same) {
        b0p = snap_raw32<SF>(sr, sh, plane, l2.slot, 0, i);
        b1p = snap_raw32<SF>(sr, sh, plane, l2.slot + 1, 0, i);
        b0q = snap_raw32<SF>(sr, sh, plane, l2.slot, 1, i);
        b1q = snap_raw32<SF>(sr, sh, plane, l2.slot + 1, 1, i);
    }
    float g11 = ld32(g11p, i), g13 = ld32(g13p, i), g33 = ld32(g33p, i);
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int k = tid + j * NT;
        if (k < N1) {
            const int rx = k % 72, rz = k / 72;
            const int xx = x0 - 4 + rx, zz = z0 - 4 + rz;
            float p1 = 0.f, q1 = 0.f, pp = 0.f, qq = 0.f;
            if (IN || (xx >= 4 && xx < nnx - 4 && zz >= 4 && zz < nnz - 4)) {
                const float tp0 = qdiv(st9_row(t.a[rz], rx + 4), dx2, bg.rx2);
                const float tq0 = qdiv(st9_col<72>(t.b, rz + 4, rx), dz2, bg.rz2);
                const float tkj = taper_of(tk[j]);
                float pn = 2.f * pb[j] + dt2 * tp0 - pa[j] * tkj;
                float qn = 2.f * qb[j] + dt2 * tq0 - qa[j] * tkj;
                if (irj[j] >= 0) {
                    pn += dt2 * inj[j];
                    qn += dt2 * inj[j];
                }
                p1 = pn * tkj;
                q1 = qn * tkj;
                pp = k11[j] * p1 + k13[j] * q1;
                qq = k13[j] * p1 + k33[j] * q1;
            }
            t.p1[rz][rx] = p1;
            t.q1[rz][rx] = q1;
            t.pp1[rz][rx] = pp;
            t.qq1[rz][rx] = qq;
        }
    }
    // level L+1 reaches level L+2 through LDS only: the snapshot and gradient
    // loads stay in flight across this barrier
    lds_barrier();
    if (!act) return;
    const float tap = taper_of(tw);
    const int rx = threadIdx.x + 4, rz = threadIdx.y + 4;
    const float tp0 = qdiv(st9_row(t.pp1[rz], rx), dx2, bg.rx2);
    const float tq0 = qdiv(st9_col<72>(t.qq1, rz, rx), dz2, bg.rz2);
    const float p1 = t.p1[rz][rx], q1 = t.q1[rz][rx];
    float pn = 2.f * p1 + dt2 * tp0 - pbc * tap;
    float qn = 2.f * q1 + dt2 * tq0 - qbc * tap;
    if (ir2 >= 0) {
        pn += dt2 * inj2;
        qn += dt2 * inj2;
    }
    pn = pn * tap;
    qn = qn * tap;
    st32(Pc, i, p1);
    st32(Qc, i, q1);
    st32(Pd, i, pn);
    st32(Qd, i, qn);
    // the segment of the wavefront's row (one segment: x0 - 4 is a multiple
    // of 64), uniform across the wave
    const long long sg = __builtin_amdgcn_readfirstlane((int)s16_seg(nnx, x0, zo));
    a0p = snap_mul<SF>(sr, sh, nnz, l1.slot, 0, sg, a0p);
    a1p = snap_mul<SF>(sr, sh, nnz, l1.slot + 1, 0, sg, a1p);
    a0q = snap_mul<SF>(sr, sh, nnz, l1.slot, 1, sg, a0q);
    a1q = snap_mul<SF>(sr, sh, nnz, l1.slot + 1, 1, sg, a1q);
    if (same) {  // the same slots, already scaled
        b0p = a0p;
        b1p = a1p;
        b0q = a0q;
        b1q = a1q;
    } else {
        b0p = snap_mul<SF>(sr, sh, nnz, l2.slot, 0, sg, b0p);
        b1p = snap_mul<SF>(sr, sh, nnz, l2.slot + 1, 0, sg, b1p);
        b0q = snap_mul<SF>(sr, sh, nnz, l2.slot, 1, sg, b0q);
        b1q = snap_mul<SF>(sr, sh, nnz, l2.slot + 1, 1, sg, b1q);
    }
    const float Dp1 = (1.f - l1.f) * a0p + l1.f * a1p;
    const float Dq1 = (1.f - l1.f) * a0q + l1.f * a1q;
    const float Dp2 = (1.f - l2.f) * b0p + l2.f * b1p;
    const float Dq2 = (1.f - l2.f) * b0q + l2.f * b1q;
    g11 += p1 * Dp1;
    g13 += p1 * Dq1 + q1 * Dp1;
    g33 += q1 * Dq1;
    g11 += pn * Dp2;
    g13 += pn * Dq2 + qn * Dp2;
    g33 += qn * Dq2;
    st32(g11p, i, g11);
    st32(g13p, i, g13);
    st32(g33p, i, g33);
}

// (16 rows: 8 waves per SIMD pinned, as for the forward)
template <int BY, int SF>
static __global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(BY == 16 ? 8 : 1, BY == 16 ? 8 : 10))) void bwd2db_t2_kernel(
    float *__restrict__ Pc, float *__restrict__ Qc, float *__restrict__ Pd, float *__restrict__ Qd,
    const float *__restrict__ Pa, const float *__restrict__ Qa, const float *__restrict__ Pb,
    const float *__restrict__ Qb, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ t0, const ShotB *__restrict__ shots,
    const int *__restrict__ tile_shot, BatchGeo bg, const int *__restrict__ recmap, const float *__restrict__ traces,
    LevelB l1, LevelB l2, SnapR sr, float *__restrict__ G11, float *__restrict__ G13, float *__restrict__ G33,
    float dx2, float dz2, float dt2)
{
    __shared__ TB2<BY, true> t;
    const int tile = xcd_tile();
    const ShotB sh = shots[tile_shot[tile]];
    const int tl = tile - sh.tile0;
    const int x0 = 4 + (tl % sh.ntx) * 64, z0 = 4 + (tl / sh.ntx) * BY;
    const long long fo = sh.f_off, co = sh.c_off;
    if (t2_inner<BY>(x0, z0, sh.nnx, bg.nnz, bg.npad))
        bwd2db_t2_body<BY, SF, true>(t, sh, x0, z0, Pc + fo, Qc + fo, Pd + fo, Qd + fo, Pa + fo, Qa + fo, Pb + fo,
                                     Qb + fo, c11 + co, c13 + co, c33 + co, t0, bg, recmap + fo, traces, l1, l2, sr,
                                     G11 + fo, G13 + fo, G33 + fo, dx2, dz2, dt2);
    else
        bwd2db_t2_body<BY, SF, false>(t, sh, x0, z0, Pc + fo, Qc + fo, Pd + fo, Qd + fo, Pa + fo, Qa + fo, Pb + fo,
                                      Qb + fo, c11 + co, c13 + co, c33 + co, t0, bg, recmap + fo, traces, l1, l2, sr,
                                      G11 + fo, G13 + fo, G33 + fo, dx2, dz2, dt2);
}

}  // namespace vti
