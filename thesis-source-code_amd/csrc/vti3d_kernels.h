// 3D VTI kernels for gfx950 (header-only templates; included by vti3d.hip).
//
// Device layout: the reference's x-fastest order (j = ix + iy*nx + iz*nx*ny,
// lib/vti-fwi-3d/gpu/objFuncGradient3d_f.cu:560-567) with the row pitch
// padded to a multiple of 64 floats, so every 64-lane wavefront row load is
// one aligned 256-byte segment.  Interior x,y: [4, n-4); z: every plane of the
// first 4*(nz/4), with zero ghosts outside, as the reference's top/bottom
// block kernels use (kernel3d.cu:76-220).
//
// Blocking (2.5D): a workgroup owns a 64 x BY tile of (x, y) columns and
// marches through a z-chunk.  sigma_x needs x and y derivatives: each plane of
// the tile plus a 4-point halo lives in LDS (double buffered, one barrier per
// plane).  sigma_z needs only the z derivative: each lane keeps a 9-deep
// register queue of its own column.
//
// Arithmetic (template parameter ARITH): kArithRef evaluates the reference's
// own expressions -- float pair sums, double weights, products and sums,
// rounded to float once per stencil and once per update (kernel3d.cu:54-69)
// -- and is the default; kArithF32 evaluates the same expressions in FP32.
// Contraction is off in both, so each mode is reproducible bit for bit by the
// oracle's matching mode.
#pragma once

#include "vti_internal.h"

#pragma clang fp contract(off)

namespace vti {

struct Geo3 {
    int nx, ny, nz;   // logical dims (reference layout)
    int zmax;         // 4*(nz/4): planes the reference updates
    int pitch;        // device row pitch in floats (multiple of 64)
    long long plane;  // pitch*ny
};

// Maps the hardware block order (round-robin over the 8 XCDs) to a logical
// tile order in which each XCD owns a contiguous range of tiles, so halo
// re-reads between neighbouring tiles hit the same L2.
__device__ __forceinline__ int xcd_logical_block(int nblocks)
{
    const int i = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int q = nblocks >> 3, r = nblocks & 7;
    const int xcd = i & 7, pos = i >> 3;
    return (xcd < r) ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// Raw buffer access (buffer_load/store_dword): a wave-uniform 128-bit
// resource in SGPRs, a 32-bit per-lane byte offset and a wave-uniform byte
// offset, so a plane's loads cost one VGPR of address per distinct lane
// offset instead of a 64-bit address pair.  Every access is also guarded by
// the kernels' own predicates; the resource's range is not relied on.
using rsrc_t = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const float *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)0xFFFFFFFF, 0x00020000);
}

__device__ __forceinline__ float bload(rsrc_t r, unsigned lane_byte, unsigned uni_byte)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_byte, (int)uni_byte, 0));
}

__device__ __forceinline__ void bstore(rsrc_t r, unsigned lane_byte, unsigned uni_byte, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)lane_byte, (int)uni_byte, 0);
}

// The same with the non-temporal cache policy (the SLC/NT bit, 2), for the
// streams a step touches once per point -- the previous level it overwrites,
// the coefficients, the accumulators -- so that they do not displace the
// sigma planes whose halo rows neighbouring workgroups read again from L2.
constexpr int kNonTemporal = 2;
__device__ __forceinline__ float bload_nt(rsrc_t r, unsigned lane_byte, unsigned uni_byte)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_byte, (int)uni_byte, kNonTemporal));
}
__device__ __forceinline__ void bstore_nt(rsrc_t r, unsigned lane_byte, unsigned uni_byte, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)lane_byte, (int)uni_byte, kNonTemporal);
}

template <int BY>
struct Tile {
    // LDS row width: 64 columns + 4 halo on each side
    static constexpr int W = 72;
    static constexpr int R = BY + 8;
    static constexpr int NYH = (512 + 64 * BY - 1) / (64 * BY);  // y-halo elements per thread (8 rows x 64)
};

// Per-thread description of what it loads into the x/y tile each plane.
template <int BY>
struct TileLoader {
    unsigned main_off;  // offset of own point within a plane (unsigned: zero-extended address offsets)
    bool main_ok;
    int main_lds;
    unsigned yh_off[Tile<BY>::NYH];
    bool yh_has[Tile<BY>::NYH];
    bool yh_ok[Tile<BY>::NYH];
    int yh_lds[Tile<BY>::NYH];
    unsigned xh_off;
    bool xh_ok;
    int xh_lds;
    bool has_xh;

    __device__ __forceinline__ void init(int tx, int ty, int x0, int y0, const Geo3 &g)
    {
        constexpr int W = Tile<BY>::W;
        const int tid = tx + 64 * ty;
        const int ix = x0 + tx, iy = y0 + ty;
        main_ok = ix < g.nx && iy < g.ny;
        main_off = (unsigned)(iy * g.pitch + ix);
        main_lds = (ty + 4) * W + tx + 4;
#pragma unroll
        for (int h = 0; h < Tile<BY>::NYH; ++h) {
            const int e = tid + h * 64 * BY;  // e < 512: one of the 8 halo rows x 64 columns
            const int r = (e >> 6) & 7, c = e & 63;
            const int hy = r < 4 ? y0 - 4 + r : y0 + BY + r - 4;
            const int hx = x0 + c;
            yh_has[h] = e < 512;
            yh_ok[h] = e < 512 && hy < g.ny && hx < g.nx;
            yh_off[h] = (unsigned)(hy * g.pitch + hx);
            yh_lds[h] = (r < 4 ? r : r + BY) * W + c + 4;
        }
        has_xh = tid < 8 * BY;
        {
            const int r = tid >> 3, c = tid & 7;
            const int hx = c < 4 ? x0 - 4 + c : x0 + 64 + c - 4;
            const int hy = y0 + r;
            xh_ok = has_xh && hx >= 0 && hx < g.nx && hy < g.ny;
            xh_off = (unsigned)(hy * g.pitch + hx);
            xh_lds = (r + 4) * W + (c < 4 ? c : 64 + c);
        }
    }
};

template <int BY>
struct PlaneRegs {
    float m;
    float yh[Tile<BY>::NYH];
    float xh;

    __device__ __forceinline__ void load(const float *__restrict__ p, const TileLoader<BY> &L)
    {
        m = L.main_ok ? p[L.main_off] : 0.f;
#pragma unroll
        for (int h = 0; h < Tile<BY>::NYH; ++h) yh[h] = L.yh_ok[h] ? p[L.yh_off[h]] : 0.f;
        xh = L.xh_ok ? p[L.xh_off] : 0.f;
    }
    __device__ __forceinline__ void loadb(rsrc_t r, unsigned uni_byte, const TileLoader<BY> &L)
    {
        m = L.main_ok ? bload(r, 4u * L.main_off, uni_byte) : 0.f;
#pragma unroll
        for (int h = 0; h < Tile<BY>::NYH; ++h) yh[h] = L.yh_ok[h] ? bload(r, 4u * L.yh_off[h], uni_byte) : 0.f;
        xh = L.xh_ok ? bload(r, 4u * L.xh_off, uni_byte) : 0.f;
    }
    __device__ __forceinline__ void store(float *lds, const TileLoader<BY> &L) const
    {
        lds[L.main_lds] = m;
#pragma unroll
        for (int h = 0; h < Tile<BY>::NYH; ++h)
            if (L.yh_has[h]) lds[L.yh_lds[h]] = yh[h];
        if (L.has_xh) lds[L.xh_lds] = xh;
    }
};

// Second derivative along x and y from the LDS tile, each scaled as the
// reference scales it (kernel3d.cu:54-61): tx*sx + ty*sy, returned as ty+tx.
__device__ __forceinline__ float lap_xy_scaled(const float *T, float sx, float sy)
{
    constexpr int W = 72;
    const float s0 = T[0];
    float tx = (VTI_C0 * s0 + VTI_C1 * (T[-1] + T[1]) + VTI_C2 * (T[-2] + T[2]) + VTI_C3 * (T[-3] + T[3]) +
                VTI_C4 * (T[-4] + T[4])) *
               sx;
    float ty = (VTI_C0 * s0 + VTI_C1 * (T[-W] + T[W]) + VTI_C2 * (T[-2 * W] + T[2 * W]) +
                VTI_C3 * (T[-3 * W] + T[3 * W]) + VTI_C4 * (T[-4 * W] + T[4 * W])) *
               sy;
    return ty + tx;
}

// Same stencils divided by dx2/dy2 (gradientKernel, kernel3d.cu:274-281).
__device__ __forceinline__ float lap_xy_div(const float *T, float dx2, float dy2)
{
    constexpr int W = 72;
    const float s0 = T[0];
    float tx = (VTI_C0 * s0 + VTI_C1 * (T[-1] + T[1]) + VTI_C2 * (T[-2] + T[2]) + VTI_C3 * (T[-3] + T[3]) +
                VTI_C4 * (T[-4] + T[4])) /
               dx2;
    float ty = (VTI_C0 * s0 + VTI_C1 * (T[-W] + T[W]) + VTI_C2 * (T[-2 * W] + T[2 * W]) +
                VTI_C3 * (T[-3 * W] + T[3 * W]) + VTI_C4 * (T[-4 * W] + T[4 * W])) /
               dy2;
    return ty + tx;
}

__device__ __forceinline__ float lap_z(const float *e)
{
    return VTI_C0 * e[4] + VTI_C1 * (e[3] + e[5]) + VTI_C2 * (e[2] + e[6]) + VTI_C3 * (e[1] + e[7]) +
           VTI_C4 * (e[0] + e[8]);
}

// ---- the reference's own 3D arithmetic (kernel3d.cu:54-69, wave3d.h:8-12):
// pair sums of float neighbours in float, products and accumulation in
// double with double weights, rounded to float once per stencil.
#define VTI_C0D (-2.84722222222)
#define VTI_C1D (1.6)
#define VTI_C2D (-0.2)
#define VTI_C3D (0.02539682539)
#define VTI_C4D (-0.00178571428)

__device__ __forceinline__ double acc9(float m, float p1, float p2, float p3, float p4)
{
    return VTI_C0D * (double)m + VTI_C1D * (double)p1 + VTI_C2D * (double)p2 + VTI_C3D * (double)p3 +
           VTI_C4D * (double)p4;
}

// Arithmetic mode of the bulk stencil: 0 = FP32 throughout, 1 = the
// reference's mixed FP64 arithmetic on FP32 storage.
enum { kArithF32 = 0, kArithRef = 1 };

template <int ARITH>
__device__ __forceinline__ float sx_scaled(const float *T, float sx, float sy)
{
    if (ARITH == kArithF32) return lap_xy_scaled(T, sx, sy);
    constexpr int W = 72;
    const float tx = (float)(acc9(T[0], T[-1] + T[1], T[-2] + T[2], T[-3] + T[3], T[-4] + T[4]) * (double)sx);
    float ty = (float)(acc9(T[0], T[-W] + T[W], T[-2 * W] + T[2 * W], T[-3 * W] + T[3 * W], T[-4 * W] + T[4 * W]) *
                       (double)sy);
    ty += tx;
    return ty;
}

template <int ARITH>
__device__ __forceinline__ float sx_div(const float *T, float dx2, float dy2)
{
    if (ARITH == kArithF32) return lap_xy_div(T, dx2, dy2);
    constexpr int W = 72;
    const float tx = (float)(acc9(T[0], T[-1] + T[1], T[-2] + T[2], T[-3] + T[3], T[-4] + T[4]) / (double)dx2);
    float ty = (float)(acc9(T[0], T[-W] + T[W], T[-2 * W] + T[2 * W], T[-3 * W] + T[3 * W], T[-4 * W] + T[4 * W]) /
                       (double)dy2);
    ty += tx;
    return ty;
}

template <int ARITH>
__device__ __forceinline__ float sz_scaled(const float *e, float s)
{
    if (ARITH == kArithF32) return lap_z(e) * s;
    return (float)(acc9(e[4], e[3] + e[5], e[2] + e[6], e[1] + e[7], e[0] + e[8]) * (double)s);
}

template <int ARITH>
__device__ __forceinline__ float sz_div(const float *e, float d2)
{
    if (ARITH == kArithF32) return lap_z(e) / d2;
    return (float)(acc9(e[4], e[3] + e[5], e[2] + e[6], e[1] + e[7], e[0] + e[8]) / (double)d2);
}

// next = a + 2*cur - prev: float in FP32 mode, double then rounded in the
// reference's arithmetic (kernel3d.cu:68-69)
template <int ARITH>
__device__ __forceinline__ float leap(float a, float cur, float prev)
{
    if (ARITH == kArithF32) return (a + 2.f * cur) - prev;
    return (float)((double)a + 2. * (double)cur - (double)prev);
}

// Effective sigma_z column for output plane 4b+R.  The reference damps each
// z-block's current level right after that block's step, before the next
// block runs (objFuncGradient3d_f.cu:660-662), so the 4-R planes of the block
// above that this point's z-stencil reaches have been damped one extra time.
template <int R, bool ABC>
__device__ __forceinline__ void effective_column(float *e, const float *q, float dxy, const float *__restrict__ dzv,
                                                 int z)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = q[k];
    if (ABC && z >= 4) {
#pragma unroll
        for (int k = 0; k < 4 - R; ++k) e[k] = q[k] * (dxy * dzv[z - 4 + k]);
    }
}

// effective_column with the plane's position in its 4-plane block known only
// at run time (r = z % 4).
template <bool ABC>
__device__ __forceinline__ void effective_column_rt(float *e, const float *q, float dxy,
                                                    const float *__restrict__ dzv, int z, int r)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = q[k];
    if (ABC && z >= 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < 4 - r) e[k] = q[k] * (dxy * dzv[z - 4 + k]);
    }
}

// ---------------------------------------------------------------------------
// Forward step (forwardKernel semantics, kernel3d.cu:3-74, plus the fused
// injectSource and abcXYZ/abcXY of objFuncGradient3d_f.cu:655-662).
// prev_x/prev_z are read at each point and overwritten with the next level.
// With ABC the stored levels are damped once; the second damping the
// reference applies to the current level is applied on read, to prev.
//
// Software pipeline along z: while plane z is computed, the loads for plane
// z+1 (prev, c11/c13/c33, the sigma_z plane entering the queue) and the
// sigma_x tile of plane z+2 are in flight, so each wave keeps about two
// planes of requests outstanding.  Plane bases are wave-uniform; per-lane
// offsets within a plane are 32-bit.
// ---------------------------------------------------------------------------
template <int ARITH, bool ABC, int BY>
__global__ __launch_bounds__(64 * BY) void fwd3d_kernel(
    float *__restrict__ nxt_x, const float *__restrict__ cur_x, float *__restrict__ nxt_z,
    const float *__restrict__ cur_z, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ dxy_plane, const float *__restrict__ dzv, Geo3 g,
    int zc, int ntx, int nty, float dt2dx2, float dt2dy2, float dt2dz2, long long src_idx, float src_val)
{
    constexpr int W = Tile<BY>::W;
    constexpr int R = Tile<BY>::R;
    __shared__ float tile[2][R * W];

    const int lb = xcd_logical_block(gridDim.x * gridDim.y * gridDim.z);
    const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = bx * 64, y0 = 4 + by * BY;
    const int zs = bz * zc;
    const int ze = min(zs + zc, g.zmax);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool inrange = ix < g.nx && iy < g.ny;
    const bool active = ix >= 4 && ix < g.nx - 4 && iy < g.ny - 4;

    TileLoader<BY> L;
    L.init(tx, ty, x0, y0, g);
    const unsigned col4 = 4u * L.main_off;
    const float dxy = (ABC && inrange) ? dxy_plane[L.main_off] : 1.f;
    // resources based at the first plane this chunk touches; the host keeps
    // (zc + 8) planes within 32-bit byte offsets
    const int zb = max(zs - 4, 0);
    const size_t base = (size_t)zb * (size_t)g.plane;
    const rsrc_t Rcx = make_rsrc(cur_x + base), Rcz = make_rsrc(cur_z + base);
    const rsrc_t Rnx = make_rsrc(nxt_x + base), Rnz = make_rsrc(nxt_z + base);
    const rsrc_t R11 = make_rsrc(c11 + base), R13 = make_rsrc(c13 + base), R33 = make_rsrc(c33 + base);
    const unsigned pb = (unsigned)(4 * g.plane);
    auto so = [&](int z) { return (unsigned)(z - zb) * pb; };

    // sigma_z queue: q[k] = sigma_z(z-4+k) at plane z; qn = sigma_z(z+5) in flight
    float q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int zz = zs - 4 + k;
        q[k] = (inrange && zz >= 0 && zz < g.zmax) ? bload(Rcz, col4, so(zz)) : 0.f;
    }
    // point data of plane z (prev and coefficients)
    float px = 0.f, pz = 0.f, a11 = 0.f, a13 = 0.f, a33 = 0.f;
    if (inrange) {
        px = bload_nt(Rnx, col4, so(zs));
        pz = bload_nt(Rnz, col4, so(zs));
        a11 = bload_nt(R11, col4, so(zs));
        a13 = bload_nt(R13, col4, so(zs));
        a33 = bload_nt(R33, col4, so(zs));
    }
    PlaneRegs<BY> P;
    P.loadb(Rcx, so(zs), L);
    P.store(tile[0], L);
    if (zs + 1 < ze) P.loadb(Rcx, so(zs + 1), L);
    __syncthreads();

    for (int z0 = zs; z0 < ze; z0 += 4) {
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
            const int z = z0 + r;
            // the tile of plane z+1 goes to LDS; plane z+2's is requested
            if (z + 1 < ze) P.store(tile[(r + 1) & 1], L);
            if (z + 2 < ze) P.loadb(Rcx, so(z + 2), L);
            const bool nx1 = inrange && z + 1 < ze;
            float npx = 0.f, npz = 0.f, n11 = 0.f, n13 = 0.f, n33 = 0.f, qn = 0.f;
            if (nx1) {
                const unsigned o1 = so(z + 1);
                npx = bload_nt(Rnx, col4, o1);
                npz = bload_nt(Rnz, col4, o1);
                n11 = bload_nt(R11, col4, o1);
                n13 = bload_nt(R13, col4, o1);
                n33 = bload_nt(R33, col4, o1);
            }
            if (inrange && z + 5 < g.zmax) qn = bload(Rcz, col4, so(z + 5));

            const float *T = &tile[r & 1][(ty + 4) * W + tx + 4];
            const float s0 = T[0];
            const float txy = sx_scaled<ARITH>(T, dt2dx2, dt2dy2);
            float e[9];
            effective_column_rt<ABC>(e, q, dxy, dzv, z, r);
            const float tz = sz_scaled<ARITH>(e, dt2dz2);
            float d = 1.f;
            if (ABC) {
                d = dxy * dzv[z];
                px = px * d;
                pz = pz * d;
            }
            float vx = leap<ARITH>(a11 * txy + a13 * tz, s0, px);
            float vz = leap<ARITH>(a13 * txy + a33 * tz, q[4], pz);
            const long long gi = (long long)z * g.plane + L.main_off;
            if (gi == src_idx) {
                vx = vx + src_val;
                vz = vz + src_val;
            }
            if (ABC) {
                vx = vx * d;
                vz = vz * d;
            }
            if (active) {
                bstore_nt(Rnx, col4, so(z), vx);
                bstore_nt(Rnz, col4, so(z), vz);
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
            q[8] = qn;
            px = npx;
            pz = npz;
            a11 = n11;
            a13 = n13;
            a33 = n33;
        }
    }
}

// ---------------------------------------------------------------------------
// Two forward levels per launch (temporal blocking).
//
// Level L-1 is read from (ax, az), level L from (bx, bz); level L+1 goes to
// (cx, cz) and L+2 to (dx_, dz_) -- four distinct buffer pairs, since the
// neighbouring workgroups still read the halos of L-1 and L.
//
// A workgroup of 64 x BY threads owns a 64 x BY column tile and a z-chunk.
// Marching down z, it advances level L+1 at plane z on the tile plus a
// 4-point x/y halo E = [x0-4, x0+68) x [y0-4, y0+BY+4), and level L+2 at
// plane z-4 on the tile.  Per point the arithmetic is that of fwd3d_kernel,
// operation for operation (same stencils, the same effective sigma_z column,
// damping on read, source, damping on write), so one launch gives the bits of
// two fwd3d_kernel launches.  Points of E outside the updated interior keep
// the value of level L-1 there (a one-level step leaves them untouched in the
// buffer it writes), out-of-range points read as 0.  The caller keeps the
// values outside the interior of (cx, cz) equal to those of (ax, az), and of
// (dx_, dz_) equal to (bx, bz).
//
// sigma_x(L) planes sit in an 80-wide LDS tile (double buffered) and
// sigma_x(L+1) planes of E in a 5-slot LDS ring, so that level L+2 at plane
// z-4 finds its plane there (fwd3d_t2b_kernel below).
// ---------------------------------------------------------------------------

template <int ARITH, int W>
__device__ __forceinline__ float sx_scaled_w(const float *T, float sx, float sy)
{
    if (ARITH == kArithF32) {
        const float s0 = T[0];
        float tx = (VTI_C0 * s0 + VTI_C1 * (T[-1] + T[1]) + VTI_C2 * (T[-2] + T[2]) + VTI_C3 * (T[-3] + T[3]) +
                    VTI_C4 * (T[-4] + T[4])) *
                   sx;
        float ty = (VTI_C0 * s0 + VTI_C1 * (T[-W] + T[W]) + VTI_C2 * (T[-2 * W] + T[2 * W]) +
                    VTI_C3 * (T[-3 * W] + T[3 * W]) + VTI_C4 * (T[-4 * W] + T[4 * W])) *
                   sy;
        return ty + tx;
    }
    const float tx = (float)(acc9(T[0], T[-1] + T[1], T[-2] + T[2], T[-3] + T[3], T[-4] + T[4]) * (double)sx);
    float ty = (float)(acc9(T[0], T[-W] + T[W], T[-2 * W] + T[2 * W], T[-3 * W] + T[3 * W], T[-4 * W] + T[4 * W]) *
                       (double)sy);
    ty += tx;
    return ty;
}

template <int BY>
struct T2Shape {
    static constexpr int NT = 64 * BY;            // threads
    static constexpr int EW = 72, EH = BY + 8;    // E: x0-4 .. x0+68, y0-4 .. y0+BY+4
    static constexpr int NRIM = EW * EH - 64 * BY;
    static constexpr int TW = 80, TH = BY + 16;   // sigma_x(L) tile: x0-8 .., y0-8 ..
};

// E coordinates of the points a thread owns: slot 0 the tile point, then rim
// points (top 4 rows, bottom 4 rows, left 4 columns, right 4 columns)
template <int BY>
__device__ __forceinline__ bool t2_point(int tid, int j, int &ex, int &ey)
{
    using S = T2Shape<BY>;
    if (j == 0) {
        ex = (tid & 63) + 4;
        ey = (tid >> 6) + 4;
        return true;
    }
    const int h = tid + (j - 1) * S::NT;
    if (h >= S::NRIM) return false;
    if (h < 288) {
        ey = h / 72;
        ex = h % 72;
    } else if (h < 576) {
        ey = BY + 4 + (h - 288) / 72;
        ex = (h - 288) % 72;
    } else if (h < 576 + 4 * BY) {
        ey = 4 + (h - 576) / 4;
        ex = (h - 576) % 4;
    } else {
        ey = 4 + (h - 576 - 4 * BY) / 4;
        ex = 68 + (h - 576 - 4 * BY) % 4;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Two forward levels per launch on 64 x 16 tiles (1024 threads).
//
// The scheme above with the arithmetic of fwd3d_kernel, arranged for few
// registers per thread so that four waves per SIMD stay resident without
// spills:
//  - a 16-row tile: E = 72 x 24 points, the tile's 1024 plus a 704-point rim,
//    so each thread owns its tile point and at most one rim point (the level
//    L+1 recompute on the rim is 0.69 of the tile's work, against 1.25 on
//    8-row tiles);
//  - a rim point needs only sigma_x(L+1) (level L+2 reads sigma_z(L+1) on
//    the tile's own columns), so it computes sigma_x alone, and the 9-deep
//    sigma_z(L) queue of its column lives in LDS, private to its thread (no
//    barrier: one lane writes and reads it);
//  - the operands of level L+2 at plane z-4 (sigma_x(L) there and the
//    coefficients) are read again from memory -- the same values, from L2 --
//    instead of being held in registers for four planes.
// Launch: blocks of (64, 16), grid ntx * nty * ntz with nty = ceil((ny-8)/16).
// ---------------------------------------------------------------------------
template <int ARITH, bool ABC>
__global__ __launch_bounds__(1024) void fwd3d_t2b_kernel(
    float *__restrict__ cx, float *__restrict__ cz, float *__restrict__ dx_, float *__restrict__ dz_,
    const float *__restrict__ ax, const float *__restrict__ az, const float *__restrict__ bx,
    const float *__restrict__ bz, const float *__restrict__ c11, const float *__restrict__ c13,
    const float *__restrict__ c33, const float *__restrict__ dxy_plane, const float *__restrict__ dzv, Geo3 g,
    int zc, int ntx, int nty, float dt2dx2, float dt2dy2, float dt2dz2, long long src_idx, float src1, float src2)
{
    constexpr int BY = 16;
    using S = T2Shape<BY>;
    constexpr int TW = S::TW, TH = S::TH, EW = S::EW, EH = S::EH, NRIM = S::NRIM;
    constexpr int TL = (TW * TH + S::NT - 1) / S::NT;  // tile elements loaded per thread
    __shared__ float tile[2][TH * TW];
    __shared__ float ring[5][EH * EW];
    __shared__ float rq[9][NRIM];   // sigma_z(L) queues of the rim columns, slot = plane mod 9
    __shared__ float q1s[9][S::NT];  // sigma_z(L+1) queues of the tile columns, slot = plane mod 9

    const int lb = xcd_logical_block(gridDim.x * gridDim.y * gridDim.z);
    const int bxi = lb % ntx, byi = (lb / ntx) % nty, bzi = lb / (ntx * nty);
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int x0 = bxi * 64, y0 = 4 + byi * BY;
    const int zs = bzi * zc, ze = min(zs + zc, g.zmax);
    const int z1s = max(zs - 4, 0), z1e = min(ze + 4, g.zmax);  // planes of level L+1

    // the tile point (slot 0) and the rim point (slot 1) of this thread
    int ex0, ey0, ex1 = 0, ey1 = 0;
    t2_point<BY>(tid, 0, ex0, ey0);
    const bool rim = t2_point<BY>(tid, 1, ex1, ey1);
    const int gx0 = x0 - 4 + ex0, gy0 = y0 - 4 + ey0, gx1 = x0 - 4 + ex1, gy1 = y0 - 4 + ey1;
    const bool inr0 = gx0 < g.nx && gy0 < g.ny;
    const bool act0 = inr0 && gx0 >= 4 && gx0 < g.nx - 4 && gy0 < g.ny - 4;
    const bool inr1 = rim && gx1 >= 0 && gx1 < g.nx && gy1 >= 0 && gy1 < g.ny;
    const bool act1 = inr1 && gx1 >= 4 && gx1 < g.nx - 4 && gy1 >= 4 && gy1 < g.ny - 4;
    const unsigned off0 = inr0 ? (unsigned)(gy0 * g.pitch + gx0) : 0u;
    const unsigned off1 = inr1 ? (unsigned)(gy1 * g.pitch + gx1) : 0u;
    const float dxy0 = (ABC && inr0) ? dxy_plane[off0] : 1.f;
    const float dxy1 = (ABC && inr1) ? dxy_plane[off1] : 1.f;
    auto P = [&](const float *base, int z) { return base + (size_t)z * (size_t)g.plane; };

    // sigma_x(L) tile elements this thread loads: offsets within a plane
    unsigned toff[TL];
    bool tok[TL];
#pragma unroll
    for (int k = 0; k < TL; ++k) {
        const int e = tid + k * S::NT;
        const int x = x0 - 8 + e % TW, y = y0 - 8 + e / TW;
        tok[k] = e < TW * TH && x >= 0 && x < g.nx && y >= 0 && y < g.ny;
        toff[k] = tok[k] ? (unsigned)(y * g.pitch + x) : 0u;
    }
    float treg[TL];
    auto tile_regs = [&](int z) {
        const float *B = P(bx, z);
#pragma unroll
        for (int k = 0; k < TL; ++k) treg[k] = tok[k] ? B[toff[k]] : 0.f;
    };
    auto tile_store = [&](float *dst) {
#pragma unroll
        for (int k = 0; k < TL; ++k)
            if (tid + k * S::NT < TW * TH) dst[tid + k * S::NT] = treg[k];
    };

    // sigma_z(L) queue of the tile column (q[k] = plane z-4+k) and of the rim
    // column (rq, slot (z-4+k) mod 9); sigma_z(L+1) of the tile column (q1s,
    // slot (z-8+k) mod 9).  The LDS queues are private to their thread.
    float q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int zz = z1s - 4 + k;
        const bool in = zz >= 0 && zz < g.zmax;
        q[k] = (inr0 && in) ? P(bz, zz)[off0] : 0.f;
        q1s[(z1s - 8 + k + 18) % 9][tid] = 0.f;
        if (rim) rq[(zz + 9) % 9][tid] = (inr1 && in) ? P(bz, zz)[off1] : 0.f;
    }

    // operands of level L+1 at one plane: the previous level and the
    // coefficients (the rim point needs what sigma_x takes)
    float pax0, paz0, k11_0, k13_0, k33_0, pax1, k11_1, k13_1;
    auto load_pts = [&](int z, float &oax0, float &oaz0, float &o11_0, float &o13_0, float &o33_0, float &oax1,
                        float &o11_1, float &o13_1) {
        const bool l0 = z < z1e && inr0, l1 = z < z1e && inr1;
        oax0 = l0 ? P(ax, z)[off0] : 0.f;
        oaz0 = l0 ? P(az, z)[off0] : 0.f;
        o11_0 = l0 && act0 ? P(c11, z)[off0] : 0.f;
        o13_0 = l0 && act0 ? P(c13, z)[off0] : 0.f;
        o33_0 = l0 && act0 ? P(c33, z)[off0] : 0.f;
        oax1 = l1 ? P(ax, z)[off1] : 0.f;
        o11_1 = l1 && act1 ? P(c11, z)[off1] : 0.f;
        o13_1 = l1 && act1 ? P(c13, z)[off1] : 0.f;
    };

    if (z1s < z1e) {
        tile_regs(z1s);
        tile_store(tile[z1s & 1]);
    }
    if (z1s + 1 < z1e) tile_regs(z1s + 1);
    load_pts(z1s, pax0, paz0, k11_0, k13_0, k33_0, pax1, k11_1, k13_1);
    __syncthreads();

    for (int z = z1s; z < ze + 4; ++z) {
        const bool step1 = z < z1e;
        const int zp = z - 4;
        const bool step2 = zp >= zs && zp < ze && act0;
        // requests for later planes, in flight while this one is computed:
        // level L+2's operands at zp, the next plane's point operands and
        // sigma_z(L) planes, then the sigma_x(L) tile two planes ahead
        float hx = 0.f, s11 = 0.f, s13 = 0.f, s33 = 0.f;
        if (step2) {
            hx = P(bx, zp)[off0];
            s11 = P(c11, zp)[off0];
            s13 = P(c13, zp)[off0];
            s33 = P(c33, zp)[off0];
        }
        float nax0, naz0, n11_0, n13_0, n33_0, nax1, n11_1, n13_1;
        load_pts(z + 1, nax0, naz0, n11_0, n13_0, n33_0, nax1, n11_1, n13_1);
        const bool qz = z + 5 < g.zmax;
        const float qn0 = (inr0 && qz) ? P(bz, z + 5)[off0] : 0.f;
        const float qn1 = (inr1 && qz) ? P(bz, z + 5)[off1] : 0.f;
        if (z + 1 < z1e) tile_store(tile[(z + 1) & 1]);
        if (z + 2 < z1e) tile_regs(z + 2);

        // ---- level L+1 at plane z on E
        if (step1) {
            const float *Tp = tile[z & 1];
            const float d1 = ABC ? dzv[z] : 1.f;
            const long long pz = (long long)z * g.plane;
            // tile point: sigma_x and sigma_z
            float vx = 0.f, vz = 0.f;
            if (act0) {
                const float *T = &Tp[(ey0 + 4) * TW + ex0 + 4];
                const float s0 = T[0];
                const float txy = sx_scaled_w<ARITH, TW>(T, dt2dx2, dt2dy2);
                float e9[9];
                effective_column_rt<ABC>(e9, q, dxy0, dzv, z, z & 3);
                const float tz = sz_scaled<ARITH>(e9, dt2dz2);
                float px = pax0, pzv = paz0, d = 1.f;
                if (ABC) {
                    d = dxy0 * d1;
                    px = px * d;
                    pzv = pzv * d;
                }
                vx = leap<ARITH>(k11_0 * txy + k13_0 * tz, s0, px);
                vz = leap<ARITH>(k13_0 * txy + k33_0 * tz, q[4], pzv);
                const long long gi = pz + off0;
                if (gi == src_idx) {
                    vx = vx + src1;
                    vz = vz + src1;
                }
                if (ABC) {
                    vx = vx * d;
                    vz = vz * d;
                }
                if (z >= zs && z < ze) {
                    cx[gi] = vx;
                    cz[gi] = vz;
                }
            } else if (inr0) {
                vx = pax0;
                vz = paz0;
            }
            ring[z % 5][ey0 * EW + ex0] = vx;
            q1s[z % 9][tid] = vz;
            // rim point: sigma_x only
            if (rim) {
                float vr = 0.f;
                if (act1) {
                    const float *T = &Tp[(ey1 + 4) * TW + ex1 + 4];
                    const float s0 = T[0];
                    const float txy = sx_scaled_w<ARITH, TW>(T, dt2dx2, dt2dy2);
                    float e9[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) e9[k] = rq[(z - 4 + k + 9) % 9][tid];
                    effective_column_rt<ABC>(e9, e9, dxy1, dzv, z, z & 3);
                    const float tz = sz_scaled<ARITH>(e9, dt2dz2);
                    float px = pax1, d = 1.f;
                    if (ABC) {
                        d = dxy1 * d1;
                        px = px * d;
                    }
                    vr = leap<ARITH>(k11_1 * txy + k13_1 * tz, s0, px);
                    if (pz + off1 == src_idx) vr = vr + src1;
                    if (ABC) vr = vr * d;
                } else if (inr1) {
                    vr = pax1;
                }
                ring[z % 5][ey1 * EW + ex1] = vr;
            }
        } else {
            q1s[z % 9][tid] = 0.f;
        }
        // ---- level L+2 at plane z-4 on the tile (its ring plane was
        // written four iterations ago)
        if (step2) {
            const float *T = &ring[zp % 5][ey0 * EW + ex0];
            const float s0 = T[0];
            const float txy = sx_scaled_w<ARITH, EW>(T, dt2dx2, dt2dy2);
            float e9[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) e9[k] = q1s[(zp - 4 + k + 9) % 9][tid];
            const float q1c = e9[4];
            effective_column_rt<ABC>(e9, e9, dxy0, dzv, zp, zp & 3);
            const float tz = sz_scaled<ARITH>(e9, dt2dz2);
            const long long gi = (long long)zp * g.plane + off0;
            float px = hx, pzv = q[0], d = 1.f;
            if (ABC) {
                d = dxy0 * dzv[zp];
                px = px * d;
                pzv = pzv * d;
            }
            float vx = leap<ARITH>(s11 * txy + s13 * tz, s0, px);
            float vz = leap<ARITH>(s13 * txy + s33 * tz, q1c, pzv);
            if (gi == src_idx) {
                vx = vx + src2;
                vz = vz + src2;
            }
            if (ABC) {
                vx = vx * d;
                vz = vz * d;
            }
            dx_[gi] = vx;
            dz_[gi] = vz;
        }
        // the next iteration overwrites the tile buffer read above and the
        // ring slot of plane z-4
        __syncthreads();
        // advance to plane z+1: queues, operands
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
        q[8] = qn0;
        if (rim) rq[(z + 5) % 9][tid] = qn1;
        pax0 = nax0;
        paz0 = naz0;
        k11_0 = n11_0;
        k13_0 = n13_0;
        k33_0 = n33_0;
        pax1 = nax1;
        k11_1 = n11_1;
        k13_1 = n13_1;
    }
}

// ---------------------------------------------------------------------------
// Fused backward step of the 3D gradient and RTM passes.
//
// MODE 0 (gradient, objFuncGradient3d_f.cu:1008-1028): per point
//   source side: gradientKernel -- sigma_next = dt2*(C*lap sigma) + 2 sigma - sigma_prev
//                (the undamped random-boundary reconstruction, run backward in
//                time) with gc11 += lx*Lxy, gc33 += lz*Lz, gc13 += lx*Lz + lz*Lxy
//   receiver side: forwardKernel on lambda with the damping of the absorbing pass
// MODE 1 (RTM, rtm3d_f.cu:585-610): forwardKernel on both fields and
//   image += (2/3 sx + 1/3 sz)(2/3 lx + 1/3 lz), with lambda at the current
//   level after its damping.
// The source injection into sigma_next is fused (own point).  In the four
// receiver planes lambda_next is stored undamped: the residual injection that
// the reference performs before damping is applied next by
// inject_damp3d_kernel, which then damps those planes.
// ---------------------------------------------------------------------------
template <int MODE, int ARITH, int BY>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void bwd3d_kernel(
    float *__restrict__ s_nx, const float *__restrict__ s_cx, float *__restrict__ s_nz,
    const float *__restrict__ s_cz, float *__restrict__ l_nx, const float *__restrict__ l_cx,
    float *__restrict__ l_nz, const float *__restrict__ l_cz, const float *__restrict__ c11,
    const float *__restrict__ c13, const float *__restrict__ c33, float *__restrict__ g11, float *__restrict__ g13,
    float *__restrict__ g33, const float *__restrict__ dxy_plane, const float *__restrict__ dzv, Geo3 g, int zc,
    int ntx, int nty, float dx2, float dy2, float dz2, float dt2, float dt2dx2, float dt2dy2, float dt2dz2,
    long long src_idx, float src_val, int zrec0)
{
    constexpr int W = Tile<BY>::W;
    constexpr int R = Tile<BY>::R;
    __shared__ float ts[2][R * W];
    __shared__ float tl[2][R * W];

    const int lb = xcd_logical_block(gridDim.x * gridDim.y * gridDim.z);
    const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = bx * 64, y0 = 4 + by * BY;
    const int zs = bz * zc;
    const int ze = min(zs + zc, g.zmax);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool inrange = ix < g.nx && iy < g.ny;
    const bool active = ix >= 4 && ix < g.nx - 4 && iy < g.ny - 4;

    TileLoader<BY> L;
    L.init(tx, ty, x0, y0, g);
    const unsigned col4 = 4u * L.main_off;
    const float dxy = inrange ? dxy_plane[L.main_off] : 1.f;
    const int zb = max(zs - 4, 0);
    const size_t base = (size_t)zb * (size_t)g.plane;
    const rsrc_t Rscx = make_rsrc(s_cx + base), Rscz = make_rsrc(s_cz + base);
    const rsrc_t Rlcx = make_rsrc(l_cx + base), Rlcz = make_rsrc(l_cz + base);
    const rsrc_t Rsnx = make_rsrc(s_nx + base), Rsnz = make_rsrc(s_nz + base);
    const rsrc_t Rlnx = make_rsrc(l_nx + base), Rlnz = make_rsrc(l_nz + base);
    const rsrc_t R11 = make_rsrc(c11 + base), R13 = make_rsrc(c13 + base), R33 = make_rsrc(c33 + base);
    const rsrc_t G11 = make_rsrc(g11 + base);
    const rsrc_t G13 = make_rsrc(MODE == 0 ? g13 + base : g11 + base);
    const rsrc_t G33 = make_rsrc(MODE == 0 ? g33 + base : g11 + base);
    const unsigned pb = (unsigned)(4 * g.plane);
    auto so = [&](int z) { return (unsigned)(z - zb) * pb; };

    // queues: qs/ql[k] = sigma_z/lambda_z(z-4+k) at plane z; the (z+5) values in flight
    float qs[9], ql[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int zz = zs - 4 + k;
        const bool ok = inrange && zz >= 0 && zz < g.zmax;
        qs[k] = ok ? bload(Rscz, col4, so(zz)) : 0.f;
        ql[k] = ok ? bload(Rlcz, col4, so(zz)) : 0.f;
    }
    // point data of plane z
    float spx = 0.f, spz = 0.f, lpx = 0.f, lpz = 0.f, a11 = 0.f, a13 = 0.f, a33 = 0.f;
    if (inrange) {
        const unsigned o0 = so(zs);
        spx = bload_nt(Rsnx, col4, o0);
        spz = bload_nt(Rsnz, col4, o0);
        lpx = bload_nt(Rlnx, col4, o0);
        lpz = bload_nt(Rlnz, col4, o0);
        a11 = bload_nt(R11, col4, o0);
        a13 = bload_nt(R13, col4, o0);
        a33 = bload_nt(R33, col4, o0);
    }
    PlaneRegs<BY> PS, PL;
    PS.loadb(Rscx, so(zs), L);
    PL.loadb(Rlcx, so(zs), L);
    PS.store(ts[0], L);
    PL.store(tl[0], L);
    if (zs + 1 < ze) {
        PS.loadb(Rscx, so(zs + 1), L);
        PL.loadb(Rlcx, so(zs + 1), L);
    }
    __syncthreads();

    for (int z0 = zs; z0 < ze; z0 += 4) {
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
            const int z = z0 + r;
            if (z + 1 < ze) {
                PS.store(ts[(r + 1) & 1], L);
                PL.store(tl[(r + 1) & 1], L);
            }
            if (z + 2 < ze) {
                PS.loadb(Rscx, so(z + 2), L);
                PL.loadb(Rlcx, so(z + 2), L);
            }
            // the image/gradient accumulators of this plane, needed last
            float o11 = 0.f, o13 = 0.f, o33 = 0.f;
            if (inrange) {
                const unsigned oz = so(z);
                o11 = bload_nt(G11, col4, oz);
                if (MODE == 0) {
                    o13 = bload_nt(G13, col4, oz);
                    o33 = bload_nt(G33, col4, oz);
                }
            }
            float n_spx = 0.f, n_spz = 0.f, n_lpx = 0.f, n_lpz = 0.f, n11 = 0.f, n13 = 0.f, n33 = 0.f;
            float qsn = 0.f, qln = 0.f;
            if (inrange && z + 1 < ze) {
                const unsigned o1 = so(z + 1);
                n_spx = bload_nt(Rsnx, col4, o1);
                n_spz = bload_nt(Rsnz, col4, o1);
                n_lpx = bload_nt(Rlnx, col4, o1);
                n_lpz = bload_nt(Rlnz, col4, o1);
                n11 = bload_nt(R11, col4, o1);
                n13 = bload_nt(R13, col4, o1);
                n33 = bload_nt(R33, col4, o1);
            }
            if (inrange && z + 5 < g.zmax) {
                qsn = bload(Rscz, col4, so(z + 5));
                qln = bload(Rlcz, col4, so(z + 5));
            }
            const float *TS = &ts[r & 1][(ty + 4) * W + tx + 4];
            const float *TL = &tl[r & 1][(ty + 4) * W + tx + 4];
            const float s0 = TS[0], l0 = TL[0];

            // ---- source side
            float svx, svz, sTxy = 0.f, sTz = 0.f;
            if (MODE == 0) {
                sTxy = sx_div<ARITH>(TS, dx2, dy2);
                sTz = sz_div<ARITH>(qs, dz2);
                svx = leap<ARITH>(dt2 * (a11 * sTxy + a13 * sTz), s0, spx);
                svz = leap<ARITH>(dt2 * (a13 * sTxy + a33 * sTz), qs[4], spz);
            } else {
                const float txy = sx_scaled<ARITH>(TS, dt2dx2, dt2dy2);
                const float tz = sz_scaled<ARITH>(qs, dt2dz2);
                svx = leap<ARITH>(a11 * txy + a13 * tz, s0, spx);
                svz = leap<ARITH>(a13 * txy + a33 * tz, qs[4], spz);
            }
            const long long gi = (long long)z * g.plane + L.main_off;
            if (gi == src_idx) {
                svx = svx + src_val;
                svz = svz + src_val;
            }

            // ---- receiver side (forwardKernel + damping)
            const float lTxy = sx_scaled<ARITH>(TL, dt2dx2, dt2dy2);
            float e[9];
            effective_column_rt<true>(e, ql, dxy, dzv, z, r);
            const float lTz = sz_scaled<ARITH>(e, dt2dz2);
            const float d = dxy * dzv[z];
            float lvx = leap<ARITH>(a11 * lTxy + a13 * lTz, l0, lpx * d);
            float lvz = leap<ARITH>(a13 * lTxy + a33 * lTz, ql[4], lpz * d);
            if (z < zrec0 || z >= zrec0 + 4) {
                lvx = lvx * d;
                lvz = lvz * d;
            }

            if (MODE == 0) {
                o11 = o11 + l0 * sTxy;
                o33 = o33 + ql[4] * sTz;
                o13 = o13 + (l0 * sTz + ql[4] * sTxy);
            } else if (ARITH == kArithF32) {
                const float ls = VTI_TWOTHIRD_F * (l0 * d) + VTI_ONETHIRD_F * (ql[4] * d);
                o11 = o11 + (VTI_TWOTHIRD_F * s0 + VTI_ONETHIRD_F * qs[4]) * ls;
            } else {
                // imagingKernel (kernel3d.cu:465): double weights, float storage
                const double ls = VTI_TWOTHIRD_D * (double)(l0 * d) + VTI_ONETHIRD_D * (double)(ql[4] * d);
                o11 = (float)((double)o11 + (VTI_TWOTHIRD_D * (double)s0 + VTI_ONETHIRD_D * (double)qs[4]) * ls);
            }
            if (active) {
                const unsigned oz = so(z);
                bstore_nt(Rsnx, col4, oz, svx);
                bstore_nt(Rsnz, col4, oz, svz);
                bstore_nt(Rlnx, col4, oz, lvx);
                bstore_nt(Rlnz, col4, oz, lvz);
                bstore_nt(G11, col4, oz, o11);
                if (MODE == 0) {
                    bstore_nt(G13, col4, oz, o13);
                    bstore_nt(G33, col4, oz, o33);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                qs[k] = qs[k + 1];
                ql[k] = ql[k + 1];
            }
            qs[8] = qsn;
            ql[8] = qln;
            spx = n_spx;
            spz = n_spz;
            lpx = n_lpx;
            lpz = n_lpz;
            a11 = n11;
            a13 = n13;
            a33 = n33;
        }
    }
}

// Residual (or observed-data) injection into lambda_next at the receivers
// (injectResidual, injectRecord3d.cu:38-46) with the time interpolation of
// interpolateResidual (injectRecord3d.cu:60-78) evaluated on the device:
// the coarse traces are time-major rows [it][ir]; i1 < 0 selects row nnt-1.
__global__ void inject_res3d_kernel(float *__restrict__ lx, float *__restrict__ lz,
                                    const long long *__restrict__ rec, const float *__restrict__ coarse, int nr,
                                    int i1, float f, float dt2)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= nr) return;
    float fine;
    if (i1 < 0) {
        fine = coarse[ir];
    } else {
        fine = (float)((1. - (double)f) * (double)coarse[(size_t)i1 * nr + ir] +
                       (double)(f * coarse[(size_t)(i1 + 1) * nr + ir]));
    }
    const float s = dt2 * fine;
    const long long i = rec[ir];
    lx[i] = (float)((double)lx[i] + VTI_TWOTHIRD_D * (double)s);
    lz[i] = (float)((double)lz[i] + VTI_ONETHIRD_D * (double)s);
}

// Damping of the four receiver planes after the injection (abcXYZ/abcXY).
__global__ void damp_planes3d_kernel(float *__restrict__ lx, float *__restrict__ lz,
                                     const float *__restrict__ dxy_plane, const float *__restrict__ dzv, Geo3 g,
                                     int z0)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int z = z0 + blockIdx.z;
    if (ix < 4 || ix >= g.nx - 4 || iy < 4 || iy >= g.ny - 4 || z >= g.zmax) return;
    const long long col = (long long)iy * g.pitch + ix;
    const long long gi = z * g.plane + col;
    const float d = dxy_plane[col] * dzv[z];
    lx[gi] = lx[gi] * d;
    lz[gi] = lz[gi] * d;
}

}  // namespace vti
