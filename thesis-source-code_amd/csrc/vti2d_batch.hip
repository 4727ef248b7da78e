// Batched 2D objective and gradient: many shots, each on its own aperture
// sub-grid of one global model, in lockstep -- one launch per time level for
// the whole batch instead of one per shot.
//
// Per shot this is objFuncGradient_f (lib/vti-fwi-2d/objFuncGradient_f.cu:
// 15-371) on the model cropped to the shot's sub-grid: the same kernels'
// arithmetic, the same schedule (snapshots of D every rate/dt levels,
// residual injected with linear time interpolation, imaging against
// interpolated snapshots).  Each shot's gradient is accumulated in FP32 on its
// sub-grid, then added in FP64 into the global gradient at the sub-grid's
// columns, shot by shot in input order.  The batch is split into waves that
// fit the device memory budget; within a wave all shots advance together.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "vti2d_batch_kernels.h"
#include "vti_amd.h"
#include "vti_ctx.h"
#include "vti_dbuf.h"
#include "vti_host.h"
#include "vti_internal.h"

using vti::BatchGeo;
using vti::DBuf;
using vti::kBY2;
using vti::ShotB;

namespace {

__global__ void set_src_b_kernel(float *__restrict__ P, float *__restrict__ Q, const ShotB *__restrict__ shots,
                                 int ns, float v)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) {
        const long long i = shots[s].f_off + shots[s].src;
        P[i] = v;
        Q[i] = v;
    }
}

// synthetic_f's level 1: dt2*w[0] at each shot's source times the taper
// there (synthetic_f.cu:86-89), the taper formed as the step kernels form it
__global__ void set_src_taper_b_kernel(float *__restrict__ P, float *__restrict__ Q, const ShotB *__restrict__ shots,
                                       int ns, float v, const float *__restrict__ t0, int npad, int nnz)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) {
        const ShotB sh = shots[s];
        const float tap = vti::taper_of(vti::taper_ld(t0, npad, sh.nnx, nnz, sh.src % sh.nnx, sh.src / sh.nnx));
        const long long i = sh.f_off + sh.src;
        P[i] = v * tap;
        Q[i] = v * tap;
    }
}

// recordData of every receiver of the batch into row `row` of its shot's traces
__global__ void record_b_kernel(float *__restrict__ tr, const float *__restrict__ P, const float *__restrict__ Q,
                                const long long *__restrict__ fidx, const long long *__restrict__ tbase,
                                const int *__restrict__ tstride, int nrec, int row)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nrec) tr[tbase[k] + (long long)row * tstride[k]] = 0.5f * (P[fidx[k]] + Q[fidx[k]]);
}

// injectData of the last trace row at the first adjoint level (fr = 0)
__global__ void inject_first_b_kernel(float *__restrict__ P, float *__restrict__ Q, const float *__restrict__ tr,
                                      const long long *__restrict__ fidx, const long long *__restrict__ tbase,
                                      const int *__restrict__ tstride, int nrec, int row, float dt2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrec) return;
    const float r0 = tr[tbase[k] + (long long)row * tstride[k]];
    const float s = 0.5f * ((1.f - 0.f) * r0 + 0.f * 0.f);
    P[fidx[k]] += dt2 * s;
    Q[fidx[k]] += dt2 * s;
}

__global__ void residual_b_kernel(float *__restrict__ tr, const float *__restrict__ obs, long long n,
                                  double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const float r = tr[k] - obs[k];
        tr[k] = r;
        acc += (double)(r * r);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// the FP64 gradient rounded to FP32 (what the one-shot driver returns)
__global__ void narrow_b_kernel(float *__restrict__ out, const double *__restrict__ a, long long n)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        out[k] = (float)a[k];
}

// one shot's FP32 sub-grid gradient into the FP64 global gradient
__global__ void accum_crop_kernel(double *__restrict__ A11, double *__restrict__ A13, double *__restrict__ A33,
                                  const float *__restrict__ G11, const float *__restrict__ G13,
                                  const float *__restrict__ G33, long long f_off, long long c_off, int nnx, int nnz,
                                  int cpitch)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y;
    if (ix >= nnx || iz >= nnz) return;
    const long long i = f_off + ix + (long long)iz * nnx, c = c_off + ix + (long long)iz * cpitch;
    A11[c] += (double)G11[i];
    A13[c] += (double)G13[i];
    A33[c] += (double)G33[i];
}

// zero n 4-byte words with a kernel (what a captured graph may record,
// unlike hipMemsetAsync)
__global__ void fill0_kernel(unsigned *__restrict__ p, long long n)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = 0u;
}
void zero_words(void *p, size_t nwords, hipStream_t st)
{
    const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((nwords + 255) / 256, 4096));
    hipLaunchKernelGGL(fill0_kernel, dim3(nb), dim3(256), 0, st, (unsigned *)p, (long long)nwords);
}

// Slot 0 of each shot's snapshots, D of the zero level 0, which the forward
// pass never writes, and its 16-bit scales (two per slot for FP16/BF16, two
// per 64-point segment for S16): cleared in each wave's sequence, because
// the waves of a captured batch share the snapshot buffers.  Grid: y = shot.
__global__ void clear_slot0_kernel(float *__restrict__ snapf, unsigned short *__restrict__ snaph,
                                   float *__restrict__ scale, const vti::ShotB *__restrict__ sb, int nnz, int fmt)
{
    const vti::ShotB sh = sb[blockIdx.y];
    const long long plane = (long long)sh.nnx * nnz;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (fmt == 0) {
        for (long long k = t; k < 2 * plane; k += stride) snapf[sh.snap_off + k] = 0.f;
        return;
    }
    for (long long k = t; k < 2 * plane; k += stride) snaph[sh.snap_off + k] = 0;
    const long long nsc = fmt == 3 ? 2 * vti::s16_nseg(sh.nnx, nnz) : 2;
    for (long long k = t; k < nsc; k += stride) scale[sh.scale_off + k] = 0.f;
}

struct ShotIn {
    int c0, nx, ix, iz, nr, r0;
};

// The call's checked arguments and derived sizes.
struct BatchArgs {
    int ns, nr, NX, nz, nt, npad, NNX, nnz, ratio, nnt, nnt_data, levels, BYT, snap_fmt;
    bool half;
    float dx2, dz2, dt2;
    std::vector<ShotIn> in;
    std::vector<long long> data_off;
    long long shot_bytes(int s) const
    {
        // 8 wavefield planes, 3 gradient planes, the receiver map, the
        // snapshots (and for FP16/BF16 two staging planes and the slice
        // scales/maxima, for S16 the segment scales) and two trace copies
        const long long nxz = (long long)(in[s].nx + 2 * npad) * nnz;
        const long long snapb = !half ? 4LL * nxz * 2 * nnt
                                : snap_fmt == 3
                                    ? 2LL * nxz * 2 * nnt + 4LL * 2 * nnt * vti::s16_nseg(in[s].nx + 2 * npad, nnz)
                                    : 2LL * nxz * 2 * nnt + 4LL * nxz * 2 + 16LL * nnt;
        return 4LL * nxz * (8 + 3 + 1) + snapb + 8LL * nnt_data * in[s].nr + 64;
    }
};

int parse_args(BatchArgs &a, const int *shots, int ns, const int *rloc, int nr, int NX, int nz, int nt, int npad,
               float dx, float dz, float dt, float rate, int levels, int tile_rows, int snap_fmt)
{
    VTI_CHECK(ns >= 1 && nr >= 0 && NX > 0 && nz > 0 && npad >= 0 && nt >= 3 && dx > 0 && dz > 0 && dt > 0,
              "vti2d_objfunc_gradient_batch: bad sizes");
    if (levels == 0) levels = 2;
    if (tile_rows == 0) tile_rows = 16;
    VTI_CHECK(levels == 1 || levels == 2, "levels per launch must be 1 or 2, got %d", levels);
    VTI_CHECK(tile_rows == 4 || tile_rows == 8 || tile_rows == 16, "tile rows must be 4, 8 or 16, got %d", tile_rows);
    VTI_CHECK(snap_fmt >= 0 && snap_fmt <= 3,
              "snapshot format must be 0 (FP32), 1 (FP16), 2 (BF16) or 3 (S16), got %d", snap_fmt);
    a.ns = ns;
    a.nr = nr;
    a.NX = NX;
    a.nz = nz;
    a.nt = nt;
    a.npad = npad;
    a.NNX = NX + 2 * npad;
    a.nnz = nz + 2 * npad;
    a.levels = levels;
    a.BYT = tile_rows;
    a.snap_fmt = snap_fmt;
    a.half = snap_fmt != 0;
    VTI_CHECK((long long)a.NNX * a.nnz < (1LL << 30),
              "vti2d_objfunc_gradient_batch: padded model of %d x %d points exceeds the 2^30 the kernels index", a.nnz,
              a.NNX);
    a.ratio = (int)(rate / dt + 0.5f);  // objFuncGradient_f.cu:30 rounds in float
    VTI_CHECK(a.ratio >= 2 && (nt - 1) % a.ratio == 0,
              "the reference's snapshot rotation lines up only for rate/dt >= 2 and (nt-1) %% (rate/dt) == 0 "
              "(got ratio %d, nt %d)", a.ratio, nt);
    a.nnt = (nt - 1) / a.ratio + 1;
    a.nnt_data = a.nnt;
    a.dx2 = dx * dx;
    a.dz2 = dz * dz;
    a.dt2 = dt * dt;
    a.in.resize(ns);
    a.data_off.resize(ns);
    long long data_total = 0;
    const int nnz = a.nnz;
    for (int s = 0; s < ns; ++s) {
        ShotIn &q = a.in[s];
        q.c0 = shots[6 * s];
        q.nx = shots[6 * s + 1];
        q.ix = shots[6 * s + 2];
        q.iz = shots[6 * s + 3];
        q.nr = shots[6 * s + 4];
        q.r0 = shots[6 * s + 5];
        const int w = q.nx + 2 * npad;
        VTI_CHECK(q.nx > 0 && q.c0 >= 0 && q.c0 + w <= a.NNX,
                  "shot %d: sub-grid [%d, %d) outside the model's %d columns", s, q.c0, q.c0 + w, a.NNX);
        VTI_CHECK(w >= 9 && nnz >= 9, "shot %d: sub-grid too small", s);
        VTI_CHECK(q.ix >= 4 && q.ix < w - 4 && q.iz >= 4 && q.iz < nnz - 4, "shot %d: source outside the interior",
                  s);
        VTI_CHECK(q.nr >= 0 && q.r0 >= 0 && q.r0 + q.nr <= nr, "shot %d: receiver range outside rloc", s);
        std::vector<char> seen((size_t)w * nnz, 0);
        for (int k = 0; k < q.nr; ++k) {
            const int x = rloc[2 * (q.r0 + k)], z = rloc[2 * (q.r0 + k) + 1];
            VTI_CHECK(x >= 4 && x < w - 4 && z >= 4 && z < nnz - 4, "shot %d receiver %d outside the interior", s, k);
            VTI_CHECK(!seen[x + (size_t)z * w], "shot %d: two receivers share a grid point", s);
            seen[x + (size_t)z * w] = 1;
        }
        a.data_off[s] = data_total;
        data_total += (long long)a.nnt_data * q.nr;
    }
    return vti::kOk;
}

// Device state of a call shared by its waves: the model, the taper profile
// (t0[npad] = 1, the factor away from the edges) and the FP64 gradient.
constexpr int kNblkR = 1024;
struct BatchShared {
    DBuf<float> C[3], t0d;
    DBuf<double> A[3];
};

int setup_shared(BatchShared &S, const BatchArgs &a, hipStream_t st)
{
    const size_t gsize = (size_t)a.NNX * a.nnz;
    int rc = vti::kOk;
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = S.C[k].alloc(gsize, st);
        if (!rc) rc = S.A[k].alloc(gsize, st);
    }
    std::vector<float> t0(a.npad + 1);
    vti::taper_profile2d(t0.data(), a.npad);
    t0[a.npad] = 1.f;
    if (!rc) rc = S.t0d.alloc(t0.size(), st);
    if (!rc) VTI_HIP(hipMemcpyAsync(S.t0d.p, t0.data(), t0.size() * 4, hipMemcpyHostToDevice, st));
    if (!rc) VTI_HIP(hipStreamSynchronize(st));  // t0 is a host temporary
    return rc;
}

int upload_model(BatchShared &S, const BatchArgs &a, const float *c11, const float *c13, const float *c33,
                 hipStream_t st)
{
    const size_t gsize = (size_t)a.NNX * a.nnz;
    const float *c[3] = {c11, c13, c33};
    for (int k = 0; k < 3; ++k) VTI_HIP(hipMemcpyAsync(S.C[k].p, c[k], gsize * 4, hipMemcpyHostToDevice, st));
    return vti::kOk;
}

// One wave of shots on the device: descriptors and tile maps (sb/tile_shot
// for the one-level kernels' 64 x 8 tiles, sbt/tile_shot_t for the
// two-level ones' 64 x tile_rows), receiver maps, wavefields, snapshots,
// traces, the observed data and the residual's partial sums.
struct BatchWave {
    int s0 = 0, nw = 0, nrec = 0;
    long long f_total = 0, snap_total = 0, tr_total = 0, max_plane = 0, scale_total = 0;
    std::vector<ShotB> sb, sbt;
    std::vector<int> tile_shot, tile_shot_t;
    DBuf<ShotB> sbd, sbtd;
    DBuf<int> tsh, tsht, recmap, tsdd;
    DBuf<long long> fidxd, tbd;
    DBuf<float> P[4], Q[4], G[3], snapf, stage, scale, tr, obs;
    DBuf<unsigned short> snaph;
    DBuf<unsigned> amax;
    DBuf<double> part;
    // residual partial sums per shot (merged requests): the one-shot
    // driver's block count for each shot's traces, and their offsets in part
    std::vector<int> pblk, poff;
};

// the one-shot driver's residual grid for n trace samples
int shot_residual_blocks(long long n)
{
    return (int)std::min<long long>(2048, std::max<long long>(1, (n + 255) / 256));
}

// The wave's shot descriptors, tile maps and buffer sizes (host only).
void size_wave(BatchWave &W, const BatchArgs &a, int s0, int s1, const int *rloc)
{
    const int nw = s1 - s0, nnz = a.nnz, npad = a.npad, nnt = a.nnt;
    const int ntz = (nnz - 8 + kBY2 - 1) / kBY2, ntz_t = (nnz - 8 + a.BYT - 1) / a.BYT;
    W.s0 = s0;
    W.nw = nw;
    W.sb.assign(nw, ShotB{});
    W.sbt.assign(nw, ShotB{});
    W.tile_shot.clear();
    W.tile_shot_t.clear();
    W.f_total = W.snap_total = W.tr_total = W.max_plane = W.scale_total = 0;
    W.nrec = 0;
    for (int j = 0; j < nw; ++j) {
        const ShotIn &q = a.in[s0 + j];
        ShotB &b = W.sb[j];
        b.nnx = q.nx + 2 * npad;
        b.ntx = (b.nnx - 8 + 63) / 64;
        b.tile0 = (int)W.tile_shot.size();
        for (int t = 0; t < b.ntx * ntz; ++t) W.tile_shot.push_back(j);
        b.nr = q.nr;
        b.f_off = W.f_total;
        b.c_off = q.c0;
        b.snap_off = W.snap_total;
        b.r_off = W.tr_total;
        b.stage_off = 2 * W.f_total;
        b.src = q.ix + q.iz * b.nnx;
        b.rz0 = nnz;
        b.rz1 = -1;
        for (int r = 0; r < q.nr; ++r) {
            const int z = rloc[2 * (q.r0 + r) + 1];
            b.rz0 = std::min(b.rz0, z);
            b.rz1 = std::max(b.rz1, z);
        }
        // FP16/BF16: 2 slice scales per slot; S16: 2 * nseg segment scales per slot
        b.scale_off = a.snap_fmt == 3 ? W.scale_total : 2LL * nnt * j;
        W.sbt[j] = b;
        W.sbt[j].tile0 = (int)W.tile_shot_t.size();
        for (int t = 0; t < b.ntx * ntz_t; ++t) W.tile_shot_t.push_back(j);
        const long long plane = (long long)b.nnx * nnz;
        W.max_plane = std::max(W.max_plane, plane);
        W.f_total += plane;
        W.snap_total += (long long)nnt * 2 * plane;
        W.scale_total += (long long)nnt * 2 * vti::s16_nseg(b.nnx, nnz);
        W.tr_total += (long long)a.nnt_data * q.nr;
        W.nrec += q.nr;
    }
}

// Device buffers that waves run one after another can share: wavefields,
// gradients, snapshots (with their scales, and for FP16/BF16 the staging
// planes and slice maxima) and
// recorded traces, sized for the largest wave.  enqueue_wave clears
// whatever a wave reads before writing it, so a captured sequence of waves
// runs them all through one set.
struct WavePool {
    long long f = 0, snap = 0, scale = 0, amax = 0, tr = 1;
    DBuf<float> P[4], Q[4], G[3], snapf, stage, scale_b, tr_b;
    DBuf<unsigned short> snaph;
    DBuf<unsigned> amax_b;
};

long long wave_scale_count(const BatchWave &W, const BatchArgs &a)
{
    return a.snap_fmt == 3 ? W.scale_total : 2LL * a.nnt * W.nw;
}

void pool_fit(WavePool &pool, const BatchWave &W, const BatchArgs &a)
{
    pool.f = std::max(pool.f, W.f_total);
    pool.snap = std::max(pool.snap, W.snap_total);
    pool.scale = std::max(pool.scale, wave_scale_count(W, a));
    pool.amax = std::max(pool.amax, 2LL * a.nnt * W.nw);
    pool.tr = std::max(pool.tr, std::max<long long>(W.tr_total, 1));
}

long long pool_bytes(const WavePool &pool, const BatchArgs &a)
{
    const long long stage = a.snap_fmt == 3 ? 0 : 4 * 2 * pool.f;
    const long long snapb = a.half ? 2 * pool.snap + stage + 4 * pool.scale + 4 * pool.amax : 4 * pool.snap;
    return 4 * pool.f * 11 + snapb + 4 * pool.tr;
}

int pool_alloc(WavePool &pool, const BatchArgs &a, hipStream_t st)
{
    int rc = vti::kOk;
    for (int k = 0; k < 4 && !rc; ++k) {
        rc = pool.P[k].alloc((size_t)pool.f, st, false);
        if (!rc) rc = pool.Q[k].alloc((size_t)pool.f, st, false);
    }
    for (int k = 0; k < 3 && !rc; ++k) rc = pool.G[k].alloc((size_t)pool.f, st, false);
    if (a.half) {
        if (!rc) rc = pool.snaph.alloc((size_t)pool.snap, st, false);
        if (!rc && a.snap_fmt != 3) rc = pool.stage.alloc((size_t)(2 * pool.f), st, false);
        if (!rc) rc = pool.scale_b.alloc((size_t)pool.scale, st, false);
        if (!rc) rc = pool.amax_b.alloc((size_t)pool.amax, st, false);
    } else {
        if (!rc) rc = pool.snapf.alloc((size_t)pool.snap, st, false);
    }
    if (!rc) rc = pool.tr_b.alloc((size_t)pool.tr, st, false);
    return rc;
}

// Allocate (or, with a pool, view) the wave's device buffers and upload its
// tables, receiver map and observed data.
int fill_wave(BatchWave &W, const BatchArgs &a, const float *data, const int *rloc, hipStream_t st,
              const WavePool *pool)
{
    const int s0 = W.s0, nw = W.nw, nnt = a.nnt;
    const int nrec = W.nrec;
    std::vector<long long> fidx(nrec), tb(nrec);
    std::vector<int> tsd(nrec);
    std::vector<int> hmap((size_t)W.f_total, -1);
    for (int j = 0, k = 0; j < nw; ++j) {
        const ShotIn &q = a.in[s0 + j];
        for (int r = 0; r < q.nr; ++r, ++k) {
            const int x = rloc[2 * (q.r0 + r)], z = rloc[2 * (q.r0 + r) + 1];
            fidx[k] = W.sb[j].f_off + x + (long long)z * W.sb[j].nnx;
            tb[k] = W.sb[j].r_off + r;
            tsd[k] = q.nr;
            hmap[(size_t)fidx[k]] = r;
        }
    }
    std::vector<float> obs_h((size_t)std::max<long long>(W.tr_total, 1));
    for (int j = 0; j < nw; ++j)
        if (a.in[s0 + j].nr)
            std::memcpy(obs_h.data() + W.sb[j].r_off, data + a.data_off[s0 + j],
                        sizeof(float) * (size_t)a.nnt_data * a.in[s0 + j].nr);
    const long long f_total = W.f_total, tr_total = W.tr_total;
    int rc = W.sbd.alloc(nw, st);
    if (!rc) rc = W.sbtd.alloc(nw, st);
    if (!rc) rc = W.tsh.alloc(W.tile_shot.size(), st);
    if (!rc) rc = W.tsht.alloc(W.tile_shot_t.size(), st);
    if (!rc) rc = W.recmap.alloc((size_t)f_total, st);
    if (!rc) rc = W.tsdd.alloc(nrec, st);
    if (!rc) rc = W.fidxd.alloc(nrec, st);
    if (!rc) rc = W.tbd.alloc(nrec, st);
    // snapshot slots 1.. are written by the forward pass before the adjoint
    // reads them; slot 0 (D of the zero level 0) is never written and is
    // cleared by enqueue_wave, so the slots themselves need no clearing
    if (pool) {
        VTI_CHECK(pool->f >= f_total && pool->snap >= W.snap_total && pool->scale >= wave_scale_count(W, a) &&
                      pool->amax >= 2LL * nnt * nw && pool->tr >= tr_total,
                  "batch wave larger than its shared buffers");
        for (int k = 0; k < 4; ++k) {
            W.P[k].borrow(pool->P[k], (size_t)f_total);
            W.Q[k].borrow(pool->Q[k], (size_t)f_total);
        }
        for (int k = 0; k < 3; ++k) W.G[k].borrow(pool->G[k], (size_t)f_total);
        if (a.half) {
            W.snaph.borrow(pool->snaph, (size_t)W.snap_total);
            if (a.snap_fmt != 3) W.stage.borrow(pool->stage, (size_t)(2 * f_total));
            W.scale.borrow(pool->scale_b, (size_t)wave_scale_count(W, a));
            W.amax.borrow(pool->amax_b, (size_t)(2 * nnt * nw));
        } else {
            W.snapf.borrow(pool->snapf, (size_t)W.snap_total);
        }
        W.tr.borrow(pool->tr_b, (size_t)std::max<long long>(tr_total, 1));
    } else {
        for (int k = 0; k < 4 && !rc; ++k) {
            rc = W.P[k].alloc((size_t)f_total, st);
            if (!rc) rc = W.Q[k].alloc((size_t)f_total, st);
        }
        for (int k = 0; k < 3 && !rc; ++k) rc = W.G[k].alloc((size_t)f_total, st);
        if (a.half) {
            if (!rc) rc = W.snaph.alloc((size_t)W.snap_total, st, false);
            if (!rc && a.snap_fmt != 3) rc = W.stage.alloc((size_t)(2 * f_total), st);
            if (!rc) rc = W.scale.alloc((size_t)wave_scale_count(W, a), st);
            if (!rc) rc = W.amax.alloc((size_t)(2 * nnt * nw), st);
        } else {
            if (!rc) rc = W.snapf.alloc((size_t)W.snap_total, st, false);
        }
        if (!rc) rc = W.tr.alloc((size_t)std::max<long long>(tr_total, 1), st);
    }
    if (!rc) rc = W.obs.alloc((size_t)std::max<long long>(tr_total, 1), st);
    W.pblk.assign(nw, 0);
    W.poff.assign(nw, 0);
    int ptot = 0;
    for (int j = 0; j < nw; ++j) {
        W.pblk[j] = shot_residual_blocks((long long)a.nnt_data * a.in[s0 + j].nr);
        W.poff[j] = ptot;
        ptot += W.pblk[j];
    }
    if (!rc) rc = W.part.alloc(std::max(kNblkR, ptot), st);
    if (rc) return rc;
    VTI_HIP(hipMemcpyAsync(W.sbd.p, W.sb.data(), sizeof(ShotB) * nw, hipMemcpyHostToDevice, st));
    VTI_HIP(hipMemcpyAsync(W.sbtd.p, W.sbt.data(), sizeof(ShotB) * nw, hipMemcpyHostToDevice, st));
    VTI_HIP(hipMemcpyAsync(W.tsh.p, W.tile_shot.data(), sizeof(int) * W.tile_shot.size(), hipMemcpyHostToDevice, st));
    VTI_HIP(hipMemcpyAsync(W.tsht.p, W.tile_shot_t.data(), sizeof(int) * W.tile_shot_t.size(),
                           hipMemcpyHostToDevice, st));
    VTI_HIP(hipMemcpyAsync(W.recmap.p, hmap.data(), sizeof(int) * hmap.size(), hipMemcpyHostToDevice, st));
    if (nrec) {
        VTI_HIP(hipMemcpyAsync(W.tsdd.p, tsd.data(), sizeof(int) * nrec, hipMemcpyHostToDevice, st));
        VTI_HIP(hipMemcpyAsync(W.fidxd.p, fidx.data(), sizeof(long long) * nrec, hipMemcpyHostToDevice, st));
        VTI_HIP(hipMemcpyAsync(W.tbd.p, tb.data(), sizeof(long long) * nrec, hipMemcpyHostToDevice, st));
    }
    if (tr_total)
        VTI_HIP(hipMemcpyAsync(W.obs.p, obs_h.data(), sizeof(float) * tr_total, hipMemcpyHostToDevice, st));
    VTI_HIP(hipStreamSynchronize(st));  // the host staging vectors end here
    return vti::kOk;
}

int setup_wave(BatchWave &W, const BatchArgs &a, int s0, int s1, const float *data, const int *rloc, hipStream_t st,
               const WavePool *pool = nullptr)
{
    size_wave(W, a, s0, s1, rloc);
    return fill_wave(W, a, data, rloc, st, pool);
}

// Enqueue one wave's whole objective + gradient on `st`: the forward pass
// with snapshots and traces, the residual (its partial sums into W.part),
// the adjoint pass with the imaging condition and the shots' gradients
// added in FP64 into S.A, shot by shot.  Every buffer a pass reads before
// writing is cleared here by a kernel, so the sequence can be captured into a
// graph and replayed.  Host work is limited to computing launch arguments.
// Events that order one wave against the previous one when consecutive
// waves run on two streams: fwd_done is recorded once the wave's forward pass
// and residual are enqueued; the FP64 accumulation waits for acc_wait (the
// previous wave's accumulation, so the global sum keeps shot order) and
// records acc_done.
// A captured plan passes `split` instead: called where fwd_done would be
// recorded, it ends the capture of the wave's forward pass and begins that
// of its adjoint.
struct WaveSync {
    hipEvent_t fwd_done = nullptr, acc_wait = nullptr, acc_done = nullptr;
    std::function<int()> split;
};

// A second stream and two pairs of WaveSync events for one call, released
// after both streams are idle.
struct StreamPair {
    hipStream_t main = nullptr, side = nullptr;
    hipEvent_t fwd[2] = {nullptr, nullptr}, acc[2] = {nullptr, nullptr};
    int init(hipStream_t st)
    {
        main = st;
        VTI_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) {
            VTI_HIP(hipEventCreateWithFlags(&fwd[k], hipEventDisableTiming));
            VTI_HIP(hipEventCreateWithFlags(&acc[k], hipEventDisableTiming));
        }
        return vti::kOk;
    }
    ~StreamPair()
    {
        if (side) (void)hipStreamSynchronize(side);
        if (main) (void)hipStreamSynchronize(main);
        for (int k = 0; k < 2; ++k) {
            if (fwd[k]) (void)hipEventDestroy(fwd[k]);
            if (acc[k]) (void)hipEventDestroy(acc[k]);
        }
        if (side) (void)hipStreamDestroy(side);
    }
};

int enqueue_wave(BatchWave &W, BatchShared &S, const BatchArgs &a, const float *wavelet, hipStream_t st,
                 bool per_shot_obj = false, const WaveSync *sync = nullptr)
{
    const int nw = W.nw, nrec = W.nrec, nt = a.nt, ratio = a.ratio, nnt = a.nnt, nnz = a.nnz, BYT = a.BYT;
    const long long f_total = W.f_total, tr_total = W.tr_total;
    const float dx2 = a.dx2, dz2 = a.dz2, dt2 = a.dt2;
    const int levels = a.levels, snap_fmt = a.snap_fmt;
    const bool half = a.half;
    DBuf<float> *P = W.P, *Q = W.Q, *G = W.G, *C = S.C;
    for (int k = 0; k < 4; ++k) {
        zero_words(P[k].p, (size_t)f_total, st);
        zero_words(Q[k].p, (size_t)f_total, st);
    }
    for (int k = 0; k < 3; ++k) zero_words(G[k].p, (size_t)f_total, st);
    zero_words(W.tr.p, W.tr.n, st);  // row 0 is never recorded
    if (half && snap_fmt != 3) zero_words(W.amax.p, W.amax.n, st);
    // FP16/BF16: the forward kernels stage D at interior points only; the
    // rest of the staging planes stays 0 (waves of a captured batch share
    // these planes).  S16 has no staging: the kernels write its 16-bit
    // values and segment scales themselves.
    if (half && snap_fmt != 3) zero_words(W.stage.p, (size_t)(2 * f_total), st);
    hipLaunchKernelGGL(clear_slot0_kernel, dim3((unsigned)std::min<long long>((2 * W.max_plane + 255) / 256, 1024),
                                                (unsigned)nw),
                       dim3(256), 0, st, W.snapf.p, W.snaph.p, W.scale.p, W.sbd.p, nnz, snap_fmt);

    BatchGeo bg;
    bg.nnz = nnz;
    bg.npad = a.npad;
    bg.cpitch = a.NNX;
    bg.rx2 = vti::div_reciprocal(dx2);
    bg.rz2 = vti::div_reciprocal(dz2);
    bg.nnxz_total = f_total;
    const dim3 grid((unsigned)W.tile_shot.size()), block(64, kBY2);
    const dim3 gridt((unsigned)W.tile_shot_t.size()), blockt(64, BYT);
    const int rb = (nrec + 255) / 256;
    vti::SnapW sw{half ? W.stage.p : W.snapf.p, half && snap_fmt != 3 ? W.amax.p : nullptr, half};
    if (snap_fmt == 3) sw = vti::SnapW{nullptr, nullptr, false, W.snaph.p, W.scale.p, true};
    const vti::SnapR sr{W.snapf.p, W.snaph.p, W.scale.p};
    const dim3 gpack((unsigned)std::min<long long>((2 * W.max_plane + 255) / 256, 1024), (unsigned)nw);
    auto pack = [&](int slot) {  // FP16/BF16: the staged slot, once its slice maxima are known
        if (!half || snap_fmt == 3 || slot < 0) return;
        hipLaunchKernelGGL(vti::pack_snap_kernel, gpack, dim3(256), 0, st, W.snaph.p, W.scale.p, W.stage.p,
                           W.amax.p, W.sbd.p, nnz, slot, snap_fmt == 2 ? 1 : 0);
    };
    auto record = [&](int it, int buf) {
        if (it % ratio == 0 && nrec)
            hipLaunchKernelGGL(record_b_kernel, dim3(rb), dim3(256), 0, st, W.tr.p, P[buf].p, Q[buf].p, W.fidxd.p,
                               W.tbd.p, W.tsdd.p, nrec, it / ratio);
    };
    auto snap_slot = [&](int lc) { return lc % ratio == 0 ? lc / ratio : -1; };
    ShotB *sbd = W.sbd.p, *sbtd = W.sbtd.p;
    int *tsh = W.tsh.p, *tsht = W.tsht.p;
    float *t0d = S.t0d.p;

    // ---- forward (objFuncGradient_f.cu:160-210 for every shot).  Buffers
    // ip/ic hold the previous and current levels; s0b/s1b are free.
    int ip = 1, ic = 0, s0b = 2, s1b = 3;
    hipLaunchKernelGGL(set_src_b_kernel, dim3((nw + 63) / 64), dim3(64), 0, st, P[ic].p, Q[ic].p, sbd, nw,
                       dt2 * wavelet[0]);
    for (int it = 2; it < nt;) {
        if (levels == 2 && it + 1 < nt) {
            // levels it and it+1, reading levels it-1 and it
            const int sl1 = snap_slot(it - 1), sl2 = snap_slot(it);
#define FWD_T2(BYV)                                                                                                 \
    hipLaunchKernelGGL(vti::fwd2db_t2_kernel<BYV>, gridt, blockt, 0, st, P[s0b].p, Q[s0b].p, P[s1b].p, Q[s1b].p,     \
                       P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p, C[1].p, C[2].p, t0d, sbtd, tsht, bg, dx2, dz2, \
                       dt2, dt2 * wavelet[it - 1], dt2 * wavelet[it], sw, sl1, sl2)
            if (BYT == 4) FWD_T2(4);
            else if (BYT == 8) FWD_T2(8);
            else FWD_T2(16);
#undef FWD_T2
            pack(sl1 >= 0 ? sl1 : sl2);
            record(it, s0b);
            record(it + 1, s1b);
            std::swap(ip, s0b);
            std::swap(ic, s1b);
            it += 2;
        } else {
            const int sl = snap_slot(it - 1);
            hipLaunchKernelGGL(vti::fwd2db_kernel, grid, block, 0, st, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p,
                               C[1].p, C[2].p, t0d, sbd, tsh, bg, dx2, dz2, dt2, dt2 * wavelet[it - 1], sw, sl);
            pack(sl);
            std::swap(ip, ic);
            record(it, ic);
            it += 1;
        }
    }
    hipLaunchKernelGGL(vti::dsnap2db_kernel, grid, block, 0, st, P[ic].p, Q[ic].p, sbd, tsh, bg, dx2, dz2, sw,
                       nnt - 1);
    pack(nnt - 1);

    // ---- residual and its partial sums (objFuncGradient_f.cu:214-219)
    if (per_shot_obj) {
        // each shot's residual on its own, with the one-shot driver's grid, so
        // that its partial sums are that driver's
        for (int j = 0; j < nw; ++j)
            if (W.sb[j].nr)
                hipLaunchKernelGGL(residual_b_kernel, dim3(W.pblk[j]), dim3(256), 0, st, W.tr.p + W.sb[j].r_off,
                                   W.obs.p + W.sb[j].r_off, (long long)a.nnt_data * W.sb[j].nr, W.part.p + W.poff[j]);
    } else if (tr_total) {
        hipLaunchKernelGGL(residual_b_kernel, dim3(kNblkR), dim3(256), 0, st, W.tr.p, W.obs.p, tr_total, W.part.p);
    }

    if (sync && sync->fwd_done) VTI_HIP(hipEventRecord(sync->fwd_done, st));
    if (sync && sync->split) {
        const int rc = sync->split();
        if (rc) return rc;
    }

    // ---- backward (objFuncGradient_f.cu:225-300 for every shot)
    for (int k = 0; k < 4; ++k) {
        zero_words(P[k].p, (size_t)f_total, st);
        zero_words(Q[k].p, (size_t)f_total, st);
    }
    ic = 0;
    ip = 1;
    s0b = 2;
    s1b = 3;
    if (nrec)
        hipLaunchKernelGGL(inject_first_b_kernel, dim3(rb), dim3(256), 0, st, P[ic].p, Q[ic].p, W.tr.p, W.fidxd.p,
                           W.tbd.p, W.tsdd.p, nrec, a.nnt_data - 1, dt2);
    {
        float f = (float)((nt - 2.) / ratio);
        const int i1 = (int)f;
        f = f - i1;
#define XCOR(SFV)                                                                                                  \
    hipLaunchKernelGGL(vti::xcor2db_kernel<SFV>, grid, block, 0, st, P[ic].p, Q[ic].p, sbd, tsh, bg, sr, i1, f, \
                       G[0].p, G[1].p, G[2].p)
        if (snap_fmt == 0) XCOR(0);
        else if (snap_fmt == 1) XCOR(1);
        else if (snap_fmt == 2) XCOR(2);
        else XCOR(3);
#undef XCOR
    }
    // trace rows and snapshot slots bracketing adjoint level it
    auto level = [&](int it) {
        vti::LevelB l;
        float fr = (float)((it + 1.) / ratio);
        l.row = (int)fr;
        l.fr = fr - l.row;
        float f = (float)it / ratio;
        l.slot = (int)f;
        l.f = f - l.slot;
        return l;
    };
    int *recmap = W.recmap.p;
    float *tr = W.tr.p;
    for (int it = nt - 3; it >= 0;) {
        if (levels == 2 && it >= 1) {
            const vti::LevelB l1 = level(it), l2 = level(it - 1);
#define BWD_T2(BYV, SFV)                                                                                         \
    vti::bwd2db_t2_kernel<BYV, SFV><<<gridt, blockt, 0, st>>>(P[s0b].p, Q[s0b].p, P[s1b].p, Q[s1b].p, P[ip].p,  \
                                                              Q[ip].p, P[ic].p, Q[ic].p, C[0].p, C[1].p, C[2].p, \
                                                              t0d, sbtd, tsht, bg, recmap, tr, l1, l2, sr, G[0].p, \
                                                              G[1].p, G[2].p, dx2, dz2, dt2)
#define BWD_T2_SF(BYV)                      \
    if (snap_fmt == 0) BWD_T2(BYV, 0);      \
    else if (snap_fmt == 1) BWD_T2(BYV, 1); \
    else if (snap_fmt == 2) BWD_T2(BYV, 2); \
    else BWD_T2(BYV, 3)
            if (BYT == 4) {
                BWD_T2_SF(4);
            } else if (BYT == 8) {
                BWD_T2_SF(8);
            } else {
                BWD_T2_SF(16);
            }
#undef BWD_T2_SF
#undef BWD_T2
            std::swap(ip, s0b);
            std::swap(ic, s1b);
            it -= 2;
        } else {
            const vti::LevelB l = level(it);
#define BWD1(SFV)                                                                                                  \
    hipLaunchKernelGGL(vti::bwd2db_kernel<SFV>, grid, block, 0, st, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p,    \
                       C[1].p, C[2].p, t0d, sbd, tsh, bg, recmap, tr, l.row, l.fr, sr, l.slot, l.f, G[0].p, G[1].p, \
                       G[2].p, dx2, dz2, dt2)
            if (snap_fmt == 0) BWD1(0);
            else if (snap_fmt == 1) BWD1(1);
            else if (snap_fmt == 2) BWD1(2);
            else BWD1(3);
#undef BWD1
            std::swap(ip, ic);
            it -= 1;
        }
    }
    // ---- per-shot FP32 gradients into the FP64 global sum, in shot order
    if (sync && sync->acc_wait) VTI_HIP(hipStreamWaitEvent(st, sync->acc_wait, 0));
    for (int j = 0; j < nw; ++j)
        hipLaunchKernelGGL(accum_crop_kernel, dim3((W.sb[j].nnx + 255) / 256, nnz), dim3(256), 0, st, S.A[0].p,
                           S.A[1].p, S.A[2].p, G[0].p, G[1].p, G[2].p, W.sb[j].f_off, W.sb[j].c_off, W.sb[j].nnx, nnz,
                           a.NNX);
    if (sync && sync->acc_done) VTI_HIP(hipEventRecord(sync->acc_done, st));
    VTI_HIP(hipGetLastError());
    return vti::kOk;
}

// Enqueue one wave of synthetic_f: the forward pass of enqueue_wave with
// synthetic_f's level 1 and no snapshots, recording every rate/dt levels.
int enqueue_model_wave(BatchWave &W, BatchShared &S, const BatchArgs &a, const float *wavelet, hipStream_t st)
{
    const int nw = W.nw, nrec = W.nrec, nt = a.nt, ratio = a.ratio, BYT = a.BYT;
    const long long f_total = W.f_total;
    const float dx2 = a.dx2, dz2 = a.dz2, dt2 = a.dt2;
    DBuf<float> *P = W.P, *Q = W.Q, *C = S.C;
    for (int k = 0; k < 4; ++k) {
        zero_words(P[k].p, (size_t)f_total, st);
        zero_words(Q[k].p, (size_t)f_total, st);
    }
    zero_words(W.tr.p, W.tr.n, st);
    BatchGeo bg;
    bg.nnz = a.nnz;
    bg.npad = a.npad;
    bg.cpitch = a.NNX;
    bg.rx2 = vti::div_reciprocal(dx2);
    bg.rz2 = vti::div_reciprocal(dz2);
    bg.nnxz_total = f_total;
    const dim3 grid((unsigned)W.tile_shot.size()), block(64, kBY2);
    const dim3 gridt((unsigned)W.tile_shot_t.size()), blockt(64, BYT);
    const int rb = (nrec + 255) / 256;
    const vti::SnapW sw{nullptr, nullptr};  // no snapshots (every slot argument is -1)
    auto record = [&](int it, int buf) {
        if (it % ratio == 0 && nrec)
            hipLaunchKernelGGL(record_b_kernel, dim3(rb), dim3(256), 0, st, W.tr.p, P[buf].p, Q[buf].p, W.fidxd.p,
                               W.tbd.p, W.tsdd.p, nrec, it / ratio);
    };
    ShotB *sbd = W.sbd.p, *sbtd = W.sbtd.p;
    int *tsh = W.tsh.p, *tsht = W.tsht.p;
    float *t0d = S.t0d.p;
    int ip = 1, ic = 0, s0b = 2, s1b = 3;
    hipLaunchKernelGGL(set_src_taper_b_kernel, dim3((nw + 63) / 64), dim3(64), 0, st, P[ic].p, Q[ic].p, sbd, nw,
                       dt2 * wavelet[0], t0d, a.npad, a.nnz);
    for (int it = 2; it < nt;) {
        if (a.levels == 2 && it + 1 < nt) {
#define FWD_T2(BYV)                                                                                                 \
    hipLaunchKernelGGL(vti::fwd2db_t2_kernel<BYV>, gridt, blockt, 0, st, P[s0b].p, Q[s0b].p, P[s1b].p, Q[s1b].p,     \
                       P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p, C[1].p, C[2].p, t0d, sbtd, tsht, bg, dx2, dz2, \
                       dt2, dt2 * wavelet[it - 1], dt2 * wavelet[it], sw, -1, -1)
            if (BYT == 4) FWD_T2(4);
            else if (BYT == 8) FWD_T2(8);
            else FWD_T2(16);
#undef FWD_T2
            record(it, s0b);
            record(it + 1, s1b);
            std::swap(ip, s0b);
            std::swap(ic, s1b);
            it += 2;
        } else {
            hipLaunchKernelGGL(vti::fwd2db_kernel, grid, block, 0, st, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p,
                               C[1].p, C[2].p, t0d, sbd, tsh, bg, dx2, dz2, dt2, dt2 * wavelet[it - 1], sw, -1);
            std::swap(ip, ic);
            record(it, ic);
            it += 1;
        }
    }
    VTI_HIP(hipGetLastError());
    return vti::kOk;
}

// the objective's share of a wave: the residual's partial sums, in order
// Adds a wave's residual partial sums, shot by shot and block by block
// (enqueue_wave with per_shot_obj), to obj[owner[s]] for batch shot s
// (obj[0] when owner is null), so that the objective does not depend on how
// the shots were split into waves.
int wave_objective(const BatchWave &W, hipStream_t st, const int *owner, double *obj)
{
    if (!W.tr_total) return vti::kOk;
    std::vector<double> hp(W.part.n);
    VTI_HIP(hipMemcpyAsync(hp.data(), W.part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
    VTI_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < W.nw; ++j)
        if (W.sb[j].nr) {
            double &o = obj[owner ? owner[W.s0 + j] : 0];
            for (int b = 0; b < W.pblk[j]; ++b) o += hp[W.poff[j] + b];
        }
    return vti::kOk;
}

int download_gradient(const BatchShared &S, const BatchArgs &a, double *gc11, double *gc13, double *gc33,
                      hipStream_t st)
{
    const size_t gsize = (size_t)a.NNX * a.nnz;
    double *out[3] = {gc11, gc13, gc33};
    for (int k = 0; k < 3; ++k) VTI_HIP(hipMemcpyAsync(out[k], S.A[k].p, gsize * 8, hipMemcpyDeviceToHost, st));
    VTI_HIP(hipStreamSynchronize(st));
    return vti::kOk;
}

// The shots [s0, s1) of each wave: consecutive shots, at most max_wave of
// them (when > 0) and within `budget` bytes (at least one shot per wave).
std::vector<std::pair<int, int>> split_waves(const BatchArgs &a, int max_wave, long long budget)
{
    std::vector<std::pair<int, int>> w;
    for (int s0 = 0; s0 < a.ns;) {
        long long used = 0;
        int s1 = s0;
        while (s1 < a.ns && (max_wave <= 0 || s1 - s0 < max_wave) && used + a.shot_bytes(s1) <= budget)
            used += a.shot_bytes(s1++);
        if (s1 == s0) s1 = s0 + 1;
        w.emplace_back(s0, s1);
        s0 = s1;
    }
    return w;
}

// Runs the waves of a batch, two in flight when there are several.  Wave k
// runs on stream k % 2 (st, or a second one made for the call) and starts its
// forward pass when wave k-1's is done, so that it overlaps wave k-1's
// adjoint (the forward is bound by VALU work, the adjoint waits on memory),
// and the host sets wave k+1 up meanwhile.  The FP64 accumulations and the
// objective keep wave order, so the result is that of one wave after the
// other, bit for bit.  Objective partial sums go to obj as wave_objective
// adds them.
int run_waves(BatchShared &S, const BatchArgs &a, const std::vector<std::pair<int, int>> &waves,
              const float *data, const int *rloc, const float *wavelet, hipStream_t st, const int *owner,
              double *obj)
{
    int rc = vti::kOk;
    const size_t n = waves.size();
    StreamPair sp;
    if (n > 1) rc = sp.init(st);
    auto stream_of = [&](size_t k) { return k % 2 ? sp.side : st; };
    // Wave k+1 is set up (its tables built on the host, then uploaded on its
    // stream behind wave k-1) while waves k-1 and k run, so that it is
    // enqueued as soon as wave k-1 is done: three waves hold memory at once.
    std::unique_ptr<BatchWave> prev, cur;
    if (!rc && n) {
        cur = std::make_unique<BatchWave>();
        rc = setup_wave(*cur, a, waves[0].first, waves[0].second, data, rloc, st);
    }
    for (size_t k = 0; k < n && !rc; ++k) {
        hipStream_t sk = stream_of(k);
        WaveSync sync;
        if (n > 1) {
            sync.fwd_done = sp.fwd[k % 2];
            sync.acc_done = sp.acc[k % 2];
            if (k >= 1) {
                sync.acc_wait = sp.acc[(k - 1) % 2];
                if (hipStreamWaitEvent(sk, sp.fwd[(k - 1) % 2], 0) != hipSuccess) {
                    vti::set_error("batched 2D waves: hipStreamWaitEvent failed");
                    rc = vti::kErrHip;
                }
            }
        }
        if (!rc) rc = enqueue_wave(*cur, S, a, wavelet, sk, true, n > 1 ? &sync : nullptr);
        std::unique_ptr<BatchWave> next;
        if (!rc && k + 1 < n) {
            next = std::make_unique<BatchWave>();
            // (on stream k+1: waits for wave k-1, which ran there)
            rc = setup_wave(*next, a, waves[k + 1].first, waves[k + 1].second, data, rloc, stream_of(k + 1));
        }
        if (!rc && prev) rc = wave_objective(*prev, stream_of(k - 1), owner, obj);  // wave k-1 done
        prev = std::move(cur);
        cur = std::move(next);
    }
    if (!rc && prev) rc = wave_objective(*prev, stream_of(n - 1), owner, obj);
    return rc;  // (sp's destructor waits for both streams before it releases them)
}

// Shots per wave whose adjoint working set stays in the 256 MB Infinity
// Cache: per shot and point 16 FP32 planes (four buffers of p and q, the
// three gradients, the two snapshot slots a level reads, the receiver map),
// 192 MB in all, at the shots' mean sub-grid size.  At full length a wave
// that fits runs about 12% faster than one that does not; the short timed
// trials of the tuner cannot see this, so the wave size is modelled.
int cache_wave(const int *shots, int ns, int npad, int nz)
{
    double mean = 0.0;
    for (int s = 0; s < ns; ++s) mean += (double)(shots[6 * s + 1] + 2 * npad) * (nz + 2 * npad);
    mean /= std::max(ns, 1);
    const int w = (int)(192e6 / (64.0 * std::max(mean, 1.0)));
    return std::max(1, std::min(w, ns));
}

}  // namespace

// A captured batch: each wave's own tables, receiver map and observed data,
// two sets of wavefield/snapshot/trace buffers (even waves share one, odd
// waves the other, each sized for its largest wave), and two graphs per
// wave, its forward pass and its adjoint.  A replay runs wave k on stream
// k % 2 (the context's, or the plan's second stream) as the direct call
// does: the forward pass of wave k after that of wave k-1, beside wave
// k-1's adjoint, and the adjoint of wave k after that of wave k-1, so the
// FP64 accumulations keep wave order.
struct vti2d_batch_plan {
    vti_ctx *ctx = nullptr;
    BatchArgs a;
    BatchShared S;
    WavePool pool[2];
    std::vector<std::unique_ptr<BatchWave>> W;
    std::vector<hipGraphExec_t> fwd, adj;
    hipStream_t side = nullptr;
    hipEvent_t ev_fwd[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr}, ev_join = nullptr;
};

extern "C" {

// Batched objective + gradient (see vti_amd.h).
int vti2d_objfunc_gradient_batch(vti_ctx *ctx, double *objfunc, double *gc11, double *gc13, double *gc33,
                                 const float *data, const float *c11, const float *c13, const float *c33,
                                 const float *wavelet, const int *shots, int ns, const int *rloc, int nr, int NX, int nz,
                                 int nt, int npad, float dx, float dz, float dt, float rate, int max_wave, int levels,
                                 int tile_rows, int snap_fmt)
{
    VTI_CHECK(ctx && objfunc && gc11 && gc13 && gc33 && data && c11 && c13 && c33 && wavelet && shots && rloc,
              "vti2d_objfunc_gradient_batch: null argument");
    BatchArgs a;
    int rc = parse_args(a, shots, ns, rloc, nr, NX, nz, nt, npad, dx, dz, dt, rate, levels, tile_rows, snap_fmt);
    if (rc) return rc;
    VTI_HIP(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipStream_t st = ctx->stream;
    BatchShared S;
    rc = setup_shared(S, a, st);
    if (!rc) rc = upload_model(S, a, c11, c13, c33, st);
    if (rc) return rc;

    // Waves within the memory budget, two in flight (run_waves).  Each takes
    // half of max_wave (rounded up), so the two in flight hold about
    // max_wave shots, the number whose working set the cache model sizes: at
    // cfg4's 256 shots, waves of 4 overlapped gave 69.4 shot-gradients/s
    // against 64.7-65.5 for waves of 7 one after the other (waves of 7
    // overlapped: 43.6-60.2).
    const size_t mfree = vti::device_free_bytes();
    const long long budget = ((long long)mfree - (long long)(2ULL << 30)) / 3;  // three waves set up at once
    VTI_CHECK(budget > a.shot_bytes(0), "not enough device memory for one shot");
    const auto waves = split_waves(a, max_wave > 0 ? (max_wave + 1) / 2 : 0, budget);
    double obj = 0.0;
    rc = run_waves(S, a, waves, data, rloc, wavelet, st, nullptr, &obj);
    if (rc) return rc;
    rc = download_gradient(S, a, gc11, gc13, gc33, st);
    if (rc) return rc;
    *objfunc = 0.5 * obj;
    return vti::kOk;
}

// Captured batch (see vti_amd.h).
int vti2d_batch_plan_create(vti_ctx *ctx, vti2d_batch_plan **plan, const float *data, const float *wavelet,
                            const int *shots, int ns, const int *rloc, int nr, int NX, int nz, int nt, int npad,
                            float dx, float dz, float dt, float rate, int max_wave, int levels, int tile_rows,
                            int snap_fmt)
{
    VTI_CHECK(ctx && plan && data && wavelet && shots && rloc, "vti2d_batch_plan_create: null argument");
    *plan = nullptr;
    auto *pl = new vti2d_batch_plan;
    pl->ctx = ctx;
    int rc = parse_args(pl->a, shots, ns, rloc, nr, NX, nz, nt, npad, dx, dz, dt, rate, levels, tile_rows, snap_fmt);
    hipStream_t st = ctx->stream;
    if (!rc) {
        const hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) {
            vti::set_error("hipSetDevice: %s", hipGetErrorString(e));
            rc = vti::kErrHip;
        }
    }
    if (!rc) {
        // waves within the memory budget, sized on the host, two in flight
        // as in the direct call (half of max_wave each); each buffer set
        // fits the largest wave that uses it, and each wave adds its own
        // tables and data
        const long long budget = (long long)vti::device_free_bytes() - (long long)(2ULL << 30);
        for (const auto &w : split_waves(pl->a, max_wave > 0 ? (max_wave + 1) / 2 : 0, budget / 2)) {
            pl->W.emplace_back(new BatchWave);
            size_wave(*pl->W.back(), pl->a, w.first, w.second, rloc);
            pool_fit(pl->pool[(pl->W.size() - 1) % 2], *pl->W.back(), pl->a);
        }
        long long need = pool_bytes(pl->pool[0], pl->a) + (pl->W.size() > 1 ? pool_bytes(pl->pool[1], pl->a) : 0);
        for (const auto &W : pl->W)
            need += 4 * W->f_total + 4 * std::max<long long>(W->tr_total, 1) + 20LL * W->nrec +
                    4LL * (long long)(W->tile_shot.size() + W->tile_shot_t.size()) + (1LL << 16);
        if (need > budget) {
            vti::set_error("vti2d_batch_plan_create: the %d shots in %d waves need %.1f GB, %.1f GB free", ns,
                           (int)pl->W.size(), need / 1e9, (budget + (2LL << 30)) / 1e9);
            rc = vti::kErrArg;
        }
    }
    if (!rc) rc = setup_shared(pl->S, pl->a, st);
    for (int k = 0; k < 2 && k < (int)pl->W.size(); ++k)
        if (!rc) rc = pool_alloc(pl->pool[k], pl->a, st);
    for (size_t k = 0; k < pl->W.size() && !rc; ++k)
        rc = fill_wave(*pl->W[k], pl->a, data, rloc, st, &pl->pool[k % 2]);
    if (!rc && pl->W.size() > 1) {
        if (hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking) != hipSuccess) rc = vti::kErrHip;
        for (int k = 0; k < 2 && !rc; ++k)
            if (hipEventCreateWithFlags(&pl->ev_fwd[k], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&pl->ev_adj[k], hipEventDisableTiming) != hipSuccess)
                rc = vti::kErrHip;
        if (!rc && hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming) != hipSuccess) rc = vti::kErrHip;
        if (rc) vti::set_error("vti2d_batch_plan_create: stream or event creation failed");
    }
    // each wave captured on its own, on one stream, as two graphs
    for (size_t k = 0; k < pl->W.size() && !rc; ++k) {
        hipGraph_t g[2] = {nullptr, nullptr};
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        bool capturing = e == hipSuccess;
        WaveSync sync;
        sync.split = [&]() -> int {
            capturing = false;
            hipError_t e1 = hipStreamEndCapture(st, &g[0]);
            if (e1 == hipSuccess) e1 = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e1 != hipSuccess) {
                vti::set_error("vti2d_batch_plan_create: graph capture: %s", hipGetErrorString(e1));
                return vti::kErrHip;
            }
            capturing = true;
            return vti::kOk;
        };
        if (capturing) rc = enqueue_wave(*pl->W[k], pl->S, pl->a, wavelet, st, true, &sync);
        if (capturing) {
            const hipError_t e2 = hipStreamEndCapture(st, &g[1]);
            if (e == hipSuccess) e = e2;
        }
        hipGraphExec_t x[2] = {nullptr, nullptr};
        for (int j = 0; j < 2 && !rc && e == hipSuccess; ++j) e = hipGraphInstantiate(&x[j], g[j], nullptr, nullptr, 0);
        for (int j = 0; j < 2; ++j)
            if (g[j]) (void)hipGraphDestroy(g[j]);
        pl->fwd.push_back(x[0]);
        pl->adj.push_back(x[1]);
        if (!rc && e != hipSuccess) {
            vti::set_error("vti2d_batch_plan_create: graph capture: %s", hipGetErrorString(e));
            rc = vti::kErrHip;
        }
    }
    if (rc) {
        vti2d_batch_plan_destroy(pl);
        return rc;
    }
    *plan = pl;
    return vti::kOk;
}

int vti2d_batch_plan_run(vti2d_batch_plan *plan, double *objfunc, double *gc11, double *gc13, double *gc33,
                         const float *c11, const float *c13, const float *c33)
{
    VTI_CHECK(plan && objfunc && gc11 && gc13 && gc33 && c11 && c13 && c33, "vti2d_batch_plan_run: null argument");
    VTI_HIP(hipSetDevice(plan->ctx->device));
    hipStream_t st = plan->ctx->stream;
    int rc = upload_model(plan->S, plan->a, c11, c13, c33, st);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) zero_words(plan->S.A[k].p, 2 * plan->S.A[k].n, st);
    const size_t nw = plan->W.size();
    if (nw > 1) {  // the second stream starts after the upload
        VTI_HIP(hipEventRecord(plan->ev_join, st));
        VTI_HIP(hipStreamWaitEvent(plan->side, plan->ev_join, 0));
    }
    for (size_t k = 0; k < nw; ++k) {
        hipStream_t sk = k % 2 ? plan->side : st;
        if (k >= 1) VTI_HIP(hipStreamWaitEvent(sk, plan->ev_fwd[(k - 1) % 2], 0));
        VTI_HIP(hipGraphLaunch(plan->fwd[k], sk));
        if (nw > 1) VTI_HIP(hipEventRecord(plan->ev_fwd[k % 2], sk));
        if (k >= 1) VTI_HIP(hipStreamWaitEvent(sk, plan->ev_adj[(k - 1) % 2], 0));
        VTI_HIP(hipGraphLaunch(plan->adj[k], sk));
        if (nw > 1) VTI_HIP(hipEventRecord(plan->ev_adj[k % 2], sk));
    }
    if (nw > 1) {  // the context's stream waits for the last adjoint
        VTI_HIP(hipStreamWaitEvent(st, plan->ev_adj[(nw - 1) % 2], 0));
    }
    double obj = 0.0;
    for (const auto &W : plan->W)
        if (!rc) rc = wave_objective(*W, st, nullptr, &obj);
    if (!rc) rc = download_gradient(plan->S, plan->a, gc11, gc13, gc33, st);
    if (rc) return rc;
    *objfunc = 0.5 * obj;
    return vti::kOk;
}

int vti2d_batch_plan_destroy(vti2d_batch_plan *plan)
{
    if (!plan) return vti::kOk;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);  // no replay still running
    if (plan->side) (void)hipStreamSynchronize(plan->side);
    for (auto x : plan->fwd)
        if (x) (void)hipGraphExecDestroy(x);
    for (auto x : plan->adj)
        if (x) (void)hipGraphExecDestroy(x);
    for (int k = 0; k < 2; ++k) {
        if (plan->ev_fwd[k]) (void)hipEventDestroy(plan->ev_fwd[k]);
        if (plan->ev_adj[k]) (void)hipEventDestroy(plan->ev_adj[k]);
    }
    if (plan->ev_join) (void)hipEventDestroy(plan->ev_join);
    if (plan->side) (void)hipStreamDestroy(plan->side);
    delete plan;  // the buffers go back to the allocator on the context's stream
    return vti::kOk;
}

}  // extern "C"

namespace vti {

// Whether vti2d_objfunc_gradient's arguments can run through grad2d_merged:
// the batched driver's schedule conditions (rate/dt >= 2, (nt-1) a multiple
// of it, ot = 0), npad >= 4 as the one-shot drivers require, the standard
// init_abc taper (compared bit for bit),
// sources and receivers in the interior, one receiver per point per shot.
bool grad2d_batchable(const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                      int npad, float dt, float rate, float ot)
{
    if (!sloc || !rloc || !taper || ns < 1 || nx <= 0 || nz <= 0 || npad < 4 || nt < 3 || !(dt > 0) || ot != 0.f)
        return false;
    const int ratio = (int)(rate / dt + 0.5f);
    if (ratio < 2 || (nt - 1) % ratio != 0) return false;
    const int nnx = nx + 2 * npad, nnz = nz + 2 * npad;
    if (nnx < 9 || nnz < 9 || (long long)nnx * nnz >= (1LL << 26)) return false;
    std::vector<float> t((size_t)nnx * nnz);
    vti::init_taper2d(t.data(), nx, nz, npad);
    if (std::memcmp(t.data(), taper, t.size() * sizeof(float)) != 0) return false;
    // one receiver per grid point per shot: the shot's linear receiver
    // indices, sorted, have no repeats (O(nr log nr) per shot)
    std::vector<long long> idx;
    for (int k = 0; k < ns; ++k) {
        const int ix = sloc[4 * k], iz = sloc[4 * k + 1], nri = sloc[4 * k + 2], r0 = sloc[4 * k + 3];
        if (nri < 0 || r0 < 0 || r0 + nri > nr) return false;
        if (ix < 4 || ix >= nnx - 4 || iz < 4 || iz >= nnz - 4) return false;
        idx.clear();
        for (int ir = 0; ir < nri; ++ir) {
            const int x = rloc[2 * (r0 + ir)], z = rloc[2 * (r0 + ir) + 1];
            if (x < 4 || x >= nnx - 4 || z < 4 || z >= nnz - 4) return false;
            idx.push_back(x + (long long)z * nnx);
        }
        std::sort(idx.begin(), idx.end());
        if (std::adjacent_find(idx.begin(), idx.end()) != idx.end()) return false;
    }
    return true;
}

// synthetic_f (vti2d_synthetic) for several shots as one batch: the forward
// pass of the batched driver with synthetic_f's level 1 (the source times
// the taper) and no snapshots.  Each shot's traces are the one-shot loop's,
// bit for bit; rows the loop never records (row 0) keep their input values.
// The caller has checked grad2d_batchable and synthetic_f's own rounding of
// rate/dt.
int synth2d_batched(vti_ctx *ctx, float *out, const float *c11, const float *c13, const float *c33,
                    const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, int nx, int nz, int nt,
                    int npad, float dx, float dz, float dt, float rate)
{
    VTI_CHECK(ctx && out && c11 && c13 && c33 && wavelet && sloc && rloc && ns >= 1, "synth2d_batched: bad arguments");
    const int nnx = nx + 2 * npad, nnz = nz + 2 * npad;
    const int ratio = (int)(rate / dt + 0.5f);
    const int nnt = (nt - 1) / ratio + 1;
    std::vector<int> shots, rl;
    for (int k = 0; k < ns; ++k) {
        const int ix = sloc[4 * k], iz = sloc[4 * k + 1], nri = sloc[4 * k + 2], r0 = sloc[4 * k + 3];
        VTI_CHECK(nri >= 0 && r0 >= 0 && r0 + nri <= nr, "shot %d: receiver range outside rloc", k);
        shots.insert(shots.end(), {0, nx, ix, iz, nri, (int)(rl.size() / 2)});
        for (int ir = 0; ir < nri; ++ir) {
            rl.push_back(rloc[2 * (r0 + ir)]);
            rl.push_back(rloc[2 * (r0 + ir) + 1]);
        }
    }
    BatchArgs a;
    int rc = parse_args(a, shots.data(), ns, rl.data(), (int)(rl.size() / 2), nx, nz, nt, npad, dx, dz, dt, rate, 2,
                        16, 0);
    if (rc) return rc;
    VTI_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    BatchShared S;
    rc = setup_shared(S, a, st);
    if (rc) return rc;
    const float *cs[3] = {c11, c13, c33};
    for (int k = 0; k < 3; ++k)
        VTI_HIP(hipMemcpyAsync(S.C[k].p, cs[k], (size_t)nnx * nnz * 4, hipMemcpyDefault, st));
    const std::vector<float> nodata((size_t)std::max<long long>((long long)nnt * (rl.size() / 2), 1), 0.f);
    const int cw = cache_wave(shots.data(), ns, npad, nz);
    const long long budget = (long long)vti::device_free_bytes() - (long long)(2ULL << 30);
    for (const auto &wv : split_waves(a, cw, budget)) {
        BatchWave W;
        rc = setup_wave(W, a, wv.first, wv.second, nodata.data(), rl.data(), st);
        if (!rc) rc = enqueue_model_wave(W, S, a, wavelet, st);
        if (rc) return rc;
        std::vector<float> h((size_t)std::max<long long>(W.tr_total, 1));
        if (W.tr_total) VTI_HIP(hipMemcpyAsync(h.data(), W.tr.p, W.tr_total * 4, hipMemcpyDeviceToHost, st));
        VTI_HIP(hipStreamSynchronize(st));
        for (int j = 0; j < W.nw; ++j) {
            const int k = wv.first + j, nri = sloc[4 * k + 2], r0 = sloc[4 * k + 3];
            for (int row = 1; row < nnt; ++row)  // row 0 is never recorded
                for (int ir = 0; ir < nri; ++ir)
                    out[(size_t)row * nr + r0 + ir] = h[W.sb[j].r_off + (size_t)row * nri + ir];
        }
    }
    return vti::kOk;
}

// Device bytes of grad2d_merged: the side-by-side model and its FP64
// gradient, and the largest wave (the max_wave largest shots).
long long grad2d_merged_bytes(const Grad2dPart *p, int n, int nz, int nt, int npad, float dt, float rate,
                              int max_wave)
{
    const int ratio = std::max(1, (int)(rate / dt + 0.5f));
    const long long nnt = (nt - 1) / ratio + 1, nnz = nz + 2 * npad;
    long long total = 0;
    std::vector<long long> shot;
    for (int r = 0; r < n; ++r) {
        const long long nxz = (long long)(p[r].nx + 2 * npad) * nnz;
        total += 4LL * nxz * 3 + 8LL * nxz * 3;
        for (int k = 0; k < p[r].ns; ++k)
            shot.push_back(4LL * nxz * 12 + 8LL * nxz * nnt + 8LL * nnt * p[r].sloc[4 * k + 2] + 64);
    }
    std::sort(shot.begin(), shot.end(), std::greater<long long>());
    for (size_t k = 0; k < shot.size() && (max_wave <= 0 || (int)k < max_wave); ++k) total += shot[k];
    return total + (1LL << 20);
}

// Several 2D objective + gradient requests of one schedule (nz, npad, nt,
// dx, dz, dt, rate, wavelet; the standard taper) in one batched run.  Their
// models sit side by side in one model of their total width; each request's
// shots become batch shots whose sub-grid is the request's whole model, so
// every shot's FP32 gradient is the one-shot driver's.  The shots' FP64 sums
// stay within their request's columns, in shot order, and each shot's
// residual runs with the one-shot driver's grid, so every request gets the
// objective and gradient vti2d_objfunc_gradient returns for it alone, bit for
// bit (FP32 outputs, as that driver's), with the snapshots stored in
// snap_fmt (0 = FP32 as the one-shot loop stores them; 1-3 the 16-bit forms).
// With overlap, two waves of half the size run at once (run_waves), in
// memory the estimate of grad2d_merged_bytes does not cover; the request
// server runs merged requests without it.
int grad2d_merged(vti_ctx *ctx, const Grad2dPart *p, int n, const float *wavelet, int nz, int nt, int npad,
                  float dx, float dz, float dt, float rate, int max_wave, int snap_fmt, bool overlap)
{
    VTI_CHECK(ctx && p && n >= 1 && wavelet, "grad2d_merged: bad arguments");
    const int nnz = nz + 2 * npad;
    std::vector<int> off(n);
    int NNX = 0;
    for (int r = 0; r < n; ++r) {
        off[r] = NNX;
        NNX += p[r].nx + 2 * npad;
    }
    const int ratio = (int)(rate / dt + 0.5f);
    VTI_CHECK(ratio >= 1, "grad2d_merged: rate/dt below 1");
    const int nnt = (nt - 1) / ratio + 1;
    std::vector<int> shots, rl, owner;
    std::vector<float> data;
    size_t ndata = 0;
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < p[r].ns; ++k) ndata += (size_t)nnt * std::max(0, p[r].sloc[4 * k + 2]);
    data.reserve(ndata);
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < p[r].ns; ++k) {
            const int ix = p[r].sloc[4 * k], iz = p[r].sloc[4 * k + 1], nri = p[r].sloc[4 * k + 2],
                      r0 = p[r].sloc[4 * k + 3];
            VTI_CHECK(nri >= 0 && r0 >= 0 && r0 + nri <= p[r].nr, "request %d shot %d: receiver range outside rloc",
                      r, k);
            shots.insert(shots.end(), {off[r], p[r].nx, ix, iz, nri, (int)(rl.size() / 2)});
            for (int ir = 0; ir < nri; ++ir) {
                rl.push_back(p[r].rloc[2 * (r0 + ir)]);
                rl.push_back(p[r].rloc[2 * (r0 + ir) + 1]);
            }
            for (int row = 0; row < nnt; ++row) {  // the shot's columns of each trace row
                const float *src = p[r].data + (size_t)row * p[r].nr + r0;
                data.insert(data.end(), src, src + nri);
            }
            owner.push_back(r);
        }
    const int ns = (int)owner.size();
    std::vector<double> obj(n, 0.0);
    if (ns) {
        BatchArgs a;
        int rc = parse_args(a, shots.data(), ns, rl.data(), (int)(rl.size() / 2), NNX - 2 * npad, nz, nt, npad, dx,
                            dz, dt, rate, 2, 16, snap_fmt);
        if (rc) return rc;
        VTI_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        BatchShared S;
        rc = setup_shared(S, a, st);
        if (rc) return rc;
        for (int r = 0; r < n; ++r) {
            const int w = p[r].nx + 2 * npad;
            for (int k = 0; k < 3; ++k)
                if (n == 1)  // one model: a plain copy (a strided one from pageable memory goes row by row)
                    VTI_HIP(hipMemcpyAsync(S.C[k].p, p[r].c[k], (size_t)w * nnz * 4, hipMemcpyDefault, st));
                else
                    VTI_HIP(hipMemcpy2DAsync(S.C[k].p + off[r], (size_t)NNX * 4, p[r].c[k], (size_t)w * 4,
                                             (size_t)w * 4, nnz, hipMemcpyDefault, st));
        }
        // at most max_wave shots, fewer when the cache model asks; with
        // overlap, two waves of half as many in flight (run_waves)
        const int cw = cache_wave(shots.data(), ns, npad, nz);
        const int wave = max_wave > 0 ? std::min(max_wave, cw) : cw;
        const size_t mfree = vti::device_free_bytes();
        const long long budget = ((long long)mfree - (long long)(2ULL << 30)) / (overlap ? 3 : 1);
        if (overlap)
            rc = run_waves(S, a, split_waves(a, (wave + 1) / 2, budget), data.data(), rl.data(), wavelet, st,
                           owner.data(), obj.data());
        else
            for (const auto &wv : split_waves(a, wave, budget)) {
                BatchWave W;
                rc = setup_wave(W, a, wv.first, wv.second, data.data(), rl.data(), st);
                if (!rc) rc = enqueue_wave(W, S, a, wavelet, st, true);
                if (!rc) rc = wave_objective(W, st, owner.data(), obj.data());
                if (rc) break;
            }
        if (rc) return rc;
        // rounded to FP32 on the device
        const size_t g = (size_t)NNX * nnz;
        DBuf<float> G32;
        rc = G32.alloc(3 * g, st);
        if (rc) return rc;
        for (int k = 0; k < 3; ++k)
            hipLaunchKernelGGL(narrow_b_kernel, dim3((unsigned)std::min<size_t>((g + 255) / 256, 4096)), dim3(256), 0,
                               st, G32.p + k * g, S.A[k].p, (long long)g);
        // one contiguous transfer (a strided copy into pageable memory goes
        // row by row), then the columns split on the host
        std::vector<float> h(3 * g);
        VTI_HIP(hipMemcpyAsync(h.data(), G32.p, 3 * g * 4, hipMemcpyDeviceToHost, st));
        VTI_HIP(hipStreamSynchronize(st));
        for (int r = 0; r < n; ++r) {
            const size_t w = (size_t)(p[r].nx + 2 * npad);
            for (int k = 0; k < 3; ++k)
                for (int z = 0; z < nnz; ++z)
                    std::memcpy(p[r].g[k] + (size_t)z * w, h.data() + k * g + (size_t)z * NNX + off[r], w * 4);
        }
    }
    for (int r = 0; r < n; ++r) {
        if (!ns)
            for (int k = 0; k < 3; ++k)
                std::fill(p[r].g[k], p[r].g[k] + (size_t)(p[r].nx + 2 * npad) * nnz, 0.f);
        // objFuncGradient_f returns 0.5f*objfunc as float (objFuncGradient_f.cu:370)
        if (p[r].objfunc) *p[r].objfunc = (float)(0.5f * obj[r]);
    }
    return vti::kOk;
}

}  // namespace vti

extern "C" {

// Launch-shape autotuner for vti2d_objfunc_gradient_batch (see vti_amd.h).
// Times short runs of the same shots (nt cut to 24 snapshot intervals) with a
// synthetic model and zero data over: levels per launch 1/2, tile rows 8/16,
// with waves sized by cache_wave.  The pick is cached per geometry for the
// life of the process.
int vti2d_batch_tune(vti_ctx *ctx, const int *shots, int ns, const int *rloc, int nr, int NX, int nz, int nt,
                     int npad, float dx, float dz, float dt, float rate, int *levels, int *tile_rows, int *max_wave,
                     double *best_ms)
{
    VTI_CHECK(ctx && shots && rloc && levels && tile_rows && max_wave, "vti2d_batch_tune: null argument");
    VTI_CHECK(ns >= 1 && NX > 0 && nz > 0 && npad >= 0 && nt >= 3 && dt > 0 && rate > 0,
              "vti2d_batch_tune: bad sizes");
    const int ratio = (int)(rate / dt + 0.5f);
    VTI_CHECK(ratio >= 2 && (nt - 1) % ratio == 0, "vti2d_batch_tune: schedule needs rate/dt >= 2 and "
                                                   "(nt-1) %% (rate/dt) == 0");
    const int nt_t = std::min(nt, 24 * ratio + 1);
    using Key = std::tuple<int, int, int, int, int, int, long long>;
    static std::mutex mu;
    static std::map<Key, std::tuple<int, int, int, double>> cache;
    long long h = 1469598103934665603LL;
    for (int k = 0; k < 6 * ns; ++k) h = (h ^ shots[k]) * 1099511628211LL;
    for (int k = 0; k < 2 * nr; ++k) h = (h ^ rloc[k]) * 1099511628211LL;
    const Key key{ctx->device, ns, NX, nz, npad, ratio, h};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            std::tie(*levels, *tile_rows, *max_wave, std::ignore) = it->second;
            if (best_ms) *best_ms = std::get<3>(it->second);
            return vti::kOk;
        }
    }
    const int NNX = NX + 2 * npad, nnz = nz + 2 * npad;
    const size_t g = (size_t)NNX * nnz;
    std::vector<float> c33(g, 2000.f * 2000.f), c11(g), c13(g);
    for (size_t i = 0; i < g; ++i) {
        c11[i] = 1.2f * c33[i];
        c13[i] = 1.05f * c33[i];
    }
    std::vector<float> w(nt_t, 0.f);
    if (nt_t > 1) w[1] = 1.f;
    const int nnt = (nt_t - 1) / ratio + 1;
    long long ndata = 0;
    for (int s = 0; s < ns; ++s) ndata += (long long)nnt * shots[6 * s + 4];
    std::vector<float> data((size_t)std::max<long long>(ndata, 1), 0.f);
    std::vector<double> gc(3 * g);
    double obj = 0;
    const int wave = cache_wave(shots, ns, npad, nz);
    int bl = 2, br = 16, bw = wave >= ns ? 0 : wave;
    double best = 1e300;
    for (int lv = 1; lv <= 2; ++lv)
        for (int rows : {8, 16})
            for (int wv : {bw}) {
                if (lv == 1 && rows != 8) continue;  // the tile height only shapes two-level launches
                VTI_HIP(hipStreamSynchronize(ctx->stream));
                const auto t0 = std::chrono::steady_clock::now();
                const int rc = vti2d_objfunc_gradient_batch(ctx, &obj, gc.data(), gc.data() + g, gc.data() + 2 * g,
                                                            data.data(), c11.data(), c13.data(), c33.data(), w.data(),
                                                            shots, ns, rloc, nr, NX, nz, nt_t, npad, dx, dz, dt, rate,
                                                            wv, lv, rows, 0);
                if (rc) return rc;
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (ms < best) {
                    best = ms;
                    bl = lv;
                    br = rows;
                    bw = wv;
                }
            }
    *levels = bl;
    *tile_rows = br;
    *max_wave = bw;
    if (best_ms) *best_ms = best;
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = std::make_tuple(bl, br, bw, best);
    return vti::kOk;
}

}  // extern "C"
