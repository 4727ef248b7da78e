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
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "vti2d_batch_kernels.h"
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

struct ShotIn {
    int c0, nx, ix, iz, nr, r0;
};

}  // namespace

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
    VTI_CHECK(ns >= 1 && nr >= 0 && NX > 0 && nz > 0 && npad >= 0 && nt >= 3 && dx > 0 && dz > 0 && dt > 0,
              "vti2d_objfunc_gradient_batch: bad sizes");
    if (levels == 0) levels = 2;
    if (tile_rows == 0) tile_rows = 16;
    VTI_CHECK(levels == 1 || levels == 2, "levels per launch must be 1 or 2, got %d", levels);
    VTI_CHECK(tile_rows == 4 || tile_rows == 8 || tile_rows == 16, "tile rows must be 4, 8 or 16, got %d", tile_rows);
    VTI_CHECK(snap_fmt >= 0 && snap_fmt <= 2, "snapshot format must be 0 (FP32), 1 (FP16) or 2 (BF16), got %d",
              snap_fmt);
    VTI_HIP(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipStream_t st = ctx->stream;
    const int NNX = NX + 2 * npad, nnz = nz + 2 * npad;
    VTI_CHECK((long long)NNX * nnz < (1LL << 30),
              "vti2d_objfunc_gradient_batch: padded model of %d x %d points exceeds the 2^30 the kernels index", nnz,
              NNX);
    const int ratio = (int)(rate / dt + 0.5f);  // objFuncGradient_f.cu:30 rounds in float
    VTI_CHECK(ratio >= 2 && (nt - 1) % ratio == 0,
              "the reference's snapshot rotation lines up only for rate/dt >= 2 and (nt-1) %% (rate/dt) == 0 "
              "(got ratio %d, nt %d)", ratio, nt);
    const int nnt = (nt - 1) / ratio + 1, nnt_data = nnt;
    const float dx2 = dx * dx, dz2 = dz * dz, dt2 = dt * dt;
    const bool half = snap_fmt != 0;

    // ---- shots: check and lay out
    std::vector<ShotIn> in(ns);
    long long data_total = 0;
    std::vector<long long> data_off(ns);
    for (int s = 0; s < ns; ++s) {
        ShotIn &a = in[s];
        a.c0 = shots[6 * s];
        a.nx = shots[6 * s + 1];
        a.ix = shots[6 * s + 2];
        a.iz = shots[6 * s + 3];
        a.nr = shots[6 * s + 4];
        a.r0 = shots[6 * s + 5];
        const int w = a.nx + 2 * npad;
        VTI_CHECK(a.nx > 0 && a.c0 >= 0 && a.c0 + w <= NNX, "shot %d: sub-grid [%d, %d) outside the model's %d columns",
                  s, a.c0, a.c0 + w, NNX);
        VTI_CHECK(w >= 9 && nnz >= 9, "shot %d: sub-grid too small", s);
        VTI_CHECK(a.ix >= 4 && a.ix < w - 4 && a.iz >= 4 && a.iz < nnz - 4, "shot %d: source outside the interior",
                  s);
        VTI_CHECK(a.nr >= 0 && a.r0 >= 0 && a.r0 + a.nr <= nr, "shot %d: receiver range outside rloc", s);
        std::vector<char> seen((size_t)w * nnz, 0);
        for (int k = 0; k < a.nr; ++k) {
            const int x = rloc[2 * (a.r0 + k)], z = rloc[2 * (a.r0 + k) + 1];
            VTI_CHECK(x >= 4 && x < w - 4 && z >= 4 && z < nnz - 4, "shot %d receiver %d outside the interior", s, k);
            VTI_CHECK(!seen[x + (size_t)z * w], "shot %d: two receivers share a grid point", s);
            seen[x + (size_t)z * w] = 1;
        }
        data_off[s] = data_total;
        data_total += (long long)nnt_data * a.nr;
    }

    // ---- global model and FP64 gradient
    const size_t gsize = (size_t)NNX * nnz;
    DBuf<float> C[3], t0d;
    DBuf<double> A[3];
    DBuf<double> part;
    int rc = vti::kOk;
    const float *csrc[3] = {c11, c13, c33};
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = C[k].alloc(gsize, st);
        if (!rc) VTI_HIP(hipMemcpyAsync(C[k].p, csrc[k], gsize * 4, hipMemcpyHostToDevice, st));
        if (!rc) rc = A[k].alloc(gsize, st);
    }
    std::vector<float> t0(npad + 1);  // t0[npad] = 1: the factor away from the edges
    vti::taper_profile2d(t0.data(), npad);
    t0[npad] = 1.f;
    if (!rc) rc = t0d.alloc(t0.size(), st);
    if (!rc) VTI_HIP(hipMemcpyAsync(t0d.p, t0.data(), t0.size() * 4, hipMemcpyHostToDevice, st));
    const int nblk_r = 1024;
    if (!rc) rc = part.alloc(nblk_r, st);
    if (rc) return rc;

    // ---- waves within the memory budget: per shot 8 wavefield planes, 3
    // gradient planes, the receiver map, the snapshots (and for 16-bit ones
    // two staging planes and the slice scales/maxima) and two trace copies
    auto shot_bytes = [&](int s) {
        const long long nxz = (long long)(in[s].nx + 2 * npad) * nnz;
        const long long snapb = half ? 2LL * nxz * 2 * nnt + 4LL * nxz * 2 + 16LL * nnt : 4LL * nxz * 2 * nnt;
        return 4LL * nxz * (8 + 3 + 1) + snapb + 8LL * nnt_data * in[s].nr + 64;
    };
    size_t mfree = 0, mtotal = 0;
    VTI_HIP(hipMemGetInfo(&mfree, &mtotal));
    const long long budget = (long long)mfree - (long long)(2ULL << 30);
    VTI_CHECK(budget > shot_bytes(0), "not enough device memory for one shot");
    double obj = 0.0;
    int s0 = 0;
    while (s0 < ns) {
        long long used = 0;
        int s1 = s0;
        while (s1 < ns && (max_wave <= 0 || s1 - s0 < max_wave) && used + shot_bytes(s1) <= budget)
            used += shot_bytes(s1++);
        if (s1 == s0) s1 = s0 + 1;
        const int nw = s1 - s0;

        // layout of the wave: sb/tile_shot for the one-level kernels (64 x 8
        // tiles), sbt/tile_shot_t for the two-level ones (64 x tile_rows)
        std::vector<ShotB> sb(nw), sbt(nw);
        std::vector<int> tile_shot, tile_shot_t;
        const int BYT = tile_rows, ntz = (nnz - 8 + kBY2 - 1) / kBY2, ntz_t = (nnz - 8 + BYT - 1) / BYT;
        long long f_total = 0, snap_total = 0, tr_total = 0, max_plane = 0;
        int nrec = 0;
        for (int j = 0; j < nw; ++j) {
            const ShotIn &a = in[s0 + j];
            ShotB &b = sb[j];
            b.nnx = a.nx + 2 * npad;
            b.ntx = (b.nnx - 8 + 63) / 64;
            b.tile0 = (int)tile_shot.size();
            for (int t = 0; t < b.ntx * ntz; ++t) tile_shot.push_back(j);
            b.nr = a.nr;
            b.f_off = f_total;
            b.c_off = a.c0;
            b.snap_off = snap_total;
            b.r_off = tr_total;
            b.stage_off = 2 * f_total;
            b.src = a.ix + a.iz * b.nnx;
            b.rz0 = nnz;
            b.rz1 = -1;
            for (int r = 0; r < a.nr; ++r) {
                const int z = rloc[2 * (a.r0 + r) + 1];
                b.rz0 = std::min(b.rz0, z);
                b.rz1 = std::max(b.rz1, z);
            }
            b.scale_off = 2 * nnt * j;
            sbt[j] = b;
            sbt[j].tile0 = (int)tile_shot_t.size();
            for (int t = 0; t < b.ntx * ntz_t; ++t) tile_shot_t.push_back(j);
            const long long plane = (long long)b.nnx * nnz;
            max_plane = std::max(max_plane, plane);
            f_total += plane;
            snap_total += (long long)nnt * 2 * plane;
            tr_total += (long long)nnt_data * a.nr;
            nrec += a.nr;
        }
        std::vector<long long> fidx(nrec), tb(nrec);
        std::vector<int> tsd(nrec);
        std::vector<int> hmap((size_t)f_total, -1);
        {
            int k = 0;
            for (int j = 0; j < nw; ++j) {
                const ShotIn &a = in[s0 + j];
                for (int r = 0; r < a.nr; ++r, ++k) {
                    const int x = rloc[2 * (a.r0 + r)], z = rloc[2 * (a.r0 + r) + 1];
                    fidx[k] = sb[j].f_off + x + (long long)z * sb[j].nnx;
                    tb[k] = sb[j].r_off + r;
                    tsd[k] = a.nr;
                    hmap[(size_t)fidx[k]] = r;
                }
            }
        }
        std::vector<float> obs_h((size_t)std::max<long long>(tr_total, 1));
        for (int j = 0; j < nw; ++j)
            if (in[s0 + j].nr)
                std::memcpy(obs_h.data() + sb[j].r_off, data + data_off[s0 + j],
                            sizeof(float) * (size_t)nnt_data * in[s0 + j].nr);

        DBuf<ShotB> sbd, sbtd;
        DBuf<int> tsh, tsht, recmap, tsdd;
        DBuf<long long> fidxd, tbd;
        DBuf<float> P[4], Q[4], G[3], snapf, stage, scale, tr, obs;
        DBuf<unsigned short> snaph;
        DBuf<unsigned> amax;
        rc = sbd.alloc(nw, st);
        if (!rc) rc = sbtd.alloc(nw, st);
        if (!rc) rc = tsh.alloc(tile_shot.size(), st);
        if (!rc) rc = tsht.alloc(tile_shot_t.size(), st);
        if (!rc) rc = recmap.alloc((size_t)f_total, st);
        if (!rc) rc = tsdd.alloc(nrec, st);
        if (!rc) rc = fidxd.alloc(nrec, st);
        if (!rc) rc = tbd.alloc(nrec, st);
        for (int k = 0; k < 4 && !rc; ++k) {
            rc = P[k].alloc((size_t)f_total, st);
            if (!rc) rc = Q[k].alloc((size_t)f_total, st);
        }
        for (int k = 0; k < 3 && !rc; ++k) rc = G[k].alloc((size_t)f_total, st);
        if (half) {
            if (!rc) rc = snaph.alloc((size_t)snap_total, st);
            if (!rc) rc = stage.alloc((size_t)(2 * f_total), st);
            if (!rc) rc = scale.alloc((size_t)(2 * nnt * nw), st);
            if (!rc) rc = amax.alloc((size_t)(2 * nnt * nw), st);
        } else {
            if (!rc) rc = snapf.alloc((size_t)snap_total, st);
        }
        if (!rc) rc = tr.alloc((size_t)std::max<long long>(tr_total, 1), st);
        if (!rc) rc = obs.alloc((size_t)std::max<long long>(tr_total, 1), st);
        if (rc) return rc;
        VTI_HIP(hipMemcpyAsync(sbd.p, sb.data(), sizeof(ShotB) * nw, hipMemcpyHostToDevice, st));
        VTI_HIP(hipMemcpyAsync(sbtd.p, sbt.data(), sizeof(ShotB) * nw, hipMemcpyHostToDevice, st));
        VTI_HIP(hipMemcpyAsync(tsh.p, tile_shot.data(), sizeof(int) * tile_shot.size(), hipMemcpyHostToDevice, st));
        VTI_HIP(hipMemcpyAsync(tsht.p, tile_shot_t.data(), sizeof(int) * tile_shot_t.size(), hipMemcpyHostToDevice,
                               st));
        VTI_HIP(hipMemcpyAsync(recmap.p, hmap.data(), sizeof(int) * hmap.size(), hipMemcpyHostToDevice, st));
        if (nrec) {
            VTI_HIP(hipMemcpyAsync(tsdd.p, tsd.data(), sizeof(int) * nrec, hipMemcpyHostToDevice, st));
            VTI_HIP(hipMemcpyAsync(fidxd.p, fidx.data(), sizeof(long long) * nrec, hipMemcpyHostToDevice, st));
            VTI_HIP(hipMemcpyAsync(tbd.p, tb.data(), sizeof(long long) * nrec, hipMemcpyHostToDevice, st));
        }
        if (tr_total)
            VTI_HIP(hipMemcpyAsync(obs.p, obs_h.data(), sizeof(float) * tr_total, hipMemcpyHostToDevice, st));

        BatchGeo bg;
        bg.nnz = nnz;
        bg.npad = npad;
        bg.cpitch = NNX;
        bg.nnxz_total = f_total;
        const dim3 grid((unsigned)tile_shot.size()), block(64, kBY2);
        const dim3 gridt((unsigned)tile_shot_t.size()), blockt(64, BYT);
        const int rb = (nrec + 255) / 256;
        const vti::SnapW sw{half ? stage.p : snapf.p, half ? amax.p : nullptr};
        const vti::SnapR sr{snapf.p, snaph.p, scale.p};
        const dim3 gpack((unsigned)std::min<long long>((2 * max_plane + 255) / 256, 1024), (unsigned)nw);
        auto pack = [&](int slot) {
            if (half && slot >= 0)
                hipLaunchKernelGGL(vti::pack_snap_kernel, gpack, dim3(256), 0, st, snaph.p, scale.p, stage.p, amax.p,
                                   sbd.p, nnz, slot, snap_fmt == 2 ? 1 : 0);
        };
        auto record = [&](int it, int buf) {
            if (it % ratio == 0 && nrec)
                hipLaunchKernelGGL(record_b_kernel, dim3(rb), dim3(256), 0, st, tr.p, P[buf].p, Q[buf].p, fidxd.p,
                                   tbd.p, tsdd.p, nrec, it / ratio);
        };
        auto snap_slot = [&](int lc) { return lc % ratio == 0 ? lc / ratio : -1; };

        // ---- forward (objFuncGradient_f.cu:160-210 for every shot).  Buffers
        // ip/ic hold the previous and current levels; s0b/s1b are free.
        int ip = 1, ic = 0, s0b = 2, s1b = 3;
        hipLaunchKernelGGL(set_src_b_kernel, dim3((nw + 63) / 64), dim3(64), 0, st, P[ic].p, Q[ic].p, sbd.p, nw,
                           dt2 * wavelet[0]);
        for (int it = 2; it < nt;) {
            if (levels == 2 && it + 1 < nt) {
                // levels it and it+1, reading levels it-1 and it
                const int sl1 = snap_slot(it - 1), sl2 = snap_slot(it);
#define FWD_T2(BYV)                                                                                                 \
    hipLaunchKernelGGL(vti::fwd2db_t2_kernel<BYV>, gridt, blockt, 0, st, P[s0b].p, Q[s0b].p, P[s1b].p, Q[s1b].p,     \
                       P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p, C[1].p, C[2].p, t0d.p, sbtd.p, tsht.p, bg, dx2, \
                       dz2, dt2, dt2 * wavelet[it - 1], dt2 * wavelet[it], sw, sl1, sl2)
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
                hipLaunchKernelGGL(vti::fwd2db_kernel, grid, block, 0, st, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p,
                                   C[0].p, C[1].p, C[2].p, t0d.p, sbd.p, tsh.p, bg, dx2, dz2, dt2,
                                   dt2 * wavelet[it - 1], sw, sl);
                pack(sl);
                std::swap(ip, ic);
                record(it, ic);
                it += 1;
            }
        }
        hipLaunchKernelGGL(vti::dsnap2db_kernel, grid, block, 0, st, P[ic].p, Q[ic].p, sbd.p, tsh.p, bg, dx2, dz2, sw,
                           nnt - 1);
        pack(nnt - 1);
        VTI_HIP(hipGetLastError());

        // ---- residual and objective (objFuncGradient_f.cu:214-219)
        if (tr_total) {
            hipLaunchKernelGGL(residual_b_kernel, dim3(nblk_r), dim3(256), 0, st, tr.p, obs.p, tr_total, part.p);
            std::vector<double> hp(nblk_r);
            VTI_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * nblk_r, hipMemcpyDeviceToHost, st));
            VTI_HIP(hipStreamSynchronize(st));
            for (double v : hp) obj += v;
        }

        // ---- backward (objFuncGradient_f.cu:225-300 for every shot)
        for (int k = 0; k < 4; ++k) {
            VTI_HIP(hipMemsetAsync(P[k].p, 0, (size_t)f_total * 4, st));
            VTI_HIP(hipMemsetAsync(Q[k].p, 0, (size_t)f_total * 4, st));
        }
        ic = 0;
        ip = 1;
        s0b = 2;
        s1b = 3;
        if (nrec)
            hipLaunchKernelGGL(inject_first_b_kernel, dim3(rb), dim3(256), 0, st, P[ic].p, Q[ic].p, tr.p, fidxd.p,
                               tbd.p, tsdd.p, nrec, nnt_data - 1, dt2);
        {
            float f = (float)((nt - 2.) / ratio);
            const int i1 = (int)f;
            f = f - i1;
#define XCOR(SFV)                                                                                                  \
    hipLaunchKernelGGL(vti::xcor2db_kernel<SFV>, grid, block, 0, st, P[ic].p, Q[ic].p, sbd.p, tsh.p, bg, sr, i1, f, \
                       G[0].p, G[1].p, G[2].p)
            if (snap_fmt == 0) XCOR(0);
            else if (snap_fmt == 1) XCOR(1);
            else XCOR(2);
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
        for (int it = nt - 3; it >= 0;) {
            if (levels == 2 && it >= 1) {
                const vti::LevelB l1 = level(it), l2 = level(it - 1);
#define BWD_T2(BYV, SFV)                                                                                             \
    vti::bwd2db_t2_kernel<BYV, SFV><<<gridt, blockt, 0, st>>>(P[s0b].p, Q[s0b].p, P[s1b].p,                         \
                       Q[s1b].p, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p, C[1].p, C[2].p, t0d.p, sbtd.p, tsht.p, \
                       bg, recmap.p, tr.p, l1, l2, sr, G[0].p, G[1].p, G[2].p, dx2, dz2, dt2)
#define BWD_T2_SF(BYV)                  \
    if (snap_fmt == 0) BWD_T2(BYV, 0);  \
    else if (snap_fmt == 1) BWD_T2(BYV, 1); \
    else BWD_T2(BYV, 2)
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
#define BWD1(SFV)                                                                                                    \
    hipLaunchKernelGGL(vti::bwd2db_kernel<SFV>, grid, block, 0, st, P[ip].p, Q[ip].p, P[ic].p, Q[ic].p, C[0].p,      \
                       C[1].p, C[2].p, t0d.p, sbd.p, tsh.p, bg, recmap.p, tr.p, l.row, l.fr, sr, l.slot, l.f, G[0].p, \
                       G[1].p, G[2].p, dx2, dz2, dt2)
                if (snap_fmt == 0) BWD1(0);
                else if (snap_fmt == 1) BWD1(1);
                else BWD1(2);
#undef BWD1
                std::swap(ip, ic);
                it -= 1;
            }
        }
        VTI_HIP(hipGetLastError());
        // ---- per-shot FP32 gradients into the FP64 global sum, in shot order
        for (int j = 0; j < nw; ++j)
            hipLaunchKernelGGL(accum_crop_kernel, dim3((sb[j].nnx + 255) / 256, nnz), dim3(256), 0, st, A[0].p,
                               A[1].p, A[2].p, G[0].p, G[1].p, G[2].p, sb[j].f_off, sb[j].c_off, sb[j].nnx, nnz, NNX);
        VTI_HIP(hipGetLastError());
        VTI_HIP(hipStreamSynchronize(st));
        s0 = s1;
    }
    double *out[3] = {gc11, gc13, gc33};
    for (int k = 0; k < 3; ++k) VTI_HIP(hipMemcpyAsync(out[k], A[k].p, gsize * 8, hipMemcpyDeviceToHost, st));
    VTI_HIP(hipStreamSynchronize(st));
    *objfunc = 0.5 * obj;
    return vti::kOk;
}

// Launch-shape autotuner for vti2d_objfunc_gradient_batch (see vti_amd.h).
// Times short runs of the same shots (nt cut to 24 snapshot intervals) with a
// synthetic model and zero data over: levels per launch 1/2, tile rows 8/16,
// waves of all shots that fit / at most 8.  The pick is cached per geometry
// for the life of the process.
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
    int bl = 2, br = 16, bw = 0;
    double best = 1e300;
    const int waves[2] = {0, 8};
    for (int lv = 1; lv <= 2; ++lv)
        for (int rows : {8, 16})
            for (int wv : waves) {
                if (lv == 1 && rows != 8) continue;  // the tile height only shapes two-level launches
                if (wv && wv >= ns) continue;
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
