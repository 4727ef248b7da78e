// 2D VTI drivers on one MI355X with the reference's signatures and results:
// synthetic_f, objFuncGradient_f and GNhessian_f of lib/vti-fwi-2d.
//
// Where the reference streams D snapshots between device and pinned host
// buffers through a three-buffer rotation (objFuncGradient_f.cu:104-300),
// all snapshots of a shot stay in HBM (cfg2: 751 x 2 x 0.65 M floats =
// 3.9 GB) and the backward pass reads the two that bracket each level.  For
// ratio >= 2 and (nt-1) % ratio == 0 -- the conditions the reference's
// rotation needs to line its buffers up -- this is the same schedule: level
// it is imaged against (1-f) D(ratio*k) + f D(ratio*(k+1)), k = it/ratio.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "vti2d_kernels.h"
#include "vti_dbuf.h"
#include "vti_ctx.h"
#include "vti_internal.h"

using vti::DBuf;
using vti::Geo2;
using vti::kBY2;

namespace {

__global__ void set2_kernel(float *a, float *b, long long i, float v)
{
    a[i] = v;
    b[i] = v;
}

// recordData (kernels.cu:317-324) into one time-major row
__global__ void record2d_kernel(float *__restrict__ row, const float *__restrict__ p, const float *__restrict__ q,
                                const long long *__restrict__ idx, int nr)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < nr) row[ir] = 0.5f * (p[idx[ir]] + q[idx[ir]]);
}

// injectData (kernels.cu:208-217) at arbitrary receiver points (first adjoint level)
__global__ void inject2d_kernel(float *__restrict__ p, float *__restrict__ q, const long long *__restrict__ idx,
                                const float *__restrict__ r0, const float *__restrict__ r1, float fr, int nr,
                                float dt2)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < nr) {
        const float s = 0.5f * ((1.f - fr) * r0[ir] + fr * r1[ir]);
        p[idx[ir]] += dt2 * s;
        q[idx[ir]] += dt2 * s;
    }
}

// residual rows r = modeled - observed and partial sums of r*r
// (objFuncGradient_f.cu:214-219); rows time-major [it][ir]
__global__ void residual2d_kernel(float *__restrict__ bg, const float *__restrict__ obs, long long n,
                                  double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const float r = bg[k] - obs[k];
        bg[k] = r;
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

__global__ void acc64_2d_kernel(double *__restrict__ acc, const float *__restrict__ v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc[i] += (double)v[i];
}

__global__ void narrow2d_kernel(float *__restrict__ out, const double *__restrict__ acc, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)acc[i];
}

// ---- rtm_f / odcig_f (lib/vti-fwi-2d/rtm_f.cu, odcig_f.cu) -----------------

// level 1 of a dipole source: injectDipoleSource into zero fields, then abc
// (rtm_f.cu:121-123)
__global__ void dipole_level1_kernel(float *p, float *q, const float *__restrict__ taper, long long s, int nnx, float v)
{
    p[s] = v * taper[s];
    q[s] = v * taper[s];
    p[s + nnx] = -v * taper[s + nnx];
    q[s + nnx] = -v * taper[s + nnx];
}

// recordWavefieldSlice (kernels.cu:335-343)
__global__ void slice2d_kernel(float *__restrict__ d, const float *__restrict__ p, const float *__restrict__ q, Geo2 g)
{
    const int ix = 4 + blockIdx.x * blockDim.x + threadIdx.x, iz = 4 + blockIdx.y;
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    d[i] = 0.5f * (p[i] + q[i]);
}

// first adjoint level: injectDipoleData of the last data row (f = 0) into
// zero fields, then abc (rtm_f.cu:160-164)
__global__ void dipole_first_kernel(float *__restrict__ p, float *__restrict__ q, const float *__restrict__ taper,
                                    const long long *__restrict__ idx, const float *__restrict__ row, int nr, int nnx,
                                    float dt2)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= nr) return;
    const long long i = idx[ir], j = i + nnx;
    const float s = 0.5f * ((1.f - 0.f) * row[ir] + 0.f * 0.f);
    const float v = dt2 * s;
    p[i] = v * taper[i];
    q[i] = v * taper[i];
    p[j] = -v * taper[j];
    q[j] = -v * taper[j];
}

// imagingCrossCor (kernels.cu:414-423) on its own, for the first adjoint level
__global__ void image2d_kernel(float *__restrict__ image, const float *__restrict__ ap, const float *__restrict__ aq,
                               const float *__restrict__ s0, const float *__restrict__ s1, float f, Geo2 g)
{
    const int ix = 4 + blockIdx.x * blockDim.x + threadIdx.x, iz = 4 + blockIdx.y;
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    const float d = (1.f - f) * s0[i] + f * s1[i];
    image[i] += 0.5f * (ap[i] + aq[i]) * d;
}

// extendedImagingCrossCor (kernels.cu:425-436) over the points whose
// contributions can be nonzero: both shifted points inside the interior
// (elsewhere the reference adds products with a zero factor)
__global__ void ximage2d_kernel(float *__restrict__ image, const float *__restrict__ ap, const float *__restrict__ aq,
                                const float *__restrict__ s0, const float *__restrict__ s1, float f, Geo2 g, int nhx,
                                int xlo)
{
    const int ix = xlo + blockIdx.x * blockDim.x + threadIdx.x, iz = 4 + blockIdx.y;
    if (ix >= g.nnx - xlo || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx, n = (long long)g.nnx * g.nnz;
    for (int h = -nhx; h <= nhx; ++h) {
        const float d = (1.f - f) * s0[i + h] + f * s1[i + h];
        image[i + (h + nhx) * n] += 0.5f * (ap[i - h] + aq[i - h]) * d;
    }
}

// One adjoint level of rtm_f/odcig_f: forwardCD (the reference propagates the
// receiver side with the forward operator, rtm_f.cu:172-173), injectDipoleData
// (recmap: +, dipmap: the point below a receiver, -), abc, and for RTM the
// imaging condition fused (image += 0.5 (p+q) * interpolated slice).
template <bool IMAGE>
__global__ __launch_bounds__(64 * kBY2) void rtm_bwd2d_kernel(
    float *__restrict__ p_n, float *__restrict__ q_n, const float *__restrict__ p_c, const float *__restrict__ q_c,
    const float *__restrict__ c11, const float *__restrict__ c13, const float *__restrict__ c33,
    const float *__restrict__ taper, const int *__restrict__ recmap, const int *__restrict__ dipmap, int inject,
    const float *__restrict__ r0, const float *__restrict__ r1, float fr, const float *__restrict__ s0,
    const float *__restrict__ s1, float f, float *__restrict__ image, Geo2 g, float dx2, float dz2, float dt2)
{
    __shared__ vti::Tiles2 t;
    int x0, z0;
    vti::tile_origin(x0, z0);
    const int ix = x0 + threadIdx.x, iz = z0 + threadIdx.y;
    vti::load_tiles<0>(t, p_c, q_c, nullptr, nullptr, nullptr, g, x0, z0);
    __syncthreads();
    if (ix >= g.nnx - 4 || iz >= g.nnz - 4) return;
    const long long i = ix + (long long)iz * g.nnx;
    const float tap = taper[i];
    const float tp = vti::dxx(t, dx2, g.rx2), tq = vti::dzz(t, dz2, g.rz2);
    float pn = 2.f * p_c[i] + dt2 * (c11[i] * tp + c13[i] * tq) - p_n[i] * tap;
    float qn = 2.f * q_c[i] + dt2 * (c13[i] * tp + c33[i] * tq) - q_n[i] * tap;
    if (inject) {
        const int ir = recmap[i];
        if (ir >= 0) {
            const float s = 0.5f * ((1.f - fr) * r0[ir] + fr * r1[ir]);
            pn += dt2 * s;
            qn += dt2 * s;
        }
        const int jr = dipmap[i];
        if (jr >= 0) {
            const float s = 0.5f * ((1.f - fr) * r0[jr] + fr * r1[jr]);
            pn -= dt2 * s;
            qn -= dt2 * s;
        }
    }
    pn = pn * tap;
    qn = qn * tap;
    p_n[i] = pn;
    q_n[i] = qn;
    if (IMAGE) {
        const float d = (1.f - f) * s0[i] + f * s1[i];
        image[i] += 0.5f * (pn + qn) * d;
    }
}

struct Call2 {
    vti_ctx *ctx;
    Geo2 g;
    int nx, nz, npad, nt, ratio, ntNeg, nnt, nnt_data;
    long long nnxz;
    float dt, ot, dx2, dz2, dt2;
    dim3 grid, block;
};

int setup(Call2 &c, vti_ctx *ctx, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot,
          bool fwd_ratio_f)
{
    VTI_CHECK(ctx, "null context");
    VTI_CHECK(nx > 0 && nz > 0 && npad >= 4 && nt >= 3, "bad grid/time sizes");
    VTI_CHECK(dx > 0 && dz > 0 && dt > 0 && rate > 0, "spacings and rates must be positive");
    c.ctx = ctx;
    c.nx = nx;
    c.nz = nz;
    c.npad = npad;
    c.nt = nt;
    c.g.nnx = nx + 2 * npad;
    c.g.nnz = nz + 2 * npad;
    c.nnxz = (long long)c.g.nnx * c.g.nnz;
    // synthetic_f rounds rate/dt+0.5 in double, objFuncGradient_f in float
    c.ratio = fwd_ratio_f ? (int)(rate / dt + 0.5f) : (int)(rate / dt + 0.5);
    c.ntNeg = (int)std::round(std::fabs(ot) / dt);
    VTI_CHECK(c.ratio >= 1, "sampling rate below the time step");
    c.nnt = (nt - 1) / c.ratio + 1;
    c.nnt_data = (nt - c.ntNeg - 1) / c.ratio + 1;
    c.dt = dt;
    c.ot = ot;
    c.dx2 = dx * dx;
    c.dz2 = dz * dz;
    c.dt2 = dt * dt;
    c.g.rx2 = vti::div_reciprocal(c.dx2);
    c.g.rz2 = vti::div_reciprocal(c.dz2);
    c.block = dim3(64, kBY2);
    c.grid = dim3((c.g.nnx - 8 + 63) / 64, (c.g.nnz - 8 + kBY2 - 1) / kBY2);
    VTI_HIP(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    return vti::kOk;
}

bool interior(const Call2 &c, int ix, int iz) { return ix >= 4 && ix < c.g.nnx - 4 && iz >= 4 && iz < c.g.nnz - 4; }

// copy to the device from host memory or, for a model kept resident by the
// caller, from device memory (the direction is taken from the pointers)
int upload(const Call2 &c, void *dst, const void *src, size_t bytes)
{
    VTI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c.ctx->stream));
    return vti::kOk;
}

int launch_blocks(long long n) { return (int)std::min<long long>(2048, std::max<long long>(1, (n + 255) / 256)); }

struct Shot2 {
    long long src;
    int ix, iz, nri, r0;
    std::vector<long long> idx;
};

int shot_of(const Call2 &c, const int *sloc, int is, const int *rloc, int nr, Shot2 &s)
{
    s.ix = sloc[4 * is];
    s.iz = sloc[4 * is + 1];
    s.nri = sloc[4 * is + 2];
    s.r0 = sloc[4 * is + 3];
    VTI_CHECK(s.nri >= 0 && s.r0 >= 0 && s.r0 + s.nri <= nr, "shot %d: receiver range outside rloc", is);
    VTI_CHECK(interior(c, s.ix, s.iz), "shot %d: source (%d,%d) outside the updated interior", is, s.ix, s.iz);
    s.src = s.ix + (long long)s.iz * c.g.nnx;
    s.idx.resize(s.nri);
    for (int ir = 0; ir < s.nri; ++ir) {
        const int x = rloc[2 * (s.r0 + ir)], z = rloc[2 * (s.r0 + ir) + 1];
        VTI_CHECK(x >= 0 && x < c.g.nnx && z >= 0 && z < c.g.nnz, "receiver %d outside the grid", s.r0 + ir);
        s.idx[ir] = x + (long long)z * c.g.nnx;
    }
    return vti::kOk;
}

struct Fields2 {
    DBuf<float> c11, c13, c33, taper, p[2], q[2];
};

int load_model(const Call2 &c, Fields2 &F, const float *c11, const float *c13, const float *c33, const float *taper)
{
    const size_t n = (size_t)c.nnxz;
    int rc = F.c11.alloc(n, c.ctx->stream);
    if (!rc) rc = F.c13.alloc(n, c.ctx->stream);
    if (!rc) rc = F.c33.alloc(n, c.ctx->stream);
    if (!rc) rc = F.taper.alloc(n, c.ctx->stream);
    for (int k = 0; k < 2 && !rc; ++k) {
        rc = F.p[k].alloc(n, c.ctx->stream);
        if (!rc) rc = F.q[k].alloc(n, c.ctx->stream);
    }
    if (rc) return rc;
    rc = upload(c, F.c11.p, c11, n * 4);
    if (!rc) rc = upload(c, F.c13.p, c13, n * 4);
    if (!rc) rc = upload(c, F.c33.p, c33, n * 4);
    if (!rc) rc = upload(c, F.taper.p, taper, n * 4);
    return rc;
}

int zero2(const Call2 &c, DBuf<float> *a, DBuf<float> *b)
{
    for (int k = 0; k < 2; ++k) {
        VTI_HIP(hipMemsetAsync(a[k].p, 0, a[k].n * 4, c.ctx->stream));
        VTI_HIP(hipMemsetAsync(b[k].p, 0, b[k].n * 4, c.ctx->stream));
    }
    return vti::kOk;
}

}  // namespace

extern "C" {

// synthetic_f, lib/vti-fwi-2d/synthetic_f.cu:13-152 (wave.h:4).  data: nr
// columns (all shots' receivers) x nnt_data rows, time-major; rows a shot
// never records keep their input values, as in the reference.
int vti2d_synthetic(vti_ctx *ctx, float *data, const float *c11, const float *c13, const float *c33,
                    const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                    int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    // Several shots with the standard taper (and rate/dt rounding the same in
    // float as in synthetic_f's double) run as one batch, bit for bit.
    if (ctx && ns >= 2 && data && c11 && c13 && c33 && wavelet && dx > 0 && dz > 0 && dt > 0 &&
        (int)(rate / dt + 0.5f) == (int)(rate / dt + 0.5) &&
        vti::grad2d_batchable(sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dt, rate, ot))
        return vti::synth2d_batched(ctx, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, nx, nz, nt, npad, dx, dz,
                                    dt, rate);
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, false);
    if (rc) return rc;
    VTI_CHECK(data && c11 && c13 && c33 && wavelet && sloc && rloc && taper && ns >= 0, "vti2d_synthetic: null argument");
    Fields2 F;
    rc = load_model(c, F, c11, c13, c33, taper);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    for (int is = 0; is < ns; ++is) {
        Shot2 s;
        rc = shot_of(c, sloc, is, rloc, nr, s);
        if (rc) return rc;
        DBuf<long long> idx;
        DBuf<float> tr;
        rc = idx.alloc(s.nri, st);
        if (!rc) rc = tr.alloc((size_t)s.nri * c.nnt_data, st);
        if (!rc && s.nri) rc = upload(c, idx.p, s.idx.data(), sizeof(long long) * s.nri);
        if (rc) return rc;
        rc = zero2(c, F.p, F.q);
        if (rc) return rc;
        // level 1: dt2*w[0] at the source, then abc (synthetic_f.cu:86-89)
        const float v0 = c.dt2 * wavelet[0];
        hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(1), 0, st, F.p[0].p, F.q[0].p, s.src, v0 * taper[s.src]);
        std::vector<char> rec_row(c.nnt_data, 0);
        auto record = [&](int row) {
            hipLaunchKernelGGL(record2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, tr.p + (size_t)row * s.nri,
                               F.p[0].p, F.q[0].p, idx.p, s.nri);
            rec_row[row] = 1;
        };
        if (c.ratio == 1 && ot == 0.f && s.nri && 1 < c.nnt_data) record(1);
        for (int it = 2; it < nt; ++it) {
            const float t = it * dt + ot;
            hipLaunchKernelGGL((vti::fwd2d_kernel<false>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                               F.q[0].p, nullptr, nullptr, nullptr, nullptr, F.c11.p, F.c13.p, F.c33.p, nullptr,
                               nullptr, nullptr, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, s.src, c.dt2 * wavelet[it - 1],
                               nullptr, nullptr, -1LL, nullptr, nullptr);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            if (t >= 0. && (it - c.ntNeg) % c.ratio == 0 && s.nri) {
                const int row = (it - c.ntNeg) / c.ratio;
                if (row < c.nnt_data) record(row);
            }
        }
        VTI_HIP(hipGetLastError());
        std::vector<float> h((size_t)s.nri * c.nnt_data);
        if (s.nri) VTI_HIP(hipMemcpyAsync(h.data(), tr.p, h.size() * 4, hipMemcpyDeviceToHost, st));
        VTI_HIP(hipStreamSynchronize(st));
        for (int row = 0; row < c.nnt_data; ++row)
            if (rec_row[row])
                for (int ir = 0; ir < s.nri; ++ir) data[(size_t)row * nr + s.r0 + ir] = h[(size_t)row * s.nri + ir];
    }
    return vti::kOk;
}

}  // extern "C"

extern "C" {

// modeling_f, lib/vti-fwi-2d/modeling_f.cu:10-86 (wave.h:14): one source at
// the grid index slocxz; level 1 is dt2*w[0] there, level it >= 2 adds
// w[it-1] with no dt2 (modeling_f.cu:66); after each recorded step the
// whole p field, tapered, goes to slot (it-ntNeg)/ratio of wavefield (nnxz
// floats per slot, (nt-ntNeg-1)/ratio+1 slots; the reference's second copy
// of p over the half-sum slice, modeling_f.cu:71-72, leaves p), and slot 1
// gets level 1 when ratio == 1 and ot == 0.  ratio = round(rate/dt) in
// float, as there.  Slots never recorded keep their input values.  The
// fields go to a device array of the recorded slots and come down in one
// copy at the end.
int vti2d_modeling(vti_ctx *ctx, float *wavefield, const float *c11, const float *c13, const float *c33,
                   const float *wavelet, int slocxz, const float *taper, int nx, int nz, int nt, int npad, float dx,
                   float dz, float dt, float rate, float ot)
{
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, false);
    if (rc) return rc;
    VTI_CHECK(wavefield && c11 && c13 && c33 && wavelet && taper, "vti2d_modeling: null argument");
    c.ratio = (int)std::round(rate / dt);
    VTI_CHECK(c.ratio >= 1, "vti2d_modeling: sampling rate below the time step");
    c.nnt_data = (nt - c.ntNeg - 1) / c.ratio + 1;
    VTI_CHECK(slocxz >= 0 && slocxz < c.nnxz && interior(c, slocxz % c.g.nnx, slocxz / c.g.nnx),
              "vti2d_modeling: source index %d outside the updated interior", slocxz);
    Fields2 F;
    rc = load_model(c, F, c11, c13, c33, taper);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    rc = zero2(c, F.p, F.q);
    if (rc) return rc;
    // the recorded slots form one run [lo, hi]
    std::vector<int> rows;
    if (c.ratio == 1 && ot == 0.f && 1 < c.nnt_data) rows.push_back(1);
    for (int it = 2; it < nt; ++it) {
        const float t = it * dt + ot;
        if (t >= 0.f && (it - c.ntNeg) % c.ratio == 0 && (it - c.ntNeg) / c.ratio < c.nnt_data)
            rows.push_back((it - c.ntNeg) / c.ratio);
    }
    const int lo = rows.empty() ? 0 : rows.front(), hi = rows.empty() ? -1 : rows.back();
    VTI_CHECK(rows.empty() || hi - lo + 1 == (int)rows.size(), "vti2d_modeling: recorded slots %d..%d not contiguous",
              lo, hi);
    const size_t n = (size_t)c.nnxz;
    DBuf<float> slots;
    rc = slots.alloc(std::max<size_t>(1, (size_t)(hi - lo + 1) * n), st, false);
    if (rc) return rc;
    auto keep = [&](int row) {
        return hipMemcpyAsync(slots.p + (size_t)(row - lo) * n, F.p[0].p, n * 4, hipMemcpyDeviceToDevice, st);
    };
    const long long src = slocxz;
    hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(1), 0, st, F.p[0].p, F.q[0].p, src, c.dt2 * wavelet[0] * taper[src]);
    if (c.ratio == 1 && ot == 0.f && 1 < c.nnt_data) VTI_HIP(keep(1));
    for (int it = 2; it < nt; ++it) {
        const float t = it * dt + ot;
        hipLaunchKernelGGL((vti::fwd2d_kernel<false>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                           F.q[0].p, nullptr, nullptr, nullptr, nullptr, F.c11.p, F.c13.p, F.c33.p, nullptr, nullptr,
                           nullptr, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, src, wavelet[it - 1], nullptr, nullptr, -1LL,
                           nullptr, nullptr);
        std::swap(F.p[0].p, F.p[1].p);
        std::swap(F.q[0].p, F.q[1].p);
        if (t >= 0.f && (it - c.ntNeg) % c.ratio == 0) {
            const int row = (it - c.ntNeg) / c.ratio;
            if (row < c.nnt_data) VTI_HIP(keep(row));
        }
    }
    VTI_HIP(hipGetLastError());
    if (!rows.empty())
        VTI_HIP(hipMemcpyAsync(wavefield + (size_t)lo * n, slots.p, (size_t)(hi - lo + 1) * n * 4,
                               hipMemcpyDeviceToHost, st));
    VTI_HIP(hipStreamSynchronize(st));
    return vti::kOk;
}

}  // extern "C"

// ---- synthetic_f without allocation ----------------------------------------
//
// vti2d_synthetic_noalloc runs the one-shot schedule of vti2d_synthetic with
// every device buffer carved out of `scratch`, a file-scope device array of
// fixed size, so that a call makes no device (or host) allocation.  Each GPU
// holds its own copy of the array until the process ends; calls on one GPU
// share it, so calls that overlap on one GPU race.
namespace {

constexpr long long kNaPoints = 1LL << 20;  // padded grid points (nnx * nnz)
constexpr long long kNaRec = 1LL << 16;     // receivers of one shot
constexpr long long kNaTrace = 1LL << 22;   // floats of one shot's traces (receivers x rows)
// layout in floats: c11 c13 c33 taper p0 p1 q0 q1 (kNaPoints each), the
// shot's receiver (ix, iz) pairs (2 kNaRec ints), its traces (kNaTrace)
constexpr long long kNaFloats = 8 * kNaPoints + 2 * kNaRec + kNaTrace;
constexpr int kNaMaxDevices = 64;

__device__ __align__(256) float scratch[kNaFloats];

// Device address of this device's `scratch` (the host-side symbol is only a
// handle), looked up once per device with that device current.
std::mutex g_na_mu;
float *g_na_ptr[kNaMaxDevices] = {};

int scratch_ptr(int device, float **out)
{
    VTI_CHECK(device >= 0 && device < kNaMaxDevices, "vti2d_synthetic_noalloc: device %d beyond %d", device,
              kNaMaxDevices);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        VTI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VTI_CHECK(p != nullptr, "vti2d_synthetic_noalloc: no device address for the scratch array");
        g_na_ptr[device] = static_cast<float *>(p);
    }
    *out = g_na_ptr[device];
    return vti::kOk;
}

__global__ void fill2_kernel(float *__restrict__ a, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = 0.f;
}

// recordData (kernels.cu:317-324) at receivers given as (ix, iz) pairs
__global__ void record2d_xz_kernel(float *__restrict__ row, const float *__restrict__ p, const float *__restrict__ q,
                                   const int *__restrict__ rxz, int nr, int nnx)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < nr) {
        const long long i = rxz[2 * ir] + (long long)rxz[2 * ir + 1] * nnx;
        row[ir] = 0.5f * (p[i] + q[i]);
    }
}

}  // namespace

extern "C" {

// vti2d_synthetic's results (the one-shot schedule, bit for bit) with no
// allocation: see vti_amd.h.
int vti2d_synthetic_noalloc(vti_ctx *ctx, float *data, const float *c11, const float *c13, const float *c33,
                            const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper,
                            int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, false);
    if (rc) return rc;
    VTI_CHECK(data && c11 && c13 && c33 && wavelet && sloc && rloc && taper && ns >= 0 && nr >= 0,
              "vti2d_synthetic_noalloc: null argument");
    VTI_CHECK(c.nnxz <= kNaPoints, "vti2d_synthetic_noalloc: %lld grid points above the scratch limit %lld", c.nnxz,
              kNaPoints);
    // every shot is checked before any work is queued
    for (int is = 0; is < ns; ++is) {
        const int ix = sloc[4 * is], iz = sloc[4 * is + 1], nri = sloc[4 * is + 2], r0 = sloc[4 * is + 3];
        VTI_CHECK(nri >= 0 && r0 >= 0 && r0 + (long long)nri <= nr, "shot %d: receiver range outside rloc", is);
        VTI_CHECK(interior(c, ix, iz), "shot %d: source (%d,%d) outside the updated interior", is, ix, iz);
        VTI_CHECK(nri <= kNaRec, "vti2d_synthetic_noalloc: shot %d has %d receivers, above the scratch limit %lld",
                  is, nri, kNaRec);
        VTI_CHECK((long long)nri * c.nnt_data <= kNaTrace,
                  "vti2d_synthetic_noalloc: shot %d needs %lld trace samples, above the scratch limit %lld", is,
                  (long long)nri * c.nnt_data, kNaTrace);
        for (int ir = 0; ir < nri; ++ir) {
            const int x = rloc[2 * (r0 + ir)], z = rloc[2 * (r0 + ir) + 1];
            VTI_CHECK(x >= 0 && x < c.g.nnx && z >= 0 && z < c.g.nnz, "receiver %d outside the grid", r0 + ir);
        }
    }
    float *base = nullptr;
    rc = scratch_ptr(ctx->device, &base);
    if (rc) return rc;
    const size_t n = (size_t)c.nnxz;
    float *m11 = base, *m13 = base + kNaPoints, *m33 = base + 2 * kNaPoints, *tap = base + 3 * kNaPoints;
    float *p[2] = {base + 4 * kNaPoints, base + 5 * kNaPoints};
    float *q[2] = {base + 6 * kNaPoints, base + 7 * kNaPoints};
    int *rxz = reinterpret_cast<int *>(base + 8 * kNaPoints);
    float *tr = base + 8 * kNaPoints + 2 * kNaRec;
    hipStream_t st = ctx->stream;
    rc = upload(c, m11, c11, n * 4);
    if (!rc) rc = upload(c, m13, c13, n * 4);
    if (!rc) rc = upload(c, m33, c33, n * 4);
    if (!rc) rc = upload(c, tap, taper, n * 4);
    if (rc) return rc;
    const int fill_blocks = launch_blocks((long long)n);
    for (int is = 0; is < ns; ++is) {
        const int ix = sloc[4 * is], iz = sloc[4 * is + 1], nri = sloc[4 * is + 2], r0 = sloc[4 * is + 3];
        const long long src = ix + (long long)iz * c.g.nnx;
        if (nri) rc = upload(c, rxz, rloc + 2 * (size_t)r0, sizeof(int) * 2 * (size_t)nri);
        if (rc) return rc;
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL(fill2_kernel, dim3(fill_blocks), dim3(256), 0, st, p[k], (long long)n);
            hipLaunchKernelGGL(fill2_kernel, dim3(fill_blocks), dim3(256), 0, st, q[k], (long long)n);
        }
        // level 1: dt2*w[0] at the source, then abc (synthetic_f.cu:86-89)
        hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(1), 0, st, p[0], q[0], src, c.dt2 * wavelet[0] * taper[src]);
        int row_lo = c.nnt_data, row_hi = -1, nrows = 0;
        auto record = [&](int row) {
            hipLaunchKernelGGL(record2d_xz_kernel, dim3((nri + 255) / 256), dim3(256), 0, st,
                               tr + (size_t)row * nri, p[0], q[0], rxz, nri, c.g.nnx);
            row_lo = std::min(row_lo, row);
            row_hi = std::max(row_hi, row);
            ++nrows;
        };
        if (c.ratio == 1 && ot == 0.f && nri && 1 < c.nnt_data) record(1);
        for (int it = 2; it < nt; ++it) {
            const float t = it * dt + ot;
            hipLaunchKernelGGL((vti::fwd2d_kernel<false>), c.grid, c.block, 0, st, p[1], q[1], p[0], q[0], nullptr,
                               nullptr, nullptr, nullptr, m11, m13, m33, nullptr, nullptr, nullptr, tap, c.g, c.dx2,
                               c.dz2, c.dt2, src, c.dt2 * wavelet[it - 1], nullptr, nullptr, -1LL, nullptr, nullptr);
            std::swap(p[0], p[1]);
            std::swap(q[0], q[1]);
            if (t >= 0. && (it - c.ntNeg) % c.ratio == 0 && nri) {
                const int row = (it - c.ntNeg) / c.ratio;
                if (row < c.nnt_data) record(row);
            }
        }
        VTI_HIP(hipGetLastError());
        // each row is recorded at most once, so the recorded rows are the
        // whole range [row_lo, row_hi] when their count fills it
        VTI_CHECK(nrows == 0 || nrows == row_hi - row_lo + 1,
                  "vti2d_synthetic_noalloc: recorded rows %d..%d not contiguous", row_lo, row_hi);
        if (nrows)
            VTI_HIP(hipMemcpy2DAsync(data + (size_t)row_lo * nr + r0, sizeof(float) * (size_t)nr,
                                     tr + (size_t)row_lo * nri, sizeof(float) * (size_t)nri,
                                     sizeof(float) * (size_t)nri, (size_t)nrows, hipMemcpyDeviceToHost, st));
        // the next shot's uploads and fills reuse the buffers
        VTI_HIP(hipStreamSynchronize(st));
    }
    return vti::kOk;
}

}  // extern "C"

namespace {

// objFuncGradient_f (born = false) and GNhessian_f (born = true).
// rtm_f (nhx < 0) / odcig_f (nhx >= 0): see vti_amd.h
int rtm2d(vti_ctx *ctx, int nhx, float *image, const float *data, const float *c11, const float *c13,
          const float *c33, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper,
          int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    const bool odcig = nhx >= 0;
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, true);
    if (rc) return rc;
    VTI_CHECK(image && data && c11 && c13 && c33 && wavelet && sloc && rloc && taper && ns >= 0, "null argument");
    VTI_CHECK((nt - 1) % c.ratio == 0,
              "(nt-1) %% (rate/dt) must be 0: otherwise the reference reads one source-wavefield slice past its "
              "array (got nt %d, ratio %d)", nt, c.ratio);
    VTI_CHECK(!odcig || 2 * nhx + 9 <= c.g.nnx, "offset half-width %d too large for %d columns", nhx, c.g.nnx);
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)c.nnxz;
    const int nh = odcig ? 2 * nhx + 1 : 1;
    Fields2 F;
    rc = load_model(c, F, c11, c13, c33, taper);
    if (rc) return rc;
    DBuf<float> slices, img;
    DBuf<double> img64;
    DBuf<int> recmap, dipmap;
    rc = slices.alloc((size_t)c.nnt * n, st);  // slot 0 stays zero, as the reference's
    if (!rc) rc = img.alloc(n * nh, st);
    if (!rc) rc = img64.alloc(n * nh, st);
    if (!rc) rc = recmap.alloc(n, st);
    if (!rc) rc = dipmap.alloc(n, st);
    if (rc) return rc;
    const dim3 rows((c.g.nnx - 8 + 255) / 256, c.g.nnz - 8);
    const int xlo = odcig ? std::max(4, nhx) : 4;
    const dim3 xrows((std::max(1, c.g.nnx - 2 * xlo) + 255) / 256, c.g.nnz - 8);
    for (int is = 0; is < ns; ++is) {
        Shot2 s;
        rc = shot_of(c, sloc, is, rloc, nr, s);
        if (rc) return rc;
        VTI_CHECK(interior(c, s.ix, s.iz + 1), "shot %d: the dipole's second pole (%d,%d) outside the interior", is,
                  s.ix, s.iz + 1);
        std::vector<int> hp(n, -1), hm(n, -1);
        for (int ir = 0; ir < s.nri; ++ir) {
            const int x = (int)(s.idx[ir] % c.g.nnx), z = (int)(s.idx[ir] / c.g.nnx);
            VTI_CHECK(interior(c, x, z) && interior(c, x, z + 1),
                      "receiver %d: (%d,%d) and the point below must lie inside the updated interior", s.r0 + ir, x,
                      z);
            VTI_CHECK(hp[s.idx[ir]] < 0, "receivers %d and %d share a grid point", hp[s.idx[ir]] + s.r0, s.r0 + ir);
            hp[s.idx[ir]] = ir;
            hm[s.idx[ir] + c.g.nnx] = ir;
        }
        for (size_t k = 0; k < n; ++k)
            VTI_CHECK(hp[k] < 0 || hm[k] < 0,
                      "receiver %d lies right below receiver %d: the reference's dipole injection races there",
                      hp[k] + s.r0, hm[k] + s.r0);
        DBuf<long long> idx;
        DBuf<float> tr;
        rc = idx.alloc(s.nri, st);
        // the shot's data rows, time-major, and one zero row past the last
        if (!rc) rc = tr.alloc((size_t)(c.nnt_data + 1) * std::max(1, s.nri), st);
        if (!rc && s.nri) rc = upload(c, idx.p, s.idx.data(), sizeof(long long) * s.nri);
        if (!rc) rc = upload(c, recmap.p, hp.data(), n * sizeof(int));
        if (!rc) rc = upload(c, dipmap.p, hm.data(), n * sizeof(int));
        std::vector<float> d((size_t)c.nnt_data * std::max(1, s.nri), 0.f);
        for (int row = 0; row < c.nnt_data; ++row)
            for (int ir = 0; ir < s.nri; ++ir) d[(size_t)row * s.nri + ir] = data[(size_t)row * nr + s.r0 + ir];
        if (!rc && s.nri) rc = upload(c, tr.p, d.data(), d.size() * 4);
        if (rc) return rc;
        VTI_HIP(hipMemsetAsync(img.p, 0, n * nh * 4, st));

        // ---- source side (rtm_f.cu:110-150)
        rc = zero2(c, F.p, F.q);
        if (rc) return rc;
        hipLaunchKernelGGL(dipole_level1_kernel, dim3(1), dim3(1), 0, st, F.p[0].p, F.q[0].p, F.taper.p, s.src,
                           c.g.nnx, c.dt2 * wavelet[0]);
        if (c.ratio == 1)
            hipLaunchKernelGGL(slice2d_kernel, rows, dim3(256), 0, st, slices.p + n, F.p[0].p, F.q[0].p, c.g);
        for (int it = 2; it < nt; ++it) {
            // odcig_f.cu:123 injects wavelet[it-1] without the dt2 of its first level
            const float sv = odcig ? wavelet[it - 1] : c.dt2 * wavelet[it - 1];
            hipLaunchKernelGGL((vti::fwd2d_kernel<false>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                               F.q[0].p, nullptr, nullptr, nullptr, nullptr, F.c11.p, F.c13.p, F.c33.p, nullptr,
                               nullptr, nullptr, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, s.src, sv, nullptr, nullptr,
                               s.src + c.g.nnx, nullptr, nullptr);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            if (it % c.ratio == 0)
                hipLaunchKernelGGL(slice2d_kernel, rows, dim3(256), 0, st, slices.p + (size_t)(it / c.ratio) * n,
                                   F.p[0].p, F.q[0].p, c.g);
        }
        VTI_HIP(hipGetLastError());

        // ---- receiver side and imaging (rtm_f.cu:152-216)
        rc = zero2(c, F.p, F.q);
        if (rc) return rc;
        if (s.nri)
            hipLaunchKernelGGL(dipole_first_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, F.p[0].p, F.q[0].p,
                               F.taper.p, idx.p, tr.p + (size_t)(c.nnt_data - 1) * s.nri, s.nri, c.g.nnx, c.dt2);
        auto slot = [&](int k) { return slices.p + (size_t)k * n; };
        {
            float f = (float)((nt - 2.) / c.ratio);
            const int i1 = (int)f;
            f = f - i1;
            if (odcig)
                hipLaunchKernelGGL(ximage2d_kernel, xrows, dim3(256), 0, st, img.p, F.p[0].p, F.q[0].p, slot(i1),
                                   slot(i1 + 1), f, c.g, nhx, xlo);
            else
                hipLaunchKernelGGL(image2d_kernel, rows, dim3(256), 0, st, img.p, F.p[0].p, F.q[0].p, slot(i1),
                                   slot(i1 + 1), f, c.g);
        }
        for (int it = nt - 3; it >= 0; --it) {
            // the data rows bracketing the level: row i1 and row i1+1 (zero past the last)
            int inj = 0, ir1 = 0;
            float fr = 0.f;
            if (it >= c.ntNeg && s.nri) {
                fr = (float)((it - c.ntNeg + 1.) / c.ratio);
                ir1 = (int)fr;
                fr = fr - ir1;
                inj = 1;
            }
            float f = (float)it / c.ratio;
            const int i1 = (int)f;
            f = f - i1;
            const float *r0 = tr.p + (size_t)ir1 * s.nri, *r1 = tr.p + (size_t)(ir1 + 1) * s.nri;
            if (odcig)
                hipLaunchKernelGGL((rtm_bwd2d_kernel<false>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, F.c11.p, F.c13.p, F.c33.p, F.taper.p, recmap.p, dipmap.p, inj, r0, r1, fr,
                                   slot(i1), slot(i1 + 1), f, img.p, c.g, c.dx2, c.dz2, c.dt2);
            else
                hipLaunchKernelGGL((rtm_bwd2d_kernel<true>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, F.c11.p, F.c13.p, F.c33.p, F.taper.p, recmap.p, dipmap.p, inj, r0, r1, fr,
                                   slot(i1), slot(i1 + 1), f, img.p, c.g, c.dx2, c.dz2, c.dt2);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            if (odcig)
                hipLaunchKernelGGL(ximage2d_kernel, xrows, dim3(256), 0, st, img.p, F.p[0].p, F.q[0].p, slot(i1),
                                   slot(i1 + 1), f, c.g, nhx, xlo);
        }
        hipLaunchKernelGGL(acc64_2d_kernel, dim3(launch_blocks((long long)(n * nh))), dim3(256), 0, st, img64.p, img.p,
                           (long long)(n * nh));
        VTI_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(narrow2d_kernel, dim3(launch_blocks((long long)(n * nh))), dim3(256), 0, st, img.p, img64.p,
                       (long long)(n * nh));
    VTI_HIP(hipMemcpyAsync(image, img.p, n * nh * 4, hipMemcpyDeviceToHost, st));
    VTI_HIP(hipStreamSynchronize(st));
    return vti::kOk;
}

int fg2d(vti_ctx *ctx, bool born, float *objfunc, float *gc11, float *gc13, float *gc33, const float *data,
         const float *c11, const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
         const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz,
         int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, true);
    if (rc) return rc;
    VTI_CHECK(gc11 && gc13 && gc33 && c11 && c13 && c33 && wavelet && sloc && rloc && taper && ns >= 0,
              "null argument");
    VTI_CHECK(born || data, "objective/gradient needs the observed data");
    VTI_CHECK(!born || (dc11 && dc13 && dc33), "Hessian-vector product needs dc11, dc13, dc33");
    VTI_CHECK(c.ratio >= 2 && (nt - 1) % c.ratio == 0 && c.ntNeg == 0,
              "the reference's snapshot rotation lines up only for rate/dt >= 2, (nt-1) %% (rate/dt) == 0 and ot == 0 "
              "(got ratio %d, nt %d, ot %g)",
              c.ratio, nt, (double)ot);
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)c.nnxz;
    Fields2 F, B;
    rc = load_model(c, F, c11, c13, c33, taper);
    if (rc) return rc;
    DBuf<float> dc[3], snap, acc[3];
    DBuf<double> acc64[3];
    DBuf<int> recmap;
    if (born) {
        for (int k = 0; k < 2 && !rc; ++k) {
            rc = B.p[k].alloc(n, st);
            if (!rc) rc = B.q[k].alloc(n, st);
        }
        const float *src[3] = {dc11, dc13, dc33};
        for (int k = 0; k < 3 && !rc; ++k) {
            rc = dc[k].alloc(n, st);
            if (!rc) rc = upload(c, dc[k].p, src[k], n * 4);
        }
    }
    // [k][p | q]: slots 1.. are written before they are read; slot 0 (D of
    // the zero level 0) is never written and is the only one cleared
    if (!rc) rc = snap.alloc((size_t)c.nnt * 2 * n, st, false);
    if (!rc) VTI_HIP(hipMemsetAsync(snap.p, 0, 2 * n * sizeof(float), st));
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = acc[k].alloc(n, st);
        if (!rc) rc = acc64[k].alloc(n, st);
    }
    if (!rc) rc = recmap.alloc(n, st);
    if (rc) return rc;
    auto snap_p = [&](int k) { return snap.p + (size_t)k * 2 * n; };
    auto snap_q = [&](int k) { return snap.p + (size_t)k * 2 * n + n; };
    double obj = 0.0;
    std::vector<int> hmap(n, -1);
    for (int is = 0; is < ns; ++is) {
        Shot2 s;
        rc = shot_of(c, sloc, is, rloc, nr, s);
        if (rc) return rc;
        for (int ir = 0; ir < s.nri; ++ir) {
            const int x = (int)(s.idx[ir] % c.g.nnx), z = (int)(s.idx[ir] / c.g.nnx);
            VTI_CHECK(interior(c, x, z), "receiver %d outside the updated interior", s.r0 + ir);
            VTI_CHECK(hmap[s.idx[ir]] < 0, "receivers %d and %d share a grid point", hmap[s.idx[ir]] + s.r0, s.r0 + ir);
            hmap[s.idx[ir]] = ir;
        }
        DBuf<long long> idx;
        DBuf<float> bg, obs, zrow;
        DBuf<double> part;
        const int nblk = launch_blocks((long long)s.nri * c.nnt_data);
        rc = idx.alloc(s.nri, st);
        if (!rc) rc = bg.alloc((size_t)s.nri * c.nnt_data, st);
        if (!rc) rc = zrow.alloc(s.nri, st);
        if (!rc) rc = part.alloc(nblk, st);
        if (!rc && s.nri) rc = upload(c, idx.p, s.idx.data(), sizeof(long long) * s.nri);
        if (!rc) rc = upload(c, recmap.p, hmap.data(), n * sizeof(int));
        if (rc) return rc;

        // ---- forward (objFuncGradient_f.cu:160-210 / hessian_f.cu:735-806)
        rc = zero2(c, F.p, F.q);
        if (!rc && born) rc = zero2(c, B.p, B.q);
        if (rc) return rc;
        hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(1), 0, st, F.p[0].p, F.q[0].p, s.src, c.dt2 * wavelet[0]);
        for (int it = 2; it < nt; ++it) {
            // the step computing level it reads level it-1: snapshot its D when due
            const int lc = it - 1;
            float *sp = nullptr, *sq = nullptr;
            if (lc % c.ratio == 0) {
                sp = snap_p(lc / c.ratio);
                sq = snap_q(lc / c.ratio);
            }
            const float sv = c.dt2 * wavelet[it - 1];
            if (born)
                hipLaunchKernelGGL((vti::fwd2d_kernel<true>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, B.p[1].p, B.q[1].p, B.p[0].p, B.q[0].p, F.c11.p, F.c13.p, F.c33.p,
                                   dc[0].p, dc[1].p, dc[2].p, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, s.src, sv, sp, sq, -1LL, nullptr, nullptr);
            else
                hipLaunchKernelGGL((vti::fwd2d_kernel<false>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, nullptr, nullptr, nullptr, nullptr, F.c11.p, F.c13.p, F.c33.p, nullptr,
                                   nullptr, nullptr, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, s.src, sv, sp, sq, -1LL, nullptr, nullptr);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            if (born) {
                std::swap(B.p[0].p, B.p[1].p);
                std::swap(B.q[0].p, B.q[1].p);
            }
            if (it % c.ratio == 0 && s.nri) {
                const DBuf<float> *rp = born ? &B.p[0] : &F.p[0];
                const DBuf<float> *rq = born ? &B.q[0] : &F.q[0];
                hipLaunchKernelGGL(record2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st,
                                   bg.p + (size_t)(it / c.ratio) * s.nri, rp->p, rq->p, idx.p, s.nri);
            }
        }
        hipLaunchKernelGGL(vti::dsnap2d_kernel, c.grid, c.block, 0, st, F.p[0].p, F.q[0].p, c.g, c.dx2, c.dz2,
                           snap_p(c.nnt - 1), snap_q(c.nnt - 1));
        VTI_HIP(hipGetLastError());

        // ---- residual and objective (objFuncGradient_f.cu:214-219)
        if (!born) {
            std::vector<float> o((size_t)s.nri * c.nnt_data);
            for (int row = 0; row < c.nnt_data; ++row)
                for (int ir = 0; ir < s.nri; ++ir) o[(size_t)row * s.nri + ir] = data[(size_t)row * nr + s.r0 + ir];
            rc = obs.alloc(o.size(), st);
            if (!rc && !o.empty()) rc = upload(c, obs.p, o.data(), o.size() * 4);
            if (rc) return rc;
            hipLaunchKernelGGL(residual2d_kernel, dim3(nblk), dim3(256), 0, st, bg.p, obs.p,
                               (long long)s.nri * c.nnt_data, part.p);
            std::vector<double> hp(nblk);
            VTI_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
            VTI_HIP(hipStreamSynchronize(st));
            for (double v : hp) obj += v;
        }

        // ---- backward (objFuncGradient_f.cu:225-300 / hessian_f.cu:834-900)
        rc = zero2(c, F.p, F.q);
        if (rc) return rc;
        if (s.nri)
            hipLaunchKernelGGL(inject2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, F.p[0].p, F.q[0].p,
                               idx.p, bg.p + (size_t)(c.nnt_data - 1) * s.nri, zrow.p, 0.f, s.nri, c.dt2);
        {
            float f = (float)((nt - 2.) / c.ratio);
            const int i1 = (int)f;
            f = f - i1;
            hipLaunchKernelGGL(vti::xcor2d_kernel, dim3((c.g.nnx - 8 + 255) / 256, c.g.nnz - 8), dim3(256), 0, st,
                               F.p[0].p, F.q[0].p, snap_p(i1), snap_q(i1), snap_p(i1 + 1), snap_q(i1 + 1), f,
                               acc[0].p, acc[1].p, acc[2].p, c.g);
        }
        for (int it = nt - 3; it >= 0; --it) {
            float fr = (float)((it + 1.) / c.ratio);
            const int ir1 = (int)fr;
            fr = fr - ir1;
            float f = (float)it / c.ratio;
            const int i1 = (int)f;
            f = f - i1;
            const float *r0 = bg.p + (size_t)ir1 * s.nri, *r1 = bg.p + (size_t)(ir1 + 1) * s.nri;
            if (born)
                hipLaunchKernelGGL((vti::bwd2d_kernel<1>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, F.c11.p, F.c13.p, F.c33.p, F.taper.p, recmap.p, r0, r1, fr, snap_p(i1),
                                   snap_q(i1), snap_p(i1 + 1), snap_q(i1 + 1), f, acc[0].p, acc[1].p, acc[2].p, c.g,
                                   c.dx2, c.dz2, c.dt2);
            else
                hipLaunchKernelGGL((vti::bwd2d_kernel<0>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                                   F.q[0].p, F.c11.p, F.c13.p, F.c33.p, F.taper.p, recmap.p, r0, r1, fr, snap_p(i1),
                                   snap_q(i1), snap_p(i1 + 1), snap_q(i1 + 1), f, acc[0].p, acc[1].p, acc[2].p, c.g,
                                   c.dx2, c.dz2, c.dt2);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
        }
        VTI_HIP(hipGetLastError());
        // per-shot FP32 gradient into the FP64 cross-shot sum
        for (int k = 0; k < 3; ++k) {
            hipLaunchKernelGGL(acc64_2d_kernel, dim3(launch_blocks((long long)n)), dim3(256), 0, st, acc64[k].p,
                               acc[k].p, (long long)n);
            VTI_HIP(hipMemsetAsync(acc[k].p, 0, n * 4, st));
        }
        for (int ir = 0; ir < s.nri; ++ir) hmap[s.idx[ir]] = -1;
        VTI_HIP(hipStreamSynchronize(st));
    }
    float *out[3] = {gc11, gc13, gc33};
    for (int k = 0; k < 3; ++k) {
        hipLaunchKernelGGL(narrow2d_kernel, dim3(launch_blocks((long long)n)), dim3(256), 0, st, acc[k].p, acc64[k].p,
                           (long long)n);
        VTI_HIP(hipMemcpyAsync(out[k], acc[k].p, n * 4, hipMemcpyDeviceToHost, st));
    }
    VTI_HIP(hipStreamSynchronize(st));
    // objFuncGradient_f returns 0.5f*objfunc as float (objFuncGradient_f.cu:370)
    if (objfunc) *objfunc = (float)(0.5f * obj);
    return vti::kOk;
}

// hessian_f (lib/vti-fwi-2d/hessian_f.cu:15-573): the full Hessian-vector
// product, per shot:
//  - forward: background and Born fields as GNhessian_f, recording both traces
//    and the D snapshots of both at every rate/dt levels;
//  - the residual r = background traces - data;
//  - backward: the residual's adjoint lambda and its perturbation dlambda
//    (driven by the Born traces and by dt2 * D(dC lambda)), both with the
//    exact adjoint backwardDC, imaged against the Born and the background
//    snapshots respectively.
// The reference routes the Born snapshots through its pipeline with the first
// host buffer filled from the background ones (hessian_f.cu:381-382); every use
// of that buffer is an image of Born slot nnt-3, so the background slot
// nnt-3 is copied over the Born one before the backward pass.
int hess2d_full(vti_ctx *ctx, float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
                const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    Call2 c;
    int rc = setup(c, ctx, nx, nz, nt, npad, dx, dz, dt, rate, ot, true);
    if (rc) return rc;
    VTI_CHECK(gc11 && gc13 && gc33 && data && c11 && c13 && c33 && dc11 && dc13 && dc33 && wavelet && sloc && rloc &&
                  taper && ns >= 0,
              "null argument");
    VTI_CHECK(c.ratio >= 2 && (nt - 1) % c.ratio == 0 && c.ntNeg == 0,
              "the reference's snapshot rotation lines up only for rate/dt >= 2, (nt-1) %% (rate/dt) == 0 and ot == 0 "
              "(got ratio %d, nt %d, ot %g)",
              c.ratio, nt, (double)ot);
    VTI_CHECK(c.nnt >= 4, "hessian_f reads snapshot slot nnt-3 of at least 4 (got nnt %d)", c.nnt);
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)c.nnxz;
    Fields2 F, B;
    rc = load_model(c, F, c11, c13, c33, taper);
    if (rc) return rc;
    DBuf<float> dc[3], snap, bsnap, acc[3];
    DBuf<double> acc64[3];
    DBuf<int> recmap;
    for (int k = 0; k < 2 && !rc; ++k) {
        rc = B.p[k].alloc(n, st);
        if (!rc) rc = B.q[k].alloc(n, st);
    }
    const float *src[3] = {dc11, dc13, dc33};
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = dc[k].alloc(n, st);
        if (!rc) rc = upload(c, dc[k].p, src[k], n * 4);
    }
    if (!rc) rc = snap.alloc((size_t)c.nnt * 2 * n, st);
    if (!rc) rc = bsnap.alloc((size_t)c.nnt * 2 * n, st);
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = acc[k].alloc(n, st);
        if (!rc) rc = acc64[k].alloc(n, st);
    }
    if (!rc) rc = recmap.alloc(n, st);
    if (rc) return rc;
    auto sp = [&](DBuf<float> &b, int k) { return b.p + (size_t)k * 2 * n; };
    auto sq = [&](DBuf<float> &b, int k) { return b.p + (size_t)k * 2 * n + n; };
    std::vector<int> hmap(n, -1);
    for (int is = 0; is < ns; ++is) {
        Shot2 s;
        rc = shot_of(c, sloc, is, rloc, nr, s);
        if (rc) return rc;
        for (int ir = 0; ir < s.nri; ++ir) {
            const int x = (int)(s.idx[ir] % c.g.nnx), z = (int)(s.idx[ir] / c.g.nnx);
            VTI_CHECK(interior(c, x, z), "receiver %d outside the updated interior", s.r0 + ir);
            VTI_CHECK(hmap[s.idx[ir]] < 0, "receivers %d and %d share a grid point", hmap[s.idx[ir]] + s.r0, s.r0 + ir);
            hmap[s.idx[ir]] = ir;
        }
        DBuf<long long> idx;
        DBuf<float> bg, bd, obs, zrow;
        DBuf<double> part;
        const int nblk = launch_blocks((long long)s.nri * c.nnt_data);
        rc = idx.alloc(s.nri, st);
        if (!rc) rc = bg.alloc((size_t)s.nri * c.nnt_data, st);
        if (!rc) rc = bd.alloc((size_t)s.nri * c.nnt_data, st);
        if (!rc) rc = zrow.alloc(s.nri, st);
        if (!rc) rc = part.alloc(nblk, st);
        if (!rc && s.nri) rc = upload(c, idx.p, s.idx.data(), sizeof(long long) * s.nri);
        if (!rc) rc = upload(c, recmap.p, hmap.data(), n * sizeof(int));
        if (rc) return rc;

        // ---- forward (hessian_f.cu:204-280)
        rc = zero2(c, F.p, F.q);
        if (!rc) rc = zero2(c, B.p, B.q);
        if (rc) return rc;
        hipLaunchKernelGGL(set2_kernel, dim3(1), dim3(1), 0, st, F.p[0].p, F.q[0].p, s.src, c.dt2 * wavelet[0]);
        for (int it = 2; it < nt; ++it) {
            const int lc = it - 1;
            const bool due = lc % c.ratio == 0;
            const float sv = c.dt2 * wavelet[it - 1];
            hipLaunchKernelGGL((vti::fwd2d_kernel<true>), c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                               F.q[0].p, B.p[1].p, B.q[1].p, B.p[0].p, B.q[0].p, F.c11.p, F.c13.p, F.c33.p, dc[0].p,
                               dc[1].p, dc[2].p, F.taper.p, c.g, c.dx2, c.dz2, c.dt2, s.src, sv,
                               due ? sp(snap, lc / c.ratio) : nullptr, due ? sq(snap, lc / c.ratio) : nullptr, -1LL,
                               due ? sp(bsnap, lc / c.ratio) : nullptr, due ? sq(bsnap, lc / c.ratio) : nullptr);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            std::swap(B.p[0].p, B.p[1].p);
            std::swap(B.q[0].p, B.q[1].p);
            if (it % c.ratio == 0 && s.nri) {
                const size_t row = (size_t)(it / c.ratio) * s.nri;
                hipLaunchKernelGGL(record2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, bg.p + row,
                                   F.p[0].p, F.q[0].p, idx.p, s.nri);
                hipLaunchKernelGGL(record2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, bd.p + row,
                                   B.p[0].p, B.q[0].p, idx.p, s.nri);
            }
        }
        hipLaunchKernelGGL(vti::dsnap2d_kernel, c.grid, c.block, 0, st, F.p[0].p, F.q[0].p, c.g, c.dx2, c.dz2,
                           sp(snap, c.nnt - 1), sq(snap, c.nnt - 1));
        hipLaunchKernelGGL(vti::dsnap2d_kernel, c.grid, c.block, 0, st, B.p[0].p, B.q[0].p, c.g, c.dx2, c.dz2,
                           sp(bsnap, c.nnt - 1), sq(bsnap, c.nnt - 1));
        VTI_HIP(hipMemcpyAsync(sp(bsnap, c.nnt - 3), sp(snap, c.nnt - 3), 2 * n * 4, hipMemcpyDeviceToDevice, st));
        VTI_HIP(hipGetLastError());

        // ---- residual (hessian_f.cu:300-304)
        {
            std::vector<float> o((size_t)s.nri * c.nnt_data);
            for (int row = 0; row < c.nnt_data; ++row)
                for (int ir = 0; ir < s.nri; ++ir) o[(size_t)row * s.nri + ir] = data[(size_t)row * nr + s.r0 + ir];
            rc = obs.alloc(o.size(), st);
            if (!rc && !o.empty()) rc = upload(c, obs.p, o.data(), o.size() * 4);
            if (rc) return rc;
            hipLaunchKernelGGL(residual2d_kernel, dim3(nblk), dim3(256), 0, st, bg.p, obs.p,
                               (long long)s.nri * c.nnt_data, part.p);
        }

        // ---- backward (hessian_f.cu:306-443)
        rc = zero2(c, F.p, F.q);
        if (!rc) rc = zero2(c, B.p, B.q);
        if (rc) return rc;
        if (s.nri) {
            hipLaunchKernelGGL(inject2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, F.p[0].p, F.q[0].p,
                               idx.p, bg.p + (size_t)(c.nnt_data - 1) * s.nri, zrow.p, 0.f, s.nri, c.dt2);
            hipLaunchKernelGGL(inject2d_kernel, dim3((s.nri + 255) / 256), dim3(256), 0, st, B.p[0].p, B.q[0].p,
                               idx.p, bd.p + (size_t)(c.nnt_data - 1) * s.nri, zrow.p, 0.f, s.nri, c.dt2);
        }
        {
            float f = (float)((nt - 2.) / c.ratio);
            const int i1 = (int)f;
            f = f - i1;
            const dim3 xg((c.g.nnx - 8 + 255) / 256, c.g.nnz - 8);
            hipLaunchKernelGGL(vti::xcor2d_kernel, xg, dim3(256), 0, st, F.p[0].p, F.q[0].p, sp(bsnap, i1),
                               sq(bsnap, i1), sp(bsnap, i1 + 1), sq(bsnap, i1 + 1), f, acc[0].p, acc[1].p, acc[2].p,
                               c.g);
            hipLaunchKernelGGL(vti::xcor2d_kernel, xg, dim3(256), 0, st, B.p[0].p, B.q[0].p, sp(snap, i1),
                               sq(snap, i1), sp(snap, i1 + 1), sq(snap, i1 + 1), f, acc[0].p, acc[1].p, acc[2].p, c.g);
        }
        for (int it = nt - 3; it >= 0; --it) {
            float fr = (float)((it + 1.) / c.ratio);
            const int ir1 = (int)fr;
            fr = fr - ir1;
            float f = (float)it / c.ratio;
            const int i1 = (int)f;
            f = f - i1;
            const size_t o0 = (size_t)ir1 * s.nri, o1 = (size_t)(ir1 + 1) * s.nri;
            hipLaunchKernelGGL(vti::hess_bwd2d_kernel, c.grid, c.block, 0, st, F.p[1].p, F.q[1].p, F.p[0].p,
                               F.q[0].p, B.p[1].p, B.q[1].p, B.p[0].p, B.q[0].p, F.c11.p, F.c13.p, F.c33.p, dc[0].p,
                               dc[1].p, dc[2].p, F.taper.p, recmap.p, bg.p + o0, bg.p + o1, bd.p + o0, bd.p + o1, fr,
                               sp(bsnap, i1), sq(bsnap, i1), sp(bsnap, i1 + 1), sq(bsnap, i1 + 1), sp(snap, i1),
                               sq(snap, i1), sp(snap, i1 + 1), sq(snap, i1 + 1), f, acc[0].p, acc[1].p, acc[2].p, c.g,
                               c.dx2, c.dz2, c.dt2);
            std::swap(F.p[0].p, F.p[1].p);
            std::swap(F.q[0].p, F.q[1].p);
            std::swap(B.p[0].p, B.p[1].p);
            std::swap(B.q[0].p, B.q[1].p);
        }
        VTI_HIP(hipGetLastError());
        for (int k = 0; k < 3; ++k) {
            hipLaunchKernelGGL(acc64_2d_kernel, dim3(launch_blocks((long long)n)), dim3(256), 0, st, acc64[k].p,
                               acc[k].p, (long long)n);
            VTI_HIP(hipMemsetAsync(acc[k].p, 0, n * 4, st));
        }
        for (int ir = 0; ir < s.nri; ++ir) hmap[s.idx[ir]] = -1;
        VTI_HIP(hipStreamSynchronize(st));
    }
    float *out[3] = {gc11, gc13, gc33};
    for (int k = 0; k < 3; ++k) {
        hipLaunchKernelGGL(narrow2d_kernel, dim3(launch_blocks((long long)n)), dim3(256), 0, st, acc[k].p, acc64[k].p,
                           (long long)n);
        VTI_HIP(hipMemcpyAsync(out[k], acc[k].p, n * 4, hipMemcpyDeviceToHost, st));
    }
    VTI_HIP(hipStreamSynchronize(st));
    return vti::kOk;
}

}  // namespace

extern "C" {

// objFuncGradient_f, lib/vti-fwi-2d/objFuncGradient_f.cu:15-371 (wave.h:12).
// gc11/gc13/gc33 (nnx*nnz) are overwritten; *objfunc = 0.5*sum r^2.
int vti2d_objfunc_gradient(vti_ctx *ctx, float *objfunc, float *gc11, float *gc13, float *gc33, const float *data,
                           const float *c11, const float *c13, const float *c33, const float *wavelet, const int *sloc,
                           int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt, int npad,
                           float dx, float dz, float dt, float rate, float ot)
{
    // Several shots with the standard taper run as one batch (two levels per
    // launch for all of a wave's shots); each shot's gradient and objective
    // are the one-shot loop's, summed in the same order, so the result is the
    // same bit for bit.
    // A context set to 16-bit snapshots sends single shots there too.
    if (ctx && (ns >= 2 || (ns == 1 && ctx->snap_fmt != 0)) && objfunc && gc11 && gc13 && gc33 && data && c11 &&
        c13 && c33 && wavelet && dx > 0 && dz > 0 &&
        vti::grad2d_batchable(sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dt, rate, ot)) {
        vti::Grad2dPart part;
        part.c[0] = c11;
        part.c[1] = c13;
        part.c[2] = c33;
        part.data = data;
        part.sloc = sloc;
        part.ns = ns;
        part.rloc = rloc;
        part.nr = nr;
        part.nx = nx;
        part.objfunc = objfunc;
        part.g[0] = gc11;
        part.g[1] = gc13;
        part.g[2] = gc33;
        return vti::grad2d_merged(ctx, &part, 1, wavelet, nz, nt, npad, dx, dz, dt, rate, 8, ctx->snap_fmt, true);
    }
    return fg2d(ctx, false, objfunc, gc11, gc13, gc33, data, c11, c13, c33, nullptr, nullptr, nullptr, wavelet, sloc,
                ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
}

// GNhessian_f, lib/vti-fwi-2d/hessian_f.cu:575-1010 (wave.h:34).  The data
// argument is unused there as well (the adjoint source is the Born data).
int vti2d_gn_hessian(vti_ctx *ctx, float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                     const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
                     const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                     int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    (void)data;
    return fg2d(ctx, true, nullptr, gc11, gc13, gc33, nullptr, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, ns,
                rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
}

// hessian_f, lib/vti-fwi-2d/hessian_f.cu:15-573 (wave.h:32).
int vti2d_hessian(vti_ctx *ctx, float *gc11, float *gc13, float *gc33, const float *data, const float *c11,
                  const float *c13, const float *c33, const float *dc11, const float *dc13, const float *dc33,
                  const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                  int nz, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    return hess2d_full(ctx, gc11, gc13, gc33, data, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, ns, rloc, nr,
                       taper, nx, nz, nt, npad, dx, dz, dt, rate, ot);
}

// rtm_f, lib/vti-fwi-2d/rtm_f.cu:13-269 (wave.h:6).  image: nnx*nnz.
int vti2d_rtm(vti_ctx *ctx, float *image, const float *data, const float *c11, const float *c13, const float *c33,
              const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz,
              int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    return rtm2d(ctx, -1, image, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz, dt,
                 rate, ot);
}

// odcig_f, lib/vti-fwi-2d/odcig_f.cu:12-247 (wave.h:16).  image: (2*nhx+1)
// planes of nnx*nnz, offset -nhx first.
int vti2d_odcig(vti_ctx *ctx, float *image, const float *data, const float *c11, const float *c13, const float *c33,
                const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx,
                int nz, int nhx, int nt, int npad, float dx, float dz, float dt, float rate, float ot)
{
    VTI_CHECK(nhx >= 0, "offset half-width must be >= 0, got %d", nhx);
    return rtm2d(ctx, nhx, image, data, c11, c13, c33, wavelet, sloc, ns, rloc, nr, taper, nx, nz, nt, npad, dx, dz,
                 dt, rate, ot);
}

}  // extern "C"
