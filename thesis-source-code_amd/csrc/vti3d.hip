// 3D VTI propagation on one MI355X: device-resident propagator, the
// reference-signature drivers (modelDataCij3d_f, objFuncCij3d_f,
// objFuncGradientCij3d_f, rtmCij3d_f) and the step-level entry points used by
// the benchmark and the serving layer.
//
// Everything a shot touches stays in HBM for the whole call: two time levels
// of (sigma_x, sigma_z) and (lambda_x, lambda_z), the coefficient volumes
// with the absorbing and with the random boundary shell, the damping
// profile, the traces and the accumulators.  This replaces the reference's
// host-streamed, temporally blocked z-block pipeline
// (objFuncGradient3d_f.cu:538-1078) while reproducing its results.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vti3d_kernels.h"
#include "vti_ctx.h"
#include "vti_host.h"
#include "vti_internal.h"

using vti::Geo3;

namespace {

constexpr int kBY = 8;  // tile rows per workgroup (64 x kBY lanes)

__global__ void set_point_kernel(float *a, float *b, long long i, float v)
{
    a[i] = v;
    b[i] = v;
}

// recordData (injectRecord3d.cu:22-28): 2/3 sx + 1/3 sz in double, stored as
// float; one time-major row of traces per call.
__global__ void record3d_kernel(float *__restrict__ row, const float *__restrict__ sx, const float *__restrict__ sz,
                                const long long *__restrict__ rec, int nr)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < nr) {
        const long long i = rec[ir];
        row[ir] = (float)(VTI_TWOTHIRD_D * (double)sx[i] + VTI_ONETHIRD_D * (double)sz[i]);
    }
}

// out[c * rows + r] = in[r * cols + c]: traces between the time-major rows
// the kernels record ([it][ir]) and the trace-major layout of the drivers'
// arguments ([ir][it]), through a 32 x 32 LDS tile (padded against bank
// conflicts); blocks of 32 x 8 threads, one block per tile, tiles numbered
// column tile fastest (transpose_tiles).
__global__ __launch_bounds__(256) void transpose_kernel(float *__restrict__ out, const float *__restrict__ in,
                                                        int rows, int cols)
{
    __shared__ float t[32][33];
    const int ntc = (cols + 31) / 32;
    const int c0 = (int)(blockIdx.x % ntc) * 32, r0 = (int)(blockIdx.x / ntc) * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) t[k][threadIdx.x] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = t[threadIdx.x][k];
    }
}

dim3 transpose_tiles(int rows, int cols) { return dim3((unsigned)(((cols + 31) / 32) * (long long)((rows + 31) / 32))); }

// First adjoint sample (objFuncGradient3d_f.cu:921-927): lambda_cur at the
// true receiver index = (2/3, 1/3) * dt2 * r[nnt-1] (assignment).
__global__ void lambda_init3d_kernel(float *__restrict__ lx, float *__restrict__ lz, const long long *__restrict__ rec,
                                     const float *__restrict__ last_row, int nr, float dt2)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir < nr) {
        const float t = dt2 * last_row[ir];
        lx[rec[ir]] = (float)(VTI_TWOTHIRD_D * (double)t);
        lz[rec[ir]] = (float)(VTI_ONETHIRD_D * (double)t);
    }
}

// Mute, optional t^p gain (gain[it] = (ot + it*rate)^tpow, formed on the host
// with the host's powf as tpower does, agc.cpp:103-115), residual against the
// observed traces and the per-block partial sums of r*r
// (objFuncGradient3d_f.cu:736-779).  Traces are
// time-major [it][ir]; obs is the reference's trace-major data file slice.
__global__ void residual3d_kernel(float *__restrict__ tr, const float *__restrict__ obs,
                                  const unsigned char *__restrict__ live, int nr, int nnt,
                                  const float *__restrict__ gain, int subtract, double *__restrict__ partial)
{
    __shared__ double red[256];
    const long long n = (long long)nr * nnt;
    double acc = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const int it = (int)(k / nr), ir = (int)(k % nr);
        float m = live[ir] ? tr[k] : 0.f;
        if (gain) m *= gain[it];
        const float r = subtract ? m - obs[(long long)ir * nnt + it] : m;
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

// residualAGC (lib/vti-fwi-3d/gpu/agc.cpp:70-101) on one receiver per
// thread: the automatic-gain-controlled residual h(d) - h(d0) with
// h(d)_i = d_i / sqrt(max(1e-6, sum of d_j^2 over the circular window
// |i-j| <= hw)), its sum of squares (fi_out[ir]) and, in place of the traces,
// the adjoint source.  Sums run in the reference's order; divisions and
// square roots are correctly rounded, as on the host.
__device__ __forceinline__ float agc_max(float a, float b) { return a < b ? b : a; }

__global__ void agc3d_kernel(float *__restrict__ tr, const float *__restrict__ obs, int nr, int nnt, int hw,
                             float *__restrict__ rsb, float *__restrict__ drb, double *__restrict__ fi_out)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= nr) return;
    const float eps = 1e-6f;
    auto wrap = [&](int j) { return j < 0 ? j + nnt : (j >= nnt ? j - nnt : j); };
    auto D = [&](int i) { return tr[(long long)i * nr + ir]; };
    auto O = [&](int i) { return obs[(long long)ir * nnt + i]; };
    double fi = 0.0;
    for (int i = 0; i < nnt; ++i) {
        float s = 0.f, s0 = 0.f;
        for (int j = i - hw; j <= i + hw; ++j) {
            const float v = D(wrap(j)), v0 = O(wrap(j));
            s += v * v;
            s0 += v0 * v0;
        }
        const float hd = D(i) / sqrtf(agc_max(eps, s));
        const float hd0 = O(i) / sqrtf(agc_max(eps, s0));
        const float res = hd - hd0;
        fi += (double)(res * res);
        const float t1 = agc_max(eps, s);
        const float t2 = res / sqrtf(t1);
        drb[(long long)i * nr + ir] = D(i) * t2 / t1;
        rsb[(long long)i * nr + ir] = t2;
    }
    for (int i = 0; i < nnt; ++i) {
        float s = 0.f;
        for (int j = i - hw; j <= i + hw; ++j) s += drb[(long long)wrap(j) * nr + ir];
        tr[(long long)i * nr + ir] = rsb[(long long)i * nr + ir] - D(i) * s;
    }
    fi_out[ir] = fi;
}

__global__ void accumulate64_kernel(double *__restrict__ acc, const float *__restrict__ v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc[i] += (double)v[i];
}

__global__ void narrow64_kernel(float *__restrict__ out, const double *__restrict__ acc, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)acc[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// Device-resident propagator
// ---------------------------------------------------------------------------
struct vti3d_prop {
    vti_ctx *ctx = nullptr;
    Geo3 g{};
    int npad = 0;
    float dx = 0, dy = 0, dz = 0, dt = 0;
    float dt2 = 0, dt2dx2 = 0, dt2dy2 = 0, dt2dz2 = 0, dx2 = 0, dy2 = 0, dz2 = 0;
    // coefficient sets: [0] the model as uploaded (absorbing-boundary shell in
    // the gradient driver), [1] the random-boundary shell (gradient driver)
    float *cm[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int active = 0;  // set the kernels read
    float *dxy = nullptr, *dzv = nullptr;
    float *sx[2] = {nullptr, nullptr}, *sz[2] = {nullptr, nullptr};  // [cur, prev]
    float *lx[2] = {nullptr, nullptr}, *lz[2] = {nullptr, nullptr};
    float *acc[3] = {nullptr, nullptr, nullptr};  // per-shot gradient (or image in acc[0])
    double *acc64[3] = {nullptr, nullptr, nullptr};
    int zc = 64;
    int arith = 1;  // kArithRef
    size_t bytes = 0;
    // two forward levels per launch: the second buffer pair it writes, and
    // whether their values outside the updated interior match (sx[1], sz[1])
    // and (sx[0], sz[0]) -- any other write to the fields clears it
    int levels = 1;
    float *sx2[2] = {nullptr, nullptr}, *sz2[2] = {nullptr, nullptr};
    bool t2_synced = false;
};

namespace {

long long vox(const Geo3 &g) { return g.plane * (long long)g.nz; }

// Largest z-chunk whose planes (plus the 4-plane halo on each side) stay
// within the 32-bit byte offsets of the kernels' buffer accesses.
int max_zchunk(const Geo3 &g)
{
    const long long planes = 0xFFFFFFFFLL / (4 * g.plane);
    return (int)std::min<long long>(1 << 20, (planes - 8) / 4 * 4);
}

// z-chunk of a launch: `requested` (64 by default) on a large grid; on a
// small grid a shorter chunk, so that a launch still has about 2048
// workgroups (8 per CU) to spread over the 256 CUs -- each chunk re-reads a
// 4-plane halo of sigma_z on either side, which the parallelism pays for
// below 64 planes -- but never below 8 planes.  `fixed` keeps `requested`.
// The result is always a multiple of 4 within [4, max_zchunk(g)], so a
// chunk's planes and halo stay within 32-bit byte offsets; 0 when even 4
// planes do not.
int choose_zchunk(const Geo3 &g, int requested, bool fixed)
{
    const int cap = max_zchunk(g);
    if (cap < 4) return 0;
    long long zc = std::max(4, requested / 4 * 4);
    if (!fixed) {
        const long long cols = (long long)((g.nx - 4 + 63) / 64) * ((g.ny - 8 + kBY - 1) / kBY);
        const long long fill = g.zmax * cols / 2048 / 4 * 4;
        zc = std::max<long long>(8, std::min<long long>(zc, fill));
    }
    return (int)std::min<long long>(zc, cap);
}

void freep(void *p, hipStream_t s) { vti::dfree(p, s); }

// stream-ordered allocation on the propagator's stream, returning on failure
#define DMALLOC(ptr, bytes)                                                      \
    do {                                                                         \
        const int rc_ = vti::dmalloc((void **)(ptr), (bytes), p->ctx->stream); \
        if (rc_) return rc_;                                                     \
    } while (0)

int prop_free(vti3d_prop *p)
{
    if (!p) return vti::kOk;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    hipStream_t s = p->ctx->stream;
    for (int k = 0; k < 3; ++k) {
        freep(p->cm[0][k], s);
        freep(p->cm[1][k], s);
        freep(p->acc[k], s);
        freep(p->acc64[k], s);
    }
    freep(p->dxy, s);
    freep(p->dzv, s);
    for (int k = 0; k < 2; ++k) {
        freep(p->sx[k], s);
        freep(p->sz[k], s);
        freep(p->lx[k], s);
        freep(p->lz[k], s);
        freep(p->sx2[k], s);
        freep(p->sz2[k], s);
    }
    (void)hipStreamSynchronize(s);
    delete p;
    return vti::kOk;
}

template <class T>
int dalloc(vti3d_prop *p, T **ptr, size_t n)
{
    if (*ptr) return vti::kOk;
    const int rc = vti::dmalloc((void **)ptr, n * sizeof(T), p->ctx->stream);
    if (rc) return rc;
    p->bytes += n * sizeof(T);
    VTI_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(T), p->ctx->stream));
    return vti::kOk;
}

// lambda fields and accumulators are only needed by the backward passes
int ensure_backward(vti3d_prop *p, bool gradient, bool random_model)
{
    const size_t nv = (size_t)vox(p->g);
    int rc = vti::kOk;
    for (int k = 0; k < 2 && rc == vti::kOk; ++k) {
        rc = dalloc(p, &p->lx[k], nv);
        if (rc == vti::kOk) rc = dalloc(p, &p->lz[k], nv);
    }
    for (int k = 0; k < (gradient ? 3 : 1) && rc == vti::kOk; ++k) {
        rc = dalloc(p, &p->acc[k], nv);
        if (rc == vti::kOk) rc = dalloc(p, &p->acc64[k], nv);
    }
    for (int k = 0; k < 3 && random_model && rc == vti::kOk; ++k) rc = dalloc(p, &p->cm[1][k], nv);
    return rc;
}

// host (reference layout, pitch nx) -> device (pitched)
bool device_ptr(const void *q)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, q) != hipSuccess) {
        (void)hipGetLastError();  // pageable host memory is not known to the runtime
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// host (reference layout, pitch nx) -> device (pitched).  src in host memory
// goes over as one contiguous copy into a stream-ordered staging block and is
// pitched on the device (a 2D copy from pageable memory moves row by row and
// took seconds for a 600 MB volume); a model kept resident by the caller in
// device memory is pitched directly.
int upload_volume(vti3d_prop *p, float *dst, const float *src)
{
    const Geo3 &g = p->g;
    const size_t row = (size_t)g.nx * 4, rows = (size_t)g.ny * g.nz;
    hipStream_t s = p->ctx->stream;
    if (device_ptr(src)) {
        VTI_HIP(hipMemcpy2DAsync(dst, (size_t)g.pitch * 4, src, row, row, rows, hipMemcpyDeviceToDevice, s));
        return vti::kOk;
    }
    float *stage = nullptr;
    DMALLOC(&stage, row * rows);
    hipError_t e = hipMemcpyAsync(stage, src, row * rows, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(dst, (size_t)g.pitch * 4, stage, row, row, rows, hipMemcpyDeviceToDevice, s);
    freep(stage, s);
    VTI_HIP(e);
    return vti::kOk;
}

// device (pitched) -> host (reference layout): unpitched on the device into a
// staging block, then one contiguous copy
int download_volume(vti3d_prop *p, float *dst, const float *src)
{
    const Geo3 &g = p->g;
    const size_t row = (size_t)g.nx * 4, rows = (size_t)g.ny * g.nz;
    hipStream_t s = p->ctx->stream;
    float *stage = nullptr;
    DMALLOC(&stage, row * rows);
    hipError_t e = hipMemcpy2DAsync(stage, row, src, (size_t)g.pitch * 4, row, rows, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, stage, row * rows, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // dst is complete when the call returns
    freep(stage, s);
    VTI_HIP(e);
    return vti::kOk;
}

long long dev_index(const Geo3 &g, long long ref_index)
{
    const long long nxy = (long long)g.nx * g.ny;
    const long long iz = ref_index / nxy, rem = ref_index % nxy;
    const long long iy = rem / g.nx, ix = rem % g.nx;
    return ix + iy * g.pitch + iz * g.plane;
}

struct Launch {
    int ntx, nty, ntz;
    dim3 grid, block;
};

Launch launch_shape(const vti3d_prop *p)
{
    Launch L;
    const Geo3 &g = p->g;
    L.ntx = (g.nx - 4 + 63) / 64;
    L.nty = (g.ny - 8 + kBY - 1) / kBY;
    L.ntz = (g.zmax + p->zc - 1) / p->zc;
    L.block = dim3(64, kBY);
    L.grid = dim3(L.ntx * L.nty * L.ntz);
    return L;
}

template <int ARITH, bool ABC>
void fwd_kernel(vti3d_prop *p, const Launch &L, long long src_dev, float src_val)
{
    hipLaunchKernelGGL((vti::fwd3d_kernel<ARITH, ABC, kBY>), L.grid, L.block, 0, p->ctx->stream, p->sx[1], p->sx[0],
                       p->sz[1], p->sz[0], p->cm[p->active][0], p->cm[p->active][1], p->cm[p->active][2], p->dxy,
                       p->dzv, p->g, p->zc, L.ntx, L.nty, p->dt2dx2, p->dt2dy2, p->dt2dz2, src_dev, src_val);
}

int fwd_launch(vti3d_prop *p, bool absorbing, long long src_dev, float src_val)
{
    p->t2_synced = false;
    const Launch L = launch_shape(p);
    if (p->arith == vti::kArithRef) {
        if (absorbing) fwd_kernel<vti::kArithRef, true>(p, L, src_dev, src_val);
        else fwd_kernel<vti::kArithRef, false>(p, L, src_dev, src_val);
    } else {
        if (absorbing) fwd_kernel<vti::kArithF32, true>(p, L, src_dev, src_val);
        else fwd_kernel<vti::kArithF32, false>(p, L, src_dev, src_val);
    }
    VTI_HIP(hipGetLastError());
    std::swap(p->sx[0], p->sx[1]);
    std::swap(p->sz[0], p->sz[1]);
    return vti::kOk;
}

// 64 x 16 tiles of the two-level kernel
Launch launch_shape_t2(const vti3d_prop *p)
{
    Launch L = launch_shape(p);
    L.nty = (p->g.ny - 8 + 15) / 16;
    L.block = dim3(64, 16);
    L.grid = dim3(L.ntx * L.nty * L.ntz);
    return L;
}

template <int ARITH, bool ABC>
void fwd_t2_kernel(vti3d_prop *p, const Launch &L, long long src_dev, float v1, float v2)
{
    hipLaunchKernelGGL((vti::fwd3d_t2b_kernel<ARITH, ABC>), L.grid, L.block, 0, p->ctx->stream, p->sx2[0],
                       p->sz2[0], p->sx2[1], p->sz2[1], p->sx[1], p->sz[1], p->sx[0], p->sz[0], p->cm[p->active][0],
                       p->cm[p->active][1], p->cm[p->active][2], p->dxy, p->dzv, p->g, p->zc, L.ntx, L.nty,
                       p->dt2dx2, p->dt2dy2, p->dt2dz2, src_dev, v1, v2);
}

// Two forward levels in one launch (sources v1 at the first, v2 at the
// second).  Allocates the second buffer pair on first use and, when the
// fields were written otherwise since, copies the current pair into it so
// that the values outside the updated interior agree.
int fwd2_launch(vti3d_prop *p, bool absorbing, long long src_dev, float v1, float v2)
{
    const size_t nv = (size_t)vox(p->g);
    for (int k = 0; k < 2; ++k) {
        int rc = dalloc(p, &p->sx2[k], nv);
        if (!rc) rc = dalloc(p, &p->sz2[k], nv);
        if (rc) return rc;
    }
    if (!p->t2_synced) {
        VTI_HIP(hipMemcpyAsync(p->sx2[0], p->sx[1], nv * 4, hipMemcpyDeviceToDevice, p->ctx->stream));
        VTI_HIP(hipMemcpyAsync(p->sz2[0], p->sz[1], nv * 4, hipMemcpyDeviceToDevice, p->ctx->stream));
        VTI_HIP(hipMemcpyAsync(p->sx2[1], p->sx[0], nv * 4, hipMemcpyDeviceToDevice, p->ctx->stream));
        VTI_HIP(hipMemcpyAsync(p->sz2[1], p->sz[0], nv * 4, hipMemcpyDeviceToDevice, p->ctx->stream));
    }
    const Launch L = launch_shape_t2(p);
    if (p->arith == vti::kArithRef) {
        if (absorbing) fwd_t2_kernel<vti::kArithRef, true>(p, L, src_dev, v1, v2);
        else fwd_t2_kernel<vti::kArithRef, false>(p, L, src_dev, v1, v2);
    } else {
        if (absorbing) fwd_t2_kernel<vti::kArithF32, true>(p, L, src_dev, v1, v2);
        else fwd_t2_kernel<vti::kArithF32, false>(p, L, src_dev, v1, v2);
    }
    VTI_HIP(hipGetLastError());
    // (prev, cur) <- (L+1, L+2); the old pair becomes the spare pair
    float *ax = p->sx[1], *az = p->sz[1], *bx = p->sx[0], *bz = p->sz[0];
    p->sx[1] = p->sx2[0];
    p->sz[1] = p->sz2[0];
    p->sx[0] = p->sx2[1];
    p->sz[0] = p->sz2[1];
    p->sx2[0] = ax;
    p->sz2[0] = az;
    p->sx2[1] = bx;
    p->sz2[1] = bz;
    p->t2_synced = true;
    return vti::kOk;
}

template <int MODE, int ARITH>
void bwd_kernel(vti3d_prop *p, const Launch &L, long long src_dev, float src_val, int zrec0)
{
    hipLaunchKernelGGL((vti::bwd3d_kernel<MODE, ARITH, kBY>), L.grid, L.block, 0, p->ctx->stream, p->sx[1], p->sx[0],
                       p->sz[1], p->sz[0], p->lx[1], p->lx[0], p->lz[1], p->lz[0], p->cm[p->active][0],
                       p->cm[p->active][1], p->cm[p->active][2], p->acc[0], MODE == 0 ? p->acc[1] : nullptr, MODE == 0 ? p->acc[2] : nullptr, p->dxy, p->dzv,
                       p->g, p->zc, L.ntx, L.nty, p->dx2, p->dy2, p->dz2, p->dt2, p->dt2dx2, p->dt2dy2, p->dt2dz2,
                       src_dev, src_val, zrec0);
}

// One backward level: the fused kernel, then (when the receiver block is
// present) the injection and damping of the receiver planes.
int bwd_launch(vti3d_prop *p, int mode, long long src_dev, float src_val, int zrec0, const long long *d_rec_eff,
               const float *coarse, int nr, int i1, float f)
{
    p->t2_synced = false;
    const Geo3 &g = p->g;
    const Launch L = launch_shape(p);
    if (p->arith == vti::kArithRef) {
        if (mode == 0) bwd_kernel<0, vti::kArithRef>(p, L, src_dev, src_val, zrec0);
        else bwd_kernel<1, vti::kArithRef>(p, L, src_dev, src_val, zrec0);
    } else {
        if (mode == 0) bwd_kernel<0, vti::kArithF32>(p, L, src_dev, src_val, zrec0);
        else bwd_kernel<1, vti::kArithF32>(p, L, src_dev, src_val, zrec0);
    }
    VTI_HIP(hipGetLastError());
    if (zrec0 >= 0) {
        if (nr > 0) {
            hipLaunchKernelGGL(vti::inject_res3d_kernel, dim3((nr + 255) / 256), dim3(256), 0, p->ctx->stream,
                               p->lx[1], p->lz[1], d_rec_eff, coarse, nr, i1, f, p->dt2);
            VTI_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(vti::damp_planes3d_kernel, dim3((g.nx + 255) / 256, g.ny, 4), dim3(256), 0,
                           p->ctx->stream, p->lx[1], p->lz[1], p->dxy, p->dzv, g, zrec0);
        VTI_HIP(hipGetLastError());
    }
    std::swap(p->sx[0], p->sx[1]);
    std::swap(p->sz[0], p->sz[1]);
    std::swap(p->lx[0], p->lx[1]);
    std::swap(p->lz[0], p->lz[1]);
    return vti::kOk;
}

int zero_fields(vti3d_prop *p, bool sigma, bool lambda)
{
    p->t2_synced = false;
    const size_t nb = (size_t)vox(p->g) * sizeof(float);
    for (int k = 0; k < 2; ++k) {
        if (sigma) {
            VTI_HIP(hipMemsetAsync(p->sx[k], 0, nb, p->ctx->stream));
            VTI_HIP(hipMemsetAsync(p->sz[k], 0, nb, p->ctx->stream));
        }
        if (lambda) {
            VTI_HIP(hipMemsetAsync(p->lx[k], 0, nb, p->ctx->stream));
            VTI_HIP(hipMemsetAsync(p->lz[k], 0, nb, p->ctx->stream));
        }
    }
    return vti::kOk;
}

// ---- shot geometry exactly as objFuncGradient3d_f.cu:545-576 computes it ---
struct Shot3 {
    int nr = 0, irbegin = 0;
    long long sou_ref = 0;
    int rec_block = 0;
    std::vector<long long> rec_true, rec_eff;  // reference-layout indices
};

int shot_setup(Shot3 &s, const float *souloc, int is, const float *recloc, int recstride, int nx, int ny, int nz,
               float ox, float oy, float oz, float dx, float dy, float dz)
{
    const long long nxy = (long long)nx * ny, nxyz = nxy * nz;
    s.nr = (int)souloc[5 * is + 3];
    s.irbegin = (int)souloc[5 * is + 4];
    VTI_CHECK(s.nr >= 0 && s.irbegin >= 0, "shot %d: bad receiver count/offset", is);
    s.rec_block = (int)((recloc[2] - oz) / dz / 4);
    s.rec_true.resize(s.nr);
    s.rec_eff.resize(s.nr);
    for (int ir = 0; ir < s.nr; ir++) {
        const int ir1 = ir + s.irbegin;
        const int ix = (int)((recloc[recstride * ir1] - ox) / dx);
        const int iy = (int)((recloc[recstride * ir1 + 1] - oy) / dy);
        const int iz = (int)((recloc[recstride * ir1 + 2] - oz) / dz);
        VTI_CHECK(ix >= 0 && ix < nx && iy >= 0 && iy < ny && iz >= 0 && iz < nz, "receiver %d outside the grid", ir1);
        const long long ixy = ix + (long long)iy * nx;
        s.rec_true[ir] = ixy + (long long)iz * nxy;
        s.rec_eff[ir] = ixy + (long long)(iz % 4) * nxy + 4LL * s.rec_block * nxy;
        VTI_CHECK(s.rec_eff[ir] < nxyz, "receiver %d: block index outside the grid", ir1);
    }
    const int sx = (int)((souloc[5 * is] - ox) / dx);
    const int sy = (int)((souloc[5 * is + 1] - oy) / dy);
    const int sz = (int)((souloc[5 * is + 2] - oz) / dz);
    VTI_CHECK(sx >= 0 && sx < nx && sy >= 0 && sy < ny && sz >= 0 && sz < nz, "shot %d: source outside the grid", is);
    s.sou_ref = sx + (long long)sy * nx + (long long)sz * nxy;
    return vti::kOk;
}

// number of levels the reference's pipeline advances (objFuncGradient3d_f.cu:539)
int levels(int nt, int ngpu)
{
    const int nround = (nt - 2) / (ngpu * 64);
    return nround * ngpu * 64 + 1;
}

// Device scratch for one shot's traces and indices.
struct ShotDev {
    long long *rec = nullptr, *rec_true = nullptr, *rec_eff = nullptr;
    float *tr = nullptr;       // time-major [nnt][nr]
    float *obs = nullptr;      // trace-major [nr][nnt]
    unsigned char *live = nullptr;
    double *partial = nullptr;
    hipStream_t s;  // the stream that uses (and releases) them
    explicit ShotDev(hipStream_t st) : s(st) {}
    ~ShotDev()
    {
        freep(rec, s);
        freep(rec_true, s);
        freep(rec_eff, s);
        freep(tr, s);
        freep(obs, s);
        freep(live, s);
        freep(partial, s);
    }
};

int to_dev_idx(vti3d_prop *p, long long **dst, const std::vector<long long> &ref)
{
    std::vector<long long> d(ref.size());
    for (size_t i = 0; i < ref.size(); ++i) d[i] = dev_index(p->g, ref[i]);
    int rc = vti::dmalloc((void **)dst, sizeof(long long) * std::max<size_t>(1, d.size()), p->ctx->stream);
    if (rc) return rc;
    if (!d.empty())
        VTI_HIP(hipMemcpyAsync(*dst, d.data(), sizeof(long long) * d.size(), hipMemcpyHostToDevice, p->ctx->stream));
    return vti::kOk;
}

// Forward pass of one shot (levels 2..lmax) with optional recording into
// sd.tr rows it/sts for it % sts == 0, it/sts < nnt.
int forward_pass(vti3d_prop *p, bool absorbing, int lmax, const float *wavelet, long long sou_ref, const ShotDev *sd,
                 int nr, int sts, int nnt)
{
    const long long src = dev_index(p->g, sou_ref);
    // record level it, held in (sx[k], sz[k])
    auto record = [&](int it, int k) -> int {
        if (sd && sd->tr && nr > 0 && it % sts == 0 && it / sts < nnt) {
            hipLaunchKernelGGL(record3d_kernel, dim3((nr + 255) / 256), dim3(256), 0, p->ctx->stream,
                               sd->tr + (size_t)(it / sts) * nr, p->sx[k], p->sz[k], sd->rec_eff, nr);
            VTI_HIP(hipGetLastError());
        }
        return vti::kOk;
    };
    for (int it = 2; it <= lmax;) {
        int rc;
        if (p->levels == 2 && it + 1 <= lmax) {
            rc = fwd2_launch(p, absorbing, src, p->dt2 * wavelet[it - 1], p->dt2 * wavelet[it]);
            if (!rc) rc = record(it, 1);
            if (!rc) rc = record(it + 1, 0);
            it += 2;
        } else {
            rc = fwd_launch(p, absorbing, src, p->dt2 * wavelet[it - 1]);
            if (!rc) rc = record(it, 0);
            it += 1;
        }
        if (rc) return rc;
    }
    return vti::kOk;
}

// Backward pass (gradient mode 0, RTM mode 1) after the random-boundary
// forward and the flip: s = 2..lmax, it = nt-1-s, source dt2*wavelet[it+1],
// receiver-side injection of the traces in sd.tr interpolated at it+1.
int backward_pass(vti3d_prop *p, int mode, int lmax, int nt, const float *wavelet, long long sou_ref,
                  const ShotDev &sd, int nr, int rec_block, int sts, int nnt)
{
    const long long src = dev_index(p->g, sou_ref);
    const int zrec0 = 4 * rec_block;
    const bool recv_ok = zrec0 + 3 < p->g.zmax;
    for (int s = 2; s <= lmax; ++s) {
        const int it = nt - 1 - s;
        const int ti = it + 1;
        float f = (float)ti / (float)sts;
        int i1 = (int)f;
        const float *coarse = sd.tr;
        if (i1 >= nnt - 1) {
            coarse = sd.tr + (size_t)(nnt - 1) * nr;
            i1 = -1;
            f = 0.f;
        } else {
            f = f - i1;
        }
        int rc = bwd_launch(p, mode, src, p->dt2 * wavelet[it + 1], recv_ok ? zrec0 : -8, sd.rec_eff, coarse, nr, i1,
                            f);
        if (rc) return rc;
    }
    return vti::kOk;
}

int launch_blocks(long long n) { return (int)std::min<long long>(4096, std::max<long long>(1, (n + 255) / 256)); }

}  // namespace

extern "C" {

int vti3d_prop_create(vti_ctx *ctx, vti3d_prop **out, int nx, int ny, int nz, int npad, float dx, float dy,
                      float dz, float dt)
{
    VTI_CHECK(ctx && out, "vti3d_prop_create: null argument");
    VTI_CHECK(nx >= 9 && ny >= 9 && nz >= 8, "vti3d_prop_create: grid %dx%dx%d too small", nx, ny, nz);
    VTI_CHECK(npad >= 1 && 2 * npad < nx && 2 * npad < ny && 2 * npad < nz, "vti3d_prop_create: bad npad %d", npad);
    VTI_CHECK(dx > 0 && dy > 0 && dz > 0 && dt > 0, "vti3d_prop_create: spacings must be positive");
    // the kernels index within one x-y plane with 32-bit offsets
    VTI_CHECK((long long)((nx + 63) / 64 * 64) * ny < (1LL << 31), "vti3d_prop_create: x-y plane %dx%d too large", nx,
              ny);
    VTI_HIP(hipSetDevice(ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    vti3d_prop *p = new vti3d_prop();
    p->ctx = ctx;
    p->arith = ctx->arith;
    Geo3 &g = p->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.zmax = 4 * (nz / 4);
    g.pitch = (nx + 63) / 64 * 64;
    g.plane = (long long)g.pitch * ny;
    p->npad = npad;
    p->dx = dx;
    p->dy = dy;
    p->dz = dz;
    p->dt = dt;
    p->dx2 = dx * dx;
    p->dy2 = dy * dy;
    p->dz2 = dz * dz;
    p->dt2 = dt * dt;
    p->dt2dx2 = p->dt2 / p->dx2;
    p->dt2dy2 = p->dt2 / p->dy2;
    p->dt2dz2 = p->dt2 / p->dz2;
    const char *zenv = getenv("VTI_ZCHUNK");
    p->zc = choose_zchunk(g, zenv ? atoi(zenv) : p->zc, zenv != nullptr);
    if (p->zc == 0) {
        delete p;
        VTI_CHECK(false, "vti3d_prop_create: x-y plane %dx%d too large for 32-bit chunk offsets", nx, ny);
    }
    const size_t nv = (size_t)vox(g);
    int rc = vti::kOk;
    for (int k = 0; k < 3 && rc == vti::kOk; ++k) rc = dalloc(p, &p->cm[0][k], nv);
    for (int k = 0; k < 2 && rc == vti::kOk; ++k) {
        rc = dalloc(p, &p->sx[k], nv);
        if (rc == vti::kOk) rc = dalloc(p, &p->sz[k], nv);
    }
    if (rc == vti::kOk) rc = dalloc(p, &p->dxy, (size_t)g.plane);
    if (rc == vti::kOk) rc = dalloc(p, &p->dzv, (size_t)nz);
    if (rc != vti::kOk) {
        prop_free(p);
        return rc;
    }
    // damping profile of abc3d.cu:23-50, computed on the host as the reference does
    std::vector<float> damp((size_t)nx * ny + nz);
    vti::init_damping3d(damp.data(), nx, ny, nz, npad);
    VTI_HIP(hipMemcpy2DAsync(p->dxy, (size_t)g.pitch * 4, damp.data(), (size_t)nx * 4, (size_t)nx * 4, ny,
                             hipMemcpyHostToDevice, ctx->stream));
    VTI_HIP(hipMemcpyAsync(p->dzv, damp.data() + (size_t)nx * ny, (size_t)nz * 4, hipMemcpyHostToDevice,
                           ctx->stream));
    VTI_HIP(hipStreamSynchronize(ctx->stream));
    *out = p;
    return vti::kOk;
}

int vti3d_prop_destroy(vti3d_prop *p) { return prop_free(p); }

size_t vti3d_prop_bytes(const vti3d_prop *p) { return p ? p->bytes : 0; }

int vti3d_prop_set_arith(vti3d_prop *p, int arith)
{
    VTI_CHECK(p && (arith == 0 || arith == 1), "vti3d_prop_set_arith: arith must be 0 (fp32) or 1 (reference)");
    p->arith = arith;
    return vti::kOk;
}

// The z-chunk vti3d_prop_create picks for a grid (host only; for tests).
int vti3d_zchunk_for(int nx, int ny, int nz, int requested, int fixed)
{
    Geo3 g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.zmax = 4 * (nz / 4);
    g.pitch = (nx + 63) / 64 * 64;
    g.plane = (long long)g.pitch * ny;
    return choose_zchunk(g, requested, fixed != 0);
}

int vti3d_prop_set_zchunk(vti3d_prop *p, int zc)
{
    VTI_CHECK(p && zc >= 4 && zc % 4 == 0, "vti3d_prop_set_zchunk: chunk must be a positive multiple of 4");
    VTI_CHECK(zc <= max_zchunk(p->g), "vti3d_prop_set_zchunk: chunk %d above %d for this plane size", zc,
              max_zchunk(p->g));
    p->zc = zc;
    return vti::kOk;
}

// Forward levels per launch of vti3d_prop_forward_steps / forward_shot: 1
// or 2 (the same bits either way).
int vti3d_prop_set_levels(vti3d_prop *p, int levels)
{
    VTI_CHECK(p && (levels == 1 || levels == 2), "vti3d_prop_set_levels: levels must be 1 or 2");
    p->levels = levels;
    return vti::kOk;
}

// Upload c11, c13, c33 (host arrays in the reference layout, or a model
// the caller keeps in device memory).
int vti3d_prop_set_model(vti3d_prop *p, const float *c11, const float *c13, const float *c33)
{
    VTI_CHECK(p && c11 && c13 && c33, "vti3d_prop_set_model: null argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    const float *src[3] = {c11, c13, c33};
    for (int k = 0; k < 3; ++k) {
        int rc = upload_volume(p, p->cm[0][k], src[k]);
        if (rc) return rc;
    }
    p->active = 0;
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

// Zero both time levels and put the reference's level-1 source:
// cur[sou] = value (objFuncGradient3d_f.cu:594-599).  sou_index < 0: no source.
int vti3d_prop_reset(vti3d_prop *p, long long sou_index, float value)
{
    VTI_CHECK(p, "null propagator");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    int rc = zero_fields(p, true, false);
    if (rc) return rc;
    if (sou_index >= 0) {
        VTI_CHECK(sou_index < (long long)p->g.nx * p->g.ny * p->g.nz, "source index out of range");
        hipLaunchKernelGGL(set_point_kernel, dim3(1), dim3(1), 0, p->ctx->stream, p->sx[0], p->sz[0],
                           dev_index(p->g, sou_index), value);
        VTI_HIP(hipGetLastError());
    }
    return vti::kOk;
}

// Load an arbitrary state (reference layout); NULL pointers leave a field as is.
int vti3d_prop_set_state(vti3d_prop *p, const float *cur_x, const float *cur_z, const float *prev_x,
                         const float *prev_z)
{
    VTI_CHECK(p, "null propagator");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    p->t2_synced = false;
    const float *src[4] = {cur_x, cur_z, prev_x, prev_z};
    float *dst[4] = {p->sx[0], p->sz[0], p->sx[1], p->sz[1]};
    for (int k = 0; k < 4; ++k)
        if (src[k]) {
            int rc = upload_volume(p, dst[k], src[k]);
            if (rc) return rc;
        }
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

int vti3d_prop_get_state(vti3d_prop *p, float *cur_x, float *cur_z, float *prev_x, float *prev_z)
{
    VTI_CHECK(p, "null propagator");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    float *dst[4] = {cur_x, cur_z, prev_x, prev_z};
    const float *src[4] = {p->sx[0], p->sz[0], p->sx[1], p->sz[1]};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) {
            int rc = download_volume(p, dst[k], src[k]);
            if (rc) return rc;
        }
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

// Load the backward passes' adjoint state (reference layout; NULL leaves a
// field as is) and clear the FP32 accumulators the fused backward kernels
// add into (gc11/gc13/gc33 for mode 0, the image for mode 1).
int vti3d_prop_set_adjoint(vti3d_prop *p, int mode, const float *cur_x, const float *cur_z, const float *prev_x,
                           const float *prev_z)
{
    VTI_CHECK(p && (mode == 0 || mode == 1), "vti3d_prop_set_adjoint: bad argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    int rc = ensure_backward(p, mode == 0, false);
    if (rc) return rc;
    const float *src[4] = {cur_x, cur_z, prev_x, prev_z};
    float *dst[4] = {p->lx[0], p->lz[0], p->lx[1], p->lz[1]};
    for (int k = 0; k < 4; ++k)
        if (src[k]) {
            rc = upload_volume(p, dst[k], src[k]);
            if (rc) return rc;
        }
    for (int k = 0; k < (mode == 0 ? 3 : 1); ++k)
        VTI_HIP(hipMemsetAsync(p->acc[k], 0, sizeof(float) * (size_t)vox(p->g), p->ctx->stream));
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

// Read the adjoint state and the FP32 accumulators back (NULL skips a field;
// acc1/acc2 exist after a gradient-mode set_adjoint only).
int vti3d_prop_get_adjoint(vti3d_prop *p, float *cur_x, float *cur_z, float *prev_x, float *prev_z, float *acc0,
                           float *acc1, float *acc2)
{
    VTI_CHECK(p && p->lx[0], "vti3d_prop_get_adjoint: no adjoint state (call vti3d_prop_set_adjoint first)");
    VTI_CHECK(!(acc1 || acc2) || p->acc[2], "vti3d_prop_get_adjoint: no gradient accumulators on this propagator");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();
    float *dst[7] = {cur_x, cur_z, prev_x, prev_z, acc0, acc1, acc2};
    const float *src[7] = {p->lx[0], p->lz[0], p->lx[1], p->lz[1], p->acc[0], p->acc[1], p->acc[2]};
    for (int k = 0; k < 7; ++k)
        if (dst[k]) {
            int rc = download_volume(p, dst[k], src[k]);
            if (rc) return rc;
        }
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

// Advance `nsteps` forward levels with no source.  When elapsed_ms is not
// NULL the launches are bracketed by HIP events on the propagator's stream
// and the average time per step is returned there.
int vti3d_prop_forward_steps(vti3d_prop *p, int nsteps, int absorbing, float *elapsed_ms)
{
    VTI_CHECK(p && nsteps >= 0, "vti3d_prop_forward_steps: bad argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        VTI_HIP(hipEventCreate(&e0));
        VTI_HIP(hipEventCreate(&e1));
        VTI_HIP(hipEventRecord(e0, p->ctx->stream));
    }
    for (int s = 0; s < nsteps;) {
        int rc;
        if (p->levels == 2 && s + 1 < nsteps) {
            rc = fwd2_launch(p, absorbing != 0, -1, 0.f, 0.f);
            s += 2;
        } else {
            rc = fwd_launch(p, absorbing != 0, -1, 0.f);
            s += 1;
        }
        if (rc) return rc;
    }
    if (elapsed_ms) {
        VTI_HIP(hipEventRecord(e1, p->ctx->stream));
        VTI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        VTI_HIP(hipEventElapsedTime(&ms, e0, e1));
        *elapsed_ms = nsteps > 0 ? ms / nsteps : 0.f;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    return vti::kOk;
}

// Advance `nsteps` fused backward levels (mode 0 gradient, 1 RTM) on the
// current state with no source and no receivers, timed like
// vti3d_prop_forward_steps.  Allocates the backward fields on first use.
int vti3d_prop_backward_steps(vti3d_prop *p, int mode, int nsteps, float *elapsed_ms)
{
    VTI_CHECK(p && nsteps >= 0 && (mode == 0 || mode == 1), "vti3d_prop_backward_steps: bad argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    int rc = ensure_backward(p, mode == 0, false);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        VTI_HIP(hipEventCreate(&e0));
        VTI_HIP(hipEventCreate(&e1));
        VTI_HIP(hipEventRecord(e0, p->ctx->stream));
    }
    for (int s = 0; s < nsteps; ++s) {
        rc = bwd_launch(p, mode, -1, 0.f, -8, nullptr, nullptr, 0, 0, 0.f);
        if (rc) return rc;
    }
    if (elapsed_ms) {
        VTI_HIP(hipEventRecord(e1, p->ctx->stream));
        VTI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        VTI_HIP(hipEventElapsedTime(&ms, e0, e1));
        *elapsed_ms = nsteps > 0 ? ms / nsteps : 0.f;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    return vti::kOk;
}

// Forward modeling of one source with the reference's level/recording rules
// (objFuncGradient3d_f.cu:594-735): levels 2..lmax, source dt2*wavelet[it-1]
// at sou_index, absorbing damping when `absorbing`, traces recorded when
// it % sts == 0 and it/sts < nnt into traces_tm (time-major, nnt rows of nr;
// rows never recorded keep their input values).  rec: reference-layout indices.
int vti3d_prop_forward_shot(vti3d_prop *p, int lmax, const float *wavelet, long long sou_index, int nr,
                            const long long *rec, int sts, int nnt, int absorbing, float *traces_tm)
{
    VTI_CHECK(p && wavelet, "vti3d_prop_forward_shot: null argument");
    VTI_CHECK(sts >= 1 && nnt >= 1 && nr >= 0, "vti3d_prop_forward_shot: bad sampling");
    VTI_HIP(hipSetDevice(p->ctx->device));
    (void)hipGetLastError();  // drop a stale error left by an unrelated earlier call
    const Geo3 &g = p->g;
    VTI_CHECK(sou_index >= 0 && sou_index < (long long)g.nx * g.ny * g.nz, "source index out of range");
    ShotDev sd(p->ctx->stream);
    if (nr > 0 && traces_tm) {
        std::vector<long long> r(rec, rec + nr);
        for (int i = 0; i < nr; ++i)
            VTI_CHECK(r[i] >= 0 && r[i] < (long long)g.nx * g.ny * g.nz, "receiver %d index out of range", i);
        int rc = to_dev_idx(p, &sd.rec_eff, r);
        if (rc) return rc;
        DMALLOC(&sd.tr, sizeof(float) * (size_t)nr * nnt);
        VTI_HIP(hipMemcpyAsync(sd.tr, traces_tm, sizeof(float) * (size_t)nr * nnt, hipMemcpyHostToDevice,
                               p->ctx->stream));
    }
    int rc = forward_pass(p, absorbing != 0, lmax, wavelet, sou_index, sd.tr ? &sd : nullptr, nr, sts, nnt);
    if (rc) return rc;
    if (sd.tr)
        VTI_HIP(hipMemcpyAsync(traces_tm, sd.tr, sizeof(float) * (size_t)nr * nnt, hipMemcpyDeviceToHost,
                               p->ctx->stream));
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-signature drivers
// ---------------------------------------------------------------------------
namespace {

struct DriverArgs {
    const float *souloc;
    const int *shotid;
    int nshot;
    const float *recloc;
    const float *wavelet;
    int nx, ny, nz, nt, npad;
    float ox, oy, oz, ot, dx, dy, dz, dt, rate;
    int ngpu;
};

int check_common(const DriverArgs &a)
{
    VTI_CHECK(a.souloc && a.recloc && a.wavelet && (a.nshot == 0 || a.shotid), "null geometry/wavelet argument");
    VTI_CHECK(a.nshot >= 0 && a.ngpu >= 1, "bad shot or GPU count");
    VTI_CHECK(a.nz / 4 >= 2, "nz must hold at least two 4-plane blocks");
    const int sts = (int)std::round(a.rate / a.dt);
    VTI_CHECK(sts >= 3, "samplingRate/dt = %d: the reference pipeline records correctly only for >= 3", sts);
    return vti::kOk;
}

// Per-shot forward with absorbing boundary and recording, then mute, gain and
// (when obs != nullptr) the residual in sd.tr; returns sum(r*r) in *ss.
int modeled_shot(vti3d_prop *p, const DriverArgs &a, const Shot3 &sh, ShotDev &sd, int lmax, int sts, int nnt,
                 const float *obs, float tpow, int objtype, int halfwidth, double *ss,
                 std::vector<float> *traces_out)
{
    int rc = zero_fields(p, true, false);
    if (rc) return rc;
    hipLaunchKernelGGL(set_point_kernel, dim3(1), dim3(1), 0, p->ctx->stream, p->sx[0], p->sz[0],
                       dev_index(p->g, sh.sou_ref), p->dt2 * a.wavelet[0]);
    VTI_HIP(hipGetLastError());
    rc = to_dev_idx(p, &sd.rec_eff, sh.rec_eff);
    if (rc) return rc;
    rc = to_dev_idx(p, &sd.rec_true, sh.rec_true);
    if (rc) return rc;
    const size_t nn = (size_t)std::max(1, sh.nr) * nnt;
    DMALLOC(&sd.tr, sizeof(float) * nn);
    VTI_HIP(hipMemsetAsync(sd.tr, 0, sizeof(float) * nn, p->ctx->stream));
    rc = forward_pass(p, true, lmax, a.wavelet, sh.sou_ref, &sd, sh.nr, sts, nnt);
    if (rc) return rc;
    std::vector<unsigned char> live(std::max(1, sh.nr));
    for (int ir = 0; ir < sh.nr; ++ir) live[ir] = a.recloc[4 * (ir + sh.irbegin) + 3] != 0.f;
    DMALLOC(&sd.live, live.size());
    VTI_HIP(hipMemcpyAsync(sd.live, live.data(), live.size(), hipMemcpyHostToDevice, p->ctx->stream));
    const int nblk = launch_blocks((long long)sh.nr * nnt);
    DMALLOC(&sd.partial, sizeof(double) * nblk);
    if (obs) {
        DMALLOC(&sd.obs, sizeof(float) * nn);
        VTI_HIP(hipMemcpyAsync(sd.obs, obs + (size_t)nnt * sh.irbegin, sizeof(float) * (size_t)sh.nr * nnt,
                               hipMemcpyHostToDevice, p->ctx->stream));
    } else {
        DMALLOC(&sd.obs, sizeof(float) * nn);
        VTI_HIP(hipMemsetAsync(sd.obs, 0, sizeof(float) * nn, p->ctx->stream));
    }
    float *gain = nullptr;
    std::vector<float> g;
    if (tpow != 0.f) {
        g.resize(nnt);
        for (int it = 0; it < nnt; ++it) {
            const float t = a.ot + it * a.rate;
            g[it] = std::pow(t, tpow);
        }
        DMALLOC(&gain, sizeof(float) * nnt);
        VTI_HIP(hipMemcpyAsync(gain, g.data(), sizeof(float) * nnt, hipMemcpyHostToDevice, p->ctx->stream));
    }
    hipLaunchKernelGGL(residual3d_kernel, dim3(nblk), dim3(256), 0, p->ctx->stream, sd.tr, sd.obs, sd.live, sh.nr,
                       nnt, gain, objtype == 0 ? 1 : 0, sd.partial);
    VTI_HIP(hipGetLastError());
    std::vector<double> part(objtype == 0 ? nblk : std::max(1, sh.nr));
    if (objtype == 0) {
        VTI_HIP(hipMemcpyAsync(part.data(), sd.partial, sizeof(double) * nblk, hipMemcpyDeviceToHost, p->ctx->stream));
    } else if (sh.nr > 0) {
        // the AGC objective (objFuncGradient3d_f.cu:770-773): traces become the adjoint source
        float *rsb = nullptr, *drb = nullptr;
        double *fi = nullptr;
        DMALLOC(&rsb, sizeof(float) * nn);
        DMALLOC(&drb, sizeof(float) * nn);
        DMALLOC(&fi, sizeof(double) * sh.nr);
        hipLaunchKernelGGL(agc3d_kernel, dim3((sh.nr + 63) / 64), dim3(64), 0, p->ctx->stream, sd.tr, sd.obs, sh.nr,
                           nnt, halfwidth, rsb, drb, fi);
        VTI_HIP(hipGetLastError());
        VTI_HIP(hipMemcpyAsync(part.data(), fi, sizeof(double) * sh.nr, hipMemcpyDeviceToHost, p->ctx->stream));
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
        freep(rsb, p->ctx->stream);
        freep(drb, p->ctx->stream);
        freep(fi, p->ctx->stream);
    }
    if (traces_out) {
        traces_out->resize((size_t)sh.nr * nnt);
        VTI_HIP(hipMemcpyAsync(traces_out->data(), sd.tr, sizeof(float) * (size_t)sh.nr * nnt, hipMemcpyDeviceToHost,
                               p->ctx->stream));
    }
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    freep(gain, p->ctx->stream);
    double s = 0;
    for (double v : part) s += v;
    *ss = s;
    return vti::kOk;
}

int put_shell_and_upload(vti3d_prop *p, float *const dst[3], const float *cij, const float *packed, int nx, int ny,
                         int nz, int npad)
{
    const long long nxyz = (long long)nx * ny * nz, nb = vti::boundary_size(nx, ny, nz, npad);
    std::vector<float> tmp((size_t)nxyz);
    for (int k = 0; k < 3; ++k) {
        std::memcpy(tmp.data(), cij + k * nxyz, sizeof(float) * nxyz);
        if (packed) vti::put_boundary(packed + k * nb, tmp.data(), nx, ny, nz, npad);
        int rc = upload_volume(p, dst[k], tmp.data());
        if (rc) return rc;
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    return vti::kOk;
}

int accumulate(vti3d_prop *p, int n)
{
    const long long nv = vox(p->g);
    for (int k = 0; k < n; ++k) {
        hipLaunchKernelGGL(accumulate64_kernel, dim3(launch_blocks(nv)), dim3(256), 0, p->ctx->stream, p->acc64[k],
                           p->acc[k], nv);
        VTI_HIP(hipGetLastError());
        VTI_HIP(hipMemsetAsync(p->acc[k], 0, sizeof(float) * nv, p->ctx->stream));
    }
    return vti::kOk;
}

int narrow_download(vti3d_prop *p, float *dst, int k)
{
    const long long nv = vox(p->g);
    hipLaunchKernelGGL(narrow64_kernel, dim3(launch_blocks(nv)), dim3(256), 0, p->ctx->stream, p->acc[k], p->acc64[k],
                       nv);
    VTI_HIP(hipGetLastError());
    return download_volume(p, dst, p->acc[k]);
}

struct PropGuard {
    vti3d_prop *p = nullptr;
    ~PropGuard() { prop_free(p); }
};

}  // namespace

extern "C" {

// modelDataCij3d_f (lib/vti-fwi-3d/gpu/modelData3d_f.cu:427-808): absorbing
// forward modeling of every shot in shotid; ddata (trace-major, nnt samples
// per receiver) receives each shot's traces at offset nnt*irbegin, dead
// traces (recloc[4*ir+3] == 0) zeroed.  ngpu only sets the number of time
// levels the reference's pipeline would advance for that many GPUs.
int vti3d_model_data(vti_ctx *ctx, float *ddata, const float *souloc, int ns, const int *shotid, int nshot,
                     const float *recloc, const float *wavelet, const float *c11, const float *c13, const float *c33,
                     int nx, int ny, int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx,
                     float dy, float dz, float dt, float samplingRate, int ngpu)
{
    DriverArgs a{souloc, shotid, nshot, recloc, wavelet, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt,
                 samplingRate, ngpu};
    int rc = check_common(a);
    if (rc) return rc;
    VTI_CHECK(ctx && ddata && c11 && c13 && c33, "vti3d_model_data: null argument");
    for (int i = 0; i < nshot; ++i) VTI_CHECK(shotid[i] >= 0 && shotid[i] < ns, "shot id %d out of range", shotid[i]);
    PropGuard G;
    rc = vti3d_prop_create(ctx, &G.p, nx, ny, nz, npad, dx, dy, dz, dt);
    if (rc) return rc;
    rc = vti3d_prop_set_model(G.p, c11, c13, c33);
    if (rc) return rc;
    const int sts = (int)std::round(samplingRate / dt), nnt = (nt - 1) / sts + 1, lmax = levels(nt, ngpu);
    for (int k = 0; k < nshot; ++k) {
        Shot3 sh;
        rc = shot_setup(sh, souloc, shotid[k], recloc, 4, nx, ny, nz, ox, oy, oz, dx, dy, dz);
        if (rc) return rc;
        ShotDev sd(ctx->stream);
        double ss = 0;
        rc = modeled_shot(G.p, a, sh, sd, lmax, sts, nnt, nullptr, 0.f, 0, 0, &ss, nullptr);
        if (rc) return rc;
        if (sh.nr > 0) {
            // time-major rows -> the trace-major block of this shot, on the device
            float *tm = nullptr;
            const size_t nn = (size_t)sh.nr * nnt;
            rc = vti::dmalloc((void **)&tm, sizeof(float) * nn, ctx->stream);
            if (rc) return rc;
            hipLaunchKernelGGL(transpose_kernel, transpose_tiles(nnt, sh.nr), dim3(32, 8), 0, ctx->stream, tm, sd.tr,
                               nnt, sh.nr);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(ddata + (size_t)nnt * sh.irbegin, tm, sizeof(float) * nn, hipMemcpyDeviceToHost,
                                   ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            vti::dfree(tm, ctx->stream);
            VTI_HIP(e);
        }
    }
    return vti::kOk;
}

// objFuncCij3d_f (objFuncGradient3d_f.cu:1272-1689) when grad == NULL and
// objFuncGradientCij3d_f (objFuncGradient3d_f.cu:361-1231) otherwise.
// fgcij: 1 + 3*nx*ny*nz floats, f at [0] = 0.5*sum r^2, then gc11, gc13, gc33.
// obs replaces the reference's read("data") (:760): trace-major, nnt samples
// per receiver over all receivers.  tpow is the reference's tpow= parameter
// (0 disables the t^p gain).  Like the reference, the call leaves cij's
// boundary shell holding the random boundary (gradient) or the absorbing
// boundary (objective only).
static int objfunc_impl(vti_ctx *ctx, float *fgcij, bool grad, const float *souloc, int ns, const int *shotid,
                        int nshot, const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij,
                        const float *randboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                        float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate, int ngpu,
                        const float *obs, float tpow, int objtype, int halfwidth)
{
    DriverArgs a{souloc, shotid, nshot, recloc, wavelet, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt,
                 samplingRate, ngpu};
    int rc = check_common(a);
    if (rc) return rc;
    VTI_CHECK(ctx && fgcij && cij && padboundaryCij && obs && (!grad || randboundaryCij),
              "objective/gradient: null argument");
    VTI_CHECK(objtype == 0 || objtype == 1, "objtype must be 0 (least squares) or 1 (AGC), got %d", objtype);
    for (int i = 0; i < nshot; ++i) VTI_CHECK(shotid[i] >= 0 && shotid[i] < ns, "shot id %d out of range", shotid[i]);
    PropGuard G;
    rc = vti3d_prop_create(ctx, &G.p, nx, ny, nz, npad, dx, dy, dz, dt);
    if (rc) return rc;
    vti3d_prop *p = G.p;
    if (grad) {
        rc = ensure_backward(p, true, true);
        if (rc) return rc;
        rc = put_shell_and_upload(p, p->cm[1], cij, randboundaryCij, nx, ny, nz, npad);
        if (rc) return rc;
    }
    rc = put_shell_and_upload(p, p->cm[0], cij, padboundaryCij, nx, ny, nz, npad);
    if (rc) return rc;
    const int sts = (int)std::round(samplingRate / dt), nnt = (nt - 1) / sts + 1, lmax = levels(nt, ngpu);
    VTI_CHECK(objtype == 0 || (halfwidth >= 0 && 2 * halfwidth < nnt),
              "AGC half width %d needs 0 <= 2*halfwidth < %d samples", halfwidth, nnt);
    double objfunc = 0.0;
    for (int k = 0; k < nshot; ++k) {
        Shot3 sh;
        rc = shot_setup(sh, souloc, shotid[k], recloc, 4, nx, ny, nz, ox, oy, oz, dx, dy, dz);
        if (rc) return rc;
        ShotDev sd(p->ctx->stream);
        double ss = 0;
        p->active = 0;
        rc = modeled_shot(p, a, sh, sd, lmax, sts, nnt, obs, tpow, objtype, halfwidth, &ss, nullptr);
        if (rc) return rc;
        objfunc += ss;
        if (!grad) continue;
        // random-boundary forward, no damping (:786-905), then the flip (:911-914)
        p->active = 1;
        rc = zero_fields(p, true, true);
        if (rc) return rc;
        hipLaunchKernelGGL(set_point_kernel, dim3(1), dim3(1), 0, p->ctx->stream, p->sx[0], p->sz[0],
                           dev_index(p->g, sh.sou_ref), p->dt2 * wavelet[0]);
        VTI_HIP(hipGetLastError());
        rc = forward_pass(p, false, lmax, wavelet, sh.sou_ref, nullptr, 0, sts, nnt);
        if (rc) return rc;
        std::swap(p->sx[0], p->sx[1]);
        std::swap(p->sz[0], p->sz[1]);
        if (sh.nr > 0)
            hipLaunchKernelGGL(lambda_init3d_kernel, dim3((sh.nr + 255) / 256), dim3(256), 0, p->ctx->stream,
                               p->lx[0], p->lz[0], sd.rec_true, sd.tr + (size_t)(nnt - 1) * sh.nr, sh.nr, p->dt2);
        VTI_HIP(hipGetLastError());
        rc = backward_pass(p, 0, lmax, nt, wavelet, sh.sou_ref, sd, sh.nr, sh.rec_block, sts, nnt);
        if (rc) return rc;
        rc = accumulate(p, 3);
        if (rc) return rc;
    }
    const long long nxyz = (long long)nx * ny * nz;
    fgcij[0] = (float)(objfunc / 2);
    if (grad) {
        for (int k = 0; k < 3; ++k) {
            rc = narrow_download(p, fgcij + 1 + k * nxyz, k);
            if (rc) return rc;
        }
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    // leave the caller's cij as the reference does (putBoundary in place)
    if (nshot > 0) {
        const long long nb = vti::boundary_size(nx, ny, nz, npad);
        const float *shell = grad ? randboundaryCij : padboundaryCij;
        for (int k = 0; k < 3; ++k) vti::put_boundary(shell + k * nb, cij + k * nxyz, nx, ny, nz, npad);
    }
    return vti::kOk;
}

int vti3d_objfunc_gradient(vti_ctx *ctx, float *fgcij, const float *souloc, int ns, const int *shotid, int nshot,
                           const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij,
                           const float *randboundaryCij, int nx, int ny, int nz, int nt, int npad, float ox, float oy,
                           float oz, float ot, float dx, float dy, float dz, float dt, float samplingRate, int ngpu,
                           const float *obs, float tpow, int objtype, int halfwidth)
{
    return objfunc_impl(ctx, fgcij, true, souloc, ns, shotid, nshot, recloc, wavelet, cij, padboundaryCij,
                        randboundaryCij, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, ngpu, obs,
                        tpow, objtype, halfwidth);
}

int vti3d_objfunc(vti_ctx *ctx, float *f, const float *souloc, int ns, const int *shotid, int nshot,
                  const float *recloc, const float *wavelet, float *cij, const float *padboundaryCij, int nx, int ny,
                  int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
                  float dt, float samplingRate, int ngpu, const float *obs, float tpow, int objtype, int halfwidth)
{
    return objfunc_impl(ctx, f, false, souloc, ns, shotid, nshot, recloc, wavelet, cij, padboundaryCij, nullptr, nx,
                        ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt, samplingRate, ngpu, obs, tpow, objtype,
                        halfwidth);
}

// rtmCij3d_f (lib/vti-fwi-3d/gpu/rtm3d_f.cu:161-830): per shot a forward pass
// without damping on c (the caller's model already carries the random
// boundary), then the backward pass that reconstructs the source field,
// propagates the observed data (recloc stride 3, rtm3d_f.cu:348) and images.
// image (nx*ny*nz) is overwritten.  obs replaces read("data") (:509).
int vti3d_rtm(vti_ctx *ctx, float *image, const float *souloc, int ns, const int *shotid, int nshot,
              const float *recloc, const float *wavelet, const float *c11, const float *c13, const float *c33, int nx,
              int ny, int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
              float dt, float samplingRate, int ngpu, const float *obs)
{
    DriverArgs a{souloc, shotid, nshot, recloc, wavelet, nx, ny, nz, nt, npad, ox, oy, oz, ot, dx, dy, dz, dt,
                 samplingRate, ngpu};
    int rc = check_common(a);
    if (rc) return rc;
    VTI_CHECK(ctx && image && c11 && c13 && c33 && obs, "vti3d_rtm: null argument");
    for (int i = 0; i < nshot; ++i) VTI_CHECK(shotid[i] >= 0 && shotid[i] < ns, "shot id %d out of range", shotid[i]);
    PropGuard G;
    rc = vti3d_prop_create(ctx, &G.p, nx, ny, nz, npad, dx, dy, dz, dt);
    if (rc) return rc;
    vti3d_prop *p = G.p;
    rc = vti3d_prop_set_model(p, c11, c13, c33);
    if (rc) return rc;
    rc = ensure_backward(p, false, false);
    if (rc) return rc;
    const int sts = (int)std::round(samplingRate / dt), nnt = (nt - 1) / sts + 1, lmax = levels(nt, ngpu);
    for (int k = 0; k < nshot; ++k) {
        Shot3 sh;
        rc = shot_setup(sh, souloc, shotid[k], recloc, 3, nx, ny, nz, ox, oy, oz, dx, dy, dz);
        if (rc) return rc;
        ShotDev sd(p->ctx->stream);
        rc = to_dev_idx(p, &sd.rec_eff, sh.rec_eff);
        if (rc) return rc;
        rc = to_dev_idx(p, &sd.rec_true, sh.rec_true);
        if (rc) return rc;
        // observed traces, time-major on the device: the shot's trace-major
        // block is uploaded as it is and transposed there
        const size_t nn = (size_t)std::max(1, sh.nr) * nnt;
        DMALLOC(&sd.tr, sizeof(float) * nn);
        if (sh.nr > 0) {
            float *trm = nullptr;
            DMALLOC(&trm, sizeof(float) * nn);
            hipError_t e = hipMemcpyAsync(trm, obs + (size_t)nnt * sh.irbegin, sizeof(float) * nn,
                                          hipMemcpyHostToDevice, p->ctx->stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(transpose_kernel, transpose_tiles(sh.nr, nnt), dim3(32, 8), 0, p->ctx->stream,
                                   sd.tr, trm, sh.nr, nnt);
                e = hipGetLastError();
            }
            freep(trm, p->ctx->stream);
            VTI_HIP(e);
        } else {
            VTI_HIP(hipMemsetAsync(sd.tr, 0, sizeof(float) * nn, p->ctx->stream));
        }
        rc = zero_fields(p, true, true);
        if (rc) return rc;
        hipLaunchKernelGGL(set_point_kernel, dim3(1), dim3(1), 0, p->ctx->stream, p->sx[0], p->sz[0],
                           dev_index(p->g, sh.sou_ref), p->dt2 * wavelet[0]);
        VTI_HIP(hipGetLastError());
        rc = forward_pass(p, false, lmax, wavelet, sh.sou_ref, nullptr, 0, sts, nnt);
        if (rc) return rc;
        std::swap(p->sx[0], p->sx[1]);
        std::swap(p->sz[0], p->sz[1]);
        if (sh.nr > 0)
            hipLaunchKernelGGL(lambda_init3d_kernel, dim3((sh.nr + 255) / 256), dim3(256), 0, p->ctx->stream,
                               p->lx[0], p->lz[0], sd.rec_true, sd.tr + (size_t)(nnt - 1) * sh.nr, sh.nr, p->dt2);
        VTI_HIP(hipGetLastError());
        rc = backward_pass(p, 1, lmax, nt, wavelet, sh.sou_ref, sd, sh.nr, sh.rec_block, sts, nnt);
        if (rc) return rc;
        rc = accumulate(p, 1);
        if (rc) return rc;
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    rc = narrow_download(p, image, 0);
    if (rc) return rc;
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

}  // extern "C"
