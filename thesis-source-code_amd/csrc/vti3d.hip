// 3D VTI propagation on one MI355X: device-resident propagator, the
// reference-signature drivers (modelDataCij3d_f, objFuncGradientCij3d_f,
// objFuncCij3d_f, rtmCij3d_f) and the step-level entry points used by the
// benchmark and the serving layer.
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
    float *c[3] = {nullptr, nullptr, nullptr};
    float *dxy = nullptr, *dzv = nullptr;
    float *sx[2] = {nullptr, nullptr}, *sz[2] = {nullptr, nullptr};  // [cur, prev]
    int zc = 64;
    size_t bytes = 0;
};

namespace {

int prop_free(vti3d_prop *p)
{
    if (!p) return vti::kOk;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
    for (auto &q : p->c) hipFree(q);
    hipFree(p->dxy);
    hipFree(p->dzv);
    for (int k = 0; k < 2; ++k) {
        hipFree(p->sx[k]);
        hipFree(p->sz[k]);
    }
    delete p;
    return vti::kOk;
}

int dalloc(vti3d_prop *p, float **ptr, size_t n)
{
    hipError_t e = hipMalloc((void **)ptr, n * sizeof(float));
    if (e != hipSuccess) {
        vti::set_error("hipMalloc(%zu bytes): %s", n * sizeof(float), hipGetErrorString(e));
        return vti::kErrNoMem;
    }
    p->bytes += n * sizeof(float);
    VTI_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(float), p->ctx->stream));
    return vti::kOk;
}

long long vox(const Geo3 &g) { return g.plane * (long long)g.nz; }

// host (reference layout, pitch nx) -> device (pitched)
int upload_volume(vti3d_prop *p, float *dst, const float *src)
{
    const Geo3 &g = p->g;
    VTI_HIP(hipMemcpy2DAsync(dst, (size_t)g.pitch * 4, src, (size_t)g.nx * 4, (size_t)g.nx * 4,
                             (size_t)g.ny * g.nz, hipMemcpyHostToDevice, p->ctx->stream));
    return vti::kOk;
}

int download_volume(vti3d_prop *p, float *dst, const float *src)
{
    const Geo3 &g = p->g;
    VTI_HIP(hipMemcpy2DAsync(dst, (size_t)g.nx * 4, src, (size_t)g.pitch * 4, (size_t)g.nx * 4,
                             (size_t)g.ny * g.nz, hipMemcpyDeviceToHost, p->ctx->stream));
    return vti::kOk;
}

long long dev_index(const Geo3 &g, long long ref_index)
{
    const long long nxy = (long long)g.nx * g.ny;
    const long long iz = ref_index / nxy, rem = ref_index % nxy;
    const long long iy = rem / g.nx, ix = rem % g.nx;
    return ix + iy * g.pitch + iz * g.plane;
}

int fwd_launch(vti3d_prop *p, bool absorbing, long long src_dev, float src_val)
{
    const Geo3 &g = p->g;
    const int ntx = (g.nx - 4 + 63) / 64;
    const int nty = (g.ny - 8 + kBY - 1) / kBY;
    const int ntz = (g.zmax + p->zc - 1) / p->zc;
    const dim3 block(64, kBY);
    const dim3 grid(ntx * nty * ntz);
    float *nx_ = p->sx[1], *nz_ = p->sz[1];
    if (absorbing)
        hipLaunchKernelGGL((vti::fwd3d_kernel<true, kBY>), grid, block, 0, p->ctx->stream, nx_, p->sx[0], nz_,
                           p->sz[0], p->c[0], p->c[1], p->c[2], p->dxy, p->dzv, g, p->zc, ntx, nty, p->dt2dx2,
                           p->dt2dy2, p->dt2dz2, src_dev, src_val);
    else
        hipLaunchKernelGGL((vti::fwd3d_kernel<false, kBY>), grid, block, 0, p->ctx->stream, nx_, p->sx[0], nz_,
                           p->sz[0], p->c[0], p->c[1], p->c[2], p->dxy, p->dzv, g, p->zc, ntx, nty, p->dt2dx2,
                           p->dt2dy2, p->dt2dz2, src_dev, src_val);
    VTI_HIP(hipGetLastError());
    std::swap(p->sx[0], p->sx[1]);
    std::swap(p->sz[0], p->sz[1]);
    return vti::kOk;
}

}  // namespace

extern "C" {

int vti3d_prop_create(vti_ctx *ctx, vti3d_prop **out, int nx, int ny, int nz, int npad, float dx, float dy,
                      float dz, float dt)
{
    VTI_CHECK(ctx && out, "vti3d_prop_create: null argument");
    VTI_CHECK(nx >= 9 && ny >= 9 && nz >= 8, "vti3d_prop_create: grid %dx%dx%d too small", nx, ny, nz);
    VTI_CHECK(npad >= 1 && 2 * npad < nx && 2 * npad < ny && 2 * npad < nz, "vti3d_prop_create: bad npad %d", npad);
    VTI_HIP(hipSetDevice(ctx->device));
    vti3d_prop *p = new vti3d_prop();
    p->ctx = ctx;
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
    if (const char *s = getenv("VTI_ZCHUNK")) p->zc = std::max(4, atoi(s) / 4 * 4);
    const size_t nv = (size_t)vox(g);
    int rc = vti::kOk;
    for (int k = 0; k < 3 && rc == vti::kOk; ++k) rc = dalloc(p, &p->c[k], nv);
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

int vti3d_prop_set_zchunk(vti3d_prop *p, int zc)
{
    VTI_CHECK(p && zc >= 4 && zc % 4 == 0, "vti3d_prop_set_zchunk: chunk must be a positive multiple of 4");
    p->zc = zc;
    return vti::kOk;
}

// Upload c11, c13, c33 (host arrays in the reference layout).
int vti3d_prop_set_model(vti3d_prop *p, const float *c11, const float *c13, const float *c33)
{
    VTI_CHECK(p && c11 && c13 && c33, "vti3d_prop_set_model: null argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    const float *src[3] = {c11, c13, c33};
    for (int k = 0; k < 3; ++k) {
        int rc = upload_volume(p, p->c[k], src[k]);
        if (rc) return rc;
    }
    VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    return vti::kOk;
}

// Zero both time levels and put the reference's level-1 source:
// cur[sou] = value (objFuncGradient3d_f.cu:594-599).  sou_index < 0: no source.
int vti3d_prop_reset(vti3d_prop *p, long long sou_index, float value)
{
    VTI_CHECK(p, "null propagator");
    VTI_HIP(hipSetDevice(p->ctx->device));
    const size_t nb = (size_t)vox(p->g) * sizeof(float);
    for (int k = 0; k < 2; ++k) {
        VTI_HIP(hipMemsetAsync(p->sx[k], 0, nb, p->ctx->stream));
        VTI_HIP(hipMemsetAsync(p->sz[k], 0, nb, p->ctx->stream));
    }
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

// Advance `nsteps` forward levels with no source.  When elapsed_ms is not
// NULL the launches are bracketed by HIP events on the propagator's stream
// and the average time per step is returned there.
int vti3d_prop_forward_steps(vti3d_prop *p, int nsteps, int absorbing, float *elapsed_ms)
{
    VTI_CHECK(p && nsteps >= 0, "vti3d_prop_forward_steps: bad argument");
    VTI_HIP(hipSetDevice(p->ctx->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        VTI_HIP(hipEventCreate(&e0));
        VTI_HIP(hipEventCreate(&e1));
        VTI_HIP(hipEventRecord(e0, p->ctx->stream));
    }
    for (int s = 0; s < nsteps; ++s) {
        int rc = fwd_launch(p, absorbing != 0, -1, 0.f);
        if (rc) return rc;
    }
    if (elapsed_ms) {
        VTI_HIP(hipEventRecord(e1, p->ctx->stream));
        VTI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        VTI_HIP(hipEventElapsedTime(&ms, e0, e1));
        *elapsed_ms = nsteps > 0 ? ms / nsteps : 0.f;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    } else {
        VTI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
    return vti::kOk;
}

// Forward modeling of one source with the reference's level/recording rules
// (objFuncGradient3d_f.cu:594-735): levels 2..lmax, source dt2*wavelet[it-1]
// at sou_index, absorbing damping when `absorbing`, traces recorded when
// it % sts == 0 and it/sts < nnt into traces_tm (time-major, nnt rows of nr;
// row 0 is left untouched).  rec: reference-layout indices.
int vti3d_prop_forward_shot(vti3d_prop *p, int lmax, const float *wavelet, long long sou_index, int nr,
                            const long long *rec, int sts, int nnt, int absorbing, float *traces_tm)
{
    VTI_CHECK(p && wavelet, "vti3d_prop_forward_shot: null argument");
    VTI_CHECK(sts >= 1 && nnt >= 1, "vti3d_prop_forward_shot: bad sampling");
    VTI_HIP(hipSetDevice(p->ctx->device));
    const Geo3 &g = p->g;
    const long long src_dev = dev_index(g, sou_index);
    long long *d_rec = nullptr;
    float *d_tr = nullptr;
    if (nr > 0 && traces_tm) {
        std::vector<long long> rd(nr);
        for (int i = 0; i < nr; ++i) {
            VTI_CHECK(rec[i] >= 0 && rec[i] < (long long)g.nx * g.ny * g.nz, "receiver %d index out of range", i);
            rd[i] = dev_index(g, rec[i]);
        }
        VTI_HIP(hipMalloc((void **)&d_rec, sizeof(long long) * nr));
        VTI_HIP(hipMalloc((void **)&d_tr, sizeof(float) * (size_t)nr * nnt));
        VTI_HIP(hipMemcpyAsync(d_rec, rd.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, p->ctx->stream));
        VTI_HIP(hipMemcpyAsync(d_tr, traces_tm, sizeof(float) * (size_t)nr * nnt, hipMemcpyHostToDevice,
                               p->ctx->stream));
    }
    const float dt2 = p->dt2;
    int rc = vti::kOk;
    for (int it = 2; it <= lmax && rc == vti::kOk; ++it) {
        rc = fwd_launch(p, absorbing != 0, src_dev, dt2 * wavelet[it - 1]);
        if (rc == vti::kOk && d_tr && it % sts == 0 && it / sts < nnt) {
            hipLaunchKernelGGL(record3d_kernel, dim3((nr + 255) / 256), dim3(256), 0, p->ctx->stream,
                               d_tr + (size_t)(it / sts) * nr, p->sx[0], p->sz[0], d_rec, nr);
            if (hipGetLastError() != hipSuccess) rc = vti::kErrHip;
        }
    }
    if (d_tr && rc == vti::kOk) {
        hipError_t e = hipMemcpyAsync(traces_tm, d_tr, sizeof(float) * (size_t)nr * nnt, hipMemcpyDeviceToHost,
                                      p->ctx->stream);
        if (e != hipSuccess) rc = vti::kErrHip;
    }
    hipStreamSynchronize(p->ctx->stream);
    hipFree(d_rec);
    hipFree(d_tr);
    return rc;
}

}  // extern "C"
