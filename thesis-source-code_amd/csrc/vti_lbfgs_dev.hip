// Device-resident vectors of the FP64 L-BFGS method (vti_lbfgs.cpp): the
// m stored steps and gradient differences, the search direction, the base
// point and the gradients stay in HBM as FP64, and every vector operation of
// the method (lbfgs_c.cpp:226-240,350-428; la.cpp:7-31) is a kernel on the
// context's stream.  Only the scalars of each dot product come back to the
// host, where the line search (lsearch.cpp) runs.
//
// Dot products are reduced in a fixed order -- a fixed grid of kRedBlocks
// blocks, each thread summing its grid-stride elements in index order, a
// tree within each block, a tree over the block sums -- so a run is
// reproducible bit for bit.  x and g may be host or device pointers; host
// arrays are staged through device float buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "vti_ctx.h"
#include "vti_internal.h"
#include "vti_lbfgs_ops.h"

namespace {

constexpr int kRedBlocks = 512;
constexpr int kThreads = 256;

int blocks_for(size_t n) { return (int)std::min<size_t>(2048, std::max<size_t>(1, (n + kThreads - 1) / kThreads)); }

__device__ __forceinline__ void block_sum_store(double acc, double *__restrict__ out)
{
    __shared__ double red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// partial[b] = sum over the block's grid-stride elements of a[i] * b[i]
template <class A, class B>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const A *__restrict__ a, const B *__restrict__ b,
                                                                long long n, double *__restrict__ partial)
{
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        acc += (double)a[i] * (double)b[i];
    block_sum_store(acc, partial + blockIdx.x);
}

// *out = sum of the kRedBlocks partials, as a tree
__global__ __launch_bounds__(kThreads) void dot_final_kernel(const double *__restrict__ partial,
                                                              double *__restrict__ out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kThreads) acc += partial[i];
    block_sum_store(acc, out);
}

__global__ void widen_kernel(double *__restrict__ d, const float *__restrict__ s, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__global__ void scale_copy_kernel(double *__restrict__ d, const double *__restrict__ s, double a, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = s[i] * a;
}

__global__ void axpy_kernel(double *__restrict__ d, double a, const double *__restrict__ s, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] += a * s[i];
}

__global__ void scale_kernel(double *__restrict__ d, double a, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = a * d[i];
}

__global__ void pair_kernel(double *__restrict__ S, double *__restrict__ Y, const double *__restrict__ s,
                            const double *__restrict__ g, const double *__restrict__ w, double stp, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        S[i] = stp * s[i];
        Y[i] = g[i] - w[i];
    }
}

__global__ void trial_kernel(float *__restrict__ x, const double *__restrict__ x0, const double *__restrict__ s,
                             double stp, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = (float)(x0[i] + stp * s[i]);
}

class DeviceOps final : public vti::LbfgsOps {
public:
    DeviceOps(vti_ctx *ctx, size_t n, int m) : ctx_(ctx), n_(n), m_(m), v_((size_t)(kPairs + 2 * m), nullptr) {}
    ~DeviceOps() override
    {
        (void)hipSetDevice(ctx_->device);
        (void)hipStreamSynchronize(ctx_->stream);
        for (double *p : v_) (void)hipFree(p);
        (void)hipFree(xs_);
        (void)hipFree(gs_);
        (void)hipFree(part_);
        (void)hipFree(res_);
        (void)hipHostFree(hres_);
    }

    int init()
    {
        VTI_HIP(hipSetDevice(ctx_->device));
        for (auto &p : v_) {
            VTI_HIP(hipMalloc(&p, sizeof(double) * n_));
            VTI_HIP(hipMemsetAsync(p, 0, sizeof(double) * n_, ctx_->stream));
        }
        VTI_HIP(hipMalloc(&xs_, sizeof(float) * n_));
        VTI_HIP(hipMalloc(&gs_, sizeof(float) * n_));
        VTI_HIP(hipMalloc(&part_, sizeof(double) * kRedBlocks));
        VTI_HIP(hipMalloc(&res_, sizeof(double)));
        VTI_HIP(hipHostMalloc(&hres_, sizeof(double)));
        VTI_HIP(hipStreamSynchronize(ctx_->stream));
        return vti::kOk;
    }

    void load_g(const float *g) override { widen(kG, g); }
    double dot(int a, int b) override
    {
        if (!begin()) return 0.0;
        hipLaunchKernelGGL((dot_partial_kernel<double, double>), dim3(kRedBlocks), dim3(kThreads), 0, ctx_->stream,
                           v_[a], v_[b], (long long)n_, part_);
        return finish_dot();
    }
    void scale_copy(int dst, int src, double a) override
    {
        if (!begin()) return;
        hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, v_[dst], v_[src],
                           a, (long long)n_);
        end();
    }
    void axpy(int dst, double a, int src) override
    {
        if (!begin()) return;
        hipLaunchKernelGGL(axpy_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, v_[dst], a, v_[src],
                           (long long)n_);
        end();
    }
    void scale(int dst, double a) override
    {
        if (!begin()) return;
        hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, v_[dst], a,
                           (long long)n_);
        end();
    }
    void copy(int dst, int src) override
    {
        if (!begin()) return;
        check(hipMemcpyAsync(v_[dst], v_[src], sizeof(double) * n_, hipMemcpyDeviceToDevice, ctx_->stream),
              "hipMemcpyAsync");
    }
    void store_pair(int slot, double stp) override
    {
        if (!begin()) return;
        hipLaunchKernelGGL(pair_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, v_[kPairs + slot],
                           v_[kPairs + m_ + slot], v_[kS], v_[kG], v_[kW], stp, (long long)n_);
        end();
    }
    void save_x0(const float *x) override { widen(kX0, x); }
    void set_trial(float *x, double stp) override
    {
        if (!begin()) return;
        const bool dev = is_device(x);
        float *dst = dev ? x : xs_;
        hipLaunchKernelGGL(trial_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, dst, v_[kX0], v_[kS],
                           stp, (long long)n_);
        if (!end()) return;
        if (!dev && !check(hipMemcpyAsync(x, xs_, sizeof(float) * n_, hipMemcpyDeviceToHost, ctx_->stream),
                           "hipMemcpyAsync"))
            return;
        check(hipStreamSynchronize(ctx_->stream), "hipStreamSynchronize");  // x is the caller's next input
    }
    double sumsq_x(const float *x) override
    {
        if (!begin()) return 0.0;
        const float *src = stage(x);
        if (!src) return 0.0;
        hipLaunchKernelGGL((dot_partial_kernel<float, float>), dim3(kRedBlocks), dim3(kThreads), 0, ctx_->stream, src,
                           src, (long long)n_, part_);
        return finish_dot();
    }

private:
    bool check(hipError_t e, const char *what)
    {
        if (e != hipSuccess && !failed_) {
            vti::set_error("device L-BFGS: %s: %s", what, hipGetErrorString(e));
            failed_ = true;
        }
        return !failed_;
    }
    bool begin()
    {
        if (failed_) return false;
        return check(hipSetDevice(ctx_->device), "hipSetDevice");
    }
    bool end() { return check(hipGetLastError(), "kernel launch"); }

    // device memory (or managed) that kernels can read and write directly
    static bool is_device(const void *p)
    {
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, p) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
    }
    // a device pointer holding x (the caller's, or the staging copy)
    const float *stage(const float *x)
    {
        if (is_device(x)) return x;
        if (!check(hipMemcpyAsync(gs_, x, sizeof(float) * n_, hipMemcpyHostToDevice, ctx_->stream), "hipMemcpyAsync"))
            return nullptr;
        return gs_;
    }
    void widen(int dst, const float *x)
    {
        if (!begin()) return;
        const float *src = stage(x);
        if (!src) return;
        hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(n_)), dim3(kThreads), 0, ctx_->stream, v_[dst], src,
                           (long long)n_);
        // the staging buffer is reused by the next call
        if (end()) check(hipStreamSynchronize(ctx_->stream), "hipStreamSynchronize");
    }
    double finish_dot()
    {
        hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kThreads), 0, ctx_->stream, part_, res_);
        if (!end()) return 0.0;
        if (!check(hipMemcpyAsync(hres_, res_, sizeof(double), hipMemcpyDeviceToHost, ctx_->stream),
                   "hipMemcpyAsync"))
            return 0.0;
        if (!check(hipStreamSynchronize(ctx_->stream), "hipStreamSynchronize")) return 0.0;
        return *hres_;
    }

    vti_ctx *ctx_;
    size_t n_;
    int m_;
    std::vector<double *> v_;
    float *xs_ = nullptr, *gs_ = nullptr;
    double *part_ = nullptr, *res_ = nullptr, *hres_ = nullptr;
};

}  // namespace

int vti::lbfgs_device_ops(vti_ctx *ctx, size_t n, int m, LbfgsOps **out)
{
    DeviceOps *d = new (std::nothrow) DeviceOps(ctx, n, m);
    VTI_CHECK(d, "vti_lbfgs_create_device: out of host memory");
    const int rc = d->init();
    if (rc) {
        delete d;
        return rc == vti::kErrHip ? vti::kErrNoMem : rc;
    }
    *out = d;
    return vti::kOk;
}
