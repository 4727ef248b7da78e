// Error reporting and the execution context shared by all entry points.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "vti_ctx.h"
#include "vti_internal.h"

namespace vti {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *get_error() { return g_err; }

namespace {
std::mutex g_pool_mu;
bool g_pool_ready[256];
}  // namespace

int dmalloc(void **p, size_t bytes, hipStream_t s)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev >= 0 && dev < 256) {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (!g_pool_ready[dev]) {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            g_pool_ready[dev] = true;
        }
    }
    const hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, s);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("hipMallocAsync(%zu bytes): %s", bytes, hipGetErrorString(e));
        return kErrNoMem;
    }
    return kOk;
}

void dfree(void *p, hipStream_t s)
{
    if (p) (void)hipFreeAsync(p, s);
}

}  // namespace vti

extern "C" {

const char *vti_last_error(void) { return vti::get_error(); }

int vti_ctx_create(vti_ctx **out, int device)
{
    VTI_CHECK(out != nullptr, "vti_ctx_create: null output pointer");
    int ndev = 0;
    VTI_HIP(hipGetDeviceCount(&ndev));
    VTI_CHECK(device >= 0 && device < ndev, "vti_ctx_create: device %d not present (%d devices)", device, ndev);
    VTI_HIP(hipSetDevice(device));
    vti_ctx *c = new vti_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        vti::set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return vti::kErrHip;
    }
    *out = c;
    return vti::kOk;
}

int vti_ctx_destroy(vti_ctx *c)
{
    if (!c) return vti::kOk;
    VTI_HIP(hipSetDevice(c->device));
    VTI_HIP(hipStreamSynchronize(c->stream));
    VTI_HIP(hipStreamDestroy(c->stream));
    delete c;
    return vti::kOk;
}

int vti_ctx_set_arith(vti_ctx *c, int arith)
{
    VTI_CHECK(c != nullptr && (arith == 0 || arith == 1), "vti_ctx_set_arith: arith must be 0 (fp32) or 1 (reference)");
    c->arith = arith;
    return vti::kOk;
}

int vti_ctx_synchronize(vti_ctx *c)
{
    VTI_CHECK(c != nullptr, "null context");
    VTI_HIP(hipSetDevice(c->device));
    VTI_HIP(hipStreamSynchronize(c->stream));
    return vti::kOk;
}

}  // extern "C"
