// Error reporting and the execution context shared by all entry points.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "vti_ctx.h"
#include "vti_internal.h"
#include "vti_pool.h"

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

// Blocks of at least kBig bytes are kept for reuse (vti_pool.h); smaller
// ones come from the device's stream-ordered pool.
constexpr size_t kBig = size_t(64) << 20;

struct HipRT {
    using Stream = hipStream_t;
    using Event = hipEvent_t;
    static bool malloc(void **p, size_t bytes)
    {
        if (hipMalloc(p, bytes) == hipSuccess) return true;
        (void)hipGetLastError();
        *p = nullptr;
        return false;
    }
    static void free(void *p) { (void)hipFree(p); }
    static bool event_create(Event *e)
    {
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess) return true;
        (void)hipGetLastError();
        return false;
    }
    static bool event_record(Event e, Stream s)
    {
        if (hipEventRecord(e, s) == hipSuccess) return true;
        (void)hipGetLastError();
        return false;
    }
    static void event_sync(Event e) { (void)hipEventSynchronize(e); }
    static void event_destroy(Event e) { (void)hipEventDestroy(e); }
    static void stream_wait(Stream s, Event e) { (void)hipStreamWaitEvent(s, e, 0); }
    static void stream_sync(Stream s) { (void)hipStreamSynchronize(s); }
};

BigPool<HipRT> &big_pool()
{
    static BigPool<HipRT> *pool = new BigPool<HipRT>();  // outlives static destructors
    return *pool;
}

int current_device()
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    return dev;
}
}  // namespace

int dmalloc(void **p, size_t bytes, hipStream_t s)
{
    const int dev = current_device();
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
    if (bytes >= kBig) {
        if (big_pool().take(p, bytes, dev, s) >= 0) return kOk;
        *p = nullptr;
        set_error("hipMalloc(%zu bytes): out of device memory", bytes);
        return kErrNoMem;
    }
    hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        big_pool().flush(dev);
        e = hipMallocAsync(p, bytes ? bytes : 1, s);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("hipMallocAsync(%zu bytes): %s", bytes, hipGetErrorString(e));
        return kErrNoMem;
    }
    return kOk;
}

size_t device_free_bytes()
{
    const int dev = current_device();
    size_t mfree = 0, mtotal = 0;
    if (hipMemGetInfo(&mfree, &mtotal) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return mfree + big_pool().cached_bytes(dev);
}

void dfree(void *p, hipStream_t s)
{
    if (!p) return;
    bool misuse = false;
    if (big_pool().give(p, s, &misuse)) {
        if (misuse) std::fprintf(stderr, "libvtiamd: device block %p released twice (ignored)\n", p);
        return;
    }
    (void)hipFreeAsync(p, s);
}

// Markstein's correction step for a quotient with a known reciprocal
// (P. Markstein, "Computation of elementary functions on the IBM RISC
// System/6000 processor", IBM J. Res. Dev. 34(1), 1990), run in FP32 with
// the operations the device uses: q0 = RN(x y), r = fma(-q0, d, x),
// q = fma(r, y, q0).  A case fails when q differs from RN(x/d) or when r
// is not the exact remainder x - q0 d (exactness is what lets the result
// scale with x's exponent).
long long div_mismatches(float d, float y)
{
    long long bad = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t b = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &b, 4);
        const float q0 = x * y;
        const float r = std::fma(-q0, d, x);
        const float q = std::fma(r, y, q0);
        const float want = x / d;
        const double rex = (double)x - (double)q0 * (double)d;
        if ((double)r != rex || std::memcmp(&q, &want, 4) != 0) ++bad;
    }
    return bad;
}

float div_reciprocal(float d)
{
    static std::mutex mu;
    static std::unordered_map<uint32_t, float> known;
    if (!(d >= 0x1p-20f && d <= 0x1p20f)) return 0.f;
    uint32_t key;
    std::memcpy(&key, &d, 4);
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
    const float y = 1.f / d;
    const float out = div_mismatches(d, y) == 0 ? y : 0.f;
    known.emplace(key, out);
    return out;
}

}  // namespace vti

extern "C" {

// The divisor proof, for tests: RN(1/d) or 0 (see vti::div_reciprocal), and
// the number of failing cases of a given reciprocal.
float vti_div_reciprocal(float d) { return vti::div_reciprocal(d); }
long long vti_div_mismatches(float d, float y) { return vti::div_mismatches(d, y); }

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

int vti_ctx_set_snapshots(vti_ctx *c, int snap_fmt)
{
    VTI_CHECK(c != nullptr && snap_fmt >= 0 && snap_fmt <= 3,
              "vti_ctx_set_snapshots: format must be 0 (FP32), 1 (FP16), 2 (BF16) or 3 (S16)");
    c->snap_fmt = snap_fmt;
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
