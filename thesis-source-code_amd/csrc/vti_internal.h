// Internal definitions shared by the HIP translation units of libvtiamd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

// Finite-difference weights: the reference's truncated decimals of -205/72,
// 8/5, -1/5, 8/315, -1/560 (lib/vti-fwi-2d/kernels.h:4-8,
// lib/vti-fwi-3d/gpu/wave3d.h:8-12), as float literals.
#define VTI_C0 (-2.84722222222f)
#define VTI_C1 (1.6f)
#define VTI_C2 (-0.2f)
#define VTI_C3 (0.02539682539f)
#define VTI_C4 (-0.00178571428f)

// wave3d.h:5-6 -- the receiver/source weights of the 3D path (doubles there)
#define VTI_ONETHIRD_D 0.3333333333
#define VTI_TWOTHIRD_D 0.6666666666
#define VTI_ONETHIRD_F ((float)VTI_ONETHIRD_D)
#define VTI_TWOTHIRD_F ((float)VTI_TWOTHIRD_D)

struct vti_ctx;

namespace vti {

// Status codes returned by every extern "C" entry point (0 = success).
enum Status : int {
    kOk = 0,
    kErrArg = -1,       // invalid argument / unsupported configuration
    kErrHip = -2,       // a HIP runtime call failed
    kErrNoMem = -3,     // device allocation failed
    kErrState = -4,     // call sequence error (e.g. no model uploaded)
};

// A workgroup barrier that orders LDS only: global loads and stores issued
// before it stay in flight (__syncthreads also waits for every outstanding
// global access).  For kernels whose threads exchange data through LDS only.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

void set_error(const char *fmt, ...);
const char *get_error();

// Device memory of the drivers, stream-ordered: allocated and released on the
// stream that uses it (hipMallocAsync / hipFreeAsync) from the device's
// default pool, whose release threshold is raised so that freed blocks stay
// cached for the next call.  Unlike hipMalloc / hipFree neither waits for the
// work of other streams, so calls on different streams overlap.  dmalloc
// returns kOk, or kErrNoMem with the error message set.
// One request of a merged 2D objective + gradient (vti2d_batch.hip): the
// arguments vti2d_objfunc_gradient takes for it (sloc {ix, iz, nr, r0} per
// shot; data nnt rows of nr), its model in any memory, its outputs on the host.
struct Grad2dPart {
    const float *c[3];
    const float *data;
    const int *sloc;
    int ns;
    const int *rloc;
    int nr;
    int nx;
    float *objfunc;
    float *g[3];
};
bool grad2d_batchable(const int *sloc, int ns, const int *rloc, int nr, const float *taper, int nx, int nz, int nt,
                      int npad, float dt, float rate, float ot);
int synth2d_batched(vti_ctx *ctx, float *out, const float *c11, const float *c13, const float *c33,
                    const float *wavelet, const int *sloc, int ns, const int *rloc, int nr, int nx, int nz, int nt,
                    int npad, float dx, float dz, float dt, float rate);
long long grad2d_merged_bytes(const Grad2dPart *p, int n, int nz, int nt, int npad, float dt, float rate,
                              int max_wave);
int grad2d_merged(vti_ctx *ctx, const Grad2dPart *p, int n, const float *wavelet, int nz, int nt, int npad,
                  float dx, float dz, float dt, float rate, int max_wave, int snap_fmt, bool overlap = false);

int dmalloc(void **p, size_t bytes, hipStream_t s);
// free device memory of the current device, counting the large blocks
// dmalloc keeps for reuse
size_t device_free_bytes();
void dfree(void *p, hipStream_t s);

// RN(1/d) when the kernels' qdiv (x / d by one multiply and two fmas) is
// shown, over every mantissa, to give the correctly rounded quotient; 0 when
// it is not, or d lies outside [2^-20, 2^20] (qdiv then divides).  The proof
// runs once per distinct d in the process (about 8M cases, some 50 ms).
float div_reciprocal(float d);
// mismatches of that sequence, with reciprocal y, over x in [1,2)
long long div_mismatches(float d, float y);

}  // namespace vti

#define VTI_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            vti::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return vti::kErrHip;                                                            \
        }                                                                                   \
    } while (0)

#define VTI_CHECK(cond, ...)                 \
    do {                                     \
        if (!(cond)) {                       \
            vti::set_error(__VA_ARGS__);     \
            return vti::kErrArg;             \
        }                                    \
    } while (0)
