// Execution context: one device and the stream every launch of a call goes on.
#pragma once

#include <hip/hip_runtime.h>

struct vti_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // arithmetic of the 3D bulk stencil for objects created on this context:
    // 1 = the reference's mixed FP64 arithmetic (default), 0 = FP32
    int arith = 1;
    // snapshot storage of 2D gradients on this context: 0 = FP32 (default),
    // 1 = FP16, 2 = BF16, 3 = S16 (16-bit integers, power-of-two scale per
    // 64 points), for the calls the batched driver can take
    int snap_fmt = 0;
};
