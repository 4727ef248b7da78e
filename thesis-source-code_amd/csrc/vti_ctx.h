// Execution context: one device and the stream every launch of a call goes on.
#pragma once

#include <hip/hip_runtime.h>

struct vti_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // arithmetic of the 3D bulk stencil for objects created on this context:
    // 1 = the reference's mixed FP64 arithmetic (default), 0 = FP32
    int arith = 1;
};
