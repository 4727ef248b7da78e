// Execution context: one device and the stream every launch of a call goes on.
#pragma once

#include <hip/hip_runtime.h>

struct vti_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};
