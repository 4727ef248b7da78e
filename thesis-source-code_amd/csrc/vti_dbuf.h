// RAII device buffer, zero-filled on allocation (host code of the drivers).
#pragma once

#include <algorithm>
#include <cstddef>

#include <hip/hip_runtime.h>

#include "vti_internal.h"

namespace vti {

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count, hipStream_t s)
    {
        n = std::max<size_t>(count, 1);
        hipError_t e = hipMalloc((void **)&p, n * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            vti::set_error("hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
            return vti::kErrNoMem;
        }
        e = hipMemsetAsync(p, 0, n * sizeof(T), s);
        if (e != hipSuccess) {
            vti::set_error("hipMemsetAsync: %s", hipGetErrorString(e));
            return vti::kErrHip;
        }
        return vti::kOk;
    }
};

}  // namespace vti
