// RAII device buffer, zero-filled on allocation (host code of the drivers);
// stream-ordered: allocated and released on the stream that uses it.
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
    hipStream_t s = nullptr;
    bool own = true;  // false: a view of another buffer's memory, not released here
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf()
    {
        if (own) dfree(p, s);
    }
    // view the first count elements of a buffer that outlives this one
    void borrow(const DBuf &from, size_t count)
    {
        p = from.p;
        n = std::max<size_t>(count, 1);
        s = from.s;
        own = false;
    }
    // zero: clear the buffer (the default); a caller that writes every
    // element before reading it, or clears what it reads, passes false
    int alloc(size_t count, hipStream_t st, bool zero = true)
    {
        n = std::max<size_t>(count, 1);
        s = st;
        own = true;
        int rc = dmalloc((void **)&p, n * sizeof(T), st);
        if (rc || !zero) return rc;
        const hipError_t e = hipMemsetAsync(p, 0, n * sizeof(T), st);
        if (e != hipSuccess) {
            vti::set_error("hipMemsetAsync: %s", hipGetErrorString(e));
            return vti::kErrHip;
        }
        return vti::kOk;
    }
};

}  // namespace vti
