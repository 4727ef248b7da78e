// The vectors of the FP64 L-BFGS method (vti_lbfgs.cpp) and the operations
// the method performs on them, on the host or resident in device memory.
#pragma once

#include <cstddef>

struct vti_ctx;

namespace vti {

// Vector ids: kG the gradient just evaluated, kW the gradient at the line
// search's base point (then the two-loop work vector), kS the search
// direction, kX0 the search's base point, kPairs + k the k-th stored step
// and kPairs + m + k the k-th stored gradient difference.  Every operation
// runs in FP64.  An operation that fails records the error (set_error) and
// sets failed(); later operations then do nothing.
class LbfgsOps {
public:
    enum { kG = 0, kW = 1, kS = 2, kX0 = 3, kPairs = 4 };
    virtual ~LbfgsOps() {}
    virtual void load_g(const float *g) = 0;                  // G = g
    virtual double dot(int a, int b) = 0;                     // sum a[i] * b[i]
    virtual void scale_copy(int dst, int src, double a) = 0;  // dst[i] = src[i] * a
    virtual void axpy(int dst, double a, int src) = 0;        // dst[i] += a * src[i]
    virtual void scale(int dst, double a) = 0;                // dst[i] = a * dst[i]
    virtual void copy(int dst, int src) = 0;                  // dst = src
    virtual void store_pair(int slot, double stp) = 0;        // S_slot = stp * S, Y_slot = G - W
    virtual void save_x0(const float *x) = 0;                 // X0 = x
    virtual void set_trial(float *x, double stp) = 0;         // x[i] = (float)(X0[i] + stp * S[i])
    virtual double sumsq_x(const float *x) = 0;               // sum x[i]^2
    bool failed() const { return failed_; }

protected:
    bool failed_ = false;
};

// Host vectors, every sum in index order.  nullptr when out of memory.
LbfgsOps *lbfgs_host_ops(size_t n, int m);
// Vectors in the device memory of ctx's device; x and g may be host or
// device pointers (host arrays are staged through device buffers).  Sums are
// reduced in a fixed order, so a run is reproducible.
int lbfgs_device_ops(vti_ctx *ctx, size_t n, int m, LbfgsOps **out);

}  // namespace vti
