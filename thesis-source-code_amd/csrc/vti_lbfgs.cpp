// Limited-memory BFGS with the More'-Thuente line search, reverse
// communication, all scalars and vectors of the method in FP64.
//
// The method is the one lib/lbfgs/lbfgs_c.cpp:11-591 and lsearch.cpp:9-606
// implement (Liu & Nocedal, Math. Prog. B 45 (1989) 503-528; More' & Thuente,
// ACM TOMS 20 (1994) 286-307, the MINPACK-2 routines MCSRCH/MCSTEP), with the
// reference's own choices kept:
//  - m corrections, diagonal H0 = (y's)/(y'y) of the newest pair;
//  - first trial step 1/|g0|; every later line search starts from the step
//    the previous one accepted (lbfgs_c.cpp:433-436: stp is not reset);
//  - a function value above 1e10 (or a NaN) halves the trial step
//    (lsearch.cpp:367-380);
//  - stop when |g| / max(1, |x|) <= 1e-5 (lbfgs_c.cpp:336-340);
//  - ftol 1e-4, gtol 0.9, xtol 1e-16, step bounds [1e-10, 1e10], 20
//    evaluations per search, extrapolation factor 4 (lbfgs.h:8-15).
// The reference computes all of it in FP32 (its dot products are FP32
// OpenMP reductions, la.cpp:7-13); here the dot products, the stored steps
// and gradient differences, the line-search state and the trial point
// x = x0 + stp*s are FP64, and x is rounded to the caller's float array.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "vti_amd.h"
#include "vti_internal.h"
#include "vti_lbfgs_ops.h"

namespace {

constexpr double kEps = 1e-5, kXtol = 1e-16, kFtol = 1e-4, kGtol = 0.9, kStpmin = 1e-10, kStpmax = 1e10;
constexpr double kXtrapf = 4.0;
constexpr int kMaxfev = 20;

// MCSTEP (lsearch.cpp:9-201): safeguarded cubic/quadratic step and update of
// the interval of uncertainty [stx, sty].
void mcstep(double &stx, double &fx, double &dx, double &sty, double &fy, double &dy, double &stp, double fp,
            double dp, bool &brackt, double stmin, double stmax, int &info)
{
    info = 0;
    if ((brackt && (stp <= std::fmin(stx, sty) || stp >= std::fmax(stx, sty))) || dx * (stp - stx) >= 0.0 ||
        stmax < stmin)
        return;
    const double sgnd = dp * (dx / std::fabs(dx));
    bool bound;
    double stpf;
    if (fp > fx) {  // higher function value: the minimum is bracketed
        info = 1;
        bound = true;
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::fmax(std::fmax(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp < stx) gamma = -gamma;
        const double p = gamma - dx + theta, q = gamma - dx + gamma + dp, r = p / q;
        const double stpc = stx + r * (stp - stx);
        const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2) * (stp - stx);
        stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2;
        brackt = true;
    } else if (sgnd < 0.0) {  // lower value, derivatives of opposite sign
        info = 2;
        bound = false;
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::fmax(std::fmax(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
        if (stp > stx) gamma = -gamma;
        const double p = gamma - dp + theta, q = gamma - dp + gamma + dx, r = p / q;
        const double stpc = stp + r * (stx - stp);
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {  // lower value, same sign, |derivative| decreases
        info = 3;
        bound = true;
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::fmax(std::fmax(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
        double gamma = s * std::sqrt(std::fmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = gamma - dp + theta, q = gamma + (dx - dp) + gamma, r = p / q;
        double stpc;
        if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
        else if (stp > stx) stpc = stmax;
        else stpc = stmin;
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        if (brackt) stpf = std::fabs(stp - stpc) < std::fabs(stp - stpq) ? stpc : stpq;
        else stpf = std::fabs(stp - stpc) > std::fabs(stp - stpq) ? stpc : stpq;
    } else {  // lower value, same sign, |derivative| does not decrease
        info = 4;
        bound = false;
        if (brackt) {
            const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
            const double s = std::fmax(std::fmax(std::fabs(theta), std::fabs(dy)), std::fabs(dp));
            double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
            if (stp > sty) gamma = -gamma;
            const double p = gamma - dp + theta, q = gamma - dp + gamma + dy, r = p / q;
            stpf = stp + r * (sty - stp);
        } else {
            stpf = stp > stx ? stmax : stmin;
        }
    }
    if (fp > fx) {
        sty = stp;
        fy = fp;
        dy = dp;
    } else {
        if (sgnd < 0.0) {
            sty = stx;
            fy = fx;
            dy = dx;
        }
        stx = stp;
        fx = fp;
        dx = dp;
    }
    stpf = std::fmax(stmin, std::fmin(stmax, stpf));
    stp = stpf;
    if (brackt && bound) {
        if (sty > stx) stp = std::fmin(stx + 0.66 * (sty - stx), stp);
        else stp = std::fmax(stx + 0.66 * (sty - stx), stp);
    }
}

// Host vectors: every loop and sum in index order.
class HostOps final : public vti::LbfgsOps {
public:
    explicit HostOps(size_t n, int m) : n_(n), v_((size_t)(kPairs + 2 * m), std::vector<double>(n, 0.0)) {}
    void load_g(const float *g) override
    {
        double *G = v_[kG].data();
        for (size_t i = 0; i < n_; ++i) G[i] = g[i];
    }
    double dot(int a, int b) override
    {
        const double *x = v_[a].data(), *y = v_[b].data();
        double s = 0.0;
        for (size_t i = 0; i < n_; ++i) s += x[i] * y[i];
        return s;
    }
    void scale_copy(int dst, int src, double a) override
    {
        double *d = v_[dst].data();
        const double *x = v_[src].data();
        for (size_t i = 0; i < n_; ++i) d[i] = x[i] * a;
    }
    void axpy(int dst, double a, int src) override
    {
        double *d = v_[dst].data();
        const double *x = v_[src].data();
        for (size_t i = 0; i < n_; ++i) d[i] += a * x[i];
    }
    void scale(int dst, double a) override
    {
        double *d = v_[dst].data();
        for (size_t i = 0; i < n_; ++i) d[i] = a * d[i];
    }
    void copy(int dst, int src) override { v_[dst] = v_[src]; }
    void store_pair(int slot, double stp) override
    {
        const int m = (int)(v_.size() - kPairs) / 2;
        double *S = v_[kPairs + slot].data(), *Y = v_[kPairs + m + slot].data();
        const double *s = v_[kS].data(), *g = v_[kG].data(), *w = v_[kW].data();
        for (size_t i = 0; i < n_; ++i) {
            S[i] = stp * s[i];
            Y[i] = g[i] - w[i];
        }
    }
    void save_x0(const float *x) override
    {
        double *X0 = v_[kX0].data();
        for (size_t i = 0; i < n_; ++i) X0[i] = x[i];
    }
    void set_trial(float *x, double stp) override
    {
        const double *X0 = v_[kX0].data(), *s = v_[kS].data();
        for (size_t i = 0; i < n_; ++i) x[i] = (float)(X0[i] + stp * s[i]);
    }
    double sumsq_x(const float *x) override
    {
        double s = 0.0;
        for (size_t i = 0; i < n_; ++i) s += (double)x[i] * (double)x[i];
        return s;
    }

private:
    size_t n_;
    std::vector<std::vector<double>> v_;
};

}  // namespace

vti::LbfgsOps *vti::lbfgs_host_ops(size_t n, int m)
{
    try {
        return new HostOps(n, m);
    } catch (...) {
        return nullptr;
    }
}

struct vti_lbfgs {
    int m = 0;
    // the method's vectors, on the host or in device memory (vti::LbfgsOps:
    // G, W, S, X0 and the last m steps and gradient differences in circular
    // order, lbfgs_c.cpp:226-240); rho[m], alpha[m] and the scalars here
    std::unique_ptr<vti::LbfgsOps> v;
    std::vector<double> rho, alpha;
    int iter = 0, point = 0, nfun = 0, bound = 0, nfev = 0, info = 0;
    double stp = 0.0;
    // line-search state (MCSRCH's SAVE variables)
    int infoc = 1;
    bool brackt = false, stage1 = true;
    double dginit = 0, finit = 0, dgtest = 0, width = 0, width1 = 0, stx = 0, fx = 0, dgx = 0, sty = 0, fy = 0,
           dgy = 0, stmin = 0, stmax = 0;
    bool active = false;
    bool unit_step = false;  // restart each line search at stp = 1 (Liu & Nocedal) instead

    using Ops = vti::LbfgsOps;
    int Sk(int k) const { return Ops::kPairs + k; }
    int Yk(int k) const { return Ops::kPairs + m + k; }

    // x = x0 + stp*s in FP64, rounded to the caller's float x
    void set_trial(float *x) { v->set_trial(x, stp); }

    // MCSRCH entry with a new direction s (lsearch.cpp:493-549), the gradient
    // at x in G; returns -1 (evaluate f, g at x) or 0 (bad input / not a
    // descent direction)
    int search_start(float *x, double f)
    {
        infoc = 1;
        if (stp <= 0.0) return 0;
        dginit = v->dot(Ops::kG, Ops::kS);
        if (dginit >= 0.0) return 0;
        brackt = false;
        stage1 = true;
        nfev = 0;
        finit = f;
        dgtest = kFtol * dginit;
        width = kStpmax - kStpmin;
        width1 = 2 * width;
        v->save_x0(x);
        stx = 0.0;
        fx = finit;
        dgx = dginit;
        sty = 0.0;
        fy = finit;
        dgy = dginit;
        return next_trial(x);
    }

    // the common tail of MCSRCH (lsearch.cpp:551-605)
    int next_trial(float *x)
    {
        if (brackt) {
            stmin = std::fmin(stx, sty);
            stmax = std::fmax(stx, sty);
        } else {
            stmin = stx;
            stmax = stp + kXtrapf * (stp - stx);
        }
        stp = std::fmin(std::fmax(stp, kStpmin), kStpmax);
        if ((brackt && (stp <= stmin || stp >= stmax)) || nfev >= kMaxfev - 1 || infoc == 0 ||
            (brackt && stmax - stmin <= kXtol * stmax))
            stp = stx;
        set_trial(x);
        return -1;
    }

    // MCSRCH re-entry with f, g (in G) at the trial point (lsearch.cpp:344-491):
    // -1 = evaluate again, 1 = step accepted, 2..6 = the reference's failures
    int search_continue(float *x, double f)
    {
        ++nfev;
        const double dg = v->dot(Ops::kG, Ops::kS);
        const double ftest1 = finit + stp * dgtest;
        // a function value above 1e10 halves the step (lsearch.cpp:367-380);
        // so does a NaN (an unstable trial model), which the reference's
        // comparison lets through
        if (!(f <= 1e10)) {
            stp = stp / 2;
            set_trial(x);
            return -1;
        }
        int res = 0;
        if ((brackt && (stp <= stmin || stp >= stmax)) || infoc == 0) res = 6;
        else if (stp == kStpmax && f <= ftest1 && dg <= dgtest) res = 5;
        else if (stp == kStpmin && (f > ftest1 || dg >= dgtest)) res = 4;
        else if (nfev >= kMaxfev) res = 3;
        else if (brackt && stmax - stmin <= kXtol * stmax) res = 2;
        else if (f <= ftest1 && std::fabs(dg) <= kGtol * (-dginit)) res = 1;
        if (res != 0) return res;
        if (stage1 && f <= ftest1 && dg >= std::fmin(kFtol, kGtol) * dginit) stage1 = false;
        if (stage1 && f <= fx && f > ftest1) {
            double fm = f - stp * dgtest, fxm = fx - stx * dgtest, fym = fy - sty * dgtest;
            double dgm = dg - dgtest, dgxm = dgx - dgtest, dgym = dgy - dgtest;
            mcstep(stx, fxm, dgxm, sty, fym, dgym, stp, fm, dgm, brackt, stmin, stmax, infoc);
            fx = fxm + stx * dgtest;
            fy = fym + sty * dgtest;
            dgx = dgxm + dgtest;
            dgy = dgym + dgtest;
        } else {
            mcstep(stx, fx, dgx, sty, fy, dgy, stp, f, dg, brackt, stmin, stmax, infoc);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
        }
        return next_trial(x);
    }

    // -H*g by the two-loop recursion (lbfgs_c.cpp:371-428) into S; the
    // diagonal H0 = (y's)/(y'y) of the newest pair is one scalar
    void direction()
    {
        const int last = (point + m - 1) % m;
        const double ys = v->dot(Yk(last), Sk(last));
        const double yy = v->dot(Yk(last), Yk(last));
        const double h0 = ys / yy;
        rho[last] = 1.0 / ys;
        v->scale_copy(Ops::kW, Ops::kG, -1.0);
        int cp = point;
        for (int k = 0; k < bound; ++k) {
            cp = (cp + m - 1) % m;
            const double sq = v->dot(Sk(cp), Ops::kW);
            alpha[cp] = rho[cp] * sq;
            v->axpy(Ops::kW, -alpha[cp], Yk(cp));
        }
        v->scale(Ops::kW, h0);
        for (int k = 0; k < bound; ++k) {
            const double yr = v->dot(Yk(cp), Ops::kW);
            const double beta = alpha[cp] - rho[cp] * yr;
            v->axpy(Ops::kW, beta, Sk(cp));
            cp = (cp + 1) % m;
        }
        v->copy(Ops::kS, Ops::kW);
    }
};

namespace {

int init_state(vti_lbfgs *s, int m, int unit_step)
{
    s->m = m;
    s->unit_step = unit_step != 0;
    s->rho.assign(m, 0.0);
    s->alpha.assign(m, 0.0);
    return vti::kOk;
}

}  // namespace

extern "C" {

int vti_lbfgs_create(vti_lbfgs **out, long long n, int m, int unit_step)
{
    VTI_CHECK(out && n > 0 && m > 0, "vti_lbfgs_create: n and m must be positive");
    vti_lbfgs *s = new (std::nothrow) vti_lbfgs();
    VTI_CHECK(s, "vti_lbfgs_create: out of host memory");
    s->v.reset(vti::lbfgs_host_ops((size_t)n, m));
    if (!s->v) {
        delete s;
        VTI_CHECK(false, "vti_lbfgs_create: out of host memory for n=%lld, m=%d", n, m);
    }
    init_state(s, m, unit_step);
    *out = s;
    return vti::kOk;
}

int vti_lbfgs_create_device(vti_lbfgs **out, vti_ctx *ctx, long long n, int m, int unit_step)
{
    VTI_CHECK(out && ctx && n > 0 && m > 0, "vti_lbfgs_create_device: null context or non-positive n, m");
    vti_lbfgs *s = new (std::nothrow) vti_lbfgs();
    VTI_CHECK(s, "vti_lbfgs_create_device: out of host memory");
    vti::LbfgsOps *ops = nullptr;
    const int rc = vti::lbfgs_device_ops(ctx, (size_t)n, m, &ops);
    if (rc) {
        delete s;
        return rc;
    }
    s->v.reset(ops);
    init_state(s, m, unit_step);
    *out = s;
    return vti::kOk;
}

int vti_lbfgs_destroy(vti_lbfgs *s)
{
    delete s;
    return vti::kOk;
}

// One reverse-communication call (lbfgs, lib/lbfgs/lbfgs_c.cpp:11; lbfgs.h:32).
int vti_lbfgs_step(vti_lbfgs *s, float *x, double f, const float *gf, int *iflag, int *iter, int *nfun)
{
    VTI_CHECK(s && x && gf && iflag, "vti_lbfgs_step: null argument");
    using Ops = vti::LbfgsOps;
    vti::LbfgsOps *v = s->v.get();
    VTI_CHECK(!v->failed(), "vti_lbfgs_step: an earlier call failed; the state is not usable");
    v->load_g(gf);
    int res;
    if (*iflag == 0) {
        // initial call (lbfgs_c.cpp:202-262): H0 = I
        s->iter = 1;
        s->nfun = 1;
        s->point = 0;
        s->bound = 0;
        s->info = 0;
        v->scale_copy(Ops::kS, Ops::kG, -1.0);
        const double gnorm = std::sqrt(v->dot(Ops::kG, Ops::kG));
        s->stp = 1.0 / gnorm;
        s->active = true;
        res = s->search_start(x, f);
        if (res == 0 && !v->failed()) {
            // a zero gradient (or no descent direction) at the start: the
            // reference's first call then returns with iflag still 0
            // (lbfgs_c.cpp:262-276), i.e. done
            s->active = false;
            if (iter) *iter = s->iter;
            if (nfun) *nfun = s->nfun;
            return vti::kOk;
        }
    } else if (*iflag == 1 && s->active) {
        res = s->search_continue(x, f);
        if (res == 1) {
            // accepted: store the step and gradient change (lbfgs_c.cpp:314-325)
            s->nfun += s->nfev;
            v->store_pair(s->point, s->stp);
            s->point = (s->point + 1) % s->m;
            const double gnorm = std::sqrt(v->dot(Ops::kG, Ops::kG));
            const double xnorm = std::fmax(1.0, std::sqrt(v->sumsq_x(x)));
            if (gnorm / xnorm <= kEps && !v->failed()) {
                s->active = false;
                *iflag = 0;
                if (iter) *iter = s->iter;
                if (nfun) *nfun = s->nfun;
                return vti::kOk;
            }
            ++s->iter;
            s->bound = std::min(s->iter - 1, s->m);
            s->direction();
            if (s->unit_step) s->stp = 1.0;  // else: the step just accepted, as the reference does
            res = s->search_start(x, f);
        }
    } else {
        VTI_CHECK(false, "vti_lbfgs_step: iflag must be 0 (start) or 1 (f, g evaluated at x), got %d", *iflag);
    }
    if (v->failed()) {
        s->active = false;
        return vti::kErrHip;  // the failing operation set the message
    }
    // the gradient at the base point of the search just started (w, lbfgs_c.cpp:246,436)
    if (res == -1 && s->nfev == 0) v->copy(Ops::kW, Ops::kG);
    if (iter) *iter = s->iter;
    if (nfun) *nfun = s->nfun;
    if (v->failed()) {
        s->active = false;
        return vti::kErrHip;
    }
    if (res == -1) {
        *iflag = 1;
        return vti::kOk;
    }
    s->active = false;
    *iflag = -1;
    vti::set_error("line search failed: info=%d (%s)", res,
                   res == 0 ? "not a descent direction or bad step" :
                   res == 2 ? "interval of uncertainty below xtol" :
                   res == 3 ? "20 function evaluations in this search" :
                   res == 4 ? "step at the lower bound" :
                   res == 5 ? "step at the upper bound" : "rounding errors prevent progress");
    return vti::kOk;
}

}  // extern "C"
