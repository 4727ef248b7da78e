// Request server on one device: 2D gradient and 3D modeling requests of
// different grid sizes, admitted as they arrive, run concurrently.
//
// `slots` worker threads each own a context (device + non-blocking stream).
// A worker takes the oldest queued request as soon as it is free and the
// request's device-memory estimate fits in the budget left after the
// resident models and the requests in flight; the request then runs through
// the ordinary drivers on the worker's stream, so requests on different
// slots overlap on the device.  Driver memory is stream-ordered pool memory
// (vti::dmalloc), which neither waits for other streams nor is returned to
// the system between requests.  Models are uploaded once and kept resident;
// a request names its model by handle and the drivers copy it device to
// device.
//
// 2D gradient requests that share a schedule (grid height, padding, steps,
// spacings, time step, snapshot rate, wavelet) and use the standard taper are
// batched at kernel level: a worker that takes such a request also takes the
// queued requests it can merge with (up to `batch` of them, within the memory
// budget) and runs them as one batched launch sequence over their models
// placed side by side (vti::grad2d_merged).  Each request still gets, bit for
// bit, what vti2d_objfunc_gradient returns for it alone.  If the merged run
// fails, its requests are run one by one.
//
// Latency of a request = completion - submission (steady clock, seconds).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vti_amd.h"
#include "vti_ctx.h"
#include "vti_host.h"
#include "vti_internal.h"
#include "vti_serve_policy.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Model {
    float *c[3] = {nullptr, nullptr, nullptr};
    long long n = 0;
    int users = 0;  // queued or running requests on it
};

enum Kind { kGrad2d = 0, kModel3d = 1 };

struct Request {
    long long id = 0;
    Kind kind = kGrad2d;
    int model = -1;
    long long bytes = 0;
    // 2D gradient
    float *objfunc = nullptr, *g[3] = {nullptr, nullptr, nullptr};
    const float *data = nullptr, *wavelet = nullptr, *taper = nullptr;
    const int *sloc = nullptr, *rloc = nullptr;
    int ns = 0, nr = 0, nx = 0, ny = 0, nz = 0, nt = 0, npad = 0;
    float dx = 0, dy = 0, dz = 0, dt = 0, rate = 0;
    // 3D modeling
    float *ddata = nullptr;
    const float *souloc = nullptr, *recloc = nullptr;
    const int *shotid = nullptr;
    int nsou = 0, nshot = 0;
    float ox = 0, oy = 0, oz = 0, ot = 0;
    bool mergeable = false;  // 2D: may join a merged batch
    // outcome
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = 0;
    std::string error;
};

}  // namespace

struct vti_server {
    int device = 0;
    long long budget = 0;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request *> queue;
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    std::unordered_map<int, Model> models;
    std::vector<std::thread> workers;
    long long next_id = 1;
    int next_model = 1;
    long long resident = 0, inflight = 0, peak = 0;
    bool stop = false;
    std::vector<double> latencies;
    Clock::time_point first_submit, last_done;
    bool any = false;
    vti_ctx *upload_ctx = nullptr;
    int batch = 8;                        // most requests merged into one batched run
    int snap_fmt = 0;                     // snapshot storage of 2D gradient requests
    bool exclusive3d = true;              // 3D modeling requests run with nothing beside them
    int running = 0;                      // runs in flight
    bool excl_running = false;            // one of them is an exclusive 3D request
    int idle = 0;                         // workers waiting for work
    long long groups = 0, grouped = 0;    // merged runs and the requests in them
};

namespace {

long long est_grad2d(const Request &r)
{
    const long long n = (long long)(r.nx + 2 * r.npad) * (r.nz + 2 * r.npad);
    const int ratio = (int)(r.rate / r.dt + 0.5f);
    const long long nnt = ratio > 0 ? (r.nt - 1) / ratio + 1 : r.nt;
    // model 3, taper 1, p/q 4, snapshots 2 per slot, FP32 and FP64 accumulators,
    // receiver map; traces
    return 4LL * n * (3 + 1 + 4 + 2 * nnt + 3 + 6 + 1) + 16LL * nnt * std::max(r.nr, 1) + (1LL << 20);
}

long long est_model3d(const Request &r)
{
    const long long pitch = (r.nx + 63) / 64 * 64;
    const long long vox = pitch * r.ny * r.nz;
    const long long nnt = (long long)((r.nt - 1) / std::max(1, (int)(r.rate / r.dt + 0.5f))) + 1;
    // model 3, two levels of two fields, damping; traces
    return 4LL * vox * 7 + 16LL * nnt * std::max(r.nr, 1) + (1LL << 20);
}

// Whether a 2D request can join a merged batch (vti::grad2d_batchable); a
// request that fails here runs alone through vti2d_objfunc_gradient, which
// reports its error.
bool mergeable_grad2d(const Request &r)
{
    return vti::grad2d_batchable(r.sloc, r.ns, r.rloc, r.nr, r.taper, r.nx, r.nz, r.nt, r.npad, r.dt, r.rate, 0.f);
}

bool same_schedule(const Request &a, const Request &b)
{
    auto bits = [](float v) {
        unsigned u;
        std::memcpy(&u, &v, 4);
        return u;
    };
    return a.nz == b.nz && a.npad == b.npad && a.nt == b.nt && bits(a.dx) == bits(b.dx) &&
           bits(a.dz) == bits(b.dz) && bits(a.dt) == bits(b.dt) && bits(a.rate) == bits(b.rate) &&
           (a.wavelet == b.wavelet || std::memcmp(a.wavelet, b.wavelet, sizeof(float) * a.nt) == 0);
}

vti::Grad2dPart part_of(const Request &r, const Model &m)
{
    vti::Grad2dPart p;
    for (int k = 0; k < 3; ++k) {
        p.c[k] = m.c[k];
        p.g[k] = r.g[k];
    }
    p.data = r.data;
    p.sloc = r.sloc;
    p.ns = r.ns;
    p.rloc = r.rloc;
    p.nr = r.nr;
    p.nx = r.nx;
    p.objfunc = r.objfunc;
    return p;
}

// shots per wave of a merged run: as many as requests it may merge
long long merged_bytes(const std::vector<Request *> &g, int wave)
{
    std::vector<vti::Grad2dPart> parts;
    const Model none;
    for (const Request *q : g) parts.push_back(part_of(*q, none));
    const Request &r = *g[0];
    return vti::grad2d_merged_bytes(parts.data(), (int)parts.size(), r.nz, r.nt, r.npad, r.dt, r.rate, wave);
}

int run(vti_ctx *ctx, const Model &m, Request &r)
{
    if (r.kind == kGrad2d)
        return vti2d_objfunc_gradient(ctx, r.objfunc, r.g[0], r.g[1], r.g[2], r.data, m.c[0], m.c[1], m.c[2],
                                      r.wavelet, r.sloc, r.ns, r.rloc, r.nr, r.taper, r.nx, r.nz, r.nt, r.npad, r.dx,
                                      r.dz, r.dt, r.rate, 0.f);
    return vti3d_model_data(ctx, r.ddata, r.souloc, r.nsou, r.shotid, r.nshot, r.recloc, r.wavelet, m.c[0], m.c[1],
                            m.c[2], r.nx, r.ny, r.nz, r.nt, r.npad, r.ox, r.oy, r.oz, r.ot, r.dx, r.dy, r.dz, r.dt,
                            r.rate, 1);
}

void worker(vti_server *s)
{
    vti_ctx *ctx = nullptr;
    const int crc = vti_ctx_create(&ctx, s->device);
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        // the oldest request, once it fits beside those in flight (or runs
        // alone); on shutdown the queue is drained first.  A 3D modeling
        // request (bound by HBM bandwidth, so sharing only stretches it)
        // waits until nothing runs and keeps the device to itself, as the
        // one-request-at-a-time path would give it
        s->idle += 1;
        s->work_cv.wait(lk, [&] {
            if (s->queue.empty()) return s->stop;
            vti::ServeState st;
            st.budget = s->budget;
            st.resident = s->resident;
            st.inflight = s->inflight;
            st.running = s->running;
            st.excl_running = s->excl_running;
            st.exclusive3d = s->exclusive3d;
            const Request *f = s->queue.front();
            return vti::may_start(st, f->kind == kModel3d, f->bytes);
        });
        s->idle -= 1;
        if (s->queue.empty()) break;
        Request *r = s->queue.front();
        s->queue.pop_front();
        const bool excl = s->exclusive3d && r->kind == kModel3d;
        s->running += 1;
        if (excl) s->excl_running = true;
        // queued requests that merge with it, oldest first, while the merged
        // run fits beside those in flight
        std::vector<Request *> grp{r};
        long long gbytes = r->bytes;
        if (s->batch > 1 && r->kind == kGrad2d && r->mergeable) {
            const long long room = s->budget - s->resident;
            long long width = r->nx + 2 * r->npad;
            // this worker's share of the mergeable requests, so that idle
            // workers take the rest as runs of their own
            int compatible = 1;
            for (const Request *q : s->queue)
                if (q->kind == kGrad2d && q->mergeable && same_schedule(*r, *q)) ++compatible;
            const int share = std::min(s->batch, (compatible + s->idle) / (s->idle + 1));
            for (auto it = s->queue.begin(); it != s->queue.end() && (int)grp.size() < share;) {
                Request *q = *it;
                const long long w2 = width + q->nx + 2 * q->npad;
                if (q->kind == kGrad2d && q->mergeable && same_schedule(*r, *q) &&
                    w2 * (r->nz + 2 * r->npad) < (1LL << 30)) {
                    grp.push_back(q);
                    const long long b = merged_bytes(grp, s->batch);
                    if (b <= room && (s->inflight == 0 || s->inflight + b <= room)) {
                        gbytes = b;
                        width = w2;
                        it = s->queue.erase(it);
                        continue;
                    }
                    grp.pop_back();
                }
                ++it;
            }
        }
        s->inflight += gbytes;
        s->peak = std::max(s->peak, s->inflight + s->resident);
        std::vector<Model> ms;
        for (Request *q : grp) {
            q->t_start = Clock::now();
            ms.push_back(s->models.at(q->model));
        }
        const int wave = s->batch;
        if (ctx) ctx->snap_fmt = s->snap_fmt;  // this worker's context only
        lk.unlock();
        std::vector<int> rcs(grp.size(), crc);
        std::vector<std::string> errs(grp.size());
        if (crc == 0) {
            // With batching on, a mergeable 2D request takes the batched
            // driver's two-level kernels even when no other request joins it
            // (10-27% less time than the one-shot driver's per-level
            // launches for one shot on 261x877 to 512x1024 grids, the same
            // bits); the rest, and everything with batching off, run alone.
            bool alone = grp.size() == 1 && !(wave > 1 && r->kind == kGrad2d && r->mergeable);
            if (!alone) {
                std::vector<vti::Grad2dPart> parts;
                for (size_t i = 0; i < grp.size(); ++i) parts.push_back(part_of(*grp[i], ms[i]));
                alone = vti::grad2d_merged(ctx, parts.data(), (int)parts.size(), r->wavelet, r->nz, r->nt, r->npad,
                                           r->dx, r->dz, r->dt, r->rate, wave, ctx->snap_fmt) != 0;
                if (!alone) std::fill(rcs.begin(), rcs.end(), 0);
            }
            if (alone)
                for (size_t i = 0; i < grp.size(); ++i) {
                    rcs[i] = run(ctx, ms[i], *grp[i]);
                    if (rcs[i]) errs[i] = vti_last_error();
                }
        } else {
            std::fill(errs.begin(), errs.end(), std::string(vti_last_error()));
        }
        lk.lock();
        const Clock::time_point t_done = Clock::now();
        for (size_t i = 0; i < grp.size(); ++i) {
            Request *q = grp[i];
            q->status = rcs[i];
            q->error = errs[i];
            q->t_done = t_done;
            q->done = true;
            s->models[q->model].users -= 1;
            s->latencies.push_back(std::chrono::duration<double>(q->t_done - q->t_submit).count());
        }
        if (grp.size() > 1) {
            s->groups += 1;
            s->grouped += (long long)grp.size();
        }
        s->inflight -= gbytes;
        s->running -= 1;
        if (excl) s->excl_running = false;
        s->last_done = std::max(s->last_done, t_done);
        s->done_cv.notify_all();
        s->work_cv.notify_all();
    }
    if (ctx) (void)vti_ctx_destroy(ctx);
}

int submit(vti_server *s, std::unique_ptr<Request> r, long long *id)
{
    std::lock_guard<std::mutex> lk(s->mu);
    VTI_CHECK(!s->stop, "server is shutting down");
    auto it = s->models.find(r->model);
    VTI_CHECK(it != s->models.end(), "no resident model with handle %d", r->model);
    VTI_CHECK(r->bytes <= s->budget - s->resident,
              "request needs about %lld bytes of device memory; the budget leaves %lld after the resident models",
              r->bytes, s->budget - s->resident);
    it->second.users += 1;
    r->id = s->next_id++;
    r->t_submit = Clock::now();
    if (!s->any) {
        s->first_submit = r->t_submit;
        s->last_done = r->t_submit;
        s->any = true;
    }
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->work_cv.notify_all();
    return vti::kOk;
}

}  // namespace

extern "C" {

int vti_serve_create(vti_server **out, int device, int slots, double budget_bytes)
{
    VTI_CHECK(out, "vti_serve_create: null output pointer");
    VTI_CHECK(slots >= 1 && slots <= 16, "slots must be in [1, 16], got %d", slots);
    vti_ctx *up = nullptr;
    int rc = vti_ctx_create(&up, device);
    if (rc) return rc;
    const size_t mfree = vti::device_free_bytes();  // vti_ctx_create made `device` current
    auto *s = new vti_server();
    s->device = device;
    s->budget = budget_bytes > 0 ? (long long)budget_bytes : (long long)mfree - (2LL << 30);
    s->upload_ctx = up;
    for (int k = 0; k < slots; ++k) s->workers.emplace_back(worker, s);
    *out = s;
    return vti::kOk;
}

int vti_serve_destroy(vti_server *s)
{
    if (!s) return vti::kOk;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->work_cv.notify_all();
    }
    for (auto &t : s->workers) t.join();
    (void)hipSetDevice(s->device);
    for (auto &kv : s->models)
        for (int k = 0; k < 3; ++k) vti::dfree(kv.second.c[k], s->upload_ctx->stream);
    (void)vti_ctx_destroy(s->upload_ctx);
    delete s;
    return vti::kOk;
}

int vti_serve_put_model(vti_server *s, const float *c11, const float *c13, const float *c33, long long n, int *handle)
{
    VTI_CHECK(s && c11 && c13 && c33 && handle && n > 0, "vti_serve_put_model: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    VTI_CHECK(s->resident + s->inflight + 12 * n <= s->budget, "model of %lld points does not fit the budget", n);
    VTI_HIP(hipSetDevice(s->device));
    Model m;
    m.n = n;
    const float *src[3] = {c11, c13, c33};
    hipStream_t st = s->upload_ctx->stream;
    for (int k = 0; k < 3; ++k) {
        int rc = vti::dmalloc((void **)&m.c[k], 4 * n, st);
        if (!rc) {
            const hipError_t e = hipMemcpyAsync(m.c[k], src[k], 4 * n, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) {
                vti::set_error("model upload: %s", hipGetErrorString(e));
                rc = vti::kErrHip;
            }
        }
        if (rc) {
            for (int j = 0; j <= k; ++j) vti::dfree(m.c[j], st);
            return rc;
        }
    }
    VTI_HIP(hipStreamSynchronize(st));
    s->resident += 12 * n;
    *handle = s->next_model++;
    s->models[*handle] = m;
    return vti::kOk;
}

int vti_serve_drop_model(vti_server *s, int handle)
{
    VTI_CHECK(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->models.find(handle);
    VTI_CHECK(it != s->models.end(), "no resident model with handle %d", handle);
    VTI_CHECK(it->second.users == 0, "model %d still has %d requests queued or running", handle, it->second.users);
    VTI_HIP(hipSetDevice(s->device));
    for (int k = 0; k < 3; ++k) vti::dfree(it->second.c[k], s->upload_ctx->stream);
    VTI_HIP(hipStreamSynchronize(s->upload_ctx->stream));
    s->resident -= 12 * it->second.n;
    s->models.erase(it);
    return vti::kOk;
}

int vti_serve_grad2d(vti_server *s, long long *id, int model, float *objfunc, float *gc11, float *gc13, float *gc33,
                     const float *data, const float *wavelet, const int *sloc, int ns, const int *rloc, int nr,
                     const float *taper, int nx, int nz, int nt, int npad, float dx, float dz, float dt, float rate)
{
    VTI_CHECK(s && id && objfunc && gc11 && gc13 && gc33 && data && wavelet && sloc && rloc && taper,
              "vti_serve_grad2d: null argument");
    VTI_CHECK(nx > 0 && nz > 0 && nt >= 3 && npad >= 0 && ns >= 1 && nr >= 0 && dt > 0 && rate > 0,
              "vti_serve_grad2d: bad sizes");
    auto r = std::make_unique<Request>();
    r->kind = kGrad2d;
    r->model = model;
    r->objfunc = objfunc;
    r->g[0] = gc11;
    r->g[1] = gc13;
    r->g[2] = gc33;
    r->data = data;
    r->wavelet = wavelet;
    r->sloc = sloc;
    r->ns = ns;
    r->rloc = rloc;
    r->nr = nr;
    r->taper = taper;
    r->nx = nx;
    r->nz = nz;
    r->nt = nt;
    r->npad = npad;
    r->dx = dx;
    r->dz = dz;
    r->dt = dt;
    r->rate = rate;
    r->bytes = est_grad2d(*r);
    r->mergeable = mergeable_grad2d(*r);
    // a mergeable request runs through the batched driver even alone
    if (r->mergeable) r->bytes = std::max(r->bytes, merged_bytes({r.get()}, 1));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->models.find(model);
        VTI_CHECK(it == s->models.end() || it->second.n == (long long)(nx + 2 * npad) * (nz + 2 * npad),
                  "model %d holds %lld points, the request's grid %d", model, it->second.n,
                  (nx + 2 * npad) * (nz + 2 * npad));
    }
    return submit(s, std::move(r), id);
}

int vti_serve_model3d(vti_server *s, long long *id, int model, float *ddata, const float *souloc, int ns,
                      const int *shotid, int nshot, const float *recloc, int nr, const float *wavelet, int nx, int ny,
                      int nz, int nt, int npad, float ox, float oy, float oz, float ot, float dx, float dy, float dz,
                      float dt, float samplingRate)
{
    VTI_CHECK(s && id && ddata && souloc && shotid && recloc && wavelet, "vti_serve_model3d: null argument");
    VTI_CHECK(nx > 0 && ny > 0 && nz > 0 && nt >= 3 && nshot >= 1 && nr >= 0 && dt > 0 && samplingRate > 0,
              "vti_serve_model3d: bad sizes");
    auto r = std::make_unique<Request>();
    r->kind = kModel3d;
    r->model = model;
    r->ddata = ddata;
    r->souloc = souloc;
    r->nsou = ns;
    r->shotid = shotid;
    r->nshot = nshot;
    r->recloc = recloc;
    r->nr = nr;
    r->wavelet = wavelet;
    r->nx = nx;
    r->ny = ny;
    r->nz = nz;
    r->nt = nt;
    r->npad = npad;
    r->ox = ox;
    r->oy = oy;
    r->oz = oz;
    r->ot = ot;
    r->dx = dx;
    r->dy = dy;
    r->dz = dz;
    r->dt = dt;
    r->rate = samplingRate;
    r->bytes = est_model3d(*r);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->models.find(model);
        VTI_CHECK(it == s->models.end() || it->second.n == (long long)nx * ny * nz,
                  "model %d holds %lld points, the request's grid %lld", model, it->second.n,
                  (long long)nx * ny * nz);
    }
    return submit(s, std::move(r), id);
}

int vti_serve_wait(vti_server *s, long long id, double *latency_s, double *service_s)
{
    VTI_CHECK(s, "null server");
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        VTI_CHECK(it != s->reqs.end(), "no request %lld (already waited for, or never submitted)", id);
        Request *rp = it->second.get();
        s->done_cv.wait(lk, [&] { return rp->done; });
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (service_s) *service_s = std::chrono::duration<double>(r->t_done - r->t_start).count();
    if (r->status) vti::set_error("request %lld: %s", id, r->error.c_str());
    return r->status;
}

int vti_serve_set_batching(vti_server *s, int max_requests)
{
    VTI_CHECK(s && max_requests >= 1 && max_requests <= 64, "vti_serve_set_batching: max_requests in [1, 64]");
    std::lock_guard<std::mutex> lk(s->mu);
    s->batch = max_requests;
    return vti::kOk;
}

int vti_serve_set_snapshots(vti_server *s, int snap_fmt)
{
    VTI_CHECK(s && snap_fmt >= 0 && snap_fmt <= 3,
              "vti_serve_set_snapshots: format must be 0 (FP32), 1 (FP16), 2 (BF16) or 3 (S16)");
    std::lock_guard<std::mutex> lk(s->mu);
    s->snap_fmt = snap_fmt;
    return vti::kOk;
}

int vti_serve_set_exclusive3d(vti_server *s, int on)
{
    VTI_CHECK(s && (on == 0 || on == 1), "vti_serve_set_exclusive3d: on must be 0 or 1");
    std::lock_guard<std::mutex> lk(s->mu);
    s->exclusive3d = on != 0;
    s->work_cv.notify_all();
    return vti::kOk;
}

int vti_serve_batch_stats(vti_server *s, long long *merged_runs, long long *merged_requests)
{
    VTI_CHECK(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (merged_runs) *merged_runs = s->groups;
    if (merged_requests) *merged_requests = s->grouped;
    return vti::kOk;
}

int vti_serve_stats(vti_server *s, double *p50, double *p99, double *req_per_s, long long *completed,
                    double *peak_bytes)
{
    VTI_CHECK(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        // nearest rank
        const size_t k = (size_t)std::min<double>(l.size() - 1, std::max(0.0, std::ceil(q * l.size()) - 1));
        return l[k];
    };
    if (p50) *p50 = pct(0.50);
    if (p99) *p99 = pct(0.99);
    const double span = s->any ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0.0;
    if (req_per_s) *req_per_s = span > 0 ? l.size() / span : 0.0;
    if (completed) *completed = (long long)l.size();
    if (peak_bytes) *peak_bytes = (double)s->peak;
    return vti::kOk;
}

}  // extern "C"
