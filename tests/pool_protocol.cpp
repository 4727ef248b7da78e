// CPU test of the large-block reuse protocol (csrc/vti_pool.h) against a
// mocked stream / event layer (test infrastructure; tests/test_pool_protocol.py
// builds and runs it).
//
// The mock keeps, per stream, the ordered list of operations submitted to it
// (uses of a block, event records, waits on an event) and how many of them
// have completed.  Streams progress independently, in random steps; a wait
// completes only after the record it captured.  It checks:
//   - every use of a block by its k-th owner runs after every use by owners
//     1..k-1 (no two users of one block overlap);
//   - an event is destroyed only after its last record completed and while no
//     wait on it is pending;
//   - a block is returned to the runtime only when no use of it is pending;
//   - a double release is reported and does not cache a block twice.
// `./pool_protocol old` runs the protocol the library used before (the event
// of a reused block destroyed right after the stream wait); the checker must
// then report a violation.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "vti_pool.h"

namespace {

int g_violations = 0;
#define VIOLATION(...)                      \
    do {                                    \
        ++g_violations;                     \
        if (g_violations <= 5) {            \
            std::fprintf(stderr, __VA_ARGS__); \
            std::fprintf(stderr, "\n");     \
        }                                   \
    } while (0)

struct Op {
    enum Kind { kUse, kRecord, kWait } kind;
    long block = 0;   // kUse
    int epoch = 0;    // kUse: owner number of the block
    int event = 0;    // kRecord / kWait
    int rs = -1;      // kWait: stream and index of the record it captured
    long ri = -1;
};

struct Sim {
    std::vector<std::vector<Op>> q;
    std::vector<long> done;
    struct Ev {
        int rs = -1;
        long ri = -1;
        bool destroyed = false;
    };
    std::unordered_map<int, Ev> ev;
    int next_ev = 1;
    long next_block = 1;
    std::unordered_map<long, int> epoch;                    // current owner number per block
    std::unordered_map<long, std::vector<int>> pending_uses;  // epochs of not-yet-run uses
    std::unordered_map<long, int> ran_max_epoch;
    std::unordered_map<long, bool> freed;

    explicit Sim(int ns) : q(ns), done(ns, 0) {}

    bool record_done(int s, long i) const { return s < 0 || done[s] > i; }

    // complete the next op of stream s if it can run; returns false otherwise
    bool step(int s)
    {
        if (done[s] >= (long)q[s].size()) return false;
        const Op &o = q[s][done[s]];
        if (o.kind == Op::kWait && !record_done(o.rs, o.ri)) return false;
        if (o.kind == Op::kUse) {
            auto &pu = pending_uses[o.block];
            for (int e : pu)
                if (e < o.epoch) VIOLATION("block %ld: owner %d runs before a use of owner %d", o.block, o.epoch, e);
            for (size_t k = 0; k < pu.size(); ++k)
                if (pu[k] == o.epoch) {
                    pu.erase(pu.begin() + (long)k);
                    break;
                }
            if (freed[o.block]) VIOLATION("block %ld used after it went back to the runtime", o.block);
        }
        ++done[s];
        return true;
    }
    void drain(int s)
    {
        // run s to its end, running other streams as its waits require
        for (int guard = 0; done[s] < (long)q[s].size() && guard < 1000000; ++guard) {
            if (step(s)) continue;
            const Op &o = q[s][done[s]];
            if (!step(o.rs)) drain(o.rs);
        }
    }
    void use(int s, long b) { q[s].push_back(Op{Op::kUse, b, epoch[b], 0, -1, -1}), pending_uses[b].push_back(epoch[b]); }
};

Sim *g_sim = nullptr;

struct MockRT {
    using Stream = int;
    using Event = int;
    static bool malloc(void **p, size_t)
    {
        const long b = g_sim->next_block++;
        g_sim->epoch[b] = 0;
        *p = reinterpret_cast<void *>(b << 8);
        return true;
    }
    static void free(void *p)
    {
        const long b = reinterpret_cast<long>(p) >> 8;
        if (!g_sim->pending_uses[b].empty()) VIOLATION("block %ld freed with %zu uses pending", b, g_sim->pending_uses[b].size());
        g_sim->freed[b] = true;
    }
    static bool event_create(Event *e)
    {
        *e = g_sim->next_ev++;
        g_sim->ev[*e] = Sim::Ev{};
        return true;
    }
    static bool event_record(Event e, Stream s)
    {
        auto &E = g_sim->ev[e];
        if (E.destroyed) VIOLATION("record of destroyed event %d", e);
        g_sim->q[s].push_back(Op{Op::kRecord, 0, 0, e, -1, -1});
        E.rs = s;
        E.ri = (long)g_sim->q[s].size() - 1;
        return true;
    }
    static void event_sync(Event e)
    {
        auto &E = g_sim->ev[e];
        if (E.rs >= 0)
            while (!g_sim->record_done(E.rs, E.ri)) {
                if (!g_sim->step(E.rs)) g_sim->drain(E.rs);
            }
    }
    static void event_destroy(Event e)
    {
        auto &E = g_sim->ev[e];
        if (!g_sim->record_done(E.rs, E.ri)) VIOLATION("event %d destroyed before its record completed", e);
        for (size_t s = 0; s < g_sim->q.size(); ++s)
            for (long i = g_sim->done[s]; i < (long)g_sim->q[s].size(); ++i)
                if (g_sim->q[s][i].kind == Op::kWait && g_sim->q[s][i].event == e)
                    VIOLATION("event %d destroyed while stream %zu still waits on it", e, s);
        E.destroyed = true;
    }
    static void stream_wait(Stream s, Event e)
    {
        auto &E = g_sim->ev[e];
        if (E.destroyed) VIOLATION("wait on destroyed event %d", e);
        g_sim->q[s].push_back(Op{Op::kWait, 0, 0, e, E.rs, E.ri});
    }
    static void stream_sync(Stream s) { g_sim->drain(s); }
};

// The protocol of the previous allocator: a fresh event per release,
// destroyed right after the taking stream was made to wait on it.
class OldPool {
public:
    int take(void **p, size_t bytes, int, int s)
    {
        for (size_t k = 0; k < cached_.size(); ++k)
            if (cached_[k].bytes >= bytes && cached_[k].bytes / 2 <= bytes) {
                MockRT::stream_wait(s, cached_[k].ev);
                MockRT::event_destroy(cached_[k].ev);
                *p = cached_[k].p;
                cached_.erase(cached_.begin() + (long)k);
                return 0;
            }
        MockRT::malloc(p, bytes);
        sizes_[*p] = bytes;
        return 1;
    }
    bool give(void *p, int s, bool *misuse)
    {
        *misuse = false;
        C c{p, sizes_[p], 0};
        MockRT::event_create(&c.ev);
        MockRT::event_record(c.ev, s);
        cached_.push_back(c);
        return true;
    }
    void flush(int)
    {
        for (auto &c : cached_) {
            MockRT::event_sync(c.ev);
            MockRT::event_destroy(c.ev);
            MockRT::free(c.p);
        }
        cached_.clear();
    }

private:
    struct C {
        void *p;
        size_t bytes;
        int ev;
    };
    std::vector<C> cached_;
    std::unordered_map<void *, size_t> sizes_;
};

template <class Pool>
void scenario(Pool &pool, unsigned seed, int nstreams, int iters)
{
    std::mt19937 rng(seed);
    struct Held {
        void *p;
        int s;
    };
    std::vector<Held> held;
    const size_t sizes[3] = {size_t(64) << 20, size_t(100) << 20, size_t(256) << 20};
    for (int it = 0; it < iters; ++it) {
        const int r = (int)(rng() % 10);
        const int s = (int)(rng() % nstreams);
        if (r < 4) {  // allocate and use on s
            void *p = nullptr;
            pool.take(&p, sizes[rng() % 3], 0, s);
            const long b = reinterpret_cast<long>(p) >> 8;
            g_sim->epoch[b] += 1;
            const int nuse = 1 + (int)(rng() % 3);
            for (int k = 0; k < nuse; ++k) g_sim->use(s, b);
            held.push_back({p, s});
        } else if (r < 7 && !held.empty()) {  // release on the stream that used it
            const size_t k = rng() % held.size();
            bool misuse = false;
            pool.give(held[k].p, held[k].s, &misuse);
            if (misuse) VIOLATION("a single release reported as a double release");
            held.erase(held.begin() + (long)k);
        } else if (r < 9) {  // let a stream run a little
            const int n = (int)(rng() % 4);
            for (int k = 0; k < n; ++k) g_sim->step(s);
        } else if (rng() % 8 == 0) {
            pool.flush(0);
        }
    }
    for (auto &h : held) {
        bool misuse = false;
        pool.give(h.p, h.s, &misuse);
    }
    for (int s = 0; s < nstreams; ++s) g_sim->drain(s);
    pool.flush(0);
}

}  // namespace

int main(int argc, char **argv)
{
    const bool old = argc > 1 && std::string(argv[1]) == "old";
    long runs = 0;
    for (unsigned seed = 1; seed <= 200; ++seed) {
        Sim sim(4);
        g_sim = &sim;
        if (old) {
            OldPool pool;
            scenario(pool, seed, 4, 400);
        } else {
            vti::BigPool<MockRT> pool;
            scenario(pool, seed, 4, 400);
            // a double release is refused and reported
            void *p = nullptr;
            pool.take(&p, size_t(64) << 20, 0, 0);
            bool misuse = false;
            pool.give(p, 0, &misuse);
            if (misuse) VIOLATION("first release reported as misuse");
            pool.give(p, 1, &misuse);
            if (!misuse) VIOLATION("double release not reported");
            void *a = nullptr, *b = nullptr;
            pool.take(&a, size_t(64) << 20, 0, 2);
            pool.take(&b, size_t(64) << 20, 0, 3);
            if (a == b) VIOLATION("one cached block handed to two users");
            pool.give(a, 2, &misuse);
            pool.give(b, 3, &misuse);
            for (int s = 0; s < 4; ++s) sim.drain(s);
            pool.flush(0);
        }
        ++runs;
    }
    std::printf("runs %ld violations %d\n", runs, g_violations);
    return g_violations ? 1 : 0;
}
