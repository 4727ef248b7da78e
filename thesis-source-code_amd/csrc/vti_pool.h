// Reuse cache for large device blocks, written against a runtime policy so
// that the same code runs on HIP (vti_common.cpp) and on a mocked stream /
// event layer in the CPU tests (tests/test_pool_protocol.py).
//
// Why: a fresh multi-gigabyte hipMalloc can stall for hundreds of
// milliseconds once earlier frees have piled up (freed memory is cleared
// before it is handed out again), which made repeated driver calls three to
// seven times slower.  Large blocks are therefore kept by the library.
//
// Protocol (every block owns ONE event for its whole life):
//  - take(bytes, stream): the smallest cached block of the device that holds
//    `bytes` without wasting more than half of itself; the taking stream is
//    made to wait on the block's event (recorded when it was last released),
//    so its new user runs after every use ordered before that release.
//    Otherwise a fresh block (and its event) from the runtime.
//  - give(p, stream): records the block's event on `stream` and caches it.
//    The caller guarantees that every use of the block, on any stream, is
//    ordered before the work already submitted to `stream` (the library uses
//    each block on the stream that frees it; the request server releases a
//    resident model only after the requests that read it have completed).
//  - flush(dev): waits for each cached block's event, then returns block and
//    event to the runtime.  An event is destroyed only there, after it has
//    completed, so no pending stream wait can refer to a destroyed event, and
//    recording it again on a later release is ordered after every earlier
//    wait on it (a wait captures the record current when it is issued).
//  - Ownership is checked: releasing a block that is not live (double free,
//    a foreign pointer) is refused, so a block can never be cached twice or
//    handed to two users.
#pragma once

#include <cstddef>
#include <mutex>
#include <unordered_map>

namespace vti {

template <class RT>
class BigPool {
public:
    using Stream = typename RT::Stream;
    using Event = typename RT::Event;

    // 0 = taken from the cache, 1 = fresh from the runtime, -1 = runtime out of memory
    int take(void **p, size_t bytes, int dev, Stream s)
    {
        std::lock_guard<std::mutex> lk(mu_);
        Block *best = nullptr;
        for (auto &kv : blocks_) {
            Block &b = kv.second;
            if (b.live || b.dev != dev || b.bytes < bytes || b.bytes / 2 > bytes) continue;
            if (!best || b.bytes < best->bytes) best = &b;
        }
        if (best) {
            if (best->recorded) RT::stream_wait(s, best->ev);
            best->live = true;
            *p = best->p;
            return 0;
        }
        void *q = nullptr;
        if (!RT::malloc(&q, bytes)) {
            flush_locked(dev);  // give the cached blocks back and try once more
            if (!RT::malloc(&q, bytes)) return -1;
        }
        Block b;
        b.p = q;
        b.bytes = bytes;
        b.dev = dev;
        b.live = true;
        b.recorded = false;
        if (!RT::event_create(&b.ev)) {
            RT::free(q);
            return -1;
        }
        blocks_[q] = b;
        *p = q;
        return 1;
    }

    // true when p is a live block of this pool (now cached); false when p is
    // not one (the caller frees it otherwise) -- *misuse is set when p is a
    // block of this pool that is already cached (a double release)
    bool give(void *p, Stream s, bool *misuse)
    {
        std::lock_guard<std::mutex> lk(mu_);
        *misuse = false;
        auto it = blocks_.find(p);
        if (it == blocks_.end()) return false;
        Block &b = it->second;
        if (!b.live) {
            *misuse = true;
            return true;
        }
        if (RT::event_record(b.ev, s)) {
            b.recorded = true;
        } else {
            // no event: make the block safe to hand out by waiting here
            RT::stream_sync(s);
            b.recorded = false;
        }
        b.live = false;
        return true;
    }

    bool owns(void *p)
    {
        std::lock_guard<std::mutex> lk(mu_);
        return blocks_.count(p) != 0;
    }

    size_t cached_bytes(int dev)
    {
        std::lock_guard<std::mutex> lk(mu_);
        size_t n = 0;
        for (auto &kv : blocks_)
            if (!kv.second.live && kv.second.dev == dev) n += kv.second.bytes;
        return n;
    }

    void flush(int dev)
    {
        std::lock_guard<std::mutex> lk(mu_);
        flush_locked(dev);
    }

private:
    struct Block {
        void *p;
        size_t bytes;
        int dev;
        bool live;      // handed out (not in the cache)
        bool recorded;  // ev holds the last release
        Event ev;
    };

    void flush_locked(int dev)
    {
        for (auto it = blocks_.begin(); it != blocks_.end();) {
            Block &b = it->second;
            if (!b.live && b.dev == dev) {
                if (b.recorded) RT::event_sync(b.ev);
                RT::event_destroy(b.ev);
                RT::free(b.p);
                it = blocks_.erase(it);
            } else {
                ++it;
            }
        }
    }

    std::mutex mu_;
    std::unordered_map<void *, Block> blocks_;
};

}  // namespace vti
