// Device memory the library keeps between calls, the debug launch mode and the thread's current stream.
//
// * Temporaries (gt355_tmp_acquire / gt355_tmp_release): the pooled equivalent of the reference's cached allocator
//   (sid/allocator.hpp:44-57), which keeps freed blocks in a static stash for reuse.  The reference launches on one
//   stream, so stream order alone protects a reused block; here a block remembers the stream it was released on.
//   It is handed out again at once on that stream (the next user is ordered after the previous one), and on another
//   stream only after an event recorded at the release has completed.  Blocks released during graph capture carry
//   no event and are reused on their own stream only.
// * Scratch of the deep-column vertical advection: one growing buffer per (device, stream).  A buffer that is
//   outgrown is retired, not freed: queued work or a captured graph may still use it.  Retired buffers are freed by
//   gt355_finalize.  Growth during stream capture is refused (a graph must not depend on an allocation made inside
//   it); run the largest shape once outside capture first.
// * Debug launch mode: with GT355_DEBUG_SYNC=1 in the environment (or gt355_debug_sync(1)) every launch is followed
//   by a device synchronisation that reports a failing kernel at its own call, as the reference's debug builds do
//   (common/cuda_util.hpp:89-93).  It is skipped while the calling thread captures a graph through gt355_graph_begin.
#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace gt355 {
    namespace {
        // Events of released blocks.  One record covers every block of one release call, so blocks share their
        // event; an event no free block refers to any more goes back to g_ev_free.  Every event_ref is created,
        // copied and dropped under g_tmp_mutex, which therefore also guards g_ev_free.  The events skip the
        // system-scope fence of a record (a cache writeback and invalidation between kernels): the pool only needs
        // to know that the work before the release has completed on this device, whose next user overwrites the
        // block.  Measured on the generic engine's vadv spec at 256x256x80 (two temporaries per run): 98.6-99.0 us
        // per run with the fence, 94.7 us without.
        std::vector<hipEvent_t> g_ev_free;
        using event_ref = std::shared_ptr<hipEvent_t>;
        event_ref new_event() {
            hipEvent_t e = nullptr;
            if (!g_ev_free.empty()) {
                e = g_ev_free.back();
                g_ev_free.pop_back();
            } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
            return event_ref(new hipEvent_t(e), [](hipEvent_t *q) {
                g_ev_free.push_back(*q);
                delete q;
            });
        }
        struct tmp_block {
            void *p = nullptr;
            size_t bytes = 0;
            int dev = 0;
            hipStream_t stream = nullptr;
            event_ref ev; // recorded on `stream` at the release (none: released during capture)
        };
        std::mutex g_tmp_mutex;
        std::vector<tmp_block> g_tmp_free;
        std::vector<tmp_block> g_tmp_used;

        struct scratch_buf {
            int dev;
            hipStream_t stream;
            void *p;
            size_t bytes;
        };
        std::mutex g_scratch_mutex;
        std::vector<scratch_buf> g_scratch;
        std::vector<void *> g_retired;

        std::atomic<int> g_debug_sync{-1};
        thread_local int t_capturing = 0;
        thread_local void *t_current_stream = nullptr;

        bool stream_capturing(hipStream_t s) {
            hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &st) != hipSuccess) {
                (void)hipGetLastError();
                return t_capturing != 0;
            }
            return st != hipStreamCaptureStatusNone || t_capturing != 0;
        }
    } // namespace

    bool debug_sync_enabled() {
        int v = g_debug_sync.load(std::memory_order_relaxed);
        if (v < 0) {
            const char *e = std::getenv("GT355_DEBUG_SYNC");
            v = (e && std::atoi(e) != 0) ? 1 : 0;
            g_debug_sync.store(v, std::memory_order_relaxed);
        }
        return v != 0;
    }

    int debug_sync_after(const char *what) {
        if (t_capturing)
            return GT355_OK;
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess)
            e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("%s (debug sync after launch): %s", what, hipGetErrorString(e));
            return GT355_EHIP;
        }
        return GT355_OK;
    }

    void set_capturing(bool on) { t_capturing = on ? 1 : 0; }

    int scratch_for(hipStream_t stream, size_t bytes, void **out) {
        const int dev = current_device();
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        scratch_buf *b = nullptr;
        for (auto &s : g_scratch)
            if (s.dev == dev && s.stream == stream)
                b = &s;
        if (b && b->bytes >= bytes) {
            *out = b->p;
            return GT355_OK;
        }
        if (stream_capturing(stream)) {
            set_error("scratch: %zu bytes are needed while the stream is being captured; run the call once with "
                      "its largest shape outside capture first",
                bytes);
            return GT355_EUNSUP;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            set_error("scratch: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
            return GT355_ENOMEM;
        }
        if (b) {
            // queued launches or a captured graph may still use the old buffer: keep it until gt355_finalize
            g_retired.push_back(b->p);
            b->p = p;
            b->bytes = bytes;
        } else {
            g_scratch.push_back(scratch_buf{dev, stream, p, bytes});
        }
        *out = p;
        return GT355_OK;
    }

    void release_pools() {
        {
            std::lock_guard<std::mutex> lock(g_scratch_mutex);
            for (auto &s : g_scratch)
                (void)hipFree(s.p);
            for (void *p : g_retired)
                (void)hipFree(p);
            g_scratch.clear();
            g_retired.clear();
        }
        std::lock_guard<std::mutex> lock(g_tmp_mutex);
        for (auto &b : g_tmp_free)
            (void)hipFree(b.p);
        g_tmp_free.clear(); // drops the free blocks' events into g_ev_free
        for (hipEvent_t e : g_ev_free)
            (void)hipEventDestroy(e);
        g_ev_free.clear();
    }
} // namespace gt355

using namespace gt355;

extern "C" {

int gt355_tmp_acquire(size_t bytes, void *stream, void **ptr) {
    GT355_CHECK(ptr != nullptr, "gt355_tmp_acquire: null output");
    *ptr = nullptr;
    if (bytes == 0)
        return GT355_OK;
    const int dev = current_device();
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_tmp_mutex);
    int best = -1;
    bool best_same = false;
    for (int n = 0; n < (int)g_tmp_free.size(); ++n) {
        tmp_block &b = g_tmp_free[n];
        if (b.dev != dev || b.bytes < bytes || b.bytes > 2 * bytes)
            continue;
        const bool same = b.stream == s;
        if (!same) {
            if (!b.ev)
                continue;
            hipError_t q = hipEventQuery(*b.ev);
            if (q != hipSuccess) {
                if (q != hipErrorNotReady)
                    (void)hipGetLastError();
                continue;
            }
        }
        // prefer a block of this stream, then the smallest
        if (best < 0 || (same && !best_same) || (same == best_same && b.bytes < g_tmp_free[best].bytes)) {
            best = n;
            best_same = same;
        }
    }
    tmp_block blk;
    if (best >= 0) {
        blk = g_tmp_free[best];
        blk.ev.reset();
        g_tmp_free.erase(g_tmp_free.begin() + best);
    } else {
        hipError_t e = hipMalloc(&blk.p, bytes);
        if (e != hipSuccess) {
            set_error("gt355_tmp_acquire(%zu): %s", bytes, hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? GT355_ENOMEM : GT355_EHIP;
        }
        blk.bytes = bytes;
        blk.dev = dev;
    }
    blk.stream = s;
    g_tmp_used.push_back(blk);
    *ptr = blk.p;
    return GT355_OK;
}

int gt355_tmp_release_n(void *const *ptrs, int n, void *stream) {
    GT355_CHECK(n >= 0 && (ptrs != nullptr || n == 0), "gt355_tmp_release_n: null list");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_tmp_mutex);
    // every pointer is checked before any block changes hands
    for (int q = 0; q < n; ++q) {
        if (!ptrs[q])
            continue;
        bool found = false;
        for (const tmp_block &b : g_tmp_used)
            found = found || b.p == ptrs[q];
        for (int r = 0; r < q && found; ++r)
            found = ptrs[r] != ptrs[q];
        if (!found) {
            set_error("gt355_tmp_release: %p was not acquired with gt355_tmp_acquire (or is listed twice)", ptrs[q]);
            return GT355_EINVAL;
        }
    }
    event_ref ev;
    bool any = false;
    for (int q = 0; q < n; ++q)
        any = any || ptrs[q] != nullptr;
    if (any && !stream_capturing(s)) {
        ev = new_event();
        if (!ev) {
            set_error("gt355_tmp_release: hipEventCreateWithFlags failed");
            return GT355_EHIP;
        }
        GT355_HIP(hipEventRecord(*ev, s));
    }
    for (int q = 0; q < n; ++q) {
        if (!ptrs[q])
            continue;
        for (size_t u = 0; u < g_tmp_used.size(); ++u) {
            if (g_tmp_used[u].p != ptrs[q])
                continue;
            tmp_block b = g_tmp_used[u];
            g_tmp_used.erase(g_tmp_used.begin() + (long)u);
            b.stream = s;
            b.ev = ev;
            g_tmp_free.push_back(b);
            break;
        }
    }
    return GT355_OK;
}

int gt355_tmp_release(void *ptr, void *stream) { return gt355_tmp_release_n(&ptr, 1, stream); }

int gt355_debug_sync(int on) {
    g_debug_sync.store(on ? 1 : 0, std::memory_order_relaxed);
    return GT355_OK;
}

int gt355_stream_set_current(void *stream) {
    t_current_stream = stream;
    return GT355_OK;
}

int gt355_stream_current(void **stream) {
    GT355_CHECK(stream != nullptr, "gt355_stream_current: null output");
    *stream = t_current_stream;
    return GT355_OK;
}

} // extern "C"
