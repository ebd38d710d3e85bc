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
#include <cstdlib>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace gt355 {
    namespace {
        struct tmp_block {
            void *p = nullptr;
            size_t bytes = 0;
            int dev = 0;
            hipStream_t stream = nullptr;
            hipEvent_t ev = nullptr;
            bool ev_valid = false;
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
        for (auto &b : g_tmp_free) {
            (void)hipFree(b.p);
            if (b.ev)
                (void)hipEventDestroy(b.ev);
        }
        g_tmp_free.clear();
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
            if (!b.ev_valid)
                continue;
            hipError_t q = hipEventQuery(b.ev);
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

int gt355_tmp_release(void *ptr, void *stream) {
    if (!ptr)
        return GT355_OK;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_tmp_mutex);
    for (size_t n = 0; n < g_tmp_used.size(); ++n) {
        if (g_tmp_used[n].p != ptr)
            continue;
        tmp_block b = g_tmp_used[n];
        g_tmp_used.erase(g_tmp_used.begin() + (long)n);
        b.stream = s;
        b.ev_valid = false;
        if (!stream_capturing(s)) {
            if (!b.ev)
                GT355_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
            GT355_HIP(hipEventRecord(b.ev, s));
            b.ev_valid = true;
        }
        g_tmp_free.push_back(b);
        return GT355_OK;
    }
    set_error("gt355_tmp_release: %p was not acquired with gt355_tmp_acquire", ptr);
    return GT355_EINVAL;
}

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
