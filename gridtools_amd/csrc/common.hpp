// Shared helpers of libgt355: error reporting, field checks, launch utilities.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gt355.h"

namespace gt355 {

    // Thread-local last error, read by gt355_last_error().
    void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
    const char *last_error();

    // The device selected by gt355_init (or the current HIP device if gt355_init was never called).
    int current_device();

    // Number of XCDs the launch is spread over (workgroups are dispatched round-robin over XCDs).
    int xcd_count();

    // Number of compute units of the current device (cached on first use).
    int cu_count();

    inline int hip_fail(hipError_t e, const char *what) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return GT355_EHIP;
    }

#define GT355_HIP(call)                                                   \
    do {                                                                  \
        hipError_t gt355_e_ = (call);                                     \
        if (gt355_e_ != hipSuccess)                                       \
            return ::gt355::hip_fail(gt355_e_, #call);                    \
    } while (0)

#define GT355_CHECK(cond, ...)                                            \
    do {                                                                  \
        if (!(cond)) {                                                    \
            ::gt355::set_error(__VA_ARGS__);                              \
            return GT355_EINVAL;                                          \
        }                                                                 \
    } while (0)

    // Checks that field f is usable for offsets [ilo, ihi) x [jlo, jhi) x [klo, khi) from its origin.
    int check_field(const gt355_field *f, const char *name, int dtype, int ilo, int ihi, int jlo, int jhi, int klo,
        int khi);

    // Debug launch mode (GT355_DEBUG_SYNC=1 or gt355_debug_sync): a device synchronisation after every launch.
    bool debug_sync_enabled();
    int debug_sync_after(const char *what);
    // Marks the calling thread as capturing a graph (gt355_graph_begin .. gt355_graph_end).
    void set_capturing(bool on);
    // The deep-column scratch buffer of (current device, stream), grown to at least `bytes` (pools.hip).
    int scratch_for(hipStream_t stream, size_t bytes, void **out);
    // Frees the pools and retired buffers (gt355_finalize).
    void release_pools();
    // Resolves, on the current device, the device address of gt355_vadv_noalloc's static scratch (vadv.hip).
    int noalloc_setup();

    // The deferred-run scope (defer.hip).  defer_active: a scope records the calling thread's calls on `stream`;
    // defer_hdiff / defer_vadv record a checked call; defer_flush runs what the scope on `stream` recorded (every
    // other entry point that launches, copies or waits on a stream calls it first, keeping the recorded order);
    // defer_flush_all does so for every scope of the thread (before device-wide waits and blocking copies);
    // release_deferred frees the kept batches (gt355_finalize).
    bool defer_active(void *stream);
    int defer_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
        int nj, int nk, void *stream);
    int defer_vadv(int precision, const gt355_field *utens_stage, const gt355_field *u_stage, const gt355_field *wcon,
        const gt355_field *u_pos, const gt355_field *utens, double dtr, int ni, int nj, int nk, void *stream);
    int defer_flush(void *stream);
    int defer_flush_all();
    void release_deferred();
#define GT355_DEFER_FLUSH(stream)                                                                                \
    do {                                                                                                         \
        if (int rc_defer_ = ::gt355::defer_flush(stream))                                                        \
            return rc_defer_;                                                                                    \
    } while (0)

    // After a launch: report a launch-configuration error as a status; in the debug launch mode also wait for the
    // kernel and report its failure here.
    inline int check_launch(const char *what) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return hip_fail(e, what);
        if (debug_sync_enabled())
            return debug_sync_after(what);
        return GT355_OK;
    }

    // Maps a linear workgroup id so that consecutive logical ids run on the same XCD.  The hardware hands
    // workgroup n to XCD n % nxcd; logical id = xcd * per + n / nxcd keeps neighbouring tiles (which share
    // halo rows) in one L2.  Ids beyond the largest multiple of nxcd keep their natural position.
    __device__ __forceinline__ int xcd_remap(int n, int total, int nxcd) {
        int per = total / nxcd;
        int full = per * nxcd;
        if (n >= full || nxcd <= 1)
            return n;
        return (n % nxcd) * per + n / nxcd;
    }

    // Buffer-resource access: a field is a buffer descriptor (base in scalar registers), a level is a uniform
    // 32-bit byte offset (soffset) and a lane's column a 32-bit byte offset (voffset).  One level offset serves
    // every field of one layout, so a load costs no per-lane or per-field address arithmetic.  Offsets are byte
    // offsets below 4 GiB (launchers check the field span).  Data format 32 (word 3 = 0x00020000) enables
    // untyped dword access; num_records 0xFFFFFFFF disables the range clamp.
    using rsrc_t = __amdgpu_buffer_rsrc_t;
    // An address the caller knows to be the same on every lane, moved to scalar registers (a descriptor built from
    // an address the compiler cannot prove uniform would be used through a loop over the lanes' values).
    __device__ __forceinline__ const void *wave_uniform(const void *p) {
        const uint64_t v = (uint64_t)(uintptr_t)p;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        return (const void *)(uintptr_t)(((uint64_t)hi << 32) | lo);
    }
    __device__ __forceinline__ rsrc_t make_rsrc(const void *base) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xFFFFFFFF, 0x00020000);
    }
    // Cache policy of a buffer access (the instruction's cache-policy bits on gfx950): CPOL_NT marks a stream
    // that is touched once -- non-temporal, so the L2 and the Infinity Cache do not keep it at the expense of
    // lines that are re-read.  Measured on vadv FP64 256x256x80 (five input streams and one output stream):
    // 54 us with the default policy, 41 us with non-temporal loads and stores.
    constexpr int CPOL_NT = 2;
    // Cache policy of the output stores: non-temporal, like the input streams.  System-scope stores (sc0 sc1,
    // written through the L2) shorten a plain 16-byte-per-lane FP32 copy of 42 MB from 7.95 to 7.27 us, but were
    // measured slower on the stencils (hdiff FP64 256x256x80 28.5 us against 22.6 us, vadv FP64 41.7 against 40.2 us).
    constexpr int CPOL_OUT = CPOL_NT;
    template <class T, int CPOL = 0>
    __device__ __forceinline__ T bload(rsrc_t r, uint32_t voff, uint32_t soff) {
        if constexpr (sizeof(T) == 4)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, CPOL));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CPOL));
    }
    template <class T>
    struct two {
        T a, b;
    };
    // two consecutive elements (columns i and i+1) in one load
    template <class T, int CPOL = 0>
    __device__ __forceinline__ two<T> bload2(rsrc_t r, uint32_t voff, uint32_t soff) {
        if constexpr (sizeof(T) == 4)
            return __builtin_bit_cast(two<T>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CPOL));
        else
            return __builtin_bit_cast(two<T>, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, CPOL));
    }
    // two consecutive elements (columns i and i+1) in one store
    template <class T, int CPOL = 0>
    __device__ __forceinline__ void bstore_pair(rsrc_t r, T a, T b, uint32_t voff, uint32_t soff) {
        const two<T> v{a, b};
        if constexpr (sizeof(T) == 4)
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff, CPOL);
        else
            __builtin_amdgcn_raw_buffer_store_b128(
                __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0)), v), r, voff, soff,
                CPOL);
    }
    // 16 bytes per lane (four FP32 elements, i to i+3) in one request
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    template <int CPOL = 0>
    __device__ __forceinline__ f32x4 bload4(rsrc_t r, uint32_t voff, uint32_t soff) {
        return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, CPOL));
    }
    template <int CPOL = 0>
    __device__ __forceinline__ void bstore4(rsrc_t r, f32x4 v, uint32_t voff, uint32_t soff) {
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0)), v), r, voff, soff, CPOL);
    }
    template <class T, int CPOL = 0>
    __device__ __forceinline__ void bstore(rsrc_t r, T v, uint32_t voff, uint32_t soff) {
        if constexpr (sizeof(T) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, CPOL);
        else
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff,
                CPOL);
    }

} // namespace gt355
