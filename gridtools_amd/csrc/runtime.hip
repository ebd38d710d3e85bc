// libgt355 runtime: device selection, errors, storage layout and allocation, copies, synthetic inputs,
// MALL flush.  Replaces the reference's CUDA runtime glue (common/cuda_util.hpp:20-95,
// common/cuda_runtime.hpp) and the gpu storage traits' allocation and mirroring (storage/gpu.hpp:69-105).
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdarg>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace gt355 {
    namespace {
        thread_local char g_err[1024] = "";
        int g_device = -1;
        std::mutex g_mutex;

        struct flush_state {
            void *buf = nullptr;
            size_t bytes = 0;
        };
        flush_state g_flush[64];
    } // namespace

    void set_error(const char *fmt, ...) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(g_err, sizeof g_err, fmt, ap);
        va_end(ap);
    }

    const char *last_error() { return g_err; }

    int current_device() {
        if (g_device >= 0)
            return g_device;
        int d = 0;
        if (hipGetDevice(&d) != hipSuccess)
            return 0;
        return d;
    }

    int cu_count() {
        static int cached = -1;
        if (cached > 0)
            return cached;
        hipDeviceProp_t p;
        int n = 1;
        if (hipGetDeviceProperties(&p, current_device()) == hipSuccess && p.multiProcessorCount > 0)
            n = p.multiProcessorCount;
        cached = n;
        return n;
    }

    int xcd_count() {
        // MI355X (gfx950) and MI300 (gfx942) spread 256/304 CUs over 8 XCDs.
        static int cached = -1;
        if (cached > 0)
            return cached;
        hipDeviceProp_t p;
        int n = 1;
        if (hipGetDeviceProperties(&p, current_device()) == hipSuccess) {
            if (std::strncmp(p.gcnArchName, "gfx950", 6) == 0 || std::strncmp(p.gcnArchName, "gfx942", 6) == 0)
                n = 8;
        }
        cached = n;
        return n;
    }

    int check_field(const gt355_field *f, const char *name, int dtype, int ilo, int ihi, int jlo, int jhi, int klo,
        int khi) {
        GT355_CHECK(f != nullptr, "field %s: null descriptor", name);
        GT355_CHECK(f->origin != nullptr, "field %s: null origin", name);
        GT355_CHECK(f->dtype == dtype, "field %s: dtype %d, expected %d", name, f->dtype, dtype);
        GT355_CHECK(f->stride[0] == 1, "field %s: i stride must be 1 (i-innermost layout), got %lld", name,
            (long long)f->stride[0]);
        GT355_CHECK(f->stride[1] > 0 && f->stride[2] > 0, "field %s: j/k strides must be positive", name);
        GT355_CHECK(f->lower[0] <= ilo && ihi <= f->upper[0],
            "field %s: i offsets [%d,%d) outside bounds [%d,%d)", name, ilo, ihi, f->lower[0], f->upper[0]);
        GT355_CHECK(f->lower[1] <= jlo && jhi <= f->upper[1],
            "field %s: j offsets [%d,%d) outside bounds [%d,%d)", name, jlo, jhi, f->lower[1], f->upper[1]);
        GT355_CHECK(f->lower[2] <= klo && khi <= f->upper[2],
            "field %s: k offsets [%d,%d) outside bounds [%d,%d)", name, klo, khi, f->lower[2], f->upper[2]);
        GT355_CHECK(uintptr_t(f->origin) % (dtype == GT355_F64 ? 8 : 4) == 0, "field %s: misaligned origin", name);
        // The bounds box must not overlap itself: rows and planes are disjoint.
        GT355_CHECK(f->stride[1] >= f->upper[0] - f->lower[0], "field %s: j stride smaller than the i extent", name);
        GT355_CHECK(f->stride[2] >= f->stride[1] * (int64_t)(f->upper[1] - f->lower[1]),
            "field %s: k stride smaller than the j plane", name);
        return GT355_OK;
    }

    // Counter-based SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) -> U(a,b) in FP64.
    __device__ __forceinline__ uint64_t mix64(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }

    template <class T>
    __global__ void __launch_bounds__(256) fill_uniform_kernel(T *base, int64_t n0, int64_t n1, int64_t n2,
        int64_t s0, int64_t s1, int64_t s2, uint64_t seed_base, double a, double w) {
        int64_t total = n0 * n1 * n2;
        for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < total;
             n += (int64_t)gridDim.x * blockDim.x) {
            int64_t i = n % n0;
            int64_t r = n / n0;
            int64_t j = r % n1;
            int64_t k = r / n1;
            uint64_t x = mix64(seed_base + (uint64_t)(n + 1) * 0x9E3779B97F4A7C15ULL);
            double u = (double)(x >> 11) * 0x1.0p-53;
            base[i * s0 + j * s1 + k * s2] = (T)(a + w * u);
        }
    }

    __global__ void __launch_bounds__(256) flush_kernel(float4 *buf, int64_t n, float v) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            buf[i] = make_float4(v, v, v, v);
    }
} // namespace gt355

using namespace gt355;

extern "C" {

const char *gt355_last_error(void) { return gt355::last_error(); }

const char *gt355_version(void) { return "gt355 0.1 (gfx950)"; }

int gt355_init(int device) {
    int n = 0;
    GT355_HIP(hipGetDeviceCount(&n));
    GT355_CHECK(device >= 0 && device < n, "gt355_init: device %d out of range (%d devices)", device, n);
    GT355_HIP(hipSetDevice(device));
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_device = device;
    }
    // the no-allocation entry point's static scratch: its address on this device, resolved once
    return noalloc_setup();
}

int gt355_finalize(void) {
    if (int rc = defer_flush_all())
        return rc;
    GT355_HIP(hipDeviceSynchronize());
    release_deferred();
    release_pools();
    std::lock_guard<std::mutex> lock(g_mutex);
    for (auto &f : g_flush) {
        if (f.buf)
            (void)hipFree(f.buf);
        f.buf = nullptr;
        f.bytes = 0;
    }
    g_device = -1;
    return GT355_OK;
}

int gt355_get_device_info(gt355_device_info *out) {
    GT355_CHECK(out != nullptr, "gt355_get_device_info: null output");
    hipDeviceProp_t p;
    GT355_HIP(hipGetDeviceProperties(&p, current_device()));
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", p.name);
    std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    out->cu_count = p.multiProcessorCount;
    out->xcd_count = xcd_count();
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    out->lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
    out->l2_bytes = (int64_t)p.l2CacheSize;
    out->clock_khz = p.clockRate;
    out->mem_clock_khz = p.memoryClockRate;
    out->mem_bus_width = p.memoryBusWidth;
    out->wavefront = p.warpSize;
    return GT355_OK;
}

int gt355_storage_layout(int64_t elem_size, const int64_t lengths[3], const int64_t halos[3], gt355_layout *out) {
    GT355_CHECK(out && lengths && halos, "gt355_storage_layout: null argument");
    GT355_CHECK(elem_size == 4 || elem_size == 8, "gt355_storage_layout: element size %lld", (long long)elem_size);
    for (int d = 0; d < 3; ++d) {
        GT355_CHECK(lengths[d] > 0, "gt355_storage_layout: length[%d] = %lld", d, (long long)lengths[d]);
        GT355_CHECK(halos[d] >= 0 && 2 * halos[d] < lengths[d], "gt355_storage_layout: halo[%d] = %lld", d,
            (long long)halos[d]);
    }
    const int64_t align_bytes = 128;
    // storage/info.hpp:38-59: innermost length rounded up to a multiple of alignment / gcd(sizeof T, alignment)
    int64_t g = elem_size; // gcd(elem_size, 128) for 4 and 8
    int64_t align_elems = align_bytes / g;
    std::memset(out, 0, sizeof *out);
    out->elem_size = elem_size;
    for (int d = 0; d < 3; ++d) {
        out->lengths[d] = lengths[d];
        out->halos[d] = halos[d];
        out->padded[d] = lengths[d];
    }
    out->padded[0] = (lengths[0] + align_elems - 1) / align_elems * align_elems;
    // layout_map<2,1,0>: i stride 1, j stride padded_i, k stride padded_i * padded_j (storage/info.hpp:61-85)
    out->stride[0] = 1;
    out->stride[1] = out->padded[0];
    out->stride[2] = out->padded[0] * out->padded[1];
    // storage/data_store.hpp:63-72: allocate length + alignment elements
    int64_t length = out->stride[2] * out->padded[2];
    out->alloc_bytes = (length + align_elems) * elem_size;
    return GT355_OK;
}

int gt355_storage_base(const gt355_layout *layout, void *raw, void **base) {
    GT355_CHECK(layout && raw && base, "gt355_storage_base: null argument");
    const uintptr_t align = 128;
    int64_t off = layout->halos[0] * layout->stride[0] + layout->halos[1] * layout->stride[1] +
                  layout->halos[2] * layout->stride[2];
    uintptr_t byte_off = (uintptr_t)(off * layout->elem_size);
    uintptr_t addr = reinterpret_cast<uintptr_t>(raw) + byte_off;
    uintptr_t aligned = (addr + align - 1) / align * align;
    *base = reinterpret_cast<void *>(aligned - byte_off);
    return GT355_OK;
}

int gt355_malloc(size_t bytes, void **ptr) {
    GT355_CHECK(ptr != nullptr, "gt355_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0)
        return GT355_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        set_error("gt355_malloc(%zu): %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? GT355_ENOMEM : GT355_EHIP;
    }
    return GT355_OK;
}

int gt355_free(void *ptr) {
    if (ptr)
        GT355_HIP(hipFree(ptr));
    return GT355_OK;
}

int gt355_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (int rc = defer_flush_all())
        return rc;
    GT355_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GT355_OK;
}

int gt355_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (int rc = defer_flush_all())
        return rc;
    GT355_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GT355_OK;
}

int gt355_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return GT355_OK;
}

int gt355_memset(void *dst, int value, size_t bytes, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return GT355_OK;
}

int gt355_synchronize(void) {
    if (int rc = defer_flush_all())
        return rc;
    GT355_HIP(hipDeviceSynchronize());
    return GT355_OK;
}

int gt355_stream_synchronize(void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GT355_OK;
}

int gt355_fill_uniform(int dtype, void *base, const int64_t lengths[3], const int64_t stride[3], uint64_t seed,
    uint64_t field_id, double a, double b, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(base && lengths && stride, "gt355_fill_uniform: null argument");
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_fill_uniform: dtype %d", dtype);
    for (int d = 0; d < 3; ++d)
        GT355_CHECK(lengths[d] > 0 && stride[d] > 0, "gt355_fill_uniform: bad length/stride in dim %d", d);
    uint64_t seed_base = seed ^ (field_id * 0xD1B54A32D192ED03ULL);
    int64_t total = lengths[0] * lengths[1] * lengths[2];
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    if (dtype == GT355_F64)
        fill_uniform_kernel<double><<<blocks, 256, 0, (hipStream_t)stream>>>((double *)base, lengths[0], lengths[1],
            lengths[2], stride[0], stride[1], stride[2], seed_base, a, b - a);
    else
        fill_uniform_kernel<float><<<blocks, 256, 0, (hipStream_t)stream>>>((float *)base, lengths[0], lengths[1],
            lengths[2], stride[0], stride[1], stride[2], seed_base, a, b - a);
    return check_launch("gt355_fill_uniform");
}

int gt355_stream_create(void **stream) {
    GT355_CHECK(stream != nullptr, "gt355_stream_create: null output");
    hipStream_t s = nullptr;
    GT355_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GT355_OK;
}

int gt355_stream_destroy(void *stream) {
    if (stream)
        GT355_HIP(hipStreamDestroy((hipStream_t)stream));
    return GT355_OK;
}

int gt355_event_create(void **event) {
    GT355_CHECK(event != nullptr, "gt355_event_create: null output");
    hipEvent_t e = nullptr;
    GT355_HIP(hipEventCreate(&e));
    *event = e;
    return GT355_OK;
}

int gt355_event_destroy(void *event) {
    if (event)
        GT355_HIP(hipEventDestroy((hipEvent_t)event));
    return GT355_OK;
}

int gt355_event_record(void *event, void *stream) {
    GT355_CHECK(event != nullptr, "gt355_event_record: null event");
    GT355_DEFER_FLUSH(stream);
    GT355_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return GT355_OK;
}

int gt355_event_synchronize(void *event) {
    GT355_CHECK(event != nullptr, "gt355_event_synchronize: null event");
    GT355_HIP(hipEventSynchronize((hipEvent_t)event));
    return GT355_OK;
}

int gt355_event_elapsed_ms(void *start, void *stop, float *ms) {
    GT355_CHECK(start && stop && ms, "gt355_event_elapsed_ms: null argument");
    GT355_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GT355_OK;
}

int gt355_graph_begin(void *stream) {
    GT355_CHECK(stream != nullptr, "gt355_graph_begin: capture needs a non-null stream");
    GT355_DEFER_FLUSH(stream); // calls recorded before the capture run before it
    GT355_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    set_capturing(true);
    return GT355_OK;
}

int gt355_graph_end(void *stream, void **graph_exec) {
    GT355_CHECK(stream != nullptr && graph_exec != nullptr, "gt355_graph_end: null argument");
    GT355_DEFER_FLUSH(stream); // calls recorded during the capture belong to the graph
    hipGraph_t g = nullptr;
    set_capturing(false);
    GT355_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess)
        return hip_fail(err, "gt355_graph_end: hipGraphInstantiate");
    *graph_exec = e;
    return GT355_OK;
}

int gt355_graph_launch(void *graph_exec, void *stream) {
    GT355_CHECK(graph_exec != nullptr, "gt355_graph_launch: null graph");
    GT355_DEFER_FLUSH(stream);
    GT355_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return GT355_OK;
}

int gt355_graph_destroy(void *graph_exec) {
    if (graph_exec)
        GT355_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return GT355_OK;
}

int gt355_range_push(const char *name) {
    GT355_CHECK(name != nullptr, "gt355_range_push: null name");
    roctxRangePushA(name);
    return GT355_OK;
}

int gt355_range_pop(void) {
    roctxRangePop();
    return GT355_OK;
}

int gt355_mall_flush(void *stream) {
    GT355_DEFER_FLUSH(stream);
    // 1 GiB: four times the 256 MB Infinity Cache of MI355X.
    const size_t bytes = size_t(1) << 30;
    int dev = current_device();
    GT355_CHECK(dev >= 0 && dev < 64, "gt355_mall_flush: device %d", dev);
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (!g_flush[dev].buf) {
            void *p = nullptr;
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess)
                return hip_fail(e, "gt355_mall_flush: hipMalloc");
            g_flush[dev].buf = p;
            g_flush[dev].bytes = bytes;
        }
    }
    static float counter = 0.f;
    counter += 1.f;
    int64_t n = (int64_t)(g_flush[dev].bytes / sizeof(float4));
    flush_kernel<<<4096, 256, 0, (hipStream_t)stream>>>((float4 *)g_flush[dev].buf, n, counter);
    return check_launch("gt355_mall_flush");
}

} // extern "C"
