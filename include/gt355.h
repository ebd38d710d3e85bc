/*
 * gt355.h -- C ABI of libgt355.so, the MI355X (gfx950) stencil backend for the GridTools stencil API.
 *
 * Every entry point takes plain pointers, sizes and the PODs below, returns 0 on success or a
 * negative gt355_status, and leaves a message for gt355_last_error() (thread-local) on failure.
 * This mirrors the reference's error convention, where every HIP/CUDA failure becomes a
 * std::runtime_error carrying the message (common/cuda_util.hpp:20-35): a C++ caller rethrows.
 *
 * Device memory is passed as raw device addresses.  Launches are asynchronous on the given HIP
 * stream (NULL = the null stream), like the reference's cuda_util::launch (common/cuda_util.hpp:77-95).
 * No entry point synchronises the device unless its name says so.
 */
#ifndef GT355_H
#define GT355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status / types */
enum gt355_status {
    GT355_OK = 0,
    GT355_EINVAL = -1,   /* bad argument (null pointer, size, stride, bounds) */
    GT355_EHIP = -2,     /* a HIP runtime call failed */
    GT355_ENOMEM = -3,   /* device allocation failed */
    GT355_EUNSUP = -4,   /* combination not supported (dtype / precision / size) */
};

enum gt355_dtype { GT355_F64 = 0, GT355_F32 = 1 };

/* How a stencil computes (the backend's Deref policy, stencil/be_api.hpp:160-169 + frontend/cartesian/stage.hpp:53-63):
 *  GT355_NATIVE : compute in the storage type (the reference's float_t instantiation, FP64 or FP32);
 *  GT355_WIDEN  : FP32 storage, every load widened to FP64, all temporaries, k-caches, Thomas pivots and
 *                 divisions in FP64, result narrowed once on store (the precision form). */
enum gt355_precision { GT355_NATIVE = 0, GT355_WIDEN = 1 };

/* A field as the backend sees it after core::shift_origin (stencil/core/backend.hpp:28-34): the SID
 * triple of origin, strides and bounds (sid/concept.hpp:36-120).  origin is the device address of grid point
 * (0,0,0); stride[0] must be 1 (i-innermost, storage/gpu.hpp:34-40) unless an entry point says otherwise (a
 * masked dimension of a selector<> storage has stride 0; gt355_copy_fields takes any i stride); lower[d] <= 0 <
 * upper[d] bound the offsets from the origin that may be touched (the allocation box).  Every entry point checks
 * the offsets its stencil reads or writes against these bounds before launching. */
typedef struct gt355_field {
    void *origin;
    int64_t stride[3];
    int32_t lower[3];
    int32_t upper[3];
    int32_t dtype; /* enum gt355_dtype */
    int32_t reserved;
} gt355_field;

/* ------------------------------------------------------------------ runtime */
/* Select and initialise a device (the reference leaves this to the CUDA runtime; gt355 keeps per-device
 * state: temporary pool, flush buffer, tuning table). */
int gt355_init(int device);
/* Waits for the device, then frees what the library keeps between calls: the temporary pool, the deep-column
 * scratch buffers (retired ones included) and the MALL-flush buffer.  The reference frees its allocator stash at
 * process exit (sid/allocator.hpp:44-57); a caller that must return the memory earlier calls this.  Graphs that use
 * library scratch must be destroyed first.  gt355_init (or any call) may follow: pools are re-created on use. */
int gt355_finalize(void);
const char *gt355_last_error(void);
const char *gt355_version(void);

typedef struct gt355_device_info {
    char name[64];
    char arch[32];       /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t cu_count;
    int32_t xcd_count;   /* 8 on MI355X; derived from the CU count and arch */
    int64_t hbm_bytes;
    int64_t lds_per_cu;
    int64_t l2_bytes;
    int32_t clock_khz;
    int32_t mem_clock_khz;
    int32_t mem_bus_width;
    int32_t wavefront;
} gt355_device_info;
int gt355_get_device_info(gt355_device_info *out);

/* ------------------------------------------------------------------ storage (storage/gpu.hpp:69-105)
 * gt355_storage_layout: the layout the reference's gpu storage traits give an N=3 field:
 * i-innermost layout_map<2,1,0> (storage/gpu.hpp:34-40), innermost length padded to a multiple of
 * 128/gcd(elem,128) elements (storage/info.hpp:38-59), strides from the padded lengths (info.hpp:61-85),
 * and an allocation of length + alignment elements so that the first interior point (halo_i, halo_j, halo_k)
 * can be placed on a 128-byte boundary (storage/data_store.hpp:63-72). */
typedef struct gt355_layout {
    int64_t lengths[3];   /* total lengths with halos */
    int64_t halos[3];
    int64_t padded[3];    /* padded lengths */
    int64_t stride[3];    /* element strides (stride[0] == 1) */
    int64_t elem_size;
    int64_t alloc_bytes;  /* bytes to request from gt355_malloc */
} gt355_layout;
int gt355_storage_layout(int64_t elem_size, const int64_t lengths[3], const int64_t halos[3], gt355_layout *out);
/* Given the raw allocation, the address of storage index (0,0,0) such that the first interior point is
 * 128-byte aligned (storage/data_store.hpp:63-72). */
int gt355_storage_base(const gt355_layout *layout, void *raw, void **base);

int gt355_malloc(size_t bytes, void **ptr);   /* storage_allocate (storage/gpu.hpp:78-81) */
/* Pooled temporaries, the cached allocator of the reference's temporaries (sid/allocator.hpp:44-57, whose static
 * stash hands freed blocks out again).  A released block is reused at once by an acquire on the same stream (stream
 * order keeps its users apart) and by an acquire on another stream only after the work queued before its release
 * has completed (an event recorded at the release).  Blocks released while the stream is captured into a graph
 * are reused on that stream only.  Blocks are freed by gt355_finalize. */
int gt355_tmp_acquire(size_t bytes, void *stream, void **ptr);
int gt355_tmp_release(void *ptr, void *stream);
/* Releases n blocks (null entries skipped) with one event for all of them, as the reference's run drops every
 * temporary of a spec together once its kernels are queued (stencil/frontend/run.hpp:34-60 with the allocator of
 * sid/allocator.hpp:44-57).  Every pointer is checked first: on an error no block is released. */
int gt355_tmp_release_n(void *const *ptrs, int n, void *stream);
int gt355_free(void *ptr);
int gt355_memcpy_h2d(void *dst, const void *src, size_t bytes); /* storage_update_target (gpu.hpp:83-90) */
int gt355_memcpy_d2h(void *dst, const void *src, size_t bytes); /* storage_update_host (gpu.hpp:92-99) */
int gt355_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int gt355_memset(void *dst, int value, size_t bytes, void *stream);
int gt355_synchronize(void);
int gt355_stream_synchronize(void *stream);

/* Synthetic inputs: counter-based SplitMix64 -> U(a,b) computed in FP64 (rounded once to FP32 for F32
 * fields) over a logical box of lengths[3] points (i fastest, then j, then k), element n at
 * base + i*stride[0] + j*stride[1] + k*stride[2]; n = i + lengths[0]*(j + lengths[1]*k). */
int gt355_fill_uniform(int dtype, void *base, const int64_t lengths[3], const int64_t stride[3], uint64_t seed,
    uint64_t field_id, double a, double b, void *stream);

/* ------------------------------------------------------------------ stencils
 * Horizontal diffusion (tests/regression/horizontal_diffusion.cpp:35-107): one launch fusing the
 * lap, flx, fly and out stages of the execute_parallel().ij_cached(lap, flx, fly) multistage, the job
 * of gpu::gridtools_backend_entry_point (stencil/gpu/entry_point.hpp:253-260) for this spec.
 * Reads in over [-2, n+2) in i and j, coeff and writes out over [0, n). */
int gt355_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni, int nj,
    int nk, void *stream);

/* copy_stencil (tests/regression/copy_stencil.cpp:23-44, run_single_stage(copy_functor)): out = in over the
 * interior.  Also the library's bandwidth yardstick (one read + one write per point). */
int gt355_copy(const gt355_field *out, const gt355_field *in, int ni, int nj, int nk, void *stream);

/* Laplacian (tests/regression/laplacian.cpp:20-38, run_single_stage): out = 4 in - (in[i+-1] + in[j+-1]). */
int gt355_laplacian(int precision, const gt355_field *out, const gt355_field *in, int ni, int nj, int nk,
    void *stream);

/* Vertical advection (tests/regression/vertical_advection_dycore.cpp:32-150): forward Thomas sweep
 * with k-caches ccol/dcol (flush) and u_stage (fill), then backward substitution with data_col, fused into
 * one launch like the reference's multi_kernel (stencil/gpu/entry_point.hpp:48-59,107-124).
 * utens_stage is read (forward) and overwritten (backward).  Reads wcon at i+1 and k+1; nk >= 2. */
int gt355_vadv(int precision, const gt355_field *utens_stage, const gt355_field *u_stage, const gt355_field *wcon,
    const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj, int nk, void *stream);

/* gt355_vadv without any allocation.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as gt355_vadv.  Where gt355_vadv takes ccol/dcol scratch from its per-stream pool
 * (nk > 80, or a field span beyond 4 GiB), this entry point uses a static device array of 64 MiB that each device
 * holds for the life of the process, processing the rows in chunks whose ccol/dcol fit it; a single row needing more
 * than 64 MiB (sizeof(compute type) * ni * (2 nk - 1) bytes) is rejected with GT355_EINVAL before any launch.  It
 * calls no hipMalloc / hipFree.  Every call on one device shares that device's array: calls that overlap in time on
 * one device (two streams, two threads) race and must be ordered by the caller. */
int gt355_vadv_noalloc(int precision, const gt355_field *utens_stage, const gt355_field *u_stage,
    const gt355_field *wcon, const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj,
    int nk, void *stream);

/* hdiff -> vadv composition (one run() with multi_pass(hdiff MSS, forward MSS, backward MSS); the reference gpu
 * backend splits it into the same two launches, stencil/gpu/entry_point.hpp:96-124): u_stage = hdiff(in, coeff),
 * then vadv(utens_stage, u_stage, wcon, u_pos, utens).  Every field is validated before the first launch. */
int gt355_hdiff_vadv(int precision, const gt355_field *in, const gt355_field *coeff, const gt355_field *u_stage,
    const gt355_field *utens_stage, const gt355_field *wcon, const gt355_field *u_pos, const gt355_field *utens,
    double dtr_stage, int ni, int nj, int nk, void *stream);

/* ------------------------------------------------------------------ further cartesian regression stencils
 * simple_hori_diff (tests/regression/simple_hori_diff.cpp:25-85): the execute_parallel().ij_cached(lap)
 * multistage of wlap_function (lap = in[i+1] + in[i-1] - 2 in + crlato (in[j+1] - in) + crlatu (in[j-1] - in))
 * and divflux_function (out = in + ((fluxx_m - fluxx) + (fluxy_m - fluxy)) coeff), as one launch.  crlato and
 * crlatu are typically j-only fields (the selector<0,1,0> storages of the reference test): stride 0 in i and k,
 * bounds checked only in j.  Full 3-D metric fields are accepted too.  Reads in over [-2, n+2) in i and j and
 * the metric over [-1, n+1); writes out over [0, n). */
int gt355_simple_hori_diff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *crlato,
    const gt355_field *crlatu, const gt355_field *coeff, int ni, int nj, int nk, void *stream);

/* expandable_run of prepare_tracers (tests/regression/advection_pdbott_prepare_tracers.cpp:23-57, the member
 * splitting of stencil/frontend/expandable_run.hpp:150-183): data[m] = rho * data_nnow[m] for m < nmembers.  The
 * reference launches one replicated spec per chunk of Factor members; here one launch covers up to 32 members and
 * reads rho once per point.  All data members share one layout, all data_nnow members another. */
int gt355_prepare_tracers(int precision, const gt355_field *data, const gt355_field *data_nnow, int nmembers,
    const gt355_field *rho, int ni, int nj, int nk, void *stream);

/* copy_stencil_tuple (tests/regression/copy_stencil_tuple.cpp:23-100 functors, :116-260 tests): out[f] = in[f] over
 * the interior for nfields field pairs of one dtype, one launch per 16 pairs.  Unlike the other entry points any
 * non-negative i stride is accepted, so each member of an array-of-structs storage (array<T, 5> or tuple<T...>
 * elements: origin + f, i stride 5) and each slice of an extra data dimension is a field; an input may have
 * masked (stride 0) dimensions, an output may not, and an output's index map must be injective. */
int gt355_copy_fields(const gt355_field *out, const gt355_field *in, int nfields, int ni, int nj, int nk,
    void *stream);

/* extended_4D (tests/regression/extended_4D.cpp:58-135): per grid point, the projection
 *   result(I,J,K) = 1/8 sum_q phi(I,J,K,q) jac(q) sum_{a,b,c} psi(a,b,c,q) f(a,b,c)      (I..c, q in {0, 1})
 * in the operation order of integration::apply.  result and f are storages with three extra dimensions of 2
 * (element (I,J,K) at the spatial field's address + I*result_extra[0] + J*result_extra[1] + K*result_extra[2]), jac
 * has one (q at + q*jac_extra): the gpu layout puts extra dimensions after k (storage/gpu.hpp:27-40).  phi / psi are
 * the global parameters' values, [((x*2 + y)*2 + z)*2 + q], in double as the reference's elemental returns them
 * (a float storage keeps sum and res in float and the products in double, like the C++ expression). */
int gt355_extended_4d(const double phi[16], const double psi[16], const gt355_field *result,
    const int64_t result_extra[3], const gt355_field *f, const int64_t f_extra[3], const gt355_field *jac,
    int64_t jac_extra, int ni, int nj, int nk, void *stream);

/* transform_layout (include/gridtools/layout_transformation.hpp:64-75, layout_transformation/gpu.hpp:72-116):
 * dst[sum_d i_d * dst_strides[d]] = src[sum_d i_d * src_strides[d]] for every index of the ndims-dimensional box
 * dims (1 <= ndims <= 8), strides in elements of elem_size bytes (1, 2, 4, 8 or 16).  Strides are non-negative and
 * the destination's index map must be injective (checked); elements outside the index map are not touched.  Both
 * pointers are device addresses covering every index of the box (the caller's allocation, as in the reference). */
int gt355_transform_layout(int elem_size, void *dst, const void *src, int ndims, const int64_t *dims,
    const int64_t *dst_strides, const int64_t *src_strides, void *stream);

/* Strided 3-D copy with element conversion between F64 and F32 (the role of layout_transformation,
 * include/gridtools/layout_transformation/gpu.hpp:72-116, plus a dtype change): dst[i + j*ds1 + k*ds2] =
 * (dst type) src[i + j*ss1 + k*ss2] over lengths[0] x lengths[1] x lengths[2]; i strides must be 1. */
int gt355_convert(int dst_dtype, void *dst, const int64_t dst_stride[3], int src_dtype, const void *src,
    const int64_t src_stride[3], const int64_t lengths[3], void *stream);

/* ------------------------------------------------------------------ boundaries
 * gt355_bc_apply: boundary::apply for the predefined functors (boundaries/boundary.hpp:57-79,
 * apply_gpu.hpp:235-309): kind 0 zero_boundary (zero.hpp:29-55), 1 value_boundary (value.hpp:27-67),
 * 2 copy_boundary dst <- src (copy.hpp:26-47).  Fields are passed as storage bases (storage index (0,0,0));
 * halo[d] = {minus, plus, begin, end, total_length} as halo_descriptor (common/halo_descriptor.hpp:44-49).
 * dir_mask bit (di+1)*9+(dj+1)*3+(dk+1) enables direction (di,dj,dk), the pre-evaluated predicate of
 * apply_gpu.hpp:52-83.  nfields fields (1..4) get the same treatment; src is used by kind 2 only. */
int gt355_bc_apply(int kind, int dtype, void *const *dst, void *const *src, int nfields, double value,
    const int32_t halo[3][5], const int64_t stride[3], uint32_t dir_mask, void *stream);

/* gt355_halo_fill: the single-rank result of gcl pack -> periodic self-exchange -> unpack
 * (gcl/high_level/descriptors_manual_gpu.hpp:342-512, gcl/low_level/proc_grids_3D.hpp:83-88) fused with a
 * zero-gradient boundary, in one launch per call.  mode[d]: 0 none, 1 zero-gradient, 2 periodic.
 * Every filled point takes the value of the interior point whose coordinates are clamped (mode 1) or
 * wrapped (mode 2) in each dimension where it lies in the halo. */
int gt355_halo_fill(int dtype, void *const *fields, int nfields, const int32_t halo[3][5], const int64_t stride[3],
    const int32_t mode[3], void *stream);

/* gt355_halo_pack / gt355_halo_unpack: copy the 26 halo-source slabs of nfields fields into one contiguous
 * buffer, ordered by direction then field, i fastest (the message layout of gcl m_pack*.hpp:15-112), and
 * back into the halos.  Pack reads the interior slab adjacent to the face/edge/corner (the data a neighbour
 * in that direction needs); unpack writes the halo slab on the opposite side.  gt355_halo_buffer_size gives
 * the bytes for a direction mask. */
int gt355_halo_buffer_size(int dtype, int nfields, const int32_t halo[3][5], uint32_t dir_mask, size_t *bytes);
int gt355_halo_pack(int dtype, void *const *fields, int nfields, const int32_t halo[3][5], const int64_t stride[3],
    uint32_t dir_mask, void *buffer, void *stream);
int gt355_halo_unpack(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, const void *buffer, void *stream);
/* The single-rank halo exchange with a periodic process grid (gcl/low_level/proc_grids_3D.hpp:83-88): pack the
 * slabs of the directions in dir_mask into `buffer` (gt355_halo_buffer_size bytes), then unpack every halo region d
 * from the segment packed for -d -- the neighbour at d is this rank.  The mask must be closed under negation and
 * the halos symmetric in the dimensions it exchanges. */
int gt355_halo_exchange_local(int dtype, void *const *fields, int nfields, const int32_t halo[3][5],
    const int64_t stride[3], uint32_t dir_mask, void *buffer, void *stream);

/* ------------------------------------------------------------------ streams and timing
 * Replaces timer_cuda (common/timer/timer_cuda.hpp:21-54): hipEvent pairs on a stream.  Streams are created
 * with default priority. */
int gt355_stream_create(void **stream);
/* The calling thread's current stream: the reference launches on the default stream and leaves propagation of a
 * stream open ("if required propagate CUDA stream", stencil/gpu/launch_kernel.hpp:158-161).  The C++ front end's
 * backend tag launches on it when the tag carries no stream of its own.  NULL (the null stream) until set. */
int gt355_stream_set_current(void *stream);
int gt355_stream_current(void **stream);
/* Debug launch mode, the reference's debug-build synchronisation after each launch (common/cuda_util.hpp:89-93):
 * on != 0 makes every entry point wait for its kernels and report a kernel failure at that call.  Also switched on
 * by GT355_DEBUG_SYNC=1 in the environment.  Not applied while the calling thread captures a graph. */
int gt355_debug_sync(int on);
int gt355_stream_destroy(void *stream);
int gt355_event_create(void **event);
int gt355_event_destroy(void *event);
int gt355_event_record(void *event, void *stream);
int gt355_event_synchronize(void *event);
int gt355_event_elapsed_ms(void *start, void *stop, float *ms);

/* ------------------------------------------------------------------ autotuner
 * No counterpart in the reference, whose tile shape is a backend template parameter fixed at compile time
 * (stencil/gpu/entry_point.hpp:147-150, gpu<64,8,1> by default).  gt355_tune_hdiff times every instantiated hdiff
 * configuration (impl 0: LDS tiles TI x TJ x KC; impl 1: one-column register sweeps JB/P; impl 2: column-pair
 * register sweeps JB/P, for an even ni and 2-element-aligned rows) on the given fields, `reps` launches each on
 * `stream` back to back, cycling through temporary copies of the three fields so that the data exceeds the MALL
 * (as a caller streaming several field sets sees it), in two passes (a candidate's time is the faster of its two),
 * records the fastest for (storage type, compute type, ni, nj, nk) so later gt355_hdiff calls with that shape use
 * it -- the measured default configuration unless a candidate beats it by more than 2% -- and returns it in
 * best[4] = {impl, a, b, c} with its mean time.  The fields' contents are
 * overwritten (out) and read (in, coeff) as by gt355_hdiff.  The table persists across processes through
 * gt355_tune_save / gt355_tune_load (text, one entry per line). */
int gt355_tune_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
    int nj, int nk, int reps, void *stream, int32_t best[4], double *best_us);
/* The tuning table by kernel key (the autotune cache lookup of the reference's backend template parameters,
 * stencil/gpu/entry_point.hpp:147-150, made a run-time table).  "hdiff": shape = {storage bytes, compute bytes, ni,
 * nj, nk}, cfg = {impl, a, b, c} as gt355_tune_hdiff returns it; "batch": shape = {domain-list signature, storage
 * bytes, compute bytes, form}, cfg = {rows per stencil item, stencil workgroups per CU, grid setting, 96-level
 * depth}.  lookup sets *found (and cfg, *us when found); store records an entry as the tuners do (an hdiff
 * configuration must be an instantiated one), for the next launch of that shape. */
int gt355_tune_lookup(const char *kernel_key, const int64_t *shape, int n, int32_t cfg[4], double *us, int *found);
int gt355_tune_store(const char *kernel_key, const int64_t *shape, int n, const int32_t cfg[4], double us);
int gt355_tune_save(const char *path);
int gt355_tune_load(const char *path, int *entries);
int gt355_tune_clear(void);
/* Launch-configuration knobs GT355_TUNE_<name> are read from the environment once per process; this reads them
 * again (for a caller that changes them at run time). */
int gt355_tune_env_refresh(void);

/* ------------------------------------------------------------------ graphs
 * Stream capture of a launch sequence into a hipGraph (the reference has no equivalent: it launches every
 * multistage eagerly on stream 0, stencil/gpu/launch_kernel.hpp:161).  Capture needs a non-null stream.  No entry
 * point records a memset into a captured graph; entry points that may allocate (the deep-column vadv scratch
 * pool) must have run once, outside capture, with the largest shape before a capture that uses them. */
int gt355_graph_begin(void *stream);
int gt355_graph_end(void *stream, void **graph_exec);
int gt355_graph_launch(void *graph_exec, void *stream);
int gt355_graph_destroy(void *graph_exec);

/* ------------------------------------------------------------------ batching
 * Many independent domains of uneven sizes, every domain in each persistent launch.  Replaces looping the
 * reference's run() (stencil/frontend/run.hpp:34-60) over domains, and generalises expandable_run
 * (stencil/frontend/expandable_run.hpp:80-183), which batches fields of ONE grid size only.  Per domain and level
 * the batch computes lap = laplacian(in) (laplacian.cpp:20-29), u_stage = hdiff(in, coeff) and
 * utens_stage = vadv(utens_stage, u_stage, wcon, u_pos, utens).  Domains: ni, nj >= 1, 2 <= nk <= 144; `in`
 * readable on [-2, n+2) in i and j, wcon on [0, ni+1) in i; every field but wcon shares the layout of `in`.  All
 * domains have the dtype of domains[0].in.  Forms (chosen at create time, GT355_TUNE_batch_impl):
 *   3 (default), two phases: one persistent launch of the laplacian and hdiff register sweep over every domain's
 *     (64-column strip, row block, level) items, then one of the vertical advection over every domain's 64-column
 *     strips of one row (GT355_TUNE_batch_flat=1: 64 consecutive entries of its column list per item instead),
 *     u_stage read back from memory; results equal gt355_laplacian + gt355_hdiff_vadv's two-launch
 *     composition bit for bit (in the precision form u_stage reaches the vertical phase rounded to FP32);
 *   2, one launch: strips evaluate hdiff and the vertical advection together, u_stage passed on in the compute
 *     type; results equal gt355_laplacian + the fused gt355_hdiff_vadv bit for bit;
 *   1, one launch over 64x4 column tiles with ccol/dcol in scratch (kept for comparison).
 * gt355_batch_create builds the work tables (longest columns first) and the per-workgroup scratch; gt355_batch_run
 * launches the persistent grids (no counters or memsets, so a run may be captured into a graph once the batch
 * exists).  Each batch owns its ccol/dcol scratch: two runs of ONE batch must not overlap in time, runs of
 * different batches may.
 * In the two-phase form a domain may take part in one phase only (`phases`: 0 both, 1 the stencil phase alone,
 * computing u_stage = hdiff(in, coeff) like gt355_hdiff; 2 the vertical phase alone, reading u_stage like
 * gt355_vadv); fields a domain's phases do not use may be null (origin NULL), and `lap` may be null in the
 * two-phase form (the laplacian is then not stored).  A stencil-phase-only domain may have any nk >= 1. */
typedef struct gt355_batch_domain {
    gt355_field in, coeff, lap, u_stage, utens_stage, wcon, u_pos, utens;
    int32_t ni, nj, nk, phases;
} gt355_batch_domain;
typedef struct gt355_batch gt355_batch;
int gt355_batch_create(int precision, const gt355_batch_domain *domains, int ndomains, int wg_per_cu,
    gt355_batch **batch);
int gt355_batch_run(gt355_batch *batch, double dtr_stage, void *stream);
int gt355_batch_info(const gt355_batch *batch, int *nitems, int *grid);
/* The batch's part of the autotuner: times `reps` runs of each launch configuration (coordinate by coordinate): the
 * persistent grid setting (1..4), for FP32 storage in the two-phase form the depth from which the vertical phase
 * keeps 96 levels in registers (81, 89, 97 or never), then the stencil phase's rows per item (8, 16, 32) and
 * workgroups per CU (2, 4, 8).  It keeps the fastest for later runs of this batch, records it in the tuning table
 * under the signature of the domain list (gt355_tune_save persists it; a batch created later for the same list of
 * domain sizes, form and types starts from it), and returns the grid setting with the mean time.  The batch's
 * outputs (and utens_stage, read and rewritten by every run) are overwritten as by that many gt355_batch_run
 * calls. */
int gt355_batch_tune(gt355_batch *batch, int reps, void *stream, int *wg_per_cu, double *best_us);
/* The launch configuration of a batch: cfg = {form, stencil-phase rows per item, stencil-phase workgroups per CU,
 * persistent grid setting, vertical-phase 96-level depth}.  set_config applies one (the form must be the batch's). */
int gt355_batch_get_config(const gt355_batch *batch, int32_t cfg[5]);
int gt355_batch_set_config(gt355_batch *batch, const int32_t cfg[5]);

/* A deferred-run scope (no counterpart in the reference, whose run() launches at once, stencil/frontend/run.hpp:
 * 34-60; it generalises expandable_run, stencil/frontend/expandable_run.hpp:80-183, to domains of different sizes).
 * Between gt355_batch_begin(stream) and gt355_batch_end(stream), calls of gt355_hdiff, gt355_vadv and
 * gt355_hdiff_vadv on that stream (the same pointer, NULL included) from the calling thread are checked and
 * recorded, not launched.  gt355_batch_end runs them as if they ran one after another in the recorded order: each
 * call is placed in the first launch after every recorded call it conflicts with (one reads or writes bytes the
 * other writes, by the fields' declared bounds), hdiff parts in stencil phases and vadv parts in vertical phases,
 * and every phase with its calls' domains runs as one persistent launch of a two-phase batch (gt355_batch_create,
 * grouped by dtype, precision and dtr_stage).  gt355_hdiff_vadv records as its hdiff and vadv parts (the
 * two-launch composition's results).  A call whose fields do not share one layout, or whose depth the vertical
 * phase does not hold (2..144 levels), runs on its own entry point at its place in the order.  The batches are
 * built at the first end with a given list of calls and kept (up to 16 lists per thread, released by
 * gt355_finalize), so a repeated scope launches only: run it once before capturing it into a graph.  Any other
 * library call that launches on the scope's stream (the generic engine's runs included) first runs the calls
 * recorded so far, keeping the order; gt355_batch_flush does the same explicitly and leaves the scope open.
 * Fields of recorded calls must not be changed by other work before the calls run. */
int gt355_batch_begin(void *stream);
int gt355_batch_flush(void *stream);
int gt355_batch_end(void *stream);
int gt355_batch_destroy(gt355_batch *batch);

/* ------------------------------------------------------------------ measurement */
/* Named host ranges for tracers (rocprofv3 --marker-trace shows them around the launches they enclose); the role
 * of the reference's timer hooks (common/timer/timer_cuda.hpp:21-54), which measure named regions with events. */
int gt355_range_push(const char *name);
int gt355_range_pop(void);

/* Writes (and reads back) a device buffer larger than the MI355X MALL so that the next kernel reads
 * HBM, not Infinity Cache.  Used between timed repetitions, outside the timed region. */
int gt355_mall_flush(void *stream);

#ifdef __cplusplus
}
#endif
#endif
