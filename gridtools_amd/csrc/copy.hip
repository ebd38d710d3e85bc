// copy_stencil (tests/regression/copy_stencil.cpp:23-44, run_single_stage(copy_functor, ...)): out = in over the
// interior.  Also the bandwidth yardstick of the library: a pure stream of one read and one write per point,
// used to calibrate the HBM byte counters and to state the copy bandwidth the box reaches.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace gt355 {
    namespace {
        // One row segment of V consecutive elements per lane, moved as one 16-byte vector when the row start is
        // 16-byte aligned (ALIGNED; the launcher checks origins and strides), else element by element.  Lanes of a
        // wavefront take consecutive segments, so one load instruction covers 1 KiB of a row.  Every value is read
        // and written once: non-temporal on both sides.
        template <class T, int V, bool ALIGNED>
        __global__ void __launch_bounds__(256) copy_kernel(T *__restrict__ out, const T *__restrict__ in, int ni,
            int nj, int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk, int segs) {
            typedef T vec __attribute__((ext_vector_type(V)));
            const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
            const int64_t rows = (int64_t)nj * nk;
            if (n >= rows * segs)
                return;
            const int seg = int(n % segs);
            const int64_t row = n / segs;
            const int j = int(row % nj), k = int(row / nj);
            const int i = seg * V;
            const T *src = in + i + j * isj + k * isk;
            T *dst = out + i + j * osj + k * osk;
            if (ALIGNED && i + V <= ni) {
                __builtin_nontemporal_store(__builtin_nontemporal_load((const vec *)src), (vec *)dst);
            } else {
                for (int v = 0; v < V && i + v < ni; ++v)
                    __builtin_nontemporal_store(__builtin_nontemporal_load(src + v), dst + v);
            }
        }
        // Strided 3-D copy with element conversion (one thread per point, i fastest).
        template <class D, class S>
        __global__ void __launch_bounds__(256) convert_kernel(D *__restrict__ dst, const S *__restrict__ src, int64_t n0,
            int64_t n1, int64_t n2, int64_t d1, int64_t d2, int64_t s1, int64_t s2) {
            const int64_t total = n0 * n1 * n2;
            for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < total;
                 n += (int64_t)gridDim.x * blockDim.x) {
                const int64_t i = n % n0, r = n / n0, j = r % n1, k = r / n1;
                dst[i + j * d1 + k * d2] = static_cast<D>(src[i + j * s1 + k * s2]);
            }
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_convert(int dst_dtype, void *dst, const int64_t dst_stride[3], int src_dtype, const void *src,
    const int64_t src_stride[3], const int64_t lengths[3], void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(dst && src && dst_stride && src_stride && lengths, "gt355_convert: null argument");
    GT355_CHECK((dst_dtype == GT355_F64 || dst_dtype == GT355_F32) && (src_dtype == GT355_F64 || src_dtype == GT355_F32),
        "gt355_convert: dtypes %d <- %d", dst_dtype, src_dtype);
    for (int d = 0; d < 3; ++d)
        GT355_CHECK(lengths[d] >= 0, "gt355_convert: negative length");
    GT355_CHECK(dst_stride[0] == 1 && src_stride[0] == 1, "gt355_convert: i strides must be 1");
    GT355_CHECK(dst_stride[1] >= lengths[0] && dst_stride[2] >= dst_stride[1] * lengths[1] &&
                    src_stride[1] >= lengths[0] && src_stride[2] >= src_stride[1] * lengths[1],
        "gt355_convert: strides do not fit the lengths");
    const int64_t total = lengths[0] * lengths[1] * lengths[2];
    if (total == 0)
        return GT355_OK;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipStream_t s = (hipStream_t)stream;
#define GT355_CONVERT(D, S)                                                                                      \
    convert_kernel<D, S><<<blocks, 256, 0, s>>>((D *)dst, (const S *)src, lengths[0], lengths[1], lengths[2],    \
        dst_stride[1], dst_stride[2], src_stride[1], src_stride[2])
    if (dst_dtype == GT355_F64 && src_dtype == GT355_F64)
        GT355_CONVERT(double, double);
    else if (dst_dtype == GT355_F64)
        GT355_CONVERT(double, float);
    else if (src_dtype == GT355_F64)
        GT355_CONVERT(float, double);
    else
        GT355_CONVERT(float, float);
#undef GT355_CONVERT
    return check_launch("gt355_convert");
}

extern "C" int gt355_copy(const gt355_field *out, const gt355_field *in, int ni, int nj, int nk, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(out && in, "gt355_copy: null field");
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0, "gt355_copy: negative size");
    if (ni == 0 || nj == 0 || nk == 0)
        return GT355_OK;
    int dtype = in->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_copy: dtype %d", dtype);
    int rc;
    if ((rc = check_field(in, "in", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    if ((rc = check_field(out, "out", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    // 16-byte segments when every row of both fields starts 16-byte aligned
    const auto a16 = [](const gt355_field *f, int64_t es) {
        return (uintptr_t)f->origin % 16 == 0 && (f->stride[1] * es) % 16 == 0 && (f->stride[2] * es) % 16 == 0;
    };
    const int64_t es = dtype == GT355_F64 ? 8 : 4;
    const bool aligned = a16(in, es) && a16(out, es) && in->stride[0] == 1 && out->stride[0] == 1;
    const int V = int(16 / es);
    const int segs = (ni + V - 1) / V;
    const int64_t total = (int64_t)segs * nj * nk;
    const int64_t blocks = (total + 255) / 256;
    GT355_CHECK(blocks < (1LL << 31), "gt355_copy: domain too large");
#define GT355_COPY(T, VV, AL)                                                                                    \
    copy_kernel<T, VV, AL><<<(int)blocks, 256, 0, s>>>((T *)out->origin, (const T *)in->origin, ni, nj, nk,      \
        out->stride[1], out->stride[2], in->stride[1], in->stride[2], segs)
    if (dtype == GT355_F64) {
        if (aligned)
            GT355_COPY(double, 2, true);
        else
            GT355_COPY(double, 2, false);
    } else {
        if (aligned)
            GT355_COPY(float, 4, true);
        else
            GT355_COPY(float, 4, false);
    }
#undef GT355_COPY
    return check_launch("gt355_copy");
}
