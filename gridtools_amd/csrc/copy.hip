// copy_stencil (tests/regression/copy_stencil.cpp:23-44, run_single_stage(copy_functor, ...)): out = in over the
// interior.  Also the bandwidth yardstick of the library: a pure stream of one read and one write per point,
// used to calibrate the HBM byte counters and to state the copy bandwidth the box reaches.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace gt355 {
    namespace {
        // One row segment of V elements per thread; rows are the i-lines of every (j, k).
        template <class T, int V>
        __global__ void __launch_bounds__(256) copy_kernel(T *__restrict__ out, const T *__restrict__ in, int ni,
            int nj, int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk, int segs) {
            int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
            int64_t rows = (int64_t)nj * nk;
            if (n >= rows * segs)
                return;
            int seg = int(n % segs);
            int64_t row = n / segs;
            int j = int(row % nj), k = int(row / nj);
            int i = seg * V;
            const T *src = in + i + j * isj + k * isk;
            T *dst = out + i + j * osj + k * osk;
            if (i + V <= ni) {
#pragma unroll
                for (int v = 0; v < V; ++v)
                    dst[v] = __builtin_nontemporal_load(src + v);
            } else {
                for (int v = 0; v < ni - i; ++v)
                    dst[v] = src[v];
            }
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_copy(const gt355_field *out, const gt355_field *in, int ni, int nj, int nk, void *stream) {
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
    constexpr int V = 4;
    int segs = (ni + V - 1) / V;
    int64_t total = (int64_t)segs * nj * nk;
    int64_t blocks = (total + 255) / 256;
    GT355_CHECK(blocks < (1LL << 31), "gt355_copy: domain too large");
    if (dtype == GT355_F64)
        copy_kernel<double, V><<<(int)blocks, 256, 0, s>>>((double *)out->origin, (const double *)in->origin, ni, nj,
            nk, out->stride[1], out->stride[2], in->stride[1], in->stride[2], segs);
    else
        copy_kernel<float, V><<<(int)blocks, 256, 0, s>>>((float *)out->origin, (const float *)in->origin, ni, nj, nk,
            out->stride[1], out->stride[2], in->stride[1], in->stride[2], segs);
    return check_launch("gt355_copy");
}
