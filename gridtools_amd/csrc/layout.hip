// layout_transformation (include/gridtools/layout_transformation.hpp:64-75, layout_transformation/gpu.hpp:72-116):
// dst[sum_d idx_d * dst_stride_d] = src[sum_d idx_d * src_stride_d] over an N-dimensional box, the operation the
// reference's C and Fortran bindings use to move arrays between a caller's layout and a storage's layout.
//
// The reference launches 8x8x8 thread blocks, one thread per (i, j, k) point, looping over the outer dimensions;
// when the two layouts have different fastest dimensions one of the two streams is uncoalesced.  Here:
//  * a = the dimension with the smallest source stride, b = the one with the smallest destination stride;
//  * a != b (a transpose in the (a, b) plane): LDS tiles of 64 (a) x TB (b) elements, TB = 64, 32 or 16 to fit the b
//    extent, read along a and written along b, so both HBM streams are coalesced; one workgroup walks tiles in a
//    grid-stride loop;
//  * with elements of 4 or 8 bytes, unit strides along a (source) and b (destination) and aligned vectors, the
//    tiles move 16 bytes per access on both sides (lt_tile_vec_kernel);
//  * a == b: a flat copy along a, 256 elements per workgroup row, the outer index decomposed once per row.
// Elements are moved as raw words of 1, 2, 4, 8 or 16 bytes, with non-temporal loads and stores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {
        constexpr int LT_MAXD = 8;

        struct lt_desc {
            int64_t dims[LT_MAXD], ds[LT_MAXD], ss[LT_MAXD];
            int nd, a, b;
        };

        struct word16 {
            uint64_t lo, hi;
        };
        // Every element is read once and written once: both streams are non-temporal, so that neither the L2 nor the
        // Infinity Cache keeps lines at the expense of others.
        template <class W>
        __device__ __forceinline__ W ld_nt(const W *p) {
            if constexpr (std::is_same<W, word16>::value)
                return word16{__builtin_nontemporal_load(&p->lo), __builtin_nontemporal_load(&p->hi)};
            else
                return __builtin_nontemporal_load(p);
        }
        template <class W>
        __device__ __forceinline__ void st_nt(W v, W *p) {
            if constexpr (std::is_same<W, word16>::value) {
                __builtin_nontemporal_store(v.lo, &p->lo);
                __builtin_nontemporal_store(v.hi, &p->hi);
            } else {
                __builtin_nontemporal_store(v, p);
            }
        }

        // offsets of outer index o over every dimension except a and b (a == b: all but a), first dimension fastest
        template <class IT>
        __device__ __forceinline__ void outer_offsets(const lt_desc &d, IT o, int64_t &so, int64_t &dof) {
            so = 0;
            dof = 0;
            for (int q = 0; q < d.nd; ++q) {
                if (q == d.a || q == d.b)
                    continue;
                const IT n = (IT)d.dims[q];
                const IT iq = o % n;
                o /= n;
                so += (int64_t)iq * d.ss[q];
                dof += (int64_t)iq * d.ds[q];
            }
        }

        // One TA x TB tile of the (a, b) plane per iteration: 256 threads load it with consecutive threads along a
        // (coalesced source rows), then store it with consecutive threads along b (coalesced destination rows).
        template <class W, int TA, int TB, class IT>
        __global__ void __launch_bounds__(256) lt_tile_kernel(W *__restrict__ dst, const W *__restrict__ src, lt_desc d,
            int64_t tiles_a, int64_t tiles_b, int64_t ntiles) {
            static_assert((TA * TB) % 256 == 0, "a tile must split evenly over the workgroup");
            __shared__ W tile[TB][TA + 1];
            constexpr int PER = TA * TB / 256;
            const int64_t na = d.dims[d.a], nb = d.dims[d.b];
            const int64_t ssa = d.ss[d.a], ssb = d.ss[d.b], dsa = d.ds[d.a], dsb = d.ds[d.b];
            for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
                const int64_t ta = t % tiles_a, r = t / tiles_a;
                const int64_t tb = r % tiles_b, o = r / tiles_b;
                int64_t so, dof;
                outer_offsets<IT>(d, (IT)o, so, dof);
                const int64_t a0 = ta * TA, b0 = tb * TB;
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int e = threadIdx.x + 256 * q;
                    const int aa = e % TA, bb = e / TA;
                    const int64_t ia = a0 + aa, ib = b0 + bb;
                    if (ia < na && ib < nb)
                        tile[bb][aa] = ld_nt(src + so + ia * ssa + ib * ssb);
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int e = threadIdx.x + 256 * q;
                    const int bb = e % TB, aa = e / TB;
                    const int64_t ia = a0 + aa, ib = b0 + bb;
                    if (ia < na && ib < nb)
                        st_nt(tile[bb][aa], dst + dof + ia * dsa + ib * dsb);
                }
                __syncthreads();
            }
        }

        // The same transpose with 16-byte accesses on both sides (V elements of 4 or 8 bytes): the source is
        // contiguous along a and the destination along b, and every vector is 16-byte aligned (both checked by the
        // launcher).  Loads: TA / V lanes per tile row, consecutive along a; stores: TB / V lanes per tile column,
        // consecutive along b.  A tile that reaches past the domain edge moves element by element.
        template <class W, int V, int TA, int TB, class IT>
        __global__ void __launch_bounds__(256) lt_tile_vec_kernel(W *__restrict__ dst, const W *__restrict__ src,
            lt_desc d, int64_t tiles_a, int64_t tiles_b, int64_t ntiles) {
            typedef W vec __attribute__((ext_vector_type(V)));
            constexpr int LPR = TA / V, LROWS = 256 / LPR, SPC = TB / V, SCOLS = 256 / SPC;
            static_assert(V * sizeof(W) == 16 && TA % V == 0 && TB % V == 0 && 256 % LPR == 0 && 256 % SPC == 0 &&
                              TB % LROWS == 0 && TA % SCOLS == 0,
                "tile shape");
            __shared__ W tile[TB][TA + 1];
            const int64_t na = d.dims[d.a], nb = d.dims[d.b];
            const int64_t ssb = d.ss[d.b], dsa = d.ds[d.a];
            const int la = (threadIdx.x % LPR) * V, lr = threadIdx.x / LPR;
            const int sb = (threadIdx.x % SPC) * V, sc = threadIdx.x / SPC;
            for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
                const int64_t ta = t % tiles_a, r = t / tiles_a;
                const int64_t tb = r % tiles_b, o = r / tiles_b;
                int64_t so, dof;
                outer_offsets<IT>(d, (IT)o, so, dof);
                const int64_t a0 = ta * TA, b0 = tb * TB;
                const bool full = a0 + TA <= na && b0 + TB <= nb;
#pragma unroll
                for (int rr = lr; rr < TB; rr += LROWS) {
                    const int64_t ia = a0 + la, ib = b0 + rr;
                    if (full) {
                        const vec v = ld_nt((const vec *)(src + so + ia + ib * ssb));
#pragma unroll
                        for (int q = 0; q < V; ++q)
                            tile[rr][la + q] = v[q];
                    } else {
#pragma unroll
                        for (int q = 0; q < V; ++q)
                            if (ia + q < na && ib < nb)
                                tile[rr][la + q] = ld_nt(src + so + ia + q + ib * ssb);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int cc = sc; cc < TA; cc += SCOLS) {
                    const int64_t ia = a0 + cc, ib = b0 + sb;
                    if (full) {
                        vec v;
#pragma unroll
                        for (int q = 0; q < V; ++q)
                            v[q] = tile[sb + q][cc];
                        st_nt(v, (vec *)(dst + dof + ia * dsa + ib));
                    } else {
#pragma unroll
                        for (int q = 0; q < V; ++q)
                            if (ia < na && ib + q < nb)
                                st_nt(tile[sb + q][cc], dst + dof + ia * dsa + ib + q);
                    }
                }
                __syncthreads();
            }
        }

        template <class W, class IT>
        __global__ void __launch_bounds__(256) lt_flat_kernel(W *__restrict__ dst, const W *__restrict__ src, lt_desc d,
            int64_t outer) {
            const int64_t na = d.dims[d.a], ssa = d.ss[d.a], dsa = d.ds[d.a];
            for (int64_t row = blockIdx.y; row < outer; row += gridDim.y) {
                int64_t so, dof;
                outer_offsets<IT>(d, (IT)row, so, dof);
                for (int64_t ia = blockIdx.x * 256 + threadIdx.x; ia < na; ia += (int64_t)gridDim.x * 256)
                    st_nt(ld_nt(src + so + ia * ssa), dst + dof + ia * dsa);
            }
        }

        template <class W>
        int launch_lt(void *dst, const void *src, const lt_desc &d, hipStream_t s) {
            int64_t outer = 1;
            for (int q = 0; q < d.nd; ++q)
                if (q != d.a && q != d.b)
                    outer *= d.dims[q];
            const bool narrow = outer < (int64_t(1) << 31);
            if (d.a != d.b) {
                // TB: the b-extent of a tile -- the largest of 64 / 32 / 16 that wastes at most an eighth of the
                // last tile row along b (nk = 80 takes 16)
                const int64_t nb = d.dims[d.b];
                auto waste_ok = [&](int64_t tb) { return ((nb + tb - 1) / tb * tb - nb) * 8 <= nb; };
                const int tbsel = sizeof(W) > 8 ? 32 : waste_ok(64) ? 64 : waste_ok(32) ? 32 : 16;
                // 16-byte accesses: 4- and 8-byte elements, unit stride along a (source) and b (destination), and
                // every vector aligned: the bases and all strides but those two in multiples of 16 bytes
                bool vec_ok = (sizeof(W) == 4 || sizeof(W) == 8) && d.ss[d.a] == 1 && d.ds[d.b] == 1 &&
                              (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0 && tune_param("layout_vec", 1) != 0;
                for (int q = 0; q < d.nd && vec_ok; ++q) {
                    if (q != d.a)
                        vec_ok = (d.ss[q] * (int64_t)sizeof(W)) % 16 == 0;
                    if (q != d.b && vec_ok)
                        vec_ok = (d.ds[q] * (int64_t)sizeof(W)) % 16 == 0;
                }
                auto go = [&](auto ta_c, auto tb_c) {
                    constexpr int TA = decltype(ta_c)::value, TB = decltype(tb_c)::value;
                    const int64_t tiles_a = (d.dims[d.a] + TA - 1) / TA, tiles_b = (nb + TB - 1) / TB;
                    const int64_t ntiles = tiles_a * tiles_b * outer;
                    const int blocks = (int)std::min<int64_t>(ntiles, 65536);
                    if constexpr (sizeof(W) == 4 || sizeof(W) == 8) {
                        constexpr int V = 16 / sizeof(W);
                        if (vec_ok) {
                            if (narrow)
                                lt_tile_vec_kernel<W, V, TA, TB, uint32_t><<<blocks, 256, 0, s>>>((W *)dst,
                                    (const W *)src, d, tiles_a, tiles_b, ntiles);
                            else
                                lt_tile_vec_kernel<W, V, TA, TB, uint64_t><<<blocks, 256, 0, s>>>((W *)dst,
                                    (const W *)src, d, tiles_a, tiles_b, ntiles);
                            return;
                        }
                    }
                    if (narrow)
                        lt_tile_kernel<W, TA, TB, uint32_t><<<blocks, 256, 0, s>>>((W *)dst, (const W *)src, d, tiles_a,
                            tiles_b, ntiles);
                    else
                        lt_tile_kernel<W, TA, TB, uint64_t><<<blocks, 256, 0, s>>>((W *)dst, (const W *)src, d, tiles_a,
                            tiles_b, ntiles);
                };
                using i16 = std::integral_constant<int, 16>;
                using i32 = std::integral_constant<int, 32>;
                using i64 = std::integral_constant<int, 64>;
                if constexpr (sizeof(W) > 8)
                    go(i32(), i32());
                else if (tbsel == 64)
                    go(i64(), i64());
                else if (tbsel == 32)
                    go(i64(), i32());
                else
                    go(i64(), i16());
            } else {
                dim3 grid((unsigned)std::min<int64_t>((d.dims[d.a] + 255) / 256, 4096),
                    (unsigned)std::min<int64_t>(outer, 65535));
                if (narrow)
                    lt_flat_kernel<W, uint32_t><<<grid, 256, 0, s>>>((W *)dst, (const W *)src, d, outer);
                else
                    lt_flat_kernel<W, uint64_t><<<grid, 256, 0, s>>>((W *)dst, (const W *)src, d, outer);
            }
            return check_launch("gt355_transform_layout");
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_transform_layout(int elem_size, void *dst, const void *src, int ndims, const int64_t *dims,
    const int64_t *dst_strides, const int64_t *src_strides, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(dst && src && dims && dst_strides && src_strides, "gt355_transform_layout: null argument");
    GT355_CHECK(ndims >= 1 && ndims <= LT_MAXD, "gt355_transform_layout: %d dimensions (1..%d)", ndims, LT_MAXD);
    GT355_CHECK(elem_size == 1 || elem_size == 2 || elem_size == 4 || elem_size == 8 || elem_size == 16,
        "gt355_transform_layout: element size %d (1, 2, 4, 8 or 16)", elem_size);
    GT355_CHECK(uintptr_t(dst) % elem_size == 0 && uintptr_t(src) % elem_size == 0,
        "gt355_transform_layout: pointers not aligned to the element size");
    lt_desc d{};
    d.nd = ndims;
    int64_t total = 1;
    for (int q = 0; q < ndims; ++q) {
        GT355_CHECK(dims[q] >= 0, "gt355_transform_layout: negative size in dimension %d", q);
        GT355_CHECK(src_strides[q] >= 0 && dst_strides[q] >= 0, "gt355_transform_layout: negative stride");
        d.dims[q] = dims[q];
        d.ds[q] = dst_strides[q];
        d.ss[q] = src_strides[q];
        total *= dims[q];
    }
    if (total == 0)
        return GT355_OK;
    // the destination index map must be injective: sorted by stride, each dimension of extent > 1 steps over the
    // whole span of the previous ones
    {
        int ord[LT_MAXD];
        int n = 0;
        for (int q = 0; q < ndims; ++q)
            if (dims[q] > 1)
                ord[n++] = q;
        std::sort(ord, ord + n, [&](int x, int y) { return dst_strides[x] < dst_strides[y]; });
        int64_t span = 1;
        for (int q = 0; q < n; ++q) {
            GT355_CHECK(dst_strides[ord[q]] >= span, "gt355_transform_layout: destination strides make points overlap");
            span = dst_strides[ord[q]] * dims[ord[q]];
        }
    }
    // a: the fastest source dimension, b: the fastest destination dimension (both among extents > 1)
    d.a = d.b = 0;
    int64_t best_s = -1, best_d = -1;
    for (int q = 0; q < ndims; ++q) {
        if (dims[q] <= 1)
            continue;
        if (best_s < 0 || src_strides[q] < best_s)
            best_s = src_strides[q], d.a = q;
        if (best_d < 0 || dst_strides[q] < best_d)
            best_d = dst_strides[q], d.b = q;
    }
    hipStream_t s = (hipStream_t)stream;
    switch (elem_size) {
    case 1:
        return launch_lt<uint8_t>(dst, src, d, s);
    case 2:
        return launch_lt<uint16_t>(dst, src, d, s);
    case 4:
        return launch_lt<uint32_t>(dst, src, d, s);
    case 8:
        return launch_lt<uint64_t>(dst, src, d, s);
    default:
        return launch_lt<word16>(dst, src, d, s);
    }
}
