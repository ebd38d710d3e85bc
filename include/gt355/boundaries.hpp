/*
 * gt355/boundaries.hpp -- user boundary-condition functors on the device, for boundaries::boundary<BC, Arch, Pred>
 * of gt355/gridtools.hpp (the reference's boundary_apply_gpu, boundaries/apply_gpu.hpp:235-310, which instantiates
 * the user's functor in a kernel).  Include it in a translation unit compiled by hipcc; the predefined zero / value /
 * copy functors do not need it.
 *
 * A functor has overloads operator()(direction<I, J, K>, fields&..., unsigned i, unsigned j, unsigned k), called
 * with storage indices of a halo point of region (I, J, K), as in boundaries/zero.hpp:29-55.  A field argument is a
 * view with operator()(i, j, k) returning a reference (writable through a const view, like the reference's data
 * views).  The predicate was evaluated per direction on the host.
 *
 * The launch: every selected region gets its own run of workgroups (a workgroup finds its region once, by a uniform
 * scan over at most 26 block offsets), points are numbered with i fastest in 32-bit arithmetic, and a thread's
 * (i, j, k) comes from multiply-high divisions with host-computed magic numbers.  The region's direction selects the
 * functor overload through a switch over the 26 direction types.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "gridtools.hpp"

#ifndef GT_FUNCTION
#define GT_FUNCTION __host__ __device__ __forceinline__
#endif

namespace gt355 {
    namespace boundaries {
        // a field of a boundary functor: storage index (i, j, k) -> element
        template <class T>
        struct bc_view {
            T *base;
            int64_t s1, s2;
            __device__ __forceinline__ T &operator()(unsigned i, unsigned j, unsigned k) const {
                return base[(int64_t)i + (int64_t)j * s1 + (int64_t)k * s2];
            }
        };

        namespace bc_impl {
            // x / d as ((mulhi(x, m) + x) >> l), the round-up method of Granlund and Montgomery (PLDI 1994)
            struct udiv {
                uint32_t m, l;
            };
            inline udiv make_udiv(uint32_t d) {
                uint32_t l = 0;
                while ((uint64_t(1) << l) < d)
                    ++l;
                return udiv{(uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1), l};
            }

            constexpr int BLOCK = 256, PER_THREAD = 4, CHUNK = BLOCK * PER_THREAD;

            struct uregion {
                int32_t lo[3], n[3];
                int32_t code;   // (di+1)*9 + (dj+1)*3 + (dk+1)
                int32_t block0; // first workgroup of the region
                udiv d0, d1;    // division by n[0] and by n[0] * n[1]
            };
            struct uregions {
                uregion r[26];
                int count, blocks;
            };

            // the halo regions of the selected directions (boundaries/apply.hpp:45-135 region boxes)
            inline uregions make_regions(std::array<halo_descriptor, 3> const &h, uint32_t mask) {
                uregions t{};
                for (int di = -1; di <= 1; ++di)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int dk = -1; dk <= 1; ++dk) {
                            const int code = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
                            if ((di == 0 && dj == 0 && dk == 0) || !((mask >> code) & 1u))
                                continue;
                            uregion r{};
                            const int dd[3] = {di, dj, dk};
                            int64_t vol = 1;
                            for (int d = 0; d < 3; ++d) {
                                halo_descriptor const &x = h[d];
                                if (dd[d] == 0) {
                                    r.lo[d] = x.begin();
                                    r.n[d] = x.end() - x.begin() + 1;
                                } else if (dd[d] < 0) {
                                    r.lo[d] = x.begin() - x.minus();
                                    r.n[d] = x.minus();
                                } else {
                                    r.lo[d] = x.end() + 1;
                                    r.n[d] = x.plus();
                                }
                                vol *= r.n[d];
                            }
                            if (vol == 0)
                                continue;
                            if (vol >= (int64_t(1) << 31))
                                throw std::invalid_argument("boundary::apply: a halo region holds 2^31 points or more");
                            r.code = code;
                            r.block0 = t.blocks;
                            r.d0 = make_udiv((uint32_t)r.n[0]);
                            r.d1 = make_udiv((uint32_t)r.n[0] * (uint32_t)r.n[1]);
                            t.blocks += (int)((vol + CHUNK - 1) / CHUNK);
                            t.r[t.count++] = r;
                        }
                return t;
            }

            template <int Code>
            using direction_of = direction<sign(Code / 9 - 1), sign(Code / 3 % 3 - 1), sign(Code % 3 - 1)>;

            template <class BC, class... V, int... Codes>
            __device__ __forceinline__ void call_at(
                int code, BC const &bc, unsigned i, unsigned j, unsigned k, std::integer_sequence<int, Codes...>,
                V &... v) {
                ((code == Codes ? (void)bc(direction_of<Codes>(), v..., i, j, k) : (void)0), ...);
            }

            template <class BC, class... V>
            __global__ void __launch_bounds__(BLOCK) user_bc_kernel(BC bc, uregions t, V... v) {
                int q = 0;
#pragma unroll 1
                for (int c = 1; c < t.count; ++c)
                    if ((int)blockIdx.x >= t.r[c].block0)
                        q = c;
                const uregion &r = t.r[q];
                const uint32_t vol = (uint32_t)r.n[0] * (uint32_t)r.n[1] * (uint32_t)r.n[2];
                const uint32_t base = (uint32_t)((int)blockIdx.x - r.block0) * CHUNK + threadIdx.x;
                for (int u = 0; u < PER_THREAD; ++u) {
                    const uint32_t loc = base + u * BLOCK;
                    if (loc >= vol)
                        break;
                    const uint32_t row = (uint32_t)(((uint64_t)__umulhi(loc, r.d0.m) + loc) >> r.d0.l);
                    const uint32_t plane = (uint32_t)(((uint64_t)__umulhi(loc, r.d1.m) + loc) >> r.d1.l);
                    const unsigned i = (unsigned)(r.lo[0] + (int)(loc - row * (uint32_t)r.n[0]));
                    const unsigned j = (unsigned)(r.lo[1] + (int)(row - plane * (uint32_t)r.n[1]));
                    const unsigned k = (unsigned)(r.lo[2] + (int)plane);
                    call_at(r.code, bc, i, j, k, std::integer_sequence<int, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
                                                     14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26>(),
                        v...);
                }
            }

            template <class T>
            bc_view<T> view_of(std::shared_ptr<storage::data_store<T>> const &ds, std::array<int, 3> const &need) {
                auto len = ds->lengths();
                for (int d = 0; d < 3; ++d)
                    if (need[d] > len[d])
                        throw std::invalid_argument("boundary::apply: field " + ds->name() +
                                                    " is smaller than the halo descriptors' total lengths");
                auto st = ds->strides();
                return {ds->get_target_ptr(), st[1], st[2]};
            }

            template <class BC, class... Stores>
            void apply_user(std::array<halo_descriptor, 3> const &halos, BC const &bc, uint32_t mask, void *stream,
                Stores const &... stores) {
                const uregions t = make_regions(halos, mask);
                if (t.blocks == 0)
                    return;
                const std::array<int, 3> need = {
                    halos[0].total_length(), halos[1].total_length(), halos[2].total_length()};
                // calls a deferred-run scope recorded on this stream run first
                if (gt355_batch_flush(stream) != GT355_OK)
                    throw std::runtime_error(std::string("boundary::apply: ") + gt355_last_error());
                user_bc_kernel<<<t.blocks, BLOCK, 0, (hipStream_t)stream>>>(bc, t, view_of(stores, need)...);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess)
                    throw std::runtime_error(std::string("boundary::apply: ") + hipGetErrorString(e));
            }
        } // namespace bc_impl
    }     // namespace boundaries
} // namespace gt355
