// Horizontal diffusion and laplacian for gfx950.
//
// The reference runs hdiff as one multistage execute_parallel().ij_cached(lap, flx, fly) with four stages
// (tests/regression/horizontal_diffusion.cpp:35-107); its gpu backend launches one 64x8 block per
// (tile, k-level), adds halo warps for the stage extents and keeps lap/flx/fly in shared memory
// (stencil/gpu/launch_kernel.hpp:82-166, gpu/ij_cache.hpp:26-47).
//
// Here one workgroup owns a TI x TJ tile and a chunk of KC consecutive k-levels:
//   * the in tile with its 2-point rim is staged in LDS once per level (loaded one level ahead into
//     registers, so the HBM loads of level k+1 overlap the arithmetic of level k);
//   * lap over the tile plus a 1-point rim is an LDS ij-cache (the lap_function extent <-1,1,-1,1>);
//   * flx/fly are evaluated where out needs them, straight from the lap and in caches, so the four
//     stages cost two barriers per level instead of four;
//   * coeff is prefetched one level ahead like in; out is written with one coalesced row store.
// Operation order is exactly the functors' (no FP contraction: the library is built with
// -ffp-contract=off), so FP64 results are bit-identical to an IEEE CPU evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "functors.hpp"
#include "sweep.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {

        template <class ST, class CT, int TI, int TJ, int KC, int NT>
        __global__ void __launch_bounds__(NT) hdiff_kernel(ST *__restrict__ out, const ST *__restrict__ in,
            const ST *__restrict__ coeff, int ni, int nj, int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk,
            int64_t csj, int64_t csk, int tiles_i, int tiles_j, int nchunks, int nxcd) {
            constexpr int IW = TI + 4, IH = TJ + 4; // in tile with 2-point rim
            constexpr int LW = TI + 2, LH = TJ + 2; // lap tile with 1-point rim
            constexpr int IN_ITERS = (IW * IH + NT - 1) / NT;
            constexpr int OUT_ITERS = (TI * TJ + NT - 1) / NT;
            static_assert((TI * TJ) % NT == 0, "out tile must split evenly over the workgroup");

            __shared__ CT s_in[IH][IW];
            __shared__ CT s_lap[LH][LW];

            const int total = tiles_i * tiles_j * nchunks;
            const int id = xcd_remap(blockIdx.x, total, nxcd);
            const int ti = id % tiles_i;
            const int tj = (id / tiles_i) % tiles_j;
            const int kc = id / (tiles_i * tiles_j);
            const int i0 = ti * TI, j0 = tj * TJ;
            const int k0 = kc * KC;
            const int k1 = min(k0 + KC, nk);
            const int tid = threadIdx.x;

            // per-thread in-tile slots (rim included), clipped to the valid box [-2, n+2)
            int in_off[IN_ITERS];
            bool in_ok[IN_ITERS];
#pragma unroll
            for (int r = 0; r < IN_ITERS; ++r) {
                int idx = tid + r * NT;
                int x = idx % IW, y = idx / IW;
                int gi = i0 + x - 2, gj = j0 + y - 2;
                in_ok[r] = idx < IW * IH && gi < ni + 2 && gj < nj + 2;
                in_off[r] = in_ok[r] ? (x - 2) + (y - 2) * (int)isj : 0; // relative to the tile origin in_t
            }
            int out_x[OUT_ITERS], out_y[OUT_ITERS];
            bool out_ok[OUT_ITERS];
#pragma unroll
            for (int r = 0; r < OUT_ITERS; ++r) {
                int idx = tid + r * NT;
                out_x[r] = idx % TI;
                out_y[r] = idx / TI;
                out_ok[r] = idx < TI * TJ && i0 + out_x[r] < ni && j0 + out_y[r] < nj;
            }

            ST pin[IN_ITERS];
            ST pco[OUT_ITERS];
            const ST *in_t = in + i0 + (int64_t)j0 * isj;
            const ST *co_t = coeff + i0 + (int64_t)j0 * csj;
            auto prefetch = [&](int k) {
#pragma unroll
                for (int r = 0; r < IN_ITERS; ++r)
                    pin[r] = in_ok[r] ? in_t[in_off[r] + k * isk] : ST(0);
#pragma unroll
                for (int r = 0; r < OUT_ITERS; ++r)
                    pco[r] = out_ok[r] ? __builtin_nontemporal_load(co_t + out_x[r] + out_y[r] * csj + k * csk)
                                       : ST(0);
            };
            if (k0 < k1)
                prefetch(k0);

            for (int k = k0; k < k1; ++k) {
                // stage the prefetched in tile; keep this level's coeff in registers
#pragma unroll
                for (int r = 0; r < IN_ITERS; ++r) {
                    int idx = tid + r * NT;
                    if (idx < IW * IH)
                        (&s_in[0][0])[idx] = static_cast<CT>(pin[r]);
                }
                CT cco[OUT_ITERS];
#pragma unroll
                for (int r = 0; r < OUT_ITERS; ++r)
                    cco[r] = static_cast<CT>(pco[r]);
                __syncthreads();
                if (k + 1 < k1)
                    prefetch(k + 1);

                // lap over the tile plus a 1-point rim (lap_function, extent <-1,1,-1,1>)
                for (int idx = tid; idx < LW * LH; idx += NT) {
                    int x = idx % LW, y = idx / LW;
                    int sx = x + 1, sy = y + 1;
                    s_lap[y][x] = lap_f<CT>(
                        s_in[sy][sx], s_in[sy][sx + 1], s_in[sy + 1][sx], s_in[sy][sx - 1], s_in[sy - 1][sx]);
                }
                __syncthreads();

                // flx, fly and out per interior point
#pragma unroll
                for (int r = 0; r < OUT_ITERS; ++r) {
                    int x = out_x[r], y = out_y[r];
                    int lx = x + 1, ly = y + 1; // lap coords
                    int sx = x + 2, sy = y + 2; // in coords
                    CT c = s_in[sy][sx];
                    CT lc = s_lap[ly][lx];
                    CT flx_c = flux_f<CT>(s_lap[ly][lx + 1], lc, s_in[sy][sx + 1], c);
                    CT flx_m = flux_f<CT>(lc, s_lap[ly][lx - 1], c, s_in[sy][sx - 1]);
                    CT fly_c = flux_f<CT>(s_lap[ly + 1][lx], lc, s_in[sy + 1][sx], c);
                    CT fly_m = flux_f<CT>(lc, s_lap[ly - 1][lx], c, s_in[sy - 1][sx]);
                    CT res = c - cco[r] * (((flx_c - flx_m) + fly_c) - fly_m);
                    if (out_ok[r])
                        __builtin_nontemporal_store(static_cast<ST>(res),
                            out + (i0 + x) + (int64_t)(j0 + y) * osj + (int64_t)k * osk);
                }
                __syncthreads();
            }
        }

        // Register sweep variant: one wavefront per 64 columns x JB rows x 1 level (sweep.hpp: hdiff_sweep_wave).
        template <class ST, class CT, int JB, int P>
        __global__ void __launch_bounds__(256) hdiff_sweep_kernel(ST *__restrict__ out, const ST *__restrict__ in,
            const ST *__restrict__ coeff, int ni, int nj, int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk,
            int64_t csj, int64_t csk, int strips_i, int strips_j, int nblocks, int nxcd) {
            // wave-uniform by construction; readfirstlane tells the compiler, so the side columns become scalar loads
            const int wid = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, nblocks, nxcd) * 4 + (threadIdx.x >> 6));
            const int total = strips_i * strips_j * nk;
            if (wid >= total)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            // strips fastest: the waves in flight cover whole rows of a level (contiguous DRAM pages); the four
            // waves of a workgroup take four neighbouring strips of one row block
            const int si_ = wid % strips_i;
            const int sj_ = (wid / strips_i) % strips_j;
            const int k = wid / (strips_j * strips_i);
            hdiff_sweep_wave<ST, CT, JB, P, false>(out, in, coeff, (ST *)nullptr, ni, nj, osj, osk, isj, isk, csj, csk,
                si_ * 64, sj_ * JB, k);
        }

        // Column-pair register sweep: one wavefront per 128 columns x JB rows x 1 level (sweep.hpp: hdiff_pair_wave).
        template <class ST, class CT, int JB, int P>
        __global__ void __launch_bounds__(256) hdiff_sweep2_kernel(ST *__restrict__ out, const ST *__restrict__ in,
            const ST *__restrict__ coeff, int ni, int nj, int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk,
            int64_t csj, int64_t csk, int strips_i, int strips_j, int nblocks, int nxcd) {
            const int wid = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, nblocks, nxcd) * 4 + (threadIdx.x >> 6));
            const int total = strips_i * strips_j * nk;
            if (wid >= total)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            const int si_ = wid % strips_i;
            const int sj_ = (wid / strips_i) % strips_j;
            const int k = wid / (strips_j * strips_i);
            hdiff_pair_wave<ST, CT, JB, P, false>(out, in, coeff, (ST *)nullptr, ni, nj, osj, osk, isj, isk, csj, csk,
                si_ * 128, sj_ * JB, k);
        }

        // laplacian.cpp:20-29, one point per thread, one k-level per grid z
        template <class ST, class CT>
        __global__ void __launch_bounds__(256) lap_kernel(ST *__restrict__ out, const ST *__restrict__ in, int ni,
            int nj, int64_t osj, int64_t osk, int64_t isj, int64_t isk) {
            int i = blockIdx.x * 64 + (threadIdx.x & 63);
            int j = blockIdx.y * 4 + (threadIdx.x >> 6);
            int k = blockIdx.z;
            if (i >= ni || j >= nj)
                return;
            const ST *p = in + i + j * isj + (int64_t)k * isk;
            CT r = lap_f<CT>(CT(p[0]), CT(p[1]), CT(p[isj]), CT(p[-1]), CT(p[-isj]));
            out[i + j * osj + (int64_t)k * osk] = static_cast<ST>(r);
        }

        template <class ST, class CT, int TI, int TJ, int KC>
        int launch_hdiff(const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni, int nj,
            int nk, hipStream_t stream) {
            constexpr int NT = 256;
            int tiles_i = (ni + TI - 1) / TI, tiles_j = (nj + TJ - 1) / TJ, nchunks = (nk + KC - 1) / KC;
            int64_t total = (int64_t)tiles_i * tiles_j * nchunks;
            GT355_CHECK(total < (1LL << 31), "gt355_hdiff: domain too large for one launch");
            hdiff_kernel<ST, CT, TI, TJ, KC, NT><<<(int)total, NT, 0, stream>>>((ST *)out->origin,
                (const ST *)in->origin, (const ST *)coeff->origin, ni, nj, nk, out->stride[1], out->stride[2],
                in->stride[1], in->stride[2], coeff->stride[1], coeff->stride[2], tiles_i, tiles_j, nchunks,
                xcd_count());
            return check_launch("gt355_hdiff");
        }

        template <class ST, class CT, int JB, int P>
        int launch_hdiff_sweep(const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
            int nj, int nk, hipStream_t stream) {
            int strips_i = (ni + 63) / 64, strips_j = (nj + JB - 1) / JB;
            int64_t waves = (int64_t)strips_i * strips_j * nk;
            int64_t blocks = (waves + 3) / 4;
            GT355_CHECK(blocks < (1LL << 31), "gt355_hdiff: domain too large for one launch");
            hdiff_sweep_kernel<ST, CT, JB, P><<<(int)blocks, 256, 0, stream>>>((ST *)out->origin,
                (const ST *)in->origin, (const ST *)coeff->origin, ni, nj, nk, out->stride[1], out->stride[2],
                in->stride[1], in->stride[2], coeff->stride[1], coeff->stride[2], strips_i, strips_j, (int)blocks,
                xcd_count());
            return check_launch("gt355_hdiff");
        }

        // The column-pair sweep needs an even ni and rows that start 2-element aligned in all three fields.
        inline bool pairs_ok(const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
            int64_t es) {
            auto al = [&](const gt355_field *f) {
                return (uintptr_t)f->origin % (2 * es) == 0 && f->stride[1] % 2 == 0 && f->stride[2] % 2 == 0;
            };
            return ni % 2 == 0 && ni >= 2 && al(out) && al(in) && al(coeff);
        }

        template <class ST, class CT, int JB, int P>
        int launch_hdiff_sweep2(const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni,
            int nj, int nk, hipStream_t stream) {
            if (!pairs_ok(out, in, coeff, ni, sizeof(ST)))
                return launch_hdiff_sweep<ST, CT, JB, P>(out, in, coeff, ni, nj, nk, stream);
            int strips_i = (ni + 127) / 128, strips_j = (nj + JB - 1) / JB;
            int64_t waves = (int64_t)strips_i * strips_j * nk;
            int64_t blocks = (waves + 3) / 4;
            GT355_CHECK(blocks < (1LL << 31), "gt355_hdiff: domain too large for one launch");
            hdiff_sweep2_kernel<ST, CT, JB, P><<<(int)blocks, 256, 0, stream>>>((ST *)out->origin,
                (const ST *)in->origin, (const ST *)coeff->origin, ni, nj, nk, out->stride[1], out->stride[2],
                in->stride[1], in->stride[2], coeff->stride[1], coeff->stride[2], strips_i, strips_j, (int)blocks,
                xcd_count());
            return check_launch("gt355_hdiff");
        }

        // Launches one hdiff configuration; GT355_EUNSUP if that shape is not instantiated.
        template <class ST, class CT>
        int launch_cfg(const hdiff_cfg &c, const gt355_field *out, const gt355_field *in, const gt355_field *coeff,
            int ni, int nj, int nk, hipStream_t stream) {
            if (c.impl == 2) {
#define GT355_HSWEEP2(JB, P)                                                                    \
    if (c.a == JB && c.b == P)                                                                  \
        return launch_hdiff_sweep2<ST, CT, JB, P>(out, in, coeff, ni, nj, nk, stream);
                GT355_HDIFF_SWEEPS2(GT355_HSWEEP2)
#undef GT355_HSWEEP2
            } else if (c.impl == 1) {
#define GT355_HSWEEP(JB, P)                                                                     \
    if (c.a == JB && c.b == P)                                                                  \
        return launch_hdiff_sweep<ST, CT, JB, P>(out, in, coeff, ni, nj, nk, stream);
                GT355_HDIFF_SWEEPS(GT355_HSWEEP)
#undef GT355_HSWEEP
            } else {
#define GT355_HDIFF_CASE(TI, TJ, KC)                                                            \
    if (c.a == TI && c.b == TJ && c.c == KC)                                                    \
        return launch_hdiff<ST, CT, TI, TJ, KC>(out, in, coeff, ni, nj, nk, stream);
                GT355_HDIFF_TILES(GT355_HDIFF_CASE)
#undef GT355_HDIFF_CASE
            }
            set_error("gt355_hdiff: configuration impl=%d (%d,%d,%d) is not instantiated", c.impl, c.a, c.b, c.c);
            return GT355_EUNSUP;
        }

        template <class ST, class CT>
        int dispatch_hdiff(const gt355_field *out, const gt355_field *in, const gt355_field *coeff, int ni, int nj,
            int nk, hipStream_t stream) {
            return launch_cfg<ST, CT>(hdiff_cfg_for(sizeof(ST), sizeof(CT), ni, nj, nk), out, in, coeff, ni, nj, nk,
                stream);
        }

        // Every instantiated configuration, in the order the tuner tries them.
        std::vector<hdiff_cfg> hdiff_candidates() {
            std::vector<hdiff_cfg> v;
#define GT355_ADD_TILE(TI, TJ, KC) v.push_back({0, TI, TJ, KC});
            GT355_HDIFF_TILES(GT355_ADD_TILE)
#undef GT355_ADD_TILE
#define GT355_ADD_SWEEP(JB, P) v.push_back({1, JB, P, 0});
            GT355_HDIFF_SWEEPS(GT355_ADD_SWEEP)
#undef GT355_ADD_SWEEP
#define GT355_ADD_SWEEP2(JB, P) v.push_back({2, JB, P, 0});
            GT355_HDIFF_SWEEPS2(GT355_ADD_SWEEP2)
#undef GT355_ADD_SWEEP2
            return v;
        }

    } // namespace

    bool hdiff_cfg_instantiated(hdiff_cfg cfg) {
        for (const hdiff_cfg &c : hdiff_candidates())
            if (c.impl == cfg.impl && c.a == cfg.a && c.b == cfg.b && (cfg.impl != 0 || c.c == cfg.c))
                return true;
        return false;
    }

    namespace {
        int check_hdiff_args(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff,
            int ni, int nj, int nk) {
            GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0, "gt355_hdiff: negative size");
            GT355_CHECK(out && in && coeff, "gt355_hdiff: null field");
            int dtype = in->dtype;
            GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_hdiff: dtype %d", dtype);
            GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dtype == GT355_F32),
                "gt355_hdiff: precision %d with dtype %d", precision, dtype);
            if (ni == 0 || nj == 0 || nk == 0)
                return GT355_OK;
            int rc;
            if ((rc = check_field(in, "in", dtype, -2, ni + 2, -2, nj + 2, 0, nk)))
                return rc;
            if ((rc = check_field(coeff, "coeff", dtype, 0, ni, 0, nj, 0, nk)))
                return rc;
            if ((rc = check_field(out, "out", dtype, 0, ni, 0, nj, 0, nk)))
                return rc;
            GT355_CHECK(out->origin != in->origin, "gt355_hdiff: out must not alias in");
            return GT355_OK;
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff,
    int ni, int nj, int nk, void *stream) {
    int rc = check_hdiff_args(precision, out, in, coeff, ni, nj, nk);
    if (rc || ni == 0 || nj == 0 || nk == 0)
        return rc;
    if (defer_active(stream)) // inside a deferred-run scope: recorded, run at its flush
        return defer_hdiff(precision, out, in, coeff, ni, nj, nk, stream);
    hipStream_t s = (hipStream_t)stream;
    if (in->dtype == GT355_F64)
        return dispatch_hdiff<double, double>(out, in, coeff, ni, nj, nk, s);
    if (precision == GT355_WIDEN)
        return dispatch_hdiff<float, double>(out, in, coeff, ni, nj, nk, s);
    return dispatch_hdiff<float, float>(out, in, coeff, ni, nj, nk, s);
}

extern "C" int gt355_tune_hdiff(int precision, const gt355_field *out, const gt355_field *in, const gt355_field *coeff,
    int ni, int nj, int nk, int reps, void *stream, int32_t best[4], double *best_us) {
    GT355_DEFER_FLUSH(stream);
    int rc = check_hdiff_args(precision, out, in, coeff, ni, nj, nk);
    if (rc)
        return rc;
    GT355_CHECK(ni > 0 && nj > 0 && nk > 0, "gt355_tune_hdiff: empty domain");
    GT355_CHECK(reps >= 1 && reps <= 10000, "gt355_tune_hdiff: reps %d", reps);
    hipStream_t s = (hipStream_t)stream;
    const int st = in->dtype == GT355_F64 ? 8 : 4, ct = (in->dtype == GT355_F64 || precision == GT355_WIDEN) ? 8 : 4;
    hipEvent_t e0, e1;
    GT355_HIP(hipEventCreate(&e0));
    GT355_HIP(hipEventCreate(&e1));
    hdiff_cfg win{};
    double win_us = 1e30;
    // Timing as a streaming caller sees it: back-to-back launches that cycle through enough copies of the fields
    // to exceed the 256 MB MALL twice (repeated launches on one small field set would run from the MALL and favour
    // configurations whose rim re-reads are only free when cached).  The copies live for this call only.
    struct span_t {
        char *base;
        size_t bytes;
    };
    auto span_of = [](const gt355_field *f) {
        const int64_t es = f->dtype == GT355_F64 ? 8 : 4;
        int64_t lo = 0, hi = 0;
        for (int d = 0; d < 3; ++d) {
            lo += (int64_t)f->lower[d] * f->stride[d];
            hi += (int64_t)(f->upper[d] - 1) * f->stride[d];
        }
        return span_t{(char *)f->origin + lo * es, (size_t)((hi - lo + 1) * es)};
    };
    const span_t sp[3] = {span_of(out), span_of(in), span_of(coeff)};
    const size_t set_bytes = sp[0].bytes + sp[1].bytes + sp[2].bytes;
    const int nsets = (int)std::min<size_t>(8, std::max<size_t>(1, ((size_t)512 << 20) / std::max<size_t>(set_bytes, 1) + 1));
    std::vector<void *> copies;
    std::vector<gt355_field> sets; // out, in, coeff per set
    sets.push_back(*out);
    sets.push_back(*in);
    sets.push_back(*coeff);
    for (int c = 1; c < nsets && !rc; ++c) {
        for (int f = 0; f < 3; ++f) {
            // each copy sits at the same address modulo 256 bytes as its field, so rows keep their alignment
            void *p = nullptr;
            hipError_t e = hipMalloc(&p, sp[f].bytes + 256);
            char *q = (char *)p + (((uintptr_t)sp[f].base - (uintptr_t)p) & 255);
            if (e == hipSuccess)
                e = hipMemcpyAsync(q, sp[f].base, sp[f].bytes, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) {
                (void)hipFree(p);
                rc = hip_fail(e, "gt355_tune_hdiff: field copies");
                break;
            }
            copies.push_back(p);
            gt355_field g = sets[f];
            g.origin = q + ((char *)g.origin - sp[f].base);
            sets.push_back(g);
        }
    }
    auto launch_set = [&](const hdiff_cfg &c, int k) {
        const gt355_field *f = &sets[3 * (k % nsets)];
        if (st == 8)
            return launch_cfg<double, double>(c, &f[0], &f[1], &f[2], ni, nj, nk, s);
        if (ct == 8)
            return launch_cfg<float, double>(c, &f[0], &f[1], &f[2], ni, nj, nk, s);
        return launch_cfg<float, float>(c, &f[0], &f[1], &f[2], ni, nj, nk, s);
    };
    // two passes over the candidates, each candidate's time the faster of its two; the measured default stays
    // unless a candidate beats it by more than 2% (single timings of candidates within noise of each other made the
    // winner a matter of chance, and a chance winner then ran slower than the default)
    const std::vector<hdiff_cfg> cands = hdiff_candidates();
    std::vector<double> best_of(cands.size(), 1e30);
    for (int pass = 0; pass < 2 && !rc; ++pass)
        for (size_t ci = 0; ci < cands.size() && !rc; ++ci) {
            const hdiff_cfg &c = cands[ci];
            for (int k = 0; k < nsets && !rc; ++k) // warm-up (and code load) on every set
                rc = launch_set(c, k);
            if (rc)
                break;
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps && !rc; ++r)
                rc = launch_set(c, r);
            (void)hipEventRecord(e1, s);
            if (rc)
                break;
            hipError_t err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess)
                err = hipEventElapsedTime(&ms, e0, e1);
            if (err != hipSuccess) {
                rc = hip_fail(err, "gt355_tune_hdiff: timing");
                break;
            }
            best_of[ci] = std::min(best_of[ci], 1e3 * ms / reps);
        }
    const hdiff_cfg dflt = hdiff_default_cfg(st, ni);
    double dflt_us = 1e30;
    for (size_t ci = 0; ci < cands.size(); ++ci) {
        const hdiff_cfg &c = cands[ci];
        if (best_of[ci] < win_us)
            win_us = best_of[ci], win = c;
        if (c.impl == dflt.impl && c.a == dflt.a && c.b == dflt.b && (c.impl != 0 || c.c == dflt.c))
            dflt_us = best_of[ci];
    }
    if (dflt_us <= win_us * 1.02)
        win = dflt, win_us = dflt_us;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!copies.empty())
        (void)hipStreamSynchronize(s);
    for (void *p : copies)
        (void)hipFree(p);
    if (rc)
        return rc;
    record_hdiff_cfg(st, ct, ni, nj, nk, win, win_us);
    if (best) {
        best[0] = win.impl;
        best[1] = win.a;
        best[2] = win.b;
        best[3] = win.c;
    }
    if (best_us)
        *best_us = win_us;
    return GT355_OK;
}

extern "C" int gt355_laplacian(int precision, const gt355_field *out, const gt355_field *in, int ni, int nj, int nk,
    void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0, "gt355_laplacian: negative size");
    GT355_CHECK(out && in, "gt355_laplacian: null field");
    if (ni == 0 || nj == 0 || nk == 0)
        return GT355_OK;
    int dtype = in->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_laplacian: dtype %d", dtype);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dtype == GT355_F32),
        "gt355_laplacian: precision %d with dtype %d", precision, dtype);
    GT355_CHECK(nk <= 65535, "gt355_laplacian: nk %d exceeds the grid z limit", nk);
    int rc;
    if ((rc = check_field(in, "in", dtype, -1, ni + 1, -1, nj + 1, 0, nk)))
        return rc;
    if ((rc = check_field(out, "out", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    dim3 grid((ni + 63) / 64, (nj + 3) / 4, nk);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == GT355_F64)
        lap_kernel<double, double><<<grid, 256, 0, s>>>((double *)out->origin, (const double *)in->origin, ni, nj,
            out->stride[1], out->stride[2], in->stride[1], in->stride[2]);
    else if (precision == GT355_WIDEN)
        lap_kernel<float, double><<<grid, 256, 0, s>>>((float *)out->origin, (const float *)in->origin, ni, nj,
            out->stride[1], out->stride[2], in->stride[1], in->stride[2]);
    else
        lap_kernel<float, float><<<grid, 256, 0, s>>>((float *)out->origin, (const float *)in->origin, ni, nj,
            out->stride[1], out->stride[2], in->stride[1], in->stride[2]);
    return check_launch("gt355_laplacian");
}
