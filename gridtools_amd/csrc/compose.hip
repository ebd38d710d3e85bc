// Stencil compositions: several multistages of one run().
//
// hdiff -> vadv: the horizontal diffusion of a field feeds the vertical advection as its u_stage
// (u_stage = hdiff(in, coeff); then vadv(utens_stage, u_stage, wcon, u_pos, utens)).  In the reference this is
// one spec with multi_pass(hdiff MSS, forward MSS, backward MSS); its gpu backend launches the parallel hdiff
// multistage and the fused vadv multistages separately (stencil/gpu/entry_point.hpp:96-124: a multistage whose
// inputs are read with a horizontal offset or a k offset cannot join the previous launch), with u_stage making an
// HBM round trip in between.
//
// hv_fused_kernel runs the whole composition in one launch.  A workgroup of four wavefronts owns a 64x4 tile of
// columns, one column per lane, with ccol/dcol of the column in registers (one wavefront per SIMD, as in the
// vadv kernel).  At vertical step k it evaluates hdiff at level k+1 for its tile: the in tile with its 2-point
// rim is staged in LDS (double-buffered), lap is an LDS ij-cache, and out is evaluated per column; that value is
// stored once as the u_stage output and passed to the Thomas sweep in the compute type, so u_stage never makes the
// HBM round trip.  In the precision form (FP32 storage, FP64 compute) this also keeps the intermediate in FP64:
// only the final outputs are rounded, which is what keeps the ill-conditioned solve within 1e-5 of FP64.
// All loads run D levels ahead through static register rings (the loops are fully unrolled over the k bucket
// NK, addresses are clamped into the valid box, so the in-order load counter stays exact).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "functors.hpp"
#include "strip.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {

        template <class ST, class CT, int NK, int D, bool UPOS_LDS, int TJ>
        __global__ void __launch_bounds__(64 * TJ) __attribute__((amdgpu_waves_per_eu(1, 1)))
        hv_fused_kernel(const ST *__restrict__ in_f, const ST *__restrict__ co_f, ST *__restrict__ ust_f,
            ST *__restrict__ uts_f, const ST *__restrict__ wc_f, const ST *__restrict__ up_f,
            const ST *__restrict__ ut_f, CT dtr, int ni, int nj, int nk, int64_t isj, int64_t isk, int64_t csj,
            int64_t csk, int64_t osj, int64_t osk, int64_t sj, int64_t sk, int64_t wsj, int64_t wsk, int tiles_i,
            int tiles_j, int nxcd) {
            constexpr int NT = 64 * TJ, TI = 64, IW = TI + 4, IH = TJ + 4, LW = TI + 2;
            constexpr int NIN = IW * IH;                  // in tile with its 2-point rim: 68 x (TJ + 4)
            constexpr int IN_PER = (NIN + NT - 1) / NT;   // tile elements per thread
            constexpr int NLU = LW * (TJ + 2);            // lap points of the tile and its 1-point rim
            constexpr int LAP_PER = (NLU + NT - 1) / NT;  // lap points per thread (surplus threads redo the last)
            __shared__ ST s_in[2][IN_PER * NT];
            __shared__ CT s_lap[NLU];
            __shared__ ST s_up[UPOS_LDS ? NK : 1][NT];

            // one wavefront (TJ = 1) or four (TJ = 4) per workgroup; with one, the workgroup barriers below
            // are wave barriers and every wavefront streams on its own
            const int tid = threadIdx.x;
            const int x = tid & 63, y = tid >> 6;
            const int id = xcd_remap(blockIdx.x, tiles_i * tiles_j, nxcd);
            const int i0 = (id % tiles_i) * TI, j0 = (id / tiles_i) * TJ;
            // the column this thread evaluates: lanes outside the domain evaluate the nearest interior column, so
            // their (identical) stores are harmless
            const int xe = min(x, ni - 1 - i0), ye = min(y, nj - 1 - j0);
            const int i = i0 + xe, j = j0 + ye;
            const int top = nk - 1;
            auto lvl = [&](int L) { return (int64_t)min(L, top); };

            // in-tile slots of this thread, clamped into the readable box [-2, n+2)
            int in_off[IN_PER];
#pragma clang loop unroll(full)
            for (int r = 0; r < IN_PER; ++r) {
                const int e = min(tid + r * NT, NIN - 1);
                const int gi = min(i0 - 2 + e % IW, ni + 1), gj = min(j0 - 2 + e / IW, nj + 1);
                in_off[r] = gi + gj * (int)isj;
            }
            const int64_t co_off = i + (int64_t)j * csj;
            const int64_t col = (int64_t)j * sj + i, wcol = (int64_t)j * wsj + i;
            const int64_t ocol = (int64_t)j * osj + i;

            // opaque per-step strides: the loads' addresses only exist from their step on (see vadv.hip)
            int64_t isk_o = isk, csk_o = csk, sk_o = sk, wsk_o = wsk;
            auto pin = [&]() {
                asm volatile("" : "+s"(isk_o));
                asm volatile("" : "+s"(csk_o));
                asm volatile("" : "+s"(sk_o));
                asm volatile("" : "+s"(wsk_o));
            };

            ST h_in[IN_PER][D], h_co[D];                      // hdiff inputs of level L (consumed at step L-1)
            ST b_uts[D], b_up[D], b_ut[D];                    // vadv inputs of level L (consumed at step L)
            ST a_wi[D], a_wip[D];                             // wcon of level L (consumed at step L-1)
            auto load_h = [&](int L) {
                const int64_t l = lvl(L);
#pragma clang loop unroll(full)
                for (int r = 0; r < IN_PER; ++r)
                    h_in[r][L % D] = in_f[l * isk_o + in_off[r]];
                h_co[L % D] = co_f[l * csk_o + co_off];
            };
            auto load_b = [&](int L) {
                const int64_t l = lvl(L);
                b_uts[L % D] = uts_f[l * sk_o + col];
                b_up[L % D] = up_f[l * sk_o + col];
                b_ut[L % D] = ut_f[l * sk_o + col];
            };
            auto load_a = [&](int L) {
                const int64_t l = lvl(L);
                a_wi[L % D] = wc_f[l * wsk_o + wcol];
                a_wip[L % D] = wc_f[l * wsk_o + wcol + 1];
            };

            // hdiff at level L for this thread's column (horizontal_diffusion.cpp:41-46, 56-60, 70-74, 86-90);
            // the value is also stored as the u_stage output
            auto hdiff_level = [&](int L) -> CT {
                const int b = L & 1;
#pragma clang loop unroll(full)
                for (int r = 0; r < IN_PER; ++r)
                    s_in[b][tid + r * NT] = h_in[r][L % D];
                const CT coeff = CT(h_co[L % D]);
                __syncthreads();
#pragma clang loop unroll(full)
                for (int q = 0; q < LAP_PER; ++q) {
                    const int idx = min(tid + q * NT, NLU - 1);
                    const int lx = idx % LW + 1, ly = idx / LW + 1; // in-tile coordinates of the lap point
                    const ST *p = &s_in[b][ly * IW + lx];
                    s_lap[idx] = lap_f<CT>(CT(p[0]), CT(p[1]), CT(p[IW]), CT(p[-1]), CT(p[-IW]));
                }
                __syncthreads();
                const ST *ip = &s_in[b][(ye + 2) * IW + xe + 2];
                const CT *lp = &s_lap[(ye + 1) * LW + xe + 1];
                const CT c = CT(ip[0]), lc = lp[0];
                const CT flx_c = flux_f<CT>(lp[1], lc, CT(ip[1]), c);
                const CT flx_m = flux_f<CT>(lc, lp[-1], c, CT(ip[-1]));
                const CT fly_c = flux_f<CT>(lp[LW], lc, CT(ip[IW]), c);
                const CT fly_m = flux_f<CT>(lc, lp[-LW], c, CT(ip[-IW]));
                const CT res = c - coeff * (((flx_c - flx_m) + fly_c) - fly_m);
                ust_f[lvl(L) * osk + ocol] = ST(res);
                return res;
            };

#pragma clang loop unroll(full)
            for (int L = 0; L < D; ++L) {
                load_h(L);
                load_b(L);
                load_a(L + 1);
            }
            CT us_c = hdiff_level(0);
            pin();
            load_h(D);

            CT c[NK], d[NK];
            CT us_m = CT(0), w_c_i = CT(0), w_c_ip = CT(0);
#pragma clang loop unroll(full)
            for (int k = 0; k < NK; ++k) {
                const bool at_top = k == top;
                // u_stage at k+1 from the diffusion of level k+1 (at the top: the top level again; not read)
                const CT us_next = hdiff_level(k + 1);
                if constexpr (UPOS_LDS)
                    s_up[k][tid] = b_up[k % D];
                const CT v_uts = CT(b_uts[k % D]), v_up = CT(b_up[k % D]), v_ut = CT(b_ut[k % D]);
                // top level: k+1 wind taken as zero and us_p = us_c, which reproduces last_level exactly
                const CT us_p = at_top ? us_c : us_next;
                const CT w_p_i = at_top ? CT(0) : CT(a_wi[(k + 1) % D]);
                const CT w_p_ip = at_top ? CT(0) : CT(a_wip[(k + 1) % D]);
                pin();
                if (k + 1 + D <= NK) // hdiff levels up to NK are consumed (the last ones clamped to the top)
                    load_h(k + 1 + D);
                if (k + D < NK)
                    load_b(k + D);
                if (k + 1 + D < NK)
                    load_a(k + 1 + D);
                if (k == 0) {
                    fwd_out<CT> r = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                    c[0] = r.c;
                    d[0] = r.d;
                } else {
                    const CT cm = c[k > 0 ? k - 1 : 0], dm = d[k > 0 ? k - 1 : 0];
                    fwd_out<CT> r = forward_body<CT>(
                        dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, v_up, v_ut, v_uts, cm, dm, at_top);
                    c[k] = r.c;
                    d[k] = r.d;
                }
                us_m = us_c;
                us_c = us_p;
                w_c_i = w_p_i;
                w_c_ip = w_p_ip;
            }

            ST r_up[D];
            if constexpr (!UPOS_LDS) {
#pragma clang loop unroll(full)
                for (int s = 0; s < D; ++s)
                    r_up[(NK - 1 - s) % D] = up_f[lvl(NK - 1 - s) * sk + col];
            }
            CT data = CT(0);
#pragma clang loop unroll(full)
            for (int k = NK - 1; k >= 0; --k) {
                CT v_up;
                if constexpr (UPOS_LDS) {
                    v_up = CT(s_up[k][tid]);
                } else {
                    v_up = CT(r_up[k % D]);
                    pin();
                    if (k - D >= 0)
                        r_up[k % D] = up_f[lvl(k - D) * sk_o + col];
                }
                // u_backward_function (vertical_advection_dycore.cpp:110-121)
                const CT dt = d[k] - c[k] * data;
                const CT nd = k == top ? d[k] : dt;
                const CT out = dtr * (nd - v_up);
                data = nd;
                uts_f[lvl(k) * sk_o + col] = ST(out);
            }
        }


        // hv_dpp_kernel: the composition with no LDS tile and no barrier (the strip sweep of strip.hpp).  One
        // wavefront per workgroup owns 64 columns of one row; NARROW (ni not a multiple of 64) lets the lanes past
        // ni of the last strip repeat column ni-1.
        template <class ST, class CT, int NK, int D, bool NARROW>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        hv_dpp_kernel(strip_fields<ST> f, CT dtr, int ni, int nj, int nk, int strips_i, int nxcd) {
            __shared__ ST s_up[NK][64];
            // strips fastest: the waves in flight cover whole rows (contiguous DRAM pages); the five waves that read
            // a row of `in` are strips_i ids apart, on one XCD
            const int id = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            const int si = id % strips_i, j = id / strips_i;
            hv_strip<ST, CT, NK, D, NARROW, false, false>(f, dtr, ni, nj, nk, si * 64, j, s_up, nullptr);
        }

        // largest byte offset a strip sweep forms in any of its fields (buffer offsets are 32-bit)
        inline int64_t strip_span(const gt355_field *f, int elem, int ni, int nj, int nk, int lo, int hi_i) {
            return (((int64_t)(nk - 1)) * f->stride[2] + (int64_t)(nj + 1 - lo) * f->stride[1] + hi_i - lo) * elem;
        }

        template <class ST, class CT>
        int launch_dpp(const gt355_field *in, const gt355_field *coeff, const gt355_field *u_stage,
            const gt355_field *uts, const gt355_field *wc, const gt355_field *up, const gt355_field *ut, CT dtr,
            int ni, int nj, int nk, hipStream_t stream) {
            const int strips_i = (ni + 63) / 64;
            const int64_t blocks = (int64_t)strips_i * nj;
            GT355_CHECK(blocks < (1LL << 31), "gt355_hdiff_vadv: domain too large");
            const int64_t lim = int64_t(1) << 32;
            GT355_CHECK(strip_span(in, sizeof(ST), ni, nj, nk, -2, ni + 2) < lim &&
                            strip_span(coeff, sizeof(ST), ni, nj, nk, 0, ni) < lim &&
                            strip_span(u_stage, sizeof(ST), ni, nj, nk, 0, ni) < lim &&
                            strip_span(uts, sizeof(ST), ni, nj, nk, 0, ni) < lim &&
                            strip_span(up, sizeof(ST), ni, nj, nk, 0, ni) < lim &&
                            strip_span(ut, sizeof(ST), ni, nj, nk, 0, ni) < lim &&
                            strip_span(wc, sizeof(ST), ni, nj, nk, 0, ni + 1) < lim,
                "gt355_hdiff_vadv: the fused strip kernel addresses fields below 4 GiB");
            strip_fields<ST> f{(const ST *)in->origin, (const ST *)coeff->origin, (const ST *)wc->origin,
                (const ST *)up->origin, (const ST *)ut->origin, (ST *)u_stage->origin, (ST *)uts->origin, nullptr,
                in->stride[1], in->stride[2], coeff->stride[1], coeff->stride[2], u_stage->stride[1],
                u_stage->stride[2], uts->stride[1], uts->stride[2], wc->stride[1], wc->stride[2]};
#define GT355_DPP(NKB)                                                                                           \
    if (nk <= NKB) {                                                                                             \
        if (ni % 64 != 0)                                                                                        \
            hv_dpp_kernel<ST, CT, NKB, 4, true><<<(int)blocks, 64, 0, stream>>>(f, dtr, ni, nj, nk, strips_i,     \
                xcd_count());                                                                                    \
        else                                                                                                     \
            hv_dpp_kernel<ST, CT, NKB, 4, false><<<(int)blocks, 64, 0, stream>>>(f, dtr, ni, nj, nk, strips_i,    \
                xcd_count());                                                                                    \
        return check_launch("gt355_hdiff_vadv");                                                                 \
    }
            GT355_DPP(16)
            GT355_DPP(32)
            GT355_DPP(48)
            GT355_DPP(64)
            GT355_DPP(80)
#undef GT355_DPP
            set_error("gt355_hdiff_vadv: nk %d exceeds the fused kernel's deepest bucket", nk);
            return GT355_EUNSUP;
        }

        template <class ST, class CT, int TJ>
        int launch_fused_tj(const gt355_field *in, const gt355_field *coeff, const gt355_field *u_stage,
            const gt355_field *uts, const gt355_field *wc, const gt355_field *up, const gt355_field *ut, CT dtr,
            int ni, int nj, int nk, hipStream_t stream) {
            const int tiles_i = (ni + 63) / 64, tiles_j = (nj + TJ - 1) / TJ;
            const int64_t blocks = (int64_t)tiles_i * tiles_j;
            GT355_CHECK(blocks < (1LL << 31), "gt355_hdiff_vadv: domain too large");
            // u_pos parked in LDS in FP32 storage (80 levels x 64 TJ columns x 4 B); in FP64 it is re-read
            constexpr bool UPOS_LDS = sizeof(ST) == 4;
#define GT355_FUSED(NKB, DD)                                                                                     \
    if (nk <= NKB) {                                                                                             \
        hv_fused_kernel<ST, CT, NKB, DD, UPOS_LDS, TJ><<<(int)blocks, 64 * TJ, 0, stream>>>(                     \
            (const ST *)in->origin, (const ST *)coeff->origin, (ST *)u_stage->origin, (ST *)uts->origin,         \
            (const ST *)wc->origin, (const ST *)up->origin, (const ST *)ut->origin, dtr, ni, nj, nk,             \
            in->stride[1], in->stride[2], coeff->stride[1], coeff->stride[2], u_stage->stride[1],                \
            u_stage->stride[2], uts->stride[1], uts->stride[2], wc->stride[1], wc->stride[2], tiles_i, tiles_j,  \
            xcd_count());                                                                                        \
        return check_launch("gt355_hdiff_vadv");                                                                 \
    }
            constexpr int DD = sizeof(ST) == 8 ? 4 : 8;
            GT355_FUSED(16, (DD < 16 ? DD : 16))
            GT355_FUSED(32, DD)
            GT355_FUSED(48, DD)
            GT355_FUSED(64, DD)
            GT355_FUSED(80, DD)
#undef GT355_FUSED
            set_error("gt355_hdiff_vadv: nk %d exceeds the fused kernel's deepest bucket", nk);
            return GT355_EUNSUP;
        }

        // tile height: 4 rows (default) or 1 row (one wavefront per workgroup, no cross-wave barriers; measured
        // slower on MI355X: with one wavefront per SIMD the per-level LDS round trips of a lone wavefront are
        // not hidden)
        template <class ST, class CT>
        int launch_fused(const gt355_field *in, const gt355_field *coeff, const gt355_field *u_stage,
            const gt355_field *uts, const gt355_field *wc, const gt355_field *up, const gt355_field *ut, CT dtr,
            int ni, int nj, int nk, hipStream_t stream) {
            if (tune_param("hv_fused", 0) == 2)
                return launch_dpp<ST, CT>(in, coeff, u_stage, uts, wc, up, ut, dtr, ni, nj, nk, stream);
            if (tune_param("hv_tj", 4) == 1)
                return launch_fused_tj<ST, CT, 1>(in, coeff, u_stage, uts, wc, up, ut, dtr, ni, nj, nk, stream);
            return launch_fused_tj<ST, CT, 4>(in, coeff, u_stage, uts, wc, up, ut, dtr, ni, nj, nk, stream);
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_hdiff_vadv(int precision, const gt355_field *in, const gt355_field *coeff,
    const gt355_field *u_stage, const gt355_field *utens_stage, const gt355_field *wcon, const gt355_field *u_pos,
    const gt355_field *utens, double dtr_stage, int ni, int nj, int nk, void *stream) {
    GT355_CHECK(in && coeff && u_stage && utens_stage && wcon && u_pos && utens, "gt355_hdiff_vadv: null field");
    GT355_CHECK(ni >= 0 && nj >= 0, "gt355_hdiff_vadv: negative size");
    GT355_CHECK(u_stage->origin != in->origin, "gt355_hdiff_vadv: u_stage must not alias in");
    GT355_CHECK(nk >= 2, "gt355_hdiff_vadv: nk = %d, the vertical sweep needs at least 2 levels", nk);
    if (ni == 0 || nj == 0)
        return GT355_OK;
    const int dt = in->dtype;
    GT355_CHECK(dt == GT355_F64 || dt == GT355_F32, "gt355_hdiff_vadv: dtype %d", dt);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dt == GT355_F32),
        "gt355_hdiff_vadv: precision %d with dtype %d", precision, dt);
    // validate every field before anything launches, so a bad call changes nothing
    int rc;
    if ((rc = check_field(in, "in", dt, -2, ni + 2, -2, nj + 2, 0, nk)) ||
        (rc = check_field(coeff, "coeff", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(u_stage, "u_stage", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(utens_stage, "utens_stage", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(wcon, "wcon", dt, 0, ni + 1, 0, nj, 0, nk)) ||
        (rc = check_field(u_pos, "u_pos", dt, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(utens, "utens", dt, 0, ni, 0, nj, 0, nk)))
        return rc;
    GT355_CHECK(utens_stage->stride[1] == u_pos->stride[1] && utens_stage->stride[2] == u_pos->stride[2] &&
                    utens_stage->stride[1] == utens->stride[1] && utens_stage->stride[2] == utens->stride[2],
        "gt355_hdiff_vadv: utens_stage, u_pos and utens must share one layout");
    if (defer_active(stream)) {
        // inside a deferred-run scope: recorded as its hdiff and vadv parts (the two-launch composition)
        if ((rc = defer_hdiff(precision, u_stage, in, coeff, ni, nj, nk, stream)))
            return rc;
        return defer_vadv(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream);
    }
    hipStream_t s = (hipStream_t)stream;
    // the two-launch sequence of the reference backend (default: on MI355X at 512x512x80 it takes 300 us in FP64
    // and 180 us in the precision form, the fused launch 337 us and 268 us, because the fused kernel holds
    // ccol/dcol in registers at one wavefront per SIMD and its per-level hdiff barriers are then not hidden), or
    // the fused single launch (GT355_TUNE_hv_fused=1), which keeps u_stage in the compute type
    if (nk <= 80 && tune_param("hv_fused", 0)) {
        if (dt == GT355_F64)
            return launch_fused<double, double>(in, coeff, u_stage, utens_stage, wcon, u_pos, utens,
                double(dtr_stage), ni, nj, nk, s);
        if (precision == GT355_WIDEN)
            return launch_fused<float, double>(in, coeff, u_stage, utens_stage, wcon, u_pos, utens,
                double(float(dtr_stage)), ni, nj, nk, s);
        return launch_fused<float, float>(in, coeff, u_stage, utens_stage, wcon, u_pos, utens, float(dtr_stage), ni,
            nj, nk, s);
    }
    rc = gt355_hdiff(precision, u_stage, in, coeff, ni, nj, nk, stream);
    if (rc)
        return rc;
    return gt355_vadv(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream);
}
