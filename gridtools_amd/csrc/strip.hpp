// One wavefront's share of the hdiff -> vadv composition: 64 columns of one row j, all levels.
//
// Used by the single-domain fused kernel (compose.hip) and by the batching form (batch.hip).  The wavefront
// sweeps k once with ccol/dcol in registers for the first NK levels (the whole column when nk <= NK), as
// vadv_onchip_kernel does.  At step k it evaluates hdiff at level k+1 for its row in registers:
//   * rows j-2..j+2 of `in` are per-lane loads of the lane's own column;
//   * the i-1 / i+1 neighbours come from DPP wavefront shifts; lanes 0 and 63 keep the side columns (i0-1 /
//     i0+64, and i0-2 / i0+65 for the side laplacian), which they load themselves;
//   * lap (the ij-cache), flx and fly stay in registers; nothing goes through LDS and there is no barrier.
// The diffusion result is stored as u_stage (and, with LAP, the laplacian as the lap output) and passed to the
// forward Thomas step in the compute type.  u_pos is parked in LDS for the backward sweep.
// TAIL: columns deeper than NK continue in a runtime loop over blocks of D levels whose ccol/dcol go to the
// wavefront's `tail` scratch (64 lanes x (c, d) per level); the backward sweep reads them back, then finishes the
// register part.  With TL > 0 the first TL tail levels use the LDS array `s_tail` instead (TL a multiple of D), so
// that only columns deeper than NK + TL move ccol/dcol through memory.
// Every load and store is unconditional, with levels clamped to the top (a step above the top recomputes or
// overwrites the top with the same bits, or with a value the top's own step overwrites later in program order),
// which keeps the code straight-line and the load counter exact.  Addresses are buffer descriptors plus 32-bit
// byte offsets (the callers check that every field spans less than 4 GiB).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "functors.hpp"

namespace gt355 {

    // v from lane src_lane (ds_bpermute; no LDS memory is used)
    template <class T>
    __device__ __forceinline__ T lane_value(T v, int src_lane) {
        if constexpr (sizeof(T) == 4) {
            return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(src_lane * 4, __builtin_bit_cast(int, v)));
        } else {
            const long long b = __builtin_bit_cast(long long, v);
            const int lo = __builtin_amdgcn_ds_bpermute(src_lane * 4, int(b));
            const int hi = __builtin_amdgcn_ds_bpermute(src_lane * 4, int(b >> 32));
            return __builtin_bit_cast(T, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
        }
    }

    // The composition's fields (origins at grid point (0,0,0)) and their element strides.  lap shares u_stage's
    // layout.
    template <class ST>
    struct strip_fields {
        const ST *in, *co, *wc, *up, *ut;
        ST *ust, *uts, *lap;
        int64_t isj, isk, csj, csk, osj, osk, sj, sk, wsj, wsk;
    };

    // Strips never overlap: utens_stage is read by the forward sweep and rewritten by the backward sweep, so a
    // column must belong to exactly one wavefront.  NARROW: in a strip that ends past ni, lanes past ni take lane
    // ni-1's results and store the same bits to column ni-1 as that lane (ds_bpermute per level); callers pass
    // NARROW = false only when every strip is full (ni a multiple of 64).
    // HD = false: the vertical advection alone, with u_stage read from memory (written beforehand by the batching
    // form's stencil phase) instead of evaluated; nothing but utens_stage is stored.
    // FLAT (with HD = false only): when `flat` is set, the wavefront's columns are 64 consecutive entries of the
    // domain's column list (column n is i = n % ni of row j = n / ni), starting at i0 (j is not used); lanes past the
    // last column repeat it, computing and storing the same bits as its own lane.  Columns are independent in the
    // vertical advection, so no lane idles at the end of a row.
    template <class ST, class CT, int NK, int D, bool NARROW, bool LAP, bool TAIL, bool HD = true, bool FLAT = false,
        int TL = 0>
    __device__ __forceinline__ void hv_strip(const strip_fields<ST> &f, CT dtr, int ni, int nj, int nk, int i0,
        int j, ST (*s_up)[64], CT *tail, bool flat = false, CT *s_tail = nullptr) {
        static_assert(NK % D == 0, "tail blocks start on a ring-slot boundary");
        static_assert(TL % D == 0, "LDS tail blocks are whole ring blocks");
        static_assert(!(FLAT && HD), "the column list form has no horizontal neighbours");
        constexpr int64_t S = sizeof(ST);
        const int lane = threadIdx.x & 63;
        const bool left = lane < 32;
        const int src_lane = min(lane, ni - 1 - i0);
        // hdiff columns: the lane's own (true up to ni+1, so valid lanes see true neighbours) and the side columns
        // lanes 0 and 63 hold; the vertical-advection column (lanes past ni repeat ni-1)
        const int ic = min(i0 + lane, ni + 1);
        const int e1c = left ? i0 - 1 : min(i0 + 64, ni + 1);
        const int e2c = left ? i0 - 2 : min(i0 + 65, ni + 1);
        int iv = min(i0 + lane, ni - 1);
        if constexpr (FLAT) {
            if (flat) {
                const int n = min(i0 + lane, ni * nj - 1);
                j = n / ni;
                iv = n - j * ni;
            }
        }
        const int top = nk - 1;
        auto lvl = [&](int L) { return (uint32_t)min(L, top); };

        // `in` is addressed from its lowest read point (-2, -2, 0), so every byte offset is non-negative
        const rsrc_t r_in = make_rsrc(f.in - 2 - 2 * f.isj), r_co = make_rsrc(f.co), r_wc = make_rsrc(f.wc),
                     r_up = make_rsrc(f.up), r_ut = make_rsrc(f.ut), r_ust = make_rsrc(f.ust),
                     r_uts = make_rsrc(f.uts), r_lap = make_rsrc(LAP ? f.lap : f.ust);
        auto in_off = [&](int dj, int col) {
            return (uint32_t)(((int64_t)(min(j + dj, nj + 1) + 2) * f.isj + col + 2) * S);
        };
        const uint32_t v_c0 = in_off(-2, ic), v_c1 = in_off(-1, ic), v_c2 = in_off(0, ic), v_c3 = in_off(1, ic),
                       v_c4 = in_off(2, ic), v_e1m = in_off(-1, e1c), v_e10 = in_off(0, e1c),
                       v_e1p = in_off(1, e1c), v_e2 = in_off(0, e2c);
        const uint32_t v_co = (uint32_t)(((int64_t)j * f.csj + iv) * S);
        const uint32_t v_col = (uint32_t)(((int64_t)j * f.sj + iv) * S);
        const uint32_t v_wc = (uint32_t)(((int64_t)j * f.wsj + iv) * S);
        const uint32_t v_o = (uint32_t)(((int64_t)j * f.osj + iv) * S);

        // byte strides per level; routed through an opaque asm at every step so that a step's loads only exist
        // from that step on (otherwise the compiler hoists every load of the unrolled sweep to the entry)
        uint32_t isk_b = (uint32_t)(f.isk * S), csk_b = (uint32_t)(f.csk * S), osk_b = (uint32_t)(f.osk * S),
                 sk_b = (uint32_t)(f.sk * S), wsk_b = (uint32_t)(f.wsk * S);
        auto pin = [&]() {
            asm volatile("" : "+s"(isk_b));
            asm volatile("" : "+s"(csk_b));
            asm volatile("" : "+s"(sk_b));
            asm volatile("" : "+s"(wsk_b));
            if constexpr (!HD)
                asm volatile("" : "+s"(osk_b));
        };

        // register rings, level L in slot L % D: hdiff inputs of level L (consumed at step L-1), vadv inputs of
        // level L (consumed at step L), wcon at i and i+1 of level L (consumed at step L-1)
        ST h_c[5][D], h_e1[3][D], h_e2[D], h_co[D];
        ST b_uts[D], b_up[D], b_ut[D];
        ST a_wi[D], a_wip[D];
        auto load_h = [&](int L, int s) {
            if constexpr (!HD) {
                // u_stage of the lane's column (default cache policy: the stencil phase has just written it)
                h_c[2][s] = bload<ST>(r_ust, v_o, lvl(L) * osk_b);
                return;
            }
            const uint32_t o = lvl(L) * isk_b;
            h_c[0][s] = bload<ST>(r_in, v_c0, o);
            h_c[1][s] = bload<ST>(r_in, v_c1, o);
            h_c[2][s] = bload<ST>(r_in, v_c2, o);
            h_c[3][s] = bload<ST>(r_in, v_c3, o);
            h_c[4][s] = bload<ST>(r_in, v_c4, o);
            h_e1[0][s] = bload<ST>(r_in, v_e1m, o);
            h_e1[1][s] = bload<ST>(r_in, v_e10, o);
            h_e1[2][s] = bload<ST>(r_in, v_e1p, o);
            h_e2[s] = bload<ST>(r_in, v_e2, o);
            h_co[s] = bload<ST, CPOL_NT>(r_co, v_co, lvl(L) * csk_b);
        };
        auto load_b = [&](int L, int s) {
            const uint32_t o = lvl(L) * sk_b;
            b_uts[s] = bload<ST, CPOL_NT>(r_uts, v_col, o);
            b_up[s] = bload<ST, CPOL_NT>(r_up, v_col, o);
            b_ut[s] = bload<ST, CPOL_NT>(r_ut, v_col, o);
        };
        auto load_a = [&](int L, int s) {
            // wcon at i and i + 1: column i0 + 64 of a strip is the first column of the next strip's line, so the
            // batch's vertical phase (HD = false) keeps wcon lines in L2 (default policy) for the neighbouring item
            // on the same XCD; non-temporal loads there fetched those lines twice (vertical-phase reads 5.29 ->
            // 5.02 GB per timestep, 2.042 -> 1.998 ms)
            const two<ST> w = HD ? bload2<ST, CPOL_NT>(r_wc, v_wc, lvl(L) * wsk_b)
                                 : bload2<ST>(r_wc, v_wc, lvl(L) * wsk_b);
            a_wi[s] = w.a;
            a_wip[s] = w.b;
        };

        // hdiff at level L for the lane's column (horizontal_diffusion.cpp:41-46, 56-60, 70-74, 86-90), inputs in
        // slot s; stored as u_stage (and lap), returned in the compute type
        auto hdiff_level = [&](int L, int s) -> CT {
            if constexpr (!HD)
                return CT(h_c[2][s]);
            const CT c0 = CT(h_c[0][s]), c1 = CT(h_c[1][s]), c2 = CT(h_c[2][s]), c3 = CT(h_c[3][s]),
                     c4 = CT(h_c[4][s]);
            const CT e1m = CT(h_e1[0][s]), e10 = CT(h_e1[1][s]), e1p = CT(h_e1[2][s]), e2 = CT(h_e2[s]);
            // in(i-1) / in(i+1) of rows j-1, j, j+1
            const CT mm = dpp_keep<DPP_WF_SR1>(e1m, c1), pm = dpp_keep<DPP_WF_SL1>(e1m, c1);
            const CT mc = dpp_keep<DPP_WF_SR1>(e10, c2), pc = dpp_keep<DPP_WF_SL1>(e10, c2);
            const CT mp = dpp_keep<DPP_WF_SR1>(e1p, c3), pp = dpp_keep<DPP_WF_SL1>(e1p, c3);
            const CT lap_m = lap_f<CT>(c1, pm, c2, mm, c0); // lap(i, j-1)
            CT lap_c = lap_f<CT>(c2, pc, c3, mc, c1);       // lap(i, j)
            const CT lap_p = lap_f<CT>(c3, pp, c4, mp, c2); // lap(i, j+1)
            // side laplacian of row j: column i0-1 on the left lanes, i0+64 on the right ones
            const CT sl = lap_f<CT>(e10, left ? c2 : e2, e1p, left ? e2 : c2, e1m);
            const CT lm = dpp_keep<DPP_WF_SR1>(sl, lap_c), lp = dpp_keep<DPP_WF_SL1>(sl, lap_c);
            const CT flx_c = flux_f<CT>(lp, lap_c, pc, c2);
            const CT flx_m = flux_f<CT>(lap_c, lm, c2, mc);
            const CT fly_c = flux_f<CT>(lap_p, lap_c, c3, c2);
            const CT fly_m = flux_f<CT>(lap_c, lap_m, c2, c1);
            CT res = c2 - CT(h_co[s]) * (((flx_c - flx_m) + fly_c) - fly_m);
            if constexpr (NARROW)
                res = lane_value(res, src_lane);
            const uint32_t o = lvl(L) * osk_b;
            bstore<ST, CPOL_OUT>(r_ust, ST(res), v_o, o);
            if constexpr (LAP) {
                if constexpr (NARROW)
                    lap_c = lane_value(lap_c, src_lane);
                bstore<ST, CPOL_OUT>(r_lap, ST(lap_c), v_o, o);
            }
            return res;
        };

#pragma clang loop unroll(full)
        for (int L = 0; L < D; ++L) {
            load_h(L, L % D);
            load_b(L, L % D);
            load_a(L + 1, (L + 1) % D);
        }
        CT us_c = hdiff_level(0, 0);
        pin();
        load_h(D, 0);

        CT c[NK], d[NK];
        CT us_m = CT(0), w_c_i = CT(0), w_c_ip = CT(0), cm = CT(0), dm = CT(0);
        // forward step of level k (slot s = k % D, s1 = (k+1) % D); refills the slots it consumed
        auto fstep = [&](int k, int s, int s1, bool first, bool more) -> fwd_out<CT> {
            const bool at_top = k == top;
            const CT us_next = hdiff_level(k + 1, s1);
            const CT v_uts = CT(b_uts[s]), v_up = CT(b_up[s]), v_ut = CT(b_ut[s]);
            // top level: k+1 wind taken as zero and us_p = us_c, which reproduces last_level exactly
            const CT us_p = at_top ? us_c : us_next;
            const CT w_p_i = at_top ? CT(0) : CT(a_wi[s1]);
            const CT w_p_ip = at_top ? CT(0) : CT(a_wip[s1]);
            pin();
            if (more) {
                load_h(k + 1 + D, s1);
                load_b(k + D, s);
                load_a(k + 1 + D, s1);
            }
            fwd_out<CT> r = first
                                ? forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts)
                                : forward_body<CT>(dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, v_up, v_ut,
                                      v_uts, cm, dm, at_top);
            cm = r.c;
            dm = r.d;
            us_m = us_c;
            us_c = us_p;
            w_c_i = w_p_i;
            w_c_ip = w_p_ip;
            return r;
        };
#pragma clang loop unroll(full)
        for (int k = 0; k < NK; ++k) {
            s_up[k][lane] = b_up[k % D];
            // loads past the last register level are only needed by the tail (and are clamped to the top)
            const fwd_out<CT> r = fstep(k, k % D, (k + 1) % D, k == 0, TAIL || k + 1 + D <= NK);
            c[k] = r.c;
            d[k] = r.d;
            // memory operations stay at their step; arithmetic may move across steps
            __builtin_amdgcn_sched_barrier(6);
        }
        // the tail level n (0-based past NK): the LDS tail for n < TL, the scratch tail beyond (a block of D levels
        // is wholly on one side; a generic pointer serves both)
        auto tail_at = [&](int n) -> CT * {
            if constexpr (TL > 0)
                if (n < TL)
                    return s_tail + (int64_t)n * 128;
            return tail + (int64_t)(n - TL) * 128;
        };
        if constexpr (TAIL) {
            for (int kb = NK; kb < nk; kb += D) {
#pragma clang loop unroll(full)
                for (int s = 0; s < D; ++s) {
                    const fwd_out<CT> r = fstep(kb + s, s, (s + 1) % D, false, true);
                    CT *t = tail_at(kb + s - NK);
                    t[lane] = r.c;
                    t[64 + lane] = r.d;
                }
            }
        }

        // backward sweep (vertical_advection_dycore.cpp:110-121), top level first
        CT data = CT(0);
        if constexpr (TAIL) {
            // the block holding the top level first (none when the column fits the register part)
            const int last = nk > NK ? NK + ((nk - NK - 1) / D) * D : NK - D;
            for (int kb = last; kb >= NK; kb -= D) {
                CT rc[D], rd[D], rp[D];
#pragma clang loop unroll(full)
                for (int s = 0; s < D; ++s) {
                    const CT *t = tail_at(kb + s - NK);
                    rc[s] = t[lane];
                    rd[s] = t[64 + lane];
                    rp[s] = CT(bload<ST, CPOL_NT>(r_up, v_col, lvl(kb + s) * sk_b));
                }
#pragma clang loop unroll(full)
                for (int s = D - 1; s >= 0; --s) {
                    const int k = kb + s;
                    const CT dt = rd[s] - rc[s] * data;
                    const CT nd = k == top ? rd[s] : dt;
                    data = nd;
                    bstore<ST, CPOL_OUT>(r_uts, ST(dtr * (nd - rp[s])), v_col, lvl(k) * sk_b);
                }
            }
        }
        // the parked u_pos read four levels ahead of its use: with one wavefront per SIMD nothing else would hide
        // the latency of each LDS read in the backward chain
        constexpr int UA = NK < 4 ? NK : 4;
        ST l_up[UA];
#pragma clang loop unroll(full)
        for (int s = 0; s < UA; ++s)
            l_up[(NK - 1 - s) % UA] = s_up[NK - 1 - s][lane];
#pragma clang loop unroll(full)
        for (int k = NK - 1; k >= 0; --k) {
            const CT v_up = CT(l_up[k % UA]);
            if (k - UA >= 0)
                l_up[k % UA] = s_up[k - UA][lane];
            const CT dt = d[k] - c[k] * data;
            const CT nd = k == top ? d[k] : dt;
            data = nd;
            bstore<ST, CPOL_OUT>(r_uts, ST(dtr * (nd - v_up)), v_col, lvl(k) * sk_b);
            __builtin_amdgcn_sched_barrier(6);
        }
    }

} // namespace gt355
