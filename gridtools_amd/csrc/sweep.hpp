// The hdiff register sweep of one wavefront (horizontal_diffusion.cpp:35-107): used by the single-domain hdiff
// kernel (hdiff.hip) and by the batching form's stencil phase (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "functors.hpp"

namespace gt355 {
    // Register sweep: no LDS, no barriers.
    //
    // One wavefront owns 64 consecutive columns [i0, i0+64) of one k-level and sweeps JB output rows in j.
    // Rows of in live in registers (a three-row window), the i-neighbours of in and lap come from DPP
    // wavefront shifts, and the two lanes at the wave edges take their out-of-wave neighbour from a "side"
    // column the wave streams with scalar loads (columns i0-1 / i0+64, plus i0-2 / i0+65 for the side
    // laplacian): the DPP bound control keeps the destination's old value for a source lane outside
    // the wave, and that old value is exactly the side column.  Loads run P rows ahead through register
    // rings; all loads are unconditional (clamped to the valid box) so the load counter stays exact.
    // The wavefront's strip is columns [i0, i0+64) x rows [j0, j0+JB) of level k (i0, j0, k wave-uniform).  LAP also
    // stores the laplacian of every output point (the lap ij-cache value there) into `lap`, laid out like `out`.
    // PO / PI: the output and input pointer types (ST * and const ST * from kernel arguments; global-address-space
    // pointers when they come from a table in memory, so that accesses are global_* rather than flat_*).
    template <class ST, class CT, int JB, int P, bool LAP, class PO = ST *, class PI = const ST *>
    __device__ __forceinline__ void hdiff_sweep_wave(PO __restrict__ out, PI __restrict__ in, PI __restrict__ coeff,
        PO __restrict__ lap, int ni, int nj, int64_t osj, int64_t osk, int64_t isj, int64_t isk, int64_t csj,
        int64_t csk, int i0, int j0, int k) {
        constexpr int NR = JB + 4; // in rows j0-2 .. j0+JB+1
        const int lane = threadIdx.x & 63;
        const bool left = lane < 32;
        // columns, clamped to the readable box [-2, ni+2): own column per lane, side columns per wave
        const int ic = min(i0 + lane, ni + 1);
        const int ieL = i0 - 1, ifL = i0 - 2;
        const int ieR = min(i0 + 64, ni + 1), ifR = min(i0 + 65, ni + 1);
        const PI inp = in + (int64_t)k * isk;
        const PI cop = coeff + (int64_t)k * csk;
        auto row = [&](int t) { return (int64_t)min(j0 - 2 + t, nj + 1); }; // absolute row of ring step t

        ST rc[P], rk[P];                 // per-lane rings: own column, coeff
        ST reL[P], rfL[P], reR[P], rfR[P]; // wave-uniform rings: side columns (scalar loads)
        auto load = [&](int t) {
            const int64_t r = row(t);
            // in rows are re-read by the neighbouring strips (rim rows) and side columns: default cache policy
            rc[t % P] = inp[r * isj + ic];
            reL[t % P] = inp[r * isj + ieL];
            rfL[t % P] = inp[r * isj + ifL];
            reR[t % P] = inp[r * isj + ieR];
            rfR[t % P] = inp[r * isj + ifR];
            // coeff of the row that is output at ring step t, R-2 = j0-4+t (clamped into the interior; only
            // steps 4..NR-1 use it)
            const int64_t ro = min(max(j0 - 4 + t, 0), nj - 1);
            rk[t % P] = __builtin_nontemporal_load(cop + ro * csj + min(ic, ni - 1)); // read once
        };
#pragma clang loop unroll(full)
        for (int t = 0; t < P; ++t)
            load(t);

        CT c0 = 0, c1 = 0, c2 = 0;       // in rows R-2, R-1, R of the own column
        CT eL0 = 0, eL1 = 0, eL2 = 0;    // the same rows of column i0-1 (uniform)
        CT eR0 = 0, eR1 = 0, eR2 = 0;    // the same rows of column i0+64 (uniform)
        CT fL1 = 0, fR1 = 0;             // columns i0-2 and i0+65 of row R-1 (uniform)
        CT m1 = 0, p1 = 0;               // in(i-1), in(i+1) of row R-1
        CT m0 = 0, p0 = 0;               // in(i-1), in(i+1) of row R-2
        CT lap1 = 0, lap0 = 0;           // lap of rows R-1 and R-2 (own column)
        CT sl1 = 0, sl0 = 0;             // side laplacian of rows R-1 and R-2 (i0-1 on lane 0, i0+64 on lane 63)
        CT fly_m = 0;                    // fly of row R-3
        // output rows through buffer descriptors based at the strip's first point (wave-uniform): the row offset is
        // a scalar, the lane's column a vector offset, and the stores take the output cache policy
        const int64_t o0 = (int64_t)k * osk + (int64_t)j0 * osj + i0;
        const rsrc_t r_o = make_rsrc(wave_uniform((const void *)(out + o0)));
        const rsrc_t r_l = make_rsrc(wave_uniform((const void *)(LAP ? lap + o0 : out + o0)));
        const uint32_t ob = (uint32_t)(lane * sizeof(ST));
        const bool col_ok = i0 + lane < ni;
#pragma clang loop unroll(full)
        for (int t = 0; t < NR; ++t) {
            // shift the windows and take row R = j0 - 2 + t from the rings
            c0 = c1; c1 = c2; c2 = CT(rc[t % P]);
            eL0 = eL1; eL1 = eL2; eL2 = CT(reL[t % P]);
            eR0 = eR1; eR1 = eR2; eR2 = CT(reR[t % P]);
            const CT kout = CT(rk[t % P]); // coeff of row R-2
            const CT fL_r = CT(rfL[t % P]), fR_r = CT(rfR[t % P]);
            if (t + P < NR)
                load(t + P);
            // in(i-1), in(i+1) of row R: lane 0 keeps column i0-1, lane 63 keeps column i0+64
            const CT m2 = dpp_keep<DPP_WF_SR1>(eL2, c2), p2 = dpp_keep<DPP_WF_SL1>(eR2, c2);
            lap0 = lap1;
            sl0 = sl1;
            if (t >= 2) {
                // lap of row R-1 (lap_function, horizontal_diffusion.cpp:41-46): needs rows R-2..R
                lap1 = lap_f<CT>(c1, p1, c2, m1, c0);
                // side laplacian at i0-1 (left lanes) or i0+64 (right lanes); the own column is the i+1
                // neighbour on the left side and the i-1 neighbour on the right side
                const CT e0 = left ? eL0 : eR0, e1 = left ? eL1 : eR1, e2 = left ? eL2 : eR2;
                const CT ip = left ? c1 : fR1, im = left ? fL1 : c1;
                sl1 = lap_f<CT>(e1, ip, e2, im, e0);
            }
            if (t >= 4) {
                // output row R-2: flx, flx[i-1], fly, fly[j-1], out (horizontal_diffusion.cpp:56-60, 70-74, 86-90)
                const CT lm = dpp_keep<DPP_WF_SR1>(sl0, lap0), lp = dpp_keep<DPP_WF_SL1>(sl0, lap0);
                const CT flx_c = flux_f<CT>(lp, lap0, p0, c0);
                const CT flx_m = flux_f<CT>(lap0, lm, c0, m0);
                const CT fly_c = flux_f<CT>(lap1, lap0, c1, c0);
                const CT res = c0 - kout * (((flx_c - flx_m) + fly_c) - fly_m);
                fly_m = fly_c;
                const int r_out = j0 + t - 4;
                if (col_ok && r_out < nj) {
                    const uint32_t rb = (uint32_t)((t - 4) * osj * (int64_t)sizeof(ST));
                    bstore<ST, CPOL_OUT>(r_o, static_cast<ST>(res), ob, rb);
                    if constexpr (LAP)
                        bstore<ST, CPOL_OUT>(r_l, static_cast<ST>(lap0), ob, rb);
                }
            } else if (t == 3) {
                // fly of row j0-1, the fly[j-1] of the first output row
                fly_m = flux_f<CT>(lap1, lap0, c1, c0);
            }
            // rows move down one: (R-1) -> R-2 for the next step's output, R -> R-1 for its laplacian
            m0 = m1;
            p0 = p1;
            m1 = m2;
            p1 = p2;
            fL1 = fL_r;
            fR1 = fR_r;
        }
    }

    // Column-pair register sweep: the register sweep above with two adjacent columns per lane, so one wavefront
    // covers 128 columns [i0, i0+128) and every row load is one 2-element vector per lane (8 bytes per lane for
    // FP32 storage: 512-byte wavefront requests, the size at which the FP64 sweep streams best; measured at
    // 256x256x80 FP32 storage: 14.2 us against 16.5 us for one column per lane).  Lane l owns columns
    // a = i0+2l and b = a+1: the i-neighbours inside a lane are the other column of the pair, across lanes they
    // come from DPP shifts (a's i-1 is the b column of lane l-1, b's i+1 the a column of lane l+1), and lanes
    // 0 / 63 keep the side columns i0-1 / i0+128 (and i0-2 / i0+129 for the side laplacian), streamed with
    // scalar loads.  flx at b-1 is flx at a, so each lane evaluates three x-fluxes, not four.  Lanes past the
    // domain load the last pair (clamped) and store nothing; since ni is even, the lane after the last valid
    // one loads exactly columns ni and ni+1, the right neighbours the last valid lane needs.  The launcher uses
    // it for even ni with 2-element-aligned rows (origin and j/k strides); other shapes take the one-column
    // sweep.
    // The wavefront's strip is columns [i0, i0+128) x rows [j0, j0+JB) of level k (wave-uniform); LAP also stores the
    // laplacian of every output point into `lap`, laid out like `out`.  PO / PI as for hdiff_sweep_wave.
    template <class ST, class CT, int JB, int P, bool LAP, class PO = ST *, class PI = const ST *>
    __device__ __forceinline__ void hdiff_pair_wave(PO __restrict__ out, PI __restrict__ in, PI __restrict__ coeff,
        PO __restrict__ lap, int ni, int nj, int64_t osj, int64_t osk, int64_t isj, int64_t isk, int64_t csj,
        int64_t csk, int i0, int j0, int k) {
        typedef ST vec2 __attribute__((ext_vector_type(2)));
        // vector pointers in the address space of the field pointers
        using VCP = decltype((const vec2 *)(PI)nullptr);
        constexpr int NR = JB + 4; // in rows j0-2 .. j0+JB+1
        const int lane = threadIdx.x & 63;
        const bool left = lane < 32;
        const int ca = i0 + 2 * lane;
        // pair starts, clamped: in is readable on [-2, ni+2), coeff on [0, ni); ni is even
        const int pin_ = min(ca, ni), pco = min(ca, ni - 2);
        const int ieL = i0 - 1, ifL = i0 - 2;
        const int ieR = min(i0 + 128, ni + 1), ifR = min(i0 + 129, ni + 1);
        const PI inp = in + (int64_t)k * isk;
        const PI cop = coeff + (int64_t)k * csk;
        auto row = [&](int t) { return (int64_t)min(j0 - 2 + t, nj + 1); };

        vec2 rc[P], rk[P];
        ST reL[P], rfL[P], reR[P], rfR[P];
        auto load = [&](int t) {
            const int64_t r = row(t);
            rc[t % P] = *(VCP)(inp + r * isj + pin_);
            reL[t % P] = inp[r * isj + ieL];
            rfL[t % P] = inp[r * isj + ifL];
            reR[t % P] = inp[r * isj + ieR];
            rfR[t % P] = inp[r * isj + ifR];
            const int64_t ro = min(max(j0 - 4 + t, 0), nj - 1);
            rk[t % P] = __builtin_nontemporal_load((VCP)(cop + ro * csj + pco)); // read once
        };
#pragma clang loop unroll(full)
        for (int t = 0; t < P; ++t)
            load(t);

        CT a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0; // in rows R-2, R-1, R of columns a and b
        CT eL0 = 0, eL1 = 0, eL2 = 0, eR0 = 0, eR1 = 0, eR2 = 0;
        CT fL1 = 0, fR1 = 0;
        CT ma1 = 0, ma0 = 0; // in(a-1) of rows R-1, R-2
        CT pb1 = 0, pb0 = 0; // in(b+1) of rows R-1, R-2
        CT la1 = 0, la0 = 0, lb1 = 0, lb0 = 0; // lap of rows R-1, R-2 at a and b
        CT sl1 = 0, sl0 = 0;                  // side laplacian (i0-1 on lane 0, i0+128 on lane 63)
        CT flya_m = 0, flyb_m = 0;            // fly of row R-3
        const int64_t o0 = (int64_t)k * osk + (int64_t)j0 * osj + i0;
        const rsrc_t r_o = make_rsrc(wave_uniform((const void *)(out + o0)));
        const rsrc_t r_l = make_rsrc(wave_uniform((const void *)(LAP ? lap + o0 : out + o0)));
        const uint32_t ob = (uint32_t)(2 * lane * sizeof(ST));
        const bool col_ok = ca < ni; // ni even: b = a+1 is inside too
#pragma clang loop unroll(full)
        for (int t = 0; t < NR; ++t) {
            const vec2 v = rc[t % P];
            a0 = a1; a1 = a2; a2 = CT(v.x);
            b0 = b1; b1 = b2; b2 = CT(v.y);
            eL0 = eL1; eL1 = eL2; eL2 = CT(reL[t % P]);
            eR0 = eR1; eR1 = eR2; eR2 = CT(reR[t % P]);
            const vec2 kv = rk[t % P];
            const CT ka = CT(kv.x), kb = CT(kv.y);
            const CT fL_r = CT(rfL[t % P]), fR_r = CT(rfR[t % P]);
            if (t + P < NR)
                load(t + P);
            const CT ma2 = dpp_keep<DPP_WF_SR1>(eL2, b2), pb2 = dpp_keep<DPP_WF_SL1>(eR2, a2);
            la0 = la1;
            lb0 = lb1;
            sl0 = sl1;
            if (t >= 2) {
                // lap of row R-1 at a and b (lap_function, horizontal_diffusion.cpp:41-46)
                la1 = lap_f<CT>(a1, b1, a2, ma1, a0);
                lb1 = lap_f<CT>(b1, pb1, b2, a1, b0);
                const CT e0 = left ? eL0 : eR0, e1 = left ? eL1 : eR1, e2 = left ? eL2 : eR2;
                const CT ip = left ? a1 : fR1, im = left ? fL1 : b1;
                sl1 = lap_f<CT>(e1, ip, e2, im, e0);
            }
            if (t >= 4) {
                // output row R-2 (horizontal_diffusion.cpp:56-60, 70-74, 86-90)
                const CT lam = dpp_keep<DPP_WF_SR1>(sl0, lb0), lbp = dpp_keep<DPP_WF_SL1>(sl0, la0);
                const CT flx_a = flux_f<CT>(lb0, la0, b0, a0);
                const CT flx_am = flux_f<CT>(la0, lam, a0, ma0);
                const CT flx_b = flux_f<CT>(lbp, lb0, pb0, b0);
                const CT fly_a = flux_f<CT>(la1, la0, a1, a0);
                const CT fly_b = flux_f<CT>(lb1, lb0, b1, b0);
                const CT ra = a0 - ka * (((flx_a - flx_am) + fly_a) - flya_m);
                const CT rb = b0 - kb * (((flx_b - flx_a) + fly_b) - flyb_m);
                flya_m = fly_a;
                flyb_m = fly_b;
                const int r_out = j0 + t - 4;
                if (col_ok && r_out < nj) {
                    const uint32_t rbo = (uint32_t)((t - 4) * osj * (int64_t)sizeof(ST));
                    bstore_pair<ST, CPOL_OUT>(r_o, static_cast<ST>(ra), static_cast<ST>(rb), ob, rbo);
                    if constexpr (LAP)
                        bstore_pair<ST, CPOL_OUT>(r_l, static_cast<ST>(la0), static_cast<ST>(lb0), ob, rbo);
                }
            } else if (t == 3) {
                // fly of row j0-1, the fly[j-1] of the first output row
                flya_m = flux_f<CT>(la1, la0, a1, a0);
                flyb_m = flux_f<CT>(lb1, lb0, b1, b0);
            }
            ma0 = ma1;
            pb0 = pb1;
            ma1 = ma2;
            pb1 = pb2;
            fL1 = fL_r;
            fR1 = fR_r;
        }
    }

} // namespace gt355
