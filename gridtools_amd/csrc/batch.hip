// Batching form: many independent domains of uneven ij x k sizes in one persistent launch.
//
// The reference batches only through expandable_run, which replicates a spec over a std::vector of fields that
// all share one grid (stencil/frontend/expandable_run.hpp:80-183); domains of different sizes need separate runs.
// Here the host lists every (domain, 64x4 column tile) as a work item, longest columns first, and a fixed grid of
// workgroups strides through the list (item = blockIdx.x + n * gridDim.x; no counters, so the launch can be
// captured in a graph and replayed).  Each item computes, for its 256 columns and every level of its domain:
//   lap     = laplacian(in)                        (laplacian.cpp:20-29; the lap ij-cache of hdiff, stored)
//   u_stage = hdiff(in, coeff)                     (horizontal_diffusion.cpp:35-107)
//   vadv(utens_stage, u_stage, wcon, u_pos, utens) (vertical_advection_dycore.cpp:32-150), u_stage taken from the
//                                                 diffusion in the compute type, not re-read
// The two-phase form runs the same chain in two persistent launches per timestep, each over every domain: the
// stencil phase (laplacian and hdiff by the one-column register sweep of sweep.hpp, one wavefront per 64-column x
// JB-row x level item; lap and u_stage stored) and the vertical phase (the strip sweep of strip.hpp with u_stage read
// back instead of evaluated).  It evaluates each laplacian once instead of once per row that needs it, at the cost
// of the u_stage round trip through memory; in the precision form u_stage reaches the vertical phase rounded to
// FP32 (as in the two-launch composition), where the single-launch forms pass it on in FP64.
// The k depth is a runtime value (up to 144 levels), so ccol/dcol go to the workgroup's own slot of a scratch
// buffer (k-major, coalesced; written and read back by the same workgroup, so they stay in L2 / Infinity Cache).
// Levels are processed in blocks of D with static ring slots inside a block; loads are unconditional (clamped to
// the domain's top level) and stores past the top land on the top level before its own value does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "functors.hpp"
#include "strip.hpp"
#include "sweep.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {
        constexpr int BTI = 64, BTJ = 4, BMAXK = 144;
        // forward steps run up to D-1 levels past a column's top; scratch rows = deepest nk + BPAST
        constexpr int BPAST = 8;

        struct bdom {
            const void *in, *coeff, *wcon, *u_pos, *utens;
            void *lap, *u_stage, *utens_stage;
            int64_t sj, sk;   // shared layout of every field but wcon (origins at grid point (0,0,0))
            int64_t wsj, wsk; // wcon layout
            int32_t ni, nj, nk, tiles_i;
        };

        struct bitem {
            int32_t dom, ti, tj, pad;
        };

        // field pointers read from the domain table are generic; marked global they load and store with global_*
        // instructions (counted by vmcnt alone) instead of flat_* (which also hold lgkmcnt, as LDS does)
        template <class T>
        using gptr = __attribute__((address_space(1))) T *;

        template <class ST, class CT, int D>
        __global__ void __launch_bounds__(256) batch_kernel(const bdom *__restrict__ doms,
            const bitem *__restrict__ items, int nitems, CT dtr, CT *__restrict__ scratch, int rows, int nxcd) {
            static_assert(D % 2 == 0, "the LDS buffer parity of a block step is static only for even D");
            constexpr int IW = BTI + 4, IH = BTJ + 4, LW = BTI + 2;
            constexpr int NIN = IW * IH, IN_PER = (NIN + 255) / 256, NLAP = 512;
            __shared__ ST s_in[2][IN_PER * 256];
            __shared__ CT s_lap[NLAP];
            const int tid = threadIdx.x, x = tid & 63, y = tid >> 6;
            static_assert(D <= BPAST, "scratch rows cover the steps past the top");
            CT *sc = scratch + (int64_t)blockIdx.x * (2 * rows * 256); // this workgroup's ccol/dcol slot
            CT *sd = sc + (int64_t)rows * 256;

            // consecutive items (neighbouring tiles, sharing rim rows) go to workgroups of one XCD
            const int wid = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            for (int it = wid; it < nitems; it += gridDim.x) {
                const bitem item = items[it];
                const bdom dm = doms[item.dom];
                const int ni = dm.ni, nj = dm.nj, nk = dm.nk, top = nk - 1;
                const int i0 = item.ti * BTI, j0 = item.tj * BTJ;
                // lanes outside the domain evaluate the nearest interior column (identical, harmless stores)
                const int xe = min(x, ni - 1 - i0), ye = min(y, nj - 1 - j0);
                const int i = i0 + xe, j = j0 + ye;
                const gptr<const ST> in_f = (gptr<const ST>)dm.in, co_f = (gptr<const ST>)dm.coeff,
                                     wc_f = (gptr<const ST>)dm.wcon, up_f = (gptr<const ST>)dm.u_pos,
                                     ut_f = (gptr<const ST>)dm.utens;
                const gptr<ST> lap_o = (gptr<ST>)dm.lap, ust_f = (gptr<ST>)dm.u_stage,
                               uts_f = (gptr<ST>)dm.utens_stage;
                int in_off[IN_PER];
#pragma clang loop unroll(full)
                for (int r = 0; r < IN_PER; ++r) {
                    const int e = min(tid + r * 256, NIN - 1);
                    const int gi = min(i0 - 2 + e % IW, ni + 1), gj = min(j0 - 2 + e / IW, nj + 1);
                    in_off[r] = gi + gj * (int)dm.sj;
                }
                const int64_t col = (int64_t)j * dm.sj + i, wcol = (int64_t)j * dm.wsj + i;
                auto lvl = [&](int L) { return (int64_t)min(L, top); };
                // opaque per-step strides keep each load at its own step (not hoisted as loop invariant)
                int64_t sk = dm.sk, wsk = dm.wsk;
                auto pin = [&]() {
                    asm volatile("" : "+s"(sk));
                    asm volatile("" : "+s"(wsk));
                };

                ST h_in[IN_PER][D], h_co[D], b_uts[D], b_up[D], b_ut[D], a_wi[D], a_wip[D];
                auto load_h = [&](int L, int slot) {
                    const int64_t l = lvl(L);
#pragma clang loop unroll(full)
                    for (int r = 0; r < IN_PER; ++r)
                        h_in[r][slot] = in_f[l * sk + in_off[r]];
                    h_co[slot] = co_f[l * sk + col];
                };
                auto load_b = [&](int L, int slot) {
                    const int64_t l = lvl(L);
                    b_uts[slot] = uts_f[l * sk + col];
                    b_up[slot] = up_f[l * sk + col];
                    b_ut[slot] = ut_f[l * sk + col];
                };
                auto load_a = [&](int L, int slot) {
                    const int64_t l = lvl(L);
                    a_wi[slot] = wc_f[l * wsk + wcol];
                    a_wip[slot] = wc_f[l * wsk + wcol + 1];
                };
                // laplacian and hdiff of level L from ring slot `slot` through LDS buffer `buf`; the lap
                // ij-cache value of the column is the laplacian output (levels past the top recompute the top)
                auto hdiff_level = [&](int L, int slot, int buf) -> CT {
#pragma clang loop unroll(full)
                    for (int r = 0; r < IN_PER; ++r)
                        s_in[buf][tid + r * 256] = h_in[r][slot];
                    const CT coeff = CT(h_co[slot]);
                    __syncthreads();
#pragma clang loop unroll(full)
                    for (int q = 0; q < NLAP / 256; ++q) {
                        const int idx = tid + q * 256;
                        const int lx = idx % LW + 1, ly = idx / LW + 1;
                        const ST *p = &s_in[buf][ly * IW + lx];
                        s_lap[idx] = lap_f<CT>(CT(p[0]), CT(p[1]), CT(p[IW]), CT(p[-1]), CT(p[-IW]));
                    }
                    __syncthreads();
                    const ST *ip = &s_in[buf][(ye + 2) * IW + xe + 2];
                    const CT *lp = &s_lap[(ye + 1) * LW + xe + 1];
                    const CT c = CT(ip[0]), lc = lp[0];
                    const CT flx_c = flux_f<CT>(lp[1], lc, CT(ip[1]), c);
                    const CT flx_m = flux_f<CT>(lc, lp[-1], c, CT(ip[-1]));
                    const CT fly_c = flux_f<CT>(lp[LW], lc, CT(ip[IW]), c);
                    const CT fly_m = flux_f<CT>(lc, lp[-LW], c, CT(ip[-IW]));
                    const CT res = c - coeff * (((flx_c - flx_m) + fly_c) - fly_m);
                    const int64_t o = lvl(L) * sk + col;
                    lap_o[o] = ST(lc);
                    ust_f[o] = ST(res);
                    return res;
                };

#pragma clang loop unroll(full)
                for (int s = 0; s < D; ++s) {
                    load_h(s, s);
                    load_b(s, s);
                    load_a(s + 1, (s + 1) % D);
                }
                CT us_c = hdiff_level(0, 0, 0);
                pin();
                load_h(D, 0);
                CT us_m = CT(0), w_c_i = CT(0), w_c_ip = CT(0), cm = CT(0), dmm = CT(0);
                // forward step of level k: its inputs are in ring slot `slot`, those of level k+1 in `nslot`
                auto fwd_step = [&](int k, int slot, int nslot, int buf, bool first) {
                    const bool at_top = k == top;
                    const CT us_next = hdiff_level(k + 1, nslot, buf);
                    const CT v_uts = CT(b_uts[slot]), v_up = CT(b_up[slot]), v_ut = CT(b_ut[slot]);
                    // top level: k+1 wind taken as zero and us_p = us_c, which reproduces last_level exactly
                    const CT us_p = at_top ? us_c : us_next;
                    const CT w_p_i = at_top ? CT(0) : CT(a_wi[nslot]);
                    const CT w_p_ip = at_top ? CT(0) : CT(a_wip[nslot]);
                    pin();
                    load_h(k + 1 + D, nslot);
                    load_b(k + D, slot);
                    load_a(k + 1 + D, nslot);
                    fwd_out<CT> r;
                    if (first)
                        r = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                    else
                        r = forward_body<CT>(
                            dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, v_up, v_ut, v_uts, cm, dmm, at_top);
                    sc[k * 256 + tid] = r.c;
                    sd[k * 256 + tid] = r.d;
                    cm = r.c;
                    dmm = r.d;
                    us_m = us_c;
                    us_c = us_p;
                    w_c_i = w_p_i;
                    w_c_ip = w_p_ip;
                };
                fwd_step(0, 0, 1 % D, 1, true);
                // levels 1.. in blocks of D: level kb+s (kb = 1 + m*D) sits in slot (1+s)%D and evaluates the
                // diffusion of level kb+s+1 through LDS buffer s&1; steps past the top compute values nobody reads
                for (int kb = 1; kb < nk; kb += D) {
#pragma clang loop unroll(full)
                    for (int s = 0; s < D; ++s)
                        fwd_step(kb + s, (1 + s) % D, (2 + s) % D, s & 1, false);
                }
                // backward sweep (vertical_advection_dycore.cpp:110-121) in blocks of D, top block first: a
                // step above the top stores into the top level before the top level's own step stores there
                const int nblk = (nk + D - 1) / D;
                CT data = CT(0);
                for (int bb = nblk - 1; bb >= 0; --bb) {
                    CT r_c[D], r_d[D], r_p[D];
                    pin();
#pragma clang loop unroll(full)
                    for (int s = 0; s < D; ++s) {
                        const int k = bb * D + s;
                        r_c[s] = sc[k * 256 + tid];
                        r_d[s] = sd[k * 256 + tid];
                        r_p[s] = CT(up_f[lvl(k) * sk + col]);
                    }
#pragma clang loop unroll(full)
                    for (int s = D - 1; s >= 0; --s) {
                        const int k = bb * D + s;
                        const CT dt = r_d[s] - r_c[s] * data;
                        const CT nd = k == top ? r_d[s] : dt;
                        data = nd;
                        uts_f[lvl(k) * sk + col] = ST(dtr * (nd - r_p[s]));
                    }
                }
                __syncthreads(); // the next item reuses the LDS tiles
            }
        }


        // Work item of the strip form: 64 columns (from i0) of row j of one domain; with `flat` set (the two-phase
        // form's vertical phase), 64 entries of the domain's column list from entry i0 (j unused).
        struct sitem {
            int32_t dom, i0, j, flat;
        };
        // load-ahead depths of the strip sweep: 4 levels in the single-launch strip form and with 32 register
        // levels (deeper rings spill there), 8 in the two-phase form's vertical phase, whose kernel then takes 476 of
        // 512 registers without spills (its one wavefront per SIMD is latency-bound at 4: five loads per level,
        // 5 MB in flight over the GPU)
        constexpr int SD = 4, SDV = 8;
        // tail levels a wavefront's scratch holds, for a register part of NK levels
        constexpr int stail_rows(int nk_reg) { return BMAXK - nk_reg + SDV; }

        // The batching form on the strip sweep (strip.hpp): one wavefront per workgroup and per item, no LDS
        // tile, no barrier, ccol/dcol in registers for the first NK levels and in the wavefront's tail scratch
        // beyond.  One sweep variant serves every depth (a column shorter than NK levels runs the register part
        // with its levels past the top clamped; separate shorter variants in the same kernel would share one
        // register allocation and spill).  Items are taken with a static stride (no counters), longest columns
        // first.
        //   NK = 80, one wavefront per SIMD: the whole column of the baseline depth in registers;
        //   NK = 32, two wavefronts per SIMD (ccol/dcol of 32 levels fit half the register file): every level
        //   past 32 goes through the tail scratch, but no column runs clamped steps past its top (the batch's
        //   depths start at 40) and two wavefronts hide each other's latency.
        // The vertical phase with FP32 storage also has a 96-level register part, taken by columns of at least
        // deep_nk levels: it halves the tail-scratch levels of the deepest columns (ccol/dcol of 96 levels still fit
        // the register file of one wavefront per SIMD; u_pos of 96 levels in LDS is 24 KiB per wavefront).
        template <class ST, class CT, int NK, int WPE, bool HD>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
        batch_strip_kernel(const bdom *__restrict__ doms, const sitem *__restrict__ items, int nitems, CT dtr,
            CT *__restrict__ scratch, int nxcd, int deep_nk) {
            constexpr bool DEEP = !HD && NK == 80 && sizeof(ST) == 4;
            constexpr int KD = !HD && NK == 80 ? SDV : SD;
            __shared__ ST s_up[DEEP ? 96 : NK][64];
            // with FP32 storage the 96-level part keeps the first LTAIL tail levels of ccol/dcol in LDS: 24 KiB of
            // u_pos + 16 KiB of tail per wavefront, the 160 KiB of a CU for its four wavefronts
            constexpr int LTAIL = DEEP && sizeof(CT) == 8 ? 16 : 0;
            __shared__ CT s_tail[LTAIL > 0 ? LTAIL * 128 : 1];
            CT *tail = scratch + (int64_t)blockIdx.x * (stail_rows(NK) * 128);
            const int wid = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            for (int it = wid; it < nitems; it += gridDim.x) {
                const sitem item = items[it];
                const bdom dm = doms[item.dom];
                const strip_fields<ST> f{(const ST *)dm.in, (const ST *)dm.coeff, (const ST *)dm.wcon,
                    (const ST *)dm.u_pos, (const ST *)dm.utens, (ST *)dm.u_stage, (ST *)dm.utens_stage,
                    (ST *)dm.lap, dm.sj, dm.sk, dm.sj, dm.sk, dm.sj, dm.sk, dm.sj, dm.sk, dm.wsj, dm.wsk};
                // the vertical phase takes items of the column list (flagged in the item) or of one row's strip
                const bool flat = item.flat != 0;
                if constexpr (!HD && NK == 80) {
                    // the vertical phase: columns of at most 48 or 64 levels run a register part of that depth
                    // (a uniform branch per item), not 80 steps with the levels past their top clamped
                    if (dm.nk <= 48)
                        hv_strip<ST, CT, 48, KD, true, true, false, HD, !HD>(
                            f, dtr, dm.ni, dm.nj, dm.nk, item.i0, item.j, s_up, tail, flat);
                    else if (dm.nk <= 64)
                        hv_strip<ST, CT, 64, KD, true, true, false, HD, !HD>(
                            f, dtr, dm.ni, dm.nj, dm.nk, item.i0, item.j, s_up, tail, flat);
                    else if (DEEP && dm.nk >= deep_nk)
                        hv_strip<ST, CT, DEEP ? 96 : NK, KD, true, true, true, HD, !HD, LTAIL>(
                            f, dtr, dm.ni, dm.nj, dm.nk, item.i0, item.j, s_up, tail, flat, s_tail);
                    else
                        hv_strip<ST, CT, NK, KD, true, true, true, HD, !HD>(
                            f, dtr, dm.ni, dm.nj, dm.nk, item.i0, item.j, s_up, tail, flat);
                } else {
                    hv_strip<ST, CT, NK, KD, true, true, true, HD, !HD>(
                        f, dtr, dm.ni, dm.nj, dm.nk, item.i0, item.j, s_up, tail, flat);
                }
            }
        }

        // Work item of the two-phase form's stencil phase: columns [i0, i0+64) x rows [j0, j0+JB) of level k.
        struct hitem {
            int32_t dom, i0, j0, k;
        };

        // The stencil phase: laplacian (stored as lap where the domain has one) and hdiff (stored as u_stage) of every
        // domain in the phase, one register
        // sweep per item, items taken by wavefronts with a static stride; the item order (domain, level, row block,
        // strip) keeps the wavefronts in flight on neighbouring strips, whose rim rows they share through L2.
        template <class ST, class CT, int JB, int P>
        __global__ void __launch_bounds__(256) batch_hdiff_kernel(const bdom *__restrict__ doms,
            const hitem *__restrict__ items, int nitems, int nxcd) {
            const int nw = gridDim.x * 4;
            const int w0 = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x, nxcd) * 4 + (threadIdx.x >> 6));
            for (int it = w0; it < nitems; it += nw) {
                const hitem item = items[it];
                const bdom dm = doms[item.dom];
                // global-address-space views of the table's pointers (global_* rather than flat_* accesses)
                const gptr<ST> ust = (gptr<ST>)dm.u_stage, lap = (gptr<ST>)dm.lap;
                const gptr<const ST> in = (gptr<const ST>)dm.in, co = (gptr<const ST>)dm.coeff;
                // the laplacian is stored where the domain has a lap field (a uniform branch per item)
                if (dm.lap)
                    hdiff_sweep_wave<ST, CT, JB, P, true, gptr<ST>, gptr<const ST>>(ust, in, co, lap, dm.ni, dm.nj,
                        dm.sj, dm.sk, dm.sj, dm.sk, dm.sj, dm.sk, item.i0, item.j0, item.k);
                else
                    hdiff_sweep_wave<ST, CT, JB, P, false, gptr<ST>, gptr<const ST>>(ust, in, co, ust, dm.ni, dm.nj,
                        dm.sj, dm.sk, dm.sj, dm.sk, dm.sj, dm.sk, item.i0, item.j0, item.k);
            }
        }
    } // namespace
} // namespace gt355

using namespace gt355;

struct gt355_batch {
    bdom *doms;    // device table
    void *items;   // device table: bitem (tile form) or sitem (strip and two-phase forms)
    void *hitems[3]; // device tables: hitem (the two-phase form's stencil phase) for 8, 16 and 32 rows per item
    void *scratch; // ccol/dcol of every workgroup the grid may have, in the compute type
    int impl;      // 1: 64x4 tiles, ccol/dcol in scratch; 2: 64-column strips, ccol/dcol in registers;
                   // 3: two phases (stencil sweep, then the vertical advection over column-list items)
    int nk_reg;    // strip form: levels of ccol/dcol held in registers (80: one wavefront per SIMD; 32: two)
    int deep_nk;   // vertical phase, FP32 storage: columns of at least this many levels keep 96 in registers
    int ndoms, nitems, dtype, precision, grid, gridmax, cus, rows;
    int w;           // the persistent grid's setting (batch_grid)
    int nhitems[3];  // the stencil phase: items per table
    int hjb, hw;     // the stencil phase: rows per item, workgroups per CU
    uint64_t sig;    // signature of the domain list (key of the tuner's table)
};

namespace gt355 {
    namespace {
        // workgroups per CU the persistent grid may use, by tuner setting w = 1..BWG_MAX: the tile form fits 3-4
        // workgroups of 256 lanes per CU (w per CU); the strip form one single-wavefront workgroup per SIMD
        // (2w per CU: 2 fills half the SIMDs, 4 all of them, 6 and 8 queue a second round)
        constexpr int BWG_MAX = 4;
        inline int batch_grid(const gt355_batch *bt, int w) {
            return std::min<int>(bt->nitems, bt->cus * (bt->impl >= 2 ? (bt->nk_reg == 32 ? 4 : 2) * w : w));
        }
        // the stencil phase's tables: index of 8, 16 or 32 rows per item
        inline int hjb_index(int hjb) { return hjb == 8 ? 0 : hjb == 32 ? 2 : 1; }
        constexpr int HJB_OF[3] = {8, 16, 32};
        // the stencil phase's grid: four wavefronts per workgroup, hw workgroups per CU (8: one round at 8
        // wavefronts per SIMD), each wavefront striding through the items
        inline int stencil_grid(const gt355_batch *bt) {
            const int64_t n = bt->nhitems[hjb_index(bt->hjb)];
            return (int)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, (int64_t)bt->cus * bt->hw));
        }
        // FNV-1a over the batch's form, types and domain sizes: the key of its tuned configuration
        inline uint64_t batch_signature(int impl, int dt, int precision, const gt355_batch_domain *d, int n) {
            uint64_t h = 1469598103934665603ULL;
            auto mix = [&](int64_t v) {
                for (int b = 0; b < 8; ++b) {
                    h ^= (uint64_t)((v >> (8 * b)) & 0xff);
                    h *= 1099511628211ULL;
                }
            };
            mix(impl), mix(dt), mix(precision), mix(n);
            for (int k = 0; k < n; ++k)
                mix(d[k].ni), mix(d[k].nj), mix(d[k].nk), mix(d[k].phases);
            return h;
        }
        inline int batch_ct(const gt355_batch *bt) {
            return (bt->dtype == GT355_F64 || bt->precision == GT355_WIDEN) ? 8 : 4;
        }
        inline int batch_st(const gt355_batch *bt) { return bt->dtype == GT355_F64 ? 8 : 4; }
    } // namespace
} // namespace gt355

namespace gt355 {
    // gt355_batch_create with the form given (0: the GT355_TUNE_batch_impl knob); the deferred-run scope builds
    // its batches in the two-phase form whatever the knob says
    int batch_create_form(int precision, const gt355_batch_domain *domains, int ndomains, int wg_per_cu,
        gt355_batch **out, int form);
}

extern "C" int gt355_batch_create(int precision, const gt355_batch_domain *domains, int ndomains, int wg_per_cu,
    gt355_batch **out) {
    return batch_create_form(precision, domains, ndomains, wg_per_cu, out, 0);
}

int gt355::batch_create_form(int precision, const gt355_batch_domain *domains, int ndomains, int wg_per_cu,
    gt355_batch **out, int form) {
    GT355_CHECK(out && domains && ndomains > 0, "gt355_batch_create: bad arguments");
    *out = nullptr;
    // the dtype: of the first field the first domain's phases use
    const int dt = domains[0].phases == 2 ? domains[0].u_stage.dtype : domains[0].in.dtype;
    GT355_CHECK(dt == GT355_F64 || dt == GT355_F32, "gt355_batch_create: dtype %d", dt);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dt == GT355_F32),
        "gt355_batch_create: precision %d with dtype %d", precision, dt);
    GT355_CHECK(wg_per_cu >= 1 && wg_per_cu <= BWG_MAX, "gt355_batch_create: wg_per_cu %d (1..%d)", wg_per_cu,
        BWG_MAX);
    // the strip form (default) or the tile form (GT355_TUNE_batch_impl=1)
    // the form (GT355_TUNE_batch_impl): two-phase by default (measured fastest on MI355X: 64 domains of seed 2024,
    // FP32 storage / FP64 compute, 2.41 ms per timestep against 2.73-2.87 ms for the single-launch strip form)
    const int impl_p = form ? form : tune_param("batch_impl", 3);
    const int impl = impl_p == 1 ? 1 : impl_p == 3 ? 3 : 2;
    // register depth of the strip form (GT355_TUNE_batch_nk: 80 or 32)
    const int nk_reg = tune_param("batch_nk", 80) == 32 ? 32 : 80;
    std::vector<bdom> doms(ndomains);
    std::vector<bitem> items;
    std::vector<sitem> sitems;
    std::vector<hitem> hitems[3];
    for (int d = 0; d < ndomains; ++d) {
        const gt355_batch_domain &b = domains[d];
        const int ni = b.ni, nj = b.nj, nk = b.nk, ph = b.phases;
        GT355_CHECK(ph >= 0 && ph <= 2, "gt355_batch_create: domain %d: phases %d", d, ph);
        GT355_CHECK(ph == 0 || impl == 3, "gt355_batch_create: domain %d: one-phase domains need the two-phase form",
            d);
        const bool sten = ph != 2, vert = ph != 1;
        GT355_CHECK(ni > 0 && nj > 0 && nk >= (vert ? 2 : 1) && (nk <= BMAXK || !vert),
            "gt355_batch_create: domain %d size %dx%dx%d", d, ni, nj, nk);
        const bool has_lap = sten && b.lap.origin != nullptr;
        GT355_CHECK(has_lap || !sten || impl == 3, "gt355_batch_create: domain %d: lap is required by this form", d);
        int rc;
        if ((sten && (rc = check_field(&b.in, "in", dt, -2, ni + 2, -2, nj + 2, 0, nk))) ||
            (sten && (rc = check_field(&b.coeff, "coeff", dt, 0, ni, 0, nj, 0, nk))) ||
            (has_lap && (rc = check_field(&b.lap, "lap", dt, 0, ni, 0, nj, 0, nk))) ||
            (rc = check_field(&b.u_stage, "u_stage", dt, 0, ni, 0, nj, 0, nk)) ||
            (vert && (rc = check_field(&b.utens_stage, "utens_stage", dt, 0, ni, 0, nj, 0, nk))) ||
            (vert && (rc = check_field(&b.wcon, "wcon", dt, 0, ni + 1, 0, nj, 0, nk))) ||
            (vert && (rc = check_field(&b.u_pos, "u_pos", dt, 0, ni, 0, nj, 0, nk))) ||
            (vert && (rc = check_field(&b.utens, "utens", dt, 0, ni, 0, nj, 0, nk))))
            return rc;
        // every used field but wcon shares one layout (the layout of u_stage, which both phases use)
        const gt355_field *same[] = {sten ? &b.in : nullptr, sten ? &b.coeff : nullptr, has_lap ? &b.lap : nullptr,
            vert ? &b.utens_stage : nullptr, vert ? &b.u_pos : nullptr, vert ? &b.utens : nullptr};
        for (const gt355_field *f : same)
            GT355_CHECK(!f || (f->stride[1] == b.u_stage.stride[1] && f->stride[2] == b.u_stage.stride[2]),
                "gt355_batch_create: domain %d: every field but wcon must share one layout", d);
        GT355_CHECK(!sten || b.u_stage.origin != b.in.origin, "gt355_batch_create: domain %d: u_stage aliases in", d);
        bdom &m = doms[d];
        m.in = sten ? b.in.origin : nullptr;
        m.coeff = sten ? b.coeff.origin : nullptr;
        m.wcon = vert ? b.wcon.origin : nullptr;
        m.u_pos = vert ? b.u_pos.origin : nullptr;
        m.utens = vert ? b.utens.origin : nullptr;
        m.lap = has_lap ? b.lap.origin : nullptr;
        m.u_stage = b.u_stage.origin;
        m.utens_stage = vert ? b.utens_stage.origin : nullptr;
        m.sj = b.u_stage.stride[1];
        m.sk = b.u_stage.stride[2];
        m.wsj = vert ? b.wcon.stride[1] : 0;
        m.wsk = vert ? b.wcon.stride[2] : 0;
        m.ni = ni;
        m.nj = nj;
        m.nk = nk;
        m.tiles_i = (ni + BTI - 1) / BTI;
        if (impl == 3 && sten) {
            // the stencil phase's item tables for every row-block height the tuner may choose
            for (int h = 0; h < 3; ++h) {
                const int jb = HJB_OF[h];
                const int strips = (ni + 63) / 64, rowblocks = (nj + jb - 1) / jb;
                for (int k = 0; k < nk; ++k)
                    for (int rb = 0; rb < rowblocks; ++rb)
                        for (int si = 0; si < strips; ++si)
                            hitems[h].push_back(hitem{d, si * 64, rb * jb, k});
            }
        }
        if (impl >= 2) {
            const int64_t lim = int64_t(1) << 32, es = dt == GT355_F64 ? 8 : 4;
            GT355_CHECK(((int64_t)(nk - 1) * m.sk + (int64_t)(nj + 3) * m.sj + ni + 4) * es < lim &&
                            ((int64_t)(nk - 1) * m.wsk + (int64_t)(nj + 1) * m.wsj + ni + 1) * es < lim,
                "gt355_batch_create: domain %d spans 4 GiB or more", d);
            if (!vert) {
                // a stencil-phase-only domain has no vertical items
            } else if (impl == 3 && tune_param("batch_flat", 0) != 0) {
                // the vertical phase on 64 consecutive entries of the domain's column list per item (no partial
                // strips at row ends; disjoint, so each column has one owner).  Measured slower than row strips on
                // 64 domains of seed 2024 (2.47 against 2.14 ms per timestep): an item's rows start at any column
                // and may span two rows, so its requests touch more, unaligned lines
                GT355_CHECK((int64_t)ni * nj < (int64_t(1) << 31), "gt355_batch_create: domain %d has too many columns",
                    d);
                for (int c0 = 0; c0 < ni * nj; c0 += 64)
                    sitems.push_back(sitem{d, c0, 0, 1});
            } else {
                const int strips = (ni + 63) / 64;
                for (int j = 0; j < nj; ++j)
                    for (int si = 0; si < strips; ++si)
                        sitems.push_back(sitem{d, si * 64, j, 0}); // disjoint strips: each column has one owner
            }
        } else {
            const int tj_n = (nj + BTJ - 1) / BTJ;
            for (int tj = 0; tj < tj_n; ++tj)
                for (int ti = 0; ti < m.tiles_i; ++ti)
                    items.push_back(bitem{d, ti, tj, 0});
        }
    }
    // longest columns first, so the strided schedule ends with short items (a greedy balance)
    std::stable_sort(items.begin(), items.end(),
        [&](const bitem &a, const bitem &b) { return doms[a.dom].nk > doms[b.dom].nk; });
    std::stable_sort(sitems.begin(), sitems.end(),
        [&](const sitem &a, const sitem &b) { return doms[a.dom].nk > doms[b.dom].nk; });
    hipDeviceProp_t p;
    GT355_HIP(hipGetDeviceProperties(&p, current_device()));
    gt355_batch *bt = new gt355_batch{};
    bt->impl = impl;
    bt->nk_reg = nk_reg;
    bt->ndoms = ndomains;
    bt->nitems = impl >= 2 ? (int)sitems.size() : (int)items.size();
    bt->dtype = dt;
    bt->precision = precision;
    bt->cus = p.multiProcessorCount;
    for (int h = 0; h < 3; ++h)
        bt->nhitems[h] = (int)hitems[h].size();
    // the launch configuration: the tuner's entry for this list of domains if one was recorded (or loaded), else
    // the measured defaults and the caller's grid setting; environment knobs override both
    bt->sig = batch_signature(impl, dt, precision, domains, ndomains);
    batch_cfg cfg{16, 8, wg_per_cu, 89};
    (void)find_batch_cfg(bt->sig, dt == GT355_F64 ? 8 : 4, batch_ct(bt), impl, &cfg);
    bt->hjb = HJB_OF[hjb_index(tune_param("batch_hjb", cfg.hjb))];
    bt->hw = std::min(8, std::max(1, tune_param("batch_hw", cfg.hw)));
    bt->w = std::min(BWG_MAX, std::max(1, cfg.w));
    bt->deep_nk = tune_param("batch_deep_nk", cfg.deep_nk);
    bt->grid = batch_grid(bt, bt->w);
    bt->gridmax = batch_grid(bt, BWG_MAX); // scratch for any grid the tuner may pick
    const size_t item_bytes = impl >= 2 ? sizeof(sitem) * sitems.size() : sizeof(bitem) * items.size();
    const void *item_src = impl >= 2 ? (const void *)sitems.data() : (const void *)items.data();
    hipError_t e = hipMalloc(&bt->doms, sizeof(bdom) * doms.size());
    if (e == hipSuccess && item_bytes)
        e = hipMalloc(&bt->items, item_bytes);
    if (e == hipSuccess)
        e = hipMemcpy(bt->doms, doms.data(), sizeof(bdom) * doms.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && item_bytes)
        e = hipMemcpy(bt->items, item_src, item_bytes, hipMemcpyHostToDevice);
    for (int h = 0; h < 3; ++h) {
        if (e == hipSuccess && !hitems[h].empty())
            e = hipMalloc(&bt->hitems[h], sizeof(hitem) * hitems[h].size());
        if (e == hipSuccess && !hitems[h].empty())
            e = hipMemcpy(bt->hitems[h], hitems[h].data(), sizeof(hitem) * hitems[h].size(), hipMemcpyHostToDevice);
    }
    // each batch owns its scratch, so batches on different streams never share it
    int maxnk = 0;
    for (const bdom &m : doms)
        maxnk = std::max(maxnk, m.nk);
    bt->rows = maxnk + BPAST;
    const size_t ct = (dt == GT355_F64 || precision == GT355_WIDEN) ? 8 : 4;
    const size_t per_wg = impl >= 2 ? (size_t)stail_rows(nk_reg) * 128 : 2 * (size_t)bt->rows * 256;
    if (e == hipSuccess && bt->gridmax > 0)
        e = hipMalloc(&bt->scratch, ct * per_wg * (size_t)bt->gridmax);
    if (e != hipSuccess) {
        (void)hipFree(bt->doms);
        (void)hipFree(bt->items);
        for (void *h : bt->hitems)
            (void)hipFree(h);
        (void)hipFree(bt->scratch);
        delete bt;
        return hip_fail(e, "gt355_batch_create");
    }
    *out = bt;
    return GT355_OK;
}

extern "C" int gt355_batch_run(gt355_batch *bt, double dtr_stage, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(bt != nullptr, "gt355_batch_run: null batch");
    void *scr = bt->scratch;
    hipStream_t s = (hipStream_t)stream;
    if (bt->impl >= 2) {
        const sitem *it = (const sitem *)bt->items;
        const int hx = hjb_index(bt->hjb);
        if (bt->impl == 3 && bt->nhitems[hx] > 0) {
            const hitem *hi = (const hitem *)bt->hitems[hx];
            const int nh = bt->nhitems[hx], hg = stencil_grid(bt);
#define GT355_BATCH_HDIFF(JB)                                                                                   \
    if (bt->dtype == GT355_F64)                                                                                 \
        batch_hdiff_kernel<double, double, JB, 4><<<hg, 256, 0, s>>>(bt->doms, hi, nh, xcd_count());            \
    else if (bt->precision == GT355_WIDEN)                                                                      \
        batch_hdiff_kernel<float, double, JB, 4><<<hg, 256, 0, s>>>(bt->doms, hi, nh, xcd_count());             \
    else                                                                                                        \
        batch_hdiff_kernel<float, float, JB, 4><<<hg, 256, 0, s>>>(bt->doms, hi, nh, xcd_count());
            if (bt->hjb == 8) {
                GT355_BATCH_HDIFF(8)
            } else if (bt->hjb == 32) {
                GT355_BATCH_HDIFF(32)
            } else {
                GT355_BATCH_HDIFF(16)
            }
#undef GT355_BATCH_HDIFF
            int rc = check_launch("gt355_batch_run");
            if (rc)
                return rc;
        }
        if (bt->nitems == 0)
            return GT355_OK; // stencil-phase domains only
        const int deep_nk = bt->deep_nk;
#define GT355_BATCH_STRIP(NK, WPE, HD)                                                                          \
    if (bt->dtype == GT355_F64)                                                                                 \
        batch_strip_kernel<double, double, NK, WPE, HD><<<bt->grid, 64, 0, s>>>(bt->doms, it, bt->nitems,          \
            dtr_stage, (double *)scr, xcd_count(), deep_nk);                                                    \
    else if (bt->precision == GT355_WIDEN)                                                                      \
        batch_strip_kernel<float, double, NK, WPE, HD><<<bt->grid, 64, 0, s>>>(bt->doms, it, bt->nitems,           \
            double(float(dtr_stage)), (double *)scr, xcd_count(), deep_nk);                                     \
    else                                                                                                        \
        batch_strip_kernel<float, float, NK, WPE, HD><<<bt->grid, 64, 0, s>>>(bt->doms, it, bt->nitems,            \
            float(dtr_stage), (float *)scr, xcd_count(), deep_nk);
        const bool hd = bt->impl == 2;
        if (bt->nk_reg == 32) {
            if (hd) {
                GT355_BATCH_STRIP(32, 2, true)
            } else {
                GT355_BATCH_STRIP(32, 2, false)
            }
        } else {
            if (hd) {
                GT355_BATCH_STRIP(80, 1, true)
            } else {
                GT355_BATCH_STRIP(80, 1, false)
            }
        }
#undef GT355_BATCH_STRIP
        return check_launch("gt355_batch_run");
    }
    const bitem *it = (const bitem *)bt->items;
    const int D = 4;
    if (bt->dtype == GT355_F64)
        batch_kernel<double, double, D><<<bt->grid, 256, 0, s>>>(bt->doms, it, bt->nitems, dtr_stage,
            (double *)scr, bt->rows, xcd_count());
    else if (bt->precision == GT355_WIDEN)
        batch_kernel<float, double, D><<<bt->grid, 256, 0, s>>>(bt->doms, it, bt->nitems,
            double(float(dtr_stage)), (double *)scr, bt->rows, xcd_count());
    else
        batch_kernel<float, float, D><<<bt->grid, 256, 0, s>>>(bt->doms, it, bt->nitems, float(dtr_stage),
            (float *)scr, bt->rows, xcd_count());
    return check_launch("gt355_batch_run");
}

extern "C" int gt355_batch_destroy(gt355_batch *bt) {
    if (!bt)
        return GT355_OK;
    (void)hipFree(bt->doms);
    (void)hipFree(bt->items);
    for (void *h : bt->hitems)
        (void)hipFree(h);
    (void)hipFree(bt->scratch);
    delete bt;
    return GT355_OK;
}

extern "C" int gt355_batch_info(const gt355_batch *bt, int *nitems, int *grid) {
    GT355_CHECK(bt != nullptr, "gt355_batch_info: null batch");
    if (nitems)
        *nitems = bt->nitems;
    if (grid)
        *grid = bt->grid;
    return GT355_OK;
}

namespace {
    // mean microseconds of `reps` runs (after one warm-up run) in the batch's current configuration
    int time_runs(gt355_batch *bt, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, double *us) {
        int rc = gt355_batch_run(bt, 0.15, s);
        if (rc)
            return rc;
        GT355_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps && rc == GT355_OK; ++r)
            rc = gt355_batch_run(bt, 0.15, s);
        if (rc)
            return rc;
        GT355_HIP(hipEventRecord(e1, s));
        GT355_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GT355_HIP(hipEventElapsedTime(&ms, e0, e1));
        *us = 1e3 * ms / reps;
        return GT355_OK;
    }
} // namespace

extern "C" int gt355_batch_tune(gt355_batch *bt, int reps, void *stream, int *wg_per_cu, double *best_us) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(bt != nullptr, "gt355_batch_tune: null batch");
    GT355_CHECK(reps >= 1, "gt355_batch_tune: reps %d", reps);
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    GT355_HIP(hipEventCreate(&e0));
    GT355_HIP(hipEventCreate(&e1));
    int rc = GT355_OK;
    double best_t = 1e300, t = 0;
    // the persistent grid of the vertical phase (or of the single-launch forms), then, for FP32 storage, the depth
    // from which the vertical phase keeps 96 levels in registers
    const int deep_opts[4] = {81, 89, 97, 1 << 20};
    const bool deep = bt->impl == 3 && bt->dtype == GT355_F32 && bt->nk_reg == 80;
    int best_w = bt->w, best_deep = bt->deep_nk;
    for (int w = 1; w <= BWG_MAX && rc == GT355_OK; ++w) {
        if (w > 1 && batch_grid(bt, w) == batch_grid(bt, w - 1))
            break; // fewer items than workgroups: larger grids are the same grid
        bt->grid = batch_grid(bt, w);
        for (int o = 0; o < (deep ? 4 : 1) && rc == GT355_OK; ++o) {
            if (deep)
                bt->deep_nk = deep_opts[o];
            rc = time_runs(bt, reps, s, e0, e1, &t);
            if (rc == GT355_OK && t < best_t)
                best_t = t, best_w = w, best_deep = bt->deep_nk;
        }
    }
    bt->w = best_w;
    bt->grid = batch_grid(bt, best_w);
    bt->deep_nk = best_deep;
    // the two-phase form's stencil phase: rows per item and workgroups per CU
    if (bt->impl == 3 && rc == GT355_OK) {
        int best_hjb = bt->hjb, best_hw = bt->hw;
        const int hws[3] = {2, 4, 8};
        for (int h = 0; h < 3 && rc == GT355_OK; ++h)
            for (int q = 0; q < 3 && rc == GT355_OK; ++q) {
                bt->hjb = HJB_OF[h];
                bt->hw = hws[q];
                rc = time_runs(bt, reps, s, e0, e1, &t);
                if (rc == GT355_OK && t < best_t)
                    best_t = t, best_hjb = bt->hjb, best_hw = bt->hw;
            }
        bt->hjb = best_hjb;
        bt->hw = best_hw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != GT355_OK)
        return rc;
    record_batch_cfg(bt->sig, batch_st(bt), batch_ct(bt), bt->impl, batch_cfg{bt->hjb, bt->hw, bt->w, bt->deep_nk},
        best_t);
    if (wg_per_cu)
        *wg_per_cu = bt->w;
    if (best_us)
        *best_us = best_t;
    return GT355_OK;
}

extern "C" int gt355_batch_get_config(const gt355_batch *bt, int32_t cfg[5]) {
    GT355_CHECK(bt != nullptr && cfg != nullptr, "gt355_batch_get_config: null argument");
    cfg[0] = bt->impl;
    cfg[1] = bt->hjb;
    cfg[2] = bt->hw;
    cfg[3] = bt->w;
    cfg[4] = bt->deep_nk;
    return GT355_OK;
}

extern "C" int gt355_batch_set_config(gt355_batch *bt, const int32_t cfg[5]) {
    GT355_CHECK(bt != nullptr && cfg != nullptr, "gt355_batch_set_config: null argument");
    GT355_CHECK(cfg[0] == bt->impl, "gt355_batch_set_config: the form (%d) is fixed at create time (%d)", cfg[0],
        bt->impl);
    GT355_CHECK(cfg[1] == 8 || cfg[1] == 16 || cfg[1] == 32, "gt355_batch_set_config: rows per item %d", cfg[1]);
    GT355_CHECK(cfg[2] >= 1 && cfg[2] <= 8, "gt355_batch_set_config: stencil workgroups per CU %d", cfg[2]);
    GT355_CHECK(cfg[3] >= 1 && cfg[3] <= BWG_MAX, "gt355_batch_set_config: grid setting %d", cfg[3]);
    GT355_CHECK(cfg[4] >= 2, "gt355_batch_set_config: deep_nk %d", cfg[4]);
    bt->hjb = cfg[1];
    bt->hw = cfg[2];
    bt->w = cfg[3];
    bt->grid = batch_grid(bt, bt->w);
    bt->deep_nk = cfg[4];
    return GT355_OK;
}
