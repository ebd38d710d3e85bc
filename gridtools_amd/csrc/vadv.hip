// Vertical advection (implicit tridiagonal Thomas solve per column) for gfx950.
//
// Reference: tests/regression/vertical_advection_dycore.cpp:32-150.  The spec is
//   multi_pass(execute_forward().k_cached(flush, ccol, dcol).k_cached(fill, u_stage).stage(u_forward...),
//              execute_backward().k_cached(data_col).stage(u_backward...))
// and the reference gpu backend fuses both multistages into one launch (stencil/gpu/entry_point.hpp:48-59,
// 107-124) with one thread per column, u_stage/data_col in register rings (gpu/k_cache.hpp:23-46) and
// ccol/dcol flushed to HBM temporaries that the backward sweep reads back (gpu/fill_flush.hpp:298-321).
//
// Here, too, one lane owns one column and both sweeps run in one launch, with the k-caches (u_stage ring,
// wcon at k and k+1, ccol/dcol at k-1, data_col at k+1) in VGPRs.  The forward sweep's ccol/dcol for the
// backward sweep are kept on chip when nk allows it (see the ONCHIP variant), otherwise they go to a
// per-launch scratch buffer in the k-major layout the lanes read coalesced.
// Arithmetic follows the functors operation by operation (built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "functors.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {

        // One column per lane; ccol/dcol of levels 0..nk-2 go to scratch (k-major, column-minor).
        template <class ST, class CT>
        __global__ void __launch_bounds__(256) vadv_scratch_kernel(ST *__restrict__ uts_f,
            const ST *__restrict__ us_f, const ST *__restrict__ wc_f, const ST *__restrict__ up_f,
            const ST *__restrict__ ut_f, CT dtr, int ni, int nj, int nk, int64_t sj, int64_t sk, int64_t wsj,
            int64_t wsk, CT *__restrict__ cd_buf, int64_t plane) {
            const int i = blockIdx.x * 64 + (threadIdx.x & 63);
            const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
            if (i >= ni || j >= nj)
                return;
            const int64_t col = (int64_t)j * sj + i;
            const int64_t wcol = (int64_t)j * wsj + i;
            const int64_t scol = (int64_t)j * ni + i; // scratch column index, plane = ni*nj
            CT *__restrict__ sc = cd_buf + scol;
            CT *__restrict__ sd = cd_buf + (int64_t)(nk - 1) * plane + scol;

            // k = 0
            CT us_c = CT(us_f[col]);
            CT us_p = CT(us_f[col + sk]);
            CT w_p_i = CT(wc_f[wcol + wsk]);
            CT w_p_ip = CT(wc_f[wcol + wsk + 1]);
            CT up = CT(up_f[col]);
            fwd_out<CT> r = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, up, CT(ut_f[col]), CT(uts_f[col]));
            sc[0] = r.c;
            sd[0] = r.d;
            CT cc_m = r.c, dc_m = r.d;
#pragma unroll 4
            for (int k = 1; k < nk - 1; ++k) {
                const int64_t p = col + (int64_t)k * sk;
                CT us_m = us_c;
                us_c = us_p;
                us_p = CT(us_f[p + sk]);
                CT w_c_i = w_p_i, w_c_ip = w_p_ip;
                w_p_i = CT(wc_f[wcol + (int64_t)(k + 1) * wsk]);
                w_p_ip = CT(wc_f[wcol + (int64_t)(k + 1) * wsk + 1]);
                up = CT(up_f[p]);
                r = forward_body<CT>(dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, up, CT(ut_f[p]),
                    CT(uts_f[p]), cc_m, dc_m, false);
                sc[(int64_t)k * plane] = r.c;
                sd[(int64_t)k * plane] = r.d;
                cc_m = r.c;
                dc_m = r.d;
            }
            // k = nk-1
            {
                const int64_t p = col + (int64_t)(nk - 1) * sk;
                CT us_m = us_c;
                us_c = us_p;
                up = CT(up_f[p]);
                CT dc = forward_last<CT>(
                    dtr, w_p_i, w_p_ip, us_m, us_c, up, CT(ut_f[p]), CT(uts_f[p]), cc_m, dc_m);
                // u_backward_function::apply(last_level) (vertical_advection_dycore.cpp:117-121)
                uts_f[p] = ST(dtr * (dc - up));
                CT data = dc;
                // u_backward_function::apply(modify<0,-1>) (vertical_advection_dycore.cpp:110-115)
#pragma unroll 4
                for (int k = nk - 2; k >= 0; --k) {
                    const int64_t q = col + (int64_t)k * sk;
                    CT cc = sc[(int64_t)k * plane];
                    CT dd = sd[(int64_t)k * plane];
                    CT dt = dd - cc * data;
                    uts_f[q] = ST(dtr * (dt - CT(up_f[q])));
                    data = dt;
                }
            }
        }

        // One wavefront per workgroup, one column per lane, k-depth bucket NK >= nk known at compile time.
        // ccol/dcol of the whole column stay in registers between the sweeps (the compiler places them in the
        // unified VGPR/AGPR file: at one wave per SIMD a lane owns 512 registers), so neither sweep touches a
        // temporary in memory.  With one wave per SIMD nothing else hides memory latency, so every input is
        // software-pipelined D levels ahead through register rings:
        //   ring A (u_stage, wcon at i and i+1) of level L is consumed at forward step L-1,
        //   ring B (utens_stage, u_pos, utens) of level L at forward step L.
        // The sweeps are straight-line code: loops are fully unrolled over NK, every load is unconditional (levels
        // past the runtime top nk-1 are clamped to it, so they re-read a cached line), and the first / interior /
        // top-level formulas are chosen by value, not by branch.  That keeps the hardware's in-order load counter
        // exact across the unrolled steps, so the wait before step k covers only level k's loads.
        // u_pos is needed again by the backward sweep.  UPOS_LDS parks the column's u_pos in LDS (NK*64*sizeof(ST)
        // bytes per wavefront; nothing is re-read from memory); otherwise the backward sweep re-loads u_pos through
        // its own D-deep ring.  Lanes past ni read the last valid column and store nothing.
        // SBM: instruction classes the per-step scheduling barrier lets through (0: none; 6: VALU and SALU, so
        // the next step's independent arithmetic can fill the latency of this step's pivot chain while memory
        // operations stay at their step)
        template <class ST, class CT, int NK, int D, bool UPOS_LDS, int SBM = 0>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        vadv_onchip_kernel(ST *__restrict__ uts_f, const ST *__restrict__ us_f, const ST *__restrict__ wc_f,
            const ST *__restrict__ up_f, const ST *__restrict__ ut_f, CT dtr, int ni, int nk, int64_t sj, int64_t sk,
            int64_t wsj, int64_t wsk) {
            static_assert(D >= 1 && D <= NK, "prefetch depth");
            const int i = min(int(blockIdx.x * 64 + threadIdx.x), ni - 1);
            const int j = blockIdx.y;
            const int top = nk - 1;
            // buffer addressing: per-field descriptors, a per-level byte offset shared by the fields of one
            // layout, the lane's column as a byte offset (the launcher checks every offset stays below 4 GiB)
            const uint32_t colb = (uint32_t)(((int64_t)j * sj + i) * (int64_t)sizeof(ST));
            const uint32_t wcolb = (uint32_t)(((int64_t)j * wsj + i) * (int64_t)sizeof(ST));
            const rsrc_t r_us = make_rsrc(us_f), r_wc = make_rsrc(wc_f), rs_up = make_rsrc(up_f),
                         r_ut = make_rsrc(ut_f), r_uts = make_rsrc(uts_f);
            auto lvl = [&](int L) { return (uint32_t)min(L, top); };
            // The loads read memory nothing in the forward sweep writes, so the compiler would treat them as
            // invariant and hoist all of them to the kernel entry (one register per in-flight value: ~800 VGPRs).
            // Routing the strides through an empty volatile asm at every step gives each step's loads an opaque
            // dependency that only exists from that step on.
            uint32_t sk_o = (uint32_t)(sk * (int64_t)sizeof(ST)), wsk_o = (uint32_t)(wsk * (int64_t)sizeof(ST));
            auto pin = [&]() {
                asm volatile("" : "+s"(sk_o));
                asm volatile("" : "+s"(wsk_o));
            };

            ST a_us[D], a_wi[D], a_wip[D]; // ring A
            ST b_uts[D], b_up[D], b_ut[D]; // ring B
            auto load_a = [&](int L) {
                const uint32_t o = lvl(L) * sk_o, w = lvl(L) * wsk_o;
                a_us[L % D] = bload<ST, CPOL_NT>(r_us, colb, o);
                const two<ST> ww = bload2<ST, CPOL_NT>(r_wc, wcolb, w); // wcon at i and i+1
                a_wi[L % D] = ww.a;
                a_wip[L % D] = ww.b;
            };
            auto load_b = [&](int L) {
                const uint32_t o = lvl(L) * sk_o;
                b_uts[L % D] = bload<ST, CPOL_NT>(r_uts, colb, o);
                b_up[L % D] = bload<ST, CPOL_NT>(rs_up, colb, o);
                b_ut[L % D] = bload<ST, CPOL_NT>(r_ut, colb, o);
            };
#pragma clang loop unroll(full)
            for (int L = 0; L < D; ++L) {
                load_b(L);
                load_a(L + 1);
            }

            __shared__ ST s_up[UPOS_LDS ? NK : 1][64];
            CT c[NK], d[NK];
            CT us_m = CT(0), us_c = CT(bload<ST, CPOL_NT>(r_us, colb, 0)), w_c_i = CT(0), w_c_ip = CT(0);
#pragma clang loop unroll(full)
            for (int k = 0; k < NK; ++k) {
                if constexpr (UPOS_LDS)
                    s_up[k][threadIdx.x] = b_up[k % D];
                const CT v_uts = CT(b_uts[k % D]), v_up = CT(b_up[k % D]), v_ut = CT(b_ut[k % D]);
                // at the top level the k+1 wind is taken as zero: with gcv = cs = c = 0 the interior formula computes
                // b = (dtr - a) - 0 and correction - 0*(us_p - us_c), i.e. exactly the last_level functor's values
                // for finite inputs (us_p is set to us_c, so the dropped term is 0*(+0) = +0 and x - (+0) == x)
                const bool at_top = k == top;
                const CT us_p = at_top ? us_c : CT(a_us[(k + 1) % D]);
                const CT w_p_i = at_top ? CT(0) : CT(a_wi[(k + 1) % D]);
                const CT w_p_ip = at_top ? CT(0) : CT(a_wip[(k + 1) % D]);
                // refill the slots just consumed
                pin();
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
                    c[k] = r.c; // never read at the top level
                    d[k] = r.d;
                }
                us_m = us_c;
                us_c = us_p;
                w_c_i = w_p_i;
                w_c_ip = w_p_ip;
                // one scheduling region per step for memory operations: the loads stay where the ring puts them
                // (D levels ahead) instead of being hoisted wholesale, which would need a register per in-flight
                // value
                __builtin_amdgcn_sched_barrier(SBM);
            }

            ST r_up[D]; // backward u_pos ring (!UPOS_LDS): level L consumed at backward step L, loaded at step L + D
            if constexpr (!UPOS_LDS) {
#pragma clang loop unroll(full)
                for (int s = 0; s < D; ++s)
                    r_up[(NK - 1 - s) % D] = bload<ST, CPOL_NT>(rs_up, colb, lvl(NK - 1 - s) * sk_o);
            }
            // UPOS_LDS: the parked u_pos is read UA levels ahead of its use, so the backward chain does not wait for
            // each LDS read (with one wavefront per SIMD nothing else would hide its latency)
            constexpr int UA = NK < 4 ? NK : 4;
            ST l_up[UA];
            if constexpr (UPOS_LDS) {
#pragma clang loop unroll(full)
                for (int s = 0; s < UA; ++s)
                    l_up[(NK - 1 - s) % UA] = s_up[NK - 1 - s][threadIdx.x];
            }
            CT data = CT(0);
#pragma clang loop unroll(full)
            for (int k = NK - 1; k >= 0; --k) {
                CT v_up;
                if constexpr (UPOS_LDS) {
                    v_up = CT(l_up[k % UA]);
                    if (k - UA >= 0)
                        l_up[k % UA] = s_up[k - UA][threadIdx.x];
                } else {
                    v_up = CT(r_up[k % D]);
                    pin();
                    if (k - D >= 0)
                        r_up[k % D] = bload<ST, CPOL_NT>(rs_up, colb, lvl(k - D) * sk_o);
                }
                // u_backward_function (vertical_advection_dycore.cpp:110-121): the top level starts the recurrence
                const CT dt = d[k] - c[k] * data;
                const CT nd = k == top ? d[k] : dt;
                const CT out = dtr * (nd - v_up);
                data = nd;
                // Unconditional store to the clamped level: a step above the top writes a value that the step at
                // the top overwrites later in program order, and lanes past ni compute column ni-1 and write its
                // bits again.  No branch, so the load counter stays exact.
                bstore<ST, CPOL_OUT>(r_uts, ST(out), colb, lvl(k) * sk_o);
                __builtin_amdgcn_sched_barrier(SBM);
            }
        }

        // Persistent form of vadv_onchip_kernel (u_pos parked in LDS) for domains with more 64-column strips than
        // one wavefront per SIMD covers at once: a grid of one wavefront per SIMD walks the strips (64 columns of one
        // row; the strips of a row consecutive) with a static stride, so the wavefronts in flight stay on
        // neighbouring strips of the same levels.  Each wavefront issues the next strip's first D levels of loads
        // between the current strip's forward and backward sweeps; the backward sweep only stores, so those loads
        // are in flight while it runs and the next forward sweep starts without a pipeline fill (one wavefront per
        // SIMD has nothing else to hide it).  The column sweep itself is vadv_onchip_kernel's, operation by
        // operation.
        template <class ST, class CT, int NK, int D>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        vadv_persist_kernel(ST *__restrict__ uts_f, const ST *__restrict__ us_f, const ST *__restrict__ wc_f,
            const ST *__restrict__ up_f, const ST *__restrict__ ut_f, CT dtr, int ni, int nk, int64_t sj, int64_t sk,
            int64_t wsj, int64_t wsk, int strips_i, int nitems, int nxcd) {
            static_assert(D >= 1 && D <= NK, "prefetch depth");
            const int top = nk - 1;
            const rsrc_t r_us = make_rsrc(us_f), r_wc = make_rsrc(wc_f), rs_up = make_rsrc(up_f),
                         r_ut = make_rsrc(ut_f), r_uts = make_rsrc(uts_f);
            auto lvl = [&](int L) { return (uint32_t)min(L, top); };
            uint32_t sk_o = (uint32_t)(sk * (int64_t)sizeof(ST)), wsk_o = (uint32_t)(wsk * (int64_t)sizeof(ST));
            auto pin = [&]() {
                asm volatile("" : "+s"(sk_o));
                asm volatile("" : "+s"(wsk_o));
            };
            // the lane's byte offsets in strip `it` (lanes past ni take column ni-1)
            auto offsets = [&](int it, uint32_t &cb, uint32_t &wb) {
                const int si = it % strips_i, j = it / strips_i;
                const int i = min(si * 64 + (int)threadIdx.x, ni - 1);
                cb = (uint32_t)(((int64_t)j * sj + i) * (int64_t)sizeof(ST));
                wb = (uint32_t)(((int64_t)j * wsj + i) * (int64_t)sizeof(ST));
            };
            ST a_us[D], a_wi[D], a_wip[D]; // ring A: u_stage, wcon at i and i+1 of level L, consumed at step L-1
            ST b_uts[D], b_up[D], b_ut[D]; // ring B: consumed at step L
            ST us0;                        // u_stage of level 0
            auto load_a = [&](int L, uint32_t cb, uint32_t wb) {
                const uint32_t o = lvl(L) * sk_o, w = lvl(L) * wsk_o;
                a_us[L % D] = bload<ST, CPOL_NT>(r_us, cb, o);
                // wcon at i and i + 1: with many strips per wavefront (the 16-level variant) the neighbouring
                // strip reads the column-64 line soon after, from L2 with the default policy (1536x1536x80 timestep
                // 1.40 -> 1.36 ms); with few (8 levels ahead) non-temporal loads measured faster (512x512x80
                // composition 156 against 160 us)
                const two<ST> ww = bload2<ST, D == 16 ? 0 : CPOL_NT>(r_wc, wb, w);
                a_wi[L % D] = ww.a;
                a_wip[L % D] = ww.b;
            };
            auto load_b = [&](int L, uint32_t cb) {
                const uint32_t o = lvl(L) * sk_o;
                b_uts[L % D] = bload<ST, CPOL_NT>(r_uts, cb, o);
                b_up[L % D] = bload<ST, CPOL_NT>(rs_up, cb, o);
                b_ut[L % D] = bload<ST, CPOL_NT>(r_ut, cb, o);
            };
            auto prologue = [&](uint32_t cb, uint32_t wb) {
                us0 = bload<ST, CPOL_NT>(r_us, cb, 0);
#pragma clang loop unroll(full)
                for (int L = 0; L < D; ++L) {
                    load_b(L, cb);
                    load_a(L + 1, cb, wb);
                }
            };
            __shared__ ST s_up[NK][64];
            int it = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            if (it >= nitems)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            uint32_t colb, wcolb;
            offsets(it, colb, wcolb);
            prologue(colb, wcolb);
            for (; it < nitems; it += gridDim.x) {
                CT c[NK], d[NK];
                CT us_m = CT(0), us_c = CT(us0), w_c_i = CT(0), w_c_ip = CT(0);
#pragma clang loop unroll(full)
                for (int k = 0; k < NK; ++k) {
                    s_up[k][threadIdx.x] = b_up[k % D];
                    const CT v_uts = CT(b_uts[k % D]), v_up = CT(b_up[k % D]), v_ut = CT(b_ut[k % D]);
                    // top level: the k+1 wind taken as zero reproduces last_level exactly (see vadv_onchip_kernel)
                    const bool at_top = k == top;
                    const CT us_p = at_top ? us_c : CT(a_us[(k + 1) % D]);
                    const CT w_p_i = at_top ? CT(0) : CT(a_wi[(k + 1) % D]);
                    const CT w_p_ip = at_top ? CT(0) : CT(a_wip[(k + 1) % D]);
                    pin();
                    if (k + D < NK)
                        load_b(k + D, colb);
                    if (k + 1 + D < NK)
                        load_a(k + 1 + D, colb, wcolb);
                    if (k == 0) {
                        fwd_out<CT> r = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                        asm volatile("" : "+v"(r.c), "+v"(r.d)); // keeps the chain at its step (no sinking)
                        c[0] = r.c;
                        d[0] = r.d;
                    } else {
                        const CT cm = c[k > 0 ? k - 1 : 0], dm = d[k > 0 ? k - 1 : 0];
                        fwd_out<CT> r = forward_body<CT>(
                            dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, v_up, v_ut, v_uts, cm, dm, at_top);
                        asm volatile("" : "+v"(r.c), "+v"(r.d));
                        c[k] = r.c;
                        d[k] = r.d;
                    }
                    us_m = us_c;
                    us_c = us_p;
                    w_c_i = w_p_i;
                    w_c_ip = w_p_ip;
                    __builtin_amdgcn_sched_barrier(0);
                }
                // the next strip's first levels, in flight during this strip's backward sweep
                const int nx = it + (int)gridDim.x;
                uint32_t ncolb = colb, nwcolb = wcolb;
                if (nx < nitems) {
                    offsets(nx, ncolb, nwcolb);
                    pin();
                    prologue(ncolb, nwcolb);
                }
                // the parked u_pos read four levels ahead of its use (see vadv_onchip_kernel)
                ST l_up[4];
#pragma clang loop unroll(full)
                for (int s = 0; s < 4; ++s)
                    l_up[(NK - 1 - s) % 4] = s_up[NK - 1 - s][threadIdx.x];
                CT data = CT(0);
#pragma clang loop unroll(full)
                for (int k = NK - 1; k >= 0; --k) {
                    const CT v_up = CT(l_up[k % 4]);
                    if (k - 4 >= 0)
                        l_up[k % 4] = s_up[k - 4][threadIdx.x];
                    // u_backward_function (vertical_advection_dycore.cpp:110-121)
                    const CT dt = d[k] - c[k] * data;
                    const CT nd = k == top ? d[k] : dt;
                    data = nd;
                    bstore<ST, CPOL_OUT>(r_uts, ST(dtr * (nd - v_up)), colb, lvl(k) * sk_o);
                    __builtin_amdgcn_sched_barrier(0);
                }
                colb = ncolb;
                wcolb = nwcolb;
            }
        }

        // FP32-storage form of the persistent sweep with 16-byte requests.  A 64-column row of one FP32 field is
        // 256 bytes, so one dwordx4 load instruction of a wavefront fetches four rows: lanes 16g..16g+15 read row g,
        // four columns each.  Per group of four levels a wavefront issues six loads (utens_stage, u_pos, utens at
        // levels 4G..4G+3; u_stage and wcon at 4G+1..4G+4, one level ahead as the forward step needs them; wcon's
        // column 64, the i+1 neighbour of the strip's last column, as one dword per row), against 20 single-dword
        // loads with one row per instruction.  The rows pass through LDS (a per-wavefront staging area; the u_pos
        // rows land in the column's u_pos park, which the backward sweep reads) and each lane reads its own column
        // back.  Outputs go the same way in reverse: four levels of a lane's column into LDS, then one dwordx4
        // store of four rows.  Fewer, larger requests keep more bytes in flight under the wavefront's limit of
        // outstanding memory instructions.  Only one wavefront uses each LDS area, whose operations it executes in
        // order, so no barrier is needed; an empty asm with a memory clobber keeps the compiler from reordering the
        // LDS writes and reads of one group.  The arithmetic is vadv_onchip_kernel's, operation by operation.
        // Requirements checked by the launcher: ni % 4 == 0 (a lane's four output columns are all inside the domain
        // or, past ni, clamped onto the last four, which it writes with the same bits as their owner); 16-byte
        // aligned rows and level planes; every row read whole (columns up to the strip's end, and column 64 of wcon)
        // inside each field's allocation.
        template <class CT, int NK, int DG>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        vadv_wide_kernel(float *__restrict__ uts_f, const float *__restrict__ us_f, const float *__restrict__ wc_f,
            const float *__restrict__ up_f, const float *__restrict__ ut_f, CT dtr, int ni, int nk, int64_t sj,
            int64_t sk, int64_t wsj, int64_t wsk, int strips_i, int nitems, int nxcd) {
            constexpr int NG = NK / 4;
            static_assert(NK % 4 == 0 && DG >= 1 && DG <= NG, "level groups");
            constexpr int RS = 68; // staging row: 64 columns + wcon's column 64, padded to 16 bytes
            const int lane = threadIdx.x, g = lane >> 4, q = lane & 15;
            const int top = nk - 1;
            const rsrc_t r_us = make_rsrc(us_f), r_wc = make_rsrc(wc_f), rs_up = make_rsrc(up_f),
                         r_ut = make_rsrc(ut_f), r_uts = make_rsrc(uts_f);
            uint32_t skb = (uint32_t)(sk * 4), wskb = (uint32_t)(wsk * 4);
            auto pin = [&]() {
                asm volatile("" : "+s"(skb));
                asm volatile("" : "+s"(wskb));
            };
            auto lds_order = []() { asm volatile("" ::: "memory"); };
            __shared__ __align__(16) float s_up[NK * 64];
            __shared__ __align__(16) float s_us[4 * RS], s_wc[4 * RS], s_uts[4 * RS], s_ut[4 * RS];
            __shared__ __align__(16) float s_out[4 * 64];

            struct group {
                f32x4 uts, up, ut, us, wc;
                float wx;
            };
            group ring[DG];
            // per strip: row-load offsets of this lane's four columns (lb, wlb), wcon's column 64 (wxb), the lane's
            // own column (colb; lanes past ni take column ni-1 through p) and its output chunk (sb, sc)
            uint32_t lb = 0, wlb = 0, wxb = 0, colb = 0, sb = 0;
            int p = 0, sc = 0;
            auto offsets = [&](int it) {
                const int si = it % strips_i, j = it / strips_i, i0 = si * 64;
                lb = (uint32_t)(((int64_t)j * sj + i0 + 4 * q) * 4);
                wlb = (uint32_t)(((int64_t)j * wsj + i0 + 4 * q) * 4);
                wxb = (uint32_t)(((int64_t)j * wsj + i0 + 64) * 4);
                p = min(lane, ni - 1 - i0);
                colb = (uint32_t)(((int64_t)j * sj + i0 + p) * 4);
                sc = min(4 * q, ni - 4 - i0);
                sb = (uint32_t)(((int64_t)j * sj + i0 + sc) * 4);
            };
            auto issue = [&](group &r, int G) {
                const uint32_t lB = (uint32_t)min(4 * G + g, top), lA = (uint32_t)min(4 * G + 1 + g, top);
                r.uts = bload4<CPOL_NT>(r_uts, lb + lB * skb, 0);
                r.up = bload4<CPOL_NT>(rs_up, lb + lB * skb, 0);
                r.ut = bload4<CPOL_NT>(r_ut, lb + lB * skb, 0);
                r.us = bload4<CPOL_NT>(r_us, lb + lA * skb, 0);
                r.wc = bload4<CPOL_NT>(r_wc, wlb + lA * wskb, 0);
                r.wx = bload<float, CPOL_NT>(r_wc, wxb + lA * wskb, 0);
            };
            auto stage = [&](group const &r, int G) {
                *(f32x4 *)&s_uts[g * RS + 4 * q] = r.uts;
                *(f32x4 *)&s_ut[g * RS + 4 * q] = r.ut;
                *(f32x4 *)&s_up[G * 256 + 4 * lane] = r.up; // rows 4G..4G+3 of the park are contiguous
                *(f32x4 *)&s_us[g * RS + 4 * q] = r.us;
                *(f32x4 *)&s_wc[g * RS + 4 * q] = r.wc;
                s_wc[g * RS + 64] = r.wx; // the 16 lanes of a row write the same value
            };
            float us0 = 0.f; // u_stage of level 0 at the lane's column
            auto prologue = [&]() {
                us0 = bload<float, CPOL_NT>(r_us, colb, 0);
#pragma clang loop unroll(full)
                for (int G = 0; G < DG; ++G)
                    issue(ring[G], G);
            };

            // a group's values of the lane's column, read from LDS one group ahead of their use so that the LDS
            // latency is hidden behind the previous group's arithmetic
            struct vals {
                float uts[4], up[4], ut[4], us[4], wi[4], wip[4];
            };
            auto lds_vals = [&](vals &v, int G) {
#pragma clang loop unroll(full)
                for (int r = 0; r < 4; ++r) {
                    v.uts[r] = s_uts[r * RS + p];
                    v.up[r] = s_up[(4 * G + r) * 64 + p];
                    v.ut[r] = s_ut[r * RS + p];
                    v.us[r] = s_us[r * RS + p];
                    v.wi[r] = s_wc[r * RS + p];
                    v.wip[r] = s_wc[r * RS + p + 1];
                }
            };

            int it = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            if (it >= nitems)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            offsets(it);
            prologue();
            for (; it < nitems; it += gridDim.x) {
                CT c[NK], d[NK];
                CT us_m = CT(0), us_c = CT(us0), w_c_i = CT(0), w_c_ip = CT(0);
                vals a, b;
                // group 0: staged and read before the sweep; its ring slot refilled with group DG
                lds_order();
                stage(ring[0], 0);
                pin();
                if (DG < NG)
                    issue(ring[0], DG);
                lds_order();
                lds_vals(a, 0);
#pragma clang loop unroll(full)
                for (int G = 0; G < NG; ++G) {
                    // stage group G+1 (the staging rows of group G were read into `a` before, in program order),
                    // refill its ring slot with group G+1+DG and read its values, all in flight while group G computes
                    if (G + 1 < NG) {
                        lds_order();
                        stage(ring[(G + 1) % DG], G + 1);
                        pin();
                        if (G + 1 + DG < NG)
                            issue(ring[(G + 1) % DG], G + 1 + DG);
                        lds_order();
                        lds_vals(b, G + 1);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
                    for (int r = 0; r < 4; ++r) {
                        const int k = 4 * G + r;
                        const CT v_uts = CT(a.uts[r]), v_up = CT(a.up[r]), v_ut = CT(a.ut[r]);
                        // top level: the k+1 wind taken as zero reproduces last_level exactly (see
                        // vadv_onchip_kernel)
                        const bool at_top = k == top;
                        const CT us_p = at_top ? us_c : CT(a.us[r]);
                        const CT w_p_i = at_top ? CT(0) : CT(a.wi[r]);
                        const CT w_p_ip = at_top ? CT(0) : CT(a.wip[r]);
                        if (k == 0) {
                            fwd_out<CT> o = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                            asm volatile("" : "+v"(o.c), "+v"(o.d));
                            c[0] = o.c;
                            d[0] = o.d;
                        } else {
                            const CT cm = c[k > 0 ? k - 1 : 0], dm = d[k > 0 ? k - 1 : 0];
                            fwd_out<CT> o = forward_body<CT>(dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p,
                                v_up, v_ut, v_uts, cm, dm, at_top);
                            asm volatile("" : "+v"(o.c), "+v"(o.d));
                            c[k] = o.c;
                            d[k] = o.d;
                        }
                        us_m = us_c;
                        us_c = us_p;
                        w_c_i = w_p_i;
                        w_c_ip = w_p_ip;
                    }
                    a = b;
                    __builtin_amdgcn_sched_barrier(0);
                }
                // the next strip's first groups, in flight during this strip's backward sweep
                const int nx = it + (int)gridDim.x;
                const uint32_t csb = sb;
                const int cp = p, csc = sc;
                if (nx < nitems) {
                    offsets(nx);
                    pin();
                    prologue();
                }
                // backward: u_pos of the next group read while this group computes; the four output rows of a group
                // read from LDS at its end and stored at the end of the next group (after the read has completed)
                float upa[4], upb[4];
#pragma clang loop unroll(full)
                for (int r = 0; r < 4; ++r)
                    upa[r] = s_up[(NK - 4 + r) * 64 + cp];
                f32x4 o_prev = {0.f, 0.f, 0.f, 0.f};
                uint32_t so_prev = 0;
                CT data = CT(0);
#pragma clang loop unroll(full)
                for (int G = NG - 1; G >= 0; --G) {
                    if (G > 0) {
#pragma clang loop unroll(full)
                        for (int r = 0; r < 4; ++r)
                            upb[r] = s_up[(4 * G - 4 + r) * 64 + cp];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
                    for (int r = 3; r >= 0; --r) {
                        const int k = 4 * G + r;
                        // u_backward_function (vertical_advection_dycore.cpp:110-121)
                        const CT dt = d[k] - c[k] * data;
                        const CT nd = k == top ? d[k] : dt;
                        data = nd;
                        s_out[r * 64 + lane] = float(dtr * (nd - CT(upa[r])));
                    }
                    // levels 4G..4G+3: row g holds level 4G+g; a row above the top is stored as the top row (the same
                    // bits its own lanes store) to the top level, which the group holding the top, stored later in
                    // program order, overwrites with the same bits or its final value
                    lds_order();
                    const int L = min(4 * G + g, top), rg = max(L - 4 * G, 0);
                    const f32x4 o = *(const f32x4 *)&s_out[rg * 64 + csc];
                    lds_order();
                    if (G < NG - 1) {
                        pin();
                        bstore4<CPOL_OUT>(r_uts, o_prev, so_prev, 0);
                    }
                    o_prev = o;
                    so_prev = csb + (uint32_t)L * skb;
#pragma clang loop unroll(full)
                    for (int r = 0; r < 4; ++r)
                        upa[r] = upb[r];
                    __builtin_amdgcn_sched_barrier(0);
                }
                pin();
                bstore4<CPOL_OUT>(r_uts, o_prev, so_prev, 0);
            }
        }

        // vadv_wide_kernel with its rows loaded straight into LDS: each dwordx4 request goes to LDS by the buffer
        // unit (buffer_load ... lds; lane l's 16 bytes land at the instruction's LDS base + 16 l, so the 16 lanes of
        // row g fill row g of a 64-column staging block), and no VGPR holds a row in flight.  The 21 registers per
        // group of the register ring are what limited vadv_wide_kernel to three groups in flight at FP64 compute
        // (four spill); here NS LDS slots hold the groups (each 4 fields x 4 rows x 64 columns), NS - 1 of them in
        // flight while a group computes.  The u_pos rows of groups loaded inside a sweep go straight to the park;
        // those of the next strip's first NS groups, loaded during the backward sweep while the park is still read,
        // take registers and are written to the park as the sweep reaches them.  wcon's column 64 (the i+1
        // neighbour of the strip's last column) comes as one dword per lane into a register, and lane 63 takes row
        // r's value from lane 16 r.  The backward sweep stages a group's output rows in the park rows of the same
        // group, whose u_pos values it has read by then.  LDS: NK x 64 x 4 bytes of park plus NS x 4 KiB of slots
        // (40 KiB at NK = 80, NS = 5: four wavefronts per CU).
        // Ordering is the kernel's own.  The compiler's waits for these LDS writes cannot be relied on: with every
        // slot in one array it waits for the newest request before any read (all groups drained), with more than
        // eight target arrays it emits no wait for the later ones, and it never waits before a refill overwrites a
        // slot whose reads may be pending.  So the six requests of a group are issued between two fences (asm
        // memory clobber + sched_barrier; their order inside the group is free); before a group's slot is read,
        // `s_waitcnt vmcnt(n)` waits until at most n requests are outstanding, n = 6 x the groups issued after it
        // (a load request's count drops only after its data is in LDS); before a slot is refilled,
        // `s_waitcnt lgkmcnt(0)` waits for the reads of its previous group.  Requests or stores issued later only
        // make a wait stricter.  One array per slot and per park group keeps the compiler's own waits no stricter
        // than these (a read then depends only on the requests into its array).  Requirements and arithmetic as
        // vadv_wide_kernel.  Measured at 256x256x80, FP64 compute: 22.7 against 22.9 us for the register ring.
        template <class CT, int NK, int NS>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        vadv_lds_kernel(float *__restrict__ uts_f, const float *__restrict__ us_f, const float *__restrict__ wc_f,
            const float *__restrict__ up_f, const float *__restrict__ ut_f, CT dtr, int ni, int nk, int64_t sj,
            int64_t sk, int64_t wsj, int64_t wsk, int strips_i, int nitems, int nxcd) {
            constexpr int NG = NK / 4;
            static_assert(NK % 4 == 0 && NS >= 2 && NS <= NG, "level groups");
            constexpr int PER_GROUP = 6; // requests per group: wcon column 64, u_pos, utens_stage, utens, u_stage, wcon
            const int lane = threadIdx.x, g = lane >> 4, q = lane & 15;
            const int top = nk - 1;
            const rsrc_t r_us = make_rsrc(us_f), r_wc = make_rsrc(wc_f), rs_up = make_rsrc(up_f),
                         r_ut = make_rsrc(ut_f), r_uts = make_rsrc(uts_f);
            uint32_t skb = (uint32_t)(sk * 4), wskb = (uint32_t)(wsk * 4);
            auto pin = [&]() {
                asm volatile("" : "+s"(skb));
                asm volatile("" : "+s"(wskb));
            };
            auto fence = []() {
                asm volatile("" ::: "memory");
                __builtin_amdgcn_sched_barrier(0);
            };
            // LDS: one array per u_pos park group (rows 4G..4G+3, row-major; a group's output rows during the
            // backward sweep) and one per slot ([utens_stage, utens, u_stage, wcon][row * 64 + column]).  Separate
            // arrays keep the compiler's own LDS-DMA tracking from ordering a read after every request in flight:
            // a read of one array can only depend on the requests into that array.
            static_assert(NG == 20 && NS == 5, "the park and slot arrays below are written out for NK = 80, NS = 5");
#define GT355_LDS_ARRAY(name, n) __shared__ __align__(16) float name##n[256];
            GT355_LDS_ARRAY(pk, 0) GT355_LDS_ARRAY(pk, 1) GT355_LDS_ARRAY(pk, 2) GT355_LDS_ARRAY(pk, 3)
            GT355_LDS_ARRAY(pk, 4) GT355_LDS_ARRAY(pk, 5) GT355_LDS_ARRAY(pk, 6) GT355_LDS_ARRAY(pk, 7)
            GT355_LDS_ARRAY(pk, 8) GT355_LDS_ARRAY(pk, 9) GT355_LDS_ARRAY(pk, 10) GT355_LDS_ARRAY(pk, 11)
            GT355_LDS_ARRAY(pk, 12) GT355_LDS_ARRAY(pk, 13) GT355_LDS_ARRAY(pk, 14) GT355_LDS_ARRAY(pk, 15)
            GT355_LDS_ARRAY(pk, 16) GT355_LDS_ARRAY(pk, 17) GT355_LDS_ARRAY(pk, 18) GT355_LDS_ARRAY(pk, 19)
#undef GT355_LDS_ARRAY
            __shared__ __align__(16) float sl0[4 * 256], sl1[4 * 256], sl2[4 * 256], sl3[4 * 256], sl4[4 * 256];
            // LDS-space pointers (a generic pointer would be converted with a null check at every request)
            typedef __attribute__((address_space(3))) float lfloat;
            typedef __attribute__((address_space(3))) f32x4 lf32x4;
#define GT355_L(a) ((lfloat *)(a))
            lfloat *const park[NG] = {GT355_L(pk0), GT355_L(pk1), GT355_L(pk2), GT355_L(pk3), GT355_L(pk4),
                GT355_L(pk5), GT355_L(pk6), GT355_L(pk7), GT355_L(pk8), GT355_L(pk9), GT355_L(pk10), GT355_L(pk11),
                GT355_L(pk12), GT355_L(pk13), GT355_L(pk14), GT355_L(pk15), GT355_L(pk16), GT355_L(pk17),
                GT355_L(pk18), GT355_L(pk19)};
            lfloat *const slot[NS] = {GT355_L(sl0), GT355_L(sl1), GT355_L(sl2), GT355_L(sl3), GT355_L(sl4)};
#undef GT355_L

            f32x4 up_r[NS]; // u_pos rows of the first NS groups of a strip (register loads)
            float wx_r[NS]; // wcon column 64 of row g = lane >> 4
            uint32_t lb = 0, wlb = 0, wxb = 0, colb = 0, sb = 0;
            int p = 0, sc = 0;
            auto offsets = [&](int it) {
                const int si = it % strips_i, j = it / strips_i, i0 = si * 64;
                lb = (uint32_t)(((int64_t)j * sj + i0 + 4 * q) * 4);
                wlb = (uint32_t)(((int64_t)j * wsj + i0 + 4 * q) * 4);
                wxb = (uint32_t)(((int64_t)j * wsj + i0 + 64) * 4);
                p = min(lane, ni - 1 - i0);
                colb = (uint32_t)(((int64_t)j * sj + i0 + p) * 4);
                sc = min(4 * q, ni - 4 - i0);
                sb = (uint32_t)(((int64_t)j * sj + i0 + sc) * 4);
            };
            auto to_lds = [](rsrc_t r, lfloat *lds, uint32_t voff) {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)lds, 16, voff, 0,
                    0, CPOL_NT);
            };
            // group G's requests into slot G % NS; u_pos to the park (in_sweep) or to up_r
            auto issue = [&](int G, bool in_sweep) {
                const int s = G % NS;
                const uint32_t lB = (uint32_t)min(4 * G + g, top), lA = (uint32_t)min(4 * G + 1 + g, top);
                fence();
                wx_r[s] = bload<float, CPOL_NT>(r_wc, wxb + lA * wskb, 0);
                if (in_sweep)
                    to_lds(rs_up, park[G], lb + lB * skb);
                else
                    up_r[s] = bload4<CPOL_NT>(rs_up, lb + lB * skb, 0);
                to_lds(r_uts, slot[s], lb + lB * skb);
                to_lds(r_ut, slot[s] + 256, lb + lB * skb);
                to_lds(r_us, slot[s] + 512, lb + lA * skb);
                to_lds(r_wc, slot[s] + 768, wlb + lA * wskb);
                fence();
            };
            // waits until group G's requests are complete, given that groups G+1..last were issued after it
            auto wait_group = [&](int G, int last) {
                const int n = PER_GROUP * (last - G);
                fence();
                // vmcnt(n): bits 3:0 and 15:14; expcnt 7 and lgkmcnt 15 leave those counters alone
                switch (n) {
#define GT355_VMWAIT(N)                                                                                          \
    case N:                                                                                                      \
        __builtin_amdgcn_s_waitcnt(((N) & 15) | (7 << 4) | (15 << 8) | (((N) >> 4) << 14));                     \
        break;
                    GT355_VMWAIT(0)
                    GT355_VMWAIT(6)
                    GT355_VMWAIT(12)
                    GT355_VMWAIT(18)
                    GT355_VMWAIT(24)
                    GT355_VMWAIT(30)
                    GT355_VMWAIT(36)
                    GT355_VMWAIT(42)
#undef GT355_VMWAIT
                default:
                    __builtin_amdgcn_s_waitcnt(0 | (7 << 4) | (15 << 8));
                }
                fence();
            };
            auto lgkm_wait = [&]() {
                fence();
                __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (0 << 8) | (3 << 14)); // lgkmcnt(0)
                fence();
            };
            float us0 = 0.f;
            auto prologue = [&]() {
                us0 = bload<float, CPOL_NT>(r_us, colb, 0);
#pragma clang loop unroll(full)
                for (int G = 0; G < NS; ++G)
                    issue(G, false);
            };

            struct vals {
                float uts[4], up[4], ut[4], us[4], wi[4], wip[4];
            };
            auto lds_vals = [&](vals &v, int G) {
                const int s = G % NS;
                const int pn = min(p + 1, 63);
#pragma clang loop unroll(full)
                for (int r = 0; r < 4; ++r) {
                    v.uts[r] = slot[s][r * 64 + p];
                    v.up[r] = park[G][r * 64 + p];
                    v.ut[r] = slot[s][256 + r * 64 + p];
                    v.us[r] = slot[s][512 + r * 64 + p];
                    v.wi[r] = slot[s][768 + r * 64 + p];
                    const float wn = slot[s][768 + r * 64 + pn];
                    const float wx = __builtin_bit_cast(float,
                        __builtin_amdgcn_readlane(__builtin_bit_cast(int, wx_r[s]), 16 * r));
                    v.wip[r] = p == 63 ? wx : wn;
                }
            };

            int it = xcd_remap(blockIdx.x, gridDim.x, nxcd);
            if (it >= nitems)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            offsets(it);
            prologue();
            for (; it < nitems; it += gridDim.x) {
                CT c[NK], d[NK];
                CT us_m = CT(0), us_c = CT(us0), w_c_i = CT(0), w_c_ip = CT(0);
                vals a, b;
                wait_group(0, NS - 1);
                *(lf32x4 *)&park[0][4 * lane] = up_r[0];
                fence();
                lds_vals(a, 0);
#pragma clang loop unroll(full)
                for (int G = 0; G < NG; ++G) {
                    if (G + 1 < NG) {
                        // slot G % NS: group G's values are in `a`; refill it with group G+NS
                        if (G + NS < NG) {
                            lgkm_wait();
                            pin();
                            issue(G + NS, true);
                        }
                        wait_group(G + 1, min(G + NS, NG - 1));
                        if (G + 1 < NS)
                            *(lf32x4 *)&park[G + 1][4 * lane] = up_r[G + 1];
                        fence();
                        lds_vals(b, G + 1);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
                    for (int r = 0; r < 4; ++r) {
                        const int k = 4 * G + r;
                        const CT v_uts = CT(a.uts[r]), v_up = CT(a.up[r]), v_ut = CT(a.ut[r]);
                        // top level: the k+1 wind taken as zero reproduces last_level exactly (see
                        // vadv_onchip_kernel)
                        const bool at_top = k == top;
                        const CT us_p = at_top ? us_c : CT(a.us[r]);
                        const CT w_p_i = at_top ? CT(0) : CT(a.wi[r]);
                        const CT w_p_ip = at_top ? CT(0) : CT(a.wip[r]);
                        if (k == 0) {
                            fwd_out<CT> o = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                            asm volatile("" : "+v"(o.c), "+v"(o.d));
                            c[0] = o.c;
                            d[0] = o.d;
                        } else {
                            const CT cm = c[k > 0 ? k - 1 : 0], dm = d[k > 0 ? k - 1 : 0];
                            fwd_out<CT> o = forward_body<CT>(dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p,
                                v_up, v_ut, v_uts, cm, dm, at_top);
                            asm volatile("" : "+v"(o.c), "+v"(o.d));
                            c[k] = o.c;
                            d[k] = o.d;
                        }
                        us_m = us_c;
                        us_c = us_p;
                        w_c_i = w_p_i;
                        w_c_ip = w_p_ip;
                    }
                    a = b;
                    __builtin_amdgcn_sched_barrier(0);
                }
                // the next strip's first groups, in flight during this strip's backward sweep (every slot's values
                // have been read and used by now)
                const int nx = it + (int)gridDim.x;
                const uint32_t csb = sb;
                const int cp = p, csc = sc;
                if (nx < nitems) {
                    lgkm_wait();
                    offsets(nx);
                    pin();
                    prologue();
                }
                float upa[4], upb[4];
#pragma clang loop unroll(full)
                for (int r = 0; r < 4; ++r)
                    upa[r] = park[NG - 1][r * 64 + cp];
                f32x4 o_prev = {0.f, 0.f, 0.f, 0.f};
                uint32_t so_prev = 0;
                CT data = CT(0);
#pragma clang loop unroll(full)
                for (int G = NG - 1; G >= 0; --G) {
                    if (G > 0) {
#pragma clang loop unroll(full)
                        for (int r = 0; r < 4; ++r)
                            upb[r] = park[G > 0 ? G - 1 : 0][r * 64 + cp];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    // the group's output rows go to its own park rows: every lane's reads of them (upa) precede
                    // these writes in program order
                    lfloat *const s_out = park[G];
#pragma clang loop unroll(full)
                    for (int r = 3; r >= 0; --r) {
                        const int k = 4 * G + r;
                        // u_backward_function (vertical_advection_dycore.cpp:110-121)
                        const CT dt = d[k] - c[k] * data;
                        const CT nd = k == top ? d[k] : dt;
                        data = nd;
                        s_out[r * 64 + lane] = float(dtr * (nd - CT(upa[r])));
                    }
                    asm volatile("" ::: "memory");
                    const int L = min(4 * G + g, top), rg = max(L - 4 * G, 0);
                    const f32x4 o = *(const lf32x4 *)&s_out[rg * 64 + csc];
                    asm volatile("" ::: "memory");
                    if (G < NG - 1) {
                        pin();
                        bstore4<CPOL_OUT>(r_uts, o_prev, so_prev, 0);
                    }
                    o_prev = o;
                    so_prev = csb + (uint32_t)L * skb;
#pragma clang loop unroll(full)
                    for (int r = 0; r < 4; ++r)
                        upa[r] = upb[r];
                    __builtin_amdgcn_sched_barrier(0);
                }
                pin();
                bstore4<CPOL_OUT>(r_uts, o_prev, so_prev, 0);
            }
        }

        // Largest linear element offset of a field's allocation box (lower/upper are the box around the origin).
        inline int64_t alloc_last(const gt355_field *f) {
            return (int64_t)(f->upper[0] - 1) + (int64_t)(f->upper[1] - 1) * f->stride[1] +
                   (int64_t)(f->upper[2] - 1) * f->stride[2];
        }
        // Whether vadv_wide_kernel may run on these FP32 fields (its requirements, see the kernel).
        inline bool wide_fits(const gt355_field *uts, const gt355_field *us, const gt355_field *wc,
            const gt355_field *up, const gt355_field *ut, int ni, int nj, int nk) {
            if (ni % 4 != 0)
                return false;
            const int64_t sj = uts->stride[1], sk = uts->stride[2], wsj = wc->stride[1], wsk = wc->stride[2];
            if (sj % 4 || sk % 4 || wsj % 4 || wsk % 4)
                return false;
            const int64_t end_i = (int64_t)((ni + 63) / 64) * 64; // rows are read to the end of the last strip
            const int64_t last = (int64_t)(nk - 1) * sk + (int64_t)(nj - 1) * sj + end_i - 1;
            const int64_t wlast = (int64_t)(nk - 1) * wsk + (int64_t)(nj - 1) * wsj + end_i; // + column 64
            for (const gt355_field *f : {uts, us, up, ut})
                if (((uintptr_t)f->origin & 15) || last > alloc_last(f))
                    return false;
            if (((uintptr_t)wc->origin & 15) || wlast > alloc_last(wc))
                return false;
            return std::max(last, wlast) * 4 < (int64_t(1) << 32);
        }

        // Largest k-depth whose ccol/dcol fit the register file next to the prefetch rings without private-memory
        // spills (FP64 NK=80: 255 VGPRs + 160 AGPRs; deeper buckets spill).
        template <class CT>
        constexpr int onchip_max_nk() {
            return 80;
        }

        // ext_buf == nullptr: the deep-column path takes its ccol/dcol buffer from the per-stream scratch pool.
        // Otherwise ext_buf (ext_bytes long) is used and nothing is allocated: the rows are processed in chunks whose
        // ccol/dcol fit it (columns are independent, so a chunked launch computes the same bits).
        template <class ST, class CT>
        int launch_vadv(const gt355_field *uts, const gt355_field *us, const gt355_field *wc, const gt355_field *up,
            const gt355_field *ut, double dtr_stage, int ni, int nj, int nk, hipStream_t stream,
            void *ext_buf = nullptr, size_t ext_bytes = 0) {
            int64_t sj = uts->stride[1], sk = uts->stride[2];
            GT355_CHECK(us->stride[1] == sj && us->stride[2] == sk && up->stride[1] == sj && up->stride[2] == sk &&
                            ut->stride[1] == sj && ut->stride[2] == sk,
                "gt355_vadv: utens_stage, u_stage, u_pos and utens must share one layout");
            // the reference keeps dtr_stage as a global_parameter of the storage type (vertical_advection_dycore.cpp:139)
            const CT dtr_c = CT(ST(dtr_stage));
            // the on-chip kernel addresses fields through buffer descriptors with 32-bit byte offsets
            const int64_t span = std::max(((int64_t)(nk - 1) * sk + (int64_t)(nj - 1) * sj + ni) * (int64_t)sizeof(ST),
                ((int64_t)(nk - 1) * wc->stride[2] + (int64_t)(nj - 1) * wc->stride[1] + ni + 1) * (int64_t)sizeof(ST));
            const bool fits32 = span < (int64_t(1) << 32);
            if (nk <= onchip_max_nk<CT>() && fits32) {
                GT355_CHECK(nj <= 65535, "gt355_vadv: nj too large");
                dim3 grid((ni + 63) / 64, nj);
#define GT355_VADV_ONCHIP(NKB)                                                                                   \
    if (nk <= NKB) {                                                                                             \
        if (upos_lds)                                                                                            \
            vadv_onchip_kernel<ST, CT, NKB, (NKB < 8 ? NKB : 8), true><<<grid, 64, 0, stream>>>(                  \
                (ST *)uts->origin, (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin,       \
                (const ST *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2]);                    \
        else                                                                                                     \
            vadv_onchip_kernel<ST, CT, NKB, (NKB < 8 ? NKB : 8), false><<<grid, 64, 0, stream>>>(                 \
                (ST *)uts->origin, (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin,       \
                (const ST *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2]);                    \
        return check_launch("gt355_vadv");                                                                       \
    }
                // FP32 storage, 65..80 levels, fewer strips than two rounds of one wavefront per SIMD: the
                // 16-byte-request form, by default with its rows loaded straight into LDS (GT355_TUNE_vadv_wide=0
                // leaves the form; vadv_wide_lds=0 takes the register-ring kernel, whose groups of four levels in flight
                // vadv_wide_dg sets, 3 to 5).  Measured at 256x256x80 (one round), FP64 compute: LDS slots 22.7 us,
                // register ring 22.9 us, the one-row kernels below 23.4 us; with more rounds the persistent form is
                // faster (1536x1536x80 timestep 1.43 against 1.49 ms)
                const bool persistent_shape = (int64_t)((ni + 63) / 64) * nj >= 2 * (int64_t)cu_count() * 4;
                if constexpr (sizeof(ST) == 4) {
                    const int wide_mode = tune_param("vadv_wide", 1); // 2: the wide form for any number of rounds
                    if (nk > 64 && (!persistent_shape || wide_mode == 2) && wide_mode != 0 &&
                        wide_fits(uts, us, wc, up, ut, ni, nj, nk)) {
                        const int strips_i = (ni + 63) / 64;
                        const int64_t items = (int64_t)strips_i * nj;
                        GT355_CHECK(items < (int64_t(1) << 31), "gt355_vadv: too many columns");
                        const int waves = (int)std::min<int64_t>(items, (int64_t)cu_count() * 4);
                        // FP64 compute fits three groups in flight without spills, FP32 compute four
                        const int dg = tune_param("vadv_wide_dg", sizeof(CT) == 8 ? 3 : 4);
#define GT355_VADV_WIDE(DGV)                                                                                     \
    if (dg == DGV) {                                                                                             \
        vadv_wide_kernel<CT, 80, DGV><<<waves, 64, 0, stream>>>((float *)uts->origin, (const float *)us->origin, \
            (const float *)wc->origin, (const float *)up->origin, (const float *)ut->origin, dtr_c, ni, nk, sj, sk, \
            wc->stride[1], wc->stride[2], strips_i, (int)items, xcd_count());                                    \
        return check_launch("gt355_vadv");                                                                       \
    }
                        if (tune_param("vadv_wide_lds", 1) != 0) {
                            // rows loaded straight into LDS: five slots, four groups in flight (write-back instead of
                            // non-temporal output stores measured 24.8-25.0 against 22.7 us)
                            vadv_lds_kernel<CT, 80, 5><<<waves, 64, 0, stream>>>((float *)uts->origin,
                                (const float *)us->origin, (const float *)wc->origin, (const float *)up->origin,
                                (const float *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2], strips_i,
                                (int)items, xcd_count());
                            return check_launch("gt355_vadv");
                        }
                        GT355_VADV_WIDE(3)
                        GT355_VADV_WIDE(5)
                        GT355_VADV_WIDE(4)
#undef GT355_VADV_WIDE
                        GT355_CHECK(false, "gt355_vadv: GT355_TUNE_vadv_wide_dg = %d (3, 4 or 5)", dg);
                    }
                }
                // more strips than two rounds of one wavefront per SIMD: the persistent form (GT355_TUNE_vadv_persist=0
                // selects the one-strip-per-wavefront grid)
                if (nk > 64 && tune_param("vadv_upos_lds", 1) != 0 && tune_param("vadv_persist", 1) != 0) {
                    const int strips_i = (ni + 63) / 64;
                    const int64_t items = (int64_t)strips_i * nj;
                    const int waves = cu_count() * 4;
                    if (items >= 2 * (int64_t)waves && items < (int64_t(1) << 31)) {
                        if constexpr (sizeof(ST) == 4) {
                            // FP32 storage: load-ahead depth 16 from 8 rounds of strips per wavefront on
                            // (GT355_TUNE_vadv_persist_d = 8 / 16 forces one).  Measured: 1536x1536x80 1005 -> 964 us
                            // (36 rounds); 512x512x80 (4 rounds) 92.0 us at depth 8 against 94.4 us at 16.
                            const int pd = tune_param("vadv_persist_d", 0);
                            if (pd == 16 || (pd == 0 && items >= 8 * (int64_t)waves)) {
                                vadv_persist_kernel<ST, CT, 80, 16><<<waves, 64, 0, stream>>>((ST *)uts->origin,
                                    (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin,
                                    (const ST *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2],
                                    strips_i, (int)items, xcd_count());
                                return check_launch("gt355_vadv");
                            }
                        }
                        vadv_persist_kernel<ST, CT, 80, 8><<<waves, 64, 0, stream>>>((ST *)uts->origin,
                            (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin,
                            (const ST *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2], strips_i,
                            (int)items, xcd_count());
                        return check_launch("gt355_vadv");
                    }
                }
                // u_pos parked in LDS costs NK*64*sizeof(ST) bytes per wavefront (40 KB for FP64 at NK=80: four
                // wavefronts still fill a CU's 160 KB) and saves the backward re-read; measured faster for FP64 and
                // FP32 storage at 256x256x80 (54.5 vs 56.3 us FP64).
                const bool upos_lds = tune_param("vadv_upos_lds", 1) != 0;
                // deepest bucket: the load-ahead depth D is selectable (GT355_TUNE_vadv_d; 8 takes the generic bucket
                // below).  16 by default: measured 0.5 us faster than 8 at 256x256x80 for both FP64 (40.45 vs 40.96 us)
                // and FP32 storage (23.8 vs 24.3 us)
                if (nk > 64 && upos_lds) {
                    const int d = tune_param("vadv_d", 16);
#define GT355_VADV_D(DD)                                                                                         \
    if (d == DD) {                                                                                               \
        vadv_onchip_kernel<ST, CT, 80, DD, true><<<grid, 64, 0, stream>>>((ST *)uts->origin,                     \
            (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin, (const ST *)ut->origin, dtr_c, \
            ni, nk, sj, sk, wc->stride[1], wc->stride[2]);                                                       \
        return check_launch("gt355_vadv");                                                                       \
    }
                    GT355_VADV_D(4)
                    GT355_VADV_D(12)
                    GT355_VADV_D(16)
#undef GT355_VADV_D
                    if (tune_param("vadv_sbm", 0) == 6) {
                        vadv_onchip_kernel<ST, CT, 80, 8, true, 6><<<grid, 64, 0, stream>>>((ST *)uts->origin,
                            (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin,
                            (const ST *)ut->origin, dtr_c, ni, nk, sj, sk, wc->stride[1], wc->stride[2]);
                        return check_launch("gt355_vadv");
                    }
                }
                GT355_VADV_ONCHIP(16)
                GT355_VADV_ONCHIP(32)
                GT355_VADV_ONCHIP(48)
                GT355_VADV_ONCHIP(64)
                GT355_VADV_ONCHIP(80)
#undef GT355_VADV_ONCHIP
            }
            // ccol/dcol of levels 0..nk-2 for every column of a row
            const size_t row_bytes = sizeof(CT) * (size_t)ni * (size_t)(2 * nk - 1);
            int rows = nj;
            void *buf = ext_buf;
            if (ext_buf) {
                GT355_CHECK(row_bytes <= ext_bytes,
                    "gt355_vadv_noalloc: ni = %d, nk = %d needs %zu bytes of ccol/dcol for one row of columns, more "
                    "than the %zu-byte static scratch holds",
                    ni, nk, row_bytes, ext_bytes);
                rows = (int)std::min<size_t>((size_t)nj, ext_bytes / row_bytes);
            } else {
                int rc = scratch_for(stream, row_bytes * (size_t)nj, &buf);
                if (rc)
                    return rc;
            }
            GT355_CHECK((std::min(rows, nj) + 3) / 4 <= 65535, "gt355_vadv: nj too large");
            for (int j0 = 0; j0 < nj; j0 += rows) {
                const int nr = std::min(rows, nj - j0);
                dim3 grid((ni + 63) / 64, (nr + 3) / 4);
                vadv_scratch_kernel<ST, CT><<<grid, 256, 0, stream>>>((ST *)uts->origin + j0 * sj,
                    (const ST *)us->origin + j0 * sj, (const ST *)wc->origin + j0 * wc->stride[1],
                    (const ST *)up->origin + j0 * sj, (const ST *)ut->origin + j0 * sj, dtr_c, ni, nr, nk, sj, sk,
                    wc->stride[1], wc->stride[2], (CT *)buf, (int64_t)ni * nr);
                int rc = check_launch("gt355_vadv");
                if (rc)
                    return rc;
            }
            return GT355_OK;
        }
        // The no-allocation entry point's ccol/dcol buffer: a file-scope device array, one copy per device for the
        // life of the process.  Host code never uses its name as an address (on the host it names a stand-in): the
        // device address comes from hipGetSymbolAddress, once per device, with that device current.
        constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
        static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

        constexpr int NOALLOC_MAX_DEVICES = 64;
        std::mutex g_noalloc_mutex;
        void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

        // The current device's address of `scratch`.
        int noalloc_scratch(void **out) {
            int dev = 0;
            GT355_HIP(hipGetDevice(&dev));
            GT355_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "gt355_vadv_noalloc: device %d", dev);
            std::lock_guard<std::mutex> lock(g_noalloc_mutex);
            if (!g_noalloc_ptr[dev]) {
                void *p = nullptr;
                hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                if (e != hipSuccess)
                    return hip_fail(e, "gt355_vadv_noalloc: hipGetSymbolAddress(scratch)");
                GT355_CHECK(p != nullptr, "gt355_vadv_noalloc: no device address for the static scratch");
                g_noalloc_ptr[dev] = p;
            }
            *out = g_noalloc_ptr[dev];
            return GT355_OK;
        }
    } // namespace

    // Called by gt355_init with the selected device current: resolves the static scratch's address there.
    int noalloc_setup() {
        void *p = nullptr;
        return noalloc_scratch(&p);
    }
} // namespace gt355

using namespace gt355;

namespace {
    int vadv_entry(int precision, const gt355_field *utens_stage, const gt355_field *u_stage, const gt355_field *wcon,
        const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj, int nk, void *stream,
        bool noalloc) {
    GT355_CHECK(ni >= 0 && nj >= 0, "gt355_vadv: negative size");
    GT355_CHECK(nk >= 2, "gt355_vadv: nk = %d, the spec needs at least 2 levels (first and last level)", nk);
    GT355_CHECK(utens_stage && u_stage && wcon && u_pos && utens, "gt355_vadv: null field");
    if (ni == 0 || nj == 0)
        return GT355_OK;
    int dtype = utens_stage->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_vadv: dtype %d", dtype);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dtype == GT355_F32),
        "gt355_vadv: precision %d with dtype %d", precision, dtype);
    int rc;
    if ((rc = check_field(utens_stage, "utens_stage", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    if ((rc = check_field(u_stage, "u_stage", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    if ((rc = check_field(wcon, "wcon", dtype, 0, ni + 1, 0, nj, 0, nk)))
        return rc;
    if ((rc = check_field(u_pos, "u_pos", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    if ((rc = check_field(utens, "utens", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    // inside a deferred-run scope gt355_vadv is recorded and runs at the scope's flush; the no-allocation entry
    // point always runs at once (the flush may build a batch), after what the scope recorded before it
    if (!noalloc && defer_active(stream))
        return defer_vadv(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream);
    if (noalloc && (rc = defer_flush(stream)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    void *buf = nullptr;
    const size_t bytes = noalloc ? NOALLOC_SCRATCH_BYTES : 0;
    if (noalloc && (rc = noalloc_scratch(&buf)))
        return rc;
    if (dtype == GT355_F64)
        return launch_vadv<double, double>(
            utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s, buf, bytes);
    if (precision == GT355_WIDEN)
        return launch_vadv<float, double>(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s, buf, bytes);
    return launch_vadv<float, float>(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s, buf, bytes);
    }
} // namespace

extern "C" int gt355_vadv(int precision, const gt355_field *utens_stage, const gt355_field *u_stage,
    const gt355_field *wcon, const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj,
    int nk, void *stream) {
    return vadv_entry(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream, false);
}

extern "C" int gt355_vadv_noalloc(int precision, const gt355_field *utens_stage, const gt355_field *u_stage,
    const gt355_field *wcon, const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj,
    int nk, void *stream) {
    return vadv_entry(precision, utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, stream, true);
}
