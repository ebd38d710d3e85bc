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

#include <mutex>

#include "common.hpp"

namespace gt355 {
    namespace {

        constexpr double BET_M = 0.5, BET_P = 0.5; // vertical_advection_defs.hpp:14-16

        template <class CT>
        struct fwd_out {
            CT c, d;
        };

        // u_forward_function::apply(first_level) (vertical_advection_dycore.cpp:83-96)
        template <class CT>
        __device__ __forceinline__ fwd_out<CT> forward_first(
            CT dtr, CT w_p_i, CT w_p_ip, CT us_c, CT us_p, CT up, CT ut, CT uts) {
            CT gcv = CT(.25) * (w_p_ip + w_p_i);
            CT cs = gcv * CT(BET_M);
            CT c = gcv * CT(BET_P);
            CT b = dtr - c;
            CT correction = -cs * (us_p - us_c);
            CT d = ((dtr * up + ut) + uts) + correction;
            CT divided = CT(1) / b;
            return {c * divided, d * divided};
        }

        // u_forward_function::apply(modify<1,-1>) (vertical_advection_dycore.cpp:50-67)
        template <class CT>
        __device__ __forceinline__ fwd_out<CT> forward_body(CT dtr, CT w_c_i, CT w_c_ip, CT w_p_i, CT w_p_ip,
            CT us_m, CT us_c, CT us_p, CT up, CT ut, CT uts, CT cc_m, CT dc_m) {
            CT gav = -CT(.25) * (w_c_ip + w_c_i);
            CT gcv = CT(.25) * (w_p_ip + w_p_i);
            CT as = gav * CT(BET_M);
            CT cs = gcv * CT(BET_M);
            CT a = gav * CT(BET_P);
            CT c = gcv * CT(BET_P);
            CT b = (dtr - a) - c;
            CT correction = -as * (us_m - us_c) - cs * (us_p - us_c);
            CT d = ((dtr * up + ut) + uts) + correction;
            CT divided = CT(1) / (b - cc_m * a);
            return {c * divided, (d - dc_m * a) * divided};
        }

        // u_forward_function::apply(last_level) (vertical_advection_dycore.cpp:69-81); ccol is not written
        template <class CT>
        __device__ __forceinline__ CT forward_last(
            CT dtr, CT w_c_i, CT w_c_ip, CT us_m, CT us_c, CT up, CT ut, CT uts, CT cc_m, CT dc_m) {
            CT gav = -CT(.25) * (w_c_ip + w_c_i);
            CT as = gav * CT(BET_M);
            CT a = gav * CT(BET_P);
            CT b = dtr - a;
            CT correction = -as * (us_m - us_c);
            CT d = ((dtr * up + ut) + uts) + correction;
            CT divided = CT(1) / (b - cc_m * a);
            return (d - dc_m * a) * divided;
        }

        // One column per lane; ccol/dcol of levels 0..nk-2 go to scratch (k-major, column-minor).
        template <class ST, class CT>
        __global__ void __launch_bounds__(256) vadv_scratch_kernel(ST *__restrict__ uts_f,
            const ST *__restrict__ us_f, const ST *__restrict__ wc_f, const ST *__restrict__ up_f,
            const ST *__restrict__ ut_f, CT dtr, int ni, int nj, int nk, int64_t sj, int64_t sk, int64_t wsj,
            int64_t wsk, CT *__restrict__ scratch, int64_t plane) {
            const int i = blockIdx.x * 64 + (threadIdx.x & 63);
            const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
            if (i >= ni || j >= nj)
                return;
            const int64_t col = (int64_t)j * sj + i;
            const int64_t wcol = (int64_t)j * wsj + i;
            const int64_t scol = (int64_t)j * ni + i; // scratch column index, plane = ni*nj
            CT *__restrict__ sc = scratch + scol;
            CT *__restrict__ sd = scratch + (int64_t)(nk - 1) * plane + scol;

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
                    CT(uts_f[p]), cc_m, dc_m);
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
        // unified VGPR/AGPR file: at one wave per SIMD a lane owns 512 registers), so the kernel moves only the
        // compulsory bytes.  With one wave per SIMD nothing else hides memory latency, so every input is
        // software-pipelined D levels ahead through register rings:
        //   ring A (u_stage, wcon at i and i+1) of level L is consumed at forward step L-1,
        //   ring B (utens_stage, u_pos, utens) of level L at forward step L,
        //   ring C (u_pos again) of level L at backward step L.
        // Loops are fully unrolled over NK with uniform guards on the runtime nk, so every ring slot and every
        // ccol/dcol element has a static register.  Lanes past ni read the last valid column and store nothing.
        template <class ST, class CT, int NK, int D>
        __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
        vadv_onchip_kernel(ST *__restrict__ uts_f, const ST *__restrict__ us_f, const ST *__restrict__ wc_f,
            const ST *__restrict__ up_f, const ST *__restrict__ ut_f, CT dtr, int ni, int nk, int64_t sj, int64_t sk,
            int64_t wsj, int64_t wsk) {
            static_assert(D >= 1 && D <= NK, "prefetch depth");
            const int i_raw = blockIdx.x * 64 + threadIdx.x;
            const bool store_ok = i_raw < ni;
            const int i = store_ok ? i_raw : ni - 1;
            const int j = blockIdx.y;
            const int64_t col = (int64_t)j * sj + i;
            const int64_t wcol = (int64_t)j * wsj + i;
            const ST *us = us_f + col;
            const ST *up = up_f + col;
            const ST *ut = ut_f + col;
            ST *uts = uts_f + col;
            const ST *wc = wc_f + wcol;

            ST a_us[D], a_wi[D], a_wip[D]; // ring A
            ST b_uts[D], b_up[D], b_ut[D]; // ring B
            auto load_a = [&](int L) {
                a_us[L % D] = us[L * sk];
                a_wi[L % D] = wc[L * wsk];
                a_wip[L % D] = wc[L * wsk + 1];
            };
            auto load_b = [&](int L) {
                b_uts[L % D] = uts[L * sk];
                b_up[L % D] = up[L * sk];
                b_ut[L % D] = ut[L * sk];
            };
#pragma unroll
            for (int L = 0; L < D; ++L) {
                if (L < nk)
                    load_b(L);
                if (L + 1 < nk)
                    load_a(L + 1);
            }

            CT c[NK], d[NK];
            CT us_m = CT(0), us_c = CT(us[0]), w_c_i = CT(0), w_c_ip = CT(0);
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                if (k < nk) {
                    const CT v_uts = CT(b_uts[k % D]), v_up = CT(b_up[k % D]), v_ut = CT(b_ut[k % D]);
                    CT us_p = CT(0), w_p_i = CT(0), w_p_ip = CT(0);
                    if (k + 1 < nk) {
                        us_p = CT(a_us[(k + 1) % D]);
                        w_p_i = CT(a_wi[(k + 1) % D]);
                        w_p_ip = CT(a_wip[(k + 1) % D]);
                    }
                    // refill the slots just consumed
                    if (k + D < NK && k + D < nk)
                        load_b(k + D);
                    if (k + 1 + D < NK && k + 1 + D < nk)
                        load_a(k + 1 + D);
                    if (k == 0) {
                        fwd_out<CT> r = forward_first<CT>(dtr, w_p_i, w_p_ip, us_c, us_p, v_up, v_ut, v_uts);
                        c[0] = r.c;
                        d[0] = r.d;
                    } else if (k < nk - 1) {
                        fwd_out<CT> r = forward_body<CT>(dtr, w_c_i, w_c_ip, w_p_i, w_p_ip, us_m, us_c, us_p, v_up,
                            v_ut, v_uts, c[k > 0 ? k - 1 : 0], d[k > 0 ? k - 1 : 0]);
                        c[k] = r.c;
                        d[k] = r.d;
                    } else {
                        d[k] = forward_last<CT>(dtr, w_c_i, w_c_ip, us_m, us_c, v_up, v_ut, v_uts,
                            c[k > 0 ? k - 1 : 0], d[k > 0 ? k - 1 : 0]);
                    }
                    us_m = us_c;
                    us_c = us_p;
                    w_c_i = w_p_i;
                    w_c_ip = w_p_ip;
                }
            }

            // backward sweep: u_pos of level L is prefetched at backward step L + D (or up front for the top D
            // levels, which no active step precedes)
            ST r_up[D];
#pragma unroll
            for (int L = 0; L < NK; ++L)
                if (L < nk && L + D >= nk)
                    r_up[L % D] = up[L * sk];
            CT data = CT(0);
#pragma unroll
            for (int k = NK - 1; k >= 0; --k) {
                if (k < nk) {
                    const CT v_up = CT(r_up[k % D]);
                    if (k - D >= 0)
                        r_up[k % D] = up[(k - D) * sk];
                    CT out;
                    if (k == nk - 1) {
                        // u_backward_function::apply(last_level) (vertical_advection_dycore.cpp:117-121)
                        out = dtr * (d[k] - v_up);
                        data = d[k];
                    } else {
                        // u_backward_function::apply(modify<0,-1>) (vertical_advection_dycore.cpp:110-115)
                        CT dt = d[k] - c[k] * data;
                        out = dtr * (dt - v_up);
                        data = dt;
                    }
                    if (store_ok)
                        uts[k * sk] = ST(out);
                }
            }
        }

        // Scratch pool (the role of the reference's cached_allocator, sid/allocator.hpp:46-56): grows, never
        // shrinks, one per device; calls on one device are serialised by the caller's stream order.
        struct scratch_pool {
            std::mutex m;
            void *ptr[64] = {};
            size_t bytes[64] = {};
        };
        scratch_pool g_pool;

        int get_scratch(size_t bytes, void **out) {
            int dev = current_device();
            GT355_CHECK(dev >= 0 && dev < 64, "scratch: device %d", dev);
            std::lock_guard<std::mutex> lock(g_pool.m);
            if (g_pool.bytes[dev] < bytes) {
                if (g_pool.ptr[dev]) {
                    // the old buffer may still be in use by queued work on any stream
                    GT355_HIP(hipDeviceSynchronize());
                    GT355_HIP(hipFree(g_pool.ptr[dev]));
                    g_pool.ptr[dev] = nullptr;
                    g_pool.bytes[dev] = 0;
                }
                void *p = nullptr;
                hipError_t e = hipMalloc(&p, bytes);
                if (e != hipSuccess) {
                    set_error("scratch: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
                    return GT355_ENOMEM;
                }
                g_pool.ptr[dev] = p;
                g_pool.bytes[dev] = bytes;
            }
            *out = g_pool.ptr[dev];
            return GT355_OK;
        }

        // Largest k-depth whose ccol/dcol fit the register file next to the prefetch rings without private-memory
        // spills (FP64 NK=80: 255 VGPRs + 160 AGPRs; deeper buckets spill).
        template <class CT>
        constexpr int onchip_max_nk() {
            return 80;
        }

        template <class ST, class CT>
        int launch_vadv(const gt355_field *uts, const gt355_field *us, const gt355_field *wc, const gt355_field *up,
            const gt355_field *ut, double dtr_stage, int ni, int nj, int nk, hipStream_t stream) {
            int64_t sj = uts->stride[1], sk = uts->stride[2];
            GT355_CHECK(us->stride[1] == sj && us->stride[2] == sk && up->stride[1] == sj && up->stride[2] == sk &&
                            ut->stride[1] == sj && ut->stride[2] == sk,
                "gt355_vadv: utens_stage, u_stage, u_pos and utens must share one layout");
            // the reference keeps dtr_stage as a global_parameter of the storage type (vertical_advection_dycore.cpp:139)
            const CT dtr_c = CT(ST(dtr_stage));
            if (nk <= onchip_max_nk<CT>()) {
                GT355_CHECK(nj <= 65535, "gt355_vadv: nj too large");
                dim3 grid((ni + 63) / 64, nj);
#define GT355_VADV_ONCHIP(NKB)                                                                                   \
    if (nk <= NKB) {                                                                                             \
        vadv_onchip_kernel<ST, CT, NKB, (NKB < 8 ? NKB : 8)><<<grid, 64, 0, stream>>>((ST *)uts->origin,         \
            (const ST *)us->origin, (const ST *)wc->origin, (const ST *)up->origin, (const ST *)ut->origin, dtr_c, \
            ni, nk, sj, sk, wc->stride[1], wc->stride[2]);                                                       \
        return check_launch("gt355_vadv");                                                                       \
    }
                GT355_VADV_ONCHIP(16)
                GT355_VADV_ONCHIP(32)
                GT355_VADV_ONCHIP(48)
                GT355_VADV_ONCHIP(64)
                GT355_VADV_ONCHIP(80)
#undef GT355_VADV_ONCHIP
            }
            int64_t plane = (int64_t)ni * nj;
            size_t bytes = sizeof(CT) * (size_t)plane * (size_t)(2 * nk - 1);
            void *scratch = nullptr;
            int rc = get_scratch(bytes, &scratch);
            if (rc)
                return rc;
            CT dtr = dtr_c;
            dim3 grid((ni + 63) / 64, (nj + 3) / 4);
            GT355_CHECK(grid.y <= 65535, "gt355_vadv: nj too large");
            vadv_scratch_kernel<ST, CT><<<grid, 256, 0, stream>>>((ST *)uts->origin, (const ST *)us->origin,
                (const ST *)wc->origin, (const ST *)up->origin, (const ST *)ut->origin, dtr, ni, nj, nk, sj, sk,
                wc->stride[1], wc->stride[2], (CT *)scratch, plane);
            return check_launch("gt355_vadv");
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_vadv(int precision, const gt355_field *utens_stage, const gt355_field *u_stage,
    const gt355_field *wcon, const gt355_field *u_pos, const gt355_field *utens, double dtr_stage, int ni, int nj,
    int nk, void *stream) {
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
    hipStream_t s = (hipStream_t)stream;
    if (dtype == GT355_F64)
        return launch_vadv<double, double>(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s);
    if (precision == GT355_WIDEN)
        return launch_vadv<float, double>(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s);
    return launch_vadv<float, float>(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, ni, nj, nk, s);
}
