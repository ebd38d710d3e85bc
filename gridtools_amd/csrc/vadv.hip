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

        template <class ST, class CT>
        int launch_vadv(const gt355_field *uts, const gt355_field *us, const gt355_field *wc, const gt355_field *up,
            const gt355_field *ut, double dtr_stage, int ni, int nj, int nk, hipStream_t stream) {
            int64_t sj = uts->stride[1], sk = uts->stride[2];
            GT355_CHECK(us->stride[1] == sj && us->stride[2] == sk && up->stride[1] == sj && up->stride[2] == sk &&
                            ut->stride[1] == sj && ut->stride[2] == sk,
                "gt355_vadv: utens_stage, u_stage, u_pos and utens must share one layout");
            int64_t plane = (int64_t)ni * nj;
            size_t bytes = sizeof(CT) * (size_t)plane * (size_t)(2 * nk - 1);
            void *scratch = nullptr;
            int rc = get_scratch(bytes, &scratch);
            if (rc)
                return rc;
            // the reference keeps dtr_stage as a global_parameter of the storage type (vertical_advection_dycore.cpp:139)
            CT dtr = CT(ST(dtr_stage));
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
