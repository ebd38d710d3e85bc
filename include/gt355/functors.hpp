// The reference functor bodies the hand-written kernels evaluate, operation for operation (the library is built with
// -ffp-contract=off, so each line below is a sequence of correctly rounded IEEE operations in this order).  Shared by
// the library's kernels and by the generic engine, which checks a user spec against them on the host before it hands
// the spec to a hand-written kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace gt355 {

    // Reference functor bodies (horizontal_diffusion.cpp:41-46, :56-60, :70-74, :86-90).
    template <class CT>
    __host__ __device__ __forceinline__ CT lap_f(CT c, CT ip, CT jp, CT im, CT jm) {
        return CT(4) * c - (((ip + jp) + im) + jm);
    }
    template <class CT>
    __host__ __device__ __forceinline__ CT flux_f(CT lap_p, CT lap_c, CT in_p, CT in_c) {
        CT res = lap_p - lap_c;
        return (res * (in_p - in_c) > CT(0)) ? CT(0) : res;
    }
    // out_function (horizontal_diffusion.cpp:86-90)
    template <class CT>
    __host__ __device__ __forceinline__ CT hdiff_out_f(CT c, CT co, CT flx_c, CT flx_m, CT fly_c, CT fly_m) {
        return c - co * (((flx_c - flx_m) + fly_c) - fly_m);
    }

    // wlap_function::apply (simple_hori_diff.cpp:33-39): C++ sums a + b - c + d + e left to right.
    template <class CT>
    __host__ __device__ __forceinline__ CT wlap_f(CT c, CT ip, CT im, CT jp, CT jm, CT cro, CT cru) {
        return (((ip + im) - CT(2) * c) + cro * (jp - c)) + cru * (jm - c);
    }
    // divflux_function::apply (simple_hori_diff.cpp:51-60); l* are the laplacians at the centre and neighbours.
    template <class CT>
    __host__ __device__ __forceinline__ CT divflux_f(CT c, CT l, CT lip, CT lim, CT ljp, CT ljm, CT cro, CT co) {
        CT fluxx = lip - l;
        CT fluxx_m = l - lim;
        CT fluxy = cro * (ljp - l);
        CT fluxy_m = cro * (l - ljm);
        return c + ((fluxx_m - fluxx) + (fluxy_m - fluxy)) * co;
    }

    constexpr double BET_M = 0.5, BET_P = 0.5; // vertical_advection_defs.hpp:14-16

    template <class CT>
    struct fwd_out {
        CT c, d;
    };

    // u_forward_function::apply(first_level) (vertical_advection_dycore.cpp:83-96)
    template <class CT>
    __host__ __device__ __forceinline__ fwd_out<CT> forward_first(
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

    // u_forward_function::apply(modify<1,-1>) (vertical_advection_dycore.cpp:50-67).
    // top: evaluate apply(last_level) (:69-81) instead, for callers that pass zero k+1 winds at the top level: then
    // c = +0, b = (dtr - a) - (+0) == dtr - a, and the k+1 correction term is replaced by +0 (x - (+0) == x for every
    // x), so the result has last_level's bits whatever u_stage holds (an infinite u_stage would make cs * (us_p -
    // us_c) a NaN).
    template <class CT>
    __host__ __device__ __forceinline__ fwd_out<CT> forward_body(CT dtr, CT w_c_i, CT w_c_ip, CT w_p_i, CT w_p_ip,
        CT us_m, CT us_c, CT us_p, CT up, CT ut, CT uts, CT cc_m, CT dc_m, bool top) {
        CT gav = -CT(.25) * (w_c_ip + w_c_i);
        CT gcv = CT(.25) * (w_p_ip + w_p_i);
        CT as = gav * CT(BET_M);
        CT cs = gcv * CT(BET_M);
        CT a = gav * CT(BET_P);
        CT c = gcv * CT(BET_P);
        CT b = (dtr - a) - c;
        CT up_term = cs * (us_p - us_c);
        CT correction = -as * (us_m - us_c) - (top ? CT(0) : up_term);
        CT d = ((dtr * up + ut) + uts) + correction;
        CT divided = CT(1) / (b - cc_m * a);
        return {c * divided, (d - dc_m * a) * divided};
    }

    // u_forward_function::apply(last_level) (vertical_advection_dycore.cpp:69-81); ccol is not written
    template <class CT>
    __host__ __device__ __forceinline__ CT forward_last(
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

    // u_backward_function::apply (vertical_advection_dycore.cpp:110-121): data = dcol - ccol * data(k+1), or dcol at
    // the top level; the output is dtr * (data - u_pos)
    template <class CT>
    __host__ __device__ __forceinline__ CT backward_data(CT cc, CT dc, CT data_above, bool top) {
        const CT dt = dc - cc * data_above;
        return top ? dc : dt;
    }

} // namespace gt355
