// The reference functor bodies the kernels evaluate, operation for operation (the library is built with
// -ffp-contract=off, so each line below is a sequence of correctly rounded IEEE operations in this order).
#pragma once

namespace gt355 {

    // Reference functor bodies (horizontal_diffusion.cpp:41-46, :56-60, :70-74, :86-90).
    template <class CT>
    __device__ __forceinline__ CT lap_f(CT c, CT ip, CT jp, CT im, CT jm) {
        return CT(4) * c - (((ip + jp) + im) + jm);
    }
    template <class CT>
    __device__ __forceinline__ CT flux_f(CT lap_p, CT lap_c, CT in_p, CT in_c) {
        CT res = lap_p - lap_c;
        return (res * (in_p - in_c) > CT(0)) ? CT(0) : res;
    }

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

} // namespace gt355
