/*
 * The GridTools reference's own stencil specs (functors of its regression and unit tests), used unmodified as inputs
 * of the generic engine's drop-in tests: each namespace holds the functors of one reference test file under their
 * reference names, with the file and lines they come from.  The test harnesses around them (storages, grids,
 * expected values) are this project's and live in generic.hip.
 *
 * These functors are taken from GridTools (https://github.com/GridTools/gridtools), which carries this notice:
 *
 *   GridTools
 *
 *   Copyright (c) 2014-2023, ETH Zurich
 *   All rights reserved.
 *
 *   Redistribution and use in source and binary forms, with or without modification, are permitted provided that
 *   the following conditions are met:
 *   1. Redistributions of source code must retain the above copyright notice, this list of conditions and the
 *      following disclaimer.
 *   2. Redistributions in binary form must reproduce the above copyright notice, this list of conditions and the
 *      following disclaimer in the documentation and/or other materials provided with the distribution.
 *   3. Neither the name of the copyright holder nor the names of its contributors may be used to endorse or promote
 *      products derived from this software without specific prior written permission.
 *
 *   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS" AND ANY EXPRESS OR IMPLIED
 *   WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR
 *   PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER OR CONTRIBUTORS BE LIABLE FOR ANY DIRECT,
 *   INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF
 *   SUBSTITUTE GOODS OR SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND ON ANY
 *   THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING
 *   IN ANY WAY OUT OF THE USE OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
 *
 *   SPDX-License-Identifier: BSD-3-Clause
 */
#pragma once

#include <gt355/stencil.hpp>

namespace reference_specs {
    using namespace gt355;
    using namespace gt355::stencil;
    using namespace gt355::stencil::cartesian;

    // tests/regression/copy_stencil.cpp:23-34
    namespace copy_stencil {
        struct copy_functor {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval) {
                eval(out()) = eval(in());
            }
        };
    } // namespace copy_stencil

    // tests/regression/laplacian.cpp:20-29
    namespace laplacian {
        struct lap {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 1>>;
            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) = 4 * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
            }
        };
    } // namespace laplacian

    // tests/regression/horizontal_diffusion.cpp:35-91 (spec at :97-107)
    namespace horizontal_diffusion {
        struct lap_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 1>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                using float_t = std::decay_t<decltype(eval(out()))>;
                eval(out()) =
                    float_t{4} * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
            }
        };

        struct flx_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<0, 1, 0, 0>>;
            using lap = in_accessor<2, extent<0, 1, 0, 0>>;

            using param_list = make_param_list<out, in, lap>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                auto res = eval(lap(1, 0)) - eval(lap(0, 0));
                eval(out()) = res * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0 : res;
            }
        };

        struct fly_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<0, 0, 0, 1>>;
            using lap = in_accessor<2, extent<0, 0, 0, 1>>;

            using param_list = make_param_list<out, in, lap>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                auto res = eval(lap(0, 1)) - eval(lap(0, 0));
                eval(out()) = res * (eval(in(0, 1)) - eval(in(0, 0))) > 0 ? 0 : res;
            }
        };

        struct out_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1>;
            using flx = in_accessor<2, extent<-1, 0, 0, 0>>;
            using fly = in_accessor<3, extent<0, 0, -1, 0>>;
            using coeff = in_accessor<4>;

            using param_list = make_param_list<out, in, flx, fly, coeff>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) =
                    eval(in()) - eval(coeff()) * (eval(flx()) - eval(flx(-1, 0)) + eval(fly()) - eval(fly(0, -1)));
            }
        };

        // the spec of get_spec() (:97-107), for storages of value type Float
        template <class Float>
        auto spec() {
            return [](auto in, auto coeff, auto out) {
                GT_DECLARE_TMP(Float, lap, flx, fly);
                return execute_parallel()
                    .ij_cached(lap, flx, fly)
                    .stage(lap_function(), lap, in)
                    .stage(flx_function(), flx, in, lap)
                    .stage(fly_function(), fly, in, lap)
                    .stage(out_function(), out, in, flx, fly, coeff);
            };
        }
    } // namespace horizontal_diffusion

    // tests/regression/vertical_advection_defs.hpp:14-16 and vertical_advection_dycore.cpp:28-122 (spec at :140-149)
    namespace vertical_advection_dycore {
        constexpr double BET_M = 0.5;
        constexpr double BET_P = 0.5;

        // This is the definition of the special regions in the "vertical" direction
        using axis_t = axis<1, axis_config::offset_limit<3>>;
        using full_t = axis_t::full_interval;

        class u_forward_function {
            using utens_stage = in_accessor<0>;
            using wcon = in_accessor<1, extent<0, 1, 0, 0, 0, 1>>;
            using u_stage = in_accessor<2, extent<0, 0, 0, 0, -1, 1>>;
            using u_pos = in_accessor<3>;
            using utens = in_accessor<4>;
            using dtr_stage = in_accessor<5>;
            using ccol = inout_accessor<6, extent<0, 0, 0, 0, -1, 0>>;
            using dcol = inout_accessor<7, extent<0, 0, 0, 0, -1, 0>>;

            template <class Eval>
            GT_FUNCTION static auto compute_d(Eval &&eval) {
                return eval(dtr_stage()) * eval(u_pos()) + eval(utens()) + eval(utens_stage());
            }

          public:
            using param_list = make_param_list<utens_stage, wcon, u_stage, u_pos, utens, dtr_stage, ccol, dcol>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::modify<1, -1>) {
                using float_t = std::decay_t<decltype(eval(ccol()))>;
                auto gav = -float_t(.25) * (eval(wcon(1, 0, 0)) + eval(wcon(0, 0, 0)));
                auto gcv = float_t(.25) * (eval(wcon(1, 0, 1)) + eval(wcon(0, 0, 1)));
                auto as = gav * float_t(BET_M);
                auto cs = gcv * float_t(BET_M);
                auto a = gav * float_t(BET_P);
                auto c = gcv * float_t(BET_P);
                auto b = eval(dtr_stage()) - a - c;
                auto correction =
                    -as * (eval(u_stage(0, 0, -1)) - eval(u_stage())) - cs * (eval(u_stage(0, 0, 1)) - eval(u_stage()));
                auto d = compute_d(eval) + correction;
                auto divided = float_t(1) / (b - eval(ccol(0, 0, -1)) * a);

                eval(ccol()) = c * divided;
                eval(dcol()) = (d - eval(dcol(0, 0, -1)) * a) * divided;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::last_level) {
                using float_t = std::decay_t<decltype(eval(ccol()))>;
                auto gav = -float_t(.25) * (eval(wcon(1, 0, 0)) + eval(wcon()));
                auto as = gav * float_t(BET_M);
                auto a = gav * float_t(BET_P);
                auto b = eval(dtr_stage()) - a;
                auto correction = -as * (eval(u_stage(0, 0, -1)) - eval(u_stage()));
                auto d = compute_d(eval) + correction;
                auto divided = float_t(1) / (b - eval(ccol(0, 0, -1)) * a);

                eval(dcol()) = (d - eval(dcol(0, 0, -1)) * a) * divided;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::first_level) {
                using float_t = std::decay_t<decltype(eval(ccol()))>;
                auto gcv = float_t(.25) * (eval(wcon(1, 0, 1)) + eval(wcon(0, 0, 1)));
                auto cs = gcv * float_t(BET_M);
                auto c = gcv * float_t(BET_P);
                auto b = eval(dtr_stage()) - c;
                auto correction = -cs * (eval(u_stage(0, 0, 1)) - eval(u_stage()));
                auto d = compute_d(eval) + correction;
                auto divided = float_t(1) / b;

                eval(ccol()) = c * divided;
                eval(dcol()) = d * divided;
            }
        };

        class u_backward_function {
            using utens_stage = inout_accessor<0>;
            using u_pos = in_accessor<1>;
            using dtr_stage = in_accessor<2>;
            using ccol = in_accessor<3>;
            using dcol = in_accessor<4>;
            using data_col = inout_accessor<5, extent<0, 0, 0, 0, 0, 1>>;

          public:
            using param_list = make_param_list<utens_stage, u_pos, dtr_stage, ccol, dcol, data_col>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::modify<0, -1>) {
                auto data = eval(dcol()) - eval(ccol()) * eval(data_col(0, 0, 1));
                eval(utens_stage()) = eval(dtr_stage()) * (data - eval(u_pos()));
                eval(data_col()) = data;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::last_level) {
                eval(utens_stage()) = eval(dtr_stage()) * (eval(dcol()) - eval(u_pos()));
                eval(data_col()) = eval(dcol());
            }
        };

        // the spec of the regression test (:140-149), for storages of value type Float
        template <class Float>
        auto spec() {
            return [](auto utens_stage, auto u_stage, auto wcon, auto u_pos, auto utens, auto dtr_stage) {
                GT_DECLARE_TMP(Float, ccol, dcol, data_col);
                return multi_pass(execute_forward()
                                      .k_cached(cache_io_policy::flush(), ccol, dcol)
                                      .k_cached(cache_io_policy::fill(), u_stage)
                                      .stage(u_forward_function(),
                                          utens_stage,
                                          wcon,
                                          u_stage,
                                          u_pos,
                                          utens,
                                          dtr_stage,
                                          ccol,
                                          dcol),
                    execute_backward().k_cached(data_col).stage(
                        u_backward_function(), utens_stage, u_pos, dtr_stage, ccol, dcol, data_col));
            };
        }
    } // namespace vertical_advection_dycore

    // tests/regression/tridiagonal.cpp:30-75 (spec at :106-110)
    namespace tridiagonal {
        using namespace expressions;

        // This is the definition of the special regions in the "vertical" direction
        using axis_t = axis<1>;
        using full_t = axis_t::full_interval;

        struct forward_thomas {
            using inf = in_accessor<0>;                               // a
            using diag = in_accessor<1>;                              // b
            using sup = inout_accessor<2, extent<0, 0, 0, 0, -1, 0>>; // c
            using rhs = inout_accessor<3, extent<0, 0, 0, 0, -1, 0>>; // d
            using param_list = make_param_list<inf, diag, sup, rhs>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<1, 0>) {
                eval(sup{}) = eval(sup{} / (diag{} - sup{0, 0, -1} * inf{}));
                eval(rhs{}) = eval((rhs{} - inf{} * rhs{0, 0, -1}) / (diag{} - sup{0, 0, -1} * inf{}));
            }

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::first_level) {
                eval(sup{}) = eval(sup{}) / eval(diag{});
                eval(rhs{}) = eval(rhs{}) / eval(diag{});
            }
        };

        struct backward_thomas {
            using out = inout_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
            using sup = in_accessor<1>; // c
            using rhs = in_accessor<2>; // d
            using param_list = make_param_list<out, sup, rhs>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<0, -1>) {
                eval(out{}) = eval(rhs{} - sup{} * out{0, 0, 1});
            }

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::last_level) {
                eval(out{}) = eval(rhs{});
            }
        };
    } // namespace tridiagonal

    // tests/regression/horizontal_diffusion_fused.cpp:23-81 (run_single_stage(out_function) at :90)
    namespace horizontal_diffusion_fused {
        struct lap_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 1>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                using float_t = std::decay_t<decltype(eval(out()))>;
                eval(out()) =
                    float_t{4} * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
            }
        };

        struct flx_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 2, -1, 1>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                auto lap_hi = call<lap_function>::with(eval, in(1, 0));
                auto lap_lo = call<lap_function>::with(eval, in(0, 0));
                auto flx = lap_hi - lap_lo;
                eval(out()) = flx * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0 : flx;
            }
        };

        struct fly_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 2>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                auto lap_hi = call<lap_function>::with(eval, in(0, 1));
                auto lap_lo = call<lap_function>::with(eval, in(0, 0));
                auto fly = lap_hi - lap_lo;
                eval(out()) = fly * (eval(in(0, 1)) - eval(in(0, 0))) > 0 ? 0 : fly;
            }
        };

        struct out_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-2, 2, -2, 2>>;
            using coeff = in_accessor<2>;

            using param_list = make_param_list<out, in, coeff>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                auto flx_hi = call<flx_function>::with(eval, in(0, 0));
                auto flx_lo = call<flx_function>::with(eval, in(-1, 0));

                auto fly_hi = call<fly_function>::with(eval, in(0, 0));
                auto fly_lo = call<fly_function>::with(eval, in(0, -1));

                eval(out()) = eval(in()) - eval(coeff()) * (flx_hi - flx_lo + fly_hi - fly_lo);
            }
        };
    } // namespace horizontal_diffusion_fused

    // tests/regression/horizontal_diffusion_functions.cpp:22-179 (spec at :181-205)
    namespace horizontal_diffusion_functions {
        struct lap_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 1>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) = 4. * eval(in()) - (eval(in(-1, 0)) + eval(in(0, -1)) + eval(in(0, 1)) + eval(in(1, 0)));
            }
        };

        template <class Eval, class Acc>
        using float_type = std::decay_t<decltype(std::declval<Eval>()(Acc()))>;

        enum class variation { monolithic, call, call_offsets, procedures, procedures_offsets };

        template <variation, class Acc>
        struct lap;

        template <typename Acc>
        struct lap<variation::monolithic, Acc> {
            template <typename Evaluation>
            GT_FUNCTION static auto do00(Evaluation &eval) {
                return float_type<Evaluation, Acc>(4) * eval(Acc()) -
                       (eval(Acc(-1, 0)) + eval(Acc(0, -1)) + eval(Acc(0, 1)) + eval(Acc(1, 0)));
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do10(Evaluation &eval) {
                return float_type<Evaluation, Acc>(4) * eval(Acc(1, 0)) -
                       (eval(Acc(0, 0)) + eval(Acc(1, -1)) + eval(Acc(1, 1)) + eval(Acc(2, 0)));
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do01(Evaluation &eval) {
                return float_type<Evaluation, Acc>(4) * eval(Acc(0, 1)) -
                       (eval(Acc(-1, 1)) + eval(Acc(0, 0)) + eval(Acc(0, 2)) + eval(Acc(1, 1)));
            }
        };

        template <typename Acc>
        struct lap<variation::call, Acc> {
            template <typename Evaluation>
            GT_FUNCTION static auto do00(Evaluation &eval) {
                return call<lap_function>::at<0, 0, 0>::with(eval, Acc());
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do10(Evaluation &eval) {
                return call<lap_function>::at<1, 0, 0>::with(eval, Acc());
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do01(Evaluation &eval) {
                return call<lap_function>::at<0, 1, 0>::with(eval, Acc());
            }
        };

        template <typename Acc>
        struct lap<variation::call_offsets, Acc> {
            template <typename Evaluation>
            GT_FUNCTION static auto do00(Evaluation &eval) {
                return call<lap_function>::with(eval, Acc(0, 0));
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do10(Evaluation &eval) {
                return call<lap_function>::with(eval, Acc(1, 0));
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do01(Evaluation &eval) {
                return call<lap_function>::with(eval, Acc(0, 1));
            }
        };

        template <typename Acc>
        struct lap<variation::procedures, Acc> {
            template <typename Evaluation>
            GT_FUNCTION static auto do00(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::at<0, 0, 0>::with(eval, res, Acc());
                return res;
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do10(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::at<1, 0, 0>::with(eval, res, Acc());
                return res;
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do01(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::at<0, 1, 0>::with(eval, res, Acc());
                return res;
            }
        };

        template <typename Acc>
        struct lap<variation::procedures_offsets, Acc> {
            template <typename Evaluation>
            GT_FUNCTION static auto do00(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::with(eval, res, Acc(0, 0));
                return res;
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do10(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::with(eval, res, Acc(1, 0));
                return res;
            }
            template <typename Evaluation>
            GT_FUNCTION static auto do01(Evaluation &eval) {
                float_type<Evaluation, Acc> res;
                call_proc<lap_function>::with(eval, res, Acc(0, 1));
                return res;
            }
        };

        template <variation Variation>
        struct flx_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 2, -1, 1>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) = lap<Variation, in>::do10(eval) - lap<Variation, in>::do00(eval);
                eval(out()) = eval(out()) * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0.0 : eval(out());
            }
        };

        template <variation Variation>
        struct fly_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 2>>;

            using param_list = make_param_list<out, in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) = lap<Variation, in>::do01(eval) - lap<Variation, in>::do00(eval);
                eval(out()) = eval(out()) * (eval(in(0, 1, 0)) - eval(in(0, 0, 0))) > 0 ? 0.0 : eval(out());
            }
        };

        struct out_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1>;
            using flx = in_accessor<2, extent<-1, 0, 0, 0>>;
            using fly = in_accessor<3, extent<0, 0, -1, 0>>;
            using coeff = in_accessor<4>;

            using param_list = make_param_list<out, in, flx, fly, coeff>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(out()) =
                    eval(in()) - eval(coeff()) * (eval(flx()) - eval(flx(-1, 0)) + eval(fly()) - eval(fly(0, -1)));
            }
        };
    } // namespace horizontal_diffusion_functions

    // tests/regression/simple_hori_diff.cpp:25-61 (spec at :73-80)
    namespace simple_hori_diff {
        struct wlap_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<-1, 1, -1, 1>>;
            using crlato = in_accessor<2>;
            using crlatu = in_accessor<3>;

            using param_list = make_param_list<out, in, crlato, crlatu>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                using float_t = std::decay_t<decltype(eval(out()))>;
                eval(out()) = eval(in(1, 0)) + eval(in(-1, 0)) - float_t{2} * eval(in()) +
                              eval(crlato()) * (eval(in(0, 1)) - eval(in())) +
                              eval(crlatu()) * (eval(in(0, -1)) - eval(in()));
            }
        };

        struct divflux_function {
            using out = inout_accessor<0>;
            using in = in_accessor<1>;
            using lap = in_accessor<2, extent<-1, 1, -1, 1>>;
            using crlato = in_accessor<3>;
            using coeff = in_accessor<4>;

            using param_list = make_param_list<out, in, lap, crlato, coeff>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation &eval) {
                auto fluxx = eval(lap(1, 0)) - eval(lap());
                auto fluxx_m = eval(lap()) - eval(lap(-1, 0));

                auto fluxy = eval(crlato()) * (eval(lap(0, 1)) - eval(lap()));
                auto fluxy_m = eval(crlato()) * (eval(lap()) - eval(lap(0, -1)));

                eval(out()) = eval(in()) + ((fluxx_m - fluxx) + (fluxy_m - fluxy)) * eval(coeff());
            }
        };
    } // namespace simple_hori_diff

    // tests/regression/positional_stencil.cpp:21-32
    namespace positional_stencil {
        struct functor {
            using out = inout_accessor<0>;
            using i_pos = in_accessor<1>;
            using j_pos = in_accessor<2>;
            using k_pos = in_accessor<3>;
            using param_list = make_param_list<out, i_pos, j_pos, k_pos>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval) {
                eval(out()) = eval(i_pos()) + eval(j_pos()) + eval(k_pos());
            }
        };
    } // namespace positional_stencil

    // tests/regression/parallel_multistage_fusion.cpp:23-52
    namespace parallel_multistage_fusion {
        using axis_t = axis<1>;
        using full_t = axis_t::full_interval;

        struct set_constant {
            using out = inout_accessor<0>;

            using param_list = make_param_list<out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval) {
                eval(out()) = 1;
            }
        };

        struct copy_with_vertical_offset {
            using in = in_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
            using out = inout_accessor<1>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::modify<0, -1>) {
                eval(out()) = eval(in(0, 0, 1));
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, full_t::last_level) {
                eval(out()) = eval(in());
            }
        };
    } // namespace parallel_multistage_fusion

    // tests/regression/expandable_parameters.cpp:22-33
    namespace expandable_parameters {
        struct copy_functor {
            using parameters_out = inout_accessor<0>;
            using parameters_in = accessor<1>;

            using param_list = make_param_list<parameters_out, parameters_in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation &eval) {
                eval(parameters_out{}) = eval(parameters_in{});
            }
        };
    } // namespace expandable_parameters

    // tests/regression/expandable_parameters_single_kernel.cpp:22-56
    namespace expandable_parameters_single_kernel {
        struct functor_single_kernel {
            using parameters1_out = inout_accessor<0>;
            using parameters2_out = inout_accessor<1>;
            using parameters3_out = inout_accessor<2>;
            using parameters4_out = inout_accessor<3>;
            using parameters5_out = inout_accessor<4>;

            using parameters1_in = in_accessor<5>;
            using parameters2_in = in_accessor<6>;
            using parameters3_in = in_accessor<7>;
            using parameters4_in = in_accessor<8>;
            using parameters5_in = in_accessor<9>;

            using param_list = make_param_list<parameters1_out,
                parameters2_out,
                parameters3_out,
                parameters4_out,
                parameters5_out,
                parameters1_in,
                parameters2_in,
                parameters3_in,
                parameters4_in,
                parameters5_in>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(parameters1_out()) = eval(parameters1_in());
                eval(parameters2_out()) = eval(parameters2_in());
                eval(parameters3_out()) = eval(parameters3_in());
                eval(parameters4_out()) = eval(parameters4_in());
                eval(parameters5_out()) = eval(parameters5_in());
            }
        };
    } // namespace expandable_parameters_single_kernel

    // tests/regression/advection_pdbott_prepare_tracers.cpp:23-35
    namespace advection_pdbott_prepare_tracers {
        struct prepare_tracers {
            using data = inout_accessor<0>;
            using data_nnow = in_accessor<1>;
            using rho = in_accessor<2>;

            using param_list = make_param_list<data, data_nnow, rho>;

            template <typename Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(data()) = eval(rho()) * eval(data_nnow());
            }
        };
    } // namespace advection_pdbott_prepare_tracers
} // namespace reference_specs

namespace reference_specs {
    // tests/unit_tests/stencil/frontend/cartesian/test_kcache_fill.cpp:20-125 (axis, functors; specs in the tests)
    namespace test_kcache_fill {
        using axis_t = axis<3, axis_config::offset_limit<3>>;
        using kfull = axis_t::full_interval;

        struct shift_acc_forward_fill {
            using in = in_accessor<0, extent<0, 0, 0, 0, -1, 1>>;
            using out = inout_accessor<1>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(out()) = eval(in()) + eval(in(0, 0, 1));
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, -1>) {
                eval(out()) = eval(in(0, 0, -1)) + eval(in()) + eval(in(0, 0, 1));
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(out()) = eval(in(0, 0, -1)) + eval(in());
            }
        };

        struct shift_acc_backward_fill {
            using in = in_accessor<0, extent<0, 0, 0, 0, -1, 1>>;
            using out = inout_accessor<1>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(out()) = eval(in()) + eval(in(0, 0, -1));
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, -1>) {
                eval(out()) = eval(in(0, 0, 1)) + eval(in()) + eval(in(0, 0, -1));
            }
            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(out()) = eval(in()) + eval(in(0, 0, 1));
            }
        };

        struct copy_fill {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull) {
                eval(out()) = eval(in());
            }
        };
    } // namespace test_kcache_fill

    // tests/unit_tests/stencil/frontend/cartesian/test_kcache_flush.cpp:20-99
    namespace test_kcache_flush {
        using axis_t = axis<3, axis_config::offset_limit<3>>;
        using kfull = axis_t::full_interval;

        struct shift_acc_forward_flush {
            using in = in_accessor<0>;
            using out = inout_accessor<1, extent<0, 0, 0, 0, -1, 0>>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(out()) = eval(in());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, 0>) {
                eval(out()) = eval(out(0, 0, -1)) + eval(in());
            }
        };

        struct shift_acc_backward_flush {
            using in = in_accessor<0>;
            using out = inout_accessor<1, extent<0, 0, 0, 0, 0, 1>>;

            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(out()) = eval(in());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<0, -1>) {
                eval(out()) = eval(out(0, 0, 1)) + eval(in());
            }
        };
    } // namespace test_kcache_flush

    // tests/unit_tests/stencil/frontend/cartesian/test_kcache_fill_and_flush.cpp:20-125
    namespace test_kcache_fill_and_flush {
        using axis_t = axis<3, axis_config::offset_limit<3>>;
        using kfull = axis_t::full_interval;

        struct shift_acc_forward_fill_and_flush {
            using in = inout_accessor<0, extent<0, 0, 0, 0, -1, 0>>;
            using param_list = make_param_list<in>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, 0>) {
                eval(in()) = eval(in()) + eval(in(0, 0, -1));
            }
            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(in()) = eval(in());
            }
        };

        struct shift_acc_backward_fill_and_flush {
            using in = inout_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
            using param_list = make_param_list<in>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<0, -1>) {
                eval(in()) = eval(in()) + eval(in(0, 0, 1));
            }
            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(in()) = eval(in());
            }
        };

        struct copy_fill {
            using in = inout_accessor<0>;
            using param_list = make_param_list<in>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull) {
                eval(in()) = eval(in());
            }
        };

        struct scale_fill {
            using in = inout_accessor<0>;
            using param_list = make_param_list<in>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull) {
                eval(in()) = 2 * eval(in());
            }
        };
    } // namespace test_kcache_fill_and_flush

    // tests/unit_tests/stencil/frontend/cartesian/test_kcache_local.cpp:20-213
    namespace test_kcache_local {
        using axis_t = axis<3, axis_config::offset_limit<3>>;
        using kfull = axis_t::full_interval;

        struct shift_acc_forward {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using buff = inout_accessor<2, extent<0, 0, 0, 0, -1, 0>>;

            using param_list = make_param_list<in, out, buff>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(buff()) = eval(in());
                eval(out()) = eval(buff());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, 0>) {
                eval(buff()) = eval(buff(0, 0, -1)) + eval(in());
                eval(out()) = eval(buff());
            }
        };

        struct shift_acc_backward {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using buff = inout_accessor<2, extent<0, 0, 0, 0, 0, 1>>;

            using param_list = make_param_list<in, out, buff>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(buff()) = eval(in());
                eval(out()) = eval(buff());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<0, -1>) {
                eval(buff()) = eval(buff(0, 0, 1)) + eval(in());
                eval(out()) = eval(buff());
            }
        };

        struct biside_large_kcache_forward {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using buff = inout_accessor<2, extent<0, 0, 0, 0, -2, 1>>;

            using param_list = make_param_list<in, out, buff>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(buff()) = eval(in());
                eval(buff(0, 0, 1)) = eval(in()) / 2;
                eval(out()) = eval(buff());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level::shift<1>) {
                eval(buff(0, 0, 1)) = eval(in()) / 2;
                eval(out()) = eval(buff()) + eval(buff(0, 0, -1)) / 4;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<2, -1>) {
                eval(buff(0, 0, 1)) = eval(in()) / 2;
                eval(out()) = eval(buff()) + eval(buff(0, 0, -1)) / 4 + eval(buff(0, 0, -2)) * .12;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(out()) = eval(buff()) + eval(buff(0, 0, -1)) / 4 + eval(buff(0, 0, -2)) * .12;
            }
        };

        struct biside_large_kcache_backward {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using buff = inout_accessor<2, extent<0, 0, 0, 0, -1, 2>>;

            using param_list = make_param_list<in, out, buff>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level) {
                eval(buff()) = eval(in());
                eval(buff(0, 0, -1)) = eval(in()) / 2;
                eval(out()) = eval(buff());
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::last_level::shift<-1>) {
                eval(buff(0, 0, -1)) = eval(in()) / 2;
                eval(out()) = eval(buff()) + eval(buff(0, 0, 1)) / 4;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::modify<1, -2>) {
                eval(buff(0, 0, -1)) = eval(in()) / 2;
                eval(out()) = eval(buff()) + eval(buff(0, 0, 1)) / 4 + eval(buff(0, 0, 2)) * .12;
            }

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, kfull::first_level) {
                eval(out()) = eval(buff()) + eval(buff(0, 0, 1)) / 4 + eval(buff(0, 0, 2)) * .12;
            }
        };
    } // namespace test_kcache_local

    // tests/unit_tests/stencil/frontend/cartesian/test_kparallel.cpp:23-68
    namespace test_kparallel {
        struct parallel_functor {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, axis<2>::get_interval<0>) {
                eval(out()) = eval(in());
            }
            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, axis<2>::get_interval<1>) {
                eval(out()) = 2 * eval(in());
            }
        };

        struct parallel_functor_on_upper_interval {
            using in = in_accessor<0>;
            using out = inout_accessor<1>;
            using param_list = make_param_list<in, out>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval, axis<3>::get_interval<1>) {
                eval(out()) = eval(in());
            }
        };
    } // namespace test_kparallel

    // tests/regression/whole_axis_access.cpp:25-38
    namespace whole_axis_access {
        struct functor {
            using out = inout_accessor<0>;
            using in = in_accessor<1, extent<>, 4>;
            using k_pos = in_accessor<2>;
            using param_list = make_param_list<out, in, k_pos>;

            template <class Eval>
            GT_FUNCTION static void apply(Eval &&eval) {
                auto k = eval(k_pos());
                std::decay_t<decltype(eval(out()))> res = 0;
                for (int kk = 0; kk < k; ++kk)
                    res += eval(in(0, 0, 0, kk));
                eval(out()) = res;
            }
        };
    } // namespace whole_axis_access
} // namespace reference_specs

#include <gt355/boundaries.hpp>

namespace reference_specs {
    // examples/boundaries/boundaries.cpp:63-110 (uint_t spelled unsigned): writes `value` into the halo of the first
    // field, except the directions (*, -1, *) and (-1, -1, *), which copy the second field from a shifted position
    namespace examples_boundaries {
        namespace bd = gt355::boundaries;

        template <typename T>
        struct direction_bc_input {
            T value;

            GT_FUNCTION
            direction_bc_input() : value(1) {}

            GT_FUNCTION
            direction_bc_input(T v) : value(v) {}

            // This is the primary template and it will be picked when all the other specializations below fail to be
            // selected.  It is important for debugging to note that, if a needed specialization i missing, this
            // version will be selected.
            template <typename Direction, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(
                Direction, DataField0 &data_field0, DataField1 const &, unsigned i, unsigned j, unsigned k) const {
                data_field0(i, j, k) = value;
            }

            template <bd::sign I, bd::sign K, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(bd::direction<I, bd::minus_, K>,
                DataField0 &data_field0,
                DataField1 const &data_field1,
                unsigned i,
                unsigned j,
                unsigned k) const {
                data_field0(i, j, k) = data_field1(i, j + 1, k);
            }

            template <bd::sign K, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(bd::direction<bd::minus_, bd::minus_, K>,
                DataField0 &data_field0,
                DataField1 const &data_field1,
                unsigned i,
                unsigned j,
                unsigned k) const {
                data_field0(i, j, k) = data_field1(i + 1, j + 1, k);
            }

            template <typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(bd::direction<bd::minus_, bd::minus_, bd::minus_>,
                DataField0 &data_field0,
                DataField1 const &data_field1,
                unsigned i,
                unsigned j,
                unsigned k) const {
                data_field0(i, j, k) = data_field1(i + 1, j + 1, k + 1);
            }
        };
    } // namespace examples_boundaries

    // tests/regression/boundary_conditions.cpp:20-59: writes the second field -- a factor times the first field's
    // value by default, fixed values on the (*, -1, *) faces, the (-1, -1, *) edges and the (-1, -1, -1) corner
    namespace boundary_conditions {
        using namespace gt355::boundaries;

        template <class Float>
        constexpr Float face_value = 88;
        template <class Float>
        constexpr Float edge_value = 77777;
        template <class Float>
        constexpr Float corner_value = 55555;
        template <class Float>
        constexpr Float factor = 2;

        template <class Float>
        struct direction_bc_input {
            template <typename Direction, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(
                Direction, DataField0 &data_field0, DataField1 const &data_field1, unsigned i, unsigned j, unsigned k) const {
                data_field1(i, j, k) = data_field0(i, j, k) * factor<Float>;
            }

            template <sign I, sign K, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(
                direction<I, minus_, K>, DataField0 &, DataField1 const &data_field1, unsigned i, unsigned j, unsigned k) const {
                data_field1(i, j, k) = face_value<Float> * factor<Float>;
            }

            template <sign K, typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(direction<minus_, minus_, K>,
                DataField0 &,
                DataField1 const &data_field1,
                unsigned i,
                unsigned j,
                unsigned k) const {
                data_field1(i, j, k) = edge_value<Float> * factor<Float>;
            }

            template <typename DataField0, typename DataField1>
            GT_FUNCTION void operator()(direction<minus_, minus_, minus_>,
                DataField0 &,
                DataField1 const &data_field1,
                unsigned i,
                unsigned j,
                unsigned k) const {
                data_field1(i, j, k) = corner_value<Float> * factor<Float>;
            }
        };
    } // namespace boundary_conditions
} // namespace reference_specs
