// The generic stencil engine (include/gt355/stencil.hpp) on user specs: the reference's own regression and
// unit-test specs (reference_specs.hpp, unmodified functors under their reference names), run on the gfx950 backend
// tag.  Expected values come from the oracle (oracle/gt_oracle.h) on the same seeded inputs (FP64 must match bit for
// bit: the functors evaluate the same operations in the same order) or from the reference tests' known answers.
//
//   gt355_generic            all cases (needs a GPU)
//   gt355_generic --bench N  times the engine on the hdiff, vadv and tridiagonal specs at 256x256x80 FP64 (one JSON
//                            line)
#include <gt355/stencil.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <functional>
#include <vector>

#include "../../oracle/gt_oracle.h"
#include "reference_specs.hpp"

using namespace gt355;
using namespace gt355::stencil;
using namespace gt355::stencil::cartesian;

namespace {
    int failures = 0;

    void report(const char *name, bool ok) {
        std::printf("[%s] %s\n", ok ? "ok" : "FAIL", name);
        std::fflush(stdout);
        if (!ok)
            ++failures;
    }

    template <class T>
    using store = std::shared_ptr<storage::data_store<T>>;

    // seeded store over the whole allocation, the library's fill_uniform sequence
    template <class T>
    store<T> seeded(int n0, int n1, int n2, int h, uint64_t seed, uint64_t fid, double lo, double hi) {
        std::vector<double> v((size_t)n0 * n1 * n2);
        gto_fill_uniform_f64(v.data(), (int64_t)v.size(), seed, fid, lo, hi);
        return storage::builder<storage::gpu>.template type<T>().dimensions(n0, n1, n2).halos(h, h, 0)
            .initializer([v, n0, n1](int i, int j, int k) { return v[i + (size_t)n0 * (j + (size_t)n1 * k)]; })
            .build();
    }

    // host copy in the oracle layout [k][j][i padded] (element (0,0,0) at index 0)
    template <class T>
    std::vector<T> mirror(store<T> const &ds) {
        auto v = ds->const_host_view();
        auto len = ds->lengths();
        std::vector<T> out((size_t)v.stride[2] * len[2]);
        for (int k = 0; k < len[2]; ++k)
            for (int j = 0; j < len[1]; ++j)
                for (int i = 0; i < len[0]; ++i)
                    out[i + j * v.stride[1] + k * v.stride[2]] = v(i, j, k);
        return out;
    }

    template <class T>
    bool same_interior(store<T> const &ds, std::vector<T> const &ref, grid const &g) {
        auto v = ds->const_host_view();
        for (int k = 0; k < g.k_size; ++k)
            for (int j = g.j_start; j < g.j_start + g.j_size; ++j)
                for (int i = g.i_start; i < g.i_start + g.i_size; ++i) {
                    T a = v(i, j, k), b = ref[i + j * v.stride[1] + k * v.stride[2]];
                    if (std::memcmp(&a, &b, sizeof(T)) != 0) {
                        std::printf("  mismatch at (%d,%d,%d): %.17g vs %.17g\n", i, j, k, (double)a, (double)b);
                        return false;
                    }
                }
        return true;
    }

    grid halo_grid(int ni, int nj, int nk, int h) {
        return make_grid(halo_descriptor(h, h, h, ni + h - 1, ni + 2 * h), halo_descriptor(h, h, h, nj + h - 1, nj + 2 * h),
            nk);
    }

    // ------------------------------------------------------------------------------------- copy and laplacian
    void copy_and_laplacian() {
        const int ni = 23, nj = 11, nk = 43, h = 1;
        auto in = seeded<double>(ni + 2, nj + 2, nk, h, 42, 1, -1, 1);
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 2, nj + 2, nk).halos(h, h, 0);
        auto out = b.build(), lap = b.build();
        auto g = halo_grid(ni, nj, nk, h);
        run_single_stage(reference_specs::copy_stencil::copy_functor(), gfx950<>(), g, in, out);
        report("run_single_stage copy", same_interior(out, mirror(in), g));
        run_single_stage(reference_specs::laplacian::lap(), gfx950<>(), g, lap, in);
        auto hin = mirror(in);
        std::vector<double> ref(hin.size(), 0.);
        auto st = in->strides();
        const size_t o = h + h * st[1];
        gto_lap_f64(ref.data() + o, hin.data() + o, ni, nj, nk, st[1], st[2]);
        report("run_single_stage laplacian (extent <-1,1,-1,1>)", same_interior(lap, ref, g));
    }

    // ------------------------------------------------------------------------------------- horizontal diffusion
    void horizontal_diffusion(int ni, int nj, int nk, bool cached = true) {
        const int h = 2;
        auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
        auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
        auto out = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0).build();
        auto g = halo_grid(ni, nj, nk, h);
        if (cached)
            run([](auto in, auto coeff, auto out) {
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .ij_cached(lap, flx, fly)
                    .stage(reference_specs::horizontal_diffusion::lap_function(), lap, in)
                    .stage(reference_specs::horizontal_diffusion::flx_function(), flx, in, lap)
                    .stage(reference_specs::horizontal_diffusion::fly_function(), fly, in, lap)
                    .stage(reference_specs::horizontal_diffusion::out_function(), out, in, flx, fly, coeff);
            },
                gfx950<>(), g, in, coeff, out);
        else
            run([](auto in, auto coeff, auto out) {
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .stage(reference_specs::horizontal_diffusion::lap_function(), lap, in)
                    .stage(reference_specs::horizontal_diffusion::flx_function(), flx, in, lap)
                    .stage(reference_specs::horizontal_diffusion::fly_function(), fly, in, lap)
                    .stage(reference_specs::horizontal_diffusion::out_function(), out, in, flx, fly, coeff);
            },
                gfx950<>(), g, in, coeff, out);
        auto hin = mirror(in), hco = mirror(coeff);
        std::vector<double> ref(hin.size(), 0.);
        auto st = in->strides();
        const size_t o = h + h * st[1];
        gto_hdiff_f64(ref.data() + o, hin.data() + o, hco.data() + o, ni, nj, nk, st[1], st[2], 4);
        char name[96];
        std::snprintf(name, sizeof name, "fused multistage hdiff, temporaries %s %dx%dx%d",
            cached ? "ij-cached in LDS" : "in HBM", ni, nj, nk);
        report(name, same_interior(out, ref, g));
    }

    // ------------------------------------------------------------------------------------- stencil functions
    // horizontal_diffusion_fused.cpp:23-91: the flux and output functors call the laplacian / flux functors as
    // stencil functions at shifted points; run as one stage
    // horizontal_diffusion_functions.cpp:22-205: flx / fly evaluate the laplacian at the point and its east or
    // north neighbour in five ways (monolithic expressions, call<>::at<>, call<> with offset accessors,
    // call_proc<>::at<>, call_proc<> with offset accessors); flx and fly are ij-cached, lap is never stored
    template <reference_specs::horizontal_diffusion_functions::variation V>
    void hdiff_functions(const char *vname, int ni, int nj, int nk) {
        const int h = 2;
        auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
        auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
        auto out = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0).build();
        auto g = halo_grid(ni, nj, nk, h);
        run([](auto coeff, auto in, auto out) {
            GT_DECLARE_TMP(double, flx, fly);
            return execute_parallel()
                .ij_cached(flx, fly)
                .stage(reference_specs::horizontal_diffusion_functions::flx_function<V>(), flx, in)
                .stage(reference_specs::horizontal_diffusion_functions::fly_function<V>(), fly, in)
                .stage(reference_specs::horizontal_diffusion_functions::out_function(), out, in, flx, fly, coeff);
        },
            gfx950<>(), g, coeff, in, out);
        // expected: the same operations in the same order on the host (the laplacian sums its neighbours as
        // i-1, j-1, j+1, i+1 here, unlike lap_function of horizontal_diffusion.cpp), and the oracle's hdiff within
        // 1e-12 relative (same mathematics, other summation order)
        auto hin = mirror(in), hco = mirror(coeff);
        auto st = in->strides();
        const int64_t sj = st[1], sk = st[2];
        auto at = [&](std::vector<double> const &a, int i, int j, int k) { return a[i + j * sj + k * sk]; };
        auto lap = [&](int i, int j, int k) {
            return 4. * at(hin, i, j, k) -
                   (at(hin, i - 1, j, k) + at(hin, i, j - 1, k) + at(hin, i, j + 1, k) + at(hin, i + 1, j, k));
        };
        auto flx = [&](int i, int j, int k) {
            const double f = lap(i + 1, j, k) - lap(i, j, k);
            return f * (at(hin, i + 1, j, k) - at(hin, i, j, k)) > 0 ? 0.0 : f;
        };
        auto fly = [&](int i, int j, int k) {
            const double f = lap(i, j + 1, k) - lap(i, j, k);
            return f * (at(hin, i, j + 1, k) - at(hin, i, j, k)) > 0 ? 0.0 : f;
        };
        std::vector<double> ref(hin.size(), 0.), ora(hin.size(), 0.);
        for (int k = 0; k < nk; ++k)
            for (int j = h; j < h + nj; ++j)
                for (int i = h; i < h + ni; ++i)
                    ref[i + j * sj + k * sk] = at(hin, i, j, k) - at(hco, i, j, k) * (flx(i, j, k) - flx(i - 1, j, k) +
                                                                                          fly(i, j, k) - fly(i, j - 1, k));
        const size_t o = h + h * sj;
        gto_hdiff_f64(ora.data() + o, hin.data() + o, hco.data() + o, ni, nj, nk, sj, sk, 4);
        auto v = out->const_host_view();
        bool same = true;
        double maxref = 0, maxdiff = 0;
        for (int k = 0; k < nk; ++k)
            for (int j = h; j < h + nj; ++j)
                for (int i = h; i < h + ni; ++i) {
                    const double got = v(i, j, k), r = ref[i + j * sj + k * sk], q = ora[i + j * sj + k * sk];
                    same = same && got == r;
                    maxref = std::max(maxref, std::fabs(q));
                    maxdiff = std::max(maxdiff, std::fabs(got - q));
                }
        char name[128];
        std::snprintf(name, sizeof name, "horizontal_diffusion_functions (%s) %dx%dx%d: bit-identical; %.1e of the oracle",
            vname, ni, nj, nk, maxdiff / maxref);
        report(name, same && maxdiff <= 1e-12 * maxref);
    }

    // expandable_parameters.cpp:22-60 and expandable_parameters_single_kernel.cpp:22-100: copies through an ij-cached
    // temporary, over five members by expandable_run (expandable temporary) and as one spec of ten fields
    void expandable_parameters(int ni, int nj, int nk) {
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk);
        auto g = make_grid(ni, nj, nk);
        auto make = [&](double v) { return b.value(v).build(); };
        auto check_copies = [&](std::vector<store<double>> const &in, std::vector<store<double>> const &out) {
            bool ok = true;
            for (size_t m = 0; m < in.size(); ++m) {
                auto vi = in[m]->const_host_view();
                auto vo = out[m]->const_host_view();
                for (int k = 0; k < nk && ok; ++k)
                    for (int j = 0; j < nj && ok; ++j)
                        for (int i = 0; i < ni && ok; ++i)
                            ok = vo(i, j, k) == vi(i, j, k);
            }
            return ok;
        };
        {
            std::vector<store<double>> out = {make(1.), make(2.), make(3.), make(4.), make(5.)};
            std::vector<store<double>> in = {make(-1.), make(-2.), make(-3.), make(-4.), make(-5.)};
            expandable_run<2>([](auto in, auto out) {
                GT_DECLARE_EXPANDABLE_TMP(double, tmp);
                return execute_parallel().ij_cached(tmp).stage(reference_specs::expandable_parameters::copy_functor(), tmp, in)
                    .stage(reference_specs::expandable_parameters::copy_functor(), out, tmp);
            },
                gfx950<>(), g, in, out);
            report("expandable_parameters: expandable_run<2> over 5 members, expandable ij-cached temporary",
                check_copies(in, out));
        }
        {
            std::vector<store<double>> out = {make(1.), make(2.), make(3.), make(4.), make(5.)};
            std::vector<store<double>> in = {make(-1.), make(-2.), make(-3.), make(-4.), make(-5.)};
            run([](auto out0, auto out1, auto out2, auto out3, auto out4, auto in0, auto in1, auto in2, auto in3,
                    auto in4) {
                GT_DECLARE_TMP(double, tmp0, tmp1, tmp2, tmp3, tmp4);
                return execute_parallel()
                    .ij_cached(tmp0, tmp1, tmp2, tmp3, tmp4)
                    .stage(reference_specs::expandable_parameters_single_kernel::functor_single_kernel(), tmp0, tmp1, tmp2, tmp3, tmp4, in0, in1, in2, in3, in4)
                    .stage(reference_specs::expandable_parameters_single_kernel::functor_single_kernel(), out0, out1, out2, out3, out4, tmp0, tmp1, tmp2, tmp3, tmp4);
            },
                gfx950<>(), g, out[0], out[1], out[2], out[3], out[4], in[0], in[1], in[2], in[3], in[4]);
            report("expandable_parameters_single_kernel: ten fields, five ij-cached temporaries", check_copies(in, out));
        }
    }

    // call_proc and at<>: a procedure writing two outputs, called one point to the east
    struct sf_pair {
        using a = inout_accessor<0>;
        using b = inout_accessor<1>;
        using in = in_accessor<2, extent<0, 1, 0, 0>>;
        using param_list = make_param_list<a, b, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(a()) = 2 * eval(in());
            eval(b()) = eval(in()) + 1;
        }
    };
    struct sf_use_proc {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<0, 1, 0, 0>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            double a = 0, b = 0;
            call_proc<sf_pair>::at<1, 0, 0>::with(eval, a, b, in());
            eval(out()) = a - b; // 2 in(i+1) - (in(i+1) + 1)
        }
    };

    void stencil_functions(int ni, int nj, int nk) {
        const int h = 2;
        auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
        auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0);
        auto out = b.build(), out2 = b.build();
        auto g = halo_grid(ni, nj, nk, h);
        run_single_stage(reference_specs::horizontal_diffusion_fused::out_function(), gfx950<>(), g, out, in, coeff);
        auto hin = mirror(in), hco = mirror(coeff);
        std::vector<double> ref(hin.size(), 0.);
        auto st = in->strides();
        const size_t o = h + h * st[1];
        gto_hdiff_f64(ref.data() + o, hin.data() + o, hco.data() + o, ni, nj, nk, st[1], st[2], 4);
        char name[96];
        std::snprintf(name, sizeof name, "stencil functions call<>: hdiff as one stage %dx%dx%d", ni, nj, nk);
        report(name, same_interior(out, ref, g));
        run_single_stage(sf_use_proc(), gfx950<>(), g, out2, in);
        auto v = out2->const_host_view();
        bool ok = true;
        for (int k = 0; k < nk && ok; ++k)
            for (int j = h; j < h + nj && ok; ++j)
                for (int i = h; i < h + ni && ok; ++i) {
                    const double x = hin[i + 1 + j * st[1] + k * st[2]];
                    ok = v(i, j, k) == 2 * x - (x + 1);
                }
        report("call_proc<>::at<1,0,0> with value outputs", ok);
    }

    // ------------------------------------------------------------------------------------- vertical advection
    void vertical_advection(int ni, int nj, int nk) {
        const int h = 3;
        auto utens_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 6, -1, 1);
        auto u_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 3, -1, 1);
        auto u_pos = seeded<double>(ni + 6, nj + 6, nk, h, 42, 4, -1, 1);
        auto utens = seeded<double>(ni + 6, nj + 6, nk, h, 42, 5, -1, 1);
        auto wcon = seeded<double>(ni + 6, nj + 6, nk + 1, h, 42, 7, -0.5, 0.5);
        auto g = halo_grid(ni, nj, nk, h);
        auto uts0 = mirror(utens_stage);
        run(reference_specs::vertical_advection_dycore::spec<double>(), gfx950<>(), g, utens_stage, u_stage, wcon, u_pos,
            utens, global_parameter(0.15));
        auto hus = mirror(u_stage), hup = mirror(u_pos), hut = mirror(utens), hw = mirror(wcon);
        auto st = u_stage->strides(), wst = wcon->strides();
        const size_t o = h + h * st[1], wo = h + h * wst[1];
        gto_vadv_f64(uts0.data() + o, hus.data() + o, hw.data() + wo, hup.data() + o, hut.data() + o, 0.15, ni, nj, nk,
            st[1], st[2], wst[1], wst[2], 4);
        char name[96];
        std::snprintf(name, sizeof name, "multi_pass forward/backward vadv with intervals %dx%dx%d", ni, nj, nk);
        report(name, same_interior(utens_stage, uts0, g));
    }

    // ------------------------------------------------------------------------------------- tridiagonal (expressions)
    void tridiagonal(int ni, int nj, int nk, bool known_answer) {
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk);
        store<double> inf, diag, sup, rhs;
        if (known_answer) {
            // tridiagonal.cpp:81-97: -1 / 3 / 1 bands and a right-hand side making the solution 1
            inf = b.value(-1.).build();
            diag = b.value(3.).build();
            sup = b.value(1.).build();
            rhs = b.initializer([nk](int, int, int k) { return k == 0 ? 4. : k == nk - 1 ? 2. : 3.; }).build();
        } else {
            inf = seeded<double>(ni, nj, nk, 0, 7, 11, -1, 1);
            diag = seeded<double>(ni, nj, nk, 0, 7, 12, 3, 4);
            sup = seeded<double>(ni, nj, nk, 0, 7, 13, -1, 1);
            rhs = seeded<double>(ni, nj, nk, 0, 7, 14, -1, 1);
        }
        auto out = b.build();
        auto hinf = mirror(inf), hdiag = mirror(diag), hsup = mirror(sup), hrhs = mirror(rhs);
        run([](auto inf, auto diag, auto sup, auto rhs, auto out) {
            return multi_pass(execute_forward().stage(reference_specs::tridiagonal::forward_thomas(), inf, diag, sup, rhs),
                execute_backward().stage(reference_specs::tridiagonal::backward_thomas(), out, sup, rhs));
        },
            gfx950<>(), make_grid(ni, nj, nk), inf, diag, sup, rhs, out);
        std::vector<double> ref(hinf.size(), 0.);
        auto st = inf->strides();
        gto_tridiagonal_f64(ref.data(), hinf.data(), hdiag.data(), hsup.data(), hrhs.data(), ni, nj, nk, st[1], st[2]);
        bool ok = same_interior(out, ref, make_grid(ni, nj, nk)) && same_interior(sup, hsup, make_grid(ni, nj, nk)) &&
                  same_interior(rhs, hrhs, make_grid(ni, nj, nk));
        if (known_answer) {
            auto v = out->const_host_view();
            for (int k = 0; k < nk && ok; ++k)
                ok = std::fabs(v(ni - 1, nj - 1, k) - 1.) < 1e-14;
        }
        char name[96];
        std::snprintf(name, sizeof name, "tridiagonal (expressions, inout fields) %dx%dx%d%s", ni, nj, nk,
            known_answer ? " known answer" : "");
        report(name, ok);
    }

    // ------------------------------------------------------------------------------------- simple_hori_diff (j-only)
    void simple_hori_diff(int ni, int nj, int nk) {
        const int h = 2;
        auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 21, -1, 1);
        auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 22, 0, 0.25);
        auto jb = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0)
                      .selector<0, 1, 0>();
        std::vector<double> cro(nj + 4), cru(nj + 4);
        gto_fill_uniform_f64(cro.data(), nj + 4, 42, 23, 0.5, 1.5);
        gto_fill_uniform_f64(cru.data(), nj + 4, 42, 24, 0.5, 1.5);
        auto crlato = jb.initializer([cro](int, int j, int) { return cro[j]; }).build();
        auto crlatu = jb.initializer([cru](int, int j, int) { return cru[j]; }).build();
        auto out = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0).build();
        auto g = halo_grid(ni, nj, nk, h);
        run([](auto coeff, auto in, auto out, auto crlato, auto crlatu) {
            GT_DECLARE_TMP(double, lap);
            return execute_parallel()
                .ij_cached(lap)
                .stage(reference_specs::simple_hori_diff::wlap_function(), lap, in, crlato, crlatu)
                .stage(reference_specs::simple_hori_diff::divflux_function(), out, in, lap, crlato, coeff);
        },
            gfx950<>(), g, coeff, in, out, crlato, crlatu);
        auto hin = mirror(in), hco = mirror(coeff);
        std::vector<double> ref(hin.size(), 0.);
        auto st = in->strides();
        const size_t o = h + h * st[1];
        gto_shd_f64(ref.data() + o, hin.data() + o, cro.data() + h, cru.data() + h, 0, 1, 0, hco.data() + o, ni, nj, nk,
            st[1], st[2]);
        char name[96];
        std::snprintf(name, sizeof name, "simple_hori_diff spec, j-only metric %dx%dx%d", ni, nj, nk);
        report(name, same_interior(out, ref, g));
    }

    // ------------------------------------------------------------------------------------- positional, fusion
    void positional_and_fusion() {
        const int ni = 12, nj = 33, nk = 61, h = 2;
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0);
        auto out = b.value(-1.).build();
        auto g = halo_grid(ni, nj, nk, h);
        run_single_stage(reference_specs::positional_stencil::functor(), gfx950<>(), g, out, positional<dim::i>(), positional<dim::j>(),
            positional<dim::k>());
        bool ok = true;
        {
            auto v = out->const_host_view();
            for (int k = 0; k < nk && ok; ++k)
                for (int j = 0; j < nj + 4 && ok; ++j)
                    for (int i = 0; i < ni + 4 && ok; ++i) {
                        const bool inside = i >= h && i < h + ni && j >= h && j < h + nj;
                        ok = v(i, j, k) == (inside ? double(i + j + k) : -1.);
                    }
        }
        // positional_stencil.cpp:36-44: the positional value is the storage index (grid origin included)
        report("positional<dim::i/j/k> (storage indices, halo untouched)", ok);
        // parallel_multistage_fusion.cpp:54-66: a temporary set to 1 by one parallel multistage, copied with a
        // vertical offset by the next
        auto out2 = b.value(0.).build();
        run([](auto out) {
            GT_DECLARE_TMP(double, tmp);
            using namespace reference_specs::parallel_multistage_fusion;
            return multi_pass(execute_parallel().stage(set_constant(), tmp),
                execute_parallel().stage(copy_with_vertical_offset(), tmp, out));
        },
            gfx950<>(), g, out2);
        ok = true;
        auto v = out2->const_host_view();
        for (int k = 0; k < nk && ok; ++k)
            for (int j = h; j < h + nj && ok; ++j)
                for (int i = h; i < h + ni && ok; ++i)
                    ok = v(i, j, k) == 1.;
        report("multi_pass of parallel multistages through a temporary (k offset, intervals)", ok);
    }

    // ------------------------------------------------------------------------------------- expandable_run, errors
    void expandable_and_errors() {
        const int ni = 12, nj = 33, nk = 61;
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk);
        std::vector<store<double>> in, out;
        for (int m = 0; m < 5; ++m) {
            out.push_back(b.build());
            in.push_back(b.value(double(m)).build());
        }
        auto rho = b.value(1.1).build();
        expandable_run<2>([](auto out, auto in, auto rho) {
            return execute_parallel().stage(reference_specs::advection_pdbott_prepare_tracers::prepare_tracers(), out, in, rho);
        },
            gfx950<>(), make_grid(ni, nj, nk), out, in, rho);
        bool ok = true;
        for (int m = 0; m < 5 && ok; ++m) {
            auto v = out[m]->const_host_view();
            ok = v(3, 4, 5) == 1.1 * m && v(ni - 1, nj - 1, nk - 1) == 1.1 * m;
        }
        report("expandable_run of a user spec x5 (1.1 * i)", ok);
        // a field too small for the extents the spec reads is rejected before launch
        bool threw = false;
        try {
            auto small = b.build(), res = b.build();
            run_single_stage(reference_specs::laplacian::lap(), gfx950<>(), make_grid(ni, nj, nk), res, small);
        } catch (std::runtime_error const &e) {
            threw = std::strstr(e.what(), "too small") != nullptr;
        }
        report("field bounds check of the extent analysis", threw);
    }
    // ------------------------------------------------------------------------------------- horizontal dependencies
    // Vertical multistages whose stages read an earlier stage's output at a horizontal offset (the reference marks
    // such stages need_sync, stencil/core/need_sync.hpp:21-56, and its gpu backend synchronises the block there,
    // stencil/gpu/make_kernel_fun.hpp:32-40).  Expected values: plain host loops in the reference's order (level by
    // level in the sweep's direction, each stage over its whole extended box), the same operations in the same
    // order, so FP64 matches bit for bit.
    namespace hdep {
        using full_t = axis<1>::full_interval;
        // tmp = in + tmp(k - 1) / 2, from in at the first level
        struct accumulate {
            using tmp = inout_accessor<0, extent<0, 0, 0, 0, -1, 0>>;
            using in = in_accessor<1>;
            using param_list = make_param_list<tmp, in>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<1, 0>) {
                eval(tmp()) = eval(in()) + 0.5 * eval(tmp{0, 0, -1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::first_level) {
                eval(tmp()) = eval(in());
            }
        };
        // out = the mean of tmp's four neighbours, less half of tmp at (i + 1, k - 1)
        struct spread {
            using out = inout_accessor<0>;
            using tmp = in_accessor<1, extent<-1, 1, -1, 1, -1, 0>>;
            using param_list = make_param_list<out, tmp>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<1, 0>) {
                eval(out()) = 0.25 * (eval(tmp{-1, 0, 0}) + eval(tmp{1, 0, 0}) + eval(tmp{0, -1, 0}) + eval(tmp{0, 1, 0})) -
                              0.5 * eval(tmp{1, 0, -1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::first_level) {
                eval(out()) = 0.25 * (eval(tmp{-1, 0, 0}) + eval(tmp{1, 0, 0}) + eval(tmp{0, -1, 0}) + eval(tmp{0, 1, 0}));
            }
        };
        // backward: mid = in + mid(k + 1) / 2, from in at the top
        struct descend {
            using mid = inout_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
            using in = in_accessor<1>;
            using param_list = make_param_list<mid, in>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<0, -1>) {
                eval(mid()) = eval(in()) + 0.5 * eval(mid{0, 0, 1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::last_level) {
                eval(mid()) = eval(in());
            }
        };
        // out = mid(i + 1) - mid(i - 1) + mid(j + 1, k + 1) / 4
        struct gradient {
            using out = inout_accessor<0>;
            using mid = in_accessor<1, extent<-1, 1, 0, 1, 0, 1>>;
            using param_list = make_param_list<out, mid>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<0, -1>) {
                eval(out()) = eval(mid{1, 0, 0}) - eval(mid{-1, 0, 0}) + 0.25 * eval(mid{0, 1, 1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::last_level) {
                eval(out()) = eval(mid{1, 0, 0}) - eval(mid{-1, 0, 0});
            }
        };
        struct twice {
            using t = inout_accessor<0>;
            using in = in_accessor<1>;
            using param_list = make_param_list<t, in>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval) {
                eval(t()) = 2.0 * eval(in());
            }
        };
        // acc = the sum of t's four neighbours + acc(k - 1) / 2 + w(i + 1, j + 1)
        struct gather {
            using acc = inout_accessor<0, extent<0, 0, 0, 0, -1, 0>>;
            using t = in_accessor<1, extent<-1, 1, -1, 1>>;
            using w = in_accessor<2, extent<0, 1, 0, 1>>;
            using param_list = make_param_list<acc, t, w>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<1, 0>) {
                eval(acc()) = eval(t{1, 0, 0}) + eval(t{-1, 0, 0}) + eval(t{0, 1, 0}) + eval(t{0, -1, 0}) +
                              0.5 * eval(acc{0, 0, -1}) + eval(w{1, 1, 0});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::first_level) {
                eval(acc()) = eval(t{1, 0, 0}) + eval(t{-1, 0, 0}) + eval(t{0, 1, 0}) + eval(t{0, -1, 0}) + eval(w{1, 1, 0});
            }
        };
        // reads tmp one level ahead of the sweep at a neighbour: rejected
        struct peek_ahead {
            using out = inout_accessor<0>;
            using tmp = in_accessor<1, extent<0, 1, 0, 0, 0, 1>>;
            using param_list = make_param_list<out, tmp>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<0, -1>) {
                eval(out()) = eval(tmp{1, 0, 1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::last_level) {
                eval(out()) = eval(tmp{1, 0, 0});
            }
        };
        // on the extended box the next stage needs, reads out at k - 1, which the next stage writes on its own
        // columns only: rejected
        struct lagged {
            using tmp = inout_accessor<0>;
            using out = in_accessor<1, extent<0, 0, 0, 0, -1, 0>>;
            using param_list = make_param_list<tmp, out>;
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::modify<1, 0>) {
                eval(tmp()) = eval(out{0, 0, -1});
            }
            template <class Evaluation>
            GT_FUNCTION static void apply(Evaluation eval, full_t::first_level) {
                eval(tmp()) = 0.0;
            }
        };
    } // namespace hdep

    // the rule that lets two consecutive vertical multistages share a launch, on hand-made access lists
    void mss_fusion_rule() {
        using engine_::access_rt;
        using engine_::ext_rt;
        using engine_::stage_rt;
        const ext_rt at0{0, 0, 0, 0, 0, 0, true}, around{-1, 1, -1, 1, 0, 0, true}, below{0, 0, 0, 0, -1, 0, true};
        auto st = [](std::vector<access_rt> a, std::vector<access_rt> b) {
            return std::vector<stage_rt>{stage_rt{0, true, std::move(a), {}}, stage_rt{1, true, std::move(b), {}}};
        };
        auto fusable = [](std::vector<stage_rt> const &v) { return engine_::vpair_fusable(v, 0, 1, 1, 1); };
        // key 1 written by the first, read by the second at the column (and at k - 1): fused
        const bool col = fusable(st({{1, true, at0, 8}, {2, false, around, 8}}, {{1, false, below, 8}, {3, true, at0, 8}}));
        // the second reads key 1 at +-1: not fused
        const bool off = fusable(st({{1, true, at0, 8}}, {{1, false, around, 8}, {3, true, at0, 8}}));
        // the first reads key 3 at +-1, which the second writes: not fused
        const bool war = fusable(st({{3, false, around, 8}, {1, true, at0, 8}}, {{3, true, at0, 8}}));
        // a field both only read, at offsets: fused
        const bool ro = fusable(st({{2, false, around, 8}, {1, true, at0, 8}}, {{2, false, around, 8}, {3, true, at0, 8}}));
        report("multistage fusion rule (column data fused; offsets on passed data, either way, refused)",
            col && !off && !war && ro);
    }

    void horizontal_dependencies(int ni, int nj, int nk) {
        const int h = 2, n0 = ni + 2 * h, n1 = nj + 2 * h;
        auto g = halo_grid(ni, nj, nk, h);
        auto in = seeded<double>(n0, n1, nk, h, 91, 1, -1, 1);
        auto w = seeded<double>(n0, n1, nk, h, 91, 2, -1, 1);
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(n0, n1, nk).halos(h, h, 0);
        const std::vector<double> vin = mirror(in), vw = mirror(w);
        const auto lay = in->const_host_view();
        const int64_t s1 = lay.stride[1], s2 = lay.stride[2];
        // grid coordinates (origin at the first interior point) to the mirror's index
        auto at = [&](int i, int j, int k) { return (size_t)((i + h) + (j + h) * s1 + k * s2); };
        auto same = [&](const char *name, store<double> const &ds, std::vector<double> const &e) {
            const std::vector<double> got = mirror(ds);
            size_t bad = 0, first = 0;
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < n1; ++j)
                    for (int i = 0; i < n0; ++i) {
                        const size_t n = (size_t)(i + j * s1 + k * s2); // row padding excluded
                        if (std::memcmp(&got[n], &e[n], sizeof(double)) != 0 && bad++ == 0)
                            first = n;
                    }
            if (bad)
                std::printf("  %s: %zu values differ, first at %zu: %.17g, expected %.17g\n", name, bad, first, got[first],
                    e[first]);
            return bad == 0;
        };
        char label[160];
        // forward: a recurrence in a temporary on the box the next stage reads at +-1 in i and j
        {
            auto out = b.value(-7.).build();
            run([](auto out, auto in) {
                GT_DECLARE_TMP(double, tmp);
                return execute_forward().stage(hdep::accumulate(), tmp, in).stage(hdep::spread(), out, tmp);
            },
                gfx950<>(), g, out, in);
            std::vector<double> tmp(vin.size(), 0.0), e(vin.size(), -7.);
            for (int k = 0; k < nk; ++k) {
                for (int j = -1; j < nj + 1; ++j)
                    for (int i = -1; i < ni + 1; ++i)
                        tmp[at(i, j, k)] = k == 0 ? vin[at(i, j, k)] : vin[at(i, j, k)] + 0.5 * tmp[at(i, j, k - 1)];
                for (int j = 0; j < nj; ++j)
                    for (int i = 0; i < ni; ++i) {
                        const double m = 0.25 * (tmp[at(i - 1, j, k)] + tmp[at(i + 1, j, k)] + tmp[at(i, j - 1, k)] +
                                                    tmp[at(i, j + 1, k)]);
                        e[at(i, j, k)] = k == 0 ? m : m - 0.5 * tmp[at(i + 1, j, k - 1)];
                    }
            }
            std::snprintf(label, sizeof label,
                "vertical multistage, forward: temporary read at +-1 in i and j by the next stage %dx%dx%d", ni, nj, nk);
            report(label, same("out", out, e));
        }
        // the same recurrence and spread in two vertical multistages: the second reads the first's temporary at +-1
        // in i and j, so the two must not share a launch (engine_::vpair_fusable refuses the pair)
        {
            auto out = b.value(-7.).build();
            run([](auto out, auto in) {
                GT_DECLARE_TMP(double, tmp);
                return multi_pass(execute_forward().stage(hdep::accumulate(), tmp, in),
                    execute_backward().stage(hdep::spread(), out, tmp));
            },
                gfx950<>(), g, out, in);
            std::vector<double> tmp(vin.size(), 0.0), e(vin.size(), -7.);
            for (int k = 0; k < nk; ++k) {
                for (int j = -1; j < nj + 1; ++j)
                    for (int i = -1; i < ni + 1; ++i)
                        tmp[at(i, j, k)] = k == 0 ? vin[at(i, j, k)] : vin[at(i, j, k)] + 0.5 * tmp[at(i, j, k - 1)];
            }
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < nj; ++j)
                    for (int i = 0; i < ni; ++i) {
                        const double m = 0.25 * (tmp[at(i - 1, j, k)] + tmp[at(i + 1, j, k)] + tmp[at(i, j - 1, k)] +
                                                    tmp[at(i, j + 1, k)]);
                        e[at(i, j, k)] = k == 0 ? m : m - 0.5 * tmp[at(i + 1, j, k - 1)];
                    }
            std::snprintf(label, sizeof label,
                "two vertical multistages, the second reading the first's temporary at +-1 (not fused) %dx%dx%d", ni, nj,
                nk);
            report(label, same("out", out, e));
        }
        // backward: a user field written on the extended box (its halo) and read at offsets and at k + 1
        {
            auto mid = b.value(3.).build(), out = b.value(-7.).build();
            run([](auto out, auto mid, auto in) {
                return execute_backward().stage(hdep::descend(), mid, in).stage(hdep::gradient(), out, mid);
            },
                gfx950<>(), g, out, mid, in);
            std::vector<double> em(vin.size(), 3.), e(vin.size(), -7.);
            for (int k = nk - 1; k >= 0; --k) {
                for (int j = 0; j < nj + 1; ++j)
                    for (int i = -1; i < ni + 1; ++i)
                        em[at(i, j, k)] = k == nk - 1 ? vin[at(i, j, k)] : vin[at(i, j, k)] + 0.5 * em[at(i, j, k + 1)];
                for (int j = 0; j < nj; ++j)
                    for (int i = 0; i < ni; ++i) {
                        const double d = em[at(i + 1, j, k)] - em[at(i - 1, j, k)];
                        e[at(i, j, k)] = k == nk - 1 ? d : d + 0.25 * em[at(i, j + 1, k + 1)];
                    }
            }
            std::snprintf(label, sizeof label,
                "vertical multistage, backward: field written on its halo, read at offsets and k + 1 %dx%dx%d", ni, nj,
                nk);
            report(label, same("mid", mid, em) && same("out", out, e));
        }
        // forward: an ij_cached temporary (LDS) and a staged read-only field, with a recurrence in the output
        {
            auto acc = b.value(-7.).build();
            run([](auto acc, auto in, auto w) {
                GT_DECLARE_TMP(double, t);
                return execute_forward().ij_cached(t).stage(hdep::twice(), t, in).stage(hdep::gather(), acc, t, w);
            },
                gfx950<>(), g, acc, in, w);
            std::vector<double> t(vin.size(), 0.0), e(vin.size(), -7.);
            for (int k = 0; k < nk; ++k) {
                for (int j = -1; j < nj + 1; ++j)
                    for (int i = -1; i < ni + 1; ++i)
                        t[at(i, j, k)] = 2.0 * vin[at(i, j, k)];
                for (int j = 0; j < nj; ++j)
                    for (int i = 0; i < ni; ++i) {
                        double v = t[at(i + 1, j, k)] + t[at(i - 1, j, k)] + t[at(i, j + 1, k)] + t[at(i, j - 1, k)];
                        if (k > 0)
                            v = v + 0.5 * e[at(i, j, k - 1)];
                        e[at(i, j, k)] = v + vw[at(i + 1, j + 1, k)];
                    }
            }
            std::snprintf(label, sizeof label,
                "vertical multistage, forward: ij_cached temporary and staged field at offsets %dx%dx%d", ni, nj, nk);
            report(label, same("acc", acc, e));
        }
        // what the tiles cannot run exactly is rejected before any launch
        auto rejects = [&](const char *name, auto spec, auto const &... fields) {
            bool threw = false;
            try {
                run(spec, gfx950<>(), g, fields...);
            } catch (std::runtime_error const &x) {
                threw = std::strstr(x.what(), "horizontal dependencies") != nullptr;
                if (!threw)
                    std::printf("  unexpected message: %s\n", x.what());
            }
            report(name, threw);
        };
        if (ni == 70) {
            auto out = b.build();
            rejects("vertical multistage: a neighbour's value one level ahead of the sweep is rejected",
                [](auto out, auto in) {
                    GT_DECLARE_TMP(double, tmp);
                    return execute_forward().stage(hdep::accumulate(), tmp, in).stage(hdep::peek_ahead(), out, tmp);
                },
                out, in);
            rejects("vertical multistage: an extended box reading a level behind that only its own columns wrote is "
                    "rejected",
                [](auto out) {
                    GT_DECLARE_TMP(double, tmp);
                    return execute_forward().stage(hdep::lagged(), tmp, out).stage(hdep::spread(), out, tmp);
                },
                out);
        }
    }

    // ------------------------------------------------------------------------------------- k-caches, axis<N>
    // the unit tests' fields (test_kcache_*.cpp: in(i, j, k) = i + j + k + 1 on a 6 x 6 grid, axis<3> intervals of 2, 6
    // and 2 levels, no halo) and their expected columns
    double kin(int i, int j, int k) { return i + j + k + 1; }

    template <bool WithOut, class Spec>
    bool kcache_case(const char *name, Spec spec, std::function<double(int, int, int, int)> expected) {
        using axis_t = reference_specs::test_kcache_fill::axis_t;
        const int d0 = 6, d1 = 6;
        auto g = make_grid(d0, d1, axis_t(2, 6, 2));
        const int nk = g.k_size;
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(d0, d1, nk);
        auto in = b.initializer(kin).build();
        auto out = b.build();
        if constexpr (WithOut)
            run(spec, gfx950<>(), g, in, out);
        else
            run(spec, gfx950<>(), g, in);
        auto v = (WithOut ? out : in)->const_host_view();
        bool ok = true;
        for (int i = 0; i < d0 && ok; ++i)
            for (int j = 0; j < d1 && ok; ++j)
                for (int k = 0; k < nk && ok; ++k) {
                    const double e = expected(i, j, k, nk);
                    if (v(i, j, k) != e) {
                        std::printf("  (%d,%d,%d): %.17g, expected %.17g\n", i, j, k, v(i, j, k), e);
                        ok = false;
                    }
                }
        report(name, ok);
        return ok;
    }

    void kcache_tests() {
        namespace kf = reference_specs::test_kcache_fill;
        namespace kfl = reference_specs::test_kcache_flush;
        namespace kff = reference_specs::test_kcache_fill_and_flush;
        namespace kl = reference_specs::test_kcache_local;
        // test_kcache_fill.cpp:52-125
        kcache_case<true>("test_kcache_fill forward (axis<3>, fill policy, k offsets -1..1)",
            [](auto in, auto out) {
                return execute_forward().k_cached(cache_io_policy::fill(), in).stage(kf::shift_acc_forward_fill(), in, out);
            },
            [](int i, int j, int k, int nk) {
                return k == 0        ? kin(i, j, 0) + kin(i, j, 1)
                       : k == nk - 1 ? kin(i, j, nk - 1) + kin(i, j, nk - 2)
                                     : kin(i, j, k - 1) + kin(i, j, k) + kin(i, j, k + 1);
            });
        kcache_case<true>("test_kcache_fill backward",
            [](auto in, auto out) {
                return execute_backward().k_cached(cache_io_policy::fill(), in).stage(kf::shift_acc_backward_fill(), in, out);
            },
            [](int i, int j, int k, int nk) {
                return k == nk - 1 ? kin(i, j, nk - 1) + kin(i, j, nk - 2)
                       : k == 0    ? kin(i, j, 1) + kin(i, j, 0)
                                   : kin(i, j, k + 1) + kin(i, j, k) + kin(i, j, k - 1);
            });
        kcache_case<true>("test_kcache_fill fill_copy_forward",
            [](auto in, auto out) {
                return execute_forward().k_cached(cache_io_policy::fill(), in).stage(kf::copy_fill(), in, out);
            },
            [](int i, int j, int k, int) { return kin(i, j, k); });
        // test_kcache_flush.cpp:38-99: a running sum kept in the cache and flushed to `out`
        auto fwd_sum = [](int i, int j, int k, int) {
            double r = 0;
            for (int kk = 0; kk <= k; ++kk)
                r = kk == 0 ? kin(i, j, 0) : r + kin(i, j, kk);
            return r;
        };
        auto bwd_sum = [](int i, int j, int k, int nk) {
            double r = 0;
            for (int kk = nk - 1; kk >= k; --kk)
                r = kk == nk - 1 ? kin(i, j, nk - 1) : r + kin(i, j, kk);
            return r;
        };
        kcache_case<true>("test_kcache_flush forward (flush policy)",
            [](auto in, auto out) {
                return execute_forward().k_cached(cache_io_policy::flush(), out).stage(kfl::shift_acc_forward_flush(), in, out);
            },
            fwd_sum);
        kcache_case<true>("test_kcache_flush backward",
            [](auto in, auto out) {
                return execute_backward()
                    .k_cached(cache_io_policy::flush(), out)
                    .stage(kfl::shift_acc_backward_flush(), in, out);
            },
            bwd_sum);
        // test_kcache_fill_and_flush.cpp:38-125: the cached field is read and written in place
        kcache_case<false>("test_kcache_fill_and_flush forward (fill + flush on one inout field)",
            [](auto in) {
                return execute_forward()
                    .k_cached(cache_io_policy::fill(), cache_io_policy::flush(), in)
                    .stage(kff::shift_acc_forward_fill_and_flush(), in);
            },
            [](int i, int j, int k, int) {
                double r = 0;
                for (int kk = 0; kk <= k; ++kk)
                    r = kk == 0 ? kin(i, j, 0) : kin(i, j, kk) + r;
                return r;
            });
        kcache_case<false>("test_kcache_fill_and_flush backward",
            [](auto in) {
                return execute_backward()
                    .k_cached(cache_io_policy::fill(), cache_io_policy::flush(), in)
                    .stage(kff::shift_acc_backward_fill_and_flush(), in);
            },
            [](int i, int j, int k, int nk) {
                double r = 0;
                for (int kk = nk - 1; kk >= k; --kk)
                    r = kk == nk - 1 ? kin(i, j, nk - 1) : r + kin(i, j, kk);
                return r;
            });
        kcache_case<false>("test_kcache_fill_and_flush copy_forward",
            [](auto in) {
                return execute_forward()
                    .k_cached(cache_io_policy::fill(), cache_io_policy::flush(), in)
                    .stage(kff::copy_fill(), in);
            },
            [](int i, int j, int k, int) { return kin(i, j, k); });
        kcache_case<false>("test_kcache_fill_and_flush scale_forward",
            [](auto in) {
                return execute_forward()
                    .k_cached(cache_io_policy::fill(), cache_io_policy::flush(), in)
                    .stage(kff::scale_fill(), in);
            },
            [](int i, int j, int k, int) { return 2 * kin(i, j, k); });
        // test_kcache_local.cpp:38-213: a k-cached temporary with no policy (registers only)
        kcache_case<true>("test_kcache_local forward (policy-free k-cached temporary)",
            [](auto in, auto out) {
                GT_DECLARE_TMP(double, tmp);
                return execute_forward().k_cached(tmp).stage(kl::shift_acc_forward(), in, out, tmp);
            },
            fwd_sum);
        kcache_case<true>("test_kcache_local backward",
            [](auto in, auto out) {
                GT_DECLARE_TMP(double, tmp);
                return execute_backward().k_cached(tmp).stage(kl::shift_acc_backward(), in, out, tmp);
            },
            bwd_sum);
        kcache_case<true>("test_kcache_local biside_forward (window -2..1, writes at k+1, first_level::shift<1>)",
            [](auto in, auto out) {
                GT_DECLARE_TMP(double, tmp);
                return execute_forward().k_cached(tmp).stage(kl::biside_large_kcache_forward(), in, out, tmp);
            },
            [](int i, int j, int k, int nk) {
                // test_kcache_local.cpp:136-158: the reference's host loop over a buffer column
                std::vector<double> buff(nk + 1), ref(nk);
                buff[0] = kin(i, j, 0);
                buff[1] = kin(i, j, 0) / 2;
                ref[0] = kin(i, j, 0);
                buff[2] = kin(i, j, 1) / 2;
                ref[1] = buff[1] + buff[0] / 4;
                for (int kk = 2; kk < nk; ++kk) {
                    if (kk != nk - 1)
                        buff[kk + 1] = kin(i, j, kk) / 2;
                    ref[kk] = buff[kk] + buff[kk - 1] / 4 + .12 * buff[kk - 2];
                }
                return ref[k];
            });
        kcache_case<true>("test_kcache_local biside_backward (window -1..2, last_level::shift<-1>)",
            [](auto in, auto out) {
                GT_DECLARE_TMP(double, tmp);
                return execute_backward().k_cached(tmp).stage(kl::biside_large_kcache_backward(), in, out, tmp);
            },
            [](int i, int j, int k, int nk) {
                // test_kcache_local.cpp:190-212
                std::vector<double> buff(nk + 1), ref(nk);
                buff[nk - 1] = kin(i, j, nk - 1);
                buff[nk - 2] = kin(i, j, nk - 1) / 2;
                ref[nk - 1] = kin(i, j, nk - 1);
                buff[nk - 3] = kin(i, j, nk - 2) / 2;
                ref[nk - 2] = buff[nk - 2] + buff[nk - 1] / 4;
                for (int kk = nk - 3; kk >= 0; --kk) {
                    if (kk != 0)
                        buff[kk - 1] = kin(i, j, kk) / 2;
                    ref[kk] = buff[kk] + buff[kk + 1] / 4 + .12 * buff[kk + 2];
                }
                return ref[k];
            });
    }

    // test_kparallel.cpp:38-68: parallel multistages over axis<2> / axis<3>, interval overloads per k
    void kparallel_tests() {
        namespace kp = reference_specs::test_kparallel;
        auto in = [](int i, int j, int k) { return i * 1000 + j * 100 + k; };
        {
            auto g = make_grid(7, 8, axis<2>(14, 16));
            auto b = storage::builder<storage::gpu>.type<double>().dimensions(7, 8, g.k_size);
            auto out = b.value(1.5).build();
            run([](auto in, auto out) {
                GT_DECLARE_TMP(double, tmp);
                return execute_parallel().stage(kp::parallel_functor(), in, tmp).stage(kp::parallel_functor(), tmp, out);
            },
                gfx950<>(), g, b.initializer(in).build(), out);
            auto v = out->const_host_view();
            bool ok = true;
            for (int i = 0; i < 7; ++i)
                for (int j = 0; j < 8; ++j)
                    for (int k = 0; k < g.k_size; ++k)
                        ok = ok && v(i, j, k) == in(i, j, k) * (k < 14 ? 1 : 4);
            report("test_kparallel with_temporary (axis<2>: 14 + 16 levels)", ok);
        }
        {
            auto g = make_grid(7, 8, axis<3>(14, 16, 18));
            auto b = storage::builder<storage::gpu>.type<double>().dimensions(7, 8, g.k_size);
            auto out = b.value(1.5).build();
            run_single_stage(kp::parallel_functor_on_upper_interval(), gfx950<>(), g, b.initializer(in).build(), out);
            auto v = out->const_host_view();
            bool ok = true;
            for (int i = 0; i < 7; ++i)
                for (int j = 0; j < 8; ++j)
                    for (int k = 0; k < g.k_size; ++k)
                        ok = ok && v(i, j, k) == (k >= 14 && k < 14 + 16 ? in(i, j, k) : 1.5);
            report("test_kparallel with_unused_intervals (axis<3>: only interval 1 written)", ok);
        }
    }

    // whole_axis_access.cpp:40-60: the k dimension renamed to data dimension 3, summed over the column with a
    // four-dimensional accessor
    void whole_axis_access(int ni, int nj, int nk) {
        using namespace literals;
        auto in = [](int i, int j, int k) { return i + j + k; };
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk);
        auto out = b.build();
        run_single_stage(reference_specs::whole_axis_access::functor(), gfx950<>(), make_grid(ni, nj, nk), out,
            sid::rename_dimensions<dim::k, decltype(3_c)>(b.initializer(in).build()), positional<dim::k>());
        auto v = out->const_host_view();
        bool ok = true;
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j)
                for (int k = 0; k < nk; ++k) {
                    int r = 0;
                    for (int kk = 0; kk < k; ++kk)
                        r += in(i, j, kk);
                    ok = ok && v(i, j, k) == r;
                }
        char name[96];
        std::snprintf(name, sizeof name, "whole_axis_access (rename_dimensions<k, 3>, 4-D accessor) %dx%dx%d", ni, nj, nk);
        report(name, ok);
    }

    // ------------------------------------------------------------------------------------- precision policy
    // gfx950<GT355_WIDEN>: float storages, read-only fields widened to double, temporaries (ccol, dcol, data_col, lap,
    // flx, fly) held in double, stores narrowed -- the oracle's mixed variants bit for bit
    void precision_policy(int ni, int nj, int nk) {
        {
            const int h = 2;
            auto in = seeded<float>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
            auto coeff = seeded<float>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
            auto out = storage::builder<storage::gpu>.type<float>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0).build();
            auto g = halo_grid(ni, nj, nk, h);
            run(reference_specs::horizontal_diffusion::spec<float>(), gfx950<GT355_WIDEN>(), g, in, coeff, out);
            auto hin = mirror(in), hco = mirror(coeff);
            std::vector<float> ref(hin.size(), 0.f);
            auto st = in->strides();
            const size_t o = h + h * st[1];
            gto_hdiff_mixed(ref.data() + o, hin.data() + o, hco.data() + o, ni, nj, nk, st[1], st[2], 4);
            char name[112];
            std::snprintf(name, sizeof name, "hdiff spec, precision policy (float storage, double compute) %dx%dx%d", ni,
                nj, nk);
            report(name, same_interior(out, ref, g));
        }
        {
            const int h = 3;
            auto utens_stage = seeded<float>(ni + 6, nj + 6, nk, h, 42, 6, -1, 1);
            auto u_stage = seeded<float>(ni + 6, nj + 6, nk, h, 42, 3, -1, 1);
            auto u_pos = seeded<float>(ni + 6, nj + 6, nk, h, 42, 4, -1, 1);
            auto utens = seeded<float>(ni + 6, nj + 6, nk, h, 42, 5, -1, 1);
            auto wcon = seeded<float>(ni + 6, nj + 6, nk + 1, h, 42, 7, -0.5, 0.5);
            auto g = halo_grid(ni, nj, nk, h);
            auto uts0 = mirror(utens_stage);
            run(reference_specs::vertical_advection_dycore::spec<float>(), gfx950<GT355_WIDEN>(), g, utens_stage,
                u_stage, wcon, u_pos, utens, global_parameter(0.15f));
            auto hus = mirror(u_stage), hup = mirror(u_pos), hut = mirror(utens), hw = mirror(wcon);
            auto st = u_stage->strides(), wst = wcon->strides();
            const size_t o = h + h * st[1], wo = h + h * wst[1];
            gto_vadv_mixed(uts0.data() + o, hus.data() + o, hw.data() + wo, hup.data() + o, hut.data() + o, 0.15f, ni,
                nj, nk, st[1], st[2], wst[1], wst[2], 4);
            char name[112];
            std::snprintf(name, sizeof name, "vadv spec, precision policy (float storage, double compute) %dx%dx%d", ni,
                nj, nk);
            report(name, same_interior(utens_stage, uts0, g));
        }
    }

    // ------------------------------------------------------------------------------------- temporaries on two streams
    // the hdiff spec with HBM temporaries on two streams at once, several rounds: a temporary released on one stream
    // is not handed to the other while the first stream's kernels may still use it
    void two_streams(int ni, int nj, int nk) {
        const int h = 2;
        hipStream_t s[2];
        for (auto &x : s)
            check(hipStreamCreateWithFlags(&x, hipStreamNonBlocking) == hipSuccess ? 0 : -2, "stream");
        std::vector<store<double>> in, coeff, out;
        for (int n = 0; n < 2; ++n) {
            in.push_back(seeded<double>(ni + 4, nj + 4, nk, h, 100 + n, 1, -1, 1));
            coeff.push_back(seeded<double>(ni + 4, nj + 4, nk, h, 100 + n, 2, 0, 0.25));
            out.push_back(storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0).build());
        }
        auto g = halo_grid(ni, nj, nk, h);
        auto spec = [](auto in, auto coeff, auto out) {
            using namespace reference_specs::horizontal_diffusion;
            GT_DECLARE_TMP(double, lap, flx, fly);
            return execute_parallel()
                .stage(lap_function(), lap, in)
                .stage(flx_function(), flx, in, lap)
                .stage(fly_function(), fly, in, lap)
                .stage(out_function(), out, in, flx, fly, coeff);
        };
        // bring the inputs to the device before the streams run
        for (int n = 0; n < 2; ++n) {
            in[n]->get_const_target_ptr();
            coeff[n]->get_const_target_ptr();
            out[n]->get_target_ptr();
        }
        check(gt355_synchronize(), "sync");
        for (int round = 0; round < 6; ++round)
            for (int n = 0; n < 2; ++n) {
                gfx950<> be;
                be.stream = s[n];
                run(spec, be, g, in[n], coeff[n], out[n]);
            }
        check(gt355_synchronize(), "sync");
        bool ok = true;
        for (int n = 0; n < 2; ++n) {
            auto hin = mirror(in[n]), hco = mirror(coeff[n]);
            std::vector<double> ref(hin.size(), 0.);
            auto st = in[n]->strides();
            const size_t o = h + h * st[1];
            gto_hdiff_f64(ref.data() + o, hin.data() + o, hco.data() + o, ni, nj, nk, st[1], st[2], 4);
            ok = ok && same_interior(out[n], ref, g);
        }
        for (auto &x : s)
            (void)hipStreamDestroy(x);
        report("hdiff spec with HBM temporaries on two streams at once", ok);
    }

    // ------------------------------------------------------------------------------------- fast path
    // the reference's hdiff and vadv specs are recognised and run on the hand-written kernels (FP64, native FP32 and
    // the precision policy); a spec that differs in one operation is not, and runs on the generic engine
    struct flx_strict { // flx_function with the limiter's comparison >= instead of > (this project's variant)
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<0, 1, 0, 0>>;
        using lap = in_accessor<2, extent<0, 1, 0, 0>>;
        using param_list = make_param_list<out, in, lap>;
        template <typename Evaluation>
        GT_FUNCTION static void apply(Evaluation eval) {
            auto res = eval(lap(1, 0)) - eval(lap(0, 0));
            eval(out()) = res * (eval(in(1, 0)) - eval(in(0, 0))) >= 0 ? 0 : res;
        }
    };
    struct backward_shifted { // u_backward_function with u_pos added instead of subtracted (this project's variant)
        using utens_stage = inout_accessor<0>;
        using u_pos = in_accessor<1>;
        using dtr_stage = in_accessor<2>;
        using ccol = in_accessor<3>;
        using dcol = in_accessor<4>;
        using data_col = inout_accessor<5, extent<0, 0, 0, 0, 0, 1>>;
        using param_list = make_param_list<utens_stage, u_pos, dtr_stage, ccol, dcol, data_col>;
        using full_t = reference_specs::vertical_advection_dycore::full_t;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full_t::modify<0, -1>) {
            auto data = eval(dcol()) - eval(ccol()) * eval(data_col(0, 0, 1));
            eval(utens_stage()) = eval(dtr_stage()) * (data + eval(u_pos()));
            eval(data_col()) = data;
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full_t::last_level) {
            eval(utens_stage()) = eval(dtr_stage()) * (eval(dcol()) + eval(u_pos()));
            eval(data_col()) = eval(dcol());
        }
    };

    void fast_path() {
        if (const char *e = std::getenv("GT355_ENGINE_FAST"); e && std::atoi(e) == 0) {
            report("fast path switched off (GT355_ENGINE_FAST=0): every spec on the generic engine",
                engine_::fast_runs() == 0);
            return;
        }
        const int ni = 70, nj = 33, nk = 21;
        // the reference specs: taken by the hand-written kernels, results equal the oracle
        {
            const long before = engine_::fast_runs();
            horizontal_diffusion(ni, nj, nk);
            vertical_advection(ni, nj, nk);
            precision_policy(ni, nj, nk);
            report("reference hdiff / vadv specs (FP64 and precision policy) ran on the hand-written kernels",
                engine_::fast_runs() == before + 4);
        }
        // one changed operation: the generic engine
        {
            const int h = 2;
            auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
            auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
            auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0);
            auto out = b.build();
            auto g = halo_grid(ni, nj, nk, h);
            const long before = engine_::fast_runs();
            run([](auto in, auto coeff, auto out) {
                using namespace reference_specs::horizontal_diffusion;
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .ij_cached(lap, flx, fly)
                    .stage(lap_function(), lap, in)
                    .stage(flx_strict(), flx, in, lap)
                    .stage(fly_function(), fly, in, lap)
                    .stage(out_function(), out, in, flx, fly, coeff);
            },
                gfx950<>(), g, in, coeff, out);
            const bool generic = engine_::fast_runs() == before;
            report("hdiff spec with a changed limiter comparison runs on the generic engine", generic);
        }
        {
            const int h = 3;
            auto utens_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 6, -1, 1);
            auto u_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 3, -1, 1);
            auto u_pos = seeded<double>(ni + 6, nj + 6, nk, h, 42, 4, -1, 1);
            auto utens = seeded<double>(ni + 6, nj + 6, nk, h, 42, 5, -1, 1);
            auto wcon = seeded<double>(ni + 6, nj + 6, nk + 1, h, 42, 7, -0.5, 0.5);
            auto g = halo_grid(ni, nj, nk, h);
            const long before = engine_::fast_runs();
            run([](auto utens_stage, auto u_stage, auto wcon, auto u_pos, auto utens, auto dtr_stage) {
                using namespace reference_specs::vertical_advection_dycore;
                GT_DECLARE_TMP(double, ccol, dcol, data_col);
                return multi_pass(execute_forward()
                                      .k_cached(cache_io_policy::flush(), ccol, dcol)
                                      .k_cached(cache_io_policy::fill(), u_stage)
                                      .stage(u_forward_function(), utens_stage, wcon, u_stage, u_pos, utens, dtr_stage,
                                          ccol, dcol),
                    execute_backward().k_cached(data_col).stage(
                        backward_shifted(), utens_stage, u_pos, dtr_stage, ccol, dcol, data_col));
            },
                gfx950<>(), g, utens_stage, u_stage, wcon, u_pos, utens, global_parameter(0.15));
            report("vadv spec with a changed backward stage runs on the generic engine", engine_::fast_runs() == before);
        }
        // inside a deferred-run scope the recognised hdiff spec is recorded; a generic-engine run (a copy spec) that
        // reads its output first runs the recorded call, so the copy sees the diffusion, as without the scope
        {
            const int h = 2;
            auto in = seeded<double>(ni + 4, nj + 4, nk, h, 42, 1, -1, 1);
            auto coeff = seeded<double>(ni + 4, nj + 4, nk, h, 42, 2, 0, 0.25);
            auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0);
            auto out_a = b.build(), copy_a = b.build(), out_b = b.build(), copy_b = b.build();
            auto g = halo_grid(ni, nj, nk, h);
            auto hdiff_spec = [](auto in, auto coeff, auto out) {
                using namespace reference_specs::horizontal_diffusion;
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .ij_cached(lap, flx, fly)
                    .stage(lap_function(), lap, in)
                    .stage(flx_function(), flx, in, lap)
                    .stage(fly_function(), fly, in, lap)
                    .stage(out_function(), out, in, flx, fly, coeff);
            };
            auto copy_spec = [](auto out, auto in) {
                return execute_parallel().stage(reference_specs::copy_stencil::copy_functor(), in, out);
            };
            const long before = engine_::fast_runs();
            {
                deferred scope(gfx950<>{});
                run(hdiff_spec, gfx950<>(), g, in, coeff, out_a);
                run(copy_spec, gfx950<>(), g, copy_a, out_a);
                scope.end();
            }
            run(hdiff_spec, gfx950<>(), g, in, coeff, out_b);
            run(copy_spec, gfx950<>(), g, copy_b, out_b);
            auto va = copy_a->const_host_view(), vb = copy_b->const_host_view();
            bool ok = engine_::fast_runs() == before + 2;
            if (!ok)
                std::printf("  hand-kernel runs: %ld, expected %ld\n", engine_::fast_runs() - before, 2L);
            for (int k = 0; k < nk && ok; ++k)
                for (int j = h; j < h + nj && ok; ++j)
                    for (int i = h; i < h + ni && ok; ++i) {
                        ok = std::memcmp(&va(i, j, k), &vb(i, j, k), sizeof(double)) == 0 && va(i, j, k) != 0.0;
                        if (!ok)
                            std::printf("  (%d,%d,%d): %.17g, expected %.17g\n", i, j, k, va(i, j, k), vb(i, j, k));
                    }
            report("deferred scope: a recorded hdiff spec runs before a generic-engine run that reads its output", ok);
        }
    }

    // ------------------------------------------------------------------------------------- boundary functors
    // the host restatement of the reference's CPU boundary_apply (boundaries/apply.hpp:45-135): every selected
    // region's halo points, the functor called with host views -- the expected values of the device kernel
    template <class T>
    struct host_field {
        storage::host_view<T> v;
        T &operator()(unsigned i, unsigned j, unsigned k) const { return v((int)i, (int)j, (int)k); }
    };
    template <class BC, class... F>
    void host_boundary_apply(
        std::array<halo_descriptor, 3> const &h, BC const &bc, uint32_t mask, F const &... f) {
        auto lo_n = [&](int d, int dir, int &lo, int &n) {
            if (dir == 0)
                lo = h[d].begin(), n = h[d].end() - h[d].begin() + 1;
            else if (dir < 0)
                lo = h[d].begin() - h[d].minus(), n = h[d].minus();
            else
                lo = h[d].end() + 1, n = h[d].plus();
        };
        auto region = [&](auto dir) {
            using D = decltype(dir);
            if (!(mask & boundaries::direction_bit(D::i, D::j, D::k)))
                return;
            int lo[3], n[3];
            lo_n(0, D::i, lo[0], n[0]), lo_n(1, D::j, lo[1], n[1]), lo_n(2, D::k, lo[2], n[2]);
            for (int k = lo[2]; k < lo[2] + n[2]; ++k)
                for (int j = lo[1]; j < lo[1] + n[1]; ++j)
                    for (int i = lo[0]; i < lo[0] + n[0]; ++i)
                        bc(dir, f..., (unsigned)i, (unsigned)j, (unsigned)k);
        };
        using namespace boundaries;
        auto row = [&](auto di) {
            constexpr sign I = decltype(di)::value;
            region(direction<I, minus_, minus_>()), region(direction<I, minus_, zero_>());
            region(direction<I, minus_, plus_>()), region(direction<I, zero_, minus_>());
            if constexpr (I != zero_)
                region(direction<I, zero_, zero_>());
            region(direction<I, zero_, plus_>()), region(direction<I, plus_, minus_>());
            region(direction<I, plus_, zero_>()), region(direction<I, plus_, plus_>());
        };
        row(std::integral_constant<sign, minus_>());
        row(std::integral_constant<sign, zero_>());
        row(std::integral_constant<sign, plus_>());
    }

    template <class T>
    bool same_everywhere(store<T> const &a, store<T> const &b) {
        auto va = a->const_host_view(), vb = b->const_host_view();
        auto len = a->lengths();
        for (int k = 0; k < len[2]; ++k)
            for (int j = 0; j < len[1]; ++j)
                for (int i = 0; i < len[0]; ++i)
                    if (std::memcmp(&va(i, j, k), &vb(i, j, k), sizeof(T)) != 0) {
                        std::printf("  mismatch at (%d,%d,%d)\n", i, j, k);
                        return false;
                    }
        return true;
    }

    // a predicate over directions (boundaries/predicate.hpp): the halo of the j-minus side only
    struct minus_j_only {
        template <class Direction>
        bool operator()(Direction) const {
            return Direction::j == boundaries::minus_;
        }
    };

    void boundary_functors() {
        using namespace boundaries;
        // examples/boundaries/boundaries.cpp:121-190: int fields with a 1-point halo in every dimension
        {
            const int d1 = 13, d2 = 11, d3 = 9;
            auto b = storage::builder<storage::gpu>.type<int>().dimensions(d1, d2, d3).halos(1, 1, 1);
            auto in_s = b.name("in").initializer([](int i, int j, int k) { return i + j + k; }).build();
            auto out_s = b.name("out").value(0).build();
            std::array<halo_descriptor, 3> halos = {halo_descriptor(1, 1, 1, d1 - 2, d1),
                halo_descriptor(1, 1, 1, d2 - 2, d2), halo_descriptor(1, 1, 1, d3 - 2, d3)};
            using bc_t = reference_specs::examples_boundaries::direction_bc_input<unsigned>;
            make_boundary<storage::gpu>(halos, bc_t(42)).apply(out_s, in_s);
            auto in = in_s->const_host_view();
            auto out = out_s->const_host_view();
            // the example's own checks
            bool ok = out(0, 0, 0) == in(1, 1, 1);
            for (int k = 1; k < d3; ++k)
                ok = ok && out(0, 0, k) == in(1, 1, k);
            for (int i = 1; i < d1; ++i)
                for (int k = 0; k < d3; ++k)
                    ok = ok && out(i, 0, k) == in(i, 1, k);
            for (int i = 0; i < d1; ++i)
                for (int j = 0; j < d2; ++j)
                    for (int k = 0; k < d3; ++k)
                        if (((i == 0 || i == d1 - 1) && j > 0) || (j > 0 && (k == 0 || k == d3 - 1)))
                            ok = ok && out(i, j, k) == 42;
            // every point against the host apply of the same functor
            auto ref = b.value(0).build();
            host_boundary_apply(halos, bc_t(42), all_directions, host_field<int>{ref->host_view()},
                host_field<int>{in_s->host_view()});
            ok = ok && same_everywhere(out_s, ref);
            report("boundary functor of examples/boundaries (int fields, direction overloads)", ok);
        }
        // tests/regression/boundary_conditions.cpp:61-143 with halo 3 in every dimension
        auto bc_case = [&](auto float_tag, int n0, int n1, int n2) {
            using F = decltype(float_tag);
            using namespace reference_specs::boundary_conditions;
            const int h = 3;
            auto b = storage::builder<storage::gpu>.type<F>().dimensions(n0 + 2 * h, n1 + 2 * h, n2 + 2 * h).halos(h, h, h);
            auto src = b.initializer([](int i, int j, int k) { return F(i + j + k); }).build();
            auto dst = b.value(F(0)).build();
            auto lengths = src->lengths();
            std::array<halo_descriptor, 3> halos;
            for (int d = 0; d < 3; ++d)
                halos[d] = halo_descriptor(h, h, h, lengths[d] - h - 1, lengths[d]);
            make_boundary<storage::gpu>(halos, direction_bc_input<F>()).apply(src, dst);
            // verify_result (boundary_conditions.cpp:61-118), as known answers
            auto sv = src->const_host_view();
            auto dv = dst->const_host_view();
            bool ok = true;
            for (int i = 0; i < lengths[0]; ++i)
                for (int j = 0; j < lengths[1]; ++j)
                    for (int k = 0; k < lengths[2]; ++k) {
                        const bool ii = i >= h && i <= lengths[0] - h - 1, ij = j >= h && j <= lengths[1] - h - 1,
                                   ik = k >= h && k <= lengths[2] - h - 1;
                        F e;
                        if (ii && ij && ik)
                            e = 0;
                        else if (i < h && j < h && k < h)
                            e = factor<F> * corner_value<F>;
                        else if (i < h && j < h)
                            e = factor<F> * edge_value<F>;
                        else if (j < h)
                            e = factor<F> * face_value<F>;
                        else
                            e = factor<F> * sv(i, j, k);
                        ok = ok && dv(i, j, k) == e;
                    }
            auto ref = b.value(F(0)).build();
            host_boundary_apply(halos, direction_bc_input<F>(), all_directions, host_field<F>{src->host_view()},
                host_field<F>{ref->host_view()});
            ok = ok && same_everywhere(dst, ref);
            char name[112];
            std::snprintf(name, sizeof name, "boundary_conditions functor (%s, halo 3 in i, j, k) %dx%dx%d",
                sizeof(F) == 8 ? "double" : "float", n0, n1, n2);
            report(name, ok);
        };
        bc_case(double(), 12, 33, 61);
        bc_case(float(), 23, 11, 43);
        // a callable predicate: only the j-minus side's regions are written
        {
            const int d1 = 40, d2 = 20, d3 = 10;
            auto b = storage::builder<storage::gpu>.type<double>().dimensions(d1, d2, d3).halos(2, 2, 1);
            auto in_s = b.initializer([](int i, int j, int k) { return 0.5 * i + j - 0.25 * k; }).build();
            auto out_s = b.value(-3.).build();
            std::array<halo_descriptor, 3> halos = {halo_descriptor(2, 2, 2, d1 - 3, d1),
                halo_descriptor(2, 2, 2, d2 - 3, d2), halo_descriptor(1, 1, 1, d3 - 2, d3)};
            using bc_t = reference_specs::examples_boundaries::direction_bc_input<double>;
            make_boundary<storage::gpu>(halos, bc_t(7.5), minus_j_only()).apply(out_s, in_s);
            auto ref = b.value(-3.).build();
            host_boundary_apply(halos, bc_t(7.5), bc_impl::mask_of(minus_j_only()), host_field<double>{ref->host_view()},
                host_field<double>{in_s->host_view()});
            bool ok = same_everywhere(out_s, ref) && out_s->const_host_view()(20, 10, 5) == -3. &&
                      out_s->const_host_view()(20, d2 - 1, 5) == -3.;
            report("boundary functor with a callable predicate (j-minus regions only)", ok);
        }
    }

    // --bench: the engine on the reference's hdiff, vadv and tridiagonal specs at 256x256x80 FP64 over 4 rotating
    // field sets (past the 256 MB MALL), mean microseconds per run from one hipEvent pair around `reps` back-to-back
    // runs
    int bench(int reps) {
        const int ni = 256, nj = 256, nk = 80, nsets = 4;
        std::vector<store<double>> hin, hco, hout, uts, us, up, ut, wc, inf, dia, sup, rhs, tout;
        for (int s = 0; s < nsets; ++s) {
            hin.push_back(seeded<double>(ni + 4, nj + 4, nk, 2, 42 + s, 1, -1, 1));
            hco.push_back(seeded<double>(ni + 4, nj + 4, nk, 2, 42 + s, 2, 0, 0.25));
            hout.push_back(storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(2, 2, 0)
                               .build());
            uts.push_back(seeded<double>(ni + 6, nj + 6, nk, 3, 42 + s, 6, -1, 1));
            us.push_back(seeded<double>(ni + 6, nj + 6, nk, 3, 42 + s, 3, -1, 1));
            up.push_back(seeded<double>(ni + 6, nj + 6, nk, 3, 42 + s, 4, -1, 1));
            ut.push_back(seeded<double>(ni + 6, nj + 6, nk, 3, 42 + s, 5, -1, 1));
            wc.push_back(seeded<double>(ni + 6, nj + 6, nk + 1, 3, 42 + s, 7, -0.5, 0.5));
            inf.push_back(seeded<double>(ni, nj, nk, 0, 42 + s, 11, -1, 1));
            dia.push_back(seeded<double>(ni, nj, nk, 0, 42 + s, 12, 3, 4));
            sup.push_back(seeded<double>(ni, nj, nk, 0, 42 + s, 13, -1, 1));
            rhs.push_back(seeded<double>(ni, nj, nk, 0, 42 + s, 14, -1, 1));
            tout.push_back(storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk).build());
        }
        auto gh = halo_grid(ni, nj, nk, 2), gv = halo_grid(ni, nj, nk, 3);
        auto gt = make_grid(ni, nj, nk);
        auto hd = [&](int s) {
            run(reference_specs::horizontal_diffusion::spec<double>(), gfx950<>(), gh, hin[s % nsets], hco[s % nsets],
                hout[s % nsets]);
        };
        auto va = [&](int s) {
            run(reference_specs::vertical_advection_dycore::spec<double>(), gfx950<>(), gv, uts[s % nsets], us[s % nsets],
                wc[s % nsets], up[s % nsets], ut[s % nsets], global_parameter(0.15));
        };
        auto tri = [&](int s) {
            run([](auto inf, auto diag, auto sup, auto rhs, auto out) {
                using namespace reference_specs::tridiagonal;
                return multi_pass(execute_forward().stage(forward_thomas(), inf, diag, sup, rhs),
                    execute_backward().stage(backward_thomas(), out, sup, rhs));
            },
                gfx950<>(), gt, inf[s % nsets], dia[s % nsets], sup[s % nsets], rhs[s % nsets], tout[s % nsets]);
        };
        // host microseconds per call of the last timed loop (enqueue only: the loop does not wait for the device)
        double host_us = 0;
        auto time_us = [&](auto &&fn) {
            for (int s = 0; s < nsets; ++s)
                fn(s);
            hipEvent_t e0, e1;
            check(hipEventCreate(&e0) == hipSuccess ? 0 : -2, "event");
            check(hipEventCreate(&e1) == hipSuccess ? 0 : -2, "event");
            check(gt355_synchronize(), "sync");
            (void)hipEventRecord(e0, nullptr);
            const auto t0 = std::chrono::steady_clock::now();
            for (int s = 0; s < reps; ++s)
                fn(s);
            host_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / reps;
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return 1e3 * ms / reps;
        };
        const double hd_us = time_us(hd), hd_host = host_us;
        const double va_us = time_us(va), va_host = host_us;
        const double tri_us = time_us(tri), tri_host = host_us;
        const double pts = double(ni) * nj * nk;
        // algorithmic bytes per point: hdiff 24 (in, coeff, out), vadv 48 (five reads, one write), tridiagonal 56
        // (inf, diag, sup, rhs read; sup, rhs, out written)
        std::printf("{\"hdiff_spec_us\": %.2f, \"hdiff_gbs\": %.1f, \"vadv_spec_us\": %.2f, \"vadv_gbs\": %.1f, "
                    "\"tridiagonal_spec_us\": %.2f, \"tridiagonal_gbs\": %.1f, \"reps\": %d, "
                    "\"host_us_per_run\": {\"hdiff\": %.2f, \"vadv\": %.2f, \"tridiagonal\": %.2f}}\n",
            hd_us, 24 * pts / (hd_us * 1e-6) / 1e9, va_us, 48 * pts / (va_us * 1e-6) / 1e9, tri_us,
            56 * pts / (tri_us * 1e-6) / 1e9, reps, hd_host, va_host, tri_host);
        return 0;
    }
} // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::strcmp(argv[1], "--bench") == 0) {
        try {
            check(gt355_init(0), "init");
            return bench(argc > 2 ? std::atoi(argv[2]) : 50);
        } catch (std::exception const &e) {
            std::printf("[FAIL] exception: %s\n", e.what());
            return 1;
        }
    }
    try {
        check(gt355_init(0), "init");
        copy_and_laplacian();
        horizontal_diffusion(23, 11, 43);
        horizontal_diffusion(64, 64, 60);
        horizontal_diffusion(130, 37, 5, false);
        horizontal_diffusion(256, 256, 80);
        vertical_advection(23, 11, 43);
        vertical_advection(70, 9, 80);
        tridiagonal(12, 33, 6, true);
        tridiagonal(23, 11, 40, false);
        simple_hori_diff(12, 33, 61);
        stencil_functions(23, 11, 43);
        hdiff_functions<reference_specs::horizontal_diffusion_functions::variation::monolithic>("monolithic", 23, 11, 43);
        hdiff_functions<reference_specs::horizontal_diffusion_functions::variation::call>("call", 23, 11, 43);
        hdiff_functions<reference_specs::horizontal_diffusion_functions::variation::call_offsets>("call_offsets", 23, 11, 43);
        hdiff_functions<reference_specs::horizontal_diffusion_functions::variation::procedures>("procedures", 23, 11, 43);
        hdiff_functions<reference_specs::horizontal_diffusion_functions::variation::procedures_offsets>("procedures_offsets", 64, 64, 8);
        expandable_parameters(23, 11, 43);
        positional_and_fusion();
        expandable_and_errors();
        kcache_tests();
        kparallel_tests();
        mss_fusion_rule();
        horizontal_dependencies(70, 37, 19);
        horizontal_dependencies(130, 70, 24);
        whole_axis_access(12, 33, 61);
        whole_axis_access(23, 11, 43);
        precision_policy(23, 11, 43);
        precision_policy(256, 256, 80);
        two_streams(130, 70, 24);
        boundary_functors();
        fast_path();
    } catch (std::exception const &e) {
        std::printf("[FAIL] exception: %s\n", e.what());
        return 1;
    }
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
