// The generic stencil engine (include/gt355/stencil.hpp) on user specs written like the reference's cartesian
// regression tests: functors with accessors, extents and interval overloads, multistage specs with temporaries, run
// on the gfx950 backend tag.  Expected values come from the oracle (oracle/gt_oracle.h) on the same seeded inputs
// (FP64 must match bit for bit: the functors evaluate the same operations in the same order) or from the reference
// tests' known answers.
//
//   gt355_generic            all cases (needs a GPU)
//   gt355_generic --bench N  times the engine's hdiff and vadv specs at 256x256x80 FP64 (one JSON line)
#include <gt355/stencil.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../oracle/gt_oracle.h"

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
    struct copy_functor {
        using in = in_accessor<0>;
        using out = inout_accessor<1>;
        using param_list = make_param_list<in, out>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval) {
            eval(out()) = eval(in());
        }
    };

    struct lap_functor {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval) {
            using T = std::decay_t<decltype(eval(out()))>;
            eval(out()) = T(4) * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
        }
    };

    void copy_and_laplacian() {
        const int ni = 23, nj = 11, nk = 43, h = 1;
        auto in = seeded<double>(ni + 2, nj + 2, nk, h, 42, 1, -1, 1);
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 2, nj + 2, nk).halos(h, h, 0);
        auto out = b.build(), lap = b.build();
        auto g = halo_grid(ni, nj, nk, h);
        run_single_stage(copy_functor(), gfx950<>(), g, in, out);
        report("run_single_stage copy", same_interior(out, mirror(in), g));
        run_single_stage(lap_functor(), gfx950<>(), g, lap, in);
        auto hin = mirror(in);
        std::vector<double> ref(hin.size(), 0.);
        auto st = in->strides();
        const size_t o = h + h * st[1];
        gto_lap_f64(ref.data() + o, hin.data() + o, ni, nj, nk, st[1], st[2]);
        report("run_single_stage laplacian (extent <-1,1,-1,1>)", same_interior(lap, ref, g));
    }

    // ------------------------------------------------------------------------------------- horizontal diffusion
    struct hd_lap {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            using T = std::decay_t<decltype(eval(out()))>;
            eval(out()) = T{4} * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
        }
    };
    struct hd_flx {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<0, 1, 0, 0>>;
        using lap = in_accessor<2, extent<0, 1, 0, 0>>;
        using param_list = make_param_list<out, in, lap>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            auto res = eval(lap(1, 0)) - eval(lap(0, 0));
            eval(out()) = res * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0 : res;
        }
    };
    struct hd_fly {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<0, 0, 0, 1>>;
        using lap = in_accessor<2, extent<0, 0, 0, 1>>;
        using param_list = make_param_list<out, in, lap>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            auto res = eval(lap(0, 1)) - eval(lap(0, 0));
            eval(out()) = res * (eval(in(0, 1)) - eval(in(0, 0))) > 0 ? 0 : res;
        }
    };
    struct hd_out {
        using out = inout_accessor<0>;
        using in = in_accessor<1>;
        using flx = in_accessor<2, extent<-1, 0, 0, 0>>;
        using fly = in_accessor<3, extent<0, 0, -1, 0>>;
        using coeff = in_accessor<4>;
        using param_list = make_param_list<out, in, flx, fly, coeff>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(out()) =
                eval(in()) - eval(coeff()) * (eval(flx()) - eval(flx(-1, 0)) + eval(fly()) - eval(fly(0, -1)));
        }
    };

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
                    .stage(hd_lap(), lap, in)
                    .stage(hd_flx(), flx, in, lap)
                    .stage(hd_fly(), fly, in, lap)
                    .stage(hd_out(), out, in, flx, fly, coeff);
            },
                gfx950<>(), g, in, coeff, out);
        else
            run([](auto in, auto coeff, auto out) {
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .stage(hd_lap(), lap, in)
                    .stage(hd_flx(), flx, in, lap)
                    .stage(hd_fly(), fly, in, lap)
                    .stage(hd_out(), out, in, flx, fly, coeff);
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
    struct sf_lap {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            using T = std::decay_t<decltype(eval(out()))>;
            eval(out()) = T{4} * eval(in()) - (eval(in(1, 0)) + eval(in(0, 1)) + eval(in(-1, 0)) + eval(in(0, -1)));
        }
    };
    struct sf_flx {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 2, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            auto hi = call<sf_lap>::with(eval, in(1, 0));
            auto lo = call<sf_lap>::with(eval, in(0, 0));
            auto flx = hi - lo;
            eval(out()) = flx * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0 : flx;
        }
    };
    struct sf_fly {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 2>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            auto hi = call<sf_lap>::with(eval, in(0, 1));
            auto lo = call<sf_lap>::with(eval, in(0, 0));
            auto fly = hi - lo;
            eval(out()) = fly * (eval(in(0, 1)) - eval(in(0, 0))) > 0 ? 0 : fly;
        }
    };
    struct sf_out {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-2, 2, -2, 2>>;
        using coeff = in_accessor<2>;
        using param_list = make_param_list<out, in, coeff>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            auto flx_hi = call<sf_flx>::with(eval, in(0, 0));
            auto flx_lo = call<sf_flx>::with(eval, in(-1, 0));
            auto fly_hi = call<sf_fly>::with(eval, in(0, 0));
            auto fly_lo = call<sf_fly>::with(eval, in(0, -1));
            eval(out()) = eval(in()) - eval(coeff()) * (flx_hi - flx_lo + fly_hi - fly_lo);
        }
    };
    // horizontal_diffusion_functions.cpp:22-205: flx / fly evaluate the laplacian at the point and its east or
    // north neighbour in five ways (monolithic expressions, call<>::at<>, call<> with offset accessors,
    // call_proc<>::at<>, call_proc<> with offset accessors); flx and fly are ij-cached, lap is never stored
    struct hf_lap {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(out()) = 4. * eval(in()) - (eval(in(-1, 0)) + eval(in(0, -1)) + eval(in(0, 1)) + eval(in(1, 0)));
        }
    };
    template <class Eval, class Acc>
    using hf_float = std::decay_t<decltype(std::declval<Eval>()(Acc()))>;
    enum class hf_var { monolithic, call, call_offsets, procedures, procedures_offsets };
    template <hf_var V, class Acc>
    struct hf_lap_at;
    template <class Acc>
    struct hf_lap_at<hf_var::monolithic, Acc> {
        template <class Eval>
        GT_FUNCTION static auto do00(Eval &eval) {
            return hf_float<Eval, Acc>(4) * eval(Acc()) -
                   (eval(Acc(-1, 0)) + eval(Acc(0, -1)) + eval(Acc(0, 1)) + eval(Acc(1, 0)));
        }
        template <class Eval>
        GT_FUNCTION static auto do10(Eval &eval) {
            return hf_float<Eval, Acc>(4) * eval(Acc(1, 0)) -
                   (eval(Acc(0, 0)) + eval(Acc(1, -1)) + eval(Acc(1, 1)) + eval(Acc(2, 0)));
        }
        template <class Eval>
        GT_FUNCTION static auto do01(Eval &eval) {
            return hf_float<Eval, Acc>(4) * eval(Acc(0, 1)) -
                   (eval(Acc(-1, 1)) + eval(Acc(0, 0)) + eval(Acc(0, 2)) + eval(Acc(1, 1)));
        }
    };
    template <class Acc>
    struct hf_lap_at<hf_var::call, Acc> {
        template <class Eval>
        GT_FUNCTION static auto do00(Eval &eval) {
            return call<hf_lap>::at<0, 0, 0>::with(eval, Acc());
        }
        template <class Eval>
        GT_FUNCTION static auto do10(Eval &eval) {
            return call<hf_lap>::at<1, 0, 0>::with(eval, Acc());
        }
        template <class Eval>
        GT_FUNCTION static auto do01(Eval &eval) {
            return call<hf_lap>::at<0, 1, 0>::with(eval, Acc());
        }
    };
    template <class Acc>
    struct hf_lap_at<hf_var::call_offsets, Acc> {
        template <class Eval>
        GT_FUNCTION static auto do00(Eval &eval) {
            return call<hf_lap>::with(eval, Acc(0, 0));
        }
        template <class Eval>
        GT_FUNCTION static auto do10(Eval &eval) {
            return call<hf_lap>::with(eval, Acc(1, 0));
        }
        template <class Eval>
        GT_FUNCTION static auto do01(Eval &eval) {
            return call<hf_lap>::with(eval, Acc(0, 1));
        }
    };
    template <class Acc>
    struct hf_lap_at<hf_var::procedures, Acc> {
        template <class Eval>
        GT_FUNCTION static auto do00(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::at<0, 0, 0>::with(eval, res, Acc());
            return res;
        }
        template <class Eval>
        GT_FUNCTION static auto do10(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::at<1, 0, 0>::with(eval, res, Acc());
            return res;
        }
        template <class Eval>
        GT_FUNCTION static auto do01(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::at<0, 1, 0>::with(eval, res, Acc());
            return res;
        }
    };
    template <class Acc>
    struct hf_lap_at<hf_var::procedures_offsets, Acc> {
        template <class Eval>
        GT_FUNCTION static auto do00(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::with(eval, res, Acc(0, 0));
            return res;
        }
        template <class Eval>
        GT_FUNCTION static auto do10(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::with(eval, res, Acc(1, 0));
            return res;
        }
        template <class Eval>
        GT_FUNCTION static auto do01(Eval &eval) {
            hf_float<Eval, Acc> res;
            call_proc<hf_lap>::with(eval, res, Acc(0, 1));
            return res;
        }
    };
    template <hf_var V>
    struct hf_flx {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 2, -1, 1>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(out()) = hf_lap_at<V, in>::do10(eval) - hf_lap_at<V, in>::do00(eval);
            eval(out()) = eval(out()) * (eval(in(1, 0)) - eval(in(0, 0))) > 0 ? 0.0 : eval(out());
        }
    };
    template <hf_var V>
    struct hf_fly {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 2>>;
        using param_list = make_param_list<out, in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(out()) = hf_lap_at<V, in>::do01(eval) - hf_lap_at<V, in>::do00(eval);
            eval(out()) = eval(out()) * (eval(in(0, 1, 0)) - eval(in(0, 0, 0))) > 0 ? 0.0 : eval(out());
        }
    };

    template <hf_var V>
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
                .stage(hf_flx<V>(), flx, in)
                .stage(hf_fly<V>(), fly, in)
                .stage(hd_out(), out, in, flx, fly, coeff);
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
    struct ep_copy {
        using parameters_out = inout_accessor<0>;
        using parameters_in = accessor<1>;
        using param_list = make_param_list<parameters_out, parameters_in>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &eval) {
            eval(parameters_out{}) = eval(parameters_in{});
        }
    };
    struct ep_single {
        using o1 = inout_accessor<0>;
        using o2 = inout_accessor<1>;
        using o3 = inout_accessor<2>;
        using o4 = inout_accessor<3>;
        using o5 = inout_accessor<4>;
        using i1 = in_accessor<5>;
        using i2 = in_accessor<6>;
        using i3 = in_accessor<7>;
        using i4 = in_accessor<8>;
        using i5 = in_accessor<9>;
        using param_list = make_param_list<o1, o2, o3, o4, o5, i1, i2, i3, i4, i5>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            eval(o1()) = eval(i1());
            eval(o2()) = eval(i2());
            eval(o3()) = eval(i3());
            eval(o4()) = eval(i4());
            eval(o5()) = eval(i5());
        }
    };

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
                return execute_parallel().ij_cached(tmp).stage(ep_copy(), tmp, in).stage(ep_copy(), out, tmp);
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
                    .stage(ep_single(), tmp0, tmp1, tmp2, tmp3, tmp4, in0, in1, in2, in3, in4)
                    .stage(ep_single(), out0, out1, out2, out3, out4, tmp0, tmp1, tmp2, tmp3, tmp4);
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
        run_single_stage(sf_out(), gfx950<>(), g, out, in, coeff);
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
    using axis_t = axis<1, axis_config::offset_limit<3>>;
    using full_t = axis_t::full_interval;

    struct va_forward {
        using utens_stage = in_accessor<0>;
        using wcon = in_accessor<1, extent<0, 1, 0, 0, 0, 1>>;
        using u_stage = in_accessor<2, extent<0, 0, 0, 0, -1, 1>>;
        using u_pos = in_accessor<3>;
        using utens = in_accessor<4>;
        using dtr_stage = in_accessor<5>;
        using ccol = inout_accessor<6, extent<0, 0, 0, 0, -1, 0>>;
        using dcol = inout_accessor<7, extent<0, 0, 0, 0, -1, 0>>;
        using param_list = make_param_list<utens_stage, wcon, u_stage, u_pos, utens, dtr_stage, ccol, dcol>;

        template <class Eval>
        GT_FUNCTION static auto rhs(Eval &&eval) {
            return eval(dtr_stage()) * eval(u_pos()) + eval(utens()) + eval(utens_stage());
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full_t::first_level) {
            using T = std::decay_t<decltype(eval(ccol()))>;
            auto gcv = T(.25) * (eval(wcon(1, 0, 1)) + eval(wcon(0, 0, 1)));
            auto cs = gcv * T(0.5);
            auto c = gcv * T(0.5);
            auto b = eval(dtr_stage()) - c;
            auto correction = -cs * (eval(u_stage(0, 0, 1)) - eval(u_stage()));
            auto d = rhs(eval) + correction;
            auto divided = T(1) / b;
            eval(ccol()) = c * divided;
            eval(dcol()) = d * divided;
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full_t::modify<1, -1>) {
            using T = std::decay_t<decltype(eval(ccol()))>;
            auto gav = -T(.25) * (eval(wcon(1, 0, 0)) + eval(wcon(0, 0, 0)));
            auto gcv = T(.25) * (eval(wcon(1, 0, 1)) + eval(wcon(0, 0, 1)));
            auto as = gav * T(0.5);
            auto cs = gcv * T(0.5);
            auto a = gav * T(0.5);
            auto c = gcv * T(0.5);
            auto b = eval(dtr_stage()) - a - c;
            auto correction =
                -as * (eval(u_stage(0, 0, -1)) - eval(u_stage())) - cs * (eval(u_stage(0, 0, 1)) - eval(u_stage()));
            auto d = rhs(eval) + correction;
            auto divided = T(1) / (b - eval(ccol(0, 0, -1)) * a);
            eval(ccol()) = c * divided;
            eval(dcol()) = (d - eval(dcol(0, 0, -1)) * a) * divided;
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full_t::last_level) {
            using T = std::decay_t<decltype(eval(ccol()))>;
            auto gav = -T(.25) * (eval(wcon(1, 0, 0)) + eval(wcon()));
            auto as = gav * T(0.5);
            auto a = gav * T(0.5);
            auto b = eval(dtr_stage()) - a;
            auto correction = -as * (eval(u_stage(0, 0, -1)) - eval(u_stage()));
            auto d = rhs(eval) + correction;
            auto divided = T(1) / (b - eval(ccol(0, 0, -1)) * a);
            eval(dcol()) = (d - eval(dcol(0, 0, -1)) * a) * divided;
        }
    };

    struct va_backward {
        using utens_stage = inout_accessor<0>;
        using u_pos = in_accessor<1>;
        using dtr_stage = in_accessor<2>;
        using ccol = in_accessor<3>;
        using dcol = in_accessor<4>;
        using data_col = inout_accessor<5, extent<0, 0, 0, 0, 0, 1>>;
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

    void vertical_advection(int ni, int nj, int nk) {
        const int h = 3;
        auto utens_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 6, -1, 1);
        auto u_stage = seeded<double>(ni + 6, nj + 6, nk, h, 42, 3, -1, 1);
        auto u_pos = seeded<double>(ni + 6, nj + 6, nk, h, 42, 4, -1, 1);
        auto utens = seeded<double>(ni + 6, nj + 6, nk, h, 42, 5, -1, 1);
        auto wcon = seeded<double>(ni + 6, nj + 6, nk + 1, h, 42, 7, -0.5, 0.5);
        auto g = halo_grid(ni, nj, nk, h);
        auto uts0 = mirror(utens_stage);
        auto spec = [](auto utens_stage, auto u_stage, auto wcon, auto u_pos, auto utens, auto dtr_stage) {
            GT_DECLARE_TMP(double, ccol, dcol, data_col);
            return multi_pass(execute_forward()
                                  .k_cached(cache_io_policy::flush(), ccol, dcol)
                                  .k_cached(cache_io_policy::fill(), u_stage)
                                  .stage(va_forward(), utens_stage, wcon, u_stage, u_pos, utens, dtr_stage, ccol, dcol),
                execute_backward().k_cached(data_col).stage(
                    va_backward(), utens_stage, u_pos, dtr_stage, ccol, dcol, data_col));
        };
        run(spec, gfx950<>(), g, utens_stage, u_stage, wcon, u_pos, utens, global_parameter(0.15));
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
    using full1_t = axis<1>::full_interval;
    struct thomas_forward {
        using inf = in_accessor<0>;
        using diag = in_accessor<1>;
        using sup = inout_accessor<2, extent<0, 0, 0, 0, -1, 0>>;
        using rhs = inout_accessor<3, extent<0, 0, 0, 0, -1, 0>>;
        using param_list = make_param_list<inf, diag, sup, rhs>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval, full1_t::modify<1, 0>) {
            eval(sup{}) = eval(sup{} / (diag{} - sup{0, 0, -1} * inf{}));
            eval(rhs{}) = eval((rhs{} - inf{} * rhs{0, 0, -1}) / (diag{} - sup{0, 0, -1} * inf{}));
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval, full1_t::first_level) {
            eval(sup{}) = eval(sup{}) / eval(diag{});
            eval(rhs{}) = eval(rhs{}) / eval(diag{});
        }
    };
    struct thomas_backward {
        using out = inout_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
        using sup = in_accessor<1>;
        using rhs = in_accessor<2>;
        using param_list = make_param_list<out, sup, rhs>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval, full1_t::modify<0, -1>) {
            eval(out{}) = eval(rhs{} - sup{} * out{0, 0, 1});
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval, full1_t::last_level) {
            eval(out{}) = eval(rhs{});
        }
    };

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
            return multi_pass(execute_forward().stage(thomas_forward(), inf, diag, sup, rhs),
                execute_backward().stage(thomas_backward(), out, sup, rhs));
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
    struct wlap {
        using out = inout_accessor<0>;
        using in = in_accessor<1, extent<-1, 1, -1, 1>>;
        using crlato = in_accessor<2>;
        using crlatu = in_accessor<3>;
        using param_list = make_param_list<out, in, crlato, crlatu>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval eval) {
            using T = std::decay_t<decltype(eval(out()))>;
            eval(out()) = eval(in(1, 0)) + eval(in(-1, 0)) - T{2} * eval(in()) +
                          eval(crlato()) * (eval(in(0, 1)) - eval(in())) + eval(crlatu()) * (eval(in(0, -1)) - eval(in()));
        }
    };
    struct divflux {
        using out = inout_accessor<0>;
        using in = in_accessor<1>;
        using lap = in_accessor<2, extent<-1, 1, -1, 1>>;
        using crlato = in_accessor<3>;
        using coeff = in_accessor<4>;
        using param_list = make_param_list<out, in, lap, crlato, coeff>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &eval) {
            auto fluxx = eval(lap(1, 0)) - eval(lap());
            auto fluxx_m = eval(lap()) - eval(lap(-1, 0));
            auto fluxy = eval(crlato()) * (eval(lap(0, 1)) - eval(lap()));
            auto fluxy_m = eval(crlato()) * (eval(lap()) - eval(lap(0, -1)));
            eval(out()) = eval(in()) + ((fluxx_m - fluxx) + (fluxy_m - fluxy)) * eval(coeff());
        }
    };

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
                .stage(wlap(), lap, in, crlato, crlatu)
                .stage(divflux(), out, in, lap, crlato, coeff);
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
    struct positional_functor {
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

    struct set_one {
        using out = inout_accessor<0>;
        using param_list = make_param_list<out>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval) {
            eval(out()) = 1;
        }
    };
    struct shift_down {
        using in = in_accessor<0, extent<0, 0, 0, 0, 0, 1>>;
        using out = inout_accessor<1>;
        using param_list = make_param_list<in, out>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full1_t::modify<0, -1>) {
            eval(out()) = eval(in(0, 0, 1));
        }
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval, full1_t::last_level) {
            eval(out()) = eval(in());
        }
    };

    void positional_and_fusion() {
        const int ni = 12, nj = 33, nk = 61, h = 2;
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni + 4, nj + 4, nk).halos(h, h, 0);
        auto out = b.value(-1.).build();
        auto g = halo_grid(ni, nj, nk, h);
        run_single_stage(positional_functor(), gfx950<>(), g, out, positional<dim::i>(), positional<dim::j>(),
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
            return multi_pass(execute_parallel().stage(set_one(), tmp), execute_parallel().stage(shift_down(), tmp, out));
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
    struct scale_functor {
        using data = inout_accessor<0>;
        using data_nnow = in_accessor<1>;
        using rho = in_accessor<2>;
        using param_list = make_param_list<data, data_nnow, rho>;
        template <class Eval>
        GT_FUNCTION static void apply(Eval &&eval) {
            eval(data()) = eval(rho()) * eval(data_nnow());
        }
    };

    void expandable_and_errors() {
        const int ni = 12, nj = 33, nk = 61;
        auto b = storage::builder<storage::gpu>.type<double>().dimensions(ni, nj, nk);
        std::vector<store<double>> in, out;
        for (int m = 0; m < 5; ++m) {
            out.push_back(b.build());
            in.push_back(b.value(double(m)).build());
        }
        auto rho = b.value(1.1).build();
        expandable_run<2>([](auto out, auto in, auto rho) { return execute_parallel().stage(scale_functor(), out, in, rho); },
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
            run_single_stage(lap_functor(), gfx950<>(), make_grid(ni, nj, nk), res, small);
        } catch (std::runtime_error const &e) {
            threw = std::strstr(e.what(), "too small") != nullptr;
        }
        report("field bounds check of the extent analysis", threw);
    }
    // --bench: the engine's hdiff and vadv specs at 256x256x80 FP64 over 4 rotating field sets (past the 256 MB
    // MALL), mean microseconds per run from one hipEvent pair around `reps` back-to-back runs
    int bench(int reps) {
        const int ni = 256, nj = 256, nk = 80, nsets = 4;
        std::vector<store<double>> hin, hco, hout, uts, us, up, ut, wc;
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
        }
        auto gh = halo_grid(ni, nj, nk, 2), gv = halo_grid(ni, nj, nk, 3);
        auto hd = [&](int s) {
            run([](auto in, auto coeff, auto out) {
                GT_DECLARE_TMP(double, lap, flx, fly);
                return execute_parallel()
                    .ij_cached(lap, flx, fly)
                    .stage(hd_lap(), lap, in)
                    .stage(hd_flx(), flx, in, lap)
                    .stage(hd_fly(), fly, in, lap)
                    .stage(hd_out(), out, in, flx, fly, coeff);
            },
                gfx950<>(), gh, hin[s % nsets], hco[s % nsets], hout[s % nsets]);
        };
        auto va = [&](int s) {
            run([](auto utens_stage, auto u_stage, auto wcon, auto u_pos, auto utens, auto dtr_stage) {
                GT_DECLARE_TMP(double, ccol, dcol, data_col);
                return multi_pass(execute_forward()
                                      .k_cached(cache_io_policy::flush(), ccol, dcol)
                                      .k_cached(cache_io_policy::fill(), u_stage)
                                      .stage(va_forward(), utens_stage, wcon, u_stage, u_pos, utens, dtr_stage, ccol, dcol),
                    execute_backward().k_cached(data_col).stage(
                        va_backward(), utens_stage, u_pos, dtr_stage, ccol, dcol, data_col));
            },
                gfx950<>(), gv, uts[s % nsets], us[s % nsets], wc[s % nsets], up[s % nsets], ut[s % nsets],
                global_parameter(0.15));
        };
        auto time_us = [&](auto &&fn) {
            for (int s = 0; s < nsets; ++s)
                fn(s);
            hipEvent_t e0, e1;
            check(hipEventCreate(&e0) == hipSuccess ? 0 : -2, "event");
            check(hipEventCreate(&e1) == hipSuccess ? 0 : -2, "event");
            check(gt355_synchronize(), "sync");
            (void)hipEventRecord(e0, nullptr);
            for (int s = 0; s < reps; ++s)
                fn(s);
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return 1e3 * ms / reps;
        };
        const double hd_us = time_us(hd), va_us = time_us(va);
        const double pts = double(ni) * nj * nk;
        std::printf("{\"hdiff_fused_lds_us\": %.2f, \"hdiff_gbs\": %.1f, \"vadv_vertical_us\": %.2f, \"vadv_gbs\": "
                    "%.1f, \"reps\": %d}\n",
            hd_us, 24 * pts / (hd_us * 1e-6) / 1e9, va_us, 48 * pts / (va_us * 1e-6) / 1e9, reps);
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
        hdiff_functions<hf_var::monolithic>("monolithic", 23, 11, 43);
        hdiff_functions<hf_var::call>("call", 23, 11, 43);
        hdiff_functions<hf_var::call_offsets>("call_offsets", 23, 11, 43);
        hdiff_functions<hf_var::procedures>("procedures", 23, 11, 43);
        hdiff_functions<hf_var::procedures_offsets>("procedures_offsets", 64, 64, 8);
        expandable_parameters(23, 11, 43);
        positional_and_fusion();
        expandable_and_errors();
    } catch (std::exception const &e) {
        std::printf("[FAIL] exception: %s\n", e.what());
        return 1;
    }
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
