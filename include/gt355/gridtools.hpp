/*
 * gt355/gridtools.hpp -- header-only C++17 front end of libgt355 with the GridTools user API for the stencils
 * of the hot path: storage::builder / data_store (storage/builder.hpp:168-326, storage/data_store.hpp:33-147),
 * with selector<I, J, K> masked dimensions, halo_descriptor (common/halo_descriptor.hpp:44-227), make_grid
 * (stencil/frontend/make_grid.hpp:19-42),
 * global_parameter, stencil::run and expandable_run (stencil/frontend/run.hpp:201-258,
 * stencil/frontend/expandable_run.hpp:150-183) with a backend tag, and boundaries::boundary /
 * make_boundary (boundaries/boundary.hpp:31-79).
 *
 * The spec argument of run() names one of the path's stencils (specs::horizontal_diffusion, ...) in place of the
 * reference's spec lambda; its fields follow the order of the reference regression specs.  Every call checks its
 * status and throws std::runtime_error with the library's message (common/cuda_util.hpp:20-35).  Host code only:
 * any C++17 compiler, link with -lgt355.
 */
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <utility>
#include <vector>

#include "../gt355.h"

namespace gt355 {

    inline void check(int rc, const char *what) {
        if (rc != GT355_OK)
            throw std::runtime_error(std::string(what) + ": " + gt355_last_error());
    }

    // common/halo_descriptor.hpp:44-227: (minus, plus, begin, end, total_length), end inclusive
    class halo_descriptor {
        int m_minus = 0, m_plus = 0, m_begin = 0, m_end = 0, m_total_length = 1;

      public:
        halo_descriptor() = default;
        // halo_descriptor(size): no halo (halo_descriptor.hpp:78-79)
        explicit halo_descriptor(int size) : m_end(size - 1), m_total_length(size) {
            if (size < 1)
                throw std::invalid_argument("halo_descriptor: size < 1");
        }
        halo_descriptor(int minus, int plus, int begin, int end, int total_length)
            : m_minus(minus), m_plus(plus), m_begin(begin), m_end(end), m_total_length(total_length) {
            // the constructor's assertions (halo_descriptor.hpp:88-93), as exceptions
            if (minus + plus + (end - begin + 1) > total_length)
                throw std::invalid_argument("halo_descriptor: minus + plus + size exceeds total_length");
            if (begin > end)
                throw std::invalid_argument("halo_descriptor: begin > end");
            if (plus > total_length - end - 1)
                throw std::invalid_argument("halo_descriptor: plus halo exceeds total_length");
            if (begin < minus)
                throw std::invalid_argument("halo_descriptor: begin < minus");
        }
        int minus() const { return m_minus; }
        int plus() const { return m_plus; }
        int begin() const { return m_begin; }
        int end() const { return m_end; }
        int total_length() const { return m_total_length; }
    };

    namespace storage {
        // Traits tag of the gpu storage (storage/gpu.hpp:69-105): i-innermost, i padded to 128 B, first interior
        // point 128-byte aligned, device memory with a lazily synchronised host mirror.
        struct gpu {};

        // the library's element type code (gt355_dtype); storages of other arithmetic types (int for example, as
        // boundary functors may use) get -1, which the library's stencil entry points reject
        template <class T>
        constexpr int dtype_of() {
            static_assert(std::is_arithmetic<T>::value && (sizeof(T) == 4 || sizeof(T) == 8),
                "gt355 storage holds 4- or 8-byte arithmetic values");
            return std::is_same<T, double>::value ? GT355_F64 : std::is_same<T, float>::value ? GT355_F32 : -1;
        }

        template <class T>
        struct host_view {
            T *base;
            int64_t stride[3];
            std::array<int, 3> len;
            T &operator()(int i, int j, int k) const { return base[i * stride[0] + j * stride[1] + k * stride[2]]; }
            std::array<int, 3> const &lengths() const { return len; }
        };

        // storage/data_store.hpp:33-147 for the gpu traits: device allocation plus host mirror; host_view() makes
        // the host copy current and the device copy stale, get_target_ptr() the converse; the const accessors
        // leave both current.
        template <class T>
        class data_store {
            gt355_layout m_layout{};
            std::array<int, 3> m_lengths{};
            std::array<bool, 3> m_sel{{true, true, true}};
            void *m_raw = nullptr;
            T *m_target = nullptr;
            std::vector<T> m_host;
            enum class state { both, host_only, target_only } m_state = state::both;
            std::string m_name;

            size_t span() const { return (size_t)m_layout.stride[2] * (size_t)m_layout.lengths[2]; }
            void to_host() {
                if (m_state == state::target_only) {
                    // launches may be pending on any stream: wait for the device before reading back
                    check(gt355_synchronize(), "data_store: update host");
                    check(gt355_memcpy_d2h(m_host.data(), m_target, span() * sizeof(T)), "data_store: update host");
                    m_state = state::both;
                }
            }
            void to_target() {
                if (m_state == state::host_only) {
                    check(gt355_memcpy_h2d(m_target, m_host.data(), span() * sizeof(T)), "data_store: update target");
                    m_state = state::both;
                }
            }

          public:
            using data_t = T;
            // sel: the selector of storage/builder.hpp:168-326 -- a masked dimension (false) holds one value for
            // every index: it is allocated with length 1 and has stride 0
            data_store(std::array<int, 3> lengths, std::array<int, 3> halos, std::string name,
                std::array<bool, 3> sel = {{true, true, true}})
                : m_lengths(lengths), m_sel(sel), m_name(std::move(name)) {
                const int64_t l[3] = {sel[0] ? lengths[0] : 1, sel[1] ? lengths[1] : 1, sel[2] ? lengths[2] : 1};
                const int64_t h[3] = {sel[0] ? halos[0] : 0, sel[1] ? halos[1] : 0, sel[2] ? halos[2] : 0};
                check(gt355_storage_layout(sizeof(T), l, h, &m_layout), "data_store: layout");
                check(gt355_malloc((size_t)m_layout.alloc_bytes, &m_raw), "data_store: allocate");
                void *base = nullptr;
                check(gt355_storage_base(&m_layout, m_raw, &base), "data_store: align");
                m_target = static_cast<T *>(base);
                m_host.assign(span(), T(0));
                check(gt355_memset(m_raw, 0, (size_t)m_layout.alloc_bytes, nullptr), "data_store: clear");
                check(gt355_synchronize(), "data_store: clear");
            }
            data_store(data_store const &) = delete;
            data_store &operator=(data_store const &) = delete;
            ~data_store() {
                if (m_raw)
                    gt355_free(m_raw);
            }

            std::string const &name() const { return m_name; }
            std::array<int, 3> lengths() const { return m_lengths; }
            std::array<bool, 3> selector() const { return m_sel; }
            std::array<int, 3> native_lengths() const {
                return {(int)m_layout.padded[0], (int)m_layout.padded[1], (int)m_layout.padded[2]};
            }
            std::array<int64_t, 3> strides() const {
                return {m_sel[0] ? m_layout.stride[0] : 0, m_sel[1] ? m_layout.stride[1] : 0,
                    m_sel[2] ? m_layout.stride[2] : 0};
            }
            std::array<int, 3> halos() const {
                return {(int)m_layout.halos[0], (int)m_layout.halos[1], (int)m_layout.halos[2]};
            }

            T *get_target_ptr() {
                to_target();
                m_state = state::target_only;
                return m_target;
            }
            T const *get_const_target_ptr() {
                to_target();
                return m_target;
            }
            storage::host_view<T> host_view() {
                to_host();
                m_state = state::host_only;
                auto st = strides();
                return {m_host.data(), {st[0], st[1], st[2]}, lengths()};
            }
            storage::host_view<T const> const_host_view() {
                to_host();
                auto st = strides();
                return {m_host.data(), {st[0], st[1], st[2]}, lengths()};
            }

            // The SID of this store after core::shift_origin (stencil/core/backend.hpp:28-34): origin at storage
            // index `origin`, bounds = the allocation box relative to it.  `write` marks the device copy as the
            // only current one.
            gt355_field sid(std::array<int, 3> origin, bool write) {
                T *p = write ? get_target_ptr() : const_cast<T *>(get_const_target_ptr());
                gt355_field f{};
                auto st = strides();
                f.origin = p + origin[0] * st[0] + origin[1] * st[1] + origin[2] * st[2];
                for (int d = 0; d < 3; ++d) {
                    f.stride[d] = st[d];
                    // a masked dimension can be indexed anywhere
                    f.lower[d] = m_sel[d] ? -origin[d] : -(1 << 30);
                    f.upper[d] = m_sel[d] ? m_lengths[d] - origin[d] : (1 << 30);
                }
                f.dtype = dtype_of<T>();
                return f;
            }
        };

        // storage/builder.hpp:168-326: builder<Traits>.type<T>().dimensions(...).halos(...)
        //     [.initializer(f) | .value(v)].name(s).build() -> std::shared_ptr<data_store<T>>
        template <class Traits, class T = void>
        class builder_type {
            std::array<int, 3> m_dims{{0, 0, 0}}, m_halos{{0, 0, 0}};
            std::array<bool, 3> m_sel{{true, true, true}};
            std::string m_name;
            std::function<T(int, int, int)> m_init;

            template <class, class>
            friend class builder_type;

          public:
            template <class U>
            builder_type<Traits, U> type() const {
                builder_type<Traits, U> b;
                b.m_dims = m_dims;
                b.m_halos = m_halos;
                b.m_sel = m_sel;
                b.m_name = m_name;
                return b;
            }
            builder_type dimensions(int i, int j, int k) const {
                builder_type b = *this;
                b.m_dims = {i, j, k};
                return b;
            }
            builder_type halos(int i, int j, int k) const {
                builder_type b = *this;
                b.m_halos = {i, j, k};
                return b;
            }
            // selector<I, J, K>(): dimensions with a 0 flag are masked (selector<0, 1, 0> is a j-only field)
            template <int I, int J, int K>
            builder_type selector() const {
                builder_type b = *this;
                b.m_sel = {I != 0, J != 0, K != 0};
                return b;
            }
            builder_type name(std::string n) const {
                builder_type b = *this;
                b.m_name = std::move(n);
                return b;
            }
            template <class F>
            builder_type initializer(F f) const {
                static_assert(!std::is_void<T>::value, "builder: call type<T>() first");
                builder_type b = *this;
                b.m_init = [f](int i, int j, int k) { return T(f(i, j, k)); };
                return b;
            }
            template <class U = T>
            builder_type value(U v) const {
                return initializer([v](int, int, int) { return v; });
            }
            std::shared_ptr<data_store<T>> build() const {
                static_assert(!std::is_void<T>::value, "builder: call type<T>() first");
                auto ds = std::make_shared<data_store<T>>(m_dims, m_halos, m_name, m_sel);
                if (m_init) {
                    // the reference initializer fills every index of the allocation box (builder.hpp:67-84); a
                    // masked dimension is evaluated at index 0
                    auto v = ds->host_view();
                    const int n0 = m_sel[0] ? m_dims[0] : 1, n1 = m_sel[1] ? m_dims[1] : 1,
                              n2 = m_sel[2] ? m_dims[2] : 1;
                    for (int k = 0; k < n2; ++k)
                        for (int j = 0; j < n1; ++j)
                            for (int i = 0; i < n0; ++i)
                                v(i, j, k) = m_init(i, j, k);
                }
                return ds;
            }
        };

        // inside a template, a dependent element type needs the template keyword: builder<gpu>.template type<T>()
        template <class Traits>
        inline const builder_type<Traits> builder = {};
    } // namespace storage

    // literals::operator""_c (common/integral_constant.hpp): 3_c is std::integral_constant<int, 3>
    namespace literals {
        namespace literals_impl {
            template <char... Cs>
            constexpr int digits() {
                int v = 0;
                ((v = v * 10 + (Cs - '0')), ...);
                return v;
            }
        } // namespace literals_impl
        template <char... Cs>
        constexpr std::integral_constant<int, literals_impl::digits<Cs...>()> operator""_c() {
            return {};
        }
    } // namespace literals

    namespace stencil {
        // ------------------------------------------------------------------------------------------ k levels
        // level<splitter, offset> (stencil/core/level.hpp): offset > 0 counts up from the splitter (1 = the
        // splitter's own level), offset < 0 counts down (-1 = the level below the splitter); 0 is not a level.
        template <int Splitter, int Offset>
        struct level {
            static_assert(Splitter >= 0 && Offset != 0, "level: splitter >= 0 and a non-zero offset");
            static_assert(Offset >= -4 && Offset <= 4, "level offsets are limited to +-4");
            static constexpr int splitter = Splitter, offset = Offset;
        };

        // offset arithmetic that skips 0 (interval_impl::add_offset, stencil/core/interval.hpp)
        constexpr int shift_level_offset(int o, int d) {
            int r = o + d;
            if (o > 0 && r <= 0)
                r -= 1;
            else if (o < 0 && r >= 0)
                r += 1;
            return r;
        }

        // interval<from, to> (stencil/core/interval.hpp): a closed range of levels; modify<l, r> moves its ends,
        // shift<d> moves both, first_level / last_level are its end levels
        template <class From, class To>
        struct interval {
            static_assert(From::splitter < To::splitter ||
                              (From::splitter == To::splitter && From::offset <= To::offset),
                "interval: the from level must not be above the to level");
            using from_t = From;
            using to_t = To;
            template <int A, int B>
            using modify = interval<level<From::splitter, shift_level_offset(From::offset, A)>,
                level<To::splitter, shift_level_offset(To::offset, B)>>;
            template <int D>
            using shift = modify<D, D>;
            using first_level = interval<From, From>;
            using last_level = interval<To, To>;
        };

        namespace axis_config {
            template <int L>
            struct offset_limit {
                static constexpr int value = L;
            };
        } // namespace axis_config

        // axis<N, offset_limit<L>> (stencil/frontend/axis.hpp:37-72): N consecutive k intervals between N + 1
        // splitters, built from the intervals' sizes; get_interval<i, ...> is interval i (or consecutive ones)
        template <int N, class Limit = axis_config::offset_limit<2>>
        class axis {
            static_assert(N >= 1 && N <= 8, "axis: 1 to 8 intervals");
            static_assert(Limit::value >= 1 && Limit::value <= 4, "axis: level offsets are limited to +-4");
            std::array<int, N> m_sizes;

            template <int I, int... Is>
            struct interval_of {
                static constexpr bool consecutive() {
                    int e = I;
                    bool ok = true;
                    ((ok = ok && Is == ++e), ...);
                    return ok;
                }
                static_assert(consecutive(), "axis::get_interval: the intervals must be consecutive");
                static_assert(I >= 0 && I + (int)sizeof...(Is) < N, "axis::get_interval: interval id out of range");
                using type = interval<level<I, 1>, level<I + (int)sizeof...(Is) + 1, -1>>;
            };

          public:
            static constexpr int n_intervals = N;
            using full_interval = interval<level<0, 1>, level<N, -1>>;
            using axis_interval_t = full_interval;

            template <class... Sizes,
                std::enable_if_t<sizeof...(Sizes) == N && (std::is_convertible<Sizes, int>::value && ...), int> = 0>
            axis(Sizes... sizes) : m_sizes{{static_cast<int>(sizes)...}} {}

            int interval_size(int i) const { return m_sizes.at(i); }
            std::array<int, N> const &interval_sizes() const { return m_sizes; }

            template <int I, int... Is>
            using get_interval = typename interval_of<I, Is...>::type;
        };

        // core::grid (stencil/core/grid.hpp:30-128): the (i, j) compute box and the k positions of the axis
        // splitters (splitter 0 at k = 0, splitter N at k = k_size)
        template <int NIntervals = 1>
        struct grid_t {
            int i_start = 0, i_size = 0, j_start = 0, j_size = 0, k_size = 0;
            std::array<int, NIntervals + 1> splitters{};
            static constexpr int n_intervals = NIntervals;
            std::array<int, 3> origin() const { return {i_start, j_start, 0}; }
        };
        using grid = grid_t<1>;

        namespace grid_impl {
            template <int N>
            grid_t<N> make(int i0, int ni, int j0, int nj, std::array<int, N> const &sizes) {
                grid_t<N> g;
                g.i_start = i0, g.i_size = ni, g.j_start = j0, g.j_size = nj;
                g.splitters[0] = 0;
                for (int n = 0; n < N; ++n) {
                    if (sizes[n] < 0)
                        throw std::invalid_argument("make_grid: negative interval size");
                    g.splitters[n + 1] = g.splitters[n] + sizes[n];
                }
                g.k_size = g.splitters[N];
                return g;
            }
        } // namespace grid_impl

        // make_grid (stencil/frontend/make_grid.hpp:19-42): from halo descriptors or sizes, with a k size (one
        // interval) or an axis
        inline grid make_grid(halo_descriptor const &hi, halo_descriptor const &hj, int k_size) {
            return grid_impl::make<1>(
                hi.begin(), hi.end() - hi.begin() + 1, hj.begin(), hj.end() - hj.begin() + 1, {{k_size}});
        }
        inline grid make_grid(int ni, int nj, int nk) { return grid_impl::make<1>(0, ni, 0, nj, {{nk}}); }
        template <int N, class L>
        grid_t<N> make_grid(halo_descriptor const &hi, halo_descriptor const &hj, axis<N, L> const &ax) {
            return grid_impl::make<N>(
                hi.begin(), hi.end() - hi.begin() + 1, hj.begin(), hj.end() - hj.begin() + 1, ax.interval_sizes());
        }
        template <int N, class L>
        grid_t<N> make_grid(int ni, int nj, axis<N, L> const &ax) {
            return grid_impl::make<N>(0, ni, 0, nj, ax.interval_sizes());
        }

        template <class T>
        struct global_parameter_t {
            T value;
        };
        template <class T>
        global_parameter_t<T> global_parameter(T v) {
            return {v};
        }

        // The backend tag (the role of gpu<> at stencil/gpu/entry_point.hpp:253-260): a stream and the precision
        // policy (GT355_NATIVE computes in the storage type; GT355_WIDEN: FP32 storage, FP64 compute).
        template <int Precision = GT355_NATIVE>
        struct gfx950 {
            void *stream = nullptr;
            static constexpr int precision = Precision;
        };
        using gfx950_widen = gfx950<GT355_WIDEN>;

        // The stream a run launches on: the tag's own, else the calling thread's current stream (gt355_stream_current)
        template <class Be>
        void *stream_of(Be const &be) {
            void *s = be.stream;
            if (!s)
                check(gt355_stream_current(&s), "run: current stream");
            return s;
        }

        // The path's stencil specs; fields in the order of the reference regression specs.
        namespace specs {
            struct horizontal_diffusion {}; // (in, coeff, out)                horizontal_diffusion.cpp:97-107
            struct vertical_advection {};   // (utens_stage, u_stage, wcon, u_pos, utens, dtr_stage)
                                            //                                 vertical_advection_dycore.cpp:135-148
            struct laplacian {};            // (out, in)                       laplacian.cpp:20-29, 37
            struct copy {};                 // (in, out)                       copy_stencil.cpp:23-41
            struct hdiff_vadv {};           // (in, coeff, u_stage, utens_stage, wcon, u_pos, utens, dtr_stage):
                                            // u_stage = hdiff(in, coeff), then vertical_advection of u_stage
            struct horizontal_diffusion_fused {}; // (out, in, coeff)      horizontal_diffusion_fused.cpp:97-105
            struct simple_hori_diff {};     // (coeff, in, out, crlato, crlatu) simple_hori_diff.cpp:73-80
            struct prepare_tracers {};      // (data, data_nnow, rho)      advection_pdbott_prepare_tracers.cpp:44-51
            struct copy_tuple {};           // (in0 .. inN-1, out0 .. outN-1) copy_stencil_tuple.cpp:116-150
        } // namespace specs

        template <class Be, class T>
        void run(specs::horizontal_diffusion, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &in,
            std::shared_ptr<storage::data_store<T>> const &coeff, std::shared_ptr<storage::data_store<T>> const &out) {
            auto o = g.origin();
            gt355_field fi = in->sid(o, false), fc = coeff->sid(o, false), fo = out->sid(o, true);
            check(gt355_hdiff(Be::precision, &fo, &fi, &fc, g.i_size, g.j_size, g.k_size, stream_of(be)),
                "run(horizontal_diffusion)");
        }

        template <class Be, class T>
        void run(specs::vertical_advection, Be be, grid const &g,
            std::shared_ptr<storage::data_store<T>> const &utens_stage,
            std::shared_ptr<storage::data_store<T>> const &u_stage, std::shared_ptr<storage::data_store<T>> const &wcon,
            std::shared_ptr<storage::data_store<T>> const &u_pos, std::shared_ptr<storage::data_store<T>> const &utens,
            global_parameter_t<T> dtr_stage) {
            auto o = g.origin();
            gt355_field fs = utens_stage->sid(o, true), fu = u_stage->sid(o, false), fw = wcon->sid(o, false),
                        fp = u_pos->sid(o, false), ft = utens->sid(o, false);
            check(gt355_vadv(Be::precision, &fs, &fu, &fw, &fp, &ft, (double)dtr_stage.value, g.i_size, g.j_size,
                      g.k_size, stream_of(be)),
                "run(vertical_advection)");
        }

        template <class Be, class T>
        void run(specs::laplacian, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &out,
            std::shared_ptr<storage::data_store<T>> const &in) {
            auto o = g.origin();
            gt355_field fo = out->sid(o, true), fi = in->sid(o, false);
            check(gt355_laplacian(Be::precision, &fo, &fi, g.i_size, g.j_size, g.k_size, stream_of(be)), "run(laplacian)");
        }

        template <class Be, class T>
        void run(specs::copy, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &in,
            std::shared_ptr<storage::data_store<T>> const &out) {
            static_assert(Be::precision == GT355_NATIVE, "copy has no compute type");
            auto o = g.origin();
            gt355_field fo = out->sid(o, true), fi = in->sid(o, false);
            check(gt355_copy(&fo, &fi, g.i_size, g.j_size, g.k_size, stream_of(be)), "run(copy)");
        }

        template <class Be, class T>
        void run(specs::hdiff_vadv, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &in,
            std::shared_ptr<storage::data_store<T>> const &coeff, std::shared_ptr<storage::data_store<T>> const &u_stage,
            std::shared_ptr<storage::data_store<T>> const &utens_stage,
            std::shared_ptr<storage::data_store<T>> const &wcon, std::shared_ptr<storage::data_store<T>> const &u_pos,
            std::shared_ptr<storage::data_store<T>> const &utens, global_parameter_t<T> dtr_stage) {
            auto o = g.origin();
            gt355_field fi = in->sid(o, false), fc = coeff->sid(o, false), fu = u_stage->sid(o, true),
                        fs = utens_stage->sid(o, true), fw = wcon->sid(o, false), fp = u_pos->sid(o, false),
                        ft = utens->sid(o, false);
            check(gt355_hdiff_vadv(Be::precision, &fi, &fc, &fu, &fs, &fw, &fp, &ft, (double)dtr_stage.value, g.i_size,
                      g.j_size, g.k_size, stream_of(be)),
                "run(hdiff_vadv)");
        }

        // horizontal_diffusion_fused: the hdiff out_function written with stencil functions (call<lap>, call<flx>,
        // call<fly>); it evaluates the same operations in the same order as the multistage hdiff, so it runs on
        // the same kernels.
        template <class Be, class T>
        void run(specs::horizontal_diffusion_fused, Be be, grid const &g,
            std::shared_ptr<storage::data_store<T>> const &out, std::shared_ptr<storage::data_store<T>> const &in,
            std::shared_ptr<storage::data_store<T>> const &coeff) {
            run(specs::horizontal_diffusion(), be, g, in, coeff, out);
        }

        template <class Be, class T>
        void run(specs::simple_hori_diff, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &coeff,
            std::shared_ptr<storage::data_store<T>> const &in, std::shared_ptr<storage::data_store<T>> const &out,
            std::shared_ptr<storage::data_store<T>> const &crlato,
            std::shared_ptr<storage::data_store<T>> const &crlatu) {
            auto o = g.origin();
            gt355_field fc = coeff->sid(o, false), fi = in->sid(o, false), fo = out->sid(o, true),
                        fa = crlato->sid(o, false), fu = crlatu->sid(o, false);
            check(gt355_simple_hori_diff(Be::precision, &fo, &fi, &fa, &fu, &fc, g.i_size, g.j_size, g.k_size,
                      stream_of(be)),
                "run(simple_hori_diff)");
        }

        template <class Be, class T>
        void run(specs::prepare_tracers, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &data,
            std::shared_ptr<storage::data_store<T>> const &data_nnow,
            std::shared_ptr<storage::data_store<T>> const &rho) {
            auto o = g.origin();
            gt355_field fd = data->sid(o, true), fn = data_nnow->sid(o, false), fr = rho->sid(o, false);
            check(gt355_prepare_tracers(Be::precision, &fd, &fn, 1, &fr, g.i_size, g.j_size, g.k_size, stream_of(be)),
                "run(prepare_tracers)");
        }

        // copy_stencil_tuple's unrolled functor: the first half of the stores are read, the second half written
        template <class Be, class T, class... Rest>
        void run(specs::copy_tuple, Be be, grid const &g, std::shared_ptr<storage::data_store<T>> const &first,
            Rest const &... rest) {
            static_assert(sizeof...(Rest) % 2 == 1, "copy_tuple takes N inputs followed by N outputs");
            std::vector<std::shared_ptr<storage::data_store<T>>> all{first, rest...};
            const size_t n = all.size() / 2;
            auto o = g.origin();
            std::vector<gt355_field> in(n), out(n);
            for (size_t f = 0; f < n; ++f) {
                in[f] = all[f]->sid(o, false);
                out[f] = all[n + f]->sid(o, true);
            }
            check(gt355_copy_fields(out.data(), in.data(), (int)n, g.i_size, g.j_size, g.k_size, stream_of(be)),
                "run(copy_tuple)");
        }

        // The batching form (gt355_batch_*): independent domains of uneven sizes, each computing lap =
        // laplacian(in), u_stage = hdiff(in, coeff) and the vertical advection of u_stage into utens_stage, in one
        // persistent launch per run().  The batch keeps the member stores alive; run() brings every member's
        // device copy up to date first (the lazy host/target synchronisation of data_store).
        template <class T>
        struct batch_member {
            std::shared_ptr<storage::data_store<T>> in, coeff, lap, u_stage, utens_stage, wcon, u_pos, utens;
            grid g;
        };

        template <class Be, class T>
        class batch {
            gt355_batch *m_batch = nullptr;
            std::vector<batch_member<T>> m_members;
            void *m_stream;

          public:
            batch(Be be, std::vector<batch_member<T>> members, int wg_per_cu = 2)
                : m_members(std::move(members)), m_stream(stream_of(be)) {
                std::vector<gt355_batch_domain> d(m_members.size());
                for (size_t i = 0; i < m_members.size(); ++i) {
                    auto &m = m_members[i];
                    auto o = m.g.origin();
                    d[i].in = m.in->sid(o, false);
                    d[i].coeff = m.coeff->sid(o, false);
                    d[i].lap = m.lap->sid(o, true);
                    d[i].u_stage = m.u_stage->sid(o, true);
                    d[i].utens_stage = m.utens_stage->sid(o, true);
                    d[i].wcon = m.wcon->sid(o, false);
                    d[i].u_pos = m.u_pos->sid(o, false);
                    d[i].utens = m.utens->sid(o, false);
                    d[i].ni = m.g.i_size;
                    d[i].nj = m.g.j_size;
                    d[i].nk = m.g.k_size;
                }
                check(gt355_batch_create(Be::precision, d.data(), (int)d.size(), wg_per_cu, &m_batch),
                    "batch: create");
            }
            batch(batch const &) = delete;
            batch &operator=(batch const &) = delete;
            ~batch() {
                if (m_batch)
                    gt355_batch_destroy(m_batch);
            }
            void run(global_parameter_t<T> dtr_stage) {
                for (auto &m : m_members) {
                    m.in->get_const_target_ptr();
                    m.coeff->get_const_target_ptr();
                    m.wcon->get_const_target_ptr();
                    m.u_pos->get_const_target_ptr();
                    m.utens->get_const_target_ptr();
                    m.lap->get_target_ptr();
                    m.u_stage->get_target_ptr();
                    m.utens_stage->get_target_ptr();
                }
                check(gt355_batch_run(m_batch, (double)dtr_stage.value, m_stream), "batch: run");
            }
            // the autotuner's choice of persistent grid size: (workgroups-per-CU setting, mean microseconds)
            std::pair<int, double> tune(int reps = 5) {
                int w = 0;
                double us = 0;
                check(gt355_batch_tune(m_batch, reps, m_stream, &w, &us), "batch: tune");
                return {w, us};
            }
        };

        namespace expandable_impl {
            template <class T>
            T const &member(T const &v, size_t) {
                return v;
            }
            template <class T>
            T const &member(std::vector<T> const &v, size_t m) {
                return v.at(m);
            }
            template <class T>
            size_t count(T const &) {
                return 0;
            }
            template <class T>
            size_t count(std::vector<T> const &v) {
                return v.size();
            }
        } // namespace expandable_impl

        // expandable_run<Factor>(spec, backend, grid, fields...) (stencil/frontend/expandable_run.hpp:150-183):
        // std::vector arguments are expanded member by member, scalar arguments are shared; every vector must
        // have the same size.  Factor is kept for source compatibility; the members run on the backend's stream
        // inside a deferred-run scope, so members that run the hdiff or vertical advection kernels go out as one
        // persistent launch per phase.
        // expandable_run of prepare_tracers: every member in one library call (one launch per 32 members) instead of
        // one run per member
        template <size_t Factor, class Be, class T>
        void expandable_run(specs::prepare_tracers, Be be, grid const &g,
            std::vector<std::shared_ptr<storage::data_store<T>>> const &data,
            std::vector<std::shared_ptr<storage::data_store<T>>> const &data_nnow,
            std::shared_ptr<storage::data_store<T>> const &rho) {
            static_assert(Factor > 0, "expandable_run: Factor must be positive");
            if (data.size() != data_nnow.size())
                throw std::invalid_argument("expandable_run: vector arguments differ in size");
            auto o = g.origin();
            std::vector<gt355_field> fd(data.size()), fn(data.size());
            for (size_t m = 0; m < data.size(); ++m) {
                fd[m] = data[m]->sid(o, true);
                fn[m] = data_nnow[m]->sid(o, false);
            }
            gt355_field fr = rho->sid(o, false);
            check(gt355_prepare_tracers(Be::precision, fd.data(), fn.data(), (int)data.size(), &fr, g.i_size, g.j_size,
                      g.k_size, stream_of(be)),
                "expandable_run(prepare_tracers)");
        }

        // A deferred-run scope (gt355_batch_begin / gt355_batch_end) on a stream: the hdiff and vertical advection
        // runs on that stream between construction and end() -- tag specs, and spec lambdas the engine recognises
        // as those kernels -- are recorded and then run together as persistent batches, with the results of running
        // them one after another.  end() reports errors; the destructor ends an open scope and drops them.
        class deferred {
            void *m_stream;
            bool m_open = true;

          public:
            explicit deferred(void *stream = nullptr) : m_stream(stream) {
                check(gt355_batch_begin(stream), "deferred");
            }
            template <class Be, class = decltype(std::declval<Be>().stream)>
            explicit deferred(Be const &be) : deferred(stream_of(be)) {}
            deferred(deferred const &) = delete;
            deferred &operator=(deferred const &) = delete;
            void flush() { check(gt355_batch_flush(m_stream), "deferred: flush"); }
            void end() {
                if (m_open) {
                    m_open = false;
                    check(gt355_batch_end(m_stream), "deferred: end");
                }
            }
            ~deferred() {
                if (m_open)
                    (void)gt355_batch_end(m_stream);
            }
        };

        template <size_t Factor, class Spec, class Be, int N, class... Args>
        void expandable_run(Spec spec, Be be, grid_t<N> const &g, Args const &... args) {
            static_assert(Factor > 0, "expandable_run: Factor must be positive");
            size_t n = 0;
            for (size_t c : {expandable_impl::count(args)...}) {
                if (c && n && c != n)
                    throw std::invalid_argument("expandable_run: vector arguments differ in size");
                n = std::max(n, c);
            }
            // the members' hdiff / vertical advection runs go out as persistent batches (a scope of its own unless
            // the caller has one open on the stream)
            struct scope_guard {
                void *s;
                bool own;
                ~scope_guard() {
                    if (own)
                        (void)gt355_batch_end(s);
                }
            } guard{stream_of(be), false};
            guard.own = gt355_batch_begin(guard.s) == GT355_OK;
            for (size_t m = 0; m < n; ++m)
                run(spec, be, g, expandable_impl::member(args, m)...);
            if (guard.own) {
                guard.own = false;
                check(gt355_batch_end(guard.s), "expandable_run: deferred runs");
            }
        }
    } // namespace stencil

    // transform_layout(dst, src, dims, dst_strides, src_strides) (include/gridtools/layout_transformation.hpp:64-75)
    // for device pointers: dst[sum i_d dst_strides[d]] = src[sum i_d src_strides[d]] over the box dims (any
    // sequence types of one length, 1 to 8 dimensions), asynchronously on `stream`.
    template <class T, class Dims, class DstStrides, class SrcStrides>
    void transform_layout(T *dst, T const *src, Dims const &dims, DstStrides const &dst_strides,
        SrcStrides const &src_strides, void *stream = nullptr) {
        std::vector<int64_t> d(std::begin(dims), std::end(dims)), ds(std::begin(dst_strides), std::end(dst_strides)),
            ss(std::begin(src_strides), std::end(src_strides));
        if (ds.size() != d.size() || ss.size() != d.size())
            throw std::invalid_argument("transform_layout: dims and strides differ in length");
        check(gt355_transform_layout((int)sizeof(T), dst, src, (int)d.size(), d.data(), ds.data(), ss.data(), stream),
            "transform_layout");
    }

    namespace boundaries {
        // boundary functors (boundaries/zero.hpp:29-55, value.hpp:27-67, copy.hpp:26-47)
        struct zero_boundary {};
        template <class T>
        struct value_boundary {
            T value;
        };
        struct copy_boundary {};

        // direction<I, J, K> (boundaries/direction.hpp:26-76): one of the 26 halo regions, each component minus_,
        // zero_ or plus_
        enum sign { any_ = -2, minus_ = -1, zero_ = 0, plus_ = 1 };
        template <sign I, sign J, sign K>
        struct direction {
            static constexpr sign i = I, j = J, k = K;
        };

        // default_predicate (boundaries/predicate.hpp:27-32): every direction
        struct default_predicate {
            template <class Direction>
            bool operator()(Direction) const {
                return true;
            }
        };

        // precomputed_pred as a 27-bit direction mask: bit (di+1)*9 + (dj+1)*3 + (dk+1) selects direction (di, dj, dk)
        constexpr uint32_t all_directions = 0x7FFFFFFu;
        constexpr uint32_t direction_bit(int di, int dj, int dk) { return 1u << ((di + 1) * 9 + (dj + 1) * 3 + (dk + 1)); }

        namespace bc_impl {
            // a predicate evaluated once per direction on the host (the reference evaluates it per region at apply
            // time, apply_gpu.hpp:52-83): a callable taking direction<I, J, K>, or a ready 27-bit mask
            template <class Pred>
            uint32_t mask_of(Pred const &pred) {
                if constexpr (std::is_integral<Pred>::value) {
                    return (uint32_t)pred;
                } else {
                    uint32_t m = 0;
                    auto one = [&](auto d) {
                        using D = decltype(d);
                        if (pred(d))
                            m |= direction_bit(D::i, D::j, D::k);
                    };
                    auto row = [&](auto di) {
                        constexpr sign I = decltype(di)::value;
                        one(direction<I, minus_, minus_>()), one(direction<I, minus_, zero_>());
                        one(direction<I, minus_, plus_>()), one(direction<I, zero_, minus_>());
                        if constexpr (I != zero_)
                            one(direction<I, zero_, zero_>());
                        one(direction<I, zero_, plus_>()), one(direction<I, plus_, minus_>());
                        one(direction<I, plus_, zero_>()), one(direction<I, plus_, plus_>());
                    };
                    row(std::integral_constant<sign, minus_>());
                    row(std::integral_constant<sign, zero_>());
                    row(std::integral_constant<sign, plus_>());
                    return m;
                }
            }
            template <class BC>
            constexpr bool is_predefined = std::is_same<BC, zero_boundary>::value ||
                                           std::is_same<BC, copy_boundary>::value;
            template <class T>
            constexpr bool is_predefined<value_boundary<T>> = true;

            // user functors run as device code: defined by gt355/boundaries.hpp, which a translation unit compiled
            // by hipcc includes
            template <class BC, class... Stores>
            void apply_user(
                std::array<halo_descriptor, 3> const &halos, BC const &bc, uint32_t mask, void *stream, Stores const &...);
        } // namespace bc_impl

        // boundary<BC, Arch, Predicate> (boundaries/boundary.hpp:31-79): apply(fields...) fills the halo points of
        // the 26 regions the halo descriptors define, for the directions the predicate selects.  zero_boundary,
        // value_boundary and copy_boundary run on the library's kernel; any other functor
        // (operator()(direction<I, J, K>, fields&..., i, j, k)) is instantiated as a device kernel by
        // gt355/boundaries.hpp.
        template <class BC, class Arch = storage::gpu, class Predicate = default_predicate>
        class boundary {
            std::array<halo_descriptor, 3> m_halos;
            BC m_bc;
            uint32_t m_mask;
            void *m_stream;

            template <class T>
            void launch(int kind, std::vector<void *> dst, std::vector<void *> src, double value,
                std::array<int64_t, 3> stride) const {
                int32_t h[3][5];
                for (int d = 0; d < 3; ++d) {
                    h[d][0] = m_halos[d].minus();
                    h[d][1] = m_halos[d].plus();
                    h[d][2] = m_halos[d].begin();
                    h[d][3] = m_halos[d].end();
                    h[d][4] = m_halos[d].total_length();
                }
                check(gt355_bc_apply(kind, storage::dtype_of<T>(), dst.data(), src.empty() ? nullptr : src.data(),
                          (int)dst.size(), value, h, stride.data(), m_mask, m_stream),
                    "boundary::apply");
            }

          public:
            boundary(std::array<halo_descriptor, 3> const &halos, BC const &bc, Predicate predicate = Predicate(),
                void *stream = nullptr)
                : m_halos(halos), m_bc(bc), m_mask(bc_impl::mask_of(predicate)), m_stream(stream) {}

            template <class T, class... Rest>
            void apply(std::shared_ptr<storage::data_store<T>> const &first, Rest const &... rest) const {
                if constexpr (!bc_impl::is_predefined<BC>) {
                    bc_impl::apply_user(m_halos, m_bc, m_mask, m_stream, first, rest...);
                } else if constexpr (std::is_same<BC, copy_boundary>::value) {
                    static_assert(sizeof...(Rest) == 1, "copy_boundary applies to (dst, src)");
                    auto src = std::get<0>(std::tie(rest...));
                    if (src->strides() != first->strides())
                        throw std::invalid_argument("copy_boundary: dst and src layouts differ");
                    launch<T>(2, {first->get_target_ptr()}, {(void *)src->get_const_target_ptr()}, 0.0,
                        first->strides());
                } else {
                    if (!((rest->strides() == first->strides()) && ...))
                        throw std::invalid_argument("boundary::apply: fields of different layouts");
                    std::vector<void *> ptrs{first->get_target_ptr()};
                    (ptrs.push_back(rest->get_target_ptr()), ...);
                    if constexpr (std::is_same<BC, zero_boundary>::value)
                        launch<T>(0, ptrs, {}, 0.0, first->strides());
                    else
                        launch<T>(1, ptrs, {}, (double)m_bc.value, first->strides());
                }
            }
        };

        // make_boundary<Arch>(halos, functor[, predicate]) (boundaries/boundary.hpp:74-79); the predicate is a
        // callable over direction<I, J, K> or a 27-bit direction mask
        template <class Arch = storage::gpu, class BC, class Predicate = default_predicate>
        boundary<BC, Arch, Predicate> make_boundary(std::array<halo_descriptor, 3> const &halos, BC const &bc,
            Predicate predicate = Predicate(), void *stream = nullptr) {
            return boundary<BC, Arch, Predicate>(halos, bc, predicate, stream);
        }

        // The single-rank effect of a periodic / zero-gradient halo update (gt355_halo_fill): mode per dimension
        enum halo_mode { none = 0, zero_gradient = 1, periodic = 2 };
        template <class T, class... Rest>
        void halo_fill(std::array<halo_descriptor, 3> const &halos, std::array<halo_mode, 3> modes, void *stream,
            std::shared_ptr<storage::data_store<T>> const &first, Rest const &... rest) {
            if (!((rest->strides() == first->strides()) && ...))
                throw std::invalid_argument("halo_fill: fields of different layouts");
            std::vector<void *> ptrs{first->get_target_ptr()};
            (ptrs.push_back(rest->get_target_ptr()), ...);
            int32_t h[3][5], m[3];
            for (int d = 0; d < 3; ++d) {
                h[d][0] = halos[d].minus();
                h[d][1] = halos[d].plus();
                h[d][2] = halos[d].begin();
                h[d][3] = halos[d].end();
                h[d][4] = halos[d].total_length();
                m[d] = modes[d];
            }
            auto st = first->strides();
            check(gt355_halo_fill(storage::dtype_of<T>(), ptrs.data(), (int)ptrs.size(), h, st.data(), m, stream),
                "halo_fill");
        }
    } // namespace boundaries
} // namespace gt355
