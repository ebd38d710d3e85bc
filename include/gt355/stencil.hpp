/*
 * gt355/stencil.hpp -- the generic stencil engine: user stencil specs written against the GridTools cartesian
 * frontend (accessors with extents, make_param_list, execute_parallel / execute_forward / execute_backward stages,
 * multi_pass, GT_DECLARE_TMP temporaries, interval-dispatched apply overloads, global_parameter, positional) run on
 * the device through stencil::run / run_single_stage with the gfx950 backend tag.
 *
 * The reference lowers a spec to be_api matrices (stencil/be_api.hpp:51-412) that its gpu backend turns into fused
 * kernels (stencil/gpu/entry_point.hpp:48-260).  This engine does the same job for specs the library has no
 * hand-written kernel for:
 *   * extent analysis: stages are walked backwards and each stage's extent is the union of what later stages read
 *     of its outputs (the reference's stage extents, frontend/cartesian/... -> be_api::plh_info::extent);
 *   * temporaries: GT_DECLARE_TMP placeholders get device fields sized by their extents, taken from a cached pool
 *     (the role of sid::cached_allocator, sid/allocator.hpp:46-56);
 *   * a parallel multistage launches one kernel per stage over the stage's extended domain, one thread per point
 *     (each stage completes before the next reads it, as the reference's parallel execution guarantees);
 *   * a forward / backward multistage launches one kernel in which a thread owns an (i, j) column and walks k in
 *     order, running every stage at each level; its k_cached placeholders live in a per-thread register window
 *     (kring) that slides with k, flushed to memory under the flush policy;
 *   * apply overloads are dispatched per k-level by interval (level<splitter, offset>, full_interval::modify<>,
 *     first_level, last_level), like the reference's interval matching.
 * A vertical multistage whose stages read earlier stages' outputs at a horizontal offset (the reference's need_sync)
 * runs tile by tile instead: a workgroup walks the levels of its tile and runs each stage over the tile's extended
 * box with a barrier between stages; specs in which workgroups could see each other's writes in an order the sweep
 * does not fix are rejected with an error.  Fields are the data_store stores of
 * gt355/gridtools.hpp; the engine is header-only and must be compiled by hipcc (it launches kernels).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

#include "functors.hpp"
#include "gridtools.hpp"

// Load-ahead depth of the register windows of vertical multistages: a filled window takes the level it uncovers
// from a load issued this many levels earlier, and the k sweep runs in unrolled blocks of this many levels.
// Measured with consecutive vertical multistages sharing a launch (256x256x80 FP64, the vadv and tridiagonal specs
// on the engine's kernels): 4 -> 88.8 / 87.6 us, 8 -> 82.5 / 72.7 us, 12 -> 86.3 / 72.5 us, 16 -> 88.9 / 74.5 us,
// 32 -> 99.7 / 83.7 us.
#ifndef GT355_ENGINE_PREFETCH
#define GT355_ENGINE_PREFETCH 8
#endif
// Non-temporal loads for the register windows of fields no multistage of the spec writes, read at one column: each
// value is read once, so it need not stay in the caches (0: every load with the default policy).
#ifndef GT355_ENGINE_NT
#define GT355_ENGINE_NT 1
#endif

#ifndef GT_FUNCTION
#define GT_FUNCTION __host__ __device__ __forceinline__
#endif

namespace gt355 {
    namespace stencil {
        // ------------------------------------------------------------------------------------------ extents
        template <int IMinus = 0, int IPlus = 0, int JMinus = 0, int JPlus = 0, int KMinus = 0, int KPlus = 0>
        struct extent {
            static constexpr int iminus = IMinus, iplus = IPlus, jminus = JMinus, jplus = JPlus, kminus = KMinus,
                                 kplus = KPlus;
        };

        namespace dim {
            struct i {};
            struct j {};
            struct k {};
        } // namespace dim

        // ------------------------------------------------------------------------------------------ k intervals
        // level, interval, axis and grid_t are the front end's (gt355/gridtools.hpp)
        namespace engine_ {
            // the k positions of a run's axis splitters: splitter s sits at pos[s] (0 at k = 0, the last at nk)
            struct ksplit {
                int pos[9];
            };
            __host__ __device__ constexpr int level_k(int s, int o, ksplit const &ks) {
                return ks.pos[s] + (o > 0 ? o - 1 : o);
            }
            template <class I>
            __host__ __device__ bool in_interval(int k, ksplit const &ks) {
                return level_k(I::from_t::splitter, I::from_t::offset, ks) <= k &&
                       k <= level_k(I::to_t::splitter, I::to_t::offset, ks);
            }
        } // namespace engine_

        // ------------------------------------------------------------------------------------------ accessors
        namespace cartesian {
            // in_accessor / inout_accessor (stencil/frontend/cartesian/accessor.hpp): offsets in i, j, k and, for
            // accessors of more than three dimensions, in up to three extra data dimensions (x[0..2])
            template <int I, class Extent = extent<>, int NDims = 3>
            struct in_accessor {
                static_assert(NDims >= 1 && NDims <= 6, "accessors have 1 to 6 dimensions");
                static constexpr int index = I;
                static constexpr bool is_const = true;
                static constexpr int ndims = NDims;
                using extent_t = Extent;
                int di = 0, dj = 0, dk = 0;
                int x[3] = {0, 0, 0};
                GT_FUNCTION constexpr in_accessor() = default;
                GT_FUNCTION constexpr in_accessor(int i, int j = 0, int k = 0, int x0 = 0, int x1 = 0, int x2 = 0)
                    : di(i), dj(j), dk(k), x{x0, x1, x2} {}
            };
            template <int I, class Extent = extent<>, int NDims = 3>
            struct inout_accessor {
                static_assert(NDims >= 1 && NDims <= 6, "accessors have 1 to 6 dimensions");
                static constexpr int index = I;
                static constexpr bool is_const = false;
                static constexpr int ndims = NDims;
                using extent_t = Extent;
                int di = 0, dj = 0, dk = 0;
                int x[3] = {0, 0, 0};
                GT_FUNCTION constexpr inout_accessor() = default;
                GT_FUNCTION constexpr inout_accessor(int i, int j = 0, int k = 0, int x0 = 0, int x1 = 0, int x2 = 0)
                    : di(i), dj(j), dk(k), x{x0, x1, x2} {}
            };
            // accessor<I, intent, extent> (stencil/frontend/cartesian/accessor.hpp): the same two accessor kinds
            enum class intent { in, inout };
            template <int I, intent Intent = intent::in, class Extent = extent<>, int NDims = 3>
            using accessor = std::conditional_t<Intent == intent::in, in_accessor<I, Extent, NDims>,
                inout_accessor<I, Extent, NDims>>;
            template <class... Accs>
            struct param_list_t {
                static constexpr int size = sizeof...(Accs);
            };
            template <class... Accs>
            using make_param_list = param_list_t<Accs...>;

            // cache policies are accepted and ignored: the engine keeps k-caches in memory
            namespace cache_io_policy {
                struct fill {};
                struct flush {};
            } // namespace cache_io_policy


            // expressions over accessors (stencil/frontend/cartesian/expressions/expr_base.hpp:27-77): eval(expr)
            // applies the operator to the evaluated operands, scalars passing through
            template <class Op, class L, class R>
            struct expr2 {
                L l;
                R r;
            };
            template <class Op, class A>
            struct expr1 {
                A a;
            };
            struct plus_f {
                template <class A, class B>
                GT_FUNCTION auto operator()(A const &a, B const &b) const {
                    return a + b;
                }
            };
            struct minus_f {
                template <class A, class B>
                GT_FUNCTION auto operator()(A const &a, B const &b) const {
                    return a - b;
                }
            };
            struct times_f {
                template <class A, class B>
                GT_FUNCTION auto operator()(A const &a, B const &b) const {
                    return a * b;
                }
            };
            struct divide_f {
                template <class A, class B>
                GT_FUNCTION auto operator()(A const &a, B const &b) const {
                    return a / b;
                }
            };
            struct negate_f {
                template <class A>
                GT_FUNCTION auto operator()(A const &a) const {
                    return -a;
                }
            };
            template <class T>
            struct is_expr_operand : std::false_type {};
            template <int I, class E, int N>
            struct is_expr_operand<in_accessor<I, E, N>> : std::true_type {};
            template <int I, class E, int N>
            struct is_expr_operand<inout_accessor<I, E, N>> : std::true_type {};
            template <class Op, class L, class R>
            struct is_expr_operand<expr2<Op, L, R>> : std::true_type {};
            template <class Op, class A>
            struct is_expr_operand<expr1<Op, A>> : std::true_type {};
            template <class L, class R>
            using expr_enable =
                std::enable_if_t<(is_expr_operand<L>::value || is_expr_operand<R>::value) &&
                                     (is_expr_operand<L>::value || std::is_arithmetic<L>::value) &&
                                     (is_expr_operand<R>::value || std::is_arithmetic<R>::value),
                    int>;
            template <class L, class R, expr_enable<L, R> = 0>
            GT_FUNCTION constexpr expr2<plus_f, L, R> operator+(L l, R r) {
                return {l, r};
            }
            template <class L, class R, expr_enable<L, R> = 0>
            GT_FUNCTION constexpr expr2<minus_f, L, R> operator-(L l, R r) {
                return {l, r};
            }
            template <class L, class R, expr_enable<L, R> = 0>
            GT_FUNCTION constexpr expr2<times_f, L, R> operator*(L l, R r) {
                return {l, r};
            }
            template <class L, class R, expr_enable<L, R> = 0>
            GT_FUNCTION constexpr expr2<divide_f, L, R> operator/(L l, R r) {
                return {l, r};
            }
            template <class A, std::enable_if_t<is_expr_operand<A>::value, int> = 0>
            GT_FUNCTION constexpr expr1<negate_f, A> operator-(A a) {
                return {a};
            }
            namespace expressions {}
        } // namespace cartesian

        // positional<dim>() (stencil/positional.hpp): the storage index of the evaluation point along one dimension
        template <class Dim>
        struct positional {};

        // sid::rename_dimensions<Old, New>(field) (sid/rename_dimensions.hpp): the field's dimension Old (i, j or k)
        // is seen as data dimension New (3, 4 or 5; an accessor of more than three dimensions indexes it with its
        // extra offsets), and the grid no longer moves along it -- whole_axis_access reads a whole k column at every
        // point this way.  The renamed dimension is not bounds-checked.
        namespace sid {
            template <class T, int From, int To>
            struct renamed {
                std::shared_ptr<storage::data_store<T>> store;
            };
            template <class D>
            constexpr int dim_index() {
                return std::is_same<D, dim::i>::value ? 0 : std::is_same<D, dim::j>::value ? 1 : 2;
            }
            template <class From, class To, class T>
            renamed<T, dim_index<From>(), (int)To::value> rename_dimensions(
                std::shared_ptr<storage::data_store<T>> const &s) {
                static_assert(To::value >= 3 && To::value <= 5, "rename_dimensions: the new dimension is 3, 4 or 5");
                return {s};
            }
            template <class T>
            std::shared_ptr<storage::data_store<T>> const &store_of(std::shared_ptr<storage::data_store<T>> const &s) {
                return s;
            }
            template <class T, int F, int To>
            std::shared_ptr<storage::data_store<T>> const &store_of(renamed<T, F, To> const &r) {
                return r.store;
            }
            template <class T>
            void rename(std::shared_ptr<storage::data_store<T>> const &, gt355_field &, int64_t *) {}
            template <class T, int F, int To>
            void rename(renamed<T, F, To> const &, gt355_field &f, int64_t *sx) {
                sx[To - 3] = f.stride[F];
                f.stride[F] = 0;
                f.lower[F] = -(1 << 30);
                f.upper[F] = 1 << 30;
            }
        } // namespace sid

        // ------------------------------------------------------------------------------------------ placeholders
        namespace engine_ {
            template <class T>
            struct field_kind {
                using data_t = T;
            };
            template <class T>
            struct global_kind {
                using data_t = T;
            };
            template <class D>
            struct positional_kind {
                using data_t = int;
                using dim_t = D;
            };
            template <int N, class Kind>
            struct arg_plh {
                static constexpr int key = N;
                static constexpr bool is_tmp = false;
                using kind = Kind;
            };
            template <int Id, class T>
            struct tmp_plh {
                static constexpr int key = 100000 + Id;
                static constexpr bool is_tmp = true;
                using kind = field_kind<T>;
            };

            struct parallel {};
            struct forward {};
            struct backward {};

            template <class F, class... Plhs>
            struct stage_desc {
                using functor = F;
            };

            template <class... Ts>
            struct tlist {};

            // a k-cached placeholder of a vertical multistage and its policies (cache_io_policy fill / flush)
            template <class P, bool Fill, bool Flush>
            struct kc {};
            template <class T>
            struct is_policy : std::false_type {};
            template <>
            struct is_policy<cartesian::cache_io_policy::fill> : std::true_type {};
            template <>
            struct is_policy<cartesian::cache_io_policy::flush> : std::true_type {};
            template <class A, class B>
            struct lcat;
            template <class... A, class... B>
            struct lcat<tlist<A...>, tlist<B...>> {
                using type = tlist<A..., B...>;
            };
            template <bool Fi, bool Fl, class... Ts>
            struct make_kcs {
                using type = tlist<>;
            };
            template <bool Fi, bool Fl, class T, class... Ts>
            struct make_kcs<Fi, Fl, T, Ts...> {
                using rest = typename make_kcs<Fi, Fl, Ts...>::type;
                using type = std::conditional_t<is_policy<T>::value, rest, typename lcat<tlist<kc<T, Fi, Fl>>, rest>::type>;
            };

            // Caches: the placeholders named by ij_cached (kept in LDS tiles when the multistage is fused) and the kc<>
            // entries of k_cached (kept in registers by a vertical multistage)
            template <class Exec, class Caches, class... Stages>
            struct mss_desc;
            template <class Exec, class... Cs, class... Stages>
            struct mss_desc<Exec, tlist<Cs...>, Stages...> {
                using exec = Exec;
                template <class F, class... Plhs>
                constexpr mss_desc<Exec, tlist<Cs...>, Stages..., stage_desc<F, Plhs...>> stage(F, Plhs...) const {
                    return {};
                }
                template <class... Ts>
                constexpr mss_desc<Exec, tlist<Cs..., Ts...>, Stages...> ij_cached(Ts...) const {
                    return {};
                }
                template <class... Ts>
                constexpr auto k_cached(Ts...) const {
                    constexpr bool fill = (std::is_same<Ts, cartesian::cache_io_policy::fill>::value || ...);
                    constexpr bool flush = (std::is_same<Ts, cartesian::cache_io_policy::flush>::value || ...);
                    using added = typename make_kcs<fill, flush, Ts...>::type;
                    return typename lcat_mss<typename lcat<tlist<Cs...>, added>::type>::type();
                }
                template <class NewCaches>
                struct lcat_mss {
                    using type = mss_desc<Exec, NewCaches, Stages...>;
                };
            };

            template <class... Msses>
            struct multi_pass_desc {};
        } // namespace engine_

        inline constexpr engine_::mss_desc<engine_::parallel, engine_::tlist<>> execute_parallel() { return {}; }
        inline constexpr engine_::mss_desc<engine_::forward, engine_::tlist<>> execute_forward() { return {}; }
        inline constexpr engine_::mss_desc<engine_::backward, engine_::tlist<>> execute_backward() { return {}; }
        template <class... Msses>
        constexpr engine_::multi_pass_desc<Msses...> multi_pass(Msses...) {
            return {};
        }

        namespace engine_ {
            template <class T, int C, size_t... Is>
            constexpr auto make_tmps(std::index_sequence<Is...>) {
                return std::tuple<tmp_plh<C * 64 + (int)Is, T>...>();
            }
        } // namespace engine_

#define GT355_NARGS_(_1, _2, _3, _4, _5, _6, _7, _8, _9, _10, _11, _12, _13, _14, _15, _16, N, ...) N
#define GT355_NARGS(...) GT355_NARGS_(__VA_ARGS__, 16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0)
// GT_DECLARE_TMP(type, names...) (stencil/frontend/cartesian/tmp_arg.hpp:23-25): temporaries of one value type
#define GT_DECLARE_TMP(type, ...)                                                                               \
    auto [__VA_ARGS__] = ::gt355::stencil::engine_::make_tmps<type, __COUNTER__>(                               \
        std::make_index_sequence<GT355_NARGS(__VA_ARGS__)>())
// GT_DECLARE_EXPANDABLE_TMP (stencil/frontend/cartesian/expandable_parameters): temporaries of a spec given to
// expandable_run.  expandable_run here runs the spec once per member, so each run's temporaries are the member's.
#define GT_DECLARE_EXPANDABLE_TMP(type, ...) GT_DECLARE_TMP(type, __VA_ARGS__)

        namespace engine_ {
            // ------------------------------------------------------------------------------------------ views
            template <class T>
            struct field_view {
                T *origin;
                int64_t s0, s1, s2;
                int64_t sx[3]; // strides of the extra data dimensions (0 where the field has none)
                __host__ __device__ T &at(int i, int j, int k) const { return origin[i * s0 + j * s1 + k * s2]; }
                __host__ __device__ T &at(int i, int j, int k, const int (&x)[3]) const {
                    return origin[i * s0 + j * s1 + k * s2 + x[0] * sx[0] + x[1] * sx[1] + x[2] * sx[2]];
                }
                __host__ __device__ void rebase(int, int) {}
            };
            template <class V>
            struct is_field_view : std::false_type {};
            template <class T>
            struct is_field_view<field_view<T>> : std::true_type {};
            template <class T>
            struct global_view {
                T v;
                __host__ __device__ T const &at(int, int, int) const { return v; }
                __host__ __device__ void rebase(int, int) {}
            };
            template <class D>
            struct positional_view {
                int origin;
                __host__ __device__ int at(int i, int j, int k) const {
                    return origin + (std::is_same<D, dim::i>::value ? i : std::is_same<D, dim::j>::value ? j : k);
                }
                __host__ __device__ void rebase(int, int) {}
            };

            // the dynamic LDS of a fused parallel multistage: its ij-cached temporaries, one tile each
            extern __shared__ double engine_lds[];

            // an ij-cached temporary: a (TI + rim) x (TJ + rim) tile in LDS, row length w, whose element (0, 0) is grid
            // point (tile_i0 + imin, tile_j0 + jmin); off is the tile's first element in the LDS buffer
            template <class T>
            struct lds_view {
                int off, w, imin, jmin, base;
                __device__ void rebase(int ti0, int tj0) { base = off - (ti0 + imin) - (tj0 + jmin) * w; }
                __device__ T &at(int i, int j, int) const { return reinterpret_cast<T *>(engine_lds)[base + i + j * w]; }
            };

            template <class Kind>
            struct view_of;
            template <class T>
            struct view_of<field_kind<T>> {
                using type = field_view<T>;
            };
            template <class T>
            struct view_of<global_kind<T>> {
                using type = global_view<T>;
            };
            template <class D>
            struct view_of<positional_kind<D>> {
                using type = positional_view<D>;
            };

            // a minimal device tuple
            template <int I, class T>
            struct vleaf {
                T val;
            };
            template <class Seq, class... Ts>
            struct vtuple_impl;
            template <int... Is, class... Ts>
            struct vtuple_impl<std::integer_sequence<int, Is...>, Ts...> : vleaf<Is, Ts>... {};
            template <class... Ts>
            using vtuple = vtuple_impl<std::make_integer_sequence<int, sizeof...(Ts)>, Ts...>;
            template <int I, class T>
            __host__ __device__ T &vget(vleaf<I, T> &l) {
                return l.val;
            }
            template <int I, class T>
            __host__ __device__ T const &vget(vleaf<I, T> const &l) {
                return l.val;
            }

            // a k-cached placeholder of a vertical multistage: its values live in the vertical kernel's register window
            // (kring) for that key; the evaluator reads level k + dk as element dk - KMIN of the window
            template <class T, int KMIN, int KMAX, int KEY, int IMIN = 0, int IMAX = 0, int JMIN = 0, int JMAX = 0>
            struct kring_ref {
                __host__ __device__ void rebase(int, int) {}
            };

            template <int KEY, int... Is, class... Rs>
            __host__ __device__ __forceinline__ auto &ring_by_key(vtuple_impl<std::integer_sequence<int, Is...>, Rs...> &rs) {
                constexpr int idx = ((Rs::key == KEY ? Is : 0) + ... + 0);
                return vget<idx>(rs);
            }

            // The precision policy of a read-only access (the backend's Deref hook, be_api.hpp:160-169,
            // frontend/cartesian/stage.hpp:53-63): with W (GT355_WIDEN) a float value is read as a double; otherwise
            // the value or reference passes through.  Writable accesses always see the storage type, so a store
            // narrows once.
            template <bool W, class X>
            __host__ __device__ __forceinline__ decltype(auto) widen_read(X &&x) {
                if constexpr (W && std::is_same<std::decay_t<X>, float>::value)
                    return double(x);
                else if constexpr (std::is_lvalue_reference<X>::value)
                    return static_cast<X>(x);
                else
                    return std::decay_t<X>(x);
            }

            template <class Views, class Rings = vtuple<>, bool W = false>
            struct evaluator {
                Views const &v;
                Rings &rs;
                int i, j, k;
                template <class View, class Acc>
                __host__ __device__ __forceinline__ decltype(auto) get(View const &view, Acc const &a) const {
                    if constexpr (is_field_view<View>::value)
                        return view.at(i + a.di, j + a.dj, k + a.dk, a.x);
                    else
                        return view.at(i + a.di, j + a.dj, k + a.dk);
                }
                template <class T, int KMIN, int KMAX, int KEY, int IMIN, int IMAX, int JMIN, int JMAX, class Acc>
                __host__ __device__ __forceinline__ T &get(
                    kring_ref<T, KMIN, KMAX, KEY, IMIN, IMAX, JMIN, JMAX> const &, Acc const &a) const {
                    return ring_by_key<KEY>(rs).r[(a.di - IMIN) * (JMAX - JMIN + 1) + (a.dj - JMIN)][a.dk - KMIN];
                }
                template <int I, class E, int N>
                __host__ __device__ __forceinline__ decltype(auto) operator()(cartesian::in_accessor<I, E, N> const &a) const {
                    return widen_read<W>(get(vget<I>(v), a));
                }
                template <int I, class E, int N>
                __host__ __device__ __forceinline__ decltype(auto) operator()(cartesian::inout_accessor<I, E, N> const &a) const {
                    return get(vget<I>(v), a);
                }
                template <class X>
                __host__ __device__ auto operand(X const &x) const {
                    if constexpr (std::is_arithmetic<X>::value)
                        return x;
                    else
                        return std::decay_t<decltype((*this)(x))>((*this)(x));
                }
                template <class Op, class L, class R>
                __host__ __device__ auto operator()(cartesian::expr2<Op, L, R> const &e) const {
                    return Op()(operand(e.l), operand(e.r));
                }
                template <class Op, class A>
                __host__ __device__ auto operator()(cartesian::expr1<Op, A> const &e) const {
                    return Op()(operand(e.a));
                }
            };

            // the register window of one k-cached placeholder: levels kcur + KMIN .. kcur + KMAX of the thread's column.
            // With FILL and PD > 0, the sweep runs in blocks of PD levels and each block starts by loading, for every
            // filled window, the PD levels its slides will uncover (pf[], one slot per step of the block, so every slot
            // index is static in the unrolled block): the block's loads are in flight together and the block waits for
            // memory once instead of once per level.  No load is pending across the loop's back edge.
            template <class T, int KMIN, int KMAX, int KEY, bool FILL, bool FLUSH, int PD = 0, int IMIN = 0, int IMAX = 0,
                int JMIN = 0, int JMAX = 0, bool NT = false>
            struct kring {
                static constexpr int key = KEY;
                static constexpr int pd = PD;
                // columns of the window: (i + di, j + dj) for di in [IMIN, IMAX], dj in [JMIN, JMAX] (a read-only field
                // read at horizontal offsets); a written window has the thread's own column only
                static constexpr int nc = (IMAX - IMIN + 1) * (JMAX - JMIN + 1);
                static_assert(!FLUSH || nc == 1, "a written window holds the thread's own column");
                static constexpr int ncol = nc;
                T r[nc][KMAX - KMIN + 1];
                T pf[nc][(FILL && PD > 0) ? PD : 1];
                int kcur;
                int fi, fj;        // the thread's column clamped into the fill box: every fill load is unconditional
                field_view<T> mem; // the placeholder's memory (fill / flush)
                int fb[4], wb[4];  // column boxes {i0, i1, j0, j1}: read (fill) and written (flush) by the multistage
                __device__ bool in(const int *b, int i, int j) const { return i >= b[0] && i < b[1] && j >= b[2] && j < b[3]; }
                // the level uncovered by the slide after step n of the sweep (the sweep starts at level k0)
                template <bool Up>
                __device__ static int uncovered(int n, int nk) {
                    return Up ? n + 1 + KMAX : nk - 2 - n + KMIN;
                }
                __device__ T load(int c, int k, int nk) const {
                    const int di = IMIN + c / (JMAX - JMIN + 1), dj = JMIN + c % (JMAX - JMIN + 1);
                    T *x = &mem.at(fi + di, fj + dj, min(max(k, 0), nk - 1));
                    if constexpr (NT)
                        return __builtin_nontemporal_load(x);
                    else
                        return *x;
                }
                // the whole window at the sweep's first level.  Loads are unconditional (column clamped into the fill
                // box, level into [0, nk)): values outside the box or the column are never read.
                template <bool Up>
                __device__ void start(int i, int j, int k, int nk) {
                    kcur = k;
                    if constexpr (FILL) {
                        fi = min(max(i, fb[0]), fb[1] - 1);
                        fj = min(max(j, fb[2]), fb[3] - 1);
#pragma unroll
                        for (int c = 0; c < nc; ++c)
#pragma unroll
                            for (int o = KMIN; o <= KMAX; ++o)
                                r[c][o - KMIN] = load(c, k + o, nk);
                    }
                }
                // step S of the block starting at step nb: the load of the level its slide uncovers
                template <bool Up, int S>
                __device__ void block_load(int nb, int nk) {
                    if constexpr (FILL && PD > 0) {
#pragma unroll
                        for (int c = 0; c < nc; ++c)
                            pf[c][S] = load(c, uncovered<Up>(nb + S, nk), nk);
                    }
                }
                __device__ void flush(int i, int j) {
                    if constexpr (FLUSH) {
                        if (in(wb, i, j))
                            mem.at(i, j, kcur) = r[0][-KMIN];
                    }
                }
                // move the window one level in the sweep direction and fill the level it uncovers (from slot S of the
                // block's loads)
                template <bool Up, int S = 0>
                __device__ void slide(int nk) {
#pragma unroll
                    for (int c = 0; c < nc; ++c) {
                        if constexpr (Up) {
#pragma unroll
                            for (int o = KMIN; o < KMAX; ++o)
                                r[c][o - KMIN] = r[c][o + 1 - KMIN];
                        } else {
#pragma unroll
                            for (int o = KMAX; o > KMIN; --o)
                                r[c][o - KMIN] = r[c][o - 1 - KMIN];
                        }
                    }
                    kcur += Up ? 1 : -1;
                    if constexpr (FILL) {
#pragma unroll
                        for (int c = 0; c < nc; ++c) {
                            T &dst = r[c][Up ? KMAX - KMIN : 0];
                            if constexpr (PD > 0)
                                dst = pf[c][S];
                            else
                                dst = load(c, Up ? kcur + KMAX : kcur + KMIN, nk);
                        }
                    }
                }
            };

            template <class P, class L>
            struct in_list : std::false_type {};
            template <class P, class... Ls>
            struct in_list<P, tlist<Ls...>> : std::bool_constant<(std::is_same<P, Ls>::value || ...)> {};

            // the k extent window of placeholder P over the stages of a multistage (union of its accessors' k extents)
            template <class P, class PL, class AL>
            struct kext_pair;
            template <class P, class... Ps, class... As>
            struct kext_pair<P, tlist<Ps...>, cartesian::param_list_t<As...>> {
                static constexpr int kmin = std::min({0, (std::is_same<P, Ps>::value ? As::extent_t::kminus : 0)...});
                static constexpr int kmax = std::max({0, (std::is_same<P, Ps>::value ? As::extent_t::kplus : 0)...});
            };
            template <class P, class Stage>
            struct stage_kext;
            template <class P, class F, class... Ps>
            struct stage_kext<P, stage_desc<F, Ps...>> : kext_pair<P, tlist<Ps...>, typename F::param_list> {};
            template <class P, class... Stages>
            struct mss_kext {
                static constexpr int kmin = std::min({0, stage_kext<P, Stages>::kmin...});
                static constexpr int kmax = std::max({0, stage_kext<P, Stages>::kmax...});
            };

            // the horizontal extent box of placeholder P over the stages of a multistage
            template <class P, class PL, class AL>
            struct ijext_pair;
            template <class P, class... Ps, class... As>
            struct ijext_pair<P, tlist<Ps...>, cartesian::param_list_t<As...>> {
                static constexpr int imin = std::min({0, (std::is_same<P, Ps>::value ? As::extent_t::iminus : 0)...});
                static constexpr int imax = std::max({0, (std::is_same<P, Ps>::value ? As::extent_t::iplus : 0)...});
                static constexpr int jmin = std::min({0, (std::is_same<P, Ps>::value ? As::extent_t::jminus : 0)...});
                static constexpr int jmax = std::max({0, (std::is_same<P, Ps>::value ? As::extent_t::jplus : 0)...});
            };
            template <class P, class Stage>
            struct stage_ijext;
            template <class P, class F, class... Ps>
            struct stage_ijext<P, stage_desc<F, Ps...>> : ijext_pair<P, tlist<Ps...>, typename F::param_list> {};
            template <class P, class... Stages>
            struct mss_ijext {
                static constexpr int imin = std::min({0, stage_ijext<P, Stages>::imin...});
                static constexpr int imax = std::max({0, stage_ijext<P, Stages>::imax...});
                static constexpr int jmin = std::min({0, stage_ijext<P, Stages>::jmin...});
                static constexpr int jmax = std::max({0, stage_ijext<P, Stages>::jmax...});
                static constexpr int area = (imax - imin + 1) * (jmax - jmin + 1);
            };

            // the kc<> entry of P in a cache list, if any
            template <class P, class L>
            struct find_kc {
                static constexpr bool found = false, fill = false, flush = false;
            };
            template <class P, bool Fi, bool Fl, class... Ls>
            struct find_kc<P, tlist<kc<P, Fi, Fl>, Ls...>> {
                static constexpr bool found = true, fill = Fi, flush = Fl;
            };
            template <class P, class L0, class... Ls>
            struct find_kc<P, tlist<L0, Ls...>> : find_kc<P, tlist<Ls...>> {};

            // the view context of a vertical multistage: its cache list and stages
            template <class Caches, class... Stages>
            struct vctx {};

            // the view a stage gets for placeholder P: an LDS tile if P is ij-cached by a fused multistage, a register
            // window if P is k-cached by a vertical multistage, else the placeholder's own view
            template <class P, class Ctx>
            struct view_for_t {
                using type = std::conditional_t<in_list<P, Ctx>::value, lds_view<typename P::kind::data_t>,
                    typename view_of<typename P::kind>::type>;
            };
            template <class P, class Caches, class... Stages>
            struct view_for_t<P, vctx<Caches, Stages...>> {
                using X = mss_ijext<P, Stages...>;
                using type = std::conditional_t<find_kc<P, Caches>::found,
                    kring_ref<typename P::kind::data_t, mss_kext<P, Stages...>::kmin, mss_kext<P, Stages...>::kmax, P::key,
                        X::imin, X::imax, X::jmin, X::jmax>,
                    typename view_of<typename P::kind>::type>;
            };
            // fused parallel multistages stage every field they only read, at a horizontal offset and without k offsets,
            // into an LDS tile per workgroup (one coalesced load), so that the stages' neighbour reads hit LDS
            template <class P, class PL, class AL>
            struct acc_props;
            template <class P, class... Ps, class... As>
            struct acc_props<P, tlist<Ps...>, cartesian::param_list_t<As...>> {
                static constexpr bool written = ((std::is_same<P, Ps>::value && !As::is_const) || ...);
                static constexpr bool ij = ((std::is_same<P, Ps>::value &&
                                                (As::extent_t::iminus || As::extent_t::iplus || As::extent_t::jminus ||
                                                    As::extent_t::jplus)) ||
                                            ...);
                static constexpr bool kofs =
                    ((std::is_same<P, Ps>::value && (As::extent_t::kminus || As::extent_t::kplus)) || ...);
                // accessed with more than three dimensions (a data dimension or a renamed axis)
                static constexpr bool nd = ((std::is_same<P, Ps>::value && As::ndims > 3) || ...);
            };
            template <class P, class Stage>
            struct stage_props;
            template <class P, class F, class... Ps>
            struct stage_props<P, stage_desc<F, Ps...>> : acc_props<P, tlist<Ps...>, typename F::param_list> {};
            template <class P, class... Stages>
            struct auto_stage {
                static constexpr bool value =
                    std::is_same<typename view_of<typename P::kind>::type, field_view<typename P::kind::data_t>>::value &&
                    !(stage_props<P, Stages>::written || ...) && (stage_props<P, Stages>::ij || ...) &&
                    !(stage_props<P, Stages>::kofs || ...);
            };

            template <class Caches, class... Stages>
            struct fctx {};
            template <class P, class Caches, class... Stages>
            struct view_for_t<P, fctx<Caches, Stages...>> {
                using type = std::conditional_t<in_list<P, Caches>::value || auto_stage<P, Stages...>::value,
                    lds_view<typename P::kind::data_t>, typename view_of<typename P::kind>::type>;
            };

            template <class P, class Ctx>
            using view_for = typename view_for_t<P, Ctx>::type;

            // the distinct placeholders of a list of stages, and those a fused multistage stages into LDS
            template <class L, class P>
            struct tset_add {
                using type = std::conditional_t<in_list<P, L>::value, L, typename lcat<L, tlist<P>>::type>;
            };
            template <class L, class... Ps>
            struct tset_add_all {
                using type = L;
            };
            template <class L, class P, class... Ps>
            struct tset_add_all<L, P, Ps...> : tset_add_all<typename tset_add<L, P>::type, Ps...> {};
            template <class L, class... Stages>
            struct plhs_of_stages {
                using type = L;
            };
            template <class L, class F, class... Ps, class... Stages>
            struct plhs_of_stages<L, stage_desc<F, Ps...>, Stages...>
                : plhs_of_stages<typename tset_add_all<L, Ps...>::type, Stages...> {};
            template <class L, class Stages>
            struct staged_of;
            template <class... Stages>
            struct staged_of<tlist<>, tlist<Stages...>> {
                using type = tlist<>;
            };
            template <class P, class... Ps, class... Stages>
            struct staged_of<tlist<P, Ps...>, tlist<Stages...>> {
                using rest = typename staged_of<tlist<Ps...>, tlist<Stages...>>::type;
                using type =
                    std::conditional_t<auto_stage<P, Stages...>::value, typename lcat<tlist<P>, rest>::type, rest>;
            };
            template <class... Stages>
            using staged_inputs = typename staged_of<typename plhs_of_stages<tlist<>, Stages...>::type, tlist<Stages...>>::type;

            // A vertical multistage needs barriers between its stages when a stage reads, at a horizontal offset, a
            // placeholder that an earlier stage of the multistage writes (the reference's need_sync,
            // stencil/core/need_sync.hpp:21-56): the values at the neighbouring columns come from other threads.
            // Such a multistage runs tile by tile, every workgroup walking the levels of its tile in the sweep's order.
            template <class Written, class... Stages>
            struct hsync : std::false_type {};
            template <class Written, class F, class... Ps, class... Stages>
            struct hsync<Written, stage_desc<F, Ps...>, Stages...> {
                using S = stage_desc<F, Ps...>;
                static constexpr bool here = ((in_list<Ps, Written>::value && stage_props<Ps, S>::ij) || ...);
                // the placeholders S writes join the written set (void stands for the ones it only reads)
                using next = typename tset_add_all<Written,
                    std::conditional_t<stage_props<Ps, S>::written, Ps, void>...>::type;
                static constexpr bool value = here || hsync<next, Stages...>::value;
            };
            // the ij_cached entries of a cache list (its kc<> entries dropped)
            template <class L>
            struct ij_entries {
                using type = tlist<>;
            };
            template <class C, class... Cs>
            struct ij_entries<tlist<C, Cs...>> {
                using type = typename lcat<tlist<C>, typename ij_entries<tlist<Cs...>>::type>::type;
            };
            template <class P, bool Fi, bool Fl, class... Cs>
            struct ij_entries<tlist<kc<P, Fi, Fl>, Cs...>> : ij_entries<tlist<Cs...>> {};

            // The register windows a vertical multistage gets besides the spec's own k-caches: every field placeholder
            // its stages access in three dimensions and either only read (at zero horizontal offset, or at up to four
            // horizontal offsets, with one window per column) or, at zero horizontal offset, write and also access at
            // a k offset (a recurrence).  Such a window is the reference's k_cache with the
            // fill policy (and flush, if written): the same values in registers, so its reads are served by the
            // load-ahead slots instead of one memory round trip per level.  A placeholder written at its own level only
            // keeps its direct stores.
            template <class P, class... Stages>
            struct mss_acc {
                static constexpr bool written = (stage_props<P, Stages>::written || ...);
                static constexpr bool ij = (stage_props<P, Stages>::ij || ...);
                static constexpr bool kofs = (stage_props<P, Stages>::kofs || ...);
                static constexpr bool nd = (stage_props<P, Stages>::nd || ...);
                static constexpr bool field =
                    std::is_same<typename view_of<typename P::kind>::type, field_view<typename P::kind::data_t>>::value;
            };
            template <class Caches, class Plhs, class... Stages>
            struct auto_vc {
                using type = Caches;
            };
            template <class Caches, class P, class... Ps, class... Stages>
            struct auto_vc<Caches, tlist<P, Ps...>, Stages...> {
                using A = mss_acc<P, Stages...>;
                // read-only fields may also be read at up to four horizontal offsets (one window per column)
                static constexpr bool take = A::field && !A::nd && !find_kc<P, Caches>::found &&
                                             (A::ij ? (!A::written && mss_ijext<P, Stages...>::area <= 4)
                                                    : (!A::written || A::kofs));
                using next = std::conditional_t<take, typename lcat<Caches, tlist<kc<P, true, A::written>>>::type, Caches>;
                using type = typename auto_vc<next, tlist<Ps...>, Stages...>::type;
            };
            template <class Caches, class... Stages>
            using auto_vcaches = typename auto_vc<Caches, typename plhs_of_stages<tlist<>, Stages...>::type, Stages...>::type;


            // one staged input: its memory and its LDS tile (first element, row length, tile box relative to the tile)
            template <class T>
            struct stage_in {
                field_view<T> src;
                int off, w, imin, imax, jmin, jmax;
            };
            template <class T>
            __device__ void load_stage_in(
                stage_in<T> const &s, int ti0, int tj0, int TI, int TJ, int ni, int nj, int k) {
                const int i_lo = ti0 + s.imin, i_hi = min(ti0 + TI + s.imax, ni + s.imax);
                const int j_lo = tj0 + s.jmin, j_hi = min(tj0 + TJ + s.jmax, nj + s.jmax);
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                T *lds = reinterpret_cast<T *>(engine_lds) + s.off;
                for (int j = j_lo + wave; j < j_hi; j += 4)
                    for (int i = i_lo + lane; i < i_hi; i += 64)
                        lds[(j - j_lo) * s.w + (i - i_lo)] = s.src.at(i, j, k);
            }
            template <int... Is, class... Ss>
            __device__ void load_stage_ins(vtuple_impl<std::integer_sequence<int, Is...>, Ss...> const &si, int ti0,
                int tj0, int TI, int TJ, int ni, int nj, int k) {
                (load_stage_in(vget<Is>(si), ti0, tj0, TI, TJ, ni, nj, k), ...);
            }
            template <class L>
            struct stage_in_tuple;
            template <class... Ps>
            struct stage_in_tuple<tlist<Ps...>> {
                using type = vtuple<stage_in<typename Ps::kind::data_t>...>;
            };

            // One level of a staged input in registers: a thread's points of the tile are its lane's column in rows
            // wave, wave + 4, ... (at most SROWS of them; the host checks that the tile fits one pass of the 64 lanes
            // and 4 * SROWS rows).  Loads are unconditional, clamped into the tile, so that none waits on a branch.
            constexpr int SROWS = 4;
            template <class T>
            struct stage_buf {
                T v[SROWS];
            };
            template <class L>
            struct stage_buf_tuple;
            template <class... Ps>
            struct stage_buf_tuple<tlist<Ps...>> {
                using type = vtuple<stage_buf<typename Ps::kind::data_t>...>;
            };
            template <class T>
            __device__ __forceinline__ void fetch_stage_in(
                stage_in<T> const &s, stage_buf<T> &b, int ti0, int tj0, int TI, int TJ, int ni, int nj, int k) {
                const int i_lo = ti0 + s.imin, i_hi = min(ti0 + TI + s.imax, ni + s.imax);
                const int j_lo = tj0 + s.jmin, j_hi = min(tj0 + TJ + s.jmax, nj + s.jmax);
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                const int i = min(i_lo + lane, i_hi - 1);
#pragma unroll
                for (int r = 0; r < SROWS; ++r)
                    b.v[r] = s.src.at(i, min(j_lo + wave + 4 * r, j_hi - 1), k);
            }
            template <class T>
            __device__ __forceinline__ void put_stage_in(
                stage_in<T> const &s, stage_buf<T> const &b, int ti0, int tj0, int TI, int TJ, int ni, int nj) {
                const int i_lo = ti0 + s.imin, i_hi = min(ti0 + TI + s.imax, ni + s.imax);
                const int j_lo = tj0 + s.jmin, j_hi = min(tj0 + TJ + s.jmax, nj + s.jmax);
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                T *lds = reinterpret_cast<T *>(engine_lds) + s.off;
#pragma unroll
                for (int r = 0; r < SROWS; ++r) {
                    const int j = j_lo + wave + 4 * r, i = i_lo + lane;
                    if (j < j_hi && i < i_hi)
                        lds[(j - j_lo) * s.w + (i - i_lo)] = b.v[r];
                }
            }
            template <int... Is, class... Ss, class Bs>
            __device__ void fetch_stage_ins(vtuple_impl<std::integer_sequence<int, Is...>, Ss...> const &si, Bs &bs,
                int ti0, int tj0, int TI, int TJ, int ni, int nj, int k) {
                (fetch_stage_in(vget<Is>(si), vget<Is>(bs), ti0, tj0, TI, TJ, ni, nj, k), ...);
            }
            template <int... Is, class... Ss, class Bs>
            __device__ void put_stage_ins(vtuple_impl<std::integer_sequence<int, Is...>, Ss...> const &si,
                Bs const &bs, int ti0, int tj0, int TI, int TJ, int ni, int nj) {
                (put_stage_in(vget<Is>(si), vget<Is>(bs), ti0, tj0, TI, TJ, ni, nj), ...);
            }

            template <class Stage, class Caches = tlist<>>
            struct stage_views;
            template <class F, class... Plhs, class Caches>
            struct stage_views<stage_desc<F, Plhs...>, Caches> {
                using type = vtuple<view_for<Plhs, Caches>...>;
            };

            template <int... Is, class... Vs>
            __device__ void rebase_views(vtuple_impl<std::integer_sequence<int, Is...>, Vs...> &t, int ti0, int tj0) {
                (vget<Is>(t).rebase(ti0, tj0), ...);
            }

            // ------------------------------------------------------------------------------------------ dispatch
            template <class F, class E, class = void>
            struct has_plain_apply : std::false_type {};
            template <class F, class E>
            struct has_plain_apply<F, E, std::void_t<decltype(F::apply(std::declval<E &>()))>> : std::true_type {};
            template <class F, class E, class I, class = void>
            struct has_interval_apply : std::false_type {};
            template <class F, class E, class I>
            struct has_interval_apply<F, E, I, std::void_t<decltype(F::apply(std::declval<E &>(), I()))>>
                : std::true_type {};

            template <class A, class B>
            struct tcat;
            template <class... A, class... B>
            struct tcat<tlist<A...>, tlist<B...>> {
                using type = tlist<A..., B...>;
            };

            // candidate intervals of an axis with NSPL splitters: every [level<s1, o1>, level<s2, o2>] with s in
            // [0, NSPL), o in +-1..+-4 and the first level not above the second.  Level n of the enumeration is
            // splitter n / 8 with offsets -4, -3, -2, -1, 1, 2, 3, 4 in k order.
            constexpr int cand_offset(int n) { return n < 4 ? n - 4 : n - 3; }
            template <int N>
            using cand_level = level<N / 8, cand_offset(N % 8)>;
            template <int M, int N>
            struct cand {
                static constexpr bool valid = N / M <= N % M;
                using type = interval<cand_level<N / M>, cand_level<N % M>>;
            };

            // runs F at level k: the plain apply everywhere, else the first interval overload whose interval
            // contains k (none: nothing to do at this level), like the reference's interval matching
            template <class F, class E, int M, int N>
            __host__ __device__ __forceinline__ void try_interval(E &e, int k, ksplit const &ks, bool &done) {
                if constexpr (cand<M, N>::valid) {
                    using I = typename cand<M, N>::type;
                    if constexpr (has_interval_apply<F, E, I>::value) {
                        if (!done && in_interval<I>(k, ks)) {
                            F::apply(e, I());
                            done = true;
                        }
                    }
                }
            }
            template <class F, class E, int M, int... Ns>
            __host__ __device__ __forceinline__ void apply_at(E &e, int k, ksplit const &ks, std::integer_sequence<int, Ns...>) {
                if constexpr (has_plain_apply<F, E>::value) {
                    F::apply(e);
                } else {
                    bool done = false;
                    (try_interval<F, E, M, Ns>(e, k, ks, done), ...);
                }
            }
            template <class F, int NSPL, bool W, class Views, class Rings>
            __host__ __device__ __forceinline__ void run_stage_at(
                Views const &v, Rings &rs, int i, int j, int k, ksplit const &ks) {
                using E = evaluator<Views, Rings, W>;
                E e{v, rs, i, j, k};
                constexpr int M = NSPL * 8;
                apply_at<F, E, M>(e, k, ks, std::make_integer_sequence<int, M * M>());
            }
            template <class F, int NSPL, bool W, class Views>
            __host__ __device__ __forceinline__ void run_stage_at(Views const &v, int i, int j, int k, ksplit const &ks) {
                vtuple<> none;
                run_stage_at<F, NSPL, W>(v, none, i, j, k, ks);
            }

            // ------------------------------------------------------------------------------------------ kernels
            template <class F, int NSPL, bool W, class Views>
            __global__ void __launch_bounds__(256) parallel_stage_kernel(
                Views v, int i0, int ni, int j0, int nj, ksplit ks) {
                const int i = i0 + (int)(blockIdx.x * 64 + (threadIdx.x & 63));
                const int j = j0 + (int)(blockIdx.y * 4 + (threadIdx.x >> 6));
                const int k = (int)blockIdx.z;
                if (i >= i0 + ni || j >= j0 + nj)
                    return;
                run_stage_at<F, NSPL, W>(v, i, j, k, ks);
            }

            // one vertical multistage: stage s runs on columns inside its own extent box dom[s] = {i0, i1, j0, j1}
            template <class Exec, int NSPL, bool W, class StageList, class ViewsTuple>
            struct vertical_runner;
            template <class Exec, int NSPL, bool W, class... Stages, class... Views>
            struct vertical_runner<Exec, NSPL, W, tlist<Stages...>, vtuple<Views...>> {
                template <class Rings, int... Ss>
                __device__ static void level(vtuple<Views...> const &v, Rings &rs, const int (*dom)[4], int i, int j, int k,
                    ksplit const &ks, std::integer_sequence<int, Ss...>) {
                    ((i >= dom[Ss][0] && i < dom[Ss][1] && j >= dom[Ss][2] && j < dom[Ss][3]
                             ? run_stage_at<typename Stages::functor, NSPL, W>(vget<Ss>(v), rs, i, j, k, ks)
                             : void()),
                        ...);
                }
            };
            template <int NS>
            struct domains {
                int d[NS][4];
            };
            // the register windows of a vertical multistage: one kring per kc<> entry of its cache list
            // Written: the placeholders some multistage of the spec writes
            template <class Written, class Entry, class... Stages>
            struct ring_of;
            template <class Written, class P, bool Fi, bool Fl, class... Stages>
            struct ring_of<Written, kc<P, Fi, Fl>, Stages...> {
                using X = mss_ijext<P, Stages...>;
                static constexpr bool nt = GT355_ENGINE_NT && Fi && X::area == 1 && !in_list<P, Written>::value;
                using type = kring<typename P::kind::data_t, mss_kext<P, Stages...>::kmin, mss_kext<P, Stages...>::kmax,
                    P::key, Fi, Fl, Fi ? GT355_ENGINE_PREFETCH : 0, X::imin, X::imax, X::jmin, X::jmax, nt>;
            };
            template <class Written, class Caches, class StageList>
            struct ring_list {
                using type = tlist<>;
            };
            template <class Written, class C0, class... Cs, class... Stages>
            struct ring_list<Written, tlist<C0, Cs...>, tlist<Stages...>> {
                using rest = typename ring_list<Written, tlist<Cs...>, tlist<Stages...>>::type;
                using type = rest;
            };
            template <class Written, class P, bool Fi, bool Fl, class... Cs, class... Stages>
            struct ring_list<Written, tlist<kc<P, Fi, Fl>, Cs...>, tlist<Stages...>> {
                using rest = typename ring_list<Written, tlist<Cs...>, tlist<Stages...>>::type;
                using type =
                    typename lcat<tlist<typename ring_of<Written, kc<P, Fi, Fl>, Stages...>::type>, rest>::type;
            };
            // the placeholders written by some stage of a list of multistages
            template <class L, class... Stages>
            struct written_by {
                using type = L;
            };
            template <class L, class F, class... Ps, class... Stages>
            struct written_by<L, stage_desc<F, Ps...>, Stages...>
                : written_by<typename tset_add_all<L,
                                 std::conditional_t<stage_props<Ps, stage_desc<F, Ps...>>::written, Ps, void>...>::type,
                      Stages...> {};
            template <class L, class Mss>
            struct mss_written;
            template <class L, class E, class C, class... Stages>
            struct mss_written<L, mss_desc<E, C, Stages...>> : written_by<L, Stages...> {};
            template <class L, class Msses>
            struct spec_written;
            template <class L>
            struct spec_written<L, tlist<>> {
                using type = L;
            };
            template <class L, class M0, class... Ms>
            struct spec_written<L, tlist<M0, Ms...>> : spec_written<typename mss_written<L, M0>::type, tlist<Ms...>> {};
            template <class L>
            struct to_vtuple;
            template <class... Ts>
            struct to_vtuple<tlist<Ts...>> {
                using type = vtuple<Ts...>;
            };

            template <class F, int... Is, class... Rs>
            __device__ void for_rings(vtuple_impl<std::integer_sequence<int, Is...>, Rs...> &rs, F &&f) {
                (f(vget<Is>(rs)), ...);
            }

            // the steps of one block of the sweep, each with its static slot
            template <bool Up, int PD, int... Ss, class F>
            __device__ __forceinline__ void step_block(std::integer_sequence<int, Ss...>, F &&f) {
                (f(std::integral_constant<int, Ss>()), ...);
            }

            // one column's sweep of a vertical multistage over every level
            template <class Exec, int NSPL, bool W, class StageList, class ViewsTuple, class RingSet, int NS>
            __device__ __forceinline__ void vertical_sweep(
                ViewsTuple const &v0, RingSet const &rs0, domains<NS> const &dom, int i, int j, ksplit const &ks) {
                using R = vertical_runner<Exec, NSPL, W, StageList, ViewsTuple>;
                constexpr bool up = !std::is_same<Exec, backward>::value;
                const int nk = ks.pos[NSPL - 1];
                RingSet rs = rs0;
                const int k0 = up ? 0 : nk - 1;
                for_rings(rs, [&](auto &r) { r.template start<up>(i, j, k0, nk); });
                // blocks of PD levels, unrolled, so that step S of a block uses the windows' load slot S
                constexpr int PD = GT355_ENGINE_PREFETCH > 0 ? GT355_ENGINE_PREFETCH : 1;
                for (int nb = 0; nb < nk; nb += PD) {
                    step_block<up, PD>(std::make_integer_sequence<int, PD>(), [&](auto sc) {
                        for_rings(rs, [&](auto &r) { r.template block_load<up, decltype(sc)::value>(nb, nk); });
                    });
                    // the block's loads stay together at its start
                    __builtin_amdgcn_sched_barrier(0);
                    step_block<up, PD>(std::make_integer_sequence<int, PD>(), [&](auto sc) {
                        constexpr int S = decltype(sc)::value;
                        const int n = nb + S;
                        if (n < nk) {
                            const int k = up ? n : nk - 1 - n;
                            R::level(v0, rs, dom.d, i, j, k, ks, std::make_integer_sequence<int, NS>());
                            for_rings(rs, [&](auto &r) { r.flush(i, j); });
                            for_rings(rs, [&](auto &r) { r.template slide<up, S>(nk); });
                        }
                    });
                }
            }

            // one vertical multistage: a thread per column of the box [i0, i0 + ni) x [j0, j0 + nj)
            template <class Exec, int NSPL, bool W, class StageList, class ViewsTuple, class RingSet, int NS>
            __global__ void __launch_bounds__(256) vertical_kernel(
                ViewsTuple v0, RingSet rs0, domains<NS> dom, int i0, int ni, int j0, int nj, ksplit ks) {
                const int i = i0 + (int)(blockIdx.x * 64 + (threadIdx.x & 63));
                const int j = j0 + (int)(blockIdx.y * 4 + (threadIdx.x >> 6));
                if (i >= i0 + ni || j >= j0 + nj)
                    return;
                vertical_sweep<Exec, NSPL, W, StageList, ViewsTuple, RingSet, NS>(v0, rs0, dom, i, j, ks);
            }

            // a column box [i0, i1) x [j0, j1)
            struct cbox {
                int i0, i1, j0, j1;
                __device__ bool in(int i, int j) const { return i >= i0 && i < i1 && j >= j0 && j < j1; }
            };
            // two consecutive vertical multistages in one launch, as the reference fuses multistages whose data pass
            // at zero horizontal offset (stencil/gpu/entry_point.hpp:107-124): a thread per column of the union of
            // their boxes runs the first multistage's sweep where its box holds the column, then the second's.  The
            // launcher fuses only when every placeholder one of the two writes and the other accesses is accessed at
            // the column itself, so a column's second sweep depends on nothing but its own first sweep.
            template <class EA, class SA, class VA, class RA, int NA, class EB, class SB, class VB, class RB, int NB,
                int NSPL, bool W>
            __global__ void __launch_bounds__(256) vertical_pair_kernel(VA va, RA ra, domains<NA> da, cbox ba, VB vb,
                RB rb, domains<NB> db, cbox bb, int i0, int ni, int j0, int nj, ksplit ks) {
                const int i = i0 + (int)(blockIdx.x * 64 + (threadIdx.x & 63));
                const int j = j0 + (int)(blockIdx.y * 4 + (threadIdx.x >> 6));
                if (i >= i0 + ni || j >= j0 + nj)
                    return;
                if (ba.in(i, j))
                    vertical_sweep<EA, NSPL, W, SA, VA, RA, NA>(va, ra, da, i, j, ks);
                if (bb.in(i, j))
                    vertical_sweep<EB, NSPL, W, SB, VB, RB, NB>(vb, rb, db, i, j, ks);
            }

            // one fused parallel multistage: a workgroup owns a TI x TJ tile of one k-level and runs every stage over the
            // tile extended by the stage's extent (clamped to the stage's domain), with a barrier between stages; the
            // ij-cached temporaries live in LDS, so stage s+1 reads what stage s left in the tile
            template <class F, int NSPL, bool W, class V>
            __device__ void stage_tile(V const &v, const int *e, int ti0, int tj0, int TI, int TJ, int ni, int nj, int k,
                ksplit const &ks) {
                const int i_lo = ti0 + e[0], i_hi = min(ti0 + TI + e[1], ni + e[1]);
                const int j_lo = tj0 + e[2], j_hi = min(tj0 + TJ + e[3], nj + e[3]);
                // 64 lanes along i (coalesced rows), the 4 wavefronts of the workgroup along j
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                for (int j = j_lo + wave; j < j_hi; j += 4)
                    for (int i = i_lo + lane; i < i_hi; i += 64)
                        run_stage_at<F, NSPL, W>(v, i, j, k, ks);
                __syncthreads();
            }
            template <int NSPL, bool W, class StageList, class ViewsTuple>
            struct fused_runner;
            template <int NSPL, bool W, class... Stages, class... Views>
            struct fused_runner<NSPL, W, tlist<Stages...>, vtuple<Views...>> {
                template <int... Ss>
                __device__ static void run(vtuple<Views...> &v, const int (*ext)[4], int ti0, int tj0, int TI, int TJ,
                    int ni, int nj, int k, ksplit const &ks, std::integer_sequence<int, Ss...>) {
                    (rebase_views(vget<Ss>(v), ti0, tj0), ...);
                    (stage_tile<typename Stages::functor, NSPL, W>(vget<Ss>(v), ext[Ss], ti0, tj0, TI, TJ, ni, nj, k, ks),
                        ...);
                }
            };
            template <int NSPL, bool W, class StageList, class ViewsTuple, class StageIns, int NS>
            __global__ void __launch_bounds__(256) fused_kernel(
                ViewsTuple v0, StageIns si, domains<NS> ext, int ni, int nj, ksplit ks, int tiles_i, int TI, int TJ) {
                const int tile = (int)blockIdx.x;
                const int ti0 = (tile % tiles_i) * TI, tj0 = (tile / tiles_i) * TJ;
                load_stage_ins(si, ti0, tj0, TI, TJ, ni, nj, (int)blockIdx.z);
                __syncthreads();
                ViewsTuple v = v0;
                fused_runner<NSPL, W, StageList, ViewsTuple>::run(
                    v, ext.d, ti0, tj0, TI, TJ, ni, nj, (int)blockIdx.z, ks, std::make_integer_sequence<int, NS>());
            }
            // The same over a chunk of KC levels per workgroup: the staged inputs of level k + 1 are fetched into
            // registers before level k's stages run and written to their LDS tiles after (the last stage of a level
            // ends with a barrier), so their loads overlap the level's compute, and a workgroup's setup is paid once
            // per chunk.
            template <int NSPL, bool W, class StageList, class ViewsTuple, class StageIns, class StageBufs, int NS>
            __global__ void __launch_bounds__(256) fused_kc_kernel(ViewsTuple v0, StageIns si, domains<NS> ext, int ni,
                int nj, ksplit ks, int tiles_i, int TI, int TJ, int KC, int nk) {
                const int tile = (int)blockIdx.x;
                const int ti0 = (tile % tiles_i) * TI, tj0 = (tile / tiles_i) * TJ;
                const int k0 = (int)blockIdx.z * KC, k1 = min(k0 + KC, nk);
                StageBufs buf;
                fetch_stage_ins(si, buf, ti0, tj0, TI, TJ, ni, nj, k0);
                put_stage_ins(si, buf, ti0, tj0, TI, TJ, ni, nj);
                __syncthreads();
                for (int k = k0; k < k1; ++k) {
                    const bool more = k + 1 < k1;
                    if (more)
                        fetch_stage_ins(si, buf, ti0, tj0, TI, TJ, ni, nj, k + 1);
                    ViewsTuple v = v0;
                    fused_runner<NSPL, W, StageList, ViewsTuple>::run(
                        v, ext.d, ti0, tj0, TI, TJ, ni, nj, k, ks, std::make_integer_sequence<int, NS>());
                    if (more) {
                        put_stage_ins(si, buf, ti0, tj0, TI, TJ, ni, nj);
                        __syncthreads();
                    }
                }
            }

            // one vertical multistage with horizontal dependencies between its stages: a workgroup owns a TI x TJ tile
            // and walks the levels in the sweep's order; at each level it runs every stage over the tile extended by
            // the stage's extent, with a barrier between stages, as the fused parallel form does for its one level.
            // Stages that run on extended boxes compute the points of the neighbouring tiles' rims again (the same
            // values); the host checks that no read can see another workgroup's write in an order the sweep does not
            // fix.
            template <class Exec, int NSPL, bool W, class StageList, class ViewsTuple, class StageIns, int NS>
            __global__ void __launch_bounds__(256) vfused_kernel(
                ViewsTuple v0, StageIns si, domains<NS> ext, int ni, int nj, ksplit ks, int tiles_i, int TI, int TJ) {
                const int tile = (int)blockIdx.x;
                const int ti0 = (tile % tiles_i) * TI, tj0 = (tile / tiles_i) * TJ;
                constexpr bool up = !std::is_same<Exec, backward>::value;
                const int nk = ks.pos[NSPL - 1];
                for (int n = 0; n < nk; ++n) {
                    const int k = up ? n : nk - 1 - n;
                    // the previous level's last stage ended with a barrier, so the staging tiles are free
                    load_stage_ins(si, ti0, tj0, TI, TJ, ni, nj, k);
                    __syncthreads();
                    ViewsTuple v = v0;
                    fused_runner<NSPL, W, StageList, ViewsTuple>::run(
                        v, ext.d, ti0, tj0, TI, TJ, ni, nj, k, ks, std::make_integer_sequence<int, NS>());
                }
            }

            // ------------------------------------------------------------------------------------------ host side
            struct ext_rt {
                int imin = 0, imax = 0, jmin = 0, jmax = 0, kmin = 0, kmax = 0;
                bool set = false;
                void merge(ext_rt const &o) {
                    if (!o.set)
                        return;
                    if (!set) {
                        *this = o;
                        return;
                    }
                    imin = std::min(imin, o.imin), imax = std::max(imax, o.imax);
                    jmin = std::min(jmin, o.jmin), jmax = std::max(jmax, o.jmax);
                    kmin = std::min(kmin, o.kmin), kmax = std::max(kmax, o.kmax);
                }
            };
            struct access_rt {
                int key;
                bool write;
                ext_rt ext;
                int elem; // element size of the placeholder's values
            };
            struct stage_rt {
                int mss;
                bool vertical;
                std::vector<access_rt> acc;
                ext_rt ext; // the stage's own extent (horizontal part used)
            };

            template <class Acc>
            access_rt make_access(int key, int elem) {
                using E = typename Acc::extent_t;
                access_rt a{key, !Acc::is_const, {}, elem};
                a.ext = {E::iminus, E::iplus, E::jminus, E::jplus, E::kminus, E::kplus, true};
                return a;
            }
            template <class F, class... Plhs, class... Accs>
            std::vector<access_rt> stage_accesses(cartesian::param_list_t<Accs...>) {
                static_assert(sizeof...(Accs) == sizeof...(Plhs), "a stage binds one placeholder per accessor");
                constexpr int keys[] = {Plhs::key...};
                constexpr int elems[] = {(int)sizeof(typename Plhs::kind::data_t)...};
                constexpr int idx[] = {Accs::index...};
                for (int n = 0; n < (int)sizeof...(Accs); ++n)
                    if (idx[n] != n)
                        throw std::invalid_argument("param_list accessors must have indices 0, 1, 2, ... in order");
                return {make_access<Accs>(keys[Accs::index], elems[Accs::index])...};
            }

            template <class Mss>
            struct mss_traits;
            template <class Exec, class Caches, class... Stages>
            struct mss_traits<mss_desc<Exec, Caches, Stages...>> {
                using exec = Exec;
                using caches = Caches;
                using stages = tlist<Stages...>;
                static constexpr int n = sizeof...(Stages);
                // a parallel multistage of several stages, or with ij-caches, runs as one fused launch
                static constexpr bool fused =
                    std::is_same<Exec, parallel>::value && (sizeof...(Stages) > 1 || !std::is_same<Caches, tlist<>>::value);
                // a vertical multistage with horizontal dependencies between its stages runs tile by tile
                static constexpr bool vsync = !std::is_same<Exec, parallel>::value && hsync<tlist<>, Stages...>::value;
                using ij_caches = typename ij_entries<Caches>::type;
            };
            template <class Spec>
            struct msses_of {
                using type = tlist<Spec>;
            };
            template <class... Msses>
            struct msses_of<multi_pass_desc<Msses...>> {
                using type = tlist<Msses...>;
            };

            // ------------------------------------------------------------------------------------------ binding
            template <class A>
            struct kind_of;
            template <class T>
            struct kind_of<std::shared_ptr<storage::data_store<T>>> {
                using type = field_kind<T>;
            };
            template <class T, int From, int To>
            struct kind_of<sid::renamed<T, From, To>> {
                using type = field_kind<T>;
            };
            template <class T>
            struct kind_of<global_parameter_t<T>> {
                using type = global_kind<T>;
            };
            template <class D>
            struct kind_of<positional<D>> {
                using type = positional_kind<D>;
            };

            struct field_rt {
                void *origin;
                int64_t s[3];
                int64_t sx[3]; // extra data dimensions (0 where absent)
            };
            // the run's compute box and axis splitters, as the host side of the engine sees them
            struct run_geom {
                int i_start, i_size, j_start, j_size, k_size;
                ksplit ks;
                std::array<int, 3> origin() const { return {i_start, j_start, 0}; }
            };
            template <class T>
            field_view<T> view_of_rt(field_rt const &f) {
                return {static_cast<T *>(f.origin), f.s[0], f.s[1], f.s[2], {f.sx[0], f.sx[1], f.sx[2]}};
            }
            struct lds_rt {
                int off, w, imin, jmin; // first element (in elements of the value type), row length, tile origin
                int imax, jmax;         // the tile's far edges relative to the workgroup's block
            };
            struct fuse_rt {
                int ti = 64, tj = 8;
                int kc = 1; // levels per workgroup (fused_kc_kernel when > 1)
                size_t bytes = 0;
            };
            struct run_ctx {
                std::map<int, field_rt> fields; // user fields and HBM temporaries by placeholder key
                std::map<int, lds_rt> lds;      // ij-cached temporaries of fused multistages
                std::vector<fuse_rt> fuse;      // tile and LDS size per multistage
                int origin[3];
            };

            template <class P, class Caches, class ArgsTuple>
            view_for<P, Caches> make_view(run_ctx const &c, ArgsTuple const &args) {
                using K = typename P::kind;
                if constexpr (std::is_same<view_for<P, Caches>, lds_view<typename K::data_t>>::value) {
                    auto const &l = c.lds.at(P::key);
                    return {l.off, l.w, l.imin, l.jmin, 0};
                } else if constexpr (!std::is_same<view_for<P, Caches>, typename view_of<K>::type>::value) {
                    return {}; // a register window of the vertical kernel
                } else if constexpr (std::is_same<typename view_of<K>::type, field_view<typename K::data_t>>::value) {
                    auto const &f = c.fields.at(P::key);
                    return view_of_rt<typename K::data_t>(f);
                } else if constexpr (std::is_same<K, global_kind<typename K::data_t>>::value) {
                    return {std::get<P::key>(args).value};
                } else {
                    using D = typename K::dim_t;
                    constexpr int d = std::is_same<D, dim::i>::value ? 0 : std::is_same<D, dim::j>::value ? 1 : 2;
                    return {c.origin[d]};
                }
            }

            template <class Stage>
            struct stage_plhs;
            template <class F, class... Plhs>
            struct stage_plhs<stage_desc<F, Plhs...>> {
                template <class Caches = tlist<>, class ArgsTuple>
                static auto views(run_ctx const &c, ArgsTuple const &args) {
                    typename stage_views<stage_desc<F, Plhs...>, Caches>::type v{};
                    assign<Caches>(v, c, args, std::make_integer_sequence<int, sizeof...(Plhs)>());
                    return v;
                }
                template <class Caches, class V, class ArgsTuple, int... Is>
                static void assign(V &v, run_ctx const &c, ArgsTuple const &args, std::integer_sequence<int, Is...>) {
                    ((vget<Is>(v) = make_view<Plhs, Caches>(c, args)), ...);
                }
                static std::vector<access_rt> accesses() {
                    return stage_accesses<F, Plhs...>(typename F::param_list());
                }
                // allocate each temporary of this stage that has no memory yet
                static void alloc_tmps(run_ctx &c, std::map<int, ext_rt> const &box, run_geom const &g, void *stream,
                    std::vector<void *> &taken) {
                    (alloc_one<Plhs>(c, box, g, stream, taken), ...);
                }
                template <class P>
                static void alloc_one(run_ctx &c, std::map<int, ext_rt> const &box, run_geom const &g, void *stream,
                    std::vector<void *> &taken) {
                    if constexpr (P::is_tmp) {
                        if (c.fields.count(P::key) || c.lds.count(P::key))
                            return;
                        using T = typename P::kind::data_t;
                        ext_rt b = box.count(P::key) ? box.at(P::key) : ext_rt{};
                        const int hi = std::max(-b.imin, b.imax), hj = std::max(-b.jmin, b.jmax),
                                  hk = std::max(-b.kmin, b.kmax);
                        gt355_layout lay{};
                        const int64_t len[3] = {g.i_size + 2 * hi, g.j_size + 2 * hj, g.k_size + 2 * hk};
                        const int64_t hal[3] = {hi, hj, hk};
                        check(gt355_storage_layout(sizeof(T), len, hal, &lay), "stencil engine: temporary layout");
                        void *raw = nullptr;
                        check(gt355_tmp_acquire((size_t)lay.alloc_bytes, stream, &raw), "stencil engine: temporary");
                        taken.push_back(raw);
                        void *base = nullptr;
                        check(gt355_storage_base(&lay, raw, &base), "stencil engine: temporary layout");
                        T *origin = static_cast<T *>(base) + hi + hj * lay.stride[1] + hk * lay.stride[2];
                        c.fields[P::key] = {origin, {1, lay.stride[1], lay.stride[2]}, {0, 0, 0}};
                    }
                }
            };

            inline void check_launch(const char *what) {
                hipError_t e = hipGetLastError();
                if (e != hipSuccess)
                    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
            }

            inline dim3 grid_2d(int ni, int nj, int nz) {
                return dim3((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)nz);
            }

            // the kernel arguments of a vertical multistage: its views, register windows, stage domains and the box
            // of columns its stages cover
            template <class Mss, class Written>
            struct vprep {
                template <class... Stages>
                struct of {
                    using T = mss_traits<Mss>;
                    using Caches = auto_vcaches<typename T::caches, Stages...>;
                    using Ctx = vctx<Caches, Stages...>;
                    using views = vtuple<typename stage_views<Stages, Ctx>::type...>;
                    using rings = typename to_vtuple<typename ring_list<Written, Caches, tlist<Stages...>>::type>::type;
                    static constexpr int NS = sizeof...(Stages);
                    views vt{};
                    rings rs{};
                    domains<NS> dom{};
                    cbox box{};
                };
                template <class... Stages, class ArgsTuple>
                static of<Stages...> make(tlist<Stages...>, std::vector<stage_rt> const &st, int first,
                    run_ctx const &c, ArgsTuple const &args, run_geom const &g) {
                    using O = of<Stages...>;
                    constexpr int NS = O::NS;
                    const int ni = g.i_size, nj = g.j_size;
                    O p;
                    ext_rt u;
                    for (int n = 0; n < NS; ++n) {
                        ext_rt const &e = st[first + n].ext;
                        p.dom.d[n][0] = e.imin, p.dom.d[n][1] = ni + e.imax, p.dom.d[n][2] = e.jmin,
                        p.dom.d[n][3] = nj + e.jmax;
                        u.merge(e);
                    }
                    assign_all<typename O::Ctx>(p.vt, c, args, tlist<Stages...>(), std::make_integer_sequence<int, NS>());
                    init_rings(p.rs, c, st, first, NS, p.dom.d);
                    p.box = cbox{u.imin, ni + u.imax, u.jmin, nj + u.jmax};
                    return p;
                }
            };

            // whether two consecutive vertical multistages (stages [fa, fa + na) and [fb, fb + nb)) may share a
            // launch: every placeholder that one of them writes and the other accesses is accessed at the column
            // itself (no horizontal offset) by both
            inline bool vpair_fusable(std::vector<stage_rt> const &st, int fa, int na, int fb, int nb) {
                struct use {
                    bool wa = false, aa = false, wb = false, ab = false, off = false;
                };
                std::map<int, use> u;
                for (int s = fa; s < fb + nb; ++s) {
                    const bool first = s < fa + na;
                    for (access_rt const &a : st[s].acc) {
                        use &x = u[a.key];
                        (first ? x.aa : x.ab) = true;
                        if (a.write)
                            (first ? x.wa : x.wb) = true;
                        x.off = x.off || a.ext.imin || a.ext.imax || a.ext.jmin || a.ext.jmax;
                    }
                }
                for (auto const &e : u) {
                    use const &x = e.second;
                    if (((x.wa && x.ab) || (x.wb && x.aa)) && x.off)
                        return false;
                }
                return true;
            }
            // GT355_ENGINE_MSS_FUSION=0 launches every multistage on its own
            inline bool mss_fusion_on() {
                static const bool on = [] {
                    const char *e = std::getenv("GT355_ENGINE_MSS_FUSION");
                    return !(e && std::atoi(e) == 0);
                }();
                return on;
            }

            template <class MA, class MB, int NSPL, bool W, class Written, class ArgsTuple>
            void launch_vpair(std::vector<stage_rt> const &st, int fa, int fb, run_ctx const &c, ArgsTuple const &args,
                run_geom const &g, hipStream_t stream) {
                auto a = vprep<MA, Written>::make(typename mss_traits<MA>::stages(), st, fa, c, args, g);
                auto b = vprep<MB, Written>::make(typename mss_traits<MB>::stages(), st, fb, c, args, g);
                using A = decltype(a);
                using B = decltype(b);
                auto empty = [](cbox const &x) { return x.i1 <= x.i0 || x.j1 <= x.j0; };
                if (empty(a.box))
                    a.box = cbox{0, 0, 0, 0};
                if (empty(b.box))
                    b.box = cbox{0, 0, 0, 0};
                cbox un = empty(a.box) ? b.box : a.box;
                if (!empty(b.box))
                    un = cbox{std::min(un.i0, b.box.i0), std::max(un.i1, b.box.i1), std::min(un.j0, b.box.j0),
                        std::max(un.j1, b.box.j1)};
                const int w = un.i1 - un.i0, h = un.j1 - un.j0;
                if (w > 0 && h > 0 && g.k_size > 0)
                    vertical_pair_kernel<typename mss_traits<MA>::exec, typename mss_traits<MA>::stages,
                        typename A::views, typename A::rings, A::NS, typename mss_traits<MB>::exec,
                        typename mss_traits<MB>::stages, typename B::views, typename B::rings, B::NS, NSPL, W>
                        <<<grid_2d(w, h, 1), 256, 0, stream>>>(a.vt, a.rs, a.dom, a.box, b.vt, b.rs, b.dom, b.box,
                            un.i0, w, un.j0, h, g.ks);
                check_launch("stencil engine: fused vertical multistages launch");
            }

            template <class Mss, int NSPL, bool W, class Written, class... Stages, class ArgsTuple>
            void launch_mss(tlist<Stages...>, std::vector<stage_rt> const &st, int first, int m, run_ctx const &c,
                ArgsTuple const &args, run_geom const &g, hipStream_t stream) {
                using T = mss_traits<Mss>;
                using Exec = typename T::exec;
                constexpr int NS = sizeof...(Stages);
                const int ni = g.i_size, nj = g.j_size, nk = g.k_size;
                if constexpr (T::fused || T::vsync) {
                    using Ctx = fctx<typename T::ij_caches, Stages...>;
                    using ViewsTuple = vtuple<typename stage_views<Stages, Ctx>::type...>;
                    using Staged = staged_inputs<Stages...>;
                    using StageIns = typename stage_in_tuple<Staged>::type;
                    StageIns si{};
                    init_stage_ins(si, Staged(), c);
                    ViewsTuple vt{};
                    domains<NS> ext{};
                    for (int n = 0; n < NS; ++n) {
                        ext_rt const &e = st[first + n].ext;
                        ext.d[n][0] = e.imin, ext.d[n][1] = e.imax, ext.d[n][2] = e.jmin, ext.d[n][3] = e.jmax;
                    }
                    assign_all<Ctx>(vt, c, args, tlist<Stages...>(), std::make_integer_sequence<int, NS>());
                    fuse_rt const &f = c.fuse[m];
                    const int tiles_i = (ni + f.ti - 1) / f.ti, tiles_j = (nj + f.tj - 1) / f.tj;
                    if constexpr (T::fused) {
                        if (f.kc > 1) {
                            using StageBufs = typename stage_buf_tuple<Staged>::type;
                            fused_kc_kernel<NSPL, W, tlist<Stages...>, ViewsTuple, StageIns, StageBufs, NS>
                                <<<dim3((unsigned)(tiles_i * tiles_j), 1, (unsigned)((nk + f.kc - 1) / f.kc)), 256,
                                    f.bytes, stream>>>(vt, si, ext, ni, nj, g.ks, tiles_i, f.ti, f.tj, f.kc, nk);
                        } else {
                            fused_kernel<NSPL, W, tlist<Stages...>, ViewsTuple, StageIns, NS>
                                <<<dim3((unsigned)(tiles_i * tiles_j), 1, (unsigned)nk), 256, f.bytes, stream>>>(
                                    vt, si, ext, ni, nj, g.ks, tiles_i, f.ti, f.tj);
                        }
                        check_launch("stencil engine: fused multistage launch");
                    } else {
                        vfused_kernel<Exec, NSPL, W, tlist<Stages...>, ViewsTuple, StageIns, NS>
                            <<<dim3((unsigned)(tiles_i * tiles_j), 1, 1), 256, f.bytes, stream>>>(
                                vt, si, ext, ni, nj, g.ks, tiles_i, f.ti, f.tj);
                        check_launch("stencil engine: vertical multistage launch (tiles)");
                    }
                } else if constexpr (std::is_same<Exec, parallel>::value) {
                    int s = first;
                    auto one = [&](auto stage) {
                        using S = decltype(stage);
                        ext_rt const &e = st[s++].ext;
                        auto v = stage_plhs<S>::views(c, args);
                        const int w = ni - e.imin + e.imax, h = nj - e.jmin + e.jmax;
                        if (w > 0 && h > 0)
                            parallel_stage_kernel<typename S::functor, NSPL, W, decltype(v)>
                                <<<grid_2d(w, h, nk), 256, 0, stream>>>(v, e.imin, w, e.jmin, h, g.ks);
                        check_launch("stencil engine: parallel stage launch");
                    };
                    (one(Stages()), ...);
                } else {
                    auto p = vprep<Mss, Written>::make(tlist<Stages...>(), st, first, c, args, g);
                    using P = decltype(p);
                    const int w = p.box.i1 - p.box.i0, h = p.box.j1 - p.box.j0;
                    if (w > 0 && h > 0 && nk > 0)
                        vertical_kernel<Exec, NSPL, W, tlist<Stages...>, typename P::views, typename P::rings, NS>
                            <<<grid_2d(w, h, 1), 256, 0, stream>>>(p.vt, p.rs, p.dom, p.box.i0, w, p.box.j0, h, g.ks);
                    check_launch("stencil engine: vertical multistage launch");
                }
            }


            // a register window's memory and column boxes: it fills where its multistage reads the placeholder and
            // flushes where it writes it
            template <class Ring>
            void init_ring(Ring &r, run_ctx const &c, std::vector<stage_rt> const &st, int first, int ns,
                const int (*dom)[4]) {
                using T = std::decay_t<decltype(r.r[0][0])>;
                bool any_r = false, any_w = false;
                for (int n = 0; n < ns; ++n)
                    for (auto const &a : st[first + n].acc) {
                        if (a.key != Ring::key)
                            continue;
                        if (Ring::ncol == 1 && (a.ext.imin || a.ext.imax || a.ext.jmin || a.ext.jmax))
                            throw std::runtime_error("run: a k_cached placeholder is accessed at a horizontal offset");
                        auto grow = [&](int *b, bool &any) {
                            if (!any) {
                                for (int q = 0; q < 4; ++q)
                                    b[q] = dom[n][q];
                                any = true;
                            } else {
                                b[0] = std::min(b[0], dom[n][0]), b[1] = std::max(b[1], dom[n][1]);
                                b[2] = std::min(b[2], dom[n][2]), b[3] = std::max(b[3], dom[n][3]);
                            }
                        };
                        grow(r.fb, any_r);
                        if (a.write)
                            grow(r.wb, any_w);
                    }
                if (!any_w)
                    r.wb[0] = r.wb[1] = r.wb[2] = r.wb[3] = 0;
                auto it = c.fields.find(Ring::key);
                if (it != c.fields.end())
                    r.mem = view_of_rt<T>(it->second);
                else
                    r.fb[0] = r.fb[1] = r.wb[0] = r.wb[1] = 0; // no memory: a purely register-held temporary
            }
            template <int... Is, class... Rs>
            void init_rings(vtuple_impl<std::integer_sequence<int, Is...>, Rs...> &rs, run_ctx const &c,
                std::vector<stage_rt> const &st, int first, int ns, const int (*dom)[4]) {
                (init_ring(vget<Is>(rs), c, st, first, ns, dom), ...);
            }
            template <class Caches, class VT, class ArgsTuple, class... Stages, int... Is>
            void assign_all(VT &vt, run_ctx const &c, ArgsTuple const &args, tlist<Stages...>,
                std::integer_sequence<int, Is...>) {
                ((vget<Is>(vt) = stage_plhs<Stages>::template views<Caches>(c, args)), ...);
            }

            template <class Mss>
            struct staged_list;
            template <class Exec, class Caches, class... Stages>
            struct staged_list<mss_desc<Exec, Caches, Stages...>> {
                using type = staged_inputs<Stages...>;
            };
            template <class P, class T>
            void init_stage_in(stage_in<T> &s, run_ctx const &c) {
                auto const &f = c.fields.at(P::key);
                auto const &l = c.lds.at(P::key);
                s.src = view_of_rt<T>(f);
                s.off = l.off, s.w = l.w, s.imin = l.imin, s.jmin = l.jmin;
                s.imax = l.imax, s.jmax = l.jmax;
            }
            template <int... Is, class... Ss, class... Ps>
            void init_stage_ins(vtuple_impl<std::integer_sequence<int, Is...>, Ss...> &si, tlist<Ps...>, run_ctx const &c) {
                (init_stage_in<Ps>(vget<Is>(si), c), ...);
            }

            // The tiled vertical form (vfused_kernel) is exact when no read can see another workgroup's write in an
            // order that the sweep does not fix.  Workgroups share the points a stage on an extended box computes
            // again for the neighbouring tiles, and they run at different levels.  For a placeholder in memory that
            // the multistage writes:
            //  * if several stages write it, each writes only its own columns (extent zero);
            //  * a read at a level where a stage at or after the reader writes it (a later stage at the same level,
            //    or any writer at a level ahead of the sweep) stays on the tile's own columns, and those writers
            //    write only their own columns;
            //  * a read of a level already written (an earlier stage at the same level, or a level behind the sweep)
            //    is covered by the box on which the last such writer runs, so the workgroup wrote every value it
            //    reads itself.
            // Placeholders in LDS (ij_cached) are private to the workgroup.  An inout accessor writes at its own
            // point and reads at the other levels of its k extent; a read of the point it writes (eval(x) += ...)
            // counts as part of the write, so a stage that does so on an extended box is not caught here.
            inline void check_vsync(int m, int dir, std::vector<stage_rt> const &st, std::set<int> const &lds_keys) {
                auto zero_h = [](ext_rt const &e) {
                    return !e.set || (e.imin == 0 && e.imax == 0 && e.jmin == 0 && e.jmax == 0);
                };
                auto covers = [](ext_rt const &o, ext_rt const &r) {
                    return o.imin <= r.imin && o.imax >= r.imax && o.jmin <= r.jmin && o.jmax >= r.jmax;
                };
                std::vector<int> ids;
                for (int q = 0; q < (int)st.size(); ++q)
                    if (st[q].mss == m)
                        ids.push_back(q);
                std::map<int, std::vector<int>> writers;
                for (int q : ids)
                    for (auto const &a : st[q].acc)
                        if (a.write && !lds_keys.count(a.key) &&
                            (writers[a.key].empty() || writers[a.key].back() != q))
                            writers[a.key].push_back(q);
                for (auto const &w : writers)
                    if (w.second.size() > 1)
                        for (int q : w.second)
                            if (!zero_h(st[q].ext))
                                throw std::runtime_error("run: a vertical multistage with horizontal dependencies "
                                                         "writes a field from several stages, one of them on an "
                                                         "extended box");
                for (int s : ids)
                    for (auto const &a : st[s].acc) {
                        if ((a.write && a.ext.kmin == 0 && a.ext.kmax == 0) || !writers.count(a.key))
                            continue;
                        auto const &ws = writers.at(a.key);
                        ext_rt r{st[s].ext.imin + a.ext.imin, st[s].ext.imax + a.ext.imax, st[s].ext.jmin + a.ext.jmin,
                            st[s].ext.jmax + a.ext.jmax, 0, 0, true};
                        const bool same = !a.write && a.ext.kmin <= 0 && a.ext.kmax >= 0;
                        const bool ahead = dir > 0 ? a.ext.kmax > 0 : a.ext.kmin < 0;
                        const bool behind = dir > 0 ? a.ext.kmin < 0 : a.ext.kmax > 0;
                        int last_before = -1, last_all = ws.back();
                        for (int t : ws) {
                            const bool after = ahead || (same && t >= s);
                            if (after && (!zero_h(st[t].ext) || !zero_h(r)))
                                throw std::runtime_error("run: a vertical multistage with horizontal dependencies "
                                                         "reads a field that the same or a later stage writes, "
                                                         "outside the workgroup's own columns");
                            if (same && t < s)
                                last_before = t;
                        }
                        if ((same && last_before >= 0 && !covers(st[last_before].ext, r)) ||
                            (behind && !covers(st[last_all].ext, r)))
                            throw std::runtime_error("run: a vertical multistage with horizontal dependencies reads, "
                                                     "at columns its writer does not compute there, a field written "
                                                     "at another level");
                    }
            }

            // LDS plan of a fused multistage: each ij-cached temporary gets a tile of the workgroup's TI x TJ block
            // extended by the temporary's extent; TJ shrinks until the tiles fit 64 KB of LDS
            template <class... Cs, class... Ins>
            void plan_lds(tlist<Cs...>, tlist<Ins...>, int m, int dir, std::vector<stage_rt> const &st,
                std::map<int, ext_rt> const &box, run_ctx &c) {
                const int keys[] = {0, Cs::key...};
                const int elems[] = {0, (int)sizeof(typename Cs::kind::data_t)...};
                const int in_keys[] = {0, Ins::key...};
                const int in_elems[] = {0, (int)sizeof(typename Ins::kind::data_t)...};
                constexpr bool all_tmp = (true && ... && Cs::is_tmp);
                static_assert(all_tmp, "ij_cached takes temporaries (GT_DECLARE_TMP)");
                for (int q = 1; q <= (int)sizeof...(Cs); ++q)
                    for (auto const &s : st)
                        for (auto const &a : s.acc)
                            if (a.key == keys[q] && (s.mss != m || a.ext.kmin || a.ext.kmax))
                                throw std::runtime_error("run: an ij_cached temporary must be used only inside its "
                                                         "multistage and without k offsets");
                // workgroups of one fused launch run concurrently: a field that lives in memory (not in LDS) and is
                // read at a horizontal offset must not be written by that stage or a later one of the multistage
                for (auto const &r : st) {
                    if (r.mss != m)
                        continue;
                    for (auto const &a : r.acc) {
                        bool cached = false;
                        for (int q = 1; q <= (int)sizeof...(Cs); ++q)
                            cached = cached || keys[q] == a.key;
                        if (cached || !(a.ext.imin || a.ext.imax || a.ext.jmin || a.ext.jmax))
                            continue;
                        for (auto const &w : st)
                            if (w.mss == m && &w >= &r)
                                for (auto const &b : w.acc)
                                    if (b.write && b.key == a.key)
                                        throw std::runtime_error("run: a stage reads a field at a horizontal offset "
                                                                 "that it or a later stage of the multistage writes");
                    }
                }
                if (dir != 0) {
                    std::set<int> lds_keys;
                    for (int q = 1; q <= (int)sizeof...(Cs); ++q)
                        lds_keys.insert(keys[q]);
                    check_vsync(m, dir, st, lds_keys);
                }
                fuse_rt f;
                // tile columns: 64 less the widest rim of the multistage's LDS tiles and stage extents, so that every row
                // a stage or a staging load walks fits one pass of the 64 lanes (hdiff: 60); at least 32
                int rim = 0;
                for (auto const &s : st)
                    if (s.mss == m)
                        rim = std::max(rim, s.ext.imax - s.ext.imin);
                auto widen = [&](int key) {
                    if (box.count(key))
                        rim = std::max(rim, box.at(key).imax - box.at(key).imin);
                };
                for (int q = 1; q <= (int)sizeof...(Cs); ++q)
                    widen(keys[q]);
                for (int q = 1; q <= (int)sizeof...(Ins); ++q)
                    widen(in_keys[q]);
                f.ti = std::max(32, 64 - rim);
                // tile rows per workgroup: GT355_ENGINE_TJ (a power of two up to 32) or 8, halved until the tiles fit
                const char *env = std::getenv("GT355_ENGINE_TJ");
                const int tj0 = env ? std::max(1, std::min(32, std::atoi(env))) : 8;
                for (f.tj = tj0;; f.tj /= 2) {
                    size_t bytes = 0;
                    std::map<int, lds_rt> lds;
                    auto place = [&](int key, int elem) {
                        if (!box.count(key))
                            return;
                        ext_rt const &b = box.at(key);
                        const int w = f.ti + b.imax - b.imin, h = f.tj + b.jmax - b.jmin;
                        bytes = (bytes + 15) / 16 * 16;
                        lds[key] = {(int)(bytes / elem), w, b.imin, b.jmin, b.imax, b.jmax};
                        bytes += (size_t)w * h * elem;
                    };
                    for (int q = 1; q <= (int)sizeof...(Cs); ++q)
                        place(keys[q], elems[q]);
                    for (int q = 1; q <= (int)sizeof...(Ins); ++q)
                        place(in_keys[q], in_elems[q]);
                    if (bytes <= 65536 || f.tj == 1) {
                        if (bytes > 65536)
                            throw std::runtime_error("run: the LDS tiles of a multistage exceed 64 KB");
                        f.bytes = bytes;
                        for (auto &l : lds)
                            c.lds[l.first] = l.second;
                        // a parallel multistage walks chunks of levels (GT355_ENGINE_KC, default 2) when its staged
                        // inputs fit the registers of one level: one pass of the 64 lanes, 4 * SROWS rows.  Measured
                        // on the hdiff spec at 256x256x80 FP64: 54.3 us one level per workgroup, 47.5 at 2, 48.5 at
                        // 4, 53.9 at 8
                        bool fits = true;
                        for (int q = 1; q <= (int)sizeof...(Ins); ++q)
                            if (box.count(in_keys[q])) {
                                ext_rt const &b = box.at(in_keys[q]);
                                fits = fits && f.ti + b.imax - b.imin <= 64 && f.tj + b.jmax - b.jmin <= 4 * SROWS;
                            }
                        const char *kenv = std::getenv("GT355_ENGINE_KC");
                        const int kc = kenv ? std::max(1, std::atoi(kenv)) : 2;
                        f.kc = dir == 0 && fits ? kc : 1;
                        break;
                    }
                }
                c.fuse[m] = f;
            }
            template <class... Msses>
            void plan_all(tlist<Msses...>, std::vector<stage_rt> const &st, std::map<int, ext_rt> const &box, run_ctx &c) {
                c.fuse.assign(sizeof...(Msses), fuse_rt{});
                int m = 0;
                auto one = [&](auto mss) {
                    using T = mss_traits<decltype(mss)>;
                    if constexpr (T::fused || T::vsync)
                        plan_lds(typename T::ij_caches(), typename staged_list<decltype(mss)>::type(), m,
                            T::vsync ? (std::is_same<typename T::exec, backward>::value ? -1 : 1) : 0, st, box, c);
                    ++m;
                };
                (one(Msses()), ...);
            }

            template <class Spec, class PList, class = void>
            struct spec_result {
                using type = void;
            };
            template <class Spec, class... Ps>
            struct spec_result<Spec, tlist<Ps...>, std::void_t<decltype(std::declval<Spec>()(Ps()...))>> {
                using type = decltype(std::declval<Spec>()(Ps()...));
            };
            template <class T>
            struct is_spec : std::false_type {};
            template <class E, class C, class... S>
            struct is_spec<mss_desc<E, C, S...>> : std::true_type {};
            template <class... M>
            struct is_spec<multi_pass_desc<M...>> : std::true_type {};

            template <class Spec, class Seq, class... Args>
            struct plhs_for;
            template <class Spec, size_t... Is, class... Args>
            struct plhs_for<Spec, std::index_sequence<Is...>, Args...> {
                using type = tlist<arg_plh<(int)Is, typename kind_of<std::decay_t<Args>>::type>...>;
            };

            template <class Spec, class... Args>
            using spec_of = typename spec_result<Spec,
                typename plhs_for<Spec, std::index_sequence_for<Args...>, Args...>::type>::type;

            // user-field bookkeeping: bounds check and host -> device synchronisation of every field argument
            template <int N, class A>
            void bind_arg(run_ctx &c, A const &a, std::vector<stage_rt> const &st, run_geom const &g) {
                if constexpr (std::is_same<typename kind_of<A>::type, field_kind<typename kind_of<A>::type::data_t>>::value) {
                    bool written = false;
                    ext_rt need;
                    for (auto const &s : st)
                        for (auto const &ac : s.acc)
                            if (ac.key == N) {
                                written = written || ac.write;
                                ext_rt r{s.ext.imin + ac.ext.imin, s.ext.imax + ac.ext.imax, s.ext.jmin + ac.ext.jmin,
                                    s.ext.jmax + ac.ext.jmax, ac.ext.kmin, ac.ext.kmax, true};
                                need.merge(r);
                            }
                    if (!need.set)
                        return;
                    auto const &store = sid::store_of(a);
                    gt355_field f = store->sid(g.origin(), written);
                    int64_t sx[3] = {0, 0, 0};
                    sid::rename(a, f, sx);
                    auto bad = [&](int d, int lo, int hi) {
                        return f.stride[d] != 0 && (lo < f.lower[d] || hi > f.upper[d]);
                    };
                    if (bad(0, need.imin, g.i_size + need.imax) || bad(1, need.jmin, g.j_size + need.jmax))
                        throw std::runtime_error("run: field argument " + std::to_string(N) + " (" + store->name() +
                                                 ") is too small for the extents the stencil reads or writes");
                    if (written && (f.stride[0] == 0 || f.stride[1] == 0 || f.stride[2] == 0))
                        throw std::runtime_error("run: a written field cannot have masked dimensions");
                    c.fields[N] = {f.origin, {f.stride[0], f.stride[1], f.stride[2]}, {sx[0], sx[1], sx[2]}};
                }
            }
            template <class ArgsTuple, size_t... Is>
            void bind_args(run_ctx &c, ArgsTuple const &args, std::vector<stage_rt> const &st, run_geom const &g,
                std::index_sequence<Is...>) {
                (bind_arg<(int)Is>(c, std::get<Is>(args), st, g), ...);
            }

            template <class... Msses>
            void collect(tlist<Msses...>, std::vector<stage_rt> &st) {
                int m = 0;
                auto one_mss = [&](auto mss) {
                    using T = mss_traits<decltype(mss)>;
                    const bool vertical = !std::is_same<typename T::exec, parallel>::value;
                    collect_stages(typename T::stages(), m, vertical, st);
                    ++m;
                };
                (one_mss(Msses()), ...);
            }
            template <class... Stages>
            void collect_stages(tlist<Stages...>, int m, bool vertical, std::vector<stage_rt> &st) {
                ((st.push_back(stage_rt{m, vertical, stage_plhs<Stages>::accesses(), {}})), ...);
            }

            // extent analysis: walks the stages backwards; a stage's extent is the union of the extents at which
            // later stages read its outputs; returns the box every placeholder is touched over (for temporaries)
            inline std::map<int, ext_rt> analyse(std::vector<stage_rt> &st) {
                std::map<int, ext_rt> req, box;
                for (int s = (int)st.size() - 1; s >= 0; --s) {
                    ext_rt e;
                    e.set = true;
                    for (auto const &a : st[s].acc)
                        if (a.write && req.count(a.key)) {
                            ext_rt r = req[a.key];
                            r.kmin = r.kmax = 0;
                            e.merge(r);
                        }
                    st[s].ext = e;
                    // every access reads (an inout accessor sees the values earlier stages left)
                    for (auto const &a : st[s].acc) {
                        ext_rt r{e.imin + a.ext.imin, e.imax + a.ext.imax, e.jmin + a.ext.jmin, e.jmax + a.ext.jmax,
                            a.ext.kmin, a.ext.kmax, true};
                        req[a.key].merge(r);
                        box[a.key].merge(r);
                    }
                }
                // vertical multistages: a stage may read, at a horizontal offset, a field of the multistage only
                // after an earlier stage wrote it (the multistage then runs tile by tile with barriers)
                for (auto const &s : st) {
                    if (!s.vertical)
                        continue;
                    for (auto const &a : s.acc) {
                        if (!(a.ext.imin || a.ext.imax || a.ext.jmin || a.ext.jmax))
                            continue;
                        for (auto const &t : st)
                            if (t.mss == s.mss && &t >= &s)
                                for (auto const &b : t.acc)
                                    if (b.write && b.key == a.key)
                                        throw std::runtime_error(
                                            "run: a vertical multistage reads, at a horizontal offset, a field that "
                                            "the same or a later stage writes");
                    }
                }
                return box;
            }

            template <int I, class... Ts>
            struct type_at;
            template <class T0, class... Ts>
            struct type_at<0, T0, Ts...> {
                using type = T0;
            };
            template <int I, class T0, class... Ts>
            struct type_at<I, T0, Ts...> : type_at<I - 1, Ts...> {};
            // a multistage that launch_mss runs as vertical_kernel (a sweep per column, no tiles)
            template <class Mss>
            constexpr bool plain_vertical() {
                using T = mss_traits<Mss>;
                return !T::fused && !T::vsync && !std::is_same<typename T::exec, parallel>::value;
            }
            // multistages I.. of the spec (first[m]: index of multistage m's first stage): two consecutive plain
            // vertical multistages that pass data only at the column itself share one launch (vertical_pair_kernel);
            // every other multistage launches on its own
            template <int NSPL, bool W, class Written, int I, class... Msses, class ArgsTuple>
            void launch_from(tlist<Msses...> l, const int *first, std::vector<stage_rt> const &st, run_ctx const &c,
                ArgsTuple const &args, run_geom const &g, hipStream_t stream) {
                constexpr int N = sizeof...(Msses);
                if constexpr (I < N) {
                    using MA = typename type_at<I, Msses...>::type;
                    if constexpr (I + 1 < N) {
                        using MB = typename type_at<I + 1, Msses...>::type;
                        if constexpr (plain_vertical<MA>() && plain_vertical<MB>()) {
                            if (mss_fusion_on() &&
                                vpair_fusable(st, first[I], mss_traits<MA>::n, first[I + 1], mss_traits<MB>::n)) {
                                launch_vpair<MA, MB, NSPL, W, Written>(st, first[I], first[I + 1], c, args, g, stream);
                                launch_from<NSPL, W, Written, I + 2>(l, first, st, c, args, g, stream);
                                return;
                            }
                        }
                    }
                    launch_mss<MA, NSPL, W, Written>(
                        typename mss_traits<MA>::stages(), st, first[I], I, c, args, g, stream);
                    launch_from<NSPL, W, Written, I + 1>(l, first, st, c, args, g, stream);
                }
            }
            template <int NSPL, bool W, class... Msses, class ArgsTuple>
            void launch_all(tlist<Msses...> l, std::vector<stage_rt> const &st, run_ctx const &c, ArgsTuple const &args,
                run_geom const &g, hipStream_t stream) {
                using Written = typename spec_written<tlist<>, tlist<Msses...>>::type;
                int first[sizeof...(Msses) + 1];
                int n = 0, q = 0;
                ((first[q++] = n, n += mss_traits<Msses>::n), ...);
                launch_from<NSPL, W, Written, 0>(l, first, st, c, args, g, stream);
            }
            // temporaries come from the library's pool (gt355_tmp_acquire: the cached allocator of
            // sid/allocator.hpp:44-57, made safe across streams) and go back to it after the run's launches are
            // queued on `stream`
            template <class... Msses>
            void alloc_all(tlist<Msses...>, run_ctx &c, std::map<int, ext_rt> const &box, run_geom const &g, void *stream,
                std::vector<void *> &taken) {
                auto one = [&](auto mss) {
                    alloc_stages(typename mss_traits<decltype(mss)>::stages(), c, box, g, stream, taken);
                };
                (one(Msses()), ...);
            }
            template <class... Stages>
            void alloc_stages(tlist<Stages...>, run_ctx &c, std::map<int, ext_rt> const &box, run_geom const &g,
                void *stream, std::vector<void *> &taken) {
                (stage_plhs<Stages>::alloc_tmps(c, box, g, stream, taken), ...);
            }
        } // namespace engine_

        // ------------------------------------------------------------------------------------------ stencil functions
        // call<F, Region, ReturnType>::[at<i, j, k>::]with(eval, args...) and call_proc<F, Region>::with(eval, args...)
        // (stencil/frontend/cartesian/stencil_functions.hpp:139-247): run functor F at the evaluation point (shifted by
        // at<>) with its accessors bound to the caller's arguments -- an accessor argument is read by the caller's
        // evaluator at the sum of both offsets, a value argument is read (or written) directly.  call returns what F
        // writes to its single inout accessor; call_proc passes every argument, outputs included.
        namespace engine_ {
            template <class T>
            struct is_accessor : std::false_type {};
            template <int I, class E, int N>
            struct is_accessor<cartesian::in_accessor<I, E, N>> : std::true_type {};
            template <int I, class E, int N>
            struct is_accessor<cartesian::inout_accessor<I, E, N>> : std::true_type {};

            template <class AL>
            struct out_index;
            template <class... As>
            struct out_index<cartesian::param_list_t<As...>> {
                static constexpr int count = ((!As::is_const ? 1 : 0) + ... + 0);
                static constexpr int value = ((!As::is_const ? As::index : 0) + ... + 0);
            };

            // Args: a vtuple of references to the arguments in the callee's parameter order (OUT: the position of the
            // result, which is not among Args; -1 for call_proc)
            template <class Eval, class Args, class Res, int OUT, int OI, int OJ, int OK>
            struct call_eval {
                Eval &e;
                Args const &args;
                Res *res;
                template <int I>
                __host__ __device__ __forceinline__ decltype(auto) param() const {
                    if constexpr (I == OUT)
                        return (*res);
                    else
                        return vget<(OUT >= 0 && I > OUT ? I - 1 : I)>(args);
                }
                template <class P, class Acc>
                __host__ __device__ __forceinline__ decltype(auto) bind(P &&p, Acc const &a) const {
                    using A = std::decay_t<P>;
                    if constexpr (is_accessor<A>::value)
                        return e(A(p.di + a.di + OI, p.dj + a.dj + OJ, p.dk + a.dk + OK, p.x[0] + a.x[0],
                            p.x[1] + a.x[1], p.x[2] + a.x[2]));
                    else
                        return std::forward<P>(p);
                }
                template <int I, class E, int N>
                __host__ __device__ __forceinline__ decltype(auto) operator()(cartesian::in_accessor<I, E, N> const &a) const {
                    return bind(param<I>(), a);
                }
                template <int I, class E, int N>
                __host__ __device__ __forceinline__ decltype(auto) operator()(cartesian::inout_accessor<I, E, N> const &a) const {
                    return bind(param<I>(), a);
                }
                template <class X>
                __host__ __device__ auto operand(X const &x) const {
                    if constexpr (std::is_arithmetic<X>::value)
                        return x;
                    else
                        return std::decay_t<decltype((*this)(x))>((*this)(x));
                }
                template <class Op, class L, class R>
                __host__ __device__ auto operator()(cartesian::expr2<Op, L, R> const &x) const {
                    return Op()(operand(x.l), operand(x.r));
                }
                template <class Op, class A>
                __host__ __device__ auto operator()(cartesian::expr1<Op, A> const &x) const {
                    return Op()(operand(x.a));
                }
            };

            template <class F, class Region, class E>
            __host__ __device__ __forceinline__ void call_functor(E &ce) {
                if constexpr (std::is_void<Region>::value)
                    F::apply(ce);
                else
                    F::apply(ce, Region());
            }
        } // namespace engine_

        namespace cartesian {
            template <class F, class Region = void, class ReturnType = void, int OI = 0, int OJ = 0, int OK = 0>
            struct call {
                template <int I, int J, int K>
                using at = call<F, Region, ReturnType, I, J, K>;
                template <class R>
                using return_type = call<F, Region, R, OI, OJ, OK>;

                template <class Eval, class A0, class... Args>
                __host__ __device__ __forceinline__ static auto with(Eval &eval, A0 &&a0, Args &&... args) {
                    using out_t = engine_::out_index<typename F::param_list>;
                    static_assert(out_t::count == 1, "call<>: the stencil function must have exactly one inout accessor");
                    using deduced = std::conditional_t<engine_::is_accessor<std::decay_t<A0>>::value,
                        std::decay_t<decltype(eval(a0))>, std::decay_t<A0>>;
                    using Res = std::conditional_t<std::is_void<ReturnType>::value, deduced, ReturnType>;
                    Res res{};
                    using Tup = engine_::vtuple<A0 &&, Args &&...>;
                    Tup t{{std::forward<A0>(a0)}, {std::forward<Args>(args)}...};
                    engine_::call_eval<Eval, Tup, Res, out_t::value, OI, OJ, OK> ce{eval, t, &res};
                    engine_::call_functor<F, Region>(ce);
                    return res;
                }
            };

            template <class F, class Region = void, int OI = 0, int OJ = 0, int OK = 0>
            struct call_proc {
                template <int I, int J, int K>
                using at = call_proc<F, Region, I, J, K>;

                template <class Eval, class... Args>
                __host__ __device__ __forceinline__ static void with(Eval &eval, Args &&... args) {
                    using Tup = engine_::vtuple<Args &&...>;
                    Tup t{{std::forward<Args>(args)}...};
                    engine_::call_eval<Eval, Tup, int, -1, OI, OJ, OK> ce{eval, t, nullptr};
                    engine_::call_functor<F, Region>(ce);
                }
            };
        } // namespace cartesian

        namespace engine_ {
            // The precision policy's temporaries: under GT355_WIDEN a temporary declared with the storage type float
            // is held in double (memory, LDS tile or register window), so k-caches, pivots and intermediate fields
            // keep the compute type and only the stores to user fields narrow.
            template <class P>
            struct promote {
                using type = P;
            };
            template <int Id>
            struct promote<tmp_plh<Id, float>> {
                using type = tmp_plh<Id, double>;
            };
            template <class P, bool Fi, bool Fl>
            struct promote<kc<P, Fi, Fl>> {
                using type = kc<typename promote<P>::type, Fi, Fl>;
            };
            template <class F, class... Ps>
            struct promote<stage_desc<F, Ps...>> {
                using type = stage_desc<F, typename promote<Ps>::type...>;
            };
            template <class... Ts>
            struct promote<tlist<Ts...>> {
                using type = tlist<typename promote<Ts>::type...>;
            };
            template <class Exec, class Caches, class... Stages>
            struct promote<mss_desc<Exec, Caches, Stages...>> {
                using type = mss_desc<Exec, typename promote<Caches>::type, typename promote<Stages>::type...>;
            };
            template <class... Msses>
            struct promote<multi_pass_desc<Msses...>> {
                using type = multi_pass_desc<typename promote<Msses>::type...>;
            };

            template <int N>
            run_geom geom_of(grid_t<N> const &g) {
                run_geom r{g.i_start, g.i_size, g.j_start, g.j_size, g.k_size, {}};
                for (int s = 0; s <= N; ++s)
                    r.ks.pos[s] = g.splitters[s];
                return r;
            }
        } // namespace engine_

        namespace engine_ {
            // ------------------------------------------------------------------------------------------ host execution
            // A spec run on host memory in the reference's semantics: a parallel multistage stage by stage over each
            // stage's extended box, a vertical multistage column by column in k order (k-caches are plain memory).
            // The engine uses it to check, once per spec, whether a spec computes exactly what one of the library's
            // hand-written kernels computes.
            constexpr int HOST_HALO = 3;

            struct host_store {
                std::vector<std::vector<unsigned char>> mem;
                template <class T>
                field_rt make(int ni, int nj, int nk) {
                    const int64_t s1 = ni + 2 * HOST_HALO, s2 = s1 * (nj + 2 * HOST_HALO);
                    mem.emplace_back((size_t)(s2 * (nk + 2 * HOST_HALO)) * sizeof(T), (unsigned char)0);
                    T *base = reinterpret_cast<T *>(mem.back().data());
                    return {base + HOST_HALO * (1 + s1 + s2), {1, s1, s2}, {0, 0, 0}};
                }
            };
            template <class T>
            T &host_at(field_rt const &f, int i, int j, int k) {
                return static_cast<T *>(f.origin)[i + j * f.s[1] + k * f.s[2]];
            }

            template <class... Ps>
            void host_tmps(tlist<Ps...>, run_ctx &c, host_store &hs, run_geom const &g) {
                auto one = [&](auto p) {
                    using P = decltype(p);
                    if constexpr (P::is_tmp)
                        if (!c.fields.count(P::key))
                            c.fields[P::key] = hs.template make<typename P::kind::data_t>(g.i_size, g.j_size, g.k_size);
                };
                (one(Ps()), ...);
            }
            template <class... Msses>
            void host_all_tmps(tlist<Msses...>, run_ctx &c, host_store &hs, run_geom const &g) {
                auto one = [&](auto mss) {
                    using St = typename mss_traits<decltype(mss)>::stages;
                    host_stage_tmps(St(), c, hs, g);
                };
                (one(Msses()), ...);
            }
            template <class... Stages>
            void host_stage_tmps(tlist<Stages...>, run_ctx &c, host_store &hs, run_geom const &g) {
                (host_tmps(typename plhs_of_stages<tlist<>, Stages>::type(), c, hs, g), ...);
            }

            template <int NSPL, bool W, class S, class ArgsTuple>
            void host_stage_box(run_ctx const &c, ArgsTuple const &args, ext_rt const &e, run_geom const &g) {
                auto v = stage_plhs<S>::views(c, args);
                for (int k = 0; k < g.k_size; ++k)
                    for (int j = e.jmin; j < g.j_size + e.jmax; ++j)
                        for (int i = e.imin; i < g.i_size + e.imax; ++i)
                            run_stage_at<typename S::functor, NSPL, W>(v, i, j, k, g.ks);
            }
            template <int NSPL, bool W, class Mss, class... Stages, class ArgsTuple, size_t... Ss>
            void host_mss(tlist<Stages...>, std::index_sequence<Ss...>, std::vector<stage_rt> const &st, int first,
                run_ctx const &c, ArgsTuple const &args, run_geom const &g) {
                using Exec = typename mss_traits<Mss>::exec;
                if constexpr (std::is_same<Exec, parallel>::value) {
                    (host_stage_box<NSPL, W, Stages>(c, args, st[first + (int)Ss].ext, g), ...);
                } else {
                    // level by level in the sweep's order, each stage over its whole box (the reference's naive
                    // backend): the semantics of both vertical forms
                    auto views = std::make_tuple(stage_plhs<Stages>::views(c, args)...);
                    constexpr bool up = !std::is_same<Exec, backward>::value;
                    for (int n = 0; n < g.k_size; ++n) {
                        const int k = up ? n : g.k_size - 1 - n;
                        auto one = [&](auto s) {
                            constexpr size_t S = decltype(s)::value;
                            ext_rt const &e = st[first + (int)S].ext;
                            for (int j = e.jmin; j < g.j_size + e.jmax; ++j)
                                for (int i = e.imin; i < g.i_size + e.imax; ++i)
                                    run_stage_at<typename std::tuple_element_t<S, std::tuple<Stages...>>::functor, NSPL,
                                        W>(std::get<S>(views), i, j, k, g.ks);
                        };
                        (one(std::integral_constant<size_t, Ss>()), ...);
                    }
                }
            }
            template <int NSPL, bool W, class... Msses, class ArgsTuple>
            void host_run(tlist<Msses...>, std::vector<stage_rt> const &st, run_ctx const &c, ArgsTuple const &args,
                run_geom const &g) {
                int first = 0;
                auto one = [&](auto mss) {
                    using T = mss_traits<decltype(mss)>;
                    host_mss<NSPL, W, decltype(mss)>(
                        typename T::stages(), std::make_index_sequence<T::n>(), st, first, c, args, g);
                    first += T::n;
                };
                (one(Msses()), ...);
            }

            // ------------------------------------------------------------------------------------------ fast path
            // A spec that computes, operation for operation, what one of the library's hand-written kernels computes
            // runs on that kernel (the job of the reference's backend entry point for a known spec,
            // stencil/gpu/entry_point.hpp:253-260).  The decision is made once per spec type: the roles of the
            // fields come from the access analysis (which field is written, at which extents each is read), then
            // the spec runs on the host on seeded random inputs and must reproduce the hand kernel's formulas
            // (gt355/functors.hpp) bit for bit at every point.  Anything else takes the generic engine.
            struct access_summary {
                bool present = false, written = false;
                ext_rt ext; // union of the accessors' extents
            };
            inline std::map<int, access_summary> summarize(std::vector<stage_rt> const &st) {
                std::map<int, access_summary> m;
                for (auto const &s : st)
                    for (auto const &a : s.acc) {
                        auto &x = m[a.key];
                        x.present = true;
                        x.written = x.written || a.write;
                        x.ext.merge(a.ext);
                    }
                return m;
            }
            inline bool ext_is(ext_rt const &e, int i0, int i1, int j0, int j1, int k0, int k1) {
                return e.set && e.imin == i0 && e.imax == i1 && e.jmin == j0 && e.jmax == j1 && e.kmin == k0 &&
                       e.kmax == k1;
            }
            // counter-based uniform values for the host check
            inline double host_uniform(uint64_t n, double a, double b) {
                uint64_t z = n * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
                z ^= z >> 31;
                return a + (b - a) * (double)(z >> 11) * 0x1.0p-53;
            }
            // quant = 0: continuous U(a, b); quant = q > 1: one of q equally spaced values a .. b, so that equal
            // neighbours, zero differences and zero products occur (a comparison that differs only at a tie, such as
            // > against >=, then changes the result)
            template <class T>
            void host_fill(
                field_rt const &f, run_geom const &g, int seed, double a, double b, int extra_k = 0, int quant = 0) {
                uint64_t n = (uint64_t)seed << 32;
                for (int k = -HOST_HALO; k < g.k_size + extra_k; ++k)
                    for (int j = -HOST_HALO; j < g.j_size + HOST_HALO; ++j)
                        for (int i = -HOST_HALO; i < g.i_size + HOST_HALO; ++i) {
                            double v = host_uniform(n++, a, b);
                            if (quant > 1)
                                v = a + (b - a) * std::floor((v - a) / (b - a) * quant) / (quant - 1);
                            host_at<T>(f, i, j, k) = T(std::min(v, b));
                        }
            }

            enum class fast_kind { none, hdiff, vadv };
            struct fast_plan {
                fast_kind kind = fast_kind::none;
                int role[6] = {-1, -1, -1, -1, -1, -1}; // hdiff: in, coeff, out; vadv: utens_stage, u_stage, wcon,
                                                        // u_pos, utens, dtr_stage
            };

            template <class T>
            constexpr bool fast_type = std::is_same<T, double>::value || std::is_same<T, float>::value;

            // the hand kernel's horizontal diffusion on host fields (gt355_hdiff's formulas)
            template <class T, class CT>
            bool hdiff_reproduced(field_rt const &in, field_rt const &co, field_rt const &out, run_geom const &g) {
                auto I = [&](int i, int j, int k) { return CT(host_at<T>(in, i, j, k)); };
                auto lap = [&](int i, int j, int k) {
                    return lap_f<CT>(I(i, j, k), I(i + 1, j, k), I(i, j + 1, k), I(i - 1, j, k), I(i, j - 1, k));
                };
                auto flx = [&](int i, int j, int k) {
                    return flux_f<CT>(lap(i + 1, j, k), lap(i, j, k), I(i + 1, j, k), I(i, j, k));
                };
                auto fly = [&](int i, int j, int k) {
                    return flux_f<CT>(lap(i, j + 1, k), lap(i, j, k), I(i, j + 1, k), I(i, j, k));
                };
                for (int k = 0; k < g.k_size; ++k)
                    for (int j = 0; j < g.j_size; ++j)
                        for (int i = 0; i < g.i_size; ++i) {
                            const T e = T(hdiff_out_f<CT>(I(i, j, k), CT(host_at<T>(co, i, j, k)), flx(i, j, k),
                                flx(i - 1, j, k), fly(i, j, k), fly(i, j - 1, k)));
                            const T got = host_at<T>(out, i, j, k);
                            if (std::memcmp(&e, &got, sizeof(T)) != 0)
                                return false;
                        }
                return true;
            }

            // the hand kernel's vertical advection on host fields (gt355_vadv's formulas); uts0 holds the input
            template <class T, class CT>
            bool vadv_reproduced(field_rt const &uts0, field_rt const &uts, field_rt const &us, field_rt const &wc,
                field_rt const &up, field_rt const &ut, T dtr_s, run_geom const &g) {
                const int nk = g.k_size, top = nk - 1;
                const CT dtr = CT(dtr_s);
                std::vector<CT> c(nk), d(nk);
                for (int j = 0; j < g.j_size; ++j)
                    for (int i = 0; i < g.i_size; ++i) {
                        auto V = [&](field_rt const &f, int ii, int k) { return CT(host_at<T>(f, ii, j, k)); };
                        for (int k = 0; k < nk; ++k) {
                            if (k == 0) {
                                auto r = forward_first<CT>(dtr, V(wc, i, 1), V(wc, i + 1, 1), V(us, i, 0), V(us, i, 1),
                                    V(up, i, 0), V(ut, i, 0), V(uts0, i, 0));
                                c[0] = r.c, d[0] = r.d;
                            } else if (k < top) {
                                auto r = forward_body<CT>(dtr, V(wc, i, k), V(wc, i + 1, k), V(wc, i, k + 1),
                                    V(wc, i + 1, k + 1), V(us, i, k - 1), V(us, i, k), V(us, i, k + 1), V(up, i, k),
                                    V(ut, i, k), V(uts0, i, k), c[k - 1], d[k - 1], false);
                                c[k] = r.c, d[k] = r.d;
                            } else {
                                d[k] = forward_last<CT>(dtr, V(wc, i, k), V(wc, i + 1, k), V(us, i, k - 1),
                                    V(us, i, k), V(up, i, k), V(ut, i, k), V(uts0, i, k), c[k - 1], d[k - 1]);
                            }
                        }
                        CT data = CT(0);
                        for (int k = top; k >= 0; --k) {
                            const CT nd = backward_data<CT>(c[k], d[k], data, k == top);
                            const T e = T(dtr * (nd - V(up, i, k)));
                            data = nd;
                            const T got = host_at<T>(uts, i, j, k);
                            if (std::memcmp(&e, &got, sizeof(T)) != 0)
                                return false;
                        }
                    }
                return true;
            }

            template <class A>
            struct arg_value {
                using type = void;
            };
            template <class T>
            struct arg_value<std::shared_ptr<storage::data_store<T>>> {
                using type = T;
            };
            template <class T>
            struct arg_value<global_parameter_t<T>> {
                using type = T;
            };
            template <class A>
            constexpr bool is_store = std::is_same<typename kind_of<A>::type,
                field_kind<typename arg_value<A>::type>>::value && !std::is_void<typename arg_value<A>::type>::value;

            template <bool W, class Msses, class... Args>
            fast_plan plan_fast(std::vector<stage_rt> const &st, Args const &... args) {
                fast_plan none;
                using T0 = typename arg_value<std::decay_t<std::tuple_element_t<0, std::tuple<Args...>>>>::type;
                constexpr bool same_t =
                    (std::is_same<typename arg_value<std::decay_t<Args>>::type, T0>::value && ...);
                if constexpr (!same_t || !fast_type<T0> || (W && !std::is_same<T0, float>::value)) {
                    return none;
                } else {
                    using CT = std::conditional_t<W, double, T0>;
                    constexpr int nargs = sizeof...(Args);
                    constexpr bool stores[] = {is_store<std::decay_t<Args>>...};
                    const auto acc = summarize(st);
                    auto args_t = std::forward_as_tuple(args...);
                    run_geom g{0, 9, 0, 7, 5, {}};
                    g.ks.pos[0] = 0, g.ks.pos[1] = g.k_size;
                    auto run_host = [&](run_ctx &c, host_store &hs) {
                        host_all_tmps(Msses(), c, hs, g);
                        host_run<2, W>(Msses(), st, c, args_t, g);
                    };
                    int written = -1, nwritten = 0;
                    for (int a = 0; a < nargs; ++a)
                        if (stores[a] && acc.count(a) && acc.at(a).written)
                            written = a, ++nwritten;
                    if (nwritten != 1)
                        return none;
                    if (nargs == 3 && stores[0] && stores[1] && stores[2]) {
                        // horizontal diffusion: out written at no offset; in read over <-2, 2, -2, 2> (through the
                        // stages' extents: the accessors read it at up to one point); coeff read at no offset
                        int in = -1, co = -1;
                        for (int a = 0; a < 3; ++a) {
                            if (a == written || !acc.count(a))
                                continue;
                            if (ext_is(acc.at(a).ext, 0, 0, 0, 0, 0, 0))
                                co = a;
                            else
                                in = a;
                        }
                        if (in < 0 || co < 0 || !ext_is(acc.at(written).ext, 0, 0, 0, 0, 0, 0))
                            return none;
                        // continuous inputs, then inputs on a few levels (ties in the limiter's comparison)
                        for (int quant : {0, 5}) {
                            run_ctx c;
                            host_store hs;
                            c.origin[0] = c.origin[1] = c.origin[2] = 0;
                            for (int a = 0; a < 3; ++a)
                                c.fields[a] = hs.template make<T0>(g.i_size, g.j_size, g.k_size);
                            host_fill<T0>(c.fields[in], g, 1, -1, 1, 0, quant);
                            host_fill<T0>(c.fields[co], g, 2, 0, 0.25, 0, quant);
                            run_host(c, hs);
                            if (!hdiff_reproduced<T0, CT>(c.fields[in], c.fields[co], c.fields[written], g))
                                return none;
                        }
                        fast_plan p;
                        p.kind = fast_kind::hdiff;
                        p.role[0] = in, p.role[1] = co, p.role[2] = written;
                        return p;
                    }
                    if (nargs == 6) {
                        // vertical advection: utens_stage read and written; u_stage read at k-1..k+1; wcon at i..i+1
                        // and k..k+1; u_pos and utens read at no offset (told apart by the check); dtr_stage a global
                        int us = -1, wc = -1, dtr = -1, zero[2] = {-1, -1}, nz = 0;
                        for (int a = 0; a < 6; ++a) {
                            if (!stores[a]) {
                                dtr = a;
                                continue;
                            }
                            if (a == written || !acc.count(a))
                                continue;
                            ext_rt const &e = acc.at(a).ext;
                            if (ext_is(e, 0, 0, 0, 0, -1, 1))
                                us = a;
                            else if (ext_is(e, 0, 1, 0, 0, 0, 1))
                                wc = a;
                            else if (ext_is(e, 0, 0, 0, 0, 0, 0) && nz < 2)
                                zero[nz++] = a;
                        }
                        if (us < 0 || wc < 0 || dtr < 0 || nz != 2 || stores[dtr] ||
                            !ext_is(acc.at(written).ext, 0, 0, 0, 0, 0, 0))
                            return none;
                        T0 dtr_value = 0;
                        {
                            int a = 0;
                            auto pick = [&](auto const &x) {
                                if constexpr (!is_store<std::decay_t<decltype(x)>>)
                                    if (a == dtr)
                                        dtr_value = x.value;
                                ++a;
                            };
                            (pick(args), ...);
                        }
                        // continuous inputs, then inputs on a few levels (zero differences and zero winds)
                        auto reproduced = [&](int up, int ut, int quant) {
                            run_ctx c;
                            host_store hs;
                            c.origin[0] = c.origin[1] = c.origin[2] = 0;
                            for (int a = 0; a < 6; ++a)
                                if (stores[a])
                                    c.fields[a] = hs.template make<T0>(g.i_size, g.j_size, g.k_size + 1);
                            host_fill<T0>(c.fields[written], g, 6, -1, 1, 0, quant);
                            host_fill<T0>(c.fields[us], g, 3, -1, 1, 0, quant);
                            host_fill<T0>(c.fields[wc], g, 7, -0.5, 0.5, 1, quant);
                            host_fill<T0>(c.fields[up], g, 4, -1, 1, 0, quant);
                            host_fill<T0>(c.fields[ut], g, 5, -1, 1, 0, quant);
                            field_rt uts0 = hs.template make<T0>(g.i_size, g.j_size, g.k_size + 1);
                            host_fill<T0>(uts0, g, 6, -1, 1, 0, quant);
                            run_host(c, hs);
                            return vadv_reproduced<T0, CT>(uts0, c.fields[written], c.fields[us], c.fields[wc],
                                c.fields[up], c.fields[ut], dtr_value, g);
                        };
                        for (int perm = 0; perm < 2; ++perm) {
                            const int up = zero[perm], ut = zero[1 - perm];
                            if (reproduced(up, ut, 0) && reproduced(up, ut, 5)) {
                                fast_plan p;
                                p.kind = fast_kind::vadv;
                                const int r[6] = {written, us, wc, up, ut, dtr};
                                for (int q = 0; q < 6; ++q)
                                    p.role[q] = r[q];
                                return p;
                            }
                        }
                    }
                    return none;
                }
            }

            // runs handed to a hand-written kernel so far (for tests and diagnostics)
            inline std::atomic<long> &fast_runs() {
                static std::atomic<long> n{0};
                return n;
            }

            // GT355_ENGINE_FAST=0 keeps every spec on the generic engine (read once per process)
            inline bool fast_path_enabled() {
                static const bool on = [] {
                    const char *e = std::getenv("GT355_ENGINE_FAST");
                    return !(e && std::atoi(e) == 0);
                }();
                return on;
            }

            template <class... Args>
            bool run_fast(fast_plan const &p, bool W, run_geom const &g, void *stream, Args const &... args) {
                constexpr int nargs = sizeof...(Args);
                gt355_field f[nargs > 0 ? nargs : 1] = {};
                double gval = 0;
                int a = 0;
                auto one = [&](auto const &x) {
                    using X = std::decay_t<decltype(x)>;
                    if constexpr (is_store<X>) {
                        const bool w = (p.kind == fast_kind::hdiff && a == p.role[2]) ||
                                       (p.kind == fast_kind::vadv && a == p.role[0]);
                        f[a] = x->sid(g.origin(), w);
                    } else if constexpr (!std::is_void<typename arg_value<X>::type>::value) {
                        gval = (double)x.value;
                    }
                    ++a;
                };
                (one(args), ...);
                const int prec = W ? GT355_WIDEN : GT355_NATIVE;
                if (p.kind == fast_kind::hdiff) {
                    check(gt355_hdiff(prec, &f[p.role[2]], &f[p.role[0]], &f[p.role[1]], g.i_size, g.j_size, g.k_size,
                              stream),
                        "run (horizontal diffusion spec)");
                    ++fast_runs();
                    return true;
                }
                if (p.kind == fast_kind::vadv && g.k_size >= 2) {
                    check(gt355_vadv(prec, &f[p.role[0]], &f[p.role[1]], &f[p.role[2]], &f[p.role[3]], &f[p.role[4]],
                              gval, g.i_size, g.j_size, g.k_size, stream),
                        "run (vertical advection spec)");
                    ++fast_runs();
                    return true;
                }
                return false;
            }
        } // namespace engine_

        // run(spec, backend, grid, fields...) (stencil/frontend/run.hpp:201-250) for a spec lambda: the spec is called
        // with one placeholder per field argument and returns execute_*().stage(...) chains or a multi_pass of them.
        // The backend tag's precision selects the Deref policy: GT355_NATIVE computes in the storage type;
        // GT355_WIDEN (float storage) reads read-only fields and global parameters as double, keeps temporaries in
        // double and narrows on the stores to user fields.
        template <class Spec, class Be, int N, class... Args,
            std::enable_if_t<engine_::is_spec<engine_::spec_of<Spec, Args...>>::value, int> = 0>
        void run(Spec, Be be, grid_t<N> const &grid, Args const &... args) {
            constexpr bool W = Be::precision == GT355_WIDEN;
            using raw_spec_t = engine_::spec_of<Spec, Args...>;
            using spec_t = std::conditional_t<W, typename engine_::promote<raw_spec_t>::type, raw_spec_t>;
            using msses = typename engine_::msses_of<spec_t>::type;
            const engine_::run_geom g = engine_::geom_of(grid);
            std::vector<engine_::stage_rt> st;
            engine_::collect(msses(), st);
            auto box = engine_::analyse(st);
            engine_::run_ctx c;
            c.origin[0] = g.i_start, c.origin[1] = g.j_start, c.origin[2] = 0;
            auto args_t = std::forward_as_tuple(args...);
            engine_::bind_args(c, args_t, st, g, std::index_sequence_for<Args...>());
            if (g.i_size <= 0 || g.j_size <= 0 || g.k_size <= 0)
                return;
            void *stream = stream_of(be);
            if constexpr (N == 1) {
                // a spec that reproduces a hand-written kernel runs on it (decided once per spec type)
                static const engine_::fast_plan plan =
                    engine_::fast_path_enabled() ? engine_::plan_fast<W, msses>(st, args...) : engine_::fast_plan();
                if (plan.kind != engine_::fast_kind::none && engine_::run_fast(plan, W, g, stream, args...))
                    return;
            }
            // calls a deferred-run scope recorded on this stream run before the engine's own launches
            check(gt355_batch_flush(stream), "run: deferred calls");
            engine_::plan_all(msses(), st, box, c);
            std::vector<void *> taken;
            engine_::alloc_all(msses(), c, box, g, stream, taken);
            engine_::launch_all<N + 1, W>(msses(), st, c, args_t, g, (hipStream_t)stream);
            // every temporary of the run back to the pool with one event (one record per run, not per temporary)
            if (!taken.empty())
                check(gt355_tmp_release_n(taken.data(), (int)taken.size(), stream), "stencil engine: temporaries");
        }

        // run_single_stage(functor, backend, grid, fields...) (stencil/frontend/run.hpp:252-258)
        template <class F, class Be, int N, class... Args>
        void run_single_stage(F, Be be, grid_t<N> const &g, Args const &... args) {
            run([](auto... p) { return execute_parallel().stage(F(), p...); }, be, g, args...);
        }
    } // namespace stencil
} // namespace gt355
