// The cartesian regression stencils beyond horizontal diffusion and vertical advection, on gfx950:
//
//  * simple_hori_diff (tests/regression/simple_hori_diff.cpp:25-85): a metric-weighted laplacian (wlap) into an
//    ij-cached temporary, then its flux divergence (divflux).  crlato / crlatu are j-only fields
//    (selector<0,1,0>, stride 0 in i and k).  One wavefront sweeps 64 columns x JB rows of one k-level with the
//    rows in registers, the i-neighbours from DPP wavefront shifts and the wave-edge neighbours from side
//    columns, like the register sweep of hdiff; the j-only metric is one scalar load per row.
//  * prepare_tracers under expandable_run (advection_pdbott_prepare_tracers.cpp:23-57,
//    frontend/expandable_run.hpp:150-183): data[m] = rho * data_nnow[m] for every member m.  The reference runs
//    one launch per chunk of Factor members; here one launch covers up to 32 members and reads rho once per point.
//  * copy_stencil_tuple (copy_stencil_tuple.cpp:23-100, 116-260): out[f] = in[f] for a tuple of fields, one
//    launch per 16 pairs.  Any positive i stride is accepted, so the members of an array-of-structs storage
//    (element f of array<T, 5> at i stride 5) and the slices of an extra data dimension are fields too.
//  * extended_4D (extended_4D.cpp:58-135): per grid point, a projection over extra storage dimensions -- the 8
//    values f(a,b,c) and 2 jac(q) in, 8 results(I,J,K) out, with the 2x2x2x2 global-parameter tables phi / psi in
//    the kernel arguments.  One thread per point.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "functors.hpp"
#include "tuning.hpp"

namespace gt355 {
    namespace {

        // A field that may have masked dimensions (stride 0: one value for every index, the storages of
        // storage::builder::selector, storage/builder.hpp:168-326).  Unmasked dimensions are checked against the
        // offset box [lo, hi); a written field must have no masked dimension and an injective index map.
        int check_sel_field(const gt355_field *f, const char *name, int dtype, const int lo[3], const int hi[3],
            bool write) {
            GT355_CHECK(f != nullptr, "field %s: null descriptor", name);
            GT355_CHECK(f->origin != nullptr, "field %s: null origin", name);
            GT355_CHECK(f->dtype == dtype, "field %s: dtype %d, expected %d", name, f->dtype, dtype);
            for (int d = 0; d < 3; ++d) {
                GT355_CHECK(f->stride[d] >= 0, "field %s: negative stride in dimension %d", name, d);
                if (f->stride[d] == 0) {
                    GT355_CHECK(!write, "field %s: a written field cannot have a masked dimension (%d)", name, d);
                    continue;
                }
                GT355_CHECK(f->lower[d] <= lo[d] && hi[d] <= f->upper[d],
                    "field %s: offsets [%d,%d) in dimension %d outside bounds [%d,%d)", name, lo[d], hi[d], d,
                    f->lower[d], f->upper[d]);
            }
            GT355_CHECK(uintptr_t(f->origin) % (dtype == GT355_F64 ? 8 : 4) == 0, "field %s: misaligned origin", name);
            if (write) {
                // sorted by stride, each dimension must step over the whole extent of the previous one
                int ord[3] = {0, 1, 2};
                std::sort(ord, ord + 3, [&](int a, int b) { return f->stride[a] < f->stride[b]; });
                for (int q = 1; q < 3; ++q) {
                    const int a = ord[q - 1], b = ord[q];
                    GT355_CHECK(f->stride[b] >= f->stride[a] * (int64_t)(f->upper[a] - f->lower[a]),
                        "field %s: strides {%lld,%lld,%lld} make points overlap", name, (long long)f->stride[0],
                        (long long)f->stride[1], (long long)f->stride[2]);
                }
            }
            return GT355_OK;
        }

        // -------------------------------------------------------------------------------------------------------
        // simple_hori_diff register sweep.  Step t of the unrolled loop takes in-row R = j0-2+t; it evaluates
        // wlap for row R-1 (own column, and the side column of the lane's half: i0-1 on lanes 0-31, i0+64 on
        // lanes 32-63) and divflux for row R-2 from the laplacians of rows R-3..R-1.  All loads are clamped into
        // the readable boxes, so they are unconditional; only the store is guarded.
        //   in:       i [-2, ni+2), j [-2, nj+2)        metric (crlato, crlatu): i [-1, ni+1), j [-1, nj+1)
        //   coeff:    i [0, ni),    j [0, nj)           out: i [0, ni), j [0, nj)
        // CJ: both metric fields are j-only (i and k strides 0), so a row's metric is one uniform scalar load.
        // -------------------------------------------------------------------------------------------------------
        template <class ST, class CT, int JB, int P, bool CJ>
        __global__ void __launch_bounds__(256) shd_sweep_kernel(ST *__restrict__ out, const ST *__restrict__ in,
            const ST *__restrict__ cro, const ST *__restrict__ cru, const ST *__restrict__ coeff, int ni, int nj,
            int nk, int64_t osj, int64_t osk, int64_t isj, int64_t isk, int64_t csj, int64_t csk, int64_t ao0,
            int64_t ao1, int64_t ao2, int64_t au0, int64_t au1, int64_t au2, int strips_i, int strips_j, int nblocks,
            int nxcd) {
            constexpr int NR = JB + 4; // in rows j0-2 .. j0+JB+1
            const int lane = threadIdx.x & 63;
            const int wid =
                __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, nblocks, nxcd) * 4 + (threadIdx.x >> 6));
            if (wid >= strips_i * strips_j * nk)
                return; // a whole wavefront leaves; there is no barrier in this kernel
            const int si_ = wid % strips_i;
            const int sj_ = (wid / strips_i) % strips_j;
            const int k = wid / (strips_j * strips_i);
            const int i0 = si_ * 64, j0 = sj_ * JB;
            const bool left = lane < 32;
            const int ic = min(i0 + lane, ni + 1);
            const int ieL = i0 - 1, ifL = i0 - 2;
            const int ieR = min(i0 + 64, ni + 1), ifR = min(i0 + 65, ni + 1);
            const int mc = min(i0 + lane, ni);                 // own metric column
            const int ms = left ? i0 - 1 : min(i0 + 64, ni);  // side metric column of the lane's half
            const ST *inp = in + (int64_t)k * isk;
            const ST *cop = coeff + (int64_t)k * csk;
            const ST *rop = cro + (int64_t)k * ao2;
            const ST *rup = cru + (int64_t)k * au2;

            ST rc[P], rk[P], reL[P], rfL[P], reR[P], rfR[P]; // in (own, side columns), coeff
            ST ro[P], ru[P], ros[P], rus[P];                 // metric of row R-1 (own column, side column)
            auto load = [&](int t) {
                const int64_t r = min(j0 - 2 + t, nj + 1);
                rc[t % P] = inp[r * isj + ic];
                reL[t % P] = inp[r * isj + ieL];
                rfL[t % P] = inp[r * isj + ifL];
                reR[t % P] = inp[r * isj + ieR];
                rfR[t % P] = inp[r * isj + ifR];
                const int64_t rout = min(max(j0 - 4 + t, 0), nj - 1); // the row output at step t
                rk[t % P] = __builtin_nontemporal_load(cop + rout * csj + min(ic, ni - 1)); // read once
                const int64_t rm = min(max(j0 - 3 + t, -1), nj); // the row whose laplacian step t evaluates
                if constexpr (CJ) {
                    ro[t % P] = rop[rm * ao1];
                    ru[t % P] = rup[rm * au1];
                } else {
                    ro[t % P] = rop[rm * ao1 + mc * ao0];
                    ru[t % P] = rup[rm * au1 + mc * au0];
                    ros[t % P] = rop[rm * ao1 + ms * ao0];
                    rus[t % P] = rup[rm * au1 + ms * au0];
                }
            };
#pragma clang loop unroll(full)
            for (int t = 0; t < P; ++t)
                load(t);

            CT c0 = 0, c1 = 0, c2 = 0;    // in rows R-2, R-1, R of the own column
            CT eL0 = 0, eL1 = 0, eL2 = 0; // the same rows of column i0-1
            CT eR0 = 0, eR1 = 0, eR2 = 0; // the same rows of column i0+64
            CT fL1 = 0, fR1 = 0;          // columns i0-2 and i0+65 of row R-1
            CT m1 = 0, p1 = 0;            // in(i-1), in(i+1) of row R-1
            CT lA = 0, lB = 0, lC = 0;    // laplacian of rows R-3, R-2, R-1 (own column)
            CT slB = 0, slC = 0;          // side-column laplacian of rows R-2, R-1
            CT o_prev = 0;                // crlato of row R-2 (own column): the divflux metric of the output row
            ST *orow = out + (int64_t)k * osk + (i0 + lane);
            const bool col_ok = i0 + lane < ni;
#pragma clang loop unroll(full)
            for (int t = 0; t < NR; ++t) {
                c0 = c1; c1 = c2; c2 = CT(rc[t % P]);
                eL0 = eL1; eL1 = eL2; eL2 = CT(reL[t % P]);
                eR0 = eR1; eR1 = eR2; eR2 = CT(reR[t % P]);
                const CT kout = CT(rk[t % P]);
                const CT fL_r = CT(rfL[t % P]), fR_r = CT(rfR[t % P]);
                const CT o_c = CT(ro[t % P]), u_c = CT(ru[t % P]);
                CT os_c = o_c, us_c = u_c;
                if constexpr (!CJ) {
                    os_c = CT(ros[t % P]);
                    us_c = CT(rus[t % P]);
                }
                if (t + P < NR)
                    load(t + P);
                // in(i-1), in(i+1) of row R: lane 0 keeps column i0-1, lane 63 keeps column i0+64
                const CT m2 = dpp_keep<DPP_WF_SR1>(eL2, c2), p2 = dpp_keep<DPP_WF_SL1>(eR2, c2);
                if (t >= 2) {
                    // wlap of row R-1 (simple_hori_diff.cpp:33-39)
                    const CT lnew = wlap_f<CT>(c1, p1, m1, c2, c0, o_c, u_c);
                    // at the side column the own column is the i+1 neighbour (left half) or the i-1 neighbour
                    // (right half)
                    const CT e0 = left ? eL0 : eR0, e1 = left ? eL1 : eR1, e2 = left ? eL2 : eR2;
                    const CT ip = left ? c1 : fR1, im = left ? fL1 : c1;
                    const CT snew = wlap_f<CT>(e1, ip, im, e2, e0, os_c, us_c);
                    lA = lB;
                    lB = lC;
                    lC = lnew;
                    slB = slC;
                    slC = snew;
                }
                if (t >= 4) {
                    // divflux of row R-2 (simple_hori_diff.cpp:51-60)
                    const CT lm = dpp_keep<DPP_WF_SR1>(slB, lB), lp = dpp_keep<DPP_WF_SL1>(slB, lB);
                    const CT res = divflux_f<CT>(c0, lB, lp, lm, lC, lA, o_prev, kout);
                    const int r_out = j0 + t - 4;
                    if (col_ok && r_out < nj)
                        __builtin_nontemporal_store(static_cast<ST>(res), orow + (int64_t)r_out * osj);
                }
                o_prev = o_c;
                m1 = m2;
                p1 = p2;
                fL1 = fL_r;
                fR1 = fR_r;
            }
        }

        template <class ST, class CT, int JB, bool CJ>
        int launch_shd(const gt355_field *out, const gt355_field *in, const gt355_field *cro, const gt355_field *cru,
            const gt355_field *coeff, int ni, int nj, int nk, hipStream_t stream) {
            constexpr int P = 4;
            const int strips_i = (ni + 63) / 64, strips_j = (nj + JB - 1) / JB;
            const int64_t waves = (int64_t)strips_i * strips_j * nk;
            const int64_t blocks = (waves + 3) / 4;
            GT355_CHECK(waves < (1LL << 31), "gt355_simple_hori_diff: domain too large for one launch");
            shd_sweep_kernel<ST, CT, JB, P, CJ><<<(int)blocks, 256, 0, stream>>>((ST *)out->origin,
                (const ST *)in->origin, (const ST *)cro->origin, (const ST *)cru->origin, (const ST *)coeff->origin,
                ni, nj, nk, out->stride[1], out->stride[2], in->stride[1], in->stride[2], coeff->stride[1],
                coeff->stride[2], cro->stride[0], cro->stride[1], cro->stride[2], cru->stride[0], cru->stride[1],
                cru->stride[2], strips_i, strips_j, (int)blocks, xcd_count());
            return check_launch("gt355_simple_hori_diff");
        }

        template <class ST, class CT>
        int dispatch_shd(const gt355_field *out, const gt355_field *in, const gt355_field *cro,
            const gt355_field *cru, const gt355_field *coeff, int ni, int nj, int nk, hipStream_t stream) {
            const bool cj = cro->stride[0] == 0 && cro->stride[2] == 0 && cru->stride[0] == 0 && cru->stride[2] == 0;
            const int jb = tune_param("shd_jb", 16);
#define GT355_SHD(JB)                                                                                            \
    if (jb == JB)                                                                                                \
        return cj ? launch_shd<ST, CT, JB, true>(out, in, cro, cru, coeff, ni, nj, nk, stream)                   \
                  : launch_shd<ST, CT, JB, false>(out, in, cro, cru, coeff, ni, nj, nk, stream);
            GT355_SHD(8)
            GT355_SHD(16)
            GT355_SHD(32)
#undef GT355_SHD
            set_error("gt355_simple_hori_diff: rows per wave %d is not instantiated (8, 16, 32)", jb);
            return GT355_EUNSUP;
        }

        // -------------------------------------------------------------------------------------------------------
        // prepare_tracers over up to TRACER_MAX members: one thread per point, rho read once, member values
        // loaded U at a time before their stores (the member pointers may alias each other, so the compiler
        // would not hoist the loads by itself).
        // -------------------------------------------------------------------------------------------------------
        constexpr int TRACER_MAX = 32;
        template <class T>
        struct tracer_table {
            T *data[TRACER_MAX];
            const T *nnow[TRACER_MAX];
        };

        template <class ST, class CT, int U>
        __global__ void __launch_bounds__(256) tracers_kernel(tracer_table<ST> tab, int nm, const ST *__restrict__ rho,
            int ni, int nj, int nk, int64_t dsj, int64_t dsk, int64_t nsj, int64_t nsk, int64_t rsj, int64_t rsk,
            int segs) {
            const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
            const int64_t w = n >> 6;
            const int i = int(w % segs) * 64 + int(n & 63);
            const int64_t row = w / segs;
            const int j = int(row % nj), k = int(row / nj);
            if (k >= nk || i >= ni)
                return;
            const CT r = CT(rho[i + j * rsj + k * rsk]);
            const int64_t od = i + j * dsj + k * dsk, on = i + j * nsj + k * nsk;
            for (int m0 = 0; m0 < nm; m0 += U) {
                CT v[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    v[u] = m0 + u < nm ? CT(__builtin_nontemporal_load(tab.nnow[m0 + u] + on)) : CT(0);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (m0 + u < nm)
                        __builtin_nontemporal_store(static_cast<ST>(r * v[u]), tab.data[m0 + u] + od);
            }
        }

        // -------------------------------------------------------------------------------------------------------
        // Field-tuple copy: one thread per point, every pair of the table in turn; W is the element word.
        // -------------------------------------------------------------------------------------------------------
        constexpr int COPY_MAX = 16;
        struct copy_pair {
            char *out;
            const char *in;
            int64_t os[3], is[3];
        };
        struct copy_table {
            copy_pair p[COPY_MAX];
        };

        template <class W>
        __global__ void __launch_bounds__(256) copy_fields_kernel(copy_table tab, int np, int ni, int nj, int nk,
            int segs) {
            const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
            const int64_t w = n >> 6;
            const int i = int(w % segs) * 64 + int(n & 63);
            const int64_t row = w / segs;
            const int j = int(row % nj), k = int(row / nj);
            if (k >= nk || i >= ni)
                return;
            for (int p = 0; p < np; ++p) {
                const copy_pair &c = tab.p[p];
                // every value is read and written once: non-temporal on both sides
                const W v = __builtin_nontemporal_load((const W *)c.in + (i * c.is[0] + j * c.is[1] + k * c.is[2]));
                __builtin_nontemporal_store(v, (W *)c.out + (i * c.os[0] + j * c.os[1] + k * c.os[2]));
            }
        }

        // -------------------------------------------------------------------------------------------------------
        // extended_4D integration, the operation order of integration::apply (extended_4D.cpp:67-90): sum and res in
        // the storage type, products promoted to double by the double-valued global parameters.
        // -------------------------------------------------------------------------------------------------------
        struct ext4d_tables {
            double phi[16], psi[16]; // [((x*2 + y)*2 + z)*2 + q]
        };

        template <class T>
        __global__ void __launch_bounds__(256) ext4d_kernel(ext4d_tables tab, T *__restrict__ result,
            const T *__restrict__ f, const T *__restrict__ jac, int ni, int nj, int nk, int64_t rsj, int64_t rsk,
            int64_t rx0, int64_t rx1, int64_t rx2, int64_t fsj, int64_t fsk, int64_t fx0, int64_t fx1, int64_t fx2,
            int64_t jsj, int64_t jsk, int64_t jq, int segs) {
            const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
            const int64_t w = n >> 6;
            const int i = int(w % segs) * 64 + int(n & 63);
            const int64_t row = w / segs;
            const int j = int(row % nj), k = int(row / nj);
            if (k >= nk || i >= ni)
                return;
            const T *fp = f + i + j * fsj + k * fsk;
            T fv[8];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        fv[(a * 2 + b) * 2 + c] = __builtin_nontemporal_load(fp + (a * fx0 + b * fx1 + c * fx2));
            const T *jp = jac + i + j * jsj + k * jsk;
            const T jv[2] = {__builtin_nontemporal_load(jp), __builtin_nontemporal_load(jp + jq)};
            // sum over (a, b, c) for each quadrature point: independent of (I, J, K)
            T sum[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                T s_ = 0;
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    s_ += tab.psi[e * 2 + q] * fv[e];
                sum[q] = s_;
            }
            T *rp = result + i + j * rsj + k * rsk;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int z = 0; z < 2; ++z) {
                        T res = 0;
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            res += tab.phi[((x * 2 + y) * 2 + z) * 2 + q] * jv[q] * sum[q];
                        __builtin_nontemporal_store(T(res / 8), rp + (x * rx0 + y * rx1 + z * rx2));
                    }
        }

        // grid of 64-lane row segments covering ni x nj x nk points
        int row_grid(const char *who, int ni, int nj, int nk, int *segs, int *blocks) {
            *segs = (ni + 63) / 64;
            const int64_t waves = (int64_t)*segs * nj * nk;
            GT355_CHECK(waves < (1LL << 33), "%s: domain too large for one launch", who);
            const int64_t b = (waves + 3) / 4;
            GT355_CHECK(b < (1LL << 31), "%s: domain too large for one launch", who);
            *blocks = (int)b;
            return GT355_OK;
        }
    } // namespace
} // namespace gt355

using namespace gt355;

extern "C" int gt355_simple_hori_diff(int precision, const gt355_field *out, const gt355_field *in,
    const gt355_field *crlato, const gt355_field *crlatu, const gt355_field *coeff, int ni, int nj, int nk,
    void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0, "gt355_simple_hori_diff: negative size");
    GT355_CHECK(out && in && crlato && crlatu && coeff, "gt355_simple_hori_diff: null field");
    const int dtype = in->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_simple_hori_diff: dtype %d", dtype);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dtype == GT355_F32),
        "gt355_simple_hori_diff: precision %d with dtype %d", precision, dtype);
    if (ni == 0 || nj == 0 || nk == 0)
        return GT355_OK;
    int rc;
    const int mlo[3] = {-1, -1, 0}, mhi[3] = {ni + 1, nj + 1, nk};
    if ((rc = check_field(in, "in", dtype, -2, ni + 2, -2, nj + 2, 0, nk)) ||
        (rc = check_field(coeff, "coeff", dtype, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(out, "out", dtype, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_sel_field(crlato, "crlato", dtype, mlo, mhi, false)) ||
        (rc = check_sel_field(crlatu, "crlatu", dtype, mlo, mhi, false)))
        return rc;
    GT355_CHECK(out->origin != in->origin, "gt355_simple_hori_diff: out must not alias in");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == GT355_F64)
        return dispatch_shd<double, double>(out, in, crlato, crlatu, coeff, ni, nj, nk, s);
    if (precision == GT355_WIDEN)
        return dispatch_shd<float, double>(out, in, crlato, crlatu, coeff, ni, nj, nk, s);
    return dispatch_shd<float, float>(out, in, crlato, crlatu, coeff, ni, nj, nk, s);
}

extern "C" int gt355_prepare_tracers(int precision, const gt355_field *data, const gt355_field *data_nnow,
    int nmembers, const gt355_field *rho, int ni, int nj, int nk, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0 && nmembers >= 0, "gt355_prepare_tracers: negative size");
    GT355_CHECK(rho && (nmembers == 0 || (data && data_nnow)), "gt355_prepare_tracers: null field");
    const int dtype = rho->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_prepare_tracers: dtype %d", dtype);
    GT355_CHECK(precision == GT355_NATIVE || (precision == GT355_WIDEN && dtype == GT355_F32),
        "gt355_prepare_tracers: precision %d with dtype %d", precision, dtype);
    if (ni == 0 || nj == 0 || nk == 0 || nmembers == 0)
        return GT355_OK;
    int rc;
    if ((rc = check_field(rho, "rho", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    for (int m = 0; m < nmembers; ++m) {
        if ((rc = check_field(&data[m], "data", dtype, 0, ni, 0, nj, 0, nk)) ||
            (rc = check_field(&data_nnow[m], "data_nnow", dtype, 0, ni, 0, nj, 0, nk)))
            return rc;
        GT355_CHECK(data[m].stride[1] == data[0].stride[1] && data[m].stride[2] == data[0].stride[2] &&
                        data_nnow[m].stride[1] == data_nnow[0].stride[1] &&
                        data_nnow[m].stride[2] == data_nnow[0].stride[2],
            "gt355_prepare_tracers: member %d has a different layout from member 0", m);
    }
    int segs, blocks;
    if ((rc = row_grid("gt355_prepare_tracers", ni, nj, nk, &segs, &blocks)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    for (int m0 = 0; m0 < nmembers; m0 += TRACER_MAX) {
        const int nm = std::min(TRACER_MAX, nmembers - m0);
#define GT355_TRACERS(ST, CT)                                                                                    \
    do {                                                                                                         \
        tracer_table<ST> tab{};                                                                                  \
        for (int m = 0; m < nm; ++m) {                                                                           \
            tab.data[m] = (ST *)data[m0 + m].origin;                                                             \
            tab.nnow[m] = (const ST *)data_nnow[m0 + m].origin;                                                  \
        }                                                                                                        \
        tracers_kernel<ST, CT, 4><<<blocks, 256, 0, s>>>(tab, nm, (const ST *)rho->origin, ni, nj, nk,           \
            data[0].stride[1], data[0].stride[2], data_nnow[0].stride[1], data_nnow[0].stride[2], rho->stride[1],   \
            rho->stride[2], segs);                                                                               \
    } while (0)
        if (dtype == GT355_F64)
            GT355_TRACERS(double, double);
        else if (precision == GT355_WIDEN)
            GT355_TRACERS(float, double);
        else
            GT355_TRACERS(float, float);
#undef GT355_TRACERS
        if ((rc = check_launch("gt355_prepare_tracers")))
            return rc;
    }
    return GT355_OK;
}

extern "C" int gt355_copy_fields(const gt355_field *out, const gt355_field *in, int nfields, int ni, int nj, int nk,
    void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0 && nfields >= 0, "gt355_copy_fields: negative size");
    GT355_CHECK(nfields == 0 || (out && in), "gt355_copy_fields: null field");
    if (ni == 0 || nj == 0 || nk == 0 || nfields == 0)
        return GT355_OK;
    const int dtype = in[0].dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_copy_fields: dtype %d", dtype);
    const int lo[3] = {0, 0, 0}, hi[3] = {ni, nj, nk};
    int rc;
    for (int f = 0; f < nfields; ++f) {
        if ((rc = check_sel_field(&in[f], "in", dtype, lo, hi, false)) ||
            (rc = check_sel_field(&out[f], "out", dtype, lo, hi, true)))
            return rc;
    }
    int segs, blocks;
    if ((rc = row_grid("gt355_copy_fields", ni, nj, nk, &segs, &blocks)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    for (int f0 = 0; f0 < nfields; f0 += COPY_MAX) {
        const int np = std::min(COPY_MAX, nfields - f0);
        copy_table tab{};
        for (int p = 0; p < np; ++p) {
            tab.p[p].out = (char *)out[f0 + p].origin;
            tab.p[p].in = (const char *)in[f0 + p].origin;
            for (int d = 0; d < 3; ++d) {
                tab.p[p].os[d] = out[f0 + p].stride[d];
                tab.p[p].is[d] = in[f0 + p].stride[d];
            }
        }
        if (dtype == GT355_F64)
            copy_fields_kernel<uint64_t><<<blocks, 256, 0, s>>>(tab, np, ni, nj, nk, segs);
        else
            copy_fields_kernel<uint32_t><<<blocks, 256, 0, s>>>(tab, np, ni, nj, nk, segs);
        if ((rc = check_launch("gt355_copy_fields")))
            return rc;
    }
    return GT355_OK;
}

extern "C" int gt355_extended_4d(const double phi[16], const double psi[16], const gt355_field *result,
    const int64_t result_extra[3], const gt355_field *f, const int64_t f_extra[3], const gt355_field *jac,
    int64_t jac_extra, int ni, int nj, int nk, void *stream) {
    GT355_DEFER_FLUSH(stream);
    GT355_CHECK(ni >= 0 && nj >= 0 && nk >= 0, "gt355_extended_4d: negative size");
    GT355_CHECK(phi && psi && result && result_extra && f && f_extra && jac, "gt355_extended_4d: null argument");
    const int dtype = result->dtype;
    GT355_CHECK(dtype == GT355_F64 || dtype == GT355_F32, "gt355_extended_4d: dtype %d", dtype);
    if (ni == 0 || nj == 0 || nk == 0)
        return GT355_OK;
    int rc;
    if ((rc = check_field(result, "result", dtype, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(f, "f", dtype, 0, ni, 0, nj, 0, nk)) ||
        (rc = check_field(jac, "jac", dtype, 0, ni, 0, nj, 0, nk)))
        return rc;
    // every extra-dimension slice must lie beyond the spatial box of the previous one (no overlapping points)
    auto vol = [](const gt355_field *x) { return x->stride[2] * (int64_t)(x->upper[2] - x->lower[2]); };
    GT355_CHECK(jac_extra >= vol(jac), "gt355_extended_4d: jac's quadrature stride overlaps its spatial box");
    for (int d = 0; d < 3; ++d) {
        GT355_CHECK(f_extra[d] >= 0, "gt355_extended_4d: negative extra stride of f");
        GT355_CHECK(result_extra[d] > 0, "gt355_extended_4d: result extra strides must be positive");
    }
    {
        int64_t r[3] = {result_extra[0], result_extra[1], result_extra[2]};
        std::sort(r, r + 3);
        GT355_CHECK(r[0] >= vol(result) && r[1] >= 2 * r[0] && r[2] >= 2 * r[1],
            "gt355_extended_4d: result's extra strides make points overlap");
    }
    int segs, blocks;
    if ((rc = row_grid("gt355_extended_4d", ni, nj, nk, &segs, &blocks)))
        return rc;
    ext4d_tables tab;
    for (int e = 0; e < 16; ++e) {
        tab.phi[e] = phi[e];
        tab.psi[e] = psi[e];
    }
    hipStream_t s = (hipStream_t)stream;
#define GT355_EXT4D(T)                                                                                           \
    ext4d_kernel<T><<<blocks, 256, 0, s>>>(tab, (T *)result->origin, (const T *)f->origin, (const T *)jac->origin,  \
        ni, nj, nk, result->stride[1], result->stride[2], result_extra[0], result_extra[1], result_extra[2],          \
        f->stride[1], f->stride[2], f_extra[0], f_extra[1], f_extra[2], jac->stride[1], jac->stride[2], jac_extra,   \
        segs)
    if (dtype == GT355_F64)
        GT355_EXT4D(double);
    else
        GT355_EXT4D(float);
#undef GT355_EXT4D
    return check_launch("gt355_extended_4d");
}
