/*
 * gt_oracle.c -- CPU restatement of the reference hdiff / vadv / laplacian / tridiagonal /
 * boundary-condition arithmetic.  TEST INFRASTRUCTURE ONLY (see gt_oracle.h): the checker
 * for the HIP path and the "port" CPU baseline, never part of the product.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off; no -ffast-math, so every
 * operation is a correctly rounded IEEE operation in source order).
 */
#include "gt_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------
 * Inputs.  Counter-based SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014): element n of a field is mix(base + (n+1)*gamma) with
 * base = seed ^ (field_id * 0xD1B54A32D192ED03); u = (x >> 11) * 2^-53; value = a + (b-a)*u.
 * The reference fills every index of the allocation with its initializer
 * (storage/builder.hpp:67-84); n runs over the logical box, i fastest, then j, then k.
 * ------------------------------------------------------------------------------------------ */
static uint64_t gto_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void gto_fill_uniform_f64(double *dst, int64_t n, uint64_t seed, uint64_t field_id, double a, double b) {
    const uint64_t base = seed ^ (field_id * 0xD1B54A32D192ED03ULL);
    const double w = b - a;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t x = gto_mix64(base + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        double u = (double)(x >> 11) * 0x1.0p-53;
        dst[i] = a + w * u;
    }
}

/* ---------------------------------- horizontal diffusion ---------------------------------- */
#define ST double
#define CT double
#define FN gto_hdiff_f64
#include "hdiff_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT float
#define FN gto_hdiff_f32
#include "hdiff_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT double
#define FN gto_hdiff_mixed
#include "hdiff_body.inc"
#undef ST
#undef CT
#undef FN

/* ---------------------------------- vertical advection ---------------------------------- */
#define ST double
#define CT double
#define FN gto_vadv_f64
#include "vadv_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT float
#define FN gto_vadv_f32
#include "vadv_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT double
#define FN gto_vadv_mixed
#include "vadv_body.inc"
#undef ST
#undef CT
#undef FN

/* ---------------------------------- laplacian ----------------------------------
 * tests/regression/laplacian.cpp:20-29: out = 4*in - (in[i+1] + in[j+1] + in[i-1] + in[j-1]). */
void gto_lap_f64(double *out, const double *in, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            const double *p = in + (ptrdiff_t)k * sk + (ptrdiff_t)j * sj;
            double *o = out + (ptrdiff_t)k * sk + (ptrdiff_t)j * sj;
            for (int i = 0; i < ni; ++i)
                o[i] = 4 * p[i] - (((p[i + 1] + p[i + sj]) + p[i - 1]) + p[i - sj]);
        }
}

/* the float_t = float instantiation of the same functor: every operation in FP32 */
void gto_lap_f32(float *out, const float *in, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            const float *p = in + (ptrdiff_t)k * sk + (ptrdiff_t)j * sj;
            float *o = out + (ptrdiff_t)k * sk + (ptrdiff_t)j * sj;
            for (int i = 0; i < ni; ++i)
                o[i] = 4 * p[i] - (((p[i + 1] + p[i + sj]) + p[i - 1]) + p[i - sj]);
        }
}

/* ---------------------------------- simple_hori_diff ---------------------------------- */
#define ST double
#define CT double
#define FN gto_shd_f64
#include "shd_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT float
#define FN gto_shd_f32
#include "shd_body.inc"
#undef ST
#undef CT
#undef FN

#define ST float
#define CT double
#define FN gto_shd_mixed
#include "shd_body.inc"
#undef ST
#undef CT
#undef FN

/* ---------------------------------- prepare_tracers ----------------------------------
 * tests/regression/advection_pdbott_prepare_tracers.cpp:23-34: data = rho * data_nnow, one member. */
void gto_prepare_tracers_f64(double *data, const double *data_nnow, const double *rho, int ni, int nj, int nk,
    ptrdiff_t sj, ptrdiff_t sk) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const ptrdiff_t p = (ptrdiff_t)k * sk + (ptrdiff_t)j * sj + i;
                data[p] = rho[p] * data_nnow[p];
            }
}

void gto_prepare_tracers_f32(float *data, const float *data_nnow, const float *rho, int ni, int nj, int nk,
    ptrdiff_t sj, ptrdiff_t sk) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const ptrdiff_t p = (ptrdiff_t)k * sk + (ptrdiff_t)j * sj + i;
                data[p] = rho[p] * data_nnow[p];
            }
}

/* ---------------------------------- extended_4D ----------------------------------
 * tests/regression/extended_4D.cpp:58-91 (integration::apply), one grid point at a time:
 *   for I, J, K in {0,1}: res = 0; for q: sum = 0; for a, b, c: sum += psi(a,b,c,q) * f(a,b,c);
 *                                     res += phi(I,J,K,q) * jac(q) * sum;
 *                         result(I,J,K) = res / 8.
 * phi / psi are the global parameters' values (the reference's elemental returns double), indexed
 * [((x*2 + y)*2 + z)*2 + q].  jac carries the quadrature index q at stride jq; f and result carry (a,b,c) /
 * (I,J,K) at strides fs[3] / rs[3].  The float instantiation keeps sum and res in float and the products in
 * double (double * float promotes), exactly as the C++ expression evaluation does. */
#define GTO_EXT4D_BODY(T)                                                                                        \
    for (int k = 0; k < nk; ++k)                                                                                 \
        for (int j = 0; j < nj; ++j)                                                                             \
            for (int i = 0; i < ni; ++i) {                                                                       \
                const ptrdiff_t p = (ptrdiff_t)k * sk + (ptrdiff_t)j * sj + i;                                   \
                for (int I = 0; I < 2; ++I)                                                                      \
                    for (int J = 0; J < 2; ++J)                                                                  \
                        for (int K = 0; K < 2; ++K) {                                                            \
                            T res = 0;                                                                           \
                            for (int q = 0; q < 2; ++q) {                                                        \
                                T sum = 0;                                                                       \
                                for (int a = 0; a < 2; ++a)                                                      \
                                    for (int b = 0; b < 2; ++b)                                                  \
                                        for (int c = 0; c < 2; ++c)                                              \
                                            sum += psi[((a * 2 + b) * 2 + c) * 2 + q] *                          \
                                                   f[p + a * fs[0] + b * fs[1] + c * fs[2]];                     \
                                res += phi[((I * 2 + J) * 2 + K) * 2 + q] * jac[p + q * jq] * sum;               \
                            }                                                                                    \
                            result[p + I * rs[0] + J * rs[1] + K * rs[2]] = res / 8;                             \
                        }                                                                                        \
            }

void gto_extended_4d_f64(double *result, const ptrdiff_t rs[3], const double *f, const ptrdiff_t fs[3],
    const double *jac, ptrdiff_t jq, const double phi[16], const double psi[16], int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk) {
    GTO_EXT4D_BODY(double)
}

void gto_extended_4d_f32(float *result, const ptrdiff_t rs[3], const float *f, const ptrdiff_t fs[3],
    const float *jac, ptrdiff_t jq, const double phi[16], const double psi[16], int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk) {
    GTO_EXT4D_BODY(float)
}

/* ---------------------------------- k-first traversals (cpu_kfirst) ---------------------------------- */
#include "kfirst.inc"

/* ---------------------------------- transform_layout ----------------------------------
 * layout_transformation/cpu.hpp:20-55 (transform_cpu_loop): every index of the box dims, the first three dimensions
 * in an (i, j, k) loop nest, the remaining ones as an outer hypercube; elements are moved as raw bytes. */
void gto_transform_layout(int elem_size, void *dst, const void *src, int ndims, const int64_t *dims,
    const int64_t *dst_strides, const int64_t *src_strides) {
    int64_t idx[16] = {0};
    int64_t total = 1;
    if (ndims < 1 || ndims > 16)
        return;
    for (int d = 0; d < ndims; ++d)
        total *= dims[d];
    for (int64_t n = 0; n < total; ++n) {
        int64_t rem = n, so = 0, dof = 0;
        for (int d = 0; d < ndims; ++d) {
            idx[d] = rem % dims[d];
            rem /= dims[d];
            so += idx[d] * src_strides[d];
            dof += idx[d] * dst_strides[d];
        }
        memcpy((char *)dst + dof * elem_size, (const char *)src + so * elem_size, (size_t)elem_size);
    }
}

/* ---------------------------------- tridiagonal ----------------------------------
 * tests/regression/tridiagonal.cpp:33-68.  forward_thomas: first_level sup/=diag, rhs/=diag;
 * modify<1,0>: sup = sup/(diag - sup[k-1]*inf), rhs = (rhs - inf*rhs[k-1])/(diag - sup[k-1]*inf).
 * backward_thomas: last_level out = rhs; modify<0,-1>: out = rhs - sup*out[k+1]. */
void gto_tridiagonal_f64(double *out, const double *inf, const double *diag, double *sup, double *rhs, int ni,
    int nj, int nk, ptrdiff_t sj, ptrdiff_t sk) {
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const ptrdiff_t c = (ptrdiff_t)j * sj + i;
            for (int k = 0; k < nk; ++k) {
                const ptrdiff_t p = c + (ptrdiff_t)k * sk;
                if (k == 0) {
                    sup[p] = sup[p] / diag[p];
                    rhs[p] = rhs[p] / diag[p];
                } else {
                    const ptrdiff_t m = p - sk;
                    sup[p] = sup[p] / (diag[p] - sup[m] * inf[p]);
                    rhs[p] = (rhs[p] - inf[p] * rhs[m]) / (diag[p] - sup[m] * inf[p]);
                }
            }
            for (int k = nk - 1; k >= 0; --k) {
                const ptrdiff_t p = c + (ptrdiff_t)k * sk;
                out[p] = (k == nk - 1) ? rhs[p] : rhs[p] - sup[p] * out[p + sk];
            }
        }
}

/* ---------------------------------- boundary conditions ---------------------------------- */
/* direction component of index x against halo_descriptor h = {minus, plus, begin, end, total}:
 * -1 on [begin-minus, begin-1], 0 on [begin, end], +1 on [end+1, end+plus], 2 = untouched. */
static int gto_dir(int x, const int h[5]) {
    if (x >= h[2] && x <= h[3])
        return 0;
    if (x >= h[2] - h[0] && x < h[2])
        return -1;
    if (x > h[3] && x <= h[3] + h[1])
        return 1;
    return 2;
}

#define GTO_BC_BODY(T)                                                                                           \
    for (int k = 0; k < halo[2][4]; ++k) {                                                                     \
        int dk = gto_dir(k, halo[2]);                                                                          \
        if (dk == 2)                                                                                           \
            continue;                                                                                          \
        for (int j = 0; j < halo[1][4]; ++j) {                                                                 \
            int dj = gto_dir(j, halo[1]);                                                                      \
            if (dj == 2)                                                                                       \
                continue;                                                                                      \
            for (int i = 0; i < halo[0][4]; ++i) {                                                             \
                int di = gto_dir(i, halo[0]);                                                                  \
                if (di == 2 || (di == 0 && dj == 0 && dk == 0))                                                \
                    continue;                                                                                  \
                int bit = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);                                              \
                if (!((dir_mask >> bit) & 1u))                                                                 \
                    continue;                                                                                  \
                ptrdiff_t p = (ptrdiff_t)k * sk + (ptrdiff_t)j * sj + i;                                       \
                dst[p] = kind == 0 ? (T)0 : kind == 1 ? value : src[p];                                        \
            }                                                                                                  \
        }                                                                                                      \
    }

void gto_bc_f64(int kind, double *dst, const double *src, double value, const int halo[3][5], ptrdiff_t sj,
    ptrdiff_t sk, uint32_t dir_mask) {
    GTO_BC_BODY(double)
}

void gto_bc_f32(int kind, float *dst, const float *src, float value, const int halo[3][5], ptrdiff_t sj,
    ptrdiff_t sk, uint32_t dir_mask) {
    GTO_BC_BODY(float)
}

/* maps x into [begin, end] for a dimension in whose halo x lies; returns -1 if x is not to be filled */
static int gto_map(int x, const int h[5], int mode, int *in_halo) {
    int d = gto_dir(x, h);
    *in_halo = 0;
    if (d == 0)
        return x;
    if (d == 2 || mode == 0)
        return -1;
    *in_halo = 1;
    if (mode == 1)
        return d < 0 ? h[2] : h[3];
    int len = h[3] - h[2] + 1;
    return d < 0 ? x + len : x - len;
}

#define GTO_FILL_BODY                                                                                            \
    for (int k = 0; k < halo[2][4]; ++k)                                                                       \
        for (int j = 0; j < halo[1][4]; ++j)                                                                   \
            for (int i = 0; i < halo[0][4]; ++i) {                                                             \
                int hi, hj, hk;                                                                                \
                int si = gto_map(i, halo[0], mode[0], &hi);                                                    \
                int sj_ = gto_map(j, halo[1], mode[1], &hj);                                                   \
                int sk_ = gto_map(k, halo[2], mode[2], &hk);                                                   \
                if (si < 0 || sj_ < 0 || sk_ < 0 || !(hi || hj || hk))                                         \
                    continue;                                                                                  \
                f[(ptrdiff_t)k * sk + (ptrdiff_t)j * sj + i] = f[(ptrdiff_t)sk_ * sk + (ptrdiff_t)sj_ * sj + si]; \
            }

void gto_halo_fill_f64(double *f, const int halo[3][5], ptrdiff_t sj, ptrdiff_t sk, const int mode[3]) {
    GTO_FILL_BODY
}

void gto_halo_fill_f32(float *f, const int halo[3][5], ptrdiff_t sj, ptrdiff_t sk, const int mode[3]) {
    GTO_FILL_BODY
}
