/*
 * gt_oracle.h -- CPU restatement of the GridTools hdiff / vadv / laplacian / boundary path and of the
 * cartesian regression stencils simple_hori_diff, prepare_tracers and extended_4D, and of transform_layout.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The product
 * (gridtools_amd, libgt355.so) never links, loads or calls it.
 *
 * Every function restates the arithmetic of the reference functor it names, in
 * the same operation order, so that an IEEE build (-O2 -ffp-contract=off, no
 * -ffast-math) produces the bits the reference cpu_ifirst / naive backends would.
 * The reference itself cannot be built in this image (it needs Boost.Preprocessor,
 * which is absent), so the restatement is pinned by the reference's own closed
 * forms and known answers plus mathematical invariants (see tests/test_oracle_*.py).
 *
 * Field convention: a field is passed as a pointer to its interior origin
 * (storage index (halo_i, halo_j, 0)) plus element strides sj, sk; the i stride is 1.
 * Halo points are reached with negative offsets.
 */
#ifndef GT_ORACLE_H
#define GT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- inputs: counter-based SplitMix64 -> U(a,b) over a logical (i fastest, j, k) box ---- */
void gto_fill_uniform_f64(double *dst, int64_t n, uint64_t seed, uint64_t field_id, double a, double b);

/* ---- horizontal diffusion: tests/regression/horizontal_diffusion.cpp:35-107 ---- */
void gto_hdiff_f64(double *out, const double *in, const double *coeff, int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk, int nthreads);
void gto_hdiff_f32(float *out, const float *in, const float *coeff, int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk, int nthreads);
/* FP32 storage, every load widened to FP64, result narrowed once on store */
void gto_hdiff_mixed(float *out, const float *in, const float *coeff, int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk, int nthreads);

/* ---- laplacian: tests/regression/laplacian.cpp:20-29 ---- */
void gto_lap_f64(double *out, const double *in, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);
void gto_lap_f32(float *out, const float *in, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);

/* ---- simple_hori_diff: tests/regression/simple_hori_diff.cpp:25-80 ----
 * crlato / crlatu (pointers at grid point (0,0,0)) have strides (ci, cj, ck): (0, 1, 0) for the j-only fields of
 * the reference test (selector<0,1,0>).  _mixed: FP32 storage, FP64 loads, lap and arithmetic, result narrowed
 * once on store. */
void gto_shd_f64(double *out, const double *in, const double *cro, const double *cru, ptrdiff_t ci, ptrdiff_t cj,
    ptrdiff_t ck, const double *coeff, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);
void gto_shd_f32(float *out, const float *in, const float *cro, const float *cru, ptrdiff_t ci, ptrdiff_t cj,
    ptrdiff_t ck, const float *coeff, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);
void gto_shd_mixed(float *out, const float *in, const float *cro, const float *cru, ptrdiff_t ci, ptrdiff_t cj,
    ptrdiff_t ck, const float *coeff, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);

/* ---- prepare_tracers: tests/regression/advection_pdbott_prepare_tracers.cpp:23-34 (data = rho * data_nnow) ---- */
void gto_prepare_tracers_f64(double *data, const double *data_nnow, const double *rho, int ni, int nj, int nk,
    ptrdiff_t sj, ptrdiff_t sk);
void gto_prepare_tracers_f32(float *data, const float *data_nnow, const float *rho, int ni, int nj, int nk,
    ptrdiff_t sj, ptrdiff_t sk);

/* ---- extended_4D: tests/regression/extended_4D.cpp:58-91 ----
 * result (I,J,K extra dims at strides rs), f (a,b,c extra dims at strides fs) and jac (q at stride jq) share the
 * spatial strides (1, sj, sk); phi / psi are the 2x2x2x2 global-parameter tables [((x*2+y)*2+z)*2+q]. */
void gto_extended_4d_f64(double *result, const ptrdiff_t rs[3], const double *f, const ptrdiff_t fs[3],
    const double *jac, ptrdiff_t jq, const double phi[16], const double psi[16], int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk);
void gto_extended_4d_f32(float *result, const ptrdiff_t rs[3], const float *f, const ptrdiff_t fs[3],
    const float *jac, ptrdiff_t jq, const double phi[16], const double psi[16], int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk);

/* ---- vertical advection: tests/regression/vertical_advection_dycore.cpp:32-149 ----
 * wcon may have its own strides (it is often allocated with nk+1 staggered levels).
 * nk >= 2 is required (first and last level are distinct intervals).            */
int gto_vadv_f64(double *utens_stage, const double *u_stage, const double *wcon, const double *u_pos,
    const double *utens, double dtr_stage, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk, ptrdiff_t wsj,
    ptrdiff_t wsk, int nthreads);
int gto_vadv_f32(float *utens_stage, const float *u_stage, const float *wcon, const float *u_pos,
    const float *utens, float dtr_stage, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk, ptrdiff_t wsj,
    ptrdiff_t wsk, int nthreads);
int gto_vadv_mixed(float *utens_stage, const float *u_stage, const float *wcon, const float *u_pos,
    const float *utens, float dtr_stage, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk, ptrdiff_t wsj,
    ptrdiff_t wsk, int nthreads);

/* ---- k-first traversals of the FP64 hdiff and vadv (the loop nest of stencil/cpu_kfirst.hpp:88-151: 8x8 (i, j)
 * blocks over OpenMP threads, k innermost inside each stage).  Same operations as gto_hdiff_f64 / gto_vadv_f64, so
 * the same bits; a second CPU baseline. */
void gto_hdiff_kfirst_f64(double *out, const double *in, const double *coeff, int ni, int nj, int nk, ptrdiff_t sj,
    ptrdiff_t sk, int nthreads);
int gto_vadv_kfirst_f64(double *utens_stage, const double *u_stage, const double *wcon, const double *u_pos,
    const double *utens, double dtr_stage, int ni, int nj, int nk, ptrdiff_t sj, ptrdiff_t sk, ptrdiff_t wsj,
    ptrdiff_t wsk, int nthreads);

/* ---- transform_layout: layout_transformation/cpu.hpp:20-55 -- dst[sum i_d dst_strides[d]] = src[sum i_d
 * src_strides[d]] over the box dims (ndims <= 16), strides in elements of elem_size bytes. */
void gto_transform_layout(int elem_size, void *dst, const void *src, int ndims, const int64_t *dims,
    const int64_t *dst_strides, const int64_t *src_strides);

/* ---- tridiagonal Thomas solve: tests/regression/tridiagonal.cpp:33-68 ----
 * inf/diag/sup/rhs/out share one layout; sup and rhs are overwritten like the reference's inout accessors. */
void gto_tridiagonal_f64(double *out, const double *inf, const double *diag, double *sup, double *rhs, int ni,
    int nj, int nk, ptrdiff_t sj, ptrdiff_t sk);

/* ---- boundary conditions: boundaries/apply.hpp:45-135 (26-direction loop), zero.hpp:29-55,
 * value.hpp:27-67, copy.hpp:26-47 ----
 * Here the field pointer is the STORAGE BASE (index (0,0,0)); halo[d] = {minus, plus, begin, end, total_length}
 * exactly as common/halo_descriptor.hpp:44-49.  A point belongs to direction (di,dj,dk) with d = -1 on
 * [begin-minus, begin-1], 0 on [begin, end], +1 on [end+1, end+plus] (halo_descriptor.hpp:72-120); points
 * outside those ranges are never touched.  Direction bit (di+1)*9+(dj+1)*3+(dk+1) of dir_mask must be set
 * (bit 13, the interior, is ignored).  kind: 0 zero, 1 value, 2 copy (dst <- src at the same point). */
void gto_bc_f64(int kind, double *dst, const double *src, double value, const int halo[3][5], ptrdiff_t sj,
    ptrdiff_t sk, uint32_t dir_mask);
void gto_bc_f32(int kind, float *dst, const float *src, float value, const int halo[3][5], ptrdiff_t sj,
    ptrdiff_t sk, uint32_t dir_mask);

/* ---- local halo fill (the single-rank effect of gcl halo pack + exchange + unpack with a periodic
 * process grid, gcl/low_level/proc_grids_3D.hpp:83-88, plus a zero-gradient boundary) ----
 * mode[d]: 0 none, 1 zero-gradient (clamp to [begin,end]), 2 periodic (wrap by end-begin+1).
 * A point is filled iff it is in the halo of at least one dimension and every dimension in whose halo it
 * lies has a mode != 0; it receives the value at the point whose coordinates are all mapped into the
 * interior, so corners are (wrap/clamp i, wrap/clamp j, wrap/clamp k) and the result is order-free. */
void gto_halo_fill_f64(double *f, const int halo[3][5], ptrdiff_t sj, ptrdiff_t sk, const int mode[3]);
void gto_halo_fill_f32(float *f, const int halo[3][5], ptrdiff_t sj, ptrdiff_t sk, const int mode[3]);

#ifdef __cplusplus
}
#endif
#endif
