"""Pins the CPU oracle's simple_hori_diff and prepare_tracers.

simple_hori_diff (tests/regression/simple_hori_diff.cpp:25-85), checks chosen so that a swapped crlato/crlatu, a
wrong sign, a dropped flux term or a j-offset mistake in the j-only fields fails at least one:
  * the reference's own known answer: horizontal_diffusion_repository's analytic input, j-only crlato/crlatu and
    pointwise closed form out_simple (tests/regression/horizontal_diffusion_repository.hpp:23-46, 68-79) at the
    reference's test sizes with halo 2 (tests/include/test_environment.hpp:69-77, 193-199) -- bit for bit in
    double (same evaluation order) and at the reference's float tolerance (tests/include/verifier.hpp:25-51);
  * the same closed form on random inputs and random j-only fields (every term non-trivial);
  * a constant input is a fixed point (every lap and flux vanishes exactly);
  * with crlato = crlatu = 1 the operator is the biharmonic in - coeff * L(L(in)), L the 5-point laplacian of
    the laplacian oracle (an independent composition, to rounding);
  * the mixed variant equals the double result on widened inputs rounded once.
prepare_tracers (tests/regression/advection_pdbott_prepare_tracers.cpp:36-57): the known answer 1.1 * i for
member i with data_nnow = i and rho = 1.1.
"""
import numpy as np
import pytest

from test_oracle_hdiff import repo_in


def repo_crlat(d2):
    """horizontal_diffusion_repository's j-only crlato / crlatu over storage j in [0, d2) (:23-27, :44-46)."""
    delta0 = (0.995156 - 0.994954) / (d2 - 1.0)
    delta1 = (0.995143 - 0.994924) / (d2 - 1.0)

    def crlat0(j):
        return 0.994954 + j * delta0

    def crlat1(j):
        return 0.994924 + j * delta1

    cro = np.array([crlat1(j) / crlat0(j) for j in range(d2)])
    cru = np.array([0.0 if j == 0 else crlat1(j - 1) / crlat0(j) for j in range(d2)])
    return cro, cru


def out_simple(inp, cro, cru, coeff, h, ni, nj):
    """horizontal_diffusion_repository::out_simple (:68-79) over the interior; inp/coeff [k, j, i] storage arrays,
    cro/cru over storage j.  Each expression is evaluated left to right like the C++ source."""
    def at(a, di, dj):
        return a[:, h + dj: h + dj + nj, h + di: h + di + ni]

    def cj(v, dj):
        if v.ndim == 3:  # a full metric field crlato(i, j, k) at the laplacian's point, like the closed form's
            return v[:, h + dj: h + dj + nj, h: h + ni]
        return v[h + dj: h + dj + nj][None, :, None]

    def cij(v, di, dj):
        return at(v, di, dj) if v.ndim == 3 else cj(v, dj)

    def lap(di, dj):
        return (((at(inp, di + 1, dj) + at(inp, di - 1, dj)) - 2 * at(inp, di, dj))
                + cij(cro, di, dj) * (at(inp, di, dj + 1) - at(inp, di, dj))) \
            + cij(cru, di, dj) * (at(inp, di, dj - 1) - at(inp, di, dj))

    fluxx = lap(1, 0) - lap(0, 0)
    fluxx_m = lap(0, 0) - lap(-1, 0)
    fluxy = cj(cro, 0) * (lap(0, 1) - lap(0, 0))
    fluxy_m = cj(cro, 0) * (lap(0, 0) - lap(0, -1))
    return at(inp, 0, 0) + ((fluxx_m - fluxx) + (fluxy_m - fluxy)) * at(coeff, 0, 0)


def run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk, mode="native"):
    out = np.zeros_like(inp)
    V = oracle.View
    oracle.simple_hori_diff(V(out, h, h), V(inp, h, h), cro, cru, V(coeff, h, h), ni, nj, nk, mode=mode)
    return out


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (23, 11, 43)])
def test_simple_hori_diff_matches_reference_repository(oracle, ni, nj, nk):
    h = 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp = repo_in(d1, d2, nk, d1, d2)
    coeff = np.full_like(inp, 0.025)
    cro, cru = repo_crlat(d2)
    got = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    ref = out_simple(inp, cro, cru, coeff, h, ni, nj)
    err = np.abs(got - ref)
    assert np.all((err < 1e-14) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-14))
    assert np.array_equal(got, ref)
    # the halo rim of out is not written
    out = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk)
    assert not out[:, :h, :].any() and not out[:, :, :h].any()


def test_simple_hori_diff_float_at_reference_tolerance(oracle):
    ni, nj, nk, h = 12, 33, 61, 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp = repo_in(d1, d2, nk, d1, d2)
    coeff = np.full_like(inp, 0.025)
    cro, cru = repo_crlat(d2)
    ref = out_simple(inp, cro, cru, coeff, h, ni, nj)
    f = np.float32
    got = run_shd(oracle, inp.astype(f), cro.astype(f), cru.astype(f), coeff.astype(f), h, ni, nj, nk)
    got = got[:, h:h + nj, h:h + ni].astype(np.float64)
    err = np.abs(got - ref)
    assert np.all((err < 1e-6) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-6))


def test_simple_hori_diff_random_closed_form(oracle):
    ni, nj, nk, h = 29, 17, 4, 2
    rng = np.random.default_rng(11)
    inp = rng.uniform(-1, 1, size=(nk, nj + 2 * h, ni + 2 * h))
    coeff = rng.uniform(0, 0.25, size=inp.shape)
    cro = rng.uniform(0.5, 1.5, size=nj + 2 * h)
    cru = rng.uniform(0.5, 1.5, size=nj + 2 * h)
    got = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    assert np.array_equal(got, out_simple(inp, cro, cru, coeff, h, ni, nj))
    # swapping the two j-only fields changes the answer (they enter asymmetrically)
    swapped = run_shd(oracle, inp, cru, cro, coeff, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    assert not np.allclose(got, swapped)


def test_simple_hori_diff_full_metric_fields(oracle):
    """The closed form evaluates crlato(ii, jj, k) as a 3-D function; full metric fields that vary in i and k follow
    it too, and a 3-D field holding the j-only values reproduces the j-only result bit for bit."""
    ni, nj, nk, h = 19, 14, 3, 2
    rng = np.random.default_rng(15)
    shape = (nk, nj + 2 * h, ni + 2 * h)
    inp = rng.uniform(-1, 1, size=shape)
    coeff = rng.uniform(0, 0.25, size=shape)
    cro3 = rng.uniform(0.5, 1.5, size=shape)
    cru3 = rng.uniform(0.5, 1.5, size=shape)
    V = oracle.View
    got = run_shd(oracle, inp, V(cro3, h, h), V(cru3, h, h), coeff, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    assert np.array_equal(got, out_simple(inp, cro3, cru3, coeff, h, ni, nj))
    cro, cru = rng.uniform(0.5, 1.5, size=(2, nj + 2 * h))
    b = [np.ascontiguousarray(np.broadcast_to(c[None, :, None], shape)) for c in (cro, cru)]
    full = run_shd(oracle, inp, V(b[0], h, h), V(b[1], h, h), coeff, h, ni, nj, nk)
    assert np.array_equal(full, run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk))


def test_simple_hori_diff_constant_is_fixed_point(oracle):
    ni, nj, nk, h = 9, 7, 3, 2
    rng = np.random.default_rng(12)
    inp = np.full((nk, nj + 2 * h, ni + 2 * h), -2.75)
    coeff = rng.uniform(0, 0.25, size=inp.shape)
    cro, cru = rng.uniform(0.5, 1.5, size=(2, nj + 2 * h))
    got = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk)
    assert np.array_equal(got[:, h:-h, h:-h], inp[:, h:-h, h:-h])


def test_simple_hori_diff_unit_metric_is_biharmonic(oracle):
    """crlato = crlatu = 1: lap = -L(in) and ((fluxx_m - fluxx) + (fluxy_m - fluxy)) = L(lap), so
    out = in - coeff * L(L(in)) with L the laplacian oracle (laplacian.cpp:20-29) applied twice."""
    ni, nj, nk, h = 21, 13, 3, 2
    rng = np.random.default_rng(13)
    inp = rng.uniform(-1, 1, size=(nk, nj + 2 * h, ni + 2 * h))
    coeff = rng.uniform(0, 0.25, size=inp.shape)
    ones = np.ones(nj + 2 * h)
    got = run_shd(oracle, inp, ones, ones, coeff, h, ni, nj, nk)[:, h:-h, h:-h]
    V = oracle.View
    l1 = np.zeros_like(inp)
    oracle.laplacian(V(l1, 1, 1), V(inp, 1, 1), ni + 2, nj + 2, nk)  # L(in) over the interior plus a 1-point rim
    l2 = np.zeros_like(inp)
    oracle.laplacian(V(l2, h, h), V(l1, h, h), ni, nj, nk)
    want = inp[:, h:-h, h:-h] - coeff[:, h:-h, h:-h] * l2[:, h:-h, h:-h]
    assert np.max(np.abs(got - want)) < 1e-13
    assert np.max(np.abs(got - inp[:, h:-h, h:-h])) > 1e-2  # the operator is not the identity here


def test_simple_hori_diff_mixed_is_double_on_widened_inputs(oracle):
    ni, nj, nk, h = 31, 20, 5, 2
    rng = np.random.default_rng(14)
    f = np.float32
    inp = rng.uniform(-1, 1, size=(nk, nj + 2 * h, ni + 2 * h)).astype(f)
    coeff = rng.uniform(0, 0.25, size=inp.shape).astype(f)
    cro, cru = rng.uniform(0.5, 1.5, size=(2, nj + 2 * h)).astype(f)
    mixed = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk, mode="mixed")
    wide = run_shd(oracle, inp.astype(np.float64), cro.astype(np.float64), cru.astype(np.float64),
                   coeff.astype(np.float64), h, ni, nj, nk)
    assert np.array_equal(mixed[:, h:-h, h:-h], wide[:, h:-h, h:-h].astype(f))
    native = run_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk)
    assert not np.array_equal(native, mixed)  # the float instantiation rounds every operation


@pytest.mark.parametrize("dtype,tol", [(np.float64, 0.0), (np.float32, 1e-6)])
def test_prepare_tracers_known_answer(oracle, dtype, tol):
    """advection_pdbott_prepare_tracers.cpp:39-56: member i holds i, rho = 1.1, expected 1.1 * i."""
    ni, nj, nk = 12, 33, 61
    rho = np.full((nk, nj, ni), 1.1, dtype=dtype)
    V = oracle.View
    for m in range(11):
        src = np.full((nk, nj, ni), m, dtype=dtype)
        dst = np.zeros_like(src)
        oracle.prepare_tracers(V(dst, 0, 0), V(src, 0, 0), V(rho, 0, 0), ni, nj, nk)
        want = 1.1 * m
        err = np.abs(dst.astype(np.float64) - want)
        assert np.all((err <= tol) | (err < abs(want) * tol)), m
        if tol == 0.0:
            assert np.all(dst == want)


# ------------------------------------------------------------------------------------------------ extended_4D
def ext4d_closed_form(f, jac, phi, psi):
    """The projection as a tensor contraction (numpy einsum, any summation order): result[I,J,K] =
    1/8 sum_q phi[I,J,K,q] jac[q] sum_abc psi[a,b,c,q] f[a,b,c] per grid point; f / result [c,b,a,k,j,i]."""
    fa = np.transpose(f, (2, 1, 0, 3, 4, 5)).astype(np.float64)  # [a,b,c,...]
    s = np.einsum("abcq,abcxyz->qxyz", psi, fa)
    r = np.einsum("IJKq,qxyz,qxyz->IJKxyz", phi, jac.astype(np.float64), s) / 8
    return np.transpose(r, (2, 1, 0, 3, 4, 5))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_extended_4d_reference_known_answer(oracle, dtype, tol):
    """extended_4D.cpp:104-134: phi = 10, psi = 11, f = 1.3, jac(q) = 1 + q; expected (jac0 + jac1) phi psi f at every
    point and every (I, J, K), at the reference's tolerance (verifier.hpp:25-51)."""
    ni, nj, nk = 12, 33, 61
    f_ = dtype(1.3)
    f = np.full((2, 2, 2, nk, nj, ni), f_, dtype=dtype)
    jac = np.stack([np.full((nk, nj, ni), 1 + q, dtype=dtype) for q in range(2)])
    res = np.zeros_like(f)
    oracle.extended_4d(res, f, jac, np.full(16, 10.0), np.full(16, 11.0), ni, nj, nk)
    want = (1.0 + 2.0) * 10 * 11 * float(f_)
    err = np.abs(res.astype(np.float64) - want)
    assert np.all((err < tol) | (err < want * tol))


def test_extended_4d_random_tables_match_contraction(oracle):
    """Random phi / psi tables and random f / jac: every index of the contraction is exercised, so a transposed
    table index, a wrong extra-dimension stride or a dropped quadrature point fails."""
    ni, nj, nk = 9, 7, 5
    rng = np.random.default_rng(21)
    f = rng.uniform(-1, 1, size=(2, 2, 2, nk, nj, ni))
    jac = rng.uniform(0.5, 1.5, size=(2, nk, nj, ni))
    phi = rng.uniform(-1, 1, size=(2, 2, 2, 2))
    psi = rng.uniform(-1, 1, size=(2, 2, 2, 2))
    res = np.zeros_like(f)
    oracle.extended_4d(res, f, jac, phi, psi, ni, nj, nk)
    want = ext4d_closed_form(f, jac, phi, psi)
    assert np.max(np.abs(res - want)) < 1e-14
    # the tables enter with distinct roles
    res2 = np.zeros_like(f)
    oracle.extended_4d(res2, f, jac, psi, phi, ni, nj, nk)
    assert np.max(np.abs(res2 - want)) > 1e-3
    # float: sums held in float, products in double
    res32 = np.zeros(f.shape, np.float32)
    oracle.extended_4d(res32, f.astype(np.float32), jac.astype(np.float32), phi, psi, ni, nj, nk)
    assert np.max(np.abs(res32 - ext4d_closed_form(f.astype(np.float32), jac.astype(np.float32), phi, psi))) < 1e-5
