"""Pins the CPU oracle's vertical advection and tridiagonal solve.

  * the reference's own known answer: vertical_advection_repository's closed-form Thomas solve
    (tests/regression/vertical_advection_repository.hpp:95-151), applied to random fields drawn at the
    repository's scales (:153-174), at the reference verifier's double tolerance 1e-14
    (tests/include/verifier.hpp:25-51) and test sizes (test_environment.hpp:69-77);
  * a textbook check independent of the recurrence: the tridiagonal system a x[k-1] + b x[k] + c x[k+1] = d
    assembled from the inputs and solved by dense LU (numpy.linalg.solve), on well-conditioned columns;
  * wcon = 0 reduces the solve to a diagonal one with a closed form;
  * the reference tridiagonal known answer (tests/regression/tridiagonal.cpp:81-97: solution == 1);
  * the mixed-precision variant equals the FP64 result on widened inputs rounded once.
"""
import numpy as np
import pytest

BET_M = BET_P = 0.5


def fields(ni, nj, nk, h, dist, seed=0, dtype=np.float64):
    """u_stage, wcon (nk+1 levels), u_pos, utens, utens_stage as [k, j, i] arrays with an ij halo h."""
    rng = np.random.default_rng(seed)
    shp = (nk, nj + 2 * h, ni + 2 * h)
    wshp = (nk + 1, nj + 2 * h, ni + 2 * h)
    if dist == "workload":
        us, up, ut, uts = (rng.uniform(-1, 1, shp) for _ in range(4))
        wc = rng.uniform(-0.5, 0.5, wshp)
    else:  # reference-like scales (vertical_advection_repository.hpp:153-174): well conditioned
        us, up = rng.uniform(5, 9, shp), rng.uniform(5, 9, shp)
        ut = rng.uniform(-3e-6, 3e-6, shp)
        uts = rng.uniform(7, 10, shp)
        wc = rng.uniform(-2e-4, 2e-4, wshp)
    return [np.ascontiguousarray(a, dtype=dtype) for a in (uts, us, wc, up, ut)]


def run(oracle, uts, us, wc, up, ut, dtr, h, ni, nj, nk, mode="native", threads=1):
    out = uts.copy()
    V = oracle.View
    oracle.vadv(V(out, h, h), V(us, h, h), V(wc, h, h), V(up, h, h), V(ut, h, h), dtr, ni, nj, nk, mode=mode,
                threads=threads)
    return out


def repository_solution(uts, us, wc, up, ut, dtr, h, ni, nj, nk):
    """vertical_advection_repository::utens_stage_out (:95-151), column by column, vectorised over (j, i)."""
    J, I = slice(h, h + nj), slice(h, h + ni)
    Ip = slice(h + 1, h + 1 + ni)
    c = np.zeros((nk, nj, ni))
    d = np.zeros((nk, nj, ni))
    res = np.zeros((nk, nj, ni))
    k = 0
    gcv = .25 * (wc[k + 1, J, Ip] + wc[k + 1, J, I])
    cs = gcv * BET_M
    c[k] = gcv * BET_P
    b = dtr - c[k]
    corr = -cs * (us[k + 1, J, I] - us[k, J, I])
    d[k] = dtr * up[k, J, I] + ut[k, J, I] + uts[k, J, I] + corr
    c[k] /= b
    d[k] /= b
    for k in range(1, nk - 1):
        gav = -.25 * (wc[k, J, Ip] + wc[k, J, I])
        gcv = .25 * (wc[k + 1, J, Ip] + wc[k + 1, J, I])
        as_ = gav * BET_M
        cs = gcv * BET_M
        a = gav * BET_P
        c[k] = gcv * BET_P
        bcol = dtr - a - c[k]
        corr = -as_ * (us[k - 1, J, I] - us[k, J, I]) - cs * (us[k + 1, J, I] - us[k, J, I])
        d[k] = dtr * up[k, J, I] + ut[k, J, I] + uts[k, J, I] + corr
        div = 1 / (bcol - (c[k - 1] * a))
        c[k] *= div
        d[k] = (d[k] - d[k - 1] * a) * div
    k = nk - 1
    gav = -.25 * (wc[k, J, Ip] + wc[k, J, I])
    as_ = gav * BET_M
    a = gav * BET_P
    b = dtr - a
    corr = -as_ * (us[k - 1, J, I] - us[k, J, I])
    d[k] = dtr * up[k, J, I] + ut[k, J, I] + uts[k, J, I] + corr
    d[k] = (d[k] - d[k - 1] * a) / (b - (c[k - 1] * a))
    data = d[k]
    res[k] = dtr * (data - up[k, J, I])
    for k in range(nk - 2, -1, -1):
        data = d[k] - c[k] * data
        res[k] = dtr * (data - up[k, J, I])
    return res


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (23, 11, 43), (32, 24, 80)])
def test_vadv_matches_reference_repository(oracle, ni, nj, nk):
    h = 3
    uts, us, wc, up, ut = fields(ni, nj, nk, h, "reference", seed=ni)
    dtr = 3.0 / 20.0
    got = run(oracle, uts, us, wc, up, ut, dtr, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    ref = repository_solution(uts, us, wc, up, ut, dtr, h, ni, nj, nk)
    err = np.abs(got - ref)
    assert np.all((err < 1e-14) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-14))


def assemble_and_solve(uts, us, wc, up, ut, dtr, h, i, j, nk):
    """Dense tridiagonal system of one column from the stencil's definitions (vertical_advection_dycore.cpp:50-96):
    row k: a_k x_{k-1} + b_k x_k + c_k x_{k+1} = d_k with a_k = gav*BET_P, c_k = gcv*BET_P, b_k = dtr - a_k - c_k,
    d_k = dtr*u_pos + utens + utens_stage + correction; first/last rows drop a/c.  Returns dtr*(x - u_pos)."""
    ii, jj = h + i, h + j
    A = np.zeros((nk, nk))
    rhs = np.zeros(nk)
    for k in range(nk):
        gav = -.25 * (wc[k, jj, ii + 1] + wc[k, jj, ii]) if k > 0 else 0.0
        gcv = .25 * (wc[k + 1, jj, ii + 1] + wc[k + 1, jj, ii]) if k < nk - 1 else 0.0
        a, c = gav * BET_P, gcv * BET_P
        corr = 0.0
        if k > 0:
            corr += -(gav * BET_M) * (us[k - 1, jj, ii] - us[k, jj, ii])
        if k < nk - 1:
            corr += -(gcv * BET_M) * (us[k + 1, jj, ii] - us[k, jj, ii])
        A[k, k] = dtr - a - c
        if k > 0:
            A[k, k - 1] = a
        if k < nk - 1:
            A[k, k + 1] = c
        rhs[k] = dtr * up[k, jj, ii] + ut[k, jj, ii] + uts[k, jj, ii] + corr
    x = np.linalg.solve(A, rhs)
    return dtr * (x - up[:, jj, ii]), np.linalg.cond(A)


def functor_order_solution(uts, us, wc, up, ut, dtr, h, ni, nj, nk):
    """A second, independent transcription of u_forward_function / u_backward_function
    (vertical_advection_dycore.cpp:50-121) keeping the functors' exact operation order: b = (dtr - a) - c,
    divided = 1 / (b - ccol[k-1]*a), ccol = c*divided, dcol = (d - dcol[k-1]*a)*divided.  The repository above
    rounds differently (c /= b, d /= b); this one must agree with the oracle bit for bit even where the system
    is ill-conditioned, which pins the operation order as written in the reference source."""
    J, I = slice(h, h + nj), slice(h, h + ni)
    Ip = slice(h + 1, h + 1 + ni)
    cc = np.zeros((nk, nj, ni))
    dc = np.zeros((nk, nj, ni))
    q = 0.25

    def cd(k):
        return ((dtr * up[k, J, I] + ut[k, J, I]) + uts[k, J, I])

    gcv = q * (wc[1, J, Ip] + wc[1, J, I])
    cs, c = gcv * BET_M, gcv * BET_P
    b = dtr - c
    d = cd(0) + (-cs) * (us[1, J, I] - us[0, J, I])
    divided = 1.0 / b
    cc[0], dc[0] = c * divided, d * divided
    for k in range(1, nk):
        gav = -q * (wc[k, J, Ip] + wc[k, J, I])
        as_, a = gav * BET_M, gav * BET_P
        if k < nk - 1:
            gcv = q * (wc[k + 1, J, Ip] + wc[k + 1, J, I])
            cs, c = gcv * BET_M, gcv * BET_P
            b = (dtr - a) - c
            corr = (-as_) * (us[k - 1, J, I] - us[k, J, I]) - cs * (us[k + 1, J, I] - us[k, J, I])
        else:
            b = dtr - a
            corr = (-as_) * (us[k - 1, J, I] - us[k, J, I])
        d = cd(k) + corr
        divided = 1.0 / (b - cc[k - 1] * a)
        if k < nk - 1:
            cc[k] = c * divided
        dc[k] = (d - dc[k - 1] * a) * divided
    res = np.zeros((nk, nj, ni))
    data = dc[nk - 1]
    res[nk - 1] = dtr * (dc[nk - 1] - up[nk - 1, J, I])
    for k in range(nk - 2, -1, -1):
        data = dc[k] - cc[k] * data
        res[k] = dtr * (data - up[k, J, I])
    return res


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (64, 16, 80)])
def test_vadv_operation_order_at_workload_distributions(oracle, ni, nj, nk):
    h = 3
    uts, us, wc, up, ut = fields(ni, nj, nk, h, "workload", seed=nk)
    got = run(oracle, uts, us, wc, up, ut, 0.15, h, ni, nj, nk)[:, h:h + nj, h:h + ni]
    ref = functor_order_solution(uts, us, wc, up, ut, 0.15, h, ni, nj, nk)
    assert np.array_equal(got, ref)
    # the repository's differently rounded closed form drifts measurably at these distributions: the parity
    # bar (1e-12 relative) is only reachable with the functors' own operation order
    rep = repository_solution(uts, us, wc, up, ut, 0.15, h, ni, nj, nk)
    assert np.max(np.abs(rep - ref)) > 0.0


def test_vadv_matches_dense_lu(oracle):
    ni, nj, nk, h = 7, 5, 24, 3
    uts, us, wc, up, ut = fields(ni, nj, nk, h, "workload", seed=11)
    wc *= 0.05  # keep every column well conditioned (cond < 1e3) so LU and Thomas agree to ~1e-12
    dtr = 0.15
    got = run(oracle, uts, us, wc, up, ut, dtr, h, ni, nj, nk)
    for j in range(nj):
        for i in range(ni):
            ref, cond = assemble_and_solve(uts, us, wc, up, ut, dtr, h, i, j, nk)
            assert cond < 1e3
            col = got[:, h + j, h + i]
            assert np.max(np.abs(col - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))) * cond


def test_vadv_zero_wind_closed_form(oracle):
    ni, nj, nk, h = 9, 4, 10, 3
    uts, us, wc, up, ut = fields(ni, nj, nk, h, "workload", seed=5)
    wc[:] = 0.0
    dtr = 0.15
    got = run(oracle, uts, us, wc, up, ut, dtr, h, ni, nj, nk)[:, h:-h, h:-h]
    s = (slice(None), slice(h, -h), slice(h, -h))
    d = ((dtr * up[s] + ut[s]) + uts[s]) + (-0.0) * (us[s] - us[s])  # correction vanishes
    exp = dtr * ((d * (1.0 / dtr)) - up[s])
    assert np.array_equal(got, exp)


def test_tridiagonal_known_answer(oracle):
    """tridiagonal.cpp:81-97: inf -1, diag 3, sup 1, rhs 4 / 3 / 2 -> solution 1 (nk = 6)."""
    for ni, nj in [(12, 33), (23, 11)]:
        nk = 6
        shp = (nk, nj, ni)
        inf = np.full(shp, -1.0)
        diag = np.full(shp, 3.0)
        sup = np.full(shp, 1.0)
        rhs = np.full(shp, 3.0)
        rhs[0] = 4.0
        rhs[5] = 2.0
        out = np.zeros(shp)
        V = oracle.View
        oracle.tridiagonal(V(out, 0, 0), V(inf, 0, 0), V(diag, 0, 0), V(sup, 0, 0), V(rhs, 0, 0), ni, nj, nk)
        err = np.abs(out - 1.0)
        assert np.all((err < 1e-14) | (err < 1e-14 * np.abs(out)))


def test_vadv_mixed_is_fp64_on_widened_inputs(oracle):
    ni, nj, nk, h = 40, 6, 30, 3
    f32 = fields(ni, nj, nk, h, "workload", seed=9, dtype=np.float32)
    dtr32 = np.float32(0.15)
    mixed = run(oracle, *f32, dtr32, h, ni, nj, nk, mode="mixed")
    wide = run(oracle, *[a.astype(np.float64) for a in f32], float(dtr32), h, ni, nj, nk)
    assert np.array_equal(mixed[:, h:-h, h:-h], wide[:, h:-h, h:-h].astype(np.float32))


def test_vadv_rejects_single_level(oracle):
    uts, us, wc, up, ut = fields(4, 4, 2, 3, "workload")
    with pytest.raises(ValueError):
        run(oracle, uts, us, wc, up, ut, 0.15, 3, 4, 4, 1)


def test_vadv_threads_do_not_change_bits(oracle):
    ni, nj, nk, h = 50, 21, 17, 3
    f = fields(ni, nj, nk, h, "workload", seed=3)
    a = run(oracle, *f, 0.15, h, ni, nj, nk, threads=1)
    b = run(oracle, *f, 0.15, h, ni, nj, nk, threads=5)
    assert np.array_equal(a, b)


def test_vadv_infinite_top_level_matches_fixture(oracle):
    """Infinite / NaN u_stage at the top level and infinite winds on the unread staggered level nk: the oracle
    equals the fixture written by an independent pure-Python transcription of the reference functors
    (tests/golden/make_vadv_inf_top.py), NaN and infinity positions included."""
    import os

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "vadv_inf_top.npz"))
    ni, nj, nk, h = (int(x) for x in z["dims"])
    fl = [np.ascontiguousarray(z[n]) for n in ("utens_stage", "u_stage", "wcon", "u_pos", "utens")]
    with np.errstate(invalid="ignore", over="ignore"):
        got = run(oracle, *fl, float(z["dtr"]), h, ni, nj, nk)
    exp = z["expected"]
    sl = (slice(None), slice(h, h + nj), slice(h, h + ni))
    assert np.isnan(exp[sl]).any() and np.isinf(exp[sl]).any()
    np.testing.assert_array_equal(got[sl], exp[sl])
