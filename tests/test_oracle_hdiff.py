"""Pins the CPU oracle's horizontal diffusion and laplacian.

Checks, each chosen so that a plausible mistake (wrong offset, swapped i/j stride, wrong limiter sign,
dropped term, wrong summation order) fails at least one:
  * the reference's own known answer: horizontal_diffusion_repository's analytic input and pointwise closed
    form (tests/regression/horizontal_diffusion_repository.hpp:32-66), verified at the reference's tolerance
    (1e-14 double, 1e-6 float: tests/include/verifier.hpp:25-51) at the reference's test sizes
    (tests/include/test_environment.hpp:69-77);
  * constant and integer-linear inputs give out == in exactly (lap vanishes);
  * a unit impulse gives hand-derived values that exercise both limiter branches;
  * the mixed-precision variant equals the FP64 result on widened inputs rounded once;
  * the laplacian known answer (tests/regression/laplacian.cpp:31-39, constant -1 -> 0).
"""
import numpy as np
import pytest

from conftest import rel_err


def make(ni, nj, nk, h, dtype=np.float64, fill=0.0):
    return np.full((nk, nj + 2 * h, ni + 2 * h), fill, dtype=dtype)


def repo_in(ni_tot, nj_tot, nk, d1, d2):
    """horizontal_diffusion_repository::in over storage indices (i, j): values between 5 and 9."""
    PI = np.arctan(1.0) * 4.0
    i = np.arange(ni_tot)[None, None, :].astype(np.float64)
    j = np.arange(nj_tot)[None, :, None].astype(np.float64)
    x = (1.0 / d1) * i
    y = (1.0 / d2) * j
    v = 5.0 + 8 * (2.0 + np.cos(PI * (x + 1.5 * y)) + np.sin(2 * PI * (x + 1.5 * y))) / 4.0
    return np.broadcast_to(v, (nk, nj_tot, ni_tot)).copy()


def repo_out(inp, coeff, h, ni, nj):
    """horizontal_diffusion_repository::out (:48-66), evaluated pointwise over the interior."""
    def I(di, dj):
        return inp[:, h + dj - 2: h + dj - 2 + nj + 4, h + di - 2: h + di - 2 + ni + 4]

    # values on a window of the interior extended by 2: index (a, b) <-> storage (h-2+b, h-2+a)
    def at(a, da, db):
        return a[:, 2 + db: 2 + db + nj, 2 + da: 2 + da + ni]

    c = I(0, 0)

    def lap(da, db):
        return 4 * at(c, da, db) - (at(c, da + 1, db) + at(c, da, db + 1) + at(c, da - 1, db) + at(c, da, db - 1))

    def flx(da, db):
        res = lap(da + 1, db) - lap(da, db)
        return np.where(res * (at(c, da + 1, db) - at(c, da, db)) > 0, 0.0, res)

    def fly(da, db):
        res = lap(da, db + 1) - lap(da, db)
        return np.where(res * (at(c, da, db + 1) - at(c, da, db)) > 0, 0.0, res)

    cf = coeff[:, h:h + nj, h:h + ni]
    return at(c, 0, 0) - cf * (flx(0, 0) - flx(-1, 0) + fly(0, 0) - fly(0, -1))


def run_hdiff(oracle, inp, coeff, h, ni, nj, nk, mode="native", threads=1):
    out = np.zeros_like(inp)
    V = oracle.View
    oracle.hdiff(V(out, h, h), V(inp, h, h), V(coeff, h, h), ni, nj, nk, mode=mode, threads=threads)
    return out


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (23, 11, 43), (64, 64, 8)])
def test_hdiff_matches_reference_repository(oracle, ni, nj, nk):
    h = 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp = repo_in(d1, d2, nk, d1, d2)
    coeff = np.full_like(inp, 0.025)
    out = run_hdiff(oracle, inp, coeff, h, ni, nj, nk)
    ref = repo_out(inp, coeff, h, ni, nj)
    got = out[:, h:h + nj, h:h + ni]
    # verifier.hpp:47-51 with the double precision 1e-14
    err = np.abs(got - ref)
    assert np.all((err < 1e-14) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-14))
    # identical evaluation order: the oracle reproduces the closed form bit for bit
    assert np.array_equal(got, ref)


def test_hdiff_float_matches_reference_repository_at_float_tolerance(oracle):
    ni, nj, nk, h = 23, 11, 43, 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp = repo_in(d1, d2, nk, d1, d2)
    coeff = np.full_like(inp, 0.025)
    ref = repo_out(inp, coeff, h, ni, nj)
    out = run_hdiff(oracle, inp.astype(np.float32), coeff.astype(np.float32), h, ni, nj, nk)
    got = out[:, h:h + nj, h:h + ni].astype(np.float64)
    err = np.abs(got - ref)
    assert np.all((err < 1e-6) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-6))


def test_hdiff_constant_and_linear_inputs_are_fixed_points(oracle):
    ni, nj, nk, h = 17, 9, 3, 2
    rng = np.random.default_rng(0)
    coeff = rng.uniform(0, 0.25, size=(nk, nj + 4, ni + 4))
    const = np.full_like(coeff, 3.25)
    out = run_hdiff(oracle, const, coeff, h, ni, nj, nk)
    assert np.array_equal(out[:, h:-h, h:-h], const[:, h:-h, h:-h])
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj + 4), np.arange(ni + 4), indexing="ij")
    lin = (3.0 * i + 5.0 * j + 7.0 * k + 11.0)  # integers: every lap is exactly 0
    out = run_hdiff(oracle, lin, coeff, h, ni, nj, nk)
    assert np.array_equal(out[:, h:-h, h:-h], lin[:, h:-h, h:-h])


def test_hdiff_unit_impulse_known_answer(oracle):
    """Impulse 1 at the centre of a 5x5 interior, coeff c everywhere.  By hand: lap(centre) = 4 and -1 at the four
    neighbours; every flux touching the centre points down its own gradient and is zeroed by the limiter, so
    out(centre) = 1; at the east neighbour flx = 1, flx[i-1] = 0, fly = 1, fly[j-1] = -1, so out = -3c; the
    same by symmetry at the other three neighbours; diagonal neighbours get -(-1 + ... ) terms:
    at (c+1, c+1) flx = lap(c+2,c+1) - lap(c+1,c+1) = 0, flx[i-1] = lap(c+1,c+1) - lap(c,c+1) = 0 - (-1) = 1,
    fly = 0, fly[j-1] = lap(c+1,c+1) - lap(c+1,c) = 1, so out = -c*(0 - 1 + 0 - 1) = 2c."""
    ni = nj = 5
    nk, h, c = 1, 2, 0.125
    inp = make(ni, nj, nk, h)
    inp[0, h + 2, h + 2] = 1.0
    coeff = np.full_like(inp, c)
    out = run_hdiff(oracle, inp, coeff, h, ni, nj, nk)[0, h:-h, h:-h]
    assert out[2, 2] == 1.0
    for (dj, di) in [(0, 1), (0, -1), (1, 0), (-1, 0)]:
        assert out[2 + dj, 2 + di] == -3 * c
    for (dj, di) in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
        assert out[2 + dj, 2 + di] == 2 * c
    # two steps away on the axes: flx(c+1) = lap(c+2) - lap(c+1) = 1 with zero gradient -> kept
    assert out[2, 4] == -c * (0 - 1 + 0 - 0)


def test_hdiff_mixed_is_fp64_on_widened_inputs(oracle):
    ni, nj, nk, h = 31, 20, 5, 2
    rng = np.random.default_rng(1)
    inp = rng.uniform(-1, 1, size=(nk, nj + 4, ni + 4)).astype(np.float32)
    coeff = rng.uniform(0, 0.25, size=inp.shape).astype(np.float32)
    mixed = run_hdiff(oracle, inp, coeff, h, ni, nj, nk, mode="mixed")
    wide = run_hdiff(oracle, inp.astype(np.float64), coeff.astype(np.float64), h, ni, nj, nk)
    assert np.array_equal(mixed[:, h:-h, h:-h], wide[:, h:-h, h:-h].astype(np.float32))


def test_hdiff_threads_do_not_change_bits(oracle):
    ni, nj, nk, h = 70, 45, 6, 2
    rng = np.random.default_rng(2)
    inp = rng.uniform(-1, 1, size=(nk, nj + 4, ni + 4))
    coeff = rng.uniform(0, 0.25, size=inp.shape)
    a = run_hdiff(oracle, inp, coeff, h, ni, nj, nk, threads=1)
    b = run_hdiff(oracle, inp, coeff, h, ni, nj, nk, threads=4)
    assert np.array_equal(a, b)
    assert rel_err(a[:, h:-h, h:-h], repo_out(inp, coeff, h, ni, nj)) == 0.0


def test_laplacian_constant_known_answer(oracle):
    ni, nj, nk, h = 12, 33, 4, 1
    inp = np.full((nk, nj + 2, ni + 2), -1.0)
    out = np.full_like(inp, 99.0)
    oracle.laplacian(oracle.View(out, h, h), oracle.View(inp, h, h), ni, nj, nk)
    assert np.array_equal(out[:, 1:-1, 1:-1], np.zeros((nk, nj, ni)))
    assert np.all(out[:, 0, :] == 99.0)  # halo untouched


def test_laplacian_fp32_known_answers_and_rounding(oracle):
    """laplacian.cpp:31-39 known answer in FP32, the harmonic-input answer (lap of a linear field is 0 in exact
    arithmetic and in FP32 for small integers), and the FP32 result is the float rounding of each operation: it
    differs from the FP64 laplacian rounded once on random inputs (a kernel computing in FP64 would fail this)."""
    ni, nj, nk, h = 20, 9, 3, 1
    inp = np.full((nk, nj + 2, ni + 2), -1.0, dtype=np.float32)
    out = np.full_like(inp, 7.0)
    oracle.laplacian(oracle.View(out, h, h), oracle.View(inp, h, h), ni, nj, nk)
    assert np.array_equal(out[:, 1:-1, 1:-1], np.zeros((nk, nj, ni), np.float32))
    assert np.all(out[:, 0, :] == 7.0)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj + 2), np.arange(ni + 2), indexing="ij")
    lin = (3 * i - 2 * j + k).astype(np.float32)
    oracle.laplacian(oracle.View(out, h, h), oracle.View(lin, h, h), ni, nj, nk)
    assert np.array_equal(out[:, 1:-1, 1:-1], np.zeros((nk, nj, ni), np.float32))
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, size=(nk, nj + 2, ni + 2)).astype(np.float32)
    o32 = np.zeros_like(x)
    oracle.laplacian(oracle.View(o32, h, h), oracle.View(x, h, h), ni, nj, nk)
    x64 = x.astype(np.float64)
    o64 = np.zeros_like(x64)
    oracle.laplacian(oracle.View(o64, h, h), oracle.View(x64, h, h), ni, nj, nk)
    a, b = o32[:, 1:-1, 1:-1], o64[:, 1:-1, 1:-1]
    assert np.max(np.abs(a - b)) < 1e-5
    assert not np.array_equal(a, b.astype(np.float32))
    with pytest.raises(TypeError):
        oracle.laplacian(oracle.View(o32, h, h), oracle.View(x64, h, h), ni, nj, nk)
