"""Pins the CPU oracle's boundary conditions and halo fill, and the input generator.

  * zero / value / copy boundaries (boundaries/zero.hpp, value.hpp, copy.hpp applied over the regions of
    boundaries/apply.hpp:45-135): per-region closed forms built independently from coordinate masks, including a
    direction predicate;
  * the reference boundary test's region bookkeeping (tests/regression/boundary_conditions.cpp:67-143: the
    interior is untouched, every halo point is written exactly once);
  * halo fill: periodic and zero-gradient as numpy index maps (np.take with mode='wrap' / clip);
  * SplitMix64: the published first output for seed 0 (0xE220A8397B1DCDAF) and the numpy twin.
"""
import numpy as np
import pytest

from gridtools_amd import inputs


def halos3(ni, nj, nk, hi, hj, hk):
    def hd(h, n):
        return (h, h, h, h + n - 1, n + 2 * h)
    return [hd(hi, ni), hd(hj, nj), hd(hk, nk)]


def dir_of(x, h):
    m, p, b, e, t = h
    d = np.full(x.shape, 2)
    d[(x >= b) & (x <= e)] = 0
    d[(x >= b - m) & (x < b)] = -1
    d[(x > e) & (x <= e + p)] = 1
    return d


def region_masks(shape_kji, halos):
    nk, nj, ni = shape_kji
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    return dir_of(i, halos[0]), dir_of(j, halos[1]), dir_of(k, halos[2])


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bc_kinds_closed_form(oracle, kind):
    ni, nj, nk = 9, 7, 5
    hal = halos3(ni, nj, nk, 3, 2, 1)
    shape = (nk + 2, nj + 4, ni + 6)
    rng = np.random.default_rng(kind)
    dst = rng.uniform(-1, 1, shape)
    src = rng.uniform(-1, 1, shape)
    before = dst.copy()
    # predicate: every direction except those pointing to +j (an arbitrary but asymmetric choice)
    mask = 0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                if (di, dj, dk) != (0, 0, 0) and dj != 1:
                    mask |= 1 << ((di + 1) * 9 + (dj + 1) * 3 + (dk + 1))
    oracle.bc(kind, dst, src, 2.5, hal, mask)
    di, dj, dk = region_masks(shape, hal)
    touched = (di != 2) & (dj != 2) & (dk != 2) & ~((di == 0) & (dj == 0) & (dk == 0)) & (dj != 1)
    exp = before.copy()
    exp[touched] = {0: 0.0, 1: 2.5, 2: None}[kind] if kind != 2 else src[touched]
    assert np.array_equal(dst, exp)
    assert touched.sum() > 0 and (~touched).sum() > 0


def test_bc_writes_every_halo_point_once_and_keeps_interior(oracle):
    """boundary_conditions.cpp:67-143 style bookkeeping: a value boundary over all 26 directions touches exactly the
    halo (minus/plus ranges) and never the interior."""
    ni, nj, nk = 6, 5, 4
    hal = halos3(ni, nj, nk, 3, 3, 0)
    shape = (nk, nj + 6, ni + 6)
    dst = np.zeros(shape)
    oracle.bc(1, dst, None, 1.0, hal, 0x7FFFFFF)
    interior = np.zeros(shape, dtype=bool)
    interior[:, 3:-3, 3:-3] = True
    assert np.all(dst[interior] == 0.0)
    assert np.all(dst[~interior] == 1.0)


def test_halo_fill_periodic_i_zero_gradient_j(oracle):
    ni, nj, nk, h = 11, 8, 3, 3
    hal = halos3(ni, nj, nk, h, h, 0)
    shape = (nk, nj + 2 * h, ni + 2 * h)
    f = np.random.default_rng(4).uniform(-1, 1, shape)
    orig = f.copy()
    oracle.halo_fill(f, hal, (2, 1, 0))
    ii = np.arange(ni + 2 * h)
    jj = np.arange(nj + 2 * h)
    src_i = h + (ii - h) % ni                   # wrap into [h, h+ni)
    src_j = np.clip(jj, h, h + nj - 1)          # clamp into [h, h+nj)
    exp = orig[:, src_j][:, :, src_i]
    assert np.array_equal(f, exp)


def test_halo_fill_none_leaves_dimension(oracle):
    ni, nj, nk, h = 6, 6, 2, 2
    hal = halos3(ni, nj, nk, h, h, 0)
    shape = (nk, nj + 2 * h, ni + 2 * h)
    f = np.random.default_rng(8).uniform(-1, 1, shape)
    orig = f.copy()
    oracle.halo_fill(f, hal, (2, 0, 0))
    # j halo rows (including their i-halo corners) untouched; i halo of interior rows wrapped
    assert np.array_equal(f[:, :h], orig[:, :h])
    assert np.array_equal(f[:, -h:], orig[:, -h:])
    assert np.array_equal(f[:, h:-h, :h], orig[:, h:-h, ni:ni + h])


def test_splitmix64_published_first_output():
    # SplitMix64 seeded with 0: first output 0xE220A8397B1DCDAF, second 0x6E789E6AA1B965F4
    # (Steele, Lea, Flood 2014; the values Java's SplittableRandom and Vigna's reference code produce).
    z = inputs.GAMMA * np.arange(1, 3, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * inputs.M1
        z = (z ^ (z >> np.uint64(27))) * inputs.M2
        z = z ^ (z >> np.uint64(31))
    assert int(z[0]) == 0xE220A8397B1DCDAF
    assert int(z[1]) == 0x6E789E6AA1B965F4


def test_uniform_numpy_matches_oracle_and_range(oracle):
    lengths = (13, 7, 5)
    n = int(np.prod(lengths))
    a = inputs.uniform(lengths, seed=42, field_id=3, a=-0.5, b=0.5)
    b = oracle.fill_uniform(n, 42, 3, -0.5, 0.5).reshape(5, 7, 13).transpose(2, 1, 0)
    assert np.array_equal(a, b)
    assert a.min() >= -0.5 and a.max() < 0.5
    # different field ids give different streams
    c = inputs.uniform(lengths, seed=42, field_id=4, a=-0.5, b=0.5)
    assert not np.array_equal(a, c)
