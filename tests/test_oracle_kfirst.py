"""The second CPU baseline: the cpu_kfirst traversal (stencil/cpu_kfirst.hpp:88-151: 8x8 (i, j) blocks, k innermost
in each stage) of the FP64 hdiff and vadv oracles.  It evaluates the same functor operations in the same order as
the i-first oracle, so it must give the same bits, on block-remainder shapes and with several threads; the i-first
oracle is pinned by the reference closed forms (test_oracle_hdiff.py, test_oracle_vadv.py)."""
import numpy as np
import pytest

from test_oracle_hdiff import repo_in, repo_out


@pytest.mark.parametrize("ni,nj,nk,threads", [(70, 45, 6, 1), (13, 9, 80, 4), (8, 8, 2, 2), (1, 17, 3, 3)])
def test_hdiff_kfirst_equals_ifirst(oracle, ni, nj, nk, threads):
    h = 2
    rng = np.random.default_rng(ni + nj + nk)
    inp = rng.uniform(-1, 1, size=(nk, nj + 2 * h, ni + 2 * h))
    coeff = rng.uniform(0, 0.25, size=inp.shape)
    V = oracle.View
    a, b = np.zeros_like(inp), np.zeros_like(inp)
    oracle.hdiff(V(a, h, h), V(inp, h, h), V(coeff, h, h), ni, nj, nk, threads=threads)
    oracle.hdiff(V(b, h, h), V(inp, h, h), V(coeff, h, h), ni, nj, nk, mode="kfirst", threads=threads)
    assert np.array_equal(a, b)


def test_hdiff_kfirst_reference_closed_form(oracle):
    ni, nj, nk, h = 23, 11, 43, 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp = repo_in(d1, d2, nk, d1, d2)
    coeff = np.full_like(inp, 0.025)
    out = np.zeros_like(inp)
    V = oracle.View
    oracle.hdiff(V(out, h, h), V(inp, h, h), V(coeff, h, h), ni, nj, nk, mode="kfirst", threads=2)
    assert np.array_equal(out[:, h:h + nj, h:h + ni], repo_out(inp, coeff, h, ni, nj))


@pytest.mark.parametrize("ni,nj,nk,threads", [(70, 45, 6, 1), (13, 9, 80, 4), (9, 8, 2, 2), (256, 8, 80, 8)])
def test_vadv_kfirst_equals_ifirst(oracle, ni, nj, nk, threads):
    """Seeded inputs of the workload's distributions (u* ~ U(-1,1), wcon ~ U(-0.5,0.5): ill-conditioned columns)."""
    h = 3
    rng = np.random.default_rng(7 * ni + nk)
    shp = (nk, nj + 2 * h, ni + 2 * h)
    us, up, ut, uts0 = (rng.uniform(-1, 1, size=shp) for _ in range(4))
    wc = rng.uniform(-0.5, 0.5, size=(nk + 1, nj + 2 * h, ni + 2 * h))
    V = oracle.View
    a, b = uts0.copy(), uts0.copy()
    oracle.vadv(V(a, h, h), V(us, h, h), V(wc, h, h), V(up, h, h), V(ut, h, h), 0.15, ni, nj, nk, threads=threads)
    oracle.vadv(V(b, h, h), V(us, h, h), V(wc, h, h), V(up, h, h), V(ut, h, h), 0.15, ni, nj, nk, mode="kfirst",
                threads=threads)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, uts0)
