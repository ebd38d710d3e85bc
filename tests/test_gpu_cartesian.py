"""GPU parity of the further cartesian regression stencils against the CPU oracle, through the C ABI.

* simple_hori_diff (tests/regression/simple_hori_diff.cpp:25-85): the reference's own case (repository input,
  j-only crlato / crlatu built with selector(0, 1, 0), closed-form expectation; halo 2 at the typed-suite sizes,
  tests/include/test_environment.hpp:69-77) and seeded random inputs at block-remainder shapes and at
  256x256x80, for FP64, native FP32 and the precision form (FP32 storage, FP64 compute), with every
  rows-per-wave variant and with full 3-D metric fields.  FP64 and native FP32 must match the oracle bit for bit
  (same operation order, IEEE operations, no contraction); the precision form must equal the oracle's mixed
  variant bit for bit.
* expandable_run of prepare_tracers (advection_pdbott_prepare_tracers.cpp:36-57): the known answer 1.1 * i over
  11 members, seeded members against the oracle, and more members than one launch holds.
* copy_stencil_tuple (copy_stencil_tuple.cpp:116-260): the unrolled five-field copy (expected i + j + k + t), an
  array-of-structs storage, an extra data dimension, a masked (j-only) input and more pairs than one launch holds.
* extended_4D (extended_4D.cpp:104-134): the reference's known answer (jac0 + jac1) phi psi f over 6-D stores, and
  random 2x2x2x2 tables against the oracle bit for bit.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 42
FP64_TOL = 1e-12


def grid_of(gt, ni, nj, nk, h):
    return gt.make_grid(gt.halo_descriptor.symmetric(h, ni + 2 * h), gt.halo_descriptor.symmetric(h, nj + 2 * h), nk)


def shd_case(gt, ni, nj, nk, dtype, h=2, metric="j", seed=SEED):
    b = gt.builder().type(dtype).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    inp = b.uniform(seed, 21, -1.0, 1.0).build()
    coeff = b.uniform(seed, 22, 0.0, 0.25).build()
    out = b.value(0).build()
    if metric == "j":
        mb = b.selector(0, 1, 0)
    else:
        mb = b
    cro = mb.uniform(seed, 23, 0.5, 1.5).build()
    cru = mb.uniform(seed, 24, 0.5, 1.5).build()
    return inp, cro, cru, coeff, out, grid_of(gt, ni, nj, nk, h)


def oracle_shd(oracle, inp, cro, cru, coeff, h, ni, nj, nk, mode):
    a_in, a_co = inp.padded_numpy(), coeff.padded_numpy()
    if cro.masked:
        # j-only store: [1, nj_storage, 1] on the device -> a 1-D array over storage j
        m = [c.padded_numpy()[0, :, 0].copy() for c in (cro, cru)]
    else:
        m = [oracle.View(c.padded_numpy(), h, h) for c in (cro, cru)]
    res = np.zeros_like(a_in)
    V = oracle.View
    oracle.simple_hori_diff(V(res, h, h), V(a_in, h, h), m[0], m[1], V(a_co, h, h), ni, nj, nk, mode=mode)
    return res[:nk, h:h + nj, h:h + ni]


SHD_SIZES = [(12, 33, 61), (23, 11, 43), (64, 64, 8), (1, 1, 2), (63, 17, 3), (130, 37, 5), (256, 256, 80)]


@pytest.mark.parametrize("ni,nj,nk", SHD_SIZES)
@pytest.mark.parametrize("dtype,precision,mode", [
    (np.float64, "native", "native"), (np.float32, "native", "native"), (np.float32, "widen", "mixed")])
def test_simple_hori_diff_parity(gpu, oracle, ni, nj, nk, dtype, precision, mode):
    from gridtools_amd import stencils

    print(f"[case] simple_hori_diff {np.dtype(dtype).name} {precision} {ni}x{nj}x{nk}", flush=True)
    inp, cro, cru, coeff, out, grid = shd_case(gpu, ni, nj, nk, dtype)
    prec = stencils.WIDEN if precision == "widen" else stencils.NATIVE
    stencils.simple_hori_diff(out, inp, cro, cru, coeff, grid, precision=prec)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    ref = oracle_shd(oracle, inp, cro, cru, coeff, 2, ni, nj, nk, mode)
    if dtype == np.float64:
        assert rel_err(got, ref) <= FP64_TOL
    assert np.array_equal(got, ref)
    full = out.padded_numpy()
    assert not full[:nk, :2, :].any() and not full[:nk, :, :2].any()  # halo rim untouched


@pytest.mark.parametrize("jb", [8, 16, 32])
@pytest.mark.parametrize("metric", ["j", "full"])
def test_simple_hori_diff_variants(gpu, oracle, monkeypatch, jb, metric):
    from gridtools_amd import stencils

    monkeypatch.setenv("GT355_TUNE_shd_jb", str(jb))
    ni, nj, nk = 150, 45, 4
    print(f"[case] simple_hori_diff jb={jb} metric={metric} {ni}x{nj}x{nk}", flush=True)
    inp, cro, cru, coeff, out, grid = shd_case(gpu, ni, nj, nk, np.float64, metric=metric)
    stencils.simple_hori_diff(out, inp, cro, cru, coeff, grid)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    assert np.array_equal(got, oracle_shd(oracle, inp, cro, cru, coeff, 2, ni, nj, nk, "native"))


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (23, 11, 43)])
def test_simple_hori_diff_reference_case(gpu, ni, nj, nk):
    """The reference test itself: repository input, j-only metric from selector<0,1,0>, coeff 0.025, verified
    against out_simple (horizontal_diffusion_repository.hpp:68-79) -- bit for bit in double."""
    from gridtools_amd import stencils
    from test_oracle_cartesian import out_simple, repo_crlat
    from test_oracle_hdiff import repo_in

    print(f"[case] simple_hori_diff reference {ni}x{nj}x{nk}", flush=True)
    h = 2
    d1, d2 = ni + 2 * h, nj + 2 * h
    inp_a = repo_in(d1, d2, nk, d1, d2)  # [k, j, i]
    cro_a, cru_a = repo_crlat(d2)
    b = gpu.builder().type(np.float64).dimensions(d1, d2, nk).halos(h, h, 0)
    inp = b.build().from_numpy(inp_a.transpose(2, 1, 0))
    coeff = b.value(0.025).build()
    jb = b.selector(0, 1, 0)
    cro = jb.initializer(lambda i, j, k: cro_a[j]).build()
    cru = jb.initializer(lambda i, j, k: cru_a[j]).build()
    out = b.value(0).build()
    stencils.simple_hori_diff(out, inp, cro, cru, coeff, grid_of(gpu, ni, nj, nk, h))
    got = out.to_numpy()[h:h + ni, h:h + nj, :].transpose(2, 1, 0)
    ref = out_simple(inp_a, cro_a, cru_a, np.full_like(inp_a, 0.025), h, ni, nj)
    err = np.abs(got - ref)
    assert np.all((err < 1e-14) | (err < np.maximum(np.abs(got), np.abs(ref)) * 1e-14))
    assert np.array_equal(got, ref)


def test_simple_hori_diff_deterministic(gpu):
    from gridtools_amd import stencils

    inp, cro, cru, coeff, out, grid = shd_case(gpu, 200, 70, 9, np.float32)
    stencils.simple_hori_diff(out, inp, cro, cru, coeff, grid, precision=stencils.WIDEN)
    a = out.padded_numpy()
    for _ in range(3):
        stencils.simple_hori_diff(out, inp, cro, cru, coeff, grid, precision=stencils.WIDEN)
        assert np.array_equal(out.padded_numpy(), a)


# ------------------------------------------------------------------------------------------------ prepare_tracers
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_prepare_tracers_reference_case(gpu, dtype):
    """advection_pdbott_prepare_tracers.cpp:39-56: 11 members, member i holds i, rho = 1.1, expected 1.1 * i (bit for
    bit in double: the product rho * i is the expected value's own evaluation)."""
    from gridtools_amd import stencils

    ni, nj, nk = 12, 33, 61
    print(f"[case] prepare_tracers reference {np.dtype(dtype).name}", flush=True)
    b = gpu.builder().type(dtype).dimensions(ni, nj, nk)
    src = [b.value(m).build() for m in range(11)]
    dst = [b.value(-1).build() for _ in range(11)]
    rho = b.value(1.1).build()
    stencils.prepare_tracers(dst, src, rho, gpu.make_grid(ni, nj, nk))
    for m in range(11):
        got = dst[m].to_numpy().astype(np.float64)
        want = 1.1 * m
        if dtype == np.float64:
            assert np.all(got == want), m
        else:
            err = np.abs(got - want)
            assert np.all((err < 1e-6) | (err < abs(want) * 1e-6)), m
            assert np.all(got == np.float32(np.float32(1.1) * np.float32(m))), m


@pytest.mark.parametrize("nmembers", [1, 11, 40])
@pytest.mark.parametrize("dtype,precision", [(np.float64, "native"), (np.float32, "native"), (np.float32, "widen")])
def test_prepare_tracers_parity(gpu, oracle, nmembers, dtype, precision):
    from gridtools_amd import stencils

    ni, nj, nk, h = 70, 19, 5, 1
    print(f"[case] prepare_tracers {nmembers} members {np.dtype(dtype).name} {precision}", flush=True)
    b = gpu.builder().type(dtype).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    src = [b.uniform(SEED + m, 31, -1.0, 1.0).build() for m in range(nmembers)]
    dst = [b.value(0).build() for _ in range(nmembers)]
    rho = b.uniform(SEED, 32, 0.5, 1.5).build()
    grid = grid_of(gpu, ni, nj, nk, h)
    stencils.prepare_tracers(dst, src, rho, grid,
                             precision=stencils.WIDEN if precision == "widen" else stencils.NATIVE)
    r = rho.padded_numpy()
    V = oracle.View
    for m in range(nmembers):
        want = np.zeros_like(r)
        oracle.prepare_tracers(V(want, h, h), V(src[m].padded_numpy(), h, h), V(r, h, h), ni, nj, nk)
        got = dst[m].padded_numpy()
        assert np.array_equal(got[:nk, h:h + nj, h:h + ni], want[:nk, h:h + nj, h:h + ni]), m
        assert not got[:nk, :h, :].any() and not got[:nk, :, :h].any()


# ------------------------------------------------------------------------------------------------ copy_stencil_tuple
def test_copy_tuple_unrolled(gpu):
    """copy_stencil_tuple_unrolled (copy_stencil_tuple.cpp:116-159): five fields in, five out, in_t = i + j + k + t."""
    from gridtools_amd import stencils

    ni, nj, nk = 12, 33, 61
    print("[case] copy_stencil_tuple unrolled", flush=True)
    b = gpu.builder().type(np.float64).dimensions(ni, nj, nk)
    ins = [b.initializer(lambda i, j, k, t=t: i + j + k + t).build() for t in range(5)]
    outs = [b.value(0).build() for _ in range(5)]
    stencils.copy_fields(outs, ins, gpu.make_grid(ni, nj, nk))
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    for t in range(5):
        assert np.array_equal(outs[t].to_numpy(), (i + j + k + t).astype(np.float64)), t


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_copy_tuple_array_of_structs_and_extra_dim(gpu, dtype):
    """copy_stencil_tuple_array / _tuple (:161-205): elements array<T, 5> in one buffer, member t at origin + t with
    i stride 5; copy_stencil_tuple_extra_dim (:207-224): a fourth dimension of 5, slice t at origin + t * stride3.
    The AoS buffer is copied into the extra-dim buffer member by member and back."""
    from gridtools_amd import _abi, stencils

    ni, nj, nk, nt = 23, 11, 43, 5
    print(f"[case] copy_stencil_tuple aos/extra-dim {np.dtype(dtype).name}", flush=True)
    es = np.dtype(dtype).itemsize
    code = gpu.F64 if dtype == np.float64 else gpu.F32
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    aos = np.zeros((nk, nj, ni, nt), dtype=dtype)
    for t in range(nt):
        aos[..., t] = (i + j + k + t).transpose(2, 1, 0)
    n = aos.size
    bufs = []
    for _ in range(3):
        p = ctypes.c_void_p()
        _abi.call("gt355_malloc", ctypes.c_size_t(n * es), ctypes.byref(p))
        bufs.append(p)
    try:
        _abi.call("gt355_memcpy_h2d", bufs[0], aos.ctypes.data, ctypes.c_size_t(n * es))
        _abi.call("gt355_memset", bufs[1], 0, ctypes.c_size_t(n * es), None)
        _abi.call("gt355_memset", bufs[2], 0, ctypes.c_size_t(n * es), None)

        def field(base, off, strides):
            f = _abi.Field()
            f.origin = base.value + off * es
            f.stride = (ctypes.c_int64 * 3)(*strides)
            f.lower = (ctypes.c_int32 * 3)(0, 0, 0)
            f.upper = (ctypes.c_int32 * 3)(ni, nj, nk)
            f.dtype = code
            return f

        aos_f = lambda base: [field(base, t, (nt, nt * ni, nt * ni * nj)) for t in range(nt)]
        ext_f = lambda base: [field(base, t * ni * nj * nk, (1, ni, ni * nj)) for t in range(nt)]
        grid = gpu.make_grid(ni, nj, nk)
        stencils.copy_fields(ext_f(bufs[1]), aos_f(bufs[0]), grid)
        stencils.copy_fields(aos_f(bufs[2]), ext_f(bufs[1]), grid)
        ext = np.zeros((nt, nk, nj, ni), dtype=dtype)
        back = np.zeros_like(aos)
        _abi.call("gt355_memcpy_d2h", ext.ctypes.data, bufs[1], ctypes.c_size_t(n * es))
        _abi.call("gt355_memcpy_d2h", back.ctypes.data, bufs[2], ctypes.c_size_t(n * es))
        for t in range(nt):
            assert np.array_equal(ext[t], (i + j + k + t).transpose(2, 1, 0).astype(dtype)), t
        assert np.array_equal(back, aos)
    finally:
        for p in bufs:
            _abi.lib().gt355_free(p)


def test_copy_fields_masked_input_and_many_pairs(gpu):
    """A j-only input broadcasts over i and k; 20 pairs take two launches."""
    from gridtools_amd import stencils

    ni, nj, nk, h = 40, 9, 3, 1
    print("[case] copy_fields masked + 20 pairs", flush=True)
    b = gpu.builder().type(np.float32).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    jonly = b.selector(0, 1, 0).initializer(lambda i, j, k: 10.0 * j).build()
    ins = [jonly] + [b.uniform(SEED + t, 41, -1.0, 1.0).build() for t in range(19)]
    outs = [b.value(0).build() for _ in range(20)]
    stencils.copy_fields(outs, ins, grid_of(gpu, ni, nj, nk, h))
    got0 = outs[0].to_numpy()[h:h + ni, h:h + nj, :]
    assert np.array_equal(got0, np.broadcast_to(10.0 * np.arange(h, h + nj)[None, :, None], got0.shape))
    for t in range(1, 20):
        assert np.array_equal(outs[t].to_numpy()[h:h + ni, h:h + nj, :], ins[t].to_numpy()[h:h + ni, h:h + nj, :])
        assert not outs[t].padded_numpy()[:nk, :h, :].any()


# ------------------------------------------------------------------------------------------------ extended_4D
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_extended_4d_reference_case(gpu, oracle, dtype, tol):
    """extended_4D.cpp:104-134: phi = 10, psi = 11, f = 1.3 over (i,j,k,a,b,c), jac(i,j,k,q) = 1 + q; expected
    (jac0 + jac1) phi psi f at the reference's tolerance, and the oracle's bits."""
    from gridtools_amd import stencils

    ni, nj, nk = 12, 33, 61
    print(f"[case] extended_4d reference {np.dtype(dtype).name}", flush=True)
    result = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2, 2, 2).build()
    f = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2, 2, 2).value(1.3).build()
    jac = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2).initializer(lambda i, j, k, q: 1 + q).build()
    stencils.extended_4d(result, f, jac, 10.0, 11.0, gpu.make_grid(ni, nj, nk))
    got = result.to_numpy().astype(np.float64)
    assert got.shape == (ni, nj, nk, 2, 2, 2)
    want = (1.0 + 2.0) * 10 * 11 * float(dtype(1.3))
    err = np.abs(got - want)
    assert np.all((err < tol) | (err < want * tol))
    ref = np.zeros_like(result.padded_numpy())
    oracle.extended_4d(ref, f.padded_numpy(), jac.padded_numpy(), np.full(16, 10.0), np.full(16, 11.0), ni, nj, nk)
    assert np.array_equal(result.padded_numpy()[..., :ni], ref[..., :ni])


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (70, 19, 5), (256, 64, 80)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_extended_4d_parity(gpu, oracle, ni, nj, nk, dtype):
    from gridtools_amd import stencils

    print(f"[case] extended_4d {np.dtype(dtype).name} {ni}x{nj}x{nk}", flush=True)
    rng = np.random.default_rng(ni * 1000 + nk)
    phi = rng.uniform(-1, 1, size=(2, 2, 2, 2))
    psi = rng.uniform(-1, 1, size=(2, 2, 2, 2))
    result = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2, 2, 2).build()
    f = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2, 2, 2).uniform(SEED, 51, -1.0, 1.0).build()
    jac = gpu.builder().type(dtype).dimensions(ni, nj, nk, 2).uniform(SEED, 52, 0.5, 1.5).build()
    stencils.extended_4d(result, f, jac, phi, psi, gpu.make_grid(ni, nj, nk))
    ref = np.zeros_like(result.padded_numpy())
    oracle.extended_4d(ref, f.padded_numpy(), jac.padded_numpy(), phi, psi, ni, nj, nk)
    got = result.padded_numpy()
    assert np.array_equal(got[..., :ni], ref[..., :ni])
    assert not got[..., ni:].any()  # the i padding is never written
