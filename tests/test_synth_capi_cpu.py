"""CPU-only checks: the synthetic generator and the C-ABI library surface.

No compute call reaches a GPU here: device-requiring entry points must return
CCSD_ERR_NO_DEVICE on a machine without one, and host-side validation errors
(null pointers, NaN, vanishing denominators) must be reported before any
device work.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ccsd_amd.h")


def test_splitmix64_known_answers():
    # published SplitMix64 outputs for state 0 (Steele, Lea & Flood 2014 reference implementation)
    h = synth.hstream(0, np.arange(3))
    assert [int(x) for x in h] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_full_tensor_symmetries():
    O, V, seed = 3, 5, 42
    g = synth.full_tensor(O, V, seed, 0.02)
    assert np.array_equal(g, -g.transpose(1, 0, 2, 3))
    assert np.array_equal(g, -g.transpose(0, 1, 3, 2))
    assert np.array_equal(g, g.transpose(2, 3, 0, 1))
    n = O + V
    for p in range(n):
        assert np.all(g[p, p] == 0.0) and np.all(g[:, :, p, p] == 0.0)


def test_packed_blocks_consistent_with_full_tensor():
    O, V, seed = 3, 6, 43
    g = synth.full_tensor(O, V, seed, 0.02)
    ref = synth.pack_dense(synth.blocks_from_full(g, O), O, V)
    prob = synth.make_problem(O, V, seed, scale=0.02)
    for b in synth.BLOCKS:
        assert np.array_equal(prob[b], ref[b]), b
    d = synth.dense_blocks(prob)
    for b, x in synth.blocks_from_full(g, O).items():
        assert np.array_equal(d[b], x), b


def test_generator_statistics():
    vals = synth.packed_block("vvvv", 4, 60, 7, 0.02)
    iu = np.triu_indices(vals.shape[0])
    x = vals[iu]
    assert abs(x.mean()) < 5 * 0.02 / np.sqrt(x.size)
    assert abs(x.std() / 0.02 - 1.0) < 0.01
    eo, ev = synth.orbital_energies(20, 180, 20180)
    assert eo.min() >= -1.5 and eo.max() <= -0.3 and ev.min() >= 0.2 and ev.max() <= 3.0
    assert np.all(np.diff(eo) >= 0) and np.all(np.diff(ev) >= 0)


# ----------------------------------------------------------------- C ABI surface

def _declared_symbols():
    """Every C-ABI function declared in include/*.h."""
    inc = os.path.dirname(HEADER)
    src = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    decl = r"^\s*(?:int|void|size_t|long|const char\s*\*)\s+(ccsd_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from ccsd_parallel_amd import capi
    lib = capi.load()
    names = _declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(capi.EXPORTED) == set(names)


def test_strerror_and_defaults():
    from ccsd_parallel_amd import capi
    for code in range(0, -9, -1):
        assert capi.strerror(code) and capi.strerror(code) != "unknown error"
    o = capi.default_options()
    assert o.diis_dim == 8 and o.final_fp64_sweeps == 3 and o.precision == 0
    assert o.switch_factor == 100.0 and o.switch_threshold == 0.0 and o.diis_reset_on_switch == 1
    assert o.strassen_min_order == 0 and o.lambda_tol == 0.0  # the last two fields: the struct sizes agree


def test_plan_fits_hbm():
    from ccsd_parallel_amd import capi
    for prec in ("fp64", "fp32", "bf16x3"):
        b = capi.plan_bytes(40, 300, prec)
        assert 3e10 < b < 288e9
    assert capi.plan_bytes(1, 10) == 0


def _problem(O=3, V=5, seed=1):
    return synth.make_problem(O, V, seed, scale=0.02)


def test_host_validation_errors():
    from ccsd_parallel_amd import capi
    lib = capi.load()
    h = C.c_void_p()
    assert lib.ccsd_create(None, None, C.byref(h)) == -1
    assert lib.ccsd_create_synthetic(1, 10, 1, 0.02, None, C.byref(h)) == -1
    assert lib.ccsd_create_synthetic(4, 10, 1, -1.0, None, C.byref(h)) == -1
    prob = _problem()
    bad = dict(prob)
    bad["oovv"] = prob["oovv"].copy()
    bad["oovv"][0, 0] = np.nan
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_problem(bad)
    assert e.value.code == -4
    deg = dict(prob)
    f = prob["f"].copy()
    f[4, 4] = f[0, 0]  # occupied and virtual degenerate
    deg["f"] = f
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_problem(deg)
    assert e.value.code == -6
    # RHF input: shape/occupation checks and non-finite inputs are rejected on the host
    eps = np.array([-1.0, -0.5, 0.5, 1.0])
    eri = np.zeros((4, 4, 4, 4))
    for n_occ, n_frozen in ((0, 0), (4, 0), (2, 2), (2, -1)):
        with pytest.raises(capi.CCSDError) as e:
            capi.Context.from_rhf(eps, eri, n_occ, n_frozen=n_frozen)
        assert e.value.code == -1, (n_occ, n_frozen)
    bad_eri = eri.copy()
    bad_eri[1, 2, 3, 0] = np.inf
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_rhf(eps, bad_eri, 2)
    assert e.value.code == -4
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_rhf(eps, eri, 2, mo_coeff=np.full((4, 4), np.nan))
    assert e.value.code == -4
    deg_eps = np.array([-1.0, 0.5, 0.5, 1.0])  # occupied 2 and virtual 3 coincide
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_rhf(deg_eps, eri, 2)
    assert e.value.code == -6
    assert lib.ccsd_create_rhf(None, None, C.byref(h)) == -1
    A = np.eye(3)
    assert lib.ccsd_ao_to_mo(0, capi._ptr(A), capi._ptr(A), capi._ptr(A), 0) == -1


def test_no_device_is_reported():
    from ccsd_parallel_amd import capi
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.synthetic(4, 8, 1, 0.02)
    assert e.value.code == -7
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_problem(_problem())
    assert e.value.code == -7
    A = np.ones((2, 2))
    with pytest.raises(capi.CCSDError) as e:
        capi.selftest_gemm(A, A, A)
    assert e.value.code == -7
    with pytest.raises(capi.CCSDError) as e:
        capi.ao_to_mo(np.eye(3), np.zeros((3, 3, 3, 3)))
    assert e.value.code == -7
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_rhf(np.array([-1.0, -0.5, 0.5, 1.0]), np.zeros((4, 4, 4, 4)), 2)
    assert e.value.code == -7
