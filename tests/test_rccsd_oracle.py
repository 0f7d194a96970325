"""Closed-shell (spin-adapted) CCSD for RHF inputs, on the CPU.

The closed-shell oracle (oracle/rccsd_ref.py) is pinned by the spin-orbital
oracle on the same RHF integrals: its residuals equal the alpha-beta and
alpha-alpha spin blocks of the spin-orbital residual at random amplitudes, its
energy functional equals the spin-orbital one, and the converged energies agree;
Crawford's published STO-3G water CCSD energy is reproduced.  A planted
mistake (one term dropped) is caught by the same comparison.  The contraction
planner and the term sequence that the device runs (ccsd_parallel_amd/csrc/
rplan.hpp) are executed on host arrays by tests/cpp/rccsd_cpu.cpp and checked
against the oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import ccsd_ref, rccsd_ref, rhf_spin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# o < v, o = 1 (no antisymmetric pairs), o > v (STO-3G water is o = 5, v = 2), v = 1
SHAPES = [(8, 3, 0, 5), (9, 4, 1, 6), (7, 2, 0, 7), (10, 5, 2, 8), (6, 1, 0, 12), (7, 3, 2, 13), (7, 5, 0, 14),
          (12, 8, 1, 15), (6, 5, 0, 16)]


def _amps(I, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal((I.o, I.v)) * scale
    t2 = rng.standard_normal((I.o, I.o, I.v, I.v)) * scale
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


def _spin_ints(eps, mo, nocc, nf):
    return ccsd_ref.Ints.from_problem(rhf_spin.problem_from_rhf(eps, mo, nocc, nf))


@pytest.mark.parametrize("n,nocc,nf,seed", SHAPES)
def test_residual_equals_spin_orbital_blocks(n, nocc, nf, seed):
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=seed, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, nocc, nf)
    J = _spin_ints(eps, mo, nocc, nf)
    t1, t2 = _amps(I, seed)
    T1, T2 = rccsd_ref.spin_blocks(t1, t2)
    R1, R2 = rccsd_ref.rhs(I, t1, t2)
    Q1, Q2 = ccsd_ref.rhs(J, T1, T2)
    scale = np.abs(Q2).max()
    assert np.abs(Q1[0::2, 0::2] - R1).max() <= 1e-14 * scale
    assert np.abs(Q1[1::2, 1::2] - R1).max() <= 1e-14 * scale
    assert np.abs(Q2[0::2, 1::2, 0::2, 1::2] - R2).max() <= 1e-14 * scale                       # alpha beta
    assert np.abs(Q2[0::2, 0::2, 0::2, 0::2] - (R2 - R2.transpose(0, 1, 3, 2))).max() <= 1e-14 * scale
    assert abs(rccsd_ref.energy(I, t1, t2) - ccsd_ref.energy(J, T1, T2)) <= 1e-15


def test_planted_mistake_is_caught():
    """Dropping any one T2 term (or flipping a sign) breaks the block equality."""
    eps, mo = rhf_spin.random_rhf(8, 3, seed=11, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, 3)
    J = _spin_ints(eps, mo, 3, 0)
    t1, t2 = _amps(I, 11)
    T1, T2 = rccsd_ref.spin_blocks(t1, t2)
    _, _, _, r2, _ = rccsd_ref.rhs(I, t1, t2, parts=True)
    ab = ccsd_ref.rhs(J, T1, T2)[1][0::2, 1::2, 0::2, 1::2]
    for k in r2:
        if k == "ints":
            continue
        wrong = sum(x for kk, x in r2.items() if kk != k)
        assert np.abs(wrong - ab).max() > 1e-6, k
        flipped = wrong - r2[k]
        assert np.abs(flipped - ab).max() > 1e-6, k


def test_converged_energy_equals_spin_orbital():
    eps, mo = rhf_spin.random_rhf(8, 3, seed=21, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, 3)
    e_r, t1, t2, _ = rccsd_ref.solve(I, toler=1e-12, max_iter=300)
    e_s = ccsd_ref.ccsd_energy(rhf_spin.problem_from_rhf(eps, mo, 3), toler=1e-12, max_iter=300, diis_dim=8)[0]
    assert abs(e_r - e_s) < 1e-10
    # the first energy is the closed-shell MP2 formula
    t1m, t2m = rccsd_ref.mp2_guess(I)
    assert abs(rccsd_ref.energy(I, t1m, t2m) - rhf_spin.mp2_closed_shell(eps, mo, 3)) < 1e-14


def test_crawford_sto3g_water_closed_shell():
    """Crawford's H2O / STO-3G CCSD correlation energy, -0.070680088376 Eh."""
    from ccsd_parallel_amd import capi
    from tests.test_rhf import WATER_BOHR
    r = capi.rhf([8, 1, 1], WATER_BOHR, "STO-3G")
    I = rccsd_ref.RInts(r["eps"], r["mo_eri"], r["nocc"])
    e = rccsd_ref.solve(I, toler=1e-12, max_iter=300)[0]
    assert abs(e - (-0.070680088376)) < 1e-10


def _host_executor(tmp_path):
    so = str(tmp_path / "librcpu.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", os.path.join(ROOT, "tests", "cpp", "rccsd_cpu.cpp"),
           "-o", so]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    lib.rcpu_rhs.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, C.POINTER(C.c_long)]
    return lib


@pytest.mark.parametrize("n,nocc,nf,seed", SHAPES)
def test_planner_and_term_sequence_on_host(tmp_path, n, nocc, nf, seed):
    lib = _host_executor(tmp_path)
    ptr = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=seed, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, nocc, nf)
    t1, t2 = _amps(I, seed)
    R1, R2 = rccsd_ref.rhs(I, t1, t2)
    r1, r2 = np.zeros_like(R1), np.zeros_like(R2)
    perm = C.c_long(0)
    rc = lib.rcpu_rhs(n, nocc, nf, ptr(np.ascontiguousarray(mo)), ptr(t1), ptr(t2), ptr(r1), ptr(r2), C.byref(perm))
    assert rc == 0
    assert np.abs(r1 - R1).max() <= 1e-14 and np.abs(r2 - R2).max() <= 1e-14
    # only amplitude-, o^3 v- and o^4-sized operands are permuted (no o v^3 or v^4 block)
    o, v = I.o, I.v
    assert perm.value <= 12 * t2.size + 4 * o ** 3 * v + 4 * o ** 4


# ------------------------------------------------------------- the planner alone

def _ein_lib(tmp_path):
    lib = _host_executor(tmp_path)
    lp = C.POINTER(C.c_long)
    dp = C.POINTER(C.c_double)
    lib.rcpu_ein.argtypes = [C.c_char_p, C.c_double, dp, C.c_int, lp, lp, dp, C.c_int, lp, lp, C.c_double, dp,
                             C.c_int, lp, lp, C.c_long, C.c_char_p, C.c_int]
    return lib


def _ein(lib, spec, alpha, A, B, beta, Cm, scratch=1 << 16):
    """Run one planned contraction on strided numpy views (any strides)."""
    def parts(x):
        base = x.base if x.base is not None else x
        e = np.array(x.shape, dtype=np.int64)
        s = np.array([st // 8 for st in x.strides], dtype=np.int64)
        p = C.cast(x.__array_interface__["data"][0], C.POINTER(C.c_double))
        return p, x.ndim, e.ctypes.data_as(C.POINTER(C.c_long)), s.ctypes.data_as(C.POINTER(C.c_long)), (e, s, base)
    pa, pb, pc = parts(A), parts(B), parts(Cm)
    err = C.create_string_buffer(256)
    rc = lib.rcpu_ein(spec.encode(), alpha, *pa[:4], *pb[:4], beta, *pc[:4], scratch, err, 256)
    return rc, err.value.decode()


@pytest.mark.parametrize("spec,shapes", [
    ("ik,kj->ij", ((5, 7), (7, 3))),                    # plain GEMM
    ("ki,kj->ji", ((7, 5), (7, 3))),                    # transposed A and C (C permuted)
    ("bik,bkj->bij", ((4, 5, 6), (4, 6, 3))),           # one batch run
    ("abik,abkj->abij", ((2, 3, 5, 4), (2, 3, 4, 6))),  # two batch labels collapsing to one run
    ("aikb,bkj->bija", ((3, 5, 4, 2), (2, 4, 6))),      # batch labels in different places: permutes
    ("ia,jb->ijab", ((4, 3), (5, 2))),                  # outer product (K empty)
    ("iakc,kc->ia", ((3, 4, 5, 6), (5, 6))),            # GEMV (N empty)
    ("kc,iakc->ia", ((5, 6), (3, 4, 5, 6))),            # M empty
    ("kcld,ilcd->ki", ((3, 4, 5, 4), (2, 5, 4, 4))),    # K over three labels
])
def test_planner_matches_einsum(tmp_path, spec, shapes):
    lib = _ein_lib(tmp_path)
    rng = np.random.default_rng(len(spec))
    A = rng.standard_normal(shapes[0])
    B = rng.standard_normal(shapes[1])
    out_shape = np.einsum(spec, A, B).shape
    C0 = rng.standard_normal(out_shape)
    for alpha, beta in ((1.0, 0.0), (-0.5, 1.5)):
        Cm = C0.copy()
        rc, err = _ein(lib, spec, alpha, A, B, beta, Cm)
        assert rc == 0, err
        ref = alpha * np.einsum(spec, A, B) + beta * C0
        assert np.abs(Cm - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (spec, alpha, beta)


def test_planner_strided_views_and_errors(tmp_path):
    lib = _ein_lib(tmp_path)
    rng = np.random.default_rng(3)
    # strided input views (a column slice, a transposed view) and a strided output view
    big = rng.standard_normal((6, 11))
    A = big[:, 2:9]                          # (6, 7), row stride 11
    B = rng.standard_normal((3, 7)).T        # (7, 3), transposed
    Cbig = rng.standard_normal((6, 8))
    C0 = Cbig.copy()
    Cv = Cbig[:, 1:4]                        # (6, 3), row stride 8
    rc, err = _ein(lib, "ik,kj->ij", 2.0, A, B, 1.0, Cv)
    assert rc == 0, err
    ref = C0.copy()
    ref[:, 1:4] = 2.0 * A @ B + C0[:, 1:4]
    assert np.abs(Cbig - ref).max() < 1e-13     # columns outside the view untouched
    # planner errors, no arithmetic
    X = rng.standard_normal((3, 4))
    Y = rng.standard_normal((4, 5))
    Z = np.zeros((3, 5))
    for bad in ("ik,kj->ijz", "ik,kk->ij", "ik,kj,->ij", "ix,kj->ij"):
        rc, err = _ein(lib, bad, 1.0, X, Y, 0.0, Z)
        assert rc != 0 and err, bad
    rc, err = _ein(lib, "ik,jk->ij", 1.0, X, np.zeros((5, 5)), 0.0, Z)   # k: 4 against 5
    assert rc != 0 and err
    # K labels in different orders in A and B cannot share one stride: a
    # permute is needed, and a 4-element scratch cannot hold it
    A = rng.standard_normal((3, 4, 5, 4))
    B = rng.standard_normal((2, 5, 4, 4))
    rc, err = _ein(lib, "kcld,ilcd->ki", 1.0, A, B, 0.0, np.zeros((3, 2)), scratch=4)
    assert rc != 0 and "scratch" in err
