"""Closed-shell (spin-adapted) CCSD contexts (ccsd_rhf_*) on the GPU.

Residuals at random amplitudes and converged energies against the
closed-shell oracle (itself pinned by the spin-orbital oracle), the same
energies against the spin-orbital GPU context on the same RHF inputs (MO and
AO input), H2O / 6-31G (config 1), and the error paths.
"""
import numpy as np
import pytest

from oracle import rccsd_ref, rhf_spin

pytestmark = pytest.mark.gpu


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _amps(o, v, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal((o, v)) * scale
    t2 = rng.standard_normal((o, o, v, v)) * scale
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


@pytest.mark.parametrize("n,nocc,nf,seed", [(8, 3, 0, 5), (9, 4, 1, 6), (7, 2, 0, 7), (20, 5, 1, 9), (26, 6, 0, 10),
                                          (6, 1, 0, 12), (7, 3, 2, 13), (7, 5, 0, 14), (12, 8, 1, 15),
                                          (6, 5, 0, 16), (40, 28, 2, 17)])
def test_residual_matches_oracle(n, nocc, nf, seed):
    capi = _capi()
    print("case:", n, nocc, nf, seed, flush=True)
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=seed, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, nocc, nf)
    t1, t2 = _amps(I.o, I.v, seed)
    R1, R2 = rccsd_ref.rhs(I, t1, t2)
    with capi.RhfContext(eps, mo, nocc, n_frozen=nf) as ctx:
        r1, r2 = ctx.residual(t1, t2)
        e = ctx.energy(t1, t2)
    scale = max(np.abs(R2).max(), np.abs(R1).max())
    assert np.abs(r1 - R1).max() <= 1e-13 * scale
    assert np.abs(r2 - R2).max() <= 1e-13 * scale
    assert abs(e - rccsd_ref.energy(I, t1, t2)) <= 1e-13


def test_solve_matches_oracle_and_spin_orbital_context():
    capi = _capi()
    n, nocc = 10, 4
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=31, scale=0.05)
    I = rccsd_ref.RInts(eps, mo, nocc)
    e_ref, t1r, t2r, _ = rccsd_ref.solve(I, toler=1e-12, max_iter=300)
    with capi.RhfContext(eps, mo, nocc, diis_dim=8) as ctx:
        r = ctx.solve(1e-12, 300)
        t1, t2 = ctx.get_amplitudes()
    with capi.Context.from_rhf(eps, mo, nocc, diis_dim=8) as cs:
        rs = cs.solve(1e-12, 300)
    assert r["converged"] == 1
    assert abs(r["e_corr"] - e_ref) < 1e-10
    assert abs(r["e_corr"] - rs["e_corr"]) < 1e-10
    assert abs(r["e_mp2"] - rhf_spin.mp2_closed_shell(eps, mo, nocc)) < 1e-12
    assert np.abs(t2 - t2r).max() < 1e-8 and np.abs(t1 - t1r).max() < 1e-8
    assert np.abs(t2 - t2.transpose(1, 0, 3, 2)).max() < 1e-12  # t2[i,j,a,b] = t2[j,i,b,a]


def test_ao_input_and_frozen_core_match_spin_orbital_context():
    """AO integrals with MO coefficients (transformed on the device), one frozen core orbital."""
    capi = _capi()
    n, nocc, nf = 14, 5, 1
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=41, scale=0.05)
    Cm = np.linalg.qr(np.random.default_rng(41).standard_normal((n, n)))[0]
    # AO integrals whose MO transform with Cm is mo: (mu nu|la si) = sum C C C C (pq|rs) with C^-1 = C^T
    ao = np.einsum("pqrs,ap,bq,cr,ds->abcd", mo, Cm, Cm, Cm, Cm, optimize=True)
    with capi.RhfContext(eps, ao, nocc, mo_coeff=Cm, n_frozen=nf) as ctx:
        r = ctx.solve(1e-12, 300)
    with capi.Context.from_rhf(eps, ao, nocc, mo_coeff=Cm, n_frozen=nf) as cs:
        rs = cs.solve(1e-12, 300)
    I = rccsd_ref.RInts(eps, mo, nocc, nf)
    e_ref = rccsd_ref.solve(I, toler=1e-12, max_iter=300)[0]
    assert abs(r["e_corr"] - e_ref) < 1e-10 and abs(rs["e_corr"] - e_ref) < 1e-10


def test_config1_h2o_631g_closed_shell():
    """H2O / 6-31G (13 basis functions, 5 doubly occupied): E_corr of the
    closed-shell context equals the spin-orbital context's to 1e-10 Eh."""
    capi = _capi()
    from tests.test_rhf import WATER_ANG
    r = capi.rhf([8, 1, 1], WATER_ANG, "6-31G", angstrom=True)
    with capi.RhfContext(r["eps"], r["mo_eri"], r["nocc"]) as ctx:
        e = ctx.solve(1e-12, 300)["e_corr"]
    with capi.Context.from_rhf(r["eps"], r["mo_eri"], r["nocc"]) as cs:
        es = cs.solve(1e-12, 300)["e_corr"]
    assert abs(e - es) < 1e-10


@pytest.mark.timeout(600)
def test_medium_size_matches_spin_orbital_context():
    """n = 60, 10 doubly occupied (spin orbitals O = 20, V = 100): tiles, split-K
    and permutes at production shapes; energies of both contexts to 1e-9."""
    capi = _capi()
    n, nocc = 60, 10
    eps, mo = rhf_spin.random_rhf(n, nocc, seed=51, scale=0.02)
    with capi.RhfContext(eps, mo, nocc) as ctx:
        r = ctx.solve(1e-11, 200)
    with capi.Context.from_rhf(eps, mo, nocc) as cs:
        rs = cs.solve(1e-11, 200)
    print("closed-shell %.3f ms/it, spin-orbital %.3f ms/it" % (1e3 * r["seconds_per_iter"], 1e3 * rs["seconds_per_iter"]))
    assert r["converged"] == 1 and rs["converged"] == 1
    assert abs(r["e_corr"] - rs["e_corr"]) < 1e-9


def test_error_paths():
    capi = _capi()
    eps, mo = rhf_spin.random_rhf(8, 3, seed=61, scale=0.05)
    with pytest.raises(capi.CCSDError) as e:
        capi.RhfContext(eps, mo, 3, precision="fp32")
    assert e.value.code == -1  # FP64 only
    with pytest.raises(capi.CCSDError) as e:
        capi.RhfContext(eps, mo, 3, hbm_budget_bytes=1 << 20)
    assert e.value.code == -3  # over the budget, refused before allocating
    with capi.RhfContext(eps, mo, 3, diis_dim=8) as ctx:
        with pytest.raises(capi.CCSDError) as e:
            ctx.solve(1e-14, 2)
        assert e.value.code == -5
        r = ctx.solve(1e-14, 2, allow_unconverged=True)
        assert r["iterations"] == 2 and r["converged"] == 0 and np.isfinite(r["e_corr"])
    assert capi.rhf_plan_bytes(8, 3) > 0 and capi.rhf_plan_bytes(8, 0) == 0
