"""GPU AO->MO transform and RHF context set-up (ccsd_ao_to_mo, ccsd_create_rhf).

The reference transforms its AO integrals on the host (Prepear_AOs,
ccsd_memory_optimized.cpp:2698-3522) and slices the antisymmetrised
spin-orbital integrals on the fly (Fill_integ*_2D, :3574-4779).  Here both run
on the device; the checks are:
  * the transform against the host RHF producer's own MO integrals for water
    (that producer is pinned by published SCF/MP2/CCSD values), against the
    four-index definition on a non-orthogonal C, and by an orthogonal round
    trip (Q then Q^T gives the input back);
  * the device-packed spin-orbital problem against the oracle's restatement
    (oracle/rhf_spin.py, pinned by the closed-shell MP2 formula, two-electron
    FCI and the frozen-core slicing identity): right-hand sides at random
    amplitudes to 1e-12 of the largest entry, MP2 and converged energies to
    1e-11 Eh, with and without a frozen core.
"""
import numpy as np
import pytest

from oracle import ccsd_ref, rhf_spin, synth

from tests.test_rhf import WATER_BOHR

pytestmark = pytest.mark.gpu


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def test_ao_to_mo_matches_host_rhf_transform():
    capi = _capi()
    r = capi.rhf([8, 1, 1], WATER_BOHR, "6-31G", want_ao_eri=True)
    mo = capi.ao_to_mo(r["C"], r["ao_eri"])
    assert np.max(np.abs(mo - r["mo_eri"])) < 1e-12


def test_ao_to_mo_definition_and_round_trip():
    capi = _capi()
    n = 9
    rng = np.random.default_rng(31)
    ao = rhf_spin.random_rhf(n, 3, 32)[1]
    Cn = rng.standard_normal((n, n))  # not orthogonal: every index really is transformed
    ref = np.einsum("abcd,ap,bq,cr,ds->pqrs", ao, Cn, Cn, Cn, Cn, optimize=True)
    got = capi.ao_to_mo(Cn, ao)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    back = capi.ao_to_mo(Q.T, capi.ao_to_mo(Q, ao))
    assert np.max(np.abs(back - ao)) < 1e-14
    assert np.array_equal(capi.ao_to_mo(np.eye(n), ao), ao)


def _rhs(ctx, O, V, seed):
    rng = np.random.default_rng(seed)
    t1 = 0.05 * rng.standard_normal((O, V))
    t2 = 0.05 * rng.standard_normal((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    t2p = synth.pack_pairs(synth.pack_pairs(t2, O, 0), V, 1)
    return ctx.residual(t1, t2p)


@pytest.mark.parametrize("n,nocc,nfrozen,seed", [(8, 3, 0, 41), (10, 4, 1, 42), (9, 2, 0, 43)])
def test_rhf_context_matches_oracle_packing(n, nocc, nfrozen, seed):
    capi = _capi()
    eps, mo = rhf_spin.random_rhf(n, nocc, seed)
    rng = np.random.default_rng(seed + 100)
    Cq = np.linalg.qr(rng.standard_normal((n, n)))[0]
    ao = np.einsum("pqrs,ap,bq,cr,ds->abcd", mo, Cq, Cq, Cq, Cq, optimize=True)  # transforms back to mo with Cq
    prob = rhf_spin.problem_from_rhf(eps, mo, nocc, nfrozen)
    O, V = prob["O"], prob["V"]
    print("case:", n, nocc, nfrozen, seed, "O", O, "V", V, flush=True)
    with capi.Context.from_problem(prob, diis_dim=8) as ref, \
            capi.Context.from_rhf(eps, ao, nocc, mo_coeff=Cq, n_frozen=nfrozen, diis_dim=8) as dev_ao, \
            capi.Context.from_rhf(eps, mo, nocc, n_frozen=nfrozen, diis_dim=8) as dev_mo:
        assert (dev_ao.O, dev_ao.V) == (O, V)
        r1, r2 = _rhs(ref, O, V, seed)
        for ctx in (dev_ao, dev_mo):
            g1, g2 = _rhs(ctx, O, V, seed)
            scale = max(1.0, np.abs(r2).max(), np.abs(r1).max())
            assert np.abs(g1 - r1).max() < 1e-12 * scale
            assert np.abs(g2 - r2).max() < 1e-12 * scale
        e_ref = ref.solve(1e-11, 200)
        for ctx in (dev_ao, dev_mo):
            e = ctx.solve(1e-11, 200)
            assert abs(e["e_mp2"] - e_ref["e_mp2"]) < 1e-12
            assert abs(e["e_corr"] - e_ref["e_corr"]) < 1e-11
    if nfrozen == 0:
        assert abs(e_ref["e_mp2"] - rhf_spin.mp2_closed_shell(eps, mo, nocc)) < 1e-12
