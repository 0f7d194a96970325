"""Pins of the FP64 oracle (CPU only).

The reference ships no tests or fixtures (its only known answer is a printed
STO-3G energy, ccsd_memory_optimized.cpp:2003), so the restatement of its
factorisation (oracle.ccsd_ref) is pinned by mathematics:
  * an independent textbook Stanton-Gauss formulation (oracle.ccsd_sg),
  * two-electron full CI, where CCSD is exact,
  * the MP2 closed form of the first energy (T1 = 0, T2 = <ab||ij>/D),
  * invariance of the converged energy under occupied-occupied and
    virtual-virtual rotations (exercises every off-diagonal Fock term),
  * the DIIS and Jacobi fixed points coinciding,
  * planted mistakes in the oracle being caught by these checks.
"""
import json
import os

import numpy as np
import pytest

from oracle import ccsd_ref, ccsd_sg, fci2, synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _problem_from_full(O, V, seed, scale, f=None):
    g = synth.full_tensor(O, V, seed, scale)
    if f is None:
        eo, ev = synth.orbital_energies(O, V, seed)
        f = np.diag(np.concatenate([eo, ev]))
    prob = dict(O=O, V=V, f=f, **synth.pack_dense(synth.blocks_from_full(g, O), O, V))
    return g, f, prob


def _random_t(O, V, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    t1 = amp * rng.standard_normal((O, V))
    t2 = amp * rng.standard_normal((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    return t1, t2 - t2.transpose(0, 1, 3, 2)


def _noncanonical(f, seed, eps=0.02):
    rng = np.random.default_rng(seed + 99)
    a = rng.standard_normal(f.shape)
    return f + eps * (a + a.T)


@pytest.mark.parametrize("O,V,seed", [(2, 6, 1), (4, 8, 2), (6, 10, 3), (3, 7, 4)])
def test_reference_factorisation_matches_textbook(O, V, seed):
    g, f, prob = _problem_from_full(O, V, seed, 0.05)
    f = _noncanonical(f, seed)
    prob["f"] = f
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = _random_t(O, V, seed)
    R1, R2 = ccsd_ref.rhs(I, t1, t2)
    S1, S2 = ccsd_sg.rhs_sg(f, g, O, t1, t2)
    assert np.max(np.abs(R1 - S1)) < 1e-14
    assert np.max(np.abs(R2 - S2)) < 1e-14
    assert abs(ccsd_ref.energy(I, t1, t2) - ccsd_sg.energy_sg(f, g, O, t1, t2)) < 1e-14


@pytest.mark.parametrize("V,seed,noncanon", [(6, 5, False), (10, 6, False), (8, 7, True)])
def test_two_electron_ccsd_is_fci(V, seed, noncanon):
    g, f, prob = _problem_from_full(2, V, seed, 0.05)
    if noncanon:
        f = _noncanonical(f, seed)
        prob["f"] = f
    e_fci = fci2.fci_two_electron(f, g)[0]
    E, _, _, info = ccsd_ref.ccsd_energy(prob, toler=1e-13, max_iter=500, diis_dim=8)
    assert info["converged"]
    assert abs(E - e_fci) < 1e-11


def test_first_energy_is_mp2():
    O, V, seed = 6, 16, 8
    prob = synth.make_problem(O, V, seed, scale=0.03)
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = ccsd_ref.mp2_guess(I)
    e = ccsd_ref.energy(I, t1, t2)
    # closed form: sum_{i<j, a<b} |<ij||ab>|^2 / (e_i + e_j - e_a - e_b)
    eo, ev = np.diag(prob["f"])[:O], np.diag(prob["f"])[O:]
    io, jo = synth.pair_list(O)
    av, bv = synth.pair_list(V)
    D = (eo[io] + eo[jo])[:, None] - (ev[av] + ev[bv])[None, :]
    assert abs(e - np.sum(prob["oovv"] ** 2 / D)) < 1e-14 * abs(e)


def _rotate(g, f, O, seed):
    rng = np.random.default_rng(seed)
    n = f.shape[0]
    U = np.zeros((n, n))
    U[:O, :O] = np.linalg.qr(rng.standard_normal((O, O)))[0]
    U[O:, O:] = np.linalg.qr(rng.standard_normal((n - O, n - O)))[0]
    g2 = np.einsum("pqrs,pa,qb,rc,sd->abcd", g, U, U, U, U, optimize=True)
    return g2, U.T @ f @ U


def test_energy_invariant_under_orbital_rotations():
    O, V, seed = 4, 8, 9
    g, f, prob = _problem_from_full(O, V, seed, 0.05)
    E0, *_ = ccsd_ref.ccsd_energy(prob, toler=1e-12, max_iter=300, diis_dim=8)
    g2, f2 = _rotate(g, f, O, seed)
    # rotated orbitals: non-canonical Fock, same correlation energy
    assert np.max(np.abs(f2 - np.diag(np.diag(f2)))) > 1e-3
    prob2 = dict(O=O, V=V, f=f2, **synth.pack_dense(synth.blocks_from_full(g2, O), O, V))
    E1, *_ = ccsd_ref.ccsd_energy(prob2, toler=1e-12, max_iter=300, diis_dim=8)
    assert abs(E0 - E1) < 1e-9


def test_diis_and_jacobi_share_the_fixed_point():
    O, V, seed = 6, 14, 10
    prob = synth.make_problem(O, V, seed, scale=1.0 / np.sqrt(O * synth.p2(V)))
    Ej, t1j, t2j, ij = ccsd_ref.ccsd_energy(prob, toler=1e-12, max_iter=2000, diis_dim=0)
    Ed, t1d, t2d, idd = ccsd_ref.ccsd_energy(prob, toler=1e-12, max_iter=300, diis_dim=8)
    assert ij["converged"] and idd["converged"]
    assert idd["iterations"] < ij["iterations"]
    assert abs(Ej - Ed) < 1e-10
    assert np.max(np.abs(t2j - t2d)) < 1e-8


@pytest.mark.parametrize("plant", ["drop_ring", "wrong_w1_half", "swap_fae_sign"])
def test_planted_mistakes_are_caught(plant, monkeypatch):
    """A plausible slip in the factorisation must fail the textbook cross-check."""
    O, V, seed = 4, 8, 12
    g, f, prob = _problem_from_full(O, V, seed, 0.05)
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = _random_t(O, V, seed)
    orig = ccsd_ref.intermediates

    def planted(I_, t1_, t2_):
        X = orig(I_, t1_, t2_)
        if plant == "drop_ring":
            X["W4"] = X["W4"] * 0.0
        elif plant == "wrong_w1_half":
            X["W1"] = X["W1"] - 0.25 * np.einsum("ijef,mnef->mnij", X["tau"], I_.oovv)
        elif plant == "swap_fae_sign":
            X["Fae"] = -X["Fae"]
        return X

    monkeypatch.setattr(ccsd_ref, "intermediates", planted)
    R1, R2 = ccsd_ref.rhs(I, t1, t2)
    S1, S2 = ccsd_sg.rhs_sg(f, g, O, t1, t2)
    assert max(np.max(np.abs(R1 - S1)), np.max(np.abs(R2 - S2))) > 1e-6


def test_golden_fixtures():
    """Regression against committed oracle outputs (tests/golden/make_golden.py)."""
    with open(os.path.join(GOLDEN, "ccsd_small.json")) as fh:
        gold = json.load(fh)
    for case in gold["cases"]:
        O, V, seed, scale = case["O"], case["V"], case["seed"], case["scale"]
        prob = synth.make_problem(O, V, seed, scale=scale)
        I = ccsd_ref.Ints.from_problem(prob)
        _, _, _, info = ccsd_ref.solve(I, toler=0.0, max_iter=len(case["jacobi_energies"]), diis_dim=0)
        assert np.max(np.abs(np.array(info["energies"]) - case["jacobi_energies"])) < 1e-12
        E, _, _, info = ccsd_ref.solve(I, toler=1e-11, max_iter=300, diis_dim=8)
        assert abs(E - case["e_corr"]) < 1e-10
        blk = synth.packed_block("vvvv", O, V, seed, scale, rows=(0, 1))
        assert abs(float(blk.sum()) - case["vvvv_row0_sum"]) < 1e-14


def test_cpu_restatement_loops_compute_the_ladder_and_ring_terms():
    """The C++ loops the CPU baseline times (oracle/ccsd_cpu.cpp) compute the
    oracle's terms: packed dot products over c<d give the pp ladder
    1/2 sum_cd <ab||cd> tau_ij^cd, and dot products over (m,e) give the ring
    sum_me t_im^ae W_mbej, at random amplitudes."""
    from oracle import cpu_build
    try:
        cpu_build.load()
    except RuntimeError:
        cpu_build.build(verbose=False)
    O, V, seed = 4, 9, 5
    prob = synth.make_problem(O, V, seed, scale=0.1)
    I = ccsd_ref.Ints.from_problem(prob)
    rng = np.random.default_rng(3)
    t1 = 0.05 * rng.standard_normal((O, V))
    t2 = 0.05 * rng.standard_normal((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    _, _, _, terms2, X = ccsd_ref.rhs(I, t1, t2, parts=True)
    tau_p = synth.pack_pairs(synth.pack_pairs(X["tau"], O, 0), V, 1)          # [i<j][c<d]
    lad = cpu_build.dot_rows(prob["vvvv"], tau_p)                              # [a<b][i<j]
    pp_ref = synth.pack_pairs(synth.pack_pairs(terms2["pp"], O, 0), V, 1)      # [i<j][a<b]
    assert np.abs(lad.T - pp_ref).max() < 1e-14
    assert np.abs(pp_ref).max() > 1e-4
    # ring: rows (i,a) of t_im^ae, columns (j,b) of W_mbej, both over (m,e)
    Tring = t2.transpose(0, 2, 1, 3).reshape(O * V, O * V)                      # [(i,a)][(m,e)]
    Wcol = X["W4"].transpose(3, 1, 0, 2).reshape(O * V, O * V)                  # [(j,b)][(m,e)] = W[m,b,e,j]
    ring = cpu_build.dot_rows(Tring, Wcol).reshape(O, V, O, V).transpose(0, 2, 1, 3)  # [i,j,a,b]
    ref = np.einsum("imae,mbej->ijab", t2, X["W4"])
    assert np.abs(ring - ref).max() < 1e-14 and np.abs(ref).max() > 1e-4
