"""Pins of the CPU adjoint oracle (oracle.adjoint_ref).

* the torch restatement of the residual equals ccsd_ref.rhs;
* the Lagrangian gradient equals central finite differences (h = 1e-5) of the
  numpy oracle's converged energy along random symmetric directions in
  (f, <pq||rs>) -- the directions respect the antisymmetry and pq<->rs symmetry
  because they come from the same generator with another seed.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_ref, ccsd_ref, synth


def _prob(O, V, seed):
    return synth.make_problem(O, V, seed, scale=1.0 / np.sqrt(O * synth.p2(V)))


def _direction(O, V, seed):
    d = synth.make_problem(O, V, seed + 1000, scale=1.0)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(d["f"].shape)
    d["f"] = 0.5 * (a + a.T)
    return d


def _energy(prob):
    E, *_ = ccsd_ref.ccsd_energy(prob, toler=1e-14, max_iter=500, diis_dim=8)
    return E


def test_torch_residual_matches_oracle():
    O, V = 4, 8
    prob = _prob(O, V, 3)
    rng = np.random.default_rng(1)
    t1 = 0.05 * rng.standard_normal((O, V))
    t2p = 0.05 * rng.standard_normal((synth.p2(O), synth.p2(V)))
    tp = adjoint_ref.TorchProblem(prob)
    R1, R2 = adjoint_ref.rhs_torch(tp, torch.tensor(t1), torch.tensor(t2p))
    I = ccsd_ref.Ints.from_problem(prob)
    S1, S2 = ccsd_ref.rhs(I, t1, synth.unpack_pairs(synth.unpack_pairs(t2p, V, 1), O, 0))
    assert np.max(np.abs(R1.numpy() - S1)) < 1e-14
    assert np.max(np.abs(R2.numpy() - synth.pack_pairs(synth.pack_pairs(S2, O, 0), V, 1))) < 1e-14
    e = adjoint_ref.energy_torch(tp, torch.tensor(t1), torch.tensor(t2p)).item()
    assert abs(e - ccsd_ref.energy(I, t1, synth.unpack_pairs(synth.unpack_pairs(t2p, V, 1), O, 0))) < 1e-15


@pytest.mark.parametrize("O,V,seed", [(2, 6, 4), (4, 8, 5)])
def test_gradient_matches_finite_differences(O, V, seed):
    prob = _prob(O, V, seed)
    E, grad, _, linfo = adjoint_ref.adjoint(prob, toler=1e-13)
    assert linfo["residual"] < 1e-12
    h = 1e-5
    for k in range(3):
        dirn = _direction(O, V, seed * 10 + k)
        pred = sum(float(np.sum(grad[b] * dirn[b])) for b in ("f",) + synth.BLOCKS)
        plus = {b: prob[b] + h * dirn[b] for b in ("f",) + synth.BLOCKS}
        minus = {b: prob[b] - h * dirn[b] for b in ("f",) + synth.BLOCKS}
        fd = (_energy(dict(prob, **plus)) - _energy(dict(prob, **minus))) / (2 * h)
        assert abs(fd - pred) <= 1e-7 * max(1.0, abs(fd)), (k, fd, pred)
