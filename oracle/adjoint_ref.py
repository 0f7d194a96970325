"""CPU FP64 adjoint of the CCSD energy (test infrastructure only).

The reference has no adjoint.  This module states the CCSD Lambda equations
and the gradient of the converged correlation energy through the Lagrangian

    L(T, Lambda; f, V) = E(T; f, V) + Lambda . R(T; f, V),
    R(T) = G(T) - D o T        (G: right-hand side of ccsd_ref.rhs; D: denominators)

* Lambda equations  dL/dT = 0:   D o Lambda = dE/dT + (dG/dT)^T Lambda
  solved by the same Jacobi/DIIS iteration as the amplitudes;
* gradient          dE/dtheta = dL/dtheta at (T*, Lambda*), theta = (f, packed <pq||rs> blocks):
  the unrelaxed one- and two-particle density contractions.

Derivatives of E and G are taken by torch autograd on a torch restatement of
``oracle.ccsd_ref`` (same factorisation, same packed amplitude coordinates:
t1 (O x V) and the unique t2 amplitudes i<j, a<b).  Gradients are returned
for the inputs in the packed block layout of ``oracle.synth`` and symmetrised
over duplicate entries (oooo, ovov, vvvv are symmetric matrices; f is
symmetric).

Pinned by: central finite differences of the numpy oracle's converged energy
along random symmetric directions (tests/test_adjoint_oracle.py), and the
residual of the torch restatement equal to ccsd_ref.rhs.
"""
import numpy as np
import torch

from . import ccsd_ref, synth

torch.set_default_dtype(torch.float64)


def _unpack_matrix(n):
    """U (n*n x P2(n)) with U @ packed = vec(antisymmetric n x n)."""
    p, q = synth.pair_list(n)
    U = torch.zeros(n * n, len(p))
    k = torch.arange(len(p))
    U[torch.as_tensor(p * n + q), k] = 1.0
    U[torch.as_tensor(q * n + p), k] = -1.0
    return U


class TorchProblem:
    """Packed inputs as torch leaves, dense views built differentiably."""

    def __init__(self, prob, requires_grad=False):
        self.O, self.V = O, V = prob["O"], prob["V"]
        self.f = torch.tensor(prob["f"], requires_grad=requires_grad)
        self.blocks = {b: torch.tensor(prob[b], requires_grad=requires_grad) for b in synth.BLOCKS}
        self.Uo, self.Uv = _unpack_matrix(O), _unpack_matrix(V)

    def dense(self):
        O, V = self.O, self.V
        Uo, Uv = self.Uo, self.Uv
        b = self.blocks
        d = {}
        d["oooo"] = (Uo @ b["oooo"] @ Uo.T).reshape(O, O, O, O)
        d["ooov"] = (Uo @ b["ooov"]).reshape(O, O, O, V)
        d["oovv"] = (Uo @ b["oovv"] @ Uv.T).reshape(O, O, V, V)
        d["ovov"] = b["ovov"].reshape(O, V, O, V)
        d["ovvv"] = (b["ovvv"] @ Uv.T).reshape(O, V, V, V)
        d["vvvv"] = (Uv @ b["vvvv"] @ Uv.T).reshape(V, V, V, V)
        return d

    def unpack_t2(self, t2p):
        O, V = self.O, self.V
        return (self.Uo @ t2p @ self.Uv.T).reshape(O, O, V, V)

    def pack_t2(self, t2):
        O, V = self.O, self.V
        io, jo = synth.pair_list(O)
        av, bv = synth.pair_list(V)
        x = t2[torch.as_tensor(io), torch.as_tensor(jo)]          # (po, V, V)
        return x[:, torch.as_tensor(av), torch.as_tensor(bv)]      # (po, pv)


def _P(x, a, b):
    return x - x.transpose(a, b)


def rhs_torch(tp, t1, t2p):
    """Same algebra as ccsd_ref.rhs (torch); returns (R1, packed R2)."""
    O, V = tp.O, tp.V
    d = tp.dense()
    f = tp.f
    foo, fov, fvv = f[:O, :O], f[:O, O:], f[O:, O:]
    foo_off = foo - torch.diag(torch.diag(foo))
    fvv_off = fvv - torch.diag(torch.diag(fvv))
    t2 = tp.unpack_t2(t2p)
    es = torch.einsum
    oooo, ooov, oovv, ovov, ovvv, vvvv = (d[k] for k in synth.BLOCKS)
    mbej = -ovov.permute(0, 1, 3, 2)
    amef = -ovvv.permute(1, 0, 2, 3)
    abej = -ovvv.permute(2, 3, 1, 0)
    mbij = ooov.permute(2, 3, 0, 1)
    tau = t2 + es("ia,jb->ijab", t1, t1) - es("ib,ja->ijab", t1, t1)
    Fme = fov + es("nf,mnef->me", t1, oovv)
    Fmi = foo_off + es("ne,mnie->mi", t1, ooov) + 0.5 * es("inef,mnef->mi", t2, oovv) + es("me,ie->mi", Fme, t1)
    Fae = fvv_off + es("mf,amef->ae", t1, amef) - 0.5 * es("mnaf,mnef->ae", t2, oovv) - es("me,ma->ae", Fme, t1)
    W1 = oooo + _P(es("je,mnie->mnij", t1, ooov), 2, 3) + 0.5 * es("ijef,mnef->mnij", tau, oovv)
    W3 = ooov + es("if,mnfe->mnie", t1, oovv)
    W2 = mbij + _P(es("ie,mbej->mbij", t1, mbej), 2, 3) + 0.5 * es("ijef,mbef->mbij", tau, ovvv)
    Wmnej = -W3.permute(0, 1, 3, 2)
    W4 = (mbej + es("jf,mbef->mbej", t1, ovvv) - es("nb,mnej->mbej", t1, Wmnej)
          - 0.5 * es("jnfb,mnef->mbej", t2, oovv))
    ring = es("imae,mbej->ijab", t2, W4)
    R2 = (oovv + _P(_P(ring, 0, 1), 2, 3) - _P(es("ma,mbij->ijab", t1, W2), 2, 3)
          + 0.5 * es("mnab,mnij->ijab", tau, W1) + _P(es("ijae,be->ijab", t2, Fae), 2, 3)
          - _P(es("imab,mj->ijab", t2, Fmi), 0, 1) + _P(es("ie,abej->ijab", t1, abej), 0, 1)
          + 0.5 * es("ijef,abef->ijab", tau, vvvv))
    R1 = (fov + es("ie,ae->ia", t1, fvv_off) - 0.5 * es("imef,maef->ia", tau, ovvv)
          - 0.5 * es("mnae,mnie->ia", t2, W3) + es("imae,me->ia", t2, Fme) - es("ma,mi->ia", t1, Fmi)
          - es("nf,naif->ia", t1, ovov))
    return R1, tp.pack_t2(R2)


def energy_torch(tp, t1, t2p):
    O = tp.O
    oovv_p = tp.blocks["oovv"]
    t2 = t2p
    io, jo = synth.pair_list(O)
    av, bv = synth.pair_list(tp.V)
    io_, jo_, av_, bv_ = (torch.as_tensor(x) for x in (io, jo, av, bv))
    tau = t2 + t1[io_][:, av_] * t1[jo_][:, bv_] - t1[io_][:, bv_] * t1[jo_][:, av_]
    return (tp.f[:O, O:] * t1).sum() + (oovv_p * tau).sum()


def denominators(tp):
    O = tp.O
    e = torch.diagonal(tp.f)
    eo, ev = e[:O], e[O:]
    io, jo = synth.pair_list(O)
    av, bv = synth.pair_list(tp.V)
    D1 = eo[:, None] - ev[None, :]
    D2 = (eo[torch.as_tensor(io)] + eo[torch.as_tensor(jo)])[:, None] - (ev[torch.as_tensor(av)] + ev[torch.as_tensor(bv)])[None, :]
    return D1, D2


def solve_lambda(prob, t1, t2p, toler=1e-12, max_iter=500, diis_dim=8):
    """Lambda1 (O x V), Lambda2 (packed) at converged amplitudes (numpy in/out)."""
    tp = TorchProblem(prob)
    T1 = torch.tensor(t1, requires_grad=True)
    T2 = torch.tensor(t2p, requires_grad=True)
    R1, R2 = rhs_torch(tp, T1, T2)
    E = energy_torch(tp, T1, T2)
    e1, e2 = torch.autograd.grad(E, (T1, T2), retain_graph=True)
    D1, D2 = (x.detach() for x in denominators(tp))
    l1, l2 = torch.zeros_like(T1), torch.zeros_like(T2)
    diis = ccsd_ref.DIIS(diis_dim) if diis_dim >= 2 else None
    it, delta = 0, np.inf
    for it in range(1, max_iter + 1):
        g1, g2 = torch.autograd.grad((R1, R2), (T1, T2), grad_outputs=(l1, l2), retain_graph=True)
        n1 = (e1 + g1) / D1
        n2 = (e2 + g2) / D2
        delta = max(float((n1 - l1).abs().max()), float((n2 - l2).abs().max()))
        if diis is not None:
            x = diis.push_and_extrapolate(torch.cat([n1.ravel(), n2.ravel()]).numpy(),
                                          torch.cat([l1.ravel(), l2.ravel()]).numpy())
            n1 = torch.tensor(x[:n1.numel()]).reshape(n1.shape)
            n2 = torch.tensor(x[n1.numel():]).reshape(n2.shape)
        l1, l2 = n1, n2
        if delta < toler:
            break
    return l1.numpy(), l2.numpy(), dict(iterations=it, residual=delta)


def gradient(prob, t1, t2p, l1, l2):
    """dE/df (N x N) and dE/d(packed blocks), symmetrised over duplicates."""
    tp = TorchProblem(prob, requires_grad=True)
    T1 = torch.tensor(t1)
    T2 = torch.tensor(t2p)
    R1, R2 = rhs_torch(tp, T1, T2)
    D1, D2 = denominators(tp)
    L = (energy_torch(tp, T1, T2) + (torch.tensor(l1) * (R1 - D1 * T1)).sum()
         + (torch.tensor(l2) * (R2 - D2 * T2)).sum())
    leaves = [tp.f] + [tp.blocks[b] for b in synth.BLOCKS]
    grads = torch.autograd.grad(L, leaves)
    out = {"f": grads[0].numpy()}
    for b, g in zip(synth.BLOCKS, grads[1:]):
        out[b] = g.numpy()
    return symmetrise(out)


def symmetrise(g):
    g = dict(g)
    g["f"] = 0.5 * (g["f"] + g["f"].T)
    for b in ("oooo", "ovov", "vvvv"):
        g[b] = 0.5 * (g[b] + g[b].T)
    return g


def adjoint(prob, toler=1e-12):
    """Converge amplitudes (numpy oracle), Lambda and the gradient."""
    I = ccsd_ref.Ints.from_problem(prob)
    E, t1, t2, info = ccsd_ref.solve(I, toler=toler, max_iter=500, diis_dim=8)
    t2p = synth.pack_pairs(synth.pack_pairs(t2, prob["O"], 0), prob["V"], 1)
    l1, l2, linfo = solve_lambda(prob, t1, t2p)
    return E, gradient(prob, t1, t2p, l1, l2), (t1, t2p, l1, l2), linfo
