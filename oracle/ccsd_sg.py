"""Textbook spin-orbital CCSD (Stanton & Gauss, J. Chem. Phys. 94, 4334 (1991)).

Test infrastructure only.  An independent formulation used to pin
``oracle.ccsd_ref`` (the restatement of the reference's factorisation): both must
give the same T1/T2 right-hand sides at random amplitudes.  It works on the full
dense <pq||rs> tensor over all spin orbitals (slices taken here, not the block
views of ccsd_ref), with the tilde-tau F intermediates and the full W_abef, and
a general (non-canonical) Fock matrix.
"""
import numpy as np


def _P(x, a, b):
    return x - np.swapaxes(x, a, b)


def rhs_sg(f, g, O, t1, t2):
    o, v = slice(0, O), slice(O, None)
    n = f.shape[0]
    fd = f - np.diag(np.diag(f))
    foo, fov, fvv = fd[o, o], f[o, v], fd[v, v]
    tau_t = t2 + 0.5 * (np.einsum("ia,jb->ijab", t1, t1) - np.einsum("ib,ja->ijab", t1, t1))
    tau = t2 + np.einsum("ia,jb->ijab", t1, t1) - np.einsum("ib,ja->ijab", t1, t1)

    Fae = (fvv - 0.5 * np.einsum("me,ma->ae", fov, t1)
           + np.einsum("mf,mafe->ae", t1, g[o, v, v, v])
           - 0.5 * np.einsum("mnaf,mnef->ae", tau_t, g[o, o, v, v]))
    Fmi = (foo + 0.5 * np.einsum("ie,me->mi", t1, fov)
           + np.einsum("ne,mnie->mi", t1, g[o, o, o, v])
           + 0.5 * np.einsum("inef,mnef->mi", tau_t, g[o, o, v, v]))
    Fme = fov + np.einsum("nf,mnef->me", t1, g[o, o, v, v])
    Wmnij = (g[o, o, o, o] + _P(np.einsum("je,mnie->mnij", t1, g[o, o, o, v]), 2, 3)
             + 0.25 * np.einsum("ijef,mnef->mnij", tau, g[o, o, v, v]))
    Wabef = (g[v, v, v, v] - _P(np.einsum("mb,amef->abef", t1, g[v, o, v, v]), 0, 1)
             + 0.25 * np.einsum("mnab,mnef->abef", tau, g[o, o, v, v]))
    Wmbej = (g[o, v, v, o] + np.einsum("jf,mbef->mbej", t1, g[o, v, v, v])
             - np.einsum("nb,mnej->mbej", t1, g[o, o, v, o])
             - np.einsum("jnfb,mnef->mbej", 0.5 * t2 + np.einsum("jf,nb->jnfb", t1, t1), g[o, o, v, v]))

    R1 = (f[o, v].copy() + np.einsum("ie,ae->ia", t1, Fae) - np.einsum("ma,mi->ia", t1, Fmi)
          + np.einsum("imae,me->ia", t2, Fme) - np.einsum("nf,naif->ia", t1, g[o, v, o, v])
          - 0.5 * np.einsum("imef,maef->ia", t2, g[o, v, v, v])
          - 0.5 * np.einsum("mnae,nmei->ia", t2, g[o, o, v, o]))

    R2 = g[o, o, v, v].copy()
    R2 = R2 + _P(np.einsum("ijae,be->ijab", t2, Fae - 0.5 * np.einsum("mb,me->be", t1, Fme)), 2, 3)
    R2 = R2 - _P(np.einsum("imab,mj->ijab", t2, Fmi + 0.5 * np.einsum("je,me->mj", t1, Fme)), 0, 1)
    R2 = R2 + 0.5 * np.einsum("mnab,mnij->ijab", tau, Wmnij)
    R2 = R2 + 0.5 * np.einsum("ijef,abef->ijab", tau, Wabef)
    ring = (np.einsum("imae,mbej->ijab", t2, Wmbej)
            - np.einsum("ie,ma,mbej->ijab", t1, t1, g[o, v, v, o]))
    R2 = R2 + _P(_P(ring, 0, 1), 2, 3)
    R2 = R2 + _P(np.einsum("ie,abej->ijab", t1, g[v, v, v, o]), 0, 1)
    R2 = R2 - _P(np.einsum("ma,mbij->ijab", t1, g[o, v, o, o]), 2, 3)
    return R1, R2


def energy_sg(f, g, O, t1, t2):
    o, v = slice(0, O), slice(O, None)
    return float(np.einsum("ia,ia->", f[o, v], t1) + 0.25 * np.einsum("ijab,ijab->", g[o, o, v, v], t2)
                 + 0.5 * np.einsum("ijab,ia,jb->", g[o, o, v, v], t1, t1))


def solve_sg(f, g, O, toler=1e-12, max_iter=500):
    """Plain Jacobi CCSD from the MP2 guess (diagonal of f in the denominators)."""
    e = np.diag(f)
    eo, ev = e[:O], e[O:]
    Dia = eo[:, None] - ev[None, :]
    Dijab = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    t1 = np.zeros((O, len(ev)))
    t2 = g[:O, :O, O:, O:] / Dijab
    E = energy_sg(f, g, O, t1, t2)
    for _ in range(max_iter):
        R1, R2 = rhs_sg(f, g, O, t1, t2)
        t1, t2 = R1 / Dia, R2 / Dijab
        En = energy_sg(f, g, O, t1, t2)
        if abs(En - E) < toler:
            return En, t1, t2
        E = En
    return E, t1, t2
