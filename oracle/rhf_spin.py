"""RHF spatial MO integrals -> spin-orbital <PQ||RS> (test infrastructure only).

Restates the reference's on-the-fly antisymmetrised slices
(Fill_integ*_2D, ccsd_memory_optimized.cpp:3574-4779 and :6045-6199):
spin orbital P has spatial orbital P//2 and spin P%2 within the occupied and
within the virtual range (:483-485); <PQ||RS> = <PQ|RS> - <PQ|SR> with
<PQ|RS> = (pr|qs) when spins match pairwise (e.g. :3728, :3870).

Frozen core (Prepear_AOs(nr_freeze), :2698): the lowest nfrozen occupied
spatial orbitals are left out, so occupied spin orbital P is spatial orbital
nfrozen + P//2.

Pinned by: spin-adapted identities in the tests (closed-shell MP2 from the
spatial formula sum (ia|jb)[2(ia|jb) - (ib|ja)]/D equals the spin-orbital MP2,
and the CCSD energy of a 2-electron RHF system equals its FCI); the frozen-core
problem equals the full problem of the integrals with the core rows and
columns removed.
"""
import numpy as np

from . import synth


def spin_orbital_tensor(eps, mo, nocc, nfrozen=0):
    """Full <PQ||RS> and Fock over all correlated spin orbitals (occupied first)."""
    n = len(eps)
    nv = n - nocc
    O, V = 2 * (nocc - nfrozen), 2 * nv
    spatial = np.array([nfrozen + P // 2 for P in range(O)] + [nocc + A // 2 for A in range(V)])
    spin = np.array([P % 2 for P in range(O)] + [A % 2 for A in range(V)])
    ch = mo[np.ix_(spatial, spatial, spatial, spatial)]  # (pr|qs) needs reordering below
    # <PQ|RS> = (PR|QS) delta(sP,sR) delta(sQ,sS)
    phys = ch.transpose(0, 2, 1, 3)  # phys[P,Q,R,S] = ch[P,R,Q,S]
    d = (spin[:, None] == spin[None, :]).astype(float)
    phys = phys * d[:, None, :, None] * d[None, :, None, :]
    g = phys - phys.transpose(0, 1, 3, 2)
    f = np.diag(np.asarray(eps)[spatial])
    return f, g, O, V


def problem_from_rhf(eps, mo, nocc, nfrozen=0):
    f, g, O, V = spin_orbital_tensor(eps, mo, nocc, nfrozen)
    return dict(O=O, V=V, f=f, **synth.pack_dense(synth.blocks_from_full(g, O), O, V))


def random_rhf(n, nocc, seed, scale=0.05):
    """Random 8-fold symmetric spatial (pq|rs) and sorted orbital energies."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n, n, n)) * scale
    x = x + x.transpose(1, 0, 2, 3)
    x = x + x.transpose(0, 1, 3, 2)
    x = x + x.transpose(2, 3, 0, 1)
    eo = np.sort(rng.uniform(-1.5, -0.3, nocc))
    ev = np.sort(rng.uniform(0.2, 3.0, n - nocc))
    return np.concatenate([eo, ev]), x / 8.0


def mp2_closed_shell(eps, mo, nocc):
    """Spatial-orbital closed-shell MP2: sum (ia|jb)[2(ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b)."""
    o, v = slice(0, nocc), slice(nocc, None)
    iajb = mo[o, v, o, v]
    eo, ev = eps[o], eps[v]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(iajb * (2 * iajb - iajb.transpose(0, 3, 2, 1)) / D))
