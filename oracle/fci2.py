"""Brute-force full CI for two electrons (test infrastructure only).

For two electrons CCSD is exact, so the CCSD correlation energy equals
E_FCI - E_ref.  The Hamiltonian is rebuilt from the Fock matrix and <pq||rs>:
h_pq = f_pq - sum_k <pk||qk> (k occupied), so that its Fock operator with respect
to the reference |0 1> is f.  Matrix elements between two-electron determinants
|pq> (p<q): <pq|H|rs> = h_pr d_qs + h_qs d_pr - h_ps d_qr - h_qr d_ps + <pq||rs>.
"""
import numpy as np


def fci_two_electron(f, g):
    n = f.shape[0]
    O = 2
    h = f - np.einsum("pkqk->pq", g[:, :O, :, :O])
    dets = [(p, q) for q in range(n) for p in range(q)]
    nd = len(dets)
    H = np.zeros((nd, nd))
    for x, (p, q) in enumerate(dets):
        for y, (r, s) in enumerate(dets):
            val = g[p, q, r, s]
            if q == s:
                val += h[p, r]
            if p == r:
                val += h[q, s]
            if q == r:
                val -= h[p, s]
            if p == s:
                val -= h[q, r]
            H[x, y] = val
    w = np.linalg.eigvalsh(H)
    e_ref = h[0, 0] + h[1, 1] + g[0, 1, 0, 1]
    return float(w[0] - e_ref), float(w[0]), float(e_ref)
