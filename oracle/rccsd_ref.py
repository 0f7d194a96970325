"""Closed-shell (spin-adapted) CCSD for RHF inputs, in spatial orbitals
(test infrastructure only).

The reference's hot path (``CCSD_Memory_optimized::CCSD``,
ccsd_memory_optimized.cpp:16-567) iterates spin-orbital amplitudes of an RHF
reference, skipping the spin-forbidden blocks (spin conservation,
``(a+i+b+j)`` even, :483-485).  For a closed-shell reference every spin block
is fixed by one spatial amplitude set:

    t1[i,a]     = t_{i alpha}^{a alpha}           (= t_{i beta}^{a beta})
    t2[i,j,a,b] = t_{i alpha, j beta}^{a alpha, b beta}

with t2[i,j,a,b] = t2[j,i,b,a]; the same-spin amplitudes are
t_{i alpha j alpha}^{a alpha b alpha} = t2[i,j,a,b] - t2[i,j,b,a].  The
equations below are the spin summation of the spin-orbital ones for a
canonical RHF Fock matrix (diagonal orbital energies; the reference takes only
``eigvals``), written with chemist-notation integrals (pq|rs), in the
closed-shell intermediate form of Scuseria, Janssen & Schaefer (J. Chem. Phys.
89, 7382 (1988)) with the F/L/W intermediates of Stanton & Gauss (J. Chem.
Phys. 94, 4334 (1991)) summed over spin:

* Fov, Foo, Fvv, Loo, Lvv  one-particle intermediates
* Woooo, Wvoov, Wvovo      two-particle intermediates
* pp ladder                sum_cd (ac|bd) tau_ij^cd, and its t1 part through
                           X[k,a,i,j] = sum_cd (kd|ac) tau_ij^cd (o^3 v^3)
* T2 update                every term as tmp + P tmp, P swapping (i,a)<->(j,b)

R1, R2 are the right-hand sides of D t = R with D = e_i - e_a (and
e_i + e_j - e_a - e_b), the convention of ``ccsd_ref.rhs``.

Pinned by: equality with the spin-orbital oracle ``ccsd_ref`` on RHF
integrals mapped by ``rhf_spin`` -- residuals at random amplitudes (all spin
blocks: alpha-beta equals R2, alpha-alpha equals R2 - R2 with a<->b), the
energy functional, and the converged CCSD energy (tests/test_rccsd_oracle.py);
the closed-shell MP2 formula for the first energy.
"""
import numpy as np


def _es(*a):
    return np.einsum(*a, optimize=True)


class RInts:
    """Chemist-notation spatial blocks of an RHF problem (occupied first)."""

    def __init__(self, eps, mo, nocc, nfrozen=0):
        eps = np.asarray(eps, dtype=np.float64)
        n = len(eps)
        oc = np.arange(nfrozen, nocc)
        vi = np.arange(nocc, n)
        self.o, self.v = len(oc), len(vi)
        self.eo, self.ev = eps[oc].copy(), eps[vi].copy()
        g = lambda a, b, c, d: np.ascontiguousarray(mo[np.ix_(a, b, c, d)])
        self.ovov = g(oc, vi, oc, vi)   # (ia|jb)
        self.oovv = g(oc, oc, vi, vi)   # (ij|ab)
        self.ovoo = g(oc, vi, oc, oc)   # (ia|jk)
        self.oooo = g(oc, oc, oc, oc)   # (ij|kl)
        self.ovvv = g(oc, vi, vi, vi)   # (ia|bc)
        self.vvvv = g(vi, vi, vi, vi)   # (ab|cd)
        self.Dia = self.eo[:, None] - self.ev[None, :]
        self.Dijab = self.Dia[:, None, :, None] + self.Dia[None, :, None, :]
        # 2 (ia|jb) - (ib|ja)
        self.L = 2.0 * self.ovov - self.ovov.transpose(0, 3, 2, 1)


def sym(x):
    """x + x with (i,a) <-> (j,b) swapped, for x[i,j,a,b]."""
    return x + x.transpose(1, 0, 3, 2)


def build_tau(t1, t2):
    return t2 + _es("ia,jb->ijab", t1, t1)


def energy(I, t1, t2):
    """E = sum [2 (ia|jb) - (ib|ja)] (t_ij^ab + t_i^a t_j^b)."""
    return float(_es("iajb,ijab->", I.L, build_tau(t1, t2)))


def intermediates(I, t1, t2):
    tau = build_tau(t1, t2)
    L = I.L                                            # L[k,c,l,d] = 2(kc|ld) - (kd|lc)
    Fov = _es("kcld,ld->kc", L, t1)
    Foo = _es("kcld,ilcd->ki", L, tau)
    Fvv = -_es("kcld,klad->ac", L, tau)
    # (ki|lc) and (kc|li) as ovoo slices: ovoo[l,c,k,i] = (lc|ki)
    Loo = Foo + 2.0 * _es("lcki,lc->ki", I.ovoo, t1) - _es("kcli,lc->ki", I.ovoo, t1)
    Lvv = Fvv + 2.0 * _es("kdac,kd->ac", I.ovvv, t1) - _es("kcad,kd->ac", I.ovvv, t1)
    Woooo = (I.oooo.transpose(0, 2, 1, 3)                     # (ki|lj) as [k,l,i,j]
             + _es("lcki,jc->klij", I.ovoo, t1) + _es("kclj,ic->klij", I.ovoo, t1)
             + _es("kcld,ijcd->klij", I.ovov, tau))
    ovvo = I.ovov.transpose(0, 1, 3, 2)                       # (kc|ai) = (kc|ia) as [k,c,a,i]
    Wvoov = (ovvo.transpose(2, 0, 3, 1)                       # (ai|kc) as [a,k,i,c]
             + _es("kcad,id->akic", I.ovvv, t1) - _es("kcli,la->akic", I.ovoo, t1)
             - 0.5 * _es("ldkc,ilda->akic", I.ovov, t2) - 0.5 * _es("lckd,ilad->akic", I.ovov, t2)
             - _es("ldkc,id,la->akic", I.ovov, t1, t1) + _es("ldkc,ilad->akic", I.ovov, t2))
    Wvovo = (I.oovv.transpose(2, 0, 3, 1)                     # (ki|ac) as [a,k,c,i]
             + _es("kdac,id->akci", I.ovvv, t1) - _es("lcki,la->akci", I.ovoo, t1)
             - 0.5 * _es("lckd,ilda->akci", I.ovov, t2) - _es("lckd,id,la->akci", I.ovov, t1, t1))
    return dict(tau=tau, Fov=Fov, Foo=Foo, Fvv=Fvv, Loo=Loo, Lvv=Lvv, Woooo=Woooo, Wvoov=Wvoov, Wvovo=Wvovo)


def rhs(I, t1, t2, parts=False):
    """Right-hand sides R1 [o,v] and R2 [o,o,v,v] of D t = R."""
    X = intermediates(I, t1, t2)
    tau, Fov = X["tau"], X["Fov"]
    r1 = dict(
        fvv=_es("ac,ic->ia", X["Fvv"], t1),
        foo=-_es("ki,ka->ia", X["Foo"], t1),
        fov=2.0 * _es("kc,kica->ia", Fov, t2) - _es("kc,ikca->ia", Fov, t2) + _es("kc,ic,ka->ia", Fov, t1, t1),
        ovvo=2.0 * _es("kcia,kc->ia", I.ovov, t1) - _es("kiac,kc->ia", I.oovv, t1),
        ovvv=2.0 * _es("kdac,ikcd->ia", I.ovvv, tau) - _es("kcad,ikcd->ia", I.ovvv, tau),
        ovoo=-2.0 * _es("lcki,klac->ia", I.ovoo, tau) + _es("kcli,klac->ia", I.ovoo, tau),
    )
    # bare pp ladder and its t1 part: X[k,a,i,j] = sum_cd (kd|ac) tau_ij^cd
    Xk = _es("kdac,ijcd->kaij", I.ovvv, tau)
    tmp2 = I.ovvv.transpose(1, 3, 0, 2) - _es("kibc,ka->abic", I.oovv, t1)   # (ai|bc) - t_ka (ki|bc) as [a,b,i,c]
    tmp3 = _es("kcai,jc->akij", I.ovov.transpose(0, 1, 3, 2), t1) + I.ovoo.transpose(1, 3, 0, 2)
    r2 = dict(
        ints=I.ovov.transpose(0, 2, 1, 3).copy(),                           # (ia|jb)
        t1a=sym(_es("abic,jc->ijab", tmp2, t1)),
        t1b=-sym(_es("akij,kb->ijab", tmp3, t1)),
        hh=_es("klij,klab->ijab", X["Woooo"], tau),
        pp=_es("acbd,ijcd->ijab", I.vvvv, tau),
        pp_t1=-sym(_es("kaij,kb->ijab", Xk, t1)),
        lvv=sym(_es("ac,ijcb->ijab", X["Lvv"], t2)),
        loo=-sym(_es("ki,kjab->ijab", X["Loo"], t2)),
        ring1=sym(2.0 * _es("akic,kjcb->ijab", X["Wvoov"], t2) - _es("akci,kjcb->ijab", X["Wvovo"], t2)),
        ring2=-sym(_es("akic,kjbc->ijab", X["Wvoov"], t2)),
        ring3=-sym(_es("bkci,kjac->ijab", X["Wvovo"], t2)),
    )
    R1 = sum(r1.values())
    R2 = sum(r2.values())
    if parts:
        return R1, R2, r1, r2, X
    return R1, R2


def mp2_guess(I):
    return np.zeros((I.o, I.v)), I.ovov.transpose(0, 2, 1, 3) / I.Dijab


def solve(I, toler=1e-10, max_iter=500, diis_dim=8, record=None):
    """Jacobi iterations with FP64 DIIS on (t1, t2), |dE| < toler."""
    from .ccsd_ref import DIIS
    t1, t2 = mp2_guess(I)
    e_old = energy(I, t1, t2)
    diis = DIIS(diis_dim) if diis_dim >= 2 else None
    it = 0
    for it in range(1, max_iter + 1):
        R1, R2 = rhs(I, t1, t2)
        n1, n2 = R1 / I.Dia, R2 / I.Dijab
        if diis is not None:
            x = diis.push_and_extrapolate(np.concatenate([n1.ravel(), n2.ravel()]),
                                          np.concatenate([t1.ravel(), t2.ravel()]))
            n1 = x[:t1.size].reshape(t1.shape)
            n2 = x[t1.size:].reshape(t2.shape)
        t1, t2 = n1, n2
        e = energy(I, t1, t2)
        if record is not None:
            record.append(e)
        if abs(e - e_old) < toler:
            return e, t1, t2, it
        e_old = e
    return e_old, t1, t2, it


def spin_blocks(t1, t2):
    """Spin-orbital amplitudes (occupied P: spatial P//2, spin P%2; same for
    virtuals) of the closed-shell amplitudes, dense [O,V] and [O,O,V,V]."""
    o, v = t1.shape
    O, V = 2 * o, 2 * v
    T1 = np.zeros((O, V))
    T2 = np.zeros((O, O, V, V))
    same = t2 - t2.transpose(0, 1, 3, 2)
    for s in (0, 1):
        T1[s::2, s::2] = t1
        T2[s::2, s::2, s::2, s::2] = same
        T2[s::2, 1 - s::2, s::2, 1 - s::2] = t2
        T2[s::2, 1 - s::2, 1 - s::2, s::2] = -t2.transpose(0, 1, 3, 2)
    return T1, T2
