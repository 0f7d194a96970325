"""Spin-orbital CCSD in the factorisation of CCSD_Memory_optimized (test infrastructure only).

Restates, in dense numpy FP64, the amplitude update of the reference's hot path
``CCSD_Memory_optimized::CCSD`` (ccsd_memory_optimized.cpp:16-567) for a general
spin-orbital problem (Fock matrix ``f`` over occupied+virtual spin orbitals and
antisymmetrised integral blocks).  The reference stores the same quantities in
RHF compact storage; the algebra is identical:

* tau            Fill_tau                      :1669-1850
* energy         Calc_Energy                   :592-648   (E = 1/4 <ij||ab> tau)
* F_me (D1)      part-1 "F1 term"              :1351, Fill_F2 :1596-1602
* F_mi (D2)      :708/793/912/964, :1309/1329, Fill_F2 :1604-1629 (+ F_me t_i^e)
* F_ae (D3)      :1076/1251, :1356/1376, Fill_F3 :1631-1667   (- F_me t_m^a)
* W_mnij (W_1)   :659-1029, with 1/2 tau <mn||ef> (:752, 770, 941, 1014)
* W_mnie (W_3)   :720/735/829/844/924/997
* W_mbij (W_2)   :1047-1295, with 1/2 <mb||ef> tau (:1097, 1161, 1202, 1264)
* W_mbej (W_4)   Fill_W4_MPI :4781-5089 (the uninitialised ``temp`` at :4837 is taken as 0)
* ring           Fill_W5 :6201-6882 (P(ab)) and Fill_t2_new :2069-2070 (P(ij))
* T2 residual    Fill_t2_new :2001-2406 (bare pp ladder 1/2 <ab||cd> tau, :2093)
* T1 residual    Fill_t2_new :2410-2457 plus part-1 1/2 <ka||cd> tau (:1400, 1409)
* update         Map_T_new :2460-2696 and ``T_1 = T_1_new / DEN_AI`` :535
* driver         MP2 guess :473-490, loop :513-557, |dE| < toler, Itermax 1000

Generalisations beyond the reference (exactly zero for canonical HF / the
synthetic diagonal Fock): the off-diagonal Fock terms (1-delta) f_ae t_i^e and
f_ia t_i^a in the energy, which the reference drops (:2446-2447, :646).

Pinned by: the independent textbook Stanton-Gauss formulation in ``ccsd_sg``
(residuals at random amplitudes), 2-electron FCI (``fci2``), the MP2 identity of
the first energy, and orbital-rotation invariance of the converged energy.
"""
import numpy as np

from .synth import dense_blocks, p2, pack_pairs, unpack_pairs


def _es(*a):
    return np.einsum(*a, optimize=True)



def P(x, a, b):
    """Antisymmetriser P(pq) X = X - X(p<->q) over axes a, b."""
    return x - np.swapaxes(x, a, b)


class Ints:
    """Dense integral views needed by the residual (all <pq||rs> classes)."""

    def __init__(self, f, d, O, V):
        self.O, self.V = O, V
        self.f = np.asarray(f, dtype=np.float64)
        self.foo = self.f[:O, :O]
        self.fov = self.f[:O, O:]
        self.fvv = self.f[O:, O:]
        self.eo = np.diag(self.foo).copy()
        self.ev = np.diag(self.fvv).copy()
        self.foo_off = self.foo - np.diag(self.eo)
        self.fvv_off = self.fvv - np.diag(self.ev)
        self.oooo = d["oooo"]
        self.ooov = d["ooov"]                                   # <mn||ie>
        self.oovv = d["oovv"]                                   # <mn||ef>
        self.ovov = d["ovov"]                                   # <ia||jb>
        self.ovvv = d["ovvv"]                                   # <ma||ef>
        self.vvvv = d["vvvv"]                                   # <ab||ef>
        self.mbej = -self.ovov.transpose(0, 1, 3, 2)            # <mb||ej>
        self.amef = -self.ovvv.transpose(1, 0, 2, 3)            # <am||ef>
        self.abej = -self.ovvv.transpose(2, 3, 1, 0)            # <ab||ej>
        self.mbij = self.ooov.transpose(2, 3, 0, 1)             # <mb||ij>
        self.mnej = -self.ooov.transpose(0, 1, 3, 2)            # <mn||ej>
        self.Dia = self.eo[:, None] - self.ev[None, :]
        self.Dijab = (self.eo[:, None, None, None] + self.eo[None, :, None, None]
                      - self.ev[None, None, :, None] - self.ev[None, None, None, :])

    @classmethod
    def from_problem(cls, prob):
        return cls(prob["f"], dense_blocks(prob), prob["O"], prob["V"])


def build_tau(t1, t2):
    """tau_ij^ab = t_ij^ab + t_i^a t_j^b - t_i^b t_j^a  (Fill_tau :1669-1850)."""
    return t2 + _es("ia,jb->ijab", t1, t1) - _es("ib,ja->ijab", t1, t1)


def energy(I, t1, t2):
    """E = f_ia t_i^a + 1/4 <ij||ab> tau_ij^ab  (Calc_Energy :592-648; f term is 0 for HF)."""
    tau = build_tau(t1, t2)
    return float(_es("ia,ia->", I.fov, t1) + 0.25 * _es("ijab,ijab->", I.oovv, tau))


def intermediates(I, t1, t2):
    """F and W intermediates in the reference's factorisation."""
    tau = build_tau(t1, t2)
    # F_me: part-1 "F1 term" (:1351) on top of FS_AI (Fill_F1 :1578-1589)
    Fme = I.fov + _es("nf,mnef->me", t1, I.oovv)
    # F_mi (D2): t1 <mn||ie> (:708..), 1/2 t2 <mn||ef> (:1309, :1329), + F_me t_i^e (Fill_F2)
    Fmi = (I.foo_off + _es("ne,mnie->mi", t1, I.ooov)
           + 0.5 * _es("inef,mnef->mi", t2, I.oovv)
           + _es("me,ie->mi", Fme, t1))
    # F_ae (D3): t1 <am||ef> (:1076, :1251), -1/2 t2 <mn||ef> (:1356, :1376), - F_me t_m^a (Fill_F3)
    Fae = (I.fvv_off + _es("mf,amef->ae", t1, I.amef)
           - 0.5 * _es("mnaf,mnef->ae", t2, I.oovv)
           - _es("me,ma->ae", Fme, t1))
    # W_1 = W_mnij with 1/2 tau (absorbs W_abef's tau tau part; pp ladder then uses bare integrals)
    W1 = (I.oooo + P(_es("je,mnie->mnij", t1, I.ooov), 2, 3)
          + 0.5 * _es("ijef,mnef->mnij", tau, I.oovv))
    # W_3 = W_mnie
    W3 = I.ooov + _es("if,mnfe->mnie", t1, I.oovv)
    # W_2 = W_mbij with 1/2 <mb||ef> tau (carries W_abef's t1 part)
    W2 = (I.mbij + P(_es("ie,mbej->mbij", t1, I.mbej), 2, 3)
          + 0.5 * _es("ijef,mbef->mbij", tau, I.ovvv))
    # W_4 = W_mbej (Fill_W4_MPI): <mb||ej> + t_j^f <mb||ef> - t_n^b W_mnej + 1/2 t2 <mn||ef>
    Wmnej = -W3.transpose(0, 1, 3, 2)
    W4 = (I.mbej + _es("jf,mbef->mbej", t1, I.ovvv)
          - _es("nb,mnej->mbej", t1, Wmnej)
          - 0.5 * _es("jnfb,mnef->mbej", t2, I.oovv))
    return dict(tau=tau, Fme=Fme, Fmi=Fmi, Fae=Fae, W1=W1, W2=W2, W3=W3, W4=W4)


def rhs(I, t1, t2, parts=False):
    """Right-hand sides of D t1 = R1 and D t2 = R2 (reference factorisation)."""
    X = intermediates(I, t1, t2)
    tau = X["tau"]
    ring = _es("imae,mbej->ijab", t2, X["W4"])
    terms2 = dict(
        ints=I.oovv.copy(),                                                       # <ab||ij>
        ring=P(P(ring, 0, 1), 2, 3),                                              # Fill_W5 + :2069
        w2=-P(_es("ma,mbij->ijab", t1, X["W2"]), 2, 3),                     # -P(ab) W_2 t1
        hh=0.5 * _es("mnab,mnij->ijab", tau, X["W1"]),                      # 1/2 W_1 tau
        fae=P(_es("ijae,be->ijab", t2, X["Fae"]), 2, 3),                    # P(ab) t2 F_ae
        fmi=-P(_es("imab,mj->ijab", t2, X["Fmi"]), 0, 1),                   # -P(ij) t2 F_mi
        abej=P(_es("ie,abej->ijab", t1, I.abej), 0, 1),                     # P(ij) <ab||cj> t
        pp=0.5 * _es("ijef,abef->ijab", tau, I.vvvv),                       # bare pp ladder
    )
    terms1 = dict(
        f=I.fov.copy() + _es("ie,ae->ia", t1, I.fvv_off),
        tau_ovvv=-0.5 * _es("imef,maef->ia", tau, I.ovvv),                  # :1400, :1409
        w3=-0.5 * _es("mnae,mnie->ia", t2, X["W3"]),                        # :2424-2436
        fme=_es("imae,me->ia", t2, X["Fme"]),                               # :2441
        fmi=-_es("ma,mi->ia", t1, X["Fmi"]),                                # :2444
        ovov=-_es("nf,naif->ia", t1, I.ovov),                               # integ10 :2450
    )
    R2 = sum(terms2.values())
    R1 = sum(terms1.values())
    if parts:
        return R1, R2, terms1, terms2, X
    return R1, R2


def mp2_guess(I):
    """T1 = 0, T2 = <ab||ij> / D  (ccsd_memory_optimized.cpp:473-490)."""
    return np.zeros((I.O, I.V)), I.oovv / I.Dijab


def jacobi_step(I, t1, t2):
    R1, R2 = rhs(I, t1, t2)
    return R1 / I.Dia, R2 / I.Dijab


class DIIS:
    """Pulay DIIS on (T1, packed T2) with the reference HF convention
    (hartree_fock_solver.cpp:1252-1326: B_ij = <e_i|e_j>, bordered by -1, rhs -1).

    Vectors: T1 (O*V) then the unique T2 amplitudes (i<j, a<b).  Error vector
    e_k = t_k(new) - t_k(old).  A solve with |c|_inf > 1e6 or a singular B falls
    back to the plain Jacobi step (reported as a fallback, not an error)."""

    def __init__(self, dim):
        self.dim = dim
        self.vecs, self.errs = [], []
        self.fallbacks = 0

    def push_and_extrapolate(self, x_new, x_old):
        self.vecs.append(x_new.copy())
        self.errs.append(x_new - x_old)
        if len(self.vecs) > self.dim:
            self.vecs.pop(0)
            self.errs.pop(0)
        m = len(self.vecs)
        if m < 2:
            return x_new
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for i in range(m):
            for j in range(m):
                B[i, j] = float(np.dot(self.errs[i], self.errs[j]))
        rhs_ = np.zeros(m + 1)
        rhs_[m] = -1.0
        c = diis_solve(B, rhs_)
        if c is None:
            self.fallbacks += 1
            return x_new
        return sum(ci * v for ci, v in zip(c[:m], self.vecs))


def diis_solve(B, rhs_):
    """Solve the bordered DIIS system after diagonal scaling; None if singular."""
    m = B.shape[0] - 1
    d = np.abs(np.diag(B)[:m])
    if not np.all(np.isfinite(B)) or d.max() <= 0.0:
        return None
    s = np.ones(m + 1)
    s[:m] = 1.0 / np.sqrt(np.where(d > 0, d, 1.0))
    Bs = B * s[:, None] * s[None, :]
    try:
        c = np.linalg.solve(Bs, rhs_ * s)
    except np.linalg.LinAlgError:
        return None
    c = c * s
    if not np.all(np.isfinite(c)) or np.max(np.abs(c[:m])) > 1e6:
        return None
    return c


def flat(t1, t2, O, V):
    return np.concatenate([t1.ravel(), pack_pairs(pack_pairs(t2, O, 0), V, 1).ravel()])


def unflat(x, O, V):
    t1 = x[:O * V].reshape(O, V)
    t2 = unpack_pairs(unpack_pairs(x[O * V:].reshape(p2(O), p2(V)), V, 1), O, 0)
    return t1, t2


def solve(I, toler=1e-10, max_iter=1000, diis_dim=0, t_init=None, record=None):
    """The reference driver loop (:473-557) with optional DIIS.

    Returns (E_corr, t1, t2, info).  ``info['energies']`` holds E after the MP2
    guess followed by one entry per iteration, as the reference prints them
    ("Energi: E Steg: n", :506, :548)."""
    O, V = I.O, I.V
    t1, t2 = mp2_guess(I) if t_init is None else t_init
    E_new = energy(I, t1, t2)
    energies = [E_new]
    it = 1
    diis = DIIS(diis_dim) if diis_dim and diis_dim >= 2 else None
    converged = False
    while True:
        E_old = E_new
        n1, n2 = jacobi_step(I, t1, t2)
        if diis is not None:
            x = diis.push_and_extrapolate(flat(n1, n2, O, V), flat(t1, t2, O, V))
            n1, n2 = unflat(x, O, V)
        t1, t2 = n1, n2
        E_new = energy(I, t1, t2)
        energies.append(E_new)
        if record is not None:
            record.append((t1.copy(), t2.copy()))
        it += 1
        if abs(E_new - E_old) < toler:
            converged = True
            break
        if it >= max_iter or not np.isfinite(E_new):
            break
    info = dict(energies=energies, iterations=it - 1, converged=converged,
                diis_fallbacks=(diis.fallbacks if diis else 0))
    return E_new, t1, t2, info


def ccsd_energy(prob, **kw):
    I = Ints.from_problem(prob)
    return solve(I, **kw)
