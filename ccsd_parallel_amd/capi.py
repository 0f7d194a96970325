"""ctypes binding of the C ABI in ``include/ccsd_amd.h``.

The product path is the HIP library ``lib/libccsd_amd.so``; there is no CPU
fallback.  :func:`load` raises if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libccsd_amd.so")

PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

ERRORS = {
    0: "CCSD_OK", -1: "CCSD_ERR_ARG", -2: "CCSD_ERR_HIP", -3: "CCSD_ERR_BUDGET", -4: "CCSD_ERR_NONFINITE",
    -5: "CCSD_ERR_NOT_CONVERGED", -6: "CCSD_ERR_DENOMINATOR", -7: "CCSD_ERR_NO_DEVICE", -8: "CCSD_ERR_STATE",
}

EXPORTED = (
    "ccsd_default_options", "ccsd_create", "ccsd_create_synthetic", "ccsd_solve", "ccsd_iterate", "ccsd_reset",
    "ccsd_residual", "ccsd_energy", "ccsd_get_amplitudes", "ccsd_set_amplitudes", "ccsd_energy_trace",
    "ccsd_backward", "ccsd_plan_bytes", "ccsd_last_timings", "ccsd_destroy", "ccsd_strerror", "ccsd_debug_get",
    "ccsd_selftest_gemm", "ccsd_update_integrals",
    "ccsd_host_alloc", "ccsd_host_free", "ccsd_create_rhf", "ccsd_ao_to_mo", "ccsd_energy_noalloc",
    # closed-shell (spin-adapted) contexts for RHF inputs
    "ccsd_rhf_create", "ccsd_rhf_plan_bytes", "ccsd_rhf_solve", "ccsd_rhf_iterate", "ccsd_rhf_reset",
    "ccsd_rhf_residual", "ccsd_rhf_energy", "ccsd_rhf_get_amplitudes", "ccsd_rhf_set_amplitudes",
    "ccsd_rhf_last_timings", "ccsd_rhf_last_error", "ccsd_rhf_destroy",
    # include/ccsd_rhf.h (host RHF input producer)
    "ccsd_basis_size", "ccsd_rhf", "ccsd_ao_one_electron",
)

ANGSTROM_TO_BOHR = 1.889725989  # the reference's factor (initializer.cpp:56-59)


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("switch_threshold", C.c_double), ("final_fp64_sweeps", C.c_int),
        ("diis_dim", C.c_int), ("diis_start", C.c_int), ("recompute_v4", C.c_int), ("verbose", C.c_int),
        ("device", C.c_int), ("hbm_budget_bytes", C.c_size_t), ("diis_reset_on_switch", C.c_int),
        ("switch_factor", C.c_double), ("lambda_switch_factor", C.c_double), ("lambda_tol", C.c_double),
        ("strassen_min_order", C.c_int),
    ]


class Problem(C.Structure):
    _fields_ = [("n_occ", C.c_int), ("n_virt", C.c_int)] + [
        (n, C.POINTER(C.c_double)) for n in ("fock", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")]


class RhfProblem(C.Structure):
    _fields_ = [("n_basis", C.c_int), ("n_occ", C.c_int), ("n_frozen", C.c_int), ("eps", C.POINTER(C.c_double)),
                ("C", C.POINTER(C.c_double)), ("eri", C.POINTER(C.c_double))]


class Result(C.Structure):
    _fields_ = [
        ("e_corr", C.c_double), ("e_mp2", C.c_double), ("last_delta_e", C.c_double), ("max_delta_t2", C.c_double),
        ("iterations", C.c_int), ("converged", C.c_int), ("diis_fallbacks", C.c_int), ("lowp_iterations", C.c_int),
        ("fp64_iterations", C.c_int), ("seconds_total", C.c_double), ("seconds_per_iter", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GradInfo(C.Structure):
    _fields_ = [("lambda_iterations", C.c_int), ("converged", C.c_int), ("lambda_residual", C.c_double),
                ("seconds", C.c_double), ("lambda_lowp_iterations", C.c_int)]


class RhfResult(C.Structure):
    _fields_ = [("e_hf", C.c_double), ("e_nuc", C.c_double), ("commutator", C.c_double), ("n_basis", C.c_int),
                ("n_occ", C.c_int), ("iterations", C.c_int), ("converged", C.c_int)]


class CCSDError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)}")


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.ccsd_default_options.argtypes = [C.POINTER(Options)]
    lib.ccsd_create.argtypes = [C.POINTER(Problem), C.POINTER(Options), C.POINTER(vp)]
    lib.ccsd_create_synthetic.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double, C.POINTER(Options), C.POINTER(vp)]
    lib.ccsd_create_rhf.argtypes = [C.POINTER(RhfProblem), C.POINTER(Options), C.POINTER(vp)]
    lib.ccsd_ao_to_mo.argtypes = [C.c_int, dp, dp, dp, C.c_int]
    lib.ccsd_solve.argtypes = [vp, C.c_double, C.c_int, C.POINTER(Result)]
    lib.ccsd_iterate.argtypes = [vp, C.c_int, C.c_int, dp]
    lib.ccsd_reset.argtypes = [vp]
    lib.ccsd_residual.argtypes = [vp, dp, dp, dp, dp]
    lib.ccsd_energy.argtypes = [vp, dp, dp, dp]
    lib.ccsd_energy_noalloc.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
    lib.ccsd_get_amplitudes.argtypes = [vp, dp, dp]
    lib.ccsd_set_amplitudes.argtypes = [vp, dp, dp]
    lib.ccsd_energy_trace.argtypes = [vp, dp, C.c_int]
    lib.ccsd_backward.argtypes = [vp] + [dp] * 7 + [C.POINTER(GradInfo)]
    lib.ccsd_plan_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ccsd_plan_bytes.restype = C.c_size_t
    lib.ccsd_last_timings.argtypes = [vp, dp, C.c_int]
    lib.ccsd_destroy.argtypes = [vp]
    lib.ccsd_destroy.restype = None
    lib.ccsd_strerror.argtypes = [C.c_int]
    lib.ccsd_strerror.restype = C.c_char_p
    lib.ccsd_debug_get.argtypes = [vp, C.c_char_p, dp, C.c_long]
    lib.ccsd_debug_get.restype = C.c_long
    lib.ccsd_update_integrals.argtypes = [vp, C.c_int, C.c_long, C.POINTER(C.c_long), dp]
    lib.ccsd_selftest_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, C.c_double,
                                       dp, C.c_int]
    lib.ccsd_rhf_create.argtypes = [C.POINTER(RhfProblem), C.POINTER(Options), C.POINTER(vp)]
    lib.ccsd_rhf_plan_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.ccsd_rhf_plan_bytes.restype = C.c_size_t
    lib.ccsd_rhf_solve.argtypes = [vp, C.c_double, C.c_int, C.POINTER(Result)]
    lib.ccsd_rhf_iterate.argtypes = [vp, C.c_int, dp]
    lib.ccsd_rhf_reset.argtypes = [vp]
    lib.ccsd_rhf_residual.argtypes = [vp, dp, dp, dp, dp]
    lib.ccsd_rhf_energy.argtypes = [vp, dp, dp, dp]
    lib.ccsd_rhf_get_amplitudes.argtypes = [vp, dp, dp]
    lib.ccsd_rhf_set_amplitudes.argtypes = [vp, dp, dp]
    lib.ccsd_rhf_last_timings.argtypes = [vp, dp, C.c_int]
    lib.ccsd_rhf_last_error.argtypes = [vp]
    lib.ccsd_rhf_last_error.restype = C.c_char_p
    lib.ccsd_rhf_destroy.argtypes = [vp]
    lib.ccsd_rhf_destroy.restype = None
    lib.ccsd_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    lib.ccsd_host_free.argtypes = [vp]
    ip = C.POINTER(C.c_int)
    lib.ccsd_basis_size.argtypes = [C.c_int, ip, C.c_char_p]
    lib.ccsd_rhf.argtypes = [C.c_int, ip, dp, C.c_char_p, C.c_int, C.c_double, C.POINTER(RhfResult), dp, dp, dp, dp,
                             C.c_int]
    lib.ccsd_ao_one_electron.argtypes = [C.c_int, ip, dp, C.c_char_p, dp, dp, dp, C.c_int]
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, what):
    if rc != 0:
        raise CCSDError(rc, what)


def default_options(**kw):
    lib = load()
    o = Options()
    lib.ccsd_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(o, k, v)
    return o


def p2(n):
    return n * (n - 1) // 2


class Context:
    """One engine context (device memory owned by the library)."""

    def __init__(self, handle, O, V):
        self.h = handle
        self.O, self.V = O, V
        self.po, self.pv, self.ov = p2(O), p2(V), O * V
        self.lib = load()

    @classmethod
    def from_problem(cls, prob, **opts):
        lib = load()
        O, V = prob["O"], prob["V"]
        keep = {k: np.ascontiguousarray(prob[k], dtype=np.float64) for k in
                ("f", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")}
        p = Problem(O, V, _ptr(keep["f"]), _ptr(keep["oooo"]), _ptr(keep["ooov"]), _ptr(keep["oovv"]),
                    _ptr(keep["ovov"]), _ptr(keep["ovvv"]), _ptr(keep["vvvv"]))
        o = default_options(**opts)
        h = C.c_void_p()
        _check(lib.ccsd_create(C.byref(p), C.byref(o), C.byref(h)), "ccsd_create")
        return cls(h, O, V)

    @classmethod
    def from_rhf(cls, eps, eri, n_occ, mo_coeff=None, n_frozen=0, **opts):
        """Context from closed-shell data: orbital energies eps[n], spatial
        integrals eri[n,n,n,n] ((pq|rs), AO basis with coefficients
        mo_coeff[mu, p], or MO basis when mo_coeff is None); transform and
        spin-orbital packing run on the device (ccsd_create_rhf)."""
        lib = load()
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        n = eps.shape[0]
        if eri.shape != (n, n, n, n):
            raise ValueError("eri must be n x n x n x n")
        Cm = None if mo_coeff is None else np.ascontiguousarray(mo_coeff, dtype=np.float64)
        p = RhfProblem(n, int(n_occ), int(n_frozen), _ptr(eps), _ptr(Cm) if Cm is not None else None, _ptr(eri))
        o = default_options(**opts)
        h = C.c_void_p()
        _check(lib.ccsd_create_rhf(C.byref(p), C.byref(o), C.byref(h)), "ccsd_create_rhf")
        return cls(h, 2 * (int(n_occ) - int(n_frozen)), 2 * (n - int(n_occ)))

    @classmethod
    def synthetic(cls, O, V, seed, scale, **opts):
        lib = load()
        o = default_options(**opts)
        h = C.c_void_p()
        _check(lib.ccsd_create_synthetic(O, V, seed, scale, C.byref(o), C.byref(h)), "ccsd_create_synthetic")
        return cls(h, O, V)

    def close(self):
        if self.h:
            self.lib.ccsd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, tol=1e-10, max_iter=1000, allow_unconverged=False):
        r = Result()
        rc = self.lib.ccsd_solve(self.h, tol, max_iter, C.byref(r))
        if rc != 0 and not (allow_unconverged and rc == -5):
            raise CCSDError(rc, "ccsd_solve")
        return r.as_dict()

    def solve_status(self, tol=1e-10, max_iter=1000):
        """(status code, result dict) of ccsd_solve without raising; the result
        is filled on CCSD_ERR_NOT_CONVERGED and on a non-finite iteration."""
        r = Result()
        rc = self.lib.ccsd_solve(self.h, tol, max_iter, C.byref(r))
        return rc, r.as_dict()

    def backward_status(self):
        """Status code of ccsd_backward into throw-away arrays (no raising)."""
        g = {k: np.zeros(sh) for k, sh in self.gradient_shapes().items()}
        info = GradInfo()
        return self.lib.ccsd_backward(self.h, *[_ptr(g[k]) for k in ("f", "oooo", "ooov", "oovv", "ovov", "ovvv",
                                                                      "vvvv")], C.byref(info))

    def iterate(self, n, precision="fp64"):
        e = np.zeros(max(n, 1))
        _check(self.lib.ccsd_iterate(self.h, n, PRECISIONS.get(precision, precision), _ptr(e)), "ccsd_iterate")
        return e[:n]

    def reset(self):
        _check(self.lib.ccsd_reset(self.h), "ccsd_reset")

    def residual(self, t1, t2p):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64)
        r1 = np.zeros(self.ov)
        r2 = np.zeros(self.po * self.pv)
        _check(self.lib.ccsd_residual(self.h, _ptr(t1), _ptr(t2p), _ptr(r1), _ptr(r2)), "ccsd_residual")
        return r1.reshape(self.O, self.V), r2.reshape(self.po, self.pv)

    def energy(self, t1, t2p):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64)
        e = C.c_double()
        _check(self.lib.ccsd_energy(self.h, _ptr(t1), _ptr(t2p), C.byref(e)), "ccsd_energy")
        return e.value

    def get_amplitudes(self):
        t1 = np.zeros(self.ov)
        t2 = np.zeros(self.po * self.pv)
        _check(self.lib.ccsd_get_amplitudes(self.h, _ptr(t1), _ptr(t2)), "ccsd_get_amplitudes")
        return t1.reshape(self.O, self.V), t2.reshape(self.po, self.pv)

    def set_amplitudes(self, t1, t2p):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64)
        _check(self.lib.ccsd_set_amplitudes(self.h, _ptr(t1), _ptr(t2p)), "ccsd_set_amplitudes")

    def energy_trace(self):
        n = self.lib.ccsd_energy_trace(self.h, None, 0)
        out = np.zeros(max(n, 1))
        self.lib.ccsd_energy_trace(self.h, _ptr(out), n)
        return out[:n]

    def timings(self):
        ms = np.zeros(5)
        self.lib.ccsd_last_timings(self.h, _ptr(ms), 5)
        return dict(zip(("pp_ladder", "ring", "wmbej_quad", "wmbij", "other"), ms.tolist()))

    def setup_timings(self):
        """Device times (ms) of a ccsd_create_rhf context's set-up: the AO->MO
        transform and the spin-orbital packing."""
        ms = np.zeros(7)
        self.lib.ccsd_last_timings(self.h, _ptr(ms), 7)
        return {"ao_to_mo": float(ms[5]), "pack": float(ms[6])}

    def debug(self, name):
        n = self.lib.ccsd_debug_get(self.h, name.encode(), None, 0)
        if n < 0:
            raise CCSDError(int(n), "ccsd_debug_get")
        out = np.zeros(n)
        rc = self.lib.ccsd_debug_get(self.h, name.encode(), _ptr(out), n)
        if rc < 0:
            raise CCSDError(int(rc), "ccsd_debug_get")
        return out

    BLOCK_IDS = {"oooo": 0, "ooov": 1, "oovv": 2, "ovov": 3, "ovvv": 4, "vvvv": 5, "f": 6}

    def update_integrals(self, block, index, delta):
        """Add delta to entries (flat indices) of one input array on the device."""
        idx = np.ascontiguousarray(index, dtype=np.int64)
        dl = np.ascontiguousarray(delta, dtype=np.float64)
        _check(self.lib.ccsd_update_integrals(self.h, self.BLOCK_IDS[block], len(idx),
                                              idx.ctypes.data_as(C.POINTER(C.c_long)), _ptr(dl)),
               "ccsd_update_integrals")

    def gradient_shapes(self):
        N = self.O + self.V
        return dict(f=(N, N), oooo=(self.po, self.po), ooov=(self.po, self.ov), oovv=(self.po, self.pv),
                    ovov=(self.ov, self.ov), ovvv=(self.ov, self.pv), vvvv=(self.pv, self.pv))

    def backward(self, allow_unconverged=False, out=None):
        """dE/df and dE/d<pq||rs> (dict of arrays in the packed layouts) and the
        Lambda statistics.  ``out`` may supply the destination arrays (C-contiguous
        float64 of gradient_shapes(), e.g. page-locked for DMA-rate copies)."""
        shapes = self.gradient_shapes()
        if out is None:
            g = {k: np.zeros(sh) for k, sh in shapes.items()}
        else:
            g = out
            for k, sh in shapes.items():
                a = g[k]
                if a.shape != sh or a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
                    raise ValueError(f"out[{k!r}] must be a C-contiguous float64 array of shape {sh}")
        info = GradInfo()
        rc = self.lib.ccsd_backward(self.h, *[_ptr(g[k]) for k in ("f", "oooo", "ooov", "oovv", "ovov", "ovvv",
                                                                    "vvvv")], C.byref(info))
        if rc != 0 and not (allow_unconverged and rc == -5):
            raise CCSDError(rc, "ccsd_backward")
        return g, {k: getattr(info, k) for k, _ in info._fields_}


def plan_bytes(O, V, precision="fp64"):
    return load().ccsd_plan_bytes(O, V, PRECISIONS.get(precision, precision))


def strerror(code):
    return load().ccsd_strerror(code).decode()


def selftest_gemm(A, B, C_in, alpha=1.0, beta=0.0, ta=False, tb=False, precision="fp64", tile=0,
                  elementwise=False, strassen=False):
    """C = alpha op(A) op(B) + beta C through the device GEMM engine; tile > 0
    forces that tile configuration (index + 1) instead of the automatic choice;
    elementwise forces single-element global loads where the FP64 engine would
    load element pairs; strassen runs one level of Strassen's method (FP64 or
    FP32, square, even size, beta 0)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    Cm = np.ascontiguousarray(C_in, dtype=np.float64).copy()
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    _check(load().ccsd_selftest_gemm(M, N, K, int(ta), int(tb), alpha, _ptr(A), _ptr(B), beta, _ptr(Cm),
                                     PRECISIONS.get(precision, precision) | (int(tile) << 8)
                                     | (int(bool(elementwise)) << 16) | (int(bool(strassen)) << 17)),
           "ccsd_selftest_gemm")
    return Cm


def ao_to_mo(Cmat, ao, device=0):
    """(pq|rs) = sum C[mu,p] C[nu,q] C[la,r] C[si,s] (mu nu|la si) on the device."""
    Cm = np.ascontiguousarray(Cmat, dtype=np.float64)
    ao = np.ascontiguousarray(ao, dtype=np.float64)
    n = Cm.shape[0]
    mo = np.empty_like(ao)
    _check(load().ccsd_ao_to_mo(n, _ptr(Cm), _ptr(ao), _ptr(mo), device), "ccsd_ao_to_mo")
    return mo


def _molecule(Z, R, angstrom):
    Zc = np.ascontiguousarray(Z, dtype=np.int32)
    Rc = np.ascontiguousarray(R, dtype=np.float64).reshape(len(Zc), 3)
    if angstrom:
        Rc = Rc * ANGSTROM_TO_BOHR
    return Zc, np.ascontiguousarray(Rc)


def basis_size(Z, basis):
    Zc = np.ascontiguousarray(Z, dtype=np.int32)
    n = load().ccsd_basis_size(len(Zc), Zc.ctypes.data_as(C.POINTER(C.c_int)), basis.encode())
    _check(min(n, 0), "ccsd_basis_size")
    return n


def rhf(Z, R, basis, n_electrons=None, tol=1e-10, angstrom=False, want_ao_eri=False, allow_unconverged=False):
    """Closed-shell RHF (host code in the library): returns a dict with e_hf,
    e_nuc, eps, C (C[mu, p]), mo_eri ((pq|rs), MO basis) and, if asked, ao_eri."""
    Zc, Rc = _molecule(Z, R, angstrom)
    n = basis_size(Zc, basis)
    ne = int(Zc.sum()) if n_electrons is None else int(n_electrons)
    res = RhfResult()
    Cm = np.zeros((n, n))
    eps = np.zeros(n)
    mo = np.zeros((n, n, n, n))
    ao = np.zeros((n, n, n, n)) if want_ao_eri else None
    rc = load().ccsd_rhf(len(Zc), Zc.ctypes.data_as(C.POINTER(C.c_int)), _ptr(Rc), basis.encode(), ne, tol,
                         C.byref(res), _ptr(Cm), _ptr(eps), _ptr(mo), _ptr(ao) if ao is not None else None, n)
    if not (rc == 0 or (allow_unconverged and rc == -5)):
        _check(rc, "ccsd_rhf")
    out = {k: getattr(res, k) for k, _ in RhfResult._fields_}
    out.update(C=Cm, eps=eps, mo_eri=mo, nocc=res.n_occ)
    if ao is not None:
        out["ao_eri"] = ao
    return out


def ao_one_electron(Z, R, basis, angstrom=False):
    """AO overlap, kinetic and nuclear-attraction matrices."""
    Zc, Rc = _molecule(Z, R, angstrom)
    n = basis_size(Zc, basis)
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    _check(load().ccsd_ao_one_electron(len(Zc), Zc.ctypes.data_as(C.POINTER(C.c_int)), _ptr(Rc), basis.encode(),
                                       _ptr(S), _ptr(T), _ptr(V), n), "ccsd_ao_one_electron")
    return S, T, V


class RhfContext:
    """Closed-shell (spin-adapted) CCSD context for RHF inputs (ccsd_rhf_*):
    amplitudes t1[o, v] and t2[o, o, v, v] (the alpha-beta block)."""

    def __init__(self, eps, eri, n_occ, mo_coeff=None, n_frozen=0, **opts):
        self.lib = lib = load()
        self.h = None
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        n = eps.shape[0]
        if eri.shape != (n, n, n, n):
            raise ValueError("eri must be n x n x n x n")
        Cm = None if mo_coeff is None else np.ascontiguousarray(mo_coeff, dtype=np.float64)
        p = RhfProblem(n, int(n_occ), int(n_frozen), _ptr(eps), _ptr(Cm) if Cm is not None else None, _ptr(eri))
        o = default_options(**opts)
        h = C.c_void_p()
        _check(lib.ccsd_rhf_create(C.byref(p), C.byref(o), C.byref(h)), "ccsd_rhf_create")
        self.h = h
        self.o, self.v = int(n_occ) - int(n_frozen), n - int(n_occ)

    def close(self):
        if self.h:
            self.lib.ccsd_rhf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, tol=1e-10, max_iter=1000, allow_unconverged=False):
        r = Result()
        rc = self.lib.ccsd_rhf_solve(self.h, tol, max_iter, C.byref(r))
        if rc != 0 and not (allow_unconverged and rc == -5):
            raise CCSDError(rc, "ccsd_rhf_solve")
        return r.as_dict()

    def iterate(self, n):
        e = np.zeros(max(n, 1))
        _check(self.lib.ccsd_rhf_iterate(self.h, n, _ptr(e)), "ccsd_rhf_iterate")
        return e[:n]

    def reset(self):
        _check(self.lib.ccsd_rhf_reset(self.h), "ccsd_rhf_reset")

    def residual(self, t1, t2):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2 = np.ascontiguousarray(t2, dtype=np.float64)
        r1, r2 = np.zeros_like(t1), np.zeros_like(t2)
        _check(self.lib.ccsd_rhf_residual(self.h, _ptr(t1), _ptr(t2), _ptr(r1), _ptr(r2)), "ccsd_rhf_residual")
        return r1, r2

    def energy(self, t1, t2):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2 = np.ascontiguousarray(t2, dtype=np.float64)
        e = C.c_double()
        _check(self.lib.ccsd_rhf_energy(self.h, _ptr(t1), _ptr(t2), C.byref(e)), "ccsd_rhf_energy")
        return e.value

    def get_amplitudes(self):
        t1 = np.zeros((self.o, self.v))
        t2 = np.zeros((self.o, self.o, self.v, self.v))
        _check(self.lib.ccsd_rhf_get_amplitudes(self.h, _ptr(t1), _ptr(t2)), "ccsd_rhf_get_amplitudes")
        return t1, t2

    def set_amplitudes(self, t1, t2):
        t1 = np.ascontiguousarray(t1, dtype=np.float64)
        t2 = np.ascontiguousarray(t2, dtype=np.float64)
        _check(self.lib.ccsd_rhf_set_amplitudes(self.h, _ptr(t1), _ptr(t2)), "ccsd_rhf_set_amplitudes")

    def timings(self):
        ms = np.zeros(1)
        self.lib.ccsd_rhf_last_timings(self.h, _ptr(ms), 1)
        return {"residual": float(ms[0])}


def rhf_plan_bytes(n_basis, n_occ, n_frozen=0):
    return load().ccsd_rhf_plan_bytes(n_basis, n_occ, n_frozen)


class PinnedArrays:
    """Page-locked float64 host arrays (ccsd_host_alloc), e.g. destinations for
    Context.backward(out=...); freed by close() or on exit of a with-block."""

    def __init__(self, shapes):
        self.arrays, self._ptrs = {}, []
        lib = load()
        try:
            for k, sh in shapes.items():
                n = int(np.prod(sh))
                p = C.c_void_p()
                _check(lib.ccsd_host_alloc(n * 8, C.byref(p)), "ccsd_host_alloc")
                self._ptrs.append(p)
                self.arrays[k] = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(n,)).reshape(sh)
        except Exception:
            self.close()
            raise

    def close(self):
        self.arrays = {}
        for p in self._ptrs:
            load().ccsd_host_free(p)
        self._ptrs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
