"""Build recipe and binding for the oracle's C++ restatement (``oracle/ccsd_cpu.cpp``).

Test and baseline infrastructure only: ``bench.py``'s CPU baseline and the CPU
tests load ``oracle/lib/libccsd_cpu.so``; the product never does.
Run ``python -m oracle.cpu_build`` (``__graft_entry__.build()`` calls it).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ccsd_cpu.cpp")
LIB = os.path.join(HERE, "lib", "libccsd_cpu.so")


def build(verbose=True):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    # AVX2/FMA, not -march=native: the library is built here and run on other x86-64 hosts
    cmd = ["g++", "-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=c++17", SRC, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("g++ failed for ccsd_cpu.cpp:\n" + p.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def load():
    import ctypes as C
    if not os.path.exists(LIB):
        raise RuntimeError(f"oracle CPU library not built: {LIB} (run python -m oracle.cpu_build)")
    lib = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    lib.cpu_dot_rows.argtypes = [dp, dp, dp, C.c_long, C.c_long, C.c_long, C.c_int]
    lib.cpu_dot_rows.restype = C.c_int
    lib.cpu_max_threads.restype = C.c_int
    lib.cpu_ccsd_residual.argtypes = [C.c_int, C.c_int] + [dp] * 11 + [C.c_int]
    lib.cpu_ccsd_residual.restype = C.c_int
    return lib


def dot_rows(A, B, threads=0):
    """out[r, q] = sum_k A[r, k] B[q, k] through the C++ dot-product loops."""
    import ctypes as C

    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.empty((A.shape[0], B.shape[0]))
    ptr = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = load().cpu_dot_rows(ptr(A), ptr(B), ptr(out), A.shape[0], B.shape[0], A.shape[1], int(threads))
    if rc != 0:
        raise RuntimeError("cpu_dot_rows failed")
    return out


def residual(prob, t1, t2p, threads=0):
    """R1 (O x V) and packed R2 (P2(O) x P2(V)) of the C++ packed FP64 residual
    (cpu_ccsd_residual) for a problem dict in the packed layouts of
    ``oracle.synth`` (keys f, oooo, ooov, oovv, ovov, ovvv, vvvv)."""
    import ctypes as C

    import numpy as np
    O, V = prob["O"], prob["V"]
    po, pv = O * (O - 1) // 2, V * (V - 1) // 2
    arr = {k: np.ascontiguousarray(prob[k], dtype=np.float64)
           for k in ("f", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")}
    t1 = np.ascontiguousarray(t1, dtype=np.float64)
    t2p = np.ascontiguousarray(t2p, dtype=np.float64)
    if t1.size != O * V or t2p.size != po * pv:
        raise ValueError("amplitude shapes")
    r1 = np.zeros((O, V))
    r2 = np.zeros((po, pv))
    ptr = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = load().cpu_ccsd_residual(O, V, *[ptr(arr[k]) for k in ("f", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")],
                                  ptr(t1), ptr(t2p), ptr(r1), ptr(r2), int(threads))
    if rc != 0:
        raise RuntimeError("cpu_ccsd_residual failed")
    return r1, r2


def solve(prob, toler=1e-10, max_iter=100, diis_dim=0, threads=0):
    """The reference driver loop (ccsd_memory_optimized.cpp:473-557: MP2 guess,
    T <- R(T)/D, stop at |dE| < toler) with optional DIIS
    (``oracle.ccsd_ref.DIIS``), on the packed C++ residual: the same iteration
    as ``ccsd_ref.solve`` at sizes the dense oracle cannot hold.  Returns
    (E, t1, packed t2, info) with info['energies'] (MP2 guess first)."""
    import numpy as np

    from . import ccsd_ref, synth
    O, V = prob["O"], prob["V"]
    f = prob["f"]
    eo, ev = np.diag(f)[:O], np.diag(f)[O:]
    ip, iq = synth.pair_list(O)
    ap, aq = synth.pair_list(V)
    D2 = (eo[ip] + eo[iq])[:, None] - (ev[ap] + ev[aq])[None, :]
    D1 = eo[:, None] - ev[None, :]
    oovv = prob["oovv"]
    fov = f[:O, O:]

    def energy(t1, t2):
        with np.errstate(over="ignore", invalid="ignore"):  # a divergent run reports a non-finite E
            tau = t2 + (t1[ip][:, ap] * t1[iq][:, aq] - t1[ip][:, aq] * t1[iq][:, ap])
            return float(np.sum(fov * t1) + np.sum(oovv * tau))

    t1, t2 = np.zeros((O, V)), oovv / D2
    E = energy(t1, t2)
    energies = [E]
    diis = ccsd_ref.DIIS(diis_dim) if diis_dim and diis_dim >= 2 else None
    converged = False
    for _ in range(max_iter):
        r1, r2 = residual(prob, t1, t2, threads)
        n1, n2 = r1 / D1, r2 / D2
        if diis is not None:
            x = diis.push_and_extrapolate(np.concatenate([n1.ravel(), n2.ravel()]),
                                          np.concatenate([t1.ravel(), t2.ravel()]))
            n1, n2 = x[:O * V].reshape(O, V), x[O * V:].reshape(n2.shape)
        t1, t2 = n1, n2
        E_old, E = E, energy(t1, t2)
        energies.append(E)
        if not np.isfinite(E):
            break
        if abs(E - E_old) < toler:
            converged = True
            break
    return E, t1, t2, dict(energies=energies, iterations=len(energies) - 1, converged=converged)


if __name__ == "__main__":
    build()
