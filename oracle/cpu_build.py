"""Build recipe for the oracle's C++ restatement (``oracle/ccsd_cpu.cpp``).

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


if __name__ == "__main__":
    build()
