"""Build the in-tree HIP library ``ccsd_parallel_amd/lib/libccsd_amd.so`` for gfx950.

Every source under ``csrc/`` is compiled with hipcc (device code for gfx950
only) and linked into one shared library exporting the C ABI declared in
``include/ccsd_amd.h`` and ``include/ccsd_rhf.h`` and the drop-in class of
``include/ccsd_memory_optimized.h``; the host programs under ``apps/`` (the
INCAR front-end) are linked against it with g++.
Run ``python -m ccsd_parallel_amd.build`` or call :func:`build`.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
APPS = os.path.join(HERE, "apps")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libccsd_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no hand-written or compiler-formed packed FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return src, obj, p.returncode, err


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, err in ex.map(_compile, srcs):
            if err and verbose:
                print(f"[{os.path.basename(src)}]\n{err}", file=sys.stderr)
            if rc != 0:
                failed.append(src)
            objs.append(obj)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(os.path.basename(f) for f in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_apps(verbose)
    return LIB


def build_apps(verbose=True):
    """Host programs under ``apps/`` (g++, linked against the library, rpath to it)."""
    for f in sorted(os.listdir(APPS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(LIBDIR, f[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), os.path.join(APPS, f),
               "-L" + LIBDIR, "-lccsd_amd", "-Wl,-rpath,$ORIGIN", "-o", exe]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"g++ failed for {f}:\n" + p.stderr)
        if verbose:
            print("built", exe)


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build()
