"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/cpp/sanitize_main.cpp drives the host routines beside the device path on
small inputs: the RHF producer (ccsd_parallel_amd/csrc/rhf.cpp), the closed-
shell contraction planner and its term sequence through the host executor
(csrc/rplan.hpp via tests/cpp/rccsd_cpu.cpp), single planned contractions with
the planner's refusals, and the packed CPU residual (oracle/ccsd_cpu.cpp).  The
build uses g++ with -fsanitize=address,undefined and aborts on the first
report (leak checking on).  A planted heap overflow in a separate three-line
program checks that the sanitizer is active with these flags.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-fopenmp"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="2")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def _build(out, sources):
    cmd = ["g++"] + FLAGS + sources + ["-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]


def test_sanitizer_reports_a_planted_overflow(tmp_path):
    src = tmp_path / "overflow.cpp"
    src.write_text("#include <vector>\nint main(int argc, char**) {\n"
                   "  std::vector<int> v(4);\n  return v.data()[4 + argc];\n}\n")
    exe = str(tmp_path / "overflow")
    _build(exe, [str(src)])
    p = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=120)
    assert p.returncode != 0 and "AddressSanitizer" in p.stderr


@pytest.mark.timeout(900)
def test_host_code_is_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_main")
    _build(exe, [os.path.join(ROOT, "tests", "cpp", "sanitize_main.cpp"),
                 os.path.join(ROOT, "tests", "cpp", "rccsd_cpu.cpp"),
                 os.path.join(ROOT, "oracle", "ccsd_cpu.cpp"),
                 os.path.join(ROOT, "ccsd_parallel_amd", "csrc", "rhf.cpp")])
    p = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert p.returncode == 0, (p.stdout[-2000:], p.stderr[-6000:])
    assert "sanitized host run: ok" in p.stdout
    assert "runtime error" not in p.stderr and "Sanitizer" not in p.stderr, p.stderr[-4000:]
