"""MI355X-native CCSD amplitude-update engine (HIP, gfx950).

Host side of the drop-in for the reference's ``CCSD_Memory_optimized::CCSD``
(ccsd_memory_optimized.cpp:16-567): the C ABI lives in ``lib/libccsd_amd.so``
(built from ``csrc/``), the C++ drop-in class in the same library, and this
package exposes the ctypes binding used by tests and the benchmark.
"""
from .capi import (CCSDError, Context, PRECISIONS, ao_one_electron, basis_size, default_options, load,  # noqa: F401
                   plan_bytes, rhf, selftest_gemm, strerror)

__all__ = ["CCSDError", "Context", "PRECISIONS", "ao_one_electron", "basis_size", "default_options", "load",
           "plan_bytes", "rhf", "selftest_gemm", "strerror"]
