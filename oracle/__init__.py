"""CPU FP64 oracle for the CCSD amplitude-update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product library imports, links or
executes anything under this package; only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules
-------
synth        counter-based synthetic spin-orbital integrals (SplitMix64 + Box-Muller)
ccsd_ref     spin-orbital CCSD restating CCSD_Memory_optimized::CCSD's factorisation
             (ccsd_memory_optimized.cpp:16-567) in dense numpy
ccsd_sg      independent textbook Stanton-Gauss CCSD (cross-check of ccsd_ref)
fci2         brute-force 2-electron FCI (CCSD is exact for two electrons)
adjoint_ref  CCSD Lambda / gradient oracle via torch autograd on ccsd_ref (FP64, CPU)
rhf_spin     RHF spatial integrals -> packed spin-orbital problem (frozen core included)
ccsd_cpu.cpp the reference's loop shapes in C++/OpenMP (one dot product per output,
             threads over output rows), built by cpu_build.py: the CPU baseline
             bench.py times beside the GPU

Every function states what pins it; see the module docstrings.
"""
