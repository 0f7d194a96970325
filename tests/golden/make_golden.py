"""Write tests/golden/ccsd_small.json from the FP64 oracle.

The reference has no fixtures for this path (no tests at all), so these are
regression fixtures of the pinned oracle: per-iteration Jacobi energies (the
reference's own update rule, ccsd_memory_optimized.cpp:513-557), the DIIS
converged correlation energy, and a generator checksum.
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ccsd_ref, synth  # noqa: E402

CASES = [(4, 8, 101, 0.05), (6, 12, 102, None), (8, 16, 20180, None)]


def main():
    out = {"generator": "SplitMix64 + Box-Muller (oracle/synth.py)", "cases": []}
    for O, V, seed, scale in CASES:
        if scale is None:
            scale = 1.0 / np.sqrt(O * synth.p2(V))
        prob = synth.make_problem(O, V, seed, scale=scale)
        I = ccsd_ref.Ints.from_problem(prob)
        _, _, _, info = ccsd_ref.solve(I, toler=0.0, max_iter=12, diis_dim=0)
        E, _, _, info2 = ccsd_ref.solve(I, toler=1e-12, max_iter=300, diis_dim=8)
        blk = synth.packed_block("vvvv", O, V, seed, scale, rows=(0, 1))
        out["cases"].append(dict(O=O, V=V, seed=seed, scale=scale, jacobi_energies=info["energies"], e_corr=E,
                                 diis_iterations=info2["iterations"], vvvv_row0_sum=float(blk.sum())))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ccsd_small.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
