"""ccsd_energy_noalloc: the allocation-free energy entry point, through the C ABI.

Checked against the existing entry point ccsd_energy (context-owned buffers,
host amplitudes) and the oracle's energy (ccsd_ref.energy) on the same
amplitudes, to 1e-13 relative (FP64, different summation order), on two calls
in a row on one device, the second with a smaller input (a stale partial from
the first call would show).  Operand buffers are device memory made by torch,
the caller's allocator.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import ccsd_ref, synth

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_BUDGET = -1, -3


def _case(O, V, seed):
    prob = synth.make_problem(O, V, seed, scale=1.0 / np.sqrt(O * synth.p2(V)))
    rng = np.random.default_rng(seed)
    # a non-canonical f_ia so the f t1 term is exercised
    N = O + V
    a = 0.01 * rng.standard_normal((N, N))
    prob["f"] = prob["f"] + np.block([[np.zeros((O, O)), a[:O, O:]], [a[:O, O:].T, np.zeros((V, V))]])
    t1 = 0.05 * rng.standard_normal((O, V))
    t2 = 0.05 * rng.standard_normal((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return prob, t1, t2


def test_energy_noalloc_matches_ccsd_energy():
    import torch
    from ccsd_parallel_amd import capi
    lib = capi.load()
    stream = torch.cuda.current_stream()
    for O, V, seed in ((12, 40, 71), (6, 14, 72)):
        print("case: ccsd_energy_noalloc O", O, "V", V, "seed", seed, flush=True)
        prob, t1, t2 = _case(O, V, seed)
        t2p = synth.pack_pairs(synth.pack_pairs(t2, O, 0), V, 1)
        with capi.Context.from_problem(prob) as ctx:
            e_ctx = ctx.energy(t1, t2p)
        e_ora = ccsd_ref.energy(ccsd_ref.Ints.from_problem(prob), t1, t2)
        dev = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")  # noqa: E731
        fov, oovv, d1, d2 = dev(prob["f"][:O, O:]), dev(prob["oovv"]), dev(t1), dev(t2p)
        e = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        rc = lib.ccsd_energy_noalloc(O, V, fov.data_ptr(), oovv.data_ptr(), d1.data_ptr(), d2.data_ptr(),
                                     e.data_ptr(), C.c_void_p(stream.cuda_stream))
        assert rc == 0
        torch.cuda.synchronize()
        got = float(e.cpu()[0])
        print("E noalloc %.15f ctx %.15f oracle %.15f" % (got, e_ctx, e_ora), flush=True)
        assert abs(got - e_ctx) <= 1e-13 * max(1.0, abs(e_ctx))
        assert abs(got - e_ora) <= 1e-13 * max(1.0, abs(e_ora))


def test_energy_noalloc_rejects_what_it_cannot_hold():
    import torch
    from ccsd_parallel_amd import capi
    lib = capi.load()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = buf.data_ptr()
    assert lib.ccsd_energy_noalloc(129, 4, p, p, p, p, p, None) == ERR_BUDGET
    assert lib.ccsd_energy_noalloc(1, 4, p, p, p, p, p, None) == ERR_ARG
    assert lib.ccsd_energy_noalloc(4, 4, p, None, p, p, p, None) == ERR_ARG
    torch.cuda.synchronize()
