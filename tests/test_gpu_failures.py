"""Failure paths of the C ABI on the device.

The reference stops silently at its iteration cap and has no error reporting
(ccsd_memory_optimized.cpp:42, :552-556; hartree_fock_solver.cpp:1301 solves
its DIIS system unguarded).  Here each condition returns a distinct status:
  * an HBM plan over the budget is refused before any device allocation;
  * max_iter reached returns CCSD_ERR_NOT_CONVERGED with the result filled;
  * a singular DIIS system falls back to the Jacobi step and is counted;
  * ccsd_backward before a converged solve returns CCSD_ERR_STATE;
  * amplitudes that overflow inside the iteration return CCSD_ERR_NONFINITE.
The backward's two four-virtual gradient modes (slabs, whole block) agree.
"""
import numpy as np
import pytest

from oracle import synth

pytestmark = pytest.mark.gpu

OK, ERR_BUDGET, ERR_NONFINITE, ERR_NOT_CONVERGED, ERR_STATE = 0, -3, -4, -5, -8


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _scale(O, V):
    return 1.0 / np.sqrt(O * synth.p2(V))


def _free_bytes():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


@pytest.mark.parametrize("how", ["synthetic", "host"])
def test_over_budget_is_refused_before_allocating(how):
    capi = _capi()
    O, V = (40, 300) if how == "synthetic" else (6, 14)
    plan = capi.plan_bytes(O, V, "fp64")
    assert plan > 0
    # a context made and freed first, so the runtime's own set-up is not counted below
    with capi.Context.synthetic(4, 8, 1, _scale(4, 8)):
        pass
    before = _free_bytes()
    with pytest.raises(capi.CCSDError) as e:
        if how == "synthetic":
            capi.Context.synthetic(O, V, 40300, _scale(O, V), hbm_budget_bytes=plan - 1)
        else:
            prob = synth.make_problem(O, V, 5, scale=_scale(O, V))
            capi.Context.from_problem(prob, hbm_budget_bytes=plan - 1)
    assert e.value.code == ERR_BUDGET
    after = _free_bytes()
    print("plan", plan, "free before", before, "after", after, flush=True)
    assert abs(before - after) < (64 << 20)
    # the same budget plus one byte is accepted
    if how == "host":
        prob = synth.make_problem(O, V, 5, scale=_scale(O, V))
        with capi.Context.from_problem(prob, hbm_budget_bytes=plan) as ctx:
            assert ctx.solve(1e-10, 200)["converged"] == 1


def test_not_converged_returns_filled_result():
    capi = _capi()
    O, V = 8, 24
    with capi.Context.synthetic(O, V, 11, _scale(O, V)) as ctx:
        rc, r = ctx.solve_status(1e-12, 2)
        assert rc == ERR_NOT_CONVERGED
        assert r["iterations"] == 2 and r["converged"] == 0
        assert np.isfinite(r["e_corr"]) and r["e_corr"] < 0 and r["e_corr"] != r["e_mp2"]
        assert r["last_delta_e"] > 1e-12 and r["fp64_iterations"] == 2
        # ... and the energy is that of the amplitudes left in the context
        t1, t2 = ctx.get_amplitudes()
        assert abs(ctx.energy(t1, t2) - r["e_corr"]) < 1e-14
        # backward needs a converged solve
        assert ctx.backward_status() == ERR_STATE


def test_backward_before_solve_is_a_state_error():
    capi = _capi()
    with capi.Context.synthetic(4, 10, 12, _scale(4, 10)) as ctx:
        assert ctx.backward_status() == ERR_STATE
        ctx.solve(1e-10, 200)
        assert ctx.backward_status() == OK
        # new amplitudes invalidate the converged state
        t1, t2 = ctx.get_amplitudes()
        ctx.set_amplitudes(t1, t2)
        assert ctx.backward_status() == ERR_STATE


def test_singular_diis_falls_back_to_jacobi():
    """All two-electron integrals zero: every residual and error vector is
    exactly 0, so each bordered DIIS system with two or more vectors is
    singular.  FP32 start: iteration 1 (low precision) switches at once, the
    DIIS subspace is cleared, and the three FP64 sweeps hold 1, 2, 3 vectors:
    two singular solves, both taken as plain Jacobi steps."""
    capi = _capi()
    O, V = 4, 8
    prob = synth.make_problem(O, V, 3, scale=_scale(O, V))
    for k in synth.BLOCKS:
        prob[k] = np.zeros_like(prob[k])
    with capi.Context.from_problem(prob, precision="fp32", diis_dim=8) as ctx:
        r = ctx.solve(1e-10, 50)
    assert r["converged"] == 1 and r["e_corr"] == 0.0
    assert r["lowp_iterations"] == 1 and r["fp64_iterations"] == 3
    assert r["diis_fallbacks"] == 2


def test_restart_at_converged_amplitudes():
    """A second solve started at the converged amplitudes: its error vectors are
    at round-off; it stops at once on the same energy."""
    capi = _capi()
    O, V = 10, 40
    with capi.Context.synthetic(O, V, 13, _scale(O, V), diis_dim=8) as ctx:
        r0 = ctx.solve(1e-11, 200)
        t1, t2 = ctx.get_amplitudes()
        ctx.set_amplitudes(t1, t2)
        r1 = ctx.solve(1e-11, 200)
    print("restart:", r1["iterations"], "iterations,", r1["diis_fallbacks"], "fallbacks", flush=True)
    assert r1["converged"] == 1 and r1["iterations"] <= 3
    assert abs(r1["e_corr"] - r0["e_corr"]) < 1e-11


def test_overflowing_amplitudes_are_reported():
    """t1 ~ 1e100 gives a finite first energy (tau ~ 1e200) and an overflow in
    the next right-hand side (tau W_mnij ~ 1e400): CCSD_ERR_NONFINITE, with the
    iteration count filled."""
    capi = _capi()
    O, V = 6, 14
    with capi.Context.synthetic(O, V, 14, _scale(O, V), diis_dim=8) as ctx:
        t1, t2 = ctx.get_amplitudes()
        rng = np.random.default_rng(1)
        ctx.set_amplitudes(1e100 * rng.standard_normal(t1.shape), t2)
        e0 = ctx.energy(*ctx.get_amplitudes())
        assert np.isfinite(e0) and abs(e0) > 1e150
        rc, r = ctx.solve_status(1e-10, 20)
    assert rc == ERR_NONFINITE
    assert r["iterations"] == 1 and not np.isfinite(r["e_corr"])


def test_backward_four_virtual_block_modes_agree():
    """recompute_v4 = 1 (row slabs through two buffers, formed on a side
    stream), 0 (the whole block in HBM) and the default (whole here: 2.1 GB
    fits) give the same gradient; the slab size is 512 MB, so O=20/V=180
    (P2(V) = 16110) runs five slabs."""
    capi = _capi()
    O, V = 20, 180
    g = {}
    for rv4 in (1, 0, -1):
        with capi.Context.synthetic(O, V, 20180, _scale(O, V), diis_dim=8, recompute_v4=rv4) as ctx:
            ctx.solve(1e-10, 200)
            g[rv4], info = ctx.backward()
            assert info["converged"] == 1
    for k in g[1]:
        scale = max(1e-300, np.abs(g[1][k]).max())
        assert np.abs(g[0][k] - g[1][k]).max() <= 1e-13 * scale, k
        assert np.abs(g[-1][k] - g[1][k]).max() <= 1e-13 * scale, k
