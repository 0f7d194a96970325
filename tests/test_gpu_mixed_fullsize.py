"""Mixed precision at the O=40 / V=300 workload (seed 40300), against FP64.

* Solve: FP32-mixed and BF16x3-mixed (ladder / ring / W_mbej / W_mbij GEMMs
  in low precision, then at least 3 FP64 sweeps) converged to |dE| < 1e-7
  land within 1e-7 Eh of the FP64 solution converged to 1e-11, and their
  amplitudes stay as close to it as the FP64 path's own amplitudes stopped at
  the same tolerance (max|dT2| and max|dT1| within 2x of the FP64 run's, plus
  a 1e-8 floor).
* Forward + adjoint: BF16x3-mixed forward (to 1e-7) and BF16x3-mixed Lambda
  iterations (BF16x3 sweeps, then FP64 passes each followed by a BF16x3
  correction solve until an FP64 pass meets the Lambda tolerance) give
  dE/df and dE/d<pq||rs> within 1e-6 relative (Frobenius norm
  over every block) of the FP64 forward (to 1e-10) and FP64 adjoint.  The
  22 GB of gradients per run land in page-locked host arrays and are compared
  in row chunks.
"""
import numpy as np
import pytest

from oracle import synth

pytestmark = pytest.mark.gpu

O, V, SEED = 40, 300, 40300
SCALE = 1.0 / np.sqrt(O * synth.p2(V))


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _solve(capi, precision, tol):
    print("case: O", O, "V", V, "seed", SEED, precision, "tol", tol, flush=True)
    with capi.Context.synthetic(O, V, SEED, SCALE, precision=precision, diis_dim=8, final_fp64_sweeps=3) as ctx:
        r = ctx.solve(tol, 100)
        t1, t2 = ctx.get_amplitudes()
    return r, t1, t2


@pytest.mark.timeout(900)
def test_mixed_solve_o40_v300_matches_fp64():
    capi = _capi()
    r64, t1_ref, t2_ref = _solve(capi, "fp64", 1e-11)
    assert r64["converged"] == 1
    r7, t1_7, t2_7 = _solve(capi, "fp64", 1e-7)
    own_t2 = np.abs(t2_7 - t2_ref).max()
    own_t1 = np.abs(t1_7 - t1_ref).max()
    print("FP64 @1e-7: |dE| %.2e max|dT2| %.2e max|dT1| %.2e" % (abs(r7["e_corr"] - r64["e_corr"]), own_t2, own_t1))
    del t1_7, t2_7
    for prec in ("fp32", "bf16x3"):
        r, t1, t2 = _solve(capi, prec, 1e-7)
        de = abs(r["e_corr"] - r64["e_corr"])
        dt2 = np.abs(t2 - t2_ref).max()
        dt1 = np.abs(t1 - t1_ref).max()
        print("%s: iterations %d low + %d FP64, |dE| %.2e, max|dT2| %.2e, max|dT1| %.2e" % (
            prec, r["lowp_iterations"], r["fp64_iterations"], de, dt2, dt1), flush=True)
        assert r["converged"] == 1
        assert r["lowp_iterations"] > 0 and r["fp64_iterations"] >= 3
        assert de <= 1e-7
        assert dt2 <= 2.0 * own_t2 + 1e-8
        assert dt1 <= 2.0 * own_t1 + 1e-8


@pytest.mark.timeout(900)
def test_mixed_forward_adjoint_o40_v300_matches_fp64():
    capi = _capi()
    bufs = {}
    try:
        for prec, tol in (("fp64", 1e-10), ("bf16x3", 1e-7)):
            print("case: forward + adjoint O", O, "V", V, prec, "tol", tol, flush=True)
            with capi.Context.synthetic(O, V, SEED, SCALE, precision=prec, diis_dim=8) as ctx:
                bufs[prec] = capi.PinnedArrays(ctx.gradient_shapes())
                r = ctx.solve(tol, 100)
                _, info = ctx.backward(out=bufs[prec].arrays)
            print(prec, "forward", r["lowp_iterations"], "+", r["fp64_iterations"], "Lambda",
                  info["lambda_lowp_iterations"], "+", info["lambda_iterations"] - info["lambda_lowp_iterations"],
                  flush=True)
            assert info["converged"] == 1
            if prec == "bf16x3":
                assert r["lowp_iterations"] > 0 and info["lambda_lowp_iterations"] > 0
                assert info["lambda_iterations"] - info["lambda_lowp_iterations"] >= 1
        a, b = bufs["fp64"].arrays, bufs["bf16x3"].arrays
        num = den = 0.0
        worst = {}
        for k in a:
            nk = dk = 0.0
            for r0 in range(0, a[k].shape[0], 1024):
                x, y = a[k][r0:r0 + 1024], b[k][r0:r0 + 1024]
                nk += float(np.sum((y - x) ** 2))
                dk += float(np.sum(x * x))
            worst[k] = np.sqrt(nk / dk) if dk > 0 else 0.0
            num += nk
            den += dk
        rel = np.sqrt(num / den)
        print("gradient relative error %.3e; per block %s" % (rel, {k: "%.1e" % v for k, v in worst.items()}))
        assert rel <= 1e-6
    finally:
        for bb in bufs.values():
            bb.close()
