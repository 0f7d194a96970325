"""GPU parity at the workload sizes themselves (O=20/V=180 and O=40/V=300).

The dense oracle cannot hold these tensors (V^4 doubles), so the checks use
what stays exact at any size:
  * the first energy: E(MP2 guess) = sum_{i<j,a<b} <ij||ab>^2 / D_ij^ab
    (ccsd_memory_optimized.cpp:473-502), from the host generator's packed
    <ij||ab> block and orbital energies;
  * the packed pp + hh ladder (Fill_t2_new, :2093): for a set of columns a<b
    at the edges and the middle of the P2(V) range,
    Rp[:, ab] = <ij||ab> + sum_{c<d} tau[ij][cd] <cd||ab> + sum_{m<n} W_mnij[mn][ij] tau[mn][ab],
    with <ab||cd> rows regenerated on the host for just those pairs and tau,
    W_mnij taken from the engine's own buffers (their construction is pinned
    against the oracle at small sizes in test_gpu_parity.py);
  * the ring GEMM (Fill_W5, :6240): rows (i,a) of X = Tring W_mbej^T, from the
    engine's Tring and W_mbej buffers;
  * the whole right-hand side at O=20/V=180 against the oracle's independent
    packed C++ restatement (oracle/ccsd_cpu.cpp) on the host generator's blocks;
  * the converged FP64 solve at O=20/V=180 to 1e-10 (the plain FP64
    configuration): its amplitudes are a fixed point, D t = R(t), to the
    reported max|dt2|, and the returned energy is E(t*).
These exercise the production tile / split-K configurations at the real
shapes (P2(O) = 190 / 780 rows, K = P2(V) = 16110 / 44850, OV = 3600 / 12000).
"""
import numpy as np
import pytest

from oracle import synth

pytestmark = pytest.mark.gpu

SHAPES = [(20, 180, 20180), (40, 300, 40300)]


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _columns(n, width=48):
    mid = n // 2
    return np.unique(np.concatenate([np.arange(width), np.arange(mid, mid + width), np.arange(n - width, n)]))


@pytest.mark.parametrize("O,V,seed", SHAPES)
def test_full_size_products_and_first_energy(O, V, seed):
    capi = _capi()
    scale = 1.0 / np.sqrt(O * synth.p2(V))
    po, pv, ov = synth.p2(O), synth.p2(V), O * V
    print("case: O", O, "V", V, "seed", seed, flush=True)
    with capi.Context.synthetic(O, V, seed, scale, diis_dim=8) as ctx:
        # first energy = closed form on the generator's <ij||ab>
        oovv = synth.packed_block("oovv", O, V, seed, scale)               # [i<j][a<b]
        eo, ev = synth.orbital_energies(O, V, seed)
        ip, iq = synth.pair_list(O)
        ap, aq = synth.pair_list(V)
        D = (eo[ip] + eo[iq])[:, None] - (ev[ap] + ev[aq])[None, :]
        e_mp2 = float(np.sum(oovv * oovv / D))
        t1m, t2m = ctx.get_amplitudes()                                     # the MP2 guess
        assert abs(ctx.energy(t1m, t2m) - e_mp2) < 1e-12 * abs(e_mp2)
        assert np.abs(t2m - oovv / D).max() < 1e-14 * np.abs(t2m).max()

        # right-hand side at random amplitudes around the MP2 guess
        rng = np.random.default_rng(seed + 1)
        t1 = 1e-3 * rng.standard_normal((O, V))
        t2 = t2m * (1.0 + 0.1 * rng.standard_normal(t2m.shape))
        ctx.residual(t1, t2)
        tau = ctx.debug("tau").reshape(po, pv)
        W1 = ctx.debug("W1").reshape(po, po)
        Rp = ctx.debug("Rp").reshape(po, pv)
        cols = _columns(pv)
        vv = np.concatenate([synth.packed_block("vvvv", O, V, seed, scale, rows=(int(c), int(c) + 1)) for c in cols])
        ref = oovv[:, cols] + tau @ vv.T + W1.T @ tau[:, cols]
        err = np.abs(Rp[:, cols] - ref).max()
        assert err < 1e-12 * np.abs(ref).max(), err

        Tring = ctx.debug("Tring").reshape(ov, ov)
        WT = ctx.debug("WT").reshape(ov, ov)
        X = ctx.debug("X").reshape(ov, ov)
        rows = _columns(ov, 32)
        ref = Tring[rows] @ WT.T
        err = np.abs(X[rows] - ref).max()
        assert err < 1e-12 * np.abs(ref).max(), err


@pytest.mark.timeout(600)
def test_residual_o20_v180_matches_cpu_oracle():
    """The whole FP64 right-hand side at O=20 / V=180 (R1 and the packed R2)
    against the oracle's independent packed C++ restatement
    (oracle/ccsd_cpu.cpp cpu_ccsd_residual, itself equal to the dense numpy
    oracle at small shapes), on the host generator's blocks, at random
    amplitudes around the MP2 guess: every entry within 1e-12 of the largest."""
    from oracle import cpu_build
    capi = _capi()
    O, V, seed = 20, 180, 20180
    scale = 1.0 / np.sqrt(O * synth.p2(V))
    print("case: residual O", O, "V", V, "seed", seed, flush=True)
    with capi.Context.synthetic(O, V, seed, scale, diis_dim=8) as ctx:
        t1m, t2m = ctx.get_amplitudes()
        rng = np.random.default_rng(seed + 2)
        t1 = 1e-3 * rng.standard_normal((O, V))
        t2 = t2m * (1.0 + 0.1 * rng.standard_normal(t2m.shape))
        g1, g2 = ctx.residual(t1, t2)
    prob = synth.make_problem(O, V, seed, scale=scale, workers=16)
    c1, c2 = cpu_build.residual(prob, t1, t2)
    e2 = np.abs(g2 - c2).max() / np.abs(c2).max()
    e1 = np.abs(g1 - c1).max() / np.abs(c1).max()
    print("R2 max|d|/max %.2e  R1 max|d|/max %.2e" % (e2, e1), flush=True)
    assert e2 < 1e-12 and e1 < 1e-12


def test_fp64_solve_o20_v180_is_a_fixed_point():
    capi = _capi()
    O, V, seed = 20, 180, 20180
    scale = 1.0 / np.sqrt(O * synth.p2(V))
    eo, ev = synth.orbital_energies(O, V, seed)
    ip, iq = synth.pair_list(O)
    ap, aq = synth.pair_list(V)
    D2 = (eo[ip] + eo[iq])[:, None] - (ev[ap] + ev[aq])[None, :]
    D1 = eo[:, None] - ev[None, :]
    with capi.Context.synthetic(O, V, seed, scale, diis_dim=8) as ctx:
        r = ctx.solve(1e-10, 200)
        assert r["converged"] and r["last_delta_e"] < 1e-10
        t1, t2 = ctx.get_amplitudes()
        assert abs(ctx.energy(t1, t2) - r["e_corr"]) < 1e-12
        R1, R2 = ctx.residual(t1, t2)
        d2 = np.abs(R2 / D2 - t2).max()
        d1 = np.abs(R1 / D1 - t1).max()
        print("E", r["e_corr"], "iterations", r["iterations"], "max|dt1|", d1, "max|dt2|", d2, flush=True)
        assert d2 < 1e-8 and d1 < 1e-8
        assert d2 <= 10 * max(r["max_delta_t2"], 1e-13)
