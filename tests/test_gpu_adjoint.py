"""GPU adjoint (ccsd_backward) against the CPU adjoint oracle and finite differences.

Gradients are with respect to the packed input arrays (the Fock matrix and the
six <pq||rs> blocks), symmetrised over duplicate entries, so a symmetric
perturbation Delta gives dE/dh = sum(g * Delta) over the stored entries.
"""
import numpy as np
import pytest

from oracle import adjoint_ref, synth

pytestmark = pytest.mark.gpu
KEYS = ("f",) + synth.BLOCKS


def _capi():
    from ccsd_parallel_amd import capi
    return capi


def _prob(O, V, seed, noncanon=False):
    prob = synth.make_problem(O, V, seed, scale=1.0 / np.sqrt(O * synth.p2(V)))
    if noncanon:
        rng = np.random.default_rng(seed)
        a = rng.standard_normal(prob["f"].shape)
        prob["f"] = prob["f"] + 0.01 * (a + a.T)
    return prob


# (3, 5) and (2, 3): odd spin-orbital counts and single-tile edges; (12, 40) and
# (14, 60) put the reverse pass's split-K and many-slice batch reductions (the
# transposes of the forward's F_mi, F_ae and T1 tau.<ma||ef> reductions) on
# their production paths, as the forward parity tests do at the same shapes
@pytest.mark.parametrize("O,V,seed,noncanon", [(3, 5, 35, True), (2, 3, 36, False), (4, 8, 31, False), (6, 12, 32, True),
                                               (8, 16, 33, False), (12, 40, 38, True), (14, 60, 39, False)])
def test_backward_matches_cpu_adjoint(O, V, seed, noncanon):
    _check_backward(O, V, seed, noncanon)


# strassen_min_order 2: the forward's (OV)^3 products and the four reverse
# ones (the ring's two cotangents, W_mbej's quadratic term and its
# t-cotangent) through one level of Strassen, half orders 3, 36 and 240
@pytest.mark.parametrize("O,V,seed,noncanon", [(2, 3, 36, False), (6, 12, 32, True), (12, 40, 38, True)])
def test_backward_strassen_matches_cpu_adjoint(O, V, seed, noncanon):
    _check_backward(O, V, seed, noncanon, strassen_min_order=2)


def _check_backward(O, V, seed, noncanon, **opts):
    capi = _capi()
    prob = _prob(O, V, seed, noncanon)
    E, gref, _, linfo = adjoint_ref.adjoint(prob, toler=1e-13)
    with capi.Context.from_problem(prob, diis_dim=8, **opts) as ctx:
        r = ctx.solve(1e-13, 300)
        g, info = ctx.backward()
    assert abs(r["e_corr"] - E) < 1e-11
    assert info["converged"] == 1
    for k in KEYS:
        scale = max(1e-3, np.abs(gref[k]).max())
        err = np.abs(g[k] - gref[k]).max()
        assert err <= 1e-8 * scale, (k, err, scale)  # Lambda converged to 1e-10


def test_backward_dense_finite_differences():
    """16 dense random symmetric directions on the host-built problem, h = 1e-5."""
    capi = _capi()
    O, V, seed = 6, 16, 34
    prob = _prob(O, V, seed)
    with capi.Context.from_problem(prob, diis_dim=8) as ctx:
        ctx.solve(1e-13, 300)
        g, _ = ctx.backward()
    h = 1e-5
    worst = 0.0
    for k in range(16):
        d = synth.make_problem(O, V, 5000 + k, scale=1.0)
        rng = np.random.default_rng(k)
        a = rng.standard_normal(d["f"].shape)
        d["f"] = 0.5 * (a + a.T)
        pred = sum(float(np.sum(g[b] * d[b])) for b in KEYS)
        es = []
        for sgn in (1, -1):
            p = dict(prob, **{b: prob[b] + sgn * h * d[b] for b in KEYS})
            with capi.Context.from_problem(p, diis_dim=8) as c2:
                es.append(c2.solve(1e-14, 300)["e_corr"])
        fd = (es[0] - es[1]) / (2 * h)
        worst = max(worst, abs(fd - pred) / max(1e-12, abs(fd)))
    assert worst <= 1e-6, worst


def _sparse_direction(rng, O, V, kind, n):
    """Flat indices and deltas of one symmetric-consistent sparse direction."""
    N = O + V
    po, pv, ov = synth.p2(O), synth.p2(V), O * V
    shapes = dict(f=(N, N), oooo=(po, po), ooov=(po, ov), oovv=(po, pv), ovov=(ov, ov), ovvv=(ov, pv),
                  vvvv=(pv, pv))
    r, c = shapes[kind]
    if kind == "f":
        i = rng.integers(0, N, size=N * N // 4)
        j = rng.integers(0, N, size=N * N // 4)
    else:
        i = rng.integers(0, r, size=n)
        j = rng.integers(0, c, size=n)
    # larger steps on the blocks with small gradients lift the difference above the
    # ~1e-9 noise of converged energies; small steps elsewhere keep O(h^2) terms small
    v = rng.standard_normal(len(i)) * dict(f=1.0, oovv=1.0, ooov=1.0, ovvv=3.0, oooo=10.0, ovov=10.0,
                                           vvvv=10.0)[kind]
    if kind in ("f", "oooo", "ovov", "vvvv"):  # symmetric blocks: add the mirrored entry too
        idx = np.concatenate([i * c + j, j * c + i])
        val = np.concatenate([v, v])
    else:
        idx, val = i * c + j, v
    return idx.astype(np.int64), val


@pytest.mark.timeout(900)
def test_backward_finite_differences_o20_v180():
    """64 seeded directions at O=20 / V=180, applied on the device with
    ccsd_update_integrals; each direction perturbs one input array (cycling
    over f and the six blocks).  Four-point central differences, h = 1e-5,
    f'(0) = [8 (E(h) - E(-h)) - (E(2h) - E(-2h))] / 12h (truncation O(h^4)).

    Every direction must agree to 1e-6 relative or to an absolute floor of
    1e-9: energies converged to |dE| < 1e-15 carry ~2e-15 Eh of round-off
    (a 3.1e6-term FP64 dot product at |E| ~ 1.9), which the stencil turns into
    about 1.5 * 2e-15 / 1e-5 = 3e-10 in the derivative."""
    capi = _capi()
    O, V, seed = 20, 180, 20180
    sc = 1.0 / np.sqrt(O * synth.p2(V))
    h = 1e-5
    kinds = list(KEYS)
    with capi.Context.synthetic(O, V, seed, sc, diis_dim=8) as ctx:
        ctx.solve(1e-13, 200)
        t1, t2 = ctx.get_amplitudes()
        g, info = ctx.backward()
        assert info["converged"] == 1
        rng = np.random.default_rng(64)
        rel, rows = [], []
        for k in range(64):
            kind = kinds[k % len(kinds)]
            idx, val = _sparse_direction(rng, O, V, kind, 20000)
            pred = float(np.sum(g[kind].ravel()[idx] * val))
            es = {}
            pos = 0
            for target in (1, 2, -1, -2):                  # steps of h: +h, +2h, -h, -2h
                ctx.update_integrals(kind, idx, (target - pos) * h * val)
                pos = target
                ctx.set_amplitudes(t1, t2)
                es[target] = ctx.solve(1e-15, 300, allow_unconverged=True)["e_corr"]
            ctx.update_integrals(kind, idx, -pos * h * val)  # restore
            fd = (8.0 * (es[1] - es[-1]) - (es[2] - es[-2])) / (12.0 * h)
            rel.append(abs(fd - pred) / max(abs(fd), 1e-30))
            rows.append((kind, fd, pred))
        rel = np.array(rel)
    for kind, fd, pred in sorted(rows, key=lambda x: -abs(x[1] - x[2]) / max(abs(x[1]), 1e-30))[:8]:
        print("worst: %s fd %.6e pred %.6e |d| %.2e rel %.2e" % (kind, fd, pred, abs(fd - pred),
                                                                abs(fd - pred) / max(abs(fd), 1e-30)))
    print("FD relative errors: max %.2e median %.2e" % (rel.max(), np.median(rel)))
    for kind, fd, pred in rows:
        assert abs(fd - pred) <= max(1e-6 * abs(fd), 1e-9), (kind, fd, pred)
    assert np.median(rel) <= 1e-7


def test_backward_mixed_precision_matches_fp64():
    """BF16x3 forward (ending with FP64 sweeps) and BF16x3 Lambda iterations
    (ending with FP64-verified refinement: the last pass is an FP64 one that
    meets the tolerance) give the FP64 gradient: relative error of the whole
    gradient below 1e-6."""
    capi = _capi()
    O, V, seed = 10, 40, 35
    sc = 1.0 / np.sqrt(O * synth.p2(V))
    out = {}
    for prec in ("fp64", "bf16x3"):
        with capi.Context.synthetic(O, V, seed, sc, precision=prec, diis_dim=8) as ctx:
            r = ctx.solve(1e-10, 200)
            g, info = ctx.backward()
        out[prec] = (r, g, info)
    (r64, g64, i64), (rmx, gmx, imx) = out["fp64"], out["bf16x3"]
    assert i64["lambda_lowp_iterations"] == 0 and imx["lambda_lowp_iterations"] > 0
    assert imx["converged"] == 1 and imx["lambda_iterations"] - imx["lambda_lowp_iterations"] >= 1
    num = np.sqrt(sum(float(np.sum((gmx[k] - g64[k]) ** 2)) for k in KEYS))
    den = np.sqrt(sum(float(np.sum(g64[k] ** 2)) for k in KEYS))
    assert num / den <= 1e-6, num / den
    assert abs(rmx["e_corr"] - r64["e_corr"]) <= 1e-8


def test_backward_mixed_switches_at_a_stall():
    """lambda_switch_factor 1e-6 puts the Lambda switch (1e-10) below the
    BF16x3 rounding floor: the Lambda iteration switches to the FP64-verified
    refinement once max|dLambda| has set no new minimum in 4 iterations, and
    the gradient still matches the FP64 one."""
    capi = _capi()
    O, V, seed = 8, 24, 37
    sc = 1.0 / np.sqrt(O * synth.p2(V))
    out = {}
    for prec, opts in (("fp64", {}), ("bf16x3", {"lambda_switch_factor": 1e-6})):
        with capi.Context.synthetic(O, V, seed, sc, precision=prec, diis_dim=8, **opts) as ctx:
            ctx.solve(1e-10, 200)
            g, info = ctx.backward()
        out[prec] = (g, info)
    (g64, _), (gmx, imx) = out["fp64"], out["bf16x3"]
    assert imx["converged"] == 1 and imx["lambda_lowp_iterations"] > 0
    assert imx["lambda_iterations"] - imx["lambda_lowp_iterations"] >= 1
    num = np.sqrt(sum(float(np.sum((gmx[k] - g64[k]) ** 2)) for k in KEYS))
    den = np.sqrt(sum(float(np.sum(g64[k] ** 2)) for k in KEYS))
    assert num / den <= 1e-6, num / den
