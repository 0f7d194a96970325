"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Bars: FP64 residuals and intermediates within 1e-12 of the largest entry;
per-iteration energies within 1e-12 Eh; converged energies within 1e-10 Eh
(the 2-electron FCI check within 1e-10 Eh).  Mixed precision: FP32 and BF16x3
runs switch to FP64 and must land within 1e-7 Eh of the FP64 energy.
"""
import numpy as np
import pytest

from oracle import ccsd_ref, fci2, synth

pytestmark = pytest.mark.gpu


def _ctx():
    from ccsd_parallel_amd import capi
    return capi


def norm_scale(O, V):
    return 1.0 / np.sqrt(O * synth.p2(V))


def random_amplitudes(O, V, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    t1 = amp * rng.standard_normal((O, V))
    t2 = amp * rng.standard_normal((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, t2


def packed(t2, O, V):
    return synth.pack_pairs(synth.pack_pairs(t2, O, 0), V, 1)


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (37, 53, 29, 0, 0), (128, 128, 64, 0, 1), (64, 300, 17, 1, 0), (200, 33, 1000, 1, 1), (1, 77, 500, 0, 0),
    (190, 257, 4000, 0, 1), (5, 5, 20000, 0, 0), (300, 1, 300, 0, 0),
    # N = 1: pair loads with full 512-element chunks and a ragged end; an odd
    # row pitch (element loads); a transposed A (element loads)
    (257, 1, 1500, 0, 0), (301, 1, 301, 0, 0), (64, 1, 700, 1, 0),
    # N = 1 over an m-contiguous A of at least 64 rows (column kernel), K not a multiple of 8
    (1000, 1, 3001, 1, 0), (130, 1, 7, 1, 1)])
def test_gemm_fp64(M, N, K, ta, tb):
    capi = _ctx()
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 0.7 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.3 * C0
    out = capi.selftest_gemm(A, B, C0, 0.7, -0.3, ta=bool(ta), tb=bool(tb), precision="fp64")
    assert np.max(np.abs(out - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n,ta,tb", [(64, 0, 1), (64, 0, 0), (998, 0, 1), (1000, 1, 0), (1000, 1, 1), (1026, 0, 0)])
def test_gemm_fp64_strassen(n, ta, tb):
    """One level of Strassen's method (Winograd's form) through the engine:
    every orientation of A and B (k- or m/n-contiguous), half sizes with ragged
    tiles (499, 513), against the classical product.  The bound allows the
    small constant-factor growth of the error over the classical product."""
    capi = _ctx()
    rng = np.random.default_rng(n + 10 * ta + tb)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, n))
    C0 = rng.standard_normal((n, n))
    ref = 0.7 * ((A.T if ta else A) @ (B.T if tb else B))
    out = capi.selftest_gemm(A, B, C0, 0.7, 0.0, ta=bool(ta), tb=bool(tb), precision="fp64", strassen=True)
    scale = np.abs(A).max() * np.abs(B).max() * n
    assert np.max(np.abs(out - ref)) <= 1e-14 * scale
    with pytest.raises(capi.CCSDError):   # beta != 0 is not a Strassen product
        capi.selftest_gemm(A, B, C0, 0.7, 0.5, precision="fp64", strassen=True)


@pytest.mark.parametrize("n", [64, 998, 1026])
def test_gemm_fp32_strassen(n):
    """One level of Strassen on FP32 operands with the FP64 result (the FP32
    ring's form), half sizes 32, 499 and 513: its error against the FP64
    product stays within a small factor of the classical FP32 product's."""
    capi = _ctx()
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, n))
    C0 = rng.standard_normal((n, n))
    ref = 0.5 * (A @ B)
    classical = capi.selftest_gemm(A, B, C0, 0.5, 0.0, precision="fp32")
    out = capi.selftest_gemm(A, B, C0, 0.5, 0.0, precision="fp32", strassen=True)
    e_c = np.max(np.abs(classical - ref))
    e_s = np.max(np.abs(out - ref))
    assert 0 < e_c and e_s <= 8 * e_c, (e_s, e_c)
    for prec, beta in (("fp32", 0.5), ("bf16x3", 0.0)):  # beta != 0, and BF16x3, are refused
        with pytest.raises(capi.CCSDError):
            capi.selftest_gemm(A, B, C0, 0.5, beta, precision=prec, strassen=True)


N_FP64_TILES = 17  # configurations of the generic MFMA engine
N_BF16_TILES = 6   # configurations of the BF16 hi/lo kernel


@pytest.mark.parametrize("elementwise", [False, True])
@pytest.mark.parametrize("tile", range(1, N_FP64_TILES + 1))
@pytest.mark.parametrize("M,N,K,ta,tb", [(40, 40, 30000, 0, 1), (97, 301, 77, 1, 0), (300, 40, 300, 0, 0),
                                         (98, 302, 77, 1, 0), (66, 34, 61, 1, 1)])
def test_gemm_fp64_every_tile(tile, M, N, K, ta, tb, elementwise):
    """Every tile shape, ragged edges (M, N and the k tail), the few-tile /
    long-K split path, and both load paths: element pairs (even pitches, all
    four operand layouts among these shapes) and single elements."""
    capi = _ctx()
    rng = np.random.default_rng(tile * 100 + M)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 1.3 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.4 * C0
    out = capi.selftest_gemm(A, B, C0, 1.3, 0.4, ta=bool(ta), tb=bool(tb), precision="fp64", tile=tile,
                             elementwise=elementwise)
    assert np.max(np.abs(out - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))) * np.sqrt(K / 1000 + 1)


@pytest.mark.parametrize("tile", [3, 15])  # 128 x 128 on 4 and on 8 waves
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("M,N,K,ta,tb", [(2900, 2950, 1000, 0, 0), (2944, 2944, 1024, 0, 1), (3000, 2900, 999, 1, 0),
                                         (2900, 2950, 1000, 1, 1)])
def test_gemm_split_tail(prec, M, N, K, ta, tb, tile):
    """128 x 128 tiles (forced) on more tiles than one full round of the GPU:
    512 tiles run whole and the rest as k-chunks whose partials are reduced
    afterwards (ragged M, N and last k-chunk; all four operand layouts)."""
    capi = _ctx()
    rng = np.random.default_rng(M + N + K + ta + 2 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 0.9 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.2 * C0
    out = capi.selftest_gemm(A, B, C0, 0.9, -0.2, ta=bool(ta), tb=bool(tb), precision=prec, tile=tile)
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.max(np.abs(out - ref)) <= tol * np.sqrt(K) * max(1.0, np.abs(ref).max() / np.sqrt(K))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("M,N,K,ta,tb", [(2100, 1000, 300, 0, 1), (1000, 2100, 257, 1, 0), (2050, 1030, 200, 0, 0)])
def test_gemm_tile_bands(prec, M, N, K, ta, tb):
    """128 x 128 tiles (forced) with at least 16 tile rows and 8 tile columns,
    fewer tiles than one full round: the tile walk goes in bands of 8 tile rows
    with a short last band (17 rows: 8 + 8 + 1) and no split tail."""
    capi = _ctx()
    rng = np.random.default_rng(3 * M + N + K + ta + 2 * tb)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 1.1 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.25 * C0
    out = capi.selftest_gemm(A, B, C0, 1.1, 0.25, ta=bool(ta), tb=bool(tb), precision=prec, tile=3)
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert np.max(np.abs(out - ref)) <= tol * np.sqrt(K) * max(1.0, np.abs(ref).max() / np.sqrt(K))


@pytest.mark.parametrize("prec,tile,tol", [("fp32", t, 2e-5) for t in (1, 9, 10, 12)] +
                         [("bf16x3", t, 5e-5) for t in range(1, N_BF16_TILES + 1)])
@pytest.mark.parametrize("M,N,K", [(300, 520, 200), (40, 45, 2000)])
def test_gemm_lowprec_every_tile(prec, tile, tol, M, N, K):
    capi = _ctx()
    rng = np.random.default_rng(M + N + K + tile)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    ref = A @ B - 0.5 * C0
    out = capi.selftest_gemm(A, B, C0, 1.0, -0.5, precision=prec, tile=tile)
    assert np.max(np.abs(out - ref)) <= tol * np.sqrt(K)


@pytest.mark.parametrize("tile", [1, 2])  # BF16 128 x 128 (552 tiles: 512 whole + 40 split), 256 x 256 (144, all split)
def test_gemm_bf16_split_tail(tile):
    capi = _ctx()
    M, N, K = 2900, 2950, 1000
    rng = np.random.default_rng(tile + 7)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    ref = A @ B - 0.5 * C0
    out = capi.selftest_gemm(A, B, C0, 1.0, -0.5, precision="bf16x3", tile=tile)
    assert np.max(np.abs(out - ref)) <= 5e-5 * np.sqrt(K)


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("bf16x3", 5e-5)])
# K = 20 and 64: one and two k-tiles (the pipeline's prologue and last-tile paths)
@pytest.mark.parametrize("M,N,K", [(130, 70, 300), (256, 256, 1024), (33, 129, 65), (70, 50, 20), (100, 90, 64)])
def test_gemm_lowprec(prec, tol, M, N, K):
    capi = _ctx()
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    ref = A @ B + 0.5 * C0
    out = capi.selftest_gemm(A, B, C0, 1.0, 0.5, precision=prec)
    # error relative to the natural scale sqrt(K) of each entry
    assert np.max(np.abs(out - ref)) <= tol * np.sqrt(K)


# (12, 40) and (14, 60): enough packed pairs that the split-K and batch
# reductions (F_mi, F_ae, the T1 tau.<ma||ef> term) take the many-slice path.
# (2, 2): one occupied and one virtual pair (every product has a unit edge);
# (3, 5), (5, 3), (2, 3): odd spin-orbital counts, and V < O.
@pytest.mark.parametrize("O,V,seed", [(2, 2, 17), (3, 5, 18), (5, 3, 19), (2, 3, 20), (4, 8, 11), (6, 10, 12),
                                      (8, 20, 13), (10, 26, 14), (12, 40, 15), (14, 60, 16)])
def test_residual_parity(O, V, seed):
    _check_residual(O, V, seed)


# strassen_min_order 2: the ring and W_mbej's quadratic term (both OV x OV x OV)
# through one level of Strassen at every even OV, here with half orders 3, 16,
# 30 (ragged against every tile), 240 and 420
@pytest.mark.parametrize("O,V,seed", [(2, 3, 20), (4, 8, 11), (6, 10, 12), (12, 40, 15), (14, 60, 16)])
def test_residual_parity_strassen(O, V, seed):
    _check_residual(O, V, seed, strassen_min_order=2)


def test_strassen_option_takes_effect():
    """strassen_min_order reaches the plan and the products: with it the plan
    grows by the three (OV/2)^2 temporaries (a budget of exactly the default
    plan is refused, CCSD_ERR_BUDGET), and the residual differs from the
    classical one in its last bits only."""
    capi = _ctx()
    O, V, seed = 12, 40, 15
    prob = synth.make_problem(O, V, seed, scale=0.05)
    plan = capi.plan_bytes(O, V, "fp64")  # default options: no Strassen at OV = 480
    with pytest.raises(capi.CCSDError) as e:
        capi.Context.from_problem(prob, diis_dim=8, hbm_budget_bytes=plan, strassen_min_order=2)
    assert e.value.code == -3
    t1, t2 = random_amplitudes(O, V, seed)
    res = {}
    for order in (-1, 2):
        with capi.Context.from_problem(prob, diis_dim=8, hbm_budget_bytes=plan + (8 << 20),
                                       strassen_min_order=order) as ctx:
            res[order] = np.concatenate([r.ravel() for r in ctx.residual(t1, packed(t2, O, V))])
    diff = np.abs(res[2] - res[-1]).max()
    assert 0.0 < diff <= 1e-13 * np.abs(res[-1]).max(), diff


def test_strassen_sums_follow_integral_updates():
    """The operand sums of Doovv (formed once with Doovv, not per iteration)
    are formed again when ccsd_update_integrals changes <mn||ef>: after a dense
    update of the whole block the Strassen residual equals the oracle's for the
    updated integrals."""
    capi = _ctx()
    O, V, seed = 6, 10, 24
    prob = synth.make_problem(O, V, seed, scale=0.05)
    rng = np.random.default_rng(seed)
    delta = 0.01 * rng.standard_normal(prob["oovv"].shape)
    t1, t2 = random_amplitudes(O, V, seed)
    upd = dict(prob, oovv=prob["oovv"] + delta)
    R1, R2 = ccsd_ref.rhs(ccsd_ref.Ints.from_problem(upd), t1, t2)[:2]
    with capi.Context.from_problem(prob, diis_dim=0, strassen_min_order=2) as ctx:
        ctx.residual(t1, packed(t2, O, V))  # an iteration's worth of use before the update
        ctx.update_integrals("oovv", np.arange(delta.size), delta.ravel())
        r1, r2 = ctx.residual(t1, packed(t2, O, V))
    scale = max(np.abs(R2).max(), np.abs(R1).max())
    assert np.max(np.abs(r2 - packed(R2, O, V))) <= 1e-12 * scale
    assert np.max(np.abs(r1 - R1)) <= 1e-12 * scale


def _check_residual(O, V, seed, **opts):
    capi = _ctx()
    prob = synth.make_problem(O, V, seed, scale=0.05)
    rng = np.random.default_rng(seed)
    f = prob["f"] + 0.01 * (lambda a: a + a.T)(rng.standard_normal(prob["f"].shape))  # non-canonical Fock
    prob = dict(prob, f=f)
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = random_amplitudes(O, V, seed)
    R1, R2, _, _, X = ccsd_ref.rhs(I, t1, t2, parts=True)
    with capi.Context.from_problem(prob, diis_dim=0, **opts) as ctx:
        r1, r2 = ctx.residual(t1, packed(t2, O, V))
        scale = max(np.abs(R2).max(), np.abs(R1).max())
        # intermediates first, so a failure names the term
        checks = {
            "Fme": (ctx.debug("Fme"), X["Fme"].ravel()),
            "Fmi": (ctx.debug("Fmi"), X["Fmi"].ravel()),
            "Fae": (ctx.debug("Fae"), X["Fae"].ravel()),
            "W1": (ctx.debug("W1"), synth.pack_pairs(synth.pack_pairs(X["W1"], O, 0), O, 1).ravel()),
            "W2": (ctx.debug("W2"), synth.pack_pairs(X["W2"], O, 2).transpose(2, 0, 1).ravel()),
            "W3": (ctx.debug("W3"), synth.pack_pairs(X["W3"], O, 0).transpose(1, 0, 2).ravel()),
            "WT": (ctx.debug("WT"), X["W4"].transpose(3, 1, 0, 2).ravel()),
        }
        for name, (dev, ref) in checks.items():
            err = np.max(np.abs(dev - ref))
            assert err <= 1e-12 * max(1.0, np.abs(ref).max()), f"{name}: {err}"
        assert np.max(np.abs(r1 - R1)) <= 1e-12 * scale
        assert np.max(np.abs(r2 - packed(R2, O, V))) <= 1e-12 * scale
        assert abs(ctx.energy(t1, packed(t2, O, V)) - ccsd_ref.energy(I, t1, t2)) <= 1e-13 * max(1, abs(ccsd_ref.energy(I, t1, t2)))


def _asymmetric_vvvv_problem(O, V, seed):
    prob = synth.make_problem(O, V, seed, scale=0.05)
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(prob["vvvv"].shape)
    return dict(prob, vvvv=prob["vvvv"] + 0.01 * (q - q.T))


def test_residual_parity_asymmetric_vvvv():
    """The pp ladder contracts tau_ij^ef with <ab||ef>: row ab, column ef of the
    packed block, as Fill_t2_new does (ccsd_memory_optimized.cpp:2093).  A
    block that is not symmetric under pair exchange must still give the
    oracle's residual (contracting it the other way round moves the ladder
    term by about half its size here)."""
    capi = _ctx()
    O, V, seed = 6, 12, 21
    prob = _asymmetric_vvvv_problem(O, V, seed)
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = random_amplitudes(O, V, seed)
    R1, R2 = ccsd_ref.rhs(I, t1, t2)[:2]
    with capi.Context.from_problem(prob, diis_dim=0) as ctx:
        r1, r2 = ctx.residual(t1, packed(t2, O, V))
    scale = max(np.abs(R2).max(), np.abs(R1).max())
    assert np.max(np.abs(r2 - packed(R2, O, V))) <= 1e-12 * scale
    assert np.max(np.abs(r1 - R1)) <= 1e-12 * scale


def test_residual_parity_asymmetric_oooo_vvvv():
    """oooo and vvvv are read as stored, <row||column>: with both made
    asymmetric under pair exchange (transposing either moves the residual by
    12-21% here) the residual and W_mnij still equal the oracle's.  ovov is
    the exception: W_mbej and the T1 ring term take <ia||jb> = <jb||ia>
    (real orbitals, as the reference)."""
    capi = _ctx()
    O, V, seed = 5, 9, 23
    prob = synth.make_problem(O, V, seed, scale=0.05)
    rng = np.random.default_rng(seed)
    for blk in ("oooo", "vvvv"):
        q = rng.standard_normal(prob[blk].shape)
        prob[blk] = prob[blk] + 0.01 * (q - q.T)
    I = ccsd_ref.Ints.from_problem(prob)
    t1, t2 = random_amplitudes(O, V, seed)
    R1, R2, _, _, X = ccsd_ref.rhs(I, t1, t2, parts=True)
    with capi.Context.from_problem(prob, diis_dim=0) as ctx:
        r1, r2 = ctx.residual(t1, packed(t2, O, V))
        for name, dev, ref in (("W1", ctx.debug("W1"), synth.pack_pairs(synth.pack_pairs(X["W1"], O, 0), O, 1)),
                               ("WT", ctx.debug("WT"), X["W4"].transpose(3, 1, 0, 2))):
            err = np.max(np.abs(dev - ref.ravel()))
            assert err <= 1e-12 * max(1.0, np.abs(ref).max()), (name, err)
    scale = max(np.abs(R2).max(), np.abs(R1).max())
    assert np.max(np.abs(r2 - packed(R2, O, V))) <= 1e-12 * scale
    assert np.max(np.abs(r1 - R1)) <= 1e-12 * scale


@pytest.mark.parametrize("prec,tol", [("fp32", 5e-6), ("bf16x3", 5e-4)])
def test_low_precision_update_asymmetric_vvvv(prec, tol):
    """The low-precision ladder contracts the packed block the same way round
    as the FP64 one: one update from the same amplitudes agrees to the
    rounding level of the mode on a block that is not pair-exchange symmetric."""
    capi = _ctx()
    O, V, seed = 6, 12, 22
    prob = _asymmetric_vvvv_problem(O, V, seed)
    t1, t2 = random_amplitudes(O, V, seed, amp=0.02)
    out = {}
    with capi.Context.from_problem(prob, diis_dim=0) as ctx:
        for p in ("fp64", prec):
            ctx.set_amplitudes(t1, packed(t2, O, V))
            ctx.iterate(1, p)
            out[p] = ctx.get_amplitudes()
    (a1, a2), (b1, b2) = out["fp64"], out[prec]
    e2 = np.max(np.abs(b2 - a2)) / np.max(np.abs(a2))
    e1 = np.max(np.abs(b1 - a1)) / np.max(np.abs(a1))
    assert e2 <= tol and e1 <= tol, (e1, e2)


def test_synthetic_generator_matches_host():
    """Device-generated integrals equal the host generator (up to libm rounding)."""
    capi = _ctx()
    O, V, seed = 6, 14, 2024
    sc = norm_scale(O, V)
    prob = synth.make_problem(O, V, seed, scale=sc)
    t1, t2 = random_amplitudes(O, V, 5, amp=0.02)
    with capi.Context.from_problem(prob) as a, capi.Context.synthetic(O, V, seed, sc) as b:
        ra = a.residual(t1, packed(t2, O, V))
        rb = b.residual(t1, packed(t2, O, V))
        for x, y in zip(ra, rb):
            assert np.max(np.abs(x - y)) <= 1e-14 * max(1.0, np.abs(x).max())


def test_energy_trace_parity_jacobi():
    """Plain Jacobi (the reference's update, no DIIS): 40 updates, energy by energy."""
    capi = _ctx()
    O, V, seed = 8, 24, 20180
    prob = synth.make_problem(O, V, seed, scale=norm_scale(O, V))
    I = ccsd_ref.Ints.from_problem(prob)
    E, t1, t2, info = ccsd_ref.solve(I, toler=0.0, max_iter=41, diis_dim=0)
    with capi.Context.from_problem(prob, diis_dim=0) as ctx:
        r = ctx.solve(0.0, 40, allow_unconverged=True)
        tr = ctx.energy_trace()
        g1, g2 = ctx.get_amplitudes()
    ref = np.array(info["energies"])
    assert r["iterations"] == 40 and len(tr) == 41 and len(ref) == 41
    assert np.max(np.abs(tr - ref)) <= 1e-11
    assert abs(r["e_mp2"] - ref[0]) <= 1e-13
    assert np.max(np.abs(g2 - packed(t2, O, V))) <= 1e-10
    assert np.max(np.abs(g1 - t1)) <= 1e-10


def test_energy_trace_parity_diis():
    capi = _ctx()
    O, V, seed = 8, 24, 20180
    prob = synth.make_problem(O, V, seed, scale=norm_scale(O, V))
    I = ccsd_ref.Ints.from_problem(prob)
    E, t1, t2, info = ccsd_ref.solve(I, toler=1e-10, max_iter=200, diis_dim=8)
    with capi.Context.from_problem(prob, diis_dim=8) as ctx:
        r = ctx.solve(1e-10, 200)
        tr = ctx.energy_trace()
        g1, g2 = ctx.get_amplitudes()
    ref = np.array(info["energies"])
    assert r["converged"] == 1 and info["converged"]
    n = min(len(ref), len(tr))
    assert np.max(np.abs(tr[:n] - ref[:n])) <= 1e-11
    assert abs(r["e_corr"] - E) <= 1e-10
    assert np.max(np.abs(g2 - packed(t2, O, V))) <= 1e-8


def test_two_electron_fci():
    capi = _ctx()
    O, V, seed = 2, 10, 77
    g = synth.full_tensor(O, V, seed, scale=0.05)
    eo, ev = synth.orbital_energies(O, V, seed)
    f = np.diag(np.concatenate([eo, ev]))
    prob = dict(O=O, V=V, f=f, **synth.pack_dense(synth.blocks_from_full(g, O), O, V))
    e_fci = fci2.fci_two_electron(f, g)[0]
    with capi.Context.from_problem(prob) as ctx:
        r = ctx.solve(1e-12, 200)
    assert abs(r["e_corr"] - e_fci) <= 1e-10


@pytest.mark.parametrize("prec,floor", [("fp32", 1e-7), ("bf16x3", 2e-6)])
def test_low_precision_fixed_point_is_near_fp64(prec, floor):
    """Never switching to FP64, the low-precision iteration must still converge
    to within its rounding floor of the FP64 energy (catches a wrong or stale
    low-precision operand, which the FP64 refinement would otherwise hide)."""
    capi = _ctx()
    O, V, seed = 10, 40, 4321
    sc = norm_scale(O, V)
    with capi.Context.synthetic(O, V, seed, sc) as ctx:
        e64 = ctx.solve(1e-11, 200)["e_corr"]
    with capi.Context.synthetic(O, V, seed, sc, precision=prec, switch_threshold=1e-30) as ctx:
        r = ctx.solve(1e-30, 40, allow_unconverged=True)
        tr = ctx.energy_trace()
    assert r["fp64_iterations"] == 0
    assert abs(tr[-1] - e64) <= floor * max(1.0, abs(e64))


# (3, 5): odd counts, every tile mostly clamped rows; (50, 20): more occupied
# than virtual orbitals, N = 50 in the batched W_mbej t1 product; (2, 330):
# long packed rows (P2(V) = 54285) with a single occupied pair
@pytest.mark.parametrize("O,V,seed", [(8, 36, 77), (3, 5, 78), (50, 20, 80), (2, 330, 81)])
@pytest.mark.parametrize("prec,tol", [("fp32", 5e-6), ("bf16x3", 5e-4)])
def test_low_precision_update_matches_fp64_update(prec, tol, O, V, seed):
    """One amplitude update from the same amplitudes (T1 nonzero everywhere) in
    FP64 and in low precision.  Every contraction of the low-precision residual
    -- the BF16/FP32 dominant GEMMs and the FP32 secondary ones -- enters the
    new amplitudes, which must agree to the rounding level of the mode."""
    _check_low_precision_update(prec, tol, O, V, seed)


# strassen_min_order 2: the FP32 ring (FP64 result) and W_mbej's quadratic term
# (FP32 result accumulated into the FP32 W_mbej) through one level of Strassen,
# half orders 144 and 500
@pytest.mark.parametrize("O,V,seed", [(8, 36, 77), (50, 20, 80)])
def test_fp32_update_with_strassen_matches_fp64_update(O, V, seed):
    _check_low_precision_update("fp32", 5e-6, O, V, seed, strassen_min_order=2)


def _check_low_precision_update(prec, tol, O, V, seed, **opts):
    capi = _ctx()
    sc = norm_scale(O, V)
    rng = np.random.default_rng(5)
    with capi.Context.synthetic(O, V, seed, sc, **opts) as ctx:
        ctx.iterate(2)
        t1, t2 = ctx.get_amplitudes()
        t1 = t1 + 0.01 * rng.standard_normal(t1.shape)
        out = {}
        for p in ("fp64", prec, "fp64"):
            ctx.set_amplitudes(t1, t2)
            ctx.iterate(1, p)
            out.setdefault(p, []).append(ctx.get_amplitudes())
    (a1, a2), (a1b, a2b) = out["fp64"]
    b1, b2 = out[prec][0]
    assert np.array_equal(a2, a2b) and np.array_equal(a1, a1b)  # FP64 after low precision is unaffected
    e2 = np.max(np.abs(b2 - a2)) / np.max(np.abs(a2))
    e1 = np.max(np.abs(b1 - a1)) / np.max(np.abs(a1))
    assert 0 < e2 <= tol and e1 <= tol, (e1, e2)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_converges_to_fp64(prec):
    capi = _ctx()
    O, V, seed = 10, 40, 4321
    sc = norm_scale(O, V)
    with capi.Context.synthetic(O, V, seed, sc) as ctx:
        r64 = ctx.solve(1e-10, 200)
        t1a, t2a = ctx.get_amplitudes()
    with capi.Context.synthetic(O, V, seed, sc, precision=prec, switch_threshold=1e-6, final_fp64_sweeps=3) as ctx:
        rmx = ctx.solve(1e-10, 200)
        t1b, t2b = ctx.get_amplitudes()
    assert rmx["lowp_iterations"] > 0 and rmx["fp64_iterations"] >= 3
    assert abs(rmx["e_corr"] - r64["e_corr"]) <= 1e-7
    assert np.max(np.abs(t2a - t2b)) <= 1e-6
