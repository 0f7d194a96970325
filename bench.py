"""Benchmark: CCSD amplitude iterations on one MI355X.

A "step" is one CCSD amplitude update on synthetic spin-orbital integrals
already resident in HBM. It covers the right-hand side, the Jacobi/DIIS update,
tau and the energy.

Workload: O=40 occupied / V=300 virtual spin orbitals, seed 40300, FP64.
`value` is seconds per FP64 iteration. The same JSON line also reports:
  * time-to-converge to |dE| <= 1e-7 Eh for FP64 and for the mixed paths
    (FP32 and BF16 hi/lo x3, each finishing with 3 FP64 sweeps);
  * per-iteration cost in each precision;
  * pp-ladder TF/s and its roofline fraction (HIP events on the engine stream,
    averaged over the timed steps);
  * mixed forward + adjoint against FP64 forward + adjoint (wall time, |dE|,
    gradient relative error);
  * the plain FP64 configuration O=20 / V=180 (seed 20180) converged to 1e-10;
  * the RHF input path at the same O/V (n = O/2 + V/2 spatial orbitals): the
    four AO->MO quarter-transform GEMMs and the spin-orbital packing in HBM;
  * the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}  # dense MFMA TFLOP/s, MI355X
METRIC = "CCSD s/iteration & time-to-converge (FP64 vs mixed) at |dE|<=1e-7 Eh; ladder TF/s"


def p2(n):
    return n * (n - 1) // 2


def flops_per_iteration(O, V):
    po, pv, ov = p2(O), p2(V), O * V
    return dict(
        pp_ladder=2.0 * pv * pv * po,
        hh_ladder=2.0 * pv * po * po,
        wmnij=2.0 * po * po * pv,
        wmbej_quad=2.0 * ov ** 3,
        ring=2.0 * ov ** 3,
        wmbij=2.0 * ov * po * pv,
        t1_ovvv=2.0 * O * O * V ** 3,
    )


def cpu_baseline(O, V, seed, scale, budget_s=12.0):
    """Reference-algorithm CPU restatement on a bounded share of the workload.

    oracle/ccsd_cpu.cpp restates the reference's loop shapes: one dot product
    per output over contiguous rows (accu(A % B), ccsd_memory_optimized.cpp),
    OpenMP threads in place of MPI ranks splitting the output rows.  Two
    shares of the O=40/V=300 iteration are timed on all host threads, about
    budget_s / 2 each:
      * ladder type: rows a<b of <ab||cd> (same generator) against all P2(O)
        rows of tau, dots of length P2(V) -- the rate of the pp ladder, the hh
        ladder, W_mnij and W_mbij;
      * ring type: all O V rows of Tring against columns (j,b) of W_mbej,
        dots of length O V -- the rate of the ring, the W_mbej t2 term and the
        t1 <mb||ef> terms.
    Each term's flops at its class rate give the extrapolated s/iteration."""
    from oracle import cpu_build, synth
    lib = cpu_build.load()
    threads = lib.cpu_max_threads()
    po, pv, ov = p2(O), p2(V), O * V
    fl = flops_per_iteration(O, V)
    rng = np.random.default_rng(0)

    def timed(A, B, budget):
        """Repeat the dot-product share until `budget` seconds have passed."""
        reps, el = 0, 0.0
        t0 = time.perf_counter()
        while el < budget:
            cpu_build.dot_rows(A, B)
            reps += 1
            el = time.perf_counter() - t0
        return reps, el

    # ladder type: 128 rows a<b of <ab||cd> (the generator's values) against every tau row
    tau = rng.standard_normal((po, pv)) * 1e-3
    rows = 128
    blk = synth.packed_block("vvvv", O, V, seed, scale, rows=(0, rows))
    reps_l, el_l = timed(blk, tau, budget_s / 2)
    rate_l = 2.0 * rows * po * pv * reps_l / el_l
    del tau, blk
    # ring type: every Tring row (i,a) against 64 columns (j,b) of W_mbej
    tring = rng.standard_normal((ov, ov)) * 1e-3
    cols = 64
    wt = rng.standard_normal((cols, ov)) * 1e-3
    reps_r, el_r = timed(tring, wt, budget_s / 2)
    rate_r = 2.0 * ov * cols * ov * reps_r / el_r
    del tring
    ladder_class = fl["pp_ladder"] + fl["hh_ladder"] + fl["wmnij"] + fl["wmbij"]
    ring_class = fl["ring"] + fl["wmbej_quad"] + fl["t1_ovvv"]
    value = ladder_class / rate_l + ring_class / rate_r
    return {
        "value": value, "unit": "s/iteration (extrapolated)", "cores": threads, "kind": "port",
        "sample": f"reference-algorithm CPU restatement (oracle/ccsd_cpu.cpp: dot products over contiguous rows, "
                  f"OpenMP over output rows; not the reference binary): ladder type {rows} rows of <ab||cd> x "
                  f"{po} tau rows (length {pv}), {reps_l} reps in {el_l:.1f}s = {rate_l / 1e9:.1f} GFLOP/s; "
                  f"ring type {ov} Tring rows x {cols} W_mbej columns (length {ov}), {reps_r} reps in "
                  f"{el_r:.1f}s = {rate_r / 1e9:.1f} GFLOP/s; "
                  f"{ladder_class:.3e} + {ring_class:.3e} flop/iteration at those rates",
    }


def ladder_traffic(O, V, precision):
    """HBM bytes per pp-ladder launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE passes, calibrated on a copy kernel), if it covers
    this workload; otherwise None."""
    path = os.path.join(ROOT, "profiles", f"pp_ladder_o{O}_v{V}_{precision}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def combined_forward_adjoint(capi, O, V, seed, scale, tol=1e-7):
    """Mixed forward + adjoint against FP64 forward + adjoint on the same
    workload: wall time of solve (to |dE| < tol) plus backward (Lambda
    equations and the density contractions into dE/df, dE/d<pq||rs>, written
    into page-locked host arrays), |dE_corr| and the relative error of the
    whole gradient (Frobenius norm over all blocks)."""
    res, bufs = {}, {}
    try:
        for p in ("fp64", "bf16x3"):
            c = capi.Context.synthetic(O, V, seed, scale, precision=p, diis_dim=8)
            bufs[p] = capi.PinnedArrays(c.gradient_shapes())
            t0 = time.perf_counter()
            r = c.solve(tol, 100)
            t1 = time.perf_counter()
            _, info = c.backward(out=bufs[p].arrays)
            t2 = time.perf_counter()
            c.close()
            res[p] = {"e_corr": r["e_corr"], "forward_s": t1 - t0, "backward_s": t2 - t1, "total_s": t2 - t0,
                      "forward_iterations": f"{r['lowp_iterations']} low + {r['fp64_iterations']} FP64",
                      "lambda_iterations": f"{info['lambda_lowp_iterations']} low + "
                                           f"{info['lambda_iterations'] - info['lambda_lowp_iterations']} FP64"}
        a, b = bufs["fp64"].arrays, bufs["bf16x3"].arrays
        num = den = 0.0
        for k in a:
            for r0 in range(0, a[k].shape[0], 1024):  # row chunks: no full-size temporaries
                x, y = a[k][r0:r0 + 1024], b[k][r0:r0 + 1024]
                num += float(np.sum((y - x) ** 2))
                den += float(np.sum(x * x))
        res["abs_dE_corr"] = abs(res["bf16x3"]["e_corr"] - res["fp64"]["e_corr"])
        res["gradient_rel_error"] = float(np.sqrt(num / den))
        res["speedup_total"] = res["fp64"]["total_s"] / res["bf16x3"]["total_s"]
    finally:
        for bb in bufs.values():
            bb.close()
    return res


def synthetic_rhf(n, nocc, seed=170, sigma=0.015, rank=8):
    """8-fold symmetric, positive semi-definite (pq|rs) = sum_k L[pq,k] L[rs,k]
    with L[pq] = L[qp] ~ N(0, sigma^2) (rank 8; sigma keeps the ring operator
    of the CCSD iteration contractive, so the problem converges), sorted
    orbital energies in the synthetic generator's ranges, and random orthogonal
    MO coefficients (AO input, transformed on the device)."""
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((n, n, rank)) * sigma
    L = 0.5 * (L + L.transpose(1, 0, 2))
    Lm = L.reshape(n * n, rank)
    eri = (Lm @ Lm.T).reshape(n, n, n, n)
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, nocc)), np.sort(rng.uniform(0.2, 3.0, n - nocc))])
    Cm = np.linalg.qr(rng.standard_normal((n, n)))[0]
    return eps, eri, Cm


def rhf_setup(capi, O, V, eps, eri, Cm, cpu_budget_s=5.0):
    """Context set-up from RHF AO data (ccsd_create_rhf) at n = O/2 + V/2
    spatial orbitals: device times of the AO->MO transform (4 GEMMs, 2 n^5
    flops each) and of the packing of the six spin-orbital blocks, the wall
    time of the whole create (host-to-device copy of the n^4 integrals
    included), and the oracle's quarter transform (numpy BLAS) on a bounded
    slab of the same tensor, extrapolated by flop count."""
    nocc, n = O // 2, O // 2 + V // 2
    t0 = time.perf_counter()
    ctx = capi.Context.from_rhf(eps, eri, nocc, mo_coeff=Cm, diis_dim=8)
    wall = time.perf_counter() - t0
    tm = ctx.setup_timings()
    ctx.close()
    flops = 4 * 2.0 * n ** 5
    # CPU: one quarter pass on row slabs of the (n^3 x n) view, as the oracle's numpy restatement would do it
    flat = eri.reshape(n ** 3, n)
    rows = min(n ** 3, 1 << 20)
    reps, el = 0, 0.0
    t0 = time.perf_counter()
    while el < cpu_budget_s and reps < 1000:
        _ = flat[:rows] @ Cm
        reps += 1
        el = time.perf_counter() - t0
    cpu_rate = 2.0 * rows * n * n * reps / el
    return {
        "n_spatial": n, "n_occ": nocc, "O": O, "V": V, "eri_bytes": eri.nbytes,
        "ao_to_mo_ms": tm["ao_to_mo"], "pack_ms": tm["pack"], "create_wall_s": wall,
        "ao_to_mo_tflops": flops / (tm["ao_to_mo"] * 1e-3) / 1e12 if tm["ao_to_mo"] > 0 else None,
        "ao_to_mo_flops": flops,
        "cpu_baseline": {"value": flops / cpu_rate, "unit": "s (AO->MO transform, extrapolated)", "kind": "port",
                         "sample": f"numpy BLAS quarter pass on a {rows} x {n} slab, {reps} reps in {el:.1f}s, "
                                   f"{cpu_rate / 1e9:.1f} GFLOP/s scaled to {flops:.3e} flop"},
    }


def closed_shell_flops(o, v):
    """GEMM flops of one closed-shell iteration (terms of rplan.hpp's rccsd_rhs
    above o^2 v^3): six ov x ov x ov ring products, the S/A pp ladder over
    packed pairs, the t1 part of the ladder (ov x o^2 x v^2), hh ladder and
    W_oooo, F_oo / F_vv, the two T1 o^2 v^3 terms."""
    ov, n1, n0 = o * v, v * (v + 1) // 2, v * (v - 1) // 2
    return (6 * 2.0 * ov ** 3 + 2.0 * (o * (o + 1) // 2) * n1 * n1 + 2.0 * (o * (o - 1) // 2) * n0 * n0
            + 2.0 * ov * o * o * v * v + 2 * 2.0 * o ** 4 * v * v + 2 * 2.0 * o ** 3 * v ** 2 + 2 * 2.0 * o ** 2 * v ** 3)


def rhf_closed_shell(capi, O, V, eps, eri, Cm, tol=1e-7, steps=5):
    """The RHF workload at n = O/2 + V/2 spatial orbitals through both paths:
    the closed-shell spin-adapted iteration (ccsd_rhf_*) and the spin-orbital
    one (ccsd_create_rhf, the O/V spin-orbital shape of the headline), each
    from the MP2 guess to |dE| < tol: seconds per iteration, time to converge,
    iterations and the energy difference."""
    nocc = O // 2
    o, v = nocc, V // 2
    t0 = time.perf_counter()
    cs = capi.RhfContext(eps, eri, nocc, mo_coeff=Cm, diis_dim=8)
    create = time.perf_counter() - t0
    cs.iterate(2)
    t0 = time.perf_counter()
    cs.iterate(steps)
    s_iter = (time.perf_counter() - t0) / steps
    cs.reset()
    r = cs.solve(tol, 200)
    cs.close()
    t0 = time.perf_counter()
    so = capi.Context.from_rhf(eps, eri, nocc, mo_coeff=Cm, diis_dim=8)
    create_so = time.perf_counter() - t0
    rs = so.solve(tol, 200)
    so.close()
    fl = closed_shell_flops(o, v)
    return {
        "workload": f"RHF input, n={o + v} spatial orbitals ({o} doubly occupied), synthetic rank-8 (pq|rs), AO basis",
        "closed_shell": {"s_per_iter": s_iter, "tflops_iteration": fl / s_iter / 1e12, "flops_per_iteration": fl,
                         "create_wall_s": create, "e_corr": r["e_corr"], "iterations": r["iterations"],
                         "seconds_total": r["seconds_total"]},
        "spin_orbital": {"s_per_iter": rs["seconds_per_iter"], "create_wall_s": create_so, "e_corr": rs["e_corr"],
                         "iterations": rs["iterations"], "seconds_total": rs["seconds_total"]},
        "abs_dE_corr": abs(r["e_corr"] - rs["e_corr"]),
        "speedup_time_to_converge": rs["seconds_total"] / r["seconds_total"],
        "tol": tol,
    }


def timed_steps(ctx, steps, warmup, precision):
    """Wall time per step and HIP-event kernel times averaged over the timed steps."""
    ctx.iterate(warmup, precision)
    acc = None
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.iterate(1, precision)
        tm = ctx.timings()
        acc = tm if acc is None else {k: acc[k] + tm[k] for k in tm}
    el = time.perf_counter() - t0
    return el / steps, {k: v / steps for k, v in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--O", type=int, default=40)
    ap.add_argument("--V", type=int, default=300)
    ap.add_argument("--seed", type=int, default=40300)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lowp", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-rhf", action="store_true")
    ap.add_argument("--only-rhf", action="store_true",
                    help="run only the RHF-input workload (set-up, closed-shell and spin-orbital solves)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (--gpus 1): this build has no multi-GPU path; "
              "the sharding across GPUs is a design only", file=sys.stderr)
        sys.exit(2)

    from ccsd_parallel_amd import capi
    O, V, seed = args.O, args.V, args.seed
    if args.only_rhf:
        eps_r, eri_r, C_r = synthetic_rhf(O // 2 + V // 2, O // 2)
        print(json.dumps({"rhf_setup": rhf_setup(capi, O, V, eps_r, eri_r, C_r, cpu_budget_s=0.5),
                          "rhf_closed_shell_1e-7": rhf_closed_shell(capi, O, V, eps_r, eri_r, C_r)}))
        return
    scale = 1.0 / np.sqrt(O * p2(V))
    fl = flops_per_iteration(O, V)
    total_flops = sum(fl.values())

    ctx = capi.Context.synthetic(O, V, seed, scale, precision="fp64", diis_dim=8)
    s_iter, tm = timed_steps(ctx, args.steps, args.warmup, args.precision)
    per_prec = {args.precision: {"s_per_iter": s_iter, "kernel_ms": tm}}
    if not args.no_lowp:
        for p in ("bf16x3", "fp32"):
            if p == args.precision:
                continue
            ctx.reset()
            s_p, tm_p = timed_steps(ctx, max(2, args.steps // 2), 1, p)
            per_prec[p] = {"s_per_iter": s_p, "kernel_ms": tm_p}
    ctx.close()
    for p, v in per_prec.items():
        ms = v["kernel_ms"]["pp_ladder"]
        v["ladder_tflops"] = fl["pp_ladder"] / (ms * 1e-3) / 1e12 if ms > 0 else None
    prec = args.precision
    ladder = per_prec[prec]["ladder_tflops"]
    out = {
        "metric": METRIC,
        "value": s_iter, "unit": "s/iteration", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": s_iter * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[prec], "data": "synthetic",
        "config": {"workload": f"CCSD amplitude update, O={O} V={V} spin orbitals, synthetic <pq||rs> seed {seed}, "
                               f"sigma=1/sqrt(O*P2(V))={scale:.3e}, DIIS(8)", "O": O, "V": V},
        "tflops_iteration": total_flops / s_iter / 1e12,
        "flops_per_iteration": total_flops,
        "per_precision": per_prec,
        "ladder_tflops": ladder,
        "roofline": {"bound": "mfma", "achieved": ladder, "peak": PEAK[prec], "unit": "TFLOP/s",
                     "frac": (ladder / PEAK[prec]) if ladder else None,
                     "traffic": ladder_traffic(O, V, prec), "traffic_unit": "bytes/launch",
                     "kernel": "pp ladder GEMM tau[P2(O) x P2(V)] x <ab||ef>^T (packed vvvv rows ab, k = ef)",
                     "algorithmic_flops": fl["pp_ladder"]},
    }

    if not args.no_converge:
        # FP64 oracle for the error statement: the FP64 GPU path converged to 1e-11
        c = capi.Context.synthetic(O, V, seed, scale, precision="fp64", diis_dim=8)
        e_tight = c.solve(1e-11, 100, allow_unconverged=True)["e_corr"]
        t1_tight, t2_tight = c.get_amplitudes()
        c.close()
        conv = {}
        for p in ("fp64", "fp32", "bf16x3"):
            c = capi.Context.synthetic(O, V, seed, scale, precision=p, diis_dim=8, final_fp64_sweeps=3)
            r = c.solve(1e-7, 100, allow_unconverged=True)
            conv[p] = {k: r[k] for k in ("e_corr", "iterations", "converged", "seconds_total", "lowp_iterations",
                                         "fp64_iterations")}
            t1, t2 = c.get_amplitudes()
            c.close()
            conv[p]["abs_dE_vs_fp64_1e-11"] = abs(conv[p]["e_corr"] - e_tight)
            conv[p]["max_dT2_vs_fp64_1e-11"] = float(np.abs(t2 - t2_tight).max())
            conv[p]["max_dT1_vs_fp64_1e-11"] = float(np.abs(t1 - t1_tight).max())
            del t1, t2
        del t1_tight, t2_tight
        for p in ("fp32", "bf16x3"):
            conv[p]["abs_dE_vs_fp64"] = abs(conv[p]["e_corr"] - conv["fp64"]["e_corr"])
            conv[p]["speedup_vs_fp64"] = conv["fp64"]["seconds_total"] / conv[p]["seconds_total"]
        out["time_to_converge_1e-7"] = conv
    if not args.no_converge:
        # plain FP64 configuration: O=20 / V=180, seed 20180, converged to 1e-10
        o2, v2, s2 = 20, 180, 20180
        c = capi.Context.synthetic(o2, v2, s2, 1.0 / np.sqrt(o2 * p2(v2)), precision="fp64", diis_dim=8)
        r = c.solve(1e-10, 200, allow_unconverged=True)
        c.close()
        out["fp64_o20_v180_1e-10"] = {k: r[k] for k in ("e_corr", "e_mp2", "iterations", "converged", "seconds_total",
                                                         "seconds_per_iter", "last_delta_e", "max_delta_t2")}
    if not args.no_adjoint:
        out["forward_adjoint_1e-7"] = combined_forward_adjoint(capi, O, V, seed, scale)
    if not args.no_rhf:
        eps_r, eri_r, C_r = synthetic_rhf(O // 2 + V // 2, O // 2)
        out["rhf_setup"] = rhf_setup(capi, O, V, eps_r, eri_r, C_r)
        out["rhf_closed_shell_1e-7"] = rhf_closed_shell(capi, O, V, eps_r, eri_r, C_r)
        del eri_r
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(O, V, seed, scale)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
