"""Benchmark: time-to-verified-1e-6 relative KKT on the 1,000,000-variable
transportation LP (1,000 sources x 1,000 sinks: m = 2,000, n = 1,000,000,
nnz = 2,000,000), on one MI355X.

One step = one full PDLP solve from the cold start x = clip(0, l, u), y = 0 to
max(eps_P, eps_D, eps_G) <= 1e-6, with the LP already resident in HBM (an
engine created once: upload, layout and SpMV autotuning are outside the timed
region).  Each result is certified afterwards by the host FP64 verifier on
the unscaled problem.  Prints ONE JSON line (see the keys below).

Besides the headline (mixed FP32-iterate / FP64-residual engine) it reports:
  * the same solve in the device-resident FP64 mode and in the plain FP64 mode
    (the reference GPU backend's design: host sync per attempt, host KKT);
  * the two iteration kernels' duration (HIP events, standalone launches at
    this LP's shapes) and achieved bandwidth from algorithmic bytes;
  * the reference's own CPU PDLP (oracle/_ref, 1 thread) on a bounded prefix of
    the same solve, extrapolated with the reference's full iteration count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "team-suprema_amd"))

import pdlp_amd as P  # noqa: E402

METRIC = "time-to-1e-6 rel KKT at 1M vars; batched LP solves/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
TOL = 1e-6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(lp, s, ell_AT, ell_A):
    """Algorithmic bytes of one launch of each iteration kernel (every array
    read or written once), value/vector size s (4: FP32 loop, 8: FP64 loop).
      primal half (rows of A^T): A^T indices + values, row pointers unless ELL,
        y once, then per column x, c, l, u read, xsum read+write, aty and xn
        written, and in the FP32 loop the (xn, xn - x) pair written
        -> nnz (4 + s) + [4 (n + 1)] + s m + 8 s n  [+ 8 n]
      dual half (rows of A): A indices + values, row pointers unless ELL, the
        gathered vector once (xn; the (xn, xn - x) pairs in the FP32 loop),
        then per row y, ax, b read, ysum read+write, axn and yn written, the
        inequality flag -> nnz (4 + s) + [4 (m + 1)] + (s | 8) n + 7 s m + m"""
    n, m, nnz = lp.n, lp.m, lp.nnz
    pair = s == 4
    primal = nnz * (4 + s) + (0 if ell_AT else 4 * (n + 1)) + s * m + 8 * s * n + (8 * n if pair else 0)
    dual = nnz * (4 + s) + (0 if ell_A else 4 * (m + 1)) + (8 if pair else s) * n + 7 * s * m + m
    return primal, dual


def cpu_baseline(lp, scaled, Dr, Dc, sample_attempts):
    """Reference CPU PDLP (oracle/_ref) for a bounded prefix of the solve,
    extrapolated with the full run's iteration count (tests/golden/
    c1_reference.json, written by the same reference build)."""
    try:
        sys.path.insert(0, ROOT)
        from oracle import Checker, available
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None
    kind = "ref" if available("ref") else "oracle"
    chk = Checker(kind)
    ref_path = os.path.join(ROOT, "tests", "golden", "c1_reference.json")
    full = None
    if os.path.exists(ref_path):
        with open(ref_path) as f:
            full = json.load(f)
    t0 = time.perf_counter()
    out = chk.pdlp(lp, scaled, Dr, Dc, sample_attempts, 64, TOL)
    dt = time.perf_counter() - t0
    per_attempt = dt / max(1, out.iterations)
    attempts = full["iterations"] if full else None
    value = per_attempt * attempts if attempts else None
    return {
        "value": value, "unit": "s", "cores": 1,
        "kind": "reference" if kind == "ref" else "port",
        "sample": (f"first {out.iterations} attempts ({out.iterations // 64} host KKT checks) of the "
                   f"{'reference' if kind == 'ref' else 'oracle'} CPU PDLP on this LP: {dt:.2f} s, "
                   f"{per_attempt * 1e3:.2f} ms/attempt, extrapolated to the full reference run's "
                   f"{attempts} attempts"),
        "host_cpus": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=256, help="attempts of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the standalone iteration-kernel timing (for profiling it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(instances shard across GPUs without communication; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    if P.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)

    log("generating transport LP 1000x1000 and scaling (host, untimed)")
    lp = P.gen_transport(1000, 1000, seed=42)
    scaled, Dr, Dc = P.scale(lp)

    if args.kernels_only:
        kb = P.bench_kernels(lp, scaled, Dr, Dc, reps=200)
        print(json.dumps({"kernels_ms": kb}))
        return

    results = {}
    for name, mode in (("mixed", P.MODE_MIXED), ("fast_fp64", P.MODE_FAST_FP64)):
        t0 = time.perf_counter()
        eng = P.Engine(lp, scaled, Dr, Dc, tol=TOL, mode=mode)
        setup_s = time.perf_counter() - t0
        for w in range(args.warmup):
            eng.solve()
        times, iters = [], []
        for k in range(args.steps):
            t0 = time.perf_counter()
            res = eng.solve()
            times.append(time.perf_counter() - t0)
            iters.append(res.iterations)
            rep = P.verify(lp, res.x, res.y)
            worst = max(rep.eps_P, rep.eps_D, rep.eps_G)
            if not res.converged or worst > TOL:
                raise SystemExit(f"{name}: step {k} not verified (converged={res.converged}, "
                                 f"eps={worst:.3e})")
            log(f"{name} step {k}: {times[-1] * 1e3:.2f} ms, {res.iterations} attempts "
                f"(fp32 {res.info['iterations_fp32']}), verified eps {worst:.2e}, obj {rep.objective:.6f}")
        eng.close()
        results[name] = dict(seconds=float(np.mean(times)), min_seconds=float(np.min(times)),
                             iterations=int(np.median(iters)), objective=rep.objective,
                             eps=[rep.eps_P, rep.eps_D, rep.eps_G], setup_seconds=setup_s,
                             fp32_iterations=res.info["iterations_fp32"])

    plain = None
    if not args.no_plain:
        log("plain FP64 mode (host-driven, reference GPU design): one solve")
        t0 = time.perf_counter()
        res = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=P.MODE_PLAIN_FP64)
        wall = time.perf_counter() - t0
        rep = P.verify(lp, res.x, res.y)
        plain = dict(seconds=res.info["seconds_solve"], wall_with_setup=wall,
                     iterations=res.iterations, converged=res.converged,
                     eps=[rep.eps_P, rep.eps_D, rep.eps_G], objective=rep.objective)
        log(f"plain: {plain['seconds']:.3f} s solve, {res.iterations} attempts")

    log("timing the iteration kernels (HIP events, 200 launches each)")
    kb = P.bench_kernels(lp, scaled, Dr, Dc, reps=200)
    prim32, dual32 = kernel_bytes(lp, 4, kb["ell_AT"] > 0, kb["ell_A"] > 0)
    prim64, dual64 = kernel_bytes(lp, 8, kb["ell_AT"] > 0, kb["ell_A"] > 0)
    kernels = {
        "k_primal_f32": dict(ms=kb["primal_ms_fp32"], bytes=prim32),
        "k_dual_f32": dict(ms=kb["dual_ms_fp32"], bytes=dual32),
        "k_primal_f64": dict(ms=kb["primal_ms_fp64"], bytes=prim64),
        "k_dual_f64": dict(ms=kb["dual_ms_fp64"], bytes=dual64),
    }
    for k in kernels.values():
        k["GBps"] = k["bytes"] / (k["ms"] * 1e-3) / 1e9
        k["frac"] = k["GBps"] / HBM_PEAK_GBS
    dom_name = max(("k_primal_f32", "k_dual_f32"), key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: reference CPU PDLP, {args.cpu_sample} attempts")
        cpu = cpu_baseline(lp, scaled, Dr, Dc, args.cpu_sample)

    head = results["mixed"]
    best_name = min(results, key=lambda k: results[k]["seconds"])
    out = {
        "metric": METRIC,
        "value": head["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["seconds"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 iterates/SpMV + f64 statistics, KKT, restarts",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "transportation LP 1000x1000: m=2000, n=1000000, nnz=2000000, "
                        "tol 1e-6 relative KKT (verified), cold start, LP resident in HBM",
            "mode": "mixed",
            "attempts": head["iterations"],
            "fp32_attempts": head["fp32_iterations"],
            "objective": head["objective"],
            "verified_eps": head["eps"],
            "engine_setup_seconds": head["setup_seconds"],
        },
        "modes": results,
        "fastest_mode": best_name,
        "plain_fp64": plain,
        "speedup_vs_plain": (plain["seconds"] / head["seconds"]) if plain else None,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": dom["GBps"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["frac"],
            "traffic": None,
            "algorithmic_bytes_per_launch": dom["bytes"],
            "ms_per_launch": dom["ms"],
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
