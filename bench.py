"""Benchmark: time-to-verified-1e-6 relative KKT on the 1,000,000-variable
transportation LP (1,000 sources x 1,000 sinks: m = 2,000, n = 1,000,000,
nnz = 2,000,000), on one MI355X.

One step = one full PDLP solve from the cold start x = clip(0, l, u), y = 0 to
max(eps_P, eps_D, eps_G) <= 1e-6, with the LP already resident in HBM (an
engine created once: upload and layout are outside the timed region).  Each
result is certified afterwards by the host FP64 verifier on the unscaled
problem.  Prints ONE JSON line (see the keys below).

The headline is the device-resident FP64 mode (the reference's arithmetic
width); beside it:
  * the mixed FP32-iterate / FP64-residual mode;
  * "one_shot": the drop-in call supr_solve_pdhg_gpu (upload, layout and
    teardown included, inputs in host memory) in each mode, against the plain
    FP64 mode's one-shot call (the reference GPU backend's design: host sync
    per attempt, host KKT);
  * the two iteration kernels' duration (HIP events, standalone launches at
    this LP's shapes) and achieved bandwidth from algorithmic bytes, with the
    rocprofv3 average of the same kernels from the committed profile;
  * the reference's own CPU PDLP (oracle/_ref, 1 thread) run to completion on
    the same LP;
  * "batched": 512 blending LPs (20-400 components x 10-200 products,
    log-uniform sizes, n 230 to 73,117) solved together in the mixed mode
    with the tuned layout, every result verified; solves/s against
    sequential plain-FP64 and sequential resident-mixed solves of a
    stratified sample of the same instances;
  * "serving": an open-loop Poisson stream of random LPs (6 nonzeros per
    column, m/n in [0.2, 0.8], nnz log-uniform over [1e4, 2e7]) through the
    continuous-batching server, every result verified; p50/p99 latency and
    solves/s, against one-at-a-time plain FP64 solves of the same requests
    (their latency from a FIFO queue over the same arrivals); "serving-solve":
    the server's runs of the narrow stream alone (for tracing its kernels);
  * "c0": the random equality LP m = 2,000, n = 5,000 in each mode, beside the
    reference's own CPU PDLP run to completion;
  * "large": the power-law LP m = 500,000, n = 2,000,000 (35.1M nonzeros):
    its kernels and one FP64 solve to a verified 1e-6;
  * "polish_1e-8": the 1M-variable LP to a verified 1e-8 (the FP64 polish);
  * each mode's objective error against the reference CPU PDLP's;
  * "ingest": the same 1M-variable LP as an MPS file, read by this library's
    reader and by the reference's (same LP bit for bit), and file -> verified
    solution through solve_lp_gpu;
  * "prep": scale() on the device against the host version and the
    reference's, bit for bit, at 2M and 7M nonzeros.
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


def nonneg_bounds(lp, s):
    """Every l_j is +0 and every u_j +inf in the loop's precision (size s):
    the single-LP primal kernel then reads neither l nor u."""
    dt = np.float32 if s == 4 else np.float64
    l, u = np.asarray(lp.l).astype(dt), np.asarray(lp.u).astype(dt)
    return bool(np.all((l == 0) & ~np.signbit(l)) and np.all(np.isposinf(u)))


def dual_form(rows, nnz, lanes, ell):
    """The dual kernel's third template argument for a launch over A with
    `rows` rows, `nnz` entries and `lanes` lanes per row (the engine's rule:
    Solver::one_pass_dual and launch_dual): 1 the one-pass form (a mean row
    fits one prefetched group of 4 per lane), 2 the flat one-pass form (long
    CSR rows, at most 1,024 blocks), 0 the grid-stride kernel."""
    staged = lanes >= 2 and rows >= (1 << 17)
    if not (lanes <= 64 and not staged and rows <= 2048 * (256 // lanes)):
        return 0
    if nnz <= 4 * lanes * rows:
        return 1
    g = min(max(1, -(-rows // (256 // lanes))), 2048)
    if g >= 8:
        g = -(-g // 8) * 8
    return 2 if not ell and g <= 1024 else 0


def primal_kernel(n, lanes, ell, fp64):
    """The rocprof name of the primal kernel for an A^T with n rows, more
    than one pass of the 2,048-block grid holds at `lanes` lanes per row (the
    1M-variable transportation LP's; the engine's rule: primal_view and
    launch_primal): FP64 row pairs that fit one pass of pairs run
    k_primal_pairs, the rest k_primal's grid-stride form (template flag
    false)."""
    pairs = ell and lanes == 1 and n % 2 == 0 and n > 2048 * 256
    if fp64 and pairs and n <= 2 * 2048 * 256:
        return "k_primal_pairs<double>"
    return f"k_primal<{lanes}, {'double' if fp64 else 'float'}, false>"


def kernel_bytes(lp, s, ell_AT, ell_A, nonneg=False, store_aty=False, rebuilt=False, compressed=True):
    """Algorithmic bytes of one launch of each iteration kernel (every array
    read or written once), value/vector size s (4: FP32 loop, 8: FP64 loop).
      primal half (rows of A^T): A^T indices + values, row pointers unless ELL,
        y once, then per column x, c, l, u read (l, u not when nonneg: every
        l = +0, u = +inf), xsum read+write, xn written (and A^T y written where
        it is stored: the batched engine), and in the FP32 loop the (xn, xn - x)
        pair written
        -> nnz (4 + s) + [4 (n + 1)] + s m + (7 | 5) s n [+ s n] [+ 8 n]
      dual half (rows of A): A indices + values, row pointers unless ELL, the
        gathered vector once (xn; the (xn, xn - x) pairs in the FP32 loop),
        then per row y, ax, b read, ysum read+write, axn and yn written, the
        inequality flag -> nnz (4 + s) + [4 (m + 1)] + (s | 8) n + 7 s m + m
    An ELL index is 2 bytes instead of 4 where the gathered vector has at most
    65,536 entries (the library's 16-bit ELL indices).  rebuilt: the FP64
    primal half forms A^T's values from Dc_j (once per column) and Dr_i a (an
    m-vector read once) instead of reading 8 bytes per nonzero.
    compressed=False: the same arrays with 4-byte indices and s-byte values
    read per nonzero (the layout before those two compressions)."""
    n, m, nnz = lp.n, lp.m, lp.nnz
    pair = s == 4
    iAT = 2 if compressed and ell_AT and m <= 65536 else 4
    iA = 2 if compressed and ell_A and n <= 65536 else 4
    vals_AT = 8 * n + 8 * m if (rebuilt and compressed) else nnz * s
    primal = (nnz * iAT + vals_AT + (0 if ell_AT else 4 * (n + 1)) + s * m + (5 if nonneg else 7) * s * n
              + (s * n if store_aty else 0) + (8 * n if pair else 0))
    dual = nnz * (iA + s) + (0 if ell_A else 4 * (m + 1)) + (8 if pair else s) * n + 7 * s * m + m
    return primal, dual


def moved_bytes(lp, attempts, attempts_fp32):
    """Algorithmic bytes the iteration kernels move over a solve of lp (CSR
    layout, as in the batched and served engines): FP32-loop attempts at the
    FP32 figure, the rest at the FP64 one."""
    p32, d32 = kernel_bytes(lp, 4, False, False, store_aty=True)
    p64, d64 = kernel_bytes(lp, 8, False, False, store_aty=True)
    return attempts_fp32 * (p32 + d32) + (attempts - attempts_fp32) * (p64 + d64)


def cpu_per_attempt(U, S, Dr, Dc, sample_attempts):
    """Seconds per attempt of the reference's CPU PDLP (oracle/_ref, one
    thread; the oracle's restatement where _ref is absent) over a bounded
    prefix of the solve, and the kind of engine timed."""
    sys.path.insert(0, ROOT)
    from oracle import Checker, available
    kind = "ref" if available("ref") else "oracle"
    t0 = time.perf_counter()
    out = Checker(kind).pdlp(U, S, Dr, Dc, sample_attempts, 64, TOL)
    return (time.perf_counter() - t0) / max(1, out.iterations), ("reference" if kind == "ref" else "port")


def cpu_baseline(lp, scaled, Dr, Dc):
    """The reference's own CPU PDLP (oracle/_ref: src/pdlp.cpp + include/
    pdlp_algo.hpp compiled from the reference's sources; the oracle's
    restatement where _ref is absent), one thread, run to completion on this
    LP from the same scaled form, timed around the engine call only."""
    try:
        sys.path.insert(0, ROOT)
        from oracle import Checker, available
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None
    kind = "ref" if available("ref") else "oracle"
    chk = Checker(kind)
    t0 = time.perf_counter()
    out = chk.pdlp(lp, scaled, Dr, Dc, 200000, 64, TOL)
    dt = time.perf_counter() - t0
    rep = P.verify(lp, out.x, out.y)
    return {
        "value": dt, "unit": "s", "cores": 1,
        "kind": "reference" if kind == "ref" else "port",
        "sample": (f"the whole {'reference' if kind == 'ref' else 'oracle'} CPU PDLP solve of this LP to "
                   f"1e-6 ({out.iterations} attempts, converged={out.converged}), one thread"),
        "attempts": out.iterations, "objective": rep.objective,
        "verified_eps": [rep.eps_P, rep.eps_D, rep.eps_G],
        "host_cpus": os.cpu_count(),
    }


def bench_c0():
    """The reference's CPU-runnable case: the random equality LP m = 2,000,
    n = 5,000 (6 nonzeros per column), solved to 1e-6 on the GPU in each mode
    (verified) and by the reference's own CPU PDLP (oracle/_ref, one thread,
    to completion)."""
    lp = P.gen_random_lp(2000, 5000, 6, seed=42)
    s, Dr, Dc = P.scale(lp)
    out = {"workload": "random equality LP m=2000 n=5000 nnz=30000 (seed 42), tol 1e-6, verified"}
    for name, mode in (("mixed", P.MODE_MIXED), ("fast_fp64", P.MODE_FAST_FP64)):
        eng = P.Engine(lp, s, Dr, Dc, tol=TOL, mode=mode)
        eng.solve()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = eng.solve()
            ts.append(time.perf_counter() - t0)
        eng.close()
        rep = P.verify(lp, r.x, r.y)
        if not r.converged or max(rep.eps_P, rep.eps_D, rep.eps_G) > TOL:
            raise SystemExit(f"c0 {name}: not verified")
        out[name] = dict(seconds=float(np.mean(ts)), attempts=r.iterations, objective=rep.objective)
    t0 = time.perf_counter()
    r = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=P.MODE_PLAIN_FP64)
    out["plain_fp64"] = dict(seconds=time.perf_counter() - t0, attempts=r.iterations)
    try:
        sys.path.insert(0, ROOT)
        from oracle import Checker, available
        kind = "ref" if available("ref") else "oracle"
        t0 = time.perf_counter()
        ref = Checker(kind).pdlp(lp, s, Dr, Dc, 200000, 64, TOL)
        dt = time.perf_counter() - t0
        ref_obj = P.verify(lp, ref.x, ref.y).objective
        out["cpu_reference"] = dict(seconds=dt, attempts=ref.iterations, cores=1, objective=ref_obj,
                                    kind="reference" if kind == "ref" else "port")
        for name in ("mixed", "fast_fp64"):
            out[name]["objective_rel_error_vs_cpu"] = abs(out[name]["objective"] - ref_obj) / (1 + abs(ref_obj))
    except Exception as e:  # pragma: no cover
        log(f"c0 CPU run unavailable: {e}")
    log(f"c0: mixed {out['mixed']['seconds'] * 1e3:.1f} ms, fast64 {out['fast_fp64']['seconds'] * 1e3:.1f} ms, "
        f"plain {out['plain_fp64']['seconds']:.2f} s, CPU reference "
        f"{out.get('cpu_reference', {}).get('seconds', float('nan')):.2f} s")
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_transport_1000x1000.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        k = json.loads(block)["kernels"][kernel]
        return k["read_bytes"] + k["write_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def copy_rates():
    """GB/s (bytes read + written) of the plain stream copy k_copy4 at
    1 GiB each way (HBM) and at 64 MB and 16 MB each way (held in the 256 MB
    last-level cache between launches): the box's attainable streaming
    rates, for context beside the 8 TB/s peak."""
    out = {}
    for label, nb, reps in (("1GiB", 1 << 30, 10), ("64MiB", 64 << 20, 100), ("16MiB", 16 << 20, 200)):
        ms = P.bench_copy(nb, reps=reps)
        out[label] = 2 * nb / (ms * 1e-3) / 1e9
    return out


def _pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=float), q))


def blend_batch(count=512):
    sizes = P.gen_blend_sizes(count, seed=2026)
    lps, forms = [], []
    for i, (K, Pp) in enumerate(sizes):
        lp = P.gen_blend(K, Pp, seed=10000 + i)
        form, meta = P.build_pdlp_form(lp)
        s, Dr, Dc = P.scale(form)
        lps.append((lp, meta))
        forms.append((form, s, Dr, Dc))
    return lps, forms


def bench_batched_solve(count=512):
    """The batch alone, its replays launched kernel by kernel (the run to
    trace the batched kernels; same bits as the graph replays): the layout
    the tuner picks in the full run (lanes bias 0.5, block cap 512)."""
    lps, forms = blend_batch(count)
    with P.BatchEngine(forms, tol=TOL, mode=P.MODE_MIXED, tune=False, lanes_bias=0.5, max_blocks=512,
                       graph_attempts=-16) as eng:
        res, secs = eng.solve()
    conv = sum(r.converged for r in res)
    log(f"batched solve (direct launches): {count} LPs in {secs:.3f} s, {conv} converged")
    return {"instances": count, "seconds": secs, "converged": conv,
            "attempts_max": int(max(r.iterations for r in res))}


def bench_batched(count=512, sample=8):
    """Batched blending LPs (mixed mode, tuned layout) against sequential
    single-LP solves of a stratified sample of the same instances."""
    lps, forms = blend_batch(count)
    t0 = time.perf_counter()
    eng = P.BatchEngine(forms, tol=TOL, mode=P.MODE_MIXED, tune=True)
    setup = time.perf_counter() - t0
    res, secs = eng.solve()
    cfg = eng.config
    eng.close()
    worst = 0.0
    for (lp, meta), r in zip(lps, res):
        x, y = P.extract_solution(meta, r.x, r.y)
        rep = P.verify(lp, x, y)
        worst = max(worst, rep.eps_P, rep.eps_D, rep.eps_G)
    conv = sum(r.converged for r in res)
    if conv != count or worst > TOL:
        raise SystemExit(f"batched: {conv}/{count} converged, worst verified eps {worst:.3e}")
    order = np.argsort([f[0].nnz for f in forms])
    pick = [int(order[i]) for i in range(count // (2 * sample), count, count // sample)]
    seq = {}
    for label, mode in (("plain_fp64", P.MODE_PLAIN_FP64), ("resident_mixed", P.MODE_MIXED)):
        ts = []
        for i in pick:
            f, s, Dr, Dc = forms[i]
            t0 = time.perf_counter()
            r = P.solve_pdhg_gpu(f, s, Dr, Dc, tol=TOL, mode=mode)
            ts.append(time.perf_counter() - t0)
            if not r.converged:
                raise SystemExit(f"batched baseline {label}: instance {i} did not converge")
        seq[label] = dict(mean_seconds=float(np.mean(ts)), solves_per_s=1.0 / float(np.mean(ts)))
    sps = count / secs
    moved = sum(moved_bytes(f[0], r.iterations, r.info["iterations_fp32"]) for f, r in zip(forms, res))
    cpu = None
    try:
        est = []
        for i in pick:
            per, kind = cpu_per_attempt(*forms[i], min(3000, res[i].iterations))
            est.append(per * res[i].iterations)
        cores = os.cpu_count() or 1
        cpu = dict(kind=kind, cores=1, solves_per_s=1.0 / float(np.mean(est)),
                   all_cores_bound_solves_per_s=cores / float(np.mean(est)), host_cpus=cores,
                   sample="the sequential sample's instances: the CPU PDLP timed over up to 3,000 attempts each, "
                          "scaled by the instance's attempt count in the batch; all_cores_bound = host_cpus x "
                          "the one-thread rate (independent instances, no contention)")
    except Exception as e:  # pragma: no cover
        log(f"batched: CPU baseline unavailable: {e}")
    log(f"batched: {count} LPs in {secs:.3f} s = {sps:.1f} solves/s; sequential plain "
        f"{seq['plain_fp64']['solves_per_s']:.2f}/s, resident mixed {seq['resident_mixed']['solves_per_s']:.2f}/s"
        + (f", CPU 1 thread {cpu['solves_per_s']:.2f}/s" if cpu else "")
        + f"; {moved / secs / 1e9:.0f} GB/s of algorithmic bytes")
    return {
        "workload": f"{count} blending LPs (K 20-400 x P 10-200 log-uniform, seed 2026), mixed mode, tol 1e-6, "
                    "every result verified on the original LP",
        "instances": count, "total_nnz": int(sum(f[0].nnz for f in forms)),
        "seconds": secs, "solves_per_s": sps, "setup_seconds": setup,
        "layout": {"lanes_bias": cfg[0], "max_blocks": cfg[1]},
        "attempts_max": int(max(r.iterations for r in res)),
        "attempts_median": int(np.median([r.iterations for r in res])),
        "worst_verified_eps": worst,
        "sequential_sample": {"instances": pick, **seq},
        "speedup_vs_sequential_plain": sps / seq["plain_fp64"]["solves_per_s"],
        "speedup_vs_sequential_resident_mixed": sps / seq["resident_mixed"]["solves_per_s"],
        "roofline": {"bound": "hbm", "achieved": moved / secs / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": moved / secs / 1e9 / HBM_PEAK_GBS,
                     "note": "algorithmic bytes of every instance's attempts (CSR, FP32 / FP64 loop) over the batch time"},
        "cpu_baseline": cpu,
    }


def bench_serving(count=10, qps=1.5, nnz_hi=1e6, seed=7, plain_sample_attempts=None, span=False,
                  saturated=True, baselines=True, server_options=None):
    """Open-loop Poisson stream through the server against one-at-a-time
    plain FP64 solves of the same requests.  plain_sample_attempts: time the
    plain solves on that many attempts each and scale by the request's
    attempt count in the server (for requests too large to solve plainly in
    a bench run; the output says so).  span: the requests' nnz are spread
    log-evenly over [1e4, nnz_hi] (both ends included), in random order.
    saturated: also serve every request submitted at once.  baselines:
    also the plain FP64 and CPU baselines (off: the server's runs alone, for
    tracing its kernels).  server_options: more P.Server options."""
    import threading
    rng = np.random.default_rng(seed)
    if span:
        nnzs = rng.permutation(np.geomspace(1e4, nnz_hi, count))
    else:
        nnzs = np.exp(rng.uniform(np.log(1e4), np.log(nnz_hi), count))
    ratios = rng.uniform(0.2, 0.8, count)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, count))
    reqs = []
    for i in range(count):
        n = max(20, int(nnzs[i] / 6))
        m = max(6, int(ratios[i] * n))
        lp = P.gen_random_lp(m, n, 6, seed=1000 + i)
        s, Dr, Dc = P.scale(lp)
        reqs.append((lp, s, Dr, Dc))
        if plain_sample_attempts:
            log(f"serving: request {i} m={m} n={n} nnz={lp.nnz}")
    caps = dict(max_cols=1 << 24, max_rows=1 << 24, max_nnz=1 << 27, max_instances=64, max_blocks=1024)

    def serve(open_loop):
        with P.Server(mode=P.MODE_MIXED, tol=TOL, **caps, **(server_options or {})) as srv:
            tickets, done, res = [None] * count, [None] * count, [None] * count
            lock = threading.Lock()
            t0 = time.perf_counter()

            def submitter():
                for i in range(count):
                    if open_loop:
                        d = t0 + arrivals[i] - time.perf_counter()
                        if d > 0:
                            time.sleep(d)
                    tk = srv.submit(*reqs[i])
                    with lock:
                        tickets[i] = tk

            th = threading.Thread(target=submitter)
            th.start()
            nd, last = 0, time.perf_counter()
            while nd < count:
                running, queued, _ = srv.step()
                if time.perf_counter() - last > 20.0:
                    log(f"serving: {nd}/{count} done, {running} running, {queued} queued")
                    last = time.perf_counter()
                with lock:
                    tk = list(tickets)
                for i in range(count):
                    if tk[i] is not None and done[i] is None:
                        r = srv.result(tk[i])
                        if r is not None:
                            done[i] = time.perf_counter() - t0
                            res[i] = r
                            nd += 1
            th.join()
            dev = srv.device_bytes
        worst = 0.0
        for i in range(count):
            rep = P.verify(reqs[i][0], res[i].x, res[i].y)
            worst = max(worst, rep.eps_P, rep.eps_D, rep.eps_G)
        if not all(r.converged for r in res) or worst > TOL:
            raise SystemExit(f"serving: not all requests verified (worst eps {worst:.3e})")
        done = np.array(done)
        lat = done - (arrivals if open_loop else 0.0)
        moved = sum(moved_bytes(reqs[i][0], res[i].iterations, res[i].info["iterations_fp32"])
                    for i in range(count))
        return dict(makespan=float(done.max()), solves_per_s=count / float(done.max()),
                    latency_p50=_pct(lat, 50), latency_p99=_pct(lat, 99),
                    worst_verified_eps=worst, device_bytes=dev,
                    attempts=[int(r.iterations) for r in res],
                    achieved_GBps=moved / float(done.max()) / 1e9,
                    frac_of_hbm_peak=moved / float(done.max()) / 1e9 / HBM_PEAK_GBS)

    ol = serve(True)
    sat = serve(False) if saturated else None
    base = sat if sat is not None else ol
    if not baselines:
        log(f"serving: " + (f"saturated {sat['solves_per_s']:.2f} solves/s; " if sat else "")
            + f"open loop {qps} qps p50 {ol['latency_p50']:.2f} s p99 {ol['latency_p99']:.2f} s")
        return {"request_nnz": [int(r[0].nnz) for r in reqs], "saturated": sat, "open_loop": ol}
    svc = []
    for i in range(count):
        t0 = time.perf_counter()
        if plain_sample_attempts:
            r = P.solve_pdhg_gpu(*reqs[i], tol=TOL, mode=P.MODE_PLAIN_FP64,
                                 max_iterations=plain_sample_attempts)
            per = (time.perf_counter() - t0) / max(1, r.iterations)
            svc.append(per * base["attempts"][i])
        else:
            r = P.solve_pdhg_gpu(*reqs[i], tol=TOL, mode=P.MODE_PLAIN_FP64)
            svc.append(time.perf_counter() - t0)
            if not r.converged:
                raise SystemExit(f"serving baseline: request {i} did not converge")
    end, lat = 0.0, []
    for i in range(count):
        end = max(arrivals[i], end) + svc[i]
        lat.append(end - arrivals[i])
    plain = dict(total_seconds=float(np.sum(svc)), solves_per_s=count / float(np.sum(svc)),
                 latency_p50=_pct(lat, 50), latency_p99=_pct(lat, 99),
                 measured=("full solves" if not plain_sample_attempts else
                           f"estimated: {plain_sample_attempts} timed attempts per request x the "
                           "request's attempt count in the server"))
    cpu = None
    try:
        est = []
        for i in range(count):
            per, kind = cpu_per_attempt(*reqs[i], min(300, base["attempts"][i]))
            est.append(per * base["attempts"][i])
        cores = os.cpu_count() or 1
        cpu = dict(kind=kind, cores=1, solves_per_s=count / float(np.sum(est)),
                   all_cores_bound_solves_per_s=cores * count / float(np.sum(est)), host_cpus=cores,
                   sample="each request's CPU PDLP timed over up to 300 attempts, scaled by its attempt count "
                          "in the server; all_cores_bound = host_cpus x the one-thread rate")
    except Exception as e:  # pragma: no cover
        log(f"serving: CPU baseline unavailable: {e}")
    log(f"serving: " + (f"saturated {sat['solves_per_s']:.2f} solves/s; " if sat else "")
        + f"open loop {qps} qps {ol['solves_per_s']:.3f} solves/s p50 "
        f"{ol['latency_p50']:.2f} s p99 {ol['latency_p99']:.2f} s; plain one-at-a-time "
        f"{plain['solves_per_s']:.3f} solves/s, p50 {plain['latency_p50']:.1f} s p99 {plain['latency_p99']:.1f} s")
    spread = "log-evenly spread over" if span else "log-uniform in"
    return {
        "workload": f"{count} random LPs (6 nnz/column, m/n U[0.2,0.8], nnz {spread} [1e4, {nnz_hi:.0e}], "
                    f"seed {seed}), Poisson arrivals at {qps} qps, mixed mode, tol 1e-6, every result verified",
        "request_nnz": [int(r[0].nnz) for r in reqs],
        "total_nnz": int(sum(r[0].nnz for r in reqs)),
        "saturated": sat, "open_loop": ol, "plain_fp64_one_at_a_time": plain,
        "speedup_solves_per_s_vs_plain": base["solves_per_s"] / plain["solves_per_s"],
        "speedup_basis": "saturated" if sat else "open loop (makespan of the stream)",
        "cpu_baseline": cpu,
    }


def bench_polish(lp, scaled, Dr, Dc, tol=1e-8):
    """The 1M-variable LP to a tighter verified tolerance: the mixed mode's
    FP32 loop hands over to the FP64 loop (the polish), which restores 1e-8
    relative KKT; the device-resident FP64 mode beside it."""
    out = {"workload": f"transportation LP 1000x1000, tol {tol:g} relative KKT (verified), LP resident in HBM"}
    for name, mode in (("mixed", P.MODE_MIXED), ("fast_fp64", P.MODE_FAST_FP64)):
        with P.Engine(lp, scaled, Dr, Dc, tol=tol, mode=mode, max_iterations=500000) as eng:
            eng.solve()
            t0 = time.perf_counter()
            r = eng.solve()
            dt = time.perf_counter() - t0
        rep = P.verify(lp, r.x, r.y)
        worst = max(rep.eps_P, rep.eps_D, rep.eps_G)
        if not r.converged or worst > tol:
            raise SystemExit(f"polish {name}: not verified at {tol:g} (eps {worst:.3e})")
        out[name] = dict(seconds=dt, attempts=r.iterations, fp32_attempts=r.info["iterations_fp32"],
                         verified_eps=[rep.eps_P, rep.eps_D, rep.eps_G], objective=rep.objective)
    log(f"polish to {tol:g}: mixed {out['mixed']['seconds'] * 1e3:.1f} ms ({out['mixed']['attempts']} attempts, "
        f"{out['mixed']['fp32_attempts']} FP32), fast64 {out['fast_fp64']['seconds'] * 1e3:.1f} ms")
    return out


def bench_ingest(S=1000, D=1000):
    """MPS ingestion of the 1M-variable transportation LP (the input format of
    the reference's published 48.4 s / 8.1 s run: tools/gen_lp.cpp writes it,
    src/solve.cpp:150-232 reads and solves it): the LP written as MPS with
    gen_lp's names (x_i_j, S<i>, D<j>), read back by this library's reader
    and by the reference's own reader (oracle/_ref, one thread), checked to
    be the same LP bit for bit; then file -> verified solution: read_mps +
    solve_lp_gpu (form, scale() on the device, the mixed solve, extract,
    verify on the LP; the reference's solve_mps without presolve)."""
    import shutil
    import tempfile
    lp = P.gen_transport(S, D, seed=42)
    cols = [f"x_{i}_{j}" for i in range(S) for j in range(D)]
    rows = [f"S{i}" for i in range(S)] + [f"D{j}" for j in range(D)]
    d = tempfile.mkdtemp(prefix="supr_mps_")
    try:
        path = os.path.join(d, f"transport_{S}x{D}.mps")
        t0 = time.perf_counter()
        P.write_mps(path, lp, col_names=cols, row_names=rows)
        t_write = time.perf_counter() - t0
        size = os.path.getsize(path)
        def best_read(threads=None):
            if threads is not None:
                os.environ["SUPR_MPS_THREADS"] = str(threads)
            try:
                ts = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    mp = P.read_mps(path, names=False)
                    ts.append(time.perf_counter() - t0)
            finally:
                os.environ.pop("SUPR_MPS_THREADS", None)
            return mp, float(min(ts))
        mp1, t_one = best_read(1)
        mp, t_read = best_read()
        ours = mp.lp
        if not all(np.asarray(getattr(ours, k)).tobytes() == np.asarray(getattr(mp1.lp, k)).tobytes()
                   for k in ("row", "col", "val", "c", "rL", "rU", "l", "u")):
            raise SystemExit("ingest: the threaded COLUMNS pass differs from the one-thread parse")
        out = {"workload": f"transportation LP {S}x{D} as MPS ({size / 1e6:.1f} MB, gen_lp names)",
               "file_bytes": size, "write_seconds": t_write, "read_seconds": t_read,
               "read_MBps": size / t_read / 1e6, "reader_threads": min(16, os.cpu_count() or 1),
               "read_seconds_one_thread": t_one, "m": ours.m, "n": ours.n, "nnz": ours.nnz}
        try:
            sys.path.insert(0, ROOT)
            from oracle import available, read_mps_ref
            if available("ref"):
                t0 = time.perf_counter()
                r = read_mps_ref(path)
                t_ref = time.perf_counter() - t0
                same = all(np.asarray(getattr(ours, k)).tobytes() == np.asarray(r[k]).tobytes()
                           for k in ("row", "col", "val", "c", "rL", "rU", "l", "u"))
                if not same:
                    raise SystemExit("ingest: the MPS reader differs from the reference reader")
                out["reference_read_seconds"] = t_ref
                out["speedup_vs_reference_reader"] = t_ref / t_read
                out["identical_to_reference_reader"] = True
        except ImportError as e:   # pragma: no cover
            log(f"ingest: reference reader unavailable: {e}")
        P.solve_lp_gpu(ours, tol=TOL)   # warm-up (device initialisation, kernel loads)
        t0 = time.perf_counter()
        mp = P.read_mps(path, names=False)
        t_e2e_read = time.perf_counter() - t0
        sol = P.solve_lp_gpu(mp.lp, tol=TOL)
        e2e = time.perf_counter() - t0
        worst = max(sol.report.eps_P, sol.report.eps_D, sol.report.eps_G)
        if sol.status != "optimal" or worst > TOL:
            raise SystemExit(f"ingest: file -> solution not verified ({sol.status}, eps {worst:.3e})")
        out["file_to_verified_solution_seconds"] = e2e
        out["file_to_solution"] = {"status": sol.status, "attempts": sol.iterations,
                                   "objective": sol.report.objective, "verified_eps": worst,
                                   "read_seconds": t_e2e_read,
                                   "solve_lp_gpu_seconds": {k: sol.info[k] for k in
                                                            ("seconds_setup", "seconds_solve", "seconds_total")}}
        log(f"ingest: {size / 1e6:.1f} MB MPS read in {t_read:.3f} s ({t_one:.3f} s on one thread)"
            + (f" (reference reader {out['reference_read_seconds']:.2f} s)" if "reference_read_seconds" in out else "")
            + f"; file -> verified solution {e2e:.2f} s")
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def bench_prep():
    """Preprocessing in front of every solve: scale() (10 Ruiz passes +
    Pock-Chambolle, src/scaling.cpp:25-84) on the device (supr_scale_gpu:
    upload, stable CSR / CSC build, scaling, download) against the host
    version (supr_scale) and the reference's own scale() (oracle/_ref, one
    thread), on the 1M-variable transportation LP and a 7M-nonzero power-law
    LP; all three results compared bit for bit."""
    out = {}
    sys.path.insert(0, ROOT)
    try:
        from oracle import Checker, available
        ref = Checker("ref") if available("ref") else None
    except ImportError:   # pragma: no cover
        ref = None
    cases = (("transport_1000x1000", lambda: P.gen_transport(1000, 1000, seed=42)),
             ("power_law_100000x400000", lambda: P.gen_power_law_lp(100000, 400000, seed=42)))
    for name, make in cases:
        lp = make()
        P.scale_gpu(lp)   # warm-up
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            dev = P.scale_gpu(lp, timing=True)
            t = time.perf_counter() - t0
            if best is None or t < best[0]:
                best = (t, dev)
        t_dev, (ds, dDr, dDc, secs) = best
        t0 = time.perf_counter()
        hs, hDr, hDc = P.scale(lp)
        t_host = time.perf_counter() - t0
        mine = [ds.val, ds.c, ds.rL, ds.rU, ds.l, ds.u, dDr, dDc]
        host = [hs.val, hs.c, hs.rL, hs.rU, hs.l, hs.u, hDr, hDc]
        same = all(a.tobytes() == b.tobytes() for a, b in zip(mine, host))
        rec = {"m": lp.m, "n": lp.n, "nnz": lp.nnz, "device_seconds": t_dev,
               "device_layout_seconds": float(secs[0]), "device_scale_seconds": float(secs[1]),
               "device_download_seconds": float(secs[2]), "host_seconds": t_host,
               "device_equals_host": same}
        if ref is not None:
            t0 = time.perf_counter()
            rs = ref.scale(lp)
            rec["reference_seconds"] = time.perf_counter() - t0
            rec["device_equals_reference"] = all(a.tobytes() == np.asarray(b).tobytes()
                                                 for a, b in zip(mine, rs))
            rec["speedup_vs_reference"] = rec["reference_seconds"] / t_dev
        if not same or not rec.get("device_equals_reference", True):
            raise SystemExit(f"prep: device scale() differs on {name}")
        out[name] = rec
        log(f"prep {name}: device scale {t_dev:.3f} s (scaling kernels {secs[1] * 1e3:.1f} ms), host "
            f"{t_host:.2f} s" + (f", reference {rec['reference_seconds']:.2f} s" if ref else ""))
    return out


def large_lp():
    lp = P.gen_power_law_lp(500000, 2000000, seed=2026)
    s, Dr, Dc = P.scale(lp)
    return lp, s, Dr, Dc


def large_kernels(lp, s, Dr, Dc, reps=50):
    kb = P.bench_kernels(lp, s, Dr, Dc, reps=reps)
    out = {"lanes_A_fp64": kb["lanes_A"], "lanes_AT_fp64": kb["lanes_AT"], "kernels": {}}
    for name, key, sz in (("k_primal_f32", "primal_ms_fp32", 4), ("k_dual_f32", "dual_ms_fp32", 4),
                          ("k_primal_f64", "primal_ms_fp64", 8), ("k_dual_f64", "dual_ms_fp64", 8)):
        prim, dual = kernel_bytes(lp, sz, False, False, nonneg_bounds(s, sz))
        b = prim if "primal" in name else dual
        out["kernels"][name] = dict(ms=kb[key], bytes=b, GBps=b / (kb[key] * 1e-3) / 1e9,
                                    frac=b / (kb[key] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    return out


def bench_large_kernels(reps=50):
    """The power-law LP's iteration kernels alone (HIP events), with their
    algorithmic bytes -- the run to profile them."""
    lp, s, Dr, Dc = large_lp()
    out = large_kernels(lp, s, Dr, Dc, reps)
    out["workload"] = f"power-law LP m=500000 n=2000000 nnz={lp.nnz}"
    return out


def bench_large(mode=P.MODE_FAST_FP64):
    """The power-law LP (m = 500,000, n = 2,000,000, Zipf 1.5 column degrees
    on [2, 200]: 35.1M nonzeros, magnitudes log-uniform in [1e-3, 1e3]): its
    iteration kernels, and one solve to a verified 1e-6 in the FP64 mode
    (progress lines every 5 s on stderr)."""
    log("large: generating and scaling the power-law LP 500000 x 2000000 (host)")
    lp, s, Dr, Dc = large_lp()
    out = {"workload": f"power-law LP m=500000 n=2000000 nnz={lp.nnz} (Zipf 1.5 on [2,200] column degrees, "
                       "magnitudes log-uniform [1e-3,1e3], seed 2026), tol 1e-6 verified"}
    out.update(large_kernels(lp, s, Dr, Dc))
    name = "fast_fp64" if mode == P.MODE_FAST_FP64 else "mixed"
    log(f"large: {name} solve")
    with P.Engine(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=400000, verbose=2) as eng:
        t0 = time.perf_counter()
        r = eng.solve()
        dt = time.perf_counter() - t0
        setup = r.info["seconds_setup"]
    rep = P.verify(lp, r.x, r.y)
    worst = max(rep.eps_P, rep.eps_D, rep.eps_G)
    if not r.converged or worst > TOL:
        raise SystemExit(f"large: not verified (converged={r.converged}, eps {worst:.3e})")
    out["mode"] = name
    out.update(converged=r.converged, attempts=r.iterations, fp32_attempts=r.info["iterations_fp32"],
               seconds=dt, seconds_per_attempt=dt / max(1, r.iterations),
               verified_eps=[rep.eps_P, rep.eps_D, rep.eps_G], objective=rep.objective,
               engine_setup_seconds=setup)
    moved = moved_bytes_single(lp, s, r.iterations, r.info["iterations_fp32"])
    out["roofline"] = {"bound": "hbm", "achieved": moved / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": moved / dt / 1e9 / HBM_PEAK_GBS,
                       "note": "algorithmic bytes of the iteration kernels over the whole solve's wall time"}
    log(f"large: {name} {dt:.1f} s, {r.iterations} attempts, eps {worst:.2e}")
    return out


def moved_bytes_single(lp, s, attempts, attempts_fp32):
    """Algorithmic bytes the single-LP iteration kernels move over a solve."""
    p32, d32 = kernel_bytes(lp, 4, False, False, nonneg_bounds(s, 4))
    p64, d64 = kernel_bytes(lp, 8, False, False, nonneg_bounds(s, 8))
    return attempts_fp32 * (p32 + d32) + (attempts - attempts_fp32) * (p64 + d64)


def bench_one_shot(lp, scaled, Dr, Dc, reps=3):
    """The drop-in call as the reference's callers make it: supr_solve_pdhg_gpu
    from host arrays, upload, layout, solve, result copy and teardown
    included, per mode; the plain FP64 mode (the reference GPU design) the
    same way, once."""
    out = {}
    for name, mode in (("fast_fp64", P.MODE_FAST_FP64), ("mixed", P.MODE_MIXED)):
        P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=mode)   # warm (allocator, code objects)
        ts, r = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            r = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=mode)
            ts.append(time.perf_counter() - t0)
        rep = P.verify(lp, r.x, r.y)
        if not r.converged or max(rep.eps_P, rep.eps_D, rep.eps_G) > TOL:
            raise SystemExit(f"one-shot {name}: not verified")
        out[name] = dict(seconds=float(np.mean(ts)), min_seconds=float(np.min(ts)),
                         setup_seconds=r.info["seconds_setup"], solve_seconds=r.info["seconds_solve"],
                         attempts=r.iterations, objective=rep.objective)
    t0 = time.perf_counter()
    r = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=P.MODE_PLAIN_FP64)
    wall = time.perf_counter() - t0
    rep = P.verify(lp, r.x, r.y)
    out["plain_fp64"] = dict(seconds=wall, setup_seconds=r.info["seconds_setup"],
                             solve_seconds=r.info["seconds_solve"], attempts=r.iterations,
                             converged=r.converged, verified_eps=[rep.eps_P, rep.eps_D, rep.eps_G],
                             objective=rep.objective)
    for name in ("fast_fp64", "mixed"):
        out[name]["speedup_vs_plain_one_shot"] = wall / out[name]["seconds"]
    log(f"one-shot: fast64 {out['fast_fp64']['seconds'] * 1e3:.1f} ms, mixed {out['mixed']['seconds'] * 1e3:.1f} ms, "
        f"plain {wall * 1e3:.0f} ms -> {out['fast_fp64']['speedup_vs_plain_one_shot']:.1f}x / "
        f"{out['mixed']['speedup_vs_plain_one_shot']:.1f}x")
    return out


def rocprof_average_us(kernel):
    """Average duration (us) of `kernel` in the committed rocprofv3 kernel
    statistics of the 1M-variable LP's iteration kernels, or None."""
    path = os.path.join(ROOT, "profiles", "iteration_kernels_transport_1000x1000.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return float(json.loads(block)["kernels"][kernel]["avg_us"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def trace_summary(d, loop=False):
    """Summary of a rocprofv3 output directory (CSV): the kernel statistics
    and, with loop=True, the per-attempt timeline of a traced solve from the
    kernel trace (an attempt = a k_primal launch followed by a k_dual launch;
    its time = start of the next k_primal - start of this one)."""
    import csv
    import glob
    import statistics as st
    out = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
        rows = sorted(csv.DictReader(open(f)), key=lambda r: -float(r["TotalDurationNs"]))
        out["kernels"] = {r["Name"].split("(")[0].replace("void ", "").replace("supr::", ""):
                          {"calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
                           "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3,
                           "total_ms": float(r["TotalDurationNs"]) / 1e6, "pct": float(r["Percentage"])}
                          for r in rows[:16]}
    if loop:
        for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
            rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Start_Timestamp"]))
            nm = [r["Kernel_Name"] for r in rows]
            s0 = [int(r["Start_Timestamp"]) for r in rows]
            e0 = [int(r["End_Timestamp"]) for r in rows]
            att, dp, dd, checks = [], [], [], []
            for i in range(len(rows) - 2):
                if "k_primal" in nm[i] and "k_dual" in nm[i + 1]:
                    dp.append(e0[i] - s0[i])
                    dd.append(e0[i + 1] - s0[i + 1])
                    if "k_primal" in nm[i + 2]:
                        att.append(s0[i + 2] - s0[i])
                if "k_check_prep" in nm[i]:
                    j = i
                    while j < len(nm) and "k_restart_fin" not in nm[j]:
                        j += 1
                    if j < len(nm):
                        checks.append(e0[j] - s0[i])
            q = lambda v, p: float(np.percentile(v, p)) / 1e3 if v else None  # noqa: E731
            out["loop"] = {"dispatches": len(rows), "attempts_timed": len(att),
                           "attempt_us": {"median": q(att, 50), "p10": q(att, 10), "p90": q(att, 90),
                                          "mean": st.mean(att) / 1e3 if att else None},
                           "k_primal_us": {"median": q(dp, 50), "mean": st.mean(dp) / 1e3 if dp else None},
                           "k_dual_us": {"median": q(dd, 50), "mean": st.mean(dd) / 1e3 if dd else None},
                           "check_group_us": {"count": len(checks), "median": q(checks, 50),
                                              "mean": st.mean(checks) / 1e3 if checks else None},
                           "span_ms": (e0[-1] - s0[0]) / 1e6}
    return out


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, kernels):
    """Bytes per dispatch from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch), each scaled by what the same counter reads
    for the 1 GiB k_copy4 of --copy-calibration (median over dispatches)."""
    import csv
    import glob
    import statistics as st

    def med(d, sub):
        f = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)[0]
        v = [float(r["Counter_Value"]) for r in csv.DictReader(open(f)) if sub in r["Kernel_Name"]]
        return st.median(v), len(v)
    cf, _ = med(copy_fetch_dir, "k_copy4")
    cw, _ = med(copy_write_dir, "k_copy4")
    ff, wf = cf / (1 << 20), cw / (1 << 20)
    out = {"calibration": {"copy_bytes_each_way": 1 << 30, "fetch_kib": cf, "write_kib": cw,
                           "fetch_factor": ff, "write_factor": wf}, "kernels": {}}
    for key, sub in kernels.items():
        f, n = med(fetch_dir, sub)
        w, _ = med(write_dir, sub)
        out["kernels"][key] = {"kernel": sub, "dispatches": n, "fetch_kib": f, "write_kib": w,
                               "read_bytes": round(f * 1024 / ff), "write_bytes": round(w * 1024 / wf)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the power-law LP leg (about 3 minutes)")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-prep", action="store_true")
    ap.add_argument("--only", choices=["c0", "batched", "batched-solve", "serving", "serving-solve", "serving-full",
                                       "large", "large-mixed",
                                       "large-kernels", "ingest", "prep", "one-shot"],
                    help="run only this section and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a plain device copy of known size (reference for HBM counters)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the standalone iteration-kernel timing (for profiling it)")
    ap.add_argument("--lanes", default=None,
                    help="A,AT lanes per row for --kernels-only (default: the shape rule)")
    ap.add_argument("--solve-only", default=None, choices=["fast_fp64", "mixed"],
                    help="only resident-engine solves of the 1M-variable LP in this mode (for tracing the loop; "
                         "the attempts are launched on the stream directly, 16 per host poll)")
    ap.add_argument("--trace-summary", default=None, metavar="DIR",
                    help="summarise a rocprofv3 --kernel-trace --stats CSV directory (no GPU needed)")
    ap.add_argument("--loop", action="store_true", help="with --trace-summary: per-attempt timeline")
    ap.add_argument("--pmc-summary", nargs=4, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="HBM bytes per dispatch of the 1M-variable LP's iteration kernels from "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, copy-calibrated (no GPU needed)")
    args = ap.parse_args()
    if args.trace_summary:
        print(json.dumps(trace_summary(args.trace_summary, args.loop), indent=1))
        return
    if args.pmc_summary:
        lanes = [32, 1]
        # the 1M-variable LP's primal launches take k_primal_pairs (FP64) and
        # k_primal's grid-stride row-pair path (FP32), its dual launches the
        # flat one-pass form over A's 2,000 rows of 1,000
        df = dual_form(2000, 2000000, lanes[0], False)
        print(json.dumps(pmc_summary(*args.pmc_summary, {
            "k_primal_f32": primal_kernel(1000000, lanes[1], True, False),
            "k_dual_f32": f"k_dual<{lanes[0]}, float, {df}>",
            "k_primal_f64": primal_kernel(1000000, lanes[1], True, True),
            "k_dual_f64": f"k_dual<{lanes[0]}, double, {df}>"}), indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent instances would shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)
    if P.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.copy_calibration:
        nbytes = 1 << 30
        ms = P.bench_copy(nbytes, reps=10)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 11,
                          "ms_per_copy": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return

    serving_full = lambda: bench_serving(count=6, qps=0.5, nnz_hi=2e7, seed=11, span=True,  # noqa: E731
                                         saturated=False, plain_sample_attempts=1000)
    if args.only and args.only != "one-shot":
        fn = {"c0": bench_c0, "batched": bench_batched, "batched-solve": bench_batched_solve,
              "serving": bench_serving,
              # the server's runs alone, replays launched kernel by kernel (the
              # run to trace its kernels; the same bits as the graph replays)
              "serving-solve": lambda: bench_serving(baselines=False, server_options=dict(graph_attempts=-16)),
              "large": bench_large,
              "large-mixed": lambda: bench_large(P.MODE_MIXED), "ingest": bench_ingest,
              "prep": bench_prep, "large-kernels": bench_large_kernels,
              "serving-full": serving_full}[args.only]
        print(json.dumps({args.only: fn()}))
        return

    log("generating transport LP 1000x1000 and scaling (host, untimed)")
    lp = P.gen_transport(1000, 1000, seed=42)
    scaled, Dr, Dc = P.scale(lp)

    if args.kernels_only:
        pin = {}
        if args.lanes:
            la, lat = (int(v) for v in args.lanes.split(","))
            pin = dict(spmv_lanes_A=la, spmv_lanes_AT=lat)
        kb = P.bench_kernels(lp, scaled, Dr, Dc, reps=200, **pin)
        print(json.dumps({"kernels_ms": kb}))
        return

    if args.solve_only:
        mode = P.MODE_FAST_FP64 if args.solve_only == "fast_fp64" else P.MODE_MIXED
        # the default options: on this LP (2M nonzeros) the loop's kernels are
        # launched on the stream directly, 16 attempts and the check group per
        # host poll, which a kernel trace can follow
        with P.Engine(lp, scaled, Dr, Dc, tol=TOL, mode=mode) as eng:
            for k in range(args.warmup + args.steps):
                t0 = time.perf_counter()
                r = eng.solve()
                log(f"{args.solve_only} solve {k}: {(time.perf_counter() - t0) * 1e3:.2f} ms, "
                    f"{r.iterations} attempts")
        print(json.dumps({"solve_only": args.solve_only, "attempts": r.iterations}))
        return

    # the process's first device use (runtime start, code-object loading) is
    # paid here, so that setup_seconds below is the engine's own setup
    warm = P.gen_random_lp(50, 120, 4, seed=1)
    ws, wdr, wdc = P.scale(warm)
    P.solve_pdhg_gpu(warm, ws, wdr, wdc, tol=TOL, mode=P.MODE_MIXED)

    if args.only == "one-shot":
        print(json.dumps({"one_shot": bench_one_shot(lp, scaled, Dr, Dc)}))
        return

    results = {}
    for name, mode in (("fast_fp64", P.MODE_FAST_FP64), ("mixed", P.MODE_MIXED)):
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
        if len(set(iters)) != 1:
            raise SystemExit(f"{name}: attempt counts differ between runs {sorted(set(iters))}")
        results[name] = dict(seconds=float(np.mean(times)), min_seconds=float(np.min(times)),
                             iterations=int(np.median(iters)), objective=rep.objective,
                             eps=[rep.eps_P, rep.eps_D, rep.eps_G], setup_seconds=setup_s,
                             fp32_iterations=res.info["iterations_fp32"],
                             lanes=[res.info["lanes_A"], res.info["lanes_AT"]],
                             us_per_attempt=float(np.mean(times)) / res.iterations * 1e6)

    ref_obj = None
    ref_path = os.path.join(ROOT, "tests", "golden", "c1_reference.json")
    if os.path.exists(ref_path):
        with open(ref_path) as f:
            ref_obj = json.load(f)["objective"]
        for r in results.values():
            r["objective_rel_error_vs_cpu_reference"] = abs(r["objective"] - ref_obj) / (1 + abs(ref_obj))

    one_shot = bench_one_shot(lp, scaled, Dr, Dc)
    plain = one_shot["plain_fp64"]
    polish = bench_polish(lp, scaled, Dr, Dc)

    log("timing the iteration kernels (HIP events, 200 launches each)")
    kb = P.bench_kernels(lp, scaled, Dr, Dc, reps=200)
    prim32, dual32 = kernel_bytes(lp, 4, kb["ell_AT"] > 0, kb["ell_A"] > 0, nonneg_bounds(scaled, 4))
    prim64, dual64 = kernel_bytes(lp, 8, kb["ell_AT"] > 0, kb["ell_A"] > 0, nonneg_bounds(scaled, 8),
                                  rebuilt=kb.get("rebuilt_AT_fp64", 0) > 0)
    u32 = kernel_bytes(lp, 4, kb["ell_AT"] > 0, kb["ell_A"] > 0, nonneg_bounds(scaled, 4), compressed=False)
    u64 = kernel_bytes(lp, 8, kb["ell_AT"] > 0, kb["ell_A"] > 0, nonneg_bounds(scaled, 8), compressed=False)
    la, lat = int(kb["lanes_A"]), int(kb["lanes_AT"])
    # bytes: 4-byte indices and s-byte values per nonzero (the per-nonzero
    # term of the SpMV figure) plus the kernel's vectors once; format_bytes:
    # what this layout moves (16-bit indices, values formed from factors)
    kernels = {
        "k_primal_f32": dict(ms=kb["primal_ms_fp32"], bytes=u32[0], format_bytes=prim32),
        "k_dual_f32": dict(ms=kb["dual_ms_fp32"], bytes=u32[1], format_bytes=dual32),
        "k_primal_f64": dict(ms=kb["primal_ms_fp64"], bytes=u64[0], format_bytes=prim64,
                             rocprof=primal_kernel(lp.n, lat, kb["ell_AT"] > 0, True),
                             rebuilt_values=bool(kb.get("rebuilt_AT_fp64", 0) > 0)),
        "k_dual_f64": dict(ms=kb["dual_ms_fp64"], bytes=u64[1], format_bytes=dual64,
                           rocprof=f"k_dual<{la}, double, {dual_form(lp.m, lp.nnz, la, kb['ell_A'] > 0)}>"),
    }
    for k in kernels.values():
        k["GBps"] = k["bytes"] / (k["ms"] * 1e-3) / 1e9
        k["frac"] = k["GBps"] / HBM_PEAK_GBS
        k["format_frac"] = k["format_bytes"] / (k["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    dom_name = max(("k_primal_f64", "k_dual_f64"), key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    rp_us = rocprof_average_us(dom["rocprof"])
    box_copy = copy_rates()
    log("plain copy: " + ", ".join(f"{k} each way {v:.0f} GB/s" for k, v in box_copy.items()))

    cpu = None
    if not args.no_cpu:
        log("CPU baseline: the reference CPU PDLP on this LP, to completion (one thread)")
        cpu = cpu_baseline(lp, scaled, Dr, Dc)
        log(f"CPU baseline: {cpu['value']:.2f} s, {cpu['attempts']} attempts")

    c0 = bench_c0()
    batched = None if args.no_batched else bench_batched()
    serving = None if args.no_serving else serving_full()
    large = None if args.no_large else bench_large()
    ingest = None if args.no_ingest else bench_ingest()
    prep = None if args.no_prep else bench_prep()

    head = results["fast_fp64"]
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
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "transportation LP 1000x1000: m=2000, n=1000000, nnz=2000000, "
                        "tol 1e-6 relative KKT (verified), cold start, LP resident in HBM",
            "mode": "fast_fp64 (device-resident loop, FP64 iterates, SpMV and residuals)",
            "attempts": head["iterations"],
            "objective": head["objective"],
            "verified_eps": head["eps"],
            "engine_setup_seconds": head["setup_seconds"],
        },
        "modes": results,
        "fastest_mode": best_name,
        "one_shot": one_shot,
        "one_shot_seconds": one_shot["fast_fp64"]["seconds"],
        "one_shot_speedup_vs_plain": one_shot["fast_fp64"]["speedup_vs_plain_one_shot"],
        "plain_fp64": plain,
        "speedup_vs_plain": plain["solve_seconds"] / head["seconds"],
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": dom["GBps"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["frac"],
            "traffic": measured_traffic(dom_name),
            "traffic_source": "profiles/hbm_traffic_transport_1000x1000.md (rocprofv3 FETCH_SIZE + "
                              "WRITE_SIZE per dispatch, copy-calibrated)",
            "algorithmic_bytes_per_launch": dom["bytes"],
            "ms_per_launch": dom["ms"],
            "rocprof_avg_us": rp_us,
            "rocprof_frac": (dom["bytes"] / (rp_us * 1e-6) / 1e9 / HBM_PEAK_GBS) if rp_us else None,
            "rocprof_source": "profiles/iteration_kernels_transport_1000x1000.md",
            "copy_GBps": box_copy,
            "format_bytes_per_launch": dom["format_bytes"],
            "format_frac": dom["format_frac"],
            "note": "achieved: 4-byte indices and 8-byte values per nonzero plus the kernel's vectors "
                    "once; format_*: the bytes this layout moves (16-bit A^T indices, A^T's FP64 "
                    "values formed from Dr_i a and Dc_j when every unscaled entry shares one value); "
                    "traffic: measured HBM bytes; copy_GBps: the plain stream copy on this box at 1 GiB "
                    "each way (HBM) and at 64 / 16 MB each way (resident in the 256 MB last-level cache, "
                    "as the iteration kernels' working set is)",
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
        "speedup_vs_cpu_baseline": (cpu["value"] / head["seconds"]) if cpu else None,
        "c0": c0,
        "batched": batched,
        "serving": serving,
        "large": large,
        "ingest": ingest,
        "prep": prep,
        "polish_1e-8": polish,
        "cpu_reference_objective": ref_obj,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
