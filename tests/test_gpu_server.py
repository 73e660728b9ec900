"""GPU tests of the serving engine (supr_server_*): requests submitted while
others run are admitted when their ranges fit the device arenas; every
result is checked against the FP64 oracle, and is the same bits as the same
request served alone on a fresh server."""
import sys
import threading

import numpy as np
import pytest

import pdlp_amd as P
from conftest import golden_lp

pytestmark = pytest.mark.gpu

TOL = 1e-6
OBJ_RTOL = 1e-5   # two 1e-6-certified points differ by a few 1e-6 on the degenerate Netlib LPs
# room for about two of the requests below at once, so admission waits
CAPS = dict(max_cols=2600, max_rows=420, max_nnz=6000, max_instances=3, max_blocks=16)


def say(*a):
    print("[server]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def requests_(oracle, golden):
    out = []
    for name in ["tiny_range", "afiro", "sc50a", "transport_30x40_s7", "sc105", "blend"]:
        case = golden[name]
        form = golden_lp(case)
        s, Dr, Dc = P.scale(form)
        out.append((name, (form, s, Dr, Dc), case["ref"]["objective_form"]))
    lp = P.gen_random_lp(60, 150, 6, seed=11)
    s, Dr, Dc = P.scale(lp)
    o = oracle.scale(lp)
    ref = oracle.pdlp(lp, lp.with_values(*o[:6]), o[6], o[7], 200000, 64, TOL)
    out.append(("random_60x150", (lp, s, Dr, Dc), oracle.verify(lp, ref.x, ref.y)[3]))
    return out


def _serve(reqs, mode, stagger, threaded=None):
    """Submits request r at step stagger*r (request `threaded` from another
    thread) and steps until all are done; returns the results in order."""
    with P.Server(mode=mode, tol=TOL, max_iterations=100000, **CAPS) as srv:
        tickets, results, step = {}, {}, 0
        while len(results) < len(reqs):
            r = step // stagger if stagger else None
            if stagger and step % stagger == 0 and r < len(reqs):
                if r == threaded:
                    box = {}
                    th = threading.Thread(target=lambda: box.update(t=srv.submit(*reqs[r])))
                    th.start()
                    th.join()
                    tickets[r] = box["t"]
                else:
                    tickets[r] = srv.submit(*reqs[r])
            elif not stagger and step == 0:
                for i, q in enumerate(reqs):
                    tickets[i] = srv.submit(*q)
            srv.step()
            step += 1
            for i, t in tickets.items():
                if i not in results:
                    res = srv.result(t)
                    if res is not None:
                        results[i] = res
            assert step < 200000, "server made no progress"
        return [results[i] for i in range(len(reqs))], step


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_server_results_match_oracle_and_solo(oracle, requests_, mode):
    reqs = [q for _, q, _ in requests_]
    res, steps = _serve(reqs, mode, stagger=3, threaded=2)
    say(f"mode {mode}: {len(res)} requests in {steps} steps; attempts", [r.iterations for r in res])
    for (name, (form, _, _, _), ref_obj), r in zip(requests_, res):
        assert r.converged, (name, r.iterations)
        rep = oracle.verify(form, r.x, r.y)
        assert max(rep[:3]) <= TOL, (name, rep[:3])
        np.testing.assert_allclose([r.eps_P, r.eps_D, r.eps_G], rep[:3], rtol=1e-6, atol=1e-12)
        assert abs(rep[3] - ref_obj) / (1 + abs(ref_obj)) <= OBJ_RTOL, (name, rep[3], ref_obj)
        assert r.info["seconds_total"] >= r.info["seconds_solve"] > 0
    # each request alone on a fresh server: the same bits
    for i, q in enumerate(reqs):
        alone, _ = _serve([q], mode, stagger=1)
        assert alone[0].iterations == res[i].iterations, requests_[i][0]
        assert np.array_equal(alone[0].x, res[i].x) and np.array_equal(alone[0].y, res[i].y)
    say(f"mode {mode}: every request identical alone and in the mix")


def test_server_all_at_once_and_refusals(requests_):
    reqs = [q for _, q, _ in requests_]
    res, _ = _serve(reqs, P.MODE_MIXED, stagger=0)
    assert all(r.converged for r in res)
    big = P.gen_random_lp(500, 3000, 6, seed=2)
    s, Dr, Dc = P.scale(big)
    with P.Server(tol=TOL, **CAPS) as srv:
        with pytest.raises(P.SupremaError) as e:
            srv.submit(big, s, Dr, Dc)
        assert "capacity" in str(e.value)
        assert srv.result(12345) is None
        assert srv.step() == (0, 0, 0)
        assert srv.device_bytes > 0
    assert P.Server.request_bytes(27, 32, 83, P.MODE_MIXED) > P.Server.request_bytes(27, 32, 83, P.MODE_FAST_FP64)


# ------------------------------------------------------------- serving mix --
# Requests from the serving workload's own family: the random equality LP
# generator (6 nonzeros per column, feasible and bounded by construction) with
# nnz log-uniform and m / n uniform in [0.2, 0.8], submitted from another
# thread at Poisson arrival times while the server steps.
C4_CAPS = dict(max_cols=1 << 22, max_rows=1 << 21, max_nnz=1 << 24, max_instances=16,
               max_blocks=1024)
C4_OBJ_RTOL = 1e-6


def c4_request(nnz, ratio, seed):
    n = max(20, int(nnz / 6))
    m = max(6, int(ratio * n))
    lp = P.gen_random_lp(m, n, 6, seed=seed)
    s, Dr, Dc = P.scale(lp)
    return lp, s, Dr, Dc


def serve_poisson(reqs, mode, qps, seed):
    """Submits reqs[i] at Poisson arrival times (another thread) while this
    thread steps the server; returns the results in order and the largest
    number of requests seen running at once."""
    import time
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / qps, len(reqs)))
    with P.Server(mode=mode, tol=TOL, max_iterations=200000, **C4_CAPS) as srv:
        tickets = [None] * len(reqs)
        lock = threading.Lock()
        t0 = time.perf_counter()

        def submitter():
            for i, q in enumerate(reqs):
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                t = srv.submit(*q)
                with lock:
                    tickets[i] = t

        th = threading.Thread(target=submitter)
        th.start()
        results, most, steps = {}, 0, 0
        try:
            while len(results) < len(reqs):
                running, _, _ = srv.step()
                most = max(most, running)
                steps += 1
                with lock:
                    tk = list(tickets)
                for i, t in enumerate(tk):
                    if t is not None and i not in results:
                        r = srv.result(t)
                        if r is not None:
                            results[i] = r
                assert steps < 2000000, "server made no progress"
        finally:
            th.join()
        return [results[i] for i in range(len(reqs))], most


@pytest.fixture(scope="module")
def c4_mix(oracle):
    rng = np.random.default_rng(2024)
    nnzs = np.exp(rng.uniform(np.log(1e4), np.log(1e5), 6))
    ratios = rng.uniform(0.2, 0.8, 6)
    reqs = [c4_request(z, r, 3000 + k) for k, (z, r) in enumerate(zip(nnzs, ratios))]
    refs = []
    for lp, _, _, _ in reqs:
        o = oracle.scale(lp)
        ref = oracle.pdlp(lp, lp.with_values(*o[:6]), o[6], o[7], 200000, 64, TOL)
        assert ref.converged
        refs.append(oracle.verify(lp, ref.x, ref.y)[3])
    say("serving mix: m x n (nnz) =", [f"{q[0].m}x{q[0].n} ({q[0].nnz})" for q in reqs])
    return reqs, refs


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_serving_mix_matches_oracle_and_alone(oracle, c4_mix, mode):
    """Six requests of nnz 1e4-1e5 arriving at 40 requests/s: each result is
    verified by the oracle (max eps <= 1e-6), its objective agrees with the
    oracle's own solve to 1e-6 relative, and it is the same bits as the same
    request served alone on a fresh server."""
    reqs, refs = c4_mix
    res, most = serve_poisson(reqs, mode, qps=40.0, seed=1)
    assert most >= 2, "the arrivals never overlapped on the device"
    errs = []
    for (lp, _, _, _), r, ref_obj in zip(reqs, res, refs):
        assert r.converged, (lp.m, lp.n, r.iterations)
        rep = oracle.verify(lp, r.x, r.y)
        assert max(rep[:3]) <= TOL, rep[:3]
        errs.append(abs(rep[3] - ref_obj) / (1 + abs(ref_obj)))
        assert errs[-1] <= C4_OBJ_RTOL, (lp.m, lp.n, rep[3], ref_obj)
    for q, r in zip(reqs, res):
        alone, _ = serve_poisson([q], mode, qps=1e9, seed=2)
        assert alone[0].iterations == r.iterations
        assert np.array_equal(alone[0].x, r.x) and np.array_equal(alone[0].y, r.y)
    say(f"mode {mode}: {len(res)} requests, up to {most} running at once, attempts "
        f"{[r.iterations for r in res]}, objective errors vs the oracle {['%.1e' % e for e in errs]}; "
        "each identical alone")


def test_serving_large_request_properties(oracle, c4_mix):
    """A request of about 1.2M nonzeros (too large for an oracle solve here)
    in the mix with two small ones: its point is verified by the oracle's
    verifier, the mixed and fast-FP64 servers agree on the objective to
    1e-6 relative, and it is the same bits alone as in the mix."""
    big = c4_request(1.2e6, 0.5, 77)
    lp = big[0]
    small = c4_mix[0][:2]
    objs = {}
    for mode in (P.MODE_MIXED, P.MODE_FAST_FP64):
        res, most = serve_poisson([small[0], big, small[1]], mode, qps=40.0, seed=3)
        r = res[1]
        assert r.converged and most >= 2
        rep = oracle.verify(lp, r.x, r.y)
        assert max(rep[:3]) <= TOL, rep[:3]
        objs[mode] = rep[3]
        alone, _ = serve_poisson([big], mode, qps=1e9, seed=4)
        assert alone[0].iterations == r.iterations
        assert np.array_equal(alone[0].x, r.x) and np.array_equal(alone[0].y, r.y)
        say(f"mode {mode}: {lp.m}x{lp.n} ({lp.nnz} nnz) in {r.iterations} attempts, eps {max(rep[:3]):.2e}")
    a, b = objs[P.MODE_MIXED], objs[P.MODE_FAST_FP64]
    assert abs(a - b) / (1 + abs(b)) <= C4_OBJ_RTOL, (a, b)

