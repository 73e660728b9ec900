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
OBJ_RTOL = 1e-5
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
    assert P.Server.request_bytes(27, 32, 83) > P.Server.request_bytes(27, 32, 83, P.MODE_FAST_FP64)
