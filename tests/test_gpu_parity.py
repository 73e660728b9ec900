"""GPU parity: the HIP engine through the C ABI against the FP64 oracle.

Tolerances (floating point; the reference path is FP64 throughout):
  * SpMV, FP64: per row |gpu - oracle| <= (nnz_row + 2) * 2^-52 * (|A||x|)_row;
    FP32 (values, x and accumulation in FP32): (nnz_row + 2) * 2^-23 * (|A||x|)_row.
  * Solves: converged, the ORACLE's verifier on the returned point gives
    max(eps_P, eps_D, eps_G) <= tol (1e-6), and the objective agrees with the
    reference engine's own 1e-6 solution to |d| / (1 + |obj|) <= 1e-5: two
    points that both meet a 1e-6 KKT tolerance can differ by a few 1e-6
    (measured: up to 2.3e-6 on blend, 1e-9 to 1e-11 on the random and
    transportation LPs).
  * Solves to 1e-9: the objective agrees with the reference engine's own 1e-9
    solution (tests/golden/tight.json) to 1e-7 relative, ten times audit_lp's
    WARN level (tools/audit_lp.cpp:261-268).
"""
import sys

import numpy as np
import pytest

import pdlp_amd as P
from conftest import golden_lp, load_fixture

pytestmark = pytest.mark.gpu

TOL = 1e-6
OBJ_RTOL = 1e-5
TIGHT_TOL = 1e-9
TIGHT_OBJ_RTOL = 1e-7


def _spmv_cases():
    rng = np.random.default_rng(5)
    cases = {
        # A^T is ELL (6 per column) with 16-bit indices (m <= 65,536) ...
        "random_c0_small": P.gen_random_lp(200, 500, 6, seed=1),
        # ... and with 32-bit ones (m > 65,536)
        "random_ell_wide": P.gen_random_lp(70000, 30000, 6, seed=3),
        "transport_30x40": P.gen_transport(30, 40, seed=7),
        "powerlaw_mag3": P.gen_random_lp(300, 700, 0, seed=2, mag_decades=3,
                                         col_deg=np.minimum(2 + rng.zipf(1.5, 700), 200).clip(2, 250)),
    }
    # rows of very different lengths plus empty rows and columns
    m, n = 90, 1500
    rows, cols = [], []
    for i in range(0, m, 3):            # every third row empty
        k = (i * 37) % 700 + 1
        cs = rng.choice(n - 10, size=k, replace=False)
        rows += [i] * k
        cols += cs.tolist()
    rows, cols = np.array(rows), np.array(cols)
    cases["ragged_empty"] = P.LP(m, n, rows, cols, rng.standard_normal(len(rows)), np.zeros(n),
                                 np.zeros(m), np.zeros(m), np.zeros(n), np.zeros(n))
    # tall enough (>= 131072 rows both ways) for the LDS-staged schedules
    mt = nt = 140000
    lens = rng.integers(0, 40, mt)
    lens[::7] = 0
    trow = np.repeat(np.arange(mt), lens)
    tcol = rng.integers(0, nt, trow.size)
    cases["tall_ragged_staged"] = P.LP(mt, nt, trow, tcol, rng.standard_normal(trow.size),
                                       np.zeros(nt), np.zeros(mt), np.zeros(mt), np.zeros(nt),
                                       np.zeros(nt))
    return cases


SPMV = _spmv_cases()


@pytest.mark.parametrize("name", sorted(SPMV))
@pytest.mark.parametrize("transposed", [False, True])
@pytest.mark.parametrize("fp32", [False, True])
def test_spmv_schedules_match_oracle(oracle, name, transposed, fp32):
    lp = SPMV[name]
    rng = np.random.default_rng(11)
    x = rng.standard_normal(lp.m if transposed else lp.n)
    absA = lp.with_values(np.abs(lp.val), lp.c, lp.rL, lp.rU, lp.l, lp.u)
    xe = x.astype(np.float32).astype(np.float64) if fp32 else x
    want = oracle.spmv(lp, xe, transposed)
    mag = oracle.spmv(absA, np.abs(xe), transposed)
    rows = lp.col if transposed else lp.row
    cnt = np.bincount(rows, minlength=lp.n if transposed else lp.m)
    unit = 2.0 ** -23 if fp32 else 2.0 ** -52
    bound = (cnt + 2) * unit * mag + 1e-300
    for lanes in (1, 2, 4, 8, 16, 32, 64, 256, 0, -1):
        got = P.spmv(lp, x, transposed=transposed, lanes=lanes, fp32=fp32)
        err = np.abs(got - want)
        assert np.all(err <= bound), f"lanes={lanes}: worst row {np.argmax(err / bound)}"
        assert np.all(got[cnt == 0] == 0)


def _solve(case, mode, **kw):
    form = golden_lp(case)
    scaled, Dr, Dc = P.scale(form)
    res = P.solve_pdhg_gpu(form, scaled, Dr, Dc, tol=TOL, max_iterations=500000, mode=mode, **kw)
    return form, res


def _check_against_reference(oracle, case, form, res):
    assert res.converged, f"{case['name']}: not converged after {res.iterations}"
    rep = oracle.verify(form, res.x, res.y)
    assert max(rep[0], rep[1], rep[2]) <= TOL, f"{case['name']}: oracle verify {rep[:3]}"
    # the engine's own eps are the oracle verifier's within FP64 reduction order
    np.testing.assert_allclose([res.eps_P, res.eps_D, res.eps_G], rep[:3], rtol=1e-6, atol=1e-12)
    # the engine's objective (summed by the device check) is the oracle's c'x
    # + offset of the returned point to FP64 summation-order rounding
    scale = 1.0 + float(np.dot(np.abs(form.c), np.abs(res.x))) + abs(form.obj_offset)
    assert abs(res.info["objective"] - rep[3]) <= 1e-13 * scale, (res.info["objective"], rep[3])
    ref_obj = case["ref"]["objective_form"]
    err = abs(rep[3] - ref_obj) / (1 + abs(ref_obj))
    print(f"[parity] {case['name']}: {res.iterations} attempts (reference {case['ref']['iterations']}), "
          f"objective error {err:.1e}", file=sys.stderr, flush=True)
    assert err <= OBJ_RTOL, (rep[3], ref_obj)
    return rep


FAST_CASES = ["afiro", "adlittle", "blend", "sc50a", "sc50b", "sc105", "kb2", "scagr7", "israel",
              "share1b", "beaconfd", "tiny_range", "med", "illcond", "random_200x500_s1",
              "random_2000x5000_s42", "transport_60x60_s42", "transport_30x40_s7",
              "power_law_1000x4000_s7", "blend_25x40_s103"]
PLAIN_CASES = ["afiro", "sc50a", "blend", "tiny_range", "transport_30x40_s7", "random_200x500_s1",
               "med", "power_law_1000x4000_s7"]


@pytest.mark.parametrize("name", FAST_CASES)
@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_device_resident_solve_matches_reference(oracle, golden, name, mode):
    case = golden[name]
    form, res = _solve(case, mode)
    _check_against_reference(oracle, case, form, res)


@pytest.fixture(scope="module")
def tight():
    import json
    import os
    from conftest import GOLDEN_DIR
    with open(os.path.join(GOLDEN_DIR, "tight.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


@pytest.mark.parametrize("name", FAST_CASES)
@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_tight_solve_objective_matches_reference(oracle, golden, tight, name, mode):
    """To 1e-9 relative KKT (the mixed mode through its FP64 polish): the
    oracle's verifier certifies the point and the objective is the
    reference engine's 1e-9 objective within 1e-7."""
    case, want = golden[name], tight[name]
    form = golden_lp(case)
    scaled, Dr, Dc = P.scale(form)
    res = P.solve_pdhg_gpu(form, scaled, Dr, Dc, tol=TIGHT_TOL, max_iterations=500000, mode=mode)
    assert res.converged, (name, res.iterations)
    rep = oracle.verify(form, res.x, res.y)
    assert max(rep[:3]) <= TIGHT_TOL, rep[:3]
    ref_obj = want["objective_form"]
    err = abs(rep[3] - ref_obj) / (1 + abs(ref_obj))
    print(f"[parity] {name} to 1e-9: {res.iterations} attempts (reference {want['iterations']}), "
          f"objective error {err:.1e}", file=sys.stderr, flush=True)
    assert err <= TIGHT_OBJ_RTOL, (rep[3], ref_obj)


@pytest.mark.parametrize("name", PLAIN_CASES)
def test_plain_fp64_solve_matches_reference(oracle, golden, name):
    case = golden[name]
    form, res = _solve(case, P.MODE_PLAIN_FP64)
    _check_against_reference(oracle, case, form, res)
    # same algorithm, same arithmetic width: the trajectory length stays close
    it_ref = case["ref"]["iterations"]
    assert 0.5 * it_ref <= res.iterations <= 1.5 * it_ref


@pytest.mark.parametrize("name", ["afiro", "blend", "israel", "share1b", "beaconfd"])
def test_netlib_original_objective_matches_published(oracle, golden, name):
    """The audit path: solve on the GPU, extract_solution, verify on the
    ORIGINAL LP; published optimum within 1e-4 (tools/audit_lp.cpp:270-284)."""
    lp, maximize = load_fixture(name)
    form, meta = P.build_pdlp_form(lp)
    scaled, Dr, Dc = P.scale(form)
    res = P.solve_pdhg_gpu(form, scaled, Dr, Dc, tol=TOL, max_iterations=500000)
    assert res.converged
    x, y = P.extract_solution(meta, res.x, res.y)
    rep = oracle.verify(lp, x, y)
    obj = -rep[3] if maximize else rep[3]
    pub = golden[name]["published_optimum"]
    assert abs(obj - pub) / (1 + abs(pub)) <= 1e-4
    assert max(rep[:3]) <= 1e-5


@pytest.mark.parametrize("name", ["afiro", "blend", "israel", "share1b", "beaconfd", "med"])
def test_mps_file_to_verified_solution(oracle, golden, tmp_path, name):
    """The solve_mps path without presolve, from the MPS file itself: this
    library's reader (tests/golden/mps/<name>.mps.gz) -> solve_lp_gpu (form,
    scale, solve, extract, verify on the LP) -> the published optimum within
    1e-4 where one exists (tools/audit_lp.cpp:270-284), the oracle's verifier
    agreeing with the device's report, and the same LP as the fixture the
    reference's reader made."""
    import gzip
    import os
    from conftest import GOLDEN_DIR
    with gzip.open(os.path.join(GOLDEN_DIR, "mps", name + ".mps.gz"), "rb") as f:
        path = tmp_path / (name + ".mps")
        path.write_bytes(f.read())
    mp = P.read_mps(str(path))
    ref_lp, maximize = load_fixture(name)
    assert mp.maximize == maximize and np.array_equal(mp.lp.val, ref_lp.val)
    sol = P.solve_lp_gpu(mp.lp, tol=TOL, max_iterations=500000)
    assert sol.status in ("optimal", "near_optimal"), sol.status
    rep = oracle.verify(mp.lp, sol.x, sol.y)
    np.testing.assert_allclose([sol.report.eps_P, sol.report.eps_D, sol.report.eps_G], rep[:3],
                               rtol=1e-9, atol=1e-13)
    assert max(rep[:3]) <= 10 * TOL
    pub = golden[name].get("published_optimum")
    if pub is not None:
        obj = -rep[3] if maximize else rep[3]
        assert abs(obj - pub) / (1 + abs(pub)) <= 1e-4


@pytest.mark.parametrize("name", ["tiny_infeasible", "tiny_unbounded"])
@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64, P.MODE_PLAIN_FP64])
def test_infeasible_and_unbounded_samples_run_to_the_limit(oracle, tmp_path, name, mode):
    """The reference's own infeasible / unbounded samples (sample_problems/):
    PDLP has no infeasibility detection, so the reference engine (and the
    oracle, checked here) runs to max_iterations and returns its best point,
    not converged.  The GPU engine must do the same in every mode -- no fault,
    no early convergence claim, a finite point -- and solve_lp_gpu must report
    iteration_limit."""
    import gzip
    import os
    from conftest import GOLDEN_DIR
    with gzip.open(os.path.join(GOLDEN_DIR, "mps", name + ".mps.gz"), "rb") as f:
        path = tmp_path / (name + ".mps")
        path.write_bytes(f.read())
    lp = P.read_mps(str(path)).lp
    form, meta = P.build_pdlp_form(lp)
    s, Dr, Dc = P.scale(form)
    ref = oracle.pdlp(form, s, Dr, Dc, 5000, 64, TOL)
    assert not ref.converged and ref.iterations == 5000
    res = P.solve_pdhg_gpu(form, s, Dr, Dc, tol=TOL, max_iterations=5000, mode=mode)
    assert not res.converged and res.iterations == 5000
    assert res.x.shape == (form.n,) and np.isfinite(res.x).all() and np.isfinite(res.y).all()
    if mode != P.MODE_PLAIN_FP64:
        sol = P.solve_lp_gpu(lp, tol=TOL, max_iterations=5000, mode=mode)
        assert sol.status == "iteration_limit" and not sol.converged


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_polish_restores_1e8(oracle, mode):
    """The FP32 loop cannot reach 1e-8; in the mixed mode the FP64 polish
    takes over and restores it.  The random equality LP m = 2,000, n = 5,000
    to 1e-8, certified by the oracle's verifier, with the objective the FP64
    oracle reaches at 1e-8."""
    lp = P.gen_random_lp(2000, 5000, 6, seed=42)
    s, Dr, Dc = P.scale(lp)
    res = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=1e-8, mode=mode, max_iterations=500000)
    assert res.converged
    rep = oracle.verify(lp, res.x, res.y)
    assert max(rep[:3]) <= 1e-8, rep[:3]
    if mode == P.MODE_MIXED:
        assert 0 < res.info["iterations_fp32"] < res.iterations   # the polish ran
    ref = oracle.pdlp(lp, s, Dr, Dc, 500000, 64, 1e-8)
    ref_obj = oracle.verify(lp, ref.x, ref.y)[3]
    assert abs(rep[3] - ref_obj) / (1 + abs(ref_obj)) <= 1e-7


@pytest.fixture(scope="module")
def transport_1m():
    lp = P.gen_transport(1000, 1000, seed=42)
    scaled, Dr, Dc = P.scale(lp)
    return lp, scaled, Dr, Dc


def test_transport_1m_verified_and_modes_agree(oracle, transport_1m):
    """1,000,000-variable transportation LP: every device-resident mode reaches
    1e-6 as measured by the oracle's verifier; the objectives agree; a second
    run with the same SpMV schedule is bit-identical (deterministic
    reductions)."""
    lp, scaled, Dr, Dc = transport_1m
    objs = []
    sched = dict(spmv_lanes_A=32, spmv_lanes_AT=1)
    for mode in (P.MODE_MIXED, P.MODE_FAST_FP64):
        res = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=mode, **sched)
        assert res.converged
        rep = oracle.verify(lp, res.x, res.y)
        assert max(rep[:3]) <= TOL
        objs.append(rep[3])
        if mode == P.MODE_MIXED:
            again = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=mode, **sched)
            assert again.iterations == res.iterations
            assert np.array_equal(again.x, res.x) and np.array_equal(again.y, res.y)
    assert abs(objs[0] - objs[1]) / (1 + abs(objs[1])) <= OBJ_RTOL
    # supplies and demands are integers: the optimal cost is an integer too
    assert abs(objs[1] - round(objs[1])) / (1 + abs(objs[1])) <= OBJ_RTOL


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64, P.MODE_PLAIN_FP64])
def test_iteration_limit_returns_best_point(oracle, mode):
    lp = P.gen_random_lp(200, 500, 6, seed=1)
    scaled, Dr, Dc = P.scale(lp)
    res = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, max_iterations=300, mode=mode)
    assert not res.converged
    assert res.iterations == 300
    assert res.info["has_point"] == 1
    rep = oracle.verify(lp, res.x, res.y)
    np.testing.assert_allclose([res.eps_P, res.eps_D, res.eps_G], rep[:3], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_PLAIN_FP64])
def test_stop_flag_cancels(mode):
    lp = P.gen_random_lp(2000, 5000, 6, seed=42)
    scaled, Dr, Dc = P.scale(lp)
    polls = {"n": 0}

    def stop():
        polls["n"] += 1
        return polls["n"] > 20

    res = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=1e-9, stop_flag=stop, mode=mode)
    assert not res.converged
    assert res.iterations < 200000
    assert polls["n"] > 20


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_stop_flag_cancels_planned_launches(oracle, mode):
    """The stop callback on an LP of 2^20+ nonzeros (direct launches with
    planned check groups): polled once per launch unit, it ends the solve
    early with the best point a check kept, whose eps the oracle's verifier
    reproduces."""
    lp = P.gen_transport(1000, 600, seed=12)
    scaled, Dr, Dc = P.scale(lp)
    polls = {"n": 0}

    def stop():
        polls["n"] += 1
        return polls["n"] > 40

    res = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=1e-12, stop_flag=stop, mode=mode)
    assert not res.converged and res.iterations < 200000
    assert polls["n"] == 41
    assert res.info["has_point"] == 1 and res.info["checks"] >= 1
    rep = oracle.verify(lp, res.x, res.y)
    np.testing.assert_allclose([res.eps_P, res.eps_D, res.eps_G], rep[:3], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_default_runs_are_bit_reproducible(oracle, mode):
    """With default options (lanes per row from the matrix's shape, no
    timing) every setup of an LP sums in the same order: the one-shot entry
    point, two solves on one resident engine and a second engine give the
    same attempts and the same bits."""
    lp = P.gen_transport(60, 60, seed=42)
    scaled, Dr, Dc = P.scale(lp)
    one = P.solve_pdhg_gpu(lp, scaled, Dr, Dc, tol=TOL, mode=mode)
    runs = []
    for _ in range(2):
        with P.Engine(lp, scaled, Dr, Dc, tol=TOL, mode=mode) as eng:
            runs += [eng.solve(), eng.solve()]
    for r in runs:
        assert r.converged and r.iterations == one.iterations and r.restarts == one.restarts
        assert np.array_equal(r.x, one.x) and np.array_equal(r.y, one.y)
    rep = oracle.verify(lp, one.x, one.y)
    assert max(rep[:3]) <= TOL


def test_cpp_drop_in_api():
    """The C++ drop-in (include/pdlp_gpu.hpp) used as the reference's callers
    use it: form, scale, solve, extract, verify; cancellation from another
    thread; std::runtime_error on bad input."""
    import json
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "team-suprema_amd", "_build", "drop_in_solve")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and out["failures"] == 0, (r.stdout, r.stderr)
    assert out["converged"] == 1 and max(out["eps"]) <= 1e-6
    assert out["cancelled_converged"] == 0 and out["threw"] == 1


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_direct_launches_match_graph_replay(mode):
    """graph_attempts < 0 launches the replay's kernel sequence on the stream
    directly, and 0 (auto) replays 8 units of 16 attempts and the check group
    per graph on an LP this small: the same kernels in the same order, so the
    same bits (x, y and the objective)."""
    lp = P.gen_random_lp(300, 800, 6, seed=21)
    s, Dr, Dc = P.scale(lp)
    runs = [P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, graph_attempts=g,
                             spmv_lanes_A=4, spmv_lanes_AT=1) for g in (16, 0, -16, -6)]
    assert runs[0].converged
    for r in runs[1:]:
        assert r.iterations == runs[0].iterations and r.restarts == runs[0].restarts
        assert np.array_equal(r.x, runs[0].x) and np.array_equal(r.y, runs[0].y)
        assert r.info["objective"] == runs[0].info["objective"]


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_planned_direct_launches_match_graph_replay(mode):
    """LPs of 2^20+ nonzeros launch the loop directly with check groups placed
    where checks are expected (the default) instead of after every 16
    attempts: the device runs the same state machine, so the solve is the
    graph replay's bit for bit -- to convergence (in the mixed mode across the
    FP32 -> FP64 hand-over), stopped by the iteration limit (a check group
    placed at the limit), and on a resident engine's repeated solves."""
    cases = [(P.gen_transport(1000, 600, seed=12), {}),
             (P.gen_power_law_lp(30000, 120000, seed=4), dict(max_iterations=6000)),
             (P.gen_transport(1000, 600, seed=12), dict(max_iterations=777))]
    for lp, kw in cases:
        assert lp.nnz >= 1 << 20
        s, Dr, Dc = P.scale(lp)
        kw = dict(dict(tol=TOL, mode=mode, max_iterations=60000), **kw)
        planned = P.solve_pdhg_gpu(lp, s, Dr, Dc, **kw)
        replayed = P.solve_pdhg_gpu(lp, s, Dr, Dc, graph_attempts=16, **kw)
        assert planned.iterations == replayed.iterations and planned.restarts == replayed.restarts
        assert planned.converged == replayed.converged
        assert np.array_equal(planned.x, replayed.x) and np.array_equal(planned.y, replayed.y)
        assert planned.info["objective"] == replayed.info["objective"]
        assert planned.info["checks"] == replayed.info["checks"]
        # a resident engine's later solves start from a state copy that the
        # previous solve left finished: they must run whole solves all the same
        with P.Engine(lp, s, Dr, Dc, **kw) as eng:
            for _ in range(2):
                again = eng.solve()
                assert again.iterations == planned.iterations
                assert np.array_equal(again.x, planned.x) and np.array_equal(again.y, planned.y)
        print(f"[parity] planned launches m={lp.m} n={lp.n} mode {mode}: {planned.iterations} attempts, "
              f"converged {planned.converged}, same bits as graph replays", file=sys.stderr, flush=True)


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_device_layout_matches_host_layout(monkeypatch, mode):
    """Caller-scaled forms of 2^16 or more nonzeros are laid out on the device
    (a stable counting sort, as the host's build_csr): with the lanes pinned,
    the solve gives the same bits as with the host layout (SUPR_HOST_LAYOUT=1),
    for CSR rows (power-law LP) and ELL rows (transportation LP's A^T)."""
    cases = [(P.gen_power_law_lp(3000, 12000, seed=8), 16, 4), (P.gen_transport(200, 400, seed=3), 32, 1)]
    for lp, la, lat in cases:
        assert lp.nnz >= 1 << 16
        s, Dr, Dc = P.scale(lp)
        def run():
            return P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=4000,
                                    spmv_lanes_A=la, spmv_lanes_AT=lat)
        dev = run()
        monkeypatch.setenv("SUPR_HOST_LAYOUT", "1")
        try:
            host = run()
        finally:
            monkeypatch.delenv("SUPR_HOST_LAYOUT")
        assert dev.iterations == host.iterations and dev.restarts == host.restarts
        assert np.array_equal(dev.x, host.x) and np.array_equal(dev.y, host.y)
        assert (dev.info["lanes_A"], dev.info["lanes_AT"]) == (host.info["lanes_A"], host.info["lanes_AT"])


def _relabel_order(deg, group=16, tile=256):
    """The column relabelling rule for skewed degrees: columns sorted by degree
    (descending, stable), groups of 16 dealt round-robin over 256-column
    tiles, the remainder at the end in sorted order.  Returns newpos."""
    n = len(deg)
    srt = np.argsort(-deg, kind="stable")
    T = n // tile
    ng = T * (tile // group)
    g = np.arange(ng)
    base = (g % T) * tile + (g // T) * group
    newpos = np.empty(n, dtype=np.int64)
    newpos[srt[:ng * group]] = (base[:, None] + np.arange(group)[None, :]).ravel()
    newpos[srt[ng * group:]] = np.arange(ng * group, n)
    return newpos


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_skewed_columns_are_relabelled_and_mapped_back(monkeypatch, mode):
    """A power-law LP (2^16+ columns, largest degree >= 4x the mean) is solved
    with its columns relabelled by degree: the one-shot call and the engine
    give exactly the solve of the LP relabelled here by the same rule (with the
    library's own relabelling off, SUPR_COLUMN_ORDER=given), x mapped back to
    the caller's order -- same bits, same attempts, the same objective (the
    device check's sum of the relabelled terms), which agrees with c'x summed
    in the caller's order to rounding."""
    lp = P.gen_power_law_lp(20000, 80000, seed=5)
    deg = np.bincount(lp.col, minlength=lp.n)
    assert lp.n >= 1 << 16 and deg.max() >= 4 * lp.nnz / lp.n
    s, Dr, Dc = P.scale(lp)
    pin = dict(tol=TOL, mode=mode, max_iterations=3000, spmv_lanes_A=16, spmv_lanes_AT=16)
    auto = P.solve_pdhg_gpu(lp, s, Dr, Dc, **pin)
    with P.Engine(lp, s, Dr, Dc, **pin) as eng:
        via_engine = eng.solve()
    newpos = _relabel_order(deg)
    order = np.empty_like(newpos)
    order[newpos] = np.arange(lp.n)

    def relabelled(q):
        return P.LP(q.m, q.n, q.row, newpos[q.col].astype(np.int32), q.val, q.c[order], q.rL, q.rU,
                    q.l[order], q.u[order], q.obj_offset, q.name)
    monkeypatch.setenv("SUPR_COLUMN_ORDER", "given")
    try:
        ref = P.solve_pdhg_gpu(relabelled(lp), relabelled(s), Dr, Dc[order], **pin)
    finally:
        monkeypatch.delenv("SUPR_COLUMN_ORDER")
    for r in (auto, via_engine):
        assert r.iterations == ref.iterations and r.restarts == ref.restarts
        assert np.array_equal(r.x, ref.x[newpos]) and np.array_equal(r.y, ref.y)
        assert r.info["objective"] == ref.info["objective"]
        obj = float(np.dot(lp.c, r.x)) + lp.obj_offset
        assert abs(r.info["objective"] - obj) <= 1e-13 * (1 + float(np.dot(np.abs(lp.c), np.abs(r.x))))


def test_cpu_gpu_race_through_the_drop_in():
    """The reference's CPU / GPU race (src/solve.cpp:76-148): the reference's
    own CPU solve_pdhg (oracle/_ref) and this engine on two threads sharing one
    stop flag.  The GPU wins a 300,000-variable LP (verified) and the CPU stops
    early; with a staggered GPU start the CPU wins a small LP and the GPU
    returns at its first poll without converging."""
    import json
    import os
    import subprocess
    from conftest import ROOT
    from oracle import PATHS, available
    if not available("ref"):
        pytest.skip("reference CPU engine (oracle/_ref) not built")
    exe = os.path.join(ROOT, "team-suprema_amd", "_build", "race")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe, PATHS["ref"]], capture_output=True, text=True, timeout=120)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(out)
    assert r.returncode == 0 and out["failures"] == 0, (r.stdout, r.stderr)
    assert out["gpu_wins"]["winner"] == 1 and out["gpu_wins"]["verified_eps"] <= 1e-6
    assert out["cpu_wins"]["winner"] == 0 and out["cpu_wins"]["gpu_converged"] == 0


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_staged_schedules_solve_a_tall_lp(oracle, mode):
    """A (140,000 rows) and A^T (200,000 rows) both run the LDS-staged
    schedules (lanes forced to 4 and 2): the solve must reach a point the
    oracle's verifier accepts, with the same objective in both modes."""
    lp = P.gen_random_lp(140000, 200000, 6, seed=5)
    s, Dr, Dc = P.scale(lp)
    res = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, max_iterations=500000, mode=mode,
                           spmv_lanes_A=4, spmv_lanes_AT=2)
    assert res.converged and (res.info["lanes_A"], res.info["lanes_AT"]) == (4, 2)
    rep = oracle.verify(lp, res.x, res.y)
    assert max(rep[0], rep[1], rep[2]) <= TOL, rep[:3]
    np.testing.assert_allclose([res.eps_P, res.eps_D, res.eps_G], rep[:3], rtol=1e-6, atol=1e-12)
    STAGED_OBJ.setdefault(mode, rep[3])
    if len(STAGED_OBJ) == 2:
        a, b = STAGED_OBJ.values()
        assert abs(a - b) <= OBJ_RTOL * (1 + abs(b))


STAGED_OBJ = {}



def _degenerate_lps():
    def mk(m, n, rows, cols, vals, c, rL, rU, l, u, name):
        return P.LP(m, n, np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32),
                    np.array(vals, dtype=float), np.array(c, dtype=float), np.array(rL, dtype=float),
                    np.array(rU, dtype=float), np.array(l, dtype=float), np.array(u, dtype=float), 0.0, name)
    inf = np.inf
    return {
        # box constraints only: the PDLP form has no rows (eta0 = 1, omega0 = 1 in the reference)
        "no_rows": mk(0, 3, [], [], [], [1, -1, 2], [], [], [0, 0, 0], [1, 2, 3], "no rows"),
        # rows but no nonzeros (a ranged row and a one-sided row that 0 satisfies)
        "no_nonzeros": mk(2, 3, [], [], [], [1, -1, 2], [-1, -inf], [1, 5], [0, 0, 0], [1, 2, 3], "no nonzeros"),
        "one_variable": mk(1, 1, [0], [0], [2.0], [1.0], [2.0], [2.0], [0.0], [inf], "one variable"),
        # a column with no entries, a free row side, an infinite bound
        "empty_column": mk(2, 3, [0, 1], [0, 0], [1.0, 1.0], [1, 1, 0], [1, -inf], [inf, 4], [0, 0, -1],
                           [inf, inf, 1], "empty column"),
    }


DEGENERATE = _degenerate_lps()


@pytest.mark.parametrize("name", sorted(DEGENERATE))
@pytest.mark.parametrize("mode", [P.MODE_FAST_FP64, P.MODE_MIXED, P.MODE_PLAIN_FP64])
def test_degenerate_lps_match_oracle(oracle, name, mode):
    """Shapes the reference's pdlp_run accepts though no generator makes them:
    a form without rows, rows without nonzeros, one variable, an empty
    column.  Each solve converges, the oracle's verifier certifies 1e-6 and
    the objective matches the oracle's own solve."""
    lp = DEGENERATE[name]
    form, _ = P.build_pdlp_form(lp)
    s, Dr, Dc = P.scale(form)
    ref = oracle.pdlp(form, form.with_values(*oracle.scale(form)[:6]), Dr, Dc, 20000, 64, TOL)
    ref_obj = oracle.verify(form, ref.x, ref.y)[3]
    r = P.solve_pdhg_gpu(form, s, Dr, Dc, tol=TOL, mode=mode)
    rep = oracle.verify(form, r.x, r.y)
    print(f"[parity] {name} mode {mode}: {r.iterations} attempts (oracle {ref.iterations}), eps {rep[:3]}, "
          f"objective {rep[3]} (oracle {ref_obj})", file=sys.stderr, flush=True)
    assert ref.converged and r.converged
    assert max(rep[:3]) <= TOL
    assert abs(rep[3] - ref_obj) <= OBJ_RTOL * (1 + abs(ref_obj))
    if mode != P.MODE_PLAIN_FP64:
        sol = P.solve_lp_gpu(lp, tol=TOL, mode=mode)
        assert sol.status == "optimal"
        orig = oracle.verify(lp, sol.x, sol.y)
        assert max(orig[:3]) <= TOL and abs(orig[3] - ref_obj) <= OBJ_RTOL * (1 + abs(ref_obj))


def test_degenerate_lps_in_a_batch(oracle):
    forms = []
    for name in sorted(DEGENERATE):
        form, _ = P.build_pdlp_form(DEGENERATE[name])
        forms.append((form,) + tuple(P.scale(form)))
    with P.BatchEngine(forms, tol=TOL, mode=P.MODE_MIXED, tune=False) as eng:
        res, _ = eng.solve()
    for (form, *_), r in zip(forms, res):
        rep = oracle.verify(form, r.x, r.y)
        assert r.converged and max(rep[:3]) <= TOL, (form.name, rep[:3])


@pytest.mark.parametrize("mode", [P.MODE_FAST_FP64, P.MODE_MIXED])
def test_rebuilt_ell_values_give_the_same_bits(monkeypatch, mode):
    """A^T's FP64 values formed from the scaling factors (every unscaled entry
    the same value: transportation LPs) instead of read: the setup check
    accepts it only if every product equals the stored value, so the solve is
    the one that reads the values, bit for bit (SUPR_ELL_VALUES=stored turns
    the rebuild off).  An LP with varied entries is refused by the check."""
    for S, D in ((60, 60), (300, 400)):
        lp = P.gen_transport(S, D, seed=5)
        s, Dr, Dc = P.scale(lp)
        assert P.bench_kernels(lp, s, Dr, Dc, reps=2)["rebuilt_AT_fp64"] == 1.0
        on = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode)
        monkeypatch.setenv("SUPR_ELL_VALUES", "stored")
        assert P.bench_kernels(lp, s, Dr, Dc, reps=2)["rebuilt_AT_fp64"] == 0.0
        off = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode)
        monkeypatch.delenv("SUPR_ELL_VALUES")
        assert on.converged and on.iterations == off.iterations and on.restarts == off.restarts
        assert np.array_equal(on.x, off.x) and np.array_equal(on.y, off.y)
        print(f"[parity] transport {S}x{D} mode {mode}: rebuilt values, {on.iterations} attempts, same bits",
              file=sys.stderr, flush=True)
    lp = P.gen_random_lp(200, 500, 6, seed=1)   # A^T is ELL, entries all different
    s, Dr, Dc = P.scale(lp)
    kb = P.bench_kernels(lp, s, Dr, Dc, reps=2)
    assert kb["ell_AT"] > 0 and kb["rebuilt_AT_fp64"] == 0.0


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_one_pass_kernels_give_the_same_bits(monkeypatch, golden, mode):
    """Launches that cover every row in one pass use the one-pass iteration
    kernels (matrix slots and epilogue operands loaded before the step
    decision); SUPR_ONE_PASS=off runs the grid-stride kernels everywhere.
    Same terms in the same order, so the same solve bit for bit: ELL A^T with
    several lanes per row (random LP), ELL A^T with values formed from factors
    (transportation), CSR on both sides with rows longer than one prefetched
    group (Netlib, power-law), bounded columns and inequality rows, lanes
    pinned from 1 to 64, and an FP64 dual over rows of 200 and 1,000 entries
    in the flat one-pass form (each lane's entries in chunks of 32, the first
    loaded before the state; transportation 200 x 1,000)."""
    cases = [("random", P.gen_random_lp(400, 1000, 6, seed=13), {}),
             ("transport", P.gen_transport(60, 80, seed=4), {}),
             ("transport_long_rows", P.gen_transport(200, 1000, seed=4), {}),
             ("power_law", P.gen_power_law_lp(1000, 4000, seed=3), {}),
             ("power_law_l1", P.gen_power_law_lp(1000, 4000, seed=3), dict(spmv_lanes_A=1, spmv_lanes_AT=2)),
             ("random_l64", P.gen_random_lp(400, 1000, 6, seed=13), dict(spmv_lanes_A=64, spmv_lanes_AT=4))]
    for name in ("afiro", "blend", "sc50a", "tiny_range", "israel"):
        cases.append((name, golden_lp(golden[name]), {}))
    for name, lp, pin in cases:
        s, Dr, Dc = P.scale(lp)
        on = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=400000, **pin)
        monkeypatch.setenv("SUPR_ONE_PASS", "off")
        off = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=400000, **pin)
        monkeypatch.delenv("SUPR_ONE_PASS")
        assert on.converged, name
        assert on.iterations == off.iterations and on.restarts == off.restarts, name
        assert np.array_equal(on.x, off.x) and np.array_equal(on.y, off.y), name
        assert on.info["objective"] == off.info["objective"], name
        print(f"[parity] one-pass {name} mode {mode}: {on.iterations} attempts, lanes "
              f"{on.info['lanes_A']}/{on.info['lanes_AT']}, same bits", file=sys.stderr, flush=True)


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_row_pairs_follow_the_one_row_per_lane_solve(monkeypatch, mode):
    """A^T in ELL with one lane per row, an even number of rows and more rows
    than one pass of the capped grid: the primal launches give each lane two
    adjacent rows (SUPR_ROW_PAIRS=off: one row per lane).  Each row's values
    are the same bits; only the grouping of the step statistic's partial sums
    differs, so 64 attempts of both schedules, with a KKT check every 8
    accepted steps and restarts to the average, agree to rounding (the random
    LP in FP64: identical after 2 attempts, 1e-15 after 4, 2e-13 after 64, the
    last-bit differences growing as the trajectories drift apart; 1e-7 after
    160) -- a wrong row, index, bound or average update would not.  Covered:
    16-bit indices with values formed from factors and x >= 0
    (transportation), the same with stored values, and 32-bit indices, stored
    values and finite upper bounds (random LP with 70,000 rows)."""
    rtol = 1e-9 if mode == P.MODE_FAST_FP64 else 1e-5
    cases = [("transport", P.gen_transport(1000, 600, seed=9), {}),
             ("transport_stored", P.gen_transport(1000, 600, seed=9), {"SUPR_ELL_VALUES": "stored"}),
             ("random_wide", P.gen_random_lp(70000, 600000, 3, seed=21), {})]
    for name, lp, env in cases:
        s, Dr, Dc = P.scale(lp)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        kb = P.bench_kernels(lp, s, Dr, Dc, reps=2)
        assert kb["ell_AT"] > 0 and kb["lanes_AT"] == 1 and lp.n % 2 == 0 and lp.n > 2048 * 256, name
        run = dict(tol=1e-12, mode=mode, max_iterations=64, check_every=8)
        on = P.solve_pdhg_gpu(lp, s, Dr, Dc, **run)
        monkeypatch.setenv("SUPR_ROW_PAIRS", "off")
        off = P.solve_pdhg_gpu(lp, s, Dr, Dc, **run)
        monkeypatch.delenv("SUPR_ROW_PAIRS")
        for k in env:
            monkeypatch.delenv(k)
        assert on.iterations == off.iterations == 64, name
        assert on.restarts == off.restarts >= 1, (name, on.restarts, off.restarts)
        dx = np.abs(on.x - off.x).max() / (1.0 + np.abs(off.x).max())
        dy = np.abs(on.y - off.y).max() / (1.0 + np.abs(off.y).max())
        dobj = abs(on.info["objective"] - off.info["objective"]) / (1.0 + abs(off.info["objective"]))
        assert dx <= rtol and dy <= rtol and dobj <= rtol, (name, dx, dy, dobj)
        print(f"[parity] row pairs {name} mode {mode}: 64 attempts, {on.restarts} restarts, max rel "
              f"diff x {dx:.1e} y {dy:.1e} objective {dobj:.1e}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_early_index_slots_give_the_same_bits(monkeypatch, mode):
    """FP64 row pairs in one pass of the grid run k_primal_pairs, which loads
    each pair's first two index slots before the decision prologue;
    SUPR_EARLY_SLOTS=off runs k_primal's row-pair path.  The same terms, the
    same epilogue and the same partial sums (each lane's two squares added by
    an explicit fma in both), so the same solve bit for bit: 16-bit indices
    with values formed from factors (transportation, to convergence), stored
    values, and 32-bit indices with finite upper bounds (random LP, to an
    iteration limit)."""
    cases = [("transport", P.gen_transport(1000, 600, seed=9), {}, 200000),
             ("transport_stored", P.gen_transport(1000, 600, seed=9), {"SUPR_ELL_VALUES": "stored"}, 3000),
             ("random_wide", P.gen_random_lp(70000, 600000, 3, seed=21), {}, 1500)]
    for name, lp, env, limit in cases:
        s, Dr, Dc = P.scale(lp)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        on = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=limit)
        monkeypatch.setenv("SUPR_EARLY_SLOTS", "off")
        off = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL, mode=mode, max_iterations=limit)
        monkeypatch.delenv("SUPR_EARLY_SLOTS")
        for k in env:
            monkeypatch.delenv(k)
        assert on.iterations == off.iterations and on.restarts == off.restarts, name
        assert np.array_equal(on.x, off.x) and np.array_equal(on.y, off.y), name
        assert on.info["objective"] == off.info["objective"], name
        print(f"[parity] early slots {name} mode {mode}: {on.iterations} attempts, same bits",
              file=sys.stderr, flush=True)


def test_device_memory_cache_reuse_and_release():
    """Device memory a call releases is kept and reused by the next call (no
    hipMalloc / hipFree between one-shot calls); results do not depend on
    whether a block was fresh or reused, and releasing the cache between calls
    changes nothing either."""
    lp = P.gen_transport(200, 300, seed=9)
    s, Dr, Dc = P.scale(lp)
    runs = [P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL) for _ in range(2)]
    P.release_cached_memory()
    runs.append(P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL))
    for r in runs[1:]:
        assert r.iterations == runs[0].iterations
        assert np.array_equal(r.x, runs[0].x) and np.array_equal(r.y, runs[0].y)
    P.release_cached_memory()


def test_device_memory_cache_off_gives_the_same_bits(tmp_path):
    """SUPR_DEVICE_CACHE_MB=0 (every block freed at once, as without the
    cache): a fresh process solves the same LP to the same bits."""
    import json
    import os
    import subprocess
    from conftest import ROOT
    lp = P.gen_transport(120, 150, seed=4)
    s, Dr, Dc = P.scale(lp)
    here = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=TOL)
    script = tmp_path / "solve.py"
    script.write_text(
        "import json, sys\n"
        f"sys.path.insert(0, {os.path.join(ROOT, 'team-suprema_amd')!r})\n"
        "import pdlp_amd as P\n"
        "lp = P.gen_transport(120, 150, seed=4)\n"
        "s, Dr, Dc = P.scale(lp)\n"
        "r = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=1e-6)\n"
        "print(json.dumps({'iterations': r.iterations, 'x': r.x.tobytes().hex()[:4096], 'sum': float(r.x.sum())}))\n")
    env = dict(os.environ, SUPR_DEVICE_CACHE_MB="0")
    out = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr[-2000:]
    got = json.loads(out.stdout.strip().splitlines()[-1])
    assert got["iterations"] == here.iterations
    assert got["x"] == here.x.tobytes().hex()[:4096] and got["sum"] == float(here.x.sum())
