"""GPU parity of the batched engine (supr_batch_*): every instance of a mixed
batch against the FP64 oracle (same tolerances as test_gpu_parity.py), and
an instance's result is the same bits alone, in the whole batch and in
another mix.  Instances: netlib fixtures, a generated transportation LP and
generated blending LPs, 10 in all, so the layout compaction (more than 8
instances, half of them finished) runs as well."""
import sys

import numpy as np
import pytest

import pdlp_amd as P
from conftest import golden_lp

pytestmark = pytest.mark.gpu

TOL = 1e-6
OBJ_RTOL = 1e-5   # two 1e-6-certified points differ by a few 1e-6 on the degenerate Netlib LPs
MAX_IT = 60000


def say(*a):
    print("[batch]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def batch_cases(oracle, golden):
    cases = []
    # random_2000x5000_s42 runs both sides on the LDS-staged tiles (>= 1024 rows, L >= 2)
    for name in ["afiro", "tiny_range", "sc50a", "sc105", "blend", "israel", "transport_30x40_s7",
                 "random_2000x5000_s42"]:
        case = golden[name]
        cases.append((name, golden_lp(case), case["ref"]["objective_form"]))
    for k, (K, Pp) in enumerate([(20, 10), (35, 12)]):
        form, _ = P.build_pdlp_form(P.gen_blend(K, Pp, seed=100 + k))
        s = oracle.scale(form)
        ref = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 200000, 64, TOL)
        assert ref.converged
        cases.append((f"blend_{K}x{Pp}", form, oracle.verify(form, ref.x, ref.y)[3]))
    out = []
    for name, form, obj in cases:
        scaled, Dr, Dc = P.scale(form)
        out.append((name, form, scaled, Dr, Dc, obj))
    return out


def _forms(cases):
    return [(f, s, dr, dc) for _, f, s, dr, dc, _ in cases]


def _check(oracle, cases, results):
    for (name, form, _, _, _, ref_obj), r in zip(cases, results):
        assert r.converged, (name, r.iterations, r.eps_P, r.eps_D, r.eps_G)
        rep = oracle.verify(form, r.x, r.y)
        assert max(rep[:3]) <= TOL, (name, rep[:3])
        np.testing.assert_allclose([r.eps_P, r.eps_D, r.eps_G], rep[:3], rtol=1e-6, atol=1e-12)
        assert abs(rep[3] - ref_obj) / (1 + abs(ref_obj)) <= OBJ_RTOL, (name, rep[3], ref_obj)
        assert abs(r.info["objective"] - rep[3]) <= 1e-9 * (1 + abs(rep[3])), name


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_batch_matches_oracle_and_is_mix_independent(oracle, batch_cases, mode):
    with P.BatchEngine(_forms(batch_cases), tol=TOL, max_iterations=MAX_IT, mode=mode, tune=False,
                       lanes_bias=1.0, max_blocks=32) as eng:
        res, secs = eng.solve()
        say(f"mode {mode}: {len(res)} instances in {secs:.3f} s, attempts",
            [r.iterations for r in res], "fp32", [r.info["iterations_fp32"] for r in res])
        assert secs > 0
        _check(oracle, batch_cases, res)
        # alone, and in a different mix, each result is the same bits
        for i in (0, 5, len(res) - 1):
            alone, _ = eng.solve([i])
            assert alone[0].iterations == res[i].iterations
            assert np.array_equal(alone[0].x, res[i].x) and np.array_equal(alone[0].y, res[i].y)
        sub = [len(res) - 1, 2, 6]
        mix, _ = eng.solve(sub)
        for r, i in zip(mix, sub):
            assert np.array_equal(r.x, res[i].x) and np.array_equal(r.y, res[i].y)
        say(f"mode {mode}: alone and sub-mix results identical")


def test_batch_direct_launches_match_graph_replay(batch_cases):
    """graph_attempts < 0 launches each replay's kernels on the stream
    directly (for kernel tracers): the same kernels in the same order, so the
    same bits as the captured graph's replays, compaction included."""
    runs = []
    for g in (16, -16):
        with P.BatchEngine(_forms(batch_cases), tol=TOL, max_iterations=MAX_IT, tune=False,
                           lanes_bias=1.0, max_blocks=32, graph_attempts=g) as eng:
            runs.append(eng.solve()[0])
    for a, b in zip(*runs):
        assert a.iterations == b.iterations and np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    say("direct launches identical to graph replays on", len(runs[0]), "instances")


def test_batch_tuned_config_stays_within_tolerance(oracle, batch_cases):
    cases = batch_cases[:4]
    with P.BatchEngine(_forms(cases), tol=TOL, max_iterations=MAX_IT, tune=True) as eng:
        bias, cap = eng.config
        assert bias in (0.5, 1.0, 2.0) and cap == 512
        res, secs = eng.solve()
        say(f"tuned bias {bias} cap {cap}: {secs:.3f} s")
        _check(oracle, cases, res)


def test_batch_iteration_limit_and_bad_ids():
    lp = P.gen_random_lp(200, 500, 6, seed=1)
    scaled, Dr, Dc = P.scale(lp)
    with P.BatchEngine([(lp, scaled, Dr, Dc)] * 3, tol=TOL, max_iterations=300, tune=False) as eng:
        res, _ = eng.solve()
        for r in res:
            assert not r.converged and r.iterations == 300 and r.info["has_point"] == 1
        assert res[0].iterations == res[2].iterations and np.array_equal(res[0].x, res[2].x)
        with pytest.raises(P.SupremaError):
            eng.solve([0, 0])
        with pytest.raises(ValueError):
            eng.solve([3])
    with P.BatchEngine([(lp, scaled, Dr, Dc)], tol=TOL, max_iterations=0, tune=False) as eng:
        res, _ = eng.solve()
        assert res[0].iterations == 0 and not res[0].converged


@pytest.mark.skip(reason="withdrawn design: a persistent kernel running the whole PDHG loop per "
                         "workgroup did not return on the GPU for a 1-instance, 300-attempt solve; "
                         "its cause was not found, so that kernel is not built; batched mode uses "
                         "the segmented device-resident loop tested above")
def test_persistent_whole_loop_batch_kernel():
    pass
