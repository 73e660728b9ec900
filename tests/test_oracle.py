"""The FP64 oracle (oracle/pdlp_oracle.c) pinned against the reference.

1. Golden fixtures written by the reference's own CPU engine (tests/golden/
   golden.json, made by tests/golden/make_golden.py): the oracle must
   reproduce them bit for bit -- scaled form, PDLP form, iteration and
   restart counts, eps, and the SHA-256 of x and y.
2. The live reference build (oracle/_ref), where present.
3. Published Netlib optima (tools/audit_lp.cpp:148-160).
4. Closed forms: dense brute-force SpMV, verifier cases with known eps,
   scaling invariants, tiny LPs with known optima.
A deliberately changed parameter must break the golden match (the fixtures
can tell a wrong trajectory from a right one).
"""
import hashlib

import numpy as np
import pytest

import pdlp_amd as P
from conftest import golden_lp, load_fixture


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def oracle_form(oracle, lp):
    dims, ri, ci, arrs, perm, signs = oracle.pdlp_form(lp)
    form = P.LP(int(dims[0]), int(dims[1]), ri, ci, *arrs, lp.obj_offset, lp.name)
    return form, dims, sha(ri, ci, *arrs, perm, signs)


def case_lp(oracle, case):
    if case["kind"] == "mps":
        lp, _ = load_fixture(case["name"])
        form, dims, fsha = oracle_form(oracle, lp)
        assert [int(v) for v in dims] == case["form_dims"]
        assert fsha == case["form_sha"], "PDLP form differs from the reference's build_pdlp_form"
        return form
    return golden_lp(case)


# Small and medium golden cases run in the default CPU suite; all are quick.
GOLDEN_CASES = ["afiro", "adlittle", "blend", "sc50a", "sc50b", "sc105", "kb2", "scagr7", "israel",
                "share1b", "beaconfd", "tiny_range", "med", "illcond", "random_200x500_s1",
                "random_2000x5000_s42", "transport_60x60_s42", "transport_30x40_s7",
                "power_law_1000x4000_s7", "blend_25x40_s103"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_reference_golden(oracle, golden, name):
    case = golden[name]
    form = case_lp(oracle, case)
    r = case["ref"]
    s = oracle.scale(form)
    assert sha(*s) == r["scaled_sha"], "scale() output differs from the reference"
    scaled = form.with_values(*s[:6])
    out = oracle.pdlp(form, scaled, s[6], s[7], r["max_iterations"], r["check_every"], r["tol"])
    assert out.iterations == r["iterations"]
    assert out.restarts == r["restarts"]
    assert out.converged == r["converged"]
    assert [out.eps_P, out.eps_D, out.eps_G] == r["eps"]
    assert sha(out.x) == r["x_sha"] and sha(out.y) == r["y_sha"]


def test_generators_match_golden(golden):
    for case in golden.values():
        if case["kind"] == "mps":
            continue
        lp = golden_lp(case)
        assert sha(lp.row, lp.col, lp.val, lp.c, lp.rL, lp.rU, lp.l, lp.u) == case["lp_sha"]


def test_golden_detects_a_changed_trajectory(oracle, golden):
    """The fixtures pin the trajectory: a different check cadence (a plausible
    porting slip) must not reproduce the reference's iteration count."""
    case = golden["random_200x500_s1"]
    form = case_lp(oracle, case)
    s = oracle.scale(form)
    out = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 200000, 63, 1e-6)
    assert (out.iterations, sha(out.x)) != (case["ref"]["iterations"], case["ref"]["x_sha"])


@pytest.mark.parametrize("name", ["afiro", "sc50a", "random_200x500_s1", "transport_30x40_s7"])
def test_oracle_matches_live_reference(oracle, ref, golden, name):
    form = case_lp(oracle, golden[name])
    so, sr = oracle.scale(form), ref.scale(form)
    assert all(np.array_equal(a, b) for a, b in zip(so, sr))
    scaled = form.with_values(*so[:6])
    a = oracle.pdlp(form, scaled, so[6], so[7], 200000, 64, 1e-6)
    b = ref.pdlp(form, scaled, so[6], so[7], 200000, 64, 1e-6)
    assert (a.iterations, a.restarts, a.converged) == (b.iterations, b.restarts, b.converged)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    x = np.linspace(-1, 2, form.n)
    y = np.linspace(3, -1, form.m)
    assert np.array_equal(oracle.verify(form, x, y), ref.verify(form, x, y))


NETLIB = ["afiro", "adlittle", "blend", "sc50a", "sc50b", "sc105", "kb2", "scagr7", "israel",
          "share1b", "beaconfd"]


@pytest.mark.parametrize("name", NETLIB)
def test_oracle_reaches_published_netlib_optimum(oracle, golden, name):
    """audit_lp's rule: |obj - published| / (1 + |published|) <= 1e-4
    (tools/audit_lp.cpp:270-277), evaluated on the original LP after
    extract_solution, with the sign convention of report_objective."""
    case = golden[name]
    lp, maximize = load_fixture(name)
    dims, ri, ci, arrs, perm, signs = oracle.pdlp_form(lp)
    form = P.LP(int(dims[0]), int(dims[1]), ri, ci, *arrs, lp.obj_offset, name)
    s = oracle.scale(form)
    out = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 500000, 64, 1e-6)
    assert out.converged
    x = out.x[:lp.n]
    y = np.zeros(lp.m)
    y[perm] = signs * out.y
    rep = oracle.verify(lp, x, y)
    obj = -rep[3] if maximize else rep[3]
    pub = case["published_optimum"]
    assert abs(obj - pub) / (1 + abs(pub)) <= 1e-4
    assert max(rep[0], rep[1], rep[2]) <= 1e-5


def test_spmv_matches_dense_brute_force(oracle):
    rng = np.random.default_rng(3)
    m, n = 37, 23
    dense = np.where(rng.random((m, n)) < 0.2, rng.standard_normal((m, n)), 0.0)
    ri, ci = np.nonzero(dense)
    perm = rng.permutation(len(ri))          # arbitrary COO order
    ri, ci = ri[perm], ci[perm]
    lp = P.LP(m, n, ri, ci, dense[ri, ci], np.zeros(n), np.zeros(m), np.zeros(m), np.zeros(n),
              np.zeros(n))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    np.testing.assert_allclose(oracle.spmv(lp, x), dense @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.spmv(lp, y, transposed=True), dense.T @ y, rtol=1e-13,
                               atol=1e-13)
    # CSR keeps the COO order inside each row (stable counting sort)
    ptr, idx, _ = oracle.csr(lp)
    for i in range(m):
        expect = ci[ri == i]
        assert np.array_equal(idx[ptr[i]:ptr[i + 1]], expect)


def _lp(A, c, rL, rU, l, u):
    A = np.asarray(A, dtype=float)
    ri, ci = np.nonzero(A)
    return P.LP(A.shape[0], A.shape[1], ri, ci, A[ri, ci], c, rL, rU, l, u)


def test_verify_closed_forms(oracle):
    inf = np.inf
    # 2 equality rows; x satisfies row 0, violates row 1 by delta
    A = [[1.0, 1.0], [1.0, -1.0]]
    b = np.array([3.0, 1.0])
    lp = _lp(A, [1.0, 2.0], b, b, [0.0, 0.0], [inf, inf])
    delta = 0.25
    x = np.array([2.0 + delta / 2, 1.0 - delta / 2])   # row0 = 3, row1 = 1 + delta
    rep = oracle.verify(lp, x, np.zeros(2))
    assert rep[0] == pytest.approx(delta / (1 + 3.0), rel=1e-15)
    assert rep[5] == pytest.approx(delta)
    # dual sign violation: column with l = -inf, z = c - A'y > 0 counts in eps_D
    lp2 = _lp([[1.0, 0.0]], [5.0, 1.0], [0.0], [0.0], [-inf, 0.0], [inf, inf])
    rep2 = oracle.verify(lp2, np.zeros(2), np.array([2.0]))   # z = (3, 1): z0 > 0 on free-below col
    assert rep2[1] == pytest.approx(3.0 / (1 + 5.0))
    # y > 0 on a row with rL = -inf (a <= row) is a sign violation
    lp3 = _lp([[1.0]], [1.0], [-inf], [4.0], [0.0], [inf])
    rep3 = oracle.verify(lp3, np.zeros(1), np.array([0.5]))
    assert rep3[1] == pytest.approx(0.5 / (1 + 1.0))
    # optimal primal-dual pair of min x1 + 2 x2 s.t. x1 + x2 = 10, x >= 0:
    # x = (10, 0), y = 1, z = (0, 1) -> all eps 0, objective 10
    lp4 = _lp([[1.0, 1.0]], [1.0, 2.0], [10.0], [10.0], [0.0, 0.0], [inf, inf])
    rep4 = oracle.verify(lp4, np.array([10.0, 0.0]), np.array([1.0]))
    assert rep4[0] == 0 and rep4[1] == 0 and rep4[2] == 0 and rep4[3] == 10.0 and rep4[4] == 10.0
    # a suboptimal dual gives the textbook gap |cx - D| / (1 + |cx| + |D|)
    rep5 = oracle.verify(lp4, np.array([10.0, 0.0]), np.array([0.5]))  # D = 5 + 0 (z = (.5, 1.5))
    assert rep5[2] == pytest.approx(abs(10 - 5) / (1 + 10 + 5))


def test_scale_invariants(oracle):
    lp = P.gen_random_lp(60, 150, 6, seed=11, mag_decades=3)
    val_s, c_s, rL_s, rU_s, l_s, u_s, Dr, Dc = oracle.scale(lp)
    assert (Dr > 0).all() and (Dc > 0).all()
    # the scaled LP is diag(Dr) A diag(Dc) with c, rows and bounds scaled to match
    np.testing.assert_allclose(val_s, Dr[lp.row] * lp.val * Dc[lp.col], rtol=1e-12)
    np.testing.assert_allclose(c_s, Dc * lp.c, rtol=1e-15)
    np.testing.assert_allclose(rL_s, Dr * lp.rL, rtol=1e-15)
    np.testing.assert_allclose(rU_s, Dr * lp.rU, rtol=1e-15)
    np.testing.assert_allclose(l_s, lp.l / Dc, rtol=1e-15)
    np.testing.assert_allclose(u_s, lp.u / Dc, rtol=1e-15)
    # Ruiz alone (no Pock-Chambolle effect on a 1x1: its row and column sums
    # are the entry itself) drives |a| to 1: a = 4 -> 1, Dr * Dc = 1/4
    one = _lp([[4.0]], [1.0], [2.0], [2.0], [0.0], [10.0])
    v, _, _, _, _, _, dr, dc = oracle.scale(one, ruiz_passes=40)
    assert v[0] == pytest.approx(1.0, rel=1e-9)
    assert dr[0] * dc[0] == pytest.approx(0.25, rel=1e-9)
    # zero passes: only Pock-Chambolle, whose factors are 1/sqrt of the abs row
    # and column sums of A
    s0 = oracle.scale(lp, ruiz_passes=0)
    rows = np.bincount(lp.row, weights=np.abs(lp.val), minlength=lp.m)
    cols = np.bincount(lp.col, weights=np.abs(lp.val), minlength=lp.n)
    np.testing.assert_allclose(s0[6], 1 / np.sqrt(rows), rtol=1e-12)
    np.testing.assert_allclose(s0[7], 1 / np.sqrt(cols), rtol=1e-12)


def test_scale_keeps_integer_columns_unscaled(oracle):
    lp = P.gen_random_lp(30, 60, 6, seed=5)
    integ = np.zeros(lp.n, dtype=np.int8)
    integ[::3] = 1
    Dc_all = oracle.scale(lp)[7]
    Dc_int = oracle.scale(lp, integer=integ)[7]
    assert not np.all(Dc_all[::3] == 1.0)
    assert np.all(Dc_int[::3] == 1.0)


def test_tiny_lps_with_known_optima(oracle):
    inf = np.inf
    # tiny_range: min x1 + 2 x2 s.t. 10 <= x1 + x2 <= 14, x >= 0 -> 10
    lp, _ = load_fixture("tiny_range")
    form, _, _ = oracle_form(oracle, lp)
    s = oracle.scale(form)
    out = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 50000, 64, 1e-9)
    assert out.converged
    assert oracle.verify(lp, out.x[:lp.n], out.y)[3] == pytest.approx(10.0, abs=1e-4)
    # a box LP with one equality: min -x1 - x2 s.t. x1 + 2 x2 = 4, 0 <= x <= 3 -> x = (3, .5), obj -3.5
    lp2 = _lp([[1.0, 2.0]], [-1.0, -1.0], [4.0], [4.0], [0.0, 0.0], [3.0, 3.0])
    s2 = oracle.scale(lp2)
    out2 = oracle.pdlp(lp2, lp2.with_values(*s2[:6]), s2[6], s2[7], 100000, 64, 1e-9)
    assert out2.converged
    np.testing.assert_allclose(out2.x, [3.0, 0.5], atol=1e-6)
    assert oracle.verify(lp2, out2.x, out2.y)[3] == pytest.approx(-3.5, abs=1e-6)
    del inf


def test_stop_at_iteration_limit_returns_best_point(oracle):
    lp = P.gen_random_lp(200, 500, 6, seed=1)
    s = oracle.scale(lp)
    out = oracle.pdlp(lp, lp.with_values(*s[:6]), s[6], s[7], 300, 64, 1e-6)
    assert not out.converged and out.iterations == 300 and out.has_point
    rep = oracle.verify(lp, out.x, out.y)
    assert [out.eps_P, out.eps_D, out.eps_G] == [rep[0], rep[1], rep[2]]


# ---------------------------------------------------------- extract_solution --

@pytest.fixture(scope="module")
def extract_golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "extract.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


EXTRACT_CASES = ["afiro", "adlittle", "blend", "sc50a", "sc50b", "sc105", "kb2", "scagr7",
                 "israel", "share1b", "beaconfd", "tiny_range", "med", "illcond"]


@pytest.mark.parametrize("name", EXTRACT_CASES)
def test_extract_solution_matches_reference_golden(oracle, golden, extract_golden, name):
    """The reference's extract_solution() on its own solution of each MPS case
    (tests/golden/extract.json, through oracle/_ref): the oracle's
    extract of the oracle's (bit-identical) solution gives the same x, y bits,
    and the oracle's verify() of them on the original LP the same report."""
    case, want = golden[name], extract_golden[name]
    lp, _ = load_fixture(name)
    dims, ri, ci, arrs, perm, signs = oracle.pdlp_form(lp)
    form = P.LP(int(dims[0]), int(dims[1]), ri, ci, *arrs, lp.obj_offset, name)
    s = oracle.scale(form)
    r = case["ref"]
    out = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], r["max_iterations"],
                      r["check_every"], r["tol"])
    x, y = oracle.extract_solution(lp.n, lp.m, perm, signs, out.x, out.y)
    assert (len(x), len(y)) == (want["n_orig"], want["m_orig"])
    assert sha(x) == want["x_sha"] and sha(y) == want["y_sha"]
    assert list(oracle.verify(lp, x, y)) == want["verify_on_lp"]


def test_extract_solution_closed_form(oracle):
    """Hand-built form metadata: 3 original columns plus 2 slack columns, 4
    original rows of which row 1 is free (dropped from the form), the others
    permuted with one sign flip.  x keeps the first n_orig entries; y[perm[p]]
    = sign[p] y2[p]; the free row's multiplier is 0."""
    perm = np.array([2, 0, 3], dtype=np.int32)
    signs = np.array([1.0, -1.0, 1.0])
    x2 = np.array([1.5, -2.0, 3.25, 9.0, 8.0])
    y2 = np.array([0.5, 7.0, -4.0])
    x, y = oracle.extract_solution(3, 4, perm, signs, x2, y2)
    assert np.array_equal(x, [1.5, -2.0, 3.25])
    assert np.array_equal(y, [-7.0, 0.0, 0.5, -4.0])
    # a slip that indexes y by the form's position instead of row_perm fails here
    assert not np.array_equal(y, np.concatenate([signs * y2, [0.0]]))


def test_extract_solution_matches_live_reference(oracle, ref):
    rng = np.random.default_rng(5)
    for m_orig, n_orig, m2, n2 in ((7, 5, 6, 9), (30, 40, 30, 44), (1, 1, 1, 1)):
        perm = rng.permutation(m_orig)[:m2].astype(np.int32)
        signs = rng.choice([-1.0, 1.0], m2)
        x2, y2 = rng.standard_normal(n2), rng.standard_normal(m2)
        a = oracle.extract_solution(n_orig, m_orig, perm, signs, x2, y2)
        b = ref.extract_solution(n_orig, m_orig, perm, signs, x2, y2)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# -------------------------------------------------------------- scalar rules --
# orc_pdlp takes every step and check decision through these functions, so the
# golden trajectories above pin them; these closed forms name what they compute.

def test_step_rule_closed_forms(oracle):
    # inter = 0: eta_bar infinite, accepted, eta grows by (1 + k1^-0.6)
    acc, eb, en = oracle.step_rule(0.5, 2.0, 0, 1.0, 1.0, 0.0)
    assert acc and eb == float("inf") and en == (1.0 + 2.0 ** -0.6) * 0.5
    # eta_bar = (omega dx2 + dy2 / omega) / (2 |inter|) = (2*3 + 8/2) / (2*5) = 1
    acc, eb, en = oracle.step_rule(0.75, 2.0, 8, 3.0, 8.0, -5.0)
    assert acc and eb == 1.0
    assert en == min((1.0 - 10.0 ** -0.3) * 1.0, (1.0 + 10.0 ** -0.6) * 0.75)
    # eta > eta_bar: rejected, and the next eta is the shrunk eta_bar
    acc, eb, en = oracle.step_rule(1.5, 2.0, 8, 3.0, 8.0, 5.0)
    assert not acc and en == (1.0 - 10.0 ** -0.3) * 1.0
    # eta == eta_bar is accepted (<=)
    assert oracle.step_rule(1.0, 2.0, 8, 3.0, 8.0, 5.0)[0]


def test_check_rule_closed_forms(oracle):
    inf = float("inf")
    # converged on the candidate with the smaller merit (the average here)
    r = oracle.check_rule([1e-5, 0, 0], [1e-7, 2e-7, 0], 1e-6, 3, 10, inf, inf, inf)
    assert r["converged"] and r["use_avg"] and r["best_which"] == 2
    # first check: merit_last_restart = inf, so "sufficient" restarts
    r = oracle.check_rule([3e-3, 4e-3, 0], [1.0, 0, 0], 1e-6, 3, 10, inf, inf, inf)
    assert not r["converged"] and not r["use_avg"] and r["restart"]
    assert r["cand_merit"] == 5e-3 and r["best_merit"] == 5e-3 and r["best_which"] == 1
    assert r["merit_prev_cand"] == 5e-3
    # no restart: merit above 0.8 x the last restart's, few steps since
    r = oracle.check_rule([0.9, 0, 0], [1.0, 0, 0], 1e-6, 3, 10, 1.0, 0.5, 0.7)
    assert not r["restart"] and r["best_which"] == 0 and r["best_merit"] == 0.5
    # necessary: <= 0.8 x last and worse than the previous candidate
    assert oracle.check_rule([0.7, 0, 0], [1.0, 0, 0], 1e-6, 3, 10, 1.0, 0.5, 0.6)["restart"]
    # artificial: since_restart >= 0.36 accepted
    assert oracle.check_rule([0.9, 0, 0], [1.0, 0, 0], 1e-6, 36, 100, 1.0, 0.5, 0.7)["restart"]
    assert not oracle.check_rule([0.9, 0, 0], [1.0, 0, 0], 1e-6, 35, 100, 1.0, 0.5, 0.7)["restart"]


def test_primal_weight_closed_forms(oracle):
    import math
    assert oracle.primal_weight(4.0, 2.0, 8.0) == math.exp(0.5 * math.log(4.0) + 0.5 * math.log(4.0))
    assert oracle.primal_weight(4.0, 1e-11, 8.0) == 4.0 and oracle.primal_weight(4.0, 2.0, 0.0) == 4.0
