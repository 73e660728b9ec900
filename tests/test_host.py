"""Host-side product functions (no GPU): scale, PDLP form, verify, extract,
generators, argument checking -- each compared with the oracle."""
import hashlib

import numpy as np
import pytest

import pdlp_amd as P
from conftest import load_fixture


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", ["afiro", "blend", "share1b", "israel", "tiny_range", "med"])
def test_scale_and_form_bit_identical_to_oracle(oracle, golden, name):
    lp, _ = load_fixture(name)
    form, meta = P.build_pdlp_form(lp)
    dims, ri, ci, arrs, perm, signs = oracle.pdlp_form(lp)
    assert (form.m, form.n, form.nnz, meta.n_slack) == tuple(int(v) for v in dims)
    assert sha(form.row, form.col, form.val, form.c, form.rL, form.rU, form.l, form.u,
               meta.row_perm, meta.row_signs) == golden[name]["form_sha"]
    scaled, Dr, Dc = P.scale(form)
    s = oracle.scale(form)
    mine = [scaled.val, scaled.c, scaled.rL, scaled.rU, scaled.l, scaled.u, Dr, Dc]
    assert all(np.array_equal(a, b) for a, b in zip(mine, s))
    assert sha(*mine) == golden[name]["ref"]["scaled_sha"]


def test_scale_integer_columns_match_oracle(oracle):
    lp = P.gen_random_lp(40, 90, 6, seed=9, mag_decades=2)
    integ = (np.arange(lp.n) % 4 == 0).astype(np.int8)
    scaled, Dr, Dc = P.scale(lp, integer=integ)
    s = oracle.scale(lp, integer=integ)
    assert np.array_equal(Dc, s[7]) and np.array_equal(scaled.val, s[0])


def test_verify_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    for name in ["afiro", "share1b", "tiny_range"]:
        lp, _ = load_fixture(name)
        x = rng.standard_normal(lp.n) * 3
        y = rng.standard_normal(lp.m)
        mine = P.verify(lp, x, y)
        ref = oracle.verify(lp, x, y)
        got = [mine.eps_P, mine.eps_D, mine.eps_G, mine.objective, mine.dual_objective,
               mine.max_row_violation, mine.max_bound_violation]
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)


def test_extract_solution_inverts_the_form():
    lp, _ = load_fixture("tiny_range")
    form, meta = P.build_pdlp_form(lp)
    assert meta.n_slack == 1 and form.n == lp.n + 1
    x2 = np.arange(form.n, dtype=float)
    y2 = np.arange(1, form.m + 1, dtype=float)
    x, y = P.extract_solution(meta, x2, y2)
    assert np.array_equal(x, x2[:lp.n])
    assert np.array_equal(y[meta.row_perm], meta.row_signs * y2)


def test_transport_generator_structure():
    S, D = 7, 9
    lp = P.gen_transport(S, D, seed=3)
    assert (lp.m, lp.n, lp.nnz) == (S + D, S * D, 2 * S * D)
    assert np.all(lp.val == 1.0)
    # column (i, j) = i*D + j has rows i and S + j
    cols = np.arange(S * D)
    assert np.array_equal(lp.row[0::2], cols // D) and np.array_equal(lp.row[1::2], S + cols % D)
    assert np.array_equal(lp.rL, lp.rU)
    assert lp.rL[:S].sum() == lp.rL[S:].sum()       # balanced
    assert np.all((lp.rL[:S] >= 50) & (lp.rL[:S] <= 150))
    assert np.all((lp.c >= 1) & (lp.c <= 100) & (lp.c == np.round(lp.c)))
    assert np.all(lp.l == 0) and np.all(np.isinf(lp.u))


def test_random_generator_is_feasible_and_bounded():
    m, n = 50, 120
    lp = P.gen_random_lp(m, n, 6, seed=4)
    assert lp.nnz == 6 * n
    for j in range(n):
        rows = lp.row[lp.col == j]
        assert len(rows) == 6 and len(set(rows.tolist())) == 6
    assert np.array_equal(lp.rL, lp.rU)
    assert np.all(lp.l == 0) and np.all(lp.u == 10)
    # a power-law degree vector and per-column upper bounds U[1, 100]
    deg = np.clip(np.arange(n) % 9 + 2, 2, 10).astype(np.int32)
    lp2 = P.gen_random_lp(m, n, 0, seed=4, u_hi=-100.0, mag_decades=3, col_deg=deg)
    assert lp2.nnz == deg.sum()
    assert np.all((lp2.u >= 1) & (lp2.u <= 100))
    mags = np.abs(lp2.val)
    assert mags.min() >= 1e-3 and mags.max() <= 1e3
    # determinism
    lp3 = P.gen_random_lp(m, n, 6, seed=4)
    assert np.array_equal(lp.val, lp3.val) and np.array_equal(lp.c, lp3.c)


def test_bad_arguments_raise():
    lp = P.gen_random_lp(20, 40, 6, seed=1)
    bad = P.LP(lp.m, lp.n, lp.row.copy(), lp.col.copy(), lp.val, lp.c, lp.rL, lp.rU, lp.l, lp.u)
    bad.row[3] = lp.m    # out of range
    with pytest.raises(P.SupremaError) as e:
        P.verify(bad, np.zeros(lp.n), np.zeros(lp.m))
    assert e.value.code == -1 and "out of range" in str(e.value)
    with pytest.raises(P.SupremaError):
        P.gen_transport(1, 5)
    with pytest.raises(ValueError):
        P.verify(lp, np.zeros(lp.n + 1), np.zeros(lp.m))


def test_blend_generator_structure(oracle):
    K, Pp = 7, 5
    lp = P.gen_blend(K, Pp, seed=9)
    assert (lp.m, lp.n, lp.nnz) == (3 * Pp + K, K * Pp, 4 * K * Pp)
    # demand rows are equalities, capacity rows are <=, quality rows are ranged
    assert np.array_equal(lp.rL[:Pp], lp.rU[:Pp])
    assert np.all(np.isinf(lp.rL[Pp:Pp + K])) and np.all(np.isfinite(lp.rU[Pp:Pp + K]))
    q = slice(Pp + K, None)
    assert np.all(lp.rL[q] < lp.rU[q]) and np.allclose(lp.rU[q] / lp.rL[q], 1.1 / 0.9)
    assert np.all((lp.c >= 10) & (lp.c <= 100))
    # the PDLP form adds one slack column per ranged row
    form, meta = P.build_pdlp_form(lp)
    assert meta.n_slack == 2 * Pp and form.n == lp.n + 2 * Pp
    # feasible and bounded: the oracle solves it
    s = oracle.scale(form)
    out = oracle.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 200000, 64, 1e-6)
    assert out.converged
    sizes = P.gen_blend_sizes(512)
    assert len(sizes) == 512 and all(20 <= k <= 400 and 10 <= p <= 200 for k, p in sizes)


@pytest.mark.parametrize("name", ["afiro", "share1b", "tiny_range", "med", "israel"])
def test_extract_solution_bit_identical_to_oracle(oracle, name):
    lp, _ = load_fixture(name)
    form, meta = P.build_pdlp_form(lp)
    rng = np.random.default_rng(11)
    x2, y2 = rng.standard_normal(form.n), rng.standard_normal(form.m)
    x, y = P.extract_solution(meta, x2, y2)
    xo, yo = oracle.extract_solution(lp.n, lp.m, meta.row_perm, meta.row_signs, x2, y2)
    assert np.array_equal(x, xo) and np.array_equal(y, yo)
