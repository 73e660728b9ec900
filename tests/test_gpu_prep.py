"""GPU tests of the device preprocessing: the stable CSR / CSC build
(supr_csr_gpu) against the stable-sort definition of the reference's to_csr
/ to_csc_as_transposed_csr (include/sparse.hpp:31-57), and scale() on the
device (supr_scale_gpu) against the oracle's scale() (src/scaling.cpp:25-84,
pinned to the reference by the golden SHA-256 digests) and against the
product's host scale(); all bit for bit.  Cases that are golden fixtures
are also checked against the reference's digest directly."""
import hashlib
import sys
from dataclasses import replace

import numpy as np
import pytest

import pdlp_amd as P
from conftest import golden_lp

pytestmark = pytest.mark.gpu


def say(*a):
    print("[prep]", *a, file=sys.stderr, flush=True)


def stable_csr(lp, transposed):
    key, other, rows = (lp.col, lp.row, lp.n) if transposed else (lp.row, lp.col, lp.m)
    perm = np.argsort(key, kind="stable").astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=rows))]).astype(np.int32)
    return ptr, other[perm].astype(np.int32), perm


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def assert_same_scaling(oracle, lp, passes=10, integer=None, golden_sha=None):
    """Device scale() equals the oracle's (the checker) and the host
    product's, bit for bit; golden_sha: the reference's digest of scale()'s
    eight outputs for this LP."""
    ds, dDr, dDc = P.scale_gpu(lp, passes, integer)
    dev = [ds.val, ds.c, ds.rL, ds.rU, ds.l, ds.u, dDr, dDc]
    want = oracle.scale(lp, passes, integer=None if integer is None else np.asarray(integer, np.int8))
    hs, hDr, hDc = P.scale(lp, passes, integer)
    host = [hs.val, hs.c, hs.rL, hs.rU, hs.l, hs.u, hDr, hDc]
    for who, ref in (("oracle", want), ("host", host)):
        for name, h, d in zip(("val", "c", "rL", "rU", "l", "u", "Dr", "Dc"), ref, dev):
            h = np.ascontiguousarray(h, dtype=np.float64)
            if not bits_equal(h, d):
                bad = np.flatnonzero(h.view(np.uint64) != d.view(np.uint64))
                raise AssertionError(f"{lp.name}: {name} differs from the {who} at {bad.size} entries, "
                                     f"first {bad[:5]}: {who} {h[bad[:3]]} device {d[bad[:3]]}")
    if golden_sha is not None:
        assert sha(*dev) == golden_sha, f"{lp.name}: device scale() differs from the reference's digest"


def shuffled(lp, seed):
    """The same LP with its COO entries in a random order (the CSR order
    within a row then differs from row-major)."""
    p = np.random.default_rng(seed).permutation(lp.nnz)
    return replace(lp, row=lp.row[p].copy(), col=lp.col[p].copy(), val=lp.val[p].copy(),
                   name=lp.name + " shuffled")


def dense_rows_lp(m=3, n=900, seed=1):
    """Rows longer than 256 entries (the one-wavefront-per-row sums),
    magnitudes over 6 decades, COO in random order."""
    rng = np.random.default_rng(seed)
    rows, cols = np.nonzero(rng.random((m, n)) < 0.7)
    p = rng.permutation(rows.size)
    base = P.gen_random_lp(m, n, 1, seed=seed)
    vals = rng.standard_normal(rows.size) * 10.0 ** rng.uniform(-3, 3, rows.size)
    return replace(base, row=rows[p].astype(np.int32), col=cols[p].astype(np.int32), val=vals,
                   name="dense rows")


GOLDEN_SCALED = ["afiro", "sc50a", "blend", "tiny_range", "sc105"]


def small_cases(golden):
    out = [golden_lp(golden[nm]) for nm in GOLDEN_SCALED]
    out.append(P.gen_random_lp(600, 2000, 6, seed=8))
    out.append(shuffled(P.gen_random_lp(600, 2000, 6, seed=9), 1))
    out.append(dense_rows_lp())
    out.append(P.gen_transport(30, 40, seed=7))
    out.append(P.build_pdlp_form(P.gen_blend(40, 20, seed=5))[0])
    return out


def test_csr_layout_is_the_stable_sort(golden):
    for lp in small_cases(golden):
        for tr in (False, True):
            got = P.csr_gpu(lp, transposed=tr)
            want = stable_csr(lp, tr)
            for nm, g, w in zip(("ptr", "idx", "perm"), got, want):
                assert np.array_equal(g, w), (lp.name, tr, nm)
    say("CSR / CSC layouts equal the stable sort on", len(small_cases(golden)), "LPs")


def test_csr_layout_edge_shapes():
    base = P.gen_random_lp(5, 9, 2, seed=3)
    empty = replace(base, row=base.row[:0], col=base.col[:0], val=base.val[:0], name="empty")
    one_row = replace(base, row=np.zeros(base.nnz, np.int32), name="one row")
    for lp in (empty, one_row, base):
        for tr in (False, True):
            got, want = P.csr_gpu(lp, transposed=tr), stable_csr(lp, tr)
            assert all(np.array_equal(g, w) for g, w in zip(got, want)), (lp.name, tr)


@pytest.mark.parametrize("passes", [0, 1, 10])
def test_scale_gpu_is_bit_identical_to_oracle(oracle, golden, passes):
    cases = small_cases(golden)
    for k, lp in enumerate(cases):
        gsha = golden[GOLDEN_SCALED[k]]["ref"]["scaled_sha"] if passes == 10 and k < len(GOLDEN_SCALED) else None
        assert_same_scaling(oracle, lp, passes, golden_sha=gsha)
    say(f"{passes} Ruiz passes: {len(cases)} LPs bit-identical to the oracle"
        + (f" ({len(GOLDEN_SCALED)} to the reference's digests)" if passes == 10 else ""))


def test_scale_gpu_golden_generated_cases(oracle, golden):
    """Generated golden cases (random, transportation, power-law, blending):
    the device scale() gives the reference's digest."""
    names = ["random_2000x5000_s42", "transport_60x60_s42", "power_law_1000x4000_s7", "blend_25x40_s103"]
    for nm in names:
        assert_same_scaling(oracle, golden_lp(golden[nm]), golden_sha=golden[nm]["ref"]["scaled_sha"])
    say(f"{len(names)} generated golden LPs: the reference's scale() digest")


def test_scale_gpu_integer_columns_and_edge_shapes(oracle):
    lp = P.gen_random_lp(80, 200, 5, seed=4)
    assert_same_scaling(oracle, lp, 10, np.arange(lp.n) % 3 == 0)
    base = P.gen_random_lp(5, 9, 2, seed=3)
    assert_same_scaling(oracle, replace(base, row=base.row[:0], col=base.col[:0], val=base.val[:0], name="empty"))
    assert_same_scaling(oracle, replace(base, row=base.row[:3], col=base.col[:3], val=base.val[:3], name="sparse"))


def test_scale_gpu_at_full_sizes(oracle):
    """The 1M-variable transportation LP and a power-law LP with 7M entries
    (long columns, 6 decades of magnitude), against the oracle."""
    for lp in (P.gen_transport(1000, 1000, seed=42), P.gen_power_law_lp(100000, 400000, seed=42)):
        _, _, _, secs = P.scale_gpu(lp, timing=True)
        assert_same_scaling(oracle, lp)
        say(f"{lp.name}: m={lp.m} n={lp.n} nnz={lp.nnz} bit-identical; device "
            f"layout {secs[0]:.3f}s scale {secs[1]:.4f}s download {secs[2]:.3f}s")


def test_scale_gpu_rejects_bad_input():
    lp = P.gen_random_lp(5, 9, 2, seed=3)
    bad = replace(lp, row=np.where(np.arange(lp.nnz) == 0, 7, lp.row).astype(np.int32))
    with pytest.raises(P.SupremaError):
        P.scale_gpu(bad)
    with pytest.raises(P.SupremaError):
        P.scale_gpu(lp, ruiz_passes=-1)


# SpMV schedules fixed: autotuned lanes-per-row are picked by timing, and
# another schedule sums in another order.
PIN = dict(spmv_lanes_A=8, spmv_lanes_AT=2)


def same_result(a, b):
    return (a.iterations == b.iterations and a.restarts == b.restarts and a.converged == b.converged
            and bits_equal(a.x, b.x) and bits_equal(a.y, b.y)
            and bits_equal([a.eps_P, a.eps_D, a.eps_G], [b.eps_P, b.eps_D, b.eps_G]))


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64, P.MODE_PLAIN_FP64])
def test_solve_form_gpu_matches_host_scaled_solve(golden, oracle, mode):
    """scale() on the device + solve == host scale() + solve_pdhg_gpu, bit
    for bit (ELL and CSR layouts: uniform columns, ragged rows, power law)."""
    forms = [golden_lp(golden[nm]) for nm in ["afiro", "sc50a", "blend", "tiny_range"]]
    forms += [P.gen_random_lp(60, 150, 6, seed=11), P.gen_transport(30, 40, seed=7),
              P.gen_power_law_lp(2000, 8000, seed=5)]
    for form in forms:
        dev = P.solve_form_gpu(form, tol=1e-6, max_iterations=100000, mode=mode, **PIN)
        host = P.solve_pdhg_gpu(form, *P.scale(form), tol=1e-6, max_iterations=100000, mode=mode,
                                **PIN)
        assert same_result(dev, host), (form.name, dev.iterations, host.iterations)
        rep = oracle.verify(form, dev.x, dev.y)
        assert dev.converged and max(rep[:3]) <= 1e-6, (form.name, rep[:3])
    say(f"mode {mode}: {len(forms)} LPs, device-scaled solve identical to the host-scaled one")


def test_solve_form_gpu_integer_columns_and_stop_before_first_check():
    lp = P.gen_random_lp(80, 200, 5, seed=4)
    flags = np.arange(lp.n) % 4 == 0
    dev = P.solve_form_gpu(lp, integer=flags, tol=1e-6, **PIN)
    host = P.solve_pdhg_gpu(lp, *P.scale(lp, integer=flags), tol=1e-6, **PIN)
    assert same_result(dev, host) and dev.converged
    # no KKT check before the limit: the point is assessed by the host verifier
    # on the unscaled form, which this engine fetches back from the device
    for mode in (P.MODE_MIXED, P.MODE_FAST_FP64):
        dev = P.solve_form_gpu(lp, tol=1e-9, max_iterations=20, mode=mode, **PIN)
        host = P.solve_pdhg_gpu(lp, *P.scale(lp), tol=1e-9, max_iterations=20, mode=mode, **PIN)
        assert not dev.converged and same_result(dev, host), mode


def test_engine_from_form_at_1m_variables():
    """The 1M-variable transportation LP: a resident engine built from the
    unscaled form (layout and scaling on the device) solves to the same bits
    as one built from the host-scaled form, and sets up faster."""
    form = P.gen_transport(1000, 1000, seed=42)
    import time
    t = time.perf_counter()
    scaled, Dr, Dc = P.scale(form)
    t_host_scale = time.perf_counter() - t
    pin = dict(spmv_lanes_A=256, spmv_lanes_AT=1)
    with P.Engine.from_form(form, tol=1e-6, **pin) as dev, \
            P.Engine(form, scaled, Dr, Dc, tol=1e-6, **pin) as host:
        a, b = dev.solve(), host.solve()
        a2 = dev.solve()
    assert a.converged and same_result(a, b) and same_result(a, a2)
    say(f"1M vars: identical ({a.iterations} attempts); setup device-scaled "
        f"{a.info['seconds_setup']:.3f}s vs host scale {t_host_scale:.3f}s + setup "
        f"{b.info['seconds_setup']:.3f}s")


def all_row_kinds_lp(seed=3):
    """Equality, ranged, <=, >= and free rows (each kind about a fifth),
    some rows longer than 256 entries."""
    rng = np.random.default_rng(seed)
    m, n = 40, 600
    dense = rng.random((m, n)) < np.where(np.arange(m)[:, None] < 3, 0.7, 0.05)
    rows, cols = np.nonzero(dense)
    p = rng.permutation(rows.size)
    rows, cols = rows[p].astype(np.int32), cols[p].astype(np.int32)
    vals = rng.standard_normal(rows.size)
    u = rng.uniform(1, 10, n)
    a = np.zeros(m)
    np.add.at(a, rows, vals * rng.uniform(0, 1, n)[cols] * u[cols])   # A x0, x0 inside the box
    kind = np.arange(m) % 5
    lo, hi = a - rng.uniform(0, 2, m), a + rng.uniform(0, 2, m)
    rL = np.where(kind == 0, a, np.where((kind == 1) | (kind == 3), lo, -np.inf))
    rU = np.where(kind == 0, a, np.where((kind == 1) | (kind == 2), hi, np.inf))
    return P.LP(m, n, rows, cols, vals, rng.uniform(-1, 1, n), rL, rU, np.zeros(n), u, 0.5,
                "all row kinds")


def original_lps(golden):
    from conftest import load_fixture
    out = [load_fixture(nm)[0] for nm in ["afiro", "sc50a", "blend", "tiny_range", "share1b"]]
    out += [P.gen_blend(30, 12, seed=3), all_row_kinds_lp()]
    return out


def test_pdlp_form_gpu_is_bit_identical(golden, oracle):
    """build_pdlp_form on the device against the oracle's (and the host
    product's); golden MPS cases also against the reference's digest."""
    for lp in original_lps(golden):
        dims, ri, ci, arrs, perm, signs = oracle.pdlp_form(lp)
        df, dm = P.build_pdlp_form_gpu(lp)
        hf, hm = P.build_pdlp_form(lp)
        want = dict(zip(("row", "col", "val", "c", "rL", "rU", "l", "u"), (ri, ci, *arrs)))
        for nm, w in want.items():
            assert np.asarray(w).tobytes() == np.asarray(getattr(df, nm)).tobytes(), (lp.name, nm)
            assert np.asarray(getattr(hf, nm)).tobytes() == np.asarray(getattr(df, nm)).tobytes(), (lp.name, nm)
        assert perm.tobytes() == dm.row_perm.tobytes() and signs.tobytes() == dm.row_signs.tobytes()
        assert (int(dims[3]), int(dims[0]), int(dims[1])) == (dm.n_slack, df.m, df.n)
        if lp.name in golden and golden[lp.name]["kind"] == "mps":
            assert sha(df.row, df.col, df.val, df.c, df.rL, df.rU, df.l, df.u, dm.row_perm,
                       dm.row_signs) == golden[lp.name]["form_sha"], lp.name
    say("build_pdlp_form on the device: bit-identical to the oracle on", len(original_lps(golden)), "LPs")


def test_verify_gpu_is_bit_identical(golden, oracle):
    """verify() with the products on the device against the oracle's verify
    (and the host product's), all seven outputs bit for bit."""
    rng = np.random.default_rng(5)
    lps = original_lps(golden) + [dense_rows_lp(), P.gen_transport(1000, 1000, seed=42)]
    for lp in lps:
        x, y = rng.standard_normal(lp.n) * 3, rng.standard_normal(lp.m)
        h, d = P.verify(lp, x, y), P.verify_gpu(lp, x, y)
        dv = [d.eps_P, d.eps_D, d.eps_G, d.objective, d.dual_objective, d.max_row_violation,
              d.max_bound_violation]
        assert bits_equal(oracle.verify(lp, x, y), dv), lp.name
        assert bits_equal([h.eps_P, h.eps_D, h.eps_G, h.objective, h.dual_objective,
                           h.max_row_violation, h.max_bound_violation], dv), lp.name


def host_pipeline(lp, tol, mode, **kw):
    """The reference's solve_mps() PDLP stage after presolve, on the host
    helpers and the host-scaled solve: form, scale, solve, extract, verify,
    10x-tighter retries (src/solve.cpp:194-224)."""
    form, meta = P.build_pdlp_form(lp)
    scaled, Dr, Dc = P.scale(form)
    total, inner = 0, tol
    for _ in range(3):
        r = P.solve_pdhg_gpu(form, scaled, Dr, Dc, tol=inner, mode=mode, **kw)
        total += r.iterations
        x, y = P.extract_solution(meta, r.x, r.y)
        rep = P.verify(lp, x, y)
        if not r.converged or max(rep.eps_P, rep.eps_D, rep.eps_G) <= tol:
            break
        inner *= 0.1
    return x, y, rep, total, r.converged


@pytest.mark.parametrize("mode", [P.MODE_MIXED, P.MODE_FAST_FP64])
def test_solve_lp_gpu_matches_host_pipeline(golden, mode):
    """Form, scale, solve, extract, z, verify and retries with every O(nnz)
    step on the device: the same x, y, report and attempt count as the host
    pipeline, bit for bit; z = c - A^T y."""
    for lp in original_lps(golden):
        sol = P.solve_lp_gpu(lp, tol=1e-6, max_iterations=200000, mode=mode, **PIN)
        x, y, rep, total, conv = host_pipeline(lp, 1e-6, mode, max_iterations=200000, **PIN)
        assert sol.iterations == total and sol.converged == conv, (lp.name, sol.iterations, total)
        assert bits_equal(sol.x, x) and bits_equal(sol.y, y), lp.name
        assert bits_equal([sol.report.eps_P, sol.report.eps_D, sol.report.eps_G, sol.report.objective],
                          [rep.eps_P, rep.eps_D, rep.eps_G, rep.objective]), lp.name
        aty = P.spmv(lp, sol.y, transposed=True)
        np.testing.assert_allclose(sol.z, lp.c - aty, rtol=1e-12, atol=1e-12)
        assert sol.status == "optimal", (lp.name, sol.status, sol.report)
    say(f"mode {mode}: solve_lp_gpu identical to the host pipeline on {len(original_lps(golden))} LPs")


def test_solve_lp_gpu_refuses_plain_mode_and_bad_input():
    lp = P.gen_random_lp(20, 50, 4, seed=1)
    with pytest.raises(P.SupremaError):
        P.solve_lp_gpu(lp, mode=P.MODE_PLAIN_FP64)
    bad = replace(lp, col=np.where(np.arange(lp.nnz) == 3, 99, lp.col).astype(np.int32))
    with pytest.raises(P.SupremaError):
        P.solve_lp_gpu(bad)
