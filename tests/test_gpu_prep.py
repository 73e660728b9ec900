"""GPU tests of the device preprocessing: the stable CSR / CSC build
(supr_csr_gpu) against the stable-sort definition of the reference's to_csr
/ to_csc_as_transposed_csr (include/sparse.hpp:31-57), and scale() on the
device (supr_scale_gpu) against the host scale() (src/scaling.cpp:25-84),
which the CPU tests pin to the reference; both bit for bit."""
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


def assert_same_scaling(lp, passes=10, integer=None):
    hs, hDr, hDc = P.scale(lp, passes, integer)
    ds, dDr, dDc = P.scale_gpu(lp, passes, integer)
    for name, h, d in [("val", hs.val, ds.val), ("c", hs.c, ds.c), ("rL", hs.rL, ds.rL),
                       ("rU", hs.rU, ds.rU), ("l", hs.l, ds.l), ("u", hs.u, ds.u),
                       ("Dr", hDr, dDr), ("Dc", hDc, dDc)]:
        if not bits_equal(h, d):
            bad = np.flatnonzero(h.view(np.uint64) != d.view(np.uint64))
            raise AssertionError(f"{lp.name}: {name} differs at {bad.size} entries, first {bad[:5]}: "
                                 f"host {h[bad[:3]]} device {d[bad[:3]]}")


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


def small_cases(golden):
    out = [golden_lp(golden[nm]) for nm in ["afiro", "sc50a", "blend", "tiny_range", "sc105"]]
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
def test_scale_gpu_is_bit_identical_to_host(golden, passes):
    cases = small_cases(golden)
    for lp in cases:
        assert_same_scaling(lp, passes)
    say(f"{passes} Ruiz passes: {len(cases)} LPs bit-identical")


def test_scale_gpu_integer_columns_and_edge_shapes():
    lp = P.gen_random_lp(80, 200, 5, seed=4)
    assert_same_scaling(lp, 10, np.arange(lp.n) % 3 == 0)
    base = P.gen_random_lp(5, 9, 2, seed=3)
    assert_same_scaling(replace(base, row=base.row[:0], col=base.col[:0], val=base.val[:0], name="empty"))
    assert_same_scaling(replace(base, row=base.row[:3], col=base.col[:3], val=base.val[:3], name="sparse"))


def test_scale_gpu_at_full_sizes():
    """The 1M-variable transportation LP and a power-law LP with 7M entries
    (long columns, 6 decades of magnitude)."""
    for lp in (P.gen_transport(1000, 1000, seed=42), P.gen_power_law_lp(100000, 400000, seed=42)):
        _, _, _, secs = P.scale_gpu(lp, timing=True)
        assert_same_scaling(lp)
        say(f"{lp.name}: m={lp.m} n={lp.n} nnz={lp.nnz} bit-identical; device "
            f"layout {secs[0]:.3f}s scale {secs[1]:.4f}s download {secs[2]:.3f}s")


def test_scale_gpu_rejects_bad_input():
    lp = P.gen_random_lp(5, 9, 2, seed=3)
    bad = replace(lp, row=np.where(np.arange(lp.nnz) == 0, 7, lp.row).astype(np.int32))
    with pytest.raises(P.SupremaError):
        P.scale_gpu(bad)
    with pytest.raises(P.SupremaError):
        P.scale_gpu(lp, ruiz_passes=-1)
