"""GPU test of supr_verify_gpu_noalloc, the entry point that allocates
nothing: verify() of an LP and a point already in device memory, its
partial sums in the library's fixed device array.  Checked against
supr_verify_gpu (the existing entry point, bit-identical to the host
verify()) on three LPs in a row -- the 1M-variable transportation LP, then a
smaller LP with equality, ranged, one-sided and free rows and finite and
infinite bounds (every branch of the verifier), then a Netlib fixture -- and
once more on the first.  Tolerances: the sums are formed per block and
added in a fixed order instead of left to right, and the row products A x,
A^T y are formed by several lanes per row instead of left to right.  So
eps_P and eps_D agree to 1e-12 relative plus the 2-norm of the products'
rounding bounds (2 (nnz_row + 2) 2^-52 (|A||x|)_row per row, the same for
A^T y on the columns, divided by the norms' denominators: at a converged
point the violations are differences of nearly equal numbers); the
objective c'x, the dual objective and eps_G to 1e-13 of the sum of their
terms' magnitudes; the largest row and bound violations to their rows'
rounding bounds.  Device memory for the inputs and the output is the test's
own, as a caller's would be: hipMalloc of the HIP runtime the library itself
is linked against (one runtime in the process; torch's bundled runtime is not
loaded)."""
import ctypes

import numpy as np
import pytest

import pdlp_amd as P
from pdlp_amd import _lib
from conftest import load_fixture

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    loaded (the caller's own device memory for the inputs and the output)."""

    def __init__(self):
        P.lib()   # loads the HIP runtime (soname libamdhip64.so.7) with the library
        self.rt = ctypes.CDLL("libamdhip64.so.7")   # the already-loaded object
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        for f in (self.rt.hipMalloc, self.rt.hipMemcpy, self.rt.hipFree, self.rt.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p.value

    def get(self, ptr, count):
        out = np.empty(count, dtype=np.float64)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def hip():
    h = Hip()
    yield h
    h.free_all()


def all_row_kinds_lp(seed=3):
    rng = np.random.default_rng(seed)
    m, n = 40, 600
    dense = rng.random((m, n)) < np.where(np.arange(m)[:, None] < 3, 0.7, 0.05)
    rows, cols = np.nonzero(dense)
    p = rng.permutation(rows.size)
    rows, cols = rows[p].astype(np.int32), cols[p].astype(np.int32)
    vals = rng.standard_normal(rows.size)
    u = np.where(np.arange(n) % 7 == 0, np.inf, rng.uniform(1, 10, n))
    l = np.where(np.arange(n) % 11 == 0, -np.inf, 0.0)
    a = np.zeros(m)
    np.add.at(a, rows, vals * rng.uniform(0, 1, n)[cols])
    kind = np.arange(m) % 5
    lo, hi = a - rng.uniform(0, 2, m), a + rng.uniform(0, 2, m)
    rL = np.where(kind == 0, a, np.where((kind == 1) | (kind == 3), lo, -np.inf))
    rU = np.where(kind == 0, a, np.where((kind == 1) | (kind == 2), hi, np.inf))
    return P.LP(m, n, rows, cols, vals, rng.uniform(-1, 1, n), rL, rU, l, u, 0.5, "all row kinds")


class DeviceLp:
    """The LP's CSR pair and vectors in device memory."""

    def __init__(self, hip, lp):
        ptrA, idxA, permA = P.csr_gpu(lp, transposed=False)
        ptrT, idxT, permT = P.csr_gpu(lp, transposed=True)
        k = [hip.put(np.asarray(a, dtype=dt)) for a, dt in (
            (ptrA, np.int32), (idxA, np.int32), (lp.val[permA], np.float64),
            (ptrT, np.int32), (idxT, np.int32), (lp.val[permT], np.float64),
            (lp.c, np.float64), (lp.rL, np.float64), (lp.rU, np.float64),
            (lp.l, np.float64), (lp.u, np.float64))]
        self.s = _lib.SuprDevLp(lp.m, lp.n, lp.nnz, *k, float(lp.obj_offset))


def verify_noalloc(hip, dlp, x, y):
    xd = hip.put(np.asarray(x, dtype=np.float64))
    yd = hip.put(np.asarray(y, dtype=np.float64))
    out = hip.put(np.full(7, np.nan))
    rc = P.lib().supr_verify_gpu_noalloc(ctypes.byref(dlp.s), xd, yd, out, None)   # the null stream
    assert rc == 0, P.lib().supr_last_error().decode()
    return hip.get(out, 7)


def magnitudes(lp, x, y):
    """Sums of the magnitudes of the terms of c'x and of the dual objective."""
    aty = P.spmv(lp, y, transposed=True)
    z = lp.c - aty
    fin = lambda v: np.where(np.isfinite(v), np.abs(v), 0.0)  # noqa: E731
    return (np.sum(np.abs(lp.c * x)),
            np.sum((fin(lp.rL) + fin(lp.rU)) * np.abs(y)) + np.sum((fin(lp.l) + fin(lp.u)) * np.abs(z)))


def rounding_bounds(lp, x, y):
    """Per row of A and of A^T: 2 (nnz + 2) 2^-52 (|A||x|), the largest
    difference two summation orders of the row product can show."""
    absA = lp.with_values(np.abs(lp.val), lp.c, lp.rL, lp.rU, lp.l, lp.u)
    cr = np.bincount(lp.row, minlength=lp.m)
    cc = np.bincount(lp.col, minlength=lp.n)
    u = 2.0 ** -52
    return (2 * (cr + 2) * u * P.spmv(absA, np.abs(x)),
            2 * (cc + 2) * u * (P.spmv(absA, np.abs(y), transposed=True) + np.abs(lp.c)))


def check_same(lp, x, y, got):
    want = P.verify_gpu(lp, x, y)
    w = np.array([want.eps_P, want.eps_D, want.eps_G, want.objective, want.dual_objective,
                  want.max_row_violation, want.max_bound_violation])
    mcx, mD = magnitudes(lp, x, y)
    br, bc = rounding_bounds(lp, x, y)
    fin = lambda v: np.max(np.abs(v[np.isfinite(v)]), initial=0.0)  # noqa: E731
    den_p = 1.0 + max(fin(lp.rL), fin(lp.rU))
    den_d = 1.0 + fin(lp.c)
    assert abs(got[0] - w[0]) <= 1e-12 * w[0] + np.linalg.norm(br) / den_p, (got[0], w[0])
    assert abs(got[1] - w[1]) <= 1e-12 * w[1] + np.linalg.norm(bc) / den_d, (got[1], w[1])
    assert abs(got[3] - w[3]) <= 1e-13 * (mcx + 1.0), (got[3], w[3])
    tol_D = 1e-13 * (mD + 1.0) + np.sum(bc * (np.where(np.isfinite(lp.l), np.abs(lp.l), 0)
                                              + np.where(np.isfinite(lp.u), np.abs(lp.u), 0)))
    assert abs(got[4] - w[4]) <= tol_D, (got[4], w[4])
    assert abs(got[2] - w[2]) <= (1e-13 * (mcx + 1.0) + tol_D) / (1 + abs(w[3]) + abs(w[4])) * 2 + 1e-15, \
        (got[2], w[2])
    assert abs(got[5] - w[5]) <= np.max(br, initial=0.0) and got[6] == w[6], (got[5:], w[5:])
    return w


def test_noalloc_verify_matches_verify_gpu_on_calls_in_a_row(hip):
    rng = np.random.default_rng(7)
    big = P.gen_transport(1000, 1000, seed=42)
    cases = [big, all_row_kinds_lp(), load_fixture("share1b")[0], big]
    devs = {}
    for k, lp in enumerate(cases):
        if id(lp) not in devs:
            devs[id(lp)] = DeviceLp(hip, lp)
        dlp = devs[id(lp)]
        x = rng.standard_normal(lp.n) * 3
        y = rng.standard_normal(lp.m)
        if k == 0:
            # a solved point too: eps near zero, violations tiny
            s, Dr, Dc = P.scale(lp)
            r = P.solve_pdhg_gpu(lp, s, Dr, Dc, tol=1e-6)
            x, y = r.x, r.y
        got = verify_noalloc(hip, dlp, x, y)
        check_same(lp, x, y, got)
        again = verify_noalloc(hip, dlp, x, y)
        assert np.array_equal(got.view(np.uint64), again.view(np.uint64))   # fixed order: same bits
        print(f"[no-alloc] call {k}: {lp.name} m={lp.m} n={lp.n} eps {got[:3]}", flush=True)


def test_noalloc_verify_rejects_bad_arguments():
    s = _lib.SuprDevLp(-1, 3, 0, *([None] * 11), 0.0)
    rc = P.lib().supr_verify_gpu_noalloc(ctypes.byref(s), None, None, None, None)
    assert rc == _lib.SUPR_EINVAL
