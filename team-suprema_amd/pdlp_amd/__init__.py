"""Python host mirror of the MI355X PDLP engine's C ABI.

Mirrors the reference's PDLP interface (include/pdlp.hpp, include/pdlp_gpu.hpp,
include/verify.hpp, include/scaling.hpp): the same names, argument meanings
and result fields (PdlpResult: x, y, iterations, converged, eps_P, eps_D,
eps_G, restarts), over numpy arrays.  Errors from the library raise
SupremaError (a RuntimeError, as the reference's std::runtime_error).
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (MODE_FAST_FP64, MODE_MIXED, MODE_PLAIN_FP64, SupremaError,  # noqa: F401
                   check, lib)

__all__ = [
    "LP", "PdlpResult", "KKTReport", "solve_pdhg_gpu", "spmv", "verify", "scale", "build_pdlp_form",
    "extract_solution", "scale_gpu", "csr_gpu", "solve_form_gpu", "build_pdlp_form_gpu", "verify_gpu", "solve_lp_gpu", "LpSolution", "certified_status", "Engine", "BatchEngine", "Server", "gen_blend", "gen_blend_sizes", "gen_transport", "gen_random_lp", "zipf_degrees", "gen_power_law_lp", "bench_kernels", "bench_copy", "device_count", "read_mps", "write_mps", "MpsProblem",
    "SupremaError", "MODE_PLAIN_FP64", "MODE_FAST_FP64", "MODE_MIXED",
]

_i32 = np.int32
_f64 = np.float64


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


def _ip(a):
    return a.ctypes.data_as(_lib.c_int32_p)


@dataclass
class LP:
    """min c'x s.t. rL <= Ax <= rU, l <= x <= u, A in COO order (the
    reference's RangedLP, include/ranged_lp.hpp:10-25)."""
    m: int
    n: int
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray
    c: np.ndarray
    rL: np.ndarray
    rU: np.ndarray
    l: np.ndarray
    u: np.ndarray
    obj_offset: float = 0.0
    name: str = "lp"

    def __post_init__(self):
        self.row = np.ascontiguousarray(self.row, dtype=_i32)
        self.col = np.ascontiguousarray(self.col, dtype=_i32)
        for k in ("val", "c", "rL", "rU", "l", "u"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=_f64))
        if self.row.shape != self.col.shape or self.row.shape != self.val.shape:
            raise ValueError("COO arrays differ in length")
        if self.c.shape != (self.n,) or self.l.shape != (self.n,) or self.u.shape != (self.n,):
            raise ValueError("column arrays must have length n")
        if self.rL.shape != (self.m,) or self.rU.shape != (self.m,):
            raise ValueError("row arrays must have length m")

    @property
    def nnz(self):
        return int(self.val.shape[0])

    def cstruct(self):
        return _lib.SuprLp(self.m, self.n, self.nnz, _ip(self.row), _ip(self.col), _dp(self.val),
                           _dp(self.c), _dp(self.rL), _dp(self.rU), _dp(self.l), _dp(self.u),
                           float(self.obj_offset))

    def with_values(self, val, c, rL, rU, l, u, name=None):
        return LP(self.m, self.n, self.row, self.col, val, c, rL, rU, l, u, self.obj_offset,
                  name or self.name)


@dataclass
class KKTReport:
    """include/verify.hpp:7-11"""
    eps_P: float
    eps_D: float
    eps_G: float
    objective: float
    dual_objective: float
    max_row_violation: float
    max_bound_violation: float


@dataclass
class PdlpResult:
    """include/pdlp.hpp:18-24, plus the engine's run information."""
    x: np.ndarray
    y: np.ndarray
    iterations: int = 0
    converged: bool = False
    eps_P: float = 0.0
    eps_D: float = 0.0
    eps_G: float = 0.0
    restarts: int = 0
    info: dict = field(default_factory=dict)


def device_count():
    return int(lib().supr_device_count())


def spmv(lp, x, transposed=False, lanes=0, fp32=False):
    """A x (or A^T x) on the device with one SpMV schedule (GpuBackend::spmv,
    gpu/pdlp_gpu_solve.cu:211-222); lanes 0 = the shape rule, -1 = timed autotune."""
    x = np.ascontiguousarray(x, dtype=_f64)
    if x.shape != ((lp.m,) if transposed else (lp.n,)):
        raise ValueError("x has the wrong length")
    y = np.zeros(lp.n if transposed else lp.m)
    s = lp.cstruct()
    check(lib().supr_spmv(ctypes.byref(s), int(bool(transposed)), int(lanes), int(bool(fp32)),
                          _dp(x), _dp(y)))
    return y


def verify(lp, x, y):
    """verify(ranged, x, y) -- src/verify.cpp:18-103."""
    x = np.ascontiguousarray(x, dtype=_f64)
    y = np.ascontiguousarray(y, dtype=_f64)
    if x.shape != (lp.n,) or y.shape != (lp.m,):
        raise ValueError("x must have length n and y length m")
    out = np.zeros(7)
    s = lp.cstruct()
    check(lib().supr_verify(ctypes.byref(s), _dp(x), _dp(y), _dp(out)))
    return KKTReport(*[float(v) for v in out])


def scale(lp, ruiz_passes=10, integer=None):
    """scale() -- src/scaling.cpp:25-84.  Returns (scaled LP, Dr, Dc)."""
    outs = [np.zeros(lp.nnz), np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.m), np.zeros(lp.n),
            np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)]
    integ = None
    if integer is not None:
        integ = np.ascontiguousarray(integer, dtype=np.int8).tobytes()
    s = lp.cstruct()
    check(lib().supr_scale(ctypes.byref(s), integ, int(ruiz_passes), *[_dp(a) for a in outs]))
    val, c, rL, rU, l, u, Dr, Dc = outs
    return lp.with_values(val, c, rL, rU, l, u, lp.name + " (scaled)"), Dr, Dc


def scale_gpu(lp, ruiz_passes=10, integer=None, device=0, timing=False):
    """scale() on the GPU: the same result as scale(), bit for bit.  With
    timing=True also returns (upload + layout, scaling, download) seconds."""
    outs = [np.zeros(lp.nnz), np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.m), np.zeros(lp.n),
            np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)]
    integ = None
    if integer is not None:
        integ = np.ascontiguousarray(integer, dtype=np.int8).tobytes()
    secs = np.zeros(3)
    s = lp.cstruct()
    check(lib().supr_scale_gpu(ctypes.byref(s), integ, int(ruiz_passes), int(device),
                               *[_dp(a) for a in outs], _dp(secs)))
    val, c, rL, rU, l, u, Dr, Dc = outs
    out = (lp.with_values(val, c, rL, rU, l, u, lp.name + " (scaled)"), Dr, Dc)
    return out + (tuple(secs),) if timing else out


def csr_gpu(lp, transposed=False, device=0):
    """CSR of A (or of A^T) built on the GPU: (ptr, idx, perm), perm[p] = the
    COO position of entry p; entries of a row keep COO order."""
    rows = lp.n if transposed else lp.m
    ptr = np.zeros(rows + 1, dtype=_i32)
    idx, perm = np.zeros(lp.nnz, dtype=_i32), np.zeros(lp.nnz, dtype=_i32)
    s = lp.cstruct()
    check(lib().supr_csr_gpu(ctypes.byref(s), int(bool(transposed)), int(device), _ip(ptr),
                             _ip(idx), _ip(perm)))
    return ptr, idx, perm


@dataclass
class PdlpFormMeta:
    """include/pdlp.hpp:11-16"""
    n_orig: int
    m_orig: int
    n_slack: int
    row_perm: np.ndarray
    row_signs: np.ndarray


def build_pdlp_form(lp):
    """build_pdlp_form() -- src/pdlp.cpp:8-82.  Returns (form LP, meta)."""
    s = lp.cstruct()
    dims = np.zeros(4, dtype=_i32)
    check(lib().supr_pdlp_form_sizes(ctypes.byref(s), _ip(dims)))
    m2, n2, nnz2, n_slack = (int(v) for v in dims)
    ri, ci = np.zeros(nnz2, dtype=_i32), np.zeros(nnz2, dtype=_i32)
    v = np.zeros(nnz2)
    c, l, u = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    rL, rU = np.zeros(m2), np.zeros(m2)
    perm, signs = np.zeros(m2, dtype=_i32), np.zeros(m2)
    check(lib().supr_pdlp_form(ctypes.byref(s), _ip(ri), _ip(ci), _dp(v), _dp(c), _dp(rL),
                               _dp(rU), _dp(l), _dp(u), _ip(perm), _dp(signs)))
    form = LP(m2, n2, ri, ci, v, c, rL, rU, l, u, lp.obj_offset, lp.name + " (pdlp-form)")
    return form, PdlpFormMeta(lp.n, lp.m, n_slack, perm, signs)


def build_pdlp_form_gpu(lp, device=0):
    """build_pdlp_form() on the GPU: the same (form LP, meta) as
    build_pdlp_form(), bit for bit."""
    s = lp.cstruct()
    dims = np.zeros(4, dtype=_i32)
    check(lib().supr_pdlp_form_sizes(ctypes.byref(s), _ip(dims)))
    m2, n2, nnz2, n_slack = (int(v) for v in dims)
    ri, ci = np.zeros(nnz2, dtype=_i32), np.zeros(nnz2, dtype=_i32)
    v = np.zeros(nnz2)
    c, l, u = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    rL, rU = np.zeros(m2), np.zeros(m2)
    perm, signs = np.zeros(m2, dtype=_i32), np.zeros(m2)
    got = np.zeros(4, dtype=_i32)
    check(lib().supr_pdlp_form_gpu(ctypes.byref(s), int(device), _ip(ri), _ip(ci), _dp(v), _dp(c),
                                   _dp(rL), _dp(rU), _dp(l), _dp(u), _ip(perm), _dp(signs),
                                   _ip(got)))
    if not np.array_equal(got, dims):
        raise SupremaError(f"device form sizes {got} differ from {dims}")
    form = LP(m2, n2, ri, ci, v, c, rL, rU, l, u, lp.obj_offset, lp.name + " (pdlp-form)")
    return form, PdlpFormMeta(lp.n, lp.m, n_slack, perm, signs)


def verify_gpu(lp, x, y, device=0):
    """verify() with its products on the GPU: the same KKTReport as verify()."""
    x = np.ascontiguousarray(x, dtype=_f64)
    y = np.ascontiguousarray(y, dtype=_f64)
    if x.shape != (lp.n,) or y.shape != (lp.m,):
        raise ValueError("x must have length n and y length m")
    out = np.zeros(7)
    s = lp.cstruct()
    check(lib().supr_verify_gpu(ctypes.byref(s), _dp(x), _dp(y), int(device), _dp(out)))
    return KKTReport(*[float(v) for v in out])


def extract_solution(meta, x2, y2):
    """extract_solution() -- src/pdlp.cpp:84-90."""
    x2 = np.ascontiguousarray(x2, dtype=_f64)
    y2 = np.ascontiguousarray(y2, dtype=_f64)
    x, y = np.zeros(meta.n_orig), np.zeros(meta.m_orig)
    check(lib().supr_extract_solution(meta.n_orig, meta.m_orig, len(meta.row_perm),
                                      _ip(meta.row_perm), _dp(meta.row_signs), _dp(x2), _dp(y2),
                                      _dp(x), _dp(y)))
    return x, y


def _options(mode, max_iterations, check_every, tol, **kw):
    o = _lib.SuprOptions()
    lib().supr_pdlp_default_options(ctypes.byref(o))
    o.mode = int(mode)
    o.max_iterations = int(max_iterations)
    o.check_every = int(check_every)
    o.tol = float(tol)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, type(getattr(o, k))(v))
    return o


def _stop_cb(stop_flag):
    if stop_flag is None:
        return _lib.STOP_FN()
    poll = stop_flag.is_set if hasattr(stop_flag, "is_set") else stop_flag
    return _lib.STOP_FN(lambda _ctx: 1 if poll() else 0)


def _result(info, x, y):
    d = {k: getattr(info, k) for k, _ in _lib.SuprInfo._fields_}
    return PdlpResult(x if info.has_point else np.zeros(0), y if info.has_point else np.zeros(0),
                      int(info.iterations), bool(info.converged), float(info.eps_P),
                      float(info.eps_D), float(info.eps_G), int(info.restarts), d)


def solve_pdhg_gpu(unscaled_form, scaled_form, Dr, Dc, eta=0.99, max_iterations=200000,
                   check_every=64, tol=1e-6, stop_flag=None, mode=MODE_FAST_FP64, **options):
    """solve_pdhg_gpu() -- include/pdlp_gpu.hpp:22-25.

    eta is accepted and ignored, as in the reference (the step size is
    adaptive).  stop_flag: a callable returning True, or an object with
    is_set() (threading.Event); polled by the host loop.  Returns PdlpResult
    with x, y in unscaled PDLP-form coordinates."""
    del eta
    Dr = np.ascontiguousarray(Dr, dtype=_f64)
    Dc = np.ascontiguousarray(Dc, dtype=_f64)
    if Dr.shape != (scaled_form.m,) or Dc.shape != (scaled_form.n,):
        raise ValueError("Dr must have length m and Dc length n")
    o = _options(mode, max_iterations, check_every, tol, **options)
    x = np.zeros(scaled_form.n)
    y = np.zeros(scaled_form.m)
    info = _lib.SuprInfo()
    cb = _stop_cb(stop_flag)
    su, ss = unscaled_form.cstruct(), scaled_form.cstruct()
    check(lib().supr_solve_pdhg_gpu(ctypes.byref(su), ctypes.byref(ss), _dp(Dr), _dp(Dc),
                                    ctypes.byref(o), cb, None, _dp(x), _dp(y),
                                    ctypes.byref(info)))
    return _result(info, x, y)


def solve_form_gpu(form, ruiz_passes=10, integer=None, max_iterations=200000, check_every=64,
                   tol=1e-6, stop_flag=None, mode=MODE_FAST_FP64, **options):
    """scale(form) then solve_pdhg_gpu(form, scaled, Dr, Dc) with the scaling
    on the device (src/solve.cpp:195-207); the same result as
    solve_pdhg_gpu(form, *scale(form)), bit for bit."""
    o = _options(mode, max_iterations, check_every, tol, **options)
    x, y = np.zeros(form.n), np.zeros(form.m)
    info = _lib.SuprInfo()
    cb = _stop_cb(stop_flag)
    sf = form.cstruct()
    check(lib().supr_solve_form_gpu(ctypes.byref(sf), _flags(integer, form.n), int(ruiz_passes),
                                    ctypes.byref(o), cb, None, _dp(x), _dp(y), ctypes.byref(info)))
    return _result(info, x, y)


@dataclass
class LpSolution:
    """The PDLP stage's answer on the original LP (src/solve.cpp:219-231):
    x, y, z = c - A^T y, the verifier's report on the LP, the certified
    status (src/solve.cpp:47-51 with opt_tol = 10 tol), total attempts."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    report: "KKTReport"
    status: str
    iterations: int
    restarts: int
    converged: bool
    info: dict


def certified_status(engine_status, eps, opt_tol):
    """src/solve.cpp:47-51"""
    if engine_status not in ("optimal", "near_optimal"):
        return engine_status
    if eps <= opt_tol:
        return engine_status
    return "near_optimal" if eps <= 1e-4 else "inaccurate"


def solve_lp_gpu(lp, max_iterations=200000, check_every=64, tol=1e-6, stop_flag=None,
                 mode=MODE_FAST_FP64, **options):
    """solve_mps()'s PDLP stage after presolve on the GPU: form, scale, solve,
    extract, z, verify on lp, and the 10x-tighter retries."""
    o = _options(mode, max_iterations, check_every, tol, **options)
    x, y, z, rep = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n), np.zeros(7)
    info = _lib.SuprInfo()
    cb = _stop_cb(stop_flag)
    s = lp.cstruct()
    check(lib().supr_solve_lp_gpu(ctypes.byref(s), ctypes.byref(o), cb, None, _dp(x), _dp(y), _dp(z),
                                  ctypes.byref(info), _dp(rep)))
    report = KKTReport(*[float(v) for v in rep])
    status = certified_status("optimal" if info.converged else "iteration_limit",
                              max(report.eps_P, report.eps_D, report.eps_G), 10.0 * tol)
    d = {k: getattr(info, k) for k, _ in _lib.SuprInfo._fields_}
    return LpSolution(x, y, z, report, status, int(info.iterations), int(info.restarts),
                      bool(info.converged), d)


def _flags(integer, n):
    if integer is None:
        return None
    f = np.ascontiguousarray(integer, dtype=np.int8)
    if f.shape != (n,):
        raise ValueError("integer flags must have length n")
    return f.tobytes()


class Engine:
    """A resident engine (supr_engine_*): the LP is uploaded and tuned once,
    then solve() runs cold-start solves with the data already in HBM.
    Engine.from_form(form) builds it from the unscaled form, with the layout
    and scale() on the device."""

    @classmethod
    def from_form(cls, form, ruiz_passes=10, integer=None, max_iterations=200000,
                  check_every=64, tol=1e-6, mode=MODE_FAST_FP64, **options):
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        self.m, self.n = form.m, form.n
        o = _options(mode, max_iterations, check_every, tol, **options)
        h = ctypes.c_void_p()
        sf = form.cstruct()
        check(lib().supr_engine_create_form(ctypes.byref(sf), _flags(integer, form.n),
                                            int(ruiz_passes), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        return self

    def __init__(self, unscaled_form, scaled_form, Dr, Dc, max_iterations=200000,
                 check_every=64, tol=1e-6, mode=MODE_FAST_FP64, **options):
        Dr = np.ascontiguousarray(Dr, dtype=_f64)
        Dc = np.ascontiguousarray(Dc, dtype=_f64)
        if Dr.shape != (scaled_form.m,) or Dc.shape != (scaled_form.n,):
            raise ValueError("Dr must have length m and Dc length n")
        self.m, self.n = scaled_form.m, scaled_form.n
        o = _options(mode, max_iterations, check_every, tol, **options)
        h = ctypes.c_void_p()
        su, ss = unscaled_form.cstruct(), scaled_form.cstruct()
        check(lib().supr_engine_create(ctypes.byref(su), ctypes.byref(ss), _dp(Dr), _dp(Dc),
                                       ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def solve(self, stop_flag=None):
        if not self._h:
            raise ValueError("engine is closed")
        x, y = np.zeros(self.n), np.zeros(self.m)
        info = _lib.SuprInfo()
        cb = _stop_cb(stop_flag)
        check(lib().supr_engine_solve(self._h, cb, None, _dp(x), _dp(y), ctypes.byref(info)))
        return _result(info, x, y)

    def close(self):
        if self._h:
            lib().supr_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchEngine:
    """Many independent LPs solved together (supr_batch_*).

    forms: list of (unscaled_form, scaled_form, Dr, Dc).  solve(ids=None)
    solves the listed instances (all by default) from the cold start and
    returns (list of PdlpResult, seconds).  tune=True times the lanes-per-row
    rule on a short run at creation (block cap 512); otherwise lanes_bias and
    max_blocks are used."""

    def __init__(self, forms, max_iterations=200000, check_every=64, tol=1e-6, mode=MODE_FAST_FP64,
                 tune=True, lanes_bias=1.0, max_blocks=512, **options):
        self._keep = []
        us, ss, drs, dcs = [], [], [], []
        self.shapes = []
        for U, S, Dr, Dc in forms:
            Dr = np.ascontiguousarray(Dr, dtype=_f64)
            Dc = np.ascontiguousarray(Dc, dtype=_f64)
            if Dr.shape != (S.m,) or Dc.shape != (S.n,):
                raise ValueError("Dr must have length m and Dc length n")
            self._keep += [U, S, Dr, Dc]
            us.append(U.cstruct())
            ss.append(S.cstruct())
            drs.append(_dp(Dr))
            dcs.append(_dp(Dc))
            self.shapes.append((S.m, S.n))
        cnt = len(us)
        if cnt == 0:
            raise ValueError("empty batch")
        self._U = (_lib.SuprLp * cnt)(*us)
        self._S = (_lib.SuprLp * cnt)(*ss)
        self._Dr = (_lib.c_double_p * cnt)(*drs)
        self._Dc = (_lib.c_double_p * cnt)(*dcs)
        o = _options(mode, max_iterations, check_every, tol, **options)
        o.spmv_lanes_A = 0 if tune else 1
        h = ctypes.c_void_p()
        check(lib().supr_batch_create(cnt, self._U, self._S, self._Dr, self._Dc, ctypes.byref(o),
                                      ctypes.byref(h)))
        self._h = h
        if not tune:
            self.configure(lanes_bias, max_blocks)

    @property
    def config(self):
        """(lanes_bias, max_blocks) in use."""
        bias, cap = ctypes.c_double(), ctypes.c_int32()
        check(lib().supr_batch_config(self._h, ctypes.byref(bias), ctypes.byref(cap)))
        return float(bias.value), int(cap.value)

    def configure(self, lanes_bias, max_blocks):
        check(lib().supr_batch_configure(self._h, float(lanes_bias), int(max_blocks)))

    def solve(self, ids=None):
        if ids is None:
            ids = list(range(len(self.shapes)))
        ids_a = np.ascontiguousarray(ids, dtype=_i32)
        k = len(ids_a)
        if k == 0 or ids_a.min() < 0 or ids_a.max() >= len(self.shapes):
            raise ValueError("instance ids must be in [0, count) and not empty")
        xs = [np.zeros(self.shapes[i][1]) for i in ids_a]
        ys = [np.zeros(self.shapes[i][0]) for i in ids_a]
        xp = (_lib.c_double_p * k)(*[_dp(a) for a in xs])
        yp = (_lib.c_double_p * k)(*[_dp(a) for a in ys])
        infos = (_lib.SuprInfo * k)()
        secs = ctypes.c_double()
        check(lib().supr_batch_solve(self._h, _ip(ids_a), k, xp, yp, infos, ctypes.byref(secs)))
        return [_result(infos[r], xs[r], ys[r]) for r in range(k)], float(secs.value)

    def close(self):
        if self._h:
            lib().supr_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching LP server (supr_server_*).

    submit(unscaled_form, scaled_form, Dr, Dc) -> ticket (any thread);
    step() -> (running, queued, finished_total) retires, admits and runs
    replays_per_step graph replays; result(ticket) -> PdlpResult or None.
    Capacities (max_cols, max_rows, max_nnz, max_instances) bound what is
    resident at once; a request that does not fit waits in the queue."""

    def __init__(self, mode=MODE_FAST_FP64, tol=1e-6, max_iterations=200000, check_every=64,
                 max_cols=1 << 24, max_rows=1 << 23, max_nnz=1 << 27, max_instances=256,
                 max_blocks=1024, lanes_bias=1.0, replays_per_step=1, **options):
        so = _lib.SuprServerOptions()
        lib().supr_server_default_options(ctypes.byref(so))
        so.pdlp = _options(mode, max_iterations, check_every, tol, **options)
        so.max_cols, so.max_rows, so.max_nnz = int(max_cols), int(max_rows), int(max_nnz)
        so.max_instances, so.max_blocks = int(max_instances), int(max_blocks)
        so.lanes_bias, so.replays_per_step = float(lanes_bias), int(replays_per_step)
        h = ctypes.c_void_p()
        check(lib().supr_server_create(ctypes.byref(so), ctypes.byref(h)))
        self._h = h

    @property
    def device_bytes(self):
        b = ctypes.c_int64()
        check(lib().supr_server_device_bytes(self._h, ctypes.byref(b)))
        return int(b.value)

    @staticmethod
    def request_bytes(m, n, nnz, mode=MODE_FAST_FP64):
        return int(lib().supr_server_request_bytes(int(m), int(n), int(nnz), int(mode)))

    def submit(self, unscaled_form, scaled_form, Dr, Dc):
        Dr = np.ascontiguousarray(Dr, dtype=_f64)
        Dc = np.ascontiguousarray(Dc, dtype=_f64)
        if Dr.shape != (scaled_form.m,) or Dc.shape != (scaled_form.n,):
            raise ValueError("Dr must have length m and Dc length n")
        U, S = unscaled_form.cstruct(), scaled_form.cstruct()
        t = ctypes.c_int64()
        check(lib().supr_server_submit(self._h, ctypes.byref(U), ctypes.byref(S), _dp(Dr), _dp(Dc),
                                       ctypes.byref(t)))
        return int(t.value)

    def step(self):
        r, q, f = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().supr_server_step(self._h, ctypes.byref(r), ctypes.byref(q), ctypes.byref(f)))
        return int(r.value), int(q.value), int(f.value)

    def result(self, ticket):
        ready, m, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().supr_server_result(self._h, int(ticket), ctypes.byref(m), ctypes.byref(n), None,
                                       None, None, ctypes.byref(ready)))
        if not ready.value:
            return None
        x, y = np.zeros(n.value), np.zeros(m.value)
        info = _lib.SuprInfo()
        check(lib().supr_server_result(self._h, int(ticket), None, None, _dp(x), _dp(y),
                                       ctypes.byref(info), ctypes.byref(ready)))
        return _result(info, x, y)

    def close(self):
        if self._h:
            lib().supr_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class MpsProblem:
    """read_mps() + to_ranged_lp() (src/mps_reader.cpp:54-308,
    include/ranged_lp.hpp:27-62): the LP, its names, the integer flags, the
    QP lower triangle (empty for an LP) and whether the file maximizes."""
    lp: LP
    col_names: list
    row_names: list
    obj_name: str
    integer: np.ndarray
    maximize: bool
    q_row: np.ndarray
    q_col: np.ndarray
    q_val: np.ndarray

    @property
    def name(self):
        return self.lp.name


def _mps_names(h, kind):
    """Every column (kind 1) or row (kind 2) name of an MPS handle, one call."""
    nb = ctypes.c_int64()
    check(lib().supr_mps_names(h, kind, None, 0, ctypes.byref(nb)))
    if nb.value == 0:
        return []
    buf = ctypes.create_string_buffer(nb.value)
    check(lib().supr_mps_names(h, kind, buf, nb.value, ctypes.byref(nb)))
    return buf.raw[:nb.value - 1].decode().split("\0")


def read_mps(path, names=True):
    """Parses an MPS file as the reference's read_mps + to_ranged_lp do.
    names=False skips building the Python name lists (large files)."""
    h = ctypes.c_void_p()
    check(lib().supr_mps_read(os.fsencode(path), ctypes.byref(h)))
    try:
        dims = [ctypes.c_int32() for _ in range(5)]
        check(lib().supr_mps_sizes(h, *[ctypes.byref(d) for d in dims]))
        m, n, nnz, qn, mx = (d.value for d in dims)
        # every array is filled by supr_mps_get
        row, col = np.empty(nnz, dtype=_i32), np.empty(nnz, dtype=_i32)
        val, c, rL, rU, l, u = (np.empty(k) for k in (nnz, n, m, m, n, n))
        integ = ctypes.create_string_buffer(max(1, n))
        qr, qc, qv = np.zeros(qn, dtype=_i32), np.zeros(qn, dtype=_i32), np.zeros(qn)
        off = ctypes.c_double(0.0)
        check(lib().supr_mps_get(h, _ip(row), _ip(col), _dp(val), _dp(c), _dp(rL), _dp(rU), _dp(l),
                                 _dp(u), ctypes.byref(off), integ, _ip(qr), _ip(qc), _dp(qv)))
        nm = lib().supr_mps_name(h, 0, 0).decode()
        cols = _mps_names(h, 1) if names else []
        rows = _mps_names(h, 2) if names else []
        obj = lib().supr_mps_name(h, 3, 0).decode()
    finally:
        lib().supr_mps_free(h)
    lp = LP(m, n, row, col, val, c, rL, rU, l, u, off.value, nm)
    flags = np.frombuffer(integ.raw[:n], dtype=np.int8).copy()
    return MpsProblem(lp, cols, rows, obj, flags, bool(mx), qr, qc, qv)


def write_mps(path, lp, col_names=None, row_names=None, integer=None):
    """Writes lp as a free-format MPS file (supr_mps_write)."""
    def arr(names, k):
        if names is None:
            return None
        if len(names) != k:
            raise ValueError("one name per column / row")
        a = (ctypes.c_char_p * k)()
        a[:] = [str(v).encode() for v in names]
        return a
    s = lp.cstruct()
    check(lib().supr_mps_write(os.fsencode(path), ctypes.byref(s), lp.name.encode(),
                               arr(col_names, lp.n), arr(row_names, lp.m),
                               _flags(integer, lp.n) if integer is not None else None))


def gen_blend(K, P, seed=1):
    """Blending LP (K components, P products, 2 quality attributes); a ranged
    LP: pass it through build_pdlp_form before solving.  See suprema_pdlp.h."""
    n, m, nnz = K * P, 3 * P + K, 4 * K * P
    ri, ci = np.zeros(nnz, dtype=_i32), np.zeros(nnz, dtype=_i32)
    v = np.zeros(nnz)
    c, rL, rU, l, u = np.zeros(n), np.zeros(m), np.zeros(m), np.zeros(n), np.zeros(n)
    check(lib().supr_gen_blend(int(K), int(P), int(seed), _ip(ri), _ip(ci), _dp(v), _dp(c),
                               _dp(rL), _dp(rU), _dp(l), _dp(u)))
    return LP(m, n, ri, ci, v, c, rL, rU, l, u, 0.0, f"blend_{K}x{P}")


def gen_blend_sizes(count, seed=2026, K=(20, 400), P=(10, 200)):
    """Per-instance (K, P), each log-uniform in its range."""
    Ks, Ps = np.zeros(count, dtype=_i32), np.zeros(count, dtype=_i32)
    check(lib().supr_gen_blend_sizes(int(count), int(seed), int(K[0]), int(K[1]), int(P[0]),
                                     int(P[1]), _ip(Ks), _ip(Ps)))
    return list(zip(Ks.tolist(), Ps.tolist()))


def bench_kernels(unscaled_form, scaled_form, Dr, Dc, reps=50, **options):
    """Per-launch time (ms) of the primal-half and dual-half iteration kernels
    (FP32 and FP64) at this LP's shapes; see supr_bench_kernels."""
    Dr = np.ascontiguousarray(Dr, dtype=_f64)
    Dc = np.ascontiguousarray(Dc, dtype=_f64)
    o = _options(MODE_MIXED, 200000, 64, 1e-6, **options)
    out = np.zeros(9)
    su, ss = unscaled_form.cstruct(), scaled_form.cstruct()
    check(lib().supr_bench_kernels(ctypes.byref(su), ctypes.byref(ss), _dp(Dr), _dp(Dc),
                                   ctypes.byref(o), int(reps), _dp(out)))
    keys = ["primal_ms_fp32", "dual_ms_fp32", "primal_ms_fp64", "dual_ms_fp64", "lanes_A",
            "lanes_AT", "ell_A", "ell_AT", "rebuilt_AT_fp64"]
    return dict(zip(keys, (float(v) for v in out)))


def release_cached_memory():
    """Frees the device memory the library keeps for reuse between calls
    (see supr_release_cached_memory)."""
    lib().supr_release_cached_memory()


def bench_copy(nbytes=1 << 30, reps=20, device=0):
    """ms per launch of a plain device copy of nbytes (supr_bench_copy)."""
    ms = ctypes.c_double()
    check(lib().supr_bench_copy(int(nbytes), int(reps), int(device), ctypes.byref(ms)))
    return float(ms.value)


def gen_transport(S, D, seed=42, cost_max=100):
    """Balanced transportation LP (tools/gen_lp.cpp family); see suprema_pdlp.h."""
    n, m = S * D, S + D
    nnz = 2 * n
    ri, ci = np.zeros(nnz, dtype=_i32), np.zeros(nnz, dtype=_i32)
    v = np.zeros(nnz)
    c, rL, rU, l, u = np.zeros(n), np.zeros(m), np.zeros(m), np.zeros(n), np.zeros(n)
    check(lib().supr_gen_transport(int(S), int(D), int(seed), int(cost_max), _ip(ri), _ip(ci),
                                   _dp(v), _dp(c), _dp(rL), _dp(rU), _dp(l), _dp(u)))
    return LP(m, n, ri, ci, v, c, rL, rU, l, u, 0.0, f"transport_{S}x{D}")


def gen_random_lp(m, n, col_nnz=6, seed=42, x0_hi=5.0, u_hi=10.0, mag_decades=0, col_deg=None):
    """Random feasible, bounded equality-form LP; see suprema_pdlp.h."""
    deg = None
    if col_deg is not None:
        deg = np.ascontiguousarray(col_deg, dtype=_i32)
        if deg.shape != (n,):
            raise ValueError("col_deg must have length n")
        nnz = int(deg.sum())
    else:
        nnz = int(n) * int(col_nnz)
    ri, ci = np.zeros(nnz, dtype=_i32), np.zeros(nnz, dtype=_i32)
    v = np.zeros(nnz)
    c, rL, rU, l, u = np.zeros(n), np.zeros(m), np.zeros(m), np.zeros(n), np.zeros(n)
    check(lib().supr_gen_random_lp(int(m), int(n), int(col_nnz), _ip(deg) if deg is not None else None,
                                   int(seed), float(x0_hi), float(u_hi), int(mag_decades),
                                   _ip(ri), _ip(ci), _dp(v), _dp(c), _dp(rL), _dp(rU), _dp(l),
                                   _dp(u)))
    return LP(m, n, ri, ci, v, c, rL, rU, l, u, 0.0, f"random_{m}x{n}")


def zipf_degrees(n, s=1.5, k_min=2, k_max=200, seed=42):
    """Column degrees with P(k) proportional to k**-s on k_min..k_max (the
    power law truncated to the range), drawn by inverse CDF from a seeded
    numpy PCG64 stream.  For s = 1.5 on [2, 200] the mean is about 17.6."""
    k = np.arange(k_min, k_max + 1, dtype=np.float64)
    cdf = np.cumsum(k ** -s)
    cdf /= cdf[-1]
    u = np.random.default_rng(seed).random(n)
    return (np.searchsorted(cdf, u, side="right") + k_min).astype(_i32)


def gen_power_law_lp(m=500000, n=2000000, seed=42):
    """The large random LP: power-law column degrees (zipf_degrees), values
    of log-uniform magnitude in [1e-3, 1e3] with random sign, box bounds
    [0, U[1, 100]], feasible and bounded as gen_random_lp."""
    deg = np.minimum(zipf_degrees(n, seed=seed), m)
    lp = gen_random_lp(m, n, 0, seed=seed, u_hi=-100.0, mag_decades=3, col_deg=deg)
    lp.name = f"power_law_{m}x{n}"
    return lp
