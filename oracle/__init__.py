"""TEST INFRASTRUCTURE ONLY -- ctypes wrappers of the FP64 oracle.

`Checker("oracle")` wraps liboracle.so (the C restatement, pdlp_oracle.c);
`Checker("ref")` wraps _ref/libsovref.so (the reference's own CPU sources,
built here when the reference tree is present).  Both expose the same
functions over the product's `pdlp_amd.LP` objects (any object with the same
fields works).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg may use this package, and only to check results against.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"oracle": os.path.join(HERE, "liboracle.so"),
         "ref": os.path.join(HERE, "_ref", "libsovref.so")}
PREFIX = {"oracle": "orc_", "ref": "ref_"}

_ip = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def available(kind):
    return os.path.exists(PATHS[kind])


class PdlpOut:
    def __init__(self, x, y, iterations, restarts, converged, has_point, eps):
        self.x, self.y = x, y
        self.iterations, self.restarts = iterations, restarts
        self.converged, self.has_point = bool(converged), bool(has_point)
        self.eps_P, self.eps_D, self.eps_G = (float(e) for e in eps)


class Checker:
    def __init__(self, kind="oracle"):
        if not available(kind):
            raise OSError(f"{PATHS[kind]} is not built")
        self.kind = kind
        self.lib = ctypes.CDLL(PATHS[kind])
        p = PREFIX[kind]
        self._scale = getattr(self.lib, p + "scale")
        self._verify = getattr(self.lib, p + "verify")
        self._form_sizes = getattr(self.lib, p + "pdlp_form_sizes")
        self._form = getattr(self.lib, p + "pdlp_form")
        self._pdlp = getattr(self.lib, p + "pdlp")
        for f in (self._scale, self._verify, self._form_sizes, self._form, self._pdlp):
            f.restype = ctypes.c_int
        self._pdlp.argtypes = ([ctypes.c_int] * 3 + [_ip, _ip] + [_dp] * 6 + [ctypes.c_double]
                               + [_dp] * 8 + [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_void_p, _dp, _dp, _ip, _dp])
        self._verify.argtypes = ([ctypes.c_int] * 3 + [_ip, _ip] + [_dp] * 6
                                 + [ctypes.c_double, _dp, _dp, _dp])
        self._scale.argtypes = ([ctypes.c_int] * 3 + [_ip, _ip] + [_dp] * 6
                                + ([ctypes.c_char_p] if kind == "oracle" else [])
                                + [ctypes.c_int] + [_dp] * 8)
        self._extract = getattr(self.lib, p + "extract_solution")
        self._extract.restype = ctypes.c_int
        self._extract.argtypes = [ctypes.c_int] * 3 + [_ip] + [_dp] * 5
        if kind == "oracle":
            self._spmv = self.lib.orc_csr_spmv
            self._spmv.restype = ctypes.c_int
            self._spmv.argtypes = [ctypes.c_int] * 3 + [_ip, _ip, _dp, ctypes.c_int, _dp, _dp]
            self._csr = self.lib.orc_csr_arrays
            self._csr.restype = ctypes.c_int
            self._csr.argtypes = [ctypes.c_int] * 3 + [_ip, _ip, _dp, _ip, _ip, _dp]
            self._step_rule = self.lib.orc_step_rule
            self._step_rule.restype = ctypes.c_int
            self._step_rule.argtypes = ([ctypes.c_double, ctypes.c_double, ctypes.c_int]
                                        + [ctypes.c_double] * 3 + [_dp, _dp])
            self._check_rule = self.lib.orc_check_rule
            self._check_rule.restype = ctypes.c_int
            self._check_rule.argtypes = ([_dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, _dp, _dp, _ip, _ip, _ip, _dp])
            self._primal_weight = self.lib.orc_primal_weight
            self._primal_weight.restype = ctypes.c_double
            self._primal_weight.argtypes = [ctypes.c_double] * 3

    @staticmethod
    def _lpargs(lp):
        return [lp.m, lp.n, lp.nnz, _i(lp.row), _i(lp.col), _d(lp.val), _d(lp.c), _d(lp.rL),
                _d(lp.rU), _d(lp.l), _d(lp.u)]

    def verify(self, lp, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(7)
        rc = self._verify(*self._lpargs(lp), float(lp.obj_offset), _d(x), _d(y), _d(out))
        assert rc == 0
        return out

    def scale(self, lp, ruiz_passes=10, integer=None):
        outs = [np.zeros(lp.nnz), np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.m), np.zeros(lp.n),
                np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n)]
        if integer is not None and self.kind != "oracle":
            raise ValueError("integer columns: oracle only")
        integ = None if integer is None else np.ascontiguousarray(integer, dtype=np.int8).tobytes()
        extra = [integ] if self.kind == "oracle" else []
        rc = self._scale(*self._lpargs(lp), *extra, int(ruiz_passes), *[_d(a) for a in outs])
        assert rc == 0
        return outs   # val_s, c_s, rL_s, rU_s, l_s, u_s, Dr, Dc

    def pdlp_form(self, lp):
        dims = np.zeros(4, dtype=np.int32)
        if self.kind == "oracle":
            rc = self._form_sizes(lp.m, lp.n, lp.nnz, _i(lp.row), _d(lp.rL), _d(lp.rU), _i(dims))
        else:
            rc = self._form_sizes(*self._lpargs(lp), _i(dims))
        assert rc == 0
        m2, n2, nnz2, _ = (int(v) for v in dims)
        ri, ci = np.zeros(nnz2, dtype=np.int32), np.zeros(nnz2, dtype=np.int32)
        arrs = [np.zeros(nnz2), np.zeros(n2), np.zeros(m2), np.zeros(m2), np.zeros(n2), np.zeros(n2)]
        perm, signs = np.zeros(m2, dtype=np.int32), np.zeros(m2)
        rc = self._form(*self._lpargs(lp), _i(ri), _i(ci), *[_d(a) for a in arrs], _i(perm), _d(signs))
        assert rc == 0
        return dims, ri, ci, arrs, perm, signs

    def extract_solution(self, n_orig, m_orig, row_perm, row_signs, x2, y2):
        """extract_solution(meta, x2, y2): x = the first n_orig entries of x2,
        y[row_perm[p]] = row_signs[p] * y2[p] (y = 0 elsewhere)."""
        row_perm = np.ascontiguousarray(row_perm, dtype=np.int32)
        row_signs = np.ascontiguousarray(row_signs, dtype=np.float64)
        x2 = np.ascontiguousarray(x2, dtype=np.float64)
        y2 = np.ascontiguousarray(y2, dtype=np.float64)
        x, y = np.zeros(n_orig), np.zeros(m_orig)
        rc = self._extract(int(n_orig), int(m_orig), len(row_perm), _i(row_perm), _d(row_signs),
                           _d(x2), _d(y2), _d(x), _d(y))
        assert rc == 0
        return x, y

    def pdlp(self, unscaled, scaled, Dr, Dc, max_iterations=200000, check_every=64, tol=1e-6):
        n, m = scaled.n, scaled.m
        x, y = np.zeros(n), np.zeros(m)
        info_i = np.zeros(4, dtype=np.int32)
        info_d = np.zeros(3)
        Dr = np.ascontiguousarray(Dr, dtype=np.float64)
        Dc = np.ascontiguousarray(Dc, dtype=np.float64)
        rc = self._pdlp(m, n, scaled.nnz, _i(scaled.row), _i(scaled.col),
                        _d(unscaled.val), _d(unscaled.c), _d(unscaled.rL), _d(unscaled.rU),
                        _d(unscaled.l), _d(unscaled.u), float(unscaled.obj_offset),
                        _d(scaled.val), _d(scaled.c), _d(scaled.rL), _d(scaled.rU),
                        _d(scaled.l), _d(scaled.u), _d(Dr), _d(Dc), int(max_iterations),
                        int(check_every), float(tol), None, _d(x), _d(y), _i(info_i), _d(info_d))
        assert rc == 0
        return PdlpOut(x, y, int(info_i[0]), int(info_i[1]), info_i[2], info_i[3], info_d)

    # The scalar rules orc_pdlp takes its decisions through (oracle only).
    def step_rule(self, eta, omega, accepted, dx2, dy2, inter):
        """include/pdlp_algo.hpp:113-131 -> (accepted?, eta_bar, eta_next)."""
        eb, en = ctypes.c_double(0.0), ctypes.c_double(0.0)
        acc = self._step_rule(float(eta), float(omega), int(accepted), float(dx2), float(dy2),
                              float(inter), ctypes.byref(eb), ctypes.byref(en))
        return bool(acc), eb.value, en.value

    def check_rule(self, rc, ra, tol, since_restart, accepted, merit_last_restart, best_merit,
                   merit_prev_cand):
        """include/pdlp_algo.hpp:94-100, 135-163 from the eps of the current
        point (rc) and the average (ra).  Returns a dict: converged,
        best_which, use_avg, restart, cand_merit, best_merit, merit_prev_cand."""
        rc = np.ascontiguousarray(rc, dtype=np.float64)
        ra = np.ascontiguousarray(ra, dtype=np.float64)
        bm, mp, cm = (ctypes.c_double(float(best_merit)), ctypes.c_double(float(merit_prev_cand)),
                      ctypes.c_double(0.0))
        bw, ua, rs = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        done = self._check_rule(_d(rc), _d(ra), float(tol), int(since_restart), int(accepted),
                                float(merit_last_restart), ctypes.byref(bm), ctypes.byref(mp),
                                ctypes.byref(bw), ctypes.byref(ua), ctypes.byref(rs), ctypes.byref(cm))
        return dict(converged=bool(done), best_which=bw.value, use_avg=bool(ua.value),
                    restart=bool(rs.value), cand_merit=cm.value, best_merit=bm.value,
                    merit_prev_cand=mp.value)

    def primal_weight(self, omega, ddx, ddy):
        """include/pdlp_algo.hpp:165-167 (ddx, ddy: restart distances)."""
        return self._primal_weight(float(omega), float(ddx), float(ddy))

    def spmv(self, lp, x, transposed=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(lp.n if transposed else lp.m)
        rc = self._spmv(lp.m, lp.n, lp.nnz, _i(lp.row), _i(lp.col), _d(lp.val), int(transposed),
                        _d(x), _d(out))
        assert rc == 0
        return out

    def csr(self, lp, transposed=False):
        rows, cols = (lp.n, lp.m) if transposed else (lp.m, lp.n)
        r, c = (lp.col, lp.row) if transposed else (lp.row, lp.col)
        ptr = np.zeros(rows + 1, dtype=np.int32)
        idx = np.zeros(lp.nnz, dtype=np.int32)
        val = np.zeros(lp.nnz)
        rc = self._csr(rows, cols, lp.nnz, _i(r), _i(c), _d(lp.val), _i(ptr), _i(idx), _d(val))
        assert rc == 0
        return ptr, idx, val


def read_mps_ref(path, full=False):
    """read_mps + to_ranged_lp of the reference (only where _ref is built).
    Returns a dict of arrays; full=True adds integer flags, the QP triangle
    and the names."""
    lib = ctypes.CDLL(PATHS["ref"])
    dims = np.zeros(3, dtype=np.int32)
    lib.ref_mps_read.restype = ctypes.c_int
    rc = lib.ref_mps_read(path.encode(), _i(dims))
    if rc != 0:
        raise RuntimeError(f"reference read_mps failed on {path}")
    m, n, nnz = (int(v) for v in dims)
    ri, ci = np.zeros(nnz, dtype=np.int32), np.zeros(nnz, dtype=np.int32)
    v, c, rL, rU, l, u = (np.zeros(k) for k in (nnz, n, m, m, n, n))
    off = ctypes.c_double(0.0)
    sense = lib.ref_mps_arrays(_i(ri), _i(ci), _d(v), _d(c), _d(rL), _d(rU), _d(l), _d(u),
                               ctypes.byref(off))
    out = dict(m=m, n=n, row=ri, col=ci, val=v, c=c, rL=rL, rU=rU, l=l, u=u,
               obj_offset=off.value, maximize=bool(sense))
    if full:
        lib.ref_mps_q_nnz.restype = ctypes.c_int
        qn = int(lib.ref_mps_q_nnz())
        integ = ctypes.create_string_buffer(max(1, n))
        qr, qc, qv = np.zeros(qn, dtype=np.int32), np.zeros(qn, dtype=np.int32), np.zeros(qn)
        lib.ref_mps_extra(integ, _i(qr), _i(qc), _d(qv))
        lib.ref_mps_name.restype = ctypes.c_char_p
        out.update(integer=np.frombuffer(integ.raw[:n], dtype=np.int8).copy(), q_row=qr, q_col=qc,
                   q_val=qv, name=lib.ref_mps_name(0, 0).decode(),
                   col_names=[lib.ref_mps_name(1, j).decode() for j in range(n)],
                   row_names=[lib.ref_mps_name(2, i).decode() for i in range(m)])
    return out


def read_mps_ref_error(path):
    """The reference read_mps's exception text on path ('' when it parses)."""
    lib = ctypes.CDLL(PATHS["ref"])
    lib.ref_mps_error.restype = ctypes.c_char_p
    return lib.ref_mps_error(path.encode()).decode()
