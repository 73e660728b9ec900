"""ctypes binding of libsuprema_pdlp.so (include/suprema_pdlp.h).

The shared library is built in-tree (team-suprema_amd/libsuprema_pdlp.so);
loading fails loudly when it is missing -- there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libsuprema_pdlp.so")

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_double_p = ctypes.POINTER(ctypes.c_double)

SUPR_OK = 0
SUPR_EINVAL = -1
SUPR_EDEVICE = -2
SUPR_ENOMEM = -3
SUPR_ENODEV = -4
SUPR_EINTERNAL = -5

MODE_PLAIN_FP64 = 0
MODE_FAST_FP64 = 1
MODE_MIXED = 2


class SuprLp(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("nnz", ctypes.c_int32),
        ("row_idx", c_int32_p), ("col_idx", c_int32_p), ("val", c_double_p),
        ("c", c_double_p), ("rL", c_double_p), ("rU", c_double_p),
        ("l", c_double_p), ("u", c_double_p), ("obj_offset", ctypes.c_double),
    ]


class SuprOptions(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
        ("check_every", ctypes.c_int32), ("tol", ctypes.c_double),
        ("switch_tol", ctypes.c_double), ("stall_checks", ctypes.c_int32),
        ("graph_attempts", ctypes.c_int32), ("spmv_lanes_A", ctypes.c_int32),
        ("spmv_lanes_AT", ctypes.c_int32), ("device", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
    ]


class SuprInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("restarts", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("has_point", ctypes.c_int32),
        ("eps_P", ctypes.c_double), ("eps_D", ctypes.c_double), ("eps_G", ctypes.c_double),
        ("objective", ctypes.c_double), ("iterations_fp32", ctypes.c_int32),
        ("checks", ctypes.c_int32), ("lanes_A", ctypes.c_int32), ("lanes_AT", ctypes.c_int32),
        ("seconds_setup", ctypes.c_double), ("seconds_solve", ctypes.c_double),
        ("seconds_total", ctypes.c_double),
    ]


STOP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)

class SuprDevLp(ctypes.Structure):
    """supr_dev_lp: an LP in device memory (every pointer a device address)."""
    _fields_ = [
        ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("nnz", ctypes.c_int32),
        ("A_ptr", ctypes.c_void_p), ("A_idx", ctypes.c_void_p), ("A_val", ctypes.c_void_p),
        ("AT_ptr", ctypes.c_void_p), ("AT_idx", ctypes.c_void_p), ("AT_val", ctypes.c_void_p),
        ("c", ctypes.c_void_p), ("rL", ctypes.c_void_p), ("rU", ctypes.c_void_p),
        ("l", ctypes.c_void_p), ("u", ctypes.c_void_p), ("obj_offset", ctypes.c_double),
    ]


# (name, restype, argtypes) of every entry point declared in suprema_pdlp.h
class SuprServerOptions(ctypes.Structure):
    _fields_ = [
        ("pdlp", SuprOptions), ("max_cols", ctypes.c_int64), ("max_rows", ctypes.c_int64),
        ("max_nnz", ctypes.c_int64), ("max_instances", ctypes.c_int32),
        ("max_blocks", ctypes.c_int32), ("lanes_bias", ctypes.c_double),
        ("replays_per_step", ctypes.c_int32),
    ]


_LP = ctypes.POINTER(SuprLp)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = [
    ("supr_last_error", ctypes.c_char_p, []),
    ("supr_version", ctypes.c_char_p, []),
    ("supr_device_count", ctypes.c_int, []),
    ("supr_pdlp_default_options", None, [ctypes.POINTER(SuprOptions)]),
    ("supr_solve_pdhg_gpu", ctypes.c_int,
     [_LP, _LP, c_double_p, c_double_p, ctypes.POINTER(SuprOptions), STOP_FN, ctypes.c_void_p,
      c_double_p, c_double_p, ctypes.POINTER(SuprInfo)]),
    ("supr_engine_create", ctypes.c_int,
     [_LP, _LP, c_double_p, c_double_p, ctypes.POINTER(SuprOptions),
      ctypes.POINTER(ctypes.c_void_p)]),
    ("supr_engine_solve", ctypes.c_int,
     [ctypes.c_void_p, STOP_FN, ctypes.c_void_p, c_double_p, c_double_p, ctypes.POINTER(SuprInfo)]),
    ("supr_engine_destroy", None, [ctypes.c_void_p]),
    ("supr_bench_kernels", ctypes.c_int,
     [_LP, _LP, c_double_p, c_double_p, ctypes.POINTER(SuprOptions), ctypes.c_int32, c_double_p]),
    ("supr_spmv", ctypes.c_int,
     [_LP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_double_p, c_double_p]),
    ("supr_verify", ctypes.c_int, [_LP, c_double_p, c_double_p, c_double_p]),
    ("supr_scale", ctypes.c_int,
     [_LP, ctypes.c_char_p, ctypes.c_int32] + [c_double_p] * 8),
    ("supr_scale_gpu", ctypes.c_int,
     [_LP, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32] + [c_double_p] * 9),
    ("supr_solve_form_gpu", ctypes.c_int,
     [_LP, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(SuprOptions), STOP_FN, ctypes.c_void_p,
      c_double_p, c_double_p, ctypes.POINTER(SuprInfo)]),
    ("supr_solve_lp_gpu", ctypes.c_int,
     [_LP, ctypes.POINTER(SuprOptions), STOP_FN, ctypes.c_void_p, c_double_p, c_double_p,
      c_double_p, ctypes.POINTER(SuprInfo), c_double_p]),
    ("supr_engine_create_form", ctypes.c_int,
     [_LP, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(SuprOptions),
      ctypes.POINTER(ctypes.c_void_p)]),
    ("supr_pdlp_form_gpu", ctypes.c_int,
     [_LP, ctypes.c_int32, c_int32_p, c_int32_p] + [c_double_p] * 6 + [c_int32_p, c_double_p,
                                                                       c_int32_p]),
    ("supr_verify_gpu", ctypes.c_int, [_LP, c_double_p, c_double_p, ctypes.c_int32, c_double_p]),
    ("supr_verify_gpu_noalloc", ctypes.c_int,
     [ctypes.POINTER(SuprDevLp), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("supr_csr_gpu", ctypes.c_int,
     [_LP, ctypes.c_int32, ctypes.c_int32, c_int32_p, c_int32_p, c_int32_p]),
    ("supr_pdlp_form_sizes", ctypes.c_int, [_LP, c_int32_p]),
    ("supr_pdlp_form", ctypes.c_int,
     [_LP, c_int32_p, c_int32_p] + [c_double_p] * 6 + [c_int32_p, c_double_p]),
    ("supr_extract_solution", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_int32_p, c_double_p, c_double_p,
      c_double_p, c_double_p, c_double_p]),
    ("supr_gen_transport", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, c_int32_p, c_int32_p]
     + [c_double_p] * 6),
    ("supr_batch_create", ctypes.c_int,
     [ctypes.c_int32, _LP, _LP, ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p),
      ctypes.POINTER(SuprOptions), ctypes.POINTER(ctypes.c_void_p)]),
    ("supr_batch_configure", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    ("supr_batch_config", ctypes.c_int,
     [ctypes.c_void_p, c_double_p, c_int32_p]),
    ("supr_batch_solve", ctypes.c_int,
     [ctypes.c_void_p, c_int32_p, ctypes.c_int32, ctypes.POINTER(c_double_p),
      ctypes.POINTER(c_double_p), ctypes.POINTER(SuprInfo), c_double_p]),
    ("supr_batch_destroy", None, [ctypes.c_void_p]),
    ("supr_server_default_options", None, [ctypes.POINTER(SuprServerOptions)]),
    ("supr_server_request_bytes", ctypes.c_int64,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("supr_server_create", ctypes.c_int,
     [ctypes.POINTER(SuprServerOptions), ctypes.POINTER(ctypes.c_void_p)]),
    ("supr_server_device_bytes", ctypes.c_int, [ctypes.c_void_p, c_int64_p]),
    ("supr_server_submit", ctypes.c_int,
     [ctypes.c_void_p, _LP, _LP, c_double_p, c_double_p, c_int64_p]),
    ("supr_server_step", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p, c_int32_p]),
    ("supr_server_result", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, c_int32_p, c_int32_p, c_double_p, c_double_p,
      ctypes.POINTER(SuprInfo), c_int32_p]),
    ("supr_server_destroy", None, [ctypes.c_void_p]),
    ("supr_bench_copy", ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, c_double_p]),
    ("supr_release_cached_memory", None, []),
    ("supr_gen_blend", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, c_int32_p, c_int32_p] + [c_double_p] * 6),
    ("supr_gen_blend_sizes", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
      ctypes.c_int32, c_int32_p, c_int32_p]),
    ("supr_mps_read", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("supr_mps_sizes", ctypes.c_int, [ctypes.c_void_p] + [c_int32_p] * 5),
    ("supr_mps_get", ctypes.c_int,
     [ctypes.c_void_p, c_int32_p, c_int32_p] + [c_double_p] * 7 + [ctypes.c_char_p, c_int32_p,
                                                                    c_int32_p, c_double_p]),
    ("supr_mps_name", ctypes.c_char_p, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("supr_mps_names", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_int64)]),
    ("supr_mps_free", None, [ctypes.c_void_p]),
    ("supr_mps_write", ctypes.c_int,
     [ctypes.c_char_p, _LP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
      ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p]),
    ("supr_gen_random_lp", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_int32_p, ctypes.c_uint64,
      ctypes.c_double, ctypes.c_double, ctypes.c_int32, c_int32_p, c_int32_p] + [c_double_p] * 6),
]

_lib = None


class SupremaError(RuntimeError):
    """A nonzero status from the C ABI (the reference raises std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(f"suprema error {code}: {msg}")
        self.code = code


def lib():
    """The loaded library (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    if rc != SUPR_OK:
        raise SupremaError(rc, lib().supr_last_error().decode(errors="replace"))
