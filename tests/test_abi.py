"""The C ABI library loads and exports every entry point include/*.h declares
(no device compute here)."""
import ctypes
import os
import re

import pytest

import pdlp_amd as P
from pdlp_amd import _lib
from conftest import ROOT, gpu_available


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(supr_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert "supr_solve_pdhg_gpu" in names and len(names) >= 12
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert names == bound, "the Python binding and the header disagree"


def test_version_and_device_count():
    assert P.lib().supr_version().decode().startswith("suprema-pdlp")
    assert P.device_count() >= 0


def test_default_options_are_the_reference_defaults():
    o = _lib.SuprOptions()
    P.lib().supr_pdlp_default_options(ctypes.byref(o))
    # the reference's defaults (include/pdlp.hpp:35-38) and its arithmetic (FP64)
    assert (o.max_iterations, o.check_every, o.tol, o.mode) == (200000, 64, 1e-6, P.MODE_FAST_FP64)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_solve_without_device_fails_loudly():
    lp = P.gen_random_lp(20, 40, 6, seed=1)
    s, Dr, Dc = P.scale(lp)
    with pytest.raises(P.SupremaError) as e:
        P.solve_pdhg_gpu(lp, s, Dr, Dc)
    assert e.value.code == _lib.SUPR_ENODEV


def test_solve_rejects_mismatched_forms():
    lp = P.gen_random_lp(20, 40, 6, seed=1)
    s, Dr, Dc = P.scale(lp)
    other = P.gen_random_lp(20, 40, 6, seed=2)
    with pytest.raises(P.SupremaError) as e:
        P.solve_pdhg_gpu(other, s, Dr, Dc)
    assert e.value.code == _lib.SUPR_EINVAL
    with pytest.raises(P.SupremaError):
        P.solve_pdhg_gpu(lp, s, Dr, Dc, check_every=0)
    with pytest.raises(P.SupremaError):
        P.solve_pdhg_gpu(lp, s, Dr, Dc, spmv_lanes_A=3)


REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present here")
def test_adapter_compiles_against_the_reference_headers():
    """team-suprema_amd/adapter/pdlp_gpu_hip.cpp defines the reference's
    solve_pdhg_gpu with the reference's own RangedLP / PdlpResult types."""
    import subprocess
    src = os.path.join(ROOT, "team-suprema_amd", "adapter", "pdlp_gpu_hip.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DSUPREMA_WITH_HIP=1",
                        "-DSOVEREIGN_WITH_CUDA=SUPREMA_WITH_HIP",
                        "-I" + REF_INCLUDE, "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_headers_compile_standalone():
    import subprocess
    for h in ("pdlp_gpu.hpp", "verify.hpp", "pdlp.hpp", "suprema_pdlp.hpp"):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++",
                            "-I" + os.path.join(ROOT, "include"), "-"],
                           input=f'#include "{h}"\nint main() {{ return 0; }}\n',
                           capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr)
