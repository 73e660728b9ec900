"""The reference's own callers and instances over the HIP engine.

1. The reference's transportation instances (big.mps: 160,000 columns;
   transport_600.mps: 360,000 columns; tests/golden/mps/*.mps.xz) from the
   file to a verified solution through solve_lp_gpu, against the
   reference's own CPU pipeline on them (tests/golden/transport_ref.json:
   read_mps, build_pdlp_form, scale(), solve_pdhg to 1e-6, extract_solution,
   verify(), compiled from the reference's sources): the oracle's verifier
   certifies 1e-6 on the LP, the objective matches the reference's to 1e-6
   and the integer optimum (integer data) to 1e-6.
2. The reference's CLI (src/main.cpp + src/*.cpp, compiled where they lie by
   oracle/Makefile's `refcli` target with the GPU gate on) with
   solve_pdhg_gpu served by team-suprema_amd/adapter/pdlp_gpu_hip.cpp:
   * `--pdlp`: solve_mps's race of the CPU solve_pdhg against the drop-in on
     two threads with one stop flag (src/solve.cpp:71-148, 194-224);
   * `--engines=pdlp`: portfolio engine 3, PDLP on the GPU to 1e-4, then
     crossover_start and the dual simplex (src/solve.cpp:553-574).
   Objectives: the published Netlib optima (tools/audit_lp.cpp:148-160) and
   the transportation LP's integer optimum; the race's answer within the
   audit's 1e-4, the simplex-finished portfolio answer within 1e-6.
Skipped where oracle/_ref/sovereign_solve was not built (no reference tree
at build time)."""
import gzip
import json
import lzma
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import pdlp_amd as P
from conftest import GOLDEN_DIR, ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "oracle", "_ref", "sovereign_solve")
TOL = 1e-6


def say(*a):
    print("[callers]", *a, file=sys.stderr, flush=True)


def fixture_file(tmp_path, name):
    src = os.path.join(GOLDEN_DIR, "mps", name + ".mps")
    path = tmp_path / (name + ".mps")
    if os.path.exists(src + ".xz"):
        with lzma.open(src + ".xz") as f:
            path.write_bytes(f.read())
    else:
        with gzip.open(src + ".gz") as f:
            path.write_bytes(f.read())
    return str(path)


@pytest.fixture(scope="module")
def transport_ref():
    with open(os.path.join(GOLDEN_DIR, "transport_ref.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


@pytest.mark.parametrize("name", ["big", "transport_600"])
@pytest.mark.parametrize("mode", [P.MODE_FAST_FP64, P.MODE_MIXED])
def test_reference_transport_instance_file_to_verified_solution(oracle, tmp_path, transport_ref, name, mode):
    ref = transport_ref[name]
    lp = P.read_mps(fixture_file(tmp_path, name), names=False).lp
    assert (lp.m, lp.n, lp.nnz) == (ref["m"], ref["n"], ref["nnz"])
    sol = P.solve_lp_gpu(lp, tol=TOL, mode=mode)
    rep = oracle.verify(lp, sol.x, sol.y)
    assert sol.status == "optimal" and max(rep[:3]) <= TOL, (sol.status, rep[:3])
    err = abs(rep[3] - ref["objective"]) / (1 + abs(ref["objective"]))
    err_int = abs(rep[3] - ref["integer_optimum"]) / (1 + abs(ref["integer_optimum"]))
    say(f"{name} mode {mode}: {sol.iterations} attempts (reference CPU {ref['iterations']} in "
        f"{ref['cpu_seconds']:.1f} s), solve_lp_gpu {sol.info['seconds_total']:.3f} s, objective {rep[3]:.6f} "
        f"(error vs the reference {err:.1e}, vs the integer optimum {err_int:.1e})")
    assert err <= 1e-6 and err_int <= 1e-6


def run_cli(args, timeout=300):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)
    fields = dict(re.findall(r"^(\w[\w ]*?):\s+(.*)$", r.stdout, flags=re.M))
    return r, fields


def optimum(golden, transport_ref, name):
    if name in transport_ref:
        return float(transport_ref[name]["integer_optimum"])
    return golden[name]["published_optimum"]


needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="oracle/_ref/sovereign_solve not built here")


@needs_cli
@pytest.mark.parametrize("name", ["afiro", "sc50a", "blend", "big", "transport_600"])
def test_reference_cli_pdlp_race_over_the_drop_in(tmp_path, golden, transport_ref, name):
    r, f = run_cli(["--pdlp", fixture_file(tmp_path, name)])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    obj, want = float(f["Objective"]), optimum(golden, transport_ref, name)
    say(f"--pdlp {name}: status {f['Status']}, engine {f['Engine']}, {f.get('Iterations')}, "
        f"objective {obj:.8g} (optimum {want:.8g}), {f.get('Time')}")
    assert f["Status"] in ("optimal", "near_optimal")
    assert abs(obj - want) / (1 + abs(want)) <= 1e-4
    assert "GPU engine failed" not in r.stderr, r.stderr[-2000:]
    if name in transport_ref:
        # 160,000 / 360,000 columns: the GPU wins the race (the reference CPU engine
        # takes 10.7 / 214 s on them alone)
        assert f["Engine"].startswith("gpu"), f["Engine"]


@needs_cli
@pytest.mark.parametrize("name", ["afiro", "sc50a", "blend", "big", "transport_600"])
def test_reference_cli_portfolio_pdlp_crossover_simplex(tmp_path, golden, transport_ref, name):
    r, f = run_cli(["--engines=pdlp", fixture_file(tmp_path, name), "240"])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    obj, want = float(f["Objective"]), optimum(golden, transport_ref, name)
    say(f"--engines=pdlp {name}: status {f['Status']}, engine {f.get('Engine')}, {f.get('Iterations')}, "
        f"objective {obj:.10g} (optimum {want:.10g}), {f.get('Time')}")
    assert f["Status"] == "optimal"
    assert abs(obj - want) / (1 + abs(want)) <= 1e-6
    assert "failed" not in r.stderr.lower(), r.stderr[-2000:]
