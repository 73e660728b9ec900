"""The library's host-side code and the oracle under sanitizers (CPU, no
device code): team-suprema_amd/_build/sanitize_asan (AddressSanitizer +
UndefinedBehaviorSanitizer) reads every MPS fixture, runs build_pdlp_form,
scale, a short oracle PDLP run, extract_solution and verify on each LP, the
generators, an MPS write -> read round trip and the argument errors;
_build/sanitize_tsan (ThreadSanitizer) reads the fixtures and the round-trip
file with the threaded COLUMNS pass on 8 threads and tiny chunks.  Both are
built by team-suprema_amd/Makefile's `sanitize` target (part of `all`)."""
import gzip
import os
import subprocess

import pytest

from conftest import GOLDEN_DIR, ROOT

BUILD = os.path.join(ROOT, "team-suprema_amd", "_build")


@pytest.fixture(scope="module")
def mps_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("sanitize_mps")
    src = os.path.join(GOLDEN_DIR, "mps")
    for f in os.listdir(src):
        if f.endswith(".mps.gz"):
            with gzip.open(os.path.join(src, f)) as g:
                (d / f[:-3]).write_bytes(g.read())
    return str(d)


def run(exe, args):
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([path] + args, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout[-3000:], r.stderr[-5000:])
    assert "runtime error" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-5000:]
    return r


def test_host_code_and_oracle_under_asan_ubsan(mps_dir):
    r = run("sanitize_asan", [mps_dir])
    assert r.stdout.count("verify eps") >= 25


def test_threaded_mps_reader_under_tsan(mps_dir):
    run("sanitize_tsan", [mps_dir, "8"])
