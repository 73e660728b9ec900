import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "team-suprema_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import pdlp_amd as P  # noqa: E402
from oracle import Checker, available  # noqa: E402

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


@pytest.fixture(scope="session")
def oracle():
    return Checker("oracle")


@pytest.fixture(scope="session")
def ref():
    if not available("ref"):
        pytest.skip("reference CPU engine (oracle/_ref) not built here")
    return Checker("ref")


def load_fixture(name):
    """A reference LP fixture (tests/golden/netlib/<name>.npz) as a pdlp_amd.LP."""
    d = np.load(os.path.join(GOLDEN_DIR, "netlib", name + ".npz"))
    return P.LP(int(d["m"]), int(d["n"]), d["row"], d["col"], d["val"], d["c"], d["rL"], d["rU"],
                d["l"], d["u"], float(d["obj_offset"]), name), bool(d["maximize"])


def golden_lp(case):
    """The PDLP-form LP of a golden case (fixture or generated)."""
    if case["kind"] == "mps":
        lp, _ = load_fixture(case["name"])
        form, _ = P.build_pdlp_form(lp)
        return form
    g = case["gen"]
    if g["kind"] == "random":
        return P.gen_random_lp(g["m"], g["n"], g["col_nnz"], seed=g["seed"])
    if g["kind"] == "power_law":
        return P.gen_power_law_lp(g["m"], g["n"], seed=g["seed"])
    if g["kind"] == "blend":
        return P.build_pdlp_form(P.gen_blend(g["K"], g["P"], seed=g["seed"]))[0]
    return P.gen_transport(g["S"], g["D"], seed=g["seed"])


def gpu_available():
    try:
        return P.device_count() > 0
    except OSError:
        return False
