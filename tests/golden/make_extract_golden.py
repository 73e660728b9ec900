"""Writes tests/golden/extract.json: the reference's own extract_solution()
(src/pdlp.cpp:84-90, through oracle/_ref's ref_extract_solution) applied to
the reference CPU engine's PDLP-form solution of every MPS golden case, and
the reference's verify() of the extracted point on the original LP.

Per case: SHA-256 of the extracted x and y, and verify()'s seven outputs
(eps_P, eps_D, eps_G, objective, dual objective, max row / bound violation).
Run where the reference sources exist and oracle/_ref/libsovref.so is built:

    python tests/golden/make_extract_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "team-suprema_amd"))

import pdlp_amd as P  # noqa: E402
from oracle import Checker  # noqa: E402


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    ref = Checker("ref")
    with open(os.path.join(HERE, "golden.json")) as f:
        cases = [c for c in json.load(f)["cases"] if c["kind"] == "mps"]
    out = {"cases": []}
    for case in cases:
        d = np.load(os.path.join(HERE, "netlib", case["name"] + ".npz"))
        lp = P.LP(int(d["m"]), int(d["n"]), d["row"], d["col"], d["val"], d["c"], d["rL"], d["rU"],
                  d["l"], d["u"], float(d["obj_offset"]), case["name"])
        dims, ri, ci, arrs, perm, signs = ref.pdlp_form(lp)
        form = P.LP(int(dims[0]), int(dims[1]), ri, ci, *arrs, lp.obj_offset, lp.name)
        s = ref.scale(form)
        r = case["ref"]
        sol = ref.pdlp(form, form.with_values(*s[:6]), s[6], s[7], r["max_iterations"],
                       r["check_every"], r["tol"])
        x, y = ref.extract_solution(lp.n, lp.m, perm, signs, sol.x, sol.y)
        rep = ref.verify(lp, x, y)
        out["cases"].append({"name": case["name"], "n_orig": lp.n, "m_orig": lp.m,
                             "x_sha": sha(x), "y_sha": sha(y),
                             "verify_on_lp": [float(v) for v in rep]})
        print(f"{case['name']:10s} obj={rep[3]:.10g} eps={max(rep[:3]):.2e}", flush=True)
    with open(os.path.join(HERE, "extract.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
