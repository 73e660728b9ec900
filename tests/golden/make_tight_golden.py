"""Writes tests/golden/tight.json: the reference CPU engine's own solve of
every golden case (PDLP form, reference scale()) to a relative KKT tolerance
of 1e-9 (max 500,000 attempts), through oracle/_ref: attempts, convergence,
eps and the objective of the PDLP form.  Two points certified at 1e-6 can
differ in objective by a few 1e-6 (measured: 2.3e-6 relative on blend); at
1e-9 both sit next to the optimum, so the GPU engine's 1e-9 objective is
compared with these to 1e-7.

Run where the reference sources exist and oracle/_ref/libsovref.so is built:

    python tests/golden/make_tight_golden.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "team-suprema_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import golden_lp  # noqa: E402
from oracle import Checker  # noqa: E402

TOL = 1e-9
MAX_ITER = 500000


def main():
    ref = Checker("ref")
    with open(os.path.join(HERE, "golden.json")) as f:
        cases = json.load(f)["cases"]
    out = {"tol": TOL, "max_iterations": MAX_ITER, "cases": []}
    for case in cases:
        form = golden_lp(case)
        s = ref.scale(form)
        t0 = time.perf_counter()
        r = ref.pdlp(form, form.with_values(*s[:6]), s[6], s[7], MAX_ITER, 64, TOL)
        dt = time.perf_counter() - t0
        rep = ref.verify(form, r.x, r.y)
        out["cases"].append({"name": case["name"], "iterations": r.iterations,
                             "converged": r.converged, "eps": [r.eps_P, r.eps_D, r.eps_G],
                             "objective_form": float(rep[3])})
        print(f"{case['name']:24s} it={r.iterations:7d} conv={r.converged} obj={rep[3]:.12g} "
              f"({dt:.1f} s)", flush=True)
    with open(os.path.join(HERE, "tight.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
