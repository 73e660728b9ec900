"""Writes tests/golden/transport_ref.json: the reference's own pipeline on its
two transportation instances (big.mps: 400 x 400 sources and sinks, 160,000
columns; transport_600.mps: 600 x 600, 360,000 columns; kept here as
mps/*.mps.xz) -- read_mps + to_ranged_lp, build_pdlp_form, scale(), the CPU
solve_pdhg to 1e-6, extract_solution and verify() on the LP, all through
oracle/_ref (the reference's sources compiled here).  Per instance: the
sizes, attempts, restarts, eps and objective, the CPU seconds, and the
integer optimum (integer supplies, demands and costs: the optimal cost is an
integer, and the 1e-6 objective rounds to it).

Run where the reference sources exist and oracle/_ref/libsovref.so is built:

    python tests/golden/make_transport_golden.py
"""
import json
import lzma
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "team-suprema_amd"))

import pdlp_amd as P  # noqa: E402
from oracle import Checker, read_mps_ref  # noqa: E402


def main():
    ref = Checker("ref")
    out = {"cases": []}
    for name in ("big", "transport_600"):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, name + ".mps")
            with lzma.open(os.path.join(HERE, "mps", name + ".mps.xz")) as f, open(path, "wb") as g:
                g.write(f.read())
            r = read_mps_ref(path)
        lp = P.LP(r["m"], r["n"], r["row"], r["col"], r["val"], r["c"], r["rL"], r["rU"], r["l"], r["u"],
                  r["obj_offset"], name)
        dims, ri, ci, arrs, perm, signs = ref.pdlp_form(lp)
        form = P.LP(int(dims[0]), int(dims[1]), ri, ci, *arrs, lp.obj_offset, name)
        s = ref.scale(form)
        t0 = time.perf_counter()
        sol = ref.pdlp(form, form.with_values(*s[:6]), s[6], s[7], 500000, 64, 1e-6)
        dt = time.perf_counter() - t0
        x, y = ref.extract_solution(lp.n, lp.m, perm, signs, sol.x, sol.y)
        rep = ref.verify(lp, x, y)
        case = {"name": name, "m": lp.m, "n": lp.n, "nnz": lp.nnz, "iterations": sol.iterations,
                "restarts": sol.restarts, "converged": sol.converged, "eps_on_lp": [float(v) for v in rep[:3]],
                "objective": float(rep[3]), "integer_optimum": round(float(rep[3])), "cpu_seconds": dt}
        out["cases"].append(case)
        print(json.dumps(case), flush=True)
    with open(os.path.join(HERE, "transport_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
