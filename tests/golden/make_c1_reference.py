"""Runs the reference's own CPU PDLP (oracle/_ref, compiled from the reference
sources) to completion on the 1,000,000-variable transportation LP
(supr_gen_transport(1000, 1000, seed 42), tol 1e-6, check_every 64) and
writes tests/golden/c1_reference.json: its iteration / restart counts, eps,
objective and single-thread wall time on the machine that ran it.  The
benchmark's CPU-baseline leg times a bounded prefix of the same run on the
benchmark machine's host cores and extrapolates with these counts."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "team-suprema_amd"))

import pdlp_amd as P  # noqa: E402
from oracle import Checker  # noqa: E402

lp = P.gen_transport(1000, 1000, seed=42)
scaled, Dr, Dc = P.scale(lp)
ref = Checker("ref")
t0 = time.time()
out = ref.pdlp(lp, scaled, Dr, Dc, 500000, 64, 1e-6)
wall = time.time() - t0
rep = ref.verify(lp, out.x, out.y)
res = {"workload": "transport 1000x1000 seed 42 (m=2000, n=1000000, nnz=2000000)",
       "tol": 1e-6, "check_every": 64, "iterations": out.iterations, "restarts": out.restarts,
       "converged": out.converged, "eps": [out.eps_P, out.eps_D, out.eps_G],
       "objective": float(rep[3]), "wall_seconds_1_thread": wall,
       "host": os.uname().machine + " " + str(os.cpu_count()) + " logical cpus"}
with open(os.path.join(HERE, "c1_reference.json"), "w") as f:
    json.dump(res, f, indent=1)
print(json.dumps(res))
