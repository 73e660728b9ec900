"""Writes the golden fixtures under tests/golden/ (run in a tree where the
reference sources exist and oracle/_ref/libsovref.so is built):

  netlib/<name>.npz   the reference's LP fixtures (sample_problems/*.mps and the
                      transport instances med.mps / illcond.mps), read by the
                      reference's own read_mps + to_ranged_lp -- input data only
  golden.json         what the reference's own CPU engine (src/pdlp.cpp +
                      include/pdlp_algo.hpp, compiled from its sources) returns
                      on each case: iterations, restarts, convergence, eps,
                      objectives and SHA-256 digests of the scaled form and of x, y

Usage: python tests/golden/make_golden.py [/path/to/reference]
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
from oracle import Checker, read_mps_ref  # noqa: E402

# Published optimal objectives of the Netlib fixtures (tools/audit_lp.cpp:148-160)
NETLIB_OPT = {
    "afiro": -464.7531428571429, "adlittle": 225494.9631162, "blend": -30.81214985,
    "israel": -896644.8218, "kb2": -1749.90023670, "sc50a": -64.57506, "sc50b": -70.0,
    "sc105": -52.20206, "scagr7": -2331389.824, "share1b": -76589.31857918584,
    "beaconfd": 33592.4858072,
}
FIXTURES = ["afiro", "adlittle", "blend", "sc50a", "sc50b", "sc105", "kb2", "scagr7", "israel",
            "share1b", "beaconfd", "tiny_range"]
TRANSPORT_FIXTURES = ["med", "illcond"]


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def lp_from(d, name):
    return P.LP(d["m"], d["n"], d["row"], d["col"], d["val"], d["c"], d["rL"], d["rU"], d["l"],
                d["u"], d["obj_offset"], name)


def run_case(ref, form, tol, max_iter, check_every=64):
    s = ref.scale(form)
    scaled = form.with_values(*s[:6])
    Dr, Dc = s[6], s[7]
    out = ref.pdlp(form, scaled, Dr, Dc, max_iter, check_every, tol)
    rep = ref.verify(form, out.x, out.y)
    return {
        "tol": tol, "max_iterations": max_iter, "check_every": check_every,
        "iterations": out.iterations, "restarts": out.restarts, "converged": out.converged,
        "eps": [out.eps_P, out.eps_D, out.eps_G], "objective_form": float(rep[3]),
        "scaled_sha": sha(*s), "x_sha": sha(out.x), "y_sha": sha(out.y),
        "x_sum": float(np.sum(out.x)), "y_sum": float(np.sum(out.y)),
    }


def main():
    refdir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    ref = Checker("ref")
    os.makedirs(os.path.join(HERE, "netlib"), exist_ok=True)
    golden = {"cases": []}
    for name in FIXTURES + TRANSPORT_FIXTURES:
        path = (os.path.join(refdir, "sample_problems", name + ".mps") if name in FIXTURES
                else os.path.join(refdir, name + ".mps"))
        d = read_mps_ref(path)
        np.savez_compressed(os.path.join(HERE, "netlib", name + ".npz"),
                            **{k: np.asarray(v) for k, v in d.items()})
        lp = lp_from(d, name)
        form_dims, ri, ci, arrs, perm, signs = ref.pdlp_form(lp)
        form = P.LP(int(form_dims[0]), int(form_dims[1]), ri, ci, *arrs, lp.obj_offset, name)
        case = {"name": name, "kind": "mps", "m": lp.m, "n": lp.n, "nnz": lp.nnz,
                "maximize": bool(d["maximize"]), "form_dims": [int(v) for v in form_dims],
                "form_sha": sha(ri, ci, *arrs, perm, signs)}
        tol = 1e-6
        max_iter = 500000
        case["ref"] = run_case(ref, form, tol, max_iter)
        if name in NETLIB_OPT:
            case["published_optimum"] = NETLIB_OPT[name]
        golden["cases"].append(case)
        r = case["ref"]
        print(f"{name:10s} m={lp.m:4d} n={lp.n:4d} it={r['iterations']:6d} rs={r['restarts']:3d} "
              f"conv={r['converged']} obj={r['objective_form']:.10g}", flush=True)
    # Synthetic cases from the product's seeded generators
    synth = [
        ("random_200x500_s1", dict(kind="random", m=200, n=500, col_nnz=6, seed=1)),
        ("random_2000x5000_s42", dict(kind="random", m=2000, n=5000, col_nnz=6, seed=42)),
        ("transport_60x60_s42", dict(kind="transport", S=60, D=60, seed=42)),
        ("transport_30x40_s7", dict(kind="transport", S=30, D=40, seed=7)),
        ("power_law_1000x4000_s7", dict(kind="power_law", m=1000, n=4000, seed=7)),
        ("blend_25x40_s103", dict(kind="blend", K=25, P=40, seed=103)),
    ]
    for name, g in synth:
        if g["kind"] == "random":
            lp = P.gen_random_lp(g["m"], g["n"], g["col_nnz"], seed=g["seed"])
        elif g["kind"] == "power_law":
            lp = P.gen_power_law_lp(g["m"], g["n"], seed=g["seed"])
        elif g["kind"] == "blend":
            lp, _ = P.build_pdlp_form(P.gen_blend(g["K"], g["P"], seed=g["seed"]))
        else:
            lp = P.gen_transport(g["S"], g["D"], seed=g["seed"])
        case = {"name": name, "kind": g["kind"], "gen": g, "m": lp.m, "n": lp.n, "nnz": lp.nnz,
                "lp_sha": sha(lp.row, lp.col, lp.val, lp.c, lp.rL, lp.rU, lp.l, lp.u)}
        case["ref"] = run_case(ref, lp, 1e-6, 200000)
        golden["cases"].append(case)
        r = case["ref"]
        print(f"{name:22s} it={r['iterations']:6d} rs={r['restarts']:3d} conv={r['converged']} "
              f"obj={r['objective_form']:.10g}", flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
