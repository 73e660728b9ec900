"""Writes tests/golden/mps/: the reference's MPS fixtures and what the
reference's own reader makes of them.

For every file in the reference's sample_problems/ plus the transport
instances med.mps / illcond.mps (data files the reference's tests and
audits read), this stores
  mps/<name>.mps.gz   the file itself, compressed (input data), and
  mps/<name>.npz      read_mps() + to_ranged_lp() of the reference
                      (src/mps_reader.cpp:54-308, include/ranged_lp.hpp:
                      27-62), run through oracle/_ref (the reference's own
                      sources compiled by oracle/Makefile): the LP arrays,
                      objective offset, sense, integer flags, QP lower
                      triangle and names (expected output).
Run from the repository root where the reference tree exists:
    python tests/golden/make_mps_golden.py [/root/reference]
"""
import glob
import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import available, read_mps_ref  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    if not available("ref"):
        raise SystemExit("oracle/_ref/libsovref.so is missing: run `make -C oracle` first")
    files = sorted(glob.glob(os.path.join(ref, "sample_problems", "*.mps"))
                   + glob.glob(os.path.join(ref, "sample_problems", "*.qps"))
                   + [os.path.join(ref, "med.mps"), os.path.join(ref, "illcond.mps")])
    out = os.path.join(HERE, "mps")
    os.makedirs(out, exist_ok=True)
    for f in files:
        base = os.path.basename(f)
        name = base.rsplit(".", 1)[0] + ("_qps" if base.endswith(".qps") else "")
        with open(f, "rb") as fh:
            raw = fh.read()
        with gzip.GzipFile(os.path.join(out, name + ".mps.gz"), "wb", mtime=0) as gz:
            gz.write(raw)
        d = read_mps_ref(f, full=True)
        np.savez_compressed(
            os.path.join(out, name + ".npz"), m=d["m"], n=d["n"], row=d["row"], col=d["col"],
            val=d["val"], c=d["c"], rL=d["rL"], rU=d["rU"], l=d["l"], u=d["u"],
            obj_offset=d["obj_offset"], maximize=d["maximize"], integer=d["integer"],
            q_row=d["q_row"], q_col=d["q_col"], q_val=d["q_val"], name=np.array(d["name"]),
            col_names=np.array(d["col_names"], dtype=str), row_names=np.array(d["row_names"], dtype=str))
        print(f"{name}: m={d['m']} n={d['n']} nnz={len(d['val'])} q={len(d['q_val'])}")


if __name__ == "__main__":
    main()
