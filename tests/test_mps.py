"""MPS ingestion (supr_mps_*): read_mps + to_ranged_lp of the reference
(src/mps_reader.cpp:54-308, include/ranged_lp.hpp:27-62) and the writer.

Pinned by
  * tests/golden/mps/: every file of the reference's sample_problems/ plus the
    med / illcond transport instances, beside what the reference's own reader
    made of them (make_mps_golden.py, through oracle/_ref): every array bit
    for bit, the names, integer flags, QP triangle and sense;
  * hand-written files whose LP follows from the format's rules (each bound
    type, ranges of both signs, fixed format, OBJSENSE, free and stray rows,
    explicit zeros, the objective constant) and the error texts;
  * the number conversion against Python's float() (correctly rounded, as
    strtod) on random decimal strings;
  * where oracle/_ref is built: the reference reader on the same hand-written
    files and on randomly generated ones, free and fixed format (differential);
  * round trips: write_mps then read_mps gives the LP back bit for bit, and the
    reference reader reads the written file to the same LP.
"""
import glob
import gzip
import os
import random

import numpy as np
import pytest

import pdlp_amd as P
from conftest import GOLDEN_DIR, load_fixture
from oracle import available

MPS_DIR = os.path.join(GOLDEN_DIR, "mps")
FIXTURES = sorted(os.path.basename(f)[:-7] for f in glob.glob(os.path.join(MPS_DIR, "*.mps.gz")))
INF = np.inf
ARRAYS = ("row", "col", "val", "c", "rL", "rU", "l", "u")


def bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def same_as_dict(mp, d, names=True):
    lp = mp.lp
    assert (lp.m, lp.n) == (d["m"], d["n"])
    for k in ARRAYS:
        assert bits(getattr(lp, k), d[k]), k
    assert bits(np.float64(lp.obj_offset), np.float64(d["obj_offset"]))
    assert mp.maximize == bool(d["maximize"])
    if "integer" in d:
        assert bits(mp.integer, np.asarray(d["integer"], dtype=np.int8))
        assert bits(mp.q_row, d["q_row"]) and bits(mp.q_col, d["q_col"]) and bits(mp.q_val, d["q_val"])
    if names and "col_names" in d:
        assert mp.col_names == [str(v) for v in d["col_names"]]
        assert mp.row_names == [str(v) for v in d["row_names"]]
        assert mp.name == str(d["name"])


@pytest.mark.parametrize("name", FIXTURES)
def test_reader_matches_the_reference_reader_on_its_fixtures(tmp_path, name):
    with gzip.open(os.path.join(MPS_DIR, name + ".mps.gz"), "rb") as f:
        path = write(tmp_path, name + ".mps", f.read().decode("latin-1"))
    d = dict(np.load(os.path.join(MPS_DIR, name + ".npz")))
    same_as_dict(P.read_mps(path), d)


def test_fixture_set_covers_the_reference_samples():
    assert len(FIXTURES) >= 27
    for name in ("afiro", "israel", "lseu", "p0033", "hs118_qps", "refinery_12x12_qp", "med"):
        assert name in FIXTURES
    # the LP fixtures the solver tests use were read by the same reader
    for name in ("afiro", "share1b", "med", "tiny_range"):
        lp, maximize = load_fixture(name)
        d = dict(np.load(os.path.join(MPS_DIR, name + ".npz")))
        assert all(bits(getattr(lp, k), d[k]) for k in ARRAYS) and maximize == bool(d["maximize"])


RULES = """NAME RULES
ROWS
 N  obj
 L  r1
 G  r2
 E  r3
 E  r4
 N  free1
COLUMNS
    x1  obj  1.5  r1  2
    x1  r2  -1  free1  7
    x2  obj  -2  r3  0
    x2  r4  3.25  ghost  9
    x3  r1  1e-3  r2  4
    x5  r3  1
    x6  r3  2
    x7  r3  3
    x8  r3  4
    MARKER  'MARKER'  'INTORG'
    x9  r3  5
    MARKER  'MARKER'  'INTEND'
    x10  r3  6
RHS
    RHS  obj  -10  r1  5
    RHS  r2  1  r3  2
    RHS  r4  -4
RANGES
    RNG  r1  2  r2  -3
    RNG  r3  1.5  r4  -0.5
BOUNDS
 UP BND x1  4
 MI BND x2
 UP BND x2  -1
 LO BND x3  -2
 UP BND x3  -1
 BV BND x4
 UP BND x5  -3
 FR BND x6
 FX BND x7  2.5
 UP BND x8  3
 PL BND x8
 LI BND x10  -1
 UI BND x10  5
ENDATA
"""


def test_format_rules_closed_form(tmp_path):
    mp = P.read_mps(write(tmp_path, "rules.mps", RULES))
    lp = mp.lp
    assert mp.name == "RULES" and mp.obj_name == "obj" and not mp.maximize
    assert mp.row_names == ["r1", "r2", "r3", "r4"]
    assert mp.col_names == ["x1", "x2", "x3", "x5", "x6", "x7", "x8", "x9", "x10", "x4"]
    # COO in file order; the objective, the free row, the explicit zero and
    # the undeclared row 'ghost' are dropped
    assert lp.row.tolist() == [0, 1, 3, 0, 1, 2, 2, 2, 2, 2, 2]
    assert lp.col.tolist() == [0, 0, 1, 2, 2, 3, 4, 5, 6, 7, 8]
    assert lp.val.tolist() == [2, -1, 3.25, 1e-3, 4, 1, 2, 3, 4, 5, 6]
    assert lp.c.tolist() == [1.5, -2, 0, 0, 0, 0, 0, 0, 0, 0]
    assert lp.obj_offset == 10.0   # an RHS on the objective row is minus the constant
    # L: [rhs - |r|, rhs]; G: [rhs, rhs + |r|]; E: r >= 0 [rhs, rhs + r], r < 0 [rhs + r, rhs]
    assert lp.rL.tolist() == [3, 1, 2, -4.5] and lp.rU.tolist() == [5, 4, 3.5, -4]
    assert lp.l.tolist() == [0, -INF, -2, -INF, -INF, 2.5, 0, 0, -1, 0]
    assert lp.u.tolist() == [4, -1, -1, -3, INF, 2.5, INF, INF, 5, 1]
    assert mp.integer.tolist() == [0, 0, 0, 0, 0, 0, 0, 1, 1, 1]


def test_objsense_max_negates_the_objective(tmp_path):
    text = ("NAME MAXLP\nOBJSENSE\n    MAX\nROWS\n N  z\n L  c1\nCOLUMNS\n    a  z  3  c1  1\n"
            "    b  z  -1  c1  1\nRHS\n    RHS  z  2.5  c1  4\nENDATA\n")
    mp = P.read_mps(write(tmp_path, "max.mps", text))
    assert mp.maximize
    assert mp.lp.c.tolist() == [-3, 1] and mp.lp.obj_offset == 2.5
    assert mp.lp.rL.tolist() == [-INF] and mp.lp.rU.tolist() == [4]
    # the same sense on the keyword line itself is ignored (the section keyword
    # consumes its line)
    mp2 = P.read_mps(write(tmp_path, "max2.mps", text.replace("OBJSENSE\n    MAX\n", "OBJSENSE MAX\n")))
    assert not mp2.maximize and mp2.lp.c.tolist() == [3, -1]


def fixed_line(*fields):
    """A fixed-format line: fields at columns 2-3, 5-12, 15-22, 25-36, 40-47, 50-61."""
    starts = (1, 4, 14, 24, 39, 49)
    out = [" "] * 61
    for s, f in zip(starts, fields):
        out[s:s + len(f)] = list(f)
    return "".join(out).rstrip()


def fixed_text():
    return "\n".join([
        "NAME          FIXED", "ROWS", fixed_line("N", "COST"), fixed_line("L", "LIM 1"),
        fixed_line("G", "LIM 2"), "COLUMNS",
        fixed_line("", "X ONE", "COST", "1.0", "LIM 1", "2.0"),
        fixed_line("", "X ONE", "LIM 2", "-1.5"),
        fixed_line("", "Y", "LIM 2", "1"),
        "RHS", fixed_line("", "", "LIM 1", "4.0", "LIM 2", "-2"),
        "BOUNDS", fixed_line("UP", "BND", "X ONE", "3.0"), fixed_line("MI", "", "Y"),
        "ENDATA", ""])


def test_fixed_format_fallback_with_spaces_in_names(tmp_path):
    mp = P.read_mps(write(tmp_path, "fixed.mps", fixed_text()))
    assert mp.row_names == ["LIM 1", "LIM 2"] and mp.col_names == ["X ONE", "Y"]
    lp = mp.lp
    assert lp.row.tolist() == [0, 1, 1] and lp.col.tolist() == [0, 0, 1]
    assert lp.val.tolist() == [2.0, -1.5, 1.0] and lp.c.tolist() == [1.0, 0.0]
    assert lp.rL.tolist() == [-INF, -2] and lp.rU.tolist() == [4, INF]
    assert lp.l.tolist() == [0, -INF] and lp.u.tolist() == [3, INF]


def test_errors_carry_the_reference_text(tmp_path):
    def err(text):
        with pytest.raises(P.SupremaError) as e:
            P.read_mps(write(tmp_path, "bad.mps", text))
        return str(e.value)
    assert "MPS parse failed: No objective (N) row found in MPS file" in err(
        "ROWS\n L  r\nCOLUMNS\n    x  r  1\nENDATA\n")
    assert "MPS parse failed: Unsupported BOUNDS type 'XX' in MPS file" in err(
        "ROWS\n N  o\nCOLUMNS\n    x  o  1\nBOUNDS\n XX BND x  3\nENDATA\n")
    assert "MPS parse failed: stod" in err("ROWS\n N  o\nCOLUMNS\n    x  o  1e-400\nENDATA\n")
    missing = str(tmp_path / "none.mps")
    with pytest.raises(P.SupremaError, match="MPS parse failed: cannot open MPS file: " + missing):
        P.read_mps(missing)


def test_number_conversion_is_correctly_rounded(tmp_path):
    """The fast decimal path and the strtod fallback against float(): random
    mantissas of 1-22 digits, decimal points anywhere, exponents -40..40,
    signs, leading zeros, plus strtod-only forms (hex, trailing text)."""
    rng = random.Random(5)
    toks = []
    for _ in range(4000):
        digits = "".join(rng.choice("0123456789") for _ in range(rng.randint(1, 22)))
        k = rng.randint(0, len(digits))
        t = digits[:k] + ("." if rng.random() < 0.7 else "") + digits[k:]
        if t == ".":
            t = "0."
        if rng.random() < 0.5:
            t += rng.choice("eE") + rng.choice(["", "+", "-"]) + str(rng.randint(0, 40))
        if rng.random() < 0.3:
            t = "-" + t
        toks.append(t)
    toks += ["0", "-0", "1.", ".5", "00012", "9007199254740993", "1e22", "1e23", "123456789012345678",
             "4.35", "0.1", "2.2250738585072014e-308", "1.7976931348623157e308"]
    lines = ["ROWS", " N  o", " E  r", "COLUMNS"]
    lines += [f"    c{j}  r  {t}" for j, t in enumerate(toks)]
    lines += ["ENDATA", ""]
    mp = P.read_mps(write(tmp_path, "nums.mps", "\n".join(lines)))
    want = np.array([float(t) for t in toks])
    keep = want != 0.0    # zero coefficients are dropped
    assert bits(mp.lp.val, want[keep])
    # strtod-only forms and trailing text behave as std::stod
    mp = P.read_mps(write(tmp_path, "odd.mps", "ROWS\n N  o\n E  r\nCOLUMNS\n    a  r  0x10\n"
                                               "    b  r  1.5abc\n    c  r  +2\n    d  r  inf\nENDATA\n"))
    assert mp.lp.val.tolist() == [16.0, 1.5, 2.0, INF]


def test_round_trip_through_the_writer(tmp_path):
    cases = [P.gen_random_lp(40, 90, 6, seed=3, mag_decades=3), P.gen_transport(12, 9, seed=4),
             P.build_pdlp_form(P.gen_blend(9, 5, seed=2))[0], P.gen_power_law_lp(300, 1200, seed=6)]
    for name in ("afiro", "israel", "share1b", "tiny_range", "med", "kb2"):
        cases.append(load_fixture(name)[0])
    rng = np.random.default_rng(1)
    lp = cases[0]
    # bounds of every kind and a ranged / free row mix on one LP
    l, u = lp.l.copy(), lp.u.copy()
    l[0:5], u[0:5] = -INF, INF
    l[5:10], u[5:10] = -INF, rng.uniform(-5, 5, 5)
    l[10:15], u[10:15] = rng.uniform(-5, 0, 5), -1.0
    l[15:20] = u[15:20] = 2.5
    rL, rU = lp.rL.copy(), lp.rU.copy()
    rL[0:4], rU[0:4] = -INF, INF
    rL[4:8] = np.round(rL[4:8] * 8) / 8 - 0.125   # dyadic bounds: rU - rL is exact
    rU[4:8] = rL[4:8] + 2.75
    cases.append(P.LP(lp.m, lp.n, lp.row, lp.col, lp.val, lp.c, rL, rU, l, u, -3.25, "mixed_kinds"))
    for k, lp in enumerate(cases):
        path = str(tmp_path / f"rt{k}.mps")
        integ = (np.arange(lp.n) % 7 == 3).astype(np.int8)
        P.write_mps(path, lp, integer=integ)
        back = P.read_mps(path)
        keep = lp.val != 0.0
        order = np.lexsort((np.arange(lp.nnz)[keep], lp.col[keep]))   # the writer groups by column
        assert back.lp.m == lp.m and back.lp.n == lp.n, lp.name
        assert bits(back.lp.row, lp.row[keep][order]) and bits(back.lp.col, lp.col[keep][order])
        assert bits(back.lp.val, lp.val[keep][order])
        for f in ("c", "rL", "rU", "l", "u"):
            assert bits(getattr(back.lp, f), getattr(lp, f)), (lp.name, f)
        assert back.lp.obj_offset == lp.obj_offset and bits(back.integer, integ)
        if available("ref"):
            from oracle import read_mps_ref
            same_as_dict(back, read_mps_ref(path, full=True))


def test_writer_names_and_refusals(tmp_path):
    lp = P.gen_transport(3, 2, seed=1)
    cols = [f"x_{i}_{j}" for i in range(3) for j in range(2)]
    rows = [f"S{i}" for i in range(3)] + [f"D{j}" for j in range(2)]
    path = str(tmp_path / "t.mps")
    P.write_mps(path, lp, col_names=cols, row_names=rows)
    back = P.read_mps(path)
    assert back.col_names == cols and back.row_names == rows and back.obj_name == "COST"
    with pytest.raises(P.SupremaError, match="white space"):
        P.write_mps(path, lp, col_names=["a b"] + cols[1:], row_names=rows)
    bad = P.LP(lp.m, lp.n, lp.row, lp.col, lp.val, lp.c, lp.rU + 1.0, lp.rU, lp.l, lp.u)
    with pytest.raises(P.SupremaError, match="no feasible range"):
        P.write_mps(path, bad)
    # a row named COST moves the objective row's name
    P.write_mps(path, lp, col_names=cols, row_names=["COST"] + rows[1:])
    assert P.read_mps(path).obj_name == "OBJ0"


# ------------------------------------------------- differential (oracle/_ref)

def random_mps(rng, fixed):
    """A random MPS text exercising the reader's rules (names short enough
    for fixed format; spaces inside names in the fixed variant)."""
    def name(prefix):
        body = "".join(rng.choice("ABCXYZ0123_") for _ in range(rng.randint(1, 5)))
        if fixed and rng.random() < 0.3:
            body = body[:3] + " " + body[3:]
        return (prefix + body)[:8]

    def num():
        r = rng.random()
        if r < 0.1:
            return "0"
        if r < 0.2:
            return str(rng.randint(-9, 9))
        return f"{rng.uniform(-100, 100):.{rng.randint(0, 9)}g}" if r < 0.8 else f"{rng.uniform(-1, 1):.6e}"

    def line(*f):
        return fixed_line(*f) if fixed else "    " + "  ".join(x for x in f[1:] if x) if f[0] == "" \
            else " " + "  ".join(x for x in f if x)

    rows = [name("R") for _ in range(rng.randint(1, 6))]
    kinds = [rng.choice("LGEEN") for _ in rows]
    cols = [name("C") for _ in range(rng.randint(1, 7))]
    out = ["NAME  RND"]
    if rng.random() < 0.3:
        out += ["OBJSENSE", "    " + rng.choice(["MAX", "MIN", "max"])]
    out += ["ROWS", line("N", "OBJ")] + [line(k, r) for k, r in zip(kinds, rows)]
    out.append("COLUMNS")
    for j, c in enumerate(cols):
        if rng.random() < 0.2:
            out.append("    MARKER  'MARKER'  '" + rng.choice(["INTORG", "INTEND"]) + "'")
        targets = ["OBJ"] + rows + ["STRAY"]
        picks = rng.sample(targets, rng.randint(1, min(3, len(targets))))
        for r in picks:
            out.append(line("", c, r, num()))
    out.append("RHS")
    for r in rng.sample(["OBJ"] + rows, rng.randint(0, len(rows))):
        out.append(line("", "RHS", r, num()))
    if rng.random() < 0.6:
        out.append("RANGES")
        for r in rng.sample(rows, rng.randint(0, len(rows))):
            out.append(line("", "RNG", r, num()))
    out.append("BOUNDS")
    for _ in range(rng.randint(0, 8)):
        t = rng.choice(["UP", "LO", "FX", "FR", "MI", "PL", "BV", "LI", "UI"])
        c = rng.choice(cols + [name("N")])
        needs = t in ("UP", "LO", "FX", "LI", "UI")
        out.append(line(t, "BND", c, num() if needs else ""))
    out.append("ENDATA")
    return "\n".join(out) + "\n"


@pytest.mark.skipif(not available("ref"), reason="reference CPU sources (oracle/_ref) not built here")
def test_differential_against_the_reference_reader(tmp_path):
    from oracle import read_mps_ref, read_mps_ref_error
    texts = [RULES, fixed_text(), "ROWS\n L  r\nCOLUMNS\n    x  r  1\nENDATA\n",
             "ROWS\n N  o\nCOLUMNS\n    x  o  1\nBOUNDS\n XX BND x  3\nENDATA\n"]
    rng = random.Random(11)
    texts += [random_mps(rng, fixed=k % 3 == 0) for k in range(300)]
    agreed = parsed = 0
    for k, text in enumerate(texts):
        path = write(tmp_path, f"d{k}.mps", text)
        ref_err = read_mps_ref_error(path)
        if ref_err:
            with pytest.raises(P.SupremaError) as e:
                P.read_mps(path)
            assert ("MPS parse failed: " + ref_err) in str(e.value) or ref_err in str(e.value), text
        else:
            same_as_dict(P.read_mps(path), read_mps_ref(path, full=True))
            parsed += 1
        agreed += 1
    assert agreed == len(texts) and parsed > 200


# ------------------------------------- the threaded COLUMNS pass (free format)

def read_threads(monkeypatch, path, threads, min_chunk=1):
    """read_mps with the COLUMNS body cut into `threads` chunks of at least
    min_chunk bytes (threads=1: the one-thread parse); the LP or the error."""
    monkeypatch.setenv("SUPR_MPS_THREADS", str(threads))
    monkeypatch.setenv("SUPR_MPS_MIN_CHUNK", str(min_chunk))
    try:
        return P.read_mps(path)
    except P.SupremaError as e:
        return str(e)
    finally:
        monkeypatch.delenv("SUPR_MPS_THREADS")
        monkeypatch.delenv("SUPR_MPS_MIN_CHUNK")


def same_parse(a, b):
    if isinstance(a, str) or isinstance(b, str):
        return a == b
    return (all(bits(getattr(a.lp, k), getattr(b.lp, k)) for k in ARRAYS)
            and bits(np.float64(a.lp.obj_offset), np.float64(b.lp.obj_offset))
            and bits(a.integer, b.integer) and bits(a.q_row, b.q_row) and bits(a.q_col, b.q_col)
            and bits(a.q_val, b.q_val) and a.maximize == b.maximize and a.col_names == b.col_names
            and a.row_names == b.row_names and a.obj_name == b.obj_name and a.name == b.name)


@pytest.mark.parametrize("threads", [3, 8])
def test_threaded_columns_on_the_fixtures(tmp_path, monkeypatch, threads):
    """Every fixture, its COLUMNS body cut into many chunks, still gives the
    reference reader's LP bit for bit (files with markers or in fixed format
    take the one-thread path inside the reader)."""
    for name in FIXTURES:
        with gzip.open(os.path.join(MPS_DIR, name + ".mps.gz"), "rb") as f:
            path = write(tmp_path, name + ".mps", f.read().decode("latin-1"))
        d = dict(np.load(os.path.join(MPS_DIR, name + ".npz")))
        mp = read_threads(monkeypatch, path, threads)
        assert not isinstance(mp, str), (name, mp)
        same_as_dict(mp, d)


EDGE_HEAD = "NAME  EDGE\nROWS\n N  obj\n L  r1\n G  r2\n E  r3\n N  free\n L  r1\nCOLUMNS\n"
EDGE_TAIL = "RHS\n    RHS  r1  4  r2  1\n    RHS  obj  2.5\nBOUNDS\n UP BND  a  3\n LO BND  z  -1\nENDATA\n"
EDGE_BODIES = {
    # a column's lines far apart, its objective split over them (summed in file order)
    "repeats": "    a  obj  0.1  r1  1\n    b  r2  2\n    c  r3  3\n    a  obj  0.2  r2  4\n"
               "    b  obj  1e-17\n    a  obj  0.3\n    d  r1  -1\n    b  r3  5\n",
    "crlf_comments_blanks": "    a  r1  1\r\n* comment\r\n\r\n   \t \n    b  obj  2\r\n    a  r2  3\r\n",
    "odd_tokens_and_zeros": "    a  r1  1  r2\n    b  r2  0  r3  7\n    c  free  9  r1  0.0\n",
    "keyword_names_not_in_column_1": "    RHS  r1  1\n    BOUNDS  r2  2\n    a  r3  3\n",
    "endata_indented": "    a  r1  1\n    b  r2  2\n    ENDATA\n    c  r3  3\n",
    "name_line_inside": "    a  r1  1\nNAME  SECOND\n    b  r2  2\n",
    "second_columns_section": "    a  r1  1\n    b  r2  2\nRHS\n    RHS  r1  7\nCOLUMNS\n    b  r3  3\n    c  r1  4\n",
    "undeclared_row": "    a  r1  1\n    b  nowhere  2\n    c  r2  3\n",
    "marker_late": "    a  r1  1\n    b  r2  2\n    MARKER  'MARKER'  'INTORG'\n    c  r3  3\n"
                   "    MARKER  'MARKER'  'INTEND'\n    d  r1  4\n",
    "bad_number_late": "    a  r1  1\n    b  r2  2\n    c  r3  x7\n",
    "trailing_text_number": "    a  r1  1.5abc\n    b  r2  0x10\n",
    "out_of_range_number": "    a  r1  1\n    b  r2  1e999\n",
    "vertical_tab_line": "    a  r1  1\n\v\n    b  r2  2\n",
    "no_endata": None,
}


@pytest.mark.parametrize("case", sorted(EDGE_BODIES))
def test_threaded_columns_edge_cases(tmp_path, monkeypatch, case):
    """Each case gives the same LP, names and errors on 2 .. 9 chunks as on
    one thread (cuts land inside a column's run, on comments and on the
    section line), and, where oracle/_ref is built, the reference reader's."""
    body = EDGE_BODIES[case]
    text = EDGE_HEAD + (body if body is not None else "    a  r1  1\n    b  r2  2\n    a  r3  3") \
        + (EDGE_TAIL if body is not None else "")
    path = write(tmp_path, case + ".mps", text)
    one = read_threads(monkeypatch, path, 1)
    for threads in range(2, 10):
        assert same_parse(read_threads(monkeypatch, path, threads), one), (case, threads)
    # the reference reader indexes the first token of a line that has none
    # (src/mps_reader.cpp:175) and crashes on the vertical-tab line: compared
    # with the one-thread parse only
    if available("ref") and case != "vertical_tab_line":
        from oracle import read_mps_ref, read_mps_ref_error
        err = read_mps_ref_error(path)
        if err:
            assert isinstance(one, str) and err in one
        else:
            same_as_dict(one, read_mps_ref(path, full=True))


def random_big_mps(rng):
    """A free-format file with a few thousand COLUMNS lines: columns in
    shuffled runs (a name comes back after others), objective entries in
    several runs of one column, free and twice-declared rows, explicit zeros,
    comments, and now and then a marker or a stray row (both send the
    reader to its one-thread path)."""
    rows = [f"R{i}" for i in range(rng.randint(5, 40))]
    kinds = [rng.choice("LGE") for _ in rows]
    out = ["NAME  BIG", "ROWS", " N  COST"] + [f" {k}  {r}" for k, r in zip(kinds, rows)]
    out += [" N  FREE", f" G  {rows[0]}", "COLUMNS"]
    ncols = rng.randint(50, 400)
    lines = []
    for _ in range(rng.randint(1500, 4000)):
        j = rng.randrange(ncols)
        pairs = []
        for _ in range(rng.choice((1, 2))):
            r = rng.random()
            tgt = "COST" if r < 0.25 else "FREE" if r < 0.3 else rng.choice(rows)
            v = "0" if rng.random() < 0.05 else f"{rng.uniform(-50, 50):.{rng.randint(1, 12)}g}"
            pairs += [tgt, v]
        lines.append(f"    C{j}  " + "  ".join(pairs))
        if rng.random() < 0.01:
            lines.append("* a comment")
    if rng.random() < 0.15:
        lines.insert(rng.randrange(len(lines)), "    MARKER  'MARKER'  'INTORG'")
    if rng.random() < 0.15:
        lines.insert(rng.randrange(len(lines)), "    C1  STRAY  1")
    out += lines + ["RHS"] + [f"    RHS  {r}  {rng.randint(-5, 5)}" for r in rows[:5]]
    out += ["BOUNDS", " UP BND  C0  4", " FR BND  C3", "ENDATA"]
    return "\n".join(out) + "\n"


def test_threaded_columns_random_files(tmp_path, monkeypatch):
    """Random files of a few thousand lines: 2 .. 16 chunks give the
    one-thread parse bit for bit."""
    rng = random.Random(5)
    for k in range(40):
        path = write(tmp_path, f"big{k}.mps", random_big_mps(rng))
        one = read_threads(monkeypatch, path, 1)
        assert not isinstance(one, str), one
        for threads in (2, 5, 16):
            assert same_parse(read_threads(monkeypatch, path, threads, min_chunk=64), one), (k, threads)


def test_file_edge_cases(tmp_path):
    """An empty file, a directory and a file whose last line has no newline
    read as the reference's stream reads them."""
    with pytest.raises(P.SupremaError, match="No objective"):
        P.read_mps(write(tmp_path, "empty.mps", ""))
    with pytest.raises(P.SupremaError, match="No objective"):
        P.read_mps(str(tmp_path))
    mp = P.read_mps(write(tmp_path, "nonl.mps", "ROWS\n N  o\n L  r\nCOLUMNS\n    x  o  1  r  2"))
    assert (mp.lp.m, mp.lp.n, mp.lp.nnz) == (1, 1, 1) and mp.lp.c[0] == 1.0


def test_bulk_names_match_the_per_index_names(tmp_path):
    """supr_mps_names (one call per kind) gives the names supr_mps_name gives
    one index at a time; a short buffer and a bad kind are errors."""
    import ctypes
    from pdlp_amd._lib import lib
    with gzip.open(os.path.join(MPS_DIR, "israel.mps.gz"), "rb") as f:
        path = write(tmp_path, "israel.mps", f.read().decode("latin-1"))
    h = ctypes.c_void_p()
    P.check(lib().supr_mps_read(os.fsencode(path), ctypes.byref(h)))
    try:
        dims = [ctypes.c_int32() for _ in range(5)]
        P.check(lib().supr_mps_sizes(h, *[ctypes.byref(d) for d in dims]))
        m, n = dims[0].value, dims[1].value
        for kind, k in ((1, n), (2, m)):
            one = [lib().supr_mps_name(h, kind, j).decode() for j in range(k)]
            assert P._mps_names(h, kind) == one
            nb = ctypes.c_int64()
            P.check(lib().supr_mps_names(h, kind, None, 0, ctypes.byref(nb)))
            assert nb.value == sum(len(x) + 1 for x in one)
            short = ctypes.create_string_buffer(nb.value - 1)
            assert lib().supr_mps_names(h, kind, short, nb.value - 1, ctypes.byref(nb)) == -1
        assert lib().supr_mps_names(h, 3, None, 0, ctypes.byref(nb)) == -1
    finally:
        lib().supr_mps_free(h)
    d = dict(np.load(os.path.join(MPS_DIR, "israel.npz")))
    assert P.read_mps(path).col_names == [str(v) for v in d["col_names"]]
