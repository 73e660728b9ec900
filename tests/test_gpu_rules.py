"""GPU tests of the engine's device-side scalar rules against the oracle.

The device-resident loop takes its step decisions (decide_step), KKT-check
decisions (kkt_fin_update) and primal-weight updates (restart_fin_update) on
the device (team-suprema_amd/csrc/pdlp_device.hpp); the single-LP and the
batched kernels call the same functions.  tests/hip/librule_probe.so runs
them on scripted inputs, one case per thread, and these tests compare each
result with the oracle's orc_step_rule / orc_check_rule / orc_primal_weight
(include/pdlp_algo.hpp:94-100, 113-131, 135-167), which the golden
trajectories pin to the reference bit for bit.

Tolerances: every decision (accept / reject, converged, candidate, new best,
restart) is compared exactly, except where the oracle's own margin is within
1e-14 relative (device code may fuse a multiply-add that the oracle rounds
twice: the operands can then differ by an ulp, and a comparison that close to
its threshold may go either way); step sizes and merits within 4 ulp (pow
differs in the last bits between the device library and glibc); primal
weights within 4 ulp times the size of exp's argument (exp turns an ulp of its
argument into that many ulps of its result); eps exactly (sqrt and division
are correctly rounded on both sides).
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
ULP4 = 4 * 2.0 ** -52


def say(*a):
    print("[rules]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def probe():
    path = os.path.join(HERE, "hip", "librule_probe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for f in (lib.probe_step_rule, lib.probe_check_rule, lib.probe_primal_weight, lib.probe_lane_sums):
        f.restype = ctypes.c_int
    return lib


def d(a):
    return a.ctypes.data_as(_dp)


def i(a):
    return a.ctypes.data_as(_ip)


def close(a, b, rel=ULP4):
    if a == b:
        return True
    if not (math.isfinite(a) and math.isfinite(b)):
        return False
    return abs(a - b) <= rel * max(abs(a), abs(b))


def step_cases(n=4000, seed=3):
    rng = np.random.default_rng(seed)
    eta = 10.0 ** rng.uniform(-4, 2, n)
    omega = 10.0 ** rng.uniform(-3, 3, n)
    acc = rng.integers(0, 200000, n).astype(float)
    acc[:50] = np.arange(50)
    dx2 = 10.0 ** rng.uniform(-20, 4, n)
    dy2 = 10.0 ** rng.uniform(-20, 4, n)
    inter = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-20, 4, n)
    dx2[rng.random(n) < 0.05] = 0.0
    dy2[rng.random(n) < 0.05] = 0.0
    inter[rng.random(n) < 0.05] = 0.0
    inter[rng.random(n) < 0.02] = -0.0
    # eta_bar = (omega dx2 + dy2 / omega) / (2 |inter|) = 1 exactly; eta = 1 is a tie (accepted)
    for k, (e, s) in enumerate([(1.0, 5.0), (1.0, -5.0), (0.999, 5.0), (1.001, 5.0)]):
        eta[100 + k], omega[100 + k], dx2[100 + k], dy2[100 + k], inter[100 + k] = e, 2.0, 3.0, 8.0, s
    # eta near eta_bar on purpose: rescale eta to eta_bar * (1 +- small)
    for k in range(200, 400):
        num = omega[k] * dx2[k] + dy2[k] / omega[k]
        if inter[k] != 0:
            eta[k] = num / (2 * abs(inter[k])) * (1 + rng.choice([-1, 1]) * 10.0 ** rng.uniform(-12, -3))
    return np.stack([eta, omega, acc, dx2, dy2, inter], axis=1)


def test_step_rule_matches_oracle(probe, oracle):
    cases = step_cases()
    n = len(cases)
    flat = np.ascontiguousarray(cases.reshape(-1))
    outs = []
    for pre in (0, 1):
        out = np.zeros(2 * n)
        pres = np.full(n, pre, dtype=np.int32)
        assert probe.probe_step_rule(n, d(flat), i(pres), d(out)) == 0
        outs.append(out.reshape(n, 2))
    # the published step powers give the same bits as the in-line pow
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    dev = outs[0]
    ties = 0
    for k in range(n):
        eta, omega, acc, dx2, dy2, inter = cases[k]
        a, eb, en = oracle.step_rule(eta, omega, int(acc), dx2, dy2, inter)
        if math.isfinite(eb) and abs(eta - eb) <= 1e-14 * eb and k not in (100, 101):
            ties += 1
            continue
        assert bool(dev[k, 1]) == a, (k, cases[k], eb)
        assert close(dev[k, 0], en), (k, cases[k], dev[k, 0], en)
    # the constructed exact tie is accepted, its neighbours are not / are
    assert dev[100, 1] == 1 and dev[101, 1] == 1 and dev[102, 1] == 1 and dev[103, 1] == 0
    say(f"step rule: {n} cases, {ties} near-ties skipped, all decisions equal the oracle's")


def check_cases(n=3000, seed=5):
    rng = np.random.default_rng(seed)
    rows = []
    iin = []
    for k in range(n):
        sn = np.zeros(8)
        sm = np.zeros(6)
        denom_p = 1 + 10.0 ** rng.uniform(0, 3)
        denom_d = 1 + 10.0 ** rng.uniform(0, 3)
        for c in range(2):
            scale = 10.0 ** rng.uniform(-8, -1)
            ep, ed = scale * 10.0 ** rng.uniform(-2, 1), scale * 10.0 ** rng.uniform(-3, 1)
            f = rng.uniform(0, 1)
            sn[4 * c + 0] = (ep * denom_p) ** 2 * f          # bound violations^2
            sm[3 * c + 0] = (ep * denom_p) ** 2 * (1 - f)    # row violations^2
            g = rng.uniform(0, 1)
            sn[4 * c + 1] = (ed * denom_d) ** 2 * g
            sm[3 * c + 1] = (ed * denom_d) ** 2 * (1 - g)
            cx = rng.uniform(-1e4, 1e4)
            gap = scale * 10.0 ** rng.uniform(-2, 1) * (1 + 2 * abs(cx))
            D = cx - rng.choice([-1, 1]) * gap
            h = rng.uniform(0, 1)
            sn[4 * c + 2], sm[3 * c + 2], sn[4 * c + 3] = D * h, D * (1 - h), cx
            if rng.random() < 0.03:
                sn[4 * c + 2] = -math.inf        # a multiplier on an infinite bound
        tol = rng.choice([1e-4, 1e-6, 1e-8])
        best = rng.choice([math.inf, 10.0 ** rng.uniform(-9, -1)])
        mlr = rng.choice([math.inf, 10.0 ** rng.uniform(-9, -1)])
        mpc = rng.choice([math.inf, 10.0 ** rng.uniform(-9, -1)])
        rows.append(np.concatenate([sn, sm, [denom_p, denom_d, tol, best, mlr, mpc, 0.5]]))
        accepted = int(rng.integers(64, 100000))
        iin += [int(rng.integers(0, accepted)), accepted]
    return np.array(rows), np.array(iin, dtype=np.int32)


def eps_of(row, c):
    sn, sm = row[:8], row[8:14]
    denom_p, denom_d = row[14], row[15]
    bv2, svc, dcol, cx = sn[4 * c: 4 * c + 4]
    rv2, svr, drow = sm[3 * c: 3 * c + 3]
    D = drow + dcol
    eg = abs(cx - D) / (1.0 + abs(cx) + abs(D)) if math.isfinite(D) else math.inf
    return [math.sqrt(rv2 + bv2) / denom_p, math.sqrt(svr + svc) / denom_d, eg], cx + row[20]


def test_check_rule_matches_oracle(probe, oracle):
    rows, iin = check_cases()
    n = len(rows)
    out = np.zeros(9 * n)
    eps = np.zeros(6 * n)
    assert probe.probe_check_rule(n, d(np.ascontiguousarray(rows.reshape(-1))), i(iin), d(out), d(eps)) == 0
    out, eps = out.reshape(n, 9), eps.reshape(n, 6)
    ties = counts = 0
    seen = {"converged": 0, "restart": 0, "avg": 0, "best": 0}
    for k in range(n):
        rc, obj_c = eps_of(rows[k], 0)
        ra, obj_a = eps_of(rows[k], 1)
        tol, best, mlr, mpc = rows[k][16:20]
        since, acc = int(iin[2 * k]), int(iin[2 * k + 1])
        r = oracle.check_rule(rc, ra, tol, since, acc, mlr, best, mpc)
        m_c = math.sqrt(rc[0] ** 2 + rc[1] ** 2 + rc[2] ** 2)
        m_a = math.sqrt(ra[0] ** 2 + ra[1] ** 2 + ra[2] ** 2)
        mc = r["cand_merit"]
        near = [(m_c, best), (m_a, min(best, m_c)), (m_a, m_c), (mc, 0.2 * mlr), (mc, 0.8 * mlr),
                (mc, mpc), (max(rc if not r["use_avg"] else ra), tol)]
        if any(math.isfinite(a) and math.isfinite(b) and abs(a - b) <= 1e-14 * max(abs(a), abs(b))
               for a, b in near):
            ties += 1
            continue
        counts += 1
        cand = ra if r["use_avg"] else rc
        assert list(eps[k, :3]) == cand, (k, eps[k, :3], cand)          # eps bit for bit
        assert eps[k, 3] == (obj_a if r["use_avg"] else obj_c)
        assert bool(out[k, 0] == 1) == r["converged"], k
        assert bool(out[k, 2]) == r["use_avg"], k
        assert close(out[k, 4], mc), (k, out[k, 4], mc)
        if r["converged"]:
            seen["converged"] += 1
            continue
        assert int(out[k, 1]) == r["best_which"], (k, out[k, 1], r)
        assert bool(out[k, 3]) == r["restart"], (k, r)
        assert close(out[k, 5], r["best_merit"]) and close(out[k, 6], r["merit_prev_cand"]), k
        seen["restart"] += r["restart"]
        seen["avg"] += r["use_avg"]
        seen["best"] += r["best_which"] > 0
    assert counts > 0.95 * n and all(v > 20 for v in seen.values()), seen
    say(f"check rule: {counts} cases compared ({ties} near-ties skipped), branches {seen}")


def test_primal_weight_matches_oracle(probe, oracle):
    rng = np.random.default_rng(9)
    n = 2000
    omega = 10.0 ** rng.uniform(-3, 3, n)
    dx2 = 10.0 ** rng.uniform(-30, 6, n)
    dy2 = 10.0 ** rng.uniform(-30, 6, n)
    flat = np.ascontiguousarray(np.stack([omega, dx2, dy2], axis=1).reshape(-1))
    out = np.zeros(n)
    assert probe.probe_primal_weight(n, d(flat), d(out)) == 0
    unchanged = 0
    for k in range(n):
        want = oracle.primal_weight(omega[k], math.sqrt(dx2[k]), math.sqrt(dy2[k]))
        if want == omega[k]:
            unchanged += 1
            assert out[k] == omega[k], k
        else:
            # exp(0.5 log(ddy / ddx) + 0.5 log omega): an ulp of the exponent is
            # a relative error of |exponent's terms| ulps in the result
            q = math.sqrt(dy2[k]) / math.sqrt(dx2[k])
            rel = ULP4 * (1 + 0.5 * abs(math.log(q)) + 0.5 * abs(math.log(omega[k])))
            assert close(out[k], want, rel), (k, out[k], want)
    assert 0 < unchanged < n
    say(f"primal weight: {n} cases ({unchanged} below the 1e-10 distance floor)")


def test_lane_sums_match_shuffles_bit_for_bit(probe):
    """The kernels' cross-lane sums (DPP / permlane moves, no LDS) add the
    same values in the same order as the __shfl_xor butterfly, for every
    group width 2..64 and every lane, in FP64 and FP32; the wavefront sum in
    lane 0 equals the __shfl_down tree.  Inputs span 40 decades with mixed
    signs, so any change of pairing shows in the low bits."""
    rng = np.random.default_rng(11)
    n = 256 * 64
    x = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-20, 20, n)
    x[rng.random(n) < 0.02] = 0.0
    x[rng.random(n) < 0.01] = -0.0
    xf = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-15, 15, n)).astype(np.float32)
    out = np.zeros(14 * n)
    outf = np.zeros(12 * n, dtype=np.float32)
    assert probe.probe_lane_sums(n, d(x), xf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), d(out),
                                 outf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    o = out.reshape(n, 14)
    of = outf.reshape(n, 12)
    for w in range(6):
        assert np.array_equal(o[:, 2 * w].view(np.uint64), o[:, 2 * w + 1].view(np.uint64)), 2 << w
        assert np.array_equal(of[:, 2 * w].view(np.uint32), of[:, 2 * w + 1].view(np.uint32)), 2 << w
    lane0 = o[::64]
    assert np.array_equal(lane0[:, 12].view(np.uint64), lane0[:, 13].view(np.uint64))
    # and the sums are sums: the 64-wide result against a sorted-order sum
    ref = np.array([math.fsum(x[k:k + 64]) for k in range(0, n, 64)])
    assert np.allclose(lane0[:, 12], ref, rtol=1e-12, atol=1e-12 * np.abs(x).reshape(-1, 64).sum(1).max())
    say(f"cross-lane sums: {n} values, widths 2-64, FP64 and FP32 identical to the shuffles")
