"""bench.py's restatements of the engine's launch rules (which kernel form a
launch takes, so that the bench and its profiles name the kernel rocprof
reports) and the committed profiles it reads them from.  CPU only."""
import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_rules_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


B = _bench()


def test_dual_form_follows_the_engine_rule():
    # the 1M-variable transportation LP's A: 2,000 CSR rows of 1,000 at 32 lanes
    # (250 blocks): the flat one-pass form
    assert B.dual_form(2000, 2_000_000, 32, False) == 2
    # transportation 200 x 1,000: 1,200 rows of 200 and 1,000 (150 blocks)
    assert B.dual_form(1200, 400_000, 32, False) == 2
    # the random 2,000 x 5,000 LP's A: 15 entries per row, one prefetched group
    assert B.dual_form(2000, 30_000, 32, False) == 1
    # the power-law LP's A: 500,000 rows, LDS-staged tiles, grid-stride
    assert B.dual_form(500_000, 35_118_140, 32, False) == 0
    # long rows on more than 1,024 blocks: grid-stride
    assert B.dual_form(12_000, 12_000_000, 32, False) == 0
    # more rows than one pass holds: grid-stride
    assert B.dual_form(40_000, 40_000_000, 32, False) == 0
    # long ELL rows never take the flat form
    assert B.dual_form(2000, 2_000_000, 32, True) == 0


def test_primal_kernel_follows_the_engine_rule():
    # the 1M-variable LP's A^T: ELL, one lane per row, 1,000,000 rows
    assert B.primal_kernel(1_000_000, 1, True, True) == "k_primal_pairs<double>"
    assert B.primal_kernel(1_000_000, 1, True, False) == "k_primal<1, float, false>"
    # more rows than one pass of pairs holds: k_primal's grid-stride path
    assert B.primal_kernel(1_200_000, 1, True, True) == "k_primal<1, double, false>"
    # an odd row count or a CSR matrix: no pairs
    assert B.primal_kernel(999_999, 1, True, True) == "k_primal<1, double, false>"
    assert B.primal_kernel(1_000_000, 1, False, True) == "k_primal<1, double, false>"
    # the power-law LP's A^T (CSR, 8 lanes)
    assert B.primal_kernel(2_000_000, 8, False, True) == "k_primal<8, double, false>"


def test_committed_profiles_name_the_kernels_the_bench_reads():
    # roofline.rocprof_avg_us and roofline.traffic come from these files
    name = B.primal_kernel(1_000_000, 1, True, True)
    us = B.rocprof_average_us(name)
    assert us is not None and 5.0 < us < 20.0, (name, us)
    dual = f"k_dual<32, double, {B.dual_form(2000, 2_000_000, 32, False)}>"
    assert B.rocprof_average_us(dual) is not None
    text = open(os.path.join(ROOT, "profiles", "hbm_traffic_transport_1000x1000.md")).read()
    block = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    assert block["kernels"]["k_primal_f64"]["kernel"] == name
    assert block["kernels"]["k_dual_f64"]["kernel"] == dual
