"""bench.py's contract pieces that need no GPU: the one-GPU guard, the
roofline arithmetic of the dominant kernel, and the traffic lookup in the
committed counter tables."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


@pytest.fixture(scope="module")
def bench():
    import bench as b
    return b


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_roofline_of_dominant_kernel(bench):
    n, cells, cand, inr = 50_000, 140_608, 270 * 50_000, 140 * 50_000
    stages = {k: 0.01 for k in bench.STAGE_BYTES}
    kernels = {"density": (("k_density<16,false>",), ("density",), 0.036),
               "xsph+force": (("k_xsph_force<32,false>",), ("xsph", "force"), 0.046)}
    roof, per_stage = bench.roofline_for(stages, n, cells, 4, cand, inr, (16, 8, 0), kernels)
    assert roof["kernel"] == "k_xsph_force<32,false>" and roof["stages"] == "xsph+force"
    alg = (44 + 72) * n + (8 + 8) * cells
    assert roof["alg_bytes_per_launch"] == alg
    assert roof["achieved"] == pytest.approx(alg / 0.046e-3 / 1e9, rel=1e-3)
    assert roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"], rel=1e-2)
    flops = (16 + 68) * inr
    assert roof["valu"]["alg_flops_per_launch"] == flops
    assert roof["valu"]["achieved"] == pytest.approx(flops / 0.046e-3 / 1e12, rel=1e-2)
    assert roof["launch_ms"] == 0.046
    # density is charged its candidate distance tests, the list walks are not
    d = roof["kernels"]["density"]
    assert d["alg_flops_per_launch"] == 24 * inr + 8 * cand
    assert set(per_stage) == set(stages)


@pytest.mark.parametrize("workload,kernel", [("dam_break_50k", "k_xsph_force<16,false>"),
                                             ("dam_break_50k", "k_density<16,false>"),
                                             ("block_1m", "k_xsph_force<16,false>"),
                                             ("block_4m", "k_xsph_force<8,false>")])
def test_traffic_tables_have_the_step_kernels(bench, workload, kernel):
    t = bench.profiled_traffic(kernel, workload)
    assert isinstance(t, int) and t > 0
