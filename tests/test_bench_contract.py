"""CPU: the parts of bench.py's contract that need no GPU.

bench.py measures one GPU and refuses --gpus N > 1 before touching a device;
the figures it quotes from the committed rocprofv3 summaries (VALU issue
fraction, HBM traffic per launch, executed VALU FLOP per launch) must be present in those summaries.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("name,key", [
    ("aak_likelihood_N2p21_n50_sq_counters.md", "valu_issue_frac modesum_stage_mixed"),
    ("aak_likelihood_N2p21_n50_hbm_counters.md", "traffic_bytes_per_launch modesum_stage_mixed"),
    ("aak_likelihood_N2p21_n50_hbm_counters.md", "traffic_bytes_per_launch fold_stage"),
    ("aak_likelihood_N2p21_n50_hbm_counters.md", "traffic_bytes_per_launch k_rates"),
    ("aak_likelihood_N2p21_n50_flops.md", "valu_flop_per_launch modesum_stage_mixed"),
])
def test_profile_figures_the_bench_quotes_are_committed(name, key):
    import bench
    v = bench.read_profile_value(name, key)
    assert v is not None and v > 0
