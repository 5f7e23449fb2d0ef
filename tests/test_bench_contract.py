"""bench.py's contract on the CPU: it refuses more than one GPU with a message, and the committed
HBM-traffic summary it reads for the roofline holds a line for every threads-per-block variant of the
FMG-level-0 relaxation that the autotuner can pick."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("nt", [64, 128, 256])
def test_traffic_line_for_every_tuned_shape(nt):
    import bench
    v = bench._profile_traffic("amr_c2_hbm_traffic.md", f"sg::amr::k_blk8<true, float, {nt}, true>")
    # 9 437 184 leaf cells x 6 B algorithmic per launch; the measured traffic sits above it
    assert v is not None and 56_623_104 < v < 4 * 56_623_104
