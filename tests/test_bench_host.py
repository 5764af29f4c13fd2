"""bench.py's host side, without a GPU: the one-GPU refusal, the committed HBM-traffic profile
it reports as roofline.traffic, and the CPU baseline's slab step."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_comes_from_the_committed_counter_profile():
    import bench
    got = np.array([bench.profiled_traffic(s) for s in ("f32", "f16", "bf16")])
    algorithmic = 512 * 256 * 256 * np.array([152, 76, 76])
    # measured HBM bytes per quiet launch: the algorithmic bytes plus the block-edge cache lines
    # that two XCDs both fetch (3.4% at FP32, 6.7% at FP16 and BF16 when measured)
    assert np.all(got >= algorithmic)
    assert np.all(got <= 1.10 * algorithmic)
    assert bench.profiled_traffic("f64") is None


def test_cpu_baseline_slab_steps_the_workload():
    import bench
    cells, seconds = bench._cpu_slab(0.2, nz=4)
    assert cells > 0 and cells % (512 * 256 * 4) == 0 and seconds > 0
