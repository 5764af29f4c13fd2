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
    # that two XCDs both fetch (3.4% at FP32, 0.1% at FP16 and BF16 with pairs of blocks per XCD)
    assert np.all(got >= algorithmic)
    assert np.all(got <= 1.10 * algorithmic)
    assert bench.profiled_traffic("f64") is None


def test_variants_warm_up_past_the_clock_transient(monkeypatch):
    """The headline storage keeps the warm-up it is given; the 16-bit variants take at least
    VARIANT_WARMUP steps (the engine-clock dip after a kernel switch), and the JSON line reports
    each variant's warm-up."""
    import bench
    calls = []

    def fake_time_gpu(storage, steps, warmup, block_x=0):
        calls.append((storage, steps, warmup))
        return {"storage": storage, "warmup": warmup, "mlups": 1.0, "ms_per_step": 1.0,
                "wall_ms_per_step": 1.0, "quiet_kernel_ms": 1.0, "quiet_launches": steps,
                "monitors": {"drag": 0.0, "lift": 0.0}, "achieved_gbs": 1.0, "ceiling_ms": 1.0,
                "ceiling_gbs": 1.0}

    monkeypatch.setattr(bench, "time_gpu", fake_time_gpu)
    monkeypatch.setattr(sys, "argv", ["bench.py", "--steps", "20", "--warmup", "5", "--no-cpu",
                                      "--no-sweep", "--no-adjoint", "--no-derived", "--no-heat",
                                      "--no-loop", "--storage", "f32,f16,bf16"])
    import io
    import json
    import contextlib
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        bench.main()
    assert calls == [("f32", 20, 5), ("f16", 20, bench.VARIANT_WARMUP),
                     ("bf16", 20, bench.VARIANT_WARMUP)]
    line = json.loads(out.getvalue().strip().splitlines()[-1])
    assert line["warmup"] == 5 and line["steps"] == 20
    assert {k: v["warmup"] for k, v in line["variants"].items()} == \
        {"f32": 5, "f16": bench.VARIANT_WARMUP, "bf16": bench.VARIANT_WARMUP}


def test_cpu_baseline_slab_steps_the_workload():
    import bench
    cells, seconds = bench._cpu_slab(0.2, nz=4)
    assert cells > 0 and cells % (512 * 256 * 4) == 0 and seconds > 0
