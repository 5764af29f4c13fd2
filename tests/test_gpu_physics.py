"""Whole-run physics of the HIP step against the FP64 C oracle.

Each case runs the same configuration on the GPU (FP32, FP16 and BF16 population storage, FP32
arithmetic) and in the FP64 oracle, sampling the monitors every `every` steps the way
collision_check.py:14-22 does, and compares the drag/lift records and the final velocity field.
The errors are printed (pytest -s) with the bound each is held to.
"""

import os

import numpy as np
import pytest

from configs import C1, C1_STEPS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle():
    from oracle import c_oracle
    if not c_oracle.available():
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    return c_oracle


def _run_gpu(cfg, steps, every, storage):
    import tlbm
    kw = dict(cfg)
    s = tlbm.Solver(**kw, storage=storage)
    drag, lift = [], []
    while s.step_count < steps:
        for _ in range(every - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(every)
        drag.append(m["drag"])
        lift.append(m["lift"])
    tlbm.sync()
    u = np.stack([s.fields[k].numpy() for k in ("ux", "uy", "uz")]).astype(np.float64)
    return np.array(drag), np.array(lift), u


def _run_oracle(cfg, steps, every):
    C = _oracle()
    s = C.CSolver(**cfg)
    drag, lift = [], []
    while s.step_count < steps:
        for _ in range(every - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(every)
        drag.append(m["drag"])
        lift.append(m["lift"])
    u = np.stack([s.fields[k] for k in ("ux", "uy", "uz")])
    return np.array(drag), np.array(lift), u


_C1_REF = {}


def _c1_reference():
    if not _C1_REF:
        _C1_REF["r"] = _run_oracle(C1, C1_STEPS, 50)
    return _C1_REF["r"]


# bounds per storage: (max |du| / u0, rms |du| / u0, max |d drag| / mean drag, max |d lift| /
# mean drag), about 3x the errors measured on MI355X (f32 3.4e-5, 1.2e-5, 1.7e-5, 2.4e-6;
# f16 1.3e-3, 3.0e-4, 5.0e-4, 3.2e-4; bf16 1.1e-2, 2.8e-3, 8.3e-3, 2.1e-3)
C1_BOUNDS = {"f32": (1e-4, 4e-5, 6e-5, 1e-5), "f16": (4e-3, 1e-3, 2e-3, 1e-3),
             "bf16": (4e-2, 1e-2, 2.5e-2, 7e-3)}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c1_bgk_cylinder_against_fp64_oracle(storage):
    """BGK 96x48x48, cylinder r = 6 at (24, 24), Re 100, seeded noise, 1000 steps."""
    d_ref, l_ref, u_ref = _c1_reference()
    d, l, u = _run_gpu(C1, C1_STEPS, 50, storage)
    u0 = C1["u0"]
    du = np.sqrt(((u - u_ref) ** 2).sum(0))
    err = (du.max() / u0, np.sqrt((du ** 2).mean()) / u0,
           np.abs(d - d_ref).max() / abs(d_ref.mean()), np.abs(l - l_ref).max() / abs(d_ref.mean()))
    print(f"\nC1 {storage}: max|du|/u0 {err[0]:.2e}  rms|du|/u0 {err[1]:.2e}  "
          f"drag {d[-1]:.5f} (oracle {d_ref[-1]:.5f}) max rel {err[2]:.2e}  lift max {err[3]:.2e}")
    for e, b in zip(err, C1_BOUNDS[storage]):
        assert e <= b, (err, C1_BOUNDS[storage])
