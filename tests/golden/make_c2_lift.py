"""FP64 reference lift/drag record of the TRT + Smagorinsky cylinder case at Re 300.

Runs the C oracle (FP64) for the whole case (256x128x128, 20000 steps, tests/configs.py C2) and
samples the monitors every 10 steps, the way collision_check.py:14-22 does every 50: 9 quiet
steps, one report step, ``monitors(10)``.  Writes tests/golden/c2_trt_re300_lift.npz (step,
drag, lift) so the GPU run of the same case can be compared without re-running the oracle
(about two hours on 8 cores).

Run:  OMP_NUM_THREADS=8 python tests/golden/make_c2_lift.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from configs import C2, C2_SAMPLE_EVERY, C2_STEPS  # noqa: E402
from oracle import c_oracle  # noqa: E402


def main():
    cfg = dict(C2)
    s = c_oracle.CSolver(**cfg)
    steps, drag, lift = [], [], []
    t0 = time.time()
    out = os.path.join(HERE, "c2_trt_re300_lift.npz")
    while s.step_count < C2_STEPS:
        for _ in range(C2_SAMPLE_EVERY - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(C2_SAMPLE_EVERY)
        steps.append(s.step_count)
        drag.append(m["drag"])
        lift.append(m["lift"])
        if s.step_count % 1000 == 0:
            print(f"step {s.step_count} drag {m['drag']:.5f} lift {m['lift']:+.5f} "
                  f"({time.time() - t0:.0f} s)", flush=True)
            np.savez(out + ".partial.npz", step=np.array(steps), drag=np.array(drag),
                     lift=np.array(lift))
    np.savez(out, step=np.array(steps), drag=np.array(drag), lift=np.array(lift),
             sample_every=np.array(C2_SAMPLE_EVERY))
    os.remove(out + ".partial.npz")
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
