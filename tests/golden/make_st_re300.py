"""FP64 reference lift/drag record of collision_check.py's Strouhal case: the default Solver
(384x128x64 channel, cylinder D = 24 at the default placement, u0 = 0.08, sponge from 0.75 nx)
at Re 300 with TRT, 16000 steps, monitors every 50 steps (collision_check.py:12-20).

Runs the C oracle (FP64) and writes tests/golden/st_trt_re300_lift.npz (step, drag, lift), so
that the GPU run of the same case can be compared (Strouhal number and mean drag) without
re-running the oracle (a few hours on 8 cores).

Run:  OMP_NUM_THREADS=8 python tests/golden/make_st_re300.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from configs import ST, ST_EVERY, ST_STEPS  # noqa: E402
from oracle import c_oracle  # noqa: E402


def main():
    s = c_oracle.CSolver(**ST)
    steps, drag, lift = [], [], []
    t0 = time.time()
    out = os.path.join(HERE, "st_trt_re300_lift.npz")
    while s.step_count < ST_STEPS:
        for _ in range(ST_EVERY - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(ST_EVERY)
        steps.append(s.step_count)
        drag.append(m["drag"])
        lift.append(m["lift"])
        if s.step_count % 1000 == 0:
            print(f"step {s.step_count} drag {m['drag']:.5f} lift {m['lift']:+.5f} "
                  f"({time.time() - t0:.0f} s)", flush=True)
            np.savez(out + ".partial.npz", step=np.array(steps), drag=np.array(drag),
                     lift=np.array(lift))
    np.savez(out, step=np.array(steps), drag=np.array(drag), lift=np.array(lift))
    os.remove(out + ".partial.npz")
    print("wrote", out)


if __name__ == "__main__":
    main()
