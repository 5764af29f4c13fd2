"""The workload configurations of the parity and physics checks, shared by tests and fixture
generators (each is a keyword set for ``Solver`` / ``OracleSolver``)."""

# Plain method: D3Q19 BGK 96x48x48 channel, cylinder (axis z) radius 6 at x = 24, y = 24,
# u0 = 0.08, Re 100 (D = 12), seeded U(+-1e-3) velocity noise, sponge over the last 8 planes,
# 1000 steps.
C1 = dict(nx=96, ny=48, nz=48, collision="bgk", obstacle="cylinder", reynolds=100.0, u0=0.08,
          diameter=12.0, center=(24.0, 24.0, 24.0), sponge_start=88.0, noise_amp=1e-3,
          noise_seed=0)
C1_STEPS = 1000

# TRT + Smagorinsky (Cs 0.1, magic 1/4), 256x128x128, cylinder radius 12 at the reference's
# default placement, u0 = 0.08, Re 300, seeded noise, 20000 steps (Strouhal from the lift).
C2 = dict(nx=256, ny=128, nz=128, collision="trt", obstacle="cylinder", reynolds=300.0,
          u0=0.08, diameter=24.0, cs=0.1, magic=0.25, noise_amp=1e-3, noise_seed=0)
C2_STEPS = 20000
C2_SAMPLE_EVERY = 10

# LES 512x256x256, sphere radius 16 at x = 96, moving lid u = 0.04 on a free-slip floor,
# Re 3000 (D = 32), report every 100 steps (the headline benchmark workload).
C3 = dict(nx=512, ny=256, nz=256, collision="les", obstacle="sphere", reynolds=3000.0, u0=0.08,
          diameter=32.0, center=(96.0, 128.0, 128.0), lateral="lid", lid_u=(0.04, 0.0, 0.0),
          noise_amp=1e-3, noise_seed=0)

# Adjoint: BGK 128x64x64 FP32, sphere radius 10 with a smoothed surface (eps 1.5) and a 32^3
# porosity block around the nominal centre (32, 32, 32), Re 200 (nu from the nominal D = 20),
# u0 = 0.08, 2000 steps from the uniform inflow state, checkpoints every 100; the objective is
# the mean x force over the last 1000 steps.
C5 = dict(nx=128, ny=64, nz=64, reynolds=200.0, u0=0.08, radius=10.0,
          nominal_center=(32.0, 32.0, 32.0), block=32, steps=2000, window=1000,
          checkpoint_every=100, eps=1.5)
C5_SEED = 0
C5_FD_STEP = 1e-3


def c5_design(seed=C5_SEED, block=32):
    """The seeded design point: radius 10, centre = nominal + U(-2, 2)^3 cells, porosity
    U(0.01, 0.02) on the block (strictly positive, so that phi > 0 on every design cell)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    center = np.array(C5["nominal_center"]) + rng.uniform(-2.0, 2.0, 3)
    porosity = 0.01 + 0.01 * rng.uniform(size=(block, block, block))
    return C5["radius"], center, porosity

# collision_check.py's Strouhal case: the default Solver (lbm.py:548-561: 384x128x64, D = 24,
# u0 = 0.08, cylinder at (0.22 nx, 0.5 ny + 1.7), sponge from 0.75 nx), TRT at Re 300, monitors
# every 50 steps (collision_check.py:12-20); the published St = 0.200 (README.md:66-67).
ST = dict(nx=384, ny=128, nz=64, collision="trt", obstacle="cylinder", reynolds=300.0, u0=0.08,
          diameter=24.0)
ST_EVERY = 50
ST_STEPS = 16000
ST_WINDOW = (8000, 16000)  # shedding saturated
