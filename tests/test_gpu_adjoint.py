"""The adjoint kernel (tlbm_adjoint_step) and its host driver against the FP64 PyTorch reference
(oracle/lbm_torch.py, autograd) and central finite differences.

The GPU forward and adjoint run in FP32; the reference runs in FP64 from the same FP32-rounded
solid fraction and initial state.  Tolerances are relative to the largest gradient component
(FP32 rounding accumulated over the run, a few 1e-6 here) and, for the full workload, the
1e-3 relative bar against finite differences.
"""

import os

import numpy as np
import pytest
import torch

from cases import random_populations
from configs import C5, C5_FD_STEP, c5_design
from oracle import lbm_torch as T

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _reference(dims, phi32, f0_32, steps, window, omega, u0, s0, sk):
    nx, ny, nz = dims
    p = torch.tensor(phi32.astype(np.float64).reshape(nz, ny, nx), requires_grad=True)
    f0 = None if f0_32 is None else T.from_numpy_populations(f0_32.astype(np.float64), dims)
    if f0 is None:
        f0 = T.equilibrium_state(dims, u0)
    J = T.mean_drag(p, f0, steps, window, omega, u0, s0, sk)
    J.backward()
    return float(J.detach()), p.grad.numpy()


def _random_case(dims, seed):
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.02, 0.7, n).astype(np.float32)
    f0 = random_populations(n, seed + 1).astype(np.float32)
    return phi, f0


# rows of 64 cells (waves within a row) and shorter ones (waves straddle rows)
@pytest.mark.parametrize("dims,steps,window,k", [((12, 9, 8), 24, 10, 5), ((7, 5, 6), 13, 13, 4),
                                                 ((33, 4, 3), 9, 3, 100), ((64, 4, 3), 9, 4, 5)])
def test_adjoint_matches_fp64_autograd_on_every_cell(dims, steps, window, k):
    """Full-lattice block: dJ/dphi of every cell, with the sponge overlapping part of the
    solid fraction (both branches of sigma = max(phi, k t^2)) and x, y, z wrapping."""
    from tlbm.adjoint import SolidFractionAdjoint
    phi, f0 = _random_case(dims, seed=dims[0])
    nu, u0, s0, sk = 0.02, 0.06, dims[0] - 4.0, 0.7
    print(f"case dims={dims} steps={steps} window={window} k={k}", flush=True)
    run = SolidFractionAdjoint(*dims, nu, u0, steps, window, k, (0, 0, 0), dims, s0, sk)
    J = run.forward(phi, f0)
    g = run.backward()
    run.close()
    omega = float(np.float32(1.0 / (3.0 * nu + 0.5)))
    J_ref, g_ref = _reference(dims, phi, f0, steps, window, omega, u0, s0, sk)
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max() / scale
    print(f"J {J:.9e} ref {J_ref:.9e}; max |g - g_ref| / max|g_ref| = {err:.2e}")
    assert abs(J - J_ref) <= 2e-5 * max(abs(J_ref), 1e-3)
    assert err < 2e-5


def test_adjoint_block_is_a_window_of_the_full_gradient():
    from tlbm.adjoint import SolidFractionAdjoint
    dims = (14, 10, 9)
    phi, f0 = _random_case(dims, seed=21)
    args = (0.015, 0.05, 20, 8, 6)
    full = SolidFractionAdjoint(*dims, *args, (0, 0, 0), dims)
    full.forward(phi, f0)
    g = full.backward()
    full.close()
    origin, size = (3, 1, 2), (9, 8, 5)
    part = SolidFractionAdjoint(*dims, *args, origin, size)
    part.forward(phi, f0)
    gb = part.backward()
    part.close()
    (ox, oy, oz), (bx, by, bz) = origin, size
    assert np.array_equal(gb, g[oz:oz + bz, oy:oy + by, ox:ox + bx])


def test_checkpoint_interval_does_not_change_the_gradient():
    from tlbm.adjoint import SolidFractionAdjoint
    dims = (10, 8, 6)
    phi, f0 = _random_case(dims, seed=5)
    out = []
    for k in (1, 3, 7, 40):
        run = SolidFractionAdjoint(*dims, 0.02, 0.05, 15, 9, k, (0, 0, 0), dims)
        out.append((run.forward(phi, f0), run.backward()))
        run.close()
    for J, g in out[1:]:
        assert J == out[0][0] and np.array_equal(g, out[0][1])


def _small_drag_problem():
    from tlbm.adjoint import DragAdjoint
    cfg = dict(nx=32, ny=24, nz=24, reynolds=60.0, u0=0.08, radius=4.0,
               nominal_center=(10.0, 12.0, 12.0), block=14, steps=80, window=40,
               checkpoint_every=16, eps=1.5)
    rng = np.random.default_rng(11)
    center = np.array(cfg["nominal_center"]) + rng.uniform(-1, 1, 3)
    por = 0.01 + 0.01 * rng.uniform(size=(14, 14, 14))
    return DragAdjoint(**cfg), cfg, 4.0, center, por


def _torch_design_gradient(cfg, radius, center, por):
    dims = (cfg["nx"], cfg["ny"], cfg["nz"])
    origin = tuple(int(round(c - cfg["block"] / 2.0)) for c in cfg["nominal_center"])
    nu = cfg["u0"] * 2.0 * cfg["radius"] / cfg["reynolds"]
    omega = float(np.float32(1.0 / (3.0 * nu + 0.5)))
    r = torch.tensor(radius, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(center, dtype=torch.float64, requires_grad=True)
    P = torch.tensor(por, dtype=torch.float64, requires_grad=True)
    phi = T.solid_fraction(dims, c, r, P, origin, cfg["eps"])
    J = T.mean_drag(phi, T.equilibrium_state(dims, cfg["u0"]), cfg["steps"], cfg["window"],
                    omega, cfg["u0"], cfg["nx"] - 8.0)
    J.backward()
    return float(J.detach()), float(r.grad), c.grad.numpy(), P.grad.numpy()


def test_drag_adjoint_matches_fp64_autograd():
    prob, cfg, radius, center, por = _small_drag_problem()
    J = prob.forward(radius, center, por)
    g = prob.backward()
    prob.close()
    J_ref, gr, gc, gP = _torch_design_gradient(cfg, radius, center, por)
    print(f"J {J:.8e} ref {J_ref:.8e}; dJ/dr {g['radius']:.8e} ref {gr:.8e}; "
          f"dJ/dc {g['center']} ref {gc}")
    assert abs(J - J_ref) <= 1e-4 * abs(J_ref)
    vec = np.concatenate([[g["radius"]], g["center"]])
    ref = np.concatenate([[gr], gc])
    assert np.abs(vec - ref).max() <= 1e-4 * np.abs(ref).max()
    assert abs(g["radius"] - gr) <= 1e-4 * abs(gr)
    assert np.abs(g["porosity"] - gP).max() <= 1e-4 * np.abs(gP).max()


def test_autograd_function_front_end():
    from tlbm.adjoint import mean_drag
    prob, cfg, radius, center, por = _small_drag_problem()
    r = torch.tensor(radius, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(center, dtype=torch.float64, requires_grad=True)
    P = torch.tensor(por, dtype=torch.float64, requires_grad=True)
    J = mean_drag(prob, r, c, P)
    (2.0 * J).backward()
    J2 = prob.forward(radius, center, por)
    g = prob.backward(2.0)
    prob.close()
    assert float(J.detach()) == J2
    assert float(r.grad) == g["radius"]
    assert np.array_equal(c.grad.numpy(), g["center"])
    assert np.array_equal(P.grad.numpy(), g["porosity"])


def test_c5_adjoint_against_fp64_fixture():
    """The full workload: 128x64x64, 2000 steps, checkpoints every 100, against the FP64
    autograd gradient and the central differences (h = 1e-3) of tests/golden/c5_adjoint.npz."""
    import hashlib

    from tlbm.adjoint import DragAdjoint
    path = os.path.join(HERE, "golden", "c5_adjoint.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/c5_adjoint.npz not generated")
    ref = np.load(path)
    radius, center, por = c5_design()
    assert hashlib.sha256(np.ascontiguousarray(por).tobytes()).hexdigest() == \
        str(ref["porosity_sha256"])
    assert float(ref["fd_step"]) == C5_FD_STEP
    prob = DragAdjoint(**C5)
    J = prob.forward(radius, center, por)
    g = prob.backward()
    fwd, bwd = prob.forward_ms, prob.backward_ms
    prob.close()
    print(f"J {J:.9e} ref {float(ref['J']):.9e}; forward {fwd:.1f} ms, backward {bwd:.1f} ms "
          f"({bwd / fwd:.2f}x)")
    assert abs(J - float(ref["J"])) <= 1e-4 * abs(float(ref["J"]))
    ours = {"radius": g["radius"], "cx": g["center"][0], "cy": g["center"][1],
            "cz": g["center"][2]}
    for cell in ref["fd_cells"]:
        ours[f"P{tuple(int(i) for i in cell)}"] = g["porosity"][tuple(cell)]
    vec = np.array([ref["grad_radius"], *ref["grad_center"]])
    for name, fd in zip(ref["fd_names"], ref["fd_values"]):
        mine = ours[str(name)]
        print(f"{name}: adjoint {mine:.8e} FD {fd:.8e} rel {abs(mine - fd) / abs(fd):.2e}")
        # each component to 1e-3 of itself, or of the (r, c) gradient's largest component
        # for the centre coordinates that symmetry makes small
        scale = abs(fd) if str(name) in ("radius",) or str(name).startswith("P") else \
            max(abs(fd), np.abs(vec).max())
        assert abs(mine - fd) <= 1e-3 * scale
    gP = ref["grad_porosity"].astype(np.float64)
    rel = np.linalg.norm(g["porosity"] - gP) / np.linalg.norm(gP)
    print(f"porosity gradient: |g - g_ref| / |g_ref| = {rel:.2e}")
    assert rel <= 1e-3
    assert abs(g["radius"] - float(ref["grad_radius"])) <= 1e-3 * abs(float(ref["grad_radius"]))
