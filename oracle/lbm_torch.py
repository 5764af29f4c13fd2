"""PyTorch FP64 restatement of the solid-fraction BGK step, periodic in x, y and z, and of the
mean-drag objective, differentiable in the solid fraction and in the obstacle parameters.

TEST INFRASTRUCTURE: the reference gradient of the adjoint tests (tests/ and the fixture script
tests/golden/make_c5_adjoint.py).  The product (tack-lbm_amd/) never imports it.

Step, term for term as ``lbm_oracle.step`` with ``Field`` geometry, ``BGK`` collision and
``PeriodicY`` side (lbm.py:397-456; BGK lbm.py:60-69; periodic pull lbm.py:342-352; the
momentum exchange lbm.py:429-430), written on (19, nz, ny, nx) tensors with ``torch.roll`` for
the periodic pull:
    g_q(x)   = p + phi(x - c_q) (b - p),  p = f_q(x - c_q),  b = f_opp(q)(x)
    F_x      = sum_x sum_q phi(x - c_q) (1 - phi(x)) 2 b c_opp(q),x
    rho, u   = sum g,  (1 - phi) sum c g / rho
    f'_q     = feq_q(rho_s, u_s) + (1 - sigma)(1 - omega)(g_q - feq_q(rho, u))
    sigma    = max(phi, k t^2),  t = max(0, (x - s0) / (nx - s0))
    rho_s    = rho + sigma (1 - rho),  u_s = (u_x + sigma (u0 (1 - phi) - u_x), (1 - sigma) u_y,
               (1 - sigma) u_z)
``sigma`` takes the derivative of ``phi`` only where phi > k t^2 (a tie passes it to k t^2), the
one-sided rule of the adjoint kernel at the kink of the max.

Pinned by tests/test_adjoint_reference.py: the step against ``lbm_oracle.step`` and the C oracle
(FP64, 1e-13), the autograd gradient against central finite differences of this same FP64
forward, and the geometry against closed forms (the volume of the smoothed sphere, the radial
derivative of the smootherstep profile).
"""

import numpy as np
import torch

from .lbm_oracle import CX, CY, CZ, OPP, Q, W

_C = [(int(CZ[q]), int(CY[q]), int(CX[q])) for q in range(Q)]


def _feq(rho, ux, uy, uz):
    usq = ux * ux + uy * uy + uz * uz
    out = []
    for q in range(Q):
        cu = 3.0 * (float(CX[q]) * ux + float(CY[q]) * uy + float(CZ[q]) * uz)
        out.append(float(W[q]) * rho * (1.0 + cu + 0.5 * cu * cu - 1.5 * usq))
    return torch.stack(out)


class Lattice:
    """Per-run constants: the solid fraction's link walls phi(x - c_q) and the sponge profile."""

    def __init__(self, phi, omega, u0, sponge_start, sponge_strength):
        self.phi = phi                                        # (nz, ny, nx), float64
        nz, ny, nx = phi.shape
        self.fluid = 1.0 - phi
        self.walls = [None] + [torch.roll(phi, shifts=_C[q], dims=(0, 1, 2)) for q in range(1, Q)]
        x = torch.arange(nx, dtype=phi.dtype, device=phi.device)
        t = torch.clamp((x - sponge_start) / (nx - sponge_start), min=0.0)
        ks = (sponge_strength * t * t).expand(nz, ny, nx)
        self.sigma = torch.where(phi > ks, phi, ks)
        self.omega = float(omega)
        self.u0 = float(u0)


def step(lat, f):
    """One step: returns (f_next, total x force of this step)."""
    fl = lat.fluid
    g = [f[0]]
    fx = torch.zeros((), dtype=f.dtype, device=f.device)
    for q in range(1, Q):
        p = torch.roll(f[q], shifts=_C[q], dims=(0, 1, 2))
        b = f[OPP[q]]
        w = lat.walls[q]
        g.append(p + w * (b - p))
        fx = fx + float(CX[OPP[q]]) * (w * fl * 2.0 * b).sum()
    g = torch.stack(g)
    rho = g.sum(0)
    cx = torch.tensor(CX, dtype=f.dtype, device=f.device).view(Q, 1, 1, 1)
    cy = torch.tensor(CY, dtype=f.dtype, device=f.device).view(Q, 1, 1, 1)
    cz = torch.tensor(CZ, dtype=f.dtype, device=f.device).view(Q, 1, 1, 1)
    ux = fl * (cx * g).sum(0) / rho
    uy = fl * (cy * g).sum(0) / rho
    uz = fl * (cz * g).sum(0) / rho
    e = _feq(rho, ux, uy, uz)
    s = lat.sigma
    rho_s = rho + s * (1.0 - rho)
    es = _feq(rho_s, ux + s * (lat.u0 * fl - ux), (1.0 - s) * uy, (1.0 - s) * uz)
    return es + ((1.0 - s) * (1.0 - lat.omega)) * (g - e), fx


def equilibrium_state(dims, u0, dtype=torch.float64, device="cpu"):
    """The uniform state feq(1, (u0, 0, 0)) (lbm.py:383-394 with no seed, no solid)."""
    nx, ny, nz = dims
    one = torch.ones((nz, ny, nx), dtype=dtype, device=device)
    return _feq(one, u0 * one, 0.0 * one, 0.0 * one)


def mean_drag(phi, f0, steps, window, omega, u0, sponge_start, sponge_strength=1.0,
              segment=0):
    """J = (1 / window) sum of the x force over the last ``window`` of ``steps`` steps from f0.

    ``segment`` > 0 recomputes the graph of each run of that many steps in the backward pass
    (torch.utils.checkpoint) so that long runs fit in memory.
    """
    lat = Lattice(phi, omega, u0, sponge_start, sponge_strength)
    first = steps - window

    def run(f, t0, t1):
        acc = torch.zeros((), dtype=f.dtype, device=f.device)
        for t in range(t0, t1):
            f, fx = step(lat, f)
            if t >= first:
                acc = acc + fx
        return f, acc

    f = f0
    total = torch.zeros((), dtype=f0.dtype, device=f0.device)
    t = 0
    while t < steps:
        t1 = steps if segment <= 0 else min(steps, t + segment)
        if segment > 0 and torch.is_grad_enabled():
            from torch.utils.checkpoint import checkpoint
            f, acc = checkpoint(run, f, t, t1, use_reentrant=False)
        else:
            f, acc = run(f, t, t1)
        total = total + acc
        t = t1
    return total / window


def smoothed_sphere(dims, center, radius, eps):
    """Solid fraction of a sphere with a quintic smootherstep surface of width 2 eps:
    s = clip(1/2 + (r - |x - c|) / (2 eps), 0, 1), phi = s^3 (10 - 15 s + 6 s^2)."""
    nx, ny, nz = dims
    dt, dev = radius.dtype, radius.device
    z, y, x = torch.meshgrid(torch.arange(nz, dtype=dt, device=dev),
                             torch.arange(ny, dtype=dt, device=dev),
                             torch.arange(nx, dtype=dt, device=dev), indexing="ij")
    d = torch.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2)
    s = torch.clamp(0.5 + (radius - d) / (2.0 * eps), 0.0, 1.0)
    return s * s * s * (10.0 + s * (-15.0 + 6.0 * s))


def solid_fraction(dims, center, radius, porosity, origin, eps):
    """phi = 1 - (1 - phi_sphere)(1 - P), P the porosity block at ``origin`` (zero elsewhere)."""
    nx, ny, nz = dims
    sph = smoothed_sphere(dims, center, radius, eps)
    bz, by, bx = porosity.shape
    P = torch.zeros((nz, ny, nx), dtype=porosity.dtype, device=porosity.device)
    P = P.index_put((torch.arange(origin[2], origin[2] + bz).view(-1, 1, 1),
                     torch.arange(origin[1], origin[1] + by).view(1, -1, 1),
                     torch.arange(origin[0], origin[0] + bx).view(1, 1, -1)), porosity)
    return 1.0 - (1.0 - sph) * (1.0 - P)


def to_numpy_populations(f):
    """(19, nz, ny, nx) tensor -> (19, n) array in the oracle's flat cell order (x fastest)."""
    return f.detach().cpu().numpy().reshape(Q, -1)


def from_numpy_populations(a, dims, dtype=torch.float64):
    nx, ny, nz = dims
    return torch.as_tensor(np.asarray(a, np.float64).reshape(Q, nz, ny, nx), dtype=dtype)
