"""NumPy FP64 restatement of the derived fields of the output grid: the velocity gradient at the
points and the vorticity magnitude and Q-criterion at the cell centres.

TEST INFRASTRUCTURE: the checker of tlbm_velocity_gradient / tlbm_vortex_fields (tests/).  The
product (tack-lbm_amd/) never imports it.

* point_gradient: derived.py:52-89 (`_jacobian`): central differences, one-sided at the grid's
  faces, zero along an axis of one point.
* cell_gradient_lattice: derived_check.py:88-97: the mean of a cell's eight corner gradients.
* cell_gradient_basis: derived_check.py:53-66: the trilinear cell's gradient at its centre, the
  mean of the four edge differences along each axis.
* vortex_fields: derived_check.py:67-72, 98-102: |curl u| and
  Q = -1/2 (G00^2 + G11^2 + G22^2) - (G01 G10 + G02 G20 + G12 G21).
Arrays are (mz, my, mx) (x fastest in memory), gradients G[c, d] = d u_c / d x_d.

Pinned by tests/test_derived.py: exact gradients of linear fields u = A x + b at interior and
face points and in both cell forms; solid-body rotation (|curl u| = 2 |Omega|, Q = |Omega|^2);
a symmetric trace-free strain (|curl u| = 0, Q = -|S|_F^2 / 2); and agreement with the
np.gradient formulation of derived_check.py:88 on a random field.
"""

import numpy as np


def point_gradient(u, h):
    """u: (3, mz, my, mx) -> J: (3, 3, mz, my, mx), J[c, d] = d u_c / d x_d (x, y, z = d)."""
    u = np.asarray(u, np.float64)
    J = np.zeros((3, 3) + u.shape[1:])
    for d, axis in enumerate((3, 2, 1)):           # x, y, z are the last, middle, first axes
        n = u.shape[axis]
        if n == 1:
            continue
        lo = np.clip(np.arange(n) - 1, 0, n - 1)
        hi = np.clip(np.arange(n) + 1, 0, n - 1)
        scale = 1.0 / (h * (hi - lo))
        shape = [1, 1, 1]
        shape[axis - 1] = n
        diff = np.take(u, hi, axis=axis) - np.take(u, lo, axis=axis)
        J[:, d] = diff * scale.reshape(shape)
    return J


def cell_gradient_lattice(J):
    """The mean of each cell's eight corner gradients: (3, 3, mz-1, my-1, mx-1)."""
    return 0.125 * sum(J[..., dk:J.shape[-3] - 1 + dk, dj:J.shape[-2] - 1 + dj,
                         di:J.shape[-1] - 1 + di]
                       for dk in (0, 1) for dj in (0, 1) for di in (0, 1))


def cell_gradient_basis(u, h):
    """The trilinear cell's gradient at its centre: (3, 3, mz-1, my-1, mx-1)."""
    u = np.asarray(u, np.float64)
    G = np.empty((3, 3, u.shape[1] - 1, u.shape[2] - 1, u.shape[3] - 1))
    ex = (u[..., 1:] - u[..., :-1]) / h                  # edges along x
    G[:, 0] = 0.25 * (ex[:, :-1, :-1] + ex[:, :-1, 1:] + ex[:, 1:, :-1] + ex[:, 1:, 1:])
    ey = (u[:, :, 1:, :] - u[:, :, :-1, :]) / h
    G[:, 1] = 0.25 * (ey[:, :-1, :, :-1] + ey[:, :-1, :, 1:] + ey[:, 1:, :, :-1] + ey[:, 1:, :, 1:])
    ez = (u[:, 1:] - u[:, :-1]) / h
    G[:, 2] = 0.25 * (ez[:, :, :-1, :-1] + ez[:, :, :-1, 1:] + ez[:, :, 1:, :-1] + ez[:, :, 1:, 1:])
    return G


def vortex_fields(G):
    """(|curl u|, Q) from cell gradients G (3, 3, ...)."""
    w = np.stack([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    vort = np.sqrt((w ** 2).sum(axis=0))
    q = -0.5 * (G[0, 0] ** 2 + G[1, 1] ** 2 + G[2, 2] ** 2) - (
        G[0, 1] * G[1, 0] + G[0, 2] * G[2, 0] + G[1, 2] * G[2, 1])
    return vort, q
