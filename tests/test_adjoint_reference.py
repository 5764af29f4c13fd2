"""The FP64 PyTorch reference of the adjoint workload (oracle/lbm_torch.py) and the host side of
the product's adjoint (geometry and chain rule), on the CPU.

The torch step is pinned against the NumPy and C oracles (the same step restated twice more),
its autograd gradient against central finite differences of the same FP64 forward, and the
product's analytic derivatives of the smoothed sphere against torch autograd of an independent
restatement of the geometry.
"""

import ctypes
import os

import numpy as np
import pytest
import torch

from cases import random_fraction, random_populations
from oracle import c_oracle as C
from oracle import lbm_oracle as O
from oracle import lbm_torch as T


def _tensor_phi(phi, dims):
    nx, ny, nz = dims
    return torch.tensor(np.asarray(phi, np.float64).reshape(nz, ny, nx))


@pytest.mark.parametrize("dims", [(10, 7, 6), (5, 4, 3)])
def test_torch_step_matches_numpy_and_c_oracles(dims):
    n = dims[0] * dims[1] * dims[2]
    phi = random_fraction(dims, seed=8)
    f = random_populations(n, 9)
    phys = O.BGK(0.02)
    geo = O.Field(phi, dims)
    flow = O.Flow(0.08, dims[0] - 3.0, 0.8, 0.0)
    diag = {k: np.zeros(n) for k in O.FIELDS}
    ref = O.step(dims, phys, geo, flow, f.copy(), diag)
    lat = T.Lattice(_tensor_phi(phi, dims), phys.omega, 0.08, dims[0] - 3.0, 0.8)
    got, fx = T.step(lat, T.from_numpy_populations(f, dims))
    assert np.abs(T.to_numpy_populations(got) - ref).max() < 1e-15
    assert abs(float(fx) - diag["fx"].sum()) < 1e-13
    if C.available():
        ref_c = C.step(dims, phys, geo, flow, f)
        assert np.abs(T.to_numpy_populations(got) - ref_c).max() < 1e-15


def test_torch_equilibrium_state_is_the_uniform_inflow():
    f = T.equilibrium_state((4, 3, 2), 0.08).numpy().reshape(19, -1)
    for q in range(19):
        assert np.allclose(f[q], O.feq(q, 1.0, 0.08, 0.0, 0.0), rtol=0, atol=1e-16)


def _tiny_problem(seed=3):
    dims = (7, 6, 5)
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.05, 0.9, dims[0] * dims[1] * dims[2])  # no ties with the sponge
    f0 = random_populations(dims[0] * dims[1] * dims[2], seed + 1)
    return dims, phi, f0


def test_torch_gradient_matches_central_differences():
    dims, phi, f0 = _tiny_problem()
    kw = dict(steps=10, window=6, omega=1.6, u0=0.08, sponge_start=4.0, sponge_strength=0.7)
    f0t = T.from_numpy_populations(f0, dims)
    p = _tensor_phi(phi, dims).requires_grad_(True)
    T.mean_drag(p, f0t, **kw).backward()
    g = p.grad.numpy().reshape(-1)
    h = 1e-6
    rng = np.random.default_rng(0)
    for i in list(rng.choice(phi.size, 6, replace=False)) + [int(np.argmax(np.abs(g)))]:
        vals = []
        for s in (h, -h):
            q = phi.copy()
            q[i] += s
            with torch.no_grad():
                vals.append(float(T.mean_drag(_tensor_phi(q, dims), f0t, **kw)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])


def test_checkpointed_gradient_equals_plain_gradient():
    dims, phi, f0 = _tiny_problem(5)
    kw = dict(steps=9, window=4, omega=1.7, u0=0.05, sponge_start=5.0)
    f0t = T.from_numpy_populations(f0, dims)
    grads = []
    for seg in (0, 4):
        p = _tensor_phi(phi, dims).requires_grad_(True)
        T.mean_drag(p, f0t, segment=seg, **kw).backward()
        grads.append(p.grad.numpy())
    assert np.allclose(grads[0], grads[1], rtol=0, atol=1e-15)


def test_smoothed_sphere_profile():
    dims = (30, 30, 30)
    c = torch.tensor([14.3, 15.1, 14.7], dtype=torch.float64)
    r = torch.tensor(8.0, dtype=torch.float64)
    phi = T.smoothed_sphere(dims, c, r, 1.5).numpy().reshape(-1)
    x, y, z = O.coords(*dims)
    d = np.sqrt((x - 14.3) ** 2 + (y - 15.1) ** 2 + (z - 14.7) ** 2)
    assert np.all(phi[d <= 8.0 - 1.5] == 1.0) and np.all(phi[d >= 8.0 + 1.5] == 0.0)
    band = (d > 6.5) & (d < 9.5)
    assert np.all((phi[band] > 0) & (phi[band] < 1))
    # the smootherstep is odd about the surface, phi(r + u) + phi(r - u) = 1: on the x axis of
    # a sphere of radius 30 at the origin, cells x = 30 - k and 30 + k pair up
    def profile(radius):
        return T.smoothed_sphere((61, 1, 1), torch.zeros(3, dtype=torch.float64),
                                 torch.tensor(radius, dtype=torch.float64), 1.5).numpy()[0, 0]
    prof = profile(30.0)
    k = np.arange(31)
    assert np.allclose(prof[30 - k] + prof[30 + k], 1.0, rtol=0, atol=1e-15)
    assert prof[30] == 0.5
    # the radial slope at the surface is S'(1/2) / (2 eps) = (15/8) / 3
    dr = 1e-5
    assert abs((profile(30.0 + dr)[30] - profile(30.0 - dr)[30]) / (2 * dr) - 0.625) < 1e-8
    # its volume is that of the sharp sphere up to the lattice sampling (~1 %)
    assert abs(phi.sum() / (4.0 / 3.0 * np.pi * 8.0 ** 3) - 1.0) < 0.02


def test_product_geometry_derivatives_match_autograd():
    """The product's analytic d phi / d(r, c, P) (tlbm.adjoint.SmoothedSphere) against torch
    autograd of the independent restatement lbm_torch.solid_fraction."""
    from tlbm.adjoint import SmoothedSphere
    dims = (40, 36, 34)
    origin, size = (6, 5, 4), (26, 24, 25)
    rng = np.random.default_rng(1)
    center = np.array([19.3, 16.8, 16.1])
    radius = 7.4
    por = 0.01 + 0.02 * rng.uniform(size=size[::-1])
    geo = SmoothedSphere(origin, size, eps=1.5)
    assert geo.fits(center, radius)
    phi_b, dr, dc, dP = geo.evaluate(center, radius, por)
    r = torch.tensor(radius, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(center, dtype=torch.float64, requires_grad=True)
    P = torch.tensor(por, dtype=torch.float64, requires_grad=True)
    phi = T.solid_fraction(dims, c, r, P, origin, 1.5)
    (ox, oy, oz), (bx, by, bz) = origin, size
    blk = phi[oz:oz + bz, oy:oy + by, ox:ox + bx]
    assert np.allclose(blk.detach().numpy(), phi_b, rtol=0, atol=1e-15)
    # nothing of the sphere outside the block
    outside = phi.detach().clone()
    outside[oz:oz + bz, oy:oy + by, ox:ox + bx] = 0.0
    assert float(outside.abs().max()) == 0.0
    a = torch.tensor(rng.standard_normal(size[::-1]))  # a random dJ/dphi
    (a * blk).sum().backward()
    an = a.numpy()
    assert np.isclose(float(r.grad), (an * dr).sum(), rtol=1e-12, atol=1e-12)
    assert np.allclose(c.grad.numpy(), [(an * dc[k]).sum() for k in range(3)], rtol=1e-12,
                       atol=1e-12)
    assert np.allclose(P.grad.numpy(), an * dP, rtol=1e-12, atol=1e-14)


def test_design_block_fit_check():
    from tlbm.adjoint import SmoothedSphere
    geo = SmoothedSphere((16, 16, 16), (32, 32, 32), eps=1.5)
    assert geo.fits((34.0, 30.0, 33.0), 10.0)
    assert not geo.fits((20.0, 32.0, 32.0), 10.0)     # reaches x = 8.5 < 15
    assert not geo.fits((32.0, 32.0, 32.0), 15.0)


def test_adjoint_entry_point_rejects_bad_arguments():
    """Argument checks of tlbm_adjoint_step run before any device work."""
    import tlbm
    from tlbm import _capi
    L = _capi.lib()
    p = _capi.Params()
    p.nx, p.ny, p.nz, p.qstride = 8, 6, 4, 8 * 6 * 4
    p.omega = 1.5
    o = (ctypes.c_int32 * 3)(0, 0, 0)
    s = (ctypes.c_int32 * 3)(8, 6, 4)
    dummy = ctypes.c_void_p(16)
    # no phi: the solid-fraction obstacle needs one
    assert L.tlbm_adjoint_step(ctypes.byref(p), dummy, dummy, dummy, ctypes.c_void_p(32), 0.0,
                               0.0, o, s, dummy, 0, None) != 0
    p.phi = 48
    assert L.tlbm_adjoint_step(ctypes.byref(p), None, dummy, dummy, ctypes.c_void_p(32), 0.0,
                               0.0, o, s, dummy, 0, None) != 0
    assert L.tlbm_adjoint_step(ctypes.byref(p), dummy, dummy, dummy, dummy, 0.0, 0.0, o, s,
                               dummy, 0, None) != 0                       # mu_in == mu_out
    bad = (ctypes.c_int32 * 3)(8, 0, 4)
    assert L.tlbm_adjoint_step(ctypes.byref(p), dummy, dummy, dummy, ctypes.c_void_p(32), 0.0,
                               0.0, o, bad, dummy, 0, None) != 0
    p.qstride = 8 * 6 * 4 - 16                                            # below the cells
    assert L.tlbm_adjoint_step(ctypes.byref(p), dummy, dummy, dummy, ctypes.c_void_p(32), 0.0,
                               0.0, o, s, dummy, 0, None) != 0
    assert "qstride" in tlbm._capi.lib().tlbm_last_error().decode()
