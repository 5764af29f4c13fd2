"""Derived fields of the output grid (vorticity magnitude, Q-criterion): the NumPy oracle pinned
by closed forms and by derived_check.py's own formulation (CPU), and the HIP kernels against the
oracle (GPU, through the C ABI and through tlbm.derived.DerivedFields)."""

import ctypes

import numpy as np
import pytest

from oracle import derived_oracle as D


def _linear(dims, A, b, h):
    mx, my, mz = dims
    z, y, x = np.meshgrid(np.arange(mz) * h, np.arange(my) * h, np.arange(mx) * h, indexing="ij")
    X = np.stack([x, y, z])
    return np.einsum("cd,d...->c...", A, X) + np.asarray(b).reshape(3, 1, 1, 1)


def test_linear_field_gradients_are_exact_everywhere():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((3, 3))
    h = 0.37
    u = _linear((6, 5, 4), A, [0.1, -0.2, 0.3], h)
    J = D.point_gradient(u, h)
    assert np.allclose(J, A[:, :, None, None, None], atol=1e-12)      # faces included
    for G in (D.cell_gradient_lattice(J), D.cell_gradient_basis(u, h)):
        assert np.allclose(G, A[:, :, None, None, None], atol=1e-12)
    vort, q = D.vortex_fields(D.cell_gradient_basis(u, h))
    w = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.allclose(vort, np.linalg.norm(w), atol=1e-12)
    # Q = (|W|^2 - |S|^2) / 2 with S, W the symmetric and antisymmetric parts of A
    S, W = 0.5 * (A + A.T), 0.5 * (A - A.T)
    assert np.allclose(q, 0.5 * ((W ** 2).sum() - (S ** 2).sum()), atol=1e-12)


def test_solid_body_rotation_and_pure_strain():
    om = np.array([0.3, -0.5, 0.7])
    R = np.array([[0, -om[2], om[1]], [om[2], 0, -om[0]], [-om[1], om[0], 0]])
    vort, q = D.vortex_fields(D.cell_gradient_lattice(D.point_gradient(
        _linear((5, 5, 5), R, [0, 0, 0], 1.0), 1.0)))
    assert np.allclose(vort, 2 * np.linalg.norm(om)) and np.allclose(q, om @ om)
    S = np.array([[0.2, 0.1, 0.0], [0.1, -0.5, 0.3], [0.0, 0.3, 0.3]])     # symmetric, trace 0
    vort, q = D.vortex_fields(D.cell_gradient_basis(_linear((4, 4, 4), S, [1, 1, 1], 0.5), 0.5))
    assert np.allclose(vort, 0.0) and np.allclose(q, -0.5 * (S ** 2).sum())


def test_point_gradient_matches_numpy_gradient_formulation():
    """derived_check.py:88: np.gradient per component (central inside, first order at faces)."""
    rng = np.random.default_rng(1)
    u = rng.standard_normal((3, 5, 6, 7))
    h = 0.25
    J = D.point_gradient(u, h)
    ref = np.stack([np.stack(np.gradient(u[i], h, axis=(2, 1, 0))) for i in range(3)])
    assert np.allclose(J, ref, rtol=0, atol=1e-12)


def test_point_gradient_of_a_single_point_axis_is_zero():
    u = np.random.default_rng(2).standard_normal((3, 1, 4, 5))
    J = D.point_gradient(u, 1.0)
    assert np.all(J[:, 2] == 0.0)


def test_derived_fields_reject_expressions_without_a_kernel():
    from tlbm.derived import DerivedFields, parse
    assert parse(["v=mag(curl(velocity))@cells", "s=speed"]) == {
        "v": ("mag(curl(velocity))", "cells"), "s": ("speed", "points")}

    class FakeSolver:
        out_dims, stride, diameter = (4, 4, 4), 1, 8.0
    with pytest.raises(NotImplementedError):
        DerivedFields(FakeSolver(), {"div": ("div(velocity)", "cells")})


def test_entry_points_reject_bad_arguments():
    from tlbm import _capi
    L = _capi.lib()
    p = ctypes.c_void_p(64)
    assert L.tlbm_velocity_gradient(None, p, p, p, 4, 4, 4, 1.0, None) != 0
    assert L.tlbm_velocity_gradient(p, p, p, p, 0, 4, 4, 1.0, None) != 0
    for basis in (0, 1):
        assert L.tlbm_vortex_fields(None, p, p, p, p, 4, 4, 4, 1.0, basis, None) != 0  # no u
        assert L.tlbm_vortex_fields(p, p, p, p, p, 1, 4, 4, 1.0, basis, None) != 0     # no cells
        assert L.tlbm_vortex_fields(p, p, p, None, None, 4, 4, 4, 1.0, basis, None) == 0


# ---- GPU ---------------------------------------------------------------------------------------
def _gpu_fields(u32, h, basis):
    import tlbm
    mz, my, mx = u32.shape[1:]
    n = mx * my * mz
    L = tlbm._capi.lib()
    dev = [tlbm.array(u32[c].reshape(-1)) for c in range(3)]
    jac = tlbm.empty((9 * n,))
    cells = (mx - 1) * (my - 1) * (mz - 1)
    vort, qc = tlbm.empty((cells,)), tlbm.empty((cells,))
    tlbm._capi.check(L.tlbm_velocity_gradient(dev[0].ptr, dev[1].ptr, dev[2].ptr, jac.ptr, mx,
                                              my, mz, 1.0 / h, None), "gradient")
    tlbm._capi.check(L.tlbm_vortex_fields(dev[0].ptr, dev[1].ptr, dev[2].ptr, vort.ptr, qc.ptr,
                                          mx, my, mz, 1.0 / h, basis, None), "cells")
    tlbm.sync()
    # one output only: the other is not written
    only = tlbm.empty((cells,))
    tlbm._capi.check(L.tlbm_vortex_fields(dev[0].ptr, dev[1].ptr, dev[2].ptr, None, only.ptr,
                                          mx, my, mz, 1.0 / h, basis, None), "q only")
    tlbm.sync()
    assert np.array_equal(only.numpy(), qc.numpy())
    return (jac.numpy().reshape(3, 3, mz, my, mx), vort.numpy().reshape(mz - 1, my - 1, mx - 1),
            qc.numpy().reshape(mz - 1, my - 1, mx - 1))


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(2, 2, 2), (7, 5, 3), (64, 33, 17), (130, 9, 4), (3, 2, 70),
                                  (66, 7, 35), (2, 6, 33)])
@pytest.mark.parametrize("basis", [0, 1])
def test_kernels_match_oracle(dims, basis):
    """Point gradients and cell fields against the oracle; the cell kernel's z chunks are 32
    cells, so the last three grids cross chunk boundaries (x and y tiles: 64 x 4 cells)."""
    rng = np.random.default_rng(sum(dims) + basis)
    mx, my, mz = dims
    u = (0.05 * rng.standard_normal((3, mz, my, mx))).astype(np.float32)
    h = 1.0 / 24.0
    J, vort, q = _gpu_fields(u, h, basis)
    Jr = D.point_gradient(u, h)
    assert np.abs(J - Jr).max() <= 1e-5 * np.abs(Jr).max()
    G = D.cell_gradient_basis(u, h) if basis else D.cell_gradient_lattice(Jr)
    vr, qr = D.vortex_fields(G)
    assert np.abs(vort - vr).max() <= 2e-5 * np.abs(vr).max()
    assert np.abs(q - qr).max() <= 2e-5 * np.abs(qr).max()


@pytest.mark.gpu
@pytest.mark.parametrize("stride", [1, 2])
def test_derived_fields_of_a_solver(stride):
    """DerivedFields on a running solver (derived_check.py's setup, smaller): the kernels read
    the report fields, coarsened by `stride`, and match the oracle on the same velocities."""
    import tlbm
    from tlbm.derived import DerivedFields
    s = tlbm.Solver(48, 32, 16, collision="les", obstacle="cylinder", reynolds=300.0,
                    stride=stride)
    for _ in range(40):
        s.step()
    s.step(report=True)
    for gradient in ("lattice", "basis"):
        d = DerivedFields(s, gradient=gradient)
        res = d.evaluate()
        s.update_output()
        tlbm.sync()
        mx, my, mz = s.out_dims
        u = np.stack([s.output[k].numpy.reshape(mz, my, mx) for k in ("ux", "uy", "uz")])
        h = stride / s.diameter
        G = D.cell_gradient_basis(u, h) if gradient == "basis" else \
            D.cell_gradient_lattice(D.point_gradient(u, h))
        vr, qr = D.vortex_fields(G)
        assert res["vorticity"][1] == "element"
        assert np.abs(res["vorticity"][0] - vr.reshape(-1)).max() <= 1e-4 * np.abs(vr).max()
        assert np.abs(res["q_criterion"][0] - qr.reshape(-1)).max() <= 1e-4 * np.abs(qr).max()
