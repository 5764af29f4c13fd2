"""Pins the CPU oracle: golden vectors, closed forms, invariants and brute force (no GPU).

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong sign
or index, a transposed table) fails it; none re-types the oracle's own formula.
"""

import itertools
import os

import numpy as np
import pytest

from oracle import lbm_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C = np.stack([O.CX, O.CY, O.CZ], axis=1)


# --- tables (lbm.py:15-35) ------------------------------------------------------------------------
def test_velocity_set_is_d3q19():
    vecs = {tuple(v) for v in C}
    expected = {v for v in itertools.product((-1, 0, 1), repeat=3) if sum(map(abs, v)) <= 2}
    assert vecs == expected and len(C) == 19
    assert tuple(C[0]) == (0, 0, 0)
    assert all(abs(C[q]).sum() == 1 for q in range(1, 7))


def test_weight_moments():
    w = O.W
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ C, 0, atol=1e-16)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, C, C), np.eye(3) / 3, atol=1e-15)
    # fourth moment isotropy of D3Q19: sum w c_x^2 c_y^2 = 1/9
    assert abs((w * C[:, 0] ** 2 * C[:, 1] ** 2).sum() - 1 / 9) < 1e-15


def test_opposite_and_mirror_by_brute_force():
    for q in range(19):
        assert (C[O.OPP[q]] == -C[q]).all()
        assert (C[O.MIRY[q]] == C[q] * np.array([1, -1, 1])).all()
    assert (O.OPP[O.OPP] == np.arange(19)).all()


# --- equilibrium (lbm.py:52-55) -----------------------------------------------------------------
def test_feq_moments_closed_form():
    rng = np.random.default_rng(0)
    rho = 1 + 0.1 * rng.standard_normal(50)
    u = 0.1 * rng.standard_normal((3, 50))
    feq = np.stack([O.feq(q, rho, *u) for q in range(19)])
    assert np.allclose(feq.sum(0), rho, rtol=0, atol=1e-14)
    assert np.allclose(C.T @ feq, rho * u, atol=1e-14)
    second = np.einsum("qa,qb,qn->abn", C, C, feq)
    expect = rho * (np.eye(3)[:, :, None] / 3 + u[:, None, :] * u[None, :, :])
    assert np.allclose(second, expect, atol=1e-14)


# --- the reference's known answer (trt_ref.py) ------------------------------------------------------
def test_step_matches_trt_ref_known_answer():
    d = np.load(os.path.join(ROOT, "tests", "golden", "trt_ref_8x6x4.npz"))
    nx, ny, nz = (int(v) for v in d["dims"])
    f_in = d["f_in"].astype(np.float64).reshape(19, -1)
    out = O.step((nx, ny, nz), O.TRT(float(d["nu"])), O.Cylinder(-100, -100, 0),
                 O.Flow(0.0, 1e9, 0.0), f_in)
    assert float(np.abs(out.reshape(-1) - d["f_out"]).max()) < 1e-13
    # the float32 restatement (the timed CPU baseline) meets the reference's own bar
    out32 = O.step((nx, ny, nz), O.TRT(float(d["nu"])), O.Cylinder(-100, -100, 0),
                   O.Flow(0.0, 1e9, 0.0), d["f_in"].reshape(19, -1))
    assert out32.dtype == np.float32
    assert float(np.abs(out32.reshape(-1) - d["f_out"]).max()) < float(d["tol"])


def test_golden_inputs_are_the_reference_recipe():
    """The fixture's inputs are trt_ref.py:11,22-25's default_rng(1) draws (hash pinned)."""
    import hashlib
    d = np.load(os.path.join(ROOT, "tests", "golden", "trt_ref_8x6x4.npz"))
    assert hashlib.sha256(d["f_in"].tobytes()).hexdigest() == (
        "de2a963bd25babfa808816091a8d19f5cd9e7d94109781c3181fea52f52723c3")


# --- collision invariants (lbm.py:60-139) ---------------------------------------------------------
def _near_eq(n=64, seed=1):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal(n)
    u = 0.05 * rng.standard_normal((3, n))
    f = np.stack([O.feq(q, rho, *u) for q in range(19)]) + 1e-3 * rng.standard_normal((19, n))
    rho = f.sum(0)
    u = (C.T @ f) / rho
    return f, rho, u


@pytest.mark.parametrize("phys", [O.BGK(0.01), O.TRT(0.002), O.LES(0.003), O.TRT(0.01, 0.25, 0.1)])
def test_collision_conserves_mass_and_momentum(phys):
    f, rho, u = _near_eq()
    post, _ = phys.collide(f, rho, *u)
    assert np.allclose(post.sum(0), rho, atol=1e-15)
    assert np.allclose(C.T @ post, rho * u, atol=1e-15)


@pytest.mark.parametrize("phys", [O.BGK(0.01), O.TRT(0.002), O.LES(0.003)])
def test_equilibrium_is_a_fixed_point(phys):
    _, rho, u = _near_eq()
    feq = np.stack([O.feq(q, rho, *u) for q in range(19)])
    post, tau = phys.collide(feq, rho, *u)
    assert np.allclose(post, feq, atol=1e-16)
    if hasattr(phys, "tau0"):                 # Pi = 0 -> tau = tau0 (lbm.py:93)
        assert np.allclose(tau, phys.tau0, atol=1e-15)


def test_les_without_subgrid_is_bgk():
    f, rho, u = _near_eq()
    nu = 0.004
    a, _ = O.LES(nu, cs=0.0).collide(f, rho, *u)
    b, _ = O.BGK(nu).collide(f, rho, *u)
    assert np.allclose(a, b, atol=1e-16)


def test_trt_with_equal_rates_is_bgk():
    """omega- = omega+ when magic = (tau - 1/2)^2 (lbm.py:115-116); with Cs = 0 tau = tau0."""
    f, rho, u = _near_eq()
    nu = 0.01
    tau0 = 3 * nu + 0.5
    a, _ = O.TRT(nu, magic=(tau0 - 0.5) ** 2, cs=0.0).collide(f, rho, *u)
    b, _ = O.BGK(nu).collide(f, rho, *u)
    assert np.allclose(a, b, atol=1e-15)


def test_trt_antisymmetric_rate_from_magic():
    """TRT relaxes the odd part with 1/(magic/(tau-1/2)+1/2): check on a pure odd perturbation."""
    _, rho, u = _near_eq(n=8)
    feq = np.stack([O.feq(q, rho, *u) for q in range(19)])
    odd = np.zeros_like(feq)
    odd[1], odd[2] = 1e-4, -1e-4              # f+ unchanged, f- perturbed
    phys = O.TRT(0.02, magic=0.25, cs=0.0)
    post, tau = phys.collide(feq + odd, rho + 0 * rho, *u)
    om_m = 1 / (0.25 / (phys.tau0 - 0.5) + 0.5)
    assert np.allclose(post - feq, (1 - om_m) * odd, atol=1e-17)


# --- streaming (lbm.py:410-426) -----------------------------------------------------------------
def _free_stream(side, dims, f):
    """One step with collision, sponge and obstacle switched off: f_out = streamed f."""
    phys = O.BGK(1.0)
    phys.omega = 0.0
    flow = O.Flow(0.0, 1e9, 0.0, 0.0, side)
    return O.step(dims, phys, O.Cylinder(-100, -100, 1.0), flow, f)


def test_periodic_pull_is_a_roll():
    dims = (5, 4, 3)
    f = O.W[:, None] + 0.01 * np.random.default_rng(2).standard_normal((19, 60))
    out = _free_stream(O.PeriodicY(), dims, f)
    for q in range(19):
        grid = f[q].reshape(3, 4, 5)
        rolled = np.roll(grid, shift=(O.CZ[q], O.CY[q], O.CX[q]), axis=(0, 1, 2))
        assert np.allclose(out[q].reshape(3, 4, 5), rolled, rtol=0, atol=1e-15)


def test_slip_walls_reflect_specularly():
    nx, ny, nz = 4, 3, 3
    f = O.W[:, None] + 0.01 * np.random.default_rng(3).standard_normal((19, nx * ny * nz))
    out = _free_stream(O.SlipY(), (nx, ny, nz), f).reshape(19, nz, ny, nx)
    g = f.reshape(19, nz, ny, nx)
    for q, z, y, x in itertools.product(range(19), range(nz), range(ny), range(nx)):
        ys = y - O.CY[q]
        qs = q
        if ys < 0 or ys >= ny:                 # leaves through a wall: mirrored, same row
            ys, qs = y, int(O.MIRY[q])
        expect = g[qs, (z - O.CZ[q]) % nz, ys, (x - O.CX[q]) % nx]
        assert abs(out[q, z, y, x] - expect) < 1e-15


def test_lid_links_bounce_with_wall_velocity():
    nx, ny, nz = 3, 3, 2
    f = O.W[:, None] + 0.01 * np.random.default_rng(4).standard_normal((19, nx * ny * nz))
    lid = (0.04, 0.01, -0.02)
    out = _free_stream(O.LidY(*lid), (nx, ny, nz), f).reshape(19, nz, ny, nx)
    g = f.reshape(19, nz, ny, nx)
    y = ny - 1
    for q in range(19):
        if O.CY[q] != -1:
            continue
        cdu = O.CX[q] * lid[0] + O.CY[q] * lid[1] + O.CZ[q] * lid[2]
        expect = g[O.OPP[q], :, y, :] + 6 * O.W[q] * cdu
        assert np.allclose(out[q, :, y, :], expect, atol=1e-15)


# --- bounce-back and momentum exchange (lbm.py:416-432) -----------------------------------------
def test_single_solid_cell_halfway_bounce_back():
    """A one-cell box: each fluid neighbour gets back its own opposite population, and the
    exchanged momentum is 2 f c_opp summed over the wall links (hand-computed)."""
    dims = (5, 5, 5)
    n = 125
    f = np.random.default_rng(5).standard_normal((19, n)) * 0.01 + O.W[:, None]
    geo = O.Box(2.0, 2.0, 2.0, 0.5, 0.5, 0.5)  # only (2,2,2) is solid
    x, y, z = O.coords(*dims)
    assert geo.solid(x, y, z).sum() == 1
    phys = O.BGK(1.0)
    phys.omega = 0.0
    fields = O.new_fields(n)
    out = O.step(dims, phys, geo, O.Flow(0.0, 1e9, 0.0), f, fields)
    solid_idx = (2 * 5 + 2) * 5 + 2
    for q in range(1, 19):
        cell = ((2 + O.CZ[q]) * 5 + 2 + O.CY[q]) * 5 + 2 + O.CX[q]   # fluid cell downstream of the solid
        assert abs(out[q, cell] - f[O.OPP[q], cell]) < 1e-16
        fx_expect = 2 * f[O.OPP[q], cell] * O.CX[O.OPP[q]]
        assert np.isclose(fields["fx"][cell], fx_expect, atol=1e-17)
    assert np.allclose(out[:, solid_idx], O.W)   # solid: sigma = 1 -> rest equilibrium


def test_rotating_cylinder_wall_term():
    """Moving wall: f_q = f_opp + 6 w_q c_q . u_w, u_w = omega * tangent at the link midpoint."""
    dims = (9, 9, 1)
    n = 81
    f = np.tile(O.W[:, None], (1, n))
    geo = O.Cylinder(4.0, 4.0, 0.9)            # one solid cell at (4, 4)
    phys = O.BGK(1.0)
    phys.omega = 0.0
    om = 0.3
    out = O.step(dims, phys, geo, O.Flow(0.0, 1e9, 0.0, om), f)
    # east neighbour (5,4): link q=1 (c=(1,0,0)) pulls from the solid; midpoint (4.5, 4)
    cell = 4 * 9 + 5
    uw = np.array([-(4 - 4.0) / 0.9, (4.5 - 4.0) / 0.9, 0.0]) * om
    expect = O.W[2] + 6 * O.W[1] * (C[1] @ uw)
    assert np.isclose(out[1, cell], expect, atol=1e-16)
    # north neighbour (4,5), link q=3 (c=(0,1,0)); midpoint (4, 4.5): u_w = (-0.5/0.9, 0) om
    cell = 5 * 9 + 4
    uw = np.array([-(4.5 - 4.0) / 0.9, 0.0, 0.0]) * om
    assert np.isclose(out[3, cell], O.W[4] + 6 * O.W[3] * (C[3] @ uw), atol=1e-16)


def test_symmetric_obstacle_at_rest_has_no_net_force():
    dims = (16, 16, 16)
    n = 16 ** 3
    f = np.tile(O.W[:, None], (1, n))
    fields = O.new_fields(n)
    O.step(dims, O.BGK(0.01), O.Sphere(8.0, 8.0, 8.0, 4.2), O.Flow(0.0, 1e9, 0.0), f, fields)
    for k in ("fx", "fy", "fz"):
        assert abs(fields[k].sum()) < 1e-14
    assert fields["fx"].any()                   # the links are there, they cancel


def test_fluid_mass_is_conserved_in_a_closed_slip_box():
    dims = (12, 10, 8)
    n = 960
    rng = np.random.default_rng(6)
    f = O.init(dims, O.Sphere(6.0, 5.0, 4.0, 2.5), O.Flow(0.05, 1e9, 0.0), 0.05) \
        + 1e-4 * rng.standard_normal((19, n))
    geo = O.Sphere(6.0, 5.0, 4.0, 2.5)
    x, y, z = O.coords(*dims)
    fluid = geo.solid(x, y, z) == 0
    flow = O.Flow(0.05, 1e9, 0.0, 0.0, O.SlipY())
    for phys in (O.BGK(0.01), O.TRT(0.003), O.LES(0.002)):
        g = f.copy()
        m0 = g[:, fluid].sum()
        for _ in range(3):
            g = O.step(dims, phys, geo, flow, g)
        assert abs(g[:, fluid].sum() - m0) < 1e-11


# --- sponge (lbm.py:446-456) --------------------------------------------------------------------------
def test_fractional_solid_interpolates_streaming_and_bounce_back():
    """phi = 0 streams, phi = 1 bounces back (the 0/1 obstacle), phi in between is the linear
    blend of the two (lbm.py:428 read with a fractional `solid`)."""
    dims = (6, 5, 4)
    n = 120
    f = O.W[:, None] + 0.01 * np.random.default_rng(11).standard_normal((19, n))
    phys = O.BGK(1.0)
    phys.omega = 0.0
    flow = O.Flow(0.0, 1e9, 0.0)
    x, y, z = O.coords(*dims)
    box = O.Box(2.0, 2.0, 1.0, 0.6, 0.6, 0.6)
    binary = box.solid(x, y, z)
    a = O.step(dims, phys, box, flow, f)
    b = O.step(dims, phys, O.Field(binary, dims), flow, f)
    assert np.abs(a - b).max() < 1e-16
    free = O.step(dims, phys, O.Field(np.zeros(n), dims), flow, f)
    half = O.step(dims, phys, O.Field(0.5 * binary, dims), flow, f)
    fluid = binary == 0
    # in fluid cells with solid neighbours: half-way between streaming and bounce-back
    assert np.allclose(half[:, fluid], 0.5 * (free[:, fluid] + a[:, fluid]), atol=1e-15)


def test_sponge_at_full_strength_gives_inflow_equilibrium():
    dims = (8, 3, 2)
    f, _, _ = _near_eq(n=48, seed=7)
    flow = O.Flow(0.08, 0.0, 4.0)              # t = x/8, sigma = 4 t^2 = 1 at x = 4
    out = O.step(dims, O.LES(0.01), O.Cylinder(-100, -100, 1.0), flow, f)
    x, _, _ = O.coords(*dims)
    col = x == 4
    expect = np.stack([O.feq(q, 1.0, 0.08, 0.0, 0.0) for q in range(19)])
    assert np.allclose(out[:, col], expect[:, None], atol=1e-16)
    assert np.array_equal(out[:, x == 0], O.step(dims, O.LES(0.01), O.Cylinder(-100, -100, 1.0),
                                                 O.Flow(0.08, 1e9, 0.0), f)[:, x == 0])


# --- report fields (lbm.py:493-513) ----------------------------------------------------------------------
def test_report_fields_from_two_consecutive_states():
    dims = (6, 5, 4)
    n = 120
    f, _, _ = _near_eq(n=n, seed=8)
    fields = O.new_fields(n)
    flow = O.Flow(0.0, 1e9, 0.0)
    geo = O.Cylinder(-100, -100, 1.0)
    g1 = O.step(dims, O.BGK(0.01), geo, flow, f, fields)
    first = {k: v.copy() for k, v in fields.items()}
    O.step(dims, O.BGK(0.01), geo, flow, g1, fields)
    # moments of the streamed populations, by an independent roll
    streamed = np.stack([np.roll(g1[q].reshape(4, 5, 6), (O.CZ[q], O.CY[q], O.CX[q]), (0, 1, 2))
                         .reshape(-1) for q in range(19)])
    rho = streamed.sum(0)
    u = (C.T @ streamed) / rho
    assert np.allclose(fields["p"], (rho - 1) / 3, atol=1e-15)
    assert np.allclose(fields["ux"], u[0], atol=1e-15)
    assert np.allclose(fields["vmag"], np.sqrt((u ** 2).sum(0)), atol=1e-15)
    assert np.allclose(fields["r_ux"], (u[0] - first["ux"]) ** 2, atol=1e-18)
    assert np.allclose(fields["r_rho"], (rho - 1 - 3 * first["p"]) ** 2, atol=1e-15)
    ke_old = 0.5 * (first["ux"] ** 2 + first["uy"] ** 2 + first["uz"] ** 2)   # no rho: lbm.py:500
    assert np.allclose(fields["r_ke"], (0.5 * rho * (u ** 2).sum(0) - ke_old) ** 2, atol=1e-18)


def test_monitors_formulas():
    n = 10
    fields = O.new_fields(n)
    fields["r_ux"][:] = 4.0
    fields["fx"][:] = 0.5
    fields["fy"][:] = -0.25
    m = O.monitors(fields, n, 50, 0.08, 24.0, 64)
    assert np.isclose(m["momentum_x"], 2.0 / 50)
    assert np.isclose(m["drag"], 5.0 / (0.5 * 0.08 ** 2 * 24 * 64))
    assert np.isclose(m["lift"], -2.5 / (0.5 * 0.08 ** 2 * 24 * 64))


# --- initialisation (lbm.py:383-394) --------------------------------------------------------------
def test_init_moments():
    dims = (16, 6, 8)
    geo = O.Cylinder(5.0, 3.0, 1.6)
    f = O.init(dims, geo, O.Flow(0.08, 12.0, 1.0), 0.05)
    x, y, z = O.coords(*dims)
    fluid = 1 - geo.solid(x, y, z)
    rho = f.sum(0)
    u = (C.T @ f) / rho
    assert np.allclose(rho, 1, atol=1e-15)
    assert np.allclose(u[0], 0.08 * fluid, atol=1e-15)
    uy = 0.05 * 0.08 * fluid * np.sin(2 * np.pi * x / 16) * np.cos(2 * np.pi * z / 8)
    assert np.allclose(u[1], uy, atol=1e-9)
    assert np.allclose(u[2], 0, atol=1e-16)


def _splitmix_reference(seed, counter):
    mask = (1 << 64) - 1
    zz = (seed * 0x9E3779B97F4A7C15 + counter * 0xD1B54A32D192ED03) & mask
    zz = ((zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9) & mask
    zz = ((zz ^ (zz >> 27)) * 0x94D049BB133111EB) & mask
    zz ^= zz >> 31
    return (zz >> 40) / float(1 << 24)


def test_noise_hash_against_python_integers():
    idx = np.array([0, 1, 2, 12345, 2 ** 25 - 1])
    for seed in (0, 1, 2 ** 40 + 3):
        for comp in range(3):
            got = O.hash_uniform(seed, idx, comp)
            ref = [_splitmix_reference(seed, int(i) * 3 + comp + 1) for i in idx]
            assert np.array_equal(got, ref)
    u = O.hash_uniform(0, np.arange(200000), 0)
    assert abs(u.mean() - 0.5) < 0.005 and u.min() >= 0 and u.max() < 1


def test_strouhal_fft_recovers_a_sinusoid():
    D, u0, every = 24.0, 0.08, 50
    st = 0.2
    t = np.arange(400) * every
    lift = np.sin(2 * np.pi * st * u0 / D * t)
    assert abs(O.strouhal_fft(lift, every, D, u0) - st) < 0.02


def test_oracle_solver_first_report_matches_construction():
    s = O.OracleSolver(24, 12, 8, "bgk", "cylinder", 100.0, 0.08, 6.0)
    m = s.monitors(1)
    assert np.isfinite(m["drag"]) and s.step_count == 0


# --- the C restatement (oracle/lbm_oracle.c) against the NumPy one --------------------------------
def _c_oracle():
    import subprocess
    from oracle import c_oracle
    if not c_oracle.available():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    return c_oracle


@pytest.mark.parametrize("collision", ["bgk", "trt", "les"])
@pytest.mark.parametrize("obstacle", ["cylinder", "sphere", "box", "field"])
@pytest.mark.parametrize("side", ["periodic", "slip", "lid"])
def test_c_oracle_matches_numpy_oracle(collision, obstacle, side):
    C_ = _c_oracle()
    dims = (24, 12, 9)
    n = 24 * 12 * 9
    f = O.W[:, None] + 0.01 * np.random.default_rng(9).standard_normal((19, n))
    phys = {"bgk": O.BGK(0.02), "trt": O.TRT(0.01, 0.25, 0.1), "les": O.LES(0.01)}[collision]
    phi = np.clip(np.random.default_rng(10).uniform(-0.5, 1.25, n), 0, 1).astype(np.float32)
    geo = {"cylinder": O.Cylinder(9.3, 6.2, 3.1), "sphere": O.Sphere(9.3, 6.2, 4.4, 3.1),
           "box": O.Box(0.4, 6.2, 4.4, 2.6, 2.1, 2.4), "field": O.Field(phi, dims)}[obstacle]
    side_obj = O.LidY(0.04, 0.01, -0.02) if side == "lid" else O.SIDES[side]()
    flow = O.Flow(0.08, 12.0, 1.0, 0.37, side_obj)
    d1, d2 = O.new_fields(n), O.new_fields(n)
    d1["ux"][:] = d2["ux"][:] = 0.01
    a = O.step(dims, phys, geo, flow, f, d1)
    b = C_.step(dims, phys, geo, flow, f, d2)
    assert np.abs(a - b).max() <= 1e-15
    for k in O.FIELDS:
        assert np.abs(d1[k] - d2[k]).max() <= 1e-15, k
    assert np.abs(d1["_rho"] - d2["_rho"]).max() <= 1e-15


def test_c_oracle_init_and_solver_match():
    C_ = _c_oracle()
    kw = dict(collision="trt", obstacle="sphere", reynolds=200.0, diameter=6.0, noise_amp=1e-3,
              noise_seed=3, lateral="slip", cs=0.1, magic=0.25)
    a = O.OracleSolver(32, 16, 12, **kw)
    b = C_.CSolver(32, 16, 12, **kw)
    assert np.abs(a.f - b.f).max() <= 1e-15
    for _ in range(3):
        a.step()
        b.step()
    a.step(report=True)
    b.step(report=True)
    assert np.abs(a.f - b.f).max() <= 1e-14
    ma, mb = a.monitors(4), b.monitors(4)
    for k in ma:
        assert abs(ma[k] - mb[k]) <= 1e-12 * max(1.0, abs(ma[k])), k


@pytest.mark.parametrize("st", [0.2213, 0.2, 0.1837])
def test_strouhal_refined_recovers_off_bin_frequencies(st):
    """A shedding-like record (fundamental + second harmonic + offset + noise), 6 periods, at
    frequencies between FFT bins: the refined estimate is within 0.2 %."""
    D, u0, every = 24.0, 0.08, 50
    period = D / (st * u0)
    t = np.arange(int(6 * period / every)) * every
    rng = np.random.default_rng(0)
    lift = (0.9 * np.sin(2 * np.pi * t / period + 0.3) + 0.2 * np.sin(4 * np.pi * t / period)
            + 0.05 + 0.01 * rng.standard_normal(len(t)))
    assert abs(O.strouhal_refined(lift, every, D, u0) / st - 1.0) < 2e-3
