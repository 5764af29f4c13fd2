"""The oracle's temperature variant (lbm.py:283-320, 457-492) pinned by closed forms: D3Q7
invariants, exact heat conservation (periodic box, adiabatic walls), the diffusivity of a
decaying sine mode, the steady linear profile between two fixed-temperature plates
(anti-bounce-back), the q_solid neighbour sum, and the HeatSink geometry by hand."""

import numpy as np
import pytest

from oracle import lbm_oracle as O


def test_d3q7_tables():
    assert abs(O.W7.sum() - 1.0) < 1e-15
    for a, ca in enumerate((O.CX7, O.CY7, O.CZ7)):
        for b, cb in enumerate((O.CX7, O.CY7, O.CZ7)):
            assert abs((O.W7 * ca * cb).sum() - (0.25 if a == b else 0.0)) < 1e-15   # c_s^2 = 1/4
    for q in range(7):
        o = O.OPP7[q]
        assert (O.CX7[o], O.CY7[o], O.CZ7[o]) == (-O.CX7[q], -O.CY7[q], -O.CZ7[q])


def test_temperature_equilibrium_moments():
    T, u = 1.3, (0.04, -0.02, 0.01)
    g = np.array([O.Heat.geq(q, T, *u) for q in range(7)])
    assert abs(g.sum() - T) < 1e-15
    for c, uc in zip((O.CX7, O.CY7, O.CZ7), u):
        assert abs((g * c).sum() - T * uc) < 1e-15


class _Geo:
    """Test geometry: solid and sink masks given per cell (flat), resting walls."""
    kind = "field"

    def __init__(self, solid, sink, dims):
        self.s, self.k = np.asarray(solid, np.float32), np.asarray(sink, np.float32)
        self.nx, self.ny, self.nz = dims

    def _at(self, a, x, y, z):
        return a[(np.asarray(z) * self.ny + np.asarray(y)) * self.nx + np.asarray(x)]

    def solid(self, x, y, z):
        return self._at(self.s, x, y, z)

    def sink(self, x, y, z):
        return self._at(self.k, x, y, z)

    def wall_vel(self, x, y, z, dt):
        zero = np.zeros_like(x)
        return zero, zero, zero


def _rest(n):
    one = np.ones(n)
    return np.stack([O.feq(q, one, 0 * one, 0 * one, 0 * one) for q in range(19)])


def _run(dims, geo, g, steps, alpha, t_solid, flow=None, f=None):
    n = dims[0] * dims[1] * dims[2]
    phys = O.BGK(0.05)
    heat = O.Heat(alpha, 0.05, t_solid)
    flow = flow or O.Flow(0.0, dims[0] - 1.0, 0.0, 0.0)
    f = _rest(n) if f is None else f
    for _ in range(steps):
        f, g = O.step_heat(dims, phys, geo, flow, f, g, heat)
    return f, g


def test_heat_is_conserved_in_a_periodic_box_and_between_adiabatic_walls():
    dims = (6, 7, 5)
    n = 6 * 7 * 5
    rng = np.random.default_rng(0)
    T = 1 + 0.2 * rng.standard_normal(n)
    g0 = np.stack([O.Heat.geq(q, T, 0, 0, 0) for q in range(7)])
    # a moving fluid (uniform u0 inflow state, no sponge): advection-diffusion conserves heat
    free = _Geo(np.zeros(n), np.zeros(n), dims)
    flow = O.Flow(0.05, dims[0] - 1.0, 0.0, 0.0)
    f = O.init(dims, free, flow, 0.0)
    _, g = _run(dims, free, g0, 5, 0.1, np.zeros(n), flow, f)
    assert abs(g.sum() - g0.sum()) < 1e-12
    # adiabatic (bounce-back) walls at y = 0 and ny - 1: the fluid's heat stays
    y = (np.arange(n) // 6) % 7
    wall = ((y == 0) | (y == 6)).astype(float)
    geo = _Geo(wall, np.zeros(n), dims)
    _, g = _run(dims, geo, g0 * (1 - wall), 8, 0.1, np.zeros(n))
    assert abs(g[:, wall == 0].sum() - (g0 * (1 - wall)).sum()) < 1e-12


def test_sine_mode_decays_at_the_diffusivity():
    """T = 1 + a sin(k x) at rest decays as exp(-alpha k^2 t) (k small: lattice terms ~ k^4)."""
    dims = (64, 2, 2)
    n = 256
    x = np.arange(n) % 64
    k = 2 * np.pi / 64
    alpha = 0.1
    T = 1 + 0.1 * np.sin(k * x)
    g0 = np.stack([O.Heat.geq(q, T, 0, 0, 0) for q in range(7)])
    steps = 150
    _, g = _run(dims, _Geo(np.zeros(n), np.zeros(n), dims), g0, steps, alpha, np.zeros(n))
    amp = np.abs(np.fft.rfft(g.sum(0)[:64]))[1] / 32
    rate = -np.log(amp / 0.1) / steps
    assert abs(rate / (alpha * k * k) - 1) < 0.01


def test_plates_at_fixed_temperature_give_a_linear_profile():
    """Sink rows y = 0 (T = 1) and y = ny - 1 (T = 0), anti-bounce-back halfway: the steady
    fluid profile is T(y) = 1 - (y - 1/2) / (ny - 2)."""
    dims = (3, 10, 3)
    n = 90
    y = (np.arange(n) // 3) % 10
    plate = ((y == 0) | (y == 9)).astype(float)
    ts = np.where(y == 0, 1.0, 0.0)
    g0 = np.zeros((7, n))
    _, g = _run(dims, _Geo(plate, plate, dims), g0, 1500, 0.25, ts)
    T = g.sum(0)
    fluid = plate == 0
    exact = 1 - (y - 0.5) / 8
    assert np.abs(T[fluid] - exact[fluid]).max() < 1e-9


def test_report_writes_temperature_and_the_sink_neighbour_sum():
    dims = (5, 5, 5)
    n = 125
    v = (2 * 5 + 2) * 5 + 2                       # the centre voxel is a sink at T = 2
    sink = np.zeros(n)
    sink[v] = 1
    rng = np.random.default_rng(3)
    T = 1 + 0.1 * rng.standard_normal(n)
    g0 = np.stack([O.Heat.geq(q, T, 0, 0, 0) for q in range(7)])
    ts = np.zeros(n)
    ts[v] = 2.0
    heat = O.Heat(0.1, 0.05, ts)
    out = {"temp": np.zeros(n), "q_solid": np.zeros(n)}
    O.step_heat(dims, O.BGK(0.05), _Geo(sink, sink, dims), O.Flow(0.0, 4.0, 0.0, 0.0), _rest(n),
                g0, heat, diag=O.new_fields(n), heat_out=out)
    assert out["temp"][v] == 2.0
    nb = [v + 1, v - 1, v + 5, v - 5, v + 25, v - 25]
    assert abs(out["q_solid"][v] - out["temp"][nb].sum()) < 1e-12
    assert np.count_nonzero(out["q_solid"]) == 1


def test_heatsink_geometry_by_hand():
    geo = O.make_obstacle("heatsink", 40, 24, 32, 12.0)   # x 8..19, z 4..27, base to y 6,
    cells = {(10, 3, 5): 1, (10, 10, 5): 1, (10, 10, 9): 0, (10, 17, 5): 0, (7, 3, 5): 0,
             (19, 16, 7): 1, (20, 3, 5): 0, (10, 6, 28): 0, (10, 6, 27): 1}
    for (x, y, z), want in cells.items():
        assert geo.sink(x, y, z) == want, (x, y, z)        # fins: z - 4 mod 12 < 4, y < 17
    assert geo.solid(5, 0, 5) == 1 and geo.sink(5, 0, 5) == 0
    assert geo.solid(5, 23, 5) == 1 and geo.solid(5, 12, 5) == 0
