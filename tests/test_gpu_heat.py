"""The temperature variant of the step (lbm.py:457-492) against the FP64 oracle, through the
C ABI: one step of every collision x obstacle x side x diagnostics combination from seeded
populations, temperatures and solid temperatures; a multi-step run of the heat-sink duct; and
the Solver's heat wiring (lbm.py:586-604, 639-647)."""

import numpy as np
import pytest

from cases import Case, random_fields, random_populations
from oracle import lbm_oracle as O

pytestmark = pytest.mark.gpu

DIMS = (24, 12, 9)


def _random_g(n, seed):
    """7 temperature populations around the equilibrium of T = 1 + 0.1 N, u = 0.05 N."""
    rng = np.random.default_rng(seed)
    T = 1 + 0.1 * rng.standard_normal(n)
    u = 0.05 * rng.standard_normal((3, n))
    g = np.stack([O.Heat.geq(q, T, *u) for q in range(7)])
    return g + 1e-3 * rng.standard_normal((7, n))


def _run(case, f64, g64, ts, report, steps=1, alpha=0.004, nu=0.006):
    import tlbm
    lat, phys, geo, flow = case.device()
    n = lat.n
    f = [tlbm.array(f64.astype(np.float32)), tlbm.empty((19 * n,))]
    g = [tlbm.array(g64.astype(np.float32)), tlbm.empty((7 * n,))]
    t_solid = tlbm.array(ts.astype(np.float32))
    q_solid = tlbm.zeros((n,))
    temp = tlbm.zeros((n,))
    heat = tlbm.Heat({"t_solid": t_solid, "q_solid": q_solid, "temp": temp}, alpha, nu)
    old = random_fields(n, 5)
    fields = {k: tlbm.array(old[k]) for k in tlbm.FIELDS}
    diag = tlbm.Report(fields) if report else tlbm.Quiet(fields)
    sums = tlbm.HostBuffer(14)
    for _ in range(steps):
        if report:
            q_solid.zero_()
        tlbm.lbm_step(lat, phys, geo, flow, diag, f[0], f[1], heat, g[0], g[1], sums=sums)
        f.reverse()
        g.reverse()
    tlbm.sync()
    got = (f[0].numpy().reshape(19, n), g[0].numpy().reshape(7, n), temp.numpy(), q_solid.numpy())
    # oracle, FP64, from the same FP32-rounded inputs
    o_phys, o_geo, o_flow = case.oracle()
    of = f64.astype(np.float32).astype(np.float64)
    og = g64.astype(np.float32).astype(np.float64)
    o_heat = O.Heat(float(np.float32(alpha)), float(np.float32(nu)),
                    ts.astype(np.float32).astype(np.float64))
    o_fields = {k: v.astype(np.float64) for k, v in old.items()} if report else None
    out = {"temp": np.zeros(n), "q_solid": np.zeros(n)}
    for _ in range(steps):
        out["q_solid"][:] = 0.0
        of, og = O.step_heat(case.dims, o_phys, o_geo, o_flow, of, og, o_heat, o_fields, out)
    return got, (of, og, out["temp"], out["q_solid"])


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


COMBOS = [(c, o, s) for c in ("bgk", "trt", "les")
          for o in ("cylinder", "sphere", "box", "field", "heatsink")
          for s in ("periodic", "slip", "lid")]


# rows of 24 cells (waves straddle rows) and of 64 (waves within a row)
@pytest.mark.parametrize("nx", [24, 64])
@pytest.mark.parametrize("report", [False, True])
@pytest.mark.parametrize("collision,obstacle,side", COMBOS)
def test_heat_step_matches_oracle(collision, obstacle, side, report, nx):
    case = Case((nx,) + DIMS[1:], collision, obstacle, side)
    n = case.n
    f = random_populations(n, seed=21)
    g = _random_g(n, seed=22)
    ts = 0.5 + np.random.default_rng(23).uniform(size=n)
    (gf, gg, gt, gq), (rf, rg, rt, rq) = _run(case, f, g, ts, report)
    assert _rel(gf, rf) < 1e-6
    assert _rel(gg, rg) < 2e-6
    if report:
        assert _rel(gt, rt) < 2e-6
        if np.abs(rq).max() > 0:
            assert _rel(gq, rq) < 2e-6          # FP32 atomics in any order
        else:
            assert np.all(gq == 0)


@pytest.mark.parametrize("side", ["periodic", "slip"])
def test_heat_sink_across_periodic_faces(side):
    from cases import WRAP_GEOMETRIES
    case = Case((20, 12, 8), "les", "heatsink", side, geometry=WRAP_GEOMETRIES["heatsink"])
    n = case.n
    (gf, gg, gt, gq), (rf, rg, rt, rq) = _run(case, random_populations(n, 24), _random_g(n, 25),
                                              np.full(n, 1.5), True)
    assert _rel(gf, rf) < 1e-6 and _rel(gg, rg) < 2e-6 and _rel(gq, rq) < 2e-6


def test_heat_sink_duct_over_many_steps():
    """30 report steps of the heat-sink duct (LES): FP32 against FP64."""
    case = Case((32, 12, 10), "les", "heatsink", "slip", u0=0.05, sponge=(0.75, 1.0))
    n = case.n
    o_phys, o_geo, o_flow = case.oracle()
    f = O.init(case.dims, o_geo, o_flow, 0.0)
    g = np.zeros((7, n))
    (gf, gg, gt, gq), (rf, rg, rt, rq) = _run(case, f, g, np.ones(n), True, steps=30)
    assert _rel(gf, rf) < 1e-5
    assert _rel(gt, rt) < 1e-5
    assert _rel(gq, rq) < 1e-5
    assert rt.max() > 0.05                                  # heat entered the fluid


def test_solver_heat_wiring():
    """heat_check.py's setup, smaller: a heat-sink duct at fixed wall temperature stays bounded,
    q_solid is cleared before every report step, temp reaches the output arrays."""
    import tlbm
    s = tlbm.Solver(48, 24, 24, collision="les", obstacle="heatsink", reynolds=500.0, heat=True,
                    diameter=6.0)
    s.set_wall_temperature(1.0)
    for k in range(3):
        for _ in range(49):
            s.step()
        s.step(report=True)
        m = s.monitors(50)
        q = s.q_solid()
        assert np.isfinite(m["drag"]) and q.max() > 0
    s.step(report=True)                                  # one step later: about the same sums,
    q2 = s.q_solid()                                     # not twice them (cleared each report)
    assert abs(q2.sum() / q.sum() - 1) < 0.2
    temp = s.output["temp"].numpy
    # at cell Peclet ~60 the D3Q7 BGK scheme over- and undershoots near fronts; heat_check.py
    # tracks the fraction of cells below -0.01 for the same reason
    assert np.all(np.isfinite(temp)) and temp.min() > -0.5 and temp.max() < 1.5
    assert (temp < -0.05).mean() < 0.05
    # solid cells (duct walls, base, fins) report their t_solid exactly (lbm.py:483)
    geo = O.make_obstacle("heatsink", 48, 24, 24, 6.0)
    solid = geo.solid(*O.coords(48, 24, 24)) == 1.0
    assert solid.any() and np.all(temp[solid] == 1.0)
    with pytest.raises(ValueError):
        tlbm.Solver(16, 12, 8, obstacle="heatsink", heat=True, storage="f16")


def test_solver_heat_with_padded_planes():
    """The temperature planes follow the populations' plane stride: a padded qstride gives the
    same populations, temperature planes and cell temperatures, bit for bit, as the unpadded
    one.  q_solid is a sum of FP32 atomic adds whose order differs from run to run, so it is
    compared to within that rounding."""
    import tlbm
    out = []
    for qstride in (None, 48 * 24 * 16 + 320):
        s = tlbm.Solver(48, 24, 16, collision="les", obstacle="heatsink", reynolds=500.0,
                        heat=True, diameter=6.0, qstride=qstride)
        s.set_wall_temperature(1.0)
        for _ in range(9):
            s.step()
        s.step(report=True)
        s.update_output()
        tlbm.sync()
        g = s.temperature_populations()
        out.append((s.populations(), g, s.output["temp"].numpy.copy(), s.q_solid()))
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)
    qa, qb = out[0][3], out[1][3]
    assert np.abs(qa).max() > 0
    assert np.abs(qa - qb).max() <= 1e-6 * np.abs(qa).max()


@pytest.mark.parametrize("nx", [64, 60])
def test_large_heat_lattice_repeats_the_small_one(nx):
    """A temperature step on a lattice large enough for the non-temporal-store kernel (nx x 64 x
    1024, 26 FP32 planes of about 4 M cells in each buffer) whose state, solid temperatures and
    geometry (a z-axis cylinder) repeat every 4 cells along z gives, bit for bit, the step of
    the nx x 64 x 4 lattice (plain stores) repeated 256 times.  nx = 64: waves within rows;
    nx = 60: waves and blocks straddling rows."""
    import tlbm
    reps = 256
    small = Case((nx, 64, 4), "les", "cylinder", "slip", geometry=(20.3, 30.2, 9.1))
    big = Case((nx, 64, 4 * reps), "les", "cylinder", "slip", geometry=(20.3, 30.2, 9.1))
    assert 2 * 26 * 4 * big.n > 512 * 2 ** 20
    f64 = random_populations(small.n, 41).astype(np.float32).reshape(19, 1, small.n)
    g64 = _random_g(small.n, 42).astype(np.float32).reshape(7, 1, small.n)
    ts = (0.5 + np.random.default_rng(43).uniform(size=small.n)).astype(np.float32)
    out = []
    for case, r in ((small, 1), (big, reps)):
        lat, phys, geo, flow = case.device()
        n = case.n
        f = [tlbm.array(np.broadcast_to(f64, (19, r, small.n))), tlbm.empty((19 * n,))]
        g = [tlbm.array(np.broadcast_to(g64, (7, r, small.n))), tlbm.empty((7 * n,))]
        heat = tlbm.Heat({"t_solid": tlbm.array(np.broadcast_to(ts, (r, small.n))),
                          "q_solid": tlbm.zeros((n,)), "temp": tlbm.zeros((n,))}, 0.004, 0.006)
        fields = {k: tlbm.zeros((n,)) for k in tlbm.FIELDS}
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Quiet(fields), f[0], f[1], heat, g[0], g[1])
        tlbm.sync()
        out.append((f[1].numpy().reshape(19, r, small.n), g[1].numpy().reshape(7, r, small.n)))
    (sf, sg), (bf, bg) = out
    assert np.array_equal(bf, np.broadcast_to(sf, bf.shape))
    assert np.array_equal(bg, np.broadcast_to(sg, bg.shape))
