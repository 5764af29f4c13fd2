"""Batched lattices (parameter sweeps): one launch steps many domains of different sizes.

Each domain of a batch must step as the single-lattice kernel steps it (the same per-cell
code: bit for bit for BGK and LES on MI355X; TRT differs by one unit in the last place where
the compiler contracts a product into an FMA differently in the two kernels) and as the FP64
oracle steps it (to FP32 rounding); the per-domain report sums must equal the single-lattice
ones; every launch shape and the hipGraph replay must give the same state as eager steps.
"""

import ctypes

import numpy as np
import pytest

from oracle import lbm_oracle as O

pytestmark = pytest.mark.gpu


def _domains():
    from tlbm.batch import Domain
    # uneven sizes (rows shorter and longer than a block, odd extents), Re and radii
    return [Domain(20, 10, 10, 150.0, 5.0), Domain(64, 32, 32, 400.0, 9.0),
            Domain(37, 12, 8, 900.0, 6.5), Domain(130, 16, 12, 250.0, 8.0),
            Domain(8, 6, 4, 100.0, 3.0)]


def _view(raw, batch, d):
    """Domain d's (19, n) populations inside a raw batch buffer (q-major planes of the batch's
    plane stride, or tiles of rows: then a copy)."""
    import tlbm
    e, _ = batch.offsets(d)
    dom, k = batch.domains[d], batch.tile_rows[d]
    if k:
        lat = tlbm.D3Q19(dom.nx, dom.ny, dom.nz)
        return tlbm.untile_populations(raw[e:e + tlbm.population_elems(lat, None, k)], lat, k)
    qs = batch.plane_strides[d]
    return raw[e:e + 19 * qs].reshape(19, qs)[:, :dom.n]


def _put(raw, batch, d, planes):
    """Write domain d's (19, n) populations into a raw batch buffer in the batch's layout."""
    import tlbm
    dom, k = batch.domains[d], batch.tile_rows[d]
    if k:
        e, _ = batch.offsets(d)
        lat = tlbm.D3Q19(dom.nx, dom.ny, dom.nz)
        tiled = tlbm.tile_populations(planes, lat, k)
        raw[e:e + tiled.size] = tiled
    else:
        _view(raw, batch, d)[:] = planes


def _random_state(batch, seed):
    import tlbm
    rng = np.random.default_rng(seed)
    raw = np.zeros(batch.elements, batch.f[0].dtype)
    for d, dom in enumerate(batch.domains):
        rho = 1 + 0.01 * rng.standard_normal(dom.n)
        u = 0.05 * rng.standard_normal((3, dom.n))
        f = np.stack([O.feq(q, rho, *u) for q in range(19)]) + 1e-3 * rng.standard_normal(
            (19, dom.n))
        _put(raw, batch, d, tlbm.encode_populations(f.astype(np.float32), batch.storage))
    batch.f[0].copy_from(raw)
    return raw


def _single_step(batch, d, raw, report):
    """Domain d stepped alone by the single-lattice kernel from the same populations."""
    import tlbm
    dom = batch.domains[d]
    lat = tlbm.D3Q19(dom.nx, dom.ny, dom.nz)
    a = tlbm.array(np.ascontiguousarray(_view(raw, batch, d)).reshape(-1), batch.storage)
    out = tlbm.empty((19 * dom.n,), batch.storage)
    fields = {k: tlbm.zeros((dom.n,)) for k in tlbm.FIELDS}
    sums = tlbm.HostBuffer(14)
    flow = tlbm.Flow(batch.u0, 0.75 * dom.nx, 1.0, batch._omega, tlbm.PeriodicY())
    diag = tlbm.Report(fields) if report else tlbm.Quiet(fields)
    tlbm.lbm_step(lat, batch.phys[d], batch.geo[d], flow, diag, a, out, sums=sums)
    tlbm.sync()
    return out.numpy(), sums.numpy.copy()


# population layouts: padded planes, packed planes, tiles of 4 rows in every domain, and the
# default (tiles of batch_tile_rows rows at FP32, padded planes at FP16/BF16)
LAYOUTS = {"pad256": (256, 0), "packed": (0, 0), "tiles4": (256, 4), "auto": (256, "auto")}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("collision", ["bgk", "trt", "les"])
@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_batch_step_matches_single_lattice_and_oracle(storage, collision, layout):
    import tlbm
    from tlbm.batch import BatchSolver
    plane_pad, tile_rows = LAYOUTS[layout]
    b = BatchSolver(_domains(), collision=collision, storage=storage, cs=0.1, magic=0.25,
                    plane_pad=plane_pad, tile_rows=tile_rows)
    if layout == "auto":
        assert b.tile_rows == [tlbm.batch.batch_tile_rows(d.nx, storage) for d in b.domains]
        assert any(b.tile_rows) == (storage == "f32")
    raw = _random_state(b, seed=5)
    b.step(report=True)
    tlbm.sync()
    got = b.f[0].numpy()
    sums = b.sums.numpy.reshape(len(b.domains), 14)
    for d, dom in enumerate(b.domains):
        single, ssums = _single_step(b, d, raw, report=True)
        a32 = tlbm.decode_populations(_view(got, b, d), storage)
        s32 = tlbm.decode_populations(single.reshape(19, -1), storage)
        diff = float(np.abs(a32 - s32).max())
        print(f"domain {d}: max |batch - single| = {diff:.3e}")
        # one unit in the last place: of f for FP32, of the stored deviation f - w otherwise
        dev = float(np.abs(s32 - np.array(O.W, np.float32)[:, None]).max())
        ulp = 2.0 ** -23 * float(np.abs(s32).max()) if storage == "f32" else \
            dev * (2.0 ** -10 if storage == "f16" else 2.0 ** -7)
        assert diff <= 2 * ulp, (d, diff)
        # report sums: the single kernel used old fields = 0; the batch ran without fields, so
        # compare the sums that do not involve residuals (u, p, |u|, forces, mass)
        for k in (0, 1, 2, 3, 4, 10, 11, 12, 13):
            assert abs(sums[d, k] - ssums[k]) <= 1e-9 * max(1.0, abs(ssums[k])), (d, k)
        if storage == "f32":
            g = b.geo[d]
            geo = O.Cylinder(g.x, g.y, g.r)
            phys = {"bgk": O.BGK, "trt": lambda nu: O.TRT(nu, 0.25, 0.1),
                    "les": lambda nu: O.LES(nu, 0.1)}[collision](b.u0 * dom.diameter / dom.reynolds)
            ref = O.step((dom.nx, dom.ny, dom.nz), phys, geo, O.Flow(b.u0, 0.75 * dom.nx, 1.0),
                         _view(raw, b, d).astype(np.float64))
            err = np.abs(_view(got, b, d) - ref).max() / np.abs(ref).max()
            assert err < 1e-6, (d, err)


@pytest.mark.parametrize("block,grid", [(64, 0), (128, 0), (256, 0), (64, 7), (256, 3),
                                        (128, 1000), (192, 0), (192, 5)])
@pytest.mark.parametrize("group", [1, 3, 8])
def test_batch_launch_shapes_agree(block, grid, group):
    """Quiet and report steps give the same state and fields at every launch shape and chunk
    order (runs of `group` consecutive chunks per XCD; persistent grids rounded to 8 blocks),
    and the same per-domain FP64 sums up to summation order."""
    import tlbm
    from tlbm.batch import BatchSolver
    out = []
    for shape, order in (((0, 0), 1), ((block, grid), group)):
        b = BatchSolver(_domains(), collision="trt", storage="f32", fields=True)
        b.set_launch(*shape)
        b.set_order(order)
        _random_state(b, seed=6)
        b.step()
        b.step(report=True)
        tlbm.sync()
        out.append((b.f[0].numpy(), {k: v.numpy() for k, v in b.fields.items()},
                    b.sums.numpy.copy()))
        b.close()
    (f0, fl0, s0), (f1, fl1, s1) = out
    assert np.array_equal(f1, f0)
    for k in fl0:
        assert np.array_equal(fl1[k], fl0[k]), k
    assert np.allclose(s1, s0, rtol=1e-12, atol=1e-12 * np.abs(s0).max())


def test_batch_graph_replays_match_eager_steps():
    """Each replay of the captured graph (with eager work and syncs between replays) leaves the
    state that the same number of eager steps leaves."""
    import tlbm
    from tlbm.batch import BatchSolver
    L = tlbm._capi.lib()
    stream = ctypes.c_void_p()
    tlbm._capi.check(L.tlbm_stream_create(ctypes.byref(stream)), "stream")
    g = BatchSolver(_domains(), collision="les", storage="f16", stream=stream.value,
                    noise_amp=1e-3)
    g.set_order(4)  # graphs hold the chunk order they were captured with
    e = BatchSolver(_domains(), collision="les", storage="f16", noise_amp=1e-3)
    for replay in range(3):
        g.run(4, report_every=2)
        for i in range(4):
            e.step(report=(i % 2 == 1))
        tlbm.sync(stream.value)
        tlbm.sync()
        assert np.array_equal(g.f[0].numpy(), e.f[0].numpy()), replay
        assert np.allclose(g.sums.numpy, e.sums.numpy, rtol=1e-12, atol=1e-12), replay
        e.f[1].zero_()  # eager work between replays
    g.close()
    L.tlbm_stream_destroy(stream)


def test_batch_autotune_keeps_state(tmp_path, monkeypatch):
    import tlbm
    from tlbm import autotune
    from tlbm.batch import BatchSolver
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    b = BatchSolver(_domains(), collision="trt", storage="f32")
    _random_state(b, seed=7)
    before = b.f[0].numpy()
    block, grid, ms = autotune.tune_batch(b, reps=2)
    assert block in (64, 128, 192, 256) and ms > 0
    assert np.array_equal(b.f[0].numpy(), before)
    again = autotune.tune_batch(b, reps=2)
    assert again[:2] == (block, grid)  # from the cache


def test_batch_layout_autotune(tmp_path, monkeypatch):
    """The layout tuner builds one batch per tile-rows candidate, keeps the fastest (tuned) one
    and closes the others; the kept batch steps as a fresh batch of its layout with its tuned
    launch shape."""
    import tlbm
    from tlbm import autotune
    from tlbm.batch import BatchSolver
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    built = []

    def make(tile_rows):
        b = BatchSolver(_domains(), collision="trt", storage="f32", tile_rows=tile_rows)
        built.append(b)
        return b
    batch, cand, (block, grid, ms) = autotune.tune_batch_layout(make, (0, 4, "auto"), reps=2)
    assert cand in (0, 4, "auto") and ms > 0 and len(built) == 3
    assert all(b._batch is None for b in built if b is not batch)   # the others are closed
    ref = make(cand)
    ref.set_launch(block, grid)
    ref.set_order(batch.group)
    for b in (batch, ref):
        _random_state(b, seed=8)
        b.step()
        b.step(report=True)
    tlbm.sync()
    assert batch.tile_rows == ref.tile_rows
    assert np.array_equal(batch.f[0].numpy(), ref.f[0].numpy())


def test_sweep_domains_are_seeded_and_in_range():
    from tlbm.batch import sweep_domains
    a = sweep_domains(64, seed=3)
    b = sweep_domains(64, seed=3)
    assert [(d.nx, d.reynolds, d.diameter) for d in a] == [(d.nx, d.reynolds, d.diameter) for d in b]
    for d in a:
        assert 64 <= d.nx <= 192 and d.ny == d.nz == d.nx // 2
        assert 100 <= d.reynolds <= 1000 and 8 <= d.diameter <= 24


def test_large_batch_domain_steps_as_the_single_lattice():
    """A batch of one 256 x 256 x 300 domain (19,660,800 cells, beyond 2^24: the index split
    divides by invariant integers, exact for any 32-bit index) plus a small one steps as the
    single-lattice kernel (LES, FP32, from the same seeded initial state): within 2 ulp, the
    bar of the batch parity test (the two kernels may contract the sponge blend of q = 0 into
    FMAs differently), and bit for bit outside the sponge."""
    import tlbm
    from tlbm.batch import BatchSolver, Domain
    big = Domain(256, 256, 300, 3000.0, 32.0, center=(96.0, 128.0, 150.0))
    assert big.n > 2 ** 24
    b = BatchSolver([Domain(20, 10, 10, 150.0, 5.0), big], collision="les", obstacle="sphere",
                    lateral="lid", storage="f32", noise_amp=1e-3)
    b.step()
    tlbm.sync()
    got = b.populations(1)
    assert b.tile_rows[1] == 16  # 16 rows of 256 cells: 16 KiB of a plane per tile
    s = tlbm.Solver(256, 256, 300, collision="les", obstacle="sphere", reynolds=3000.0,
                    diameter=32.0, center=(96.0, 128.0, 150.0), lateral="lid", noise_amp=1e-3,
                    qstride=b.plane_strides[1])
    # the Solver takes one report step at construction (lbm.py:583-584); restart it from the
    # batch's initial state
    tlbm.lbm_init(s.lat, s.geo, s.flow, s.f[0], 0.05, 1e-3, 0, tile_rows=s.tile_rows)
    s.step()
    tlbm.sync()
    ref = s.populations()
    assert np.abs(got - ref).max() <= 2 * 2.0 ** -23 * np.abs(ref).max()
    x = np.arange(big.n) % big.nx
    fluid_side = x < 0.75 * big.nx                    # upstream of the sponge
    assert np.array_equal(got[:, fluid_side], ref[:, fluid_side])
    b.close()
