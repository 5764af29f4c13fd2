"""Parity of the HIP step (through the C ABI) with the CPU oracle and the reference's known answer.

Tolerances:
  * FP32 storage, one step: max |gpu - oracle| <= 1e-6 * max |oracle| (FP32 arithmetic against
    the FP64 restatement; the reference's own bar is 1e-5 absolute, trt_ref.py:60);
  * FP16/BF16 storage: the stored deviation may differ from the oracle's exactly rounded
    deviation by one unit in the last place (the FP32 pre-rounding value can straddle a
    rounding boundary) on top of the FP32 arithmetic error, i.e.
    |gpu - oracle| <= |oracle - w| * 2^-10 (f16) / 2^-7 (bf16) + 2^-24 + 1e-6 max |oracle|.
"""

import os

import numpy as np
import pytest

from oracle import lbm_oracle as O
from cases import Case, WRAP_GEOMETRIES, random_fields, random_populations

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def test_trt_known_answer():
    """trt_ref.py: one TRT(0.002) step on 8x6x4, periodic, no obstacle, no sponge."""
    import tlbm
    d = np.load(os.path.join(ROOT, "tests", "golden", "trt_ref_8x6x4.npz"))
    nx, ny, nz = (int(v) for v in d["dims"])
    lat = tlbm.D3Q19(nx, ny, nz)
    fields = {k: tlbm.ManagedArray(lat.n) for k in tlbm.FIELDS}
    a_in = tlbm.array(d["f_in"])
    a_out = tlbm.empty((19 * lat.n,))
    tlbm.lbm_step(lat, tlbm.TRT(float(d["nu"])), tlbm.Cylinder(-100.0, -100.0, 0.0),
                  tlbm.Flow(0.0, 1e9, 0.0), tlbm.Quiet(fields), a_in, a_out)
    got = a_out.numpy()
    assert float(np.abs(got - d["f_out"]).max()) < float(d["tol"])
    assert _rel(got, d["f_out"]) < 1e-6


def _run_step(case, f64, report, storage="f32", seed=3, block_x=0):
    """One step on the GPU and in the oracle from the same populations (and old fields),
    launched with `block_x` threads per block (0 = the library's default)."""
    import tlbm
    lat, phys, geo, flow = case.device()
    diag_kind = tlbm.Report({k: None for k in tlbm.FIELDS}) if report else \
        tlbm.Quiet({k: None for k in tlbm.FIELDS})
    tlbm.set_launch(phys, geo, diag_kind, flow.side, storage, block_x)
    try:
        return _run_step_inner(case, f64, report, storage, seed, lat, phys, geo, flow)
    finally:
        tlbm.set_launch(phys, geo, diag_kind, flow.side, storage, 0)


def _run_step_inner(case, f64, report, storage, seed, lat, phys, geo, flow):
    import tlbm
    raw = tlbm.encode_populations(f64.astype(np.float32), storage)
    f_in_dev = tlbm.array(raw, storage)
    f_out_dev = tlbm.empty((19 * lat.n,), storage)
    old = random_fields(lat.n, seed)
    fields = {k: tlbm.ManagedArray(lat.n) for k in tlbm.FIELDS}
    for k in tlbm.FIELDS:
        fields[k].numpy[:] = old[k]
    sums = tlbm.HostBuffer(14)
    diag = tlbm.Report(fields) if report else tlbm.Quiet(fields)
    tlbm.lbm_step(lat, phys, geo, flow, diag, f_in_dev, f_out_dev, sums=sums)
    tlbm.sync()
    got = tlbm.decode_populations(f_out_dev.numpy().reshape(19, lat.n), storage)
    o_phys, o_geo, o_flow = case.oracle()
    o_in = tlbm.decode_populations(raw, storage).astype(np.float64)
    o_fields = {k: v.astype(np.float64) for k, v in old.items()} if report else None
    ref = O.step(case.dims, o_phys, o_geo, o_flow, o_in, o_fields)
    gfields = {k: fields[k].numpy.copy() for k in tlbm.FIELDS}
    return got, ref, gfields, o_fields, sums.numpy.copy()


COMBOS = [(c, o, s) for c in ("bgk", "trt", "les")
          for o in ("cylinder", "sphere", "box", "field", "heatsink")
          for s in ("periodic", "slip", "lid")]


# Rows of 24 cells (waves straddle rows) and of 64 (waves within a row).
MAPPING_NX = [24, 64]


@pytest.mark.parametrize("nx", MAPPING_NX)
@pytest.mark.parametrize("block_x", [0, 64])
@pytest.mark.parametrize("collision,obstacle,side", COMBOS)
def test_quiet_step_f32(collision, obstacle, side, block_x, nx):
    case = Case((nx, 12, 9), collision, obstacle, side)
    f = random_populations(case.n, seed=11)
    got, ref, *_ = _run_step(case, f, report=False, block_x=block_x)
    assert _rel(got, ref) < 1e-6


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("dims", [(256, 3, 2), (512, 3, 2), (1024, 2, 2), (96, 4, 3), (16, 5, 4),
                                  (320, 3, 2), (200, 2, 3)])
@pytest.mark.parametrize("side", ["periodic", "lid"])
@pytest.mark.parametrize("block_x", [0, 64, 192, 256])
def test_rows_across_blocks(storage, dims, side, block_x):
    """Rows of one to sixteen blocks (x streaming across block edges and the periodic wrap) and
    rows shorter than a block; the obstacle straddles x = 0."""
    case = Case(dims, "les", "sphere", side, geometry=(1.3, 1.4, 1.2, 2.2))
    f = random_populations(case.n, seed=21)
    got, ref, *_ = _run_step(case, f, report=True, storage=storage, block_x=block_x)
    if storage == "f32":
        assert _rel(got, ref) < 1e-6
    else:
        bits = 10 if storage == "f16" else 7
        w = np.array(O.W)[:, None]
        bound = np.abs(ref - w) * 2.0 ** -bits + 2.0 ** -24 + 1e-6 * np.abs(ref).max()
        assert bool(np.all(np.abs(got - ref) <= bound))


@pytest.mark.parametrize("nx", MAPPING_NX)
@pytest.mark.parametrize("block_x", [0, 64])
@pytest.mark.parametrize("collision,obstacle,side", COMBOS)
def test_report_step_f32(collision, obstacle, side, block_x, nx):
    case = Case((nx, 12, 9), collision, obstacle, side)
    f = random_populations(case.n, seed=12)
    got, ref, g, o, sums = _run_step(case, f, report=True, block_x=block_x)
    assert _rel(got, ref) < 1e-6
    for k in O.FIELDS:
        scale = max(float(np.abs(o[k]).max()), 1e-12)
        assert float(np.abs(g[k] - o[k]).max()) <= 2e-5 * scale + 1e-9, k
    # in-kernel FP64 sums equal FP64 sums of the fields the kernel wrote
    for i, k in enumerate(O.FIELDS):
        exact = float(g[k].astype(np.float64).sum())
        assert abs(sums[i] - exact) <= 1e-12 * max(1.0, float(np.abs(g[k]).astype(np.float64).sum())), k
    # per-cell rho is FP32 on the GPU, FP64 in the oracle: bound by FP32 rounding per cell
    assert abs(sums[13] - float(o["_rho"].sum())) < 2e-7 * case.n


@pytest.mark.parametrize("obstacle", ["cylinder", "sphere", "box", "heatsink"])
@pytest.mark.parametrize("side", ["periodic", "slip"])
def test_obstacle_across_periodic_x(obstacle, side):
    dims = (20, 12, 8) if obstacle == "heatsink" else (20, 11, 8)
    case = Case(dims, "les", obstacle, side, geometry=WRAP_GEOMETRIES[obstacle])
    f = random_populations(case.n, seed=13)
    got, ref, *_ = _run_step(case, f, report=True)
    assert _rel(got, ref) < 1e-6


@pytest.mark.parametrize("dims", [(70, 5, 3), (130, 3, 2), (257, 2, 2), (1, 4, 3), (3, 1, 1)])
def test_ragged_and_degenerate_grids(dims):
    """x extents that are not multiples of the block, and 1-cell-thick axes (self-periodic)."""
    case = Case(dims, "trt", "sphere", "periodic", geometry=(1.5, 1.2, 1.1, 1.3))
    f = random_populations(case.n, seed=14)
    got, ref, g, o, sums = _run_step(case, f, report=True)
    assert _rel(got, ref) < 1e-6


def _edge_reps(nx, storage):
    """The most repeats of an nx x 64 x 4 slab whose two population buffers (19 planes of n
    cells, packed) still span under 2^32 bytes together with one plane and a row: the largest
    lattice the kernels address with 32-bit buffer offsets."""
    size = 4 if storage == "f32" else 2
    small = nx * 64 * 4
    reps = ((1 << 32) // size - nx) // (20 * small)
    assert (20 * small * reps + nx) * size < (1 << 32) <= (20 * small * (reps + 1) + nx) * size
    return reps


@pytest.mark.parametrize("nx", [64, 60])
@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("size", ["beyond", "edge"])
def test_planes_beyond_32_bit_offsets(storage, nx, size):
    """A lattice of 2^27 cells (64 x 64 x 32768; with nx = 60, 15/16 of that): its population
    planes end beyond 2^31 elements and beyond 2^32 bytes ("beyond": per-lane 64-bit addresses);
    or the largest lattice whose buffers span under 2^32 bytes ("edge": one buffer resource,
    plane offsets up to 4 GB).  State, old fields and geometry (a rotating z-axis cylinder)
    repeat every 4 cells along z, so a report step and then a quiet step must give, bit for bit,
    the steps of the nx x 64 x 4 lattice repeated, and FP64 sums that many times that lattice's.
    The large lattice's quiet step takes the non-temporal-store kernel, the small one's the
    plain one; with nx = 60 blocks of 256 cells straddle rows and planes."""
    import tlbm
    reps = 8192 if size == "beyond" else _edge_reps(nx, storage)
    small = Case((nx, 64, 4), "les", "cylinder", "slip", geometry=(20.3, 30.2, 9.1))
    big = Case((nx, 64, 4 * reps), "les", "cylinder", "slip", geometry=(20.3, 30.2, 9.1))
    if size == "beyond":
        assert 19 * big.n > 2 ** 31
    raw = tlbm.encode_populations(random_populations(small.n, seed=31).astype(np.float32),
                                  storage).reshape(19, 1, small.n)
    old = random_fields(small.n, 32)
    results = []
    for case, r in ((small, 1), (big, reps)):
        lat, phys, geo, flow = case.device()
        f_in = tlbm.array(np.broadcast_to(raw, (19, r, small.n)), storage)
        f_out = tlbm.empty((19 * case.n,), storage)
        fields = {k: tlbm.array(np.broadcast_to(old[k], (r, small.n))) for k in tlbm.FIELDS}
        sums = tlbm.HostBuffer(14)
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Report(fields), f_in, f_out, sums=sums)
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Quiet(fields), f_out, f_in)
        tlbm.sync()
        results.append((f_out, f_in, fields, sums.numpy.copy()))
    (s_out, s_q, s_fields, s_sums), (b_out, b_q, b_fields, b_sums) = results
    for big_f, small_f in ((b_out, s_out), (b_q, s_q)):
        assert np.array_equal(big_f.numpy().reshape(19, reps, small.n),
                              np.broadcast_to(small_f.numpy().reshape(19, 1, small.n),
                                              (19, reps, small.n)))
    for k in tlbm.FIELDS:
        assert np.array_equal(b_fields[k].numpy().reshape(reps, small.n),
                              np.broadcast_to(s_fields[k].numpy(), (reps, small.n))), k
    assert np.all(np.abs(b_sums - reps * s_sums) <= 1e-9 * np.abs(reps * s_sums))


def test_report_steps_on_two_streams_keep_their_own_sums():
    """Report steps queued at once on two streams (a large lattice on one, a small one on the
    other) give bit for bit the FP64 sums each gives when run alone."""
    import ctypes
    import tlbm
    L = tlbm._capi.lib()
    streams = []
    for _ in range(2):
        st = ctypes.c_void_p()
        tlbm._capi.check(L.tlbm_stream_create(ctypes.byref(st)), "stream")
        streams.append(st.value)

    def solvers():
        return (tlbm.Solver(256, 128, 96, collision="les", obstacle="sphere", reynolds=500.0,
                            noise_amp=1e-3, stream=streams[0]),
                tlbm.Solver(96, 48, 32, collision="trt", obstacle="cylinder", reynolds=200.0,
                            noise_amp=1e-3, stream=streams[1]))
    together, alone = [], []
    a, b = solvers()
    for _ in range(5):
        a.step(report=True)
        b.step(report=True)
        for st in streams:
            tlbm.sync(st)
        together.append((a.sums.numpy.copy(), b.sums.numpy.copy()))
    a2, b2 = solvers()
    for _ in range(5):
        a2.step(report=True)
        tlbm.sync(streams[0])
        b2.step(report=True)
        tlbm.sync(streams[1])
        alone.append((a2.sums.numpy.copy(), b2.sums.numpy.copy()))
    for (ta, tb), (sa, sb) in zip(together, alone):
        assert np.array_equal(ta, sa) and np.array_equal(tb, sb)
    del a, b, a2, b2
    for st in streams:
        L.tlbm_stream_destroy(ctypes.c_void_p(st))


@pytest.mark.parametrize("storage,bits", [("f16", 10), ("bf16", 7)])
@pytest.mark.parametrize("collision", ["bgk", "trt", "les"])
@pytest.mark.parametrize("side", ["periodic", "slip", "lid"])
@pytest.mark.parametrize("nx,obstacle", [(24, "cylinder"), (64, "cylinder"), (130, "cylinder"),
                                         (24, "field")])
def test_shifted_storage(storage, bits, collision, side, nx, obstacle):
    case = Case((nx, 12, 9), collision, obstacle, side)
    f = random_populations(case.n, seed=15)
    got, ref, *_ = _run_step(case, f, report=False, storage=storage)
    w = np.array(O.W)[:, None]
    bound = np.abs(ref - w) * 2.0 ** -bits + 2.0 ** -24 + 1e-6 * np.abs(ref).max()
    assert bool(np.all(np.abs(got - ref) <= bound))


def test_solid_cells_hold_rest_equilibrium():
    """sigma = 1 in solids (lbm.py:449): f_out = feq(1, 0) = w_q exactly, deviation 0."""
    import tlbm
    case = Case((24, 12, 9), "les", "sphere", "periodic")
    f = random_populations(case.n, seed=16)
    for storage in ("f32", "f16", "bf16"):
        lat, phys, geo, flow = case.device()
        raw = tlbm.encode_populations(f.astype(np.float32), storage)
        out = tlbm.empty((19 * lat.n,), storage)
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Quiet({k: None for k in tlbm.FIELDS}),
                      tlbm.array(raw, storage), out)
        x, y, z = O.coords(*case.dims)
        solid = O.Sphere(*case.geometry).solid(x, y, z) > 0
        got = out.numpy().reshape(19, lat.n)[:, solid]
        if storage == "f32":
            assert np.array_equal(got, np.repeat(np.float32(O.W)[:, None].astype(np.float32),
                                                 solid.sum(), axis=1))
        else:
            assert not got.any()


@pytest.mark.parametrize("obstacle", ["cylinder", "sphere", "box", "field"])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_init(obstacle, storage):
    """lbm.py:383-394 with the seeded velocity noise."""
    import tlbm
    case = Case((40, 14, 10), obstacle=obstacle)
    lat, phys, geo, flow = case.device()
    f = tlbm.empty((19 * lat.n,), storage)
    tlbm.lbm_init(lat, geo, flow, f, 0.05, noise_amp=1e-3, noise_seed=42)
    got = tlbm.decode_populations(f.numpy().reshape(19, lat.n), storage)
    _, o_geo, o_flow = case.oracle()
    ref = O.init(case.dims, o_geo, o_flow, 0.05, 1e-3, 42)
    if storage == "f32":
        assert _rel(got, ref) < 1e-6
    else:
        bits = 10 if storage == "f16" else 7
        w = np.array(O.W)[:, None]
        bound = np.abs(ref - w) * 2.0 ** -bits + 2.0 ** -24 + 1e-6 * np.abs(ref).max()
        assert bool(np.all(np.abs(got - ref) <= bound))


def test_coarsen():
    import tlbm
    lat = tlbm.D3Q19(13, 9, 7)
    src = np.random.default_rng(0).standard_normal(lat.n).astype(np.float32)
    s = 3
    out_dims = tuple((d + s - 1) // s for d in (13, 9, 7))
    dst = tlbm.zeros((int(np.prod(out_dims)),))
    tlbm.coarsen(lat, s, tlbm.array(src), dst, *out_dims)
    assert np.array_equal(dst.numpy(), O.coarsen((13, 9, 7), s, src, out_dims))


def test_step_rejects_bad_arguments():
    import tlbm
    case = Case((24, 12, 9))
    lat, phys, geo, flow = case.device()
    a = tlbm.empty((19 * lat.n,))
    with pytest.raises(RuntimeError, match="distinct"):
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Quiet({k: None for k in tlbm.FIELDS}), a, a)
    with pytest.raises(ValueError):
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Quiet({k: None for k in tlbm.FIELDS}), a,
                      tlbm.empty((19 * lat.n - 19,)))


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_solver_run_graph_replays_match_eager_steps(storage):
    """Solver.run replays a captured hipGraph: every replay (with eager work and syncs between
    replays) leaves the state, report fields and FP64 sums that the same steps taken one by one
    leave; after set_physics the next run captures the new collision."""
    import ctypes
    import tlbm
    L = tlbm._capi.lib()
    st = ctypes.c_void_p()
    tlbm._capi.check(L.tlbm_stream_create(ctypes.byref(st)), "stream")

    def make():
        return tlbm.Solver(96, 48, 32, collision="les", obstacle="sphere", reynolds=400.0,
                           lateral="lid", noise_amp=1e-3, storage=storage, stream=st.value)
    g, e = make(), make()
    for rep in range(4):
        if rep == 2:
            g.set_physics("trt", 250.0)
            e.set_physics("trt", 250.0)
        g.run(20, report_every=10)
        for i in range(20):
            e.step(report=(i + 1) % 10 == 0)
        assert g.step_count == e.step_count == 20 * (rep + 1)
        mg, me = g.monitors(10), e.monitors(10)        # eager work + sync between replays
        assert mg == me, rep
        assert np.array_equal(g.populations(), e.populations()), rep
        for k in tlbm.FIELDS:
            assert np.array_equal(g.fields[k].numpy, e.fields[k].numpy), (rep, k)
    assert len(g._graphs) == 1                       # the TRT graph; the LES one was dropped
    with pytest.raises(ValueError):
        g.run(3)
    with pytest.raises(ValueError):
        tlbm.Solver(16, 8, 8).run(2)                 # no created stream
    del g, e
    L.tlbm_stream_destroy(st)


def test_capture_refuses_what_a_graph_must_not_hold():
    """Under capture, tlbm_memset and a report step whose stream has no reduction workspace yet
    are refused with a message (nothing allocates or records a memset node); the capture still
    ends cleanly, and the stream works eagerly afterwards."""
    import ctypes
    import tlbm
    L = tlbm._capi.lib()
    st = ctypes.c_void_p()
    tlbm._capi.check(L.tlbm_stream_create(ctypes.byref(st)), "stream")
    case = Case((24, 12, 9), "bgk", "cylinder", "periodic")
    lat, phys, geo, flow = case.device()
    f = [tlbm.array(random_populations(case.n, 51).astype(np.float32)), tlbm.empty((19 * case.n,))]
    buf = tlbm.empty((64,))
    sums = tlbm.HostBuffer(14)
    fields = {k: tlbm.zeros((case.n,)) for k in tlbm.FIELDS}
    out = ctypes.c_void_p()
    tlbm._capi.check(L.tlbm_capture_begin(st), "capture begin")
    assert L.tlbm_memset(buf.ptr, 0, 256, st) != 0
    assert "capture" in tlbm._capi.last_error()
    with pytest.raises(RuntimeError, match="eagerly"):
        tlbm.lbm_step(lat, phys, geo, flow, tlbm.Report(fields), f[0], f[1], sums=sums,
                      stream=st.value)
    tlbm._capi.check(L.tlbm_capture_end(st, ctypes.byref(out)), "capture end")
    L.tlbm_graph_destroy(out)
    # eager use of the same stream: the report step sizes the workspace and runs
    tlbm.lbm_step(lat, phys, geo, flow, tlbm.Report(fields), f[0], f[1], sums=sums,
                  stream=st.value)
    tlbm.sync(st.value)
    assert np.isfinite(sums.numpy).all() and sums.numpy[13] > 0
    L.tlbm_stream_destroy(st)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_quiet_chunk_orders_step_as_plain_order(storage):
    """Every chunk order (blocks to XCDs in runs of 1..64, grids rounded up to whole runs, the
    extra slots idle) and block size steps every cell exactly as plain order does; 750 blocks
    of 256 cells is not a multiple of any run of 8 groups."""
    import tlbm
    kw = dict(collision="les", obstacle="sphere", reynolds=400.0, diameter=10.0,
              noise_amp=1e-3, storage=storage)
    ref = tlbm.Solver(200, 40, 24, **kw)
    ref.set_launch(256, 1)
    for _ in range(3):
        ref.step()
    want = ref.populations()
    for block, group in ((256, 2), (256, 4), (256, 64), (64, 8), (128, 2), (192, 16), (0, 0)):
        print(f"[order] {storage} block {block} group {group}", flush=True)
        s = tlbm.Solver(200, 40, 24, **kw)
        s.set_launch(block, group)
        for _ in range(3):
            s.step()
        assert np.array_equal(s.populations(), want), (block, group)
    with pytest.raises(ValueError):
        s.set_launch(256, 3)


def test_solver_autotune_keeps_state_and_caches(tmp_path, monkeypatch):
    """tune_solver times each block size on the solver's own buffers without changing its
    state, applies the fastest, and answers the second call from its cache; it refuses a
    solver with the temperature variant."""
    import tlbm
    from tlbm import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    s = tlbm.Solver(200, 40, 24, collision="les", obstacle="sphere", reynolds=400.0,
                    diameter=10.0, noise_amp=1e-3)
    before = s.populations()
    block, ms, group = autotune.tune_solver(s, reps=2)
    assert block in (64, 128, 192, 256) and group in (1, 2, 4) and ms > 0
    assert (s.block_x, s.xcd_group) == (block, group)
    assert np.array_equal(s.populations(), before)
    assert autotune.tune_solver(s, reps=2)[::2] == (block, group)
    s.step()                                  # the tuned launch steps as the default does
    ref = tlbm.Solver(200, 40, 24, collision="les", obstacle="sphere", reynolds=400.0,
                      diameter=10.0, noise_amp=1e-3)
    ref.step()
    assert np.array_equal(s.populations(), ref.populations())
    h = tlbm.Solver(48, 24, 16, obstacle="heatsink", heat=True, diameter=6.0)
    with pytest.raises(NotImplementedError):
        autotune.tune_solver(h)
