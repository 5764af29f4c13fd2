"""The tiled population layout on the GPU: every path of the step (16-bit interior waves whose
source rows are addressed per wave, the per-lane path, walls, lid and slip faces, report fields
and FP64 sums, the temperature variant, non-temporal stores, graph replays, the no-allocation
entry point and the access-pattern copy) gives bit for bit what the q-major layout gives, whose
parity with the oracle the other GPU tests establish.  Same arithmetic, different addresses:
any address slip shows up as a mismatch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # collision, obstacle, lateral, dims, tile rows
    ("les", "sphere", "lid", (128, 12, 10), 4),        # rows of whole waves: per-wave rows
    ("trt", "cylinder", "periodic", (128, 10, 6), 16),  # last tile part-filled (60 rows)
    ("bgk", "box", "slip", (96, 10, 9), 4),            # rows of 1.5 waves: per-lane path
    ("les", "cylinder", "periodic", (64, 8, 8), 1),
]


def _pair(collision, obstacle, lateral, dims, k, storage, **kw):
    import tlbm
    nx, ny, nz = dims
    common = dict(collision=collision, obstacle=obstacle, reynolds=200.0, diameter=6.0,
                  lateral=lateral, storage=storage, noise_amp=1e-3, noise_seed=3, **kw)
    a = tlbm.Solver(nx, ny, nz, tile_rows=k, **common)
    b = tlbm.Solver(nx, ny, nz, tile_rows=0, **common)
    assert a.tile_rows == k and b.tile_rows == 0
    return a, b


def _same_state(a, b):
    import tlbm
    tlbm.sync()
    assert np.array_equal(a.populations(), b.populations())
    for name in tlbm.FIELDS:
        assert np.array_equal(a.fields[name].numpy, b.fields[name].numpy), name
    assert np.array_equal(a.sums.numpy, b.sums.numpy)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}-{c[4]}" for c in CASES])
def test_tiled_solver_steps_as_the_planes_layout(case, storage):
    print(f"[tiles] {case} {storage}", flush=True)
    a, b = _pair(*case, storage)
    if case[1] == "cylinder":          # moving-wall term on
        a.set_actuation(0.01)
        b.set_actuation(0.01)
    _same_state(a, b)                  # lbm_init + the construction's report step
    for s in (a, b):
        for i in range(5):
            s.step(report=(i == 2))
    _same_state(a, b)


@pytest.mark.parametrize("dims,k", [((128, 24, 16), 4), ((256, 128, 96), 32)])
def test_tiled_heat_solver_steps_as_the_planes_layout(dims, k):
    """Populations and temperature planes tiled in one solver, q-major in the other; the larger
    lattice (the default tiles, 26 FP32 planes of 3.1 M cells per buffer) takes the
    non-temporal-store kernels.  q_solid is a sum of FP32 atomic adds whose order may differ."""
    import tlbm
    print(f"[tiles] heat sink {dims} {k}", flush=True)
    if dims[0] >= 256:
        assert tlbm.default_tile_rows(*dims) == k
        assert 2 * 26 * 4 * dims[0] * dims[1] * dims[2] > 512 * 2 ** 20
    a, b = _pair("les", "heatsink", "periodic", dims, k, "f32", heat=True)
    for s in (a, b):
        s.set_wall_temperature(1.0)
        for i in range(4):
            s.step(report=(i == 3))
    tlbm.sync()
    _same_state(a, b)
    assert a.g[0].size == tlbm.population_elems(a.lat, None, k, 7)
    assert not np.array_equal(a.g[0].numpy(), b.g[0].numpy())  # the layouts differ ...
    assert np.array_equal(a.temperature_populations(), b.temperature_populations())  # ... the planes not
    assert np.array_equal(a.heat_fields["temp"].numpy, b.heat_fields["temp"].numpy)
    qa, qb = a.q_solid(), b.q_solid()
    assert np.abs(qa).max() > 0
    assert np.abs(qa - qb).max() <= 1e-6 * np.abs(qa).max()


@pytest.mark.parametrize("storage,dims", [("f32", (256, 128, 128)), ("f16", (256, 256, 128))])
def test_default_tiles_on_a_large_lattice(storage, dims):
    """A lattice large enough for the default tiles and the non-temporal-store kernels."""
    import tlbm
    print(f"[tiles] default tiles {storage} {dims}", flush=True)
    k = tlbm.default_tile_rows(*dims, storage)
    assert k == (32 if storage == "f32" else 16)
    a, b = _pair("les", "sphere", "lid", dims, k, storage)
    for s in (a, b):
        for i in range(4):
            s.step(report=(i == 3))
    _same_state(a, b)


def test_tiled_graph_replay_matches_eager_steps():
    import tlbm
    from tlbm import _capi
    print("[tiles] graph replay", flush=True)
    st = ctypes.c_void_p()
    _capi.check(_capi.lib().tlbm_stream_create(ctypes.byref(st)), "stream")
    # tiles replayed from a graph against q-major planes stepped eagerly
    g, e = _pair("trt", "cylinder", "periodic", (128, 10, 6), 4, "f16", stream=st.value)
    for _ in range(3):
        g.run(4, 2)
        for i in range(4):
            e.step(report=(i % 2 == 1))
        _same_state(g, e)


def test_tiled_no_alloc_entry_point_matches_tlbm_step():
    import tlbm
    from tlbm import _capi
    print("[tiles] no-alloc entry point", flush=True)
    L = _capi.lib()
    s = tlbm.Solver(128, 12, 10, collision="les", obstacle="sphere", reynolds=200.0,
                    diameter=6.0, lateral="lid", storage="f16", noise_amp=1e-3, tile_rows=4)
    p = tlbm.make_params(s.lat, s.phys, s.geo, s.flow, s.qstride, s.tile_rows)
    h = s._handle(report=True)
    # one zeroed field set per call: a report step's residuals read the fields it overwrites
    fs = [{k: tlbm.DeviceArray(s.lat.n).zero_() for k in tlbm.FIELDS} for _ in range(2)]
    outs = [tlbm.empty((s.f[0].size,), "f16") for _ in range(2)]
    sums = [tlbm.HostBuffer(_capi.TLBM_NSUMS) for _ in range(2)]
    _capi.check(L.tlbm_step(h, s.f[0].ptr, outs[0].ptr, ctypes.byref(p),
                            ctypes.byref(tlbm.lbm._fields_struct(fs[0])), sums[0].ptr, None),
                "step")
    _capi.check(L.tlbm_step_noalloc(h, s.f[0].ptr, outs[1].ptr, ctypes.byref(p),
                                    ctypes.byref(tlbm.lbm._fields_struct(fs[1])), sums[1].ptr,
                                    None), "step_noalloc")
    tlbm.sync()
    assert np.array_equal(outs[0].numpy(), outs[1].numpy())
    assert np.array_equal(sums[0].numpy, sums[1].numpy)
    for k in tlbm.FIELDS:
        assert np.array_equal(fs[0][k].numpy(), fs[1][k].numpy()), k


@pytest.mark.parametrize("storage,dims,k", [("f32", (128, 10, 8), 4), ("f16", (128, 10, 8), 4),
                                            ("f16", (96, 7, 5), 2)])
def test_tiled_pull_copy_moves_what_the_planes_copy_moves(storage, dims, k):
    import tlbm
    from tlbm import _capi
    print(f"[tiles] pull copy {storage} {dims} {k}", flush=True)
    L = _capi.lib()
    lat = tlbm.D3Q19(*dims)
    rng = np.random.default_rng(11)
    # raw element bits (compared as integers: random 16-bit patterns include NaNs)
    bits = np.uint32 if storage == "f32" else np.uint16
    planes = rng.integers(0, 1 << (8 * np.dtype(bits).itemsize), (19, lat.n),
                          dtype=np.uint64).astype(bits)
    geo, flow = tlbm.Cylinder(-100.0, -100.0, 1.0), tlbm.Flow(0.05, 1e9, 0.0)
    code = tlbm.STORAGES[storage]
    outs = []
    for tiles in (0, k):
        src = planes.reshape(-1) if tiles == 0 else tlbm.tile_populations(planes, lat, tiles)
        a = tlbm.empty((src.size,), storage)
        a.copy_from(src.view(a.dtype))
        b = tlbm.empty((src.size,), storage)
        p = tlbm.make_params(lat, None, geo, flow, lat.n, tiles)
        _capi.check(L.tlbm_pull_copy(a.ptr, b.ptr, ctypes.byref(p), code, None), "pull copy")
        tlbm.sync()
        raw = b.numpy().view(bits)
        outs.append(raw.reshape(19, lat.n) if tiles == 0
                    else tlbm.untile_populations(raw, lat, tiles))
    assert np.array_equal(outs[0], outs[1])
    # and the copy is the pull: plane q of cell (x, y, z) comes from (x, y, z) - c_q (periodic)
    _check_pull(outs[0], planes, dims)


def _check_pull(out, planes, dims):
    import tlbm
    nx, ny, nz = dims
    for q in range(19):
        want = np.roll(planes[q].reshape(nz, ny, nx), (tlbm.CZ[q], tlbm.CY[q], tlbm.CX[q]),
                       axis=(0, 1, 2))
        assert np.array_equal(out[q].reshape(nz, ny, nx), want), q


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_large_pull_copy_is_the_pull(storage):
    """512 x 256 x 32 in tiles of 4 rows: population buffers beyond 512 MB (FP32), so the copy
    takes the quiet step's large-lattice build (non-temporal stores and unshifted loads, rows
    addressed per wave); every plane equals the periodic pull of the input."""
    import tlbm
    from tlbm import _capi
    dims, k = (512, 256, 32), 4
    print(f"[tiles] large pull copy {storage} {dims} {k}", flush=True)
    L = _capi.lib()
    lat = tlbm.D3Q19(*dims)
    rng = np.random.default_rng(12)
    bits = np.uint32 if storage == "f32" else np.uint16
    planes = rng.integers(0, 1 << (8 * np.dtype(bits).itemsize), (19, lat.n),
                          dtype=np.uint64).astype(bits)
    src = tlbm.tile_populations(planes, lat, k)
    a = tlbm.empty((src.size,), storage)
    a.copy_from(src.view(a.dtype))
    b = tlbm.empty((src.size,), storage)
    p = tlbm.make_params(lat, None, tlbm.Cylinder(-100.0, -100.0, 1.0), tlbm.Flow(0.05, 1e9, 0.0),
                         lat.n, k)
    _capi.check(L.tlbm_pull_copy(a.ptr, b.ptr, ctypes.byref(p), tlbm.STORAGES[storage], None),
                "pull copy")
    tlbm.sync()
    _check_pull(tlbm.untile_populations(b.numpy().view(bits), lat, k), planes, dims)
