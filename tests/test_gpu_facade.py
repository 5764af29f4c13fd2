"""The host facade driven the way the reference's own scripts drive lbm.py.

Each test reproduces one call shape a tack-lbm caller uses (trt_ref.py's hand-built step with a
bare Solver for the heat arguments and ManagedArray.tack views, bench.py's and cht.py's reads of
solver fields through ``.numpy``), plus the Solver's two field layouts and its graph replays
against the step-by-step path.
"""

import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stream():
    import tlbm
    st = ctypes.c_void_p()
    tlbm._capi.check(tlbm._capi.lib().tlbm_stream_create(ctypes.byref(st)), "stream")
    return st.value


def test_direct_step_with_bare_solver_heat_and_managed_views():
    """trt_ref.py:27-38: plugins built by hand, report fields as ``ManagedArray(n).tack`` views,
    and heat, g_in and g_out taken from ``Solver.__new__(Solver)`` after ``_setup_heat(n)``; the
    step must reproduce the reference's known answer (the committed trt_ref fixture)."""
    import tlbm
    d = np.load(os.path.join(ROOT, "tests", "golden", "trt_ref_8x6x4.npz"))
    nx, ny, nz = (int(v) for v in d["dims"])
    lat = tlbm.D3Q19(nx, ny, nz)
    n = nx * ny * nz
    views = {k: tlbm.ManagedArray(n).tack for k in tlbm.FIELDS}
    a_in = tlbm.array(d["f_in"])
    a_out = tlbm.empty((19 * n,))
    bare = tlbm.Solver.__new__(tlbm.Solver)
    bare.heat_enabled = False
    bare.nx = nx
    bare._setup_heat(n)
    assert bare.heat.enabled == 0 and len(bare.g) == 2
    tlbm.lbm_step(lat, tlbm.TRT(float(d["nu"])), tlbm.Cylinder(-100.0, -100.0, 0.0),
                  tlbm.Flow(0.0, 1e9, 0.0), tlbm.Quiet(views), a_in, a_out, bare.heat,
                  bare.g[0], bare.g[1])
    got = a_out.numpy()
    assert float(np.abs(got - d["f_out"]).max()) < 1e-6
    # a report step through the same views writes the managed memory behind them
    tlbm.lbm_step(lat, tlbm.TRT(float(d["nu"])), tlbm.Cylinder(-100.0, -100.0, 0.0),
                  tlbm.Flow(0.0, 1e9, 0.0), tlbm.Report(views), a_in, a_out, bare.heat,
                  bare.g[0], bare.g[1])
    tlbm.sync()
    managed = views["ux"]._owner.numpy
    assert np.array_equal(managed, views["ux"].numpy()) and np.abs(managed).max() > 0


def test_solver_fields_read_through_numpy_attribute():
    """bench.py:29: ``ux = s.fields["ux"].numpy`` then ``ux.min()``: a host array of the field
    the last report step wrote (stream synchronised), read-only because it is a copy."""
    import tlbm
    st = _stream()
    s = tlbm.Solver(64, 32, 16, collision="les", obstacle="cylinder", reynolds=200.0,
                    diameter=8.0, noise_amp=1e-3, stream=st)
    for _ in range(9):
        s.step()
    s.step(report=True)
    m = s.monitors(10)
    ux = s.fields["ux"].numpy
    assert ux.shape == (s.lat.n,) and ux.dtype == np.float32
    assert 0.0 < float(ux.max()) < 0.2 and float(ux.min()) > -0.2 and not np.isnan(ux).any()
    assert np.array_equal(ux, tlbm.DeviceArray.numpy(s.fields["ux"]))
    with pytest.raises(ValueError):
        ux[0] = 1.0
    # the FP64 monitor sums are the sums of these fields
    assert abs(float(s.fields["fx"].numpy.astype(np.float64).sum())
               / (0.5 * s.u0 ** 2 * s.diameter * s.nz) - m["drag"]) < 1e-9 * max(1, abs(m["drag"]))


def test_heat_fields_in_the_conjugate_heat_call_shape():
    """cht.py:72 writes ``heat_fields["t_solid"].numpy[voxels]`` and cht.py:85 reads
    ``heat_fields["q_solid"].numpy[voxels]`` after a report step and ``monitors``."""
    import tlbm
    st = _stream()
    s = tlbm.Solver(48, 24, 16, collision="les", obstacle="heatsink", reynolds=200.0,
                    diameter=6.0, heat=True, stream=st)
    voxels = np.arange(0, s.lat.n, 7)
    s.heat_fields["t_solid"].numpy[voxels] = 1.0           # managed: host writes reach the kernel
    for _ in range(4):
        s.step()
    s.step(report=True)
    s.monitors(5)
    q = s.heat_fields["q_solid"].numpy[voxels].astype(np.float64)
    assert np.array_equal(q, s.q_solid()[voxels])
    assert q.max() > 0.0                                    # fluid next to the hot sink heated up
    temp = s.output["temp"].numpy
    assert np.array_equal(temp, s.heat_fields["temp"].numpy)


@pytest.mark.parametrize("stride", [1, 2])
def test_managed_field_memory_matches_device_fields(stride):
    """``Solver(field_memory="managed")`` (the reference's layout: the kernel writes managed
    fields, which are the output arrays at stride 1) steps, reports and monitors exactly as the
    default HBM layout does."""
    import tlbm
    kw = dict(nx=96, ny=48, nz=24, collision="trt", obstacle="sphere", reynolds=300.0,
              diameter=10.0, lateral="slip", noise_amp=1e-3, stride=stride)
    a = tlbm.Solver(**kw, field_memory="managed")
    b = tlbm.Solver(**kw)
    if stride == 1:
        assert all(a.output[k] is a.fields[k] for k in a.OUTPUT)
    for rep in range(3):
        for s in (a, b):
            for _ in range(6):
                s.step()
            s.step(report=True)
        ma, mb = a.monitors(7), b.monitors(7)
        assert ma == mb, rep
        assert np.array_equal(a.populations(), b.populations())
        for k in tlbm.FIELDS:
            assert np.array_equal(a.fields[k].numpy, b.fields[k].numpy), (rep, k)
        for k in a.OUTPUT:
            assert np.array_equal(a.output[k].numpy, b.output[k].numpy), (rep, k)


def test_output_arrays_refresh_when_read_not_in_monitors():
    """At stride 1 with HBM fields, ``monitors`` moves no field; an output array brings itself
    up to date when read after a report step, keeping its address (zero-copy readers)."""
    import tlbm
    st = _stream()
    s = tlbm.Solver(64, 32, 16, obstacle="sphere", reynolds=300.0, diameter=8.0,
                    noise_amp=1e-3, stream=st)
    out = s.output["ux"]
    view = out.numpy
    assert np.array_equal(view, s.fields["ux"].numpy)       # the construction report step
    for _ in range(5):
        s.step()
    s.step(report=True)
    s.monitors(6)
    assert s._output_stale                                   # monitors copied nothing
    before = view.copy()
    assert out.numpy is view                                 # same array, same address
    assert np.array_equal(view, s.fields["ux"].numpy) and not np.array_equal(view, before)
    assert not s._output_stale


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_graph_replays_after_an_odd_number_of_steps(storage):
    """run / one eager report step / run again: the second replay starts from the buffer the
    eager step wrote (graphs are keyed on the ping-pong order), matching step-by-step runs."""
    import tlbm
    st = _stream()

    def make():
        return tlbm.Solver(80, 40, 24, collision="les", obstacle="cylinder", reynolds=250.0,
                           diameter=8.0, noise_amp=1e-3, storage=storage, stream=st)
    g, e = make(), make()
    schedule = [("run", 20, 10), ("step", True), ("run", 20, 10), ("step", False),
                ("step", False), ("step", True), ("run", 20, 10), ("run", 20, 10)]
    for item in schedule:
        if item[0] == "run":
            g.run(item[1], report_every=item[2])
            for i in range(item[1]):
                e.step(report=(i + 1) % item[2] == 0)
        else:
            g.step(report=item[1])
            e.step(report=item[1])
        assert g.step_count == e.step_count
        assert g.monitors(10) == e.monitors(10), item
        assert np.array_equal(g.populations(), e.populations()), item
    assert len(g._graphs) == 2                               # one graph per buffer order


def test_graph_keeps_its_report_workspace_when_the_stream_outgrows_it():
    """A small solver's graph (report steps included) is captured; a larger solver's eager
    report step on the same stream then needs a larger reduction workspace.  The graph keeps
    the buffer it was captured with, so its replays still give the step-by-step sums."""
    import tlbm
    st = _stream()
    small = tlbm.Solver(48, 24, 16, collision="bgk", obstacle="cylinder", reynolds=100.0,
                        diameter=6.0, noise_amp=1e-3, stream=st)
    ref = tlbm.Solver(48, 24, 16, collision="bgk", obstacle="cylinder", reynolds=100.0,
                      diameter=6.0, noise_amp=1e-3, stream=st)
    small.run(10, report_every=5)
    for i in range(10):
        ref.step(report=(i + 1) % 5 == 0)
    big = tlbm.Solver(160, 96, 64, collision="les", obstacle="sphere", reynolds=300.0,
                      diameter=12.0, noise_amp=1e-3, stream=st)  # grows the workspace
    for _ in range(3):
        big.step(report=True)
        small.run(10, report_every=5)
        for i in range(10):
            ref.step(report=(i + 1) % 5 == 0)
        assert small.monitors(5) == ref.monitors(5)
        assert np.array_equal(small.populations(), ref.populations())
    assert np.isfinite(big.monitors(1)["drag"])


def test_launch_shape_belongs_to_one_solver():
    """set_launch on one solver leaves another solver of the same specialisation at the default
    launch; both step to the same state (the cell update does not depend on the block size)."""
    import tlbm
    kw = dict(nx=96, ny=32, nz=16, collision="les", obstacle="sphere", reynolds=300.0,
              diameter=8.0, noise_amp=1e-3)
    a, b = tlbm.Solver(**kw), tlbm.Solver(**kw)
    a.set_launch(64)
    assert a.block_x == 64 and b.block_x == 0
    for s in (a, b):
        for _ in range(4):
            s.step()
        s.step(report=True)
    assert np.array_equal(a.populations(), b.populations())
    ma, mb = a.monitors(5), b.monitors(5)
    for k in ma:
        assert abs(ma[k] - mb[k]) <= 1e-12 * max(1.0, abs(mb[k])), k
    with pytest.raises(ValueError):
        a.set_launch(100)
