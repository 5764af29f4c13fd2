"""The Catalyst channel tree of a Solver (run_lbm.py:38-121): the Conduit paths the reference
builds, zero-copy vertex fields, SI scaling, the monitor history, with and without heat."""

import numpy as np
import pytest


class _Managed:
    def __init__(self, n, value):
        self.numpy = np.full(n, value, np.float32)


class _FakeSolver:
    collision_name, reynolds, u0, diameter, stride = "les", 300.0, 0.08, 24.0, 2
    out_dims = (4, 3, 2)
    omega_rot = 0.04

    def __init__(self, heat):
        names = ("ux", "uy", "uz", "p", "vmag") + (("temp",) if heat else ())
        self.output = {k: _Managed(24, i) for i, k in enumerate(names)}


# run_lbm.py:62-115: every leaf the reference sets for the grid and monitor channels
GRID = ["catalyst/channels/grid/type",
        "catalyst/channels/grid/data/coordsets/coords/type",
        *[f"catalyst/channels/grid/data/coordsets/coords/dims/{a}" for a in "ijk"],
        *[f"catalyst/channels/grid/data/coordsets/coords/origin/{a}" for a in "xyz"],
        *[f"catalyst/channels/grid/data/coordsets/coords/spacing/d{a}" for a in "xyz"],
        "catalyst/channels/grid/data/topologies/mesh/type",
        "catalyst/channels/grid/data/topologies/mesh/coordset",
        "catalyst/channels/grid/data/fields/velocity/association",
        "catalyst/channels/grid/data/fields/velocity/topology",
        *[f"catalyst/channels/grid/data/fields/velocity/values/{a}" for a in "xyz"],
        *[f"catalyst/channels/grid/data/fields/{f}/{leaf}" for f in ("pressure", "speed")
          for leaf in ("association", "topology", "values")]]
STATE = ["catalyst/state/timestep", "catalyst/state/time", "catalyst/state/parameters"]


@pytest.mark.parametrize("heat", [False, True])
def test_channel_tree_has_the_reference_paths_and_zero_copy_fields(heat):
    from tlbm.catalyst import MONITOR_NAMES, CatalystChannels, _get, paths
    s = _FakeSolver(heat)
    ch = CatalystChannels(s)
    tree = ch.execute(100, 0.33, {"drag": 1.5, "lift": 0.1})
    want = STATE + GRID + ["catalyst/channels/monitors/type"] + \
        [f"catalyst/channels/monitors/data/values/{n}" for n in ("iteration", *MONITOR_NAMES)]
    if heat:
        want += [f"catalyst/channels/grid/data/fields/temperature/{leaf}"
                 for leaf in ("association", "topology", "values")]
    assert paths(tree) == sorted(want)
    mesh = "catalyst/channels/grid/data"
    assert _get(tree, f"{mesh}/fields/velocity/values/x") is s.output["ux"].numpy
    assert _get(tree, f"{mesh}/fields/speed/values") is s.output["vmag"].numpy
    assert _get(tree, f"{mesh}/coordsets/coords/spacing/dx") == 2 / 24.0
    assert [_get(tree, f"{mesh}/coordsets/coords/dims/{a}") for a in "ijk"] == [4, 3, 2]
    assert _get(tree, "catalyst/state/parameters") == ["collision=les", "reynolds=300.0",
                                                       "rotation=0.5"]
    assert _get(tree, f"{mesh}/fields/velocity/association") == "vertex"


def test_scales_history_and_steering_values():
    from tlbm.catalyst import CatalystChannels, _get, steering_monitors
    s = _FakeSolver(False)
    ch = CatalystChannels(s, scales={"length": 0.01, "pressure": (2.0, 1.0)}, monitor_history=2)
    for step in (10, 20, 30):
        m = steering_monitors(s, {"drag": step / 10.0}, step)
        tree = ch.execute(step, m["time"], m)
    mesh = "catalyst/channels/grid/data"
    assert _get(tree, f"{mesh}/coordsets/coords/spacing/dy") == 2 * 0.01
    assert np.all(_get(tree, f"{mesh}/fields/pressure/values") == 2.0 * 3 + 1.0)
    vals = "catalyst/channels/monitors/data/values"
    assert list(_get(tree, f"{vals}/iteration")) == [20.0, 30.0]
    assert list(_get(tree, f"{vals}/drag")) == [2.0, 3.0]
    assert list(_get(tree, f"{vals}/collision")) == [2.0, 2.0]
    assert abs(_get(tree, f"{vals}/time")[-1] - 30 * 0.08 / 24.0) < 1e-15
    assert np.isnan(_get(tree, f"{vals}/energy")).all()


def test_to_conduit_needs_conduit():
    from tlbm.catalyst import to_conduit
    try:
        import conduit  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            to_conduit({"a": 1})


@pytest.mark.gpu
def test_channels_of_a_running_solver():
    import tlbm
    from tlbm.catalyst import CatalystChannels, _get, steering_monitors
    from tlbm.derived import DerivedFields
    s = tlbm.Solver(48, 24, 16, collision="les", obstacle="cylinder", reynolds=300.0, stride=2)
    d = DerivedFields(s)
    ch = CatalystChannels(s, derived=d)
    for _ in range(20):
        s.step()
    s.step(report=True)
    m = steering_monitors(s, s.monitors(20), s.step_count, d)
    tree = ch.execute(s.step_count, m["time"], m)
    mesh = "catalyst/channels/grid/data"
    ux = _get(tree, f"{mesh}/fields/velocity/values/x")
    assert np.shares_memory(ux, s.output["ux"].numpy) and ux.size == 24 * 12 * 8
    assert np.abs(ux).max() > 0
    vort = _get(tree, f"{mesh}/fields/vorticity")
    assert vort["association"] == "element" and vort["values"].size == 23 * 11 * 7
    assert list(_get(tree, "catalyst/channels/derived_status/data/values/message")) == \
        list(b"ok: no steered fields")


class _SteerableSolver(_FakeSolver):
    """Records what steering asks of a Solver (run_lbm.py:34, lbm.py:613-620)."""

    def __init__(self):
        super().__init__(False)
        self.step_count = 0
        self.calls = []
        self.omega_rot = 0.0

    def set_physics(self, collision, reynolds):
        self.calls.append(("set_physics", collision, reynolds))
        self.collision_name, self.reynolds = collision, reynolds

    def set_actuation(self, omega_rot):
        self.calls.append(("set_actuation", omega_rot))
        self.omega_rot = omega_rot


def _results(parts):
    """A results tree with one steerable part holding `parts` (key -> array)."""
    fields = {k: {"values": np.atleast_1d(np.asarray(v, np.float64))} for k, v in parts.items()}
    return {"catalyst": {"steerable": {"data": {"proxy0": {"fields": fields}}}}}


def test_read_steering_takes_both_layouts_and_text():
    from tlbm.catalyst import read_steering
    tree = _results({"reynolds": 450.0, "collision": 1.0})
    tree["catalyst"]["steerable"]["data"]["proxy1"] = {
        "state": {"fields": {"rotation": np.array([0.75]), "derived": np.array([0])}}}
    tree["catalyst"]["derived"] = {"data": {"p": {"fields": {"derived": {
        "values": np.array([ord(c) for c in " w=mag(velocity) "] + [0], np.int32)}}}}}
    got = read_steering(tree)
    assert got == {"reynolds": 450.0, "collision": 1.0, "rotation": 0.75, "derived": ""}
    del tree["catalyst"]["steerable"]
    assert read_steering(tree) == {"derived": "w=mag(velocity)"}
    assert read_steering({}) == {}


def test_steer_applies_only_what_changed():
    """run_lbm.py:123-144: unchanged values do nothing, a rotation change only sets the
    actuation, a Reynolds or collision change calls set_physics with both."""
    from tlbm.catalyst import CatalystChannels, read_steering
    s = _SteerableSolver()
    ch = CatalystChannels(s)
    assert ch.steer(read_steering(_results({"reynolds": 300.0, "collision": 2.0}))) == {}
    assert s.calls == []
    assert ch.steer({"rotation": 0.5}) == {"rotation": 0.5}
    assert s.calls == [("set_actuation", 0.5 * s.u0)]
    assert ch.steer({"reynolds": 1000.0, "collision": 0.0, "rotation": 0.5}) == {
        "reynolds": 1000.0, "collision": 0.0}
    assert s.calls[-1] == ("set_physics", "bgk", 1000.0)
    assert ch.steer({"capture_stride": 500.0}) == {"capture_stride": 500.0}
    assert s.calls[-1] == ("set_physics", "bgk", 1000.0) and len(s.calls) == 2
    assert ch.steer({"power": 3.0}) == {}             # not a value this solver steers
    assert ch.steer({"collision": 7.0}) == {"collision": 7.0}
    assert s.calls[-1] == ("set_physics", "bgk", 1000.0)   # unknown code keeps the model


def test_rotation_schedule_fires_each_event_once_in_step_order():
    from tlbm.catalyst import RotationSchedule
    s = _SteerableSolver()
    sched = RotationSchedule("200:-1.5, 100:0.5,,300:2")
    assert sched.apply(s) == []
    s.step_count = 150
    assert sched.apply(s) == [(100, 0.5)] and s.omega_rot == 0.5 * s.u0
    s.step_count = 400
    assert sched.apply(s) == [(200, -1.5), (300, 2.0)] and s.omega_rot == 2.0 * s.u0
    assert sched.apply(s) == []
    with pytest.raises(ValueError):
        RotationSchedule("100")


class _Node:
    """The part of conduit.Node's API that read_steering uses, over a nested dict."""

    def __init__(self, tree):
        self.tree = tree

    def _get(self, path):
        node = self.tree
        for part in path.split("/"):
            node = node[part]
        return node

    def has_path(self, path):
        try:
            self._get(path)
            return True
        except (KeyError, TypeError):
            return False

    def __getitem__(self, path):
        v = self._get(path)
        return _Node(v) if isinstance(v, dict) else _Leaf(v)

    def number_of_children(self):
        return len(self.tree)

    def child(self, i):
        return _Node(list(self.tree.values())[i])


class _Leaf:
    def __init__(self, v):
        self.v = v

    def number_of_children(self):
        return 0

    def value(self):
        return self.v


def test_read_steering_takes_a_conduit_node():
    """catalyst.results fills a conduit.Node: the same values as from the dict tree."""
    from tlbm.catalyst import read_steering
    tree = _results({"reynolds": 450.0, "collision": 1.0, "rotation": -0.25})
    tree["catalyst"]["derived"] = {"data": {"p": {"state": {"fields": {
        "derived": np.array([ord(c) for c in "v=mag(velocity)"], np.int32)}}}}}
    assert read_steering(_Node(tree)) == read_steering(tree) == {
        "reynolds": 450.0, "collision": 1.0, "rotation": -0.25, "derived": "v=mag(velocity)"}
    assert read_steering(_Node({})) == {}
