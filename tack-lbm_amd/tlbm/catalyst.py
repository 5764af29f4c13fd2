"""The Catalyst channels of a running Solver (run_lbm.py:38-121), as Conduit node paths.

``CatalystChannels.execute(step, time, monitors)`` builds the tree the reference's
``CatalystAdaptor.execute`` hands to ``catalyst.execute``: the state (timestep, time,
parameters), the ``grid`` channel (a Blueprint uniform mesh of the solver's output grid whose
vertex fields are zero-copy views of ``Solver.output``: velocity x/y/z, pressure, speed and
temperature with heat), the optional derived fields, the ``monitors`` table (the history of the
monitor values) and the ``derived_status`` text table.  The tree is a nested dict keyed by path
components; ``to_conduit`` copies it into a ``conduit.Node`` (arrays with ``set_external``) when
Conduit is importable.  Catalyst itself (initialize/execute/results) and ParaView are not part
of this build.
"""

from __future__ import annotations

import numpy as np

COLLISION_CODES = {"bgk": 0, "trt": 1, "les": 2}  # steering.py:7
COLLISION_NAMES = {code: name for name, code in COLLISION_CODES.items()}
# the scalar properties a steering client may send (steering.py:141)
STEERABLE = ("reynolds", "collision", "power", "gust", "stiffness", "rotation", "capture_stride")
# run_lbm.py:25-26
MONITOR_NAMES = ("continuity", "momentum_x", "momentum_y", "momentum_z", "energy", "drag", "lift",
                 "reynolds", "collision", "rotation", "derived_error", "time")


def steering_monitors(solver, monitors, step, derived=None):
    """run_lbm.py:218-224: the steering values and the convective time added to the monitors."""
    monitors["reynolds"] = solver.reynolds
    monitors["collision"] = float(COLLISION_CODES[solver.collision_name])
    monitors["rotation"] = rotation_rate(solver)
    monitors["derived_error"] = float(getattr(derived, "error_column", 0))
    monitors["time"] = step * solver.u0 / solver.diameter
    return monitors


def _put(tree, path, value):
    node = tree
    parts = path.split("/")
    for p in parts[:-1]:
        node = node.setdefault(p, {})
    node[parts[-1]] = value
    return value


def _get(tree, path):
    node = tree
    for p in path.split("/"):
        node = node[p]
    return node


def rotation_rate(solver):
    """run_lbm.py:29-31: cylinder surface speed over inflow speed."""
    return getattr(solver, "omega_rot", 0.0) / solver.u0


def set_rotation_rate(solver, alpha):
    """run_lbm.py:33-34: set the cylinder surface speed to alpha times the inflow speed."""
    solver.set_actuation(float(alpha) * solver.u0)


class RotationSchedule:
    """run_lbm.py:206-213: "step:alpha,step:alpha,..." events, applied once the solver has
    taken at least that many steps (each event once, in step order)."""

    def __init__(self, spec=""):
        events = []
        for item in spec.split(","):
            item = item.strip()
            if not item:
                continue
            step, sep, alpha = item.partition(":")
            if not sep:
                raise ValueError(f"a rotation event is STEP:ALPHA, not {item!r}")
            events.append((int(step), float(alpha)))
        self.events = sorted(events)

    def apply(self, solver):
        """Apply every due event; returns the (step, alpha) pairs applied."""
        done = []
        while self.events and solver.step_count >= self.events[0][0]:
            event = self.events.pop(0)
            set_rotation_rate(solver, event[1])
            done.append(event)
        return done


def _text(codes):
    """steering.py:116-121: text sent as an array of UTF-8 byte values (0 alone: empty)."""
    values = np.atleast_1d(np.asarray(codes)).reshape(-1)
    return bytes(int(c) for c in values if 0 < int(c) < 256).decode("utf-8", "replace").strip()


def _first(value):
    return float(np.atleast_1d(np.asarray(value, dtype=np.float64)).reshape(-1)[0])


def _is_node(obj):
    """A conduit.Node (what catalyst.results fills), told apart by its path API."""
    return hasattr(obj, "has_path") and hasattr(obj, "number_of_children")


def _lookup(tree, path):
    """The value at `path` of a nested dict or a conduit.Node; None where it is missing."""
    if _is_node(tree):
        if not tree.has_path(path):
            return None
        node = tree[path]
        return node if node.number_of_children() else node.value()
    node = tree
    for part in path.split("/"):
        if not isinstance(node, dict) or part not in node:
            return None
        node = node[part]
    return node


def _parts(data):
    """The children of a mesh's data node (a dict's values or a conduit.Node's children)."""
    if _is_node(data):
        return [data.child(i) for i in range(data.number_of_children())]
    return list(data.values()) if isinstance(data, dict) else []


def read_steering(results):
    """steering.py:124-147 over a results tree: a conduit.Node (what ``catalyst.results``
    fills) or nested dicts keyed by path component (as ``CatalystChannels.execute`` builds
    them).  Returns the derived-field request text and the steered scalars of every part under
    ``catalyst/derived/data`` and ``catalyst/steerable/data``, each from
    ``state/fields/<key>`` or ``fields/<key>/values``; missing values are left out."""
    values = {}
    for mesh in ("catalyst/derived/data", "catalyst/steerable/data"):
        for part in _parts(_lookup(results, mesh)):
            for path in ("state/fields/derived", "fields/derived/values"):
                v = _lookup(part, path)
                if v is not None:
                    values["derived"] = _text(v)
    for part in _parts(_lookup(results, "catalyst/steerable/data")):
        for key in STEERABLE:
            for path in (f"state/fields/{key}", f"fields/{key}/values"):
                v = _lookup(part, path)
                if v is not None:
                    values[key] = _first(v)
    return values


class CatalystChannels:
    """run_lbm.py:38-121 without the Catalyst calls.  ``scales``: optional SI output,
    {"length": metres per cell, quantity: (factor, offset)} (run_lbm.py:51-52)."""

    def __init__(self, solver, monitor_names=MONITOR_NAMES, derived=None, scales=None,
                 monitor_history=0):
        self.solver = solver
        self.monitor_names = tuple(monitor_names)
        self.derived = derived
        self.scales = scales
        self.monitor_history = monitor_history
        self.history = {name: [] for name in ("iteration", *self.monitor_names)}
        # the values a steering client last set (run_lbm.py:55-56)
        self.steered = {"reynolds": solver.reynolds,
                        "collision": float(COLLISION_CODES[solver.collision_name]),
                        "rotation": rotation_rate(solver), "capture_stride": 0.0}

    def steer(self, values):
        """run_lbm.py:123-144: apply steered values (``read_steering``'s dict).  A changed
        rotation rate sets the cylinder's surface speed; a changed Reynolds number or collision
        model calls ``set_physics`` (a collision change selects another kernel).  Returns the
        values that changed."""
        values = dict(values)
        text = values.pop("derived", None)
        if text is not None and self.derived is not None and hasattr(self.derived, "steer"):
            self.derived.steer(text)
        s = self.solver
        changed = {k: float(v) for k, v in values.items()
                   if abs(float(v) - self.steered.get(k, float(v))) > 1e-9}
        if not changed:
            return changed
        self.steered.update(changed)
        if "rotation" in changed:
            set_rotation_rate(s, self.steered["rotation"])
        if changed.keys() - {"rotation", "capture_stride"}:
            collision = COLLISION_NAMES.get(int(self.steered["collision"]), s.collision_name)
            s.set_physics(collision, self.steered["reynolds"])
        return changed

    def _values(self, key, quantity):
        arr = self.solver.output[key].numpy
        if self.scales and quantity in self.scales:
            factor, offset = self.scales[quantity]
            return arr * factor + offset
        return arr                                          # zero-copy (set_external)

    def execute(self, step, time_value, monitors):
        """The node tree of one Catalyst call (run_lbm.py:58-121)."""
        self.history["iteration"].append(float(step))
        for name in self.monitor_names:
            self.history[name].append(monitors.get(name, np.nan))
        s = self.solver
        tree = {}
        _put(tree, "catalyst/state/timestep", step)
        _put(tree, "catalyst/state/time", time_value)
        _put(tree, "catalyst/state/parameters",
             [f"collision={s.collision_name}", f"reynolds={s.reynolds}",
              f"rotation={rotation_rate(s)}"])
        grid = "catalyst/channels/grid"
        _put(tree, f"{grid}/type", "mesh")
        mesh = f"{grid}/data"
        mx, my, mz = s.out_dims
        for axis, dim in zip("ijk", (mx, my, mz)):
            _put(tree, f"{mesh}/coordsets/coords/dims/{axis}", dim)
        _put(tree, f"{mesh}/coordsets/coords/type", "uniform")
        spacing = s.stride * self.scales["length"] if self.scales else s.stride / s.diameter
        for axis in "xyz":
            _put(tree, f"{mesh}/coordsets/coords/origin/{axis}", 0.0)
            _put(tree, f"{mesh}/coordsets/coords/spacing/d{axis}", spacing)
        _put(tree, f"{mesh}/topologies/mesh/type", "uniform")
        _put(tree, f"{mesh}/topologies/mesh/coordset", "coords")
        _put(tree, f"{mesh}/fields/velocity/association", "vertex")
        _put(tree, f"{mesh}/fields/velocity/topology", "mesh")
        for axis in "xyz":
            _put(tree, f"{mesh}/fields/velocity/values/{axis}",
                 self._values(f"u{axis}", "velocity"))
        for name, key, quantity in (("pressure", "p", "pressure"), ("speed", "vmag", "velocity"),
                                    ("temperature", "temp", "temperature")):
            if key not in s.output:
                continue
            _put(tree, f"{mesh}/fields/{name}/association", "vertex")
            _put(tree, f"{mesh}/fields/{name}/topology", "mesh")
            _put(tree, f"{mesh}/fields/{name}/values", self._values(key, quantity))
        if self.derived:
            self.derived.add_to(_get(tree, mesh))
        _put(tree, "catalyst/channels/monitors/type", "table")
        for name, values in self.history.items():
            tail = values[-self.monitor_history:] if self.monitor_history else values
            _put(tree, f"catalyst/channels/monitors/data/values/{name}", np.asarray(tail))
        if self.derived:
            codes = list(getattr(self.derived, "message", "ok").encode("utf-8")) or [0]
            _put(tree, "catalyst/channels/derived_status/type", "table")
            _put(tree, "catalyst/channels/derived_status/data/values/message",
                 np.asarray(codes, dtype=np.int32))
        return tree


def paths(tree, prefix=""):
    """Every leaf path of a tree, sorted."""
    out = []
    for k, v in tree.items():
        p = f"{prefix}/{k}" if prefix else k
        out.extend(paths(v, p) if isinstance(v, dict) else [p])
    return sorted(out)


def to_conduit(tree, node=None):
    """Copy a tree into a conduit.Node (arrays by reference); needs the conduit module."""
    import conduit  # noqa: F401  (absent in this image)
    node = conduit.Node() if node is None else node
    for p in paths(tree):
        v = _get(tree, p)
        if isinstance(v, np.ndarray):
            node[p].set_external(v)
        elif isinstance(v, list):
            for item in v:
                node[p].append().set(item)
        else:
            node[p] = v
    return node
