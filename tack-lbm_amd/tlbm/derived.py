"""Derived fields of the solver's output grid on the GPU (derived.py:48-93).

The reference evaluates VTK-calculator expressions (tack-expr) over the output grid; its two
default fields are the vorticity magnitude ``mag(curl(velocity))`` and the Q-criterion
``qcrit(velocity)``, both at cell centres (derived.py:26-29).  This module computes those two
with one hand-written kernel (``tlbm_vortex_fields``, reading the velocity once), with the
solver's gradient (central differences at the points, one-sided at the faces, a cell taking the
mean of its eight corners; ``gradient="lattice"``, the default) or the trilinear basis's
(``gradient="basis"``).  Other expressions need the expression compiler, which is not part of
this build: they raise ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes
import re

import numpy as np

from ._capi import check, lib
from .lbm import coarsen, empty, sync

DEFAULTS = {
    "vorticity": ("mag(curl(velocity))", "cells"),
    "q_criterion": ("qcrit(velocity)", "cells"),
}


_SPEC = re.compile(r"([^=]*)=(.*?)(?:@(cells|points))?", re.S)


def parse(specs):
    """Derived-field specifications NAME=EXPRESSION[@cells|@points] (derived.py:32-44) into
    {name: (expression, location)}: the name is what precedes the first '=', a trailing
    '@cells' or '@points' picks the location (points when absent), whitespace around the
    name and expression is dropped."""
    fields = {}
    for text in specs:
        m = _SPEC.fullmatch(text)
        name, expr = (m.group(1).strip(), m.group(2).strip()) if m else ("", "")
        if not name or not expr:
            raise ValueError(f"a derived field is NAME=EXPRESSION[@cells|@points], not {text!r}")
        fields[name] = (expr, m.group(3) or "points")
    return fields


def _kind(expr, at):
    e = "".join(expr.split()).lower()
    if at == "cells" and e == "mag(curl(velocity))":
        return "vorticity"
    if at == "cells" and e == "qcrit(velocity)":
        return "qcrit"
    return None


class DerivedFields:
    """derived.py:92-215 for the default expressions: ``evaluate()`` returns
    ``{name: (values, "element")}`` from the solver's last report step."""

    def __init__(self, solver, specs=None, gradient="lattice"):
        if gradient not in ("lattice", "basis"):
            raise ValueError(f"gradient is 'lattice' or 'basis', not {gradient!r}")
        specs = dict(DEFAULTS if specs is None else specs)
        self.kinds = {}
        for name, (expr, at) in specs.items():
            kind = _kind(expr, at)
            if kind is None:
                raise NotImplementedError(
                    f"derived field {name}={expr}@{at}: only mag(curl(velocity))@cells and "
                    "qcrit(velocity)@cells are built (no expression compiler)")
            self.kinds[name] = kind
        self.solver = solver
        self.gradient = gradient
        self.mx, self.my, self.mz = solver.out_dims
        if min(self.mx, self.my, self.mz) < 2:
            raise ValueError(f"output grid {solver.out_dims} has no cells")
        h = solver.stride / solver.diameter
        self.inv_h = 1.0 / h
        self.n = self.mx * self.my * self.mz
        self.cells = (self.mx - 1) * (self.my - 1) * (self.mz - 1)
        if solver.stride > 1:
            self._vel = [empty((self.n,)) for _ in range(3)]
        else:
            self._vel = None
        self._out = {k: empty((self.cells,)) for k in set(self.kinds.values())}
        # the steering status the Catalyst channel reports (derived.py:113-115); no steered
        # expressions without the expression compiler
        self.error_column = 0
        self.message = "ok: no steered fields"

    @property
    def names(self):
        return list(self.kinds)

    def _velocity(self):
        f = self.solver.fields
        if self._vel is None:
            return [f[k] for k in ("ux", "uy", "uz")]
        for k, dst in zip(("ux", "uy", "uz"), self._vel):
            coarsen(self.solver.lat, self.solver.stride, f[k], dst, self.mx, self.my, self.mz,
                    self.solver.stream)
        return self._vel

    def compute(self):
        """Launch the kernel on the solver's stream (device results in ``self._out``)."""
        L = lib()
        u = self._velocity()
        st = self.solver.stream
        vort = self._out.get("vorticity")
        qc = self._out.get("qcrit")
        check(L.tlbm_vortex_fields(u[0].ptr, u[1].ptr, u[2].ptr,
                                   vort.ptr if vort is not None else None,
                                   qc.ptr if qc is not None else None,
                                   self.mx, self.my, self.mz, self.inv_h,
                                   1 if self.gradient == "basis" else 0, st), "vortex fields")

    def evaluate(self):
        """{name: (values, 'element')} for the current report fields (derived.py:185-198)."""
        self.compute()
        sync(self.solver.stream)
        host = {k: a.numpy() for k, a in self._out.items()}
        return {name: (host[kind], "element") for name, kind in self.kinds.items()}

    def add_to(self, mesh):
        """Add the fields to a Conduit Blueprint mesh node or a nested dict (derived.py:200-212)."""
        for name, (values, association) in self.evaluate().items():
            fld = mesh[f"fields/{name}"] if hasattr(mesh, "fetch") else \
                mesh.setdefault("fields", {}).setdefault(name, {})
            fld["association"] = association
            fld["topology"] = "mesh"
            if hasattr(fld, "set"):
                fld["values"].set(np.ascontiguousarray(values))
            else:
                fld["values"] = np.ascontiguousarray(values)
