"""Parameter sweeps: many independent lattices stepped together in one launch per step.

The reference steps one lattice per ``lbm_step`` launch (lbm.py:639-647); a sweep over Reynolds
number and obstacle size would run one ``Solver`` after another.  ``BatchSolver`` packs every
domain of a sweep into one population buffer and steps them all with one persistent kernel
launch (``tlbm_batch_step``), optionally replayed from a captured hipGraph.  Each domain is a
channel with its own size, Reynolds number and obstacle, set up as ``Solver`` sets up its one
lattice (lbm.py:551-584, 613-637), and reports its own FP64 monitors.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import check, lib
from .lbm import (COLLISIONS, FIELDS, LidY, PeriodicY, SlipY, Box, Cylinder, Sphere,
                  STORAGES, D3Q19, Flow, HostBuffer, decode_populations, make_params,
                  kernel_handle, sync, untile_populations, zeros)

_OBSTACLES = {"cylinder": (_capi.CYLINDER, Cylinder), "sphere": (_capi.SPHERE, Sphere),
              "box": (_capi.BOX, Box)}
_SIDES = {"periodic": PeriodicY, "slip": SlipY, "lid": LidY}


class Domain:
    """One lattice of a sweep: size, Reynolds number, obstacle diameter (and optional centre)."""

    def __init__(self, nx, ny, nz, reynolds, diameter, center=None, sponge_start=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.reynolds = float(reynolds)
        self.diameter = float(diameter)
        self.center = center
        self.sponge_start = sponge_start

    @property
    def n(self):
        return self.nx * self.ny * self.nz


class BatchSolver:
    """All domains of a sweep, one specialisation (collision x obstacle x side x storage).

    Domain d's populations sit at ``offsets(d)[0]`` in ``f``.  ``tile_rows`` = 0: q-major with
    plane stride ``plane_strides[d]``, its cell count rounded up to 64 plus ``plane_pad`` (0:
    packed); measured on MI355X over the 64-channel sweep, a 256-cell pad streams 17% faster
    (FP32) and 12% faster (FP16) than packed planes.  ``tile_rows`` = k > 0: tiles of k rows in
    every domain (19 planes of k nx cells per tile, as a Solver's ``tile_rows``); "auto" (the
    default): per domain ``batch_tile_rows`` (tiles at FP32, planes at FP16/BF16); a sequence:
    one value per domain.  ``tile_rows[d]`` holds each domain's choice.
    """

    def __init__(self, domains, collision="trt", obstacle="cylinder", lateral="periodic",
                 u0=0.08, storage="f32", cs=None, magic=None, seed_amp=0.05, noise_amp=0.0,
                 noise_seed=0, lid_u=(0.04, 0.0, 0.0), sponge_strength=1.0, stream=None,
                 fields=False, plane_pad=256, tile_rows="auto"):
        if storage not in STORAGES:
            raise ValueError(f"storage must be one of {sorted(STORAGES)}")
        self.domains = list(domains)
        if not self.domains:
            raise ValueError("a batch needs at least one domain")
        self.u0 = u0
        self.storage = storage
        self.collision = collision
        self.stream = stream
        side = LidY(*lid_u) if lateral == "lid" else _SIDES[lateral]()
        obs_code, obs_cls = _OBSTACLES[obstacle]
        self.handle = kernel_handle(COLLISIONS[collision].code, obs_code, _capi.QUIET, side.code,
                                    STORAGES[storage])
        self._omega = zeros((1,))
        params = (_capi.Params * len(self.domains))()
        self.phys, self.geo, self.plane_strides, self.tile_rows = [], [], [], []
        for i, d in enumerate(self.domains):
            nu = u0 * d.diameter / d.reynolds
            kw = {}
            if collision in ("trt", "les") and cs is not None:
                kw["cs"] = cs
            if collision == "trt" and magic is not None:
                kw["magic"] = magic
            phys = COLLISIONS[collision](nu, **kw)
            r = d.diameter / 2
            x0, y0, z0 = (0.22 * d.nx, 0.5 * d.ny + 1.7, 0.5 * d.nz) if d.center is None \
                else d.center
            geo = {"cylinder": lambda: Cylinder(x0, y0, r), "sphere": lambda: Sphere(x0, y0, z0, r),
                   "box": lambda: Box(x0, y0, z0, r, r, 0.3 * d.nz)}[obstacle]()
            flow = Flow(u0, 0.75 * d.nx if d.sponge_start is None else d.sponge_start,
                        sponge_strength, self._omega, side)
            if isinstance(tile_rows, str):
                if tile_rows != "auto":
                    raise ValueError(f"tile_rows {tile_rows!r} is not 0, a power of two or 'auto'")
                k = batch_tile_rows(d.nx, storage)
            elif np.ndim(tile_rows):
                k = int(tile_rows[i])
            else:
                k = int(tile_rows)
            if k < 0 or k & (k - 1):
                raise ValueError(f"tile_rows {k} is not 0, a power of two or 'auto'")
            # plane stride: packed, or padded past a 64-cell boundary by `plane_pad` cells
            # (q-major planes only; tiles have none)
            qs = d.n if plane_pad <= 0 or k else (d.n + 63) // 64 * 64 + int(plane_pad)
            params[i] = make_params(D3Q19(d.nx, d.ny, d.nz), phys, geo, flow, qs, k)
            self.plane_strides.append(qs)
            self.tile_rows.append(k)
            self.phys.append(phys)
            self.geo.append(geo)
        h = ctypes.c_void_p()
        check(lib().tlbm_batch_create(self.handle, params, len(self.domains), ctypes.byref(h)),
              "batch create")
        self._batch = h
        self.elements, self.cells = self._layout(len(self.domains))[2:]
        # zero-filled: the plane padding is never written and reads as zero
        self.f = [zeros((self.elements,), storage), zeros((self.elements,), storage)]
        self.sums = HostBuffer(_capi.TLBM_NSUMS * len(self.domains))
        self.fields = None
        self._fields_struct = None
        if fields:
            self.fields = {k: zeros((self.cells,)) for k in FIELDS}
            fs = _capi.Fields()
            for k in FIELDS:
                setattr(fs, k, self.fields[k].ptr)
            self._fields_struct = fs
        check(lib().tlbm_batch_init(self._batch, self.f[0].ptr, seed_amp, noise_amp, noise_seed,
                                    stream), "batch init")
        self.step_count = 0
        self.group = 1
        self._graphs = {}

    def _layout(self, d):
        e, c, te, tc = (ctypes.c_uint64() for _ in range(4))
        check(lib().tlbm_batch_layout(self._batch, d, ctypes.byref(e), ctypes.byref(c),
                                      ctypes.byref(te), ctypes.byref(tc)), "batch layout")
        return e.value, c.value, te.value, tc.value

    def offsets(self, d):
        """(population element offset, cell offset) of domain d."""
        e, c, _, _ = self._layout(d)
        return e, c

    def set_launch(self, block_x=0, grid_blocks=0):
        """Threads per block and persistent grid size (0: one block per chunk)."""
        for g in self._graphs.values():
            lib().tlbm_graph_destroy(g)
        self._graphs.clear()
        check(lib().tlbm_batch_set_launch(self._batch, int(block_x), int(grid_blocks)),
              "batch set launch")

    def set_order(self, group=1):
        """Chunk order: 1 plain, g > 1 runs of g consecutive chunks per XCD (tlbm_batch_set_order).
        Captured graphs hold the order they were captured with, so they are dropped."""
        for g in self._graphs.values():
            lib().tlbm_graph_destroy(g)
        self._graphs.clear()
        check(lib().tlbm_batch_set_order(self._batch, int(group)), "batch set order")
        self.group = int(group)

    def step(self, report=False):
        """One step of every domain (lbm.py:639-647 for all of them at once)."""
        fs = ctypes.byref(self._fields_struct) if (report and self._fields_struct) else None
        check(lib().tlbm_batch_step(self._batch, self.f[0].ptr, self.f[1].ptr, int(report), fs,
                                    self.sums.ptr if report else None, self.stream),
              "batch step")
        self.f.reverse()
        self.step_count += 1

    def run(self, steps, report_every=0):
        """`steps` (even) steps from a captured hipGraph (captured once per (steps,
        report_every)); needs a created stream."""
        if self.stream is None:
            raise ValueError("graph replay needs a created stream (BatchSolver(stream=...))")
        key = (int(steps), int(report_every))
        g = self._graphs.get(key)
        if g is None:
            out = ctypes.c_void_p()
            fs = ctypes.byref(self._fields_struct) if self._fields_struct else None
            check(lib().tlbm_batch_graph(self._batch, self.f[0].ptr, self.f[1].ptr, key[0],
                                         key[1], fs, self.sums.ptr, self.stream,
                                         ctypes.byref(out)), "batch graph")
            g = self._graphs[key] = out.value
        check(lib().tlbm_graph_launch(g, self.stream), "graph launch")
        self.step_count += key[0]

    def monitors(self, interval):
        """Per-domain monitors from the last report step (lbm.py:649-661 for each domain)."""
        sync(self.stream)
        s = self.sums.numpy.reshape(len(self.domains), _capi.TLBM_NSUMS)
        out = []
        for i, d in enumerate(self.domains):
            total = {k: float(s[i, j]) for j, k in enumerate(FIELDS)}
            res = {name: float(np.sqrt(max(total[key], 0.0) / d.n)) / max(interval, 1)
                   for name, key in (("continuity", "r_rho"), ("momentum_x", "r_ux"),
                                     ("momentum_y", "r_uy"), ("momentum_z", "r_uz"),
                                     ("energy", "r_ke"))}
            scale = 0.5 * self.u0 * self.u0 * d.diameter * d.nz
            res["drag"] = total["fx"] / scale
            res["lift"] = total["fy"] / scale
            res["mass"] = float(s[i, 13])
            out.append(res)
        return out

    def populations(self, d):
        """Domain d's populations as float32, shaped (19, n)."""
        sync(self.stream)
        e, _ = self.offsets(d)
        e1, _ = self.offsets(d + 1)
        dom, k = self.domains[d], self.tile_rows[d]
        flat = self.f[0].numpy()[e:e1]
        if k:
            raw = untile_populations(flat, D3Q19(dom.nx, dom.ny, dom.nz), k)
        else:
            qs = self.plane_strides[d]
            raw = flat[:19 * qs].reshape(19, qs)[:, :dom.n]
        return decode_populations(raw, self.storage)

    def close(self):
        for g in self._graphs.values():
            lib().tlbm_graph_destroy(g)
        self._graphs.clear()
        if getattr(self, "_batch", None):
            lib().tlbm_batch_destroy(self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_tile_rows(nx, storage="f32"):
    """Rows per tile of a batched domain with rows of nx cells (0: q-major planes).  FP32: the
    largest power of two k with k nx <= 4096 cells (16 KiB of a plane per tile).  Measured on
    MI355X over the 64-channel TRT sweep (nx in [64, 192]): 1.158-1.160 ms per step with 16 to
    32 KiB per plane (or 16 or 32 rows for every domain), 1.237 ms with 8 KiB, 1.205-1.212 ms
    with padded q-major planes.  FP16/BF16: 0, since tiles make the 16-bit loads of rows that
    are not whole waves form a 64-bit address per lane (0.744-0.768 ms against 0.705-0.721)."""
    if storage != "f32":
        return 0
    if nx > 4096:
        return 0
    k = 1
    while 2 * k * nx <= 4096:
        k *= 2
    return k


def sweep_domains(count=64, seed=0, nx_range=(64, 192), re_range=(100.0, 1000.0),
                  radius_range=(4.0, 12.0)):
    """A seeded sweep of channels: nx uniform in nx_range (multiple of 2), ny = nz = nx / 2,
    Reynolds number and cylinder radius uniform in their ranges."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(rng.integers(nx_range[0] // 2, nx_range[1] // 2 + 1)) * 2
        re = float(rng.uniform(*re_range))
        r = float(rng.uniform(*radius_range))
        out.append(Domain(nx, nx // 2, nx // 2, re, 2 * r))
    return out
