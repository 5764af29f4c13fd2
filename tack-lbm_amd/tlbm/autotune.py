"""Launch-shape autotuner for the step kernels.

Times each candidate launch shape on the device with HIP events and keeps the fastest, caching
the choice in a JSON file keyed by device and workload (``$TLBM_AUTOTUNE_CACHE``, default
``~/.cache/tlbm/autotune.json``).  Tuning steps from the current state into the second
population buffer without swapping, so it does not advance or alter the simulation.
"""

from __future__ import annotations

import ctypes
import json
import os

from . import _capi
from ._capi import check, lib

CACHE = os.environ.get("TLBM_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "tlbm", "autotune.json"))


def _load():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(table):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass  # a read-only home: tuning still applies to this process


def device_key():
    buf = ctypes.create_string_buffer(256)
    check(lib().tlbm_device_name(buf, 256), "device name")
    cus = ctypes.c_int()
    check(lib().tlbm_device_cus(ctypes.byref(cus)), "device CUs")
    return f"{buf.value.decode(errors='replace')}/{cus.value}cu"


def _time(launch, stream, reps):
    L = lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    check(L.tlbm_event_create(ctypes.byref(a)), "event")
    check(L.tlbm_event_create(ctypes.byref(b)), "event")
    try:
        launch()
        launch()
        check(L.tlbm_event_record(a, stream), "record")
        for _ in range(reps):
            launch()
        check(L.tlbm_event_record(b, stream), "record")
        ms = ctypes.c_float()
        check(L.tlbm_event_elapsed_ms(a, b, ctypes.byref(ms)), "elapsed")
        return ms.value / reps
    finally:
        L.tlbm_event_destroy(a)
        L.tlbm_event_destroy(b)


def tune_batch(batch, reps=5, use_cache=True, candidates=None):
    """Pick (threads per block, persistent grid, chunk order) for a BatchSolver; applies them and
    returns (block, grid, ms); the chunk order chosen is ``batch.group``."""
    L = lib()
    cus = ctypes.c_int()
    check(L.tlbm_device_cus(ctypes.byref(cus)), "device CUs")
    key = (f"batch|{device_key()}|h{batch.handle}|{len(batch.domains)}dom|{batch.cells}cells|"
           f"{batch.elements}elems|" + ",".join(f"{d.nx}x{d.ny}x{d.nz}" for d in batch.domains[:4])
           + "|t" + ",".join(str(k) for k in getattr(batch, "tile_rows", [])[:4]))
    table = _load() if use_cache else {}
    if key in table:
        block, grid, ms = table[key][:3]
        group = table[key][3] if len(table[key]) > 3 else 1
        batch.set_launch(block, grid)
        batch.set_order(group)
        return block, grid, ms
    if candidates is None:
        candidates = []
        for block in (64, 128, 192, 256):
            waves = block // 64
            for per_simd in (0, 2, 4, 8):  # 0: one block per chunk (not persistent)
                grid = 0 if per_simd == 0 else cus.value * 4 * per_simd // waves
                for group in (1, 4, 8):
                    candidates.append((block, grid, group))
    def launch():
        check(L.tlbm_batch_step(batch._batch, batch.f[0].ptr, batch.f[1].ptr, 0, None, None,
                                batch.stream), "batch step")

    def timed(cand, n):
        batch.set_launch(cand[0], cand[1])
        batch.set_order(cand[2] if len(cand) > 2 else 1)
        return _time(launch, batch.stream, n)

    # a short pass over every candidate, then the six fastest again with four times the steps
    # (candidates within a few percent are told apart only by the longer timing)
    first = sorted((timed(c, reps), i) for i, c in enumerate(candidates))
    best = None
    for _, i in first[:6]:
        cand = candidates[i]
        ms = timed(cand, 4 * reps)
        if best is None or ms < best[2]:
            best = (cand[0], cand[1], ms, cand[2] if len(cand) > 2 else 1)
    batch.set_launch(best[0], best[1])
    batch.set_order(best[3])
    if use_cache:
        table[key] = list(best)
        _store(table)
    return best[:3]


def tune_batch_layout(make_batch, candidates=(0, "auto"), reps=5, use_cache=True):
    """Pick a batch's population layout as well as its launch shape: ``make_batch(tile_rows)``
    builds the BatchSolver for one layout candidate (``BatchSolver(..., tile_rows=...)``: 0 for
    padded q-major planes, "auto", a power of two, or one value per domain); each candidate's
    launch shape is tuned with ``tune_batch`` and the fastest batch is kept, the others closed.
    Returns (batch, candidate, (block, grid, ms))."""
    best = None
    for cand in candidates:
        batch = make_batch(cand)
        shape = tune_batch(batch, reps=reps, use_cache=use_cache)
        if best is None or shape[2] < best[2][2]:
            if best is not None:
                best[0].close()
            best = (batch, cand, shape)
        else:
            batch.close()
    return best  # tune_batch left the kept batch with its best launch shape and chunk order


def tune_solver(solver, reps=5, use_cache=True, blocks=(64, 128, 192, 256), groups=(1, 2, 4)):
    """Pick threads per block and the chunk order (``xcd_group``) of a Solver's quiet kernel
    (the report kernel takes the same block size); applies them with ``solver.set_launch`` and
    returns (block, ms, xcd_group)."""
    if getattr(solver, "heat_enabled", False):
        raise NotImplementedError("tune_solver times the flow step; this solver has heat")
    L = lib()
    lat = solver.lat
    key = (f"solver|{device_key()}|{solver.collision_name}|{type(solver.geo).__name__}|"
           f"{type(solver.flow.side).__name__}|{solver.storage}|{lat.nx}x{lat.ny}x{lat.nz}|"
           f"t{solver.tile_rows}")
    table = _load() if use_cache else {}
    if key in table and len(table[key]) == 3:
        block, ms, group = table[key]
        solver.set_launch(block, group)
        return block, ms, group
    from .lbm import make_params
    params = make_params(lat, solver.phys, solver.geo, solver.flow, solver.qstride,
                         solver.tile_rows)
    h = solver._handle(report=False)

    def timed(block, group, n):
        params.block_x, params.xcd_group = block, group

        def launch():
            check(L.tlbm_step(h, solver.f[0].ptr, solver.f[1].ptr, ctypes.byref(params), None,
                              None, solver.stream), "step")
        return _time(launch, solver.stream, n)

    cands = [(b, g) for b in blocks for g in groups]
    first = sorted((timed(b, g, reps), i) for i, (b, g) in enumerate(cands))
    best = None
    for _, i in first[:4]:  # the four fastest again, timed four times longer
        b, g = cands[i]
        ms = timed(b, g, 4 * reps)
        if best is None or ms < best[1]:
            best = (b, ms, g)
    solver.set_launch(best[0], best[2])
    if use_cache:
        table[key] = list(best)
        _store(table)
    return best
