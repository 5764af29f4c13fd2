"""MLUPS and HBM-roofline fraction of the fused D3Q19 step on one MI355X.

Workload (the headline configuration): D3Q19 LES (Smagorinsky, Cs 0.16) on 512x256x256 =
33,554,432 cells, sphere of radius 16 centred at (96, 128, 128), moving top lid u = (0.04, 0, 0)
over a free-slip floor, inflow u0 = 0.08, Re 3000 (nu = u0 D / Re, D = 32), sponge over the last
quarter in x, rho = 1 with seeded U(-1e-3, 1e-3) velocity noise, and a report step (drag/lift,
residuals, FP64 sums) every 100 steps.  A "step" is one lattice update of the whole grid.

value = whole-grid MLUPS over the timed steps (report steps included) with FP32 populations;
the FP16 shifted-storage variant is timed the same way and reported under "variants" (BF16 with
--storage f32,f16,bf16: it misses the 1% Strouhal/drag bar against the FP64 oracle at Re 300).
The 16-bit variants warm up for at least VARIANT_WARMUP steps (their "warmup" field): when the
16-bit step starts after other work, the GPU's engine clock drops under its power management
(measured with GRBM_GUI_ACTIVE per dispatch: from ~2.3 GHz to 1.65-1.9 GHz for the first ~40
launches, 0.45-0.52 ms each against 0.41-0.42 ms once settled), and the 16-bit step, whose VALU
work is nearly as long as its memory time, slows with it; a few warm-up steps would time that
transient instead of the step.
Each variant also reports its pattern ceiling: the step's 19-load/19-store access pattern with
no arithmetic (tlbm_pull_copy) on the same buffers.
roofline: the quiet FP32 step kernel, 152 algorithmic bytes per cell (19 x 4 B read + 19 x 4 B
written), over its mean duration from HIP events on the launch stream.
cpu_baseline: the NumPy FP32 restatement of the same step (the oracle in float32), on a 512x256x4
slab of the same workload, single-threaded.

sweep: the batched parameter sweep (64 uneven TRT channels, one persistent launch per step from
a hipGraph), FP32 and FP16 storage, reported under "sweep".
adjoint: the gradient workload (BGK 128x64x64, 2000 steps checkpointed every 100), forward and
backward pass times and the adjoint kernel against the forward step kernel, under "adjoint".
derived: vorticity magnitude and Q-criterion of the 512x256x256 output grid, under "derived".
heat: the temperature variant's step against the flow step (heat-sink lattice), under "heat".
driver_loop: the reference's loop shape (99 quiet steps, a report step, monitors), wall clock.
graph (with --graph): a small lattice stepped one step() at a time and by Solver.run hipGraph
replays, under "graph".

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-sweep]
                       [--no-adjoint] [--no-derived] [--no-heat] [--no-loop] [--graph]
                       [--storage f32,f16[,bf16]] [--block N]
                       [--calibrate] [--sweep-steps S] [--sweep-launch BLOCK,GRID,GROUP]
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tack-lbm_amd"))
sys.path.insert(0, ROOT)

NX, NY, NZ = 512, 256, 256
RADIUS = 16.0
CENTER = (96.0, NY / 2, NZ / 2)
RE = 3000.0
U0 = 0.08
LID = (0.04, 0.0, 0.0)
REPORT_EVERY = 100
PEAK_GBS = 8000.0
BYTES_PER_CELL = {"f32": 152, "f16": 76, "bf16": 76}


def make_solver(storage, stream, nx=NX, ny=NY, nz=NZ):
    import tlbm
    return tlbm.Solver(nx, ny, nz, collision="les", obstacle="sphere", reynolds=RE, u0=U0,
                       diameter=2 * RADIUS, lateral="lid", lid_u=LID,
                       center=(CENTER[0], ny / 2, nz / 2), storage=storage, noise_amp=1e-3,
                       noise_seed=0, stream=stream)


class Events:
    def __init__(self, lib):
        self.lib = lib

    def new(self):
        e = ctypes.c_void_p()
        from tlbm._capi import check
        check(self.lib.tlbm_event_create(ctypes.byref(e)), "event create")
        return e

    def elapsed(self, a, b):
        ms = ctypes.c_float()
        from tlbm._capi import check
        check(self.lib.tlbm_event_elapsed_ms(a, b, ctypes.byref(ms)), "event elapsed")
        return ms.value


VARIANT_WARMUP = 100


def time_gpu(storage, steps, warmup, block_x=0):
    """Times `steps` steps (report every 100th) and the quiet-kernel mean from HIP events, after
    `warmup` untimed steps."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    s = make_solver(storage, stream.value)
    s.set_launch(block_x)
    for _ in range(warmup):
        s.step(report=(s.step_count % REPORT_EVERY == REPORT_EVERY - 1))
    ev = Events(L)
    tlbm.sync(stream.value)
    check(L.tlbm_device_sync(), "device sync")
    # quiet runs between report steps are bracketed by events on the launch stream
    marks = []
    t0 = time.perf_counter()
    start = ev.new()
    check(L.tlbm_event_record(start, stream), "record")
    run_start, run_len = None, 0
    for _ in range(steps):
        report = s.step_count % REPORT_EVERY == REPORT_EVERY - 1
        if report:
            if run_len:
                e = ev.new()
                check(L.tlbm_event_record(e, stream), "record")
                marks.append((run_start, e, run_len))
                run_start, run_len = None, 0
            s.step(report=True)
        else:
            if run_start is None:
                run_start = ev.new()
                check(L.tlbm_event_record(run_start, stream), "record")
            s.step()
            run_len += 1
    if run_len:
        e = ev.new()
        check(L.tlbm_event_record(e, stream), "record")
        marks.append((run_start, e, run_len))
    stop = ev.new()
    check(L.tlbm_event_record(stop, stream), "record")
    tlbm.sync(stream.value)
    check(L.tlbm_device_sync(), "device sync")
    wall = time.perf_counter() - t0
    total_ms = ev.elapsed(start, stop)
    quiet_ms = sum(ev.elapsed(a, b) for a, b, _ in marks)
    quiet_n = sum(k for _, _, k in marks)
    m = s.monitors(REPORT_EVERY)
    n = s.lat.n
    ceiling_ms = time_ceiling(s, stream)
    out = {
        "storage": storage,
        "warmup": warmup,
        "mlups": n * steps / (total_ms * 1e-3) / 1e6,
        "ms_per_step": total_ms / steps,
        "wall_ms_per_step": wall * 1e3 / steps,
        "quiet_kernel_ms": quiet_ms / max(quiet_n, 1),
        "quiet_launches": quiet_n,
        "monitors": {k: float(v) for k, v in m.items()},
    }
    out["achieved_gbs"] = BYTES_PER_CELL[storage] * n / (out["quiet_kernel_ms"] * 1e-3) / 1e9
    out["ceiling_ms"] = ceiling_ms
    out["ceiling_gbs"] = BYTES_PER_CELL[storage] * n / (ceiling_ms * 1e-3) / 1e9
    del s
    L.tlbm_stream_destroy(stream)
    return out


def time_loop(storage="f32", intervals=3):
    """The reference's driver loop shape (run_lbm.py:207-217, collision_check.py:14-17) on the
    headline workload: per interval 99 quiet steps, one report step and ``monitors(100)``, wall
    clock around `intervals` intervals; then the same with an in-situ reader taking the five
    output arrays (OutputArray.numpy: the report fields copied into managed memory)."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    s = make_solver(storage, stream.value)
    out = {}
    for reader in (False, True):
        for _ in range(REPORT_EVERY):                       # warm-up interval
            s.step(report=(s.step_count + 1) % REPORT_EVERY == 0)
        s.monitors(REPORT_EVERY)
        t0 = time.perf_counter()
        for _ in range(intervals):
            for _ in range(REPORT_EVERY - 1):
                s.step()
            s.step(report=True)
            s.monitors(REPORT_EVERY)
            if reader:
                for k in s.OUTPUT:
                    s.output[k].numpy
        ms = (time.perf_counter() - t0) * 1e3 / (intervals * REPORT_EVERY)
        out["with_output_read" if reader else "monitors_only"] = ms
    del s
    L.tlbm_stream_destroy(stream)
    return {"workload": "headline, 99 quiet steps + 1 report step + monitors(100) per interval",
            "storage": storage, "ms_per_step": round(out["monitors_only"], 4),
            "ms_per_step_with_output_read": round(out["with_output_read"], 4)}


def time_ceiling(s, stream, reps=50):
    """The step's access pattern without its arithmetic (tlbm_pull_copy: 19 pull loads and 19
    stores per cell, the same mapping, addressing and store policy) on the solver's own two
    population buffers, after its timed steps: mean ms per launch from HIP events.  The
    buffers' contents are overwritten."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    p = tlbm.make_params(s.lat, s.phys, s.geo, s.flow, s.qstride, s.tile_rows)
    code = tlbm.STORAGES[s.storage]
    for _ in range(5):
        check(L.tlbm_pull_copy(s.f[0].ptr, s.f[1].ptr, ctypes.byref(p), code, stream), "pull copy")
    ev = Events(L)
    e0, e1 = ev.new(), ev.new()
    check(L.tlbm_event_record(e0, stream), "record")
    for i in range(reps):
        a, b = (s.f[0], s.f[1]) if i % 2 == 0 else (s.f[1], s.f[0])
        check(L.tlbm_pull_copy(a.ptr, b.ptr, ctypes.byref(p), code, stream), "pull copy")
    check(L.tlbm_event_record(e1, stream), "record")
    return ev.elapsed(e0, e1) / reps


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles",
                               "lbm_step_les_sphere_lid_512x256x256_hbm_traffic.txt")


def profiled_traffic(storage):
    """HBM bytes per quiet launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (calibrated on a
    copy of known size) and committed under profiles/; None if not measured for `storage`."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == storage:
                    return float(parts[2])
    except OSError:
        pass
    return None


def calibrate_copy(gib=1, reps=5):
    """A plain copy of `gib` GiB (read + write) with the library's copy kernel: the known-size
    reference for HBM counters and the achievable streaming bandwidth on this device."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    nbytes = gib << 30
    a = tlbm.empty((nbytes // 4,))
    b = tlbm.empty((nbytes // 4,))
    check(L.tlbm_copy(b.ptr, a.ptr, nbytes, None), "copy")
    ev = Events(L)
    e0, e1 = ev.new(), ev.new()
    check(L.tlbm_event_record(e0, None), "record")
    for _ in range(reps):
        check(L.tlbm_copy(b.ptr, a.ptr, nbytes, None), "copy")
    check(L.tlbm_event_record(e1, None), "record")
    ms = ev.elapsed(e0, e1) / reps
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
            "gbs": 2 * nbytes / (ms * 1e-3) / 1e9}


def time_sweep(storage, steps=5000, graph_steps=100, seq_steps=20, launch=None):
    """The parameter sweep: 64 seeded TRT + Smagorinsky channels (nx in [64, 192], ny = nz =
    nx / 2, Re in [100, 1000], cylinder radius in [4, 12], u0 = 0.08, seeded 1e-3 noise) stepped
    `steps` times by one persistent launch per step, replayed from a hipGraph of `graph_steps`
    steps with a report step at its end; the BatchSolver's default population layout (tiles of
    rows at FP32, padded planes at FP16), launch shape from the autotuner.  For comparison, the
    same domains stepped one launch per domain (one Solver each) for `seq_steps` steps.
    `launch` = (block, persistent grid, chunk group) pins the launch shape instead."""
    import tlbm
    from tlbm import autotune
    from tlbm.batch import BatchSolver, sweep_domains
    from tlbm._capi import check, lib
    L = lib()
    domains = sweep_domains(64, seed=0)
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    b = BatchSolver(domains, collision="trt", storage=storage, noise_amp=1e-3,
                    stream=stream.value)
    if launch is None:
        block, grid, tune_ms = autotune.tune_batch(b)
    else:
        block, grid, group = launch
        b.set_launch(block, grid)
        b.set_order(group)
        tune_ms = 0.0
    group = b.group
    b.run(graph_steps, report_every=graph_steps)  # capture + warm-up
    ev = Events(L)
    e0, e1 = ev.new(), ev.new()
    tlbm.sync(stream.value)
    check(L.tlbm_event_record(e0, stream), "record")
    for _ in range(steps // graph_steps):
        b.run(graph_steps, report_every=graph_steps)
    check(L.tlbm_event_record(e1, stream), "record")
    ms = ev.elapsed(e0, e1)
    mons = b.monitors(graph_steps)
    cells = b.cells
    tiles = sorted(set(b.tile_rows))
    b.close()
    del b
    # one launch per domain per step
    seq = [tlbm.Solver(d.nx, d.ny, d.nz, collision="trt", obstacle="cylinder",
                       reynolds=d.reynolds, u0=0.08, diameter=d.diameter, storage=storage,
                       stream=stream.value) for d in domains]
    for s in seq:
        s.step()
    tlbm.sync(stream.value)
    check(L.tlbm_event_record(e0, stream), "record")
    for _ in range(seq_steps):
        for s in seq:
            s.step()
    check(L.tlbm_event_record(e1, stream), "record")
    seq_ms = ev.elapsed(e0, e1)
    del seq
    L.tlbm_stream_destroy(stream)
    bpc = BYTES_PER_CELL[storage]
    return {"domains": len(domains), "cells": cells, "steps": steps,
            "mlups": cells * steps / (ms * 1e-3) / 1e6, "ms_per_step": ms / steps,
            "achieved_gbs": bpc * cells * steps / (ms * 1e-3) / 1e9,
            "launch": {"block": block, "persistent_grid": grid, "xcd_chunk_group": group,
                       "tuned_ms": round(tune_ms, 4)},
            "tile_rows": tiles,
            "one_launch_per_domain_mlups": cells * seq_steps / (seq_ms * 1e-3) / 1e6,
            "drag_range": [min(m["drag"] for m in mons), max(m["drag"] for m in mons)]}


def time_adjoint(reps=100):
    """The adjoint workload: BGK 128x64x64 FP32, smoothed sphere (radius 10, centre seeded
    within +-2 cells of (32, 32, 32)) times a 32^3 porosity block, Re 200, 2000 steps with
    checkpoints every 100, d(mean drag over the last 1000 steps)/d(radius, centre, porosity).
    Reports the whole forward and backward passes (HIP events), and the adjoint kernel against
    the forward quiet step kernel on the same lattice (`reps` back-to-back launches each)."""
    import numpy as np
    import tlbm  # noqa: F401
    from tlbm.adjoint import DragAdjoint
    from tlbm.adjoint import _elapsed, _event
    rng = np.random.default_rng(0)
    center = np.array([32.0, 32.0, 32.0]) + rng.uniform(-2.0, 2.0, 3)
    porosity = 0.01 + 0.01 * rng.uniform(size=(32, 32, 32))
    prob = DragAdjoint(nx=128, ny=64, nz=64, reynolds=200.0, u0=0.08, radius=10.0,
                       nominal_center=(32.0, 32.0, 32.0), block=32, steps=2000, window=1000,
                       checkpoint_every=100, eps=1.5)
    prob.forward(10.0, center, porosity)
    prob.backward()  # warm-up (allocates the segment buffers)
    J = prob.forward(10.0, center, porosity)
    g = prob.backward()
    fwd, bwd = prob.forward_ms, prob.backward_ms
    run = prob.run
    e0 = _event(run.stream)
    for _ in range(reps):
        run._step(run.ckpt[0], run._seg[0])
    step_ms = _elapsed(e0, _event(run.stream)) / reps
    e0 = _event(run.stream)
    for _ in range(reps):
        run._adjoint(run.ckpt[0], run._seg[0], run._mu[0], run._mu[1], 1e-3, 1e-3)
    adj_ms = _elapsed(e0, _event(run.stream)) / reps
    prob.close()
    steps = 2000
    return {"workload": "D3Q19 BGK 128x64x64 FP32, smoothed sphere r10 + 32^3 porosity, Re 200, "
                        "2000 steps, checkpoints every 100, J = mean drag over the last 1000",
            "J": J, "dJ_dradius": g["radius"], "dJ_dcenter": [float(v) for v in g["center"]],
            "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
            "forward_ms_per_step": fwd / steps, "backward_ms_per_step": bwd / steps,
            "adjoint_kernel_ms": adj_ms, "forward_step_kernel_ms": step_ms,
            "adjoint_kernel_over_step_kernel": adj_ms / step_ms}


def time_derived(reps=20):
    """The default derived fields (vorticity magnitude and Q-criterion at the cell centres, the
    solver's gradient) of the headline workload's output grid at stride 1, against one quiet
    FP32 step: HIP events over `reps` evaluations (kernels only, no host copy)."""
    import tlbm
    from tlbm._capi import check, lib
    from tlbm.derived import DerivedFields
    L = lib()
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    s = make_solver("f32", stream.value)
    for _ in range(5):
        s.step()
    s.step(report=True)
    d = DerivedFields(s)
    d.compute()
    ev = Events(L)
    e0, e1, e2 = ev.new(), ev.new(), ev.new()
    tlbm.sync(stream.value)
    check(L.tlbm_event_record(e0, stream), "record")
    for _ in range(reps):
        d.compute()
    check(L.tlbm_event_record(e1, stream), "record")
    for _ in range(reps):
        s.step()
    check(L.tlbm_event_record(e2, stream), "record")
    tlbm.sync(stream.value)
    ms, step_ms = ev.elapsed(e0, e1) / reps, ev.elapsed(e1, e2) / reps
    m, cells = d.n, d.cells
    # one pass: 3 velocity reads per point, 2 writes per cell
    algorithmic = 4 * (3 * m + 2 * cells)
    del d, s
    L.tlbm_stream_destroy(stream)
    return {"grid": f"{NX}x{NY}x{NZ} points", "ms": ms, "step_ms": step_ms,
            "achieved_gbs": algorithmic / (ms * 1e-3) / 1e9, "algorithmic_bytes": algorithmic}


def time_graph(steps=1000, every=50):
    """Host-launch cost on a small lattice: BGK 96x48x48 with a cylinder (r 6, Re 100), 1000
    steps with a report step every 50, taken one step() at a time and as Solver.run(50, 50)
    hipGraph replays; HIP events around each loop on the solver's stream."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    s = tlbm.Solver(96, 48, 48, collision="bgk", obstacle="cylinder", reynolds=100.0,
                    diameter=12.0, center=(24.0, 24.0, 24.0), noise_amp=1e-3,
                    stream=stream.value)
    s.run(every, every)                                  # capture + warm-up
    ev = Events(L)
    e0, e1, e2 = ev.new(), ev.new(), ev.new()
    tlbm.sync(stream.value)
    check(L.tlbm_event_record(e0, stream), "record")
    for i in range(steps):
        s.step(report=(i + 1) % every == 0)
    check(L.tlbm_event_record(e1, stream), "record")
    for _ in range(steps // every):
        s.run(every, every)
    check(L.tlbm_event_record(e2, stream), "record")
    tlbm.sync(stream.value)
    eager, graph = ev.elapsed(e0, e1), ev.elapsed(e1, e2)
    cells = 96 * 48 * 48
    del s
    L.tlbm_stream_destroy(stream)
    return {"lattice": "BGK 96x48x48, cylinder r6, Re 100, report every 50", "steps": steps,
            "eager_ms_per_step": eager / steps, "graph_ms_per_step": graph / steps,
            "eager_mlups": cells * steps / (eager * 1e-3) / 1e6,
            "graph_mlups": cells * steps / (graph * 1e-3) / 1e6}


def time_heat(reps=20):
    """The temperature variant (heat=True, lbm.py:457-492) on the headline lattice with a heat
    sink in place of the sphere (LES, 512x256x256, the Solver's heat-sink placement), against the
    flow-only step of the same lattice: quiet-kernel means from HIP events."""
    import tlbm
    from tlbm._capi import check, lib
    L = lib()
    stream = ctypes.c_void_p()
    check(L.tlbm_stream_create(ctypes.byref(stream)), "stream create")
    out = {}
    for heat in (False, True):
        s = tlbm.Solver(NX, NY, NZ, collision="les", obstacle="heatsink", reynolds=RE, u0=U0,
                        diameter=2 * RADIUS, heat=heat, stream=stream.value)
        if heat:
            s.set_wall_temperature(1.0)
        for _ in range(3):
            s.step()
        ev = Events(L)
        e0, e1 = ev.new(), ev.new()
        tlbm.sync(stream.value)
        check(L.tlbm_event_record(e0, stream), "record")
        for _ in range(reps):
            s.step()
        check(L.tlbm_event_record(e1, stream), "record")
        tlbm.sync(stream.value)
        out["heat" if heat else "flow"] = ev.elapsed(e0, e1) / reps
        del s
    L.tlbm_stream_destroy(stream)
    n = NX * NY * NZ
    return {"workload": "D3Q19 LES + D3Q7 temperature 512x256x256, heat sink, Re 3000",
            "flow_step_ms": out["flow"], "heat_step_ms": out["heat"],
            "heat_over_flow": out["heat"] / out["flow"],
            "achieved_gbs": (152 + 56) * n / (out["heat"] * 1e-3) / 1e9,
            "algorithmic_bytes_per_cell": 152 + 56}


def _cpu_slab(seconds, nz=4):
    """One host process: the NumPy float32 oracle stepping a 512x256x`nz` slab of the workload
    for about `seconds` (at most 100 steps); returns (cell updates, seconds)."""
    import numpy as np
    from oracle import lbm_oracle as O
    s = O.OracleSolver(NX, NY, nz, "les", "sphere", RE, U0, 2 * RADIUS, lateral="lid",
                       dtype=np.float32, center=(CENTER[0], NY / 2, nz / 2), lid_u=LID,
                       noise_amp=1e-3, construct_report=False)
    s.step()  # warm-up (index caches)
    steps, t0 = 0, time.perf_counter()
    while True:
        s.step()
        steps += 1
        if time.perf_counter() - t0 > seconds or steps >= 100:
            break
    return s.n * steps, time.perf_counter() - t0


def cpu_baseline(max_seconds=15.0):
    """The NumPy float32 restatement on 512x256x4 slabs of the workload, one process per usable
    host core (NumPy's elementwise kernels are single-threaded), run together; the value is
    their summed cell updates over the longest process's time.  Worker processes are started
    fresh (spawn): they run NumPy only and never touch the GPU."""
    import multiprocessing as mp
    usable = len(os.sched_getaffinity(0))
    workers = max(1, min(usable, 16))
    with mp.get_context("spawn").Pool(workers) as pool:
        runs = pool.starmap(_cpu_slab, [(max_seconds,)] * workers)
    cells = sum(c for c, _ in runs)
    dt = max(t for _, t in runs)
    single = runs[0][0] / runs[0][1] / 1e6
    return {"value": cells / dt / 1e6, "unit": "MLUPS", "cores": workers, "kind": "port",
            "sample": f"NumPy float32 oracle step on a {NX}x{NY}x4 slab of the workload (LES, "
                      f"sphere, lid) in each of {workers} processes at once, {dt:.1f} s; "
                      f"{single:.2f} MLUPS per process; host has {os.cpu_count()} CPUs, "
                      f"{usable} usable"}


def _launch_triple(text):
    parts = [int(v) for v in text.split(",")]
    if len(parts) != 3:
        raise argparse.ArgumentTypeError("expected BLOCK,GRID,GROUP")
    return tuple(parts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f32,f16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0 = default)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched parameter sweep")
    ap.add_argument("--sweep-steps", type=int, default=5000,
                    help="timed steps of the sweep (a multiple of 100)")
    ap.add_argument("--sweep-launch", type=_launch_triple, default=None,
                    help="BLOCK,GRID,GROUP: the sweep's launch shape and chunk order, not tuned")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--no-derived", action="store_true", help="skip the derived-field timing")
    ap.add_argument("--no-heat", action="store_true", help="skip the temperature-variant timing")
    ap.add_argument("--no-loop", action="store_true",
                    help="skip the driver-loop timing (report step + monitors every 100)")
    ap.add_argument("--graph", action="store_true",
                    help="add the small-lattice step-by-step vs hipGraph timing")
    ap.add_argument("--calibrate", action="store_true",
                    help="also time a plain 1 GiB copy (counter calibration, streaming ceiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is "
              "described, not built", file=sys.stderr)
        sys.exit(2)
    if args.sweep_steps <= 0 or args.sweep_steps % 100:
        ap.error("--sweep-steps must be a positive multiple of 100")
    storages = [s for s in args.storage.split(",") if s]
    results = {}
    head_storage = "f32" if "f32" in storages else storages[0]
    for st in storages:
        warm = args.warmup if st == head_storage else max(args.warmup, VARIANT_WARMUP)
        results[st] = time_gpu(st, args.steps, warm, args.block)
        r = results[st]
        print(f"[bench] {st}: {r['mlups']:.0f} MLUPS, quiet kernel {r['quiet_kernel_ms']:.3f} ms "
              f"= {r['achieved_gbs']:.0f} GB/s; the same access pattern without arithmetic "
              f"{r['ceiling_ms']:.3f} ms = {r['ceiling_gbs']:.0f} GB/s", file=sys.stderr,
              flush=True)
    sweep = {}
    if not args.no_sweep:
        for st in ("f32", "f16"):
            sweep[st] = time_sweep(st, steps=args.sweep_steps, launch=args.sweep_launch)
            print(f"[bench] sweep {st}: {sweep[st]['mlups']:.0f} MLUPS over {sweep[st]['cells']} "
                  f"cells in {sweep[st]['domains']} domains (one launch per domain: "
                  f"{sweep[st]['one_launch_per_domain_mlups']:.0f})", file=sys.stderr, flush=True)
    adjoint = None
    if not args.no_adjoint:
        adjoint = time_adjoint()
        print(f"[bench] adjoint: forward {adjoint['forward_ms']:.1f} ms, backward "
              f"{adjoint['backward_ms']:.1f} ms ({adjoint['backward_over_forward']:.2f}x); "
              f"adjoint kernel {adjoint['adjoint_kernel_ms'] * 1e3:.1f} us vs step kernel "
              f"{adjoint['forward_step_kernel_ms'] * 1e3:.1f} us", file=sys.stderr, flush=True)
    derived = None
    if not args.no_derived:
        derived = time_derived()
        print(f"[bench] derived fields: {derived['ms']:.3f} ms ({derived['achieved_gbs']:.0f} "
              f"GB/s) vs one step {derived['step_ms']:.3f} ms", file=sys.stderr, flush=True)
    heat = None
    if not args.no_heat:
        heat = time_heat()
        print(f"[bench] heat: step {heat['heat_step_ms']:.3f} ms vs flow {heat['flow_step_ms']:.3f} "
              f"ms ({heat['heat_over_flow']:.2f}x, {heat['achieved_gbs']:.0f} GB/s)",
              file=sys.stderr, flush=True)
    loop = None
    if not args.no_loop:
        loop = time_loop()
        print(f"[bench] driver loop (99 quiet + report + monitors): {loop['ms_per_step']:.4f} ms "
              f"per step; with the output grid read each interval "
              f"{loop['ms_per_step_with_output_read']:.4f} ms", file=sys.stderr, flush=True)
    graph = None
    if args.graph:
        graph = time_graph()
        print(f"[bench] small lattice (96x48x48): {graph['eager_mlups']:.0f} MLUPS step by step, "
              f"{graph['graph_mlups']:.0f} MLUPS from hipGraph replays", file=sys.stderr,
              flush=True)
    copy = calibrate_copy() if args.calibrate else None
    if copy:
        print(f"[bench] copy kernel: {copy['gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    head = results.get("f32") or results[storages[0]]
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "MLUPS & %HBM roofline, D3Q19 LES 512x256x256, FP32 vs FP16-store",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["storage"] == "f32" else head["storage"] + "-store/f32-compute",
        "data": "synthetic",
        "config": {"workload": "D3Q19 LES 512x256x256, sphere r16 at x=96, lid u=0.04, Re 3000, "
                               "seeded 1e-3 velocity noise, report every 100 steps",
                   "cells": NX * NY * NZ, "storage": head["storage"]},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1),
                     "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["achieved_gbs"] / PEAK_GBS, 4),
                     "traffic": profiled_traffic(head["storage"]),
                     "algorithmic_bytes": BYTES_PER_CELL[head["storage"]] * NX * NY * NZ},
        "cpu_baseline": cpu,
        "variants": {st: {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
                          "warmup": r["warmup"],
                          "quiet_kernel_ms": round(r["quiet_kernel_ms"], 4),
                          "achieved_gbs": round(r["achieved_gbs"], 1),
                          "roofline_frac": round(r["achieved_gbs"] / PEAK_GBS, 4),
                          "traffic": profiled_traffic(st),
                          "pattern_ceiling": {"ms": round(r["ceiling_ms"], 4),
                                              "gbs": round(r["ceiling_gbs"], 1),
                                              "step_frac_of_ceiling": round(
                                                  r["ceiling_ms"] / r["quiet_kernel_ms"], 4)},
                          "drag": r["monitors"]["drag"], "lift": r["monitors"]["lift"]}
                     for st, r in results.items()},
    }
    if sweep:
        line["sweep"] = {st: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                         for st, r in sweep.items()}
    if adjoint:
        line["adjoint"] = {k: (round(v, 6) if isinstance(v, float) else v)
                           for k, v in adjoint.items()}
    if heat:
        line["heat"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in heat.items()}
    if loop:
        line["driver_loop"] = loop
    if graph:
        line["graph"] = {k: (round(v, 4) if isinstance(v, float) else v)
                         for k, v in graph.items()}
    if derived:
        line["derived"] = {k: (round(v, 4) if isinstance(v, float) else v)
                           for k, v in derived.items()}
    if copy:
        line["copy_calibration"] = {k: round(v, 4) if isinstance(v, float) else v
                                    for k, v in copy.items()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
