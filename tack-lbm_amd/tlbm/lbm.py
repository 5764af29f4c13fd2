"""Host mirror of the reference's solver API, driving the HIP step library through ctypes.

Same names, argument meanings and layout as the reference module (lbm.py): lattice, collision,
obstacle, diagnostics, side and flow "plugins" are plain Python objects whose *types* pick one
pre-compiled kernel specialisation (the reference's Tack compiles one per type combination,
lbm.py:1-7, README.md:55-56) and whose instance attributes become the kernel's runtime
arguments.  ``lbm_step`` / ``lbm_init`` / ``coarsen`` keep the reference's kernel signatures;
``Solver`` keeps its constructor, ``step``, ``monitors`` and field layout (lbm.py:548-661).

Differences, by design:
  * the step runs on libtlbm.so (hand-written HIP for gfx950); there is no CPU path;
  * ``Solver.monitors`` reads FP64 sums reduced inside the report kernel instead of summing the
    13 managed fields on the host (lbm.py:654);
  * the 13 report fields live in HBM by default (``field_memory="device"``): managed memory
    on an MI355X without XNACK is host-resident, and writing the fields there each report step
    would cost ~30x the step itself.  Host code reads them as it reads managed arrays
    (``DeviceField.numpy`` is a synchronised host copy).  ``Solver.output`` stays a set of
    managed arrays that in-situ readers use zero-copy (``OutputArray``): reading one after a
    report step refreshes it from the device field.  ``field_memory="managed"`` keeps the
    reference's layout (the kernel writes managed memory);
  * population storage can be FP16 or BF16 deviations f_q - w_q (``storage="f16"/"bf16"``);
  * a Solver of 2^20 cells or more pads the population planes (``default_qstride``): plane q of
    ``Solver.f`` starts at ``q * qstride``, not ``q * n`` (the kernels take the stride); and,
    when its rows are a multiple of 64 cells, keeps its populations in tiles of a few rows
    (``default_tile_rows``; ``lbm_step``/``lbm_init`` take ``tile_rows``): each tile holds its
    19 planes together (and, with ``heat=True``, each temperature tile its 7), which streams
    faster than planes a lattice apart.  ``populations()`` and ``temperature_populations()``
    return the q-major planes in either layout;
  * ``LidY`` adds a moving lid; obstacle centre, sponge start, Cs/magic and a seeded velocity
    noise can be overridden;
  * with ``heat=True`` the temperature field follows the report fields: ``heat_fields["temp"]``
    is device memory (its managed copy is ``output["temp"]``), ``q_solid`` is device memory
    (the kernel adds to it atomically; read it with ``q_solid()``), ``t_solid`` stays a managed
    array written from the host as in the reference; the temperature variant needs FP32
    populations.
"""

from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np

from . import _capi
from ._capi import check, lib

# ---- lattice tables (lbm.py:15-35) ---------------------------------------------------------------
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]
W = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12


def _opposite():
    return [next(p for p in range(19) if (CX[p], CY[p], CZ[p]) == (-CX[q], -CY[q], -CZ[q]))
            for q in range(19)]


def _mirror_y():
    return [next(p for p in range(19) if (CX[p], CY[p], CZ[p]) == (CX[q], -CY[q], CZ[q]))
            for q in range(19)]


W32 = np.array(W, np.float32)

STORAGES = {"f32": _capi.F32, "f16": _capi.F16, "bf16": _capi.BF16}
_RAW_DTYPE = {"f32": np.float32, "f16": np.float16, "bf16": np.uint16}


# ---- device memory (tack.array / tack.empty / tack.zeros, lbm.py:45-50, 566) ---------------------
class DeviceArray:
    """A flat device allocation of ``size`` elements in one of the storage formats.

    ``f32`` holds float32 values; ``f16``/``bf16`` population buffers hold deviations
    f_q - w_q in binary16 / bfloat16 (raw uint16 on the host side for bf16).
    """

    def __init__(self, size, storage="f32", ptr=None):
        if storage not in STORAGES:
            raise ValueError(f"storage must be one of {sorted(STORAGES)}, got {storage!r}")
        self.size = int(size)
        self.storage = storage
        self.dtype = np.dtype(_RAW_DTYPE[storage])
        self.nbytes = self.size * self.dtype.itemsize
        # ptr given: a view of memory someone else owns (ManagedArray.tack); never freed here
        self._owned = ptr is None
        if ptr is None:
            p = ctypes.c_void_p()
            check(lib().tlbm_alloc_device(self.nbytes, ctypes.byref(p)), "device allocation")
            ptr = p.value
        self.ptr = ptr

    def numpy(self):
        """A host copy of the raw contents."""
        out = np.empty(self.size, self.dtype)
        check(lib().tlbm_memcpy(out.ctypes.data, self.ptr, self.nbytes), "device-to-host copy")
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        if a.size != self.size:
            raise ValueError(f"size {a.size} != {self.size}")
        check(lib().tlbm_memcpy(self.ptr, a.ctypes.data, self.nbytes), "host-to-device copy")
        return self

    def zero_(self, stream=None):
        check(lib().tlbm_memset(self.ptr, 0, self.nbytes, stream), "memset")
        return self

    def free(self):
        if getattr(self, "ptr", None):
            if self._owned:
                lib().tlbm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceField(DeviceArray):
    """An n-float field in HBM that host code reads as it reads a ``ManagedArray``: ``numpy``
    is a property that waits for ``stream`` (None: the device) and returns a host copy, so the
    reference's ``solver.fields["ux"].numpy.min()`` (bench.py:29) and
    ``heat_fields["q_solid"].numpy[voxels]`` (cht.py:85) read the field after the last step.
    The copy is read-only: a write to it would not reach the device (write device fields with
    ``copy_from``)."""

    def __init__(self, size, stream=None):
        super().__init__(size, "f32")
        self.stream = stream
        self.zero_(stream)

    @property
    def numpy(self):
        sync(self.stream)
        out = DeviceArray.numpy(self)
        out.flags.writeable = False
        return out

    @property
    def tack(self):
        """The device array itself (the kernels take it as they take a ``ManagedArray.tack``)."""
        return self


def array(a, storage="f32"):
    """tack.array: upload a NumPy array (raw storage values)."""
    a = np.asarray(a)
    return DeviceArray(a.size, storage).copy_from(a)


def empty(shape, storage="f32"):
    size = int(np.prod(shape))
    return DeviceArray(size, storage)


def zeros(shape, storage="f32"):
    return empty(shape, storage).zero_()


def encode_populations(f, storage, q_axis=0):
    """Float populations (19 along ``q_axis``) -> raw storage values (f32, or f - w_q for
    f16/bf16 with round-to-nearest-even), the format the step kernels read and write."""
    f = np.asarray(f, np.float32)
    if storage == "f32":
        return f.copy()
    shape = [1] * f.ndim
    shape[q_axis] = 19
    dev = (f - W32.reshape(shape)).astype(np.float32)
    if storage == "f16":
        return dev.astype(np.float16)
    u = dev.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def decode_populations(raw, storage, q_axis=0):
    """Raw storage values -> float32 populations (inverse of ``encode_populations``)."""
    raw = np.asarray(raw)
    if storage == "f32":
        return raw.astype(np.float32)
    shape = [1] * raw.ndim
    shape[q_axis] = 19
    if storage == "f16":
        dev = raw.astype(np.float32)
    else:
        dev = (raw.astype(np.uint32) << 16).view(np.float32)
    return (W32.reshape(shape) + dev).astype(np.float32)


class ManagedArray:
    """A managed allocation seen by the kernels and, zero-copy, as a NumPy array (lbm.py:525-541).

    The pages prefer the GPU; host reads fault them over, so read ``numpy`` only after the
    stream that wrote it has been synchronised.
    """

    def __init__(self, n, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        p = ctypes.c_void_p()
        check(lib().tlbm_alloc_managed(self.size * self.dtype.itemsize, ctypes.byref(p)),
              "hipMallocManaged")
        self.ptr = p.value
        ctype = {np.dtype(np.float32): ctypes.c_float, np.dtype(np.float64): ctypes.c_double}[
            self.dtype]
        self._view = np.ctypeslib.as_array((ctype * self.size).from_address(self.ptr))

    @property
    def numpy(self):
        """The NumPy view of the allocation (zero-copy; the same array object every time)."""
        return self._view

    @property
    def tack(self):
        """A device-array view of the same memory (lbm.py:539: the Tack view the kernels take),
        as accepted by ``lbm_step`` and the other kernels; it does not own the memory."""
        if self.dtype != np.float32:
            raise TypeError("only float32 managed arrays have a device-array view")
        view = DeviceArray(self.size, "f32", ptr=self.ptr)
        view._owner = self  # the managed allocation lives at least as long as its view
        return view

    def free(self):
        if getattr(self, "ptr", None):
            self._view = None
            lib().tlbm_free(self.ptr)
            self.ptr = None


class OutputArray(ManagedArray):
    """A managed output array that brings itself up to date when read: ``numpy`` first asks its
    solver to refresh the output arrays if a report step has run since their last refresh (a
    copy or coarsening of the device field it mirrors) and waits for the solver's stream.  The
    array object and its address stay the same, so readers that take it zero-copy (Conduit's
    ``set_external``, run_lbm.py:88-93) read it again after each report step."""

    def __init__(self, n, refresh):
        super().__init__(n)
        self._refresh = refresh  # weakref.WeakMethod of the solver's refresh

    @property
    def numpy(self):
        refresh = self._refresh() if self._refresh is not None else None
        if refresh is not None:
            refresh()
        return self._view


class HostBuffer:
    """Pinned host memory the GPU can write directly (used for the report-step FP64 sums)."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        p = ctypes.c_void_p()
        check(lib().tlbm_alloc_host(self.size * self.dtype.itemsize, ctypes.byref(p)),
              "hipHostMalloc")
        self.ptr = p.value
        ctype = {np.dtype(np.float64): ctypes.c_double, np.dtype(np.float32): ctypes.c_float}[
            self.dtype]
        self.numpy = np.ctypeslib.as_array((ctype * self.size).from_address(self.ptr))
        self.numpy[:] = 0

    def free(self):
        if getattr(self, "ptr", None):
            self.numpy = None
            lib().tlbm_free_host(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    if a is None:
        return None
    return a.ptr


def sync(stream=None):
    """tack.sync (lbm.py:651): wait for the stream (None: the whole device)."""
    if stream is None:
        check(lib().tlbm_device_sync(), "device synchronize")
    else:
        check(lib().tlbm_stream_sync(stream), "stream synchronize")


# ---- lattice and physics plugins (lbm.py:38-380) ---------------------------------------------------
class D3Q19:
    """lbm.py:38-55.  The velocity tables are compiled into the kernels."""
    Q = 19

    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz
        self.cx = np.array(CX, np.int32)
        self.cy = np.array(CY, np.int32)
        self.cz = np.array(CZ, np.int32)
        self.w = np.array(W, np.float32)
        self.opp = np.array(_opposite(), np.int32)
        self.miry = np.array(_mirror_y(), np.int32)


class BGK:
    """lbm.py:60-69."""
    code = _capi.BGK

    def __init__(self, nu):
        self.omega = 1.0 / (3.0 * nu + 0.5)


class TRT:
    """lbm.py:96-123: TRT with a Smagorinsky subgrid viscosity; ``magic`` and ``cs`` are class
    constants as in the reference and may be overridden per instance."""
    code = _capi.TRT
    magic = 0.1875
    cs = 0.16

    def __init__(self, nu, magic=None, cs=None):
        self.tau0 = 3.0 * nu + 0.5
        if magic is not None:
            self.magic = magic
        if cs is not None:
            self.cs = cs


class SmagorinskyLES:
    """lbm.py:126-139."""
    code = _capi.LES
    cs = 0.16

    def __init__(self, nu, cs=None):
        self.tau0 = 3.0 * nu + 0.5
        if cs is not None:
            self.cs = cs


class Cylinder:
    """lbm.py:144-169 (axis z)."""
    code = _capi.CYLINDER

    def __init__(self, x, y, r):
        self.x, self.y, self.r = x, y, r

    def obs(self):
        return (self.x, self.y, self.r, 0.0, 0.0, 0.0)


class Sphere:
    """lbm.py:172-198."""
    code = _capi.SPHERE

    def __init__(self, x, y, z, r):
        self.x, self.y, self.z, self.r = x, y, z, r

    def obs(self):
        return (self.x, self.y, self.z, self.r, 0.0, 0.0)


class Box:
    """lbm.py:201-227."""
    code = _capi.BOX

    def __init__(self, x, y, z, hx, hy, hz):
        self.x, self.y, self.z = x, y, z
        self.hx, self.hy, self.hz = hx, hy, hz

    def obs(self):
        return (self.x, self.y, self.z, self.hx, self.hy, self.hz)


class SolidField:
    """A per-cell solid fraction phi in [0, 1] (a DeviceArray of nx*ny*nz float32) in place of
    an analytic obstacle: every link is a partial bounce-back f = pulled + phi (bounced -
    pulled), and phi = 1 cells hold the rest equilibrium (lbm.py:403, 416, 428, 449 with
    fractional `solid`).  No counterpart in the reference; the obstacle of the adjoint."""
    code = _capi.FIELD

    def __init__(self, phi):
        self.phi = phi

    def obs(self):
        return (0.0,) * 6


class HeatSink:
    """lbm.py:231-268: a finned heat sink on the floor of a duct (sizes in cells).  The base
    plate and the fins (along x, ``fin_thickness`` cells every ``fin_pitch`` along z) are the
    sink; the duct's floor and ceiling rows are adiabatic walls."""
    code = _capi.HEATSINK

    def __init__(self, ny, x0, length, z0, width, base, fin_height, fin_thickness, fin_pitch):
        if int(fin_pitch) < 1 or int(fin_thickness) < 0:
            raise ValueError("fin_pitch must be >= 1 and fin_thickness >= 0")
        self.ny = ny
        self.x0, self.x1 = float(x0), float(x0 + length)
        self.z0, self.z1 = float(z0), float(z0 + width)
        self.y_base = 1.0 + base
        self.y_top = 1.0 + base + fin_height
        self.fin_thickness = int(fin_thickness)
        self.fin_pitch = int(fin_pitch)

    def obs(self):
        return (self.x0, self.x1, self.z0, self.z1, self.y_base, self.y_top,
                float(self.fin_thickness), float(self.fin_pitch))


class NoHeat:
    """lbm.py:282-308: the thermal variant compiled out (``enabled = 0``)."""
    enabled = 0
    Q = 7
    pr_t = 0.85

    def __init__(self, arrays=None):
        arrays = arrays or {}
        self.t_solid = arrays.get("t_solid")
        self.q_solid = arrays.get("q_solid")
        self.temp = arrays.get("temp")
        self.alpha = 0.0
        self.nu = 0.0


class Heat(NoHeat):
    """lbm.py:311-319: the D3Q7 temperature variant, diffusivity ``alpha`` (nu / Pr) plus the
    subgrid part max(0, (tau - 1/2) / 3 - nu) / pr_t.  ``arrays``: "t_solid", "q_solid", "temp"
    (n floats each, device or managed)."""
    enabled = 1

    def __init__(self, arrays, alpha, nu):
        super().__init__(arrays)
        self.alpha = alpha
        self.nu = nu


class Quiet:
    """lbm.py:325-334: diagnostics compiled out."""
    enabled = 0

    def __init__(self, fields):
        self.fields = fields
        for name in _capi.FIELD_NAMES:
            setattr(self, name, fields[name])


class Report(Quiet):
    """lbm.py:337-339: moments, residuals and forces compiled in."""
    enabled = 1


class PeriodicY:
    """lbm.py:342-352."""
    code = _capi.PERIODIC_Y


class SlipY:
    """lbm.py:355-369."""
    code = _capi.SLIP_Y


class LidY:
    """A moving no-slip lid halfway outside y = ny - 1 over a free-slip floor (no counterpart in
    the reference).  Ceiling links take f_q = f_opp + 6 w_q c_q . u."""
    code = _capi.LID_Y

    def __init__(self, ux=0.04, uy=0.0, uz=0.0):
        self.u = (float(ux), float(uy), float(uz))


class Flow:
    """lbm.py:372-380.  ``omega_rot`` is a one-element device/managed array so that host writes
    reach the kernel."""

    def __init__(self, u0, sponge_start, sponge_strength, omega_rot=None, side=None):
        self.u0 = u0
        self.sponge_start = sponge_start
        self.sponge_strength = sponge_strength
        self.omega_rot = zeros((1,)) if omega_rot is None else omega_rot
        self.side = PeriodicY() if side is None else side


# ---- kernels (lbm.py:383-522) ----------------------------------------------------------------------
_HANDLES = {}


def kernel_handle(collision, obstacle, diag, side, storage, heat=0):
    """The specialisation for this combination of plugin types (cached like Tack's cache)."""
    key = (collision, obstacle, diag, side, storage, heat)
    h = _HANDLES.get(key)
    if h is None:
        spec = _capi.Spec(collision, obstacle, diag, side, storage, heat)
        out = ctypes.c_int(-1)
        check(lib().tlbm_kernel_lookup(ctypes.byref(spec), ctypes.byref(out)), "kernel lookup")
        h = _HANDLES[key] = out.value
    return h


def set_launch(phys, geo, diag, side, storage, block_x=0, heat=None):
    """Threads per block along x of one specialisation (process-wide; 0 = the default)."""
    h = kernel_handle(_code(phys, "collision"), _code(geo, "obstacle"),
                      _capi.REPORT if getattr(diag, "enabled", 0) == 1 else _capi.QUIET,
                      _code(side, "side"), STORAGES[storage],
                      1 if getattr(heat, "enabled", 0) == 1 else 0)
    check(lib().tlbm_set_launch(h, int(block_x)), "set launch")


def _code(obj, what):
    code = getattr(type(obj), "code", None)
    if code is None:
        raise TypeError(f"{type(obj).__name__} is not a known {what} plugin")
    return code


def make_params(lat, phys, geo, flow, qstride=None, tile_rows=0):
    """The runtime arguments of a step: instance attributes of the plugins."""
    if isinstance(geo, HeatSink) and int(geo.ny) != lat.ny:
        raise ValueError(f"HeatSink(ny={geo.ny}) on a lattice with ny = {lat.ny}: its duct walls "
                         "are the lattice's first and last rows")
    p = _capi.Params()
    p.nx, p.ny, p.nz = lat.nx, lat.ny, lat.nz
    p.qstride = lat.n if qstride is None else int(qstride)
    p.tile_rows = int(tile_rows)
    if phys is not None:
        p.omega = getattr(phys, "omega", 0.0)
        p.tau0 = getattr(phys, "tau0", 0.0)
        p.cs = getattr(phys, "cs", 0.0)
        p.magic = getattr(phys, "magic", 0.0)
    for i, v in enumerate(geo.obs()):
        p.obs[i] = v
    p.u0 = flow.u0
    p.sponge_start = flow.sponge_start
    p.sponge_strength = flow.sponge_strength
    lid = getattr(flow.side, "u", (0.0, 0.0, 0.0))
    for i in range(3):
        p.lid[i] = lid[i]
    p.omega_rot = _ptr(flow.omega_rot)
    p.phi = _ptr(getattr(geo, "phi", None))
    return p


def default_qstride(n):
    """The population plane stride a Solver uses for n cells.  Large lattices get planes padded
    by 65792 cells (257 KiB at FP32) past a 128-cell boundary: measured on MI355X, a
    power-of-two plane size (512x256x256, 256x128x128, 384x128x64) streams 3-5% slower than a
    padded one, in FP32 and FP16, and other sizes are unaffected.  Small lattices keep n."""
    if n < (1 << 20):
        return n
    return (n + 127) // 128 * 128 + 65792


def default_tile_rows(nx, ny, nz, storage="f32"):
    """Rows per population tile for a Solver of nx x ny x nz cells (0: q-major planes): the
    largest power of two k with k nx cells of one plane in at most 32 KiB at FP32 (k nx <= 8192
    cells) and 8 KiB at FP16/BF16 (k nx <= 4096).  Measured on MI355X (LES, sphere, lid, quiet
    step, FP32 with non-temporal device-scope stores): 512x256x256 0.806 ms with 16 rows against
    0.814-0.819 with 4, 0.827 with 8 and 0.834 with 32; 256x256x256 0.403 ms with 32 rows against
    0.414-0.419 with 4 to 16; 1024x256x128 0.802-0.805 with 8 against 0.820-0.826 with 2 or 4;
    512x128x128 0.2045 with 16 against 0.206-0.209; FP16 512x256x256 0.417-0.418 ms with 8, 16
    or 32 rows (0.423 with 4, 0.451 with 64), 1024x256x128 0.415-0.419 with 4 against 0.421 with
    8 and 0.430 with 16.  Q-major planes: 0.891-0.904 ms (FP32), 0.458-0.461 (FP16) at
    512x256x256 before the per-wave row addressing.  Only for large lattices whose rows are
    whole waves (nx a multiple of 64), where the step addresses a wave's source rows once for
    all its lanes."""
    per = 8192 if storage == "f32" else 4096
    if nx * ny * nz < (1 << 20) or nx % 64 or nx > per:
        return 0
    k = 1
    while 2 * k * nx <= per:
        k *= 2
    return k


def population_elems(lat, qstride=None, tile_rows=0, q=19):
    """Elements of one population buffer of ``q`` planes (19, or 7 for the temperature): q planes
    of ``qstride`` (default n), or for tiles of k = ``tile_rows`` rows, q planes of k nx cells per
    tile (the last tile full size)."""
    if tile_rows:
        k = int(tile_rows)
        return q * k * lat.nx * -(-(lat.ny * lat.nz) // k)
    return q * (lat.n if qstride is None else int(qstride))


def untile_populations(raw, lat, tile_rows, q=19):
    """A tiled buffer of ``q`` planes (flat, ``population_elems`` long) as q-major planes (q, n)."""
    k = int(tile_rows)
    tiles = -(-(lat.ny * lat.nz) // k)
    a = np.asarray(raw)[:q * k * lat.nx * tiles].reshape(tiles, q, k * lat.nx)
    return a.transpose(1, 0, 2).reshape(q, tiles * k * lat.nx)[:, :lat.n]


def tile_populations(planes, lat, tile_rows):
    """q-major planes (q, n) as a flat tiled buffer (inverse of ``untile_populations``; the rows
    past the lattice in the last tile are zero)."""
    k = int(tile_rows)
    tiles = -(-(lat.ny * lat.nz) // k)
    planes = np.asarray(planes)
    q = planes.shape[0]
    out = np.zeros((q, tiles * k * lat.nx), planes.dtype)
    out[:, :lat.n] = planes.reshape(q, lat.n)
    return out.reshape(q, tiles, k * lat.nx).transpose(1, 0, 2).reshape(-1)


def _check_tiled(lat, f, tile_rows):
    need = population_elems(lat, None, tile_rows)
    if f.size < need:
        raise ValueError(f"population buffer holds {f.size} < {need} elements (tiles of "
                         f"{tile_rows} rows)")


def _qstride(lat, f):
    if f.size % 19:
        raise ValueError(f"population buffer of {f.size} elements is not 19 planes")
    qs = f.size // 19
    if qs < lat.n:
        raise ValueError(f"population buffer holds {f.size} < 19 * {lat.n} elements")
    return qs


def lbm_init(lat, geo, flow, f, seed_amp, noise_amp=0.0, noise_seed=0, stream=None,
             tile_rows=0):
    """lbm.py:383-394, plus the optional seeded velocity noise U(-noise_amp, noise_amp).
    ``tile_rows`` > 0: ``f`` holds tiles of that many rows (``population_elems``)."""
    h = kernel_handle(_capi.BGK, _code(geo, "obstacle"), _capi.QUIET, _capi.PERIODIC_Y,
                      STORAGES[f.storage])
    if tile_rows:
        _check_tiled(lat, f, tile_rows)
        p = make_params(lat, None, geo, flow, lat.n, tile_rows)
    else:
        p = make_params(lat, None, geo, flow, _qstride(lat, f))
    check(lib().tlbm_init(h, f.ptr, ctypes.byref(p), seed_amp, noise_amp, noise_seed, stream),
          "lbm_init")


def _fields_struct(fields):
    fs = _capi.Fields()
    for name in _capi.FIELD_NAMES:
        setattr(fs, name, fields[name].ptr)
    return fs


def lbm_step(lat, phys, geo, flow, diag, f_in, f_out, heat=None, g_in=None, g_out=None,
             sums=None, stream=None, block_x=0, tile_rows=0, xcd_group=0):
    """lbm.py:397-513: one step from ``f_in`` to ``f_out``.

    ``diag`` is ``Quiet``/``Report``; a ``Report`` step writes its 13 fields and, when ``sums``
    (a ``HostBuffer``/``ManagedArray`` of 14 float64) is given, their FP64 totals plus the total
    mass.  With ``heat.enabled == 1`` (a ``Heat``) the temperature populations step from ``g_in``
    to ``g_out`` too (FP32), and a report step writes ``heat.temp`` and adds to
    ``heat.q_solid``.  ``block_x``: threads per block of this launch (0: the specialisation's
    default).  ``tile_rows`` > 0: the population buffers hold tiles of that many rows
    (``population_elems``), and so do the temperature buffers (7 planes per tile,
    ``population_elems(lat, None, tile_rows, 7)``); otherwise the temperature planes share the
    populations' plane stride.  ``xcd_group``: a quiet step's chunk order (0: the library's
    choice).
    """
    if f_in.storage != f_out.storage:
        raise ValueError(f"f_in is {f_in.storage}, f_out is {f_out.storage}")
    report = getattr(diag, "enabled", 0) == 1
    hot = heat is not None and getattr(heat, "enabled", 0) == 1
    h = kernel_handle(_code(phys, "collision"), _code(geo, "obstacle"),
                      _capi.REPORT if report else _capi.QUIET, _code(flow.side, "side"),
                      STORAGES[f_in.storage], 1 if hot else 0)
    if tile_rows:
        _check_tiled(lat, f_in, tile_rows)
        _check_tiled(lat, f_out, tile_rows)
        qs = lat.n
    else:
        qs = _qstride(lat, f_in)
        if _qstride(lat, f_out) != qs:
            raise ValueError("f_in and f_out have different plane strides")
    p = make_params(lat, phys, geo, flow, qs, tile_rows)
    p.block_x = int(block_x)
    p.xcd_group = int(xcd_group)
    fs = _fields_struct(diag.fields) if report else None
    if not hot:
        check(lib().tlbm_step(h, f_in.ptr, f_out.ptr, ctypes.byref(p),
                              ctypes.byref(fs) if fs is not None else None,
                              _ptr(sums) if report else None, stream), "lbm_step")
        return
    if g_in is None or g_out is None:
        raise ValueError("the temperature variant needs g_in and g_out")
    need = population_elems(lat, qs, tile_rows, 7)
    for g in (g_in, g_out):
        if g.storage != "f32" or g.size < need:
            raise ValueError(f"g buffers need {need} FP32 elements (got {g.size} {g.storage})")
    ht = _capi.Heat(_ptr(heat.t_solid), _ptr(heat.q_solid), _ptr(heat.temp), heat.alpha,
                    heat.nu, heat.pr_t)
    check(lib().tlbm_step_heat(h, f_in.ptr, f_out.ptr, g_in.ptr, g_out.ptr, ctypes.byref(p),
                               ctypes.byref(ht), ctypes.byref(fs) if fs is not None else None,
                               _ptr(sums) if report else None, stream), "lbm_step (heat)")


def coarsen(lat, s, src, dst, mx, my, mz, stream=None):
    """lbm.py:516-522."""
    check(lib().tlbm_coarsen(src.ptr, dst.ptr, lat.nx, lat.ny, lat.nz, s, mx, my, mz, stream),
          "coarsen")


COLLISIONS = {"bgk": BGK, "trt": TRT, "les": SmagorinskyLES}
FIELDS = ("ux", "uy", "uz", "p", "vmag", "r_rho", "r_ux", "r_uy", "r_uz", "r_ke", "fx", "fy", "fz")
SIDES = {"periodic": PeriodicY, "slip": SlipY, "lid": LidY}


def _select_device(device):
    if device in ("cuda", "hip", None):
        index = None
    elif isinstance(device, str) and device.split(":")[0] in ("cuda", "hip") and ":" in device:
        index = int(device.split(":")[1])
    else:
        raise ValueError(f"device must be 'cuda' or 'cuda:N', got {device!r}")
    if _capi.device_count() < 1:
        raise RuntimeError("no HIP device is visible")
    if index is not None:
        check(lib().tlbm_set_device(index), "set device")


class Solver:
    """lbm.py:548-661: owns the two population buffers, the 13 managed fields and the plugins."""
    OUTPUT = ("ux", "uy", "uz", "p", "vmag")

    def __init__(self, nx=384, ny=128, nz=64, collision="les", obstacle="cylinder",
                 reynolds=300.0, u0=0.08, diameter=24.0, device="cuda", stride=1,
                 heat=False, prandtl=0.7, lateral="periodic", *, storage="f32",
                 lid_u=(0.04, 0.0, 0.0), center=None, sponge_start=None, sponge_strength=1.0,
                 cs=None, magic=None, seed_amp=0.05, noise_amp=0.0, noise_seed=0,
                 qstride=None, tile_rows=None, stream=None, field_memory="device"):
        if heat and storage != "f32":
            raise ValueError("the temperature variant needs FP32 populations (storage='f32')")
        _select_device(device)
        self.stream = stream
        self.lateral = lateral
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lat = D3Q19(nx, ny, nz)
        self.u0, self.diameter = u0, diameter
        self.storage = storage
        self.step_count = 0
        self.block_x = 0
        self.xcd_group = 0
        self._omega_state = ManagedArray(1)
        side = LidY(*lid_u) if lateral == "lid" else SIDES[lateral]()
        self.flow = Flow(u0, 0.75 * nx if sponge_start is None else sponge_start,
                         sponge_strength, self._omega_state, side)
        self._center = center
        self._cs, self._magic = cs, magic
        self.set_obstacle(obstacle)
        self.set_physics(collision, reynolds)
        n = self.lat.n
        self.qstride = default_qstride(n) if qstride is None else int(qstride)
        if self.qstride < n:
            raise ValueError(f"qstride {self.qstride} < cells {n}")
        self.tile_rows = (default_tile_rows(nx, ny, nz, storage) if tile_rows is None
                          else int(tile_rows))
        if self.tile_rows < 0 or self.tile_rows & (self.tile_rows - 1):
            raise ValueError(f"tile_rows {self.tile_rows} is not 0 or a power of two")
        elems = population_elems(self.lat, self.qstride, self.tile_rows)
        self.f = [empty((elems,), storage), empty((elems,), storage)]
        # Report fields.  "device": HBM, written by the report kernel at full bandwidth, with
        # host-visible managed copies of the output fields refreshed by update_output().
        # "managed": the reference's layout (lbm.py:567), the kernel writes managed memory
        # directly; on devices without XNACK/HMM that memory is host-resident and slow to write.
        if field_memory not in ("device", "managed"):
            raise ValueError(f"field_memory must be 'device' or 'managed', got {field_memory!r}")
        self.field_memory = field_memory
        if field_memory == "device":
            self.fields = {name: DeviceField(n, stream) for name in FIELDS}
        else:
            self.fields = {name: ManagedArray(n) for name in FIELDS}
        self.quiet = Quiet(self.fields)
        self.report = Report(self.fields)
        self.sums = HostBuffer(_capi.TLBM_NSUMS)
        self.heat_enabled = bool(heat)
        self.prandtl = prandtl
        self._setup_heat(n)
        self.stride = stride
        self.out_dims = tuple((d + stride - 1) // stride for d in (nx, ny, nz))
        outputs = self.OUTPUT + (("temp",) if self.heat_enabled else ())
        # Output arrays: the fields themselves when they are managed and not coarsened (the
        # reference's layout, lbm.py:577-580); otherwise managed copies that refresh when read.
        self._output_stale = False
        if stride > 1 or field_memory == "device":
            m = self.out_dims[0] * self.out_dims[1] * self.out_dims[2]
            refresh = weakref.WeakMethod(self._refresh_output)
            self.output = {name: OutputArray(m, refresh) for name in outputs}
        else:
            self.output = {name: self.heat_fields.get(name) or self.fields[name]
                           for name in outputs}
        lbm_init(self.lat, self.geo, self.flow, self.f[0], seed_amp, noise_amp, noise_seed,
                 self.stream, self.tile_rows)
        self.step(report=True)
        self.step_count = 0

    def _setup_heat(self, n):
        """lbm.py:586-600: the temperature fields (n floats each when heat is on, else one)
        and the D3Q7 population buffers, zero (lbm.py:594).

        Also callable on a bare ``Solver.__new__(Solver)`` that has only ``heat_enabled`` set,
        as trt_ref.py:33-36 does to get ``heat``, ``g[0]`` and ``g[1]`` for a direct
        ``lbm_step`` call: unset attributes take the reference's layout (managed temperature,
        planes of n cells)."""
        small = n if self.heat_enabled else 1
        stream = getattr(self, "stream", None)
        managed = getattr(self, "field_memory", "managed") == "managed"
        temp = ManagedArray(small) if managed else DeviceField(small, stream)
        # q_solid takes float atomic adds from the kernel: device memory in every layout
        self.heat_fields = {"t_solid": ManagedArray(small), "q_solid": DeviceField(small, stream),
                            "temp": temp}
        # the temperature planes take the populations' layout: their plane stride, or their tiles
        # of rows with 7 planes per tile
        tiles = getattr(self, "tile_rows", 0) if self.heat_enabled else 0
        if tiles:
            elems = population_elems(self.lat, None, tiles, 7)
        else:
            elems = 7 * (getattr(self, "qstride", n) if self.heat_enabled else 1)
        self.g = [zeros((elems,)), zeros((elems,))]
        self._make_heat()

    def _make_heat(self):
        if self.heat_enabled:
            self.heat = Heat(self.heat_fields, self.nu / self.prandtl, self.nu)
        else:
            self.heat = NoHeat(self.heat_fields)

    def set_wall_temperature(self, value):
        """lbm.py:602-604: fill every solid voxel's temperature (t_solid is a managed array;
        set parts of it through ``heat_fields["t_solid"].numpy``)."""
        self.heat_fields["t_solid"].numpy[:] = value

    def q_solid(self):
        """A host copy of q_solid: per sink voxel, the sum of the temperatures of the fluid cells
        facing it, accumulated by the last report step (lbm.py:484-492)."""
        return self.heat_fields["q_solid"].numpy.copy()

    def update_output(self):
        """lbm.py:606-611: refresh the host-visible output arrays (coarsened when stride > 1),
        enqueued on the solver's stream."""
        def source(name):
            return self.heat_fields[name] if name == "temp" else self.fields[name]
        if self.stride > 1:
            mx, my, mz = self.out_dims
            for name in self.output:
                coarsen(self.lat, self.stride, source(name), self.output[name], mx, my, mz,
                        self.stream)
        elif self.field_memory == "device":
            for name in self.output:
                src = source(name)
                check(lib().tlbm_memcpy_async(self.output[name].ptr, src.ptr,
                                              4 * self.lat.n, self.stream), "output copy")
        self._output_stale = False

    def _refresh_output(self):
        """Bring the output arrays up to date with the last report step and wait for them."""
        if self._output_stale:
            self.update_output()
        sync(self.stream)

    def set_launch(self, block_x=0, xcd_group=0):
        """Threads per block of this solver's steps (64, 128, 192 or 256; 0: the default) and
        the quiet steps' chunk order (blocks go to the XCDs in runs of ``xcd_group``, a power of
        two up to 64; 0: the library's choice for the storage).  The choice belongs to this
        solver only (it travels in its step parameters)."""
        if block_x not in (0, 64, 128, 192, 256):
            raise ValueError(f"block_x must be 0, 64, 128, 192 or 256, got {block_x}")
        if xcd_group < 0 or xcd_group > 64 or xcd_group & (xcd_group - 1):
            raise ValueError(f"xcd_group must be 0 or a power of two up to 64, got {xcd_group}")
        self._drop_graphs()
        self.block_x = int(block_x)
        self.xcd_group = int(xcd_group)

    def set_physics(self, collision, reynolds):
        """lbm.py:613-616: nu = u0 D / Re; a collision change selects another kernel."""
        self._drop_graphs()
        self.collision_name, self.reynolds = collision, float(reynolds)
        self.nu = self.u0 * self.diameter / self.reynolds
        kw = {}
        if collision in ("trt", "les") and self._cs is not None:
            kw["cs"] = self._cs
        if collision == "trt" and self._magic is not None:
            kw["magic"] = self._magic
        self.phys = COLLISIONS[collision](self.nu, **kw)
        if getattr(self, "heat_enabled", False):
            # the reference means to rebuild Heat here (lbm.py:625-626, unreachable after the
            # return of omega_rot); the diffusivity follows the new viscosity
            self._make_heat()

    def set_actuation(self, omega_rot):
        """lbm.py:618-620: cylinder surface speed in lattice units."""
        self._omega_state.numpy[0] = float(omega_rot)

    @property
    def omega_rot(self):
        return float(self._omega_state.numpy[0])

    def set_obstacle(self, kind):
        """lbm.py:628-637."""
        self._drop_graphs()
        nx, ny, nz, r = self.nx, self.ny, self.nz, self.diameter / 2
        if self._center is None:
            x0, y0, z0 = 0.22 * nx, 0.5 * ny + 1.7, 0.5 * nz
        else:
            x0, y0, z0 = self._center
        self.obstacle_name = kind
        makers = {"cylinder": lambda: Cylinder(x0, y0, r),
                  "sphere": lambda: Sphere(x0, y0, z0, r),
                  "box": lambda: Box(x0, y0, z0, r, r, 0.3 * nz),
                  "heatsink": lambda: HeatSink(ny, int(0.2 * nx), int(0.3 * nx), int(0.125 * nz),
                                               int(0.75 * nz), round(self.diameter / 2),
                                               int(0.45 * ny), round(self.diameter / 3),
                                               round(self.diameter))}
        if kind not in makers:
            raise KeyError(f"obstacle {kind!r} is not built (have {sorted(makers)})")
        self.geo = makers[kind]()

    def _handle(self, report=False):
        """The kernel handle of this solver's quiet or report step."""
        return kernel_handle(_code(self.phys, "collision"), _code(self.geo, "obstacle"),
                             _capi.REPORT if report else _capi.QUIET,
                             _code(self.flow.side, "side"), STORAGES[self.storage],
                             1 if self.heat_enabled else 0)

    def step(self, report=False):
        """lbm.py:639-647: one step, then swap the ping-pong buffers."""
        diag = self.report if report else self.quiet
        if report and self.heat_enabled:
            self.heat_fields["q_solid"].zero_(self.stream)
        lbm_step(self.lat, self.phys, self.geo, self.flow, diag, self.f[0], self.f[1],
                 self.heat, self.g[0], self.g[1], sums=self.sums if report else None,
                 stream=self.stream, block_x=getattr(self, "block_x", 0),
                 tile_rows=getattr(self, "tile_rows", 0),
                 xcd_group=getattr(self, "xcd_group", 0))
        self.f.reverse()
        self.g.reverse()
        self.step_count += 1
        if report:
            self._output_stale = True

    def run(self, steps, report_every=0):
        """``steps`` steps (even), a report step every ``report_every`` of them (0: none), as
        ``step()`` would take them, replayed from a hipGraph captured on first use for this
        (steps, report_every).  Needs a created stream (``Solver(stream=...)``).  Graphs hold
        the plugins' parameters as they were at capture: ``set_physics``, ``set_obstacle`` and
        ``set_launch`` drop them (``set_actuation`` and wall temperatures are read from memory
        at every step and need no new capture)."""
        steps, report_every = int(steps), int(report_every)
        if self.stream is None:
            raise ValueError("graph replay needs a created stream (Solver(stream=...))")
        if steps <= 0 or steps % 2:
            raise ValueError(f"steps must be even and > 0, got {steps}")
        if report_every < 0:
            raise ValueError(f"report_every must be >= 0, got {report_every}")
        if self.heat_enabled and report_every:
            raise ValueError("report steps of the temperature variant clear q_solid with a "
                             "memset, which a graph must not hold: use step()")
        # a graph fixes which ping-pong buffer each of its steps reads: key it on the current
        # order too (an odd number of step() calls between two runs swaps it)
        key = (steps, report_every, self.f[0].ptr)
        graph = self._graphs.get(key)
        if graph is None:
            graph = self._graphs[key] = self._capture(steps, report_every)
        check(lib().tlbm_graph_launch(graph, self.stream), "graph launch")
        self.step_count += steps
        if report_every:
            self._output_stale = True

    def _capture(self, steps, report_every):
        L = lib()
        order, count = (list(self.f), list(self.g)), self.step_count
        out = ctypes.c_void_p()
        check(L.tlbm_capture_begin(self.stream), "capture begin")
        try:
            for i in range(steps):
                self.step(report=report_every > 0 and (i + 1) % report_every == 0)
        except Exception:
            if L.tlbm_capture_end(self.stream, ctypes.byref(out)) == 0:
                L.tlbm_graph_destroy(out)
            raise
        finally:
            self.f, self.g = order
            self.step_count = count
        check(L.tlbm_capture_end(self.stream, ctypes.byref(out)), "capture end")
        return out.value

    def _drop_graphs(self):
        for graph in getattr(self, "_graphs", {}).values():
            lib().tlbm_graph_destroy(graph)
        self._graphs = {}

    def __del__(self):
        try:
            self._drop_graphs()
        except Exception:
            pass

    def monitors(self, interval):
        """lbm.py:649-661, from the FP64 sums of the last report step.  As in the reference,
        the output grid is coarsened here when stride > 1; at stride 1 the output arrays copy
        the device fields only when read (``OutputArray``), so monitors itself moves no field."""
        if self.stride > 1:
            self.update_output()
        sync(self.stream)
        n = float(self.lat.n)
        s = self.sums.numpy
        total = {key: float(s[i]) for i, key in enumerate(FIELDS)}
        res = {name: float(np.sqrt(total[key] / n)) / max(interval, 1)
               for name, key in (("continuity", "r_rho"), ("momentum_x", "r_ux"),
                                 ("momentum_y", "r_uy"), ("momentum_z", "r_uz"),
                                 ("energy", "r_ke"))}
        scale = 0.5 * self.u0 * self.u0 * self.diameter * self.nz
        res["drag"] = total["fx"] / scale
        res["lift"] = total["fy"] / scale
        res["mass"] = float(s[13])
        return res

    def populations(self):
        """The current populations as float32, shaped (19, n) (unshifted for f16/bf16)."""
        sync(self.stream)
        if self.tile_rows:
            raw = untile_populations(self.f[0].numpy(), self.lat, self.tile_rows)
        else:
            raw = self.f[0].numpy().reshape(19, self.qstride)[:, :self.lat.n]
        return decode_populations(raw, self.storage)

    def temperature_populations(self):
        """The current D3Q7 temperature populations (heat=True) as float32, shaped (7, n)."""
        if not self.heat_enabled:
            raise ValueError("this solver has no temperature populations (heat=False)")
        sync(self.stream)
        if self.tile_rows:
            return untile_populations(self.g[0].numpy(), self.lat, self.tile_rows, 7)
        return self.g[0].numpy().reshape(7, self.qstride)[:, :self.lat.n].copy()
