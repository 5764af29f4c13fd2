"""Gradient of the time-averaged drag on a smoothed sphere with a porous shell, by the discrete
adjoint of the fused step (no counterpart in the reference).

The forward model is the solid-fraction BGK step (lbm.py:397-456 read with a fractional
``solid``; periodic in x, y and z; FP32 populations).  The design variables are the sphere's
radius r and centre c and a porosity block P (bx x by x bz cells at ``origin``):

    s(x)     = clip(1/2 + (r - |x - c|) / (2 eps), 0, 1)
    phi_s(x) = s^3 (10 - 15 s + 6 s^2)                        (quintic smootherstep)
    phi(x)   = 1 - (1 - phi_s(x)) (1 - P(x))                 (P = 0 outside the block)

and the objective is the mean x force on the solid over the last ``window`` of ``steps`` steps,
J = (1 / window) sum_t F_x(t) (the momentum exchange of lbm.py:429-430).  The initial state is
the uniform inflow equilibrium, independent of the design.

The forward pass keeps the state every ``checkpoint_every`` steps.  The backward pass walks the
segments from last to first: it recomputes a segment's states from its checkpoint, runs the
adjoint kernel (``tlbm_adjoint_step``) from the segment's last step to its first, and finally
adds the links of step 0.  The kernel leaves dJ/dphi on the block in FP64; the chain rule to
(r, c, P) is the analytic derivative of phi above.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import check, lib
from .lbm import (BGK, D3Q19, DeviceArray, Flow, HostBuffer, PeriodicY, SolidField, array,
                  kernel_handle, lbm_init, make_params, sync, zeros)

_SUMS = _capi.TLBM_NSUMS
_FX = 10  # the x force in the report sums (lbm.py:545 FIELDS order)


class _F64:
    """A device buffer of FP64 values (the dJ/dphi accumulator)."""

    def __init__(self, n):
        self.size = int(n)
        p = ctypes.c_void_p()
        check(lib().tlbm_alloc_device(8 * self.size, ctypes.byref(p)), "device allocation")
        self.ptr = p.value

    def zero_(self, stream=None):
        check(lib().tlbm_memset(self.ptr, 0, 8 * self.size, stream), "memset")

    def numpy(self):
        out = np.empty(self.size, np.float64)
        check(lib().tlbm_memcpy(out.ctypes.data, self.ptr, 8 * self.size), "device-to-host copy")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().tlbm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SmoothedSphere:
    """The solid fraction of the design block and its derivatives (FP64, block-shaped (bz, by,
    bx), x fastest)."""

    def __init__(self, origin, size, eps=1.5):
        self.origin = tuple(int(v) for v in origin)
        self.size = tuple(int(v) for v in size)
        self.eps = float(eps)
        bx, by, bz = self.size
        ox, oy, oz = self.origin
        self.z, self.y, self.x = np.meshgrid(np.arange(oz, oz + bz, dtype=np.float64),
                                             np.arange(oy, oy + by, dtype=np.float64),
                                             np.arange(ox, ox + bx, dtype=np.float64),
                                             indexing="ij")

    def fits(self, center, radius):
        """True when the sphere's smoothed surface (radius + eps) lies inside the block, so that
        phi_s = 0 on every cell outside it."""
        reach = radius + self.eps
        for a in range(3):
            lo = self.origin[a]
            hi = self.origin[a] + self.size[a] - 1
            if center[a] - reach <= lo - 1 or center[a] + reach >= hi + 1:
                return False
        return True

    def evaluate(self, center, radius, porosity):
        """Returns (phi, dphi/dr, dphi/dc (3, ...), dphi/dP)."""
        cx, cy, cz = (float(v) for v in center)
        dx, dy, dz = self.x - cx, self.y - cy, self.z - cz
        d = np.sqrt(dx * dx + dy * dy + dz * dz)
        s = np.clip(0.5 + (float(radius) - d) / (2.0 * self.eps), 0.0, 1.0)
        sph = s * s * s * (10.0 + s * (-15.0 + 6.0 * s))
        ds = np.where((s > 0.0) & (s < 1.0), 30.0 * s * s * (1.0 - s) * (1.0 - s), 0.0)
        dsph_dr = ds / (2.0 * self.eps)
        inv_d = np.where(d > 0.0, 1.0 / np.maximum(d, 1e-300), 0.0)
        dsph_dc = np.stack([dsph_dr * dx * inv_d, dsph_dr * dy * inv_d, dsph_dr * dz * inv_d])
        P = np.asarray(porosity, np.float64).reshape(self.size[::-1])
        phi = 1.0 - (1.0 - sph) * (1.0 - P)
        return phi, dsph_dr * (1.0 - P), dsph_dc * (1.0 - P), 1.0 - sph


class SolidFractionAdjoint:
    """J = (1 / window) sum of the x force over the last ``window`` of ``steps`` steps of the
    solid-fraction BGK lattice, and dJ/dphi over a block of cells, for any solid fraction phi
    and initial state f0 (the design-independent part of the problem).

    ``forward(phi, f0=None)`` runs the steps, keeping the state every ``checkpoint_every``
    steps, and returns J (f0 = None: the uniform inflow equilibrium).  ``backward()`` returns
    dJ/dphi on the block (FP64, shape (bz, by, bx)) for the last forward run.
    """

    def __init__(self, nx, ny, nz, nu, u0, steps, window, checkpoint_every, block_origin,
                 block_size, sponge_start=None, sponge_strength=1.0, stream=None):
        if not 0 < window <= steps:
            raise ValueError(f"window {window} must be in (0, steps={steps}]")
        if checkpoint_every <= 0:
            raise ValueError("checkpoint_every must be > 0")
        self.lat = D3Q19(nx, ny, nz)
        self.n = self.lat.n
        self.u0 = float(u0)
        self.phys = BGK(nu)
        self.steps, self.window, self.k = int(steps), int(window), int(checkpoint_every)
        self.stream = stream
        self.origin = tuple(int(v) for v in block_origin)
        self.size = tuple(int(v) for v in block_size)
        for a, (o, b, m) in enumerate(zip(self.origin, self.size, (nx, ny, nz))):
            if b <= 0 or o < 0 or o + b > m:
                raise ValueError(f"block [{o}, {o + b}) is not inside [0, {m}) on axis {a}")
        self.sponge_start = float(nx - 8 if sponge_start is None else sponge_start)
        self.sponge_strength = float(sponge_strength)
        self._omega_rot = array(np.zeros(1, np.float32))
        self.flow = Flow(self.u0, self.sponge_start, self.sponge_strength, self._omega_rot,
                         PeriodicY())
        self.phi = zeros((self.n,))
        self.geo = SolidField(self.phi)
        # population planes padded past a 64-cell boundary (power-of-two plane sizes stream
        # slower on MI355X); all state and adjoint buffers share this stride
        self.qstride = self.n if self.n < 4096 else (self.n + 63) // 64 * 64 + 256
        self.params = make_params(self.lat, self.phys, self.geo, self.flow, self.qstride)
        self.quiet = kernel_handle(_capi.BGK, _capi.FIELD, _capi.QUIET, _capi.PERIODIC_Y,
                                   _capi.F32)
        self.report = kernel_handle(_capi.BGK, _capi.FIELD, _capi.REPORT, _capi.PERIODIC_Y,
                                    _capi.F32)
        nseg = (self.steps + self.k - 1) // self.k
        self.ckpt = [DeviceArray(19 * self.qstride) for _ in range(nseg)]
        self.f = [DeviceArray(19 * self.qstride), DeviceArray(19 * self.qstride)]
        self.sums = HostBuffer(self.window * _SUMS)
        self._seg = None
        self._mu = None
        self._gphi = None
        self._ran = False
        self.drag_history = None
        self.forward_ms = None
        self.backward_ms = None

    def lam(self, t):
        """dJ/dF_x(t): 1 / window inside the averaging window, 0 elsewhere (and at t = steps)."""
        return 1.0 / self.window if self.steps - self.window <= t < self.steps else 0.0

    # ---- launches --------------------------------------------------------------------------------
    def _step(self, fin, fout, sums_row=None):
        L = lib()
        if sums_row is None:
            rc = L.tlbm_step(self.quiet, fin.ptr, fout.ptr, ctypes.byref(self.params), None,
                             None, self.stream)
        else:
            rc = L.tlbm_step(self.report, fin.ptr, fout.ptr, ctypes.byref(self.params), None,
                             self.sums.ptr + 8 * _SUMS * sums_row, self.stream)
        check(rc, "adjoint forward step")

    def _adjoint(self, f_t, f_t1, mu_in, mu_out, lam_t, lam_t1, final=False):
        o = (ctypes.c_int32 * 3)(*self.origin)
        s = (ctypes.c_int32 * 3)(*self.size)
        check(lib().tlbm_adjoint_step(ctypes.byref(self.params),
                                      f_t.ptr if f_t is not None else None, f_t1.ptr, mu_in.ptr,
                                      mu_out.ptr if mu_out is not None else None, lam_t, lam_t1,
                                      o, s, self._gphi.ptr, 1 if final else 0, self.stream),
              "adjoint step")

    def _copy(self, dst, src):
        check(lib().tlbm_memcpy_async(dst.ptr, src.ptr, src.nbytes, self.stream), "copy")

    # ---- forward ---------------------------------------------------------------------------------
    def forward(self, phi, f0=None):
        phi = np.ascontiguousarray(phi, np.float32).reshape(-1)
        if phi.size != self.n:
            raise ValueError(f"phi has {phi.size} cells, the lattice {self.n}")
        self.phi.copy_from(phi)
        if f0 is None:
            zero_geo = SolidField(zeros((self.n,)))  # the initial state ignores phi
            lbm_init(self.lat, zero_geo, self.flow, self.f[0], 0.0, stream=self.stream)
            sync(self.stream)
            zero_geo.phi.free()
        else:
            f0 = np.ascontiguousarray(f0, np.float32).reshape(-1)
            if f0.size != 19 * self.n:
                raise ValueError(f"f0 has {f0.size} values, the lattice needs {19 * self.n}")
            padded = np.zeros((19, self.qstride), np.float32)
            padded[:, :self.n] = f0.reshape(19, self.n)
            self.f[0].copy_from(padded)
        first = self.steps - self.window
        t0 = _event(self.stream)
        for t in range(self.steps):
            if t % self.k == 0:
                self._copy(self.ckpt[t // self.k], self.f[0])
            self._step(self.f[0], self.f[1], t - first if t >= first else None)
            self.f.reverse()
        t1 = _event(self.stream)
        sync(self.stream)
        self.forward_ms = _elapsed(t0, t1)
        fx = self.sums.numpy.reshape(self.window, _SUMS)[:, _FX]
        self.drag_history = fx.copy()
        self._ran = True
        return float(fx.sum() / self.window)

    # ---- backward --------------------------------------------------------------------------------
    def backward(self, grad_output=1.0):
        if not self._ran:
            raise RuntimeError("backward() needs a forward() first")
        if self._seg is None:
            self._seg = [DeviceArray(19 * self.qstride) for _ in range(min(self.k, self.steps))]
            self._mu = [DeviceArray(19 * self.qstride), DeviceArray(19 * self.qstride)]
            self._gphi = _F64(int(np.prod(self.size)))
        self._mu[0].zero_(self.stream)
        self._gphi.zero_(self.stream)
        t0 = _event(self.stream)
        for s in reversed(range(len(self.ckpt))):
            a = s * self.k
            b = min(self.steps, a + self.k)
            prev = self.ckpt[s]
            for i in range(b - a):                      # recompute f_{a+1} .. f_b
                self._step(prev, self._seg[i])
                prev = self._seg[i]
            for t in reversed(range(a, b)):             # mu^{t+1} -> mu^t
                i = t - a
                f_t = self.ckpt[s] if i == 0 else self._seg[i - 1]
                self._adjoint(f_t, self._seg[i], self._mu[0], self._mu[1], self.lam(t),
                              self.lam(t + 1))
                self._mu.reverse()
        self._adjoint(None, self.ckpt[0], self._mu[0], None, 0.0, self.lam(0), final=True)
        t1 = _event(self.stream)
        sync(self.stream)
        self.backward_ms = _elapsed(t0, t1)
        return self._gphi.numpy().reshape(self.size[::-1]) * float(grad_output)

    def close(self):
        for buf in self.ckpt + self.f + (self._seg or []) + (self._mu or []):
            buf.free()
        if self._gphi is not None:
            self._gphi.free()
        self.phi.free()
        self._omega_rot.free()
        self.sums.free()


class DragAdjoint:
    """d(mean drag)/d(radius, centre, porosity) of a smoothed sphere with a porosity block.

    ``forward(radius, center, porosity)`` returns J; ``backward()`` returns the gradient of the
    last forward run as a dict: radius, center (3,), porosity and phi (block-shaped (bz, by, bx)).
    The viscosity is fixed by the nominal radius: nu = u0 * 2 radius / reynolds.
    """

    def __init__(self, nx=128, ny=64, nz=64, reynolds=200.0, u0=0.08, radius=10.0,
                 nominal_center=None, block=32, steps=2000, window=1000, checkpoint_every=100,
                 eps=1.5, sponge_start=None, sponge_strength=1.0, stream=None):
        if nominal_center is None:
            nominal_center = (nx / 4.0, ny / 2.0, nz / 2.0)
        self.nominal_center = tuple(float(v) for v in nominal_center)
        size = (int(block),) * 3 if np.isscalar(block) else tuple(int(b) for b in block)
        origin = tuple(int(round(c - b / 2.0)) for c, b in zip(self.nominal_center, size))
        self.nu = float(u0) * 2.0 * float(radius) / float(reynolds)
        self.run = SolidFractionAdjoint(nx, ny, nz, self.nu, u0, steps, window,
                                        checkpoint_every, origin, size, sponge_start,
                                        sponge_strength, stream)
        self.geometry = SmoothedSphere(origin, size, eps)
        self.lat = self.run.lat
        self._parts = None

    def solid_fraction(self, radius, center, porosity):
        """The full-lattice phi (float32, flat) and its derivative parts on the block."""
        center = tuple(float(v) for v in np.asarray(center, np.float64).reshape(3))
        radius = float(radius)
        if not self.geometry.fits(center, radius):
            raise ValueError(f"sphere at {center} radius {radius} (+ eps {self.geometry.eps}) "
                             f"leaves the design block {self.geometry.origin} + "
                             f"{self.geometry.size}")
        phi_b, dr, dc, dP = self.geometry.evaluate(center, radius, porosity)
        phi = np.zeros((self.lat.nz, self.lat.ny, self.lat.nx), np.float32)
        (ox, oy, oz), (bx, by, bz) = self.geometry.origin, self.geometry.size
        phi[oz:oz + bz, oy:oy + by, ox:ox + bx] = phi_b
        return phi.reshape(-1), (dr, dc, dP)

    def forward(self, radius, center, porosity):
        phi, self._parts = self.solid_fraction(radius, center, porosity)
        return self.run.forward(phi)

    def backward(self, grad_output=1.0):
        if self._parts is None:
            raise RuntimeError("backward() needs a forward() first")
        g = self.run.backward(grad_output)
        dr, dc, dP = self._parts
        return {"radius": float((g * dr).sum()),
                "center": np.array([(g * dc[a]).sum() for a in range(3)]),
                "porosity": g * dP, "phi": g}

    @property
    def forward_ms(self):
        return self.run.forward_ms

    @property
    def backward_ms(self):
        return self.run.backward_ms

    def close(self):
        self.run.close()


def _event(stream):
    e = ctypes.c_void_p()
    check(lib().tlbm_event_create(ctypes.byref(e)), "event")
    check(lib().tlbm_event_record(e, stream), "event record")
    return e


def _elapsed(a, b):
    ms = ctypes.c_float()
    check(lib().tlbm_event_elapsed_ms(a, b, ctypes.byref(ms)), "elapsed")
    lib().tlbm_event_destroy(a)
    lib().tlbm_event_destroy(b)
    return float(ms.value)


def mean_drag(problem, radius, center, porosity):
    """torch.autograd.Function front end: J(radius, center, porosity) as an FP64 tensor whose
    backward runs the adjoint (inputs: 0-d, (3,) and the block-shaped porosity tensors)."""
    import torch

    class _MeanDrag(torch.autograd.Function):
        @staticmethod
        def forward(ctx, r, c, P):
            J = problem.forward(float(r.detach().cpu()), c.detach().cpu().numpy(),
                                P.detach().cpu().numpy())
            ctx.meta = (r.dtype, r.device, c.dtype, P.dtype)
            return torch.tensor(J, dtype=torch.float64, device=r.device)

        @staticmethod
        def backward(ctx, gJ):
            rdt, dev, cdt, Pdt = ctx.meta
            g = problem.backward(float(gJ.detach().cpu()))
            return (torch.tensor(g["radius"], dtype=rdt, device=dev),
                    torch.as_tensor(g["center"], dtype=cdt, device=dev),
                    torch.as_tensor(g["porosity"], dtype=Pdt, device=dev))

    return _MeanDrag.apply(radius, center, porosity)
