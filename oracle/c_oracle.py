"""ctypes front end of the C oracle (oracle/lbm_oracle.c), FP64, OpenMP.

TEST INFRASTRUCTURE: used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
checker for runs too long for the NumPy oracle.  Same interface as ``lbm_oracle.step`` /
``lbm_oracle.init``, taking the NumPy oracle's plugin objects.  Build: ``make -C oracle``.
"""

import ctypes
import os

import numpy as np

from . import lbm_oracle as O

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblbm_oracle.so")


class Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("collision", ctypes.c_int32), ("obstacle", ctypes.c_int32),
                ("side", ctypes.c_int32), ("omega", ctypes.c_double), ("tau0", ctypes.c_double),
                ("cs", ctypes.c_double), ("magic", ctypes.c_double),
                ("obs", ctypes.c_double * 6), ("u0", ctypes.c_double),
                ("sponge_start", ctypes.c_double), ("sponge_strength", ctypes.c_double),
                ("lid", ctypes.c_double * 3), ("omega_rot", ctypes.c_double),
                ("phi", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built (make -C oracle)")
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.oracle_step.argtypes = [ctypes.POINTER(Params), dp, dp, dp, dp]
        _lib.oracle_init.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_uint64]
    return _lib


def available():
    return os.path.exists(_LIB_PATH)


def params(dims, phys, geo, flow):
    p = Params()
    p.nx, p.ny, p.nz = dims
    p.collision = {"bgk": 0, "trt": 1, "les": 2}[phys.kind] if phys is not None else 0
    p.obstacle = {"cylinder": 0, "sphere": 1, "box": 2, "field": 3}[geo.kind]
    p.side = {"periodic": 0, "slip": 1, "lid": 2}[flow.side.kind]
    if phys is not None:
        p.omega = getattr(phys, "omega", 0.0)
        p.tau0 = getattr(phys, "tau0", 0.0)
        p.cs = getattr(phys, "cs", 0.0)
        p.magic = getattr(phys, "magic", 0.0)
    vals = {"cylinder": ("x", "y", "r"), "sphere": ("x", "y", "z", "r"),
            "box": ("x", "y", "z", "hx", "hy", "hz"), "field": ()}[geo.kind]
    for i, name in enumerate(vals):
        p.obs[i] = float(getattr(geo, name))
    if geo.kind == "field":
        p.phi = geo.phi.ctypes.data  # kept alive by geo
    p.u0 = flow.u0
    p.sponge_start = flow.sponge_start
    p.sponge_strength = flow.sponge_strength
    lid = getattr(flow.side, "u", (0.0, 0.0, 0.0))
    for i in range(3):
        p.lid[i] = lid[i]
    p.omega_rot = float(flow.omega_rot)
    return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def step(dims, phys, geo, flow, f_in, diag=None):
    """As ``lbm_oracle.step`` (FP64): returns f_out; ``diag`` (13 fields) is updated in place."""
    n = dims[0] * dims[1] * dims[2]
    f_in = np.ascontiguousarray(f_in, np.float64).reshape(19, n)
    f_out = np.empty_like(f_in)
    p = params(dims, phys, geo, flow)
    fields = None
    rho = None
    if diag is not None:
        fields = np.stack([np.asarray(diag[k], np.float64) for k in O.FIELDS])
        rho = np.empty(n)
    lib().oracle_step(ctypes.byref(p), _dp(f_in), _dp(f_out),
                      _dp(fields) if fields is not None else None,
                      _dp(rho) if rho is not None else None)
    if diag is not None:
        for i, k in enumerate(O.FIELDS):
            diag[k][:] = fields[i]
        diag["_rho"] = rho
    return f_out


def init(dims, geo, flow, seed_amp=0.05, noise_amp=0.0, noise_seed=0):
    n = dims[0] * dims[1] * dims[2]
    f = np.empty((19, n))
    p = params(dims, None, geo, flow)
    lib().oracle_init(ctypes.byref(p), _dp(f), seed_amp, noise_amp, noise_seed)
    return f


class CSolver(O.OracleSolver):
    """``OracleSolver`` (the CPU mirror of lbm.py's Solver) on the C oracle, FP64."""

    def __init__(self, *args, **kw):
        kw["dtype"] = np.float64
        construct_report = kw.pop("construct_report", True)
        seed_amp = kw.get("seed_amp", 0.05)
        noise_amp = kw.get("noise_amp", 0.0)
        noise_seed = kw.get("noise_seed", 0)
        super().__init__(*args, construct_report=False, **kw)
        self.f = init(self.dims, self.geo, self.flow, seed_amp, noise_amp, noise_seed)
        if construct_report:
            self.step(report=True)
            self.step_count = 0

    def step(self, report=False):
        self.f = step(self.dims, self.phys, self.geo, self.flow, self.f,
                      self.fields if report else None)
        self.step_count += 1
