"""CPU oracle for the D3Q19 lattice Boltzmann step: a NumPy restatement of the reference solver.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker (or, in the bench, as the timed CPU
baseline).  The product path (``tack-lbm_amd/``) never imports it; the HIP library is the only
thing that computes a product result.

Every function cites the reference file:line it restates.  Arithmetic runs in the dtype of the
population array handed in: float64 is the oracle proper, float32 is "the reference CPU path"
timed beside the GPU.  Geometry (``solid``) is always evaluated in float32 with one rounding per
multiply and per add, so that cell classification matches a float32 device evaluation bit for
bit (the reference evaluates it in the kernel's float32, ``lbm.py:150-153``).

Layout follows the reference: population ``q`` of cell ``idx = (z*ny + y)*nx + x`` sits at
``f[q*n + idx]`` (``lbm.py:386-388,394,426,456``); arrays here are shaped ``(19, n)``.

What pins each part (checked in ``tests/test_oracle.py``):
  * tables, ``opp``, ``miry``: brute force over the velocity set and the moment identities
    (sum w = 1, sum w c_a c_b = delta_ab / 3);
  * ``feq``: its first three moments in closed form;
  * BGK/LES/TRT ``collide``: ``trt_ref.py``'s golden inputs (regenerated from
    ``default_rng(1)``, committed under ``tests/golden/``), mass and momentum conservation,
    ``feq`` as a fixed point, TRT with equal rates reducing to BGK, Smagorinsky with Cs = 0
    reducing to the molecular tau;
  * streaming: brute-force pull on small periodic grids; specular slip by the mirror identity;
  * bounce-back / moving wall / momentum exchange: hand-computed single-link cases, zero net
    force at rest, mass conservation in a closed slip box;
  * sponge: sigma = 1 gives feq(rho_s, u_s) exactly; solid cells end at w_q;
  * diagnostics / monitors: closed forms from two consecutive states.
"""

from __future__ import annotations

import math

import numpy as np

# --- lattice tables: lbm.py:15-18 (D3Q19 ordering: rest, 6 faces, 12 edges) -------------------
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0], np.int64)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1], np.int64)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1], np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float64)
Q = 19


def opposite():
    """lbm.py:28-30: index of -c_q, by search over the velocity set."""
    return np.array([next(p for p in range(Q) if (CX[p], CY[p], CZ[p]) == (-CX[q], -CY[q], -CZ[q]))
                     for q in range(Q)], np.int64)


def mirror_y():
    """lbm.py:33-35: index of (c_x, -c_y, c_z)."""
    return np.array([next(p for p in range(Q) if (CX[p], CY[p], CZ[p]) == (CX[q], -CY[q], CZ[q]))
                     for q in range(Q)], np.int64)


OPP = opposite()
MIRY = mirror_y()

# D3Q7 for temperature (lbm.py:20-25): rest plus the six faces, c_s^2 = 1/4
CX7 = np.array([0, 1, -1, 0, 0, 0, 0], np.int64)
CY7 = np.array([0, 0, 0, 1, -1, 0, 0], np.int64)
CZ7 = np.array([0, 0, 0, 0, 0, 1, -1], np.int64)
W7 = np.array([1 / 4] + [1 / 8] * 6, np.float64)
OPP7 = np.array([0, 2, 1, 4, 3, 6, 5], np.int64)


def _c(v, dt):
    return dt.type(v)


# --- equilibrium: lbm.py:52-55 -------------------------------------------------------------------
def feq(q, rho, ux, uy, uz):
    """w_q rho (1 + cu + cu^2/2 - 1.5 u.u), cu = 3 c_q.u  (lbm.py:52-55)."""
    dt = np.result_type(rho, ux)
    cu = 3.0 * (float(CX[q]) * ux + float(CY[q]) * uy + float(CZ[q]) * uz)
    w = _c(W[q], dt)
    return w * rho * (1.0 + cu + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy + uz * uz))


# --- collision models: lbm.py:60-139 -------------------------------------------------------------
class BGK:
    """lbm.py:60-69: single relaxation time, omega = 1 / (3 nu + 1/2)."""
    kind = "bgk"

    def __init__(self, nu):
        self.omega = 1.0 / (3.0 * nu + 0.5)

    def collide(self, f, rho, ux, uy, uz):
        dt = f.dtype
        om = _c(self.omega, dt)
        post = np.empty_like(f)
        for q in range(Q):
            post[q] = f[q] - om * (f[q] - feq(q, rho, ux, uy, uz))
        return post, (1.0 / om if om else np.inf)


def smagorinsky_tau(f, rho, ux, uy, uz, tau0, cs):
    """lbm.py:72-93: tau from the non-equilibrium stress Pi_ab = sum_q c_a c_b (f_q - feq_q)."""
    dt = f.dtype
    pxx = np.zeros_like(rho)
    pyy = np.zeros_like(rho)
    pzz = np.zeros_like(rho)
    pxy = np.zeros_like(rho)
    pxz = np.zeros_like(rho)
    pyz = np.zeros_like(rho)
    for q in range(Q):
        neq = f[q] - feq(q, rho, ux, uy, uz)
        ex, ey, ez = float(CX[q]), float(CY[q]), float(CZ[q])
        pxx += ex * ex * neq
        pyy += ey * ey * neq
        pzz += ez * ez * neq
        pxy += ex * ey * neq
        pxz += ex * ez * neq
        pyz += ey * ez * neq
    pi = np.sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz))
    t0 = _c(tau0, dt)
    c = _c(cs, dt)
    return 0.5 * (t0 + np.sqrt(t0 * t0 + _c(25.456, dt) * c * c * pi / rho))


class TRT:
    """lbm.py:96-123: two relaxation times with a Smagorinsky subgrid viscosity.

    ``magic`` (0.1875) and ``cs`` (0.16) are class constants in the reference; here they are
    constructor arguments with those defaults.
    """
    kind = "trt"

    def __init__(self, nu, magic=0.1875, cs=0.16):
        self.tau0 = 3.0 * nu + 0.5
        self.magic = magic
        self.cs = cs

    def collide(self, f, rho, ux, uy, uz):
        dt = f.dtype
        tau = smagorinsky_tau(f, rho, ux, uy, uz, self.tau0, self.cs)
        omega_p = 1.0 / tau
        omega_m = 1.0 / (_c(self.magic, dt) / (tau - 0.5) + 0.5)
        post = np.empty_like(f)
        for q in range(Q):
            o = OPP[q]
            eq = feq(q, rho, ux, uy, uz)
            eo = feq(o, rho, ux, uy, uz)
            post[q] = (f[q] - omega_p * (0.5 * (f[q] + f[o]) - 0.5 * (eq + eo))
                       - omega_m * (0.5 * (f[q] - f[o]) - 0.5 * (eq - eo)))
        return post, tau


class LES:
    """lbm.py:126-139 (SmagorinskyLES): BGK with the Smagorinsky tau."""
    kind = "les"

    def __init__(self, nu, cs=0.16):
        self.tau0 = 3.0 * nu + 0.5
        self.cs = cs

    def collide(self, f, rho, ux, uy, uz):
        tau = smagorinsky_tau(f, rho, ux, uy, uz, self.tau0, self.cs)
        omega = 1.0 / tau
        post = np.empty_like(f)
        for q in range(Q):
            post[q] = f[q] - omega * (f[q] - feq(q, rho, ux, uy, uz))
        return post, tau


COLLISIONS = {"bgk": BGK, "trt": TRT, "les": LES}


# --- obstacles: lbm.py:144-227 -------------------------------------------------------------------
def _f32(v):
    return np.asarray(v, dtype=np.float32)


class Cylinder:
    """lbm.py:144-169: z-axis disk; wall velocity is the unit tangent (-(y-yc)/r, (x-xc)/r, 0)."""
    kind = "cylinder"

    def __init__(self, x, y, r):
        self.x, self.y, self.r = np.float32(x), np.float32(y), np.float32(r)

    def solid(self, x, y, z):
        dx = _f32(x) - self.x
        dy = _f32(y) - self.y
        return (dx * dx + dy * dy < self.r * self.r).astype(np.float32)

    def sink(self, x, y, z):
        """lbm.py:156, 185, 214: the whole obstacle holds the solid temperature."""
        return self.solid(x, y, z)

    def wall_vel(self, x, y, z, dt):
        r = _c(self.r, dt)
        return (-(y - _c(self.y, dt)) / r, (x - _c(self.x, dt)) / r, np.zeros_like(x))


class Sphere:
    """lbm.py:172-198: sphere; resting wall."""
    kind = "sphere"

    def __init__(self, x, y, z, r):
        self.x, self.y, self.z, self.r = (np.float32(v) for v in (x, y, z, r))

    def solid(self, x, y, z):
        dx = _f32(x) - self.x
        dy = _f32(y) - self.y
        dz = _f32(z) - self.z
        return (dx * dx + dy * dy + dz * dz < self.r * self.r).astype(np.float32)

    def sink(self, x, y, z):
        """lbm.py:156, 185, 214: the whole obstacle holds the solid temperature."""
        return self.solid(x, y, z)

    def wall_vel(self, x, y, z, dt):
        zero = np.zeros_like(x)
        return zero, zero, zero


class Box:
    """lbm.py:201-227: axis-aligned box, max-norm signed distance; resting wall."""
    kind = "box"

    def __init__(self, x, y, z, hx, hy, hz):
        self.x, self.y, self.z = (np.float32(v) for v in (x, y, z))
        self.hx, self.hy, self.hz = (np.float32(v) for v in (hx, hy, hz))

    def solid(self, x, y, z):
        d = np.maximum(np.maximum(np.abs(_f32(x) - self.x) - self.hx,
                                  np.abs(_f32(y) - self.y) - self.hy),
                       np.abs(_f32(z) - self.z) - self.hz)
        return (d < np.float32(0.0)).astype(np.float32)

    def sink(self, x, y, z):
        """lbm.py:156, 185, 214: the whole obstacle holds the solid temperature."""
        return self.solid(x, y, z)

    def wall_vel(self, x, y, z, dt):
        zero = np.zeros_like(x)
        return zero, zero, zero


class Field:
    """A per-cell solid fraction phi (float32, flat over idx) used as ``solid`` (no counterpart
    in the reference; lbm.py:403, 416, 428-429 and 449 read with a fractional value give partial
    bounce-back).  Resting wall."""
    kind = "field"

    def __init__(self, phi, dims):
        self.phi = np.asarray(phi, np.float32).reshape(-1)
        self.nx, self.ny, self.nz = dims

    def solid(self, x, y, z):
        return self.phi[(np.asarray(z) * self.ny + np.asarray(y)) * self.nx + np.asarray(x)]

    def sink(self, x, y, z):
        """lbm.py:156, 185, 214: the whole obstacle holds the solid temperature."""
        return self.solid(x, y, z)

    def wall_vel(self, x, y, z, dt):
        zero = np.zeros_like(x)
        return zero, zero, zero


class HeatSink:
    """lbm.py:231-268: a finned heat sink on the floor of a duct.  The base plate and the fins
    (parallel to x, ``fin_thickness`` cells every ``fin_pitch`` along z) are the sink; the duct's
    floor and ceiling rows (y = 0 and y = ny - 1) are adiabatic walls.  Resting walls."""
    kind = "heatsink"

    def __init__(self, ny, x0, length, z0, width, base, fin_height, fin_thickness, fin_pitch):
        self.ny = int(ny)
        self.x0, self.x1 = np.float32(x0), np.float32(x0 + length)
        self.z0, self.z1 = np.float32(z0), np.float32(z0 + width)
        self.y_base = np.float32(1.0 + base)
        self.y_top = np.float32(1.0 + base + fin_height)
        self.fin_thickness = int(fin_thickness)
        self.fin_pitch = int(fin_pitch)

    def sink(self, x, y, z):
        fx, fy, fz = _f32(x), _f32(y), _f32(z)
        inside = (fx >= self.x0) & (fx < self.x1) & (fz >= self.z0) & (fz < self.z1) & (fy >= 1.0)
        base = fy < self.y_base
        fin = (fy < self.y_top) & ((np.asarray(z) - int(self.z0)) % self.fin_pitch
                                    < self.fin_thickness)
        return (inside & (base | fin)).astype(np.float32)

    def solid(self, x, y, z):
        y = np.asarray(y)
        wall = ((y <= 0) | (y >= self.ny - 1)).astype(np.float32)
        return np.maximum(wall, self.sink(x, y, z))

    def wall_vel(self, x, y, z, dt):
        zero = np.zeros_like(x)
        return zero, zero, zero


# --- lateral sides: lbm.py:342-369, plus the moving top lid --------------------------------------
class PeriodicY:
    """lbm.py:342-352: y wraps."""
    kind = "periodic"

    def source(self, y, cy, ny, q):
        return (y - cy + ny) % ny, q, None


class SlipY:
    """lbm.py:355-369: specular free-slip walls halfway outside y = 0 and y = ny - 1."""
    kind = "slip"

    def source(self, y, cy, ny, q):
        yr = y - cy
        out = ((yr < 0) | (yr >= ny)).astype(np.int64)
        return yr + out * (y - yr), q + out * (MIRY[q] - q), None


class LidY:
    """Moving no-slip lid halfway outside y = ny - 1, free-slip floor halfway outside y = 0.

    Not in the reference (it only moves the cylinder wall, lbm.py:159-169, 417-428).  The lid
    uses the same moving-wall bounce-back term as the reference's obstacle links:
    f_q = f*_opp(x) + 6 w_q c_q . u_lid  (the ``wall = 1`` case of lbm.py:428).  Lid links do
    not enter the obstacle's momentum-exchange force.
    """
    kind = "lid"

    def __init__(self, ux=0.04, uy=0.0, uz=0.0):
        self.u = (float(ux), float(uy), float(uz))

    def source(self, y, cy, ny, q):
        yr = y - cy
        low = (yr < 0).astype(np.int64)
        top = yr >= ny
        out = low + top.astype(np.int64)
        return yr + out * (y - yr), q + low * (MIRY[q] - q), top


SIDES = {"periodic": PeriodicY, "slip": SlipY, "lid": LidY}


class Flow:
    """lbm.py:372-380: inflow speed, sponge, cylinder surface speed and lateral side."""

    def __init__(self, u0, sponge_start, sponge_strength, omega_rot=0.0, side=None):
        self.u0 = u0
        self.sponge_start = sponge_start
        self.sponge_strength = sponge_strength
        self.omega_rot = omega_rot
        self.side = PeriodicY() if side is None else side


# --- grid helpers ---------------------------------------------------------------------------------
def coords(nx, ny, nz):
    """lbm.py:386-388 / 400-402: x fastest, then y, then z."""
    idx = np.arange(nx * ny * nz, dtype=np.int64)
    return idx % nx, (idx // nx) % ny, idx // (nx * ny)


_SRC_CACHE = {}


def _sources(nx, ny, nz, side):
    """Per-q pull sources (flat index, source population, lid mask) for lbm.py:410-415, 426."""
    key = (nx, ny, nz, side.kind)
    if key not in _SRC_CACHE:
        x, y, z = coords(nx, ny, nz)
        n = nx * ny * nz
        out = []
        for q in range(Q):
            xs = (x - CX[q] + nx) % nx
            ys, qs, lid = side.source(y, CY[q], ny, q)
            zs = (z - CZ[q] + nz) % nz
            flat = np.asarray(qs, np.int64) * n + (zs * ny + ys) * nx + xs
            out.append((xs, ys, zs, flat, lid))
        _SRC_CACHE.clear()
        _SRC_CACHE[key] = out
    return _SRC_CACHE[key]


# --- initialisation: lbm.py:383-394 ---------------------------------------------------------------
def hash_uniform(seed, idx, comp):
    """Counter-based uniform in [0, 1): splitmix64 of (seed, 3*idx + comp), top 24 bits.

    The build's seeded velocity noise (not in the reference).  Integer arithmetic, so a float32
    device evaluation gives the same (2u - 1) bit for bit.
    """
    with np.errstate(over="ignore"):
        s = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
        k = (np.asarray(idx, np.uint64) * np.uint64(3) + np.uint64(comp) + np.uint64(1))
        z = s + k * np.uint64(0xD1B54A32D192ED03)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def init(dims, geo, flow, seed_amp=0.05, noise_amp=0.0, noise_seed=0, dtype=np.float64):
    """lbm.py:383-394: f = feq(1, u0 fluid, seed_amp u0 fluid sin(2 pi x/nx) cos(2 pi z/nz), 0).

    ``noise_amp`` adds the seeded uniform velocity noise U(-a, a) on fluid cells to all three
    components (build-only; 0 reproduces the reference exactly).
    """
    nx, ny, nz = dims
    dt = np.dtype(dtype)
    x, y, z = coords(nx, ny, nz)
    fluid = (1.0 - geo.solid(x, y, z)).astype(dt)
    u0 = _c(flow.u0, dt)
    two_pi = _c(6.2831853, dt)
    ux = u0 * fluid
    uy = (_c(seed_amp, dt) * u0 * fluid * np.sin(two_pi * x.astype(dt) / _c(nx, dt))
          * np.cos(two_pi * z.astype(dt) / _c(nz, dt)))
    uz = np.zeros_like(ux)
    if noise_amp:
        idx = np.arange(nx * ny * nz, dtype=np.int64)
        a = _c(noise_amp, dt)
        ux = ux + fluid * (a * (2.0 * hash_uniform(noise_seed, idx, 0) - 1.0).astype(dt))
        uy = uy + fluid * (a * (2.0 * hash_uniform(noise_seed, idx, 1) - 1.0).astype(dt))
        uz = uz + fluid * (a * (2.0 * hash_uniform(noise_seed, idx, 2) - 1.0).astype(dt))
    one = np.ones_like(ux)
    return np.stack([feq(q, one, ux, uy, uz) for q in range(Q)]).astype(dt)


FIELDS = ("ux", "uy", "uz", "p", "vmag", "r_rho", "r_ux", "r_uy", "r_uz", "r_ke", "fx", "fy", "fz")


def new_fields(n, dtype=np.float64):
    return {k: np.zeros(n, dtype) for k in FIELDS}


# --- the step: lbm.py:397-456 and 493-513 ---------------------------------------------------------
def step(dims, phys, geo, flow, f_in, diag=None):
    """One fused pull-stream + bounce-back + collide + sponge step (lbm.py:397-456).

    ``f_in`` is ``(19, n)`` in the arithmetic dtype.  ``diag`` (a dict of the 13 fields) makes it
    the report variant (lbm.py:493-513): the fields are updated in place.  Returns ``f_out``.
    """
    return _step_core(dims, phys, geo, flow, f_in, diag)[0]


def _step_core(dims, phys, geo, flow, f_in, diag=None):
    """``step``, also returning the cell quantities the temperature step reads."""
    nx, ny, nz = dims
    n = nx * ny * nz
    dt = f_in.dtype
    F = f_in.reshape(Q * n)
    x, y, z = coords(nx, ny, nz)
    xf, yf, zf = x.astype(dt), y.astype(dt), z.astype(dt)
    here = geo.solid(x, y, z).astype(dt)
    fluid = 1.0 - here
    omega = _c(flow.omega_rot, dt)
    side = flow.side
    lid_u = getattr(side, "u", None)
    f = np.empty((Q, n), dt)
    mx = np.zeros(n, dt)
    my = np.zeros(n, dt)
    mz = np.zeros(n, dt)
    for q, (xs, ys, zs, flat, lid) in enumerate(_sources(nx, ny, nz, side)):
        o = OPP[q]
        cx, cy, cz = float(CX[q]), float(CY[q]), float(CZ[q])
        w6 = 6.0 * _c(W[q], dt)
        wall = geo.solid(xs, ys, zs).astype(dt)
        # wall velocity at the link midpoint (lbm.py:418-425).  It only matters on wall links;
        # elsewhere the reference multiplies it by wall = 0.  It is zeroed off the wall links so
        # that a non-finite wall velocity (r = 0 in trt_ref.py:28's disabled cylinder, which
        # divides by r in lbm.py:161,165) does not turn 0 * inf into NaN.
        with np.errstate(divide="ignore", invalid="ignore"):
            uwx, uwy, uwz = geo.wall_vel(xf - 0.5 * cx, yf - 0.5 * cy, zf - 0.5 * cz, dt)
            c_dot_uw = cx * (omega * uwx) + cy * (omega * uwy) + cz * (omega * uwz)
        c_dot_uw = np.where(wall > 0, c_dot_uw, _c(0.0, dt))
        pulled = F[flat]
        bounced = F[o * n:(o + 1) * n]
        fq = pulled + wall * (bounced - pulled + w6 * c_dot_uw)            # lbm.py:428
        hit = wall * fluid * (2.0 * bounced + w6 * c_dot_uw)                 # lbm.py:429
        if lid is not None:
            c_dot_lid = _c(cx * lid_u[0] + cy * lid_u[1] + cz * lid_u[2], dt)
            fq = np.where(lid, bounced + w6 * c_dot_lid, fq)
            hit = np.where(lid, _c(0.0, dt), hit)
        f[q] = fq
        mx += hit * float(CX[o])
        my += hit * float(CY[o])
        mz += hit * float(CZ[o])
    rho = np.zeros(n, dt)
    jx = np.zeros(n, dt)
    jy = np.zeros(n, dt)
    jz = np.zeros(n, dt)
    for q in range(Q):                                                        # lbm.py:437-441
        rho += f[q]
        jx += f[q] * float(CX[q])
        jy += f[q] * float(CY[q])
        jz += f[q] * float(CZ[q])
    ux = fluid * jx / rho
    uy = fluid * jy / rho
    uz = fluid * jz / rho
    post, tau = phys.collide(f, rho, ux, uy, uz)
    # sponge toward the inflow state (lbm.py:446-456)
    s0 = _c(flow.sponge_start, dt)
    t = np.maximum(_c(0.0, dt), (xf - s0) / (_c(nx, dt) - s0))
    sigma = np.maximum(here, _c(flow.sponge_strength, dt) * t * t)
    u0 = _c(flow.u0, dt)
    rho_s = rho + sigma * (1.0 - rho)
    ux_s = ux + sigma * (u0 * fluid - ux)
    uy_s = (1.0 - sigma) * uy
    uz_s = (1.0 - sigma) * uz
    f_out = np.empty((Q, n), dt)
    for q in range(Q):
        neq = post[q] - feq(q, rho, ux, uy, uz)
        f_out[q] = feq(q, rho_s, ux_s, uy_s, uz_s) + (1.0 - sigma) * neq
    if diag is not None:                                                      # lbm.py:493-513
        p = (rho - 1.0) / 3.0
        ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz)
        ox, oy, oz, op = (diag[k].astype(dt) for k in ("ux", "uy", "uz", "p"))
        ok = 0.5 * (ox * ox + oy * oy + oz * oz)
        diag["r_rho"][:] = (3.0 * (p - op)) * (3.0 * (p - op))
        diag["r_ux"][:] = (ux - ox) * (ux - ox)
        diag["r_uy"][:] = (uy - oy) * (uy - oy)
        diag["r_uz"][:] = (uz - oz) * (uz - oz)
        diag["r_ke"][:] = (ke - ok) * (ke - ok)
        diag["ux"][:] = ux
        diag["uy"][:] = uy
        diag["uz"][:] = uz
        diag["p"][:] = p
        diag["vmag"][:] = np.sqrt(ux * ux + uy * uy + uz * uz)
        diag["fx"][:] = mx
        diag["fy"][:] = my
        diag["fz"][:] = mz
        diag["_rho"] = rho
    aux = dict(here=here, fluid=fluid, tau=tau, sigma=sigma, ux=ux, uy=uy, uz=uz, ux_s=ux_s,
               uy_s=uy_s, uz_s=uz_s)
    return f_out, aux


# --- the temperature variant: lbm.py:283-320, 457-492 ---------------------------------------------
class Heat:
    """lbm.py:283-320 with ``enabled = 1``: D3Q7 advection-diffusion of the temperature with
    diffusivity alpha + max(0, (tau - 1/2)/3 - nu) / pr_t (molecular plus subgrid, turbulent
    Prandtl number 0.85).  ``t_solid``: the temperature of every solid voxel (n)."""
    Q = 7
    pr_t = 0.85

    def __init__(self, alpha, nu, t_solid):
        self.alpha, self.nu = alpha, nu
        self.t_solid = np.asarray(t_solid)

    @staticmethod
    def geq(q, temp, ux, uy, uz):
        """lbm.py:300-303: w_q T (1 + 4 c_q . u)."""
        cu = float(CX7[q]) * ux + float(CY7[q]) * uy + float(CZ7[q]) * uz
        return W7[q] * temp * (1.0 + 4.0 * cu)

    def diffusivity(self, tau):
        """lbm.py:305-308."""
        return self.alpha + np.maximum(0.0, (tau - 0.5) / 3.0 - self.nu) / self.pr_t


def step_heat(dims, phys, geo, flow, f_in, g_in, heat, diag=None, heat_out=None):
    """``step`` with the temperature variant (lbm.py:457-492), FP64.  Returns (f_out, g_out).

    ``g_in``: (7, n).  The temperature populations are pulled periodically in x, y and z; a link
    from a solid voxel bounces back (adiabatic) or, from a sink voxel, anti-bounces to its
    temperature t_solid (lbm.py:468-473).  With ``diag``, ``heat_out`` (a dict with "temp" and
    "q_solid", n each) receives the cell temperature and, for every wetted sink voxel, the sum of
    its fluid neighbours' temperatures (lbm.py:482-492; q_solid is added to, not cleared).
    """
    nx, ny, nz = dims
    n = nx * ny * nz
    f_out, aux = _step_core(dims, phys, geo, flow, f_in, diag)
    g_in = np.asarray(g_in, np.float64).reshape(7, n)
    x, y, z = coords(nx, ny, nz)
    idx = np.arange(n)
    ts = heat.t_solid.astype(np.float64)
    g = np.empty((7, n))
    for q in range(7):
        xs = (x - CX7[q] + nx) % nx
        ys = (y - CY7[q] + ny) % ny
        zs = (z - CZ7[q] + nz) % nz
        nidx = (zs * ny + ys) * nx + xs
        wall = geo.solid(xs, ys, zs).astype(np.float64)
        hot = geo.sink(xs, ys, zs).astype(np.float64)
        pulled = g_in[q, nidx]
        bounced = g_in[OPP7[q], idx]
        at_wall = bounced * (1.0 - 2.0 * hot) + hot * 2.0 * W7[q] * ts[nidx]
        g[q] = pulled + wall * (at_wall - pulled)
    temp = g.sum(axis=0)
    alpha = heat.diffusivity(aux["tau"])
    omega_g = 1.0 / (4.0 * alpha + 0.5)
    sigma = aux["sigma"]
    temp_s = (1.0 - sigma) * temp
    g_out = np.empty((7, n))
    for q in range(7):
        eq = heat.geq(q, temp, aux["ux"], aux["uy"], aux["uz"])
        post = g[q] - omega_g * (g[q] - eq)
        g_out[q] = heat.geq(q, temp_s, aux["ux_s"], aux["uy_s"], aux["uz_s"]) + \
            (1.0 - sigma) * (post - eq)
    if diag is not None and heat_out is not None:
        heat_out["temp"][:] = aux["fluid"] * temp + aux["here"] * ts
        for q in range(1, 7):
            xs = (x - CX7[q] + nx) % nx
            ys = (y - CY7[q] + ny) % ny
            zs = (z - CZ7[q] + nz) % nz
            nidx = (zs * ny + ys) * nx + xs
            hot = geo.sink(xs, ys, zs).astype(np.float64)
            np.add.at(heat_out["q_solid"], nidx, aux["fluid"] * hot * temp)
    return f_out, g_out


def monitors(fields, n, interval, u0, diameter, nz):
    """lbm.py:649-661: FP64 sums of the report fields, residual norms, drag and lift."""
    total = {k: float(np.asarray(fields[k]).sum(dtype=np.float64)) for k in FIELDS}
    res = {name: float(np.sqrt(total[key] / float(n))) / max(interval, 1)
           for name, key in (("continuity", "r_rho"), ("momentum_x", "r_ux"),
                             ("momentum_y", "r_uy"), ("momentum_z", "r_uz"), ("energy", "r_ke"))}
    scale = 0.5 * u0 * u0 * diameter * nz
    res["drag"] = total["fx"] / scale
    res["lift"] = total["fy"] / scale
    return res


def coarsen(dims, s, src, out_dims):
    """lbm.py:516-522: dst[i,j,k] = src[i s, j s, k s]."""
    nx, ny, nz = dims
    mx, my, mz = out_dims
    a = np.asarray(src).reshape(nz, ny, nx)
    return np.ascontiguousarray(a[:mz * s:s, :my * s:s, :mx * s:s]).reshape(-1)


# --- Solver-level setup: lbm.py:551-584, 613-637 -------------------------------------------------
def make_obstacle(kind, nx, ny, nz, diameter, center=None):
    """lbm.py:628-637: centre (0.22 nx, 0.5 ny + 1.7[, 0.5 nz]), radius D/2, box half-sizes
    (r, r, 0.3 nz).  ``center`` overrides the placement (build-only)."""
    r = diameter / 2
    x0, y0, z0 = (0.22 * nx, 0.5 * ny + 1.7, 0.5 * nz) if center is None else center
    if kind == "cylinder":
        return Cylinder(x0, y0, r)
    if kind == "sphere":
        return Sphere(x0, y0, z0, r)
    if kind == "box":
        return Box(x0, y0, z0, r, r, 0.3 * nz)
    if kind == "heatsink":                                        # lbm.py:635-637
        return HeatSink(ny, int(0.2 * nx), int(0.3 * nx), int(0.125 * nz), int(0.75 * nz),
                        round(diameter / 2), int(0.45 * ny), round(diameter / 3), round(diameter))
    raise KeyError(kind)


class OracleSolver:
    """A CPU mirror of ``Solver`` (lbm.py:548-661) driving ``init``/``step``/``monitors``."""

    def __init__(self, nx=384, ny=128, nz=64, collision="les", obstacle="cylinder",
                 reynolds=300.0, u0=0.08, diameter=24.0, lateral="periodic", dtype=np.float64,
                 center=None, sponge_start=None, sponge_strength=1.0, lid_u=(0.04, 0.0, 0.0),
                 cs=None, magic=None, seed_amp=0.05, noise_amp=0.0, noise_seed=0,
                 construct_report=True):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dims = (nx, ny, nz)
        self.n = nx * ny * nz
        self.u0, self.diameter = u0, diameter
        self.dtype = np.dtype(dtype)
        side = LidY(*lid_u) if lateral == "lid" else SIDES[lateral]()
        self.flow = Flow(u0, 0.75 * nx if sponge_start is None else sponge_start,
                         sponge_strength, 0.0, side)
        self.geo = make_obstacle(obstacle, nx, ny, nz, diameter, center)
        self._cs, self._magic = cs, magic
        self.set_physics(collision, reynolds)
        self.fields = new_fields(self.n, self.dtype)
        self.f = init(self.dims, self.geo, self.flow, seed_amp, noise_amp, noise_seed, self.dtype)
        self.step_count = 0
        if construct_report:                                      # lbm.py:582-584
            self.step(report=True)
            self.step_count = 0

    def set_physics(self, collision, reynolds):
        """lbm.py:613-616: nu = u0 D / Re and the named collision model."""
        self.collision_name, self.reynolds = collision, float(reynolds)
        self.nu = self.u0 * self.diameter / self.reynolds
        kw = {}
        if self._cs is not None and collision in ("trt", "les"):
            kw["cs"] = self._cs
        if self._magic is not None and collision == "trt":
            kw["magic"] = self._magic
        self.phys = COLLISIONS[collision](self.nu, **kw)

    def set_actuation(self, omega_rot):
        self.flow.omega_rot = float(omega_rot)

    def step(self, report=False):
        self.f = step(self.dims, self.phys, self.geo, self.flow, self.f,
                      self.fields if report else None)
        self.step_count += 1

    def monitors(self, interval):
        return monitors(self.fields, self.n, interval, self.u0, self.diameter, self.nz)


def strouhal_fft(lift, sample_every, diameter, u0):
    """collision_check.py:27-31: second half of the lift record, rfft arg-max (bin resolution
    D / (u0 * sample_every * N/2))."""
    cl = np.array(lift[len(lift) // 2:], dtype=np.float64)
    cl -= cl.mean()
    spec = np.abs(np.fft.rfft(cl))
    freq = np.fft.rfftfreq(len(cl), d=sample_every)[np.argmax(spec[1:]) + 1]
    return float(freq * diameter / u0)


def strouhal_refined(lift, sample_every, diameter, u0, pad=16):
    """Bin-free Strouhal number of a lift record (all of it; pass the window to use): the peak
    of the Hann-windowed spectrum zero-padded ``pad`` times, refined by a parabola through the
    log magnitudes of the peak bin and its two neighbours.  Not in the reference, whose
    collision_check.py:27-31 estimator quantises St to D / (u0 * sample_every * N/2)."""
    cl = np.asarray(lift, np.float64)
    cl = cl - cl.mean()
    n = len(cl)
    spec = np.abs(np.fft.rfft(cl * np.hanning(n), pad * n))
    k = int(np.argmax(spec[1:-1])) + 1
    a, b, c = np.log(spec[k - 1:k + 2] + 1e-300)
    shift = 0.5 * (a - c) / (a - 2.0 * b + c) if a - 2.0 * b + c != 0 else 0.0
    freq = (k + shift) / (pad * n * sample_every)
    return float(freq * diameter / u0)


def _selfcheck():
    assert abs(W.sum() - 1.0) < 1e-15
    assert math.isclose(float((W * CX * CX).sum()), 1 / 3)


_selfcheck()
