"""Matched plugin sets for the HIP step (tlbm) and the CPU oracle, plus seeded inputs.

Shared by the parity tests: one case description builds the same physics on both sides.
"""

import numpy as np

from oracle import lbm_oracle as O

GEOMETRIES = {
    # obstacle kind -> constructor arguments (centre off the lattice points, walls present)
    "cylinder": (9.3, 6.2, 3.1),
    "sphere": (9.3, 6.2, 4.4, 3.1),
    "box": (9.3, 6.2, 4.4, 2.6, 2.1, 2.4),
    # HeatSink(ny, x0, length, z0, width, base, fin_height, fin_thickness, fin_pitch): for a
    # lattice 12 cells high
    "heatsink": (12, 4, 9, 1, 7, 2, 4, 1, 3),
}
# an obstacle that straddles the periodic x boundary (pull sources wrap into the solid)
WRAP_GEOMETRIES = {
    "cylinder": (0.4, 5.3, 2.9),
    "sphere": (0.4, 5.3, 3.6, 2.9),
    "box": (0.4, 5.3, 3.6, 2.2, 1.9, 2.1),
    "heatsink": (12, 0, 30, 0, 20, 1, 3, 2, 4),   # touching the x and z faces
}
LID = (0.04, 0.01, -0.02)


def random_populations(n, seed, noise=1e-3):
    """trt_ref.py:22-24's recipe: feq(1 + 0.01 N, 0.05 N) + noise N, FP64, shape (19, n)."""
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal(n)
    ux, uy, uz = (0.05 * rng.standard_normal(n) for _ in range(3))
    f = np.stack([O.feq(q, rho, ux, uy, uz) for q in range(19)])
    return f + noise * rng.standard_normal((19, n))


def random_fields(n, seed):
    rng = np.random.default_rng(seed)
    out = {k: np.zeros(n, np.float32) for k in O.FIELDS}
    for k in ("ux", "uy", "uz"):
        out[k] = (0.05 * rng.standard_normal(n)).astype(np.float32)
    out["p"] = (0.003 * rng.standard_normal(n)).astype(np.float32)
    return out


def random_fraction(dims, seed=4):
    """A seeded solid-fraction field: a third of the cells fluid (0), a fifth solid (1), the rest
    fractional."""
    rng = np.random.default_rng(seed)
    return np.clip(rng.uniform(-0.5, 1.25, dims[0] * dims[1] * dims[2]), 0, 1).astype(np.float32)


class Case:
    def __init__(self, dims, collision="bgk", obstacle="cylinder", side="periodic",
                 nu=0.02, u0=0.08, sponge=(0.5, 1.0), omega_rot=0.37, geometry=None, cs=0.16,
                 magic=0.1875):
        self.dims = dims
        self.collision, self.obstacle, self.side = collision, obstacle, side
        self.nu, self.u0, self.omega_rot = nu, u0, omega_rot
        self.sponge_start = sponge[0] * dims[0]
        self.sponge_strength = sponge[1]
        if obstacle == "field":
            self.geometry = random_fraction(dims) if geometry is None else geometry
        else:
            self.geometry = GEOMETRIES[obstacle] if geometry is None else geometry
        self.cs, self.magic = cs, magic

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    # --- oracle side ---
    def oracle(self):
        phys = {"bgk": lambda: O.BGK(self.nu), "trt": lambda: O.TRT(self.nu, self.magic, self.cs),
                "les": lambda: O.LES(self.nu, self.cs)}[self.collision]()
        if self.obstacle == "field":
            geo = O.Field(self.geometry, self.dims)
        else:
            geo = {"cylinder": O.Cylinder, "sphere": O.Sphere, "box": O.Box,
                   "heatsink": O.HeatSink}[self.obstacle](*self.geometry)
        side = O.LidY(*LID) if self.side == "lid" else O.SIDES[self.side]()
        flow = O.Flow(self.u0, self.sponge_start, self.sponge_strength, self.omega_rot, side)
        return phys, geo, flow

    # --- HIP side ---
    def device(self):
        import tlbm
        phys = {"bgk": lambda: tlbm.BGK(self.nu),
                "trt": lambda: tlbm.TRT(self.nu, magic=self.magic, cs=self.cs),
                "les": lambda: tlbm.SmagorinskyLES(self.nu, cs=self.cs)}[self.collision]()
        if self.obstacle == "field":
            geo = tlbm.SolidField(tlbm.array(self.geometry))
        else:
            geo = {"cylinder": tlbm.Cylinder, "sphere": tlbm.Sphere, "box": tlbm.Box,
                   "heatsink": tlbm.HeatSink}[self.obstacle](*self.geometry)
        side = tlbm.LidY(*LID) if self.side == "lid" else {"periodic": tlbm.PeriodicY,
                                                          "slip": tlbm.SlipY}[self.side]()
        omega = tlbm.array(np.array([self.omega_rot], np.float32))
        flow = tlbm.Flow(self.u0, self.sponge_start, self.sponge_strength, omega, side)
        return tlbm.D3Q19(*self.dims), phys, geo, flow
