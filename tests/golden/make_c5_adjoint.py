"""FP64 reference values of the adjoint workload (tests/configs.py C5).

Evaluates J = mean x force over the last 1000 of 2000 steps of the smoothed-sphere + porosity
lattice with the PyTorch FP64 restatement (oracle/lbm_torch.py), and writes
tests/golden/c5_adjoint.npz with:
  J             J at the seeded design point (c5_design)
  grad_radius, grad_center, grad_porosity (float32, (32, 32, 32))
                the autograd gradient of the FP64 model (graph recomputed per 25 steps)
  fd_names, fd_values
                central finite differences (J(x + h) - J(x - h)) / 2h, h = 1e-3, for the
                radius, the three centre coordinates and three porosity cells (the ones with the
                largest autograd gradient; their block indices in fd_cells)
  porosity_sha256
                checksum of the seeded porosity block the values belong to
Each FP64 forward run is about 15 k torch operations per 100 steps; use a GPU torch device
(about a minute in all) or expect hours on the CPU.

Run:  python tests/golden/make_c5_adjoint.py --device cuda --out tests/golden/c5_adjoint.npz
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from configs import C5, C5_FD_STEP, c5_design  # noqa: E402
from oracle import lbm_torch as T  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--device", default="cpu")
    ap.add_argument("--out", default=os.path.join(HERE, "c5_adjoint.npz"))
    ap.add_argument("--segment", type=int, default=25)
    args = ap.parse_args()
    dev = torch.device(args.device)
    dims = (C5["nx"], C5["ny"], C5["nz"])
    block = C5["block"]
    origin = tuple(int(round(c - block / 2.0)) for c in C5["nominal_center"])
    nu = C5["u0"] * 2.0 * C5["radius"] / C5["reynolds"]
    omega = 1.0 / (3.0 * nu + 0.5)
    s0 = float(C5["nx"] - 8)
    radius, center, porosity = c5_design()
    f0 = T.equilibrium_state(dims, C5["u0"], device=dev)

    def J(r, c, P, grad=False):
        with torch.set_grad_enabled(grad):
            phi = T.solid_fraction(dims, c, r, P, origin, C5["eps"])
            return T.mean_drag(phi, f0, C5["steps"], C5["window"], omega, C5["u0"], s0, 1.0,
                               segment=args.segment if grad else 0)

    def tensors():
        return (torch.tensor(radius, dtype=torch.float64, device=dev),
                torch.tensor(center, dtype=torch.float64, device=dev),
                torch.tensor(porosity, dtype=torch.float64, device=dev))

    t0 = time.time()
    r, c, P = tensors()
    for v in (r, c, P):
        v.requires_grad_(True)
    J0 = J(r, c, P, grad=True)
    J0.backward()
    g_r, g_c, g_P = (float(r.grad), c.grad.cpu().numpy(), P.grad.cpu().numpy())
    print(f"J = {float(J0.detach()):.12e}  dJ/dr = {g_r:.6e}  dJ/dc = {g_c}  "
          f"({time.time() - t0:.1f} s)", flush=True)

    h = C5_FD_STEP
    cells = [tuple(int(i) for i in np.unravel_index(k, g_P.shape))
             for k in np.argsort(-np.abs(g_P).reshape(-1))[:3]]
    names, values = [], []
    for name in ["radius", "cx", "cy", "cz"] + [f"P{cell}" for cell in cells]:
        vals = []
        for sign in (1.0, -1.0):
            r, c, P = tensors()
            if name == "radius":
                r = r + sign * h
            elif name in ("cx", "cy", "cz"):
                c = c.clone()
                c["xyz".index(name[1])] += sign * h
            else:
                P = P.clone()
                P[cells[len(names) - 4]] += sign * h
            vals.append(float(J(r, c, P)))
        fd = (vals[0] - vals[1]) / (2.0 * h)
        names.append(name)
        values.append(fd)
        print(f"FD {name}: {fd:.9e}  ({time.time() - t0:.1f} s)", flush=True)
    np.savez(args.out, J=float(J0.detach()), grad_radius=g_r, grad_center=g_c,
             grad_porosity=g_P.astype(np.float32), fd_names=np.array(names),
             fd_values=np.array(values), fd_cells=np.array(cells, np.int64), fd_step=h,
             porosity_sha256=hashlib.sha256(np.ascontiguousarray(porosity).tobytes()).hexdigest())
    print("wrote", args.out)


if __name__ == "__main__":
    main()
