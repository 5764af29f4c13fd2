"""Regenerate the reference's one-step TRT known-answer case as a fixture.

Follows trt_ref.py:8-25 for the inputs (8x6x4 periodic grid, ``default_rng(1)``,
rho = 1 + 0.01 N, u = 0.05 N, f = feq + 1e-3 N, cast to float32) and trt_ref.py:40-57 for the
expected output: pull streaming by explicit source indices, Pi_ab by einsum, tau from the
Smagorinsky closure with TRT(0.002) (tau0 = 0.506, magic 0.1875, Cs 0.16), and the TRT update
written with the symmetric/antisymmetric split.  That formulation differs from lbm.py's per-q
loop, which is what the oracle restates, so agreement between the two pins the oracle.
The reference accepts |tack - numpy| < 1e-5 (trt_ref.py:60).

Run:  python tests/golden/make_trt_ref.py   (writes tests/golden/trt_ref_8x6x4.npz)
"""
import hashlib
import os

import numpy as np

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]
W = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12

nx, ny, nz = 8, 6, 4
n = nx * ny * nz
c = np.array([CX, CY, CZ]).T
w = np.array(W)
opp = np.array([next(p for p in range(19) if (c[p] == -c[q]).all()) for q in range(19)])


def feq(rho, u):
    cu = 3.0 * (u @ c.T)
    return w * rho[:, None] * (1 + cu + 0.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])


def make():
    rng = np.random.default_rng(1)
    rho0 = 1 + 0.01 * rng.standard_normal(n)
    u0 = 0.05 * rng.standard_normal((n, 3))
    f = feq(rho0, u0) + 1e-3 * rng.standard_normal((n, 19))
    f_in = np.ascontiguousarray(f.T.reshape(-1)).astype(np.float32)

    nu = 0.002
    tau0 = 3.0 * nu + 0.5
    magic, cs = 0.1875, 0.16
    idx = np.arange(n)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    fs = np.empty((n, 19))
    for q in range(19):
        src = (((z - c[q, 2]) % nz) * ny + (y - c[q, 1]) % ny) * nx + (x - c[q, 0]) % nx
        fs[:, q] = f_in.reshape(19, n)[q, src]
    rho = fs.sum(1)
    u = (fs @ c) / rho[:, None]
    eq = feq(rho, u)
    neq = fs - eq
    pi_ab = np.einsum("qa,qb,nq->nab", c, c, neq)
    pi = np.sqrt((pi_ab ** 2).sum(axis=(1, 2)))
    tau = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + 25.456 * cs ** 2 * pi / rho))
    omega_p = (1 / tau)[:, None]
    omega_m = (1 / (magic / (tau - 0.5) + 0.5))[:, None]
    fp, fm = 0.5 * (fs + fs[:, opp]), 0.5 * (fs - fs[:, opp])
    ep, em = 0.5 * (eq + eq[:, opp]), 0.5 * (eq - eq[:, opp])
    ref = fs - omega_p * (fp - ep) - omega_m * (fm - em)
    f_out = np.ascontiguousarray(ref.T.reshape(-1))          # back to the q-major layout
    return f_in, f_out, nu


if __name__ == "__main__":
    f_in, f_out, nu = make()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trt_ref_8x6x4.npz")
    np.savez(out, f_in=f_in, f_out=f_out, dims=np.array([nx, ny, nz]), nu=np.array(nu),
             tol=np.array(1e-5))
    print("f_in sha256:", hashlib.sha256(f_in.tobytes()).hexdigest())
    print("f_out sha256:", hashlib.sha256(f_out.tobytes()).hexdigest())
    print("wrote", out)
