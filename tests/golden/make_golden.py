"""Writes tests/golden/reference_vectors.npz from the reference's own solver.

Runs in the development container only: it calls oracle/_ref/libref{2,3}d.so,
which oracle/Makefile compiles from the reference sources under
/root/reference (Local_code/level-set_MPI_2D/include, .../level-set_MPI_3D/include)
with -ffp-contract=off.  The fixture holds inputs and the reference's outputs:

  <case>/G, u, v[, w]   seeded inputs (G ghost-filled by the reference BCs)
  <case>/rhs            computeRHS / computeRHS3D
  <case>/rk3            one RK3 step (rk3TimeStep composed from the reference
                        stage, RHS and BC functions), dt = <case>/dt
  <case>/hcr2           HCR-2 (2 iterations) of 1.7*G
  <case>/adv            11 steps of the time loop (HCR-2 every 3, sliver last step)
  <case>/meta           [dim, nx, ny, nz, bcx, bcy, bcz, s_l, dt, t_final]

usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.pyoracle import Reference, make_grid  # noqa: E402
from helpers import bumpy_sdf, rand_vel  # noqa: E402

CASES = [
    ("2d_px_zgy_sl0", 2, (23, 19, 1), (0, 1, 0), 0.0),
    ("2d_px_zgy_sl", 2, (23, 19, 1), (0, 1, 0), 0.3),
    ("2d_zgx_zgy_sl", 2, (21, 17, 1), (1, 1, 0), 0.1),
    ("2d_px_exy_sl", 2, (19, 22, 1), (0, 2, 0), 0.2),
    ("2d_zgx_exy_sl0", 2, (18, 20, 1), (1, 2, 0), 0.0),
    ("3d_periodic_sl0", 3, (11, 9, 8), (0, 0, 0), 0.0),
    ("3d_periodic_sl", 3, (11, 9, 8), (0, 0, 0), 0.25),
]


def main():
    out = {}
    for name, dim, n, bc, s_l in CASES:
        R = Reference(dim)
        g = make_grid(dim, n[0], n[1], n[2], bc=bc, s_l=s_l)
        h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
        seed = sum(map(ord, name))
        G = R.fill_ghosts(g, bumpy_sdf(g.shape, h, seed, r=0.25, noise=0.02))
        vel = rand_vel(g.shape, seed, 0.7, n=dim)
        w = vel[2] if dim == 3 else None
        dt = 0.35 * min(g.dx, g.dy)
        t_final = 10.37 * dt
        out[name + "/G"] = G
        out[name + "/u"] = vel[0]
        out[name + "/v"] = vel[1]
        if dim == 3:
            out[name + "/w"] = w
        out[name + "/rhs"] = R.rhs(g, G, vel[0], vel[1], w)
        out[name + "/rk3"] = R.rk3_step(g, G, vel[0], vel[1], w, dt=dt)
        out[name + "/hcr2"] = R.hcr2(g, R.fill_ghosts(g, 1.7 * G))
        adv, steps, t, dl = R.advance(g, G, vel[0].copy(), vel[1].copy(),
                                      None if w is None else w.copy(), dt=dt, t_final=t_final,
                                      reinit_interval=3)
        assert steps == 11
        out[name + "/adv"] = adv
        out[name + "/meta"] = np.array([dim, n[0], n[1], n[2], bc[0], bc[1], bc[2], s_l, dt, t_final])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
