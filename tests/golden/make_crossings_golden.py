"""Writes tests/golden/crossings_vectors.npz from the reference's own
computeInterfaceCrossings / computeInterfaceCrossings3D.

Runs in the development container only: it calls oracle/_ref/libref{2,3}d.so,
which oracle/Makefile compiles from the reference sources under
/root/reference (Local_code/level-set_MPI_2D/include/reinitialization.h:75-130,
Local_code/level-set_MPI_3D/include/reinitialization.h:46-89).  The fixture
holds inputs and the reference's outputs:

  <case>/p0     phi_0 (padded, ghosts filled), with planted cells below
  <case>/rt     r_tilde on interior cells (ghost entries 0)
  <case>/flag   crossing bit mask on interior cells (ghost entries 0)
  <case>/meta   [dim, nx, ny, nz, bcx, bcy, bcz]

Planted cells (build_p0): a crossing-neighbour sum of -9e-16 (below the
1e-15 gate: r~ = 0), one of -1.1e-15 (above it), a cell that is exactly 0 and
one that is -0.0 (a zero product is not a crossing), a neighbour that is
exactly 0, and a cell with every neighbour across the front.

usage: python tests/golden/make_crossings_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.pyoracle import Oracle, Reference, make_grid  # noqa: E402
from helpers import bumpy_sdf  # noqa: E402

CASES = [
    ("2d_px_zgy", 2, (23, 19, 1), (0, 1, 0)),
    ("2d_zgx_zgy", 2, (21, 17, 1), (1, 1, 0)),
    ("2d_px_exy", 2, (19, 22, 1), (0, 2, 0)),
    ("2d_zgx_exy", 2, (18, 20, 1), (1, 2, 0)),
    ("3d_periodic", 3, (11, 9, 8), (0, 0, 0)),
    ("3d_zg", 3, (10, 12, 9), (1, 1, 1)),
    ("3d_mixed", 3, (9, 10, 11), (0, 1, 1)),
]


def build_p0(dim, n, bc, seed):
    """Seeded bumpy SDF with planted cells, ghosts filled by the oracle's
    ghost fill (bitwise the reference's, pinned by test_oracle)."""
    g = make_grid(dim, n[0], n[1], n[2], bc=bc)
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    o = Oracle()
    p0 = bumpy_sdf(g.shape, h, seed, r=0.25, noise=0.02)
    # neighbour values below are listed per array axis (z, y, x in 3D; y, x
    # in 2D), low side first
    c = tuple(s // 2 for s in g.shape)          # interior centre
    off = [np.eye(dim, dtype=int)[a] for a in range(dim)]

    def at(base, *d):
        return tuple(int(b + x) for b, x in zip(base, d))

    def plant(base, centre, nbrs):
        p0[base] = centre
        for a in range(dim):
            p0[at(base, *(-off[a]))] = nbrs[2 * a]
            p0[at(base, *off[a])] = nbrs[2 * a + 1]
    far = tuple(3 + 2 for _ in range(dim))       # a cell near the low corner
    # planted stencils are 3 cells apart along an axis, so none shares a cell
    nb = 2 * dim
    # sum of crossing neighbours -9e-16 (|sum| <= 1e-15): r~ = 0, flags set
    plant(c, 0.3, [-4e-16, 0.1, 0.2, -5e-16] + [0.3, 0.4][:nb - 4])
    # sum -1.1e-15: r~ = 0.3 / -1.1e-15
    c2 = at(c, *([0] * (dim - 1) + [-3]))
    plant(c2, 0.3, [-6e-16, 0.1, 0.2, -5e-16] + [0.3, 0.4][:nb - 4])
    # centre exactly 0 and -0.0, a neighbour exactly 0, all neighbours across
    c3 = at(c, *([0] * (dim - 1) + [3]))
    plant(c3, 0.0, [-0.1, 0.1, 0.2, -0.2] + [0.3, -0.4][:nb - 4])
    c4 = at(far, *([0] * dim))
    plant(c4, -0.0, [0.1, 0.1, -0.2, 0.2] + [0.3, 0.4][:nb - 4])
    c5 = at(c, *([0] * (dim - 2) + [3, 0]))
    plant(c5, -0.05, [0.0, 0.1, 0.2, 0.25] + [0.3, 0.35][:nb - 4])
    return g, o.fill_ghosts(g, p0)


def main():
    out = {}
    for name, dim, n, bc in CASES:
        R = Reference(dim)
        g, p0 = build_p0(dim, n, bc, sum(map(ord, name)))
        if dim == 2 or bc == (0, 0, 0):
            assert np.array_equal(R.fill_ghosts(g, p0), p0)
        rt, fl = R.crossings(g, p0)
        assert (fl != 0).any() and (rt == 0).any()
        out[name + "/p0"] = p0
        out[name + "/rt"] = rt
        out[name + "/flag"] = fl
        out[name + "/meta"] = np.array([dim, n[0], n[1], n[2], bc[0], bc[1], bc[2]])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crossings_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
