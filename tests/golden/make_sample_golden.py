"""Writes tests/golden/sample_vectors.npz from the reference's own sample/ solver.

Runs in the development container only: it calls oracle/_ref/libsample.so,
which oracle/Makefile compiles from the reference sources under
/root/reference/sample with the sample's own flags (-O3 -std=c++17 -fopenmp).
The fixture holds inputs and the reference's outputs:

  <case>/G, G0, u, v     seeded inputs (G0 is the anchoring field)
  <case>/rhs             compute_geqn_rhs
  <case>/rk3             RK3_step (dt, S_L from meta)
  <case>/periodic        Field2D::apply_periodic of G
  <case>/anchored        Field2D::apply_anchored_bc of G
  <case>/reinit          Field2D::reinitialize of G, 3 iterations
  <case>/fo              compute_first_order_upwind
  <case>/meta            [nx, ny, dx, dy, dt, S_L]

usage: python tests/golden/make_sample_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.pyoracle import SampleOracle, SampleReference  # noqa: E402

CASES = [("bunsen_12x10", 12, 10, 0.4, True), ("random_17x21", 17, 21, 0.4, False),
         ("bunsen_24x24_sl0", 24, 24, 0.0, True)]


def inputs(nx, ny, bunsen, seed):
    rng = np.random.default_rng(seed)
    dx, dy = 0.02 / (nx - 1), 0.02 / (ny - 1)
    if bunsen:
        G0 = SampleOracle().bunsen(nx, dx, dy) if nx == ny else None
        if G0 is None:
            G0 = np.zeros((ny + 6, nx + 6))
            O = SampleOracle()
            full = O.bunsen(max(nx, ny), dx, dy)
            G0[:, :] = full[:ny + 6, :nx + 6]
    else:
        G0 = 1e-3 * rng.standard_normal((ny + 6, nx + 6))
    G = G0 + 2e-4 * rng.standard_normal(G0.shape)
    u = 0.1 * rng.standard_normal(G0.shape)
    v = 0.8 + 0.1 * rng.standard_normal(G0.shape)
    return dx, dy, G, G0, u, v


def main():
    R = SampleReference()
    out = {}
    for k, (name, nx, ny, S_L, bunsen) in enumerate(CASES):
        dx, dy, G, G0, u, v = inputs(nx, ny, bunsen, 100 + k)
        dt = 2.5e-6
        out[name + "/G"], out[name + "/G0"], out[name + "/u"], out[name + "/v"] = G, G0, u, v
        out[name + "/rhs"] = R.rhs(G, u, v, dx, dy, S_L)
        out[name + "/rk3"] = R.rk3_step(G, u, v, dx, dy, dt, S_L, G0)
        out[name + "/periodic"] = R.apply_periodic(G)
        out[name + "/anchored"] = R.apply_anchored(G, G0)
        out[name + "/reinit"] = R.reinitialize(G, G0, dx, dy, 3)
        out[name + "/fo"] = R.first_order_upwind(G, u, v, dx, dy, dt, S_L)
        out[name + "/meta"] = np.array([nx, ny, dx, dy, dt, S_L])
    path = os.path.join(ROOT, "tests", "golden", "sample_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
