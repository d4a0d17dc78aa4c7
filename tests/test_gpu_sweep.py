"""Seeded random sweeps of RK3 steps against the oracle: exact mode bit for
bit, fast mode to 1e-12 relative in G and in the zero-level-set position.

Each case draws the dimension, the interior shape (3D boxes from 7 to 59
cells a side; 2D grids small enough for the per-cell kernel or tall enough
for the y-march, 2100-6000 rows), the boundary kinds (periodic, zero-gradient,
linear extrapolation in y), S_L (off, scalar, or a per-cell field that dips
below the 1e-10 gate), the stencil noise, the velocity scale and dt.  Together
the cases cover partial tiles, chunk and strip edges, ghost fusion in every
boundary mode and both division paths of the stage kernels."""
import numpy as np
import pytest

from helpers import bumpy_sdf, front_parity, rand_vel, rel_linf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu


def _case(rng):
    dim = int(rng.choice([2, 3]))
    if dim == 3:
        n = tuple(int(x) for x in rng.integers(7, 60, 3))
        bc = tuple(int(x) for x in rng.choice([0, 1], 3))
        if rng.random() < 0.2:
            bc = (bc[0], 2, bc[2])
    else:
        if rng.random() < 0.5:
            n = (int(rng.integers(60, 200)), int(rng.integers(2100, 6000)), 1)
        else:
            n = (int(rng.integers(7, 300)), int(rng.integers(7, 300)), 1)
        bc = (int(rng.choice([0, 1])), int(rng.choice([0, 1, 2])), 0)
    return dim, n, bc, float(rng.choice([0.0, 0.05, 0.3])), bool(rng.random() < 0.3)


def _run_case(ls, oracle, cuda, rng, mode, tag):
    d = lambda a: None if a is None else cuda.from_numpy(np.ascontiguousarray(a)).cuda()
    dim, n, bc, s_l, use_field = _case(rng)
    nx, ny, nz = n
    print("%s: dim %d n %s bc %s s_l %s field %s" % (tag, dim, n, bc, s_l, use_field), flush=True)
    g = make_grid(dim, nx, ny, nz, bc=bc, s_l=s_l)
    p = ls.default_params(dim, nx, nx=nx, ny=ny, nz=nz if dim == 3 else 1, bc=bc, s_l=s_l, mode=mode)
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    cs = int(rng.integers(1 << 30))
    G = bumpy_sdf(g.shape, h, cs, noise=float(rng.choice([0.0, 0.01, 0.1])))
    vel = rand_vel(g.shape, cs, float(rng.choice([0.3, 1.0, 3.0])), n=3 if dim == 3 else 2)
    G = oracle.fill_ghosts(g, G)
    sl = rng.uniform(-0.05, 0.3, g.shape) if use_field else None
    dt = float(rng.uniform(0.05, 0.5)) * min(g.dx, g.dy)
    s = ls.Solver(p)
    out = s.empty()
    w = vel[2] if dim == 3 else None
    s.rk3_step(d(G), out, d(vel[0]), d(vel[1]), d(w), dt=dt, sl=d(sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], w, dt=dt, sl=sl)
    got = out.cpu().numpy()
    s.close()
    return g, h, got, ref


@pytest.mark.parametrize("seed", [201, 202])
def test_random_fast_steps_within_tolerance(ls, oracle, cuda, seed):
    rng = np.random.default_rng(seed)
    for c in range(30):
        g, h, got, ref = _run_case(ls, oracle, cuda, rng, 1, "fast sweep seed %d case %d" % (seed, c))
        assert rel_linf(got, ref, g.interior) <= 1e-12, "seed %d case %d" % (seed, c)
        front_parity(got, ref, h, 1e-12, g.interior)


@pytest.mark.parametrize("seed", [101, 102, 103, 104])
def test_random_exact_steps_bitwise(ls, oracle, cuda, seed):
    rng = np.random.default_rng(seed)
    for c in range(30):
        g, h, got, ref = _run_case(ls, oracle, cuda, rng, 0, "exact sweep seed %d case %d" % (seed, c))
        assert np.array_equal(got, ref), "seed %d case %d" % (seed, c)
