"""The launch variants of the 3D z-march stage kernel on grids whose tiles are
all full (nx a multiple of 64, ny of 8: every tile of both modes is full), so
the variants without bounds tests run: scalar S_L and a per-cell S_L field,
stage 1 with G_s == G_n and with distinct arrays, stages 2 and 3 with the same
array passed as G_s and G_n.  Exact mode: bit for bit the oracle.  Fast mode:
1e-12 relative in G and in zero-level-set position against the oracle."""
import numpy as np
import pytest

from helpers import bumpy_sdf, front_parity, rand_vel, rel_linf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12
FULL = [((128, 16, 12), (0, 0, 0)), ((64, 24, 9), (1, 1, 1)), ((192, 8, 20), (0, 1, 0))]


def _dev(cuda, a):
    return None if a is None else cuda.from_numpy(np.ascontiguousarray(a)).cuda()


def _setup(ls, oracle, n, bc, s_l, mode, seed):
    g = make_grid(3, *n, bc=bc, s_l=s_l)
    p = ls.default_params(3, n[0], nx=n[0], ny=n[1], nz=n[2], bc=bc, s_l=s_l, mode=mode)
    assert p.nx % 64 == 0 and p.ny % 8 == 0
    h = (p.dz, p.dy, p.dx)
    G = oracle.fill_ghosts(g, bumpy_sdf(g.shape, h, seed))
    vel = rand_vel(g.shape, seed, 1.0, n=3)
    return g, p, h, G, vel


def _check(mode, got, ref, g, h):
    if mode == 0:
        assert np.array_equal(got, ref), "rel Linf %.3e" % rel_linf(got, ref)
    else:
        assert rel_linf(got, ref, g.interior) <= TOL
        front_parity(got, ref, h, TOL, g.interior)


@pytest.mark.parametrize("n,bc", FULL)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("field", [False, True])
def test_full_tiles_step(ls, oracle, cuda, n, bc, mode, field):
    """rk3_step (stage 1 with G_s == G_n, stages 2 and 3 with distinct arrays)
    with a scalar S_L or a per-cell S_L field (values below the 1e-10 gate
    included)."""
    g, p, h, G, vel = _setup(ls, oracle, n, bc, 0.0 if field else 0.15, mode, 31)
    sl = np.random.default_rng(3).uniform(-0.05, 0.2, g.shape) if field else None
    s = ls.Solver(p)
    out = s.empty()
    dt = 0.3 * g.dx
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, vel[2]), dt=dt,
               sl=_dev(cuda, sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], vel[2], dt=dt, sl=sl)
    _check(mode, out.cpu().numpy(), ref, g, h)
    s.close()


@pytest.mark.parametrize("n,bc", FULL[:2])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("same", [False, True])
def test_full_tiles_stage(ls, oracle, cuda, n, bc, mode, k, same):
    """ls_rk3_stage k with G_s and G_n distinct, or the same array passed as
    both (stage 1 then skips the G_n load; stages 2 and 3 load it anyway)."""
    g, p, h, Gs, vel = _setup(ls, oracle, n, bc, 0.2, mode, 41)
    Gn = Gs if same else oracle.fill_ghosts(g, bumpy_sdf(g.shape, h, 42))
    dt = 1.5e-3
    s = ls.Solver(p)
    out = s.empty()
    dGs = _dev(cuda, Gs)
    dGn = dGs if same else _dev(cuda, Gn)
    s.rk3_stage(k, dGs, dGn, out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, vel[2]), dt=dt)
    r = oracle.rhs(g, Gs, vel[0], vel[1], vel[2])
    if k == 1:
        e = Gn + dt * r
    elif k == 2:
        e = 0.75 * Gn + 0.25 * Gs + 0.25 * dt * r
    else:
        e = (1.0 / 3.0) * Gn + (2.0 / 3.0) * Gs + (2.0 / 3.0) * dt * r
    ref = Gn.copy()
    ref[g.interior] = e[g.interior]
    ref = oracle.fill_ghosts(g, ref)
    _check(mode, out.cpu().numpy(), ref, g, h)
    s.close()
