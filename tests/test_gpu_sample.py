"""GPU sample/ scheme (ls_sample_*) against the restatement (oracle/ls_sample.c,
itself bitwise equal to the reference's sample code): every entry point
bit for bit, partial tiles, and the driver loop at the reference's own size
(800 x 800)."""
import math

import numpy as np
import pytest

from oracle.pyoracle import SampleOracle

pytestmark = pytest.mark.gpu

SHAPES = [(12, 10), (33, 17), (70, 9), (130, 66)]


def _inputs(nx, ny, seed):
    rng = np.random.default_rng(seed)
    dx, dy = 0.02 / (nx - 1), 0.02 / (ny - 1)
    O = SampleOracle()
    G0 = O.bunsen(max(nx, ny), dx, dy)[:ny + 6, :nx + 6].copy()
    G = G0 + 2e-4 * rng.standard_normal(G0.shape)
    u = 0.1 * rng.standard_normal(G0.shape)
    v = 0.8 + 0.1 * rng.standard_normal(G0.shape)
    return dx, dy, G, G0, u, v


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("S_L", [0.0, 0.4])
def test_sample_rhs_and_rk3_bitwise(ls, cuda, nx, ny, S_L):
    torch = cuda
    print("sample rhs/rk3", nx, ny, S_L, flush=True)
    dx, dy, G, G0, u, v = _inputs(nx, ny, nx + ny)
    O = SampleOracle()
    s = ls.Sample(nx, ny, dx, dy)
    d = lambda a: torch.from_numpy(a).cuda()
    r = s.rhs(d(G), d(u), d(v), S_L).cpu().numpy()
    assert np.array_equal(r, O.rhs(G, u, v, dx, dy, S_L))
    Gd = d(G)
    for k in range(3):
        s.rk3_step(Gd, d(u), d(v), 2.5e-6, S_L, d(G0))
        G = O.rk3_step(G, u, v, dx, dy, 2.5e-6, S_L, G0)
        assert np.array_equal(Gd.cpu().numpy(), G), k
    s.close()


@pytest.mark.parametrize("S_L", [0.4, 1e20])
def test_sample_rk3_outside_fma_division_class_bitwise(ls, cuda, S_L):
    """Planted G values outside the class for which the sample y-march
    divides with FMA sequences (|G| above 2^50, nonzero |G| below 2^-200),
    and an S_L outside it: the wavefronts concerned take the IEEE divisions
    and the step stays bitwise equal to the oracle."""
    torch = cuda
    print("sample unsafe-class", S_L, flush=True)
    nx, ny = 130, 66
    dx, dy, G, G0, u, v = _inputs(nx, ny, 5)
    rng = np.random.default_rng(6)
    flat = G.reshape(-1)
    idx = rng.choice(flat.size, size=16, replace=False)
    flat[idx[:8]] *= 1e20
    flat[idx[8:]] = 1e-250
    O = SampleOracle()
    s = ls.Sample(nx, ny, dx, dy)
    d = lambda a: torch.from_numpy(a).cuda()
    Gd = d(G)
    s.rk3_step(Gd, d(u), d(v), 2.5e-6, S_L, d(G0))
    ref = O.rk3_step(G, u, v, dx, dy, 2.5e-6, S_L, G0)
    assert np.isfinite(ref).all()
    assert np.array_equal(Gd.cpu().numpy(), ref)
    s.close()


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_sample_boundaries_reinit_first_order_bitwise(ls, cuda, nx, ny):
    torch = cuda
    print("sample bc/reinit/fo", nx, ny, flush=True)
    dx, dy, G, G0, u, v = _inputs(nx, ny, 7 * nx + ny)
    O = SampleOracle()
    s = ls.Sample(nx, ny, dx, dy)
    d = lambda a: torch.from_numpy(a).cuda()
    a = d(G)
    s.apply_periodic(a)
    assert np.array_equal(a.cpu().numpy(), O.apply_periodic(G))
    a = d(G)
    s.apply_anchored_bc(a, d(G0))
    assert np.array_equal(a.cpu().numpy(), O.apply_anchored(G, G0))
    for it in (1, 4):
        a = d(G)
        s.reinitialize(a, d(G0), it)
        assert np.array_equal(a.cpu().numpy(), O.reinitialize(G, G0, dx, dy, it)), it
    a = d(G)
    s.first_order_upwind(a, d(u), d(v), 1e-5, 0.4)
    assert np.array_equal(a.cpu().numpy(), O.first_order_upwind(G, u, v, dx, dy, 1e-5, 0.4))
    s.close()


def _oracle_loop(O, G, G0, nx, ny, dx, dy, steps, freq, dt=2.5e-6, S_L=0.4, u0=0.0, v0=0.8, eu=0.05, ev=0.05):
    """sample/main.cpp:126-148 with the restatement's functions."""
    t = 0.0
    u = np.empty_like(G)
    v = np.empty_like(G)
    for _ in range(steps):
        G = O.apply_periodic(O.apply_anchored(G, G0))
        # libm cos through math.cos, as the reference's std::cos
        uu = u0 + eu * v0 * math.cos(freq * 2 * math.pi * (t))
        u.fill(uu)
        v.fill(v0 + ev * v0 * math.cos(freq * 2 * math.pi * (t)))
        G = O.rk3_step(G, u, v, dx, dy, dt, S_L, G0)
        G = O.apply_anchored(O.apply_periodic(G), G0)
        t += dt
    return G, t, uu


@pytest.mark.parametrize("n,steps", [(96, 60), (800, 4)])
def test_sample_driver_loop_bitwise(ls, cuda, n, steps):
    torch = cuda
    print("sample loop", n, steps, flush=True)
    O = SampleOracle()
    s = ls.Sample(n, n)
    G0d = s.bunsen()
    G0 = O.bunsen(n, s.dx, s.dy)
    assert np.array_equal(G0d.cpu().numpy(), G0)
    G = G0d.clone()
    u, v = s.empty(), s.empty()
    t = s.advance(G, u, v, G0d, steps, freq=10)
    Go, to, u_last = _oracle_loop(O, G0.copy(), G0, n, n, s.dx, s.dy, steps, 10)
    assert t == to
    assert np.array_equal(G.cpu().numpy(), Go)
    assert float(u[5, 5]) == u_last  # velocity of the last step
    s.close()


def test_sample_write_G(ls, cuda, tmp_path):
    s = ls.Sample(20, 14)
    G = s.bunsen().cpu().numpy()
    p = tmp_path / "itr_G_field_0.bin"
    ls.sample_write_G(G, p)
    assert np.array_equal(np.fromfile(p, dtype=np.float64), G[3:-3, 3:-3].ravel())
    s.close()


def test_sample_rejects_small_grids(ls, cuda):
    with pytest.raises(ls.LSError):
        ls.Sample(5, 20)
