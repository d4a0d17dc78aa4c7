"""GPU parity of the fast mode.

The fast mode evaluates the same scheme with reorganised arithmetic (product-
form WENO weights, one refined reciprocal per cell, FMA), so it is held to a
tolerance rather than bit equality: 1e-12 relative L-infinity in G
over the interior and 1e-12 (relative to the box length) in the interpolated
zero-level-set position on every grid edge.  Reference values: the CPU oracle
at sizes it finishes in seconds, and the exact mode (bitwise equal to the
oracle) at the full 256^3 and 512^3 sizes.
"""
import numpy as np
import pytest

from helpers import bumpy_sdf, front_parity, rand_vel, rel_linf
from oracle.pyoracle import Oracle, make_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(cuda, a):
    return cuda.from_numpy(np.ascontiguousarray(a)).cuda()


def _h(p):
    return (p.dz, p.dy, p.dx) if p.dim == 3 else (p.dy, p.dx)


CASES = [
    (2, (23, 19, 1), (0, 1, 0)),
    (2, (23, 19, 1), (1, 1, 0)),
    (2, (17, 29, 1), (1, 2, 0)),
    (3, (21, 17, 13), (0, 0, 0)),
    (3, (19, 11, 15), (0, 2, 1)),
    (3, (70, 9, 40), (1, 1, 1)),    # partial x and y tiles, z chunking, zero-gradient images
    (3, (130, 6, 5), (0, 0, 0)),    # three x tiles, ny and nz below the stencil width
    (2, (130, 5500, 1), (0, 1, 0)),  # tall enough for the 2D y-march: partial strip, chunk edges
    (2, (130, 5500, 1), (1, 2, 0)),  # y-march with a separate ghost fill (linear extrapolation)
]


@pytest.mark.parametrize("dim,n,bc", CASES)
@pytest.mark.parametrize("s_l", [0.0, 0.3])
def test_fast_step_and_hcr2(ls, oracle, cuda, dim, n, bc, s_l):
    g = make_grid(dim, *n, bc=bc, s_l=s_l)
    p = ls.default_params(dim, n[0], nx=n[0], ny=n[1], nz=n[2] if dim == 3 else 1, bc=bc, s_l=s_l,
                          mode=ls.MODE_FAST)
    h = _h(p)
    G = oracle.fill_ghosts(g, bumpy_sdf(g.shape, h, 17))
    vel = rand_vel(g.shape, 17, 1.0, n=dim)
    w = vel[2] if dim == 3 else None
    s = ls.Solver(p)
    out = s.empty()
    dt = 0.3 * min(g.dx, g.dy)
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]),
               None if w is None else _dev(cuda, w), dt=dt)
    ref = oracle.rk3_step(g, G, vel[0], vel[1], w, dt=dt)
    I = g.interior
    assert rel_linf(out.cpu().numpy(), ref, I) <= TOL
    # ghost cells written inside the stage kernels equal a separate ghost fill
    refilled = out.clone()
    s.fill_ghosts(refilled)
    assert cuda.equal(refilled, out)
    H = oracle.fill_ghosts(g, 1.7 * G)
    dH = _dev(cuda, H)
    s.reinit(dH)
    refH = oracle.hcr2(g, H)
    assert rel_linf(dH.cpu().numpy(), refH, I) <= TOL
    front_parity(dH.cpu().numpy(), refH, h, TOL, I)


@pytest.mark.parametrize("n", [(23, 19, 1), (130, 5500, 1)])  # one thread per cell / y-march
def test_fast_sl_field_2d(ls, oracle, cuda, n):
    """A per-cell S_L field (values below the 1e-10 gate included)."""
    bc = (0, 1, 0)
    g = make_grid(2, *n, bc=bc, s_l=0.0)
    p = ls.default_params(2, n[0], nx=n[0], ny=n[1], bc=bc, s_l=0.0, mode=ls.MODE_FAST)
    G = oracle.fill_ghosts(g, bumpy_sdf(g.shape, _h(p), 23))
    vel = rand_vel(g.shape, 23, 1.0, n=2)
    sl = np.random.default_rng(6).uniform(-0.05, 0.3, g.shape)
    s = ls.Solver(p)
    out = s.empty()
    dt = 0.3 * min(g.dx, g.dy)
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), None, dt=dt, sl=_dev(cuda, sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], None, dt=dt, sl=sl)
    assert rel_linf(out.cpu().numpy(), ref, g.interior) <= TOL


def test_fast_circle2d_256_full(ls, oracle, cuda):
    p = ls.default_params(2, 256, s_l=0.1, reinit_interval=10, bc=(1, 1, 0), mode=ls.MODE_FAST)
    G, u, v, _ = ls.init_case(p, ls.CASE_CIRCLE2D)
    s = ls.Solver(p)
    s.fill_ghosts(G)
    Gh, uh, vh = G.cpu().numpy(), u.cpu().numpy(), v.cpu().numpy()
    dt = ls.cfl_dt(p, 0.5, 0.5, 0.0, 0.0)
    s.advance(G, u, v, dt=dt, t_final=1e30, max_steps=200)
    g = make_grid(2, 256, 256, bc=(1, 1, 0), s_l=0.1)
    ref, _, _, _ = oracle.advance(g, Gh, uh, vh, dt=dt, t_final=1e30, max_steps=200, reinit_interval=10)
    got = G.cpu().numpy()
    assert rel_linf(got, ref, g.interior) <= TOL
    front_parity(got, ref, _h(p), TOL, g.interior)


@pytest.mark.parametrize("case,n,steps", [("sphere_swirl", 64, 100), ("plane_modes", 48, 40)])
def test_fast_3d_workloads_reduced(ls, oracle, cuda, case, n, steps):
    s_l = 0.05 if case == "sphere_swirl" else 0.08
    p = ls.default_params(3, n, s_l=s_l, reinit_interval=5, mode=ls.MODE_FAST)
    cid = ls.CASE_SPHERE_SWIRL3D if case == "sphere_swirl" else ls.CASE_PLANE_MODES3D
    G, u, v, w = ls.init_case(p, cid, seed=3)
    s = ls.Solver(p)
    s.fill_ghosts(G)
    host = [a.cpu().numpy() for a in (G, u, v, w)]
    dt = 0.5 * p.dx / 1.6
    s.advance(G, u, v, w, dt=dt, t_final=1e30, max_steps=steps)
    g = make_grid(3, n, n, n, bc=(0, 0, 0), s_l=s_l)
    ref, _, _, _ = oracle.advance(g, *host, dt=dt, t_final=1e30, max_steps=steps, reinit_interval=5)
    got = G.cpu().numpy()
    assert rel_linf(got, ref, g.interior) <= TOL
    front_parity(got, ref, _h(p), TOL, g.interior)


@pytest.mark.parametrize("case,n,steps", [("sphere_swirl", 256, 60), ("plane_modes", 512, 100)])
def test_fast_full_size_every_step(ls, cuda, case, n, steps):
    """The full-size 3D workloads (sphere in swirl 256^3, plane with modes 512^3):
    fast mode against the exact mode (bitwise the oracle; the oracle itself is
    compared at these sizes in test_gpu_workloads.py) after EVERY step of the
    horizon, HCR-2 after every 5th step, steps counted from 1 as the benchmark
    counts them (its timed region covers steps 4..23 or 6..25).  Both 1e-12
    bars hold at every step: relative L-infinity in G and the interpolated
    front position on every grid edge.  The gap is not monotonic in time:
    discrete events near zero (an upwind choice at u_eff ~ 0, an HCR-2
    stability test) open local differences that the scheme damps again; the
    largest over the horizon is printed."""
    from helpers import front_parity_torch
    s_l = 0.05 if case == "sphere_swirl" else 0.08
    cid = ls.CASE_SPHERE_SWIRL3D if case == "sphere_swirl" else ls.CASE_PLANE_MODES3D
    run = []
    for mode in (ls.MODE_EXACT, ls.MODE_FAST):
        p = ls.default_params(3, n, s_l=s_l, reinit_interval=5, mode=mode)
        G, u, v, w = ls.init_case(p, cid, seed=3)
        s = ls.Solver(p)
        s.fill_ghosts(G)
        run.append((s, G, u, v, w))
    speed = (run[0][2].abs() + run[0][3].abs() + run[0][4].abs()).max().item()
    dt = 0.5 * p.dx / (speed + s_l)
    h = (p.dz, p.dy, p.dx)
    I = (slice(3, 3 + n),) * 3
    worst_g, worst_f, at = 0.0, 0.0, 0
    for k in range(1, steps + 1):
        for s, G, u, v, w in run:
            done, _, _ = s.advance(G, u, v, w, dt=dt, t_final=1e30, step0=k, max_steps=1)
            assert done == 1
        ge, gf = run[0][1][I], run[1][1][I]
        err = float((gf - ge).abs().max()) / float(ge.abs().max())
        fr, _ = front_parity_torch(gf, ge, h, TOL)
        if err > worst_g:
            worst_g, at = err, k
        worst_f = max(worst_f, fr)
        assert err <= TOL, "step %d: fast vs exact %.3e" % (k, err)
    print("fast vs exact %s %d^3, steps 1..%d: max rel L-inf in G %.3e (step %d), max front offset %.3e L"
          % (case, n, steps, worst_g, at, worst_f))
    for s, *_ in run:
        s.close()
    del run
    cuda.cuda.empty_cache()


def test_fast_log_known_answers(ls, cuda):
    """The fast mode reproduces the reference's logged metrics too."""
    N, T = 201, 6.283185
    g = make_grid(2, N, N, bc=(0, 1, 0))
    G0, u, v = Oracle(fma=True).ic_zalesak2d(g, T)
    p = ls.default_params(2, N, reinit_interval=0, mode=ls.MODE_FAST)
    s = ls.Solver(p)
    dG, dG0 = _dev(cuda, G0), _dev(cuda, G0)
    s.fill_ghosts(dG)
    s.fill_ghosts(dG0)
    steps, _, _ = s.advance(dG, _dev(cuda, u), _dev(cuda, v), dt=ls.cfl_dt(p, 0.2, 1.0), t_final=T)
    l2, area = s.diag(dG, dG0)
    assert steps == 6284 and "%.6e" % l2 == "8.576633e-02" and "%.6f" % area == "0.058800"
    g3 = make_grid(3, 64, 64, 64, bc=(0, 0, 0))
    o = Oracle(fma=True)
    G0 = o.ic_sphere3d(g3)
    uu, vv, ww = o.deformation_velocity(g3, 0.0, 1.5)
    p3 = ls.default_params(3, 64, reinit_interval=0, velocity=ls.VEL_DEFORM3D, vel_period=1.5,
                           mode=ls.MODE_FAST)
    s3 = ls.Solver(p3)
    dG, dG0 = _dev(cuda, G0), _dev(cuda, G0)
    s3.fill_ghosts(dG)
    s3.fill_ghosts(dG0)
    steps, _, _ = s3.advance(dG, _dev(cuda, uu), _dev(cuda, vv), _dev(cuda, ww), dt=0.001, t_final=1.5)
    l2, vol = s3.diag(dG, dG0)
    assert steps == 1501 and "%.6e" % l2 == "9.158634e-04" and "%.6f" % vol == "0.014245"


@pytest.mark.parametrize("n,bc", [((70, 9, 40), (1, 1, 1)),    # partial tiles, zero-gradient images
                                  ((66, 13, 37), (0, 0, 0)),   # periodic, odd chunk lengths
                                  ((64, 8, 70), (0, 1, 1)),    # mixed axes, z chunking
                                  ((130, 6, 5), (1, 0, 0))])   # three x tiles, ny and nz below a tile
@pytest.mark.parametrize("iters", [2, 3])
def test_fast_hcr2_pair_bitwise(ls, oracle, cuda, n, bc, iters):
    """The first two HCR-2 iterations run as one pass over phi_0.  That pass
    must give bit for bit the two single iterations with a ghost fill between
    them (ls_hcr2_step), over the whole padded array, ghost cells included."""
    g = make_grid(3, *n, bc=bc, s_l=0.0)
    p = ls.default_params(3, n[0], nx=n[0], ny=n[1], nz=n[2], bc=bc, mode=ls.MODE_FAST, reinit_iters=iters)
    G = oracle.fill_ghosts(g, 1.7 * bumpy_sdf(g.shape, (p.dz, p.dy, p.dx), 23, noise=0.05))
    s = ls.Solver(p)
    fused = _dev(cuda, G)
    s.reinit(fused)
    phi0 = _dev(cuda, G)
    cur = phi0
    for _ in range(iters):
        nxt = s.empty()
        s.hcr2_step(cur, phi0, nxt)
        cur = nxt
    assert not cuda.equal(cur, phi0)
    assert cuda.equal(fused, cur)
