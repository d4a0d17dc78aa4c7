"""GPU parity of the exact mode against the CPU oracle (through the C ABI).

The exact mode keeps the reference's operation order with no FMA contraction;
the oracle is the contraction-free build of the C restatement, itself proven
bitwise equal to the reference's compiled sources (tests/test_oracle.py).  The
bar here is therefore bit-for-bit equality of every padded value.
"""
import numpy as np
import pytest

from helpers import bumpy_sdf, rand_vel, rel_linf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu


def _grid_and_params(ls, dim, n, bc, s_l, mode=0, ri=0):
    nx, ny, nz = n
    g = make_grid(dim, nx, ny, nz, bc=bc, s_l=s_l)
    p = ls.default_params(dim, nx, nx=nx, ny=ny, nz=nz if dim == 3 else 1, bc=bc, s_l=s_l,
                          mode=mode, reinit_interval=ri)
    assert tuple(p.shape) == tuple(g.shape)
    assert (p.dx, p.dy, p.reinit_dtau) == (g.dx, g.dy, g.reinit_dtau)
    return g, p


def _inputs(g, seed, vel_scale=1.0):
    h = (g.dz, g.dy, g.dx) if g.dim == 3 else (g.dy, g.dx)
    G = bumpy_sdf(g.shape, h, seed)
    vel = rand_vel(g.shape, seed, vel_scale, n=3 if g.dim == 3 else 2)
    return G, vel


def _dev(cuda, a):
    return None if a is None else cuda.from_numpy(np.ascontiguousarray(a)).cuda()


CASES = [
    (2, (23, 19, 1), (0, 1, 0)),   # reference 2D default: periodic x, zero-gradient y
    (2, (23, 19, 1), (1, 1, 0)),   # zero-gradient all sides (the 2D circle workload)
    (2, (23, 19, 1), (0, 2, 0)),   # linear extrapolation in y
    (2, (17, 29, 1), (1, 2, 0)),
    (2, (70, 9, 1), (0, 0, 0)),    # fully periodic (no reference counterpart in 2D)
    (3, (21, 17, 13), (0, 0, 0)),  # reference 3D default: periodic
    (3, (67, 9, 7), (1, 1, 1)),
    (3, (19, 11, 15), (0, 2, 1)),
    (2, (130, 5500, 1), (0, 1, 0)),  # tall enough for the 2D y-march: partial strip, chunk edges
    (2, (130, 5500, 1), (1, 2, 0)),  # y-march with a separate ghost fill (linear extrapolation)
]


@pytest.mark.parametrize("dim,n,bc", CASES)
@pytest.mark.parametrize("s_l", [0.0, 0.3])
def test_rk3_step_bitwise(ls, oracle, cuda, dim, n, bc, s_l):
    g, p = _grid_and_params(ls, dim, n, bc, s_l)
    G, vel = _inputs(g, seed=7 + dim)
    G = oracle.fill_ghosts(g, G)
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    dv = [_dev(cuda, a) for a in vel] + ([None] if dim == 2 else [])
    out = s.empty()
    dt = 0.3 * min(g.dx, g.dy)
    s.rk3_step(dG, out, dv[0], dv[1], dv[2], dt=dt)
    ref = oracle.rk3_step(g, G, *vel, dt=dt) if dim == 3 else oracle.rk3_step(g, G, vel[0], vel[1], None, dt=dt)
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), "rel Linf %.3e" % rel_linf(got, ref)


@pytest.mark.parametrize("dim,n,bc", [CASES[0], CASES[5], CASES[8]])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_rk3_stage_distinct_fields_bitwise(ls, oracle, cuda, dim, n, bc, k):
    """ls_rk3_stage with G_s and G_n distinct (also for stage 1): interior =
    the reference's stage form (rk3.h:103-150) of G_n, G_s and RHS(G_s), then
    the ghost fill."""
    g, p = _grid_and_params(ls, dim, n, bc, 0.2)
    Gs, vel = _inputs(g, seed=13)
    Gs = oracle.fill_ghosts(g, Gs)
    Gn = oracle.fill_ghosts(g, _inputs(g, seed=14)[0])
    w = vel[2] if dim == 3 else None
    dt = 1.5e-3
    s = ls.Solver(p)
    out = s.empty()
    s.rk3_stage(k, _dev(cuda, Gs), _dev(cuda, Gn), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w),
                dt=dt)
    r = oracle.rhs(g, Gs, vel[0], vel[1], w)
    if k == 1:
        e = Gn + dt * r
    elif k == 2:
        e = 0.75 * Gn + 0.25 * Gs + 0.25 * dt * r
    else:
        e = (1.0 / 3.0) * Gn + (2.0 / 3.0) * Gs + (2.0 / 3.0) * dt * r
    I = g.interior
    ref = Gn.copy()
    ref[I] = e[I]
    ref = oracle.fill_ghosts(g, ref)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("dim,n,bc", [CASES[0], CASES[1], CASES[5], CASES[8]])
def test_rk3_step_sl_field_bitwise(ls, oracle, cuda, dim, n, bc):
    g, p = _grid_and_params(ls, dim, n, bc, 0.0)
    G, vel = _inputs(g, seed=11)
    G = oracle.fill_ghosts(g, G)
    rng = np.random.default_rng(5)
    sl = rng.uniform(-0.05, 0.2, g.shape)  # includes cells below the 1e-10 gate
    s = ls.Solver(p)
    out = s.empty()
    w = vel[2] if dim == 3 else None
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w), dt=2e-3,
               sl=_dev(cuda, sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], w, dt=2e-3, sl=sl)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("dim,n,bc", CASES)
def test_fill_ghosts_bitwise(ls, oracle, cuda, dim, n, bc):
    g, p = _grid_and_params(ls, dim, n, bc, 0.0)
    G, _ = _inputs(g, seed=3)
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    s.fill_ghosts(dG)
    assert np.array_equal(dG.cpu().numpy(), oracle.fill_ghosts(g, G))


@pytest.mark.parametrize("dim,n,bc", CASES)
@pytest.mark.parametrize("iters", [1, 2, 3])
def test_hcr2_bitwise(ls, oracle, cuda, dim, n, bc, iters):
    g, p = _grid_and_params(ls, dim, n, bc, 0.0)
    g.reinit_iters = iters
    p.reinit_iters = iters
    G, _ = _inputs(g, seed=21)
    G = oracle.fill_ghosts(g, 1.7 * G)  # |grad| != 1 so the eikonal term acts
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    s.reinit(dG)
    ref = oracle.hcr2(g, G)
    got = dG.cpu().numpy()
    assert np.array_equal(got, ref), "rel Linf %.3e" % rel_linf(got, ref)


@pytest.mark.parametrize("dim,n,bc", [CASES[0], CASES[1], CASES[2], CASES[5], CASES[6], CASES[8]])
def test_advance_loop_bitwise(ls, oracle, cuda, dim, n, bc):
    """Time loop: sliver last step, HCR-2 after steps ri+1, 2ri+1, ... ."""
    g, p = _grid_and_params(ls, dim, n, bc, 0.1, ri=3)
    G, vel = _inputs(g, seed=31, vel_scale=0.5)
    G = oracle.fill_ghosts(g, G)
    dt = 0.4 * min(g.dx, g.dy) / 2.0
    t_final = 10.37 * dt  # 11 steps, the last one a sliver
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    w = vel[2] if dim == 3 else None
    steps, t, dl = s.advance(dG, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w), dt=dt,
                             t_final=t_final)
    ref, n_ref, t_ref, dl_ref = oracle.advance(g, G, vel[0], vel[1], w, dt=dt, t_final=t_final,
                                               reinit_interval=3)
    assert (steps, t, dl) == (n_ref, t_ref, dl_ref) and steps == 11
    assert np.array_equal(dG.cpu().numpy(), ref)


def test_diag_and_finite(ls, oracle, cuda):
    g, p = _grid_and_params(ls, 3, (21, 17, 13), (0, 0, 0), 0.0)
    G, _ = _inputs(g, seed=41)
    G0 = G + 1e-3 * np.sin(np.arange(G.size)).reshape(G.shape)
    s = ls.Solver(p)
    l2, vol = s.diag(_dev(cuda, G), _dev(cuda, G0))
    assert abs(l2 - oracle.l2(g, G, G0)) <= 1e-14 * oracle.l2(g, G, G0)
    assert abs(vol - oracle.area(g, G)) <= 1e-12 * oracle.area(g, G)
    assert s.check_finite(_dev(cuda, G)) == 0
    bad = G.copy()
    bad[5, 6, 7] = np.nan
    assert s.check_finite(_dev(cuda, bad)) == ls.LS_ERR_NONFINITE
    bad = G.copy()
    bad[0, 0, 0] = np.inf  # ghost cells are not checked
    assert s.check_finite(_dev(cuda, bad)) == 0


def test_deformation_velocity_in_loop(ls, oracle, cuda):
    """LS_VEL_DEFORM3D recomputes the reference deformation field every step.

    sin/cos on the device may differ from glibc in the last bit, so this case
    is held to 1e-12 relative rather than bit equality."""
    g, p = _grid_and_params(ls, 3, (24, 24, 24), (0, 0, 0), 0.0)
    p.velocity = ls.VEL_DEFORM3D
    p.vel_period = 1.5
    G = oracle.ic_sphere3d(g)
    u, v, w = oracle.deformation_velocity(g, 0.0, 1.5)
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    du, dv, dw = _dev(cuda, u), _dev(cuda, v), _dev(cuda, w)
    s.advance(dG, du, dv, dw, dt=0.01, t_final=0.2)
    ref, n, _, _ = oracle.advance(g, G, u, v, w, dt=0.01, t_final=0.2, deform_T=1.5)
    assert rel_linf(du.cpu().numpy(), u) < 1e-14
    assert rel_linf(dG.cpu().numpy(), ref) < 1e-12


def _plant_unsafe(G, seed):
    """A few values outside the class for which the stage kernels divide with
    FMA sequences (|G| above 2^50 or nonzero below 2^-200): the blocks and
    wavefronts that load them take the IEEE divisions, the others do not."""
    rng = np.random.default_rng(seed)
    flat = G.reshape(-1)
    idx = rng.choice(flat.size, size=max(4, flat.size // 2000), replace=False)
    h = len(idx) // 2
    flat[idx[:h]] *= 1e20
    flat[idx[h:]] = np.where(flat[idx[h:]] < 0, -1e-250, 1e-250)
    return G


@pytest.mark.parametrize("dim,n,bc", [CASES[0], CASES[5], CASES[8]])
def test_rk3_step_outside_fma_division_class_bitwise(ls, oracle, cuda, dim, n, bc):
    """Planted G and S_L values outside the FMA-division class: the step is
    still bitwise equal to the oracle (both division paths give the same
    bits, and the planted blocks must take the IEEE one)."""
    print("exact unsafe-class", dim, n, bc, flush=True)
    g, p = _grid_and_params(ls, dim, n, bc, 0.0)
    G, vel = _inputs(g, seed=17)
    G = oracle.fill_ghosts(g, _plant_unsafe(G, 3))
    rng = np.random.default_rng(4)
    sl = rng.uniform(0.0, 0.2, g.shape)
    sl.reshape(-1)[rng.choice(sl.size, size=max(2, sl.size // 3000), replace=False)] = 1e20
    s = ls.Solver(p)
    out = s.empty()
    w = vel[2] if dim == 3 else None
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w), dt=1e-24,
               sl=_dev(cuda, sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], w, dt=1e-24, sl=sl)
    got = out.cpu().numpy()
    assert np.isfinite(ref).all()
    assert np.array_equal(got, ref), "rel Linf %.3e" % rel_linf(got, ref)


def test_rk3_step_tiny_spacing_bitwise(ls, oracle, cuda):
    """Grid spacing below 2^-60: the host keeps every stage quotient on the
    IEEE divisions."""
    dim, n, bc = CASES[5]
    g, p = _grid_and_params(ls, dim, n, bc, 0.1)
    scale = 2.0 ** -70
    g.dx, g.dy, g.dz = g.dx * scale, g.dy * scale, g.dz * scale
    p.dx, p.dy, p.dz = g.dx, g.dy, g.dz
    G, vel = _inputs(g, seed=19)
    G = oracle.fill_ghosts(g, G * scale)
    s = ls.Solver(p)
    out = s.empty()
    dt = 0.3 * g.dx
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, vel[2]), dt=dt)
    ref = oracle.rk3_step(g, G, *vel, dt=dt)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("scale", [2.0 ** 40, 2.0 ** -150])
@pytest.mark.parametrize("dim,n,bc", [CASES[5], CASES[8]])
def test_rk3_step_extreme_magnitudes_bitwise(ls, oracle, cuda, dim, n, bc, scale):
    """G (and the S_L field) scaled towards the ends of the range in which
    the stage kernels divide with FMA sequences (|G| up to 2^50, down to
    2^-200): the quotients then reach far into the sequences' proven domain,
    and the step stays bitwise equal to the oracle."""
    print("exact extreme magnitudes", dim, n, bc, scale, flush=True)
    g, p = _grid_and_params(ls, dim, n, bc, 0.0)
    G, vel = _inputs(g, seed=23)
    G = oracle.fill_ghosts(g, G * scale)
    rng = np.random.default_rng(12)
    sl = rng.uniform(0.0, 0.2, g.shape) * scale
    s = ls.Solver(p)
    out = s.empty()
    w = vel[2] if dim == 3 else None
    dt = 0.3 * min(g.dx, g.dy)
    s.rk3_step(_dev(cuda, G), out, _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w), dt=dt, sl=_dev(cuda, sl))
    ref = oracle.rk3_step(g, G, vel[0], vel[1], w, dt=dt, sl=sl)
    got = out.cpu().numpy()
    assert np.isfinite(ref).all()
    assert np.array_equal(got, ref), "rel Linf %.3e" % rel_linf(got, ref)


@pytest.mark.parametrize("dim,n,bc", [CASES[0], CASES[5], CASES[8]])
def test_advance_graph_replays_bitwise(ls, oracle, cuda, dim, n, bc):
    """ls_advance without reinitialisation replays a captured graph of 16 RK3
    steps.  Three calls in a row (the second reuses the graph, the third has
    a new dt and recaptures), with eager stage launches and a synchronisation
    between them, each bitwise equal to the same steps as eager rk3_step
    launches; the first also to the oracle's loop, sliver step included."""
    print("advance graph", dim, n, bc, flush=True)
    g, p = _grid_and_params(ls, dim, n, bc, 0.1, ri=0)
    G, vel = _inputs(g, seed=37, vel_scale=0.5)
    G = oracle.fill_ghosts(g, G)
    w = vel[2] if dim == 3 else None
    s = ls.Solver(p)
    dG = _dev(cuda, G)
    du, dv, dw = _dev(cuda, vel[0]), _dev(cuda, vel[1]), _dev(cuda, w)
    eager = dG.clone()
    tmp = s.empty()
    t, step = 0.0, 0
    h = min(g.dx, g.dy)
    for call, (dt, nsteps) in enumerate([(0.3 * h, 37), (0.3 * h, 21), (0.2 * h, 40)]):
        t_final = t + (nsteps - 0.63) * dt  # the last step is a sliver
        k, t_gpu, dl = s.advance(dG, du, dv, dw, dt=dt, t_final=t_final, step0=step, t0=t)
        assert k == nsteps and dl < dt
        for q in range(k):  # the same steps as eager launches, the sliver last
            s.rk3_step(eager, tmp, du, dv, dw, dt=dt if q < k - 1 else dl)
            eager, tmp = tmp, eager
        cuda.cuda.synchronize()
        got = dG.cpu().numpy()
        assert np.isfinite(got).all()
        assert np.array_equal(got, eager.cpu().numpy()), "call %d: graph vs eager" % call
        if call == 0:
            ref, k_ref, t_ref, dl_ref = oracle.advance(g, G, vel[0], vel[1], w, dt=dt, t_final=t_final,
                                                       reinit_interval=0)
            assert (k, t_gpu, dl) == (k_ref, t_ref, dl_ref)
            assert np.array_equal(got, ref), "vs oracle"
        t, step = t_gpu, step + k
    s.close()
