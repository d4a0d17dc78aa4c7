"""Pins the CPU oracle (oracle/ls_oracle.c) before it is trusted as the checker.

1. Golden vectors written by tests/golden/make_golden.py from the reference's
   own solver compiled from its sources (contraction off): bit-for-bit.
2. Live comparison against that reference build when oracle/_ref is present.
3. The reference's published known answers (run logs), which pin WENO5 + RK3
   + ghost fill + time loop with S_L = 0 and no reinitialisation:
   2D Zalesak 201^2 (Local_code/level-set_MPI_2D/log/201x201.log:20-34) and the
   step counts of every logged run (sliver last step).
4. Closed forms and invariants that a plausible slip would break: constant and
   linear fields, the ENO property on a step, the u_eff formula, HCR-2 fixed
   points, periodic translation, zero-gradient ghost copies.
"""
import math
import os

import numpy as np
import pytest

from helpers import bumpy_sdf, rand_vel
from oracle.pyoracle import Oracle, Reference, make_grid, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def _cases():
    z = np.load(GOLDEN)
    names = sorted({k.split("/")[0] for k in z.files})
    return z, names


@pytest.mark.parametrize("name", _cases()[1])
def test_golden_reference_vectors(oracle, name):
    z = np.load(GOLDEN)
    m = z[name + "/meta"]
    dim = int(m[0])
    g = make_grid(dim, int(m[1]), int(m[2]), int(m[3]), bc=(int(m[4]), int(m[5]), int(m[6])), s_l=m[7])
    dt, t_final = float(m[8]), float(m[9])
    G, u, v = z[name + "/G"], z[name + "/u"], z[name + "/v"]
    w = z[name + "/w"] if dim == 3 else None
    assert np.array_equal(oracle.fill_ghosts(g, G), G)
    assert np.array_equal(oracle.rhs(g, G, u, v, w), z[name + "/rhs"])
    assert np.array_equal(oracle.rk3_step(g, G, u, v, w, dt=dt), z[name + "/rk3"])
    assert np.array_equal(oracle.hcr2(g, oracle.fill_ghosts(g, 1.7 * G)), z[name + "/hcr2"])
    adv, steps, _, _ = oracle.advance(g, G, u.copy(), v.copy(), None if w is None else w.copy(),
                                      dt=dt, t_final=t_final, reinit_interval=3)
    assert steps == 11
    assert np.array_equal(adv, z[name + "/adv"])


@pytest.mark.skipif(not (ref_available(2) and ref_available(3)), reason="reference build absent")
@pytest.mark.parametrize("dim,bc", [(2, (0, 1, 0)), (2, (1, 1, 0)), (2, (0, 2, 0)), (2, (1, 2, 0)),
                                    (3, (0, 0, 0))])
@pytest.mark.parametrize("s_l", [0.0, 0.15])
def test_live_reference_bitwise(oracle, dim, bc, s_l):
    R = Reference(dim)
    n = (17, 13, 11)
    g = make_grid(dim, *n, bc=bc, s_l=s_l)
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    G = R.fill_ghosts(g, bumpy_sdf(g.shape, h, 99 + dim))
    vel = rand_vel(g.shape, 5, 1.0, n=dim)
    w = vel[2] if dim == 3 else None
    assert np.array_equal(oracle.rk3_step(g, G, vel[0], vel[1], w, dt=1e-3),
                          R.rk3_step(g, G, vel[0], vel[1], w, dt=1e-3))
    assert np.array_equal(oracle.hcr2(g, G), R.hcr2(g, G))
    # the reference's own rk3TimeStep / reinitializeWithSwap, called directly
    if (dim == 2 and bc == (0, 1, 0)) or dim == 3:
        assert np.array_equal(oracle.rk3_step(g, G, vel[0], vel[1], w, dt=1e-3),
                              R.rk3_step(g, G, vel[0], vel[1], w, dt=1e-3, direct=True))
        assert np.array_equal(oracle.hcr2(g, G), R.hcr2(g, G, direct=True))
    a = oracle.advance(g, G, vel[0], vel[1], w, dt=1e-3, t_final=0.0203, reinit_interval=4)
    b = R.advance(g, G, vel[0].copy(), vel[1].copy(), None if w is None else w.copy(), dt=1e-3,
                  t_final=0.0203, reinit_interval=4)
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


@pytest.mark.skipif(not ref_available(2, fma=True), reason="reference build absent")
def test_initial_conditions_match_reference(oracle):
    for fma in (False, True):
        o, R = Oracle(fma=fma), Reference(2, fma=fma)
        g = make_grid(2, 201, 201, bc=(0, 1, 0))
        Gr, ur, vr, _ = R.init_case(g, 2, t_final=6.283185)
        Go, uo, vo = o.ic_zalesak2d(g, 6.283185)
        assert np.array_equal(Gr, Go) and np.array_equal(ur, uo) and np.array_equal(vr, vo)
    R3 = Reference(3)
    g = make_grid(3, 24, 24, 24, bc=(0, 0, 0))
    Gr, ur, vr, wr = R3.init_case(g, 0, t_final=1.5)
    assert np.array_equal(Gr, oracle.ic_sphere3d(g))
    for a, b in zip((ur, vr, wr), oracle.deformation_velocity(g, 0.0, 1.5)):
        assert np.array_equal(a, b)


def _count_steps(T, dt):
    # main.cpp:314-336: while (t < T) { if (t + dt > T) dt = T - t; ...; t += dt; step++; }
    t, n = 0.0, 0
    while t < T:
        if t + dt > T:
            dt = T - t
        t += dt
        n += 1
    return n


def test_logged_step_counts():
    """Every logged run's step count follows from the sliver-step loop."""
    for N, steps in ((201, 6284), (401, 12567), (801, 25133), (1601, 50266)):
        dx = (1.0 - 0.0) / (N - 1)
        assert _count_steps(6.283185, 0.2 * dx / 1.0) == steps
    assert _count_steps(1.5, 0.001) == 1501  # Local_code/level-set_MPI_3D/log/simulation.log:31


def test_zalesak_201_log_known_answer():
    """Local_code/level-set_MPI_2D/log/201x201.log:20-28 (32 MPI ranks, -O3 -march=native)."""
    o = Oracle(fma=True)
    g = make_grid(2, 201, 201, bc=(0, 1, 0))
    T = 6.283185
    G0, u, v = o.ic_zalesak2d(g, T)
    G0 = o.fill_ghosts(g, G0)
    assert "%.6f" % o.area(g, G0) == "0.057850"
    dt = 0.2 * g.dx / 1.0
    for flavour in (Oracle(fma=False), o):
        G, n, t, _ = flavour.advance(g, G0, u, v, dt=dt, t_final=T)
        assert n == 6284
        assert "%.6e" % flavour.l2(g, G, G0) == "8.576633e-02"
        assert "%.6f" % flavour.area(g, G) == "0.058800"


@pytest.mark.slow
def test_sphere_deformation_64_log_known_answer():
    """Local_code/level-set_MPI_3D/log/simulation.log:9-34 (opt in: several minutes on CPU)."""
    o = Oracle(fma=True)
    g = make_grid(3, 64, 64, 64, bc=(0, 0, 0))
    G0 = o.ic_sphere3d(g)
    u, v, w = o.deformation_velocity(g, 0.0, 1.5)
    G, n, _, _ = o.advance(g, G0, u, v, w, dt=0.001, t_final=1.5, deform_T=1.5)
    assert n == 1501
    assert "%.6e" % o.l2(g, G, G0) == "9.158634e-04"
    assert "%.6f" % o.area(g, G0) == "0.014165" and "%.6f" % o.area(g, G) == "0.014245"


# ---- closed forms and invariants -------------------------------------------

def test_weno_linear_exact_and_eno(oracle):
    v = 0.3 + 0.7 * np.arange(5.0)           # linear data: all beta equal, faces exact
    assert abs(oracle.weno_left(v) - (0.3 + 0.7 * 2.5)) < 1e-14
    assert abs(oracle.weno_right(v) - (0.3 + 0.7 * 1.5)) < 1e-14
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0])
    # ENO: the left-biased face value of a step comes from the flat sub-stencil
    assert abs(oracle.weno_left(step)) < 1e-9
    # right-biased face (left face of v2) of the mirrored step sees the flat side
    assert abs(oracle.weno_right(step[::-1].copy()) - 0.0) < 1e-9
    # the reference's right-biased weights: (0.3, 0.6, 0.1) on the sub-stencils
    # (v2,v3,v4), (v1,v2,v3), (v0,v1,v2) [weno5.h:66-68]; on smooth quadratic
    # data all beta equal, so the result is the fixed linear combination
    q = np.arange(5.0) ** 2 * 1e-3 + 0.5   # beta_k = 13/12 (2e-3)^2 + 1/4 (4e-3 (k+?))^2 differ;
    # use data with equal beta: symmetric |second difference| and equal spans
    lin = 2.0 - 0.25 * np.arange(5.0)
    want = (-0.1 * lin[0] + 1.7 * lin[1] + 6.5 * lin[2] - 2.7 * lin[3] + 0.6 * lin[4]) / 6.0
    assert abs(oracle.weno_right(lin) - want) < 1e-14
    del q


def test_rhs_constant_and_plane(oracle):
    g = make_grid(3, 12, 10, 9, bc=(1, 1, 1), s_l=0.2)
    sh = g.shape
    z = np.zeros(sh)
    rng = np.random.default_rng(0)
    u, v, w = (rng.standard_normal(sh) for _ in range(3))
    R = oracle.rhs(g, np.full(sh, 0.37), u, v, w)
    # constant field: upwinded differences vanish, gradient gate gives u_eff = u
    assert np.count_nonzero(R) == 0
    # plane G = a.x: every stencil is linear, so RHS = -(u_eff . a) with
    # u_eff = u - S_L a / sqrt(|a|^2 + eps) (rk3.h:31-81 applied to exact data)
    a = np.array([0.6, -0.8, 0.0])
    kk, jj, ii = np.meshgrid(*[np.arange(s, dtype=float) for s in sh], indexing="ij")
    G = a[0] * ii * g.dx + a[1] * jj * g.dy + a[2] * kk * g.dz
    R = oracle.rhs(g, G, u, v, w)
    m = math.sqrt(a @ a + 1e-10)
    ue = u - 0.2 * a[0] / m
    ve = v - 0.2 * a[1] / m
    want = -(ue * a[0] + ve * a[1])
    I = g.interior
    assert np.abs(R[I] - want[I]).max() < 1e-12
    del z


def test_rk3_plane_translation_closed_form(oracle):
    """Linear G with uniform u and S_L: every stage is exact on linear data, so
    G_n = G_0 - n dt (u_eff . a) with u_eff = u - S_L a/sqrt(|a|^2 + eps) away
    from the zero-gradient edges (their kink travels 3 cells per stage)."""
    g = make_grid(2, 80, 80, bc=(1, 1, 0), s_l=0.1)
    sh = g.shape
    jj, ii = np.meshgrid(*[np.arange(s, dtype=float) for s in sh], indexing="ij")
    a = (0.5, -0.3)
    G0 = a[0] * (ii - 3) * g.dx + a[1] * (jj - 3) * g.dy + 0.25
    u = np.full(sh, 0.3)
    v = np.full(sh, -0.7)
    dt = 0.2 * g.dx
    G, n, _, _ = oracle.advance(g, G0, u, v, dt=dt, t_final=1e30, max_steps=2)
    m = math.sqrt(a[0] ** 2 + a[1] ** 2 + 1e-10)
    rate = (0.3 - 0.1 * a[0] / m) * a[0] + (-0.7 - 0.1 * a[1] / m) * a[1]
    want = G0 - 2 * dt * rate
    mid = (slice(3 + 20, 3 + 60), slice(3 + 20, 3 + 60))
    assert n == 2
    assert np.abs(G[mid] - want[mid]).max() < 1e-13


def test_hcr2_fixed_point_plane(oracle):
    """A plane SDF has |grad phi| = 1 and r~ sum(phi) = phi at interface cells,
    so HCR-2 leaves it unchanged to rounding (reinitialization.h:200-263)."""
    g = make_grid(3, 16, 14, 12, bc=(1, 1, 1))
    sh = g.shape
    kk, jj, ii = np.meshgrid(*[np.arange(s, dtype=float) for s in sh], indexing="ij")
    nrm = np.array([1.0, 0.0, 0.0])
    G = (ii - 3) * g.dx - 0.4037
    G = oracle.fill_ghosts(g, G)
    H = oracle.hcr2(g, G)
    I = (slice(4, 4 + 10), slice(4, 4 + 10), slice(4, 4 + 10))
    assert np.abs(H[I] - G[I]).max() < 1e-14
    del nrm, kk, jj


def test_hcr2_no_sign_flip_and_distance_restored(oracle):
    """HCR-2 pulls a stretched SDF towards |grad| = 1 without moving the front."""
    g = make_grid(2, 64, 64, bc=(1, 1, 0))
    sh = g.shape
    jj, ii = np.meshgrid(*[np.arange(s, dtype=float) for s in sh], indexing="ij")
    d = np.sqrt(((ii - 3) * g.dx - 0.5) ** 2 + ((jj - 3) * g.dy - 0.5) ** 2) - 0.25
    G = oracle.fill_ghosts(g, 3.0 * d)
    H = G.copy()
    for _ in range(40):
        H = oracle.hcr2(g, H)
    I = g.interior
    band = np.abs(d[I]) < 3 * g.dx
    assert np.array_equal(np.sign(H[I][band]), np.sign(G[I][band]))
    gy, gx = np.gradient(H, g.dy, g.dx)
    gm = np.sqrt(gx ** 2 + gy ** 2)[I]
    ring = (np.abs(d[I]) > 3 * g.dx) & (np.abs(d[I]) < 0.15)
    assert abs(np.median(gm[ring]) - 1.0) < 0.05


@pytest.mark.parametrize("dim", [2, 3])
def test_periodic_translation_equivariance(oracle, dim):
    n = (14, 12, 10)
    g = make_grid(dim, *n, bc=(0, 0, 0), s_l=0.1)
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    G = oracle.fill_ghosts(g, bumpy_sdf(g.shape, h, 3))
    vel = [oracle.fill_ghosts(g, a) for a in rand_vel(g.shape, 4, 0.5, n=dim)]
    I = g.interior
    shift = (2, -3, 5)[:dim]
    axes = tuple(range(dim))

    def roll(a):
        out = a.copy()
        out[I] = np.roll(a[I], shift, axes)
        return oracle.fill_ghosts(g, out)
    w = vel[2] if dim == 3 else None
    A, _, _, _ = oracle.advance(g, G, vel[0], vel[1], w, dt=2e-3, t_final=1e30, max_steps=4,
                                reinit_interval=2)
    B, _, _, _ = oracle.advance(g, roll(G), roll(vel[0]), roll(vel[1]),
                                None if w is None else roll(w), dt=2e-3, t_final=1e30,
                                max_steps=4, reinit_interval=2)
    assert np.array_equal(roll(A)[I], B[I])


def test_ghost_fill_rules(oracle):
    g = make_grid(2, 7, 6, bc=(0, 2, 0))
    G = np.arange(13 * 12, dtype=float).reshape(12, 13) ** 1.5
    F = oracle.fill_ghosts(g, G)
    I = g.interior
    assert np.array_equal(F[I], G[I])
    # periodic x: ghost i <- i + nx, ghost nx+3+q <- 3+q (period nx cells)
    assert np.array_equal(F[3:9, 0:3], F[3:9, 7:10]) and np.array_equal(F[3:9, 10:13], F[3:9, 3:6])
    # linear extrapolation in y on x-filled columns
    assert np.allclose(F[2], 2 * F[3] - F[4], rtol=0, atol=1e-12)
    assert np.allclose(F[0], F[3] + 3 * (F[3] - F[4]), rtol=0, atol=1e-12)
    assert np.allclose(F[9], 2 * F[8] - F[7], rtol=0, atol=1e-12)
    g2 = make_grid(2, 7, 6, bc=(1, 1, 0))
    F2 = oracle.fill_ghosts(g2, G)
    assert np.array_equal(F2[0], F2[3]) and np.array_equal(F2[:, 12], F2[:, 9])


def test_write_G_format(oracle, tmp_path):
    g = make_grid(2, 7, 6, bc=(0, 1, 0))
    G = np.random.default_rng(1).standard_normal(g.shape)
    p = tmp_path / "g.bin"
    oracle.write_G(g, G, p)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [7, 6, 3]
    body = np.frombuffer(raw[12:], np.float64).reshape(g.shape)
    assert np.array_equal(body[3:9], G[3:9])
    assert np.array_equal(body[0], G[3]) and np.array_equal(body[11], G[8])
    if ref_available(2):
        q = tmp_path / "r.bin"
        Reference(2).write_G(g, G, q)
        assert q.read_bytes() == raw
    g3 = make_grid(3, 5, 4, 3, bc=(0, 0, 0))
    G3 = np.random.default_rng(2).standard_normal(g3.shape)
    oracle.write_G(g3, G3, p)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:16], np.int32).tolist() == [5, 4, 3, 3]
    if ref_available(3):
        q = tmp_path / "r3.bin"
        Reference(3).write_G(g3, G3, q)
        assert q.read_bytes() == raw


CROSSINGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "crossings_vectors.npz")


def _crossing_cases():
    z = np.load(CROSSINGS)
    return sorted({k.split("/")[0] for k in z.files})


@pytest.mark.parametrize("name", _crossing_cases())
def test_golden_crossings(oracle, name):
    """computeInterfaceCrossings[3D] (Local_code/level-set_MPI_2D/include/reinitialization.h:75-130,
    3D :46-89): the oracle's r~ and bit mask equal the reference's bit for bit,
    including the planted cells of make_crossings_golden.py."""
    z = np.load(CROSSINGS)
    m = z[name + "/meta"]
    dim = int(m[0])
    g = make_grid(dim, int(m[1]), int(m[2]), int(m[3]), bc=(int(m[4]), int(m[5]), int(m[6])))
    p0 = z[name + "/p0"]
    rt, fl = oracle.crossings(g, p0)
    assert np.array_equal(rt, z[name + "/rt"]) and np.array_equal(fl, z[name + "/flag"])
    # the planted centre cell: two crossing neighbours summing to -9e-16, below
    # the 1e-15 gate, so r~ = 0 while both bits are set
    c = tuple(s // 2 for s in g.shape)
    lo = [1 << (2 * (dim - 1 - a)) for a in range(dim)]   # the low-side bit of array axis a
    hi = [b << 1 for b in lo]
    assert rt[c] == 0.0 and fl[c] == lo[0] | hi[1]
    # 3 cells to the low x side: the sum -1.1e-15 passes the gate
    c2 = c[:-1] + (c[-1] - 3,)
    assert fl[c2] == lo[0] | hi[1] and rt[c2] == 0.3 / (-6e-16 + -5e-16)
    # a centre of exactly 0 has no crossing (the product is 0, not < 0)
    c3 = c[:-1] + (c[-1] + 3,)
    assert fl[c3] == 0 and rt[c3] == 0.0
    # every ghost entry is left at 0
    mask = np.zeros(g.shape, bool)
    mask[g.interior] = True
    assert not fl[~mask].any() and not rt[~mask].any()
    # flagged neighbours have the opposite sign of phi_c, so r~ <= 0 everywhere
    assert (rt <= 0.0).all() and (rt < 0.0).any()


@pytest.mark.skipif(not (ref_available(2) and ref_available(3)), reason="reference build absent")
@pytest.mark.parametrize("dim", [2, 3])
def test_live_reference_crossings(oracle, dim):
    R = Reference(dim)
    n = (19, 14, 12)
    g = make_grid(dim, *n, bc=(0, 0, 0) if dim == 3 else (0, 1, 0))
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    for seed in range(3):
        p0 = R.fill_ghosts(g, bumpy_sdf(g.shape, h, 50 + seed, noise=0.05))
        a, b = oracle.crossings(g, p0), R.crossings(g, p0)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert (a[1] != 0).sum() > 10


@pytest.mark.parametrize("fma", [False, True])
def test_timed_loop_is_the_advance_loop(fma):
    """The CPU baseline's timed loop runs the same steps as the oracle's time
    loop (HCR-2 after step > 0 with step % ri == 0), with the in-loop
    diagnostics of the reference on top."""
    o = Oracle(fma=fma)
    g = make_grid(3, 12, 10, 9, bc=(0, 0, 0), s_l=0.1)
    h = (g.dz, g.dy, g.dx)
    G = o.fill_ghosts(g, bumpy_sdf(g.shape, h, 7))
    vel = rand_vel(g.shape, 7, 0.5, n=3)
    ref, n, _, _ = o.advance(g, G, *[a.copy() for a in vel], dt=2e-3, t_final=1e30, max_steps=11, reinit_interval=5)
    for threads in (1, 3):
        sec, got, l2, area = o.timed_loop(g, G, G, *[a.copy() for a in vel], dt=2e-3, nsteps=11, reinit_interval=5,
                                          diag_interval=5, threads=threads)
        assert sec > 0.0 and np.array_equal(got, ref)
    # the last diagnostics were taken after step 10
    mid, _, _, _ = o.advance(g, G, *[a.copy() for a in vel], dt=2e-3, t_final=1e30, max_steps=10, reinit_interval=5)
    assert l2 == o.l2(g, mid, G) and area == o.area(g, mid)
