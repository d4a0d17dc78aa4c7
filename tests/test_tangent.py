"""CPU pins of the forward-mode checker (oracle/ls_tangent.cpp) that the GPU
adjoint is tested against: its value part must be bitwise the C oracle's, its
tangent must match central finite differences of the oracle, and a constant
shift of G0 must pass through RK3 unchanged (WENO-5 differences are
shift-invariant), which fixes dL in closed form."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, Tangent, heaviside, make_grid


def _case(bc=(0, 0, 0), sl_field=True, seed=3):
    g = make_grid(3, 16, 14, 12, bc=bc, s_l=0.05)
    sh = g.shape
    z, y, x = np.meshgrid(*[np.arange(s) - 3.0 for s in sh], indexing="ij")
    X, Y, Z = x * g.dx, y * g.dy, z * g.dz
    rng = np.random.default_rng(seed)
    G0 = np.sqrt((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) - 0.25 + 0.01 * np.sin(7 * X)
    G0 = Oracle().fill_ghosts(g, G0)
    u = 0.5 + 0.2 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal(sh)
    v = 0.3 * np.cos(2 * np.pi * Z)
    w = -0.2 + 0.1 * np.sin(2 * np.pi * X)
    sl = 0.04 + 0.05 * (X + Y) if sl_field else None
    return g, G0, u, v, w, sl


@pytest.mark.parametrize("bc,sl_field,ns,ri", [((0, 0, 0), True, 6, 2), ((1, 1, 1), False, 5, 0),
                                                ((0, 1, 0), True, 7, 3)])
def test_value_part_is_the_oracle(bc, sl_field, ns, ri):
    g, G0, u, v, w, sl = _case(bc, sl_field)
    O = Oracle()
    Go, n, _, _ = O.advance(g, G0, u, v, w, sl=sl, dt=0.2 * g.dx, t_final=1e9, max_steps=ns,
                            reinit_interval=ri)
    GT, L, _ = Tangent().run(g, G0, u, v, w, G0, 1.5 * g.dx, 0.2 * g.dx, ns, ri, sl=sl)
    assert n == ns and np.array_equal(GT, Go)
    I = g.interior
    ref = np.sum((heaviside(GT[I], 1.5 * g.dx) - heaviside(G0[I], 1.5 * g.dx)) ** 2)
    assert abs(L - ref) <= 1e-12 * ref


@pytest.mark.parametrize("ri", [0, 2])
def test_tangent_matches_central_differences(ri):
    g, G0, u, v, w, sl = _case()
    O, T = Oracle(), Tangent()
    dt, e, ns = 0.2 * g.dx, 1.5 * g.dx, 6
    Gt = G0 + 0.003
    I = g.interior

    def loss(G0x=G0, ux=u, slx=sl):
        G, _, _, _ = O.advance(g, O.fill_ghosts(g, G0x), ux, v, w, sl=slx, dt=dt, t_final=1e9,
                               max_steps=ns, reinit_interval=ri)
        return np.sum((heaviside(G[I], e) - heaviside(Gt[I], e)) ** 2)

    rng = np.random.default_rng(7)
    h = 1e-5
    for name in ("G0", "u", "sl"):
        d = 0.01 * rng.standard_normal(g.shape)
        kw = {"d" + name: d}
        _, _, dL = T.run(g, G0, u, v, w, Gt, e, dt, ns, ri, sl=sl, **kw)
        arg = {"G0": "G0x", "u": "ux", "sl": "slx"}[name]
        base = {"G0": G0, "u": u, "sl": sl}[name]
        fd = (loss(**{arg: base + h * d}) - loss(**{arg: base - h * d})) / (2 * h)
        assert abs(dL - fd) <= 1e-7 * abs(fd), (name, dL, fd)


def test_constant_shift_of_G0_passes_through_rk3():
    g, G0, u, v, w, sl = _case()
    e, dt, ns = 1.5 * g.dx, 0.2 * g.dx, 5
    Gt = G0 + 0.003
    GT, L, dL = Tangent().run(g, G0, u, v, w, Gt, e, dt, ns, 0, sl=sl, dG0=np.ones(g.shape))
    I = g.interior
    x = GT[I]
    r = heaviside(x, e) - heaviside(Gt[I], e)
    dH = np.where(np.abs(x) > e, 0.0, (1.0 + np.cos(np.pi * x / e)) / (2 * e))
    want = np.sum(2.0 * r * dH)
    assert abs(dL - want) <= 1e-10 * np.sum(np.abs(2.0 * r * dH))
