"""CPU pins of the sample/ restatement (oracle/ls_sample.c): bitwise against
the reference's own sample code (live when oracle/_ref/libsample.so is built,
and through the committed fixtures tests/golden/sample_vectors.npz written by
make_sample_golden.py), plus closed forms a plausible mistake would break."""
import os

import numpy as np
import pytest

from oracle.pyoracle import SampleOracle, SampleReference, sample_ref_available

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "sample_vectors.npz")


def _cases():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    return z, names


def test_restatement_matches_reference_fixtures():
    z, names = _cases()
    O = SampleOracle()
    assert len(names) == 3
    for n in names:
        nx, ny, dx, dy, dt, S_L = z[n + "/meta"]
        G, G0, u, v = z[n + "/G"], z[n + "/G0"], z[n + "/u"], z[n + "/v"]
        assert np.array_equal(O.rhs(G, u, v, dx, dy, S_L), z[n + "/rhs"]), n
        assert np.array_equal(O.rk3_step(G, u, v, dx, dy, dt, S_L, G0), z[n + "/rk3"]), n
        assert np.array_equal(O.apply_periodic(G), z[n + "/periodic"]), n
        assert np.array_equal(O.apply_anchored(G, G0), z[n + "/anchored"]), n
        assert np.array_equal(O.reinitialize(G, G0, dx, dy, 3), z[n + "/reinit"]), n
        assert np.array_equal(O.first_order_upwind(G, u, v, dx, dy, dt, S_L), z[n + "/fo"]), n


@pytest.mark.skipif(not sample_ref_available(), reason="reference sample/ build absent (oracle/_ref)")
@pytest.mark.parametrize("nx,ny,S", [(9, 13, 0.4), (40, 31, 0.4), (26, 26, 0.0)])
def test_restatement_matches_live_reference(nx, ny, S):
    O, R = SampleOracle(), SampleReference()
    rng = np.random.default_rng(nx * ny)
    dx, dy = 0.02 / (nx - 1), 0.02 / (ny - 1)
    G0 = 1e-3 * rng.standard_normal((ny + 6, nx + 6))
    G = G0 + 1e-4 * rng.standard_normal(G0.shape)
    u = 0.2 * rng.standard_normal(G.shape)
    v = 0.8 + 0.2 * rng.standard_normal(G.shape)
    for _ in range(3):  # a few chained steps
        Go, Gr = O.rk3_step(G, u, v, dx, dy, 2.5e-6, S, G0), R.rk3_step(G, u, v, dx, dy, 2.5e-6, S, G0)
        assert np.array_equal(Go, Gr)
        G = Go
    assert np.array_equal(O.reinitialize(G, G0, dx, dy, 5), R.reinitialize(G, G0, dx, dy, 5))
    assert np.array_equal(O.first_order_upwind(G, u, v, dx, dy, 1e-5, S),
                          R.first_order_upwind(G, u, v, dx, dy, 1e-5, S))


def test_constant_field_has_zero_rhs():
    O = SampleOracle()
    G = np.full((16, 18), 0.37)
    u = np.full(G.shape, 0.3)
    v = np.full(G.shape, -0.8)
    assert not O.rhs(G, u, v, 1e-3, 1e-3, 0.4).any()


def test_linear_field_closed_form():
    """G = a x + b y: all smoothness indicators of a face are equal, the
    weights are the linear ones and the faces exact, so both derivatives are
    (a, b), the dissipation vanishes and rhs = -(u a + v b - S_L sqrt(a^2+b^2))."""
    O = SampleOracle()
    nx, ny, dx, dy = 20, 16, 1e-3, 1.5e-3
    j, i = np.meshgrid(np.arange(ny + 6) - 3.0, np.arange(nx + 6) - 3.0, indexing="ij")
    a, b = 0.8, -0.6
    G = a * i * dx + b * j * dy
    u = np.full(G.shape, 0.25)
    v = np.full(G.shape, 0.9)
    r = O.rhs(G, u, v, dx, dy, 0.4)
    want = -(0.25 * a + 0.9 * b - 0.4 * np.hypot(a, b))
    I = (slice(3, ny + 3), slice(3, nx + 3))
    assert np.abs(r[I] - want).max() <= 1e-12
    assert not r[:3].any() and not r[:, :3].any()


def test_right_derivative_is_left_derivative_one_cell_over():
    """weno5_right(v0..v4) equals weno5_left(v4..v0) bit for bit; since the
    reference hands weno5_right the stencil already reversed, its dG_R at i
    is dG_L at i+1 (the identity the GPU kernel's face sharing relies on)."""
    O = SampleOracle()
    beta = O.lib.lss_beta
    wl, wr = O.lib.lss_weno_left, O.lib.lss_weno_right
    import ctypes as C
    for f in (beta, wl, wr):
        f.restype = C.c_double
    rng = np.random.default_rng(3)
    for _ in range(50):
        s = rng.standard_normal(7)
        # weno5_right(v0..v4) == weno5_left(v4..v0) bitwise
        r = wr(*[C.c_double(x) for x in s[:5]])
        l = wl(*[C.c_double(x) for x in s[4::-1]])
        assert r == l
    # Jiang-Shu beta_0 = 13/12(f0-2f1+f2)^2 + 1/4(f0-4f1+3f2)^2 (sample's beta for every sub-stencil)
    assert beta(C.c_double(1.0), C.c_double(2.0), C.c_double(4.0)) == (13.0 / 12.0) * 1.0 + 0.25 * 25.0


def test_write_G_is_interior_rows(tmp_path):
    O = SampleOracle()
    G = np.arange(14 * 17, dtype=np.float64).reshape(14, 17)
    p = tmp_path / "g.bin"
    assert O.write_G(G, p) == 0
    got = np.fromfile(p, dtype=np.float64)
    assert np.array_equal(got, G[3:-3, 3:-3].ravel())
