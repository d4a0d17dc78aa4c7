"""Batched engine (many uneven 2D flames per launch) against the oracle.

Each member is advanced by the batched kernels; the reference result of a
member is the oracle's single-grid time loop on the same inputs (the
reference's rk3TimeStep semantics, Local_code/level-set_MPI_2D/include/rk3.h:168-187,
fixed step count, no reinitialisation)."""
import numpy as np
import pytest

from helpers import bumpy_sdf, rand_vel, rel_linf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu

MEMBERS = [  # nx, ny, bc, s_l
    (24, 20, (1, 1), 0.1),
    (33, 17, (0, 1), 0.0),
    (70, 9, (0, 0), 0.2),
    (48, 40, (1, 1), 0.05),
    (130, 12, (1, 0), 0.15),
    (16, 64, (0, 1), 0.12),
]


def _build(ls, cuda, mode):
    members, inputs = [], []
    for q, (nx, ny, bc, s_l) in enumerate(MEMBERS):
        g = make_grid(2, nx, ny, bc=(bc[0], bc[1], 0), s_l=s_l)
        dt = 0.3 * min(g.dx, g.dy)
        members.append(dict(nx=nx, ny=ny, bc=bc, dx=g.dx, dy=g.dy, s_l=s_l, dt=dt))
        G = bumpy_sdf(g.shape, (g.dy, g.dx), 50 + q)
        u, v = rand_vel(g.shape, 50 + q, 0.8, n=2)
        inputs.append((g, G, u, v, dt))
    b = ls.Batch(members, mode=mode)
    G, U, V = b.empty(), b.empty(), b.empty()
    for q, (g, Gh, uh, vh, _) in enumerate(inputs):
        b.view(G, q).copy_(cuda.from_numpy(Gh))
        b.view(U, q).copy_(cuda.from_numpy(uh))
        b.view(V, q).copy_(cuda.from_numpy(vh))
    b.fill_ghosts(G)
    for q, (g, *_rest) in enumerate(inputs):
        inputs[q] = (g, b.view(G, q).cpu().numpy(), *_rest[1:])
    return b, G, U, V, inputs


@pytest.mark.parametrize("graph", [False, True])
def test_batch_exact_bitwise(ls, oracle, cuda, graph):
    b, G, U, V, inputs = _build(ls, cuda, ls.MODE_EXACT)
    steps = 12
    b.advance(G, U, V, steps, graph=graph)
    cuda.cuda.synchronize()
    for q, (g, Gh, uh, vh, dt) in enumerate(inputs):
        ref, n, _, _ = oracle.advance(g, Gh, uh, vh, dt=dt, t_final=1e30, max_steps=steps)
        assert n == steps
        assert np.array_equal(b.view(G, q).cpu().numpy(), ref), "member %d" % q


def test_batch_fast_within_tolerance(ls, oracle, cuda):
    b, G, U, V, inputs = _build(ls, cuda, ls.MODE_FAST)
    b.advance(G, U, V, 12)
    cuda.cuda.synchronize()
    for q, (g, Gh, uh, vh, dt) in enumerate(inputs):
        ref, _, _, _ = oracle.advance(g, Gh, uh, vh, dt=dt, t_final=1e30, max_steps=12)
        assert rel_linf(b.view(G, q).cpu().numpy(), ref, g.interior) <= 1e-12, "member %d" % q


def test_batch_graph_replays(ls, cuda):
    """Replays of the captured graph, with eager work between them, match the
    eager launches each time."""
    b, G, U, V, _ = _build(ls, cuda, ls.MODE_FAST)
    G0 = G.clone()
    eager = G0.clone()
    graphed = G0.clone()
    for rep in range(3):
        b.advance(eager, U, V, 4, graph=False)
        b.advance(graphed, U, V, 4, graph=True)
        scratch = cuda.randn(1 << 20, device="cuda", dtype=cuda.float64)
        _ = (scratch * 2).sum().item()
        cuda.cuda.synchronize()
        assert cuda.equal(eager, graphed), "replay %d" % rep


def test_batch_exact_all_shapes_bitwise(ls, oracle, cuda):
    """Exact mode on every schedule shape (chunk length x register budget):
    bitwise the oracle each time."""
    b, G, U, V, inputs = _build(ls, cuda, ls.MODE_EXACT)
    steps = 5
    refs = [oracle.advance(g, Gh, uh, vh, dt=dt, t_final=1e30, max_steps=steps)[0]
            for (g, Gh, uh, vh, dt) in inputs]
    for shape in range(8):
        Gs = G.clone()
        b.set_tile(shape)
        b.advance(Gs, U, V, steps, graph=False)
        cuda.cuda.synchronize()
        for q in range(len(inputs)):
            assert np.array_equal(b.view(Gs, q).cpu().numpy(), refs[q]), "shape %d member %d" % (shape, q)


def test_batch_exact_outside_fma_division_class_bitwise(ls, oracle, cuda):
    """One member with planted values outside the class for which the exact
    y-march divides with FMA sequences (|G| above 2^50, nonzero |G| below
    2^-200): its wavefronts that load them take the IEEE divisions; every
    member stays bitwise equal to the oracle."""
    b, G, U, V, inputs = _build(ls, cuda, ls.MODE_EXACT)
    q = 3
    g, Gh, uh, vh, dt = inputs[q]
    rng = np.random.default_rng(8)
    Gp = Gh.copy()
    flat = Gp.reshape(-1)
    idx = rng.choice(flat.size, size=12, replace=False)
    flat[idx[:6]] *= 1e20
    flat[idx[6:]] = 1e-250
    b.view(G, q).copy_(cuda.from_numpy(Gp))
    b.fill_ghosts(G)
    inputs[q] = (g, b.view(G, q).cpu().numpy(), uh, vh, dt)
    steps = 3
    b.advance(G, U, V, steps, graph=False)
    cuda.cuda.synchronize()
    for m, (gm, Gm, um, vm, dtm) in enumerate(inputs):
        ref = oracle.advance(gm, Gm, um, vm, dt=dtm, t_final=1e30, max_steps=steps)[0]
        assert np.isfinite(ref).all()
        assert np.array_equal(b.view(G, m).cpu().numpy(), ref), "member %d" % m


def test_batch_autotune_keeps_results(ls, cuda):
    b, G, U, V, _ = _build(ls, cuda, ls.MODE_EXACT)
    G2 = G.clone()
    b.set_tile(0)
    b.advance(G, U, V, 5, graph=False)
    shape, ms = b.autotune(G2, U, V, probe_steps=2)
    assert 0 <= shape < 8 and ms > 0
    b.advance(G2, U, V, 5, graph=True)
    cuda.cuda.synchronize()
    assert cuda.equal(G, G2)  # the arithmetic does not depend on the chunk length


def test_batch_fast_shapes_match_serial(ls, cuda):
    """Fast mode: every schedule shape (chunk length, register budget) gives the same
    bits, and so does the single-grid solver run member by member: one thread
    per cell for the small members, its own y-march for the tall one."""
    spec = MEMBERS + [(130, 5500, (0, 1), 0.1)]
    members, params = [], []
    for nx, ny, bc, s_l in spec:
        p = ls.default_params(2, nx, nx=nx, ny=ny, bc=(bc[0], bc[1], 0), s_l=s_l, reinit_interval=0,
                              mode=ls.MODE_FAST)
        params.append(p)
        members.append(dict(nx=nx, ny=ny, bc=bc, dx=p.dx, dy=p.dy, s_l=s_l, dt=0.3 * min(p.dx, p.dy)))
    b = ls.Batch(members, mode=ls.MODE_FAST)
    G, U, V = b.empty(), b.empty(), b.empty()
    for q, p in enumerate(params):
        h = (p.dy, p.dx)
        b.view(G, q).copy_(cuda.from_numpy(bumpy_sdf(p.shape, h, 70 + q)))
        uh, vh = rand_vel(p.shape, 70 + q, 0.8, n=2)
        b.view(U, q).copy_(cuda.from_numpy(uh))
        b.view(V, q).copy_(cuda.from_numpy(vh))
    b.fill_ghosts(G)
    steps = 6
    results = []
    for shape in range(8):
        Gs = G.clone()
        b.set_tile(shape)
        b.advance(Gs, U, V, steps, graph=False)
        results.append(Gs)
    cuda.cuda.synchronize()
    for shape in range(1, 8):
        assert cuda.equal(results[0], results[shape]), "shape %d" % shape
    for q, p in enumerate(params):
        s = ls.Solver(p)
        Gq = b.view(G, q).clone()
        s.advance(Gq, b.view(U, q).clone(), b.view(V, q).clone(), dt=members[q]["dt"], t_final=1e30,
                  max_steps=steps)
        cuda.cuda.synchronize()
        assert cuda.equal(Gq, b.view(results[0], q)), "member %d" % q


def test_batch_rejects_bad_members(ls):
    with pytest.raises(ls.LSError):
        ls.Batch([dict(nx=4, ny=20, dx=0.1, dy=0.1, s_l=0.0, dt=1e-3)])
    with pytest.raises(ls.LSError):
        ls.Batch([dict(nx=20, ny=20, bc=(2, 1), dx=0.1, dy=0.1, s_l=0.0, dt=1e-3)])
