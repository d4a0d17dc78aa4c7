"""The batched-ensemble workload itself: 96 seeded 2D flames with edge lengths
in {128, 192, 256, 384, 512}, S_L ~ U[0.05, 0.2], circle/plane G0, uniform +
8-mode velocity, zero-gradient edges, 300 RK3 steps (the package's
``ensemble()``, which the benchmark also uses).

1. The batch (autotuned schedule, hipGraph replay, as benchmarked) is bitwise
   equal to the serial per-member loop through the single-grid solver, for
   every member, in both arithmetic modes.  The single-grid path it batches is
   rk3TimeStep (Local_code/level-set_MPI_2D/include/rk3.h:168-187) run per
   member.
2. Members of edge 128, 256 and 512 from that batch, after 20 steps: bitwise
   equal to the CPU oracle's loop in exact mode, and within 1e-12 in G and
   front position in fast mode.
"""
import numpy as np
import pytest

from helpers import front_parity, rel_linf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu


def _batch(ls, cuda, mode):
    members, params, fields = ls.ensemble(96, 4, mode)
    b = ls.Batch(members, mode=mode)
    G, U, V = b.empty(), b.empty(), b.empty()
    for q, (g, u, v) in enumerate(fields):
        b.view(G, q).copy_(g)
        b.view(U, q).copy_(u)
        b.view(V, q).copy_(v)
    b.fill_ghosts(G)
    cuda.cuda.synchronize()
    return b, members, params, G, U, V


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_ensemble_300_steps_bitwise_serial(ls, cuda, mode):
    m = ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT
    b, members, params, G, U, V = _batch(ls, cuda, m)
    assert len(members) == 96 and {mm["nx"] for mm in members} == set(ls.ENSEMBLE_SIZES)
    G0 = G.clone()
    shape, _ = b.autotune(G, U, V, probe_steps=4)
    assert cuda.equal(G, G0)  # the autotuner works on a scratch copy
    b.advance(G, U, V, 300, graph=True)
    cuda.cuda.synchronize()
    bad = []
    for q, p in enumerate(params):
        s = ls.Solver(p)
        g = b.view(G0, q).clone()
        n, _, _ = s.advance(g, b.view(U, q).clone(), b.view(V, q).clone(), dt=members[q]["dt"], t_final=1e30,
                            max_steps=300)
        assert n == 300
        if not cuda.equal(g, b.view(G, q)):
            bad.append(q)
        s.close()
    print("ensemble %s: tile shape %d, 96 members bitwise equal to the serial loop: %s"
          % (mode, shape, not bad))
    assert not bad, "members differ from the serial loop: %s" % bad
    assert cuda.isfinite(G).all()
    b.close()


def test_ensemble_members_vs_oracle(ls, oracle, cuda):
    steps = 20
    picks = None
    for mode in (ls.MODE_EXACT, ls.MODE_FAST):
        b, members, params, G, U, V = _batch(ls, cuda, mode)
        if picks is None:
            picks = [next(q for q, mm in enumerate(members) if mm["nx"] == n) for n in (128, 256, 512)]
            host = {q: tuple(b.view(t, q).cpu().numpy() for t in (G, U, V)) for q in picks}
        b.advance(G, U, V, steps, graph=True)
        cuda.cuda.synchronize()
        for q in picks:
            mm = members[q]
            g = make_grid(2, mm["nx"], mm["ny"], bc=(1, 1, 0), s_l=mm["s_l"])
            assert g.dx == mm["dx"] and g.dy == mm["dy"]
            Gh, uh, vh = host[q]
            ref, n, _, _ = oracle.advance(g, Gh, uh, vh, dt=mm["dt"], t_final=1e30, max_steps=steps)
            assert n == steps
            got = b.view(G, q).cpu().numpy()
            if mode == ls.MODE_EXACT:
                assert np.array_equal(got, ref), "member %d (%d^2)" % (q, mm["nx"])
            else:
                I = g.interior
                assert rel_linf(got, ref, I) <= 1e-12, "member %d" % q
                front_parity(got, ref, (g.dy, g.dx), 1e-12, I)
        b.close()
