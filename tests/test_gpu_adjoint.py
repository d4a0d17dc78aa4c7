"""GPU discrete adjoint (ls_adjoint_run, ls_flame_*) against the CPU checkers.

* Dot-product tests: for random directions e, <grad_adjoint, e> must equal the
  forward-mode derivative of the restatement (oracle/ls_tangent.cpp, pinned
  bitwise to the C oracle and to finite differences by tests/test_tangent.py)
  to 1e-10 relative (exact mode; 1e-8 with the fast-mode forward).
* The adjoint's loss-only sweep and the plain solver give the same loss.
* The flame-inversion parameter gradient (S_L on a coarse grid, velocity-mode
  amplitudes) matches central finite differences (h = 1e-6) through the plain
  GPU solver to 1e-6 relative, and the torch.autograd binding returns it.
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle, Tangent, heaviside


pytestmark = pytest.mark.gpu


def _fields(p, seed, sl_field=True):
    sh = p.shape
    z, y, x = np.meshgrid(*[np.arange(s) - 3.0 for s in sh], indexing="ij")
    X, Y, Z = x * p.dx, y * p.dy, z * p.dz
    rng = np.random.default_rng(seed)
    G0 = np.sqrt((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2) - 0.25 + 0.01 * np.sin(7 * X + seed)
    u = 0.5 + 0.2 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal(sh)
    v = 0.3 * np.cos(2 * np.pi * Z) + 0.05 * rng.standard_normal(sh)
    w = -0.2 + 0.1 * np.sin(2 * np.pi * X)
    sl = 0.04 + 0.05 * (X + Y) if sl_field else None
    return G0, u, v, w, sl


def _grid(p):
    from oracle.pyoracle import make_grid
    g = make_grid(3, p.nx, p.ny, p.nz, bc=tuple(p.bc), s_l=p.s_l, reinit_iters=p.reinit_iters)
    assert (g.dx, g.reinit_dtau) == (p.dx, p.reinit_dtau)
    return g


CASES = [
    # (name, n, bc, sl_field, nsteps, reinit interval, ckpt, mode, tol)
    ("periodic_reinit", (20, 18, 16), (0, 0, 0), True, 7, 2, 3, 0, 1e-10),
    ("zero_grad_scalar_sl", (16, 17, 18), (1, 1, 1), False, 5, 0, 2, 0, 1e-10),
    ("mixed_bc_reinit", (18, 16, 20), (0, 1, 0), True, 6, 3, 6, 0, 1e-10),
    ("fast_forward", (20, 18, 16), (0, 0, 0), True, 7, 2, 3, 1, 1e-8),
    # two x segments of the stage adjoint (spills between them), ny % 4 == 3
    # (ghost rows inside the last 4-row tile), two z chunks
    ("two_segments_reinit", (70, 19, 13), (0, 1, 0), True, 5, 2, 3, 0, 1e-10),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_adjoint_dot_product_vs_tangent(ls, cuda, case):
    torch = cuda
    name, (nx, ny, nz), bc, slf, ns, ri, ck, mode, tol = case
    print("adjoint dot test", name, flush=True)
    p = ls.default_params(3, nx, nx=nx, ny=ny, nz=nz, bc=bc, s_l=0.05, reinit_interval=ri,
                          mode=mode)
    g = _grid(p)
    G0, u, v, w, sl = _fields(p, seed=11, sl_field=slf)
    G0 = Oracle().fill_ghosts(g, G0)
    Gt = G0 + 0.004 * np.cos(3 * np.arange(G0.size).reshape(G0.shape) / G0.size)
    dt, e = 0.2 * p.dx, 1.5 * p.dx
    d = lambda t: torch.from_numpy(t).cuda()
    s = ls.Solver(p)
    adj = ls.Adjoint(s, ns, ck)
    L, gr = adj.run(d(G0), d(u), d(v), d(w), d(Gt), e, dt, sl=d(sl) if slf else None)
    T = Tangent()
    rng = np.random.default_rng(5)
    I = g.interior
    for key in ("G0", "u", "v", "w") + (("sl",) if slf else ()):
        dirn = rng.standard_normal(G0.shape)
        _, Lt, dLt = T.run(g, G0, u, v, w, Gt, e, dt, ns, ri, sl=sl, **{"d" + key: dirn})
        ga = gr[key].cpu().numpy()
        dot = float(np.sum(ga[I] * dirn[I]))
        scale = float(np.sum(np.abs(ga[I] * dirn[I])))
        assert abs(dot - dLt) <= tol * scale, (name, key, dot, dLt, scale)
        assert abs(L - Lt) <= 1e-12 * Lt
    if not slf:
        # scalar s_l: the per-cell field sums to dL/ds_l
        _, _, dLs = T.run(g, G0, u, v, w, Gt, e, dt, ns, ri, sl=np.full(G0.shape, p.s_l),
                          dsl=np.ones(G0.shape))
        tot = float(gr["sl"].cpu().numpy()[I].sum())
        assert abs(tot - dLs) <= tol * max(abs(dLs), 1e-300) * 10
    # ghost cells of the field adjoints are never touched, the G0 adjoint is folded
    for key in ("u", "G0"):
        a = gr[key].cpu().numpy().copy()
        a[I] = 0.0
        assert not a.any(), key
    # loss-only sweep = the same loss
    L2, none = adj.run(d(G0), d(u), d(v), d(w), d(Gt), e, dt, sl=d(sl) if slf else None, grad=False)
    assert none is None and L2 == L
    # a second run reuses the context and gives the same gradient bit for bit
    L3, gr3 = adj.run(d(G0), d(u), d(v), d(w), d(Gt), e, dt, sl=d(sl) if slf else None)
    assert L3 == L and all(torch.equal(gr[k], gr3[k]) for k in gr)
    adj.close()
    s.close()


def test_adjoint_loss_matches_plain_solver(ls, cuda):
    torch = cuda
    p = ls.default_params(3, 24, s_l=0.05, reinit_interval=2)
    G0, u, v, w, sl = _fields(p, seed=3)
    d = lambda t: torch.from_numpy(t).cuda()
    s = ls.Solver(p)
    G0d = d(G0)
    s.fill_ghosts(G0d)
    dt = 0.2 * p.dx
    adj = ls.Adjoint(s, 9, 4)
    Gt = G0d + 0.002
    L, _ = adj.run(G0d, d(u), d(v), d(w), Gt, 1.5 * p.dx, dt, sl=d(sl), grad=False)
    G = G0d.clone()
    n, _, _ = s.advance(G, d(u), d(v), d(w), sl=d(sl), dt=dt, t_final=9.5 * dt, max_steps=9)
    assert n == 9
    assert ls.adjoint_loss(s, G, Gt, 1.5 * p.dx) == L
    I = p.interior
    Gh, Th = G.cpu().numpy()[I], Gt.cpu().numpy()[I]
    ref = np.sum((heaviside(Gh, 1.5 * p.dx) - heaviside(Th, 1.5 * p.dx)) ** 2)
    assert abs(L - ref) <= 1e-12 * ref


def test_adjoint_rejects_bad_input(ls, cuda):
    p2 = ls.default_params(2, 32)
    s2 = ls.Solver(p2)
    with pytest.raises(ls.LSError):
        ls.Adjoint(s2, 4, 2)
    p = ls.default_params(3, 16)
    s = ls.Solver(p)
    with pytest.raises(ls.LSError):
        ls.Adjoint(s, 0, 2)
    p3 = ls.default_params(3, 16, bc=(0, 2, 0))
    with pytest.raises(ls.LSError):
        ls.Adjoint(ls.Solver(p3), 4, 2)


def test_flame_maps_are_transposes(ls, cuda):
    """<fields(theta), f> = <theta, project(f)> for random theta, f; constant
    and affine coarse values are reproduced exactly; zero amplitudes give the
    swirl field of LS_CASE_SPHERE_SWIRL3D."""
    torch = cuda
    p = ls.default_params(3, 30, nx=30, ny=26, nz=22)
    nc, nm, seed = 5, 16, 77
    rng = np.random.default_rng(1)
    th_s = torch.tensor(rng.standard_normal(nc ** 3), device="cuda")
    th_a = torch.tensor(rng.standard_normal(nm), device="cuda")
    mk = lambda: torch.zeros(p.shape, dtype=torch.float64, device="cuda")
    sl, u, v, w = mk(), mk(), mk(), mk()
    ls.flame_fields(p, seed, th_s, th_a, sl, u, v, w)
    f = [torch.tensor(rng.standard_normal(p.shape), device="cuda") for _ in range(4)]
    # velocity map is affine (swirl offset): subtract the zero-amplitude field
    u0, v0, w0 = mk(), mk(), mk()
    ls.flame_fields(p, seed, None, torch.zeros(nm, dtype=torch.float64, device="cuda"), None, u0, v0, w0)
    I = p.interior
    gc, ga = ls.flame_project(p, seed, nc, nm, f[0], f[1], f[2], f[3])
    lhs_s = float((sl[I] * f[0][I]).sum())
    lhs_a = float(((u - u0)[I] * f[1][I] + (v - v0)[I] * f[2][I] + (w - w0)[I] * f[3][I]).sum())
    assert abs(lhs_s - float((th_s * gc).sum())) <= 1e-12 * float((sl[I] * f[0][I]).abs().sum())
    assert abs(lhs_a - float((th_a * ga).sum())) <= 1e-12 * abs(lhs_a) + 1e-9
    # constant and affine coarse values
    const = torch.full((nc ** 3,), 0.07, dtype=torch.float64, device="cuda")
    ls.flame_fields(p, seed, const, None, sl)
    assert float((sl[I] - 0.07).abs().max()) <= 1e-16
    Lx = (p.nx - 1) * p.dx
    xs = torch.arange(nc, dtype=torch.float64, device="cuda") * (Lx / (nc - 1))
    aff = (0.05 + 0.3 * xs).repeat(nc * nc)  # x fastest
    ls.flame_fields(p, seed, aff.contiguous(), None, sl)
    X = torch.arange(p.nx, dtype=torch.float64, device="cuda") * p.dx
    assert float((sl[I] - (0.05 + 0.3 * X)).abs().max()) <= 1e-15
    # swirl
    G, us, vs, ws = ls.init_case(p, ls.CASE_SPHERE_SWIRL3D)
    assert torch.equal(u0, us) and torch.equal(v0, vs) and torch.equal(w0, ws)


def _fd_check(ls, torch, prob, idx_s, idx_a, h=1e-6):
    L, gc, ga = prob.loss_and_grad(prob.sl_c, prob.amps_c)
    assert L > 0
    fd_s, fd_a = [], []
    for q in idx_s:
        sp, sm = prob.sl_c.clone(), prob.sl_c.clone()
        sp[q] += h
        sm[q] -= h
        fd_s.append((prob.loss_plain(sp, prob.amps_c) - prob.loss_plain(sm, prob.amps_c)) / (2 * h))
    for q in idx_a:
        ap, am = prob.amps_c.clone(), prob.amps_c.clone()
        ap[q] += h
        am[q] -= h
        fd_a.append((prob.loss_plain(prob.sl_c, ap) - prob.loss_plain(prob.sl_c, am)) / (2 * h))
    adj = np.concatenate([gc.cpu().numpy()[idx_s], ga.cpu().numpy()[idx_a]])
    fd = np.array(fd_s + fd_a)
    return L, adj, fd


def test_flame_gradient_vs_central_differences(ls, cuda):
    torch = cuda
    prob = ls.FlameInversion(n=32, nc=4, nmodes=6, nsteps=12, reinit_interval=5, ckpt=5)
    prob.sl_c, prob.amps_c = prob.draw(3)
    L, adj, fd = _fd_check(ls, torch, prob, list(range(64)), list(range(6)))
    err = np.abs(adj - fd).max() / np.abs(fd).max()
    print("flame 32^3: loss %.6e  max rel err vs FD %.3e" % (L, err), flush=True)
    assert err <= 1e-6, err
    # the adjoint's own forward sweep gives the plain solver's loss
    assert prob.loss(prob.sl_c, prob.amps_c) == prob.loss_plain(prob.sl_c, prob.amps_c)
    prob.close()


def test_flame_autograd_function(ls, cuda):
    torch = cuda
    prob = ls.FlameInversion(n=24, nc=3, nmodes=4, nsteps=6, reinit_interval=5, ckpt=3)
    sl_c, amps = prob.draw(4)
    a = sl_c.clone().requires_grad_(True)
    b = amps.clone().requires_grad_(True)
    loss = ls.flame_loss(a, b, prob)
    (3.0 * loss).backward()
    L, gc, ga = prob.loss_and_grad(sl_c, amps)
    assert float(loss.detach()) == L
    assert torch.equal(a.grad, 3.0 * gc) and torch.equal(b.grad, 3.0 * ga)
    with torch.no_grad():
        assert float(ls.flame_loss(sl_c, amps, prob)) == L
    # one gradient-descent step lowers the loss
    with torch.no_grad():
        lr = 0.1 * float(L) / float((gc * gc).sum() + (ga * ga).sum())
        L2 = prob.loss(sl_c - lr * gc, amps - lr * ga)
    assert L2 < L
    prob.close()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_flame_batch_bitwise_per_set(ls, cuda, mode):
    """ls_adjoint_run_sets through FlameBatch: 7 parameter sets over 3
    concurrent contexts (three waves, the last one partial) give for every set
    the loss and gradients of loss_and_grad run set by set, bit for bit."""
    torch = cuda
    m = ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT
    prob = ls.FlameInversion(n=32, nc=4, nmodes=6, nsteps=12, reinit_interval=5, ckpt=5, mode=m)
    sets = [prob.draw(200 + q) for q in range(7)]
    fb = ls.FlameBatch(prob, 3)
    got = fb.loss_and_grad(sets)
    for q, (sl_c, amps) in enumerate(sets):
        L, gc, ga = prob.loss_and_grad(sl_c, amps)
        assert got[q][0] == L, "set %d loss" % q
        assert torch.equal(got[q][1], gc) and torch.equal(got[q][2], ga), "set %d gradient" % q
    assert len({g[0] for g in got}) == len(sets)  # the sets differ
    # a second call on the same contexts reuses their workspace
    again = fb.loss_and_grad(sets[:2])
    assert again[0][0] == got[0][0] and torch.equal(again[1][1], got[1][1])
    fb.close()
    prob.close()


def test_adjoint_run_sets_rejects_shared_context(ls, cuda):
    import ctypes as C
    prob = ls.FlameInversion(n=16, nc=2, nmodes=2, nsteps=3, reinit_interval=5, ckpt=3)
    ctx = (C.c_void_p * 2)(prob.adj.h.value, prob.adj.h.value)
    sets = (ls.AdjointSet * 1)()
    assert ls.lib().ls_adjoint_run_sets(ctx, None, 2, sets, 1) == ls.LS_ERR_INVALID
    prob.close()


@pytest.mark.slow
def test_flame_full_size_gradient_vs_central_differences(ls, cuda):
    """The adjoint workload at full size: 128^3, 8^3 S_L nodes + 16 modes, 100
    steps, HCR-2 every 5, checkpoints every 10; a seeded sample of 12
    parameters against central differences (h = 1e-6)."""
    torch = cuda
    prob = ls.FlameInversion(n=128, nc=8, nmodes=16, nsteps=100, reinit_interval=5, ckpt=10)
    prob.sl_c, prob.amps_c = prob.draw(1)
    rng = np.random.default_rng(0)
    idx_s = sorted(rng.choice(512, 8, replace=False).tolist())
    L, adj, fd = _fd_check(ls, torch, prob, idx_s, [0, 5, 10, 15])
    err = np.abs(adj - fd).max() / np.abs(fd).max()
    print("flame 128^3: loss %.6e max rel err vs FD %.3e" % (L, err), flush=True)
    assert err <= 1e-6, err
    prob.close()
