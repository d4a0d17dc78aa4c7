"""CPU checks of the C ABI (no GPU needed): the library loads, exports every
entry point include/ls_api.h declares, validates parameters before touching a
device, and its host-only functions match the reference formulas."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle.pyoracle import make_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(ls\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol(ls):
    names = declared_symbols()
    assert len(names) >= 15, names
    L = ls.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_struct_layout(ls):
    # ls_params_t: 10 int32 (40 B) + 6 doubles + 2 doubles + 2 int32 + 3 doubles = 136 B
    assert C.sizeof(ls.Params) == 136


def test_default_params_match_reference(ls):
    p = ls.default_params(2, 201)
    # config.h:43-44 (dx = L/(N-1)), :205 (dtau = 0.25 min(dx,dy)), :84, :90, :93
    assert p.dx == (1.0 - 0.0) / (201 - 1) and p.reinit_dtau == 0.25 * p.dx
    assert (p.reinit_iters, p.reinit_beta, p.eps, p.nghost) == (2, 0.5, 1e-10, 3)
    assert tuple(p.bc) == (ls.BC_PERIODIC, ls.BC_ZERO_GRAD, ls.BC_PERIODIC)
    p3 = ls.default_params(3, 64)
    assert tuple(p3.bc) == (0, 0, 0) and p3.dz == 1.0 / 63


def test_cfl_dt_matches_reference(ls):
    p = ls.default_params(2, 201, s_l=0.1)
    # 2D calculateTimeStep: cfl*min(dx,dy)/(|u|+|v|+S_L), floor eps (config.h:251-263)
    assert ls.cfl_dt(p, 0.5, 0.5, -0.2) == 0.5 * p.dx / (0.5 + 0.2 + 0.1)
    q = ls.default_params(2, 201, s_l=0.0)
    assert ls.cfl_dt(q, 0.2, 0.0, 0.0) == 0.2 * q.dx / 1e-10
    # 3D: floor 2.0 for the deformation test (config.h:272-293)
    r = ls.default_params(3, 64, s_l=0.0)
    assert ls.cfl_dt(r, 0.2, 0.0, 0.0, 0.0) == 0.2 * r.dx / 2.0


@pytest.mark.parametrize("bad", [
    dict(nghost=2), dict(dim=4), dict(nx=2), dict(dx=0.0), dict(mode=7),
    dict(bc=(2, 0, 0)), dict(bc=(0, 0, 5)), dict(velocity=1), dict(reinit_iters=-1)])
def test_create_rejects_bad_params_without_gpu(ls, bad):
    p = ls.default_params(2, 16)
    for k, v in bad.items():
        if k == "bc":
            for a in range(3):
                p.bc[a] = v[a]
        else:
            setattr(p, k, v)
    h = C.c_void_p()
    rc = ls.lib().ls_create(C.byref(p), C.byref(h))
    assert rc == ls.LS_ERR_INVALID and not h.value
    assert len(ls.lib().ls_last_error()) > 0


def test_null_arguments_rejected(ls):
    L = ls.lib()
    assert L.ls_fill_ghosts(None, None, None) == ls.LS_ERR_INVALID
    assert L.ls_reinit_hcr2(None, None, None) == ls.LS_ERR_INVALID
    assert L.ls_create(None, None) == ls.LS_ERR_INVALID


@pytest.mark.parametrize("dim", [2, 3])
def test_write_G_matches_oracle_bytes(ls, oracle, tmp_path, dim):
    n = (9, 7, 5)
    p = ls.default_params(dim, 9, nx=n[0], ny=n[1], nz=n[2] if dim == 3 else 1)
    g = make_grid(dim, *n, bc=tuple(p.bc))
    G = np.random.default_rng(dim).standard_normal(g.shape)
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    ls.write_G(p, G, a)
    oracle.write_G(g, G, b)
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("dim", [2, 3])
def test_write_G_vtk_matches_reference(ls, oracle, tmp_path, dim):
    """saveFieldVTK / saveFieldVTK3D: byte-equal to the restatement and, where
    oracle/_ref is built, to the reference's own writer (origin 0 there)."""
    from oracle.pyoracle import Reference, ref_available
    n = (9, 7, 5)
    p = ls.default_params(dim, 9, nx=n[0], ny=n[1], nz=n[2] if dim == 3 else 1)
    p.x0 = p.y0 = p.z0 = 0.0
    g = make_grid(dim, *n, bc=tuple(p.bc))
    assert (g.dx, g.dy) == (p.dx, p.dy)
    G = np.random.default_rng(10 + dim).standard_normal(g.shape)
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    ls.write_G_vtk(p, G, a)
    oracle.write_vtk(g, G, b)
    assert a.read_bytes() == b.read_bytes()
    if ref_available(dim):
        c = tmp_path / "c.vtk"
        assert Reference(dim).write_vtk(g, G, c) == 0
        assert a.read_bytes() == c.read_bytes()
    lines = a.read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[4] == ("DIMENSIONS 9 7 5" if dim == 3 else "DIMENSIONS 9 7 1")
    vals = np.array([float(x) for x in lines[10:]])
    assert np.allclose(vals, G[g.interior].ravel(), rtol=1e-6, atol=0)


def test_write_G_vtk_origin_and_errors(ls, tmp_path):
    p = ls.default_params(3, 6)
    p.x0, p.y0, p.z0 = -1.5, 0.25, 2.0
    a = tmp_path / "o.vtk"
    ls.write_G_vtk(p, np.zeros(p.shape), a)
    assert a.read_text().splitlines()[5] == "ORIGIN -1.500000 0.250000 2.000000"
    with pytest.raises(ls.LSError) as e:
        ls.write_G_vtk(p, np.zeros(p.shape), tmp_path / "no_such_dir" / "x.vtk")
    assert e.value.code == ls.LS_ERR_IO
    with pytest.raises(ValueError):
        ls.write_G_vtk(p, np.zeros((3, 3)), a)


def test_write_G_io_error(ls, tmp_path):
    p = ls.default_params(2, 9)
    with pytest.raises(ls.LSError) as e:
        ls.write_G(p, np.zeros(p.shape), tmp_path / "no_such_dir" / "x.bin")
    assert e.value.code == ls.LS_ERR_IO
