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


def test_write_G_io_error(ls, tmp_path):
    p = ls.default_params(2, 9)
    with pytest.raises(ls.LSError) as e:
        ls.write_G(p, np.zeros(p.shape), tmp_path / "no_such_dir" / "x.bin")
    assert e.value.code == ls.LS_ERR_IO
