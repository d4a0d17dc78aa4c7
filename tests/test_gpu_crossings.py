"""GPU parity of ls_hcr2_crossings, the drop-in for computeInterfaceCrossings
(Local_code/level-set_MPI_2D/include/reinitialization.h:75-130) and
computeInterfaceCrossings3D (Local_code/level-set_MPI_3D/include/reinitialization.h:46-89).

1. The golden vectors of the reference's own function
   (tests/golden/crossings_vectors.npz, make_crossings_golden.py), with the
   planted cells there: crossing sums just below and just above the 1e-15
   gate, centres of 0 and -0.0, a neighbour of exactly 0.  r~ and the bit
   mask are compared bit for bit; ghost entries must be left untouched.
2. Seeded fields over several tiles, on every boundary kind, against the oracle.
3. The shim's computeInterfaceCrossings3D as the CLI driver calls it inside
   reinitializeWithSwap3D: the arrays it leaves behind after the last
   reinitialisation are the oracle's crossings of the phi_0 it used.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import bumpy_sdf
from oracle.pyoracle import make_grid

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "crossings_vectors.npz")
SENTINEL_RT, SENTINEL_FLAG = 12345.5, -7


def _dev(cuda, a):
    return cuda.from_numpy(np.ascontiguousarray(a)).cuda()


def _gpu_crossings(ls, cuda, p, p0):
    s = ls.Solver(p)
    rt = cuda.full(p.shape, SENTINEL_RT, dtype=cuda.float64, device="cuda")
    fl = cuda.full(p.shape, SENTINEL_FLAG, dtype=cuda.int32, device="cuda")
    s.hcr2_crossings(_dev(cuda, p0), rt, fl)
    cuda.cuda.synchronize()
    out = rt.cpu().numpy(), fl.cpu().numpy()
    s.close()
    return out


def _check(p, rt, fl, want_rt, want_fl):
    I = p.interior
    assert np.array_equal(rt[I], want_rt[I]), "r_tilde differs"
    assert np.array_equal(fl[I], want_fl[I]), "flag differs"
    # the signs of zeros too: r~ = 0 is +0.0 in the reference
    assert np.array_equal(np.signbit(rt[I]), np.signbit(want_rt[I]))
    ghost = np.ones(p.shape, bool)
    ghost[I] = False
    assert (rt[ghost] == SENTINEL_RT).all() and (fl[ghost] == SENTINEL_FLAG).all(), "ghost entries written"


def _names():
    return sorted({k.split("/")[0] for k in np.load(GOLDEN).files})


@pytest.mark.parametrize("name", _names())
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_crossings_golden(ls, cuda, name, mode):
    z = np.load(GOLDEN)
    m = z[name + "/meta"]
    dim = int(m[0])
    bc = (int(m[4]), int(m[5]), int(m[6]))
    p = ls.default_params(dim, int(m[1]), nx=int(m[1]), ny=int(m[2]), nz=int(m[3]) if dim == 3 else 1, bc=bc,
                          mode=ls.MODE_FAST if mode == "fast" else ls.MODE_EXACT)
    rt, fl = _gpu_crossings(ls, cuda, p, z[name + "/p0"])
    _check(p, rt, fl, z[name + "/rt"], z[name + "/flag"])


@pytest.mark.parametrize("dim,n,bc", [
    (2, (130, 70, 1), (0, 1, 0)), (2, (130, 70, 1), (1, 1, 0)), (2, (97, 41, 1), (0, 2, 0)),
    (2, (97, 41, 1), (1, 2, 0)), (3, (70, 9, 40), (0, 0, 0)), (3, (70, 9, 40), (1, 1, 1)),
    (3, (66, 13, 17), (0, 1, 1)), (3, (65, 30, 8), (1, 0, 0))])
def test_crossings_vs_oracle(ls, oracle, cuda, dim, n, bc):
    g = make_grid(dim, *n, bc=bc)
    p = ls.default_params(dim, n[0], nx=n[0], ny=n[1], nz=n[2] if dim == 3 else 1, bc=bc)
    h = (g.dz, g.dy, g.dx) if dim == 3 else (g.dy, g.dx)
    for seed in (1, 2):
        p0 = oracle.fill_ghosts(g, bumpy_sdf(g.shape, h, 300 + seed, noise=0.03))
        # cells exactly on the front and tiny crossing sums around them
        flat = p0.reshape(-1)
        rng = np.random.default_rng(seed)
        pick = rng.choice(flat.size, 40, replace=False)
        flat[pick[:20]] = 0.0
        flat[pick[20:]] = rng.choice([-1, 1], 20) * rng.uniform(1e-17, 2e-15, 20)
        p0 = oracle.fill_ghosts(g, p0)
        want_rt, want_fl = oracle.crossings(g, p0)
        rt, fl = _gpu_crossings(ls, cuda, p, p0)
        _check(p, rt, fl, want_rt, want_fl)
        assert (want_fl != 0).sum() > 50


EXE = os.path.join(ROOT, "level-set-method-code_amd", "ls_solve")


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_shim_crossings_through_cli(ls, oracle, cuda, tmp_path, mode):
    """ls_solve runs the sphere-in-swirl case with HCR-2 every 4 steps through
    the shim (reinitializeWithSwap3D -> computeInterfaceCrossings3D) and dumps
    phi_0, r_tilde and flag of the last reinitialisation."""
    n = 24
    r = subprocess.run([EXE, "-dim", "3", "-n", str(n), "-t", "swirl", "-sl", "0.05", "-T", "0.06", "-reinit",
                        "-ri", "4", "-q", "-no-output", "-dump-crossings", "-mode", mode, "-o",
                        str(tmp_path / "o")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    g = make_grid(3, n, n, n, bc=(0, 0, 0))
    shape = g.shape
    p0 = np.fromfile(tmp_path / "o" / "phi_0.raw", np.float64).reshape(shape)
    rt = np.fromfile(tmp_path / "o" / "r_tilde.raw", np.float64).reshape(shape)
    fl = np.fromfile(tmp_path / "o" / "flag.raw", np.int32).reshape(shape)
    # phi_0 is the state after a ghost fill (reinitialization.h:333-336, 3D :211-213)
    assert np.array_equal(oracle.fill_ghosts(g, p0), p0)
    want_rt, want_fl = oracle.crossings(g, p0)
    I = g.interior
    assert np.array_equal(rt[I], want_rt[I]) and np.array_equal(fl[I], want_fl[I])
    assert (fl[I] != 0).sum() > 100
