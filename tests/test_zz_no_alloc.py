"""ls_diag_noalloc, the allocation-free variant of ls_diag (no solver handle,
static device scratch, results in device memory, no synchronisation), called
through the C ABI and checked against ls_diag / ls_check_finite on the same
fields: l2, area and the non-finite count are the same bits.  Two calls in a
row share the GPU's scratch array; the second, on a smaller grid, must not see
anything of the first."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ls, cuda, p, G, G0, out):
    rc = ls.lib().ls_diag_noalloc(C.byref(p), C.c_void_p(G.data_ptr()),
                                  None if G0 is None else C.c_void_p(G0.data_ptr()),
                                  C.c_void_p(out.data_ptr()),
                                  C.c_void_p(cuda.cuda.current_stream().cuda_stream))
    assert rc == ls.LS_OK, ls.lib().ls_last_error()
    cuda.cuda.synchronize()
    return out.cpu().numpy().copy()


def _fields(cuda, p, seed, nan=0, inf=0):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal(p.shape)
    G0 = G + 1e-3 * rng.standard_normal(p.shape)
    I = p.interior
    sub = G[I]
    flat = sub.reshape(-1)
    pick = rng.choice(flat.size, nan + inf, replace=False)
    flat[pick[:nan]] = np.nan
    flat[pick[nan:]] = np.inf
    G[I] = flat.reshape(sub.shape)
    return cuda.from_numpy(G).cuda(), cuda.from_numpy(G0).cuda()


def test_diag_noalloc_matches_diag_two_calls(ls, cuda):
    out = cuda.full((3,), -1.0, dtype=cuda.float64, device="cuda")
    # first call: a 3D grid of several tiles, with and without G0
    p = ls.default_params(3, 70, nx=70, ny=33, nz=40)
    G, G0 = _fields(cuda, p, 1)
    s = ls.Solver(p)
    l2, area = s.diag(G, G0)
    got = _noalloc(ls, cuda, p, G, G0, out)
    assert got[0] == l2 and got[1] == area and got[2] == 0.0
    assert s.check_finite(G) == ls.LS_OK
    got = _noalloc(ls, cuda, p, G, None, out)
    _, area_only = s.diag(G)
    assert got[0] == 0.0 and got[1] == area_only and got[2] == 0.0
    s.close()
    # second call: a smaller 2D grid with planted non-finite values
    q = ls.default_params(2, 50, nx=50, ny=37)
    H, H0 = _fields(cuda, q, 2, nan=3, inf=2)
    s2 = ls.Solver(q)
    _, area2 = s2.diag(H)
    got2 = _noalloc(ls, cuda, q, H, None, out)
    assert got2[1] == area2 and got2[2] == 5.0
    assert s2.check_finite(H) == ls.LS_ERR_NONFINITE
    # and a finite small field with G0: l2 bitwise again
    F, F0 = _fields(cuda, q, 3)
    l2f, areaf = s2.diag(F, F0)
    got3 = _noalloc(ls, cuda, q, F, F0, out)
    assert got3[0] == l2f and got3[1] == areaf and got3[2] == 0.0
    s2.close()
    # the Python wrapper reaches the same entry point
    w = ls.diag_noalloc(q, F, F0)
    assert np.array_equal(w.cpu().numpy(), got3)


def test_diag_noalloc_rejects_bad_input(ls, cuda):
    p = ls.default_params(2, 16)
    out = cuda.zeros(3, dtype=cuda.float64, device="cuda")
    G = cuda.zeros(p.shape, dtype=cuda.float64, device="cuda")
    assert ls.lib().ls_diag_noalloc(C.byref(p), None, None, C.c_void_p(out.data_ptr()), None) == ls.LS_ERR_INVALID
    p.nghost = 2
    assert ls.lib().ls_diag_noalloc(C.byref(p), C.c_void_p(G.data_ptr()), None, C.c_void_p(out.data_ptr()),
                                    None) == ls.LS_ERR_INVALID
