"""Python binding of the gfx950 G-equation level-set solver (ctypes over the C ABI).

The compute path is ``liblsamd.so`` (hand-written HIP kernels behind the
``extern "C"`` entry points declared in ``include/ls_api.h``).  This module only
marshals parameters and device pointers; PyTorch supplies device memory and
streams.  There is no CPU fallback: if the library is missing, importing a
solver raises.

The directory name is not a Python identifier, so callers load it by path::

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "lsamd", "<repo>/level-set-method-code_amd/__init__.py")
    lsamd = importlib.util.module_from_spec(spec); sys.modules["lsamd"] = lsamd
    spec.loader.exec_module(lsamd)
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# LSAMD_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("LSAMD_LIB", os.path.join(HERE, "liblsamd.so"))

LS_OK = 0
LS_ERR_INVALID, LS_ERR_NOMEM, LS_ERR_DEVICE, LS_ERR_IO, LS_ERR_NONFINITE = -1, -2, -3, -4, -5
BC_PERIODIC, BC_ZERO_GRAD, BC_LIN_EXTRAP = 0, 1, 2
MODE_EXACT, MODE_FAST = 0, 1
VEL_ARRAYS, VEL_DEFORM3D = 0, 1
CASE_CIRCLE2D, CASE_SPHERE_SWIRL3D, CASE_PLANE_MODES3D = 1, 2, 3
CASE_ZALESAK2D, CASE_SPHERE_DEFORM3D, CASE_FLAME2D = 4, 5, 6


class Params(C.Structure):
    """Mirror of ``ls_params_t`` (include/ls_api.h)."""
    _fields_ = [("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("nghost", C.c_int32), ("bc", C.c_int32 * 3), ("mode", C.c_int32),
                ("velocity", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("s_l", C.c_double), ("eps", C.c_double),
                ("reinit_interval", C.c_int32), ("reinit_iters", C.c_int32),
                ("reinit_dtau", C.c_double), ("reinit_beta", C.c_double),
                ("vel_period", C.c_double)]

    @property
    def shape(self):
        g = self.nghost
        if self.dim == 3:
            return (self.nz + 2 * g, self.ny + 2 * g, self.nx + 2 * g)
        return (self.ny + 2 * g, self.nx + 2 * g)

    @property
    def interior(self):
        g = self.nghost
        if self.dim == 3:
            return (slice(g, g + self.nz), slice(g, g + self.ny), slice(g, g + self.nx))
        return (slice(g, g + self.ny), slice(g, g + self.nx))

    @property
    def cells(self):
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)


class BatchMember(C.Structure):
    """Mirror of ``ls_batch_member_t``."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("bc", C.c_int32 * 2),
                ("dx", C.c_double), ("dy", C.c_double), ("s_l", C.c_double), ("dt", C.c_double)]


class LSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ls error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load liblsamd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("liblsamd.so not built (run __graft_entry__.build()): " + LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(Params)
    vp = C.c_void_p
    L.ls_default_params.argtypes = [P, C.c_int32, C.c_int32]
    L.ls_cfl_dt.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double]
    L.ls_cfl_dt.restype = C.c_double
    L.ls_create.argtypes = [P, C.POINTER(vp)]
    L.ls_destroy.argtypes = [vp]
    L.ls_field_size.argtypes = [P]
    L.ls_field_size.restype = C.c_int64
    L.ls_fill_ghosts.argtypes = [vp, vp, vp]
    L.ls_rk3_step.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_double, vp]
    L.ls_rk3_stage.argtypes = [vp, C.c_int32, vp, vp, vp, vp, vp, vp, vp, C.c_double, vp]
    L.ls_reinit_hcr2.argtypes = [vp, vp, vp]
    L.ls_hcr2_step.argtypes = [vp, vp, vp, vp, vp]
    L.ls_hcr2_crossings.argtypes = [vp, vp, vp, vp, vp]
    L.ls_advance.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int64, C.c_int64,
                             C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double), vp]
    L.ls_diag.argtypes = [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double), vp]
    L.ls_check_finite.argtypes = [vp, vp, vp]
    L.ls_write_G.argtypes = [P, vp, C.c_char_p]
    L.ls_init_case.argtypes = [P, C.c_int32, C.c_uint64, C.c_double, vp, vp, vp, vp, vp]
    L.ls_last_error.restype = C.c_char_p
    BM = C.POINTER(BatchMember)
    L.ls_batch_create.argtypes = [BM, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.ls_batch_destroy.argtypes = [vp]
    L.ls_batch_size.argtypes = [vp]
    L.ls_batch_size.restype = C.c_int64
    L.ls_batch_offset.argtypes = [vp, C.c_int32]
    L.ls_batch_offset.restype = C.c_int64
    L.ls_batch_fill_ghosts.argtypes = [vp, vp, vp]
    L.ls_batch_advance.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, vp]
    L.ls_batch_set_tile.argtypes = [vp, C.c_int32]
    L.ls_batch_get_tile.argtypes = [vp]
    L.ls_batch_get_tile.restype = C.c_int32
    L.ls_batch_autotune.argtypes = [vp, vp, vp, vp, C.c_int32, C.POINTER(C.c_double), vp]
    _lib = L
    return L


def check(rc):
    if rc != LS_OK:
        raise LSError(rc, lib().ls_last_error().decode())
    return rc


def default_params(dim, n, **kw):
    """Reference defaults (initDefaultParams) with keyword overrides.

    ``bc`` may be given as a 3-tuple; other keys are ``ls_params_t`` fields."""
    p = Params()
    check(lib().ls_default_params(C.byref(p), dim, n))
    for k, v in kw.items():
        if k == "bc":
            for a in range(3):
                p.bc[a] = v[a]
        else:
            setattr(p, k, v)
    if "dx" not in kw and ("nx" in kw or "ny" in kw or "nz" in kw):
        p.dx = 1.0 / (p.nx - 1)
        p.dy = 1.0 / (p.ny - 1)
        if dim == 3:
            p.dz = 1.0 / (p.nz - 1)
        h = min(p.dx, p.dy, p.dz) if dim == 3 else min(p.dx, p.dy)
        if "reinit_dtau" not in kw:
            p.reinit_dtau = 0.25 * h
    return p


def cfl_dt(p, cfl, u=0.0, v=0.0, w=0.0):
    return lib().ls_cfl_dt(C.byref(p), cfl, u, v, w)


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


class Solver:
    """Owns an ``ls_solver_t``; fields are float64 CUDA tensors of shape ``p.shape``."""

    def __init__(self, params):
        self.p = params
        self.h = C.c_void_p()
        check(lib().ls_create(C.byref(self.p), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ls_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, *ts):
        for t in ts:
            if t is None:
                continue
            if not t.is_cuda or t.dtype.itemsize != 8 or not t.is_contiguous():
                raise ValueError("fields must be contiguous float64 device tensors")
            if tuple(t.shape) != tuple(self.p.shape):
                raise ValueError("field shape %s != %s" % (tuple(t.shape), self.p.shape))

    def empty(self):
        import torch
        return torch.zeros(self.p.shape, dtype=torch.float64, device="cuda")

    def fill_ghosts(self, G, stream=None):
        self._chk(G)
        check(lib().ls_fill_ghosts(self.h, _ptr(G), _stream(stream)))

    def rk3_step(self, G_n, G_new, u, v, w=None, dt=1e-3, sl=None, stream=None):
        self._chk(G_n, G_new, u, v, w, sl)
        check(lib().ls_rk3_step(self.h, _ptr(G_n), _ptr(G_new), _ptr(u), _ptr(v), _ptr(w),
                                _ptr(sl), dt, _stream(stream)))

    def rk3_stage(self, k, G_s, G_n, out, u, v, w=None, dt=1e-3, sl=None, stream=None):
        self._chk(G_s, G_n, out, u, v, w, sl)
        check(lib().ls_rk3_stage(self.h, k, _ptr(G_s), _ptr(G_n), _ptr(out), _ptr(u), _ptr(v),
                                 _ptr(w), _ptr(sl), dt, _stream(stream)))

    def hcr2_step(self, phi_n, phi_0, phi_new, stream=None):
        self._chk(phi_n, phi_0, phi_new)
        check(lib().ls_hcr2_step(self.h, _ptr(phi_n), _ptr(phi_0), _ptr(phi_new), _stream(stream)))

    def hcr2_crossings(self, phi_0, r_tilde, flag, stream=None):
        self._chk(phi_0, r_tilde)
        if flag.dtype.itemsize != 4 or tuple(flag.shape) != tuple(self.p.shape):
            raise ValueError("flag must be an int32 field")
        check(lib().ls_hcr2_crossings(self.h, _ptr(phi_0), _ptr(r_tilde), _ptr(flag), _stream(stream)))

    def reinit(self, G, stream=None):
        self._chk(G)
        check(lib().ls_reinit_hcr2(self.h, _ptr(G), _stream(stream)))

    def advance(self, G, u, v, w=None, sl=None, dt=1e-3, t_final=1.0, step0=0, max_steps=None,
                t0=0.0, stream=None):
        """Reference time loop on G in place; returns (steps, t, dt_last)."""
        self._chk(G, u, v, w, sl)
        if max_steps is None:
            max_steps = int((t_final - t0) / dt) + 100
        t = C.c_double(t0)
        n = C.c_int64()
        dl = C.c_double()
        check(lib().ls_advance(self.h, _ptr(G), _ptr(u), _ptr(v), _ptr(w), _ptr(sl), dt, t_final,
                               step0, max_steps, C.byref(t), C.byref(n), C.byref(dl),
                               _stream(stream)))
        return n.value, t.value, dl.value

    def diag(self, G, G0=None, stream=None):
        """Returns (l2, area): RMS of G-G0 over the interior, area/volume of G<0."""
        self._chk(G, G0)
        l2 = C.c_double(float("nan"))
        area = C.c_double()
        check(lib().ls_diag(self.h, _ptr(G), _ptr(G0), C.byref(l2), C.byref(area), _stream(stream)))
        return l2.value, area.value

    def check_finite(self, G, stream=None):
        self._chk(G)
        return lib().ls_check_finite(self.h, _ptr(G), _stream(stream))


def init_case(p, case_id, seed=0, period=1.0, stream=None):
    """Device-generated inputs: returns (G, u, v, w) tensors (w is None in 2D)."""
    import torch
    mk = lambda: torch.zeros(p.shape, dtype=torch.float64, device="cuda")
    G, u, v = mk(), mk(), mk()
    w = mk() if p.dim == 3 else None
    check(lib().ls_init_case(C.byref(p), case_id, seed, period, _ptr(G), _ptr(u), _ptr(v), _ptr(w),
                             _stream(stream)))
    return G, u, v, w


def write_G(p, host_array, path):
    import numpy as np
    a = np.ascontiguousarray(host_array, dtype=np.float64)
    check(lib().ls_write_G(C.byref(p), C.c_void_p(a.ctypes.data), str(path).encode()))


class Batch:
    """Batched engine over packed float64 device buffers (see ls_api.h)."""

    def __init__(self, members, mode=MODE_EXACT):
        arr = (BatchMember * len(members))()
        for q, m in enumerate(members):
            arr[q].nx, arr[q].ny = m["nx"], m["ny"]
            arr[q].bc[0], arr[q].bc[1] = m.get("bc", (BC_PERIODIC, BC_ZERO_GRAD))
            arr[q].dx, arr[q].dy = m["dx"], m["dy"]
            arr[q].s_l, arr[q].dt = m["s_l"], m["dt"]
        self.members = members
        self.h = C.c_void_p()
        check(lib().ls_batch_create(arr, len(members), mode, C.byref(self.h)))
        self.size = lib().ls_batch_size(self.h)
        self.offsets = [lib().ls_batch_offset(self.h, q) for q in range(len(members))]

    def close(self):
        if self.h:
            lib().ls_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def empty(self):
        import torch
        return torch.zeros(self.size, dtype=torch.float64, device="cuda")

    def view(self, packed, q):
        """Member q of a packed buffer as a (ny+6, nx+6) view."""
        m = self.members[q]
        n = (m["ny"] + 6) * (m["nx"] + 6)
        o = self.offsets[q]
        return packed[o:o + n].view(m["ny"] + 6, m["nx"] + 6)

    def fill_ghosts(self, G, stream=None):
        check(lib().ls_batch_fill_ghosts(self.h, _ptr(G), _stream(stream)))

    def advance(self, G, u, v, nsteps, graph=True, stream=None):
        check(lib().ls_batch_advance(self.h, _ptr(G), _ptr(u), _ptr(v), nsteps, 1 if graph else 0,
                                     _stream(stream)))

    def set_tile(self, shape):
        check(lib().ls_batch_set_tile(self.h, shape))

    def tile(self):
        return lib().ls_batch_get_tile(self.h)

    def autotune(self, G, u, v, probe_steps=3, stream=None):
        ms = C.c_double()
        check(lib().ls_batch_autotune(self.h, _ptr(G), _ptr(u), _ptr(v), probe_steps, C.byref(ms),
                                      _stream(stream)))
        return self.tile(), ms.value
