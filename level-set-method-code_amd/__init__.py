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


class AdjointSet(C.Structure):
    """Mirror of ``ls_adjoint_set_t``."""
    _fields_ = [("G0", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("sl_field", C.c_void_p), ("G_target", C.c_void_p), ("eps_h", C.c_double), ("dt", C.c_double),
                ("loss", C.POINTER(C.c_double)), ("gG0", C.c_void_p), ("gu", C.c_void_p), ("gv", C.c_void_p),
                ("gw", C.c_void_p), ("gsl", C.c_void_p)]


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
    L.ls_div_check.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, vp]
    L.ls_diag_noalloc.argtypes = [P, vp, vp, vp, vp]
    L.ls_write_G.argtypes = [P, vp, C.c_char_p]
    L.ls_write_G_vtk.argtypes = [P, vp, C.c_char_p]
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
    L.ls_adjoint_create.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.ls_adjoint_destroy.argtypes = [vp]
    L.ls_adjoint_workspace_bytes.argtypes = [vp]
    L.ls_adjoint_workspace_bytes.restype = C.c_int64
    L.ls_adjoint_loss.argtypes = [vp, vp, vp, C.c_double, C.POINTER(C.c_double), vp]
    L.ls_adjoint_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_double,
                                 C.POINTER(C.c_double), vp, vp, vp, vp, vp, vp]
    L.ls_adjoint_run_sets.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int32, C.POINTER(AdjointSet), C.c_int32]
    L.ls_flame_fields.argtypes = [P, C.c_uint64, C.c_int32, vp, C.c_int32, vp, vp, vp, vp, vp, vp]
    L.ls_flame_project.argtypes = [P, C.c_uint64, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
    L.ls_sample_field_size.argtypes = [C.c_int32, C.c_int32]
    L.ls_sample_field_size.restype = C.c_int64
    L.ls_sample_create.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.POINTER(vp)]
    L.ls_sample_destroy.argtypes = [vp]
    L.ls_sample_rhs.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp]
    L.ls_sample_rk3_step.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, vp, vp]
    L.ls_sample_apply_periodic.argtypes = [vp, vp, vp]
    L.ls_sample_apply_anchored_bc.argtypes = [vp, vp, vp, vp]
    L.ls_sample_reinitialize.argtypes = [vp, vp, vp, C.c_int32, vp]
    L.ls_sample_first_order_upwind.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, vp]
    L.ls_sample_bunsen.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, vp, vp]
    L.ls_sample_forced_velocity.argtypes = [vp, C.c_double, C.c_int32, C.c_double, C.c_double, C.c_double,
                                            C.c_double, vp, vp, vp]
    L.ls_sample_advance.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_int64,
                                    C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), vp]
    L.ls_sample_write_G.argtypes = [C.c_int32, C.c_int32, vp, C.c_char_p]
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


def diag_noalloc(p, G, G0=None, out=None, stream=None):
    """ls_diag_noalloc: (l2, area, non-finite count) of G as a 3-element
    float64 device tensor, stream-ordered, no solver handle, no allocation in
    the library (``out`` may be preallocated by the caller)."""
    import torch
    if out is None:
        out = torch.empty(3, dtype=torch.float64, device="cuda")
    for t in (G, G0):
        if t is not None and (not t.is_cuda or t.dtype != torch.float64 or tuple(t.shape) != tuple(p.shape)
                              or not t.is_contiguous()):
            raise ValueError("fields must be contiguous float64 device tensors of shape %s" % (p.shape,))
    if not out.is_cuda or out.dtype != torch.float64 or out.numel() < 3 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 device tensor of 3 elements")
    check(lib().ls_diag_noalloc(C.byref(p), _ptr(G), _ptr(G0), _ptr(out), _stream(stream)))
    return out


def div_check(a, b, kind, stream=None):
    """ls_div_check: the device sequence ``kind`` (0 rdiv_u, 1 qdiv_u on the
    reciprocal rdiv_u(1, b), 2 div6_u; 3 sqrt_u and 4 sqrt_u(max(a, 2^-767)) - 1,
    the fast HCR-2's square root; kinds 2-4 ignore b) on contiguous float64
    device tensors; returns the results as a new device tensor."""
    import torch
    for t in (a, b):
        if t is not None and (not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous()):
            raise ValueError("operands must be contiguous float64 device tensors")
    if b is not None and b.numel() != a.numel():
        raise ValueError("a and b differ in length")
    out = torch.empty_like(a)
    check(lib().ls_div_check(_ptr(a), _ptr(b), _ptr(out), a.numel(), kind, _stream(stream)))
    return out


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


def write_G_vtk(p, host_array, path):
    """Legacy ASCII VTK of the interior (the reference's saveFieldVTK[3D])."""
    import numpy as np
    a = np.ascontiguousarray(host_array, dtype=np.float64)
    if tuple(a.shape) != tuple(p.shape):
        raise ValueError("field shape %s != %s" % (tuple(a.shape), tuple(p.shape)))
    check(lib().ls_write_G_vtk(C.byref(p), C.c_void_p(a.ctypes.data), str(path).encode()))


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

    def autotune(self, G, u, v, probe_steps=10, stream=None):
        ms = C.c_double()
        check(lib().ls_batch_autotune(self.h, _ptr(G), _ptr(u), _ptr(v), probe_steps, C.byref(ms),
                                      _stream(stream)))
        return self.tile(), ms.value


ENSEMBLE_SIZES = (128, 192, 256, 384, 512)


def ensemble(count=96, seed=4, mode=MODE_FAST):
    """The batched-ensemble workload: ``count`` 2D flames with seeded edge
    lengths from ENSEMBLE_SIZES, S_L ~ U[0.05, 0.2], zero-gradient edges and
    the seeded circle/plane G0 and uniform + 8-mode velocity of
    LS_CASE_FLAME2D (member q uses seed 1000*seed + q); dt from CFL 0.5 on
    the member's largest |u| + |v|.  Returns (members, params, fields) with
    fields[q] = (G, u, v) padded device tensors as generated (ghost cells
    hold the analytic fields; fill G's ghosts before stepping)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    members, params, fields = [], [], []
    for q in range(count):
        n = int(ENSEMBLE_SIZES[rng.integers(0, len(ENSEMBLE_SIZES))])
        s_l = float(rng.uniform(0.05, 0.2))
        p = default_params(2, n, s_l=s_l, bc=(BC_ZERO_GRAD, BC_ZERO_GRAD, 0), reinit_interval=0, mode=mode)
        params.append(p)
        members.append(dict(nx=n, ny=n, bc=(BC_ZERO_GRAD, BC_ZERO_GRAD), dx=p.dx, dy=p.dy, s_l=s_l, dt=0.0))
    for q, p in enumerate(params):
        G, u, v, _ = init_case(p, CASE_FLAME2D, seed=1000 * seed + q)
        speed = (u.abs() + v.abs()).max().item()
        members[q]["dt"] = 0.5 * min(p.dx, p.dy) / (speed + p.s_l)
        fields.append((G, u, v))
    torch.cuda.synchronize()
    return members, params, fields


# ---------------------------------------------------------------------------
# Discrete adjoint (3D)


class Adjoint:
    """Owns an ``ls_adjoint_t`` bound to a 3D :class:`Solver`: loss
    sum (H_e(G_T) - H_e(G_target))^2 after ``nsteps`` RK3 steps (HCR-2 as in
    the solver's params) and its gradient, checkpointed every ``ckpt`` steps."""

    def __init__(self, solver, nsteps, ckpt=10):
        self.solver = solver
        self.nsteps = nsteps
        self.h = C.c_void_p()
        check(lib().ls_adjoint_create(solver.h, nsteps, ckpt, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ls_adjoint_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def workspace_bytes(self):
        return lib().ls_adjoint_workspace_bytes(self.h)

    def run(self, G0, u, v, w, G_target, eps_h, dt, sl=None, grad=True, out=None, stream=None):
        """Returns (loss, grads); grads maps G0, u, v, w, sl to device fields
        (None when grad is False).  ``out`` may supply the gradient tensors."""
        s = self.solver
        s._chk(G0, u, v, w, sl, G_target)
        loss = C.c_double()
        g = None
        if grad:
            g = out if out is not None else {k: s.empty() for k in ("G0", "u", "v", "w", "sl")}
            s._chk(*g.values())
        gp = (lambda k: _ptr(g[k]) if g is not None else None)
        check(lib().ls_adjoint_run(self.h, _ptr(G0), _ptr(u), _ptr(v), _ptr(w), _ptr(sl),
                                   _ptr(G_target), eps_h, dt, C.byref(loss), gp("G0"), gp("u"),
                                   gp("v"), gp("w"), gp("sl"), _stream(stream)))
        return loss.value, g


def adjoint_loss(solver, G, G_target, eps_h, stream=None):
    """The adjoint's loss for a given final field (no stepping)."""
    solver._chk(G, G_target)
    loss = C.c_double()
    check(lib().ls_adjoint_loss(solver.h, _ptr(G), _ptr(G_target), eps_h, C.byref(loss),
                                _stream(stream)))
    return loss.value


def flame_fields(p, mode_seed, sl_coarse=None, amps=None, sl=None, u=None, v=None, w=None,
                 stream=None):
    """S_L field from nc^3 coarse values (trilinear) and/or the velocity
    swirl + sum amps[q] * mode_q, written into the given padded tensors."""
    nc = round(sl_coarse.numel() ** (1.0 / 3.0)) if sl_coarse is not None else 2
    if sl_coarse is not None and nc ** 3 != sl_coarse.numel():
        raise ValueError("sl_coarse must hold nc^3 values")
    nm = amps.numel() if amps is not None else 0
    check(lib().ls_flame_fields(C.byref(p), mode_seed, nc, _ptr(sl_coarse), nm, _ptr(amps),
                                _ptr(sl), _ptr(u), _ptr(v), _ptr(w), _stream(stream)))


def flame_project(p, mode_seed, nc, nmodes, gsl, gu, gv, gw, stream=None):
    """Transposed parameter maps: (dL/d sl_coarse [nc^3], dL/d amps [nmodes])."""
    import torch
    gc = torch.zeros(nc ** 3, dtype=torch.float64, device="cuda")
    ga = torch.zeros(max(nmodes, 1), dtype=torch.float64, device="cuda")
    check(lib().ls_flame_project(C.byref(p), mode_seed, nc, nmodes, _ptr(gsl), _ptr(gu), _ptr(gv),
                                 _ptr(gw), _ptr(gc), _ptr(ga), _stream(stream)))
    return gc, ga[:nmodes]


class FlameInversion:
    """Front-matching problem of the adjoint workload: a sphere flame
    (r = 0.2 at the centre) on an n^3 grid in the swirl field plus
    ``nmodes`` seeded solenoidal modes; parameters are S_L on an nc^3 coarse
    grid (trilinear) and the mode amplitudes.  The target front comes from a
    forward run with ``theta_true``; dt is fixed from the true fields' maxima
    (CFL ``cfl``) so that the loss is a function of the parameters only."""

    def __init__(self, n=128, nc=8, nmodes=16, nsteps=100, reinit_interval=5, ckpt=10,
                 mode=MODE_EXACT, mode_seed=2024, true_seed=0, cfl=0.5, eps_factor=1.5):
        import torch
        self.p = default_params(3, n, reinit_interval=reinit_interval, mode=mode)
        self.n, self.nc, self.nmodes, self.nsteps = n, nc, nmodes, nsteps
        self.mode_seed = mode_seed
        self.eps_h = eps_factor * self.p.dx
        self.solver = Solver(self.p)
        self.ckpt = ckpt
        self.adj = Adjoint(self.solver, nsteps, ckpt)
        G0, _, _, _ = init_case(self.p, CASE_SPHERE_SWIRL3D)
        self.solver.fill_ghosts(G0)
        self.G0 = G0
        self.sl_true, self.amps_true = self.draw(true_seed)
        self.sl, self.u, self.v, self.w = (self.solver.empty() for _ in range(4))
        self.fields(self.sl_true, self.amps_true)
        I = self.p.interior
        vmax = sum(float(t[I].abs().max()) for t in (self.u, self.v, self.w))
        self.dt = cfl * self.p.dx / (vmax + float(self.sl_true.max()))
        self.G_target = self.forward(self.sl_true, self.amps_true)
        self.grads = {k: self.solver.empty() for k in ("G0", "u", "v", "w", "sl")}
        torch.cuda.synchronize()

    def draw(self, seed):
        """Seeded parameter set: S_L nodes ~ U[0.04, 0.12], amplitudes ~ U[-0.1, 0.1]."""
        import numpy as np
        import torch
        rng = np.random.default_rng(seed)
        sl = torch.tensor(rng.uniform(0.04, 0.12, self.nc ** 3), dtype=torch.float64, device="cuda")
        amps = torch.tensor(rng.uniform(-0.1, 0.1, self.nmodes), dtype=torch.float64, device="cuda")
        return sl, amps

    def fields(self, sl_coarse, amps):
        flame_fields(self.p, self.mode_seed, sl_coarse.contiguous(), amps.contiguous(), self.sl,
                     self.u, self.v, self.w)

    def forward(self, sl_coarse, amps):
        """G_T from the plain solver loop (ls_advance)."""
        self.fields(sl_coarse, amps)
        G = self.G0.clone()
        steps, _, _ = self.solver.advance(G, self.u, self.v, self.w, sl=self.sl, dt=self.dt,
                                          t_final=(self.nsteps + 0.5) * self.dt,
                                          max_steps=self.nsteps)
        assert steps == self.nsteps
        return G

    def loss_plain(self, sl_coarse, amps):
        """Loss through the plain forward solver: the finite-difference path."""
        return adjoint_loss(self.solver, self.forward(sl_coarse, amps), self.G_target, self.eps_h)

    def loss(self, sl_coarse, amps):
        """Loss through the adjoint's forward sweep (no reverse sweep)."""
        self.fields(sl_coarse, amps)
        L, _ = self.adj.run(self.G0, self.u, self.v, self.w, self.G_target, self.eps_h, self.dt,
                            sl=self.sl, grad=False)
        return L

    def loss_and_grad(self, sl_coarse, amps):
        """(loss, dL/d sl_coarse, dL/d amps)."""
        self.fields(sl_coarse, amps)
        L, g = self.adj.run(self.G0, self.u, self.v, self.w, self.G_target, self.eps_h, self.dt,
                            sl=self.sl, grad=True, out=self.grads)
        gc, ga = flame_project(self.p, self.mode_seed, self.nc, self.nmodes, g["sl"], g["u"],
                               g["v"], g["w"])
        return L, gc, ga

    def close(self):
        self.adj.close()
        self.solver.close()


class FlameBatch:
    """Gradients of a :class:`FlameInversion` for many parameter sets at once.
    ``nctx`` workers, each with its own solver, adjoint context, parameter
    fields, gradient fields and stream, take the sets in waves of ``nctx``:
    the maps to fields on each worker's stream, then one ls_adjoint_run_sets
    call (one host thread per context inside the library, so the sets of a
    wave run concurrently on the GPU), then the transposed maps.  Results are
    those of ``problem.loss_and_grad`` set by set."""

    def __init__(self, problem, nctx=4):
        import torch
        self.prob = problem
        self.workers = []
        for _ in range(nctx):
            s = Solver(problem.p)
            w = dict(solver=s, adj=Adjoint(s, problem.nsteps, problem.ckpt), stream=torch.cuda.Stream(),
                     fields=[s.empty() for _ in range(4)],
                     grads={k: s.empty() for k in ("u", "v", "w", "sl")}, loss=C.c_double())
            self.workers.append(w)
        self.ctx = (C.c_void_p * nctx)(*[w["adj"].h.value for w in self.workers])
        self.streams = (C.c_void_p * nctx)(*[w["stream"].cuda_stream for w in self.workers])

    def loss_and_grad(self, sets):
        """sets: list of (sl_coarse, amps).  Returns a list of (loss, dL/d sl_coarse, dL/d amps)."""
        import torch
        P = self.prob
        nctx = len(self.workers)
        out = []
        for w0 in range(0, len(sets), nctx):
            wave = sets[w0:w0 + nctx]
            arr = (AdjointSet * len(wave))()
            main = torch.cuda.current_stream()
            for c, (sl_c, amps) in enumerate(wave):
                wk = self.workers[c]
                wk["stream"].wait_stream(main)
                sl, u, v, w = wk["fields"]
                with torch.cuda.stream(wk["stream"]):
                    flame_fields(P.p, P.mode_seed, sl_c.contiguous(), amps.contiguous(), sl, u, v, w,
                                 stream=wk["stream"].cuda_stream)
                g = wk["grads"]
                arr[c] = AdjointSet(P.G0.data_ptr(), u.data_ptr(), v.data_ptr(), w.data_ptr(), sl.data_ptr(),
                                    P.G_target.data_ptr(), P.eps_h, P.dt, C.pointer(wk["loss"]), None,
                                    g["u"].data_ptr(), g["v"].data_ptr(), g["w"].data_ptr(), g["sl"].data_ptr())
            check(lib().ls_adjoint_run_sets(self.ctx, self.streams, nctx, arr, len(wave)))
            for c in range(len(wave)):
                wk = self.workers[c]
                g = wk["grads"]
                with torch.cuda.stream(wk["stream"]):
                    gc, ga = flame_project(P.p, P.mode_seed, P.nc, P.nmodes, g["sl"], g["u"], g["v"], g["w"],
                                           stream=wk["stream"].cuda_stream)
                main.wait_stream(wk["stream"])
                out.append((wk["loss"].value, gc, ga))
        return out

    def close(self):
        for w in self.workers:
            w["adj"].close()
            w["solver"].close()
        self.workers = []


def _flame_loss_function():
    import torch

    class FlameLoss(torch.autograd.Function):
        """loss = FlameLoss.apply(sl_coarse, amps, problem): forward and
        reverse sweeps run together in ``forward`` when a gradient is
        needed; ``backward`` scales the stored parameter gradients."""

        @staticmethod
        def forward(ctx, sl_coarse, amps, problem):
            if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
                L, gc, ga = problem.loss_and_grad(sl_coarse.detach(), amps.detach())
                ctx.save_for_backward(gc.view_as(sl_coarse), ga.view_as(amps))
            else:
                L = problem.loss(sl_coarse.detach(), amps.detach())
            return torch.tensor(L, dtype=torch.float64, device=sl_coarse.device)

        @staticmethod
        def backward(ctx, gL):
            gc, ga = ctx.saved_tensors
            return gL * gc, gL * ga, None

    return FlameLoss


_FLAME_LOSS = None


def flame_loss(sl_coarse, amps, problem):
    """Differentiable loss of a :class:`FlameInversion` (torch.autograd)."""
    global _FLAME_LOSS
    if _FLAME_LOSS is None:
        _FLAME_LOSS = _flame_loss_function()
    return _FLAME_LOSS.apply(sl_coarse, amps, problem)


# ---------------------------------------------------------------------------
# The sample/ scheme (2D, local Lax-Friedrichs)


class Sample:
    """Owns an ``ls_sample_t``: the reference's ``sample/`` solver on float64
    device tensors of shape (ny+6, nx+6).  Defaults are sample/config.hpp's
    (800 x 800 on a 0.02 m box, dt 2.5e-6 s, S_L 0.4, v0 0.8, 5% forcing)."""

    def __init__(self, nx=800, ny=800, dx=None, dy=None, lx=0.02, ly=0.02):
        self.nx, self.ny = nx, ny
        self.dx = lx / (nx - 1) if dx is None else dx
        self.dy = ly / (ny - 1) if dy is None else dy
        self.shape = (ny + 6, nx + 6)
        self.h = C.c_void_p()
        check(lib().ls_sample_create(nx, ny, self.dx, self.dy, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ls_sample_destroy(self.h)
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
            if not t.is_cuda or t.dtype.itemsize != 8 or not t.is_contiguous() or tuple(t.shape) != self.shape:
                raise ValueError("fields must be contiguous float64 device tensors of shape %s" % (self.shape,))

    def empty(self):
        import torch
        return torch.zeros(self.shape, dtype=torch.float64, device="cuda")

    def rhs(self, G, u, v, S_L, stream=None):
        self._chk(G, u, v)
        r = self.empty()
        check(lib().ls_sample_rhs(self.h, _ptr(G), _ptr(u), _ptr(v), S_L, _ptr(r), _stream(stream)))
        return r

    def rk3_step(self, G, u, v, dt, S_L, G_initial, stream=None):
        self._chk(G, u, v, G_initial)
        check(lib().ls_sample_rk3_step(self.h, _ptr(G), _ptr(u), _ptr(v), dt, S_L, _ptr(G_initial),
                                       _stream(stream)))

    def apply_periodic(self, G, stream=None):
        self._chk(G)
        check(lib().ls_sample_apply_periodic(self.h, _ptr(G), _stream(stream)))

    def apply_anchored_bc(self, G, G_initial, stream=None):
        self._chk(G, G_initial)
        check(lib().ls_sample_apply_anchored_bc(self.h, _ptr(G), _ptr(G_initial), _stream(stream)))

    def reinitialize(self, G, G_initial, num_iters, stream=None):
        self._chk(G, G_initial)
        check(lib().ls_sample_reinitialize(self.h, _ptr(G), _ptr(G_initial), num_iters, _stream(stream)))

    def first_order_upwind(self, G, u, v, dt, S_L, stream=None):
        self._chk(G, u, v)
        check(lib().ls_sample_first_order_upwind(self.h, _ptr(G), _ptr(u), _ptr(v), dt, S_L,
                                                 _stream(stream)))

    def bunsen(self, R=0.005, angle=30.0, X0=0.01, Y0=0.0, stream=None):
        G = self.empty()
        check(lib().ls_sample_bunsen(self.h, R, angle, X0, Y0, _ptr(G), _stream(stream)))
        return G

    def forced_velocity(self, t, freq, u, v, u0=0.0, v0=0.8, eps_u=0.05, eps_v=0.05, stream=None):
        self._chk(u, v)
        check(lib().ls_sample_forced_velocity(self.h, t, freq, u0, v0, eps_u, eps_v, _ptr(u), _ptr(v),
                                              _stream(stream)))

    def advance(self, G, u, v, G_initial, nsteps, freq=10, t0=0.0, dt=2.5e-6, S_L=0.4, u0=0.0, v0=0.8,
                eps_u=0.05, eps_v=0.05, stream=None):
        """nsteps iterations of the sample driver loop; returns the final time."""
        self._chk(G, u, v, G_initial)
        t = C.c_double()
        check(lib().ls_sample_advance(self.h, _ptr(G), _ptr(u), _ptr(v), _ptr(G_initial), S_L, dt, t0,
                                      nsteps, freq, u0, v0, eps_u, eps_v, C.byref(t), _stream(stream)))
        return t.value


def sample_write_G(host_array, path):
    """save_G_to_bin of a host (ny+6, nx+6) array."""
    import numpy as np
    a = np.ascontiguousarray(host_array, dtype=np.float64)
    ny, nx = a.shape[0] - 6, a.shape[1] - 6
    check(lib().ls_sample_write_G(nx, ny, C.c_void_p(a.ctypes.data), str(path).encode()))
