"""ctypes front end to the CPU checkers (test infrastructure only).

Loads ``oracle/liblsoracle.so`` (the C restatement in ``ls_oracle.c``) and,
when present, ``oracle/_ref/libref{2,3}d.so`` (the reference's own solver
compiled from its sources by ``oracle/Makefile``).  Only the tests, the smoke
entry point and the bench's CPU-baseline leg import this module; the product
path never does.

All fields are numpy float64 arrays in the padded layout, x fastest:
shape (nyt, nxt) in 2D and (nzt, nyt, nxt) in 3D with 3 ghost layers.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

BC_PERIODIC, BC_ZERO_GRAD, BC_LIN_EXTRAP = 0, 1, 2


class Grid(C.Structure):
    """Mirror of ``lso_grid`` (ls_oracle.c) and ``ref_grid`` (ref_harness.cpp)."""
    _fields_ = [("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("ng", C.c_int32), ("bc", C.c_int32 * 3),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("s_l", C.c_double), ("eps", C.c_double),
                ("reinit_iters", C.c_int32), ("pad0", C.c_int32),
                ("reinit_dtau", C.c_double), ("reinit_beta", C.c_double)]

    @property
    def shape(self):
        if self.dim == 3:
            return (self.nz + 2 * self.ng, self.ny + 2 * self.ng, self.nx + 2 * self.ng)
        return (self.ny + 2 * self.ng, self.nx + 2 * self.ng)

    @property
    def interior(self):
        g = self.ng
        if self.dim == 3:
            return (slice(g, g + self.nz), slice(g, g + self.ny), slice(g, g + self.nx))
        return (slice(g, g + self.ny), slice(g, g + self.nx))


def make_grid(dim, nx, ny, nz=1, L=(1.0, 1.0, 1.0), bc=(0, 1, 0), s_l=0.0,
              reinit_iters=2, eps=1e-10, dtau_factor=0.25, beta=0.5):
    """Node-centred spacing dx = L/(N-1) (Local_code/level-set_MPI_2D/include/config.h:43-44)."""
    g = Grid()
    g.dim, g.nx, g.ny, g.nz, g.ng = dim, nx, ny, (nz if dim == 3 else 1), 3
    for a in range(3):
        g.bc[a] = bc[a]
    g.dx = L[0] / (nx - 1)
    g.dy = L[1] / (ny - 1)
    g.dz = L[2] / (nz - 1) if dim == 3 else 1.0
    g.s_l, g.eps = s_l, eps
    g.reinit_iters = reinit_iters
    h = min(g.dx, g.dy, g.dz) if dim == 3 else min(g.dx, g.dy)
    g.reinit_dtau = dtau_factor * h
    g.reinit_beta = beta
    return g


_P = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


class _Lib:
    """Common wrapper: both libraries export the same operations."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix

    def f(self, name):
        return getattr(self.lib, self.prefix + name)


class Oracle(_Lib):
    """The C restatement (ls_oracle.c)."""

    def __init__(self, fma=False):
        super().__init__(os.path.join(HERE, "liblsoracle_fma.so" if fma else "liblsoracle.so"), "lso_")
        L = self.lib
        L.lso_padded_size.restype = C.c_int64
        L.lso_weno_left.restype = C.c_double
        L.lso_weno_right.restype = C.c_double
        L.lso_l2.restype = C.c_double
        L.lso_area.restype = C.c_double
        L.lso_advance.restype = C.c_int64
        L.lso_advance.argtypes = [C.POINTER(Grid), _P, _P, _P, _P, _P, C.c_double, C.c_double,
                                  C.c_int64, C.c_int32, C.c_double, _P, _I, _P, _P]
        L.lso_rk3_step.argtypes = [C.POINTER(Grid), _P, _P, _P, _P, _P, _P, C.c_double, _P]
        L.lso_deformation_velocity.argtypes = [C.POINTER(Grid), C.c_double, C.c_double, C.c_double,
                                               C.c_double, C.c_double, _P, _P, _P]
        L.lso_stage.argtypes = [C.c_int, C.c_int64, _P, _P, _P, _P, C.c_double]

    def weno_left(self, v):
        return self.lib.lso_weno_left(_p(np.ascontiguousarray(v, np.float64)))

    def weno_right(self, v):
        return self.lib.lso_weno_right(_p(np.ascontiguousarray(v, np.float64)))

    def rhs(self, g, G, u, v, w=None, sl=None):
        R = np.zeros_like(G)
        self.lib.lso_rhs(C.byref(g), _p(G), _p(R), _p(u), _p(v), _p(w), _p(sl))
        return R

    def fill_ghosts(self, g, G):
        G = G.copy()
        self.lib.lso_fill_ghosts(C.byref(g), _p(G))
        return G

    def rk3_step(self, g, G, u, v, w=None, dt=1e-3, sl=None):
        out = np.zeros_like(G)
        work = np.zeros((3,) + G.shape)
        self.lib.lso_rk3_step(C.byref(g), _p(G), _p(out), _p(u), _p(v), _p(w), _p(sl),
                              C.c_double(dt), _p(work))
        return out

    def crossings(self, g, p0):
        rt = np.zeros_like(p0)
        fl = np.zeros(p0.shape, np.int32)
        self.lib.lso_crossings(C.byref(g), _p(p0), _p(rt), _pi(fl))
        return rt, fl

    def hcr2(self, g, G):
        G = G.copy()
        work = np.zeros((3,) + G.shape)
        fl = np.zeros(G.shape, np.int32)
        self.lib.lso_hcr2(C.byref(g), _p(G), _p(work), _pi(fl))
        return G

    def advance(self, g, G, u, v, w=None, sl=None, dt=1e-3, t_final=1.0, max_steps=None,
                reinit_interval=0, deform_T=0.0):
        """Returns (G_final, steps, t_final_reached, dt_last); u/v/w may be updated in place."""
        G = G.copy()
        if max_steps is None:
            max_steps = int(t_final / dt) + 100
        work = np.zeros((6,) + G.shape)
        fl = np.zeros(G.shape, np.int32)
        t = C.c_double()
        dl = C.c_double()
        n = self.lib.lso_advance(C.byref(g), _p(G), _p(u), _p(v), _p(w), _p(sl), dt, t_final,
                                 max_steps, reinit_interval, deform_T, _p(work), _pi(fl),
                                 C.byref(t), C.byref(dl))
        return G, n, t.value, dl.value

    def timed_loop(self, g, G, G0, u, v, w=None, dt=1e-3, nsteps=10, reinit_interval=0, diag_interval=0,
                   threads=1):
        """lso_timed_loop: the reference's timed region on `threads` OpenMP
        threads.  Returns (seconds, G_final, last_l2, last_area)."""
        L = self.lib
        L.lso_timed_loop.restype = C.c_double
        L.lso_timed_loop.argtypes = [C.POINTER(Grid), _P, _P, _P, _P, _P, C.c_double, C.c_int64, C.c_int32,
                                     C.c_int32, C.c_int32, _P, _I, _P]
        G = G.copy()
        work = np.zeros((6,) + G.shape)
        fl = np.zeros(G.shape, np.int32)
        d = np.zeros(2)
        sec = L.lso_timed_loop(C.byref(g), _p(G), _p(G0), _p(u), _p(v), _p(w), dt, nsteps, reinit_interval,
                               diag_interval, threads, _p(work), _pi(fl), _p(d))
        return sec, G, float(d[0]), float(d[1])

    def deformation_velocity(self, g, t, T):
        u = np.zeros(g.shape)
        v = np.zeros(g.shape)
        w = np.zeros(g.shape)
        self.lib.lso_deformation_velocity(C.byref(g), 0.0, 0.0, 0.0, t, T, _p(u), _p(v), _p(w))
        return u, v, w

    def l2(self, g, G, G0):
        return self.lib.lso_l2(C.byref(g), _p(G), _p(G0))

    def ic_zalesak2d(self, g, T):
        G, u, v = np.zeros(g.shape), np.zeros(g.shape), np.zeros(g.shape)
        self.lib.lso_ic_zalesak2d(C.byref(g), C.c_double(T), _p(G), _p(u), _p(v))
        return G, u, v

    def ic_sphere3d(self, g, c=(0.35, 0.35, 0.35), r=0.15):
        G = np.zeros(g.shape)
        self.lib.lso_ic_sphere3d(C.byref(g), C.c_double(c[0]), C.c_double(c[1]), C.c_double(c[2]),
                                 C.c_double(r), _p(G))
        return G

    def area(self, g, G):
        return self.lib.lso_area(C.byref(g), _p(G))

    def write_G(self, g, G, path):
        return self.lib.lso_write_G(C.byref(g), _p(G), str(path).encode())

    def write_vtk(self, g, G, path, origin=(0.0, 0.0, 0.0)):
        o = [C.c_double(x) for x in origin]
        return self.lib.lso_write_vtk(C.byref(g), _p(G), o[0], o[1], o[2], str(path).encode())


def ref_available(dim=2, fma=False):
    return os.path.exists(os.path.join(HERE, "_ref", "libref%dd%s.so" % (dim, "_fma" if fma else "")))


class Reference(_Lib):
    """The reference's own functions compiled from its sources (oracle/_ref)."""

    def __init__(self, dim, fma=False):
        super().__init__(os.path.join(HERE, "_ref", "libref%dd%s.so" % (dim, "_fma" if fma else "")), "ref_")
        L = self.lib
        L.ref_l2.restype = C.c_double
        L.ref_area.restype = C.c_double
        L.ref_advance.restype = C.c_int64
        L.ref_advance.argtypes = [C.POINTER(Grid), _P, _P, _P, _P, C.c_double, C.c_double,
                                  C.c_int64, C.c_int32, C.c_double, _P, _I, _P, _P]
        L.ref_rk3_step.argtypes = [C.POINTER(Grid), _P, _P, _P, _P, _P, C.c_double, _P, C.c_int]
        L.ref_init_case.argtypes = [C.POINTER(Grid), C.c_int, C.c_double, C.c_double, C.c_double,
                                    _P, _P, _P, _P]

    def rhs(self, g, G, u, v, w=None):
        R = np.zeros_like(G)
        rc = self.lib.ref_rhs(C.byref(g), _p(G), _p(R), _p(u), _p(v), _p(w))
        assert rc == 0
        return R

    def fill_ghosts(self, g, G):
        G = G.copy()
        rc = self.lib.ref_fill_ghosts(C.byref(g), _p(G))
        assert rc == 0
        return G

    def rk3_step(self, g, G, u, v, w=None, dt=1e-3, direct=False):
        G = G.copy()
        out = np.zeros_like(G)
        work = np.zeros((3,) + G.shape)
        rc = self.lib.ref_rk3_step(C.byref(g), _p(G), _p(out), _p(u), _p(v), _p(w),
                                   C.c_double(dt), _p(work), 1 if direct else 0)
        assert rc == 0
        return out

    def hcr2(self, g, G, direct=False):
        G = G.copy()
        work = np.zeros((3,) + G.shape)
        fl = np.zeros(G.shape, np.int32)
        rc = self.lib.ref_hcr2(C.byref(g), _p(G), _p(work), _pi(fl), 1 if direct else 0)
        assert rc == 0
        return G

    def crossings(self, g, p0):
        """computeInterfaceCrossings[3D]: (r_tilde, flag); ghost entries stay 0."""
        rt = np.zeros_like(p0)
        fl = np.zeros(p0.shape, np.int32)
        rc = self.lib.ref_crossings(C.byref(g), _p(np.ascontiguousarray(p0)), _p(rt), _pi(fl))
        assert rc == 0
        return rt, fl

    def advance(self, g, G, u, v, w=None, dt=1e-3, t_final=1.0, max_steps=None,
                reinit_interval=0, deform_T=0.0):
        G = G.copy()
        if max_steps is None:
            max_steps = int(t_final / dt) + 100
        work = np.zeros((7,) + G.shape)
        fl = np.zeros(G.shape, np.int32)
        t = C.c_double()
        dl = C.c_double()
        n = self.lib.ref_advance(C.byref(g), _p(G), _p(u), _p(v), _p(w), dt, t_final, max_steps,
                                 reinit_interval, deform_T, _p(work), _pi(fl), C.byref(t),
                                 C.byref(dl))
        assert n >= 0
        return G, n, t.value, dl.value

    def init_case(self, g, which, t_final=1.0, u_c=0.0, v_c=0.0):
        G = np.zeros(g.shape)
        u = np.zeros(g.shape)
        v = np.zeros(g.shape)
        w = np.zeros(g.shape) if g.dim == 3 else None
        rc = self.lib.ref_init_case(C.byref(g), which, t_final, u_c, v_c, _p(G), _p(u), _p(v), _p(w))
        assert rc == 0
        return G, u, v, w

    def l2(self, g, G, G0):
        return self.lib.ref_l2(C.byref(g), _p(G), _p(G0))

    def area(self, g, G):
        return self.lib.ref_area(C.byref(g), _p(G))

    def write_G(self, g, G, path):
        return self.lib.ref_write_G(C.byref(g), _p(G), str(path).encode())

    def write_vtk(self, g, G, path):
        return self.lib.ref_write_vtk(C.byref(g), _p(G), str(path).encode())


class Tangent:
    """Forward-mode restatement of the 3D path (ls_tangent.cpp): the loss and
    its derivative along one direction of (G0, u, v, w, S_L)."""

    def __init__(self):
        self.lib = C.CDLL(os.path.join(HERE, "liblstangent.so"))
        self.lib.lst_run.restype = C.c_int

    def run(self, g, G0, u, v, w, G_target, eps_h, dt, nsteps, reinit_interval, sl=None,
            dG0=None, du=None, dv=None, dw=None, dsl=None):
        """Returns (G_T, L, dL) with dL the directional derivative."""
        out = np.zeros(g.shape)
        L = C.c_double()
        dL = C.c_double()
        rc = self.lib.lst_run(C.byref(g), _p(G0), _p(dG0), _p(u), _p(du), _p(v), _p(dv), _p(w),
                              _p(dw), _p(sl), _p(dsl), _p(G_target), C.c_double(eps_h),
                              C.c_double(dt), C.c_int32(nsteps), C.c_int32(reinit_interval),
                              _p(out), C.byref(L), C.byref(dL))
        assert rc == 0, "lst_run rejected the grid"
        return out, L.value, dL.value


def heaviside(x, e):
    """Smoothed Heaviside of the adjoint loss, numpy."""
    x = np.asarray(x, np.float64)
    h = 0.5 * (1.0 + x / e + np.sin(np.pi * x / e) / np.pi)
    return np.where(x < -e, 0.0, np.where(x > e, 1.0, h))


class _Sample:
    """The sample/ scheme (2D, LLF Hamiltonian) on padded (ny+6, nx+6) arrays."""

    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.pre = prefix
        for name in ("rhs", "rk3_step", "apply_periodic", "apply_anchored", "reinitialize",
                     "first_order_upwind"):
            getattr(self.lib, prefix + name).restype = None

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    @staticmethod
    def _n(a):
        ny, nx = a.shape
        return nx - 6, ny - 6

    def rhs(self, G, u, v, dx, dy, S_L):
        nx, ny = self._n(G)
        r = np.zeros_like(G)
        self._f("rhs")(nx, ny, _p(G), _p(u), _p(v), _p(r), C.c_double(dx), C.c_double(dy), C.c_double(S_L))
        return r

    def apply_periodic(self, G):
        nx, ny = self._n(G)
        G = G.copy()
        self._f("apply_periodic")(nx, ny, _p(G))
        return G

    def apply_anchored(self, G, init):
        nx, ny = self._n(G)
        G = G.copy()
        self._f("apply_anchored")(nx, ny, _p(G), _p(init))
        return G


class SampleOracle(_Sample):
    """C restatement of sample/ (ls_sample.c, inside liblsoracle.so)."""

    def __init__(self, fma=False):
        super().__init__(os.path.join(HERE, "liblsoracle_fma.so" if fma else "liblsoracle.so"), "lss_")
        self.lib.lss_write_G.restype = C.c_int
        self.lib.lss_weno_left.restype = C.c_double
        self.lib.lss_weno_right.restype = C.c_double
        self.lib.lss_beta.restype = C.c_double

    def rk3_step(self, G, u, v, dx, dy, dt, S_L, G_initial):
        nx, ny = self._n(G)
        G = G.copy()
        work = np.zeros((3,) + G.shape)
        self.lib.lss_rk3_step(nx, ny, _p(G), _p(u), _p(v), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                              C.c_double(S_L), _p(G_initial), _p(work))
        return G

    def reinitialize(self, G, init, dx, dy, iters):
        nx, ny = self._n(G)
        G = G.copy()
        work = np.zeros(G.shape)
        self.lib.lss_reinitialize(nx, ny, _p(G), _p(init), C.c_double(dx), C.c_double(dy), iters, _p(work))
        return G

    def first_order_upwind(self, G, u, v, dx, dy, dt, S_L):
        nx, ny = self._n(G)
        G = G.copy()
        work = np.zeros(G.shape)
        self.lib.lss_first_order_upwind(nx, ny, _p(G), _p(u), _p(v), C.c_double(dx), C.c_double(dy),
                                        C.c_double(dt), C.c_double(S_L), _p(work))
        return G

    def bunsen(self, n, dx, dy, R=0.005, angle=30.0, X0=0.01, Y0=0.0):
        G = np.zeros((n + 6, n + 6))
        self.lib.lss_bunsen(n, n, C.c_double(dx), C.c_double(dy), C.c_double(R), C.c_double(angle),
                            C.c_double(X0), C.c_double(Y0), _p(G))
        return G

    def write_G(self, G, path):
        nx, ny = self._n(G)
        return self.lib.lss_write_G(nx, ny, _p(G), str(path).encode())


def sample_ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libsample.so"))


class SampleReference(_Sample):
    """The reference's own sample/ functions compiled from its sources (oracle/_ref)."""

    def __init__(self):
        super().__init__(os.path.join(HERE, "_ref", "libsample.so"), "refs_")

    def rk3_step(self, G, u, v, dx, dy, dt, S_L, G_initial):
        nx, ny = self._n(G)
        G = G.copy()
        self.lib.refs_rk3_step(nx, ny, _p(G), _p(u), _p(v), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                               C.c_double(S_L), _p(G_initial))
        return G

    def reinitialize(self, G, init, dx, dy, iters):
        nx, ny = self._n(G)
        G = G.copy()
        self.lib.refs_reinitialize(nx, ny, _p(G), _p(init), C.c_double(dx), C.c_double(dy), iters)
        return G

    def first_order_upwind(self, G, u, v, dx, dy, dt, S_L):
        nx, ny = self._n(G)
        G = G.copy()
        self.lib.refs_first_order_upwind(nx, ny, _p(G), _p(u), _p(v), C.c_double(dx), C.c_double(dy),
                                         C.c_double(dt), C.c_double(S_L))
        return G
