"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / the timed CPU baseline.  It wraps
oracle/liboracle.so (built by oracle/Makefile) -- the plain-C restatement of
the reference's one-rank vlp4d_mpi OpenMP step (see vlp4d_oracle.h).

Arrays cross the boundary as float64 numpy arrays.  4-D halo arrays have shape
(N3+6, N2+6, N1+6, N0+6) in numpy order (x fastest in memory), position p of
dimension d at index p+3.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
EXE_PATH = os.path.join(HERE, "vlp4d_oracle")

VO_OK, VO_E_FILE, VO_E_CASE, VO_E_CFL, VO_E_ARG, VO_E_SIZE = range(6)


class VoConfig(ctypes.Structure):
    _fields_ = [
        ("nxmax", ctypes.c_int * 4),
        ("minPhy", ctypes.c_double * 4),
        ("maxPhy", ctypes.c_double * 4),
        ("dx", ctypes.c_double * 4),
        ("idcase", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("nbiter", ctypes.c_int),
        ("ifreq", ctypes.c_int),
        ("fxvx", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    vp = ctypes.c_void_p
    C = P(VoConfig)
    lib.vo_parse_deck.argtypes = [ctypes.c_char_p, C]
    lib.vo_config_check.argtypes = [C]
    lib.vo_finish_config.argtypes = [C]
    lib.vo_halo_size.argtypes = [C]
    lib.vo_halo_size.restype = ctypes.c_size_t
    lib.vo_init_case.argtypes = [C, dp]
    lib.vo_exchange_halo.argtypes = [C, dp]
    lib.vo_spline_xy.argtypes = [C, dp]
    lib.vo_spline_vxvy.argtypes = [C, dp]
    lib.vo_advect_2d.argtypes = [C, dp, dp, ctypes.c_double]
    lib.vo_advect_4d.argtypes = [C, dp, dp, dp, dp, ctypes.c_double]
    lib.vo_feet.argtypes = [C, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
    lib.vo_field_rho.argtypes = [C, dp, dp]
    lib.vo_field_poisson.argtypes = [C, dp, dp, dp]
    lib.vo_diags_field.argtypes = [C, dp, dp, dp, dp]
    lib.vo_l2norm.argtypes = [C, dp]
    lib.vo_l2norm.restype = ctypes.c_double
    lib.vo_spline_line.argtypes = [ctypes.c_int, dp, dp]
    lib.vo_rfft2.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    lib.vo_irfft2.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    lib.vo_create.argtypes = [C, P(vp)]
    lib.vo_step.argtypes = [vp]
    lib.vo_run.argtypes = [vp, ctypes.c_int]
    lib.vo_destroy.argtypes = [vp]
    lib.vo_iter.argtypes = [vp]
    lib.vo_fn.argtypes = [vp]
    lib.vo_fn.restype = dp
    lib.vo_field.argtypes = [vp, ctypes.c_int]
    lib.vo_field.restype = dp
    lib.vo_diag_series.argtypes = [vp, ctypes.c_int]
    lib.vo_diag_series.restype = dp
    lib.vo_timer_seconds.argtypes = [vp, ctypes.c_int]
    lib.vo_timer_seconds.restype = ctypes.c_double
    lib.vo_timer_calls.argtypes = [vp, ctypes.c_int]
    lib.vo_save.argtypes = [vp, ctypes.c_char_p]
    lib.vo_get_interior.argtypes = [C, dp, dp]
    lib.vo_set_interior.argtypes = [C, dp, dp]
    lib.vl_lag_basis.argtypes = [ctypes.c_double, dp]
    lib.vl_advect_xy.argtypes = [C, ctypes.c_int, dp, dp, ctypes.c_double]
    lib.vl_advect_v.argtypes = [C, ctypes.c_int, dp, dp, dp, ctypes.c_double]
    lib.vl_field_rho.argtypes = [C, dp, dp]
    lib.vl_init_case.argtypes = [C, dp]
    lib.vl_create.argtypes = [C, P(vp)]
    lib.vl_step.argtypes = [vp]
    lib.vl_run.argtypes = [vp, ctypes.c_int]
    lib.vl_destroy.argtypes = [vp]
    lib.vl_iter.argtypes = [vp]
    lib.vl_f.argtypes = [vp]
    lib.vl_f.restype = dp
    lib.vl_set_f.argtypes = [vp, dp]
    lib.vl_field.argtypes = [vp, ctypes.c_int]
    lib.vl_field.restype = dp
    lib.vl_diag_series.argtypes = [vp, ctypes.c_int]
    lib.vl_diag_series.restype = dp
    _lib = lib
    return lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(n, minphy, maxphy, idcase, dt, nbiter, ifreq=1, fxvx=0):
    c = VoConfig()
    n = list(n) if hasattr(n, "__len__") else [n] * 4
    for i in range(4):
        c.nxmax[i] = int(n[i])
        c.minPhy[i] = float(minphy[i])
        c.maxPhy[i] = float(maxphy[i])
    c.idcase, c.dt, c.nbiter, c.ifreq, c.fxvx = int(idcase), float(dt), int(nbiter), int(ifreq), int(fxvx)
    load().vo_finish_config(ctypes.byref(c))
    return c


def parse_deck(path):
    c = VoConfig()
    rc = load().vo_parse_deck(os.fsencode(path), ctypes.byref(c))
    if rc:
        raise RuntimeError(f"vo_parse_deck failed ({rc})")
    return c


def halo_shape(c):
    return tuple(c.nxmax[d] + 6 for d in (3, 2, 1, 0))


def interior_shape(c):
    return tuple(c.nxmax[d] for d in (3, 2, 1, 0))


def halo_zeros(c):
    return np.zeros(halo_shape(c), dtype=np.float64)


def interior_of(h):
    """Interior view (positions 0..N-1) of a halo array."""
    return h[3:-3, 3:-3, 3:-3, 3:-3]


def ext_view(h, lo=-2, hi=1):
    """Positions lo..N+hi of every dimension of a halo array."""
    return h[3 + lo:h.shape[0] - 3 + hi + 1, 3 + lo:h.shape[1] - 3 + hi + 1,
             3 + lo:h.shape[2] - 3 + hi + 1, 3 + lo:h.shape[3] - 3 + hi + 1]


# ---- thin wrappers over the phase functions ---------------------------------

def init_case(c):
    h = halo_zeros(c)
    rc = load().vo_init_case(ctypes.byref(c), _p(h))
    if rc:
        raise RuntimeError(f"vo_init_case failed ({rc})")
    return h


def exchange_halo(c, h):
    load().vo_exchange_halo(ctypes.byref(c), _p(h))


def spline_xy(c, h):
    load().vo_spline_xy(ctypes.byref(c), _p(h))


def spline_vxvy(c, h):
    load().vo_spline_vxvy(ctypes.byref(c), _p(h))


def advect_2d(c, h, dt):
    tmp = np.empty_like(h)
    return load().vo_advect_2d(ctypes.byref(c), _p(h), _p(tmp), dt)


def advect_4d(c, ex, ey, h, dt):
    tmp = np.empty_like(h)
    return load().vo_advect_4d(ctypes.byref(c), _p(ex), _p(ey), _p(h), _p(tmp), dt)


def feet(c, ex, ey, ix, iy, ivx, ivy, dt):
    out = np.zeros(4)
    load().vo_feet(ctypes.byref(c), _p(ex), _p(ey), ix, iy, ivx, ivy, dt, _p(out))
    return out


def field_rho(c, h):
    rho = np.zeros((c.nxmax[1], c.nxmax[0]))
    load().vo_field_rho(ctypes.byref(c), _p(h), _p(rho))
    return rho


def field_poisson(c, rho):
    """Returns (phi_or_rho, ex, ey); rho is copied first."""
    r = np.array(rho, dtype=np.float64, copy=True, order="C")
    ex = np.zeros_like(r)
    ey = np.zeros_like(r)
    load().vo_field_poisson(ctypes.byref(c), _p(r), _p(ex), _p(ey))
    return r, ex, ey


def diags_field(c, rho, ex, ey):
    out = np.zeros(4)
    load().vo_diags_field(ctypes.byref(c), _p(rho), _p(ex), _p(ey), _p(out))
    return out


def l2norm(c, h):
    return load().vo_l2norm(ctypes.byref(c), _p(h))


def spline_line(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    c = np.zeros(f.size + 4)
    rc = load().vo_spline_line(f.size, _p(f), _p(c))
    if rc:
        raise RuntimeError(f"vo_spline_line failed ({rc})")
    return c


def rfft2(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ny, nx = x.shape
    out = np.zeros((ny, nx // 2 + 1, 2))
    load().vo_rfft2(nx, ny, _p(x), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def irfft2(X, nx):
    ny = X.shape[0]
    buf = np.ascontiguousarray(np.stack([X.real, X.imag], axis=-1), dtype=np.float64)
    out = np.zeros((ny, nx))
    load().vo_irfft2(nx, ny, _p(buf), _p(out))
    return out


class State:
    """vo_state: the oracle driver (init + onetimestep + swap)."""

    def __init__(self, c):
        self.lib = load()
        self.c = c
        h = ctypes.c_void_p()
        rc = self.lib.vo_create(ctypes.byref(c), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"vo_create failed ({rc})")
        self.h = h

    def step(self, n=1):
        rc = self.lib.vo_run(self.h, n)
        if rc:
            raise RuntimeError(f"vo_run failed ({rc})")

    @property
    def iter(self):
        return self.lib.vo_iter(self.h)

    def fn(self):
        n = int(np.prod(halo_shape(self.c)))
        p = self.lib.vo_fn(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(halo_shape(self.c)).copy()

    def f_interior(self):
        return np.ascontiguousarray(interior_of(self.fn()))

    def set_f(self, f):
        """Replace the interior of f^n (x fastest, [vy][vx][y][x]); the next
        step starts from it (its halo is rebuilt by the exchange)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == interior_shape(self.c)
        self.lib.vo_set_interior(ctypes.byref(self.c), self.lib.vo_fn(self.h), _p(f))

    def field(self, which):
        n = self.c.nxmax[0] * self.c.nxmax[1]
        p = self.lib.vo_field(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.c.nxmax[1], self.c.nxmax[0]).copy()

    def diags(self):
        n = self.c.nbiter + 1
        out = {}
        for k, name in enumerate(["nrj", "nrjx", "nrjy", "mass", "l2norm"]):
            p = self.lib.vo_diag_series(self.h, k)
            out[name] = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return out

    def timer(self, k):
        return self.lib.vo_timer_seconds(self.h, k), self.lib.vo_timer_calls(self.h, k)

    def close(self):
        if self.h:
            self.lib.vo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- Lagrange split-step vlp4d (miniapps/vlp4d/openmp) ----------------------

def lag_basis(px):
    out = np.zeros(6)
    load().vl_lag_basis(px, _p(out))
    return out


def lag_advect_xy(c, d, f, dt):
    """One 1-D advection along x (d=0) or y (d=1) by dt*v; f has no halo."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    load().vl_advect_xy(ctypes.byref(c), d, _p(f), _p(out), dt)
    return out


def lag_advect_v(c, d, f, e, dt):
    """One 1-D advection along vx (d=2) or vy (d=3) by dt*E(x, y)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    out = np.empty_like(f)
    load().vl_advect_v(ctypes.byref(c), d, _p(f), _p(out), _p(e), dt)
    return out


def lag_field_rho(c, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    rho = np.zeros((c.nxmax[1], c.nxmax[0]))
    load().vl_field_rho(ctypes.byref(c), _p(f), _p(rho))
    return rho


def lag_init_case(c):
    f = np.zeros(interior_shape(c))
    rc = load().vl_init_case(ctypes.byref(c), _p(f))
    if rc:
        raise RuntimeError(f"vl_init_case failed ({rc})")
    return f


class LagState:
    """vl_state: the Lagrange split-step driver (init + onetimestep)."""

    def __init__(self, c):
        self.lib = load()
        self.c = c
        h = ctypes.c_void_p()
        rc = self.lib.vl_create(ctypes.byref(c), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"vl_create failed ({rc})")
        self.h = h

    def step(self, n=1):
        rc = self.lib.vl_run(self.h, n)
        if rc:
            raise RuntimeError(f"vl_run failed ({rc})")

    @property
    def iter(self):
        return self.lib.vl_iter(self.h)

    def f(self):
        n = int(np.prod(interior_shape(self.c)))
        return np.ctypeslib.as_array(self.lib.vl_f(self.h), shape=(n,)).reshape(interior_shape(self.c)).copy()

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == int(np.prod(interior_shape(self.c)))
        self.lib.vl_set_f(self.h, _p(f))

    def field(self, which):
        n = self.c.nxmax[0] * self.c.nxmax[1]
        p = self.lib.vl_field(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.c.nxmax[1], self.c.nxmax[0]).copy()

    def diags(self):
        n = self.c.nbiter + 1
        return {name: np.ctypeslib.as_array(self.lib.vl_diag_series(self.h, k), shape=(n,)).copy()
                for k, name in enumerate(["nrj", "mass"])}

    def close(self):
        if self.h:
            self.lib.vl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
