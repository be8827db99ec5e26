"""Python mirror of the vlp4d driver interface over the C-ABI of libvlp4d.so.

The reference (miniapps/vlp4d_mpi/openmp) drives its hot path from C++:
``init`` (VMO/Init.cpp:324), ``onetimestep`` (VMO/Timestep.hpp:34), the
component functions ``Spline::computeCoeff_xy/vxvy``, ``Advection::advect_2D_xy
/advect_4D``, ``field_rho/field_reduce/field_poisson`` and ``Diags::compute/
computeL2norm/save``.  :class:`Vlp4d` exposes the same operations on one GPU:
the whole step (:meth:`Vlp4d.step`) and the four phases of one step.

Errors are raised as :class:`Vlp4dError` carrying the C-ABI status code; the
reference's ``exit(0)`` on a too-large time step is :data:`E_CFL` here.

The directory name is not a Python identifier; import it with
``importlib`` (see ``__graft_entry__.py``) or load ``libvlp4d.so`` directly.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvlp4d.so")

OK, E_FILE, E_CASE, E_CFL, E_ARG, E_HIP, E_SIZE = range(7)
FP64, FP32, FP32_ACC = 64, 32, 33
SCHEME_SPLINE, SCHEME_LAGRANGE = 0, 1
BUF_F, BUF_CXY, BUF_C4, BUF_RHO, BUF_EX, BUF_EY = range(6)
# extension initial conditions (no reference counterpart), include/vlp4d.h
CASE_STRONG_LANDAU, CASE_TWO_STREAM = 30, 31
TIMER_LABELS = ["total", "MainLoop", "pack", "comm", "unpack", "advec2D", "advec4D", "field",
                "all_reduce", "Fourier", "diag", "splinecoeff_xy", "splinecoeff_vxvy"]

# Every symbol declared in include/vlp4d.h.
EXPORTS = [
    "vlp_version", "vlp_parse_deck", "vlp_create", "vlp_create_config", "vlp_destroy", "vlp_step",
    "vlp_iter", "vlp_get_config", "vlp_diags", "vlp_save", "vlp_get_f", "vlp_set_f",
    "vlp_buffer_size", "vlp_get_buffer", "vlp_spline_xy", "vlp_advect2d_poisson",
    "vlp_spline_vxvy", "vlp_advect4d_poisson", "vlp_timer", "vlp_timer_label", "vlp_last_error",
    "vlp_kernel_variants", "vlp_autotune", "vlp_tuned_tile", "vlp_ensemble_create", "vlp_ensemble_step",
    "vlp_ensemble_size", "vlp_ensemble_destroy", "vlp_adjoint_step", "vlp_adjoint_steps", "vlp_create_lagrange", "vlp_write_fxvx",
    "vlp_field_solve", "vlp_adjoint_field_energy", "vlp_field_solve_noalloc",
]


class VlpConfig(ctypes.Structure):
    """``vlp_config`` of include/vlp4d.h (Domain of VMO/Config.hpp:34-48)."""
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

    @classmethod
    def make(cls, n, minphy, maxphy, idcase, dt, nbiter, ifreq=1, fxvx=0):
        c = cls()
        n = list(n) if hasattr(n, "__len__") else [n] * 4
        for i in range(4):
            c.nxmax[i] = int(n[i])
            c.minPhy[i] = float(minphy[i])
            c.maxPhy[i] = float(maxphy[i])
            c.dx[i] = (c.maxPhy[i] - c.minPhy[i]) / c.nxmax[i]
        c.idcase, c.dt, c.nbiter, c.ifreq, c.fxvx = int(idcase), float(dt), int(nbiter), int(ifreq), int(fxvx)
        return c

    def as_dict(self):
        return {"nxmax": list(self.nxmax), "minPhy": list(self.minPhy), "maxPhy": list(self.maxPhy),
                "dx": list(self.dx), "idcase": self.idcase, "dt": self.dt, "nbiter": self.nbiter,
                "ifreq": self.ifreq, "fxvx": self.fxvx}


class Vlp4dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vlp4d error {code}: {msg}")
        self.code = code


_lib = None


def load(path=LIB_PATH):
    """Load libvlp4d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built; run p3-miniapps_amd/build.py")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    dp = P(ctypes.c_double)
    lib.vlp_version.restype = ctypes.c_char_p
    lib.vlp_parse_deck.argtypes = [ctypes.c_char_p, P(VlpConfig)]
    lib.vlp_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(vp)]
    lib.vlp_create_config.argtypes = [P(VlpConfig), ctypes.c_int, ctypes.c_int, P(vp)]
    lib.vlp_destroy.argtypes = [vp]
    lib.vlp_step.argtypes = [vp, ctypes.c_int]
    lib.vlp_iter.argtypes = [vp]
    lib.vlp_get_config.argtypes = [vp, P(VlpConfig)]
    lib.vlp_diags.argtypes = [vp, dp, dp, dp, dp, dp, ctypes.c_int]
    lib.vlp_save.argtypes = [vp, ctypes.c_char_p]
    lib.vlp_get_f.argtypes = [vp, dp, ctypes.c_size_t]
    lib.vlp_set_f.argtypes = [vp, dp, ctypes.c_size_t]
    lib.vlp_buffer_size.argtypes = [vp, ctypes.c_int]
    lib.vlp_buffer_size.restype = ctypes.c_size_t
    lib.vlp_get_buffer.argtypes = [vp, ctypes.c_int, dp, ctypes.c_size_t]
    for name in ("vlp_spline_xy", "vlp_advect2d_poisson", "vlp_spline_vxvy", "vlp_advect4d_poisson"):
        getattr(lib, name).argtypes = [vp]
    lib.vlp_timer.argtypes = [vp, ctypes.c_int, dp, P(ctypes.c_int)]
    lib.vlp_timer_label.argtypes = [ctypes.c_int]
    lib.vlp_timer_label.restype = ctypes.c_char_p
    lib.vlp_last_error.argtypes = [vp]
    lib.vlp_last_error.restype = ctypes.c_char_p
    lib.vlp_kernel_variants.argtypes = [vp]
    lib.vlp_autotune.argtypes = [vp, ctypes.c_char_p]
    lib.vlp_tuned_tile.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int)]
    lib.vlp_ensemble_create.argtypes = [P(vp), ctypes.c_int, P(vp)]
    lib.vlp_ensemble_step.argtypes = [vp, ctypes.c_int]
    lib.vlp_ensemble_size.argtypes = [vp]
    lib.vlp_ensemble_destroy.argtypes = [vp]
    lib.vlp_adjoint_step.argtypes = [vp, dp, dp, ctypes.c_size_t]
    lib.vlp_adjoint_steps.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_size_t]
    lib.vlp_adjoint_field_energy.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_size_t]
    lib.vlp_create_lagrange.argtypes = [P(VlpConfig), ctypes.c_int, P(vp)]
    lib.vlp_write_fxvx.argtypes = [vp, ctypes.c_char_p]
    lib.vlp_field_solve.argtypes = [P(VlpConfig), vp, ctypes.c_int, vp, vp, vp, vp, vp]
    lib.vlp_field_solve_noalloc.argtypes = [P(VlpConfig), vp, ctypes.c_int, vp, vp, vp, vp, vp]
    _lib = lib
    return lib


def field_solve(cfg, f_ptr, f_bits, phi_ptr, ex_ptr, ey_ptr, diag_ptr, stream=None):
    """vlp_field_solve on device pointers (integers): the density, Poisson
    solve, field diagnostics and L2 norm of a device-resident f."""
    lib = load()
    rc = lib.vlp_field_solve(ctypes.byref(cfg), f_ptr, f_bits, phi_ptr, ex_ptr, ey_ptr, diag_ptr, stream)
    if rc:
        raise Vlp4dError(rc, lib.vlp_last_error(None).decode())


ENSEMBLE_SHAPES = [(32, 32, 32, 32)] * 16 + [(48, 48, 64, 64)] * 16 + [(64, 64, 64, 64)] * 16


def landau_member_setup(n, alpha, k, vmax=6.0):
    """(config arguments, initial f) of one ensemble member: Landau damping
    f0 = (1 + alpha cos(k x) cos(k y)) exp(-|v|^2 / 2) / (2 pi) on the box
    [0, 2 pi / k)^2 x [-vmax, vmax]^2, dt = dx / (2 vmax); returns
    ((n, minphy, maxphy, idcase, dt), f0 as [vy][vx][y][x])."""
    lx = 2 * math.pi / k
    dt = 0.5 * (lx / max(n[0], n[1])) / vmax
    x = np.arange(n[0]) * (lx / n[0])
    y = np.arange(n[1]) * (lx / n[1])
    vx = -vmax + np.arange(n[2]) * (2 * vmax / n[2])
    vy = -vmax + np.arange(n[3]) * (2 * vmax / n[3])
    sp = 1 + alpha * np.cos(k * x)[None, :] * np.cos(k * y)[:, None]           # [y][x]
    mv = np.exp(-0.5 * (vx[None, :] ** 2 + vy[:, None] ** 2)) / (2 * math.pi)  # [vy][vx]
    return (list(n), [0.0, 0.0, -vmax, -vmax], [lx, lx, vmax, vmax], 10, dt), mv[:, :, None, None] * sp[None, None]


def landau_ensemble_params(seed, shapes=None):
    """The batched Landau workload: for each grid (16 x 32^4, 16 x 48^2 x 64^2,
    16 x 64^4 by default) a perturbation amplitude alpha ~ U[0.01, 0.5] and a
    wave number k in {0.3, 0.5}, drawn in order from a seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for n in (shapes or ENSEMBLE_SHAPES):
        alpha = float(rng.uniform(0.01, 0.5))
        k = float(rng.choice([0.3, 0.5]))
        out.append((tuple(n), alpha, k))
    return out


def trig_interp_matrix(n, m):
    """(m x n) periodic trigonometric interpolation from n to m equally spaced
    samples of one period (n even): sample i at i/n, output j at j/m,
    P[j, i] = (1/n) [1 + 2 sum_{k=1}^{n/2-1} cos(2 pi k s) + cos(pi n s)],
    s = j/m - i/n -- the interpolant band-limited to |k| < n/2 with the
    Nyquist mode split evenly between +-n/2.  It maps a density perturbation
    given on a coarser periodic grid to the solver's grid (64 x 64 -> 96 x 96
    as P @ delta @ P.T); its transpose carries the gradient back."""
    if n % 2:
        raise ValueError("n must be even")
    s = np.arange(m)[:, None] / m - np.arange(n)[None, :] / n
    k = np.arange(1, n // 2)
    return (1.0 + 2.0 * np.cos(2 * np.pi * s[..., None] * k).sum(axis=-1) + np.cos(np.pi * n * s)) / n


def field_solve_noalloc(cfg, f_ptr, f_bits, phi_ptr, ex_ptr, ey_ptr, diag_ptr, stream=None):
    """vlp_field_solve_noalloc: the same field solve with no device
    allocation (scratch in the library's static device array); it only
    enqueues the work on `stream`."""
    lib = load()
    rc = lib.vlp_field_solve_noalloc(ctypes.byref(cfg), f_ptr, f_bits, phi_ptr, ex_ptr, ey_ptr, diag_ptr, stream)
    if rc:
        raise Vlp4dError(rc, lib.vlp_last_error(None).decode())


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def parse_deck(path):
    """``import`` of VMO/Init.cpp:12-90."""
    lib = load()
    cfg = VlpConfig()
    rc = lib.vlp_parse_deck(os.fsencode(path), ctypes.byref(cfg))
    if rc:
        raise Vlp4dError(rc, lib.vlp_last_error(None).decode())
    return cfg


class Vlp4d:
    """One simulation on one GPU (context of the C-ABI)."""

    def __init__(self, cfg_or_path, precision=FP64, device=0, scheme=SCHEME_SPLINE):
        """scheme: SCHEME_SPLINE (the spline step of miniapps/vlp4d_mpi) or
        SCHEME_LAGRANGE (the split-step solver of miniapps/vlp4d, FP64)."""
        self.lib = load()
        h = ctypes.c_void_p()
        if scheme == SCHEME_LAGRANGE:
            if precision != FP64:
                raise Vlp4dError(E_ARG, "the Lagrange split-step solver runs in FP64")
            if isinstance(cfg_or_path, (str, bytes, os.PathLike)):
                cfg_or_path = parse_deck(cfg_or_path)
            rc = self.lib.vlp_create_lagrange(ctypes.byref(cfg_or_path), device, ctypes.byref(h))
        elif isinstance(cfg_or_path, (str, bytes, os.PathLike)):
            rc = self.lib.vlp_create(os.fsencode(cfg_or_path), precision, device, ctypes.byref(h))
        else:
            rc = self.lib.vlp_create_config(ctypes.byref(cfg_or_path), precision, device, ctypes.byref(h))
        if rc:
            raise Vlp4dError(rc, self.lib.vlp_last_error(None).decode())
        self.h = h
        self.precision = precision
        self.cfg = VlpConfig()
        self.lib.vlp_get_config(self.h, ctypes.byref(self.cfg))
        self.shape4 = tuple(self.cfg.nxmax[i] for i in range(4))

    def _check(self, rc):
        if rc:
            raise Vlp4dError(rc, self.lib.vlp_last_error(self.h).decode())

    def close(self):
        """Release the context.  Raises (and keeps the context) while it is a
        member of a live Ensemble, whose captured graph uses its buffers."""
        if getattr(self, "h", None):
            self._check(self.lib.vlp_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def iter(self):
        return self.lib.vlp_iter(self.h)

    def step(self, n=1):
        self._check(self.lib.vlp_step(self.h, n))

    # phases of onetimestep (VMO/Timestep.hpp:41-96)
    def spline_xy(self):
        self._check(self.lib.vlp_spline_xy(self.h))

    def advect2d_poisson(self):
        self._check(self.lib.vlp_advect2d_poisson(self.h))

    def spline_vxvy(self):
        self._check(self.lib.vlp_spline_vxvy(self.h))

    def advect4d_poisson(self):
        self._check(self.lib.vlp_advect4d_poisson(self.h))

    def get_f(self):
        """Interior f (x fastest) as an array of shape (Nvy, Nvx, Ny, Nx)."""
        n = int(np.prod(self.shape4))
        out = np.empty(n, dtype=np.float64)
        self._check(self.lib.vlp_get_f(self.h, _dptr(out), n))
        return out.reshape(self.shape4[::-1])

    def set_f(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).ravel()
        self._check(self.lib.vlp_set_f(self.h, _dptr(a), a.size))

    def write_fxvx(self, directory="."):
        """fxvx (and, for the spline solver, fyvy) cross sections of the
        current f as CSV files in an existing directory."""
        self._check(self.lib.vlp_write_fxvx(self.h, os.fsencode(directory)))

    def adjoint_step(self, g):
        """Vector-Jacobian product of one step at the current f (FP64
        contexts): (d f^{n+1} / d f^n)^T g, same shape as get_f().  The
        context is not advanced."""
        a = np.ascontiguousarray(g, dtype=np.float64).ravel()
        out = np.empty_like(a)
        self._check(self.lib.vlp_adjoint_step(self.h, _dptr(a), _dptr(out), a.size))
        return out.reshape(self.shape4[::-1])

    def adjoint_steps(self, nsteps, g):
        """Vector-Jacobian product of nsteps steps from the current f (FP64
        contexts, checkpointed reverse sweep); the context ends where it
        started."""
        a = np.ascontiguousarray(g, dtype=np.float64).ravel()
        out = np.empty_like(a)
        self._check(self.lib.vlp_adjoint_steps(self.h, int(nsteps), _dptr(a), _dptr(out), a.size))
        return out.reshape(self.shape4[::-1])

    def field_energy_gradient(self, nsteps, vweight=None, checkpoint_every=0):
        """(J, dJ/df) for J = dt sum_{k=1..nsteps} sum_{x,y} |E(f^{n+k})|^2 dx dy
        (FP64 contexts).  With vweight (array [Nvy][Nvx]) the gradient is the
        moment sum_v w(v) dJ/df(x, y, v), shape [Ny][Nx]; otherwise the full
        [Nvy][Nvx][Ny][Nx] array.  checkpoint_every: states kept by the
        forward sweep (0: ceil(sqrt(nsteps))).  The context ends where it
        started."""
        J = ctypes.c_double()
        nx, ny = self.shape4[0], self.shape4[1]
        if vweight is None:
            out = np.empty(self.shape4[::-1])
            self._check(self.lib.vlp_adjoint_field_energy(self.h, int(nsteps), int(checkpoint_every), None, ctypes.byref(J), _dptr(out),
                                                          out.size))
        else:
            w = np.ascontiguousarray(vweight, dtype=np.float64)
            assert w.shape == (self.shape4[3], self.shape4[2])
            out = np.empty((ny, nx))
            self._check(self.lib.vlp_adjoint_field_energy(self.h, int(nsteps), int(checkpoint_every), _dptr(w), ctypes.byref(J), _dptr(out),
                                                          out.size))
        return J.value, out

    def buffer(self, which):
        n = self.lib.vlp_buffer_size(self.h, which)
        out = np.empty(n, dtype=np.float64)
        self._check(self.lib.vlp_get_buffer(self.h, which, _dptr(out), n))
        N = self.shape4
        if which == BUF_CXY:
            return out.reshape(N[3], N[2], N[1] + 4, N[0] + 4)
        if which == BUF_C4:
            return out.reshape(N[3] + 4, N[2] + 4, N[1] + 4, N[0] + 4)
        if which == BUF_F:
            return out.reshape(N[::-1])
        return out.reshape(N[1], N[0])

    def diags(self, n=None):
        """dict of nrj, nrjx, nrjy, mass, l2norm for iterations 0..n-1."""
        n = self.iter + 1 if n is None else n
        arrs = [np.zeros(n) for _ in range(5)]
        self._check(self.lib.vlp_diags(self.h, *[_dptr(a) for a in arrs], n))
        return dict(zip(["nrj", "nrjx", "nrjy", "mass", "l2norm"], arrs))

    def save(self, directory="."):
        self._check(self.lib.vlp_save(self.h, os.fsencode(directory)))

    def kernel_variants(self):
        """Which kernels this context runs (bit layout in include/vlp4d.h at
        vlp_kernel_variants): e.g. bit 7 the y-streaming 4-D advection, bits
        8-11 its vx tile, bit 12 the fused x/y spline + 2-D advection."""
        return self.lib.vlp_kernel_variants(self.h)

    def autotune(self, csv_path=None):
        """Time every workgroup shape of the 4-D advection this geometry
        admits on this context's buffers and keep the fastest; returns it
        as (vx_tile, plane_buffers)."""
        self._check(self.lib.vlp_autotune(self.h, os.fsencode(csv_path) if csv_path else None))
        return self.tuned_tile()

    def tuned_tile(self):
        t, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.vlp_tuned_tile(self.h, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def timers(self):
        out = {}
        for k, lab in enumerate(TIMER_LABELS):
            s = ctypes.c_double()
            c = ctypes.c_int()
            self.lib.vlp_timer(self.h, k, ctypes.byref(s), ctypes.byref(c))
            out[lab] = (s.value, c.value)
        return out


class Ensemble:
    """Independent simulations (Vlp4d contexts on one device, any grids and
    precisions) advanced together through one captured HIP graph per two
    steps.  Close the ensemble before its members (closing a member of a
    live ensemble raises)."""

    def __init__(self, members):
        self.lib = load()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = ctypes.c_void_p()
        rc = self.lib.vlp_ensemble_create(arr, len(self.members), ctypes.byref(h))
        if rc:
            raise Vlp4dError(rc, self.lib.vlp_last_error(None).decode())
        self.h = h

    def step(self, nsteps=1):
        rc = self.lib.vlp_ensemble_step(self.h, int(nsteps))
        if rc:
            raise Vlp4dError(rc, self.lib.vlp_last_error(None).decode())

    def __len__(self):
        return self.lib.vlp_ensemble_size(self.h)

    def close(self):
        if self.h:
            self.lib.vlp_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
