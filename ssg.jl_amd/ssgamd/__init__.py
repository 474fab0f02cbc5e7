"""Python mirror of SSG.jl's time-step API over the HIP C ABI (include/ssg.h).

The names follow the reference so that host code and tests read like its own:

=========================================  ==========================================
reference (SSG.jl)                          here
=========================================  ==========================================
``make_domain`` (src/domain.jl:217-230)     :func:`make_domain` -> :class:`Domain`
``allocate_fields`` (src/fields.jl:72)      :func:`allocate_fields` -> :class:`Fields`
``TimeParams`` (src/timestep.jl:19-40)      :class:`TimeParams`
``TimeState`` (src/timestep.jl:45-86)       :class:`TimeState`
``timestep!`` (src/timestep.jl:267-285)     :func:`timestep`
``compute_tendency!`` (src/nonlinear.jl:22) :func:`compute_tendency`
``compute_jacobian!`` (src/nonlinear.jl:42) :func:`compute_jacobian`
``rfft_2d!``/``irfft_2d!`` (transforms:42)  :func:`rfft_2d` / :func:`irfft_2d`
``apply_spectral_filter!`` (timestep:247)   :func:`apply_spectral_filter`
=========================================  ==========================================

All compute runs in ``libssg_hip.so``; there is no CPU fallback.  Importing
works without a GPU (so the ABI can be inspected), but creating a domain
needs the library and a HIP device and raises otherwise.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# SSG_LIB selects another build of the library (e.g. one compiled with
# different launch constants) for side-by-side measurement
LIB_PATH = os.environ.get("SSG_LIB") or os.path.normpath(os.path.join(_PKG, "..", "lib", "libssg_hip.so"))

SSG_OK = 0
SSG_ERR_ARG = -1
SSG_ERR_HIP = -2
SSG_ERR_NONFINITE = -3
SSG_ERR_UNSUPPORTED = -4
SSG_ERR_NOT_CONVERGED = -5

SSG_FP64 = 0
SSG_FP32_MIXED = 1

FIELD_B, FIELD_PHI, FIELD_U, FIELD_V = 0, 1, 2, 3

# Every symbol include/ssg.h declares
EXPORTED_SYMBOLS = (
    "ssg_create", "ssg_destroy", "ssg_last_error", "ssg_sync", "ssg_set_field", "ssg_get_field",
    "ssg_set_field_device", "ssg_get_field_device", "ssg_timestep", "ssg_tendency", "ssg_jacobian",
    "ssg_rfft_2d", "ssg_irfft_2d", "ssg_apply_exponential_filter", "ssg_diagnostics", "ssg_last_timing",
    "ssg_profile", "ssg_profile_read", "ssg_tuning", "ssg_ma_solve", "ssg_get_phi3d", "ssg_set_coupling",
    "ssg_last_ma_diag", "ssg_copy_probe", "ssg_adjoint", "ssg_ensemble_create", "ssg_ensemble_destroy",
    "ssg_ensemble_member", "ssg_ensemble_last_error", "ssg_ensemble_timestep", "ssg_ensemble_tuning",
    "ssg_ma_adjoint", "ssg_adjoint_coupled", "ssg_set_b", "ssg_cfl", "ssg_state_size", "ssg_save_state",
    "ssg_load_state", "ssg_timestep_prepare", "ssg_tendency_noalloc", "ssg_set_vertical_grid",
    "ssg_vertical_velocity",
)

MA_BODY = 0      # as wired: b on every level, surface slope 0
MA_SURFACE = 1   # documented physics: surface slope = b
MA_STAGE_TOL32 = 1e-6   # relative residual of the coupled step's FP32 stage solves (SSG_MA_STAGE_TOL32)


class TimeScheme(enum.IntEnum):
    """``@enum TimeScheme`` (src/timestep.jl:10-14)."""
    AB2_LowStorage = 0
    RK3 = 1
    RK3_LowStorage = 2


AB2_LowStorage = TimeScheme.AB2_LowStorage
RK3 = TimeScheme.RK3
RK3_LowStorage = TimeScheme.RK3_LowStorage


class SSGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ssg error {code}: {msg}")
        self.code = code


class _DomainDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
                ("precision", ctypes.c_int32), ("device", ctypes.c_int32)]


class _TimeParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("scheme", ctypes.c_int32), ("filter_freq", ctypes.c_int32),
                ("filter_strength", ctypes.c_double), ("filter_order", ctypes.c_int32),
                ("filter_cutoff", ctypes.c_double), ("ab2_reference_aliasing", ctypes.c_int32),
                ("check_finite", ctypes.c_int32)]


class _TimeState(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("step", ctypes.c_int64), ("dt_actual", ctypes.c_double)]


class _Diag(ctypes.Structure):
    _fields_ = [("sum_b", ctypes.c_double), ("sum_b2", ctypes.c_double), ("parseval_b", ctypes.c_double),
                ("ke_surface", ctypes.c_double), ("nonfinite", ctypes.c_int64), ("ke_true", ctypes.c_double)]


class _MaParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("tol", ctypes.c_double), ("maxiter", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("warm", ctypes.c_int32), ("stage_fp32", ctypes.c_int32)]


class _MaDiag(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("iterations", ctypes.c_int32), ("final_residual", ctypes.c_double),
                ("relative_residual", ctypes.c_double), ("history", ctypes.c_double * 64),
                ("solve_ms", ctypes.c_double), ("iterations_fp32", ctypes.c_int32), ("not_elliptic", ctypes.c_int32)]

    def as_dict(self):
        n = min(self.iterations + 1, 64)
        return {"converged": bool(self.converged), "iterations": self.iterations,
                "final_residual": self.final_residual, "relative_residual": self.relative_residual,
                "convergence_history": list(self.history[:n]), "solve_time": self.solve_ms / 1e3,
                "iterations_fp32": self.iterations_fp32, "not_elliptic": bool(self.not_elliptic)}


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load_library(path: str | None = None):
    """Load libssg_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    lib.ssg_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_DomainDesc)]
    lib.ssg_destroy.argtypes = [vp]
    lib.ssg_last_error.restype = ctypes.c_char_p
    lib.ssg_last_error.argtypes = [vp]
    lib.ssg_sync.argtypes = [vp]
    for n in ("ssg_set_field", "ssg_get_field", "ssg_set_field_device", "ssg_get_field_device"):
        getattr(lib, n).argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int64]
    lib.ssg_timestep.argtypes = [vp, ctypes.POINTER(_TimeParams), ctypes.POINTER(_TimeState), ctypes.c_int64]
    lib.ssg_timestep_prepare.argtypes = [vp, ctypes.POINTER(_TimeParams), ctypes.POINTER(_TimeState),
                                         ctypes.c_int64]
    lib.ssg_tendency.argtypes = [vp, _dp, ctypes.c_int32]
    lib.ssg_set_vertical_grid.argtypes = [vp, _dp, _dp, ctypes.c_int32]
    lib.ssg_vertical_velocity.argtypes = [vp, ctypes.POINTER(_MaParams), _dp, ctypes.c_int64, ctypes.POINTER(_MaDiag)]
    lib.ssg_tendency_noalloc.argtypes = [ctypes.POINTER(_DomainDesc), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    lib.ssg_jacobian.argtypes = [vp, _dp]
    lib.ssg_rfft_2d.argtypes = [vp, _dp, _dp]
    lib.ssg_irfft_2d.argtypes = [vp, _dp, _dp]
    lib.ssg_apply_exponential_filter.argtypes = [vp, ctypes.c_double, ctypes.c_int32, ctypes.c_double]
    lib.ssg_diagnostics.argtypes = [vp, ctypes.POINTER(_Diag)]
    lib.ssg_last_timing.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int64)]
    lib.ssg_profile.argtypes = [vp, ctypes.c_int32]
    lib.ssg_tuning.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _dp]
    lib.ssg_ma_solve.argtypes = [vp, ctypes.POINTER(_MaParams), ctypes.POINTER(_MaDiag)]
    lib.ssg_get_phi3d.argtypes = [vp, _dp, ctypes.c_int64]
    lib.ssg_set_coupling.argtypes = [vp, ctypes.POINTER(_MaParams)]
    lib.ssg_last_ma_diag.argtypes = [vp, ctypes.POINTER(_MaDiag)]
    lib.ssg_copy_probe.argtypes = [vp, ctypes.c_int64, ctypes.c_int32]
    lib.ssg_adjoint.argtypes = [vp, ctypes.POINTER(_TimeParams), ctypes.c_int64, ctypes.c_int64, _dp, _dp]
    lib.ssg_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(_DomainDesc)]
    lib.ssg_ensemble_destroy.argtypes = [vp]
    lib.ssg_ensemble_member.argtypes = [vp, ctypes.c_int32]
    lib.ssg_ensemble_last_error.argtypes = [vp]
    lib.ssg_ensemble_timestep.argtypes = [vp, ctypes.POINTER(_TimeParams), _dp, ctypes.POINTER(_TimeState),
                                          ctypes.c_int64]
    lib.ssg_ensemble_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), _dp]
    lib.ssg_profile_read.argtypes = [vp, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int64), _dp]
    lib.ssg_ma_adjoint.argtypes = [vp, ctypes.POINTER(_MaParams), _dp, _dp, ctypes.POINTER(_MaDiag)]
    lib.ssg_set_b.argtypes = [vp, _dp, ctypes.c_int64, ctypes.POINTER(_MaParams), ctypes.POINTER(_MaDiag)]
    lib.ssg_cfl.argtypes = [vp, ctypes.c_double, _dp, _dp, _dp]
    lib.ssg_state_size.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.ssg_save_state.argtypes = [vp, ctypes.POINTER(_TimeState), vp, ctypes.c_size_t]
    lib.ssg_load_state.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(_TimeState)]
    lib.ssg_adjoint_coupled.argtypes = [vp, ctypes.POINTER(_TimeParams), ctypes.POINTER(_MaParams), ctypes.c_int64,
                                        ctypes.c_int64, ctypes.c_int32, _dp, _dp]
    for n in EXPORTED_SYMBOLS:
        getattr(lib, n).restype = ctypes.c_char_p if n in ("ssg_last_error", "ssg_ensemble_last_error") else ctypes.c_int
    lib.ssg_ensemble_member.restype = vp
    if path is None:
        _lib = lib
    return lib


def _f64(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Domain:
    """Grid + device context (``Domain``, src/domain.jl:33-186).

    Real fields are numpy arrays of shape ``(ny, nx)`` (x fastest, the
    reference's column-major ``(nx, ny)``).  Spectra in the reference layout
    are complex ``(ny//2+1, nx)`` arrays.
    """

    def __init__(self, nx: int, ny: int, nz: int = 1, Lx: float = 2 * np.pi, Ly: float = 2 * np.pi,
                 Lz: float = 1.0, precision: int = SSG_FP64, device: int = 0, _borrowed=None):
        self.lib = load_library()
        self.Nx, self.Ny, self.Nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.precision = int(precision)
        self.coupled = False
        self._owned = _borrowed is None
        if _borrowed is not None:
            h = ctypes.c_void_p(_borrowed)
        else:
            desc = _DomainDesc(self.Nx, self.Ny, self.Nz, self.Lx, self.Ly, self.Lz, self.precision, int(device))
            h = ctypes.c_void_p()
            rc = self.lib.ssg_create(ctypes.byref(h), ctypes.byref(desc))
            if rc != SSG_OK:
                raise SSGError(rc, "ssg_create failed (see stderr)")
        self._h = h
        self.nyc = self.Ny // 2 + 1
        # src/domain.jl:145-162
        self.x = (self.Lx / self.Nx) * np.arange(self.Nx)
        self.y = (self.Ly / self.Ny) * np.arange(self.Ny)
        self.kx = 2 * np.pi * np.fft.fftfreq(self.Nx, self.Lx / self.Nx)
        self.ky = 2 * np.pi * np.fft.rfftfreq(self.Ny, self.Ly / self.Ny)

    @property
    def handle(self):
        return self._h

    def check(self, rc: int):
        if rc != SSG_OK:
            msg = self.lib.ssg_last_error(self._h)
            raise SSGError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                self.lib.ssg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- field transfer ----
    def set_field(self, field: int, a):
        a = _f64(a, (self.Ny, self.Nx))
        self.check(self.lib.ssg_set_field(self._h, field, a.ctypes.data, a.size))

    def get_field(self, field: int) -> np.ndarray:
        out = np.empty((self.Ny, self.Nx))
        self.check(self.lib.ssg_get_field(self._h, field, out.ctypes.data, out.size))
        return out

    def set_field_device(self, field: int, ptr: int, n: int):
        self.check(self.lib.ssg_set_field_device(self._h, field, ctypes.c_void_p(ptr), n))

    def get_field_device(self, field: int, ptr: int, n: int):
        self.check(self.lib.ssg_get_field_device(self._h, field, ctypes.c_void_p(ptr), n))

    def profile(self, enable: bool):
        self.check(self.lib.ssg_profile(self._h, int(bool(enable))))

    def profile_read(self, kind: int):
        """(milliseconds, launches, algorithmic bytes) of one kernel class."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.lib.ssg_profile_read(self._h, int(kind), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by)))
        return ms.value, n.value, by.value

    def tuning(self, path: int | None = None):
        """Autotuned elements-per-thread per kernel class and candidate timings (us)."""
        path = (1 if self.precision == SSG_FP32_MIXED else 0) if path is None else path
        ept = (ctypes.c_int32 * 3)()
        us = (ctypes.c_double * 9)()
        self.check(self.lib.ssg_tuning(self._h, int(path), ept, us))
        names = ("k_rows_fwd", "k_cols", "k_rows_inv")
        return {names[k]: {"ept": ept[k], "cand_us": {c: us[3 * k + i] for i, c in enumerate((4, 8, 16))}}
                for k in range(3)}

    def sync(self):
        self.check(self.lib.ssg_sync(self._h))

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.lib.ssg_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def make_domain(Nx: int, Ny: int, Nz: int = 1, *, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0,
                precision: int = SSG_FP64, device: int = 0) -> Domain:
    """``make_domain`` (src/domain.jl:217-230) on one GPU."""
    return Domain(Nx, Ny, Nz, Lx, Ly, Lz, precision, device)


class Fields:
    """Device-resident ``Fields`` (src/fields.jl:28-60): ``bs`` and ``phis``
    read and write through the ABI."""

    def __init__(self, domain: Domain):
        self.domain = domain

    @property
    def bs(self) -> np.ndarray:
        return self.domain.get_field(FIELD_B)

    @bs.setter
    def bs(self, a):
        self.domain.set_field(FIELD_B, a)

    @property
    def phis(self) -> np.ndarray:
        return self.domain.get_field(FIELD_PHI)

    @phis.setter
    def phis(self, a):
        self.domain.set_field(FIELD_PHI, a)

    @property
    def u(self) -> np.ndarray:
        return self.domain.get_field(FIELD_U)

    @property
    def v(self) -> np.ndarray:
        return self.domain.get_field(FIELD_V)


def allocate_fields(domain: Domain) -> Fields:
    """``allocate_fields`` (src/fields.jl:72-116)."""
    return Fields(domain)


@dataclass
class TimeParams:
    """``TimeParams`` (src/timestep.jl:19-40).  ``filter_order`` and
    ``filter_cutoff`` default to the legacy path's fixed 4 and 0.65
    (src/filter.jl:525-532)."""
    dt: float
    scheme: TimeScheme = AB2_LowStorage
    filter_freq: int = 10
    filter_strength: float = 0.1
    cfl_safety: float = 0.5
    max_dt: float = 0.1
    min_dt: float = 1e-6
    adaptive_dt: bool = False
    filter_order: int = 4
    filter_cutoff: float = 0.65
    ab2_reference_aliasing: bool = False
    check_finite: bool = False

    def _c(self, dt: float | None = None) -> _TimeParams:
        return _TimeParams(float(self.dt if dt is None else dt), int(self.scheme), int(self.filter_freq),
                           float(self.filter_strength), int(self.filter_order), float(self.filter_cutoff),
                           int(bool(self.ab2_reference_aliasing)), int(bool(self.check_finite)))


class TimeState:
    """``TimeState`` scalars (src/timestep.jl:45-86); the stage buffers and
    the AB2 history live on the device."""

    def __init__(self, initial_time: float = 0.0, fields: Fields | None = None):
        self._c = _TimeState(float(initial_time), 0, 0.0)
        self.cfl_max = 0.0

    @property
    def t(self):
        return self._c.t

    @t.setter
    def t(self, v):
        self._c.t = float(v)

    @property
    def step(self):
        return self._c.step

    @step.setter
    def step(self, v):
        self._c.step = int(v)

    @property
    def dt_actual(self):
        return self._c.dt_actual

    @dt_actual.setter
    def dt_actual(self, v):
        self._c.dt_actual = float(v)


def adaptive_timestep(fields: Fields, domain: Domain, params: TimeParams, state: TimeState) -> float:
    """``adaptive_timestep`` (src/timestep.jl:122-147): params.dt scaled by the
    CFL number of the current surface velocities."""
    if not params.adaptive_dt:
        return params.dt
    cfl = compute_cfl_number(fields, domain, params.dt)
    state.cfl_max = cfl
    if cfl > params.cfl_safety:
        return max(params.dt * params.cfl_safety / cfl, params.min_dt)
    if cfl < 0.5 * params.cfl_safety:
        return min(params.dt * 1.1, params.max_dt)
    return params.dt


def timestep(fields: Fields, domain: Domain, params: TimeParams, state: TimeState, nsteps: int = 1) -> float:
    """``timestep!`` (src/timestep.jl:267-285), ``nsteps`` times.  Returns dt.

    With ``params.adaptive_dt`` each step's dt comes from
    :func:`adaptive_timestep`; with φₛ frozen the velocities, and so dt, do
    not change within a call, so the call stays one batch on the device."""
    if params.adaptive_dt and nsteps > 0:
        if domain.coupled:
            for _ in range(int(nsteps)):
                timestep(fields, domain, TimeParams(adaptive_timestep(fields, domain, params, state),
                                                    scheme=params.scheme, filter_freq=params.filter_freq,
                                                    filter_strength=params.filter_strength,
                                                    filter_order=params.filter_order,
                                                    filter_cutoff=params.filter_cutoff,
                                                    ab2_reference_aliasing=params.ab2_reference_aliasing,
                                                    check_finite=params.check_finite), state, 1)
            return state.dt_actual
        p = params._c(adaptive_timestep(fields, domain, params, state))
    else:
        p = params._c()
    domain.check(domain.lib.ssg_timestep(domain.handle, ctypes.byref(p), ctypes.byref(state._c), int(nsteps)))
    return state.dt_actual


def timestep_prepare(domain: Domain, params: TimeParams, state: TimeState, nsteps: int = 1) -> None:
    """Launch set-up of a following ``timestep(..., nsteps)`` from ``state``:
    autotuning and the step graphs it will replay, without stepping (no
    reference counterpart)."""
    p = params._c()
    domain.check(domain.lib.ssg_timestep_prepare(domain.handle, ctypes.byref(p), ctypes.byref(state._c),
                                                 int(nsteps)))


def compute_tendency(fields: Fields, domain: Domain, ab2_reference_aliasing: bool = False) -> np.ndarray:
    """``compute_tendency!`` (src/nonlinear.jl:22-36) of the current fields."""
    out = np.empty((domain.Ny, domain.Nx))
    domain.check(domain.lib.ssg_tendency(domain.handle, out.ctypes.data_as(_dp), int(ab2_reference_aliasing)))
    return out


def compute_tendency_noalloc(b_ptr: int, phi_ptr: int, out_ptr: int, Nx: int, Ny: int, *, Lx=2 * np.pi,
                             Ly=2 * np.pi, ab2_reference_aliasing: bool = False, stream: int = 0,
                             device: int = 0) -> None:
    """``compute_tendency!`` of device-resident fields without any device
    allocation (``ssg_tendency_noalloc``): ``b_ptr``, ``phi_ptr`` and
    ``out_ptr`` are device addresses of Nx*Ny FP64 arrays (x fastest); the
    kernels are enqueued on ``stream`` and the call does not synchronise."""
    lib = load_library()
    desc = _DomainDesc(int(Nx), int(Ny), 1, float(Lx), float(Ly), 1.0, SSG_FP64, int(device))
    rc = lib.ssg_tendency_noalloc(ctypes.byref(desc), b_ptr, phi_ptr, out_ptr, int(ab2_reference_aliasing), stream)
    if rc != SSG_OK:
        raise SSGError(rc, lib.ssg_last_error(None).decode(errors="replace"))


def compute_jacobian(fields: Fields, domain: Domain) -> np.ndarray:
    """``compute_jacobian!`` (src/nonlinear.jl:42-57): -J(phi, b)."""
    out = np.empty((domain.Ny, domain.Nx))
    domain.check(domain.lib.ssg_jacobian(domain.handle, out.ctypes.data_as(_dp)))
    return out


def rfft_2d(domain: Domain, a) -> np.ndarray:
    """``rfft_2d!`` (src/transforms.jl:42-45): reference layout (ny//2+1, nx)."""
    a = _f64(a, (domain.Ny, domain.Nx))
    out = np.empty((domain.nyc, domain.Nx), dtype=np.complex128)
    domain.check(domain.lib.ssg_rfft_2d(domain.handle, a.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    return out


def irfft_2d(domain: Domain, s) -> np.ndarray:
    """``irfft_2d!`` (src/transforms.jl:53-56)."""
    s = np.ascontiguousarray(s, dtype=np.complex128).reshape(domain.nyc, domain.Nx)
    out = np.empty((domain.Ny, domain.Nx))
    domain.check(domain.lib.ssg_irfft_2d(domain.handle, s.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    return out


def apply_spectral_filter(fields: Fields, domain: Domain, strength: float, order: int = 4,
                          cutoff: float = 0.65) -> None:
    """``apply_spectral_filter!`` (src/timestep.jl:247-260) on ``bs``."""
    domain.check(domain.lib.ssg_apply_exponential_filter(domain.handle, float(strength), int(order), float(cutoff)))


def diagnostics(domain: Domain) -> dict:
    d = _Diag()
    domain.check(domain.lib.ssg_diagnostics(domain.handle, ctypes.byref(d)))
    return {"sum_b": d.sum_b, "sum_b2": d.sum_b2, "parseval_b": d.parseval_b, "ke_surface": d.ke_surface,
            "nonfinite": d.nonfinite, "ke_true": d.ke_true}


def _ma_params(eps, tol, maxiter, variant, warm, stage_fp32=False):
    return _MaParams(float(eps), float(tol), int(maxiter), int(variant), int(warm), int(bool(stage_fp32)))


def solve_monge_ampere(fields: Fields, domain: Domain, *, eps: float = 0.1, tol: float = 1e-10, maxiter: int = 50,
                       variant: int = MA_BODY, warm: bool = False) -> dict:
    """``solve_monge_ampere_fields!`` (src/poisson.jl:1068-1090) on the GPU.

    Solves the reference residual operator for the current ``bs`` and sets
    ``phis`` (and the surface velocities) from the top level.  Returns the
    reference's diagnostics tuple as a dict; ``converged`` is False (no
    exception) when ``maxiter`` is reached, as in ``solve_ssg_equation``."""
    p = _ma_params(eps, tol, maxiter, variant, warm)
    d = _MaDiag()
    rc = domain.lib.ssg_ma_solve(domain.handle, ctypes.byref(p), ctypes.byref(d))
    if rc not in (SSG_OK, SSG_ERR_NOT_CONVERGED):
        domain.check(rc)
    out = d.as_dict()
    out["epsilon_parameter"] = float(eps)
    return out


def make_vertical_grid(Nz: int, Lz: float = 1.0, z_grid: str = "uniform", z_coords=None):
    """``make_vertical_grid`` (src/domain.jl:238-281) for the non-periodic
    grids: ``uniform`` (z = range(0, Lz, Nz), dz = Lz/(Nz-1) with the last
    entry duplicated) and ``custom`` (z = z_coords; dz centred inside, one-sided
    at both ends).  Returns (z, dz)."""
    if z_grid == "uniform":
        z = np.linspace(0.0, Lz, Nz)
        dz = np.full(Nz, Lz / (Nz - 1))
        return z, dz
    if z_grid == "custom":
        if z_coords is None:
            raise ValueError("For :custom grid, provide z_coords")
        z = np.array(z_coords, dtype=np.float64)
        if z.size != Nz:
            raise ValueError(f"Custom z_coords length ({z.size}) must match Nz ({Nz})")
        dz = np.zeros(Nz)
        dz[1:-1] = 0.5 * (z[2:] - z[:-2])
        dz[0] = z[1] - z[0]
        dz[-1] = z[-1] - z[-2]
        return z, dz
    raise ValueError(f"Unknown z_grid type: {z_grid}. Use uniform or custom")


def set_vertical_grid(domain: Domain, z, dz=None) -> None:
    """The Monge-Ampère problem's vertical grid (``ssg_set_vertical_grid``):
    levels z and spacings dz as a Domain holds them (dz None: the :custom
    rule applied to z)."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    dzp = None
    if dz is not None:
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        dzp = dz.ctypes.data_as(_dp)
    domain.check(domain.lib.ssg_set_vertical_grid(domain.handle, z.ctypes.data_as(_dp), dzp, int(z.size)))
    domain._z = z.copy()
    domain._dz = dz.copy() if dz is not None else make_vertical_grid(z.size, float(z[-1] - z[0]), "custom", z)[1]


def compute_vertical_velocity(fields: Fields, domain: Domain, *, eps: float = 0.1, tol: float = 1e-10,
                              maxiter: int = 50, variant: int = MA_SURFACE) -> tuple:
    """``compute_vertical_velocity`` (src/extras/omega_eq.jl:574-652) of the
    current surface buoyancy: returns (w with shape (nz, ny, nx), the
    Monge-Ampère solve's diagnostics)."""
    p = _ma_params(eps, tol, maxiter, variant, 0)
    d = _MaDiag()
    w = np.empty((domain.Nz, domain.Ny, domain.Nx))
    domain.check(domain.lib.ssg_vertical_velocity(domain.handle, ctypes.byref(p), w.ctypes.data_as(_dp), w.size,
                                                  ctypes.byref(d)))
    return w, d.as_dict()


def get_phi3d(domain: Domain) -> np.ndarray:
    """``fields.φ`` after a Monge-Ampère solve, shape (nz, ny, nx)."""
    out = np.empty((domain.Nz, domain.Ny, domain.Nx))
    domain.check(domain.lib.ssg_get_phi3d(domain.handle, out.ctypes.data_as(_dp), out.size))
    return out


def set_coupling(domain: Domain, *, eps: float | None = 0.1, tol: float = 1e-10, maxiter: int = 50,
                 variant: int = MA_BODY, warm: int = 2, stage_fp32: bool = False) -> None:
    """Coupled stepping: re-solve the Monge-Ampère problem at every RK stage
    (``eps=None`` restores the reference's frozen φₛ).  ``warm``: 0 cold
    starts, 1 starts each solve from the last D(Φ), 2 extrapolates over the
    same stage of the last three steps.  ``stage_fp32`` (FP32-mixed domains):
    the solves inside a ``timestep`` call run in FP32 to a relative residual
    of ``MA_STAGE_TOL32``; the solve for the b the call returns stays FP64 to
    ``tol``."""
    if eps is None:
        domain.check(domain.lib.ssg_set_coupling(domain.handle, None))
        domain.coupled = False
        return
    p = _ma_params(eps, tol, maxiter, variant, warm, stage_fp32)
    domain.check(domain.lib.ssg_set_coupling(domain.handle, ctypes.byref(p)))
    domain.coupled = True


def set_b(fields: Fields, domain: Domain, b, *, eps: float | None = 0.1, tol: float = 1e-10, maxiter: int = 20,
          variant: int = MA_BODY) -> dict | None:
    """``set_b!`` (src/setICs.jl:21-88): the k = 0 mode of ``b`` removed and the
    2/3 mask applied, then (``eps`` not None) the Monge-Ampère solve for φₛ and
    the surface velocities.  Returns the solve's diagnostics (None without a
    solve); ``converged`` is False, without an exception, when ``maxiter`` is
    reached, as in the reference."""
    b = _f64(b, (domain.Ny, domain.Nx))
    d = _MaDiag()
    p = None if eps is None else _ma_params(eps, tol, maxiter, variant, False)
    rc = domain.lib.ssg_set_b(domain.handle, b.ctypes.data_as(_dp), b.size,
                              None if p is None else ctypes.byref(p), ctypes.byref(d))
    if rc not in (SSG_OK, SSG_ERR_NOT_CONVERGED):
        domain.check(rc)
    if p is None:
        return None
    out = d.as_dict()
    out["epsilon_parameter"] = float(eps)
    return out


def compute_cfl_number(fields: Fields, domain: Domain, dt: float) -> float:
    """``compute_cfl_number`` (src/timestep.jl:91-117) from the surface
    velocities: max(max|u| dt/dx, max|v| dt/dy)."""
    cfl, um, vm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    domain.check(domain.lib.ssg_cfl(domain.handle, float(dt), ctypes.byref(cfl), ctypes.byref(um), ctypes.byref(vm)))
    return cfl.value


def save_state(domain: Domain, state: "TimeState") -> bytes:
    """The exact-restart state (``ssg_save_state``) as bytes: b, φₛ, the AB2
    history, t and step, the last Monge-Ampère solve's Φ̂ and, in coupled
    mode, the coupling parameters and warm-start history.  Loading it into a
    context of the same grid and precision resumes bitwise, frozen or
    coupled."""
    n = ctypes.c_size_t()
    domain.check(domain.lib.ssg_state_size(domain.handle, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    domain.check(domain.lib.ssg_save_state(domain.handle, ctypes.byref(state._c), buf, n.value))
    return buf.raw


_STATE_HEADER_BYTES = 56


def load_state(domain: Domain, blob: bytes, state: "TimeState") -> None:
    buf = ctypes.create_string_buffer(blob, len(blob))
    domain.check(domain.lib.ssg_load_state(domain.handle, buf, len(blob), ctypes.byref(state._c)))
    # the blob restores the coupling mode; mirror it
    n2 = domain.Nx * domain.Ny
    ma = _STATE_HEADER_BYTES + n2 * (16 + (4 if domain.precision == SSG_FP32_MIXED else 8))
    if blob[:8] in (b"SSGSTAT2", b"SSGSTAT3") and len(blob) >= ma + 28:
        domain.coupled = bool(int.from_bytes(blob[ma + 24:ma + 28], "little"))


def save_checkpoint(path: str, fields: Fields, domain: Domain, state: "TimeState",
                    params: "TimeParams | None" = None) -> str:
    """``save_simulation_state_full`` (src/output.jl:318-448) with the
    reference's keys (time, step, dt, grid/*, fields/surface/*, fields/3d/*,
    timestepper/*), written as a numpy ``.npz`` archive (no JLD2/HDF5 writer
    is available here), plus ``restart/state``, the exact-restart bytes of
    :func:`save_state`.  Arrays keep this library's (Ny, Nx[, Nz-first])
    C order, which is the reference's column-major (Nx, Ny[, Nz]) memory."""
    d = domain
    z, dz = (d._z, d._dz) if getattr(d, "_z", None) is not None else make_vertical_grid(d.Nz, d.Lz)
    out = {
        "time": np.float64(state.t), "step": np.int64(state.step),
        "dt": np.float64(state.dt_actual if state.dt_actual > 0 else (params.dt if params else 0.0)),
        "grid/Nx": np.int64(d.Nx), "grid/Ny": np.int64(d.Ny), "grid/Nz": np.int64(d.Nz),
        "grid/Lx": np.float64(d.Lx), "grid/Ly": np.float64(d.Ly), "grid/Lz": np.float64(d.Lz),
        "grid/x": d.Lx / d.Nx * np.arange(d.Nx), "grid/y": d.Ly / d.Ny * np.arange(d.Ny),
        "grid/z": np.asarray(z), "grid/dz": np.asarray(dz),
        "fields/surface/buoyancy": fields.bs, "fields/surface/streamfunction": fields.phis,
        "restart/state": np.frombuffer(save_state(d, state), dtype=np.uint8),
    }
    u, v = fields.u, fields.v
    out["fields/3d/u_velocity"] = np.broadcast_to(u, (d.Nz,) + u.shape).copy()   # every level (src/setICs.jl:161-192)
    out["fields/3d/v_velocity"] = np.broadcast_to(v, (d.Nz,) + v.shape).copy()
    try:
        out["fields/3d/streamfunction"] = get_phi3d(d)
    except SSGError:
        pass   # no Monge-Ampère solve yet
    if params is not None:
        out.update({"timestepper/scheme": np.array(["AB2_LowStorage", "RK3", "RK3_LowStorage"][params.scheme]),
                    "timestepper/dt": np.float64(params.dt), "timestepper/adaptive_dt": np.bool_(params.adaptive_dt),
                    "timestepper/filter_freq": np.int64(params.filter_freq),
                    "timestepper/filter_strength": np.float64(params.filter_strength),
                    "timestepper/cfl_safety": np.float64(params.cfl_safety),
                    "timestepper/max_dt": np.float64(params.max_dt), "timestepper/min_dt": np.float64(params.min_dt)})
    with open(path, "wb") as fh:
        np.savez(fh, **out)
    return path


def load_checkpoint(path: str, domain: Domain, state: "TimeState") -> dict:
    """Restores the exact-restart state saved by :func:`save_checkpoint` into
    ``domain`` and ``state``; returns the archive's other entries."""
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    load_state(domain, data.pop("restart/state").tobytes(), state)
    return data


def step_until(fields: Fields, domain: Domain, params: "TimeParams", state: "TimeState", stop_time: float) -> int:
    """``step_until!`` (src/timestep.jl:408-449): steps of min(dt, time left)
    until ``stop_time``; the clock ends exactly at ``stop_time``.  Returns
    the number of steps taken."""
    if not stop_time > state.t:
        raise ValueError(f"stop_time ({stop_time}) must be greater than current time ({state.t})")
    n = 0
    while state.t < stop_time:
        dt_use = min(params.dt, stop_time - state.t)
        p = TimeParams(dt_use, scheme=params.scheme, filter_freq=params.filter_freq,
                       filter_strength=params.filter_strength, filter_order=params.filter_order,
                       filter_cutoff=params.filter_cutoff, ab2_reference_aliasing=params.ab2_reference_aliasing,
                       check_finite=params.check_finite)
        timestep(fields, domain, p, state, 1)
        n += 1
    state.t = stop_time
    return n


class DiagnosticTimeSeries:
    """``DiagnosticTimeSeries`` (src/timestep.jl:289-298): times, kinetic
    energy, enstrophy, total buoyancy, max divergence and max CFL, one entry
    per :func:`update_diagnostics` call.  As in the reference, enstrophy and
    max divergence are recorded as 0 placeholders."""

    def __init__(self):
        self.times, self.kinetic_energy, self.enstrophy = [], [], []
        self.total_buoyancy, self.max_divergence, self.max_cfl = [], [], []


def update_diagnostics(diag: DiagnosticTimeSeries, fields: Fields, domain: Domain, state: "TimeState"):
    """``update_diagnostics!`` (src/timestep.jl:303-330) from one
    ``ssg_diagnostics`` reduction on the device: the surface kinetic energy
    (the ``compute_surface_kinetic_energy`` formula), Σb and ``state.cfl_max``."""
    d = diagnostics(domain)
    diag.times.append(state.t)
    diag.kinetic_energy.append(d["ke_surface"])
    diag.enstrophy.append(0.0)
    diag.total_buoyancy.append(d["sum_b"])
    diag.max_divergence.append(0.0)
    diag.max_cfl.append(getattr(state, "cfl_max", 0.0))
    return diag


def ma_adjoint(domain: Domain, phibar, *, eps: float = 0.1, tol: float = 1e-10, maxiter: int = 50,
               variant: int = MA_BODY) -> tuple:
    """Vector-Jacobian product of the last Monge-Ampère solve:
    (∂φₛ/∂bₛ)ᵀ phibar, and the transposed iteration's diagnostics."""
    phibar = _f64(phibar, (domain.Ny, domain.Nx))
    out = np.empty_like(phibar)
    p = _ma_params(eps, tol, maxiter, variant, False)
    d = _MaDiag()
    domain.check(domain.lib.ssg_ma_adjoint(domain.handle, ctypes.byref(p), phibar.ctypes.data_as(_dp),
                                           out.ctypes.data_as(_dp), ctypes.byref(d)))
    return out, d.as_dict()


def timestep_adjoint_coupled(domain: Domain, params: TimeParams, step_start: int, nsteps: int, lam, *,
                             eps: float = 0.1, tol: float = 1e-10, maxiter: int = 50, variant: int = MA_BODY,
                             warm: int = 2, checkpoint_every: int = 10) -> np.ndarray:
    """Vector-Jacobian product of ``nsteps`` coupled RK3 steps (Monge-Ampère
    solve per stage) from the domain's current ``bs`` at step ``step_start``:
    returns (∂b_end/∂b_start)ᵀ lam.  Checkpoints every ``checkpoint_every``
    steps; ``bs`` is left unchanged."""
    lam = _f64(lam, (domain.Ny, domain.Nx))
    out = np.empty_like(lam)
    p = params._c()
    mp = _ma_params(eps, tol, maxiter, variant, warm)
    domain.check(domain.lib.ssg_adjoint_coupled(domain.handle, ctypes.byref(p), ctypes.byref(mp), int(step_start),
                                                int(nsteps), int(checkpoint_every), lam.ctypes.data_as(_dp),
                                                out.ctypes.data_as(_dp)))
    return out


def timestep_adjoint(domain: Domain, params: TimeParams, step_end: int, nsteps: int, lam) -> np.ndarray:
    """Vector-Jacobian product of ``nsteps`` RK3 ``timestep!`` calls (frozen φₛ)
    that ended at step ``step_end``: returns (∂b_end/∂b_start)ᵀ lam."""
    lam = _f64(lam, (domain.Ny, domain.Nx))
    out = np.empty_like(lam)
    p = params._c()
    domain.check(domain.lib.ssg_adjoint(domain.handle, ctypes.byref(p), int(step_end), int(nsteps),
                                        lam.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    return out


class Ensemble:
    """Members on their own grids stepped together with one launch per kernel
    class per RK stage (``ssg_ensemble_*``).  ``members`` holds
    ``(nx, ny[, nz])`` tuples; ``member(i)`` is a :class:`Domain` view for
    setting fields and running per-member solves."""

    def __init__(self, members, precision: int = SSG_FP64, device: int = 0, Lx=2 * np.pi, Ly=2 * np.pi, Lz=1.0):
        self.lib = load_library()
        descs = (_DomainDesc * len(members))()
        for i, m in enumerate(members):
            nx, ny = m[0], m[1]
            nz = m[2] if len(m) > 2 else 1
            descs[i] = _DomainDesc(nx, ny, nz, Lx, Ly, Lz, int(precision), int(device))
        h = ctypes.c_void_p()
        rc = self.lib.ssg_ensemble_create(ctypes.byref(h), len(members), descs)
        if rc != SSG_OK:
            raise SSGError(rc, "ssg_ensemble_create failed (see stderr)")
        self._h = h
        self.n = len(members)
        self.precision = int(precision)
        self.members = [Domain(d.nx, d.ny, d.nz, Lx, Ly, Lz, precision, device,
                               _borrowed=self.lib.ssg_ensemble_member(h, i)) for i, d in enumerate(descs)]
        self.states = [TimeState(0.0) for _ in range(self.n)]

    def member(self, i: int) -> Domain:
        return self.members[i]

    def timestep(self, params: TimeParams, dts, nsteps: int = 1):
        p = params._c()
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        st = (_TimeState * self.n)(*[s._c for s in self.states])
        rc = self.lib.ssg_ensemble_timestep(self._h, ctypes.byref(p), dt.ctypes.data_as(_dp), st, int(nsteps))
        for i in range(self.n):
            self.states[i]._c = st[i]
        if rc != SSG_OK:
            msg = self.lib.ssg_ensemble_last_error(self._h)
            raise SSGError(rc, msg.decode() if msg else "")

    def tuning(self):
        ept = ctypes.c_int32()
        us = (ctypes.c_double * 3)()
        self.lib.ssg_ensemble_tuning(self._h, ctypes.byref(ept), us)
        return {"ept": ept.value, "cand_us": {4: us[0], 8: us[1], 16: us[2]}}

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            self.lib.ssg_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
