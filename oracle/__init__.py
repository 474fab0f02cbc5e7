"""ORACLE — test infrastructure only.

ctypes bindings to the FP64 CPU restatement of the SSG.jl time-step path
(``oracle/ssg_oracle.c``) plus an independent numpy witness
(``oracle/witness.py``).  Only tests, the smoke check and the benchmark's CPU
baseline use this package; the HIP library never links or calls it.

Array conventions (the reference's, ``src/fields.jl:30-60``): real 2D fields are
numpy arrays of shape ``(ny, nx)`` in C order, i.e. column-major ``(nx, ny)``
with x fastest; spectral 2D fields are complex ``(ny//2+1, nx)`` arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SSG_ORACLE_LIB selects another build of the same source (the sanitizer
# build of `make -C oracle sanitize`, used by oracle/run_sanitized_tests.sh)
_LIB_PATH = os.environ.get("SSG_ORACLE_LIB") or os.path.join(_HERE, "libssg_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the C restatement (gcc + OpenMP) in place."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    lib.ora_create.restype = vp
    lib.ora_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3
    lib.ora_destroy.argtypes = [vp]
    lib.ora_set_vertical_grid.argtypes = [vp, _dp, _dp]
    lib.ora_get_wavenumbers.argtypes = [vp, _dp, _dp, ctypes.POINTER(ctypes.c_ubyte)]
    for name in ("ora_rfft2", "ora_irfft2", "ora_ddx2", "ora_ddy2"):
        getattr(lib, name).argtypes = [vp, _dp, _dp]
    lib.ora_dealias2.argtypes = [vp, _dp]
    lib.ora_tendency.argtypes = [vp, _dp, _dp, _dp, ctypes.c_int]
    lib.ora_jacobian.argtypes = [vp, _dp, _dp, _dp]
    lib.ora_set_linear_hook.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.ora_exponential_transfer.restype = ctypes.c_double
    lib.ora_exponential_transfer.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.ora_filter_spectral.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.ora_filter.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.ora_timestep.restype = ctypes.c_int
    lib.ora_timestep.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                 _dp, ctypes.POINTER(ctypes.c_int64)]
    lib.ora_parsevalsum2.restype = ctypes.c_double
    lib.ora_parsevalsum2.argtypes = [vp, _dp, ctypes.c_int]
    lib.ora_ma_apply_bc.argtypes = [vp, _dp, _dp, _dp]
    lib.ora_ma_residual.argtypes = [vp, _dp, _dp, _dp, ctypes.c_double, _dp]
    lib.ora_ma_d_operator.argtypes = [vp, _dp, _dp]
    lib.ora_norm.restype = ctypes.c_double
    lib.ora_norm.argtypes = [_dp, ctypes.c_int64]
    lib.ora_num_threads.restype = ctypes.c_int
    lib.ora_set_num_threads.argtypes = [ctypes.c_int]
    _lib = lib
    return lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.complex128:
        return a.ctypes.data_as(_dp)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(_dp)


def set_num_threads(n: int) -> None:
    _load().ora_set_num_threads(int(n))


def num_threads() -> int:
    return int(_load().ora_num_threads())


def exponential_transfer(k_norm: float, strength: float, order: int, cutoff: float) -> float:
    """``exponential_transfer`` (src/filter.jl:119-126)."""
    return float(_load().ora_exponential_transfer(k_norm, strength, order, cutoff))


class Oracle:
    """One reference ``Domain`` + ``Fields`` scratch + ``TimeState`` on the CPU.

    Mirrors ``make_domain`` (src/domain.jl:217-230) with a uniform,
    non-periodic vertical grid (src/domain.jl:238-249).
    """

    def __init__(self, nx: int, ny: int, nz: int = 1, lx: float = 2 * np.pi, ly: float = 2 * np.pi,
                 lz: float = 1.0):
        self.lib = _load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nyc = ny // 2 + 1
        self.lx, self.ly, self.lz = lx, ly, lz
        self.h = self.lib.ora_create(nx, ny, nz, lx, ly, lz)
        if not self.h:
            raise ValueError("invalid domain")
        self.t = ctypes.c_double(0.0)
        self.step = ctypes.c_int64(0)

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- grid ----
    def wavenumbers(self):
        kx = np.zeros(self.nx)
        ky = np.zeros(self.nyc)
        mask = np.zeros((self.nyc, self.nx), dtype=np.uint8)
        self.lib.ora_get_wavenumbers(self.h, _p(kx), _p(ky), mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return kx, ky, mask.astype(bool)

    def set_vertical_grid(self, z: np.ndarray, dz: np.ndarray):
        z = np.ascontiguousarray(z, dtype=np.float64)
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        self.lib.ora_set_vertical_grid(self.h, _p(z), _p(dz))

    # ---- transforms (src/transforms.jl:42-56, :135-166; src/domain.jl:522-535) ----
    def rfft2(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.ny, self.nx)
        out = np.zeros((self.nyc, self.nx), dtype=np.complex128)
        self.lib.ora_rfft2(self.h, _p(a), _p(out))
        return out

    def irfft2(self, s: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(s, dtype=np.complex128).reshape(self.nyc, self.nx)
        out = np.zeros((self.ny, self.nx))
        self.lib.ora_irfft2(self.h, _p(s), _p(out))
        return out

    def ddx2(self, s: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(s, dtype=np.complex128)
        out = np.zeros_like(s)
        self.lib.ora_ddx2(self.h, _p(s), _p(out))
        return out

    def ddy2(self, s: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(s, dtype=np.complex128)
        out = np.zeros_like(s)
        self.lib.ora_ddy2(self.h, _p(s), _p(out))
        return out

    def dealias2(self, s: np.ndarray) -> np.ndarray:
        s = np.array(s, dtype=np.complex128, order="C")
        self.lib.ora_dealias2(self.h, _p(s))
        return s

    # ---- physics (src/nonlinear.jl:22-57, src/filter.jl) ----
    def jacobian(self, phi: np.ndarray, b: np.ndarray) -> np.ndarray:
        """``compute_jacobian!``: the negated flux-form Jacobian."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros((self.ny, self.nx))
        self.lib.ora_jacobian(self.h, _p(out), _p(phi), _p(b))
        return out

    def tendency(self, phi: np.ndarray, b: np.ndarray, quirk: bool = False) -> np.ndarray:
        """``compute_tendency!``; ``quirk`` routes through the AB2 call site's buffer."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros((self.ny, self.nx))
        self.lib.ora_tendency(self.h, _p(out), _p(phi), _p(b), int(quirk))
        return out

    def set_linear_hook(self, on: bool, lam: float = 0.0):
        self.lib.ora_set_linear_hook(self.h, int(on), float(lam))

    def filter_spectral(self, s: np.ndarray, strength: float, order: int = 4, cutoff: float = 0.65):
        s = np.array(s, dtype=np.complex128, order="C")
        self.lib.ora_filter_spectral(self.h, _p(s), strength, order, cutoff)
        return s

    def filter(self, b: np.ndarray, strength: float, order: int = 4, cutoff: float = 0.65) -> np.ndarray:
        b = np.array(b, dtype=np.float64, order="C")
        self.lib.ora_filter(self.h, _p(b), strength, order, cutoff)
        return b

    def timestep(self, b: np.ndarray, phi: np.ndarray, scheme: int, dt: float, filter_freq: int = 10,
                 filter_strength: float = 0.1, filter_order: int = 4, filter_cutoff: float = 0.65,
                 ab2_mode: int = 0) -> int:
        """``timestep!`` (src/timestep.jl:267-285) on ``b`` in place."""
        assert b.dtype == np.float64 and b.flags["C_CONTIGUOUS"]
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        return int(self.lib.ora_timestep(self.h, scheme, dt, filter_freq, filter_strength, filter_order,
                                         filter_cutoff, ab2_mode, _p(b), _p(phi), ctypes.byref(self.t),
                                         ctypes.byref(self.step)))

    def parsevalsum2(self, s: np.ndarray, nlev: int = 1) -> float:
        s = np.ascontiguousarray(s, dtype=np.complex128)
        return float(self.lib.ora_parsevalsum2(self.h, _p(s), nlev))

    # ---- Monge-Ampere residual operator (src/poisson.jl:143-364) ----
    def ma_residual(self, phi3d: np.ndarray, b3d: np.ndarray, bs: np.ndarray | None, eps: float):
        """Returns (r, phi_after_bc); phi is copied, not modified."""
        phi = np.array(phi3d, dtype=np.float64, order="C")
        b3d = np.ascontiguousarray(b3d, dtype=np.float64)
        r = np.zeros_like(phi)
        bsp = _p(np.ascontiguousarray(bs, dtype=np.float64)) if bs is not None else None
        self.lib.ora_ma_residual(self.h, _p(phi), _p(b3d), bsp, float(eps), _p(r))
        return r, phi

    def ma_apply_bc(self, phi3d: np.ndarray, bs: np.ndarray | None):
        phi = np.array(phi3d, dtype=np.float64, order="C")
        bsp = _p(np.ascontiguousarray(bs, dtype=np.float64)) if bs is not None else None
        self.lib.ora_ma_apply_bc(self.h, _p(phi), None, bsp)
        return phi

    def ma_d_operator(self, phi3d: np.ndarray) -> np.ndarray:
        phi = np.ascontiguousarray(phi3d, dtype=np.float64)
        out = np.zeros_like(phi)
        self.lib.ora_ma_d_operator(self.h, _p(phi), _p(out))
        return out
