"""ORACLE — test infrastructure only.

Restatement of the vertical-velocity (ω equation) path of
``compute_vertical_velocity`` (src/extras/omega_eq.jl:574-652) on the
discrete vertical grid of the Monge-Ampère problem, in numpy:

* ``divq_h`` follows ``compute_divergence_Q!`` (:326-405): per level
  Q₁ = ∂x u_g ∂x b + ∂x v_g ∂y b, Q₂ = ∂y u_g ∂x b + ∂y v_g ∂y b with
  u_g = -∂yΦ, v_g = ∂xΦ and the surface buoyancy b, then
  -2ε (i kx Q̂₁ + i ky Q̂₂) without its x- and y-Nyquist bins.  Spectral
  derivatives use the r2c-along-x transform of that file (``rfftplan`` over
  (nkr, nl)); it and the step path's r2c-along-y transform differ only in
  the Nyquist bins, which the forcing leaves out.
* ``vertical_solve`` solves (∂zz - |k|²) w* = f per horizontal wavenumber
  with w* = 0 on the bottom and top levels, the equation the file documents
  (:524-537), discretised with the three-point d2dz2! stencil
  (src/transforms.jl:798-850) on the same dz as the Monge-Ampère operator.
  The file's own integrals (:408-521) are not used: their kernel
  (e^{k(z-z')} + e^{-k(z-z')})/2k is a cosh, not the sinh the equation needs,
  so they neither satisfy it nor vanish at the bottom.
* ``vertical_velocity`` divides by the inverse Jacobian 1 - ε(Φxx + Φyy)
  (:612-646).

Arrays: fields are (ny, nx) and (nz, ny, nx) in C order (x fastest), the
reference's column-major (nx, ny[, nz]) memory.
"""
from __future__ import annotations

import numpy as np


def _k(nx, ny, lx, ly):
    kx = 2 * np.pi / lx * np.fft.rfftfreq(nx, 1.0 / nx)[None, :]
    ky = 2 * np.pi / ly * np.fft.fftfreq(ny, 1.0 / ny)[:, None]
    return kx, ky


def hessian(phi2: np.ndarray, lx: float = 2 * np.pi, ly: float = 2 * np.pi):
    """Φxx, Φyy, Φxy of one level."""
    ny, nx = phi2.shape
    kx, ky = _k(nx, ny, lx, ly)
    p = np.fft.rfft2(phi2)
    inv = lambda s: np.fft.irfft2(s * p, s=(ny, nx))
    return inv(-kx * kx), inv(-ky * ky), inv(-kx * ky)


def divq_h(phi3d: np.ndarray, b: np.ndarray, eps: float, lx: float = 2 * np.pi, ly: float = 2 * np.pi):
    """-2ε ∇·Q per level, forward spectrum (nz, ny, nx//2 + 1)."""
    nz, ny, nx = phi3d.shape
    kx, ky = _k(nx, ny, lx, ly)
    bh = np.fft.rfft2(b)
    bx = np.fft.irfft2(1j * kx * bh, s=(ny, nx))
    by = np.fft.irfft2(1j * ky * bh, s=(ny, nx))
    out = np.empty((nz, ny, nx // 2 + 1), dtype=complex)
    for k in range(nz):
        xx, yy, xy = hessian(phi3d[k], lx, ly)
        dx_u, dy_u, dx_v, dy_v = -xy, -yy, xx, xy
        q1 = dx_u * bx + dx_v * by
        q2 = dy_u * bx + dy_v * by
        out[k] = -2 * eps * (1j * kx * np.fft.rfft2(q1) + 1j * ky * np.fft.rfft2(q2))
    # the x- and y-Nyquist bins are left out: the reference's two transform
    # conventions (r2c along x here, along y in the step path) disagree there
    out[:, :, nx // 2] = 0.0
    out[:, ny // 2, :] = 0.0
    return out


def vertical_solve(f: np.ndarray, dz: np.ndarray, ksq: np.ndarray) -> np.ndarray:
    """(∂zz - |k|²) w = f on interior levels, w = 0 on levels 0 and nz-1;
    f: (nz, ...) per wavenumber, ksq: |k|² broadcast against f[0]."""
    nz = f.shape[0]
    shp = f.shape[1:]
    ks = np.broadcast_to(ksq, shp).ravel()
    fl = f.reshape(nz, -1)
    m = nz - 2
    A = np.zeros((ks.size, m, m))
    for i in range(m):
        k = i + 1
        dzm, dzp = dz[k - 1], dz[k]
        al, be, ga = 2 / (dzm * (dzm + dzp)), -2 / (dzm * dzp), 2 / (dzp * (dzm + dzp))
        A[:, i, i] = be - ks
        if i > 0:
            A[:, i, i - 1] = al
        if i < m - 1:
            A[:, i, i + 1] = ga
    rhs = fl[1:-1].T[:, :, None]
    w = np.zeros((nz, ks.size), dtype=np.result_type(f, float))
    w[1:-1] = np.linalg.solve(A, rhs)[:, :, 0].T
    return w.reshape((nz,) + shp)


def vertical_velocity(phi3d: np.ndarray, b: np.ndarray, eps: float, dz: np.ndarray, lx: float = 2 * np.pi,
                      ly: float = 2 * np.pi) -> np.ndarray:
    """w (nz, ny, nx) from Φ (nz, ny, nx) and the surface buoyancy b (ny, nx)."""
    nz, ny, nx = phi3d.shape
    kx, ky = _k(nx, ny, lx, ly)
    ws = vertical_solve(divq_h(phi3d, b, eps, lx, ly), np.asarray(dz, dtype=float), kx * kx + ky * ky)
    w = np.empty((nz, ny, nx))
    for k in range(nz):
        xx, yy, _ = hessian(phi3d[k], lx, ly)
        w[k] = np.fft.irfft2(ws[k], s=(ny, nx)) / (1.0 - eps * (xx + yy))
    return w
