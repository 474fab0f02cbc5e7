"""ORACLE — test infrastructure only.

Brute-force solve of the reference's Monge-Ampère problem at tiny sizes:
Newton's method with a finite-difference dense Jacobian of the literal
residual operator (``Oracle.ma_residual`` = ``compute_ssg_residual!`` +
``apply_ssg_boundary_conditions!``, src/poisson.jl:143-364).  It shares no
algorithm with the GPU solver (spectral Picard with per-wavenumber
tridiagonal solves), only the residual definition.

Equations: the interior residual rows r = 0, and on the two boundary levels
Φ = BC(Φ) (the one-sided Neumann overwrite).  The constant mode is a null
direction; the least-squares Newton step leaves it at the initial guess's.
"""
from __future__ import annotations

import numpy as np

from . import Oracle


def ma_equations(o: Oracle, phi: np.ndarray, b3d: np.ndarray, bs: np.ndarray | None, eps: float) -> np.ndarray:
    r, phi_bc = o.ma_residual(phi, b3d, bs, eps)
    f = r.copy()
    f[0] = phi[0] - phi_bc[0]
    f[-1] = phi[-1] - phi_bc[-1]
    return f.ravel()


def ma_newton(o: Oracle, b3d: np.ndarray, bs: np.ndarray | None, eps: float, phi0: np.ndarray | None = None,
              tol: float = 1e-12, maxit: int = 12, h: float = 1e-3):
    """The residual is quadratic in Φ, so the central difference below is
    the exact Jacobian up to roundoff for any h."""
    shape = (o.nz, o.ny, o.nx)
    phi = np.zeros(shape) if phi0 is None else np.array(phi0, dtype=float)
    hist = []
    for _ in range(maxit):
        f0 = ma_equations(o, phi, b3d, bs, eps)
        nrm = np.linalg.norm(f0)
        hist.append(nrm)
        if nrm < tol:
            break
        n = f0.size
        J = np.empty((n, n))
        flat = phi.ravel()
        for j in range(n):
            save = flat[j]
            flat[j] = save + h
            fp = ma_equations(o, flat.reshape(shape), b3d, bs, eps)
            flat[j] = save - h
            fm = ma_equations(o, flat.reshape(shape), b3d, bs, eps)
            J[:, j] = (fp - fm) / (2 * h)
            flat[j] = save
        step = np.linalg.lstsq(J, -f0, rcond=1e-12)[0]
        phi = (flat + step).reshape(shape)
    return phi, hist
