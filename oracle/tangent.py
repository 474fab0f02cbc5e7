"""ORACLE — test infrastructure only.

Forward-mode (tangent-linear) derivatives of the CPU restatement's coupled
step: the reference gradient that ForwardDiff would give from the Julia CPU
code, obtained here by propagating tangents through the oracle's own
operators instead of by dual numbers.

Every derivative is exact for the oracle's discrete model up to the solve
tolerance:

* The Monge-Ampère equations F(Φ; b) = 0 are the oracle's literal residual
  rows (``ma_equations``: ``compute_ssg_residual!`` on interior levels plus
  the one-sided Neumann rows, src/poisson.jl:143-364), body-forced variant
  (b on every level, src/poisson.jl:1096-1113).  F is quadratic in Φ, so the
  Jacobian-vector product (F(Φ + v) - F(Φ - v)) / 2 is exact, and Newton-Krylov
  (GMRES) solves F = 0 and the tangent system J(Φ) δΦ = δb to roundoff-level
  tolerance.  The preconditioner is the inverse of the ε = 0 operator per
  horizontal wavenumber; it only speeds GMRES up and does not enter the answer.
* The k = 0 (horizontal-mean) profile of Φ is seen by no horizontal
  derivative, so neither D(Φ) nor the surface velocities depend on it, and the
  equations of the other modes do not involve it.  The solves therefore work on
  the modes k ≠ 0 only (each level's horizontal mean projected out of Φ and of
  the equations), which also removes the k = 0 compatibility condition of the
  Neumann problem (its D̄ term is not exactly zero where the Nyquist modes
  carry content).
* The tendency T(φₛ, b) (src/nonlinear.jl:22-36) is bilinear, so its tangent is
  T(δφₛ, b) + T(φₛ, δb); the filter (src/filter.jl:283-329) is linear.
* Heun RK3 (src/timestep.jl:195-242) with the filter cadence of
  src/timestep.jl:267-285 (after the step counter is incremented).
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres

from . import Oracle
from .ma_newton import ma_equations


class _LinearInverse:
    """Inverse of the ε = 0 Monge-Ampère operator per horizontal wavenumber
    (interior rows Φzz + (kx² + ky²) symbol, one-sided Neumann rows on a
    uniform grid); the (0, 0) mode maps to zero."""

    def __init__(self, o: Oracle):
        nx, ny, nz = o.nx, o.ny, o.nz
        dz = o.lz / (nz - 1)
        kx = 2 * np.pi / o.lx * np.fft.rfftfreq(nx, 1.0 / nx)
        ky = 2 * np.pi / o.ly * np.fft.fftfreq(ny, 1.0 / ny)
        hs = -(ky[:, None] ** 2 + kx[None, :] ** 2)           # (ny, nx//2+1)
        nw = hs.size
        A = np.zeros((nw, nz, nz))
        for k in range(1, nz - 1):
            A[:, k, k - 1] = 1.0 / dz ** 2
            A[:, k, k] = -2.0 / dz ** 2 + hs.ravel()
            A[:, k, k + 1] = 1.0 / dz ** 2
        # Φ_edge + (b Φ_next + c Φ_next2) / a with a = 3/(2dz), b = -2/dz, c = 1/(2dz)
        A[:, 0, 0], A[:, 0, 1], A[:, 0, 2] = 1.0, -4.0 / 3.0, 1.0 / 3.0
        A[:, -1, -1], A[:, -1, -2], A[:, -1, -3] = 1.0, -4.0 / 3.0, 1.0 / 3.0
        A[0] = np.eye(nz)                                    # (0, 0) mode: zeroed below
        self.inv = np.linalg.inv(A)
        self.inv[0] = 0.0
        self.shape = (nz, ny, nx)
        self.nx = nx

    def __call__(self, f: np.ndarray) -> np.ndarray:
        f = f.reshape(self.shape)
        fh = np.fft.rfft2(f).reshape(self.shape[0], -1)      # (nz, nw)
        xh = np.einsum("wij,jw->iw", self.inv, fh)
        return np.fft.irfft2(xh.reshape(self.shape[0], self.shape[1], -1), s=self.shape[1:]).ravel()


class CoupledTangent:
    """Coupled (Monge-Ampère solve per RK stage) RK3 steps of the oracle and
    their tangent-linear model."""

    def __init__(self, o: Oracle, eps: float, tol: float = 1e-13):
        self.o, self.eps, self.tol = o, eps, tol
        self.M = _LinearInverse(o)
        self.n = o.nz * o.ny * o.nx
        self.iterations = []

    # -- Monge-Ampère ------------------------------------------------------
    def _b3d(self, b: np.ndarray) -> np.ndarray:
        return np.stack([b - b.mean()] * self.o.nz)

    def _proj(self, f: np.ndarray) -> np.ndarray:
        f = f.reshape(self.o.nz, self.o.ny * self.o.nx)
        return (f - f.mean(axis=1, keepdims=True)).ravel()

    def _F(self, phi: np.ndarray, b3d: np.ndarray) -> np.ndarray:
        return self._proj(ma_equations(self.o, phi.reshape(self.o.nz, self.o.ny, self.o.nx), b3d, None, self.eps))

    def _jvp_op(self, phi: np.ndarray, b3d: np.ndarray) -> LinearOperator:
        def mv(v):
            v = self._proj(np.asarray(v, dtype=float))
            s = np.linalg.norm(phi) / max(np.linalg.norm(v), 1e-300)
            return (self._F(phi + s * v, b3d) - self._F(phi - s * v, b3d)) / (2 * s)
        return LinearOperator((self.n, self.n), matvec=mv, dtype=float)

    def _solve_linear(self, phi, b3d, rhs, x0=None):
        J = self._jvp_op(phi, b3d)
        M = LinearOperator((self.n, self.n), matvec=self.M, dtype=float)
        scale = max(np.linalg.norm(rhs), 1e-300)
        x, info = gmres(J, rhs, x0=x0, rtol=0.0, atol=self.tol * scale, restart=60, maxiter=20, M=M)
        x = self._proj(x)
        res = np.linalg.norm(J.matvec(x) - rhs) / scale
        if info != 0 or res > 100 * self.tol:
            raise RuntimeError(f"GMRES did not converge (info {info}, relative residual {res:.1e})")
        return x

    def solve(self, b: np.ndarray, phi0: np.ndarray | None = None) -> np.ndarray:
        """Newton-Krylov solve of F(Φ; b) = 0; returns Φ (nz, ny, nx)."""
        b3d = self._b3d(b)
        phi = self.M(-self._F(np.zeros(self.n), b3d)) if phi0 is None else self._proj(np.array(phi0, dtype=float))
        ref = np.linalg.norm(b3d)
        for it in range(20):
            f = self._F(phi, b3d)
            if np.linalg.norm(f) <= self.tol * ref:
                self.iterations.append(it)
                return phi.reshape(self.o.nz, self.o.ny, self.o.nx)
            phi = phi + self._solve_linear(phi, b3d, -f)
        raise RuntimeError("Newton-Krylov Monge-Ampère solve did not converge")

    def tangent(self, phi: np.ndarray, b: np.ndarray, db: np.ndarray) -> np.ndarray:
        """δΦ solving J(Φ) δΦ = δb on the interior rows (δF/δb = -δb there)."""
        rhs = self._b3d(db)
        rhs[0] = 0.0
        rhs[-1] = 0.0
        return self._solve_linear(phi.ravel(), self._b3d(b), self._proj(rhs)).reshape(phi.shape)

    # -- the step ---------------------------------------------------------
    @staticmethod
    def _psi(phi: np.ndarray) -> np.ndarray:
        top = phi[-1]
        return top - top.mean()

    def run(self, b0: np.ndarray, directions, dt: float, nsteps: int, filter_freq: int = 0,
            filter_strength: float = 36.0, filter_order: int = 8, filter_cutoff: float = 0.65):
        """RK3 from step 0: returns b(T) and the tangents δb(T) of each
        direction in ``directions``."""
        o = self.o
        b = np.array(b0, dtype=float)
        db = [np.array(v, dtype=float) for v in directions]
        phi = None
        for step in range(1, nsteps + 1):
            ks, dks = [], []
            stage_b, stage_db = b, db
            for s in range(3):
                phi = self.solve(stage_b, phi)
                psi = self._psi(phi)
                ks.append(o.tendency(psi, stage_b))
                dk = []
                for d in stage_db:
                    dpsi = self._psi(self.tangent(phi, stage_b, d))
                    dk.append(o.tendency(dpsi, stage_b) + o.tendency(psi, d))
                dks.append(dk)
                if s == 0:
                    stage_b = b + dt / 3 * ks[0]
                    stage_db = [x + dt / 3 * y for x, y in zip(db, dks[0])]
                elif s == 1:
                    stage_b = b + 2 * dt / 3 * ks[1]
                    stage_db = [x + 2 * dt / 3 * y for x, y in zip(db, dks[1])]
            b = b + dt / 4 * (ks[0] + 3 * ks[2])
            db = [x + dt / 4 * (y + 3 * z) for x, y, z in zip(db, dks[0], dks[2])]
            if filter_freq > 0 and step % filter_freq == 0:
                b = o.filter(b, filter_strength, filter_order, filter_cutoff)
                db = [o.filter(x, filter_strength, filter_order, filter_cutoff) for x in db]
        return b, db
