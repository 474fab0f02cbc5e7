"""ORACLE — test infrastructure only.

Independent numpy (pocketfft) restatement of the same reference routines as
``ssg_oracle.c``.  It shares no code with the C oracle, so agreement between
the two pins the C oracle's own FFT, layout and scratch handling.

Arrays: real fields ``(ny, nx)`` C order (x fastest); spectra ``(ny//2+1, nx)``.
"""
from __future__ import annotations

import numpy as np


class Witness:
    def __init__(self, nx, ny, lx=2 * np.pi, ly=2 * np.pi):
        self.nx, self.ny, self.nyc = nx, ny, ny // 2 + 1
        self.lx, self.ly = lx, ly
        # src/domain.jl:155-162
        self.kx = 2 * np.pi * np.fft.fftfreq(nx, lx / nx)
        self.ky = 2 * np.pi * np.fft.rfftfreq(ny, ly / ny)
        # src/domain.jl:498-514
        m = np.ones((self.nyc, nx), dtype=bool)
        kxm = nx // 3
        m[:, kxm:nx - kxm] = False
        kym = 2 * self.nyc // 3
        m[kym:, :] = False
        self.mask = m

    # src/transforms.jl:42-56: r2c along y (axis 0), c2c along x (axis 1)
    def rfft2(self, a):
        return np.fft.fft(np.fft.rfft(a, axis=0), axis=1)

    def irfft2(self, s):
        return np.fft.irfft(np.fft.ifft(s, axis=1), n=self.ny, axis=0)

    def ddx(self, s):
        return (1j * self.kx)[None, :] * s

    def ddy(self, s):
        return (1j * self.ky)[:, None] * s

    def dealias(self, s):
        return np.where(self.mask, s, 0)

    def jacobian_terms(self, a, b):
        """The two flux-divergence terms of jacobian_2d! (src/transforms.jl:430-465)."""
        u = -self.irfft2(self.ddy(self.rfft2(a)))
        t1 = self.irfft2(self.ddx(self.rfft2(u * b)))
        v = self.irfft2(self.ddx(self.rfft2(a)))
        t2 = self.irfft2(self.ddy(self.rfft2(v * b)))
        return t1, t2

    def tendency(self, phi, b, quirk=False):
        """compute_tendency! (src/nonlinear.jl:22-36); quirk = AB2 aliasing."""
        t1, t2 = self.jacobian_terms(phi, b)
        j = (t2 + t2) if quirk else (t1 + t2)
        return self.irfft2(self.dealias(self.rfft2(-j)))

    def filter(self, b, strength, order=4, cutoff=0.65):
        s = self.rfft2(b)
        kmax = np.hypot(np.pi / (self.lx / self.nx), np.pi / (self.ly / self.ny))
        kn = np.hypot(self.kx[None, :], self.ky[:, None]) / kmax
        ratio = (kn - cutoff) / (1 - cutoff)
        sig = np.where(kn <= cutoff, 1.0, np.exp(-strength * np.maximum(ratio, 0) ** order))
        return self.irfft2(s * sig)

    def rk3(self, b, phi, dt):
        k1 = self.tendency(phi, b)
        k2 = self.tendency(phi, b + dt / 3 * k1)
        k3 = self.tendency(phi, b + 2 * dt / 3 * k2)
        return b + dt / 4 * (k1 + 3 * k3)
