"""Deterministic synthetic initial conditions for the benchmark and tests.

Surface buoyancy b0: random-phase Gaussian field with isotropic shell spectrum
E(k) ∝ k²·exp(-(k/k0)²).  Per-mode amplitude is sqrt(E(k)/k) (a 2D shell of
radius k holds ∝ k modes), phases are uniform, keyed by SplitMix64 of
(seed, kx, ky) so they do not depend on layout or resolution.  The k = 0 mode
is zero, the reference's 2/3 mask (src/domain.jl:498-514) is applied, the
field is transformed with the reference's inverse convention and rescaled to
rms 1.  The reference generates its own ICs with Julia's RNG
(test/test_ssg_jfm_example.jl:107-108); that stream is not reproduced.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014), vectorised on uint64."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_keyed(seed: int, kx: np.ndarray, ky: np.ndarray, stream: int = 0) -> np.ndarray:
    """Uniform [0, 1) numbers keyed by (seed, stream, kx, ky)."""
    kx = np.asarray(kx, dtype=np.int64).astype(np.uint64)
    ky = np.asarray(ky, dtype=np.int64).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        h = splitmix64(h ^ (kx * np.uint64(0xD6E8FEB86659FD93)))
        h = splitmix64(h ^ (ky * np.uint64(0xA0761D6478BD642F)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def twothirds_mask(nx: int, ny: int) -> np.ndarray:
    """The reference's mask on its (ny//2+1, nx) half plane."""
    nyc = ny // 2 + 1
    m = np.ones((nyc, nx), dtype=bool)
    kxm = nx // 3
    m[:, kxm:nx - kxm] = False
    m[2 * nyc // 3:, :] = False
    return m


def spectral_b0(nx: int, ny: int, seed: int = 0, k0: float = 6.0, lx: float = 2 * np.pi,
                ly: float = 2 * np.pi) -> np.ndarray:
    """Random-phase b0, shape (ny, nx), rms 1."""
    nyc = ny // 2 + 1
    kxi = np.round(np.fft.fftfreq(nx, 1.0 / nx)).astype(np.int64)      # signed integer kx
    kyi = np.arange(nyc, dtype=np.int64)
    KXI, KYI = np.meshgrid(kxi, kyi)                                      # (nyc, nx)
    kx = 2 * np.pi / lx * KXI
    ky = 2 * np.pi / ly * KYI
    k = np.hypot(kx, ky)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(k ** 2 * np.exp(-(k / k0) ** 2) / k), 0.0)
    # phases keyed on the canonical member of each conjugate pair: on the
    # ky = 0 line the pair (kx, 0) ~ (-kx, 0) shares |kx|
    canon_kx = np.where(KYI == 0, np.abs(KXI), KXI)
    ph = 2 * np.pi * uniform_keyed(seed, canon_kx, KYI)
    spec = amp * np.exp(1j * ph)
    # Hermitian ky = 0 line: S(-kx, 0) = conj(S(kx, 0))
    neg = (KYI == 0) & (KXI < 0)
    spec[neg] = np.conj(spec[neg])
    spec[(KYI == 0) & (KXI == 0)] = 0.0
    spec[(KYI == 0) & (np.abs(KXI) == nx // 2)] = 0.0
    spec = np.where(twothirds_mask(nx, ny), spec, 0.0)
    # the mask keeps kx = -(nx//3) but not +(nx//3); on the ky = 0 line a real
    # field holds both, so drop that pair to stay inside the mask
    spec[(KYI == 0) & (np.abs(KXI) >= nx // 3)] = 0.0
    b = np.fft.irfft(np.fft.ifft(spec, axis=1), n=ny, axis=0)
    b /= np.sqrt(np.mean(b * b))
    return np.ascontiguousarray(b)


def streamfunction_from_b(b: np.ndarray, lx: float = 2 * np.pi, ly: float = 2 * np.pi) -> np.ndarray:
    """Frozen surface streamfunction for the Mode R step: the ε = 0 solution
    of the as-wired Monge-Ampère problem, φ̂ = -b̂/k² (k = 0 mode zero), i.e.
    ``solve_poisson_simple`` (src/poisson.jl:1002-1059) on the surface level."""
    ny, nx = b.shape
    s = np.fft.fft(np.fft.rfft(b, axis=0), axis=1)
    kx = 2 * np.pi * np.fft.fftfreq(nx, lx / nx)
    ky = 2 * np.pi * np.fft.rfftfreq(ny, ly / ny)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(k2 > 1e-14, -s / k2, 0.0)
    return np.ascontiguousarray(np.fft.irfft(np.fft.ifft(s, axis=1), n=ny, axis=0))
