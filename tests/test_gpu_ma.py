"""GPU Monge-Ampère solve against the oracle.

The anchor is the reference residual operator restated in the oracle
(compute_ssg_residual! + apply_ssg_boundary_conditions!,
src/poisson.jl:143-364): the GPU solution must zero it on interior levels
and satisfy the one-sided Neumann rows; it is also compared with an
independent dense Newton solve (oracle/ma_newton.py) and, at ε = 0, with
closed forms.  The horizontal-mean (k = 0) profile is compared separately:
its Neumann-Neumann problem is singular and is pinned by a gauge.
"""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from oracle.ma_newton import ma_newton

pytestmark = pytest.mark.gpu


def grid(nx, ny, lx=2 * np.pi, ly=2 * np.pi):
    return np.meshgrid(lx / nx * np.arange(nx), ly / ny * np.arange(ny))


def demean(phi):
    return phi - phi.mean(axis=(1, 2), keepdims=True)


def smooth_b(n, seed, k0):
    from ssgamd.ics import spectral_b0
    return spectral_b0(n, n, seed=seed, k0=k0)


def gpu_solve(gpu, b, nz, eps, variant, tol=1e-12, maxiter=60, warm=False):
    n = b.shape[0]
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    f.bs = b
    diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=tol, maxiter=maxiter, variant=variant, warm=warm)
    return d, f, diag


def test_k12_body_eps0_closed_form(gpu):
    """ε = 0, b on every level: Φ̂ = -b̂/k² on every level (src/poisson.jl:1002-1059)."""
    X, Y = grid(32, 32)
    b = np.cos(2 * X) * np.sin(Y) + 0.3 * np.sin(3 * X + Y)
    d, f, diag = gpu_solve(gpu, b, 8, 0.0, gpu.MA_BODY)
    assert diag["converged"] and diag["iterations"] == 0
    phi = gpu.get_phi3d(d)
    expect = -(np.cos(2 * X) * np.sin(Y) / 5 + 0.3 * np.sin(3 * X + Y) / 10)
    assert np.abs(phi - expect[None]).max() < 1e-13
    assert np.abs(f.phis - expect).max() < 1e-13


def dense_surface_solution(b, nz, lz=1.0):
    """ε = 0, surface slope b: per-wavenumber dense solve of the interior
    d2dz2 rows + both one-sided Neumann rows (numpy.linalg.solve)."""
    n = b.shape[0]
    dz = lz / (nz - 1)
    bh = np.fft.fft2(b)
    k = np.fft.fftfreq(n, 1.0 / n)
    K2 = k[None, :] ** 2 + k[:, None] ** 2
    out = np.zeros((nz, n, n), dtype=complex)
    for iy in range(n):
        for ix in range(n):
            A = np.zeros((nz, nz))
            rhs = np.zeros(nz, dtype=complex)
            for z in range(1, nz - 1):
                A[z, z - 1] = A[z, z + 1] = 1 / dz ** 2
                A[z, z] = -2 / dz ** 2 - K2[iy, ix]
            A[0, :3] = [-3 / (2 * dz), 2 / dz, -1 / (2 * dz)]
            A[-1, -3:] = [1 / (2 * dz), -2 / dz, 3 / (2 * dz)]
            rhs[-1] = bh[iy, ix]
            if ix == 0 and iy == 0:
                A[0, :] = 0
                A[0, 0] = 1
                rhs[:] = 0
            out[:, iy, ix] = np.linalg.solve(A, rhs)
    return np.real(np.fft.ifft2(out, axes=(1, 2)))


def test_k13_surface_eps0_dense(gpu):
    b = smooth_b(32, 2, 3.0)
    d, f, diag = gpu_solve(gpu, b, 7, 0.0, gpu.MA_SURFACE)
    assert diag["converged"]
    phi = gpu.get_phi3d(d)
    ref = dense_surface_solution(b, 7)
    assert np.abs(phi - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nz", [3, 5, 6])
@pytest.mark.parametrize("variant", [0, 1])
def test_eps_newton_parity_tiny(gpu, variant, nz):
    """Against dense Newton at 8²; nz = 3, 5, 6 cover the per-wavenumber
    solve's two-lane split at its smallest, odd and even sizes."""
    n, eps = 8, 0.1
    b = smooth_b(n, 4, 1.2)
    d, f, diag = gpu_solve(gpu, b, nz, eps, variant)
    assert diag["converged"], diag
    phi = gpu.get_phi3d(d)
    o = Oracle(n, n, nz)
    b3d = np.stack([b] * nz) if variant == 0 else np.zeros((nz, n, n))
    g = None if variant == 0 else b
    ref, hist = ma_newton(o, b3d, g, eps)
    assert np.abs(demean(phi) - demean(ref)).max() < 1e-9 * np.abs(ref).max()
    # the reference residual of the GPU solution, interior levels
    r, _ = o.ma_residual(phi, b3d, g, eps)
    assert np.abs(r).max() < 1e-9 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("variant", [0, 1])
def test_reference_residual_256(gpu, variant):
    """256²×16, ε = 0.1: converged to 1e-10 relative; the oracle's residual
    operator agrees at the returned Φ."""
    n, nz, eps = 256, 16, 0.1
    # the surface-forced problem stays elliptic (ε·max|∇∇Φ| < 1) only for
    # weaker forcing than the rms-1 body-forced case
    b = smooth_b(n, 0, 6.0) * (1.0 if variant == 0 else 0.2)
    d, f, diag = gpu_solve(gpu, b, nz, eps, variant, tol=1e-10)
    print("MA 256^2x16 variant", variant, diag["iterations"], diag["relative_residual"], flush=True)
    assert diag["converged"] and diag["relative_residual"] <= 1e-10
    h = diag["convergence_history"]
    assert all(h[i + 1] < h[i] for i in range(len(h) - 1))
    phi = gpu.get_phi3d(d)
    o = Oracle(n, n, nz)
    b3d = np.stack([b] * nz) if variant == 0 else np.zeros((nz, n, n))
    g = None if variant == 0 else b
    r, phi_bc = o.ma_residual(phi, b3d, g, eps)
    rhs = np.linalg.norm(b3d[1:-1]) if variant == 0 else np.linalg.norm(phi[1:-1]) * 0 + 1.0
    assert abs(np.linalg.norm(r) - diag["final_residual"]) < 1e-6 * max(diag["final_residual"], 1e-12) + 1e-10
    assert np.linalg.norm(r) < 1e-9 * rhs
    # boundary rows hold: the reference's overwrite leaves the top level unchanged
    assert np.abs(phi_bc[-1] - phi[-1]).max() < 1e-9 * np.abs(phi).max()
    # surface outputs
    assert np.abs(f.phis - phi[-1]).max() < 1e-12 * np.abs(phi).max()
    if variant == 0:
        assert np.abs(phi - phi[:1]).max() < 1e-9 * np.abs(phi).max()   # z-independent fixed point


def test_loss_of_ellipticity_is_reported(gpu):
    """rms-1 surface forcing at ε = 0.1 leaves the elliptic range: the solve
    must stop and report non-convergence, not return garbage as converged."""
    b = smooth_b(128, 0, 6.0)
    d, f, diag = gpu_solve(gpu, b, 8, 0.1, gpu.MA_SURFACE, tol=1e-10, maxiter=50)
    assert not diag["converged"]


def test_not_converged_returns_flag(gpu):
    b = smooth_b(64, 0, 6.0)
    d, f, diag = gpu_solve(gpu, b, 8, 0.1, gpu.MA_SURFACE, tol=1e-15, maxiter=2)
    assert not diag["converged"] and diag["iterations"] == 2


@pytest.mark.parametrize("warm", [1, 2])
def test_coupled_rk3_against_oracle(gpu, warm):
    """Coupled step: the MA problem re-solved for every RK stage (φₛ follows b),
    against the oracle tendency fed by dense-Newton solves (warm starts from
    the last D(Φ) and from the extrapolation of the last two)."""
    n, nz, eps, dt = 8, 4, 0.1, 1e-2
    b0 = smooth_b(n, 5, 1.2)
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-13, maxiter=80, variant=gpu.MA_BODY, warm=warm)
    st = gpu.TimeState(0.0, f)
    nsteps = 4   # with warm = 2 the fourth step's solves use the per-stage quadratic extrapolation
    gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=0), st, nsteps)
    g = f.bs
    o = Oracle(n, n, nz)

    def psi_of(b):
        phi, _ = ma_newton(o, np.stack([b] * nz), None, eps)
        return phi[-1] - phi[-1].mean()

    b = b0.copy()
    for _ in range(nsteps):
        k1 = o.tendency(psi_of(b), b)
        b1 = b + dt / 3 * k1
        k2 = o.tendency(psi_of(b1), b1)
        b2 = b + 2 * dt / 3 * k2
        k3 = o.tendency(psi_of(b2), b2)
        b = b + dt / 4 * (k1 + 3 * k3)
    assert np.abs(g - b).max() < 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("variant", [0, 1])
def test_mixed_precision_solve(gpu, variant):
    """FP32-mixed contexts run the leading iterations with FP32 transforms,
    products and D̂, then FP64 iterations to `tol`: the answer must satisfy
    the oracle's residual operator like the FP64 solve and equal it."""
    n, nz, eps = 256, 16, 0.1
    b = smooth_b(n, 0, 6.0) * (1.0 if variant == 0 else 0.2)
    sols = {}
    for prec in (gpu.SSG_FP64, gpu.SSG_FP32_MIXED):
        d = gpu.make_domain(n, n, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b
        diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=variant)
        print("MA precision", prec, "variant", variant, diag["iterations"], diag["iterations_fp32"],
              diag["convergence_history"], flush=True)
        assert diag["converged"] and diag["relative_residual"] <= 1e-10, diag
        sols[prec] = (gpu.get_phi3d(d), f.phis, diag)
        d.close()
    phi32, phis32, diag32 = sols[gpu.SSG_FP32_MIXED]
    phi64, phis64, diag64 = sols[gpu.SSG_FP64]
    assert diag32["iterations_fp32"] >= 2 and diag64["iterations_fp32"] == 0
    assert diag32["iterations"] > diag32["iterations_fp32"]   # the last evaluations are FP64
    assert np.abs(phi32 - phi64).max() < 1e-9 * np.abs(phi64).max()
    assert np.abs(phis32 - phi32[-1]).max() < 1e-12 * np.abs(phi32).max()
    o = Oracle(n, n, nz)
    b3d = np.stack([b] * nz) if variant == 0 else np.zeros((nz, n, n))
    g = None if variant == 0 else b
    r, _ = o.ma_residual(phi32, b3d, g, eps)
    rhs = np.linalg.norm(b3d[1:-1]) if variant == 0 else 1.0
    assert np.linalg.norm(r) < 1e-9 * rhs


def test_mixed_precision_warm_start(gpu):
    n, nz, eps = 128, 8, 0.1
    b = smooth_b(n, 3, 6.0)
    d = gpu.make_domain(n, n, nz, precision=gpu.SSG_FP32_MIXED)
    f = gpu.allocate_fields(d)
    f.bs = b
    cold = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
    f.bs = b * 1.001
    warm = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=True)
    assert cold["converged"] and warm["converged"]
    assert warm["iterations"] < cold["iterations"]
    r, _ = Oracle(n, n, nz).ma_residual(gpu.get_phi3d(d), np.stack([b * 1.001] * nz), None, eps)
    assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(np.stack([b] * (nz - 2)))
    d.close()


def test_coupled_rk3_mixed_against_fp64(gpu):
    """Coupled step in both precisions from the same state: per-output errors
    of the FP32-mixed form (b, φₛ, surface kinetic energy) stay within 1e-5."""
    n, nz, eps, dt, steps = 128, 8, 0.1, 1e-3, 10
    b0 = smooth_b(n, 6, 6.0)
    out = {}
    for prec in (gpu.SSG_FP64, gpu.SSG_FP32_MIXED):
        d = gpu.make_domain(n, n, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=5, filter_strength=36.0, filter_order=8,
                                          filter_cutoff=0.65), st, steps)
        out[prec] = (f.bs, f.phis, gpu.diagnostics(d)["ke_surface"])
        d.close()
    (b32, p32, k32), (b64, p64, k64) = out[gpu.SSG_FP32_MIXED], out[gpu.SSG_FP64]
    eb = np.linalg.norm(b32 - b64) / np.linalg.norm(b64)
    ep = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    ek = abs(k32 - k64) / abs(k64)
    print(f"coupled mixed vs fp64 after {steps} steps: b {eb:.2e} phis {ep:.2e} ke {ek:.2e}", flush=True)
    assert eb < 1e-5 and ep < 1e-5 and ek < 1e-5
    assert eb > 0   # the FP32 path really ran


@pytest.mark.parametrize("nx,ny,nz,variant,k0", [(16, 8, 4, 1, 1.0), (8, 16, 4, 0, 1.5), (8, 16, 4, 1, 1.5),
                                                  (8, 32, 3, 0, 1.5)])
def test_rectangular_grids_newton(gpu, nx, ny, nz, variant, k0):
    """Non-square horizontal grids: the row pass (along x) and the column
    pass (along y) have different lengths; nz = 3 is the smallest column
    (one interior level between the two boundary rows)."""
    eps = 0.1
    from ssgamd.ics import spectral_b0
    b = spectral_b0(nx, ny, seed=8, k0=k0) * (1.0 if variant == 0 else 0.3)
    d = gpu.make_domain(nx, ny, nz)
    f = gpu.allocate_fields(d)
    f.bs = b
    diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-12, maxiter=80, variant=variant)
    assert diag["converged"], diag
    phi = gpu.get_phi3d(d)
    o = Oracle(nx, ny, nz)
    b3d = np.stack([b] * nz) if variant == 0 else np.zeros((nz, ny, nx))
    g = None if variant == 0 else b
    ref, _ = ma_newton(o, b3d, g, eps)
    assert np.abs(demean(phi) - demean(ref)).max() < 1e-9 * np.abs(ref).max()
    r, _ = o.ma_residual(phi, b3d, g, eps)
    assert np.abs(r).max() < 1e-9 * max(1.0, np.abs(b).max())
    d.close()


@pytest.mark.parametrize("warm", [1, 2])
def test_coupled_ab2_against_oracle(gpu, warm):
    """Coupled AB2 (forward Euler first step): the Monge-Ampère problem
    re-solved before every tendency, against the oracle tendency fed by
    dense-Newton solves."""
    n, nz, eps, dt = 8, 4, 0.1, 1e-2
    b0 = smooth_b(n, 9, 1.2)
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-13, maxiter=80, variant=gpu.MA_BODY, warm=warm)
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.AB2_LowStorage, filter_freq=0), st, 5)
    g = f.bs
    o = Oracle(n, n, nz)

    def psi_of(b):
        phi, _ = ma_newton(o, np.stack([b] * nz), None, eps)
        return phi[-1] - phi[-1].mean()

    b = b0.copy()
    fold = None
    for k in range(5):
        t = o.tendency(psi_of(b), b)
        b = b + dt * t if k == 0 else b + dt * (1.5 * t - 0.5 * fold)
        fold = t
    assert np.abs(g - b).max() < 1e-9 * np.abs(b).max()
    d.close()


@pytest.mark.parametrize("prec", ["SSG_FP64", "SSG_FP32_MIXED"])
def test_extrapolated_warm_start(gpu, prec):
    """Coupled RK3 at 128²×8 with warm=2 (first iterate from 2 D̂ₖ - D̂ₖ₋₁)
    against warm=1: the same trajectory to the solve tolerance, and no more
    iterations in the last solve."""
    n, nz, eps, dt, steps = 128, 8, 0.1, 1e-3, 10
    b0 = smooth_b(n, 6, 6.0)
    out = {}
    for warm in (1, 2):
        d = gpu.make_domain(n, n, nz, precision=getattr(gpu, prec))
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=warm)
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=5, filter_strength=36.0, filter_order=8,
                                          filter_cutoff=0.65), st, steps)
        last = gpu._MaDiag()
        d.check(d.lib.ssg_last_ma_diag(d.handle, ctypes.byref(last)))
        out[warm] = (f.bs, last.iterations, last.iterations_fp32, last.relative_residual, bool(last.converged))
        d.close()
    (b1, it1, f1, r1, c1), (b2, it2, f2, r2, c2) = out[1], out[2]
    e = np.linalg.norm(b2 - b1) / np.linalg.norm(b1)
    print(f"{prec}: warm=1 last solve {it1} its ({f1} fp32) r={r1:.1e}; warm=2 {it2} its ({f2} fp32) r={r2:.1e};"
          f" b diff {e:.1e}", flush=True)
    assert c1 and c2
    assert e < 1e-8
    assert it2 <= it1


@pytest.mark.parametrize("scheme", [1, 0])
@pytest.mark.parametrize("form", ["fp64", "stage_fp32"])
def test_fused_extrapolation_is_bitwise(gpu, scheme, form, monkeypatch):
    """The quadratic warm start formed inside the first per-wavenumber solve
    (k_ma_tridiag's extrapolating form, the default) gives the same bits as
    the k_extrap pass followed by the plain solve (SSG_MA_EXTRAP_FUSED=0):
    coupled RK3 and AB2 steps at 64²×8 with warm=2, FP64 and FP32 stage
    solves, b and φₛ compared after each of two calls.  The profile counts
    show which form ran."""
    n, nz, eps, dt, calls = 64, 8, 0.1, 1e-3, (6, 6)
    b0 = smooth_b(n, 6, 6.0)
    s32 = form == "stage_fp32"
    prec = gpu.SSG_FP32_MIXED if s32 else gpu.SSG_FP64
    kind = 13 if s32 else 5  # profile class of the extrapolating solve
    res = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("SSG_MA_EXTRAP_FUSED", fused)
        d = gpu.make_domain(n, n, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2, stage_fp32=s32)
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(dt, scheme=scheme, filter_freq=5, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
        d.profile(True)
        outs = []
        for k in calls:
            gpu.timestep(f, d, p, st, k)
            outs.append((f.bs, f.phis))
        launches = d.profile_read(kind)[1]
        d.profile(False)
        d.close()
        res[fused] = (outs, launches)
    (o1, n1), (o0, n0) = res["1"], res["0"]
    print(f"{form} scheme {scheme}: extrapolating solves {n1} (fused), {n0} (separate)", flush=True)
    assert n1 > 0 and n0 == 0
    for (b1, p1), (b0_, p0) in zip(o1, o0):
        assert np.array_equal(b1, b0_)
        assert np.array_equal(p1, p0)


def test_warm_start_history_across_mode_changes(gpu):
    """The extrapolated warm start's solve history across every change of
    pattern: coupled RK3 (three solves per step), coupled AB2 (one), a
    standalone solve, a new b written mid-run, the coupling switched off and
    on.  Each pattern change must drop or reset the history; the result must
    follow the same trajectory as the plain warm start to the solve
    tolerance."""
    n, nz, eps, dt = 64, 8, 0.1, 1e-3
    b0 = smooth_b(n, 4, 6.0)
    b1 = smooth_b(n, 5, 6.0)
    out = {}
    for warm in (1, 2):
        d = gpu.make_domain(n, n, nz)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.set_coupling(d, eps=eps, tol=1e-12, maxiter=80, variant=gpu.MA_BODY, warm=warm)
        st = gpu.TimeState(0.0, f)
        rk3 = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=4, filter_strength=36.0, filter_order=8,
                             filter_cutoff=0.65)
        ab2 = gpu.TimeParams(dt, scheme=gpu.AB2_LowStorage, filter_freq=0)
        gpu.timestep(f, d, rk3, st, 6)
        gpu.timestep(f, d, ab2, gpu.TimeState(0.0, f), 5)
        diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-12, maxiter=80, variant=gpu.MA_BODY, warm=warm)
        assert diag["converged"]
        gpu.timestep(f, d, rk3, st, 4)
        f.bs = 0.5 * (f.bs + b1)            # a new b mid-run
        gpu.timestep(f, d, rk3, st, 4)
        gpu.set_coupling(d, eps=None)
        gpu.set_coupling(d, eps=eps, tol=1e-12, maxiter=80, variant=gpu.MA_BODY, warm=warm)
        gpu.timestep(f, d, rk3, st, 4)
        out[warm] = f.bs
        d.close()
    e = np.abs(out[2] - out[1]).max() / np.abs(out[1]).max()
    print(f"history across mode changes: max rel diff {e:.1e}", flush=True)
    assert e < 1e-9


@pytest.mark.parametrize("prec", ["SSG_FP64", "SSG_FP32_MIXED"])
def test_streaming_plane_access_is_bitwise_identical(gpu, prec, monkeypatch):
    """Nontemporal access to the Monge-Ampère planes (chosen by size: only
    planes above 128 MiB use it) changes cache hints, not arithmetic.  Forced
    on and off at 128²×8 through SSG_MA_STREAM, a cold solve and coupled RK3
    steps with the extrapolated warm start give bitwise the same b and φₛ."""
    n, nz, eps, dt, steps = 128, 8, 0.1, 1e-3, 8
    b0 = smooth_b(n, 7, 6.0)
    out = {}
    for stream in ("0", "1"):
        monkeypatch.setenv("SSG_MA_STREAM", stream)
        d = gpu.make_domain(n, n, nz, precision=getattr(gpu, prec))
        f = gpu.allocate_fields(d)
        f.bs = b0
        diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        assert diag["converged"]
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2)
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=4, filter_strength=36.0, filter_order=8,
                                          filter_cutoff=0.65), st, steps)
        out[stream] = (f.bs, f.phis)
        d.close()
    assert np.array_equal(out["0"][0], out["1"][0])
    assert np.array_equal(out["0"][1], out["1"][1])


def test_eps_closed_form_known_answer(gpu):
    """ε = 0.1, body variant: Φ = A cos x + B cos y on every level solves the
    residual operator for b = -A cos x - B cos y - εAB cos x cos y (its
    Φxx Φyy - Φxy² is AB cos x cos y and Φzz = 0).  The GPU solve recovers Φ
    on every level (horizontal mean 0, the gauge's value for this b)."""
    n, nz, eps, A, B = 64, 8, 0.1, 0.8, -0.6
    X, Y = grid(n, n)
    b = -A * np.cos(X) - B * np.cos(Y) - eps * A * B * np.cos(X) * np.cos(Y)
    d, f, diag = gpu_solve(gpu, b, nz, eps, gpu.MA_BODY, tol=1e-13, maxiter=80)
    assert diag["converged"] and not diag["not_elliptic"]
    expect = A * np.cos(X) + B * np.cos(Y)
    phi = gpu.get_phi3d(d)
    assert np.abs(phi - expect[None]).max() < 1e-11
    assert np.abs(f.phis - expect).max() < 1e-11
    d.close()


def test_not_elliptic_is_rejected_up_front(gpu):
    """Body variant with ε·max(b) >= 1: no elliptic solution exists
    (det(I - εH) = 1 - εb at the z-independent fixed point), so the solve is
    rejected before iterating, through ssg_ma_solve and through set_b!.  Just
    inside the bound the same field solves normally."""
    n, nz, eps = 64, 8, 0.3
    b0 = smooth_b(n, 7, 6.0)
    for scale, elliptic in ((1.05 / (eps * b0.max()), False), (0.5 / (eps * b0.max()), True)):
        b = scale * b0
        d = gpu.make_domain(n, n, nz)
        f = gpu.allocate_fields(d)
        f.bs = b
        p = gpu._ma_params(eps, 1e-10, 60, gpu.MA_BODY, 0)
        dg = gpu._MaDiag()
        rc = d.lib.ssg_ma_solve(d.handle, ctypes.byref(p), ctypes.byref(dg))
        if elliptic:
            assert rc == 0 and dg.converged and not dg.not_elliptic
        else:
            assert rc == gpu.SSG_ERR_NOT_CONVERGED
            assert dg.not_elliptic == 1 and dg.iterations == 0 and not dg.converged
            assert b"not elliptic" in d.lib.ssg_last_error(d.handle)
            diag = gpu.set_b(f, d, b, eps=eps, tol=1e-10, maxiter=60)
            assert diag["not_elliptic"] and not diag["converged"]
        d.close()


def _coupled_run(gpu, b0, nz, eps, dt, scheme, steps, prec="SSG_FP64"):
    nx = b0.shape[1]
    ny = b0.shape[0]
    d = gpu.make_domain(nx, ny, nz, precision=getattr(gpu, prec))
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-13, maxiter=80, variant=gpu.MA_BODY, warm=2)
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, gpu.TimeParams(dt, scheme=scheme, filter_freq=3, filter_strength=36.0, filter_order=8,
                                      filter_cutoff=0.65), st, steps)
    return d, f, st


def _oracle_coupled(o, b0, eps, dt, scheme, steps, filter_freq=3):
    from oracle.tangent import CoupledTangent
    ct = CoupledTangent(o, eps, tol=1e-13)
    b = b0.copy()
    phi = None
    fold = None
    for step in range(1, steps + 1):
        if scheme == 1:
            ks = []
            sb = b
            for s in range(3):
                phi = ct.solve(sb, phi)
                ks.append(o.tendency(ct._psi(phi), sb))
                sb = b + dt / 3 * ks[0] if s == 0 else b + 2 * dt / 3 * ks[1]
            b = b + dt / 4 * (ks[0] + 3 * ks[2])
        else:
            phi = ct.solve(b, phi)
            t = o.tendency(ct._psi(phi), b)
            b = b + dt * t if step == 1 else b + dt * (1.5 * t - 0.5 * fold)
            fold = t
        if step % filter_freq == 0:
            b = o.filter(b, 36.0, 8, 0.65)
    phi = ct.solve(b, phi)
    return b, ct._psi(phi)


@pytest.mark.parametrize("scheme", [1, 0])
def test_coupled_rectangular_against_oracle(gpu, scheme):
    """Coupled RK3 and AB2 on a non-square 32×16×6 grid (multi-thread line
    transforms along both axes), extrapolated warm starts (warm = 2) over
    enough steps to reach the same-stage quadratic extrapolation, filter every
    3 steps, against the oracle's step fed by Newton-Krylov solves of its
    literal residual equations.  After the call φₛ belongs to the returned b:
    it equals the oracle's solve for that b and a cold GPU solve of it."""
    nx, ny, nz, eps, dt, steps = 32, 16, 6, 0.1, 5e-3, 6
    b0 = spectral_b0_rect(nx, ny, 3, 2.0)
    d, f, _ = _coupled_run(gpu, b0, nz, eps, dt, scheme, steps)
    g, gphis = f.bs, f.phis
    o = Oracle(nx, ny, nz)
    b, psi = _oracle_coupled(o, b0, eps, dt, scheme, steps)
    assert np.abs(g - b).max() < 1e-9 * np.abs(b).max()
    assert np.abs((gphis - gphis.mean()) - psi).max() < 1e-9 * np.abs(psi).max()
    # the same φₛ as a cold standalone solve of the returned b
    d2, f2, diag = gpu_solve_rect(gpu, g, nz, eps)
    assert diag["converged"]
    assert np.abs(f2.phis - gphis).max() < 1e-10 * np.abs(gphis).max()
    d.close()
    d2.close()


def spectral_b0_rect(nx, ny, seed, k0):
    from ssgamd.ics import spectral_b0
    return spectral_b0(nx, ny, seed=seed, k0=k0)


def gpu_solve_rect(gpu, b, nz, eps):
    d = gpu.make_domain(b.shape[1], b.shape[0], nz)
    f = gpu.allocate_fields(d)
    f.bs = b
    diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-13, maxiter=80, variant=gpu.MA_BODY)
    return d, f, diag


@pytest.mark.parametrize("scheme", [1, 0])
@pytest.mark.parametrize("prec", ["SSG_FP64", "SSG_FP32_MIXED"])
def test_coupled_outputs_follow_returned_b(gpu, scheme, prec):
    """After coupled steps φₛ, the velocities (CFL) and the surface kinetic
    energy belong to the returned b, in both precisions: they equal those of
    a separate context that solves the returned b cold.  A second call
    continues from the solve that ended the first."""
    n, nz, eps, dt = 128, 8, 0.1, 1e-3
    b0 = smooth_b(n, 4, 6.0)
    d, f, st = _coupled_run(gpu, b0, nz, eps, dt, scheme, 4, prec)
    gpu.timestep(f, d, gpu.TimeParams(dt, scheme=scheme, filter_freq=3, filter_strength=36.0, filter_order=8,
                                      filter_cutoff=0.65), st, 3)
    b = f.bs
    d2, f2, diag = gpu_solve_rect(gpu, b, nz, eps)
    assert diag["converged"]
    tol = 1e-9
    assert np.abs(f.phis - f2.phis).max() < tol * np.abs(f2.phis).max()
    k1, k2 = gpu.diagnostics(d)["ke_surface"], gpu.diagnostics(d2)["ke_surface"]
    assert abs(k1 - k2) < tol * k2
    c1, c2 = gpu.compute_cfl_number(f, d, dt), gpu.compute_cfl_number(f2, d2, dt)
    assert abs(c1 - c2) < tol * c2
    d.close()
    d2.close()


def _stretched_z(nz):
    """Levels on [0, 1], finer toward the surface (tanh stretching)."""
    s = np.linspace(0.0, 1.0, nz)
    return 1.0 + np.tanh(2.0 * (s - 1.0)) / np.tanh(2.0)


@pytest.mark.parametrize("rule", ["true_spacing", "custom"])
def test_custom_vertical_grid_surface_newton(gpu, rule):
    """The surface-forced problem (where the vertical grid matters) on a
    stretched grid: ε = 0 at 16²×6 and ε = 0.1 at 8²×5, against the dense
    Newton solve of the oracle's residual equations on the same z, dz (the
    spacings given explicitly, or dz from make_vertical_grid's :custom
    rule)."""
    for n, nz, eps, amp in ((16, 6, 0.0, 1.0), (8, 5, 0.1, 0.2)):
        z = _stretched_z(nz)
        if rule == "custom":
            _, dz = gpu.make_vertical_grid(nz, 1.0, "custom", z)
        else:
            dz = np.append(np.diff(z), z[-1] - z[-2])
        b = smooth_b(n, 4, 1.5) * amp
        d = gpu.make_domain(n, n, nz)
        f = gpu.allocate_fields(d)
        gpu.set_vertical_grid(d, z, None if rule == "custom" else dz)
        f.bs = b
        diag = gpu.solve_monge_ampere(f, d, eps=eps, tol=1e-13, maxiter=100, variant=gpu.MA_SURFACE)
        assert diag["converged"], diag
        phi = gpu.get_phi3d(d)
        d.close()
        o = Oracle(n, n, nz)
        o.set_vertical_grid(z, dz)
        ref, _ = ma_newton(o, np.zeros((nz, n, n)), b, eps)
        assert np.abs(demean(phi) - demean(ref)).max() < 1e-9 * np.abs(ref).max(), (n, nz, eps)
        r, _ = o.ma_residual(phi, np.zeros((nz, n, n)), b, eps)
        assert np.abs(r).max() < 1e-9


def test_vertical_grid_changes_the_solution_and_rejects(gpu):
    """A stretched grid gives a different surface solution from the uniform
    one (the spacings reach the operator); bad grids are rejected."""
    n, nz = 16, 6
    b = smooth_b(n, 4, 1.5)
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    f.bs = b
    gpu.solve_monge_ampere(f, d, eps=0.0, tol=1e-13, variant=gpu.MA_SURFACE)
    uni = f.phis
    gpu.set_vertical_grid(d, _stretched_z(nz))
    gpu.solve_monge_ampere(f, d, eps=0.0, tol=1e-13, variant=gpu.MA_SURFACE)
    assert np.abs(f.phis - uni).max() > 1e-3 * np.abs(uni).max()
    z_uni, dz_uni = gpu.make_vertical_grid(nz, 1.0)
    gpu.set_vertical_grid(d, z_uni, dz_uni)
    gpu.solve_monge_ampere(f, d, eps=0.0, tol=1e-13, variant=gpu.MA_SURFACE)
    assert np.array_equal(f.phis, uni)
    lib = d.lib
    z = _stretched_z(nz)
    bad = z[::-1].copy()
    with pytest.raises(gpu.SSGError):
        gpu.set_vertical_grid(d, bad)
    with pytest.raises(gpu.SSGError):
        gpu.set_vertical_grid(d, z[:-1])
    with pytest.raises(gpu.SSGError):
        gpu.set_vertical_grid(d, z, -np.ones(nz))
    d.close()


@pytest.mark.parametrize("nx,ny,nz,stretched", [(64, 64, 8, False), (128, 64, 8, False), (64, 64, 9, True)])
def test_vertical_velocity_against_oracle(gpu, nx, ny, nz, stretched):
    """ω equation (src/extras/omega_eq.jl:574-652): w from the GPU against
    the oracle's numpy restatement fed the same Φ and surface b, for the
    surface-forced Monge-Ampère solution at ε = 0.1 (square, rectangular and
    stretched vertical grids).  The grids leave the spectrum (peak k0 = 2)
    room, so the Nyquist bins, where the two paths' transform conventions
    differ, hold only roundoff (at 32 points per axis the aliased content of
    Φ there reaches 1e-10 and shows at 1e-8 in w)."""
    from oracle.omega import vertical_velocity
    from ssgamd.ics import spectral_b0
    eps = 0.1
    b = 0.2 * spectral_b0(nx, ny, seed=6, k0=2.0)
    d = gpu.make_domain(nx, ny, nz)
    f = gpu.allocate_fields(d)
    if stretched:
        z = _stretched_z(nz)
        dz = np.append(np.diff(z), z[-1] - z[-2])
        gpu.set_vertical_grid(d, z, dz)
    else:
        _, dz = gpu.make_vertical_grid(nz, 1.0)
    f.bs = b
    w, diag = gpu.compute_vertical_velocity(f, d, eps=eps, tol=1e-12, maxiter=80)
    assert diag["converged"]
    phi = gpu.get_phi3d(d)
    d.close()
    ref = vertical_velocity(phi, b, eps, dz)
    scale = np.abs(ref).max()
    assert scale > 1e-6
    assert np.abs(w[0]).max() < 1e-14 * scale + 1e-300 and np.abs(w[-1]).max() < 1e-14 * scale + 1e-300
    assert np.abs(w - ref).max() < 1e-10 * scale, np.abs(w - ref).max() / scale


def _coupled_stage_run(gpu, b0, nz, eps, dt, scheme, calls, prec, stage_fp32, tol=1e-10):
    """Coupled steps in several ssg_timestep calls (each ends with the FP64
    solve for the b it returns); returns b, φₛ and the surface kinetic energy
    after each call."""
    n = b0.shape[0]
    d = gpu.make_domain(n, n, nz, precision=prec)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=tol, maxiter=60, variant=gpu.MA_BODY, warm=2, stage_fp32=stage_fp32)
    st = gpu.TimeState(0.0, f)
    p = gpu.TimeParams(dt, scheme=scheme, filter_freq=5, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    outs = []
    for k in calls:
        gpu.timestep(f, d, p, st, k)
        outs.append((f.bs, f.phis, gpu.diagnostics(d)["ke_surface"]))
    return d, f, outs


@pytest.mark.parametrize("scheme", [1, 0])
def test_coupled_stage_fp32_against_fp64(gpu, scheme, monkeypatch, capfd):
    """FP32 stage solves (stage_fp32): the solves inside a timestep call run in
    FP32 to SSG_MA_STAGE_TOL32, most on their first evaluation (read from the
    solve trace); the call's last solve is FP64.  After 3 calls of
    7 steps, b, φₛ and the surface kinetic energy stay within 1e-5 of the
    FP64 coupled step, and φₛ and the kinetic energy after each call are
    those of an FP64 solve of the returned b (1e-9 of a cold solve)."""
    n, nz, eps, dt, calls = 128, 8, 0.1, 1e-3, (7, 7, 7)
    b0 = smooth_b(n, 6, 6.0)
    d64, _, ref = _coupled_stage_run(gpu, b0, nz, eps, dt, scheme, calls, gpu.SSG_FP64, False)
    d64.close()
    monkeypatch.setenv("SSG_MA_TRACE", "1")
    capfd.readouterr()
    d, f, got = _coupled_stage_run(gpu, b0, nz, eps, dt, scheme, calls, gpu.SSG_FP32_MIXED, True)
    monkeypatch.delenv("SSG_MA_TRACE")
    trace = [ln.split() for ln in capfd.readouterr().err.splitlines() if ln.startswith("ssg_ma_trace")]
    fp32 = [t for t in trace if t[1] == "fp32"]
    finals = [t for t in trace if t[1] == "fp64-final"]
    solves_per_step = 3 if scheme == 1 else 1
    # every stage solve after the first call's start is FP32; one FP64 solve ends each call
    assert len(finals) == len(calls)
    assert len(fp32) >= solves_per_step * sum(calls) - len(calls) - 2, trace
    assert all(t[6] == "conv=1" for t in fp32)
    evals = [int(t[2].split("=")[1]) for t in fp32]
    # the first solves of a run extrapolate from a short history
    assert sum(e == 1 for e in evals) >= 0.8 * len(evals) and max(evals) <= 4, evals
    for (b32, p32, k32), (b64, p64, k64) in zip(got, ref):
        eb = np.linalg.norm(b32 - b64) / np.linalg.norm(b64)
        ep = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
        ek = abs(k32 - k64) / abs(k64)
        print(f"stage_fp32 vs fp64: b {eb:.2e} phis {ep:.2e} ke {ek:.2e}", flush=True)
        assert eb < 1e-5 and ep < 1e-5 and ek < 1e-5
        assert eb > 0
    d2, f2, diag = gpu_solve(gpu, f.bs, nz, eps, gpu.MA_BODY, tol=1e-12)
    assert diag["converged"]
    assert np.abs(f.phis - f2.phis).max() < 1e-9 * np.abs(f2.phis).max()
    k1, k2 = gpu.diagnostics(d)["ke_surface"], gpu.diagnostics(d2)["ke_surface"]
    assert abs(k1 - k2) < 1e-9 * k2
    d.close()
    d2.close()


def test_coupled_stage_fp32_rejects(gpu):
    """stage_fp32 needs an FP32-mixed context and a warm start; other values
    than 0 and 1 are rejected."""
    d = gpu.make_domain(32, 32, 4)
    with pytest.raises(gpu.SSGError):
        gpu.set_coupling(d, eps=0.1, stage_fp32=True)
    d.close()
    d = gpu.make_domain(32, 32, 4, precision=gpu.SSG_FP32_MIXED)
    with pytest.raises(gpu.SSGError):
        gpu.set_coupling(d, eps=0.1, warm=0, stage_fp32=True)
    p = gpu._ma_params(0.1, 1e-10, 50, gpu.MA_BODY, 2)
    p.stage_fp32 = 2
    assert d.lib.ssg_set_coupling(d.handle, ctypes.byref(p)) == gpu.SSG_ERR_ARG
    gpu.set_coupling(d, eps=0.1, stage_fp32=True)
    d.close()


@pytest.mark.parametrize("order", ["2", "3"])
def test_stage_fp32_warm_start_order(gpu, order, monkeypatch, capfd):
    """The FP32 stage solves' same-stage warm start at 1024²×32, ε = 0.1,
    dt = 2.5e-4, 60 RK3 steps in two calls.  Linear (order 2, the default):
    once two steps of history are held every FP32 stage solve converges on
    its first evaluation.  The quadratic (order 3, SSG_MA_EXTRAP32=3) is
    kept as an option; its accepted one-evaluation iterates feed an error
    recurrence that grows from step to step until a solve needs a second
    evaluation, so it takes at least as many.  Both stay within 1e-5 of the
    FP64 coupled step (b, φₛ, surface kinetic energy)."""
    n, nz, eps, dt = 1024, 32, 0.1, 2.5e-4
    b0 = smooth_b(n, 0, 6.0)

    def run(prec, s32):
        d = gpu.make_domain(n, n, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2, stage_fp32=s32)
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                           filter_cutoff=0.65)
        for k in (30, 30):
            gpu.timestep(f, d, p, st, k)
        out = (f.bs, f.phis, gpu.diagnostics(d)["ke_surface"])
        d.close()
        return out

    b64, p64, k64 = run(gpu.SSG_FP64, False)
    monkeypatch.setenv("SSG_MA_TRACE", "1")
    if order == "3":
        monkeypatch.setenv("SSG_MA_EXTRAP32", "3")
    capfd.readouterr()
    b32, p32, k32 = run(gpu.SSG_FP32_MIXED, True)
    trace = [ln.split() for ln in capfd.readouterr().err.splitlines() if ln.startswith("ssg_ma_trace")]
    fp32 = [t for t in trace if t[1] == "fp32"]
    assert all(t[6] == "conv=1" for t in fp32)
    evals = [int(t[2].split("=")[1]) for t in fp32]
    print(f"order {order}: {len(evals)} FP32 stage solves, {sum(evals)} evaluations, "
          f"{sum(e > 1 for e in evals)} with more than one", flush=True)
    if order == "2":
        # the first call's first two steps build the history (5 FP32 solves)
        assert all(e == 1 for e in evals[6:]), evals
    eb = np.linalg.norm(b32 - b64) / np.linalg.norm(b64)
    ep = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    ek = abs(k32 - k64) / abs(k64)
    print(f"order {order}: b {eb:.2e} phis {ep:.2e} ke {ek:.2e}", flush=True)
    assert 0 < eb < 1e-5 and ep < 1e-5 and ek < 1e-5


def test_coupled_stage_fp32_full_size(gpu):
    """The precision configuration's grid, 1024²×32, ε = 0.1, dt = 2.5e-4:
    60 coupled RK3 steps (two calls) with FP32 stage solves against the FP64
    coupled step from the same cold solve.  b, φₛ and the surface kinetic
    energy stay within 1e-5, the total buoyancy is conserved to roundoff, and
    the solve that ends each call converges to the FP64 tolerance."""
    import ctypes
    n, nz, eps, dt = 1024, 32, 0.1, 2.5e-4
    b0 = smooth_b(n, 0, 6.0)
    out = {}
    for prec, s32 in ((gpu.SSG_FP64, False), (gpu.SSG_FP32_MIXED, True)):
        d = gpu.make_domain(n, n, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2, stage_fp32=s32)
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
        for k in (30, 30):
            gpu.timestep(f, d, p, st, k)
            last = gpu._MaDiag()
            d.lib.ssg_last_ma_diag(d.handle, ctypes.byref(last))
            assert last.converged and last.relative_residual <= 1e-10
        out[s32] = (f.bs, f.phis, gpu.diagnostics(d)["ke_surface"])
        d.close()
    (b32, p32, k32), (b64, p64, k64) = out[True], out[False]
    eb = np.linalg.norm(b32 - b64) / np.linalg.norm(b64)
    ep = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    ek = abs(k32 - k64) / abs(k64)
    print(f"1024^2x32 stage_fp32 vs fp64 after 60 steps: b {eb:.2e} phis {ep:.2e} ke {ek:.2e}", flush=True)
    assert 0 < eb < 1e-5 and ep < 1e-5 and ek < 1e-5
    assert abs(b32.sum() - b0.sum()) < 1e-9 * np.abs(b0).sum()
