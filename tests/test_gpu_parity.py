"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle on
identical seeded inputs.  FP64 bar: relative max error 1e-11 (FFT roundoff
of a different factorisation; bitwise equality is not defined across FFT
algorithms).  FP32-mixed bar: relative error 1e-5 (the precision metric's
stated tolerance)."""
import numpy as np
import pytest

from oracle import Oracle
from oracle.witness import Witness

pytestmark = pytest.mark.gpu

TOL64 = 1e-11
TOL32 = 1e-5


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def grid(nx, ny, lx=2 * np.pi, ly=2 * np.pi):
    x = lx / nx * np.arange(nx)
    y = ly / ny * np.arange(ny)
    return np.meshgrid(x, y)


SIZES = [(8, 8), (16, 16), (32, 16), (16, 64), (64, 64), (128, 128), (256, 128)]


@pytest.mark.parametrize("nx,ny", SIZES)
def test_rfft2_irfft2(gpu, nx, ny):
    print(f"rfft2 {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx + 3 * ny)
    d = gpu.make_domain(nx, ny, 1)
    o = Oracle(nx, ny)
    a = rng.standard_normal((ny, nx))
    s = gpu.rfft_2d(d, a)
    assert rel(s, o.rfft2(a)) < TOL64
    # irfft of an arbitrary (non-Hermitian) reference-layout array: the
    # reference's c2r convention must be reproduced
    z = rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape)
    assert rel(gpu.irfft_2d(d, z), o.irfft2(z)) < TOL64
    assert rel(gpu.irfft_2d(d, s), a) < TOL64


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("prec", [0, 1])
def test_tendency_and_jacobian(gpu, nx, ny, prec):
    print(f"tendency {nx}x{ny} prec={prec}", flush=True)
    rng = np.random.default_rng(11 * nx + ny)
    d = gpu.make_domain(nx, ny, 1, precision=prec)
    f = gpu.allocate_fields(d)
    o = Oracle(nx, ny)
    # deliberately not band-limited: exercises the Nyquist rows and columns
    psi = rng.standard_normal((ny, nx))
    b = rng.standard_normal((ny, nx))
    f.phis = psi
    f.bs = b
    tol = TOL64 if prec == 0 else TOL32 * 10
    w = Witness(nx, ny)
    assert rel(f.u, -w.irfft2(w.ddy(w.rfft2(psi)))) < (TOL64 if True else tol)
    assert rel(f.v, w.irfft2(w.ddx(w.rfft2(psi)))) < TOL64
    assert rel(gpu.compute_tendency(f, d), o.tendency(psi, b)) < tol
    assert rel(gpu.compute_tendency(f, d, True), o.tendency(psi, b, True)) < tol
    assert rel(gpu.compute_jacobian(f, d), o.jacobian(psi, b)) < tol


def test_k1_k2_closed_forms(gpu):
    X, Y = grid(32, 32)
    d = gpu.make_domain(32, 32, 4)
    f = gpu.allocate_fields(d)
    f.phis = np.sin(X) * np.cos(Y)
    f.bs = np.cos(X) * np.sin(Y)
    expect = np.sin(X) ** 2 * np.sin(Y) ** 2 - np.cos(X) ** 2 * np.cos(Y) ** 2
    assert np.abs(gpu.compute_jacobian(f, d) - expect).max() < 1e-13
    X, Y = grid(16, 16)
    d = gpu.make_domain(16, 16, 4)
    f = gpu.allocate_fields(d)
    f.phis = 0.1 * np.sin(X) * np.cos(Y)
    f.bs = 0.1 * np.cos(X) * np.sin(Y)
    expect = -0.01 * (np.cos(X) ** 2 * np.cos(Y) ** 2 - np.sin(X) ** 2 * np.sin(Y) ** 2)
    assert np.abs(gpu.compute_tendency(f, d) - expect).max() < 1e-15


@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 32), (128, 128)])
def test_filter(gpu, nx, ny):
    rng = np.random.default_rng(5)
    d = gpu.make_domain(nx, ny, 1)
    f = gpu.allocate_fields(d)
    o = Oracle(nx, ny)
    b = rng.standard_normal((ny, nx))
    f.bs = b
    gpu.apply_spectral_filter(f, d, 36.0, 8, 0.65)
    assert rel(f.bs, o.filter(b, 36.0, 8, 0.65)) < TOL64
    f.bs = b
    gpu.apply_spectral_filter(f, d, 1.0)
    assert rel(f.bs, o.filter(b, 1.0)) < TOL64


def run_pair(gpu, nx, ny, scheme, nsteps, dt, prec=0, filter_freq=10, strength=36.0, order=8, cutoff=0.65,
             aliasing=False, seed=0, chunks=(None,)):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    b0 = spectral_b0(nx, ny, seed=seed)
    psi = streamfunction_from_b(b0)
    d = gpu.make_domain(nx, ny, 1, precision=prec)
    f = gpu.allocate_fields(d)
    f.phis = psi
    f.bs = b0
    params = gpu.TimeParams(dt, scheme=scheme, filter_freq=filter_freq, filter_strength=strength,
                            filter_order=order, filter_cutoff=cutoff, ab2_reference_aliasing=aliasing)
    st = gpu.TimeState(0.0, f)
    done = 0
    for ch in chunks:
        n = nsteps - done if ch is None else ch
        gpu.timestep(f, d, params, st, n)
        done += n
    assert done == nsteps and st.step == nsteps
    o = Oracle(nx, ny)
    bo = b0.copy()
    for _ in range(nsteps):
        assert o.timestep(bo, psi, int(scheme), dt, filter_freq, strength, order, cutoff, int(aliasing)) == 0
    assert abs(st.t - o.t.value) < 1e-12
    return f.bs, bo


@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 128), (128, 128)])
def test_rk3_steps_match_oracle(gpu, nx, ny):
    print(f"rk3 {nx}x{ny}", flush=True)
    g, o = run_pair(gpu, nx, ny, gpu.RK3, 25, 1e-3, chunks=(7, 1, None))
    assert rel(g, o) < TOL64


@pytest.mark.parametrize("nx,ny", [(4096, 8), (8, 4096), (2048, 16)])
@pytest.mark.parametrize("prec", [0, 1])
def test_longest_lines_match_oracle(gpu, nx, ny, prec):
    """Lines of 4096 points (the largest supported axis) along x and along y."""
    print(f"long lines {nx}x{ny} prec={prec}", flush=True)
    g, o = run_pair(gpu, nx, ny, gpu.RK3, 6, 1e-4, prec=prec, filter_freq=3)
    assert rel(g, o) < (TOL64 if prec == 0 else TOL32)


@pytest.mark.parametrize("aliasing", [False, True])
def test_ab2_steps_match_oracle(gpu, aliasing):
    g, o = run_pair(gpu, 64, 64, gpu.AB2_LowStorage, 23, 2.5e-4, aliasing=aliasing, chunks=(1, 5, None))
    assert rel(g, o) < TOL64


def test_legacy_filter_cadence(gpu):
    """timestep! with the reference defaults: filter every 10 steps, α = 0.1, p = 4, c = 0.65."""
    g, o = run_pair(gpu, 64, 64, gpu.RK3, 20, 1e-3, strength=0.1, order=4, cutoff=0.65)
    assert rel(g, o) < TOL64


def test_fp32_mixed_rk3_against_fp64(gpu):
    g, o = run_pair(gpu, 128, 128, gpu.RK3, 50, 1e-3, prec=1)
    err = np.linalg.norm(g - o) / np.linalg.norm(o)
    print("fp32-mixed rel-L2 after 50 steps", err)
    assert err < TOL32


def test_diagnostics(gpu):
    rng = np.random.default_rng(9)
    nx, ny = 64, 32
    d = gpu.make_domain(nx, ny, 4)
    f = gpu.allocate_fields(d)
    o = Oracle(nx, ny, 4)
    b = rng.standard_normal((ny, nx))
    psi = rng.standard_normal((ny, nx))
    f.bs = b
    f.phis = psi
    dg = gpu.diagnostics(d)
    assert abs(dg["sum_b"] - b.sum()) < 1e-10
    assert abs(dg["sum_b2"] - (b * b).sum()) < 1e-9
    assert abs(dg["parseval_b"] - o.parsevalsum2(o.rfft2(b))) < 1e-10 * abs(dg["parseval_b"])
    u, v = f.u, f.v
    ke = 0.5 * o.parsevalsum2(o.rfft2(u)) + 0.5 * o.parsevalsum2(o.rfft2(v))
    assert abs(dg["ke_surface"] - ke) < 1e-10 * ke
    assert dg["nonfinite"] == 0
    area = (2 * np.pi) ** 2 / (nx * ny)
    assert abs(dg["ke_true"] - 0.5 * np.sum(u * u + v * v) * area) < 1e-12 * dg["ke_true"]
    # closed form: ψ = sin x cos 2y gives u = 2 sin x sin 2y, v = cos x cos 2y and
    # ½∫∫(u² + v²) = ½ (4π² + π²); the reference's formula counts x-modes 1..Nx/2-1
    # twice and multiplies by Lz
    X, Y = np.meshgrid(2 * np.pi / nx * np.arange(nx), 2 * np.pi / ny * np.arange(ny))
    f.phis = np.sin(X) * np.cos(2 * Y)
    dg = gpu.diagnostics(d)
    assert abs(dg["ke_true"] - 2.5 * np.pi ** 2) < 1e-12 * dg["ke_true"]


def test_error_paths(gpu):
    d = gpu.make_domain(16, 16, 1)
    f = gpu.allocate_fields(d)
    st = gpu.TimeState(0.0, f)
    with pytest.raises(gpu.SSGError) as e:
        gpu.timestep(f, d, gpu.TimeParams(0.1, scheme=gpu.RK3_LowStorage), st)
    assert e.value.code == gpu.SSG_ERR_UNSUPPORTED
    with pytest.raises(gpu.SSGError) as e:
        gpu.timestep(f, d, gpu.TimeParams(0.1, scheme=gpu.RK3, filter_strength=-1.0), st)
    assert e.value.code == gpu.SSG_ERR_ARG
    b = np.zeros((16, 16))
    b[3, 4] = np.nan
    f.bs = b
    with pytest.raises(gpu.SSGError) as e:
        gpu.timestep(f, d, gpu.TimeParams(0.01, scheme=gpu.RK3, check_finite=True), st)
    assert e.value.code == gpu.SSG_ERR_NONFINITE
    with pytest.raises(gpu.SSGError):
        gpu.make_domain(12, 16, 1)
    z = np.zeros(17)
    assert d.lib.ssg_set_field(d.handle, gpu.FIELD_B, z.ctypes.data, 17) == gpu.SSG_ERR_ARG
    assert d.lib.ssg_set_field(d.handle, 9, np.zeros(256).ctypes.data, 256) == gpu.SSG_ERR_ARG


@pytest.mark.parametrize("nx,nz,k0,seed,dt", [(1024, 32, 6.0, 0, 2.5e-4), (2048, 64, 12.0, 7, 1e-4)])
def test_full_size_properties(gpu, nx, nz, k0, seed, dt):
    """The precision configuration's grid (1024²) and the stiff
    configuration's (2048², k0 = 12): Σb is conserved to roundoff, a rerun is
    bitwise identical and the FP32-mixed form stays within the metric's
    tolerance of the FP64 one."""
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    ny = nx
    b0 = spectral_b0(nx, ny, seed=seed, k0=k0)
    psi = streamfunction_from_b(b0)
    outs = []
    for prec in (0, 0, 1):
        d = gpu.make_domain(nx, ny, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0,
                                          filter_order=8, filter_cutoff=0.65), st, 20)
        outs.append(f.bs)
        d.close()
    assert np.array_equal(outs[0], outs[1])
    assert abs(outs[0].sum() - b0.sum()) < 1e-9 * nx / 1024
    assert np.linalg.norm(outs[2] - outs[0]) / np.linalg.norm(outs[0]) < TOL32


def test_full_size_coupled_properties(gpu):
    """The coupled step at the precision configuration's full size
    (1024²×32, Monge-Ampère solve per RK stage with the extrapolated warm
    start, dt = 2.5e-4): every stage solve converges, Σb is conserved to
    roundoff, a rerun is bitwise identical, and the FP32-mixed form stays
    within the metric's tolerance of the FP64 one."""
    import ctypes
    from ssgamd.ics import spectral_b0
    nx, nz, dt, steps = 1024, 32, 2.5e-4, 20
    b0 = spectral_b0(nx, nx, seed=0, k0=6.0)
    outs = []
    for prec in (0, 0, 1):
        d = gpu.make_domain(nx, nx, nz, precision=prec)
        f = gpu.allocate_fields(d)
        f.bs = b0
        gpu.set_coupling(d, eps=0.1, tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0,
                                          filter_order=8, filter_cutoff=0.65), st, steps)
        last = gpu._MaDiag()
        d.check(d.lib.ssg_last_ma_diag(d.handle, ctypes.byref(last)))
        assert last.converged and last.relative_residual <= 1e-10
        outs.append(f.bs)
        d.close()
    assert np.array_equal(outs[0], outs[1])
    assert abs(outs[0].sum() - b0.sum()) < 1e-9
    e = np.linalg.norm(outs[2] - outs[0]) / np.linalg.norm(outs[0])
    print(f"coupled 1024^2x32, {steps} steps: FP32-mixed vs FP64 b rel-L2 {e:.1e}", flush=True)
    assert e < TOL32


def test_graph_replay_matches_eager(gpu, monkeypatch):
    """The step graphs (whole filter cycles and the partial cycles at both ends
    of a 23-step call from every phase), replayed many times with eager work
    and synchronisations in between, some captured ahead of the call by
    ssg_timestep_prepare, give bitwise the eager launches' result."""
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    nx = ny = 256
    b0 = spectral_b0(nx, ny, seed=3)
    psi = streamfunction_from_b(b0)
    params = gpu.TimeParams(1e-3, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    runs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("SSG_GRAPH", mode)
        monkeypatch.setenv("SSG_EPT", "8,4,8")
        d = gpu.make_domain(nx, ny, 1)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        snaps = []
        for i in range(4):
            if mode == "1" and i % 2:
                gpu.timestep_prepare(d, params, st, 23)   # graphs captured ahead of the call
            gpu.timestep(f, d, params, st, 23)
            snaps.append(f.bs)
            gpu.diagnostics(d)           # eager work between replays
        runs[mode] = snaps
        d.close()
    for a, b in zip(runs["1"], runs["0"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("scheme", [1, 0])
def test_precision_configuration_against_oracle(gpu, scheme):
    """The precision configuration at full size: 1024², 1000 steps of RK3
    (the metric's scheme) or AB2 (the configuration's), dt = 2.5e-4, filter
    α = 36, p = 8, c = 0.65 every 10 steps.  Both GPU forms against the
    FP64 oracle on the same inputs: surface buoyancy rel-L2 (≤ 1e-5 for
    FP32-mixed, the metric's bar) and the buoyancy variance ½Σb²."""
    from oracle import set_num_threads
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    import os
    n, steps, dt = 1024, 1000, 2.5e-4
    b0 = spectral_b0(n, n, seed=0)
    psi = streamfunction_from_b(b0)
    params = gpu.TimeParams(dt, scheme=scheme, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    outs = {}
    for prec in (0, 1):
        d = gpu.make_domain(n, n, 32, precision=prec)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, params, st, steps)
        outs[prec] = f.bs
        d.close()
    set_num_threads(min(16, os.cpu_count() or 1))
    o = Oracle(n, n)
    bo = b0.copy()
    for _ in range(steps):
        assert o.timestep(bo, psi, scheme, dt, 10, 36.0, 8, 0.65, 0) == 0
    e64 = np.linalg.norm(outs[0] - bo) / np.linalg.norm(bo)
    e32 = np.linalg.norm(outs[1] - bo) / np.linalg.norm(bo)
    v32 = abs(np.sum(outs[1] ** 2) - np.sum(bo ** 2)) / np.sum(bo ** 2)
    print(f"scheme {scheme}, {steps} steps at {n}^2: FP64 rel-L2 {e64:.2e}, FP32-mixed rel-L2 {e32:.2e}, "
          f"variance {v32:.2e}", flush=True)
    assert e64 < 1e-11
    assert e32 < TOL32 and v32 < TOL32
