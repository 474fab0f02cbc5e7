"""Setup, drivers and restart around the step (the callers either side of
timestep!): set_b! (src/setICs.jl:21-88), compute_cfl_number /
adaptive_timestep (src/timestep.jl:91-147), step_until!
(src/timestep.jl:408-449) and an exact restart of the step state."""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu


def test_set_b_matches_oracle_pipeline(gpu):
    """b <- irfft(mask · rfft(b)) with the k = 0 mode removed, then the
    Monge-Ampère solve for φₛ and the velocities."""
    n, nz = 64, 8
    rng = np.random.default_rng(12)
    b = rng.standard_normal((n, n)) + 3.0          # not band-limited, nonzero mean
    o = Oracle(n, n, nz)
    s = o.rfft2(b)
    s[0, 0] = 0.0
    expect = o.irfft2(o.dealias2(s))
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    assert gpu.set_b(f, d, b, eps=None) is None
    assert np.abs(f.bs - expect).max() < 1e-12 * np.abs(expect).max()
    assert abs(f.bs.mean()) < 1e-14
    diag = gpu.set_b(f, d, b, eps=0.1, tol=1e-12, maxiter=60)
    assert diag["converged"]
    phis = f.phis
    # the same as an explicit solve on the set buoyancy
    d2 = gpu.make_domain(n, n, nz)
    f2 = gpu.allocate_fields(d2)
    f2.bs = expect
    gpu.solve_monge_ampere(f2, d2, eps=0.1, tol=1e-12, maxiter=60)
    assert np.abs(phis - f2.phis).max() < 1e-10 * np.abs(phis).max()
    assert np.abs(f.u - f2.u).max() < 1e-10 * np.abs(f.u).max()
    d.close()
    d2.close()


def test_cfl_and_adaptive_dt(gpu):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    n = 128
    d = gpu.make_domain(n, n, 1)
    f = gpu.allocate_fields(d)
    f.phis = streamfunction_from_b(spectral_b0(n, n, seed=2))
    dt = 0.05
    cfl = gpu.compute_cfl_number(f, d, dt)
    dx = 2 * np.pi / n
    expect = max(np.abs(f.u).max() * dt / dx, np.abs(f.v).max() * dt / dx)
    assert abs(cfl - expect) < 1e-14 * expect
    # adaptive_timestep: dt scaled to the CFL safety bound
    p = gpu.TimeParams(dt, scheme=gpu.RK3, adaptive_dt=True, cfl_safety=0.5 * cfl, min_dt=1e-8, max_dt=1.0)
    st = gpu.TimeState(0.0, f)
    f.bs = spectral_b0(n, n, seed=3)
    used = gpu.timestep(f, d, p, st, 4)
    assert abs(used - dt * 0.5) < 1e-15 and abs(st.t - 4 * used) < 1e-14 and abs(st.cfl_max - cfl) < 1e-14
    d.close()


def test_step_until_lands_on_stop_time(gpu):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    n = 64
    d = gpu.make_domain(n, n, 1)
    f = gpu.allocate_fields(d)
    b0 = spectral_b0(n, n, seed=4)
    f.phis = streamfunction_from_b(b0)
    f.bs = b0
    st = gpu.TimeState(0.0, f)
    p = gpu.TimeParams(1e-3, scheme=gpu.RK3, filter_freq=4, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    nsteps = gpu.step_until(f, d, p, st, 0.0105)
    assert nsteps == 11 and st.step == 11 and st.t == 0.0105
    # the oracle: ten steps of dt and one of 0.5 dt
    o = Oracle(n, n)
    bo = b0.copy()
    for k in range(11):
        o.timestep(bo, f.phis, 1, 1e-3 if k < 10 else 0.0105 - o.t.value, 4, 36.0, 8, 0.65, 0)
    assert np.abs(f.bs - bo).max() < 1e-11 * np.abs(bo).max()
    d.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1])
def test_restart_is_bitwise(gpu, prec, scheme):
    """Saved after 7 steps and loaded into a fresh context, 5 more steps give
    bitwise the uninterrupted run (the AB2 history travels with the state)."""
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    n = 128
    b0 = spectral_b0(n, n, seed=5)
    psi = streamfunction_from_b(b0)
    p = gpu.TimeParams(1e-3, scheme=scheme, filter_freq=3, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    d = gpu.make_domain(n, n, 1, precision=prec)
    f = gpu.allocate_fields(d)
    f.phis = psi
    f.bs = b0
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, p, st, 7)
    blob = gpu.save_state(d, st)
    gpu.timestep(f, d, p, st, 5)
    ref = f.bs
    d2 = gpu.make_domain(n, n, 1, precision=prec)
    f2 = gpu.allocate_fields(d2)
    st2 = gpu.TimeState(0.0, f2)
    gpu.load_state(d2, blob, st2)
    assert st2.step == 7
    gpu.timestep(f2, d2, p, st2, 5)
    assert np.array_equal(f2.bs, ref) and st2.step == st.step and st2.t == st.t
    with pytest.raises(gpu.SSGError):     # a blob of another grid is rejected
        d3 = gpu.make_domain(64, 64, 1, precision=prec)
        gpu.load_state(d3, blob, gpu.TimeState())
    d.close()
    d2.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1])
def test_coupled_restart_is_bitwise(gpu, prec, scheme, tmp_path):
    """Coupled stepping (a Monge-Ampère solve per stage, extrapolated warm
    starts): checkpointed after 7 steps, when the same-stage extrapolation
    ring is full, and loaded into a fresh context, 6 more steps give bitwise
    the uninterrupted run's b, φₛ and 3D φ.  The checkpoint carries the
    reference's keys (src/output.jl:356-409); re-stating the same coupling
    after the load keeps the restored history."""
    from ssgamd.ics import spectral_b0
    n, nz, eps = 128, 8, 0.1
    b0 = spectral_b0(n, n, seed=5)
    p = gpu.TimeParams(1e-3, scheme=scheme, filter_freq=3, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    d = gpu.make_domain(n, n, nz, precision=prec)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2)
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, p, st, 7)
    path = gpu.save_checkpoint(str(tmp_path / "ck.npz"), f, d, st, p)
    gpu.timestep(f, d, p, st, 6)
    ref_b, ref_phis, ref_phi = f.bs, f.phis, gpu.get_phi3d(d)
    d.close()
    for restate in (False, True):
        d2 = gpu.make_domain(n, n, nz, precision=prec)
        f2 = gpu.allocate_fields(d2)
        st2 = gpu.TimeState(0.0, f2)
        data = gpu.load_checkpoint(path, d2, st2)
        assert d2.coupled and st2.step == 7
        for k in ("time", "step", "dt", "grid/Nx", "grid/z", "grid/dz", "fields/surface/buoyancy",
                  "fields/surface/streamfunction", "fields/3d/streamfunction", "fields/3d/u_velocity",
                  "timestepper/scheme"):
            assert k in data, k
        assert data["fields/3d/streamfunction"].shape == (nz, n, n)
        if restate:
            gpu.set_coupling(d2, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2)
        gpu.timestep(f2, d2, p, st2, 6)
        assert np.array_equal(f2.bs, ref_b), restate
        assert np.array_equal(f2.phis, ref_phis), restate
        assert np.array_equal(gpu.get_phi3d(d2), ref_phi), restate
        d2.close()


@pytest.mark.parametrize("scheme", [0, 1])
def test_coupled_stage_fp32_restart_is_bitwise(gpu, scheme):
    """With FP32 stage solves the restart blob carries their FP32 history
    (widened to FP64 exactly): saved after 7 steps and loaded into a fresh
    context, 6 more steps give bitwise the uninterrupted run's b, φₛ and 3D φ."""
    from ssgamd.ics import spectral_b0
    n, nz, eps = 128, 8, 0.1
    b0 = spectral_b0(n, n, seed=5)
    p = gpu.TimeParams(1e-3, scheme=scheme, filter_freq=3, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    d = gpu.make_domain(n, n, nz, precision=gpu.SSG_FP32_MIXED)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-10, maxiter=60, variant=gpu.MA_BODY, warm=2, stage_fp32=True)
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, p, st, 7)
    blob = gpu.save_state(d, st)
    assert blob[:8] == b"SSGSTAT3"
    gpu.timestep(f, d, p, st, 6)
    ref_b, ref_phis, ref_phi = f.bs, f.phis, gpu.get_phi3d(d)
    d.close()
    d2 = gpu.make_domain(n, n, nz, precision=gpu.SSG_FP32_MIXED)
    f2 = gpu.allocate_fields(d2)
    st2 = gpu.TimeState(0.0, f2)
    gpu.load_state(d2, blob, st2)
    assert d2.coupled and st2.step == 7
    gpu.timestep(f2, d2, p, st2, 6)
    assert np.array_equal(f2.bs, ref_b)
    assert np.array_equal(f2.phis, ref_phis)
    assert np.array_equal(gpu.get_phi3d(d2), ref_phi)
    d2.close()


def test_coupled_adjoint_from_restart(gpu):
    """The coupled adjoint started after loading a checkpoint equals the one
    started in the run that wrote it (same start state, bitwise)."""
    from ssgamd.ics import spectral_b0
    n, nz, eps, dt = 64, 6, 0.1, 2e-3
    b0 = spectral_b0(n, n, seed=2, k0=4.0)
    lam = spectral_b0(n, n, seed=3, k0=4.0)
    p = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=2, filter_strength=36.0, filter_order=8, filter_cutoff=0.65)
    d = gpu.make_domain(n, n, nz)
    f = gpu.allocate_fields(d)
    f.bs = b0
    gpu.set_coupling(d, eps=eps, tol=1e-12, maxiter=100, variant=gpu.MA_BODY, warm=2)
    st = gpu.TimeState(0.0, f)
    gpu.timestep(f, d, p, st, 3)
    blob = gpu.save_state(d, st)
    gpu.set_coupling(d, eps=None)
    g1 = gpu.timestep_adjoint_coupled(d, p, 3, 4, lam, eps=eps, tol=1e-12, maxiter=100, variant=gpu.MA_BODY,
                                      warm=2, checkpoint_every=2)
    d.close()
    d2 = gpu.make_domain(n, n, nz)
    f2 = gpu.allocate_fields(d2)
    st2 = gpu.TimeState(0.0, f2)
    gpu.load_state(d2, blob, st2)
    gpu.set_coupling(d2, eps=None)
    g2 = gpu.timestep_adjoint_coupled(d2, p, st2.step, 4, lam, eps=eps, tol=1e-12, maxiter=100,
                                      variant=gpu.MA_BODY, warm=2, checkpoint_every=2)
    d2.close()
    assert np.array_equal(g1, g2)


def test_diagnostic_time_series(gpu):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    n = 64
    d = gpu.make_domain(n, n, 1)
    f = gpu.allocate_fields(d)
    b0 = spectral_b0(n, n, seed=6)
    f.phis = streamfunction_from_b(b0)
    f.bs = b0
    st = gpu.TimeState(0.0, f)
    ts = gpu.DiagnosticTimeSeries()
    p = gpu.TimeParams(1e-3, scheme=gpu.RK3, filter_freq=0)
    for _ in range(3):
        gpu.timestep(f, d, p, st, 5)
        gpu.update_diagnostics(ts, f, d, st)
    assert len(ts.times) == 3 and abs(ts.times[-1] - 0.015) < 1e-15
    # frozen φₛ: the kinetic energy is constant and Σb is conserved
    assert max(ts.kinetic_energy) - min(ts.kinetic_energy) < 1e-12 * ts.kinetic_energy[0]
    assert max(abs(x - b0.sum()) for x in ts.total_buoyancy) < 1e-10
    d.close()
