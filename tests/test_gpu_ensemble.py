"""Batched ensemble (one launch per kernel class per RK stage over members of
uneven resolution) against single-member runs and the oracle."""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu

SIZES = [128, 256, 128, 512]


def member_ics(n, seed):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    b0 = spectral_b0(n, n, seed=seed)
    return b0, streamfunction_from_b(b0)


def run_ensemble(gpu, prec, nsteps, params, chunks=None):
    ens = gpu.Ensemble([(n, n) for n in SIZES], precision=prec)
    dts = [1e-3 * 128 / n for n in SIZES]
    for i, n in enumerate(SIZES):
        b0, psi = member_ics(n, i)
        m = ens.member(i)
        m.set_field(gpu.FIELD_PHI, psi)
        m.set_field(gpu.FIELD_B, b0)
    for ch in (chunks or [nsteps]):
        ens.timestep(params, dts, ch)
    out = [ens.member(i).get_field(gpu.FIELD_B) for i in range(len(SIZES))]
    tune = ens.tuning()
    steps = [s.step for s in ens.states]
    ens.close()
    return out, tune, steps


def test_ensemble_matches_single_member_runs(gpu):
    params = gpu.TimeParams(1.0, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    # the first call ends on a filter step (the filter's no-emission form)
    out, tune, steps = run_ensemble(gpu, 0, 25, params, chunks=[10, 15])
    print("ensemble tuning", tune, flush=True)
    assert steps == [25] * len(SIZES)
    for i, n in enumerate(SIZES):
        b0, psi = member_ics(n, i)
        d = gpu.make_domain(n, n, 1)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(1e-3 * 128 / n, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                           filter_cutoff=0.65)
        gpu.timestep(f, d, p, st, 25)
        ref = f.bs
        assert np.abs(out[i] - ref).max() < 1e-11 * np.abs(ref).max(), i
        if n == 128 and i == 0:
            o = Oracle(n, n)
            bo = b0.copy()
            for _ in range(25):
                o.timestep(bo, psi, 1, 1e-3, 10, 36.0, 8, 0.65, 0)
            assert np.abs(out[i] - bo).max() < 1e-11 * np.abs(bo).max()


def test_ensemble_fp32_mixed_against_fp64(gpu):
    params = gpu.TimeParams(1.0, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    o64, _, _ = run_ensemble(gpu, 0, 40, params)
    o32, tune, _ = run_ensemble(gpu, 1, 40, params)
    for a, b in zip(o32, o64):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5


def test_ensemble_rejects_bad_members(gpu):
    with pytest.raises(gpu.SSGError):
        gpu.Ensemble([(128, 128), (64, 64)])


def test_ensemble_ab2_matches_single_member_runs(gpu):
    """AB2 (forward Euler on the first step, then the two-level update) in
    the batched form, chunked so that a later call continues the history."""
    params = gpu.TimeParams(1.0, scheme=gpu.AB2_LowStorage, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    out, tune, steps = run_ensemble(gpu, 0, 33, params, chunks=[1, 12, 20])
    assert steps == [33] * len(SIZES)
    for i, n in enumerate(SIZES):
        b0, psi = member_ics(n, i)
        d = gpu.make_domain(n, n, 1)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(1e-3 * 128 / n, scheme=gpu.AB2_LowStorage, filter_freq=10, filter_strength=36.0,
                           filter_order=8, filter_cutoff=0.65)
        gpu.timestep(f, d, p, st, 33)
        ref = f.bs
        assert np.abs(out[i] - ref).max() < 1e-11 * np.abs(ref).max(), i
        d.close()


def _c4_members():
    """The batched configuration: 32 x 128²×16, 24 x 256²×16, 8 x 512²×32,
    ICs from seeds 0-63, ε_i ~ U[0.05, 0.2] (SplitMix64, stream 1)."""
    from ssgamd.ics import uniform_keyed
    sizes = [128] * 32 + [256] * 24 + [512] * 8
    nzs = [16] * 56 + [32] * 8
    eps = [0.05 + 0.15 * float(uniform_keyed(i, np.array(0), np.array(0), stream=1)) for i in range(64)]
    return sizes, nzs, eps


def _c4_run(gpu, prec, nsteps):
    from ssgamd.ics import spectral_b0
    sizes, nzs, eps = _c4_members()
    ens = gpu.Ensemble([(n, n, z) for n, z in zip(sizes, nzs)], precision=prec)
    phis = []
    for i, n in enumerate(sizes):
        m = ens.member(i)
        f = gpu.allocate_fields(m)
        f.bs = spectral_b0(n, n, seed=i)
        diag = gpu.solve_monge_ampere(f, m, eps=eps[i], tol=1e-10, maxiter=60, variant=gpu.MA_BODY)
        assert diag["converged"], (i, diag["relative_residual"])
        phis.append(f.phis)
    dts = [1e-3 * 128 / n for n in sizes]
    params = gpu.TimeParams(1.0, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    ens.timestep(params, dts, nsteps)
    out = [ens.member(i).get_field(gpu.FIELD_B) for i in range(len(sizes))]
    assert all(s.step == nsteps for s in ens.states)
    ens.close()
    return out, phis


def test_ensemble_c4_configuration(gpu):
    """The full 64-member configuration with each member's φₛ from its own
    body-forced Monge-Ampère solve at its ε, 50 RK3 steps in FP64 and in
    FP32-mixed: one member of each size equals its single-context run
    (1e-11), the 128² member equals the oracle, and FP32-mixed stays within
    1e-5 of FP64 on every member."""
    from ssgamd.ics import spectral_b0
    nsteps = 50
    sizes, nzs, _ = _c4_members()
    out64, phis = _c4_run(gpu, gpu.SSG_FP64, nsteps)
    out32, _ = _c4_run(gpu, gpu.SSG_FP32_MIXED, nsteps)
    for i in (0, 32, 56):
        n = sizes[i]
        print(f"C4 member {i}: {n}x{n}x{nzs[i]} against its single-context run", flush=True)
        b0 = spectral_b0(n, n, seed=i)
        d = gpu.make_domain(n, n, nzs[i])
        f = gpu.allocate_fields(d)
        f.phis = phis[i]
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        p = gpu.TimeParams(1e-3 * 128 / n, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                           filter_cutoff=0.65)
        gpu.timestep(f, d, p, st, nsteps)
        ref = f.bs
        d.close()
        assert np.abs(out64[i] - ref).max() < 1e-11 * np.abs(ref).max(), i
        if n == 128:
            o = Oracle(n, n)
            bo = b0.copy()
            for _ in range(nsteps):
                o.timestep(bo, phis[i], 1, 1e-3, 10, 36.0, 8, 0.65, 0)
            assert np.abs(out64[i] - bo).max() < 1e-11 * np.abs(bo).max()
    worst = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(out32, out64))
    print(f"C4 FP32-mixed vs FP64, worst member rel-L2 {worst:.2e}", flush=True)
    assert worst < 1e-5


def test_ensemble_with_1024_member(gpu):
    """A 1024² member (the largest batched line length and LDS footprint)
    beside a 128² one: each equals its single-context run."""
    sizes = [1024, 128]
    ens = gpu.Ensemble([(n, n) for n in sizes], precision=gpu.SSG_FP64)
    ics = [member_ics(n, 11 + i) for i, n in enumerate(sizes)]
    for i, (b0, psi) in enumerate(ics):
        m = ens.member(i)
        m.set_field(gpu.FIELD_PHI, psi)
        m.set_field(gpu.FIELD_B, b0)
    dts = [2.5e-4 * 1024 / n for n in sizes]
    params = gpu.TimeParams(1.0, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    ens.timestep(params, dts, 23)
    out = [ens.member(i).get_field(gpu.FIELD_B) for i in range(len(sizes))]
    ens.close()
    for i, n in enumerate(sizes):
        b0, psi = ics[i]
        d = gpu.make_domain(n, n, 1)
        f = gpu.allocate_fields(d)
        f.phis = psi
        f.bs = b0
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, gpu.TimeParams(dts[i], scheme=gpu.RK3, filter_freq=10, filter_strength=36.0,
                                          filter_order=8, filter_cutoff=0.65), st, 23)
        ref = f.bs
        d.close()
        assert np.abs(out[i] - ref).max() < 1e-11 * np.abs(ref).max(), n
