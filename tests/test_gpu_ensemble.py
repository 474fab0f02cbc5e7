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
