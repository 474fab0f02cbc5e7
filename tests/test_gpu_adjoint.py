"""Discrete adjoint of the RK3 timestep! (frozen φₛ) through the C ABI.

Pins: (1) the dot-product (transpose) identity against the ORACLE's forward
step, ⟨S d, λ⟩ = ⟨d, Sᵀ λ⟩, so the adjoint is the transpose of the
reference's own step, filter cadence included; (2) the gradient of
J = ½∫(b(T) − b_tgt)² dA against central finite differences of GPU forward
runs along random unit directions (the adjoint configuration's check)."""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu


def setup(gpu, n, seed=0):
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    b0 = spectral_b0(n, n, seed=seed)
    psi = streamfunction_from_b(b0)
    d = gpu.make_domain(n, n, 1)
    f = gpu.allocate_fields(d)
    f.phis = psi
    return d, f, b0, psi


@pytest.mark.parametrize("filter_freq", [0, 4])
def test_dot_product_against_oracle_step(gpu, filter_freq):
    n, nsteps, dt = 64, 9, 2e-2
    d, f, b0, psi = setup(gpu, n)
    params = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=filter_freq, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((n, n))
    lam = rng.standard_normal((n, n))
    o = Oracle(n, n)
    y = x.copy()
    for _ in range(nsteps):
        o.timestep(y, psi, 1, dt, filter_freq, 36.0, 8, 0.65, 0)
    # forward steps 3..3+nsteps: the adjoint needs the same filter phase
    o2 = Oracle(n, n)
    o2.step.value = 3
    y2 = x.copy()
    for _ in range(nsteps):
        o2.timestep(y2, psi, 1, dt, filter_freq, 36.0, 8, 0.65, 0)
    adj = gpu.timestep_adjoint(d, params, 3 + nsteps, nsteps, lam)
    lhs, rhs = np.sum(y2 * lam), np.sum(x * adj)
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), np.sqrt(np.sum(y2 ** 2) * np.sum(lam ** 2)))
    adj0 = gpu.timestep_adjoint(d, params, nsteps, nsteps, lam)
    assert abs(np.sum(y * lam) - np.sum(x * adj0)) < 1e-11 * np.sqrt(np.sum(y ** 2) * np.sum(lam ** 2))


def test_gradient_against_finite_differences(gpu):
    """256², RK3 dt = 1e-3, 100 steps, J = ½∫(b(T) − b_tgt)² dA, b_tgt from
    seed 1; FD along 8 random unit directions with step 1e-6 (FP64)."""
    from ssgamd.ics import spectral_b0
    n, nsteps, dt = 256, 100, 1e-3
    d, f, b0, psi = setup(gpu, n, seed=0)
    b_tgt = spectral_b0(n, n, seed=1)
    dA = (2 * np.pi / n) ** 2
    params = gpu.TimeParams(dt, scheme=gpu.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)

    def forward(b):
        f.bs = b
        st = gpu.TimeState(0.0, f)
        gpu.timestep(f, d, params, st, nsteps)
        return f.bs

    bT = forward(b0)
    grad = gpu.timestep_adjoint(d, params, nsteps, nsteps, (bT - b_tgt) * dA)
    rng = np.random.default_rng(2)
    h = 1e-6
    for k in range(8):
        # unit-RMS direction (unit amplitude per grid point).  With unit
        # 2-norm the per-point perturbation h/256 ~ 4e-9 sits within a
        # factor 1e6 of the forward run's accumulated FP64 roundoff, so the
        # difference quotient itself carries ~1e-6 relative noise
        v = rng.standard_normal((n, n))
        v /= np.sqrt(np.mean(v * v))
        # central difference of J, with J(+) - J(-) formed as the exact
        # difference of squares ½Σ(r₊ - r₋)(r₊ + r₋) dA: the same finite
        # difference without cancelling two O(10) values against each other
        rp = forward(b0 + h * v) - b_tgt
        rm = forward(b0 - h * v) - b_tgt
        fd = 0.5 * np.sum((rp - rm) * (rp + rm)) * dA / (2 * h)
        an = np.sum(grad * v)
        print(f"direction {k}: adjoint {an:.12e}  fd {fd:.12e}  rel {abs(fd - an) / abs(an):.2e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an)


@pytest.mark.parametrize("aliasing", [False, True])
@pytest.mark.parametrize("start", [0, 3])
def test_ab2_dot_product_against_oracle_step(gpu, aliasing, start):
    """AB2 windows from step 0 (forward Euler first) and from step 3 (the
    previous tendency held fixed: the oracle starts from a zero history),
    plain and with the aliased tendency of the reference's call site."""
    n, nsteps, dt, ff = 64, 8, 2e-2, 4
    d, f, b0, psi = setup(gpu, n)
    params = gpu.TimeParams(dt, scheme=gpu.AB2_LowStorage, filter_freq=ff, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65, ab2_reference_aliasing=aliasing)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((n, n))
    lam = rng.standard_normal((n, n))
    o = Oracle(n, n)
    o.step.value = start
    y = x.copy()
    for _ in range(nsteps):
        o.timestep(y, psi, 0, dt, ff, 36.0, 8, 0.65, int(aliasing))
    adj = gpu.timestep_adjoint(d, params, start + nsteps, nsteps, lam)
    lhs, rhs = np.sum(y * lam), np.sum(x * adj)
    assert abs(lhs - rhs) < 1e-11 * np.sqrt(np.sum(y ** 2) * np.sum(lam ** 2))


def test_adjoint_rejects_unsupported(gpu):
    d, f, b0, psi = setup(gpu, 32)
    with pytest.raises(gpu.SSGError):
        gpu.timestep_adjoint(d, gpu.TimeParams(1e-3, scheme=gpu.RK3_LowStorage), 4, 4, np.zeros((32, 32)))
    with pytest.raises(gpu.SSGError):
        gpu.timestep_adjoint(d, gpu.TimeParams(1e-3, scheme=gpu.RK3), 2, 4, np.zeros((32, 32)))
