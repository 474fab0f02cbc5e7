"""Pins the FP64 CPU oracle (oracle/ssg_oracle.c) before it is trusted.

Each check is chosen so that a plausible slip (wrong sign, dropped term,
transposed axis, wrong normalisation, wrong Nyquist handling) fails it:
brute-force DFTs, the independent numpy witness, and closed-form
known-answer tests restated from the reference's own test setups.
"""
import numpy as np
import pytest

from oracle import Oracle, exponential_transfer
from oracle.witness import Witness


def grid(nx, ny, lx=2 * np.pi, ly=2 * np.pi):
    x = lx / nx * np.arange(nx)
    y = ly / ny * np.arange(ny)
    X, Y = np.meshgrid(x, y)  # (ny, nx): X[j, i] = x_i
    return X, Y


def brute_rfft2(a):
    """O(N^4) DFT: S[j, i] = sum_{x,y} a[y, x] exp(-2πi (i x/nx + j y/ny)), j <= ny/2."""
    ny, nx = a.shape
    xs = np.arange(nx)
    ys = np.arange(ny)
    out = np.zeros((ny // 2 + 1, nx), dtype=complex)
    for j in range(ny // 2 + 1):
        for i in range(nx):
            ph = np.exp(-2j * np.pi * (i * xs[None, :] / nx + j * ys[:, None] / ny))
            out[j, i] = np.sum(a * ph)
    return out


@pytest.mark.parametrize("nx,ny", [(8, 8), (16, 8), (8, 16), (12, 10), (6, 9)])
def test_rfft2_matches_brute_force_dft(nx, ny):
    rng = np.random.default_rng(nx * 100 + ny)
    a = rng.standard_normal((ny, nx))
    o = Oracle(nx, ny)
    s = o.rfft2(a)
    ref = brute_rfft2(a)
    assert np.abs(s - ref).max() < 1e-12 * np.abs(ref).max()
    # round trip is the identity with the reference's 1/(nx*ny) normalisation
    assert np.abs(o.irfft2(s) - a).max() < 1e-13


@pytest.mark.parametrize("nx,ny", [(8, 8), (32, 16), (64, 64), (12, 10)])
def test_oracle_matches_numpy_witness(nx, ny):
    rng = np.random.default_rng(7)
    o, w = Oracle(nx, ny), Witness(nx, ny)
    a = rng.standard_normal((ny, nx))
    p = rng.standard_normal((ny, nx))
    s = rng.standard_normal((ny // 2 + 1, nx)) + 1j * rng.standard_normal((ny // 2 + 1, nx))
    assert np.abs(o.rfft2(a) - w.rfft2(a)).max() < 1e-12 * nx * ny
    # non-Hermitian input: the c2r along y drops Im at the DC/Nyquist bins
    assert np.abs(o.irfft2(s) - w.irfft2(s)).max() < 1e-14
    scale = np.abs(w.tendency(p, a)).max()
    assert np.abs(o.tendency(p, a) - w.tendency(p, a)).max() < 1e-13 * scale
    assert np.abs(o.tendency(p, a, True) - w.tendency(p, a, True)).max() < 1e-13 * scale
    assert np.abs(o.filter(a, 1.0) - w.filter(a, 1.0)).max() < 1e-13


def test_nyquist_convention_ignores_imaginary_dc_and_nyquist():
    nx, ny = 8, 8
    o = Oracle(nx, ny)
    s = np.zeros((5, 8), dtype=complex)
    s[0, 0] = 3.0 + 5.0j      # Im of y-DC bin is discarded
    s[4, 0] = 1.0 + 7.0j      # Im of y-Nyquist bin is discarded
    out = o.irfft2(s)
    y = np.arange(ny)
    expect = (3.0 + 1.0 * (-1.0) ** y)[:, None] / (nx * ny) * np.ones((1, nx))
    assert np.abs(out - expect).max() < 1e-15


def test_wavenumbers_and_mask():
    o = Oracle(128, 128)
    kx, ky, mask = o.wavenumbers()
    assert kx[64] == -64.0 and kx[63] == 63.0 and kx[127] == -1.0     # fftfreq, Nyquist negative
    assert ky[64] == 64.0 and ky.shape == (65,)                           # rfftfreq
    # kept kx in [-42, 41], kept ky in [0, 42]  (src/domain.jl:498-514)
    kept_kx = sorted(int(round(k)) for k, m in zip(kx, mask[0]) if m)
    assert kept_kx == list(range(-42, 42))
    kept_ky = [j for j in range(65) if mask[j, 0]]
    assert kept_ky == list(range(0, 43))


def test_k1_jacobian_closed_form():
    """test/test_nonlinear.jl:14-84 setup; expectation restated: -J = sin²x sin²y - cos²x cos²y."""
    nx = ny = 32
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny)
    out = o.jacobian(np.sin(X) * np.cos(Y), np.cos(X) * np.sin(Y))
    expect = np.sin(X) ** 2 * np.sin(Y) ** 2 - np.cos(X) ** 2 * np.cos(Y) ** 2
    assert np.abs(out - expect).max() < 1e-13


def test_k2_tendency_closed_form():
    """test/test_nonlinear.jl:150-192 (0.1x amplitudes on 16²)."""
    nx = ny = 16
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny)
    out = o.tendency(0.1 * np.sin(X) * np.cos(Y), 0.1 * np.cos(X) * np.sin(Y))
    expect = -0.01 * (np.cos(X) ** 2 * np.cos(Y) ** 2 - np.sin(X) ** 2 * np.sin(Y) ** 2)
    assert np.abs(out - expect).max() < 1e-15


def test_k3_sum_b_conserved_rk3():
    """test/test_nonlinear.jl:194-246: Σb constant over 10 RK3 steps."""
    nx = ny = 32
    rng = np.random.default_rng(3)
    o = Oracle(nx, ny)
    b = np.ascontiguousarray(rng.standard_normal((ny, nx)))
    psi = rng.standard_normal((ny, nx))
    s0 = b.sum()
    for _ in range(10):
        assert o.timestep(b, psi, 1, 1e-3, filter_freq=0) == 0
    assert abs(b.sum() - s0) <= 1e-12 * max(1.0, np.abs(b).sum())


def test_k4_rk3_linear_decay():
    """test/test_timestep.jl:13-82: db/dt = -b, dt=0.01, 10 steps."""
    o = Oracle(8, 8)
    o.set_linear_hook(True, -1.0)
    b = np.ones((8, 8))
    for _ in range(10):
        o.timestep(b, np.zeros((8, 8)), 1, 0.01, filter_freq=0)
    h = 0.01
    expect = (1 - h + h * h / 2 - h ** 3 / 6) ** 10
    assert abs(expect - 0.904837414235516) < 1e-14   # the restated closed form
    assert np.abs(b - expect).max() < 1e-15
    assert abs(o.t.value - 0.1) < 1e-15 and o.step.value == 10


def test_k5_ab2_linear_recurrence():
    """test/test_timestep.jl:84-131: damping λ = -0.1; forward Euler first."""
    o = Oracle(16, 16)
    lam, h = -0.1, 0.001
    o.set_linear_hook(True, lam)
    rng = np.random.default_rng(5)
    b0 = 0.01 * rng.standard_normal((16, 16))
    b = b0.copy()
    for _ in range(500):
        o.timestep(b, np.zeros((16, 16)), 0, h, filter_freq=0)
    prev, cur = b0.copy(), (1 + h * lam) * b0
    for _ in range(499):
        prev, cur = cur, cur + h * (1.5 * lam * cur - 0.5 * lam * prev)
    assert np.abs(b - cur).max() < 1e-15


def test_k6_legacy_filter_mode_15_15():
    """test/test_timestep.jl:168-212 with the legacy ExponentialFilter(1, 4, 0.65)."""
    nx = ny = 32
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny)
    out = o.filter(np.sin(X) + 0.5 * np.sin(15 * X) * np.cos(15 * Y), 1.0)
    sig = np.exp(-(((15 / 16) - 0.65) / 0.35) ** 4)
    assert abs(sig - 0.6342699) < 1e-6
    assert np.abs(out - (np.sin(X) + 0.5 * sig * np.sin(15 * X) * np.cos(15 * Y))).max() < 1e-14


def test_k7_transfer_values():
    """test/test_filter.jl:43-74."""
    assert exponential_transfer(0.3, 1.0, 2, 0.5) == 1.0
    assert abs(exponential_transfer(0.8, 1.0, 2, 0.5) - np.exp(-0.36)) < 1e-15


def test_k8_low_mode_passes_exponential_filter():
    """test/test_filter.jl:126-173 setup (64², cos x sin y), exponential path."""
    X, Y = grid(64, 64)
    o = Oracle(64, 64)
    f = np.cos(X) * np.sin(Y)
    assert np.abs(o.filter(f, 36.0, 8, 0.65) - f).max() < 1e-14


def test_k10_dc_conserved():
    """test/test_filter.jl:222-259: constant field, ExponentialFilter(0.5, 4, 0.6)."""
    o = Oracle(32, 32)
    out = o.filter(np.ones((32, 32)), 0.5, 4, 0.6)
    assert abs(out.sum() - 1024.0) / 1024.0 < 1e-14


def test_k11_d_operator():
    """D(cos x cos y) = cos²x cos²y - sin²x sin²y on every level."""
    nx, ny, nz = 16, 16, 4
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny, nz)
    phi = np.stack([np.cos(X) * np.cos(Y)] * nz)
    D = o.ma_d_operator(phi)
    expect = np.cos(X) ** 2 * np.cos(Y) ** 2 - np.sin(X) ** 2 * np.sin(Y) ** 2
    assert np.abs(D - expect[None]).max() < 5e-14


def test_k12_ma_residual_zero_at_poisson_solution():
    """ε = 0, b on every level, g = 0: Φ̂ = -b̂/k² (src/poisson.jl:1002-1059) is a
    fixed point of the residual operator (src/poisson.jl:228-253)."""
    nx, ny, nz = 16, 16, 6
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny, nz)
    b2 = np.cos(2 * X) * np.sin(Y) + 0.3 * np.sin(3 * X + Y)
    phi2 = -(np.cos(2 * X) * np.sin(Y) / 5 + 0.3 * np.sin(3 * X + Y) / 10)
    phi = np.stack([phi2] * nz)
    r, phi_bc = o.ma_residual(phi, np.stack([b2] * nz), None, 0.0)
    assert np.abs(r).max() < 1e-14
    assert np.abs(phi_bc - phi).max() < 1e-14


def test_ma_residual_eps_closed_form():
    """ε > 0, b on every level, g = 0 (src/poisson.jl:228-253): Φ = A cos x +
    B cos y on every level has Φxx Φyy - Φxy² = AB cos x cos y and Φzz = 0,
    so b = -A cos x - B cos y - εAB cos x cos y makes the residual vanish.
    The same Φ and b at -ε leave r = 2εAB cos x cos y on interior levels,
    which pins the sign of the ε term."""
    nx, ny, nz, eps, A, B = 16, 16, 6, 0.1, 0.8, -0.6
    X, Y = grid(nx, ny)
    o = Oracle(nx, ny, nz)
    phi2 = A * np.cos(X) + B * np.cos(Y)
    det = A * B * np.cos(X) * np.cos(Y)
    b2 = -A * np.cos(X) - B * np.cos(Y) - eps * det
    phi = np.stack([phi2] * nz)
    b3 = np.stack([b2] * nz)
    assert np.abs(o.ma_d_operator(phi) - det[None]).max() < 1e-13
    r, phi_bc = o.ma_residual(phi, b3, None, eps)
    assert np.abs(r).max() < 1e-13
    assert np.abs(phi_bc - phi).max() < 1e-14
    r_flip, _ = o.ma_residual(phi, b3, None, -eps)
    assert np.abs(r_flip[1:-1] - 2 * eps * det[None]).max() < 1e-13
    assert np.abs(r_flip[0]).max() == 0 and np.abs(r_flip[-1]).max() == 0   # BC rows zero r


def test_stiff_configuration_is_not_elliptic():
    """The stiff configuration (2048², seed 7, spectrum peak k0 = 12, rms 1,
    ε = 0.3): the body-forced problem's z-independent fixed point has
    det(I - εH) = 1 - εb, so an elliptic solution needs ε·max(b) < 1.  Its
    b0 exceeds 1/ε at a few hundred points, so no elliptic solution exists
    (the solver rejects such a solve up front)."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ssg.jl_amd"))
    from ssgamd.ics import spectral_b0
    b = spectral_b0(2048, 2048, seed=7, k0=12.0)
    assert abs(np.sqrt(np.mean(b * b)) - 1.0) < 1e-12
    assert 0.3 * b.max() > 1.04
    assert 300 < int((0.3 * b >= 1.0).sum()) < 400
    # the identity itself, on a smooth field: det(I - εH) = 1 - ε(tr H - ε det H)
    nx = 64
    X, Y = grid(nx, nx)
    phi = 0.7 * np.cos(X) * np.sin(2 * Y) + 0.2 * np.sin(3 * X)
    hxx = -0.7 * np.cos(X) * np.sin(2 * Y) - 1.8 * np.sin(3 * X)
    hyy = -2.8 * np.cos(X) * np.sin(2 * Y)
    hxy = -1.4 * np.sin(X) * np.cos(2 * Y)
    eps = 0.3
    lhs = (1 - eps * hxx) * (1 - eps * hyy) - (eps * hxy) ** 2
    bb = (hxx + hyy) - eps * (hxx * hyy - hxy ** 2)
    assert np.abs(lhs - (1 - eps * bb)).max() < 1e-13
    assert phi.shape == (nx, nx)


def test_ma_boundary_rows_one_sided_neumann():
    """BC rows (src/poisson.jl:309-364): a quadratic in z with slope g at the top and
    0 at the bottom satisfies both one-sided formulas exactly."""
    nx, ny, nz = 8, 8, 5
    o = Oracle(nx, ny, nz, lz=1.0)
    z = np.linspace(0, 1, nz)
    g = 0.7
    prof = 0.5 * g * z ** 2          # φ'(0) = 0, φ'(1) = g
    phi = np.ones((nz, ny, nx)) * prof[:, None, None]
    phi_bad = phi.copy()
    phi_bad[0] += 3.0
    phi_bad[-1] -= 2.0
    out = o.ma_apply_bc(phi_bad, np.full((ny, nx), g))
    assert np.abs(out - phi).max() < 1e-14


def test_vertical_grid_rules_and_nonuniform_rows():
    """make_vertical_grid (src/domain.jl:238-281): uniform dz = Lz/(Nz-1)
    with the last entry duplicated; :custom dz centred inside and one-sided
    at the ends.  On a non-uniform grid given its true spacings, the
    d2dz2! stencil and both one-sided Neumann rows are exact for a quadratic
    profile: Φ = a z² has Φzz = 2a and slopes 0 at z = 0, 2a at z = 1."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ssg.jl_amd"))
    from ssgamd import make_vertical_grid
    z, dz = make_vertical_grid(5, 2.0)
    assert np.allclose(z, [0, 0.5, 1.0, 1.5, 2.0]) and np.allclose(dz, 0.5)
    z, dz = make_vertical_grid(5, 1.0, "custom", [0.0, 0.1, 0.3, 0.6, 1.0])
    assert np.allclose(dz, [0.1, 0.15, 0.25, 0.35, 0.4])
    nx, ny, nz, a = 8, 8, 6, 0.7
    z = np.array([0.0, 0.1, 0.25, 0.45, 0.7, 1.0])
    h = np.append(np.diff(z), z[-1] - z[-2])
    o = Oracle(nx, ny, nz)
    o.set_vertical_grid(z, h)
    phi = np.ones((nz, ny, nx)) * (a * z ** 2)[:, None, None]
    r, phi_bc = o.ma_residual(phi, np.full((nz, ny, nx), 2 * a), np.full((ny, nx), 2 * a), 0.0)
    assert np.abs(r).max() < 1e-12
    assert np.abs(phi_bc - phi).max() < 1e-13
    # the centred :custom spacings are not the true spacings on this grid:
    # the same profile then leaves a residual
    _, dzc = make_vertical_grid(nz, 1.0, "custom", z)
    o.set_vertical_grid(z, dzc)
    r, _ = o.ma_residual(phi, np.full((nz, ny, nx), 2 * a), None, 0.0)
    assert np.abs(r).max() > 1e-3


def test_omega_divq_closed_form():
    """-2ε∇·Q (src/extras/omega_eq.jl:326-405) for Φ = cos x on every level and
    b = cos 2y: ∂x v_g = Φxx = -cos x, the other gradients of u_g, v_g vanish,
    ∂y b = -2 sin 2y, so Q = (2 cos x sin 2y, 0), ∇·Q = -2 sin x sin 2y and
    -2ε∇·Q = 4ε sin x sin 2y."""
    from oracle.omega import divq_h
    nx, ny, nz, eps = 16, 16, 4, 0.3
    X, Y = grid(nx, ny)
    phi = np.stack([np.cos(X)] * nz)
    f = divq_h(phi, np.cos(2 * Y), eps)
    expect = 4 * eps * np.sin(X) * np.sin(2 * Y)
    for k in range(nz):
        assert np.abs(np.fft.irfft2(f[k], s=(ny, nx)) - expect).max() < 1e-13


def test_omega_vertical_solve_converges_to_closed_form():
    """(∂zz - k²) w = f with w = 0 at both ends: for f = sin(πz) on z ∈ [0, 1]
    the solution is -sin(πz)/(k² + π²); the three-point solve converges to it
    at second order (error ratio 4 per halving of dz), on a uniform grid and
    with the stencil weights of a stretched one."""
    from oracle.omega import vertical_solve
    k2 = np.array([0.0, 4.0, 25.0])
    errs = []
    for nz in (17, 33, 65):
        z = np.linspace(0, 1, nz)
        dz = np.full(nz, 1.0 / (nz - 1))
        f = np.sin(np.pi * z)[:, None] * np.ones(3)[None, :]
        w = vertical_solve(f, dz, k2)
        exact = -np.sin(np.pi * z)[:, None] / (k2[None, :] + np.pi ** 2)
        errs.append(np.abs(w - exact).max())
    assert 3.8 < errs[0] / errs[1] < 4.2 and 3.8 < errs[1] / errs[2] < 4.2
    # stretched grid given its true spacings: still consistent (error falls with nz)
    e2 = []
    for nz in (33, 129):
        s = np.linspace(0, 1, nz)
        z = 1 + np.tanh(1.5 * (s - 1)) / np.tanh(1.5)
        dz = np.append(np.diff(z), z[-1] - z[-2])
        f = np.sin(np.pi * z)[:, None] * np.ones(1)[None, :]
        w = vertical_solve(f, dz, np.array([4.0]))
        e2.append(np.abs(w[:, 0] + np.sin(np.pi * z) / (4 + np.pi ** 2)).max())
    assert e2[1] < e2[0] / 5


def test_omega_reference_integrals_do_not_solve_the_equation():
    """The vertical integrals of src/extras/omega_eq.jl:408-521 (the path this
    restatement replaces) leave w(-H) != 0 and (∂zz - k²) w != f."""
    k, nz = 2.0, 2001
    z = np.linspace(-1.0, 0.0, nz)
    f = np.sin(np.pi * (z + 1.0)) + 0.3 * z
    tr = lambda x, y: np.trapezoid(y, x)
    i1 = np.array([-tr(z[i:], np.exp(-k * z[i:]) * f[i:] / (2 * k)) * np.exp(k * z[i]) if i < nz - 1 else 0.0
                   for i in range(nz)])
    i2 = np.array([-tr(z[i:], np.exp(k * z[i:]) * f[i:] / (2 * k)) * np.exp(-k * z[i]) if i < nz - 1 else 0.0
                   for i in range(nz)])
    H = 1.0
    i2v = -np.exp(-k * H) * tr(z, np.exp(-k * z) * f / (2 * k)) + np.exp(k * H) * tr(z, np.exp(k * z) * f / (2 * k))
    w = i1 + i2 + 2 * i2v / (-2 * np.sinh(k * H)) * np.sinh(k * z)
    dz = z[1] - z[0]
    res = (w[2:] - 2 * w[1:-1] + w[:-2]) / dz ** 2 - k * k * w[1:-1] - f[1:-1]
    assert abs(w[0]) > 0.05 and np.abs(res).max() > 1.0


def test_parsevalsum2_reference_weights():
    """src/transforms.jl:230-274: weights 1 at x-index 0 and Nx/2, 2 for 0 < i < Nx/2, 0 above."""
    nx, ny = 8, 8
    o = Oracle(nx, ny, 1)
    s = np.zeros((5, 8), dtype=complex)
    s[1, 0] = 1.0      # x-index 0: weight 1
    s[2, 4] = 2.0      # x-index Nx/2: weight 1
    s[0, 2] = 1.0j     # 0 < i < Nx/2: weight 2
    s[3, 6] = 5.0      # i > Nx/2: weight 0
    expect = (1 + 4 + 2) * (2 * np.pi) ** 2 * 1.0 / (nx * nx * ny * ny * 1)
    assert abs(o.parsevalsum2(s) - expect) < 1e-15


def test_rk3_lowstorage_is_rejected():
    """src/timestep.jl:275-277 raises for RK3_LowStorage."""
    o = Oracle(8, 8)
    b = np.zeros((8, 8))
    assert o.timestep(b, np.zeros((8, 8)), 2, 0.1) == -1


def test_ab2_quirk_is_minus_two_dy_vb():
    """Appendix-A aliasing: the AB2 call site's tendency is -2·Dealias[∂y(v b)]."""
    nx = ny = 32
    rng = np.random.default_rng(11)
    w = Witness(nx, ny)
    o = Oracle(nx, ny)
    p = rng.standard_normal((ny, nx))
    b = rng.standard_normal((ny, nx))
    _, t2 = w.jacobian_terms(p, b)
    expect = w.irfft2(w.dealias(w.rfft2(-2 * t2)))
    assert np.abs(o.tendency(p, b, True) - expect).max() < 1e-12 * np.abs(expect).max()
    assert np.abs(o.tendency(p, b, True) - o.tendency(p, b, False)).max() > 1e-3


def _smooth_field(n, seed, k0):
    """Random-phase field with spectrum ∝ k² exp(-(k/k0)²), unit RMS."""
    rng = np.random.default_rng(seed)
    k = np.sqrt(np.fft.fftfreq(n, 1.0 / n)[:, None] ** 2 + np.fft.rfftfreq(n, 1.0 / n)[None, :] ** 2)
    amp = k * np.exp(-0.5 * (k / k0) ** 2)
    spec = amp * np.exp(2j * np.pi * rng.random(amp.shape))
    spec[0, 0] = 0.0
    f = np.fft.irfft2(spec, s=(n, n))
    return f / np.sqrt(np.mean(f * f))


def test_tangent_oracle_pins():
    """The forward-mode oracle (Newton-Krylov Monge-Ampère solve and the
    tangent-linear coupled RK3 step): its solve agrees with the dense-Newton
    oracle, and its tangents with central differences of its own solves and
    runs (the truncation error of the differences, about h² times the third
    derivative, and their roundoff, about 1e-16/h, both sit near 1e-10)."""
    from oracle.ma_newton import ma_newton
    from oracle.tangent import CoupledTangent

    def top(p):
        return p[-1] - p[-1].mean()

    eps = 0.1
    o = Oracle(8, 8, 4)
    b = _smooth_field(8, 5, 1.2)
    ct = CoupledTangent(o, eps)
    phi = ct.solve(b)
    phin, _ = ma_newton(o, np.stack([b] * 4), None, eps)
    assert np.abs(top(phi) - top(phin)).max() < 1e-9 * np.abs(top(phin)).max()

    o = Oracle(16, 16, 4)
    b0 = _smooth_field(16, 0, 3.0)
    ct = CoupledTangent(o, eps)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((16, 16))
    phi = ct.solve(b0)
    dphi = ct.tangent(phi, b0, v)
    h = 1e-5
    fd = (top(ct.solve(b0 + h * v, phi)) - top(ct.solve(b0 - h * v, phi))) / (2 * h)
    assert np.abs(top(dphi) - fd).max() < 1e-8 * np.abs(fd).max()

    bT, (dbT,) = ct.run(b0, [v], 2e-3, 3, filter_freq=2)
    h = 1e-6
    bp, _ = ct.run(b0 + h * v, [], 2e-3, 3, filter_freq=2)
    bm, _ = ct.run(b0 - h * v, [], 2e-3, 3, filter_freq=2)
    fd = (bp - bm) / (2 * h)
    assert np.abs(dbT - fd).max() < 1e-7 * np.abs(fd).max()
