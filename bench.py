"""SSG time-step benchmark on one MI355X.

Workload: the precision configuration's grid, Nx = Ny = 1024 (Nz = 32 carried
by the domain), surface buoyancy b0 from the seeded random-phase spectrum
k²·exp(-(k/6)²) at rms 1, frozen φₛ = ε = 0 inversion of b0, RK3 steps of
dt = 2.5e-4 with the exponential filter (α = 36, p = 8, c = 0.65) every 10
steps.  A step is one `timestep!` (three tendency evaluations + stage
updates, plus the filter on every 10th step).

`value` is FP32-mixed RK3 steps/s with all inputs resident in HBM; the plain
FP64 drop-in is measured in the same run and reported beside it with the
rel-L2 difference of b between the two.  The roofline object covers the
dominant kernel class, timed with HIP events on the library's stream in a
second, instrumented pass over the same number of steps.  The CPU baseline
is the FP64 oracle (a C restatement of the reference step) on the host
cores, on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ssg.jl_amd"))

METRIC = "SSG RK3 steps/s @1024²×32: FP32-mixed vs FP64 at b rel-err≤1e-5; %HBM peak"
HBM_PEAK_GBS = 8000.0
KIND_NAMES = {0: "k_rows_fwd", 1: "k_cols", 2: "k_rows_inv"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def settle(ssg, d, f, params, b0, ms):
    """Untimed whole filter cycles from step 0 for about `ms` milliseconds, so
    the GPU's clocks are at their steady state when the warm-up starts; b is
    then reset to b0 and the run starts again from step 0.  Returns the number
    of steps run."""
    if ms <= 0:
        return 0
    F = params.filter_freq if params.filter_freq > 0 else 16
    st = ssg.TimeState(0.0, f)
    n = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < ms:
        ssg.timestep(f, d, params, st, F)
        d.sync()
        n += F
    f.bs = b0
    return n


def run_gpu(ssg, prec, b0, psi, nx, ny, nz, params, warmup, steps, profile, settle_ms=0.0):
    d = ssg.make_domain(nx, ny, nz, precision=prec)
    f = ssg.allocate_fields(d)
    f.phis = psi
    f.bs = b0
    settled = settle(ssg, d, f, params, b0, settle_ms)
    st = ssg.TimeState(0.0, f)
    # launch set-up (autotuning, the step graphs the warm-up and the timed
    # run replay) before the warm-up, so it stays out of the timed region and
    # the warm-up runs right before it: the rate does not depend on --steps
    ssg.timestep_prepare(d, params, st, warmup)
    ab2_later = params.scheme == ssg.AB2_LowStorage and warmup > 0   # needs the warm-up's history first
    if not ab2_later:
        at_start = ssg.TimeState(0.0)
        at_start.step = warmup
        ssg.timestep_prepare(d, params, at_start, steps)
    if warmup:
        ssg.timestep(f, d, params, st, warmup)
    if ab2_later:
        ssg.timestep_prepare(d, params, st, steps)
    d.sync()
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps)
    d.sync()
    wall = time.perf_counter() - t0
    gpu_ms, launches = d.last_timing()
    tune = d.tuning()
    b_end = f.bs
    prof = None
    if profile:
        # the same number of steps again with every launch timestamped by its
        # own dispatch (hipExtLaunchKernelGGL events on the library's stream)
        d.profile(True)
        d.sync()
        t0 = time.perf_counter()
        ssg.timestep(f, d, params, st, steps)
        d.sync()
        prof = {k: d.profile_read(k) for k in KIND_NAMES}
        prof["wall"] = time.perf_counter() - t0
        d.profile(False)
    d.close()
    return wall, gpu_ms, launches, b_end, prof, tune, settled


def load_traffic(workload):
    """Measured HBM bytes per launch of the dominant kernel (rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE, scaled by a copy-kernel calibration), if a
    summary for this workload is committed under profiles/."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            data = json.load(fh)
        return data.get(workload)
    except Exception:
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(b0, psi, nx, ny, nz, dt, budget_s):
    """The FP64 C oracle (a restatement of the reference step, not the
    reference itself) on the host cores: OpenMP over all threads it is given,
    then one thread, each on a bounded sample of the same workload."""
    from oracle import Oracle, num_threads, set_num_threads
    o = Oracle(nx, ny, nz)

    def rate(budget, max_steps):
        b = b0.copy()
        n = 0
        t0 = time.perf_counter()
        while True:
            o.timestep(b, psi, 1, dt, 10, 36.0, 8, 0.65, 0)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget or n >= max_steps:
                return n, el

    threads = num_threads()
    n, el = rate(budget_s, 1000)
    set_num_threads(1)
    n1, el1 = rate(0.6 * budget_s, 200)
    set_num_threads(threads)
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"value": n / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} RK3 steps of the same {nx}x{ny} workload (FP64 C oracle, OpenMP, {threads} threads), "
                      f"{el:.1f} s; one thread: {n1} steps, {el1:.1f} s",
            "one_core_steps_per_s": n1 / el1, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "affinity_cpus": affinity}


def measure_c1(ssg, steps):
    """The reference's own CPU-runnable case: 128²×16, seed 0, ε = 0.1:
    set_b! (the Monge-Ampère solve, body-forced as wired, relative residual
    1e-10), then `steps` RK3 steps of dt = 1e-3 with φₛ frozen (the
    reference's timestep!), filter every 10 steps, in FP64.  Returns the GPU
    figures and the inputs (b after set_b!, φₛ) and result the CPU leg uses."""
    from ssgamd.ics import spectral_b0
    n, nz, dt = 128, 16, 1e-3
    b0 = spectral_b0(n, n, seed=0, k0=6.0)
    params = ssg.TimeParams(dt, scheme=ssg.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    d = ssg.make_domain(n, n, nz)
    f = ssg.allocate_fields(d)
    t0 = time.perf_counter()
    diag = ssg.set_b(f, d, b0, eps=0.1, tol=1e-10, maxiter=60)
    d.sync()
    setup_s = time.perf_counter() - t0
    bs0, psi = f.bs, f.phis
    st = ssg.TimeState(0.0, f)
    ssg.timestep_prepare(d, params, st, steps)
    d.sync()
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps)
    d.sync()
    wall = time.perf_counter() - t0
    b_end = f.bs
    d.close()
    out = {"grid": f"{n}x{n}x{nz}", "steps": steps, "dt": dt, "eps": 0.1,
           "set_b": {"converged": diag["converged"], "iterations": diag["iterations"],
                     "relative_residual": diag["relative_residual"],
                     "absolute_residual": diag["final_residual"], "s": setup_s},
           "gpu_fp64_steps_per_s": steps / wall}
    return out, (bs0, psi, b_end)


def cpu_c1(inputs, steps):
    """The C oracle on the reference's CPU case from the same b and φₛ as
    the GPU run (all threads, then one thread): steps/s and the rel-L2 of b
    between the two after `steps` steps."""
    from oracle import Oracle, num_threads, set_num_threads
    bs0, psi, b_gpu = inputs
    n = bs0.shape[0]

    def run():
        o = Oracle(n, n, 16)
        b = bs0.copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            o.timestep(b, psi, 1, 1e-3, 10, 36.0, 8, 0.65, 0)
        return b, time.perf_counter() - t0

    threads = num_threads()
    b, el = run()
    set_num_threads(1)
    _, el1 = run()
    set_num_threads(threads)
    return {"value": steps / el, "unit": "steps/s", "cores": threads, "one_core_steps_per_s": steps / el1,
            "b_rel_l2_gpu_vs_cpu": rel_l2(b_gpu, b)}


def measure_coupled(ssg, b0, nx, ny, nz, dt, steps, prec, warm=2, step_tol=1e-10, stage_fp32=False, scheme=None):
    """Coupled step (Monge-Ampère solve at every RK stage, body-forced
    variant, ε = 0.1, relative residual `step_tol`; with `stage_fp32` the
    solves inside a call run in FP32 and the call's last solve in FP64):
    steps/s, the cold solve's iterations and time, a per-kernel-class profile
    of 10 further coupled steps (one call), and the final surface fields (b,
    φₛ) and surface kinetic energy for the cross-precision comparison."""
    import ctypes
    d = ssg.make_domain(nx, ny, nz, precision=prec)
    f = ssg.allocate_fields(d)
    f.bs = b0
    diag = ssg.solve_monge_ampere(f, d, eps=0.1, tol=1e-10, maxiter=60, variant=ssg.MA_BODY)
    cold = {"iterations": diag["iterations"], "iterations_fp32": diag["iterations_fp32"],
            "ms": diag["solve_time"] * 1e3, "relative_residual": diag["relative_residual"],
            "absolute_residual": diag["final_residual"], "converged": diag["converged"]}
    ssg.set_coupling(d, eps=0.1, tol=step_tol, maxiter=60, variant=ssg.MA_BODY, warm=warm, stage_fp32=stage_fp32)
    st = ssg.TimeState(0.0, f)
    params = ssg.TimeParams(dt, scheme=ssg.RK3 if scheme is None else scheme, filter_freq=10, filter_strength=36.0,
                            filter_order=8, filter_cutoff=0.65)
    ssg.timestep(f, d, params, st, 1)
    d.sync()
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps)
    d.sync()
    wall = time.perf_counter() - t0
    last = ssg._MaDiag()
    d.lib.ssg_last_ma_diag(d.handle, ctypes.byref(last))
    fields = {"b": f.bs, "phis": f.phis, "ke": ssg.diagnostics(d)["ke_surface"]}
    # kernel classes of stepping itself, timed per launch
    d.profile(True)
    ssg.timestep(f, d, params, st, 10)
    names = {0: "k_rows_fwd", 1: "k_cols (tendency and MA ops)", 2: "k_rows_inv", 3: "k_ma_rows",
             4: "k_ma_tridiag", 5: "k_ma_tridiag (extrapolating)", 11: "k_ma_rows (FP32)",
             12: "k_ma_tridiag (FP32)", 13: "k_ma_tridiag (FP32, extrapolating)"}
    kernels = {}
    for k, nm in names.items():
        ms, n, by = d.profile_read(k)
        if n:
            kernels[nm] = {"avg_us": 1e3 * ms / n, "launches": n, "GBps": by / max(ms, 1e-12) / 1e6,
                           "bytes_per_launch": by / n}
    d.profile(False)
    d.close()
    # roofline of the coupled step's dominant kernel, the Monge-Ampère row
    # pass in the precision the stage solves run in: algorithmic bytes per
    # launch / duration
    dom = "k_ma_rows (FP32)" if stage_fp32 else "k_ma_rows"
    kd = kernels[dom]
    form = "fp32" if stage_fp32 else "fp64"
    roof = {"bound": "hbm", "kernel": dom, "achieved": kd["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kd["GBps"] / HBM_PEAK_GBS, "traffic": load_traffic(f"k_ma_rows:{nx}x{nz}:{form}"),
            "bytes_per_launch": kd["bytes_per_launch"], "avg_launch_us": kd["avg_us"]}
    return {"steps_per_s": steps / wall, "steps": steps, "variant": "body (as wired)", "eps": 0.1, "tol": step_tol,
            "stage_solves": "fp32" if stage_fp32 else ("fp64" if prec == ssg.SSG_FP64 else "fp32 phase, fp64 residual"),
            "cold_solve": cold, "warm_start": warm, "last_call_final_solve_iterations": last.iterations,
            "last_call_final_solve_relative_residual": last.relative_residual, "ma_kernels": kernels,
            "roofline": roof}, fields


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def measure_ab2(ssg, b0, psi, nx, ny, nz, dt, steps):
    """AB2 (the precision configuration's scheme) at the same grid: FP64 and
    FP32-mixed steps/s and the rel-L2 of b between them after `steps`."""
    params = ssg.TimeParams(dt, scheme=ssg.AB2_LowStorage, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    out = {}
    ends = {}
    for name, prec in (("fp64", ssg.SSG_FP64), ("fp32", ssg.SSG_FP32_MIXED)):
        wall, gpu_ms, launches, b_end, _, _, _ = run_gpu(ssg, prec, b0, psi, nx, ny, nz, params, 0, steps, False,
                                                         settle_ms=50.0)
        out[f"{name}_steps_per_s"] = steps / wall
        ends[name] = b_end
    out["fp32_over_fp64"] = out["fp32_steps_per_s"] / out["fp64_steps_per_s"]
    out["b_rel_l2_fp32_vs_fp64"] = rel_l2(ends["fp32"], ends["fp64"])
    out["steps"] = steps
    return out


def measure_adjoint(ssg, steps=100, every=10):
    """Adjoint configuration, 256²×16, ε = 0.1, RK3 dt = 1e-3, J = ½∫(b(T) -
    b_tgt)² dA: wall time of one gradient (vector-Jacobian product through
    `steps` steps) for the frozen-φₛ step and for the coupled step
    (Monge-Ampère solve per stage, checkpoints every `every` steps), beside
    the forward run's time."""
    from ssgamd.ics import spectral_b0, streamfunction_from_b
    n, nz, dt, eps = 256, 16, 1e-3, 0.1
    b0 = spectral_b0(n, n, seed=0)
    b_tgt = spectral_b0(n, n, seed=1)
    dA = (2 * np.pi / n) ** 2
    params = ssg.TimeParams(dt, scheme=ssg.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    out = {"grid": f"{n}x{n}x{nz}", "steps": steps, "checkpoint_every": every, "eps": eps}
    d = ssg.make_domain(n, n, nz)
    f = ssg.allocate_fields(d)
    # frozen φₛ
    f.phis = streamfunction_from_b(b0)
    f.bs = b0
    st = ssg.TimeState(0.0, f)
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps)
    bT = f.bs
    t1 = time.perf_counter()
    ssg.timestep_adjoint(d, params, steps, steps, (bT - b_tgt) * dA)
    t2 = time.perf_counter()
    out["frozen_forward_s"] = t1 - t0
    out["frozen_adjoint_s"] = t2 - t1
    # coupled
    f.bs = b0
    ssg.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=100, variant=ssg.MA_BODY)
    ssg.set_coupling(d, eps=eps, tol=1e-10, maxiter=100, variant=ssg.MA_BODY, warm=2)
    st = ssg.TimeState(0.0, f)
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps)
    bT = f.bs
    t1 = time.perf_counter()
    ssg.set_coupling(d, eps=None)
    f.bs = b0
    g = ssg.timestep_adjoint_coupled(d, params, 0, steps, (bT - b_tgt) * dA, eps=eps, tol=1e-10, maxiter=100,
                                     variant=ssg.MA_BODY, warm=2, checkpoint_every=every)
    t2 = time.perf_counter()
    d.close()
    out["coupled_forward_s"] = t1 - t0
    out["coupled_adjoint_s"] = t2 - t1
    out["coupled_adjoint_over_forward"] = (t2 - t1) / (t1 - t0)
    out["grad_norm"] = float(np.linalg.norm(g))
    return out


def measure_stiff(ssg, steps, adj_steps, eps=0.25):
    """Large stiff configuration: 2048²×64, spectrum peak k0 = 12, seed 7, RK3
    dt = 1e-4.  Its ε = 0.3 leaves no elliptic solution for this b0 (at the
    body-forced problem's z-independent fixed point det(I - εH) = 1 - εb, and
    ε·max(b0) = 1.05): that solve is attempted and its rejection recorded, and
    the run proceeds at ε = 0.25 (ε·max(b0) = 0.875).  Measured: the cold
    Monge-Ampère solves, FP32-mixed steps with φₛ frozen at the ε = 0.25
    solution for `steps` steps, then the configuration's forward itself:
    `steps` FP32-mixed coupled steps from b0 (a solve per RK stage, FP32
    stage solves), and the FP64 checkpointed adjoint of the last `adj_steps`
    of them (checkpoints every 10 steps)."""
    from ssgamd.ics import spectral_b0
    n, nz, dt = 2048, 64, 1e-4
    b0 = spectral_b0(n, n, seed=7, k0=12.0)
    params = ssg.TimeParams(dt, scheme=ssg.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    out = {"grid": f"{n}x{n}x{nz}", "k0": 12, "eps_max_b0_at_0.3": 0.3 * float(b0.max()), "eps": eps,
           "steps": steps, "coupled_steps": adj_steps}
    d = ssg.make_domain(n, n, nz, precision=ssg.SSG_FP32_MIXED)
    f = ssg.allocate_fields(d)
    f.bs = b0
    diag = ssg.solve_monge_ampere(f, d, eps=0.3, tol=1e-10, maxiter=100, variant=ssg.MA_BODY)
    out["ma_eps_0.3"] = {"converged": diag["converged"], "not_elliptic": diag["not_elliptic"],
                         "iterations": diag["iterations"], "ms": diag["solve_time"] * 1e3}
    diag = ssg.solve_monge_ampere(f, d, eps=eps, tol=1e-10, maxiter=100, variant=ssg.MA_BODY)
    out["ma"] = {"converged": diag["converged"], "iterations": diag["iterations"],
                 "iterations_fp32": diag["iterations_fp32"], "relative_residual": diag["relative_residual"],
                 "ms": diag["solve_time"] * 1e3}
    if not diag["converged"]:
        d.close()
        return out
    st = ssg.TimeState(0.0, f)
    ssg.timestep(f, d, params, st, 10)
    ssg.timestep_prepare(d, params, st, steps - 10)
    d.sync()
    t0 = time.perf_counter()
    ssg.timestep(f, d, params, st, steps - 10)
    d.sync()
    out["frozen_fp32_steps_per_s"] = (steps - 10) / (time.perf_counter() - t0)
    if adj_steps > 0:
        adj_steps = min(adj_steps, steps)
        ssg.set_coupling(d, eps=eps, tol=1e-10, maxiter=100, variant=ssg.MA_BODY, warm=2, stage_fp32=True)
        f.bs = b0
        st = ssg.TimeState(0.0, f)
        t0 = time.perf_counter()
        ssg.timestep(f, d, params, st, steps - adj_steps)
        d.sync()
        t_lead = time.perf_counter() - t0
        b_start = f.bs  # the checkpoint the adjoint starts from
        t0 = time.perf_counter()
        ssg.timestep(f, d, params, st, adj_steps)
        d.sync()
        t_tail = time.perf_counter() - t0
        out["coupled_forward_steps"] = steps
        out["coupled_fp32_steps_per_s"] = steps / (t_lead + t_tail)
        out["coupled_fp32_steps_per_s_last"] = adj_steps / t_tail
        bT = f.bs
        ssg.set_coupling(d, eps=None)
        f.bs = b_start
        lam = bT * (2 * np.pi / n) ** 2
        t0 = time.perf_counter()
        g = ssg.timestep_adjoint_coupled(d, params, steps - adj_steps, adj_steps, lam, eps=eps, tol=1e-10, maxiter=100,
                                         variant=ssg.MA_BODY, warm=2, checkpoint_every=10)
        out["coupled_adjoint_s"] = time.perf_counter() - t0
        out["coupled_adjoint_steps_per_s"] = adj_steps / out["coupled_adjoint_s"]
        out["grad_finite"] = bool(np.isfinite(g).all())
    d.close()
    return out


def measure_ensemble(ssg, steps, warmup):
    """Batched configuration: 64 members (32 x 128², 24 x 256², 8 x 512²),
    ICs from seeds 0-63, per-member ε ~ U[0.05, 0.2] (SplitMix64, stream 1)
    used in each member's body-forced Monge-Ampère solve for its frozen φₛ,
    RK3 dt = 1e-3·128/Nx, FP32-mixed, filter every 10 steps."""
    from ssgamd.ics import spectral_b0, uniform_keyed
    sizes = [128] * 32 + [256] * 24 + [512] * 8
    nzs = [16] * 32 + [16] * 24 + [32] * 8
    ens = ssg.Ensemble([(n, n, z) for n, z in zip(sizes, nzs)], precision=ssg.SSG_FP32_MIXED)
    eps = []
    t0 = time.perf_counter()
    iters = []
    for i, n in enumerate(sizes):
        e_i = 0.05 + 0.15 * float(uniform_keyed(i, np.array(0), np.array(0), stream=1))
        eps.append(e_i)
        m = ens.member(i)
        f = ssg.allocate_fields(m)
        f.bs = spectral_b0(n, n, seed=i)
        diag = ssg.solve_monge_ampere(f, m, eps=e_i, tol=1e-10, maxiter=60, variant=ssg.MA_BODY)
        iters.append(diag["iterations"])
    setup_s = time.perf_counter() - t0
    dts = [1e-3 * 128 / n for n in sizes]
    params = ssg.TimeParams(1.0, scheme=ssg.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    ens.timestep(params, dts, warmup)
    t0 = time.perf_counter()
    ens.timestep(params, dts, steps)
    wall = time.perf_counter() - t0
    tune = ens.tuning()
    ens.close()
    points = sum(n * n for n in sizes)
    return {"members": len(sizes), "steps": steps, "ensemble_steps_per_s": steps / wall,
            "member_steps_per_s": len(sizes) * steps / wall, "grid_points": points,
            "Mpoint_steps_per_s": points * steps / wall / 1e6, "autotune": tune,
            "setup_ma_solves_s": setup_s, "setup_ma_iterations_mean": float(np.mean(iters)),
            "eps_range": [min(eps), max(eps)]}


def fft_compare(ssg, n, reps):
    """Spectral round trip at n x n: this library's filter application
    (k_rows_fwd + k_cols + k_rows_inv, HIP-event kernel times summed) against
    rocFFT R2C + scale + C2R (HIP events around back-to-back round trips)."""
    import ctypes
    out = {"n": n, "reps": reps}
    lib = os.path.join(ROOT, "ssg.jl_amd", "lib", "libssg_rocfft_cmp.so")
    cmp = ctypes.CDLL(lib)
    cmp.ssg_cmp_rocfft.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
    for name, prec in (("fp64", ssg.SSG_FP64), ("fp32", ssg.SSG_FP32_MIXED)):
        d = ssg.make_domain(n, n, 1, precision=prec)
        f = ssg.allocate_fields(d)
        f.bs = np.random.default_rng(0).standard_normal((n, n))
        for _ in range(5):
            ssg.apply_spectral_filter(f, d, 36.0, 8, 0.65)
        d.profile(True)
        for _ in range(reps):
            ssg.apply_spectral_filter(f, d, 36.0, 8, 0.65)
        ours = sum(d.profile_read(k)[0] for k in KIND_NAMES) * 1e3 / reps
        d.profile(False)
        d.close()
        us = (ctypes.c_double * 3)()
        rc = cmp.ssg_cmp_rocfft(n, n, 1 if name == "fp32" else 0, reps, us)
        out[name] = {"ours_roundtrip_us": ours, "rocfft_roundtrip_us": us[0] if rc == 0 else None,
                     "rocfft_fwd_us": us[1] if rc == 0 else None, "rocfft_inv_us": us[2] if rc == 0 else None}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed steps for about this long before the warm-up (clocks to steady state); "
                         "the run then restarts from b0 at step 0")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--dt", type=float, default=2.5e-4)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None)
    ap.add_argument("--coupled-ab2-steps", type=int, default=1000,
                    help="coupled AB2 steps per form (the precision configuration); 0 = skip")
    ap.add_argument("--coupled-steps", type=int, default=1000,
                    help="also time this many coupled steps (Monge-Ampère solve per stage); 0 = skip")
    ap.add_argument("--ma-warm", type=int, default=2, choices=(0, 1, 2),
                    help="warm start of the coupled step's solves: 0 cold, 1 last D(Φ), 2 extrapolated")
    ap.add_argument("--ab2-steps", type=int, default=1000,
                    help="also time this many AB2 steps per precision; 0 = skip")
    ap.add_argument("--adjoint-steps", type=int, default=100,
                    help="also time one gradient through this many steps of the adjoint configuration; 0 = skip")
    ap.add_argument("--stiff-steps", type=int, default=2000,
                    help="also run the 2048²×64 stiff configuration for this many FP32-mixed steps, with φₛ "
                         "frozen and coupled; 0 = skip")
    ap.add_argument("--stiff-coupled-steps", type=int, default=50,
                    help="the last this many coupled steps of the stiff configuration get a checkpointed adjoint; "
                         "0 = no coupled run")
    ap.add_argument("--c1-steps", type=int, default=200,
                    help="also run the reference's own CPU case (128²×16, set_b! then RK3 dt = 1e-3) for this "
                         "many steps on the GPU and, unless --no-cpu, on the CPU oracle; 0 = skip")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="also time the 64-member batched configuration for this many steps; 0 = skip")
    ap.add_argument("--fft-compare", action="store_true",
                    help="only time the spectral round trip (rfft2 -> diagonal -> irfft2) of this library "
                         "against rocFFT through hipFFT, at --nx")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB three times with a plain device copy kernel "
                         "(reference point for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import ssgamd as ssg
    from ssgamd.ics import spectral_b0, streamfunction_from_b

    nx = ny = args.nx
    if args.fft_compare:
        print(json.dumps(fft_compare(ssg, nx, reps=200)), flush=True)
        return
    nz = args.nz
    b0 = spectral_b0(nx, ny, seed=0, k0=6.0)
    psi = streamfunction_from_b(b0)
    params = ssg.TimeParams(args.dt, scheme=ssg.RK3, filter_freq=10, filter_strength=36.0, filter_order=8,
                            filter_cutoff=0.65)
    res = {}
    for name, prec in (("fp64", ssg.SSG_FP64), ("fp32", ssg.SSG_FP32_MIXED)):
        if args.only and name != args.only:
            continue
        log(f"bench: {name} RK3 {nx}x{ny} warmup={args.warmup} steps={args.steps}")
        res[name] = run_gpu(ssg, prec, b0, psi, nx, ny, nz, params, args.warmup, args.steps,
                            profile=(not args.no_profile) and name == "fp32", settle_ms=args.settle_ms)
        log(f"bench: {name} {args.steps / res[name][0]:.1f} steps/s (wall), gpu {res[name][1]:.2f} ms")

    main_name = "fp32" if "fp32" in res else "fp64"
    wall, gpu_ms, launches, b_end, prof, tune, settled = res[main_name]
    value = args.steps / wall
    kernel_ms_per_step = gpu_ms / args.steps   # HIP events around the timed call's launches
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_steps": settled,
        "ms_per_step": 1e3 * wall / args.steps,
        "device_ms_per_step": kernel_ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 state)" if main_name == "fp32" else "f64",
        "data": "synthetic (seeded random-phase b0, frozen psi)",
        "config": {"workload": f"RK3 timestep! {nx}x{ny}x{nz} (precision config grid), Mode R (frozen psi), "
                               f"dt={args.dt}, exp filter a=36 p=8 c=0.65 every 10 steps",
                   "nx": nx, "ny": ny, "nz": nz, "scheme": "RK3", "mode": "R", "form": main_name},
    }
    if "fp64" in res:
        out["fp64_steps_per_s"] = args.steps / res["fp64"][0]
    if "fp64" in res and "fp32" in res:
        out["fp32_over_fp64"] = out["value"] / out["fp64_steps_per_s"]
        bf, b6 = res["fp32"][3], res["fp64"][3]
        out["b_rel_l2_fp32_vs_fp64"] = rel_l2(bf, b6)
        out["b_variance_rel_fp32_vs_fp64"] = float(abs(np.sum(bf * bf) - np.sum(b6 * b6)) / np.sum(b6 * b6))
        out["b_rel_l2_steps"] = args.warmup + args.steps
    if prof:
        out["instrumented_steps_per_s"] = args.steps / prof.pop("wall")
        kinds = {KIND_NAMES[k]: {"ms": v[0], "launches": v[1], "bytes": v[2]} for k, v in prof.items()}
        dom = max(kinds, key=lambda k: kinds[k]["ms"])
        kd = kinds[dom]
        avg_s = kd["ms"] / 1e3 / max(kd["launches"], 1)
        per_launch = kd["bytes"] / max(kd["launches"], 1)
        achieved = per_launch / avg_s / 1e9
        traffic = load_traffic(f"{dom}:{nx}:{main_name}")
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                           "bytes_per_launch": per_launch, "avg_launch_us": avg_s * 1e6}
        tot_ms = sum(v["ms"] for v in kinds.values())
        tot_b = sum(v["bytes"] for v in kinds.values())
        out["kernels"] = {k: {"avg_us": 1e3 * v["ms"] / max(v["launches"], 1), "launches": v["launches"],
                              "GBps": v["bytes"] / max(v["ms"], 1e-12) / 1e6} for k, v in kinds.items()}
        out["step_algorithmic_GBps"] = tot_b / (tot_ms / 1e3) / 1e9 if tot_ms > 0 else None
    out["autotune"] = {k: v[5] for k, v in res.items()}
    if args.ab2_steps > 0 and not args.only:
        log(f"bench: AB2 {nx}x{ny}, {args.ab2_steps} steps per precision")
        out["ab2"] = measure_ab2(ssg, b0, psi, nx, ny, nz, args.dt, args.ab2_steps)
    if args.coupled_steps > 0 and not args.only:
        log(f"bench: coupled RK3 {nx}x{ny}x{nz}, {args.coupled_steps} steps per precision")
        c64, f64 = measure_coupled(ssg, b0, nx, ny, nz, args.dt, args.coupled_steps, ssg.SSG_FP64, args.ma_warm)
        # the precision form: FP32 stage solves, FP64 solve for the returned b
        c32, f32 = measure_coupled(ssg, b0, nx, ny, nz, args.dt, args.coupled_steps, ssg.SSG_FP32_MIXED,
                                   args.ma_warm, stage_fp32=True)
        # FP32-mixed tendency with every stage solve checked on an FP64
        # residual at 1e-10
        c32s, f32s = measure_coupled(ssg, b0, nx, ny, nz, args.dt, args.coupled_steps, ssg.SSG_FP32_MIXED,
                                     args.ma_warm)

        def errs(a, ref, nsteps):
            return {"b_rel_l2": rel_l2(a["b"], ref["b"]), "phis_rel_l2": rel_l2(a["phis"], ref["phis"]),
                    "ke_surface_rel": abs(a["ke"] - ref["ke"]) / abs(ref["ke"]), "steps": nsteps}
        out["coupled_fp64"] = c64
        out["coupled_fp32"] = c32
        out["coupled_fp32_over_fp64"] = c32["steps_per_s"] / c64["steps_per_s"]
        out["coupled_fp32_errors"] = errs(f32, f64, args.coupled_steps + 1)
        out["coupled_fp32_strict"] = c32s
        out["coupled_fp32_strict_over_fp64"] = c32s["steps_per_s"] / c64["steps_per_s"]
        out["coupled_fp32_strict_errors"] = errs(f32s, f64, args.coupled_steps + 1)
        out["roofline_coupled"] = dict(c32["roofline"], form="fp32 stage solves")
        out["roofline_coupled_fp64"] = dict(c64["roofline"], form="fp64")
        # Mode C at a glance: the metric's grid with the Monge-Ampère solve at
        # every RK stage, the precision form against the FP64 GPU step
        out["mode_c"] = {"workload": f"coupled RK3 {nx}x{ny}x{nz} (Monge-Ampere solve per stage, body-forced, "
                                     f"eps=0.1, tol 1e-10), dt={args.dt}",
                         "fp32_steps_per_s": c32["steps_per_s"], "fp64_steps_per_s": c64["steps_per_s"],
                         "fp32_over_fp64": c32["steps_per_s"] / c64["steps_per_s"],
                         "b_rel_l2_fp32_vs_fp64": out["coupled_fp32_errors"]["b_rel_l2"],
                         "steps": args.coupled_steps, "roofline": out["roofline_coupled"]}
        # the precision configuration as specified: coupled AB2, dt = 2.5e-4,
        # 1000 steps, the precision form against the FP64 GPU step
        if args.coupled_ab2_steps > 0:
            log(f"bench: coupled AB2 {nx}x{ny}x{nz}, {args.coupled_ab2_steps} steps per form")
            a64, g64 = measure_coupled(ssg, b0, nx, ny, nz, args.dt, args.coupled_ab2_steps, ssg.SSG_FP64,
                                       args.ma_warm, scheme=ssg.AB2_LowStorage)
            a32, g32 = measure_coupled(ssg, b0, nx, ny, nz, args.dt, args.coupled_ab2_steps, ssg.SSG_FP32_MIXED,
                                       args.ma_warm, stage_fp32=True, scheme=ssg.AB2_LowStorage)
            out["coupled_ab2"] = {"fp64_steps_per_s": a64["steps_per_s"], "fp32_steps_per_s": a32["steps_per_s"],
                                  "fp32_over_fp64": a32["steps_per_s"] / a64["steps_per_s"],
                                  "errors": errs(g32, g64, args.coupled_ab2_steps + 1),
                                  "steps": args.coupled_ab2_steps, "form": "fp32 stage solves, fp64 solve of the returned b"}
    if args.adjoint_steps > 0 and not args.only:
        log(f"bench: adjoint configuration 256x256x16, {args.adjoint_steps} steps")
        out["adjoint"] = measure_adjoint(ssg, args.adjoint_steps)
    if args.stiff_steps > 0 and not args.only:
        log(f"bench: stiff configuration 2048x2048x64, {args.stiff_steps} frozen and {args.stiff_steps} coupled "
            f"steps, the adjoint of the last {args.stiff_coupled_steps}")
        out["stiff"] = measure_stiff(ssg, args.stiff_steps, args.stiff_coupled_steps)
    if args.ensemble_steps > 0 and not args.only:
        log(f"bench: ensemble 64 members, {args.ensemble_steps} steps")
        out["ensemble_fp32"] = measure_ensemble(ssg, args.ensemble_steps, 20)
    c1_inputs = None
    if args.c1_steps > 0 and not args.only:
        log(f"bench: reference CPU case 128x128x16, set_b! and {args.c1_steps} RK3 steps")
        out["c1"], c1_inputs = measure_c1(ssg, args.c1_steps)
    if not args.no_cpu:
        log("bench: CPU baseline (oracle)")
        dt = args.dt
        out["cpu_baseline"] = cpu_baseline(b0, psi, nx, ny, nz, dt, args.cpu_seconds)
        if c1_inputs is not None:
            out["cpu_baseline"]["c1"] = cpu_c1(c1_inputs, args.c1_steps)
    if args.copy_calibration:
        d = ssg.make_domain(64, 64, 1)
        nbytes = args.copy_calibration * (1 << 20)
        d.check(d.lib.ssg_copy_probe(d.handle, nbytes, 3))
        d.close()
        out["copy_calibration_bytes"] = nbytes
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
