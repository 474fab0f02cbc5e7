// C ABI of the SSG time-step library: context, field transfer, the step
// driver and the spectral sub-boundary.  Declarations: include/ssg.h.
#include "ssg_internal.hpp"

namespace ssg_impl {

struct TuneEntry {
    int ept[3];
    double us[3][3];
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, TuneEntry> g_tune_cache;

template <typename R>
int autotune(ssg_ctx* c) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    if (c->tuned[p]) return SSG_OK;
    c->tuned[p] = true;
    const auto key = std::make_tuple(c->nx, c->ny, p, c->device);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end() && !getenv("SSG_EPT")) {
            for (int k = 0; k < 3; ++k) {
                c->ept[p][k] = it->second.ept[k];
                for (int i = 0; i < 3; ++i) c->tune_us[p][k][i] = it->second.us[k][i];
            }
            return SSG_OK;
        }
    }
    if (const char* env = getenv("SSG_EPT")) {
        int e0 = 0, e1 = 0, e2 = 0;
        if (sscanf(env, "%d,%d,%d", &e0, &e1, &e2) == 3) {
            c->ept[p][0] = e0;
            c->ept[p][1] = e1;
            c->ept[p][2] = e2;
        }
        return SSG_OK;
    }
    const char* tv = getenv("SSG_TUNE");
    if ((tv && tv[0] == '0') || (long)c->nx * c->ny < 256L * 256L) return SSG_OK;
    const bool prof_saved = c->prof;
    const int64_t launches_saved = c->launches;
    c->prof = false;
    const int cands[3] = {4, 8, 16};
    int rc = SSG_OK;
    // three interleaved rounds of 8 timed launches per candidate; each
    // candidate keeps its fastest round (drift and one-off stalls fall out)
    for (int kind = 0; kind < 3 && rc == SSG_OK; ++kind) {
        const int n = kind == 1 ? c->ny : c->nx;
        double best_per[3] = {1e30, 1e30, 1e30};
        for (int round = 0; round < 3 && rc == SSG_OK; ++round) {
            for (int ci = 0; ci < 3; ++ci) {
                const int e = cands[ci];
                if (!ept_valid(n, e)) continue;
                c->ept[p][kind] = e;
                if (round == 0)
                    for (int i = 0; i < 3 && rc == SSG_OK; ++i) rc = run_kind<R>(c, kind);
                if (rc) break;
                if (hipEventRecord(c->ev0, c->stream) != hipSuccess) {
                    rc = SSG_ERR_HIP;
                    break;
                }
                for (int i = 0; i < 8 && rc == SSG_OK; ++i) rc = run_kind<R>(c, kind);
                if (rc) break;
                float ms = 0;
                if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
                    hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
                    rc = set_err(c, SSG_ERR_HIP, "autotune timing failed");
                    break;
                }
                best_per[ci] = std::min(best_per[ci], 1e3 * ms / 8.0);
            }
        }
        int best = 0;
        double best_us = 1e30;
        for (int ci = 0; ci < 3; ++ci) {
            c->tune_us[p][kind][ci] = best_per[ci] < 1e29 ? best_per[ci] : -1.0;
            if (best_per[ci] < best_us) {
                best_us = best_per[ci];
                best = cands[ci];
            }
        }
        c->ept[p][kind] = best;
    }
    c->prof = prof_saved;
    c->launches = launches_saved;
    if (rc == SSG_OK) {
        TuneEntry e{};
        for (int k = 0; k < 3; ++k) {
            e.ept[k] = c->ept[p][k];
            for (int i = 0; i < 3; ++i) e.us[k][i] = c->tune_us[p][k][i];
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache.emplace(key, e);
    }
    return rc;
}

int ma_solve_dev(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d);

// Coupled step: the Monge-Ampère problem is re-solved for the stage
// buoyancy before every tendency evaluation, so φₛ and (u, v) follow b.
// Each step ends with the solve for the buoyancy it returns, which is also
// the next step's first-stage solve: after the call φₛ, u and v belong to
// the returned b, and a following step reuses that solve (ma_phi_current).
template <typename R>
int run_steps_coupled(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    const double dt = p->dt;
    ssg_ma_diag d{};
    int rc;
    // solves per step: the extrapolated warm start takes each stage's guess
    // from the same stage of the previous steps
    struct PeriodScope {
        ssg_ctx* c;
        ~PeriodScope() { c->ma_period = 0; }
    } period_scope{c};
    c->ma_period = p->scheme == SSG_RK3 ? 3 : 1;
    auto solve = [&](const double* bsrc, bool final_solve) -> int {
        c->ma_phi_current = false;
        int r = ma_solve_coupled(c, bsrc, &c->ma_par, &d, final_solve);
        if (r) return r;
        if (!d.converged)
            return ma_not_converged(c, d);
        return SSG_OK;
    };
    auto stage_tendency = [&](const double* bsrc, bool solved) -> int {
        int r;
        if (!solved && (r = solve(bsrc, false))) return r;
        if ((r = flux_forward<R>(c, c->ncols_tend, bsrc))) return r;
        return tendency_cols<R>(c, p->scheme == SSG_AB2_LOWSTORAGE ? p->ab2_reference_aliasing : 0);
    };
    for (int64_t it = 0; it < nsteps; ++it) {
        const int64_t step_after = s->step + 1;
        const bool filt = p->filter_freq > 0 && (step_after % p->filter_freq) == 0;
        // the solve for b itself ended the previous step
        const bool solved = c->ma_phi_current;
        if (p->scheme == SSG_RK3) {
            if ((rc = stage_tendency(c->b, solved))) return rc;
            if ((rc = stage_inverse<R, EPI_RK1>(c, dt / 3.0, 0, 0, c->bst))) return rc;
            if ((rc = stage_tendency(c->bst, false))) return rc;
            if ((rc = stage_inverse<R, EPI_RK2>(c, 2.0 * dt / 3.0, 0, 0, c->bst))) return rc;
            if ((rc = stage_tendency(c->bst, false))) return rc;
            if ((rc = stage_inverse<R, EPI_RK3>(c, dt / 4.0, 0, 0))) return rc;
        } else {
            if ((rc = stage_tendency(c->b, solved))) return rc;
            const int first = (s->step == 0) ? 1 : 0;
            if ((rc = stage_inverse<R, EPI_AB2>(c, dt, 0, first))) return rc;
            c->ab2_have_history = true;
        }
        s->t += dt;
        s->step = step_after;
        s->dt_actual = dt;
        if (filt && (rc = filter_b<R>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0))) return rc;
        if ((rc = solve(c->b, it + 1 == nsteps))) return rc;
        c->ma_phi_current = true;
    }
    return SSG_OK;
}

// Kernels of one timestep! (src/timestep.jl:267-285) starting from step
// number `step_before`; `last` suppresses the next-stage flux transform that
// nothing would read.
template <typename R>
int issue_step(ssg_ctx* c, const ssg_time_params* p, int64_t step_before, bool last) {
    int rc;
    const double dt = p->dt;
    const int64_t step_after = step_before + 1;
    const bool filt = p->filter_freq > 0 && (step_after % p->filter_freq) == 0;
    const int emit_final = (!filt && !last) ? 1 : 0;
    if (p->scheme == SSG_RK3) {
        // src/timestep.jl:195-242
        if ((rc = tendency_cols<R>(c, 0))) return rc;
        if ((rc = stage_inverse<R, EPI_RK1>(c, dt / 3.0, 1, 0))) return rc;
        if ((rc = tendency_cols<R>(c, 0))) return rc;
        if ((rc = stage_inverse<R, EPI_RK2>(c, 2.0 * dt / 3.0, 1, 0))) return rc;
        if ((rc = tendency_cols<R>(c, 0))) return rc;
        if ((rc = stage_inverse<R, EPI_RK3>(c, dt / 4.0, emit_final, 0))) return rc;
    } else {
        // src/timestep.jl:152-190
        if ((rc = tendency_cols<R>(c, p->ab2_reference_aliasing))) return rc;
        const int first = (step_before == 0) ? 1 : 0;
        if ((rc = stage_inverse<R, EPI_AB2>(c, dt, emit_final, first))) return rc;
    }
    if (filt) {
        // src/timestep.jl:280-281
        if ((rc = filter_b<R>(c, p->filter_strength, p->filter_order, p->filter_cutoff, last ? 0 : 1))) return rc;
    }
    return SSG_OK;
}


// Step runs are replayed from hipGraphs.  A graph holds `len` consecutive
// steps that start at phase `phase` of the filter cycle (phase = step % F,
// F = filter_freq, or 16 when the filter is off) and stay within one cycle,
// so every run of steps is a chain of at most one partial cycle, whole
// cycles and one partial cycle, whatever its length and starting step.
// Graphs are captured on first use (or by ssg_timestep_prepare) and kept per
// (phase, len); kernel arguments are fixed by the parameters, so all of them
// are dropped when the parameters change.  The AB2 step from step 0 (forward
// Euler) runs eagerly.
inline int cycle_len(const ssg_time_params* p) { return p->filter_freq > 0 ? p->filter_freq : 16; }

template <typename R>
int step_graph(ssg_ctx* c, const ssg_time_params* p, int phase, int len, int64_t step0, hipGraphExec_t* out) {
    const int path = std::is_same<R, float>::value ? 1 : 0;
    auto& gc = c->gcache[path];
    if (!gc.have_p || !same_params(gc.p, *p)) {
        for (auto& kv : gc.exec) (void)hipGraphExecDestroy(kv.second);
        gc.exec.clear();
        gc.p = *p;
        gc.have_p = true;
    }
    const auto key = std::make_pair(phase, len);
    auto it = gc.exec.find(key);
    if (it != gc.exec.end()) {
        *out = it->second;
        return SSG_OK;
    }
    HIP_OK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = SSG_OK;
    const int64_t l0 = c->launches;
    for (int i = 0; i < len && rc == SSG_OK; ++i) rc = issue_step<R>(c, p, step0 + i, false);
    c->launches = l0;  // captured, not launched
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_err(c, SSG_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_err(c, SSG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    // upload now, so the first replay costs what every later one does
    e = hipGraphUpload(ex, c->stream);
    if (e != hipSuccess) {
        (void)hipGraphExecDestroy(ex);
        return set_err(c, SSG_ERR_HIP, "hipGraphUpload: %s", hipGetErrorString(e));
    }
    gc.exec.emplace(key, ex);
    *out = ex;
    return SSG_OK;
}

inline bool graphs_enabled(const ssg_ctx* c) {
    const char* genv = getenv("SSG_GRAPH");
    return !(genv && genv[0] == '0') && !c->prof;
}

// Walk the pieces of a run of nsteps steps from step `step`: fn(step, n,
// graph) for a piece of n steps replayed from a graph, or n = 1 and graph =
// false for an eager step.
template <typename Fn>
int for_each_piece(const ssg_ctx* c, const ssg_time_params* p, int64_t step, int64_t nsteps, Fn&& fn) {
    const bool graphs = graphs_enabled(c);
    const int F = cycle_len(p);
    int64_t it = 0;
    while (it < nsteps) {
        const bool ab2_first = p->scheme == SSG_AB2_LOWSTORAGE && step == 0;
        int64_t n = 1;
        bool g = false;
        if (graphs && !ab2_first) {
            const int phase = (int)(step % F);
            n = std::min<int64_t>(F - phase, nsteps - it);
            g = true;
        }
        int rc = fn(step, n, g, it + n == nsteps);
        if (rc) return rc;
        step += n;
        it += n;
    }
    return SSG_OK;
}

template <typename R>
int run_steps(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    if (c->coupled) return run_steps_coupled<R>(c, p, s, nsteps);
    c->ma_phi_current = false;
    int rc = flux_forward<R>(c, c->ncols_tend);
    if (rc) return rc;
    const int F = cycle_len(p);
    return for_each_piece(c, p, s->step, nsteps, [&](int64_t step, int64_t n, bool graph, bool last) -> int {
        int r;
        if (graph) {
            hipGraphExec_t ex = nullptr;
            if ((r = step_graph<R>(c, p, (int)(step % F), (int)n, step, &ex))) return r;
            HIP_OK(c, hipGraphLaunch(ex, c->stream));
            for (int64_t k = step; k < step + n; ++k)
                c->launches += (p->scheme == SSG_RK3 ? 6 : 2) +
                               ((p->filter_freq > 0 && (k + 1) % p->filter_freq == 0) ? 3 : 0);
        } else if ((r = issue_step<R>(c, p, step, last))) {
            return r;
        }
        for (int64_t k = 0; k < n; ++k) s->t += p->dt;  // the same sum as one step at a time
        s->step += n;
        s->dt_actual = p->dt;
        if (p->scheme == SSG_AB2_LOWSTORAGE) c->ab2_have_history = true;
        return SSG_OK;
    });
}

// Instantiate the graphs that run_steps would replay for nsteps from `step`.
template <typename R>
int prepare_steps(ssg_ctx* c, const ssg_time_params* p, int64_t step, int64_t nsteps) {
    if (c->coupled) return SSG_OK;
    const int F = cycle_len(p);
    return for_each_piece(c, p, step, nsteps, [&](int64_t st, int64_t n, bool graph, bool) -> int {
        if (!graph) return SSG_OK;
        hipGraphExec_t ex = nullptr;
        return step_graph<R>(c, p, (int)(st % F), (int)n, st, &ex);
    });
}


int check_domain(const ssg_domain_desc* d, char* msg, size_t len) {
    if (!d) return snprintf(msg, len, "null domain descriptor"), SSG_ERR_ARG;
    if (!pow2_in_range(d->nx) || !pow2_in_range(d->ny))
        return snprintf(msg, len, "nx, ny must be powers of two in [8, 4096] (got %d, %d)", d->nx, d->ny),
               SSG_ERR_UNSUPPORTED;
    if (d->nz < 1) return snprintf(msg, len, "nz must be >= 1"), SSG_ERR_ARG;
    if (!(d->lx > 0) || !(d->ly > 0) || !(d->lz > 0)) return snprintf(msg, len, "domain lengths must be > 0"), SSG_ERR_ARG;
    if (d->precision != SSG_FP64 && d->precision != SSG_FP32_MIXED)
        return snprintf(msg, len, "unknown precision %d", d->precision), SSG_ERR_ARG;
    return SSG_OK;
}

template <typename R>
int alloc_path(ssg_ctx* c, PathBuffers<R>& B) {
    const size_t n = (size_t)c->nx * c->ny;
    const size_t planes = (size_t)c->ncols_full * c->g.pitch;
    HIP_OK(c, hipMalloc(&B.w, n * sizeof(cpx<R>)));
    HIP_OK(c, hipMalloc(&B.k1, n * sizeof(R)));
    HIP_OK(c, hipMalloc(&B.fold, n * sizeof(R)));
    HIP_OK(c, hipMalloc(&B.tmp, n * sizeof(R)));
    HIP_OK(c, hipMalloc(&B.Fx, planes * sizeof(cpx<R>)));
    HIP_OK(c, hipMalloc(&B.Fy, planes * sizeof(cpx<R>)));
    HIP_OK(c, hipMalloc(&B.Tp, planes * sizeof(cpx<R>)));
    HIP_OK(c, hipMemsetAsync(B.w, 0, n * sizeof(cpx<R>), c->stream));
    HIP_OK(c, hipMemsetAsync(B.k1, 0, n * sizeof(R), c->stream));
    HIP_OK(c, hipMemsetAsync(B.fold, 0, n * sizeof(R), c->stream));
    HIP_OK(c, hipMemsetAsync(B.Fx, 0, planes * sizeof(cpx<R>), c->stream));
    HIP_OK(c, hipMemsetAsync(B.Fy, 0, planes * sizeof(cpx<R>), c->stream));
    HIP_OK(c, hipMemsetAsync(B.Tp, 0, planes * sizeof(cpx<R>), c->stream));
    // twiddles exp(-2 pi i m / n), computed in FP64
    for (int axis = 0; axis < 2; ++axis) {
        const int n1 = axis == 0 ? c->nx : c->ny;
        std::vector<cpx<R>> h(n1);
        for (int m = 0; m < n1; ++m) {
            const double a = -2.0 * M_PI * (double)m / (double)n1;
            h[m] = cpx<R>{(R)std::cos(a), (R)std::sin(a)};
        }
        cpx<R>** dst = axis == 0 ? &B.twx : &B.twy;
        HIP_OK(c, hipMalloc(dst, n1 * sizeof(cpx<R>)));
        HIP_OK(c, hipMemcpy(*dst, h.data(), n1 * sizeof(cpx<R>), hipMemcpyHostToDevice));
    }
    return SSG_OK;
}

template <typename R>
void free_path(PathBuffers<R>& B) {
    (void)hipFree(B.w);
    (void)hipFree(B.k1);
    (void)hipFree(B.fold);
    (void)hipFree(B.tmp);
    (void)hipFree(B.Fx);
    (void)hipFree(B.Fy);
    (void)hipFree(B.Tp);
    (void)hipFree(B.twx);
    (void)hipFree(B.twy);
    B = PathBuffers<R>{};
}

int finite_check(ssg_ctx* c, int64_t* nonfinite, double* sum, double* sum2) {
    const long n = (long)c->nx * c->ny;
    const int nb = c->npartial / 3;
    hipLaunchKernelGGL(k_stats, dim3(nb), dim3(256), 0, c->stream, c->b, n, c->partial);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(c->partial_host, c->partial, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double s = 0, s2 = 0, nf = 0;
    for (int i = 0; i < nb; ++i) {
        s += c->partial_host[3 * i];
        s2 += c->partial_host[3 * i + 1];
        nf += c->partial_host[3 * i + 2];
    }
    if (nonfinite) *nonfinite = (int64_t)nf;
    if (sum) *sum = s;
    if (sum2) *sum2 = s2;
    return SSG_OK;
}

// rfft2 of a device FP64 field into dst (default c->d.Tp), this library's
// layout, FP64
int rfft2_device(ssg_ctx* c, const double* src, cpx<double>* dst) {
    PathBuffers<double>& D = c->d;
    if (!dst) dst = D.Tp;
    int rc = launch_rows_fwd<double, double, SRC_PAIR>(c, src, nullptr, dst, nullptr, c->ncols_full);
    if (rc) return rc;
    return launch_cols<double, OP_FWD>(c, dst, nullptr, dst, nullptr, c->ncols_full, 1.0);
}

double parseval_of_tp(ssg_ctx* c, int* rc_out) {
    const int nb = c->npartial;
    hipLaunchKernelGGL((k_parseval_ref<double>), dim3(nb), dim3(256), 0, c->stream, c->d.Tp, c->g, c->partial);
    c->launches++;
    *rc_out = SSG_OK;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(c->partial_host, c->partial, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        *rc_out = set_err(c, SSG_ERR_HIP, "parseval reduction failed");
        return 0.0;
    }
    double s = 0;
    for (int i = 0; i < nb; ++i) s += c->partial_host[i];
    // src/transforms.jl:271: Lx*Ly*Lz / (Nx^2 Ny^2 Nz)
    return s * (c->lx * c->ly * c->lz) / ((double)c->nx * c->nx * (double)c->ny * c->ny * (double)c->nz);
}

}  // namespace ssg_impl

extern "C" {


int ssg_create(ssg_ctx** out, const ssg_domain_desc* desc) {
    if (!out) return SSG_ERR_ARG;
    *out = nullptr;
    char msg[256];
    int rc = check_domain(desc, msg, sizeof(msg));
    if (rc) {
        fprintf(stderr, "ssg_create: %s\n", msg);
        return rc;
    }
    ssg_ctx* c = new (std::nothrow) ssg_ctx();
    if (!c) return SSG_ERR_ARG;
    c->nx = desc->nx;
    c->ny = desc->ny;
    c->nz = desc->nz;
    c->lx = desc->lx;
    c->ly = desc->ly;
    c->lz = desc->lz;
    c->precision = desc->precision;
    c->device = desc->device;
    c->g.nx = c->nx;
    c->g.ny = c->ny;
    c->g.pitch = c->ny;
    c->g.K = c->nx / 3;
    c->g.Ky = 2 * (c->ny / 2 + 1) / 3;
    c->g.dkx = 2.0 * M_PI / c->lx;
    c->g.dky = 2.0 * M_PI / c->ly;
    c->ncols_full = c->nx / 2 + 1;
    c->ncols_tend = c->g.K + 1;
    auto fail = [&](int code) {
        ssg_destroy(c);
        return code;
    };
    if (hipSetDevice(c->device) != hipSuccess) {
        fprintf(stderr, "ssg_create: hipSetDevice(%d) failed\n", c->device);
        return fail(SSG_ERR_HIP);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(SSG_ERR_HIP);
    const size_t n = (size_t)c->nx * c->ny;
    const size_t nspec = (size_t)c->nx * (c->ny / 2 + 1);
    if (hipMalloc(&c->b, n * sizeof(double)) != hipSuccess || hipMalloc(&c->phi, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->dscratch, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->spec, nspec * sizeof(double2)) != hipSuccess ||
        hipMalloc(&c->partial, c->npartial * 3 * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->partial_host, c->npartial * 3 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(SSG_ERR_HIP);
    (void)hipMemsetAsync(c->b, 0, n * sizeof(double), c->stream);
    (void)hipMemsetAsync(c->phi, 0, n * sizeof(double), c->stream);
    if ((rc = alloc_path<double>(c, c->d))) {
        fprintf(stderr, "ssg_create: %s\n", c->err);
        return fail(rc);
    }
    if (c->precision == SSG_FP32_MIXED && (rc = alloc_path<float>(c, c->f))) {
        fprintf(stderr, "ssg_create: %s\n", c->err);
        return fail(rc);
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(SSG_ERR_HIP);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(SSG_ERR_HIP);
    *out = c;
    return SSG_OK;
}

int ssg_destroy(ssg_ctx* c) {
    if (!c) return SSG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_path(c->d);
    free_path(c->f);
    ma_free(c);
    (void)hipFree(c->adj_lam);
    (void)hipFree(c->adj_acc);
    (void)hipFree(c->adj_s);
    adjc_free(c);
    (void)hipFree(c->b);
    (void)hipFree(c->phi);
    (void)hipFree(c->dscratch);
    (void)hipFree(c->spec);
    (void)hipFree(c->partial);
    if (c->partial_host) (void)hipHostFree(c->partial_host);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto& g : c->gcache)
        for (auto& kv : g.exec) (void)hipGraphExecDestroy(kv.second);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SSG_OK;
}

// With a null context: the message of this thread's last failed
// context-free call (ssg_tendency_noalloc), or "" if it succeeded.
const char* ssg_last_error(const ssg_ctx* c) { return c ? c->err : noalloc_last_error(); }

int ssg_sync(ssg_ctx* c) {
    if (!c) return SSG_ERR_ARG;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

static int field_io(ssg_ctx* c, int32_t field, const double* src, double* dst, int64_t n, bool device) {
    if (!c) return SSG_ERR_ARG;
    const int64_t need = (int64_t)c->nx * c->ny;
    if (n != need) return set_err(c, SSG_ERR_ARG, "field size %lld != nx*ny = %lld", (long long)n, (long long)need);
    HIP_OK(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)need * sizeof(double);
    if (src) {
        if (!src) return set_err(c, SSG_ERR_ARG, "null pointer");
        const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (field == SSG_FIELD_B) {
            HIP_OK(c, hipMemcpyAsync(c->b, src, bytes, kind, c->stream));
            c->ab2_have_history = false;
        } else if (field == SSG_FIELD_PHI) {
            c->ma_phi_current = false;
            c->ma_phis_from_phi = false;
            HIP_OK(c, hipMemcpyAsync(c->phi, src, bytes, kind, c->stream));
            int rc = compute_velocities(c);
            if (rc) return rc;
        } else {
            return set_err(c, SSG_ERR_ARG, "field %d cannot be set", field);
        }
    } else {
        if (!dst) return set_err(c, SSG_ERR_ARG, "null pointer");
        const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (field == SSG_FIELD_B) {
            HIP_OK(c, hipMemcpyAsync(dst, c->b, bytes, kind, c->stream));
        } else if (field == SSG_FIELD_PHI) {
            HIP_OK(c, hipMemcpyAsync(dst, c->phi, bytes, kind, c->stream));
        } else if (field == SSG_FIELD_U || field == SSG_FIELD_V) {
            const long nn = (long)need;
            double* u = field == SSG_FIELD_U ? c->dscratch : c->d.tmp;
            double* v = field == SSG_FIELD_U ? c->d.tmp : c->dscratch;
            hipLaunchKernelGGL((k_w_split<double>), dim3(grid1d(nn)), dim3(256), 0, c->stream, c->d.w, u, v, nn);
            HIP_OK(c, hipGetLastError());
            HIP_OK(c, hipMemcpyAsync(dst, c->dscratch, bytes, kind, c->stream));
        } else {
            return set_err(c, SSG_ERR_ARG, "unknown field %d", field);
        }
    }
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_set_field(ssg_ctx* c, int32_t field, const double* host, int64_t n) {
    if (!host) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (c && field == SSG_FIELD_B) ma_history_reset(c);  // b no longer continues the solves' sequence
    return field_io(c, field, host, nullptr, n, false);
}
int ssg_get_field(ssg_ctx* c, int32_t field, double* host, int64_t n) {
    if (!host) return set_err(c, SSG_ERR_ARG, "null pointer");
    return field_io(c, field, nullptr, host, n, false);
}
int ssg_set_field_device(ssg_ctx* c, int32_t field, const double* dev, int64_t n) {
    if (!dev) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (c && field == SSG_FIELD_B) ma_history_reset(c);
    return field_io(c, field, dev, nullptr, n, true);
}
int ssg_get_field_device(ssg_ctx* c, int32_t field, double* dev, int64_t n) {
    if (!dev) return set_err(c, SSG_ERR_ARG, "null pointer");
    return field_io(c, field, nullptr, dev, n, true);
}

static int check_step(ssg_ctx* c, const ssg_time_params* p, const ssg_time_state* s, int64_t nsteps) {
    if (!p || !s) return set_err(c, SSG_ERR_ARG, "null params/state");
    if (nsteps < 0) return set_err(c, SSG_ERR_ARG, "nsteps < 0");
    if (p->scheme == SSG_RK3_LOWSTORAGE)
        return set_err(c, SSG_ERR_UNSUPPORTED, "Time integration scheme RK3_LowStorage not implemented yet");
    if (p->scheme != SSG_RK3 && p->scheme != SSG_AB2_LOWSTORAGE)
        return set_err(c, SSG_ERR_ARG, "unknown scheme %d", p->scheme);
    if (!(p->dt > 0) || !std::isfinite(p->dt)) return set_err(c, SSG_ERR_ARG, "dt must be finite and > 0");
    if (p->filter_freq > 0) {
        // ExponentialFilter constructor checks, src/filter.jl:51-53
        if (!(p->filter_strength >= 0)) return set_err(c, SSG_ERR_ARG, "Filter strength must be non-negative");
        if (p->filter_order < 1) return set_err(c, SSG_ERR_ARG, "Filter order must be positive");
        if (!(p->filter_cutoff > 0 && p->filter_cutoff <= 1))
            return set_err(c, SSG_ERR_ARG, "Cutoff must be between 0 and 1");
    }
    if (p->scheme == SSG_AB2_LOWSTORAGE && s->step > 0 && !c->ab2_have_history)
        return set_err(c, SSG_ERR_ARG, "AB2 step %lld needs the previous tendency; none is held", (long long)s->step);
    HIP_OK(c, hipSetDevice(c->device));
    return SSG_OK;
}

// autotune and instantiate the step graphs of nsteps from s->step
static int prepare(ssg_ctx* c, const ssg_time_params* p, const ssg_time_state* s, int64_t nsteps) {
    if (c->precision == SSG_FP32_MIXED) {
        int rc = autotune<float>(c);
        return rc ? rc : prepare_steps<float>(c, p, s->step, nsteps);
    }
    int rc = autotune<double>(c);
    return rc ? rc : prepare_steps<double>(c, p, s->step, nsteps);
}

int ssg_timestep_prepare(ssg_ctx* c, const ssg_time_params* p, const ssg_time_state* s, int64_t nsteps) {
    if (!c) return SSG_ERR_ARG;
    int rc = check_step(c, p, s, nsteps);
    if (rc || nsteps == 0) return rc;
    return prepare(c, p, s, nsteps);
}

int ssg_timestep(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    if (!c) return SSG_ERR_ARG;
    {
        int crc = check_step(c, p, s, nsteps);
        if (crc || nsteps == 0) return crc;
        // graphs are captured before the timed region, so last_timing holds
        // kernel time only
        if ((crc = prepare(c, p, s, nsteps))) return crc;
    }
    const int64_t l0 = c->launches;
    HIP_OK(c, hipEventRecord(c->ev0, c->stream));
    int rc = c->precision == SSG_FP32_MIXED ? run_steps<float>(c, p, s, nsteps) : run_steps<double>(c, p, s, nsteps);
    if (rc) return rc;
    HIP_OK(c, hipEventRecord(c->ev1, c->stream));
    // wait by polling the end event: a blocking stream synchronisation adds
    // its wake-up latency to every call, which shows on short step runs
    for (;;) {
        const hipError_t q = hipEventQuery(c->ev1);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) return set_err(c, SSG_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
        __builtin_ia32_pause();
    }
    float ms = 0;
    HIP_OK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = c->launches - l0;
    if (p->check_finite) {
        int64_t nf = 0;
        if ((rc = finite_check(c, &nf, nullptr, nullptr))) return rc;
        if (nf) return set_err(c, SSG_ERR_NONFINITE, "%lld non-finite buoyancy values after step %lld", (long long)nf,
                               (long long)s->step);
    }
    return SSG_OK;
}

int ssg_cfl(ssg_ctx* c, double dt, double* cfl, double* umax, double* vmax) {
    if (!c) return SSG_ERR_ARG;
    HIP_OK(c, hipSetDevice(c->device));
    const long n = (long)c->nx * c->ny;
    const int nb = std::min(c->npartial * 3 / 2, 1024);
    hipLaunchKernelGGL(k_wmax, dim3(nb), dim3(256), 0, c->stream, c->d.w, n, c->partial);
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(c->partial_host, c->partial, (size_t)nb * 2 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double mu = 0, mv = 0;
    for (int i = 0; i < nb; ++i) {
        mu = std::max(mu, c->partial_host[2 * i]);
        mv = std::max(mv, c->partial_host[2 * i + 1]);
    }
    const double dx = c->lx / c->nx, dy = c->ly / c->ny;
    if (umax) *umax = mu;
    if (vmax) *vmax = mv;
    if (cfl) *cfl = std::max(mu * dt / dx, mv * dt / dy);
    return SSG_OK;
}

// Step-state blob: header, b, φₛ (FP64), the AB2 history in the path's
// precision, then the Monge-Ampère part (the last solve's Φ̂, the coupled
// run's warm-start state and coupling parameters).  On load the velocities
// are recomputed from φₛ, as set_field does, or from Φ̂ when φₛ came from
// it, so a restart reproduces the uninterrupted run bitwise, frozen or
// coupled.
struct StateHeader {
    char magic[8];
    int32_t nx, ny, nz, precision;
    int32_t have_ab2, reserved;
    double t;
    int64_t step;
    double dt_actual;
};

static size_t state_bytes_base(const ssg_ctx* c) {
    const size_t n2 = (size_t)c->nx * c->ny;
    return sizeof(StateHeader) + 2 * n2 * sizeof(double) +
           n2 * (c->precision == SSG_FP32_MIXED ? sizeof(float) : sizeof(double));
}

static size_t state_bytes(const ssg_ctx* c) { return state_bytes_base(c) + ma_state_bytes(c); }

int ssg_state_size(ssg_ctx* c, size_t* bytes) {
    if (!c || !bytes) return SSG_ERR_ARG;
    *bytes = state_bytes(c);
    return SSG_OK;
}

int ssg_save_state(ssg_ctx* c, const ssg_time_state* s, void* buf, size_t bytes) {
    if (!c) return SSG_ERR_ARG;
    if (!s || !buf) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (bytes < state_bytes(c)) return set_err(c, SSG_ERR_ARG, "state buffer too small (%zu < %zu)", bytes,
                                               state_bytes(c));
    HIP_OK(c, hipSetDevice(c->device));
    StateHeader h{};
    memcpy(h.magic, "SSGSTAT3", 8);
    h.nx = c->nx;
    h.ny = c->ny;
    h.nz = c->nz;
    h.precision = c->precision;
    h.have_ab2 = c->ab2_have_history ? 1 : 0;
    h.t = s->t;
    h.step = s->step;
    h.dt_actual = s->dt_actual;
    char* p = static_cast<char*>(buf);
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    const size_t n2 = (size_t)c->nx * c->ny;
    HIP_OK(c, hipMemcpyAsync(p, c->b, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    p += n2 * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(p, c->phi, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    p += n2 * sizeof(double);
    if (c->precision == SSG_FP32_MIXED) {
        HIP_OK(c, hipMemcpyAsync(p, c->f.fold, n2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        p += n2 * sizeof(float);
    } else {
        HIP_OK(c, hipMemcpyAsync(p, c->d.fold, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        p += n2 * sizeof(double);
    }
    int rc = ma_state_save(c, p);
    if (rc) return rc;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_load_state(ssg_ctx* c, const void* buf, size_t bytes, ssg_time_state* s) {
    if (!c) return SSG_ERR_ARG;
    if (!s || !buf) return set_err(c, SSG_ERR_ARG, "null pointer");
    StateHeader h{};
    if (bytes < sizeof(h)) return set_err(c, SSG_ERR_ARG, "state buffer too small");
    memcpy(&h, buf, sizeof(h));
    const int version = memcmp(h.magic, "SSGSTAT1", 8) == 0   ? 1
                        : memcmp(h.magic, "SSGSTAT2", 8) == 0 ? 2
                        : memcmp(h.magic, "SSGSTAT3", 8) == 0 ? 3
                                                              : 0;
    if (!version) return set_err(c, SSG_ERR_ARG, "not a saved step state");
    const bool v1 = version == 1;
    if (h.nx != c->nx || h.ny != c->ny || h.precision != c->precision)
        return set_err(c, SSG_ERR_ARG, "state of a %dx%d grid (precision %d) does not fit this %dx%d context (%d)",
                       h.nx, h.ny, h.precision, c->nx, c->ny, c->precision);
    if (bytes < state_bytes_base(c)) return set_err(c, SSG_ERR_ARG, "state buffer truncated");
    HIP_OK(c, hipSetDevice(c->device));
    ma_history_reset(c);
    const size_t n2 = (size_t)c->nx * c->ny;
    const char* p = static_cast<const char*>(buf) + sizeof(h);
    int rc = field_io(c, SSG_FIELD_B, reinterpret_cast<const double*>(p), nullptr, (int64_t)n2, false);
    if (rc) return rc;
    p += n2 * sizeof(double);
    if ((rc = field_io(c, SSG_FIELD_PHI, reinterpret_cast<const double*>(p), nullptr, (int64_t)n2, false))) return rc;
    p += n2 * sizeof(double);
    if (c->precision == SSG_FP32_MIXED) {
        HIP_OK(c, hipMemcpyAsync(c->f.fold, p, n2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        p += n2 * sizeof(float);
    } else {
        HIP_OK(c, hipMemcpyAsync(c->d.fold, p, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        p += n2 * sizeof(double);
    }
    HIP_OK(c, hipStreamSynchronize(c->stream));
    if (!v1) {
        const size_t used = (size_t)(p - static_cast<const char*>(buf));
        if ((rc = ma_state_load(c, p, bytes - used, version))) return rc;
    }
    c->ab2_have_history = h.have_ab2 != 0;
    s->t = h.t;
    s->step = h.step;
    s->dt_actual = h.dt_actual;
    return SSG_OK;
}

int ssg_profile(ssg_ctx* c, int32_t enable) {
    if (!c) return SSG_ERR_ARG;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    c->prof = enable != 0;
    c->prof_recs.clear();
    return SSG_OK;
}

int ssg_profile_read(ssg_ctx* c, int32_t kind, double* total_ms, int64_t* count, double* bytes) {
    if (!c) return SSG_ERR_ARG;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double ms = 0, by = 0;
    int64_t n = 0;
    for (const auto& r : c->prof_recs) {
        if (r.kind != kind) continue;
        float t = 0;
        HIP_OK(c, hipEventElapsedTime(&t, c->ev_pool[r.e0], c->ev_pool[r.e1]));
        ms += t;
        by += r.bytes;
        ++n;
    }
    if (total_ms) *total_ms = ms;
    if (count) *count = n;
    if (bytes) *bytes = by;
    return SSG_OK;
}

int ssg_copy_probe(ssg_ctx* c, int64_t bytes, int32_t reps) {
    if (!c || bytes <= 0 || reps < 1) return SSG_ERR_ARG;
    HIP_OK(c, hipSetDevice(c->device));
    const long n = (long)(bytes / (int64_t)sizeof(double));
    double *a = nullptr, *b = nullptr;
    HIP_OK(c, hipMalloc(&a, n * sizeof(double)));
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return set_err(c, SSG_ERR_HIP, "copy probe allocation failed");
    }
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipError_t e = hipMemsetAsync(a, 0, n * sizeof(double), c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipLaunchKernelGGL((k_convert<double, double>), dim3(blocks), dim3(256), 0, c->stream, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return set_err(c, SSG_ERR_HIP, "copy probe: %s", hipGetErrorString(e));
    return SSG_OK;
}

int ssg_tuning(ssg_ctx* c, int32_t path, int32_t* ept, double* us) {
    if (!c || path < 0 || path > 1) return SSG_ERR_ARG;
    for (int k = 0; k < 3; ++k) {
        const int n = k == 1 ? c->ny : c->nx;
        if (ept) ept[k] = (c->ept[path][k] && ept_valid(n, c->ept[path][k])) ? c->ept[path][k] : ept_default(n);
        for (int i = 0; i < 3; ++i)
            if (us) us[3 * k + i] = c->tuned[path] ? c->tune_us[path][k][i] : -1.0;
    }
    return SSG_OK;
}

int ssg_last_timing(ssg_ctx* c, double* ms, int64_t* launches) {
    if (!c) return SSG_ERR_ARG;
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    return SSG_OK;
}

}  // extern "C"

template <typename R>
static int tendency_impl(ssg_ctx* c, double* out, int quirk, int jac) {
    PathBuffers<R>& B = bufs<R>(c);
    const int nc = jac ? c->ncols_full : c->ncols_tend;
    int rc = flux_forward<R>(c, nc);
    if (rc) return rc;
    const R scale = (R)(1.0 / ((double)c->nx * c->ny));
    if (jac)
        rc = launch_cols<R, OP_JAC>(c, B.Fx, B.Fy, B.Tp, nullptr, nc, scale);
    else
        rc = tendency_cols<R>(c, quirk);
    if (rc) return rc;
    RowsInvArgs<R> a{};
    a.P0 = B.Tp;
    a.ncols_in = nc;
    a.out_r = B.tmp;
    if ((rc = launch_rows_inv<R, MODE_PAIR, EPI_STORE>(c, a))) return rc;
    const long n = (long)c->nx * c->ny;
    const double* src = nullptr;
    if (std::is_same<R, double>::value) {
        src = reinterpret_cast<const double*>(B.tmp);
    } else {
        hipLaunchKernelGGL((k_convert<R, double>), dim3(grid1d(n)), dim3(256), 0, c->stream, B.tmp, c->dscratch, n);
        HIP_OK(c, hipGetLastError());
        src = c->dscratch;
    }
    HIP_OK(c, hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

extern "C" {

int ssg_tendency(ssg_ctx* c, double* out, int32_t quirk) {
    if (!c) return SSG_ERR_ARG;
    if (!out) return set_err(c, SSG_ERR_ARG, "null pointer");
    HIP_OK(c, hipSetDevice(c->device));
    return c->precision == SSG_FP32_MIXED ? tendency_impl<float>(c, out, quirk, 0) : tendency_impl<double>(c, out, quirk, 0);
}

int ssg_jacobian(ssg_ctx* c, double* out) {
    if (!c) return SSG_ERR_ARG;
    if (!out) return set_err(c, SSG_ERR_ARG, "null pointer");
    HIP_OK(c, hipSetDevice(c->device));
    return c->precision == SSG_FP32_MIXED ? tendency_impl<float>(c, out, 0, 1) : tendency_impl<double>(c, out, 0, 1);
}

int ssg_rfft_2d(ssg_ctx* c, const double* real_in, double* spec_out) {
    if (!c) return SSG_ERR_ARG;
    if (!real_in || !spec_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    HIP_OK(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->nx * c->ny, nspec = (size_t)c->nx * (c->ny / 2 + 1);
    HIP_OK(c, hipMemcpyAsync(c->dscratch, real_in, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = rfft2_device(c, c->dscratch);
    if (rc) return rc;
    hipLaunchKernelGGL((k_to_ref<double>), dim3(grid1d((long)nspec)), dim3(256), 0, c->stream, c->d.Tp, c->spec, c->g);
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(spec_out, c->spec, nspec * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_irfft_2d(ssg_ctx* c, const double* spec_in, double* real_out) {
    if (!c) return SSG_ERR_ARG;
    if (!spec_in || !real_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    HIP_OK(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->nx * c->ny, nspec = (size_t)c->nx * (c->ny / 2 + 1);
    PathBuffers<double>& D = c->d;
    HIP_OK(c, hipMemcpyAsync(c->spec, spec_in, nspec * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    const long total = (long)c->ncols_full * c->ny;
    hipLaunchKernelGGL((k_from_ref<double>), dim3(grid1d(total)), dim3(256), 0, c->stream, c->spec, D.Tp, c->g);
    HIP_OK(c, hipGetLastError());
    int rc = launch_cols<double, OP_INV>(c, D.Tp, nullptr, D.Tp, nullptr, c->ncols_full, 1.0 / ((double)c->nx * c->ny));
    if (rc) return rc;
    RowsInvArgs<double> a{};
    a.P0 = D.Tp;
    a.ncols_in = c->ncols_full;
    a.out_r = D.tmp;
    if ((rc = launch_rows_inv<double, MODE_PAIR, EPI_STORE>(c, a))) return rc;
    HIP_OK(c, hipMemcpyAsync(real_out, D.tmp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_apply_exponential_filter(ssg_ctx* c, double strength, int32_t order, double cutoff) {
    if (!c) return SSG_ERR_ARG;
    if (!(strength >= 0)) return set_err(c, SSG_ERR_ARG, "Filter strength must be non-negative");
    if (order < 1) return set_err(c, SSG_ERR_ARG, "Filter order must be positive");
    if (!(cutoff > 0 && cutoff <= 1)) return set_err(c, SSG_ERR_ARG, "Cutoff must be between 0 and 1");
    HIP_OK(c, hipSetDevice(c->device));
    c->ma_phi_current = false;
    int rc = c->precision == SSG_FP32_MIXED ? filter_b<float>(c, strength, order, cutoff, 0)
                                            : filter_b<double>(c, strength, order, cutoff, 0);
    if (rc) return rc;
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_diagnostics(ssg_ctx* c, ssg_diag* out) {
    if (!c) return SSG_ERR_ARG;
    if (!out) return set_err(c, SSG_ERR_ARG, "null pointer");
    HIP_OK(c, hipSetDevice(c->device));
    int rc = finite_check(c, &out->nonfinite, &out->sum_b, &out->sum_b2);
    if (rc) return rc;
    if ((rc = rfft2_device(c, c->b))) return rc;
    out->parseval_b = parseval_of_tp(c, &rc);
    if (rc) return rc;
    // surface KE from u, v (src/setICs.jl:225-246): ½ parsevalsum2(û) + ½ parsevalsum2(v̂)
    const long n = (long)c->nx * c->ny;
    hipLaunchKernelGGL((k_w_split<double>), dim3(grid1d(n)), dim3(256), 0, c->stream, c->d.w, c->dscratch, c->d.tmp, n);
    HIP_OK(c, hipGetLastError());
    if ((rc = rfft2_device(c, c->dscratch))) return rc;
    const double ku = parseval_of_tp(c, &rc);
    if (rc) return rc;
    if ((rc = rfft2_device(c, c->d.tmp))) return rc;
    const double kv = parseval_of_tp(c, &rc);
    if (rc) return rc;
    out->ke_surface = 0.5 * ku + 0.5 * kv;
    // ½ ∫∫ |w|² dx dy in physical space (fixed-order host sum of the block partials)
    const int nb = c->npartial;
    hipLaunchKernelGGL(k_w_energy, dim3(nb), dim3(256), 0, c->stream, (const double2*)c->d.w, n, c->partial);
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(c->partial_host, c->partial, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double e2 = 0;
    for (int i = 0; i < nb; ++i) e2 += c->partial_host[i];
    out->ke_true = 0.5 * e2 * (c->lx * c->ly) / ((double)c->nx * c->ny);
    return SSG_OK;
}

}  // extern "C"
