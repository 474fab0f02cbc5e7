// Batched ensemble: members of different grids advanced with one launch per
// kernel class per RK stage.  Declarations: include/ssg.h.
#include "ssg_internal.hpp"


// ===========================================================================
// Ensemble: members of (possibly) different grids advanced together with one
// launch per kernel class per RK stage (k_cols_batch, k_rows_inv_batch).
// Each member is a full context (set_field, Monge-Ampère solve, diagnostics
// work on it as usual); the batched launches read the members' buffers
// through per-member argument structs and a block -> (member, block) table.
// ===========================================================================
struct ssg_ensemble {
    int n = 0;
    int precision = SSG_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<ssg_ctx*> m;
    int ept = 0;
    bool tuned = false;
    double tune_us[3] = {-1.0, -1.0, -1.0};
    // device argument arrays: [0] cols TEND, rows RK1, RK2, RK3 emit, RK3 no emit
    void* d_cols = nullptr;
    // rows argument sets: [0] RK1, [1] RK2, [2] RK3 emit, [3] RK3 no emit,
    // [4] AB2 emit, [5] AB2 no emit, [6] AB2 forward Euler emit, [7] no emit
    void* d_rows[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int2* d_wcols = nullptr;
    int2* d_wrows = nullptr;
    int ncb = 0, nrb = 0;
    size_t lds_cols = 0, lds_rows = 0;
    // batched filter (apply_spectral_filter! on every member at once): row
    // transform of b, filter column pass, inverse rows with the update
    // ([0] refreshing the next step's fluxes, [1] not)
    void* d_ffwd = nullptr;
    void* d_fcols = nullptr;
    void* d_frows[2] = {nullptr, nullptr};
    int2* d_wffwd = nullptr;
    int2* d_wfcols = nullptr;
    int nffb = 0, nfcb = 0;
    size_t lds_ffwd = 0, lds_fcols = 0;
    double built_filter[3] = {0.0, 0.0, 0.0};
    int built_ept = 0;
    std::vector<double> built_dt;
    hipGraphExec_t graph = nullptr;
    int graph_len = 0;
    ssg_time_params graph_p{};
    std::vector<double> graph_dt;
    char err[512] = {0};
};

namespace {

int ens_err(ssg_ensemble* e, int code, const char* fmt, ...) {
    if (e) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->err, sizeof(e->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define ENS_OK(e, call)                                                                                      \
    do {                                                                                                     \
        hipError_t err_ = (call);                                                                            \
        if (err_ != hipSuccess) return ens_err(e, SSG_ERR_HIP, "%s: %s", #call, hipGetErrorString(err_));    \
    } while (0)

bool ens_size_ok(int n) { return n == 128 || n == 256 || n == 512 || n == 1024; }

void ens_free_tables(ssg_ensemble* e) {
    (void)hipFree(e->d_cols);
    for (auto& p : e->d_rows) {
        (void)hipFree(p);
        p = nullptr;
    }
    (void)hipFree(e->d_wcols);
    (void)hipFree(e->d_wrows);
    (void)hipFree(e->d_ffwd);
    (void)hipFree(e->d_fcols);
    (void)hipFree(e->d_frows[0]);
    (void)hipFree(e->d_frows[1]);
    (void)hipFree(e->d_wffwd);
    (void)hipFree(e->d_wfcols);
    e->d_ffwd = e->d_fcols = e->d_frows[0] = e->d_frows[1] = nullptr;
    e->d_wffwd = e->d_wfcols = nullptr;
    e->d_cols = nullptr;
    e->d_wcols = e->d_wrows = nullptr;
    e->built_ept = 0;
}

// per-member argument structs and the block tables for elements-per-thread E
template <typename R>
int ens_build(ssg_ensemble* e, int E, const double* dt, const ssg_time_params* p) {
    ens_free_tables(e);
    std::vector<ColsArgs<R>> ca(e->n);
    std::vector<RowsFwdArgs<R, double>> fa(e->n);
    std::vector<ColsArgs<R>> fc(e->n);
    std::vector<RowsInvArgs<R>> fr[2];
    for (auto& v : fr) v.resize(e->n);
    std::vector<int2> wf, wfc;
    size_t lf = 0, lfc = 0;
    const int delta = std::is_same<R, float>::value ? 1 : 0;
    std::vector<RowsInvArgs<R>> ra[8];
    for (auto& v : ra) v.resize(e->n);
    std::vector<int2> wc, wr;
    size_t lc = 0, lr = 0;
    for (int i = 0; i < e->n; ++i) {
        ssg_ctx* c = e->m[i];
        PathBuffers<R>& B = bufs<R>(c);
        const int Gc = 256 * E / c->ny, Gr = 256 * E / c->nx;
        if (Gc < 1 || Gr < 1 || 2 * Gr > c->ny)
            return ens_err(e, SSG_ERR_UNSUPPORTED, "member %d (%dx%d) does not fit E = %d", i, c->nx, c->ny, E);
        ColsArgs<R> a{};
        a.inA = B.Fx;
        a.inB = B.Fy;
        a.outA = B.Tp;
        a.ncols = c->ncols_tend;
        a.G = Gc;
        a.scale = (R)(1.0 / ((double)c->nx * c->ny));
        a.twy = B.twy;
        a.g = c->g;
        a.nlev = 1;
        ca[i] = a;
        const int nbc = (c->ncols_tend + Gc - 1) / Gc;
        for (int j = 0; j < nbc; ++j) wc.push_back(int2{i, j});
        lc = std::max(lc, (size_t)Gc * 2 * lds_line_elems(c->ny) * sizeof(cpx<R>));
        const double coef[8] = {dt[i] / 3.0, 2.0 * dt[i] / 3.0, dt[i] / 4.0, dt[i] / 4.0,
                                dt[i],       dt[i],             dt[i],       dt[i]};
        for (int k = 0; k < 8; ++k) {
            RowsInvArgs<R> r{};
            r.P0 = B.Tp;
            r.ncols_in = c->ncols_tend;
            r.b = c->b;
            r.k1 = B.k1;
            r.fold = B.fold;
            r.c1 = coef[k];
            r.emit = (k == 3 || k == 5 || k == 7) ? 0 : 1;
            r.first = k >= 6 ? 1 : 0;
            r.w = B.w;
            r.Fx = B.Fx;
            r.Fy = B.Fy;
            r.ncols_out = c->ncols_tend;
            r.G = Gr;
            r.nblk = c->ny / (2 * Gr);
            r.twx = B.twx;
            r.g = c->g;
            ra[k][i] = r;
        }
        for (int j = 0; j < c->ny / (2 * Gr); ++j) wr.push_back(int2{i, j});
        lr = std::max(lr, (size_t)Gr * 2 * lds_line_elems(c->nx) * sizeof(cpx<R>));
        // filter: the same launch shapes as filter_b
        RowsFwdArgs<R, double> f{c->b, nullptr, B.Tp, nullptr, c->ncols_full, Gr, c->ny / (2 * Gr), B.twx, c->g};
        fa[i] = f;
        for (int j = 0; j < c->ny / (2 * Gr); ++j) wf.push_back(int2{i, j});
        lf = std::max(lf, (size_t)Gr * lds_line_elems(c->nx) * sizeof(cpx<R>));
        ColsArgs<R> q{};
        q.inA = B.Tp;
        q.outA = B.Tp;
        q.ncols = c->ncols_full;
        q.G = Gc;
        q.scale = (R)(1.0 / ((double)c->nx * c->ny));
        q.fstrength = (R)p->filter_strength;
        q.forder = p->filter_order;
        q.fcutoff = (R)p->filter_cutoff;
        q.fdelta = delta;
        const double kmx = M_PI * c->nx / c->lx, kmy = M_PI * c->ny / c->ly;
        q.inv_kmax = (R)(1.0 / std::sqrt(kmx * kmx + kmy * kmy));
        q.twy = B.twy;
        q.g = c->g;
        q.nlev = 1;
        fc[i] = q;
        const int nfc = (c->ncols_full + Gc - 1) / Gc;
        for (int j = 0; j < nfc; ++j) wfc.push_back(int2{i, j});
        lfc = std::max(lfc, (size_t)Gc * lds_line_elems(c->ny) * sizeof(cpx<R>));
        for (int k = 0; k < 2; ++k) {
            RowsInvArgs<R> r{};
            r.P0 = B.Tp;
            r.ncols_in = c->ncols_full;
            r.b = c->b;
            r.emit = k == 0 ? 1 : 0;
            r.w = B.w;
            r.Fx = B.Fx;
            r.Fy = B.Fy;
            r.ncols_out = c->ncols_tend;
            r.G = Gr;
            r.nblk = c->ny / (2 * Gr);
            r.twx = B.twx;
            r.g = c->g;
            fr[k][i] = r;
        }
    }
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        ENS_OK(e, hipMalloc(dst, bytes));
        ENS_OK(e, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return SSG_OK;
    };
    int urc;
    if ((urc = upload(&e->d_ffwd, fa.data(), fa.size() * sizeof(fa[0])))) return urc;
    if ((urc = upload(&e->d_fcols, fc.data(), fc.size() * sizeof(fc[0])))) return urc;
    for (int k = 0; k < 2; ++k)
        if ((urc = upload(&e->d_frows[k], fr[k].data(), fr[k].size() * sizeof(fr[k][0])))) return urc;
    if ((urc = upload((void**)&e->d_wffwd, wf.data(), wf.size() * sizeof(int2)))) return urc;
    if ((urc = upload((void**)&e->d_wfcols, wfc.data(), wfc.size() * sizeof(int2)))) return urc;
    e->nffb = (int)wf.size();
    e->nfcb = (int)wfc.size();
    e->lds_ffwd = lf;
    e->lds_fcols = lfc;
    e->built_filter[0] = p->filter_strength;
    e->built_filter[1] = p->filter_order;
    e->built_filter[2] = p->filter_cutoff;
    ENS_OK(e, hipMalloc(&e->d_cols, ca.size() * sizeof(ColsArgs<R>)));
    ENS_OK(e, hipMemcpy(e->d_cols, ca.data(), ca.size() * sizeof(ColsArgs<R>), hipMemcpyHostToDevice));
    for (int k = 0; k < 8; ++k) {
        ENS_OK(e, hipMalloc(&e->d_rows[k], ra[k].size() * sizeof(RowsInvArgs<R>)));
        ENS_OK(e, hipMemcpy(e->d_rows[k], ra[k].data(), ra[k].size() * sizeof(RowsInvArgs<R>), hipMemcpyHostToDevice));
    }
    ENS_OK(e, hipMalloc(&e->d_wcols, wc.size() * sizeof(int2)));
    ENS_OK(e, hipMemcpy(e->d_wcols, wc.data(), wc.size() * sizeof(int2), hipMemcpyHostToDevice));
    ENS_OK(e, hipMalloc(&e->d_wrows, wr.size() * sizeof(int2)));
    ENS_OK(e, hipMemcpy(e->d_wrows, wr.data(), wr.size() * sizeof(int2), hipMemcpyHostToDevice));
    e->ncb = (int)wc.size();
    e->nrb = (int)wr.size();
    e->lds_cols = lc;
    e->lds_rows = lr;
    e->built_ept = E;
    e->built_dt.assign(dt, dt + e->n);
    return SSG_OK;
}

template <typename R, int E>
int ens_launch_stage(ssg_ensemble* e, int stage_k) {
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)k_cols_batch<R, E, OP_TEND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_RK1>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_RK2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_RK3>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_AB2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((k_cols_batch<R, E, OP_TEND>), dim3(e->ncb), dim3(256), e->lds_cols, e->stream,
                       (const ColsArgs<R>*)e->d_cols, e->d_wcols);
    ENS_OK(e, hipGetLastError());
    const RowsInvArgs<R>* ra = (const RowsInvArgs<R>*)e->d_rows[stage_k];
    if (stage_k == 0)
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_RK1>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    else if (stage_k == 1)
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_RK2>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    else if (stage_k >= 4)
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_AB2>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    else
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_RK3>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    ENS_OK(e, hipGetLastError());
    return SSG_OK;
}

template <typename R>
int ens_stage(ssg_ensemble* e, int stage_k) {
    switch (e->built_ept) {
        case 4: return ens_launch_stage<R, 4>(e, stage_k);
        case 8: return ens_launch_stage<R, 8>(e, stage_k);
        case 16: return ens_launch_stage<R, 16>(e, stage_k);
        default: return ens_err(e, SSG_ERR_ARG, "ensemble tables not built");
    }
}

// apply_spectral_filter! on every member: three launches for the ensemble
template <typename R, int E>
int ens_launch_filter(ssg_ensemble* e, int emit) {
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)k_rows_fwd_batch<R, E, SRC_PAIR>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_cols_batch<R, E, OP_FILT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_ADD>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)k_rows_inv_batch<R, E, EPI_SET>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((k_rows_fwd_batch<R, E, SRC_PAIR>), dim3(e->nffb), dim3(256), e->lds_ffwd, e->stream,
                       (const RowsFwdArgs<R, double>*)e->d_ffwd, e->d_wffwd);
    ENS_OK(e, hipGetLastError());
    hipLaunchKernelGGL((k_cols_batch<R, E, OP_FILT>), dim3(e->nfcb), dim3(256), e->lds_fcols, e->stream,
                       (const ColsArgs<R>*)e->d_fcols, e->d_wfcols);
    ENS_OK(e, hipGetLastError());
    const RowsInvArgs<R>* ra = (const RowsInvArgs<R>*)e->d_frows[emit ? 0 : 1];
    if (std::is_same<R, float>::value)
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_ADD>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    else
        hipLaunchKernelGGL((k_rows_inv_batch<R, E, EPI_SET>), dim3(e->nrb), dim3(256), e->lds_rows, e->stream, ra,
                           e->d_wrows);
    ENS_OK(e, hipGetLastError());
    return SSG_OK;
}

template <typename R>
int ens_filter(ssg_ensemble* e, int emit) {
    switch (e->built_ept) {
        case 4: return ens_launch_filter<R, 4>(e, emit);
        case 8: return ens_launch_filter<R, 8>(e, emit);
        case 16: return ens_launch_filter<R, 16>(e, emit);
        default: return ens_err(e, SSG_ERR_ARG, "ensemble tables not built");
    }
}

// point every member's launches at the ensemble stream for the duration
struct StreamSwap {
    ssg_ensemble* e;
    std::vector<hipStream_t> saved;
    explicit StreamSwap(ssg_ensemble* en) : e(en) {
        for (ssg_ctx* c : e->m) {
            saved.push_back(c->stream);
            c->stream = e->stream;
        }
    }
    ~StreamSwap() {
        for (size_t i = 0; i < e->m.size(); ++i) e->m[i]->stream = saved[i];
    }
};

template <typename R>
int ens_issue_step(ssg_ensemble* e, const ssg_time_params* p, int64_t step_before, bool last) {
    const int64_t step_after = step_before + 1;
    const bool filt = p->filter_freq > 0 && (step_after % p->filter_freq) == 0;
    const bool emit = !filt && !last;
    int rc;
    if (p->scheme == SSG_AB2_LOWSTORAGE) {
        // one tendency per step; forward Euler on the step from step 0
        // (src/timestep.jl:152-190)
        if ((rc = ens_stage<R>(e, 4 + (step_before == 0 ? 2 : 0) + (emit ? 0 : 1)))) return rc;
    } else {
        if ((rc = ens_stage<R>(e, 0))) return rc;
        if ((rc = ens_stage<R>(e, 1))) return rc;
        if ((rc = ens_stage<R>(e, emit ? 2 : 3))) return rc;
    }
    if (filt && (rc = ens_filter<R>(e, last ? 0 : 1))) return rc;
    return SSG_OK;
}

template <typename R>
int ens_prologue(ssg_ensemble* e) {
    for (ssg_ctx* c : e->m) {
        int rc = flux_forward<R>(c, c->ncols_tend);
        if (rc) return ens_err(e, rc, "member prologue: %s", c->err);
    }
    return SSG_OK;
}

// autotuned tile shape: elements per thread of the batched launches, timed
// over one full RK3 step (6 launches); the probe only writes scratch planes
template <typename R>
int ens_tune(ssg_ensemble* e, const double* dt, const ssg_time_params* p) {
    if (e->tuned) return SSG_OK;
    e->tuned = true;
    if (const char* env = getenv("SSG_ENS_EPT")) {
        e->ept = atoi(env);
        return SSG_OK;
    }
    hipEvent_t a, b;
    ENS_OK(e, hipEventCreate(&a));
    ENS_OK(e, hipEventCreate(&b));
    const int cand[3] = {4, 8, 16};
    double best = 1e30;
    int rc = SSG_OK;
    for (int i = 0; i < 3 && rc == SSG_OK; ++i) {
        e->tune_us[i] = -1.0;
        if ((rc = ens_build<R>(e, cand[i], dt, p))) {
            rc = SSG_OK;  // shape does not fit every member: skip
            continue;
        }
        for (int w = 0; w < 2 && rc == SSG_OK; ++w) rc = ens_stage<R>(e, 1);
        if (rc) break;
        (void)hipEventRecord(a, e->stream);
        for (int k = 0; k < 3 && rc == SSG_OK; ++k) rc = ens_stage<R>(e, 1);
        (void)hipEventRecord(b, e->stream);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        e->tune_us[i] = 1e3 * ms;
        if (e->tune_us[i] < best) {
            best = e->tune_us[i];
            e->ept = cand[i];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ens_free_tables(e);
    return rc;
}

template <typename R>
int ens_run(ssg_ensemble* e, const ssg_time_params* p, const double* dt, ssg_time_state* st, int64_t nsteps) {
    StreamSwap swap(e);
    int rc;
    if ((rc = ens_tune<R>(e, dt, p))) return rc;
    const bool same_dt = e->built_dt.size() == (size_t)e->n && std::equal(dt, dt + e->n, e->built_dt.begin());
    const bool same_filter = e->built_filter[0] == p->filter_strength && e->built_filter[1] == p->filter_order &&
                             e->built_filter[2] == p->filter_cutoff;
    if (e->built_ept != e->ept || !same_dt || !same_filter) {
        if (e->graph) {
            (void)hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        if ((rc = ens_build<R>(e, e->ept, dt, p))) return rc;
    }
    if ((rc = ens_prologue<R>(e))) return rc;
    const char* genv = getenv("SSG_GRAPH");
    const bool graphs = !(genv && genv[0] == '0');
    const int F = p->filter_freq > 0 ? p->filter_freq : 16;
    int64_t step = st[0].step;
    int64_t it = 0;
    while (it < nsteps) {
        const int64_t remaining = nsteps - it;
        int adv = 1;
        // the forward-Euler first AB2 step is never part of a replayed cycle
        const bool euler_ahead = p->scheme == SSG_AB2_LOWSTORAGE && step == 0;
        if (graphs && !euler_ahead && (p->filter_freq <= 0 || step % F == 0) && remaining > F) {
            if (!e->graph || e->graph_len != F || !same_params(e->graph_p, *p)) {
                if (e->graph) {
                    (void)hipGraphExecDestroy(e->graph);
                    e->graph = nullptr;
                }
                ENS_OK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                int crc = SSG_OK;
                for (int i = 0; i < F && crc == SSG_OK; ++i) crc = ens_issue_step<R>(e, p, step + i, false);
                hipGraph_t g = nullptr;
                hipError_t ee = hipStreamEndCapture(e->stream, &g);
                if (crc) {
                    if (g) (void)hipGraphDestroy(g);
                    return crc;
                }
                if (ee != hipSuccess) return ens_err(e, SSG_ERR_HIP, "capture: %s", hipGetErrorString(ee));
                ee = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ee != hipSuccess) return ens_err(e, SSG_ERR_HIP, "instantiate: %s", hipGetErrorString(ee));
                e->graph_len = F;
                e->graph_p = *p;
            }
            ENS_OK(e, hipGraphLaunch(e->graph, e->stream));
            adv = F;
        } else {
            if ((rc = ens_issue_step<R>(e, p, step, it == nsteps - 1))) return rc;
        }
        for (int k = 0; k < e->n; ++k)
            for (int i = 0; i < adv; ++i) {
                st[k].t += dt[k];
                st[k].step += 1;
            }
        for (int k = 0; k < e->n; ++k) st[k].dt_actual = dt[k];
        if (p->scheme == SSG_AB2_LOWSTORAGE)
            for (ssg_ctx* c : e->m) c->ab2_have_history = true;
        step += adv;
        it += adv;
    }
    ENS_OK(e, hipStreamSynchronize(e->stream));
    return SSG_OK;
}

}  // namespace

extern "C" {

int ssg_ensemble_create(ssg_ensemble** out, int32_t n, const ssg_domain_desc* descs) {
    if (!out || n < 1 || !descs) return SSG_ERR_ARG;
    *out = nullptr;
    ssg_ensemble* e = new (std::nothrow) ssg_ensemble();
    if (!e) return SSG_ERR_ARG;
    e->n = n;
    e->precision = descs[0].precision;
    e->device = descs[0].device;
    for (int i = 0; i < n; ++i) {
        if (descs[i].precision != e->precision || descs[i].device != e->device || !ens_size_ok(descs[i].nx) ||
            !ens_size_ok(descs[i].ny)) {
            fprintf(stderr, "ssg_ensemble_create: member %d: same precision/device and nx, ny in {128, 256, 512, 1024} "
                    "required\n", i);
            ssg_ensemble_destroy(e);
            return SSG_ERR_UNSUPPORTED;
        }
        ssg_ctx* c = nullptr;
        int rc = ssg_create(&c, &descs[i]);
        if (rc) {
            ssg_ensemble_destroy(e);
            return rc;
        }
        e->m.push_back(c);
    }
    if (hipSetDevice(e->device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        ssg_ensemble_destroy(e);
        return SSG_ERR_HIP;
    }
    *out = e;
    return SSG_OK;
}

int ssg_ensemble_destroy(ssg_ensemble* e) {
    if (!e) return SSG_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    ens_free_tables(e);
    for (ssg_ctx* c : e->m) ssg_destroy(c);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return SSG_OK;
}

ssg_ctx* ssg_ensemble_member(ssg_ensemble* e, int32_t i) {
    if (!e || i < 0 || i >= e->n) return nullptr;
    return e->m[i];
}

const char* ssg_ensemble_last_error(const ssg_ensemble* e) { return e ? e->err : "null ensemble"; }

int ssg_ensemble_timestep(ssg_ensemble* e, const ssg_time_params* p, const double* dt, ssg_time_state* st,
                          int64_t nsteps) {
    if (!e) return SSG_ERR_ARG;
    if (!p || !dt || !st) return ens_err(e, SSG_ERR_ARG, "null pointer");
    if (p->scheme != SSG_RK3 && p->scheme != SSG_AB2_LOWSTORAGE)
        return ens_err(e, SSG_ERR_UNSUPPORTED, "the ensemble step covers RK3 and AB2");
    if (p->scheme == SSG_AB2_LOWSTORAGE && p->ab2_reference_aliasing)
        return ens_err(e, SSG_ERR_UNSUPPORTED, "the ensemble's AB2 uses the corrected tendency");
    if (nsteps < 0) return ens_err(e, SSG_ERR_ARG, "nsteps < 0");
    for (int i = 0; i < e->n; ++i) {
        if (!(dt[i] > 0) || !std::isfinite(dt[i])) return ens_err(e, SSG_ERR_ARG, "dt[%d] must be finite and > 0", i);
        if (st[i].step != st[0].step) return ens_err(e, SSG_ERR_ARG, "members must be at the same step number");
        if (e->m[i]->coupled) return ens_err(e, SSG_ERR_UNSUPPORTED, "member %d is in coupled mode", i);
        if (p->scheme == SSG_AB2_LOWSTORAGE && st[i].step > 0 && !e->m[i]->ab2_have_history)
            return ens_err(e, SSG_ERR_ARG, "AB2 step %lld of member %d needs the previous tendency; none is held",
                           (long long)st[i].step, i);
    }
    if (p->filter_freq > 0 &&
        (!(p->filter_strength >= 0) || p->filter_order < 1 || !(p->filter_cutoff > 0 && p->filter_cutoff <= 1)))
        return ens_err(e, SSG_ERR_ARG, "invalid filter parameters");
    ENS_OK(e, hipSetDevice(e->device));
    if (nsteps == 0) return SSG_OK;
    for (ssg_ctx* c : e->m) ENS_OK(e, hipStreamSynchronize(c->stream));
    return e->precision == SSG_FP32_MIXED ? ens_run<float>(e, p, dt, st, nsteps) : ens_run<double>(e, p, dt, st, nsteps);
}

int ssg_ensemble_tuning(ssg_ensemble* e, int32_t* ept, double* us) {
    if (!e) return SSG_ERR_ARG;
    if (ept) *ept = e->ept;
    if (us)
        for (int i = 0; i < 3; ++i) us[i] = e->tune_us[i];
    return SSG_OK;
}

}  // extern "C"
