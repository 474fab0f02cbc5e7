// Monge-Ampère solve, coupled stepping and the discrete adjoints (host
// side of ssg_ma.hpp); their C entry points.  Declarations: include/ssg.h.
#include "ssg_internal.hpp"

namespace ssg_impl {

template <typename C>
void hist_free(SolveHistoryT<C>& h);

// ---------------------------------------------------------------------------
// Monge-Ampère solve (FP64)
// ---------------------------------------------------------------------------
// Nontemporal access to the Monge-Ampère planes (ld_plane) pays when an
// FP64 plane (C3) exceeds kMaStreamBytes: then the passes stream planes that
// the caches cannot hold.  Smaller planes stay cache-resident between passes
// with plain access (measured: the 256²×16 coupled gradient is 7 % slower
// with nontemporal access, the 1024²×32 coupled step 2 % faster).  The
// environment variable SSG_MA_STREAM=0/1, read when a context first solves,
// overrides the choice; both give bitwise the same results.
constexpr size_t kMaStreamBytes = 128u << 20;

// Stencil coefficients of the vertical operator for the context's dz (the
// uniform grid's Lz/(Nz-1) unless ssg_set_vertical_grid gave one): per
// interior level k the d2dz2! weights (α, β, γ) with h₋ = dz[k-1],
// h₊ = dz[k] (src/transforms.jl:798-850), then the one-sided Neumann rows
// of the bottom and the surface (src/poisson.jl:309-364).
std::vector<double> ma_coefficients(const ssg_ctx* c) {
    const int nz = c->nz;
    std::vector<double> dz = c->dz;
    if ((int)dz.size() != nz) dz.assign(nz, c->lz / (nz - 1));
    std::vector<double> cf(3 * nz + 6, 0.0);
    for (int k = 1; k <= nz - 2; ++k) {
        const double dzm = dz[k - 1], dzp = dz[k];
        cf[3 * k] = 2.0 / (dzm * (dzm + dzp));
        cf[3 * k + 1] = -2.0 / (dzm * dzp);
        cf[3 * k + 2] = 2.0 / (dzp * (dzm + dzp));
    }
    {  // apply_neumann_bottom!, src/poisson.jl:343-364
        const double h1 = dz[0], h2 = dz[1], al = h1 + h2;
        cf[3 * nz] = -(2 * h1 + h2) / (h1 * al);
        cf[3 * nz + 1] = al / (h1 * h2);
        cf[3 * nz + 2] = -h1 / (h2 * al);
    }
    {  // apply_neumann_surface!, src/poisson.jl:309-336
        const double h1 = dz[nz - 2], h2 = dz[nz - 3], al = h1 + h2;
        cf[3 * nz + 3] = (2 * h1 + h2) / (h1 * al);
        cf[3 * nz + 4] = -al / (h1 * h2);
        cf[3 * nz + 5] = h1 / (h2 * al);
    }
    return cf;
}

int ma_upload_coef(ssg_ctx* c) {
    const std::vector<double> cf = ma_coefficients(c);
    HIP_OK(c, hipMemcpyAsync(c->ma_coef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ma_alloc(ssg_ctx* c) {
    if (c->ma_ready) return SSG_OK;
    if (c->nz < 3 || c->nz > 128)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the Monge-Ampere solve needs 3 <= nz <= 128 (nz = %d)", c->nz);
    // the y-pass transforms three lines in lockstep: 3 FP64 line buffers of
    // length ny must fit one CU's 160 KiB of LDS
    if (3 * (size_t)lds_line_elems(c->ny) * sizeof(cpx<double>) > 160 * 1024)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the Monge-Ampere solve supports ny <= 2048 (ny = %d)", c->ny);
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    if (const char* env = getenv("SSG_MA_STREAM"))
        c->ma_stream = atoi(env) != 0 ? 1 : 0;
    else
        c->ma_stream = c3 * sizeof(cpx<double>) > kMaStreamBytes ? 1 : 0;
    // SSG_MA_EXTRAP_FUSED=0: the quadratic warm start in a pass of its own
    // (k_extrap) before the first solve; the same bits either way
    if (const char* env = getenv("SSG_MA_EXTRAP_FUSED")) c->ma_ext_fused = atoi(env) != 0;
    HIP_OK(c, hipMalloc(&c->ma_phi, c3 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->ma_dprev, c3 * sizeof(cpx<double>)));
    for (int i = 0; i < 3; ++i) HIP_OK(c, hipMalloc(&c->ma_p[i], c3 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->ma_b2, (size_t)c->ncols_full * c->ny * sizeof(cpx<double>)));
    if (c->precision == SSG_FP32_MIXED) {
        HIP_OK(c, hipMalloc(&c->ma32_phi, c3 * sizeof(cpx<float>)));
        HIP_OK(c, hipMalloc(&c->ma32_d, c3 * sizeof(cpx<float>)));
    }
    c->ma_partial_cap = c->ncols_full * c->nz + 64;
    HIP_OK(c, hipMalloc(&c->ma_partial, (size_t)c->ma_partial_cap * 2 * sizeof(double)));
    HIP_OK(c, hipHostMalloc(&c->ma_sum_host, sizeof(MaSum), hipHostMallocMapped | hipHostMallocCoherent));
    c->ma_sum_host->r2 = c->ma_sum_host->h2 = kMaSumUnset;
    HIP_OK(c, hipHostGetDevicePointer((void**)&c->ma_sum_dev, c->ma_sum_host, 0));
    HIP_OK(c, hipMalloc(&c->bst, (size_t)c->nx * c->ny * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->ma_coef, (3 * (size_t)c->nz + 6) * sizeof(double)));
    int rc = ma_upload_coef(c);
    if (rc) return rc;
    c->g3 = c->g;
    c->g3.pitch = c->nz * c->ny;
    c->ma_ready = true;
    return SSG_OK;
}

void om_free(ssg_ctx* c) {
    for (int i = 0; i < 3; ++i) (void)hipFree(c->om_pd[i]);
    (void)hipFree(c->om_q);
    (void)hipFree(c->om_gb);
    (void)hipFree(c->om_w);
}

void ma_free(ssg_ctx* c) {
    om_free(c);
    (void)hipFree(c->ma_phi);
    (void)hipFree(c->ma_dprev);
    for (int i = 0; i < 3; ++i) (void)hipFree(c->ma_p[i]);
    (void)hipFree(c->ma_b2);
    (void)hipFree(c->ma32_phi);
    (void)hipFree(c->ma32_d);
    hist_free(c->ma_hist);
    hist_free(c->ma_hist32);
    (void)hipFree(c->ma_coef);
    (void)hipFree(c->ma_partial);
    if (c->ma_sum_host) (void)hipHostFree(c->ma_sum_host);
    (void)hipFree(c->bst);
}

// ---------------------------------------------------------------------------
// Solve history for extrapolated warm starts (struct SolveHistory).  The
// ring holds the iterates of the last converged solves, newest first; the
// buffers rotate between the ring, the solver's working buffer and the
// spares, so no iterate is ever copied.
// ---------------------------------------------------------------------------
int hist_cap(int period) {
    return period == 3 ? 9 : 3;
}

template <typename C>
void hist_free(SolveHistoryT<C>& h) {
    for (int i = 0; i < h.n; ++i) (void)hipFree(h.ring[i]);
    for (int i = 0; i < h.nspare; ++i) (void)hipFree(h.spare[i]);
    h = SolveHistoryT<C>{};
}

// Drop the history when the working buffer itself holds the newest iterate.
template <typename C>
void hist_drop(SolveHistoryT<C>& h) {
    for (int i = 0; i < h.n; ++i) {
        h.spare[h.nspare++] = h.ring[i];
        h.ring[i] = nullptr;
    }
    h.n = 0;
}

// Drop the history, keeping the last iterate as the working buffer (the
// plain warm start's invariant).  Pointer moves only: no device work.
template <typename C>
void hist_reset(SolveHistoryT<C>& h, C*& work) {
    if (h.n == 0) return;
    h.spare[h.nspare++] = work;
    work = h.ring[0];
    for (int i = 1; i < h.n; ++i) h.spare[h.nspare++] = h.ring[i];
    for (int i = 0; i < h.n; ++i) h.ring[i] = nullptr;
    h.n = 0;
}

// The working buffer (holding the newest iterate) becomes ring[0]; the
// working buffer is replaced by the evicted oldest entry, a spare or a new
// buffer.  Entries pushed under another period are dropped first, so the
// ring only holds consecutive solves of one pattern (their lags then mean
// the same stage of earlier steps).
template <typename C>
int hist_push(ssg_ctx* c, SolveHistoryT<C>& h, C*& work, int period) {
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    if (h.period != period) {
        hist_drop(h);
        h.period = period;
    }
    const int cap = hist_cap(period);
    while (h.n > cap) {
        h.spare[h.nspare++] = h.ring[--h.n];
        h.ring[h.n] = nullptr;
    }
    C* freed = nullptr;
    if (h.n == cap) freed = h.ring[--h.n];
    for (int i = h.n; i > 0; --i) h.ring[i] = h.ring[i - 1];
    h.ring[0] = work;
    h.n++;
    if (!freed && h.nspare > 0) freed = h.spare[--h.nspare];
    if (!freed && hipMalloc(&freed, c3 * sizeof(C)) != hipSuccess) {
        // no memory for more history: undo the push and keep the plain warm
        // start (the working buffer keeps the newest iterate)
        (void)hipGetLastError();
        work = h.ring[0];
        for (int i = 0; i + 1 < h.n; ++i) h.ring[i] = h.ring[i + 1];
        h.ring[--h.n] = nullptr;
        hist_drop(h);
        return SSG_OK;
    }
    work = freed;
    return SSG_OK;
}

static int ma_stream_planes(const ssg_ctx* c) { return c->ma_stream; }

// Extrapolated first iterate into the working buffer (needs h.n >= 1).  With
// a stepping period P (solves per step) and three steps of history, the
// guess is the quadratic extrapolation over the same stage of the last three
// steps, 3 x[-P] - 3 x[-2P] + x[-3P]: its error is O(dt³), and it carries no
// error from the RK stage states' own O(dt²) offsets from the trajectory.
// order = 2 takes the same-stage linear 2 x[-P] - x[-2P] instead.  With less
// history: the linear 2 x[-1] - x[-2], else x[-1].
template <typename C>
int hist_extrapolate(ssg_ctx* c, const SolveHistoryT<C>& h, C* work, int period, int order = 3) {
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    const int n = h.n, P = period;
    C* const* r = h.ring;
    if (P > 0 && order == 3 && n >= 3 * P)
        hipLaunchKernelGGL(k_extrap<C>, dim3(grid_all((long)c3)), dim3(256), 0, c->stream, work, r[P - 1], 3.0,
                           r[2 * P - 1], -3.0, r[3 * P - 1], 1.0, (long)c3, ma_stream_planes(c));
    else if (P > 0 && order == 2 && n >= 2 * P)
        hipLaunchKernelGGL(k_extrap<C>, dim3(grid_all((long)c3)), dim3(256), 0, c->stream, work, r[P - 1], 2.0,
                           r[2 * P - 1], -1.0, (const C*)nullptr, 0.0, (long)c3, ma_stream_planes(c));
    else if (n >= 2)
        hipLaunchKernelGGL(k_extrap<C>, dim3(grid_all((long)c3)), dim3(256), 0, c->stream, work, r[0], 2.0, r[1],
                           -1.0, (const C*)nullptr, 0.0, (long)c3, ma_stream_planes(c));
    else {
        HIP_OK(c, hipMemcpyAsync(work, r[0], c3 * sizeof(C), hipMemcpyDeviceToDevice, c->stream));
        return SSG_OK;
    }
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

// b no longer continues the sequence of solves (it was written, or the
// parameters changed): drop the extrapolation history, and φₛ no longer
// counts as solved for b by the coupled stepper
void ma_history_reset(ssg_ctx* c) {
    hist_reset(c->ma_hist, c->ma_dprev);
    hist_drop(c->ma_hist32);
    c->ma_have_d32 = false;
    c->ma_phi_current = false;
}

// the extrapolating form of k_ma_tridiag holds three history entries per
// level in registers: up to 64 levels (beyond, it would spill)
constexpr int kMaExtMaxNz = 64;

// Φ̂ = L⁻¹(b̂ + ε D̂) per wavenumber (D̂ term omitted when use_d = 0); with
// `ext`, D̂ is the warm start formed in the solve and stored to Dh (the
// extrapolating form of k_ma_tridiag): the quadratic 3 ext[0] - 3 ext[1] +
// ext[2] (order 3) or the linear 2 ext[0] - ext[1] (order 2).
template <typename R, typename P>
int ma_tridiag(ssg_ctx* c, const ssg_ma_params* p, int use_d, const cpx<R>* Dh, cpx<P>* Phi,
               const cpx<R>* const* ext = nullptr, int order = 3) {
    if (ext && c->nz > kMaExtMaxNz) return set_err(c, SSG_ERR_ARG, "extrapolating solve with nz > %d", kMaExtMaxNz);
    MaTriArgs<R, P> a{};
    a.Dh = Dh;
    a.B2 = c->ma_b2;
    a.Phi = Phi;
    a.coef = c->ma_coef;
    a.nz = c->nz;
    a.ncols = c->ncols_full;
    a.eps = p->eps;
    a.body = p->variant == SSG_MA_BODY ? 1 : 0;
    a.top = p->variant == SSG_MA_SURFACE ? 1 : 0;
    a.use_d = use_d;
    a.nt = ma_stream_planes(c);
    a.g = c->g;
    if (ext) {
        for (int i = 0; i < 3; ++i) a.ring[i] = ext[i];
        a.next = order == 2 ? 2 : 3;
        a.rc[0] = order == 2 ? 2.0 : 3.0;
        a.rc[1] = order == 2 ? -1.0 : -3.0;
        a.rc[2] = order == 2 ? 0.0 : 1.0;
        a.Dg = const_cast<cpx<R>*>(Dh);
        a.use_d = 1;
    }
    const long total = (long)c->ncols_full * c->ny;
    // two lanes per wavenumber
    const int nblk = (int)((2 * total + 255) / 256);
    const double bytes =
        (double)total * (c->nz * ((ext ? (order == 2 ? 3 : 4) : 1) * sizeof(cpx<R>) + sizeof(cpx<P>)) +
                         sizeof(cpx<double>));
    int slot, rc;
    if ((rc = prof_begin(c, &slot))) return rc;
    auto go = [&](auto kern) { launch_on(c, slot, kern, dim3(nblk), dim3(256), 0, a); };
    if (ext) {  // nz <= kMaExtMaxNz (ma_tridiag_warm)
        if (c->nz <= 8) go(k_ma_tridiag<R, P, 4, true>);
        else if (c->nz <= 16) go(k_ma_tridiag<R, P, 8, true>);
        else if (c->nz <= 32) go(k_ma_tridiag<R, P, 16, true>);
        else go(k_ma_tridiag<R, P, 32, true>);
    } else {
        if (c->nz <= 8) go(k_ma_tridiag<R, P, 4>);
        else if (c->nz <= 16) go(k_ma_tridiag<R, P, 8>);
        else if (c->nz <= 32) go(k_ma_tridiag<R, P, 16>);
        else if (c->nz <= 64) go(k_ma_tridiag<R, P, 32>);
        else go(k_ma_tridiag<R, P, 64>);
    }
    c->launches++;
    HIP_OK(c, hipGetLastError());
    const int kind = ext ? (std::is_same<R, float>::value ? 13 : 5) : (std::is_same<R, float>::value ? 12 : 4);
    return prof_end(c, slot, kind, bytes);
}

// The first iterate of a solve warm-started from history h (of `period`
// solves per step, the working buffer `work`): with `order` steps of
// history the same-stage extrapolation of that order (3 quadratic, 2 linear)
// is formed inside the solve, otherwise hist_extrapolate runs first (as does
// no extrapolation without history).
template <typename R, typename P>
int ma_tridiag_warm(ssg_ctx* c, const ssg_ma_params* p, const SolveHistoryT<cpx<R>>& h, int period, cpx<R>* work,
                    cpx<P>* Phi, int order = 3) {
    const int P3 = period;
    if (c->ma_ext_fused && c->nz <= kMaExtMaxNz && P3 > 0 && h.n >= order * P3) {
        const cpx<R>* const ext[3] = {h.ring[P3 - 1], h.ring[2 * P3 - 1],
                                      h.ring[(order == 2 ? 2 : 3) * P3 - 1]};
        return ma_tridiag<R, P>(c, p, 1, work, Phi, ext, order);
    }
    int rc;
    if (h.n > 0 && (rc = hist_extrapolate(c, h, work, period, order))) return rc;
    return ma_tridiag<R, P>(c, p, 1, work, Phi);
}

// x-inverse of the three derivative planes, D = Φxx Φyy - Φxy², x-FFT of D
// into planes[0]
// elements per thread of k_ma_rows (its own choice: the tendency's row-pass
// choice does not carry over to this strided, bandwidth-bound pass)
template <typename R>
int ma_ept_of(ssg_ctx* c) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    const int e = c->ma_ept[p];
    if (e && ept_valid(c->nx, e)) return e;
    return ept_valid(c->nx, 8) ? 8 : ept_default(c->nx);
}

// elements per thread of the Monge-Ampère column passes (OP_MA_DERIV,
// OP_MA_DIFF): their own tuned choice, else the tendency's column-pass one
template <typename R>
int ma_cols_ept_of(ssg_ctx* c) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    const int e = c->ma_cols_ept[p];
    return (e && ept_valid(c->ny, e)) ? e : ept_of<R>(c, 1, c->ny);
}

// log2 of the row tile of the Monge-Ampère derivative and D planes: the 2G
// rows one k_ma_rows block of path R owns (its line-pair count G follows
// from its elements per thread, as in ma_rows)
template <typename R>
int ma_tile_log2(ssg_ctx* c) {
    // at least 8 rows (64 B runs per kx in FP32, 128 B in FP64), whatever a
    // row-pass block covers; a block of more rows gets a tile of its own
    const int g = lines_per_block<R>(c->nx, ma_ept_of<R>(c), c->nz * c->ny / 2, 1);
    const int yt = std::max(8, 2 * g);
    int s = 0;
    while ((1 << s) < yt) ++s;
    return s;
}

template <typename R>
int ma_rows(ssg_ctx* c, cpx<R>* const (&planes)[3]) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ma_ept_of<R>(c);
    const int rows = c->nz * c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<R>(N, E, rows / 2, 1);
        const int nblk = rows / (2 * G);
        MaRowsArgs<R> a{planes[0], planes[1], planes[2], planes[0], {nullptr, nullptr, nullptr}, c->ncols_full, G, nblk,
                        B.twx, c->g3, ma_tile_log2<R>(c)};
        const size_t lds = (size_t)G * lds_line_elems(N) * sizeof(cpx<R>);
        const double bytes = 4.0 * c->ncols_full * rows * sizeof(cpx<R>);
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_ma_rows<R, N, E>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, std::is_same<R, float>::value ? 11 : 3, bytes);
    });
}

template <typename A, typename B>
int cpx_convert(ssg_ctx* c, const cpx<A>* in, cpx<B>* out, size_t n) {
    hipLaunchKernelGGL((k_cpx_convert<A, B>), dim3(grid1d((long)n)), dim3(256), 0, c->stream, in, out, (long)n);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

std::map<std::tuple<int, int, int, int, int>, int> g_ma_tune_cache;
std::map<std::tuple<int, int, int, int, int>, int> g_ma_cols_tune_cache;

// The Monge-Ampère column passes' elements per thread for path R: one
// derivative pass and one residual pass per candidate (4, 8, 16), timed
// twice after an untimed round, on scratch planes (the residual pass writes
// planes[1], not the solver's D̂).  SSG_MA_COLS_EPT=e forces a choice.
template <typename R>
int ma_tune_cols(ssg_ctx* c, const cpx<R>* phi) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    if (c->ma_cols_ept[p]) return SSG_OK;
    if (const char* env = getenv("SSG_MA_COLS_EPT")) {
        c->ma_cols_ept[p] = atoi(env);
        return SSG_OK;
    }
    const char* tv = getenv("SSG_TUNE");
    if ((tv && tv[0] == '0') || (long)c->nx * c->ny * c->nz < 256L * 256L * 8L) return SSG_OK;
    const auto key = std::make_tuple(c->nx, c->ny, c->nz, p, c->device);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_ma_cols_tune_cache.find(key);
        if (it != g_ma_cols_tune_cache.end()) {
            c->ma_cols_ept[p] = it->second;
            return SSG_OK;
        }
    }
    const bool prof_saved = c->prof;
    c->prof = false;
    cpx<R>* const pl[3] = {(cpx<R>*)c->ma_p[0], (cpx<R>*)c->ma_p[1], (cpx<R>*)c->ma_p[2]};
    const int lines = c->ncols_full * c->nz;
    ColsArgs<R> da{};
    da.inA = phi;
    da.outA = pl[0];
    da.outB = pl[1];
    da.outC = pl[2];
    da.scale = (R)1;
    da.nt = ma_stream_planes(c);
    da.yt_log2 = ma_tile_log2<R>(c);
    ColsArgs<R> ra{};
    ra.inA = pl[0];
    ra.outA = pl[1];
    ra.B2 = c->ma_b2;
    ra.eps = 0.1;
    ra.body = 1;
    ra.partial = c->ma_partial;
    ra.nt = da.nt;
    ra.yt_log2 = da.yt_log2;
    const int cands[3] = {4, 8, 16};
    int best = 0, rc = SSG_OK;
    double best_ms = 1e30;
    for (int ci = 0; ci < 3 && rc == SSG_OK; ++ci) {
        const int e = cands[ci];
        if (!ept_valid(c->ny, e)) continue;
        auto pass = [&]() -> int {
            int r = launch_cols_args<R, OP_MA_DERIV>(c, da, lines, c->nz, e);
            return r ? r : launch_cols_args<R, OP_MA_DIFF>(c, ra, lines, c->nz, e);
        };
        if ((rc = pass())) break;
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) {
            rc = set_err(c, SSG_ERR_HIP, "tuning timing failed");
            break;
        }
        for (int i = 0; i < 2 && rc == SSG_OK; ++i) rc = pass();
        if (rc) break;
        float ms = 0;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
            rc = set_err(c, SSG_ERR_HIP, "tuning timing failed");
            break;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = e;
        }
    }
    c->prof = prof_saved;
    c->ma_cols_ept[p] = best;
    if (rc == SSG_OK && best) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_ma_cols_tune_cache.emplace(key, best);
    }
    return rc;
}

// Pick k_ma_rows' elements per thread for path R on this grid: two timed
// launches per candidate after one untimed one, on the (allocated) planes.
// SSG_MA_EPT=e forces a choice, SSG_TUNE=0 keeps the default.
template <typename R>
int ma_tune(ssg_ctx* c) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    if (c->ma_ept[p]) return SSG_OK;
    if (const char* env = getenv("SSG_MA_EPT")) {
        c->ma_ept[p] = atoi(env);
        return SSG_OK;
    }
    const char* tv = getenv("SSG_TUNE");
    if ((tv && tv[0] == '0') || (long)c->nx * c->ny * c->nz < 256L * 256L * 8L) {
        c->ma_ept[p] = ma_ept_of<R>(c);
        return SSG_OK;
    }
    const auto key = std::make_tuple(c->nx, c->ny, c->nz, p, c->device);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_ma_tune_cache.find(key);
        if (it != g_ma_tune_cache.end()) {
            c->ma_ept[p] = it->second;
            return SSG_OK;
        }
    }
    const bool prof_saved = c->prof;
    c->prof = false;
    cpx<R>* const pl[3] = {(cpx<R>*)c->ma_p[0], (cpx<R>*)c->ma_p[1], (cpx<R>*)c->ma_p[2]};
    const int cands[3] = {4, 8, 16};
    int best = 0, rc = SSG_OK;
    double best_us = 1e30;
    for (int ci = 0; ci < 3 && rc == SSG_OK; ++ci) {
        if (!ept_valid(c->nx, cands[ci])) continue;
        c->ma_ept[p] = cands[ci];
        if ((rc = ma_rows<R>(c, pl))) break;
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) {
            rc = set_err(c, SSG_ERR_HIP, "tuning timing failed");
            break;
        }
        for (int i = 0; i < 2 && rc == SSG_OK; ++i) rc = ma_rows<R>(c, pl);
        if (rc) break;
        float ms = 0;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
            rc = set_err(c, SSG_ERR_HIP, "tuning timing failed");
            break;
        }
        if (ms < best_us) {
            best_us = ms;
            best = cands[ci];
        }
    }
    c->prof = prof_saved;
    c->ma_ept[p] = best;
    if (rc == SSG_OK && best) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_ma_tune_cache.emplace(key, best);
    }
    return rc;
}

// Sums (Σ|r̂|², Σ|b̂ + εD̂|²) of the last OP_MA_DIFF / OP_MA_ADJ pass, reduced on
// the device and published to pinned memory (struct MaSum).  The host polls
// the two words rather than copying the partials and synchronising the
// stream, so it resumes as soon as the sums land.  Every 1024 polls it
// queries the stream: an error ends the wait, and a drained stream without
// the sums is an error, so the wait always ends.
// `ahead` (may be null) enqueues work behind the reduction before the host
// waits for it, so the device does not idle through the host's decision.
static int ma_read_sums(ssg_ctx* c, double* r2, double* h2, int (*ahead)(ssg_ctx*) = nullptr) {
    MaSum* h = c->ma_sum_host;
    __atomic_store_n(&h->r2, kMaSumUnset, __ATOMIC_RELAXED);
    __atomic_store_n(&h->h2, kMaSumUnset, __ATOMIC_RELAXED);
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
    hipLaunchKernelGGL(k_ma_sum, dim3(1), dim3(kMaSumThreads), 0, c->stream, (const double2*)c->ma_partial,
                       c->ma_partial_used, c->ma_sum_dev);
    HIP_OK(c, hipGetLastError());
    if (ahead) {
        const int rc = ahead(c);
        if (rc) return rc;
    }
    unsigned long long a = kMaSumUnset, b = kMaSumUnset;
    auto ready = [&]() {
        a = __atomic_load_n(&h->r2, __ATOMIC_ACQUIRE);
        b = __atomic_load_n(&h->h2, __ATOMIC_ACQUIRE);
        return a != kMaSumUnset && b != kMaSumUnset;
    };
    for (unsigned polls = 1; !ready(); ++polls) {
        if (polls % 1024) {
            __builtin_ia32_pause();
            continue;
        }
        const hipError_t q = hipStreamQuery(c->stream);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return set_err(c, SSG_ERR_HIP, "residual read-back: %s", hipGetErrorString(q));
        if (ready()) break;
        return set_err(c, SSG_ERR_HIP, "residual read-back: the sums were not published");
    }
    memcpy(r2, &a, sizeof(double));
    memcpy(h2, &b, sizeof(double));
    return SSG_OK;
}

// Residual evaluation of the iterate Φ̂ = L⁻¹(b̂ + ε D̂_prev): D̂ of Φ into
// `dh` (which holds D̂_prev on entry) and the residual ε(D̂_prev - D̂) on
// interior levels, reduced over blocks and read back.  `rabs` is
// norm_field(r) (src/poisson.jl:986-995) via Parseval, `rrel` is
// ‖r‖ / ‖b + εD‖.
template <typename R>
int ma_eval(ssg_ctx* c, const ssg_ma_params* p, const cpx<R>* phi, cpx<R>* const (&planes)[3], cpx<R>* dh,
            double* rabs, double* rrel, int (*ahead)(ssg_ctx*) = nullptr) {
    const double scale = 1.0 / ((double)c->nx * c->ny);
    const int lines = c->ncols_full * c->nz;
    ColsArgs<R> a{};
    a.inA = phi;
    a.outA = planes[0];
    a.outB = planes[1];
    a.outC = planes[2];
    a.scale = (R)scale;
    a.nt = ma_stream_planes(c);
    a.yt_log2 = ma_tile_log2<R>(c);
    int rc;
    if ((rc = launch_cols_args<R, OP_MA_DERIV>(c, a, lines, c->nz, ma_cols_ept_of<R>(c)))) return rc;
    if ((rc = ma_rows<R>(c, planes))) return rc;
    ColsArgs<R> b{};
    b.inA = planes[0];
    b.outA = dh;
    b.B2 = c->ma_b2;
    b.eps = p->eps;
    b.body = p->variant == SSG_MA_BODY ? 1 : 0;
    b.partial = c->ma_partial;
    b.nt = a.nt;
    b.yt_log2 = a.yt_log2;
    if ((rc = launch_cols_args<R, OP_MA_DIFF>(c, b, lines, c->nz, ma_cols_ept_of<R>(c)))) return rc;
    double r2 = 0, h2 = 0;
    if ((rc = ma_read_sums(c, &r2, &h2, ahead))) return rc;
    // Parseval: sum of r² over the grid = (1/(nx ny)) sum of |r̂|² per level
    *rabs = std::sqrt(r2 * scale);
    *rrel = h2 > 0 ? std::sqrt(r2 / h2) : std::sqrt(r2 * scale);
    return SSG_OK;
}

// The FP32 phase stops once its residual estimate reaches this level (or
// stops falling); below it FP32 rounding of D̂ dominates the residual.
constexpr double kMaFp32Floor = 1e-6;

// Picard / defect-correction iteration Φ̂ <- L⁻¹(b̂ + ε D̂(Φ)); the residual
// of iterate n is ε(D̂ⁿ⁻¹ - D̂ⁿ) on interior levels, so every iteration
// reports the reference residual of the Φ it holds.
//
// Mixed precision (SSG_FP32_MIXED contexts): the iteration first runs with
// FP32 transforms, products and D̂ (the per-wavenumber solve stays FP64
// arithmetic, rounded to FP32 on output).  The switch solves
// Φ̂ = L⁻¹(b̂ + ε D̂₃₂) in FP64 and carries D̂₃₂ over as D̂_prev, so the first
// FP64 evaluation yields the exact FP64 residual of that Φ; FP64 iterations
// then continue to `tol`.  Convergence is only ever declared on an FP64
// residual.
int ma_iterate(ssg_ctx* c, const ssg_ma_params* p, ssg_ma_diag* d) {
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    int rc0 = ma_tune<double>(c);
    if (rc0 == SSG_OK && c->precision == SSG_FP32_MIXED) rc0 = ma_tune<float>(c);
    if (rc0 == SSG_OK) rc0 = ma_tune_cols<double>(c, c->ma_phi);
    if (rc0 == SSG_OK && c->precision == SSG_FP32_MIXED) rc0 = ma_tune_cols<float>(c, c->ma32_phi);
    if (rc0) return rc0;
    const bool warm = p->warm && c->ma_have_d;
    const bool extrap = warm && p->warm >= 2;
    // A warm solve whose predecessor already started near the FP32 floor
    // skips the FP32 phase: it could not lower the residual, and its passes
    // and precision switch would only add to the FP64 iterations
    const bool phase32 = c->precision == SSG_FP32_MIXED && !(warm && c->ma_last_r0 <= 10.0 * kMaFp32Floor);
    if (!extrap) {
        ma_history_reset(c);  // the last D̂ back in the working buffer
    } else if (phase32) {
        // extrapolated warm start (with an empty ring the working buffer
        // holds the last D̂, a plain warm start); the residual stays exact,
        // since the first iterate is L⁻¹ of whatever D̂ it was built from.
        // Without the FP32 phase the first solve forms it (ma_tridiag_warm).
        if (c->ma_hist.n > 0 && (rc0 = hist_extrapolate(c, c->ma_hist, c->ma_dprev, c->ma_period))) return rc0;
    }
    memset(d, 0, sizeof(*d));
    int it = 0, rc;
    double best = 1e300;
    double ra = 0, rr = 0;
    // record one evaluation; returns true when the iteration must stop
    auto record = [&](bool fp64) -> bool {
        d->final_residual = ra;
        d->relative_residual = rr;
        if (it < 64) d->history[it] = rr;
        d->iterations = it;
        ++it;
        if (!std::isfinite(rr)) return true;
        // divergence (loss of ellipticity): stop once the residual has grown
        // a hundredfold over the best iterate
        best = std::min(best, rr);
        if (rr > 100.0 * best) return true;
        if (fp64 && rr <= p->tol) {
            d->converged = 1;
            return true;
        }
        return it > p->maxiter;
    };
    if (phase32) {
        cpx<float>* const pl[3] = {(cpx<float>*)c->ma_p[0], (cpx<float>*)c->ma_p[1], (cpx<float>*)c->ma_p[2]};
        if (warm) {
            if ((rc = cpx_convert<double, float>(c, c->ma_dprev, c->ma32_d, c3))) return rc;
        } else {
            HIP_OK(c, hipMemsetAsync(c->ma32_d, 0, c3 * sizeof(cpx<float>), c->stream));
        }
        if ((rc = ma_tridiag<float, float>(c, p, warm ? 1 : 0, c->ma32_d, c->ma32_phi))) return rc;
        double prev = 1e300;
        bool stop = false;
        for (;;) {
            if ((rc = ma_eval<float>(c, p, c->ma32_phi, pl, c->ma32_d, &ra, &rr))) return rc;
            d->iterations_fp32++;
            stop = record(false);
            if (stop || rr <= std::max(p->tol, kMaFp32Floor) || rr > 0.7 * prev) break;
            prev = rr;
            if ((rc = ma_tridiag<float, float>(c, p, 1, c->ma32_d, c->ma32_phi))) return rc;
        }
        // switch to FP64; a maxiter stop still gets one FP64 evaluation, so
        // the reported residual is always an FP64 one
        if ((rc = cpx_convert<float, double>(c, c->ma32_d, c->ma_dprev, c3))) return rc;
        if ((rc = ma_tridiag<float, double>(c, p, 1, c->ma32_d, c->ma_phi))) return rc;
        if (stop && (!std::isfinite(rr) || rr > 100.0 * best)) {
            hist_drop(c->ma_hist);
            c->ma_have_d = false;
            c->ma_last_r0 = 1e300;
            return SSG_OK;
        }
    } else {
        if (!warm) HIP_OK(c, hipMemsetAsync(c->ma_dprev, 0, c3 * sizeof(cpx<double>), c->stream));
        if (extrap)
            rc = ma_tridiag_warm<double, double>(c, p, c->ma_hist, c->ma_period, c->ma_dprev, c->ma_phi);
        else
            rc = ma_tridiag<double, double>(c, p, warm ? 1 : 0, c->ma_dprev, c->ma_phi);
        if (rc) return rc;
    }
    cpx<double>* const pl[3] = {c->ma_p[0], c->ma_p[1], c->ma_p[2]};
    for (;;) {
        if ((rc = ma_eval<double>(c, p, c->ma_phi, pl, c->ma_dprev, &ra, &rr))) return rc;
        if (record(true)) break;
        if ((rc = ma_tridiag<double, double>(c, p, 1, c->ma_dprev, c->ma_phi))) return rc;
    }
    if (!std::isfinite(rr) || rr > 100.0 * best) {  // diverged: keep no warm start
        hist_drop(c->ma_hist);
        c->ma_have_d = false;
        c->ma_last_r0 = 1e300;
        return SSG_OK;
    }
    // the working buffer holds the final iterate's D̂
    c->ma_have_d = true;
    c->ma_last_r0 = d->history[0];
    if (p->warm >= 2 && d->converged) return hist_push(c, c->ma_hist, c->ma_dprev, c->ma_period);
    hist_drop(c->ma_hist);
    return SSG_OK;
}

// φₛ = Φ[:,:,Nz] (src/poisson.jl:1119-1131) and u = -∂yφₛ, v = ∂xφₛ
// (src/setICs.jl:161-192), taken from the top level of Φ̂
int ma_surface_outputs(ssg_ctx* c) {
    PathBuffers<double>& D = c->d;
    const double scale = 1.0 / ((double)c->nx * c->ny);
    // the top level's lines are read in place: kx stride nz * ny in Φ̂
    const cpx<double>* top = c->ma_phi + (size_t)(c->nz - 1) * c->ny;
    const long top_stride = (long)c->nz * c->ny;
    ColsArgs<double> a{};
    a.inA = top;
    a.in_stride = top_stride;
    a.outA = D.Fx;
    a.outB = D.Fy;
    a.scale = scale;
    int rc = launch_cols_args<double, OP_VEL_SPEC>(c, a, c->ncols_full, 1);
    if (rc) return rc;
    RowsInvArgs<double> w{};
    w.P0 = D.Fx;
    w.P1 = D.Fy;
    w.ncols_in = c->ncols_full;
    w.w_out = D.w;
    if ((rc = launch_rows_inv<double, MODE_PLANES, EPI_STORE_W>(c, w))) return rc;
    ColsArgs<double> v{};
    v.inA = top;
    v.in_stride = top_stride;
    v.outA = D.Tp;
    v.scale = scale;
    if ((rc = launch_cols_args<double, OP_INV>(c, v, c->ncols_full, 1))) return rc;
    RowsInvArgs<double> f{};
    f.P0 = D.Tp;
    f.ncols_in = c->ncols_full;
    f.out_r = c->phi;
    if ((rc = launch_rows_inv<double, MODE_PAIR, EPI_STORE>(c, f))) return rc;
    if (c->precision == SSG_FP32_MIXED) {
        const long n = (long)c->nx * c->ny;
        hipLaunchKernelGGL((k_w_convert<float>), dim3(grid1d(n)), dim3(256), 0, c->stream, D.w, c->f.w, n);
        c->launches++;
        HIP_OK(c, hipGetLastError());
    }
    return SSG_OK;
}

int ma_check_params(ssg_ctx* c, const ssg_ma_params* p) {
    if (!p) return set_err(c, SSG_ERR_ARG, "null Monge-Ampere parameters");
    if (!(p->eps >= 0) || !std::isfinite(p->eps)) return set_err(c, SSG_ERR_ARG, "eps must be finite and >= 0");
    if (!(p->tol > 0)) return set_err(c, SSG_ERR_ARG, "tol must be > 0");
    if (p->maxiter < 0) return set_err(c, SSG_ERR_ARG, "maxiter must be >= 0");
    if (p->variant != SSG_MA_BODY && p->variant != SSG_MA_SURFACE)
        return set_err(c, SSG_ERR_ARG, "unknown Monge-Ampere variant %d", p->variant);
    if (p->warm < 0 || p->warm > 2) return set_err(c, SSG_ERR_ARG, "warm must be 0, 1 or 2 (got %d)", p->warm);
    if (p->stage_fp32 != 0 && p->stage_fp32 != 1)
        return set_err(c, SSG_ERR_ARG, "stage_fp32 must be 0 or 1 (got %d)", p->stage_fp32);
    return SSG_OK;
}

// Ellipticity of the body-forced problem.  With b on every level and zero
// surface slope, the Picard iterates (and so the fixed point the solver
// seeks) are independent of z, and each level solves tr H - ε det H = b with
// H the horizontal Hessian of Φ.  Then det(I - εH) = 1 - ε tr H + ε² det H
// = 1 - εb pointwise, and I - εH is positive definite (the problem
// elliptic) only where εb < 1.  So εb >= 1 at any point leaves no elliptic
// solution: such a solve is rejected without iterating (d->not_elliptic).
// Checked on cold solves only; a warm solve of a stepping sequence that loses
// ellipticity diverges and is retried cold, which then checks.
int ma_elliptic(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d, bool* ok) {
    *ok = true;
    if (p->variant != SSG_MA_BODY || !(p->eps > 0)) return SSG_OK;
    const long n = (long)c->nx * c->ny;
    const int nb = std::min(c->npartial, (int)((n + 255) / 256));
    hipLaunchKernelGGL(k_bmax, dim3(nb), dim3(256), 0, c->stream, bs, n, c->partial);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    HIP_OK(c, hipMemcpyAsync(c->partial_host, c->partial, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double m = -INFINITY;
    for (int i = 0; i < nb; ++i) m = (c->partial_host[i] > m || c->partial_host[i] != c->partial_host[i]) ? c->partial_host[i] : m;
    if (p->eps * m < 1.0) return SSG_OK;  // NaN b is left to the iteration, which reports it
    memset(d, 0, sizeof(*d));
    d->not_elliptic = 1;
    d->final_residual = d->relative_residual = INFINITY;
    *ok = false;
    set_err(c, SSG_ERR_NOT_CONVERGED,
            "Monge-Ampere problem not elliptic: eps*max(b) = %.6g >= 1 (det(I - eps*H) = 1 - eps*b at the "
            "z-independent solution)",
            p->eps * m);
    return SSG_OK;
}

// full solve for the buoyancy `bs` (device, FP64)
int ma_solve_dev(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d) {
    int rc = ma_alloc(c);
    if (rc) return rc;
    bool ok = true;
    const bool cold = !(p->warm && c->ma_have_d);
    if (cold && (rc = ma_elliptic(c, bs, p, d, &ok))) return rc;
    if (!ok) {
        hist_drop(c->ma_hist);
        c->ma_have_d = false;
        c->last_diag = *d;
        return SSG_OK;
    }
    if ((rc = rfft2_device(c, bs, c->ma_b2))) return rc;
    if ((rc = ma_iterate(c, p, d))) return rc;
    if (!d->converged && p->warm) {
        // a warm start that did not lead to convergence is retried cold
        ssg_ma_params cold_p = *p;
        cold_p.warm = 0;
        if (!cold && (rc = ma_elliptic(c, bs, p, d, &ok))) return rc;
        if (!ok) {
            hist_drop(c->ma_hist);
            c->ma_have_d = false;
            c->last_diag = *d;
            return SSG_OK;
        }
        if ((rc = ma_iterate(c, &cold_p, d))) return rc;
    }
    if ((rc = ma_surface_outputs(c))) return rc;
    c->ma_have_phi = true;
    c->ma_phis_from_phi = true;
    c->last_diag = *d;
    return SSG_OK;
}

// (u, v) = (-∂y, ∂x) of Φ̂'s top level in path precision R, into the path's
// velocity buffer (the tendency's operand); φₛ itself is not formed.
template <typename R>
int ma_surface_w(ssg_ctx* c, const cpx<R>* phi) {
    PathBuffers<R>& B = bufs<R>(c);
    ColsArgs<R> a{};
    a.inA = phi + (size_t)(c->nz - 1) * c->ny;
    a.in_stride = (long)c->nz * c->ny;
    a.outA = B.Fx;
    a.outB = B.Fy;
    a.scale = (R)(1.0 / ((double)c->nx * c->ny));
    int rc = launch_cols_args<R, OP_VEL_SPEC>(c, a, c->ncols_full, 1);
    if (rc) return rc;
    RowsInvArgs<R> w{};
    w.P0 = B.Fx;
    w.P1 = B.Fy;
    w.ncols_in = c->ncols_full;
    w.w_out = B.w;
    return launch_rows_inv<R, MODE_PLANES, EPI_STORE_W>(c, w);
}

// After an FP64 solve in a run with FP32 stage solves: its D̂, rounded,
// continues the FP32 history (it takes that solve's place in the stage
// pattern), and the FP64 history keeps only the newest D̂ as its working
// buffer (pointer moves).
static int ma_seed32(ssg_ctx* c, const ssg_ma_params* p) {
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    const cpx<double>* newest = c->ma_hist.n > 0 ? c->ma_hist.ring[0] : c->ma_dprev;
    int rc = cpx_convert<double, float>(c, newest, c->ma32_d, c3);
    if (rc) return rc;
    hist_reset(c->ma_hist, c->ma_dprev);
    if (p->warm >= 2 && (rc = hist_push(c, c->ma_hist32, c->ma32_d, c->ma_period))) return rc;
    c->ma_have_d32 = true;
    return SSG_OK;
}

// FP32 stage solve (ssg_ma_params.stage_fp32): warm start from the FP32
// history (extrapolated as in hist_extrapolate), then Picard iterations whose
// transforms, products, D̂ and residual sums are FP32 (the per-wavenumber
// solve keeps FP64 arithmetic), to the relative residual
// max(tol, SSG_MA_STAGE_TOL32).  Leaves (u, v) of the FP32 path; φₛ, the
// FP64 velocities and the FP64 Φ̂ keep the last FP64 solve.  d->converged = 0
// when it fails (the caller then solves in FP64).
static int ma_solve_stage32(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d) {
    int rc;
    if ((rc = rfft2_device(c, bs, c->ma_b2))) return rc;
    memset(d, 0, sizeof(*d));
    const double tol32 = std::max(p->tol, (double)SSG_MA_STAGE_TOL32);
    cpx<float>* const pl[3] = {(cpx<float>*)c->ma_p[0], (cpx<float>*)c->ma_p[1], (cpx<float>*)c->ma_p[2]};
    if (!c->ma_extrap32) {
        const char* e = getenv("SSG_MA_EXTRAP32");
        c->ma_extrap32 = (e && e[0] == '3') ? 3 : 2;
    }
    if ((rc = ma_tridiag_warm<float, float>(c, p, c->ma_hist32, c->ma_period, c->ma32_d, c->ma32_phi,
                                            c->ma_extrap32)))
        return rc;
    double best = 1e300, ra = 0, rr = 0;
    // the velocities of this iterate are enqueued behind each evaluation's
    // residual reduction, ahead of the host's decision: the iterate is the
    // solution whenever that evaluation converges (almost always the first),
    // and a later iterate enqueues its own
    auto surface_ahead = [](ssg_ctx* cc) -> int { return ma_surface_w<float>(cc, cc->ma32_phi); };
    for (int it = 0;; ++it) {
        if ((rc = ma_eval<float>(c, p, c->ma32_phi, pl, c->ma32_d, &ra, &rr, +surface_ahead))) return rc;
        d->final_residual = ra;
        d->relative_residual = rr;
        if (it < 64) d->history[it] = rr;
        d->iterations = it;
        d->iterations_fp32 = it + 1;
        if (!std::isfinite(rr)) break;
        best = std::min(best, rr);
        if (rr > 100.0 * best) break;
        if (rr <= tol32) {
            d->converged = 1;
            break;
        }
        if (it >= p->maxiter) break;
        if ((rc = ma_tridiag<float, float>(c, p, 1, c->ma32_d, c->ma32_phi))) return rc;
    }
    if (!d->converged) {
        // the FP32 history no longer leads anywhere; the FP64 solve that
        // follows reseeds it
        hist_drop(c->ma_hist32);
        c->ma_have_d32 = false;
        return SSG_OK;
    }
    c->ma_last_r0 = d->history[0];
    if (p->warm >= 2 && (rc = hist_push(c, c->ma_hist32, c->ma32_d, c->ma_period))) return rc;
    return SSG_OK;  // (u, v) of the converged iterate are already enqueued
}

// SSG_MA_TRACE=1 in the environment prints one line per coupled solve to
// stderr: FP32 or FP64, evaluations, first and last relative residual.
static void ma_trace(const char* kind, const ssg_ma_diag& d) {
    const char* e = getenv("SSG_MA_TRACE");
    if (e && e[0] == '1')
        fprintf(stderr, "ssg_ma_trace %s evals=%d fp32=%d r0=%.3e r=%.3e conv=%d\n", kind, d.iterations + 1,
                d.iterations_fp32, d.history[0], d.relative_residual, d.converged);
}

// The coupled stepper's solve for the buoyancy bs.  With FP32 stage solves
// (FP32-mixed context, params->stage_fp32) every solve but the call's last
// (final_solve) runs in FP32 once a warm start is held; the others, and an
// FP32 solve that fails, run in FP64 from the FP32 history's guess and then
// continue that history.
int ma_solve_coupled(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d, bool final_solve) {
    const bool s32 = p->stage_fp32 && c->precision == SSG_FP32_MIXED;
    int rc;
    if (!s32) {
        rc = ma_solve_dev(c, bs, p, d);
        if (!rc) ma_trace("fp64", *d);
        return rc;
    }
    if ((rc = ma_alloc(c))) return rc;
    if (!final_solve && c->ma_have_d32) {
        if ((rc = ma_solve_stage32(c, bs, p, d))) return rc;
        ma_trace("fp32", *d);
        if (d->converged) {
            c->last_diag = *d;
            return SSG_OK;
        }
    }
    if (c->ma_have_d32) {
        // FP64 solve warm-started from the FP32 history's guess
        const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
        if (c->ma_hist32.n > 0 && (rc = hist_extrapolate(c, c->ma_hist32, c->ma32_d, c->ma_period))) return rc;
        hist_reset(c->ma_hist, c->ma_dprev);
        if ((rc = cpx_convert<float, double>(c, c->ma32_d, c->ma_dprev, c3))) return rc;
        c->ma_have_d = true;
    }
    if ((rc = ma_solve_dev(c, bs, p, d))) return rc;
    ma_trace(final_solve ? "fp64-final" : "fp64", *d);
    if (d->converged && (rc = ma_seed32(c, p))) return rc;
    return SSG_OK;
}

// ---------------------------------------------------------------------------
// Monge-Ampère part of a saved step state: the last solve's Φ̂ (the 3D
// streamfunction of the reference's checkpoint, src/output.jl:376-379) and
// the warm-start state a coupled run continues from (the extrapolation
// ring, or the last D̂), with the coupling parameters.  Restored, a coupled
// run continues bitwise; the surface outputs are recomputed from Φ̂ by the
// same passes that produced them.
// ---------------------------------------------------------------------------
struct MaStateHeader {
    int32_t nz;
    int32_t have_phi;     // Φ̂ follows (C3)
    int32_t have_d;       // a warm start is held: ring entries, or the working D̂ when the ring is empty
    int32_t ring_n;
    int32_t ring_period;
    int32_t phi_current;  // the coupled stepper's end-of-step solve is for the saved b
    int32_t coupled;
    int32_t phis_from_phi;  // φₛ and the velocities came from Φ̂'s top level (not set since)
    double last_r0;
    ssg_ma_params par;
    // from blob version 3: the warm-start state is the FP32 stage solves'
    // history (entries widened to FP64, exactly)
    int32_t ring32;
    int32_t reserved;
};
// the header of version-2 blobs ends before ring32
constexpr size_t kMaStateHeaderV2 = offsetof(MaStateHeader, ring32);

static size_t ma_c3_bytes(const ssg_ctx* c) { return (size_t)c->ncols_full * c->nz * c->ny * sizeof(cpx<double>); }

// the FP32 stage solves' history is what a coupled run continues from
static bool ma_state_ring32(const ssg_ctx* c) { return c->ma_ready && c->ma_par.stage_fp32 && c->ma_have_d32; }

static int ma_state_nbuf(const ssg_ctx* c) {
    if (!c->ma_ready) return 0;
    int n = c->ma_have_phi ? 1 : 0;
    if (ma_state_ring32(c))
        n += c->ma_hist32.n > 0 ? c->ma_hist32.n : 1;
    else if (c->ma_have_d)
        n += c->ma_hist.n > 0 ? c->ma_hist.n : 1;
    return n;
}

size_t ma_state_bytes(const ssg_ctx* c) { return sizeof(MaStateHeader) + ma_state_nbuf(c) * ma_c3_bytes(c); }

// header and device buffers into p (device-to-host copies on the context
// stream; the caller synchronises)
int ma_state_save(ssg_ctx* c, char* p) {
    MaStateHeader h{};
    h.nz = c->nz;
    h.have_phi = c->ma_ready && c->ma_have_phi;
    h.ring32 = ma_state_ring32(c) ? 1 : 0;
    h.have_d = h.ring32 || (c->ma_ready && c->ma_have_d);
    h.ring_n = h.ring32 ? c->ma_hist32.n : (h.have_d ? c->ma_hist.n : 0);
    h.ring_period = h.ring32 ? c->ma_hist32.period : c->ma_hist.period;
    h.phi_current = c->ma_phi_current;
    h.coupled = c->coupled;
    h.phis_from_phi = c->ma_phis_from_phi;
    h.last_r0 = c->ma_last_r0;
    h.par = c->ma_par;
    memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    const size_t c3b = ma_c3_bytes(c);
    auto put = [&](const cpx<double>* src) -> int {
        HIP_OK(c, hipMemcpyAsync(p, src, c3b, hipMemcpyDeviceToHost, c->stream));
        p += c3b;
        return SSG_OK;
    };
    // an FP32 entry is widened into the scratch plane ma_p[0] first (stream
    // order keeps each copy ahead of the next widening)
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    auto put32 = [&](const cpx<float>* src) -> int {
        int r = cpx_convert<float, double>(c, src, c->ma_p[0], c3);
        return r ? r : put(c->ma_p[0]);
    };
    int rc;
    if (h.have_phi && (rc = put(c->ma_phi))) return rc;
    if (h.ring32) {
        if (h.ring_n > 0) {
            for (int i = 0; i < h.ring_n; ++i)
                if ((rc = put32(c->ma_hist32.ring[i]))) return rc;
        } else if ((rc = put32(c->ma32_d))) {
            return rc;
        }
    } else if (h.have_d) {
        if (h.ring_n > 0) {
            for (int i = 0; i < h.ring_n; ++i)
                if ((rc = put(c->ma_hist.ring[i]))) return rc;
        } else if ((rc = put(c->ma_dprev))) {
            return rc;
        }
    }
    return SSG_OK;
}

int ma_state_load(ssg_ctx* c, const char* p, size_t avail, int version) {
    MaStateHeader h{};
    const size_t hb = version >= 3 ? sizeof(h) : kMaStateHeaderV2;
    if (avail < hb) return set_err(c, SSG_ERR_ARG, "state buffer truncated (Monge-Ampere part)");
    memcpy(&h, p, hb);
    p += hb;
    avail -= hb;
    ma_history_reset(c);
    c->ma_have_d = false;
    c->ma_have_phi = false;
    c->coupled = false;
    const int nbuf = (h.have_phi ? 1 : 0) + (h.have_d ? (h.ring_n > 0 ? h.ring_n : 1) : 0);
    if (nbuf == 0 && !h.coupled) return SSG_OK;
    if (h.nz != c->nz)
        return set_err(c, SSG_ERR_ARG, "state of nz = %d does not fit this context (nz = %d)", h.nz, c->nz);
    if (h.ring_n < 0 || h.ring_n > 9) return set_err(c, SSG_ERR_ARG, "corrupt state (ring of %d)", h.ring_n);
    int rc = ma_alloc(c);
    if (rc) return rc;
    const size_t c3b = ma_c3_bytes(c);
    if (avail < nbuf * c3b) return set_err(c, SSG_ERR_ARG, "state buffer truncated (Monge-Ampere planes)");
    auto get = [&](cpx<double>* dst) -> int {
        HIP_OK(c, hipMemcpyAsync(dst, p, c3b, hipMemcpyHostToDevice, c->stream));
        p += c3b;
        return SSG_OK;
    };
    if (h.have_phi && (rc = get(c->ma_phi))) return rc;
    SolveHistory& hs = c->ma_hist;
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    // an FP32 history entry: through the scratch plane ma_p[0], narrowed
    // (exactly: it was widened from FP32)
    auto get32 = [&](cpx<float>* dst) -> int {
        int r = get(c->ma_p[0]);
        return r ? r : cpx_convert<double, float>(c, c->ma_p[0], dst, c3);
    };
    if (h.ring32 && h.have_d) {
        if (c->precision != SSG_FP32_MIXED)
            return set_err(c, SSG_ERR_ARG, "state with FP32 stage solves needs an FP32-mixed context");
        SolveHistory32& h32 = c->ma_hist32;
        if (h.ring_n > 0) {
            for (int i = 0; i < h.ring_n; ++i) {
                cpx<float>* buf = nullptr;
                if (h32.nspare > 0)
                    buf = h32.spare[--h32.nspare];
                else
                    HIP_OK(c, hipMalloc(&buf, c3 * sizeof(cpx<float>)));
                h32.ring[i] = buf;
                h32.n = i + 1;
                if ((rc = get32(buf))) return rc;
            }
            h32.period = h.ring_period;
        } else if ((rc = get32(c->ma32_d))) {
            return rc;
        }
        c->ma_have_d32 = true;
    } else if (h.have_d) {
        if (h.ring_n > 0) {
            for (int i = 0; i < h.ring_n; ++i) {
                cpx<double>* buf = nullptr;
                if (hs.nspare > 0)
                    buf = hs.spare[--hs.nspare];
                else
                    HIP_OK(c, hipMalloc(&buf, c3b));
                hs.ring[i] = buf;
                hs.n = i + 1;
                if ((rc = get(buf))) return rc;
            }
            hs.period = h.ring_period;
        } else if ((rc = get(c->ma_dprev))) {
            return rc;
        }
        c->ma_have_d = true;
    }
    c->ma_last_r0 = h.last_r0;
    if (h.coupled) {
        if ((rc = ma_check_params(c, &h.par))) return rc;
        c->ma_par = h.par;
        c->coupled = true;
    }
    if (h.have_phi) {
        c->ma_have_phi = true;
        if (h.phis_from_phi) {
            // φₛ and the velocities by the passes that made them after the solve
            if ((rc = ma_surface_outputs(c))) return rc;
            c->ma_phis_from_phi = true;
            c->ma_phi_current = h.phi_current != 0;
        }
    }
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}


// ---------------------------------------------------------------------------
// Discrete adjoint of timestep! with frozen φₛ (FP64).  The step is linear in
// b, so its transpose needs no trajectory: for RK3
//   s3 = λ:              g3 = (3 dt/4) Tᵀ s3,  acc = λ + g3,  s2 = (2dt/3) g3
//   s2:                  g2 = Tᵀ s2,           acc += g2,     s1 = (dt/4) λ + (dt/3) g2
//   s1:                  g1 = Tᵀ s1,           λ ← acc + g1
// with Tᵀ s = u Dx(M s) + v Dy(M s).  The filter E[σ] (real, even σ) is
// self-adjoint and is applied to λ before the RK3 transpose of a filter step.
// ---------------------------------------------------------------------------
int adj_alloc(ssg_ctx* c) {
    if (c->adj_lam) return SSG_OK;
    const size_t n = (size_t)c->nx * c->ny;
    HIP_OK(c, hipMalloc(&c->adj_lam, n * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->adj_acc, n * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->adj_s, n * sizeof(double)));
    return SSG_OK;
}

template <int EPI>
int adj_stage(ssg_ctx* c, const double* sin, double gscale, double dt) {
    PathBuffers<double>& D = c->d;
    int rc = launch_rows_fwd<double, double, SRC_PAIR>(c, sin, nullptr, D.Tp, nullptr, c->ncols_tend);
    if (rc) return rc;
    ColsArgs<double> a{};
    a.inA = D.Tp;
    a.outA = D.Fx;
    a.outB = D.Fy;
    a.scale = 1.0 / ((double)c->nx * c->ny);
    if ((rc = launch_cols_args<double, OP_ADJ>(c, a, c->ncols_tend, 1))) return rc;
    RowsInvArgs<double> r{};
    r.P0 = D.Fx;
    r.P1 = D.Fy;
    r.ncols_in = c->ncols_tend;
    r.w = D.w;
    r.lam = c->adj_lam;
    r.acc = c->adj_acc;
    r.sbuf = c->adj_s;
    r.c1 = dt / 3.0;
    r.c2 = 2.0 * dt / 3.0;
    r.c3 = dt / 4.0;
    r.gscale = gscale;
    return launch_rows_inv<double, MODE_PLANES, EPI>(c, r);
}

int adjoint_steps(ssg_ctx* c, const ssg_time_params* p, int64_t step_end, int64_t nsteps) {
    const double dt = p->dt;
    int rc;
    for (int64_t n = step_end; n > step_end - nsteps; --n) {
        // step n took b from step n-1 to n; the filter ran after it when n % F == 0
        if (p->filter_freq > 0 && (n % p->filter_freq) == 0)
            if ((rc = filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0, c->adj_lam)))
                return rc;
        if ((rc = adj_stage<EPI_ADJ3>(c, c->adj_lam, 3.0 * dt / 4.0, dt))) return rc;
        if ((rc = adj_stage<EPI_ADJ2>(c, c->adj_s, 1.0, dt))) return rc;
        if ((rc = adj_stage<EPI_ADJ1>(c, c->adj_s, 1.0, dt))) return rc;
    }
    return SSG_OK;
}

int lincomb(ssg_ctx* c, double* out, double ca, const double* x, double cb, const double* y) {
    const long n = (long)c->nx * c->ny;
    hipLaunchKernelGGL(k_lincomb, dim3(grid1d(n)), dim3(256), 0, c->stream, out, ca, x, cb, y, n);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

// AB2 (frozen φₛ).  State (b, f_old); step n: b' = b + dt(1.5 T b - 0.5 f_old),
// f_old' = T b, with forward Euler b' = b + dt T b on the step from step 0
// (src/timestep.jl:152-190).  Transpose with adjoint state (λ_b, λ_f):
//   s = c λ_b + λ_f  (c = 1.5 dt, or dt on the Euler step)
//   λ_f <- -0.5 dt λ_b (0 on the Euler step),  λ_b <- λ_b + Tᵀ s
// The aliased call site's tendency T_Q b = -2 M Dy(v b) transposes to
// 2 v Dy(M s).  λ_f starts at 0: the objective reads b only.
int adj_tt(ssg_ctx* c, const double* sin, double* out, int quirk) {
    PathBuffers<double>& D = c->d;
    int rc = launch_rows_fwd<double, double, SRC_PAIR>(c, sin, nullptr, D.Tp, nullptr, c->ncols_tend);
    if (rc) return rc;
    ColsArgs<double> a{};
    a.inA = D.Tp;
    a.outA = D.Fx;
    a.outB = D.Fy;
    a.scale = 1.0 / ((double)c->nx * c->ny);
    if ((rc = launch_cols_args<double, OP_ADJ>(c, a, c->ncols_tend, 1))) return rc;
    RowsInvArgs<double> r{};
    r.P0 = D.Fx;
    r.P1 = D.Fy;
    r.ncols_in = c->ncols_tend;
    r.w = D.w;
    r.out_r = out;
    r.gscale = 1.0;
    r.adj_quirk = quirk;
    return launch_rows_inv<double, MODE_PLANES, EPI_ADJG>(c, r);
}

int adjoint_steps_ab2(ssg_ctx* c, const ssg_time_params* p, int64_t step_end, int64_t nsteps) {
    const double dt = p->dt;
    const int quirk = p->ab2_reference_aliasing ? 1 : 0;
    double* lf = c->adj_acc;
    double* sb = c->adj_s;
    double* g = c->dscratch;
    HIP_OK(c, hipMemsetAsync(lf, 0, (size_t)c->nx * c->ny * sizeof(double), c->stream));
    int rc;
    for (int64_t n = step_end; n > step_end - nsteps; --n) {
        if (p->filter_freq > 0 && (n % p->filter_freq) == 0)
            if ((rc = filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0, c->adj_lam)))
                return rc;
        const bool euler = (n - 1) == 0;
        if ((rc = lincomb(c, sb, euler ? dt : 1.5 * dt, c->adj_lam, 1.0, lf))) return rc;
        if ((rc = adj_tt(c, sb, g, quirk))) return rc;
        if ((rc = lincomb(c, lf, euler ? 0.0 : -0.5 * dt, c->adj_lam, 0.0, nullptr))) return rc;
        if ((rc = lincomb(c, c->adj_lam, 1.0, c->adj_lam, 1.0, g))) return rc;
    }
    return SSG_OK;
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the coupled step (Monge-Ampère solve at every stage),
// FP64.  Stage s maps b_s -> Φ_s = MA(b_s) -> w_s = (-Dy φ_s, Dx φ_s) with
// φ_s the top level -> k_s = T(w_s, b_s).  Its transpose, for an incoming
// k̄:
//   b̄_s = T_wᵀ k̄                      (frozen-velocity part)
//       + Bᵀ Gᵀ (I - ε Nᵀ Gᵀ)⁻¹ φ̄     (through the linearised MA solve)
//   φ̄ = Dy ū - Dx v̄,  (ū, v̄) = b_s (Dx M k̄, Dy M k̄)
// where G is the per-wavenumber solve of k_ma_tridiag, N = D'(Φ_s) the
// linearised D operator on interior rows and B places b on the right-hand
// side (every interior row, or the top Neumann row for the surface
// variant).  The bracket is solved by the transposed Picard iteration
// F̄ <- Gᵀ(φ̄ + ε Nᵀ F̄), whose residual is ε Nᵀ(F̄ⁿ - F̄ⁿ⁻¹) -- the mirror
// of the forward iteration.  The forward trajectory is recomputed from
// checkpoints.
// ---------------------------------------------------------------------------
void adjc_free(ssg_ctx* c) {
    hist_free(c->adjc_hist);
    for (int i = 0; i < 3; ++i) {
        (void)hipFree(c->adjc_pd[i]);
        (void)hipFree(c->adjc_bst[i]);
        (void)hipFree(c->adjc_phis[i]);
    }
    (void)hipFree(c->adjc_fbar);
    (void)hipFree(c->adjc_q);
    (void)hipFree(c->adjc_lam);
    (void)hipFree(c->adjc_phibar);
    (void)hipFree(c->adjc_bbar);
    (void)hipFree(c->adjc_wbar);
    (void)hipFree(c->adjc_g);
    (void)hipFree(c->adjc_k1);
    (void)hipFree(c->adjc_tmp);
}

int adjc_alloc(ssg_ctx* c) {
    int rc = ma_alloc(c);
    if (rc) return rc;
    if ((rc = adj_alloc(c))) return rc;
    if (c->adjc_ready) return SSG_OK;
    const size_t n2 = (size_t)c->nx * c->ny, r3 = n2 * c->nz;
    const size_t c2 = (size_t)c->ncols_full * c->ny, c3 = c2 * c->nz;
    for (int i = 0; i < 3; ++i) {
        HIP_OK(c, hipMalloc(&c->adjc_pd[i], r3 * sizeof(double)));
        HIP_OK(c, hipMalloc(&c->adjc_bst[i], n2 * sizeof(double)));
        HIP_OK(c, hipMalloc(&c->adjc_phis[i], c3 * sizeof(cpx<double>)));
    }
    HIP_OK(c, hipMalloc(&c->adjc_fbar, c3 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_q, c3 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_lam, c3 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_phibar, c2 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_bbar, c2 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_wbar, n2 * sizeof(cpx<double>)));
    HIP_OK(c, hipMalloc(&c->adjc_g, n2 * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->adjc_k1, n2 * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->adjc_tmp, n2 * sizeof(double)));
    c->adjc_ready = true;
    return SSG_OK;
}


// inverse 2D transform of a [kx][ky] spectrum (unnormalized forward
// convention) into a real field
int irfft2_device(ssg_ctx* c, cpx<double>* spec, double* out) {
    PathBuffers<double>& D = c->d;
    const double scale = 1.0 / ((double)c->nx * c->ny);
    int rc = launch_cols<double, OP_INV>(c, spec, nullptr, D.Tp, nullptr, c->ncols_full, scale);
    if (rc) return rc;
    RowsInvArgs<double> f{};
    f.P0 = D.Tp;
    f.ncols_in = c->ncols_full;
    f.out_r = out;
    return launch_rows_inv<double, MODE_PAIR, EPI_STORE>(c, f);
}

// Φxx, Φyy, Φxy of the current ma_phi in physical space (adjc_pd)
int ma_phys_derivs(ssg_ctx* c) {
    PathBuffers<double>& D = c->d;
    const int lines = c->ncols_full * c->nz;
    ColsArgs<double> a{};
    a.inA = c->ma_phi;
    a.outA = c->ma_p[0];
    a.outB = c->ma_p[1];
    a.outC = c->ma_p[2];
    a.scale = 1.0 / ((double)c->nx * c->ny);
    a.yt_log2 = ma_tile_log2<double>(c);
    int rc = launch_cols_args<double, OP_MA_DERIV>(c, a, lines, c->nz);
    if (rc) return rc;
    const int ept = ma_ept_of<double>(c);
    const int rows = c->nz * c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<double>(N, E, rows / 2, 1);
        const int nblk = rows / (2 * G);
        MaRowsArgs<double> m{c->ma_p[0], c->ma_p[1], c->ma_p[2], nullptr,
                             {c->adjc_pd[0], c->adjc_pd[1], c->adjc_pd[2]}, c->ncols_full, G, nblk, D.twx, c->g3,
                             ma_tile_log2<double>(c)};
        const size_t lds = (size_t)G * lds_line_elems(N) * sizeof(cpx<double>);
        hipLaunchKernelGGL((k_ma_rows<double, N, E, MA_ROWS_PHYS>), dim3(nblk), dim3(G * T), lds, c->stream, m);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return SSG_OK;
    });
}

int ma_adj_rows(ssg_ctx* c) {
    PathBuffers<double>& D = c->d;
    const int ept = ept_of<double>(c, 2, c->nx);
    const int rows = c->nz * c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<double>(N, E, rows / 2, 3);
        const int nblk = rows / (2 * G);
        MaAdjRowsArgs<double> a{c->adjc_lam, c->adjc_pd[0], c->adjc_pd[1], c->adjc_pd[2],
                                {c->ma_p[0], c->ma_p[1], c->ma_p[2]}, c->ncols_full, G, nblk, D.twx, c->g3};
        const size_t lds = (size_t)G * 3 * lds_line_elems(N) * sizeof(cpx<double>);
        hipLaunchKernelGGL((k_ma_adj_rows<double, N, E>), dim3(nblk), dim3(G * T), lds, c->stream, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return SSG_OK;
    });
}

int ma_tridiag_adj(ssg_ctx* c, const ssg_ma_params* p, int use_q) {
    MaTriAdjArgs a{};
    a.q = c->adjc_q;
    a.phibar = c->adjc_phibar;
    a.Fbar = c->adjc_fbar;
    a.bbar = c->adjc_bbar;
    a.coef = c->ma_coef;
    a.nz = c->nz;
    a.ncols = c->ncols_full;
    a.eps = p->eps;
    a.body = p->variant == SSG_MA_BODY ? 1 : 0;
    a.top = p->variant == SSG_MA_SURFACE ? 1 : 0;
    a.use_q = use_q;
    a.g = c->g;
    const long total = (long)c->ncols_full * c->ny;
    const int nblk = (int)((total + 127) / 128);
    if (c->nz <= 8)
        hipLaunchKernelGGL((k_ma_tridiag_adj<8>), dim3(nblk), dim3(128), 0, c->stream, a);
    else if (c->nz <= 16)
        hipLaunchKernelGGL((k_ma_tridiag_adj<16>), dim3(nblk), dim3(128), 0, c->stream, a);
    else if (c->nz <= 32)
        hipLaunchKernelGGL((k_ma_tridiag_adj<32>), dim3(nblk), dim3(128), 0, c->stream, a);
    else if (c->nz <= 64)
        hipLaunchKernelGGL((k_ma_tridiag_adj<64>), dim3(nblk), dim3(128), 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_ma_tridiag_adj<128>), dim3(nblk), dim3(128), 0, c->stream, a);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

// Transposed linearised Monge-Ampère solve at the current ma_phi (whose
// physical derivative fields are in adjc_pd): adjc_phibar = F(φ̄) in,
// adjc_bbar = F(b̄) out.  With `warm`, inside an adjoint sweep, the first
// iterate is extrapolated from the same stage of the last steps (the sweep
// runs backward, and so does the extrapolation); otherwise it starts at 0.
int ma_adjoint_solve(ssg_ctx* c, const ssg_ma_params* p, ssg_ma_diag* d, bool warm) {
    const double scale = 1.0 / ((double)c->nx * c->ny);
    const int lines = c->ncols_full * c->nz;
    const size_t c3 = (size_t)lines * c->ny;
    memset(d, 0, sizeof(*d));
    int rc;
    const bool start = warm && c->adjc_hist.n > 0;
    if (start) {
        if ((rc = hist_extrapolate(c, c->adjc_hist, c->adjc_q, c->adjc_period))) return rc;
    } else {
        hist_reset(c->adjc_hist, c->adjc_q);
        HIP_OK(c, hipMemsetAsync(c->adjc_q, 0, c3 * sizeof(cpx<double>), c->stream));
    }
    if ((rc = ma_tridiag_adj(c, p, start ? 1 : 0))) return rc;
    double best = 1e300;
    for (int it = 0;; ++it) {
        ColsArgs<double> a{};
        a.inA = c->adjc_fbar;
        a.outA = c->adjc_lam;
        a.scale = scale;
        if ((rc = launch_cols_args<double, OP_INV>(c, a, lines, c->nz))) return rc;
        if ((rc = ma_adj_rows(c))) return rc;
        ColsArgs<double> b{};
        b.inA = c->ma_p[0];
        b.inB = c->ma_p[1];
        b.inC = c->ma_p[2];
        b.outA = c->adjc_q;
        b.B2 = c->adjc_phibar;
        b.eps = p->eps;
        b.partial = c->ma_partial;
        if ((rc = launch_cols_args<double, OP_MA_ADJ>(c, b, lines, c->nz))) return rc;
        double r2 = 0, h2 = 0;
        if ((rc = ma_read_sums(c, &r2, &h2))) return rc;
        d->final_residual = std::sqrt(r2 * scale);
        d->relative_residual = h2 > 0 ? std::sqrt(r2 / h2) : 0.0;
        if (it < 64) d->history[it] = d->relative_residual;
        d->iterations = it;
        if (!std::isfinite(d->relative_residual)) break;
        best = std::min(best, d->relative_residual);
        if (d->relative_residual > 100.0 * best) break;
        if (d->relative_residual <= p->tol) {
            d->converged = 1;
            break;
        }
        if (it >= p->maxiter) break;
        if ((rc = ma_tridiag_adj(c, p, 1))) return rc;
    }
    c->adjc_diag = *d;
    if (!d->converged) {
        hist_drop(c->adjc_hist);
        return set_err(c, SSG_ERR_NOT_CONVERGED, "adjoint Monge-Ampere solve did not converge (rel. residual %.3e)",
                       d->relative_residual);
    }
    // adjc_q holds the last iterate
    if (warm) return hist_push(c, c->adjc_hist, c->adjc_q, c->adjc_period);
    hist_drop(c->adjc_hist);
    return SSG_OK;
}

// One coupled RK3 step in FP64 from c->b, recording the stage buoyancies
// and their Monge-Ampère solutions; c->b ends at the next step's state.
int coupled_step_record(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step_after) {
    const double dt = p->dt;
    const size_t n2 = (size_t)c->nx * c->ny;
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    ssg_ma_diag d{};
    int rc;
    HIP_OK(c, hipMemcpyAsync(c->adjc_bst[0], c->b, n2 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    const int epi_next[3] = {EPI_RK1, EPI_RK2, EPI_RK3};
    for (int st = 0; st < 3; ++st) {
        if ((rc = ma_solve_dev(c, c->adjc_bst[st], mp, &d))) return rc;
        if (!d.converged)
            return ma_not_converged(c, d);
        HIP_OK(c, hipMemcpyAsync(c->adjc_phis[st], c->ma_phi, c3 * sizeof(cpx<double>), hipMemcpyDeviceToDevice,
                                 c->stream));
        if ((rc = flux_forward<double>(c, c->ncols_tend, c->adjc_bst[st]))) return rc;
        if ((rc = tendency_cols<double>(c, 0))) return rc;
        if (epi_next[st] == EPI_RK1)
            rc = stage_inverse<double, EPI_RK1>(c, dt / 3.0, 0, 0, c->adjc_bst[1]);
        else if (epi_next[st] == EPI_RK2)
            rc = stage_inverse<double, EPI_RK2>(c, 2.0 * dt / 3.0, 0, 0, c->adjc_bst[2]);
        else
            rc = stage_inverse<double, EPI_RK3>(c, dt / 4.0, 0, 0);
        if (rc) return rc;
    }
    if (p->filter_freq > 0 && (step_after % p->filter_freq) == 0)
        return filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0);
    return SSG_OK;
}

// b̄ of one stage into adjc_g, for the incoming tendency adjoint `kbar`
int adjc_stage(ssg_ctx* c, const ssg_ma_params* mp, const double* kbar, int st) {
    PathBuffers<double>& D = c->d;
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    const double scale = 1.0 / ((double)c->nx * c->ny);
    // the stage's Monge-Ampère solution, its velocities and derivative fields
    HIP_OK(c, hipMemcpyAsync(c->ma_phi, c->adjc_phis[st], c3 * sizeof(cpx<double>), hipMemcpyDeviceToDevice,
                             c->stream));
    int rc = ma_surface_outputs(c);
    if (rc) return rc;
    if ((rc = ma_phys_derivs(c))) return rc;
    // frozen-velocity transpose and the velocity adjoint (ū, v̄)
    if ((rc = launch_rows_fwd<double, double, SRC_PAIR>(c, kbar, nullptr, D.Tp, nullptr, c->ncols_tend))) return rc;
    ColsArgs<double> a{};
    a.inA = D.Tp;
    a.outA = D.Fx;
    a.outB = D.Fy;
    a.scale = scale;
    if ((rc = launch_cols_args<double, OP_ADJ>(c, a, c->ncols_tend, 1))) return rc;
    RowsInvArgs<double> r{};
    r.P0 = D.Fx;
    r.P1 = D.Fy;
    r.ncols_in = c->ncols_tend;
    r.w = D.w;
    r.b = c->adjc_bst[st];
    r.out_r = c->adjc_g;
    r.w_out = c->adjc_wbar;
    r.gscale = 1.0;
    if ((rc = launch_rows_inv<double, MODE_PLANES, EPI_ADJW>(c, r))) return rc;
    // F(φ̄) = F(Dy ū - Dx v̄)
    if ((rc = launch_rows_fwd<double, double, SRC_W>(c, nullptr, c->adjc_wbar, D.Fx, D.Fy, c->ncols_full))) return rc;
    ColsArgs<double> ph{};
    ph.inA = D.Fx;
    ph.inB = D.Fy;
    ph.outA = c->adjc_phibar;
    ph.scale = 1.0;
    if ((rc = launch_cols_args<double, OP_PHIBAR>(c, ph, c->ncols_full, 1))) return rc;
    // through the linearised Monge-Ampère solve
    ssg_ma_diag d{};
    if ((rc = ma_adjoint_solve(c, mp, &d, mp->warm >= 2 && c->adjc_period > 0))) return rc;
    if ((rc = irfft2_device(c, c->adjc_bbar, c->adjc_tmp))) return rc;
    return lincomb(c, c->adjc_g, 1.0, c->adjc_g, 1.0, c->adjc_tmp);
}

// Reverse sweep over steps [step0 + 1, step0 + nsteps] with the forward
// states b(step0 + j) provided by `state_at(j)` (which leaves it in c->b).
template <typename StateAt>
int adjoint_coupled_sweep(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step0,
                          int64_t nsteps, StateAt&& state_at) {
    const double dt = p->dt;
    int rc;
    for (int64_t j = nsteps; j >= 1; --j) {
        const int64_t n = step0 + j;  // step n took b(n-1) to b(n)
        if (p->filter_freq > 0 && (n % p->filter_freq) == 0)
            if ((rc = filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0, c->adj_lam)))
                return rc;
        if ((rc = state_at(j - 1))) return rc;
        if ((rc = coupled_step_record(c, p, mp, n))) return rc;
        // λ̄ of b(n-1): acc = λ + Σ stage terms
        if ((rc = lincomb(c, c->adj_acc, 1.0, c->adj_lam, 0.0, nullptr))) return rc;
        if ((rc = lincomb(c, c->adj_s, 0.75 * dt, c->adj_lam, 0.0, nullptr))) return rc;      // k̄3
        if ((rc = lincomb(c, c->adjc_k1, 0.25 * dt, c->adj_lam, 0.0, nullptr))) return rc;    // k̄1
        if ((rc = adjc_stage(c, mp, c->adj_s, 2))) return rc;                                 // b̄2
        if ((rc = lincomb(c, c->adj_acc, 1.0, c->adj_acc, 1.0, c->adjc_g))) return rc;
        if ((rc = lincomb(c, c->adj_s, 2.0 * dt / 3.0, c->adjc_g, 0.0, nullptr))) return rc;  // k̄2
        if ((rc = adjc_stage(c, mp, c->adj_s, 1))) return rc;                                 // b̄1
        if ((rc = lincomb(c, c->adj_acc, 1.0, c->adj_acc, 1.0, c->adjc_g))) return rc;
        if ((rc = lincomb(c, c->adjc_k1, 1.0, c->adjc_k1, dt / 3.0, c->adjc_g))) return rc;
        if ((rc = adjc_stage(c, mp, c->adjc_k1, 0))) return rc;                               // b̄0 part
        if ((rc = lincomb(c, c->adj_lam, 1.0, c->adj_acc, 1.0, c->adjc_g))) return rc;
    }
    return SSG_OK;
}

// One coupled AB2 step in FP64 from (c->b, the held tendency), recording the
// step's buoyancy and Monge-Ampère solution in slot 0 (src/timestep.jl:152-190
// with φₛ re-solved before the tendency).
int coupled_step_record_ab2(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step_after) {
    const size_t n2 = (size_t)c->nx * c->ny;
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    ssg_ma_diag d{};
    int rc;
    HIP_OK(c, hipMemcpyAsync(c->adjc_bst[0], c->b, n2 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if ((rc = ma_solve_dev(c, c->adjc_bst[0], mp, &d))) return rc;
    if (!d.converged)
        return ma_not_converged(c, d);
    HIP_OK(c, hipMemcpyAsync(c->adjc_phis[0], c->ma_phi, c3 * sizeof(cpx<double>), hipMemcpyDeviceToDevice,
                             c->stream));
    if ((rc = flux_forward<double>(c, c->ncols_tend, c->adjc_bst[0]))) return rc;
    if ((rc = tendency_cols<double>(c, 0))) return rc;
    if ((rc = stage_inverse<double, EPI_AB2>(c, p->dt, 0, (step_after - 1) == 0 ? 1 : 0))) return rc;
    if (p->filter_freq > 0 && (step_after % p->filter_freq) == 0)
        return filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0);
    return SSG_OK;
}

// Reverse sweep for coupled AB2 with adjoint state (λ_b, λ_f = c->adj_acc):
// s = c λ_b + λ_f (c = 1.5 dt, or dt on the Euler step), λ_f <- -0.5 dt λ_b
// (0 on the Euler step), λ_b <- λ_b + J(b_{n-1})ᵀ s, where J is the
// linearised coupled tendency at the step's starting state.
template <typename StateAt>
int adjoint_coupled_sweep_ab2(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step0,
                              int64_t nsteps, StateAt&& state_at) {
    const double dt = p->dt;
    const size_t n2 = (size_t)c->nx * c->ny;
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
    ssg_ma_diag d{};
    int rc;
    for (int64_t j = nsteps; j >= 1; --j) {
        const int64_t n = step0 + j;
        if (p->filter_freq > 0 && (n % p->filter_freq) == 0)
            if ((rc = filter_b<double>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0, c->adj_lam)))
                return rc;
        if ((rc = state_at(j - 1))) return rc;
        HIP_OK(c, hipMemcpyAsync(c->adjc_bst[0], c->b, n2 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if ((rc = ma_solve_dev(c, c->adjc_bst[0], mp, &d))) return rc;
        if (!d.converged)
            return ma_not_converged(c, d);
        HIP_OK(c, hipMemcpyAsync(c->adjc_phis[0], c->ma_phi, c3 * sizeof(cpx<double>), hipMemcpyDeviceToDevice,
                                 c->stream));
        const bool euler = (n - 1) == 0;
        if ((rc = lincomb(c, c->adj_s, euler ? dt : 1.5 * dt, c->adj_lam, 1.0, c->adj_acc))) return rc;
        if ((rc = adjc_stage(c, mp, c->adj_s, 0))) return rc;
        if ((rc = lincomb(c, c->adj_acc, euler ? 0.0 : -0.5 * dt, c->adj_lam, 0.0, nullptr))) return rc;
        if ((rc = lincomb(c, c->adj_lam, 1.0, c->adj_lam, 1.0, c->adjc_g))) return rc;
    }
    return SSG_OK;
}

// checkpointed adjoint of nsteps coupled steps from the context's current b
// (at step number step0): checkpoints every `every` steps on the forward
// pass, then each segment is recomputed from its checkpoint, storing the
// per-step states, and swept in reverse.  AB2 checkpoints also hold the
// previous tendency, which the recomputation needs.  The context's b and
// AB2 history are left at the start state.
int adjoint_coupled(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step0, int64_t nsteps,
                    int every) {
    const bool ab2 = p->scheme == SSG_AB2_LOWSTORAGE;
    const size_t n2 = (size_t)c->nx * c->ny, bytes = n2 * sizeof(double);
    const int64_t nck = (nsteps + every - 1) / every;
    const int per = ab2 ? 2 : 1;  // checkpoint fields: b (and the held tendency)
    double* ck = nullptr;
    double* seg = nullptr;
    HIP_OK(c, hipMalloc(&ck, (size_t)nck * per * bytes));
    if (hipMalloc(&seg, (size_t)every * bytes) != hipSuccess) {
        (void)hipFree(ck);
        return set_err(c, SSG_ERR_HIP, "checkpoint allocation failed");
    }
    const bool had_history = c->ab2_have_history;
    // solves per step, for the extrapolated warm starts' same-stage history;
    // the transposed solves run as one backward sequence over all segments
    const int saved_period = c->ma_period;
    c->ma_period = ab2 ? 1 : 3;
    c->adjc_period = ab2 ? 1 : 3;
    hist_reset(c->adjc_hist, c->adjc_q);
    auto step = [&](int64_t step_after) -> int {
        return ab2 ? coupled_step_record_ab2(c, p, mp, step_after) : coupled_step_record(c, p, mp, step_after);
    };
    auto save_ck = [&](int64_t k) -> int {
        HIP_OK(c, hipMemcpyAsync(ck + (size_t)k * per * n2, c->b, bytes, hipMemcpyDeviceToDevice, c->stream));
        if (ab2)
            HIP_OK(c, hipMemcpyAsync(ck + ((size_t)k * per + 1) * n2, c->d.fold, bytes, hipMemcpyDeviceToDevice,
                                     c->stream));
        return SSG_OK;
    };
    auto load_ck = [&](int64_t k) -> int {
        HIP_OK(c, hipMemcpyAsync(c->b, ck + (size_t)k * per * n2, bytes, hipMemcpyDeviceToDevice, c->stream));
        if (ab2)
            HIP_OK(c, hipMemcpyAsync(c->d.fold, ck + ((size_t)k * per + 1) * n2, bytes, hipMemcpyDeviceToDevice,
                                     c->stream));
        return SSG_OK;
    };
    auto body = [&]() -> int {
        int rc;
        if (ab2) {
            HIP_OK(c, hipMemsetAsync(c->adj_acc, 0, bytes, c->stream));  // λ_f
            // the FP64 recomputation starts from the held tendency, which an
            // FP32-mixed context keeps in FP32
            if (c->precision == SSG_FP32_MIXED) {
                hipLaunchKernelGGL((k_convert<float, double>), dim3(grid1d((long)n2)), dim3(256), 0, c->stream,
                                   c->f.fold, c->d.fold, (long)n2);
                HIP_OK(c, hipGetLastError());
            }
        }
        // The solves below run in three unbroken sequences: the forward pass,
        // each segment's recomputation (forward from its checkpoint), and
        // each segment's sweep (backward, one step at a time).  The
        // extrapolated warm start's history is reset at every break, so it
        // only ever extrapolates along one sequence, in its own direction.
        ma_history_reset(c);
        for (int64_t j = 0; j < nsteps; ++j) {
            if (j % every == 0 && (rc = save_ck(j / every))) return rc;
            if ((rc = step(step0 + j + 1))) return rc;
        }
        for (int64_t sg = nck - 1; sg >= 0; --sg) {
            const int64_t j0 = sg * every;
            const int64_t len = std::min<int64_t>(every, nsteps - j0);
            // recompute the segment's per-step states from its checkpoint
            if ((rc = load_ck(sg))) return rc;
            ma_history_reset(c);
            for (int64_t k = 0; k < len; ++k) {
                HIP_OK(c, hipMemcpyAsync(seg + (size_t)k * n2, c->b, bytes, hipMemcpyDeviceToDevice, c->stream));
                if (k + 1 < len && (rc = step(step0 + j0 + k + 1))) return rc;
            }
            auto state_at = [&](int64_t jj) -> int {
                HIP_OK(c, hipMemcpyAsync(c->b, seg + (size_t)jj * n2, bytes, hipMemcpyDeviceToDevice, c->stream));
                return SSG_OK;
            };
            ma_history_reset(c);
            rc = ab2 ? adjoint_coupled_sweep_ab2(c, p, mp, step0 + j0, len, state_at)
                     : adjoint_coupled_sweep(c, p, mp, step0 + j0, len, state_at);
            if (rc) return rc;
        }
        // leave b (and the AB2 history) at the start state
        return load_ck(0);
    };
    const int rc = nsteps > 0 ? body() : SSG_OK;
    (void)hipStreamSynchronize(c->stream);
    c->ma_period = saved_period;
    c->adjc_period = 0;
    hist_reset(c->adjc_hist, c->adjc_q);
    c->ab2_have_history = had_history;
    (void)hipFree(ck);
    (void)hipFree(seg);
    return rc;
}

// ---------------------------------------------------------------------------
// Vertical velocity from the ω equation (src/extras/omega_eq.jl:574-652,
// compute_vertical_velocity) for the current surface buoyancy: Φ from the
// Monge-Ampère solve, Q and -2ε∇·Q per level, the per-wavenumber vertical
// solve with w* = 0 on the bottom and top levels, w = w*/(1 - ε∇ₕ²Φ).
// ---------------------------------------------------------------------------
int omega_alloc(ssg_ctx* c) {
    if (c->om_w) return SSG_OK;
    const size_t n2 = (size_t)c->nx * c->ny, n3 = n2 * c->nz;
    for (int i = 0; i < 3; ++i) HIP_OK(c, hipMalloc(&c->om_pd[i], n3 * sizeof(double)));
    HIP_OK(c, hipMalloc(&c->om_q, n3 * sizeof(double2)));
    HIP_OK(c, hipMalloc(&c->om_gb, n2 * sizeof(double2)));
    HIP_OK(c, hipMalloc(&c->om_w, n3 * sizeof(double)));
    return SSG_OK;
}

int omega_w(ssg_ctx* c, double eps) {
    PathBuffers<double>& D = c->d;
    const double scale = 1.0 / ((double)c->nx * c->ny);
    const int lines = c->ncols_full * c->nz, rows = c->nz * c->ny;
    const long n2 = (long)c->nx * c->ny, n3 = n2 * c->nz;
    // Φxx, Φyy, Φxy of every level in physical space
    ColsArgs<double> a{};
    a.inA = c->ma_phi;
    a.outA = c->ma_p[0];
    a.outB = c->ma_p[1];
    a.outC = c->ma_p[2];
    a.scale = scale;
    a.yt_log2 = ma_tile_log2<double>(c);
    int rc = launch_cols_args<double, OP_MA_DERIV>(c, a, lines, c->nz);
    if (rc) return rc;
    rc = with_size_ept(c->nx, ma_ept_of<double>(c), [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<double>(N, E, rows / 2, 1);
        const int nblk = rows / (2 * G);
        MaRowsArgs<double> m{c->ma_p[0], c->ma_p[1], c->ma_p[2], nullptr, {c->om_pd[0], c->om_pd[1], c->om_pd[2]},
                             c->ncols_full, G, nblk, D.twx, c->g3, ma_tile_log2<double>(c)};
        const size_t lds = (size_t)G * lds_line_elems(N) * sizeof(cpx<double>);
        hipLaunchKernelGGL((k_ma_rows<double, N, E, MA_ROWS_PHYS>), dim3(nblk), dim3(G * T), lds, c->stream, m);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return SSG_OK;
    });
    if (rc) return rc;
    // (-∂y b, ∂x b) of the surface buoyancy, as the velocities are made from φₛ
    if ((rc = launch_rows_fwd<double, double, SRC_PAIR>(c, c->b, nullptr, D.Fx, nullptr, c->ncols_full))) return rc;
    if ((rc = launch_cols<double, OP_VEL>(c, D.Fx, nullptr, D.Tp, D.Fy, c->ncols_full, scale))) return rc;
    RowsInvArgs<double> g{};
    g.P0 = D.Tp;
    g.P1 = D.Fy;
    g.ncols_in = c->ncols_full;
    g.w_out = c->om_gb;
    if ((rc = launch_rows_inv<double, MODE_PLANES, EPI_STORE_W>(c, g))) return rc;
    // Q per level, x-FFT into two planes, y-FFT and -2ε ∇·Q
    hipLaunchKernelGGL(k_omega_q, dim3(grid_all(n3)), dim3(256), 0, c->stream, c->om_pd[0], c->om_pd[1], c->om_pd[2],
                       (const double2*)c->om_gb, c->om_q, n2, n3);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    if ((rc = launch_rows_fwd<double, double, SRC_W>(c, nullptr, c->om_q, c->ma_p[0], c->ma_p[1], c->ncols_full, rows,
                                                      &c->g3)))
        return rc;
    ColsArgs<double> dq{};
    dq.inA = c->ma_p[0];
    dq.inB = c->ma_p[1];
    dq.outA = c->ma_p[2];
    dq.eps = eps;
    if ((rc = launch_cols_args<double, OP_DIVQ>(c, dq, lines, c->nz))) return rc;
    // per-wavenumber vertical solve
    OmegaTriArgs t{};
    t.F = c->ma_p[2];
    t.W = c->ma_p[1];
    t.coef = c->ma_coef;
    t.nz = c->nz;
    t.ncols = c->ncols_full;
    t.g = c->g;
    const long total = (long)c->ncols_full * c->ny;
    const int nblk = (int)((total + 127) / 128);
    if (c->nz <= 8)
        hipLaunchKernelGGL(k_omega_tridiag<8>, dim3(nblk), dim3(128), 0, c->stream, t);
    else if (c->nz <= 16)
        hipLaunchKernelGGL(k_omega_tridiag<16>, dim3(nblk), dim3(128), 0, c->stream, t);
    else if (c->nz <= 32)
        hipLaunchKernelGGL(k_omega_tridiag<32>, dim3(nblk), dim3(128), 0, c->stream, t);
    else if (c->nz <= 64)
        hipLaunchKernelGGL(k_omega_tridiag<64>, dim3(nblk), dim3(128), 0, c->stream, t);
    else
        hipLaunchKernelGGL(k_omega_tridiag<128>, dim3(nblk), dim3(128), 0, c->stream, t);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    // w* per level in physical space, then w = w*/J⁻¹
    ColsArgs<double> iv{};
    iv.inA = c->ma_p[1];
    iv.outA = c->ma_p[0];
    iv.scale = scale;
    if ((rc = launch_cols_args<double, OP_INV>(c, iv, lines, c->nz))) return rc;
    RowsInvArgs<double> r{};
    r.P0 = c->ma_p[0];
    r.ncols_in = c->ncols_full;
    r.out_r = c->om_w;
    if ((rc = launch_rows_inv<double, MODE_PAIR, EPI_STORE>(c, r, rows, &c->g3))) return rc;
    hipLaunchKernelGGL(k_omega_final, dim3(grid_all(n3)), dim3(256), 0, c->stream, c->om_w, c->om_pd[0], c->om_pd[1],
                       eps, n3);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

}  // namespace ssg_impl

extern "C" {

int ssg_vertical_velocity(ssg_ctx* c, const ssg_ma_params* p, double* w, int64_t n, ssg_ma_diag* diag) {
    if (!c) return SSG_ERR_ARG;
    if (!w) return set_err(c, SSG_ERR_ARG, "null pointer");
    int rc = ma_check_params(c, p);
    if (rc) return rc;
    const int64_t need = (int64_t)c->nx * c->ny * c->nz;
    if (n != need) return set_err(c, SSG_ERR_ARG, "size %lld != nx*ny*nz = %lld", (long long)n, (long long)need);
    HIP_OK(c, hipSetDevice(c->device));
    ssg_ma_diag d{};
    c->ma_phi_current = false;
    if ((rc = ma_solve_dev(c, c->b, p, &d))) return rc;
    c->last_diag = d;
    if (diag) *diag = d;
    if (!d.converged) return ma_not_converged(c, d);
    if ((rc = omega_alloc(c))) return rc;
    if ((rc = omega_w(c, p->eps))) return rc;
    HIP_OK(c, hipMemcpyAsync(w, c->om_w, (size_t)need * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_set_b(ssg_ctx* c, const double* host, int64_t n, const ssg_ma_params* p, ssg_ma_diag* d) {
    if (!c) return SSG_ERR_ARG;
    if (!host) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (n != (int64_t)c->nx * c->ny)
        return set_err(c, SSG_ERR_ARG, "field size %lld != nx*ny", (long long)n);
    int rc;
    if (p && (rc = ma_check_params(c, p))) return rc;
    if ((rc = ssg_set_field(c, SSG_FIELD_B, host, n))) return rc;
    // zero mean and 2/3 dealias in spectral space (src/setICs.jl:30-52)
    PathBuffers<double>& D = c->d;
    if ((rc = launch_rows_fwd<double, double, SRC_PAIR>(c, c->b, nullptr, D.Tp, nullptr, c->ncols_full))) return rc;
    if ((rc = launch_cols<double, OP_SETB>(c, D.Tp, nullptr, D.Tp, nullptr, c->ncols_full,
                                           1.0 / ((double)c->nx * c->ny))))
        return rc;
    RowsInvArgs<double> a{};
    a.P0 = D.Tp;
    a.ncols_in = c->ncols_full;
    a.b = c->b;
    if ((rc = launch_rows_inv<double, MODE_PAIR, EPI_SET>(c, a))) return rc;
    if (!p) {
        HIP_OK(c, hipStreamSynchronize(c->stream));
        return SSG_OK;
    }
    // the Monge-Ampère solve for φₛ and the surface velocities
    // (src/setICs.jl:55-62)
    ssg_ma_diag dd{};
    HIP_OK(c, hipEventRecord(c->ev0, c->stream));
    rc = ma_solve_dev(c, c->b, p, &dd);
    HIP_OK(c, hipEventRecord(c->ev1, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_OK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    dd.solve_ms = ms;
    c->last_diag = dd;
    if (d) *d = dd;
    if (rc) return rc;
    if (!dd.converged)
        return ma_not_converged(c, dd);
    return SSG_OK;
}

int ssg_ma_solve(ssg_ctx* c, const ssg_ma_params* p, ssg_ma_diag* d) {
    if (!c) return SSG_ERR_ARG;
    int rc = ma_check_params(c, p);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    ssg_ma_diag local{};
    if (!d) d = &local;
    c->ma_phi_current = false;
    HIP_OK(c, hipEventRecord(c->ev0, c->stream));
    if ((rc = ma_solve_dev(c, c->b, p, d))) return rc;
    HIP_OK(c, hipEventRecord(c->ev1, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_OK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    d->solve_ms = ms;
    c->last_diag = *d;
    if (!d->converged)
        return ma_not_converged(c, *d);
    return SSG_OK;
}

int ssg_get_phi3d(ssg_ctx* c, double* host, int64_t n) {
    if (!c) return SSG_ERR_ARG;
    if (!host) return set_err(c, SSG_ERR_ARG, "null pointer");
    const int64_t need = (int64_t)c->nx * c->ny * c->nz;
    if (n != need) return set_err(c, SSG_ERR_ARG, "size %lld != nx*ny*nz = %lld", (long long)n, (long long)need);
    if (!c->ma_ready) return set_err(c, SSG_ERR_ARG, "no Monge-Ampere solution yet");
    HIP_OK(c, hipSetDevice(c->device));
    const int lines = c->ncols_full * c->nz;
    ColsArgs<double> a{};
    a.inA = c->ma_phi;
    a.outA = c->ma_p[0];
    a.scale = 1.0 / ((double)c->nx * c->ny);
    int rc = launch_cols_args<double, OP_INV>(c, a, lines, c->nz);
    if (rc) return rc;
    RowsInvArgs<double> r{};
    r.P0 = c->ma_p[0];
    r.ncols_in = c->ncols_full;
    r.out_r = reinterpret_cast<double*>(c->ma_p[1]);
    if ((rc = launch_rows_inv<double, MODE_PAIR, EPI_STORE>(c, r, c->nz * c->ny, &c->g3))) return rc;
    HIP_OK(c, hipMemcpyAsync(host, c->ma_p[1], (size_t)need * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_set_vertical_grid(ssg_ctx* c, const double* z, const double* dz, int32_t nz) {
    if (!c) return SSG_ERR_ARG;
    if (!z) return set_err(c, SSG_ERR_ARG, "null z");
    if (nz != c->nz) return set_err(c, SSG_ERR_ARG, "Custom z_coords length (%d) must match Nz (%d)", nz, c->nz);
    if (nz < 3) return set_err(c, SSG_ERR_UNSUPPORTED, "a vertical grid needs nz >= 3 (nz = %d)", nz);
    for (int k = 0; k + 1 < nz; ++k)
        if (!(z[k + 1] > z[k]) || !std::isfinite(z[k]) || !std::isfinite(z[k + 1]))
            return set_err(c, SSG_ERR_ARG, "z must be finite and strictly increasing (z[%d] = %g, z[%d] = %g)", k, z[k],
                           k + 1, z[k + 1]);
    std::vector<double> h(nz);
    if (dz) {
        for (int k = 0; k < nz; ++k) {
            if (!(dz[k] > 0) || !std::isfinite(dz[k]))
                return set_err(c, SSG_ERR_ARG, "dz[%d] = %g must be finite and > 0", k, dz[k]);
            h[k] = dz[k];
        }
    } else {
        // the :custom rule of make_vertical_grid (src/domain.jl:258-275)
        for (int k = 1; k + 1 < nz; ++k) h[k] = 0.5 * (z[k + 1] - z[k - 1]);
        h[0] = z[1] - z[0];
        h[nz - 1] = z[nz - 1] - z[nz - 2];
    }
    HIP_OK(c, hipSetDevice(c->device));
    c->dz = h;
    // the operator changed: no earlier solve continues into this one
    ma_history_reset(c);
    c->ma_have_d = false;
    if (c->ma_ready) return ma_upload_coef(c);
    return SSG_OK;
}

int ssg_set_coupling(ssg_ctx* c, const ssg_ma_params* p) {
    if (!c) return SSG_ERR_ARG;
    // the same parameters again (e.g. after a restart that restored them)
    // leave the coupled run's history in place
    if (p && c->coupled && p->eps == c->ma_par.eps && p->tol == c->ma_par.tol && p->maxiter == c->ma_par.maxiter &&
        p->variant == c->ma_par.variant && p->warm == c->ma_par.warm && p->stage_fp32 == c->ma_par.stage_fp32)
        return SSG_OK;
    ma_history_reset(c);
    if (!p) {
        c->coupled = false;
        return SSG_OK;
    }
    int rc = ma_check_params(c, p);
    if (rc) return rc;
    if (p->stage_fp32 && c->precision != SSG_FP32_MIXED)
        return set_err(c, SSG_ERR_ARG, "FP32 stage solves (stage_fp32) need an FP32-mixed context");
    if (p->stage_fp32 && p->warm < 1)
        return set_err(c, SSG_ERR_ARG, "FP32 stage solves (stage_fp32) start from the held history: need warm >= 1");
    HIP_OK(c, hipSetDevice(c->device));
    if ((rc = ma_alloc(c))) return rc;
    c->ma_par = *p;
    c->coupled = true;
    return SSG_OK;
}

int ssg_adjoint(ssg_ctx* c, const ssg_time_params* p, int64_t step_end, int64_t nsteps, const double* lam_in,
                double* lam_out) {
    if (!c) return SSG_ERR_ARG;
    if (!p || !lam_in || !lam_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (p->scheme != SSG_RK3 && p->scheme != SSG_AB2_LOWSTORAGE)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the adjoint covers the RK3 and AB2 schemes");
    if (c->coupled)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the coupled step's adjoint is ssg_adjoint_coupled");
    if (nsteps < 0 || step_end < nsteps) return set_err(c, SSG_ERR_ARG, "need 0 <= nsteps <= step_end");
    if (!(p->dt > 0)) return set_err(c, SSG_ERR_ARG, "dt must be > 0");
    HIP_OK(c, hipSetDevice(c->device));
    int rc = adj_alloc(c);
    if (rc) return rc;
    const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(c->adj_lam, lam_in, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = p->scheme == SSG_RK3 ? adjoint_steps(c, p, step_end, nsteps)
                                   : adjoint_steps_ab2(c, p, step_end, nsteps)))
        return rc;
    HIP_OK(c, hipMemcpyAsync(lam_out, c->adj_lam, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_adjoint_coupled(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step_start,
                        int64_t nsteps, int32_t checkpoint_every, const double* lam_in, double* lam_out) {
    if (!c) return SSG_ERR_ARG;
    if (!p || !mp || !lam_in || !lam_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (p->scheme != SSG_RK3 && p->scheme != SSG_AB2_LOWSTORAGE)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the adjoint covers the RK3 and AB2 schemes");
    if (p->scheme == SSG_AB2_LOWSTORAGE && p->ab2_reference_aliasing)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the coupled AB2 adjoint uses the corrected tendency");
    if (nsteps < 0 || step_start < 0) return set_err(c, SSG_ERR_ARG, "need nsteps >= 0 and step_start >= 0");
    if (p->scheme == SSG_AB2_LOWSTORAGE && step_start > 0 && !c->ab2_have_history)
        return set_err(c, SSG_ERR_ARG, "AB2 step %lld needs the previous tendency; none is held",
                       (long long)step_start);
    if (checkpoint_every < 1) return set_err(c, SSG_ERR_ARG, "checkpoint_every must be >= 1");
    if (!(p->dt > 0)) return set_err(c, SSG_ERR_ARG, "dt must be > 0");
    int rc = ma_check_params(c, mp);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    if ((rc = adjc_alloc(c))) return rc;
    const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(c->adj_lam, lam_in, bytes, hipMemcpyHostToDevice, c->stream));
    if (nsteps > 0 && (rc = adjoint_coupled(c, p, mp, step_start, nsteps, checkpoint_every))) return rc;
    ma_history_reset(c);  // the recomputation's solves do not continue a forward sequence
    if (nsteps > 0) {
        // b is back at the start state; φₛ and the velocities are those of
        // the last recomputed stage: solve them for b again
        ssg_ma_diag d{};
        if ((rc = ma_solve_dev(c, c->b, mp, &d))) return rc;
        if (!d.converged)
            return set_err(c, SSG_ERR_NOT_CONVERGED, "Monge-Ampere solve for the start state did not converge");
    }
    HIP_OK(c, hipMemcpyAsync(lam_out, c->adj_lam, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_ma_adjoint(ssg_ctx* c, const ssg_ma_params* mp, const double* phibar, double* bbar, ssg_ma_diag* diag) {
    if (!c) return SSG_ERR_ARG;
    if (!mp || !phibar || !bbar) return set_err(c, SSG_ERR_ARG, "null pointer");
    int rc = ma_check_params(c, mp);
    if (rc) return rc;
    if (!c->ma_ready || !c->ma_have_d)
        return set_err(c, SSG_ERR_ARG, "no Monge-Ampere solution to linearise about (call ssg_ma_solve first)");
    HIP_OK(c, hipSetDevice(c->device));
    if ((rc = adjc_alloc(c))) return rc;
    const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(c->adjc_tmp, phibar, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = rfft2_device(c, c->adjc_tmp, c->adjc_phibar))) return rc;
    if ((rc = ma_phys_derivs(c))) return rc;
    ssg_ma_diag d{};
    rc = ma_adjoint_solve(c, mp, &d, false);
    if (diag) *diag = d;
    if (rc) return rc;
    if ((rc = irfft2_device(c, c->adjc_bbar, c->adjc_tmp))) return rc;
    HIP_OK(c, hipMemcpyAsync(bbar, c->adjc_tmp, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_last_ma_diag(ssg_ctx* c, ssg_ma_diag* d) {
    if (!c || !d) return SSG_ERR_ARG;
    *d = c->last_diag;
    return SSG_OK;
}

}  // extern "C"
