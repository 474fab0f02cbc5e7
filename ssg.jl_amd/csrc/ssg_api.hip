// C ABI of the SSG time-step library: context, field transfer, the step
// driver and the spectral sub-boundary.  Declarations: include/ssg.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/ssg.h"
#include "ssg_kernels.hpp"
#include "ssg_ma.hpp"

using namespace ssg;

namespace {

constexpr int kMaxLdsBytes = 64 * 1024;

template <typename R>
struct PathBuffers {
    cpx<R>* w = nullptr;    // (u, v)
    R* k1 = nullptr;
    R* fold = nullptr;
    R* tmp = nullptr;       // real scratch
    cpx<R>* Fx = nullptr;   // planes, (nx/2+1) * ny
    cpx<R>* Fy = nullptr;
    cpx<R>* Tp = nullptr;
    cpx<R>* twx = nullptr;
    cpx<R>* twy = nullptr;
};

}  // namespace

struct ssg_ctx {
    int nx = 0, ny = 0, nz = 0;
    double lx = 0, ly = 0, lz = 0;
    int precision = SSG_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    Grid g{};
    int ncols_full = 0;   // nx/2 + 1
    int ncols_tend = 0;   // K + 1
    double* b = nullptr;
    double* phi = nullptr;
    double* dscratch = nullptr;   // nx*ny FP64
    double2* spec = nullptr;      // nx*(ny/2+1) reference-layout spectrum
    double* partial = nullptr;    // reduction partials
    double* partial_host = nullptr;
    int npartial = 1024;
    PathBuffers<double> d;
    PathBuffers<float> f;
    bool ab2_have_history = false;
    // elements per thread of the line FFTs, per kernel class (0 rows_fwd,
    // 1 cols, 2 rows_inv) and per path (0 FP64, 1 FP32); chosen by autotune()
    int ept[2][3] = {{0, 0, 0}, {0, 0, 0}};
    bool tuned[2] = {false, false};
    double tune_us[2][3][3] = {};
    // Monge-Ampère solver state (allocated on first use)
    bool ma_ready = false;
    cpx<double>* ma_phi = nullptr;    // Φ̂ [kx][lev][ky]
    cpx<double>* ma_dprev = nullptr;  // D̂ of the current iterate
    cpx<double>* ma_p[3] = {nullptr, nullptr, nullptr};  // Φxx, Φyy, Φxy planes (y physical)
    cpx<double>* ma_b2 = nullptr;     // forcing spectrum [kx][ky]
    double* ma_coef = nullptr;
    double* ma_partial = nullptr;
    double* ma_partial_host = nullptr;
    int ma_partial_cap = 0;
    int ma_partial_used = 0;
    bool ma_have_d = false;
    // FP32 phase of the mixed-precision solve: Φ̂ and D̂ in FP32; the three
    // derivative planes alias ma_p (an FP64 plane holds two FP32 planes)
    cpx<float>* ma32_phi = nullptr;
    cpx<float>* ma32_d = nullptr;
    Grid g3{};                        // 3D plane geometry: pitch = nz * ny
    // coupled stepping (MA solve at every stage)
    bool coupled = false;
    ssg_ma_params ma_par{};
    double* bst = nullptr;            // stage buoyancy (coupled mode)
    ssg_ma_diag last_diag{};
    double* adj_lam = nullptr;        // adjoint state
    double* adj_acc = nullptr;
    double* adj_s = nullptr;
    // coupled-step adjoint (allocated on first use)
    bool adjc_ready = false;
    double* adjc_pd[3] = {nullptr, nullptr, nullptr};        // Φxx, Φyy, Φxy physical, [lev][y][x]
    cpx<double>* adjc_fbar = nullptr;                        // F̄ [kx][row][ky]
    cpx<double>* adjc_q = nullptr;                           // D'ᵀλ of the current iterate
    cpx<double>* adjc_lam = nullptr;                         // λ planes (y physical)
    cpx<double>* adjc_phibar = nullptr;                      // F(φ̄) [kx][ky]
    cpx<double>* adjc_bbar = nullptr;                        // F(b̄) [kx][ky]
    cpx<double>* adjc_wbar = nullptr;                        // (ū, v̄)
    double* adjc_g = nullptr;
    double* adjc_k1 = nullptr;
    double* adjc_tmp = nullptr;
    double* adjc_bst[3] = {nullptr, nullptr, nullptr};       // stage buoyancies of the step being reversed
    cpx<double>* adjc_phis[3] = {nullptr, nullptr, nullptr}; // their Monge-Ampère solutions
    ssg_ma_diag adjc_diag{};
    struct GraphCache {
        hipGraphExec_t exec = nullptr;
        int len = 0;
        ssg_time_params p{};
    } gcache[2];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int64_t last_launches = 0;
    int64_t launches = 0;
    // per-launch HIP event timing (ssg_profile)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    struct ProfRec {
        int kind;
        double bytes;
        int e0, e1;
    };
    std::vector<ProfRec> prof_recs;
    char err[512] = {0};
};

namespace {

int set_err(ssg_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIP_OK(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return set_err(ctx, SSG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

bool pow2_in_range(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

// run f(std::integral_constant<int, N>) for the power of two n
template <typename Fn>
int with_size(int n, Fn&& fn) {
    switch (n) {
        case 8: return fn(std::integral_constant<int, 8>{});
        case 16: return fn(std::integral_constant<int, 16>{});
        case 32: return fn(std::integral_constant<int, 32>{});
        case 64: return fn(std::integral_constant<int, 64>{});
        case 128: return fn(std::integral_constant<int, 128>{});
        case 256: return fn(std::integral_constant<int, 256>{});
        case 512: return fn(std::integral_constant<int, 512>{});
        case 1024: return fn(std::integral_constant<int, 1024>{});
        case 2048: return fn(std::integral_constant<int, 2048>{});
        case 4096: return fn(std::integral_constant<int, 4096>{});
        default: return SSG_ERR_UNSUPPORTED;
    }
}

// Elements per thread E of a line FFT of length n: candidates 4, 8, 16 with
// E <= n and n/E <= 256 threads per line.  Small lines use min(16, n) only.
bool ept_valid(int n, int e) {
    if (n < 64) return e == (n < 16 ? n : 16);
    return (e == 4 || e == 8 || e == 16) && n / e <= 256;
}
int ept_default(int n) {
    if (n < 64) return n < 16 ? n : 16;
    return n / 8 <= 256 ? 8 : 16;
}

// run fn(integral_constant<N>, integral_constant<E>)
template <typename Fn>
int with_size_ept(int n, int e, Fn&& fn) {
    if (!ept_valid(n, e)) return SSG_ERR_UNSUPPORTED;
    return with_size(n, [&](auto ntag) -> int {
        constexpr int N = decltype(ntag)::value;
        if constexpr (N < 64) {
            return fn(ntag, std::integral_constant<int, (N < 16 ? N : 16)>{});
        } else {
            if (e == 4) {
                if constexpr (N / 4 <= 256) return fn(ntag, std::integral_constant<int, 4>{});
                return SSG_ERR_UNSUPPORTED;
            }
            if (e == 8) {
                if constexpr (N / 8 <= 256) return fn(ntag, std::integral_constant<int, 8>{});
                return SSG_ERR_UNSUPPORTED;
            }
            return fn(ntag, std::integral_constant<int, 16>{});
        }
    });
}

// lines per block: at most 256 threads and kMaxLdsBytes of LDS, a power of
// two that divides `rows`
template <typename R>
int lines_per_block(int n, int e, int rows, int nbuf = 1) {
    const int T = n / e;
    const int le = lds_line_elems(n) * (int)sizeof(cpx<R>) * nbuf;
    int G = 1;
    while (2 * G * T <= 256 && 2 * G * le <= kMaxLdsBytes && rows % (2 * G) == 0) G *= 2;
    return G;
}

int prof_begin(ssg_ctx* c, int* slot);
int prof_end(ssg_ctx* c, int slot, int kind, double bytes);

// Launch on the context stream.  With a profiling slot the dispatch itself
// timestamps the slot's event pair at kernel start and end
// (hipExtLaunchKernelGGL), so the measured time is the kernel's own
// duration, without the event packets of a separate record.
template <typename K, typename... Args>
void launch_on(ssg_ctx* c, int slot, K kernel, dim3 grid, dim3 block, size_t lds, Args... args) {
    if (slot >= 0)
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, c->stream, c->ev_pool[slot],
                              c->ev_pool[slot + 1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, c->stream, args...);
}

template <typename R>
PathBuffers<R>& bufs(ssg_ctx* c);
template <>
PathBuffers<double>& bufs<double>(ssg_ctx* c) { return c->d; }
template <>
PathBuffers<float>& bufs<float>(ssg_ctx* c) { return c->f; }

template <typename R>
int ept_of(ssg_ctx* c, int kind, int n) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    const int e = c->ept[p][kind];
    return (e && ept_valid(n, e)) ? e : ept_default(n);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, typename S, int SRC>
int launch_rows_fwd(ssg_ctx* c, const S* f, const cpx<R>* w, cpx<R>* PA, cpx<R>* PB, int ncols) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_of<R>(c, 0, c->nx);
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int rows_per_line = SRC == SRC_PAIR ? 2 : 1;
        const int G = lines_per_block<R>(N, E, c->ny / rows_per_line);
        const int nblk = c->ny / (rows_per_line * G);
        RowsFwdArgs<R, S> a{f, w, PA, PB, ncols, G, nblk, B.twx, c->g};
        const size_t lds = (size_t)G * lds_line_elems(N) * sizeof(cpx<R>);
        const double n2 = (double)c->nx * c->ny, plane = (double)ncols * c->ny * sizeof(cpx<R>);
        const double bytes = SRC == SRC_FLUX ? n2 * (sizeof(double) + sizeof(cpx<R>)) + 2 * plane
                             : SRC == SRC_W  ? n2 * sizeof(cpx<R>) + 2 * plane
                                             : n2 * sizeof(S) + plane;
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_rows_fwd<R, S, N, E, SRC>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, 0, bytes);
    });
}

// General column-pass launcher: `a` carries the operands; lines are
// (kx, level) pairs when nlev > 1 (kx = line / nlev).
template <typename R, int OP>
int launch_cols_args(ssg_ctx* c, ColsArgs<R> a, int nlines, int nlev) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_of<R>(c, 1, c->ny);
    return with_size_ept(c->ny, ept, [&](auto tag, auto etag) -> int {
        constexpr int M = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = M / E;
        const int le = lds_line_elems(M) * (int)sizeof(cpx<R>) * cols_nbuf(OP);
        int G = 1;
        while (2 * G * T <= 256 && 2 * G * le <= kMaxLdsBytes && 2 * G <= nlines) G *= 2;
        const int nblk = (nlines + G - 1) / G;
        const double kmx = M_PI * c->nx / c->lx, kmy = M_PI * c->ny / c->ly;
        a.ncols = nlines;
        a.G = G;
        a.inv_kmax = (R)(1.0 / std::sqrt(kmx * kmx + kmy * kmy));
        a.twy = B.twy;
        a.g = c->g;
        a.nlev = nlev;
        if ((OP == OP_MA_DIFF || OP == OP_MA_ADJ) && nblk > c->ma_partial_cap)
            return set_err(c, SSG_ERR_ARG, "reduction buffer too small (%d blocks)", nblk);
        const size_t lds = (size_t)G * le;
        const int nin = OP == OP_MA_ADJ ? 3
                        : (OP == OP_TEND || OP == OP_TEND_Q || OP == OP_JAC || OP == OP_MA_DIFF || OP == OP_PHIBAR) ? 2
                                                                                                                  : 1;
        const int nout = (OP == OP_VEL || OP == OP_VEL_SPEC) ? 2 : (OP == OP_MA_DERIV ? 3 : 1);
        const double bytes = (double)(nin + nout) * nlines * c->ny * sizeof(cpx<R>);
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_cols<R, M, E, OP>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        if (OP == OP_MA_DIFF || OP == OP_MA_ADJ) c->ma_partial_used = nblk;
        return prof_end(c, slot, 1, bytes);
    });
}

template <typename R, int OP>
int launch_cols(ssg_ctx* c, const cpx<R>* inA, const cpx<R>* inB, cpx<R>* outA, cpx<R>* outB, int ncols,
                R scale, R fstrength = 0, int forder = 1, R fcutoff = 1, int fdelta = 0) {
    ColsArgs<R> a{};
    a.inA = inA;
    a.inB = inB;
    a.outA = outA;
    a.outB = outB;
    a.scale = scale;
    a.fstrength = fstrength;
    a.fcutoff = fcutoff;
    a.forder = forder;
    a.fdelta = fdelta;
    return launch_cols_args<R, OP>(c, a, ncols, 1);
}

template <typename R, int MODE, int EPI>
int launch_rows_inv(ssg_ctx* c, RowsInvArgs<R> a, int rows = 0, const Grid* grid = nullptr) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_of<R>(c, 2, c->nx);
    if (rows <= 0) rows = c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int rows_per_line = MODE == MODE_PAIR ? 2 : 1;
        const int G = lines_per_block<R>(N, E, rows / rows_per_line, 2);
        const int nblk = rows / (rows_per_line * G);
        a.G = G;
        a.nblk = nblk;
        a.twx = B.twx;
        a.g = grid ? *grid : c->g;
        const size_t lds = (size_t)G * 2 * lds_line_elems(N) * sizeof(cpx<R>);
        const double n2 = (double)c->nx * rows, sc = sizeof(cpx<R>);
        double bytes = (MODE == MODE_PAIR ? 1.0 : 2.0) * a.ncols_in * rows * sc;
        if (EPI == EPI_STORE) bytes += n2 * sizeof(R);
        if (EPI == EPI_STORE_W) bytes += n2 * sc;
        if (EPI == EPI_RK1) bytes += n2 * (sizeof(double) + sizeof(R));
        if (EPI == EPI_RK2) bytes += n2 * sizeof(double);
        if (EPI == EPI_RK3) bytes += n2 * (2 * sizeof(double) + sizeof(R));
        if (EPI == EPI_AB2) bytes += n2 * (2 * sizeof(double) + 2 * sizeof(R));
        if (EPI == EPI_SET) bytes += n2 * sizeof(double);
        if (EPI == EPI_ADD) bytes += n2 * 2 * sizeof(double);
        if (EPI >= EPI_RK1 && a.emit) bytes += n2 * sc + 2.0 * a.ncols_out * rows * sc;
        if (a.bstage) bytes += n2 * sizeof(double);
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_rows_inv<R, N, E, MODE, EPI>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, 2, bytes);
    });
}

// Event pair of one launch when profiling is on (recorded by launch_on).
// `bytes` is the launch's algorithmic traffic: every input read once,
// every output written once.
int prof_begin(ssg_ctx* c, int* slot) {
    *slot = -1;
    if (!c->prof) return SSG_OK;
    const int need = (int)c->prof_recs.size() * 2 + 2;
    while ((int)c->ev_pool.size() < need) {
        hipEvent_t e;
        HIP_OK(c, hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    *slot = (int)c->prof_recs.size() * 2;
    return SSG_OK;
}

int prof_end(ssg_ctx* c, int slot, int kind, double bytes) {
    if (slot < 0) return SSG_OK;
    c->prof_recs.push_back({kind, bytes, slot, slot + 1});
    return SSG_OK;
}

int grid1d(long n) {
    long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------------------
// path pieces
// ---------------------------------------------------------------------------

// surface velocities (src/setICs.jl:161-192) in FP64, and their FP32 copy
int compute_velocities(ssg_ctx* c) {
    PathBuffers<double>& D = c->d;
    int rc = launch_rows_fwd<double, double, SRC_PAIR>(c, c->phi, nullptr, D.Fx, nullptr, c->ncols_full);
    if (rc) return rc;
    rc = launch_cols<double, OP_VEL>(c, D.Fx, nullptr, D.Tp, D.Fy, c->ncols_full, 1.0 / ((double)c->nx * c->ny));
    if (rc) return rc;
    RowsInvArgs<double> a{};
    a.P0 = D.Tp;
    a.P1 = D.Fy;
    a.ncols_in = c->ncols_full;
    a.w_out = D.w;
    rc = launch_rows_inv<double, MODE_PLANES, EPI_STORE_W>(c, a);
    if (rc) return rc;
    if (c->precision == SSG_FP32_MIXED) {
        const long n = (long)c->nx * c->ny;
        hipLaunchKernelGGL((k_w_convert<float>), dim3(grid1d(n)), dim3(256), 0, c->stream, D.w, c->f.w, n);
        c->launches++;
        HIP_OK(c, hipGetLastError());
    }
    return SSG_OK;
}

// fluxes of the current b into (Fx, Fy) planes
template <typename R>
int flux_forward(ssg_ctx* c, int ncols, const double* src = nullptr) {
    PathBuffers<R>& B = bufs<R>(c);
    return launch_rows_fwd<R, double, SRC_FLUX>(c, src ? src : c->b, B.w, B.Fx, B.Fy, ncols);
}

template <typename R, int EPI>
int stage_inverse(ssg_ctx* c, double c1, int emit, int first, double* bstage = nullptr) {
    PathBuffers<R>& B = bufs<R>(c);
    RowsInvArgs<R> a{};
    a.bstage = bstage;
    a.P0 = B.Tp;
    a.ncols_in = c->ncols_tend;
    a.b = c->b;
    a.k1 = B.k1;
    a.fold = B.fold;
    a.c1 = c1;
    a.first = first;
    a.emit = emit;
    a.w = B.w;
    a.Fx = B.Fx;
    a.Fy = B.Fy;
    a.ncols_out = c->ncols_tend;
    return launch_rows_inv<R, MODE_PAIR, EPI>(c, a);
}

template <typename R>
int tendency_cols(ssg_ctx* c, int quirk) {
    PathBuffers<R>& B = bufs<R>(c);
    const R scale = (R)(1.0 / ((double)c->nx * c->ny));
    if (quirk) return launch_cols<R, OP_TEND_Q>(c, B.Fx, B.Fy, B.Tp, nullptr, c->ncols_tend, scale);
    return launch_cols<R, OP_TEND>(c, B.Fx, B.Fy, B.Tp, nullptr, c->ncols_tend, scale);
}

template <typename R>
int run_kind(ssg_ctx* c, int kind) {
    if (kind == 0) return flux_forward<R>(c, c->ncols_tend);
    if (kind == 1) return tendency_cols<R>(c, 0);
    return stage_inverse<R, EPI_RK2>(c, 0.0, 1, 0);
}

// apply_spectral_filter! on b; FP64 sets b = irfft(σ b̂); FP32-mixed adds
// the FP32 defect irfft((σ-1) b̂) to the FP64 b.  emit: refresh fluxes.
template <typename R>
int filter_b(ssg_ctx* c, double strength, int order, double cutoff, int emit, double* target = nullptr) {
    PathBuffers<R>& B = bufs<R>(c);
    double* tb = target ? target : c->b;
    int rc = launch_rows_fwd<R, double, SRC_PAIR>(c, tb, nullptr, B.Tp, nullptr, c->ncols_full);
    if (rc) return rc;
    const int delta = std::is_same<R, float>::value ? 1 : 0;
    rc = launch_cols<R, OP_FILT>(c, B.Tp, nullptr, B.Tp, nullptr, c->ncols_full, (R)(1.0 / ((double)c->nx * c->ny)),
                                 (R)strength, order, (R)cutoff, delta);
    if (rc) return rc;
    RowsInvArgs<R> a{};
    a.P0 = B.Tp;
    a.ncols_in = c->ncols_full;
    a.b = tb;
    a.emit = emit;
    a.w = B.w;
    a.Fx = B.Fx;
    a.Fy = B.Fy;
    a.ncols_out = c->ncols_tend;
    if (delta) return launch_rows_inv<R, MODE_PAIR, EPI_ADD>(c, a);
    return launch_rows_inv<R, MODE_PAIR, EPI_SET>(c, a);
}

// Autotuner: for each kernel class time every valid elements-per-thread
// candidate on this context's buffers (3 warm-up + 8 timed launches) and keep
// the fastest.  The probed launches only write scratch planes, never the
// prognostic state.  SSG_EPT="e_fwd,e_cols,e_inv" forces a choice and
// SSG_TUNE=0 keeps the defaults.
template <typename R>
int run_kind(ssg_ctx* c, int kind);

// Tuning results are shared by every context of the process with the same
// grid, path and device, so that contexts agree on launch shapes and
// therefore on roundoff (bitwise-reproducible results within a process;
// SSG_EPT pins the shapes across processes).
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
    for (int kind = 0; kind < 3 && rc == SSG_OK; ++kind) {
        const int n = kind == 1 ? c->ny : c->nx;
        int best = 0;
        double best_us = 1e30;
        for (int ci = 0; ci < 3; ++ci) {
            const int e = cands[ci];
            c->tune_us[p][kind][ci] = -1.0;
            if (!ept_valid(n, e)) continue;
            c->ept[p][kind] = e;
            for (int i = 0; i < 3 && rc == SSG_OK; ++i) rc = run_kind<R>(c, kind);
            if (rc) break;
            if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { rc = SSG_ERR_HIP; break; }
            for (int i = 0; i < 8 && rc == SSG_OK; ++i) rc = run_kind<R>(c, kind);
            if (rc) break;
            float ms = 0;
            if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
                rc = set_err(c, SSG_ERR_HIP, "autotune timing failed");
                break;
            }
            const double us = 1e3 * ms / 8.0;
            c->tune_us[p][kind][ci] = us;
            if (us < best_us) {
                best_us = us;
                best = e;
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
int rfft2_device(ssg_ctx* c, const double* src, cpx<double>* dst = nullptr);

// Coupled step: the Monge-Ampère problem is re-solved for the stage
// buoyancy before every tendency evaluation, so φₛ and (u, v) follow b.
template <typename R>
int run_steps_coupled(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    const double dt = p->dt;
    ssg_ma_diag d{};
    int rc;
    auto stage_tendency = [&](const double* bsrc) -> int {
        int r = ma_solve_dev(c, bsrc, &c->ma_par, &d);
        if (r) return r;
        if (!d.converged)
            return set_err(c, SSG_ERR_NOT_CONVERGED, "Monge-Ampere solve did not converge (rel. residual %.3e)",
                           d.relative_residual);
        if ((r = flux_forward<R>(c, c->ncols_tend, bsrc))) return r;
        return tendency_cols<R>(c, p->scheme == SSG_AB2_LOWSTORAGE ? p->ab2_reference_aliasing : 0);
    };
    for (int64_t it = 0; it < nsteps; ++it) {
        const int64_t step_after = s->step + 1;
        const bool filt = p->filter_freq > 0 && (step_after % p->filter_freq) == 0;
        if (p->scheme == SSG_RK3) {
            if ((rc = stage_tendency(c->b))) return rc;
            if ((rc = stage_inverse<R, EPI_RK1>(c, dt / 3.0, 0, 0, c->bst))) return rc;
            if ((rc = stage_tendency(c->bst))) return rc;
            if ((rc = stage_inverse<R, EPI_RK2>(c, 2.0 * dt / 3.0, 0, 0, c->bst))) return rc;
            if ((rc = stage_tendency(c->bst))) return rc;
            if ((rc = stage_inverse<R, EPI_RK3>(c, dt / 4.0, 0, 0))) return rc;
        } else {
            if ((rc = stage_tendency(c->b))) return rc;
            const int first = (s->step == 0) ? 1 : 0;
            if ((rc = stage_inverse<R, EPI_AB2>(c, dt, 0, first))) return rc;
            c->ab2_have_history = true;
        }
        s->t += dt;
        s->step = step_after;
        s->dt_actual = dt;
        if (filt && (rc = filter_b<R>(c, p->filter_strength, p->filter_order, p->filter_cutoff, 0))) return rc;
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

bool same_params(const ssg_time_params& a, const ssg_time_params& b) {
    return a.dt == b.dt && a.scheme == b.scheme && a.filter_freq == b.filter_freq &&
           a.filter_strength == b.filter_strength && a.filter_order == b.filter_order &&
           a.filter_cutoff == b.filter_cutoff && a.ab2_reference_aliasing == b.ab2_reference_aliasing;
}

// A hipGraph of one filter cycle (filter_freq steps, the filter on the last;
// 16 steps when the filter is off), captured on first use and replayed for
// every aligned cycle.  Kernel arguments are fixed by the parameters, so the
// graph is rebuilt when they change.
template <typename R>
int cycle_graph(ssg_ctx* c, const ssg_time_params* p, int len, int64_t step0, hipGraphExec_t* out) {
    const int path = std::is_same<R, float>::value ? 1 : 0;
    auto& gc = c->gcache[path];
    if (gc.exec && gc.len == len && same_params(gc.p, *p)) {
        *out = gc.exec;
        return SSG_OK;
    }
    if (gc.exec) {
        (void)hipGraphExecDestroy(gc.exec);
        gc.exec = nullptr;
    }
    HIP_OK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = SSG_OK;
    for (int i = 0; i < len && rc == SSG_OK; ++i) rc = issue_step<R>(c, p, step0 + i, false);
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
    gc.exec = ex;
    gc.len = len;
    gc.p = *p;
    *out = ex;
    return SSG_OK;
}

template <typename R>
int run_steps(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    if (c->coupled) return run_steps_coupled<R>(c, p, s, nsteps);
    int rc = flux_forward<R>(c, c->ncols_tend);
    if (rc) return rc;
    const char* genv = getenv("SSG_GRAPH");
    const bool graphs = !(genv && genv[0] == '0') && !c->prof;
    const int F = p->filter_freq > 0 ? p->filter_freq : 16;
    int64_t it = 0;
    while (it < nsteps) {
        const int64_t remaining = nsteps - it;
        const bool aligned = p->filter_freq <= 0 || (s->step % F) == 0;
        const bool ab2_first = p->scheme == SSG_AB2_LOWSTORAGE && s->step == 0;
        int adv = 1;
        if (graphs && aligned && !ab2_first && remaining > F) {
            hipGraphExec_t ex = nullptr;
            if ((rc = cycle_graph<R>(c, p, F, s->step, &ex))) return rc;
            HIP_OK(c, hipGraphLaunch(ex, c->stream));
            c->launches += (p->scheme == SSG_RK3 ? 6 : 2) * F + (p->filter_freq > 0 ? 3 : 0);
            adv = F;
        } else {
            if ((rc = issue_step<R>(c, p, s->step, it == nsteps - 1))) return rc;
        }
        for (int i = 0; i < adv; ++i) {
            s->t += p->dt;
            s->step += 1;
        }
        s->dt_actual = p->dt;
        if (p->scheme == SSG_AB2_LOWSTORAGE) c->ab2_have_history = true;
        it += adv;
    }
    return SSG_OK;
}


// ---------------------------------------------------------------------------
// Monge-Ampère solve (FP64)
// ---------------------------------------------------------------------------
int ma_alloc(ssg_ctx* c) {
    if (c->ma_ready) return SSG_OK;
    if (c->nz < 3 || c->nz > 128)
        return set_err(c, SSG_ERR_UNSUPPORTED, "the Monge-Ampere solve needs 3 <= nz <= 128 (nz = %d)", c->nz);
    const size_t c3 = (size_t)c->ncols_full * c->nz * c->ny;
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
    HIP_OK(c, hipHostMalloc(&c->ma_partial_host, (size_t)c->ma_partial_cap * 2 * sizeof(double), hipHostMallocDefault));
    HIP_OK(c, hipMalloc(&c->bst, (size_t)c->nx * c->ny * sizeof(double)));
    // stencil coefficients (uniform grid dz of the context)
    const int nz = c->nz;
    std::vector<double> dz(nz, c->lz / (nz - 1));
    std::vector<double> cf(3 * nz + 6, 0.0);
    for (int k = 1; k <= nz - 2; ++k) {  // d2dz2!, src/transforms.jl:812-819
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
    HIP_OK(c, hipMalloc(&c->ma_coef, cf.size() * sizeof(double)));
    HIP_OK(c, hipMemcpy(c->ma_coef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice));
    c->g3 = c->g;
    c->g3.pitch = c->nz * c->ny;
    c->ma_ready = true;
    return SSG_OK;
}

void ma_free(ssg_ctx* c) {
    (void)hipFree(c->ma_phi);
    (void)hipFree(c->ma_dprev);
    for (int i = 0; i < 3; ++i) (void)hipFree(c->ma_p[i]);
    (void)hipFree(c->ma_b2);
    (void)hipFree(c->ma32_phi);
    (void)hipFree(c->ma32_d);
    (void)hipFree(c->ma_coef);
    (void)hipFree(c->ma_partial);
    if (c->ma_partial_host) (void)hipHostFree(c->ma_partial_host);
    (void)hipFree(c->bst);
}

// Φ̂ = L⁻¹(b̂ + ε D̂) per wavenumber (D̂ term omitted when use_d = 0)
template <typename R, typename P>
int ma_tridiag(ssg_ctx* c, const ssg_ma_params* p, int use_d, const cpx<R>* Dh, cpx<P>* Phi) {
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
    a.g = c->g;
    const long total = (long)c->ncols_full * c->ny;
    const int nblk = (int)((total + 127) / 128);
    const double bytes = (double)total * (c->nz * (sizeof(cpx<R>) + sizeof(cpx<P>)) + sizeof(cpx<double>));
    int slot, rc;
    if ((rc = prof_begin(c, &slot))) return rc;
    if (c->nz <= 8)
        launch_on(c, slot, k_ma_tridiag<R, P, 8>, dim3(nblk), dim3(128), 0, a);
    else if (c->nz <= 16)
        launch_on(c, slot, k_ma_tridiag<R, P, 16>, dim3(nblk), dim3(128), 0, a);
    else if (c->nz <= 32)
        launch_on(c, slot, k_ma_tridiag<R, P, 32>, dim3(nblk), dim3(128), 0, a);
    else if (c->nz <= 64)
        launch_on(c, slot, k_ma_tridiag<R, P, 64>, dim3(nblk), dim3(128), 0, a);
    else
        launch_on(c, slot, k_ma_tridiag<R, P, 128>, dim3(nblk), dim3(128), 0, a);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return prof_end(c, slot, 4, bytes);
}

// x-inverse of the three derivative planes, D = Φxx Φyy - Φxy², x-FFT of D
// into planes[0]
template <typename R>
int ma_rows(ssg_ctx* c, cpx<R>* const (&planes)[3]) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_of<R>(c, 2, c->nx);
    const int rows = c->nz * c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<R>(N, E, rows / 2, 3);
        const int nblk = rows / (2 * G);
        MaRowsArgs<R> a{planes[0], planes[1], planes[2], planes[0], {nullptr, nullptr, nullptr}, c->ncols_full, G, nblk,
                        B.twx, c->g3};
        const size_t lds = (size_t)G * 3 * lds_line_elems(N) * sizeof(cpx<R>);
        const double bytes = 4.0 * c->ncols_full * rows * sizeof(cpx<R>);
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_ma_rows<R, N, E>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, 3, bytes);
    });
}

template <typename A, typename B>
int cpx_convert(ssg_ctx* c, const cpx<A>* in, cpx<B>* out, size_t n) {
    hipLaunchKernelGGL((k_cpx_convert<A, B>), dim3(grid1d((long)n)), dim3(256), 0, c->stream, in, out, (long)n);
    c->launches++;
    HIP_OK(c, hipGetLastError());
    return SSG_OK;
}

// Residual evaluation of the iterate Φ̂ = L⁻¹(b̂ + ε D̂_prev): D̂ of Φ into
// `dh` (which holds D̂_prev on entry) and the residual ε(D̂_prev - D̂) on
// interior levels, reduced over blocks and read back.  `rabs` is
// norm_field(r) (src/poisson.jl:986-995) via Parseval, `rrel` is
// ‖r‖ / ‖b + εD‖.
template <typename R>
int ma_eval(ssg_ctx* c, const ssg_ma_params* p, const cpx<R>* phi, cpx<R>* const (&planes)[3], cpx<R>* dh,
            double* rabs, double* rrel) {
    const double scale = 1.0 / ((double)c->nx * c->ny);
    const int lines = c->ncols_full * c->nz;
    ColsArgs<R> a{};
    a.inA = phi;
    a.outA = planes[0];
    a.outB = planes[1];
    a.outC = planes[2];
    a.scale = (R)scale;
    int rc;
    if ((rc = launch_cols_args<R, OP_MA_DERIV>(c, a, lines, c->nz))) return rc;
    if ((rc = ma_rows<R>(c, planes))) return rc;
    ColsArgs<R> b{};
    b.inA = planes[0];
    b.outA = dh;
    b.B2 = c->ma_b2;
    b.eps = p->eps;
    b.body = p->variant == SSG_MA_BODY ? 1 : 0;
    b.partial = c->ma_partial;
    if ((rc = launch_cols_args<R, OP_MA_DIFF>(c, b, lines, c->nz))) return rc;
    const int nb = c->ma_partial_used;
    HIP_OK(c, hipMemcpyAsync(c->ma_partial_host, c->ma_partial, (size_t)nb * 2 * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    double r2 = 0, h2 = 0;
    for (int i = 0; i < nb; ++i) {
        r2 += c->ma_partial_host[2 * i];
        h2 += c->ma_partial_host[2 * i + 1];
    }
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
    const bool warm = p->warm && c->ma_have_d;
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
    if (c->precision == SSG_FP32_MIXED) {
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
            c->ma_have_d = false;
            return SSG_OK;
        }
    } else {
        if (!warm) HIP_OK(c, hipMemsetAsync(c->ma_dprev, 0, c3 * sizeof(cpx<double>), c->stream));
        if ((rc = ma_tridiag<double, double>(c, p, warm ? 1 : 0, c->ma_dprev, c->ma_phi))) return rc;
    }
    cpx<double>* const pl[3] = {c->ma_p[0], c->ma_p[1], c->ma_p[2]};
    for (;;) {
        if ((rc = ma_eval<double>(c, p, c->ma_phi, pl, c->ma_dprev, &ra, &rr))) return rc;
        if (record(true)) break;
        if ((rc = ma_tridiag<double, double>(c, p, 1, c->ma_dprev, c->ma_phi))) return rc;
    }
    c->ma_have_d = true;
    return SSG_OK;
}

// φₛ = Φ[:,:,Nz] (src/poisson.jl:1119-1131) and u = -∂yφₛ, v = ∂xφₛ
// (src/setICs.jl:161-192), taken from the top level of Φ̂
int ma_surface_outputs(ssg_ctx* c) {
    PathBuffers<double>& D = c->d;
    const size_t lb = (size_t)c->ny * sizeof(cpx<double>);
    HIP_OK(c, hipMemcpy2DAsync(D.Tp, lb, c->ma_phi + (size_t)(c->nz - 1) * c->ny, lb * c->nz, lb, c->ncols_full,
                               hipMemcpyDeviceToDevice, c->stream));
    const double scale = 1.0 / ((double)c->nx * c->ny);
    ColsArgs<double> a{};
    a.inA = D.Tp;
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
    if ((rc = launch_cols<double, OP_INV>(c, D.Tp, nullptr, D.Tp, nullptr, c->ncols_full, scale))) return rc;
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
    return SSG_OK;
}

// full solve for the buoyancy `bs` (device, FP64)
int ma_solve_dev(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d) {
    int rc = ma_alloc(c);
    if (rc) return rc;
    if ((rc = rfft2_device(c, bs, c->ma_b2))) return rc;
    if ((rc = ma_iterate(c, p, d))) return rc;
    if ((rc = ma_surface_outputs(c))) return rc;
    c->last_diag = *d;
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

int lincomb(ssg_ctx* c, double* out, double ca, const double* x, double cb, const double* y) {
    const long n = (long)c->nx * c->ny;
    hipLaunchKernelGGL(k_lincomb, dim3(grid1d(n)), dim3(256), 0, c->stream, out, ca, x, cb, y, n);
    c->launches++;
    HIP_OK(c, hipGetLastError());
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
    int rc = launch_cols_args<double, OP_MA_DERIV>(c, a, lines, c->nz);
    if (rc) return rc;
    const int ept = ept_of<double>(c, 2, c->nx);
    const int rows = c->nz * c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int G = lines_per_block<double>(N, E, rows / 2, 3);
        const int nblk = rows / (2 * G);
        MaRowsArgs<double> m{c->ma_p[0], c->ma_p[1], c->ma_p[2], nullptr,
                             {c->adjc_pd[0], c->adjc_pd[1], c->adjc_pd[2]}, c->ncols_full, G, nblk, D.twx, c->g3};
        const size_t lds = (size_t)G * 3 * lds_line_elems(N) * sizeof(cpx<double>);
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
// adjc_bbar = F(b̄) out.
int ma_adjoint_solve(ssg_ctx* c, const ssg_ma_params* p, ssg_ma_diag* d) {
    const double scale = 1.0 / ((double)c->nx * c->ny);
    const int lines = c->ncols_full * c->nz;
    const size_t c3 = (size_t)lines * c->ny;
    memset(d, 0, sizeof(*d));
    HIP_OK(c, hipMemsetAsync(c->adjc_q, 0, c3 * sizeof(cpx<double>), c->stream));
    int rc = ma_tridiag_adj(c, p, 0);
    if (rc) return rc;
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
        const int nb = c->ma_partial_used;
        HIP_OK(c, hipMemcpyAsync(c->ma_partial_host, c->ma_partial, (size_t)nb * 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_OK(c, hipStreamSynchronize(c->stream));
        double r2 = 0, h2 = 0;
        for (int i = 0; i < nb; ++i) {
            r2 += c->ma_partial_host[2 * i];
            h2 += c->ma_partial_host[2 * i + 1];
        }
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
    if (!d->converged)
        return set_err(c, SSG_ERR_NOT_CONVERGED, "adjoint Monge-Ampere solve did not converge (rel. residual %.3e)",
                       d->relative_residual);
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
            return set_err(c, SSG_ERR_NOT_CONVERGED, "Monge-Ampere solve did not converge (rel. residual %.3e)",
                           d.relative_residual);
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
    if ((rc = ma_adjoint_solve(c, mp, &d))) return rc;
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

// checkpointed adjoint of nsteps coupled steps from the context's current b
// (at step number step0): checkpoints every `every` steps on the forward
// pass, then each segment is recomputed from its checkpoint, storing the
// per-step states, and swept in reverse.
int adjoint_coupled(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step0, int64_t nsteps,
                    int every) {
    const size_t n2 = (size_t)c->nx * c->ny, bytes = n2 * sizeof(double);
    const int64_t nck = (nsteps + every - 1) / every;
    double* ck = nullptr;
    double* seg = nullptr;
    HIP_OK(c, hipMalloc(&ck, (size_t)nck * bytes));
    if (hipMalloc(&seg, (size_t)every * bytes) != hipSuccess) {
        (void)hipFree(ck);
        return set_err(c, SSG_ERR_HIP, "checkpoint allocation failed");
    }
    auto body = [&]() -> int {
        int rc;
        // forward pass with checkpoints; c->b is restored at the end
        for (int64_t j = 0; j < nsteps; ++j) {
            if (j % every == 0)
                HIP_OK(c, hipMemcpyAsync(ck + (size_t)(j / every) * n2, c->b, bytes, hipMemcpyDeviceToDevice,
                                         c->stream));
            if ((rc = coupled_step_record(c, p, mp, step0 + j + 1))) return rc;
        }
        for (int64_t sg = nck - 1; sg >= 0; --sg) {
            const int64_t j0 = sg * every;
            const int64_t len = std::min<int64_t>(every, nsteps - j0);
            // recompute the segment's per-step states from its checkpoint
            HIP_OK(c, hipMemcpyAsync(c->b, ck + (size_t)sg * n2, bytes, hipMemcpyDeviceToDevice, c->stream));
            for (int64_t k = 0; k < len; ++k) {
                HIP_OK(c, hipMemcpyAsync(seg + (size_t)k * n2, c->b, bytes, hipMemcpyDeviceToDevice, c->stream));
                if (k + 1 < len && (rc = coupled_step_record(c, p, mp, step0 + j0 + k + 1))) return rc;
            }
            auto state_at = [&](int64_t jj) -> int {
                HIP_OK(c, hipMemcpyAsync(c->b, seg + (size_t)jj * n2, bytes, hipMemcpyDeviceToDevice, c->stream));
                return SSG_OK;
            };
            if ((rc = adjoint_coupled_sweep(c, p, mp, step0 + j0, len, state_at))) return rc;
        }
        // leave b at the start state
        HIP_OK(c, hipMemcpyAsync(c->b, ck, bytes, hipMemcpyDeviceToDevice, c->stream));
        return SSG_OK;
    };
    const int rc = body();
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(ck);
    (void)hipFree(seg);
    return rc;
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

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
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
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SSG_OK;
}

const char* ssg_last_error(const ssg_ctx* c) { return c ? c->err : "null context"; }

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
    return field_io(c, field, host, nullptr, n, false);
}
int ssg_get_field(ssg_ctx* c, int32_t field, double* host, int64_t n) {
    if (!host) return set_err(c, SSG_ERR_ARG, "null pointer");
    return field_io(c, field, nullptr, host, n, false);
}
int ssg_set_field_device(ssg_ctx* c, int32_t field, const double* dev, int64_t n) {
    if (!dev) return set_err(c, SSG_ERR_ARG, "null pointer");
    return field_io(c, field, dev, nullptr, n, true);
}
int ssg_get_field_device(ssg_ctx* c, int32_t field, double* dev, int64_t n) {
    if (!dev) return set_err(c, SSG_ERR_ARG, "null pointer");
    return field_io(c, field, nullptr, dev, n, true);
}

int ssg_timestep(ssg_ctx* c, const ssg_time_params* p, ssg_time_state* s, int64_t nsteps) {
    if (!c) return SSG_ERR_ARG;
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
    if (nsteps == 0) return SSG_OK;
    {
        int trc = c->precision == SSG_FP32_MIXED ? autotune<float>(c) : autotune<double>(c);
        if (trc) return trc;
    }
    const int64_t l0 = c->launches;
    HIP_OK(c, hipEventRecord(c->ev0, c->stream));
    int rc = c->precision == SSG_FP32_MIXED ? run_steps<float>(c, p, s, nsteps) : run_steps<double>(c, p, s, nsteps);
    if (rc) return rc;
    HIP_OK(c, hipEventRecord(c->ev1, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
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

int ssg_ma_solve(ssg_ctx* c, const ssg_ma_params* p, ssg_ma_diag* d) {
    if (!c) return SSG_ERR_ARG;
    int rc = ma_check_params(c, p);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    ssg_ma_diag local{};
    if (!d) d = &local;
    HIP_OK(c, hipEventRecord(c->ev0, c->stream));
    if ((rc = ma_solve_dev(c, c->b, p, d))) return rc;
    HIP_OK(c, hipEventRecord(c->ev1, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_OK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    d->solve_ms = ms;
    c->last_diag = *d;
    if (!d->converged)
        return set_err(c, SSG_ERR_NOT_CONVERGED, "not converged after %d iterations (rel. residual %.3e)", d->iterations,
                       d->relative_residual);
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

int ssg_set_coupling(ssg_ctx* c, const ssg_ma_params* p) {
    if (!c) return SSG_ERR_ARG;
    if (!p) {
        c->coupled = false;
        return SSG_OK;
    }
    int rc = ma_check_params(c, p);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    if ((rc = ma_alloc(c))) return rc;
    c->ma_par = *p;
    c->coupled = true;
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

int ssg_adjoint(ssg_ctx* c, const ssg_time_params* p, int64_t step_end, int64_t nsteps, const double* lam_in,
                double* lam_out) {
    if (!c) return SSG_ERR_ARG;
    if (!p || !lam_in || !lam_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (p->scheme != SSG_RK3) return set_err(c, SSG_ERR_UNSUPPORTED, "the adjoint covers the RK3 scheme");
    if (c->coupled) return set_err(c, SSG_ERR_UNSUPPORTED, "the adjoint covers the frozen-φₛ step");
    if (nsteps < 0 || step_end < nsteps) return set_err(c, SSG_ERR_ARG, "need 0 <= nsteps <= step_end");
    if (!(p->dt > 0)) return set_err(c, SSG_ERR_ARG, "dt must be > 0");
    HIP_OK(c, hipSetDevice(c->device));
    int rc = adj_alloc(c);
    if (rc) return rc;
    const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(c->adj_lam, lam_in, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = adjoint_steps(c, p, step_end, nsteps))) return rc;
    HIP_OK(c, hipMemcpyAsync(lam_out, c->adj_lam, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(c, hipStreamSynchronize(c->stream));
    return SSG_OK;
}

int ssg_adjoint_coupled(ssg_ctx* c, const ssg_time_params* p, const ssg_ma_params* mp, int64_t step_start,
                        int64_t nsteps, int32_t checkpoint_every, const double* lam_in, double* lam_out) {
    if (!c) return SSG_ERR_ARG;
    if (!p || !mp || !lam_in || !lam_out) return set_err(c, SSG_ERR_ARG, "null pointer");
    if (p->scheme != SSG_RK3) return set_err(c, SSG_ERR_UNSUPPORTED, "the adjoint covers the RK3 scheme");
    if (nsteps < 0 || step_start < 0) return set_err(c, SSG_ERR_ARG, "need nsteps >= 0 and step_start >= 0");
    if (checkpoint_every < 1) return set_err(c, SSG_ERR_ARG, "checkpoint_every must be >= 1");
    if (!(p->dt > 0)) return set_err(c, SSG_ERR_ARG, "dt must be > 0");
    int rc = ma_check_params(c, mp);
    if (rc) return rc;
    HIP_OK(c, hipSetDevice(c->device));
    if ((rc = adjc_alloc(c))) return rc;
    const size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    HIP_OK(c, hipMemcpyAsync(c->adj_lam, lam_in, bytes, hipMemcpyHostToDevice, c->stream));
    if (nsteps > 0 && (rc = adjoint_coupled(c, p, mp, step_start, nsteps, checkpoint_every))) return rc;
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
    rc = ma_adjoint_solve(c, mp, &d);
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
    return SSG_OK;
}

}  // extern "C"

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
    void* d_rows[4] = {nullptr, nullptr, nullptr, nullptr};
    int2* d_wcols = nullptr;
    int2* d_wrows = nullptr;
    int ncb = 0, nrb = 0;
    size_t lds_cols = 0, lds_rows = 0;
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
    e->d_cols = nullptr;
    e->d_wcols = e->d_wrows = nullptr;
    e->built_ept = 0;
}

// per-member argument structs and the block tables for elements-per-thread E
template <typename R>
int ens_build(ssg_ensemble* e, int E, const double* dt) {
    ens_free_tables(e);
    std::vector<ColsArgs<R>> ca(e->n);
    std::vector<RowsInvArgs<R>> ra[4];
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
        const double coef[4] = {dt[i] / 3.0, 2.0 * dt[i] / 3.0, dt[i] / 4.0, dt[i] / 4.0};
        for (int k = 0; k < 4; ++k) {
            RowsInvArgs<R> r{};
            r.P0 = B.Tp;
            r.ncols_in = c->ncols_tend;
            r.b = c->b;
            r.k1 = B.k1;
            r.fold = B.fold;
            r.c1 = coef[k];
            r.emit = k == 3 ? 0 : 1;
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
    }
    ENS_OK(e, hipMalloc(&e->d_cols, ca.size() * sizeof(ColsArgs<R>)));
    ENS_OK(e, hipMemcpy(e->d_cols, ca.data(), ca.size() * sizeof(ColsArgs<R>), hipMemcpyHostToDevice));
    for (int k = 0; k < 4; ++k) {
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
    int rc;
    if ((rc = ens_stage<R>(e, 0))) return rc;
    if ((rc = ens_stage<R>(e, 1))) return rc;
    if ((rc = ens_stage<R>(e, (!filt && !last) ? 2 : 3))) return rc;
    if (filt) {
        for (ssg_ctx* c : e->m) {
            if ((rc = filter_b<R>(c, p->filter_strength, p->filter_order, p->filter_cutoff, last ? 0 : 1)))
                return ens_err(e, rc, "member filter: %s", c->err);
        }
    }
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
int ens_tune(ssg_ensemble* e, const double* dt) {
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
        if ((rc = ens_build<R>(e, cand[i], dt))) {
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
    if ((rc = ens_tune<R>(e, dt))) return rc;
    const bool same_dt = e->built_dt.size() == (size_t)e->n && std::equal(dt, dt + e->n, e->built_dt.begin());
    if (e->built_ept != e->ept || !same_dt) {
        if (e->graph) {
            (void)hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        if ((rc = ens_build<R>(e, e->ept, dt))) return rc;
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
        if (graphs && (p->filter_freq <= 0 || step % F == 0) && remaining > F) {
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
    if (p->scheme != SSG_RK3) return ens_err(e, SSG_ERR_UNSUPPORTED, "the ensemble step covers RK3");
    if (nsteps < 0) return ens_err(e, SSG_ERR_ARG, "nsteps < 0");
    for (int i = 0; i < e->n; ++i) {
        if (!(dt[i] > 0) || !std::isfinite(dt[i])) return ens_err(e, SSG_ERR_ARG, "dt[%d] must be finite and > 0", i);
        if (st[i].step != st[0].step) return ens_err(e, SSG_ERR_ARG, "members must be at the same step number");
        if (e->m[i]->coupled) return ens_err(e, SSG_ERR_UNSUPPORTED, "member %d is in coupled mode", i);
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
