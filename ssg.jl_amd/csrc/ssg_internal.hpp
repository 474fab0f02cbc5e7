// Internal declarations shared by the translation units of libssg_hip.so:
// the context, launch helpers and the templates every part instantiates.
#pragma once
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

namespace ssg_impl {

constexpr int kMaxLdsBytes = 64 * 1024;

// Ping-pong line buffers for the tendency, filter and transform passes
// (LineFFT's PPREQ): one block barrier per FFT stage instead of two, for
// twice the LDS.  Used where a line spans several wavefronts and the doubled
// LDS of the widest block (256 threads) stays within kMaxPingPongLds, so two
// blocks still share a CU.  The arithmetic is unchanged: results are
// bitwise the same either way.  Build with -DSSG_FFT_PINGPONG=0 to turn it
// off.
#ifndef SSG_FFT_PINGPONG
#define SSG_FFT_PINGPONG 1
#endif
constexpr int kMaxPingPongLds = 80 * 1024;
template <typename R, int N, int E, int NBUF>
constexpr bool pingpong_fits() {
    return SSG_FFT_PINGPONG && (N / E) > 64 &&
           2L * (256 / (N / E)) * NBUF * lds_line_elems(N) * (long)sizeof(cpx<R>) <= kMaxPingPongLds;
}

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

}  // namespace ssg_impl

using namespace ssg_impl;

// The iterates of the last converged solves of one fixed-point iteration, for
// extrapolated warm starts (ring[0] newest).  While the ring is non-empty
// ring[0] holds the last iterate and the solver's working buffer is scratch;
// otherwise the working buffer holds it.  Buffers rotate between the ring,
// the working buffer and the spares; no iterate is ever copied.
template <typename C>
struct SolveHistoryT {
    C* ring[9] = {};
    int n = 0;
    C* spare[10] = {};
    int nspare = 0;
    int period = -1;  // the solves-per-step pattern under which the entries were pushed
};
using SolveHistory = SolveHistoryT<cpx<double>>;
using SolveHistory32 = SolveHistoryT<cpx<float>>;  // FP32 stage solves of the coupled step

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
    MaSum* ma_sum_host = nullptr;     // pinned: reduced residual sums written by the device
    MaSum* ma_sum_dev = nullptr;      // device address of ma_sum_host
    int ma_stream = 0;                // nontemporal access to the C3 planes
    bool ma_ext_fused = true;         // the first solve of a warm start forms its extrapolation (k_ma_tridiag)
    int ma_partial_cap = 0;
    int ma_partial_used = 0;
    bool ma_have_d = false;
    bool ma_phi_current = false;      // coupled stepping: φₛ, u, v were solved for the current b with ma_par
    bool ma_have_phi = false;         // ma_phi holds the last solve's Φ̂
    bool ma_phis_from_phi = false;    // φₛ and the velocities are the top level of ma_phi (not set since)
    SolveHistory ma_hist;             // D̂ of the last converged solves (working buffer: ma_dprev)
    int ma_period = 0;                // solves per step of the running coupled stepper (RK3 3, AB2 1), else 0
    double ma_last_r0 = 1e300;        // first relative residual of the last solve
    // FP32 phase of the mixed-precision solve: Φ̂ and D̂ in FP32; the three
    // derivative planes alias ma_p (an FP64 plane holds two FP32 planes)
    cpx<float>* ma32_phi = nullptr;
    cpx<float>* ma32_d = nullptr;
    // FP32 stage solves of the coupled step (ssg_ma_params.stage_fp32): the
    // history of their D̂ (working buffer: ma32_d), and whether ma32_d or
    // that history holds a warm start
    SolveHistory32 ma_hist32;
    bool ma_have_d32 = false;
    // order of their same-stage warm start: 2 (linear, the default) or 3
    // (quadratic); SSG_MA_EXTRAP32=3 in the environment selects the quadratic
    // when a context first solves in FP32.  0: not chosen yet
    int ma_extrap32 = 0;
    int ma_ept[2] = {0, 0};           // elements per thread of k_ma_rows per path (0 FP64, 1 FP32)
    int ma_cols_ept[2] = {0, 0};      // ... of the Monge-Ampère column passes (0: the tendency's choice)
    Grid g3{};                        // 3D plane geometry: pitch = nz * ny
    std::vector<double> dz;           // vertical spacings (Nz entries) of a custom grid; empty: uniform Lz/(Nz-1)
    // vertical velocity (allocated on first use): Φxx, Φyy, Φxy per level,
    // (Q₁, Q₂) per level, the surface pair (-∂y b, ∂x b), w
    double* om_pd[3] = {nullptr, nullptr, nullptr};
    double2* om_q = nullptr;
    double2* om_gb = nullptr;
    double* om_w = nullptr;
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
    SolveHistory adjc_hist;           // q of the last transposed solves (working buffer: adjc_q)
    int adjc_period = 0;              // transposed solves per step of the running adjoint sweep, else 0
    // hipGraphs of step runs within one filter cycle, keyed by (first step's
    // phase in the cycle, number of steps), for the parameters `p`
    struct GraphCache {
        std::map<std::pair<int, int>, hipGraphExec_t> exec;
        bool have_p = false;
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

namespace ssg_impl {


inline int set_err(ssg_ctx* c, int code, const char* fmt, ...) {
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

inline bool pow2_in_range(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

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
inline bool ept_valid(int n, int e) {
    if (n < 64) return e == (n < 16 ? n : 16);
    return (e == 4 || e == 8 || e == 16) && n / e <= 256;
}
inline int ept_default(int n) {
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

inline int prof_begin(ssg_ctx* c, int* slot);
inline int prof_end(ssg_ctx* c, int slot, int kind, double bytes);

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
inline PathBuffers<double>& bufs<double>(ssg_ctx* c) { return c->d; }
template <>
inline PathBuffers<float>& bufs<float>(ssg_ctx* c) { return c->f; }

template <typename R>
int ept_of(ssg_ctx* c, int kind, int n) {
    const int p = std::is_same<R, float>::value ? 1 : 0;
    const int e = c->ept[p][kind];
    return (e && ept_valid(n, e)) ? e : ept_default(n);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// rows > 0 with `grid` transforms that many rows of a stacked field (the
// levels of a 3D field, grid pitch nz * ny)
template <typename R, typename S, int SRC>
int launch_rows_fwd(ssg_ctx* c, const S* f, const cpx<R>* w, cpx<R>* PA, cpx<R>* PB, int ncols, int rows = 0,
                    const Grid* grid = nullptr) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_of<R>(c, 0, c->nx);
    if (rows <= 0) rows = c->ny;
    return with_size_ept(c->nx, ept, [&](auto tag, auto etag) -> int {
        constexpr int N = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = N / E;
        const int rows_per_line = SRC == SRC_PAIR ? 2 : 1;
        const int G = lines_per_block<R>(N, E, rows / rows_per_line);
        const int nblk = rows / (rows_per_line * G);
        RowsFwdArgs<R, S> a{f, w, PA, PB, ncols, G, nblk, B.twx, grid ? *grid : c->g};
        constexpr bool PP = pingpong_fits<R, N, E, 1>();
        const size_t lds = (PP ? 2 : 1) * (size_t)G * lds_line_elems(N) * sizeof(cpx<R>);
        const double n2 = (double)c->nx * rows, plane = (double)ncols * rows * sizeof(cpx<R>);
        const double bytes = SRC == SRC_FLUX ? n2 * (sizeof(double) + sizeof(cpx<R>)) + 2 * plane
                             : SRC == SRC_W  ? n2 * sizeof(cpx<R>) + 2 * plane
                                             : n2 * sizeof(S) + plane;
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_rows_fwd<R, S, N, E, SRC, PP>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, 0, bytes);
    });
}

// General column-pass launcher: `a` carries the operands; lines are
// (kx, level) pairs when nlev > 1 (kx = line / nlev).
template <typename R, int OP>
int launch_cols_args(ssg_ctx* c, ColsArgs<R> a, int nlines, int nlev, int ept_force = 0) {
    PathBuffers<R>& B = bufs<R>(c);
    const int ept = ept_force ? ept_force : ept_of<R>(c, 1, c->ny);
    return with_size_ept(c->ny, ept, [&](auto tag, auto etag) -> int {
        constexpr int M = decltype(tag)::value;
        constexpr int E = decltype(etag)::value;
        constexpr int T = M / E;
        const int le = lds_line_elems(M) * (int)sizeof(cpx<R>) * cols_nbuf(OP);
        // OP_MA_DIFF: nlines = kxn * nlev, blocks of G kx of one level (G <= kxn),
        // ceil(kxn / G) groups per level
        const int glim = OP == OP_MA_DIFF ? nlines / nlev : nlines;
        int G = 1;
        while (2 * G * T <= 256 && 2 * G * le <= kMaxLdsBytes && 2 * G <= glim) G *= 2;
        const int nblk = OP == OP_MA_DIFF ? (glim + G - 1) / G * nlev : (nlines + G - 1) / G;
        const double kmx = M_PI * c->nx / c->lx, kmy = M_PI * c->ny / c->ly;
        a.ncols = nlines;
        a.G = G;
        a.inv_kmax = (R)(1.0 / std::sqrt(kmx * kmx + kmy * kmy));
        a.twy = B.twy;
        a.g = c->g;
        a.nlev = nlev;
        if ((OP == OP_MA_DIFF || OP == OP_MA_ADJ) && nblk > c->ma_partial_cap)
            return set_err(c, SSG_ERR_ARG, "reduction buffer too small (%d blocks)", nblk);
        constexpr bool PP = pingpong_fits<R, M, E, cols_nbuf(OP)>();
        const size_t lds = (PP ? 2 : 1) * (size_t)G * le;
        const int nin = OP == OP_MA_ADJ ? 3
                        : (OP == OP_TEND || OP == OP_TEND_Q || OP == OP_JAC || OP == OP_MA_DIFF || OP == OP_PHIBAR) ? 2
                                                                                                                  : 1;
        const int nout = (OP == OP_VEL || OP == OP_VEL_SPEC) ? 2 : (OP == OP_MA_DERIV ? 3 : 1);
        const double bytes = (double)(nin + nout) * nlines * c->ny * sizeof(cpx<R>);
        int slot, rc;
        if ((rc = prof_begin(c, &slot))) return rc;
        launch_on(c, slot, k_cols<R, M, E, OP, PP>, dim3(nblk), dim3(G * T), lds, a);
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
        constexpr bool PP = pingpong_fits<R, N, E, 2>();
        const size_t lds = (PP ? 2 : 1) * (size_t)G * 2 * lds_line_elems(N) * sizeof(cpx<R>);
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
        launch_on(c, slot, k_rows_inv<R, N, E, MODE, EPI, PP>, dim3(nblk), dim3(G * T), lds, a);
        c->launches++;
        HIP_OK(c, hipGetLastError());
        return prof_end(c, slot, 2, bytes);
    });
}

// Event pair of one launch when profiling is on (recorded by launch_on).
// `bytes` is the launch's algorithmic traffic: every input read once,
// every output written once.
inline int prof_begin(ssg_ctx* c, int* slot) {
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

inline int prof_end(ssg_ctx* c, int slot, int kind, double bytes) {
    if (slot < 0) return SSG_OK;
    c->prof_recs.push_back({kind, bytes, slot, slot + 1});
    return SSG_OK;
}

// one thread per element, 256-thread blocks
inline unsigned grid_all(long n) { return (unsigned)((n + 255) / 256); }

inline int grid1d(long n) {
    long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------------------
// path pieces
// ---------------------------------------------------------------------------

// surface velocities (src/setICs.jl:161-192) in FP64, and their FP32 copy
inline int compute_velocities(ssg_ctx* c) {
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
inline bool same_params(const ssg_time_params& a, const ssg_time_params& b) {
    return a.dt == b.dt && a.scheme == b.scheme && a.filter_freq == b.filter_freq &&
           a.filter_strength == b.filter_strength && a.filter_order == b.filter_order &&
           a.filter_cutoff == b.filter_cutoff && a.ab2_reference_aliasing == b.ab2_reference_aliasing;
}

// SSG_ERR_NOT_CONVERGED for a failed Monge-Ampère solve; a solve rejected as
// not elliptic keeps the message that says so
inline int ma_not_converged(ssg_ctx* c, const ssg_ma_diag& d) {
    if (d.not_elliptic) return SSG_ERR_NOT_CONVERGED;
    return set_err(c, SSG_ERR_NOT_CONVERGED, "Monge-Ampere solve did not converge after %d iterations (rel. residual %.3e)",
                   d.iterations, d.relative_residual);
}

// shared across translation units
extern std::mutex g_tune_mu;
const char* noalloc_last_error();
int ma_solve_dev(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d);
int ma_solve_coupled(ssg_ctx* c, const double* bs, const ssg_ma_params* p, ssg_ma_diag* d, bool final_solve);
int rfft2_device(ssg_ctx* c, const double* src, cpx<double>* dst = nullptr);
int ma_surface_outputs(ssg_ctx* c);
void ma_free(ssg_ctx* c);
void ma_history_reset(ssg_ctx* c);
size_t ma_state_bytes(const ssg_ctx* c);
int ma_state_save(ssg_ctx* c, char* p);
int ma_state_load(ssg_ctx* c, const char* p, size_t avail, int version);
void adjc_free(ssg_ctx* c);
int adj_alloc(ssg_ctx* c);

}  // namespace ssg_impl
