// Monge-Ampère solve of the as-wired SSG problem (src/poisson.jl:143-364):
//   r = Φxx + Φyy + Φzz - ε(Φxx Φyy - Φxy²) - b = 0 on interior levels,
//   one-sided Neumann rows at the bottom (slope 0) and top (slope g).
// Φ is held as its horizontal spectrum Φ̂[kx][level][ky] (kx <= nx/2).  The
// linear part is diagonal in (kx, ky) and tridiagonal in z, so it is
// inverted exactly per wavenumber (k_ma_tridiag); the nonlinear term is a
// defect-correction (Picard) iteration Φ̂ <- L⁻¹(b̂ + ε D̂(Φ)).
//   k_cols<OP_MA_DERIV> : Φ̂ -> y-inverse of Φxx, Φyy, Φxy symbols
//   k_ma_rows           : x-inverse of the three, D = Φxx Φyy - Φxy², x-FFT of D
//   k_cols<OP_MA_DIFF>  : y-FFT of D, residual ε(D_prev - D) sums, D_prev <- D
//   k_ma_tridiag        : per (kx, ky) tridiagonal solve -> Φ̂
#pragma once
#include "ssg_kernels.hpp"

namespace ssg {

template <typename R>
struct MaRowsArgs {
    const cpx<R>* Pxx;
    const cpx<R>* Pyy;
    const cpx<R>* Pxy;
    cpx<R>* Dx;
    R* phys[3];  // MA_ROWS_PHYS: the three fields in physical space, [row][x]
    int ncols;   // nx/2 + 1
    int G;
    int nblk;
    const cpx<R>* twx;
    Grid g;      // pitch = nz * ny
    int yt_log2; // row tile of the three input planes and of Dx (ma_tile_off)
};

// gather the packed spectrum of rows (A, B) of one plane into the line buffers
template <typename R, int N, int E, int MAXIT>
__device__ __forceinline__ void ma_load(const cpx<R>* P, int pitch, int y0, int G, int total, cpx<R> (&q0)[MAXIT],
                                        cpx<R> (&q1)[MAXIT]) {
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
        const int idx = threadIdx.x + it * (int)blockDim.x;
        if (idx < total) {
            const int kx = idx >> gsh(G), ll = idx & (G - 1);
            const size_t o = (size_t)kx * pitch + y0 + ll;
            q0[it] = P[o];
            q1[it] = P[o + G];
        }
    }
}

// Gather from a row-tiled plane (ma_tile_off, tiles of 2^s rows of kxn
// columns) for a block whose 2G rows y0.. lie in one tile (2G <= 2^s, y0 a
// multiple of 2G): `tp` points at the block's first row in its tile
// (ma_tile_off(y0, 0, ...)), so row y0 + j of column kx is tp[(kx << s) + j];
// rows A = y0 + l and B = y0 + G + l.  The plane holds every column
// kx = 0..N/2 (the Monge-Ampère planes): E/2 whole passes over kx < N/2,
// with no bounds test.  The thread of kx = 0 also fetches the Nyquist
// column's real parts into the imaginary slots of its DC pair: the Hermitian
// unpack discards both imaginary parts (the DC and Nyquist bins of a real
// row are real), so no pass and no registers go to the one column past N/2.
template <typename R, int N, int E>
__device__ __forceinline__ void ma_load_tile_full(const cpx<R>* tp, int s, int G, cpx<R> (&q0)[E / 2],
                                                  cpx<R> (&q1)[E / 2]) {
    constexpr int T = N / E;
#pragma unroll
    for (int it = 0; it < E / 2; ++it) {
        const int idx = threadIdx.x + it * G * T;
        const unsigned o = ((unsigned)(idx >> gsh(G)) << s) + (unsigned)(idx & (G - 1));
        q0[it] = tp[o];
        q1[it] = tp[o + G];
    }
    if ((int)threadIdx.x < G) {
        const unsigned o = ((unsigned)(N / 2) << s) + threadIdx.x;
        q0[0].y = tp[o].x;
        q1[0].y = tp[o + G].x;
    }
}

// Hermitian unpack of ma_load_tile_full's registers into the line buffers:
// line ll receives X = A + iB of its rows A (re) and B (im)
template <typename R, int N, int E>
__device__ __forceinline__ void ma_scatter_full(cpx<R>* lds, int G, const cpx<R> (&q0)[E / 2],
                                                const cpx<R> (&q1)[E / 2]) {
    using C = cpx<R>;
    constexpr int T = N / E;
    constexpr int LE = lds_line_elems(N);
#pragma unroll
    for (int it = 0; it < E / 2; ++it) {
        const int idx = threadIdx.x + it * G * T;
        const int kx = idx >> gsh(G), ll = idx & (G - 1);
        const C p0 = q0[it], p1 = q1[it];
        C* bl = lds + ll * LE;
        if (it == 0 && kx == 0) {
            bl[lpad(0)] = C{p0.x, p1.x};
            bl[lpad(N / 2)] = C{p0.y, p1.y};
        } else {
            bl[lpad(kx)] = C{p0.x - p1.y, p0.y + p1.x};
            bl[lpad(N - kx)] = C{p0.x + p1.y, p1.x - p0.y};
        }
    }
}

template <typename R, int N, int E, int MAXIT>
__device__ __forceinline__ void ma_scatter(cpx<R>* lds, int G, int total, const cpx<R> (&q0)[MAXIT],
                                           const cpx<R> (&q1)[MAXIT], int lstride) {
    using C = cpx<R>;
    constexpr int LE = lds_line_elems(N);
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
        const int idx = threadIdx.x + it * (int)blockDim.x;
        if (idx < total) {
            const int kx = idx >> gsh(G), ll = idx & (G - 1);
            C p0 = q0[it], p1 = q1[it];
            C* bl = lds + ll * lstride * LE;
            const bool edge = (kx == 0) || (2 * kx == N);
            if (edge) {
                p0.y = 0;
                p1.y = 0;
            }
            bl[lpad(kx)] = C{p0.x - p1.y, p0.y + p1.x};
            if (!edge) bl[lpad(N - kx)] = C{p0.x + p1.y, p1.x - p0.y};
        }
    }
}

enum { MA_ROWS_D = 0, MA_ROWS_PHYS = 1 };

// MA_ROWS_D: D = Φxx Φyy - Φxy² and its x-FFT; MA_ROWS_PHYS: the three
// derivative fields stored in physical space (for the adjoint).  The three
// inverse transforms run one after another through a single LDS line
// buffer: a third of the LDS of a lockstep form, so three times the blocks
// per CU, which is what this strided, bandwidth-bound pass needs.  All
// global loads are issued before the first transform.  The planes hold every
// column kx = 0..nx/2 (a.ncols = nx/2 + 1, the context's ncols_full).
template <typename R, int N, int E, int MODE = MA_ROWS_D>
__global__ __launch_bounds__(256) void k_ma_rows(MaRowsArgs<R> a) {
    using C = cpx<R>;
    using F = LineFFT<R, N, E>;
    constexpr int T = F::T;
    constexpr int LE = lds_line_elems(N);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int G = a.G;
    const int blk = xcd_block(blockIdx.x, a.nblk);
    const int y0 = blk * 2 * G;
    typename F::Twiddles tw;
    tw.load(t, a.twx);
    C* const buf = lds + l * LE;
    // the planes are row-tiled (ma_tile_off): per kx this block's 2G rows are
    // one run, and consecutive kx are adjacent within a tile
    const int ts = a.yt_log2;
    const size_t t0 = (size_t)ma_tile_off(y0, 0, a.ncols, ts);  // this block's rows, one tile
    C qa0[E / 2], qa1[E / 2], qb0[E / 2], qb1[E / 2], qc0[E / 2], qc1[E / 2];
    ma_load_tile_full<R, N, E>(a.Pxx + t0, ts, G, qa0, qa1);
    ma_load_tile_full<R, N, E>(a.Pyy + t0, ts, G, qb0, qb1);
    if constexpr (MODE == MA_ROWS_PHYS) {
        ma_load_tile_full<R, N, E>(a.Pxy + t0, ts, G, qc0, qc1);
        C w3[3][E];
        ma_scatter_full<R, N, E>(lds, G, qa0, qa1);
        F::template run_from_lds<+1>(w3[0], buf, t, tw);
        __syncthreads();
        ma_scatter_full<R, N, E>(lds, G, qb0, qb1);
        F::template run_from_lds<+1>(w3[1], buf, t, tw);
        __syncthreads();
        ma_scatter_full<R, N, E>(lds, G, qc0, qc1);
        F::template run_from_lds<+1>(w3[2], buf, t, tw);
        const size_t rA = (size_t)(y0 + l) * a.g.nx, rB = (size_t)(y0 + G + l) * a.g.nx;
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int m = 0; m < E; ++m) {
                a.phys[f][rA + t + T * m] = w3[f][m].x;
                a.phys[f][rB + t + T * m] = w3[f][m].y;
            }
        return;
    } else {
        // D = Φxx Φyy - Φxy² for rows A (re) and B (im): compute_d_operator!
        // (src/poisson.jl:183-223), formed as the transforms finish, so one
        // product and one transform are live rather than three transforms
        C v[E], w[E];
        ma_scatter_full<R, N, E>(lds, G, qa0, qa1);
        // the third plane's loads wait until the first plane's registers are
        // free; they are in flight through the first and second transforms
        ma_load_tile_full<R, N, E>(a.Pxy + t0, ts, G, qc0, qc1);
        F::template run_from_lds<+1>(v, buf, t, tw);
        __syncthreads();
        ma_scatter_full<R, N, E>(lds, G, qb0, qb1);
        F::template run_from_lds<+1>(w, buf, t, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = C{v[m].x * w[m].x, v[m].y * w[m].y};
        __syncthreads();
        ma_scatter_full<R, N, E>(lds, G, qc0, qc1);
        F::template run_from_lds<+1>(w, buf, t, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = C{v[m].x - w[m].x * w[m].x, v[m].y - w[m].y * w[m].y};
        F::template run<-1>(v, buf, t, tw);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) buf[lpad(t + T * m)] = v[m];
        __syncthreads();
        // D's x-spectrum into the row-tiled output plane, rows A = y0 + l and
        // B = y0 + G + l: E/2 whole passes over kx < N/2, and the Nyquist
        // column by the threads of kx = 0
        C* const dt0 = a.Dx + t0;
        auto store = [&](int kx, int ll) {
            const C* bl = lds + ll * LE;
            const C z1 = bl[lpad(kx)];
            const C z2 = cconj(bl[lpad((N - kx) & (N - 1))]);
            const unsigned o = ((unsigned)kx << ts) + (unsigned)ll;
            dt0[o] = C{(z1.x + z2.x) * (R)0.5, (z1.y + z2.y) * (R)0.5};
            dt0[o + G] = C{(z1.y - z2.y) * (R)0.5, -(z1.x - z2.x) * (R)0.5};
        };
#pragma unroll
        for (int it = 0; it < E / 2; ++it) {
            const int idx = threadIdx.x + it * G * T;
            store(idx >> gsh(G), idx & (G - 1));
        }
        if ((int)threadIdx.x < G) store(N / 2, threadIdx.x);
    }
}

// Adjoint row pass: λ (x-spectral) -> physical; the products Φyy λ, Φxx λ,
// Φxy λ with the stored physical derivative fields of the linearisation
// point; x-FFT of the three products.
template <typename R>
struct MaAdjRowsArgs {
    const cpx<R>* lam;      // [kx][row], pitch = nz * ny
    const R* pxx;           // physical [row][x]
    const R* pyy;
    const R* pxy;
    cpx<R>* out[3];         // x-spectra of Φyy λ, Φxx λ, Φxy λ
    int ncols;
    int G;
    int nblk;
    const cpx<R>* twx;
    Grid g;
};

template <typename R, int N, int E>
__global__ __launch_bounds__(256) void k_ma_adj_rows(MaAdjRowsArgs<R> a) {
    using C = cpx<R>;
    using F = LineFFT<R, N, E>;
    constexpr int T = F::T;
    constexpr int LE = lds_line_elems(N);
    constexpr int MAXIT = (N / 2 + 1 + T - 1) / T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int G = a.G;
    const int blk = xcd_block(blockIdx.x, a.nblk);
    const int y0 = blk * 2 * G;
    const int pitch = a.g.pitch;
    const int total = a.ncols * G;
    typename F::Twiddles tw;
    tw.load(t, a.twx);
    const size_t rA = (size_t)(y0 + l) * a.g.nx, rB = (size_t)(y0 + G + l) * a.g.nx;
    R fA[3][E], fB[3][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int x = t + T * m;
        fA[0][m] = a.pyy[rA + x];
        fB[0][m] = a.pyy[rB + x];
        fA[1][m] = a.pxx[rA + x];
        fB[1][m] = a.pxx[rB + x];
        fA[2][m] = a.pxy[rA + x];
        fB[2][m] = a.pxy[rB + x];
    }
    C q0[MAXIT], q1[MAXIT];
    ma_load<R, N, E, MAXIT>(a.lam, pitch, y0, G, total, q0, q1);
    ma_scatter<R, N, E, MAXIT>(lds, G, total, q0, q1, 3);
    C* const b3[3] = {lds + (3 * l) * LE, lds + (3 * l + 1) * LE, lds + (3 * l + 2) * LE};
    C v[E];
    F::template run_from_lds<+1>(v, b3[0], t, tw);
    C w3[3][E];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int m = 0; m < E; ++m) w3[f][m] = C{fA[f][m] * v[m].x, fB[f][m] * v[m].y};
    F::template run_n<-1, 3>(w3, b3, t, tw);
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int m = 0; m < E; ++m) b3[f][lpad(t + T * m)] = w3[f][m];
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 3; ++f)
        store_split_stride<R, N, E>(lds + f * LE, G, 3, a.ncols, a.out[f], a.out[f], pitch, y0, y0 + G);
}

// Per-level stencil data, 0-based levels, nz >= 3 (src/transforms.jl:798-850,
// src/poisson.jl:309-364):
//   c[3k+0..2] = (alpha, beta, gamma) of d2dz2 at interior level k
//   c[3nz+0..2] = (A, B, C) bottom row  A Φ0 + B Φ1 + C Φ2 = 0
//   c[3nz+3..5] = (a, b, c) top row     a Φn-1 + b Φn-2 + c Φn-3 = g
// R: precision of the D̂ input, P: precision of the Φ̂ output; the solve
// itself is FP64 arithmetic in every case
template <typename R, typename P = double>
struct MaTriArgs {
    const cpx<R>* Dh;           // [kx][lev][ky]
    const cpx<double>* B2;      // [kx][ky] forcing spectrum
    cpx<P>* Phi;                // [kx][lev][ky]
    const double* coef;
    int nz;
    int ncols;
    double eps;
    int body;                   // B2 on interior rows (variant A)
    int top;                    // B2 as top Neumann data (variant B)
    int use_d;                  // 0: D term omitted (first solve without history)
    int nt;                     // nontemporal plane access (ld_plane)
    Grid g;
    // extrapolating form: D̂ is the warm start ring[0] rc[0] + ring[1] rc[1]
    // (+ ring[2] rc[2] when next = 3), formed here and also stored to Dg
    const cpx<R>* ring[3];
    double rc[3];
    int next;                   // history entries read: 2 (linear) or 3 (quadratic)
    cpx<R>* Dg;
};

// Two lanes per wavenumber (a twisted factorisation): the even lane
// eliminates upward from the bottom row through levels 0..m-1, the odd lane
// downward from the top row through levels nz-1..m (m = nz/2), both with the
// same code on mirrored coefficients; the two interface rows are solved
// after exchanging their reduced forms (lane shuffles), and each lane
// back-substitutes its own half.  Half the registers and half the dependent
// chain of one Thomas sweep per wavenumber.  NZH >= max(2, ceil(nz/2)).
// EXT: the extrapolating form, which reads the two or three history entries
// in place of D̂, forms the warm start itself and stores it (the residual
// pass needs it as D̂_prev): one pass over the planes fewer than k_extrap
// followed by this solve, with the same bits.
template <typename R, typename P, int NZH, bool EXT = false>
__global__ __launch_bounds__(256) void k_ma_tridiag(MaTriArgs<R, P> a) {
    // the stencil table (3 nz + 6 values, nz <= 2 NZH) in LDS: its reads
    // cost no vector-memory waits in the sweep
    __shared__ double cf[3 * 2 * NZH + 6];
    for (int i = threadIdx.x; i < 3 * a.nz + 6; i += blockDim.x) cf[i] = a.coef[i];
    __syncthreads();
    const long total = (long)a.ncols * a.g.ny;
    const long q = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 1;
    if (q >= total) return;  // both lanes of a pair leave together
    const bool up = (threadIdx.x & 1) != 0;
    const int ny = a.g.ny, nz = a.nz, m = nz / 2;
    const int cnt = up ? nz - m : m;
    const int kx = (int)(q / ny), ky = (int)(q - (long)kx * ny);
    const int s = signed_k(ky, ny);
    const double hs = (double)sym_dxx<double>(a.g, kx) + (double)sym_dyy<double>(a.g, s);
    const cpx<double> b2 = a.B2[(size_t)kx * ny + ky];
    const size_t base = (size_t)kx * nz * ny + ky;
    // level of sweep position j: 0, 1, ... from the bottom (even lane) or
    // nz-1, nz-2, ... from the top (odd lane); j >= 1 are interior rows
    auto lev = [&](int j) { return up ? nz - 1 - j : j; };
    double cc[NZH];
    double2 dd[NZH];
    // every D̂ value the lane needs in one straight run of loads: slot j >= 1
    // holds level lev(j) (dd[j] holds its right-hand side until the sweep
    // overwrites it; slot 1 is also the interior row next to the boundary
    // row); positions past this lane's rows load a clamped, in-bounds level
    // and are unused.  Slot 0 (the boundary level, which carries no D̂) is
    // loaded only by the extrapolating form, which stores the warm start on
    // every level.
    cpx<R> dv[NZH];
    if (EXT || a.use_d) {
        int lv[NZH];
#pragma unroll
        for (int j = 0; j < NZH; ++j) {
            const int k = lev(j);
            lv[j] = k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k);
        }
        constexpr int J0 = EXT ? 0 : 1;
        using V = typename vec2_t<R>::type;
        auto load_all = [&](const cpx<R>* col, cpx<R> (&out)[NZH]) {
            if (a.nt) {
#pragma unroll
                for (int j = J0; j < NZH; ++j) {
                    const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(col + (size_t)lv[j] * ny));
                    out[j] = cpx<R>{w.x, w.y};
                }
            } else {
#pragma unroll
                for (int j = J0; j < NZH; ++j) out[j] = col[(size_t)lv[j] * ny];
            }
        };
        if constexpr (EXT) {
            cpx<R> x1[NZH];
            load_all(a.ring[0] + base, dv);
            load_all(a.ring[1] + base, x1);
            if (a.next == 3) {
                cpx<R> x2[NZH];
                load_all(a.ring[2] + base, x2);
#pragma unroll
                for (int j = 0; j < NZH; ++j) dv[j] = extrap3(dv[j], a.rc[0], x1[j], a.rc[1], x2[j], a.rc[2]);
            } else {
#pragma unroll
                for (int j = 0; j < NZH; ++j) dv[j] = extrap2(dv[j], a.rc[0], x1[j], a.rc[1]);
            }
            cpx<R>* const g = a.Dg + base;
            if (a.nt) {
#pragma unroll
                for (int j = 0; j < NZH; ++j)
                    if (j < cnt)
                        __builtin_nontemporal_store(V{dv[j].x, dv[j].y}, reinterpret_cast<V*>(g + (size_t)lev(j) * ny));
            } else {
#pragma unroll
                for (int j = 0; j < NZH; ++j)
                    if (j < cnt) g[(size_t)lev(j) * ny] = dv[j];
            }
        } else {
            load_all(a.Dh + base, dv);
        }
    }
    auto rhs_of = [&](int j) -> double2 {
        double2 f = {0.0, 0.0};
        if (a.body) f = b2;
        if (EXT || a.use_d) {
            f.x += a.eps * (double)dv[j].x;
            f.y += a.eps * (double)dv[j].y;
        }
        return f;
    };
#pragma unroll
    for (int j = 1; j < NZH; ++j) dd[j] = rhs_of(j);
    // the boundary row of this lane, with the next level eliminated through
    // the adjacent interior row (src/poisson.jl:309-364)
    const double2 fb = rhs_of(1);
    {
        double diag, next;
        double2 f;
        if (!up) {  // bottom: A0 Φ0 + B0 Φ1 + C0 Φ2 = 0
            const double A0 = cf[3 * nz], B0 = cf[3 * nz + 1], C0 = cf[3 * nz + 2];
            const double al = cf[3], be = cf[4] + hs, ga = cf[5];
            diag = A0 - C0 * al / ga;
            next = B0 - C0 * be / ga;
            f = {-C0 * fb.x / ga, -C0 * fb.y / ga};
            if (kx == 0 && ky == 0) {  // gauge of the Neumann-Neumann mean mode
                diag = 1.0;
                next = 0.0;
                f = {0.0, 0.0};
            }
        } else {  // top: a Φn-1 + b Φn-2 + c Φn-3 = g
            const int ki = nz - 2;
            const double at = cf[3 * nz + 3], bt = cf[3 * nz + 4], ct = cf[3 * nz + 5];
            const double al = cf[3 * ki], be = cf[3 * ki + 1] + hs, ga = cf[3 * ki + 2];
            diag = at - ct * ga / al;
            next = bt - ct * be / al;
            double2 g = {0.0, 0.0};
            if (a.top) g = b2;
            f = {g.x - ct * fb.x / al, g.y - ct * fb.y / al};
        }
        cc[0] = next / diag;
        dd[0] = {f.x / diag, f.y / diag};
    }
    // interior rows in sweep order: `pv` couples to the level already
    // eliminated, `nx` to the next one
    // Every position is swept, also those past this lane's rows (j >= cnt,
    // levels clamped into the stencil table, results unused): without a
    // per-row predicate the stencil loads can all be issued ahead of the
    // dependent chain.
#pragma unroll
    for (int j = 1; j < NZH; ++j) {
        const int k0 = lev(j), k = k0 < 1 ? 1 : (k0 > nz - 2 ? nz - 2 : k0);
        const double al = cf[3 * k], be = cf[3 * k + 1] + hs, ga = cf[3 * k + 2];
        const double pv = up ? ga : al, nx = up ? al : ga;
        const double2 f = dd[j];
        const double rm = 1.0 / (be - pv * cc[j - 1]);
        cc[j] = nx * rm;
        dd[j] = {(f.x - pv * dd[j - 1].x) * rm, (f.y - pv * dd[j - 1].y) * rm};
    }
    // interface rows: Φ(m-1) + cp Φ(m) = dp (even lane), Φ(m) + cq Φ(m-1) = dq
    // (odd lane); both lanes solve the pair with the same operations
    double cl = 0.0;
    double2 dl = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NZH; ++j)
        if (j == cnt - 1) {
            cl = cc[j];
            dl = dd[j];
        }
    const double co = __shfl_xor(cl, 1);
    const double2 dox = {__shfl_xor(dl.x, 1), __shfl_xor(dl.y, 1)};
    const double cp = up ? co : cl, cq = up ? cl : co;
    const double2 dp = up ? dox : dl, dq = up ? dl : dox;
    const double den = 1.0 / (1.0 - cq * cp);
    const double2 xm = {(dq.x - cq * dp.x) * den, (dq.y - cq * dp.y) * den};
    const double2 xm1 = {dp.x - cp * xm.x, dp.y - cp * xm.y};
    // back-substitution of this lane's half, from the interface outward
    double2 xn = up ? xm1 : xm;
    cpx<P> xo[NZH];
#pragma unroll
    for (int j = NZH - 1; j >= 0; --j) {
        if (j < cnt) {  // (the positions past this lane's rows hold no solution)
            const double2 x = {dd[j].x - cc[j] * xn.x, dd[j].y - cc[j] * xn.y};
            xo[j] = cpx<P>{(P)x.x, (P)x.y};
            xn = x;
        }
    }
    cpx<P>* out = a.Phi + base;
    if (a.nt) {
        using V = typename vec2_t<P>::type;
#pragma unroll
        for (int j = 0; j < NZH; ++j)
            if (j < cnt) __builtin_nontemporal_store(V{xo[j].x, xo[j].y}, reinterpret_cast<V*>(out + (size_t)lev(j) * ny));
    } else {
#pragma unroll
        for (int j = 0; j < NZH; ++j)
            if (j < cnt) out[(size_t)lev(j) * ny] = xo[j];
    }
}

}  // namespace ssg

namespace ssg {

// Transpose of the per-wavenumber solve Φ̂ = L⁻¹ f̃ of k_ma_tridiag.  That
// solve is A' Φ = E f̃ with A' tridiagonal (the one-sided Neumann rows after
// the elimination through their adjacent interior rows) and E the
// elimination (r0 = -C0 f1/γ1, r_top = g - ct f_{n-2}/α_{n-2}; the gauged
// mean mode replaces row 0 by Φ0 = 0), so its transpose is Eᵀ A'⁻ᵀ.
// Input y per level: ε q on interior levels (use_q) plus φ̄ on the top
// level; output F̄ = Eᵀ A'⁻ᵀ y per row (interior rows carry f, the top row g)
// and b̄ = Σ interior F̄ (body variant) or F̄_top (surface variant).
struct MaTriAdjArgs {
    const cpx<double>* q;       // [kx][lev][ky]
    const cpx<double>* phibar;  // [kx][ky]
    cpx<double>* Fbar;          // [kx][lev][ky]
    cpx<double>* bbar;          // [kx][ky]
    const double* coef;
    int nz;
    int ncols;
    double eps;
    int body;
    int top;
    int use_q;
    Grid g;
};

template <int NZMAX>
__global__ __launch_bounds__(128) void k_ma_tridiag_adj(MaTriAdjArgs a) {
    const long total = (long)a.ncols * a.g.ny;
    const long qi = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (qi >= total) return;
    const int ny = a.g.ny, nz = a.nz;
    const int kx = (int)(qi / ny), ky = (int)(qi - (long)kx * ny);
    const int s = signed_k(ky, ny);
    const double hs = (double)sym_dxx<double>(a.g, kx) + (double)sym_dyy<double>(a.g, s);
    const double* cf = a.coef;
    const size_t base = (size_t)kx * nz * ny + ky;
    const bool gauge = (kx == 0 && ky == 0);
    // rows of A': D_k diagonal, L_k sub-diagonal (coefficient of Φ_{k-1}),
    // U_k super-diagonal (coefficient of Φ_{k+1})
    const double A0 = cf[3 * nz], B0 = cf[3 * nz + 1], C0 = cf[3 * nz + 2];
    const double ga1 = cf[5];
    const double D0 = gauge ? 1.0 : A0 - C0 * cf[3] / ga1;
    const double U0 = gauge ? 0.0 : B0 - C0 * (cf[4] + hs) / ga1;
    const int ki = nz - 2;
    const double alT = cf[3 * ki], beT = cf[3 * ki + 1] + hs, gaT = cf[3 * ki + 2];
    const double at = cf[3 * nz + 3], bt = cf[3 * nz + 4], ct = cf[3 * nz + 5];
    const double LN = bt - ct * beT / alT, DN = at - ct * gaT / alT;
    auto Dk = [&](int k) -> double { return k == 0 ? D0 : (k == nz - 1 ? DN : cf[3 * k + 1] + hs); };
    auto Lk = [&](int k) -> double { return k == nz - 1 ? LN : cf[3 * k]; };       // 1 <= k <= nz-1
    auto Uk = [&](int k) -> double { return k == 0 ? U0 : cf[3 * k + 2]; };        // 0 <= k <= nz-2
    // Thomas on A'ᵀ: row k has sub U_{k-1}, diagonal D_k, super L_{k+1}.
    // dp[k] first holds the right-hand side y_k, every value loaded before
    // the sweep in one straight run (levels clamped into the plane, values
    // off the interior levels unused), then the sweep overwrites it in place.
    double cp[NZMAX];
    double2 dp[NZMAX];
    if (a.use_q) {
#pragma unroll
        for (int k = 0; k < NZMAX; ++k) {
            const int kk = k > nz - 1 ? nz - 1 : k;
            const cpx<double> qq = a.q[base + (size_t)kk * ny];
            const bool in = k >= 1 && k <= nz - 2;
            dp[k] = {in ? a.eps * qq.x : 0.0, in ? a.eps * qq.y : 0.0};
        }
    } else {
#pragma unroll
        for (int k = 0; k < NZMAX; ++k) dp[k] = {0.0, 0.0};
    }
    const cpx<double> pb = a.phibar[(size_t)kx * ny + ky];
    auto yk = [&](int k) -> double2 {
        double2 y = dp[k];
        if (k == nz - 1) {
            y.x += pb.x;
            y.y += pb.y;
        }
        return y;
    };
    {
        const double d0 = Dk(0);
        const double2 y0 = yk(0);
        cp[0] = Lk(1) / d0;
        dp[0] = {y0.x / d0, y0.y / d0};
    }
#pragma unroll
    for (int k = 1; k < NZMAX; ++k) {
        if (k < nz) {
            const double sub = Uk(k - 1);
            const double m = Dk(k) - sub * cp[k - 1];
            const double2 y = yk(k);
            cp[k] = (k + 1 < nz) ? Lk(k + 1) / m : 0.0;
            dp[k] = {(y.x - sub * dp[k - 1].x) / m, (y.y - sub * dp[k - 1].y) / m};
        }
    }
    // back substitution (μ overwrites dp); F̄ = Eᵀ μ adds the eliminations'
    // transposes: F̄_1 -= (C0/γ1) μ_0, F̄_{n-2} -= (ct/α_{n-2}) μ_{n-1}
    double2 corr = {0.0, 0.0};
    double2 x = {0.0, 0.0};
#pragma unroll
    for (int k = NZMAX - 1; k >= 0; --k) {
        if (k <= nz - 1) {
            x = {dp[k].x - cp[k] * x.x, dp[k].y - cp[k] * x.y};
            dp[k] = x;
            if (k == nz - 1) corr = {ct / alT * x.x, ct / alT * x.y};
        }
    }
    double2 (&mu)[NZMAX] = dp;
    if (!gauge) {
        mu[1].x -= C0 / ga1 * mu[0].x;
        mu[1].y -= C0 / ga1 * mu[0].y;
    }
    double2 bsum = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NZMAX; ++k) {
        if (k <= nz - 1) {
            double2 f = mu[k];
            if (k == nz - 2) {
                f.x -= corr.x;
                f.y -= corr.y;
            }
            a.Fbar[base + (size_t)k * ny] = f;
            if (a.body && k >= 1 && k <= nz - 2) {
                bsum.x += f.x;
                bsum.y += f.y;
            }
            if (a.top && k == nz - 1) bsum = f;
        }
    }
    a.bbar[(size_t)kx * ny + ky] = bsum;
}

// Residual sums of one evaluation, in pinned host memory that the device
// writes.  The host sets both words to kMaSumUnset before the reduction is
// launched; the device writes each with a relaxed system-scope atomic store
// (single-copy atomic, performed at the system coherence point, so no cache
// writeback is needed), and the host waits until neither word is unset.
// Sums of squares are never negative, so the unset pattern (-1.0) is never a
// result; a NaN result differs from it and is passed on.
struct MaSum {
    unsigned long long r2;  // bits of a double
    unsigned long long h2;
};
constexpr unsigned long long kMaSumUnset = 0xBFF0000000000000ull;  // -1.0

// Final reduction of the per-block residual sums of an OP_MA_DIFF / OP_MA_ADJ
// pass (two doubles per block), in a fixed order, published to `out`.  One
// block of kMaSumThreads; each thread's loads are issued together (unrolled
// by 8) so the pass costs about one memory latency.
constexpr int kMaSumThreads = 1024;
static __global__ __launch_bounds__(kMaSumThreads) void k_ma_sum(const double2* __restrict__ part, int nb,
                                                                 MaSum* out) {
    __shared__ double s0[kMaSumThreads], s1[kMaSumThreads];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int i0 = t; i0 < nb; i0 += 8 * kMaSumThreads) {
        double2 v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = i0 + k * kMaSumThreads;
            v[k] = i < nb ? part[i] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a += v[k].x;
            b += v[k].y;
        }
    }
    s0[t] = a;
    s1[t] = b;
    __syncthreads();
    for (int w = kMaSumThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            s0[t] += s0[t + w];
            s1[t] += s1[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        __hip_atomic_store(&out->r2, (unsigned long long)__double_as_longlong(s0[0]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&out->h2, (unsigned long long)__double_as_longlong(s1[0]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace ssg

namespace ssg {

// ---------------------------------------------------------------------------
// Vertical velocity (ω equation, src/extras/omega_eq.jl:326-652) on the
// Monge-Ampère planes.
//   Q = (Q₁, Q₂),  Q₁ = ∂x u_g ∂x b + ∂x v_g ∂y b,  Q₂ = ∂y u_g ∂x b + ∂y v_g ∂y b
//   with u_g = -∂yΦ, v_g = ∂xΦ per level and b the surface buoyancy
//   (compute_divergence_Q!, :326-405);
//   (∂xx + ∂yy + ∂zz) w* = -2ε ∇·Q with w* = 0 on the bottom and top levels,
//   solved per horizontal wavenumber on the Monge-Ampère vertical stencil;
//   w = w* / (1 - ε (Φxx + Φyy))  (the inverse Jacobian, :612-635).
// ---------------------------------------------------------------------------

// (Q₁, Q₂) at every point of every level from the physical derivative
// fields [lev][y][x] and the surface pair gb = (-∂y b, ∂x b) [y][x]
static __global__ __launch_bounds__(256) void k_omega_q(const double* __restrict__ pxx, const double* __restrict__ pyy,
                                                        const double* __restrict__ pxy,
                                                        const double2* __restrict__ gb, double2* __restrict__ q,
                                                        long n2, long n3) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n3) return;
    const double2 g = gb[i % n2];
    const double bx = g.y, by = -g.x;
    const double xx = pxx[i], yy = pyy[i], xy = pxy[i];
    // ∂x u_g = -Φxy, ∂x v_g = Φxx, ∂y u_g = -Φyy, ∂y v_g = Φxy
    q[i] = double2{-xy * bx + xx * by, -yy * bx + xy * by};
}

struct OmegaTriArgs {
    const cpx<double>* F;  // -2ε ∇·Q, [kx][lev][ky] (forward spectrum)
    cpx<double>* W;        // w*, same layout
    const double* coef;    // the vertical stencil (k_ma_tridiag's table)
    int nz;
    int ncols;
    Grid g;
};

// Per (kx, ky): α_k w_{k-1} + (β_k - |k|²) w_k + γ_k w_{k+1} = f_k on the
// interior levels, w = 0 on the bottom and top levels (Thomas sweep).
template <int NZMAX>
__global__ __launch_bounds__(128) void k_omega_tridiag(OmegaTriArgs a) {
    const long total = (long)a.ncols * a.g.ny;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int ny = a.g.ny, nz = a.nz;
    const int kx = (int)(q / ny), ky = (int)(q - (long)kx * ny);
    const int s = signed_k(ky, ny);
    const double hs = (double)sym_dxx<double>(a.g, kx) + (double)sym_dyy<double>(a.g, s);
    const double* cf = a.coef;
    const size_t base = (size_t)kx * nz * ny + ky;
    double cp[NZMAX];
    double2 dp[NZMAX];
    // every right-hand side in one straight run of loads (levels clamped into
    // the plane; positions past the interior are unused)
#pragma unroll
    for (int k = 1; k < NZMAX; ++k) {
        const int kk = k > nz - 1 ? nz - 1 : k;
        dp[k] = a.F[base + (size_t)kk * ny];
    }
    cp[0] = 0.0;
    dp[0] = {0.0, 0.0};
#pragma unroll
    for (int k = 1; k < NZMAX; ++k) {
        if (k < nz - 1) {
            const double al = cf[3 * k], be = cf[3 * k + 1] + hs, ga = cf[3 * k + 2];
            const double2 f = dp[k];
            // row 1 sees w_0 = 0 (cp[0] = dp[0] = 0); the last interior row's γ term meets w_{nz-1} = 0
            const double rm = 1.0 / (be - al * cp[k - 1]);
            cp[k] = (k < nz - 2) ? ga * rm : 0.0;
            dp[k] = {(f.x - al * dp[k - 1].x) * rm, (f.y - al * dp[k - 1].y) * rm};
        }
    }
    double2 x = {0.0, 0.0};
    a.W[base + (size_t)(nz - 1) * ny] = cpx<double>{0.0, 0.0};
#pragma unroll
    for (int k = NZMAX - 1; k >= 1; --k) {
        if (k < nz - 1) {
            x = {dp[k].x - cp[k] * x.x, dp[k].y - cp[k] * x.y};
            a.W[base + (size_t)k * ny] = x;
        }
    }
    a.W[base] = cpx<double>{0.0, 0.0};
}

// w = w* / (1 - ε (Φxx + Φyy)), in place
static __global__ __launch_bounds__(256) void k_omega_final(double* __restrict__ w, const double* __restrict__ pxx,
                                                            const double* __restrict__ pyy, double eps, long n3) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n3) return;
    w[i] = w[i] / (1.0 - eps * (pxx[i] + pyy[i]));
}

}  // namespace ssg
