// HIP kernels of the SSG surface time step.
//
// Layout in HBM
//   real 2D fields      : a[x + nx*y]                 (the reference's column-major order)
//   velocity pair w     : w[x + nx*y] = (u, v)        (one complex-typed load per point)
//   spectral planes     : P[kx*pitch + y], kx in [0, ncols), y in [0, ny)
//       after the row pass the y axis is still physical; after the column
//       pass it holds ky in FFT order.  Only kx <= nx/2 is stored (the real
//       transform is taken along the contiguous x axis) and only the first
//       `ncols` columns that the consumer can see are written.
//
// Three passes make a spectral operator E[S] f = irfft2(S * rfft2 f):
//   rows_fwd : per row, x-FFT of two packed real lines, split, store [kx][y]
//   cols     : per kx column, y-FFT, symbol, inverse y-FFT
//   rows_inv : per row pair, inverse x-FFT (c2r via two-for-one packing) and a
//              fused epilogue (stage update, next-stage fluxes + x-FFT)
// The symbols are the full-plane operators that the reference's r2c-along-y
// transforms realise (see `sym_*` below), so results match the reference's
// irfft2(S .* rfft2(f)) chains without reproducing its layout.
#pragma once
#include "fft_line.hpp"

namespace ssg {

enum { SRC_FLUX = 0, SRC_PAIR = 1, SRC_W = 2 };
enum { MODE_PAIR = 0, MODE_PLANES = 1 };
enum {
    EPI_STORE = 0,   // out_r[row] = value
    EPI_STORE_W = 1, // w[row] = (value0, value1)   (MODE_PLANES)
    EPI_RK1 = 2,
    EPI_RK2 = 3,
    EPI_RK3 = 4,
    EPI_AB2 = 5,
    EPI_SET = 6,     // b = value        (filter, FP64)
    EPI_ADD = 7,     // b += value       (filter, defect form)
    // discrete adjoint of the RK3 stage chain (MODE_PLANES: the two planes are
    // Dx M s and Dy M s of the stage input s; g = gscale (u Gx + v Gy) = gscale T^T s)
    EPI_ADJ3 = 8,    // acc = lam + g;  s = c2 g
    EPI_ADJ2 = 9,    // acc += g;       s = c3 lam + c1 g
    EPI_ADJ1 = 10,   // lam = acc + g
    // coupled-step adjoint (MODE_PLANES): out_r = u A + v B (the frozen-w
    // transpose), w_out = (b A, b B) (the adjoint of the velocities)
    EPI_ADJW = 11,
    // transposed tendency alone (MODE_PLANES): out_r = u A + v B, or 2 v B for
    // the AB2 call site's aliased tendency
    EPI_ADJG = 12
};
enum {
    OP_TEND = 0,     // -M (Dx Fx + Dy Fy)
    OP_TEND_Q = 1,   // -2 M Dy Fy        (AB2 buffer aliasing of the reference)
    OP_JAC = 2,      // -(Dx Fx + Dy Fy)  (compute_jacobian!)
    OP_VEL = 3,      // U = -Dy P, V = Dx P
    OP_FILT = 4,     // sigma P  (or (sigma-1) P)
    OP_FWD = 5,      // y-FFT only
    OP_INV = 6,      // inverse y-FFT only
    OP_DX = 7,       // Dx P
    OP_DY = 8,       // Dy P
    OP_MASK = 9,     // M P
    OP_MA_DERIV = 10,  // Monge-Ampere: P -> (Pxx, Pyy, Pxy), P spectral
    OP_MA_DIFF = 11,   // Monge-Ampere: y-FFT of D, residual sums against the previous D
    OP_VEL_SPEC = 12,  // U = -Dy P, V = Dx P with P already spectral in y
    OP_ADJ = 13,       // A = Dx M P, B = Dy M P (transposed tendency, spectral part)
    OP_PHIBAR = 14,    // y-FFT of (U, V) -> Dy U - Dx V, spectral (adjoint of u = -Dy φ, v = Dx φ)
    OP_MA_ADJ = 15,    // y-FFT of three products -> Dxx P1 + Dyy P2 - 2 Dxy P3 on interior levels, residual sums
    OP_SETB = 16,      // M P with the k = 0 mode removed (set_b!'s zero mean and dealias)
    OP_DIVQ = 17       // y-FFT of (Q₁, Q₂) -> -2ε (Dx Q₁ + Dy Q₂), left spectral (ω equation)
};

struct Grid {
    int nx, ny;
    int pitch;      // y pitch of the spectral planes (elements)
    int K;          // nx / 3
    int Ky;         // 2 * (ny/2 + 1) / 3
    double dkx, dky;
};

// ---------------------------------------------------------------------------
// Full-plane symbols.  c = kx column (0..nx/2), s = signed ky index in FFT
// order (row r -> r < ny/2 ? r : r - ny, so the y-Nyquist is -ny/2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int signed_k(int r, int n) { return r < n / 2 ? r : r - n; }

// ddx_2d! (src/transforms.jl:135-150) through irfft2: i*kx, except on the
// x-Nyquist column where the reference's kx = -nx/2 meets the Hermitian
// extension of its y-half-plane: +i*kN for ky > 0, -i*kN for ky < 0, 0 on
// ky = 0 and on the y-Nyquist row.
template <typename R>
__device__ __forceinline__ R sym_dx(const Grid& g, int c, int s) {
    if (2 * c < g.nx) return (R)(c * g.dkx);
    const R kn = (R)(-(g.nx / 2) * g.dkx);
    if (s > 0) return kn;
    if (s < 0 && s != -(g.ny / 2)) return -kn;
    return (R)0;
}
// ddy_2d! (src/transforms.jl:152-166) through irfft2: i*ky, zero on the
// y-Nyquist row (its i*ky*f is discarded by the c2r along y).
template <typename R>
__device__ __forceinline__ R sym_dy(const Grid& g, int s) {
    if (2 * s == -g.ny) return (R)0;
    return (R)(s * g.dky);
}
// twothirds_mask (src/domain.jl:498-514) seen through irfft2: kept kx are
// [-K, K-1] on the stored ky >= 0 half, so ky < 0 keeps kx in [-K+1, K] and
// the ky = 0 line keeps |kx| < K at weight 1 and |kx| = K at weight 1/2.
template <typename R>
__device__ __forceinline__ R sym_mask(const Grid& g, int c, int s) {
    if (s > 0) return (s <= g.Ky - 1 && c <= g.K - 1) ? (R)1 : (R)0;
    if (s < 0) return (-s <= g.Ky - 1 && c <= g.K) ? (R)1 : (R)0;
    if (c <= g.K - 1) return (R)1;
    return c == g.K ? (R)0.5 : (R)0;
}
// d2/dxdy through irfft2 (ddx then ddy, src/poisson.jl:190-193): -kx*ky;
// zero on the y-Nyquist row except on the x-Nyquist column, where the
// reference's kx = -nx/2 gives +(nx/2)|ky| on both ky half planes.
template <typename R>
__device__ __forceinline__ R sym_dxy(const Grid& g, int c, int s) {
    if (2 * c < g.nx) return 2 * s == -g.ny ? (R)0 : (R)(-(c * g.dkx) * (s * g.dky));
    return (R)((g.nx / 2) * g.dkx * (s < 0 ? -s : s) * g.dky);
}
// d2/dx2 and d2/dy2 through irfft2: -kx^2, -ky^2 (Nyquist entries included)
template <typename R>
__device__ __forceinline__ R sym_dxx(const Grid& g, int c) {
    const double k = (2 * c < g.nx ? c : g.nx / 2) * g.dkx;
    return (R)(-k * k);
}
template <typename R>
__device__ __forceinline__ R sym_dyy(const Grid& g, int s) {
    const double k = s * g.dky;
    return (R)(-k * k);
}

// exponential_transfer (src/filter.jl:119-126) at k_norm = |k| / kmax
template <typename R>
__device__ __forceinline__ R sym_filter(const Grid& g, int c, int s, R strength, int order, R cutoff,
                                        R inv_kmax) {
    const R kx = (R)(c * g.dkx), ky = (R)(s * g.dky);
    const R kn = sqrt(kx * kx + ky * ky) * inv_kmax;
    if (kn <= cutoff) return (R)1;
    const R ratio = (kn - cutoff) / ((R)1 - cutoff);
    R p = (R)1;
    for (int i = 0; i < order; ++i) p *= ratio;
    return exp(-strength * p);
}

// ---------------------------------------------------------------------------
// Row pass, forward: each line packs two real sequences a + i*b, takes the
// x-FFT and stores the two half spectra (kx < ncols) to [kx][y] planes.
//   SRC_FLUX: line l <-> row y0+l, a = u*b, b = v*b  -> (PA, PB) at row y0+l
//   SRC_PAIR: line l <-> rows (y0+l, y0+G+l) of f   -> PA at both rows
//   SRC_W   : line l <-> row y0+l, a = w.x, b = w.y   -> (PA, PB) at row y0+l
// ---------------------------------------------------------------------------
template <typename R, typename S>
struct RowsFwdArgs {
    const S* f;            // SRC_PAIR source, or the buoyancy for SRC_FLUX
    const cpx<R>* w;       // (u, v) for SRC_FLUX and SRC_W
    cpx<R>* PA;
    cpx<R>* PB;
    int ncols;
    int G;                 // lines per block
    int nblk;              // grid size (for the XCD-aware block order)
    const cpx<R>* twx;
    Grid g;
};

// Blocks are dealt to the 8 XCDs round robin; renumber them so that the
// blocks resident on one XCD own neighbouring rows (shared L2 lines).
__device__ __forceinline__ int xcd_block(int b, int nblk) {
    if ((nblk & 7) != 0) return b;
    return (b & 7) * (nblk >> 3) + (b >> 3);
}

template <typename R, int N, int E>
__device__ __forceinline__ void store_split(cpx<R>* lds, int G, int ncols, cpx<R>* PA, cpx<R>* PB,
                                            int pitch, int rowA0, int rowB0) {
    using C = cpx<R>;
    constexpr int LE = lds_line_elems(N);
    const int total = ncols * G;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int kx = idx >> gsh(G), l = idx & (G - 1);
        const C* buf = lds + l * LE;
        const C z1 = buf[lpad(kx)];
        const C z2 = cconj(buf[lpad((N - kx) & (N - 1))]);
        const C a = C{(z1.x + z2.x) * (R)0.5, (z1.y + z2.y) * (R)0.5};
        // (z1 - z2) / (2i)
        const C d = C{(z1.y - z2.y) * (R)0.5, -(z1.x - z2.x) * (R)0.5};
        PA[(size_t)kx * pitch + rowA0 + l] = a;
        PB[(size_t)kx * pitch + rowB0 + l] = d;
    }
}

// store_split over buffers at a line stride (buffer of line l at l*lstride)
template <typename R, int N, int E>
__device__ __forceinline__ void store_split_stride(cpx<R>* lds, int G, int lstride, int ncols, cpx<R>* PA,
                                                   cpx<R>* PB, int pitch, int rowA0, int rowB0) {
    using C = cpx<R>;
    constexpr int LE = lds_line_elems(N);
    const int total = ncols * G;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int kx = idx >> gsh(G), l = idx & (G - 1);
        const C* buf = lds + l * lstride * LE;
        const C z1 = buf[lpad(kx)];
        const C z2 = cconj(buf[lpad((N - kx) & (N - 1))]);
        PA[(size_t)kx * pitch + rowA0 + l] = C{(z1.x + z2.x) * (R)0.5, (z1.y + z2.y) * (R)0.5};
        PB[(size_t)kx * pitch + rowB0 + l] = C{(z1.y - z2.y) * (R)0.5, -(z1.x - z2.x) * (R)0.5};
    }
}

template <typename R, typename S, int N, int E, int SRC, bool PP = false>
__device__ __forceinline__ void rows_fwd_body(const RowsFwdArgs<R, S>& a, const int blk) {
    using C = cpx<R>;
    using F = LineFFT<R, N, E, PP>;
    constexpr int T = F::T;
    constexpr int LE = lds_line_elems(N);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int G = a.G;
    const int nx = a.g.nx;
    typename F::Twiddles tw;
    tw.load(t, a.twx);
    C v[E];
    int rowA0, rowB0;
    if (SRC == SRC_FLUX || SRC == SRC_W) {
        const int y = blk * G + l;
        rowA0 = rowB0 = blk * G;
        const size_t base = (size_t)nx * y;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int x = t + T * m;
            const C uv = a.w[base + x];
            if (SRC == SRC_W) {
                v[m] = uv;
            } else {
                const R bb = (R)a.f[base + x];
                v[m] = C{uv.x * bb, uv.y * bb};
            }
        }
    } else {
        const int yA = blk * 2 * G + l, yB = yA + G;
        rowA0 = blk * 2 * G;
        rowB0 = rowA0 + G;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int x = t + T * m;
            v[m] = C{(R)a.f[(size_t)nx * yA + x], (R)a.f[(size_t)nx * yB + x]};
        }
    }
    C* buf = lds + l * LE;
    F::template run<-1>(v, buf, t, tw, G * LE);  // ping-pong partners after the G line buffers
    F::to_lds(v, buf, t);
    store_split<R, N, E>(lds, G, a.ncols, a.PA, SRC == SRC_PAIR ? a.PA : a.PB, a.g.pitch, rowA0, rowB0);
}

template <typename R, typename S, int N, int E, int SRC, bool PP>
__global__ __launch_bounds__(256) void k_rows_fwd(RowsFwdArgs<R, S> a) {
    rows_fwd_body<R, S, N, E, SRC, PP>(a, xcd_block(blockIdx.x, a.nblk));
}

// Access to the Monge-Ampère planes.  With `nt` set (planes larger than the
// caches: each pass streams them once) loads and stores are nontemporal;
// otherwise plain, so planes that fit the caches stay there between passes.
template <typename R> struct vec2_t;
template <> struct vec2_t<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct vec2_t<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename R>
__device__ __forceinline__ cpx<R> ld_plane(const cpx<R>* p, bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt) {
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
        return cpx<R>{v.x, v.y};
    }
    return *p;
}
template <typename R>
__device__ __forceinline__ void st_plane(cpx<R>* p, cpx<R> x, bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt)
        __builtin_nontemporal_store(V{x.x, x.y}, reinterpret_cast<V*>(p));
    else
        *p = x;
}
// Row-tiled Monge-Ampère planes (the derivative planes and D between the
// column and row passes): element (kx, row r = level*ny + y) of a plane of
// kxn columns at ((r >> s) * kxn + kx) << s | (r & (2^s - 1)), s = log2 of the
// tile's rows.  A row-pass block owns one tile, [kx][row in tile], a single
// contiguous run; a column-pass block of consecutive kx writes runs of
// (lines x tile rows) elements.
__device__ __forceinline__ long ma_tile_off(long r, int kx, int kxn, int s) {
    return (((r >> s) * kxn + kx) << s) | (r & ((1L << s) - 1));
}

// E elements p[T*m], m = 0..E-1, in one straight run of loads (stores): the
// access-mode branch sits outside the unrolled loop, so no load waits for
// the one before it.
template <typename R, int E>
__device__ __forceinline__ void ld_plane_n(cpx<R> (&v)[E], const cpx<R>* p, int T, bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + T * m));
            v[m] = cpx<R>{w.x, w.y};
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = p[T * m];
    }
}
// the same at element offsets off[m]
template <typename R, int E>
__device__ __forceinline__ void ld_plane_at(cpx<R> (&v)[E], const cpx<R>* p, const long (&off)[E], bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + off[m]));
            v[m] = cpx<R>{w.x, w.y};
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = p[off[m]];
    }
}
template <typename R, int E>
__device__ __forceinline__ void st_plane_at(cpx<R>* p, const cpx<R> (&v)[E], const long (&off)[E], bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt) {
#pragma unroll
        for (int m = 0; m < E; ++m) __builtin_nontemporal_store(V{v[m].x, v[m].y}, reinterpret_cast<V*>(p + off[m]));
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) p[off[m]] = v[m];
    }
}
template <typename R, int E>
__device__ __forceinline__ void st_plane_n(cpx<R>* p, const cpx<R> (&v)[E], int T, bool nt) {
    using V = typename vec2_t<R>::type;
    if (nt) {
#pragma unroll
        for (int m = 0; m < E; ++m) __builtin_nontemporal_store(V{v[m].x, v[m].y}, reinterpret_cast<V*>(p + T * m));
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) p[T * m] = v[m];
    }
}

// ---------------------------------------------------------------------------
// Column pass: per kx column, y-FFT -> symbol -> inverse y-FFT.
// ---------------------------------------------------------------------------
template <typename R>
struct ColsArgs {
    const cpx<R>* inA;
    const cpx<R>* inB;
    const cpx<R>* inC;     // OP_MA_ADJ third input
    cpx<R>* outA;
    cpx<R>* outB;
    int ncols;
    int G;
    R scale;
    R fstrength, fcutoff, inv_kmax;
    int forder;
    int fdelta;            // filter returns (sigma - 1) * P
    const cpx<R>* twy;
    Grid g;
    long in_stride;         // elements between the lines of inA (0: ny, contiguous lines)
    int nt;                 // OP_MA_DERIV / OP_MA_DIFF: nontemporal plane access (ld_plane)
    int yt_log2;            // OP_MA_DERIV out / OP_MA_DIFF in: log2 of the row tile of those planes
    // Monge-Ampere extensions (lines are (kx, level) pairs, kx = line / nlev)
    int nlev;
    cpx<R>* outC;
    const cpx<double>* B2;  // 2D forcing spectrum [kx][ky] (variant A body force)
    double eps;
    int body;               // 1: B2 is the interior right-hand side on every level
    double* partial;        // per-block (sum |eps dD|^2, sum |rhs|^2) over interior levels
                            // (OP_MA_ADJ: sum |rhs|^2 is that of B2 on the top level)
};

// per-block sums of W values into out[blockIdx.x * W + k]; blockDim is a
// power of two <= 256 (tree reduction in LDS, no partial-wave shuffles)
template <int W>
__device__ __forceinline__ void block_sum_g(double (&v)[W], double* out, int bidx) {
    __shared__ double red_g[W][256];
    const int tid = threadIdx.x, nt = blockDim.x;
#pragma unroll
    for (int k = 0; k < W; ++k) red_g[k][tid] = v[k];
    __syncthreads();
    for (int st = nt >> 1; st > 0; st >>= 1) {
        if (tid < st) {
#pragma unroll
            for (int k = 0; k < W; ++k) red_g[k][tid] += red_g[k][tid + st];
        }
        __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < W; ++k) out[(size_t)bidx * W + k] = red_g[k][0];
    }
}

// LDS line buffers per column line for each operation
__host__ __device__ constexpr int cols_nbuf(int op) {
    return (op == OP_MA_DERIV || op == OP_MA_ADJ)
               ? 3
               : ((op == OP_TEND || op == OP_TEND_Q || op == OP_JAC || op == OP_VEL || op == OP_VEL_SPEC ||
                   op == OP_ADJ || op == OP_PHIBAR || op == OP_DIVQ)
                      ? 2
                      : 1);
}

template <typename R, int M, int E, int OP, bool PP = false>
__device__ __forceinline__ void cols_body(const ColsArgs<R>& a, const int bidx) {
    using C = cpx<R>;
    using F = LineFFT<R, M, E, PP>;
    constexpr int T = F::T;
    constexpr int LE = lds_line_elems(M);
    constexpr int NBUF = cols_nbuf(OP);
    const int alt = a.G * NBUF * LE;  // ping-pong partners after the block's line buffers
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int line = bidx * a.G + l;
    const int pitch = a.g.ny;                         // lines are contiguous runs of ny
    constexpr bool MA_PLANES = (OP == OP_MA_DERIV || OP == OP_MA_DIFF);
    // Monge-Ampère passes: the lines of a block are G consecutive kx of one
    // level.  OP_MA_DERIV runs them level-major (line = level * kxn + kx);
    // OP_MA_DIFF runs blocks kx-group-major (block = group * nlev + level,
    // kx = group * G + l), so the nlev blocks of a group are consecutive
    // (k_cols places them on one XCD) and share the forcing columns they
    // read through that XCD's L2.  Lines past the end stand in for kx 0 of
    // level 0 (in bounds, nothing stored).
    const int kxn = (a.g.nx >> 1) + 1;
    const int dgrp = OP == OP_MA_DIFF ? bidx / a.nlev : 0;
    const bool live = OP == OP_MA_DIFF ? dgrp * a.G + l < kxn : line < a.ncols;
    const int mlev = OP == OP_MA_DIFF ? (live ? bidx - dgrp * a.nlev : 0)
                                      : (MA_PLANES ? (live ? line : 0) / kxn : 0);
    const int c = OP == OP_MA_DIFF ? (live ? dgrp * a.G + l : 0)
                                   : (MA_PLANES ? (live ? line : 0) - mlev * kxn
                                                : (a.nlev > 1 ? line / a.nlev : line));  // kx column of this line
    // offset of this (kx, level) line in the [kx][level][y] planes, and of its
    // elements in the row-tiled planes
    const size_t mline = ((size_t)c * a.nlev + mlev) * pitch;
    long toff[MA_PLANES ? E : 1];
    if constexpr (MA_PLANES) {
        const long r0 = (long)mlev * pitch + t;
        if ((T >> a.yt_log2) > 0) {
            // T rows are whole tiles: element m sits T * kxn entries after element 0
            const long o0 = ma_tile_off(r0, c, kxn, a.yt_log2);
#pragma unroll
            for (int m = 0; m < E; ++m) toff[m] = o0 + (long)m * T * kxn;
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) toff[m] = ma_tile_off(r0 + T * m, c, kxn, a.yt_log2);
        }
    }
    C* const buf1[1] = {lds + (l * NBUF) * LE};
    typename F::Twiddles tw;
    tw.load(t, a.twy);
    constexpr bool TWO_IN =
        (OP == OP_TEND || OP == OP_TEND_Q || OP == OP_JAC || OP == OP_PHIBAR || OP == OP_MA_ADJ || OP == OP_DIVQ);
    C vv[2][E];
    C (&va)[E] = vv[0];
    C (&vb)[E] = vv[1];
    const C* srcA = a.inA + (size_t)(live ? line : 0) * (a.in_stride ? (size_t)a.in_stride : (size_t)pitch);
    // OP_MA_DIFF: D̂_prev of this line, loaded with the input (it does not
    // depend on the transform)
    C prev[OP == OP_MA_DIFF ? E : 1];
    cpx<double> bfor[OP == OP_MA_DIFF ? E : 1];  // and the forcing of its kx column (body variant)
    if constexpr (OP == OP_MA_DERIV) {
        ld_plane_n<R, E>(va, a.inA + mline + t, T, a.nt);
    } else if constexpr (OP == OP_MA_DIFF) {
        ld_plane_at<R, E>(va, a.inA, toff, a.nt);
        ld_plane_n<R, E>(prev, a.outA + mline + t, T, a.nt);
        if (a.body) {
            const cpx<double>* b2p = a.B2 + (size_t)c * a.g.ny + t;
#pragma unroll
            for (int m = 0; m < E; ++m) bfor[m] = b2p[T * m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) va[m] = live ? srcA[t + T * m] : C{0, 0};
    }
    if constexpr (TWO_IN) {
        const C* srcB = a.inB + (size_t)(live ? line : 0) * pitch;
#pragma unroll
        for (int m = 0; m < E; ++m) vb[m] = live ? srcB[t + T * m] : C{0, 0};
    }
    auto store = [&](C* dst, const C (&v)[E]) {
        if (live) {
            if constexpr (OP == OP_MA_DERIV) {
                st_plane_at<R, E>(dst, v, toff, a.nt);
            } else if constexpr (OP == OP_MA_DIFF) {
                st_plane_n<R, E>(dst + mline + t, v, T, a.nt);
            } else {
                C* d = dst + (size_t)line * pitch;
#pragma unroll
                for (int m = 0; m < E; ++m) d[t + T * m] = v[m];
            }
        }
    };

    if constexpr (OP == OP_MA_ADJ) {
        // transpose of the linearised D operator, D'(Φ)ᵀ λ = Dxx(Φyy λ) +
        // Dyy(Φxx λ) - 2 Dxy(Φxy λ): the three products arrive x-spectral;
        // all three symbols are real and even, so each operator is its own
        // transpose.  Only interior levels carry D.
        C w3[3][E];
        const C* srcC = a.inC + (size_t)(live ? line : 0) * pitch;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            w3[0][m] = va[m];
            w3[1][m] = vb[m];
            w3[2][m] = live ? srcC[t + T * m] : C{0, 0};
        }
        C* const b3[3] = {lds + (l * NBUF) * LE, lds + (l * NBUF + 1) * LE, lds + (l * NBUF + 2) * LE};
        F::template run_n<-1, 3>(w3, b3, t, tw, alt);
        const int lev = a.nlev > 1 ? line - c * a.nlev : 0;
        const bool interior = live && lev >= 1 && lev <= a.nlev - 2;
        const bool top = live && lev == a.nlev - 1;
        const double wgt = (c == 0 || 2 * c == a.g.nx) ? 1.0 : 2.0;
        double acc[2] = {0.0, 0.0};
        C* dp = a.outA + (size_t)(live ? line : 0) * pitch;
        const cpx<double>* b2p = a.B2 + (size_t)c * a.g.ny;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = t + T * m;
            const int s = signed_k(y, a.g.ny);
            C q = C{0, 0};
            if (interior) {
                const R sxx = sym_dxx<R>(a.g, c), syy = sym_dyy<R>(a.g, s), sxy = sym_dxy<R>(a.g, c, s);
                q = C{sxx * w3[0][m].x + syy * w3[1][m].x - (R)2 * sxy * w3[2][m].x,
                      sxx * w3[0][m].y + syy * w3[1][m].y - (R)2 * sxy * w3[2][m].y};
                const C prev = dp[y];
                const double rx = a.eps * ((double)prev.x - (double)q.x);
                const double ry = a.eps * ((double)prev.y - (double)q.y);
                acc[0] += wgt * (rx * rx + ry * ry);
            }
            if (top) {
                const cpx<double> bb = b2p[y];
                acc[1] += wgt * (bb.x * bb.x + bb.y * bb.y);
            }
            if (live) dp[y] = q;
        }
        block_sum_g<2>(acc, a.partial, bidx);
        return;
    } else if constexpr (OP == OP_DIVQ) {
        // -2ε ∇·Q (src/extras/omega_eq.jl:400-402), left spectral: -2ε (i kx Q̂₁ + i ky Q̂₂),
        // without the x- and y-Nyquist bins (where the transform conventions of
        // that file and of the step path disagree)
        C* const b2[2] = {lds + (l * NBUF) * LE, lds + (l * NBUF + 1) * LE};
        F::template run_n<-1, 2>(vv, b2, t, tw, alt);
        const R f = (R)(-2.0 * a.eps);
        const bool xnyq = 2 * c == a.g.nx;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int s = signed_k(t + T * m, a.g.ny);
            const C px = cmul_i(va[m], (R)(c * a.g.dkx)), py = cmul_i(vb[m], (R)(s * a.g.dky));
            const R g = (xnyq || 2 * s == -a.g.ny) ? (R)0 : f;
            va[m] = C{g * (px.x + py.x), g * (px.y + py.y)};
        }
        store(a.outA, va);
        return;
    } else if constexpr (OP == OP_PHIBAR) {
        // φ̄ = Dy U - Dx V (u = -Dy φ, v = Dx φ; Dx, Dy are antisymmetric),
        // left spectral: F(φ̄) = i sdy F(U) - i sdx F(V)
        C* const b2[2] = {lds + (l * NBUF) * LE, lds + (l * NBUF + 1) * LE};
        F::template run_n<-1, 2>(vv, b2, t, tw, alt);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int s = signed_k(t + T * m, a.g.ny);
            const C pu = cmul_i(va[m], sym_dy<R>(a.g, s)), pv = cmul_i(vb[m], sym_dx<R>(a.g, c, s));
            va[m] = C{pu.x - pv.x, pu.y - pv.y};
        }
        store(a.outA, va);
        return;
    } else if constexpr (OP == OP_MA_DERIV) {
        C w3[3][E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int s = signed_k(t + T * m, a.g.ny);
            const C p = cscale(va[m], a.scale);
            w3[0][m] = cscale(p, sym_dxx<R>(a.g, c));
            w3[1][m] = cscale(p, sym_dyy<R>(a.g, s));
            w3[2][m] = cscale(p, sym_dxy<R>(a.g, c, s));
        }
        C* const b3[3] = {lds + (l * NBUF) * LE, lds + (l * NBUF + 1) * LE, lds + (l * NBUF + 2) * LE};
        F::template run_n<+1, 3>(w3, b3, t, tw, alt);
        store(a.outA, w3[0]);
        store(a.outB, w3[1]);
        store(a.outC, w3[2]);
        return;
    } else if constexpr (OP == OP_INV) {
        F::template run<+1>(va, buf1[0], t, tw, alt);
#pragma unroll
        for (int m = 0; m < E; ++m) va[m] = cscale(va[m], a.scale);
        store(a.outA, va);
        return;
    } else {
        C* const b2[2] = {lds + (l * NBUF) * LE, lds + (l * NBUF + (NBUF > 1 ? 1 : 0)) * LE};
        // forward y-transform(s)
        if constexpr (OP != OP_VEL_SPEC) {
            if constexpr (TWO_IN) {
                F::template run_n<-1, 2>(vv, b2, t, tw, alt);
            } else {
                F::template run<-1>(va, buf1[0], t, tw, alt);
            }
        }
        if constexpr (OP == OP_FWD) {
            store(a.outA, va);
            return;
        } else if constexpr (OP == OP_MA_DIFF) {
            // va = D-hat of this (kx, level) line; the residual of the iterate
            // that produced it is eps * (D_prev - D) on interior levels
            const bool interior = live && mlev >= 1 && mlev <= a.nlev - 2;
            const double wgt = (c == 0 || 2 * c == a.g.nx) ? 1.0 : 2.0;
            double acc[2] = {0.0, 0.0};
            if (interior) {
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const double rx = a.eps * ((double)prev[m].x - (double)va[m].x);
                    const double ry = a.eps * ((double)prev[m].y - (double)va[m].y);
                    double hx = a.eps * (double)va[m].x, hy = a.eps * (double)va[m].y;
                    if (a.body) {
                        hx += bfor[m].x;
                        hy += bfor[m].y;
                    }
                    acc[0] += wgt * (rx * rx + ry * ry);
                    acc[1] += wgt * (hx * hx + hy * hy);
                }
            }
            store(a.outA, va);
            block_sum_g<2>(acc, a.partial, bidx);
            return;
        } else if constexpr (OP == OP_ADJ) {
            // transpose of the tendency's spectral part: the symbols are
            // Hermitian, so E[S]^T = E[conj S]; M is real, Dx and Dy are
            // imaginary: (-M (Dx U + Dy V))^T = (U Dx + V Dy) M
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int s = signed_k(t + T * m, a.g.ny);
                const C p = cscale(va[m], sym_mask<R>(a.g, c, s) * a.scale);
                vb[m] = cmul_i(p, sym_dy<R>(a.g, s));
                va[m] = cmul_i(p, sym_dx<R>(a.g, c, s));
            }
            F::template run_n<+1, 2>(vv, b2, t, tw, alt);
            store(a.outA, va);
            store(a.outB, vb);
            return;
        } else if constexpr (OP == OP_VEL || OP == OP_VEL_SPEC) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int s = signed_k(t + T * m, a.g.ny);
                const R dy = sym_dy<R>(a.g, s), dx = sym_dx<R>(a.g, c, s);
                vb[m] = cmul_i(va[m], dx * a.scale);   // V = Dx P
                va[m] = cmul_i(va[m], -dy * a.scale);  // U = -Dy P
            }
            F::template run_n<+1, 2>(vv, b2, t, tw, alt);
            store(a.outA, va);
            store(a.outB, vb);
            return;
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int s = signed_k(t + T * m, a.g.ny);
                if constexpr (TWO_IN) {
                    const R dy = sym_dy<R>(a.g, s);
                    C r;
                    if (OP == OP_TEND_Q) {
                        r = cmul_i(vb[m], dy * (R)2);
                    } else {
                        const R dx = sym_dx<R>(a.g, c, s);
                        // i*dx*A + i*dy*B
                        r = C{-(dx * va[m].y + dy * vb[m].y), dx * va[m].x + dy * vb[m].x};
                    }
                    R f = -a.scale;
                    if (OP != OP_JAC) f *= sym_mask<R>(a.g, c, s);
                    va[m] = cscale(r, f);
                } else if constexpr (OP == OP_FILT) {
                    R sg = sym_filter<R>(a.g, c, s, a.fstrength, a.forder, a.fcutoff, a.inv_kmax);
                    if (a.fdelta) sg -= (R)1;
                    va[m] = cscale(va[m], sg * a.scale);
                } else if constexpr (OP == OP_DX) {
                    va[m] = cmul_i(va[m], sym_dx<R>(a.g, c, s) * a.scale);
                } else if constexpr (OP == OP_DY) {
                    va[m] = cmul_i(va[m], sym_dy<R>(a.g, s) * a.scale);
                } else if constexpr (OP == OP_MASK) {
                    va[m] = cscale(va[m], sym_mask<R>(a.g, c, s) * a.scale);
                } else if constexpr (OP == OP_SETB) {
                    const R mk = (c == 0 && s == 0) ? (R)0 : sym_mask<R>(a.g, c, s);
                    va[m] = cscale(va[m], mk * a.scale);
                }
            }
            F::template run<+1>(va, buf1[0], t, tw, alt);
            store(a.outA, va);
        }
    }
}

template <typename R, int M, int E, int OP, bool PP>
__global__ __launch_bounds__(256) void k_cols(ColsArgs<R> a) {
    cols_body<R, M, E, OP, PP>(a, OP == OP_MA_DIFF ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Row pass, inverse + epilogue.
//   MODE_PAIR  : line l <-> rows (y0+l, y0+G+l) of plane P0
//   MODE_PLANES: line l <-> row y0+l of planes (P0, P1)
// The c2r along x ignores the imaginary parts of the kx = 0 and kx = nx/2
// bins (the same convention the reference applies along y).
// ---------------------------------------------------------------------------
template <typename R>
struct RowsInvArgs {
    const cpx<R>* P0;
    const cpx<R>* P1;
    int ncols_in;
    int G;
    int nblk;
    // epilogue
    double* b;             // prognostic buoyancy (FP64 in every precision)
    R* k1;                 // RK3 first-stage tendency
    R* fold;               // AB2 previous tendency
    R* out_r;              // EPI_STORE
    cpx<R>* w_out;         // EPI_STORE_W
    double* bstage;        // RK1/RK2: also store the stage buoyancy (coupled mode)
    double* lam;           // adjoint: the adjoint state
    double* acc;           // adjoint: accumulated lambda of the step
    double* sbuf;          // adjoint: next stage's input
    double c1, c2, c3;     // stage coefficients
    double gscale;         // adjoint: scale of T^T s
    int adj_quirk;         // EPI_ADJG: transpose of the aliased AB2 tendency
    int first;             // AB2: forward Euler on step 0
    // next-stage flux emission
    int emit;
    const cpx<R>* w;
    cpx<R>* Fx;
    cpx<R>* Fy;
    int ncols_out;
    const cpx<R>* twx;
    Grid g;
};

template <typename R, int N, int E, int MODE, int EPI, bool PP = false>
__device__ __forceinline__ void rows_inv_body(const RowsInvArgs<R>& a, const int blk) {
    using C = cpx<R>;
    using F = LineFFT<R, N, E, PP>;
    constexpr int T = F::T;
    constexpr int LE = lds_line_elems(N);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int G = a.G;
    const int pitch = a.g.pitch, nx = a.g.nx;
    const int rows_per_blk = (MODE == MODE_PAIR) ? 2 * G : G;
    const int y0 = blk * rows_per_blk;
    const int rowA = y0 + l;
    const int rowB = (MODE == MODE_PAIR) ? y0 + G + l : rowA;
    const size_t bA = (size_t)nx * rowA, bB = (size_t)nx * rowB;
    typename F::Twiddles tw;
    tw.load(t, a.twx);

    // epilogue operands do not depend on the transform: load them first
    constexpr bool STATE = (EPI >= EPI_RK1 && EPI <= EPI_ADD);
    constexpr bool NEED_B = (EPI == EPI_RK1 || EPI == EPI_RK2 || EPI == EPI_RK3 || EPI == EPI_AB2 || EPI == EPI_ADD);
    double pbA[STATE ? E : 1], pbB[STATE ? E : 1];
    R pkA[STATE ? E : 1], pkB[STATE ? E : 1];
    C pwA[STATE ? E : 1], pwB[STATE ? E : 1];
    if constexpr (STATE) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int x = t + T * m;
            if (NEED_B) {
                pbA[m] = a.b[bA + x];
                pbB[m] = a.b[bB + x];
            }
            if (EPI == EPI_RK3) {
                pkA[m] = a.k1[bA + x];
                pkB[m] = a.k1[bB + x];
            }
            if (EPI == EPI_AB2) {
                pkA[m] = a.fold[bA + x];
                pkB[m] = a.fold[bB + x];
            }
            if (a.emit) {
                pwA[m] = a.w[bA + x];
                pwB[m] = a.w[bB + x];
            }
        }
    }

    // cooperative gather of the packed spectra into the per-line buffers;
    // all loads are issued before any LDS store (bounded, unrolled loop)
    {
        const int nc = a.ncols_in;
        const int total = nc * G;
        constexpr int MAXIT = (N / 2 + 1 + T - 1) / T;
        C q0[MAXIT], q1[MAXIT];
#pragma unroll
        for (int it = 0; it < MAXIT; ++it) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (idx < total) {
                const int kx = idx >> gsh(G), ll = idx & (G - 1);
                const size_t o = (size_t)kx * pitch + y0 + ll;
                if (MODE == MODE_PAIR) {
                    q0[it] = a.P0[o];
                    q1[it] = a.P0[o + G];
                } else {
                    q0[it] = a.P0[o];
                    q1[it] = a.P1[o];
                }
            }
        }
#pragma unroll
        for (int it = 0; it < MAXIT; ++it) {
            const int idx = threadIdx.x + it * (int)blockDim.x;
            if (idx < total) {
                const int kx = idx >> gsh(G), ll = idx & (G - 1);
                C p0 = q0[it], p1 = q1[it];
                C* bl = lds + ll * LE;
                const bool edge = (kx == 0) || (2 * kx == N);
                if (edge) {
                    p0.y = 0;
                    p1.y = 0;
                }
                bl[lpad(kx)] = C{p0.x - p1.y, p0.y + p1.x};
                if (!edge) bl[lpad(N - kx)] = C{p0.x + p1.y, p1.x - p0.y};
            }
        }
        // zero the band no input column covers: kx in [nc, N - nc]
        const int gap = N - 2 * nc + 1;
        if (gap > 0) {
            const int tz = gap * G;
            for (int idx = threadIdx.x; idx < tz; idx += blockDim.x) {
                const int q = idx >> gsh(G), ll = idx & (G - 1);
                lds[ll * LE + lpad(nc + q)] = C{0, 0};
            }
        }
    }
    C* buf = lds + l * LE;
    C v[E];
    F::template run_from_lds<+1>(v, buf, t, tw, 2 * G * LE);

    // epilogue
    R sA[E], sB[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int x = t + T * m;
        const R tA = v[m].x, tB = v[m].y;
        if constexpr (EPI == EPI_ADJ3 || EPI == EPI_ADJ2 || EPI == EPI_ADJ1) {
            // MODE_PLANES: tA = (Dx M s)(x), tB = (Dy M s)(x) of row A
            const C uv = a.w[bA + x];
            const double g = a.gscale * ((double)uv.x * (double)tA + (double)uv.y * (double)tB);
            if (EPI == EPI_ADJ3) {
                a.acc[bA + x] = a.lam[bA + x] + g;
                a.sbuf[bA + x] = a.c2 * g;
            } else if (EPI == EPI_ADJ2) {
                a.acc[bA + x] += g;
                a.sbuf[bA + x] = a.c3 * a.lam[bA + x] + a.c1 * g;
            } else {
                a.lam[bA + x] = a.acc[bA + x] + g;
            }
        } else if constexpr (EPI == EPI_ADJG) {
            const C uv = a.w[bA + x];
            const double g = a.adj_quirk ? 2.0 * (double)uv.y * (double)tB
                                         : (double)uv.x * (double)tA + (double)uv.y * (double)tB;
            a.out_r[bA + x] = (R)(a.gscale * g);
        } else if constexpr (EPI == EPI_ADJW) {
            const C uv = a.w[bA + x];
            const double bb = a.b[bA + x];
            a.out_r[bA + x] = (R)(a.gscale * ((double)uv.x * (double)tA + (double)uv.y * (double)tB));
            a.w_out[bA + x] = C{(R)(bb * (double)tA), (R)(bb * (double)tB)};
        } else if (EPI == EPI_STORE) {
            a.out_r[bA + x] = tA;
            a.out_r[bB + x] = tB;
        } else if (EPI == EPI_STORE_W) {
            a.w_out[bA + x] = C{tA, tB};
        } else if constexpr (STATE) {
            double nbA, nbB;
            if (EPI == EPI_RK1) {
                a.k1[bA + x] = tA;
                a.k1[bB + x] = tB;
                nbA = pbA[m] + a.c1 * (double)tA;
                nbB = pbB[m] + a.c1 * (double)tB;
            } else if (EPI == EPI_RK2) {
                nbA = pbA[m] + a.c1 * (double)tA;
                nbB = pbB[m] + a.c1 * (double)tB;
            } else if (EPI == EPI_RK3) {
                nbA = pbA[m] + a.c1 * ((double)pkA[m] + 3.0 * (double)tA);
                nbB = pbB[m] + a.c1 * ((double)pkB[m] + 3.0 * (double)tB);
                a.b[bA + x] = nbA;
                a.b[bB + x] = nbB;
            } else if (EPI == EPI_AB2) {
                if (a.first) {
                    nbA = pbA[m] + a.c1 * (double)tA;
                    nbB = pbB[m] + a.c1 * (double)tB;
                } else {
                    nbA = pbA[m] + a.c1 * (1.5 * (double)tA - 0.5 * (double)pkA[m]);
                    nbB = pbB[m] + a.c1 * (1.5 * (double)tB - 0.5 * (double)pkB[m]);
                }
                a.fold[bA + x] = tA;
                a.fold[bB + x] = tB;
                a.b[bA + x] = nbA;
                a.b[bB + x] = nbB;
            } else if (EPI == EPI_SET) {
                nbA = (double)tA;
                nbB = (double)tB;
                a.b[bA + x] = nbA;
                a.b[bB + x] = nbB;
            } else {  // EPI_ADD
                nbA = pbA[m] + (double)tA;
                nbB = pbB[m] + (double)tB;
                a.b[bA + x] = nbA;
                a.b[bB + x] = nbB;
            }
            if ((EPI == EPI_RK1 || EPI == EPI_RK2) && a.bstage) {
                a.bstage[bA + x] = nbA;
                a.bstage[bB + x] = nbB;
            }
            sA[m] = (R)nbA;
            sB[m] = (R)nbB;
        }
    }
    if constexpr (!STATE) return;
    if (!a.emit) return;  // uniform across the grid

    // next-stage fluxes of both rows, x-FFT of (u*b) + i(v*b), the two rows
    // in lockstep: row A in buffer l, row B in buffer G + l, so buffer j holds
    // row y0 + j and one store pass covers all 2G rows of the block
    static_assert(MODE == MODE_PAIR || !STATE, "state epilogues pair rows");
    C vv[2][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        vv[0][m] = C{pwA[m].x * sA[m], pwA[m].y * sA[m]};
        vv[1][m] = C{pwB[m].x * sB[m], pwB[m].y * sB[m]};
    }
    C* const bb[2] = {lds + l * LE, lds + (G + l) * LE};
    F::template run_n<-1, 2>(vv, bb, t, tw, 2 * G * LE);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) {
        bb[0][lpad(t + T * m)] = vv[0][m];
        bb[1][lpad(t + T * m)] = vv[1][m];
    }
    __syncthreads();
    store_split<R, N, E>(lds, 2 * G, a.ncols_out, a.Fx, a.Fy, pitch, y0, y0);
}

template <typename R, int N, int E, int MODE, int EPI, bool PP>
__global__ __launch_bounds__(256) void k_rows_inv(RowsInvArgs<R> a) {
    rows_inv_body<R, N, E, MODE, EPI, PP>(a, xcd_block(blockIdx.x, a.nblk));
}

// ---------------------------------------------------------------------------
// Layout conversion between this library's spectrum (kx <= nx/2 planes,
// full ky) and the reference's rfft2 array (nx, ny/2+1), x fastest.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_to_ref(const cpx<R>* __restrict__ P, double2* __restrict__ out, Grid g) {
    const int nyc = g.ny / 2 + 1;
    const long total = (long)g.nx * nyc;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nx), j = (int)(q / g.nx);
        cpx<R> z;
        if (2 * i <= g.nx) {
            z = P[(size_t)i * g.pitch + j];
        } else {
            z = cconj(P[(size_t)(g.nx - i) * g.pitch + ((g.ny - j) % g.ny)]);
        }
        out[q] = double2{(double)z.x, (double)z.y};
    }
}

// Reference half-plane -> Hermitian full plane in this library's layout:
// ky rows 0 and ny/2 take (S(i) + conj S(-i)) / 2, which is exactly what the
// reference's c2r along y keeps of them.
template <typename R>
__global__ void k_from_ref(const double2* __restrict__ in, cpx<R>* __restrict__ P, Grid g) {
    const int ncol = g.nx / 2 + 1;
    const long total = (long)ncol * g.ny;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q / g.ny), r = (int)(q % g.ny);
        const int cm = (g.nx - c) % g.nx;
        double2 z;
        if (r == 0 || 2 * r == g.ny) {
            const double2 p = in[(size_t)g.nx * r + c], m = in[(size_t)g.nx * r + cm];
            z = double2{0.5 * (p.x + m.x), 0.5 * (p.y - m.y)};
        } else if (2 * r < g.ny) {
            z = in[(size_t)g.nx * r + c];
        } else {
            const double2 p = in[(size_t)g.nx * (g.ny - r) + cm];
            z = double2{p.x, -p.y};
        }
        P[(size_t)c * g.pitch + r] = cpx<R>{(R)z.x, (R)z.y};
    }
}

// ---------------------------------------------------------------------------
// Reductions (FP64, two pass: per-block partials, summed on the host in a
// fixed order so results are deterministic).
// ---------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ void block_sum(double (&v)[W], double* out) {
    __shared__ double red[W][256 / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        double s = v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            double s = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[k][i];
            out[(size_t)blockIdx.x * W + k] = s;
        }
    }
}

// sum b, sum b^2, sum non-finite count
static __global__ __launch_bounds__(256) void k_stats(const double* __restrict__ b, long n, double* partial) {
    double v[3] = {0, 0, 0};
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double x = b[q];
        if (isfinite(x)) {
            v[0] += x;
            v[1] += x * x;
        } else {
            v[2] += 1.0;
        }
    }
    block_sum<3>(v, partial);
}

// parsevalsum2 weights (src/transforms.jl:239-261) on this library's planes:
// sum over ky in [0, ny/2] and kx in [0, nx/2] of w(kx) |P|^2, w = 1 at kx = 0
// and nx/2, 2 otherwise.
template <typename R>
__global__ __launch_bounds__(256) void k_parseval_ref(const cpx<R>* __restrict__ P, Grid g, double* partial) {
    double v[1] = {0};
    const int ncol = g.nx / 2 + 1, nrow = g.ny / 2 + 1;
    const long total = (long)ncol * nrow;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
        const int c = (int)(q / nrow), r = (int)(q % nrow);
        const cpx<R> z = P[(size_t)c * g.pitch + r];
        const double wgt = (c == 0 || 2 * c == g.nx) ? 1.0 : 2.0;
        v[0] += wgt * ((double)z.x * z.x + (double)z.y * z.y);
    }
    block_sum<1>(v, partial);
}

// ---------------------------------------------------------------------------
// small elementwise helpers
// ---------------------------------------------------------------------------
template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ in, B* __restrict__ out, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = (B)in[q];
}

// per-block Σ (u² + v²) of the interleaved velocities
static __global__ __launch_bounds__(256) void k_w_energy(const double2* __restrict__ w, long n, double* partial) {
    double v[1] = {0.0};
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double2 z = w[q];
        v[0] += z.x * z.x + z.y * z.y;
    }
    block_sum<1>(v, partial);
}

// per-block max of b (-inf for a block with no points; NaN propagates)
static __global__ __launch_bounds__(256) void k_bmax(const double* __restrict__ b, long n, double* partial) {
    __shared__ double red[256];
    double m = -INFINITY;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double x = b[q];
        m = (x > m || x != x) ? x : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const double o = red[threadIdx.x + st];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// per-block max |u| and max |v| of the interleaved velocities
static __global__ __launch_bounds__(256) void k_wmax(const double2* __restrict__ w, long n, double* partial) {
    __shared__ double red[2][256];
    double mu = 0.0, mv = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double2 z = w[q];
        mu = fmax(mu, fabs(z.x));
        mv = fmax(mv, fabs(z.y));
    }
    red[0][threadIdx.x] = mu;
    red[1][threadIdx.x] = mv;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + st]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red[0][0];
        partial[2 * blockIdx.x + 1] = red[1][0];
    }
}

// ca x + cb y + cc z of one plane element in FP64 with explicit fused
// multiply-adds and one rounding: every kernel that forms the quadratic
// warm-start extrapolation (k_extrap, k_ma_tridiag's extrapolating form)
// produces the same bits.
template <typename C>
__device__ __forceinline__ C extrap3(C x, double ca, C y, double cb, C z, double cc) {
    using R = typename cpx_real_t<C>::type;
    const double rx = fma(cc, (double)z.x, fma(cb, (double)y.x, ca * (double)x.x));
    const double ry = fma(cc, (double)z.y, fma(cb, (double)y.y, ca * (double)x.y));
    return C{(R)rx, (R)ry};
}
// ca x + cb y, as k_extrap forms it with no third entry
template <typename C>
__device__ __forceinline__ C extrap2(C x, double ca, C y, double cb) {
    using R = typename cpx_real_t<C>::type;
    const double rx = ca * (double)x.x + cb * (double)y.x, ry = ca * (double)x.y + cb * (double)y.y;
    return C{(R)rx, (R)ry};
}

// warm-start extrapolation: out = ca a + cb b + cc c (c may be null).  One
// element per thread (grid of ceil(n / 256) blocks); plane access as in
// ld_plane / st_plane.
// C = double2 or float2 (FP32 stage solves: FP64 arithmetic, one rounding on
// the store).
template <typename C>
__global__ __launch_bounds__(256) void k_extrap(C* __restrict__ out, const C* __restrict__ a, double ca,
                                                const C* __restrict__ b, double cb, const C* __restrict__ c,
                                                double cc, long n, int nt) {
    using R = typename cpx_real_t<C>::type;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const C x = ld_plane<R>(a + q, nt), y = ld_plane<R>(b + q, nt);
    if (c) {
        st_plane<R>(out + q, extrap3(x, ca, y, cb, ld_plane<R>(c + q, nt), cc), nt);
        return;
    }
    st_plane<R>(out + q, extrap2(x, ca, y, cb), nt);
}

// out = a x + b y (y may be null)
// (out may alias x or y: each element is read before it is written)
static __global__ void k_lincomb(double* out, double ca, const double* x, double cb, const double* y, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = ca * x[q] + (y ? cb * y[q] : 0.0);
}

// complex array precision change (FP32 <-> FP64)
template <typename A, typename B>
__global__ void k_cpx_convert(const cpx<A>* __restrict__ in, cpx<B>* __restrict__ out, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = cpx<B>{(B)in[q].x, (B)in[q].y};
}

template <typename R>
__global__ void k_w_convert(const double2* __restrict__ in, cpx<R>* __restrict__ out, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = cpx<R>{(R)in[q].x, (R)in[q].y};
}

template <typename R>
__global__ void k_w_split(const cpx<R>* __restrict__ w, double* __restrict__ u, double* __restrict__ v, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        u[q] = (double)w[q].x;
        v[q] = (double)w[q].y;
    }
}

}  // namespace ssg

namespace ssg {

// ---------------------------------------------------------------------------
// Batched (ensemble) launches: one launch covers every member; block b works
// on work[b] = (member, block within that member) with the member's own
// argument struct, and dispatches on the member's line length.  Every member
// uses G = 256 / T lines per block, so the block shape is the same for all.
// ---------------------------------------------------------------------------
template <typename R, int E, int OP>
__global__ __launch_bounds__(256) void k_cols_batch(const ColsArgs<R>* __restrict__ args, const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const ColsArgs<R>& a = args[wk.x];
    switch (a.g.ny) {
        case 128: cols_body<R, 128, E, OP>(a, wk.y); break;
        case 256: cols_body<R, 256, E, OP>(a, wk.y); break;
        case 512: cols_body<R, 512, E, OP>(a, wk.y); break;
        case 1024: cols_body<R, 1024, E, OP>(a, wk.y); break;
        default: break;
    }
}

template <typename R, int E, int SRC>
__global__ __launch_bounds__(256) void k_rows_fwd_batch(const RowsFwdArgs<R, double>* __restrict__ args,
                                                        const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const RowsFwdArgs<R, double>& a = args[wk.x];
    switch (a.g.nx) {
        case 128: rows_fwd_body<R, double, 128, E, SRC>(a, wk.y); break;
        case 256: rows_fwd_body<R, double, 256, E, SRC>(a, wk.y); break;
        case 512: rows_fwd_body<R, double, 512, E, SRC>(a, wk.y); break;
        case 1024: rows_fwd_body<R, double, 1024, E, SRC>(a, wk.y); break;
        default: break;
    }
}

template <typename R, int E, int EPI>
__global__ __launch_bounds__(256) void k_rows_inv_batch(const RowsInvArgs<R>* __restrict__ args,
                                                        const int2* __restrict__ work) {
    const int2 wk = work[blockIdx.x];
    const RowsInvArgs<R>& a = args[wk.x];
    switch (a.g.nx) {
        case 128: rows_inv_body<R, 128, E, MODE_PAIR, EPI>(a, wk.y); break;
        case 256: rows_inv_body<R, 256, E, MODE_PAIR, EPI>(a, wk.y); break;
        case 512: rows_inv_body<R, 512, E, MODE_PAIR, EPI>(a, wk.y); break;
        case 1024: rows_inv_body<R, 1024, E, MODE_PAIR, EPI>(a, wk.y); break;
        default: break;
    }
}

}  // namespace ssg
