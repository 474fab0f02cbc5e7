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
    int ncols;   // nx/2 + 1
    int G;
    int nblk;
    const cpx<R>* twx;
    Grid g;      // pitch = nz * ny
};

// gather the packed spectrum of rows (A, B) of one plane into the line buffers
template <typename R, int N, int E, int MAXIT>
__device__ __forceinline__ void ma_load(const cpx<R>* P, int pitch, int y0, int G, int total, cpx<R> (&q0)[MAXIT],
                                        cpx<R> (&q1)[MAXIT]) {
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
        const int idx = threadIdx.x + it * (int)blockDim.x;
        if (idx < total) {
            const int kx = idx / G, ll = idx - kx * G;
            const size_t o = (size_t)kx * pitch + y0 + ll;
            q0[it] = P[o];
            q1[it] = P[o + G];
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
            const int kx = idx / G, ll = idx - kx * G;
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

template <typename R, int N, int E>
__global__ __launch_bounds__(256) void k_ma_rows(MaRowsArgs<R> a) {
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
    // the three fields go to buffers 3l, 3l+1, 3l+2 and are inverted in
    // lockstep; all global loads are issued before the first LDS store
    C* const b3[3] = {lds + (3 * l) * LE, lds + (3 * l + 1) * LE, lds + (3 * l + 2) * LE};
    C qa0[MAXIT], qa1[MAXIT], qb0[MAXIT], qb1[MAXIT], qc0[MAXIT], qc1[MAXIT];
    ma_load<R, N, E, MAXIT>(a.Pxx, pitch, y0, G, total, qa0, qa1);
    ma_load<R, N, E, MAXIT>(a.Pyy, pitch, y0, G, total, qb0, qb1);
    ma_load<R, N, E, MAXIT>(a.Pxy, pitch, y0, G, total, qc0, qc1);
    ma_scatter<R, N, E, MAXIT>(lds, G, total, qa0, qa1, 3);
    ma_scatter<R, N, E, MAXIT>(lds + LE, G, total, qb0, qb1, 3);
    ma_scatter<R, N, E, MAXIT>(lds + 2 * LE, G, total, qc0, qc1, 3);
    C w3[3][E];
    F::template run_n_from_lds<+1, 3>(w3, b3, t, tw);
    // D = Φxx Φyy - Φxy² for rows A (re) and B (im): compute_d_operator!
    // (src/poisson.jl:183-223)
    C v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const C xx = w3[0][m], yy = w3[1][m], xy = w3[2][m];
        v[m] = C{xx.x * yy.x - xy.x * xy.x, xx.y * yy.y - xy.y * xy.y};
    }
    C* buf = b3[0];
    F::template run<-1>(v, buf, t, tw);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) buf[lpad(t + T * m)] = v[m];
    __syncthreads();
    // line l's result sits in buffer 3l: store_split with a stride-3 view
    store_split_stride<R, N, E>(lds, G, 3, a.ncols, a.Dx, a.Dx, pitch, y0, y0 + G);
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
    Grid g;
};

template <typename R, typename P, int NZMAX>
__global__ __launch_bounds__(128) void k_ma_tridiag(MaTriArgs<R, P> a) {
    const long total = (long)a.ncols * a.g.ny;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int ny = a.g.ny, nz = a.nz;
    const int kx = (int)(q / ny), ky = (int)(q - (long)kx * ny);
    const int s = signed_k(ky, ny);
    const double hs = (double)sym_dxx<double>(a.g, kx) + (double)sym_dyy<double>(a.g, s);
    const double* cf = a.coef;
    const cpx<double> b2 = a.B2[(size_t)kx * ny + ky];
    const size_t base = (size_t)kx * nz * ny + ky;
    auto rhs_at = [&](int k) -> double2 {
        double2 f = {0.0, 0.0};
        if (a.body) f = b2;
        if (a.use_d) {
            const cpx<R> d = a.Dh[base + (size_t)k * ny];
            f.x += a.eps * (double)d.x;
            f.y += a.eps * (double)d.y;
        }
        return f;
    };
    double cp[NZMAX];
    double2 dp[NZMAX];
    // row 0: bottom Neumann row with Φ2 eliminated through interior row 1
    {
        const double A0 = cf[3 * nz], B0 = cf[3 * nz + 1], C0 = cf[3 * nz + 2];
        const double al = cf[3], be = cf[4] + hs, ga = cf[5];
        const double2 f1 = rhs_at(1);
        double a0 = A0 - C0 * al / ga, b0 = B0 - C0 * be / ga;
        double2 r0 = {-C0 * f1.x / ga, -C0 * f1.y / ga};
        if (kx == 0 && ky == 0) {  // gauge of the Neumann-Neumann mean mode
            a0 = 1.0;
            b0 = 0.0;
            r0 = {0.0, 0.0};
        }
        cp[0] = b0 / a0;
        dp[0] = {r0.x / a0, r0.y / a0};
    }
    // interior rows, then the top Neumann row with Φn-3 eliminated through
    // interior row n-2 (indices stay compile-time so cp/dp live in registers)
#pragma unroll
    for (int k = 1; k < NZMAX; ++k) {
        if (k < nz - 1) {
            const double al = cf[3 * k], be = cf[3 * k + 1] + hs, ga = cf[3 * k + 2];
            const double2 f = rhs_at(k);
            const double m = be - al * cp[k - 1];
            cp[k] = ga / m;
            dp[k] = {(f.x - al * dp[k - 1].x) / m, (f.y - al * dp[k - 1].y) / m};
        } else if (k == nz - 1) {
            const int ki = k - 1;
            const double at = cf[3 * nz + 3], bt = cf[3 * nz + 4], ct = cf[3 * nz + 5];
            const double al = cf[3 * ki], be = cf[3 * ki + 1] + hs, ga = cf[3 * ki + 2];
            const double2 f = rhs_at(ki);
            double2 g = {0.0, 0.0};
            if (a.top) g = b2;
            const double lo = bt - ct * be / al, di = at - ct * ga / al;
            const double2 rr = {g.x - ct * f.x / al, g.y - ct * f.y / al};
            const double m = di - lo * cp[k - 1];
            cp[k] = 0.0;
            dp[k] = {(rr.x - lo * dp[k - 1].x) / m, (rr.y - lo * dp[k - 1].y) / m};
        }
    }
    double2 x = {0.0, 0.0};
#pragma unroll
    for (int k = NZMAX - 1; k >= 0; --k) {
        if (k <= nz - 1) {
            x = {dp[k].x - cp[k] * x.x, dp[k].y - cp[k] * x.y};
            a.Phi[base + (size_t)k * ny] = cpx<P>{(P)x.x, (P)x.y};
        }
    }
}

}  // namespace ssg
