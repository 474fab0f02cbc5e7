// Device-side building blocks: complex arithmetic, small in-register DFTs and
// a Stockham line FFT that keeps each thread's elements in registers and
// exchanges them through LDS between stages.
//
// Register layout (invariant across a whole line FFT): a line of N complex
// points is owned by T = N/E consecutive threads; thread t holds elements
// t + T*m, m = 0..E-1.  A forward or inverse transform maps that layout to
// itself in natural order, so a symbol multiply can sit between a forward and
// an inverse transform without touching memory.
#pragma once
#include <hip/hip_runtime.h>

namespace ssg {

template <typename R> struct cpx_t;
template <> struct cpx_t<double> { using type = double2; };
template <> struct cpx_t<float> { using type = float2; };
template <typename R> using cpx = typename cpx_t<R>::type;
// the real type of a complex type
template <typename C> struct cpx_real_t;
template <> struct cpx_real_t<double2> { using type = double; };
template <> struct cpx_real_t<float2> { using type = float; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
template <typename C, typename R> __device__ __forceinline__ C cscale(C a, R s) { return C{a.x * s, a.y * s}; }
// multiply by i*s (s real)
template <typename C, typename R> __device__ __forceinline__ C cmul_i(C a, R s) { return C{-a.y * s, a.x * s}; }

// cos(2*pi*m/16), m taken mod 16
__device__ __forceinline__ constexpr double cos16(int m) {
    switch (m & 15) {
        case 0: return 1.0;
        case 1: return 0.92387953251128675613;
        case 2: return 0.70710678118654752440;
        case 3: return 0.38268343236508977173;
        case 4: return 0.0;
        case 5: return -0.38268343236508977173;
        case 6: return -0.70710678118654752440;
        case 7: return -0.92387953251128675613;
        case 8: return -1.0;
        case 9: return -0.92387953251128675613;
        case 10: return -0.70710678118654752440;
        case 11: return -0.38268343236508977173;
        case 12: return 0.0;
        case 13: return 0.38268343236508977173;
        case 14: return 0.70710678118654752440;
        default: return 0.92387953251128675613;
    }
}
__device__ __forceinline__ constexpr double sin16(int m) { return cos16(m - 4); }

// o * exp(DIR * 2*pi*i * m/16); m is a compile-time constant after unrolling
template <typename R, int DIR>
__device__ __forceinline__ cpx<R> tw16(cpx<R> o, int m) {
    using C = cpx<R>;
    m &= 15;
    if (m == 0) return o;
    if (m == 8) return C{-o.x, -o.y};
    if (m == 4) return DIR < 0 ? C{o.y, -o.x} : C{-o.y, o.x};
    if (m == 12) return DIR < 0 ? C{-o.y, o.x} : C{o.y, -o.x};
    const R c = (R)cos16(m);
    const R s = (R)(DIR * sin16(m));
    return C{o.x * c - o.y * s, o.x * s + o.y * c};
}

// In-register DFT of length L (power of two, L <= 16), natural order in/out.
template <typename R, int L, int DIR> struct SmallDFT {
    __device__ __forceinline__ static void run(cpx<R>* v) {
        cpx<R> e[L / 2], o[L / 2];
#pragma unroll
        for (int i = 0; i < L / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        SmallDFT<R, L / 2, DIR>::run(e);
        SmallDFT<R, L / 2, DIR>::run(o);
#pragma unroll
        for (int k = 0; k < L / 2; ++k) {
            cpx<R> t = tw16<R, DIR>(o[k], k * (16 / L));
            v[k] = cadd(e[k], t);
            v[k + L / 2] = csub(e[k], t);
        }
    }
};
template <typename R, int DIR> struct SmallDFT<R, 1, DIR> {
    __device__ __forceinline__ static void run(cpx<R>*) {}
};

__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// log2 of a power of two (the lines per block G of every row and column
// pass): index arithmetic by shifts and masks, not integer division
__device__ __forceinline__ int gsh(int g) { return __builtin_ctz((unsigned)g); }

// LDS index padding: one spare slot every 2^SSG_LDS_PAD_SHIFT entries keeps
// the stride-E writes of the first stage off a single bank group.
#ifndef SSG_LDS_PAD_SHIFT
#define SSG_LDS_PAD_SHIFT 4
#endif
__device__ __forceinline__ int lpad(int i) { return i + (i >> SSG_LDS_PAD_SHIFT); }
__host__ __device__ constexpr int lds_line_elems(int n) { return n + (n >> SSG_LDS_PAD_SHIFT) + 1; }

// number of twiddle factors a thread applies in the stages after the first,
// and the offset of stage NS within that list
__host__ __device__ constexpr int tw_count_from(int n, int e, int ns) {
    return ns >= n ? 0 : (cmin(e, n / ns) - 1) * (e / cmin(e, n / ns)) + tw_count_from(n, e, ns * cmin(e, n / ns));
}
__host__ __device__ constexpr int tw_offset(int n, int e, int ns) {
    return tw_count_from(n, e, e) - tw_count_from(n, e, ns);
}

// Synchronisation between the stages of one line's transform.  When a line
// is owned by threads of a single wavefront (T <= 64, T a power of two, lines
// laid out contiguously) the exchange stays inside the wave: LDS operations
// of a wave complete in order, so a wave-scope fence and a scheduling
// barrier replace the block barrier and lines of other waves run
// uncoupled.  Cooperative phases across lines keep __syncthreads().
template <int T>
__device__ __forceinline__ void line_sync() {
    if constexpr (T <= 64) {
        __syncwarp();
    } else {
        __syncthreads();
    }
}

// Stockham line FFT.  v[] holds x[t + T*m] on entry and X[t + T*m] on exit.
// `buf` is this line's LDS buffer (lds_line_elems(N) entries).  With PPREQ
// and lines wider than a wavefront, the stages alternate between buf and a
// second buffer `alt` entries further on (ping-pong): a stage reads one
// buffer and writes the other, so the barrier that keeps a stage's writes
// behind every thread's reads of the same buffer goes, one block barrier
// per stage instead of two.  The caller then provides LDS for both.  The
// twiddle
// factors a thread needs depend only on t, so they are fetched once per
// kernel into registers (Twiddles::load, from tw[m] = exp(-2*pi*i*m/N)) and
// shared by every forward and inverse transform of the kernel; loading them
// up front keeps global-memory latency out of the barrier-separated stages.
// Every thread of the block must call run() (it contains block barriers).
template <typename R, int N, int E, bool PPREQ = false> struct LineFFT {
    static constexpr int T = N / E;
    static constexpr int NTW = tw_count_from(N, E, E);
    // ping-pong only where stages are separated by block barriers
    static constexpr bool PP = PPREQ && (T > 64);
    using C = cpx<R>;

    struct Twiddles {
        C w[NTW > 0 ? NTW : 1];

        template <int NS>
        __device__ __forceinline__ void load_stage(int t, const C* __restrict__ tab) {
            if constexpr (NS < N) {
                constexpr int RAD = cmin(E, N / NS);
                constexpr int Q = E / RAD;
                constexpr int OFF = tw_offset(N, E, NS);
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int k = (t + T * q) % NS;
#pragma unroll
                    for (int r = 1; r < RAD; ++r) w[OFF + q * (RAD - 1) + r - 1] = tab[(r * k) * (N / (NS * RAD))];
                }
                load_stage<NS * RAD>(t, tab);
            }
        }
        __device__ __forceinline__ void load(int t, const C* __restrict__ tab) {
            if constexpr (T > 1) load_stage<E>(t, tab);
        }
    };

    template <int DIR, int NS, bool PAR = false>
    __device__ __forceinline__ static void stage(C* v, C* buf, int t, const Twiddles& tw, int alt) {
        constexpr int RAD = cmin(E, N / NS);
        constexpr int Q = E / RAD;
        constexpr bool LAST = (NS * RAD == N);
        constexpr int OFF = tw_offset(N, E, NS);
        const C* src = (PP && PAR) ? buf + alt : buf;
        C* dst = PP ? (PAR ? buf : buf + alt) : buf;
        C u[E];
        line_sync<T>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + T * q;
#pragma unroll
            for (int r = 0; r < RAD; ++r) u[q * RAD + r] = src[lpad(j + r * (N / RAD))];
        }
        if constexpr (!PP) line_sync<T>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + T * q;
            const int k = j % NS;
            C* w = u + q * RAD;
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                C f = tw.w[OFF + q * (RAD - 1) + r - 1];
                if (DIR > 0) f.y = -f.y;
                w[r] = cmul(w[r], f);
            }
            SmallDFT<R, RAD, DIR>::run(w);
            if (LAST) {
#pragma unroll
                for (int r = 0; r < RAD; ++r) v[q + r * Q] = w[r];
            } else {
                const int base = (j / NS) * NS * RAD + k;
#pragma unroll
                for (int r = 0; r < RAD; ++r) dst[lpad(base + r * NS)] = w[r];
            }
        }
        if constexpr (!LAST) stage<DIR, NS * RAD, !PAR>(v, buf, t, tw, alt);
    }

    // NB independent lines of the same thread group transformed in lockstep:
    // one barrier pair per stage for all of them, NB-fold independent work
    // between barriers.
    template <int DIR, int NS, int NB, bool PAR = false>
    __device__ __forceinline__ static void stage_n(C (&v)[NB][E], C* const (&buf)[NB], int t, const Twiddles& tw,
                                                   int alt) {
        constexpr int RAD = cmin(E, N / NS);
        constexpr int Q = E / RAD;
        constexpr bool LAST = (NS * RAD == N);
        constexpr int OFF = tw_offset(N, E, NS);
        const int soff = (PP && PAR) ? alt : 0;
        const int doff = PP ? (PAR ? 0 : alt) : 0;
        C u[NB][E];
        line_sync<T>();
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int j = t + T * q;
#pragma unroll
                for (int r = 0; r < RAD; ++r) u[b][q * RAD + r] = buf[b][soff + lpad(j + r * (N / RAD))];
            }
        if constexpr (!PP) line_sync<T>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + T * q;
            const int k = j % NS;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                C* w = u[b] + q * RAD;
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    C f = tw.w[OFF + q * (RAD - 1) + r - 1];
                    if (DIR > 0) f.y = -f.y;
                    w[r] = cmul(w[r], f);
                }
                SmallDFT<R, RAD, DIR>::run(w);
                if (LAST) {
#pragma unroll
                    for (int r = 0; r < RAD; ++r) v[b][q + r * Q] = w[r];
                } else {
                    const int base = (j / NS) * NS * RAD + k;
#pragma unroll
                    for (int r = 0; r < RAD; ++r) buf[b][doff + lpad(base + r * NS)] = w[r];
                }
            }
        }
        if constexpr (!LAST) stage_n<DIR, NS * RAD, NB, !PAR>(v, buf, t, tw, alt);
    }

    // `alt`: offset (in elements) of each buffer's ping-pong partner; read
    // only when PP
    template <int DIR, int NB>
    __device__ __forceinline__ static void run_n(C (&v)[NB][E], C* const (&buf)[NB], int t, const Twiddles& tw,
                                                 int alt = 0) {
#pragma unroll
        for (int b = 0; b < NB; ++b) SmallDFT<R, E, DIR>::run(v[b]);
        if constexpr (T > 1) {
            line_sync<T>();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < E; ++r) buf[b][lpad(t * E + r)] = v[b][r];
            stage_n<DIR, E, NB>(v, buf, t, tw, alt);
        }
    }

    // NB lines whose inputs sit in LDS in natural order
    template <int DIR, int NB>
    __device__ __forceinline__ static void run_n_from_lds(C (&v)[NB][E], C* const (&buf)[NB], int t,
                                                          const Twiddles& tw, int alt = 0) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int m = 0; m < E; ++m) v[b][m] = buf[b][lpad(t + T * m)];
        run_n<DIR, NB>(v, buf, t, tw, alt);
    }

    // Transform from registers.
    template <int DIR>
    __device__ __forceinline__ static void run(C* v, C* buf, int t, const Twiddles& tw, int alt = 0) {
        SmallDFT<R, E, DIR>::run(v);
        if constexpr (T > 1) {
            line_sync<T>();  // earlier readers of buf are done
#pragma unroll
            for (int r = 0; r < E; ++r) buf[lpad(t * E + r)] = v[r];
            stage<DIR, E>(v, buf, t, tw, alt);
        }
    }

    // Transform whose input sits in LDS in natural order (buf[lpad(i)] = x[i]).
    template <int DIR>
    __device__ __forceinline__ static void run_from_lds(C* v, C* buf, int t, const Twiddles& tw, int alt = 0) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[lpad(t + T * m)];
        run<DIR>(v, buf, t, tw, alt);
    }

    // Leave the result in LDS in natural order.
    __device__ __forceinline__ static void to_lds(const C* v, C* buf, int t) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) buf[lpad(t + T * m)] = v[m];
        __syncthreads();
    }
};

}  // namespace ssg
