/*
 * ORACLE — test infrastructure only.
 *
 * FP64 CPU restatement of the SSG.jl surface time-step path. It is the
 * checker that the HIP library is compared against, and the CPU baseline the
 * benchmark times beside the GPU. Nothing in the product library links it.
 *
 * Every routine follows the reference literally (same operation sequence,
 * same scratch aliasing) and cites the reference file:line it restates.
 * Conventions fixed here (the reference's FFT backend is absent):
 *   - real fields are column-major, x fastest: a[i + nx*j];
 *   - spectral 2D fields are (nx, ny/2+1), x fastest, interleaved complex;
 *   - forward = r2c along y, then c2c along x, unnormalised;
 *   - inverse = c2c along x, then c2r along y, scaled by 1/(nx*ny); the c2r
 *     ignores the imaginary part of the y-DC and y-Nyquist bins.
 * Parity pins: brute-force DFT and numpy cross-checks, closed-form KATs
 * (tests/test_oracle.py).
 *
 * Build: make -C oracle   (gcc, OpenMP)
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* 1D complex FFT: iterative radix-2 (powers of two), direct DFT otherwise.  */
/* sign = -1 forward, +1 backward; unnormalised.                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int pow2;
    cplx *tw;   /* tw[m] = exp(-2 pi i m / n), m < n */
    int *rev;   /* bit reversal permutation (pow2 only) */
} fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n;
    p->pow2 = (n > 0) && ((n & (n - 1)) == 0);
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int m = 0; m < n; ++m) {
        double a = -2.0 * M_PI * (double)m / (double)n;
        p->tw[m] = cos(a) + I * sin(a);
    }
    p->rev = NULL;
    if (p->pow2) {
        p->rev = (int *)malloc(sizeof(int) * (size_t)n);
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        for (int m = 0; m < n; ++m) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (m & (1 << b)) r |= 1 << (lg - 1 - b);
            p->rev[m] = r;
        }
    }
}

static void plan_free(fftplan *p) {
    free(p->tw);
    free(p->rev);
}

/* in-place on contiguous buf[0..n); work is scratch of n entries */
static void fft1d(const fftplan *p, cplx *buf, cplx *work, int sign) {
    const int n = p->n;
    if (n <= 1) return;
    if (p->pow2) {
        for (int m = 0; m < n; ++m) work[p->rev[m]] = buf[m];
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int s = 0; s < n; s += len) {
                for (int k = 0; k < half; ++k) {
                    cplx w = p->tw[k * step];
                    if (sign > 0) w = conj(w);
                    cplx a = work[s + k], b = work[s + k + half] * w;
                    work[s + k] = a + b;
                    work[s + k + half] = a - b;
                }
            }
        }
        memcpy(buf, work, sizeof(cplx) * (size_t)n);
    } else {
        for (int k = 0; k < n; ++k) {
            cplx acc = 0;
            for (int m = 0; m < n; ++m) {
                cplx w = p->tw[(int)(((int64_t)k * m) % n)];
                if (sign > 0) w = conj(w);
                acc += buf[m] * w;
            }
            work[k] = acc;
        }
        memcpy(buf, work, sizeof(cplx) * (size_t)n);
    }
}

/* ------------------------------------------------------------------------ */
/* Context: the reference Domain + the scratch of Fields and TimeState.       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz, nyc;
    double lx, ly, lz;
    double *kx, *ky;       /* src/domain.jl:155-162 */
    unsigned char *mask;   /* src/domain.jl:498-514, (nx, nyc) */
    double *z, *dz;        /* src/domain.jl:238-249 (uniform) */
    fftplan px, py;
    /* Fields scratch, src/fields.jl:74-105 */
    double *tmp, *tmp2;
    cplx *tmpc, *tmpc2, *bshat;
    /* TimeState, src/timestep.jl:45-86 */
    double *b_stage, *k1, *k2, *k3, *db_dt_old;
    /* optional linear test tendency (lambda * b) for the KAT hooks */
    int linear_hook;
    double linear_lambda;
} ora_ctx;


ora_ctx *ora_create(int nx, int ny, int nz, double lx, double ly, double lz) {
    if (nx < 2 || ny < 2 || nz < 1) return NULL;
    ora_ctx *c = (ora_ctx *)calloc(1, sizeof(ora_ctx));
    c->nx = nx; c->ny = ny; c->nz = nz; c->nyc = ny / 2 + 1;
    c->lx = lx; c->ly = ly; c->lz = lz;
    c->kx = (double *)malloc(sizeof(double) * (size_t)nx);
    c->ky = (double *)malloc(sizeof(double) * (size_t)c->nyc);
    /* kx = 2pi * fftfreq(Nx, Nx/Lx): [0..Nx/2-1, -Nx/2..-1] * 2pi/Lx
       (AbstractFFTs puts the even-n Nyquist at -n/2). */
    for (int i = 0; i < nx; ++i) {
        int s = (i < nx / 2 || (nx % 2 == 1 && i <= nx / 2)) ? i : i - nx;
        if (nx % 2 == 0 && i == nx / 2) s = -nx / 2;
        c->kx[i] = 2.0 * M_PI / lx * (double)s;
    }
    /* ky = 2pi * rfftfreq(Ny, Ny/Ly): [0..Ny/2] * 2pi/Ly */
    for (int j = 0; j < c->nyc; ++j) c->ky[j] = 2.0 * M_PI / ly * (double)j;
    /* twothirds_mask: rows i in [Nx÷3+1, Nx-Nx÷3] and cols j in
       [2*Nyc÷3+1, Nyc] (1-based) are zero. */
    c->mask = (unsigned char *)malloc((size_t)nx * c->nyc);
    int kxmax = nx / 3, kymax = 2 * c->nyc / 3;
    for (int j = 0; j < c->nyc; ++j)
        for (int i = 0; i < nx; ++i) {
            int i1 = i + 1, j1 = j + 1;
            int zero = (i1 >= kxmax + 1 && i1 <= nx - kxmax) || (j1 >= kymax + 1);
            c->mask[i + (size_t)nx * j] = zero ? 0 : 1;
        }
    /* uniform non-periodic vertical grid: z = range(0, Lz, Nz),
       dz = fill(Lz/(Nz-1), Nz-1) with the last entry duplicated */
    c->z = (double *)malloc(sizeof(double) * (size_t)nz);
    c->dz = (double *)malloc(sizeof(double) * (size_t)nz);
    for (int k = 0; k < nz; ++k) {
        c->z[k] = nz > 1 ? lz * (double)k / (double)(nz - 1) : 0.0;
        c->dz[k] = nz > 1 ? lz / (double)(nz - 1) : lz;
    }
    plan_init(&c->px, nx);
    plan_init(&c->py, ny);
    size_t r2 = (size_t)nx * ny, c2 = (size_t)nx * c->nyc;
    c->tmp = (double *)calloc(r2, sizeof(double));
    c->tmp2 = (double *)calloc(r2, sizeof(double));
    c->tmpc = (cplx *)calloc(c2, sizeof(cplx));
    c->tmpc2 = (cplx *)calloc(c2, sizeof(cplx));
    c->bshat = (cplx *)calloc(c2, sizeof(cplx));
    c->b_stage = (double *)calloc(r2, sizeof(double));
    c->k1 = (double *)calloc(r2, sizeof(double));
    c->k2 = (double *)calloc(r2, sizeof(double));
    c->k3 = (double *)calloc(r2, sizeof(double));
    c->db_dt_old = (double *)calloc(r2, sizeof(double));
    return c;
}

void ora_destroy(ora_ctx *c) {
    if (!c) return;
    free(c->kx); free(c->ky); free(c->mask); free(c->z); free(c->dz);
    plan_free(&c->px); plan_free(&c->py);
    free(c->tmp); free(c->tmp2); free(c->tmpc); free(c->tmpc2); free(c->bshat);
    free(c->b_stage); free(c->k1); free(c->k2); free(c->k3); free(c->db_dt_old);
    free(c);
}

int ora_set_vertical_grid(ora_ctx *c, const double *z, const double *dz) {
    memcpy(c->z, z, sizeof(double) * (size_t)c->nz);
    memcpy(c->dz, dz, sizeof(double) * (size_t)c->nz);
    return 0;
}

void ora_get_wavenumbers(const ora_ctx *c, double *kx, double *ky, unsigned char *mask) {
    if (kx) memcpy(kx, c->kx, sizeof(double) * (size_t)c->nx);
    if (ky) memcpy(ky, c->ky, sizeof(double) * (size_t)c->nyc);
    if (mask) memcpy(mask, c->mask, (size_t)c->nx * c->nyc);
}

/* ------------------------------------------------------------------------ */
/* 2D transforms, src/transforms.jl:42-56 (plan src/domain.jl:136-141)       */
/* ------------------------------------------------------------------------ */
void ora_rfft2(const ora_ctx *c, const double *in, double *out_ri) {
    const int nx = c->nx, ny = c->ny, nyc = c->nyc;
    cplx *out = (cplx *)out_ri;
    /* r2c along y for every x */
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)(ny > nx ? ny : nx));
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(ny > nx ? ny : nx));
#pragma omp for schedule(static)
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) buf[j] = in[i + (size_t)nx * j];
            fft1d(&c->py, buf, work, -1);
            for (int j = 0; j < nyc; ++j) out[i + (size_t)nx * j] = buf[j];
        }
        /* c2c along x for every ky */
#pragma omp for schedule(static)
        for (int j = 0; j < nyc; ++j) fft1d(&c->px, out + (size_t)nx * j, work, -1);
        free(buf);
        free(work);
    }
}

void ora_irfft2(const ora_ctx *c, const double *in_ri, double *out) {
    const int nx = c->nx, ny = c->ny, nyc = c->nyc;
    const cplx *in = (const cplx *)in_ri;
    cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)nx * nyc);
    memcpy(t, in, sizeof(cplx) * (size_t)nx * nyc);
    const double scale = 1.0 / ((double)nx * (double)ny);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)(ny > nx ? ny : nx));
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(ny > nx ? ny : nx));
        /* inverse c2c along x for every ky */
#pragma omp for schedule(static)
        for (int j = 0; j < nyc; ++j) fft1d(&c->px, t + (size_t)nx * j, work, +1);
        /* c2r along y: Hermitian extension, imaginary parts of the DC and
           (even ny) Nyquist bins ignored */
#pragma omp for schedule(static)
        for (int i = 0; i < nx; ++i) {
            buf[0] = creal(t[i]);
            for (int j = 1; j < nyc; ++j) {
                cplx v = t[i + (size_t)nx * j];
                if (ny % 2 == 0 && j == ny / 2) {
                    buf[j] = creal(v);
                } else {
                    buf[j] = v;
                    buf[ny - j] = conj(v);
                }
            }
            fft1d(&c->py, buf, work, +1);
            for (int j = 0; j < ny; ++j) out[i + (size_t)nx * j] = creal(buf[j]) * scale;
        }
        free(buf);
        free(work);
    }
    free(t);
}

/* ddx_2d!, ddy_2d!: src/transforms.jl:135-166 (in place allowed) */
void ora_ddx2(const ora_ctx *c, const double *in_ri, double *out_ri) {
    const cplx *in = (const cplx *)in_ri;
    cplx *out = (cplx *)out_ri;
    for (int j = 0; j < c->nyc; ++j)
        for (int i = 0; i < c->nx; ++i) {
            size_t q = i + (size_t)c->nx * j;
            out[q] = (I * c->kx[i]) * in[q];
        }
}

void ora_ddy2(const ora_ctx *c, const double *in_ri, double *out_ri) {
    const cplx *in = (const cplx *)in_ri;
    cplx *out = (cplx *)out_ri;
    for (int j = 0; j < c->nyc; ++j)
        for (int i = 0; i < c->nx; ++i) {
            size_t q = i + (size_t)c->nx * j;
            out[q] = (I * c->ky[j]) * in[q];
        }
}

/* dealias_2d!: src/domain.jl:522-535 */
void ora_dealias2(const ora_ctx *c, double *a_ri) {
    cplx *a = (cplx *)a_ri;
    size_t n = (size_t)c->nx * c->nyc;
    for (size_t q = 0; q < n; ++q)
        if (!c->mask[q]) a[q] = 0;
}

/* ------------------------------------------------------------------------ */
/* jacobian_2d!: src/transforms.jl:430-465, literal sequence. Pointer        */
/* aliasing between output and tmp_real2 is honoured (AB2 path).              */
/* ------------------------------------------------------------------------ */
static void jacobian_2d(const ora_ctx *c, double *output, const double *a, const double *b,
                        cplx *tmp_spec1, cplx *tmp_spec2, double *tmp_real1, double *tmp_real2) {
    const size_t n = (size_t)c->nx * c->ny;
    /* Step 1: u_x = -d a/dy */
    ora_rfft2(c, a, (double *)tmp_spec1);
    ora_ddy2(c, (double *)tmp_spec1, (double *)tmp_spec2);
    ora_irfft2(c, (double *)tmp_spec2, tmp_real1);
    for (size_t q = 0; q < n; ++q) tmp_real1[q] *= -1.0;
    /* Step 2: flux u_x * b */
    for (size_t q = 0; q < n; ++q) tmp_real1[q] *= b[q];
    /* Step 3: d(u_x b)/dx -> output */
    ora_rfft2(c, tmp_real1, (double *)tmp_spec1);
    ora_ddx2(c, (double *)tmp_spec1, (double *)tmp_spec2);
    ora_irfft2(c, (double *)tmp_spec2, output);
    /* Step 4: u_y = d a/dx */
    ora_rfft2(c, a, (double *)tmp_spec1);
    ora_ddx2(c, (double *)tmp_spec1, (double *)tmp_spec2);
    ora_irfft2(c, (double *)tmp_spec2, tmp_real1);
    /* Step 5: flux u_y * b */
    for (size_t q = 0; q < n; ++q) tmp_real1[q] *= b[q];
    /* Step 6: d(u_y b)/dy -> tmp_real2 */
    ora_rfft2(c, tmp_real1, (double *)tmp_spec1);
    ora_ddy2(c, (double *)tmp_spec1, (double *)tmp_spec2);
    ora_irfft2(c, (double *)tmp_spec2, tmp_real2);
    /* Step 7: output += tmp_real2 */
    for (size_t q = 0; q < n; ++q) output[q] += tmp_real2[q];
}

/* compute_jacobian!: src/nonlinear.jl:42-57 */
static void compute_jacobian(ora_ctx *c, double *db_dt, const double *phi, const double *b) {
    jacobian_2d(c, db_dt, phi, b, c->tmpc, c->tmpc2, c->tmp, c->tmp2);
    const size_t n = (size_t)c->nx * c->ny;
    for (size_t q = 0; q < n; ++q) db_dt[q] *= -1.0;
}

/* compute_tendency!: src/nonlinear.jl:22-36.  db_dt may be c->tmp2 (the
   AB2 call site src/timestep.jl:159), which reproduces that aliasing. */
static void compute_tendency_buf(ora_ctx *c, double *db_dt, const double *phi, const double *b) {
    if (c->linear_hook) {
        const size_t n = (size_t)c->nx * c->ny;
        for (size_t q = 0; q < n; ++q) db_dt[q] = c->linear_lambda * b[q];
        return;
    }
    compute_jacobian(c, db_dt, phi, b);
    ora_rfft2(c, db_dt, (double *)c->tmpc);
    ora_dealias2(c, (double *)c->tmpc);
    ora_irfft2(c, (double *)c->tmpc, db_dt);
}

/* Public: tendency into out (quirk=1 routes through tmp2 like the AB2 path) */
void ora_tendency(ora_ctx *c, double *out, const double *phi, const double *b, int quirk) {
    const size_t n = (size_t)c->nx * c->ny;
    if (quirk) {
        compute_tendency_buf(c, c->tmp2, phi, b);
        memcpy(out, c->tmp2, sizeof(double) * n);
    } else {
        compute_tendency_buf(c, out, phi, b);
    }
}

/* Public: compute_jacobian! (negated flux-form Jacobian) */
void ora_jacobian(ora_ctx *c, double *out, const double *phi, const double *b) {
    compute_jacobian(c, out, phi, b);
}

void ora_set_linear_hook(ora_ctx *c, int on, double lambda) {
    c->linear_hook = on;
    c->linear_lambda = lambda;
}

/* ------------------------------------------------------------------------ */
/* Exponential filter: src/filter.jl:119-126, :283-309; legacy chain         */
/* src/timestep.jl:247-260 -> src/filter.jl:525-532                           */
/* ------------------------------------------------------------------------ */
double ora_exponential_transfer(double k_norm, double strength, int order, double cutoff) {
    if (k_norm <= cutoff) return 1.0;
    double ratio = (k_norm - cutoff) / (1.0 - cutoff);
    return exp(-strength * pow(ratio, (double)order));
}

void ora_filter_spectral(const ora_ctx *c, double *spec_ri, double strength, int order, double cutoff) {
    cplx *s = (cplx *)spec_ri;
    const double dx = c->lx / c->nx, dy = c->ly / c->ny;
    const double kmx = M_PI / dx, kmy = M_PI / dy;
    const double kmax = sqrt(kmx * kmx + kmy * kmy);
    for (int j = 0; j < c->nyc; ++j) {
        const double ky = c->ky[j];
        for (int i = 0; i < c->nx; ++i) {
            const double kx = c->kx[i];
            double kn = sqrt(kx * kx + ky * ky) / kmax;
            s[i + (size_t)c->nx * j] *= ora_exponential_transfer(kn, strength, order, cutoff);
        }
    }
}

/* apply_spectral_filter! on b in place (rfft -> filter -> irfft) */
void ora_filter(ora_ctx *c, double *b, double strength, int order, double cutoff) {
    ora_rfft2(c, b, (double *)c->bshat);
    ora_filter_spectral(c, (double *)c->bshat, strength, order, cutoff);
    ora_irfft2(c, (double *)c->bshat, b);
}

/* ------------------------------------------------------------------------ */
/* Time stepping: src/timestep.jl:152-285                                     */
/* scheme: 0 = AB2_LowStorage, 1 = RK3, 2 = RK3_LowStorage (unsupported)     */
/* ab2_mode: 0 = corrected tendency, 1 = reference aliasing (tmp2 buffer)    */
/* ------------------------------------------------------------------------ */
static void timestep_ab2(ora_ctx *c, double dt, double *b, const double *phi, int64_t step, int ab2_mode) {
    const size_t n = (size_t)c->nx * c->ny;
    double *f = c->k1; /* corrected path uses a private buffer */
    if (ab2_mode) f = c->tmp2;
    compute_tendency_buf(c, f, phi, b);
    if (step == 0) {
        for (size_t q = 0; q < n; ++q) b[q] += dt * f[q];
    } else {
        for (size_t q = 0; q < n; ++q) b[q] += dt * (1.5 * f[q] - 0.5 * c->db_dt_old[q]);
    }
    memcpy(c->db_dt_old, f, sizeof(double) * n);
}

static void timestep_rk3(ora_ctx *c, double dt, double *b, const double *phi) {
    const size_t n = (size_t)c->nx * c->ny;
    memcpy(c->b_stage, b, sizeof(double) * n);
    compute_tendency_buf(c, c->k1, phi, b);
    for (size_t q = 0; q < n; ++q) b[q] = c->b_stage[q] + (dt / 3.0) * c->k1[q];
    compute_tendency_buf(c, c->k2, phi, b);
    for (size_t q = 0; q < n; ++q) b[q] = c->b_stage[q] + (2.0 * dt / 3.0) * c->k2[q];
    compute_tendency_buf(c, c->k3, phi, b);
    for (size_t q = 0; q < n; ++q) b[q] = c->b_stage[q] + (dt / 4.0) * (c->k1[q] + 3.0 * c->k3[q]);
}

/* timestep!: returns 0 on success, -1 for an unsupported scheme */
int ora_timestep(ora_ctx *c, int scheme, double dt, int filter_freq, double filter_strength,
                 int filter_order, double filter_cutoff, int ab2_mode,
                 double *b, const double *phi, double *t, int64_t *step) {
    if (scheme == 0) {
        timestep_ab2(c, dt, b, phi, *step, ab2_mode);
    } else if (scheme == 1) {
        timestep_rk3(c, dt, b, phi);
    } else {
        return -1;
    }
    *t += dt;
    *step += 1;
    if (filter_freq > 0 && (*step % filter_freq) == 0)
        ora_filter(c, b, filter_strength, filter_order, filter_cutoff);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                                */
/* ------------------------------------------------------------------------ */
/* parsevalsum2 for a 2D rfft field: src/transforms.jl:230-274 (nz = 1 slice;
   the normalisation still carries Lz and the domain Nz) */
double ora_parsevalsum2(const ora_ctx *c, const double *spec_ri, int nlev) {
    const cplx *u = (const cplx *)spec_ri;
    double acc = 0.0;
    for (int k = 0; k < nlev; ++k) {
        const cplx *s = u + (size_t)k * c->nx * c->nyc;
        for (int j = 0; j < c->nyc; ++j) {
            double a0 = creal(s[(size_t)c->nx * j]) * creal(s[(size_t)c->nx * j]) +
                        cimag(s[(size_t)c->nx * j]) * cimag(s[(size_t)c->nx * j]);
            acc += a0;
        }
        int inyq = c->nx / 2; /* 0-based index of Nx÷2+1 */
        for (int j = 0; j < c->nyc; ++j) {
            cplx v = s[inyq + (size_t)c->nx * j];
            acc += creal(v) * creal(v) + cimag(v) * cimag(v);
        }
        for (int i = 1; i < inyq; ++i)
            for (int j = 0; j < c->nyc; ++j) {
                cplx v = s[i + (size_t)c->nx * j];
                acc += 2.0 * (creal(v) * creal(v) + cimag(v) * cimag(v));
            }
    }
    double norm = (c->lx * c->ly * c->lz) / ((double)c->nx * c->nx * (double)c->ny * c->ny * (double)c->nz);
    return acc * norm;
}

/* ------------------------------------------------------------------------ */
/* Monge-Ampere residual operator: src/poisson.jl:143-364                     */
/* 3D fields are (nx, ny, nz), x fastest. phi is modified in place by the     */
/* boundary rows exactly as apply_ssg_boundary_conditions! does.              */
/* ------------------------------------------------------------------------ */
static void spec_level_apply(const ora_ctx *c, const double *in, double *out, int which) {
    /* which: 0 = d2/dx2, 1 = d2/dy2, 2 = d2/dxdy; literal rfft -> ddx/ddy twice -> irfft */
    size_t c2 = (size_t)c->nx * c->nyc;
    cplx *h = (cplx *)malloc(sizeof(cplx) * c2);
    cplx *t = (cplx *)malloc(sizeof(cplx) * c2);
    ora_rfft2(c, in, (double *)h);
    if (which == 0) {
        ora_ddx2(c, (double *)h, (double *)t);
        ora_ddx2(c, (double *)t, (double *)t);
    } else if (which == 1) {
        ora_ddy2(c, (double *)h, (double *)t);
        ora_ddy2(c, (double *)t, (double *)t);
    } else {
        ora_ddx2(c, (double *)h, (double *)t);
        ora_ddy2(c, (double *)t, (double *)t);
    }
    ora_irfft2(c, (double *)t, out);
    free(h);
    free(t);
}

/* apply_ssg_boundary_conditions!: src/poisson.jl:262-364 (nz >= 3 rows) */
void ora_ma_apply_bc(const ora_ctx *c, double *phi, double *r, const double *bs_surface) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t pl = (size_t)nx * ny;
    const double *dz = c->dz;
    for (size_t q = 0; q < pl; ++q) {
        if (nz >= 3) {
            int k = nz - 1; /* 0-based surface */
            double h1 = dz[k - 1], h2 = dz[k - 2];
            double al = h1 + h2;
            double a = (2 * h1 + h2) / (h1 * al), b = -al / (h1 * h2), cc = h1 / (h2 * al);
            double g = bs_surface ? bs_surface[q] : 0.0;
            phi[q + pl * k] = (g - b * phi[q + pl * (k - 1)] - cc * phi[q + pl * (k - 2)]) / a;
            if (r) r[q + pl * k] = 0.0;
            /* bottom: h1 = dz[1], h2 = dz[2] in 1-based terms */
            double H1 = dz[0], H2 = dz[1];
            double AL = H1 + H2;
            double A = -(2 * H1 + H2) / (H1 * AL), B = AL / (H1 * H2), C = -H1 / (H2 * AL);
            phi[q] = -(B * phi[q + pl] + C * phi[q + 2 * pl]) / A;
            if (r) r[q] = 0.0;
        } else if (nz == 2) {
            double g = bs_surface ? bs_surface[q] : 0.0;
            phi[q + pl] = phi[q] + dz[0] * g;
            if (r) r[q + pl] = 0.0;
            phi[q] = phi[q + pl];
            if (r) r[q] = 0.0;
        }
    }
}

/* compute_ssg_residual!: src/poisson.jl:228-253 with compute_3d_laplacian!
   (:143-178), d2dz2! (src/transforms.jl:798-850) and compute_d_operator!
   (:183-223). r = lap - eps*D - b, then boundary rows. */
void ora_ma_residual(const ora_ctx *c, double *phi, const double *b3d, const double *bs_surface,
                     double eps, double *r) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t pl = (size_t)nx * ny;
    double *pxx = (double *)malloc(sizeof(double) * pl);
    double *pyy = (double *)malloc(sizeof(double) * pl);
    double *pxy = (double *)malloc(sizeof(double) * pl);
    for (int k = 0; k < nz; ++k) {
        const double *lv = phi + pl * k;
        spec_level_apply(c, lv, pxx, 0);
        spec_level_apply(c, lv, pyy, 1);
        spec_level_apply(c, lv, pxy, 2);
        double al = 0, be = 0, ga = 0;
        int interior = (k >= 1 && k <= nz - 2);
        if (interior) {
            double dzm = c->dz[k - 1], dzp = c->dz[k];
            al = 2.0 / (dzm * (dzm + dzp));
            be = -2.0 / (dzm * dzp);
            ga = 2.0 / (dzp * (dzm + dzp));
        }
        for (size_t q = 0; q < pl; ++q) {
            double pzz = interior ? al * phi[q + pl * (k - 1)] + be * lv[q] + ga * phi[q + pl * (k + 1)] : 0.0;
            double lap = pxx[q] + pyy[q] + pzz;
            double D = pxx[q] * pyy[q] - pxy[q] * pxy[q];
            r[q + pl * k] = lap - eps * D - b3d[q + pl * k];
        }
    }
    free(pxx); free(pyy); free(pxy);
    ora_ma_apply_bc(c, phi, r, bs_surface);
}

/* D operator alone (K11) */
void ora_ma_d_operator(const ora_ctx *c, const double *phi, double *out) {
    const size_t pl = (size_t)c->nx * c->ny;
    double *pxx = (double *)malloc(sizeof(double) * pl);
    double *pyy = (double *)malloc(sizeof(double) * pl);
    double *pxy = (double *)malloc(sizeof(double) * pl);
    for (int k = 0; k < c->nz; ++k) {
        spec_level_apply(c, phi + pl * k, pxx, 2);
        memcpy(pxy, pxx, sizeof(double) * pl);
        spec_level_apply(c, phi + pl * k, pxx, 0);
        spec_level_apply(c, phi + pl * k, pyy, 1);
        for (size_t q = 0; q < pl; ++q) out[q + pl * k] = pxx[q] * pyy[q] - pxy[q] * pxy[q];
    }
    free(pxx); free(pyy); free(pxy);
}

/* L2 norm, norm_field: src/poisson.jl:986-995 */
double ora_norm(const double *a, int64_t n) {
    double s = 0.0;
    for (int64_t q = 0; q < n; ++q) s += a[q] * a[q];
    return sqrt(s);
}

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ora_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
