/*
 * ssg.h — C ABI of the MI355X SSG time-step library (libssg_hip.so).
 *
 * Drop-in for the time-step hot path of SSG.jl: the reference's
 * `timestep!(fields, domain, params, state)` and the spectral routines under
 * it.  A Julia host binds these with `ccall`; the Python mirror in
 * ssg.jl_amd/ssgamd binds them with ctypes.
 *
 * Conventions
 *   - Real 2D fields cross the ABI as FP64, column-major with x fastest:
 *     element (i, j) at i + nx*j (src/fields.jl:30-60).
 *   - Reference-layout spectra are (nx, ny/2+1) interleaved complex FP64,
 *     x fastest (src/domain.jl:117-141).
 *   - Every call returns 0 on success and a negative SSG_ERR_* code on
 *     failure; ssg_last_error() gives the message.  No C++ exception,
 *     abort or exit crosses the ABI.
 *   - The context owns all device memory.  Host pointers are borrowed for
 *     the duration of a call.  A context is used by one host thread at a
 *     time; calls are synchronous.
 *   - Grid limits of this build: nx and ny powers of two in [8, 4096].
 */
#ifndef SSG_H
#define SSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ssg_ctx ssg_ctx;

enum {
    SSG_OK = 0,
    SSG_ERR_ARG = -1,
    SSG_ERR_HIP = -2,
    SSG_ERR_NONFINITE = -3,
    SSG_ERR_UNSUPPORTED = -4,
    SSG_ERR_NOT_CONVERGED = -5
};

/* arithmetic of the hot path */
enum { SSG_FP64 = 0, SSG_FP32_MIXED = 1 };

/* TimeScheme, src/timestep.jl:10-14 (same integer values) */
enum { SSG_AB2_LOWSTORAGE = 0, SSG_RK3 = 1, SSG_RK3_LOWSTORAGE = 2 };

/* field identifiers for ssg_set_field / ssg_get_field */
enum {
    SSG_FIELD_B = 0,   /* fields.bₛ  (nx*ny)            src/fields.jl:30 */
    SSG_FIELD_PHI = 1, /* fields.φₛ  (nx*ny)            src/fields.jl:31 */
    SSG_FIELD_U = 2,   /* surface u = -∂yφₛ (get only)  src/setICs.jl:161-192 */
    SSG_FIELD_V = 3    /* surface v =  ∂xφₛ (get only) */
};

/* Domain, src/domain.jl:33-74 (uniform z grid, src/domain.jl:238-249) */
typedef struct {
    int32_t nx, ny, nz;
    double lx, ly, lz;
    int32_t precision; /* SSG_FP64 or SSG_FP32_MIXED */
    int32_t device;    /* HIP device ordinal */
} ssg_domain_desc;

/* TimeParams, src/timestep.jl:19-40, plus the filter shape that the
   reference fixes in its legacy path (order 4, cutoff 0.65,
   src/filter.jl:525-532) */
typedef struct {
    double dt;
    int32_t scheme;
    int32_t filter_freq;            /* 0 = never */
    double filter_strength;         /* α */
    int32_t filter_order;           /* p (reference legacy path: 4) */
    double filter_cutoff;           /* c (reference legacy path: 0.65) */
    int32_t ab2_reference_aliasing; /* 1: reproduce the tendency of
                                       src/timestep.jl:159 exactly */
    int32_t check_finite;           /* 1: fail with SSG_ERR_NONFINITE if b
                                       is not finite after the call */
} ssg_time_params;

/* TimeState scalars, src/timestep.jl:45-86 (the stage buffers and the AB2
   history live in the context) */
typedef struct {
    double t;
    int64_t step;
    double dt_actual;
} ssg_time_state;

/* FP64 diagnostics */
typedef struct {
    double sum_b;          /* Σ b */
    double sum_b2;         /* Σ b² */
    double parseval_b;     /* parsevalsum2(rfft2(b)), src/transforms.jl:230-274 */
    double ke_surface;     /* compute_surface_kinetic_energy, src/setICs.jl:225-246: the reference's
                              formula, parsevalsum2 weights along x and the factor Lz included */
    int64_t nonfinite;     /* count of non-finite b entries */
    double ke_true;        /* ½ ∫∫ (u² + v²) dx dy of the surface velocities, the energy the reference's
                              formula stands for (no Lz, every Fourier mode counted once) */
} ssg_diag;

/* Monge-Ampère problem variants (src/poisson.jl:1068-1131):
   SSG_MA_BODY    as wired: b₃d[:,:,k] = bₛ on every level, surface slope 0
                  (extend_2d_to_3d, src/poisson.jl:1096-1113)
   SSG_MA_SURFACE documented physics: no body term, surface slope ∂Φ/∂z = bₛ
                  (src/poisson.jl:2-6) */
enum { SSG_MA_BODY = 0, SSG_MA_SURFACE = 1 };

typedef struct {
    double eps;        /* ε (src/poisson.jl:1072 default 0.1) */
    double tol;        /* relative residual target ‖r‖₂ / ‖b + εD‖₂ on interior levels */
    int32_t maxiter;   /* src/poisson.jl:874 default 50 */
    int32_t variant;   /* SSG_MA_BODY or SSG_MA_SURFACE */
    int32_t warm;      /* 0: cold start; 1: start from the previous solve's D(Φ); 2: extrapolate from the
                          last converged solves: in coupled stepping quadratically over the same stage of
                          the last three steps, otherwise 2 D̂_k - D̂_{k-1}.  Writing b or changing the
                          coupling resets that history; a failed warm start is retried cold. */
    int32_t stage_fp32; /* coupled stepping of an SSG_FP32_MIXED context only (0 elsewhere): 1 runs the
                          solves inside a ssg_timestep call (the RK stages and the step ends before the
                          last) entirely in FP32 — transforms, products, D̂, warm-start history — to the
                          relative residual SSG_MA_STAGE_TOL32, from which the tendency takes (u, v); the
                          solve for the b the call returns stays FP64 to `tol`, so φₛ, (u, v), the
                          diagnostics and a saved state after the call are FP64 solves.  No counterpart
                          in the reference, whose timestep! keeps φₛ frozen. */
} ssg_ma_params;

/* relative residual of the FP32 stage solves (ssg_ma_params.stage_fp32) */
#define SSG_MA_STAGE_TOL32 1e-6

/* solve_ssg_equation diagnostics (src/poisson.jl:928-958) */
typedef struct {
    int32_t converged;
    int32_t iterations;
    double final_residual;     /* norm_field(r) = sqrt(Σ r²) over all points, src/poisson.jl:986-995 */
    double relative_residual;  /* ‖r‖₂ / ‖b + εD‖₂ over interior levels */
    double history[64];        /* relative residual per iteration (first 64) */
    double solve_ms;
    int32_t iterations_fp32;   /* leading iterations run in FP32 arithmetic (FP32-mixed contexts) */
    int32_t not_elliptic;      /* 1: body variant with eps*max(b) >= 1, rejected before iterating: at its
                                  z-independent fixed point det(I - eps*H) = 1 - eps*b (H the horizontal
                                  Hessian of Φ), so no elliptic solution exists */
} ssg_ma_diag;

/* make_domain + allocate_fields + TimeState (src/domain.jl:217-230,
   src/fields.jl:72-116, src/timestep.jl:45-86) */
int ssg_create(ssg_ctx** out, const ssg_domain_desc* desc);
int ssg_destroy(ssg_ctx* ctx);
/* ctx = NULL gives the message of this thread's last failed context-free
   call (ssg_tendency_noalloc) */
const char* ssg_last_error(const ssg_ctx* ctx);
int ssg_sync(ssg_ctx* ctx);

/* copy a field in / out; n = nx*ny.  Setting PHI recomputes the surface
   velocities (src/setICs.jl:161-192).  Setting B resets the AB2 history. */
int ssg_set_field(ssg_ctx* ctx, int32_t field, const double* host, int64_t n);
int ssg_get_field(ssg_ctx* ctx, int32_t field, double* host, int64_t n);
/* same with device pointers (FP64, nx*ny), stream-ordered on the context */
int ssg_set_field_device(ssg_ctx* ctx, int32_t field, const double* dev, int64_t n);
int ssg_get_field_device(ssg_ctx* ctx, int32_t field, double* dev, int64_t n);

/* timestep! (src/timestep.jl:267-285) applied nsteps times, including the
   filter cadence `step % filter_freq == 0`.  Updates state->t, ->step,
   ->dt_actual.  SSG_RK3_LOWSTORAGE returns SSG_ERR_UNSUPPORTED, as the
   reference raises (src/timestep.jl:275-277). */
int ssg_timestep(ssg_ctx* ctx, const ssg_time_params* params, ssg_time_state* state, int64_t nsteps);
/* Set-up half of ssg_timestep (no reference counterpart; the reference has
   no launch set-up): runs the autotuner and instantiates the hipGraphs that
   ssg_timestep(ctx, params, state, nsteps) will replay from this state,
   without advancing it.  Step runs are replayed as graphs of up to one
   filter cycle each, keyed by their phase in the cycle and their length, so
   a prepared call launches no capture.  ssg_timestep prepares by itself;
   calling this first keeps the capture out of a timed call. */
int ssg_timestep_prepare(ssg_ctx* ctx, const ssg_time_params* params, const ssg_time_state* state, int64_t nsteps);

/* compute_tendency! (src/nonlinear.jl:22-36) of the current b, φₛ into
   out (nx*ny).  ab2_reference_aliasing = 1 gives the tendency the AB2 call
   site produces (src/timestep.jl:159). */
int ssg_tendency(ssg_ctx* ctx, double* out, int32_t ab2_reference_aliasing);
/* Allocation-free variant of ssg_tendency.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under stream or graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  It needs no context: b, phi and out are device pointers to
   nx*ny FP64 fields (x fastest), desc gives the grid (FP64 only; nz is
   ignored) and stream the HIP stream (NULL: the default stream).  The
   kernels are enqueued and the call returns without synchronising; it calls
   no hipMalloc/hipFree.  All intermediates live in one file-scope device
   array of 40 MiB per GPU (held until the process ends), which fits FP64
   grids up to 1024 x 1024 points; larger grids return SSG_ERR_UNSUPPORTED.
   Every call on a GPU shares that array, so calls that overlap on one GPU
   (other streams, other host threads) race; the array's contents are
   rewritten by each call before they are read.  Errors: ssg_last_error(NULL). */
int ssg_tendency_noalloc(const ssg_domain_desc* desc, const double* b, const double* phi, double* out,
                         int32_t ab2_reference_aliasing, void* stream);
/* compute_jacobian! (src/nonlinear.jl:42-57): -J(φₛ, b), not dealiased */
int ssg_jacobian(ssg_ctx* ctx, double* out);

/* rfft_2d! / irfft_2d! (src/transforms.jl:42-56) in the reference layout;
   spec has 2*nx*(ny/2+1) doubles */
int ssg_rfft_2d(ssg_ctx* ctx, const double* real_in, double* spec_out);
int ssg_irfft_2d(ssg_ctx* ctx, const double* spec_in, double* real_out);

/* apply_spectral_filter! on b (src/timestep.jl:247-260 with
   ExponentialFilter(strength, order, cutoff), src/filter.jl:45-56) */
int ssg_apply_exponential_filter(ssg_ctx* ctx, double strength, int32_t order, double cutoff);

int ssg_diagnostics(ssg_ctx* ctx, ssg_diag* out);

/* solve_monge_ampere_fields! (src/poisson.jl:1068-1090, with the maxiter
   keyword repaired): solves the residual operator of
   compute_ssg_residual! (src/poisson.jl:228-253) with the one-sided
   Neumann rows of apply_ssg_boundary_conditions! (src/poisson.jl:262-364)
   for the current bₛ, then sets φₛ = Φ[:,:,Nz] (extract_surface_to_2d!,
   src/poisson.jl:1119-1131) and the surface velocities.  Needs nz >= 3.
   Returns SSG_ERR_NOT_CONVERGED (with diag filled) when maxiter is hit, when
   the iteration diverges, and, without iterating, when a cold solve of the
   body variant has eps*max(b) >= 1 (diag->not_elliptic = 1). */
int ssg_ma_solve(ssg_ctx* ctx, const ssg_ma_params* params, ssg_ma_diag* diag);
/* Vertical grid of the Monge-Ampère problem: the z (nz levels, strictly
   increasing, surface last) and dz (nz spacings) a Domain carries
   (src/domain.jl:238-281).  dz = NULL applies make_vertical_grid's :custom
   rule (src/domain.jl:258-275: dz[1] = z[2]-z[1], dz[i] = (z[i+1]-z[i-1])/2,
   dz[Nz] = z[Nz]-z[Nz-1]).  The vertical operator uses dz exactly as
   d2dz2! (src/transforms.jl:798-850) and the one-sided Neumann rows
   (src/poisson.jl:309-364) do.  Without this call the grid is uniform,
   dz = Lz/(Nz-1) (src/domain.jl:238-249). */
int ssg_set_vertical_grid(ssg_ctx* ctx, const double* z, const double* dz, int32_t nz);
/* Vertical velocity from the ω equation (compute_vertical_velocity,
   src/extras/omega_eq.jl:574-652) for the current bₛ: Φ from the
   Monge-Ampère solve with params (the ω equation's Φ is the surface-forced
   problem, SSG_MA_SURFACE), Q = (∂x u_g ∂x b + ∂x v_g ∂y b,
   ∂y u_g ∂x b + ∂y v_g ∂y b) per level (compute_divergence_Q!, :326-405),
   (∇ₕ² + ∂zz) w* = -2ε ∇·Q per horizontal wavenumber on the solve's vertical
   stencil with w* = 0 on the bottom and top levels, and w = w* / (1 - ε∇ₕ²Φ)
   (:612-646).  w: nx*ny*nz, x fastest then y then z.  The reference's own
   vertical integrals (:408-521) do not satisfy that equation; this is the
   equation its documentation states. */
int ssg_vertical_velocity(ssg_ctx* ctx, const ssg_ma_params* params, double* w, int64_t n, ssg_ma_diag* diag);
/* fields.φ (nx*ny*nz, x fastest then y then z) after ssg_ma_solve */
int ssg_get_phi3d(ssg_ctx* ctx, double* host, int64_t n);
/* Coupled stepping: with params != NULL, ssg_timestep re-solves the
   Monge-Ampère problem for the stage buoyancy before every tendency
   evaluation (φₛ then follows b); NULL restores the reference's frozen φₛ. */
int ssg_set_coupling(ssg_ctx* ctx, const ssg_ma_params* params);
/* Discrete adjoint (vector-Jacobian product) of `nsteps` timestep! calls
   (RK3, src/timestep.jl:195-242, or AB2, :152-190, including the aliased
   tendency of its call site) with frozen φₛ that ended at step number
   `step_end`: lam_out = (∂b(step_end)/∂b(step_end - nsteps))ᵀ lam_in, with
   the same filter cadence; for AB2 the previous tendency at the window's
   start is held fixed.  The step is linear in b, so no forward trajectory
   is needed.  lam_in, lam_out: host, nx*ny FP64 (may alias).  FP64. */
int ssg_adjoint(ssg_ctx* ctx, const ssg_time_params* params, int64_t step_end, int64_t nsteps, const double* lam_in,
                double* lam_out);

/* Vector-Jacobian product of the Monge-Ampère solve at the solution of the
   last ssg_ma_solve (same eps/variant in params): bbar = (∂φₛ/∂bₛ)ᵀ phibar,
   the transposed linearised solve of solve_monge_ampere_fields!
   (src/poisson.jl:1068-1090; residual operator src/poisson.jl:143-364).
   phibar, bbar: host, nx*ny FP64.  diag (optional) reports the transposed
   iteration; SSG_ERR_NOT_CONVERGED if it misses params->tol. */
int ssg_ma_adjoint(ssg_ctx* ctx, const ssg_ma_params* params, const double* phibar, double* bbar, ssg_ma_diag* diag);
/* Discrete adjoint of `nsteps` coupled steps (RK3, or AB2 with the
   corrected tendency; Monge-Ampère solve before every tendency, params
   `ma`), starting from the context's current b (and AB2 history) at step
   number step_start: lam_out = (∂b(step_start+nsteps)/∂b(step_start))ᵀ lam_in.
   The trajectory is recomputed in FP64 from checkpoints taken every
   `checkpoint_every` steps; the context's b and AB2 history are left at the
   start state.  The step is timestep! (src/timestep.jl:152-242, :267-285)
   with φₛ re-solved before each tendency. */
int ssg_adjoint_coupled(ssg_ctx* ctx, const ssg_time_params* params, const ssg_ma_params* ma, int64_t step_start,
                        int64_t nsteps, int32_t checkpoint_every, const double* lam_in, double* lam_out);

/* set_b! (src/setICs.jl:21-88): b with its k = 0 mode removed and the 2/3
   mask applied (rfft_2d!, dealias_2d!, irfft_2d!), then, when params is not
   NULL, solve_monge_ampere_fields! for φₛ and the surface velocities
   (compute_surface_geostrophic_velocities!).  SSG_ERR_NOT_CONVERGED if the
   solve misses params->tol (b and φₛ are still set). */
int ssg_set_b(ssg_ctx* ctx, const double* host, int64_t n, const ssg_ma_params* params, ssg_ma_diag* diag);
/* compute_cfl_number (src/timestep.jl:91-117) on the surface velocities:
   cfl = max(max|u| dt/dx, max|v| dt/dy); umax, vmax optional. */
int ssg_cfl(ssg_ctx* ctx, double dt, double* cfl, double* umax, double* vmax);
/* Exact restart.  The reference's checkpoint (src/output.jl:318-448) keeps
   t, step, dt, the grid, bₛ, φₛ and the 3D φ but not the AB2 history; this
   byte blob holds the step state (b, φₛ, the AB2 history, t, step), the
   last Monge-Ampère solve's Φ̂ (the 3D φ) and, for coupled stepping, the
   coupling parameters and the warm-start state (the extrapolation ring of
   converged D̂, up to nine nx/2+1 by ny by nz complex FP64 planes).
   Loading it into a context of the same grid and precision restores the
   coupling mode and resumes the frozen-φₛ or the coupled step bitwise; an
   ssg_adjoint_coupled started after the load sees the same start state.
   The size depends on what the context holds (ssg_state_size). */
int ssg_state_size(ssg_ctx* ctx, size_t* bytes);
int ssg_save_state(ssg_ctx* ctx, const ssg_time_state* state, void* buf, size_t bytes);
int ssg_load_state(ssg_ctx* ctx, const void* buf, size_t bytes, ssg_time_state* state);

/* diagnostics of the most recent Monge-Ampère solve */
int ssg_last_ma_diag(ssg_ctx* ctx, ssg_ma_diag* diag);

/* Kernel timing of the last ssg_timestep call (HIP events on the context
   stream): total milliseconds and the number of launches. */
int ssg_last_timing(ssg_ctx* ctx, double* ms, int64_t* launches);

/* Per-launch HIP event timing for roofline reporting.  ssg_profile(1)
   starts recording an event pair around every kernel launch (and clears
   earlier records); ssg_profile_read sums, for one kernel class, the
   measured milliseconds, the launch count and the algorithmic bytes (every
   input read once, every output written once).  Classes: 0 = row pass
   forward (k_rows_fwd), 1 = column pass (k_cols), 2 = row pass inverse with
   fused stage epilogue (k_rows_inv), 3 = Monge-Ampère row pass (k_ma_rows),
   4 = Monge-Ampère per-wavenumber solve (k_ma_tridiag), 5 = the same solve
   forming its extrapolated warm start from the solve history; 11, 12 and 13
   = the FP32 launches of classes 3, 4 and 5. */
int ssg_profile(ssg_ctx* ctx, int32_t enable);

/* Autotuned launch shape of the three line-FFT kernel classes for one path
   (0 = FP64, 1 = FP32-mixed): ept[3] receives the elements per thread in
   use; us[9] (if not NULL) the measured microseconds per launch of the
   candidates 4, 8, 16 per class (-1 where not timed).  Tuning runs on the
   first ssg_timestep of a path. */
int ssg_tuning(ssg_ctx* ctx, int32_t path, int32_t* ept, double* us);

/* Measurement aid with no reference counterpart: runs a plain device copy
   kernel (k_convert<double,double>) of `bytes` bytes `reps` times, the
   calibration point for hardware byte counters (rocprofv3 FETCH_SIZE /
   WRITE_SIZE). */
int ssg_copy_probe(ssg_ctx* ctx, int64_t bytes, int32_t reps);
int ssg_profile_read(ssg_ctx* ctx, int32_t kind, double* total_ms, int64_t* count, double* bytes);

/* ------------------------------------------------------------------------
   Ensemble (batched form): members with their own grids, advanced together
   with one launch per kernel class per RK stage.  No reference counterpart:
   the reference runs one domain per MPI job.
   ------------------------------------------------------------------------ */
typedef struct ssg_ensemble ssg_ensemble;

/* members share precision and device; nx, ny in {128, 256, 512, 1024} */
int ssg_ensemble_create(ssg_ensemble** out, int32_t nmembers, const ssg_domain_desc* members);
int ssg_ensemble_destroy(ssg_ensemble* ens);
/* member context (owned by the ensemble): set fields, solve Monge-Ampère,
   read diagnostics through the single-domain calls */
ssg_ctx* ssg_ensemble_member(ssg_ensemble* ens, int32_t index);
const char* ssg_ensemble_last_error(const ssg_ensemble* ens);
/* nsteps RK3 timestep! calls (frozen φₛ) of every member; dt[i] per member
   (params->dt is ignored), states[i] updated; members at the same step */
int ssg_ensemble_timestep(ssg_ensemble* ens, const ssg_time_params* params, const double* dt, ssg_time_state* states,
                          int64_t nsteps);
/* autotuned elements per thread of the batched launches and the candidate
   timings (us per RK stage pair, -1 where not valid) for 4, 8, 16 */
int ssg_ensemble_tuning(ssg_ensemble* ens, int32_t* ept, double* us);

#ifdef __cplusplus
}
#endif

#endif /* SSG_H */
