/*
 * gdhip.h -- C ABI of the MI355X-native pseudo-spectral CNAB2 step.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * GD_ERR_* code; gd_last_error() returns a thread-local message for the last
 * failure on the calling thread.  Unless GD_HOST is set in `flags`, array
 * arguments are device pointers and work is queued on the context's stream
 * (asynchronous); with GD_HOST the arrays are host memory, copied in and out,
 * and the call returns after the results are in host memory.  The library
 * never retains caller pointers after a call returns.
 *
 * Layouts are the reference's (GeoDynamo.jl) boundary layouts:
 *   spectral  re/im planes of (nlm, 1, nr) column-major, l-major lm order:
 *             element (lm, r) at lm + nlm*r                 (fields.jl:6-15, shtnskit_transforms.jl:177-188)
 *   physical  (nlat, nlon, nr) column-major, colatitude fastest, north first
 *                                                           (fields.jl:18-24)
 *   band      (2bw+1, N) column-major, A[i][j] = ab[bw+i-j + (2bw+1)*j]
 *                                                           (linear_algebra.jl:3-21)
 *   radial    r table (N, 7) column-major, column p = r^(p-4) (fields.jl:108-131)
 *   step      5 fields in the order u_tor, u_pol, B_tor, B_pol, T, each a
 *             spectral (nlm, nr) block: field f at offset f*nlm*nr.
 */
#ifndef GDHIP_H
#define GDHIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    GD_OK = 0,
    GD_ERR_ARG = 1,      /* bad argument / shape */
    GD_ERR_HIP = 2,      /* HIP runtime failure */
    GD_ERR_NUMERIC = 3,  /* zero or non-finite pivot in a no-pivot band LU */
    GD_ERR_MISSING_L = 4,/* an lm column's l has no implicit matrix (timestep.jl:646) */
    GD_ERR_NOMEM = 5
};
enum { GD_PREC_FP64 = 0, GD_PREC_FP32 = 1, GD_PREC_BF16X3 = 2 };
enum { GD_HOST = 1 };
enum { GD_FD_TRANSPOSE = 1, GD_FD_LITERAL = 0 };

typedef struct gd_ctx gd_ctx;
typedef struct gd_sht_plan gd_sht_plan;
typedef struct gd_cnab2_mats gd_cnab2_mats;
typedef struct gd_step gd_step;

const char *gd_last_error(void);
const char *gd_version(void);

/* Device context: one HIP device and one stream (default priority). */
int gd_ctx_create(int device, gd_ctx **out);
/* Use an existing stream (e.g. the caller's framework stream); 0 = library's own. */
int gd_ctx_set_stream(gd_ctx *ctx, void *hip_stream);
int gd_ctx_synchronize(gd_ctx *ctx);
int gd_ctx_destroy(gd_ctx *ctx);
/* Device memory helpers for callers without their own allocator (no
 * counterpart in the reference): hipMalloc/hipFree and a plain 16-byte-vector
 * copy kernel on the context's stream (bytes a multiple of 16). */
int gd_device_alloc(gd_ctx *ctx, size_t bytes, void **out);
int gd_device_free(gd_ctx *ctx, void *p);
int gd_device_copy(gd_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- radial domain and FD operators (host FP64 set-up) -------------------
 * replaces create_radial_domain (fields.jl:108-131),
 * create_derivative_matrix (linear_algebra.jl:90-132),
 * create_radial_laplacian (linear_algebra.jl:134-151).
 * fd_mode GD_FD_TRANSPOSE solves V^T w = k! e_k; GD_FD_LITERAL is V \ rhs. */
int gd_radial_grid(int nr, double rratio, double *r /* nr x 7 */);
int gd_derivative_matrix(int order, int nr, int bw, const double *r, int fd_mode, double *band);
int gd_radial_laplacian(int nr, int bw, const double *r, int fd_mode, double *band);

/* Seeded spectral state of the configurations (host): splitmix64 -> Box-Muller
 * normals drawn in the order field, level, lm, (re, im); value
 * amp * n * (l+1)^-power, m = 0 real, rows 0 and nr-1 zeroed when
 * zero_boundaries.  No counterpart in the reference (its initial conditions
 * use Julia's RNG stream, InitialConditions.jl); the configurations define
 * their states this way.  re/im: nfields blocks of (nlm, nr). */
int gd_seeded_state(unsigned long long seed, int lmax, int mmax, int nr, int nfields, double amp,
                    double power, int zero_boundaries, double *re, double *im);

/* ---- banded LU without pivoting (FP64, GPU) --------------------------------
 * replaces factorize_banded (linear_algebra.jl:23-57) for nsys systems
 * (ab and lu: nsys consecutive (2bw+1) x n bands) and solve_banded!
 * (linear_algebra.jl:59-88) for nrhs right-hand sides x (n x nrhs, column
 * j uses system sys_of_rhs[j]; x may alias b).  Factor returns
 * GD_ERR_NUMERIC if a pivot is zero or non-finite. */
int gd_banded_factor(gd_ctx *ctx, int n, int bw, int nsys, const double *ab, double *lu, int flags);
int gd_banded_solve(gd_ctx *ctx, int n, int bw, int nsys, const double *lu,
                    const int *sys_of_rhs, int nrhs, double *x, int flags);

/* ---- CNAB2 implicit operators (FP64, GPU) ---------------------------------
 * replaces create_shtns_timestepping_matrices (timestep.jl:498-543): for
 * l = 0..lmax, L_l = nu(Lap - l(l+1)/r^2), S_l = I/dt - theta L_l, LU(S_l);
 * matrices stay device resident.  bc_rows = 0 is the reference operator (no
 * boundary rows); 1 replaces rows 0 and nr-1 of S_l by Dirichlet rows and
 * sets those rhs rows to 0 (zero boundary values). */
int gd_cnab2_mats_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu,
                         double dt, double theta, int fd_mode, int bc_rows, gd_cnab2_mats **out);
int gd_cnab2_mats_get(gd_cnab2_mats *M, double *L, double *S, double *LU /* host, (lmax+1) bands each */);
int gd_cnab2_mats_destroy(gd_cnab2_mats *M);
/* replaces build_rhs_cnab2! (timestep.jl:610-690):
 * rhs = u/dt + 1.5 nl - 0.5 nl_prev + (1-theta) L_l u, for spectral fields with
 * the given (lmax, mmax); columns whose l exceeds the matrices' lmax give
 * GD_ERR_MISSING_L, as the reference errors. */
int gd_build_rhs_cnab2(gd_cnab2_mats *M, int lmax, int mmax, int nr,
                       const double *u_re, const double *u_im,
                       const double *nl_re, const double *nl_im,
                       const double *np_re, const double *np_im,
                       double *rhs_re, double *rhs_im, int flags);
/* replaces solve_implicit_step! (timestep.jl:692-730): sol = S_l^{-1} rhs per
 * (lm, re/im); columns whose l has no matrix are skipped, as the reference does. */
int gd_solve_implicit_step(gd_cnab2_mats *M, int lmax, int mmax, int nr,
                           const double *rhs_re, const double *rhs_im,
                           double *sol_re, double *sol_im, int flags);
/* fused build_rhs_cnab2! + solve_implicit_step! in place on u. */
int gd_cnab2_rhs_solve(gd_cnab2_mats *M, int lmax, int mmax, int nr,
                       double *u_re, double *u_im,
                       const double *nl_re, const double *nl_im,
                       const double *np_re, const double *np_im, int flags);

/* ---- spherical-harmonic transforms ----------------------------------------
 * replaces create_shtnskit_config (shtnskit_transforms.jl:102-200):
 * orthonormal Y_lm with Condon-Shortley phase on a Gauss grid (nlat even,
 * nlon > 2 mmax, nlon = 2^a 3^b), P_lm by FP64 recurrence on the GPU.
 * prec selects the Legendre contraction / FFT arithmetic. */
int gd_sht_plan_create(gd_ctx *ctx, int lmax, int mmax, int nlat, int nlon, int prec,
                       gd_sht_plan **out);
int gd_sht_plan_destroy(gd_sht_plan *p);
int gd_sht_nlm(const gd_sht_plan *p);
/* host copies of the grid (cos theta, Gauss weights, theta) and FP64 tables
 * P_lm, dP_lm/dtheta in l-major (nlm x nlat) */
int gd_sht_grid(gd_sht_plan *p, double *x, double *w, double *theta);
int gd_sht_tables(gd_sht_plan *p, double *P, double *dP);
/* replaces shtnskit_spectral_to_physical! (shtnskit_field_functions.jl:11-21) */
int gd_sht_scalar_synth(gd_sht_plan *p, const double *re, const double *im, int nr,
                        double *phys, int flags);
/* replaces shtnskit_physical_to_spectral! (shtnskit_field_functions.jl:131-141);
 * imag(m=0) := 0 as the reference's store does (:432-435) */
int gd_sht_scalar_analysis(gd_sht_plan *p, const double *phys, int nr,
                           double *re, double *im, int flags);
/* replaces shtnskit_vector_synthesis! (shtnskit_field_functions.jl:252-282):
 * (v_theta, v_phi) from spheroidal S := pol and toroidal T := tor */
int gd_sht_vector_synth(gd_sht_plan *p, const double *tor_re, const double *tor_im,
                        const double *pol_re, const double *pol_im, int nr,
                        double *vt, double *vp, int flags);
/* replaces shtnskit_vector_analysis! (shtnskit_field_functions.jl:291-321) */
int gd_sht_vector_analysis(gd_sht_plan *p, const double *vt, const double *vp, int nr,
                           double *tor_re, double *tor_im, double *pol_re, double *pol_im,
                           int flags);
/* Transposes of the four transforms above (the adjoint of the SHT pair
 * shtnskit_field_functions.jl:11-21, 131-141, 252-321, whose derivatives the
 * reference takes through SHTnsKit's ChainRulesCore extension,
 * Manifest.toml:657-669): for the real inner products sum re*re' + im*im'
 * on spectral planes and sum f*f' on grid values, <A x, y> = <x, A^T y>.
 * Arguments mirror the forward calls with inputs and outputs exchanged. */
int gd_sht_scalar_synth_adjoint(gd_sht_plan *p, const double *phys, int nr, double *re, double *im,
                                int flags);
int gd_sht_scalar_analysis_adjoint(gd_sht_plan *p, const double *re, const double *im, int nr,
                                   double *phys, int flags);
int gd_sht_vector_synth_adjoint(gd_sht_plan *p, const double *vt, const double *vp, int nr,
                                double *tor_re, double *tor_im, double *pol_re, double *pol_im,
                                int flags);
int gd_sht_vector_analysis_adjoint(gd_sht_plan *p, const double *tor_re, const double *tor_im,
                                   const double *pol_re, const double *pol_im, int nr,
                                   double *vt, double *vp, int flags);

/* ---- the device-resident CNAB2 step ----------------------------------------
 * replaces the bench step of scripts/bench_timestepping.jl:6-18: the
 * nonlinear terms of thermal.jl:149-190, velocity.jl:312-332 and
 * magnetic.jl:131-159, then apply_master_implicit_step!
 * (simulation.jl:1151-1330). */
typedef struct {
    double E, Pm, Pr, Ra, rratio, dt, theta;
    int lmax, mmax, nlat, nlon, nr, bw;
    int prec;           /* GD_PREC_*: Legendre/FFT/products arithmetic; solves stay FP64 */
    int fd_mode;        /* GD_FD_TRANSPOSE (default) or GD_FD_LITERAL */
    int radials;        /* 0: radial components stay 0 as in the reference (default);
                           1: X_r = l(l+1) P_X / r^2 from the poloidal scalars */
    int coriolis_index; /* 0: reference indexing theta_{min(j,nlat)}; 1: theta_i */
    int bc_rows;        /* Dirichlet rows (value 0) in S_l: 1 all fields (default), 2 temperature
                           only, 0 none (the reference operator, unstable at these parameters) */
    int full_products;  /* 0 (default): with radials = 0 the tangential parts of u x omega,
                           j x B and u x B and the u_r dT/dr term are identically zero, so only
                           u and grad T are transformed (same result: the dropped terms are
                           exact zeros); 1: transform every field and form every product */
    int radial_force;   /* 0 (default): the reference's data flow, where the radial force
                           (buoyancy on adv_r, velocity.jl:534-537) reaches no equation; 1: the
                           buoyancy F_r = (Pm/E)(Pm/Pr) Ra r T enters the poloidal equation as the
                           poloidal scalar whose radial component is F_r under the reference's
                           definition u_r = l(l+1) P / r^2 (magnetic.jl:257-259):
                           nl(u_pol) += r^2 F_r / (l(l+1)), l >= 1 */
} gd_step_desc;

int gd_step_create(gd_ctx *ctx, const gd_step_desc *desc, gd_step **out);
int gd_step_destroy(gd_step *s);
/* state: 5 x (nlm, nr) re/im in the step layout above; resets the step counter */
int gd_step_set_state(gd_step *s, const double *re, const double *im, int flags);
int gd_step_get_state(gd_step *s, double *re, double *im, int flags);
/* evaluates the nonlinear terms of the current state (no time advance) */
int gd_step_nonlinear(gd_step *s, double *nl_re, double *nl_im, int flags);
/* advances nsteps CNAB2 steps (AB1 bootstrap at step 0, as simulation.jl:1153) */
int gd_step_run(gd_step *s, int nsteps);
/* same, replaying a captured hipGraph of one step */
int gd_step_run_graph(gd_step *s, int nsteps);
int gd_step_count(const gd_step *s);
/* FP64 energies: KE and ME (1/2 sum (T^2+P^2)/max(l(l+1),1), magnetic.jl:504-543)
 * and thermal 1/2 sum |T|^2 (thermal.jl:413-438); host out[3] */
int gd_step_energies(gd_step *s, double *out3);
/* Measurement hook with no counterpart in the reference: average device ms per
 * step of each kernel class over nsteps eager steps, timed with HIP events on
 * the step's stream.  out[GD_NKCLASS]: 0 spectral prep, 1 Legendre synthesis
 * (vector), 2 Legendre synthesis (scalar), 3 ring FFT + products, 4 Legendre
 * analysis (vector), 5 Legendre analysis (scalar), 6 CNAB2 rhs + banded solve. */
#define GD_NKCLASS 7
int gd_step_profile(gd_step *s, int nsteps, double *out);
/* Batches of independent shells (no counterpart in the reference, whose shells
 * are separate runs): steps of any (lmax, nr, grid) on one context advanced
 * together.  Each graph replay holds every shell's step, the shells spread
 * round-robin over nstreams streams forked from and joined to the context
 * stream.  All shells must be at the same step count; step 0 (the AB1
 * bootstrap) runs eagerly.  A shell's results equal its own gd_step_run_graph.
 * A member whose captured work changes (conditions, composition) makes the
 * batch recapture; member steps must outlive the batch. */
typedef struct gd_batch gd_batch;
int gd_batch_create(gd_ctx *ctx, int n, gd_step *const *steps, int nstreams, gd_batch **out);
int gd_batch_run(gd_batch *b, int nsteps);
int gd_batch_destroy(gd_batch *b);
/* Launch form of the batch graphs (default GD_BATCH_STREAMS, as above).
 * GD_BATCH_FUSED: one launch per kernel form and stage -- Legendre synthesis,
 * ring FFT + products, vector analysis, scalar analysis, CNAB2 column solves --
 * over the blocks of every shell that shares the form (synthesis tile, ring
 * length, analysis node tile, band width), largest shells first, all on the
 * context stream.  Shells whose step has no fused form (composition, full
 * products, radials) run their own step in the same graph.  Results equal
 * each shell's own gd_step_run_graph bit for bit. */
#define GD_BATCH_STREAMS 0 /* FP64 shells only (GD_ERR_ARG otherwise) */
#define GD_BATCH_FUSED 1
int gd_batch_set_mode(gd_batch *b, int mode);
/* Legendre-synthesis tile of each shell in the fused form, tiles[n]:
 * 0 = 3 waves x 8 levels with 16-degree chunks, 1 = 3 waves with 32-degree
 * chunks, 2 = 6 waves x 16 levels with 16-degree chunks (FP32 and BF16x3
 * shells only).  Every tile gives the same bits. */
int gd_batch_set_tiles(gd_batch *b, const int *tiles);
int gd_batch_get_tiles(const gd_batch *b, int *tiles);
/* Autotuner of the fused form's synthesis tiles on this device: each shell's
 * synthesis timed alone in every tile (mean of reps launches) gives a per-shell
 * choice; that assignment and each uniform one (every shell on tile 0, 1, 2
 * where available) are timed as the batch's fused synthesis launches, and the
 * fastest is kept (gd_batch_get_tiles reads it).  ms_out (host, 4 entries, or
 * NULL): the mean ms of those four assignments in that order.  Writes only
 * the shells' transform scratch; the states are untouched. */
int gd_batch_tune(gd_batch *b, int reps, double *ms_out);
/* Lanes of the fused form (default 1): the fused shells split over that many of
 * the batch's streams (1 .. nstreams), balanced by Legendre work, each lane
 * running its own fused stages concurrently with the others.  Same bits for
 * every lane count. */
int gd_batch_set_lanes(gd_batch *b, int lanes);
int gd_batch_get_lanes(const gd_batch *b, int *lanes);
/* Lane autotuner (fused mode): lanes 1, 2, 4, 8 (up to nstreams and the shell
 * count), each timed over nsteps of the batch's own steps after 2 + nsteps
 * untimed ones; keeps the fastest.  The shells advance by 2 nsteps + 2 steps
 * per candidate.  ms_out (host, 4 entries, or NULL): ms per step of each
 * candidate, -1 where not run. */
int gd_batch_tune_lanes(gd_batch *b, int nsteps, double *ms_out);

/* ---- boundary conditions of a scalar field ----------------------------------
 * replaces apply_temperature_boundary_conditions_spectral! (thermal.jl:262-313)
 * and apply_composition_boundary_conditions_spectral! (compositional.jl:318-370)
 * with apply_scalar_flux_bc_spectral!(...; method=:tau)
 * (scalar_field_common.jl:1166-1199, :866-929).  Per (l, m): bc type 1 sets
 * the boundary row to boundary_values (imag 0); type 2 adds the tau-function
 * correction that removes the flux error D1 p - value of the real profile at
 * that boundary (both boundaries: T_{N-1}, T_N with analytic end derivatives;
 * one boundary: T_{N-1} - a T_N, its derivative through D1), to the imaginary
 * profile too when one of its entries exceeds 1e-12 (the reference's rule).
 * r: the nr x 7 radial table (gd_radial_grid); D1 from gd_derivative_matrix
 * with bw and fd_mode.  bc_inner/bc_outer: nlm ints; values: Julia
 * boundary_values[2, nlm] (inner, outer per mode), l-major modes. */
typedef struct gd_scalar_bc gd_scalar_bc;
int gd_scalar_bc_create(gd_ctx *ctx, int lmax, int mmax, int nr, int bw, const double *r, int fd_mode,
                        gd_scalar_bc **out);
int gd_scalar_bc_set(gd_scalar_bc *b, const int *bc_inner, const int *bc_outer, const double *values);
/* re, im: (nlm, nr) planes as gd_build_rhs_cnab2; dirichlet = 1 applies the
 * type-1 rows first (the temperature / composition functions), 0 only the flux
 * correction (apply_scalar_flux_bc_spectral! alone) */
int gd_scalar_bc_apply(gd_scalar_bc *b, double *re, double *im, int dirichlet, int flags);
/* host copies: tab[4][nr] = T_{N-1}, T_N, inner and outer tau functions;
 * scalars[6] = end derivatives dT_{N-1}(r_i), dT_{N-1}(r_o), dT_N(r_i), dT_N(r_o),
 * D1 derivative of the inner function at r_i and of the outer one at r_o */
int gd_scalar_bc_tau(const gd_scalar_bc *b, double *tab, double *scalars);
int gd_scalar_bc_destroy(gd_scalar_bc *b);

/* ---- restart files ---------------------------------------------------------
 * The step's state and its AB2 history.  gd_step_get_history returns the
 * nonlinear terms of the previous step (the reference keeps them as
 * prev_nonlinear, simulation.jl:1312-1329; GD_ERR_ARG before the first step).
 * gd_step_restore sets state, history and step counter together (step 0: no
 * history, the next step is the AB1 bootstrap); a run continued from a restore
 * takes exactly the uninterrupted run's steps.  Planes as gd_step_set_state. */
int gd_step_get_history(gd_step *s, double *nl_re, double *nl_im, int flags);
int gd_step_restore(gd_step *s, const double *re, const double *im, const double *nl_re,
                    const double *nl_im, int step, int flags);
int gd_step_shape(const gd_step *s, int *lmax, int *mmax, int *nr);
/* Per-mode temperature conditions of the step: the temperature field's
 * bc_type_inner / bc_type_outer / boundary_values[2, nlm] (l-major, as
 * gd_scalar_bc_set), applied as apply_temperature_boundary_conditions_spectral!
 * does (thermal.jl:262-313, tau flux correction for type 2) to the state
 * between the nonlinear terms and the implicit step.  With bc_rows 1 or 2 the
 * temperature's Dirichlet rows of S_l take the boundary values the conditions
 * leave in the state.  NULL arrays: back to the default (boundary rows read as
 * 0, the reference's default Dirichlet 0).  Drops the step's captured graphs;
 * a batch holding the step recaptures its graphs before its next run. */
int gd_step_set_temperature_bc(gd_step *s, const int *bc_inner, const int *bc_outer, const double *values);
/* Composition (compositional.jl): a passive scalar C of the tangential step
 * (radials = 0, full_products = 0), advected by u (-u.grad C, as the
 * temperature), CNAB2 with diffusivity Pm/Sc (simulation.jl:130-131), its
 * conditions applied before its solve (default: no flux at both boundaries,
 * value 0, compositional.jl:110-115; set with gd_step_set_composition_bc).
 * Planes (nlm, nr) re/im, l-major; set after gd_step_set_state.  Restart files
 * carry it as composition_spectral (with its history). */
int gd_step_enable_composition(gd_step *s, double Sc);
int gd_step_set_composition(gd_step *s, const double *re, const double *im, int flags);
int gd_step_get_composition(gd_step *s, double *re, double *im, int flags);
int gd_step_set_composition_bc(gd_step *s, const int *bc_inner, const int *bc_outer, const double *values);
int gd_step_has_composition(const gd_step *s);
/* its AB2 history (GD_ERR_ARG before the first step) and its restore after a
 * gd_step_restore (same planes; NULL history when that restore was at step 0) */
int gd_step_get_composition_history(gd_step *s, double *nl_re, double *nl_im, int flags);
int gd_step_restore_composition(gd_step *s, const double *re, const double *im, const double *nl_re,
                                const double *nl_im, int flags);
/* radial grid r_i of the step's shell (fields.jl:108-131), host out[nr] */
int gd_step_radii(const gd_step *s, double *r);

/* Restart metadata: the TimeTracker and run fields that write_restart!
 * (outputs_writer.jl:1221-1273) stores beside the fields; the step number is
 * the step's own counter. */
typedef struct {
    double time;             /* simulation time: "time", "current_time", "restart_time" */
    double last_output_time; /* "last_output_time" */
    int output_count, restart_count, grid_file_written;
    int mpi_rank, mpi_total_ranks;
    char geometry[32];       /* "geometry" ("shell" when empty) */
} gd_restart_meta;

/* replaces write_restart! (outputs_writer.jl:1221-1273): NetCDF classic
 * (CDF-2) file with the reference's dimensions, variables and attributes
 * (velocity_toroidal_real ... temperature_spectral_imag as (r, spectral_mode),
 * l_values, m_values, r, time, step, the restart scalars), plus the AB2 history
 * as <component>_prev_nonlinear_{real,imag} when the step count is > 0.
 * precision: GD_PREC_FP64 or GD_PREC_FP32 (config.output_precision). */
int gd_restart_write(gd_step *s, const char *path, const gd_restart_meta *meta, int precision);
/* replaces read_restart! (outputs_writer.jl:1275-1370) for one file: checks
 * spectral_mode and r against the step, maps modes by l_values/m_values,
 * restores state and history (a file without history restores at step 0);
 * meta (may be NULL) receives time, tracker scalars and geometry. */
int gd_restart_read(gd_step *s, const char *path, gd_restart_meta *meta);
/* The composition's planes of a restart file, (nlm, nr) each, l-major
 * (composition_spectral_{real,imag} and its _prev_nonlinear history).  Write:
 * re/im NULL = none; the history is written with the fields' history.  Read:
 * present / has_history report the file, the planes are filled when given. */
typedef struct {
    double *re, *im, *nl_re, *nl_im;
    int present, has_history;
} gd_restart_comp;

/* host-only forms of the two on l-major planes 5 x (nlm, nr) (no GPU used);
 * nl_re/nl_im may be NULL (no history written / not returned); comp may be
 * NULL (no composition) */
int gd_restart_write_arrays(const char *path, const gd_restart_meta *meta, int lmax, int mmax, int nr,
                            const double *r, int step, const double *re, const double *im,
                            const double *nl_re, const double *nl_im, int precision,
                            const gd_restart_comp *comp);
int gd_restart_read_arrays(const char *path, int lmax, int mmax, int nr, double *r, int *step,
                           double *re, double *im, double *nl_re, double *nl_im, int *has_history,
                           gd_restart_meta *meta, gd_restart_comp *comp);
/* replaces generate_filename (outputs_writer.jl:338-361):
 * <dir>/<prefix>_<type>_geom_<geometry>_rank_XXXX_time_<t>.nc with t = %.6f, "." -> "p", or
 * time before rank when time_first = 1 (TIME_RANK naming) */
int gd_restart_filename(const char *output_dir, const char *prefix, const char *file_type,
                        const char *geometry, double time, int rank, int time_first, char *buf,
                        int buflen);

/* ---- adjoint of the step ---------------------------------------------------
 * Reverse-mode derivatives of the tangential step (radials = 0,
 * full_products = 0; no temperature conditions or composition): the
 * transposed CNAB2 solve (U^T then L^T of the stored no-pivot LU,
 * linear_algebra.jl:59-88, timestep.jl:610-730) and the transposed
 * nonlinear terms through the transposed SHT pair
 * (shtnskit_field_functions.jl:252-321) and the transposed pointwise products
 * (velocity.jl:512-527, scalar_field_common.jl:401-407).  The reference
 * takes such derivatives through SHTnsKit.jl's ChainRulesCore extension
 * (Manifest.toml:657-669, SHTnsKit AdvancedADExt); the Julia binding wraps
 * the entries below in rrules.  Planes as gd_step_set_state; a cotangent is
 * (dJ/d re, dJ/d im) per coefficient. */
/* lam = N'(x)^T mu at the step's current state x (N: gd_step_nonlinear) */
int gd_step_nonlinear_adjoint(gd_step *s, const double *mu_re, const double *mu_im, double *lam_re,
                              double *lam_im, int flags);
/* the transposed step at the current state (the AB1 bootstrap when the step
 * count is 0): given the cotangents lam of the new state and eta of this
 * step's nonlinear terms (the history the next step reads), returns the
 * cotangents of the state (olam) and of the previous step's nonlinear terms
 * (oeta; zero at the bootstrap) */
int gd_step_vjp(gd_step *s, const double *lam_re, const double *lam_im, const double *eta_re,
                const double *eta_im, double *olam_re, double *olam_im, double *oeta_re, double *oeta_im,
                int flags);
/* The adjoint of an nsteps run from step 0 with checkpoints (state and AB2
 * history) every `every` steps; the backward pass recomputes each segment's
 * states from its checkpoint.  forward: runs the steps from the step's
 * current state (it must be at step 0) and keeps the checkpoints.  misfit:
 * J = 1/2 sum over the nlat x nlon grid of (X - target)^2, X the field
 * `kind` (0 B_r = l(l+1) B_pol / r^2, magnetic.jl:257-259; 1 u_r from u_pol;
 * 2 T) at radial level `level` of the final state, synthesised in FP64
 * (colatitude fastest, as gd_sht_scalar_synth with nr = 1); seeds the final
 * cotangent.  seed: the final cotangent given directly.  backward: the
 * gradient of the seeded functional with respect to the initial state; the
 * step is left at the end of the forward run. */
typedef struct gd_adjoint gd_adjoint;
int gd_adjoint_create(gd_step *s, int nsteps, int every, gd_adjoint **out);
int gd_adjoint_destroy(gd_adjoint *a);
int gd_adjoint_forward(gd_adjoint *a);
int gd_adjoint_misfit(gd_adjoint *a, int kind, int level, const double *target, double *J, int flags);
int gd_adjoint_seed(gd_adjoint *a, const double *re, const double *im, int flags);
int gd_adjoint_backward(gd_adjoint *a, double *g_re, double *g_im, int flags);

/* replaces compute_timestep_error (timestep.jl:733-753) on device arrays */
int gd_timestep_error(gd_ctx *ctx, long n, const double *a_re, const double *a_im,
                      const double *b_re, const double *b_im, double *out, int flags);
/* No counterpart in the reference: gd_timestep_error for callers that cannot
 * allocate (code under stream/graph capture, tight serving loops).  Device
 * arrays only; the result is written to the device double `out` on the
 * context's stream, asynchronously.  The reduction's partial sums use a fixed
 * 2 KB static device scratch per GPU (no hipMalloc or other allocation), so
 * any n >= 0 is accepted; results equal gd_timestep_error's bit for bit.
 * Calls that overlap on one GPU share the scratch and race. */
int gd_timestep_error_noalloc(gd_ctx *ctx, long n, const double *a_re, const double *a_im,
                              const double *b_re, const double *b_im, double *out);

/* ---- exponential integrators: EAB2 and staged ERK2 (FP64) -----------------
 * A_l = nu (d2/dr2 + (2/r) d/dr - l(l+1)/r^2) for l = 0..lmax, dense nr x nr
 * (no boundary rows), from the band Laplacian of gd_radial_laplacian.  Matrix
 * functions are computed on the host (exp by [13/13] Pade scaling and
 * squaring) and kept on the device; every matrix is Julia's Matrix layout,
 * nr x nr column-major, one per l in l order.  Field arguments are l-major
 * spectral re/im planes (nlm, nr) of a field with the given (lmax, mmax). */
typedef struct gd_etd_cache gd_etd_cache;
typedef struct gd_erk2_cache gd_erk2_cache;
typedef struct gd_erk2_buffers gd_erk2_buffers;
enum { GD_EXP_ETD = 0, GD_EXP_ERK2 = 1 };

/* Host only (no device, no context): the matrices the caches below hold.
 * GD_EXP_ETD: E = exp(dt A_l), phi1 = A_l^-1 (E - I)/dt      (timestep.jl:37-65)
 * GD_EXP_ERK2: exp(dt/2 A_l), exp(dt A_l), phi1(dt/2 A_l), phi1(dt A_l),
 *   phi2(dt A_l) with the reference's series fallbacks      (timestep.jl:896-1187)
 * out: 2 (ETD) or 5 (ERK2) blocks of (lmax+1) matrices.  GD_ERR_NUMERIC when
 * an exponential fails or (ETD) A_l is exactly singular, where the reference's
 * exp/lu throw. */
int gd_exp_functions(int lmax, int nr, int bw, const double *r, double nu, double dt, int fd_mode,
                     int which, double *out);

/* replaces create_etd_cache (timestep.jl:37-65) */
int gd_etd_cache_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu, double dt,
                        int fd_mode, gd_etd_cache **out);
int gd_etd_cache_get(gd_etd_cache *c, double *E, double *phi1 /* host, (lmax+1) matrices each */);
int gd_etd_cache_destroy(gd_etd_cache *c);
/* replaces eab2_update! (timestep.jl:425-487): u = E_l u + dt phi1_l (3/2 nl - 1/2 nl_prev)
 * in place per (lm, re/im); GD_ERR_MISSING_L if the field has an l the cache lacks. */
int gd_eab2_update(gd_etd_cache *c, int lmax, int mmax, int nr, double *u_re, double *u_im,
                   const double *nl_re, const double *nl_im, const double *np_re, const double *np_im,
                   double dt, int flags);

/* replaces create_erk2_cache (timestep.jl:896-968; dense path, use_krylov = false) */
int gd_erk2_cache_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu, double dt,
                         int fd_mode, gd_erk2_cache **out);
/* the ERK2Cache constructor (timestep.jl:830-851): nl caller-given l values and
 * their five nl-matrix sets (host, column-major nr x nr each) */
int gd_erk2_cache_from_matrices(gd_ctx *ctx, int nl, const int *l_values, int nr, double dt,
                                const double *E_half, const double *E_full, const double *phi1_half,
                                const double *phi1_full, const double *phi2_full, gd_erk2_cache **out);
int gd_erk2_cache_get(gd_erk2_cache *c, double *E_half, double *E_full, double *phi1_half,
                      double *phi1_full, double *phi2_full /* host, nslots matrices each */);
int gd_erk2_cache_destroy(gd_erk2_cache *c);
/* replaces ERK2FieldBuffers (timestep.jl:1283-1313): per-field device planes */
int gd_erk2_buffers_create(gd_erk2_cache *c, int lmax, int mmax, int nr, gd_erk2_buffers **out);
int gd_erk2_buffers_destroy(gd_erk2_buffers *b);
/* replaces erk2_prepare_field! (timestep.jl:1316-1438): n_current = nl,
 * linear = E_full u, k1 = phi1_full nl, stage = E_half u + dt/2 phi1_half nl */
int gd_erk2_prepare_field(gd_erk2_buffers *b, const double *u_re, const double *u_im,
                          const double *nl_re, const double *nl_im, double dt, int flags);
/* replaces erk2_apply_stage! (timestep.jl:1440-1444): u = stage */
int gd_erk2_apply_stage(gd_erk2_buffers *b, double *u_re, double *u_im, int flags);
/* replaces erk2_store_stage_nonlinear! (timestep.jl:1446-1452): stage_nl = nl */
int gd_erk2_store_stage_nonlinear(gd_erk2_buffers *b, const double *nl_re, const double *nl_im, int flags);
/* replaces erk2_finalize_field! (timestep.jl:1449-1555):
 * u = linear + dt (k1 + phi2_full (stage_nl - n_current)) */
int gd_erk2_finalize_field(gd_erk2_buffers *b, double *u_re, double *u_im, double dt, int flags);

#ifdef __cplusplus
}
#endif
#endif
