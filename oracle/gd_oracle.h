/*
 * gd_oracle.h -- FP64 CPU restatement of GeoDynamo.jl's pseudo-spectral CNAB2
 * step.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and
 * the timed CPU baseline.  No product code links against it.
 *
 * Layouts follow the reference's boundary (fields.jl:6-24, Appendix of the
 * survey):
 *   spectral  re[lm + nlm*r], im[lm + nlm*r], lm l-major:
 *             idx(l,m) = sum_{l'<l}(min(l',mmax)+1) + m     (shtnskit_transforms.jl:177-188)
 *   physical  f[i + nlat*(j + nlon*r)], i = colatitude (north first), j = longitude
 *   band      ab[row + (2bw+1)*col], A[i][j] = ab[bw+i-j, j] (linear_algebra.jl:19-21)
 */
#ifndef GD_ORACLE_H
#define GD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- radial grid and banded operators ---------------------------------- */
void gdo_radial_grid(int N, double rratio, double *r /* N x 7 col-major */);
int  gdo_fd_weights(int order, int N, int bw, const double *r /* N x 7 */,
                    int transpose, double *band /* (2bw+1) x N */);
void gdo_radial_laplacian(int N, int bw, const double *r, int transpose, double *lap);
int  gdo_band_factor(int N, int bw, double *ab);
void gdo_band_solve(int N, int bw, const double *lu, double *x);
void gdo_band_apply(int N, int bw, const double *ab, const double *v, double *out);

/* CNAB2 per-l operators for l = 0..lmax (timestep.jl:498-543).
 * Lmat, Smat, LU: (lmax+1) blocks of (2bw+1)*N. */
int  gdo_cnab2_matrices(int N, int bw, const double *r, double nu, double dt,
                        double theta, int lmax, int fd_transpose, int bc_rows,
                        double *Lmat, double *Smat, double *LU);
void gdo_cnab2_rhs(int nlm, int N, int bw, const int *l_of_lm, const double *Lmat,
                   double dt, double theta, int bc_rows,
                   const double *u_re, const double *u_im,
                   const double *nl_re, const double *nl_im,
                   const double *np_re, const double *np_im,
                   double *rhs_re, double *rhs_im);
void gdo_cnab2_solve(int nlm, int N, int bw, const int *l_of_lm, const double *LU,
                     const double *rhs_re, const double *rhs_im,
                     double *sol_re, double *sol_im);

/* ---- spherical-harmonic transform ------------------------------------- */
int  gdo_nlm(int lmax, int mmax);
void gdo_lm_tables(int lmax, int mmax, int *l_of_lm, int *m_of_lm);
void gdo_gauss(int nlat, double *x /* cos(theta), north first */, double *w);
/* Orthonormal P_lm (Condon-Shortley phase) and dP_lm/dtheta at the nodes,
 * P[lm*nlat + i]. */
void gdo_plm(int lmax, int mmax, int nlat, const double *x, double *P, double *dP);

typedef struct gdo_sht gdo_sht;
gdo_sht *gdo_sht_create(int lmax, int mmax, int nlat, int nlon);
void gdo_sht_destroy(gdo_sht *);
void gdo_sht_grid(const gdo_sht *, double *x, double *w, double *theta);
void gdo_scalar_synth(const gdo_sht *, const double *re, const double *im, int nr, double *phys);
void gdo_scalar_analysis(const gdo_sht *, const double *phys, int nr, double *re, double *im);
void gdo_vector_synth(const gdo_sht *, const double *tor_re, const double *tor_im,
                      const double *pol_re, const double *pol_im, int nr,
                      double *vt, double *vp);
void gdo_vector_analysis(const gdo_sht *, const double *vt, const double *vp, int nr,
                         double *tor_re, double *tor_im, double *pol_re, double *pol_im);

/* ---- seeded state ----------------------------------------------------- */
void gdo_seeded_normals(unsigned long long seed, long n, double *out);
void gdo_seeded_state(unsigned long long seed, int lmax, int mmax, int nr, int nfields,
                      double amp, double power, int zero_boundaries,
                      double *re, double *im /* nfields*nlm*nr each */);

/* ---- the step -------------------------------------------------------- */
typedef struct {
    double E, Pm, Pr, Ra, rratio, dt, theta;
    int lmax, mmax, nlat, nlon, nr, bw;
    int fd_transpose;    /* 1: V^T w = k! e_k (well-posed); 0: literal V \ rhs */
    int radials;         /* 1: synthesise X_r = l(l+1) P_X / r^2; 0: radials stay 0 (reference data flow) */
    int coriolis_index;  /* 0: reference indexing theta_{min(j,nlat)}; 1: theta_i */
    int bc_rows;         /* Dirichlet rows (value 0) in S_l and the rhs: 1 all fields, 2 temperature
                            only, 0 none (the reference operator) */
    int radial_force;    /* 1: the buoyancy (Pm/E)(Pm/Pr) Ra r T (velocity.jl:534-537) enters the
                            poloidal equation as the poloidal scalar r^2 F_r / (l(l+1)) whose
                            radial component is F_r (magnetic.jl:257-259); 0: it reaches no
                            equation (the reference's data flow) */
} gdo_params;

typedef struct gdo_model gdo_model;
/* field order everywhere: 0 u_tor, 1 u_pol, 2 B_tor, 3 B_pol, 4 T */
gdo_model *gdo_model_create(const gdo_params *p);
void gdo_model_destroy(gdo_model *);
const double *gdo_model_r(const gdo_model *);
/* state/nl/nl_prev: 5 fields x nlm x nr, re and im planes */
void gdo_nonlinear(gdo_model *, const double *st_re, const double *st_im,
                   double *nl_re, double *nl_im);
void gdo_step(gdo_model *, int step_index, double *st_re, double *st_im,
              double *nl_re, double *nl_im, double *np_re, double *np_im);
void gdo_energies(const gdo_model *, const double *st_re, const double *st_im, double out[3]);
double gdo_timestep_error(long n, const double *a_re, const double *a_im,
                          const double *b_re, const double *b_im);
/* diagnostics: physical values of the 15 synthesised fields and 5 products
 * at one radial level (for product/unit tests); order documented in .c */
void gdo_physical_level(gdo_model *, const double *st_re, const double *st_im, int r,
                        double *fields15, double *prods5);
/* OpenMP thread count of the parallel loops (n > 0 sets it); returns it */
int gdo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
