/*
 * gd_oracle.c -- FP64 CPU restatement of GeoDynamo.jl's pseudo-spectral CNAB2
 * step.  TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline);
 * the HIP product never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference is non-executable or ill-posed, the closest well-posed reading is
 * taken behind a flag (gdo_params): fd_transpose, radials, coriolis_index.
 *
 * Parity pins (tests/test_oracle_*.py): Gauss nodes vs numpy leggauss; P_lm vs
 * scipy sph_harm_y; dP_lm vs the derivative identity and finite differences;
 * SHT round trips and single-mode deltas; vector closed forms; FD weights
 * exact on polynomials; band LU vs dense numpy solve; CNAB2 eigen-mode closed
 * form; full step vs an independent numpy restatement.  The exact phase and
 * normalisation of SHTnsKit.jl 1.1.10 (absent here) is parity unpinned.
 */
#define _GNU_SOURCE
#include "gd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BAND(ab, bw, row, col) ((ab)[(row) + (2 * (bw) + 1) * (col)])

/* OpenMP threads of the parallel loops below (the CPU baseline is timed on
 * all host cores and on one); returns the count in effect */
int gdo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* radial grid: fields.jl:108-131                                          */
/* ---------------------------------------------------------------------- */
void gdo_radial_grid(int N, double rratio, double *r)
{
    double ri = rratio / (1.0 - rratio);
    for (int n = 0; n < N; ++n) {
        /* r[n,4] = 0.5(1+cos(pi(N-n)/(N-1))) + ri, n 1-based */
        double v = 0.5 * (1.0 + cos(M_PI * (double)(N - (n + 1)) / (double)(N - 1))) + ri;
        r[n + N * 3] = v;
    }
    for (int p = 0; p < 7; ++p) {
        if (p == 3) continue;
        for (int n = 0; n < N; ++n) r[n + N * p] = pow(r[n + N * 3], (double)(p - 3));
    }
}

/* small dense solve with partial pivoting, A row-major s x s */
static int dense_solve(int s, double *A, double *b)
{
    for (int k = 0; k < s; ++k) {
        int p = k;
        for (int i = k + 1; i < s; ++i)
            if (fabs(A[i * s + k]) > fabs(A[p * s + k])) p = i;
        if (A[p * s + k] == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < s; ++j) { double t = A[k * s + j]; A[k * s + j] = A[p * s + j]; A[p * s + j] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < s; ++i) {
            double f = A[i * s + k] / A[k * s + k];
            for (int j = k; j < s; ++j) A[i * s + j] -= f * A[k * s + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = s - 1; i >= 0; --i) {
        double acc = b[i];
        for (int j = i + 1; j < s; ++j) acc -= A[i * s + j] * b[j];
        b[i] = acc / A[i * s + i];
    }
    return 0;
}

/* FD derivative weights on the local stencil: linear_algebra.jl:90-132.
 * transpose=1 solves V^T w = k! e_k (the well-posed reading); transpose=0 is
 * the literal V \ rhs. */
int gdo_fd_weights(int order, int N, int bw, const double *r, int transpose, double *band)
{
    const double *r4 = r + N * 3;
    memset(band, 0, sizeof(double) * (size_t)(2 * bw + 1) * N);
    double V[64], VT[64], rhs[8];
    for (int n = 0; n < N; ++n) {
        int left = n - bw < 0 ? 0 : n - bw;
        int right = n + bw > N - 1 ? N - 1 : n + bw;
        int s = right - left + 1;
        if (s > 8) return -1;
        for (int i = 0; i < s; ++i) {
            V[i * s + 0] = 1.0;
            for (int j = 1; j < s; ++j) V[i * s + j] = V[i * s + j - 1] * (r4[left + i] - r4[n]);
        }
        for (int i = 0; i < s; ++i) rhs[i] = 0.0;
        if (order <= s - 1) {
            double f = 1.0;
            for (int k = 2; k <= order; ++k) f *= k;
            rhs[order] = f;
        }
        if (transpose) {
            for (int i = 0; i < s; ++i)
                for (int j = 0; j < s; ++j) VT[i * s + j] = V[j * s + i];
            if (dense_solve(s, VT, rhs)) return -2;
        } else {
            if (dense_solve(s, V, rhs)) return -2;
        }
        for (int i = 0; i < s; ++i) {
            int idx = left + i;
            int row = bw + n - idx;
            if (row >= 0 && row <= 2 * bw) BAND(band, bw, row, idx) = rhs[i];
        }
    }
    return 0;
}

/* Lap = D2 + diag(2/r) D1: linear_algebra.jl:134-151 */
void gdo_radial_laplacian(int N, int bw, const double *r, int transpose, double *lap)
{
    double *d1 = (double *)malloc(sizeof(double) * (2 * bw + 1) * N);
    gdo_fd_weights(2, N, bw, r, transpose, lap);
    gdo_fd_weights(1, N, bw, r, transpose, d1);
    for (int n = 0; n < N; ++n) {
        double rinv = r[n + N * 2];
        int j0 = n - bw < 0 ? 0 : n - bw, j1 = n + bw > N - 1 ? N - 1 : n + bw;
        for (int j = j0; j <= j1; ++j) {
            int row = bw + n - j;
            BAND(lap, bw, row, j) += 2.0 * rinv * BAND(d1, bw, row, j);
        }
    }
    free(d1);
}

/* no-pivot band LU: linear_algebra.jl:23-57 */
int gdo_band_factor(int N, int bw, double *lu)
{
    for (int k = 0; k < N - 1; ++k) {
        double piv = BAND(lu, bw, bw, k);
        int imax = k + bw < N - 1 ? k + bw : N - 1;
        for (int i = k + 1; i <= imax; ++i) {
            int row = bw + i - k;
            double L = BAND(lu, bw, row, k) / piv;
            BAND(lu, bw, row, k) = L;
            for (int j = k + 1; j <= imax; ++j) {
                int col = bw + i - j;
                int urow = bw + k - j;
                BAND(lu, bw, col, j) -= L * BAND(lu, bw, urow, j);
            }
        }
    }
    for (int k = 0; k < N; ++k)
        if (BAND(lu, bw, bw, k) == 0.0 || !isfinite(BAND(lu, bw, bw, k))) return -1;
    return 0;
}

/* unit-L forward, U back substitution, in place: linear_algebra.jl:59-88 */
void gdo_band_solve(int N, int bw, const double *lu, double *x)
{
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        int j0 = i - bw < 0 ? 0 : i - bw;
        for (int j = j0; j < i; ++j) s += BAND(lu, bw, bw + i - j, j) * x[j];
        x[i] = x[i] - s;
    }
    for (int i = N - 1; i >= 0; --i) {
        double s = 0.0;
        int j1 = i + bw > N - 1 ? N - 1 : i + bw;
        for (int j = i + 1; j <= j1; ++j) s += BAND(lu, bw, bw + i - j, j) * x[j];
        x[i] = (x[i] - s) / BAND(lu, bw, bw, i);
    }
}

/* y = B v, column sweep: timestep.jl:89-101 */
void gdo_band_apply(int N, int bw, const double *ab, const double *v, double *out)
{
    for (int i = 0; i < N; ++i) out[i] = 0.0;
    for (int j = 0; j < N; ++j) {
        int i0 = j - bw < 0 ? 0 : j - bw, i1 = j + bw > N - 1 ? N - 1 : j + bw;
        for (int i = i0; i <= i1; ++i) out[i] += BAND(ab, bw, bw + i - j, j) * v[j];
    }
}

/* timestep.jl:498-543 */
int gdo_cnab2_matrices(int N, int bw, const double *r, double nu, double dt, double theta,
                       int lmax, int fd_transpose, int bc_rows, double *Lmat, double *Smat, double *LU)
{
    int bs = (2 * bw + 1) * N;
    double *lap = (double *)malloc(sizeof(double) * bs);
    gdo_radial_laplacian(N, bw, r, fd_transpose, lap);
    int rc = 0;
    for (int l = 0; l <= lmax; ++l) {
        double *Ll = Lmat + (size_t)l * bs, *Sl = Smat + (size_t)l * bs, *LUl = LU + (size_t)l * bs;
        for (int q = 0; q < bs; ++q) Ll[q] = nu * lap[q];
        double lf = (double)l * (double)(l + 1);
        for (int n = 0; n < N; ++n) BAND(Ll, bw, bw, n) -= nu * lf * r[n + N * 1];
        for (int q = 0; q < bs; ++q) Sl[q] = Ll[q] * (-theta);
        for (int n = 0; n < N; ++n) BAND(Sl, bw, bw, n) += 1.0 / dt;
        if (bc_rows) {
            /* Dirichlet rows: S[0][j] = S[N-1][j] = delta (boundary values 0) */
            for (int j = 0; j <= bw && j < N; ++j) BAND(Sl, bw, bw + 0 - j, j) = (j == 0) ? 1.0 : 0.0;
            for (int j = N - 1 - bw; j < N; ++j)
                if (j >= 0) BAND(Sl, bw, bw + (N - 1) - j, j) = (j == N - 1) ? 1.0 : 0.0;
        }
        memcpy(LUl, Sl, sizeof(double) * bs);
        if (gdo_band_factor(N, bw, LUl)) rc = -1;
    }
    free(lap);
    return rc;
}

/* rhs = u/dt + 1.5 nl - 0.5 nl_prev + (1-theta) L_l u: timestep.jl:610-690 */
void gdo_cnab2_rhs(int nlm, int N, int bw, const int *l_of_lm, const double *Lmat,
                   double dt, double theta, int bc_rows,
                   const double *u_re, const double *u_im,
                   const double *nl_re, const double *nl_im,
                   const double *np_re, const double *np_im,
                   double *rhs_re, double *rhs_im)
{
    int bs = (2 * bw + 1) * N;
    double inv_dt = 1.0 / dt, lw = 1.0 - theta;
#pragma omp parallel
    {
        double *ur = (double *)malloc(sizeof(double) * N * 4);
        double *ui = ur + N, *lr = ur + 2 * N, *li = ur + 3 * N;
#pragma omp for schedule(static)
        for (int lm = 0; lm < nlm; ++lm) {
            const double *Ll = Lmat + (size_t)l_of_lm[lm] * bs;
            for (int r = 0; r < N; ++r) { ur[r] = u_re[lm + (size_t)nlm * r]; ui[r] = u_im[lm + (size_t)nlm * r]; }
            if (lw != 0.0) { gdo_band_apply(N, bw, Ll, ur, lr); gdo_band_apply(N, bw, Ll, ui, li); }
            for (int r = 0; r < N; ++r) {
                size_t o = lm + (size_t)nlm * r;
                double vr = inv_dt * u_re[o] + 1.5 * nl_re[o] - 0.5 * np_re[o];
                double vi = inv_dt * u_im[o] + 1.5 * nl_im[o] - 0.5 * np_im[o];
                if (lw != 0.0) { vr += lw * lr[r]; vi += lw * li[r]; }
                if (bc_rows && (r == 0 || r == N - 1)) { vr = 0.0; vi = 0.0; }
                rhs_re[o] = vr; rhs_im[o] = vi;
            }
        }
        free(ur);
    }
}

/* timestep.jl:692-730 */
void gdo_cnab2_solve(int nlm, int N, int bw, const int *l_of_lm, const double *LU,
                     const double *rhs_re, const double *rhs_im, double *sol_re, double *sol_im)
{
    int bs = (2 * bw + 1) * N;
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * N * 2), *ti = tr + N;
#pragma omp for schedule(static)
        for (int lm = 0; lm < nlm; ++lm) {
            const double *LUl = LU + (size_t)l_of_lm[lm] * bs;
            for (int r = 0; r < N; ++r) { tr[r] = rhs_re[lm + (size_t)nlm * r]; ti[r] = rhs_im[lm + (size_t)nlm * r]; }
            gdo_band_solve(N, bw, LUl, tr);
            gdo_band_solve(N, bw, LUl, ti);
            for (int r = 0; r < N; ++r) { sol_re[lm + (size_t)nlm * r] = tr[r]; sol_im[lm + (size_t)nlm * r] = ti[r]; }
        }
        free(tr);
    }
}

/* ---------------------------------------------------------------------- */
/* spherical harmonics                                                     */
/* ---------------------------------------------------------------------- */
int gdo_nlm(int lmax, int mmax)
{
    int n = 0;
    for (int l = 0; l <= lmax; ++l) n += (l < mmax ? l : mmax) + 1;
    return n;
}

/* l-major ordering: shtnskit_transforms.jl:177-188 */
void gdo_lm_tables(int lmax, int mmax, int *l_of_lm, int *m_of_lm)
{
    int idx = 0;
    for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= (l < mmax ? l : mmax); ++m) { l_of_lm[idx] = l; m_of_lm[idx] = m; ++idx; }
}

/* Gauss-Legendre nodes by Newton iteration on P_n; north (x -> +1) first. */
void gdo_gauss(int nlat, double *x, double *w)
{
    for (int i = 0; i < (nlat + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (nlat + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= nlat; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            if (nlat == 1) { p1 = z; p0 = 1.0; }
            pp = nlat * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) {
                /* refresh derivative at the converged node */
                p0 = 1.0; p1 = z;
                for (int k = 2; k <= nlat; ++k) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                pp = nlat * (z * p1 - p0) / (z * z - 1.0);
                break;
            }
        }
        x[i] = z;
        x[nlat - 1 - i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[nlat - 1 - i] = w[i];
    }
    if (nlat % 2 == 1) x[nlat / 2] = 0.0;
}

/* Orthonormal P_lm with Condon-Shortley phase by the standard three-term
 * recurrence in l, with a base-2 exponent carried alongside the mantissa so
 * sin^m(theta) cannot underflow before the recurrence grows it back.
 * dP/dtheta = (l x P_l - sqrt((2l+1)(l^2-m^2)/(2l-1)) P_{l-1}) / sin(theta). */
void gdo_plm(int lmax, int mmax, int nlat, const double *x, double *P, double *dP)
{
    int *loff = (int *)malloc(sizeof(int) * (lmax + 2));
    loff[0] = 0;
    for (int l = 0; l <= lmax; ++l) loff[l + 1] = loff[l] + (l < mmax ? l : mmax) + 1;
    for (int i = 0; i < nlat; ++i) {
        double xi = x[i], s = sqrt(1.0 - xi * xi);
        double lgs = log(s);
        double lgacc = 0.5 * log(1.0 / (4.0 * M_PI)); /* log of sqrt(prod)/sqrt(4pi) w/o (2m+1) */
        for (int m = 0; m <= mmax; ++m) {
            if (m > 0) lgacc += 0.5 * log((2.0 * m - 1.0) / (2.0 * m));
            double lg = lgacc + 0.5 * log(2.0 * m + 1.0) + m * lgs;
            int e = (int)floor(lg / M_LN2);
            double val = exp(lg - e * M_LN2);
            if (m & 1) val = -val;
            double pm2 = 0.0, pm1 = val; /* mantissas, common exponent e */
            double tprev = 0.0;          /* true P_{l-1} */
            for (int l = m; l <= lmax; ++l) {
                double cur;
                if (l == m) cur = pm1;
                else if (l == m + 1) { cur = sqrt(2.0 * m + 3.0) * xi * pm1; pm2 = pm1; pm1 = cur; }
                else {
                    double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                    double b = sqrt((((double)l - 1) * ((double)l - 1) - (double)m * m) / (4.0 * ((double)l - 1) * ((double)l - 1) - 1.0));
                    cur = a * (xi * pm1 - b * pm2);
                    pm2 = pm1; pm1 = cur;
                }
                if (fabs(pm1) > 0x1p600) { pm1 = ldexp(pm1, -600); pm2 = ldexp(pm2, -600); e += 600; }
                double t = ldexp(pm1, e);
                int lm = loff[l] + m; /* l-major index */
                P[(size_t)lm * nlat + i] = t;
                double c = (l == m) ? 0.0 : sqrt((2.0 * l + 1.0) * ((double)l * l - (double)m * m) / (2.0 * l - 1.0));
                dP[(size_t)lm * nlat + i] = (l * xi * t - c * tprev) / s;
                tprev = t;
            }
        }
    }
    free(loff);
}

/* ---- FFT: recursive mixed-radix complex DFT (sign -1 forward, +1 inverse,
 * unnormalised) ------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static void fft_rec(const cplx *in, int stride, cplx *out, int n, const cplx *root, int rstride, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = 0;
    for (int f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
    if (!p) p = n;
    int m = n / p;
    for (int q = 0; q < p; ++q) fft_rec(in + (size_t)q * stride, stride * p, out + (size_t)q * m, m, root, rstride * p, tmp + n);
    /* combine: X[k + s m] = sum_q W_n^{q(k+sm)} Y_q[k] */
    for (int k = 0; k < m; ++k) {
        for (int s = 0; s < p; ++s) {
            int kk = k + s * m;
            double ar = 0.0, ai = 0.0;
            for (int q = 0; q < p; ++q) {
                long e = ((long)q * kk) % n;
                cplx w = root[e * rstride];
                cplx y = out[q * m + k];
                ar += w.re * y.re - w.im * y.im;
                ai += w.re * y.im + w.im * y.re;
            }
            tmp[kk].re = ar; tmp[kk].im = ai;
        }
    }
    memcpy(out, tmp, sizeof(cplx) * n);
}

typedef struct { int n; cplx *fwd, *inv; } fftplan;

static void fftplan_init(fftplan *pl, int n)
{
    pl->n = n;
    pl->fwd = (cplx *)malloc(sizeof(cplx) * n);
    pl->inv = (cplx *)malloc(sizeof(cplx) * n);
    for (int k = 0; k < n; ++k) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        pl->fwd[k].re = cos(a); pl->fwd[k].im = -sin(a);
        pl->inv[k].re = cos(a); pl->inv[k].im = sin(a);
    }
}

static void fft_exec(const fftplan *pl, const cplx *in, cplx *out, int inverse, cplx *work /* >= 2n*log */)
{
    fft_rec(in, 1, out, pl->n, inverse ? pl->inv : pl->fwd, 1, work);
}

struct gdo_sht {
    int lmax, mmax, nlat, nlon, nlm;
    double *x, *w, *theta, *P, *dP;
    int *l_of_lm, *m_of_lm, *lm0; /* lm0[m] unused; kept for clarity */
    int *lmidx;                   /* lmidx[l*(mmax+1)+m] */
    fftplan fp;
};

gdo_sht *gdo_sht_create(int lmax, int mmax, int nlat, int nlon)
{
    gdo_sht *c = (gdo_sht *)calloc(1, sizeof(gdo_sht));
    c->lmax = lmax; c->mmax = mmax; c->nlat = nlat; c->nlon = nlon;
    c->nlm = gdo_nlm(lmax, mmax);
    c->x = (double *)malloc(sizeof(double) * nlat);
    c->w = (double *)malloc(sizeof(double) * nlat);
    c->theta = (double *)malloc(sizeof(double) * nlat);
    gdo_gauss(nlat, c->x, c->w);
    for (int i = 0; i < nlat; ++i) c->theta[i] = acos(c->x[i]);
    c->P = (double *)malloc(sizeof(double) * (size_t)c->nlm * nlat);
    c->dP = (double *)malloc(sizeof(double) * (size_t)c->nlm * nlat);
    gdo_plm(lmax, mmax, nlat, c->x, c->P, c->dP);
    c->l_of_lm = (int *)malloc(sizeof(int) * c->nlm);
    c->m_of_lm = (int *)malloc(sizeof(int) * c->nlm);
    gdo_lm_tables(lmax, mmax, c->l_of_lm, c->m_of_lm);
    c->lmidx = (int *)malloc(sizeof(int) * (lmax + 1) * (mmax + 1));
    for (int q = 0; q < c->nlm; ++q) c->lmidx[c->l_of_lm[q] * (mmax + 1) + c->m_of_lm[q]] = q;
    fftplan_init(&c->fp, nlon);
    return c;
}

void gdo_sht_destroy(gdo_sht *c)
{
    if (!c) return;
    free(c->x); free(c->w); free(c->theta); free(c->P); free(c->dP);
    free(c->l_of_lm); free(c->m_of_lm); free(c->lmidx);
    free(c->fp.fwd); free(c->fp.inv);
    free(c);
}

void gdo_sht_grid(const gdo_sht *c, double *x, double *w, double *theta)
{
    memcpy(x, c->x, sizeof(double) * c->nlat);
    memcpy(w, c->w, sizeof(double) * c->nlat);
    memcpy(theta, c->theta, sizeof(double) * c->nlat);
}

/* per-thread scratch for one radial level */
typedef struct { cplx *F, *G, *z, *Z, *work; } lvl_ws;

static lvl_ws ws_alloc(const gdo_sht *c)
{
    lvl_ws w;
    int nm = c->mmax + 1;
    w.F = (cplx *)malloc(sizeof(cplx) * (size_t)nm * c->nlat * 2);
    w.G = (cplx *)malloc(sizeof(cplx) * (size_t)nm * c->nlat * 2);
    w.z = (cplx *)malloc(sizeof(cplx) * c->nlon);
    w.Z = (cplx *)malloc(sizeof(cplx) * c->nlon);
    w.work = (cplx *)malloc(sizeof(cplx) * c->nlon * 64);
    return w;
}
static void ws_free(lvl_ws *w) { free(w->F); free(w->G); free(w->z); free(w->Z); free(w->work); }

/* ring synthesis: f_j = sum_m c_m Re[F_m e^{i m phi_j}], c_0=1, c_m=2 */
static void ring_synth(const gdo_sht *c, const cplx *Fm /* stride nlat */, int i, double *out /* stride nlat */, lvl_ws *w)
{
    int n = c->nlon;
    for (int k = 0; k < n; ++k) { w->Z[k].re = 0.0; w->Z[k].im = 0.0; }
    w->Z[0].re = Fm[i].re;
    for (int m = 1; m <= c->mmax; ++m) {
        cplx v = Fm[(size_t)m * c->nlat + i];
        w->Z[m].re += v.re; w->Z[m].im += v.im;
        w->Z[n - m].re += v.re; w->Z[n - m].im -= v.im;
    }
    fft_exec(&c->fp, w->Z, w->z, 1, w->work);
    for (int j = 0; j < n; ++j) out[(size_t)j * c->nlat] = w->z[j].re;
}

/* ring analysis: G_m = (1/nlon) sum_j f_j e^{-i m phi_j} */
static void ring_analysis(const gdo_sht *c, const double *in /* stride nlat */, int i, cplx *Gm /* stride nlat */, lvl_ws *w)
{
    int n = c->nlon;
    for (int j = 0; j < n; ++j) { w->z[j].re = in[(size_t)j * c->nlat]; w->z[j].im = 0.0; }
    fft_exec(&c->fp, w->z, w->Z, 0, w->work);
    for (int m = 0; m <= c->mmax; ++m) {
        Gm[(size_t)m * c->nlat + i].re = w->Z[m].re / n;
        Gm[(size_t)m * c->nlat + i].im = w->Z[m].im / n;
    }
}

/* one level scalar synthesis: a_lm -> f(theta, phi) */
static void level_scalar_synth(const gdo_sht *c, const double *re, const double *im, double *phys, lvl_ws *w)
{
    int nlat = c->nlat;
    for (int m = 0; m <= c->mmax; ++m)
        for (int i = 0; i < nlat; ++i) {
            double sr = 0.0, si = 0.0;
            for (int l = m; l <= c->lmax; ++l) {
                int q = c->lmidx[l * (c->mmax + 1) + m];
                double p = c->P[(size_t)q * nlat + i];
                sr += re[q] * p; si += im[q] * p;
            }
            w->F[(size_t)m * nlat + i].re = sr; w->F[(size_t)m * nlat + i].im = si;
        }
    for (int i = 0; i < nlat; ++i) ring_synth(c, w->F, i, phys + i, w);
}

/* shtnskit_field_functions.jl:432-435: imag(m=0) := 0 on every store */
static void level_scalar_analysis(const gdo_sht *c, const double *phys, double *re, double *im, lvl_ws *w)
{
    int nlat = c->nlat;
    for (int i = 0; i < nlat; ++i) ring_analysis(c, phys + i, i, w->G, w);
    for (int q = 0; q < c->nlm; ++q) {
        int l = c->l_of_lm[q], m = c->m_of_lm[q];
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < nlat; ++i) {
            double p = c->P[(size_t)q * nlat + i] * c->w[i];
            sr += p * w->G[(size_t)m * nlat + i].re;
            si += p * w->G[(size_t)m * nlat + i].im;
        }
        (void)l;
        re[q] = 2.0 * M_PI * sr;
        im[q] = (m == 0) ? 0.0 : 2.0 * M_PI * si;
    }
}

/* v_theta = sum c_m Re[(S dP + i m T P/s) e^{imphi}],
 * v_phi   = sum c_m Re[(i m S P/s - T dP) e^{imphi}]   (S := pol, T := tor) */
static void level_vector_synth(const gdo_sht *c, const double *tre, const double *tim,
                               const double *pre, const double *pim, double *vt, double *vp, lvl_ws *w)
{
    int nlat = c->nlat;
    cplx *Ft = w->F, *Fp = w->F + (size_t)(c->mmax + 1) * nlat;
    for (int m = 0; m <= c->mmax; ++m)
        for (int i = 0; i < nlat; ++i) {
            double s = sqrt(1.0 - c->x[i] * c->x[i]);
            double atr = 0, ati = 0, apr = 0, api = 0;
            for (int l = m; l <= c->lmax; ++l) {
                int q = c->lmidx[l * (c->mmax + 1) + m];
                double dp = c->dP[(size_t)q * nlat + i];
                double qm = m * c->P[(size_t)q * nlat + i] / s;
                /* theta: S dP + i Q T */
                atr += pre[q] * dp - qm * tim[q];
                ati += pim[q] * dp + qm * tre[q];
                /* phi: i Q S - T dP */
                apr += -qm * pim[q] - tre[q] * dp;
                api += qm * pre[q] - tim[q] * dp;
            }
            Ft[(size_t)m * nlat + i].re = atr; Ft[(size_t)m * nlat + i].im = ati;
            Fp[(size_t)m * nlat + i].re = apr; Fp[(size_t)m * nlat + i].im = api;
        }
    for (int i = 0; i < nlat; ++i) { ring_synth(c, Ft, i, vt + i, w); ring_synth(c, Fp, i, vp + i, w); }
}

/* S = 2pi/(l(l+1)) sum_i w_i [G_t dP - i Q G_p];  T = 2pi/(l(l+1)) sum_i w_i [-i Q G_t - dP G_p] */
static void level_vector_analysis(const gdo_sht *c, const double *vt, const double *vp,
                                  double *tre, double *tim, double *pre, double *pim, lvl_ws *w)
{
    int nlat = c->nlat;
    cplx *Gt = w->G, *Gp = w->G + (size_t)(c->mmax + 1) * nlat;
    for (int i = 0; i < nlat; ++i) { ring_analysis(c, vt + i, i, Gt, w); ring_analysis(c, vp + i, i, Gp, w); }
    for (int q = 0; q < c->nlm; ++q) {
        int l = c->l_of_lm[q], m = c->m_of_lm[q];
        if (l == 0) { tre[q] = tim[q] = pre[q] = pim[q] = 0.0; continue; }
        double Sr = 0, Si = 0, Tr = 0, Ti = 0;
        for (int i = 0; i < nlat; ++i) {
            double s = sqrt(1.0 - c->x[i] * c->x[i]);
            double dp = c->dP[(size_t)q * nlat + i] * c->w[i];
            double qm = m * c->P[(size_t)q * nlat + i] / s * c->w[i];
            cplx gt = Gt[(size_t)m * nlat + i], gp = Gp[(size_t)m * nlat + i];
            /* G_t dP - i Q G_p */
            Sr += gt.re * dp + qm * gp.im;
            Si += gt.im * dp - qm * gp.re;
            /* -i Q G_t - dP G_p */
            Tr += qm * gt.im - dp * gp.re;
            Ti += -qm * gt.re - dp * gp.im;
        }
        double f = 2.0 * M_PI / ((double)l * (l + 1));
        pre[q] = f * Sr; pim[q] = (m == 0) ? 0.0 : f * Si;
        tre[q] = f * Tr; tim[q] = (m == 0) ? 0.0 : f * Ti;
    }
}

/* shtnskit_field_functions.jl:11-21,102-123 */
void gdo_scalar_synth(const gdo_sht *c, const double *re, const double *im, int nr, double *phys)
{
#pragma omp parallel
    {
        lvl_ws w = ws_alloc(c);
#pragma omp for schedule(static)
        for (int r = 0; r < nr; ++r)
            level_scalar_synth(c, re + (size_t)r * c->nlm, im + (size_t)r * c->nlm,
                               phys + (size_t)r * c->nlat * c->nlon, &w);
        ws_free(&w);
    }
}

/* shtnskit_field_functions.jl:131-141,218-239 */
void gdo_scalar_analysis(const gdo_sht *c, const double *phys, int nr, double *re, double *im)
{
#pragma omp parallel
    {
        lvl_ws w = ws_alloc(c);
#pragma omp for schedule(static)
        for (int r = 0; r < nr; ++r)
            level_scalar_analysis(c, phys + (size_t)r * c->nlat * c->nlon,
                                  re + (size_t)r * c->nlm, im + (size_t)r * c->nlm, &w);
        ws_free(&w);
    }
}

/* shtnskit_field_functions.jl:252-282 (SHsphtor_to_spat with S := pol, T := tor) */
void gdo_vector_synth(const gdo_sht *c, const double *tre, const double *tim,
                      const double *pre, const double *pim, int nr, double *vt, double *vp)
{
#pragma omp parallel
    {
        lvl_ws w = ws_alloc(c);
#pragma omp for schedule(static)
        for (int r = 0; r < nr; ++r) {
            size_t so = (size_t)r * c->nlm, po = (size_t)r * c->nlat * c->nlon;
            level_vector_synth(c, tre + so, tim + so, pre + so, pim + so, vt + po, vp + po, &w);
        }
        ws_free(&w);
    }
}

/* shtnskit_field_functions.jl:291-321 (spat_to_SHsphtor -> (pol, tor)) */
void gdo_vector_analysis(const gdo_sht *c, const double *vt, const double *vp, int nr,
                         double *tre, double *tim, double *pre, double *pim)
{
#pragma omp parallel
    {
        lvl_ws w = ws_alloc(c);
#pragma omp for schedule(static)
        for (int r = 0; r < nr; ++r) {
            size_t so = (size_t)r * c->nlm, po = (size_t)r * c->nlat * c->nlon;
            level_vector_analysis(c, vt + po, vp + po, tre + so, tim + so, pre + so, pim + so, &w);
        }
        ws_free(&w);
    }
}

/* ---------------------------------------------------------------------- */
/* seeded state: splitmix64 -> uniform(0,1) -> Box-Muller                   */
/* ---------------------------------------------------------------------- */
static unsigned long long splitmix64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void gdo_seeded_normals(unsigned long long seed, long n, double *out)
{
    unsigned long long s = seed;
    for (long k = 0; k < n; k += 2) {
        double u1 = ((double)(splitmix64(&s) >> 11) + 0.5) * 0x1p-53;
        double u2 = ((double)(splitmix64(&s) >> 11) + 0.5) * 0x1p-53;
        double rad = sqrt(-2.0 * log(u1));
        out[k] = rad * cos(2.0 * M_PI * u2);
        if (k + 1 < n) out[k + 1] = rad * sin(2.0 * M_PI * u2);
    }
}

/* draws in order field, r, lm (l-major), (re, im); value = amp * n * (l+1)^-power;
 * m = 0 kept real; rows r = 0 and r = nr-1 zeroed when zero_boundaries. */
void gdo_seeded_state(unsigned long long seed, int lmax, int mmax, int nr, int nfields,
                      double amp, double power, int zero_boundaries, double *re, double *im)
{
    int nlm = gdo_nlm(lmax, mmax);
    int *lo = (int *)malloc(sizeof(int) * nlm), *mo = (int *)malloc(sizeof(int) * nlm);
    gdo_lm_tables(lmax, mmax, lo, mo);
    long n = 2L * nfields * nlm * nr;
    double *g = (double *)malloc(sizeof(double) * n);
    gdo_seeded_normals(seed, n, g);
    long k = 0;
    for (int f = 0; f < nfields; ++f)
        for (int r = 0; r < nr; ++r)
            for (int q = 0; q < nlm; ++q) {
                double sc = amp * pow((double)lo[q] + 1.0, -power);
                size_t o = (size_t)f * nlm * nr + (size_t)r * nlm + q;
                re[o] = g[k++] * sc;
                im[o] = g[k++] * sc;
                if (mo[q] == 0) im[o] = 0.0;
                if (zero_boundaries && (r == 0 || r == nr - 1)) { re[o] = 0.0; im[o] = 0.0; }
            }
    free(g); free(lo); free(mo);
}

/* ---------------------------------------------------------------------- */
/* the model: one pseudo-spectral CNAB2 step                               */
/* ---------------------------------------------------------------------- */
struct gdo_model {
    gdo_params p;
    gdo_sht *sht;
    int nlm, N, bw;
    double *r;            /* N x 7 */
    double *D1, *D2;      /* band */
    double *Lm[3], *Sm[3], *LUm[3]; /* nu index: 0 velocity (Pm), 1 magnetic (1), 2 temperature (Pm/Pr) */
    int *l_of_lm, *m_of_lm;
};

gdo_model *gdo_model_create(const gdo_params *p)
{
    gdo_model *M = (gdo_model *)calloc(1, sizeof(gdo_model));
    M->p = *p;
    M->sht = gdo_sht_create(p->lmax, p->mmax, p->nlat, p->nlon);
    M->nlm = M->sht->nlm; M->N = p->nr; M->bw = p->bw;
    int N = p->nr, bw = p->bw, bs = (2 * bw + 1) * N;
    M->r = (double *)malloc(sizeof(double) * N * 7);
    gdo_radial_grid(N, p->rratio, M->r);
    M->D1 = (double *)malloc(sizeof(double) * bs);
    M->D2 = (double *)malloc(sizeof(double) * bs);
    gdo_fd_weights(1, N, bw, M->r, p->fd_transpose, M->D1);
    gdo_fd_weights(2, N, bw, M->r, p->fd_transpose, M->D2);
    double nus[3] = {p->Pm, 1.0, p->Pm / p->Pr}; /* simulation.jl:229-231 */
    for (int k = 0; k < 3; ++k) {
        M->Lm[k] = (double *)malloc(sizeof(double) * bs * (p->lmax + 1));
        M->Sm[k] = (double *)malloc(sizeof(double) * bs * (p->lmax + 1));
        M->LUm[k] = (double *)malloc(sizeof(double) * bs * (p->lmax + 1));
        int bc = (p->bc_rows == 1) || (p->bc_rows == 2 && k == 2);
        gdo_cnab2_matrices(N, bw, M->r, nus[k], p->dt, p->theta, p->lmax, p->fd_transpose, bc,
                           M->Lm[k], M->Sm[k], M->LUm[k]);
    }
    M->l_of_lm = M->sht->l_of_lm; M->m_of_lm = M->sht->m_of_lm;
    return M;
}

void gdo_model_destroy(gdo_model *M)
{
    if (!M) return;
    for (int k = 0; k < 3; ++k) { free(M->Lm[k]); free(M->Sm[k]); free(M->LUm[k]); }
    free(M->r); free(M->D1); free(M->D2);
    gdo_sht_destroy(M->sht);
    free(M);
}

const double *gdo_model_r(const gdo_model *M) { return M->r; }

/* Spectral-space preparation (per lm column):
 *   omega_tor = l(l+1)/r^2 P - D2 P - (2/r) D1 P, omega_pol = -l(l+1)/r^2 T   velocity.jl:427-435
 *   j likewise from B                                                         magnetic.jl:335-343
 *   radials X_r = l(l+1) P_X / r^2                                            magnetic.jl:257-259
 *   T/r (gradient synthesis source) and dT/dr = D1 T                          scalar_field_common.jl:233-327
 * spec: 14 complex fields x (nlm x nr):
 *   0 w_tor 1 w_pol 2 j_tor 3 j_pol 4 u_r 5 w_r 6 B_r 7 j_r 8 T/r 9 dT/dr */
#define NSPEC 10
static void spectral_prep(const gdo_model *M, const double *st_re, const double *st_im,
                          double *sp_re, double *sp_im)
{
    int nlm = M->nlm, N = M->N, bw = M->bw;
    size_t fs = (size_t)nlm * N;
    const double *r = M->r;
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * N * 6);
#pragma omp for schedule(static)
        for (int q = 0; q < nlm; ++q) {
            double lf = (double)M->l_of_lm[q] * (M->l_of_lm[q] + 1);
            for (int ri = 0; ri < 2; ++ri) {
                const double *st = ri ? st_im : st_re;
                double *sp = ri ? sp_im : sp_re;
                for (int src = 0; src < 2; ++src) { /* 0: velocity, 1: magnetic */
                    const double *tor = st + (size_t)(2 * src) * fs;
                    const double *pol = st + (size_t)(2 * src + 1) * fs;
                    double *col = buf, *d1 = buf + N, *d2 = buf + 2 * N;
                    for (int n = 0; n < N; ++n) col[n] = pol[q + (size_t)nlm * n];
                    gdo_band_apply(N, bw, M->D1, col, d1);
                    gdo_band_apply(N, bw, M->D2, col, d2);
                    for (int n = 0; n < N; ++n) {
                        size_t o = q + (size_t)nlm * n;
                        double rinv = r[n + N * 2], rinv2 = r[n + N * 1];
                        double ctor = lf * rinv2 * col[n] - d2[n] - 2.0 * rinv * d1[n];
                        double cpol = -lf * rinv2 * tor[o];
                        sp[(size_t)(2 * src) * fs + o] = ctor;
                        sp[(size_t)(2 * src + 1) * fs + o] = cpol;
                        if (M->p.radials) {
                            /* u_r / B_r from the field's poloidal scalar, w_r / j_r from the curl's */
                            sp[(size_t)(4 + 2 * src) * fs + o] = lf * col[n] * rinv2;
                            sp[(size_t)(5 + 2 * src) * fs + o] = lf * cpol * rinv2;
                        } else {
                            sp[(size_t)(4 + 2 * src) * fs + o] = 0.0;
                            sp[(size_t)(5 + 2 * src) * fs + o] = 0.0;
                        }
                    }
                }
                /* temperature */
                const double *T = st + (size_t)4 * fs;
                double *col = buf, *d1 = buf + N;
                for (int n = 0; n < N; ++n) col[n] = T[q + (size_t)nlm * n];
                gdo_band_apply(N, bw, M->D1, col, d1);
                for (int n = 0; n < N; ++n) {
                    size_t o = q + (size_t)nlm * n;
                    sp[(size_t)8 * fs + o] = col[n] * r[n + N * 2];
                    sp[(size_t)9 * fs + o] = d1[n];
                }
            }
        }
        free(buf);
    }
}

/* fields15 order: u_t u_p w_t w_p B_t B_p j_t j_p (theta/phi pairs), gT_t gT_p,
 *                 u_r w_r B_r j_r, dT_r.  prods5: adv_t adv_p uxB_t uxB_p advT */
static void level_physical(const gdo_model *M, const double *st_re, const double *st_im,
                           const double *sp_re, const double *sp_im, int r, double *F, double *Pr, lvl_ws *w)
{
    const gdo_sht *c = M->sht;
    int nlm = M->nlm, N = M->N, ng = c->nlat * c->nlon;
    size_t fs = (size_t)nlm * N, so = (size_t)r * nlm;
    static const double zero_dummy = 0.0;
    (void)zero_dummy;
    /* vector syntheses: (S := pol, T := tor) */
    level_vector_synth(c, st_re + 0 * fs + so, st_im + 0 * fs + so, st_re + 1 * fs + so, st_im + 1 * fs + so, F + 0 * ng, F + 1 * ng, w);
    level_vector_synth(c, sp_re + 0 * fs + so, sp_im + 0 * fs + so, sp_re + 1 * fs + so, sp_im + 1 * fs + so, F + 2 * ng, F + 3 * ng, w);
    level_vector_synth(c, st_re + 2 * fs + so, st_im + 2 * fs + so, st_re + 3 * fs + so, st_im + 3 * fs + so, F + 4 * ng, F + 5 * ng, w);
    level_vector_synth(c, sp_re + 2 * fs + so, sp_im + 2 * fs + so, sp_re + 3 * fs + so, sp_im + 3 * fs + so, F + 6 * ng, F + 7 * ng, w);
    /* gradient synthesis: S = T/r, T = 0 */
    double *zeros = (double *)calloc(nlm, sizeof(double));
    level_vector_synth(c, zeros, zeros, sp_re + 8 * fs + so, sp_im + 8 * fs + so, F + 8 * ng, F + 9 * ng, w);
    free(zeros);
    for (int k = 0; k < 4; ++k)
        level_scalar_synth(c, sp_re + (4 + k) * fs + so, sp_im + (4 + k) * fs + so, F + (10 + k) * ng, w);
    level_scalar_synth(c, sp_re + 9 * fs + so, sp_im + 9 * fs + so, F + 14 * ng, w);

    const double ro = M->p.Pm / M->p.E; /* velocity.jl:455 rossby_factor, :645 lorentz_factor */
    int nlat = c->nlat, nlon = c->nlon;
    for (int j = 0; j < nlon; ++j) {
        for (int i = 0; i < nlat; ++i) {
            int ti = M->p.coriolis_index ? i : (j < nlat - 1 ? j : nlat - 1); /* velocity.jl:494 */
            double st = sin(c->theta[ti]), ct = cos(c->theta[ti]);
            int g = i + nlat * j;
            double ut = F[0 * ng + g], up = F[1 * ng + g], ur = F[10 * ng + g];
            double wt = F[2 * ng + g], wp = F[3 * ng + g], wr = F[11 * ng + g];
            double Bt = F[4 * ng + g], Bp = F[5 * ng + g], Br = F[12 * ng + g];
            double jt = F[6 * ng + g], jp = F[7 * ng + g], jr = F[13 * ng + g];
            double gt = F[8 * ng + g], gp = F[9 * ng + g], gr = F[14 * ng + g];
            double adv_coeff = -ro;
            /* velocity.jl:512-527 */
            double at = adv_coeff * (up * wr - ur * wp);
            double ap = adv_coeff * (ur * wt - ut * wr);
            double cor_t = -ro * (-ct * up);
            double cor_p = -ro * (ct * ut + st * ur);
            at = at + cor_t;
            ap = ap + cor_p;
            /* velocity.jl:664-670 Lorentz */
            at += ro * (jp * Br - jr * Bp);
            ap += ro * (jr * Bt - jt * Br);
            Pr[0 * ng + g] = at;
            Pr[1 * ng + g] = ap;
            /* magnetic.jl:399-401 */
            Pr[2 * ng + g] = up * Br - ur * Bp;
            Pr[3 * ng + g] = ur * Bt - ut * Br;
            /* scalar_field_common.jl:401-407 */
            Pr[4 * ng + g] = -(ur * gr + ut * gt + up * gp);
        }
    }
    (void)N;
}

void gdo_physical_level(gdo_model *M, const double *st_re, const double *st_im, int r,
                        double *fields15, double *prods5)
{
    size_t fs = (size_t)M->nlm * M->N;
    double *sp_re = (double *)malloc(sizeof(double) * fs * NSPEC * 2), *sp_im = sp_re + fs * NSPEC;
    spectral_prep(M, st_re, st_im, sp_re, sp_im);
    lvl_ws w = ws_alloc(M->sht);
    level_physical(M, st_re, st_im, sp_re, sp_im, r, fields15, prods5, &w);
    ws_free(&w);
    free(sp_re);
}

/* nl order: 0 u_tor 1 u_pol 2 B_tor 3 B_pol 4 T.
 * velocity.jl:331 (adv -> nl_tor/nl_pol), magnetic.jl:437-440 (curl of u x B),
 * thermal.jl:164-180 (-u.grad T analysis) */
void gdo_nonlinear(gdo_model *M, const double *st_re, const double *st_im, double *nl_re, double *nl_im)
{
    const gdo_sht *c = M->sht;
    int nlm = M->nlm, N = M->N, ng = c->nlat * c->nlon;
    size_t fs = (size_t)nlm * N;
    double *sp_re = (double *)malloc(sizeof(double) * fs * NSPEC * 2), *sp_im = sp_re + fs * NSPEC;
    spectral_prep(M, st_re, st_im, sp_re, sp_im);
#pragma omp parallel
    {
        lvl_ws w = ws_alloc(c);
        double *F = (double *)malloc(sizeof(double) * (size_t)ng * 20), *Pr = F + (size_t)ng * 15;
        double *a = (double *)malloc(sizeof(double) * nlm * 4);
        double *atr = a, *ati = a + nlm, *apr = a + 2 * nlm, *api = a + 3 * nlm;
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < N; ++r) {
            size_t so = (size_t)r * nlm;
            level_physical(M, st_re, st_im, sp_re, sp_im, r, F, Pr, &w);
            /* velocity: (pol, tor) = vector analysis of adv */
            level_vector_analysis(c, Pr + 0 * ng, Pr + 1 * ng, atr, ati, apr, api, &w);
            memcpy(nl_re + 0 * fs + so, atr, sizeof(double) * nlm);
            memcpy(nl_im + 0 * fs + so, ati, sizeof(double) * nlm);
            memcpy(nl_re + 1 * fs + so, apr, sizeof(double) * nlm);
            memcpy(nl_im + 1 * fs + so, api, sizeof(double) * nlm);
            /* magnetic: nl_tor = l(l+1) S(uxB), nl_pol = -l(l+1) T(uxB) */
            level_vector_analysis(c, Pr + 2 * ng, Pr + 3 * ng, atr, ati, apr, api, &w);
            for (int q = 0; q < nlm; ++q) {
                double lf = (double)M->l_of_lm[q] * (M->l_of_lm[q] + 1);
                nl_re[2 * fs + so + q] = lf * apr[q];
                nl_im[2 * fs + so + q] = lf * api[q];
                nl_re[3 * fs + so + q] = -lf * atr[q];
                nl_im[3 * fs + so + q] = -lf * ati[q];
            }
            /* temperature */
            level_scalar_analysis(c, Pr + 4 * ng, nl_re + 4 * fs + so, nl_im + 4 * fs + so, &w);
            /* radial forcing: the buoyancy F_r = (Pm/E)(Pm/Pr) Ra r T as the poloidal
             * scalar r^2 F_r / (l(l+1)) (u_r = l(l+1) P / r^2, magnetic.jl:257-259) */
            if (M->p.radial_force) {
                const double rr = M->r[(size_t)3 * N + r];
                const double cb = (M->p.Pm / M->p.E) * (M->p.Pm / M->p.Pr) * M->p.Ra;
                for (int q = 0; q < nlm; ++q) {
                    const int l = M->l_of_lm[q];
                    if (l == 0) continue;
                    const double f = cb * rr * rr * rr / ((double)l * (l + 1));
                    nl_re[1 * fs + so + q] += f * st_re[4 * fs + so + q];
                    nl_im[1 * fs + so + q] += f * st_im[4 * fs + so + q];
                }
            }
        }
        free(a); free(F);
        ws_free(&w);
    }
    free(sp_re);
}

/* One CNAB2 step: nonlinear terms from the current state, temperature
 * Dirichlet rows (thermal.jl:262-313), AB1 bootstrap at step 0
 * (simulation.jl:1153-1170), rhs + banded solve per field
 * (simulation.jl:1183-1265), history roll (simulation.jl:1312-1329). */
void gdo_step(gdo_model *M, int step_index, double *st_re, double *st_im,
              double *nl_re, double *nl_im, double *np_re, double *np_im)
{
    int nlm = M->nlm, N = M->N, bw = M->bw;
    size_t fs = (size_t)nlm * N;
    gdo_nonlinear(M, st_re, st_im, nl_re, nl_im);
    /* temperature boundary rows := boundary values (0), imag := 0 */
    for (int q = 0; q < nlm; ++q) {
        st_re[4 * fs + q] = 0.0; st_im[4 * fs + q] = 0.0;
        st_re[4 * fs + (size_t)(N - 1) * nlm + q] = 0.0; st_im[4 * fs + (size_t)(N - 1) * nlm + q] = 0.0;
    }
    if (step_index == 0) { memcpy(np_re, nl_re, sizeof(double) * fs * 5); memcpy(np_im, nl_im, sizeof(double) * fs * 5); }
    double *rr = (double *)malloc(sizeof(double) * fs * 2), *ri = rr + fs;
    static const int nu_of_field[5] = {0, 0, 1, 1, 2};
    for (int f = 0; f < 5; ++f) {
        int k = nu_of_field[f];
        size_t o = (size_t)f * fs;
        int bc = (M->p.bc_rows == 1) || (M->p.bc_rows == 2 && k == 2);
        gdo_cnab2_rhs(nlm, N, bw, M->l_of_lm, M->Lm[k], M->p.dt, M->p.theta, bc,
                      st_re + o, st_im + o, nl_re + o, nl_im + o, np_re + o, np_im + o, rr, ri);
        gdo_cnab2_solve(nlm, N, bw, M->l_of_lm, M->LUm[k], rr, ri, st_re + o, st_im + o);
    }
    free(rr);
    memcpy(np_re, nl_re, sizeof(double) * fs * 5);
    memcpy(np_im, nl_im, sizeof(double) * fs * 5);
}

/* E = 1/2 sum (T^2 + P^2)/max(l(l+1),1): magnetic.jl:504-543 (used for KE too) */
void gdo_energies(const gdo_model *M, const double *st_re, const double *st_im, double out[3])
{
    int nlm = M->nlm, N = M->N;
    size_t fs = (size_t)nlm * N;
    for (int k = 0; k < 3; ++k) out[k] = 0.0;
    for (int pair = 0; pair < 2; ++pair) {
        double acc = 0.0;
        for (int r = 0; r < N; ++r)
            for (int q = 0; q < nlm; ++q) {
                double lf = (double)M->l_of_lm[q] * (M->l_of_lm[q] + 1);
                double wgt = 1.0 / (lf > 1.0 ? lf : 1.0);
                size_t o = (size_t)r * nlm + q;
                const double *tr = st_re + (size_t)(2 * pair) * fs, *ti = st_im + (size_t)(2 * pair) * fs;
                const double *pr = st_re + (size_t)(2 * pair + 1) * fs, *pi = st_im + (size_t)(2 * pair + 1) * fs;
                acc += wgt * (tr[o] * tr[o] + ti[o] * ti[o] + pr[o] * pr[o] + pi[o] * pi[o]);
            }
        out[pair] = 0.5 * acc;
    }
    /* thermal.jl:413-438: 1/2 sum |T|^2 */
    double acc = 0.0;
    for (size_t o = 0; o < fs; ++o) {
        double a = st_re[4 * fs + o], b = st_im[4 * fs + o];
        acc += a * a + b * b;
    }
    out[2] = 0.5 * acc;
}

/* timestep.jl:733-753 */
double gdo_timestep_error(long n, const double *a_re, const double *a_im, const double *b_re, const double *b_im)
{
    double e = 0.0;
    for (long k = 0; k < n; ++k) {
        double dr = a_re[k] - b_re[k], di = a_im[k] - b_im[k];
        e += dr * dr + di * di;
    }
    return sqrt(e);
}
