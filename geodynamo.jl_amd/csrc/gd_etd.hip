// gd_etd.hip -- exponential time differencing for the radial diffusion
// operator: the EAB2 update (timestep.jl:20-65, 425-487) and the staged ERK2
// update (timestep.jl:825-1555), on the reference's l-major spectral planes.
//
// Per degree l the operator is A_l = nu (d2/dr2 + (2/r) d/dr - l(l+1)/r^2)
// (timestep.jl:45-53, 912-919; no boundary rows), dense nr x nr.  Its matrix
// functions are computed once on the host in FP64:
//   exp(h A)       scaling and squaring with the [13/13] Pade approximant
//                  (N. J. Higham, SIAM J. Matrix Anal. Appl. 26 (2005)
//                  1179-1193, Algorithm 2.3 with m = 13)
//   EAB2           phi1 = A^-1 (exp(dt A) - I) / dt by a pivoted LU solve, as
//                  create_etd_cache does (timestep.jl:58-61; no fallback)
//   ERK2           phi1(h A) = (h A)^-1 (exp(h A) - I) and
//                  phi2(h A) = (h A)^-2 (exp(h A) - I - h A) with the
//                  reference's fallbacks to the Taylor series (||hA||_2 < 1e-2,
//                  reciprocal 1-norm condition number below sqrt(eps), a
//                  non-finite result; timestep.jl:970-1187)
// Matrices are kept on the device per l as Julia's Matrix (nr x nr,
// column-major).  The updates are per-(lm, re/im) matrix-vector products: one
// block per (l, plane) stages its l's matrix in LDS and sweeps the
// min(l, mmax) + 1 columns of that degree, contiguous in the l-major layout, m
// fastest across threads so that every global access is coalesced.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gd_internal.hpp"

// ---------------------------------------------------------------- host dense linear algebra
namespace {
constexpr double kEps = 2.220446049250313e-16;
using Mat = std::vector<double>; // n x n column-major: A(i, j) = A[i + n j]

inline double &at(Mat &A, int n, int i, int j) { return A[(size_t)i + (size_t)n * j]; }
inline double at(const Mat &A, int n, int i, int j) { return A[(size_t)i + (size_t)n * j]; }

Mat matmul(const Mat &A, const Mat &B, int n)
{
    Mat C((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            const double b = at(B, n, k, j);
            if (b == 0.0) continue;
            for (int i = 0; i < n; ++i) at(C, n, i, j) += at(A, n, i, k) * b;
        }
    return C;
}

Mat eye(int n)
{
    Mat I((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) at(I, n, i, i) = 1.0;
    return I;
}

bool all_finite(const Mat &A)
{
    for (double x : A)
        if (!std::isfinite(x)) return false;
    return true;
}

double norm1(const Mat &A, int n)
{
    double m = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(at(A, n, i, j));
        m = std::max(m, s);
    }
    return m;
}

// ||A||_2: square root of the largest eigenvalue of A^T A by power iteration
// (it only decides the reference's series branch and stopping rule)
double norm2(const Mat &A, int n)
{
    std::vector<double> v(n), w(n), z(n);
    for (int i = 0; i < n; ++i) v[i] = 1.0 + 0.1 * i / n;
    double s = 0.0;
    for (int it = 0; it < 200; ++it) {
        double nv = 0.0;
        for (double x : v) nv += x * x;
        nv = std::sqrt(nv);
        if (nv == 0.0) return 0.0;
        for (double &x : v) x /= nv;
        for (int i = 0; i < n; ++i) {
            double a = 0.0;
            for (int j = 0; j < n; ++j) a += at(A, n, i, j) * v[j];
            w[i] = a;
        }
        for (int j = 0; j < n; ++j) {
            double a = 0.0;
            for (int i = 0; i < n; ++i) a += at(A, n, i, j) * w[i];
            z[j] = a;
        }
        double nz = 0.0;
        for (double x : z) nz += x * x;
        const double s2 = std::sqrt(std::sqrt(nz)); // ||A^T A v|| -> sigma_max^2
        v = z;
        if (std::fabs(s2 - s) <= 1e-10 * s2) return s2;
        s = s2;
    }
    return s;
}

// LU with partial pivoting (the row order of LAPACK getrf); ok = false on an
// exactly zero pivot, where Julia's lu(A) throws SingularException
struct LU {
    int n = 0;
    Mat a;
    std::vector<int> piv;
    bool ok = false;
};

LU lu_factor(const Mat &A, int n)
{
    LU f;
    f.n = n;
    f.a = A;
    f.piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(at(f.a, n, i, k)) > std::fabs(at(f.a, n, p, k))) p = i;
        f.piv[k] = p;
        if (at(f.a, n, p, k) == 0.0) return f;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(at(f.a, n, k, j), at(f.a, n, p, j));
        const double d = at(f.a, n, k, k);
        for (int i = k + 1; i < n; ++i) at(f.a, n, i, k) /= d;
        for (int j = k + 1; j < n; ++j) {
            const double u = at(f.a, n, k, j);
            if (u != 0.0)
                for (int i = k + 1; i < n; ++i) at(f.a, n, i, j) -= at(f.a, n, i, k) * u;
        }
    }
    f.ok = true;
    return f;
}

// B := A^-1 B (n x n)
void lu_solve(const LU &f, Mat &B)
{
    const int n = f.n;
    for (int k = 0; k < n; ++k)
        if (f.piv[k] != k)
            for (int j = 0; j < n; ++j) std::swap(at(B, n, k, j), at(B, n, f.piv[k], j));
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < n; ++k) {
            const double b = at(B, n, k, j);
            if (b != 0.0)
                for (int i = k + 1; i < n; ++i) at(B, n, i, j) -= at(f.a, n, i, k) * b;
        }
        for (int k = n - 1; k >= 0; --k) {
            const double b = at(B, n, k, j) /= at(f.a, n, k, k);
            if (b != 0.0)
                for (int i = 0; i < k; ++i) at(B, n, i, j) -= at(f.a, n, i, k) * b;
        }
    }
}

// reciprocal 1-norm condition number 1 / (||A||_1 ||A^-1||_1), with A^-1
// formed explicitly (LAPACK gecon estimates the same quantity)
double rcond1(const LU &f, const Mat &A, int n)
{
    Mat inv = eye(n);
    lu_solve(f, inv);
    const double a = norm1(A, n), b = norm1(inv, n);
    return (a == 0.0 || !(b > 0.0) || !std::isfinite(b)) ? 0.0 : 1.0 / (a * b);
}

// exp(A): scaling and squaring with the [13/13] Pade approximant (Higham 2005)
bool expm(const Mat &A0, int n, Mat &E)
{
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    const double theta13 = 5.371920351148152;
    const double nA = norm1(A0, n);
    if (!std::isfinite(nA)) return false;
    int s = 0;
    if (nA > theta13) s = std::max(0, (int)std::ceil(std::log2(nA / theta13)));
    Mat A = A0;
    const double sc = std::ldexp(1.0, -s);
    for (double &x : A) x *= sc;
    const Mat I = eye(n);
    const Mat A2 = matmul(A, A, n), A4 = matmul(A2, A2, n), A6 = matmul(A4, A2, n);
    const size_t nn = (size_t)n * n;
    Mat t(nn);
    for (size_t q = 0; q < nn; ++q) t[q] = b[13] * A6[q] + b[11] * A4[q] + b[9] * A2[q];
    Mat U1 = matmul(A6, t, n);
    for (size_t q = 0; q < nn; ++q) U1[q] += b[7] * A6[q] + b[5] * A4[q] + b[3] * A2[q] + b[1] * I[q];
    const Mat U = matmul(A, U1, n);
    for (size_t q = 0; q < nn; ++q) t[q] = b[12] * A6[q] + b[10] * A4[q] + b[8] * A2[q];
    Mat V = matmul(A6, t, n);
    for (size_t q = 0; q < nn; ++q) V[q] += b[6] * A6[q] + b[4] * A4[q] + b[2] * A2[q] + b[0] * I[q];
    Mat P(nn), Q(nn);
    for (size_t q = 0; q < nn; ++q) {
        P[q] = V[q] + U[q];
        Q[q] = V[q] - U[q];
    }
    const LU f = lu_factor(Q, n);
    if (!f.ok) return false;
    lu_solve(f, P); // r13 = Q^-1 P
    E = P;
    for (int k = 0; k < s; ++k) E = matmul(E, E, n);
    return all_finite(E);
}

// Taylor series sum_{j>=0} A^j / (j + k)! with the reference's stopping rule:
// at most 15 powers, stop once ||term||_2 < 100 eps (for k = 2 the first
// power's term is added unchecked, timestep.jl:1086-1099)
Mat phi_series(const Mat &A, int n, int k)
{
    Mat result = eye(n), P = eye(n);
    double fact = (k == 1) ? 1.0 : 2.0;
    for (double &x : result) x /= fact;
    for (int j = 1; j <= 15; ++j) {
        P = matmul(P, A, n);
        fact *= (double)(j + k);
        Mat term = P;
        for (double &x : term) x /= fact;
        for (size_t q = 0; q < result.size(); ++q) result[q] += term[q];
        if (!(k == 2 && j == 1) && norm2(term, n) < kEps * 100.0) break;
    }
    return result;
}

// compute_phi1_function (timestep.jl:970-1065)
Mat phi1_of(const Mat &A, const Mat &E, int n)
{
    if (!all_finite(A) || !all_finite(E)) return eye(n);
    if (norm2(A, n) < 1e-2) return phi_series(A, n, 1);
    const LU f = lu_factor(A, n);
    if (!f.ok || rcond1(f, A, n) < std::sqrt(kEps)) return phi_series(A, n, 1);
    Mat R = E;
    for (int i = 0; i < n; ++i) at(R, n, i, i) -= 1.0;
    lu_solve(f, R);
    if (!all_finite(R)) return phi_series(A, n, 1);
    return R;
}

// compute_phi2_function (timestep.jl:1073-1187)
Mat phi2_of(const Mat &A, const Mat &E, int n)
{
    if (!all_finite(A) || !all_finite(E)) return Mat((size_t)n * n, 0.0);
    if (norm2(A, n) < 1e-2) return phi_series(A, n, 2);
    const LU f = lu_factor(A, n);
    if (!f.ok || rcond1(f, A, n) < std::sqrt(kEps)) return phi_series(A, n, 2);
    Mat R = E;
    for (size_t q = 0; q < R.size(); ++q) R[q] -= A[q];
    for (int i = 0; i < n; ++i) at(R, n, i, i) -= 1.0;
    lu_solve(f, R);
    lu_solve(f, R);
    if (!all_finite(R)) return phi_series(A, n, 2);
    return R;
}

// A_l = nu (Lap - l(l+1)/r^2 on the diagonal), dense, from the band Laplacian
// lap[(2bw+1) j + bw + i - j] = Lap(i, j); r is the (nr, 7) radial table,
// column 1 = 1/r^2
Mat operator_l(const std::vector<double> &lap, const double *r, int n, int bw, double nu, int l)
{
    const int W = 2 * bw + 1;
    Mat A((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = std::max(0, j - bw); i <= std::min(n - 1, j + bw); ++i)
            at(A, n, i, j) = nu * lap[(size_t)j * W + bw + i - j];
    const double lf = (double)l * (double)(l + 1);
    for (int i = 0; i < n; ++i) at(A, n, i, i) -= nu * lf * r[i + n];
    return A;
}
} // namespace

extern "C" int gd_exp_functions(int lmax, int nr, int bw, const double *r, double nu, double dt, int fd_mode,
                                int which, double *out)
{
    GD_CHECK(lmax >= 0 && nr >= 2 && bw >= 1 && r && out && dt > 0.0 && (which == GD_EXP_ETD || which == GD_EXP_ERK2),
             GD_ERR_ARG, "gd_exp_functions: bad arguments");
    std::vector<double> lap((size_t)(2 * bw + 1) * nr);
    GD_TRY(gd_radial_laplacian(nr, bw, r, fd_mode, lap.data()));
    const size_t nn = (size_t)nr * nr, blk = (size_t)(lmax + 1) * nn;
    for (int l = 0; l <= lmax; ++l) {
        const Mat A = operator_l(lap, r, nr, bw, nu, l);
        auto put = [&](int k, const Mat &M) { std::copy(M.begin(), M.end(), out + k * blk + (size_t)l * nn); };
        if (which == GD_EXP_ETD) {
            Mat hA = A, E;
            for (double &x : hA) x *= dt;
            GD_CHECK(expm(hA, nr, E), GD_ERR_NUMERIC, "create_etd_cache: exp(dt A) failed at l=%d", l);
            const LU f = lu_factor(A, nr);
            GD_CHECK(f.ok, GD_ERR_NUMERIC, "create_etd_cache: A_l singular at l=%d", l);
            Mat F = E;
            for (int i = 0; i < nr; ++i) at(F, nr, i, i) -= 1.0;
            for (double &x : F) x /= dt;
            lu_solve(f, F);
            put(0, E);
            put(1, F);
        } else {
            Mat hh = A, hf = A, Eh, Ef;
            for (double &x : hh) x *= dt / 2.0;
            for (double &x : hf) x *= dt;
            GD_CHECK(expm(hh, nr, Eh) && expm(hf, nr, Ef), GD_ERR_NUMERIC,
                     "create_erk2_cache: exp(h A) failed at l=%d", l);
            put(0, Eh);
            put(1, Ef);
            put(2, phi1_of(hh, Eh, nr));
            put(3, phi1_of(hf, Ef, nr));
            put(4, phi2_of(hf, Ef, nr));
        }
    }
    return GD_OK;
}

// ---------------------------------------------------------------- device: per-l dense products
namespace {
constexpr int ETD_THREADS = 256;
constexpr int RG = 8; // rows per work item

// LDS bytes of the staged matrix: nr rows padded to a multiple of RG, row stride nr + 1
inline size_t dense_lds(int nr) { return sizeof(double) * (size_t)((nr + RG - 1) / RG * RG) * (nr + 1); }

// out = b0 y0 + b1 y1 + sum over passes p of alpha_p M_p,l (ca_p x_p + cb_p xb_p)
// on l-major planes (element (lm, r) at lm + nlm r).  A null M_p skips pass p,
// a null y or xb skips that term.  Block (l, plane).
struct DenseArgs {
    const double *M[2]; // [slot] column-major nr x nr
    double alpha[2], ca[2], cb[2];
    const double *x[2][2]; // [pass][plane]
    const double *xb[2][2];
    const double *y0[2], *y1[2];
    double b0, b1;
    double *out[2];
    const int *slot_of_l; // l -> matrix slot (-1: none)
    const int *col0;      // l-major start of each l, lmax + 2 entries
    int nr, nlm;
    int *status;
};

// Work item (row group g, column): RG rows of one column, so that every x
// load feeds RG multiply-adds.  S holds the matrix padded to a multiple of RG
// rows with zeros (no row guard in the inner loop).
__global__ void __launch_bounds__(ETD_THREADS) dense_lmajor_kernel(DenseArgs a)
{
    extern __shared__ double S[]; // npad x (nr + 1), S[r (nr+1) + j] = M(r, j)
    const int l = blockIdx.x, pl = blockIdx.y, n = a.nr, ld = n + 1;
    const int slot = a.slot_of_l[l];
    if (slot < 0) {
        // the reference errors on an l without a cache entry (timestep.jl:1366-1368)
        if (threadIdx.x == 0) atomicOr(a.status, 4);
        return;
    }
    const int ng = (n + RG - 1) / RG, npad = ng * RG;
    const int c0 = a.col0[l], nc = a.col0[l + 1] - c0, total = nc * ng;
    double *out = a.out[pl];
    const size_t nlm = (size_t)a.nlm;
    bool first = true;
    for (int p = 0; p < 2; ++p) {
        if (!a.M[p]) continue;
        const double *Mg = a.M[p] + (size_t)slot * n * n;
        __syncthreads(); // the previous pass has finished reading S
        for (int e = threadIdx.x; e < npad * n; e += ETD_THREADS) {
            const int j = e / npad, r = e - j * npad;
            S[r * ld + j] = r < n ? Mg[r + (size_t)n * j] : 0.0;
        }
        __syncthreads();
        const double *x = a.x[p][pl], *xb = a.xb[p][pl];
        const double ca = a.ca[p], cb = a.cb[p], alpha = a.alpha[p];
        for (int e = threadIdx.x; e < total; e += ETD_THREADS) {
            const int g = e / nc, mc = e - g * nc;
            const size_t col = (size_t)c0 + mc;
            const double *Sg = S + g * RG * ld;
            double acc[RG];
#pragma unroll
            for (int q = 0; q < RG; ++q) acc[q] = 0.0;
            for (int j = 0; j < n; ++j) {
                double v = ca * x[col + nlm * j];
                if (xb) v += cb * xb[col + nlm * j];
#pragma unroll
                for (int q = 0; q < RG; ++q) acc[q] += Sg[q * ld + j] * v;
            }
#pragma unroll
            for (int q = 0; q < RG; ++q) {
                const int r = g * RG + q;
                if (r >= n) break;
                const size_t o = col + nlm * r;
                double v;
                if (first) {
                    v = 0.0;
                    if (a.y0[pl]) v += a.b0 * a.y0[pl][o];
                    if (a.y1[pl]) v += a.b1 * a.y1[pl][o];
                } else {
                    v = out[o]; // written by this same thread in the previous pass
                }
                out[o] = v + alpha * acc[q];
            }
        }
        first = false;
    }
}
} // namespace

// ---------------------------------------------------------------- cache objects
struct gd_exp_cache {
    gd_ctx *ctx = nullptr;
    int kind = 0; // GD_EXP_ETD or GD_EXP_ERK2
    int nslots = 0, nr = 0, lmax = -1, nmat = 0;
    double dt = 0.0;
    std::vector<int> l_values;
    DevBuf mats;      // nmat blocks of nslots column-major nr x nr matrices
    DevBuf slot_of_l; // lmax + 1 ints
    DevBuf status;
    DevBuf col0; // l-major column starts for (c_lmax, c_mmax)
    int c_lmax = -1, c_mmax = -1, c_nlm = 0;
    DevBuf work, hin; // result planes; host staging
    const double *mat(int k) const { return mats.as<double>() + (size_t)k * nslots * nr * nr; }
    void release()
    {
        for (DevBuf *b : {&mats, &slot_of_l, &status, &col0, &work, &hin}) b->release();
    }
};
struct gd_etd_cache : gd_exp_cache {};  // ETDCache: E, phi1 (timestep.jl:24-29)
struct gd_erk2_cache : gd_exp_cache {}; // ERK2Cache: E_half, E_full, phi1_half, phi1_full, phi2_full (timestep.jl:830-851)

namespace {
int cache_upload(gd_ctx *ctx, gd_exp_cache *c, int kind, const std::vector<int> &lvals, int nr, double dt,
                 const std::vector<const double *> &host_mats)
{
    c->ctx = ctx;
    c->kind = kind;
    c->nslots = (int)lvals.size();
    c->nr = nr;
    c->dt = dt;
    c->l_values = lvals;
    c->nmat = (int)host_mats.size();
    c->lmax = -1;
    for (int l : lvals) {
        GD_CHECK(l >= 0, GD_ERR_ARG, "exponential cache: negative l");
        c->lmax = std::max(c->lmax, l);
    }
    std::vector<int> slot(c->lmax + 1, -1);
    for (int s = 0; s < c->nslots; ++s)
        if (slot[lvals[s]] < 0) slot[lvals[s]] = s; // first match, as findfirst
    const size_t per = (size_t)c->nslots * nr * nr;
    GD_HIP(hipSetDevice(ctx->device));
    GD_TRY(c->mats.ensure(sizeof(double) * per * c->nmat));
    for (int k = 0; k < c->nmat; ++k)
        GD_HIP(hipMemcpyAsync(c->mats.as<double>() + k * per, host_mats[k], sizeof(double) * per,
                              hipMemcpyHostToDevice, ctx->stream));
    GD_TRY(c->slot_of_l.ensure(sizeof(int) * slot.size()));
    GD_HIP(hipMemcpyAsync(c->slot_of_l.p, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    GD_TRY(c->status.ensure(sizeof(int)));
    GD_HIP(hipMemsetAsync(c->status.p, 0, sizeof(int), ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

template <class C>
int cache_create(gd_ctx *ctx, int kind, const std::vector<int> &lv, int nr, double dt,
                 const std::vector<const double *> &m, C **out)
{
    C *c = new C();
    const int rc = cache_upload(ctx, c, kind, lv, nr, dt, m);
    if (rc != GD_OK) {
        c->release();
        delete c;
        return rc;
    }
    *out = c;
    return GD_OK;
}

int cache_get(const gd_exp_cache *c, int k, double *host)
{
    if (!host) return GD_OK;
    const size_t per = (size_t)c->nslots * c->nr * c->nr;
    GD_HIP(hipSetDevice(c->ctx->device));
    GD_HIP(hipMemcpyAsync(host, c->mat(k), sizeof(double) * per, hipMemcpyDeviceToHost, c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream));
    return GD_OK;
}

int cache_destroy(gd_exp_cache *c)
{
    if (!c) return GD_OK;
    (void)hipSetDevice(c->ctx->device);
    (void)hipStreamSynchronize(c->ctx->stream);
    c->release();
    return GD_OK;
}

// l-major column starts of a (lmax, mmax) field: l occupies min(l, mmax) + 1
// columns (shtnskit_transforms.jl:177-188)
int col_starts(gd_exp_cache *c, int lmax, int mmax)
{
    if (c->c_lmax == lmax && c->c_mmax == mmax) return GD_OK;
    std::vector<int> c0(lmax + 2);
    c0[0] = 0;
    for (int l = 0; l <= lmax; ++l) c0[l + 1] = c0[l] + std::min(l, mmax) + 1;
    GD_TRY(c->col0.ensure(sizeof(int) * c0.size()));
    GD_HIP(hipMemcpyAsync(c->col0.p, c0.data(), sizeof(int) * c0.size(), hipMemcpyHostToDevice, c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream)); // c0 is a stack array
    c->c_lmax = lmax;
    c->c_mmax = mmax;
    c->c_nlm = c0[lmax + 1];
    return GD_OK;
}

int check_field(const gd_exp_cache *c, int lmax, int mmax, int nr, const char *what)
{
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax, GD_ERR_ARG, "%s: bad lmax=%d mmax=%d", what, lmax, mmax);
    GD_CHECK(nr == c->nr, GD_ERR_ARG, "%s: field nr=%d, cache nr=%d", what, nr, c->nr);
    GD_CHECK(lmax <= c->lmax, GD_ERR_MISSING_L, "%s: missing exponential cache entry for l=%d", what, c->lmax + 1);
    GD_CHECK(dense_lds(c->nr) <= 64 * 1024, GD_ERR_ARG, "%s: nr=%d too large for an LDS-staged matrix", what, c->nr);
    return GD_OK;
}

int launch_dense(gd_exp_cache *c, DenseArgs a, int lmax)
{
    a.slot_of_l = c->slot_of_l.as<int>();
    a.status = c->status.as<int>();
    a.col0 = c->col0.as<int>();
    a.nr = c->nr;
    a.nlm = c->c_nlm;
    hipLaunchKernelGGL(dense_lmajor_kernel, dim3(lmax + 1, 2), dim3(ETD_THREADS), dense_lds(c->nr), c->ctx->stream, a);
    return gd_check_launch("dense_lmajor_kernel");
}

// synchronises; GD_ERR_MISSING_L if a block met an l without a cache entry
int read_status(gd_exp_cache *c)
{
    int st = 0;
    GD_HIP(hipMemcpyAsync(&st, c->status.p, sizeof(int), hipMemcpyDeviceToHost, c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream));
    if (st) {
        GD_HIP(hipMemsetAsync(c->status.p, 0, sizeof(int), c->ctx->stream));
        GD_HIP(hipStreamSynchronize(c->ctx->stream));
        gd_set_error("missing exponential cache entry for an l of the field");
        return GD_ERR_MISSING_L;
    }
    return GD_OK;
}

// device views of caller planes; GD_HOST inputs are staged into c->hin
int stage_in(gd_exp_cache *c, size_t plane, int flags, int n, const double *const *src, const double **dev)
{
    if (!(flags & GD_HOST)) {
        for (int k = 0; k < n; ++k) dev[k] = src[k];
        return GD_OK;
    }
    GD_TRY(c->hin.ensure(sizeof(double) * plane * 6));
    for (int k = 0; k < n; ++k) {
        double *d = c->hin.as<double>() + (size_t)k * plane;
        GD_HIP(hipMemcpyAsync(d, src[k], sizeof(double) * plane, hipMemcpyHostToDevice, c->ctx->stream));
        dev[k] = d;
    }
    return GD_OK;
}

int copy_out(gd_exp_cache *c, size_t plane, int flags, const double *src, double *dst)
{
    GD_HIP(hipMemcpyAsync(dst, src, sizeof(double) * plane,
                          (flags & GD_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->ctx->stream));
    return GD_OK;
}
} // namespace

// ---------------------------------------------------------------- EAB2
extern "C" int gd_etd_cache_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu, double dt,
                                   int fd_mode, gd_etd_cache **out)
{
    GD_CHECK(ctx && out, GD_ERR_ARG, "gd_etd_cache_create: null argument");
    GD_CHECK(lmax >= 0 && nr >= 2, GD_ERR_ARG, "gd_etd_cache_create: bad lmax=%d nr=%d", lmax, nr);
    const size_t blk = (size_t)(lmax + 1) * nr * nr;
    std::vector<double> m(2 * blk);
    GD_TRY(gd_exp_functions(lmax, nr, bw, r, nu, dt, fd_mode, GD_EXP_ETD, m.data()));
    std::vector<int> lv(lmax + 1);
    for (int l = 0; l <= lmax; ++l) lv[l] = l;
    return cache_create(ctx, GD_EXP_ETD, lv, nr, dt, {m.data(), m.data() + blk}, out);
}

extern "C" int gd_etd_cache_get(gd_etd_cache *c, double *E, double *phi1)
{
    GD_CHECK(c, GD_ERR_ARG, "gd_etd_cache_get: null cache");
    GD_TRY(cache_get(c, 0, E));
    return cache_get(c, 1, phi1);
}

extern "C" int gd_etd_cache_destroy(gd_etd_cache *c)
{
    const int rc = cache_destroy(c);
    delete c;
    return rc;
}

extern "C" int gd_eab2_update(gd_etd_cache *c, int lmax, int mmax, int nr, double *u_re, double *u_im,
                              const double *nl_re, const double *nl_im, const double *np_re, const double *np_im,
                              double dt, int flags)
{
    GD_CHECK(c && u_re && u_im && nl_re && nl_im && np_re && np_im, GD_ERR_ARG, "gd_eab2_update: null argument");
    GD_TRY(check_field(c, lmax, mmax, nr, "gd_eab2_update"));
    gd_ctx *ctx = c->ctx;
    GD_HIP(hipSetDevice(ctx->device));
    GD_TRY(col_starts(c, lmax, mmax));
    const size_t plane = (size_t)c->c_nlm * nr;
    const double *src[6] = {u_re, u_im, nl_re, nl_im, np_re, np_im}, *d[6];
    GD_TRY(stage_in(c, plane, flags, 6, src, d));
    GD_TRY(c->work.ensure(sizeof(double) * plane * 2));
    // u = E u + dt phi1 (3/2 nl - 1/2 nl_prev)
    DenseArgs a{};
    a.M[0] = c->mat(0);
    a.M[1] = c->mat(1);
    a.alpha[0] = 1.0;
    a.ca[0] = 1.0;
    a.alpha[1] = dt;
    a.ca[1] = 1.5;
    a.cb[1] = -0.5;
    for (int p = 0; p < 2; ++p) {
        a.x[0][p] = d[p];
        a.x[1][p] = d[2 + p];
        a.xb[1][p] = d[4 + p];
        a.out[p] = c->work.as<double>() + p * plane;
    }
    GD_TRY(launch_dense(c, a, lmax));
    GD_TRY(read_status(c));
    GD_TRY(copy_out(c, plane, flags, c->work.as<double>(), u_re));
    GD_TRY(copy_out(c, plane, flags, c->work.as<double>() + plane, u_im));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

// ---------------------------------------------------------------- ERK2
extern "C" int gd_erk2_cache_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu, double dt,
                                    int fd_mode, gd_erk2_cache **out)
{
    GD_CHECK(ctx && out, GD_ERR_ARG, "gd_erk2_cache_create: null argument");
    GD_CHECK(lmax >= 0 && nr >= 2, GD_ERR_ARG, "gd_erk2_cache_create: bad lmax=%d nr=%d", lmax, nr);
    const size_t blk = (size_t)(lmax + 1) * nr * nr;
    std::vector<double> m(5 * blk);
    GD_TRY(gd_exp_functions(lmax, nr, bw, r, nu, dt, fd_mode, GD_EXP_ERK2, m.data()));
    std::vector<int> lv(lmax + 1);
    for (int l = 0; l <= lmax; ++l) lv[l] = l;
    return cache_create(ctx, GD_EXP_ERK2, lv, nr, dt,
                        {m.data(), m.data() + blk, m.data() + 2 * blk, m.data() + 3 * blk, m.data() + 4 * blk}, out);
}

extern "C" int gd_erk2_cache_from_matrices(gd_ctx *ctx, int nl, const int *l_values, int nr, double dt,
                                           const double *E_half, const double *E_full, const double *phi1_half,
                                           const double *phi1_full, const double *phi2_full, gd_erk2_cache **out)
{
    GD_CHECK(ctx && out && l_values && nl >= 1 && nr >= 1 && E_half && E_full && phi1_half && phi1_full &&
                 phi2_full,
             GD_ERR_ARG, "gd_erk2_cache_from_matrices: bad arguments");
    std::vector<int> lv(l_values, l_values + nl);
    return cache_create(ctx, GD_EXP_ERK2, lv, nr, dt, {E_half, E_full, phi1_half, phi1_full, phi2_full}, out);
}

extern "C" int gd_erk2_cache_get(gd_erk2_cache *c, double *E_half, double *E_full, double *phi1_half,
                                 double *phi1_full, double *phi2_full)
{
    GD_CHECK(c, GD_ERR_ARG, "gd_erk2_cache_get: null cache");
    double *o[5] = {E_half, E_full, phi1_half, phi1_full, phi2_full};
    for (int k = 0; k < 5; ++k) GD_TRY(cache_get(c, k, o[k]));
    return GD_OK;
}

extern "C" int gd_erk2_cache_destroy(gd_erk2_cache *c)
{
    const int rc = cache_destroy(c);
    delete c;
    return rc;
}

// ERK2FieldBuffers (timestep.jl:1283-1313): the linear, k1, stage, n_current
// and stage_nl planes (re, im) of one field, device resident
struct gd_erk2_buffers {
    gd_erk2_cache *cache = nullptr;
    int lmax = 0, mmax = 0, nr = 0;
    size_t plane = 0;
    DevBuf planes; // 10 planes: linear, k1, stage, n_current, stage_nl; re then im
    double *pl(int k) { return planes.as<double>() + (size_t)k * plane; }
};

extern "C" int gd_erk2_buffers_create(gd_erk2_cache *c, int lmax, int mmax, int nr, gd_erk2_buffers **out)
{
    GD_CHECK(c && out, GD_ERR_ARG, "gd_erk2_buffers_create: null argument");
    GD_CHECK(c->nslots > 0, GD_ERR_ARG, "ERK2 cache has no precomputed matrices");
    GD_TRY(check_field(c, lmax, mmax, nr, "gd_erk2_buffers_create"));
    GD_HIP(hipSetDevice(c->ctx->device));
    gd_erk2_buffers *b = new gd_erk2_buffers();
    b->cache = c;
    b->lmax = lmax;
    b->mmax = mmax;
    b->nr = nr;
    b->plane = (size_t)gd_nlm_of(lmax, mmax) * nr;
    int rc = b->planes.ensure(sizeof(double) * 10 * b->plane);
    if (rc == GD_OK && (hipMemsetAsync(b->planes.p, 0, sizeof(double) * 10 * b->plane, c->ctx->stream) != hipSuccess ||
                        hipStreamSynchronize(c->ctx->stream) != hipSuccess)) {
        gd_set_error("gd_erk2_buffers_create: HIP failure");
        rc = GD_ERR_HIP;
    }
    if (rc != GD_OK) {
        b->planes.release();
        delete b;
        return rc;
    }
    *out = b;
    return GD_OK;
}

extern "C" int gd_erk2_buffers_destroy(gd_erk2_buffers *b)
{
    if (!b) return GD_OK;
    (void)hipSetDevice(b->cache->ctx->device);
    (void)hipStreamSynchronize(b->cache->ctx->stream);
    b->planes.release();
    delete b;
    return GD_OK;
}

extern "C" int gd_erk2_prepare_field(gd_erk2_buffers *b, const double *u_re, const double *u_im,
                                     const double *nl_re, const double *nl_im, double dt, int flags)
{
    GD_CHECK(b && u_re && u_im && nl_re && nl_im, GD_ERR_ARG, "gd_erk2_prepare_field: null argument");
    gd_erk2_cache *c = b->cache;
    GD_HIP(hipSetDevice(c->ctx->device));
    GD_TRY(col_starts(c, b->lmax, b->mmax));
    const double *src[4] = {u_re, u_im, nl_re, nl_im}, *d[4];
    GD_TRY(stage_in(c, b->plane, flags, 4, src, d));
    for (int p = 0; p < 2; ++p) // n_current = nl
        GD_HIP(hipMemcpyAsync(b->pl(6 + p), d[2 + p], sizeof(double) * b->plane, hipMemcpyDeviceToDevice,
                              c->ctx->stream));
    // matrices: 0 E_half, 1 E_full, 2 phi1_half, 3 phi1_full, 4 phi2_full
    DenseArgs a{};
    a.alpha[0] = 1.0;
    a.ca[0] = 1.0;
    a.M[0] = c->mat(1); // linear = E_full u
    for (int p = 0; p < 2; ++p) {
        a.x[0][p] = d[p];
        a.out[p] = b->pl(0 + p);
    }
    GD_TRY(launch_dense(c, a, b->lmax));
    a.M[0] = c->mat(3); // k1 = phi1_full nl
    for (int p = 0; p < 2; ++p) {
        a.x[0][p] = b->pl(6 + p);
        a.out[p] = b->pl(2 + p);
    }
    GD_TRY(launch_dense(c, a, b->lmax));
    a.M[0] = c->mat(0); // stage = E_half u + dt/2 phi1_half nl
    a.M[1] = c->mat(2);
    a.alpha[1] = dt / 2.0;
    a.ca[1] = 1.0;
    for (int p = 0; p < 2; ++p) {
        a.x[0][p] = d[p];
        a.x[1][p] = b->pl(6 + p);
        a.out[p] = b->pl(4 + p);
    }
    GD_TRY(launch_dense(c, a, b->lmax));
    return read_status(c);
}

extern "C" int gd_erk2_apply_stage(gd_erk2_buffers *b, double *u_re, double *u_im, int flags)
{
    GD_CHECK(b && u_re && u_im, GD_ERR_ARG, "gd_erk2_apply_stage: null argument");
    gd_exp_cache *c = b->cache;
    GD_HIP(hipSetDevice(c->ctx->device));
    GD_TRY(copy_out(c, b->plane, flags, b->pl(4), u_re));
    GD_TRY(copy_out(c, b->plane, flags, b->pl(5), u_im));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(c->ctx->stream));
    return GD_OK;
}

extern "C" int gd_erk2_store_stage_nonlinear(gd_erk2_buffers *b, const double *nl_re, const double *nl_im, int flags)
{
    GD_CHECK(b && nl_re && nl_im, GD_ERR_ARG, "gd_erk2_store_stage_nonlinear: null argument");
    gd_exp_cache *c = b->cache;
    GD_HIP(hipSetDevice(c->ctx->device));
    const hipMemcpyKind kind = (flags & GD_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    GD_HIP(hipMemcpyAsync(b->pl(8), nl_re, sizeof(double) * b->plane, kind, c->ctx->stream));
    GD_HIP(hipMemcpyAsync(b->pl(9), nl_im, sizeof(double) * b->plane, kind, c->ctx->stream));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(c->ctx->stream));
    return GD_OK;
}

extern "C" int gd_erk2_finalize_field(gd_erk2_buffers *b, double *u_re, double *u_im, double dt, int flags)
{
    GD_CHECK(b && u_re && u_im, GD_ERR_ARG, "gd_erk2_finalize_field: null argument");
    gd_erk2_cache *c = b->cache;
    GD_HIP(hipSetDevice(c->ctx->device));
    GD_TRY(col_starts(c, b->lmax, b->mmax));
    GD_TRY(c->work.ensure(sizeof(double) * b->plane * 2));
    // u = linear + dt k1 + dt phi2_full (stage_nl - n_current)
    DenseArgs a{};
    a.M[0] = c->mat(4);
    a.alpha[0] = dt;
    a.ca[0] = 1.0;
    a.cb[0] = -1.0;
    a.b0 = 1.0;
    a.b1 = dt;
    for (int p = 0; p < 2; ++p) {
        a.x[0][p] = b->pl(8 + p);
        a.xb[0][p] = b->pl(6 + p);
        a.y0[p] = b->pl(0 + p);
        a.y1[p] = b->pl(2 + p);
        a.out[p] = c->work.as<double>() + p * b->plane;
    }
    GD_TRY(launch_dense(c, a, b->lmax));
    GD_TRY(read_status(c));
    GD_TRY(copy_out(c, b->plane, flags, c->work.as<double>(), u_re));
    GD_TRY(copy_out(c, b->plane, flags, c->work.as<double>() + b->plane, u_im));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(c->ctx->stream));
    return GD_OK;
}
