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
constexpr int TC = 32; // lm columns per block (two 16-column MFMA tiles)
constexpr int XB = 8;  // x-tile loads per batch
constexpr int MAX_PASS = 4;

// Device matrices: per slot nr columns of ldm = nr rounded up to even doubles
// (row nr, when present, is zero).
inline int ldm_of(int nr) { return nr + (nr & 1); }

// One pass: acc += M_l (alpha x) on one re/im plane; with out set,
// out = b0 y0 + b1 y1 + acc is stored after the pass and acc restarts; xcopy,
// if set, receives a copy of x.  A pass's x must not be an earlier pass's out.
// Consecutive passes with the same M stage it once.  Planes are l-major:
// element (lm, r) at lm + nlm r.
struct DensePass {
    const double *M; // [slot][column][ldm]
    double alpha;
    const double *x[2];
    double *xcopy[2];
    double *out[2];
    const double *y0[2], *y1[2];
    double b0, b1;
};
struct DenseProg {
    DensePass p[MAX_PASS];
    int npass;
    const int *slot_of_l; // l -> matrix slot (the host has checked every l of the field)
    const int4 *tiles;    // (l, first lm column, column count, 0)
    int nr, nlm, ldm;
    int nch;              // row chunks per (tile, plane)
};

// Launch shape: wave w of a block computes the 16-row tile w of the block's
// rows against both 16-column tiles.  nr <= 68: one block holds every row
// (W = ceil(nr/16) waves; in-place outputs are safe); 69 <= nr <= 127: one
// wave per 16-row chunk, one block per chunk.
inline int waves_of(int nr) { return nr <= 68 ? (nr + 15) / 16 : 1; }
inline int chunks_of(int nr) { return nr <= 68 ? 1 : (nr + 15) / 16; }
__host__ __device__ inline int kpad_of(int nr) { return (nr + 3) & ~3; }
// LDS: x tile (kpad x 32) then the block's matrix rows (kpad columns of 16 W)
inline size_t dense_lds(int nr) { return sizeof(double) * (size_t)kpad_of(nr) * (TC + 16 * waves_of(nr)); }

// Block (tile of <= 32 columns of one l, plane, row chunk), from a 1-D grid
// in XCD-aware order: the blocks of one tile (both planes, every chunk) and
// the tiles of one l are consecutive logical blocks, so the blocks that read
// one matrix share an XCD's L2.  Per pass the x tile is loaded (coalesced, two rows per wave
// load; x columns past the tile are clamped copies whose results are not
// stored) and, when it changes, the block's matrix rows are copied global ->
// LDS by the lanes directly (global_load_lds, 16 bytes per lane, no
// registers), then each wave runs FP64
// MFMA 16x16x4 over k: A = M rows (16 x 4), B = alpha x (4 x 16).  Columns
// and rows k >= nr of the tiles are zero.  Stores follow a block barrier, so
// with one chunk an output may alias the pass's input.
template <int W> __global__ void __launch_bounds__(64 * W) dense_mfma_kernel(DenseProg a)
{
    constexpr int RB = 16 * W; // rows per block
    extern __shared__ double sm[];
    const int n = a.nr, ldm = a.ldm, kp = kpad_of(n);
    double *Sx = sm;           // [k][TC]
    double *Sm = sm + kp * TC; // [k][RB]: Sm[k RB + rr] = M(r0 + rr, k)
    const int L = xcd_logical(blockIdx.x, gridDim.x);
    const int ch = L % a.nch, pl = (L / a.nch) & 1;
    const int4 t = a.tiles[L / (2 * a.nch)];
    const int r0 = ch * RB, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, ncol = t.z;
    const size_t nlm = (size_t)a.nlm, col0 = (size_t)t.y;
    const size_t moff = (size_t)a.slot_of_l[t.x] * n * ldm;
    // the k padding (rows n..kp-1 of x, columns n..kp-1 of M) stays zero
    for (int e = tid; e < (kp - n) * (TC + RB); e += 64 * W) {
        if (e < (kp - n) * TC)
            Sx[n * TC + e] = 0.0;
        else
            Sm[n * RB + e - (kp - n) * TC] = 0.0;
    }
    const double *Mslot0 = nullptr;
    doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int pi = 0; pi < MAX_PASS; ++pi) {
        if (pi >= a.npass) break;
        const DensePass &P = a.p[pi];
        __syncthreads(); // the previous pass is done with the LDS tiles
        {
            // x tile: a wave load covers rows 2i and 2i + 1 (lane -> row 2i + lane / 32,
            // column lane % 32, columns past the tile clamped); XB loads are issued
            // before their LDS stores
            const double *xs = P.x[pl] + col0 + min(lane & (TC - 1), ncol - 1);
            const int xr = lane >> 5;
            for (int i0 = wave; 2 * i0 < n; i0 += W * XB) {
                double v[XB];
#pragma unroll
                for (int b = 0; b < XB; ++b) v[b] = xs[nlm * min(2 * (i0 + b * W) + xr, n - 1)];
#pragma unroll
                for (int b = 0; b < XB; ++b) {
                    const int k = 2 * (i0 + b * W) + xr;
                    if (k < n) Sx[k * TC + (lane & (TC - 1))] = v[b];
                }
            }
        }
        if (P.M != Mslot0) {
            // matrix rows r0 .. r0 + RB - 1 of columns 0 .. n - 1 in 16-byte row pairs:
            // pair q of the flattened [k][RB] block (columns start 16-byte aligned,
            // ldm is even); row pairs past the matrix clamped (not stored)
            const double *Ms = P.M + moff;
            const int nq = n * RB / 2;
            for (int i = wave; i * 64 < nq; i += W) {
                const int q = i * 64 + lane;
                if (q < nq) {
                    const int k = q / (RB / 2), w2 = q - k * (RB / 2);
                    const int row = min(r0 + 2 * w2, ldm - 2);
                    __builtin_amdgcn_global_load_lds(Ms + (size_t)k * ldm + row, Sm + i * 128, 16, 0, 0);
                }
            }
            Mslot0 = P.M;
        }
        __builtin_amdgcn_s_waitcnt(0); // this wave's direct LDS copies have landed
        __syncthreads();
        if (P.xcopy[pl] && ch == 0) {
            for (int e = tid; e < n * TC; e += 64 * W)
                if ((e & (TC - 1)) < ncol) P.xcopy[pl][col0 + (e & (TC - 1)) + nlm * (e >> 5)] = Sx[e];
        }
        // MFMA 16x16x4: lane holds A[lane % 16][lane / 16], B[lane / 16][lane % 16]
        const double al = P.alpha;
        const double *Sa = Sm + wave * 16 + (lane & 15) + (lane >> 4) * RB;
        const double *Sb = Sx + (lane & 15) + (lane >> 4) * TC;
#pragma unroll 4
        for (int k0 = 0; k0 < kp; k0 += 4) {
            const double av = Sa[k0 * RB];
            const double b0 = al * Sb[k0 * TC], b1 = al * Sb[k0 * TC + 16];
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1, acc1, 0, 0, 0);
        }
        if (P.out[pl]) {
            __syncthreads(); // every wave's reads of this pass are done (in-place outputs)
            // accumulator element v of lane: row (lane / 16) + 4 v, column lane % 16
            const double *y0 = P.y0[pl], *y1 = P.y1[pl];
            if (y0 || y1) {
                // the y terms of all 8 outputs loaded first (at clamped rows and
                // columns; the pass's choice of terms is uniform), so the
                // epilogue waits once, not once per guarded load
                double t0[8], t1[8];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = min(r0 + wave * 16 + (lane >> 4) + 4 * v, n - 1);
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const size_t o = col0 + min(ct * 16 + (lane & 15), ncol - 1) + nlm * r;
                        t0[2 * v + ct] = y0 ? y0[o] : 0.0;
                        t1[2 * v + ct] = y1 ? y1[o] : 0.0;
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = r0 + wave * 16 + (lane >> 4) + 4 * v;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const int cc = ct * 16 + (lane & 15);
                        if (r < n && cc < ncol) {
                            double val = ct ? acc1[v] : acc0[v];
                            if (y0) val += P.b0 * t0[2 * v + ct];
                            if (y1) val += P.b1 * t1[2 * v + ct];
                            P.out[pl][col0 + cc + nlm * r] = val;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = r0 + wave * 16 + (lane >> 4) + 4 * v;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const int cc = ct * 16 + (lane & 15);
                        if (r < n && cc < ncol) P.out[pl][col0 + cc + nlm * r] = ct ? acc1[v] : acc0[v];
                    }
                }
            }
            acc0 = doublex4{0.0, 0.0, 0.0, 0.0};
            acc1 = doublex4{0.0, 0.0, 0.0, 0.0};
        }
    }
}
} // namespace

// ---------------------------------------------------------------- cache objects
struct gd_exp_cache {
    gd_ctx *ctx = nullptr;
    int kind = 0; // GD_EXP_ETD or GD_EXP_ERK2
    int nslots = 0, nr = 0, ldm = 0, lmax = -1, nmat = 0;
    double dt = 0.0;
    std::vector<int> l_values;
    DevBuf mats;                // nmat blocks of nslots matrices, columns padded to ldm
    std::vector<int> slot_host; // lmax + 1 ints: l -> slot, -1 if the cache has no matrix for l
    DevBuf slot_of_l;           // its device copy
    DevBuf tiles;               // column tiles of a (t_lmax, t_mmax) field
    int t_lmax = -1, t_mmax = -1, ntiles = 0;
    DevBuf hin, hout; // host staging
    size_t per() const { return (size_t)nslots * nr * ldm; }
    const double *mat(int k) const { return mats.as<double>() + (size_t)k * per(); }
    void release()
    {
        for (DevBuf *b : {&mats, &slot_of_l, &tiles, &hin, &hout}) b->release();
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
    c->ldm = ldm_of(nr);
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
    // columns padded from nr to ldm doubles
    std::vector<double> padded(c->per() * c->nmat, 0.0);
    for (int k = 0; k < c->nmat; ++k)
        for (size_t sj = 0; sj < (size_t)c->nslots * nr; ++sj)
            std::copy(host_mats[k] + sj * nr, host_mats[k] + (sj + 1) * nr, padded.begin() + k * c->per() + sj * c->ldm);
    GD_HIP(hipSetDevice(ctx->device));
    GD_TRY(c->mats.ensure(sizeof(double) * padded.size()));
    GD_HIP(hipMemcpyAsync(c->mats.p, padded.data(), sizeof(double) * padded.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    GD_TRY(c->slot_of_l.ensure(sizeof(int) * slot.size()));
    GD_HIP(hipMemcpyAsync(c->slot_of_l.p, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    c->slot_host = slot;
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

// host copy of matrix set k, nr x nr column-major per slot
int cache_get(const gd_exp_cache *c, int k, double *host)
{
    if (!host) return GD_OK;
    std::vector<double> padded(c->per());
    GD_HIP(hipSetDevice(c->ctx->device));
    GD_HIP(hipMemcpyAsync(padded.data(), c->mat(k), sizeof(double) * c->per(), hipMemcpyDeviceToHost,
                          c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream));
    for (size_t sj = 0; sj < (size_t)c->nslots * c->nr; ++sj)
        std::copy(padded.begin() + sj * c->ldm, padded.begin() + sj * c->ldm + c->nr, host + sj * c->nr);
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

// column tiles of a (lmax, mmax) field: l occupies the min(l, mmax) + 1
// l-major columns after those of l - 1 (shtnskit_transforms.jl:177-188),
// cut into tiles of at most 32
int field_tiles(gd_exp_cache *c, int lmax, int mmax)
{
    if (c->t_lmax == lmax && c->t_mmax == mmax) return GD_OK;
    std::vector<int4> t;
    int c0 = 0;
    for (int l = 0; l <= lmax; ++l) {
        const int nc = std::min(l, mmax) + 1;
        for (int s = 0; s < nc; s += TC) t.push_back(make_int4(l, c0 + s, std::min(TC, nc - s), 0));
        c0 += nc;
    }
    GD_TRY(c->tiles.ensure(sizeof(int4) * t.size()));
    GD_HIP(hipMemcpyAsync(c->tiles.p, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice, c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream)); // t is a host temporary
    c->t_lmax = lmax;
    c->t_mmax = mmax;
    c->ntiles = (int)t.size();
    return GD_OK;
}

// shape checks; with check_l also that every l of the field has a matrix
int check_field(const gd_exp_cache *c, int lmax, int mmax, int nr, const char *what, bool check_l = true)
{
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax, GD_ERR_ARG, "%s: bad lmax=%d mmax=%d", what, lmax, mmax);
    GD_CHECK(nr == c->nr, GD_ERR_ARG, "%s: field nr=%d, cache nr=%d", what, nr, c->nr);
    // the reference errors on an l without a cache entry (timestep.jl:447-449, 1357, 1475)
    for (int l = 0; check_l && l <= lmax; ++l)
        GD_CHECK(l <= c->lmax && c->slot_host[l] >= 0, GD_ERR_MISSING_L, "%s: missing exponential cache entry for l=%d",
                 what, l);
    GD_CHECK(c->nr <= 127, GD_ERR_ARG, "%s: nr=%d above the supported 127", what, c->nr);
    return GD_OK;
}

int launch_prog(gd_exp_cache *c, DenseProg a, int lmax, int mmax)
{
    GD_TRY(field_tiles(c, lmax, mmax));
    a.slot_of_l = c->slot_of_l.as<int>();
    a.tiles = c->tiles.as<int4>();
    a.nr = c->nr;
    a.nlm = gd_nlm_of(lmax, mmax);
    a.ldm = c->ldm;
    const int W = waves_of(c->nr);
    a.nch = chunks_of(c->nr);
    const dim3 grid(c->ntiles * 2 * a.nch), block(64 * W);
    const size_t lds = dense_lds(c->nr);
    hipStream_t st = c->ctx->stream;
    switch (W) {
    case 1: hipLaunchKernelGGL(dense_mfma_kernel<1>, grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL(dense_mfma_kernel<2>, grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL(dense_mfma_kernel<3>, grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL(dense_mfma_kernel<4>, grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL(dense_mfma_kernel<5>, grid, block, lds, st, a); break;
    }
    return gd_check_launch("dense_mfma_kernel");
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

// output planes: the caller's (device) or c->hout (GD_HOST, copied back by finish_out)
int out_planes(gd_exp_cache *c, size_t plane, int flags, double *re, double *im, double **dev)
{
    if (!(flags & GD_HOST)) {
        dev[0] = re;
        dev[1] = im;
        return GD_OK;
    }
    GD_TRY(c->hout.ensure(sizeof(double) * plane * 2));
    dev[0] = c->hout.as<double>();
    dev[1] = c->hout.as<double>() + plane;
    return GD_OK;
}

int finish_out(gd_exp_cache *c, size_t plane, int flags, double *re, double *im)
{
    if (!(flags & GD_HOST)) return GD_OK;
    GD_HIP(hipMemcpyAsync(re, c->hout.p, sizeof(double) * plane, hipMemcpyDeviceToHost, c->ctx->stream));
    GD_HIP(hipMemcpyAsync(im, c->hout.as<double>() + plane, sizeof(double) * plane, hipMemcpyDeviceToHost,
                          c->ctx->stream));
    GD_HIP(hipStreamSynchronize(c->ctx->stream));
    return GD_OK;
}

DensePass pass(const double *M, double alpha, const double *x0, const double *x1)
{
    DensePass p{};
    p.M = M;
    p.alpha = alpha;
    p.x[0] = x0;
    p.x[1] = x1;
    return p;
}

void set_out(DensePass &p, double *o0, double *o1)
{
    p.out[0] = o0;
    p.out[1] = o1;
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
    GD_HIP(hipSetDevice(c->ctx->device));
    const size_t plane = (size_t)gd_nlm_of(lmax, mmax) * nr;
    const double *src[6] = {u_re, u_im, nl_re, nl_im, np_re, np_im}, *d[6];
    double *o[2];
    GD_TRY(stage_in(c, plane, flags, 6, src, d));
    // in place when one block holds every row of its columns; otherwise through c->hout
    const bool via_hout = (flags & GD_HOST) || chunks_of(nr) > 1;
    GD_TRY(out_planes(c, plane, via_hout ? GD_HOST : 0, u_re, u_im, o));
    // u = E u + dt phi1 (3/2 nl - 1/2 nl_prev): the phi1 product as two passes
    // over the same staged matrix
    DenseProg a{};
    a.npass = 3;
    a.p[0] = pass(c->mat(0), 1.0, d[0], d[1]);
    a.p[1] = pass(c->mat(1), 1.5 * dt, d[2], d[3]);
    a.p[2] = pass(c->mat(1), -0.5 * dt, d[4], d[5]);
    set_out(a.p[2], o[0], o[1]);
    GD_TRY(launch_prog(c, a, lmax, mmax));
    if (via_hout && !(flags & GD_HOST)) {
        GD_HIP(hipMemcpyAsync(u_re, o[0], sizeof(double) * plane, hipMemcpyDeviceToDevice, c->ctx->stream));
        GD_HIP(hipMemcpyAsync(u_im, o[1], sizeof(double) * plane, hipMemcpyDeviceToDevice, c->ctx->stream));
        return GD_OK;
    }
    return finish_out(c, plane, flags, u_re, u_im);
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
    // ERK2FieldBuffers only sizes the planes; a missing l errors in prepare/finalize as in the reference
    GD_TRY(check_field(c, lmax, mmax, nr, "gd_erk2_buffers_create", false));
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
    GD_TRY(check_field(c, b->lmax, b->mmax, b->nr, "erk2_prepare_field"));
    GD_HIP(hipSetDevice(c->ctx->device));
    const double *src[4] = {u_re, u_im, nl_re, nl_im}, *d[4];
    GD_TRY(stage_in(c, b->plane, flags, 4, src, d));
    // matrices: 0 E_half, 1 E_full, 2 phi1_half, 3 phi1_full, 4 phi2_full
    DenseProg a{};
    a.npass = 4;
    a.p[0] = pass(c->mat(1), 1.0, d[0], d[1]); // linear = E_full u
    set_out(a.p[0], b->pl(0), b->pl(1));
    a.p[1] = pass(c->mat(3), 1.0, d[2], d[3]); // k1 = phi1_full nl; n_current = nl
    a.p[1].xcopy[0] = b->pl(6);
    a.p[1].xcopy[1] = b->pl(7);
    set_out(a.p[1], b->pl(2), b->pl(3));
    a.p[2] = pass(c->mat(0), 1.0, d[0], d[1]); // stage = E_half u + dt/2 phi1_half nl
    a.p[3] = pass(c->mat(2), dt / 2.0, d[2], d[3]);
    set_out(a.p[3], b->pl(4), b->pl(5));
    GD_TRY(launch_prog(c, a, b->lmax, b->mmax));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(c->ctx->stream)); // staged host inputs are read
    return GD_OK;
}

extern "C" int gd_erk2_apply_stage(gd_erk2_buffers *b, double *u_re, double *u_im, int flags)
{
    GD_CHECK(b && u_re && u_im, GD_ERR_ARG, "gd_erk2_apply_stage: null argument");
    gd_exp_cache *c = b->cache;
    GD_HIP(hipSetDevice(c->ctx->device));
    const hipMemcpyKind kind = (flags & GD_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    GD_HIP(hipMemcpyAsync(u_re, b->pl(4), sizeof(double) * b->plane, kind, c->ctx->stream));
    GD_HIP(hipMemcpyAsync(u_im, b->pl(5), sizeof(double) * b->plane, kind, c->ctx->stream));
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
    GD_TRY(check_field(c, b->lmax, b->mmax, b->nr, "erk2_finalize_field"));
    GD_HIP(hipSetDevice(c->ctx->device));
    double *o[2];
    GD_TRY(out_planes(c, b->plane, flags, u_re, u_im, o));
    // u = linear + dt k1 + dt phi2_full (stage_nl - n_current): two passes over phi2
    DenseProg a{};
    a.npass = 2;
    a.p[0] = pass(c->mat(4), dt, b->pl(8), b->pl(9));
    a.p[1] = pass(c->mat(4), -dt, b->pl(6), b->pl(7));
    set_out(a.p[1], o[0], o[1]);
    a.p[1].b0 = 1.0;
    a.p[1].b1 = dt;
    a.p[1].y0[0] = b->pl(0);
    a.p[1].y0[1] = b->pl(1);
    a.p[1].y1[0] = b->pl(2);
    a.p[1].y1[1] = b->pl(3);
    GD_TRY(launch_prog(c, a, b->lmax, b->mmax));
    return finish_out(c, b->plane, flags, u_re, u_im);
}
