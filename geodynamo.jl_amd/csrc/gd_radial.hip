// gd_radial.hip -- FP64 banded radial operators on the GPU:
//   * no-pivot band LU, one 64-lane wavefront per system with the band in LDS
//     (factorize_banded, linear_algebra.jl:23-57)
//   * band solve (solve_banded!, linear_algebra.jl:59-88)
//   * CNAB2 operator assembly per l (create_shtns_timestepping_matrices,
//     timestep.jl:498-543)
//   * rhs / solve / fused rhs+solve in the reference l-major layout, one
//     wavefront per l with L_l and LU(S_l) in LDS (build_rhs_cnab2!,
//     solve_implicit_step!, timestep.jl:610-730)
#include "gd_internal.hpp"
#include "gd_radial.hpp"

#define BANDIDX(bw, row, col) ((row) + (2 * (bw) + 1) * (col))

// ---------------------------------------------------------------- band LU
// One block of 64 threads per system. lanes (i, j) in [1,bw]^2 do the
// rank-1 update of step k; lane (i, 1) stores the multiplier L[i][k].
__global__ void __launch_bounds__(64) band_lu_kernel(const double *__restrict__ ab, double *__restrict__ lu,
                                                     double *__restrict__ lu_t, int n, int bw, int nsys,
                                                     int *__restrict__ status)
{
    extern __shared__ double sb[];
    const int sys = blockIdx.x;
    const int w = 2 * bw + 1;
    const double *src = ab + (size_t)sys * w * n;
    for (int q = threadIdx.x; q < w * n; q += 64) sb[q] = src[q];
    __syncthreads();
    const int li = threadIdx.x / bw + 1, lj = threadIdx.x % bw + 1;
    const bool act = threadIdx.x < bw * bw;
    for (int k = 0; k < n - 1; ++k) {
        const double piv = sb[BANDIDX(bw, bw, k)];
        const int i = k + li, j = k + lj;
        double L = 0.0;
        if (act && i < n) L = sb[BANDIDX(bw, bw + li, k)] / piv;
        __syncthreads();
        if (act && i < n) {
            if (j < n) sb[BANDIDX(bw, bw + i - j, j)] -= L * sb[BANDIDX(bw, bw + k - j, j)];
            if (lj == 1) sb[BANDIDX(bw, bw + li, k)] = L;
        }
        __syncthreads();
    }
    double *dst = lu + (size_t)sys * w * n;
    for (int q = threadIdx.x; q < w * n; q += 64) {
        dst[q] = sb[q];
        // transposed copy: lu_t[(col*w + row)*nsys + sys]
        if (lu_t) lu_t[(size_t)q * nsys + sys] = sb[q];
    }
    for (int c = threadIdx.x; c < n; c += 64) {
        double d = sb[BANDIDX(bw, bw, c)];
        if (d == 0.0 || !isfinite(d)) atomicOr(status, 1);
    }
}

int gd_launch_band_lu(hipStream_t st, const double *ab, double *lu, double *lu_t, int n, int bw, int nsys,
                      int *d_status)
{
    size_t lds = sizeof(double) * (2 * bw + 1) * n;
    GD_CHECK(lds <= 64 * 1024, GD_ERR_ARG, "band too large for LDS (n=%d bw=%d)", n, bw);
    GD_CHECK(bw * bw <= 64, GD_ERR_ARG, "bw=%d too large", bw);
    hipLaunchKernelGGL(band_lu_kernel, dim3(nsys), dim3(64), lds, st, ab, lu, lu_t, n, bw, nsys, d_status);
    return gd_check_launch("band_lu_kernel");
}

// ---------------------------------------------------------------- band solve (generic)
__global__ void band_solve_kernel(const double *__restrict__ lu, const int *__restrict__ sys_of_rhs, int n,
                                  int bw, int nsys, int nrhs, double *__restrict__ x, int *__restrict__ status)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nrhs) return;
    int s = sys_of_rhs ? sys_of_rhs[c] : 0;
    if (s < 0 || s >= nsys) {
        atomicOr(status, 2);
        return;
    }
    const int w = 2 * bw + 1;
    const double *L = lu + (size_t)s * w * n;
    double *xc = x + (size_t)c * n;
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        int j0 = i - bw < 0 ? 0 : i - bw;
        for (int j = j0; j < i; ++j) acc += L[BANDIDX(bw, bw + i - j, j)] * xc[j];
        xc[i] = xc[i] - acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        double acc = 0.0;
        int j1 = i + bw > n - 1 ? n - 1 : i + bw;
        for (int j = i + 1; j <= j1; ++j) acc += L[BANDIDX(bw, bw + i - j, j)] * xc[j];
        xc[i] = (xc[i] - acc) / L[BANDIDX(bw, bw, i)];
    }
}

// ---------------------------------------------------------------- CNAB2 assembly
// L_l = nu*Lap; L_l[bw, n] -= nu l(l+1)/r^2; S_l = -theta L_l; S_l[bw,n] += 1/dt;
// bc_rows: rows 0 and n-1 of S_l become Dirichlet rows (identity)
__global__ void cnab2_assemble_kernel(const double *__restrict__ lap, const double *__restrict__ rinv2, double nu,
                                      double dt, double theta, int nl, int n, int bw, int bc_rows,
                                      double *__restrict__ L, double *__restrict__ S, double *__restrict__ L_t)
{
    const int w = 2 * bw + 1;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot = (long)nl * w * n;
    if (q >= tot) return;
    int l = (int)(q / ((long)w * n));
    int e = (int)(q % ((long)w * n));
    int row = e % w, col = e / w;
    double v = nu * lap[e];
    if (row == bw) v -= nu * ((double)l * (double)(l + 1)) * rinv2[col];
    L[q] = v;
    if (L_t) L_t[(size_t)e * nl + l] = v;
    double s = v * (-theta);
    if (row == bw) s += 1.0 / dt;
    const int i = col + row - bw; // matrix row of this band entry
    if (bc_rows && (i == 0 || i == n - 1)) s = (i == col) ? 1.0 : 0.0;
    S[q] = s;
}

int gd_cnab2_build(hipStream_t st, const double *d_lap, const double *d_rinv2, double nu, double dt,
                   double theta, int nl, int n, int bw, int bc_rows, double *L, double *S, double *LU, double *L_t,
                   double *LU_t, int *d_status)
{
    long tot = (long)nl * (2 * bw + 1) * n;
    hipLaunchKernelGGL(cnab2_assemble_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, d_lap, d_rinv2,
                       nu, dt, theta, nl, n, bw, bc_rows, L, S, L_t);
    GD_TRY(gd_check_launch("cnab2_assemble_kernel"));
    return gd_launch_band_lu(st, S, LU, LU_t, n, bw, nl, d_status);
}

// ---------------------------------------------------------------- l-major CNAB2 (reference layout)
// One block (64 lanes) per l; columns (m, re/im) of that l are spread over the
// lanes; L_l and LU_l are staged in LDS.  mode: 0 rhs only, 1 solve only,
// 2 fused rhs+solve in place on u.
__global__ void __launch_bounds__(64) cnab2_lmajor_kernel(
    const double *__restrict__ Lall, const double *__restrict__ LUall, int nl_mats, int lmax, int mmax, int n,
    int bw, int nlm, double dt, double theta, int mode, int bc_rows, const double *u_re, const double *u_im,
    const double *__restrict__ nl_re, const double *__restrict__ nl_im, const double *__restrict__ np_re,
    const double *__restrict__ np_im, double *out_re, double *out_im, int *__restrict__ status)
{
    extern __shared__ double sm[];
    const int l = blockIdx.x;
    const int w = 2 * bw + 1;
    const int ml = l < mmax ? l : mmax;
    const int lm0 = l <= mmax ? l * (l + 1) / 2 : (mmax + 1) * (mmax + 2) / 2 + (l - mmax - 1) * (mmax + 1);
    if (l >= nl_mats) {
        // build_rhs_cnab2! errors on a missing l (timestep.jl:646);
        // solve_implicit_step! skips it (timestep.jl:709)
        if (mode != 1 && threadIdx.x == 0) atomicOr(status, 4);
        return;
    }
    double *sL = sm, *sLU = sm + w * n;
    for (int q = threadIdx.x; q < w * n; q += 64) {
        sL[q] = Lall[(size_t)l * w * n + q];
        sLU[q] = LUall[(size_t)l * w * n + q];
    }
    __syncthreads();
    const double inv_dt = 1.0 / dt, lw = 1.0 - theta;
    const int ncol = 2 * (ml + 1);
    for (int c = threadIdx.x; c < ncol; c += 64) {
        const int reim = c / (ml + 1), m = c % (ml + 1);
        const size_t base = (size_t)lm0 + m;
        const double *u = reim ? u_im : u_re;
        const double *nlv = reim ? nl_im : nl_re;
        const double *npv = reim ? np_im : np_re;
        double *o = reim ? out_im : out_re;
        if (mode == 0 || mode == 2) {
            // rhs, and for the fused mode the forward substitution on the fly;
            // a window of original u keeps in-place updates safe.
            double win[16];
            for (int q = 0; q < 2 * bw + 1; ++q) {
                int j = q - bw;
                win[q] = (j >= 0 && j < n) ? u[base + (size_t)nlm * j] : 0.0;
            }
            for (int i = 0; i < n; ++i) {
                double lin = 0.0;
                if (lw != 0.0) {
                    int j0 = i - bw < 0 ? 0 : i - bw, j1 = i + bw > n - 1 ? n - 1 : i + bw;
                    for (int j = j0; j <= j1; ++j) lin += sL[BANDIDX(bw, bw + i - j, j)] * win[j - i + bw];
                }
                const size_t oi = base + (size_t)nlm * i;
                double v = inv_dt * win[bw] + 1.5 * nlv[oi] - 0.5 * npv[oi];
                if (lw != 0.0) v += lw * lin;
                if (bc_rows && (i == 0 || i == n - 1)) v = 0.0; // Dirichlet value
                if (mode == 2) {
                    double acc = 0.0;
                    int j0 = i - bw < 0 ? 0 : i - bw;
                    for (int j = j0; j < i; ++j) acc += sLU[BANDIDX(bw, bw + i - j, j)] * o[base + (size_t)nlm * j];
                    v = v - acc;
                }
                // shift window before overwriting (o may alias u)
                for (int q = 0; q < 2 * bw; ++q) win[q] = win[q + 1];
                int jn = i + bw + 1;
                win[2 * bw] = (jn < n) ? u[base + (size_t)nlm * jn] : 0.0;
                o[oi] = v;
            }
        }
        if (mode == 1 || mode == 2) {
            const double *src = (mode == 1) ? u : o;
            if (mode == 1) {
                for (int i = 0; i < n; ++i) {
                    double acc = 0.0;
                    int j0 = i - bw < 0 ? 0 : i - bw;
                    for (int j = j0; j < i; ++j) acc += sLU[BANDIDX(bw, bw + i - j, j)] * o[base + (size_t)nlm * j];
                    o[base + (size_t)nlm * i] = src[base + (size_t)nlm * i] - acc;
                }
            }
            for (int i = n - 1; i >= 0; --i) {
                double acc = 0.0;
                int j1 = i + bw > n - 1 ? n - 1 : i + bw;
                for (int j = i + 1; j <= j1; ++j) acc += sLU[BANDIDX(bw, bw + i - j, j)] * o[base + (size_t)nlm * j];
                o[base + (size_t)nlm * i] = (o[base + (size_t)nlm * i] - acc) / sLU[BANDIDX(bw, bw, i)];
            }
        }
    }
}

int gd_launch_cnab2_lmajor(hipStream_t st, const double *L, const double *LU, int nl_mats, int lmax, int mmax,
                           int n, int bw, double dt, double theta, int mode, int bc_rows, const double *u_re,
                           const double *u_im, const double *nl_re, const double *nl_im, const double *np_re,
                           const double *np_im, double *out_re, double *out_im, int *d_status)
{
    int nlm = gd_nlm_of(lmax, mmax);
    size_t lds = sizeof(double) * 2 * (2 * bw + 1) * n;
    GD_CHECK(lds <= 64 * 1024 && bw <= 7, GD_ERR_ARG, "band too large (n=%d bw=%d)", n, bw);
    hipLaunchKernelGGL(cnab2_lmajor_kernel, dim3(lmax + 1), dim3(64), lds, st, L, LU, nl_mats, lmax, mmax, n, bw,
                       nlm, dt, theta, mode, bc_rows, u_re, u_im, nl_re, nl_im, np_re, np_im, out_re, out_im,
                       d_status);
    return gd_check_launch("cnab2_lmajor_kernel");
}

// ---------------------------------------------------------------- m-major fused CNAB2 for the step
// One thread per (device idx, field) column pair (re, im).  Band values come
// from the l-fastest transposed tables L_t / LU_t[(col*w + row)*nl + l] so a
// wavefront's lanes (consecutive idx of one m, l stepping by 2) read nearly
// contiguous words.  nu index of field f: {0,0,1,1,2}.  Field 4 (temperature)
// has its boundary rows set to 0 before the rhs (thermal.jl:262-313).
template <int BW>
__global__ void __launch_bounds__(256) cnab2_step_kernel(const double *__restrict__ Lt, const double *__restrict__ LUt,
                                                         const int *__restrict__ l_of_idx, int nl, int n, int nlm,
                                                         double dt, double theta, int bc_mask, double *__restrict__ st,
                                                         const double *__restrict__ nl_arr,
                                                         const double *__restrict__ np_arr)
{
    constexpr int W = 2 * BW + 1;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (idx >= nlm) return;
    const int nu_of[5] = {0, 0, 1, 1, 2};
    const int k = nu_of[f];
    const int l = l_of_idx[idx];
    const size_t mat = (size_t)k * W * n * nl;
    const double *Lk = Lt + mat, *LUk = LUt + mat;
    const size_t fs = (size_t)nlm * n;
    double *ure = st + (size_t)(2 * f) * fs, *uim = st + (size_t)(2 * f + 1) * fs;
    const double *nre = nl_arr + (size_t)(2 * f) * fs, *nim = nl_arr + (size_t)(2 * f + 1) * fs;
    const double *pre = np_arr + (size_t)(2 * f) * fs, *pim = np_arr + (size_t)(2 * f + 1) * fs;
    const bool tbc = (f == 4);
    const bool bcr = (bc_mask >> f) & 1; // Dirichlet rows in this field's operator
    const double inv_dt = 1.0 / dt, lw = 1.0 - theta;
    double wr[W], wi[W];   // window of original u, rows i-BW..i+BW
    double yr[BW], yi[BW]; // last BW forward-substituted values
#pragma unroll
    for (int q = 0; q < W; ++q) {
        int j = q - BW;
        bool ok = j >= 0 && j < n && !(tbc && (j == 0 || j == n - 1));
        wr[q] = ok ? ure[idx + (size_t)nlm * j] : 0.0;
        wi[q] = ok ? uim[idx + (size_t)nlm * j] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < BW; ++q) yr[q] = yi[q] = 0.0;
    for (int i = 0; i < n; ++i) {
        double lr = 0.0, li = 0.0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int j = i + q - BW;
            if (j >= 0 && j < n) {
                double c = Lk[((size_t)j * W + (BW + i - j)) * nl + l];
                lr += c * wr[q];
                li += c * wi[q];
            }
        }
        const size_t oi = idx + (size_t)nlm * i;
        double vr = inv_dt * wr[BW] + 1.5 * nre[oi] - 0.5 * pre[oi];
        double vi = inv_dt * wi[BW] + 1.5 * nim[oi] - 0.5 * pim[oi];
        if (lw != 0.0) {
            vr += lw * lr;
            vi += lw * li;
        }
        if (bcr && (i == 0 || i == n - 1)) vr = vi = 0.0; // Dirichlet value
        // forward substitution with the unit-lower factor
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            int j = i - BW + q;
            if (j >= 0) {
                double c = LUk[((size_t)j * W + (BW + i - j)) * nl + l];
                vr -= c * yr[q];
                vi -= c * yi[q];
            }
        }
#pragma unroll
        for (int q = 0; q < BW - 1; ++q) {
            yr[q] = yr[q + 1];
            yi[q] = yi[q + 1];
        }
        yr[BW - 1] = vr;
        yi[BW - 1] = vi;
#pragma unroll
        for (int q = 0; q < W - 1; ++q) {
            wr[q] = wr[q + 1];
            wi[q] = wi[q + 1];
        }
        int jn = i + BW + 1;
        bool ok = jn < n && !(tbc && jn == n - 1);
        wr[W - 1] = ok ? ure[idx + (size_t)nlm * jn] : 0.0;
        wi[W - 1] = ok ? uim[idx + (size_t)nlm * jn] : 0.0;
        ure[oi] = vr;
        uim[oi] = vi;
    }
    // back substitution
    double xr[BW], xi[BW];
#pragma unroll
    for (int q = 0; q < BW; ++q) xr[q] = xi[q] = 0.0;
    for (int i = n - 1; i >= 0; --i) {
        const size_t oi = idx + (size_t)nlm * i;
        double vr = ure[oi], vi = uim[oi];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            int j = i + 1 + q;
            if (j < n) {
                double c = LUk[((size_t)j * W + (BW + i - j)) * nl + l];
                vr -= c * xr[q];
                vi -= c * xi[q];
            }
        }
        double d = LUk[((size_t)i * W + BW) * nl + l];
        vr /= d;
        vi /= d;
#pragma unroll
        for (int q = BW - 1; q > 0; --q) {
            xr[q] = xr[q - 1];
            xi[q] = xi[q - 1];
        }
        xr[0] = vr;
        xi[0] = vi;
        ure[oi] = vr;
        uim[oi] = vi;
    }
}

int gd_launch_cnab2_step(hipStream_t st, const double *Lt, const double *LUt, const int *l_of_idx, int nl, int n,
                         int bw, int nlm, double dt, double theta, int bc_mask, double *state, const double *nl_arr,
                         const double *np_arr)
{
    dim3 grid((nlm + 255) / 256, 5);
    switch (bw) {
#define CASE(B)                                                                                             \
    case B:                                                                                                 \
        hipLaunchKernelGGL(cnab2_step_kernel<B>, grid, dim3(256), 0, st, Lt, LUt, l_of_idx, nl, n, nlm, dt, \
                           theta, bc_mask, state, nl_arr, np_arr);                                          \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        gd_set_error("bw=%d unsupported (1..7)", bw);
        return GD_ERR_ARG;
    }
    return gd_check_launch("cnab2_step_kernel");
}

// ---------------------------------------------------------------- FP64 reductions
// energies over an m-major state: out[0] KE, out[1] ME, out[2] thermal
__global__ void __launch_bounds__(256) energy_partial_kernel(const double *__restrict__ st,
                                                             const int *__restrict__ l_of_idx, int nlm, int n,
                                                             double *__restrict__ part)
{
    __shared__ double red[3][256];
    double acc[3] = {0.0, 0.0, 0.0};
    const size_t fs = (size_t)nlm * n;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < fs; q += (size_t)gridDim.x * blockDim.x) {
        int idx = (int)(q % nlm);
        double lf = (double)l_of_idx[idx] * (l_of_idx[idx] + 1);
        double wgt = 1.0 / (lf > 1.0 ? lf : 1.0);
        for (int pair = 0; pair < 2; ++pair) {
            double s = 0.0;
            for (int f = 2 * pair; f < 2 * pair + 2; ++f) {
                double a = st[(size_t)(2 * f) * fs + q], b = st[(size_t)(2 * f + 1) * fs + q];
                s += a * a + b * b;
            }
            acc[pair] += wgt * s;
        }
        double a = st[(size_t)8 * fs + q], b = st[(size_t)9 * fs + q];
        acc[2] += a * a + b * b;
    }
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[(size_t)blockIdx.x * 3 + k] = red[k][0];
}

__global__ void __launch_bounds__(256) sum_partials_kernel(const double *__restrict__ part, int nblk, int ncomp,
                                                           double scale, int do_sqrt, double *__restrict__ out)
{
    __shared__ double red[256];
    for (int k = 0; k < ncomp; ++k) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256) a += part[(size_t)b * ncomp + k];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = do_sqrt ? sqrt(scale * red[0]) : scale * red[0];
        __syncthreads();
    }
}

int gd_launch_energies(hipStream_t st, const double *state, const int *l_of_idx, int nlm, int n, double *part,
                       int nblk, double *out3)
{
    hipLaunchKernelGGL(energy_partial_kernel, dim3(nblk), dim3(256), 0, st, state, l_of_idx, nlm, n, part);
    GD_TRY(gd_check_launch("energy_partial_kernel"));
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part, nblk, 3, 0.5, 0, out3);
    return gd_check_launch("sum_partials_kernel");
}

__global__ void __launch_bounds__(256) sqdiff_partial_kernel(long n, const double *__restrict__ ar,
                                                             const double *__restrict__ ai,
                                                             const double *__restrict__ br,
                                                             const double *__restrict__ bi, double *__restrict__ part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double dr = ar[q] - br[q], di = ai[q] - bi[q];
        acc += dr * dr + di * di;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int gd_launch_sqdiff(hipStream_t st, long n, const double *ar, const double *ai, const double *br, const double *bi,
                     double *part, int nblk, double *out1)
{
    hipLaunchKernelGGL(sqdiff_partial_kernel, dim3(nblk), dim3(256), 0, st, n, ar, ai, br, bi, part);
    GD_TRY(gd_check_launch("sqdiff_partial_kernel"));
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part, nblk, 1, 1.0, 1, out1);
    return gd_check_launch("sum_partials_kernel");
}

int gd_launch_band_solve(hipStream_t st, const double *lu, const int *sys_of_rhs, int n, int bw, int nsys,
                         int nrhs, double *x, int *d_status)
{
    hipLaunchKernelGGL(band_solve_kernel, dim3((nrhs + 63) / 64), dim3(64), 0, st, lu, sys_of_rhs, n, bw, nsys,
                       nrhs, x, d_status);
    return gd_check_launch("band_solve_kernel");
}
