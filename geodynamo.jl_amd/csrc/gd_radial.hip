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
// One thread per (device idx, field): the re and im columns share the band
// values.  Inputs are read from st_in and the result written to st_out
// (distinct buffers), so the loads of a block of D levels are independent of
// the stores and are issued a block ahead (software pipelining; a column is a
// sequential recurrence in r and a wavefront holds few columns, so memory
// parallelism must come from within the thread).
//   explicit part: L_l u = nu Lap u - nu l(l+1)/r^2 u; nu*Lap is the same for
//                  every l and is read with uniform (scalar) loads
//   LU(S_l):       from the l-fastest transposed table LU_t[(col*W + row)*nl + l]
// nu index of field f: {0,0,1,1,2}.  Field 4 (temperature) has its boundary
// rows set to 0 before the rhs (thermal.jl:262-313) unless bc_mask bit 5 says
// a boundary-condition stage already set them; fields in bc_mask have
// Dirichlet rows (rhs rows 0 and n-1 := 0, the stored value for such a
// temperature).
namespace {
template <int BW, int D> struct FwdBlk {
    double ur[D], ui[D], nr_[D], ni[D], pr[D], pi[D], fl[D][BW];
};
template <int BW, int D> struct BwdBlk {
    double yr[D], yi[D], bu[D][BW], dg[D];
};
} // namespace

template <int BW, int D>
__global__ void __launch_bounds__(256) cnab2_step_kernel(const double *__restrict__ nuLap, const double *__restrict__ rinv2,
                                                         const double *__restrict__ nus, const double *__restrict__ LUt,
                                                         const int *__restrict__ l_of_idx, int nl, int n, int nlm,
                                                         double dt, double theta, int bc_mask,
                                                         const double *__restrict__ st_in, double *__restrict__ st_out,
                                                         const double *__restrict__ nl_arr,
                                                         const double *__restrict__ np_arr)
{
    constexpr int W = 2 * BW + 1;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (idx >= nlm) return;
    const int k = f < 2 ? 0 : (f < 4 ? 1 : 2);
    const int l = l_of_idx[idx];
    const double nulf = nus[k] * ((double)l * (double)(l + 1));
    const double *Lap = nuLap + (size_t)k * W * n;
    const double *LU = LUt + (size_t)k * W * n * nl + l;
    const size_t fs = (size_t)nlm * n;
    const double *ure = st_in + (size_t)(2 * f) * fs + idx, *uim = ure + fs;
    double *ore = st_out + (size_t)(2 * f) * fs + idx, *oim = ore + fs;
    const double *nre = nl_arr + (size_t)(2 * f) * fs + idx, *nim = nre + fs;
    const double *pre = np_arr + (size_t)(2 * f) * fs + idx, *pim = pre + fs;
    // bc_mask bit 5: temperature rows already set by a boundary-condition
    // stage (read as stored; Dirichlet rhs rows take the stored value)
    const bool tkeep = (f == 4) && ((bc_mask >> 5) & 1);
    const bool tbc = (f == 4) && !tkeep;
    const bool bcr = (bc_mask >> f) & 1;
    const double inv_dt = 1.0 / dt, lw = 1.0 - theta;
    auto urow = [&](int j) { return j >= 0 && j < n && !(tbc && (j == 0 || j == n - 1)); };

    double wr[W], wi[W]; // window of (boundary-adjusted) u, rows i-BW..i+BW
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int j = q - BW;
        wr[q] = urow(j) ? ure[(size_t)nlm * j] : 0.0;
        wi[q] = urow(j) ? uim[(size_t)nlm * j] : 0.0;
    }
    double yr[BW], yi[BW];
#pragma unroll
    for (int q = 0; q < BW; ++q) yr[q] = yi[q] = 0.0;

    auto fload = [&](FwdBlk<BW, D> &b, int i0) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 + d, jn = i + BW + 1;
            b.ur[d] = urow(jn) ? ure[(size_t)nlm * jn] : 0.0;
            b.ui[d] = urow(jn) ? uim[(size_t)nlm * jn] : 0.0;
            if (i < n) {
                b.nr_[d] = nre[(size_t)nlm * i];
                b.ni[d] = nim[(size_t)nlm * i];
                b.pr[d] = pre[(size_t)nlm * i];
                b.pi[d] = pim[(size_t)nlm * i];
#pragma unroll
                for (int q = 0; q < BW; ++q) {
                    const int j = i - BW + q;
                    b.fl[d][q] = j >= 0 ? LU[((size_t)j * W + (BW + i - j)) * nl] : 0.0;
                }
            }
        }
    };
    auto fcompute = [&](const FwdBlk<BW, D> &b, int i0) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 + d;
            if (i >= n) break;
            double lr = 0.0, li = 0.0;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                const int j = i + q - BW;
                if (j >= 0 && j < n) {
                    const double c = (q == BW) ? Lap[BW + W * i] - nulf * rinv2[i] : Lap[(BW + i - j) + W * j];
                    lr += c * wr[q];
                    li += c * wi[q];
                }
            }
            double vr = inv_dt * wr[BW] + 1.5 * b.nr_[d] - 0.5 * b.pr[d];
            double vi = inv_dt * wi[BW] + 1.5 * b.ni[d] - 0.5 * b.pi[d];
            if (lw != 0.0) {
                vr += lw * lr;
                vi += lw * li;
            }
            if (bcr && (i == 0 || i == n - 1)) { // Dirichlet value
                vr = tkeep ? wr[BW] : 0.0;
                vi = tkeep ? wi[BW] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < BW; ++q) {
                vr -= b.fl[d][q] * yr[q];
                vi -= b.fl[d][q] * yi[q];
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
            wr[W - 1] = b.ur[d];
            wi[W - 1] = b.ui[d];
            ore[(size_t)nlm * i] = vr;
            oim[(size_t)nlm * i] = vi;
        }
    };
    {
        FwdBlk<BW, D> A, B;
        fload(A, 0);
        for (int i0 = 0; i0 < n; i0 += 2 * D) {
            fload(B, i0 + D);
            fcompute(A, i0);
            fload(A, i0 + 2 * D);
            fcompute(B, i0 + D);
        }
    }

    // back substitution, blocks of D rows going up
    double xr[BW], xi[BW];
#pragma unroll
    for (int q = 0; q < BW; ++q) xr[q] = xi[q] = 0.0;
    auto bload = [&](BwdBlk<BW, D> &b, int i0) { // rows i0, i0-1, ..., i0-D+1
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 - d;
            if (i >= 0) {
                b.yr[d] = ore[(size_t)nlm * i];
                b.yi[d] = oim[(size_t)nlm * i];
                b.dg[d] = LU[((size_t)i * W + BW) * nl];
#pragma unroll
                for (int q = 0; q < BW; ++q) {
                    const int j = i + 1 + q;
                    b.bu[d][q] = j < n ? LU[((size_t)j * W + (BW + i - j)) * nl] : 0.0;
                }
            }
        }
    };
    auto bcompute = [&](const BwdBlk<BW, D> &b, int i0) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 - d;
            if (i < 0) break;
            double vr = b.yr[d], vi = b.yi[d];
#pragma unroll
            for (int q = 0; q < BW; ++q) {
                vr -= b.bu[d][q] * xr[q];
                vi -= b.bu[d][q] * xi[q];
            }
            vr /= b.dg[d];
            vi /= b.dg[d];
#pragma unroll
            for (int q = BW - 1; q > 0; --q) {
                xr[q] = xr[q - 1];
                xi[q] = xi[q - 1];
            }
            xr[0] = vr;
            xi[0] = vi;
            ore[(size_t)nlm * i] = vr;
            oim[(size_t)nlm * i] = vi;
        }
    };
    {
        BwdBlk<BW, D> A, B;
        bload(A, n - 1);
        for (int i0 = n - 1; i0 >= 0; i0 -= 2 * D) {
            bload(B, i0 - D);
            bcompute(A, i0);
            bload(A, i0 - 2 * D);
            bcompute(B, i0 - D);
        }
    }
}

int gd_launch_cnab2_step(hipStream_t st, const double *nuLap, const double *rinv2, const double *nus,
                         const double *LUt, const int *l_of_idx, int nl, int n, int bw, int nlm, double dt,
                         double theta, int bc_mask, const double *st_in, double *st_out, const double *nl_arr,
                         const double *np_arr)
{
    GD_CHECK(st_in != st_out, GD_ERR_ARG, "cnab2_step: in and out state must differ");
    dim3 grid((nlm + 255) / 256, 5);
    switch (bw) {
#define CASE(B)                                                                                                \
    case B:                                                                                                    \
        hipLaunchKernelGGL((cnab2_step_kernel<B, 4>), grid, dim3(256), 0, st, nuLap, rinv2, nus, LUt, l_of_idx, \
                           nl, n, nlm, dt, theta, bc_mask, st_in, st_out, nl_arr, np_arr);                      \
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

// ---------------------------------------------------------------- plain device copy
// 16-byte vector copy (grid-stride); also the known-size reference kernel for
// HBM byte counters.
__global__ void __launch_bounds__(256) copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n16; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = src[q];
}

extern "C" int gd_device_copy(gd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    GD_CHECK(ctx && dst && src && bytes % 16 == 0, GD_ERR_ARG, "gd_device_copy: bytes must be a multiple of 16");
    GD_HIP(hipSetDevice(ctx->device));
    const size_t n16 = bytes / 16;
    const unsigned grid = (unsigned)((n16 + 255) / 256 < 65536 ? (n16 + 255) / 256 : 65536);
    if (n16) hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const uint4 *)src, (uint4 *)dst, n16);
    return gd_check_launch("copy_kernel");
}

extern "C" int gd_device_alloc(gd_ctx *ctx, size_t bytes, void **out)
{
    GD_CHECK(ctx && out, GD_ERR_ARG, "gd_device_alloc: null argument");
    GD_HIP(hipSetDevice(ctx->device));
    GD_HIP(hipMalloc(out, bytes));
    return GD_OK;
}

extern "C" int gd_device_free(gd_ctx *ctx, void *p)
{
    GD_CHECK(ctx, GD_ERR_ARG, "gd_device_free: null ctx");
    GD_HIP(hipSetDevice(ctx->device));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    GD_HIP(hipFree(p));
    return GD_OK;
}
