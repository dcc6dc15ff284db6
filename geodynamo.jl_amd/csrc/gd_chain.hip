// gd_chain.hip -- per-column CNAB2 kernel of the step (FP64 rhs and banded
// solve, timestep.jl:610-730, fused with the next step's spectral
// preparation).
#include <type_traits>

#include "gd_radial.hpp"

// ---------------------------------------------------------------- per-column CNAB2 for the step
// One thread per (device idx, field, re/im) column: 10 columns per idx give
// twice the wavefronts of a thread per (idx, field).  A column is a
// recurrence in r, so memory parallelism comes from within the thread: the
// loop runs over blocks of D levels and, while level i is computed, the loads
// of level i + D are issued into the slot level i has just freed.  Loads read
// clamped (valid) rows and are masked by selects instead of branches, so all
// of them stay in flight across the block.  Same arithmetic, in the same
// order, as cnab2_step_kernel; the forward-substitution result y is kept in
// st_out and overwritten by x in the back substitution.
// With c.synV set, the back substitution also writes the next step's
// Legendre-synthesis columns of this field (the per-field part of
// spectral_pre_kernel in gd_step.hip: radial derivatives by the D1/D2 bands,
// curl and radial-component factors): when x[i] is known, row i + BW has its
// whole band window x[i..i+2BW], so the state is not read again.
namespace {
template <int BW, int D, typename TS, bool DIRECT> __global__ void __launch_bounds__(64) cnab2_chain_kernel(ChainArgs c)
{
    constexpr int W = 2 * BW + 1;
    const int n = c.n, nlm = c.nlm;
    const int idx = blockIdx.x * 64 + threadIdx.x;
    // field, 0 re / 1 im; field 5 is the composition: laid out and treated
    // as the temperature (field 4) of its own buffers, diffusivity slot 3,
    // boundary rows preset by its condition stage
    const int fq = c.f0 + blockIdx.y, ri = blockIdx.z;
    const bool isC = fq == 5;
    const int f = isC ? 4 : fq;
    const int k = isC ? 3 : (f < 2 ? 0 : (f < 4 ? 1 : 2));
    // level-uniform tables in LDS: the explicit operator by row (ER doubles per
    // row i: the coefficient of u row i-BW+q in slot q, 0 where that row does
    // not exist, and 1/r_i^2 in slot W), D1 and D2 bands, 1/r, 1/r^2
    constexpr int ER = (W + 2) & ~1;
    extern __shared__ double smem[];
    double *Ex = smem, *sD1 = Ex + ER * n, *sD2 = sD1 + W * n, *rinv_ = sD2 + W * n, *rinv2_ = rinv_ + n;
    const double *Lapk = c.nuLap + (size_t)k * W * n;
    if constexpr (!DIRECT) // the radial derivatives serve only the general synthesis columns
        for (int t = threadIdx.x; t < W * n; t += 64) {
            sD1[t] = c.D1[t];
            sD2[t] = c.D2[t];
        }
    for (int t = threadIdx.x; t < ER * n; t += 64) {
        const int i = t / ER, q = t - i * ER, j = i - BW + q;
        double v = 0.0;
        if (q == BW) v = Lapk[BW + W * i];
        else if (q < W && j >= 0 && j < n) v = Lapk[(2 * BW - q) + W * j];
        else if (q == W) v = c.rtab[i + n];
        Ex[t] = v;
    }
    for (int t = threadIdx.x; t < n; t += 64) {
        rinv_[t] = c.rtab[t + 2 * n];
        rinv2_[t] = c.rtab[t + n];
    }
    __syncthreads();
    if (idx >= nlm) return;
    const int l = c.l_of_idx[idx];
    const double lf = (double)l * (double)(l + 1);
    const double nulf = c.nus[k] * lf;
    const double *LU = c.LUt + (size_t)k * W * n * c.nl + l;
    const double *rdg = c.rdiag + (size_t)k * n * c.nl + l;
    const size_t fs = (size_t)nlm * n, off = (size_t)(2 * f + ri) * fs + idx;
    const double *u = (isC ? c.c_in : c.st_in) + off;
    // a field whose nonlinear terms are known to be zero reads one zero row (L2-resident)
    const bool nlz = !isC && ((c.nl_zero_mask >> f) & 1);
    const double *pn = nlz ? c.zero_row + idx : (isC ? c.c_nl : c.nl_arr) + off;
    const double *pp = nlz ? c.zero_row + idx : (isC ? c.c_np : c.np_arr) + off;
    const size_t nstride = nlz ? 0 : (size_t)nlm;
    double *o = (isC ? c.c_out : c.st_out) + off;
    // bit 5 of bc_mask: a boundary-condition stage has already set the
    // temperature rows in st_in; read them as stored, and Dirichlet rhs rows
    // of the temperature take the stored boundary value
    const bool tkeep = isC || ((f == 4) && ((c.bc_mask >> 5) & 1));
    const bool tbc = (f == 4) && !tkeep;
    const bool bcr = isC ? c.c_bcr != 0 : ((c.bc_mask >> f) & 1);
    const double inv_dt = 1.0 / c.dt, lw = 1.0 - c.theta;
    auto cl = [&](int j) { return j < 0 ? 0 : (j > n - 1 ? n - 1 : j); }; // a valid row
    auto urow = [&](int j) { return j >= 0 && j < n && !(tbc && (j == 0 || j == n - 1)); };
    // band element (row i, column j), |i - j| <= BW, of the transposed LU table at valid rows
    auto lu = [&](int i, int j) { return LU[((size_t)cl(j) * W + (BW + i - j)) * c.nl]; };

    // ---- forward: rhs row and L-substitution
    double w[W]; // u rows i-BW..i+BW (temperature boundary rows read as 0)
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int j = q - BW;
        const double v = u[(size_t)nlm * cl(j)];
        w[q] = urow(j) ? v : 0.0;
    }
    double su[D], sn[D], sp[D], sl[D][BW];
    auto fload = [&](int s, int i) {
        const int jn = i + BW + 1, ic = cl(i);
        const double v = u[(size_t)nlm * cl(jn)];
        su[s] = urow(jn) ? v : 0.0;
        sn[s] = pn[nstride * ic];
        sp[s] = pp[nstride * ic];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            const int j = i - BW + q;
            const double e = lu(ic, ic - BW + q);
            sl[s][q] = j >= 0 ? e : 0.0;
        }
    };
    double yw[BW]; // y rows i-BW..i-1
#pragma unroll
    for (int q = 0; q < BW; ++q) yw[q] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) fload(d, d);
    for (int i0 = 0; i0 < n; i0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 + d, ic = cl(i);
            const double *er = Ex + ER * ic;
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < W; ++q) lr += ((q == BW) ? er[BW] - nulf * er[W] : er[q]) * w[q];
            double v = inv_dt * w[BW] + 1.5 * sn[d] - 0.5 * sp[d];
            v += lw * lr;
            v = (bcr && (i == 0 || i == n - 1)) ? (tkeep ? w[BW] : 0.0) : v; // Dirichlet value
#pragma unroll
            for (int q = 0; q < BW; ++q) v -= sl[d][q] * yw[q];
#pragma unroll
            for (int q = 0; q < BW - 1; ++q) yw[q] = yw[q + 1];
            yw[BW - 1] = v;
#pragma unroll
            for (int q = 0; q < W - 1; ++q) w[q] = w[q + 1];
            w[W - 1] = su[d];
            if (i < n) o[(size_t)nlm * i] = v;
            fload(d, i + D);
        }
    }

    // ---- back substitution (and the synthesis columns of row i + BW)
    const bool syn = c.synV != nullptr;
    static constexpr int colA[5] = {2, 0, 10, 8, 16}, colB[5] = {4, 6, 12, 14, 0}, colS[5] = {2, 0, 6, 4, 8};
    const bool pol = (f == 1 || f == 3);
    const bool wA = !c.syn_tang || f == 0 || f == 1 || f == 4;
    const bool wB = f < 4 && !c.syn_tang, wS = !c.syn_tang && ((f == 4) || c.radials);
    const bool need1 = (f == 4) ? wS : (pol && wB), need2 = pol && wB; // radial derivatives used
    TS *pA = nullptr, *pB = nullptr, *pS = nullptr;
    if (syn) {
        TS *sv = static_cast<TS *>(c.synV), *ss = static_cast<TS *>(c.synS);
        pA = sv + (size_t)((isC ? SYNV_COLS - 2 : colA[f]) + ri) * nlm + idx;
        pB = sv + (size_t)(colB[f] + ri) * nlm + idx;
        pS = ss + (size_t)(colS[f] + ri) * nlm + idx;
    }
    double by[D], bd[D], bu[D][BW];
    auto bload = [&](int s, int i) {
        const int ic = cl(i);
        by[s] = o[(size_t)nlm * ic];
        bd[s] = rdg[(size_t)ic * c.nl];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            const int j = i + 1 + q;
            const double e = lu(ic, ic + 1 + q > n - 1 ? ic : ic + 1 + q);
            bu[s][q] = (j < n) ? e : 0.0;
        }
    };
    double xw[W]; // x rows i..i+2BW after the shift (rows outside 0..n-1 are 0)
#pragma unroll
    for (int q = 0; q < W; ++q) xw[q] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) bload(d, n - 1 - d);
    // the tangential step needs no radial derivatives: its columns (u itself,
    // T/r) are written as x is produced; otherwise they trail the solve by BW
    // DIRECT (compile-time: the launcher sets it exactly when synV is set and
    // syn_tang is 1) keeps the branch out of the software-pipelined loop
    const int iend = (syn && !DIRECT) ? -BW : 0;
    for (int i0 = n - 1; i0 >= iend; i0 -= D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 - d;
            double v = by[d];
#pragma unroll
            for (int q = 0; q < BW; ++q) v -= bu[d][q] * xw[q];
            v *= bd[d]; // reciprocal of the U diagonal
            v = i >= 0 ? v : 0.0;
#pragma unroll
            for (int q = W - 1; q > 0; --q) xw[q] = xw[q - 1];
            xw[0] = v;
            if (i >= 0) o[(size_t)nlm * i] = v;
            if constexpr (DIRECT) {
                if (wA && i >= 0) pA[(size_t)i * SYNV_COLS * nlm] = (TS)(f == 4 ? v * rinv_[i] : v);
            } else {
            const int r = i + BW; // xw[q] = x[r - BW + q]
            if (syn && wA && r >= 0 && r < n) {
                const double rinv = rinv_[r], rinv2 = rinv2_[r];
                const double x = xw[BW];
                double d1 = 0.0, d2 = 0.0;
                if (need1) { // uniform
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        const int j = r - BW + q;
                        const bool ok = j >= 0 && j < n;
                        d1 += (ok ? sD1[(2 * BW - q) + W * cl(j)] : 0.0) * xw[q];
                    }
                }
                if (need2) {
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        const int j = r - BW + q;
                        const bool ok = j >= 0 && j < n;
                        d2 += (ok ? sD2[(2 * BW - q) + W * cl(j)] : 0.0) * xw[q];
                    }
                }
                double a, bv, sv;
                if (f == 4) {
                    a = x * rinv;
                    bv = 0.0;
                    sv = d1;
                } else if (pol) {
                    a = x;
                    bv = lf * rinv2 * x - d2 - 2.0 * rinv * d1;
                    sv = lf * x * rinv2;
                } else {
                    a = x;
                    bv = -lf * rinv2 * x;
                    sv = lf * bv * rinv2;
                }
                const size_t ov = (size_t)r * SYNV_COLS * nlm, os = (size_t)r * 10 * nlm;
                pA[ov] = (TS)a;
                if (wB) pB[ov] = (TS)bv;
                if (wS) pS[os] = (TS)sv;
            }
            }
            bload(d, i - D);
        }
    }
}
} // namespace

__global__ void recip_diag_kernel(const double *__restrict__ LUt, int nk, int n, int bw, int nl, double *__restrict__ rd)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nk * n * nl) return;
    const int l = e % nl, i = (e / nl) % n, k = e / (nl * n), W = 2 * bw + 1;
    rd[e] = 1.0 / LUt[((size_t)k * W * n + (size_t)i * W + bw) * nl + l];
}

int gd_launch_recip_diag(hipStream_t st, const double *LUt, int nk, int n, int bw, int nl, double *rdiag)
{
    const int total = nk * n * nl;
    hipLaunchKernelGGL(recip_diag_kernel, dim3((total + 255) / 256), dim3(256), 0, st, LUt, nk, n, bw, nl, rdiag);
    return gd_check_launch("recip_diag_kernel");
}

int gd_cnab2_chain_supported(int n, int bw)
{
    return bw >= 1 && bw <= 7 && n >= 2 &&
           sizeof(double) * (size_t)((((2 * bw + 1) + 2) & ~1) + 2 * (2 * bw + 1) + 2) * n <= 64 * 1024;
}

int gd_launch_cnab2_chain(hipStream_t st, const ChainArgs &c, bool syn_fp64)
{
    GD_CHECK(c.st_in != c.st_out, GD_ERR_ARG, "cnab2_chain: in and out state must differ");
    GD_CHECK(c.rdiag, GD_ERR_ARG, "cnab2_chain: reciprocal diagonal table missing");
    GD_CHECK(!c.nl_zero_mask || c.zero_row, GD_ERR_ARG, "cnab2_chain: nl_zero_mask needs zero_row");
    GD_CHECK(!c.synV || (c.synS && c.D1 && c.D2 && c.rtab), GD_ERR_ARG, "cnab2_chain: incomplete synthesis outputs");
    GD_CHECK(gd_cnab2_chain_supported(c.n, c.bw), GD_ERR_ARG, "cnab2_chain: n=%d bw=%d unsupported", c.n, c.bw);
    const int nfield = c.nf ? c.nf : (c.c_in ? 6 : 5);
    GD_CHECK(c.f0 >= 0 && c.nf >= 0 && c.f0 + nfield <= (c.c_in ? 6 : 5), GD_ERR_ARG, "cnab2_chain: fields %d+%d",
             c.f0, nfield);
    GD_CHECK(!c.c_in || (c.c_out && c.c_nl && c.c_np && (!c.synV || c.syn_tang)), GD_ERR_ARG,
             "cnab2_chain: incomplete composition buffers");
    dim3 grid((c.nlm + 63) / 64, nfield, 2);
    const size_t lds = sizeof(double) * (size_t)((((2 * c.bw + 1) + 2) & ~1) + 2 * (2 * c.bw + 1) + 2) * c.n;
    GD_CHECK(lds <= 64 * 1024, GD_ERR_ARG, "cnab2_chain: n=%d bw=%d needs %zu B of LDS", c.n, c.bw, lds);
    const bool direct = c.synV && c.syn_tang;
#define GD_CHAIN(B)                                                                                              \
    if (c.bw == B) {                                                                                             \
        if (syn_fp64 && direct) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, double, true>), grid, dim3(64), lds, st, c); \
        else if (syn_fp64) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, double, false>), grid, dim3(64), lds, st, c); \
        else if (direct) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, float, true>), grid, dim3(64), lds, st, c); \
        else hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, float, false>), grid, dim3(64), lds, st, c);           \
        return gd_check_launch("cnab2_chain_kernel");                                                            \
    }
    GD_CHAIN(1) GD_CHAIN(2) GD_CHAIN(3) GD_CHAIN(4) GD_CHAIN(5) GD_CHAIN(6) GD_CHAIN(7)
#undef GD_CHAIN
    gd_set_error("cnab2_chain: no kernel for n=%d bw=%d", c.n, c.bw);
    return GD_ERR_ARG;
}

