// gd_chain.hip -- per-column CNAB2 kernel of the step (FP64 rhs and banded
// solve, timestep.jl:610-730, fused with the next step's spectral
// preparation).
#include <algorithm>
#include <cmath>
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
// IX: the type of the row, synthesis-column and table offsets from the
// per-column base pointers -- unsigned where they fit 32 bits (each offset is
// then one scalar multiply instead of a 64-bit product), size_t otherwise.
// (blk_x, blk_y, blk_z): the block's column block, field offset and re/im (the
// launch's own grid, or its share of a batched launch)
template <int BW, int D, typename TS, bool DIRECT, typename IX, class AT>
__device__ __forceinline__ void cnab2_chain_body(const AT &c, int blk_x, int blk_y, int blk_z)
{
    constexpr int W = 2 * BW + 1;
    const int n = c.n, nlm = c.nlm;
    const int idx = blk_x * 64 + threadIdx.x;
    // field, 0 re / 1 im; field 5 is the composition: laid out and treated
    // as the temperature (field 4) of its own buffers, diffusivity slot 3,
    // boundary rows preset by its condition stage
    const int fq = c.f0 + blk_y, ri = blk_z;
    const bool isC = fq == 5;
    const int f = isC ? 4 : fq;
    const int k = isC ? 3 : (f < 2 ? 0 : (f < 4 ? 1 : 2));
    // level-uniform tables in LDS: the explicit operator by row (ER doubles per
    // row i: the coefficient of u row i-BW+q in slot q, 0 where that row does
    // not exist, and 1/r_i^2 in slot W), D1 and D2 bands, 1/r, 1/r^2
    constexpr int ER = (W + 2) & ~1;
    extern __shared__ double smem[];
    double *Ex = smem, *sD1 = Ex + ER * n, *sD2 = sD1 + W * n, *rinv_ = sD2 + W * n, *rinv2_ = rinv_ + n;
    // The prologue issues its loads together (the lane's column clamped to a
    // valid one until the range check below): values that a condition would
    // zero are read from the zero row instead (the load address is selected,
    // never the loaded value), because a select of a loaded value on a
    // wave-uniform condition becomes a branch around the load and a wait for
    // it, one memory latency per load.
    const int idc = idx < nlm ? idx : nlm - 1;
    const int l = c.l_of_idx[idc];
    const size_t fs = (size_t)nlm * n, off = (size_t)(2 * f + ri) * fs + idc;
    const double *u = (isC ? c.c_in : c.st_in) + off;
    // bit 5 of bc_mask: a boundary-condition stage has already set the
    // temperature rows in st_in; read them as stored, and Dirichlet rhs rows
    // of the temperature take the stored boundary value
    const bool tkeep = isC || ((f == 4) && ((c.bc_mask >> 5) & 1));
    const bool tbc = (f == 4) && !tkeep;
    auto cl = [&](int j) { return j < 0 ? 0 : (j > n - 1 ? n - 1 : j); }; // a valid row
    auto urow = [&](int j) { return j >= 0 && j < n && !(tbc && (j == 0 || j == n - 1)); };
    double w[W]; // u rows i-BW..i+BW of the first forward step (temperature boundary rows read as 0)
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const int j = q - BW;
        w[q] = *(urow(j) ? u + IX(nlm) * IX(j) : c.zero_row + idc);
    }
    const double *Lapk = c.nuLap + (size_t)k * W * n;
    constexpr int UF = 4; // table entries per thread and round, loads before stores
    if constexpr (!DIRECT) // the radial derivatives serve only the general synthesis columns
        for (int t0 = threadIdx.x; t0 < W * n; t0 += UF * 64) {
            double v1[UF], v2[UF];
#pragma unroll
            for (int e = 0; e < UF; ++e) {
                const int t = min(t0 + e * 64, W * n - 1);
                v1[e] = c.D1[t];
                v2[e] = c.D2[t];
            }
#pragma unroll
            for (int e = 0; e < UF; ++e) {
                const int t = min(t0 + e * 64, W * n - 1); // past the end: entry W n - 1 again
                sD1[t] = v1[e];
                sD2[t] = v2[e];
            }
        }
    for (int t0 = threadIdx.x; t0 < ER * n; t0 += UF * 64) {
        double v[UF];
#pragma unroll
        for (int e = 0; e < UF; ++e) {
            const int t = min(t0 + e * 64, ER * n - 1), i = t / ER, q = t - i * ER, j = i - BW + q;
            const bool bnd = q < W && j >= 0 && j < n; // q = BW: the diagonal, j = i
            v[e] = *(bnd ? Lapk + (2 * BW - q) + W * j : (q == W ? c.rtab + i + n : c.zero_row));
        }
#pragma unroll
        for (int e = 0; e < UF; ++e) Ex[min(t0 + e * 64, ER * n - 1)] = v[e]; // past the end: the last entry again
    }
    for (int t = threadIdx.x; t < n; t += 64) {
        rinv_[t] = c.rtab[t + 2 * n];
        rinv2_[t] = c.rtab[t + n];
    }
    __syncthreads();
    if (idx >= nlm) return;
    const double lf = (double)l * (double)(l + 1);
    const double nulf = c.nus[k] * lf;
    const double *LU = c.LUt + (size_t)k * W * n * c.nl + l;
    const double *rdg = c.rdiag + (size_t)k * n * c.nl + l;
    // a field whose nonlinear terms are known to be zero reads one zero row (L2-resident)
    const bool nlz = !isC && ((c.nl_zero_mask >> f) & 1);
    const double *pn = nlz ? c.zero_row + idx : (isC ? c.c_nl : c.nl_arr) + off;
    const double *pp = nlz ? c.zero_row + idx : (isC ? c.c_np : c.np_arr) + off;
    const IX nstride = nlz ? IX(0) : IX(nlm);
    double *o = (isC ? c.c_out : c.st_out) + off;
    const bool bcr = isC ? c.c_bcr != 0 : ((c.bc_mask >> f) & 1);
    const double inv_dt = 1.0 / c.dt, lw = 1.0 - c.theta;
    // band element (row i, column j), |i - j| <= BW, of the transposed LU table at valid rows
    auto lu = [&](int i, int j) { return LU[IX(cl(j) * W + (BW + i - j)) * IX(c.nl)]; };

    // ---- forward: rhs row and L-substitution
    double su[D], sn[D], sp[D], sl[D][BW];
    auto fload = [&](int s, int i) {
        const int jn = i + BW + 1, ic = cl(i);
        const double v = u[IX(nlm) * IX(cl(jn))];
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
            if (i < n) o[IX(nlm) * IX(i)] = v;
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
        by[s] = o[IX(nlm) * IX(ic)];
        bd[s] = rdg[IX(ic) * IX(c.nl)];
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
            if (i >= 0) o[IX(nlm) * IX(i)] = v;
            if constexpr (DIRECT) {
                if (wA && i >= 0) pA[IX(i * SYNV_COLS) * IX(nlm)] = (TS)(f == 4 ? v * rinv_[i] : v);
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
                const IX ov = IX(r * SYNV_COLS) * IX(nlm), os = IX(r * 10) * IX(nlm);
                pA[ov] = (TS)a;
                if (wB) pB[ov] = (TS)bv;
                if (wS) pS[os] = (TS)sv;
            }
            }
            bload(d, i - D);
        }
    }
}

template <int BW, int D, typename TS, bool DIRECT, typename IX>
__global__ void __launch_bounds__(64) cnab2_chain_kernel(ChainArgs c)
{
    cnab2_chain_body<BW, D, TS, DIRECT, IX>(c, blockIdx.x, blockIdx.y, blockIdx.z);
}

// the block's (column block, field, re/im) from its index within a shell's
// share of a batched launch (column block fastest, as the shell's own grid)
// a batched launch's block loc of one shell -> (column block, field, re/im)
// of that shell's own grid (gx, nfield, 2 / NC)
template <int NC, class AT> __device__ __forceinline__ void chain_decode(const AT &c, int loc, int &bx, int &by, int &bz)
{
    const int gx = (c.nlm + 63) / 64, nfield = c.nf ? c.nf : (c.c_in ? 6 : 5);
    bx = loc % gx;
    by = (loc / gx) % nfield;
    bz = loc / (gx * nfield);
}

// one grid over the CNAB2 solves of several shells (a batched step)
template <int BW, int D, typename TS, bool DIRECT, typename IX>
__global__ void __launch_bounds__(64) cnab2_chain_batch_kernel(BatchRef<ChainArgs> b)
{
    const int sh = b.shell_of_block[blockIdx.x];
    if (sh == BATCH_PAD) return;
    const auto &c = batch_args(b, sh);
    int bx, by, bz;
    chain_decode<1>(c, blockIdx.x - b.block_off[sh], bx, by, bz);
    cnab2_chain_body<BW, D, TS, DIRECT, IX>(c, bx, by, bz);
}

// Twisted form of the tangential (DIRECT) column kernel.  A block of two
// waves takes 64 columns.  Wave 0 takes rows [0, n1): top-down rhs and L1
// substitution, then upward back substitution with U1; wave 1 takes rows
// [n1, n) of the same columns with the bottom block's UL factors: bottom-up
// rhs and U2 substitution, then downward substitution with L2.  The half is
// wave-uniform, so every row index and table offset stays on the scalar unit
// as in the one-lane form.  The waves exchange their last BW intermediate
// values through LDS (one barrier), solve the interface rows with R_l, and
// finish outward: the serial chain per column is about half of the one-lane
// form.  The rhs, the Dirichlet rows and the synthesis columns are those of
// cnab2_chain_kernel.  Lanes past nlm repeat column nlm-1: they compute the
// same values as its own lane and store them to the same places, which keeps
// both waves on the barrier and the loops free of per-lane store guards.
// NC = 2: each lane solves the re and the im column of its (idx, field)
// (blk_z = 0): the two columns share every row offset, band-table load and
// interface matrix, so the scalar address work and the table loads per column
// halve and the two independent recurrences interleave.  Each column gets the
// operations of NC = 1; the compiler may fuse them into FMAs differently, so
// the two forms agree to rounding, not always bit for bit.
template <int BW, int D, typename TS, int NC, class AT>
__device__ __forceinline__ void cnab2_twist_body(const AT &c, int blk_x, int blk_y, int blk_z)
{
    constexpr int W = 2 * BW + 1, M2 = 2 * BW;
    const int n = c.n, nlm = c.nlm, n1 = (n + 1) / 2, nb = n - n1;
    const int lane = threadIdx.x & 63;
    const bool bot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) != 0;
    const int idx = min(blk_x * 64 + lane, nlm - 1);
    const int fq = c.f0 + blk_y;
    const bool isC = fq == 5;
    const int f = isC ? 4 : fq;
    const int k = isC ? 3 : (f < 2 ? 0 : (f < 4 ? 1 : 2));
    constexpr int ER = (W + 2) & ~1;
    extern __shared__ double smem[];
    // Ex: row i of the explicit operator at [ER*i + q] (column i - BW + q; q = W: 1/r^2);
    // ExB: the same rows mirrored (column i + BW - q at q), so both halves read er[q]
    double *Ex = smem, *ExB = Ex + ER * n, *rinv_ = ExB + ER * n, *xch = rinv_ + n; // xch: [col][half][BW][64]
    // ys: the first c.ylds forward results y of each half, [half][s][col][64]
    // (the rest go through st_out as in the one-column form)
    double *ys = xch + NC * 2 * BW * 64;
    // The prologue issues its loads together: every value that a condition
    // would zero is read from the zero row instead (the load address is
    // selected, never the loaded value), because a select of a loaded value
    // on a wave-uniform condition becomes a branch around the load and a wait
    // for it, one memory latency per load.
    const int l = c.l_of_idx[idx];
    const size_t fs = (size_t)nlm * n;
    const bool nlz = !isC && ((c.nl_zero_mask >> f) & 1);
    const unsigned nstride = nlz ? 0u : (unsigned)nlm;
    // column cc of this lane: re/im ri = blk_z (NC = 1) or cc (NC = 2)
    const double *u[NC], *pn[NC], *pp[NC];
    double *o[NC];
    TS *pA[NC];
    const bool wA = f == 0 || f == 1 || f == 4;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        const int ri = NC == 1 ? blk_z : cc;
        const size_t off = (size_t)(2 * f + ri) * fs + idx;
        u[cc] = (isC ? c.c_in : c.st_in) + off;
        pn[cc] = nlz ? c.zero_row + idx : (isC ? c.c_nl : c.nl_arr) + off;
        pp[cc] = nlz ? c.zero_row + idx : (isC ? c.c_np : c.np_arr) + off;
        o[cc] = (isC ? c.c_out : c.st_out) + off;
        pA[cc] = static_cast<TS *>(c.synV) +
                 (size_t)((isC ? SYNV_COLS - 2 : (f == 0 ? 2 : (f == 1 ? 0 : 16))) + ri) * nlm + idx;
    }
    const bool tkeep = isC || ((f == 4) && ((c.bc_mask >> 5) & 1));
    const bool tbc = (f == 4) && !tkeep;
    const bool bcr = isC ? c.c_bcr != 0 : ((c.bc_mask >> f) & 1);
    const double inv_dt = 1.0 / c.dt, lw = 1.0 - c.theta;
    const int dir = bot ? -1 : 1;
    auto cl = [&](int j) { return j < 0 ? 0 : (j > n - 1 ? n - 1 : j); };
    auto urow = [&](int j) { return j >= 0 && j < n && !(tbc && (j == 0 || j == n - 1)); };

    // ---- outward-in: rhs rows and the unit-triangular substitution
    const int nfw = bot ? nb : n1; // valid steps of this lane
    double w[NC][W];               // u rows (top: i-BW..i+BW; bottom: i+BW..i-BW)
    {
        const int i0 = bot ? n - 1 : 0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const int j = i0 + dir * (q - BW);
            const bool keep = urow(j);
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) w[cc][q] = *(keep ? u[cc] + (unsigned)nlm * j : c.zero_row + idx);
        }
    }
    // the operator rows (UF entries per thread and round, loads before stores)
    const double *Lapk = c.nuLap + (size_t)k * W * n;
    constexpr int UF = 4;
    for (int t0 = threadIdx.x; t0 < ER * n; t0 += UF * 128) {
        double v[UF];
#pragma unroll
        for (int e = 0; e < UF; ++e) {
            const int t = min(t0 + e * 128, ER * n - 1), i = t / ER, q = t - i * ER, j = i - BW + q;
            const bool bnd = q < W && j >= 0 && j < n; // q = BW: the diagonal, j = i
            v[e] = *(bnd ? Lapk + (2 * BW - q) + W * j : (q == W ? c.rtab + i + n : c.zero_row));
        }
#pragma unroll
        for (int e = 0; e < UF; ++e) {
            const int t = t0 + e * 128;
            if (t < ER * n) {
                const int i = t / ER, q = t - i * ER;
                Ex[t] = v[e];
                ExB[ER * i + (q < W ? 2 * BW - q : q)] = v[e];
            }
        }
    }
    for (int t = threadIdx.x; t < n; t += 128) rinv_[t] = c.rtab[t + 2 * n];
    __syncthreads();
    const double lf = (double)l * (double)(l + 1);
    const double nulf = c.nus[k] * lf;
    const double *T = (bot ? c.ULt : c.LUt) + (size_t)k * W * n * c.nl + l;
    const double *rdg = (bot ? c.rdiag2 : c.rdiag) + (size_t)k * n * c.nl + l;
    const double *Rl = c.Rint + ((size_t)k * c.nl + l) * (M2 * M2);
    // band entry (i, j) of this half's table; j outside 0..n-1 reads the
    // (finite) diagonal instead.  No mask is needed: the forward rows outside
    // 0..n-1 meet yw entries that start at 0, and the back substitution's rows
    // stay inside its half.  Loads stay unconditional and their values are
    // used D steps later, so the prefetch is never drained by a select.
    auto band = [&](int ic, int j) {
        const int jc = (j >= 0 && j < n) ? j : ic;
        return T[(unsigned)(jc * W + (BW + ic - jc)) * (unsigned)c.nl];
    };
    auto put = [&](int cc, int i, double v) { // x row i: state and (u, T/r, C/r) synthesis column
        o[cc][(unsigned)nlm * i] = v;
        if (wA) pA[cc][(unsigned)(i * SYNV_COLS) * (unsigned)nlm] = (TS)(f == 4 ? v * rinv_[i] : v);
    };
    double su[NC][D], sn[NC][D], sp[NC][D], sl[D][BW];
    auto fload = [&](int sl_, int s) {
        const int i = bot ? n - 1 - s : s, ic = cl(i), jn = i + dir * (BW + 1);
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            su[cc][sl_] = u[cc][(unsigned)nlm * cl(jn)]; // masked with urow at its use
            sn[cc][sl_] = pn[cc][nstride * ic];
            sp[cc][sl_] = pp[cc][nstride * ic];
        }
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            sl[sl_][q] = band(ic, i + dir * (q - BW)); // the q-th of the BW rows already eliminated
        }
    };
    double yw[NC][BW];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int q = 0; q < BW; ++q) yw[cc][q] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) fload(d, d);
    // step s of this half from prefetch slot d; `valid` is false only in the
    // tail block, so full blocks shift the windows as plain register renames
    // y rows kept in LDS: the first Rw of this half (a multiple of D, so the
    // loop below switches stores between whole blocks), never the last BW
    // (the interface rows, which the back substitution does not read)
    const int Rw = NC == 2 ? (min(c.ylds, nfw - BW) / D) * D : 0;
    double *yh = ys + (size_t)(bot ? c.ylds : 0) * NC * 64 + lane;
    auto fstep = [&](int d, int s, bool valid, bool yl) {
        const int i = bot ? n - 1 - s : s, ic = cl(i);
        const double *er = (bot ? ExB : Ex) + ER * ic;
        const bool unext = urow(i + dir * (BW + 1));
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            double lr = 0.0;
#pragma unroll
            for (int q = 0; q < W; ++q) lr += ((q == BW) ? er[BW] - nulf * er[W] : er[q]) * w[cc][q];
            double v = inv_dt * w[cc][BW] + 1.5 * sn[cc][d] - 0.5 * sp[cc][d];
            v += lw * lr;
            v = (bcr && (i == 0 || i == n - 1)) ? (tkeep ? w[cc][BW] : 0.0) : v; // Dirichlet value
#pragma unroll
            for (int q = 0; q < BW; ++q) v -= sl[d][q] * yw[cc][q];
            if (valid) {
#pragma unroll
                for (int q = 0; q < BW - 1; ++q) yw[cc][q] = yw[cc][q + 1];
                yw[cc][BW - 1] = v;
#pragma unroll
                for (int q = 0; q < W - 1; ++q) w[cc][q] = w[cc][q + 1];
                w[cc][W - 1] = unext ? su[cc][d] : 0.0;
                if (yl) yh[(s * NC + cc) * 64] = v;
                else o[cc][(unsigned)nlm * i] = v;
            }
        }
        fload(d, s + D);
    };
    int s0 = 0;
    for (; s0 < Rw; s0 += D) { // y to LDS
#pragma unroll
        for (int d = 0; d < D; ++d) fstep(d, s0 + d, true, true);
    }
    for (; s0 + D <= nfw; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) fstep(d, s0 + d, true, false);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) fstep(d, s0 + d, s0 + d < nfw, false); // the last nfw % D steps

    // ---- interface: top yw[q] = y(n1-BW+q), bottom yw[q] = z(n1+BW-1-q)
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int q = 0; q < BW; ++q) xch[((cc * 2 + (bot ? 1 : 0)) * BW + q) * 64 + lane] = yw[cc][q];
    __syncthreads();
    double xw[NC][BW]; // the BW rows solved last, nearest first
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        double g[M2];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            const double other = xch[((cc * 2 + (bot ? 0 : 1)) * BW + q) * 64 + lane];
            const double mine = yw[cc][q];
            // g[t] = y(n1-BW+t); g[BW+t] = z(n1+t) = bottom's yw[BW-1-t]
            if (bot) {
                g[q] = other;
                g[M2 - 1 - q] = mine;
            } else {
                g[q] = mine;
                g[M2 - 1 - q] = other;
            }
        }
        double xb[BW]; // top: x(n1-BW+t); bottom: x(n1+t)
#pragma unroll
        for (int t = 0; t < BW; ++t) {
            const int a = (bot ? BW : 0) + t;
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < M2; ++b) v += Rl[a * M2 + b] * g[b];
            xb[t] = v;
        }
#pragma unroll
        for (int t = 0; t < BW; ++t) put(cc, bot ? n1 + t : n1 - BW + t, xb[t]);
#pragma unroll
        for (int q = 0; q < BW; ++q) xw[cc][q] = bot ? xb[BW - 1 - q] : xb[q];
    }

    // ---- inside-out: back substitution (top: U1 upward; bottom: L2 downward)
    // step t is forward step s = nbk - 1 - t of this half: its y is in LDS
    // from t = T1 on
    const int nbk = bot ? nb - BW : n1 - BW; // valid steps of this lane
    const int T1 = nbk - Rw;
    double by[NC][D], bd[D], bu[D][BW];
    auto bload = [&](int sl_, int t, bool yl) {
        const int i = bot ? n1 + BW + t : n1 - BW - 1 - t, ic = cl(i);
        if (!yl)
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) by[cc][sl_] = o[cc][(unsigned)nlm * ic];
        bd[sl_] = rdg[(unsigned)ic * (unsigned)c.nl];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            bu[sl_][q] = band(ic, i + dir * (1 + q)); // the q-th nearest row solved before (top: below, bottom: above)
        }
    };
#pragma unroll
    for (int d = 0; d < D; ++d) bload(d, d, false);
    auto bstep = [&](int d, int t, bool valid, bool yl) { // as fstep: only the tail blocks are guarded
        const int i = bot ? n1 + BW + t : n1 - BW - 1 - t;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            double v = yl ? yh[((nbk - 1 - t) * NC + cc) * 64] : by[cc][d];
#pragma unroll
            for (int q = 0; q < BW; ++q) v -= bu[d][q] * xw[cc][q];
            v *= bd[d];
            if (valid) {
#pragma unroll
                for (int q = BW - 1; q > 0; --q) xw[cc][q] = xw[cc][q - 1];
                xw[cc][0] = v;
                put(cc, i, v);
            }
        }
        bload(d, t + D, yl);
    };
    int t0 = 0;
    for (; t0 + D <= T1; t0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) bstep(d, t0 + d, true, false);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) bstep(d, t0 + d, t0 + d < T1, false); // the last T1 % D steps
    if (Rw > 0) {
        // the LDS rows: restart the table prefetch at T1, y read at its use
        t0 = T1;
#pragma unroll
        for (int d = 0; d < D; ++d) bload(d, t0 + d, true);
        for (; t0 + D <= nbk; t0 += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) bstep(d, t0 + d, true, true);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) bstep(d, t0 + d, t0 + d < nbk, true);
    }
}

// three waves per SIMD hold the one-column form's registers; the paired form
// launches half the waves and takes up to 256 VGPRs at two
template <int BW, int D, typename TS, int NC>
__global__ void __launch_bounds__(128, NC == 1 ? 3 : 2) cnab2_twist_kernel(ChainArgs c)
{
    cnab2_twist_body<BW, D, TS, NC>(c, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <int BW, int D, typename TS, int NC>
__global__ void __launch_bounds__(128, NC == 1 ? 3 : 2) cnab2_twist_batch_kernel(BatchRef<ChainArgs> b)
{
    const int sh = b.shell_of_block[blockIdx.x];
    if (sh == BATCH_PAD) return;
    const auto &c = batch_args(b, sh);
    int bx, by, bz;
    chain_decode<NC>(c, blockIdx.x - b.block_off[sh], bx, by, bz);
    cnab2_twist_body<BW, D, TS, NC>(c, bx, by, bz);
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

void gd_cnab2_chain_form(const ChainArgs &c, ChainForm *f)
{
    f->lds_lane = sizeof(double) * (size_t)((((2 * c.bw + 1) + 2) & ~1) + 2 * (2 * c.bw + 1) + 2) * c.n;
    f->direct = c.synV && c.syn_tang;
    f->lds_twist = sizeof(double) * ((size_t)(2 * (((2 * c.bw + 1) + 2) & ~1) + 1) * c.n +
                                     2 * (size_t)c.bw * 64 * (c.pair ? 2 : 1) +
                                     (c.pair ? 2 * (size_t)c.ylds * 2 * 64 : 0));
    // row, synthesis-column and table offsets from the per-column base pointers
    // fit 32 bits (fewer scalar instructions per level); the twisted kernel needs it
    f->off32 = !c.off64 && (size_t)SYNV_COLS * c.nlm * c.n < ((size_t)1 << 32) &&
               (size_t)(2 * c.bw + 1) * c.n * c.nl < ((size_t)1 << 32);
    f->twist = f->direct && c.ULt && c.rdiag2 && c.Rint && c.bw <= 3 && c.n >= 2 * c.bw + 2 &&
               f->lds_twist <= 64 * 1024 && f->off32;
    f->pair = f->twist ? c.pair : 0;
}

int gd_chain_ylds(const ChainArgs &c, size_t lds_budget)
{
    ChainArgs b = c;
    b.ylds = 0;
    ChainForm fm;
    gd_cnab2_chain_form(b, &fm);
    if (!fm.pair || lds_budget <= fm.lds_twist) return 0;
    const size_t per_row = sizeof(double) * 2 * 2 * 64; // both halves, re and im, 64 lanes
    const int rows = (int)((lds_budget - fm.lds_twist) / per_row);
    const int n1 = (c.n + 1) / 2;
    return std::max(0, std::min(rows, n1 - c.bw));
}

int gd_cnab2_chain_blocks(const ChainArgs &c)
{
    const int nfield = c.nf ? c.nf : (c.c_in ? 6 : 5);
    ChainForm fm;
    gd_cnab2_chain_form(c, &fm);
    return ((c.nlm + 63) / 64) * nfield * (fm.pair ? 1 : 2);
}

// several shells' column solves in one grid, every shell in form `fm` (as
// gd_cnab2_chain_form gives it for each of them); lds: the largest shell's
int gd_launch_cnab2_chain_batch(hipStream_t st, const BatchRef<ChainArgs> &b, int nblocks, int bw, const ChainForm &fm,
                                size_t lds, bool syn_fp64)
{
    if (nblocks <= 0) return GD_OK;
    GD_CHECK(lds <= 64 * 1024, GD_ERR_ARG, "cnab2_chain batch: %zu B of LDS", lds);
    if (fm.twist) {
#define GD_TWISTB_NC(B, NC, DD)                                                                                         \
    {                                                                                                                  \
        if (syn_fp64) hipLaunchKernelGGL((cnab2_twist_batch_kernel<B, DD, double, NC>), dim3(nblocks), dim3(128), lds, st, b); \
        else hipLaunchKernelGGL((cnab2_twist_batch_kernel<B, DD, float, NC>), dim3(nblocks), dim3(128), lds, st, b);    \
    }
#define GD_TWISTB(B)                                                                                                   \
    if (bw == B) {                                                                                                     \
        if (fm.pair == 3) GD_TWISTB_NC(B, 2, 3) else if (fm.pair) GD_TWISTB_NC(B, 2, 2) else GD_TWISTB_NC(B, 1, 4)                                                   \
        return gd_check_launch("cnab2_twist_batch_kernel");                                                            \
    }
        GD_TWISTB(1) GD_TWISTB(2) GD_TWISTB(3)
#undef GD_TWISTB
#undef GD_TWISTB_NC
        GD_CHECK(false, GD_ERR_ARG, "cnab2_chain batch: twisted form with bw=%d", bw);
    }
    GD_CHECK(fm.direct, GD_ERR_ARG, "cnab2_chain batch: the tangential (direct) form only");
#define GD_CHAINB_IX(B, IX)                                                                                            \
    {                                                                                                                  \
        if (syn_fp64) hipLaunchKernelGGL((cnab2_chain_batch_kernel<B, 4, double, true, IX>), dim3(nblocks), dim3(64), lds, st, b); \
        else hipLaunchKernelGGL((cnab2_chain_batch_kernel<B, 4, float, true, IX>), dim3(nblocks), dim3(64), lds, st, b);   \
    }
#define GD_CHAINB(B)                                                                                                   \
    if (bw == B) {                                                                                                     \
        if (fm.off32) GD_CHAINB_IX(B, unsigned) else GD_CHAINB_IX(B, size_t)                                           \
        return gd_check_launch("cnab2_chain_batch_kernel");                                                            \
    }
    GD_CHAINB(1) GD_CHAINB(2) GD_CHAINB(3) GD_CHAINB(4) GD_CHAINB(5) GD_CHAINB(6) GD_CHAINB(7)
#undef GD_CHAINB
#undef GD_CHAINB_IX
    gd_set_error("cnab2_chain batch: no kernel for bw=%d", bw);
    return GD_ERR_ARG;
}

int gd_launch_cnab2_chain(hipStream_t st, const ChainArgs &c, bool syn_fp64)
{
    GD_CHECK(c.st_in != c.st_out, GD_ERR_ARG, "cnab2_chain: in and out state must differ");
    GD_CHECK(c.rdiag, GD_ERR_ARG, "cnab2_chain: reciprocal diagonal table missing");
    // the kernels' prologues read every value a condition zeroes from zero_row
    GD_CHECK(c.zero_row, GD_ERR_ARG, "cnab2_chain: zero_row (nlm zeros) missing");
    GD_CHECK(!c.synV || (c.synS && c.D1 && c.D2 && c.rtab), GD_ERR_ARG, "cnab2_chain: incomplete synthesis outputs");
    GD_CHECK(gd_cnab2_chain_supported(c.n, c.bw), GD_ERR_ARG, "cnab2_chain: n=%d bw=%d unsupported", c.n, c.bw);
    const int nfield = c.nf ? c.nf : (c.c_in ? 6 : 5);
    GD_CHECK(c.f0 >= 0 && c.nf >= 0 && c.f0 + nfield <= (c.c_in ? 6 : 5), GD_ERR_ARG, "cnab2_chain: fields %d+%d",
             c.f0, nfield);
    GD_CHECK(!c.c_in || (c.c_out && c.c_nl && c.c_np && (!c.synV || c.syn_tang)), GD_ERR_ARG,
             "cnab2_chain: incomplete composition buffers");
    dim3 grid((c.nlm + 63) / 64, nfield, 2);
    ChainForm fm;
    gd_cnab2_chain_form(c, &fm);
    const size_t lds = fm.lds_lane, tlds = fm.lds_twist;
    GD_CHECK(lds <= 64 * 1024, GD_ERR_ARG, "cnab2_chain: n=%d bw=%d needs %zu B of LDS", c.n, c.bw, lds);
    const bool direct = fm.direct, off32 = fm.off32;
    if (fm.twist) {
        // twisted form: 64 columns per block, one wave per half
        const dim3 tgrid((c.nlm + 63) / 64, grid.y, fm.pair ? 1 : 2);
#define GD_TWIST_NC(B, NC, DD)                                                                                    \
    {                                                                                                            \
        if (syn_fp64) hipLaunchKernelGGL((cnab2_twist_kernel<B, DD, double, NC>), tgrid, dim3(128), tlds, st, c); \
        else hipLaunchKernelGGL((cnab2_twist_kernel<B, DD, float, NC>), tgrid, dim3(128), tlds, st, c);           \
    }
#define GD_TWIST(B)                                                                                              \
    if (c.bw == B) {                                                                                             \
        if (fm.pair == 3) GD_TWIST_NC(B, 2, 3) else if (fm.pair) GD_TWIST_NC(B, 2, 2) else GD_TWIST_NC(B, 1, 4)                                               \
        return gd_check_launch("cnab2_twist_kernel");                                                            \
    }
        GD_TWIST(1) GD_TWIST(2) GD_TWIST(3) // wider bands spill at 3 waves per SIMD
#undef GD_TWIST
#undef GD_TWIST_NC
    }
#define GD_CHAIN_IX(B, IX)                                                                                       \
    {                                                                                                            \
        if (syn_fp64 && direct) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, double, true, IX>), grid, dim3(64), lds, st, c); \
        else if (syn_fp64) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, double, false, IX>), grid, dim3(64), lds, st, c); \
        else if (direct) hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, float, true, IX>), grid, dim3(64), lds, st, c); \
        else hipLaunchKernelGGL((cnab2_chain_kernel<B, 4, float, false, IX>), grid, dim3(64), lds, st, c);       \
    }
#define GD_CHAIN(B)                                                                                              \
    if (c.bw == B) {                                                                                             \
        if (off32) GD_CHAIN_IX(B, unsigned) else GD_CHAIN_IX(B, size_t)                                          \
        return gd_check_launch("cnab2_chain_kernel");                                                            \
    }
    GD_CHAIN(1) GD_CHAIN(2) GD_CHAIN(3) GD_CHAIN(4) GD_CHAIN(5) GD_CHAIN(6) GD_CHAIN(7)
#undef GD_CHAIN
#undef GD_CHAIN_IX
    gd_set_error("cnab2_chain: no kernel for n=%d bw=%d", c.n, c.bw);
    return GD_ERR_ARG;
}


// ---------------------------------------------------------------- twisted column solve (tables)
// The tangential column kernel can split each column at n1 = ceil(n/2): the
// top rows keep the LU of S_l (its top-left block is the LU of S_l[0:n1,
// 0:n1]), the bottom block S_l[n1:, n1:] gets a UL factorisation (eliminated
// from the bottom up: unit upper U2, lower L2), and the 2BW unknowns next to
// the split solve a 2BW x 2BW interface system
//   [ U(top BW x BW)      E1 = L1^-1 B1 (its last BW rows) ] [x(n1-BW..n1-1)]   [y(n1-BW..n1-1)]
//   [ F2 = U2^-1 C2 (first BW rows)   L2(first BW x BW)    ] [x(n1..n1+BW-1)] = [z(n1..n1+BW-1)]
// with B1, C2 the coupling blocks of S_l.  Host FP64, once per operator.
int gd_chain_twist_tables(const double *S, const double *LU, int n, int bw, int nl, double *ul_t, double *rd2,
                          double *R)
{
    GD_CHECK(n >= 2 * bw + 2, GD_ERR_ARG, "twist tables: n=%d too small for bw=%d", n, bw);
    const int W = 2 * bw + 1, n1 = (n + 1) / 2, m2 = 2 * bw;
    std::vector<double> A((size_t)W * n), E(n), M((size_t)m2 * m2), Inv((size_t)m2 * m2);
    auto bi = [&](int i, int j) { return (size_t)j * W + (bw + i - j); }; // band index, |i - j| <= bw
    auto inb = [&](int i, int j) { return i >= 0 && j >= 0 && i < n && j < n && i - j <= bw && j - i <= bw; };
    for (int l = 0; l < nl; ++l) {
        const double *Sl = S + (size_t)l * W * n, *Ll = LU + (size_t)l * W * n;
        // UL of the bottom block (band copy, entries outside the block ignored)
        for (size_t q = 0; q < A.size(); ++q) A[q] = Sl[q];
        for (int kc = n - 1; kc > n1; --kc) {
            const double piv = A[bi(kc, kc)];
            GD_CHECK(piv != 0.0 && std::isfinite(piv), GD_ERR_NUMERIC, "twist tables: zero pivot (l=%d)", l);
            for (int i = std::max(n1, kc - bw); i < kc; ++i) {
                const double fct = A[bi(i, kc)] / piv;
                for (int j = std::max(n1, kc - bw); j < kc; ++j)
                    if (inb(i, j)) A[bi(i, j)] -= fct * A[bi(kc, j)];
                A[bi(i, kc)] = fct; // the unit-upper multiplier
            }
        }
        for (int j = 0; j < n; ++j)
            for (int r = 0; r < W; ++r) {
                const int i = j + r - bw;
                const bool keep = i >= n1 && j >= n1 && i < n;
                ul_t[((size_t)j * W + r) * nl + l] = keep ? A[(size_t)j * W + r] : 0.0;
            }
        for (int i = 0; i < n; ++i) {
            const double d = A[bi(i, i)];
            if (i >= n1) GD_CHECK(d != 0.0 && std::isfinite(d), GD_ERR_NUMERIC, "twist tables: zero pivot (l=%d)", l);
            rd2[(size_t)i * nl + l] = i >= n1 ? 1.0 / d : 0.0;
        }
        // interface matrix
        for (int c = 0; c < bw; ++c) {
            // E1 column c: forward substitution of S(:, n1 + c) over rows < n1 with unit L1
            for (int i = 0; i < n1; ++i) {
                double v = inb(i, n1 + c) ? Sl[bi(i, n1 + c)] : 0.0;
                for (int j = std::max(0, i - bw); j < i; ++j) v -= Ll[bi(i, j)] * E[j];
                E[i] = v;
            }
            for (int t = 0; t < bw; ++t) M[(size_t)t * m2 + bw + c] = E[n1 - bw + t];
            // F2 column c: back substitution of S(:, n1 - bw + c) over rows >= n1 with unit U2
            for (int i = n - 1; i >= n1; --i) {
                double v = inb(i, n1 - bw + c) ? Sl[bi(i, n1 - bw + c)] : 0.0;
                for (int j = i + 1; j <= std::min(n - 1, i + bw); ++j) v -= A[bi(i, j)] * E[j];
                E[i] = v;
            }
            for (int t = 0; t < bw; ++t) M[(size_t)(bw + t) * m2 + c] = E[n1 + t];
        }
        for (int t = 0; t < bw; ++t)
            for (int s = 0; s < bw; ++s) {
                const int it = n1 - bw + t, is = n1 - bw + s;
                M[(size_t)t * m2 + s] = s >= t && inb(it, is) ? Ll[bi(it, is)] : 0.0;          // U of the LU
                const int jt = n1 + t, js = n1 + s;
                M[(size_t)(bw + t) * m2 + bw + s] = s <= t && inb(jt, js) ? A[bi(jt, js)] : 0.0; // L2
            }
        // inverse by Gauss-Jordan with partial pivoting
        for (int a = 0; a < m2; ++a)
            for (int b = 0; b < m2; ++b) Inv[(size_t)a * m2 + b] = a == b ? 1.0 : 0.0;
        for (int col = 0; col < m2; ++col) {
            int p = col;
            for (int r = col + 1; r < m2; ++r)
                if (std::fabs(M[(size_t)r * m2 + col]) > std::fabs(M[(size_t)p * m2 + col])) p = r;
            GD_CHECK(M[(size_t)p * m2 + col] != 0.0, GD_ERR_NUMERIC, "twist tables: singular interface (l=%d)", l);
            for (int b = 0; b < m2; ++b) {
                std::swap(M[(size_t)p * m2 + b], M[(size_t)col * m2 + b]);
                std::swap(Inv[(size_t)p * m2 + b], Inv[(size_t)col * m2 + b]);
            }
            const double d = M[(size_t)col * m2 + col];
            for (int b = 0; b < m2; ++b) {
                M[(size_t)col * m2 + b] /= d;
                Inv[(size_t)col * m2 + b] /= d;
            }
            for (int r = 0; r < m2; ++r)
                if (r != col) {
                    const double fct = M[(size_t)r * m2 + col];
                    for (int b = 0; b < m2; ++b) {
                        M[(size_t)r * m2 + b] -= fct * M[(size_t)col * m2 + b];
                        Inv[(size_t)r * m2 + b] -= fct * Inv[(size_t)col * m2 + b];
                    }
                }
        }
        for (int q = 0; q < m2 * m2; ++q) R[(size_t)l * m2 * m2 + q] = Inv[q];
    }
    return GD_OK;
}
