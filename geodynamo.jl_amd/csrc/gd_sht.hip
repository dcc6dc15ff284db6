// gd_sht.hip -- spherical-harmonic transform kernels for gfx950.
//
//  plm_kernel        FP64 orthonormal P_lm, dP_lm/dtheta, m P_lm/sin(theta) by
//                    the three-term recurrence with a carried base-2 exponent
//  synth_kernel      per-m Legendre contraction on MFMA (16x16x4 f32 / f64):
//                    Fourier(i, col) = sum_l Tab(l, i) a(l, col), split into
//                    (l-m) even / odd partial sums over the northern nodes;
//                    the epilogue forms north/south rings and, for vector
//                    fields, v_theta = dP S + i Q T, v_phi = i Q S - dP T
//  ana_kernel        the transpose contraction over the northern nodes with
//                    G+- = Gn +- Gs, epilogue S = X_t - i Y_p, T = -i Y_t - X_p
//  ring_kernel       longitude FFTs of a northern ring and its southern mirror
//                    as one complex FFT (Stockham, radices 2,3,4,6,8, in LDS);
//                    RING_STEP also evaluates the pointwise nonlinear products
//                    between the inverse and forward FFTs, so physical-space
//                    values never leave LDS.
#include <cmath>

#include <cstdint>
#include <cstdlib>
#include <type_traits>

#include "gd_sht.hpp"

// ====================================================================== P_lm
__global__ void plm_kernel(const double *__restrict__ x, const int *__restrict__ moff, int lmax, int mmax, int nlat2,
                           int nlm, double *__restrict__ tab64, float *__restrict__ tab32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (i >= nlat2 || m > mmax) return;
    const double xi = x[i], s = sqrt(1.0 - xi * xi);
    double lg = 0.5 * log(1.0 / (4.0 * M_PI)) + 0.5 * log(2.0 * m + 1.0) + m * log(s);
    for (int k = 1; k <= m; ++k) lg += 0.5 * log((2.0 * k - 1.0) / (2.0 * k));
    int e = (int)floor(lg / M_LN2);
    double val = exp(lg - e * M_LN2);
    if (m & 1) val = -val;
    double pm2 = 0.0, pm1 = val, tprev = 0.0;
    const int K = lmax - m + 1, ne = (K + 1) / 2;
    for (int l = m; l <= lmax; ++l) {
        if (l == m + 1) {
            double cur = sqrt(2.0 * m + 3.0) * xi * pm1;
            pm2 = pm1;
            pm1 = cur;
        } else if (l > m + 1) {
            double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            double b = sqrt((((double)l - 1) * ((double)l - 1) - (double)m * m) /
                            (4.0 * ((double)l - 1) * ((double)l - 1) - 1.0));
            double cur = a * (xi * pm1 - b * pm2);
            pm2 = pm1;
            pm1 = cur;
        }
        if (fabs(pm1) > 0x1p600) {
            pm1 = ldexp(pm1, -600);
            pm2 = ldexp(pm2, -600);
            e += 600;
        }
        const double t = ldexp(pm1, e);
        const double c = (l == m) ? 0.0 : sqrt((2.0 * l + 1.0) * ((double)l * l - (double)m * m) / (2.0 * l - 1.0));
        const double dp = (l * xi * t - c * tprev) / s;
        const double q = m * t / s;
        tprev = t;
        const int d = l - m;
        const int idx = moff[m] + ((d & 1) ? ne + (d - 1) / 2 : d / 2);
        const size_t o = (size_t)idx * nlat2 + i;
        tab64[o] = t;
        tab64[(size_t)nlm * nlat2 + o] = dp;
        tab64[(size_t)2 * nlm * nlat2 + o] = q;
        if (tab32) {
            tab32[o] = (float)t;
            tab32[(size_t)nlm * nlat2 + o] = (float)dp;
            tab32[(size_t)2 * nlm * nlat2 + o] = (float)q;
        }
    }
}

int gd_launch_plm(hipStream_t st, const double *x, const int *moff, int lmax, int mmax, int nlat2, double *tab64,
                  float *tab32, int nlm)
{
    hipLaunchKernelGGL(plm_kernel, dim3((nlat2 + 63) / 64, mmax + 1), dim3(64), 0, st, x, moff, lmax, mmax, nlat2, nlm,
                       tab64, tab32);
    return gd_check_launch("plm_kernel");
}

// ====================================================================== Legendre GEMMs
namespace {
constexpr int KT = 16; // K tile staged in LDS
constexpr int TM = 32; // rows per tile (two 16-row MFMA blocks)

template <typename T> using acc_t = typename Acc4<T>::type;

template <typename T> __device__ __forceinline__ T shfl_xor3(T v) { return __shfl_xor(v, 3); }

// XCD-aware block order.  Workgroup b of a launch goes to XCD b % 8, and each
// XCD has its own L2; the blocks of one order m share their Legendre table
// tile (and their spectral / Fourier inputs), so the logical index is built
// so that consecutive logical blocks run on one XCD: XCD x takes the x-th
// contiguous range of the logical order.  The m dimension is ordered
// 0, mmax, 1, mmax-1, ... so that equal ranges carry equal work (the degree
// count of m falls linearly with m).  (xcd_logical: gd_internal.hpp)
__device__ __forceinline__ int paired_m(int k, int nm) { return (k & 1) ? nm - 1 - (k >> 1) : (k >> 1); }

// 16-byte vector of T (stores of the staged output tiles)
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

// Tile shape of a synthesis block: NF fields, VC virtual columns per level
// (vector: S_re S_im T_re T_im per field, scalar: re im), NW waves of 16
// columns each, NTILE = 16 NW columns = LT levels; RS output slots per
// (node, level) record.
// NG gradient fields (vector only: S alone, T = 0) follow the NF full fields
// with 2 columns each: both components come from the S columns, so the zero
// T columns are not contracted.  A full field's 4 columns must start at a
// multiple of 4 (its epilogue exchanges values within a lane quad); when a
// level's column count is not a multiple of 4, two levels form one group:
// their full fields first, then their gradient fields.
template <bool VEC, int NF, int NW, int NG = 0> struct SynthTile {
    static constexpr int VC = VEC ? 4 * NF + 2 * NG : 2 * NF; // columns per level
    static constexpr int PAIR = (VEC && VC % 4 != 0) ? 2 : 1;  // levels per column group
    static constexpr int GC = VC * PAIR;                        // columns per group
    static constexpr int NTILE = 16 * NW;
    static constexpr int LT = NTILE / GC * PAIR;
    static constexpr int RS = VEC ? 8 * (NF + NG) : 4 * NF;
    static constexpr int NT = VEC ? 2 : 1; // tables: (dP, mP/sin) or P
    // virtual column -> (level within the tile, column within the level)
    __device__ static void decode(int vcol, int &lv, int &fv)
    {
        const int g = vcol / GC, w = vcol - g * GC;
        if constexpr (PAIR == 1) {
            lv = g;
            fv = w;
        } else if (w < 8 * NF) {
            lv = 2 * g + w / (4 * NF);
            fv = w % (4 * NF);
        } else {
            const int w2 = w - 8 * NF;
            lv = 2 * g + w2 / (2 * NG);
            fv = 4 * NF + w2 % (2 * NG);
        }
    }
};

// Synthesis contraction for one (m, node tile, level tile), both parity
// groups: the K chunks (KC degrees) of the even group, then of the odd group,
// run as one software-pipelined sequence -- chunk c+1 is loaded into
// registers while the MFMAs consume chunk c.
//   A (tables, shared by the NW waves): double-buffered in LDS with a padded
//     row stride TMA = TM + 4 (the 4-row MFMA step then hits disjoint bank
//     halves); loaded as 16-byte vectors through per-thread pointers when
//     VL (nlat/2 and the node tile are multiples of 4), else element-wise.
//   B (spectral columns, one wave's 16 columns): straight into registers in
//     the MFMA operand layout.  Within a 16-degree chunk, lane group g
//     (lane / 16) takes degrees 4g .. 4g+3 over the chunk's 4 MFMA steps (the
//     same permutation indexes A), so each lane reads 4 consecutive degrees.
constexpr int TMA = TM + 4;
template <typename T, bool VEC, int NF, int NW, int KC, bool VL, int NG, int MM, class AT>
__device__ __forceinline__ void synth_pipeline(const AT &a, const T *__restrict__ in, const T *__restrict__ tab0,
                                               const T *__restrict__ tab1, T *sA, const int *scol, int m,
                                               int i0, int r0, acc_t<T> (&accE)[SynthTile<VEC, NF, NW, NG>::NT][2],
                                               acc_t<T> (&accO)[SynthTile<VEC, NF, NW, NG>::NT][2])
{
    using TL = SynthTile<VEC, NF, NW, NG>;
    static_assert(TL::NTILE % TL::GC == 0, "a block tile holds whole column groups");
    static_assert(KC % 16 == 0, "chunks are whole 16-degree MFMA groups");
    constexpr int NT = TL::NT, VC = TL::VC;
    constexpr int nthr = 64 * NW;
    constexpr int sAsz = NT * KC * TMA;
    constexpr int EA = VL ? 16 / (int)sizeof(T) : 1;       // elements per A load
    constexpr int NAE = NT * KC * TM / EA;                 // A loads per chunk
    constexpr int RA = (NAE + nthr - 1) / nthr;            // per thread
    constexpr int RB = KC / 4;                             // B registers per chunk
    using VA = typename std::conditional<VL, typename Vec16<T>::type, T>::type;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int K = a.moff[m + 1] - a.moff[m], ne = a.ne[m];
    const int nE = (ne + KC - 1) / KC, nC = nE + (K - ne + KC - 1) / KC;
    const int kb0 = a.moff[m];

    // per-thread A element: table t, degree row k of the chunk, node offset i
    const T *pa[RA];
    int ka[RA], sa[RA];
    bool va[RA];
#pragma unroll
    for (int q = 0; q < RA; ++q) {
        const int e = tid + q * nthr;
        const int t = e / (KC * TM / EA), rem = e - t * (KC * TM / EA);
        const int k = rem / (TM / EA), i = (rem - k * (TM / EA)) * EA;
        va[q] = e < NAE && i0 + i < a.nlat2;
        pa[q] = (t == 0 || !va[q] ? tab0 : tab1) + i0 + i;
        ka[q] = k;
        sa[q] = (t * KC + k) * TMA + i;
    }
    // this lane's B column
    const int c16 = lane & 15, g = lane >> 4;
    const int vcol = wave * 16 + c16;
    int lv, fv;
    TL::decode(vcol, lv, fv);
    const int col = scol[fv], r = r0 + lv;
    const bool vb = col >= 0 && r < a.nr;
    const T *pb = in + (size_t)(vb ? r * a.cin + col : 0) * a.nlm + kb0;

    VA ra[RA];
    T rb[RB];
    auto chunk = [&](int c, int &kbase, int &k0, int &Kp) {
        const int p = c < nE ? 0 : 1;
        k0 = (p ? c - nE : c) * KC;
        Kp = p ? K - ne : ne;
        kbase = p ? ne : 0; // relative to moff[m]
    };
    auto gload = [&](int c) {
        int kbase, k0, Kp;
        chunk(c, kbase, k0, Kp);
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            VA v;
            T *pv = reinterpret_cast<T *>(&v);
#pragma unroll
            for (int u = 0; u < EA; ++u) pv[u] = T(0);
            if (va[q] && k0 + ka[q] < Kp)
                v = *reinterpret_cast<const VA *>(pa[q] + (size_t)(kb0 + kbase + k0 + ka[q]) * a.nlat2);
            ra[q] = v;
        }
#pragma unroll
        for (int s = 0; s < RB; ++s) {
            // degree of step s: group h = s / 4 of 16, lane group g, step s % 4
            const int k = (s >> 2) * 16 + 4 * g + (s & 3);
            rb[s] = (vb && k0 + k < Kp) ? pb[kbase + k0 + k] : T(0);
        }
    };
    auto sstore = [&](int buf) {
        T *A = sA + buf * sAsz;
#pragma unroll
        for (int q = 0; q < RA; ++q)
            if (tid + q * nthr < NAE) *reinterpret_cast<VA *>(A + sa[q]) = ra[q];
    };
    auto mma = [&](int buf, const T (&bv)[RB], acc_t<T>(&acc)[NT][2]) {
        const T *A = sA + buf * sAsz;
        if constexpr (MM) {
            // BF16x3: the 4 steps of each 16-degree group as one split-precision
            // 16x16x16 BF16 contraction (same k order: lane group g, degrees 4g+j)
#pragma unroll
            for (int h = 0; h < KC / 16; ++h) {
                const Bf16Pair bp = bf16_split4(bv[4 * h], bv[4 * h + 1], bv[4 * h + 2], bv[4 * h + 3]);
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int rb_ = 0; rb_ < 2; ++rb_) {
                        const T *Ar = A + (t * KC + h * 16 + 4 * g) * TMA + rb_ * 16 + c16;
                        const Bf16Pair ap = bf16_split4(Ar[0], Ar[TMA], Ar[2 * TMA], Ar[3 * TMA]);
                        acc[t][rb_] = mfma_bf16x3(ap, bp, acc[t][rb_]);
                    }
            }
            return;
        }
#pragma unroll
        for (int s = 0; s < RB; ++s) {
            const int k = (s >> 2) * 16 + 4 * g + (s & 3);
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int rb_ = 0; rb_ < 2; ++rb_) {
                    const T av = A[(t * KC + k) * TMA + rb_ * 16 + c16];
                    acc[t][rb_] = mfma16x16x4(av, bv[s], acc[t][rb_]);
                }
        }
    };
    // chunk c: stage it, start loading c + 1, contract it into acc.  The even
    // chunks (c < nE) and the odd ones run as two loops, each with its own
    // accumulator set: a per-chunk choice between the sets made the compiler
    // copy every accumulator register on each iteration.
    auto step = [&](int c, acc_t<T>(&acc)[NT][2]) {
        sstore(c & 1);
        T bcur[RB];
#pragma unroll
        for (int s = 0; s < RB; ++s) bcur[s] = rb[s];
        __syncthreads();
        if (c + 1 < nC) gload(c + 1);
        mma(c & 1, bcur, acc);
    };
    if (nC > 0) gload(0);
    int c = 0;
    for (; c < nE; ++c) step(c, accE);
    for (; c < nC; ++c) step(c, accO);
}

template <typename T, bool VEC, int NF, int NW, int KC, int NG = 0> constexpr size_t synth_lds_bytes()
{
    using TL = SynthTile<VEC, NF, NW, NG>;
    const size_t pipe = 2 * (size_t)TL::NT * KC * TMA, tile = (size_t)TM * TL::LT * TL::RS;
    return 128 + sizeof(T) * (pipe > tile ? pipe : tile);
}

// The block's output tile -- TM nodes x LT levels, RS contiguous slots of each
// (node, level, m) record -- is staged in LDS and stored as 16-byte vectors,
// one record run after another.
// blk, nblk: the block's index and the block count of this synthesis (the
// launch's own, or its share of a batched launch over several shells)
template <typename T, bool VEC, int NF, int NW, int KC, bool VL, int NG, int MM, class AT>
__device__ __forceinline__ void synth_body(const AT &a, int blk, int nblk)
{
    static_assert(!MM || sizeof(T) == 4, "the BF16x3 contraction takes FP32 operands");
    using TL = SynthTile<VEC, NF, NW, NG>;
    constexpr int NT = TL::NT, VC = TL::VC, LT = TL::LT, RS = TL::RS;
    extern __shared__ __align__(16) char smem_raw[];
    int *scol = reinterpret_cast<int *>(smem_raw);
    T *sA = reinterpret_cast<T *>(smem_raw + 128);
    const int gx = (a.nlat2 + TM - 1) / TM, gy = (a.nr + LT - 1) / LT, nm = a.mmax + 1;
    const int L = xcd_logical(blk, nblk);
    const int bx = L % gx, by = (L / gx) % gy;
    const int m = paired_m(L / (gx * gy), nm), i0 = bx * TM, r0 = by * LT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < 20) scol[tid] = a.colmap[tid];
    __syncthreads();
    const T *in = static_cast<const T *>(a.in);
    const T *tab0 = static_cast<const T *>(a.tab0);
    const T *tab1 = static_cast<const T *>(a.tab1);
    acc_t<T> accE[NT][2], accO[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
            accE[t][rb] = acc_t<T>{0, 0, 0, 0};
            accO[t][rb] = acc_t<T>{0, 0, 0, 0};
        }
    synth_pipeline<T, VEC, NF, NW, KC, VL, NG, MM>(a, in, tab0, tab1, sA, scol, m, i0, r0, accE, accO);
    __syncthreads(); // the tile below reuses the pipeline's LDS

    T *tile = sA; // [node][level][slot]
    const int vcol = wave * 16 + (lane & 15);
    int lv, fv;
    TL::decode(vcol, lv, fv);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int il = rb * 16 + acc_row<T>(lane, v);
            T on, os;
            int srel;
            if constexpr (VEC) {
                // north / south parity combination: dP flips sign once more than P, Q
                const T dPn = accE[0][rb][v] + accO[0][rb][v], dPs = accO[0][rb][v] - accE[0][rb][v];
                const T Qn = accE[1][rb][v] + accO[1][rb][v], Qs = accE[1][rb][v] - accO[1][rb][v];
                const T Qxn = shfl_xor3(Qn), Qxs = shfl_xor3(Qs);
                const int q = lane & 3; // 0 S_re, 1 S_im, 2 T_re, 3 T_im
                if (NG > 0 && fv >= 4 * NF) {
                    // gradient field: theta = dP S, phi = i Q S (Q S_re -> phi_im,
                    // -Q S_im -> phi_re); the values of the full form with T = 0
                    const int gq = fv - 4 * NF, q2 = gq & 1;
                    T *tg = tile + (il * LT + lv) * RS + 8 * (NF + (gq >> 1));
                    tg[q2] = dPn;
                    tg[4 + q2] = dPs;
                    tg[q2 ? 2 : 3] = q2 ? -Qn : Qn;
                    tg[q2 ? 6 : 7] = q2 ? -Qs : Qs;
                    continue;
                }
                if (q == 0) {
                    on = dPn - Qxn;
                    os = dPs - Qxs;
                } else if (q == 1) {
                    on = dPn + Qxn;
                    os = dPs + Qxs;
                } else if (q == 2) {
                    on = -Qxn - dPn;
                    os = -Qxs - dPs;
                } else {
                    on = Qxn - dPn;
                    os = Qxs - dPs;
                }
                srel = 8 * (fv >> 2) + q; // + 4 for south
            } else {
                on = accE[0][rb][v] + accO[0][rb][v];
                os = accE[0][rb][v] - accO[0][rb][v];
                srel = 4 * (fv >> 1) + (fv & 1); // + 2 for south
            }
            T *t = tile + (il * LT + lv) * RS + srel;
            t[0] = on;
            t[VEC ? 4 : 2] = os;
        }
    }
    __syncthreads();
    using V = typename Vec16<T>::type;
    constexpr int EV = 16 / sizeof(T);
    constexpr int NV = TM * LT * RS / EV;
    T *out = static_cast<T *>(a.out);
    for (int t = tid; t < NV; t += 64 * NW) {
        const int e = t * EV;
        const int rec = e / RS, so = e - rec * RS;
        const int il = rec / LT, l2 = rec - il * LT;
        const int i = i0 + il, r = r0 + l2;
        if (i < a.nlat2 && r < a.nr) {
            T *dst = out + (((size_t)i * a.nr + r) * nm + m) * a.nslot + a.slot0 + so;
            *reinterpret_cast<V *>(dst) = *reinterpret_cast<const V *>(tile + e);
        }
    }
}

template <typename T, bool VEC, int NF, int NW, int KC, bool VL, int NG = 0, int MM = 0>
__global__ void __launch_bounds__(64 * NW) synth_kernel(SynthArgs a)
{
    synth_body<T, VEC, NF, NW, KC, VL, NG, MM>(a, blockIdx.x, gridDim.x);
}

// one grid over the syntheses of several shells (a batched step)
template <typename T, bool VEC, int NF, int NW, int KC, bool VL, int NG = 0, int MM = 0>
__global__ void __launch_bounds__(64 * NW) synth_batch_kernel(BatchRef<SynthArgs> b)
{
    const int sh = b.shell_of_block[blockIdx.x];
    if (sh == BATCH_PAD) return;
    synth_body<T, VEC, NF, NW, KC, VL, NG, MM>(batch_args(b, sh), blockIdx.x - b.block_off[sh], b.block_cnt[sh]);
}

// ---------------------------------------------------------------- analysis
template <typename T, bool VEC, int NF> __global__ void __launch_bounds__(64 * NF) ana_kernel(AnaArgs a)
{
    constexpr int NT = VEC ? 2 : 1;
    extern __shared__ __align__(16) char smem_raw[];
    constexpr int VC = VEC ? 4 * NF : 2 * NF;
    constexpr int NTILE = 16 * NF;
    constexpr int LT = NTILE / VC;
    static_assert(NTILE * (TM + 1) <= 2 * NT * KT * (TM + NTILE), "analysis tile exceeds the pipeline LDS");
    T *sA = reinterpret_cast<T *>(smem_raw);
    T *sB = sA + 2 * NT * KT * TM; // A and B are double-buffered
    const int gx = (a.maxgroup + TM - 1) / TM, gy = (a.nr + LT - 1) / LT, nmz = a.mmax + 1;
    const int L = xcd_logical(blockIdx.x, gx * gy * 2 * nmz);
    const int bx = L % gx, by = (L / gx) % gy, bz = L / (gx * gy);
    const int m = paired_m(bz >> 1, nmz), p = bz & 1;
    const int K = a.moff[m + 1] - a.moff[m], ne = a.ne[m];
    const int Kp = p ? K - ne : ne;
    const int kbase = a.moff[m] + (p ? ne : 0);
    const int row0 = bx * TM, r0 = by * LT;
    if (row0 >= Kp) return; // uniform over the block
    const int tid = threadIdx.x, nthr = blockDim.x, wave = tid >> 6, lane = tid & 63;
    const T *gb = static_cast<const T *>(a.gb);
    const T *tab0 = static_cast<const T *>(a.tab0), *tab1 = static_cast<const T *>(a.tab1);
    acc_t<T> acc[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) acc[t][rb] = acc_t<T>{0, 0, 0, 0};

    // software-pipelined K loop over the northern nodes (see synth_pipeline)
    constexpr int RA = (NT * KT * TM + 64 * NF - 1) / (64 * NF);
    constexpr int RB = NT * KT * 16 / 64;
    constexpr int sAsz = NT * KT * TM, sBsz = NT * KT * NTILE;
    const int nC = (a.nlat2 + KT - 1) / KT;
    T ra[RA], rb[RB];
    auto gload = [&](int c) {
        const int k0 = c * KT;
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int e = tid + q * nthr;
            T v = T(0);
            if (e < sAsz) {
                const int kk = e % KT, i = (e / KT) % TM, t = e / (KT * TM);
                const int row = row0 + i;
                if (row < Kp && k0 + kk < a.nlat2) v = (t == 0 ? tab0 : tab1)[(size_t)(kbase + row) * a.nlat2 + k0 + kk];
            }
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int e = tid + q * nthr;
            T v = T(0);
            if (e < sBsz) {
                const int vv = e % NTILE, kk = (e / NTILE) % KT, t = e / (NTILE * KT);
                const int lv = vv / VC, fv = vv - lv * VC, r = r0 + lv;
                int comp, reim, pm;
                if (VEC) {
                    const int V = fv >> 2, q4 = fv & 3;
                    comp = 2 * V + (q4 >> 1);
                    reim = q4 & 1;
                    // dP has southern sign -(-1)^(l+m): even group uses G-, odd G+; Q the reverse
                    pm = (t == 0) ? (p == 0 ? 1 : 0) : p;
                } else {
                    comp = fv >> 1;
                    reim = fv & 1;
                    pm = p;
                }
                if (r < a.nr && k0 + kk < a.nlat2)
                    v = gb[(((size_t)m * a.nlat2 + k0 + kk) * a.nr + r) * a.nslot_g + a.gslot0 + 4 * comp + 2 * pm + reim];
            }
            rb[q] = v;
        }
    };
    auto sstore = [&](int b) {
        T *A = sA + b * sAsz, *B = sB + b * sBsz;
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int e = tid + q * nthr;
            if (e < sAsz) {
                const int kk = e % KT, i = (e / KT) % TM, t = e / (KT * TM);
                A[(t * KT + kk) * TM + i] = ra[q];
            }
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int e = tid + q * nthr;
            if (e < sBsz) B[e] = rb[q];
        }
    };
    if (nC > 0) gload(0);
    for (int c = 0; c < nC; ++c) {
        sstore(c & 1);
        __syncthreads();
        if (c + 1 < nC) gload(c + 1);
        const T *A = sA + (c & 1) * sAsz, *B = sB + (c & 1) * sBsz;
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const T b = B[(t * KT + ks + (lane >> 4)) * NTILE + wave * 16 + (lane & 15)];
#pragma unroll
                for (int rb_ = 0; rb_ < 2; ++rb_) {
                    const T av = A[(t * KT + ks + (lane >> 4)) * TM + rb_ * 16 + (lane & 15)];
                    acc[t][rb_] = mfma16x16x4(av, b, acc[t][rb_]);
                }
            }
        }
    }

    // stage the (row x column) tile in LDS, padded against bank conflicts,
    // then store each output column's rows (consecutive idx) contiguously
    __syncthreads(); // the tile reuses the pipeline's LDS
    constexpr int TMP = TM + 1;
    T *tile = sA; // [column][row]
    const int vcol = wave * 16 + (lane & 15);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = rb * 16 + acc_row<T>(lane, v);
            T o;
            if constexpr (VEC) {
                const T X = acc[0][rb][v], Y = acc[1][rb][v];
                const T Yx = shfl_xor3(Y);
                const int q = lane & 3;
                if (q == 0) o = X + Yx;       // S_re
                else if (q == 1) o = X - Yx;  // S_im
                else if (q == 2) o = Yx - X;  // T_re
                else o = -Yx - X;             // T_im
            } else {
                o = acc[0][rb][v];
            }
            tile[vcol * TMP + row] = o;
        }
    }
    __syncthreads();
    const size_t fs = (size_t)a.nr * a.nlm;
    for (int t = tid; t < NTILE * TM; t += nthr) {
        const int col = t / TM, row = t - col * TM;
        const int lv = col / VC, fv = col - lv * VC, r = r0 + lv;
        const int grow = row0 + row;
        if (grow < Kp && r < a.nr) {
            int outk, reim;
            if constexpr (VEC) {
                outk = 2 * (fv >> 2) + ((fv & 3) >> 1);
                reim = fv & 1;
            } else {
                outk = fv >> 1;
                reim = fv & 1;
            }
            double val = (double)tile[col * TMP + row];
            const int idx = kbase + grow;
            const int l = a.l_of_idx[idx];
            const int sc = outk == 0 ? a.scale[0] : outk == 1 ? a.scale[1] : outk == 2 ? a.scale[2] : a.scale[3];
            double *op = outk == 0 ? a.out[0] : outk == 1 ? a.out[1] : outk == 2 ? a.out[2] : a.out[3];
            if (sc == 0) val = (l > 0) ? val / ((double)l * (double)(l + 1)) : 0.0;
            else if (sc == 2) val = -val;
            if (m == 0 && reim) val = 0.0; // imag(m=0) := 0 (shtnskit_field_functions.jl:432-435)
            op[(size_t)reim * fs + (size_t)r * a.nlm + idx] = val;
        }
    }
}
// Analysis, table-resident form: one block of 4 waves per (row tile of TM
// degrees, m, parity).  The block's table tile -- NT tables x TM degrees x all
// KN (>= nlat/2, zero-padded) northern nodes -- is read from HBM once, in one
// batch of loads, into LDS ([t][node][degree], padded rows).  The waves then
// take the level tiles in turn (wave w: tiles w, w+4, ...); a wave keeps the
// Fourier operand G of its tile in registers, one element per lane and
// 4-node step, and reloads each register for its next tile as soon as the
// MFMA that used it has issued, so the next tile's loads overlap this tile's
// MFMAs.  STAGE: results go through a per-wave LDS tile and each output
// column's consecutive idx are stored contiguously.
// waves per analysis block (8 was measured slower at lmax 127 for FP32, 41 ->
// 48 us: a wave's next-tile operand prefetch then spans fewer tiles; for FP64
// it would exceed the register budget of a 512-thread block)
template <typename T> constexpr int ana_waves() { return 4; }
template <typename T, bool VEC, int NF, int KN, int MM, class AT>
__device__ __forceinline__ void ana_res_body(const AT &a, int blk, int nblk)
{
    static_assert(!MM || sizeof(T) == 4, "the BF16x3 contraction takes FP32 operands");
    constexpr int ANA_WAVES = ana_waves<T>();
    constexpr int NT = VEC ? 2 : 1;
    constexpr int VC = VEC ? 4 * NF : 2 * NF;
    constexpr int LT = 16 / VC; // levels per wave tile (16 columns)
    constexpr int KS = KN / 4;
    constexpr int TMP = TM + 1; // output staging row stride
    constexpr bool STAGE = sizeof(T) == 4;
    constexpr int NA = NT * TM * KN / (64 * ANA_WAVES);
    static_assert(NT * TM * KN % (64 * ANA_WAVES) == 0, "table tile must split evenly over the block");
    static_assert(KN % 16 == 0, "nodes come in 16-node MFMA groups");
    extern __shared__ __align__(16) char smem_raw[];
    T *sA = reinterpret_cast<T *>(smem_raw);    // [t][node][TMA]
    T *stg = sA + NT * KN * TMA;                 // [wave][16][TMP] when STAGE
    __shared__ double sinv[TM];                  // 1 / (l(l+1)) of the tile's rows (0 at l = 0)
    const int gx = (a.maxgroup + TM - 1) / TM, nmz = a.mmax + 1;
    const int L = xcd_logical(blk, nblk);
    const int bx = L % gx, bz = L / gx;
    const int m = paired_m(bz >> 1, nmz), p = bz & 1;
    const int K = a.moff[m + 1] - a.moff[m], ne = a.ne[m];
    const int Kp = p ? K - ne : ne;
    const int kbase = a.moff[m] + (p ? ne : 0);
    const int row0 = bx * TM;
    if (row0 >= Kp) return; // uniform over the block
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const T *gb = static_cast<const T *>(a.gb);
    const T *tab0 = static_cast<const T *>(a.tab0), *tab1 = static_cast<const T *>(a.tab1);
    const int k0 = a.k0, nk = a.nk; // this launch's node chunk
    const bool accum = a.accum != 0;

    if (tid < TM) {
        const int l = row0 + tid < Kp ? a.l_of_idx[kbase + row0 + tid] : 0;
        sinv[tid] = l > 0 ? 1.0 / ((double)l * (double)(l + 1)) : 0.0;
    }
    // table tile: all loads in flight together, then into LDS
    // 16-byte loads of EV consecutive nodes where every table row is 16-byte
    // aligned (nlat/2 a multiple of EV, the chunk start a multiple of 16)
    constexpr int EV = 16 / (int)sizeof(T);
    constexpr int NAV = (NT * TM * KN / EV + 64 * ANA_WAVES - 1) / (64 * ANA_WAVES);
    using VT = typename Vec16<T>::type;
    const bool vl = a.nlat2 % EV == 0 && nk % EV == 0 && ((uintptr_t)tab0 % 16) == 0 &&
                    (NT == 1 || ((uintptr_t)tab1 % 16) == 0);
    if (vl) {
        VT ra[NAV];
#pragma unroll
        for (int q = 0; q < NAV; ++q) {
            const int e = tid + q * 64 * ANA_WAVES;
            const int t = e / (TM * KN / EV), rem = e - t * (TM * KN / EV);
            const int i = rem / (KN / EV), k = (rem - i * (KN / EV)) * EV;
            VT v;
            T *pv = reinterpret_cast<T *>(&v);
#pragma unroll
            for (int u = 0; u < EV; ++u) pv[u] = T(0);
            if (e < NT * TM * KN / EV && row0 + i < Kp && k < nk)
                v = *reinterpret_cast<const VT *>((t == 0 ? tab0 : tab1) + (size_t)(kbase + row0 + i) * a.nlat2 + k0 + k);
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < NAV; ++q) {
            const int e = tid + q * 64 * ANA_WAVES;
            const int t = e / (TM * KN / EV), rem = e - t * (TM * KN / EV);
            const int i = rem / (KN / EV), k = (rem - i * (KN / EV)) * EV;
            const T *pv = reinterpret_cast<const T *>(&ra[q]);
            if (e < NT * TM * KN / EV)
#pragma unroll
                for (int u = 0; u < EV; ++u) sA[(t * KN + k + u) * TMA + i] = pv[u];
        }
    } else {
        T ra[NA];
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + q * 64 * ANA_WAVES;
            const int t = e / (TM * KN), rem = e - t * (TM * KN);
            const int i = rem / KN, k = rem - i * KN;
            T v = T(0);
            if (row0 + i < Kp && k < nk) v = (t == 0 ? tab0 : tab1)[(size_t)(kbase + row0 + i) * a.nlat2 + k0 + k];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + q * 64 * ANA_WAVES;
            const int t = e / (TM * KN), rem = e - t * (TM * KN);
            const int i = rem / KN, k = rem - i * KN;
            sA[(t * KN + k) * TMA + i] = ra[q];
        }
    }

    // this lane's column of a wave tile, and its G slot for table t
    const int c = lane & 15, kq = lane >> 4;
    const int lv = c / VC, fv = c - lv * VC;
    int slot[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        int comp, reim, pm;
        if constexpr (VEC) {
            const int V = fv >> 2, q4 = fv & 3;
            comp = 2 * V + (q4 >> 1);
            reim = q4 & 1;
            // dP has southern sign -(-1)^(l+m): even group uses G-, odd G+; Q the reverse
            pm = (t == 0) ? (p == 0 ? 1 : 0) : p;
        } else {
            comp = fv >> 1;
            reim = fv & 1;
            pm = p;
        }
        slot[t] = a.gslot0 + 4 * comp + 2 * pm + reim;
    }
    const int ntiles = (a.nr + LT - 1) / LT;
    const int niter = (ntiles + ANA_WAVES - 1) / ANA_WAVES;
    // node of MFMA step ks for this lane: within each group of 16 nodes, lane
    // group kq takes nodes 4kq .. 4kq+3 over 4 steps (A is indexed the same way)
    auto node = [&](int ks) { return (ks >> 2) * 16 + 4 * kq + (ks & 3); };
    const size_t nstride = (size_t)a.nr * a.nslot_g; // one node further in GB
    const T *gm = gb + (size_t)m * a.nlat2 * nstride;
    // Unconditional at clamped indices (a guarded load becomes a branch, and
    // after branches that may skip loads the compiler waits for all of them
    // at the next use): a node past the chunk meets table rows that are zero
    // in LDS, and a level past nr (or a tile past the last) belongs to an
    // output column that is never stored.
    auto gload = [&](int tile, int t, int ks) {
        const int r = min(tile * LT + lv, a.nr - 1), k = min(node(ks), nk - 1);
        return gm[(size_t)(k0 + k) * nstride + (size_t)r * a.nslot_g + slot[t]];
    };
    T b[NT][KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int t = 0; t < NT; ++t) b[t][ks] = gload(wave, t, ks);
    __syncthreads(); // table tile in LDS

    const size_t fs = (size_t)a.nr * a.nlm;
    T *wst = stg + wave * 16 * TMP;
    for (int it = 0; it < niter; ++it) {
        const int tile = it * ANA_WAVES + wave, next = tile + ANA_WAVES;
        acc_t<T> acc[NT][2];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t][0] = acc[t][1] = acc_t<T>{0, 0, 0, 0};
        if constexpr (MM) {
            // BF16x3: each 16-node group (4 steps, lane group kq on nodes 4kq+j)
            // as one split-precision 16x16x16 BF16 contraction
#pragma unroll
            for (int h = 0; h < KS / 4; ++h) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const Bf16Pair bp = bf16_split4(b[t][4 * h], b[t][4 * h + 1], b[t][4 * h + 2], b[t][4 * h + 3]);
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb) {
                        const T *A0 = sA + (t * KN + 16 * h + 4 * kq) * TMA + rb * 16 + c;
                        const Bf16Pair ap = bf16_split4(A0[0], A0[TMA], A0[2 * TMA], A0[3 * TMA]);
                        acc[t][rb] = mfma_bf16x3(ap, bp, acc[t][rb]);
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) b[t][4 * h + j] = gload(next, t, 4 * h + j);
                }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const T *Ak = sA + (t * KN + node(ks)) * TMA;
                    acc[t][0] = mfma16x16x4(Ak[c], b[t][ks], acc[t][0]);
                    acc[t][1] = mfma16x16x4(Ak[16 + c], b[t][ks], acc[t][1]);
                    b[t][ks] = gload(next, t, ks);
                }
            }
        }
        if (tile >= ntiles) continue; // uniform over the wave
        const int r = tile * LT + lv;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb * 16 + acc_row<T>(lane, v);
                T o;
                if constexpr (VEC) {
                    const T X = acc[0][rb][v], Y = acc[1][rb][v];
                    const T Yx = shfl_xor3(Y);
                    const int q = lane & 3;
                    if (q == 0) o = X + Yx;       // S_re
                    else if (q == 1) o = X - Yx;  // S_im
                    else if (q == 2) o = Yx - X;  // T_re
                    else o = -Yx - X;             // T_im
                } else {
                    o = acc[0][rb][v];
                }
                if constexpr (STAGE) {
                    wst[c * TMP + row] = o;
                } else {
                    const int grow = row0 + row;
                    if (grow < Kp && r < a.nr) {
                        int outk, reim;
                        if constexpr (VEC) {
                            outk = 2 * (fv >> 2) + ((fv & 3) >> 1);
                            reim = fv & 1;
                        } else {
                            outk = fv >> 1;
                            reim = fv & 1;
                        }
                        double val = (double)o;
                        const int idx = kbase + grow;
                        const int sc = a.scale[outk];
                        if (sc == 0) val *= sinv[row];
                        else if (sc == 2) val = -val;
                        if (m == 0 && reim) val = 0.0; // imag(m=0) := 0 (shtnskit_field_functions.jl:432-435)
                        double *dst = a.out[outk] + (size_t)reim * fs + (size_t)r * a.nlm + idx;
                        *dst = accum ? *dst + val : val;
                    }
                }
            }
        }
        if constexpr (STAGE) {
            // the wave reads back what its own lanes wrote: LDS executes one
            // wave's accesses in order; the fence keeps the compiler from
            // reordering them
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < 16 * TM / 64; ++q) {
                const int e = lane + 64 * q;
                const int col = e / TM, row = e - col * TM;
                const int lv2 = col / VC, fv2 = col - lv2 * VC, r2 = tile * LT + lv2;
                const int grow = row0 + row;
                if (grow < Kp && r2 < a.nr) {
                    int outk, reim;
                    if constexpr (VEC) {
                        outk = 2 * (fv2 >> 2) + ((fv2 & 3) >> 1);
                        reim = fv2 & 1;
                    } else {
                        outk = fv2 >> 1;
                        reim = fv2 & 1;
                    }
                    double val = (double)wst[col * TMP + row];
                    const int idx = kbase + grow;
                    const int sc = a.scale[outk];
                    if (sc == 0) val *= sinv[row];
                    else if (sc == 2) val = -val;
                    if (m == 0 && reim) val = 0.0; // imag(m=0) := 0 (shtnskit_field_functions.jl:432-435)
                    double *dst = a.out[outk] + (size_t)reim * fs + (size_t)r2 * a.nlm + idx;
                    *dst = accum ? *dst + val : val;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
}

template <typename T, bool VEC, int NF, int KN, int MM = 0>
__global__ void __launch_bounds__(64 * ana_waves<T>()) ana_res_kernel(AnaArgs a)
{
    ana_res_body<T, VEC, NF, KN, MM>(a, blockIdx.x, gridDim.x);
}

// one grid over the analyses of several shells (a batched step)
template <typename T, bool VEC, int NF, int KN, int MM = 0>
__global__ void __launch_bounds__(64 * ana_waves<T>()) ana_res_batch_kernel(BatchRef<AnaArgs> b)
{
    const int sh = b.shell_of_block[blockIdx.x];
    if (sh == BATCH_PAD) return;
    ana_res_body<T, VEC, NF, KN, MM>(batch_args(b, sh), blockIdx.x - b.block_off[sh], b.block_cnt[sh]);
}

template <typename T, bool VEC, int NF, int KN> constexpr size_t ana_res_lds()
{
    constexpr int NT = VEC ? 2 : 1;
    return sizeof(T) * ((size_t)NT * KN * TMA + (sizeof(T) == 4 ? (size_t)ana_waves<T>() * 16 * (TM + 1) : 0));
}

} // namespace

template <typename T, bool VEC, int NF, int NW, int KC, int NG, int MM>
static void launch_synth_mm(hipStream_t st, const SynthArgs &a)
{
    using TL = SynthTile<VEC, NF, NW, NG>;
    dim3 grid(((a.nlat2 + TM - 1) / TM) * ((a.nr + TL::LT - 1) / TL::LT) * (a.mmax + 1));
    constexpr size_t lds = synth_lds_bytes<T, VEC, NF, NW, KC, NG>();
    // 16-byte table loads need every node row and tile start 16-byte aligned
    if (a.nlat2 % 4 == 0 && ((uintptr_t)a.tab0 % 16) == 0 && ((uintptr_t)a.tab1 % 16) == 0)
        hipLaunchKernelGGL((synth_kernel<T, VEC, NF, NW, KC, true, NG, MM>), grid, dim3(64 * NW), lds, st, a);
    else
        hipLaunchKernelGGL((synth_kernel<T, VEC, NF, NW, KC, false, NG, MM>), grid, dim3(64 * NW), lds, st, a);
}

// a.mm = 1: the BF16x3 contraction (FP32 operands only)
template <typename T, bool VEC, int NF, int NW, int KC, int NG = 0>
static void launch_synth(hipStream_t st, const SynthArgs &a)
{
    if constexpr (sizeof(T) == 4) {
        if (a.mm) {
            launch_synth_mm<T, VEC, NF, NW, KC, NG, 1>(st, a);
            return;
        }
    }
    launch_synth_mm<T, VEC, NF, NW, KC, NG, 0>(st, a);
}

// The staged tile is stored in 16-byte vectors: record stride and first slot
// must be multiples of 4 elements.
static int synth_check(const SynthArgs &a, const char *what)
{
    GD_CHECK(a.nslot % 4 == 0 && a.slot0 % 4 == 0, GD_ERR_ARG, "%s: nslot=%d slot0=%d not multiples of 4", what,
             a.nslot, a.slot0);
    return GD_OK;
}

template <typename T> int gd_launch_synth_vec(hipStream_t st, const SynthArgs &a)
{
    GD_CHECK(!a.mm || sizeof(T) == 4, GD_ERR_ARG, "synth_vec: the BF16x3 contraction takes FP32 operands");
    GD_TRY(synth_check(a, "synth_vec"));
    GD_CHECK(a.slot0 + 8 * (a.nf + a.ng) <= a.nslot, GD_ERR_ARG, "synth_vec: slots %d+%d exceed %d", a.slot0,
             8 * (a.nf + a.ng), a.nslot);
    // one full field with 1 or 2 gradient fields: 6 columns per level (two
    // levels per 12-column group, 3 waves = 8 levels) or 8 (4 waves = 8 levels)
    if (a.ng == 1 && a.nf == 1) launch_synth<T, true, 1, 3, 16, 1>(st, a);
    else if (a.ng == 2 && a.nf == 1) launch_synth<T, true, 1, 4, 16, 2>(st, a);
    else if (a.ng == 1 && a.nf == 2) launch_synth<T, true, 2, 5, 16, 1>(st, a); // 10 columns, 5 waves = 8 levels
    else if (a.ng != 0) GD_CHECK(false, GD_ERR_ARG, "synth_vec: nf=%d ng=%d", a.nf, a.ng);
    else if (a.nf == 5) launch_synth<T, true, 5, 5, 16>(st, a);
    else if (a.nf == 3) launch_synth<T, true, 3, 3, 16>(st, a); // 12 columns per level, 3 waves = 4 levels
    else if (a.nf == 2) launch_synth<T, true, 2, 4, 16>(st, a);
    else if (a.nf == 1) launch_synth<T, true, 1, 1, 16>(st, a);
    else GD_CHECK(false, GD_ERR_ARG, "synth_vec: nf=%d (1, 2, 3 or 5)", a.nf);
    return gd_check_launch("synth_kernel<vec>");
}

template <typename T> int gd_launch_synth_scal(hipStream_t st, const SynthArgs &a)
{
    GD_CHECK(!a.mm || sizeof(T) == 4, GD_ERR_ARG, "synth_scal: the BF16x3 contraction takes FP32 operands");
    GD_TRY(synth_check(a, "synth_scal"));
    GD_CHECK(a.slot0 + 4 * a.nf <= a.nslot, GD_ERR_ARG, "synth_scal: slots %d+%d exceed %d", a.slot0, 4 * a.nf, a.nslot);
    if (a.nf == 5) launch_synth<T, false, 5, 5, 16>(st, a);
    else if (a.nf == 1) launch_synth<T, false, 1, 1, 16>(st, a);
    else GD_CHECK(false, GD_ERR_ARG, "synth_scal: nf=%d (1 or 5)", a.nf);
    return gd_check_launch("synth_kernel<scal>");
}

// grid rows cover the largest parity group (maxgroup = ceil((lmax+1)/2))
// table-resident analysis for nlat/2 <= KN; false when no instance applies
// nodes in chunks of at most KN: the first launch stores, the others add (the
// sum over nodes is linear; the per-row factors, signs and imag(m=0) := 0
// apply to every partial alike)
template <typename T, bool VEC, int NF, int KN, int MM = 0> static bool try_ana_res(hipStream_t st, const AnaArgs &a0)
{
    constexpr size_t lds = ana_res_lds<T, VEC, NF, KN>();
    if (lds > 64 * 1024) return false;
    const int gx = (a0.maxgroup + TM - 1) / TM;
    const int nchunk = (a0.nlat2 + KN - 1) / KN;
    for (int c = 0; c < nchunk; ++c) {
        AnaArgs a = a0;
        a.k0 = c * KN;
        a.nk = min(KN, a0.nlat2 - a.k0);
        a.accum = c > 0;
        hipLaunchKernelGGL((ana_res_kernel<T, VEC, NF, KN, MM>), dim3(gx * 2 * (a.mmax + 1)),
                           dim3(64 * ana_waves<T>()), lds, st, a);
    }
    return true;
}

template <typename T, bool VEC, int NF, int MM> static bool ana_res_mm(hipStream_t st, const AnaArgs &a)
{
    if (a.nlat2 <= 32) return try_ana_res<T, VEC, NF, 32, MM>(st, a);
    if (a.nlat2 <= 48) return try_ana_res<T, VEC, NF, 48, MM>(st, a);
    if (a.nlat2 <= 96) return try_ana_res<T, VEC, NF, 96, MM>(st, a);
    if constexpr (sizeof(T) == 4) // the FP64 operand registers of a 128-node tile would spill
        if (a.nlat2 <= 128) return try_ana_res<T, VEC, NF, 128, MM>(st, a);
    return try_ana_res<T, VEC, NF, 96, MM>(st, a); // node chunks of 96
}

template <typename T, bool VEC, int NF> static bool ana_res(hipStream_t st, const AnaArgs &a)
{
    if constexpr (sizeof(T) == 4)
        if (a.mm) return ana_res_mm<T, VEC, NF, 1>(st, a);
    return ana_res_mm<T, VEC, NF, 0>(st, a);
}

template <typename T> static int launch_ana(hipStream_t st, const AnaArgs &a, bool vec, int maxgroup)
{
    if (vec && a.nf == 2 && ana_res<T, true, 2>(st, a)) return gd_check_launch("ana_res_kernel<vec>");
    if (vec && a.nf == 1 && ana_res<T, true, 1>(st, a)) return gd_check_launch("ana_res_kernel<vec>");
    if (!vec && a.nf == 1 && ana_res<T, false, 1>(st, a)) return gd_check_launch("ana_res_kernel<scal>");
    if (!vec && a.nf == 2 && ana_res<T, false, 2>(st, a)) return gd_check_launch("ana_res_kernel<scal>");
    // streaming form (large nlat, or LDS beyond 64 KB)
    GD_CHECK(!a.mm, GD_ERR_ARG, "ana: the BF16x3 contraction has no streaming form (nlat/2=%d)", a.nlat2);
    const int VC = vec ? 4 * a.nf : 2 * a.nf;
    const int LT = 16 * a.nf / VC;
    dim3 grid(((maxgroup + TM - 1) / TM) * ((a.nr + LT - 1) / LT) * 2 * (a.mmax + 1));
    const int NT = vec ? 2 : 1;
    size_t lds = sizeof(T) * 2 * (NT * KT * TM + NT * KT * 16 * a.nf);
    if (vec && a.nf == 2) hipLaunchKernelGGL((ana_kernel<T, true, 2>), grid, dim3(128), lds, st, a);
    else if (vec && a.nf == 1) hipLaunchKernelGGL((ana_kernel<T, true, 1>), grid, dim3(64), lds, st, a);
    else if (!vec && a.nf == 1) hipLaunchKernelGGL((ana_kernel<T, false, 1>), grid, dim3(64), lds, st, a);
    else GD_CHECK(false, GD_ERR_ARG, "ana: nf=%d", a.nf);
    return gd_check_launch(vec ? "ana_kernel<vec>" : "ana_kernel<scal>");
}

template <typename T> int gd_launch_ana(hipStream_t st, const AnaArgs &a, bool vec, int maxgroup)
{
    GD_CHECK(!a.mm || sizeof(T) == 4, GD_ERR_ARG, "ana: the BF16x3 contraction takes FP32 operands");
    GD_CHECK(a.nf >= 1 && a.nf <= (vec ? 2 : 5), GD_ERR_ARG, "ana: nf=%d", a.nf);
    AnaArgs b = a;
    b.maxgroup = maxgroup;
    return launch_ana<T>(st, b, vec, maxgroup);
}

// ---------------------------------------------------------------- batched step launches
// The tangential step's synthesis (u as a full field, grad T as a gradient
// field) in one grid over several shells, with a tile chosen per group:
//   tile 0: 3 waves (8 levels), K chunks of 16 degrees (the step's own form)
//   tile 1: 3 waves, K chunks of 32 degrees
//   tile 2: 6 waves (16 levels), K chunks of 16 (FP32 only: the FP64 output
//           tile would exceed 64 KB of LDS)
static int synth_tile_nw(int tile) { return tile == 2 ? 6 : 3; }
static int synth_tile_kc(int tile) { return tile == 1 ? 32 : 16; }

int gd_synth_tang_blocks(const SynthArgs &a, int tile)
{
    const int nw = synth_tile_nw(tile), lt = 16 * nw / 6; // 6 columns per level, two levels per 12-column group
    return ((a.nlat2 + TM - 1) / TM) * ((a.nr + lt - 1) / lt) * (a.mmax + 1);
}

bool gd_synth_tile_ok(int tile, bool fp64) { return tile == 0 || tile == 1 || (tile == 2 && !fp64); }

template <typename T, int NW, int KC, int MM>
static void synth_batch_launch(hipStream_t st, const BatchRef<SynthArgs> &b, int nblocks)
{
    constexpr size_t lds = synth_lds_bytes<T, true, 1, NW, KC, 1>();
    hipLaunchKernelGGL((synth_batch_kernel<T, true, 1, NW, KC, true, 1, MM>), dim3(nblocks), dim3(64 * NW), lds, st, b);
}

template <typename T>
int gd_launch_synth_tang_batch(hipStream_t st, const BatchRef<SynthArgs> &b, int nblocks, int tile, int mm)
{
    GD_CHECK(gd_synth_tile_ok(tile, sizeof(T) == 8), GD_ERR_ARG, "synth batch: tile %d", tile);
    GD_CHECK(!mm || sizeof(T) == 4, GD_ERR_ARG, "synth batch: the BF16x3 contraction takes FP32 operands");
    if (nblocks <= 0) return GD_OK;
#define GD_SB(NW, KC)                                                                   \
    do {                                                                                \
        if constexpr (sizeof(T) == 4) {                                                 \
            if (mm) synth_batch_launch<T, NW, KC, 1>(st, b, nblocks);                   \
            else synth_batch_launch<T, NW, KC, 0>(st, b, nblocks);                      \
        } else {                                                                        \
            synth_batch_launch<T, NW, KC, 0>(st, b, nblocks);                           \
        }                                                                               \
    } while (0)
    if (tile == 0) GD_SB(3, 16);
    else if (tile == 1) GD_SB(3, 32);
    else if constexpr (sizeof(T) == 4) GD_SB(6, 16);
#undef GD_SB
    return gd_check_launch("synth_batch_kernel");
}
template int gd_launch_synth_tang_batch<float>(hipStream_t, const BatchRef<SynthArgs> &, int, int, int);
template int gd_launch_synth_tang_batch<double>(hipStream_t, const BatchRef<SynthArgs> &, int, int, int);

// one shell's synthesis in tile form `tile` (the autotuner times these alone)
template <typename T> int gd_launch_synth_tang_tile(hipStream_t st, const SynthArgs &a, int tile)
{
    GD_CHECK(gd_synth_tile_ok(tile, sizeof(T) == 8), GD_ERR_ARG, "synth: tile %d", tile);
    GD_CHECK(a.nf == 1 && a.ng == 1, GD_ERR_ARG, "synth tile: the tangential form (nf=1, ng=1)");
    if (tile == 0) launch_synth<T, true, 1, 3, 16, 1>(st, a);
    else if (tile == 1) launch_synth<T, true, 1, 3, 32, 1>(st, a);
    else if constexpr (sizeof(T) == 4) launch_synth<T, true, 1, 6, 16, 1>(st, a);
    return gd_check_launch("synth_kernel<tile>");
}
template int gd_launch_synth_tang_tile<float>(hipStream_t, const SynthArgs &, int);
template int gd_launch_synth_tang_tile<double>(hipStream_t, const SynthArgs &, int);

// a shell's tangential synthesis can join a batched launch: the checks of
// its own launch, and the 16-byte table form (VL) that the batched kernel has
int gd_synth_tang_batchable(const SynthArgs &a)
{
    GD_TRY(synth_check(a, "synth batch"));
    GD_CHECK(a.nf == 1 && a.ng == 1 && a.slot0 + 16 <= a.nslot, GD_ERR_ARG, "synth batch: nf=%d ng=%d slots %d/%d",
             a.nf, a.ng, a.slot0, a.nslot);
    GD_CHECK(a.nlat2 % 4 == 0 && (uintptr_t)a.tab0 % 16 == 0 && (uintptr_t)a.tab1 % 16 == 0, GD_ERR_ARG,
             "synth batch: nlat/2=%d needs 16-byte table rows", a.nlat2);
    return GD_OK;
}

// node tile of the table-resident analysis for nlat/2 (0: not in one chunk)
int gd_ana_kn(int nlat2, bool fp64)
{
    if (nlat2 <= 32) return 32;
    if (nlat2 <= 48) return 48;
    if (nlat2 <= 96) return 96;
    if (!fp64 && nlat2 <= 128) return 128;
    return 0;
}

int gd_ana_blocks(const AnaArgs &a) { return ((a.maxgroup + TM - 1) / TM) * 2 * (a.mmax + 1); }

template <typename T, bool VEC, int KN, int MM>
static void ana_batch_launch(hipStream_t st, const BatchRef<AnaArgs> &b, int nblocks)
{
    constexpr size_t lds = ana_res_lds<T, VEC, 1, KN>();
    hipLaunchKernelGGL((ana_res_batch_kernel<T, VEC, 1, KN, MM>), dim3(nblocks), dim3(64 * ana_waves<T>()), lds, st, b);
}

template <typename T, bool VEC, int KN> static void ana_batch_mm(hipStream_t st, const BatchRef<AnaArgs> &b, int nblocks, int mm)
{
    if constexpr (sizeof(T) == 4) {
        if (mm) {
            ana_batch_launch<T, VEC, KN, 1>(st, b, nblocks);
            return;
        }
    }
    ana_batch_launch<T, VEC, KN, 0>(st, b, nblocks);
}

template <typename T>
int gd_launch_ana_batch(hipStream_t st, const BatchRef<AnaArgs> &b, int nblocks, bool vec, int kn, int mm)
{
    GD_CHECK(!mm || sizeof(T) == 4, GD_ERR_ARG, "ana batch: the BF16x3 contraction takes FP32 operands");
    if (nblocks <= 0) return GD_OK;
    switch (kn) {
    case 32: vec ? ana_batch_mm<T, true, 32>(st, b, nblocks, mm) : ana_batch_mm<T, false, 32>(st, b, nblocks, mm); break;
    case 48: vec ? ana_batch_mm<T, true, 48>(st, b, nblocks, mm) : ana_batch_mm<T, false, 48>(st, b, nblocks, mm); break;
    case 96: vec ? ana_batch_mm<T, true, 96>(st, b, nblocks, mm) : ana_batch_mm<T, false, 96>(st, b, nblocks, mm); break;
    case 128:
        if constexpr (sizeof(T) == 4) {
            vec ? ana_batch_mm<T, true, 128>(st, b, nblocks, mm) : ana_batch_mm<T, false, 128>(st, b, nblocks, mm);
            break;
        }
        [[fallthrough]];
    default: GD_CHECK(false, GD_ERR_ARG, "ana batch: node tile %d", kn);
    }
    return gd_check_launch("ana_res_batch_kernel");
}
template int gd_launch_ana_batch<float>(hipStream_t, const BatchRef<AnaArgs> &, int, bool, int, int);
template int gd_launch_ana_batch<double>(hipStream_t, const BatchRef<AnaArgs> &, int, bool, int, int);

// ====================================================================== FFT ring kernel
namespace {
template <typename T> struct C2 {
    T x, y;
};
template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by i*dir
template <typename T> __device__ __forceinline__ C2<T> cmuli(C2<T> a, int dir)
{
    return dir > 0 ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

template <typename T> __device__ __forceinline__ void dft2(C2<T> &a, C2<T> &b)
{
    C2<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
template <typename T> __device__ __forceinline__ void dft3(C2<T> &v0, C2<T> &v1, C2<T> &v2, int dir)
{
    const T s3 = T(0.86602540378443864676);
    C2<T> t1 = cadd(v1, v2);
    C2<T> t2 = {v0.x - T(0.5) * t1.x, v0.y - T(0.5) * t1.y};
    C2<T> d = csub(v1, v2);
    C2<T> t3 = cmuli(C2<T>{s3 * d.x, s3 * d.y}, dir);
    v0 = cadd(v0, t1);
    v1 = cadd(t2, t3);
    v2 = csub(t2, t3);
}
template <typename T> __device__ __forceinline__ void dft4(C2<T> &v0, C2<T> &v1, C2<T> &v2, C2<T> &v3, int dir)
{
    C2<T> a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = cmuli(csub(v1, v3), dir);
    v0 = cadd(a, c);
    v2 = csub(a, c);
    v1 = cadd(b, d);
    v3 = csub(b, d);
}

template <int R, typename T> __device__ __forceinline__ void dftR(C2<T> *v, int dir)
{
    if constexpr (R == 2) {
        dft2(v[0], v[1]);
    } else if constexpr (R == 3) {
        dft3(v[0], v[1], v[2], dir);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], dir);
    } else if constexpr (R == 6) {
        // 6 = 2 x 3: E = DFT3(v0,v2,v4), O = DFT3(v1,v3,v5), X_k = E_k + w6^k O_k
        C2<T> e0 = v[0], e1 = v[2], e2 = v[4], o0 = v[1], o1 = v[3], o2 = v[5];
        dft3(e0, e1, e2, dir);
        dft3(o0, o1, o2, dir);
        const T h = T(0.5), s3 = T(0.86602540378443864676);
        C2<T> w1 = {h, dir > 0 ? s3 : -s3}, w2 = {-h, dir > 0 ? s3 : -s3};
        o1 = cmul(o1, w1);
        o2 = cmul(o2, w2);
        v[0] = cadd(e0, o0);
        v[3] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[4] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[5] = csub(e2, o2);
    } else if constexpr (R == 8) {
        C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4(e0, e1, e2, e3, dir);
        dft4(o0, o1, o2, o3, dir);
        const T r2 = T(0.70710678118654752440);
        const T sg = dir > 0 ? T(1) : T(-1);
        o1 = cmul(o1, C2<T>{r2, sg * r2});
        o2 = cmuli(o2, dir);
        o3 = cmul(o3, C2<T>{-r2, sg * r2});
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3);
        v[7] = csub(e3, o3);
    }
}

constexpr int MAXC = 32; // complex values a thread holds across one stage's barrier

// LDS index skew: one pad slot per 16 complex values, so the stride-R writes
// of the early Stockham stages (and stride-Ns groups later) spread over banks.
__device__ __forceinline__ int sk(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int ring_ld(int n) { return n + (n >> 4) + 1; } // per-transform LDS stride

// One Stockham stage over nfft transforms of length n (stride ld in buf); in
// place, with every butterfly read before any is written.  Runtime radix plan.
template <int R, typename T>
__device__ __noinline__ void fft_stage(C2<T> *buf, const C2<T> *tw, int nfft, int n, int ld, int Ns, int dir)
{
    constexpr int IT = MAXC / R;
    const int nb = n / R, total = nfft * nb;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int twstep = n / (Ns * R);
    C2<T> v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int g = tid + it * nthr;
        if (g < total) {
            const int f = g / nb, j = g - f * nb;
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = buf[f * ld + sk(j + r * nb)];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2<T> w = tw[k * r * twstep];
                if (dir > 0) w.y = -w.y;
                v[it][r] = cmul(v[it][r], w);
            }
            dftR<R, T>(v[it], dir);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int g = tid + it * nthr;
        if (g < total) {
            const int f = g / nb, j = g - f * nb;
            const int k = j % Ns;
            const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[f * ld + sk(idxD + r * Ns)] = v[it][r];
        }
    }
    __syncthreads();
}

template <typename T>
__device__ void fft_batch(C2<T> *buf, const C2<T> *tw, int nfft, int n, int ld, const int *radix, int nstages, int dir)
{
    int Ns = 1;
    for (int s = 0; s < nstages; ++s) {
        const int R = radix[s];
        switch (R) {
        case 8: fft_stage<8, T>(buf, tw, nfft, n, ld, Ns, dir); break;
        case 6: fft_stage<6, T>(buf, tw, nfft, n, ld, Ns, dir); break;
        case 4: fft_stage<4, T>(buf, tw, nfft, n, ld, Ns, dir); break;
        case 3: fft_stage<3, T>(buf, tw, nfft, n, ld, Ns, dir); break;
        default: fft_stage<2, T>(buf, tw, nfft, n, ld, Ns, dir); break;
        }
        Ns *= R;
    }
}

// ---- compile-time plans: every divisor and twiddle stride is a constant
constexpr int ct_next_radix(int n)
{
    return n % 8 == 0 ? 8 : n % 4 == 0 ? 4 : n % 6 == 0 ? 6 : n % 3 == 0 ? 3 : n % 2 == 0 ? 2 : 0;
}
constexpr int ct_nstages(int n)
{
    int s = 0;
    while (n > 1) {
        n /= ct_next_radix(n);
        ++s;
    }
    return s;
}
constexpr int ct_radix(int n, int stage)
{
    for (int s = 0; s < stage; ++s) n /= ct_next_radix(n);
    return ct_next_radix(n);
}

constexpr int RING_THREADS = 256; // ring_kernel block size

// The Stockham write of a butterfly goes to idxD + r NS with idxD = q NS R + k,
// k < NS.  sk(idxD + r NS) - sk(idxD) = r NS + [(idxD + r NS) / 16 - idxD / 16]
// is a constant of r when the bracket is the same for every residue idxD mod
// 16 the butterflies can have: then the R addresses are one base and
// constant offsets.
constexpr bool sk_write_const(int NS, int R)
{
    for (int r = 0; r < R; ++r) {
        int val = -1;
        for (int q = 0; q < 16; ++q)
            for (int k = 0; k < NS && k < 16; ++k) {
                const int res = (q * NS * R + k) % 16;
                const int d = (res + r * NS) / 16;
                if (val < 0) val = d;
                else if (d != val) return false;
            }
    }
    return true;
}
// the offset of write r: the bracket at residue 0 (q = k = 0), which every
// residue shares when sk_write_const holds
constexpr int sk_write_off(int NS, int r) { return r * NS + (r * NS) / 16; }

// NFFT > 0: the transform count is a compile-time constant (nfft is ignored)
template <int R, int NS, int N, int DIR, typename T, int NFFT>
__device__ __forceinline__ void fft_stage_ct(C2<T> *buf, const C2<T> *tw, int nfft)
{
    constexpr int nb = N / R, twstep = N / (NS * R), LD = ring_ld(N);
    constexpr int nthr = RING_THREADS;
    constexpr int IT = NFFT > 0 ? (NFFT * nb + nthr - 1) / nthr : MAXC / R;
    const int total = (NFFT > 0 ? NFFT : nfft) * nb;
    const int tid = threadIdx.x;
    // sk(e + r * s) = sk(e) + r * (s + s / 16) when s is a multiple of 16: the
    // R addresses of a butterfly are then one base and constant offsets
    constexpr bool RD16 = nb % 16 == 0, WR16 = sk_write_const(NS, R);
    C2<T> v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int g = tid + it * nthr;
        if (g < total) {
            const int f = g / nb, j = g - f * nb;
            const int k = j % NS;
            const int rb = f * LD + sk(j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = buf[RD16 ? rb + r * (nb + nb / 16) : f * LD + sk(j + r * nb)];
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C2<T> w = tw[k * r * twstep];
                    if (DIR > 0) w.y = -w.y;
                    v[it][r] = cmul(v[it][r], w);
                }
            }
            dftR<R, T>(v[it], DIR);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int g = tid + it * nthr;
        if (g < total) {
            const int f = g / nb, j = g - f * nb;
            const int k = j % NS;
            const int idxD = (j / NS) * NS * R + k;
            const int wb = f * LD + sk(idxD);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[WR16 ? wb + sk_write_off(NS, r) : f * LD + sk(idxD + r * NS)] = v[it][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int DIR, int NFFT = 0, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_ct(C2<T> *buf, const C2<T> *tw, int nfft)
{
    if constexpr (S < ct_nstages(N)) {
        constexpr int R = ct_radix(N, S);
        fft_stage_ct<R, NS, N, DIR, T, NFFT>(buf, tw, nfft);
        fft_ct<T, N, DIR, NFFT, S + 1, NS * R>(buf, tw, nfft);
    }
}

// One block per (northern node i, level r).
// NC > 0: the ring length is the compile-time constant NC (nlon); NC = 0 uses
// the runtime plan.
// (i, r): the block's northern node and level (the launch's own grid, or its
// share of a batched launch)
template <typename T, int MODE, int NC, class AT> __device__ __forceinline__ void ring_body(const AT &a, int i, int r)
{
    extern __shared__ char smem_raw[];
    const int n = NC > 0 ? NC : a.nlon, nm = a.mmax + 1;
    const int ld = ring_ld(n);
    C2<T> *tw = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *buf = tw + n;
    __shared__ int srad[12];
    __shared__ int sfn[16], sfs[16];
    // The boundary layouts (RING_SYNTH writes, RING_ANALYSIS reads) are
    // colatitude-fastest: a 128-byte line holds 16 consecutive rings of one
    // (longitude, level).  Consecutive blocks go to different XCDs (block b to
    // XCD b mod 8, each with its own L2), so those modes renumber the blocks
    // for each XCD to take a contiguous range of (level, ring) and a line's
    // pieces meet in one L2.  The step modes read and write ring-contiguous
    // records and keep the plain order.
    const int tid = threadIdx.x;
    constexpr int nthr = RING_THREADS;
    const int is = a.nlat - 1 - i; // southern mirror
    const C2<T> *twg = static_cast<const C2<T> *>(a.tw);
    for (int k = tid; k < n; k += nthr) tw[k] = twg[k];
    if (tid < 12) srad[tid] = a.radix[tid];
    if (tid < 16) {
        sfn[tid] = a.fn_off[tid];
        sfs[tid] = a.fs_off[tid];
    }
    __syncthreads();
    // fields in / products out: compile-time for the step modes
    constexpr int NFC = MODE == RING_STEP ? 15 : MODE == RING_STEP_TANG ? 4 : MODE == RING_STEP_TANG_C ? 6
                      : MODE == RING_ADJ_TANG ? 7 : 0;
    constexpr int NPC = MODE == RING_STEP ? 5 : MODE == RING_STEP_TANG ? 3 : MODE == RING_STEP_TANG_C ? 4
                      : MODE == RING_ADJ_TANG ? 4 : 0;
    const int NF = NFC ? NFC : a.nfield;

    if (MODE != RING_ANALYSIS) {
        // Hermitian spectra of the north (X) and south (Y) rings packed as Z = X + iY
        const T *src = static_cast<const T *>(a.fb) + ((size_t)i * a.nr + r) * nm * a.nslot_f;
        // fixed trip count, fully unrolled: every thread's global loads are in
        // flight together before the LDS stores consume them.  With a
        // compile-time ring length the first MAXU iterations (enough for the
        // nm <= n/3 orders of the dealiased grids) load unconditionally,
        // clamped to the last element: a guarded load becomes a branch whose
        // join waits for it.  Larger nm take the guarded iterations after.
        constexpr int MAXI = 8;
        constexpr int MAXU0 = (NC > 0 && NFC > 0) ? (NFC * (NC / 3) + nthr - 1) / nthr : 0;
        constexpr int MAXU = MAXU0 < MAXI ? MAXU0 : MAXI;
        const int ne = NF * nm;
        T xr[MAXI], xi[MAXI], yr[MAXI], yi[MAXI];
#pragma unroll
        for (int it = 0; it < MAXI; ++it) {
            const int e0 = tid + it * nthr;
            if (it < MAXU || e0 < ne) {
                const int e = e0 < ne ? e0 : ne - 1;
                const int F = e % NF, m = e / NF;
                const T *s = src + (size_t)m * a.nslot_f;
                xr[it] = s[sfn[F]];
                xi[it] = s[sfn[F] + 1];
                yr[it] = s[sfs[F]];
                yi[it] = s[sfs[F] + 1];
            }
        }
#pragma unroll
        for (int it = 0; it < MAXI; ++it) {
            const int e = tid + it * nthr;
            if (e < NF * nm) {
                const int F = e % NF, m = e / NF;
                if (m == 0) {
                    buf[F * ld] = C2<T>{xr[it], yr[it]};
                } else {
                    buf[F * ld + sk(m)] = C2<T>{xr[it] - yi[it], xi[it] + yr[it]};
                    buf[F * ld + sk(n - m)] = C2<T>{xr[it] + yi[it], yr[it] - xi[it]};
                }
            }
        }
        // zero the modes nm .. n-nm of every field
        for (int k = nm + tid; k <= n - nm; k += nthr)
            for (int F = 0; F < NF; ++F) buf[F * ld + sk(k)] = C2<T>{T(0), T(0)};
        __syncthreads();
        if constexpr (NC > 0) fft_ct<T, NC, +1, NFC>(buf, tw, NF);
        else fft_batch<T>(buf, tw, NF, n, ld, srad, a.nstages, +1);
    }

    if (MODE == RING_SYNTH) {
        for (int j = tid; j < n; j += nthr)
            for (int F = 0; F < NF; ++F) {
                double *o = F == 0 ? a.phys_out[0] : a.phys_out[1];
                const C2<T> z = buf[F * ld + sk(j)];
                o[((size_t)r * n + j) * a.nlat + i] = (double)z.x;
                o[((size_t)r * n + j) * a.nlat + is] = (double)z.y;
            }
        return;
    }

    int NP;
    if constexpr (MODE == RING_STEP_TANG || MODE == RING_STEP_TANG_C) {
        // the products of RING_STEP with u_r = omega_r = B_r = j_r = 0: the
        // tangential parts of u x omega, j x B and u x B vanish and u_r dT/dr
        // drops out; the Coriolis term and the tangential advection of T stay,
        // evaluated in the same order as there
        // fields: 0 u_t 1 u_p 2 gT_t 3 gT_p (4 gC_t 5 gC_p: -u.grad C as product 3)
        NP = NPC;
        const T ro = (T)a.ro;
        const T *costh = static_cast<const T *>(a.costh);
        for (int j = tid; j < n; j += nthr) {
            C2<T> f[NFC];
#pragma unroll
            for (int F = 0; F < NFC; ++F) f[F] = buf[F * ld + sk(j)];
            C2<T> o[NPC];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ti = a.coriolis_index ? (h ? is : i) : (j < a.nlat - 1 ? j : a.nlat - 1);
                const T cth = costh[ti];
#define FV(F) (h ? f[F].y : f[F].x)
                const T ut = FV(0), up = FV(1), gt = FV(2), gp = FV(3);
#undef FV
                const T at = -ro * (-cth * up);
                const T ap = -ro * (cth * ut);
                const T aT = -(ut * gt + up * gp);
                if (h) {
                    o[0].y = at;
                    o[1].y = ap;
                    o[2].y = aT;
                } else {
                    o[0].x = at;
                    o[1].x = ap;
                    o[2].x = aT;
                }
                if constexpr (NFC == 6) {
                    const T ct = h ? f[4].y : f[4].x, cp = h ? f[5].y : f[5].x;
                    const T aC = -(ut * ct + up * cp);
                    if (h) o[3].y = aC;
                    else o[3].x = aC;
                }
            }
#pragma unroll
            for (int P = 0; P < NPC; ++P) buf[P * ld + sk(j)] = o[P];
        }
        __syncthreads();
    } else if constexpr (MODE == RING_ADJ_TANG) {
        // transposed Jacobian of the products above at (u, grad T): with
        // a_t = ro cth u_p, a_p = -ro cth u_t, a_T = -(u_t g_t + u_p g_p),
        //   u_t* = -ro cth a_p* - g_t a_T*,  u_p* = ro cth a_t* - g_p a_T*,
        //   g_t* = -u_t a_T*,                g_p* = -u_p a_T*
        // fields: 0 u_t 1 u_p 2 a_t* 3 a_p* 4 g_t 5 g_p 6 a_T*
        NP = NPC;
        const T ro = (T)a.ro;
        const T *costh = static_cast<const T *>(a.costh);
        for (int j = tid; j < n; j += nthr) {
            C2<T> f[NFC];
#pragma unroll
            for (int F = 0; F < NFC; ++F) f[F] = buf[F * ld + sk(j)];
            C2<T> o[NPC];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ti = a.coriolis_index ? (h ? is : i) : (j < a.nlat - 1 ? j : a.nlat - 1);
                const T rc = ro * costh[ti];
#define FV(F) (h ? f[F].y : f[F].x)
                const T ut = FV(0), up = FV(1), bt = FV(2), bp = FV(3), gt = FV(4), gp = FV(5), bT = FV(6);
#undef FV
                const T vals[NPC] = {-rc * bp - gt * bT, rc * bt - gp * bT, -ut * bT, -up * bT};
#pragma unroll
                for (int P = 0; P < NPC; ++P) {
                    if (h) o[P].y = vals[P];
                    else o[P].x = vals[P];
                }
            }
#pragma unroll
            for (int P = 0; P < NPC; ++P) buf[P * ld + sk(j)] = o[P];
        }
        __syncthreads();
    } else if (MODE == RING_ANALYSIS) {
        NP = NF;
        for (int j = tid; j < n; j += nthr)
            for (int F = 0; F < NF; ++F) {
                const double *p = F == 0 ? a.phys_in[0] : a.phys_in[1];
                buf[F * ld + sk(j)] = C2<T>{(T)p[((size_t)r * n + j) * a.nlat + i], (T)p[((size_t)r * n + j) * a.nlat + is]};
            }
        __syncthreads();
    } else {
        // pointwise nonlinear terms on the north (.x) and south (.y) rings
        // fields: 0 u_t 1 u_p 2 w_t 3 w_p 4 B_t 5 B_p 6 j_t 7 j_p 8 gT_t 9 gT_p
        //         10 u_r 11 w_r 12 B_r 13 j_r 14 dT_r
        NP = 5;
        const T ro = (T)a.ro;
        const T *sinth = static_cast<const T *>(a.sinth), *costh = static_cast<const T *>(a.costh);
        for (int j = tid; j < n; j += nthr) {
            C2<T> f[15];
#pragma unroll
            for (int F = 0; F < 15; ++F) f[F] = buf[F * ld + sk(j)];
            C2<T> o[5];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                int ti;
                T sth, cth;
                if (a.coriolis_index) {
                    ti = h ? is : i;
                } else {
                    ti = j < a.nlat - 1 ? j : a.nlat - 1; // velocity.jl:494 theta_{min(j,nlat)}
                }
                sth = sinth[ti];
                cth = costh[ti];
#define FV(F) (h ? f[F].y : f[F].x)
                const T ut = FV(0), up = FV(1), wt = FV(2), wp = FV(3), Bt = FV(4), Bp = FV(5);
                const T jt = FV(6), jp = FV(7), gt = FV(8), gp = FV(9);
                const T ur = FV(10), wr = FV(11), Br = FV(12), jr = FV(13), gr = FV(14);
#undef FV
                // velocity.jl:512-527, 664-670; magnetic.jl:399-401; scalar_field_common.jl:401-407
                T at = -ro * (up * wr - ur * wp) + (-ro * (-cth * up));
                T ap = -ro * (ur * wt - ut * wr) + (-ro * (cth * ut + sth * ur));
                at += ro * (jp * Br - jr * Bp);
                ap += ro * (jr * Bt - jt * Br);
                const T bt = up * Br - ur * Bp;
                const T bp = ur * Bt - ut * Br;
                const T aT = -(ur * gr + ut * gt + up * gp);
                const T vals[5] = {at, ap, bt, bp, aT};
#pragma unroll
                for (int P = 0; P < 5; ++P) {
                    if (h) o[P].y = vals[P];
                    else o[P].x = vals[P];
                }
            }
#pragma unroll
            for (int P = 0; P < 5; ++P) buf[P * ld + sk(j)] = o[P];
        }
        __syncthreads();
    }

    if constexpr (NC > 0) fft_ct<T, NC, -1, NPC>(buf, tw, NP);
    else fft_batch<T>(buf, tw, NP, n, ld, srad, a.nstages, -1);

    // split Z into north X and south Y, form G+- = (X +- Y) * 2 pi w_i / n
    const T c = static_cast<const T *>(a.wq)[i];
    T *gb = static_cast<T *>(a.gb);
    const int NPn = NPC ? NPC : NP;
    // products np1.. go to their own records in gb2 (the scalar analysis then
    // reads only its own slots); np1 = 0: every product in gb
    const int np1 = a.gb2 ? a.np1 : NPn;
    for (int e = tid; e < NPn * nm; e += nthr) {
        const int P = e % NPn, m = e / NPn;
        const C2<T> Z = buf[P * ld + sk(m)];
        C2<T> X, Y;
        if (m == 0) {
            X = C2<T>{Z.x, T(0)};
            Y = C2<T>{Z.y, T(0)};
        } else {
            const C2<T> Zc = buf[P * ld + sk(n - m)];
            X = C2<T>{T(0.5) * (Z.x + Zc.x), T(0.5) * (Z.y - Zc.y)};
            Y = C2<T>{T(0.5) * (Z.y + Zc.y), T(0.5) * (Zc.x - Z.x)};
        }
        const size_t rec = ((size_t)m * a.nlat2 + i) * a.nr + r;
        T *d = P < np1 ? gb + rec * a.nslot_g + 4 * P : static_cast<T *>(a.gb2) + rec * a.nslot_g2 + 4 * (P - np1);
        d[0] = (X.x + Y.x) * c;
        d[1] = (X.y + Y.y) * c;
        d[2] = (X.x - Y.x) * c;
        d[3] = (X.y - Y.y) * c;
    }
}

template <typename T, int MODE, int NC> __global__ void __launch_bounds__(256) ring_kernel(RingArgs a)
{
    int i = blockIdx.x, r = blockIdx.y;
    if (MODE == RING_SYNTH || MODE == RING_ANALYSIS) {
        const int L = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
        i = L % gridDim.x;
        r = L / gridDim.x;
    }
    ring_body<T, MODE, NC>(a, i, r);
}

// The boundary passes (RING_SYNTH: Fourier -> physical rings, RING_ANALYSIS:
// physical rings -> Fourier) with RG consecutive ring pairs per block.  The
// physical arrays are colatitude-fastest, so one ring's points lie a whole
// row of nlat doubles apart: with one ring pair per block every lane's load or
// store is its own cache line.  Here lanes take (longitude, ring) with the
// ring fastest, so a wave's access covers runs of RG consecutive colatitudes
// (and their southern mirrors); the RG * NF transforms of the block share the
// twiddles and one batched FFT.  Each transform gets the operations of the
// one-ring form.
template <typename T, int MODE, int NC, int RG>
__device__ __forceinline__ void ring_io_body(const RingArgs &a, int g, int r)
{
    static_assert(MODE == RING_SYNTH || MODE == RING_ANALYSIS, "boundary passes only");
    extern __shared__ char smem_raw[];
    const int n = NC > 0 ? NC : a.nlon, nm = a.mmax + 1;
    const int ld = ring_ld(n);
    C2<T> *tw = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *buf = tw + n;
    __shared__ int srad[12];
    __shared__ int sfn[16], sfs[16];
    const int tid = threadIdx.x;
    constexpr int nthr = RING_THREADS;
    const int NF = a.nfield;
    const int i0 = g * RG, nq = min(RG, a.nlat2 - i0); // this block's ring pairs i0 .. i0 + nq - 1
    const int NT = nq * NF;                             // transform t = q * NF + F
    const C2<T> *twg = static_cast<const C2<T> *>(a.tw);
    for (int k = tid; k < n; k += nthr) tw[k] = twg[k];
    if (tid < 12) srad[tid] = a.radix[tid];
    if (tid < 16) {
        sfn[tid] = a.fn_off[tid];
        sfs[tid] = a.fs_off[tid];
    }
    __syncthreads();
    if (MODE == RING_SYNTH) {
        // Hermitian spectra of the north (X) and south (Y) rings packed as Z = X + iY
        for (int e0 = 0; e0 < NT * nm; e0 += 8 * nthr) {
            T xr[8], xi[8], yr[8], yi[8];
            // loads of 8 elements in flight before their stores; unconditional
            // (clamped to the last element): a guarded load becomes a branch
            // whose join waits for it
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int e = min(e0 + tid + it * nthr, NT * nm - 1);
                const int F = e % NF, m = (e / NF) % nm, q = e / (NF * nm);
                const T *s = static_cast<const T *>(a.fb) + (((size_t)(i0 + q) * a.nr + r) * nm + m) * a.nslot_f;
                xr[it] = s[sfn[F]];
                xi[it] = s[sfn[F] + 1];
                yr[it] = s[sfs[F]];
                yi[it] = s[sfs[F] + 1];
            }
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int e = e0 + tid + it * nthr;
                if (e < NT * nm) {
                    const int F = e % NF, m = (e / NF) % nm, q = e / (NF * nm);
                    C2<T> *b = buf + (q * NF + F) * ld;
                    if (m == 0) {
                        b[0] = C2<T>{xr[it], yr[it]};
                    } else {
                        b[sk(m)] = C2<T>{xr[it] - yi[it], xi[it] + yr[it]};
                        b[sk(n - m)] = C2<T>{xr[it] + yi[it], yr[it] - xi[it]};
                    }
                }
            }
        }
        for (int k = nm + tid; k <= n - nm; k += nthr)
            for (int t = 0; t < NT; ++t) buf[t * ld + sk(k)] = C2<T>{T(0), T(0)};
        __syncthreads();
        if constexpr (NC > 0) fft_ct<T, NC, +1, 0>(buf, tw, NT);
        else fft_batch<T>(buf, tw, NT, n, ld, srad, a.nstages, +1);
        for (int e = tid; e < n * nq; e += nthr) {
            const int q = e % nq, j = e / nq, i = i0 + q, is = a.nlat - 1 - i;
            for (int F = 0; F < NF; ++F) {
                double *o = a.phys_out[F == 0 ? 0 : 1];
                const C2<T> z = buf[(q * NF + F) * ld + sk(j)];
                o[((size_t)r * n + j) * a.nlat + i] = (double)z.x;
                o[((size_t)r * n + j) * a.nlat + is] = (double)z.y;
            }
        }
        return;
    }
    // the rings' points: UR elements per thread and round, all loads issued
    // (clamped to the last element) before any LDS store and conversion, so
    // a round waits once, not once per element
    constexpr int UR = 4;
    for (int e0 = tid; e0 < n * nq; e0 += UR * nthr) {
        double px[UR][2], py[UR][2];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            const int e = min(e0 + u * nthr, n * nq - 1);
            const int q = e % nq, j = e / nq, i = i0 + q, is = a.nlat - 1 - i;
            const size_t o = ((size_t)r * n + j) * a.nlat;
            px[u][0] = a.phys_in[0][o + i];
            py[u][0] = a.phys_in[0][o + is];
            if (NF > 1) {
                px[u][1] = a.phys_in[1][o + i];
                py[u][1] = a.phys_in[1][o + is];
            }
        }
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            const int e = e0 + u * nthr;
            if (e < n * nq) {
                const int q = e % nq, j = e / nq;
                buf[(q * NF) * ld + sk(j)] = C2<T>{(T)px[u][0], (T)py[u][0]};
                if (NF > 1) buf[(q * NF + 1) * ld + sk(j)] = C2<T>{(T)px[u][1], (T)py[u][1]};
            }
        }
    }
    __syncthreads();
    if constexpr (NC > 0) fft_ct<T, NC, -1, 0>(buf, tw, NT);
    else fft_batch<T>(buf, tw, NT, n, ld, srad, a.nstages, -1);
    // split Z into north X and south Y, form G+- = (X +- Y) * 2 pi w_i / n
    T *gb = static_cast<T *>(a.gb);
    for (int e = tid; e < NT * nm; e += nthr) {
        const int F = e % NF, m = (e / NF) % nm, q = e / (NF * nm), i = i0 + q;
        const T c = static_cast<const T *>(a.wq)[i];
        const C2<T> *b = buf + (q * NF + F) * ld;
        const C2<T> Z = b[sk(m)];
        C2<T> X, Y;
        if (m == 0) {
            X = C2<T>{Z.x, T(0)};
            Y = C2<T>{Z.y, T(0)};
        } else {
            const C2<T> Zc = b[sk(n - m)];
            X = C2<T>{T(0.5) * (Z.x + Zc.x), T(0.5) * (Z.y - Zc.y)};
            Y = C2<T>{T(0.5) * (Z.y + Zc.y), T(0.5) * (Zc.x - Z.x)};
        }
        const size_t rec = ((size_t)m * a.nlat2 + i) * a.nr + r;
        T *d = gb + rec * a.nslot_g + 4 * F;
        d[0] = (X.x + Y.x) * c;
        d[1] = (X.y + Y.y) * c;
        d[2] = (X.x - Y.x) * c;
        d[3] = (X.y - Y.y) * c;
    }
}

// blocks (ring group, level) in XCD-aware order, as the one-ring boundary form
template <typename T, int MODE, int NC, int RG> __global__ void __launch_bounds__(256) ring_io_kernel(RingArgs a)
{
    const int L = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    ring_io_body<T, MODE, NC, RG>(a, L % gridDim.x, L / gridDim.x);
}

// one grid over the ring passes of several shells (step modes): a shell's
// blocks are its (node, level) pairs, node fastest
template <typename T, int MODE, int NC> __global__ void __launch_bounds__(256) ring_batch_kernel(BatchRef<RingArgs> b)
{
    const int sh = b.shell_of_block[blockIdx.x];
    if (sh == BATCH_PAD) return;
    const int loc = blockIdx.x - b.block_off[sh];
    const auto &a = batch_args(b, sh);
    ring_body<T, MODE, NC>(a, loc % a.nlat2, loc / a.nlat2);
}
} // namespace

template <typename T, int MODE, int NC> static void ring_launch(hipStream_t st, dim3 grid, size_t lds, const RingArgs &a)
{
    hipLaunchKernelGGL((ring_kernel<T, MODE, NC>), grid, dim3(RING_THREADS), lds, st, a);
}

template <typename T, int MODE> static void ring_dispatch(hipStream_t st, dim3 grid, size_t lds, const RingArgs &a)
{
    switch (a.nlon) {
    case 48: ring_launch<T, MODE, 48>(st, grid, lds, a); break;
    case 96: ring_launch<T, MODE, 96>(st, grid, lds, a); break;
    case 144: ring_launch<T, MODE, 144>(st, grid, lds, a); break;
    case 192: ring_launch<T, MODE, 192>(st, grid, lds, a); break;
    case 288: ring_launch<T, MODE, 288>(st, grid, lds, a); break;
    case 384: ring_launch<T, MODE, 384>(st, grid, lds, a); break;
    case 768: ring_launch<T, MODE, 768>(st, grid, lds, a); break;
    default: ring_launch<T, MODE, 0>(st, grid, lds, a); break;
    }
}

static int ring_check(int mode, const RingArgs &a, int *nf_out)
{
    const int nthr = RING_THREADS;
    int nf = (mode == RING_STEP) ? 15 : (mode == RING_STEP_TANG) ? 4 : (mode == RING_STEP_TANG_C) ? 6
           : (mode == RING_ADJ_TANG) ? 7 : a.nfield;
    GD_CHECK(nf >= 1 && nf <= 15, GD_ERR_ARG, "ring: nfield=%d", nf);
    GD_CHECK((long)nf * (a.mmax + 1) <= 8L * nthr, GD_ERR_ARG, "ring: %d fields x %d orders exceed one load batch",
             nf, a.mmax + 1);
    for (int s = 0; s < a.nstages; ++s) {
        int R = a.radix[s];
        long per = ((long)nf * (a.nlon / R) + nthr - 1) / nthr;
        GD_CHECK(per * R <= MAXC, GD_ERR_ARG, "ring: nlon=%d too large for %d fields", a.nlon, nf);
    }
    *nf_out = nf;
    return GD_OK;
}

// a shell's tangential ring pass can join a batched launch (its own launch's checks)
int gd_ring_tang_batchable(const RingArgs &a, bool fp64)
{
    int nf;
    GD_TRY(ring_check(RING_STEP_TANG, a, &nf));
    GD_CHECK((fp64 ? 16 : 8) * ((size_t)ring_ld(a.nlon) * nf + a.nlon) <= 160 * 1024, GD_ERR_ARG,
             "ring batch: nlon=%d needs too much LDS", a.nlon);
    return GD_OK;
}

template <typename T, int MODE, int RG> static void ring_io_nc(hipStream_t st, dim3 grid, size_t lds, const RingArgs &a)
{
    switch (a.nlon) {
    case 384: hipLaunchKernelGGL((ring_io_kernel<T, MODE, 384, RG>), grid, dim3(RING_THREADS), lds, st, a); break;
    case 768: hipLaunchKernelGGL((ring_io_kernel<T, MODE, 768, RG>), grid, dim3(RING_THREADS), lds, st, a); break;
    default: hipLaunchKernelGGL((ring_io_kernel<T, MODE, 0, RG>), grid, dim3(RING_THREADS), lds, st, a); break;
    }
}

// ring pairs per block of the boundary passes: 2 for FP64 rings, 4 for FP32
// (measured at lmax=255, 384 x 768, 97 levels: FP64 round trip 5.78 / 5.40 /
// 5.58 / 6.10 ms and FP32 4.08 / 3.61 / 3.54 / 3.73 ms for 1 / 2 / 4 / 8;
// GD_RING_RG = 1, 2, 4 or 8 overrides, a measurement aid), halved until the
// block's transforms fit the FFT's per-thread registers (RG nfield nlon <=
// MAXC threads) and 160 KB of LDS
template <typename T> static int ring_io_rg(const RingArgs &a, int nf)
{
    static const int env = [] {
        const char *e = getenv("GD_RING_RG");
        const int v = e ? atoi(e) : 0;
        return (v == 1 || v == 2 || v == 4 || v == 8) ? v : 0;
    }();
    int rg = env ? env : (sizeof(T) == 8 ? 2 : 4);
    while (rg > 1 && ((long)rg * nf * a.nlon > (long)MAXC * RING_THREADS ||
                      sizeof(C2<T>) * ((size_t)ring_ld(a.nlon) * nf * rg + a.nlon) > 160 * 1024))
        rg /= 2;
    return rg;
}

template <typename T> int gd_launch_ring(hipStream_t st, int mode, const RingArgs &a)
{
    int nf;
    GD_TRY(ring_check(mode, a, &nf));
    if (mode == RING_SYNTH || mode == RING_ANALYSIS) {
        // the boundary passes take one scalar or one (theta, phi) pair: phys_in / phys_out [0], [1]
        GD_CHECK(nf <= 2, GD_ERR_ARG, "ring: %d fields in a boundary pass", nf);
        const int rg = ring_io_rg<T>(a, nf);
        if (rg > 1) {
            const size_t lds = sizeof(C2<T>) * ((size_t)ring_ld(a.nlon) * nf * rg + a.nlon);
            const dim3 grid((a.nlat2 + rg - 1) / rg, a.nr);
#define GD_RIO(MODE)                                                                 \
    {                                                                                \
        if (rg == 8) ring_io_nc<T, MODE, 8>(st, grid, lds, a);                       \
        else if (rg == 4) ring_io_nc<T, MODE, 4>(st, grid, lds, a);                  \
        else ring_io_nc<T, MODE, 2>(st, grid, lds, a);                               \
    }
            if (mode == RING_SYNTH) GD_RIO(RING_SYNTH) else GD_RIO(RING_ANALYSIS)
#undef GD_RIO
            return gd_check_launch("ring_io_kernel");
        }
    }
    size_t lds = sizeof(C2<T>) * ((size_t)ring_ld(a.nlon) * nf + a.nlon);
    GD_CHECK(lds <= 160 * 1024, GD_ERR_ARG, "ring: LDS %zu too large", lds);
    dim3 grid(a.nlat2, a.nr);
    switch (mode) {
    case RING_SYNTH: ring_dispatch<T, RING_SYNTH>(st, grid, lds, a); break;
    case RING_ANALYSIS: ring_dispatch<T, RING_ANALYSIS>(st, grid, lds, a); break;
    case RING_STEP_TANG: ring_dispatch<T, RING_STEP_TANG>(st, grid, lds, a); break;
    case RING_STEP_TANG_C: ring_dispatch<T, RING_STEP_TANG_C>(st, grid, lds, a); break;
    case RING_ADJ_TANG: ring_dispatch<T, RING_ADJ_TANG>(st, grid, lds, a); break;
    default: ring_dispatch<T, RING_STEP>(st, grid, lds, a); break;
    }
    return gd_check_launch("ring_kernel");
}

// The tangential step's ring pass in one grid over several shells of one
// ring length (the compile-time FFT plan of nlon where there is one)
template <typename T, int NC>
static void ring_batch_launch(hipStream_t st, const BatchRef<RingArgs> &b, int nblocks, size_t lds)
{
    hipLaunchKernelGGL((ring_batch_kernel<T, RING_STEP_TANG, NC>), dim3(nblocks), dim3(RING_THREADS), lds, st, b);
}

template <typename T> int gd_launch_ring_tang_batch(hipStream_t st, const BatchRef<RingArgs> &b, int nblocks, int nlon)
{
    if (nblocks <= 0) return GD_OK;
    const size_t lds = sizeof(C2<T>) * ((size_t)ring_ld(nlon) * 4 + nlon);
    GD_CHECK(lds <= 160 * 1024, GD_ERR_ARG, "ring batch: LDS %zu too large", lds);
    switch (nlon) {
    case 48: ring_batch_launch<T, 48>(st, b, nblocks, lds); break;
    case 96: ring_batch_launch<T, 96>(st, b, nblocks, lds); break;
    case 144: ring_batch_launch<T, 144>(st, b, nblocks, lds); break;
    case 192: ring_batch_launch<T, 192>(st, b, nblocks, lds); break;
    case 288: ring_batch_launch<T, 288>(st, b, nblocks, lds); break;
    case 384: ring_batch_launch<T, 384>(st, b, nblocks, lds); break;
    case 768: ring_batch_launch<T, 768>(st, b, nblocks, lds); break;
    default: ring_batch_launch<T, 0>(st, b, nblocks, lds); break;
    }
    return gd_check_launch("ring_batch_kernel");
}
template int gd_launch_ring_tang_batch<float>(hipStream_t, const BatchRef<RingArgs> &, int, int);
template int gd_launch_ring_tang_batch<double>(hipStream_t, const BatchRef<RingArgs> &, int, int);

// ====================================================================== permutations
template <typename T>
__global__ void pack_kernel(const double *__restrict__ re, const double *__restrict__ im,
                            const int *__restrict__ lmaj_of_idx, int nlm, int nr, int cin, int c0, T *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const int lm = lmaj_of_idx[idx];
    out[(size_t)(r * cin + c0) * nlm + idx] = (T)re[(size_t)r * nlm + lm];
    out[(size_t)(r * cin + c0 + 1) * nlm + idx] = (T)im[(size_t)r * nlm + lm];
}

template <typename T>
int gd_launch_pack(hipStream_t st, const double *re, const double *im, const int *lmaj_of_idx, int nlm, int nr,
                   int cin, int c0, T *out)
{
    hipLaunchKernelGGL((pack_kernel<T>), dim3((nlm + 255) / 256, nr), dim3(256), 0, st, re, im, lmaj_of_idx, nlm, nr,
                       cin, c0, out);
    return gd_check_launch("pack_kernel");
}

// dst[r][j] = src[r][map[j]]
__global__ void perm_kernel(const double *__restrict__ src, const int *__restrict__ map, int nlm, int nr,
                            double *__restrict__ dst)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (j >= nlm) return;
    dst[(size_t)r * nlm + j] = src[(size_t)r * nlm + map[j]];
}

int gd_launch_perm(hipStream_t st, const double *src, const int *map, int nlm, int nr, double *dst)
{
    hipLaunchKernelGGL(perm_kernel, dim3((nlm + 255) / 256, nr), dim3(256), 0, st, src, map, nlm, nr, dst);
    return gd_check_launch("perm_kernel");
}

int gd_launch_unpack(hipStream_t st, const double *src, const int *idx_of_lmaj, int nlm, int nr, double *dst)
{
    return gd_launch_perm(st, src, idx_of_lmaj, nlm, nr, dst);
}

// explicit instantiations
template int gd_launch_synth_vec<float>(hipStream_t, const SynthArgs &);
template int gd_launch_synth_vec<double>(hipStream_t, const SynthArgs &);
template int gd_launch_synth_scal<float>(hipStream_t, const SynthArgs &);
template int gd_launch_synth_scal<double>(hipStream_t, const SynthArgs &);
template int gd_launch_ana<float>(hipStream_t, const AnaArgs &, bool, int);
template int gd_launch_ana<double>(hipStream_t, const AnaArgs &, bool, int);
template int gd_launch_ring<float>(hipStream_t, int, const RingArgs &);
template int gd_launch_ring<double>(hipStream_t, int, const RingArgs &);
template int gd_launch_pack<float>(hipStream_t, const double *, const double *, const int *, int, int, int, int,
                                   float *);
template int gd_launch_pack<double>(hipStream_t, const double *, const double *, const int *, int, int, int, int,
                                    double *);
