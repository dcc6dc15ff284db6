// gd_adjoint.hip -- reverse-mode derivative (adjoint) of the tangential CNAB2
// step, checkpointed over many steps, and the surface misfit of a run.
//
// One forward step k (gd_step.hip; timestep.jl:610-730, simulation.jl:1151-1330):
//   nl_k    = N(x_k)                                  (nonlinear terms of the state)
//   rhs_k   = R (E x_k + c1 nl_k - c0 nl_{k-1})       c1 = 3/2, c0 = 1/2; AB1 at k = 0: c1 = 1, c0 = 0
//   x_{k+1} = S_l^-1 rhs_k                            per (l, m, re/im) column
// with E = I/dt + (1 - theta) L_l, R the Dirichlet rows (rhs boundary rows
// set to 0) and the temperature's boundary rows read as 0 in E x.  Given the
// cotangent lam_{k+1} of x_{k+1} and mu_k of nl_k (the contribution of step
// k+1 through its history term), the transposed step is
//   r       = R S_l^-T lam_{k+1}                      (U^T then L^T of the stored LU)
//   lam_k   = E^T r + N'(x_k)^T mu_k,   mu_k += c1 r,  mu_{k-1} = -c0 r.
// N is the tangential step's force and advection: the vector analysis of the
// Coriolis term (Pm/E)(cos u_phi, -cos u_theta) and the scalar analysis of
// -u.grad T, with u = synth(u_pol, u_tor) and grad T = synth_grad(T/r)
// (velocity.jl:512-527, scalar_field_common.jl:401-407).  Its transposed
// Jacobian uses the transposes of the transforms on the Gauss grid:
//   synthesis^T = C A W^-1 (scalar),  C L A_v W^-1 (vector)
//   analysis^T  = W S C^-1 (scalar),  W S_v C^-1 L^-1 (vector)
// with W the quadrature weights 2 pi w_i / nlon of each grid point, C the
// factor c_m (1 at m = 0, 2 otherwise) of the real-field synthesis and L =
// l(l+1).  W is a per-point factor and the products' Jacobian acts pointwise,
// so W^-1 (pointwise)^T W = (pointwise)^T and
//   N'(x)^T mu = C [L] A( P'(x)^T S(C^-1 [L^-1] mu) )
// is one more pass through the forward machinery: a synthesis of the state
// and of the scaled cotangents, a ring pass applying the products' transposed
// Jacobian between the inverse and forward FFTs (RING_ADJ_TANG), the
// weighted analysis, and the C (and 1/r) factors.
// The magnetic fields' nonlinear terms are identically zero in the tangential
// step, so their cotangents pass only through the solves.
#include <cmath>

#include "gd_step.hpp"

namespace {

// ---------------------------------------------------------------- transposed CNAB2 solve
struct AdjChainArgs {
    const double *nuLap, *rinv2, *nus, *LUt, *rdiag;
    const int *l_of_idx;
    int nl, n, nlm;
    double dt, theta;
    int bc_mask;          // bit f: Dirichlet rows in field f's S_l
    const double *lam_in; // cotangent of x_{k+1} (10 planes, m-major)
    double *lam_out;      // E^T r (stored)
    double *mu_cur;       // += c1 r
    double *mu_prev;      // = -c0 r (null at the bootstrap)
    double c1, c0;
    // bit f: field f's nonlinear terms and history are not read by the forward
    // solve (the magnetic fields of the tangential step, whose nonlinear terms
    // are identically zero): no cotangent flows to them
    int nl_zero_mask;
};

// One thread per (device idx, field, re/im) column.  U^T z = lam runs upward
// through the rows (U^T is lower triangular), L^T r = z downward (unit upper);
// z is kept in lam_out and overwritten by E^T r, written BW rows behind the
// downward sweep once r of every row of its band is known.  Both sweeps are
// recurrences in r, so memory parallelism comes from within the thread: the
// loads of row i + D (upward) or i - D (downward) are issued while row i is
// computed, into the slot row i frees; they read clamped valid rows and are
// masked where used.  The level-uniform band of nu Lap (column-major, as in
// the step's tables) and 1/r^2 sit in LDS.
template <int BW, int D> __global__ void __launch_bounds__(64) cnab2_adj_kernel(AdjChainArgs c)
{
    constexpr int W = 2 * BW + 1;
    const int n = c.n, nlm = c.nlm;
    const int idx = blockIdx.x * 64 + threadIdx.x;
    const int f = blockIdx.y, ri = blockIdx.z;
    const int k = f < 2 ? 0 : (f < 4 ? 1 : 2);
    extern __shared__ double smem[];
    double *Lc = smem, *r2 = smem + W * n;
    const double *Lap = c.nuLap + (size_t)k * W * n;
    // the column's degree and the table fill are loaded together (UF entries
    // per thread and round, loads before the LDS stores; past the end a
    // thread stores the last entry again), not one wait per entry
    const int l = c.l_of_idx[idx < nlm ? idx : nlm - 1];
    constexpr int UF = 4;
    for (int t0 = threadIdx.x; t0 < W * n; t0 += UF * 64) {
        double v[UF];
#pragma unroll
        for (int e = 0; e < UF; ++e) v[e] = Lap[min(t0 + e * 64, W * n - 1)];
#pragma unroll
        for (int e = 0; e < UF; ++e) Lc[min(t0 + e * 64, W * n - 1)] = v[e];
    }
    for (int t = threadIdx.x; t < n; t += 64) r2[t] = c.rinv2[t];
    __syncthreads();
    if (idx >= nlm) return;
    const double nulf = c.nus[k] * (double)l * (double)(l + 1);
    const double *LU = c.LUt + (size_t)k * W * n * c.nl + l;
    const double *rdg = c.rdiag + (size_t)k * n * c.nl + l;
    const size_t fs = (size_t)nlm * n, off = (size_t)(2 * f + ri) * fs + idx;
    const double *lam = c.lam_in + off;
    double *o = c.lam_out + off;
    double *mu = c.mu_cur + off;
    double *eta = c.mu_prev ? c.mu_prev + off : nullptr;
    const bool tzero = f == 4;              // the temperature's boundary rows enter E x as 0
    const bool bcr = (c.bc_mask >> f) & 1;  // Dirichlet rows: rhs rows 0 and n-1 are constant
    const bool nlz = (c.nl_zero_mask >> f) & 1;
    const double c1 = nlz ? 0.0 : c.c1, c0 = nlz ? 0.0 : c.c0;
    const double inv_dt = 1.0 / c.dt, lw = 1.0 - c.theta;
    const size_t NL = (size_t)c.nl, NS = (size_t)nlm;
    auto cl = [&](int j) { return j < 0 ? 0 : (j > n - 1 ? n - 1 : j); };

    // ---- U^T z = lam, rows upward; su[s][q] = U(i-1-q, i)
    double sl[D], sd[D], su[D][BW];
    auto uload = [&](int s, int i) {
        const int ic = cl(i);
        sl[s] = lam[NS * ic];
        sd[s] = rdg[NL * ic];
#pragma unroll
        for (int q = 0; q < BW; ++q) su[s][q] = LU[NL * (size_t)(ic * W + BW - 1 - q)];
    };
    double zw[BW]; // zw[q] = z_{i-1-q}
#pragma unroll
    for (int q = 0; q < BW; ++q) zw[q] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) uload(d, d);
    for (int i0 = 0; i0 < n; i0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 + d;
            double v = sl[d];
#pragma unroll
            for (int q = 0; q < BW; ++q) v -= (i - 1 - q >= 0 ? su[d][q] : 0.0) * zw[q];
            v *= sd[d];
#pragma unroll
            for (int q = BW - 1; q > 0; --q) zw[q] = zw[q - 1];
            zw[0] = v;
            if (i < n) o[NS * i] = v;
            uload(d, i + D);
        }
    }

    // ---- L^T r = z, rows downward; row j = i + BW of E^T r once r_{j-BW..j+BW} is known
    double sz[D], sm[D], sL[D][BW];
    auto lload = [&](int s, int i) {
        const int ic = cl(i);
        sz[s] = o[NS * ic];
        sm[s] = mu[NS * cl(i + BW)];
#pragma unroll
        for (int q = 0; q < BW; ++q) sL[s][q] = LU[NL * (size_t)(ic * W + BW + 1 + q)]; // L(i+1+q, i)
    };
    double rw[BW]; // raw r_{i+1+q}
    double xw[W];  // effective r (0 at constant rhs rows) of rows i .. i+2BW
#pragma unroll
    for (int q = 0; q < BW; ++q) rw[q] = 0.0;
#pragma unroll
    for (int q = 0; q < W; ++q) xw[q] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) lload(d, n - 1 - d);
    for (int i0 = n - 1; i0 >= -BW; i0 -= D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 - d;
            double re = 0.0;
            if (i >= 0) {
                double v = sz[d];
#pragma unroll
                for (int q = 0; q < BW; ++q) v -= (i + 1 + q < n ? sL[d][q] : 0.0) * rw[q];
#pragma unroll
                for (int q = BW - 1; q > 0; --q) rw[q] = rw[q - 1];
                rw[0] = v;
                re = (bcr && (i == 0 || i == n - 1)) ? 0.0 : v;
            }
#pragma unroll
            for (int q = W - 1; q > 0; --q) xw[q] = xw[q - 1];
            xw[0] = re;
            const int j = i + BW; // xw[q] = r_eff of row j - BW + q
            if (j >= 0 && j < n) {
                const double *col = Lc + W * j; // nu Lap[j - BW + q][j] at col[q]
                double sum = 0.0;
#pragma unroll
                for (int q = 0; q < W; ++q) sum += xw[q] * col[q]; // rows outside 0..n-1 carry r = 0
                const double rj = xw[BW];
                double lo = inv_dt * rj + lw * (sum - nulf * r2[j] * rj);
                if (tzero && (j == 0 || j == n - 1)) lo = 0.0;
                o[NS * j] = lo;
                mu[NS * j] = sm[d] + c1 * rj;
                if (eta) eta[NS * j] = -c0 * rj;
            }
            lload(d, i - D);
        }
    }
}

template <int BW> void launch_adj_chain(hipStream_t st, const AdjChainArgs &c)
{
    const size_t lds = sizeof(double) * (size_t)(2 * BW + 2) * c.n;
    hipLaunchKernelGGL((cnab2_adj_kernel<BW, 4>), dim3((c.nlm + 63) / 64, 5, 2), dim3(64), lds, st, c);
}

int launch_adj_chain_bw(hipStream_t st, const AdjChainArgs &c, int bw)
{
    GD_CHECK(sizeof(double) * (size_t)(2 * bw + 2) * c.n <= 64 * 1024, GD_ERR_ARG,
             "adjoint: nr=%d bw=%d exceeds the LDS band table", c.n, bw);
    switch (bw) {
    case 1: launch_adj_chain<1>(st, c); break;
    case 2: launch_adj_chain<2>(st, c); break;
    case 3: launch_adj_chain<3>(st, c); break;
    case 4: launch_adj_chain<4>(st, c); break;
    case 5: launch_adj_chain<5>(st, c); break;
    case 6: launch_adj_chain<6>(st, c); break;
    case 7: launch_adj_chain<7>(st, c); break;
    default: GD_CHECK(false, GD_ERR_ARG, "adjoint: bw=%d", bw);
    }
    return gd_check_launch("cnab2_adj_kernel");
}

// ---------------------------------------------------------------- nonlinear adjoint
// Legendre-synthesis columns per level (ADJ_COLS): 0, 1 u_pol; 2, 3 u_tor;
// 4, 5 mu(u_pol) / (c_m l(l+1)); 6, 7 mu(u_tor) / (c_m l(l+1)); 8, 9 T / r;
// 10, 11 mu(T) / c_m   (re, im each; l = 0 cotangent columns 0)
constexpr int ADJ_COLS = 12;
// Fourier records: u (0..7), the force cotangent (8..15), grad T (16..23),
// the advection cotangent (24..27)
constexpr int ADJ_NSLOT = 28;
constexpr int ADJ_NSLOT_G = 16;

template <typename T>
__global__ void __launch_bounds__(256) adj_pre_kernel(const double *__restrict__ st, const double *__restrict__ mu,
                                                      const int *__restrict__ l_of_idx, const int *__restrict__ m_of_idx,
                                                      const double *__restrict__ rtab, int nlm, int nr,
                                                      T *__restrict__ syn)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)r * nlm + idx;
    const int l = l_of_idx[idx], m = m_of_idx[idx];
    const double cinv = m ? 0.5 : 1.0;
    const double clinv = l ? cinv / ((double)l * (double)(l + 1)) : 0.0;
    const double rinv = rtab[r + nr * 2];
    T *v = syn + (size_t)r * ADJ_COLS * nlm + idx;
#pragma unroll
    for (int reim = 0; reim < 2; ++reim) {
        v[(0 + reim) * (size_t)nlm] = (T)st[(2 + reim) * fs + o];
        v[(2 + reim) * (size_t)nlm] = (T)st[(0 + reim) * fs + o];
        v[(4 + reim) * (size_t)nlm] = (T)(mu[(2 + reim) * fs + o] * clinv);
        v[(6 + reim) * (size_t)nlm] = (T)(mu[(0 + reim) * fs + o] * clinv);
        v[(8 + reim) * (size_t)nlm] = (T)(st[(8 + reim) * fs + o] * rinv);
        v[(10 + reim) * (size_t)nlm] = (T)(mu[(8 + reim) * fs + o] * cinv);
    }
}

// lam += c_m (S(u*) -> u_pol, T(u*) -> u_tor, S(g*) / r -> T); tmp holds the
// analysis outputs [field][re/im][level][idx] of u* (S, T) and g* (S, T).
// With the radial forcing (cb != 0) the transpose of its spectral term
// nl(u_pol) += cb r^3 T / (l(l+1)) adds cb r^3 / (l(l+1)) mu(u_pol) to T.
__global__ void __launch_bounds__(256) adj_post_kernel(const double *__restrict__ tmp, const int *__restrict__ m_of_idx,
                                                       const int *__restrict__ l_of_idx, const double *__restrict__ rtab,
                                                       const double *__restrict__ mu, double cb, int nlm, int nr,
                                                       double *__restrict__ lam)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)r * nlm + idx;
    const double c = m_of_idx[idx] ? 2.0 : 1.0;
    const double rinv = rtab[r + nr * 2];
    const int l = l_of_idx[idx];
    const double rr = rtab[r + nr * 3];
    const double fb = (cb != 0.0 && l > 0) ? cb * rr * rr * rr / ((double)l * (double)(l + 1)) : 0.0;
#pragma unroll
    for (int reim = 0; reim < 2; ++reim) {
        lam[(2 + reim) * fs + o] += c * tmp[(0 + reim) * fs + o];
        lam[(0 + reim) * fs + o] += c * tmp[(2 + reim) * fs + o];
        lam[(8 + reim) * fs + o] += c * rinv * tmp[(4 + reim) * fs + o] + fb * mu[(2 + reim) * fs + o];
    }
}

} // namespace

// work buffers of the nonlinear adjoint, owned by the step (allocated on first use)
struct gd_adj_work {
    DevBuf syn, fb, gb, tmp, m_of_idx;
};

static int adj_work(gd_step *s, gd_adj_work **out)
{
    if (!s->adjw) {
        gd_adj_work *w = new gd_adj_work();
        const size_t esz = s->d.prec == GD_PREC_FP64 ? sizeof(double) : sizeof(float);
        const size_t nm = s->d.mmax + 1, nr = s->d.nr;
        int rc = w->syn.ensure(esz * s->fs * ADJ_COLS);
        if (rc == GD_OK) rc = w->fb.ensure(esz * (size_t)s->sht.nlat2 * nr * nm * ADJ_NSLOT);
        if (rc == GD_OK) rc = w->gb.ensure(esz * nm * s->sht.nlat2 * nr * ADJ_NSLOT_G);
        if (rc == GD_OK) rc = w->tmp.ensure(sizeof(double) * s->fs * 8);
        if (rc == GD_OK) rc = w->m_of_idx.ensure(sizeof(int) * s->nlm);
        if (rc == GD_OK &&
            hipMemcpy(w->m_of_idx.p, s->sht.maps.m_of_idx.data(), sizeof(int) * s->nlm, hipMemcpyHostToDevice) !=
                hipSuccess) {
            gd_set_error("adjoint: upload of the order table failed");
            rc = GD_ERR_HIP;
        }
        if (rc != GD_OK) {
            for (DevBuf *b : {&w->syn, &w->fb, &w->gb, &w->tmp, &w->m_of_idx}) b->release();
            delete w;
            return rc;
        }
        s->adjw = w;
    }
    *out = s->adjw;
    return GD_OK;
}

void gd_adj_work_release(gd_step *s)
{
    if (!s->adjw) return;
    gd_adj_work *w = s->adjw;
    for (DevBuf *b : {&w->syn, &w->fb, &w->gb, &w->tmp, &w->m_of_idx}) b->release();
    delete w;
    s->adjw = nullptr;
}

// lam += N'(xs)^T mu (m-major planes)
template <typename T>
static int nonlinear_adjoint_T(gd_step *s, const double *xs, const double *mu, double *lam)
{
    gd_adj_work *w = nullptr;
    GD_TRY(adj_work(s, &w));
    hipStream_t st = s->ctx->stream;
    gd_sht_plan *p = &s->sht;
    const int nr = s->d.nr, nlm = s->nlm;
    const size_t fs = s->fs;
    hipLaunchKernelGGL((adj_pre_kernel<T>), dim3((nlm + 255) / 256, nr), dim3(256), 0, st, xs, mu,
                       p->d_l_of_idx.as<int>(), w->m_of_idx.as<int>(), s->rtab.as<double>(), nlm, nr, w->syn.as<T>());
    GD_TRY(gd_check_launch("adj_pre_kernel"));
    SynthArgs a{};
    a.mm = p->prec == GD_PREC_BF16X3;
    a.moff = p->d_moff.as<int>();
    a.ne = p->d_ne.as<int>();
    a.nlat2 = p->nlat2;
    a.nr = nr;
    a.nlm = nlm;
    a.mmax = p->mmax;
    a.out = w->fb.p;
    a.in = w->syn.p;
    a.cin = ADJ_COLS;
    a.nslot = ADJ_NSLOT;
    // u and the force cotangent as full fields, grad T as a gradient field
    // (S = T/r only: its zero toroidal columns are not contracted)
    a.nf = 2;
    a.ng = 1;
    for (int k = 0; k < 20; ++k) a.colmap[k] = k < 10 ? k : -1;
    a.tab0 = gd_sht_tab(p, 1);
    a.tab1 = gd_sht_tab(p, 2);
    a.slot0 = 0;
    GD_TRY(gd_launch_synth_vec<T>(st, a));
    a.nf = 1;
    for (int k = 0; k < 20; ++k) a.colmap[k] = k < 2 ? 10 + k : -1;
    a.tab0 = gd_sht_tab(p, 0);
    a.tab1 = nullptr;
    a.slot0 = 24;
    GD_TRY(gd_launch_synth_scal<T>(st, a));
    RingArgs r{};
    gd_sht_fill_ring(p, r, nr);
    r.fb = w->fb.p;
    r.gb = w->gb.p;
    r.nslot_f = ADJ_NSLOT;
    r.nslot_g = ADJ_NSLOT_G;
    r.nfield = 7;
    static const int fn[7] = {0, 2, 8, 10, 16, 18, 24}, fsou[7] = {4, 6, 12, 14, 20, 22, 26};
    for (int F = 0; F < 7; ++F) {
        r.fn_off[F] = fn[F];
        r.fs_off[F] = fsou[F];
    }
    r.ro = s->d.Pm / s->d.E;
    r.coriolis_index = s->d.coriolis_index;
    GD_TRY(gd_launch_ring<T>(st, RING_ADJ_TANG, r));
    AnaArgs g{};
    g.mm = p->prec == GD_PREC_BF16X3;
    g.gb = w->gb.p;
    g.nslot_g = ADJ_NSLOT_G;
    g.gslot0 = 0;
    g.nf = 2;
    g.moff = p->d_moff.as<int>();
    g.ne = p->d_ne.as<int>();
    g.l_of_idx = p->d_l_of_idx.as<int>();
    g.nlat2 = p->nlat2;
    g.nr = nr;
    g.nlm = nlm;
    g.mmax = p->mmax;
    g.tab0 = gd_sht_tab(p, 1);
    g.tab1 = gd_sht_tab(p, 2);
    double *tmp = w->tmp.as<double>();
    for (int k = 0; k < 4; ++k) {
        g.out[k] = tmp + 2 * k * fs;
        g.scale[k] = 1;
    }
    GD_TRY(gd_launch_ana<T>(st, g, true, (p->lmax + 2) / 2));
    hipLaunchKernelGGL(adj_post_kernel, dim3((nlm + 255) / 256, nr), dim3(256), 0, st, tmp, w->m_of_idx.as<int>(),
                       p->d_l_of_idx.as<int>(), s->rtab.as<double>(), mu, gd_step_buoyancy_factor(s), nlm, nr, lam);
    return gd_check_launch("adj_post_kernel");
}

static int nonlinear_adjoint(gd_step *s, const double *xs, const double *mu, double *lam)
{
    if (s->d.prec == GD_PREC_FP64) return nonlinear_adjoint_T<double>(s, xs, mu, lam);
    return nonlinear_adjoint_T<float>(s, xs, mu, lam);
}

static int adjoint_supported(const gd_step *s, const char *what)
{
    GD_CHECK(gd_step_tangential(s), GD_ERR_ARG,
             "%s: the adjoint is built for the tangential step (radials = 0, full_products = 0)", what);
    GD_CHECK(!s->tbc && !s->comp, GD_ERR_ARG,
             "%s: the adjoint does not cover temperature conditions or composition", what);
    return GD_OK;
}

// The transposed step k at state xs: lam_out = E^T r + N'(xs)^T mu_cur,
// mu_cur += c1 r, mu_prev = -c0 r (not written at the bootstrap)
static int step_vjp(gd_step *s, const double *xs, bool boot, const double *lam_in, double *lam_out, double *mu_cur,
                    double *mu_prev, bool nonlinear_terms = true)
{
    AdjChainArgs c{};
    c.nuLap = s->nuLap.as<double>();
    c.rinv2 = s->rtab.as<double>() + s->d.nr;
    c.nus = s->nus.as<double>();
    c.LUt = s->LUt.as<double>();
    c.rdiag = s->rdiag.as<double>();
    c.l_of_idx = s->sht.d_l_of_idx.as<int>();
    c.nl = s->nl;
    c.n = s->d.nr;
    c.nlm = s->nlm;
    c.dt = s->d.dt;
    c.theta = s->d.theta;
    c.bc_mask = s->d.bc_rows == 1 ? 0x1f : (s->d.bc_rows == 2 ? 0x10 : 0);
    c.lam_in = lam_in;
    c.lam_out = lam_out;
    c.mu_cur = mu_cur;
    c.mu_prev = boot ? nullptr : mu_prev;
    c.c1 = boot ? 1.0 : 1.5;
    c.c0 = boot ? 0.0 : 0.5;
    c.nl_zero_mask = 0x0c; // as the forward tangential step (gd_step.hip, ChainArgs::nl_zero_mask)
    GD_TRY(launch_adj_chain_bw(s->ctx->stream, c, s->d.bw));
    return nonlinear_terms ? nonlinear_adjoint(s, xs, mu_cur, lam_out) : GD_OK;
}

// l-major (nlm, nr) planes of the five fields <-> m-major device planes
static int planes_to_device(gd_step *s, const double *re, const double *im, double *dev)
{
    const int *map = s->sht.d_lmaj_of_idx.as<int>();
    for (int f = 0; f < 5; ++f) {
        const size_t o = (size_t)f * s->fs;
        GD_TRY(gd_launch_perm(s->ctx->stream, re + o, map, s->nlm, s->d.nr, dev + (2 * f) * s->fs));
        GD_TRY(gd_launch_perm(s->ctx->stream, im + o, map, s->nlm, s->d.nr, dev + (2 * f + 1) * s->fs));
    }
    return GD_OK;
}

static int planes_from_device(gd_step *s, const double *dev, double *re, double *im)
{
    const int *map = s->sht.d_idx_of_lmaj.as<int>();
    for (int f = 0; f < 5; ++f) {
        const size_t o = (size_t)f * s->fs;
        GD_TRY(gd_launch_perm(s->ctx->stream, dev + (2 * f) * s->fs, map, s->nlm, s->d.nr, re + o));
        GD_TRY(gd_launch_perm(s->ctx->stream, dev + (2 * f + 1) * s->fs, map, s->nlm, s->d.nr, im + o));
    }
    return GD_OK;
}

// host or device planes in (GD_HOST staged through buf), device planes out
static int stage_planes(gd_step *s, DevBuf &buf, const double *re, const double *im, int flags, const double **dre,
                        const double **dim)
{
    const size_t n = s->fs * 5;
    if (!(flags & GD_HOST)) {
        *dre = re;
        *dim = im;
        return GD_OK;
    }
    GD_TRY(buf.ensure(sizeof(double) * n * 2));
    GD_HIP(hipMemcpyAsync(buf.p, re, sizeof(double) * n, hipMemcpyHostToDevice, s->ctx->stream));
    GD_HIP(hipMemcpyAsync(buf.as<double>() + n, im, sizeof(double) * n, hipMemcpyHostToDevice, s->ctx->stream));
    *dre = buf.as<double>();
    *dim = *dre + n;
    return GD_OK;
}

static int unstage_planes(gd_step *s, DevBuf &buf, const double *dev, double *re, double *im, int flags)
{
    const size_t n = s->fs * 5;
    if (!(flags & GD_HOST)) return planes_from_device(s, dev, re, im);
    GD_TRY(buf.ensure(sizeof(double) * n * 2));
    GD_TRY(planes_from_device(s, dev, buf.as<double>(), buf.as<double>() + n));
    GD_TRY(gd_from_device(s->ctx, re, buf.p, sizeof(double) * n));
    GD_TRY(gd_from_device(s->ctx, im, buf.as<double>() + n, sizeof(double) * n));
    return GD_OK;
}

static double *step_state(gd_step *s) { return s->state[s->step & 1].as<double>(); }

extern "C" int gd_step_nonlinear_adjoint(gd_step *s, const double *mu_re, const double *mu_im, double *lam_re,
                                         double *lam_im, int flags)
{
    GD_CHECK(s && mu_re && mu_im && lam_re && lam_im, GD_ERR_ARG, "gd_step_nonlinear_adjoint: null argument");
    GD_TRY(adjoint_supported(s, "gd_step_nonlinear_adjoint"));
    GD_HIP(hipSetDevice(s->ctx->device));
    const size_t pb = sizeof(double) * s->fs * 10;
    DevBuf io, mu, lam;
    const double *a = nullptr, *b = nullptr;
    int rc = stage_planes(s, io, mu_re, mu_im, flags, &a, &b);
    if (rc == GD_OK) rc = mu.ensure(pb);
    if (rc == GD_OK) rc = lam.ensure(pb);
    if (rc == GD_OK && hipMemsetAsync(lam.p, 0, pb, s->ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    if (rc == GD_OK) rc = planes_to_device(s, a, b, mu.as<double>());
    if (rc == GD_OK) rc = nonlinear_adjoint(s, step_state(s), mu.as<double>(), lam.as<double>());
    if (rc == GD_OK) rc = unstage_planes(s, io, lam.as<double>(), lam_re, lam_im, flags);
    if (rc == GD_OK && hipStreamSynchronize(s->ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    io.release();
    mu.release();
    lam.release();
    if (rc == GD_ERR_HIP) gd_set_error("gd_step_nonlinear_adjoint: HIP failure");
    return rc;
}

extern "C" int gd_step_vjp(gd_step *s, const double *lam_re, const double *lam_im, const double *eta_re,
                           const double *eta_im, double *olam_re, double *olam_im, double *oeta_re, double *oeta_im,
                           int flags)
{
    GD_CHECK(s && lam_re && lam_im && eta_re && eta_im && olam_re && olam_im && oeta_re && oeta_im, GD_ERR_ARG,
             "gd_step_vjp: null argument");
    GD_TRY(adjoint_supported(s, "gd_step_vjp"));
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const size_t pb = sizeof(double) * s->fs * 10;
    const bool boot = s->step == 0;
    DevBuf io, lin, lout, mu, mp;
    const double *a = nullptr, *b = nullptr;
    int rc = lin.ensure(pb);
    if (rc == GD_OK) rc = lout.ensure(pb);
    if (rc == GD_OK) rc = mu.ensure(pb);
    if (rc == GD_OK) rc = mp.ensure(pb);
    if (rc == GD_OK && hipMemsetAsync(mp.p, 0, pb, st) != hipSuccess) rc = GD_ERR_HIP;
    if (rc == GD_OK) rc = stage_planes(s, io, lam_re, lam_im, flags, &a, &b);
    if (rc == GD_OK) rc = planes_to_device(s, a, b, lin.as<double>());
    if (rc == GD_OK) rc = stage_planes(s, io, eta_re, eta_im, flags, &a, &b);
    if (rc == GD_OK) rc = planes_to_device(s, a, b, mu.as<double>());
    if (rc == GD_OK)
        rc = step_vjp(s, step_state(s), boot, lin.as<double>(), lout.as<double>(), mu.as<double>(), mp.as<double>());
    if (rc == GD_OK) rc = unstage_planes(s, io, lout.as<double>(), olam_re, olam_im, flags);
    if (rc == GD_OK) rc = unstage_planes(s, io, mp.as<double>(), oeta_re, oeta_im, flags);
    if (rc == GD_OK && hipStreamSynchronize(st) != hipSuccess) rc = GD_ERR_HIP;
    for (DevBuf *x : {&io, &lin, &lout, &mu, &mp}) x->release();
    if (rc == GD_ERR_HIP) gd_set_error("gd_step_vjp: HIP failure");
    return rc;
}

// ---------------------------------------------------------------- checkpointed adjoint of a run
struct gd_adjoint {
    gd_step *s = nullptr;
    int K = 0, P = 0, nseg = 0;
    std::vector<DevBuf> ck_state, ck_hist; // state and AB2 history at steps 0, P, 2P, ...
    DevBuf final_state, final_hist;        // x_K and its history, restored after the backward pass
    DevBuf seg;                            // states x_{sP} .. x_{sP+P-1} of the segment being reversed
    DevBuf lam[2], mu[2];                  // cotangents of x_k (lam[k & 1]) and nl_k (mu[k & 1])
    bool forward_done = false, seeded = false;
    bool last_ok = false;                  // seg holds the last segment's states (from the forward pass)
    // activity of the seeded cotangent: whether it reaches u or T.  In the
    // tangential step u and T are coupled through the nonlinear terms while
    // B_tor, B_pol only through their own (linear) solves, so a cotangent
    // with no u or T part keeps none: the transposed nonlinear terms are then
    // exact zeros and the states need not be recomputed.
    bool active_uT = true;
    // surface misfit: FP64 transform plan of the step's grid, grid buffers
    gd_sht_plan mplan;
    bool mplan_ok = false;
    DevBuf mcol, mgrid, ones, mout;
    std::vector<double> hgrid;
};

static int adj_release(gd_adjoint *a)
{
    for (DevBuf &b : a->ck_state) b.release();
    for (DevBuf &b : a->ck_hist) b.release();
    for (DevBuf *b : {&a->final_state, &a->final_hist, &a->seg, &a->lam[0], &a->lam[1], &a->mu[0], &a->mu[1],
                      &a->mcol, &a->mgrid, &a->ones, &a->mout})
        b->release();
    if (a->mplan_ok) gd_sht_plan_release(&a->mplan);
    a->mplan_ok = false;
    return GD_OK;
}

extern "C" int gd_adjoint_create(gd_step *s, int nsteps, int every, gd_adjoint **out)
{
    GD_CHECK(s && out && nsteps >= 1 && every >= 1, GD_ERR_ARG, "gd_adjoint_create: bad arguments");
    GD_TRY(adjoint_supported(s, "gd_adjoint_create"));
    GD_HIP(hipSetDevice(s->ctx->device));
    gd_adjoint *a = new gd_adjoint();
    a->s = s;
    a->K = nsteps;
    a->P = every < nsteps ? every : nsteps;
    a->nseg = (nsteps + a->P - 1) / a->P;
    const size_t pb = sizeof(double) * s->fs * 10;
    a->ck_state.resize(a->nseg);
    a->ck_hist.resize(a->nseg);
    int rc = GD_OK;
    for (int k = 0; k < a->nseg && rc == GD_OK; ++k) {
        rc = a->ck_state[k].ensure(pb);
        if (rc == GD_OK && k > 0) rc = a->ck_hist[k].ensure(pb);
    }
    for (DevBuf *b : {&a->final_state, &a->final_hist, &a->lam[0], &a->lam[1], &a->mu[0], &a->mu[1]})
        if (rc == GD_OK) rc = b->ensure(pb);
    if (rc == GD_OK) rc = a->seg.ensure(pb * a->P);
    gd_adj_work *w = nullptr;
    if (rc == GD_OK) rc = adj_work(s, &w);
    if (rc != GD_OK) {
        adj_release(a);
        delete a;
        return rc;
    }
    *out = a;
    return GD_OK;
}

extern "C" int gd_adjoint_destroy(gd_adjoint *a)
{
    if (!a) return GD_OK;
    (void)hipSetDevice(a->s->ctx->device);
    (void)hipStreamSynchronize(a->s->ctx->stream);
    adj_release(a);
    delete a;
    return GD_OK;
}

static int copy_planes(gd_step *s, void *dst, const void *src)
{
    GD_HIP(hipMemcpyAsync(dst, src, sizeof(double) * s->fs * 10, hipMemcpyDeviceToDevice, s->ctx->stream));
    return GD_OK;
}

// the step at step k with state and history copied in (its synthesis columns
// are rebuilt by its next step)
static int set_step_at(gd_step *s, int k, const void *state, const void *hist)
{
    s->step = k;
    s->syn_ok = false;
    GD_TRY(copy_planes(s, s->state[k & 1].p, state));
    if (k > 0) GD_TRY(copy_planes(s, s->nlbuf[(k & 1) ^ 1].p, hist));
    return GD_OK;
}

extern "C" int gd_adjoint_forward(gd_adjoint *a)
{
    GD_CHECK(a, GD_ERR_ARG, "gd_adjoint_forward: null argument");
    gd_step *s = a->s;
    GD_CHECK(s->step == 0, GD_ERR_ARG, "gd_adjoint_forward: the step must be at step 0 (it is at %d)", s->step);
    GD_TRY(adjoint_supported(s, "gd_adjoint_forward"));
    GD_HIP(hipSetDevice(s->ctx->device));
    // the last segment's states are kept as they are produced: the backward
    // pass reverses that segment first and needs no recomputation for it
    const int klast = (a->nseg - 1) * a->P;
    const size_t np = s->fs * 10;
    for (int k = 0; k < a->K; ++k) {
        if (k % a->P == 0) {
            GD_TRY(copy_planes(s, a->ck_state[k / a->P].p, s->state[k & 1].p));
            if (k > 0) GD_TRY(copy_planes(s, a->ck_hist[k / a->P].p, s->nlbuf[(k & 1) ^ 1].p));
        }
        if (k >= klast) GD_TRY(copy_planes(s, a->seg.as<double>() + (size_t)(k - klast) * np, s->state[k & 1].p));
        GD_TRY(gd_step_enqueue(s, k & 1, k == 0));
        ++s->step;
    }
    GD_TRY(copy_planes(s, a->final_state.p, s->state[a->K & 1].p));
    GD_TRY(copy_planes(s, a->final_hist.p, s->nlbuf[(a->K & 1) ^ 1].p));
    a->forward_done = true;
    a->seeded = false;
    a->last_ok = true;
    return GD_OK;
}

extern "C" int gd_adjoint_seed(gd_adjoint *a, const double *re, const double *im, int flags)
{
    GD_CHECK(a && re && im, GD_ERR_ARG, "gd_adjoint_seed: null argument");
    gd_step *s = a->s;
    GD_HIP(hipSetDevice(s->ctx->device));
    DevBuf io;
    const double *x = nullptr, *y = nullptr;
    int rc = stage_planes(s, io, re, im, flags, &x, &y);
    if (rc == GD_OK) rc = planes_to_device(s, x, y, a->lam[a->K & 1].as<double>());
    if (rc == GD_OK && (flags & GD_HOST) && hipStreamSynchronize(s->ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    io.release();
    if (rc == GD_ERR_HIP) gd_set_error("gd_adjoint_seed: HIP failure");
    GD_TRY(rc);
    a->seeded = true;
    a->active_uT = true;
    return GD_OK;
}

namespace {
// the level `lev` of plane pair (p0, p0 + 1) times fac_l (l(l+1)/r^2 for the
// radial component of a poloidal field, 1 for a scalar) as two spectral
// columns [re/im][idx]
__global__ void misfit_pack_kernel(const double *__restrict__ st, int p0, int lev, const int *__restrict__ l_of_idx,
                                   double rinv2, int radial, int nlm, int nr, double *__restrict__ col)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)lev * nlm + idx;
    const int l = l_of_idx[idx];
    const double f = radial ? (double)l * (double)(l + 1) * rinv2 : 1.0;
    col[idx] = f * st[(size_t)p0 * fs + o];
    col[nlm + idx] = f * st[(size_t)(p0 + 1) * fs + o];
}

// lam = 0 except level lev of planes (p0, p0 + 1): fac_l c_m times the
// unweighted analysis out [re/im][idx] of the residual (the synthesis transposed)
__global__ void misfit_seed_kernel(const double *__restrict__ out, int p0, int lev, const int *__restrict__ l_of_idx,
                                   const int *__restrict__ m_of_idx, double rinv2, int radial, int nlm, int nr,
                                   double *__restrict__ lam)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)lev * nlm + idx;
    const int l = l_of_idx[idx];
    const double f = (radial ? (double)l * (double)(l + 1) * rinv2 : 1.0) * (m_of_idx[idx] ? 2.0 : 1.0);
    lam[(size_t)p0 * fs + o] = f * out[idx];
    lam[(size_t)(p0 + 1) * fs + o] = f * out[nlm + idx];
}
} // namespace

// J = 1/2 sum over the grid of (X - target)^2 with X the field `kind` at level
// `lev` synthesised on the step's Gauss grid in FP64; seeds lam_K = dJ/dx_K
extern "C" int gd_adjoint_misfit(gd_adjoint *a, int kind, int level, const double *target, double *J, int flags)
{
    GD_CHECK(a && target && J, GD_ERR_ARG, "gd_adjoint_misfit: null argument");
    GD_CHECK(a->forward_done, GD_ERR_ARG, "gd_adjoint_misfit: run gd_adjoint_forward first");
    gd_step *s = a->s;
    GD_CHECK(kind >= 0 && kind <= 2, GD_ERR_ARG, "gd_adjoint_misfit: kind %d (0 B_r, 1 u_r, 2 T)", kind);
    GD_CHECK(level >= 0 && level < s->d.nr, GD_ERR_ARG, "gd_adjoint_misfit: level %d outside 0..%d", level,
             s->d.nr - 1);
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const int nlm = s->nlm, nlat = s->d.nlat, nlon = s->d.nlon, nr = s->d.nr;
    if (!a->mplan_ok) {
        GD_TRY(gd_sht_plan_build(s->ctx, s->d.lmax, s->d.mmax, nlat, nlon, GD_PREC_FP64, &a->mplan));
        a->mplan_ok = true;
        GD_TRY(a->mcol.ensure(sizeof(double) * 2 * nlm));
        GD_TRY(a->mout.ensure(sizeof(double) * 2 * nlm));
        GD_TRY(a->mgrid.ensure(sizeof(double) * nlat * nlon));
        std::vector<double> one(nlat / 2, 1.0);
        GD_TRY(a->ones.ensure(sizeof(double) * one.size()));
        GD_HIP(hipMemcpy(a->ones.p, one.data(), sizeof(double) * one.size(), hipMemcpyHostToDevice));
        a->hgrid.resize((size_t)nlat * nlon);
    }
    gd_sht_plan *p = &a->mplan;
    gd_adj_work *w = nullptr;
    GD_TRY(adj_work(s, &w));
    std::vector<double> r((size_t)nr * 7);
    GD_TRY(gd_radial_grid(nr, s->d.rratio, r.data()));
    const double rinv2 = r[(size_t)nr + level];
    const int p0 = kind == 0 ? 6 : (kind == 1 ? 2 : 8), radial = kind < 2;
    const double *xK = a->final_state.as<double>();
    hipLaunchKernelGGL(misfit_pack_kernel, dim3((nlm + 255) / 256), dim3(256), 0, st, xK, p0, level,
                       p->d_l_of_idx.as<int>(), rinv2, radial, nlm, nr, a->mcol.as<double>());
    GD_TRY(gd_check_launch("misfit_pack_kernel"));
    const int nm = p->mmax + 1;
    GD_TRY(p->fb.ensure(sizeof(double) * (size_t)p->nlat2 * nm * 4));
    GD_TRY(p->gb.ensure(sizeof(double) * (size_t)nm * p->nlat2 * 4));
    SynthArgs sa{};
    sa.in = a->mcol.p;
    sa.cin = 2;
    sa.nf = 1;
    for (int k = 0; k < 20; ++k) sa.colmap[k] = k < 2 ? k : -1;
    sa.moff = p->d_moff.as<int>();
    sa.ne = p->d_ne.as<int>();
    sa.nlat2 = p->nlat2;
    sa.nr = 1;
    sa.nlm = nlm;
    sa.mmax = p->mmax;
    sa.out = p->fb.p;
    sa.nslot = 4;
    sa.slot0 = 0;
    sa.tab0 = gd_sht_tab(p, 0);
    GD_TRY(gd_launch_synth_scal<double>(st, sa));
    RingArgs rg{};
    gd_sht_fill_ring(p, rg, 1);
    rg.fb = p->fb.p;
    rg.nslot_f = 4;
    rg.nfield = 1;
    rg.fn_off[0] = 0;
    rg.fs_off[0] = 2;
    rg.phys_out[0] = a->mgrid.as<double>();
    GD_TRY(gd_launch_ring<double>(st, RING_SYNTH, rg));
    GD_TRY(gd_from_device(s->ctx, a->hgrid.data(), a->mgrid.p, sizeof(double) * nlat * nlon));
    const size_t ng = (size_t)nlat * nlon;
    std::vector<double> tgt;
    const double *th = target;
    if (!(flags & GD_HOST)) {
        tgt.resize(ng);
        GD_TRY(gd_from_device(s->ctx, tgt.data(), target, sizeof(double) * ng));
        th = tgt.data();
    }
    double acc = 0.0;
    for (size_t q = 0; q < ng; ++q) {
        const double d = a->hgrid[q] - th[q];
        a->hgrid[q] = d;
        acc += d * d;
    }
    *J = 0.5 * acc;
    GD_HIP(hipMemcpyAsync(a->mgrid.p, a->hgrid.data(), sizeof(double) * ng, hipMemcpyHostToDevice, st));
    // synthesis transposed: the analysis with unit weights, then c_m and fac_l
    RingArgs ra{};
    gd_sht_fill_ring(p, ra, 1);
    ra.wq = a->ones.p;
    ra.gb = p->gb.p;
    ra.nslot_g = 4;
    ra.nfield = 1;
    ra.phys_in[0] = a->mgrid.as<double>();
    GD_TRY(gd_launch_ring<double>(st, RING_ANALYSIS, ra));
    AnaArgs g{};
    g.gb = p->gb.p;
    g.nslot_g = 4;
    g.gslot0 = 0;
    g.nf = 1;
    g.moff = p->d_moff.as<int>();
    g.ne = p->d_ne.as<int>();
    g.l_of_idx = p->d_l_of_idx.as<int>();
    g.nlat2 = p->nlat2;
    g.nr = 1;
    g.nlm = nlm;
    g.mmax = p->mmax;
    g.tab0 = gd_sht_tab(p, 0);
    g.out[0] = a->mout.as<double>();
    g.scale[0] = 1;
    GD_TRY(gd_launch_ana<double>(st, g, false, (p->lmax + 2) / 2));
    double *lamK = a->lam[a->K & 1].as<double>();
    GD_HIP(hipMemsetAsync(lamK, 0, sizeof(double) * s->fs * 10, st));
    hipLaunchKernelGGL(misfit_seed_kernel, dim3((nlm + 255) / 256), dim3(256), 0, st, a->mout.as<double>(), p0, level,
                       p->d_l_of_idx.as<int>(), w->m_of_idx.as<int>(), rinv2, radial, nlm, nr, lamK);
    GD_TRY(gd_check_launch("misfit_seed_kernel"));
    GD_HIP(hipStreamSynchronize(st));
    a->seeded = true;
    a->active_uT = kind != 0; // B_r seeds B_pol only
    return GD_OK;
}

extern "C" int gd_adjoint_backward(gd_adjoint *a, double *g_re, double *g_im, int flags)
{
    GD_CHECK(a && g_re && g_im, GD_ERR_ARG, "gd_adjoint_backward: null argument");
    GD_CHECK(a->forward_done && a->seeded, GD_ERR_ARG,
             "gd_adjoint_backward: run gd_adjoint_forward and seed the final cotangent first");
    gd_step *s = a->s;
    GD_TRY(adjoint_supported(s, "gd_adjoint_backward"));
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const size_t pb = sizeof(double) * s->fs * 10;
    // mu_{K-1}: the final history enters no functional
    GD_HIP(hipMemsetAsync(a->mu[(a->K - 1) & 1].p, 0, pb, st));
    double *seg = a->seg.as<double>();
    const size_t np = s->fs * 10;
    for (int sg = a->nseg - 1; sg >= 0; --sg) {
        const int k0 = sg * a->P, k1 = k0 + a->P < a->K ? k0 + a->P : a->K;
        // recompute the segment's states x_{k0} .. x_{k1-1} (the last segment's
        // are still in seg from the forward pass)
        const bool kept = sg == a->nseg - 1 && a->last_ok;
        const bool need = a->active_uT; // the states enter only the transposed nonlinear terms
        if (need && !kept) GD_TRY(copy_planes(s, seg, a->ck_state[sg].p));
        if (need && !kept && k1 - k0 > 1) {
            GD_TRY(set_step_at(s, k0, a->ck_state[sg].p, a->ck_hist[sg].p));
            for (int k = k0; k < k1 - 1; ++k) {
                GD_TRY(gd_step_enqueue(s, k & 1, k == 0));
                ++s->step;
                GD_TRY(copy_planes(s, seg + (size_t)(k + 1 - k0) * np, s->state[(k + 1) & 1].p));
            }
        }
        for (int k = k1 - 1; k >= k0; --k)
            GD_TRY(step_vjp(s, seg + (size_t)(k - k0) * np, k == 0, a->lam[(k + 1) & 1].as<double>(),
                            a->lam[k & 1].as<double>(), a->mu[k & 1].as<double>(),
                            k > 0 ? a->mu[(k - 1) & 1].as<double>() : nullptr, need));
    }
    // seg now holds segment 0's states (the last one's only if it is segment 0,
    // or when no segment was recomputed)
    a->last_ok = a->last_ok && (a->nseg == 1 || !a->active_uT);
    // the step returns to the end of the forward run
    GD_TRY(set_step_at(s, a->K, a->final_state.p, a->final_hist.p));
    DevBuf io;
    int rc = unstage_planes(s, io, a->lam[0].as<double>(), g_re, g_im, flags);
    if (rc == GD_OK && hipStreamSynchronize(st) != hipSuccess) rc = GD_ERR_HIP;
    io.release();
    if (rc == GD_ERR_HIP) gd_set_error("gd_adjoint_backward: HIP failure");
    a->seeded = false;
    return rc;
}
