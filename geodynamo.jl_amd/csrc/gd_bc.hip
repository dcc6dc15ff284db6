// gd_bc.hip -- boundary conditions of a scalar field (temperature,
// composition) in spectral space:
//   apply_temperature_boundary_conditions_spectral! (thermal.jl:262-313),
//   apply_composition_boundary_conditions_spectral! (compositional.jl:318-370),
//   apply_scalar_flux_bc_spectral!(...; method=:tau)
//     (scalar_field_common.jl:1166-1199, apply_flux_bc_tau! :866-929).
// Per mode: bc type 1 sets the boundary row to its value (imag 0); type 2
// adds a combination of the tau functions (the two highest Chebyshev
// polynomials on the grid) to the radial profile so that the flux error of
// the real profile at the boundaries (D1 rows 0 and N-1) is removed -- a
// rank-2 update of the column.  The imaginary profile receives the same
// correction, and only when one of its entries exceeds 1e-12 in magnitude
// (the reference computes the coefficients from the real profile alone).
//
// Tau tables (FP64, host, once per grid): tau1 = T_{N-1}(x), tau2 = T_N(x) on
// x = 2 (r - r_i)/(r_o - r_i) - 1; end derivatives of T_n by the closed form
// n^2 sign(x)^(n+1) * 2/(r_o - r_i) (evaluate_chebyshev_derivative, :596-622);
// single-boundary tau functions T_{N-1} - a T_N with a the ratio of end
// derivatives at the other boundary (:733-765), whose own derivative at their
// boundary is taken through D1 (:771-790).
//
// Device: tiles of 16 consecutive (l, m) columns x 16 level lanes of the
// reference's (nlm, nr) planes; each column is read and written once, the
// imaginary one read twice (HBM-bound: 4 * 8 * nr bytes per mode with a flux
// condition).
#include <cmath>
#include <vector>

#include "gd_internal.hpp"

namespace {
struct BcArgs {
    const int *types;     // [nlm][2] inner, outer
    const double *vals;   // [nlm][2] inner, outer (boundary_values[:, lm])
    const double *tab;    // [4][nr]: tau1, tau2, tau_in, tau_out
    const double *d1;     // D1 row 0 (columns 0..bw), then row nr-1 (columns nr-1-bw..nr-1)
    int nlm, nr, bw, dirichlet, inner_row; // inner_row 0: r_i = 0 (no inner Dirichlet row)
    double dt1i, dt1o, dt2i, dt2o, dtau_in, dtau_out;
};

// Block: BM = 16 consecutive modes x 16 level lanes (256 threads), so the
// grid has nlm/16 blocks and each row segment is 128 contiguous bytes.
//   A  every thread scans its levels of the imaginary plane (rows a type-1
//      condition zeroes count as 0) for the 1e-12 rule;
//   B  one thread per mode forms the flux coefficients from the D1 boundary
//      rows of the real profile (with the type-1 rows already set);
//   C  every thread writes its levels: type-1 rows, then + correction.
constexpr int BM = 16, BL = 16;

__global__ void __launch_bounds__(256) scalar_bc_kernel(BcArgs a, double *__restrict__ re, double *__restrict__ im)
{
    __shared__ double sc1[BM], sc2[BM];
    __shared__ int skind[BM];
    __shared__ int sflag[BL][BM];
    const int tx = threadIdx.x % BM, ty = threadIdx.x / BM;
    const int k = blockIdx.x * BM + tx;
    const bool valid = k < a.nlm;
    const size_t s = (size_t)a.nlm;
    const int n = a.nr;
    int ti = 0, to = 0;
    double vi = 0.0, vo = 0.0;
    if (valid) {
        ti = a.types[2 * k];
        to = a.types[2 * k + 1];
        vi = a.vals[2 * k];
        vo = a.vals[2 * k + 1];
    }
    const bool dir_in = a.dirichlet && a.inner_row && ti == 1, dir_out = a.dirichlet && to == 1;
    const bool ai = ti == 2, ao = to == 2, flux = valid && (ai || ao);
    // A: imaginary entries above 1e-12 (after the type-1 rows are zeroed)
    int f = 0;
    if (flux)
        for (int i = ty; i < n; i += BL) {
            const bool zeroed = (i == 0 && dir_in) || (i == n - 1 && dir_out);
            f |= (!zeroed && fabs(im[(size_t)i * s + k]) > 1e-12) ? 1 : 0;
        }
    sflag[ty][tx] = f;
    // B: coefficients of the tau functions
    if (ty == 0) {
        int kind = 0;
        double c1 = 0.0, c2 = 0.0;
        if (flux) {
            double ci = 0.0, co = 0.0; // D1 rows 0 and n-1 of the real profile, ascending columns
            for (int j = 0; j <= a.bw; ++j) {
                const double v = (j == 0 && dir_in) ? vi : re[(size_t)j * s + k];
                ci += a.d1[j] * v;
            }
            for (int j = 0; j <= a.bw; ++j) {
                const int row = n - 1 - a.bw + j;
                const double v = (row == n - 1 && dir_out) ? vo : re[(size_t)row * s + k];
                co += a.d1[a.bw + 1 + j] * v;
            }
            const double fi = ai ? vi : 0.0, fo = ao ? vo : 0.0;
            if (ai && ao) {
                const double ei = fi - ci, eo = fo - co;
                const double det = a.dt1i * a.dt2o - a.dt1o * a.dt2i;
                if (fabs(det) > 1e-12) {
                    c1 = (ei * a.dt2o - eo * a.dt2i) / det;
                    c2 = (eo * a.dt1i - ei * a.dt1o) / det;
                }
                kind = 1;
            } else if (ai) {
                c1 = (fi - ci) / a.dtau_in;
                kind = 2;
            } else {
                c1 = (fo - co) / a.dtau_out;
                kind = 3;
            }
        }
        sc1[tx] = c1;
        sc2[tx] = c2;
        skind[tx] = kind;
    }
    __syncthreads();
    if (!valid) return;
    const int kind = skind[tx];
    if (kind == 0 && !dir_in && !dir_out) return; // nothing changes in this column
    int imnz = 0;
    for (int q = 0; q < BL; ++q) imnz |= sflag[q][tx];
    const double c1 = sc1[tx], c2 = sc2[tx];
    const double *t1 = kind == 3 ? a.tab + 3 * n : (kind == 2 ? a.tab + 2 * n : a.tab);
    const double *t2 = a.tab + n;
    // C: type-1 rows, then the correction (c1 tau1 + c2 tau2 rounded as
    // written, no fused multiply-add, then added)
    for (int i = ty; i < n; i += BL) {
        const size_t o = (size_t)i * s + k;
        double r = re[o], m = im[o];
        bool w = false;
        if ((i == 0 && dir_in) || (i == n - 1 && dir_out)) {
            r = i == 0 ? vi : vo;
            m = 0.0;
            w = true;
        }
        if (kind) {
            const double corr = kind == 1 ? __dadd_rn(__dmul_rn(c1, t1[i]), __dmul_rn(c2, t2[i])) : __dmul_rn(c1, t1[i]);
            r += corr;
            if (imnz) m += corr;
            w = true;
        }
        if (w) {
            re[o] = r;
            im[o] = m;
        }
    }
}

// nr <= 16 * ML: every thread loads its ML levels of both planes first (all
// loads in flight together), the D1 boundary windows of the real profile go
// through LDS to the mode's coefficient thread, and the update is applied to
// the registers.  Same arithmetic as scalar_bc_kernel.
template <int ML>
__global__ void __launch_bounds__(256) scalar_bc_reg_kernel(BcArgs a, double *__restrict__ re, double *__restrict__ im)
{
    constexpr int MAXW = 16; // 2 (bw + 1), bw <= 7
    __shared__ double sc1[BM], sc2[BM], swin[BM][MAXW];
    __shared__ int skind[BM], simnz[BM];
    __shared__ int sflag[BL][BM];
    const int tx = threadIdx.x % BM, ty = threadIdx.x / BM;
    const int k = blockIdx.x * BM + tx;
    const bool valid = k < a.nlm;
    const size_t s = (size_t)a.nlm;
    const int n = a.nr, bw = a.bw;
    int ti = 0, to = 0;
    double vi = 0.0, vo = 0.0;
    if (valid) {
        ti = a.types[2 * k];
        to = a.types[2 * k + 1];
        vi = a.vals[2 * k];
        vo = a.vals[2 * k + 1];
    }
    const bool dir_in = a.dirichlet && a.inner_row && ti == 1, dir_out = a.dirichlet && to == 1;
    const bool ai = ti == 2, ao = to == 2, flux = valid && (ai || ao);
    const bool touched = valid && (flux || dir_in || dir_out);
    // clamped (valid) addresses and no branches around the loads, so that
    // all of them are in flight before the first use
    const int kc = valid ? k : a.nlm - 1;
    double rv[ML], iv[ML];
#pragma unroll
    for (int j = 0; j < ML; ++j) {
        const int i = ty + BL * j, ic = i < n ? i : n - 1;
        rv[j] = re[(size_t)ic * s + kc];
        iv[j] = im[(size_t)ic * s + kc];
    }
#pragma unroll
    for (int j = 0; j < ML; ++j) {
        const bool ok = touched && ty + BL * j < n;
        rv[j] = ok ? rv[j] : 0.0;
        iv[j] = ok ? iv[j] : 0.0;
    }
    int f = 0;
#pragma unroll
    for (int j = 0; j < ML; ++j) {
        const int i = ty + BL * j;
        if ((i == 0 && dir_in) || (i == n - 1 && dir_out)) {
            rv[j] = i == 0 ? vi : vo;
            iv[j] = 0.0;
        }
        f |= (i < n && fabs(iv[j]) > 1e-12) ? 1 : 0;
        if (i <= bw) swin[tx][i] = rv[j];
        if (i < n && i >= n - 1 - bw) swin[tx][bw + 1 + i - (n - 1 - bw)] = rv[j];
    }
    sflag[ty][tx] = f;
    __syncthreads();
    if (ty == 0) {
        int kind = 0, imnz = 0;
        double c1 = 0.0, c2 = 0.0;
        if (flux) {
            double ci = 0.0, co = 0.0;
            for (int j = 0; j <= bw; ++j) ci += a.d1[j] * swin[tx][j];
            for (int j = 0; j <= bw; ++j) co += a.d1[bw + 1 + j] * swin[tx][bw + 1 + j];
            const double fi = ai ? vi : 0.0, fo = ao ? vo : 0.0;
            if (ai && ao) {
                const double ei = fi - ci, eo = fo - co;
                const double det = a.dt1i * a.dt2o - a.dt1o * a.dt2i;
                if (fabs(det) > 1e-12) {
                    c1 = (ei * a.dt2o - eo * a.dt2i) / det;
                    c2 = (eo * a.dt1i - ei * a.dt1o) / det;
                }
                kind = 1;
            } else if (ai) {
                c1 = (fi - ci) / a.dtau_in;
                kind = 2;
            } else {
                c1 = (fo - co) / a.dtau_out;
                kind = 3;
            }
            for (int q = 0; q < BL; ++q) imnz |= sflag[q][tx];
        }
        sc1[tx] = c1;
        sc2[tx] = c2;
        skind[tx] = kind;
        simnz[tx] = imnz;
    }
    __syncthreads();
    if (!touched) return;
    const int kind = skind[tx], imnz = simnz[tx];
    const double c1 = sc1[tx], c2 = sc2[tx];
    const double *t1 = kind == 3 ? a.tab + 3 * n : (kind == 2 ? a.tab + 2 * n : a.tab);
    const double *t2 = a.tab + n;
#pragma unroll
    for (int j = 0; j < ML; ++j) {
        const int i = ty + BL * j;
        if (i >= n) break;
        double r = rv[j], m = iv[j];
        if (kind) {
            const double corr = kind == 1 ? __dadd_rn(__dmul_rn(c1, t1[i]), __dmul_rn(c2, t2[i])) : __dmul_rn(c1, t1[i]);
            r += corr;
            if (imnz) m += corr;
        }
        re[(size_t)i * s + k] = r;
        im[(size_t)i * s + k] = m;
    }
}

double cheb(int n, double x) { return std::cos(n * std::acos(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x))); }

double cheb_deriv(int n, double x, double scale)
{
    double d;
    if (std::fabs(std::fabs(x) - 1.0) < 1e-12) d = (double)n * n * std::pow(x < 0 ? -1.0 : 1.0, n + 1);
    else {
        const double th = std::acos(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x));
        d = n * std::sin(n * th) / std::sin(th);
    }
    return d * scale;
}
} // namespace

extern "C" int gd_scalar_bc_destroy(gd_scalar_bc *b);

struct gd_scalar_bc {
    gd_ctx *ctx = nullptr;
    int lmax = 0, mmax = 0, nr = 0, bw = 0, nlm = 0;
    BcArgs args{};
    DevBuf tab, d1, types, vals, scratch;
    std::vector<double> tab_host; // tau1, tau2, tau_in, tau_out
    double rinner = 0.0;
    bool set = false;
};

extern "C" int gd_scalar_bc_create(gd_ctx *ctx, int lmax, int mmax, int nr, int bw, const double *r, int fd_mode,
                                   gd_scalar_bc **out)
{
    GD_CHECK(ctx && r && out, GD_ERR_ARG, "gd_scalar_bc_create: null argument");
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax && nr >= 3 && bw >= 1 && 2 * bw + 1 <= nr, GD_ERR_ARG,
             "gd_scalar_bc_create: lmax=%d mmax=%d nr=%d bw=%d", lmax, mmax, nr, bw);
    GD_HIP(hipSetDevice(ctx->device));
    gd_scalar_bc *b = new gd_scalar_bc();
    b->ctx = ctx;
    b->lmax = lmax;
    b->mmax = mmax;
    b->nr = nr;
    b->bw = bw;
    for (int l = 0; l <= lmax; ++l) b->nlm += (l < mmax ? l : mmax) + 1;
    const int W = 2 * bw + 1;
    std::vector<double> band((size_t)nr * W);
    int rc = gd_derivative_matrix(1, nr, bw, r, fd_mode, band.data());
    if (rc != GD_OK) {
        delete b;
        return rc;
    }
    // dense element (i, j) of the band: data[j][bw + i - j]
    auto D1 = [&](int i, int j) { return (j < i - bw || j > i + bw) ? 0.0 : band[(size_t)j * W + bw + i - j]; };
    const double *rr = r + (size_t)nr * 3;
    const double ri = rr[0], ro = rr[nr - 1], scale = 2.0 / (ro - ri);
    b->rinner = ri;
    std::vector<double> &t = b->tab_host;
    t.assign((size_t)4 * nr, 0.0);
    for (int i = 0; i < nr; ++i) {
        const double x = 2.0 * (rr[i] - ri) / (ro - ri) - 1.0;
        t[i] = cheb(nr - 1, x);
        t[nr + i] = cheb(nr, x);
    }
    BcArgs &a = b->args;
    a.dt1i = cheb_deriv(nr - 1, 2.0 * (ri - ri) / (ro - ri) - 1.0, scale);
    a.dt1o = cheb_deriv(nr - 1, 2.0 * (ro - ri) / (ro - ri) - 1.0, scale);
    a.dt2i = cheb_deriv(nr, 2.0 * (ri - ri) / (ro - ri) - 1.0, scale);
    a.dt2o = cheb_deriv(nr, 2.0 * (ro - ri) / (ro - ri) - 1.0, scale);
    const double al = std::fabs(a.dt2o) > 1e-12 ? a.dt1o / a.dt2o : 0.0;
    const double be = std::fabs(a.dt2i) > 1e-12 ? a.dt1i / a.dt2i : 0.0;
    for (int i = 0; i < nr; ++i) {
        t[2 * nr + i] = t[i] - al * t[nr + i];
        t[3 * nr + i] = t[i] - be * t[nr + i];
    }
    // derivative of each single-boundary tau function at its boundary through D1
    double din = 0.0, dout = 0.0;
    for (int j = 0; j < nr; ++j) {
        din += D1(0, j) * t[2 * nr + j];
        dout += D1(nr - 1, j) * t[3 * nr + j];
    }
    a.dtau_in = din;
    a.dtau_out = dout;
    std::vector<double> rows((size_t)2 * (bw + 1));
    for (int j = 0; j <= bw; ++j) {
        rows[j] = D1(0, j);
        rows[bw + 1 + j] = D1(nr - 1, nr - 1 - bw + j);
    }
    rc = b->tab.ensure(sizeof(double) * t.size());
    if (rc == GD_OK) rc = b->d1.ensure(sizeof(double) * rows.size());
    if (rc == GD_OK) rc = b->types.ensure(sizeof(int) * 2 * b->nlm);
    if (rc == GD_OK) rc = b->vals.ensure(sizeof(double) * 2 * b->nlm);
    if (rc == GD_OK && hipMemcpy(b->tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = GD_ERR_HIP;
    if (rc == GD_OK && hipMemcpy(b->d1.p, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = GD_ERR_HIP;
    if (rc != GD_OK) {
        gd_scalar_bc_destroy(b);
        return rc;
    }
    a.tab = b->tab.as<double>();
    a.d1 = b->d1.as<double>();
    a.types = b->types.as<int>();
    a.vals = b->vals.as<double>();
    a.nlm = b->nlm;
    a.nr = nr;
    a.bw = bw;
    a.inner_row = ri > 0.0 ? 1 : 0;
    *out = b;
    return GD_OK;
}

// column k of the planes holds l-major mode lm_of_pos[k] (identity when null)
int gd_scalar_bc_set_perm(gd_scalar_bc *b, const int *bc_inner, const int *bc_outer, const double *values,
                          const int *lm_of_pos)
{
    GD_CHECK(b && bc_inner && bc_outer && values, GD_ERR_ARG, "gd_scalar_bc_set: null argument");
    GD_HIP(hipSetDevice(b->ctx->device));
    std::vector<int> t((size_t)2 * b->nlm);
    std::vector<double> v((size_t)2 * b->nlm);
    for (int k = 0; k < b->nlm; ++k) {
        const int lm = lm_of_pos ? lm_of_pos[k] : k;
        t[2 * k] = bc_inner[lm];
        t[2 * k + 1] = bc_outer[lm];
        v[2 * k] = values[2 * lm];
        v[2 * k + 1] = values[2 * lm + 1];
    }
    // the host arrays are read before the call returns
    GD_HIP(hipMemcpy(b->types.p, t.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice));
    GD_HIP(hipMemcpy(b->vals.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    b->set = true;
    return GD_OK;
}

extern "C" int gd_scalar_bc_set(gd_scalar_bc *b, const int *bc_inner, const int *bc_outer, const double *values)
{
    return gd_scalar_bc_set_perm(b, bc_inner, bc_outer, values, nullptr);
}

// device planes (nr, nlm), in place on stream st; no synchronisation
int gd_scalar_bc_launch(gd_scalar_bc *b, hipStream_t st, double *re, double *im, int dirichlet)
{
    GD_CHECK(b->set, GD_ERR_ARG, "gd_scalar_bc_apply: no boundary conditions set (gd_scalar_bc_set)");
    BcArgs a = b->args;
    a.dirichlet = dirichlet ? 1 : 0;
    const dim3 grid((b->nlm + BM - 1) / BM), block(BM * BL);
    if (b->nr <= BL * 5 && b->bw <= 7) hipLaunchKernelGGL(scalar_bc_reg_kernel<5>, grid, block, 0, st, a, re, im);
    else if (b->nr <= BL * 8 && b->bw <= 7) hipLaunchKernelGGL(scalar_bc_reg_kernel<8>, grid, block, 0, st, a, re, im);
    else hipLaunchKernelGGL(scalar_bc_kernel, grid, block, 0, st, a, re, im);
    return gd_check_launch("scalar_bc_kernel");
}

extern "C" int gd_scalar_bc_apply(gd_scalar_bc *b, double *re, double *im, int dirichlet, int flags)
{
    GD_CHECK(b && re && im, GD_ERR_ARG, "gd_scalar_bc_apply: null argument");
    GD_CHECK(b->set, GD_ERR_ARG, "gd_scalar_bc_apply: no boundary conditions set (gd_scalar_bc_set)");
    GD_HIP(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const size_t n = (size_t)b->nlm * b->nr;
    double *dre = re, *dim = im;
    if (flags & GD_HOST) {
        GD_TRY(b->scratch.ensure(sizeof(double) * n * 2));
        dre = b->scratch.as<double>();
        dim = dre + n;
        GD_HIP(hipMemcpyAsync(dre, re, sizeof(double) * n, hipMemcpyHostToDevice, st));
        GD_HIP(hipMemcpyAsync(dim, im, sizeof(double) * n, hipMemcpyHostToDevice, st));
    }
    GD_TRY(gd_scalar_bc_launch(b, st, dre, dim, dirichlet));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(b->ctx, re, dre, sizeof(double) * n));
        GD_TRY(gd_from_device(b->ctx, im, dim, sizeof(double) * n));
    }
    return GD_OK;
}

extern "C" int gd_scalar_bc_tau(const gd_scalar_bc *b, double *tab, double *scalars)
{
    GD_CHECK(b && tab && scalars, GD_ERR_ARG, "gd_scalar_bc_tau: null argument");
    std::copy(b->tab_host.begin(), b->tab_host.end(), tab);
    const BcArgs &a = b->args;
    const double s[6] = {a.dt1i, a.dt1o, a.dt2i, a.dt2o, a.dtau_in, a.dtau_out};
    std::copy(s, s + 6, scalars);
    return GD_OK;
}

extern "C" int gd_scalar_bc_destroy(gd_scalar_bc *b)
{
    if (!b) return GD_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (DevBuf *d : {&b->tab, &b->d1, &b->types, &b->vals, &b->scratch}) d->release();
    delete b;
    return GD_OK;
}
