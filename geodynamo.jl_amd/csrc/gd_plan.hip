// gd_plan.hip -- C-ABI entry points for the transforms (SHT plans) and the
// banded / CNAB2 radial operators.
#include <cmath>

#include "gd_plan.hpp"
#include "gd_radial.hpp"

// ---------------------------------------------------------------- Gauss grid (host FP64)
// Newton iteration on P_nlat from cos(pi (i+3/4)/(nlat+1/2)); nodes north first.
static void gauss_nodes(int nlat, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(nlat, 0.0);
    w.assign(nlat, 0.0);
    for (int i = 0; i < (nlat + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (nlat + 0.5)), pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= nlat; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            pp = nlat * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) {
                p0 = 1.0;
                p1 = z;
                for (int k = 2; k <= nlat; ++k) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
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
}

template <typename T> static int upload(gd_ctx *ctx, DevBuf &b, const std::vector<T> &v)
{
    GD_TRY(b.ensure(sizeof(T) * v.size()));
    GD_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, ctx->stream));
    return GD_OK;
}

int gd_sht_plan_build(gd_ctx *ctx, int lmax, int mmax, int nlat, int nlon, int prec, gd_sht_plan *p)
{
    GD_CHECK(lmax >= 1 && mmax >= 0 && mmax <= lmax, GD_ERR_ARG, "sht: lmax=%d mmax=%d", lmax, mmax);
    GD_CHECK(nlat % 2 == 0 && nlat >= 2 && nlat > lmax, GD_ERR_ARG, "sht: nlat=%d must be even and > lmax", nlat);
    GD_CHECK(nlon > 2 * mmax, GD_ERR_ARG, "sht: nlon=%d must exceed 2*mmax", nlon);
    GD_CHECK(prec == GD_PREC_FP64 || prec == GD_PREC_FP32 || prec == GD_PREC_BF16X3, GD_ERR_ARG,
             "sht: precision %d not available (FP64=0, FP32=1, BF16X3=2)", prec);
    p->ctx = ctx;
    p->lmax = lmax;
    p->mmax = mmax;
    p->nlat = nlat;
    p->nlon = nlon;
    p->nlat2 = nlat / 2;
    p->prec = prec;
    p->maps.build(lmax, mmax);
    p->nlm = p->maps.nlm;
    GD_TRY(p->fft.build(nlon));
    gauss_nodes(nlat, p->x, p->w);
    p->theta.resize(nlat);
    for (int i = 0; i < nlat; ++i) p->theta[i] = std::acos(p->x[i]);
    GD_HIP(hipSetDevice(ctx->device));
    GD_TRY(upload(ctx, p->d_moff, p->maps.moff));
    GD_TRY(upload(ctx, p->d_ne, p->maps.ne));
    GD_TRY(upload(ctx, p->d_l_of_idx, p->maps.l_of_idx));
    GD_TRY(upload(ctx, p->d_lmaj_of_idx, p->maps.lmaj_of_idx));
    GD_TRY(upload(ctx, p->d_idx_of_lmaj, p->maps.idx_of_lmaj));
    GD_TRY(upload(ctx, p->d_m_of_idx, p->maps.m_of_idx));
    std::vector<double> xn(p->x.begin(), p->x.begin() + p->nlat2);
    GD_TRY(upload(ctx, p->d_x, xn));
    const size_t ntab = (size_t)3 * p->nlm * p->nlat2;
    GD_TRY(p->tab64.ensure(sizeof(double) * ntab));
    float *t32 = nullptr;
    if (prec != GD_PREC_FP64) {
        GD_TRY(p->tab32.ensure(sizeof(float) * ntab));
        t32 = p->tab32.as<float>();
    }
    GD_TRY(gd_launch_plm(ctx->stream, p->d_x.as<double>(), p->d_moff.as<int>(), lmax, mmax, p->nlat2,
                         p->tab64.as<double>(), t32, p->nlm));
    // twiddles, trig tables, quadrature weights
    std::vector<double> tw(2 * nlon), s(nlat), c(nlat), wq(p->nlat2);
    for (int k = 0; k < nlon; ++k) {
        double a = 2.0 * M_PI * (double)k / (double)nlon;
        tw[2 * k] = std::cos(a);
        tw[2 * k + 1] = -std::sin(a);
    }
    for (int i = 0; i < nlat; ++i) {
        s[i] = std::sin(p->theta[i]);
        c[i] = std::cos(p->theta[i]);
    }
    for (int i = 0; i < p->nlat2; ++i) wq[i] = 2.0 * M_PI * p->w[i] / nlon;
    GD_TRY(upload(ctx, p->wq64, wq));
    if (prec == GD_PREC_FP64) {
        GD_TRY(upload(ctx, p->wone, std::vector<double>(p->nlat2, 1.0)));
        GD_TRY(upload(ctx, p->tw, tw));
        GD_TRY(upload(ctx, p->sinth, s));
        GD_TRY(upload(ctx, p->costh, c));
        GD_TRY(upload(ctx, p->wq, wq));
    } else {
        GD_TRY(upload(ctx, p->wone, std::vector<float>(p->nlat2, 1.0f)));
        GD_TRY(upload(ctx, p->tw, std::vector<float>(tw.begin(), tw.end())));
        GD_TRY(upload(ctx, p->sinth, std::vector<float>(s.begin(), s.end())));
        GD_TRY(upload(ctx, p->costh, std::vector<float>(c.begin(), c.end())));
        GD_TRY(upload(ctx, p->wq, std::vector<float>(wq.begin(), wq.end())));
    }
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

void gd_sht_plan_release(gd_sht_plan *p)
{
    DevBuf *bufs[] = {&p->d_moff, &p->d_ne, &p->d_l_of_idx, &p->d_lmaj_of_idx, &p->d_idx_of_lmaj, &p->d_x,
                      &p->d_m_of_idx, &p->wone, &p->wq64,
                      &p->tab64, &p->tab32, &p->tw, &p->sinth, &p->costh, &p->wq, &p->syn, &p->fb,
                      &p->gb, &p->outm, &p->h_in0, &p->h_in1, &p->h_in2, &p->h_in3, &p->h_out0,
                      &p->h_out1, &p->h_out2, &p->h_out3};
    for (DevBuf *b : bufs) b->release();
}

const void *gd_sht_tab(const gd_sht_plan *p, int t)
{
    const size_t ntab = (size_t)p->nlm * p->nlat2;
    if (p->prec == GD_PREC_FP64) return p->tab64.as<double>() + t * ntab;
    return p->tab32.as<float>() + t * ntab;
}

void gd_sht_fill_ring(const gd_sht_plan *p, RingArgs &a, int nr)
{
    a.tw = p->tw.p;
    a.sinth = p->sinth.p;
    a.costh = p->costh.p;
    a.wq = p->wq.p;
    a.nlat = p->nlat;
    a.nlat2 = p->nlat2;
    a.nlon = p->nlon;
    a.nr = nr;
    a.mmax = p->mmax;
    a.nstages = p->fft.nstages;
    for (int s = 0; s < 12; ++s) a.radix[s] = p->fft.radix[s];
}

extern "C" int gd_sht_plan_create(gd_ctx *ctx, int lmax, int mmax, int nlat, int nlon, int prec, gd_sht_plan **out)
{
    GD_CHECK(ctx && out, GD_ERR_ARG, "gd_sht_plan_create: null argument");
    gd_sht_plan *p = new gd_sht_plan();
    int rc = gd_sht_plan_build(ctx, lmax, mmax, nlat, nlon, prec, p);
    if (rc != GD_OK) {
        gd_sht_plan_release(p);
        delete p;
        return rc;
    }
    *out = p;
    return GD_OK;
}

extern "C" int gd_sht_plan_destroy(gd_sht_plan *p)
{
    if (!p) return GD_OK;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    gd_sht_plan_release(p);
    delete p;
    return GD_OK;
}

extern "C" int gd_sht_nlm(const gd_sht_plan *p) { return p ? p->nlm : -1; }

extern "C" int gd_sht_grid(gd_sht_plan *p, double *x, double *w, double *theta)
{
    GD_CHECK(p, GD_ERR_ARG, "null plan");
    for (int i = 0; i < p->nlat; ++i) {
        if (x) x[i] = p->x[i];
        if (w) w[i] = p->w[i];
        if (theta) theta[i] = p->theta[i];
    }
    return GD_OK;
}

// FP64 tables in l-major (nlm x nlat), southern nodes by parity
extern "C" int gd_sht_tables(gd_sht_plan *p, double *P, double *dP)
{
    GD_CHECK(p && P && dP, GD_ERR_ARG, "gd_sht_tables: null argument");
    const size_t ntab = (size_t)p->nlm * p->nlat2;
    std::vector<double> h(2 * ntab);
    GD_HIP(hipSetDevice(p->ctx->device));
    GD_HIP(hipMemcpyAsync(h.data(), p->tab64.p, sizeof(double) * 2 * ntab, hipMemcpyDeviceToHost, p->ctx->stream));
    GD_HIP(hipStreamSynchronize(p->ctx->stream));
    for (int idx = 0; idx < p->nlm; ++idx) {
        const int lm = p->maps.lmaj_of_idx[idx], l = p->maps.l_of_idx[idx], m = p->maps.m_of_idx[idx];
        const double sg = ((l + m) & 1) ? -1.0 : 1.0;
        for (int i = 0; i < p->nlat2; ++i) {
            double pv = h[(size_t)idx * p->nlat2 + i], dv = h[ntab + (size_t)idx * p->nlat2 + i];
            P[(size_t)lm * p->nlat + i] = pv;
            P[(size_t)lm * p->nlat + p->nlat - 1 - i] = sg * pv;
            dP[(size_t)lm * p->nlat + i] = dv;
            dP[(size_t)lm * p->nlat + p->nlat - 1 - i] = -sg * dv;
        }
    }
    return GD_OK;
}

// ---------------------------------------------------------------- transform pipelines
// ---------------------------------------------------------------- transposes of the transforms
// On the Gauss grid with real fields f = sum c_m Re(a_lm Y_lm) (c_0 = 1, c_m = 2)
// the analysis is A = C^-1 S^T W (scalar) and A_v = L^-1 C^-1 S_v^T W (vector,
// L = l(l+1)), W the quadrature weight 2 pi w_i / nlon of each grid point, S
// the synthesis.  The transposes therefore run the forward pipelines with
// the weights moved:
//   S^T = C A W^-1:   the analysis with unit weights, then x c_m (and the
//                     vector one without its 1/(l(l+1)) factor);
//   A^T = W S C^-1:   the synthesis of the columns / c_m (/ l(l+1) for the
//                     vector one, 0 at l = 0), then x W per grid point.
namespace {
template <typename T>
__global__ void colscale_kernel(T *__restrict__ syn, const int *__restrict__ l_of_idx, const int *__restrict__ m_of_idx,
                                int nlm, int ll)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
    if (idx >= nlm) return;
    const int l = l_of_idx[idx];
    double f = m_of_idx[idx] ? 0.5 : 1.0;
    if (ll) f = l ? f / ((double)l * (double)(l + 1)) : 0.0;
    syn[(size_t)col * nlm + idx] = (T)((double)syn[(size_t)col * nlm + idx] * f);
}

// grid values (nlat, nlon, nr), colatitude fastest, times their ring's weight
__global__ void wscale_kernel(double *__restrict__ phys, const double *__restrict__ w, int nlat, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = (int)(q % (size_t)nlat), i2 = i < nlat / 2 ? i : nlat - 1 - i;
    phys[q] *= w[i2];
}

// m-major planes [plane][idx] times c_m
__global__ void cm_kernel(double *__restrict__ planes, const int *__restrict__ m_of_idx, int nlm)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    if (m_of_idx[idx]) planes[(size_t)blockIdx.y * nlm + idx] *= 2.0;
}
} // namespace

template <typename T>
static int run_synth(gd_sht_plan *p, bool vec, const double *a_re, const double *a_im, const double *b_re,
                     const double *b_im, int nr, double *o0, double *o1, bool adj = false)
{
    hipStream_t st = p->ctx->stream;
    const int nlm = p->nlm, nm = p->mmax + 1;
    const int cin = vec ? 4 : 2, nslot = vec ? 8 : 4;
    GD_TRY(p->syn.ensure(sizeof(T) * (size_t)nr * cin * nlm));
    GD_TRY(p->fb.ensure(sizeof(T) * (size_t)p->nlat2 * nr * nm * nslot));
    T *syn = p->syn.as<T>();
    if (vec) {
        // S := pol (b) in columns 0,1; T := tor (a) in columns 2,3
        GD_TRY(gd_launch_pack<T>(st, b_re, b_im, p->d_lmaj_of_idx.as<int>(), nlm, nr, cin, 0, syn));
        GD_TRY(gd_launch_pack<T>(st, a_re, a_im, p->d_lmaj_of_idx.as<int>(), nlm, nr, cin, 2, syn));
    } else {
        GD_TRY(gd_launch_pack<T>(st, a_re, a_im, p->d_lmaj_of_idx.as<int>(), nlm, nr, cin, 0, syn));
    }
    if (adj) {
        hipLaunchKernelGGL((colscale_kernel<T>), dim3((nlm + 255) / 256, nr * cin), dim3(256), 0, st, syn,
                           p->d_l_of_idx.as<int>(), p->d_m_of_idx.as<int>(), nlm, vec ? 1 : 0);
        GD_TRY(gd_check_launch("colscale_kernel"));
    }
    SynthArgs s{};
    s.mm = p->prec == GD_PREC_BF16X3;
    s.in = syn;
    s.cin = cin;
    s.nf = 1;
    for (int k = 0; k < 20; ++k) s.colmap[k] = k < cin ? k : -1;
    s.moff = p->d_moff.as<int>();
    s.ne = p->d_ne.as<int>();
    s.nlat2 = p->nlat2;
    s.nr = nr;
    s.nlm = nlm;
    s.mmax = p->mmax;
    s.out = p->fb.p;
    s.nslot = nslot;
    s.slot0 = 0;
    if (vec) {
        s.tab0 = gd_sht_tab(p, 1);
        s.tab1 = gd_sht_tab(p, 2);
        GD_TRY(gd_launch_synth_vec<T>(st, s));
    } else {
        s.tab0 = gd_sht_tab(p, 0);
        s.tab1 = nullptr;
        GD_TRY(gd_launch_synth_scal<T>(st, s));
    }
    RingArgs r{};
    gd_sht_fill_ring(p, r, nr);
    r.fb = p->fb.p;
    r.nslot_f = nslot;
    if (vec) {
        r.nfield = 2;
        r.fn_off[0] = 0;
        r.fs_off[0] = 4;
        r.fn_off[1] = 2;
        r.fs_off[1] = 6;
        r.phys_out[0] = o0;
        r.phys_out[1] = o1;
    } else {
        r.nfield = 1;
        r.fn_off[0] = 0;
        r.fs_off[0] = 2;
        r.phys_out[0] = o0;
    }
    GD_TRY(gd_launch_ring<T>(st, RING_SYNTH, r));
    if (adj) {
        const size_t n = (size_t)p->nlat * p->nlon * nr;
        for (double *o : {o0, o1}) {
            if (!o) continue;
            hipLaunchKernelGGL(wscale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, o,
                               p->wq64.as<double>(), p->nlat, n);
            GD_TRY(gd_check_launch("wscale_kernel"));
        }
    }
    return GD_OK;
}

template <typename T>
static int run_analysis(gd_sht_plan *p, bool vec, const double *f0, const double *f1, int nr, double *o0_re,
                        double *o0_im, double *o1_re, double *o1_im, bool adj = false)
{
    hipStream_t st = p->ctx->stream;
    const int nlm = p->nlm, nm = p->mmax + 1;
    const int nslot = vec ? 8 : 4;
    GD_TRY(p->gb.ensure(sizeof(T) * (size_t)nm * p->nlat2 * nr * nslot));
    const size_t fs = (size_t)nr * nlm;
    GD_TRY(p->outm.ensure(sizeof(double) * fs * 2 * (vec ? 2 : 1)));
    RingArgs r{};
    gd_sht_fill_ring(p, r, nr);
    r.gb = p->gb.p;
    r.nslot_g = nslot;
    r.nfield = vec ? 2 : 1;
    r.phys_in[0] = f0;
    r.phys_in[1] = f1;
    if (adj) r.wq = p->wone.p;
    GD_TRY(gd_launch_ring<T>(st, RING_ANALYSIS, r));
    AnaArgs a{};
    a.mm = p->prec == GD_PREC_BF16X3;
    a.gb = p->gb.p;
    a.nslot_g = nslot;
    a.gslot0 = 0;
    a.nf = 1;
    a.moff = p->d_moff.as<int>();
    a.ne = p->d_ne.as<int>();
    a.l_of_idx = p->d_l_of_idx.as<int>();
    a.nlat2 = p->nlat2;
    a.nr = nr;
    a.nlm = nlm;
    a.mmax = p->mmax;
    double *om = p->outm.as<double>();
    const int maxgroup = (p->lmax + 2) / 2;
    if (vec) {
        a.tab0 = gd_sht_tab(p, 1);
        a.tab1 = gd_sht_tab(p, 2);
        a.out[0] = om;          // S -> pol
        a.out[1] = om + 2 * fs; // T -> tor
        a.scale[0] = a.scale[1] = adj ? 1 : 0;
        GD_TRY(gd_launch_ana<T>(st, a, true, maxgroup));
        if (adj) {
            hipLaunchKernelGGL(cm_kernel, dim3((nlm + 255) / 256, 4 * nr), dim3(256), 0, st, om,
                               p->d_m_of_idx.as<int>(), nlm);
            GD_TRY(gd_check_launch("cm_kernel"));
        }
        // o0 = tor, o1 = pol
        GD_TRY(gd_launch_unpack(st, om + 2 * fs, p->d_idx_of_lmaj.as<int>(), nlm, nr, o0_re));
        GD_TRY(gd_launch_unpack(st, om + 3 * fs, p->d_idx_of_lmaj.as<int>(), nlm, nr, o0_im));
        GD_TRY(gd_launch_unpack(st, om, p->d_idx_of_lmaj.as<int>(), nlm, nr, o1_re));
        GD_TRY(gd_launch_unpack(st, om + fs, p->d_idx_of_lmaj.as<int>(), nlm, nr, o1_im));
    } else {
        a.tab0 = gd_sht_tab(p, 0);
        a.tab1 = nullptr;
        a.out[0] = om;
        a.scale[0] = 1;
        GD_TRY(gd_launch_ana<T>(st, a, false, maxgroup));
        if (adj) {
            hipLaunchKernelGGL(cm_kernel, dim3((nlm + 255) / 256, 2 * nr), dim3(256), 0, st, om,
                               p->d_m_of_idx.as<int>(), nlm);
            GD_TRY(gd_check_launch("cm_kernel"));
        }
        GD_TRY(gd_launch_unpack(st, om, p->d_idx_of_lmaj.as<int>(), nlm, nr, o0_re));
        GD_TRY(gd_launch_unpack(st, om + fs, p->d_idx_of_lmaj.as<int>(), nlm, nr, o0_im));
    }
    return GD_OK;
}

// stage host arrays (GD_HOST) to device buffers
static const double *stage_in(gd_sht_plan *p, DevBuf &b, const double *h, size_t n, int flags, int *rc)
{
    if (!(flags & GD_HOST)) return h;
    *rc = gd_to_device(p->ctx, b, h, sizeof(double) * n);
    return b.as<double>();
}
static double *stage_out(DevBuf &b, double *h, size_t n, int flags, int *rc)
{
    if (!(flags & GD_HOST)) return h;
    *rc = b.ensure(sizeof(double) * n);
    return b.as<double>();
}

extern "C" int gd_sht_scalar_synth(gd_sht_plan *p, const double *re, const double *im, int nr, double *phys,
                                   int flags)
{
    GD_CHECK(p && re && im && phys && nr >= 1, GD_ERR_ARG, "gd_sht_scalar_synth: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *dre = stage_in(p, p->h_in0, re, ns, flags, &rc);
    GD_TRY(rc);
    const double *dim = stage_in(p, p->h_in1, im, ns, flags, &rc);
    GD_TRY(rc);
    double *dph = stage_out(p->h_out0, phys, np, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_synth<double>(p, false, dre, dim, nullptr, nullptr, nr, dph, nullptr));
    else GD_TRY(run_synth<float>(p, false, dre, dim, nullptr, nullptr, nr, dph, nullptr));
    if (flags & GD_HOST) GD_TRY(gd_from_device(p->ctx, phys, dph, sizeof(double) * np));
    return GD_OK;
}

extern "C" int gd_sht_vector_synth(gd_sht_plan *p, const double *tor_re, const double *tor_im, const double *pol_re,
                                   const double *pol_im, int nr, double *vt, double *vp, int flags)
{
    GD_CHECK(p && tor_re && tor_im && pol_re && pol_im && vt && vp && nr >= 1, GD_ERR_ARG,
             "gd_sht_vector_synth: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *a = stage_in(p, p->h_in0, tor_re, ns, flags, &rc);
    GD_TRY(rc);
    const double *b = stage_in(p, p->h_in1, tor_im, ns, flags, &rc);
    GD_TRY(rc);
    const double *c = stage_in(p, p->h_in2, pol_re, ns, flags, &rc);
    GD_TRY(rc);
    const double *d = stage_in(p, p->h_in3, pol_im, ns, flags, &rc);
    GD_TRY(rc);
    double *o0 = stage_out(p->h_out0, vt, np, flags, &rc);
    GD_TRY(rc);
    double *o1 = stage_out(p->h_out1, vp, np, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_synth<double>(p, true, a, b, c, d, nr, o0, o1));
    else GD_TRY(run_synth<float>(p, true, a, b, c, d, nr, o0, o1));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, vt, o0, sizeof(double) * np));
        GD_TRY(gd_from_device(p->ctx, vp, o1, sizeof(double) * np));
    }
    return GD_OK;
}

extern "C" int gd_sht_scalar_analysis(gd_sht_plan *p, const double *phys, int nr, double *re, double *im, int flags)
{
    GD_CHECK(p && phys && re && im && nr >= 1, GD_ERR_ARG, "gd_sht_scalar_analysis: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *f = stage_in(p, p->h_in0, phys, np, flags, &rc);
    GD_TRY(rc);
    double *ore = stage_out(p->h_out0, re, ns, flags, &rc);
    GD_TRY(rc);
    double *oim = stage_out(p->h_out1, im, ns, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_analysis<double>(p, false, f, nullptr, nr, ore, oim, nullptr, nullptr));
    else GD_TRY(run_analysis<float>(p, false, f, nullptr, nr, ore, oim, nullptr, nullptr));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, re, ore, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, im, oim, sizeof(double) * ns));
    }
    return GD_OK;
}

extern "C" int gd_sht_vector_analysis(gd_sht_plan *p, const double *vt, const double *vp, int nr, double *tor_re,
                                      double *tor_im, double *pol_re, double *pol_im, int flags)
{
    GD_CHECK(p && vt && vp && tor_re && tor_im && pol_re && pol_im && nr >= 1, GD_ERR_ARG,
             "gd_sht_vector_analysis: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *f0 = stage_in(p, p->h_in0, vt, np, flags, &rc);
    GD_TRY(rc);
    const double *f1 = stage_in(p, p->h_in1, vp, np, flags, &rc);
    GD_TRY(rc);
    double *a = stage_out(p->h_out0, tor_re, ns, flags, &rc);
    GD_TRY(rc);
    double *b = stage_out(p->h_out1, tor_im, ns, flags, &rc);
    GD_TRY(rc);
    double *c = stage_out(p->h_out2, pol_re, ns, flags, &rc);
    GD_TRY(rc);
    double *d = stage_out(p->h_out3, pol_im, ns, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_analysis<double>(p, true, f0, f1, nr, a, b, c, d));
    else GD_TRY(run_analysis<float>(p, true, f0, f1, nr, a, b, c, d));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, tor_re, a, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, tor_im, b, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, pol_re, c, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, pol_im, d, sizeof(double) * ns));
    }
    return GD_OK;
}

// ---- transposed transforms (the adjoint of the transform pair)
extern "C" int gd_sht_scalar_synth_adjoint(gd_sht_plan *p, const double *phys, int nr, double *re, double *im,
                                           int flags)
{
    GD_CHECK(p && phys && re && im && nr >= 1, GD_ERR_ARG, "gd_sht_scalar_synth_adjoint: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *f = stage_in(p, p->h_in0, phys, np, flags, &rc);
    GD_TRY(rc);
    double *ore = stage_out(p->h_out0, re, ns, flags, &rc);
    GD_TRY(rc);
    double *oim = stage_out(p->h_out1, im, ns, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_analysis<double>(p, false, f, nullptr, nr, ore, oim, nullptr, nullptr, true));
    else GD_TRY(run_analysis<float>(p, false, f, nullptr, nr, ore, oim, nullptr, nullptr, true));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, re, ore, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, im, oim, sizeof(double) * ns));
    }
    return GD_OK;
}

extern "C" int gd_sht_scalar_analysis_adjoint(gd_sht_plan *p, const double *re, const double *im, int nr,
                                              double *phys, int flags)
{
    GD_CHECK(p && re && im && phys && nr >= 1, GD_ERR_ARG, "gd_sht_scalar_analysis_adjoint: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *dre = stage_in(p, p->h_in0, re, ns, flags, &rc);
    GD_TRY(rc);
    const double *dim = stage_in(p, p->h_in1, im, ns, flags, &rc);
    GD_TRY(rc);
    double *dph = stage_out(p->h_out0, phys, np, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_synth<double>(p, false, dre, dim, nullptr, nullptr, nr, dph, nullptr, true));
    else GD_TRY(run_synth<float>(p, false, dre, dim, nullptr, nullptr, nr, dph, nullptr, true));
    if (flags & GD_HOST) GD_TRY(gd_from_device(p->ctx, phys, dph, sizeof(double) * np));
    return GD_OK;
}

extern "C" int gd_sht_vector_synth_adjoint(gd_sht_plan *p, const double *vt, const double *vp, int nr,
                                           double *tor_re, double *tor_im, double *pol_re, double *pol_im, int flags)
{
    GD_CHECK(p && vt && vp && tor_re && tor_im && pol_re && pol_im && nr >= 1, GD_ERR_ARG,
             "gd_sht_vector_synth_adjoint: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *f0 = stage_in(p, p->h_in0, vt, np, flags, &rc);
    GD_TRY(rc);
    const double *f1 = stage_in(p, p->h_in1, vp, np, flags, &rc);
    GD_TRY(rc);
    double *a = stage_out(p->h_out0, tor_re, ns, flags, &rc);
    GD_TRY(rc);
    double *b = stage_out(p->h_out1, tor_im, ns, flags, &rc);
    GD_TRY(rc);
    double *c = stage_out(p->h_out2, pol_re, ns, flags, &rc);
    GD_TRY(rc);
    double *d = stage_out(p->h_out3, pol_im, ns, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_analysis<double>(p, true, f0, f1, nr, a, b, c, d, true));
    else GD_TRY(run_analysis<float>(p, true, f0, f1, nr, a, b, c, d, true));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, tor_re, a, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, tor_im, b, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, pol_re, c, sizeof(double) * ns));
        GD_TRY(gd_from_device(p->ctx, pol_im, d, sizeof(double) * ns));
    }
    return GD_OK;
}

extern "C" int gd_sht_vector_analysis_adjoint(gd_sht_plan *p, const double *tor_re, const double *tor_im,
                                              const double *pol_re, const double *pol_im, int nr, double *vt,
                                              double *vp, int flags)
{
    GD_CHECK(p && tor_re && tor_im && pol_re && pol_im && vt && vp && nr >= 1, GD_ERR_ARG,
             "gd_sht_vector_analysis_adjoint: bad args");
    GD_HIP(hipSetDevice(p->ctx->device));
    const size_t ns = (size_t)p->nlm * nr, np = (size_t)p->nlat * p->nlon * nr;
    int rc = GD_OK;
    const double *a = stage_in(p, p->h_in0, tor_re, ns, flags, &rc);
    GD_TRY(rc);
    const double *b = stage_in(p, p->h_in1, tor_im, ns, flags, &rc);
    GD_TRY(rc);
    const double *c = stage_in(p, p->h_in2, pol_re, ns, flags, &rc);
    GD_TRY(rc);
    const double *d = stage_in(p, p->h_in3, pol_im, ns, flags, &rc);
    GD_TRY(rc);
    double *o0 = stage_out(p->h_out0, vt, np, flags, &rc);
    GD_TRY(rc);
    double *o1 = stage_out(p->h_out1, vp, np, flags, &rc);
    GD_TRY(rc);
    if (p->prec == GD_PREC_FP64) GD_TRY(run_synth<double>(p, true, a, b, c, d, nr, o0, o1, true));
    else GD_TRY(run_synth<float>(p, true, a, b, c, d, nr, o0, o1, true));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(p->ctx, vt, o0, sizeof(double) * np));
        GD_TRY(gd_from_device(p->ctx, vp, o1, sizeof(double) * np));
    }
    return GD_OK;
}

// ---------------------------------------------------------------- banded LU / CNAB2 matrices
static int read_status(gd_ctx *ctx, DevBuf &st, int *out)
{
    GD_HIP(hipMemcpyAsync(out, st.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

static int zero_status(gd_ctx *ctx, DevBuf &st)
{
    GD_TRY(st.ensure(sizeof(int)));
    GD_HIP(hipMemsetAsync(st.p, 0, sizeof(int), ctx->stream));
    return GD_OK;
}

extern "C" int gd_banded_factor(gd_ctx *ctx, int n, int bw, int nsys, const double *ab, double *lu, int flags)
{
    GD_CHECK(ctx && ab && lu && n >= 1 && bw >= 1 && nsys >= 1, GD_ERR_ARG, "gd_banded_factor: bad args");
    GD_HIP(hipSetDevice(ctx->device));
    const size_t nb = (size_t)(2 * bw + 1) * n * nsys;
    DevBuf din, dout, status;
    const double *a = ab;
    double *o = lu;
    if (flags & GD_HOST) {
        GD_TRY(gd_to_device(ctx, din, ab, sizeof(double) * nb));
        GD_TRY(dout.ensure(sizeof(double) * nb));
        a = din.as<double>();
        o = dout.as<double>();
    }
    int rc = zero_status(ctx, status);
    if (rc == GD_OK) rc = gd_launch_band_lu(ctx->stream, a, o, nullptr, n, bw, nsys, status.as<int>());
    int st = 0;
    if (rc == GD_OK) rc = read_status(ctx, status, &st);
    if (rc == GD_OK && (flags & GD_HOST)) rc = gd_from_device(ctx, lu, o, sizeof(double) * nb);
    din.release();
    dout.release();
    status.release();
    GD_TRY(rc);
    GD_CHECK(st == 0, GD_ERR_NUMERIC, "gd_banded_factor: zero or non-finite pivot");
    return GD_OK;
}

extern "C" int gd_banded_solve(gd_ctx *ctx, int n, int bw, int nsys, const double *lu, const int *sys_of_rhs,
                               int nrhs, double *x, int flags)
{
    GD_CHECK(ctx && lu && x && n >= 1 && bw >= 1 && nsys >= 1 && nrhs >= 0, GD_ERR_ARG,
             "gd_banded_solve: bad args");
    GD_HIP(hipSetDevice(ctx->device));
    const size_t nb = (size_t)(2 * bw + 1) * n * nsys, nx = (size_t)n * nrhs;
    DevBuf dlu, dmap, dx, status;
    const double *L = lu;
    const int *map = sys_of_rhs;
    double *X = x;
    int rc = GD_OK;
    if (flags & GD_HOST) {
        rc = gd_to_device(ctx, dlu, lu, sizeof(double) * nb);
        if (rc == GD_OK && sys_of_rhs) rc = gd_to_device(ctx, dmap, sys_of_rhs, sizeof(int) * nrhs);
        if (rc == GD_OK) rc = gd_to_device(ctx, dx, x, sizeof(double) * nx);
        L = dlu.as<double>();
        map = sys_of_rhs ? dmap.as<int>() : nullptr;
        X = dx.as<double>();
    }
    if (rc == GD_OK) rc = zero_status(ctx, status);
    if (rc == GD_OK && nrhs > 0) rc = gd_launch_band_solve(ctx->stream, L, map, n, bw, nsys, nrhs, X, status.as<int>());
    int st = 0;
    if (rc == GD_OK) rc = read_status(ctx, status, &st);
    if (rc == GD_OK && (flags & GD_HOST)) rc = gd_from_device(ctx, x, X, sizeof(double) * nx);
    dlu.release();
    dmap.release();
    dx.release();
    status.release();
    GD_TRY(rc);
    GD_CHECK(st == 0, GD_ERR_ARG, "gd_banded_solve: sys_of_rhs out of range");
    return GD_OK;
}

extern "C" int gd_cnab2_mats_create(gd_ctx *ctx, int lmax, int nr, int bw, const double *r, double nu, double dt,
                                    double theta, int fd_mode, int bc_rows, gd_cnab2_mats **out)
{
    GD_CHECK(ctx && r && out && lmax >= 0 && nr >= 2 && bw >= 1 && bw <= 7 && dt > 0.0, GD_ERR_ARG,
             "gd_cnab2_mats_create: bad args");
    GD_HIP(hipSetDevice(ctx->device));
    const int W = 2 * bw + 1, nl = lmax + 1;
    std::vector<double> lap((size_t)W * nr), rinv2(r + nr, r + 2 * nr);
    GD_TRY(gd_radial_laplacian(nr, bw, r, fd_mode, lap.data()));
    gd_cnab2_mats *M = new gd_cnab2_mats();
    M->ctx = ctx;
    M->lmax = lmax;
    M->nr = nr;
    M->bw = bw;
    M->nu = nu;
    M->dt = dt;
    M->theta = theta;
    M->bc_rows = bc_rows ? 1 : 0;
    const size_t nb = (size_t)W * nr * nl;
    DevBuf dlap, dr2;
    int rc = upload(ctx, dlap, lap);
    if (rc == GD_OK) rc = upload(ctx, dr2, rinv2);
    if (rc == GD_OK) rc = M->L.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = M->S.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = M->LU.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = zero_status(ctx, M->status);
    if (rc == GD_OK)
        rc = gd_cnab2_build(ctx->stream, dlap.as<double>(), dr2.as<double>(), nu, dt, theta, nl, nr, bw,
                            M->bc_rows, M->L.as<double>(), M->S.as<double>(), M->LU.as<double>(), nullptr, nullptr,
                            M->status.as<int>());
    int st = 0;
    if (rc == GD_OK) rc = read_status(ctx, M->status, &st);
    dlap.release();
    dr2.release();
    if (rc == GD_OK && st != 0) {
        gd_set_error("gd_cnab2_mats_create: zero or non-finite pivot in LU(S_l)");
        rc = GD_ERR_NUMERIC;
    }
    if (rc != GD_OK) {
        gd_cnab2_mats_destroy(M);
        return rc;
    }
    *out = M;
    return GD_OK;
}

extern "C" int gd_cnab2_mats_get(gd_cnab2_mats *M, double *L, double *S, double *LU)
{
    GD_CHECK(M, GD_ERR_ARG, "null mats");
    const size_t nb = (size_t)(2 * M->bw + 1) * M->nr * (M->lmax + 1);
    if (L) GD_TRY(gd_from_device(M->ctx, L, M->L.p, sizeof(double) * nb));
    if (S) GD_TRY(gd_from_device(M->ctx, S, M->S.p, sizeof(double) * nb));
    if (LU) GD_TRY(gd_from_device(M->ctx, LU, M->LU.p, sizeof(double) * nb));
    return GD_OK;
}

extern "C" int gd_cnab2_mats_destroy(gd_cnab2_mats *M)
{
    if (!M) return GD_OK;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    M->L.release();
    M->S.release();
    M->LU.release();
    M->status.release();
    for (DevBuf *b : {&M->h0, &M->h1, &M->h2, &M->h3, &M->h4, &M->h5, &M->h6, &M->h7}) b->release();
    delete M;
    return GD_OK;
}

static int cnab2_call(gd_cnab2_mats *M, int lmax, int mmax, int nr, int mode, const double *u_re, const double *u_im,
                      const double *nl_re, const double *nl_im, const double *np_re, const double *np_im,
                      double *o_re, double *o_im, int flags)
{
    GD_CHECK(M && lmax >= 0 && mmax >= 0 && mmax <= lmax && nr == M->nr, GD_ERR_ARG,
             "cnab2: bad args (nr=%d, matrices nr=%d)", nr, M->nr);
    gd_ctx *ctx = M->ctx;
    GD_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)gd_nlm_of(lmax, mmax) * nr;
    const double *a = u_re, *b = u_im, *c = nl_re, *d = nl_im, *e = np_re, *f = np_im;
    double *oa = o_re, *ob = o_im;
    if (flags & GD_HOST) {
        GD_TRY(gd_to_device(ctx, M->h0, u_re, sizeof(double) * n));
        GD_TRY(gd_to_device(ctx, M->h1, u_im, sizeof(double) * n));
        a = M->h0.as<double>();
        b = M->h1.as<double>();
        if (mode != 1) {
            GD_TRY(gd_to_device(ctx, M->h2, nl_re, sizeof(double) * n));
            GD_TRY(gd_to_device(ctx, M->h3, nl_im, sizeof(double) * n));
            GD_TRY(gd_to_device(ctx, M->h4, np_re, sizeof(double) * n));
            GD_TRY(gd_to_device(ctx, M->h5, np_im, sizeof(double) * n));
            c = M->h2.as<double>();
            d = M->h3.as<double>();
            e = M->h4.as<double>();
            f = M->h5.as<double>();
        }
        if (mode == 2) {
            oa = M->h0.as<double>();
            ob = M->h1.as<double>();
        } else {
            GD_TRY(M->h6.ensure(sizeof(double) * n));
            GD_TRY(M->h7.ensure(sizeof(double) * n));
            oa = M->h6.as<double>();
            ob = M->h7.as<double>();
            if (mode == 1) {
                // solve_implicit_step! leaves columns without a matrix untouched
                GD_HIP(hipMemcpyAsync(oa, o_re, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
                GD_HIP(hipMemcpyAsync(ob, o_im, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
            }
        }
    }
    GD_TRY(zero_status(ctx, M->status));
    GD_TRY(gd_launch_cnab2_lmajor(ctx->stream, M->L.as<double>(), M->LU.as<double>(), M->lmax + 1, lmax, mmax, nr,
                                  M->bw, M->dt, M->theta, mode, M->bc_rows, a, b, c, d, e, f, oa, ob,
                                  M->status.as<int>()));
    int st = 0;
    GD_TRY(read_status(ctx, M->status, &st));
    if (flags & GD_HOST) {
        double *ha = (mode == 2) ? const_cast<double *>(u_re) : o_re;
        double *hb = (mode == 2) ? const_cast<double *>(u_im) : o_im;
        GD_TRY(gd_from_device(ctx, ha, oa, sizeof(double) * n));
        GD_TRY(gd_from_device(ctx, hb, ob, sizeof(double) * n));
    }
    GD_CHECK(!(st & 4), GD_ERR_MISSING_L, "Missing implicit matrix for l > %d", M->lmax);
    return GD_OK;
}

extern "C" int gd_build_rhs_cnab2(gd_cnab2_mats *M, int lmax, int mmax, int nr, const double *u_re,
                                  const double *u_im, const double *nl_re, const double *nl_im, const double *np_re,
                                  const double *np_im, double *rhs_re, double *rhs_im, int flags)
{
    GD_CHECK(u_re && u_im && nl_re && nl_im && np_re && np_im && rhs_re && rhs_im, GD_ERR_ARG,
             "gd_build_rhs_cnab2: null array");
    return cnab2_call(M, lmax, mmax, nr, 0, u_re, u_im, nl_re, nl_im, np_re, np_im, rhs_re, rhs_im, flags);
}

extern "C" int gd_solve_implicit_step(gd_cnab2_mats *M, int lmax, int mmax, int nr, const double *rhs_re,
                                      const double *rhs_im, double *sol_re, double *sol_im, int flags)
{
    GD_CHECK(rhs_re && rhs_im && sol_re && sol_im, GD_ERR_ARG, "gd_solve_implicit_step: null array");
    return cnab2_call(M, lmax, mmax, nr, 1, rhs_re, rhs_im, nullptr, nullptr, nullptr, nullptr, sol_re, sol_im,
                      flags);
}

extern "C" int gd_cnab2_rhs_solve(gd_cnab2_mats *M, int lmax, int mmax, int nr, double *u_re, double *u_im,
                                  const double *nl_re, const double *nl_im, const double *np_re, const double *np_im,
                                  int flags)
{
    GD_CHECK(u_re && u_im && nl_re && nl_im && np_re && np_im, GD_ERR_ARG, "gd_cnab2_rhs_solve: null array");
    return cnab2_call(M, lmax, mmax, nr, 2, u_re, u_im, nl_re, nl_im, np_re, np_im, u_re, u_im, flags);
}

extern "C" int gd_timestep_error(gd_ctx *ctx, long n, const double *a_re, const double *a_im, const double *b_re,
                                 const double *b_im, double *out, int flags)
{
    GD_CHECK(ctx && a_re && a_im && b_re && b_im && out && n >= 0, GD_ERR_ARG, "gd_timestep_error: bad args");
    GD_HIP(hipSetDevice(ctx->device));
    DevBuf h[4], part, res;
    const double *p[4] = {a_re, a_im, b_re, b_im};
    int rc = GD_OK;
    if (flags & GD_HOST)
        for (int k = 0; k < 4 && rc == GD_OK; ++k) {
            rc = gd_to_device(ctx, h[k], p[k], sizeof(double) * n);
            p[k] = h[k].as<double>();
        }
    const int nblk = 256;
    if (rc == GD_OK) rc = part.ensure(sizeof(double) * nblk);
    if (rc == GD_OK) rc = res.ensure(sizeof(double));
    if (rc == GD_OK) rc = gd_launch_sqdiff(ctx->stream, n, p[0], p[1], p[2], p[3], part.as<double>(), nblk, res.as<double>());
    if (rc == GD_OK) {
        if (flags & GD_HOST) rc = gd_from_device(ctx, out, res.p, sizeof(double));
        else if (hipMemcpyAsync(out, res.p, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            rc = GD_ERR_HIP;
        if (!(flags & GD_HOST) && rc == GD_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    }
    for (auto &b : h) b.release();
    part.release();
    res.release();
    return rc;
}
