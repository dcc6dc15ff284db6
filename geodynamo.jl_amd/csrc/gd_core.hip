// gd_core.hip -- context, error state, index maps, FFT plans and the host FP64
// radial set-up (radial grid, FD weights, Laplacian).
#include <cmath>
#include <cstdarg>

#include "gd_internal.hpp"
#include "gd_radial.hpp"

static thread_local char g_err[1024] = "";

void gd_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *gd_last_error(void) { return g_err; }
extern "C" const char *gd_version(void) { return "gdhip 0.1 (gfx950)"; }

int gd_check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        gd_set_error("launch %s: %s", what, hipGetErrorString(e));
        return GD_ERR_HIP;
    }
    return GD_OK;
}

int DevBuf::ensure(size_t nbytes)
{
    if (nbytes <= bytes && p) return GD_OK;
    release();
    if (nbytes == 0) return GD_OK;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        gd_set_error("hipMalloc(%zu): %s", nbytes, hipGetErrorString(e));
        return GD_ERR_NOMEM;
    }
    bytes = nbytes;
    return GD_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int gd_to_device(gd_ctx *ctx, DevBuf &buf, const void *host, size_t bytes)
{
    GD_TRY(buf.ensure(bytes));
    if (bytes) GD_HIP(hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    return GD_OK;
}

int gd_from_device(gd_ctx *ctx, void *host, const void *dev, size_t bytes)
{
    if (bytes) GD_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

// ---------------------------------------------------------------- allocation-free scratch
// The partial sums of gd_timestep_error_noalloc: a fixed block of
// GD_NOALLOC_PARTIALS doubles (2 KB) in every GPU's copy of this array, held
// until the process ends.  In host code the array's name is not a device
// address: each context resolves its device's copy once, at creation, with
// that device current (hipGetSymbolAddress), and kernels receive that address.
constexpr int GD_NOALLOC_PARTIALS = 256;
static __device__ __align__(256) double scratch[GD_NOALLOC_PARTIALS];

// ---------------------------------------------------------------- context
extern "C" int gd_ctx_create(int device, gd_ctx **out)
{
    GD_CHECK(out, GD_ERR_ARG, "gd_ctx_create: null out");
    int n = 0;
    GD_HIP(hipGetDeviceCount(&n));
    GD_CHECK(device >= 0 && device < n, GD_ERR_ARG, "gd_ctx_create: device %d of %d", device, n);
    GD_HIP(hipSetDevice(device));
    gd_ctx *c = new gd_ctx();
    c->device = device;
    // a blocking stream (default priority): ordered with the legacy default
    // stream, so a caller whose arrays come from work on that stream (a framework
    // on its default stream) needs no extra synchronisation
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault);
    if (e != hipSuccess) {
        delete c;
        gd_set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return GD_ERR_HIP;
    }
    c->stream = c->own_stream;
    void *sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
        (void)hipStreamDestroy(c->own_stream);
        delete c;
        gd_set_error("gd_ctx_create: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        return GD_ERR_HIP;
    }
    c->noalloc_scratch = static_cast<double *>(sp);
    *out = c;
    return GD_OK;
}

// compute_timestep_error (timestep.jl:733-753) without any allocation: device
// arrays in, the result written to the device double `out` on the context's
// stream (asynchronous).  The reduction's 256 partial sums go to the
// context's device copy of `scratch` (every call writes all of them before
// they are read, so nothing needs clearing); the block count is that of
// gd_timestep_error, so the two return the same bits.  Any n >= 0: the
// partial count does not depend on n.  Calls that overlap on one GPU share
// the array and race.
extern "C" int gd_timestep_error_noalloc(gd_ctx *ctx, long n, const double *a_re, const double *a_im,
                                         const double *b_re, const double *b_im, double *out)
{
    GD_CHECK(ctx && a_re && a_im && b_re && b_im && out, GD_ERR_ARG, "gd_timestep_error_noalloc: null argument");
    GD_CHECK(n >= 0, GD_ERR_ARG, "gd_timestep_error_noalloc: n = %ld < 0", n);
    GD_CHECK(ctx->noalloc_scratch, GD_ERR_ARG, "gd_timestep_error_noalloc: context without scratch");
    GD_HIP(hipSetDevice(ctx->device));
    return gd_launch_sqdiff(ctx->stream, n, a_re, a_im, b_re, b_im, ctx->noalloc_scratch, GD_NOALLOC_PARTIALS,
                            out);
}

extern "C" int gd_ctx_set_stream(gd_ctx *ctx, void *s)
{
    GD_CHECK(ctx, GD_ERR_ARG, "null ctx");
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return GD_OK;
}

extern "C" int gd_ctx_synchronize(gd_ctx *ctx)
{
    GD_CHECK(ctx, GD_ERR_ARG, "null ctx");
    GD_HIP(hipSetDevice(ctx->device));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

extern "C" int gd_ctx_destroy(gd_ctx *ctx)
{
    if (!ctx) return GD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) {
        (void)hipStreamSynchronize(ctx->own_stream);
        (void)hipStreamDestroy(ctx->own_stream);
    }
    delete ctx;
    return GD_OK;
}

// ---------------------------------------------------------------- lm maps
int gd_nlm_of(int lmax, int mmax)
{
    int n = 0;
    for (int l = 0; l <= lmax; ++l) n += (l < mmax ? l : mmax) + 1;
    return n;
}

void LmMaps::build(int lmax_, int mmax_)
{
    lmax = lmax_;
    mmax = mmax_;
    nlm = gd_nlm_of(lmax, mmax);
    moff.assign(mmax + 2, 0);
    ne.assign(mmax + 1, 0);
    lmaj_of_idx.assign(nlm, 0);
    idx_of_lmaj.assign(nlm, 0);
    l_of_idx.assign(nlm, 0);
    m_of_idx.assign(nlm, 0);
    // l-major offsets (shtnskit_transforms.jl:177-188)
    std::vector<int> loff(lmax + 2, 0);
    for (int l = 0; l <= lmax; ++l) loff[l + 1] = loff[l] + (l < mmax ? l : mmax) + 1;
    int idx = 0;
    for (int m = 0; m <= mmax; ++m) {
        moff[m] = idx;
        int K = lmax - m + 1;
        ne[m] = (K + 1) / 2;
        for (int par = 0; par < 2; ++par)
            for (int l = m + par; l <= lmax; l += 2) {
                int lm = loff[l] + m;
                lmaj_of_idx[idx] = lm;
                idx_of_lmaj[lm] = idx;
                l_of_idx[idx] = l;
                m_of_idx[idx] = m;
                ++idx;
            }
    }
    moff[mmax + 1] = idx;
}

// ---------------------------------------------------------------- FFT plan
int FftPlan::build(int n_)
{
    n = n_;
    nstages = 0;
    int rem = n;
    // prefer radix 8, 4, then 6, 3, 2
    static const int order[] = {8, 4, 6, 3, 2};
    while (rem > 1) {
        bool found = false;
        for (int r : order) {
            if (rem % r == 0) {
                // keep a factor 3 available: use 8/4 only while a power of two remains
                if (nstages >= 12) return GD_ERR_ARG;
                radix[nstages++] = r;
                rem /= r;
                found = true;
                break;
            }
        }
        if (!found) {
            gd_set_error("nlon=%d must factor into 2 and 3", n);
            return GD_ERR_ARG;
        }
    }
    return GD_OK;
}

// ---------------------------------------------------------------- radial host set-up
// fields.jl:108-131
extern "C" int gd_radial_grid(int nr, double rratio, double *r)
{
    GD_CHECK(nr >= 2 && r && rratio > 0.0 && rratio < 1.0, GD_ERR_ARG, "gd_radial_grid: bad args");
    double ri = rratio / (1.0 - rratio);
    for (int n = 0; n < nr; ++n)
        r[n + nr * 3] = 0.5 * (1.0 + std::cos(M_PI * (double)(nr - (n + 1)) / (double)(nr - 1))) + ri;
    for (int p = 0; p < 7; ++p) {
        if (p == 3) continue;
        for (int n = 0; n < nr; ++n) r[n + nr * p] = std::pow(r[n + nr * 3], (double)(p - 3));
    }
    return GD_OK;
}

static int small_solve(int s, double *A, double *b)
{
    for (int k = 0; k < s; ++k) {
        int p = k;
        for (int i = k + 1; i < s; ++i)
            if (std::fabs(A[i * s + k]) > std::fabs(A[p * s + k])) p = i;
        if (A[p * s + k] == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < s; ++j) std::swap(A[k * s + j], A[p * s + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < s; ++i) {
            double f = A[i * s + k] / A[k * s + k];
            for (int j = k; j < s; ++j) A[i * s + j] -= f * A[k * s + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = s - 1; i >= 0; --i) {
        double acc = b[i];
        for (int j = i + 1; j < s; ++j) acc -= A[i * s + j] * b[j];
        b[i] = acc / A[i * s + i];
    }
    return 0;
}

// linear_algebra.jl:90-132: weights on the local stencil max(1,n-bw):min(N,n+bw)
extern "C" int gd_derivative_matrix(int order, int nr, int bw, const double *r, int fd_mode,
                                    double *band)
{
    GD_CHECK(nr >= 2 && bw >= 1 && bw <= 7 && r && band && order >= 0, GD_ERR_ARG,
             "gd_derivative_matrix: bad args");
    const double *r4 = r + nr * 3;
    std::memset(band, 0, sizeof(double) * (size_t)(2 * bw + 1) * nr);
    double V[256], VT[256], rhs[16];
    for (int n = 0; n < nr; ++n) {
        int left = n - bw < 0 ? 0 : n - bw;
        int right = n + bw > nr - 1 ? nr - 1 : n + bw;
        int s = right - left + 1;
        for (int i = 0; i < s; ++i) {
            V[i * s] = 1.0;
            for (int j = 1; j < s; ++j) V[i * s + j] = V[i * s + j - 1] * (r4[left + i] - r4[n]);
        }
        for (int i = 0; i < s; ++i) rhs[i] = 0.0;
        if (order <= s - 1) {
            double f = 1.0;
            for (int k = 2; k <= order; ++k) f *= k;
            rhs[order] = f;
        }
        int rc;
        if (fd_mode == GD_FD_TRANSPOSE) {
            for (int i = 0; i < s; ++i)
                for (int j = 0; j < s; ++j) VT[i * s + j] = V[j * s + i];
            rc = small_solve(s, VT, rhs);
        } else {
            rc = small_solve(s, V, rhs);
        }
        GD_CHECK(rc == 0, GD_ERR_NUMERIC, "gd_derivative_matrix: singular stencil at n=%d", n);
        for (int i = 0; i < s; ++i) {
            int idx = left + i, row = bw + n - idx;
            if (row >= 0 && row <= 2 * bw) band[row + (2 * bw + 1) * idx] = rhs[i];
        }
    }
    return GD_OK;
}

// linear_algebra.jl:134-151
extern "C" int gd_radial_laplacian(int nr, int bw, const double *r, int fd_mode, double *band)
{
    std::vector<double> d1((size_t)(2 * bw + 1) * nr);
    GD_TRY(gd_derivative_matrix(2, nr, bw, r, fd_mode, band));
    GD_TRY(gd_derivative_matrix(1, nr, bw, r, fd_mode, d1.data()));
    for (int n = 0; n < nr; ++n) {
        double rinv = r[n + nr * 2];
        int j0 = n - bw < 0 ? 0 : n - bw, j1 = n + bw > nr - 1 ? nr - 1 : n + bw;
        for (int j = j0; j <= j1; ++j) {
            int row = bw + n - j;
            band[row + (2 * bw + 1) * j] += 2.0 * rinv * d1[row + (2 * bw + 1) * j];
        }
    }
    return GD_OK;
}

// ---------------------------------------------------------------- seeded initial state
// splitmix64 -> uniform (0,1) -> Box-Muller normals; draws in the order
// field, level, lm (l-major), (re, im); value = amp * n * (l+1)^-power, m = 0
// kept real, rows 0 and nr-1 zeroed when zero_boundaries (no-slip,
// isothermal, insulating rows of the configurations).
static unsigned long long splitmix64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

extern "C" int gd_seeded_state(unsigned long long seed, int lmax, int mmax, int nr, int nfields, double amp,
                               double power, int zero_boundaries, double *re, double *im)
{
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax && nr >= 1 && nfields >= 1 && re && im, GD_ERR_ARG,
             "gd_seeded_state: bad args");
    const int nlm = gd_nlm_of(lmax, mmax);
    std::vector<int> lo(nlm), mo(nlm);
    int q = 0;
    for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= (l < mmax ? l : mmax); ++m) {
            lo[q] = l;
            mo[q] = m;
            ++q;
        }
    unsigned long long st = seed;
    double spare = 0.0;
    bool have = false;
    auto normal = [&]() {
        if (have) {
            have = false;
            return spare;
        }
        double u1 = ((double)(splitmix64(&st) >> 11) + 0.5) * 0x1p-53;
        double u2 = ((double)(splitmix64(&st) >> 11) + 0.5) * 0x1p-53;
        double rad = std::sqrt(-2.0 * std::log(u1));
        spare = rad * std::sin(2.0 * M_PI * u2);
        have = true;
        return rad * std::cos(2.0 * M_PI * u2);
    };
    for (int f = 0; f < nfields; ++f)
        for (int r = 0; r < nr; ++r)
            for (int k = 0; k < nlm; ++k) {
                const double sc = amp * std::pow((double)lo[k] + 1.0, -power);
                const size_t o = (size_t)f * nlm * nr + (size_t)r * nlm + k;
                double a = normal() * sc, b = normal() * sc;
                if (mo[k] == 0) b = 0.0;
                if (zero_boundaries && (r == 0 || r == nr - 1)) a = b = 0.0;
                re[o] = a;
                im[o] = b;
            }
    return GD_OK;
}
