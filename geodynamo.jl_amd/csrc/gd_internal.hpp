// gd_internal.hpp -- shared internals of libgdhip (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gdhip.h"

// ---------------------------------------------------------------- errors
void gd_set_error(const char *fmt, ...);

#define GD_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            gd_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return GD_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

#define GD_CHECK(cond, code, ...)        \
    do {                                 \
        if (!(cond)) {                   \
            gd_set_error(__VA_ARGS__);   \
            return (code);               \
        }                                \
    } while (0)

#define GD_TRY(call)                \
    do {                            \
        int rc_ = (call);           \
        if (rc_ != GD_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- context
struct gd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    // device address of this device's copy of the allocation-free entry
    // points' scratch array (gd_core.hip), resolved once at creation
    double *noalloc_scratch = nullptr;
};

// device buffer helper (raw, explicit lifetime)
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t nbytes);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// host<->device staging for GD_HOST calls
int gd_to_device(gd_ctx *ctx, DevBuf &buf, const void *host, size_t bytes);
int gd_from_device(gd_ctx *ctx, void *host, const void *dev, size_t bytes);

// ---------------------------------------------------------------- lm maps
// m-major "parity-grouped" spectral index used on the device:
//   for m = 0..mmax the block [moff[m], moff[m+1]) holds l = m..lmax, first the
//   ne[m] entries with (l-m) even in increasing l, then the odd ones.
struct LmMaps {
    int lmax = 0, mmax = 0, nlm = 0;
    std::vector<int> moff, ne;          // mmax+2, mmax+1
    std::vector<int> lmaj_of_idx;       // device idx -> l-major lm
    std::vector<int> idx_of_lmaj;       // l-major lm -> device idx
    std::vector<int> l_of_idx, m_of_idx;
    void build(int lmax_, int mmax_);
};
int gd_nlm_of(int lmax, int mmax);

// ---------------------------------------------------------------- vector types
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct Acc4;
template <> struct Acc4<float> { using type = floatx4; };
template <> struct Acc4<double> { using type = doublex4; };

// 16x16x4 MFMA: lane l holds A[l%16][l/16], B[l/16][l%16]; accumulator
// element v of lane l holds D[row][l%16] with row = 4*(l/16)+v for f32 and
// row = l/16 + 4*v for f64 (measured on gfx950).
__device__ __forceinline__ floatx4 mfma16x16x4(float a, float b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16x16x4(double a, double b, doublex4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Split-precision BF16 ("BF16x3") contraction of FP32 operands on the
// 16x16x16 BF16 MFMA: x = hi + lo + r with hi = bf16(x), lo = bf16(x - hi)
// (round to nearest even), |r| <= 2^-17 |x|; a b is taken as
// hi_a hi_b + hi_a lo_b + lo_a hi_b with FP32 accumulation, so each product
// carries a relative error of about 2^-16.  Lane l supplies A[l%16][4(l/16)+j]
// and B[4(l/16)+j][l%16], j = 0..3 (the k order of four 16x16x4 FP32 steps
// s = j whose lane group l/16 takes k = 4(l/16)+s), and receives the FP32
// accumulator layout of mfma16x16x4.
typedef short shortx4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned bf16_rne(float x)
{
    const unsigned u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
struct Bf16Pair {
    shortx4 hi, lo;
};
__device__ __forceinline__ Bf16Pair bf16_split4(float x0, float x1, float x2, float x3)
{
    Bf16Pair p;
    const float x[4] = {x0, x1, x2, x3};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const unsigned h = bf16_rne(x[j]);
        p.hi[j] = (short)h;
        p.lo[j] = (short)bf16_rne(x[j] - __uint_as_float(h << 16));
    }
    return p;
}
__device__ __forceinline__ floatx4 mfma_bf16x3(const Bf16Pair &a, const Bf16Pair &b, floatx4 c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, c, 0, 0, 0);
}

template <typename T> __device__ __forceinline__ int acc_row(int lane, int v);
template <> __device__ __forceinline__ int acc_row<float>(int lane, int v) { return 4 * (lane >> 4) + v; }
template <> __device__ __forceinline__ int acc_row<double>(int lane, int v) { return (lane >> 4) + 4 * v; }

template <typename T> struct Cplx { T re, im; };

// XCD-aware block order: workgroup b of an n-block launch runs on XCD b % 8,
// and each XCD has its own L2, so a kernel whose consecutive logical blocks
// share operands takes logical block xcd_logical(b, n): XCD x gets the x-th
// contiguous range of the logical order (a bijection of 0 .. n-1).
__device__ __forceinline__ int xcd_logical(int b, int n)
{
    constexpr int NX = 8;
    const int q = n / NX, rem = n % NX, x = b % NX, s = b / NX;
    return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + s;
}

// ---------------------------------------------------------------- FFT plan
struct FftPlan {
    int n = 0;
    int nstages = 0;
    int radix[12] = {0};
    int build(int n_);
};

// A batched launch: one grid over the blocks that several shells' launches of
// one kernel form (the same compile-time form for all); block b belongs to
// shell shell_of_block[b] and is that shell's block b - block_off[shell] of
// block_cnt[shell], and args[shell] are that shell's arguments (device
// arrays).  Each shell's run of blocks starts at a multiple of 8 (the XCD
// count, so a shell's block i runs on XCD i % 8 as in its own launch); the
// padding blocks between runs carry shell BATCH_PAD and exit at once.
constexpr unsigned short BATCH_PAD = 0xFFFF;
template <class A> struct BatchRef {
    const A *args;
    const unsigned short *shell_of_block;
    const int *block_off, *block_cnt;
};

// A batched kernel reads its shell's arguments through the constant address
// space: the table is written before the launch and never during it, so, as
// with a kernel's own arguments, the compiler may load each field once into
// scalar registers.  Through a generic pointer every store of the body may
// alias the table, and the fields are loaded again inside the loops (each
// reload waits on the scalar and LDS counter together).  The kernel bodies
// take the arguments as a deduced type for this.
template <class A> using ConstArgs = const __attribute__((address_space(4))) A;
template <class A> __device__ __forceinline__ ConstArgs<A> &batch_args(const BatchRef<A> &b, int sh)
{
    return *(ConstArgs<A> *)(b.args + sh);
}

// launch helper: returns GD_ERR_HIP on launch failure
int gd_check_launch(const char *what);

// scalar boundary conditions (gd_bc.hip) for callers inside the library:
// per-mode types/values given in l-major order, stored for planes whose column
// k holds l-major mode lm_of_pos[k]; launch in place on device planes
int gd_scalar_bc_set_perm(gd_scalar_bc *b, const int *bc_inner, const int *bc_outer, const double *values,
                          const int *lm_of_pos);
int gd_scalar_bc_launch(gd_scalar_bc *b, hipStream_t st, double *re, double *im, int dirichlet);
