// gd_sht.hpp -- SHT device kernels (gd_sht.hip): argument blocks and launchers.
//
// Device spectral layout ("m-major"): index idx in [0, nlm) with, for each m,
// the block [moff[m], moff[m+1]) holding l = m..lmax, first the ne[m] degrees
// with l-m even, then those with l-m odd (LmMaps).  A spectral column set is
// stored [col][idx] with col = level*cin + c.
//
// Legendre tables Tab[t][idx][i], i in [0, nlat/2) the northern Gauss nodes,
// t = 0 P, 1 dP/dtheta, 2 m P / sin(theta).  Southern values follow from the
// parity P(pi-theta) = (-1)^(l+m) P(theta) (dP flips sign once more).
//
// Fourier layouts (complex of the real longitude transform, m = 0..mmax):
//   FB[i][level][m][slot]  synthesis output, north and south rings of node i
//   GB[m][i][level][slot]  analysis input G+ = Gn + Gs, G- = Gn - Gs, already
//                          weighted by 2 pi w_i / nlon
#pragma once
#include "gd_internal.hpp"

struct SynthArgs {
    const void *in;       // T*, [col][idx], col = level*cin + c
    int cin;              // input columns per level
    int nf;               // vector: fields (4 virtual cols each); scalar: fields (2 each)
    int colmap[20];       // virtual col within a level -> input col c, -1 = zero
    const void *tab0;     // vector: dP, scalar: P
    const void *tab1;     // vector: Q = m P / sin
    const int *moff, *ne;
    int nlat2, nr, nlm, mmax;
    void *out;            // T*, FB
    int nslot, slot0;     // vector field f -> slot0 + 8f + 4h + (2c + reim); scalar -> slot0 + 4f + 2h + reim
    int ng;               // vector: gradient fields (S only, T = 0) after the nf full ones; colmap
                          // then lists 4 nf + 2 ng columns (S_re S_im per gradient field)
    int mm;               // 1: BF16x3 split-precision contraction (FP32 operands), 0: native MFMA
};

struct AnaArgs {
    const void *gb;       // T*, GB
    int nslot_g, gslot0;  // component k of field f: gslot0 + 4*(comp index) + 2*pm + reim
    int nf;               // vector fields (components 2V, 2V+1) or scalar fields
    const void *tab0, *tab1;
    const int *moff, *ne, *l_of_idx;
    int nlat2, nr, nlm, mmax;
    int maxgroup;         // largest parity group, ceil((lmax+1)/2); set by gd_launch_ana
    // outputs, FP64 m-major [reim][level][idx]; vector field V: out[2V] gets S, out[2V+1] gets T
    double *out[4];
    int scale[4];         // 0: /l(l+1) (0 at l=0), 1: x1, 2: x(-1)
    // table-resident analysis in node chunks (set by its launcher): nodes
    // k0 .. k0 + nk - 1, accum = 1 adds to out instead of storing
    int k0, nk, accum;
    int mm;               // 1: BF16x3 split-precision contraction (FP32 operands), 0: native MFMA
};

struct RingArgs {
    const void *fb;       // T*
    void *gb;             // T*
    int nslot_f, nslot_g;
    // optional second analysis input: products np1 .. (the scalar analysis's)
    // in their own records GB2[m][i][level][nslot_g2]; gb2 = null keeps every
    // product in GB
    void *gb2;
    int nslot_g2, np1;
    int nfield;           // real fields to synthesise (SYNTH/STEP) or analyse (ANALYSIS)
    int fn_off[16], fs_off[16]; // FB slot (re) of north / south spectra of each field
    const double *phys_in[2];
    double *phys_out[2];
    const void *tw;       // T2[nlon], e^{-2 pi i k / nlon}
    const void *sinth, *costh, *wq; // T[nlat] sin, cos of colatitude; T[nlat2] 2 pi w / nlon
    int nlat, nlat2, nlon, nr, mmax;
    int nstages, radix[12];
    double ro;            // Pm / E
    int coriolis_index;
};

// RING_STEP: 15 fields in (u, omega, B, j, grad T tangential; u_r, omega_r,
// B_r, j_r, dT/dr), 5 products out.  RING_STEP_TANG: the step when every
// radial component is zero -- 4 fields in (u_theta, u_phi, (grad T)_theta,
// (grad T)_phi), 3 products out (the force and -u.grad T; u x B vanishes).
// RING_STEP_TANG_C: the same with the composition's gradient as fields 4, 5
// and -u.grad C as a fourth product.
// RING_ADJ_TANG: the transposed Jacobian of RING_STEP_TANG's products at the
// state (u, grad T) applied to the cotangents of its products -- 7 fields in
// (u_theta, u_phi, cotangent force theta / phi, (grad T)_theta, (grad T)_phi,
// cotangent of -u.grad T), 4 out (cotangents of u_theta, u_phi,
// (grad T)_theta, (grad T)_phi).
enum { RING_SYNTH = 0, RING_ANALYSIS = 1, RING_STEP = 2, RING_STEP_TANG = 3, RING_STEP_TANG_C = 4,
       RING_ADJ_TANG = 5 };

int gd_launch_plm(hipStream_t st, const double *x, const int *moff, int lmax, int mmax, int nlat2, double *tab64,
                  float *tab32, int nlm);
template <typename T> int gd_launch_synth_vec(hipStream_t st, const SynthArgs &a);
template <typename T> int gd_launch_synth_scal(hipStream_t st, const SynthArgs &a);
template <typename T> int gd_launch_ana(hipStream_t st, const AnaArgs &a, bool vec, int maxgroup);
template <typename T> int gd_launch_ring(hipStream_t st, int mode, const RingArgs &a);
template <typename T>
int gd_launch_pack(hipStream_t st, const double *re, const double *im, const int *lmaj_of_idx, int nlm, int nr,
                   int cin, int c0, T *out);
int gd_launch_unpack(hipStream_t st, const double *src, const int *idx_of_lmaj, int nlm, int nr, double *dst);
int gd_launch_perm(hipStream_t st, const double *src, const int *map, int nlm, int nr, double *dst);

// ---- batched step launches (gd_sht.hip): one grid over several shells'
// launches of the tangential step's kernels, all in one compile-time form.
// Block counts of one shell: gd_synth_tang_blocks (tile 0, 1, or 2 = FP32
// only), nlat/2 * nr for the ring, gd_ana_blocks (args with maxgroup set) for
// the analyses, whose node tile gd_ana_kn gives (0: several node chunks, not
// batched).
int gd_synth_tang_blocks(const SynthArgs &a, int tile);
bool gd_synth_tile_ok(int tile, bool fp64);
template <typename T>
int gd_launch_synth_tang_batch(hipStream_t st, const BatchRef<SynthArgs> &b, int nblocks, int tile, int mm);
template <typename T> int gd_launch_synth_tang_tile(hipStream_t st, const SynthArgs &a, int tile);
template <typename T> int gd_launch_ring_tang_batch(hipStream_t st, const BatchRef<RingArgs> &b, int nblocks, int nlon);
int gd_synth_tang_batchable(const SynthArgs &a);
int gd_ring_tang_batchable(const RingArgs &a, bool fp64);
int gd_ana_kn(int nlat2, bool fp64);
int gd_ana_blocks(const AnaArgs &a);
template <typename T>
int gd_launch_ana_batch(hipStream_t st, const BatchRef<AnaArgs> &b, int nblocks, bool vec, int kn, int mm);
