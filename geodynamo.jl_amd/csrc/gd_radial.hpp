// gd_radial.hpp -- launchers of the FP64 radial kernels (gd_radial.hip)
#pragma once
#include "gd_internal.hpp"

int gd_launch_band_lu(hipStream_t st, const double *ab, double *lu, double *lu_t, int n, int bw, int nsys,
                      int *d_status);
int gd_launch_band_solve(hipStream_t st, const double *lu, const int *sys_of_rhs, int n, int bw, int nsys,
                         int nrhs, double *x, int *d_status);
int gd_cnab2_build(hipStream_t st, const double *d_lap, const double *d_rinv2, double nu, double dt,
                   double theta, int nl, int n, int bw, int bc_rows, double *L, double *S, double *LU, double *L_t,
                   double *LU_t, int *d_status);
int gd_launch_cnab2_lmajor(hipStream_t st, const double *L, const double *LU, int nl_mats, int lmax, int mmax,
                           int n, int bw, double dt, double theta, int mode, int bc_rows, const double *u_re,
                           const double *u_im, const double *nl_re, const double *nl_im, const double *np_re,
                           const double *np_im, double *out_re, double *out_im, int *d_status);
int gd_launch_cnab2_step(hipStream_t st, const double *nuLap, const double *rinv2, const double *nus,
                         const double *LUt, const int *l_of_idx, int nl, int n, int bw, int nlm, double dt,
                         double theta, int bc_mask, const double *st_in, double *st_out, const double *nl_arr,
                         const double *np_arr);
// Arguments of the column-register CNAB2 kernel (gd_radial.hip).  synV/synS
// (element type float or double): when set, the next step's synthesis columns
// are written from the new state.
// columns per radial level of the step's Legendre-synthesis input synV:
// 0..17 as listed in gd_step.hip, 18, 19 the composition's C/r (re, im)
constexpr int SYNV_COLS = 20;

struct ChainArgs {
    const double *nuLap, *rinv2, *nus, *LUt;
    const int *l_of_idx;
    int nl, n, bw, nlm;
    double dt, theta;
    int bc_mask;
    const double *st_in;
    double *st_out;
    const double *nl_arr, *np_arr;
    int nl_zero_mask;          // bit f: nl and nl_prev of field f are identically zero
    const double *zero_row;    // nlm zeros, read in their place
    const double *D1, *D2, *rtab;
    const double *rdiag;       // [nu][row][l] = 1 / U_l[row][row] (gd_launch_recip_diag)
    int radials;
    int syn_tang; // only the u and T/r synthesis columns are needed (the tangential step)
    void *synV, *synS;
    int f0, nf;   // fields f0 .. f0 + nf - 1 (nf = 0: all five)
    // field 5 (the composition, when c_in is set): five-field buffers of which
    // planes 8, 9 are used, diffusivity slot 3, rows preset by its conditions,
    // Dirichlet rows of S_l when c_bcr; nf = 0 then means all six
    const double *c_in;
    double *c_out;
    const double *c_nl, *c_np;
    int c_bcr;                 // its synthesis columns: SYNV_COLS - 2, - 1 of synV
    // twisted column solve of the tangential variant (null: one lane per
    // column): UL factors of the bottom blocks [k][(col*W+row)][l], their
    // reciprocal diagonals [k][row][l], interface inverses [k][l][2BW][2BW]
    const double *ULt, *rdiag2, *Rint;
    int off64; // nonzero: 64-bit offsets even where 32 bits suffice (one lane per column)
    // twisted form: 2 or 3 = each lane solves the re and the im column of its
    // (idx, field), prefetching that many levels ahead; 0 = one column per lane
    int pair;
    // paired twisted form: forward results y of the first ylds rows of each
    // half kept in LDS instead of st_out (gd_chain_ylds picks it)
    int ylds;
};
int gd_cnab2_chain_supported(int n, int bw);
// host: the twisted-solve tables of one diffusivity from S_l and LU(S_l) in
// band layout [l][col * W + row]; needs n >= 2 bw + 2
int gd_chain_twist_tables(const double *S, const double *LU, int n, int bw, int nl, double *ul_t, double *rd2,
                          double *R);
// rdiag[(k*n + i)*nl + l] = 1 / LU_t[k][(i*W + bw)][l] for nk diffusivities
int gd_launch_recip_diag(hipStream_t st, const double *LUt, int nk, int n, int bw, int nl, double *rdiag);
int gd_launch_cnab2_chain(hipStream_t st, const ChainArgs &c, bool syn_fp64);
// the kernel form gd_launch_cnab2_chain picks for c, and its LDS bytes
struct ChainForm {
    bool twist, off32, direct;
    int pair;
    // paired twisted form: forward results y of the first ylds rows of each
    // half kept in LDS instead of st_out (gd_chain_ylds picks it)
    int ylds; // the twisted form's columns per lane (0: one, 2 or 3: re and im, prefetch depth)
    size_t lds_lane, lds_twist;
};
void gd_cnab2_chain_form(const ChainArgs &c, ChainForm *f);
int gd_cnab2_chain_blocks(const ChainArgs &c);
// LDS rows of y per half for the paired twisted form of c (c.ylds unset):
// as many as keep a block within lds_budget bytes
int gd_chain_ylds(const ChainArgs &c, size_t lds_budget);
int gd_launch_cnab2_chain_batch(hipStream_t st, const BatchRef<ChainArgs> &b, int nblocks, int bw, const ChainForm &fm,
                                size_t lds, bool syn_fp64);
int gd_launch_energies(hipStream_t st, const double *state, const int *l_of_idx, int nlm, int n, double *part,
                       int nblk, double *out3);
int gd_launch_sqdiff(hipStream_t st, long n, const double *ar, const double *ai, const double *br, const double *bi,
                     double *part, int nblk, double *out1);
