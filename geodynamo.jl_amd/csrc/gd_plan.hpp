// gd_plan.hpp -- SHT plan and CNAB2 operator objects shared by the C ABI and
// the device-resident step.
#pragma once
#include "gd_internal.hpp"
#include "gd_sht.hpp"

struct gd_sht_plan {
    gd_ctx *ctx = nullptr;
    int lmax = 0, mmax = 0, nlat = 0, nlon = 0, nlat2 = 0, nlm = 0, prec = 0;
    LmMaps maps;
    FftPlan fft;
    std::vector<double> x, w, theta; // host FP64 grid (cos theta, Gauss weights, colatitude)
    DevBuf d_moff, d_ne, d_l_of_idx, d_lmaj_of_idx, d_idx_of_lmaj, d_x, d_m_of_idx;
    DevBuf tab64, tab32;             // [3][nlm][nlat2]: P, dP/dtheta, m P / sin
    DevBuf tw, sinth, costh, wq;     // in the plan's arithmetic type
    DevBuf wone, wq64;               // nlat/2 ones (plan type) and the weights 2 pi w_i / nlon in FP64
    DevBuf syn, fb, gb, outm;        // scratch for the stand-alone transforms
    DevBuf h_in0, h_in1, h_in2, h_in3, h_out0, h_out1, h_out2, h_out3; // GD_HOST staging
};

struct gd_cnab2_mats {
    gd_ctx *ctx = nullptr;
    int lmax = 0, nr = 0, bw = 0, bc_rows = 0;
    double nu = 0, dt = 0, theta = 0;
    DevBuf L, S, LU, status;
    DevBuf h0, h1, h2, h3, h4, h5, h6, h7;
};

int gd_sht_plan_build(gd_ctx *ctx, int lmax, int mmax, int nlat, int nlon, int prec, gd_sht_plan *p);
void gd_sht_plan_release(gd_sht_plan *p);
const void *gd_sht_tab(const gd_sht_plan *p, int t);
void gd_sht_fill_ring(const gd_sht_plan *p, RingArgs &a, int nr);
