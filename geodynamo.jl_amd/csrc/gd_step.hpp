// gd_step.hpp -- the device-resident CNAB2 step object, shared by the step
// (gd_step.hip) and its adjoint (gd_adjoint.hip).
#pragma once
#include "gd_plan.hpp"
#include "gd_radial.hpp"

struct gd_step {
    gd_ctx *ctx = nullptr;
    gd_step_desc d{};
    gd_sht_plan sht;
    int nlm = 0, nl = 0, W = 0;
    size_t fs = 0;                 // nlm * nr
    DevBuf state[2], nlbuf[2];     // m-major FP64 [field][re/im][r][idx]; step k reads state[k&1]
    DevBuf Lt, LUt;                // [nu][(col*W+row)][l], nu = velocity, magnetic, temperature
    DevBuf nuLap, nus;             // [nu][W*nr] nu * Laplacian band; the three diffusivities
    DevBuf D1, D2, rtab;
    DevBuf synV, synS, fb, gb;
    DevBuf gb2;                    // the scalar analysis's Fourier records (-u.grad T, -u.grad C)
    DevBuf part, eout, scratch;
    DevBuf status;
    DevBuf zrow;                   // nlm zeros
    DevBuf rdiag;                  // [nu][row][l] reciprocal U diagonals
    int step = 0;
    bool chain = false;  // the column-register CNAB2 kernel applies (it also refreshes synV/synS)
    bool syn_ok = false; // synV/synS hold the synthesis columns of the current state
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // incremented whenever the step's captured work changes (its graphs are
    // dropped): batches holding the step's launches in their own graphs
    // recapture when it moved
    unsigned gen = 0;
    // second stream of the tangential step (default priority): the magnetic
    // solves, then the scalar analysis and the temperature solve, run beside
    // the synthesis / ring / vector analysis / velocity solves
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_ring = nullptr, ev_join = nullptr;
    // 0: one stream; 1: magnetic and temperature solves and the scalar analysis on
    // the side stream; 2: the scalar analysis only.  Measured at lmax 127, nr 65:
    // 2 is fastest on the FP32 path (4848 vs 4638 steps/s), 0 on the FP64 path
    // (2807 vs 2721); 1 is slower on both (three partial CNAB2 launches)
    int overlap = 0;
    // per-mode temperature conditions (gd_step_set_temperature_bc); null: the
    // boundary rows of T are read as 0 (thermal.jl defaults: Dirichlet, value 0)
    gd_scalar_bc *tbc = nullptr;
    // synthesis of the tangential step with grad T (and grad C) as gradient
    // fields (1) or as full fields with zero toroidal columns (0, same values)
    int synth_grad = 1;
    // CNAB2 column kernels with 64-bit offsets where 32 bits suffice (1; the
    // form large shells use, the same bits) or as sized (0)
    int chain_off64 = 0;
    int chain_pair = 2; // twisted solve: re and im columns in one lane (prefetch depth 2)
    int chain_ylds = -1; // paired solve: LDS rows of y (0: none, -1: as many as fit, k: at most k)
    // twisted column solve (gd_chain_twist_tables): tables for the four
    // diffusivity slots; twist = 0 keeps one lane per column
    DevBuf ULt, rdiag2, Rint;
    int twist = 0;
    // optional passive composition (gd_step_enable_composition, tangential
    // step): state and nonlinear history in the planes 8, 9 (field 4) of
    // five-field buffers whose other planes stay zero, so the CNAB2 kernel
    // treats it as a temperature-like field with diffusivity slot 3 (Pm/Sc);
    // its own synthesis columns
    bool comp = false;
    double Sc = 1.0;
    DevBuf cstate[2], cnl[2];
    gd_scalar_bc *cbc = nullptr;
    bool csyn_ok = false;
    // work buffers of the adjoint (gd_adjoint.hip), allocated on its first use
    struct gd_adj_work *adjw = nullptr;
};

// internal entry points of gd_step.hip
// one full step on the step's stream: state[parity] -> state[parity ^ 1],
// nl into nlbuf[parity]; boot = the AB1 bootstrap (step 0)
int gd_step_enqueue(gd_step *s, int parity, bool boot);
// whether the step is the tangential one (radials = 0, full_products = 0)
bool gd_step_tangential(const gd_step *s);
// drops the step's captured graphs
void gd_step_drop_graphs(gd_step *s);
// releases the adjoint's work buffers of the step (gd_adjoint.hip)
void gd_adj_work_release(gd_step *s);
// the buoyancy coefficient (Pm/E)(Pm/Pr) Ra of the radial forcing, 0 when it is off
double gd_step_buoyancy_factor(const gd_step *s);
