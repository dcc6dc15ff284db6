// gd_step.hip -- the device-resident pseudo-spectral CNAB2 step.
//
// One step (scripts/bench_timestepping.jl:6-18 semantics, all nonlinear terms
// from the current state):
//   1. spectral_pre_kernel  curl of u and B (velocity.jl:427-435,
//      magnetic.jl:335-343), radial components X_r = l(l+1) P_X / r^2
//      (magnetic.jl:257-259), T/r and dT/dr (scalar_field_common.jl:233-327),
//      written as Legendre-GEMM input columns in the plan's arithmetic type.
//      Only after set_state and on the first step: afterwards the CNAB2
//      kernel of step 5 writes these columns for the new state.
//   2. synth_kernel (vector: u, omega, B, j, grad T; scalar: u_r, omega_r,
//      B_r, j_r, dT/dr, or dT/dr alone when the radial components are zero)
//      -> Fourier rings FB
//   3. ring_kernel<STEP>: inverse FFTs, u x omega + Coriolis + j x B, u x B,
//      -u.grad T (velocity.jl:512-527,664-670; magnetic.jl:399-401;
//      scalar_field_common.jl:401-407), forward FFTs -> GB
//   4. ana_kernel: vector analysis of the force and of u x B (with the curl
//      factors of magnetic.jl:437-440), scalar analysis of -u.grad T -> nl
//   5. cnab2_chain_kernel (cnab2_step_kernel where the column does not fit in
//      registers): temperature Dirichlet rows, AB2 (AB1 at step 0,
//      simulation.jl:1153), rhs + FP64 banded solve per column
//      (timestep.jl:610-730); the nl history rolls by buffer swap
//      (simulation.jl:1312-1329).
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "gd_step.hpp"


// ---------------------------------------------------------------- spectral preparation
// One thread per (idx, level).  synV columns per level (c = 2k + re/im):
//   k: 0 u_pol 1 u_tor 2 w_pol 3 w_tor 4 B_pol 5 B_tor 6 j_pol 7 j_tor 8 T/r
//   9 C/r (the composition, when enabled; written by comp_pre_kernel and the
//   CNAB2 kernel)
// synS columns per level: k: 0 u_r 1 w_r 2 B_r 3 j_r 4 dT/dr
template <typename T>
__global__ void __launch_bounds__(256) spectral_pre_kernel(const double *__restrict__ st, const int *__restrict__ l_of_idx,
                                                           const double *__restrict__ D1, const double *__restrict__ D2,
                                                           const double *__restrict__ rtab, int nlm, int nr, int bw,
                                                           int radials, T *__restrict__ synV, T *__restrict__ synS)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr;
    const int W = 2 * bw + 1;
    const double lf = (double)l_of_idx[idx] * (l_of_idx[idx] + 1);
    const double rinv = rtab[r + nr * 2], rinv2 = rtab[r + nr * 1];
    const int j0 = r - bw < 0 ? 0 : r - bw, j1 = r + bw > nr - 1 ? nr - 1 : r + bw;
#pragma unroll
    for (int reim = 0; reim < 2; ++reim) {
        const double *utor = st + (size_t)(0 + reim) * fs, *upol = st + (size_t)(2 + reim) * fs;
        const double *Btor = st + (size_t)(4 + reim) * fs, *Bpol = st + (size_t)(6 + reim) * fs;
        const double *Tt = st + (size_t)(8 + reim) * fs;
        double d1u = 0, d2u = 0, d1B = 0, d2B = 0, d1T = 0;
        for (int j = j0; j <= j1; ++j) {
            const int b = bw + r - j + W * j;
            const size_t o = (size_t)j * nlm + idx;
            const double c1 = D1[b], c2 = D2[b];
            const double up = upol[o], Bp = Bpol[o];
            d1u += c1 * up;
            d2u += c2 * up;
            d1B += c1 * Bp;
            d2B += c2 * Bp;
            d1T += c1 * Tt[o];
        }
        const size_t o = (size_t)r * nlm + idx;
        const double up = upol[o], ut = utor[o], Bp = Bpol[o], Bt = Btor[o], tv = Tt[o];
        const double wtor = lf * rinv2 * up - d2u - 2.0 * rinv * d1u;
        const double wpol = -lf * rinv2 * ut;
        const double jtor = lf * rinv2 * Bp - d2B - 2.0 * rinv * d1B;
        const double jpol = -lf * rinv2 * Bt;
        T *v = synV + (size_t)r * SYNV_COLS * nlm + idx;
        v[(0 + reim) * (size_t)nlm] = (T)up;
        v[(2 + reim) * (size_t)nlm] = (T)ut;
        v[(4 + reim) * (size_t)nlm] = (T)wpol;
        v[(6 + reim) * (size_t)nlm] = (T)wtor;
        v[(8 + reim) * (size_t)nlm] = (T)Bp;
        v[(10 + reim) * (size_t)nlm] = (T)Bt;
        v[(12 + reim) * (size_t)nlm] = (T)jpol;
        v[(14 + reim) * (size_t)nlm] = (T)jtor;
        v[(16 + reim) * (size_t)nlm] = (T)(tv * rinv);
        T *s = synS + (size_t)r * 10 * nlm + idx;
        if (radials) { // otherwise the radial components stay zero (cleared at creation)
            s[(0 + reim) * (size_t)nlm] = (T)(lf * up * rinv2);
            s[(2 + reim) * (size_t)nlm] = (T)(lf * wpol * rinv2);
            s[(4 + reim) * (size_t)nlm] = (T)(lf * Bp * rinv2);
            s[(6 + reim) * (size_t)nlm] = (T)(lf * jpol * rinv2);
        }
        s[(8 + reim) * (size_t)nlm] = (T)d1T;
    }
}

// ---------------------------------------------------------------- orchestration
static double *cur_state(gd_step *s) { return s->state[s->step & 1].as<double>(); }

// ev (optional): events recorded after each kernel class, ev[0] before the first
#define GD_MARK(k)                                                  \
    do {                                                            \
        if (ev) GD_HIP(hipEventRecord(ev[k], s->ctx->stream));      \
    } while (0)

// With every radial component zero (radials = 0) the tangential parts of
// u x omega, j x B and u x B vanish identically and u_r dT/dr drops out, so
// only u and grad T need transforms; full_products = 1 keeps every product.
static bool tangential(const gd_step *s) { return !s->d.radials && !s->d.full_products; }

// side (tangential step only): the scalar analysis runs on s->side after the
// ring kernel, the vector analysis on the main stream
// the Legendre-synthesis columns of a state
template <typename T> static int spectral_pre_into(gd_step *s, const double *state, DevBuf &synV, DevBuf &synS)
{
    gd_sht_plan *p = &s->sht;
    const int nr = s->d.nr, nlm = s->nlm;
    hipLaunchKernelGGL((spectral_pre_kernel<T>), dim3((nlm + 255) / 256, nr), dim3(256), 0, s->ctx->stream, state,
                       p->d_l_of_idx.as<int>(), s->D1.as<double>(), s->D2.as<double>(), s->rtab.as<double>(), nlm, nr,
                       s->d.bw, s->d.radials, synV.as<T>(), synS.as<T>());
    return gd_check_launch("spectral_pre_kernel");
}

// The buoyancy of velocity.jl:534-537, F_r = (Pm/E)(Pm/Pr) Ra r T, into the
// poloidal equation (radial_force = 1): the poloidal scalar with radial
// component F_r (u_r = l(l+1) P / r^2, magnetic.jl:257-259) is
// r^2 F_r / (l(l+1)), so nl(u_pol) += c r^3 T / (l(l+1)) for l >= 1 (l = 0 has
// no poloidal flow).  Linear in T and spectral: no transform is needed.
__global__ void __launch_bounds__(256) buoyancy_kernel(const double *__restrict__ st, const int *__restrict__ l_of_idx,
                                                       const double *__restrict__ rtab, double c, int nlm, int nr,
                                                       double *__restrict__ nl)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const int l = l_of_idx[idx];
    if (l == 0) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)r * nlm + idx;
    const double rr = rtab[r + 3 * nr];
    const double f = c * rr * rr * rr / ((double)l * (double)(l + 1));
    nl[2 * fs + o] += f * st[8 * fs + o];
    nl[3 * fs + o] += f * st[9 * fs + o];
}

double gd_step_buoyancy_factor(const gd_step *s)
{
    return s->d.radial_force ? (s->d.Pm / s->d.E) * (s->d.Pm / s->d.Pr) * s->d.Ra : 0.0;
}

template <typename T> static int spectral_pre(gd_step *s)
{
    GD_TRY(spectral_pre_into<T>(s, cur_state(s), s->synV, s->synS));
    s->syn_ok = true;
    return GD_OK;
}

// the composition's synthesis columns C/r (SYNV_COLS - 2, - 1 of synV) from
// planes 8, 9 of its state buffer, as the CNAB2 kernel writes them each step
template <typename T>
__global__ void __launch_bounds__(256) comp_pre_kernel(const double *__restrict__ cst, const double *__restrict__ rtab,
                                                       int nlm, int nr, T *__restrict__ synV)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (idx >= nlm) return;
    const size_t fs = (size_t)nlm * nr, o = (size_t)r * nlm + idx;
    const double rinv = rtab[r + nr * 2];
    T *v = synV + (size_t)r * SYNV_COLS * nlm + idx;
    v[(size_t)(SYNV_COLS - 2) * nlm] = (T)(cst[8 * fs + o] * rinv);
    v[(size_t)(SYNV_COLS - 1) * nlm] = (T)(cst[9 * fs + o] * rinv);
}

template <typename T> static int spectral_pre_comp(gd_step *s)
{
    const int nr = s->d.nr, nlm = s->nlm;
    hipLaunchKernelGGL((comp_pre_kernel<T>), dim3((nlm + 255) / 256, nr), dim3(256), 0, s->ctx->stream,
                       s->cstate[s->step & 1].as<double>(), s->rtab.as<double>(), nlm, nr, s->synV.as<T>());
    GD_TRY(gd_check_launch("comp_pre_kernel"));
    s->csyn_ok = true;
    return GD_OK;
}

// Graphs of the chain path are captured once the synthesis columns are valid
// and do not contain the spectral preparation; after a restore the columns
// are stale, so they are rebuilt eagerly before the next replay.
static int ensure_syn(gd_step *s)
{
    if (!s->chain) return GD_OK;
    const bool f64 = s->d.prec == GD_PREC_FP64;
    if (!s->syn_ok) GD_TRY(f64 ? spectral_pre<double>(s) : spectral_pre<float>(s));
    if (s->comp && !s->csyn_ok) GD_TRY(f64 ? spectral_pre_comp<double>(s) : spectral_pre_comp<float>(s));
    return GD_OK;
}

// The transform arguments common to both step forms: the synthesis a (the
// vector columns synV into FB), the ring pass r (FB into GB, the scalar
// analysis's products into GB2) and the analysis g (GB into nl_out).
static void base_args(gd_step *s, SynthArgs &a, RingArgs &r, AnaArgs &g)
{
    gd_sht_plan *p = &s->sht;
    const int nr = s->d.nr, nlm = s->nlm;
    a = SynthArgs{};
    a.mm = p->prec == GD_PREC_BF16X3;
    a.moff = p->d_moff.as<int>();
    a.ne = p->d_ne.as<int>();
    a.nlat2 = p->nlat2;
    a.nr = nr;
    a.nlm = nlm;
    a.mmax = p->mmax;
    a.out = s->fb.p;
    a.in = s->synV.p;
    a.cin = SYNV_COLS;
    a.tab0 = gd_sht_tab(p, 1);
    a.tab1 = gd_sht_tab(p, 2);
    a.slot0 = 0;
    r = RingArgs{};
    gd_sht_fill_ring(p, r, nr);
    r.fb = s->fb.p;
    r.gb = s->gb.p;
    // the scalar analysis's products in records of their own, so that neither
    // analysis reads the other's slots (cache lines hold whole records)
    r.gb2 = s->gb2.p;
    r.ro = s->d.Pm / s->d.E;
    r.coriolis_index = s->d.coriolis_index;
    g = AnaArgs{};
    g.mm = p->prec == GD_PREC_BF16X3;
    g.gb = s->gb.p;
    g.moff = p->d_moff.as<int>();
    g.ne = p->d_ne.as<int>();
    g.l_of_idx = p->d_l_of_idx.as<int>();
    g.nlat2 = p->nlat2;
    g.nr = nr;
    g.nlm = nlm;
    g.mmax = p->mmax;
    g.maxgroup = (p->lmax + 2) / 2;
}

// The tangential step's arguments: u (S = u_pol, T = u_tor) and grad T
// (S = T/r) only; FB records of 16 slots, 24 with the composition's gradient
// (S = C/r) in slots 16..23.  gv: the vector analysis of the force, gs: the
// scalar analysis of -u.grad T (and -u.grad C), into nl_out; *mode: the ring
// pass's form.
static void tang_args(gd_step *s, double *nl_out, SynthArgs &a, RingArgs &r, AnaArgs &gv, AnaArgs &gs, int *mode)
{
    gd_sht_plan *p = &s->sht;
    const size_t fs = s->fs;
    AnaArgs g;
    base_args(s, a, r, g);
    const int nslot = s->comp ? 24 : 16;
    a.nslot = nslot;
    if (s->synth_grad) {
        // u as a full field, grad T (and grad C) as gradient fields: their
        // zero toroidal columns are not contracted
        a.nf = 1;
        a.ng = s->comp ? 2 : 1;
        static const int gmap[8] = {0, 1, 2, 3, 16, 17, SYNV_COLS - 2, SYNV_COLS - 1};
        for (int k = 0; k < 20; ++k) a.colmap[k] = k < 4 + 2 * a.ng ? gmap[k] : -1;
    } else {
        a.nf = s->comp ? 3 : 2;
        static const int vmap[12] = {0, 1, 2, 3, 16, 17, -1, -1, SYNV_COLS - 2, SYNV_COLS - 1, -1, -1};
        for (int k = 0; k < 20; ++k) a.colmap[k] = k < 4 * a.nf ? vmap[k] : -1;
    }
    r.nslot_f = nslot;
    // with composition the ring also takes grad C (fields 4, 5) and forms
    // -u.grad C (compute_composition_nonlinear!, compositional.jl:219-260)
    // as a fourth product; GB records of 12 slots, 16 with it
    r.nfield = s->comp ? 6 : 4;
    r.nslot_g = 8;                    // the force (products 0, 1)
    r.np1 = 2;
    r.nslot_g2 = s->comp ? 8 : 4;     // -u.grad T (and -u.grad C)
    static const int fn[6] = {0, 2, 8, 10, 16, 18}, fsou[6] = {4, 6, 12, 14, 20, 22};
    for (int F = 0; F < r.nfield; ++F) {
        r.fn_off[F] = fn[F];
        r.fs_off[F] = fsou[F];
    }
    *mode = s->comp ? RING_STEP_TANG_C : RING_STEP_TANG;
    // force -> (pol, tor) of velocity nl (velocity.jl:331); the magnetic nl
    // is u x B = 0 here and its planes stay zero
    gv = g;
    gv.nslot_g = r.nslot_g;
    gv.nf = 1;
    gv.gslot0 = 0;
    gv.tab0 = gd_sht_tab(p, 1);
    gv.tab1 = gd_sht_tab(p, 2);
    gv.out[0] = nl_out + 2 * fs; // S(adv) -> u_pol
    gv.out[1] = nl_out + 0 * fs; // T(adv) -> u_tor
    gv.scale[0] = 0;
    gv.scale[1] = 0;
    gs = g;
    gs.gb = s->gb2.p;
    gs.nslot_g = r.nslot_g2;
    gs.gslot0 = 0;
    gs.nf = 1;
    gs.tab0 = gd_sht_tab(p, 0);
    gs.tab1 = nullptr;
    gs.out[0] = nl_out + 8 * fs; // -u.grad T -> T
    gs.scale[0] = 1;
    if (s->comp) { // and -u.grad C -> the composition's nl (GB slots 12..15)
        gs.nf = 2;
        gs.out[1] = s->cnl[s->step & 1].as<double>() + 8 * fs;
        gs.scale[1] = 1;
    }
}

template <typename T> static int radial_force(gd_step *s, hipStream_t st, double *nl_out)
{
    if (!s->d.radial_force) return GD_OK;
    // after the vector analysis wrote nl(u_pol), on its stream
    hipLaunchKernelGGL(buoyancy_kernel, dim3((s->nlm + 255) / 256, s->d.nr), dim3(256), 0, st, cur_state(s),
                       s->sht.d_l_of_idx.as<int>(), s->rtab.as<double>(), gd_step_buoyancy_factor(s), s->nlm,
                       s->d.nr, nl_out);
    return gd_check_launch("buoyancy_kernel");
}

template <typename T> static int nonlinear_T(gd_step *s, double *nl_out, hipEvent_t *ev, bool side)
{
    hipStream_t st = s->ctx->stream;
    GD_MARK(0);
    gd_sht_plan *p = &s->sht;
    const int nr = s->d.nr;
    if (!s->syn_ok) GD_TRY(spectral_pre<T>(s));
    GD_MARK(1);
    const size_t fs = s->fs;
    const int maxgroup = (p->lmax + 2) / 2;

    if (tangential(s)) {
        if (s->comp && !s->csyn_ok) GD_TRY(spectral_pre_comp<T>(s));
        SynthArgs a;
        RingArgs r;
        AnaArgs gv, gs;
        int mode;
        tang_args(s, nl_out, a, r, gv, gs, &mode);
        GD_TRY(gd_launch_synth_vec<T>(st, a));
        GD_MARK(2);
        GD_MARK(3);
        GD_TRY(gd_launch_ring<T>(st, mode, r));
        GD_MARK(4);
        if (side) {
            GD_HIP(hipEventRecord(s->ev_ring, st));
            GD_HIP(hipStreamWaitEvent(s->side, s->ev_ring, 0));
        }
        GD_TRY(gd_launch_ana<T>(st, gv, true, maxgroup));
        GD_MARK(5);
        GD_TRY(gd_launch_ana<T>(side ? s->side : st, gs, false, maxgroup));
        GD_TRY(radial_force<T>(s, st, nl_out));
        GD_MARK(6);
        return GD_OK;
    }
    SynthArgs a;
    RingArgs r;
    AnaArgs g;
    base_args(s, a, r, g);
    {
        // vector fields u, omega, B, j (S,T) and grad T (S only)
        a.nslot = 60;
        a.nf = 5;
        static const int vmap[20] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, -1, -1};
        for (int k = 0; k < 20; ++k) a.colmap[k] = vmap[k];
        GD_TRY(gd_launch_synth_vec<T>(st, a));
        GD_MARK(2);
        a.in = s->synS.p;
        a.cin = 10;
        a.tab0 = gd_sht_tab(p, 0);
        a.tab1 = nullptr;
        if (s->d.radials) {
            a.nf = 5;
            for (int k = 0; k < 20; ++k) a.colmap[k] = k < 10 ? k : -1;
            a.slot0 = 40;
        } else {
            // the radial components are identically zero (their FB slots were
            // cleared at creation and are never written): only dT/dr is synthesised
            a.nf = 1;
            for (int k = 0; k < 20; ++k) a.colmap[k] = k < 2 ? 8 + k : -1;
            a.slot0 = 56;
        }
        GD_TRY(gd_launch_synth_scal<T>(st, a));
        GD_MARK(3);
        r.nslot_f = 60;
        r.nslot_g = 16;                   // the force and u x B (products 0..3)
        r.np1 = 4;
        r.nslot_g2 = 4;                   // -u.grad T
        r.nfield = 15;
        for (int F = 0; F < 10; ++F) {
            const int f = F / 2, c = F % 2;
            r.fn_off[F] = 8 * f + 2 * c;
            r.fs_off[F] = 8 * f + 4 + 2 * c;
        }
        for (int F = 10; F < 15; ++F) {
            r.fn_off[F] = 40 + 4 * (F - 10);
            r.fs_off[F] = 40 + 4 * (F - 10) + 2;
        }
        GD_TRY(gd_launch_ring<T>(st, RING_STEP, r));
        GD_MARK(4);
        // force -> (pol, tor) of velocity nl (velocity.jl:331); u x B -> curl (magnetic.jl:437-440)
        g.nslot_g = 16;
        g.nf = 2;
        g.gslot0 = 0;
        g.tab0 = gd_sht_tab(p, 1);
        g.tab1 = gd_sht_tab(p, 2);
        g.out[0] = nl_out + 2 * fs; // S(adv)   -> u_pol
        g.out[1] = nl_out + 0 * fs; // T(adv)   -> u_tor
        g.out[2] = nl_out + 4 * fs; // l(l+1) S(uxB)  -> B_tor
        g.out[3] = nl_out + 6 * fs; // -l(l+1) T(uxB) -> B_pol
        g.scale[0] = 0;
        g.scale[1] = 0;
        g.scale[2] = 1;
        g.scale[3] = 2;
        GD_TRY(gd_launch_ana<T>(st, g, true, maxgroup));
        GD_MARK(5);
        g.gb = s->gb2.p;
        g.nslot_g = r.nslot_g2;
        g.gslot0 = 0;
    }
    g.nf = 1;
    g.tab0 = gd_sht_tab(p, 0);
    g.tab1 = nullptr;
    g.out[0] = nl_out + 8 * fs; // -u.grad T -> T
    g.scale[0] = 1;
    GD_TRY(gd_launch_ana<T>(side ? s->side : st, g, false, maxgroup));
    GD_TRY(radial_force<T>(s, st, nl_out));
    GD_MARK(6);
    return GD_OK;
}

static int nonlinear(gd_step *s, double *nl_out, hipEvent_t *ev = nullptr, bool side = false)
{
    if (s->d.prec == GD_PREC_FP64) return nonlinear_T<double>(s, nl_out, ev, side);
    return nonlinear_T<float>(s, nl_out, ev, side);
}

// fields whose implicit operator carries Dirichlet rows (bit f for field f)
static int bc_mask(const gd_step *s)
{
    const int t = s->tbc ? 0x20 : 0; // temperature rows set by the condition stage
    if (s->d.bc_rows == 1) return 0x1f | t;
    if (s->d.bc_rows == 2) return 0x10 | t;
    return t;
}

// the temperature conditions of thermal.jl:262-313 on the state the solve
// reads, after the nonlinear terms (which see the state before them)
static int temperature_bc(gd_step *s, int parity, hipStream_t st)
{
    if (!s->tbc) return GD_OK;
    double *cur = s->state[parity].as<double>();
    return gd_scalar_bc_launch(s->tbc, st, cur + 8 * s->fs, cur + 9 * s->fs, 1);
}

// the composition's conditions (compositional.jl:318-370; default no flux at
// both boundaries) on the composition state the solve reads
static int composition_bc(gd_step *s, int parity, hipStream_t st)
{
    if (!s->comp) return GD_OK;
    double *cc = s->cstate[parity].as<double>();
    return gd_scalar_bc_launch(s->cbc, st, cc + 8 * s->fs, cc + 9 * s->fs, 1);
}

// one full step with the nl buffers chosen by parity; boot = AB1 bootstrap.
// Tangential step with the chain kernel (and no per-class timing), s->overlap:
//   2  the scalar analysis runs on the side stream beside the vector analysis
//      (both read the ring output; they write disjoint nl planes), joined
//      before the one CNAB2 launch;
//   1  as 2, and the magnetic fields' solves (their nonlinear terms are
//      identically zero, so they depend only on the previous state) run on the
//      side stream from the start of the step, the temperature solve after the
//      scalar analysis there, the velocity solves on the main stream.  Written
//      buffers are disjoint between the streams (state fields, nl planes,
//      synthesis columns of u and of T/r).
// the CNAB2 column solves' arguments of a step of the given parity (the chain form)
static void chain_args(gd_step *s, int parity, bool boot, ChainArgs &c)
{
    double *cur = s->nlbuf[parity].as<double>();
    double *prev = boot ? cur : s->nlbuf[parity ^ 1].as<double>();
    c = ChainArgs{};
    c.nuLap = s->nuLap.as<double>();
    c.rinv2 = s->rtab.as<double>() + s->d.nr;
    c.nus = s->nus.as<double>();
    c.LUt = s->LUt.as<double>();
    c.l_of_idx = s->sht.d_l_of_idx.as<int>();
    c.nl = s->nl;
    c.n = s->d.nr;
    c.bw = s->d.bw;
    c.nlm = s->nlm;
    c.dt = s->d.dt;
    c.theta = s->d.theta;
    c.bc_mask = bc_mask(s);
    c.st_in = s->state[parity].as<double>();
    c.st_out = s->state[parity ^ 1].as<double>();
    c.nl_arr = cur;
    c.np_arr = prev;
    // the magnetic nonlinear terms vanish identically in the tangential step
    c.nl_zero_mask = tangential(s) ? 0x0c : 0;
    c.zero_row = s->zrow.as<double>();
    c.rdiag = s->rdiag.as<double>();
    c.D1 = s->D1.as<double>();
    c.D2 = s->D2.as<double>();
    c.rtab = s->rtab.as<double>();
    c.radials = s->d.radials;
    c.synV = s->synV.p;
    c.synS = s->synS.p;
    c.syn_tang = tangential(s) ? 1 : 0;
    c.off64 = s->chain_off64;
    c.pair = s->chain_pair;
    c.ylds = 0;
    if (s->twist) {
        c.ULt = s->ULt.as<double>();
        c.rdiag2 = s->rdiag2.as<double>();
        c.Rint = s->Rint.as<double>();
    }
    if (s->comp) { // the composition as column group 5 of the same launch
        c.c_in = s->cstate[parity].as<double>();
        c.c_out = s->cstate[parity ^ 1].as<double>();
        c.c_nl = s->cnl[parity].as<double>();
        c.c_np = boot ? c.c_nl : s->cnl[parity ^ 1].as<double>();
        c.c_bcr = s->d.bc_rows == 1;
    }
    // the paired twisted solve keeps as many forward results in LDS as leave
    // the launch in one round of resident blocks (at most 64 KB a block).
    // The budget is a whole number of 2 KB granules: at lmax=127, nr=65 (645
    // blocks, three per CU) 18 rows (50.6 KB a block) ran in one round and 19
    // rows (52.6 KB, below a third of 160 KB) in two, 49 against 67 us
    if (s->chain_ylds != 0) {
        const long blocks = (long)gd_cnab2_chain_blocks(c), per_cu = std::max(1L, (blocks + 255) / 256);
        size_t budget = std::min<size_t>(64 * 1024, (160 * 1024 / per_cu) & ~(size_t)2047);
        c.ylds = gd_chain_ylds(c, budget);
        if (s->chain_ylds > 0) c.ylds = std::min(c.ylds, s->chain_ylds);
    }
}

static int enqueue_step(gd_step *s, int parity, bool boot, hipEvent_t *ev = nullptr)
{
    double *cur = s->nlbuf[parity].as<double>();
    double *prev = boot ? cur : s->nlbuf[parity ^ 1].as<double>();
    hipStream_t st = s->ctx->stream;
    const bool side_ok = s->chain && tangential(s) && !ev && s->side;
    const bool split = side_ok && s->overlap == 1 && !s->tbc, ana_side = side_ok && s->overlap >= 1;
    ChainArgs c{};
    if (s->chain) chain_args(s, parity, boot, c);
    const bool fp64 = s->d.prec == GD_PREC_FP64;
    if (split) {
        GD_HIP(hipEventRecord(s->ev_fork, st));
        GD_HIP(hipStreamWaitEvent(s->side, s->ev_fork, 0));
        ChainArgs cb = c; // B_tor, B_pol
        cb.f0 = 2;
        cb.nf = 2;
        GD_TRY(gd_launch_cnab2_chain(s->side, cb, fp64));
    }
    GD_TRY(nonlinear(s, cur, ev, ana_side));
    if (s->chain) {
        if (split) {
            ChainArgs cu = c, ct = c; // u_tor, u_pol; T
            cu.f0 = 0;
            cu.nf = 2;
            ct.f0 = 4;
            ct.nf = 1;
            GD_TRY(gd_launch_cnab2_chain(st, cu, fp64));
            if (s->comp) {
                ChainArgs cc = c;
                cc.f0 = 5;
                cc.nf = 1;
                GD_TRY(composition_bc(s, parity, st));
                GD_TRY(gd_launch_cnab2_chain(st, cc, fp64));
            }
            GD_TRY(gd_launch_cnab2_chain(s->side, ct, fp64));
            GD_HIP(hipEventRecord(s->ev_join, s->side));
            GD_HIP(hipStreamWaitEvent(st, s->ev_join, 0));
        } else {
            if (ana_side) { // the scalar analysis ran on the side stream
                GD_HIP(hipEventRecord(s->ev_join, s->side));
                GD_HIP(hipStreamWaitEvent(st, s->ev_join, 0));
            }
            GD_TRY(temperature_bc(s, parity, st));
            GD_TRY(composition_bc(s, parity, st));
            GD_TRY(gd_launch_cnab2_chain(st, c, fp64));
        }
        s->syn_ok = true; // synV/synS now describe state[parity ^ 1]
        if (s->comp) s->csyn_ok = true;
    } else {
        GD_TRY(temperature_bc(s, parity, st));
        GD_TRY(gd_launch_cnab2_step(st, s->nuLap.as<double>(), s->rtab.as<double>() + s->d.nr,
                                    s->nus.as<double>(), s->LUt.as<double>(), s->sht.d_l_of_idx.as<int>(), s->nl,
                                    s->d.nr, s->d.bw, s->nlm, s->d.dt, s->d.theta, bc_mask(s),
                                    s->state[parity].as<double>(), s->state[parity ^ 1].as<double>(), cur, prev));
        s->syn_ok = false;
    }
    GD_MARK(7);
    return GD_OK;
}

static void destroy_graphs(gd_step *s)
{
    for (int k = 0; k < 2; ++k) {
        if (s->gexec[k]) (void)hipGraphExecDestroy(s->gexec[k]);
        if (s->graph[k]) (void)hipGraphDestroy(s->graph[k]);
        s->gexec[k] = nullptr;
        s->graph[k] = nullptr;
    }
    ++s->gen;
}

int gd_step_enqueue(gd_step *s, int parity, bool boot) { return enqueue_step(s, parity, boot); }
bool gd_step_tangential(const gd_step *s) { return tangential(s); }
void gd_step_drop_graphs(gd_step *s) { destroy_graphs(s); }

static int step_release(gd_step *s)
{
    destroy_graphs(s);
    gd_adj_work_release(s);
    DevBuf *bufs[] = {&s->state[0], &s->state[1], &s->nlbuf[0], &s->nlbuf[1], &s->Lt, &s->LUt, &s->ULt, &s->rdiag2, &s->Rint, &s->nuLap, &s->nus,
                      &s->D1, &s->D2, &s->rtab, &s->synV,
                      &s->synS, &s->fb, &s->gb, &s->gb2, &s->part, &s->eout, &s->scratch, &s->status, &s->zrow, &s->rdiag};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t *e : {&s->ev_fork, &s->ev_ring, &s->ev_join})
        if (*e) (void)hipEventDestroy(*e);
    if (s->tbc) (void)gd_scalar_bc_destroy(s->tbc);
    s->tbc = nullptr;
    if (s->cbc) (void)gd_scalar_bc_destroy(s->cbc);
    s->cbc = nullptr;
    for (DevBuf *b : {&s->cstate[0], &s->cstate[1], &s->cnl[0], &s->cnl[1]})
        b->release();
    if (s->side) (void)hipStreamDestroy(s->side);
    s->side = nullptr;
    gd_sht_plan_release(&s->sht);
    return GD_OK;
}

template <typename T> static int upload_vec(gd_ctx *ctx, DevBuf &b, const std::vector<T> &v)
{
    GD_TRY(b.ensure(sizeof(T) * v.size()));
    GD_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, ctx->stream));
    return GD_OK;
}

// the twisted-solve tables of diffusivity slot k from the operator just built
// in S (band) and LU (band) on the device
static int twist_slot(gd_step *s, int k, const DevBuf &S, const DevBuf &LU)
{
    if (!s->twist) return GD_OK;
    const int nr = s->d.nr, bw = s->d.bw, nl = s->nl, m2 = 2 * bw;
    const size_t nb = (size_t)s->W * nr * nl;
    std::vector<double> hs(nb), hl(nb), ul(nb), rd((size_t)nr * nl), R((size_t)nl * m2 * m2);
    GD_HIP(hipStreamSynchronize(s->ctx->stream));
    GD_HIP(hipMemcpy(hs.data(), S.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    GD_HIP(hipMemcpy(hl.data(), LU.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    GD_TRY(gd_chain_twist_tables(hs.data(), hl.data(), nr, bw, nl, ul.data(), rd.data(), R.data()));
    GD_HIP(hipMemcpy(s->ULt.as<double>() + k * nb, ul.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
    GD_HIP(hipMemcpy(s->rdiag2.as<double>() + (size_t)k * nr * nl, rd.data(), sizeof(double) * rd.size(),
                     hipMemcpyHostToDevice));
    GD_HIP(hipMemcpy(s->Rint.as<double>() + (size_t)k * nl * m2 * m2, R.data(), sizeof(double) * R.size(),
                     hipMemcpyHostToDevice));
    return GD_OK;
}

static int step_build(gd_ctx *ctx, const gd_step_desc *d, gd_step *s)
{
    s->ctx = ctx;
    s->d = *d;
    GD_CHECK(d->nr >= 2 && d->bw >= 1 && d->bw <= 7 && d->dt > 0 && d->E != 0.0 && d->Pr != 0.0, GD_ERR_ARG,
             "gd_step_create: bad parameters");
    GD_CHECK(d->prec == GD_PREC_FP64 || d->prec == GD_PREC_FP32 || d->prec == GD_PREC_BF16X3, GD_ERR_ARG,
             "gd_step_create: precision %d not available", d->prec);
    GD_HIP(hipSetDevice(ctx->device));
    GD_TRY(gd_sht_plan_build(ctx, d->lmax, d->mmax, d->nlat, d->nlon, d->prec, &s->sht));
    const int nr = d->nr, bw = d->bw;
    s->nlm = s->sht.nlm;
    s->nl = d->lmax + 1;
    s->W = 2 * bw + 1;
    s->fs = (size_t)s->nlm * nr;
    // radial grid and FD operators (host FP64, fields.jl:108-131, linear_algebra.jl:90-151)
    std::vector<double> r((size_t)nr * 7), d1((size_t)s->W * nr), d2((size_t)s->W * nr), lap((size_t)s->W * nr);
    GD_TRY(gd_radial_grid(nr, d->rratio, r.data()));
    GD_TRY(gd_derivative_matrix(1, nr, bw, r.data(), d->fd_mode, d1.data()));
    GD_TRY(gd_derivative_matrix(2, nr, bw, r.data(), d->fd_mode, d2.data()));
    GD_TRY(gd_radial_laplacian(nr, bw, r.data(), d->fd_mode, lap.data()));
    GD_TRY(upload_vec(ctx, s->rtab, r));
    GD_TRY(upload_vec(ctx, s->D1, d1));
    GD_TRY(upload_vec(ctx, s->D2, d2));
    // implicit operators per diffusivity (simulation.jl:229-231), FP64 on the GPU
    const size_t nb = (size_t)s->W * nr * s->nl;
    // room for four diffusivities: velocity, magnetic, temperature and (once
    // enabled) composition
    GD_TRY(s->Lt.ensure(sizeof(double) * nb * 4));
    GD_TRY(s->LUt.ensure(sizeof(double) * nb * 4));
    DevBuf dlap, tmpL, tmpS, tmpLU;
    int rc = upload_vec(ctx, dlap, lap);
    if (rc == GD_OK) rc = tmpL.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = tmpS.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = tmpLU.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = s->status.ensure(sizeof(int));
    if (rc == GD_OK && hipMemsetAsync(s->status.p, 0, sizeof(int), ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    const double nus[3] = {d->Pm, 1.0, d->Pm / d->Pr};
    {
        std::vector<double> nl3(nus, nus + 3), nlap((size_t)4 * s->W * nr, 0.0);
        nl3.push_back(0.0);
        for (int k = 0; k < 3; ++k)
            for (size_t q = 0; q < (size_t)s->W * nr; ++q) nlap[k * s->W * nr + q] = nus[k] * lap[q];
        if (rc == GD_OK) rc = upload_vec(ctx, s->nus, nl3);
        if (rc == GD_OK) rc = upload_vec(ctx, s->nuLap, nlap);
    }
    GD_CHECK(d->bc_rows >= 0 && d->bc_rows <= 2, GD_ERR_ARG, "gd_step_create: bc_rows=%d", d->bc_rows);
    GD_CHECK(d->radial_force == 0 || d->radial_force == 1, GD_ERR_ARG, "gd_step_create: radial_force=%d",
             d->radial_force);
    // the twisted column solve with bw <= 3, every precision (at lmax=127,
    // nr=65, bw=3, paired columns: FP64-synthesis path 81.3 -> 68.8 us against
    // one lane per column); GD_CHAIN_TWIST = 0 keeps one lane per column
    const char *tw = getenv("GD_CHAIN_TWIST");
    s->twist = (!(tw && tw[0] == '0') && bw <= 3 && nr >= 2 * bw + 2 &&
                gd_cnab2_chain_supported(nr, bw)) ? 1 : 0;
    if (s->twist && rc == GD_OK) {
        rc = s->ULt.ensure(sizeof(double) * nb * 4);
        if (rc == GD_OK) rc = s->rdiag2.ensure(sizeof(double) * 4 * nr * s->nl);
        if (rc == GD_OK) rc = s->Rint.ensure(sizeof(double) * 4 * s->nl * 4 * bw * bw);
    }
    for (int k = 0; k < 3 && rc == GD_OK; ++k) {
        rc = gd_cnab2_build(ctx->stream, dlap.as<double>(), s->rtab.as<double>() + nr, nus[k], d->dt, d->theta, s->nl,
                            nr, bw, d->bc_rows == 1 || (d->bc_rows == 2 && k == 2), tmpL.as<double>(), tmpS.as<double>(), tmpLU.as<double>(),
                            s->Lt.as<double>() + k * nb, s->LUt.as<double>() + k * nb, s->status.as<int>());
        if (rc == GD_OK) rc = twist_slot(s, k, tmpS, tmpLU);
    }
    int stv = 0;
    if (rc == GD_OK && hipMemcpyAsync(&stv, s->status.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = GD_ERR_HIP;
    if (rc == GD_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = GD_ERR_HIP;
    dlap.release();
    tmpL.release();
    tmpS.release();
    tmpLU.release();
    GD_TRY(rc);
    GD_CHECK(stv == 0, GD_ERR_NUMERIC, "gd_step_create: zero or non-finite pivot in LU(S_l)");
    // state and work buffers
    const size_t esz = d->prec == GD_PREC_FP64 ? sizeof(double) : sizeof(float);
    const int nm = d->mmax + 1;
    GD_TRY(s->state[0].ensure(sizeof(double) * s->fs * 10));
    GD_TRY(s->state[1].ensure(sizeof(double) * s->fs * 10));
    GD_TRY(s->nlbuf[0].ensure(sizeof(double) * s->fs * 10));
    GD_TRY(s->nlbuf[1].ensure(sizeof(double) * s->fs * 10));
    GD_TRY(s->synV.ensure(esz * s->fs * SYNV_COLS));
    GD_TRY(s->synS.ensure(esz * s->fs * 10));
    GD_TRY(s->fb.ensure(esz * (size_t)s->sht.nlat2 * nr * nm * 60));
    GD_TRY(s->gb.ensure(esz * (size_t)nm * s->sht.nlat2 * nr * 16));
    GD_TRY(s->gb2.ensure(esz * (size_t)nm * s->sht.nlat2 * nr * 8));
    GD_TRY(s->part.ensure(sizeof(double) * 3 * 512));
    GD_TRY(s->eout.ensure(sizeof(double) * 4));
    GD_TRY(s->zrow.ensure(sizeof(double) * s->nlm));
    GD_TRY(s->rdiag.ensure(sizeof(double) * 4 * nr * s->nl));
    GD_TRY(gd_launch_recip_diag(ctx->stream, s->LUt.as<double>(), 3, nr, bw, s->nl, s->rdiag.as<double>()));
    GD_HIP(hipMemsetAsync(s->zrow.p, 0, sizeof(double) * s->nlm, ctx->stream));
    GD_HIP(hipMemsetAsync(s->state[0].p, 0, sizeof(double) * s->fs * 10, ctx->stream));
    GD_HIP(hipMemsetAsync(s->state[1].p, 0, sizeof(double) * s->fs * 10, ctx->stream));
    GD_HIP(hipMemsetAsync(s->nlbuf[0].p, 0, sizeof(double) * s->fs * 10, ctx->stream));
    GD_HIP(hipMemsetAsync(s->nlbuf[1].p, 0, sizeof(double) * s->fs * 10, ctx->stream));
    GD_HIP(hipMemsetAsync(s->synS.p, 0, esz * s->fs * 10, ctx->stream));
    GD_HIP(hipMemsetAsync(s->fb.p, 0, esz * (size_t)s->sht.nlat2 * nr * nm * 60, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    s->step = 0;
    s->chain = gd_cnab2_chain_supported(nr, bw) != 0;
    s->syn_ok = false;
    GD_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    for (hipEvent_t *e : {&s->ev_fork, &s->ev_ring, &s->ev_join})
        GD_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    // GD_STEP_OVERLAP = 0 / 1 / 2 overrides the stream layout (measurement aid)
    s->overlap = d->prec != GD_PREC_FP64 ? 2 : 0;
    const char *ov = getenv("GD_STEP_OVERLAP");
    if (ov && ov[0] >= '0' && ov[0] <= '2') s->overlap = ov[0] - '0';
    // GD_SYNTH_GRAD = 0 contracts the zero toroidal columns of the gradient
    // fields as well (the earlier form; the results are the same bits)
    const char *sg = getenv("GD_SYNTH_GRAD");
    s->synth_grad = !(sg && sg[0] == '0');
    // GD_CHAIN_OFF64 = 1 runs the CNAB2 column kernel with 64-bit offsets
    const char *o64 = getenv("GD_CHAIN_OFF64");
    s->chain_off64 = (o64 && o64[0] == '1') ? 1 : 0;
    // GD_CHAIN_PAIR = 0 solves one column per lane in the twisted kernel too;
    // 2 or 3: re and im per lane with that prefetch depth (measurement aid)
    const char *cp = getenv("GD_CHAIN_PAIR");
    s->chain_pair = (cp && (cp[0] == '0' || cp[0] == '3')) ? cp[0] - '0' : 2;
    // GD_CHAIN_YLDS: forward results y of the paired twisted solve in LDS --
    // unset or -1 as many rows as keep the launch in one round of resident
    // blocks, k > 0 at most k rows per half, 0 all through st_out
    const char *yl = getenv("GD_CHAIN_YLDS");
    s->chain_ylds = yl ? atoi(yl) : -1;
    return GD_OK;
}

extern "C" int gd_step_create(gd_ctx *ctx, const gd_step_desc *desc, gd_step **out)
{
    GD_CHECK(ctx && desc && out, GD_ERR_ARG, "gd_step_create: null argument");
    gd_step *s = new gd_step();
    int rc = step_build(ctx, desc, s);
    if (rc != GD_OK) {
        step_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return GD_OK;
}

extern "C" int gd_step_destroy(gd_step *s)
{
    if (!s) return GD_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->side) (void)hipStreamSynchronize(s->side);
    step_release(s);
    delete s;
    return GD_OK;
}

// l-major user layout <-> m-major device planes
static int convert_state(gd_step *s, const double *re, const double *im, double *dev, bool to_device)
{
    const int nr = s->d.nr, nlm = s->nlm;
    const int *map = to_device ? s->sht.d_lmaj_of_idx.as<int>() : s->sht.d_idx_of_lmaj.as<int>();
    for (int f = 0; f < 5; ++f) {
        const size_t o = (size_t)f * s->fs;
        if (to_device) {
            GD_TRY(gd_launch_perm(s->ctx->stream, re + o, map, nlm, nr, dev + (2 * f) * s->fs));
            GD_TRY(gd_launch_perm(s->ctx->stream, im + o, map, nlm, nr, dev + (2 * f + 1) * s->fs));
        } else {
            GD_TRY(gd_launch_perm(s->ctx->stream, dev + (2 * f) * s->fs, map, nlm, nr, const_cast<double *>(re) + o));
            GD_TRY(gd_launch_perm(s->ctx->stream, dev + (2 * f + 1) * s->fs, map, nlm, nr, const_cast<double *>(im) + o));
        }
    }
    return GD_OK;
}

extern "C" int gd_step_set_state(gd_step *s, const double *re, const double *im, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_set_state: null argument");
    GD_HIP(hipSetDevice(s->ctx->device));
    const size_t n = s->fs * 5;
    const double *a = re, *b = im;
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * n * 2));
        GD_HIP(hipMemcpyAsync(s->scratch.p, re, sizeof(double) * n, hipMemcpyHostToDevice, s->ctx->stream));
        GD_HIP(hipMemcpyAsync(s->scratch.as<double>() + n, im, sizeof(double) * n, hipMemcpyHostToDevice,
                              s->ctx->stream));
        a = s->scratch.as<double>();
        b = a + n;
    }
    if (s->comp && (s->step & 1)) { // the composition state moves with the step parity
        const size_t pb = sizeof(double) * s->fs * 2;
        GD_HIP(hipMemcpyAsync(s->cstate[0].as<double>() + 8 * s->fs, s->cstate[1].as<double>() + 8 * s->fs, pb,
                              hipMemcpyDeviceToDevice, s->ctx->stream));
    }
    s->step = 0;
    s->syn_ok = false;
    s->csyn_ok = false;
    GD_TRY(convert_state(s, a, b, cur_state(s), true));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(s->ctx->stream));
    return GD_OK;
}

static int get_planes(gd_step *s, const double *dev, double *re, double *im, int flags)
{
    GD_HIP(hipSetDevice(s->ctx->device));
    const size_t n = s->fs * 5;
    double *a = re, *b = im;
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * n * 2));
        a = s->scratch.as<double>();
        b = a + n;
    }
    GD_TRY(convert_state(s, a, b, const_cast<double *>(dev), false));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(s->ctx, re, a, sizeof(double) * n));
        GD_TRY(gd_from_device(s->ctx, im, b, sizeof(double) * n));
    }
    return GD_OK;
}

extern "C" int gd_step_get_state(gd_step *s, double *re, double *im, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_get_state: null argument");
    return get_planes(s, cur_state(s), re, im, flags);
}

// the AB2 history: the nonlinear terms of the previous step, which the next
// step combines with its own (simulation.jl:1312-1329 keeps them as
// prev_nonlinear)
static double *history(gd_step *s) { return s->nlbuf[(s->step & 1) ^ 1].as<double>(); }

extern "C" int gd_step_get_history(gd_step *s, double *nl_re, double *nl_im, int flags)
{
    GD_CHECK(s && nl_re && nl_im, GD_ERR_ARG, "gd_step_get_history: null argument");
    GD_CHECK(s->step > 0, GD_ERR_ARG, "gd_step_get_history: no nonlinear history before the first step");
    return get_planes(s, history(s), nl_re, nl_im, flags);
}

extern "C" int gd_step_restore(gd_step *s, const double *re, const double *im, const double *nl_re,
                               const double *nl_im, int step, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_restore: null argument");
    GD_CHECK(step >= 0, GD_ERR_ARG, "gd_step_restore: step %d < 0", step);
    GD_CHECK(step == 0 || (nl_re && nl_im), GD_ERR_ARG, "gd_step_restore: step %d > 0 needs the nonlinear history",
             step);
    GD_HIP(hipSetDevice(s->ctx->device));
    const size_t n = s->fs * 5;
    const int nplanes = step > 0 ? 4 : 2;
    const double *src[4] = {re, im, nl_re, nl_im};
    const double *dev[4] = {re, im, nl_re, nl_im};
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * n * nplanes));
        for (int k = 0; k < nplanes; ++k) {
            dev[k] = s->scratch.as<double>() + k * n;
            GD_HIP(hipMemcpyAsync(const_cast<double *>(dev[k]), src[k], sizeof(double) * n, hipMemcpyHostToDevice,
                                  s->ctx->stream));
        }
    }
    if (s->comp && ((s->step ^ step) & 1)) {
        // the composition state moves with the step parity (as in
        // gd_step_set_state); gd_step_restore_composition normally follows
        const size_t pb = sizeof(double) * s->fs * 2;
        GD_HIP(hipMemcpyAsync(s->cstate[step & 1].as<double>() + 8 * s->fs,
                              s->cstate[s->step & 1].as<double>() + 8 * s->fs, pb, hipMemcpyDeviceToDevice,
                              s->ctx->stream));
    }
    s->step = step;
    s->syn_ok = false;
    s->csyn_ok = false;
    GD_TRY(convert_state(s, dev[0], dev[1], cur_state(s), true));
    if (step > 0) GD_TRY(convert_state(s, dev[2], dev[3], history(s), true));
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(s->ctx->stream));
    return GD_OK;
}

extern "C" int gd_step_nonlinear(gd_step *s, double *nl_re, double *nl_im, int flags)
{
    GD_CHECK(s && nl_re && nl_im, GD_ERR_ARG, "gd_step_nonlinear: null argument");
    GD_HIP(hipSetDevice(s->ctx->device));
    // the buffer the next step would fill; the history buffer is untouched
    double *cur = s->nlbuf[s->step & 1].as<double>();
    GD_TRY(nonlinear(s, cur));
    return get_planes(s, cur, nl_re, nl_im, flags);
}

extern "C" int gd_step_run(gd_step *s, int nsteps)
{
    GD_CHECK(s && nsteps >= 0, GD_ERR_ARG, "gd_step_run: bad args");
    GD_HIP(hipSetDevice(s->ctx->device));
    for (int k = 0; k < nsteps; ++k) {
        GD_TRY(enqueue_step(s, s->step & 1, s->step == 0));
        ++s->step;
    }
    return GD_OK;
}

static int capture(gd_step *s, int parity)
{
    hipStream_t st = s->ctx->stream;
    GD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(s, parity, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    GD_TRY(rc);
    if (e != hipSuccess) {
        gd_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return GD_ERR_HIP;
    }
    s->graph[parity] = g;
    GD_HIP(hipGraphInstantiate(&s->gexec[parity], g, nullptr, nullptr, 0));
    return GD_OK;
}

extern "C" int gd_step_run_graph(gd_step *s, int nsteps)
{
    GD_CHECK(s && nsteps >= 0, GD_ERR_ARG, "gd_step_run_graph: bad args");
    GD_HIP(hipSetDevice(s->ctx->device));
    for (int k = 0; k < nsteps; ++k) {
        if (s->step == 0) {
            GD_TRY(enqueue_step(s, 0, true)); // AB1 bootstrap is not part of the graphs
        } else {
            const int par = s->step & 1;
            GD_TRY(ensure_syn(s));
            if (!s->gexec[par]) GD_TRY(capture(s, par));
            GD_HIP(hipGraphLaunch(s->gexec[par], s->ctx->stream));
        }
        ++s->step;
    }
    return GD_OK;
}

extern "C" int gd_step_count(const gd_step *s) { return s ? s->step : -1; }

extern "C" int gd_step_shape(const gd_step *s, int *lmax, int *mmax, int *nr)
{
    GD_CHECK(s && lmax && mmax && nr, GD_ERR_ARG, "gd_step_shape: null argument");
    *lmax = s->d.lmax;
    *mmax = s->d.mmax;
    *nr = s->d.nr;
    return GD_OK;
}

// Per-mode temperature conditions for the step (the reference's
// bc_type_inner / bc_type_outer / boundary_values of the temperature field,
// applied by apply_temperature_boundary_conditions_spectral!, thermal.jl:262-313,
// between the nonlinear terms and the implicit step); null arrays return to
// the default (boundary rows read as 0).  Captured graphs are dropped.
extern "C" int gd_step_set_temperature_bc(gd_step *s, const int *bc_inner, const int *bc_outer, const double *values)
{
    GD_CHECK(s, GD_ERR_ARG, "gd_step_set_temperature_bc: null step");
    GD_CHECK((bc_inner && bc_outer && values) || (!bc_inner && !bc_outer && !values), GD_ERR_ARG,
             "gd_step_set_temperature_bc: give all three arrays or none");
    GD_HIP(hipSetDevice(s->ctx->device));
    GD_HIP(hipStreamSynchronize(s->ctx->stream));
    destroy_graphs(s);
    if (!bc_inner) {
        if (s->tbc) (void)gd_scalar_bc_destroy(s->tbc);
        s->tbc = nullptr;
        return GD_OK;
    }
    if (!s->tbc) {
        std::vector<double> r((size_t)s->d.nr * 7);
        GD_TRY(gd_radial_grid(s->d.nr, s->d.rratio, r.data()));
        GD_TRY(gd_scalar_bc_create(s->ctx, s->d.lmax, s->d.mmax, s->d.nr, s->d.bw, r.data(), s->d.fd_mode, &s->tbc));
    }
    // the step's planes are in the device (m-major) mode order
    return gd_scalar_bc_set_perm(s->tbc, bc_inner, bc_outer, values, s->sht.maps.lmaj_of_idx.data());
}

// Composition (compositional.jl): a passive scalar advected by u,
// -u.grad C through its own gradient synthesis, ring pass and scalar
// analysis, CNAB2 with diffusivity Pm/Sc (simulation.jl:130-131), conditions
// applied before its solve (default: no flux at both boundaries, value 0,
// compositional.jl:110-115).  Tangential step with the column kernel only.
extern "C" int gd_step_enable_composition(gd_step *s, double Sc)
{
    GD_CHECK(s && Sc > 0.0 && std::isfinite(Sc), GD_ERR_ARG, "gd_step_enable_composition: bad arguments");
    GD_CHECK(tangential(s), GD_ERR_ARG,
             "gd_step_enable_composition: built for the tangential step (radials = 0, full_products = 0)");
    GD_CHECK(s->chain, GD_ERR_ARG, "gd_step_enable_composition: nr=%d bw=%d has no column CNAB2 kernel", s->d.nr,
             s->d.bw);
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    GD_HIP(hipStreamSynchronize(st));
    destroy_graphs(s);
    const gd_step_desc *d = &s->d;
    const int nr = d->nr, bw = d->bw;
    const size_t nb = (size_t)s->W * nr * s->nl;
    // operator of slot 3
    std::vector<double> r((size_t)nr * 7), lap((size_t)s->W * nr);
    GD_TRY(gd_radial_grid(nr, d->rratio, r.data()));
    GD_TRY(gd_radial_laplacian(nr, bw, r.data(), d->fd_mode, lap.data()));
    const double nu = d->Pm / Sc;
    std::vector<double> nlap(lap.size());
    for (size_t q = 0; q < lap.size(); ++q) nlap[q] = nu * lap[q];
    GD_HIP(hipMemcpy(s->nus.as<double>() + 3, &nu, sizeof(double), hipMemcpyHostToDevice));
    GD_HIP(hipMemcpy(s->nuLap.as<double>() + (size_t)3 * s->W * nr, nlap.data(), sizeof(double) * nlap.size(),
                     hipMemcpyHostToDevice));
    DevBuf dlap, tmpL, tmpS, tmpLU;
    int rc = upload_vec(s->ctx, dlap, lap);
    if (rc == GD_OK) rc = tmpL.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = tmpS.ensure(sizeof(double) * nb);
    if (rc == GD_OK) rc = tmpLU.ensure(sizeof(double) * nb);
    if (rc == GD_OK && hipMemsetAsync(s->status.p, 0, sizeof(int), st) != hipSuccess) rc = GD_ERR_HIP;
    if (rc == GD_OK)
        rc = gd_cnab2_build(st, dlap.as<double>(), s->rtab.as<double>() + nr, nu, d->dt, d->theta, s->nl, nr, bw,
                            d->bc_rows == 1, tmpL.as<double>(), tmpS.as<double>(), tmpLU.as<double>(),
                            s->Lt.as<double>() + 3 * nb, s->LUt.as<double>() + 3 * nb, s->status.as<int>());
    if (rc == GD_OK) rc = twist_slot(s, 3, tmpS, tmpLU);
    if (rc == GD_OK) rc = gd_launch_recip_diag(st, s->LUt.as<double>(), 4, nr, bw, s->nl, s->rdiag.as<double>());
    int stv = 0;
    if (rc == GD_OK && hipMemcpyAsync(&stv, s->status.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = GD_ERR_HIP;
    if (rc == GD_OK && hipStreamSynchronize(st) != hipSuccess) rc = GD_ERR_HIP;
    dlap.release();
    tmpL.release();
    tmpS.release();
    tmpLU.release();
    GD_TRY(rc);
    GD_CHECK(stv == 0, GD_ERR_NUMERIC, "gd_step_enable_composition: zero or non-finite pivot in LU(S_l)");
    // buffers (zeroed: the planes other than 8, 9 must stay zero)
    const size_t esz = d->prec == GD_PREC_FP64 ? sizeof(double) : sizeof(float);
    for (DevBuf *b : {&s->cstate[0], &s->cstate[1], &s->cnl[0], &s->cnl[1]}) {
        GD_TRY(b->ensure(sizeof(double) * s->fs * 10));
        GD_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * s->fs * 10, st));
    }
    if (!s->cbc) GD_TRY(gd_scalar_bc_create(s->ctx, d->lmax, d->mmax, nr, bw, r.data(), d->fd_mode, &s->cbc));
    std::vector<int> two((size_t)s->nlm, 2);
    std::vector<double> zero((size_t)2 * s->nlm, 0.0);
    GD_TRY(gd_scalar_bc_set_perm(s->cbc, two.data(), two.data(), zero.data(), s->sht.maps.lmaj_of_idx.data()));
    GD_HIP(hipStreamSynchronize(st));
    s->Sc = Sc;
    s->comp = true;
    s->csyn_ok = false;
    return GD_OK;
}

// composition planes (nlm, nr) re/im in the reference's l-major layout
extern "C" int gd_step_set_composition(gd_step *s, const double *re, const double *im, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_set_composition: null argument");
    GD_CHECK(s->comp, GD_ERR_ARG, "gd_step_set_composition: composition not enabled");
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const double *a = re, *b = im;
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * s->fs * 2));
        GD_HIP(hipMemcpyAsync(s->scratch.p, re, sizeof(double) * s->fs, hipMemcpyHostToDevice, st));
        GD_HIP(hipMemcpyAsync(s->scratch.as<double>() + s->fs, im, sizeof(double) * s->fs, hipMemcpyHostToDevice,
                              st));
        a = s->scratch.as<double>();
        b = a + s->fs;
    }
    double *dst = s->cstate[s->step & 1].as<double>();
    const int *map = s->sht.d_lmaj_of_idx.as<int>();
    GD_TRY(gd_launch_perm(st, a, map, s->nlm, s->d.nr, dst + 8 * s->fs));
    GD_TRY(gd_launch_perm(st, b, map, s->nlm, s->d.nr, dst + 9 * s->fs));
    s->csyn_ok = false;
    if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(st));
    return GD_OK;
}

extern "C" int gd_step_get_composition(gd_step *s, double *re, double *im, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_get_composition: null argument");
    GD_CHECK(s->comp, GD_ERR_ARG, "gd_step_get_composition: composition not enabled");
    GD_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    double *a = re, *b = im;
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * s->fs * 2));
        a = s->scratch.as<double>();
        b = a + s->fs;
    }
    const double *src = s->cstate[s->step & 1].as<double>();
    const int *map = s->sht.d_idx_of_lmaj.as<int>();
    GD_TRY(gd_launch_perm(st, src + 8 * s->fs, map, s->nlm, s->d.nr, a));
    GD_TRY(gd_launch_perm(st, src + 9 * s->fs, map, s->nlm, s->d.nr, b));
    if (flags & GD_HOST) {
        GD_TRY(gd_from_device(s->ctx, re, a, sizeof(double) * s->fs));
        GD_TRY(gd_from_device(s->ctx, im, b, sizeof(double) * s->fs));
    }
    return GD_OK;
}

extern "C" int gd_step_has_composition(const gd_step *s) { return s && s->comp ? 1 : 0; }

// planes 8, 9 of a composition buffer <-> (nlm, nr) l-major host or device planes
static int comp_planes(gd_step *s, double *buf, double *re, double *im, int flags, bool to_device)
{
    hipStream_t st = s->ctx->stream;
    double *a = re, *b = im;
    if (flags & GD_HOST) {
        GD_TRY(s->scratch.ensure(sizeof(double) * s->fs * 2));
        a = s->scratch.as<double>();
        b = a + s->fs;
        if (to_device) {
            GD_HIP(hipMemcpyAsync(a, re, sizeof(double) * s->fs, hipMemcpyHostToDevice, st));
            GD_HIP(hipMemcpyAsync(b, im, sizeof(double) * s->fs, hipMemcpyHostToDevice, st));
        }
    }
    if (to_device) {
        const int *map = s->sht.d_lmaj_of_idx.as<int>();
        GD_TRY(gd_launch_perm(st, a, map, s->nlm, s->d.nr, buf + 8 * s->fs));
        GD_TRY(gd_launch_perm(st, b, map, s->nlm, s->d.nr, buf + 9 * s->fs));
        if (flags & GD_HOST) GD_HIP(hipStreamSynchronize(st));
    } else {
        const int *map = s->sht.d_idx_of_lmaj.as<int>();
        GD_TRY(gd_launch_perm(st, buf + 8 * s->fs, map, s->nlm, s->d.nr, a));
        GD_TRY(gd_launch_perm(st, buf + 9 * s->fs, map, s->nlm, s->d.nr, b));
        if (flags & GD_HOST) {
            GD_TRY(gd_from_device(s->ctx, re, a, sizeof(double) * s->fs));
            GD_TRY(gd_from_device(s->ctx, im, b, sizeof(double) * s->fs));
        }
    }
    return GD_OK;
}

extern "C" int gd_step_get_composition_history(gd_step *s, double *nl_re, double *nl_im, int flags)
{
    GD_CHECK(s && nl_re && nl_im, GD_ERR_ARG, "gd_step_get_composition_history: null argument");
    GD_CHECK(s->comp, GD_ERR_ARG, "gd_step_get_composition_history: composition not enabled");
    GD_CHECK(s->step > 0, GD_ERR_ARG, "gd_step_get_composition_history: no history before the first step");
    GD_HIP(hipSetDevice(s->ctx->device));
    return comp_planes(s, s->cnl[(s->step & 1) ^ 1].as<double>(), nl_re, nl_im, flags, false);
}

// after gd_step_restore: the composition at the restored step and, for a step
// > 0, its AB2 history
extern "C" int gd_step_restore_composition(gd_step *s, const double *re, const double *im, const double *nl_re,
                                           const double *nl_im, int flags)
{
    GD_CHECK(s && re && im, GD_ERR_ARG, "gd_step_restore_composition: null argument");
    GD_CHECK(s->comp, GD_ERR_ARG, "gd_step_restore_composition: composition not enabled");
    GD_CHECK(s->step == 0 || (nl_re && nl_im), GD_ERR_ARG,
             "gd_step_restore_composition: step %d > 0 needs the composition's history", s->step);
    GD_HIP(hipSetDevice(s->ctx->device));
    GD_TRY(comp_planes(s, s->cstate[s->step & 1].as<double>(), const_cast<double *>(re), const_cast<double *>(im),
                       flags, true));
    if (s->step > 0)
        GD_TRY(comp_planes(s, s->cnl[(s->step & 1) ^ 1].as<double>(), const_cast<double *>(nl_re),
                           const_cast<double *>(nl_im), flags, true));
    s->csyn_ok = false;
    return GD_OK;
}

// the composition field's bc_type_inner / bc_type_outer / boundary_values
// (compositional.jl:48-50, 318-370), l-major; captured graphs are dropped
extern "C" int gd_step_set_composition_bc(gd_step *s, const int *bc_inner, const int *bc_outer, const double *values)
{
    GD_CHECK(s && bc_inner && bc_outer && values, GD_ERR_ARG, "gd_step_set_composition_bc: null argument");
    GD_CHECK(s->comp, GD_ERR_ARG, "gd_step_set_composition_bc: composition not enabled");
    GD_HIP(hipSetDevice(s->ctx->device));
    GD_HIP(hipStreamSynchronize(s->ctx->stream));
    destroy_graphs(s);
    return gd_scalar_bc_set_perm(s->cbc, bc_inner, bc_outer, values, s->sht.maps.lmaj_of_idx.data());
}

// the radial grid r_i of the step's shell (fields.jl:108-131), host out[nr]
extern "C" int gd_step_radii(const gd_step *s, double *r)
{
    GD_CHECK(s && r, GD_ERR_ARG, "gd_step_radii: null argument");
    const int nr = s->d.nr;
    std::vector<double> t((size_t)nr * 7);
    GD_TRY(gd_radial_grid(nr, s->d.rratio, t.data()));
    for (int i = 0; i < nr; ++i) r[i] = t[(size_t)nr * 3 + i];
    return GD_OK;
}

extern "C" int gd_step_energies(gd_step *s, double *out3)
{
    GD_CHECK(s && out3, GD_ERR_ARG, "gd_step_energies: null argument");
    GD_HIP(hipSetDevice(s->ctx->device));
    GD_TRY(gd_launch_energies(s->ctx->stream, cur_state(s), s->sht.d_l_of_idx.as<int>(), s->nlm, s->d.nr,
                              s->part.as<double>(), 512, s->eout.as<double>()));
    return gd_from_device(s->ctx, out3, s->eout.p, sizeof(double) * 3);
}

// Per-kernel-class device time with HIP events on the step's stream:
// out[k] = average ms per step of class k over nsteps eager steps, classes
// 0 spectral_pre, 1 synth vector, 2 synth scalar, 3 ring (FFT + products),
// 4 analysis vector, 5 analysis scalar, 6 CNAB2 rhs + banded solve.  Where
// the column-register CNAB2 kernel runs, it also prepares the next step's
// synthesis columns and class 0 is empty after the first step.
extern "C" int gd_step_profile(gd_step *s, int nsteps, double *out)
{
    GD_CHECK(s && out && nsteps >= 1, GD_ERR_ARG, "gd_step_profile: bad args");
    GD_HIP(hipSetDevice(s->ctx->device));
    hipEvent_t ev[8];
    for (int k = 0; k < 8; ++k) GD_HIP(hipEventCreate(&ev[k]));
    for (int k = 0; k < GD_NKCLASS; ++k) out[k] = 0.0;
    int rc = GD_OK;
    for (int it = 0; it < nsteps && rc == GD_OK; ++it) {
        rc = enqueue_step(s, s->step & 1, s->step == 0, ev);
        if (rc != GD_OK) break;
        ++s->step;
        if (hipEventSynchronize(ev[7]) != hipSuccess) {
            rc = GD_ERR_HIP;
            break;
        }
        for (int k = 0; k < GD_NKCLASS; ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) rc = GD_ERR_HIP;
            out[k] += ms / nsteps;
        }
    }
    for (int k = 0; k < 8; ++k) (void)hipEventDestroy(ev[k]);
    if (rc == GD_ERR_HIP) gd_set_error("gd_step_profile: event failure");
    return rc;
}

// ---------------------------------------------------------------- batches of independent shells
// Several steps (shells of any lmax, nr, grid) advanced together: one graph
// per parity holds every shell's step, the shells spread round-robin over the
// batch's streams (forked from and joined to the context stream), so small
// shells run side by side instead of one after another.  No counterpart in the
// reference (its shells are separate runs).
//
// Fused mode (GD_BATCH_FUSED): the graph of a parity holds one launch per
// kernel form instead of one branch per shell.  The tangential step's five
// stages -- synthesis, ring pass, vector analysis, scalar analysis, CNAB2
// column solves -- each run as one grid over the blocks of every shell that
// shares the stage's compile-time form (synthesis tile, ring length, node
// tile, band width), the shells of a grid in decreasing order of cost so the
// largest start first, every shell's blocks in the order of its own launch
// (so each output element gets the same operations as in the shell's own
// step: the results are the same bits).  A shell whose step has no fused form
// (composition, the full products, ...) runs its own step in the same graph.
// With several lanes the fused shells are split over that many of the batch
// streams (balanced by cost, largest first onto the least loaded lane), each
// lane running its own fused stages, so one lane's stage tail overlaps the
// others' work.
enum { BK_SYNTH = 0, BK_RING, BK_ANAV, BK_ANAS, BK_CHAIN, BK_NKIND };

struct BatchGroup {
    int lane = 0;
    int kind = 0, fp64 = 0, mm = 0, key = 0; // key: synthesis tile, ring length, node tile or band width
    ChainForm fm{};
    size_t lds = 0;
    int nblocks = 0;
    size_t off_args = 0, off_sob = 0, off_boff = 0, off_bcnt = 0; // in the parity's table
    std::vector<int> shells;                                      // batch indices, by decreasing cost
};

struct gd_batch {
    gd_ctx *ctx = nullptr;
    std::vector<gd_step *> steps;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // each member's generation when the graphs were captured: a member whose
    // captured work changed since (conditions set or dropped, composition
    // enabled, ...) makes the batch recapture before its next replay
    std::vector<unsigned> gens[2];
    int mode = GD_BATCH_STREAMS;
    int lanes = 1;                         // fused mode: streams the fused shells are split over
    std::vector<int> tiles;                // synthesis tile of each shell (fused mode)
    std::vector<BatchGroup> groups[2];     // fused launches of each parity
    std::vector<int> solo[2];              // shells that run their own step
    DevBuf table[2];                       // the groups' argument and block tables
};

static void batch_drop_graphs(gd_batch *b)
{
    for (int k = 0; k < 2; ++k) {
        if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->gexec[k] = nullptr;
        b->graph[k] = nullptr;
        b->gens[k].clear();
        // the fused tables are rebuilt with the graphs (they hold the
        // members' buffer addresses)
        b->groups[k].clear();
        b->solo[k].clear();
    }
}

static bool batch_stale(const gd_batch *b, int par)
{
    if (b->gens[par].size() != b->steps.size()) return true;
    for (size_t i = 0; i < b->steps.size(); ++i)
        if (b->steps[i]->gen != b->gens[par][i]) return true;
    return false;
}

static void batch_release(gd_batch *b)
{
    batch_drop_graphs(b);
    b->table[0].release();
    b->table[1].release();
    for (hipEvent_t e : b->joins)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipStream_t s : b->streams)
        if (s) (void)hipStreamDestroy(s);
}

extern "C" int gd_batch_create(gd_ctx *ctx, int n, gd_step *const *steps, int nstreams, gd_batch **out)
{
    GD_CHECK(ctx && steps && out && n >= 1 && nstreams >= 1 && nstreams <= 64, GD_ERR_ARG,
             "gd_batch_create: bad arguments");
    for (int i = 0; i < n; ++i) {
        GD_CHECK(steps[i] && steps[i]->ctx == ctx, GD_ERR_ARG, "gd_batch_create: step %d is null or on another context",
                 i);
        for (int j = 0; j < i; ++j)
            GD_CHECK(steps[j] != steps[i], GD_ERR_ARG, "gd_batch_create: step %d appears twice (also %d)", i, j);
    }
    GD_HIP(hipSetDevice(ctx->device));
    gd_batch *b = new gd_batch();
    b->ctx = ctx;
    b->steps.assign(steps, steps + n);
    b->tiles.assign(n, 0);
    hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int k = 0; k < nstreams && e == hipSuccess; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t j = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking); // default priority
        if (e == hipSuccess) b->streams.push_back(s);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(j);
    }
    if (e != hipSuccess) {
        gd_set_error("gd_batch_create: %s", hipGetErrorString(e));
        batch_release(b);
        delete b;
        return GD_ERR_HIP;
    }
    *out = b;
    return GD_OK;
}

extern "C" int gd_batch_destroy(gd_batch *b)
{
    if (!b) return GD_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    delete b;
    return GD_OK;
}

// ---- fused mode: the launch groups of one parity
struct ShellArgs {
    SynthArgs a;
    RingArgs r;
    AnaArgs gv, gs;
    ChainArgs c;
    ChainForm fm;
    bool fp64;
    int kn;
};

// the shell's fused-form arguments for a step of the given parity; false when
// its step has no fused form
static bool fused_args(gd_step *s, int parity, ShellArgs &x)
{
    if (!tangential(s) || !s->chain || s->comp || !s->synth_grad) return false;
    int mode;
    tang_args(s, s->nlbuf[parity].as<double>(), x.a, x.r, x.gv, x.gs, &mode);
    chain_args(s, parity, false, x.c);
    if (x.c.ylds > 0) {
        // a batched grid holds many rounds of blocks: LDS for the four
        // resident blocks per CU that the paired kernel's registers allow
        x.c.ylds = gd_chain_ylds(x.c, 40 * 1024);
        if (s->chain_ylds > 0) x.c.ylds = std::min(x.c.ylds, s->chain_ylds);
    }
    gd_cnab2_chain_form(x.c, &x.fm);
    x.fp64 = s->d.prec == GD_PREC_FP64;
    x.kn = gd_ana_kn(x.a.nlat2, x.fp64);
    if (mode != RING_STEP_TANG || x.kn == 0 || !x.fm.direct) return false;
    // the analyses in one node chunk of all nlat/2 nodes (the grid its own
    // launch has when nlat/2 <= the node tile)
    for (AnaArgs *g : {&x.gv, &x.gs}) {
        g->k0 = 0;
        g->nk = x.a.nlat2;
        g->accum = 0;
    }
    if (x.fm.lds_lane > 64 * 1024 || (x.fm.twist && x.fm.lds_twist > 64 * 1024)) return false;
    if (gd_synth_tang_batchable(x.a) != GD_OK || gd_ring_tang_batchable(x.r, x.fp64) != GD_OK) return false;
    return true;
}

static size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

static int fused_build(gd_batch *b, int parity)
{
    const int n = (int)b->steps.size();
    std::vector<ShellArgs> xs(n);
    std::vector<char> ok(n, 0);
    auto &groups = b->groups[parity];
    auto &solo = b->solo[parity];
    groups.clear();
    solo.clear();
    for (int i = 0; i < n; ++i) {
        ok[i] = fused_args(b->steps[i], parity, xs[i]);
        if (!ok[i]) solo.push_back(i);
    }
    // group key of each stage; costs for the order within a group
    // lanes: largest shells first onto the least loaded lane (cost: the
    // shell's Legendre work, which the transform stages scale with)
    std::vector<int> lane_of(n, 0);
    {
        std::vector<int> ord;
        for (int i = 0; i < n; ++i)
            if (ok[i]) ord.push_back(i);
        auto work = [&](int i) { return (double)xs[i].a.nlm * xs[i].a.nlat2 * xs[i].a.nr; };
        std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return work(p) > work(q); });
        std::vector<double> load(b->lanes, 0.0);
        for (int i : ord) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            lane_of[i] = l;
            load[l] += work(i);
        }
    }
    int cur_lane = 0;
    auto find = [&](int kind, int fp64, int mm, int key, const ChainForm *fm) -> BatchGroup & {
        for (auto &g : groups)
            if (g.lane == cur_lane && g.kind == kind && g.fp64 == fp64 && g.mm == mm && g.key == key &&
                (!fm || (g.fm.twist == fm->twist && g.fm.off32 == fm->off32 && g.fm.pair == fm->pair)))
                return g;
        groups.emplace_back();
        BatchGroup &g = groups.back();
        g.lane = cur_lane;
        g.kind = kind;
        g.fp64 = fp64;
        g.mm = mm;
        g.key = key;
        if (fm) g.fm = *fm;
        return g;
    };
    for (int i = 0; i < n; ++i) {
        if (!ok[i]) continue;
        const ShellArgs &x = xs[i];
        const int t = b->tiles[i];
        cur_lane = lane_of[i];
        GD_CHECK(gd_synth_tile_ok(t, x.fp64), GD_ERR_ARG, "gd_batch: shell %d: synthesis tile %d not available", i, t);
        find(BK_SYNTH, x.fp64, x.a.mm, t, nullptr).shells.push_back(i);
        find(BK_RING, x.fp64, 0, x.r.nlon, nullptr).shells.push_back(i);
        find(BK_ANAV, x.fp64, x.gv.mm, x.kn, nullptr).shells.push_back(i);
        find(BK_ANAS, x.fp64, x.gs.mm, x.kn, nullptr).shells.push_back(i);
        BatchGroup &gc = find(BK_CHAIN, x.fp64, 0, x.c.bw, &x.fm);
        gc.shells.push_back(i);
        gc.lds = std::max(gc.lds, x.fm.twist ? x.fm.lds_twist : x.fm.lds_lane);
    }
    auto blocks = [&](const BatchGroup &g, int i) -> int {
        const ShellArgs &x = xs[i];
        switch (g.kind) {
        case BK_SYNTH: return gd_synth_tang_blocks(x.a, g.key);
        case BK_RING: return x.r.nlat2 * x.r.nr;
        case BK_ANAV: return gd_ana_blocks(x.gv);
        case BK_ANAS: return gd_ana_blocks(x.gs);
        default: return gd_cnab2_chain_blocks(x.c);
        }
    };
    auto cost = [&](const BatchGroup &g, int i) -> double {
        const ShellArgs &x = xs[i];
        if (g.kind == BK_RING) return (double)x.r.nlat2 * x.r.nr * x.r.nlon * std::log2((double)x.r.nlon);
        if (g.kind == BK_CHAIN) return (double)x.c.nlm * x.c.n;
        return (double)x.a.nlm * x.a.nlat2 * x.a.nr;
    };
    // stage order, then the table: per group the args array, then the block
    // tables; each shell's blocks start at a multiple of 8
    std::stable_sort(groups.begin(), groups.end(), [](const BatchGroup &p, const BatchGroup &q) {
        return p.lane != q.lane ? p.lane < q.lane : p.kind < q.kind;
    });
    std::vector<char> blob;
    auto put = [&](const void *src, size_t bytes) -> size_t {
        const size_t o = align16(blob.size());
        blob.resize(o + bytes);
        std::memcpy(blob.data() + o, src, bytes);
        return o;
    };
    for (auto &g : groups) {
        std::stable_sort(g.shells.begin(), g.shells.end(), [&](int p, int q) { return cost(g, p) > cost(g, q); });
        const int ng = (int)g.shells.size();
        GD_CHECK(ng < (int)BATCH_PAD, GD_ERR_ARG, "gd_batch: %d shells in one launch", ng);
        std::vector<int> boff(ng), bcnt(ng);
        std::vector<unsigned short> sob;
        for (int k = 0; k < ng; ++k) {
            bcnt[k] = blocks(g, g.shells[k]);
            while (sob.size() % 8) sob.push_back(BATCH_PAD);
            boff[k] = (int)sob.size();
            sob.insert(sob.end(), (size_t)bcnt[k], (unsigned short)k);
        }
        GD_CHECK(sob.size() < (size_t)1 << 31, GD_ERR_ARG, "gd_batch: %zu blocks in one launch", sob.size());
        g.nblocks = (int)sob.size();
        switch (g.kind) {
        case BK_SYNTH: {
            std::vector<SynthArgs> v;
            for (int i : g.shells) v.push_back(xs[i].a);
            g.off_args = put(v.data(), sizeof(SynthArgs) * v.size());
            break;
        }
        case BK_RING: {
            std::vector<RingArgs> v;
            for (int i : g.shells) v.push_back(xs[i].r);
            g.off_args = put(v.data(), sizeof(RingArgs) * v.size());
            break;
        }
        case BK_ANAV:
        case BK_ANAS: {
            std::vector<AnaArgs> v;
            for (int i : g.shells) v.push_back(g.kind == BK_ANAV ? xs[i].gv : xs[i].gs);
            g.off_args = put(v.data(), sizeof(AnaArgs) * v.size());
            break;
        }
        default: {
            std::vector<ChainArgs> v;
            for (int i : g.shells) v.push_back(xs[i].c);
            g.off_args = put(v.data(), sizeof(ChainArgs) * v.size());
            break;
        }
        }
        g.off_sob = put(sob.data(), sizeof(unsigned short) * sob.size());
        g.off_boff = put(boff.data(), sizeof(int) * ng);
        g.off_bcnt = put(bcnt.data(), sizeof(int) * ng);
    }
    if (blob.empty()) return GD_OK;
    GD_TRY(b->table[parity].ensure(blob.size()));
    GD_HIP(hipMemcpy(b->table[parity].p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    return GD_OK;
}

template <class A> static BatchRef<A> group_ref(const gd_batch *b, int parity, const BatchGroup &g)
{
    const char *base = static_cast<const char *>(b->table[parity].p);
    return BatchRef<A>{reinterpret_cast<const A *>(base + g.off_args),
                       reinterpret_cast<const unsigned short *>(base + g.off_sob),
                       reinterpret_cast<const int *>(base + g.off_boff),
                       reinterpret_cast<const int *>(base + g.off_bcnt)};
}

template <typename T> static int launch_group(gd_batch *b, int parity, const BatchGroup &g, hipStream_t st)
{
    switch (g.kind) {
    case BK_SYNTH: return gd_launch_synth_tang_batch<T>(st, group_ref<SynthArgs>(b, parity, g), g.nblocks, g.key, g.mm);
    case BK_RING: return gd_launch_ring_tang_batch<T>(st, group_ref<RingArgs>(b, parity, g), g.nblocks, g.key);
    case BK_ANAV:
    case BK_ANAS:
        return gd_launch_ana_batch<T>(st, group_ref<AnaArgs>(b, parity, g), g.nblocks, g.kind == BK_ANAV, g.key, g.mm);
    default:
        return gd_launch_cnab2_chain_batch(st, group_ref<ChainArgs>(b, parity, g), g.nblocks, g.key, g.fm, g.lds,
                                           g.fp64 != 0);
    }
}

// every shell's step of the given parity: the shells without a fused form
// first (their own steps, on the context stream), then each lane's fused
// stages (one lane: on the context stream; more: on batch streams forked
// from and joined to it)
static int fused_lane(gd_batch *b, int parity, int lane, hipStream_t st)
{
    bool extras = false;
    for (const BatchGroup &g : b->groups[parity]) {
        if (g.lane != lane) continue;
        if (g.kind == BK_CHAIN && !extras) {
            // per-shell work between the analyses and the solves: the radial
            // forcing (after nl(u_pol) is written) and the temperature
            // conditions on the state the solve reads
            extras = true;
            for (const BatchGroup &h : b->groups[parity]) {
                if (h.lane != lane || h.kind != BK_CHAIN) continue;
                for (int i : h.shells) {
                    gd_step *s = b->steps[i];
                    GD_TRY(radial_force<double>(s, st, s->nlbuf[parity].as<double>()));
                    GD_TRY(temperature_bc(s, parity, st));
                }
            }
        }
        GD_TRY(g.fp64 ? launch_group<double>(b, parity, g, st) : launch_group<float>(b, parity, g, st));
    }
    return GD_OK;
}

static int fused_enqueue(gd_batch *b, int parity)
{
    hipStream_t main = b->ctx->stream;
    GD_CHECK(!b->groups[parity].empty() || !b->solo[parity].empty(), GD_ERR_ARG, "gd_batch: fused tables not built");
    for (int i : b->solo[parity]) {
        gd_step *s = b->steps[i];
        const int ov = s->overlap;
        s->overlap = 0;
        const int rc = enqueue_step(s, parity, false);
        s->overlap = ov;
        GD_TRY(rc);
    }
    const int L = b->lanes;
    if (L == 1) {
        GD_TRY(fused_lane(b, parity, 0, main));
    } else {
        GD_CHECK(L <= (int)b->streams.size(), GD_ERR_ARG, "gd_batch: %d lanes, %zu streams", L, b->streams.size());
        GD_HIP(hipEventRecord(b->fork, main));
        for (int l = 0; l < L; ++l) GD_HIP(hipStreamWaitEvent(b->streams[l], b->fork, 0));
        int rc = GD_OK;
        for (int l = 0; l < L && rc == GD_OK; ++l) rc = fused_lane(b, parity, l, b->streams[l]);
        // every forked stream joins back, also after an error (the capture
        // then ends with the error already set)
        for (int l = 0; l < L; ++l) {
            GD_HIP(hipEventRecord(b->joins[l], b->streams[l]));
            GD_HIP(hipStreamWaitEvent(main, b->joins[l], 0));
        }
        GD_TRY(rc);
    }
    for (const BatchGroup &g : b->groups[parity])
        if (g.kind == BK_CHAIN)
            for (int i : g.shells) b->steps[i]->syn_ok = true; // synV describes the new state
    return GD_OK;
}

// every shell's step of the given parity, forked over the batch streams
static int batch_enqueue(gd_batch *b, int parity)
{
    hipStream_t main = b->ctx->stream;
    const int ns = (int)b->streams.size();
    GD_HIP(hipEventRecord(b->fork, main));
    for (int k = 0; k < ns; ++k) GD_HIP(hipStreamWaitEvent(b->streams[k], b->fork, 0));
    int rc = GD_OK;
    for (size_t i = 0; i < b->steps.size() && rc == GD_OK; ++i) {
        gd_step *st = b->steps[i];
        b->ctx->stream = b->streams[i % ns]; // the step's launches follow the context stream
        // no second stream inside a batch stream: a capture that forked each
        // FP32 shell's side stream from its batch stream ended in a host
        // segmentation fault in this function; the cause was not found, so
        // batches keep every shell's step on its batch stream
        const int ov = st->overlap;
        st->overlap = 0;
        rc = enqueue_step(st, parity, false);
        st->overlap = ov;
    }
    b->ctx->stream = main;
    GD_TRY(rc);
    for (int k = 0; k < ns; ++k) {
        GD_HIP(hipEventRecord(b->joins[k], b->streams[k]));
        GD_HIP(hipStreamWaitEvent(main, b->joins[k], 0));
    }
    return GD_OK;
}

extern "C" int gd_batch_run(gd_batch *b, int nsteps)
{
    GD_CHECK(b && nsteps >= 0, GD_ERR_ARG, "gd_batch_run: bad args");
    GD_HIP(hipSetDevice(b->ctx->device));
    const int s0 = b->steps[0]->step;
    for (gd_step *s : b->steps) GD_CHECK(s->step == s0, GD_ERR_ARG, "gd_batch_run: shells at different steps");
    // the per-shell stream branches run FP64 shells only: with FP32 shells
    // their graph capture once ended in a host fault whose cause was not
    // found (the fused form runs every precision)
    if (b->mode == GD_BATCH_STREAMS)
        for (gd_step *s : b->steps)
            GD_CHECK(s->d.prec == GD_PREC_FP64, GD_ERR_ARG,
                     "gd_batch_run: the stream-branch form takes FP64 shells; use GD_BATCH_FUSED for FP32 and BF16x3");
    for (int k = 0; k < nsteps; ++k) {
        const int step = b->steps[0]->step;
        if (step == 0) {
            for (gd_step *s : b->steps) { // AB1 bootstraps, eager, one stream per shell as in the graphs
                const int ov = s->overlap;
                s->overlap = 0;
                const int rc = enqueue_step(s, 0, true);
                s->overlap = ov;
                GD_TRY(rc);
            }
        } else {
            const int par = step & 1;
            for (gd_step *s : b->steps) GD_TRY(ensure_syn(s));
            if (b->gexec[par] && batch_stale(b, par)) {
                // a member changed its captured work: the graph launches of
                // both parities may hold stale arguments (freed condition
                // tables, missing composition work)
                GD_HIP(hipStreamSynchronize(b->ctx->stream));
                batch_drop_graphs(b);
            }
            if (!b->gexec[par]) {
                hipStream_t main = b->ctx->stream;
                if (b->mode == GD_BATCH_FUSED && b->groups[par].empty() && b->solo[par].empty()) {
                    GD_HIP(hipStreamSynchronize(main)); // the table may be in use by a replay
                    GD_TRY(fused_build(b, par));
                }
                GD_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
                const int rc = b->mode == GD_BATCH_FUSED ? fused_enqueue(b, par) : batch_enqueue(b, par);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(main, &g);
                GD_TRY(rc);
                if (e != hipSuccess) {
                    gd_set_error("gd_batch_run: hipStreamEndCapture: %s", hipGetErrorString(e));
                    return GD_ERR_HIP;
                }
                b->graph[par] = g;
                GD_HIP(hipGraphInstantiate(&b->gexec[par], g, nullptr, nullptr, 0));
                b->gens[par].clear();
                for (gd_step *s : b->steps) b->gens[par].push_back(s->gen);
            }
            GD_HIP(hipGraphLaunch(b->gexec[par], b->ctx->stream));
        }
        for (gd_step *s : b->steps) ++s->step;
    }
    return GD_OK;
}

extern "C" int gd_batch_set_mode(gd_batch *b, int mode)
{
    GD_CHECK(b && (mode == GD_BATCH_STREAMS || mode == GD_BATCH_FUSED), GD_ERR_ARG, "gd_batch_set_mode: bad args");
    if (mode == b->mode) return GD_OK;
    GD_HIP(hipSetDevice(b->ctx->device));
    GD_HIP(hipStreamSynchronize(b->ctx->stream));
    batch_drop_graphs(b);
    b->mode = mode;
    return GD_OK;
}

extern "C" int gd_batch_set_tiles(gd_batch *b, const int *tiles)
{
    GD_CHECK(b && tiles, GD_ERR_ARG, "gd_batch_set_tiles: null argument");
    for (size_t i = 0; i < b->steps.size(); ++i)
        GD_CHECK(gd_synth_tile_ok(tiles[i], b->steps[i]->d.prec == GD_PREC_FP64), GD_ERR_ARG,
                 "gd_batch_set_tiles: shell %zu: tile %d not available", i, tiles[i]);
    GD_HIP(hipSetDevice(b->ctx->device));
    GD_HIP(hipStreamSynchronize(b->ctx->stream));
    batch_drop_graphs(b);
    b->tiles.assign(tiles, tiles + b->steps.size());
    return GD_OK;
}

extern "C" int gd_batch_get_tiles(const gd_batch *b, int *tiles)
{
    GD_CHECK(b && tiles, GD_ERR_ARG, "gd_batch_get_tiles: null argument");
    std::copy(b->tiles.begin(), b->tiles.end(), tiles);
    return GD_OK;
}

// mean time of reps launches of fn on the context stream, after one untimed launch
template <class F> static int time_ms(gd_batch *b, int reps, F fn, double *ms)
{
    hipStream_t st = b->ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GD_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    int rc = e == hipSuccess ? fn() : GD_ERR_HIP;
    if (rc == GD_OK && (e = hipEventRecord(e0, st)) != hipSuccess) rc = GD_ERR_HIP;
    for (int k = 0; k < reps && rc == GD_OK; ++k) rc = fn();
    if (rc == GD_OK && (e = hipEventRecord(e1, st)) != hipSuccess) rc = GD_ERR_HIP;
    if (rc == GD_OK && (e = hipEventSynchronize(e1)) != hipSuccess) rc = GD_ERR_HIP;
    float t = 0;
    if (rc == GD_OK && (e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) rc = GD_ERR_HIP;
    if (rc == GD_ERR_HIP && e != hipSuccess) gd_set_error("gd_batch_tune: %s", hipGetErrorString(e));
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

// The synthesis tiles of the fused mode, measured on this device: each shell's
// synthesis alone in every tile form it has (its fastest per shell), then the
// fused synthesis launches of the whole batch with those per-shell choices and
// with each tile for every shell; the fastest assignment is kept.  The
// synthesis writes only the shells' transform scratch (FB), which every step
// rewrites before reading.
extern "C" int gd_batch_tune(gd_batch *b, int reps, double *ms_out)
{
    GD_CHECK(b && reps >= 1, GD_ERR_ARG, "gd_batch_tune: bad args");
    GD_HIP(hipSetDevice(b->ctx->device));
    GD_HIP(hipStreamSynchronize(b->ctx->stream));
    const int n = (int)b->steps.size();
    std::vector<int> best(n, 0);
    for (int i = 0; i < n; ++i) {
        gd_step *s = b->steps[i];
        ShellArgs x;
        if (!fused_args(s, 0, x)) continue;
        double tbest = 0;
        for (int t = 0; t < 3; ++t) {
            if (!gd_synth_tile_ok(t, x.fp64)) continue;
            double ms = 0;
            GD_TRY(time_ms(b, reps, [&]() -> int {
                return x.fp64 ? gd_launch_synth_tang_tile<double>(b->ctx->stream, x.a, t)
                              : gd_launch_synth_tang_tile<float>(b->ctx->stream, x.a, t);
            }, &ms));
            if (t == 0 || ms < tbest) {
                tbest = ms;
                best[i] = t;
            }
        }
    }
    std::vector<std::vector<int>> cands{best};
    for (int t = 0; t < 3; ++t) {
        std::vector<int> u(n);
        for (int i = 0; i < n; ++i) u[i] = gd_synth_tile_ok(t, b->steps[i]->d.prec == GD_PREC_FP64) ? t : 0;
        cands.push_back(u);
    }
    batch_drop_graphs(b);
    double tmin = 0;
    std::vector<int> pick = best;
    for (size_t c = 0; c < cands.size(); ++c) {
        b->tiles = cands[c];
        GD_TRY(fused_build(b, 0));
        double ms = 0;
        GD_TRY(time_ms(b, reps, [&]() -> int {
            for (const BatchGroup &g : b->groups[0])
                if (g.kind == BK_SYNTH) GD_TRY(g.fp64 ? launch_group<double>(b, 0, g, b->ctx->stream)
                                                      : launch_group<float>(b, 0, g, b->ctx->stream));
            return GD_OK;
        }, &ms));
        if (ms_out) ms_out[c] = ms;
        if (c == 0 || ms < tmin) {
            tmin = ms;
            pick = cands[c];
        }
    }
    GD_HIP(hipStreamSynchronize(b->ctx->stream));
    batch_drop_graphs(b);
    b->tiles = pick;
    return GD_OK;
}

extern "C" int gd_batch_set_lanes(gd_batch *b, int lanes)
{
    GD_CHECK(b && lanes >= 1 && lanes <= (int)b->streams.size(), GD_ERR_ARG,
             "gd_batch_set_lanes: lanes must be 1 .. the batch's stream count");
    if (lanes == b->lanes) return GD_OK;
    GD_HIP(hipSetDevice(b->ctx->device));
    GD_HIP(hipStreamSynchronize(b->ctx->stream));
    batch_drop_graphs(b);
    b->lanes = lanes;
    return GD_OK;
}

extern "C" int gd_batch_get_lanes(const gd_batch *b, int *lanes)
{
    GD_CHECK(b && lanes, GD_ERR_ARG, "gd_batch_get_lanes: null argument");
    *lanes = b->lanes;
    return GD_OK;
}

// Lane count of the fused mode, measured: for each candidate (1, 2, 4, 8 up to
// the stream count and the shell count) the batch runs 2 steps (the
// captures), nsteps untimed and nsteps timed steps; the fastest is kept.
// These are the batch's real steps (every lane count gives the same bits), so
// the shells advance by 2 nsteps + 2 steps per candidate.
extern "C" int gd_batch_tune_lanes(gd_batch *b, int nsteps, double *ms_out)
{
    GD_CHECK(b && nsteps >= 1 && b->mode == GD_BATCH_FUSED, GD_ERR_ARG, "gd_batch_tune_lanes: bad args (fused mode)");
    GD_HIP(hipSetDevice(b->ctx->device));
    const int n = (int)b->steps.size();
    int best = 1, k = 0;
    double tbest = 0;
    for (int L = 1; L <= 8; L *= 2, ++k) {
        double ms = -1;
        if (L <= (int)b->streams.size() && L <= n) {
            GD_TRY(gd_batch_set_lanes(b, L));
            GD_TRY(gd_batch_run(b, 2));
            GD_TRY(time_ms(b, 1, [&]() -> int { return gd_batch_run(b, nsteps); }, &ms));
            ms /= nsteps;
            if (L == 1 || ms < tbest) {
                tbest = ms;
                best = L;
            }
        }
        if (ms_out) ms_out[k] = ms;
    }
    return gd_batch_set_lanes(b, best);
}
