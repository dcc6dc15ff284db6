#!/usr/bin/env python3
"""CNAB2 steps/s at lmax=127, nr=65: mixed-precision vs FP64 GPU path.

Workload (one MI355X): the full pseudo-spectral CNAB2 step of the rotating
Boussinesq MHD shell, lmax=mmax=127 on a 192 x 384 Gauss grid, nr=65, radius
ratio 0.35, E=1e-3, Pr=1, Pm=5, Ra=500, dt=1e-5, band half-width 3; seeded
(l+1)^-2 state of amplitude 1e-3.  A "step" is one CNAB2 step of all five
fields with the state resident in HBM.

  value          mixed path (FP32 Legendre / FFT / products, FP64 radial
                 solves and state) steps/s over --steps graph-replayed steps
  fp64_*         the all-FP64 GPU path, same protocol
  accuracy       per-field max relative error of the mixed state against the
                 FP64 GPU state after --err-steps steps, and of the FP64 GPU
                 state against the FP64 CPU oracle after one step
  roofline       dominant kernel of the mixed step, its algorithmic bytes or
                 flops per launch over its HIP-event duration
  cpu_baseline   the FP64 oracle step (restatement of the reference step) on
                 the host cores, a bounded sample of steps
  full_products, exp_integrators, sht_c1, batch_c3, restart_io, scalar_bc
                 secondary measurements (every product formed; EAB2 / ERK2
                 updates; the lmax=255 transform round trip; a batch of
                 uneven shells; the restart file write / read; the scalar
                 boundary conditions with the tau flux correction)

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "geodynamo.jl_amd"))

METRIC = "CNAB2 steps/s at lmax=127 nr=65, mixed vs FP64 GPU at ≤1e-6 rel err"
CFG = dict(lmax=127, nlat=192, nlon=384, nr=65, E=1e-3, Pm=5.0, Pr=1.0, Ra=500.0, rratio=0.35,
           dt=1e-5, theta=0.5, bw=3)
SEED, AMP = 2024, 1e-3
PEAK = {"hbm": 8000.0, "fp32": 157.3, "fp64": 78.6, "bf16": 2500.0}  # GB/s, TFLOP/s (MI355X dense)


def kernel_costs(prec_bytes, tangential=True):
    """Algorithmic bytes and flops per launch of each kernel class of one step.

    Legendre flops count 2*nlat*nlm per scalar-equivalent contraction per level
    (the even/odd-symmetric count); a vector synthesis or analysis is 4 such
    contractions (S and T against dP/dtheta and m P/sin).  Bytes are the
    compulsory reads and writes of each kernel's HBM operands (tables
    excluded: they stay in L2/MALL).

    tangential (the default step: radial components zero, so the tangential
    parts of u x omega, j x B, u x B and u_r dT/dr are identically zero):
      synthesis: u (4) and grad T from T/r (2); ring: 4 fields in, 3 products
      out; analysis: the force (4) and -u.grad T (1)
    full products: synthesis of u, omega, B, j (16) and grad T (2) plus dT/dr
      (1); ring: 15 fields in (60 slots per record, 16 of them zero), 5
      products out; analysis: force and u x B (8) and -u.grad T (1)
    The CNAB2 kernel reads u, nl, nl_prev and writes the new state (64 B per
    complex coefficient and field; 32 B for the two magnetic fields of the
    tangential step, whose nonlinear terms are identically zero and are not
    read) and also writes the next step's synthesis columns (6 or 20 per
    level), so the separate spectral preparation does not run in the steady
    state."""
    L, nlat, nr = CFG["lmax"], CFG["nlat"], CFG["nr"]
    nlm = (L + 1) * (L + 2) // 2
    nm, nlat2 = L + 1, nlat // 2
    es = prec_bytes
    fs = nlm * nr
    four = nlat2 * nr * nm  # Fourier (i, level, m) points
    contraction = 2.0 * nlat * nlm * nr
    if tangential:
        return {
            "spectral_pre": dict(bytes=5 * 16 * fs + 20 * es * fs, flops=0.0),
            "synth_vector": dict(bytes=6 * es * fs + 16 * es * four, flops=6 * contraction),
            "synth_scalar": dict(bytes=0.0, flops=0.0),
            "ring_fft_products": dict(bytes=16 * es * four + 12 * es * four, flops=0.0),
            "analysis_vector": dict(bytes=8 * es * four + 2 * 16 * fs, flops=4 * contraction),
            "analysis_scalar": dict(bytes=4 * es * four + 16 * fs, flops=1 * contraction),
            "cnab2_solve": dict(bytes=(3 * 64 + 2 * 32) * fs + 6 * es * fs, flops=0.0),
        }
    return {
        "spectral_pre": dict(bytes=5 * 16 * fs + 20 * es * fs, flops=0.0),
        "synth_vector": dict(bytes=18 * es * fs + 40 * es * four, flops=18 * contraction),
        "synth_scalar": dict(bytes=2 * es * fs + 4 * es * four, flops=1 * contraction),
        "ring_fft_products": dict(bytes=60 * es * four + 20 * es * four, flops=0.0),
        "analysis_vector": dict(bytes=16 * es * four + 4 * 16 * fs, flops=8 * contraction),
        "analysis_scalar": dict(bytes=4 * es * four + 16 * fs, flops=1 * contraction),
        "cnab2_solve": dict(bytes=5 * 64 * fs + 20 * es * fs, flops=0.0),
    }


# kernel of each class on the headline (tangential) path, as rocprofv3 names it
# (synthesis: u and grad T as gradient fields, or, with GD_SYNTH_GRAD=0, as
# full fields; CNAB2: one lane per column on the FP64 path, the twisted form
# on the FP32-synthesis path, re and im columns paired in a lane by default)
TANG_KERNELS = {"spectral_pre": "spectral_pre_kernel<{T}>",
                "synth_vector": ("synth_kernel<{T}, true, 1, 3, 16, true, 1, 0>",
                                 "synth_kernel<{T}, true, 2, 4, 16, true, 0, 0>"),
                "ring_fft_products": "ring_kernel<{T}, 3, 384>", "analysis_vector": "ana_res_kernel<{T}, true, 1, 96, 0>",
                "analysis_scalar": "ana_res_kernel<{T}, false, 1, 96, 0>",
                "cnab2_solve": ("cnab2_chain_kernel<3, 4, {T}, true,", "cnab2_twist_kernel<3, 2, {T}, 2>",
                                "cnab2_twist_kernel<3, 4, {T}, 1>")}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pmc_step_lmax127_nr65.json")
COPY_PROBE_BYTES = 268435456


def derive_traffic(prof_dir):
    """HBM bytes per launch of each kernel class from two rocprofv3 --pmc runs
    of `bench.py --no-full --copy-probe 268435456` (FETCH_SIZE in
    prof_dir/fetch, WRITE_SIZE in prof_dir/write; KiB per dispatch), each
    scaled by what the counter reads for copy_kernel moving a known number of
    bytes.  Writes TRAFFIC_JSON."""
    import collections
    import csv

    val = {}
    for cnt in ("fetch", "write"):
        agg = collections.defaultdict(list)
        with open(os.path.join(prof_dir, cnt, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                agg[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024)
        val[cnt] = {k: sum(v) / len(v) for k, v in agg.items()}
    copy = [k for k in val["fetch"] if k.startswith("copy_kernel")][0]
    fscale = COPY_PROBE_BYTES / val["fetch"][copy]
    wscale = COPY_PROBE_BYTES / val["write"][copy]
    out = {"fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4), "fp32": {}, "fp64": {}}
    for prec, T in (("fp32", "float"), ("fp64", "double")):
        for cls, pat in TANG_KERNELS.items():
            pats = pat if isinstance(pat, tuple) else (pat,)
            ks = [k for k in val["fetch"] if any(p.format(T=T) in k for p in pats)]
            if not ks:
                continue
            rd, wr = val["fetch"][ks[0]] * fscale, val["write"].get(ks[0], 0.0) * wscale
            out[prec][cls] = {"kernel": ks[0], "read_bytes": round(rd), "write_bytes": round(wr),
                              "hbm_bytes_per_launch": round(rd + wr)}
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(out, f, indent=1)
    return out


def summarize_trace(prof_dir):
    """Markdown table of a rocprofv3 --kernel-trace --stats run of this bench
    (prof_dir/trace/run_kernel_stats.csv) beside the HBM traffic of the
    headline kernels (TRAFFIC_JSON)."""
    import csv

    with open(os.path.join(prof_dir, "trace", "run_kernel_stats.csv")) as f:
        rows = list(csv.DictReader(f))
    traffic = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as f:
            tj = json.load(f)
        for prec in ("fp32", "fp64"):
            for cls, rec in tj.get(prec, {}).items():
                traffic[rec["kernel"]] = (cls, rec)
    lines = ["| kernel | calls | avg us | min us | max us | class | HBM read MB | HBM write MB |",
             "|---|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        name = r["Name"].replace("(anonymous namespace)::", "").replace("void ", "")
        cls, rd, wr = "", "", ""
        if r["Name"] in traffic:
            cls, rec = traffic[r["Name"]]
            rd, wr = "%.1f" % (rec["read_bytes"] / 1e6), "%.1f" % (rec["write_bytes"] / 1e6)
        lines.append("| `%s` | %s | %.2f | %.2f | %.2f | %s | %s | %s |" % (
            name.split("(")[0], r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
            float(r["MaxNs"]) / 1e3, cls, rd, wr))
    return "\n".join(lines)


def step_roofline(sec_per_step):
    """The whole mixed step against HBM: the algorithmic bytes of the kernels a
    steady-state step launches (kernel_costs: synthesis, ring pass, both
    analyses, CNAB2 solve; the spectral preparation runs only after a new
    state is set) over the measured step time, beside the bytes of the fused
    step the design survey names as the bandwidth target (0.27 GB of FP32
    Fourier-space arrays and 0.2 GB of spectral and solve traffic per step at
    lmax=127, nr=65)."""
    costs = kernel_costs(4)
    b = sum(costs[k]["bytes"] for k in ("synth_vector", "ring_fft_products", "analysis_vector", "analysis_scalar",
                                        "cnab2_solve"))
    ach = b / sec_per_step / 1e9
    h2 = 0.47e9
    return {"algorithmic_bytes": b, "achieved": round(ach, 1), "unit": "GB/s", "frac": round(ach / PEAK["hbm"], 4),
            "ceiling_steps_per_s_at_peak": round(PEAK["hbm"] * 1e9 / b, 1),
            "fused_target_bytes": h2, "fused_target_steps_per_s_at_peak": round(PEAK["hbm"] * 1e9 / h2, 1)}


def roofline(prof_ms, prec):
    es = 8 if prec == "fp64" else 4
    costs = kernel_costs(es)
    name = max(prof_ms, key=prof_ms.get)
    t = prof_ms[name] * 1e-3
    c = costs[name]
    peak_f = PEAK["fp64" if prec == "fp64" else "fp32"]
    ridge = peak_f * 1e12 / (PEAK["hbm"] * 1e9)
    mfma = c["flops"] > 0 and c["flops"] / c["bytes"] > ridge
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # runs (scaled by a copy kernel of known size), when they cover this kernel
    traffic = None
    pmc = TRAFFIC_JSON
    if os.path.exists(pmc):
        with open(pmc) as f:
            rec = json.load(f).get(prec, {}).get(name)
        if rec:
            traffic = round(rec["hbm_bytes_per_launch"])
    if mfma:
        ach = c["flops"] / t / 1e12
        return name, {"bound": "mfma", "achieved": round(ach, 3), "peak": peak_f, "unit": "TFLOP/s",
                      "frac": round(ach / peak_f, 4), "traffic": traffic, "kernel": name,
                      "kernel_ms": round(prof_ms[name], 5), "algorithmic_flops": c["flops"],
                      "algorithmic_bytes": c["bytes"]}
    ach = c["bytes"] / t / 1e9
    return name, {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK["hbm"], "unit": "GB/s",
                  "frac": round(ach / PEAK["hbm"], 4), "traffic": traffic, "kernel": name,
                  "kernel_ms": round(prof_ms[name], 5), "algorithmic_bytes": c["bytes"]}


def rel_err(a, b):
    import numpy as np

    return [float(np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300)) for k in range(5)]


def timed(step, st, steps, warmup):
    step.set_state_complex(st)
    step.run(warmup, graph=True)
    step.synchronize()
    t0 = time.perf_counter()
    step.run(steps, graph=True)
    step.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(st, n_steps, gpu_states):
    """The FP64 oracle step (restatement of the reference step) on the host
    cores, following scripts/bench_timestepping.jl:26-33: the first step (the
    AB1 bootstrap) is the warm-up, then the median of 10 individually timed
    steps on all host cores; the run continues to n_steps so that the GPU
    states after n_steps are compared with the oracle's.  Then one warm-up and
    the median of 3 steps on one core.  gpu_states: {name: (state after one
    step, state after n_steps, energies after n_steps)}."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import oracle as O

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(cores)
    m = O.Model(CFG["lmax"], CFG["nlat"], CFG["nlon"], CFG["nr"], E=CFG["E"], Pm=CFG["Pm"], Pr=CFG["Pr"],
                Ra=CFG["Ra"], rratio=CFG["rratio"], dt=CFG["dt"], theta=CFG["theta"], bw=CFG["bw"])
    t0 = time.perf_counter()
    s, nl, nlp = m.step(st, None, 0)
    t_warm = time.perf_counter() - t0
    first = s
    times = []
    for k in range(1, n_steps):
        t0 = time.perf_counter()
        s, nl, nlp = m.step(s, nlp, k)
        if k <= 10:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    e_or = m.energies(s)
    acc = {}
    for name, (one, fin, e) in gpu_states.items():
        acc["%s_vs_oracle_max_rel_per_field_after_1_step" % name] = [float("%.3e" % x) for x in rel_err(one, first)]
        acc["%s_vs_oracle_max_rel_per_field_after_%d_steps" % (name, n_steps)] = [
            float("%.3e" % x) for x in rel_err(fin, s)]
        acc["%s_energies_vs_oracle_rel_after_%d_steps" % (name, n_steps)] = [
            float("%.3e" % (abs(a - b) / abs(b))) for a, b in zip(e, e_or)]
    # one core: warm-up, then the median of 3 steps
    O.set_threads(1)
    s1, nl1, np1 = m.step(st, None, 0)
    t1 = []
    for k in range(1, 4):
        t0 = time.perf_counter()
        s1, nl1, np1 = m.step(s1, np1, k)
        t1.append(time.perf_counter() - t0)
    O.set_threads(cores)
    med1 = float(np.median(t1))
    return {"value": round(1.0 / med, 4), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": ("FP64 oracle CNAB2 step at lmax=127 nr=65 on %d threads: 1 warm-up step (%.2f s), median "
                       "of the next 10 steps (%.3f s per step); run on to step %d for the accuracy comparison"
                       % (cores, t_warm, med, n_steps)),
            "one_core": {"value": round(1.0 / med1, 4), "unit": "steps/s", "cores": 1,
                         "sample": "1 warm-up step, median of 3 steps (%.2f s per step)" % med1}}, acc


def exp_integrators(reps):
    """EAB2 (timestep.jl:425-487) and the staged ERK2 calls (timestep.jl:1316-1555)
    at this shell, nu = 1, dt = 1e-5, fields device resident.  reps consecutive
    calls are captured in one graph and replayed (the per-call host time of the
    ctypes boundary would otherwise hide the device time); the replay is checked
    bit for bit against the same calls made eagerly."""
    import torch

    import gdhip

    lmax, nr, dt = CFG["lmax"], CFG["nr"], CFG["dt"]
    cfg = gdhip.create_shtnskit_config(lmax, lmax, nlat=CFG["nlat"], nlon=CFG["nlon"])
    dom = gdhip.create_radial_domain(nr, CFG["rratio"])
    t0 = time.perf_counter()
    etd = gdhip.create_etd_cache(cfg, dom, 1.0, dt)
    t_etd = time.perf_counter() - t0
    t0 = time.perf_counter()
    erk = gdhip.create_erk2_cache(cfg, dom, 1.0, dt)
    t_erk = time.perf_counter() - t0
    gen = torch.Generator(device="cuda").manual_seed(5)

    def field():
        f = gdhip.SHTnsSpectralField(cfg, nr)
        f.data_real = torch.randn(nr, cfg.nlm, dtype=torch.float64, device="cuda", generator=gen)
        f.data_imag = torch.randn(nr, cfg.nlm, dtype=torch.float64, device="cuda", generator=gen)
        return f

    u, nl, npv = field(), field(), field()
    buf = gdhip.ERK2FieldBuffers(u, nl, erk)
    u0 = (u.data_real.clone(), u.data_imag.clone())
    calls = {
        "eab2_update": lambda: gdhip.eab2_update(u, nl, npv, etd, cfg, dt),
        "erk2_prepare_field": lambda: gdhip.erk2_prepare_field(buf, u, nl, erk, cfg, dt),
        "erk2_finalize_field": lambda: gdhip.erk2_finalize_field(buf, u, erk, cfg, dt),
    }
    plane = cfg.nlm * nr * 8
    mats = (lmax + 1) * nr * nr * 8
    # algorithmic HBM bytes per call: planes read + written once, each matrix set once
    nbytes = {"eab2_update": 8 * plane + 2 * mats, "erk2_prepare_field": 12 * plane + 4 * mats,
              "erk2_finalize_field": 10 * plane + 1 * mats}
    out = {"workload": "lmax=mmax=127 nr=65 nu=1 dt=1e-5, one field (re, im planes) in HBM",
           "cache_build_s": {"etd": round(t_etd, 3), "erk2": round(t_erk, 3)}}
    s = torch.cuda.Stream()
    for name, fn in calls.items():
        def reset():
            u.data_real.copy_(u0[0])
            u.data_imag.copy_(u0[1])
        reset()
        fn()  # first call sizes the library's work buffers outside the capture
        torch.cuda.synchronize()
        reset()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        eager = (u.data_real.clone(), u.data_imag.clone())
        reset()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                fn()
        reset()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        same = bool(torch.equal(u.data_real, eager[0]) and torch.equal(u.data_imag, eager[1]))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        nrep = 5
        e0.record()
        for _ in range(nrep):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / (nrep * reps)
        out[name] = {"ms_per_call": round(ms, 5), "algorithmic_MB": round(nbytes[name] / 1e6, 2),
                     "GB_per_s": round(nbytes[name] / (ms * 1e-3) / 1e9, 1),
                     "graph_replay_equals_eager": same}
    return out


def adjoint_c4(nsteps=20, every=5, reps=3):
    """The adjoint configuration: lmax=mmax=63 on a 96 x 192 Gauss grid, nr=65,
    dt=1e-5, the seeded (l+1)^-2 state of amplitude 1e-3; nsteps CNAB2 steps
    with checkpoints every `every` steps, the misfit 1/2 ||B_r(r) - B_r^target||^2
    on the surface grid (B_r = l(l+1) B_pol / r^2 at the outermost level the
    solve determines) and its gradient with respect to the initial state.
    Reported per path (FP64; mixed = FP32 transforms and products, FP64
    solves): the forward run alone (nsteps eager steps, as the adjoint's
    forward enqueues them), the forward with checkpoints + misfit + backward
    pass, their ratio, and the mixed gradient's error against the FP64 one.
    The B_r cotangent never reaches u or T (B is coupled to them through no
    term of the tangential step), so its reverse pass is the transposed solves
    alone; the temperature misfit at mid depth times the full reverse pass
    (segments recomputed, the transposed nonlinear terms at every step)."""
    import numpy as np
    import torch

    import gdhip

    lmax, nlat, nlon, nr = 63, 96, 192, 65
    kw = dict(E=CFG["E"], Pm=CFG["Pm"], Pr=CFG["Pr"], Ra=CFG["Ra"], rratio=CFG["rratio"], dt=CFG["dt"],
              theta=CFG["theta"], bw=CFG["bw"])
    x0 = gdhip.seeded_state(SEED, lmax, nr, amp=AMP)
    # target: B_r of a seeded (l+1)^-3 poloidal field at the outer radius
    P = gdhip.seeded_state(99, lmax, nr, nfields=1, amp=AMP, power=3.0, zero_boundaries=False)[0, nr - 1]
    cfg = gdhip.create_shtnskit_config(lmax, lmax, nlat=nlat, nlon=nlon)
    ll = cfg.l_values * (cfg.l_values + 1.0)
    ro = gdhip.create_radial_domain(nr, CFG["rratio"]).r[3, -1]
    tspec = gdhip.SHTnsSpectralField(cfg, 1, np.ascontiguousarray((ll * P / ro ** 2).real[None, :]),
                                     np.ascontiguousarray((ll * P / ro ** 2).imag[None, :]))
    tphys = gdhip.SHTnsPhysicalField(cfg, 1)
    gdhip.shtnskit_spectral_to_physical(tspec, tphys)
    target = tphys.data.ravel()
    target_T = np.random.default_rng(7).standard_normal(nlat * nlon) * 1e-3
    out = {"workload": "lmax=mmax=63 nlat=96 nlon=192 nr=65 dt=1e-5, %d steps, checkpoints every %d, "
                       "misfit 1/2||B_r - target||^2 at the outermost solved level" % (nsteps, every)}
    grads = {}
    # state and gradient stay in HBM (device tensors): no host transfer is timed
    x_re = torch.from_numpy(np.ascontiguousarray(x0.real)).cuda()
    x_im = torch.from_numpy(np.ascontiguousarray(x0.imag)).cuda()
    g_re, g_im = torch.empty_like(x_re), torch.empty_like(x_im)
    for prec in ("fp64", "fp32"):
        g = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision=prec, **kw)
        adj = gdhip.CNAB2Adjoint(g, nsteps, every=every)
        tf, ta, tt = [], [], []
        for r in range(reps + 1):
            g.set_state(x_re, x_im)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.run(nsteps)
            torch.cuda.synchronize()
            tf.append(time.perf_counter() - t0)
            g.set_state(x_re, x_im)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            adj.forward()
            J = adj.misfit(target)
            adj.backward(out=(g_re, g_im))
            torch.cuda.synchronize()
            ta.append(time.perf_counter() - t0)
            grads[prec] = (g_re.cpu().numpy() + 1j * g_im.cpu().numpy())
            # a functional whose cotangent reaches u and T: the full reverse pass
            # (recomputed segments, transposed nonlinear terms at every step)
            g.set_state(x_re, x_im)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            adj.forward()
            adj.misfit(target_T, kind="T", level=nr // 2)
            adj.backward(out=(g_re, g_im))
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - t0)
        tf, ta, tt = min(tf[1:]), min(ta[1:]), min(tt[1:])
        name = "mixed" if prec == "fp32" else "fp64"
        out[name] = {"forward_ms": round(tf * 1e3, 3), "forward_adjoint_ms": round(ta * 1e3, 3),
                     "ratio": round(ta / tf, 3), "misfit": float("%.10e" % J),
                     "temperature_misfit_forward_adjoint_ms": round(tt * 1e3, 3),
                     "temperature_misfit_ratio": round(tt / tf, 3)}
        adj.close()
        g.close()
    e = np.abs(grads["fp32"] - grads["fp64"]).max() / np.abs(grads["fp64"]).max()
    out["mixed_gradient_vs_fp64_max_rel"] = float("%.3e" % e)
    return out


def c3_shells(nshell, seed=3):
    """The uneven-batch workload: lmax drawn from {31, 47, 63, 95, 127}, nr
    from {33, 49, 65, 97}, grid 3(L+1)/2 x 3(L+1), radius ratio U[0.2, 0.5]
    (numpy seed 3)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    shells = []
    for _ in range(nshell):
        L = int(rng.choice([31, 47, 63, 95, 127]))
        nr = int(rng.choice([33, 49, 65, 97]))
        shells.append((L, nr, 3 * (L + 1) // 2, 3 * (L + 1), float(rng.uniform(0.2, 0.5))))
    return shells


def batch_c3_prec(shells, steps, nstreams, prec, with_streams=True):
    """Aggregate shell-steps/s of the shells advanced (a) one after another by
    their own step graphs, (b) by one batch graph of fused launches (one grid
    per kernel form and stage over the shells of each lane; synthesis tiles
    from the tuning table, lane count from the lane autotuner), (c)
    with_streams: by one batch graph of per-shell branches over nstreams
    streams; each timed over `steps` steps after the bootstrap and both
    captures.  States of (b) and (c) are compared bit for bit with (a)."""
    import numpy as np

    import gdhip

    seq, bat = [], []
    for k, (L, nr, nlat, nlon, rr) in enumerate(shells):
        st = gdhip.seeded_state(500 + k, L, nr, amp=1e-3)
        for g in (seq, bat):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=500.0, dt=1e-5, precision=prec)
            s.set_state_complex(st)
            g.append(s)

    def timed(fn):
        seq[0].synchronize()
        t0 = time.perf_counter()
        fn()
        seq[0].synchronize()
        return time.perf_counter() - t0

    def seq_run(k):
        for _ in range(k):
            for s in seq:
                s.run(1, graph=True)

    def same():
        return all(np.array_equal(a.get_state_complex(), b.get_state_complex()) for a, b in zip(bat, seq))

    n = len(shells)
    fused = gdhip.StepBatch(bat, nstreams=nstreams, mode="fused")
    lanes = fused.tune_lanes(3)  # real steps of the batch
    fused.run(2)  # captures with the chosen lanes
    seq_run(bat[0].step_count - seq[0].step_count)
    tf = timed(lambda: fused.run(steps))
    ts = timed(lambda: seq_run(steps))
    out = {"sequential_shell_steps_per_s": round(n * steps / ts, 1),
           "fused_shell_steps_per_s": round(n * steps / tf, 1),
           "fused_over_sequential": round(ts / tf, 3), "fused_equal_bitwise": same(), "tiles": fused.tiles(),
           "lanes": lanes, "lanes_ms_per_step": [round(v, 4) for v in fused.lanes_ms]}
    fused.close()
    if with_streams:
        streams = gdhip.StepBatch(bat, nstreams=nstreams, mode="streams")
        streams.run(2)  # both captures
        seq_run(2)
        tb = timed(lambda: streams.run(steps))
        seq_run(steps)
        out.update({"streams_shell_steps_per_s": round(n * steps / tb, 1),
                    "streams_over_sequential": round(ts / tb, 3), "fused_over_streams": round(tb / tf, 3),
                    "streams_equal_bitwise": same()})
        streams.close()
    return out


def batch_c3(nshell, steps, nstreams, precs=("fp64", "fp32")):
    """Uneven batch: nshell independent shells of the workload c3_shells
    describes, (l+1)^-2 states of amplitude 1e-3, in FP64 and in FP32 (the
    mixed step).  The stream-branch form of FP32 shells stays off the GPU (a
    host fault in an earlier form of its capture): fused and sequential only."""
    shells = c3_shells(nshell)
    out = {"workload": "%d uneven shells, lmax {31,47,63,95,127}, nr {33,49,65,97}, 3(L+1)/2 x 3(L+1)" % nshell,
           "shells_lmax_nr": [[L, nr] for L, nr, _, _, _ in shells], "streams": nstreams, "steps": steps}
    for prec in precs:
        out[prec] = batch_c3_prec(shells, steps, nstreams, prec, with_streams=prec == "fp64")
    return out


def restart_io(step, reps):
    """Restart file of a step at its current state (after the timed runs): the
    five fields and the AB2 history in FP64, NetCDF CDF-2, written to and read
    back from a temporary directory on the local disk (median of reps).  Also
    times the device -> host part alone (state and history planes through the
    l-major permutation and the copy), and checks that the restored state and
    history equal the written ones bit for bit."""
    import shutil
    import tempfile

    import numpy as np

    import gdhip

    d = tempfile.mkdtemp(prefix="gd_restart_")
    try:
        cfg = gdhip.OutputConfig(output_dir=d, filename_prefix="bench")
        tr = gdhip.TimeTracker()
        st0, h0 = step.get_state_complex(), step.get_history()
        tw, tr_, tg = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            step.get_state()
            step.get_history()
            tg.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            path = gdhip.write_restart(step, tr, {"current_time": 1.0}, cfg)
            tw.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            gdhip.read_restart(step, tr, d, 1.0, cfg)
            step.synchronize()
            tr_.append(time.perf_counter() - t0)
        same = bool(np.array_equal(step.get_state_complex(), st0) and np.array_equal(step.get_history(), h0))
        nbytes = os.path.getsize(path)
        w, r, g = float(np.median(tw)), float(np.median(tr_)), float(np.median(tg))
        return {"workload": "restart file of the lmax=127 nr=65 step: 5 fields + AB2 history, FP64, NetCDF CDF-2, "
                            "local disk", "file_bytes": nbytes, "write_s": round(w, 4), "read_s": round(r, 4),
                "write_GB_per_s": round(nbytes / w / 1e9, 3), "read_GB_per_s": round(nbytes / r / 1e9, 3),
                "device_to_host_s": round(g, 4), "restored_bitwise": same, "reps": reps}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def scalar_bc(reps):
    """Boundary conditions of one scalar field at lmax=127, nr=65 (the
    temperature / composition functions with the tau flux correction): every
    mode with a flux condition at both boundaries, random profiles with
    non-zero imaginary parts, device tensors resident in HBM; reps
    consecutive applications captured in one graph, replayed five times and
    timed with events.
    Algorithmic bytes: both planes read and written once (4 x 8 x nr x nlm)."""
    import numpy as np
    import torch

    import gdhip

    lmax, nr = CFG["lmax"], CFG["nr"]
    cfg = gdhip.create_shtnskit_config(lmax, nlat=CFG["nlat"], nlon=CFG["nlon"])
    dom = gdhip.create_radial_domain(nr, CFG["rratio"])
    nlm = cfg.nlm
    g = torch.Generator(device="cuda").manual_seed(7)
    re = torch.randn(nr, nlm, dtype=torch.float64, device="cuda", generator=g)
    im = torch.randn(nr, nlm, dtype=torch.float64, device="cuda", generator=g)
    bc = gdhip.ScalarBoundaryConditions(cfg, dom, bw=CFG["bw"])
    two = np.full(nlm, 2, dtype=np.int32)
    bc.set(two, two, np.random.default_rng(8).standard_normal((2, nlm)))
    bc.apply(re, im)  # warm-up, eager
    torch.cuda.synchronize()
    # reps applications captured in one graph, so the launch rate of the host
    # loop does not enter the kernel time
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(reps):
            bc.apply(re, im)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nrep = 5
    e0.record()
    for _ in range(nrep):
        g.replay()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / (reps * nrep)
    nbytes = 4 * 8 * nr * nlm
    bc.close()
    return {"workload": "scalar field lmax=127 nr=65, flux (tau) conditions at both boundaries on every mode, "
                        "device planes", "ms_per_apply": round(ms, 5), "algorithmic_bytes": nbytes,
            "GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(nbytes / (ms * 1e-3) / 8e12, 4), "reps": reps * nrep}


C1 = dict(lmax=255, nlat=384, nlon=768, nr=97)


def sht_c1(reps):
    """Transform round trip at lmax = mmax = 255, 384 x 768, 97 levels: three
    scalar fields and one (tor, pol) pair, seeded normal x (l+1)^-1.5,
    synthesis then analysis of each, FP64, FP32 and BF16x3 (split-precision
    BF16 Legendre contractions, FP32 accumulation), device tensors.  Per
    precision: relative L2 error of the recovered coefficients (worst field),
    ms per round trip (events around reps round trips) and the Legendre work
    rate: 14 scalar-equivalent contractions (3 + 3 scalar, 4 + 4 for the
    pair) x 2 nlat nlm nr flops (the even/odd count) = 34.3 GFLOP per round
    trip, over the whole round trip (FFT and layout passes included), against
    the MFMA peak of the arithmetic type."""
    import numpy as np
    import torch

    import gdhip

    L, nlat, nlon, nr = C1["lmax"], C1["nlat"], C1["nlon"], C1["nr"]
    out = {"workload": "SHT round trip lmax=mmax=255 nlat=384 nlon=768, 97 levels, 3 scalars + 1 (tor,pol) pair"}
    st = gdhip.seeded_state(2025, L, nr, nfields=5, amp=1.0, power=1.5, zero_boundaries=False)
    for prec, peak in (("fp64", PEAK["fp64"]), ("fp32", PEAK["fp32"]), ("bf16x3", PEAK["bf16"])):
        cfg = gdhip.create_shtnskit_config(L, nlat=nlat, nlon=nlon, precision=prec)
        st[3:, :, cfg.l_values == 0] = 0.0
        nlm = cfg.nlm

        def z():
            return torch.zeros((nr, nlm), dtype=torch.float64, device="cuda")

        def ph():
            return gdhip.SHTnsPhysicalField(cfg, nr, data=torch.empty((nr, nlon, nlat), dtype=torch.float64,
                                                                      device="cuda"))

        ins = [gdhip.SHTnsSpectralField(cfg, nr, torch.from_numpy(np.ascontiguousarray(st[f].real)).cuda(),
                                        torch.from_numpy(np.ascontiguousarray(st[f].imag)).cuda()) for f in range(5)]
        outs = [gdhip.SHTnsSpectralField(cfg, nr, z(), z()) for _ in range(5)]
        phs = [ph() for _ in range(3)]
        vec = gdhip.SHTnsVectorField(cfg, nr, comps=[ph(), ph(), ph()])

        def round_trip():
            for f in range(3):
                gdhip.shtnskit_spectral_to_physical(ins[f], phs[f])
                gdhip.shtnskit_physical_to_spectral(phs[f], outs[f])
            gdhip.shtnskit_vector_synthesis(ins[3], ins[4], vec)
            gdhip.shtnskit_vector_analysis(vec, outs[3], outs[4])

        s = torch.cuda.Stream()  # the library adopts this stream: the events bracket its work
        with torch.cuda.stream(s):
            round_trip()
        torch.cuda.synchronize()
        err = 0.0
        for f in range(5):
            d = (outs[f].data_real.cpu().numpy() - st[f].real) ** 2 + (outs[f].data_imag.cpu().numpy() - st[f].imag) ** 2
            err = max(err, float(np.sqrt(d.sum() / (np.abs(st[f]) ** 2).sum())))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):
            e0.record(s)
            for _ in range(reps):
                round_trip()
            e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gf = 14 * 2.0 * nlat * nlm * nr / 1e9
        out[prec] = {"round_trip_rel_l2_max": float("%.3e" % err), "ms_per_round_trip": round(ms, 3),
                     "legendre_gflop": round(gf, 2), "legendre_tflops": round(gf / ms, 2),
                     "fraction_of_mfma_peak": round(gf / ms / peak, 4)}
        if prec == "bf16x3":
            # each product is three BF16 MFMA products (hi hi, hi lo, lo hi)
            out[prec]["issued_bf16_fraction_of_peak"] = round(3 * gf / ms / peak, 4)
        del ins, outs, phs, vec, cfg
        torch.cuda.empty_cache()
    return out


def copy_probe(nbytes):
    """One plain copy kernel (copy_kernel) of nbytes, outside the timed region."""
    import ctypes

    from gdhip._arrays import context
    from gdhip._lib import check, lib

    ctx = context(0)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().gd_device_alloc(ctx, nbytes, ctypes.byref(a)))
    check(lib().gd_device_alloc(ctx, nbytes, ctypes.byref(b)))
    check(lib().gd_device_copy(ctx, b, a, nbytes))
    check(lib().gd_ctx_synchronize(ctx))
    check(lib().gd_device_free(ctx, a))
    check(lib().gd_device_free(ctx, b))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--err-steps", type=int, default=100,
                    help="steps after which the mixed state is compared with the FP64 one and both with the "
                         "oracle (the CPU leg runs this many oracle steps; >= 11)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one plain device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--no-full", action="store_true",
                    help="skip the full-products comparison (profiling runs of the headline path)")
    ap.add_argument("--no-exp", action="store_true", help="skip the EAB2 / ERK2 measurements")
    ap.add_argument("--exp-reps", type=int, default=20)
    ap.add_argument("--only-exp", action="store_true", help="print only the EAB2 / ERK2 measurements")
    ap.add_argument("--no-c1", action="store_true", help="skip the lmax=255 transform round trip")
    ap.add_argument("--no-c3", action="store_true", help="skip the uneven shell batch")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint run (lmax=63, 20 steps)")
    ap.add_argument("--only-adjoint", action="store_true", help="print only the adjoint run")
    ap.add_argument("--only-c3", action="store_true", help="print only the uneven shell batch")
    ap.add_argument("--only-c1", action="store_true", help="print only the lmax=255 transform round trip")
    ap.add_argument("--no-restart", action="store_true", help="skip the restart file write / read timing")
    ap.add_argument("--no-bc", action="store_true", help="skip the scalar boundary-condition timing")
    ap.add_argument("--only-bc", action="store_true", help="print only the scalar boundary-condition timing")
    ap.add_argument("--derive-traffic", metavar="DIR",
                    help="only turn rocprofv3 FETCH_SIZE/WRITE_SIZE runs under DIR into the traffic JSON")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="only print a markdown table of the rocprofv3 kernel trace under DIR")
    args = ap.parse_args()
    if args.err_steps < 11:
        ap.error("--err-steps must be at least 11 (the CPU leg times 10 oracle steps after a warm-up)")
    if args.derive_traffic:
        print(json.dumps(derive_traffic(args.derive_traffic), indent=1))
        return
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the CNAB2 step on a single MI355X "
              "(the multi-GPU sharding is a design note only); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    if not (args.no_exp and args.no_c1 and args.no_bc) or args.only_bc:
        # torch's bundled HIP runtime must be the process's HIP runtime before
        # libgdhip loads, or the two do not see the same device
        import torch  # noqa: F401
    if args.only_exp:
        print(json.dumps(exp_integrators(args.exp_reps)), flush=True)
        return
    if args.only_c1:
        print(json.dumps(sht_c1(3)), flush=True)
        return
    if args.only_bc:
        print(json.dumps(scalar_bc(50)), flush=True)
        return
    if args.only_c3:
        print(json.dumps(batch_c3(64, 20, 8)), flush=True)
        return
    if args.only_adjoint:
        print(json.dumps(adjoint_c4()), flush=True)
        return

    import gdhip

    st = gdhip.seeded_state(SEED, CFG["lmax"], CFG["nr"], amp=AMP)
    kw = dict(CFG)
    lmax, nlat, nlon, nr = kw.pop("lmax"), kw.pop("nlat"), kw.pop("nlon"), kw.pop("nr")
    mixed = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp32", **kw)
    fp64 = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp64", **kw)
    if not args.no_full:
        full_mixed = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp32", full_products=True, **kw)
        full_fp64 = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp64", full_products=True, **kw)

    # accuracy: mixed vs FP64 GPU after err_steps; both against the oracle
    # after one step and after --cpu-steps steps (CPU leg below)
    fp64.set_state_complex(st)
    fp64.run(1)
    one_step64 = fp64.get_state_complex()
    fp64.run(args.err_steps - 1, graph=True)
    ref = fp64.get_state_complex()
    e_ref = fp64.energies()
    mixed.set_state_complex(st)
    mixed.run(1)
    one_step_mx = mixed.get_state_complex()
    mixed.run(args.err_steps - 1, graph=True)
    st_mixed = mixed.get_state_complex()
    err_mixed = rel_err(st_mixed, ref)
    e_mix = mixed.energies()
    if not args.no_full:
        full_mixed.set_state_complex(st)
        full_mixed.run(args.err_steps, graph=True)
        err_full = rel_err(full_mixed.get_state_complex(), st_mixed)
    # the step with BF16x3 Legendre contractions (FFT and products FP32, solves FP64)
    bf = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="bf16x3", **kw)
    bf.set_state_complex(st)
    bf.run(1)
    one_step_bf = bf.get_state_complex()
    bf.run(args.err_steps - 1, graph=True)
    st_bf = bf.get_state_complex()
    e_bf = bf.energies()

    if args.copy_probe:
        copy_probe(args.copy_probe)
    t64 = timed(fp64, st, args.steps, args.warmup)
    tmx = timed(mixed, st, args.steps, args.warmup)
    tbf = timed(bf, st, args.steps, args.warmup)
    prof = mixed.profile(5)
    prof64 = fp64.profile(5)
    if not args.no_full:
        # the same step with every product formed (no pruning of the exact zeros)
        tfm = timed(full_mixed, st, args.steps, args.warmup)
        tf64 = timed(full_fp64, st, args.steps, args.warmup)
        prof_full = full_mixed.profile(5)
    kname, roof = roofline(prof, "fp32")

    out = {
        "metric": METRIC,
        "value": round(args.steps / tmx, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (Legendre/FFT/products) + fp64 (state, radial solves)",
        "data": "synthetic (seeded (l+1)^-2 state, amplitude 1e-3)",
        "config": {"workload": "CNAB2 step lmax=mmax=127 nlat=192 nlon=384 nr=65 Ra=500 dt=1e-5, "
                               "5 fields, mixed precision", "lmax": lmax, "nlat": nlat, "nlon": nlon,
                   "nr": nr, "parallelism": "single GPU"},
        "fp64_steps_per_s": round(args.steps / t64, 2),
        "fp64_ms_per_step": round(t64 / args.steps * 1e3, 4),
        "mixed_over_fp64": round(t64 / tmx, 3),
        "accuracy": {"mixed_vs_fp64_max_rel_per_field_after_%d_steps" % args.err_steps:
                     [float("%.3e" % e) for e in err_mixed],
                     "energies_fp64": [float("%.10e" % e) for e in e_ref],
                     "energies_mixed_rel_err": [float("%.3e" % (abs(a - b) / abs(b))) for a, b in zip(e_mix, e_ref)]},
        "bf16x3": {"steps_per_s": round(args.steps / tbf, 2), "over_fp64": round(t64 / tbf, 3),
                   "vs_fp64_max_rel_per_field_after_%d_steps" % args.err_steps:
                   [float("%.3e" % e) for e in rel_err(st_bf, ref)],
                   "kernel_ms": {k: round(v, 5) for k, v in bf.profile(5).items()}},
        "kernel_ms_mixed": {k: round(v, 5) for k, v in prof.items()},
        "kernel_ms_fp64": {k: round(v, 5) for k, v in prof64.items()},
        "roofline": roof,
        "step_roofline": step_roofline(tmx / args.steps),
    }
    if not args.no_full:
        out["full_products"] = {"mixed_steps_per_s": round(args.steps / tfm, 2),
                                "vs_default_mixed_max_rel_per_field_after_%d_steps" % args.err_steps:
                                [float("%.3e" % e) for e in err_full],
                                "fp64_steps_per_s": round(args.steps / tf64, 2),
                                "kernel_ms_mixed": {k: round(v, 5) for k, v in prof_full.items()}}
    if not args.no_exp:
        try:
            out["exp_integrators"] = exp_integrators(args.exp_reps)
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["exp_integrators"] = {"error": repr(exc)[:300]}
    if not args.no_c1:
        try:
            out["sht_c1"] = sht_c1(3)
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["sht_c1"] = {"error": repr(exc)[:300]}
    if not args.no_c3:
        try:
            out["batch_c3"] = batch_c3(64, 20, 8)
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["batch_c3"] = {"error": repr(exc)[:300]}
    if not args.no_adjoint:
        try:
            out["adjoint"] = adjoint_c4()
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["adjoint"] = {"error": repr(exc)[:300]}
    if not args.no_bc:
        try:
            out["scalar_bc"] = scalar_bc(50)
            # the mixed step with flux (tau) temperature conditions on every mode
            # at both boundaries (value 0): the condition stage inside the step
            bstep = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp32", **kw)
            two = np.full(bstep.nlm, 2)
            bstep.set_temperature_bc(two, two, np.zeros((2, bstep.nlm)))
            tb = timed(bstep, st, args.steps, args.warmup)
            out["scalar_bc"]["step_with_temperature_flux_bc"] = {
                "mixed_steps_per_s": round(args.steps / tb, 2),
                "vs_default_mixed": round(tmx / tb, 3),
                "state_finite": bool(np.all(np.isfinite(bstep.get_state_complex())))}
            bstep.close()
            # the mixed step with composition (Sc = 1, no-flux conditions)
            cstep = gdhip.CNAB2Step(lmax, nlat, nlon, nr, precision="fp32", **kw)
            cstep.enable_composition(1.0)
            cstep.set_state_complex(st)
            cstep.set_composition(gdhip.seeded_state(SEED + 1, lmax, nr, nfields=1, amp=AMP)[0])
            cstep.run(args.warmup, graph=True)
            cstep.synchronize()
            t0 = time.perf_counter()
            cstep.run(args.steps, graph=True)
            cstep.synchronize()
            tc = time.perf_counter() - t0
            out["scalar_bc"]["step_with_composition"] = {
                "mixed_steps_per_s": round(args.steps / tc, 2), "vs_default_mixed": round(tmx / tc, 3),
                "composition_finite": bool(np.all(np.isfinite(cstep.get_composition())))}
            cstep.close()
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["scalar_bc"] = {"error": repr(exc)[:300]}
    if not args.no_restart:
        try:
            out["restart_io"] = restart_io(fp64, 3)
        except Exception as exc:  # a secondary measurement never hides the headline line
            out["restart_io"] = {"error": repr(exc)[:300]}
    if not args.no_cpu_baseline:
        cb, acc = cpu_baseline(st, args.err_steps, {"fp64": (one_step64, ref, e_ref),
                                                    "mixed": (one_step_mx, st_mixed, e_mix),
                                                    "bf16x3": (one_step_bf, st_bf, e_bf)})
        out["cpu_baseline"] = cb
        out["accuracy"].update(acc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
