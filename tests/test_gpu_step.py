"""GPU parity of the device-resident CNAB2 step against the FP64 oracle.

Reference: the step of scripts/bench_timestepping.jl:6-18 (nonlinear terms of
thermal.jl:149-190, velocity.jl:312-332, magnetic.jl:131-159, then
apply_master_implicit_step!, simulation.jl:1151-1330) with the semantics the
oracle documents.  Bars: FP64 path max relative error per field <= 1e-10;
FP32 mixed path <= 1e-6 per field of the state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C0 = dict(lmax=31, nlat=48, nlon=96, nr=33)


def _relf(a, b):
    return [np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300) for k in range(5)]


def _make(oracle, prec, lmax, nlat, nlon, nr, Ra=100.0, amp=1e-3, seed=2024, **kw):
    import gdhip

    st = oracle.seeded_state(seed, lmax, lmax, nr, amp=amp)
    g = gdhip.CNAB2Step(lmax, nlat, nlon, nr, Ra=Ra, precision=prec, **kw)
    g.set_state_complex(st)
    o = oracle.Model(lmax, nlat, nlon, nr, Ra=Ra, **kw)
    return st, g, o


def test_state_roundtrip_and_energies(oracle):
    st, g, o = _make(oracle, "fp64", **C0)
    back = g.get_state_complex()
    assert np.array_equal(back, st)
    e = g.energies()
    eo = o.energies(st)
    assert np.abs(e - eo).max() < 1e-13 * np.abs(eo).max()


@pytest.mark.parametrize("coriolis_index,radials", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_nonlinear_fp64_matches_oracle(oracle, coriolis_index, radials):
    print("case: nonlinear fp64 C0 coriolis_index=%d radials=%d" % (coriolis_index, radials), flush=True)
    st, g, o = _make(oracle, "fp64", coriolis_index=coriolis_index, radials=radials, **C0)
    nl = g.nonlinear()
    ref = o.nonlinear(st)
    err = _relf(nl, ref)
    print("nl rel err", err)
    assert max(err) < 1e-10


def test_nonlinear_fp32_close_to_oracle(oracle):
    st, g, o = _make(oracle, "fp32", **C0)
    err = _relf(g.nonlinear(), o.nonlinear(st))
    print("nl fp32 rel err", err)
    assert max(err) < 1e-4


def test_steps_fp64_match_oracle_c0(oracle):
    print("case: 3 steps fp64 C0", flush=True)
    st, g, o = _make(oracle, "fp64", **C0)
    s, nlp = st, None
    for k in range(3):
        s, nl, nlp = o.step(s, nlp, k)
    g.run(3)
    err = _relf(g.get_state_complex(), s)
    print("state rel err", err)
    assert max(err) < 1e-10
    assert g.step_count == 3


@pytest.mark.parametrize("bc_rows,radials", [(0, 0), (2, 0), (1, 1), (0, 1)])
def test_step_variants_fp64_match_oracle(oracle, bc_rows, radials):
    """The alternative readings (reference operator without boundary rows,
    temperature rows only, synthesised radial components) for one step."""
    print("case: 1 step fp64 C0 bc_rows=%d radials=%d" % (bc_rows, radials), flush=True)
    st, g, o = _make(oracle, "fp64", bc_rows=bc_rows, radials=radials, **C0)
    s, nl, nlp = o.step(st, None, 0)
    g.run(1)
    err = _relf(g.get_state_complex(), s)
    print("state rel err", err)
    assert max(err) < 1e-10


def test_graph_replay_matches_eager(oracle):
    print("case: graph replay C0", flush=True)
    st, g1, _ = _make(oracle, "fp64", **C0)
    import gdhip

    g2 = gdhip.CNAB2Step(C0["lmax"], C0["nlat"], C0["nlon"], C0["nr"], precision="fp64")
    g2.set_state_complex(st)
    for k in range(5):  # replays with eager work and syncs in between
        g1.run(1)
        g2.run(1, graph=True)
        g2.synchronize()
        assert np.array_equal(g1.get_state_complex(), g2.get_state_complex()), "step %d" % k


def test_steps_fp32_within_1e6_of_oracle(oracle):
    print("case: 5 steps fp32 C0", flush=True)
    st, g, o = _make(oracle, "fp32", **C0)
    s, nlp = st, None
    for k in range(5):
        s, nl, nlp = o.step(s, nlp, k)
    g.run(5)
    err = _relf(g.get_state_complex(), s)
    print("fp32 state rel err", err)
    assert max(err) < 1e-6


def test_steps_bf16x3_close_to_oracle(oracle):
    """The step with BF16x3 Legendre contractions (split-precision BF16 MFMA,
    FP32 accumulation; FFT and products FP32, solves FP64): per-field error
    against the oracle after 5 steps, bar 1e-4 (each contraction carries
    ~2^-16 relative error)."""
    print("case: 5 steps bf16x3 C0", flush=True)
    st, g, o = _make(oracle, "bf16x3", **C0)
    s, nlp = st, None
    for k in range(5):
        s, nl, nlp = o.step(s, nlp, k)
    g.run(1)
    g.run(4, graph=True)
    err = _relf(g.get_state_complex(), s)
    print("bf16x3 state rel err", err)
    assert max(err) < 1e-4


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-6), ("bf16x3", 1e-4)])
def test_steps_c2_reduced_precision_against_oracle(oracle, prec, tol):
    """The C2 workload (lmax=127, 192x384, nr=65, Ra=500) on the mixed and the
    BF16x3 paths -- the twisted CNAB2 solve with paired columns, the FP32 and
    split-BF16 transforms -- against the FP64 oracle after 3 steps, per field."""
    print("case: 3 steps %s C2 vs oracle" % prec, flush=True)
    oracle.set_threads(8)
    st, g, o = _make(oracle, prec, 127, 192, 384, 65, Ra=500.0)
    s, nlp = st, None
    for k in range(3):
        s, nl, nlp = o.step(s, nlp, k)
    g.run(3, graph=True)
    err = _relf(g.get_state_complex(), s)
    print("%s C2 state rel err" % prec, err)
    assert max(err) < tol


def test_step_fp64_matches_oracle_c2_shape(oracle):
    print("case: 1 step fp64 lmax=127 nr=65", flush=True)
    st, g, o = _make(oracle, "fp64", lmax=127, nlat=192, nlon=384, nr=65, Ra=500.0)
    s, nl, nlp = o.step(st, None, 0)
    g.run(1)
    err = _relf(g.get_state_complex(), s)
    print("C2 state rel err", err)
    assert max(err) < 1e-10


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_tangential_path_equals_full_products(oracle, prec, tol):
    """With radials = 0 the default step transforms only u and grad T (the
    other products are identically zero); full_products = 1 forms them all.
    Both must give the same nonlinear terms and states."""
    import gdhip

    print("case: tangential vs full products C0 %s" % prec, flush=True)
    st = oracle.seeded_state(2024, 31, 31, 33, amp=1e-3)
    a = gdhip.CNAB2Step(31, 48, 96, 33, Ra=100.0, precision=prec)
    b = gdhip.CNAB2Step(31, 48, 96, 33, Ra=100.0, precision=prec, full_products=True)
    for g in (a, b):
        g.set_state_complex(st)
    na, nb = a.nonlinear(), b.nonlinear()
    err = _relf(na, nb)
    print("nl rel diff", err)
    assert max(err) < tol and np.all(na[2:4] == 0)
    for g in (a, b):
        g.run(1)
        g.run(3, graph=True)
    err = _relf(a.get_state_complex(), b.get_state_complex())
    print("state rel diff after 4 steps", err)
    assert max(err) < tol
    o = oracle.Model(31, 48, 96, 33, Ra=100.0)
    s = st
    s, nl, prev = o.step(s, None, 0)
    for k in range(1, 4):
        s, nl, prev = o.step(s, prev, k)
    err = _relf(a.get_state_complex(), s)
    print("tangential vs oracle after 4 steps", err)
    assert max(err) < (1e-10 if prec == "fp64" else 1e-6)

