"""Batches of independent shells (the uneven-batch configuration): shells of
different lmax, nr, grid and radius ratio advanced together by one graph per
step parity must give each shell exactly the state of its own run, in both
launch forms (a branch per shell over streams; one fused launch per kernel
form and stage), and the fused form at the full 64-shell workload also against
the FP64 oracle on sampled shells."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHELLS = [  # lmax, nr, nlat, nlon, rratio
    (15, 9, 24, 48, 0.35),
    (31, 17, 48, 96, 0.25),
    (23, 33, 36, 72, 0.40),
    (31, 33, 48, 96, 0.30),
]


FP32_SKIP = ("a batch of FP32 shells ended in a host segmentation fault inside gd_batch_run on the GPU "
             "(each shell's second stream was forked from its batch stream during capture); the cause was "
             "not found and batches no longer fork a second stream; the FP32 case stays off the GPU")


@pytest.mark.parametrize("prec", ["fp64", pytest.param("fp32", marks=pytest.mark.skip(reason=FP32_SKIP))])
def test_batch_equals_separate_runs(prec):
    import gdhip

    print("case: batch of %d shells %s" % (len(SHELLS), prec), flush=True)
    alone, batched = [], []
    for k, (L, nr, nlat, nlon, rr) in enumerate(SHELLS):
        st = gdhip.seeded_state(100 + k, L, nr, amp=1e-3)
        for group in (alone, batched):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=100.0, precision=prec)
            s.set_state_complex(st)
            group.append(s)
    for s in alone:
        s.run(6, graph=True)
    batch = gdhip.StepBatch(batched, nstreams=3, mode="streams")
    batch.run(6)
    batch.synchronize()
    for a, b in zip(alone, batched):
        assert b.step_count == a.step_count == 6
        sa, sb = a.get_state_complex(), b.get_state_complex()
        assert np.all(np.isfinite(sb))
        assert np.array_equal(sa, sb)
    batch.close()


def test_stream_form_refuses_reduced_precision_shells():
    """The stream-branch form takes FP64 shells only; FP32 shells run in the
    fused form (the refusal comes before any capture or launch)."""
    import gdhip

    s = gdhip.CNAB2Step(15, 24, 48, 9, precision="fp32")
    s.set_state_complex(gdhip.seeded_state(3, 15, 9, amp=1e-3))
    batch = gdhip.StepBatch([s], mode="streams")
    with pytest.raises(gdhip.GDError):
        batch.run(1)
    assert s.step_count == 0
    batch.close()


def test_batch_rejects_shells_at_different_steps():
    import gdhip

    a = gdhip.CNAB2Step(15, 24, 48, 9, precision="fp64")
    b = gdhip.CNAB2Step(15, 24, 48, 9, precision="fp64")
    a.set_state_complex(gdhip.seeded_state(1, 15, 9, amp=1e-3))
    b.set_state_complex(gdhip.seeded_state(2, 15, 9, amp=1e-3))
    a.run(1)
    batch = gdhip.StepBatch([a, b])
    with pytest.raises(gdhip.GDError):
        batch.run(1)


def _pairs(shells, prec, seed0=100, **kw):
    import gdhip

    alone, batched = [], []
    for k, (L, nr, nlat, nlon, rr) in enumerate(shells):
        st = gdhip.seeded_state(seed0 + k, L, nr, amp=1e-3)
        for group in (alone, batched):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=100.0, precision=prec, **kw)
            s.set_state_complex(st)
            group.append(s)
    return alone, batched


def _assert_same(alone, batched, nsteps):
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert b.step_count == a.step_count == nsteps
        sa, sb = a.get_state_complex(), b.get_state_complex()
        assert np.all(np.isfinite(sb)), k
        assert np.array_equal(sa, sb), k


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_fused_batch_equals_separate_runs(prec):
    """One launch per stage over all shells gives each shell its own bits;
    mixed synthesis tiles in one batch (tile 2 on FP32 operands only)."""
    import gdhip

    print("case: fused batch of %d shells %s" % (len(SHELLS), prec), flush=True)
    alone, batched = _pairs(SHELLS, prec)
    for s in alone:
        s.run(6, graph=True)
    batch = gdhip.StepBatch(batched, mode="fused", tiles=[0, 1, 0 if prec == "fp64" else 2, 1])
    batch.run(6)
    batch.synchronize()
    _assert_same(alone, batched, 6)
    batch.close()


def test_fused_batch_mixed_precisions_and_band_widths():
    """One fused batch of FP64, FP32 and BF16x3 shells with band half-widths
    1, 2 and 3: every kernel form gets its own launch group (precision,
    split-precision flag, band width, solve form) and each shell its own bits."""
    import gdhip

    print("case: fused batch of mixed precisions and band widths", flush=True)
    kinds = [("fp64", 3), ("fp32", 2), ("bf16x3", 3), ("fp32", 1), ("fp64", 2)]
    shells = SHELLS + [(15, 17, 24, 48, 0.3)]
    alone, batched = [], []
    for k, ((L, nr, nlat, nlon, rr), (prec, bw)) in enumerate(zip(shells, kinds)):
        st = gdhip.seeded_state(200 + k, L, nr, amp=1e-3)
        for group in (alone, batched):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=100.0, precision=prec, bw=bw)
            s.set_state_complex(st)
            group.append(s)
    for s in alone:
        s.run(5, graph=True)
    batch = gdhip.StepBatch(batched, nstreams=2, mode="fused", lanes=2)
    batch.run(5)
    batch.synchronize()
    _assert_same(alone, batched, 5)
    batch.close()


@pytest.mark.parametrize("prec,lanes", [("fp64", 2), ("fp64", 3), ("fp32", 2)])
def test_fused_lanes_equal_separate_runs(prec, lanes):
    """The fused stages split over several streams (lanes) give the same bits."""
    import gdhip

    print("case: fused batch %s lanes %d" % (prec, lanes), flush=True)
    alone, batched = _pairs(SHELLS, prec)
    for s in alone:
        s.run(5, graph=True)
    batch = gdhip.StepBatch(batched, nstreams=4, mode="fused", lanes=lanes)
    assert batch.lanes() == lanes
    batch.run(5)
    batch.synchronize()
    _assert_same(alone, batched, 5)
    batch.close()


def test_fused_lane_tuner_runs_real_steps():
    import gdhip

    print("case: fused batch lane tuner fp64", flush=True)
    alone, batched = _pairs(SHELLS, "fp64")
    batch = gdhip.StepBatch(batched, nstreams=4, mode="fused")
    lanes = batch.tune_lanes(2)
    assert lanes in (1, 2, 4)
    assert batch.lanes_ms[3] == -1.0 and all(v > 0 for v in batch.lanes_ms[:3])
    n = batched[0].step_count
    assert n == 3 * (2 * 2 + 2)
    for s in alone:
        s.run(n + 2, graph=True)
    batch.run(2)
    batch.synchronize()
    _assert_same(alone, batched, n + 2)
    batch.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_tiles_give_the_same_bits(prec):
    import gdhip

    tiles = [0, 1] + ([] if prec == "fp64" else [2])
    states = []
    for t in tiles:
        print("case: fused batch %s uniform tile %d" % (prec, t), flush=True)
        _, batched = _pairs(SHELLS[:2], prec)
        batch = gdhip.StepBatch(batched, mode="fused", tiles=[t] * 2)
        batch.run(3)
        batch.synchronize()
        states.append([s.get_state_complex() for s in batched])
        batch.close()
    for other in states[1:]:
        for a, b in zip(states[0], other):
            assert np.array_equal(a, b)


def test_fused_tuner_keeps_states_and_picks_valid_tiles():
    import gdhip

    print("case: fused batch autotuner fp32", flush=True)
    alone, batched = _pairs(SHELLS, "fp32")
    batch = gdhip.StepBatch(batched, mode="fused", tiles="tune")
    tiles = batch.tiles()
    assert all(t in (0, 1, 2) for t in tiles)
    assert len(batch.tune_ms) == 4 and all(v > 0 for v in batch.tune_ms)
    for a, b in zip(alone, batched):  # the tuner writes transform scratch only
        assert np.array_equal(a.get_state_complex(), b.get_state_complex())
    for s in alone:
        s.run(4, graph=True)
    batch.run(4)
    batch.synchronize()
    _assert_same(alone, batched, 4)
    batch.close()


def test_fused_batch_with_shells_of_their_own_form():
    """Shells whose step has no fused form (composition, full products) run
    their own steps inside the fused graph; the radial forcing and per-mode
    temperature conditions run per shell between the fused stages."""
    import gdhip

    print("case: fused batch with solo shells fp64", flush=True)
    shells = SHELLS[:3]
    kinds = [dict(radial_force=True), dict(full_products=True), {}]
    alone, batched = [], []
    for k, ((L, nr, nlat, nlon, rr), kw) in enumerate(zip(shells, kinds)):
        st = gdhip.seeded_state(300 + k, L, nr, amp=1e-3)
        for group in (alone, batched):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=400.0, precision="fp64", **kw)
            s.set_state_complex(st)
            group.append(s)
    L2, nr2 = shells[2][:2]
    C = gdhip.seeded_state(399, L2, nr2, amp=1e-3)[4]
    nlm0 = batched[0].nlm
    ti, to = np.ones(nlm0, np.int32), np.full(nlm0, 2, np.int32)  # Dirichlet inside, flux outside
    vals = np.zeros((2, nlm0))
    vals[0, 0] = 1.0
    for group in (alone, batched):
        group[2].enable_composition()
        group[2].set_composition(C)
        group[0].set_temperature_bc(ti, to, vals)
    for s in alone:
        s.run(5, graph=True)
    batch = gdhip.StepBatch(batched, mode="fused")
    batch.run(5)
    batch.synchronize()
    _assert_same(alone, batched, 5)
    assert np.array_equal(alone[2].get_composition(), batched[2].get_composition())
    batch.close()


def c3_shells(nshell=64, seed=3):
    """The uneven-batch workload: lmax uniform in {31, 47, 63, 95, 127}, nr in
    {33, 49, 65, 97}, grid 3(L+1)/2 x 3(L+1), radius ratio U[0.2, 0.5]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nshell):
        L = int(rng.choice([31, 47, 63, 95, 127]))
        nr = int(rng.choice([33, 49, 65, 97]))
        out.append((L, nr, 3 * (L + 1) // 2, 3 * (L + 1), float(rng.uniform(0.2, 0.5))))
    return out


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_fused_batch_c3_64_shells(oracle, prec, tol):
    """The full 64-shell workload in the fused form (tiles from the tuning
    table), FP64 and mixed: every shell equal to its own step graph after 4
    steps, and the first two steps (the eager AB1 bootstrap, then the first
    fused graph) of three sampled shells (the smallest, a middle and the
    largest) within tol of the FP64 oracle per field."""
    import gdhip

    shells = c3_shells()
    print("case: fused batch C3 64 shells %s" % prec, flush=True)
    alone, batched = [], []
    for k, (L, nr, nlat, nlon, rr) in enumerate(shells):
        st = oracle.seeded_state(500 + k, L, L, nr, amp=1e-3)
        for group in (alone, batched):
            s = gdhip.CNAB2Step(L, nlat, nlon, nr, rratio=rr, Ra=500.0, precision=prec)
            s.set_state_complex(st)
            group.append(s)
    order = sorted(range(len(shells)), key=lambda k: shells[k][0] * shells[k][0] * shells[k][1])
    sample = [order[0], order[len(order) // 2], order[-1]]
    batch = gdhip.StepBatch(batched, mode="fused")
    batch.run(2)
    batch.synchronize()
    oracle.set_threads(8)
    for k in sample:
        L, nr, nlat, nlon, rr = shells[k]
        o = oracle.Model(L, nlat, nlon, nr, Ra=500.0, rratio=rr)
        ref, nlp = oracle.seeded_state(500 + k, L, L, nr, amp=1e-3), None
        for step in range(2):
            ref, _, nlp = o.step(ref, nlp, step)
        got = batched[k].get_state_complex()
        err = [np.abs(got[f] - ref[f]).max() / max(np.abs(ref[f]).max(), 1e-300) for f in range(5)]
        print("shell", k, shells[k][:2], "rel err vs oracle", err, flush=True)
        assert max(err) < tol
    batch.run(2)
    for s in alone:
        s.run(4, graph=True)
    batch.synchronize()
    _assert_same(alone, batched, 4)
    batch.close()
