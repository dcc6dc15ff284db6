"""Batches of independent shells (the uneven-batch configuration): shells of
different lmax, nr, grid and radius ratio advanced together by one graph per
step parity must give each shell exactly the state of its own run."""
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
    batch = gdhip.StepBatch(batched, nstreams=3)
    batch.run(6)
    batch.synchronize()
    for a, b in zip(alone, batched):
        assert b.step_count == a.step_count == 6
        sa, sb = a.get_state_complex(), b.get_state_complex()
        assert np.all(np.isfinite(sb))
        assert np.array_equal(sa, sb)
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
