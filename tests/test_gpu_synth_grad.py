"""The tangential step's synthesis with grad T (and grad C) as gradient
fields -- S columns only, both components from them -- gives the same bits as
the form that also contracts their zero toroidal columns (GD_SYNTH_GRAD=0):
each MFMA output column depends only on its own input column, and the
dropped terms are exact zeros.  Checked on whole steps, FP64 and FP32, with
and without composition, at a small shell and at lmax=127, nr=65."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(prec, L, nlat, nlon, nr, comp):
    import gdhip

    out = []
    for flag in ("0", None):
        if flag is None:
            os.environ.pop("GD_SYNTH_GRAD", None)
        else:
            os.environ["GD_SYNTH_GRAD"] = flag
        s = gdhip.CNAB2Step(L, nlat, nlon, nr, Ra=500.0, dt=1e-5, precision=prec)
        if comp:
            s.enable_composition(2.0)
        out.append(s)
    os.environ.pop("GD_SYNTH_GRAD", None)
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("L,nlat,nlon,nr", [(31, 48, 96, 17), (127, 192, 384, 65)])
def test_gradient_synthesis_same_bits(prec, comp, L, nlat, nlon, nr):
    import gdhip

    print("case: gradient synthesis %s comp=%s lmax=%d nr=%d" % (prec, comp, L, nr), flush=True)
    st = gdhip.seeded_state(51, L, nr, amp=1e-3)
    old, new = _pair(prec, L, nlat, nlon, nr, comp)
    for s in (old, new):
        s.set_state_complex(st)
        if comp:
            s.set_composition(gdhip.seeded_state(52, L, nr, nfields=1, amp=1e-3)[0])
        s.run(3, graph=True)
    assert np.array_equal(old.get_state_complex(), new.get_state_complex())
    if comp:
        assert np.array_equal(old.get_composition(), new.get_composition())
