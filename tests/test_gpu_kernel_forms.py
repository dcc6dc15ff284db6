"""Kernel forms of the tangential step that must give the same bits as the
forms they replaced, checked on whole steps (FP64 and FP32, with and without
composition, a small shell and lmax=127, nr=65):
  * the synthesis with grad T (and grad C) as gradient fields -- S columns
    only -- against contracting their zero toroidal columns too
    (GD_SYNTH_GRAD=0): each MFMA output column depends only on its own input
    column, and the dropped terms are exact zeros."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(var, prec, L, nlat, nlon, nr, comp):
    import gdhip

    out = []
    for flag in ("0", None):
        if flag is None:
            os.environ.pop(var, None)
        else:
            os.environ[var] = flag
        s = gdhip.CNAB2Step(L, nlat, nlon, nr, Ra=500.0, dt=1e-5, precision=prec)
        if comp:
            s.enable_composition(2.0)
        out.append(s)
    os.environ.pop(var, None)
    return out


@pytest.mark.parametrize("var", ["GD_SYNTH_GRAD"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("L,nlat,nlon,nr", [(31, 48, 96, 17), (127, 192, 384, 65)])
def test_kernel_forms_same_bits(var, prec, comp, L, nlat, nlon, nr):
    import gdhip

    print("case: %s=0 vs default %s comp=%s lmax=%d nr=%d" % (var, prec, comp, L, nr), flush=True)
    st = gdhip.seeded_state(51, L, nr, amp=1e-3)
    old, new = _pair(var, prec, L, nlat, nlon, nr, comp)
    for s in (old, new):
        s.set_state_complex(st)
        if comp:
            s.set_composition(gdhip.seeded_state(52, L, nr, nfields=1, amp=1e-3)[0])
        s.run(3, graph=True)
    assert np.array_equal(old.get_state_complex(), new.get_state_complex())
    if comp:
        assert np.array_equal(old.get_composition(), new.get_composition())
