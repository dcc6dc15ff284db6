"""Kernel forms of the tangential step checked against the forms they
replaced, on whole steps (FP64 and FP32, with and without
composition, a small shell and lmax=127, nr=65):
  * same bits: the synthesis with grad T (and grad C) as gradient fields -- S columns
    only -- against contracting their zero toroidal columns too
    (GD_SYNTH_GRAD=0): each MFMA output column depends only on its own input
    column, and the dropped terms are exact zeros;
  * equal to rounding: the twisted CNAB2 column solve (a top and a bottom
    wave per 64 columns meeting at an interface system; both paths use it)
    against one lane per column (GD_CHAIN_TWIST=0): the same linear system
    solved in another order, so equal to rounding (max relative difference
    <= 1e-6 after 3 steps of the FP32 path, whose nonlinear terms are FP32,
    and <= 1e-13 on the FP64 path);
  * equal to rounding: the twisted solve with the re and im columns of an
    (idx, field) in one lane (the default) against one column per lane
    (GD_CHAIN_PAIR=0);
  * same bits: the paired twisted solve with the first rows of each half's
    forward results in LDS (the default) against all of them through the
    state buffer (GD_CHAIN_YLDS=0): the same operations, another store;
  * same bits: the one-lane CNAB2 column kernel with 64-bit offsets
    (GD_CHAIN_OFF64=1, the form shells with SYNV_COLS*nlm*nr >= 2^32 run)
    against the 32-bit offsets used where they fit, both with
    GD_CHAIN_TWIST=0 (the twisted solve takes 32-bit offsets only)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


ALT = {"GD_SYNTH_GRAD": "0", "GD_CHAIN_TWIST": "0", "GD_CHAIN_OFF64": "1", "GD_CHAIN_PAIR": "0",
       "GD_CHAIN_YLDS": "0"}  # non-default forms
BASE = {"GD_CHAIN_OFF64": {"GD_CHAIN_TWIST": "0"}}  # set for both sides of a pair
BITWISE = ("GD_SYNTH_GRAD", "GD_CHAIN_YLDS", "GD_CHAIN_OFF64")


def _pair(var, prec, L, nlat, nlon, nr, comp):
    import gdhip

    base = BASE.get(var, {})
    out = []
    for flag in (ALT[var], None):
        os.environ.update(base)
        if flag is None:
            os.environ.pop(var, None)
        else:
            os.environ[var] = flag
        s = gdhip.CNAB2Step(L, nlat, nlon, nr, Ra=500.0, dt=1e-5, precision=prec)
        if comp:
            s.enable_composition(2.0)
        out.append(s)
    for k in [var] + list(base):
        os.environ.pop(k, None)
    return out


@pytest.mark.parametrize("var", sorted(ALT))
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("L,nlat,nlon,nr", [(31, 48, 96, 17), (127, 192, 384, 65)])
def test_kernel_forms_agree(var, prec, comp, L, nlat, nlon, nr):
    import gdhip

    print("case: %s=%s vs default %s comp=%s lmax=%d nr=%d" % (var, ALT[var], prec, comp, L, nr), flush=True)
    st = gdhip.seeded_state(51, L, nr, amp=1e-3)
    old, new = _pair(var, prec, L, nlat, nlon, nr, comp)
    for s in (old, new):
        s.set_state_complex(st)
        if comp:
            s.set_composition(gdhip.seeded_state(52, L, nr, nfields=1, amp=1e-3)[0])
        s.run(3, graph=True)
    a, b = old.get_state_complex(), new.get_state_complex()
    bitwise = var in BITWISE
    if bitwise:
        assert np.array_equal(a, b)
    else:
        tol = 1e-13 if prec == "fp64" else 1e-6
        for f in range(5):
            scale = max(np.abs(a[f]).max(), 1e-300)
            assert np.abs(a[f] - b[f]).max() <= tol * scale, (f, np.abs(a[f] - b[f]).max() / scale)
    if comp:
        ca, cb = old.get_composition(), new.get_composition()
        if bitwise:
            assert np.array_equal(ca, cb)
        else:
            assert np.abs(ca - cb).max() <= (1e-13 if prec == "fp64" else 1e-6) * np.abs(ca).max()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("pair", ["2", "0"])
@pytest.mark.parametrize("nr,bw", [(8, 3), (9, 3), (12, 3), (13, 3), (6, 2), (11, 2), (4, 1), (17, 1), (33, 2)])
def test_twisted_solve_edge_shapes(nr, bw, pair, prec):
    """The twisted solve (both column layouts) at the smallest shell it takes
    (nr = 2 bw + 2), odd and even nr where only the guarded tail blocks run,
    and every band width it is built for (bw 1..3), with composition: equal
    to the one-lane solve after 3 steps, to the FP32 path's rounding bound
    (1e-6) or the FP64 path's (1e-13)."""
    import gdhip

    print("case: twisted solve edge shape nr=%d bw=%d pair=%s %s" % (nr, bw, pair, prec), flush=True)
    tol = 1e-6 if prec == "fp32" else 1e-13
    os.environ["GD_CHAIN_PAIR"] = pair
    L, nlat, nlon = 15, 24, 48
    st = gdhip.seeded_state(53, L, nr, amp=1e-3)
    C0 = gdhip.seeded_state(54, L, nr, nfields=1, amp=1e-3)[0]
    out = []
    for flag in ("0", None):
        if flag is None:
            os.environ.pop("GD_CHAIN_TWIST", None)
        else:
            os.environ["GD_CHAIN_TWIST"] = flag
        s = gdhip.CNAB2Step(L, nlat, nlon, nr, Ra=500.0, dt=1e-5, precision=prec, bw=bw)
        s.enable_composition(2.0)
        out.append(s)
    os.environ.pop("GD_CHAIN_TWIST", None)
    os.environ.pop("GD_CHAIN_PAIR", None)
    for s in out:
        s.set_state_complex(st)
        s.set_composition(C0)
        s.run(3, graph=True)
    a, b = out[0].get_state_complex(), out[1].get_state_complex()
    for f in range(5):
        scale = max(np.abs(a[f]).max(), 1e-300)
        assert np.abs(a[f] - b[f]).max() <= tol * scale, (f, np.abs(a[f] - b[f]).max() / scale)
    ca, cb = out[0].get_composition(), out[1].get_composition()
    assert np.abs(ca - cb).max() <= tol * np.abs(ca).max()
    assert np.all(np.isfinite(b))
    # the two solves round differently somewhere: the twisted form did run
    assert not np.array_equal(a, b)
