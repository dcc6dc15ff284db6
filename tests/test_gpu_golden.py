"""GPU parity against the committed golden fixtures (tests/golden/*.npz):
the FP64 transforms, the CNAB2 operators and banded solve, and two 5-field
steps, each through the C ABI.  Bars: FP64 1e-12 relative (transforms
1e-13 absolute at unit amplitude); FP32 step 1e-6 relative per field."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_sht_fp64_matches_fixture():
    import gdhip

    print("case: golden sht lmax=7 fp64", flush=True)
    g = _load("sht_l7")
    cfg = gdhip.create_shtnskit_config(7, nlat=12, nlon=24, precision="fp64")
    a = g["a"][None]
    spec = gdhip.SHTnsSpectralField(cfg, 1, np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag))
    phys = gdhip.SHTnsPhysicalField(cfg, 1)
    gdhip.shtnskit_spectral_to_physical(spec, phys)
    assert np.abs(phys.data[0] - g["f"]).max() < 1e-13
    back = gdhip.SHTnsSpectralField(cfg, 1)
    gdhip.shtnskit_physical_to_spectral(gdhip.SHTnsPhysicalField(cfg, 1, g["f"][None]), back)
    assert np.abs(back.complex()[0] - g["a_back"]).max() < 1e-13
    tor, pol = g["tor"][None], g["pol"][None]
    ts = gdhip.SHTnsSpectralField(cfg, 1, np.ascontiguousarray(tor.real), np.ascontiguousarray(tor.imag))
    ps = gdhip.SHTnsSpectralField(cfg, 1, np.ascontiguousarray(pol.real), np.ascontiguousarray(pol.imag))
    vec = gdhip.SHTnsVectorField(cfg, 1)
    gdhip.shtnskit_vector_synthesis(ts, ps, vec)
    assert np.abs(vec.θ_component.data[0] - g["vt"]).max() < 1e-13
    assert np.abs(vec.φ_component.data[0] - g["vp"]).max() < 1e-13


def test_cnab2_operators_and_solve_match_fixture():
    import gdhip

    print("case: golden cnab2 operators lmax=127 nr=65", flush=True)
    g = _load("band_c2")
    cfg = gdhip.create_shtnskit_config(127, nlat=192, nlon=384)
    dom = gdhip.create_radial_domain(65, 0.35)
    assert np.allclose(dom.r, g["r"], rtol=4e-16, atol=0)  # host libm may differ by an ulp
    ls = list(g["ls"])
    for k, nu in enumerate((5.0, 1.0, 5.0)):
        M = gdhip.create_shtns_timestepping_matrices(cfg, dom, nu, 1e-5, theta=0.5, bw=3, bc_rows=True)
        L, S, LU = M._fetch()
        assert _rel(L[ls], g["L%d" % k]) < 1e-14
        assert _rel(S[ls], g["S%d" % k]) < 1e-14
        assert _rel(LU[ls], g["LU%d" % k]) < 1e-13
        lus = [gdhip.BandedLU(np.ascontiguousarray(LU[l]), 3, 65) for l in ls]
        for j, lu in enumerate(lus):
            x = np.zeros(65)
            gdhip.solve_banded(x, lu, g["rhs"])
            assert _rel(x, g["x%d" % k][j]) < 1e-12
        M.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_two_steps_match_fixture(prec, tol):
    import gdhip

    print("case: golden two steps lmax=7 nr=9 %s" % prec, flush=True)
    g = _load("step_l7")
    s = gdhip.CNAB2Step(7, 12, 24, 9, Ra=100.0, precision=prec)
    s.set_state_complex(g["state0"])
    if prec == "fp64":
        nl = s.nonlinear()
        assert max(_rel(nl[k], g["nl0"][k]) for k in range(5)) < 1e-12
    s.run(1)
    st1 = s.get_state_complex()
    s.run(1, graph=True)
    st2 = s.get_state_complex()
    e1 = [_rel(st1[k], g["state1"][k]) for k in range(5)]
    e2 = [_rel(st2[k], g["state2"][k]) for k in range(5)]
    print("rel err step1", e1, "step2", e2)
    assert max(e1) < tol and max(e2) < tol
    if prec == "fp64":
        assert _rel(s.energies(), g["energies2"]) < 1e-12
    s.close()
