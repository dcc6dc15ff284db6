"""The radial forcing of the poloidal equation (radial_force = 1): the
buoyancy F_r = (Pm/E)(Pm/Pr) Ra r T (velocity.jl:534-537) enters nl(u_pol)
as the poloidal scalar r^2 F_r / (l(l+1)) whose radial component is F_r
(u_r = l(l+1) P / r^2, magnetic.jl:257-259).  GPU against the oracle (FP64
1e-10), Ra changing the state only with the forcing on, and the adjoint of
the forced step against the oracle adjoint."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C0 = dict(lmax=31, nlat=48, nlon=96, nr=33)


def _relf(a, b):
    return [np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300) for k in range(5)]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_forced_steps_match_oracle(oracle, prec, tol):
    import gdhip

    print("case: radial forcing C0 %s" % prec, flush=True)
    st = oracle.seeded_state(31, C0["lmax"], C0["lmax"], C0["nr"], amp=1e-3)
    g = gdhip.CNAB2Step(**C0, Ra=500.0, precision=prec, radial_force=True)
    g.set_state_complex(st)
    o = oracle.Model(**C0, Ra=500.0, radial_force=1)
    nl = g.nonlinear()
    err = _relf(nl, o.nonlinear(st))
    print("forced nl rel err", err)
    assert max(err) < (tol if prec == "fp64" else 1e-4)
    s, nlp = st, None
    for k in range(3):
        s, nlo, nlp = o.step(s, nlp, k)
    g.run(1)
    g.run(2, graph=True)
    err = _relf(g.get_state_complex(), s)
    print("forced state rel err after 3 steps", err)
    assert max(err) < tol


def test_ra_changes_the_state_only_with_the_forcing(oracle):
    import gdhip

    print("case: Ra dependence with and without the radial forcing", flush=True)
    st = oracle.seeded_state(32, C0["lmax"], C0["lmax"], C0["nr"], amp=1e-3)
    out = {}
    for rf in (False, True):
        for Ra in (100.0, 500.0):
            g = gdhip.CNAB2Step(**C0, Ra=Ra, precision="fp64", radial_force=rf)
            g.set_state_complex(st)
            g.run(3)
            out[rf, Ra] = g.get_state_complex()
    assert np.array_equal(out[False, 100.0], out[False, 500.0])
    d = np.abs(out[True, 100.0][1] - out[True, 500.0][1]).max()
    print("u_pol change between Ra=100 and Ra=500 with the forcing: %.3e" % d)
    assert d > 1e-6 * np.abs(out[True, 500.0][1]).max()


def test_forced_adjoint_matches_oracle_adjoint(oracle):
    import gdhip
    from oracle import np_adjoint, np_oracle

    print("case: adjoint of the forced step vs oracle adjoint", flush=True)
    kw = dict(lmax=7, nlat=12, nlon=24, nr=9)
    a = np_adjoint.NpAdjoint(np_oracle.NpModel(**kw, dt=1e-3, Ra=400.0, radial_force=1))
    g = gdhip.CNAB2Step(**kw, dt=1e-3, Ra=400.0, precision="fp64", radial_force=True)
    rng = np.random.default_rng(6)
    x = oracle.seeded_state(33, 7, 7, 9, amp=1.0, zero_boundaries=False)
    mu = rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)
    g.set_state_complex(x)
    got = gdhip.nonlinear_adjoint(g, mu)
    ref = a.nonlinear_vjp(x, mu)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("forced nonlinear adjoint rel err", err)
    assert err < 1e-10
    target = rng.standard_normal((24, 12))
    g.set_state_complex(x)
    adj = gdhip.CNAB2Adjoint(g, 3, every=2)
    adj.forward()
    adj.misfit(target, kind="u_r", level=7)
    grad = adj.backward()
    xs = a.run(x, 3)
    gr, _ = a.gradient(x, 3, a.misfit_seed(xs[-1], "u_r", 7, target))
    err = np.abs(grad - gr).max() / np.abs(gr).max()
    print("forced run gradient rel err", err)
    assert err < 1e-10
    # the forcing carries the u_r cotangent into T
    assert np.abs(grad[4]).max() > 0
