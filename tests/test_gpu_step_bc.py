"""Per-mode temperature conditions inside the step (gd_step_set_temperature_bc:
apply_temperature_boundary_conditions_spectral!, thermal.jl:262-313, between
the nonlinear terms and the implicit step, simulation.jl:1151-1330).

  * Dirichlet 0 on every mode is the step's default reading of the reference
    (temperature boundary rows read as 0): the two give bitwise equal states,
    FP64 and FP32, with and without Dirichlet rows in S_l, graph and eager.
  * Flux conditions: one AB1 step equals the oracle composite -- the GPU's
    nonlinear terms, the oracle's boundary conditions (oracle/np_bc.py) on the
    temperature, the oracle's CNAB2 rhs and band solve (gd_oracle.c) with the
    stored boundary values in the Dirichlet rows -- to 1e-10, and leaves the
    other fields bitwise as the default step does.
"""
import numpy as np
import pytest

from oracle import np_bc, np_oracle

pytestmark = pytest.mark.gpu

L, NLAT, NLON, NR = 15, 24, 48, 17
PAR = dict(E=1e-3, Pm=5.0, Pr=1.0, Ra=200.0, rratio=0.35, dt=1e-4, theta=0.5, bw=3)


def _step(prec, bc_rows):
    import gdhip

    return gdhip.CNAB2Step(L, NLAT, NLON, NR, precision=prec, bc_rows=bc_rows, **PAR)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("bc_rows", [1, 0])
@pytest.mark.parametrize("graph", [True, False])
def test_dirichlet_zero_equals_default(prec, bc_rows, graph):
    import gdhip

    print("case: step temperature bc dirichlet-0 %s bc_rows=%d graph=%s" % (prec, bc_rows, graph), flush=True)
    st = gdhip.seeded_state(31, L, NR, amp=1e-3, zero_boundaries=False)
    a, b = _step(prec, bc_rows), _step(prec, bc_rows)
    nlm = a.nlm
    b.set_temperature_bc(np.ones(nlm), np.ones(nlm), np.zeros((2, nlm)))
    for s in (a, b):
        s.set_state_complex(st)
        s.run(4, graph=graph)
    assert np.array_equal(a.get_state_complex(), b.get_state_complex())
    # back to the default
    b.set_temperature_bc()
    a.run(2, graph=graph)
    b.run(2, graph=graph)
    assert np.array_equal(a.get_state_complex(), b.get_state_complex())


@pytest.mark.parametrize("bc_rows", [0, 1])
def test_flux_conditions_match_oracle_composite(oracle, bc_rows):
    import gdhip

    print("case: step temperature bc flux vs oracle composite bc_rows=%d" % bc_rows, flush=True)
    st = gdhip.seeded_state(32, L, NR, amp=1e-3, zero_boundaries=False)
    s, ref = _step("fp64", bc_rows), _step("fp64", bc_rows)
    nlm = s.nlm
    rng = np.random.default_rng(33)
    ti = rng.integers(1, 3, nlm)
    to = rng.integers(1, 3, nlm)
    vals = 1e-3 * rng.standard_normal((2, nlm))
    s.set_state_complex(st)
    ref.set_state_complex(st)
    nl = s.nonlinear()
    s.set_temperature_bc(ti, to, vals)
    s.run(1, graph=False)
    ref.run(1, graph=False)
    out, out_ref = s.get_state_complex(), ref.get_state_complex()
    assert np.array_equal(out[:4], out_ref[:4])

    O = oracle
    r7 = O.radial_grid(NR, PAR["rratio"])
    r = r7[3]
    D1 = np_oracle.fd_dense(1, r, PAR["bw"])
    ure, uim = np_bc.apply_scalar_bc(st[4].real, st[4].imag, ti, to, vals, r, D1)
    uT = ure + 1j * uim
    l_of_lm, _ = O.lm_tables(L, L)
    Lm, _, LU, rc = O.cnab2_matrices(r7, PAR["bw"], PAR["Pm"] / PAR["Pr"], PAR["dt"], PAR["theta"], L,
                                     bc_rows=bc_rows in (1, 2))
    assert rc == 0
    rhs = O.cnab2_rhs(l_of_lm, Lm, PAR["dt"], PAR["theta"], uT, nl[4], nl[4], bc_rows=False)
    if bc_rows in (1, 2):  # identity rows of S_l carry the stored boundary values
        rhs[0], rhs[-1] = uT[0], uT[-1]
    xT = O.cnab2_solve(l_of_lm, LU, rhs)
    err = np.max(np.abs(out[4] - xT)) / np.max(np.abs(xT))
    print("flux bc step: max rel err of T vs oracle composite %.3e" % err, flush=True)
    assert err < 1e-10
    # the conditions changed the temperature (the test is not vacuous)
    assert np.max(np.abs(out[4] - out_ref[4])) > 1e-6 * np.max(np.abs(out_ref[4]))
