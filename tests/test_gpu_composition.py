"""Composition as a passive scalar of the step (compositional.jl:219-260
compute_composition_nonlinear!, :318-370 conditions; simulation.jl:130-131
diffusivity Pm/Sc, :1279-1286 its CNAB2 update).

  * Enabling composition leaves the five fields bit for bit as without it.
  * Steps of C equal the oracle composite: the oracle's -u.grad T with T := C
    (the same operator, gd_oracle.c), the oracle's conditions (np_bc, default
    no flux at both boundaries), the oracle's CNAB2 rhs and band solve with
    diffusivity Pm/Sc and AB2 history -- one step and three steps, FP64, to
    1e-10 / 1e-9; the FP32 path within 1e-6 of FP64 after 5 steps.
"""
import numpy as np
import pytest

from oracle import np_bc, np_oracle

pytestmark = pytest.mark.gpu

L, NLAT, NLON, NR = 15, 24, 48, 17
PAR = dict(E=1e-3, Pm=5.0, Pr=1.0, Ra=200.0, rratio=0.35, dt=1e-4, theta=0.5, bw=3)
SC = 3.0


def _step(prec, bc_rows=1, comp=True):
    import gdhip

    s = gdhip.CNAB2Step(L, NLAT, NLON, NR, precision=prec, bc_rows=bc_rows, **PAR)
    if comp:
        s.enable_composition(SC)
    return s


def _inputs():
    import gdhip

    st = gdhip.seeded_state(41, L, NR, amp=1e-3)
    C0 = gdhip.seeded_state(42, L, NR, nfields=1, amp=1e-3, zero_boundaries=False)[0]
    return st, C0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [True, False])
def test_other_fields_unchanged(prec, graph):
    print("case: composition leaves the five fields %s graph=%s" % (prec, graph), flush=True)
    st, C0 = _inputs()
    a, b = _step(prec, comp=False), _step(prec)
    a.set_state_complex(st)
    b.set_state_complex(st)
    b.set_composition(C0)
    a.run(4, graph=graph)
    b.run(4, graph=graph)
    assert np.array_equal(a.get_state_complex(), b.get_state_complex())
    assert np.all(np.isfinite(b.get_composition()))


def _oracle_steps(O, states, C0, nsteps, bc_rows):
    """C after nsteps by the oracle composite; states[k] is the five-field
    state at step k (the composition does not feed back)."""
    r7 = O.radial_grid(NR, PAR["rratio"])
    r = r7[3]
    D1 = np_oracle.fd_dense(1, r, PAR["bw"])
    l_of_lm, _ = O.lm_tables(L, L)
    Lm, _, LU, rc = O.cnab2_matrices(r7, PAR["bw"], PAR["Pm"] / SC, PAR["dt"], PAR["theta"], L,
                                     bc_rows=bc_rows == 1)
    assert rc == 0
    model = O.Model(L, NLAT, NLON, NR, bc_rows=bc_rows, **PAR)
    nlm = C0.shape[1]
    two, zero = np.full(nlm, 2), np.zeros((2, nlm))
    C, prev = C0, None
    for k in range(nsteps):
        s5 = states[k].copy()
        s5[4] = C
        nl = model.nonlinear(s5)[4]  # -u.grad C
        cre, cim = np_bc.apply_scalar_bc(C.real, C.imag, two, two, zero, r, D1)
        Cb = cre + 1j * cim
        rhs = O.cnab2_rhs(l_of_lm, Lm, PAR["dt"], PAR["theta"], Cb, nl, nl if prev is None else prev, bc_rows=False)
        if bc_rows == 1:
            rhs[0], rhs[-1] = Cb[0], Cb[-1]
        C, prev = O.cnab2_solve(l_of_lm, LU, rhs), nl
    return C


@pytest.mark.parametrize("bc_rows,nsteps,tol", [(1, 1, 1e-10), (0, 1, 1e-10), (1, 3, 1e-9)])
def test_composition_matches_oracle_composite(oracle, bc_rows, nsteps, tol):
    print("case: composition vs oracle composite bc_rows=%d steps=%d" % (bc_rows, nsteps), flush=True)
    st, C0 = _inputs()
    s = _step("fp64", bc_rows=bc_rows)
    s.set_state_complex(st)
    s.set_composition(C0)
    states = []
    for _ in range(nsteps):
        states.append(s.get_state_complex())
        s.run(1)
    C = s.get_composition()
    Co = _oracle_steps(oracle, states, C0, nsteps, bc_rows)
    err = np.max(np.abs(C - Co)) / np.max(np.abs(Co))
    print("composition: max rel err vs oracle composite %.3e" % err, flush=True)
    assert err < tol
    assert np.max(np.abs(C - C0)) > 1e-8 * np.max(np.abs(C0))  # it moved


def test_fp32_close_to_fp64_and_errors():
    import gdhip

    print("case: composition fp32 vs fp64, argument errors", flush=True)
    st, C0 = _inputs()
    out = {}
    for prec in ("fp64", "fp32"):
        s = _step(prec)
        s.set_state_complex(st)
        s.set_composition(C0)
        s.run(5, graph=True)
        out[prec] = s.get_composition()
    err = np.max(np.abs(out["fp32"] - out["fp64"])) / np.max(np.abs(out["fp64"]))
    assert err < 1e-6, err
    full = gdhip.CNAB2Step(L, NLAT, NLON, NR, full_products=True, **PAR)
    with pytest.raises(gdhip.GDError):
        full.enable_composition(SC)
    plain = _step("fp64", comp=False)
    with pytest.raises(gdhip.GDError):
        plain.get_composition()
    s = _step("fp64")
    s.set_state_complex(st)
    with pytest.raises(gdhip.GDError):
        s.get_composition_history()  # no history before the first step


def test_restart_carries_composition(tmp_path):
    """write_restart! / read_restart! with composition_spectral and its AB2
    history: 4 steps, write, read into a fresh step, 3 more steps equal 7
    uninterrupted steps bit for bit (fields and composition)."""
    import gdhip

    print("case: restart with composition", flush=True)
    st, C0 = _inputs()
    a = _step("fp64")
    a.set_state_complex(st)
    a.set_composition(C0)
    a.run(7, graph=True)
    b = _step("fp64")
    b.set_state_complex(st)
    b.set_composition(C0)
    b.run(4, graph=True)
    cfg = gdhip.OutputConfig(output_dir=str(tmp_path), filename_prefix="gdc")
    gdhip.write_restart(b, gdhip.TimeTracker(), {"current_time": 4e-4}, cfg)
    c = _step("fp64")
    md = gdhip.read_restart(c, gdhip.TimeTracker(), str(tmp_path), 4e-4, cfg)
    assert md["current_step"] == 4
    assert np.array_equal(c.get_composition(), b.get_composition())
    assert np.array_equal(c.get_composition_history(), b.get_composition_history())
    c.run(3, graph=True)
    assert np.array_equal(c.get_state_complex(), a.get_state_complex())
    assert np.array_equal(c.get_composition(), a.get_composition())
    # a file without composition cannot restore a step that has one
    plain = _step("fp64", comp=False)
    plain.set_state_complex(st)
    plain.run(2)
    cfg2 = gdhip.OutputConfig(output_dir=str(tmp_path / "plain"), filename_prefix="gdp")
    gdhip.write_restart(plain, gdhip.TimeTracker(), {"current_time": 2e-4}, cfg2)
    with pytest.raises(gdhip.GDError):
        gdhip.read_restart(_step("fp64"), gdhip.TimeTracker(), str(tmp_path / "plain"), 2e-4, cfg2)
