"""Pins of the oracle's nonlinear terms and full step (CPU).

The C oracle is compared with the independent numpy/scipy restatement
(oracle/np_oracle.py) for every flag of the adopted reading, plus checks the
reference's data flow fixes: with the radial components left 0
(shtnskit_field_functions.jl:264-278 never writes them) the tangential
u x B, u x omega and j x B vanish, so nl_B == 0 and the Lorentz force has no
effect; energies follow magnetic.jl:504-543 and thermal.jl:413-438.
"""
import numpy as np
import pytest

TINY = dict(lmax=7, nlat=12, nlon=24, nr=9)


def _relf(a, b):
    return max(np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300) for k in range(5))


@pytest.mark.parametrize("radials,coriolis_index", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_nonlinear_equals_numpy_restatement(oracle, radials, coriolis_index):
    from oracle import np_oracle

    st = oracle.seeded_state(7, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1e-3)
    m = oracle.Model(**TINY, radials=radials, coriolis_index=coriolis_index)
    n = np_oracle.NpModel(**TINY, radials=radials, coriolis_index=coriolis_index)
    assert _relf(m.nonlinear(st), n.nonlinear(st)) < 1e-12


@pytest.mark.parametrize("bc_rows,radials", [(1, 0), (0, 0), (2, 0), (1, 1)])
def test_two_steps_equal_numpy_restatement(oracle, bc_rows, radials):
    from oracle import np_oracle

    st = oracle.seeded_state(7, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1e-3)
    m = oracle.Model(**TINY, bc_rows=bc_rows, radials=radials)
    n = np_oracle.NpModel(**TINY, bc_rows=bc_rows, radials=radials)
    a, _, ap = m.step(st, None, 0)
    b, _, bp = n.step(st, None, 0)
    assert _relf(a, b) < 1e-10
    a, _, _ = m.step(a, ap, 1)
    b, _, _ = n.step(b, bp, 1)
    assert _relf(a, b) < 1e-10


def test_reference_data_flow_without_radials(oracle):
    st = oracle.seeded_state(11, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1e-3)
    m = oracle.Model(**TINY, radials=False)
    nl = m.nonlinear(st)
    assert np.all(nl[2] == 0) and np.all(nl[3] == 0)  # curl(u x B) of tangential u, B
    # Lorentz and u x omega are radial: scaling B or changing E's advective part leaves nl_u
    st2 = st.copy()
    st2[2:4] *= 3.0
    assert np.array_equal(m.nonlinear(st2)[:2], nl[:2])


def test_products_brute_force_one_level(oracle):
    """gdo_physical_level products against the formulas of velocity.jl:512-527,
    664-670, magnetic.jl:399-401, scalar_field_common.jl:401-407, evaluated in numpy."""
    p = dict(TINY, radials=True, coriolis_index=1)
    m = oracle.Model(**p)
    st = oracle.seeded_state(3, p["lmax"], p["lmax"], p["nr"], amp=1e-2)
    F, P = m.physical_level(st, 4)
    ut, up, wt, wp, Bt, Bp, jt, jp, gt, gp, ur, wr, Br, jr, gr = F
    sht = oracle.SHT(p["lmax"], p["lmax"], p["nlat"], p["nlon"])
    th = sht.grid()[2][None, :]
    ro = 5.0 / 1e-3
    at = -ro * (up * wr - ur * wp) + ro * np.cos(th) * up + ro * (jp * Br - jr * Bp)
    ap = -ro * (ur * wt - ut * wr) - ro * (np.cos(th) * ut + np.sin(th) * ur) + ro * (jr * Bt - jt * Br)
    ref = [at, ap, up * Br - ur * Bp, ur * Bt - ut * Br, -(ur * gr + ut * gt + up * gp)]
    for k in range(5):
        assert np.abs(P[k] - ref[k]).max() < 1e-12 * max(1.0, np.abs(ref[k]).max())


def test_energies_single_mode(oracle):
    m = oracle.Model(**TINY)
    n = m.nlm
    st = np.zeros((5, TINY["nr"], n), complex)
    lo, _ = oracle.lm_tables(TINY["lmax"], TINY["lmax"])
    q = 7
    st[0, 2, q] = 3.0 + 4.0j   # u_tor
    st[3, 5, q] = 2.0          # B_pol
    st[4, 1, 0] = 6.0          # T
    e = m.energies(st)
    ll = max(lo[q] * (lo[q] + 1), 1)
    assert abs(e[0] - 0.5 * 25.0 / ll) < 1e-14
    assert abs(e[1] - 0.5 * 4.0 / ll) < 1e-14
    assert abs(e[2] - 0.5 * 36.0) < 1e-14


def test_adopted_reading_is_stable_over_100_steps(oracle):
    """Dirichlet rows + radials from the reference data flow: energies decay."""
    p = dict(lmax=15, nlat=24, nlon=48, nr=17)
    m = oracle.Model(**p)
    st = oracle.seeded_state(2024, p["lmax"], p["lmax"], p["nr"], amp=1e-3)
    e0 = m.energies(st)
    s, nlp = st, None
    for k in range(100):
        s, nl, nlp = m.step(s, nlp, k)
    e = m.energies(s)
    assert np.all(np.isfinite(e)) and np.all(e < e0)


def test_seeded_normals_statistics(oracle):
    g = oracle.seeded_normals(99, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.01


def test_radial_force_matches_numpy_restatement(oracle):
    """With the radial forcing on, the C oracle's nonlinear terms equal the
    numpy restatement's (the u_pol term differs by the buoyancy's poloidal
    scalar, the rest are unchanged)."""
    from oracle import np_oracle

    st = oracle.seeded_state(17, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1.0, zero_boundaries=False)
    on = oracle.Model(**TINY, Ra=300.0, radial_force=1).nonlinear(st)
    off = oracle.Model(**TINY, Ra=300.0).nonlinear(st)
    ref = np_oracle.NpModel(**TINY, Ra=300.0, radial_force=1).nonlinear(st)
    assert _relf(on, ref) < 1e-12
    assert np.abs(on[1] - off[1]).max() > 1e-3 * np.abs(off[1]).max()
    for f in (0, 2, 3, 4):
        assert np.array_equal(on[f], off[f])


def test_radial_force_has_the_buoyancy_as_radial_component(oracle):
    """The poloidal increment P_F the forcing adds to nl(u_pol) has, under
    u_r = l(l+1) P / r^2 (magnetic.jl:257-259), the radial component of the
    buoyancy (Pm/E)(Pm/Pr) Ra r T (velocity.jl:534-537): synthesised on the
    grid at every level, l(l+1) P_F / r^2 equals the synthesised buoyancy
    without its l = 0 part (which has no poloidal flow)."""
    L, nlat, nlon, nr = TINY["lmax"], TINY["nlat"], TINY["nlon"], TINY["nr"]
    E, Pm, Pr, Ra = 1e-3, 5.0, 1.0, 250.0
    st = oracle.seeded_state(18, L, L, nr, amp=1.0, zero_boundaries=False)
    dP = (oracle.Model(**TINY, E=E, Pm=Pm, Pr=Pr, Ra=Ra, radial_force=1).nonlinear(st)[1]
          - oracle.Model(**TINY, E=E, Pm=Pm, Pr=Pr, Ra=Ra).nonlinear(st)[1])
    sht = oracle.SHT(L, L, nlat, nlon)
    r = oracle.radial_grid(nr, 0.35)[3]
    l = np.array([ll for ll in range(L + 1) for m in range(ll + 1)], dtype=float)
    Fr_spec = (Pm / E) * (Pm / Pr) * Ra * r[:, None] * st[4]
    Fr_spec[:, l == 0] = 0.0
    ur_spec = l * (l + 1) * dP / (r ** 2)[:, None]
    a = sht.scalar_synth(ur_spec)
    b = sht.scalar_synth(Fr_spec)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_radial_force_makes_ra_matter(oracle):
    """Ra enters only through the buoyancy: without the forcing two Rayleigh
    numbers give the same state; with it they differ."""
    st = oracle.seeded_state(19, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1e-3)
    out = {}
    for rf in (0, 1):
        for Ra in (100.0, 500.0):
            m = oracle.Model(**TINY, Ra=Ra, radial_force=rf)
            s, nl, nlp = m.step(st, None, 0)
            s, nl, nlp = m.step(s, nlp, 1)
            out[rf, Ra] = s
    assert np.array_equal(out[0, 100.0], out[0, 500.0])
    assert np.abs(out[1, 100.0][1] - out[1, 500.0][1]).max() > 1e-6 * np.abs(out[1, 500.0][1]).max()
