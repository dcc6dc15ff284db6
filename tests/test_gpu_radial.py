"""GPU parity of the FP64 banded radial operators against the oracle.

Reference: linear_algebra.jl:23-88 (band LU / solve), timestep.jl:498-730
(CNAB2 operators, rhs, solve).  Bar: FP64, max relative error <= 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_band_lu_and_solve_match_oracle(oracle):
    import gdhip

    print("case: band LU n=33 bw=3 nsys=5", flush=True)
    rng = np.random.default_rng(11)
    N, bw, ns = 33, 3, 5
    mats = []
    for _ in range(ns):
        d = rng.normal(size=(N, 2 * bw + 1))
        d[:, bw] += 10.0
        for j in range(N):  # zero entries outside the matrix
            for row in range(2 * bw + 1):
                i = j + row - bw
                if i < 0 or i >= N:
                    d[j, row] = 0.0
        mats.append(gdhip.BandedMatrix(d, bw, N))
    lus = gdhip.factorize_banded(mats)
    for m, lu in zip(mats, lus):
        ref, rc = oracle.band_factor(m.data)
        assert rc == 0
        assert _rel(lu.lu, ref) < 1e-13
        b = rng.normal(size=N)
        x = np.zeros(N)
        gdhip.solve_banded(x, lu, b)
        assert _rel(x, oracle.band_solve(ref, b)) < 1e-13
        assert _rel(gdhip.banded_to_dense(m) @ x, b) < 1e-12


def test_band_lu_zero_pivot_raises():
    import gdhip

    N, bw = 8, 2
    d = np.zeros((N, 2 * bw + 1))
    d[:, bw] = 1.0
    d[3, bw] = 0.0
    with pytest.raises(gdhip.GDError):
        gdhip.factorize_banded(gdhip.BandedMatrix(d, bw, N))


@pytest.mark.parametrize("lmax,nr", [(31, 33), (127, 65)])
def test_cnab2_matrices_match_oracle(oracle, lmax, nr):
    import gdhip

    print("case: cnab2 matrices lmax=%d nr=%d" % (lmax, nr), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, nlat=3 * (lmax + 1) // 2, nlon=3 * (lmax + 1))
    dom = gdhip.create_radial_domain(nr, 0.35)
    r = oracle.radial_grid(nr, 0.35)
    assert np.allclose(dom.r, r, rtol=2e-16, atol=0)
    for nu in (5.0, 1.0):
        M = gdhip.create_shtns_timestepping_matrices(cfg, dom, nu, 1e-5, theta=0.5, bw=3)
        L, S, LU, rc = oracle.cnab2_matrices(r, 3, nu, 1e-5, 0.5, lmax)
        assert rc == 0
        Lg, Sg, LUg = M._fetch()
        assert _rel(Lg, L) < 1e-14
        assert _rel(Sg, S) < 1e-14
        assert _rel(LUg, LU) < 1e-12


@pytest.mark.parametrize("lmax,nr,bc", [(31, 33, False), (127, 65, False), (31, 33, True), (127, 65, True)])
def test_cnab2_rhs_solve_match_oracle(oracle, lmax, nr, bc):
    import gdhip

    print("case: cnab2 rhs/solve lmax=%d nr=%d bc_rows=%d" % (lmax, nr, bc), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, nlat=3 * (lmax + 1) // 2, nlon=3 * (lmax + 1))
    dom = gdhip.create_radial_domain(nr, 0.35)
    dt, nu = 1e-5, 5.0
    M = gdhip.create_shtns_timestepping_matrices(cfg, dom, nu, dt, bc_rows=bc)
    st = oracle.seeded_state(3, lmax, lmax, nr, nfields=3, amp=1e-3)
    u, nl, nlp = st[0], st[1] * 100, st[2] * 100
    lo, _ = oracle.lm_tables(lmax, lmax)
    L, S, LU, _ = oracle.cnab2_matrices(oracle.radial_grid(nr, 0.35), 3, nu, dt, 0.5, lmax, bc_rows=bc)
    if bc:
        Sg = M._fetch()[1]
        assert _rel(Sg, S) < 1e-14
    rhs_ref = oracle.cnab2_rhs(lo, L, dt, 0.5, u, nl, nlp, bc_rows=bc)
    sol_ref = oracle.cnab2_solve(lo, LU, rhs_ref)

    def F(z):
        f = gdhip.SHTnsSpectralField(cfg, nr)
        f.data_real[...] = z.real
        f.data_imag[...] = z.imag
        return f

    fu, fn, fp = F(u), F(nl), F(nlp)
    rhs = gdhip.SHTnsSpectralField(cfg, nr)
    gdhip.build_rhs_cnab2(rhs, fu, fn, fp, dt, M)
    assert _rel(rhs.complex(), rhs_ref) < 1e-13
    sol = gdhip.SHTnsSpectralField(cfg, nr)
    gdhip.solve_implicit_step(sol, rhs, M)
    assert _rel(sol.complex(), sol_ref) < 1e-12
    # fused in place
    from gdhip._arrays import prepare
    from gdhip._lib import check, lib

    ur, ui = np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag)
    ctx, flags, p = prepare(ur, ui, np.ascontiguousarray(nl.real), np.ascontiguousarray(nl.imag),
                            np.ascontiguousarray(nlp.real), np.ascontiguousarray(nlp.imag))
    check(lib().gd_cnab2_rhs_solve(M._h, lmax, lmax, nr, *p, flags))
    assert _rel(ur + 1j * ui, sol_ref) < 1e-12


def test_missing_l_errors_in_rhs_and_skips_in_solve(oracle):
    """timestep.jl:646 raises for a missing l; :709 skips it."""
    import gdhip

    lmax_f, lmax_m, nr = 12, 10, 17
    cfg_f = gdhip.create_shtnskit_config(lmax_f, nlat=20, nlon=48)
    cfg_m = gdhip.create_shtnskit_config(lmax_m, nlat=16, nlon=36)
    dom = gdhip.create_radial_domain(nr, 0.35)
    M = gdhip.create_shtns_timestepping_matrices(cfg_m, dom, 1.0, 1e-4)
    st = oracle.seeded_state(5, lmax_f, lmax_f, nr, nfields=1)[0]
    f = gdhip.SHTnsSpectralField(cfg_f, nr, np.ascontiguousarray(st.real), np.ascontiguousarray(st.imag))
    rhs = gdhip.SHTnsSpectralField(cfg_f, nr)
    with pytest.raises(gdhip.MissingLError):
        gdhip.build_rhs_cnab2(rhs, f, f, f, 1e-4, M)
    sol = gdhip.SHTnsSpectralField(cfg_f, nr)
    sol.data_real[...] = 7.0
    gdhip.solve_implicit_step(sol, f, M)
    lvals = cfg_f.l_values
    assert np.all(sol.data_real[:, lvals > lmax_m] == 7.0)
    assert np.all(sol.data_real[:, lvals <= lmax_m] != 7.0)


def test_timestep_error(oracle):
    import gdhip

    cfg = gdhip.create_shtnskit_config(7, nlat=12, nlon=24)
    a = oracle.seeded_state(1, 7, 7, 5, nfields=2)
    f0 = gdhip.SHTnsSpectralField(cfg, 5, np.ascontiguousarray(a[0].real), np.ascontiguousarray(a[0].imag))
    f1 = gdhip.SHTnsSpectralField(cfg, 5, np.ascontiguousarray(a[1].real), np.ascontiguousarray(a[1].imag))
    e = gdhip.compute_timestep_error(f0, f1)
    assert abs(e - np.sqrt(np.sum(np.abs(a[0] - a[1]) ** 2))) < 1e-13 * e
