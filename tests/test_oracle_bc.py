"""Pins of the scalar boundary-condition oracle (oracle/np_bc.py), CPU.

The tau method of scalar_field_common.jl:573-929 is checked against what the
mathematics fixes, by machinery independent of the restatement:
  * Chebyshev values on the Gauss-Lobatto grid in closed form ((-1)^k for
    T_{N-1}) and by numpy.polynomial.chebyshev.chebval;
  * end and interior derivatives by numpy.polynomial.chebyshev.chebder;
  * a single-boundary correction makes the D1 flux at that boundary exactly
    the prescribed value; a two-boundary correction is a combination of
    T_{N-1} and T_N (recovered by least squares) whose analytic end
    derivatives are the two flux errors;
  * each single-boundary tau function is T_{N-1} - a T_N with zero analytic
    derivative at the other boundary;
  * Dirichlet rows, the 1e-12 rule for the imaginary profile, and modes
    without a condition left untouched.
"""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

from oracle import np_bc, np_oracle


def _grid(nr, rratio=0.35, bw=3):
    r = np_oracle.radial_grid(nr, rratio)
    return r, np_oracle.fd_dense(1, r, bw)


def _x(r):
    return 2.0 * (r - r[0]) / (r[-1] - r[0]) - 1.0


@pytest.mark.parametrize("nr", [9, 17, 33, 65])
def test_chebyshev_tables(nr):
    r, D1 = _grid(nr)
    tc = np_bc.TauCache(r, D1)
    k = np.arange(nr)
    assert np.max(np.abs(tc.tau1 - (-1.0) ** (nr - 1 - k))) < 1e-12
    e = np.zeros(nr + 1)
    e[nr] = 1.0
    assert np.max(np.abs(tc.tau2 - C.chebval(_x(r), e))) < 1e-11
    s = 2.0 / (r[-1] - r[0])
    for n, di, do in ((nr - 1, tc.dt1i, tc.dt1o), (nr, tc.dt2i, tc.dt2o)):
        c = np.zeros(n + 1)
        c[n] = 1.0
        d = C.chebder(c)
        assert di == pytest.approx(C.chebval(-1.0, d) * s, rel=1e-12)
        assert do == pytest.approx(C.chebval(1.0, d) * s, rel=1e-12)
        # the interior branch n sin(n th) / sin(th)
        xm = 0.3
        rv = r[0] + (xm + 1.0) * (r[-1] - r[0]) / 2.0
        assert np_bc.chebyshev_derivative(n, rv, r) == pytest.approx(C.chebval(xm, d) * s, rel=1e-10)


def _profiles(seed, nr, nlm, imag_scale=1.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nr, nlm)), imag_scale * rng.standard_normal((nr, nlm))


@pytest.mark.parametrize("nr", [9, 33])
def test_single_boundary_flux_is_exact(nr):
    r, D1 = _grid(nr)
    nlm = 6
    re, im = _profiles(1, nr, nlm)
    vals = np.random.default_rng(2).standard_normal((2, nlm))
    ti = np.array([2, 0, 2, 1, 0, 2])
    to = np.array([0, 2, 0, 2, 1, 1])
    re2, im2 = np_bc.apply_scalar_bc(re, im, ti, to, vals, r, D1)
    d = D1 @ re2
    for k in range(nlm):
        if ti[k] == 2 and to[k] != 2:
            assert d[0, k] == pytest.approx(vals[0, k], abs=1e-10 * max(1.0, np.abs(D1).max()))
        if to[k] == 2 and ti[k] != 2:
            assert d[-1, k] == pytest.approx(vals[1, k], abs=1e-10 * max(1.0, np.abs(D1).max()))
    # type 1 at the outer boundary of mode 4 and no flux condition: row set, rest untouched
    assert re2[-1, 4] == vals[1, 4] and im2[-1, 4] == 0.0
    assert np.array_equal(re2[:-1, 4], re[:-1, 4]) and np.array_equal(im2[:-1, 4], im[:-1, 4])
    # mode 5: inner flux after the outer Dirichlet row (the reference's order): the
    # single inner tau function then moves that row too
    corr = re2[:, 5] - np.r_[re[:-1, 5], vals[1, 5]]
    tc = np_bc.TauCache(r, D1)
    c = corr[0] / tc.tau_in[0]
    assert np.max(np.abs(corr - c * tc.tau_in)) < 1e-12 * max(1.0, np.abs(corr).max())


@pytest.mark.parametrize("nr", [9, 16, 17, 32, 65])
def test_two_boundary_correction_is_the_chebyshev_pair(nr):
    r, D1 = _grid(nr)
    nlm = 3
    re, im = _profiles(3, nr, nlm)
    vals = np.random.default_rng(4).standard_normal((2, nlm))
    t = np.full(nlm, 2)
    re2, _ = np_bc.apply_scalar_bc(re, im, t, t, vals, r, D1)
    x = _x(r)
    s = 2.0 / (r[-1] - r[0])
    e1, e2 = np.zeros(nr), np.zeros(nr + 1)
    e1[nr - 1], e2[nr] = 1.0, 1.0
    T1, T2 = C.chebval(x, e1), C.chebval(x, e2)
    d1, d2 = C.chebder(e1), C.chebder(e2)
    flux = D1 @ re
    for k in range(nlm):
        corr = re2[:, k] - re[:, k]
        (c1, c2), *_ = np.linalg.lstsq(np.stack([T1, T2], axis=1), corr, rcond=None)
        assert np.max(np.abs(corr - c1 * T1 - c2 * T2)) < 1e-9 * max(1.0, np.abs(corr).max())
        gi = (c1 * C.chebval(-1.0, d1) + c2 * C.chebval(-1.0, d2)) * s
        go = (c1 * C.chebval(1.0, d1) + c2 * C.chebval(1.0, d2)) * s
        assert gi == pytest.approx(vals[0, k] - flux[0, k], rel=1e-8, abs=1e-9)
        assert go == pytest.approx(vals[1, k] - flux[-1, k], rel=1e-8, abs=1e-9)


@pytest.mark.parametrize("nr", [9, 16, 33])
def test_single_boundary_functions_are_flat_at_the_other_end(nr):
    """compute_inner_tau_function (:733-748): T_{N-1} - a T_N with zero analytic
    derivative at the outer boundary; the outer function at the inner one."""
    r, D1 = _grid(nr)
    tc = np_bc.TauCache(r, D1)
    x = _x(r)
    e1, e2 = np.zeros(nr), np.zeros(nr + 1)
    e1[nr - 1], e2[nr] = 1.0, 1.0
    A = np.stack([C.chebval(x, e1), C.chebval(x, e2)], axis=1)
    for tau, end in ((tc.tau_in, 1.0), (tc.tau_out, -1.0)):
        (c1, c2), *_ = np.linalg.lstsq(A, tau, rcond=None)
        assert c1 == pytest.approx(1.0, abs=1e-9)
        slope = c1 * C.chebval(end, C.chebder(e1)) + c2 * C.chebval(end, C.chebder(e2))
        assert abs(slope) < 1e-8 * nr ** 2
        other = c1 * C.chebval(-end, C.chebder(e1)) + c2 * C.chebval(-end, C.chebder(e2))
        assert abs(other) > 1.0


def test_imaginary_rule_and_untouched_modes():
    nr = 17
    r, D1 = _grid(nr)
    re, _ = _profiles(5, nr, 4)
    im = np.zeros((nr, 4))
    im[:, 1] = 1e-13          # below the threshold: left alone
    im[:, 2] = 1e-3           # above: receives the real profile's correction
    im[3, 3] = 2e-12
    vals = np.full((2, 4), 0.5)
    t = np.full(4, 2)
    re2, im2 = np_bc.apply_scalar_bc(re, im, t, t, vals, r, D1)
    assert np.array_equal(im2[:, 0], im[:, 0]) and np.array_equal(im2[:, 1], im[:, 1])
    for k in (2, 3):
        assert np.max(np.abs((im2[:, k] - im[:, k]) - (re2[:, k] - re[:, k]))) < 1e-13
    none = np.zeros(4, dtype=int)
    re3, im3 = np_bc.apply_scalar_bc(re, im, none, none, vals, r, D1)
    assert np.array_equal(re3, re) and np.array_equal(im3, im)
    # flux-only call (apply_scalar_flux_bc_spectral!): type-1 rows untouched
    one = np.ones(4, dtype=int)
    re4, im4 = np_bc.apply_scalar_bc(re, im, one, one, vals, r, D1, dirichlet=False)
    assert np.array_equal(re4, re) and np.array_equal(im4, im)
