"""CPU pins of the exponential-integrator oracle (oracle/np_etd.py) and of the
library's host matrix functions (gd_exp_functions) against it.

Oracle pins, each chosen so that a plausible slip fails one of them:
  * test/erk2.jl known answers (1x1 cache, dt 0.1, lambda 0.5): constant
    forcing -> exp(-lambda dt) u0 + (1 - exp(-lambda dt)) c / lambda, and the
    stage-recomputed N = beta u -> exp((beta - lambda) dt) u0, atol 1e-4 as
    the reference tests them; constant forcing is exact, so also to 1e-15;
  * scalar closed forms exp(z), (e^z - 1)/z, (e^z - 1 - z)/z^2 on diagonal
    operators, in the solve branch, the small-norm series branch and the
    ill-conditioned series fallback (a wrong factorial offset fails these);
  * the block-triangular exponential identity for phi1, phi2 on the real
    radial operators where the reference takes its solve branch;
  * EAB2 with constant forcing equals the exact solution of du/dt = A u + c
    from the exponential of the augmented matrix [[A, c], [0, 0]];
  * ERK2 is second order: halving dt divides the one-step error by ~8.
Library vs oracle: the reference's algorithm is followed branch for branch;
where it solves with A_l the tolerance is 100 eps / rcond(A_l) (the radial
operator without boundary rows is nearly singular at small l), elsewhere
1e-12.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import np_etd as X

EPS = np.finfo(float).eps


def _scalar_cache(dt, lam):
    z = -lam * dt
    eh, ef = math.exp(z / 2), math.exp(z)
    mats = [np.array([[v]]) for v in (eh, ef, (eh - 1) / (z / 2), (ef - 1) / z, (ef - 1 - z) / z ** 2)]
    return [[m] for m in mats]


def test_erk2_known_answers_of_reference_test():
    dt, lam = 0.1, 0.5
    mats = _scalar_cache(dt, lam)
    lofl = [0]
    # scenario 1: constant forcing c
    u0, c = 0.3, 0.2
    u = np.array([[u0 + 0j]])
    nl = np.array([[c + 0j]])
    lin, k1, st, ncur = X.erk2_prepare(u, nl, mats, lofl, dt)
    u1 = X.erk2_finalize(lin, k1, ncur, nl, mats, lofl, dt)  # stage nl = c
    expected = math.exp(-lam * dt) * u0 + (1 - math.exp(-lam * dt)) * c / lam
    assert abs(u1[0, 0].real - expected) < 1e-4
    assert abs(u1[0, 0].real - expected) < 1e-15
    assert abs(u1[0, 0].imag) < 1e-4
    # scenario 2: N = beta u, recomputed at the stage
    u0, beta = 0.45, 0.15
    u = np.array([[u0 + 0j]])
    nl = beta * u
    lin, k1, st, ncur = X.erk2_prepare(u, nl, mats, lofl, dt)
    u1 = X.erk2_finalize(lin, k1, ncur, beta * st, mats, lofl, dt)
    assert abs(u1[0, 0].real - math.exp((beta - lam) * dt) * u0) < 1e-4
    assert abs(u1[0, 0].imag) < 1e-4


def test_erk2_one_step_error_constant():
    """The reference's ERK2 takes its stage at dt/2 (E_half, dt/2 phi1_half) but
    weights the correction by dt phi2 (N(U) - N_n), the weight of a full-step
    stage; Taylor expansion of those formulas for u' = -lambda u + beta u gives
    the one-step error u0 dt^2 beta (lambda - beta) / 4 + O(dt^3) (a first
    order scheme).  A wrong stage, weight or sign changes the constant."""
    lam, beta, u0 = 0.5, 0.15, 0.45
    for dt in (0.1, 0.05, 0.025):
        mats = _scalar_cache(dt, lam)
        u = np.array([[u0 + 0j]])
        lin, k1, st, ncur = X.erk2_prepare(u, beta * u, mats, [0], dt)
        u1 = X.erk2_finalize(lin, k1, ncur, beta * st, mats, [0], dt)
        err = u1[0, 0].real - math.exp((beta - lam) * dt) * u0
        lead = u0 * dt ** 2 * beta * (lam - beta) / 4
        assert abs(err / lead - 1) < 0.5 * dt, (dt, err, lead)


def _scalar_phis(z):
    """exp(z), (e^z - 1)/z, (e^z - 1 - z)/z^2; for |z| < 1e-4 the first three
    Taylor terms (exact to rounding there; the closed forms cancel)."""
    z = np.asarray(z, float)
    small = np.abs(z) < 1e-4
    zs = np.where(small, 1.0, z)
    p1 = np.where(small, 1 + z / 2 + z * z / 6, np.expm1(zs) / zs)
    p2 = np.where(small, 0.5 + z / 6 + z * z / 24, (np.expm1(zs) - zs) / zs ** 2)
    return np.exp(z), p1, p2


@pytest.mark.parametrize("diag,branch", [
    ([-0.5, -2.0, -7.0, -0.1], "solve"),
    ([-1e-3, -5e-3, 2e-3], "series"),
    ([-1e-12, -1.0, -0.3], "ill-conditioned series"),
])
def test_phi_functions_closed_forms(diag, branch):
    z = np.array(diag)
    A = np.diag(z)
    E = sl.expm(A)
    e, p1, p2 = _scalar_phis(z)
    assert np.allclose(np.diag(E), e, rtol=1e-14, atol=0)
    assert np.allclose(X.phi1_function(A, E), np.diag(p1), rtol=1e-13, atol=1e-15)
    assert np.allclose(X.phi2_function(A, E), np.diag(p2), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("nr,l", [(9, 7), (9, 31), (17, 31), (17, 127)])
def test_phi_functions_block_identity_on_radial_operators(oracle, nr, l):
    r = oracle.radial_grid(nr, 0.35)
    A = X.operator_dense(oracle.radial_laplacian(r, 3), r, 1.0, l)
    rc = X._rcond1(A)
    assert rc > 1e-6  # the reference's solve branch
    for dt in (1e-3, 1e-4):
        eb, p1b, p2b = X.phi_block(dt * A)
        Eh, Ef, p1h, p1f, p2f = X.erk2_matrices(A, dt)
        # first-order perturbation bounds: phi1 = (hA)^-1 (E - I) and
        # phi2 = (hA)^-2 (E - I - hA), whose numerator carries eps ||hA||
        kinv = np.linalg.norm(np.linalg.inv(dt * A), 1)
        tol1 = 100 * EPS * max(1 / rc, kinv)
        tol2 = 100 * EPS * (1 / rc) * max(1.0, kinv)
        assert np.abs(Ef - eb).max() < 1e-13 * np.abs(eb).max()
        assert np.abs(p1f - p1b).max() < tol1 * np.abs(p1b).max()
        assert np.abs(p2f - p2b).max() < tol2 * np.abs(p2b).max()
        E, phi1 = X.etd_matrices(A, dt)
        assert np.abs(phi1 - p1b).max() < tol1 * np.abs(p1b).max()


def test_eab2_constant_forcing_is_exact(oracle):
    nr, l, dt = 9, 15, 1e-3
    r = oracle.radial_grid(nr, 0.35)
    A = X.operator_dense(oracle.radial_laplacian(r, 3), r, 1.0, l)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    c = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    E, phi1 = X.etd_matrices(A, dt)
    got = X.eab2_update(u0[:, None], c[:, None], c[:, None], {l: E}, {l: phi1}, [l], dt)[:, 0]
    aug = np.zeros((nr + 1, nr + 1))
    aug[:nr, :nr] = A
    exact = np.zeros(nr, complex)
    for part in (np.real, np.imag):
        aug[:nr, nr] = part(c)
        v = sl.expm(dt * aug) @ np.concatenate([part(u0), [1.0]])
        exact += v[:nr] * (1 if part is np.real else 1j)
    assert np.abs(got - exact).max() < 1e3 * EPS / X._rcond1(A) * np.abs(exact).max()


# ---------------------------------------------------------------- library host functions
def _lib_functions(nr, lmax, dt, scheme, nu=1.0):
    import gdhip

    dom = gdhip.create_radial_domain(nr, 0.35)
    return dom, gdhip.exp_functions(dom, nu, dt, lmax, scheme)


def _rc(A):
    try:
        return X._rcond1(A)
    except np.linalg.LinAlgError:
        return 0.0


@pytest.mark.parametrize("nr,lmax,dt,nu", [(9, 15, 1e-3, 1.0), (33, 31, 1e-4, 5.0), (65, 127, 1e-5, 1.0)])
def test_library_erk2_matrices_match_oracle(oracle, nr, lmax, dt, nu):
    dom, got = _lib_functions(nr, lmax, dt, "erk2", nu)
    r = oracle.radial_grid(nr, 0.35)
    lap = oracle.radial_laplacian(r, 3)
    ls = sorted({0, 1, 2, 3, 4, 5, 7, 10, 15, lmax // 2, lmax})
    for l in ls:
        A = X.operator_dense(lap, r, nu, l)
        rc = _rc(A)
        if abs(rc / math.sqrt(EPS) - 1) < 0.5:
            continue  # either side of the reference's branch threshold
        want = X.erk2_matrices(A, dt)
        for k in range(5):
            if k < 2 or rc < math.sqrt(EPS):  # exponentials; series fallback
                tol = 1e-12
            else:  # solve branch: the perturbation bounds of phi1 (k 2, 3) and phi2 (k 4)
                h = dt / 2 if k == 2 else dt
                kinv = np.linalg.norm(np.linalg.inv(h * A), 1)
                tol = 100 * EPS * (max(1 / rc, kinv) if k < 4 else max(1.0, kinv) / rc)
                tol = max(1e-12, tol)
            err = np.abs(got[k][l] - want[k]).max() / np.abs(want[k]).max()
            assert err < tol, (l, k, err, tol, rc)


@pytest.mark.parametrize("nr,lmax,dt", [(9, 15, 1e-3), (33, 31, 1e-4)])
def test_library_etd_matrices_match_oracle(oracle, nr, lmax, dt):
    dom, (E, phi1) = _lib_functions(nr, lmax, dt, "etd")
    r = oracle.radial_grid(nr, 0.35)
    lap = oracle.radial_laplacian(r, 3)
    checked = 0
    for l in range(lmax + 1):
        A = X.operator_dense(lap, r, 1.0, l)
        Eo, Po = X.etd_matrices(A, dt)
        assert np.abs(E[l] - Eo).max() < 1e-13 * np.abs(Eo).max()
        rc = _rc(A)
        if 100 * EPS / rc < 1e-6:  # phi1 = A \ ... is rounding-dominated below this
            assert np.abs(phi1[l] - Po).max() < 100 * EPS / rc * np.abs(Po).max(), l
            checked += 1
    assert checked >= 8


def test_library_exp_functions_bad_arguments():
    import ctypes

    from gdhip import _lib

    L = _lib.lib()
    out = np.zeros(10)
    r = np.zeros((7, 9))
    assert L.gd_exp_functions(3, 9, 3, r.ctypes.data, 1.0, -1.0, 1, 1, out.ctypes.data) == _lib.GD_ERR_ARG
    assert L.gd_exp_functions(3, 9, 3, r.ctypes.data, 1.0, 1e-3, 1, 7, out.ctypes.data) == _lib.GD_ERR_ARG
    assert b"gd_exp_functions" in L.gd_last_error()
    del ctypes
