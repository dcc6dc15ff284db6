"""Pins of the oracle's radial part (CPU).

fields.jl:108-131 (grid), linear_algebra.jl:23-151 (band LU / solve, FD
weights, Laplacian), timestep.jl:498-730 (CNAB2 operators, rhs, solve),
simulation.jl:1153 (AB1 bootstrap).  Each check is a closed form or a dense
numpy computation, so a dropped term, wrong sign or transposed operand fails.
"""
import math

import numpy as np
import pytest


def dense(band):
    N, w = band.shape
    bw = (w - 1) // 2
    D = np.zeros((N, N))
    for j in range(N):
        for row in range(w):
            i = j + row - bw
            if 0 <= i < N:
                D[i, j] = band[j, row]
    return D


@pytest.mark.parametrize("N", [9, 33, 65])
def test_radial_grid_closed_form(oracle, N):
    r = oracle.radial_grid(N, 0.35)
    ri = 0.35 / 0.65
    n = np.arange(1, N + 1)
    r4 = 0.5 * (1 + np.cos(np.pi * (N - n) / (N - 1))) + ri
    assert np.allclose(r[3], r4, rtol=1e-15)
    assert abs(r[3, 0] - ri) < 1e-15 and abs(r[3, -1] - (ri + 1)) < 1e-15
    for p in range(7):
        assert np.allclose(r[p], r4 if p == 3 else r4 ** (p - 3), rtol=1e-14)


@pytest.mark.parametrize("order", [1, 2])
def test_fd_weights_exact_on_polynomials(oracle, order):
    N, bw = 33, 3
    r = oracle.radial_grid(N, 0.35)
    D = dense(oracle.fd_weights(order, r, bw, transpose=True))
    x = r[3]
    for k in range(0, 4):  # every stencil has >= bw+1 = 4 points
        f = x ** k
        exact = (k * x ** (k - 1) if order == 1 else k * (k - 1) * x ** (k - 2)) if k >= order else 0 * x
        assert np.abs(D @ f - exact).max() < 1e-8 * max(1.0, np.abs(exact).max())
    # interior rows are exact up to degree 2bw = 6
    for k in range(4, 7):
        f = x ** k
        exact = k * x ** (k - 1) if order == 1 else k * (k - 1) * x ** (k - 2)
        assert np.abs((D @ f - exact)[bw:-bw]).max() < 1e-6 * np.abs(exact).max()


def test_fd_literal_reading_is_ill_posed(oracle):
    """linear_algebra.jl:118 solves V \\ rhs; weights need V^T.  The literal
    weights are not a derivative (regression guard for the chosen reading)."""
    N, bw = 33, 3
    r = oracle.radial_grid(N, 0.35)
    D = dense(oracle.fd_weights(1, r, bw, transpose=False))
    x = r[3]
    assert np.abs(D @ x - 1.0).max() > 1e3


def test_laplacian_closed_form(oracle):
    N, bw = 33, 3
    r = oracle.radial_grid(N, 0.35)
    Lap = dense(oracle.radial_laplacian(r, bw))
    x = r[3]
    for k in range(0, 4):
        exact = k * (k + 1) * x ** (k - 2) if k >= 1 else 0 * x
        assert np.abs(Lap @ x ** k - exact).max() < 1e-8 * max(1, np.abs(exact).max())


def _random_band(rng, N, bw, dom=10.0):
    d = rng.normal(size=(N, 2 * bw + 1))
    d[:, bw] += dom
    for j in range(N):
        for row in range(2 * bw + 1):
            if not 0 <= j + row - bw < N:
                d[j, row] = 0.0
    return d


def test_band_lu_solve_apply_vs_dense(oracle):
    rng = np.random.default_rng(0)
    N, bw = 40, 3
    ab = _random_band(rng, N, bw)
    A = dense(ab)
    lu, rc = oracle.band_factor(ab)
    assert rc == 0
    LU = dense(lu)
    L = np.tril(LU, -1) + np.eye(N)
    U = np.triu(LU)
    assert np.abs(L @ U - A).max() < 1e-12 * np.abs(A).max()
    b = rng.normal(size=N)
    assert np.abs(oracle.band_solve(lu, b) - np.linalg.solve(A, b)).max() < 1e-12
    assert np.abs(oracle.band_apply(ab, b) - A @ b).max() < 1e-12


def test_band_factor_reports_zero_pivot(oracle):
    ab = np.zeros((6, 3))
    ab[:, 1] = 1.0
    ab[2, 1] = 0.0
    _, rc = oracle.band_factor(ab)
    assert rc != 0


@pytest.mark.parametrize("bc_rows", [False, True])
def test_cnab2_matrices_definition(oracle, bc_rows):
    N, bw, nu, dt, th, lmax = 17, 3, 5.0, 1e-4, 0.5, 6
    r = oracle.radial_grid(N, 0.35)
    L, S, LU, rc = oracle.cnab2_matrices(r, bw, nu, dt, th, lmax, bc_rows=bc_rows)
    assert rc == 0
    Lap = dense(oracle.radial_laplacian(r, bw))
    for l in range(lmax + 1):
        Ld = nu * (Lap - l * (l + 1) * np.diag(r[1]))
        assert np.abs(dense(L[l]) - Ld).max() < 1e-12 * np.abs(Ld).max()
        Sd = np.eye(N) / dt - th * Ld
        if bc_rows:
            Sd[0], Sd[-1] = 0, 0
            Sd[0, 0] = Sd[-1, -1] = 1
        assert np.abs(dense(S[l]) - Sd).max() < 1e-12 * np.abs(Sd).max()


def test_cnab2_eigenmode_closed_form(oracle):
    """L v = lam v, nl = nl_prev = 0  =>  u+ = (1/dt + (1-theta) lam) / (1/dt - theta lam) v"""
    N, bw, nu, dt, th, lmax = 17, 3, 1.0, 1e-3, 0.5, 3
    r = oracle.radial_grid(N, 0.35)
    L, S, LU, _ = oracle.cnab2_matrices(r, bw, nu, dt, th, lmax)
    lo, _ = oracle.lm_tables(lmax, lmax)
    for l in (0, 2, 3):
        w, V = np.linalg.eig(dense(L[l]))
        k = int(np.argmin(np.abs(w.imag) + (np.abs(w.real) < 1.0) * 1e9))
        lam, v = w[k].real, V[:, k].real
        u = np.zeros((N, lo.size), complex)
        q = int(np.where(lo == l)[0][0])
        u[:, q] = v * (1 + 0.5j)
        z = np.zeros_like(u)
        out = oracle.cnab2_solve(lo, LU, oracle.cnab2_rhs(lo, L, dt, th, u, z, z))
        g = (1 / dt + (1 - th) * lam) / (1 / dt - th * lam)
        assert np.abs(out[:, q] - g * u[:, q]).max() < 1e-9 * np.abs(u[:, q]).max()


def test_cnab2_ab2_constant_forcing(oracle):
    """With nl = nl_prev = c the AB2 extrapolation gives exactly c."""
    N, bw, nu, dt, th, lmax = 13, 3, 2.0, 1e-3, 0.5, 2
    r = oracle.radial_grid(N, 0.35)
    L, S, LU, _ = oracle.cnab2_matrices(r, bw, nu, dt, th, lmax)
    lo, _ = oracle.lm_tables(lmax, lmax)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(N, lo.size)) + 1j * rng.normal(size=(N, lo.size))
    c = rng.normal(size=(N, lo.size)) + 1j * rng.normal(size=(N, lo.size))
    out = oracle.cnab2_solve(lo, LU, oracle.cnab2_rhs(lo, L, dt, th, u, c, c))
    for q, l in enumerate(lo):
        Ld = dense(L[l])
        ref = np.linalg.solve(np.eye(N) / dt - th * Ld, u[:, q] / dt + (1 - th) * Ld @ u[:, q] + c[:, q])
        assert np.abs(out[:, q] - ref).max() < 1e-10 * np.abs(ref).max()


def test_ab1_bootstrap_ignores_history(oracle):
    """simulation.jl:1153: at step 0 nl_prev := nl."""
    m = oracle.Model(5, 8, 16, 7)
    st = oracle.seeded_state(1, 5, 5, 7, amp=1e-3)
    a, _, _ = m.step(st, None, 0)
    b, _, _ = m.step(st, st * 1e6, 0)
    assert np.array_equal(a, b)
    c, _, _ = m.step(st, st * 1e6, 1)
    assert not np.allclose(a, c)


def test_timestep_error_closed_form(oracle):
    a = np.arange(6.0)
    b = np.ones(6)
    e = oracle.lib().gdo_timestep_error(6, a, -a, b, b)
    assert abs(e - math.sqrt(np.sum((a - b) ** 2 + (-a - b) ** 2))) < 1e-12
