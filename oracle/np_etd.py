"""numpy/scipy restatement of the reference's exponential integrators.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of libgdhip's
EAB2 and staged ERK2 entry points, never as the thing measured or shipped.

  operator_dense   A_l = nu (Lap - l(l+1)/r^2 on the diagonal), dense, from the
                   C oracle's band Laplacian (timestep.jl:45-53, 912-919)
  etd_matrices     E = exp(dt A), phi1 = A \\ ((E - I)/dt)        (timestep.jl:37-65)
  phi1_function    compute_phi1_function with its series fallbacks
                   (timestep.jl:970-1065)
  phi2_function    compute_phi2_function (timestep.jl:1073-1187)
  erk2_matrices    create_erk2_cache, dense branch (timestep.jl:896-968)
  eab2_update      eab2_update! (timestep.jl:425-487)
  erk2_prepare / erk2_finalize   erk2_prepare_field! / erk2_finalize_field!
                   (timestep.jl:1316-1438, 1449-1555)

Matrix exponentials use scipy.linalg.expm (Al-Mohy & Higham 2009), solves use
LAPACK getrf/getrs through scipy.linalg.lu_factor / lu_solve, the 2-norm is
numpy's SVD-based norm and rcond the exact reciprocal 1-norm condition number:
machinery independent of the library's own host code.

Pinned by (tests/test_oracle_etd.py): the known answers of test/erk2.jl,
the scalar closed forms exp(z), (e^z - 1)/z, (e^z - 1 - z)/z^2 on diagonal
operators, the block-triangular exponential identity
    exp([[hA, I, 0], [0, 0, I], [0, 0, 0]]) = [[e^hA, phi1(hA), phi2(hA)], ...]
(Higham, Functions of Matrices, 2008, Thm 10.40) on the real operators, and
the exact solution of du/dt = A u + c for EAB2 with constant forcing.
"""
import math

import numpy as np
import scipy.linalg as sl

EPS = np.finfo(np.float64).eps


def band_to_dense(band):
    """band (N, 2bw+1) C-order == Julia (2bw+1, N): A[i, j] = band[j, bw + i - j]."""
    N, w = band.shape
    bw = (w - 1) // 2
    A = np.zeros((N, N))
    for j in range(N):
        for i in range(max(0, j - bw), min(N, j + bw + 1)):
            A[i, j] = band[j, bw + i - j]
    return A


def operator_dense(lap_band, r, nu, l):
    """r: (7, N) radial table, r[1] = 1/r^2."""
    A = nu * band_to_dense(lap_band)
    A[np.diag_indices_from(A)] -= nu * float(l * (l + 1)) * r[1]
    return A


def etd_matrices(A, dt):
    E = sl.expm(dt * A)
    F = (E - np.eye(A.shape[0])) / dt
    return E, sl.lu_solve(sl.lu_factor(A), F)


def _series(A, k):
    """sum_j A^j / (j+k)!, at most 15 powers, stop when ||term||_2 < 100 eps
    (the first power unchecked for k = 2, timestep.jl:1086-1099)."""
    n = A.shape[0]
    result = np.eye(n) / math.factorial(k)
    P = np.eye(n)
    for j in range(1, 16):
        P = P @ A
        term = P / math.factorial(j + k)
        result = result + term
        if not (k == 2 and j == 1) and np.linalg.norm(term, 2) < EPS * 100:
            break
    return result


def _rcond1(A):
    return 1.0 / (np.linalg.norm(A, 1) * np.linalg.norm(np.linalg.inv(A), 1))


def phi1_function(A, expA):
    n = A.shape[0]
    if not (np.all(np.isfinite(A)) and np.all(np.isfinite(expA))):
        return np.eye(n)
    if np.linalg.norm(A, 2) < 1e-2:
        return _series(A, 1)
    try:
        rc = _rcond1(A)
    except np.linalg.LinAlgError:
        return _series(A, 1)
    if rc < math.sqrt(EPS):
        return _series(A, 1)
    R = sl.lu_solve(sl.lu_factor(A), expA - np.eye(n))
    return R if np.all(np.isfinite(R)) else _series(A, 1)


def phi2_function(A, expA):
    n = A.shape[0]
    if not (np.all(np.isfinite(A)) and np.all(np.isfinite(expA))):
        return np.zeros((n, n))
    if np.linalg.norm(A, 2) < 1e-2:
        return _series(A, 2)
    try:
        rc = _rcond1(A)
    except np.linalg.LinAlgError:
        return _series(A, 2)
    if rc < math.sqrt(EPS):
        return _series(A, 2)
    f = sl.lu_factor(A)
    R = sl.lu_solve(f, sl.lu_solve(f, expA - np.eye(n) - A))
    return R if np.all(np.isfinite(R)) else _series(A, 2)


def erk2_matrices(A, dt):
    """(E_half, E_full, phi1_half, phi1_full, phi2_full) of one l."""
    Ah, Af = (dt / 2) * A, dt * A
    Eh, Ef = sl.expm(Ah), sl.expm(Af)
    return Eh, Ef, phi1_function(Ah, Eh), phi1_function(Af, Ef), phi2_function(Af, Ef)


def phi_block(hA):
    """(e^hA, phi1(hA), phi2(hA)) from one exponential of the 3n x 3n block
    matrix [[hA, I, 0], [0, 0, I], [0, 0, 0]]: an independent evaluation."""
    n = hA.shape[0]
    B = np.zeros((3 * n, 3 * n))
    B[:n, :n] = hA
    B[:n, n:2 * n] = np.eye(n)
    B[n:2 * n, 2 * n:] = np.eye(n)
    X = sl.expm(B)
    return X[:n, :n], X[:n, n:2 * n], X[:n, 2 * n:]


def eab2_update(u, nl, nlp, E, phi1, l_of_lm, dt):
    """u, nl, nlp complex (nr, nlm); E, phi1 per l.  Returns the new u."""
    out = np.empty_like(u)
    for lm, l in enumerate(l_of_lm):
        nrn = 1.5 * nl[:, lm] - 0.5 * nlp[:, lm]
        out[:, lm] = E[l] @ u[:, lm] + dt * (phi1[l] @ nrn)
    return out


def erk2_prepare(u, nl, mats, l_of_lm, dt):
    """Returns (linear, k1, stage, n_current) complex (nr, nlm)."""
    Eh, Ef, p1h, p1f, _ = mats
    lin, k1, st = np.empty_like(u), np.empty_like(u), np.empty_like(u)
    for lm, l in enumerate(l_of_lm):
        lin[:, lm] = Ef[l] @ u[:, lm]
        k1[:, lm] = p1f[l] @ nl[:, lm]
        st[:, lm] = Eh[l] @ u[:, lm] + (dt / 2) * (p1h[l] @ nl[:, lm])
    return lin, k1, st, nl.copy()


def erk2_finalize(lin, k1, n_cur, stage_nl, mats, l_of_lm, dt):
    phi2 = mats[4]
    out = np.empty_like(lin)
    for lm, l in enumerate(l_of_lm):
        out[:, lm] = lin[:, lm] + dt * (k1[:, lm] + phi2[l] @ (stage_nl[:, lm] - n_cur[:, lm]))
    return out
