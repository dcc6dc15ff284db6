"""Banded radial operators mirroring linear_algebra.jl.

Band storage is the reference's (2bw+1, N) column-major layout, held here as a
C-order (N, 2bw+1) array: data[j, bw + i - j] = A[i, j] (linear_algebra.jl:19-21).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._arrays import prepare
from ._lib import GD_FD_LITERAL, GD_FD_TRANSPOSE, check, lib

DEFAULT_BW = 3  # i_KL of the configurations (parameters.jl:22 default is 4)


@dataclass
class BandedMatrix:
    """linear_algebra.jl:3-7"""
    data: np.ndarray
    bandwidth: int
    size: int


@dataclass
class BandedLU:
    """linear_algebra.jl:13-17 (in-place LU factors, unit L below the diagonal row)"""
    lu: object
    bandwidth: int
    size: int


def _fd_mode(literal):
    return GD_FD_LITERAL if literal else GD_FD_TRANSPOSE


def create_derivative_matrix(order, domain, bw=DEFAULT_BW, literal=False):
    """linear_algebra.jl:90-132.  literal=True reproduces the reference's
    V \\ rhs solve; the default solves V^T w = k! e_k (exact on polynomials)."""
    N = domain.N
    band = np.zeros((N, 2 * bw + 1))
    r = np.ascontiguousarray(domain.r, dtype=np.float64)  # kept alive across the call
    if r.shape != (7, N):
        raise ValueError("domain.r must be (7, N) = (7, %d), got %s" % (N, r.shape))
    check(lib().gd_derivative_matrix(order, N, bw, r.ctypes.data, _fd_mode(literal), band.ctypes.data))
    return BandedMatrix(band, bw, N)


def create_radial_laplacian(domain, bw=DEFAULT_BW, literal=False):
    """linear_algebra.jl:134-151: d2/dr2 + (2/r) d/dr."""
    N = domain.N
    band = np.zeros((N, 2 * bw + 1))
    r = np.ascontiguousarray(domain.r, dtype=np.float64)  # kept alive across the call
    if r.shape != (7, N):
        raise ValueError("domain.r must be (7, N) = (7, %d), got %s" % (N, r.shape))
    check(lib().gd_radial_laplacian(N, bw, r.ctypes.data, _fd_mode(literal), band.ctypes.data))
    return BandedMatrix(band, bw, N)


def factorize_banded(A, device=0):
    """linear_algebra.jl:23-57 (no pivoting), computed on the GPU in FP64.
    A may be one BandedMatrix or a list of equal-shape ones (factored together)."""
    mats = A if isinstance(A, (list, tuple)) else [A]
    bw, N = mats[0].bandwidth, mats[0].size
    ab = np.ascontiguousarray(np.stack([np.asarray(m.data, dtype=np.float64) for m in mats]))
    lu = np.zeros_like(ab)
    ctx, flags, (pa, pl) = prepare(ab, lu, device=device)
    check(lib().gd_banded_factor(ctx, N, bw, len(mats), pa, pl, flags))
    out = [BandedLU(lu[k], bw, N) for k in range(len(mats))]
    return out if isinstance(A, (list, tuple)) else out[0]


def solve_banded(x, lu, b, device=0):
    """linear_algebra.jl:59-88: solves LU x = b (b may be (N,) or (nrhs, N));
    writes x and returns it."""
    bb = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, lu.size).copy()
    luarr = np.ascontiguousarray(lu.lu, dtype=np.float64)
    ctx, flags, (pl, px) = prepare(luarr, bb, device=device)
    check(lib().gd_banded_solve(ctx, lu.size, lu.bandwidth, 1, pl, None, bb.shape[0], px, flags))
    x[...] = bb.reshape(np.shape(x))
    return x


def apply_banded_matrix(out, matrix, v):
    """linear_algebra.jl:154-219 (y = B v along the last axis); host helper."""
    N, bw = matrix.size, matrix.bandwidth
    v = np.asarray(v)
    acc = np.zeros(v.shape[:-1] + (N,), dtype=np.result_type(v, np.float64))
    for j in range(N):
        for i in range(max(0, j - bw), min(N, j + bw + 1)):
            acc[..., i] += matrix.data[j, bw + i - j] * v[..., j]
    out[...] = acc
    return out


def apply_derivative_matrix(out, matrix, v):
    """linear_algebra.jl:203-219: y = D v for one radial vector (the same band
    product as apply_banded_matrix)."""
    return apply_banded_matrix(out, matrix, v)


def banded_to_dense(matrix):
    """timestep.jl:545-559"""
    N, bw = matrix.size, matrix.bandwidth
    D = np.zeros((N, N))
    for j in range(N):
        for i in range(max(0, j - bw), min(N, j + bw + 1)):
            D[i, j] = matrix.data[j, bw + i - j]
    return D
