"""Time-stepping operators mirroring timestep.jl (FP64 on the GPU): the CNAB2
implicit operators (timestep.jl:489-753) and the exponential EAB2 / staged
ERK2 integrators (timestep.jl:20-65, 425-487, 825-1555)."""
import ctypes

import numpy as np

from ._arrays import context, prepare
from ._lib import GD_EXP_ERK2, GD_EXP_ETD, GD_FD_LITERAL, GD_FD_TRANSPOSE, check, lib
from .linear_algebra import DEFAULT_BW, BandedLU, BandedMatrix


class SHTnsImplicitMatrices:
    """timestep.jl:489-496.  Holds L_l, S_l = I/dt - theta L_l and LU(S_l) for
    l = 0..lmax on the device; host copies on demand."""

    def __init__(self, config, domain, diffusivity, dt, theta=0.5, bw=DEFAULT_BW, literal_fd=False, bc_rows=False):
        self.config, self.domain = config, domain
        self.diffusivity, self.dt, self.theta, self.bw = diffusivity, dt, theta, bw
        self.lmax = int(config.lmax)
        self.nr = domain.N
        self._h = ctypes.c_void_p()
        check(lib().gd_cnab2_mats_create(context(config.device), self.lmax, self.nr, bw,
                                         np.ascontiguousarray(domain.r).ctypes.data, float(diffusivity),
                                         float(dt), float(theta),
                                         GD_FD_LITERAL if literal_fd else GD_FD_TRANSPOSE, int(bc_rows),
                                         ctypes.byref(self._h)))
        self.l_values = np.arange(self.lmax + 1)
        self.lookup = {l: l for l in range(self.lmax + 1)}
        self._host = None

    def _fetch(self):
        if self._host is None:
            shape = (self.lmax + 1, self.nr, 2 * self.bw + 1)
            L, S, LU = np.zeros(shape), np.zeros(shape), np.zeros(shape)
            check(lib().gd_cnab2_mats_get(self._h, L.ctypes.data, S.ctypes.data, LU.ctypes.data))
            self._host = (L, S, LU)
        return self._host

    @property
    def linear_matrices(self):
        return [BandedMatrix(m, self.bw, self.nr) for m in self._fetch()[0]]

    @property
    def system_matrices(self):
        return [BandedMatrix(m, self.bw, self.nr) for m in self._fetch()[1]]

    @property
    def factorizations(self):
        return [BandedLU(m, self.bw, self.nr) for m in self._fetch()[2]]

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_cnab2_mats_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_shtns_timestepping_matrices(config, domain, diffusivity, dt, theta=0.5, bw=DEFAULT_BW,
                                       literal_fd=False, bc_rows=False):
    """timestep.jl:498-543 (bc_rows=True adds Dirichlet rows, zero boundary values)"""
    return SHTnsImplicitMatrices(config, domain, diffusivity, dt, theta, bw, literal_fd, bc_rows)


def build_rhs_cnab2(rhs, un, nl, nl_prev, dt, matrices):
    """timestep.jl:610-690: rhs = un/dt + (1-theta) L un + 1.5 nl - 0.5 nl_prev.
    dt must match the matrices' dt (the reference rebuilds them when it changes)."""
    if abs(dt - matrices.dt) > 1e-15 * abs(dt):
        raise ValueError("dt differs from the implicit matrices' dt; rebuild them")
    cfg = un.config
    ctx, flags, p = prepare(un.data_real, un.data_imag, nl.data_real, nl.data_imag, nl_prev.data_real,
                            nl_prev.data_imag, rhs.data_real, rhs.data_imag, device=cfg.device)
    check(lib().gd_build_rhs_cnab2(matrices._h, cfg.lmax, cfg.mmax, un.nr, *p, flags))
    return rhs


def solve_implicit_step(solution, rhs, matrices):
    """timestep.jl:692-730"""
    cfg = rhs.config
    ctx, flags, p = prepare(rhs.data_real, rhs.data_imag, solution.data_real, solution.data_imag,
                            device=cfg.device)
    check(lib().gd_solve_implicit_step(matrices._h, cfg.lmax, cfg.mmax, rhs.nr, *p, flags))
    return solution


def apply_explicit_operator(output, inp, nonlinear, domain, diffusivity, dt, nl_prev=None, matrices=None):
    """timestep.jl:561-602 (the CNAB2 branch)."""
    if nl_prev is not None and matrices is not None:
        return build_rhs_cnab2(output, inp, nonlinear, nl_prev, dt, matrices)
    output.data_real[...] = inp.data_real / dt + nonlinear.data_real
    output.data_imag[...] = inp.data_imag / dt + nonlinear.data_imag
    return output


def compute_timestep_error(new_field, old_field):
    """timestep.jl:733-753"""
    cfg = new_field.config
    out = np.zeros(1)
    ctx, flags, p = prepare(new_field.data_real, new_field.data_imag, old_field.data_real,
                            old_field.data_imag, device=cfg.device)
    n = int(np.prod(new_field.data_real.shape))
    if flags:
        check(lib().gd_timestep_error(ctx, n, *p, out.ctypes.data, flags))
        return float(out[0])
    import torch

    dev = torch.zeros(1, dtype=torch.float64, device=new_field.data_real.device)
    check(lib().gd_timestep_error(ctx, n, *p, ctypes.c_void_p(dev.data_ptr()), flags))
    return float(dev.item())


# ---------------------------------------------------------------- exponential integrators
def _colmajor(mats, nr):
    """(k, nr, nr) numpy matrices -> contiguous column-major blocks (Julia Matrix layout)."""
    a = np.asarray(mats, dtype=np.float64).reshape(-1, nr, nr)
    return np.ascontiguousarray(a.transpose(0, 2, 1))


def _from_colmajor(buf, nr):
    return np.ascontiguousarray(buf.reshape(-1, nr, nr).transpose(0, 2, 1))


def exp_functions(domain, diffusivity, dt, lmax, scheme="erk2", bw=DEFAULT_BW, literal_fd=False):
    """The host FP64 matrix functions of A_l = nu (Lap - l(l+1)/r^2), l = 0..lmax
    (no device work).  scheme 'etd': (E, phi1) as create_etd_cache
    (timestep.jl:37-65); 'erk2': (E_half, E_full, phi1_half, phi1_full, phi2_full)
    as create_erk2_cache (timestep.jl:896-968).  Each is an (lmax+1, nr, nr) array."""
    which = {"etd": GD_EXP_ETD, "erk2": GD_EXP_ERK2}[scheme]
    nmat = 2 if which == GD_EXP_ETD else 5
    nr = domain.N
    out = np.zeros(nmat * (lmax + 1) * nr * nr)
    check(lib().gd_exp_functions(int(lmax), nr, bw, np.ascontiguousarray(domain.r).ctypes.data,
                                 float(diffusivity), float(dt),
                                 GD_FD_LITERAL if literal_fd else GD_FD_TRANSPOSE, which, out.ctypes.data))
    blocks = out.reshape(nmat, -1)
    return tuple(_from_colmajor(b, nr) for b in blocks)


class ETDCache:
    """timestep.jl:24-29: dt, l_values, E = exp(dt A_l), phi1 = A_l^-1 (E - I)/dt,
    device resident; host copies on demand."""

    def __init__(self, config, domain, diffusivity, dt, bw=DEFAULT_BW, literal_fd=False):
        self.dt = float(dt)
        self.nr = domain.N
        self.lmax = int(config.lmax)
        self.l_values = list(range(self.lmax + 1))
        self._h = ctypes.c_void_p()
        check(lib().gd_etd_cache_create(context(config.device), self.lmax, self.nr, bw,
                                        np.ascontiguousarray(domain.r).ctypes.data, float(diffusivity),
                                        self.dt, GD_FD_LITERAL if literal_fd else GD_FD_TRANSPOSE,
                                        ctypes.byref(self._h)))
        self._host = None

    def _fetch(self):
        if self._host is None:
            n = len(self.l_values) * self.nr * self.nr
            E, P = np.zeros(n), np.zeros(n)
            check(lib().gd_etd_cache_get(self._h, E.ctypes.data, P.ctypes.data))
            self._host = (_from_colmajor(E, self.nr), _from_colmajor(P, self.nr))
        return self._host

    @property
    def E(self):
        return list(self._fetch()[0])

    @property
    def phi1(self):
        return list(self._fetch()[1])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_etd_cache_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_etd_cache(config, domain, diffusivity, dt, bw=DEFAULT_BW, literal_fd=False):
    """timestep.jl:37-65"""
    return ETDCache(config, domain, diffusivity, dt, bw, literal_fd)


def eab2_update(u, nl, nl_prev, etd, config, dt):
    """timestep.jl:425-487: u = E u + dt phi1 (3/2 nl - 1/2 nl_prev) per (l, m), in place."""
    ctx, flags, p = prepare(u.data_real, u.data_imag, nl.data_real, nl.data_imag, nl_prev.data_real,
                            nl_prev.data_imag, device=config.device)
    check(lib().gd_eab2_update(etd._h, config.lmax, config.mmax, u.nr, *p, float(dt), flags))
    return u


class ERK2Cache:
    """timestep.jl:830-851.  Either built by create_erk2_cache or, as the
    reference's constructor, from explicit per-l matrices:
    ERK2Cache(dt, l_values, E_half, E_full, phi1_half, phi1_full, phi2_full)."""

    def __init__(self, dt, l_values, E_half, E_full, phi1_half, phi1_full, phi2_full,
                 use_krylov=False, krylov_m=20, krylov_tol=1e-8, mpi_consistent=True, device=0):
        if use_krylov:
            raise ValueError("Krylov-based ERK2 caches are not supported in staged integration")
        self.dt = float(dt)
        self.l_values = [int(l) for l in l_values]
        self.nr = int(np.asarray(E_full[0]).shape[0])
        self.use_krylov, self.krylov_m, self.krylov_tol = False, krylov_m, krylov_tol
        self.mpi_consistent = mpi_consistent
        mats = [_colmajor(m, self.nr) for m in (E_half, E_full, phi1_half, phi1_full, phi2_full)]
        lv = np.ascontiguousarray(self.l_values, dtype=np.int32)
        self._h = ctypes.c_void_p()
        check(lib().gd_erk2_cache_from_matrices(context(device), len(self.l_values), lv.ctypes.data, self.nr,
                                                self.dt, *[m.ctypes.data for m in mats], ctypes.byref(self._h)))
        self._host = None

    @classmethod
    def _created(cls, handle, dt, lmax, nr):
        self = cls.__new__(cls)
        self.dt, self.l_values, self.nr = float(dt), list(range(lmax + 1)), nr
        self.use_krylov, self.krylov_m, self.krylov_tol, self.mpi_consistent = False, 20, 1e-8, True
        self._h, self._host = handle, None
        return self

    def _fetch(self):
        if self._host is None:
            n = len(self.l_values) * self.nr * self.nr
            bufs = [np.zeros(n) for _ in range(5)]
            check(lib().gd_erk2_cache_get(self._h, *[b.ctypes.data for b in bufs]))
            self._host = tuple(_from_colmajor(b, self.nr) for b in bufs)
        return self._host

    E_half = property(lambda self: list(self._fetch()[0]))
    E_full = property(lambda self: list(self._fetch()[1]))
    phi1_half = property(lambda self: list(self._fetch()[2]))
    phi1_full = property(lambda self: list(self._fetch()[3]))
    phi2_full = property(lambda self: list(self._fetch()[4]))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_erk2_cache_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_erk2_cache(config, domain, diffusivity, dt, use_krylov=False, m=20, tol=1e-8, bw=DEFAULT_BW,
                      literal_fd=False):
    """timestep.jl:896-968 (dense matrix functions)."""
    if use_krylov:
        raise ValueError("Krylov-based ERK2 caches are not supported in staged integration")
    h = ctypes.c_void_p()
    check(lib().gd_erk2_cache_create(context(config.device), int(config.lmax), domain.N, bw,
                                     np.ascontiguousarray(domain.r).ctypes.data, float(diffusivity), float(dt),
                                     GD_FD_LITERAL if literal_fd else GD_FD_TRANSPOSE, ctypes.byref(h)))
    return ERK2Cache._created(h, dt, int(config.lmax), domain.N)


class ERK2FieldBuffers:
    """timestep.jl:1283-1313: linear, k1, stage, n_current, stage_nl planes of
    one field, device resident."""

    def __init__(self, u, nl, cache):
        cfg = u.config
        self.cache = cache
        self.nr = cache.nr
        self._h = ctypes.c_void_p()
        check(lib().gd_erk2_buffers_create(cache._h, cfg.lmax, cfg.mmax, u.nr, ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_erk2_buffers_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def erk2_prepare_field(buffers, u, nl, cache, config, dt):
    """timestep.jl:1316-1438"""
    ctx, flags, p = prepare(u.data_real, u.data_imag, nl.data_real, nl.data_imag, device=config.device)
    check(lib().gd_erk2_prepare_field(buffers._h, *p, float(dt), flags))
    return buffers


def erk2_apply_stage(buffers, u):
    """timestep.jl:1440-1444: u = stage"""
    ctx, flags, p = prepare(u.data_real, u.data_imag, device=u.config.device)
    check(lib().gd_erk2_apply_stage(buffers._h, *p, flags))
    return u


def erk2_store_stage_nonlinear(buffers, nl):
    """timestep.jl:1446-1452: stage_nl = nl"""
    ctx, flags, p = prepare(nl.data_real, nl.data_imag, device=nl.config.device)
    check(lib().gd_erk2_store_stage_nonlinear(buffers._h, *p, flags))
    return buffers


def erk2_finalize_field(buffers, u, cache, config, dt):
    """timestep.jl:1449-1555: u = linear + dt (k1 + phi2_full (stage_nl - n_current))"""
    ctx, flags, p = prepare(u.data_real, u.data_imag, device=config.device)
    check(lib().gd_erk2_finalize_field(buffers._h, *p, float(dt), flags))
    return u
