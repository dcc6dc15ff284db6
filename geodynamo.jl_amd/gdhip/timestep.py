"""CNAB2 implicit operators mirroring timestep.jl:489-753 (FP64 on the GPU)."""
import ctypes

import numpy as np

from ._arrays import context, prepare
from ._lib import GD_FD_LITERAL, GD_FD_TRANSPOSE, check, lib
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
