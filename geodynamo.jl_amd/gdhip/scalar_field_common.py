"""Boundary conditions of a scalar field in spectral space (thermal.jl:262-313,
compositional.jl:318-370, scalar_field_common.jl:866-929, 1166-1199).

The field's bc_type_inner / bc_type_outer (1 Dirichlet, 2 flux) and
boundary_values[2, nlm] drive a HIP kernel (gd_scalar_bc_*): type 1 sets the
boundary row, type 2 adds the tau-method correction of the flux at that
boundary.  Only method=:tau is built (the one thermal.jl and compositional.jl
call).
"""
import ctypes

import numpy as np

from ._arrays import context, prepare
from ._lib import check, lib
from .linear_algebra import DEFAULT_BW, _fd_mode


class ScalarBoundaryConditions:
    """The device tables of one (grid, band): tau functions, D1 boundary rows,
    and the per-mode condition types and values."""

    def __init__(self, config, domain, bw=DEFAULT_BW, literal=False, device=0):
        self.config, self.nr, self.bw = config, domain.N, bw
        self.nlm = config.nlm
        self.device = device
        self._h = ctypes.c_void_p()
        r = np.ascontiguousarray(domain.r, dtype=np.float64)
        check(lib().gd_scalar_bc_create(context(device), config.lmax, config.mmax, domain.N, bw, r.ctypes.data,
                                        _fd_mode(literal), ctypes.byref(self._h)))
        self._key = None

    def set(self, bc_type_inner, bc_type_outer, boundary_values):
        ti = np.ascontiguousarray(bc_type_inner, dtype=np.int32)
        to = np.ascontiguousarray(bc_type_outer, dtype=np.int32)
        # Julia boundary_values[2, nlm] is (inner, outer) per mode in memory
        v = np.ascontiguousarray(np.asarray(boundary_values, dtype=np.float64).T)
        if ti.shape != (self.nlm,) or to.shape != (self.nlm,) or v.shape != (self.nlm, 2):
            raise ValueError("bc types need nlm entries and boundary_values shape (2, nlm)")
        check(lib().gd_scalar_bc_set(self._h, ti.ctypes.data, to.ctypes.data, v.ctypes.data))

    def apply(self, re, im, dirichlet=True):
        """In place on (nr, nlm) planes: numpy host arrays or torch device tensors."""
        for x in (re, im):
            if tuple(x.shape) != (self.nr, self.nlm):
                raise ValueError("planes must be (nr, nlm) = (%d, %d), got %s" % (self.nr, self.nlm, tuple(x.shape)))
        ctx, flags, (a, b) = prepare(re, im, device=self.device)
        check(lib().gd_scalar_bc_apply(self._h, a, b, int(bool(dirichlet)), flags))

    def tau_tables(self):
        tab, sc = np.zeros((4, self.nr)), np.zeros(6)
        check(lib().gd_scalar_bc_tau(self._h, tab.ctypes.data, sc.ctypes.data))
        return tab, sc

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_scalar_bc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _bc_for(field, domain, bw, literal):
    key = (id(domain), domain.N, bw, literal)
    bc = getattr(field, "_scalar_bc", None)
    if bc is None or bc._key != key:
        bc = ScalarBoundaryConditions(field.config, domain, bw=bw, literal=literal, device=field.config.device)
        bc._key = key
        field._scalar_bc = bc
    bc.set(field.bc_type_inner, field.bc_type_outer, field.boundary_values)
    return bc


def apply_scalar_flux_bc_spectral(field, domain, method="tau", bw=DEFAULT_BW, literal=False):
    """scalar_field_common.jl:1166-1199: the flux correction of every mode
    with a type-2 condition, in place on field.data_real / data_imag."""
    if method != "tau":
        raise NotImplementedError("only method='tau' is built (the one thermal.jl and compositional.jl use)")
    _bc_for(field, domain, bw, literal).apply(field.data_real, field.data_imag, dirichlet=False)


def apply_temperature_boundary_conditions_spectral(field, domain, bw=DEFAULT_BW, literal=False):
    """thermal.jl:262-313: Dirichlet rows of type-1 modes, then the tau flux
    correction of type-2 modes (a shell: r_i > 0)."""
    _bc_for(field, domain, bw, literal).apply(field.data_real, field.data_imag, dirichlet=True)


# compositional.jl:318-370 applies the same two stages
apply_composition_boundary_conditions_spectral = apply_temperature_boundary_conditions_spectral
