"""Field containers mirroring fields.jl:6-48.

Array shapes are C-order views of the reference's column-major arrays:
  SHTnsSpectralField.data_real / data_imag : (nr, nlm)   == Julia (nlm, 1, nr)
  SHTnsPhysicalField.data                  : (nr, nlon, nlat) == Julia (nlat, nlon, nr)
  RadialDomain.r                           : (7, N)      == Julia r (N, 7), r[p] = r^(p-3)
Arrays are numpy (host) or torch device tensors; the transforms accept either.
"""
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib


class SHTnsSpectralField:
    """fields.jl:6-15 (data_real, data_imag; boundary values default 0, Dirichlet)."""

    def __init__(self, config, nr, data_real=None, data_imag=None):
        self.config = config
        self.nlm = config.nlm
        self.nr = nr
        self.data_real = np.zeros((nr, config.nlm)) if data_real is None else data_real
        self.data_imag = np.zeros((nr, config.nlm)) if data_imag is None else data_imag
        self.bc_type_inner = np.ones(config.nlm, dtype=np.int32)
        self.bc_type_outer = np.ones(config.nlm, dtype=np.int32)
        self.boundary_values = np.zeros((2, config.nlm))

    def complex(self):
        return np.asarray(self.data_real) + 1j * np.asarray(self.data_imag)


class SHTnsPhysicalField:
    """fields.jl:18-24."""

    def __init__(self, config, nr, data=None):
        self.config = config
        self.nlat, self.nlon = config.nlat, config.nlon
        self.data = np.zeros((nr, config.nlon, config.nlat)) if data is None else data


class SHTnsVectorField:
    """fields.jl:28-32: three physical components (r, theta, phi)."""

    def __init__(self, config, nr, comps=None):
        if comps is None:
            comps = [SHTnsPhysicalField(config, nr) for _ in range(3)]
        self.r_component, self.θ_component, self.φ_component = comps

    @property
    def theta_component(self):
        return self.θ_component

    @property
    def phi_component(self):
        return self.φ_component


@dataclass
class RadialDomain:
    """fields.jl:41-48 (r table and grid size)."""
    N: int
    r: np.ndarray
    local_range: range = field(default=None)

    def __post_init__(self):
        if self.local_range is None:
            self.local_range = range(1, self.N + 1)


def create_radial_domain(nr, rratio=0.35):
    """fields.jl:108-131: Gauss-Lobatto Chebyshev radii in [ri, ri+1], ri = rratio/(1-rratio)."""
    r = np.zeros((7, nr))
    check(lib().gd_radial_grid(nr, float(rratio), r.ctypes.data))
    return RadialDomain(N=nr, r=r)


def seeded_state(seed, lmax, nr, nfields=5, amp=1.0, power=2.0, mmax=None, zero_boundaries=True):
    """Seeded (l+1)^-power complex state of the configurations, shape
    (nfields, nr, nlm) in the l-major layout (see gd_seeded_state)."""
    mmax = lmax if mmax is None else mmax
    nlm = sum(min(l, mmax) + 1 for l in range(lmax + 1))
    re, im = np.zeros((nfields, nr, nlm)), np.zeros((nfields, nr, nlm))
    check(lib().gd_seeded_state(seed, lmax, mmax, nr, nfields, float(amp), float(power),
                                int(zero_boundaries), re.ctypes.data, im.ctypes.data))
    return re + 1j * im
