"""SHT configuration mirroring shtnskit_transforms.jl:61-200.

The transform is the library's own: orthonormal Y_lm with Condon-Shortley
phase on a Gauss-Legendre grid (colatitude north first, longitude
phi_j = 2 pi j / nlon), real fields f = sum_{l,m>=0} c_m Re(a_lm Y_lm) with
c_0 = 1, c_m = 2.  P_lm are generated on the GPU in FP64.
"""
import ctypes

import numpy as np

from ._arrays import context
from ._lib import GD_PREC_BF16X3, GD_PREC_FP32, GD_PREC_FP64, check, lib

PRECISIONS = {"fp64": GD_PREC_FP64, "fp32": GD_PREC_FP32, "bf16x3": GD_PREC_BF16X3}


class SHTnsKitConfig:
    """shtnskit_transforms.jl:61-93 (grid, (l,m) ordering, GPU plan)."""

    def __init__(self, lmax, mmax, nlat, nlon, precision="fp64", device=0):
        self.lmax, self.mmax, self.nlat, self.nlon = lmax, mmax, nlat, nlon
        self.precision = precision
        self.device = device
        self._plan = ctypes.c_void_p()
        check(lib().gd_sht_plan_create(context(device), lmax, mmax, nlat, nlon,
                                       PRECISIONS[precision], ctypes.byref(self._plan)))
        self.nlm = lib().gd_sht_nlm(self._plan)
        x, w, th = np.zeros(nlat), np.zeros(nlat), np.zeros(nlat)
        check(lib().gd_sht_grid(self._plan, x.ctypes.data, w.ctypes.data, th.ctypes.data))
        self.cos_theta, self.gauss_weights, self.theta_grid = x, w, th
        self.phi_grid = 2.0 * np.pi * np.arange(nlon) / nlon
        # l-major ordering (shtnskit_transforms.jl:177-188)
        self.l_values = np.array([l for l in range(lmax + 1) for m in range(min(l, mmax) + 1)], dtype=np.int64)
        self.m_values = np.array([m for l in range(lmax + 1) for m in range(min(l, mmax) + 1)], dtype=np.int64)

    def plm_tables(self):
        """FP64 (P_lm, dP_lm/dtheta) at all nodes, shape (nlm, nlat), l-major."""
        P, dP = np.zeros((self.nlm, self.nlat)), np.zeros((self.nlm, self.nlat))
        check(lib().gd_sht_tables(self._plan, P.ctypes.data, dP.ctypes.data))
        return P, dP

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            lib().gd_sht_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_shtnskit_config(lmax, mmax=None, nlat=None, nlon=None, precision="fp64", device=0):
    """shtnskit_transforms.jl:102-200.  Defaults follow the reference:
    nlat = max(lmax+2, ...), nlon = max(2 lmax + 1, 4); nlat is rounded up to
    even and nlon to the next 2^a 3^b length the ring FFT supports."""
    mmax = lmax if mmax is None else mmax
    if nlat is None:
        nlat = lmax + 2
        nlat += nlat % 2
    if nlon is None:
        nlon = max(2 * lmax + 1, 4)
        while True:
            k = nlon
            for p in (2, 3):
                while k % p == 0:
                    k //= p
            if k == 1:
                break
            nlon += 1
    return SHTnsKitConfig(lmax, mmax, nlat, nlon, precision, device)
