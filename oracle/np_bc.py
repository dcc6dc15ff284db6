"""Scalar boundary conditions in spectral space: numpy restatement.

TEST INFRASTRUCTURE ONLY (the checker of libgdhip's gd_scalar_bc_*; nothing
in the product path imports it).

Follows, literally:
  * apply_temperature_boundary_conditions_spectral! (thermal.jl:262-313) and
    apply_composition_boundary_conditions_spectral! (compositional.jl:318-370):
    bc type 1 sets the boundary row to the prescribed value (imag 0); then, if
    any mode has type 2, the flux correction below;
  * apply_scalar_flux_bc_spectral!(...; method=:tau)
    (scalar_field_common.jl:1166-1199) and apply_flux_bc_tau! (:866-929):
    per mode with a flux condition, the flux error of the REAL profile at the
    boundaries (D1 rows 1 and N, compute_boundary_fluxes :813-820) fixes the
    coefficients of the tau functions (compute_tau_coefficients_both :656-683
    with analytic Chebyshev end derivatives, _inner :690-706, _outer :712-727
    with derivatives of the single tau function through D1, :771-790); the
    correction is added to the real profile and, only when some |imag| >
    1e-12, the same correction to the imaginary profile (:917-920);
  * compute_chebyshev_polynomial (:573-590), evaluate_chebyshev_derivative
    (:596-622), _get_tau_cache (:629-650), compute_inner/outer_tau_function
    (:733-765), get_flux_value (:827-860: the boundary_values matrix).
Arrays: profiles (nr, nlm) C-order (the reference's (nlm, 1, nr)); values
(2, nlm) with row 0 inner, row 1 outer; D1 dense (nr, nr).
"""
import math

import numpy as np


def chebyshev_polynomial(n, r):
    ri, ro = r[0], r[-1]
    x = 2.0 * (r - ri) / (ro - ri) - 1.0
    return np.cos(n * np.arccos(np.clip(x, -1.0, 1.0)))


def chebyshev_derivative(n, rv, r):
    ri, ro = r[0], r[-1]
    x = 2.0 * (rv - ri) / (ro - ri) - 1.0
    if abs(abs(x) - 1.0) < 1e-12:
        d = n ** 2 * math.copysign(1.0, x) ** (n + 1)
    else:
        th = math.acos(min(max(x, -1.0), 1.0))
        d = n * math.sin(n * th) / math.sin(th)
    return d * (2.0 / (ro - ri))


class TauCache:
    def __init__(self, r, D1):
        nr = r.size
        self.tau1 = chebyshev_polynomial(nr - 1, r)
        self.tau2 = chebyshev_polynomial(nr, r)
        self.dt1i = chebyshev_derivative(nr - 1, r[0], r)
        self.dt1o = chebyshev_derivative(nr - 1, r[-1], r)
        self.dt2i = chebyshev_derivative(nr, r[0], r)
        self.dt2o = chebyshev_derivative(nr, r[-1], r)
        a = self.dt1o / self.dt2o if abs(self.dt2o) > 1e-12 else 0.0
        b = self.dt1i / self.dt2i if abs(self.dt2i) > 1e-12 else 0.0
        self.tau_in = self.tau1 - a * self.tau2
        self.tau_out = self.tau1 - b * self.tau2
        self.dtau_in = (D1 @ self.tau_in)[0]
        self.dtau_out = (D1 @ self.tau_out)[-1]
        self.D1 = D1


def boundary_fluxes(profile, D1):
    d = D1 @ profile
    return d[0], d[-1]


def flux_correction(tc, profile_real, apply_inner, apply_outer, flux_inner, flux_outer):
    """The correction vector apply_flux_bc_tau! adds (scalar_field_common.jl:896-916)."""
    fi = flux_inner if apply_inner else 0.0
    fo = flux_outer if apply_outer else 0.0
    ci, co = boundary_fluxes(profile_real, tc.D1)
    if apply_inner and apply_outer:
        ei, eo = fi - ci, fo - co
        det = tc.dt1i * tc.dt2o - tc.dt1o * tc.dt2i
        if abs(det) > 1e-12:
            c1 = (ei * tc.dt2o - eo * tc.dt2i) / det
            c2 = (eo * tc.dt1i - ei * tc.dt1o) / det
        else:
            c1 = c2 = 0.0
        return c1 * tc.tau1 + c2 * tc.tau2
    if apply_inner:
        return ((fi - ci) / tc.dtau_in) * tc.tau_in
    return ((fo - co) / tc.dtau_out) * tc.tau_out


def apply_scalar_bc(re, im, bc_inner, bc_outer, values, r, D1, dirichlet=True):
    """Returns corrected copies of re, im (nr, nlm)."""
    re, im = np.array(re, dtype=float), np.array(im, dtype=float)
    nr, nlm = re.shape
    tc = TauCache(r, D1)
    if dirichlet:
        for k in range(nlm):
            if r[0] > 0 and bc_inner[k] == 1:
                re[0, k] = values[0, k]
                im[0, k] = 0.0
            if bc_outer[k] == 1:
                re[nr - 1, k] = values[1, k]
                im[nr - 1, k] = 0.0
    for k in range(nlm):
        ai, ao = bc_inner[k] == 2, bc_outer[k] == 2
        if not (ai or ao):
            continue
        corr = flux_correction(tc, re[:, k], ai, ao, values[0, k], values[1, k])
        re[:, k] = re[:, k] + corr
        if np.any(np.abs(im[:, k]) > 1e-12):
            im[:, k] = im[:, k] + corr
    return re, im
