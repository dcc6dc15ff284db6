"""Reverse-mode derivatives of the CNAB2 step (libgdhip's gd_step_vjp,
gd_step_nonlinear_adjoint and gd_adjoint_*).

The reference differentiates through SHTnsKit.jl's ChainRulesCore extension
(Manifest.toml:657-669); here the transposed step is hand-written: the
transposed banded solves (linear_algebra.jl:59-88, timestep.jl:610-730) and
the transposed nonlinear terms through the transposed transform pair
(shtnskit_field_functions.jl:252-321).  Arrays are complex (5, nr, nlm) in the
l-major layout of CNAB2Step; a cotangent holds dJ/d(re) + i dJ/d(im).
"""
import ctypes

import numpy as np

from ._arrays import prepare
from ._lib import check, lib

MISFIT_KINDS = {"B_r": 0, "u_r": 1, "T": 2}


def _split(z):
    return np.ascontiguousarray(z.real, dtype=np.float64), np.ascontiguousarray(z.imag, dtype=np.float64)


def _check_shape(step, z, what):
    if z.shape != step.shape:
        raise ValueError("%s: shape %s, the step's planes are %s" % (what, z.shape, step.shape))


def nonlinear_adjoint(step, mu):
    """N'(x)^T mu at the step's current state x (N: CNAB2Step.nonlinear)."""
    _check_shape(step, mu, "nonlinear_adjoint")
    a, b = _split(mu)
    re, im = np.zeros(step.shape), np.zeros(step.shape)
    ctx, flags, p = prepare(a, b, re, im, device=step.device)
    check(lib().gd_step_nonlinear_adjoint(step._h, p[0], p[1], p[2], p[3], flags))
    return re + 1j * im


def step_vjp(step, lam, eta):
    """The transposed step at the current state: (cotangent of the new state,
    cotangent of this step's nonlinear terms) -> (cotangent of the state,
    cotangent of the previous step's nonlinear terms; zero at step 0)."""
    _check_shape(step, lam, "step_vjp")
    _check_shape(step, eta, "step_vjp")
    a, b = _split(lam)
    c, d = _split(eta)
    out = [np.zeros(step.shape) for _ in range(4)]
    ctx, flags, p = prepare(a, b, c, d, *out, device=step.device)
    check(lib().gd_step_vjp(step._h, *p, flags))
    return out[0] + 1j * out[1], out[2] + 1j * out[3]


def surface_level(step, kind="B_r"):
    """The outermost radial level the implicit solve determines for the
    field: nr-2 where the field's S_l has Dirichlet rows (its level nr-1 is
    then 0 after every step), nr-1 otherwise."""
    bc = step.desc.bc_rows
    dirichlet = bc == 1 or (bc == 2 and kind == "T")
    return step.nr - 2 if dirichlet else step.nr - 1


class CNAB2Adjoint:
    """Gradient of a functional of the state after `nsteps` CNAB2 steps with
    respect to the initial state, checkpointing every `every` steps."""

    def __init__(self, step, nsteps, every=5):
        self.step, self.nsteps, self.every = step, int(nsteps), int(every)
        self._h = ctypes.c_void_p()
        check(lib().gd_adjoint_create(step._h, self.nsteps, self.every, ctypes.byref(self._h)))

    def forward(self):
        """Runs the steps from the step's current state (step count 0)."""
        check(lib().gd_adjoint_forward(self._h))

    def misfit(self, target, kind="B_r", level=None):
        """J = 1/2 ||X(level) - target||^2 on the (nlat, nlon) grid
        (colatitude fastest: target[j, i] at longitude j, colatitude i);
        seeds the final cotangent and returns J."""
        s = self.step
        level = surface_level(s, kind) if level is None else int(level)
        t = np.ascontiguousarray(target, dtype=np.float64)
        if t.size != s.nlat * s.nlon:
            raise ValueError("target needs nlat * nlon = %d values" % (s.nlat * s.nlon))
        J = ctypes.c_double()
        ctx, flags, (p,) = prepare(t, device=s.device)
        check(lib().gd_adjoint_misfit(self._h, MISFIT_KINDS[kind], level, p, ctypes.byref(J), flags))
        return J.value

    def seed(self, g):
        _check_shape(self.step, g, "seed")
        a, b = _split(g)
        ctx, flags, p = prepare(a, b, device=self.step.device)
        check(lib().gd_adjoint_seed(self._h, p[0], p[1], flags))

    def backward(self, out=None):
        """The gradient with respect to the initial state: complex (5, nr, nlm)
        on the host, or, with out = (re, im) float64 planes (host arrays or
        device tensors of the step's shape), written there."""
        if out is None:
            re, im = np.zeros(self.step.shape), np.zeros(self.step.shape)
        else:
            re, im = out
        ctx, flags, p = prepare(re, im, device=self.step.device, shapes=[self.step.shape] * 2)
        check(lib().gd_adjoint_backward(self._h, p[0], p[1], flags))
        return (re + 1j * im) if out is None else out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gd_adjoint_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
