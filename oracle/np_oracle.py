"""Independent numpy/scipy restatement of the CNAB2 step at tiny sizes.

TEST INFRASTRUCTURE ONLY.  It cross-checks the C oracle (gd_oracle.c) with
different machinery, so that the oracle is not only self-consistent:
  * Gauss nodes/weights from numpy.polynomial.legendre.leggauss
  * orthonormal Y_lm (Condon-Shortley phase) from scipy.special.sph_harm_y;
    d/dtheta through the ladder identity
        dY_l^m/dtheta = m cot(theta) Y_l^m + sqrt((l-m)(l+m+1)) e^{-i phi} Y_l^{m+1}
  * longitude transforms with numpy.fft.rfft / irfft
  * FD weights with numpy.linalg.solve(V.T, k! e_k), dense radial operators and
    dense pivoted solves in place of the reference's no-pivot band LU.
Semantics follow the same reference lines as gd_oracle.c (velocity.jl:427-435,
512-527, 664-670; magnetic.jl:335-343, 399-401, 437-440;
scalar_field_common.jl:401-407; timestep.jl:498-543, 610-730;
simulation.jl:1151-1330; thermal.jl:262-313).
"""
import math

import numpy as np
from scipy.special import sph_harm_y


def lm_list(lmax, mmax):
    return [(l, m) for l in range(lmax + 1) for m in range(min(l, mmax) + 1)]


def radial_grid(N, rratio):
    n = np.arange(1, N + 1)
    r = 0.5 * (1.0 + np.cos(np.pi * (N - n) / (N - 1))) + rratio / (1.0 - rratio)
    return r


def fd_dense(order, r, bw):
    N = r.size
    D = np.zeros((N, N))
    for n in range(N):
        lo, hi = max(0, n - bw), min(N - 1, n + bw)
        h = r[lo:hi + 1] - r[n]
        V = np.vander(h, increasing=True)
        rhs = np.zeros(h.size)
        if order < h.size:
            rhs[order] = math.factorial(order)
        D[n, lo:hi + 1] = np.linalg.solve(V.T, rhs)
    return D


class NpSHT:
    def __init__(self, lmax, mmax, nlat, nlon):
        self.lmax, self.mmax, self.nlat, self.nlon = lmax, mmax, nlat, nlon
        x, w = np.polynomial.legendre.leggauss(nlat)
        order = np.argsort(-x)  # north first
        self.x, self.w = x[order], w[order]
        self.theta = np.arccos(self.x)
        self.lms = lm_list(lmax, mmax)
        th = self.theta
        self.P = np.array([sph_harm_y(l, m, th, 0.0).real for l, m in self.lms])
        dP = []
        for l, m in self.lms:
            y = m / np.tan(th) * sph_harm_y(l, m, th, 0.0).real
            if m + 1 <= l:
                y = y + math.sqrt((l - m) * (l + m + 1)) * sph_harm_y(l, m + 1, th, 0.0).real
            dP.append(y)
        self.dP = np.array(dP)
        s = np.sin(th)
        self.Q = np.array([m * self.P[q] / s for q, (l, m) in enumerate(self.lms)])
        self.m = np.array([m for l, m in self.lms])
        self.l = np.array([l for l, m in self.lms])

    def _ring_synth(self, F):  # F: (mmax+1, nlat) complex
        Z = np.zeros((self.nlon // 2 + 1, self.nlat), complex)
        Z[: self.mmax + 1] = F
        Z[0] = Z[0].real
        return np.fft.irfft(Z, n=self.nlon, axis=0) * self.nlon  # (nlon, nlat)

    def _ring_analysis(self, f):  # f: (nlon, nlat)
        return np.fft.rfft(f, axis=0)[: self.mmax + 1] / self.nlon

    def _per_m(self, a, tab):
        F = np.zeros((self.mmax + 1, self.nlat), complex)
        for q, (l, m) in enumerate(self.lms):
            F[m] += a[q] * tab[q]
        return F

    def scalar_synth(self, a):
        return self._ring_synth(self._per_m(a, self.P))

    def scalar_analysis(self, f):
        G = self._ring_analysis(f)
        out = np.array([2 * np.pi * np.sum(self.w * self.P[q] * G[m]) for q, (l, m) in enumerate(self.lms)])
        out[self.m == 0] = out[self.m == 0].real
        return out

    def vector_synth(self, tor, pol):
        Ft = self._per_m(pol, self.dP) + 1j * self._per_m(tor, self.Q)
        Fp = 1j * self._per_m(pol, self.Q) - self._per_m(tor, self.dP)
        return self._ring_synth(Ft), self._ring_synth(Fp)

    def vector_analysis(self, vt, vp):
        Gt, Gp = self._ring_analysis(vt), self._ring_analysis(vp)
        S = np.zeros(len(self.lms), complex)
        T = np.zeros(len(self.lms), complex)
        for q, (l, m) in enumerate(self.lms):
            if l == 0:
                continue
            f = 2 * np.pi / (l * (l + 1))
            S[q] = f * np.sum(self.w * (Gt[m] * self.dP[q] - 1j * self.Q[q] * Gp[m]))
            T[q] = f * np.sum(self.w * (-1j * self.Q[q] * Gt[m] - self.dP[q] * Gp[m]))
        S[self.m == 0] = S[self.m == 0].real
        T[self.m == 0] = T[self.m == 0].real
        return T, S


class NpModel:
    """Field order: u_tor, u_pol, B_tor, B_pol, T; state shape (5, nr, nlm) complex."""

    def __init__(self, lmax, nlat, nlon, nr, E=1e-3, Pm=5.0, Pr=1.0, Ra=100.0, rratio=0.35,
                 dt=1e-5, theta=0.5, bw=3, radials=False, coriolis_index=0, bc_rows=1, radial_force=0):
        self.sht = NpSHT(lmax, lmax, nlat, nlon)
        self.nr, self.E, self.Pm, self.Pr, self.dt, self.th = nr, E, Pm, Pr, dt, theta
        self.radials, self.coriolis_index, self.bc_rows = radials, coriolis_index, bc_rows
        # buoyancy coefficient of velocity.jl:534-537 when the radial forcing is on
        self.radial_force = radial_force
        self.cb = (Pm / E) * (Pm / Pr) * Ra if radial_force else 0.0
        self.r = radial_grid(nr, rratio)
        self.D1 = fd_dense(1, self.r, bw)
        self.D2 = fd_dense(2, self.r, bw)
        self.lap = self.D2 + np.diag(2.0 / self.r) @ self.D1
        ll = self.sht.l * (self.sht.l + 1.0)
        self.ll = ll
        self.L, self.S = {}, {}
        for k, nu in enumerate((Pm, 1.0, Pm / Pr)):
            for l in range(lmax + 1):
                Ll = nu * self.lap - nu * l * (l + 1) * np.diag(self.r ** -2.0)
                self.L[k, l] = Ll
                S = np.eye(nr) / dt - theta * Ll
                if bc_rows == 1 or (bc_rows == 2 and k == 2):  # Dirichlet rows, value 0
                    S[0, :] = 0.0
                    S[-1, :] = 0.0
                    S[0, 0] = S[-1, -1] = 1.0
                self.S[k, l] = S

    def nonlinear(self, st):
        sht, r, ll = self.sht, self.r, self.ll
        utor, upol, Btor, Bpol, T = st
        r2 = (r ** -2.0)[:, None]
        wtor = ll * r2 * upol - self.D2 @ upol - 2.0 / r[:, None] * (self.D1 @ upol)
        wpol = -ll * r2 * utor
        jtor = ll * r2 * Bpol - self.D2 @ Bpol - 2.0 / r[:, None] * (self.D1 @ Bpol)
        jpol = -ll * r2 * Btor
        f = 1.0 if self.radials else 0.0
        ur, wr, Br, jr = (f * ll * upol * r2, f * ll * wpol * r2, f * ll * Bpol * r2, f * ll * jpol * r2)
        Tg = T / r[:, None]
        dT = self.D1 @ T
        ro = self.Pm / self.E
        nl = np.zeros_like(st)
        nlat, nlon = sht.nlat, sht.nlon
        jj = np.arange(nlon)
        ti = np.minimum(jj, nlat - 1)
        if self.coriolis_index:
            sth = np.broadcast_to(np.sin(sht.theta)[None, :], (nlon, nlat))
            cth = np.broadcast_to(np.cos(sht.theta)[None, :], (nlon, nlat))
        else:
            sth = np.broadcast_to(np.sin(sht.theta[ti])[:, None], (nlon, nlat))
            cth = np.broadcast_to(np.cos(sht.theta[ti])[:, None], (nlon, nlat))
        zero = np.zeros(len(sht.lms), complex)
        for k in range(self.nr):
            ut, up = sht.vector_synth(utor[k], upol[k])
            wt, wp = sht.vector_synth(wtor[k], wpol[k])
            Bt, Bp = sht.vector_synth(Btor[k], Bpol[k])
            jt, jp = sht.vector_synth(jtor[k], jpol[k])
            gt, gp = sht.vector_synth(zero, Tg[k])
            urr, wrr, Brr, jrr = (sht.scalar_synth(x[k]) for x in (ur, wr, Br, jr))
            gr = sht.scalar_synth(dT[k])
            at = -ro * (up * wrr - urr * wp) - ro * (-cth * up) + ro * (jp * Brr - jrr * Bp)
            ap = -ro * (urr * wt - ut * wrr) - ro * (cth * ut + sth * urr) + ro * (jrr * Bt - jt * Brr)
            bt = up * Brr - urr * Bp
            bp = urr * Bt - ut * Brr
            aT = -(urr * gr + ut * gt + up * gp)
            tor, pol = sht.vector_analysis(at, ap)
            nl[0, k], nl[1, k] = tor, pol
            tor, pol = sht.vector_analysis(bt, bp)
            nl[2, k] = ll * pol
            nl[3, k] = -ll * tor
            nl[4, k] = sht.scalar_analysis(aT)
        if self.radial_force:
            nl[1] += self.buoyancy_poloidal(T)
        return nl

    def buoyancy_poloidal(self, T):
        """The poloidal scalar whose radial component u_r = l(l+1) P / r^2
        (magnetic.jl:257-259) is the buoyancy F_r = cb r T: r^2 F_r / (l(l+1)),
        l >= 1 (l = 0 carries no poloidal flow)."""
        ll = self.ll
        inv = np.where(ll > 0, 1.0 / np.where(ll > 0, ll, 1.0), 0.0)
        return self.cb * (self.r ** 3)[:, None] * T * inv[None, :]

    def step(self, st, nl_prev, step_index):
        st = st.copy()
        nl = self.nonlinear(st)
        st[4, 0] = 0.0
        st[4, -1] = 0.0
        if step_index == 0:
            nl_prev = nl.copy()
        nu_of = (0, 0, 1, 1, 2)
        out = np.zeros_like(st)
        for f in range(5):
            k = nu_of[f]
            for q, l in enumerate(self.sht.l):
                u = st[f, :, q]
                rhs = u / self.dt + 1.5 * nl[f, :, q] - 0.5 * nl_prev[f, :, q] + (1 - self.th) * (self.L[k, l] @ u)
                if self.bc_rows == 1 or (self.bc_rows == 2 and k == 2):
                    rhs[0] = rhs[-1] = 0.0
                out[f, :, q] = np.linalg.solve(self.S[k, l], rhs)
        return out, nl, nl.copy()
