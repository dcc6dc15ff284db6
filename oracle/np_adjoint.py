"""FP64 CPU adjoint of the CNAB2 step at tiny sizes (numpy).

TEST INFRASTRUCTURE ONLY: the checker for libgdhip's adjoint, never part of
the product path.  It differentiates the numpy restatement of the step
(np_oracle.NpModel, the reading the C oracle and the GPU implement) in
reverse mode with its own machinery:

  * every transform is a dense real matrix assembled from the numpy forward
    map column by column (the maps are linear), and its transpose is that
    matrix's transpose -- no transpose identity of the transforms is used;
  * the products' Jacobians are written out pointwise for the tangential step
    (radials = 0): a_t = ro cos u_p, a_p = -ro cos u_t, a_T = -(u_t g_t + u_p g_p)
    (velocity.jl:512-527, scalar_field_common.jl:401-407);
  * the implicit solves are transposed dense solves S_l^T (timestep.jl:610-730,
    the reference's rhs with the Dirichlet rows of the reading).

Pinned in tests/test_oracle_adjoint.py by dot-product tests against
central differences of the C oracle's step (exact for the step, which is
quadratic in the state and linear in the history) and by finite differences
of a surface misfit over several steps.

Layouts: complex (5, nr, nlm), fields u_tor, u_pol, B_tor, B_pol, T, l-major
modes; a cotangent holds dJ/d(re) + i dJ/d(im).
"""
import numpy as np

from .np_oracle import NpModel


def _rv(z):
    return np.concatenate([z.real.ravel(), z.imag.ravel()])


def _zv(v):
    n = v.size // 2
    return v[:n] + 1j * v[n:]


class NpAdjoint:
    def __init__(self, model: NpModel):
        if model.radials:
            raise ValueError("the adjoint restatement covers the tangential step (radials = 0)")
        self.m = model
        sht = model.sht
        nlm, nlat, nlon = len(sht.lms), sht.nlat, sht.nlon
        ng = nlat * nlon
        self.nlm, self.ng = nlm, ng
        zero = np.zeros(nlm, complex)
        # scalar synthesis: [re; im] (2 nlm) -> grid (ng)
        self.Ms = np.zeros((ng, 2 * nlm))
        for k in range(2 * nlm):
            e = np.zeros(2 * nlm)
            e[k] = 1.0
            self.Ms[:, k] = sht.scalar_synth(_zv(e)).ravel()
        # vector synthesis: [tor re; tor im; pol re; pol im] -> [v_t; v_p]
        self.Mv = np.zeros((2 * ng, 4 * nlm))
        for k in range(4 * nlm):
            e = np.zeros(4 * nlm)
            e[k] = 1.0
            vt, vp = sht.vector_synth(_zv(e[: 2 * nlm]), _zv(e[2 * nlm:]))
            self.Mv[:, k] = np.concatenate([vt.ravel(), vp.ravel()])
        # scalar analysis: grid -> [re; im]; vector analysis: [v_t; v_p] -> [tor; pol]
        self.As = np.zeros((2 * nlm, ng))
        self.Av = np.zeros((4 * nlm, 2 * ng))
        for k in range(ng):
            e = np.zeros(ng)
            e[k] = 1.0
            f = e.reshape(nlon, nlat)
            self.As[:, k] = _rv(sht.scalar_analysis(f))
            T, S = sht.vector_analysis(f, 0 * f)
            self.Av[:, k] = np.concatenate([_rv(T), _rv(S)])
            T, S = sht.vector_analysis(0 * f, f)
            self.Av[:, ng + k] = np.concatenate([_rv(T), _rv(S)])
        del zero
        nlat_, th = sht.nlat, sht.theta
        jj = np.arange(nlon)
        if model.coriolis_index:
            cth = np.broadcast_to(np.cos(th)[None, :], (nlon, nlat_))
        else:
            cth = np.broadcast_to(np.cos(th[np.minimum(jj, nlat_ - 1)])[:, None], (nlon, nlat_))
        self.rc = (model.Pm / model.E) * cth.ravel()

    # ---- forward pieces at one level
    def _fields(self, st, k):
        m, nlm, ng = self.m, self.nlm, self.ng
        u = self.Mv @ np.concatenate([_rv(st[0, k]), _rv(st[1, k])])
        g = self.Mv @ np.concatenate([np.zeros(2 * nlm), _rv(st[4, k] / m.r[k])])
        return u[:ng], u[ng:], g[:ng], g[ng:]

    def nonlinear(self, st):
        """The tangential step's nonlinear terms (magnetic ones zero)."""
        nlm, ng = self.nlm, self.ng
        out = np.zeros_like(st)
        for k in range(self.m.nr):
            ut, up, gt, gp = self._fields(st, k)
            at, ap, aT = self.rc * up, -self.rc * ut, -(ut * gt + up * gp)
            v = self.Av @ np.concatenate([at, ap])
            out[0, k], out[1, k] = _zv(v[: 2 * nlm]), _zv(v[2 * nlm:])
            out[4, k] = _zv(self.As @ aT)
        if self.m.radial_force:
            out[1] += self.m.buoyancy_poloidal(st[4])
        return out

    def nonlinear_vjp(self, st, mu):
        """N'(st)^T mu."""
        nlm, ng, r = self.nlm, self.ng, self.m.r
        lam = np.zeros_like(st)
        for k in range(self.m.nr):
            ut, up, gt, gp = self._fields(st, k)
            b = self.Av.T @ np.concatenate([_rv(mu[0, k]), _rv(mu[1, k])])
            bt, bp = b[:ng], b[ng:]
            bT = self.As.T @ _rv(mu[4, k])
            ubt = -self.rc * bp - gt * bT
            ubp = self.rc * bt - gp * bT
            gbt, gbp = -ut * bT, -up * bT
            x = self.Mv.T @ np.concatenate([ubt, ubp])
            lam[0, k] += _zv(x[: 2 * nlm])
            lam[1, k] += _zv(x[2 * nlm:])
            y = self.Mv.T @ np.concatenate([gbt, gbp])
            lam[4, k] += _zv(y[2 * nlm:]) / r[k]
        if self.m.radial_force:
            # the forcing term is linear in T with a real per-(l, r) factor
            lam[4] += self.m.buoyancy_poloidal(mu[1])
        return lam

    def step(self, st, nl_prev, boot):
        """The forward step of the reading (as NpModel.step, tangential products)."""
        m = self.m
        x = st.copy()
        nl = self.nonlinear(x)
        x[4, 0] = 0.0
        x[4, -1] = 0.0
        if boot:
            nl_prev = nl
        out = np.zeros_like(st)
        for f in range(5):
            kk = (0, 0, 1, 1, 2)[f]
            for q, l in enumerate(m.sht.l):
                u = x[f, :, q]
                rhs = u / m.dt + 1.5 * nl[f, :, q] - 0.5 * nl_prev[f, :, q] + (1 - m.th) * (m.L[kk, l] @ u)
                if m.bc_rows == 1 or (m.bc_rows == 2 and kk == 2):
                    rhs[0] = rhs[-1] = 0.0
                out[f, :, q] = np.linalg.solve(m.S[kk, l], rhs)
        return out, nl

    def step_vjp(self, st, boot, lam, eta):
        """(cotangent of the new state, of this step's nl) -> (cotangent of st,
        of the previous step's nl)."""
        m = self.m
        c1, c0 = (1.0, 0.0) if boot else (1.5, 0.5)
        lam_out = np.zeros_like(st)
        rbar = np.zeros_like(st)
        for f in range(5):
            kk = (0, 0, 1, 1, 2)[f]
            dirichlet = m.bc_rows == 1 or (m.bc_rows == 2 and kk == 2)
            for q, l in enumerate(m.sht.l):
                r = np.linalg.solve(m.S[kk, l].T, lam[f, :, q])
                if dirichlet:
                    r[0] = r[-1] = 0.0
                rbar[f, :, q] = r
                E = np.eye(m.nr) / m.dt + (1 - m.th) * m.L[kk, l]
                lo = E.T @ r
                if f == 4:
                    lo[0] = lo[-1] = 0.0
                lam_out[f, :, q] = lo
        mu = eta + c1 * rbar
        lam_out += self.nonlinear_vjp(st, mu)
        return lam_out, -c0 * rbar

    def run(self, x0, nsteps):
        """States x_0 .. x_nsteps."""
        xs = [x0.copy()]
        nlp = None
        for k in range(nsteps):
            x, nl = self.step(xs[-1], nlp, k == 0)
            xs.append(x)
            nlp = nl
        return xs

    def gradient(self, x0, nsteps, seed):
        """d<seed, x_nsteps>/d x0 (seed: the cotangent of the final state)."""
        xs = self.run(x0, nsteps)
        lam = seed.copy()
        eta = np.zeros_like(x0)
        for k in range(nsteps - 1, -1, -1):
            lam, eta = self.step_vjp(xs[k], k == 0, lam, eta)
        return lam, xs

    # ---- surface misfit
    def level_field(self, x, kind, level):
        """X on the grid (nlon, nlat): B_r = l(l+1) B_pol / r^2, u_r likewise from
        u_pol (magnetic.jl:257-259), or T, at one radial level."""
        m = self.m
        ll = m.sht.l * (m.sht.l + 1.0)
        if kind == "T":
            a = x[4, level]
        else:
            a = ll * x[3 if kind == "B_r" else 1, level] / m.r[level] ** 2
        return (self.Ms @ _rv(a)).reshape(m.sht.nlon, m.sht.nlat), a

    def misfit(self, x, kind, level, target):
        X, _ = self.level_field(x, kind, level)
        return 0.5 * np.sum((X - target) ** 2)

    def misfit_seed(self, x, kind, level, target):
        m = self.m
        X, _ = self.level_field(x, kind, level)
        g = _zv(self.Ms.T @ (X - target).ravel())
        seed = np.zeros_like(x)
        ll = m.sht.l * (m.sht.l + 1.0)
        if kind == "T":
            seed[4, level] = g
        else:
            seed[3 if kind == "B_r" else 1, level] = ll * g / m.r[level] ** 2
        return seed
