"""Pins of the FP64 oracle adjoint (oracle/np_adjoint.py), CPU.

The adjoint restatement is checked against the C oracle's forward step
(gd_oracle.c), not against itself:
  * its forward pieces equal the C oracle's nonlinear terms and steps;
  * dot-product tests <w, J v> = <J^T w, v> where J v is a central difference
    of the C oracle's step or nonlinear terms (exact up to rounding: the step is
    quadratic in the state and linear in the AB2 history, so the difference of
    F(x + v) and F(x - v) is 2 J v);
  * central finite differences of a surface misfit over several steps of the
    C oracle against the adjoint gradient;
  * planted mistakes in a copy of the restatement fail these checks.
"""
import numpy as np
import pytest

TINY = dict(lmax=7, nlat=12, nlon=24, nr=9)


def _dot(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def _rand(rng, lmax, nr, nlm, scale=1.0):
    z = rng.standard_normal((5, nr, nlm)) + 1j * rng.standard_normal((5, nr, nlm))
    return scale * z


@pytest.fixture(scope="module")
def adj():
    from oracle import np_adjoint, np_oracle

    return {bc: np_adjoint.NpAdjoint(np_oracle.NpModel(**TINY, bc_rows=bc)) for bc in (1, 0, 2)}


@pytest.mark.parametrize("bc_rows", [1, 0, 2])
def test_forward_pieces_equal_c_oracle(oracle, adj, bc_rows):
    a = adj[bc_rows]
    st = oracle.seeded_state(5, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1.0, zero_boundaries=False)
    m = oracle.Model(**TINY, bc_rows=bc_rows)
    ref = m.nonlinear(st)
    got = a.nonlinear(st)
    for f in (0, 1, 4):
        assert np.abs(got[f] - ref[f]).max() <= 1e-12 * np.abs(ref[f]).max()
    assert np.abs(ref[2]).max() == 0.0 and np.abs(ref[3]).max() == 0.0
    x1, nl0 = a.step(st, None, True)
    c1, cnl, _ = m.step(st, None, 0)
    assert np.abs(x1 - c1).max() <= 1e-11 * np.abs(c1).max()
    x2, _ = a.step(x1, nl0, False)
    c2, _, _ = m.step(c1, cnl, 1)
    assert np.abs(x2 - c2).max() <= 1e-11 * np.abs(c2).max()


def test_nonlinear_vjp_dot_product(oracle, adj):
    """<mu, N'(x) v> = <N'(x)^T mu, v> with N from the C oracle."""
    a = adj[1]
    rng = np.random.default_rng(11)
    nlm = oracle.nlm(TINY["lmax"], TINY["lmax"])
    m = oracle.Model(**TINY)
    for trial in range(3):
        x = _rand(rng, TINY["lmax"], TINY["nr"], nlm)
        v = _rand(rng, TINY["lmax"], TINY["nr"], nlm)
        mu = _rand(rng, TINY["lmax"], TINY["nr"], nlm)
        jv = 0.5 * (m.nonlinear(x + v) - m.nonlinear(x - v))
        lhs = _dot(mu, jv)
        rhs = _dot(a.nonlinear_vjp(x, mu), v)
        assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(jv).max() * np.abs(mu).sum())


@pytest.mark.parametrize("bc_rows", [1, 0, 2])
@pytest.mark.parametrize("boot", [True, False])
def test_step_vjp_dot_product(oracle, adj, bc_rows, boot):
    """<lam, dx'> + <eta, dnl> = <lam_out, v> + <eta_out, w> for the C oracle's step."""
    a = adj[bc_rows]
    rng = np.random.default_rng(3 + bc_rows + 10 * boot)
    nlm = oracle.nlm(TINY["lmax"], TINY["lmax"])
    m = oracle.Model(**TINY, bc_rows=bc_rows, dt=1e-2)
    sh = (TINY["lmax"], TINY["nr"], nlm)
    x, h = _rand(rng, *sh), _rand(rng, *sh)
    v, w = _rand(rng, *sh), _rand(rng, *sh)
    lam, eta = _rand(rng, *sh), _rand(rng, *sh)
    k = 0 if boot else 1

    def F(xx, hh):
        xn, nl, _ = m.step(xx, None if boot else hh, k)
        return xn, nl

    xp, nlp = F(x + v, h + w)
    xm, nlm_ = F(x - v, h - w)
    lhs = _dot(lam, 0.5 * (xp - xm)) + _dot(eta, 0.5 * (nlp - nlm_))
    import oracle.np_adjoint as npa  # noqa: F401  (the model of `a` has dt = 1e-5; rebuild at dt = 1e-2)
    from oracle import np_oracle

    b = npa.NpAdjoint(np_oracle.NpModel(**TINY, bc_rows=bc_rows, dt=1e-2)) if a.m.dt != 1e-2 else a
    lo, eo = b.step_vjp(x, boot, lam, eta)
    rhs = _dot(lo, v) + (0.0 if boot else _dot(eo, w))
    scale = np.abs(lam).sum() * np.abs(xp - xm).max() + np.abs(eta).sum() * np.abs(nlp - nlm_).max()
    assert abs(lhs - rhs) <= 1e-12 * scale
    if boot:
        assert np.abs(eo).max() == 0.0


@pytest.mark.parametrize("kind", ["B_r", "T", "u_r"])
def test_misfit_gradient_finite_differences(oracle, adj, kind):
    """Central differences of J(x0) = 1/2 ||X(x_K) - target||^2 over 3 steps of
    the C oracle against the adjoint gradient, on 6 random directions."""
    from oracle import np_adjoint, np_oracle

    a = np_adjoint.NpAdjoint(np_oracle.NpModel(**TINY, dt=1e-3))
    rng = np.random.default_rng(21)
    nlm = oracle.nlm(TINY["lmax"], TINY["lmax"])
    x0 = oracle.seeded_state(9, TINY["lmax"], TINY["lmax"], TINY["nr"], amp=1.0)
    level = TINY["nr"] - 2
    target = rng.standard_normal((TINY["nlon"], TINY["nlat"]))
    K = 3
    m = oracle.Model(**TINY, dt=1e-3)

    def J(x):
        nlp = None
        for k in range(K):
            x, nl, nlp = m.step(x, nlp, k)
        return a.misfit(x, kind, level, target)

    xs = a.run(x0, K)
    g, _ = a.gradient(x0, K, a.misfit_seed(xs[-1], kind, level, target))
    assert np.abs(g).max() > 0
    for t in range(6):
        d = _rand(rng, TINY["lmax"], TINY["nr"], nlm)
        d[:, 0] = d[:, -1] = 0
        eps = 1e-5
        fd = (J(x0 + eps * d) - J(x0 - eps * d)) / (2 * eps)
        ad = _dot(g, d)
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-8 * np.abs(g).sum()), (t, fd, ad)


def test_planted_mistakes_fail(oracle, adj):
    """A dropped advection term, a wrong Coriolis sign or an untransposed solve
    in the restatement each break the dot-product test."""
    from oracle import np_adjoint, np_oracle

    rng = np.random.default_rng(5)
    nlm = oracle.nlm(TINY["lmax"], TINY["lmax"])
    m = oracle.Model(**TINY, dt=1e-2)
    sh = (TINY["lmax"], TINY["nr"], nlm)
    x, v, lam = _rand(rng, *sh), _rand(rng, *sh), _rand(rng, *sh)
    eta = np.zeros_like(x)
    xp, nlp, _ = m.step(x + v, None, 0)
    xm, nlm_, _ = m.step(x - v, None, 0)
    lhs = _dot(lam, 0.5 * (xp - xm))

    class Sign(np_adjoint.NpAdjoint):
        def nonlinear_vjp(self, st, mu):
            self.rc = -self.rc
            out = super().nonlinear_vjp(st, mu)
            self.rc = -self.rc
            return out

    class NoTranspose(np_adjoint.NpAdjoint):
        def step_vjp(self, st, boot, lam, eta):
            S = self.m.S
            self.m.S = {k: s.T for k, s in S.items()}
            try:
                return super().step_vjp(st, boot, lam, eta)
            finally:
                self.m.S = S

    class NoAdvection(np_adjoint.NpAdjoint):
        def nonlinear_vjp(self, st, mu):
            mu = mu.copy()
            mu[4] = 0
            return super().nonlinear_vjp(st, mu)

    good = np_adjoint.NpAdjoint(np_oracle.NpModel(**TINY, dt=1e-2))
    lo, _ = good.step_vjp(x, True, lam, eta)
    scale = np.abs(lam).sum() * np.abs(xp - xm).max()
    assert abs(lhs - _dot(lo, v)) <= 1e-12 * scale
    for cls in (Sign, NoTranspose, NoAdvection):
        bad = cls(np_oracle.NpModel(**TINY, dt=1e-2))
        lo, _ = bad.step_vjp(x, True, lam, eta)
        assert abs(lhs - _dot(lo, v)) > 1e-8 * scale, cls.__name__


def test_step_vjp_dot_product_with_radial_force(oracle):
    """The adjoint restatement with the radial forcing against the C oracle's
    step with it (dot-product test)."""
    from oracle import np_adjoint, np_oracle

    rng = np.random.default_rng(13)
    nlm = oracle.nlm(TINY["lmax"], TINY["lmax"])
    m = oracle.Model(**TINY, dt=1e-2, Ra=400.0, radial_force=1)
    a = np_adjoint.NpAdjoint(np_oracle.NpModel(**TINY, dt=1e-2, Ra=400.0, radial_force=1))
    sh = (TINY["lmax"], TINY["nr"], nlm)
    x, h, v, w = (_rand(rng, *sh) for _ in range(4))
    lam, eta = _rand(rng, *sh), _rand(rng, *sh)
    xp, nlp, _ = m.step(x + v, h + w, 1)
    xm, nlm_, _ = m.step(x - v, h - w, 1)
    lhs = _dot(lam, 0.5 * (xp - xm)) + _dot(eta, 0.5 * (nlp - nlm_))
    lo, eo = a.step_vjp(x, False, lam, eta)
    rhs = _dot(lo, v) + _dot(eo, w)
    scale = np.abs(lam).sum() * np.abs(xp - xm).max() + np.abs(eta).sum() * np.abs(nlp - nlm_).max()
    assert abs(lhs - rhs) <= 1e-12 * scale
