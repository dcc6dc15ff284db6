"""GPU parity of the adjoint of the CNAB2 step (gd_step_vjp,
gd_step_nonlinear_adjoint, gd_adjoint_*), FP64 unless stated.

Checks:
  * dot-product tests <w, J v> = <J^T w, v> at lmax=31 / 48 x 96 / nr=33 with
    J v a central difference of the GPU's own forward (the step is quadratic in
    the state and linear in the history, so the difference is exact up to
    rounding): the nonlinear terms, the step at the bootstrap and after it;
  * the transposed step and nonlinear terms against the FP64 oracle adjoint
    (oracle/np_adjoint.py) at lmax=7 and 15;
  * whole-run gradients against the oracle adjoint (4 steps, checkpoints
    every 2, so segments are recomputed);
  * the lmax=63 / 96 x 192 / nr=65 configuration: 20 steps, checkpoints every 5,
    the surface B_r misfit's gradient with respect to the initial B_pol against
    central differences on 32 random directions (eps 1e-6), and the mixed path's
    gradient against the FP64 one;
  * a temperature misfit whose gradient passes through the products, against
    central differences with respect to u_pol, u_tor and T.
Bars: dot products 1e-12 relative; oracle agreement 1e-10; finite
differences 1e-6 relative (they carry eps^2 and rounding terms).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C0 = dict(lmax=31, nlat=48, nlon=96, nr=33)
C4 = dict(lmax=63, nlat=96, nlon=192, nr=65)


def _dot(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def _rand(rng, shape, scale=1.0):
    return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _step(prec="fp64", dt=1e-5, **kw):
    import gdhip

    return gdhip.CNAB2Step(precision=prec, dt=dt, **kw)


def test_nonlinear_adjoint_dot_product_c0(oracle):
    print("case: nonlinear adjoint dot product C0 fp64", flush=True)
    import gdhip

    g = _step(**C0)
    rng = np.random.default_rng(1)
    for t in range(3):
        x = oracle.seeded_state(30 + t, C0["lmax"], C0["lmax"], C0["nr"], amp=1.0, zero_boundaries=False)
        v = _rand(rng, g.shape)
        mu = _rand(rng, g.shape)
        g.set_state_complex(x + v)
        np_ = g.nonlinear()
        g.set_state_complex(x - v)
        nm = g.nonlinear()
        jv = 0.5 * (np_ - nm)
        g.set_state_complex(x)
        lam = gdhip.nonlinear_adjoint(g, mu)
        lhs, rhs = _dot(mu, jv), _dot(lam, v)
        scale = np.abs(mu).sum() * np.abs(jv).max()
        print("dot", lhs, rhs, abs(lhs - rhs) / scale)
        assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("boot", [True, False])
@pytest.mark.parametrize("bc_rows", [1, 0])
def test_step_vjp_dot_product_c0(oracle, boot, bc_rows):
    print("case: step vjp dot product C0 fp64 boot=%d bc_rows=%d" % (boot, bc_rows), flush=True)
    import gdhip

    g = _step(bc_rows=bc_rows, dt=1e-3, **C0)
    rng = np.random.default_rng(7 + boot)
    sh = g.shape
    x = oracle.seeded_state(40, C0["lmax"], C0["lmax"], C0["nr"], amp=1.0, zero_boundaries=False)
    h = _rand(rng, sh)
    v, w = _rand(rng, sh), _rand(rng, sh)
    lam, eta = _rand(rng, sh), _rand(rng, sh)
    # the magnetic fields' nonlinear terms are identically zero in the
    # tangential step: the forward neither forms nor reads them, so their
    # history and its cotangent stay zero
    h[2:4] = w[2:4] = eta[2:4] = 0

    def F(xx, hh):
        if boot:
            g.set_state_complex(xx)
        else:
            g.restore(xx, hh, step=1)
        g.run(1)
        return g.get_state_complex(), g.get_history()

    xp, np_ = F(x + v, h + w)
    xm, nm = F(x - v, h - w)
    lhs = _dot(lam, 0.5 * (xp - xm)) + _dot(eta, 0.5 * (np_ - nm))
    if boot:
        g.set_state_complex(x)
    else:
        g.restore(x, h, step=1)
    lo, eo = gdhip.step_vjp(g, lam, eta)
    rhs = _dot(lo, v) + (0.0 if boot else _dot(eo, w))
    scale = np.abs(lam).sum() * np.abs(xp - xm).max() + np.abs(eta).sum() * np.abs(np_ - nm).max()
    print("dot", lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-12 * scale
    if boot:
        assert np.abs(eo).max() == 0.0
    assert np.abs(eo[2:4]).max() == 0.0


@pytest.mark.parametrize("lmax,nlat,nlon,nr", [(7, 12, 24, 9), (15, 24, 48, 17)])
def test_vjp_matches_oracle_adjoint(oracle, lmax, nlat, nlon, nr):
    print("case: vjp vs oracle adjoint lmax=%d nr=%d" % (lmax, nr), flush=True)
    import gdhip
    from oracle import np_adjoint, np_oracle

    kw = dict(lmax=lmax, nlat=nlat, nlon=nlon, nr=nr)
    a = np_adjoint.NpAdjoint(np_oracle.NpModel(**kw, dt=1e-3))
    g = _step(dt=1e-3, **kw)
    rng = np.random.default_rng(lmax)
    x = oracle.seeded_state(50, lmax, lmax, nr, amp=1.0, zero_boundaries=False)
    mu = _rand(rng, g.shape)
    g.set_state_complex(x)
    got = gdhip.nonlinear_adjoint(g, mu)
    ref = a.nonlinear_vjp(x, mu)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("nonlinear adjoint rel err", err)
    assert err < 1e-10
    for boot in (True, False):
        lam, eta = _rand(rng, g.shape), _rand(rng, g.shape)
        h = _rand(rng, g.shape)
        if boot:
            g.set_state_complex(x)
        else:
            g.restore(x, h, step=1)
        eta[2:4] = 0  # identically zero in the tangential step (not read by its solve)
        lo, eo = gdhip.step_vjp(g, lam, eta)
        rl, re = a.step_vjp(x, boot, lam, eta)
        e1 = np.abs(lo - rl).max() / np.abs(rl).max()
        # the oracle reads the (zero) magnetic history and so returns its
        # cotangent; the GPU solve does not read it and returns zero there
        f3 = [0, 1, 4]
        e2 = 0.0 if boot else np.abs(eo[f3] - re[f3]).max() / np.abs(re[f3]).max()
        assert np.abs(eo[2:4]).max() == 0.0
        print("step vjp boot=%d rel err %.3e %.3e" % (boot, e1, e2))
        assert e1 < 1e-10 and e2 < 1e-10


@pytest.mark.parametrize("kind", ["B_r", "T"])
def test_run_gradient_matches_oracle_adjoint(oracle, kind):
    print("case: run gradient vs oracle adjoint kind=%s" % kind, flush=True)
    import gdhip
    from oracle import np_adjoint, np_oracle

    kw = dict(lmax=15, nlat=24, nlon=48, nr=17)
    a = np_adjoint.NpAdjoint(np_oracle.NpModel(**kw, dt=1e-3))
    g = _step(dt=1e-3, **kw)
    rng = np.random.default_rng(3)
    x0 = oracle.seeded_state(60, 15, 15, 17, amp=1.0)
    target = rng.standard_normal((48, 24))
    level = gdhip.surface_level(g, kind)
    g.set_state_complex(x0)
    adj = gdhip.CNAB2Adjoint(g, 4, every=2)
    adj.forward()
    J = adj.misfit(target, kind=kind, level=level)
    grad = adj.backward()
    xs = a.run(x0, 4)
    Jr = a.misfit(xs[-1], kind, level, target)
    gr, _ = a.gradient(x0, 4, a.misfit_seed(xs[-1], kind, level, target))
    print("J", J, Jr, "grad rel err", np.abs(grad - gr).max() / np.abs(gr).max())
    assert abs(J - Jr) <= 1e-10 * abs(Jr)
    assert np.abs(grad - gr).max() <= 1e-10 * np.abs(gr).max()
    # the step is left at the end of the forward run
    assert g.step_count == 4
    assert np.abs(g.get_state_complex() - xs[-1]).max() <= 1e-10 * np.abs(xs[-1]).max()
    # a second functional of the same forward run (every segment recomputed now)
    J2 = adj.misfit(target, kind=kind, level=level)
    grad2 = adj.backward()
    assert J2 == J and np.array_equal(grad2, grad)


def _c4_target(oracle, g):
    """B_r^target = l(l+1) P / r_o^2 on the surface grid from a seeded (l+1)^-3
    poloidal field."""
    import gdhip

    lmax, nr = g.lmax, g.nr
    P = oracle.seeded_state(99, lmax, lmax, nr, nfields=1, amp=1e-3, power=3.0, zero_boundaries=False)[0, nr - 1]
    ll = np.array([l * (l + 1.0) for l in range(lmax + 1) for m in range(l + 1)])
    ro = g.radii()[-1]
    cfg = gdhip.create_shtnskit_config(lmax=lmax, nlat=g.nlat, nlon=g.nlon, precision="fp64")
    a = ll * P / ro ** 2
    phys = np.zeros(g.nlat * g.nlon)
    from gdhip._arrays import prepare
    from gdhip._lib import check, lib

    re, im = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
    ctx, flags, p = prepare(re, im, phys)
    check(lib().gd_sht_scalar_synth(cfg._plan, p[0], p[1], 1, p[2], flags))
    return phys


def test_c4_gradient_central_differences(oracle):
    print("case: C4 B_r misfit gradient vs central differences (20 steps, checkpoints 5)", flush=True)
    import gdhip

    g = _step(**C4)
    x0 = oracle.seeded_state(2024, C4["lmax"], C4["lmax"], C4["nr"], amp=1e-3)
    target = _c4_target(oracle, g)
    g.set_state_complex(x0)
    adj = gdhip.CNAB2Adjoint(g, 20, every=5)
    adj.forward()
    J0 = adj.misfit(target)
    grad = adj.backward()
    assert np.abs(grad[3]).max() > 0
    # B is decoupled from u and T in the tangential step: only B_pol's cotangent
    # reaches it from a B_r misfit, and only B_pol's gradient is nonzero
    assert np.abs(grad[[0, 1, 2, 4]]).max() == 0.0
    rng = np.random.default_rng(8)
    eps = 1e-6
    worst = 0.0
    for t in range(32):
        d = np.zeros(g.shape, complex)
        d[3] = oracle.seeded_state(500 + t, C4["lmax"], C4["lmax"], C4["nr"], nfields=1, amp=1.0)[0]
        Js = []
        for sgn in (1, -1):
            g.set_state_complex(x0 + sgn * eps * d)
            adj.forward()
            Js.append(adj.misfit(target))
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = _dot(grad, d)
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
    print("C4: J %.6e, worst relative FD error over 32 directions %.3e" % (J0, worst))
    assert worst < 1e-6
    # the mixed path (FP32 transforms, FP64 solves) against the FP64 gradient
    m = _step(prec="fp32", **C4)
    m.set_state_complex(x0)
    am = gdhip.CNAB2Adjoint(m, 20, every=5)
    am.forward()
    Jm = am.misfit(target)
    gm = am.backward()
    e = np.abs(gm[3] - grad[3]).max() / np.abs(grad[3]).max()
    print("C4 mixed: J rel err %.3e, gradient rel err %.3e" % (abs(Jm - J0) / J0, e))
    assert e < 1e-4


def test_temperature_misfit_through_products(oracle):
    print("case: T misfit gradient through the products vs central differences (C0, 5 steps)", flush=True)
    import gdhip

    g = _step(dt=1e-4, **C0)
    x0 = oracle.seeded_state(77, C0["lmax"], C0["lmax"], C0["nr"], amp=1.0)
    rng = np.random.default_rng(9)
    target = rng.standard_normal(C0["nlat"] * C0["nlon"])
    level = C0["nr"] // 2
    g.set_state_complex(x0)
    adj = gdhip.CNAB2Adjoint(g, 5, every=2)
    adj.forward()
    adj.misfit(target, kind="T", level=level)
    grad = adj.backward()
    for f in (0, 1, 4):
        assert np.abs(grad[f]).max() > 0, f
    eps = 1e-5
    for t in range(8):
        d = np.zeros(g.shape, complex)
        d[[0, 1, 4]] = oracle.seeded_state(600 + t, C0["lmax"], C0["lmax"], C0["nr"], nfields=3, amp=1.0)
        Js = []
        for sgn in (1, -1):
            g.set_state_complex(x0 + sgn * eps * d)
            adj.forward()
            Js.append(adj.misfit(target, kind="T", level=level))
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = _dot(grad, d)
        print("T misfit direction %d: fd %.12e ad %.12e" % (t, fd, ad))
        assert abs(fd - ad) <= 1e-6 * abs(ad)


@pytest.mark.parametrize("prec,cfgkw,tol", [("fp64", dict(lmax=31, nlat=48, nlon=96), 1e-13),
                                             ("fp64", dict(lmax=255, nlat=384, nlon=768), 1e-12),
                                             ("fp32", dict(lmax=31, nlat=48, nlon=96), 1e-5)])
def test_sht_adjoints_dot_product(prec, cfgkw, tol):
    """<A x, y> = <x, A^T y> for the four transforms and their transposes."""
    print("case: SHT adjoint dot products %s %s" % (prec, cfgkw), flush=True)
    import gdhip

    cfg = gdhip.create_shtnskit_config(precision=prec, **cfgkw)
    nr = 3
    rng = np.random.default_rng(4)
    nlm, nlat, nlon = cfg.nlm, cfg.nlat, cfg.nlon

    def spec(re=None, im=None):
        return gdhip.SHTnsSpectralField(cfg, nr, re, im)

    def phys(d=None):
        return gdhip.SHTnsPhysicalField(cfg, nr, d)

    def rspec():
        return spec(rng.standard_normal((nr, nlm)), rng.standard_normal((nr, nlm)))

    def rphys():
        return phys(rng.standard_normal((nr, nlon, nlat)))

    def sdot(a, b):
        return float(np.sum(a.data_real * b.data_real + a.data_imag * b.data_imag))

    # scalar synthesis
    x, y = rspec(), rphys()
    sx, ty = phys(), spec()
    gdhip.shtnskit_spectral_to_physical(x, sx)
    gdhip.shtnskit_spectral_to_physical_adjoint(y, ty)
    lhs, rhs = float(np.sum(sx.data * y.data)), sdot(x, ty)
    print("scalar synth", lhs, rhs)
    assert abs(lhs - rhs) <= tol * np.abs(sx.data).max() * np.abs(y.data).sum()
    # scalar analysis
    x, y = rphys(), rspec()
    ax, ty = spec(), phys()
    gdhip.shtnskit_physical_to_spectral(x, ax)
    gdhip.shtnskit_physical_to_spectral_adjoint(y, ty)
    lhs, rhs = sdot(ax, y), float(np.sum(x.data * ty.data))
    print("scalar analysis", lhs, rhs)
    assert abs(lhs - rhs) <= tol * np.abs(ty.data).max() * np.abs(x.data).sum()
    # vector synthesis
    xt, xp = rspec(), rspec()
    y = gdhip.SHTnsVectorField(cfg, nr, [phys(), rphys(), rphys()])
    sv = gdhip.SHTnsVectorField(cfg, nr)
    gdhip.shtnskit_vector_synthesis(xt, xp, sv)
    tt, tp = spec(), spec()
    gdhip.shtnskit_vector_synthesis_adjoint(y, tt, tp)
    lhs = float(np.sum(sv.θ_component.data * y.θ_component.data + sv.φ_component.data * y.φ_component.data))
    rhs = sdot(xt, tt) + sdot(xp, tp)
    print("vector synth", lhs, rhs)
    assert abs(lhs - rhs) <= tol * np.abs(sv.θ_component.data).max() * 2 * np.abs(y.θ_component.data).sum()
    # vector analysis
    x = gdhip.SHTnsVectorField(cfg, nr, [phys(), rphys(), rphys()])
    at, ap = spec(), spec()
    gdhip.shtnskit_vector_analysis(x, at, ap)
    yt, yp = rspec(), rspec()
    tv = gdhip.SHTnsVectorField(cfg, nr)
    gdhip.shtnskit_vector_analysis_adjoint(yt, yp, tv)
    lhs = sdot(at, yt) + sdot(ap, yp)
    rhs = float(np.sum(x.θ_component.data * tv.θ_component.data + x.φ_component.data * tv.φ_component.data))
    print("vector analysis", lhs, rhs)
    assert abs(lhs - rhs) <= tol * np.abs(tv.θ_component.data).max() * 2 * np.abs(x.θ_component.data).sum()


def test_batched_synthesis_equals_single_calls():
    """batch_spectral_to_physical stacks fields of one configuration into one
    launch; the result equals the one-by-one syntheses bit for bit."""
    print("case: batched synthesis", flush=True)
    import gdhip

    cfg = gdhip.create_shtnskit_config(lmax=31, nlat=48, nlon=96)
    cfg2 = gdhip.create_shtnskit_config(lmax=15, nlat=24, nlon=48)
    rng = np.random.default_rng(5)
    specs, physs, ref = [], [], []
    for c, nr in [(cfg, 3), (cfg, 5), (cfg2, 2), (cfg, 1)]:
        s = gdhip.SHTnsSpectralField(c, nr, rng.standard_normal((nr, c.nlm)), rng.standard_normal((nr, c.nlm)))
        p = gdhip.SHTnsPhysicalField(c, nr)
        q = gdhip.SHTnsPhysicalField(c, nr)
        gdhip.shtnskit_spectral_to_physical(s, q)
        specs.append(s)
        physs.append(p)
        ref.append(q.data)
    gdhip.batch_shtnskit_transforms(specs, physs)
    for p, q in zip(physs, ref):
        assert np.array_equal(p.data, q)
