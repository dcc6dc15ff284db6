"""GPU parity of the exponential integrators (EAB2, staged ERK2) with the oracle.

Reference: timestep.jl:20-65, 425-487 (ETDCache, eab2_update!) and
timestep.jl:825-1555 (ERK2Cache, ERK2FieldBuffers, erk2_prepare_field!,
erk2_apply_stage!, erk2_store_stage_nonlinear!, erk2_finalize_field!), and
the known answers of test/erk2.jl.  The device updates are compared with the
oracle's per-(l, m) products using the cache's own matrices: FP64, max
relative error <= 1e-13.  The matrices themselves are checked against the
oracle on the CPU (tests/test_oracle_etd.py).
"""
import math
import types

import numpy as np
import pytest

from oracle import np_etd as X

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _torch():
    import torch

    return torch


def _field(cfg, nr, z, device):
    import gdhip

    f = gdhip.SHTnsSpectralField(cfg, nr)
    re, im = np.ascontiguousarray(z.real), np.ascontiguousarray(z.imag)
    if device:
        t = _torch()
        re, im = t.from_numpy(re).cuda(), t.from_numpy(im).cuda()
    f.data_real, f.data_imag = re, im
    return f


def _get(f):
    re, im = f.data_real, f.data_imag
    if not isinstance(re, np.ndarray):
        _torch().cuda.synchronize()
        re, im = re.cpu().numpy(), im.cpu().numpy()
    return re + 1j * im


def _rand(rng, nr, nlm, scale=1.0):
    return scale * (rng.standard_normal((nr, nlm)) + 1j * rng.standard_normal((nr, nlm)))


@pytest.mark.parametrize("device", [False, True])
def test_erk2_reference_known_answers(device):
    """test/erk2.jl: lmax = mmax = 0, one radial level, a 1x1 cache."""
    import gdhip

    print("case: erk2 known answers device=%s" % device, flush=True)
    cfg = types.SimpleNamespace(lmax=0, mmax=0, nlm=1, device=0, l_values=np.array([0]))
    dt, lam = 0.1, 0.5
    z = -lam * dt
    eh, ef = math.exp(z / 2), math.exp(z)
    cache = gdhip.ERK2Cache(dt, [0], [np.full((1, 1), eh)], [np.full((1, 1), ef)],
                            [np.full((1, 1), (eh - 1) / (z / 2))], [np.full((1, 1), (ef - 1) / z)],
                            [np.full((1, 1), (ef - 1 - z) / z ** 2)], False, 20, 1e-8, True)
    u0, c = 0.3, 0.2
    u = _field(cfg, 1, np.array([[u0 + 0j]]), device)
    nl = _field(cfg, 1, np.array([[c + 0j]]), device)
    buf = gdhip.ERK2FieldBuffers(u, nl, cache)
    gdhip.erk2_prepare_field(buf, u, nl, cache, cfg, dt)
    gdhip.erk2_apply_stage(buf, u)
    gdhip.erk2_store_stage_nonlinear(buf, nl)
    gdhip.erk2_finalize_field(buf, u, cache, cfg, dt)
    got = _get(u)[0, 0]
    expected = math.exp(-lam * dt) * u0 + (1 - math.exp(-lam * dt)) * c / lam
    assert abs(got.real - expected) < 1e-4 and abs(got.imag) < 1e-4
    assert abs(got.real - expected) < 1e-15  # constant forcing: the scheme is exact

    u0, beta = 0.45, 0.15
    u = _field(cfg, 1, np.array([[u0 + 0j]]), device)
    nl = _field(cfg, 1, np.array([[beta * u0 + 0j]]), device)
    buf2 = gdhip.ERK2FieldBuffers(u, nl, cache)
    gdhip.erk2_prepare_field(buf2, u, nl, cache, cfg, dt)
    gdhip.erk2_apply_stage(buf2, u)
    u_stage = _get(u)[0, 0].real
    nl = _field(cfg, 1, np.array([[beta * u_stage + 0j]]), device)
    gdhip.erk2_store_stage_nonlinear(buf2, nl)
    gdhip.erk2_finalize_field(buf2, u, cache, cfg, dt)
    got = _get(u)[0, 0]
    assert abs(got.real - math.exp((beta - lam) * dt) * u0) < 1e-4 and abs(got.imag) < 1e-4
    # the oracle's restatement of the same four calls
    mats = [[m] for m in (np.full((1, 1), eh), np.full((1, 1), ef), np.full((1, 1), (eh - 1) / (z / 2)),
                          np.full((1, 1), (ef - 1) / z), np.full((1, 1), (ef - 1 - z) / z ** 2))]
    ou = np.array([[u0 + 0j]])
    lin, k1, st, ncur = X.erk2_prepare(ou, beta * ou, mats, [0], dt)
    want = X.erk2_finalize(lin, k1, ncur, beta * st, mats, [0], dt)[0, 0]
    assert abs(got - want) < 1e-15


# nr picks the launch shape: 9 (1 wave), 17 (2), 33 (3), 57 (4), 65 (5), 97 (rows split over blocks)
@pytest.mark.parametrize("lmax,mmax,nr,dt,device", [(15, 15, 33, 1e-4, True), (15, 7, 17, 1e-3, False),
                                                    (127, 127, 65, 1e-5, True), (31, 31, 9, 1e-3, True),
                                                    (31, 20, 57, 1e-4, True), (23, 23, 97, 1e-5, True),
                                                    (23, 11, 97, 1e-5, False)])
def test_eab2_update_matches_oracle(lmax, mmax, nr, dt, device):
    import gdhip

    print("case: eab2 lmax=%d mmax=%d nr=%d device=%s" % (lmax, mmax, nr, device), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, mmax, nlat=2 * (lmax // 2 + 1), nlon=3 * (lmax + 1))
    dom = gdhip.create_radial_domain(nr, 0.35)
    etd = gdhip.create_etd_cache(cfg, dom, 1.0, dt)
    E, P = np.array(etd.E), np.array(etd.phi1)
    Eh, Ph = gdhip.exp_functions(dom, 1.0, dt, lmax, "etd")
    assert np.array_equal(E, Eh) and np.array_equal(P, Ph)  # the device holds the host matrices
    rng = np.random.default_rng(lmax + nr)
    u, nl, npv = (_rand(rng, nr, cfg.nlm) for _ in range(3))
    want = X.eab2_update(u, nl, npv, E, P, cfg.l_values, dt)
    uf = _field(cfg, nr, u, device)
    gdhip.eab2_update(uf, _field(cfg, nr, nl, device), _field(cfg, nr, npv, device), etd, cfg, dt)
    assert _rel(_get(uf), want) < 1e-13


@pytest.mark.parametrize("lmax,mmax,nr,dt,nu,device", [(31, 31, 33, 1e-4, 5.0, True), (15, 5, 17, 1e-3, 1.0, False),
                                                        (127, 127, 65, 1e-5, 1.0, True), (47, 47, 9, 1e-3, 1.0, True),
                                                        (23, 23, 97, 1e-5, 1.0, True)])
def test_erk2_staged_matches_oracle(lmax, mmax, nr, dt, nu, device):
    import gdhip

    print("case: erk2 lmax=%d mmax=%d nr=%d device=%s" % (lmax, mmax, nr, device), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, mmax, nlat=2 * (lmax // 2 + 1), nlon=3 * (lmax + 1))
    dom = gdhip.create_radial_domain(nr, 0.35)
    cache = gdhip.create_erk2_cache(cfg, dom, nu, dt)
    mats = [np.array(m) for m in (cache.E_half, cache.E_full, cache.phi1_half, cache.phi1_full, cache.phi2_full)]
    host = gdhip.exp_functions(dom, nu, dt, lmax, "erk2")
    assert all(np.array_equal(a, b) for a, b in zip(mats, host))
    rng = np.random.default_rng(7 * lmax + nr)
    u0, n0, n1 = _rand(rng, nr, cfg.nlm), _rand(rng, nr, cfg.nlm, 10.0), _rand(rng, nr, cfg.nlm, 10.0)
    lin, k1, st, ncur = X.erk2_prepare(u0, n0, mats, cfg.l_values, dt)
    want = X.erk2_finalize(lin, k1, ncur, n1, mats, cfg.l_values, dt)
    u = _field(cfg, nr, u0, device)
    buf = gdhip.ERK2FieldBuffers(u, _field(cfg, nr, n0, device), cache)
    gdhip.erk2_prepare_field(buf, u, _field(cfg, nr, n0, device), cache, cfg, dt)
    gdhip.erk2_apply_stage(buf, u)
    assert _rel(_get(u), st) < 1e-13
    gdhip.erk2_store_stage_nonlinear(buf, _field(cfg, nr, n1, device))
    gdhip.erk2_finalize_field(buf, u, cache, cfg, dt)
    assert _rel(_get(u), want) < 1e-13


def test_missing_l_raises():
    """The reference errors on an l without a cache entry (timestep.jl:1357, 1475, 447-449)."""
    import gdhip

    print("case: missing l", flush=True)
    nr = 5
    eye = np.eye(nr)
    cache = gdhip.ERK2Cache(1e-3, [0, 2], [eye, eye], [eye, eye], [eye, eye], [eye, eye], [eye, eye])
    cfg = types.SimpleNamespace(lmax=2, mmax=2, nlm=6, device=0, l_values=np.array([0, 1, 1, 2, 2, 2]))
    u = _field(cfg, nr, np.ones((nr, 6), complex), False)
    buf = gdhip.ERK2FieldBuffers(u, u, cache)
    with pytest.raises(gdhip.MissingLError):
        gdhip.erk2_prepare_field(buf, u, u, cache, cfg, 1e-3)
    big = types.SimpleNamespace(lmax=3, mmax=3, nlm=10, device=0, l_values=None)
    ub = _field(big, nr, np.ones((nr, 10), complex), False)
    bufb = gdhip.ERK2FieldBuffers(ub, ub, cache)  # sizes only, as the reference's constructor
    with pytest.raises(gdhip.MissingLError):
        gdhip.erk2_prepare_field(bufb, ub, ub, cache, big, 1e-3)
    with pytest.raises(gdhip.MissingLError):
        gdhip.erk2_finalize_field(bufb, ub, cache, big, 1e-3)
    etd_cfg = types.SimpleNamespace(lmax=1, mmax=1, nlm=3, device=0)
    etd = gdhip.create_etd_cache(etd_cfg, gdhip.create_radial_domain(nr, 0.35), 1.0, 1e-3)
    with pytest.raises(gdhip.MissingLError):
        gdhip.eab2_update(u, u, u, etd, cfg, 1e-3)
    # a cache with every l present still works after the error
    cache2 = gdhip.ERK2Cache(1e-3, [0, 1, 2], [eye] * 3, [2 * eye] * 3, [eye] * 3, [eye] * 3, [eye] * 3)
    buf2 = gdhip.ERK2FieldBuffers(u, u, cache2)
    gdhip.erk2_prepare_field(buf2, u, u, cache2, cfg, 1e-3)
    gdhip.erk2_apply_stage(buf2, u)
    assert np.allclose(_get(u), 1.0 + 0.5e-3)  # E_half u + dt/2 phi1_half nl
