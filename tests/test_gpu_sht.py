"""GPU parity of the spherical-harmonic transforms against the oracle.

Reference entry points: shtnskit_field_functions.jl:11-21 (scalar synthesis),
:131-141 (scalar analysis), :252-282 (vector synthesis), :291-321 (vector
analysis); P_lm tables shtnskit_transforms.jl:120-134.
Bars: FP64 max relative error <= 1e-12 (tables 1e-13); FP32 <= 2e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# BF16x3: each product carries ~2^-16 relative error (hi/lo bf16 halves,
# lo*lo dropped), so the transforms sit near 1e-5 relative
TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _coeffs(oracle, lmax, nr, seed, power=1.5, nfields=2):
    st = oracle.seeded_state(seed, lmax, lmax, nr, nfields=nfields, amp=1.0, power=power, zero_boundaries=False)
    lo, mo = oracle.lm_tables(lmax, lmax)
    st[:, :, lo == 0] = 0.0  # vector fields have no l=0 part
    return st


@pytest.mark.parametrize("lmax,nlat", [(31, 48), (127, 192), (255, 384)])
def test_plm_tables_match_oracle(oracle, lmax, nlat):
    import gdhip

    print("case: plm lmax=%d nlat=%d" % (lmax, nlat), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, nlat=nlat, nlon=3 * (lmax + 1))
    P, dP = cfg.plm_tables()
    x, w = oracle.gauss(nlat)
    assert np.abs(cfg.cos_theta - x).max() < 1e-15
    assert np.abs(cfg.gauss_weights - w).max() < 1e-15
    Po, dPo = oracle.plm(lmax, lmax, x)
    # recurrence round-off grows with l: 2e-14 * (lmax+1) relative to max|P|
    assert np.abs(P - Po).max() < 2e-14 * (lmax + 1) * np.abs(Po).max()
    assert np.abs(dP - dPo).max() < 1e-10 * max(1.0, np.abs(dPo).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("lmax,nlat,nlon,nr", [(6, 12, 24, 6), (31, 48, 96, 5), (127, 192, 384, 3)])
def test_scalar_transforms_match_oracle(oracle, prec, lmax, nlat, nlon, nr):
    import gdhip

    print("case: scalar sht %s lmax=%d" % (prec, lmax), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, nlat=nlat, nlon=nlon, precision=prec)
    osht = oracle.SHT(lmax, lmax, nlat, nlon)
    a = _coeffs(oracle, lmax, nr, 21, nfields=1)[0]
    spec = gdhip.SHTnsSpectralField(cfg, nr, np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag))
    phys = gdhip.SHTnsPhysicalField(cfg, nr)
    gdhip.shtnskit_spectral_to_physical(spec, phys)
    ref = osht.scalar_synth(a)
    assert _rel(phys.data, ref) < TOL[prec]
    back = gdhip.SHTnsSpectralField(cfg, nr)
    gdhip.shtnskit_physical_to_spectral(gdhip.SHTnsPhysicalField(cfg, nr, ref), back)
    aref = osht.scalar_analysis(ref)
    assert _rel(back.complex(), aref) < TOL[prec]
    assert np.all(back.data_imag[:, cfg.m_values == 0] == 0.0)
    # round trip (test/shtnskit_roundtrip.jl:54 pins sum|e|^2/sum|x|^2 < 1e-7)
    assert _rel(back.complex(), a) < TOL[prec] * 10


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("lmax,nlat,nlon,nr", [(7, 12, 24, 4), (31, 48, 96, 5), (127, 192, 384, 3)])
def test_vector_transforms_match_oracle(oracle, prec, lmax, nlat, nlon, nr):
    import gdhip

    print("case: vector sht %s lmax=%d" % (prec, lmax), flush=True)
    cfg = gdhip.create_shtnskit_config(lmax, nlat=nlat, nlon=nlon, precision=prec)
    osht = oracle.SHT(lmax, lmax, nlat, nlon)
    tp = _coeffs(oracle, lmax, nr, 33)
    tor, pol = tp[0], tp[1]
    ts = gdhip.SHTnsSpectralField(cfg, nr, np.ascontiguousarray(tor.real), np.ascontiguousarray(tor.imag))
    ps = gdhip.SHTnsSpectralField(cfg, nr, np.ascontiguousarray(pol.real), np.ascontiguousarray(pol.imag))
    vec = gdhip.SHTnsVectorField(cfg, nr)
    gdhip.shtnskit_vector_synthesis(ts, ps, vec)
    vt, vp = osht.vector_synth(tor, pol)
    assert _rel(vec.θ_component.data, vt) < TOL[prec]
    assert _rel(vec.φ_component.data, vp) < TOL[prec]
    vin = gdhip.SHTnsVectorField(cfg, nr)
    vin.θ_component.data[...] = vt
    vin.φ_component.data[...] = vp
    t2 = gdhip.SHTnsSpectralField(cfg, nr)
    p2 = gdhip.SHTnsSpectralField(cfg, nr)
    gdhip.shtnskit_vector_analysis(vin, t2, p2)
    tref, pref = osht.vector_analysis(vt, vp)
    assert _rel(t2.complex(), tref) < TOL[prec]
    assert _rel(p2.complex(), pref) < TOL[prec]
    assert _rel(t2.complex(), tor) < TOL[prec] * 10
    assert _rel(p2.complex(), pol) < TOL[prec] * 10


def test_device_tensor_path_matches_host_path(oracle):
    import torch

    import gdhip

    lmax, nlat, nlon, nr = 31, 48, 96, 4
    cfg = gdhip.create_shtnskit_config(lmax, nlat=nlat, nlon=nlon)
    a = _coeffs(oracle, lmax, nr, 5, nfields=1)[0]
    host = gdhip.SHTnsPhysicalField(cfg, nr)
    gdhip.shtnskit_spectral_to_physical(
        gdhip.SHTnsSpectralField(cfg, nr, np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)), host)
    re = torch.tensor(a.real, dtype=torch.float64, device="cuda")
    im = torch.tensor(a.imag, dtype=torch.float64, device="cuda")
    out = torch.zeros((nr, nlon, nlat), dtype=torch.float64, device="cuda")
    gdhip.shtnskit_spectral_to_physical(gdhip.SHTnsSpectralField(cfg, nr, re, im),
                                        gdhip.SHTnsPhysicalField(cfg, nr, out))
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), host.data)
