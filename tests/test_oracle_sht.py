"""Pins of the oracle's spherical-harmonic transform (CPU).

The transform behind shtnskit_field_functions.jl (SHTnsKit.jl 1.1.10, absent
here) is restated as orthonormal Y_lm with Condon-Shortley phase on a Gauss
grid.  Pins: numpy leggauss, scipy sph_harm_y, finite differences,
orthonormality, Parseval, single-mode deltas, the reference's own round-trip
test sizes (test/shtnskit_roundtrip.jl:21-54) and vector closed forms.  The
exact SHTnsKit phase/normalisation is parity unpinned.
"""
import numpy as np
import pytest
from scipy.special import sph_harm_y


def test_gauss_nodes_and_weights(oracle):
    for n in (12, 48, 192):
        x, w = oracle.gauss(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        o = np.argsort(-xr)
        assert np.abs(x - xr[o]).max() < 1e-15
        assert np.abs(w - wr[o]).max() < 5e-14  # numpy's own weights carry ~3e-14
        assert abs(w.sum() - 2.0) < 1e-14
        # exactness: sum_i w_i P_l(x_i) = 2 delta_l0 for every l < 2n
        for l in range(1, 2 * n, 5):
            c = np.zeros(l + 1)
            c[l] = 1
            assert abs(np.sum(w * np.polynomial.legendre.legval(x, c))) < 2e-15


def test_plm_against_scipy_and_derivative(oracle):
    lmax, nlat = 20, 30
    x, _ = oracle.gauss(nlat)
    th = np.arccos(x)
    P, dP = oracle.plm(lmax, lmax, x)
    lo, mo = oracle.lm_tables(lmax, lmax)
    h = 1e-6
    for q, (l, m) in enumerate(zip(lo, mo)):
        ref = sph_harm_y(l, m, th, 0.0).real
        assert np.abs(P[q] - ref).max() < 1e-13
        fd = (sph_harm_y(l, m, th + h, 0.0).real - sph_harm_y(l, m, th - h, 0.0).real) / (2 * h)
        assert np.abs(dP[q] - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())


@pytest.mark.parametrize("m", [0, 1, 100, 200, 255])
def test_plm_orthonormal_at_high_degree(oracle, m):
    lmax, nlat = 255, 384
    x, w = oracle.gauss(nlat)
    P, _ = oracle.plm(lmax, lmax, x)
    lo, mo = oracle.lm_tables(lmax, lmax)
    rows = P[mo == m]
    G = 2 * np.pi * (rows * w) @ rows.T
    assert np.abs(G - np.eye(G.shape[0])).max() < 1e-12


def _real_field_of_mode(l, m, th, phi, scale=1.0 + 0.5j):
    Y = sph_harm_y(l, m, th[None, :], phi[:, None])
    f = (scale * Y).real * (1 if m == 0 else 2)
    return f  # (nlon, nlat)


def test_single_mode_synthesis_and_delta_analysis(oracle):
    lmax, nlat, nlon = 7, 12, 24
    sht = oracle.SHT(lmax, lmax, nlat, nlon)
    x, w, th = sht.grid()
    phi = 2 * np.pi * np.arange(nlon) / nlon
    lo, mo = oracle.lm_tables(lmax, lmax)
    for q in (0, 3, 10, len(lo) - 1):
        a = np.zeros(len(lo), complex)
        a[q] = 1.0 + (0.5j if mo[q] else 0.0)
        f = sht.scalar_synth(a)[0]
        ref = _real_field_of_mode(lo[q], mo[q], th, phi, a[q])
        assert np.abs(f - ref).max() < 1e-13
        back = sht.scalar_analysis(f)[0]
        assert np.abs(back - a).max() < 1e-13


def test_reference_roundtrip_sizes(oracle):
    """test/shtnskit_roundtrip.jl: lmax=6, nlat=12, nlon=24, nr=6; the
    reference accepts sum|e|^2/sum|x|^2 < 1e-7."""
    lmax, nlat, nlon, nr = 6, 12, 24, 6
    sht = oracle.SHT(lmax, lmax, nlat, nlon)
    a = oracle.seeded_state(1234, lmax, lmax, nr, nfields=1, zero_boundaries=False)[0]
    back = sht.scalar_analysis(sht.scalar_synth(a))
    ratio = np.sum(np.abs(back - a) ** 2) / np.sum(np.abs(a) ** 2)
    assert ratio < 1e-28


def test_parseval(oracle):
    lmax, nlat, nlon = 15, 24, 48
    sht = oracle.SHT(lmax, lmax, nlat, nlon)
    x, w, th = sht.grid()
    a = oracle.seeded_state(5, lmax, lmax, 1, nfields=1, zero_boundaries=False)[0, 0]
    f = sht.scalar_synth(a)[0]
    integral = (2 * np.pi / nlon) * np.sum(w[None, :] * f ** 2)
    _, mo = oracle.lm_tables(lmax, lmax)
    spec = np.sum(np.abs(a[mo == 0]) ** 2) + 2 * np.sum(np.abs(a[mo > 0]) ** 2)
    assert abs(integral - spec) < 1e-12 * spec


def test_vector_closed_forms(oracle):
    """S = Y_lm: v_theta = dY/dtheta, v_phi = (i m / sin) Y;
    T = Y_lm: v_theta = (i m / sin) Y, v_phi = -dY/dtheta (real-field convention)."""
    lmax, nlat, nlon = 6, 12, 24
    sht = oracle.SHT(lmax, lmax, nlat, nlon)
    x, w, th = sht.grid()
    phi = 2 * np.pi * np.arange(nlon) / nlon
    lo, mo = oracle.lm_tables(lmax, lmax)
    h = 1e-6
    for q in (1, 4, 9, len(lo) - 2):
        l, m = lo[q], mo[q]
        c = 1.0 + (0.5j if m else 0.0)
        Y = c * sph_harm_y(l, m, th[None, :], phi[:, None])
        dY = c * (sph_harm_y(l, m, th[None, :] + h, phi[:, None]) - sph_harm_y(l, m, th[None, :] - h, phi[:, None])) / (2 * h)
        k = 1 if m == 0 else 2
        a = np.zeros(len(lo), complex)
        a[q] = c
        z = np.zeros_like(a)
        vt, vp = sht.vector_synth(z, a)  # S = pol
        assert np.abs(vt[0] - k * dY.real).max() < 1e-7
        assert np.abs(vp[0] - k * (1j * m / np.sin(th)[None, :] * Y).real).max() < 1e-12
        vt, vp = sht.vector_synth(a, z)  # T = tor
        assert np.abs(vt[0] - k * (1j * m / np.sin(th)[None, :] * Y).real).max() < 1e-12
        assert np.abs(vp[0] + k * dY.real).max() < 1e-7


def test_vector_roundtrip(oracle):
    lmax, nlat, nlon = 12, 20, 48
    sht = oracle.SHT(lmax, lmax, nlat, nlon)
    st = oracle.seeded_state(9, lmax, lmax, 3, nfields=2, zero_boundaries=False)
    lo, _ = oracle.lm_tables(lmax, lmax)
    st[:, :, lo == 0] = 0
    vt, vp = sht.vector_synth(st[0], st[1])
    tor, pol = sht.vector_analysis(vt, vp)
    assert np.abs(tor - st[0]).max() < 1e-13
    assert np.abs(pol - st[1]).max() < 1e-13


def test_c_oracle_equals_numpy_restatement(oracle):
    from oracle import np_oracle

    lmax, nlat, nlon = 7, 12, 24
    c = oracle.SHT(lmax, lmax, nlat, nlon)
    n = np_oracle.NpSHT(lmax, lmax, nlat, nlon)
    P, dP = oracle.plm(lmax, lmax, c.grid()[0])
    assert np.abs(P - n.P).max() < 1e-14 and np.abs(dP - n.dP).max() < 1e-13
    st = oracle.seeded_state(4, lmax, lmax, 1, nfields=2, zero_boundaries=False)
    f = c.scalar_synth(st[0])[0]
    assert np.abs(f - n.scalar_synth(st[0, 0])).max() < 1e-13
    assert np.abs(c.scalar_analysis(f)[0] - n.scalar_analysis(f)).max() < 1e-13
    vt, vp = c.vector_synth(st[0], st[1])
    nt, npp = n.vector_synth(st[0, 0], st[1, 0])
    assert np.abs(vt[0] - nt).max() < 1e-12 and np.abs(vp[0] - npp).max() < 1e-12
