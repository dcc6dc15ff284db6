#!/usr/bin/env python3
"""Write the golden fixtures of this directory (small .npz files, no pickles).

Every fixture is produced by the oracle and, where an independent path exists,
checked against it before it is written:

  seeded.npz      splitmix64 -> Box-Muller normals (seed 2024, first 64) and a
                  tiny seeded state (lmax 3, nr 4, amplitude 1, (l+1)^-2, m = 0
                  real, boundary rows zeroed): the generator that the C oracle,
                  the HIP library and the Python host share bit for bit.
  sht_l7.npz      lmax = mmax = 7 on a 12 x 24 Gauss grid: the numpy/scipy
                  (sph_harm_y) transform of a seeded scalar and of a seeded
                  (tor, pol) pair, synthesis and analysis; the C oracle must
                  agree to 1e-13 before the file is written.
  band_c2.npz     the CNAB2 operators of the lmax = 127, nr = 65 shell (radius
                  ratio 0.35, band half-width 3, dt 1e-5, theta 0.5, Dirichlet
                  rows) for the three diffusivities (Pm, 1, Pm/Pr with Pm = 5,
                  Pr = 1) at l in {0, 1, 63, 127}: L, S = I/dt - theta L, the
                  no-pivot band LU of S and the solve of a seeded right-hand
                  side; checked against a dense numpy solve of S.
  step_l7.npz     two CNAB2 steps (the AB1 bootstrap and one AB2 step) of the
                  5-field model at lmax 7, 12 x 24, nr 9, Ra 100, amplitude
                  1e-3, seed 7: states, nonlinear terms; the C oracle and the
                  numpy restatement must agree to 1e-10 first.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import np_oracle  # noqa: E402
from oracle import oracle as O  # noqa: E402

SHT_L7 = dict(lmax=7, nlat=12, nlon=24)
STEP_L7 = dict(lmax=7, nlat=12, nlon=24, nr=9)
BAND_C2 = dict(lmax=127, nr=65, bw=3, rratio=0.35, dt=1e-5, theta=0.5, Pm=5.0, Pr=1.0)
BAND_LS = (0, 1, 63, 127)


def seeded():
    normals = O.seeded_normals(2024, 64)
    st = O.seeded_state(2024, 3, 3, 4, nfields=5, amp=1.0)
    return dict(normals=normals, state=st)


def sht_l7():
    L, nlat, nlon = SHT_L7["lmax"], SHT_L7["nlat"], SHT_L7["nlon"]
    n = np_oracle.NpSHT(L, L, nlat, nlon)
    c = O.SHT(L, L, nlat, nlon)
    st = O.seeded_state(11, L, L, 1, nfields=3, amp=1.0, zero_boundaries=False)[:, 0]
    a, tor, pol = st[0], st[1].copy(), st[2].copy()
    lo, _ = O.lm_tables(L, L)
    tor[lo == 0] = 0  # the l = 0 vector harmonic vanishes
    pol[lo == 0] = 0
    f = n.scalar_synth(a)  # (nlon, nlat)
    back = n.scalar_analysis(f)
    vt, vp = n.vector_synth(tor, pol)
    t2, p2 = n.vector_analysis(vt, vp)
    assert np.abs(c.scalar_synth(a[None])[0] - f).max() < 1e-13
    assert np.abs(c.scalar_analysis(f[None])[0] - back).max() < 1e-13
    cvt, cvp = c.vector_synth(tor[None], pol[None])
    assert np.abs(cvt[0] - vt).max() < 1e-13 and np.abs(cvp[0] - vp).max() < 1e-13
    x, w, th = c.grid()
    assert np.abs(x - n.x).max() < 1e-15
    return dict(a=a, f=f, a_back=back, tor=tor, pol=pol, vt=vt, vp=vp, tor_back=t2, pol_back=p2, x=n.x, w=n.w)


def band_c2():
    p = BAND_C2
    r = O.radial_grid(p["nr"], p["rratio"])
    out = dict(r=r, ls=np.array(BAND_LS))
    rhs = O.seeded_state(5, 0, 0, p["nr"], nfields=1, amp=1.0, zero_boundaries=False)[0, :, 0].real
    out["rhs"] = rhs
    for k, nu in enumerate((p["Pm"], 1.0, p["Pm"] / p["Pr"])):
        L, S, LU, rc = O.cnab2_matrices(r, p["bw"], nu, p["dt"], p["theta"], p["lmax"], bc_rows=True)
        assert rc == 0
        sel = list(BAND_LS)
        out["L%d" % k], out["S%d" % k], out["LU%d" % k] = L[sel], S[sel], LU[sel]
        xs = []
        for j, l in enumerate(sel):
            x = O.band_solve(LU[l], rhs)
            N, bw = p["nr"], p["bw"]
            dense = np.zeros((N, N))  # band element A[i][j] = ab[j, bw + i - j]
            for i in range(N):
                for jj in range(max(0, i - bw), min(N, i + bw + 1)):
                    dense[i, jj] = S[l][jj, bw + i - jj]
            xd = np.linalg.solve(dense, rhs)
            assert np.abs(x - xd).max() < 1e-12 * np.abs(xd).max()
            xs.append(x)
        out["x%d" % k] = np.array(xs)
    return out


def step_l7():
    p = STEP_L7
    st0 = O.seeded_state(7, p["lmax"], p["lmax"], p["nr"], amp=1e-3)
    m = O.Model(p["lmax"], p["nlat"], p["nlon"], p["nr"], Ra=100.0)
    nm = np_oracle.NpModel(p["lmax"], p["nlat"], p["nlon"], p["nr"], Ra=100.0)
    nl0 = m.nonlinear(st0)
    st1, nl_a, prev1 = m.step(st0, None, 0)
    st2, nl_b, prev2 = m.step(st1, prev1, 1)
    # the numpy restatement, same two steps
    s1n, _, p1n = nm.step(st0, None, 0)
    s2n, _, _ = nm.step(s1n, p1n, 1)
    for a, b in ((st1, s1n), (st2, s2n)):
        assert max(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max() for k in range(5)) < 1e-10
    return dict(state0=st0, nl0=nl0, state1=st1, state2=st2, energies2=m.energies(st2))


def main():
    for name, fn in (("seeded", seeded), ("sht_l7", sht_l7), ("band_c2", band_c2), ("step_l7", step_l7)):
        data = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        print("%s: %d arrays, %d bytes" % (path, len(data), os.path.getsize(path)))


if __name__ == "__main__":
    main()
