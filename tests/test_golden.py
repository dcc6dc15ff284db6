"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the oracle, the numpy restatement and
the host side of the C ABI (CPU only).

seeded.npz    splitmix64/Box-Muller generator: bit-exact in the oracle and in
              libgdhip's host gd_seeded_state.
sht_l7.npz    the scipy sph_harm_y transform (shtnskit_field_functions.jl:11-321
              semantics) at lmax 7, 12 x 24; the C oracle agrees to 1e-13 and
              both round trips close.
band_c2.npz   CNAB2 operators of the lmax 127, nr 65 shell (timestep.jl:498-543,
              linear_algebra.jl:23-88) at l in {0, 1, 63, 127}; the stored
              solve equals a dense numpy solve of the stored S.
step_l7.npz   two 5-field CNAB2 steps (simulation.jl:1151-1330 as the oracle
              restates it) at lmax 7, nr 9; the numpy model agrees to 1e-10.
"""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_seeded_generator_fixture(oracle):
    import gdhip

    g = _load("seeded")
    assert np.array_equal(oracle.seeded_normals(2024, 64), g["normals"])
    assert np.array_equal(oracle.seeded_state(2024, 3, 3, 4, nfields=5, amp=1.0), g["state"])
    assert np.array_equal(gdhip.seeded_state(2024, 3, 4, nfields=5, amp=1.0), g["state"])
    st = g["state"]
    assert np.all(st[:, 0] == 0) and np.all(st[:, -1] == 0)  # boundary rows zeroed
    lo, mo = oracle.lm_tables(3, 3)
    assert np.all(st[..., mo == 0].imag == 0)  # m = 0 real


def test_sht_fixture(oracle):
    from oracle import np_oracle

    g = _load("sht_l7")
    L, nlat, nlon = 7, 12, 24
    x, w = np.polynomial.legendre.leggauss(nlat)
    o = np.argsort(-x)
    assert np.abs(g["x"] - x[o]).max() < 1e-15 and np.abs(g["w"] - w[o]).max() < 5e-14
    c = oracle.SHT(L, L, nlat, nlon)
    assert np.abs(c.scalar_synth(g["a"][None])[0] - g["f"]).max() < 1e-13
    assert np.abs(c.scalar_analysis(g["f"][None])[0] - g["a_back"]).max() < 1e-13
    vt, vp = c.vector_synth(g["tor"][None], g["pol"][None])
    assert np.abs(vt[0] - g["vt"]).max() < 1e-13 and np.abs(vp[0] - g["vp"]).max() < 1e-13
    t, p = c.vector_analysis(g["vt"][None], g["vp"][None])
    assert np.abs(t[0] - g["tor_back"]).max() < 1e-13 and np.abs(p[0] - g["pol_back"]).max() < 1e-13
    # round trips close (test/shtnskit_roundtrip.jl:54 accepts 1e-7 on squared norms)
    assert np.abs(g["a_back"] - g["a"]).max() < 1e-13
    assert np.abs(g["tor_back"] - g["tor"]).max() < 1e-13 and np.abs(g["pol_back"] - g["pol"]).max() < 1e-13
    n = np_oracle.NpSHT(L, L, nlat, nlon)
    assert np.abs(n.scalar_synth(g["a"]) - g["f"]).max() < 1e-14


def _dense(ab):
    N, W = ab.shape
    bw = (W - 1) // 2
    A = np.zeros((N, N))
    for j in range(N):
        for i in range(max(0, j - bw), min(N, j + bw + 1)):
            A[i, j] = ab[j, bw + i - j]
    return A


def test_band_fixture(oracle):
    g = _load("band_c2")
    r = oracle.radial_grid(65, 0.35)
    assert np.array_equal(r, g["r"])
    ls = list(g["ls"])
    for k, nu in enumerate((5.0, 1.0, 5.0)):
        L, S, LU, rc = oracle.cnab2_matrices(r, 3, nu, 1e-5, 0.5, 127, bc_rows=True)
        assert rc == 0
        assert _rel(L[ls], g["L%d" % k]) < 1e-14
        assert _rel(S[ls], g["S%d" % k]) < 1e-14
        assert _rel(LU[ls], g["LU%d" % k]) < 1e-14
        for j in range(len(ls)):
            A = _dense(g["S%d" % k][j])
            assert np.all(A[0, 1:] == 0) and A[0, 0] == 1.0  # Dirichlet row
            xd = np.linalg.solve(A, g["rhs"])
            assert _rel(g["x%d" % k][j], xd) < 1e-12


def test_step_fixture(oracle):
    from oracle import np_oracle

    g = _load("step_l7")
    m = oracle.Model(7, 12, 24, 9, Ra=100.0)
    s1, _, prev = m.step(g["state0"], None, 0)
    s2, _, _ = m.step(s1, prev, 1)
    for k in range(5):
        assert _rel(s1[k], g["state1"][k]) < 1e-13
        assert _rel(s2[k], g["state2"][k]) < 1e-13
    assert _rel(m.nonlinear(g["state0"]), g["nl0"]) < 1e-13
    assert _rel(m.energies(g["state2"]), g["energies2"]) < 1e-14
    nm = np_oracle.NpModel(7, 12, 24, 9, Ra=100.0)
    n1, _, p1 = nm.step(g["state0"], None, 0)
    n2, _, _ = nm.step(n1, p1, 1)
    assert max(_rel(n2[k], g["state2"][k]) for k in range(5)) < 1e-10
