"""CPU checks of the C ABI: libgdhip.so loads, exports every symbol that
include/gdhip.h declares, and its host-side FP64 set-up (radial grid, FD
weights, Laplacian) equals the oracle's (fields.jl:108-131,
linear_algebra.jl:90-151).  No GPU call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gdhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from gdhip import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_version_string():
    from gdhip import _lib

    assert b"gfx950" in _lib.lib().gd_version()


@pytest.mark.parametrize("N", [9, 33, 65, 97])
def test_radial_setup_matches_oracle(oracle, N):
    import gdhip

    dom = gdhip.create_radial_domain(N, 0.35)
    r = oracle.radial_grid(N, 0.35)
    assert np.allclose(dom.r, r, rtol=2e-16, atol=0)
    for order in (1, 2):
        for literal in (False, True):
            g = gdhip.create_derivative_matrix(order, dom, bw=3, literal=literal).data
            o = oracle.fd_weights(order, r, 3, transpose=not literal)
            assert np.allclose(g, o, rtol=1e-12, atol=1e-12 * np.abs(o).max())
    lap = gdhip.create_radial_laplacian(dom, bw=3).data
    assert np.allclose(lap, oracle.radial_laplacian(r, 3), rtol=1e-12, atol=1e-9)


def test_bad_arguments_are_reported():
    from gdhip import _lib

    L = _lib.lib()
    assert L.gd_radial_grid(1, 0.35, None) == _lib.GD_ERR_ARG
    assert b"gd_radial_grid" in L.gd_last_error()


def test_seeded_state_matches_oracle_generator(oracle):
    import gdhip

    a = gdhip.seeded_state(2024, 15, 9, nfields=5, amp=1e-3)
    b = oracle.seeded_state(2024, 15, 15, 9, nfields=5, amp=1e-3)
    # same draws; host libm of hipcc vs gcc may differ in the last ulp
    assert np.allclose(a, b, rtol=4e-16, atol=0)
    assert np.all(a[:, [0, -1]] == 0)
    lo, mo = oracle.lm_tables(15, 15)
    assert np.all(a.imag[..., mo == 0] == 0)
