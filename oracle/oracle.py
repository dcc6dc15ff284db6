"""ctypes binding of the FP64 CPU oracle (oracle/gd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.  The oracle restates GeoDynamo.jl's CNAB2 step; each C
function cites the reference file:line it follows.

Array conventions (the reference's boundary layouts):
  spectral   (nr, nlm) C-order == Julia (nlm, 1, nr) column-major, l-major lm
  physical   (nr, nlon, nlat) C-order == Julia (nlat, nlon, nr) column-major
  band       (N, 2bw+1) C-order == Julia (2bw+1, N) column-major
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libgdoracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with the Makefile beside this file."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_double, c_long = ctypes.c_int, ctypes.c_double, ctypes.c_long
        vp = ctypes.c_void_p
        sig = {
            "gdo_radial_grid": (None, [c_int, c_double, _dp]),
            "gdo_fd_weights": (c_int, [c_int, c_int, c_int, _dp, c_int, _dp]),
            "gdo_radial_laplacian": (None, [c_int, c_int, _dp, c_int, _dp]),
            "gdo_band_factor": (c_int, [c_int, c_int, _dp]),
            "gdo_band_solve": (None, [c_int, c_int, _dp, _dp]),
            "gdo_band_apply": (None, [c_int, c_int, _dp, _dp, _dp]),
            "gdo_cnab2_matrices": (c_int, [c_int, c_int, _dp, c_double, c_double, c_double,
                                           c_int, c_int, c_int, _dp, _dp, _dp]),
            "gdo_cnab2_rhs": (None, [c_int, c_int, c_int, _ip, _dp, c_double, c_double, c_int,
                                     _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "gdo_cnab2_solve": (None, [c_int, c_int, c_int, _ip, _dp, _dp, _dp, _dp, _dp]),
            "gdo_nlm": (c_int, [c_int, c_int]),
            "gdo_lm_tables": (None, [c_int, c_int, _ip, _ip]),
            "gdo_gauss": (None, [c_int, _dp, _dp]),
            "gdo_plm": (None, [c_int, c_int, c_int, _dp, _dp, _dp]),
            "gdo_sht_create": (vp, [c_int, c_int, c_int, c_int]),
            "gdo_sht_destroy": (None, [vp]),
            "gdo_sht_grid": (None, [vp, _dp, _dp, _dp]),
            "gdo_scalar_synth": (None, [vp, _dp, _dp, c_int, _dp]),
            "gdo_scalar_analysis": (None, [vp, _dp, c_int, _dp, _dp]),
            "gdo_vector_synth": (None, [vp, _dp, _dp, _dp, _dp, c_int, _dp, _dp]),
            "gdo_vector_analysis": (None, [vp, _dp, _dp, c_int, _dp, _dp, _dp, _dp]),
            "gdo_seeded_normals": (None, [ctypes.c_ulonglong, c_long, _dp]),
            "gdo_seeded_state": (None, [ctypes.c_ulonglong, c_int, c_int, c_int, c_int,
                                        c_double, c_double, c_int, _dp, _dp]),
            "gdo_model_create": (vp, [vp]),
            "gdo_model_destroy": (None, [vp]),
            "gdo_nonlinear": (None, [vp, _dp, _dp, _dp, _dp]),
            "gdo_step": (None, [vp, c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
            "gdo_energies": (None, [vp, _dp, _dp, _dp]),
            "gdo_timestep_error": (c_double, [c_long, _dp, _dp, _dp, _dp]),
            "gdo_physical_level": (None, [vp, _dp, _dp, c_int, _dp, _dp]),
            "gdo_set_threads": (c_int, [c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    """OpenMP threads of the oracle's loops (n > 0 sets them); returns the count."""
    return lib().gdo_set_threads(int(n))


# ---------------------------------------------------------------- radial
def radial_grid(N, rratio):
    r = np.zeros((7, N))
    lib().gdo_radial_grid(N, rratio, r)
    return r  # r[p, n] == Julia r[n, p+1]


def fd_weights(order, r, bw, transpose=True):
    N = r.shape[1]
    band = np.zeros((N, 2 * bw + 1))
    rc = lib().gdo_fd_weights(order, N, bw, _c(r), int(transpose), band)
    if rc:
        raise RuntimeError("fd weights failed")
    return band


def radial_laplacian(r, bw, transpose=True):
    N = r.shape[1]
    lap = np.zeros((N, 2 * bw + 1))
    lib().gdo_radial_laplacian(N, bw, _c(r), int(transpose), lap)
    return lap


def band_factor(ab):
    lu = _c(ab).copy()
    N, w = lu.shape
    rc = lib().gdo_band_factor(N, (w - 1) // 2, lu)
    return lu, rc


def band_solve(lu, b):
    x = _c(b).copy()
    N, w = lu.shape
    lib().gdo_band_solve(N, (w - 1) // 2, _c(lu), x)
    return x


def band_apply(ab, v):
    N, w = ab.shape
    out = np.zeros(N)
    lib().gdo_band_apply(N, (w - 1) // 2, _c(ab), _c(v), out)
    return out


def cnab2_matrices(r, bw, nu, dt, theta, lmax, fd_transpose=True, bc_rows=False):
    N = r.shape[1]
    shape = (lmax + 1, N, 2 * bw + 1)
    L, S, LU = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    rc = lib().gdo_cnab2_matrices(N, bw, _c(r), nu, dt, theta, lmax, int(fd_transpose), int(bc_rows), L, S, LU)
    return L, S, LU, rc


def cnab2_rhs(l_of_lm, Lmat, dt, theta, u, nl, nlp, bc_rows=False):
    """u, nl, nlp: complex (nr, nlm)."""
    nr, nlm = u.shape
    bw = (Lmat.shape[2] - 1) // 2
    rr, ri = np.zeros((nr, nlm)), np.zeros((nr, nlm))
    lib().gdo_cnab2_rhs(nlm, nr, bw, np.ascontiguousarray(l_of_lm, np.int32), _c(Lmat), dt, theta, int(bc_rows),
                        _c(u.real), _c(u.imag), _c(nl.real), _c(nl.imag), _c(nlp.real), _c(nlp.imag), rr, ri)
    return rr + 1j * ri


def cnab2_solve(l_of_lm, LU, rhs):
    nr, nlm = rhs.shape
    bw = (LU.shape[2] - 1) // 2
    sr, si = np.zeros((nr, nlm)), np.zeros((nr, nlm))
    lib().gdo_cnab2_solve(nlm, nr, bw, np.ascontiguousarray(l_of_lm, np.int32), _c(LU),
                          _c(rhs.real), _c(rhs.imag), sr, si)
    return sr + 1j * si


# ---------------------------------------------------------------- SHT
def nlm(lmax, mmax):
    return lib().gdo_nlm(lmax, mmax)


def lm_tables(lmax, mmax):
    n = nlm(lmax, mmax)
    lo, mo = np.zeros(n, np.int32), np.zeros(n, np.int32)
    lib().gdo_lm_tables(lmax, mmax, lo, mo)
    return lo, mo


def gauss(nlat):
    x, w = np.zeros(nlat), np.zeros(nlat)
    lib().gdo_gauss(nlat, x, w)
    return x, w


def plm(lmax, mmax, x):
    n = nlm(lmax, mmax)
    x = _c(x)
    P, dP = np.zeros((n, x.size)), np.zeros((n, x.size))
    lib().gdo_plm(lmax, mmax, x.size, x, P, dP)
    return P, dP


class SHT:
    """Oracle transform plan (shtnskit_transforms.jl:102-200 semantics)."""

    def __init__(self, lmax, mmax, nlat, nlon):
        self.lmax, self.mmax, self.nlat, self.nlon = lmax, mmax, nlat, nlon
        self.nlm = nlm(lmax, mmax)
        self._h = lib().gdo_sht_create(lmax, mmax, nlat, nlon)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gdo_sht_destroy(self._h)
            self._h = None

    def grid(self):
        x, w, th = np.zeros(self.nlat), np.zeros(self.nlat), np.zeros(self.nlat)
        lib().gdo_sht_grid(self._h, x, w, th)
        return x, w, th

    def scalar_synth(self, a):
        a = np.atleast_2d(a)
        nr = a.shape[0]
        out = np.zeros((nr, self.nlon, self.nlat))
        lib().gdo_scalar_synth(self._h, _c(a.real), _c(a.imag), nr, out)
        return out

    def scalar_analysis(self, f):
        f = _c(f).reshape(-1, self.nlon, self.nlat)
        nr = f.shape[0]
        re, im = np.zeros((nr, self.nlm)), np.zeros((nr, self.nlm))
        lib().gdo_scalar_analysis(self._h, f, nr, re, im)
        return re + 1j * im

    def vector_synth(self, tor, pol):
        tor, pol = np.atleast_2d(tor), np.atleast_2d(pol)
        nr = tor.shape[0]
        vt, vp = np.zeros((nr, self.nlon, self.nlat)), np.zeros((nr, self.nlon, self.nlat))
        lib().gdo_vector_synth(self._h, _c(tor.real), _c(tor.imag), _c(pol.real), _c(pol.imag), nr, vt, vp)
        return vt, vp

    def vector_analysis(self, vt, vp):
        vt = _c(vt).reshape(-1, self.nlon, self.nlat)
        vp = _c(vp).reshape(-1, self.nlon, self.nlat)
        nr = vt.shape[0]
        a = [np.zeros((nr, self.nlm)) for _ in range(4)]
        lib().gdo_vector_analysis(self._h, vt, vp, nr, *a)
        return a[0] + 1j * a[1], a[2] + 1j * a[3]  # (tor, pol)


# ---------------------------------------------------------------- state
def seeded_normals(seed, n):
    out = np.zeros(n)
    lib().gdo_seeded_normals(seed, n, out)
    return out


def seeded_state(seed, lmax, mmax, nr, nfields=5, amp=1.0, power=2.0, zero_boundaries=True):
    n = nlm(lmax, mmax)
    re, im = np.zeros((nfields, nr, n)), np.zeros((nfields, nr, n))
    lib().gdo_seeded_state(seed, lmax, mmax, nr, nfields, amp, power, int(zero_boundaries), re, im)
    return re + 1j * im


# ---------------------------------------------------------------- model
class _Params(ctypes.Structure):
    _fields_ = [("E", ctypes.c_double), ("Pm", ctypes.c_double), ("Pr", ctypes.c_double),
                ("Ra", ctypes.c_double), ("rratio", ctypes.c_double), ("dt", ctypes.c_double),
                ("theta", ctypes.c_double),
                ("lmax", ctypes.c_int), ("mmax", ctypes.c_int), ("nlat", ctypes.c_int),
                ("nlon", ctypes.c_int), ("nr", ctypes.c_int), ("bw", ctypes.c_int),
                ("fd_transpose", ctypes.c_int), ("radials", ctypes.c_int),
                ("coriolis_index", ctypes.c_int), ("bc_rows", ctypes.c_int), ("radial_force", ctypes.c_int)]


class Model:
    """Oracle CNAB2 model: field order u_tor, u_pol, B_tor, B_pol, T."""

    def __init__(self, lmax, nlat, nlon, nr, E=1e-3, Pm=5.0, Pr=1.0, Ra=100.0, rratio=0.35,
                 dt=1e-5, theta=0.5, bw=3, mmax=None, fd_transpose=True, radials=False,
                 coriolis_index=0, bc_rows=1, radial_force=0):
        mmax = lmax if mmax is None else mmax
        self.p = _Params(E, Pm, Pr, Ra, rratio, dt, theta, lmax, mmax, nlat, nlon, nr, bw,
                         int(fd_transpose), int(radials), int(coriolis_index), int(bc_rows), int(radial_force))
        self.lmax, self.mmax, self.nlat, self.nlon, self.nr = lmax, mmax, nlat, nlon, nr
        self.nlm = nlm(lmax, mmax)
        self._h = lib().gdo_model_create(ctypes.addressof(self.p))
        self.step_index = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gdo_model_destroy(self._h)
            self._h = None

    def nonlinear(self, st):
        sr, si = _c(st.real), _c(st.imag)
        nr_, ni_ = np.zeros_like(sr), np.zeros_like(si)
        lib().gdo_nonlinear(self._h, sr, si, nr_, ni_)
        return nr_ + 1j * ni_

    def step(self, st, nl_prev=None, step_index=None):
        """One CNAB2 step; returns (new_state, nl, nl_prev_after_roll)."""
        k = self.step_index if step_index is None else step_index
        sr, si = _c(st.real).copy(), _c(st.imag).copy()
        nlr, nli = np.zeros_like(sr), np.zeros_like(si)
        if nl_prev is None:
            npr, npi = np.zeros_like(sr), np.zeros_like(si)
        else:
            npr, npi = _c(nl_prev.real).copy(), _c(nl_prev.imag).copy()
        lib().gdo_step(self._h, k, sr, si, nlr, nli, npr, npi)
        self.step_index = k + 1
        return sr + 1j * si, nlr + 1j * nli, npr + 1j * npi

    def energies(self, st):
        out = np.zeros(3)
        lib().gdo_energies(self._h, _c(st.real), _c(st.imag), out)
        return out

    def physical_level(self, st, r):
        ng = self.nlat * self.nlon
        F, P = np.zeros(15 * ng), np.zeros(5 * ng)
        lib().gdo_physical_level(self._h, _c(st.real), _c(st.imag), r, F, P)
        return F.reshape(15, self.nlon, self.nlat), P.reshape(5, self.nlon, self.nlat)
