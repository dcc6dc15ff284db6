"""Scalar boundary conditions on the GPU (gd_scalar_bc_*, through the C ABI)
against the oracle (oracle/np_bc.py: thermal.jl:262-313,
compositional.jl:318-370, scalar_field_common.jl:866-929, 1166-1199), FP64.
Bar: max |GPU - oracle| <= 1e-12 x (1 + max |profile|) per plane; the tau
tables to 1e-13 relative."""
import numpy as np
import pytest

from oracle import np_bc, np_oracle

pytestmark = pytest.mark.gpu


def _case(seed, lmax, nr, mmax=None):
    import gdhip

    rng = np.random.default_rng(seed)
    cfg = gdhip.create_shtnskit_config(lmax, mmax=mmax, nlat=2 * (lmax + 1), nlon=4 * (lmax + 1))
    dom = gdhip.create_radial_domain(nr, 0.35)
    nlm = cfg.nlm
    re = rng.standard_normal((nr, nlm))
    im = rng.standard_normal((nr, nlm))
    im[:, rng.random(nlm) < 0.3] = 0.0            # m = 0-like columns: no imaginary correction
    tiny = rng.random(nlm) < 0.1
    im[:, tiny] = 5e-13                           # below the 1e-12 threshold
    ti = rng.integers(0, 3, nlm).astype(np.int32)  # 0: no condition, 1: Dirichlet, 2: flux
    to = rng.integers(0, 3, nlm).astype(np.int32)
    vals = rng.standard_normal((2, nlm))
    return cfg, dom, re, im, ti, to, vals


def _oracle(dom, re, im, ti, to, vals, dirichlet):
    r = dom.r[3]
    D1 = np_oracle.fd_dense(1, r, 3)
    return np_bc.apply_scalar_bc(re, im, ti, to, vals, r, D1, dirichlet=dirichlet)


# nr <= 80 and nr <= 128 take the register kernels (5 and 8 levels per thread), 129 the loop kernel
@pytest.mark.parametrize("lmax,nr,dirichlet", [(7, 9, True), (31, 17, True), (63, 16, False), (127, 65, True),
                                               (47, 33, False), (31, 97, True), (15, 129, True), (15, 129, False)])
def test_scalar_bc_matches_oracle(lmax, nr, dirichlet):
    import gdhip

    print("case: scalar bc lmax=%d nr=%d dirichlet=%s" % (lmax, nr, dirichlet), flush=True)
    cfg, dom, re, im, ti, to, vals = _case(lmax + nr, lmax, nr)
    bc = gdhip.ScalarBoundaryConditions(cfg, dom)
    bc.set(ti, to, vals)
    gre, gim = re.copy(), im.copy()
    bc.apply(gre, gim, dirichlet=dirichlet)
    ore, oim = _oracle(dom, re, im, ti, to, vals, dirichlet)
    for g, o in ((gre, ore), (gim, oim)):
        assert np.max(np.abs(g - o)) <= 1e-12 * (1.0 + np.abs(o).max())
    # modes without a condition (and no Dirichlet stage for them) are bit-identical
    free = (ti == 0) & (to == 0)
    assert np.array_equal(gre[:, free], re[:, free]) and np.array_equal(gim[:, free], im[:, free])
    # the tau tables
    tc = np_bc.TauCache(dom.r[3], np_oracle.fd_dense(1, dom.r[3], 3))
    tab, sc = bc.tau_tables()
    ref_tab = np.stack([tc.tau1, tc.tau2, tc.tau_in, tc.tau_out])
    assert np.max(np.abs(tab - ref_tab)) <= 1e-13 * np.abs(ref_tab).max()
    ref_sc = np.array([tc.dt1i, tc.dt1o, tc.dt2i, tc.dt2o, tc.dtau_in, tc.dtau_out])
    assert np.max(np.abs(sc - ref_sc) / np.abs(ref_sc)) <= 1e-12
    bc.close()


def test_device_tensors_and_field_functions():
    import torch

    import gdhip

    print("case: scalar bc device tensors + field functions lmax=31 nr=17", flush=True)
    cfg, dom, re, im, ti, to, vals = _case(5, 31, 17)
    f = gdhip.SHTnsSpectralField(cfg, 17, torch.from_numpy(re.copy()).cuda(), torch.from_numpy(im.copy()).cuda())
    f.bc_type_inner, f.bc_type_outer, f.boundary_values = ti, to, vals
    gdhip.apply_temperature_boundary_conditions_spectral(f, dom)
    torch.cuda.synchronize()
    ore, oim = _oracle(dom, re, im, ti, to, vals, True)
    assert np.max(np.abs(f.data_real.cpu().numpy() - ore)) <= 1e-12 * (1.0 + np.abs(ore).max())
    assert np.max(np.abs(f.data_imag.cpu().numpy() - oim)) <= 1e-12 * (1.0 + np.abs(oim).max())
    # host field, flux stage only, mmax < lmax
    cfg2, dom2, re2, im2, ti2, to2, vals2 = _case(6, 31, 17, mmax=9)
    h = gdhip.SHTnsSpectralField(cfg2, 17, re2.copy(), im2.copy())
    h.bc_type_inner, h.bc_type_outer, h.boundary_values = ti2, to2, vals2
    gdhip.apply_scalar_flux_bc_spectral(h, dom2)
    ore2, oim2 = _oracle(dom2, re2, im2, ti2, to2, vals2, False)
    assert np.max(np.abs(h.data_real - ore2)) <= 1e-12 * (1.0 + np.abs(ore2).max())
    assert np.max(np.abs(h.data_imag - oim2)) <= 1e-12 * (1.0 + np.abs(oim2).max())
    with pytest.raises(NotImplementedError):
        gdhip.apply_scalar_flux_bc_spectral(h, dom2, method="influence_matrix")
    # the composition function is the same two stages
    c = gdhip.SHTnsSpectralField(cfg2, 17, re2.copy(), im2.copy())
    c.bc_type_inner, c.bc_type_outer, c.boundary_values = ti2, to2, vals2
    gdhip.apply_composition_boundary_conditions_spectral(c, dom2)
    ore3, _ = _oracle(dom2, re2, im2, ti2, to2, vals2, True)
    assert np.max(np.abs(c.data_real - ore3)) <= 1e-12 * (1.0 + np.abs(ore3).max())
