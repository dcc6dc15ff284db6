"""Round trip of the transforms at the C1 size: lmax = mmax = 255 on a
384 x 768 Gauss grid, 97 radial levels, 3 scalar fields and one (tor, pol)
pair, coefficients seeded normal x (l+1)^-1.5 (m = 0 real, l = 0 of the
vector pair zero).  Synthesis then analysis through the C ABI with device
tensors; the relative L2 error of the recovered coefficients is what the
configuration reports (shtnskit_field_functions.jl:11-21, 131-141, 252-321).
Bars: FP64 <= 1e-12, FP32 <= 1e-5, BF16x3 (split-precision BF16 Legendre
contractions with FP32 accumulation) <= 1e-4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LMAX, NLAT, NLON, NR = 255, 384, 768, 97
TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4}


def _rel_l2(re, im, ref):
    d = (re - ref.real) ** 2 + (im - ref.imag) ** 2
    return float(np.sqrt(d.sum() / (np.abs(ref) ** 2).sum()))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_c1_round_trip(prec):
    import torch

    import gdhip

    print("case: C1 round trip lmax=%d %dx%d nr=%d %s" % (LMAX, NLAT, NLON, NR, prec), flush=True)
    cfg = gdhip.create_shtnskit_config(LMAX, nlat=NLAT, nlon=NLON, precision=prec)
    st = gdhip.seeded_state(2025, LMAX, NR, nfields=5, amp=1.0, power=1.5, zero_boundaries=False)
    lo = cfg.l_values
    st[3:, :, lo == 0] = 0.0  # the vector pair has no l = 0 part

    def dev(x):
        return torch.from_numpy(np.ascontiguousarray(x)).cuda()

    def phys():
        return gdhip.SHTnsPhysicalField(cfg, NR, data=torch.empty((NR, NLON, NLAT), dtype=torch.float64,
                                                                  device="cuda"))

    def spec(z=None):
        if z is None:
            return gdhip.SHTnsSpectralField(cfg, NR, torch.zeros((NR, cfg.nlm), dtype=torch.float64, device="cuda"),
                                            torch.zeros((NR, cfg.nlm), dtype=torch.float64, device="cuda"))
        return gdhip.SHTnsSpectralField(cfg, NR, dev(z.real), dev(z.imag))

    errs = []
    for f in range(3):
        p, back = phys(), spec()
        gdhip.shtnskit_spectral_to_physical(spec(st[f]), p)
        gdhip.shtnskit_physical_to_spectral(p, back)
        torch.cuda.synchronize()
        errs.append(_rel_l2(back.data_real.cpu().numpy(), back.data_imag.cpu().numpy(), st[f]))
        del p, back
    vec = gdhip.SHTnsVectorField(cfg, NR, comps=[phys(), phys(), phys()])
    tor, pol = spec(), spec()
    gdhip.shtnskit_vector_synthesis(spec(st[3]), spec(st[4]), vec)
    gdhip.shtnskit_vector_analysis(vec, tor, pol)
    torch.cuda.synchronize()
    errs.append(_rel_l2(tor.data_real.cpu().numpy(), tor.data_imag.cpu().numpy(), st[3]))
    errs.append(_rel_l2(pol.data_real.cpu().numpy(), pol.data_imag.cpu().numpy(), st[4]))
    print("C1 %s relative L2 round-trip errors (3 scalars, tor, pol): %s" % (prec, ["%.2e" % e for e in errs]),
          flush=True)
    assert max(errs) < TOL[prec], errs
