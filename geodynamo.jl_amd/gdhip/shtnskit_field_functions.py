"""Transform entry points mirroring shtnskit_field_functions.jl.

Each call runs the whole radial stack of a field in one batched GPU pass
(Legendre contraction on MFMA + LDS ring FFT) instead of the reference's
per-level loop.
"""
import numpy as np

from ._arrays import prepare
from ._lib import check, lib


def _nr(spec):
    return spec.data_real.shape[0]


def _shapes(cfg, nr, kinds):
    """Exact array shapes of a call: "s" a spectral plane (nr, nlm), "p" a
    physical field (nr, nlon, nlat)."""
    return [(nr, cfg.nlm) if k == "s" else (nr, cfg.nlon, cfg.nlat) for k in kinds]


def shtnskit_spectral_to_physical(spec, phys):
    """shtnskit_field_functions.jl:11-21 (scalar synthesis)."""
    cfg = spec.config
    ctx, flags, (a, b, o) = prepare(spec.data_real, spec.data_imag, phys.data, device=cfg.device,
                            shapes=_shapes(cfg, _nr(spec), "ssp"))
    check(lib().gd_sht_scalar_synth(cfg._plan, a, b, _nr(spec), o, flags))
    return None


def shtnskit_physical_to_spectral(phys, spec):
    """shtnskit_field_functions.jl:131-141 (scalar analysis; imag(m=0) := 0)."""
    cfg = spec.config
    ctx, flags, (f, a, b) = prepare(phys.data, spec.data_real, spec.data_imag, device=cfg.device,
                            shapes=_shapes(cfg, _nr(spec), "pss"))
    check(lib().gd_sht_scalar_analysis(cfg._plan, f, _nr(spec), a, b, flags))
    return None


def shtnskit_vector_synthesis(tor_spec, pol_spec, vec_phys):
    """shtnskit_field_functions.jl:252-282: (v_theta, v_phi) from S := pol, T := tor.
    The r component is not written (as in the reference)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(tor_spec.data_real, tor_spec.data_imag, pol_spec.data_real,
                            pol_spec.data_imag, vec_phys.θ_component.data, vec_phys.φ_component.data,
                            device=cfg.device, shapes=_shapes(cfg, _nr(tor_spec), "sssspp"))
    check(lib().gd_sht_vector_synth(cfg._plan, p[0], p[1], p[2], p[3], _nr(tor_spec), p[4], p[5], flags))
    return None


def shtnskit_vector_analysis(vec_phys, tor_spec, pol_spec):
    """shtnskit_field_functions.jl:291-321: (pol, tor) = spat_to_SHsphtor(v_theta, v_phi)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(vec_phys.θ_component.data, vec_phys.φ_component.data, tor_spec.data_real,
                            tor_spec.data_imag, pol_spec.data_real, pol_spec.data_imag, device=cfg.device,
                            shapes=_shapes(cfg, _nr(tor_spec), "ppssss"))
    check(lib().gd_sht_vector_analysis(cfg._plan, p[0], p[1], _nr(tor_spec), p[2], p[3], p[4], p[5], flags))
    return None


def shtnskit_spectral_to_physical_adjoint(phys, spec):
    """Transpose of shtnskit_spectral_to_physical: spec = S^T phys."""
    cfg = spec.config
    ctx, flags, (f, a, b) = prepare(phys.data, spec.data_real, spec.data_imag, device=cfg.device,
                            shapes=_shapes(cfg, _nr(spec), "pss"))
    check(lib().gd_sht_scalar_synth_adjoint(cfg._plan, f, _nr(spec), a, b, flags))
    return None


def shtnskit_physical_to_spectral_adjoint(spec, phys):
    """Transpose of shtnskit_physical_to_spectral: phys = A^T spec."""
    cfg = spec.config
    ctx, flags, (a, b, o) = prepare(spec.data_real, spec.data_imag, phys.data, device=cfg.device,
                            shapes=_shapes(cfg, _nr(spec), "ssp"))
    check(lib().gd_sht_scalar_analysis_adjoint(cfg._plan, a, b, _nr(spec), o, flags))
    return None


def shtnskit_vector_synthesis_adjoint(vec_phys, tor_spec, pol_spec):
    """Transpose of shtnskit_vector_synthesis: (tor, pol) = S_v^T (v_theta, v_phi)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(vec_phys.θ_component.data, vec_phys.φ_component.data, tor_spec.data_real,
                            tor_spec.data_imag, pol_spec.data_real, pol_spec.data_imag, device=cfg.device,
                            shapes=_shapes(cfg, _nr(tor_spec), "ppssss"))
    check(lib().gd_sht_vector_synth_adjoint(cfg._plan, p[0], p[1], _nr(tor_spec), p[2], p[3], p[4], p[5], flags))
    return None


def shtnskit_vector_analysis_adjoint(tor_spec, pol_spec, vec_phys):
    """Transpose of shtnskit_vector_analysis: (v_theta, v_phi) = A_v^T (tor, pol)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(tor_spec.data_real, tor_spec.data_imag, pol_spec.data_real, pol_spec.data_imag,
                            vec_phys.θ_component.data, vec_phys.φ_component.data, device=cfg.device,
                            shapes=_shapes(cfg, _nr(tor_spec), "sssspp"))
    check(lib().gd_sht_vector_analysis_adjoint(cfg._plan, p[0], p[1], p[2], p[3], _nr(tor_spec), p[4], p[5], flags))
    return None


def _concat(arrays):
    if type(arrays[0]).__module__.startswith("torch"):
        import torch

        return torch.cat([a.contiguous() for a in arrays], dim=0).contiguous()
    return np.ascontiguousarray(np.concatenate(arrays, axis=0))


def batch_spectral_to_physical(specs, physs):
    """shtnskit_field_functions.jl:674: the syntheses of many fields.  Fields
    sharing a configuration are stacked along the radial axis and synthesised
    in one batched pass (one Legendre launch and one ring launch for the
    group), then split back; results equal the one-by-one calls."""
    if len(specs) != len(physs):
        raise ValueError("batch_spectral_to_physical: %d spectral vs %d physical fields" % (len(specs), len(physs)))
    groups = {}
    for s, p in zip(specs, physs):
        groups.setdefault(id(s.config), []).append((s, p))
    for pairs in groups.values():
        cfg = pairs[0][0].config
        if len(pairs) == 1:
            shtnskit_spectral_to_physical(*pairs[0])
            continue
        re = _concat([s.data_real for s, _ in pairs])
        im = _concat([s.data_imag for s, _ in pairs])
        nr = sum(_nr(s) for s, _ in pairs)
        out = _concat([p.data for _, p in pairs])
        ctx, flags, (a, b, o) = prepare(re, im, out, device=cfg.device)
        check(lib().gd_sht_scalar_synth(cfg._plan, a, b, nr, o, flags))
        k = 0
        for s, p in pairs:
            n = _nr(s)
            p.data[...] = out[k:k + n]
            k += n


def batch_shtnskit_transforms(specs, physs):
    """shtnskit_field_functions.jl:651-663 (the reference runs the pairs'
    syntheses concurrently; here they share batched launches)."""
    batch_spectral_to_physical(specs, physs)
