"""Transform entry points mirroring shtnskit_field_functions.jl.

Each call runs the whole radial stack of a field in one batched GPU pass
(Legendre contraction on MFMA + LDS ring FFT) instead of the reference's
per-level loop.
"""
from ._arrays import prepare
from ._lib import check, lib


def _nr(spec):
    return spec.data_real.shape[0]


def shtnskit_spectral_to_physical(spec, phys):
    """shtnskit_field_functions.jl:11-21 (scalar synthesis)."""
    cfg = spec.config
    ctx, flags, (a, b, o) = prepare(spec.data_real, spec.data_imag, phys.data, device=cfg.device)
    check(lib().gd_sht_scalar_synth(cfg._plan, a, b, _nr(spec), o, flags))
    return None


def shtnskit_physical_to_spectral(phys, spec):
    """shtnskit_field_functions.jl:131-141 (scalar analysis; imag(m=0) := 0)."""
    cfg = spec.config
    ctx, flags, (f, a, b) = prepare(phys.data, spec.data_real, spec.data_imag, device=cfg.device)
    check(lib().gd_sht_scalar_analysis(cfg._plan, f, _nr(spec), a, b, flags))
    return None


def shtnskit_vector_synthesis(tor_spec, pol_spec, vec_phys):
    """shtnskit_field_functions.jl:252-282: (v_theta, v_phi) from S := pol, T := tor.
    The r component is not written (as in the reference)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(tor_spec.data_real, tor_spec.data_imag, pol_spec.data_real,
                            pol_spec.data_imag, vec_phys.θ_component.data, vec_phys.φ_component.data,
                            device=cfg.device)
    check(lib().gd_sht_vector_synth(cfg._plan, p[0], p[1], p[2], p[3], _nr(tor_spec), p[4], p[5], flags))
    return None


def shtnskit_vector_analysis(vec_phys, tor_spec, pol_spec):
    """shtnskit_field_functions.jl:291-321: (pol, tor) = spat_to_SHsphtor(v_theta, v_phi)."""
    cfg = tor_spec.config
    ctx, flags, p = prepare(vec_phys.θ_component.data, vec_phys.φ_component.data, tor_spec.data_real,
                            tor_spec.data_imag, pol_spec.data_real, pol_spec.data_imag, device=cfg.device)
    check(lib().gd_sht_vector_analysis(cfg._plan, p[0], p[1], _nr(tor_spec), p[2], p[3], p[4], p[5], flags))
    return None


def batch_spectral_to_physical(specs, physs):
    """shtnskit_field_functions.jl:674 (one synthesis per pair)."""
    assert len(specs) == len(physs)
    for s, p in zip(specs, physs):
        shtnskit_spectral_to_physical(s, p)


def batch_shtnskit_transforms(specs, physs):
    """shtnskit_field_functions.jl:651-663."""
    batch_spectral_to_physical(specs, physs)
