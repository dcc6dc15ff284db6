"""gdhip -- MI355X-native pseudo-spectral CNAB2 step behind GeoDynamo.jl's
transform and timestep entry points.

The compute path is libgdhip.so (hand-written HIP for gfx950, C ABI in
include/gdhip.h).  Names mirror the reference's Julia functions with the
trailing '!' dropped.
"""
from ._lib import GDError, MissingLError, lib  # noqa: F401
from .fields import (RadialDomain, SHTnsPhysicalField, SHTnsSpectralField,  # noqa: F401
                     SHTnsVectorField, create_radial_domain, seeded_state)
from .linear_algebra import (BandedLU, BandedMatrix, apply_banded_matrix, apply_derivative_matrix,  # noqa: F401
                             banded_to_dense,
                             create_derivative_matrix, create_radial_laplacian, factorize_banded,
                             solve_banded)
from .shtnskit_field_functions import (batch_shtnskit_transforms, batch_spectral_to_physical,  # noqa: F401
                                       shtnskit_physical_to_spectral, shtnskit_physical_to_spectral_adjoint,
                                       shtnskit_spectral_to_physical, shtnskit_spectral_to_physical_adjoint,
                                       shtnskit_vector_analysis, shtnskit_vector_analysis_adjoint,
                                       shtnskit_vector_synthesis, shtnskit_vector_synthesis_adjoint)
from .shtnskit_transforms import SHTnsKitConfig, create_shtnskit_config  # noqa: F401
from .step import CNAB2Step, StepBatch  # noqa: F401
from .timestep import (ERK2Cache, ERK2FieldBuffers, ETDCache, SHTnsImplicitMatrices,  # noqa: F401
                       apply_explicit_operator, build_rhs_cnab2, compute_timestep_error, create_erk2_cache,
                       create_etd_cache, create_shtns_timestepping_matrices, eab2_update, erk2_apply_stage,
                       erk2_finalize_field, erk2_prepare_field, erk2_store_stage_nonlinear, exp_functions,
                       solve_implicit_step)
from .outputs_writer import (RANK_TIME, TIME_RANK, OutputConfig, TimeTracker, find_restart_files,  # noqa: F401
                             generate_filename, read_restart, read_restart_arrays, write_restart,
                             write_restart_arrays)
from .scalar_field_common import (ScalarBoundaryConditions, apply_composition_boundary_conditions_spectral,  # noqa: F401
                                  apply_scalar_flux_bc_spectral, apply_temperature_boundary_conditions_spectral)
from .adjoint import CNAB2Adjoint, nonlinear_adjoint, step_vjp, surface_level  # noqa: F401
