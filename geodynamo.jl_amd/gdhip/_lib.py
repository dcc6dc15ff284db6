"""ctypes binding of libgdhip.so (the C ABI declared in include/gdhip.h).

The library must be built in-tree (``python geodynamo.jl_amd/gdhip/_build.py``
or ``__graft_entry__.build()``); loading fails loudly when it is missing.
There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgdhip.so")

GD_OK, GD_ERR_ARG, GD_ERR_HIP, GD_ERR_NUMERIC, GD_ERR_MISSING_L, GD_ERR_NOMEM = range(6)
GD_PREC_FP64, GD_PREC_FP32, GD_PREC_BF16X3 = 0, 1, 2
GD_HOST = 1
GD_FD_LITERAL, GD_FD_TRANSPOSE = 0, 1
GD_EXP_ETD, GD_EXP_ERK2 = 0, 1

# every symbol include/gdhip.h declares
EXPORTS = [
    "gd_last_error", "gd_version", "gd_ctx_create", "gd_ctx_set_stream", "gd_ctx_synchronize",
    "gd_ctx_destroy", "gd_device_alloc", "gd_device_free", "gd_device_copy", "gd_seeded_state", "gd_radial_grid", "gd_derivative_matrix", "gd_radial_laplacian",
    "gd_banded_factor", "gd_banded_solve", "gd_cnab2_mats_create", "gd_cnab2_mats_get",
    "gd_cnab2_mats_destroy", "gd_build_rhs_cnab2", "gd_solve_implicit_step", "gd_cnab2_rhs_solve",
    "gd_sht_plan_create", "gd_sht_plan_destroy", "gd_sht_nlm", "gd_sht_grid", "gd_sht_tables",
    "gd_sht_scalar_synth", "gd_sht_scalar_analysis", "gd_sht_vector_synth", "gd_sht_vector_analysis",
    "gd_step_create", "gd_step_destroy", "gd_step_set_state", "gd_step_get_state",
    "gd_step_nonlinear", "gd_step_run", "gd_step_run_graph", "gd_step_count", "gd_step_energies",
    "gd_step_profile",
    "gd_timestep_error",
    "gd_exp_functions", "gd_etd_cache_create", "gd_etd_cache_get", "gd_etd_cache_destroy", "gd_eab2_update",
    "gd_erk2_cache_create", "gd_erk2_cache_from_matrices", "gd_erk2_cache_get", "gd_erk2_cache_destroy",
    "gd_erk2_buffers_create", "gd_erk2_buffers_destroy", "gd_erk2_prepare_field", "gd_erk2_apply_stage",
    "gd_erk2_store_stage_nonlinear", "gd_erk2_finalize_field",
    "gd_batch_create", "gd_batch_run", "gd_batch_destroy", "gd_batch_set_mode", "gd_batch_set_tiles",
    "gd_batch_get_tiles", "gd_batch_tune", "gd_batch_set_lanes", "gd_batch_get_lanes", "gd_batch_tune_lanes",
    "gd_step_get_history", "gd_step_restore", "gd_step_shape", "gd_step_radii", "gd_step_set_temperature_bc",
    "gd_step_enable_composition", "gd_step_set_composition", "gd_step_get_composition", "gd_step_set_composition_bc",
    "gd_step_has_composition", "gd_step_get_composition_history", "gd_step_restore_composition",
    "gd_restart_write", "gd_restart_read", "gd_restart_write_arrays", "gd_restart_read_arrays",
    "gd_restart_filename",
    "gd_scalar_bc_create", "gd_scalar_bc_set", "gd_scalar_bc_apply", "gd_scalar_bc_tau", "gd_scalar_bc_destroy",
    "gd_step_nonlinear_adjoint", "gd_step_vjp", "gd_adjoint_create", "gd_adjoint_destroy", "gd_adjoint_forward",
    "gd_adjoint_misfit", "gd_adjoint_seed", "gd_adjoint_backward",
    "gd_sht_scalar_synth_adjoint", "gd_sht_scalar_analysis_adjoint", "gd_sht_vector_synth_adjoint",
    "gd_sht_vector_analysis_adjoint", "gd_timestep_error_noalloc",
]


class GDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gdhip error %d: %s" % (code, msg))
        self.code = code


class MissingLError(GDError):
    pass


class RestartMeta(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("last_output_time", ctypes.c_double),
                ("output_count", ctypes.c_int), ("restart_count", ctypes.c_int),
                ("grid_file_written", ctypes.c_int), ("mpi_rank", ctypes.c_int),
                ("mpi_total_ranks", ctypes.c_int), ("geometry", ctypes.c_char * 32)]


class RestartComp(ctypes.Structure):
    _fields_ = [("re", ctypes.c_void_p), ("im", ctypes.c_void_p), ("nl_re", ctypes.c_void_p),
                ("nl_im", ctypes.c_void_p), ("present", ctypes.c_int), ("has_history", ctypes.c_int)]


class StepDesc(ctypes.Structure):
    _fields_ = [("E", ctypes.c_double), ("Pm", ctypes.c_double), ("Pr", ctypes.c_double),
                ("Ra", ctypes.c_double), ("rratio", ctypes.c_double), ("dt", ctypes.c_double),
                ("theta", ctypes.c_double),
                ("lmax", ctypes.c_int), ("mmax", ctypes.c_int), ("nlat", ctypes.c_int),
                ("nlon", ctypes.c_int), ("nr", ctypes.c_int), ("bw", ctypes.c_int),
                ("prec", ctypes.c_int), ("fd_mode", ctypes.c_int), ("radials", ctypes.c_int),
                ("coriolis_index", ctypes.c_int), ("bc_rows", ctypes.c_int), ("full_products", ctypes.c_int),
                ("radial_force", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgdhip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d, l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        pp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "gd_last_error": (ctypes.c_char_p, []),
            "gd_version": (ctypes.c_char_p, []),
            "gd_ctx_create": (i, [i, pp]),
            "gd_ctx_set_stream": (i, [vp, vp]),
            "gd_ctx_synchronize": (i, [vp]),
            "gd_ctx_destroy": (i, [vp]),
            "gd_device_alloc": (i, [vp, ctypes.c_size_t, pp]),
            "gd_device_free": (i, [vp, vp]),
            "gd_device_copy": (i, [vp, vp, vp, ctypes.c_size_t]),
            "gd_radial_grid": (i, [i, d, vp]),
            "gd_seeded_state": (i, [ctypes.c_ulonglong, i, i, i, i, d, d, i, vp, vp]),
            "gd_derivative_matrix": (i, [i, i, i, vp, i, vp]),
            "gd_radial_laplacian": (i, [i, i, vp, i, vp]),
            "gd_banded_factor": (i, [vp, i, i, i, vp, vp, i]),
            "gd_banded_solve": (i, [vp, i, i, i, vp, vp, i, vp, i]),
            "gd_cnab2_mats_create": (i, [vp, i, i, i, vp, d, d, d, i, i, pp]),
            "gd_cnab2_mats_get": (i, [vp, vp, vp, vp]),
            "gd_cnab2_mats_destroy": (i, [vp]),
            "gd_build_rhs_cnab2": (i, [vp, i, i, i] + [vp] * 8 + [i]),
            "gd_solve_implicit_step": (i, [vp, i, i, i] + [vp] * 4 + [i]),
            "gd_cnab2_rhs_solve": (i, [vp, i, i, i] + [vp] * 6 + [i]),
            "gd_sht_plan_create": (i, [vp, i, i, i, i, i, pp]),
            "gd_sht_plan_destroy": (i, [vp]),
            "gd_sht_nlm": (i, [vp]),
            "gd_sht_grid": (i, [vp, vp, vp, vp]),
            "gd_sht_tables": (i, [vp, vp, vp]),
            "gd_sht_scalar_synth": (i, [vp, vp, vp, i, vp, i]),
            "gd_sht_scalar_analysis": (i, [vp, vp, i, vp, vp, i]),
            "gd_sht_vector_synth": (i, [vp, vp, vp, vp, vp, i, vp, vp, i]),
            "gd_sht_vector_analysis": (i, [vp, vp, vp, i, vp, vp, vp, vp, i]),
            "gd_step_create": (i, [vp, ctypes.POINTER(StepDesc), pp]),
            "gd_step_destroy": (i, [vp]),
            "gd_step_set_state": (i, [vp, vp, vp, i]),
            "gd_step_get_state": (i, [vp, vp, vp, i]),
            "gd_step_nonlinear": (i, [vp, vp, vp, i]),
            "gd_step_run": (i, [vp, i]),
            "gd_step_run_graph": (i, [vp, i]),
            "gd_step_count": (i, [vp]),
            "gd_step_energies": (i, [vp, vp]),
            "gd_step_profile": (i, [vp, i, vp]),
            "gd_timestep_error": (i, [vp, l, vp, vp, vp, vp, vp, i]),
            "gd_timestep_error_noalloc": (i, [vp, l, vp, vp, vp, vp, vp]),
            "gd_exp_functions": (i, [i, i, i, vp, d, d, i, i, vp]),
            "gd_etd_cache_create": (i, [vp, i, i, i, vp, d, d, i, pp]),
            "gd_etd_cache_get": (i, [vp, vp, vp]),
            "gd_etd_cache_destroy": (i, [vp]),
            "gd_eab2_update": (i, [vp, i, i, i] + [vp] * 6 + [d, i]),
            "gd_erk2_cache_create": (i, [vp, i, i, i, vp, d, d, i, pp]),
            "gd_erk2_cache_from_matrices": (i, [vp, i, vp, i, d] + [vp] * 5 + [pp]),
            "gd_erk2_cache_get": (i, [vp] * 6),
            "gd_erk2_cache_destroy": (i, [vp]),
            "gd_erk2_buffers_create": (i, [vp, i, i, i, pp]),
            "gd_erk2_buffers_destroy": (i, [vp]),
            "gd_erk2_prepare_field": (i, [vp] * 5 + [d, i]),
            "gd_erk2_apply_stage": (i, [vp, vp, vp, i]),
            "gd_erk2_store_stage_nonlinear": (i, [vp, vp, vp, i]),
            "gd_erk2_finalize_field": (i, [vp, vp, vp, d, i]),
            "gd_batch_create": (i, [vp, i, vp, i, pp]),
            "gd_batch_run": (i, [vp, i]),
            "gd_batch_destroy": (i, [vp]),
            "gd_batch_set_mode": (i, [vp, i]),
            "gd_batch_set_tiles": (i, [vp, vp]),
            "gd_batch_get_tiles": (i, [vp, vp]),
            "gd_batch_tune": (i, [vp, i, vp]),
            "gd_batch_set_lanes": (i, [vp, i]),
            "gd_batch_get_lanes": (i, [vp, vp]),
            "gd_batch_tune_lanes": (i, [vp, i, vp]),
            "gd_step_get_history": (i, [vp, vp, vp, i]),
            "gd_step_restore": (i, [vp] * 5 + [i, i]),
            "gd_step_shape": (i, [vp, vp, vp, vp]),
            "gd_step_radii": (i, [vp, vp]),
            "gd_step_set_temperature_bc": (i, [vp, vp, vp, vp]),
            "gd_step_enable_composition": (i, [vp, d]),
            "gd_step_set_composition": (i, [vp, vp, vp, i]),
            "gd_step_get_composition": (i, [vp, vp, vp, i]),
            "gd_step_set_composition_bc": (i, [vp, vp, vp, vp]),
            "gd_step_has_composition": (i, [vp]),
            "gd_step_get_composition_history": (i, [vp, vp, vp, i]),
            "gd_step_restore_composition": (i, [vp] * 5 + [i]),
            "gd_restart_write": (i, [vp, ctypes.c_char_p, ctypes.POINTER(RestartMeta), i]),
            "gd_restart_read": (i, [vp, ctypes.c_char_p, ctypes.POINTER(RestartMeta)]),
            "gd_restart_write_arrays": (i, [ctypes.c_char_p, ctypes.POINTER(RestartMeta), i, i, i, vp, i]
                                        + [vp] * 4 + [i, ctypes.POINTER(RestartComp)]),
            "gd_restart_read_arrays": (i, [ctypes.c_char_p, i, i, i, vp, vp] + [vp] * 4
                                       + [vp, ctypes.POINTER(RestartMeta), ctypes.POINTER(RestartComp)]),
            "gd_restart_filename": (i, [ctypes.c_char_p] * 4 + [d, i, i, ctypes.c_char_p, i]),
            "gd_scalar_bc_create": (i, [vp, i, i, i, i, vp, i, pp]),
            "gd_scalar_bc_set": (i, [vp, vp, vp, vp]),
            "gd_scalar_bc_apply": (i, [vp, vp, vp, i, i]),
            "gd_scalar_bc_tau": (i, [vp, vp, vp]),
            "gd_scalar_bc_destroy": (i, [vp]),
            "gd_step_nonlinear_adjoint": (i, [vp] * 5 + [i]),
            "gd_step_vjp": (i, [vp] * 9 + [i]),
            "gd_adjoint_create": (i, [vp, i, i, pp]),
            "gd_adjoint_destroy": (i, [vp]),
            "gd_adjoint_forward": (i, [vp]),
            "gd_adjoint_misfit": (i, [vp, i, i, vp, ctypes.POINTER(d), i]),
            "gd_adjoint_seed": (i, [vp, vp, vp, i]),
            "gd_adjoint_backward": (i, [vp, vp, vp, i]),
            "gd_sht_scalar_synth_adjoint": (i, [vp, vp, i, vp, vp, i]),
            "gd_sht_scalar_analysis_adjoint": (i, [vp, vp, vp, i, vp, i]),
            "gd_sht_vector_synth_adjoint": (i, [vp, vp, vp, i, vp, vp, vp, vp, i]),
            "gd_sht_vector_analysis_adjoint": (i, [vp, vp, vp, vp, vp, i, vp, vp, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != GD_OK:
        msg = lib().gd_last_error().decode(errors="replace")
        if rc == GD_ERR_MISSING_L:
            raise MissingLError(rc, msg)
        raise GDError(rc, msg)
    return rc
