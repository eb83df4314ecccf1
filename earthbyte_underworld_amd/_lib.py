"""ctypes binding of libust.so (the C ABI in include/ust.h).

The library must have been built (``python -m earthbyte_underworld_amd.build``); loading fails loudly
otherwise -- there is no CPU fallback.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libust.so")

UST_OK = 0
PREC_FP64, PREC_MIXED = 0, 1
SCHUR_UW, SCHUR_LEARNED, SCHUR_SCALED, SCHUR_GKGDIAG = 0, 1, 2, 3
MAT_K, MAT_G = 0, 1
VEC_F, VEC_H, VEC_U, VEC_P = 0, 1, 2, 3
FIELD_VELOCITY, FIELD_PRESSURE, FIELD_TEMPERATURE = 0, 1, 2
RH_MATERIALVISCOSITY, RH_ARRHENIUS, RH_VISCOSITYLIMITER, RH_VONMISES, RH_DRUCKERPRAGER = 1, 2, 3, 4, 5
SHAPE_EVERYWHERE, SHAPE_BELOWCOSINEPLANE, SHAPE_SPHERE, SHAPE_BOX = 0, 1, 2, 3
TERM_CONSTITUTIVE, TERM_GRADIENT, TERM_UZAWA_PC, TERM_BUOYANCY = 1, 2, 3, 4
MAX_MATERIALS = 64


class Rheology(C.Structure):
    _fields_ = [("type", C.c_int), ("p", C.c_double * 8)]


class Material(C.Structure):
    _fields_ = [("nrh", C.c_int), ("rh", Rheology * 4), ("density", C.c_double), ("alpha", C.c_double),
                ("ref_temperature", C.c_double), ("beta", C.c_double), ("ref_depth", C.c_double)]


class Buoyancy(C.Structure):
    _fields_ = [("vertical_axis", C.c_int), ("use_direction", C.c_int), ("direction", C.c_double * 3),
                ("adjust", C.c_int), ("hydrostatic", C.c_int)] + \
        [(k, C.c_double) for k in ("hs_upper_density", "hs_upper_alpha", "hs_lower_density", "hs_lower_alpha",
                                   "hs_height", "hs_material_boundary", "hs_T0", "hs_linear_coefficient",
                                   "hs_exponential_coefficient1", "hs_exponential_coefficient2", "hs_v",
                                   "hs_width", "hs_time")]


class Shape(C.Structure):
    _fields_ = [("type", C.c_int), ("material", C.c_int), ("p", C.c_double * 6)]


class StrainWeakening(C.Structure):
    _fields_ = [("softening_strain", C.c_double), ("healing_rate", C.c_double),
                ("initial_damage_fraction", C.c_double), ("initial_damage_wavenumber", C.c_double),
                ("wavenumber_sin", C.c_double * 3), ("wavenumber_cos", C.c_double * 3),
                ("initial_damage_factor", C.c_double), ("random_seed", C.c_long), ("has_shape", C.c_int),
                ("initial_strain_shape", Shape)]


class Pcdvc(C.Structure):
    _fields_ = [("res", C.c_int * 3), ("upper_t", C.c_double), ("lower_t", C.c_double),
                ("max_deletions", C.c_int), ("max_splits", C.c_int), ("split_in_interface_cells", C.c_int),
                ("delete_in_interface_cells", C.c_int), ("allow_splitting_in_boundary_interface_cells", C.c_int),
                ("inflow", C.c_int)]


class PcdvcOpts(C.Structure):
    _fields_ = [("upper_t", C.c_double), ("lower_t", C.c_double), ("max_deletions", C.c_int),
                ("max_splits", C.c_int), ("split_in_interface_cells", C.c_int),
                ("delete_in_interface_cells", C.c_int)]


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nel", C.c_int * 3), ("h", C.c_double * 3), ("xmin", C.c_double * 3),
                ("ymax", C.c_double), ("xmax", C.c_double * 3)]


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("schur_pc", C.c_int), ("mg_levels", C.c_int),
                ("smoother_degree", C.c_int), ("cheby_ratio", C.c_double), ("outer_rtol", C.c_double),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int), ("max_inner", C.c_int),
                ("remove_nullspace", C.c_int), ("stall_window", C.c_int), ("inner_relax", C.c_int),
                ("inner_relax_cap", C.c_double), ("stokes_rtol", C.c_double), ("verbose", C.c_int),
                ("coarse_degree", C.c_int), ("coarsest_degree", C.c_int), ("vcycle_bf16", C.c_int),
                ("inner_fp32", C.c_int)]


class UzawaOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int), ("min_iterations", C.c_int),
                ("use_absolute_tolerance", C.c_int), ("preconditioner", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("outer_its", C.c_int), ("inner_its", C.c_int), ("inner_solves", C.c_int),
                ("refinements", C.c_int), ("fallback_used", C.c_int), ("mg_levels", C.c_int),
                ("stokes_residual", C.c_double), ("momentum_residual", C.c_double),
                ("continuity_residual", C.c_double), ("t_assemble_ms", C.c_double), ("t_setup_ms", C.c_double),
                ("t_solve_ms", C.c_double), ("restarts", C.c_int), ("nullspace_removed", C.c_int),
                ("nullspace_ratio", C.c_double), ("converged", C.c_int), ("hhat_rescaled", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


vp = C.c_void_p
I = C.c_int
D = C.c_double

# name -> (restype, argtypes)
_SIGS = {
    "ust_create": (I, [C.POINTER(vp), I]),
    "ust_destroy": (None, [vp]),
    "ust_error": (C.c_char_p, [vp]),
    "ust_get_stream": (vp, [vp]),
    "ust_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "ust_set_mesh": (I, [vp, I, vp, vp, vp]),
    "ust_set_materials": (I, [vp, I, vp]),
    "ust_rheology_from_xml": (I, [C.c_char_p, I, vp, vp, C.POINTER(Rheology)]),
    "ust_swarm_upload": (I, [vp, I, vp, vp, vp, vp, vp]),
    "ust_swarm_upload_dev": (I, [vp, I, vp, vp, vp, vp, vp, vp, vp, I]),
    "ust_swarm_random": (I, [vp, vp, I, C.c_long]),
    "ust_swarm_assign_materials": (I, [vp, I, vp, vp]),
    "ust_swarm_cohesion_noise": (I, [vp, C.c_float, C.c_ulonglong]),
    "ust_swarm_assign_materials_field": (I, [vp, vp, D, D, I, I]),
    "ust_swarm_download": (I, [vp, vp, vp, vp, vp]),
    "ust_swarm_count": (I, [vp]),
    "ust_swarm_download_integration_points": (I, [vp, vp, vp]),
    "ust_swarm_advect": (I, [vp, D, I, I, C.POINTER(I), C.POINTER(I)]),
    "ust_swarm_spacefiller": (I, [vp, D]),
    "ust_swarm_dvc_weights": (I, [vp, C.POINTER(I), I]),
    "ust_write_petsc_matrix": (I, [vp, I, I, C.c_char_p]),
    "ust_write_petsc_vector": (I, [vp, I, I, C.c_char_p]),
    "ust_save_field_hdf5": (I, [vp, I, I, C.c_char_p]),
    "ust_load_field_hdf5": (I, [vp, I, C.c_char_p]),
    "ust_hdf5_field_info": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_long),
                                C.POINTER(I)]),
    "ust_hdf5_field_read": (I, [C.c_char_p, vp, C.c_long, I]),
    "ust_hdf5_field_write": (I, [C.c_char_p, I, C.POINTER(I), C.c_long, I, vp]),
    "ust_hdf5_error": (C.c_char_p, []),
    "ust_set_temperature": (I, [vp, vp]),
    "ust_set_buoyancy": (I, [vp, C.POINTER(Buoyancy)]),
    "ust_strain_weakening_default": (None, [C.POINTER(StrainWeakening)]),
    "ust_pcdvc_default": (None, [C.POINTER(Pcdvc)]),
    "ust_swarm_pcdvc": (I, [vp, C.POINTER(Pcdvc), C.POINTER(I), C.POINTER(I)]),
    "ust_set_strain_weakening": (I, [vp, C.POINTER(StrainWeakening)]),
    "ust_strain_weakening_initialise": (I, [vp]),
    "ust_strain_weakening_step": (I, [vp, C.c_double]),
    "ust_get_strain_weakening": (I, [vp, vp, vp]),
    "ust_set_strain_weakening_state": (I, [vp, vp, vp]),
    "ust_get_tensile_failure": (I, [vp, vp]),
    "ust_set_solution": (I, [vp, vp, vp]),
    "ust_set_previous_solution_exists": (I, [vp, I]),
    "ust_set_bcs": (I, [vp, vp, vp]),
    "ust_assemble": (I, [vp, D]),
    "ust_get_element_matrices": (I, [vp, I, I, vp]),
    "ust_get_element_matrices_dev": (I, [vp, I, I, vp]),
    "ust_get_element_G": (I, [vp, I, I, vp]),
    "ust_get_element_f": (I, [vp, I, I, vp]),
    "ust_get_approxS": (I, [vp, vp]),
    "ust_get_particle_viscosity": (I, [vp, vp, vp]),
    "ust_get_weights": (I, [vp, vp, vp, vp]),
    "ust_get_rhs": (I, [vp, vp, vp]),
    "ust_get_device_views": (I, [vp, C.POINTER(Grid), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "ust_k_apply_K_ref_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp]),
    "ust_k_apply_K_ref_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp]),
    "ust_k_copy_f64_to_f32": (I, [vp, vp, C.c_size_t, vp]),
    "ust_k_weights_len": (C.c_size_t, [C.POINTER(Grid)]),
    "ust_apply_K": (I, [vp, vp, vp]),
    "ust_apply_K_dev": (I, [vp, vp, vp]),
    "ust_apply_K_f32_dev": (I, [vp, vp, vp]),
    "ust_apply_K_f32": (I, [vp, vp, vp]),
    "ust_get_cohesion_scale": (I, [vp, vp]),
    "ust_apply_G": (I, [vp, vp, vp]),
    "ust_apply_Gt": (I, [vp, vp, vp]),
    "ust_solve": (I, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ust_uzawa_opts_default": (None, [C.POINTER(UzawaOpts)]),
    "ust_solve_uzawa": (I, [vp, C.POINTER(UzawaOpts), C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ust_get_solution": (I, [vp, vp, vp]),
    "ust_get_schur_preconditioner": (I, [vp, C.POINTER(SolveOpts), vp]),
    "ust_get_solution_dev": (I, [vp, vp, vp]),
    "ust_velocity_solve": (I, [vp, C.POINTER(SolveOpts), vp, vp, C.POINTER(I)]),
    "ust_picard": (I, [vp, D, D, I, C.POINTER(SolveOpts), C.POINTER(I), C.POINTER(D), C.POINTER(SolveStats)]),
    "ust_mlp_init": (I, [vp, I, C.c_ulonglong]),
    "ust_mlp_set_params": (I, [vp, I, vp, vp]),
    "ust_mlp_get_params": (I, [vp, vp, vp]),
    "ust_mlp_param_counts": (I, [I, I, C.POINTER(I), C.POINTER(I)]),
    "ust_mlp_features": (I, [vp, vp]),
    "ust_mlp_probe_target": (I, [vp, C.POINTER(SolveOpts), I, C.c_ulonglong, vp]),
    "ust_set_schur_scaling": (I, [vp, vp]),
    "ust_mlp_train": (I, [vp, I, vp, vp, I, I, C.c_float, C.c_ulonglong, C.POINTER(C.c_float)]),
    "ust_mlp_predict": (I, [vp, I, vp, vp]),
    "ust_mlp_gradients": (I, [vp, I, vp, vp, vp, vp, vp]),
    "ust_stg_bind_term": (I, [vp, vp, I]),
    "ust_ConstitutiveMatrixCartesian_AssembleElement": (None, [vp, vp, I, vp, vp, vp]),
    "ust_GradientStiffnessMatrixTerm_AssembleElement": (None, [vp, vp, I, vp, vp, vp]),
    "ust_UzawaPreconditionerTerm_AssembleElement": (None, [vp, vp, I, vp, vp, vp]),
    "ust_BuoyancyForceTerm_AssembleElement": (None, [vp, vp, I, vp]),
    # thin kernel ABI used directly by tests and the benchmark timers
    "ust_k_apply_work_len": (C.c_size_t, [C.POINTER(Grid)]),
    "ust_k_apply_K_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_f64_noalloc": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp]),
    "ust_k_apply_noalloc_max_work": (C.c_size_t, []),
    "ust_k_apply_K_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_residual_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_cheby_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, D, D, I, vp, vp]),
    "ust_k_apply_K_residual_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_cheby_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, I, vp, vp]),
    "ust_k_apply_K_cheby_prolong_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, D, vp, vp, vp]),
    "ust_k_apply_K_cheby_prolong_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp]),
    "ust_k_apply_K_residual_first_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, C.c_float, vp, vp]),
    "ust_k_apply_K_cheby_oop_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_int,
                                        vp, vp]),
    "ust_k_apply_K_cheby_prolong_oop_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp]),
    "ust_k_cheby_first_f32": (I, [vp, vp, vp, vp, C.c_float, C.c_size_t, vp]),
    "ust_k_dot_f32": (I, [vp, vp, C.c_size_t, vp, vp, vp]),
    "ust_k_cg_p_guarded_f32": (I, [vp, vp, vp, C.c_size_t, vp]),
    "ust_k_cg_xr_rr_guarded_f32": (I, [vp, vp, vp, vp, vp, vp, vp, C.c_size_t, vp, vp, vp]),
    "ust_k_apply_K_f32_bf16w": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_residual_f32_bf16w": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp]),
    "ust_k_apply_K_cheby_f32_bf16w": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, I, vp,
                                          vp]),
    "ust_k_apply_K_cheby_prolong_f32_bf16w": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp]),
    "ust_k_copy_f64_to_bf16": (I, [vp, vp, C.c_size_t, vp]),
    "ust_k_restrict_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp]),
    "ust_k_restrict_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp]),
    "ust_k_restrict_cheby_first_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp, D, vp]),
    "ust_k_restrict_cheby_first_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp, vp, vp, vp, C.c_float, vp]),
    "ust_k_prolong_add_f64": (I, [C.POINTER(Grid), vp, vp, vp, vp]),
    "ust_k_prolong_add_f32": (I, [C.POINTER(Grid), vp, vp, vp, vp]),
    "ust_k_particle_weights": (I, [C.POINTER(Grid), vp, I, vp, vp, vp, vp, vp, vp, vp, I, vp, vp, vp, I, vp, vp,
                                   vp, vp, vp, vp, vp, vp]),
    "ust_k_strain_weakening_step": (I, [vp, vp, I, C.c_double, vp]),
    "ust_k_permute_f64": (I, [vp, vp, vp, I, vp]),
    "ust_k_pcdvc_plan": (I, [C.POINTER(Grid), vp, vp, vp, C.POINTER(PcdvcOpts), vp, vp, vp, vp]),
    "ust_k_pcdvc_build": (I, [I, I, I, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                              vp, vp]),
    "ust_event_create": (I, [C.POINTER(vp)]),
    "ust_event_destroy": (I, [vp]),
    "ust_event_record": (I, [vp, vp]),
    "ust_event_elapsed_ms": (I, [vp, vp, C.POINTER(C.c_float)]),
    "ust_stream_sync": (I, [vp]),
    "ust_dev_malloc": (I, [C.POINTER(vp), C.c_size_t]),
    "ust_dev_free": (I, [vp]),
    "ust_memcpy_h2d": (I, [vp, vp, C.c_size_t, vp]),
    "ust_memcpy_d2h": (I, [vp, vp, C.c_size_t, vp]),
    "ust_memcpy_d2d": (I, [vp, vp, C.c_size_t, vp]),
    "ust_k_fill_hash_f64": (I, [vp, C.c_size_t, C.c_ulonglong, vp]),
    "ust_k_dense_assemble": (I, [C.POINTER(Grid), vp, vp, vp, vp]),
    "ust_k_dense_spd_inverse": (I, [vp, I, vp, vp]),
    "ust_k_dense_matvec_f64": (I, [vp, I, vp, vp, vp]),
    "ust_k_dense_matvec_f32": (I, [vp, I, vp, vp, vp]),
    "ust_dev_error_string": (C.c_char_p, []),
}

_lib = None


def load():
    """Load libust.so; raises if it is missing (the HIP path is the only path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m earthbyte_underworld_amd.build` "
                           "(the gfx950 HIP library is the only implementation; there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = _Declared(lib)
    return _lib


class _Declared:
    """The loaded library restricted to the entry points with declared signatures: ctypes would otherwise
    pass 64-bit pointers and sizes of an undeclared call as 32-bit ints."""

    def __init__(self, cdll):
        self._cdll = cdll

    def __getattr__(self, name):
        if name not in _SIGS:
            raise AttributeError(f"{name}: no signature declared in earthbyte_underworld_amd/_lib.py")
        return getattr(self._cdll, name)


def exported_symbols():
    return list(_SIGS)
