"""Loader for the gfx950 library (codim-ipc_amd/libcipc.so) and its C ABI (include/cipc.h).

The library is the only compute path: if it is missing, every entry point raises.
There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libcipc.so"))

D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
F32 = ctypes.POINTER(ctypes.c_float)

CIPC_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "HIP runtime failure",
    -3: "non-positive contact distance",
    -4: "zero time of impact",
    -5: "unsupported configuration",
    -6: "no convergence",
    -7: "buffer too small",
}


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_vertices", ctypes.c_int32),
        ("x0", D),
        ("v0", D),
        ("n_triangles", ctypes.c_int32),
        ("triangles", I32),
        ("n_dbc", ctypes.c_int32),
        ("dbc_index", I32),
        ("dbc_target", D),
        ("density", ctypes.c_double),
        ("youngs", ctypes.c_double),
        ("poisson", ctypes.c_double),
        ("thickness", ctypes.c_double),
        ("bending_stiff_mult", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("dhat", ctypes.c_double),
        ("offset", ctypes.c_double),
        ("newton_tol", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("kappa_stiff_mult", ctypes.c_double),
        ("pcg_tol", ctypes.c_double),
        ("pcg_max_iter", ctypes.c_int32),
        ("with_collision", ctypes.c_int32),
        ("n_rod_segments", ctypes.c_int32),
        ("rod_segments", I32),
        ("rod_density", ctypes.c_double),
        ("rod_youngs", ctypes.c_double),
        ("rod_bend_stiff_mult", ctypes.c_double),
        ("rod_thickness", ctypes.c_double),
        ("n_particles", ctypes.c_int32),
        ("particles", I32),
        ("particle_density", ctypes.c_double),
        ("particle_thickness", ctypes.c_double),
        ("friction_mu", ctypes.c_double),
        ("friction_epsv2", ctypes.c_double),
        ("friction_iter_amt", ctypes.c_int32),
        ("strain_limit_kappa", ctypes.c_double * 2),
        ("strain_limit_s", ctypes.c_double * 2),
        ("strain_limit_shat", ctypes.c_double * 2),
        ("fiber_stiff_mult", ctypes.c_double * 4),
        ("fiber_limit", ctypes.c_double * 3),
        ("n_tets", ctypes.c_int32),
        ("tets", I32),
        ("tet_density", ctypes.c_double),
        ("tet_youngs", ctypes.c_double),
        ("tet_poisson", ctypes.c_double),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("newton_iters", ctypes.c_int32),
        ("contacts", ctypes.c_int32),
        ("pcg_iters", ctypes.c_int32),
        ("gd_fallbacks", ctypes.c_int32),
        ("ls_trials", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("kappa", ctypes.c_double),
        ("last_residual", ctypes.c_double),
        ("energy", ctypes.c_double),
        ("t_total_ms", ctypes.c_double),
        ("t_constraint_set_ms", ctypes.c_double),
        ("t_ccd_ms", ctypes.c_double),
        ("t_hessian_ms", ctypes.c_double),
        ("t_solve_ms", ctypes.c_double),
        ("t_energy_ms", ctypes.c_double),
        ("t_gradient_ms", ctypes.c_double),
        ("kappa_s", ctypes.c_double),
        ("kappa_s_doublings", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [
        ("spmv_grid", ctypes.c_int32),
        ("spmv_lanes", ctypes.c_int32),
        ("spmv_unroll", ctypes.c_int32),
        ("factor_grid", ctypes.c_int32),
        ("factor_lanes", ctypes.c_int32),
        ("tuning_pcg", ctypes.c_int32),
        ("tuning_cell", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("cell_mult", ctypes.c_double),
    ]


# exported symbols and their prototypes (restype, argtypes)
P = ctypes.c_void_p
PROTOTYPES = {
    "cipc_create": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "cipc_destroy": (None, [P]),
    "cipc_last_error": (ctypes.c_char_p, [P]),
    "cipc_get_sizes": (ctypes.c_int, [P, I32]),
    "cipc_get_scalars": (ctypes.c_int, [P, D]),
    "cipc_get_setup": (ctypes.c_int, [P, D, I32, D, I32]),
    "cipc_get_codim_sizes": (ctypes.c_int, [P, I32]),
    "cipc_set_precond_reuse": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_set_warm_start": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_autotune": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_get_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
    "cipc_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
    "cipc_batch_step": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, D]),
    "cipc_get_codim_setup": (ctypes.c_int, [P, D, I32, D, I32]),
    "cipc_set_state": (ctypes.c_int, [P, D, D]),
    "cipc_get_state": (ctypes.c_int, [P, D, D]),
    "cipc_set_kappa": (ctypes.c_int, [P, ctypes.c_double]),
    "cipc_step": (ctypes.c_int, [P, ctypes.POINTER(StepStats)]),
    "cipc_constraint_set": (ctypes.c_int, [P, D, I32, ctypes.c_int32, I32]),
    "cipc_barrier_energy": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, D]),
    "cipc_barrier_energy_noalloc": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, D]),
    "cipc_barrier_gradient": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, D]),
    "cipc_elastic_energy": (ctypes.c_int, [P, D, D]),
    "cipc_elastic_gradient": (ctypes.c_int, [P, D, D]),
    "cipc_hessian": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, I32, I32, D, ctypes.c_int32, I32]),
    "cipc_min_dist2": (ctypes.c_int, [P, D, I32, ctypes.c_int32, D, D]),
    "cipc_ccd_stepsize": (ctypes.c_int, [P, D, D, ctypes.c_double, D, D]),
    "cipc_set_friction": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_double, ctypes.c_int32]),
    "cipc_set_friction_components": (ctypes.c_int, [P, ctypes.c_int32, I32, D]),
    "cipc_get_tet_sizes": (ctypes.c_int, [P, I32]),
    "cipc_set_graphs": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_set_static_solve": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_set_output_folder": (ctypes.c_int, [P, ctypes.c_char_p]),
    "cipc_inversion_free_stepsize": (ctypes.c_int, [P, D, D, ctypes.c_double, D]),
    "cipc_set_strain_limit": (ctypes.c_int, [P, D, D, D]),
    "cipc_get_strain_limit": (ctypes.c_int, [P, D, D, D]),
    "cipc_set_fiber": (ctypes.c_int, [P, D, D]),
    "cipc_dirichlet_add_motion": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, D, D, D, ctypes.c_double]),
    "cipc_step_dirichlet": (ctypes.c_int, [P]),
    "cipc_write_obj": (ctypes.c_int, [P, ctypes.c_char_p]),
    "cipc_turn_dirichlet": (ctypes.c_int, [P]),
    "cipc_get_dirichlet_targets": (ctypes.c_int, [P, D]),
    "cipc_set_dirichlet_targets": (ctypes.c_int, [P, D]),
    "cipc_strain_limit_energy": (ctypes.c_int, [P, D, D, I32]),
    "cipc_friction_basis": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, I32, D, ctypes.c_int32, I32]),
    "cipc_friction_energy": (ctypes.c_int, [P, D, D, D]),
    "cipc_friction_gradient": (ctypes.c_int, [P, D, D, D]),
    "cipc_friction_hessian": (ctypes.c_int, [P, D, D, I32, I32, D, ctypes.c_int32, I32]),
    "cipc_linear_solve": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, D, D, ctypes.c_double, I32, D]),
    "cipc_set_precision": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_get_solver_info": (ctypes.c_int, [P, I32]),
    "cipc_set_linear_solver": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_set_coarse": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "cipc_get_coarse_info": (ctypes.c_int, [P, D]),
    "cipc_precond_num_params": (ctypes.c_int32, []),
    "cipc_precond_init": (ctypes.c_int, [P, ctypes.c_uint64]),
    "cipc_precond_set_weights": (ctypes.c_int, [P, F32, ctypes.c_int32]),
    "cipc_precond_get_weights": (ctypes.c_int, [P, F32, ctypes.c_int32]),
    "cipc_set_preconditioner": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32]),
    "cipc_precond_loss_grad": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_uint64, D, F32]),
    "cipc_precond_train_step": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_uint64,
                                               ctypes.c_double, D]),
    "cipc_precond_factor": (ctypes.c_int, [P, D, I32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, F32, I32]),
    "cipc_precond_get_adam": (ctypes.c_int, [P, F32, F32, I32, ctypes.c_int32]),
    "cipc_precond_set_adam": (ctypes.c_int, [P, F32, F32, ctypes.c_int32, ctypes.c_int32]),
    "cipc_precond_info": (ctypes.c_int, [P, I32]),
    "cipc_set_profiling": (ctypes.c_int, [P, ctypes.c_int32]),
    "cipc_get_profile": (ctypes.c_int, [P, D]),
}

_LIB = None


def load():
    """Load libcipc.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcipc.so not found at %s: build it first (__graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
