"""ctypes binding of libcmg.so (include/cmg.h).

The library is built in-tree (cardiomechanics_amd/libcmg.so) by
cardiomechanics_amd/build.py.  Loading fails loudly when it is missing: there is
no fallback implementation of the hot path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmg.so")

SUCCESS, NOTHING_DONE, REPEAT, FAILED = 2, 0, -2, -3
INFINITIVE, NOT_A_NUMBER, CORRUPT_ELEMENT, SYSTEM_WRONG_STATE = -4, -5, -6, -8
ERR_INVALID, ERR_HIP, ERR_NOMEM = -100, -101, -102
PREC_FP64, PREC_MIXED = 0, 1

# every symbol include/cmg.h declares
EXPORTED = [
    "cmg_default_options", "cmg_last_error", "cmg_device_count", "cmg_create", "cmg_destroy", "cmg_sizes",
    "cmg_pattern", "cmg_set_pressures", "cmg_synchronize", "cmg_stream", "cmg_residual", "cmg_tangent",
    "cmg_element_residuals", "cmg_element_tangents", "cmg_surface_forces", "cmg_cavity_volumes",
    "cmg_mass_matrix", "cmg_element_export", "cmg_spmv_tuning", "cmg_residual_device", "cmg_residual_tangent_device", "cmg_element_kernel_device",
    "cmg_spmv", "cmg_spmv_layout", "cmg_linear_solve", "cmg_set_state", "cmg_get_state", "cmg_set_initial_guess",
    "cmg_static_step", "cmg_newmark_step", "cmg_newmark_energies", "cmg_newmark_solve", "cmg_trial_volumes", "cmg_commit_step",
    "cmg_copy_device", "cmg_residual_device_noalloc",
    "cmg_set_land17_parameters", "cmg_get_land17_state", "cmg_set_land17_state",
    "cmg_batch_create", "cmg_batch_destroy", "cmg_batch_configure", "cmg_batch_info_get", "cmg_batch_static_step",
    "cmg_batch_get_nodes", "cmg_batch_set_nodes", "cmg_batch_residual", "cmg_batch_cavity_volumes",
    "cmg_mlp_num_weights", "cmg_mlp_create", "cmg_mlp_destroy", "cmg_mlp_get_weights", "cmg_mlp_set_weights",
    "cmg_mlp_set_scales", "cmg_mlp_features", "cmg_mlp_targets", "cmg_mlp_forward", "cmg_mlp_loss_grad",
    "cmg_mlp_adam_step", "cmg_mlp_propose_guess",
]


class CmgMaterial(ctypes.Structure):
    _fields_ = [("law", ctypes.c_int32), ("ignore_corrupt", ctypes.c_int32), ("tension_fn", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("density", ctypes.c_double), ("params", ctypes.c_double * 12),
                ("tension_max", ctypes.c_double), ("fn", ctypes.c_double * 12)]


class CmgMesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_elements", ctypes.c_int64), ("n_surfaces", ctypes.c_int64),
                ("nodes_per_element", ctypes.c_int32), ("nodes_per_surface", ctypes.c_int32),
                ("n_groups", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("ref_coords", ctypes.c_void_p), ("elements", ctypes.c_void_p), ("bases", ctypes.c_void_p),
                ("material_index", ctypes.c_void_p), ("activation_time", ctypes.c_void_p),
                ("bc", ctypes.c_void_p), ("surfaces", ctypes.c_void_p), ("surface_group", ctypes.c_void_p)]


class CmgOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("transpose_tangent", ctypes.c_int32),
                ("pressure_tangent", ctypes.c_int32), ("consistent_mass", ctypes.c_int32),
                ("newton_maxit", ctypes.c_int32), ("krylov_restart", ctypes.c_int32),
                ("krylov_maxit", ctypes.c_int32), ("device", ctypes.c_int32),
                ("element_variant", ctypes.c_int32), ("krylov_method", ctypes.c_int32),
                ("newton_atol", ctypes.c_double), ("newton_rtol", ctypes.c_double),
                ("newton_stol", ctypes.c_double), ("krylov_rtol", ctypes.c_double),
                ("newmark_beta", ctypes.c_double), ("newmark_gamma", ctypes.c_double),
                ("fd_epsilon", ctypes.c_double), ("rayleigh_alpha", ctypes.c_double),
                ("rayleigh_beta", ctypes.c_double), ("preconditioner", ctypes.c_int32),
                ("precond_reuse", ctypes.c_int32), ("krylov_eta_max", ctypes.c_double),
                ("newton_stall", ctypes.c_int32), ("spmv_choice", ctypes.c_int32)]


class CmgSolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double)]


class CmgStepStats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int32), ("krylov_iterations", ctypes.c_int32),
                ("converged_reason", ctypes.c_int32), ("corrupt_element", ctypes.c_int32),
                ("residual_norm", ctypes.c_double), ("initial_residual_norm", ctypes.c_double),
                ("step_norm", ctypes.c_double), ("wall_seconds", ctypes.c_double),
                ("tangent_seconds", ctypes.c_double), ("krylov_seconds", ctypes.c_double),
                ("guess_source", ctypes.c_int32), ("reserved_s", ctypes.c_int32),
                ("guess_residual_ratio", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CmgBatchInfo(ctypes.Structure):
    _fields_ = [("n_sets", ctypes.c_int32), ("element_variant", ctypes.c_int32), ("spmv_lanes", ctypes.c_int32),
                ("cg_chunk", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("tuned", ctypes.c_int32),
                ("preconditioner", ctypes.c_int32), ("reserved", ctypes.c_int32), ("element_ms", ctypes.c_double * 8),
                ("spmv_ms", ctypes.c_double * 3), ("tune_seconds", ctypes.c_double)]

    def as_dict(self):
        return {"n_sets": self.n_sets, "element_variant": self.element_variant, "spmv_lanes": self.spmv_lanes,
                "cg_chunk": self.cg_chunk, "use_graph": self.use_graph, "tuned": self.tuned,
                "preconditioner": "vertex-patch Schwarz" if self.preconditioner == 2 else "block Jacobi",
                "element_ms": list(self.element_ms), "spmv_ms": list(self.spmv_ms)}


_lib = None


def load():
    """Load libcmg.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cardiomechanics_amd.build` "
                           "(the mechanics path has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "cmg_default_options": (None, [P]),
        "cmg_last_error": (ctypes.c_char_p, []),
        "cmg_device_count": (ctypes.c_int, []),
        "cmg_create": (ctypes.c_int, [P, P, i32, P, ctypes.POINTER(P)]),
        "cmg_destroy": (None, [P]),
        "cmg_sizes": (ctypes.c_int, [P, P, P, P, P]),
        "cmg_pattern": (ctypes.c_int, [P, P, P]),
        "cmg_set_pressures": (ctypes.c_int, [P, P]),
        "cmg_synchronize": (ctypes.c_int, [P]),
        "cmg_stream": (P, [P]),
        "cmg_residual": (ctypes.c_int, [P, P, dbl, i32, P]),
        "cmg_tangent": (ctypes.c_int, [P, P, dbl, i32, P]),
        "cmg_element_residuals": (ctypes.c_int, [P, P, dbl, P]),
        "cmg_element_tangents": (ctypes.c_int, [P, P, dbl, P]),
        "cmg_surface_forces": (ctypes.c_int, [P, P, P, P]),
        "cmg_cavity_volumes": (ctypes.c_int, [P, P, P]),
        "cmg_mass_matrix": (ctypes.c_int, [P, P]),
        "cmg_element_export": (ctypes.c_int, [P, P, P]),
        "cmg_spmv_tuning": (ctypes.c_int, [P, P, P, P]),
        "cmg_residual_device": (ctypes.c_int, [P, P, dbl, i32, P]),
        "cmg_residual_device_noalloc": (ctypes.c_int, [P, P, dbl, i32, P]),
        "cmg_residual_tangent_device": (ctypes.c_int, [P, P, dbl, i32, P, P]),
        "cmg_element_kernel_device": (ctypes.c_int, [P, P, dbl, i32]),
        "cmg_spmv": (ctypes.c_int, [P, P, P, P]),
        "cmg_spmv_layout": (ctypes.c_int, [P, P, P, P, i32, i32]),
        "cmg_linear_solve": (ctypes.c_int, [P, P, P, P, dbl, P]),
        "cmg_set_state": (ctypes.c_int, [P, P, P, P]),
        "cmg_get_state": (ctypes.c_int, [P, P, P, P, P]),
        "cmg_set_initial_guess": (ctypes.c_int, [P, P]),
        "cmg_set_land17_parameters": (ctypes.c_int, [P, i32, P]),
        "cmg_get_land17_state": (ctypes.c_int, [P, P, P]),
        "cmg_set_land17_state": (ctypes.c_int, [P, P, dbl]),
        "cmg_static_step": (ctypes.c_int, [P, dbl, dbl, P]),
        "cmg_newmark_step": (ctypes.c_int, [P, dbl, dbl, P]),
        "cmg_newmark_energies": (ctypes.c_int, [P, P, P]),
        "cmg_newmark_solve": (ctypes.c_int, [P, dbl, dbl, ctypes.c_int32, P]),
        "cmg_trial_volumes": (ctypes.c_int, [P, P]),
        "cmg_commit_step": (ctypes.c_int, [P]),
        "cmg_copy_device": (ctypes.c_int, [P, P, P, i64]),
        "cmg_batch_create": (ctypes.c_int, [P, i32, P, ctypes.POINTER(P)]),
        "cmg_batch_destroy": (ctypes.c_int, [P]),
        "cmg_batch_configure": (ctypes.c_int, [P, i32, i32, i32, i32]),
        "cmg_batch_info_get": (ctypes.c_int, [P, P]),
        "cmg_batch_static_step": (ctypes.c_int, [P, dbl, dbl, P]),
        "cmg_batch_get_nodes": (ctypes.c_int, [P, P]),
        "cmg_batch_set_nodes": (ctypes.c_int, [P, P]),
        "cmg_batch_residual": (ctypes.c_int, [P, P, dbl, P]),
        "cmg_batch_cavity_volumes": (ctypes.c_int, [P, P]),
        "cmg_mlp_num_weights": (i32, [i32]),
        "cmg_mlp_create": (ctypes.c_int, [P, i32, ctypes.c_uint64, ctypes.POINTER(P)]),
        "cmg_mlp_destroy": (ctypes.c_int, [P]),
        "cmg_mlp_get_weights": (ctypes.c_int, [P, P]),
        "cmg_mlp_set_weights": (ctypes.c_int, [P, P]),
        "cmg_mlp_set_scales": (ctypes.c_int, [P, dbl, dbl, dbl]),
        "cmg_mlp_features": (ctypes.c_int, [P, dbl, dbl, P]),
        "cmg_mlp_targets": (ctypes.c_int, [P, P]),
        "cmg_mlp_forward": (ctypes.c_int, [P, P, i64, i32, P]),
        "cmg_mlp_loss_grad": (ctypes.c_int, [P, P, P, i64, P, P]),
        "cmg_mlp_adam_step": (ctypes.c_int, [P, ctypes.c_float]),
        "cmg_mlp_propose_guess": (ctypes.c_int, [P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().cmg_last_error().decode(errors="replace")
