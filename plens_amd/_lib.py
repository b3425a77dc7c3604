"""ctypes binding of libplens_gpu.so (include/plens_gpu.h and include/plens_host.h).

The library is loaded from this package's directory only. If it is missing the import of the
solver fails loudly: there is no CPU fallback for the GPU path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplens_gpu.so")

PLENS_OK = 0
PLENS_E_INVALID = -1
PLENS_E_HIP = -2
PLENS_E_POSITIVITY = -3
PLENS_E_NOMEM = -4
PLENS_E_NODEVICE = -5

FLUX = {"HLLC": 0, "Rusanov": 1, "AUSM+-up": 2, "Chandrashekhar": 3, "Rusanov-HLLC": 4,
        "Rusanov-AUSM+-up": 5, "Modified SW": 6, "KG": 7, "Rusanov-KG": 8, "IR": 9}
VOL_FLUX = {"Chandrashekhar": 0, "KG": 1, "IR": 2}
BC_PERIODIC, BC_UNIFORM_INFLOW, BC_OUTFLOW, BC_UNIFORM_TEMP_WALL = 0, 1, 2, 3
BC_FREE, BC_SYMMETRY, BC_EMPTY, BC_INSULATED_WALL, BC_ZPG_INFLOW, BC_VARYING_INFLOW = 4, 5, 6, 7, 8, 9
PREC_FP64, PREC_FP32_BULK = 0, 1
GRAD_CONSERVATIVE, GRAD_ENTROPY = 0, 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("degree", C.c_int32), ("Ja", _dp), ("detJ", _dp),
                ("face_normals", _dp), ("min_vertex_dist", _dp), ("face_nbr", _ip), ("face_bid", _ip),
                ("vertices", _dp), ("node_coords", _dp)]


class Physics(C.Structure):
    _fields_ = [("gamma", C.c_double), ("molecular_weight", C.c_double), ("prandtl", C.c_double),
                ("mu0", C.c_double), ("T0", C.c_double), ("S", C.c_double),
                ("inv_surf_flux", C.c_int32), ("precision", C.c_int32),
                ("threshold_factor", C.c_double), ("threshold_exp_factor", C.c_double),
                ("sharpness_factor", C.c_double), ("alpha_min", C.c_double), ("alpha_max", C.c_double),
                ("wall_blender_limit", C.c_double), ("gradient_variables", C.c_int32),
                ("inv_vol_flux", C.c_int32)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int32), ("p", C.c_double), ("T", C.c_double), ("vel", C.c_double * 3),
                ("p_expr", C.c_char_p), ("T_expr", C.c_char_p), ("vel_expr", C.c_char_p)]


class BoxParams(C.Structure):
    _fields_ = [("degree", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("y_stretch_ratio", C.c_double),
                ("perturb_frac", C.c_double), ("seed", C.c_uint64), ("periodic", C.c_int32 * 3),
                ("side_bid", C.c_int32 * 6), ("periodic_bid", C.c_int32 * 3),
                ("transpose_y_face_normals", C.c_int32)]


class ManifoldDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("axis_dir", C.c_double * 3), ("axis_point", C.c_double * 3),
                ("nose_center", C.c_double * 3)]


MANIFOLD_CYLINDER, MANIFOLD_NOSE_CYLINDER = 1, 2


# exported symbols (checked by the CPU tests against include/*.h)
GPU_SYMBOLS = [
    "plens_gpu_create", "plens_gpu_destroy", "plens_gpu_set_state", "plens_gpu_get_state",
    "plens_gpu_calc_rhs", "plens_gpu_calc_dt", "plens_gpu_rk4_step", "plens_gpu_rk4_step_dt", "plens_gpu_run",
    "plens_gpu_last_failure", "plens_gpu_set_tuning", "plens_gpu_kernel_times", "plens_gpu_strerror",
    "plens_gpu_last_run_ms", "plens_gpu_blended_cell_stages", "plens_gpu_run_profiled", "plens_gpu_stream_copy", "plens_gpu_time_kernel",
    "plens_gpu_set_local_stepping", "plens_gpu_ss_error", "plens_gpu_ss_error_noalloc", "plens_gpu_output_fields", "plens_gpu_write_checkpoint",
    "plens_gpu_read_checkpoint", "plens_gpu_set_time", "plens_gpu_set_expressions",
    "plens_gpu_batch_create", "plens_gpu_batch_destroy", "plens_gpu_batch_run", "plens_gpu_batch_get_state",
    "plens_gpu_batch_last_run_ms", "plens_gpu_batch_last_failure", "plens_gpu_batch_n_groups",
    "plens_gpu_batch_group_info",
]
HOST_SYMBOLS = [
    "plens_mesh_build_box", "plens_mesh_destroy", "plens_mesh_fill_desc", "plens_mesh_n_cells",
    "plens_mesh_degree", "plens_mesh_vertices", "plens_mesh_node_coords", "plens_mesh_Ja",
    "plens_mesh_detJ", "plens_mesh_face_normals", "plens_mesh_face_nbr", "plens_mesh_face_bid",
    "plens_mesh_min_vertex_dist", "plens_basis_1d", "plens_prim_to_cons", "plens_ic_density_wave",
    "plens_ic_normal_shock", "plens_ic_uniform", "plens_face_node", "plens_mesh_build_curved",
    "plens_mesh_mapping_degree", "plens_mesh_support_points", "plens_mesh_output_points", "plens_output_write",
    "plens_expr_eval",
]
MLP_SYMBOLS = [
    "plens_mlp_n_params", "plens_mlp_create", "plens_mlp_destroy", "plens_mlp_get_params",
    "plens_mlp_device_bf16_params", "plens_mlp_predict", "plens_mlp_grad", "plens_mlp_step", "plens_mlp_train",
    "plens_gpu_set_learned_indicator", "plens_gpu_indicator_stats", "plens_gpu_indicator_samples",
]

_lib = None


def load():
    """Load libplens_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m plens_amd.build` "
                           "(or __graft_entry__.build()); the GPU path has no fallback")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "plens_gpu_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Physics), C.POINTER(BC), C.c_int32,
                                       C.POINTER(vp)]),
        "plens_gpu_destroy": (None, [vp]),
        "plens_gpu_set_state": (C.c_int, [vp, _dp]),
        "plens_gpu_get_state": (C.c_int, [vp, _dp]),
        "plens_gpu_calc_rhs": (C.c_int, [vp, _dp, _dp, _dp]),
        "plens_gpu_calc_dt": (C.c_int, [vp, C.c_double, _dp, _dp]),
        "plens_gpu_rk4_step_dt": (C.c_int, [vp, C.c_double, C.c_double]),
        "plens_gpu_rk4_step": (C.c_int, [vp, C.c_double, _dp]),
        "plens_gpu_run": (C.c_int, [vp, C.c_int32, C.c_double, _dp]),
        "plens_gpu_run_profiled": (C.c_int, [vp, C.c_int32, C.c_double, _dp]),
        "plens_gpu_last_failure": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
        "plens_gpu_set_tuning": (C.c_int, [vp, C.c_int32]),
        "plens_gpu_kernel_times": (C.c_int, [vp, _dp, C.c_int32]),
        "plens_gpu_strerror": (C.c_char_p, [C.c_int]),
        "plens_gpu_last_run_ms": (C.c_int, [vp, _dp]),
        "plens_gpu_blended_cell_stages": (C.c_int, [vp, C.POINTER(C.c_int64), C.c_int32]),
        "plens_gpu_stream_copy": (C.c_int, [C.c_int64, C.c_int32, _dp]),
        "plens_gpu_time_kernel": (C.c_int, [vp, C.c_int32, C.c_int32, _dp]),
        "plens_gpu_set_local_stepping": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double]),
        "plens_gpu_ss_error": (C.c_int, [vp, _dp]),
        "plens_gpu_ss_error_noalloc": (C.c_int, [vp, _dp]),
        "plens_gpu_output_fields": (C.c_int, [vp, C.c_double, C.c_double, _dp, _dp]),
        "plens_gpu_set_time": (C.c_int, [vp, C.c_double]),
        "plens_gpu_set_expressions": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_char_p]),
        "plens_expr_eval": (C.c_int, [C.c_char_p, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
        "plens_gpu_write_checkpoint": (C.c_int, [vp, C.c_char_p, C.c_double, C.c_int64, C.c_int64]),
        "plens_gpu_read_checkpoint": (C.c_int, [vp, C.c_char_p, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "plens_gpu_batch_create": (C.c_int, [C.POINTER(vp), C.c_int32, C.c_char_p, C.POINTER(vp)]),
        "plens_gpu_batch_destroy": (None, [vp]),
        "plens_gpu_batch_run": (C.c_int, [vp, C.c_int32, _dp, _dp]),
        "plens_gpu_batch_get_state": (C.c_int, [vp, C.c_int32, _dp]),
        "plens_gpu_batch_last_run_ms": (C.c_int, [vp, _dp]),
        "plens_gpu_batch_last_failure": (C.c_int, [vp, _ip, C.POINTER(C.c_int64), _ip]),
        "plens_gpu_batch_n_groups": (C.c_int, [vp]),
        "plens_gpu_batch_group_info": (C.c_int, [vp, C.c_int32, _ip, _ip, C.POINTER(C.c_int64), _ip, _ip]),
        "plens_mesh_build_box": (C.c_int, [C.POINTER(BoxParams), C.POINTER(vp)]),
        "plens_mesh_destroy": (None, [vp]),
        "plens_mesh_fill_desc": (C.c_int, [vp, C.POINTER(MeshDesc)]),
        "plens_mesh_n_cells": (C.c_int32, [vp]),
        "plens_mesh_degree": (C.c_int32, [vp]),
        "plens_mesh_vertices": (_dp, [vp]),
        "plens_mesh_node_coords": (_dp, [vp]),
        "plens_mesh_Ja": (_dp, [vp]),
        "plens_mesh_detJ": (_dp, [vp]),
        "plens_mesh_face_normals": (_dp, [vp]),
        "plens_mesh_face_nbr": (_ip, [vp]),
        "plens_mesh_face_bid": (_ip, [vp]),
        "plens_mesh_min_vertex_dist": (_dp, [vp]),
        "plens_mesh_build_curved": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _ip, _ip, _ip,
                                              C.POINTER(ManifoldDesc), C.c_int32, C.POINTER(vp)]),
        "plens_mesh_mapping_degree": (C.c_int32, [vp]),
        "plens_mesh_support_points": (_dp, [vp]),
        "plens_mesh_output_points": (C.c_int, [vp, _dp]),
        "plens_output_write": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_int32, C.c_double, C.c_double, _dp, _dp,
                                         _dp, C.c_int32]),
        "plens_basis_1d": (C.c_int, [C.c_int32, _dp, _dp, _dp, _dp]),
        "plens_prim_to_cons": (None, [C.c_double, C.c_double, _dp, C.c_double, _dp]),
        "plens_ic_density_wave": (C.c_int, [vp, C.c_double, C.c_double, _dp, C.c_double, C.c_double,
                                            C.c_double, C.c_uint64, _ip, _dp]),
        "plens_ic_normal_shock": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_uint64, _dp]),
        "plens_ic_uniform": (C.c_int, [vp, C.c_double, C.c_double, _dp, C.c_double, _dp]),
        "plens_face_node": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32]),
        "plens_mlp_n_params": (C.c_int32, []),
        "plens_mlp_create": (C.c_int, [_fp, C.POINTER(vp)]),
        "plens_mlp_destroy": (C.c_int, [vp]),
        "plens_mlp_get_params": (C.c_int, [vp, _fp]),
        "plens_mlp_device_bf16_params": (C.c_void_p, [vp]),
        "plens_mlp_predict": (C.c_int, [vp, _fp, C.c_int64, _fp]),
        "plens_mlp_grad": (C.c_int, [vp, _fp, _fp, C.c_int32, _fp, _fp]),
        "plens_mlp_step": (C.c_int, [vp, _fp, _fp, C.c_int32, C.c_float]),
        "plens_mlp_train": (C.c_int, [vp, _fp, _fp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]),
        "plens_gpu_set_learned_indicator": (C.c_int, [vp, C.c_void_p, C.c_double, C.c_double]),
        "plens_gpu_indicator_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int32]),
        "plens_gpu_indicator_samples": (C.c_int, [vp, _fp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def strerror(code):
    return load().plens_gpu_strerror(code).decode()


class PlensError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})")


def check(code, what):
    if code != PLENS_OK:
        raise PlensError(code, what)
