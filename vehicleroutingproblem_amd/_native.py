"""ctypes binding of the C ABI in include/aco_hip.h and include/aco_host.h.

The shared library is built in-tree (``vehicleroutingproblem_amd/build/libaco_hip.so``)
by ``__graft_entry__.build()`` or ``make -C vehicleroutingproblem_amd/csrc``.
There is no fallback: if the library is missing, importing the solver fails.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "build")
LIB_PATH = os.path.join(BUILD_DIR, "libaco_hip.so")
CLI_PATH = os.path.join(BUILD_DIR, "aco_vrp_hip.out")

ACO_HIP_OK = 0
ACO_HIP_ERR_INVALID_INPUT = 1
ACO_HIP_ERR_ALLOCATION = 2
ACO_HIP_ERR_NO_SOLUTION = 3
ACO_HIP_ERR_BACKEND = 4

KNN_EXACT = 0
KNN_REF_F32 = 1

DIST_NINT = 0
DIST_REAL = 1
REAL_COST_UNIT = 1.0 / 8388608.0  # ACO_HIP_REAL_COST_UNIT: real length per cost unit

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
u16p = C.POINTER(C.c_uint16)


class AcoHipInstance(C.Structure):
    _fields_ = [("n", C.c_int), ("K", C.c_int), ("capacity", C.c_int),
                ("x", i32p), ("y", i32p), ("demand", i32p), ("seed", C.c_uint32),
                ("cand_k", C.c_int)]


class AcoHipBucketInfo(C.Structure):
    _fields_ = [("n1_min", C.c_int), ("n1_max", C.c_int), ("instances", C.c_int),
                ("colonies", C.c_int), ("ants_per_group", C.c_int), ("stage_lds", C.c_int),
                ("lds_bytes", C.c_int), ("tuned_ms", C.c_float), ("ants_per_wave", C.c_int),
                ("state_bytes", C.c_int), ("lanes", C.c_int)]


class AcoHipParams(C.Structure):
    _fields_ = [("cand_k", C.c_int), ("ants", C.c_int), ("colonies", C.c_int),
                ("alpha", C.c_float), ("beta", C.c_float), ("rho", C.c_float),
                ("tau0", C.c_float), ("Q", C.c_float), ("depot_close_weight", C.c_float),
                ("knn_key", C.c_int), ("ants_per_group", C.c_int), ("stage_lds", C.c_int),
                ("fallback_matrix", C.c_int), ("dist_mode", C.c_int)]


class AcoHipColonyState(C.Structure):
    _fields_ = [("epoch", C.c_int32), ("iter_since_best", C.c_int32), ("halted", C.c_int32),
                ("found", C.c_int32), ("best_ant", C.c_int32), ("improved", C.c_int32),
                ("best_len", C.c_int32), ("feasible_ants", C.c_int32),
                ("iter_best_cost", C.c_int64), ("global_best_cost", C.c_int64),
                ("dep_iter", C.c_float), ("dep_best", C.c_float)]


class AcoHipRunConfig(C.Structure):
    _fields_ = [("timeout_seconds", C.c_double), ("stagnation_epochs", C.c_int),
                ("min_rel_improvement", C.c_double), ("fixed_epochs", C.c_int),
                ("max_epochs", C.c_int), ("progress_interval_seconds", C.c_double),
                ("print_progress", C.c_int), ("chunk_epochs", C.c_int), ("use_graph", C.c_int)]


class AcoHipRunStats(C.Structure):
    _fields_ = [("epochs", C.c_int), ("seconds", C.c_double), ("ant_tours_per_s", C.c_double)]


class AcoHipTiming(C.Structure):
    _fields_ = [("construct_ms", C.c_double), ("construct_launches", C.c_int),
                ("best_ms", C.c_double), ("best_launches", C.c_int),
                ("evaporate_ms", C.c_double), ("evaporate_launches", C.c_int),
                ("deposit_ms", C.c_double), ("deposit_launches", C.c_int),
                ("knn_ms", C.c_double), ("knn_launches", C.c_int)]


class AcoHipStats(C.Structure):
    _fields_ = [("decisions", C.c_uint64), ("fallbacks", C.c_uint64),
                ("fallback_cycles", C.c_uint64), ("ant_cycles", C.c_uint64),
                ("ants", C.c_uint64), ("staging_cycles", C.c_uint64),
                ("updates", C.c_uint64), ("evap_ticks", C.c_uint64),
                ("select_ticks", C.c_uint64), ("deposit_ticks", C.c_uint64),
                ("wall_clock_hz", C.c_uint64)]


class VrpInstance(C.Structure):
    _fields_ = [("n", C.c_int), ("K", C.c_int), ("capacity", C.c_int),
                ("x", i32p), ("y", i32p), ("demand", i32p), ("unit_demand", C.c_int),
                ("name", C.c_char * 128)]


class AcoTauTables(C.Structure):
    _fields_ = [("log_tau_min", C.c_float), ("log_tau_step", C.c_float),
                ("q_tau0", C.c_uint8), ("q_evap", C.c_uint8),
                ("deq", C.c_float * 256), ("thr", C.c_float * 255), ("taupow", C.c_float * 256)]


ctx_p = C.c_void_p

# name -> (restype, argtypes): every symbol declared in include/aco_hip.h and
# the host surface of include/aco_host.h used from Python.
HIP_SYMBOLS = {
    "aco_hip_create": (C.c_int, [C.POINTER(ctx_p), C.c_int]),
    "aco_hip_destroy": (None, [ctx_p]),
    "aco_hip_last_error": (C.c_char_p, [ctx_p]),
    "aco_hip_sync": (C.c_int, [ctx_p]),
    "aco_hip_load": (C.c_int, [ctx_p, C.POINTER(AcoHipInstance), C.c_int,
                               C.POINTER(AcoHipParams)]),
    "aco_hip_num_colonies": (C.c_int, [ctx_p]),
    "aco_hip_get_shape": (C.c_int, [ctx_p, i32p, i32p, i32p, i32p, i32p]),
    "aco_hip_build_candidates": (C.c_int, [ctx_p]),
    "aco_hip_init_tau": (C.c_int, [ctx_p]),
    "aco_hip_construct": (C.c_int, [ctx_p]),
    "aco_hip_select_best": (C.c_int, [ctx_p]),
    "aco_hip_evaporate": (C.c_int, [ctx_p]),
    "aco_hip_deposit_iter": (C.c_int, [ctx_p]),
    "aco_hip_deposit_best": (C.c_int, [ctx_p]),
    "aco_hip_epochs": (C.c_int, [ctx_p, C.c_int]),
    "aco_hip_run": (C.c_int, [ctx_p, C.POINTER(AcoHipRunConfig), C.POINTER(AcoHipRunStats)]),
    "aco_hip_get_candidates": (C.c_int, [ctx_p, C.c_int, i32p, i32p, f32p]),
    "aco_hip_set_candidate_eta": (C.c_int, [ctx_p, C.c_int, f32p]),
    "aco_hip_get_tau": (C.c_int, [ctx_p, C.c_int, u8p, u8p]),
    "aco_hip_set_tau": (C.c_int, [ctx_p, C.c_int, u8p]),
    "aco_hip_get_scores": (C.c_int, [ctx_p, C.c_int, f32p]),
    "aco_hip_get_fallback_weights": (C.c_int, [ctx_p, C.c_int, f32p]),
    "aco_hip_get_tours": (C.c_int, [ctx_p, C.c_int, i32p, i32p, i64p, i32p]),
    "aco_hip_get_colony": (C.c_int, [ctx_p, C.c_int, C.POINTER(AcoHipColonyState)]),
    "aco_hip_set_epoch": (C.c_int, [ctx_p, C.c_int, C.c_int]),
    "aco_hip_get_best": (C.c_int, [ctx_p, C.c_int, i32p, i32p, i64p]),
    "aco_hip_deposit_tour": (C.c_int, [ctx_p, C.c_int, i32p, C.c_int, C.c_float]),
    "aco_hip_route_costs": (C.c_int, [ctx_p, C.c_int, i32p, i32p, C.c_int, C.c_int, i64p]),
    "aco_hip_route_costs_noalloc": (C.c_int, [ctx_p, C.c_int, i32p, i32p, C.c_int, C.c_int,
                                              i64p]),
    "aco_hip_get_tables": (C.c_int, [ctx_p, f32p, f32p, f32p, f32p, i32p, u8p, u8p]),
    "aco_hip_set_stats": (C.c_int, [ctx_p, C.c_int]),
    "aco_hip_get_stats": (C.c_int, [ctx_p, C.c_int, C.c_void_p]),
    "aco_hip_set_timing": (C.c_int, [ctx_p, C.c_int]),
    "aco_hip_get_timing": (C.c_int, [ctx_p, C.POINTER(AcoHipTiming)]),
    "aco_vrp_hip_with_capacity": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, i32p, i32p, i32p,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_uint, C.c_int, i32p, i32p, i64p]),
    "aco_hip_num_buckets": (C.c_int, [ctx_p]),
    "aco_hip_get_bucket": (C.c_int, [ctx_p, C.c_int, C.POINTER(AcoHipBucketInfo)]),
    "aco_hip_set_bucket_shape": (C.c_int, [ctx_p, C.c_int, C.c_int, C.c_int]),
    "aco_hip_set_bucket_layout": (C.c_int, [ctx_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "aco_hip_autotune": (C.c_int, [ctx_p, C.c_int, C.POINTER(AcoHipBucketInfo), C.c_int]),
    "aco_hip_prior_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint32]),
    "aco_hip_prior_destroy": (None, [C.c_void_p]),
    "aco_hip_prior_last_error": (C.c_char_p, [C.c_void_p]),
    "aco_hip_prior_get_params": (C.c_int, [C.c_void_p, f32p]),
    "aco_hip_prior_set_params": (C.c_int, [C.c_void_p, f32p]),
    "aco_hip_edge_features": (C.c_int, [ctx_p, C.c_int, u16p]),
    "aco_hip_prior_forward": (C.c_int, [C.c_void_p, u16p, C.c_int64, f32p]),
    "aco_hip_prior_gradient": (C.c_int, [C.c_void_p, u16p, u8p, C.c_int64, f32p, f32p]),
    "aco_hip_prior_add_instance": (C.c_int, [C.c_void_p, ctx_p, C.c_int, i32p, C.c_int]),
    "aco_hip_prior_clear": (C.c_int, [C.c_void_p]),
    "aco_hip_prior_size": (C.c_int64, [C.c_void_p]),
    "aco_hip_prior_train": (C.c_int, [C.c_void_p, C.c_int, C.c_float, f32p]),
    "aco_hip_prior_apply": (C.c_int, [C.c_void_p, ctx_p, C.c_int, C.c_float]),
}

# include/aco_hip_pp.h: device buffers are passed as integer addresses
# (c_void_p), the stream as a hipStream_t handle or None.
vp = C.c_void_p
PP_SYMBOLS = {
    "aco_hip_pp_init_tau": (C.c_int, [vp, C.c_int, C.c_uint8, vp]),
    "aco_hip_pp_evaporate_tau": (C.c_int, [vp, C.c_int, C.c_uint8, C.c_uint8, vp]),
    "aco_hip_pp_build_candidate_lists": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                   vp, vp, vp]),
    "aco_hip_pp_eta_table": (C.c_int, [C.c_int, C.c_float, f32p, i32p]),
    "aco_hip_pp_deposit_solution": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                                              C.c_float, C.c_int, vp]),
    "aco_hip_pp_deposit_flat_solution": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_float,
                                                   vp]),
    "aco_hip_pp_deposit_tour": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_float, vp]),
    "aco_hip_pp_route_costs": (C.c_int, [vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int,
                                         vp, vp]),
    "aco_hip_pp_sync": (C.c_int, [vp]),
    "aco_hip_pp_last_error": (C.c_char_p, []),
}

HOST_SYMBOLS = {
    "vrp_read": (C.c_int, [C.c_char_p, C.POINTER(VrpInstance), C.c_int]),
    "vrp_write": (C.c_int, [C.c_char_p, C.POINTER(VrpInstance)]),
    "vrp_free": (None, [C.POINTER(VrpInstance)]),
    "vrp_alloc": (C.c_int, [C.POINTER(VrpInstance), C.c_int]),
    "vrp_generate_uniform": (C.c_int, [C.POINTER(VrpInstance), C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_uint32]),
    "vrp_generate_clustered": (C.c_int, [C.POINTER(VrpInstance), C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_uint32]),
    "vrp_capacity_for": (C.c_int, [C.c_longlong, C.c_int, C.c_int]),
    "vrp_nint_dist": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "vrp_real_dist": (C.c_double, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "aco_tau_tables_build": (None, [C.POINTER(AcoTauTables), C.c_float, C.c_float, C.c_float]),
    "aco_eta_table_build": (None, [f32p, C.c_int, C.c_float]),
    "aco_eta_table_build_real": (None, [f32p, C.c_int, C.c_float]),
    "aco_real_dist_batch": (None, [C.POINTER(C.c_uint32), C.c_int64, f32p]),
    "solution_cost_real_units": (C.c_int64, [C.c_void_p, C.c_void_p]),
    "aco_tau_quantize": (C.c_uint8, [C.POINTER(AcoTauTables), C.c_float]),
    "aco_parse_min_rel_improvement_percent": (C.c_double, [C.c_char_p, C.c_double]),
    "aco_is_significant_improvement": (C.c_int, [C.c_double, C.c_double, C.c_double]),
    "cli_parse_int_arg": (C.c_int, [C.c_char_p, i32p]),
    "cli_parse_uint_arg": (C.c_uint, [C.c_char_p, i32p]),
    "aco_status_string": (C.c_char_p, [C.c_int]),
    "aco_philox_host": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32)]),
    "aco_runtime_config_load_env": (None, [C.c_void_p]),
}


class AcoRuntimeConfig(C.Structure):
    _fields_ = [("timeout_seconds", C.c_double), ("stagnation_epochs", C.c_int),
                ("min_rel_improvement", C.c_double), ("fixed_epochs", C.c_int),
                ("candidate_k", C.c_int), ("ants", C.c_int), ("seed", C.c_uint),
                ("progress_interval_seconds", C.c_double)]

_lib = None


def load() -> C.CDLL:
    """Load libaco_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
            "`make -C vehicleroutingproblem_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for table in (HIP_SYMBOLS, PP_SYMBOLS, HOST_SYMBOLS):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (no copy)."""
    return arr.ctypes.data_as(C.POINTER(ctype))
