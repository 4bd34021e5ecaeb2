"""ctypes binding of ``libgzgrav.so`` (the C ABI declared in ``include/gz_grav.h``).

The library is built in-tree by ``gizmo_shared_memory_amd/csrc/Makefile`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgzgrav.so")

GZ_OK = 0
STATUS_NAMES = {0: "GZ_OK", 1: "GZ_ERR_INVALID", 2: "GZ_ERR_HIP", 3: "GZ_ERR_NOMEM",
                4: "GZ_ERR_UNSUPPORTED", 5: "GZ_ERR_ARENA", 6: "GZ_ERR_INTERNAL"}

MLP_FEATURES = 16
MLP_PARAMS = 5313

PREC_FP64 = 0
PREC_FP32_REL = 1
LAUNCH_QUEUE = 0
LAUNCH_PER_GROUP = 1
CRIT_REFERENCE = 0
CRIT_LEARNED = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class GzError(RuntimeError):
    """A non-zero gz_status, with the library's message."""

    def __init__(self, status: int, where: str, message: str):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {message}")


class gz_particles(C.Structure):
    _fields_ = [("n", C.c_size_t), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("m", C.c_void_p), ("soft", C.c_void_p), ("zeta", C.c_void_p),
                ("type", C.c_void_p)]


class gz_build_opts(C.Structure):
    _fields_ = [("randomize_seed", C.c_longlong), ("shift", C.c_double * 3),
                ("max_nodes", C.c_int), ("cluster_start", C.POINTER(C.c_size_t)),
                ("n_clusters", C.c_int), ("vx", C.c_void_p), ("vy", C.c_void_p),
                ("vz", C.c_void_p)]


class gz_build_times(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("bbox", "keys", "sort", "recurse", "links", "pack",
                                          "total")]


class gz_tree_info(C.Structure):
    _fields_ = [("n", C.c_longlong), ("n_live", C.c_longlong), ("nnodes", C.c_longlong),
                ("nleaves", C.c_longlong), ("root", C.c_longlong), ("depth", C.c_longlong),
                ("arena_grows", C.c_longlong), ("root_centre", C.c_double * 3),
                ("root_side", C.c_double), ("times", gz_build_times)]


class gz_walk_opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("criterion", C.c_int), ("group_size", C.c_int),
                ("launch", C.c_int), ("count", C.c_int), ("stack_limit", C.c_int),
                ("model", C.c_void_p), ("waves_per_cu", C.c_int)]


TUNE_RETUNE = 1
TUNE_MAX_POINTS = 16


class gz_tune_point(C.Structure):
    _fields_ = [("group_size", C.c_int), ("stack_limit", C.c_int), ("waves_per_cu", C.c_int),
                ("ms", C.c_double)]


class gz_tune_report(C.Structure):
    _fields_ = [("n_points", C.c_int), ("from_cache", C.c_int),
                ("points", gz_tune_point * TUNE_MAX_POINTS), ("best", gz_tune_point)]


class gz_walk_counts(C.Structure):
    _fields_ = [("nodes_visited", C.c_longlong), ("multipole", C.c_longlong),
                ("pairs", C.c_longlong), ("groups", C.c_longlong), ("items", C.c_longlong),
                ("walk_ms", C.c_double), ("prep_ms", C.c_double), ("max_stack", C.c_longlong),
                ("retries", C.c_longlong), ("fallback_groups", C.c_longlong)]


_PROTOS = {
    "gz_last_error": (C.c_char_p, []),
    "gz_status_string": (C.c_char_p, [C.c_int]),
    "gz_ctx_create": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "gz_ctx_destroy": (C.c_int, [C.c_void_p]),
    "gz_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "gz_tree_build": (C.c_int, [C.c_void_p, C.POINTER(gz_particles), C.POINTER(gz_build_opts),
                                C.POINTER(C.c_void_p)]),
    "gz_tree_free": (C.c_int, [C.c_void_p]),
    "gz_tree_get_info": (C.c_int, [C.c_void_p, C.POINTER(gz_tree_info)]),
    "gz_tree_download_nodes": (C.c_int, [C.c_void_p] + [_ip] * 7 + [_dp] * 11 + [_fp, _fp]),
    "gz_tree_download_particles": (C.c_int, [C.c_void_p, _up, _up, _ip, _u64p]),
    "gz_tree_device_arrays": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "gz_tree_update_sources": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles)]),
    "gz_tree_accel_grouped_tidal": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles),
                                              C.c_void_p, C.c_size_t, C.c_double, C.c_double,
                                              C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "gz_tree_ngb_search": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_size_t, C.c_double, C.c_void_p, C.c_void_p,
                                     C.c_size_t, C.POINTER(C.c_size_t)]),
    "gz_tree_init_hmax": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.c_size_t]),
    "gz_tree_raise_hmax": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_void_p]),
    "gz_tree_ngb_search_reverse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_size_t, C.c_double, C.c_void_p, C.c_void_p,
                                             C.c_size_t, C.POINTER(C.c_size_t)]),
    "gz_tree_potential": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles), C.c_void_p,
                                    C.c_size_t, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "gz_tree_kick_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_double]),
    "gz_tree_drift": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    "gz_tree_refresh_walk_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    "gz_tree_raise_soft": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_void_p]),
    "gz_tree_download_reuse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    "gz_tree_mass_events": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
                            + [C.c_void_p] * 7 + [C.c_double, C.c_double]),
    "gz_tree_shift_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
                            + [C.c_void_p] * 4),
    "gz_tree_raise_pad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.c_void_p]),
    "gz_tree_raise_vmax": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_void_p]),
    "gz_tree_download_pads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gz_tree_density": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles), C.c_void_p,
                                  C.c_size_t, C.c_double, C.c_double, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "gz_tree_accel_grouped": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles),
                                        C.c_void_p, C.c_size_t, C.c_double, C.c_double, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(gz_walk_opts), C.POINTER(gz_walk_counts)]),
    "gz_tree_accel": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles), C.c_void_p,
                                C.c_size_t, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
    "gz_walk_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles), C.c_void_p,
                                   C.c_size_t, C.c_double, C.c_double, C.c_void_p,
                                   C.POINTER(gz_walk_opts), C.c_int, C.c_int,
                                   C.POINTER(gz_walk_opts), C.POINTER(gz_tune_report)]),
    "gz_direct_sum": (C.c_int, [C.c_void_p, C.POINTER(gz_particles), C.c_void_p, C.c_size_t,
                                C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gz_direct_sum_ex": (C.c_int, [C.c_void_p, C.POINTER(gz_particles), C.c_void_p,
                                   C.c_size_t, C.c_double, C.c_int, C.c_double, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "gz_model_create": (C.c_int, [C.c_void_p, _fp, C.POINTER(C.c_void_p)]),
    "gz_model_free": (C.c_int, [C.c_void_p]),
    "gz_model_set_params": (C.c_int, [C.c_void_p, _fp]),
    "gz_model_get_params": (C.c_int, [C.c_void_p, _fp]),
    "gz_model_set_calibration": (C.c_int, [C.c_void_p, _fp, _fp, C.c_float, C.c_float,
                                           C.c_float, C.c_float]),
    "gz_model_set_input_map": (C.c_int, [C.c_void_p, _fp, _fp]),
    "gz_model_set_loss": (C.c_int, [C.c_void_p, C.c_float]),
    "gz_model_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "gz_model_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                C.c_void_p, _fp]),
    "gz_model_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                      C.c_float, _fp]),
    "gz_tree_group_features": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles),
                                         C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_size_t)]),
    "gz_walk_plan_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_particles),
                                      C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.c_int,
                                      C.POINTER(gz_walk_opts), C.POINTER(C.c_void_p)]),
    "gz_walk_plan_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gz_walk_plan_status": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gz_walk_counts)]),
    "gz_walk_plan_free": (C.c_int, [C.c_void_p]),
    "gz_graph_begin": (C.c_int, [C.c_void_p]),
    "gz_graph_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "gz_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gz_graph_free": (C.c_int, [C.c_void_p]),
    "gz_order_targets": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                   C.POINTER(C.c_size_t)]),
    "gz_order_targets_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.c_void_p, C.POINTER(C.c_size_t)]),
    "gz_aold_from_accel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_double, C.c_double, C.c_void_p]),
    "gz_scatter_accel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "gz_drift": (C.c_int, [C.c_void_p, C.c_size_t, C.c_double, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gz_kick": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
}

_lib = None


def exported_names() -> list[str]:
    return list(_PROTOS)


def load(path: str = LIB_PATH):
    """Load the library once. Raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; "
                           f"g.build()'` (the device path has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, where: str):
    if status != GZ_OK:
        msg = _lib.gz_last_error().decode() if _lib is not None else ""
        raise GzError(status, where, msg)
