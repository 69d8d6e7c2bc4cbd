"""ctypes binding of libautoplace_gp.so (include/autoplace_gp.h).

Loading fails loudly: there is no CPU fallback in this package. Build the library first
(`python -m autoplace_amd.build_lib` or `__graft_entry__.build()`).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libautoplace_gp.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double

AP_OK, AP_ERR_ARG, AP_ERR_HIP, AP_ERR_UNSUPPORTED = 0, 1, 2, 3
AP_POISSON_AUTO, AP_POISSON_GEMM, AP_POISSON_FFT = 0, 1, 2
AP_AUTOTUNE_ONLINE = 1
PROF_CLASSES = ("nesterov", "wa", "wa_cells", "density_bin", "density_tiles", "overflow_tiles", "poisson", "gather",
                "control")

_SIGNATURES = {
    "ap_last_error": ([], ctypes.c_char_p),
    "ap_version": ([], I),
    "ap_cached_device_bytes": ([], ctypes.c_longlong),
    "ap_release_cached_memory": ([], I),
    "ap_wa_wirelength": ([P, P, P, P, P, P, I, I, P, P, P, P, I, P], I),
    "ap_wa_plan_create": ([P, P, P, I, I, ctypes.POINTER(P)], I),
    "ap_wa_plan_destroy": ([P], I),
    "ap_wa_wirelength_planned": ([P, P, P, P, P, P, P, I, I, P, P, P, P, I, P], I),
    "ap_density_map": ([P, P, P, I, I, F, F, I, I, P, P, F, P, P, P], I),
    "ap_density_map_work_bytes": ([I, I, I, I], ctypes.c_longlong),
    "ap_density_map_ws": ([P, P, P, I, I, F, F, I, I, P, F, P, P, P, ctypes.c_longlong, P], I),
    "ap_nesterov_step": ([P, P, P, P, P, P, P, I, P, P, P], I),
    "ap_density_grad": ([P, P, P, P, P, P, P, I, I, P, P, F, F, I, I, P, P], I),
    "ap_control": ([P, P, P, P, P, P, P, P, P], I),
    "ap_select_step": ([P, P, P, P, P, P, P, P, I, P], I),
    "ap_net_hpwl": ([P, P, P, P, I, I, P, P, P], I),
    "ap_net_hpwl_noalloc": ([P, P, P, P, I, I, P, P], I),
    "ap_poisson_create": ([I, I, D, D, I, ctypes.POINTER(P)], I),
    "ap_poisson_destroy": ([P], I),
    "ap_poisson_method": ([P], I),
    "ap_poisson_solve": ([P, P, I, P, P, I, P], I),
    "ap_poisson_energy": ([P, P, I, P, P], I),
    "ap_level_create": ([P, ctypes.POINTER(P)], I),
    "ap_level_destroy": ([P], I),
    "ap_level_run": ([P, I, I, P], I),
    "ap_level_prepare": ([P, I], I),
    "ap_level_start_begin": ([P, D, I, I], I),
    "ap_level_start_run": ([P, I, I, P], I),
    "ap_level_start_state": ([P, ctypes.POINTER(D), ctypes.POINTER(D), I], I),
    "ap_level_last_ms": ([P, ctypes.POINTER(F)], I),
    "ap_level_profile": ([P, I, ctypes.POINTER(D), P], I),
    "ap_level_time_class": ([P, I], I),
    "ap_level_class_ms": ([P, ctypes.POINTER(D), ctypes.POINTER(I)], I),
    "ap_level_objective": ([P, P, P, P, P], I),
    "ap_level_density": ([P, P, P], I),
    "ap_mlp_param_count": ([], I),
    "ap_mlp_work_floats": ([I], ctypes.c_longlong),
    "ap_mlp_forward": ([P, P, I, P, P], I),
    "ap_mlp_grad": ([P, P, P, P, I, P, P, P, P], I),
    "ap_mlp_adam": ([P, P, P, P, I, F, F, F, F, P], I),
    "ap_level_set_controller": ([P, P, F, ctypes.c_ulonglong, I, F, P, I, I], I),
    "ap_level_controller_state": ([P, ctypes.POINTER(D)], I),
    "ap_level_autotune": ([P, I, I, P], I),
    "ap_level_set_tile_layout": ([P, I, I, I, I], I),
    "ap_level_set_bucket_room": ([P, I, I], I),
    "ap_level_launch_shape": ([P, P], I),
    "ap_level_set_launch_shape": ([P, I, I, I], I),
    "ap_level_renumber": ([P, I, I, I, I, ctypes.c_double, ctypes.c_double] + [P] * 3 + [I, I] + [P] * 7, I),
    "ap_level_time_every": ([P, I], I),
    "ap_level_bucket_spills": ([P, P], I),
    "ap_level_set_tile_split": ([P, I, I], I),
    "ap_level_tile_splits": ([P, P], I),
    "ap_batch_create": ([P, I, ctypes.POINTER(P)], I),
    "ap_batch_destroy": ([P], I),
    "ap_batch_prepare": ([P, I], I),
    "ap_batch_run": ([P, I, I, P], I),
    "ap_batch_last_ms": ([P, ctypes.POINTER(F)], I),
    "ap_batch_autotune": ([P, I, P], I),
    "ap_batch_tile_layout": ([P, P], I),
    "ap_cluster_create": ([I, P, P, I, P, I, P, ctypes.POINTER(P)], I),
    "ap_cluster_round": ([P, P, F, ctypes.c_longlong, I, ctypes.POINTER(I), P], I),
    "ap_cluster_seed_heads": ([P, ctypes.c_uint], I),
    "ap_cluster_round_seeded": ([P, F, ctypes.c_longlong, I, ctypes.POINTER(I), P], I),
    "ap_head_draws": ([ctypes.c_uint, I, P], I),
    "ap_cluster_num": ([P], I),
    "ap_cluster_edges": ([P], I),
    "ap_cluster_result": ([P, P, P, P], I),
    "ap_cluster_destroy": ([P], I),
    "ap_free_fraction": ([P, I, P, I, P, D, D, D, D, I, I, P, P], I),
}


class LevelDesc(ctypes.Structure):
    """ap_level_desc."""
    _fields_ = [("num_nodes", I), ("num_mov", I), ("num_pins", I), ("num_nets", I),
                ("pin_cell", P), ("pin_off", P), ("net_start", P), ("cell_start", P), ("cell_pin", P),
                ("size", P), ("ratio", P), ("pins", P), ("area", P), ("lo", P), ("hi", P),
                ("nbx", I), ("nby", I), ("bx", D), ("by", D), ("fixed", P), ("target", F),
                ("bin_area", D), ("mov_area", D),
                ("has_judge", I), ("njx", I), ("njy", I), ("jbx", D), ("jby", D), ("judge_bin_area", D),
                ("judge_size", P), ("judge_ratio", P), ("judge_fixed", P),
                ("poisson_method", I),
                ("u", P), ("v", P), ("g", P), ("best", P),
                ("S", P), ("kp", P), ("c", P),
                ("poisson_plan", P)]


_lib = None


def lib():
    """The loaded library (loaded once). Raises RuntimeError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libautoplace_gp.so is not built (%s): run `python -m autoplace_amd.build_lib`"
                               % LIB_PATH)
        h = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def check(status, what=""):
    if status != AP_OK:
        msg = lib().ap_last_error().decode(errors="replace")
        kind = {AP_ERR_ARG: ValueError, AP_ERR_UNSUPPORTED: NotImplementedError}.get(status, RuntimeError)
        raise kind("%s failed (status %d): %s" % (what, status, msg))


def exported_symbols():
    return list(_SIGNATURES)
