"""ctypes binding of libstwo_hip.so (the C-ABI declared in include/stwo_hip.h).

The library must have been built (``python -m proving_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise - there is no
CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstwo_hip.so")

c_int, c_u32, c_u64, c_vp, c_char_p = (ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                      ctypes.c_void_p, ctypes.c_char_p)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> (restype, argtypes)
SIGNATURES = {
    "sh_ctx_create": (c_int, [c_int, ctypes.POINTER(c_vp)]),
    "sh_ctx_destroy": (c_int, [c_vp]),
    "sh_last_error": (c_char_p, [c_vp]),
    "sh_sync": (c_int, [c_vp]),
    "sh_profile_enable": (c_int, [c_vp, c_int]),
    "sh_profile_report": (c_int, [c_vp, c_char_p, c_u64]),
    "sh_timer_start": (c_int, [c_vp]),
    "sh_timer_stop": (c_int, [c_vp, c_f32p]),
    "sh_alloc": (c_int, [c_vp, c_u64, ctypes.POINTER(c_vp)]),
    "sh_free": (c_int, [c_vp, c_vp]),
    "sh_upload": (c_int, [c_vp, c_vp, c_vp, c_u64]),
    "sh_download": (c_int, [c_vp, c_vp, c_vp, c_u64]),
    "sh_memset": (c_int, [c_vp, c_vp, c_int, c_u64]),
    "sh_copy": (c_int, [c_vp, c_vp, c_vp, c_u64]),
    "sh_pool_trim": (c_int, [c_vp]),
    "sh_bit_reverse_column": (c_int, [c_vp, c_vp, c_u32]),
    "sh_twiddles_create": (c_int, [c_vp, c_u32, c_u32, ctypes.POINTER(c_vp)]),
    "sh_twiddles_destroy": (c_int, [c_vp, c_vp]),
    "sh_twiddles_download": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "sh_interpolate": (c_int, [c_vp, c_vp, c_vp, c_u32]),
    "sh_interpolate_columns": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64]),
    "sh_evaluate": (c_int, [c_vp, c_vp, c_vp, c_u32, c_vp, c_u32]),
    "sh_evaluate_polynomials": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_vp]),
    "sh_split_at_mid": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "sh_eval_at_points": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp, c_u32, c_vp]),
    "sh_eval_at_dpoint": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp]),
    "sh_wide_fibonacci_composition": (c_int, [c_vp, c_vp, c_u64, c_u32, c_u32, c_u32, c_vp, c_vp]),
    "sh_wide_fibonacci_composition_dev": (c_int, [c_vp, c_vp, c_u64, c_u32, c_u32, c_u32, c_vp, c_vp]),
    "sh_wide_fibonacci_oods_check": (c_int, [c_vp, c_vp, c_u64, c_vp, c_vp, c_u32, c_vp]),
    "sh_wide_fibonacci_trace": (c_int, [c_vp, c_vp, c_vp, c_u32, c_u64, c_vp]),
    "sh_merkle_build_leaves": (c_int, [c_vp, c_vp, c_vp, c_u64, c_u32, c_vp]),
    "sh_merkle_build_next_layer": (c_int, [c_vp, c_vp, c_u32, c_vp]),
    "sh_merkle_commit_lifted": (c_int, [c_vp, c_vp, c_vp, c_u64, c_u32, c_vp, c_vp]),
    "sh_merkle_commit_on_layer": (c_int, [c_vp, c_u32, c_vp, c_vp, c_u64, c_vp]),
    "sh_merkle_commit_mixed": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp, c_vp]),
    "sh_commit_trace": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp]),
    "sh_commit_trace_mixed": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp]),
    "sh_eval_at_point": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "sh_barycentric_weights": (c_int, [c_vp, c_u32, c_vp, c_vp]),
    "sh_barycentric_eval_at_point": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "sh_fold_line": (c_int, [c_vp, c_vp, c_vp, c_u32, c_u32, c_vp, c_u32, c_vp]),
    "sh_fold_circle_into_line": (c_int, [c_vp, c_vp, c_vp, c_u32, c_vp, c_vp]),
    "sh_fri_decompose": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "sh_accumulate_numerators": (c_int, [c_vp, c_vp, c_u64, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_u64, c_vp, c_vp]),
    "sh_compute_quotients_and_combine": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_u32,
                                                 c_vp]),
    "sh_compute_quotients_and_combine_noalloc": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32,
                                                         c_u32, c_vp]),
    "sh_lift_and_accumulate": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp]),
    "sh_accumulate": (c_int, [c_vp, c_vp, c_vp, c_u32]),
    "sh_grind": (c_int, [c_vp, c_vp, c_u32, c_u64p]),
    "sh_grind_dchannel": (c_int, [c_vp, c_vp, c_u32, c_vp]),
    "sh_set_fft_tuning": (c_int, [c_vp, c_u32, c_u32, c_u32, c_u32, c_u32]),
    "sh_set_fft_fusion": (c_int, [c_vp, c_u32]),
    "sh_get_fft_tuning": (c_int, [c_vp, c_u32, c_vp]),
    "sh_autotune_fft": (c_int, [c_vp, c_vp, c_vp, c_u64, c_u32]),
    "sh_commit_trace_graph_create": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_vp, c_vp,
                                             ctypes.POINTER(c_vp)]),
    "sh_commit_trace_mixed_graph_create": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u32, c_vp, c_vp,
                                             ctypes.POINTER(c_vp)]),
    "sh_graph_create": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "sh_graph_warmup_begin": (c_int, [c_vp, c_vp]),
    "sh_graph_warmup_end": (c_int, [c_vp, c_vp]),
    "sh_graph_capture_begin": (c_int, [c_vp, c_vp]),
    "sh_graph_capture_end": (c_int, [c_vp, c_vp]),
    "sh_graph_launch": (c_int, [c_vp, c_vp]),
    "sh_graph_destroy": (c_int, [c_vp, c_vp]),
    # stwo_prover.h
    "sh_channel_mix_u32s": (None, [c_vp, c_vp, c_u64]),
    "sh_channel_mix_u64": (None, [c_vp, c_u64]),
    "sh_channel_mix_felts": (None, [c_vp, c_vp, c_u64]),
    "sh_channel_mix_root": (None, [c_vp, c_vp]),
    "sh_channel_draw_u32s": (None, [c_vp, c_vp]),
    "sh_channel_draw_secure_felt": (None, [c_vp, c_vp]),
    "sh_channel_draw_random_point": (None, [c_vp, c_vp]),
    "sh_channel_draw_queries": (None, [c_vp, c_u32, c_u32, c_vp]),
    "sh_generate_secure_powers": (None, [c_vp, c_u64, c_vp]),
    "sh_dchannel_load": (c_int, [c_vp, c_vp, c_vp, c_u32]),
    "sh_dchannel_mix": (c_int, [c_vp, c_vp, c_vp, c_u64]),
    "sh_dchannel_draw_secure_felts": (c_int, [c_vp, c_vp, c_u32, c_vp]),
    "sh_dchannel_draw_random_point": (c_int, [c_vp, c_vp, c_vp]),
    "sh_secure_powers_dev": (c_int, [c_vp, c_vp, c_u64, c_int, c_vp]),
    "sh_merkle_decommit": (c_int, [c_vp, c_vp, c_u32, c_vp, c_u64, c_vp, c_u64, c_u64p]),
    "sh_gather": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp]),
    "sh_tree_decommit": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp, c_u32, c_vp, c_u64, c_vp, c_vp, c_u64, c_u64p]),
    "sh_fri_prove": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp, c_u64, c_u64p, c_vp, c_u64p]),
    "sh_fri_quotients": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp]),
    "sh_fri_quotients_dpoint": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp]),
    "sh_fri_prove_dch": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_u32, c_vp, c_vp, c_u32,
                                 c_vp, c_u64, c_u64p, c_vp, c_u64p]),
    "sh_fri_commit_dch": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_u32, ctypes.POINTER(c_vp)]),
    "sh_fri_decommit_dch": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp, c_u64, c_u64p, c_vp, c_u64p]),
    "sh_fri_run_free": (c_int, [c_vp, c_vp]),
}

_lib = None


class HipError(RuntimeError):
    pass


def load():
    """Loads libstwo_hip.so and declares every C-ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m proving_amd.build` "
            "(the HIP backend has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
