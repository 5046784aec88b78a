"""ctypes bindings to libvkzg.so, the gfx950 build of the vkzg C ABI.

The shared library is loaded from this package directory only; when it is
missing the import fails loudly (there is no CPU fallback).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvkzg.so")

# The eight symbols of include/vkzg.h (reference include/vkzg.h:53-121)
# and the additive ones of include/vkzg_amd.h.
ABI_SYMBOLS = [
    "vkzg_error_string",
    "vkzg_options_default",
    "vkzg_prover_new",
    "vkzg_prover_free",
    "vkzg_prover_device_name",
    "vkzg_prover_gpu_core_count",
    "vkzg_compute_proofs_batch",
    "vkzg_recover_cells_batch",
]
EXT_SYMBOLS = [
    "vkzg_amd_prover_new_with_setup",
    "vkzg_amd_compute_proofs_device",
    "vkzg_amd_recover_cells_device",
    "vkzg_amd_recover_cells_device_noalloc",
    "vkzg_amd_prover_info",
    "vkzg_amd_set_stage_timing",
    "vkzg_amd_last_stage_times",
    "vkzg_amd_debug_fp_mul",
    "vkzg_amd_debug_fp_inv",
    "vkzg_amd_debug_fr_mul",
    "vkzg_amd_server_options_default",
    "vkzg_amd_server_new",
    "vkzg_amd_server_free",
    "vkzg_amd_server_submit_proofs",
    "vkzg_amd_server_submit_recovery",
    "vkzg_amd_server_poll",
    "vkzg_amd_server_wait",
    "vkzg_amd_server_stats",
    "vkzg_amd_compute_cells_and_proofs",
    "vkzg_amd_recover_cells_and_proofs",
    "vkzg_amd_get_tuning",
    "vkzg_amd_set_tuning",
    "vkzg_amd_autotune",
    "vkzg_amd_set_graphs",
    "vkzg_amd_load_window_table",
    "vkzg_amd_load_g2_setup",
    "vkzg_amd_verify_cell_kzg_proof_batch",
    "vkzg_amd_blob_to_commitments",
    "vkzg_amd_debug_pairing",
    "vkzg_amd_debug_g1_msm",
    "vkzg_amd_debug_verify_terms",
    "vkzg_amd_debug_fp_mul_rate",
    "vkzg_amd_debug_force_checked",
    "vkzg_amd_debug_mad_rate",
    "vkzg_amd_write_v1_tables",
    "vkzg_amd_verify_challenge",
    "vkzg_amd_sha256",
    "vkzg_amd_sha256_hw",
]


class VkzgOptions(ctypes.Structure):
    _fields_ = [("max_batch_size", ctypes.c_uint32)]


class VkzgAmdInfo(ctypes.Structure):
    _fields_ = [
        ("window_bits", ctypes.c_uint32),
        ("windows", ctypes.c_uint32),
        ("table_bytes", ctypes.c_uint64),
        ("working_bytes", ctypes.c_uint64),
        ("max_batch_size", ctypes.c_uint32),
        ("setup_ms", ctypes.c_double),
        ("last_splits", ctypes.c_uint32),
        ("graphs_enabled", ctypes.c_uint32),
        ("graph_captures", ctypes.c_uint64),
        ("graph_replays", ctypes.c_uint64),
        ("autotuned", ctypes.c_uint32),
        ("autotune_ms", ctypes.c_double),
        ("phase_b_items", ctypes.c_uint32),
        ("resident_table_bytes", ctypes.c_uint64),
        ("extra_table_windows", ctypes.c_uint32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("phase_a_groups", ctypes.c_uint32),
        ("phase_a_blocks_per_cu", ctypes.c_uint32),
        ("normalize_chunk", ctypes.c_uint32),
        ("splits", ctypes.c_uint32),
        ("recovery_waves", ctypes.c_uint32),
        ("ladder_splits", ctypes.c_uint32),
        ("reduce_lanes", ctypes.c_uint32),
        ("recovery_radix", ctypes.c_uint32),
        ("phase_a_window_bits", ctypes.c_uint32),
    ]


class ServerOptions(ctypes.Structure):
    _fields_ = [("batch_blobs", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_uint64), ("start_ns", ctypes.c_uint64),
                ("done_ns", ctypes.c_uint64), ("batches", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("blobs", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("batch_blobs", ctypes.c_uint32), ("slot_bytes", ctypes.c_uint64)]


assert ctypes.sizeof(VkzgOptions) == 4

_lib = None


def load():
    """Loads and prototypes libvkzg.so (raises OSError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libvkzg.so is not built (run python3 vkzg_amd/build.py): " + LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u8p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p
    L.vkzg_error_string.restype = ctypes.c_char_p
    L.vkzg_error_string.argtypes = [ctypes.c_int]
    L.vkzg_options_default.restype = None
    L.vkzg_options_default.argtypes = [ctypes.POINTER(VkzgOptions)]
    L.vkzg_prover_new.restype = ctypes.c_int
    L.vkzg_prover_new.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(VkzgOptions)]
    L.vkzg_prover_free.restype = None
    L.vkzg_prover_free.argtypes = [vp]
    L.vkzg_prover_device_name.restype = ctypes.c_char_p
    L.vkzg_prover_device_name.argtypes = [vp]
    L.vkzg_prover_gpu_core_count.restype = ctypes.c_uint32
    L.vkzg_prover_gpu_core_count.argtypes = [vp]
    L.vkzg_compute_proofs_batch.restype = ctypes.c_int
    L.vkzg_compute_proofs_batch.argtypes = [vp, vp, vp, sz]
    L.vkzg_recover_cells_batch.restype = ctypes.c_int
    L.vkzg_recover_cells_batch.argtypes = [vp, vp, vp, vp, sz]
    L.vkzg_amd_prover_new_with_setup.restype = ctypes.c_int
    L.vkzg_amd_prover_new_with_setup.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(VkzgOptions),
                                                 u8p, sz]
    L.vkzg_amd_compute_proofs_device.restype = ctypes.c_int
    L.vkzg_amd_compute_proofs_device.argtypes = [vp, vp, vp, sz, vp]
    L.vkzg_amd_recover_cells_device.restype = ctypes.c_int
    L.vkzg_amd_recover_cells_device.argtypes = [vp, vp, vp, vp, vp, sz, vp]
    L.vkzg_amd_recover_cells_device_noalloc.restype = ctypes.c_int
    L.vkzg_amd_recover_cells_device_noalloc.argtypes = [vp, vp, vp, vp, vp, sz, vp]
    L.vkzg_amd_prover_info.restype = ctypes.c_int
    L.vkzg_amd_prover_info.argtypes = [vp, ctypes.POINTER(VkzgAmdInfo)]
    L.vkzg_amd_set_stage_timing.restype = ctypes.c_int
    L.vkzg_amd_set_stage_timing.argtypes = [vp, ctypes.c_int]
    L.vkzg_amd_last_stage_times.restype = ctypes.c_int
    L.vkzg_amd_last_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                            ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    L.vkzg_amd_server_options_default.restype = None
    L.vkzg_amd_server_options_default.argtypes = [ctypes.POINTER(ServerOptions)]
    L.vkzg_amd_server_new.restype = ctypes.c_int
    L.vkzg_amd_server_new.argtypes = [ctypes.POINTER(vp), vp, ctypes.POINTER(ServerOptions)]
    L.vkzg_amd_server_free.restype = None
    L.vkzg_amd_server_free.argtypes = [vp]
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.vkzg_amd_server_submit_proofs.restype = ctypes.c_int
    L.vkzg_amd_server_submit_proofs.argtypes = [vp, vp, vp, sz, u64p]
    L.vkzg_amd_server_submit_recovery.restype = ctypes.c_int
    L.vkzg_amd_server_submit_recovery.argtypes = [vp, vp, vp, vp, sz, u64p]
    L.vkzg_amd_server_poll.restype = ctypes.c_int
    L.vkzg_amd_server_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.vkzg_amd_server_wait.restype = ctypes.c_int
    L.vkzg_amd_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(RequestInfo)]
    L.vkzg_amd_server_stats.restype = ctypes.c_int
    L.vkzg_amd_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.vkzg_amd_compute_cells_and_proofs.restype = ctypes.c_int
    L.vkzg_amd_compute_cells_and_proofs.argtypes = [vp, vp, vp, vp, sz]
    L.vkzg_amd_recover_cells_and_proofs.restype = ctypes.c_int
    L.vkzg_amd_recover_cells_and_proofs.argtypes = [vp, vp, vp, vp, vp, sz]
    L.vkzg_amd_get_tuning.restype = ctypes.c_int
    L.vkzg_amd_get_tuning.argtypes = [vp, sz, ctypes.POINTER(Tuning)]
    L.vkzg_amd_set_tuning.restype = ctypes.c_int
    L.vkzg_amd_set_tuning.argtypes = [vp, sz, ctypes.POINTER(Tuning)]
    L.vkzg_amd_set_graphs.restype = ctypes.c_int
    L.vkzg_amd_set_graphs.argtypes = [vp, ctypes.c_int]
    L.vkzg_amd_debug_mad_rate.restype = ctypes.c_int
    L.vkzg_amd_debug_mad_rate.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.vkzg_amd_load_window_table.restype = ctypes.c_int
    L.vkzg_amd_load_window_table.argtypes = [vp, ctypes.c_uint32]
    L.vkzg_amd_debug_force_checked.restype = ctypes.c_int
    L.vkzg_amd_debug_force_checked.argtypes = [vp, ctypes.c_int]
    L.vkzg_amd_autotune.restype = ctypes.c_int
    L.vkzg_amd_autotune.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.vkzg_amd_debug_fp_inv.restype = ctypes.c_int
    L.vkzg_amd_debug_fp_inv.argtypes = [vp, vp, vp, sz]
    L.vkzg_amd_load_g2_setup.restype = ctypes.c_int
    L.vkzg_amd_load_g2_setup.argtypes = [vp, u8p, sz]
    L.vkzg_amd_verify_cell_kzg_proof_batch.restype = ctypes.c_int
    L.vkzg_amd_verify_cell_kzg_proof_batch.argtypes = [vp, ctypes.POINTER(ctypes.c_int), u8p, u64p, u8p, u8p,
                                                       sz]
    L.vkzg_amd_blob_to_commitments.restype = ctypes.c_int
    L.vkzg_amd_blob_to_commitments.argtypes = [vp, vp, vp, sz]
    L.vkzg_amd_debug_pairing.restype = ctypes.c_int
    L.vkzg_amd_debug_pairing.argtypes = [vp, vp, u8p, u8p, sz]
    L.vkzg_amd_debug_g1_msm.restype = ctypes.c_int
    L.vkzg_amd_debug_g1_msm.argtypes = [vp, vp, u8p, u8p, sz]
    L.vkzg_amd_debug_fp_mul_rate.restype = ctypes.c_int
    L.vkzg_amd_debug_fp_mul_rate.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.vkzg_amd_write_v1_tables.restype = ctypes.c_int
    L.vkzg_amd_write_v1_tables.argtypes = [vp, ctypes.c_char_p]
    L.vkzg_amd_debug_verify_terms.restype = ctypes.c_int
    L.vkzg_amd_debug_verify_terms.argtypes = [vp, ctypes.POINTER(ctypes.c_int), vp, u8p, u64p, u8p, u8p, sz]
    L.vkzg_amd_verify_challenge.restype = ctypes.c_int
    L.vkzg_amd_verify_challenge.argtypes = [vp, u8p, u64p, u8p, u8p, sz, ctypes.c_int]
    L.vkzg_amd_sha256.restype = ctypes.c_int
    L.vkzg_amd_sha256.argtypes = [vp, u8p, sz, ctypes.c_int]
    L.vkzg_amd_sha256_hw.restype = ctypes.c_int
    L.vkzg_amd_sha256_hw.argtypes = []
    for name in ("vkzg_amd_debug_fp_mul", "vkzg_amd_debug_fr_mul"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, vp, vp, sz]
    _lib = L
    return L
