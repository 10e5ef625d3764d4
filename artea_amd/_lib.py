"""ctypes binding of libartea_hip.so (the C-ABI in include/artea_hip.h).

The library is loaded from this package directory only; a missing or
unloadable library raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libartea_hip.so")

P = C.c_void_p
u32, u64, f32, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_int
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); mirrors include/artea_hip.h one to one.
SIGNATURES = {
    "artea_last_error": (C.c_char_p, []),
    "artea_version": (C.c_char_p, []),
    "artea_graph_from_csr": (i32, [u32, u32, u32, u32, u32, P, P, P, u64, PP]),
    "artea_graph_restore": (i32, [C.c_char_p, PP]),
    "artea_graph_snapshot": (i32, [P, C.c_char_p]),
    "artea_graph_info": (i32, [P, P, P, P, P, P, P]),
    "artea_graph_vertex_info": (C.POINTER(u32), [P]),
    "artea_graph_group_offsets": (C.POINTER(u64), [P, u32, P]),
    "artea_graph_nbrs": (C.POINTER(u32), [P]),
    "artea_graph_level_nbrs": (C.POINTER(u32), [P, u32, u32, P]),
    "artea_graph_destroy": (i32, [P]),
    "artea_index_set_graph": (i32, [P, P]),
    "artea_index_get_graph": (i32, [P, PP]),
    "artea_index_create": (i32, [P, u32, u32, i32, PP]),
    "artea_index_create_device": (i32, [P, u32, u32, i32, PP]),
    "artea_index_set_graph_csr": (i32, [P, u32, u32, u32, u32, P, P, P, u64]),
    "artea_index_set_graph_hgra": (i32, [P, C.c_char_p]),
    "artea_index_set_graph_hgra_segmented": (i32, [P, C.c_char_p, u32, u32]),
    "artea_index_hgra_seconds": (i32, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "artea_index_layout": (i32, [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64)]),
    "artea_index_snapshot_hgra": (i32, [P, C.c_char_p]),
    "artea_index_info": (i32, [P, P, P, P, P, P, P]),
    "artea_index_device_vectors": (P, [P]),
    "artea_index_set_u8_rows": (i32, [P, i32, P]),
    "artea_index_u8_rows": (i32, [P]),
    "artea_index_destroy": (i32, [P]),
    "artea_router_create": (i32, [P, u32, u32, PP]),
    "artea_router_destroy": (i32, [P]),
    "artea_router_batch_query": (i32, [P, P, u32, u32, P, P, P]),
    "artea_router_batch_query_device": (i32, [P, P, u32, u32, P, P, P, P]),
    "artea_router_last_stats": (i32, [P, P, P, P]),
    "artea_router_last_overflows": (i32, [P, u32, P, C.POINTER(u32)]),
    "artea_batched_router_last_overflows": (i32, [P, u32, P, C.POINTER(u32)]),
    "artea_batched_query_noalloc": (i32, [P, u32, u32, P, u32, P, P, u32, P, P, P, P]),
    "artea_batched_router_create": (i32, [P, u32, u32, PP]),
    "artea_batched_router_destroy": (i32, [P]),
    "artea_batched_router_query_device": (i32, [P, P, u32, P, P, u32, P, P, P, P]),
    "artea_batched_router_last_stats": (i32, [P, P, P, P]),
    "artea_batched_router_capture": (i32, [P, P, u32, P, P, u32, P, P, P, PP]),
    "artea_batched_router_autotune": (i32, [P, P, u32, P, P, u32, P, P, P, P, C.POINTER(i32)]),
    "artea_batched_router_set_variant": (i32, [P, i32]),
    "artea_batched_router_get_variant": (i32, [P, C.POINTER(i32)]),
    "artea_launch_graph_launch": (i32, [P, P]),
    "artea_launch_graph_destroy": (i32, [P]),
    "artea_server_create": (i32, [P, u32, u32, u32, u32, PP]),
    "artea_server_create_budget": (i32, [P, u32, u32, u32, u64, u32, PP]),
    "artea_server_admission": (i32, [P, P, P, C.POINTER(u32), C.POINTER(u64)]),
    "artea_server_destroy": (i32, [P]),
    "artea_server_submit": (i32, [P, P, u32, P, P, C.POINTER(u64)]),
    "artea_server_poll": (i32, [P, u32, P, P, P, C.POINTER(u32)]),
    "artea_server_poll_into": (i32, [P, u64, u64, P, P, u32, P, C.POINTER(u32)]),
    "artea_server_status": (i32, [P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
    "artea_server_overflows": (i32, [P, C.POINTER(u64), C.POINTER(u64)]),
    "artea_distance_matrix_device": (i32, [i32, P, u32, P, u32, u32, P, P]),
    "artea_generate": (i32, [i32, u64, u32, u64, u64, u32, u32, P, i32]),
    "artea_bruteforce_knn_device": (i32, [i32, P, u32, P, u32, u32, u32, P, P, P]),
    "artea_bruteforce_knn_mfma_device": (i32, [i32, P, u32, P, u32, u32, u32, P, P, P, P]),
    "artea_device_alloc": (i32, [C.c_size_t, PP]),
    "artea_device_free": (i32, [P]),
    "artea_memcpy_htod": (i32, [P, P, C.c_size_t, P]),
    "artea_memcpy_dtoh": (i32, [P, P, C.c_size_t, P]),
    "artea_memset_device": (i32, [P, i32, C.c_size_t, P]),
    "artea_copy_device": (i32, [P, P, C.c_size_t, P]),
    "artea_event_query": (i32, [P, C.POINTER(i32)]),
    "artea_host_alloc": (i32, [C.c_size_t, PP]),
    "artea_host_free": (i32, [P]),
    "artea_stream_create": (i32, [PP]),
    "artea_stream_destroy": (i32, [P]),
    "artea_stream_synchronize": (i32, [P]),
    "artea_event_create": (i32, [PP]),
    "artea_event_destroy": (i32, [P]),
    "artea_event_record": (i32, [P, P]),
    "artea_event_elapsed_ms": (i32, [P, P, C.POINTER(f32)]),
    "artea_device_info": (i32, [C.c_char_p, C.c_size_t, C.POINTER(i32), C.POINTER(C.c_size_t)]),
    "artea_refine_create": (i32, [P, u32, u32, i32, u32, P, u32, u64, PP]),
    "artea_refine_destroy": (i32, [P]),
    "artea_refine_upload_rows": (i32, [P, P, P]),
    "artea_refine_download_rows": (i32, [P, P, P]),
    "artea_refine_prefill": (i32, [P, u32, u32, u32, u32, u32, u64]),
    "artea_refine_triangle": (i32, [P, f32, f32]),
    "artea_refine_triangle_range": (i32, [P, f32, f32, u32, u32]),
    "artea_refine_reverse_range": (i32, [P, u32, u32]),
    "artea_refine_truncate": (i32, [P]),
    "artea_refine_reverse": (i32, [P]),
    "artea_refine_schedule": (i32, [P, u32, u32, f32, f32]),
    "artea_refine_set_join": (i32, [P, i32]),
    "artea_refine_set_recheck_queue": (i32, [P, u32]),
    "artea_refine_last_stats": (i32, [P, P, P, P]),
    "artea_build_full": (i32, [P, u32, u32, u32, u32, u32, f32, f32, f32, f32, u64, i32, f32, P]),
    "artea_build_full_routing": (i32, [P, u32, u32, u32, u32, u32, f32, f32, f32, f32, u64, i32, f32, u32, u32, u32,
                                       P]),
    "artea_refine_routing_loops": (i32, [P, u32, u32, u32, f32]),
    "artea_rnet_build": (i32, [P, i32, P, P]),
    "artea_rnet_info": (i32, [P, P, P, P, P, P]),
    "artea_rnet_levels": (i32, [P, P]),
    "artea_rnet_level_rows": (i32, [P, u32, P, P, u64, P]),
    "artea_rnet_destroy": (i32, [P]),
    "artea_build_full_rnet": (i32, [P, P, u32, u32, f32, f32, i32, u32, u32, u32, P]),
    "artea_refine_prune": (i32, [P, f32]),
    "artea_refine_prune_range": (i32, [P, f32, u32, u32]),
    "artea_refine_routing_stats": (i32, [P, P, P, P]),
    "artea_index_compact_graph": (i32, [P, u32, u32, u32, P, P, P, P, i32]),
    "artea_recall_at_k_device": (i32, [P, u32, P, u32, u32, u32, C.POINTER(u64), P]),
    "artea_adr_at_1_device": (i32, [i32, P, u32, P, u32, u32, P, u32, P, u32, C.POINTER(C.c_double),
                                    C.POINTER(u64), P]),
    "artea_probe_knn_device": (i32, [i32, P, u32, u32, P, u32, P, C.POINTER(f32), P]),
    "artea_probe_farthest_device": (i32, [i32, P, u32, u32, P, u32, P, P]),
    "artea_distance_rows_device": (i32, [i32, P, u32, P, u32, P, u32, u32, u32, P, P]),
}

STATUS = {0: "ARTEA_OK", 1: "ARTEA_EINVAL", 2: "ARTEA_EIO", 3: "ARTEA_EHIP", 4: "ARTEA_ENOMEM", 5: "ARTEA_EOVERFLOW"}

_lib = None


class ArteaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libartea_hip.so is not built (run __graft_entry__.build() or artea_amd/build.py)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc: int):
    if rc != 0:
        msg = lib().artea_last_error().decode(errors="replace")
        if rc == 5:
            raise OverflowError(msg)
        raise ArteaError(rc, "%s: %s" % (STATUS.get(rc, rc), msg))


def exported_symbols():
    l = lib()
    return [n for n in SIGNATURES if getattr(l, n, None) is not None]
