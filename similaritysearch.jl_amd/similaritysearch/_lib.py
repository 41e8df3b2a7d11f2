"""Loader of libssg_hip.so (the C ABI declared in include/ssg.h).

The product path has no CPU fallback: if the library is missing or cannot be
loaded, every GPU index raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SSG_HIP_LIB", os.path.join(_HERE, "..", "lib", "libssg_hip.so"))

SSG_SQL2, SSG_L2, SSG_NORMCOS = 0, 1, 2

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64


SSG_OPT_QUERY_LAYOUT = 1
SSG_OPT_PREC_SLICES = 2
SSG_OPT_PREC_SAMPLE = 3


class BeamParams(C.Structure):
    _fields_ = [("bsize", C.c_int32), ("delta", C.c_float), ("maxvisits", C.c_int64),
                ("beamcap", C.c_int32), ("precision", C.c_int32), ("rerank", C.c_int32),
                ("expand", C.c_int32)]


class QueryOpts(C.Structure):
    _fields_ = [("k", _p), ("bsize", _p), ("hints", _p), ("nhints", _p),
                ("hint_stride", C.c_int32), ("order", _p), ("visited", _p)]


class BuildParams(C.Structure):
    _fields_ = [("nb_logbase", C.c_float), ("nb_minsize", C.c_int32),
                ("parallel_block", C.c_int32), ("hints_logbase", C.c_float),
                ("cb_logbase", C.c_float), ("cb_starting", C.c_int32),
                ("opt_minrecall", C.c_float), ("opt_numqueries", C.c_int32),
                ("opt_ksearch", C.c_int32), ("beamcap", C.c_int32), ("seed", C.c_uint64),
                ("verbose", C.c_int32)]


class LaunchParams(C.Structure):
    _fields_ = [("slots_per_cu", C.c_int32), ("rounds", C.c_int32), ("vbuckets", C.c_int32)]


# name -> (restype, argtypes); the exported surface of include/ssg.h
SIGNATURES = {
    "ssg_dev_malloc": (C.c_int, [C.POINTER(_p), C.c_size_t]),
    "ssg_dev_free": (C.c_int, [_p]),
    "ssg_memcpy": (C.c_int, [_p, _p, C.c_size_t, _i32]),
    "ssg_device_synchronize": (C.c_int, []),
    "ssg_stream_create": (C.c_int, [C.POINTER(_p)]),
    "ssg_stream_destroy": (C.c_int, [_p]),
    "ssg_stream_synchronize": (C.c_int, [_p]),
    "ssg_event_create": (C.c_int, [C.POINTER(_p)]),
    "ssg_event_destroy": (C.c_int, [_p]),
    "ssg_event_record": (C.c_int, [_p, _p]),
    "ssg_event_elapsed_ms": (C.c_int, [_p, _p, C.POINTER(C.c_float)]),
    "ssg_mlp_create": (C.c_int, [C.POINTER(_p), C.c_int, _i32, _i32, _i32, _i32, _i32,
                                 C.c_uint64]),
    "ssg_mlp_destroy": (C.c_int, [_p]),
    "ssg_mlp_get": (C.c_int, [_p, _i32, _p, _p]),
    "ssg_mlp_set": (C.c_int, [_p, _i32, _p, _p]),
    "ssg_mlp_grad": (C.c_int, [_p, _i32, _p, _p]),
    "ssg_mlp_step": (C.c_int, [_p, _p, _p, _p, _i32, C.c_float, C.c_float, _i32,
                               C.POINTER(C.c_float)]),
    "ssg_mlp_train_dev": (C.c_int, [_p, _p, _p, _p, _i64, _i32, _i32, C.c_float, C.c_float,
                                    C.c_uint64, _p]),
    "ssg_mlp_infer_dev": (C.c_int, [_p, _p, _i64, _p, _i32, C.c_float, _p, _p, _p, _p]),
    "ssg_learned_plan_dev": (C.c_int, [_p, _p, _p, _i64, _i32, C.c_float, _p, _i32, _i32, _p,
                                       _p, _p]),
    "ssg_learned_search_dev": (C.c_int, [_p, _p, _p, _p, _i64, _i32, _i32,
                                         C.POINTER(BeamParams), _i32, C.c_float, C.c_float, _p,
                                         _i32, _p, _p, _p, _p, _p]),
    "ssg_fft": (C.c_int, [_p, _i32, _i64, _p, _p, _p, _p]),
    "ssg_build_graph": (C.c_int, [_p, C.POINTER(BuildParams)]),
    "ssg_graph_info": (C.c_int, [_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(BeamParams)]),
    "ssg_download_graph": (C.c_int, [_p, _p, _p, _p]),
    "ssg_optimize_minrecall": (C.c_int, [_p, _p, _i64, _i32, C.c_float, C.POINTER(BeamParams),
                                         C.POINTER(C.c_double)]),
    "ssg_version": (C.c_char_p, []),
    "ssg_last_error": (C.c_char_p, [_p]),
    "ssg_create": (C.c_int, [C.POINTER(_p), C.c_int]),
    "ssg_destroy": (C.c_int, [_p]),
    "ssg_upload_db": (C.c_int, [_p, _p, _i32, _i64, _i32]),
    "ssg_upload_graph": (C.c_int, [_p, _p, _i64, _p, _i64, _p, _i64]),
    "ssg_exhaustive_knn": (C.c_int, [_p, _p, _i64, _i32, _p, _p]),
    "ssg_exhaustive_knn_dev": (C.c_int, [_p, _p, _i64, _i32, _p, _p, _p]),
    "ssg_beamsearch": (C.c_int, [_p, _p, _i64, _i32, _i32, C.POINTER(BeamParams),
                                 C.POINTER(QueryOpts), _p, _p, _p, _p]),
    "ssg_beamsearch_dev": (C.c_int, [_p, _p, _i64, _i32, _i32, C.POINTER(BeamParams),
                                     C.POINTER(QueryOpts), _p, _p, _p, _p, _p]),
    "ssg_beamsearch_noalloc_dev": (C.c_int, [_p, _p, _i64, _i32, _i32, C.POINTER(BeamParams),
                                             C.POINTER(QueryOpts), _p, _p, _p, _p, _p]),
    "ssg_set_launch": (C.c_int, [_p, C.POINTER(LaunchParams)]),
    "ssg_get_launch": (C.c_int, [_p, C.POINTER(LaunchParams)]),
    "ssg_optimize_launch": (C.c_int, [_p, _p, _i64, _i32, _i32, C.POINTER(BeamParams),
                                      C.POINTER(QueryOpts), _p, _i32, C.c_float,
                                      C.POINTER(LaunchParams), C.POINTER(C.c_double),
                                      C.POINTER(C.c_float)]),
    "ssg_allknn_graph": (C.c_int, [_p, _i32, C.POINTER(BeamParams), _p, _p]),
    "ssg_allknn_exhaustive": (C.c_int, [_p, _i32, _p, _p]),
    "ssg_order_by_bsize_dev": (C.c_int, [_p, _p, _i64, _p, _p]),
    "ssg_set_option": (C.c_int, [_p, _i32, _i64]),
    "ssg_capture_begin": (C.c_int, [_p]),
    "ssg_capture_end": (C.c_int, [_p, C.POINTER(_p)]),
    "ssg_graph_launch": (C.c_int, [_p, _p]),
    "ssg_graph_destroy": (C.c_int, [_p]),
    "ssg_rerank": (C.c_int, [_p, _p, _i64, _i32, _p, _p]),
    "ssg_append_db": (C.c_int, [_p, _p, _i64]),
    "ssg_append_items_graph": (C.c_int, [_p, _p, _i64]),
    "ssg_neardup": (C.c_int, [_p, C.c_float, _i32, _i32, _p, _p, _p, _p, _p]),
    "ssg_neardup_graph": (C.c_int, [_p, _p, C.c_float, _i32, _i32, _i32, C.POINTER(BuildParams),
                                    _p, _p, _p, _p, _p]),
    "ssg_rerank_dev": (C.c_int, [_p, _p, _i64, _i32, _p, _p, _p]),
    "ssg_hsp_queries": (C.c_int, [_p, _i64, _i32, _p, _p, _p, _p]),
    "ssg_hsp_queries_dev": (C.c_int, [_p, _i64, _i32, _p, _p, _p, _p, _p]),
    "ssg_exhaustive_knn_prec": (C.c_int, [_p, _p, _i64, _i32, _i32, _i32, _p, _p, _p, _p]),
    "ssg_exhaustive_knn_prec_dev": (C.c_int, [_p, _p, _i64, _i32, _i32, _i32, _p, _p, _p, _p, _p]),
    "ssg_allknn_exhaustive_prec": (C.c_int, [_p, _i32, _i32, _i32, _p, _p, _p, _p]),
    "ssg_allknn_exhaustive_prec_dev": (C.c_int, [_p, _i32, _i32, _i32, _p, _p, _p, _p, _p]),
}

_lib = None


class SSGError(RuntimeError):
    """A failed libssg_hip call (the message comes from ssg_last_error)."""


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SSGError(f"libssg_hip.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = load().ssg_last_error(None)
        msg = msg.decode() if msg else ""
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise SSGError(f"{what} failed ({rc}): {msg}")
