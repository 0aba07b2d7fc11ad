"""ctypes loading of the in-tree libraries and the error convention.

Every C entry point returns 0 or a negative SCD_E* code; a failure raises
ScdError carrying the code and the library's last-error message.  The HIP
engine library is required: there is no CPU fallback, and loading fails
loudly when the library has not been built.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")

SCD_OK = 0
SCD_EINVAL = -1
SCD_EIO = -2
SCD_ENOMEM = -3
SCD_EDEVICE = -4
SCD_ELIMIT = -5
SCD_ESTATE = -6
SCD_EBUSY = -7
INVALID_LABEL = 0xFFFFFFFF

_NAMES = {
    SCD_EINVAL: "EINVAL", SCD_EIO: "EIO", SCD_ENOMEM: "ENOMEM", SCD_EDEVICE: "EDEVICE",
    SCD_ELIMIT: "ELIMIT", SCD_ESTATE: "ESTATE", SCD_EBUSY: "EBUSY",
}


class ScdConvertInfoC(ctypes.Structure):
    _fields_ = [("edges_read", ctypes.c_uint64), ("id_bound", ctypes.c_uint32), ("vertices", ctypes.c_uint32),
                ("edges_simple", ctypes.c_uint64), ("edges_final", ctypes.c_uint32),
                ("vertices_final", ctypes.c_uint32), ("ms_reindex", ctypes.c_double), ("ms_sort", ctypes.c_double),
                ("ms_dedup", ctypes.c_double), ("ms_triangles", ctypes.c_double), ("ms_relabel", ctypes.c_double)]


class ScdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


_cache = {}

u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


def _load(name):
    if name in _cache:
        return _cache[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ScdError(SCD_ESTATE, f"{path} is not built; run `python het-scd_amd/build.py`")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib


class ScdGraphC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("indices", u32p), ("adj", u32p),
                ("tri", i32p), ("tri_deg", i32p)]


class ScdEdgesC(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("u", u32p), ("v", u32p)]


class ScdTuningC(ctypes.Structure):
    _fields_ = [("lds_tile", ctypes.c_uint32), ("merge_chunk", ctypes.c_uint32), ("wave_max", ctypes.c_uint32),
                ("hub_table", ctypes.c_uint32), ("lds_hash", ctypes.c_uint32), ("lanes", ctypes.c_uint32),
                ("dense_wcc", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def host():
    """libscd_host.so with argtypes set."""
    if "host_ready" in _cache:
        return _cache["libscd_host.so"]
    lib = _load("libscd_host.so")
    G = ctypes.POINTER(ScdGraphC)
    Ed = ctypes.POINTER(ScdEdgesC)
    sig = {
        "scd_host_last_error": (ctypes.c_char_p, []),
        "scd_graph_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(G)]),
        "scd_graph_write": (ctypes.c_int, [G, ctypes.c_char_p]),
        "scd_graph_from_arrays": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, i32p, i32p,
                                                 ctypes.POINTER(G)]),
        "scd_graph_free": (None, [G]),
        "scd_graph_clustering_coef": (ctypes.c_double, [G]),
        "scd_gen_planted_partition": (ctypes.c_int, [ctypes.c_uint32, u32p, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_uint64, ctypes.POINTER(Ed)]),
        "scd_gen_planted_random": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                                  ctypes.c_double, u32p, u32p, ctypes.POINTER(Ed)]),
        "scd_gen_rmat": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(Ed)]),
        "scd_edges_from_arrays": (ctypes.c_int, [ctypes.c_uint64, u32p, u32p, ctypes.POINTER(Ed)]),
        "scd_edges_free": (None, [Ed]),
        "scd_edges_write_text": (ctypes.c_int, [Ed, ctypes.c_char_p]),
        "scd_edges_read_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Ed)]),
        "scd_convert": (ctypes.c_int, [Ed, ctypes.POINTER(G)]),
        "scd_graph_concat": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(G), ctypes.POINTER(G), u64p]),
        "scd_labels_printed_count": (ctypes.c_uint32, [ctypes.c_uint32, u32p]),
        "scd_labels_compact": (ctypes.c_int, [ctypes.c_uint32, u32p, u32p, u32p]),
        "scd_labels_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, u32p]),
        "scd_labels_read": (ctypes.c_int, [ctypes.c_char_p, u32p, u32p, ctypes.POINTER(u32p)]),
        "scd_labels_free": (None, [u32p]),
        "scd_gen_planted_batch": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                                 ctypes.POINTER(G), u64p, f64p, u64p]),
    }
    for k, (res, args) in sig.items():
        f = getattr(lib, k)
        f.restype = res
        f.argtypes = args
    _cache["host_ready"] = True
    return lib


def hip():
    """libscd_hip.so with argtypes set."""
    if "hip_ready" in _cache:
        return _cache["libscd_hip.so"]
    host()
    lib = _load("libscd_hip.so")
    C = vp
    sig = {
        "scd_last_error": (ctypes.c_char_p, []),
        "scd_device_count": (ctypes.c_int, []),
        "scd_init": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(C)]),
        "scd_destroy": (None, [C]),
        "scd_upload": (ctypes.c_int, [C, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, i32p, i32p,
                                      ctypes.c_double]),
        "scd_count_triangles": (ctypes.c_int, [C, i32p, i32p, ctypes.c_int]),
        "scd_count_triangles_noalloc": (ctypes.c_int, [C, ctypes.c_void_p, ctypes.c_void_p]),
        "scd_set_triangle_split": (ctypes.c_int, [C, ctypes.c_uint32, ctypes.c_uint32]),
        "scd_convert_edges": (ctypes.c_int, [C, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             u32p, u32p, f64p]),
        "scd_download_graph": (ctypes.c_int, [C, u32p, u32p, i32p, i32p]),
        "scd_graph_shape": (ctypes.c_int, [C, u32p, u32p, f64p]),
        "scd_convert_last_info": (ctypes.c_int, [C, ctypes.POINTER(ScdConvertInfoC)]),
        "scd_initial_partition": (ctypes.c_int, [C, u32p, u32p]),
        "scd_labels_to_device": (ctypes.c_int, [C, u32p]),
        "scd_set_partition": (ctypes.c_int, [C, u32p, ctypes.c_uint32, i32p, i32p, i32p]),
        "scd_improve": (ctypes.c_int, [C, f64p, u32p]),
        "scd_store_best": (ctypes.c_int, [C]),
        "scd_labels_from_device": (ctypes.c_int, [C, u32p]),
        "scd_current_labels": (ctypes.c_int, [C, u32p]),
        "scd_refine": (ctypes.c_int, [C, ctypes.c_double, ctypes.c_uint32, f64p, u32p, f64p]),
        "scd_compute_wcc": (ctypes.c_int, [C, f64p, f64p, i32p, i32p]),
        "scd_get_stats": (ctypes.c_int, [C, u32p, i32p, i32p, i32p]),
        "scd_neighbor_label_pairs": (ctypes.c_int, [C, u32p, u32p, u32p, u32p, u64p]),
        "scd_get_tuning": (ctypes.c_int, [C, ctypes.POINTER(ScdTuningC)]),
        "scd_set_tuning": (ctypes.c_int, [C, ctypes.POINTER(ScdTuningC)]),
        "scd_last_counters": (ctypes.c_int, [C, u64p]),
        "scd_dense_wcc_info": (ctypes.c_int, [C, u32p]),
        "scd_prof_count": (ctypes.c_int, [C]),
        "scd_prof_get": (ctypes.c_int, [C, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), u64p, f64p]),
        "scd_prof_reset": (ctypes.c_int, [C]),
        "scd_device_bytes": (ctypes.c_uint64, [C]),
        "scd_bin_info": (ctypes.c_int, [C, ctypes.c_int, ctypes.c_void_p, u64p, u64p, u64p]),
        "scd_snapshot": (ctypes.c_int, [C]),
        "scd_set_graph_mode": (ctypes.c_int, [C, ctypes.c_int]),
        "scd_next_step_captures": (ctypes.c_int, [C]),
        "scd_set_profiling": (ctypes.c_int, [C, ctypes.c_int]),
        "scd_set_phase_timing": (ctypes.c_int, [C, ctypes.c_int]),
        "scd_capture_iterations": (ctypes.c_int, [C]),
        "scd_last_phase_ms": (ctypes.c_int, [C, f64p]),
        "scd_set_counting": (ctypes.c_int, [C, ctypes.c_int]),
        "scd_restore": (ctypes.c_int, [C]),
        "scd_calibration_copy": (ctypes.c_int, [C, ctypes.c_uint64, ctypes.c_int, f64p]),
        "scd_batch_upload": (ctypes.c_int, [C, ctypes.c_uint32, u64p, u32p, u32p, i32p, i32p, f64p]),
        "scd_batch_initial_partition": (ctypes.c_int, [C, u32p, u32p]),
        "scd_batch_compute_wcc": (ctypes.c_int, [C, f64p]),
        "scd_batch_refine": (ctypes.c_int, [C, f64p, ctypes.c_uint32, f64p, u32p, f64p, f64p]),
        "scd_batch_labels": (ctypes.c_int, [C, u32p]),
        "scd_batch_current_labels": (ctypes.c_int, [C, u32p]),
        "scd_batch_upload_parts": (ctypes.c_int, [C, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
        "scd_batch_export_labels": (ctypes.c_int, [C, ctypes.c_void_p, ctypes.c_void_p]),
        "scd_batch_graph_labels": (ctypes.c_int, [C, ctypes.c_uint32, ctypes.c_int, u32p]),
        "scd_batch_graphs_labels": (ctypes.c_int, [C, ctypes.c_uint32, u32p, ctypes.c_int, ctypes.POINTER(u32p)]),
        "scd_batch_step": (ctypes.c_int, [C, u32p, f64p]),
        "scd_autotune": (ctypes.c_int, [C, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ScdTuningC), f64p]),
        "scd_batch_store_best": (ctypes.c_int, [C, u32p]),
        "scd_eval_improvement": (ctypes.c_int, [C, ctypes.c_uint32, i32p, i32p, i32p, i32p, f64p, f64p, f64p]),
        "scd_eval_candidate": (ctypes.c_int, [C, ctypes.c_uint32, i32p, i32p, i32p, u32p, u32p, u32p, u32p,
                                              ctypes.c_double, f64p]),
    }
    for k, (res, args) in sig.items():
        f = getattr(lib, k, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _cache["hip_ready"] = True
    return lib


def check_host(rc):
    if rc != SCD_OK:
        raise ScdError(rc, host().scd_host_last_error().decode())
    return rc


def check_hip(rc):
    if rc < 0:
        raise ScdError(rc, hip().scd_last_error().decode())
    return rc


def ptr(a, t):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(t)


class ScdServerOptsC(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget", ctypes.c_uint64), ("max_graphs", ctypes.c_uint32),
                ("join_wait_ms", ctypes.c_double), ("max_iters", ctypes.c_uint32), ("graph_mode", ctypes.c_int),
                ("intake_graphs", ctypes.c_uint32), ("intake_bytes", ctypes.c_uint64)]


class ScdRequestC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("indices", u32p), ("adj", u32p), ("tri", i32p),
                ("tri_deg", i32p), ("clustering_coef", ctypes.c_double), ("arrival_ms", ctypes.c_double)]


class ScdResultC(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_uint32), ("initial_wcc", ctypes.c_double),
                ("best_wcc", ctypes.c_double), ("arrival_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("joined_arena_graphs", ctypes.c_uint32)]


class ScdServerStatsC(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("joins", ctypes.c_uint64), ("busy_ms", ctypes.c_double),
                ("peak_bytes", ctypes.c_uint64), ("peak_graphs", ctypes.c_uint32), ("admit_ms", ctypes.c_double),
                ("prep_ms", ctypes.c_double), ("assemble_ms", ctypes.c_double), ("step_ms", ctypes.c_double)]


def serve():
    """libscd_serve.so with argtypes set."""
    if "serve_ready" in _cache:
        return _cache["libscd_serve.so"]
    hip()
    lib = _load("libscd_serve.so")
    S = vp
    sig = {
        "scd_server_last_error": (ctypes.c_char_p, []),
        "scd_server_create": (ctypes.c_int, [ctypes.POINTER(ScdServerOptsC), ctypes.POINTER(S)]),
        "scd_server_destroy": (None, [S]),
        "scd_server_submit": (ctypes.c_int, [S, ctypes.POINTER(ScdRequestC), u64p]),
        "scd_server_run": (ctypes.c_int, [S]),
        "scd_server_result": (ctypes.c_int, [S, ctypes.c_uint64, u32p, ctypes.POINTER(ScdResultC)]),
        "scd_server_get_stats": (ctypes.c_int, [S, ctypes.POINTER(ScdServerStatsC)]),
        "scd_server_footprint": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
    }
    for k, (res, args) in sig.items():
        f = getattr(lib, k)
        f.restype = res
        f.argtypes = args
    _cache["serve_ready"] = True
    return lib


def check_serve(rc):
    if rc < 0:
        raise ScdError(rc, serve().scd_server_last_error().decode())
    return rc

