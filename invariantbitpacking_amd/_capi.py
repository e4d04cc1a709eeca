"""ctypes binding of libibp.so (the C ABI declared in include/ibp.h).

The library is loaded from this package directory.  There is no fallback: if
libibp.so is missing, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libibp.so")

_P = ctypes.c_void_p
_I32 = ctypes.c_int
_I64 = ctypes.c_int64
_SZ = ctypes.c_size_t
_F = ctypes.c_float


class FetchTask(ctypes.Structure):
    """ibp_fetch_task."""
    _fields_ = [("comp", _P), ("mask", _P), ("bitval", _P), ("bitmask", _P), ("index", _P), ("out", _P),
                ("num_vecs", _I64), ("vec_size", _I64), ("n", _I64), ("elem_bytes", ctypes.c_int32),
                ("index_bytes", ctypes.c_int32), ("comp_len", ctypes.c_int32), ("comp_loc", ctypes.c_int32),
                ("offsets", _P), ("stream_bytes", _I64)]


class Layout(ctypes.Structure):
    """ibp_layout: condensed offsets and/or per-row mask rows (include/ibp.h)."""
    _fields_ = [("offsets", _P), ("stream_bytes", _I64), ("clusters", _P), ("num_clusters", _I64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _SZ), ("num_streams", _I32), ("max_batch_requests", _I32),
                ("max_batch_rows", _I64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", _I64), ("launches", _I64), ("rows", _I64), ("p50_us", ctypes.c_double),
                ("p99_us", ctypes.c_double), ("mean_batch_requests", ctypes.c_double),
                ("peak_arena_bytes", _SZ), ("waited_admissions", _I64)]


# name -> (restype, argtypes)
SIGNATURES = {
    "ibp_last_error": (ctypes.c_char_p, []),
    "ibp_version": (ctypes.c_char_p, []),
    "ibp_set_debug": (None, [_I32]),
    "ibp_index_errors": (_I32, [ctypes.POINTER(ctypes.c_uint64), _I32]),
    "ibp_preprocess_workspace_size": (_SZ, [_I32, _I64, _F]),
    "ibp_preprocess": (_I32, [_I32, _P, _I64, _I64, _F, _P, _P, _P, _P, _SZ, _P]),
    "ibp_preprocess_kmeans": (_I32, [_I32, _P, _I64, _I64, _I32, _F, _I64, _P, _P, _P, _P]),
    "ibp_get_compress_size": (_I32, [_I32, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P, _P, _P]),
    "ibp_compress": (_I32, [_I32, _P, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P, _P]),
    "ibp_decompress_fetch": (_I32, [_I32, _P, _P, _I64, _I64, _P, _P, _P, _P, _I32, _I64, _I32, _I32, _I32,
                                    _I32, _I32, _P]),
    "ibp_get_compress_size_ex": (_I32, [_I32, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P, _P,
                                        ctypes.POINTER(Layout), _P]),
    "ibp_condensed_offsets": (_I32, [_I32, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P, ctypes.POINTER(Layout), _P]),
    "ibp_condensed_offsets_noalloc": (_I32, [_I32, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P,
                                             ctypes.POINTER(Layout), _P]),
    "ibp_compress_ex": (_I32, [_I32, _P, _P, _I64, _I64, _P, _P, _P, _I32, _I64, _P, ctypes.POINTER(Layout), _P]),
    "ibp_decompress_fetch_ex": (_I32, [_I32, _P, _P, _I64, _I64, _P, _P, _P, _P, _I32, _I64, _I32, _I32, _I32,
                                       _I32, _I32, ctypes.POINTER(Layout), _P]),
    "ibp_read_file": (_I32, [ctypes.c_char_p, _P, _SZ, _I32, ctypes.POINTER(_SZ)]),
    "ibp_set_fetch_tuning": (_I32, [_I32, _I32]),
    "ibp_plan_create": (_I32, [ctypes.POINTER(FetchTask), _I32, ctypes.POINTER(_P)]),
    "ibp_plan_destroy": (_I32, [_P]),
    "ibp_plan_launch": (_I32, [_P, _P]),
    "ibp_plan_autotune": (_I32, [_P, _P, ctypes.POINTER(_F)]),
    "ibp_plan_capture_graph": (_I32, [_P, _P]),
    "ibp_plan_graph_launch": (_I32, [_P, _P]),
    "ibp_plan_describe": (_I32, [_P, ctypes.c_char_p, _SZ]),
    "ibp_plan_get_config": (_I32, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "ibp_plan_set_config": (_I32, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "ibp_server_create": (_I32, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "ibp_server_destroy": (_I32, [_P]),
    "ibp_server_register": (_I32, [_P, _I32, _P, _I64, _I64, _P, _P, _P, _I32, _I32, ctypes.POINTER(_I32)]),
    "ibp_server_submit": (_I32, [_P, _I32, _P, _I64, ctypes.POINTER(_I64)]),
    "ibp_server_wait": (_I32, [_P, _I64, _I32, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double)]),
    "ibp_server_release": (_I32, [_P, _I64]),
    "ibp_server_release_after": (_I32, [_P, _I64, _P]),
    "ibp_server_submit_at": (_I32, [_P, _I32, _P, _I64, ctypes.c_double, ctypes.POINTER(_I64)]),
    "ibp_server_clock_us": (ctypes.c_double, []),
    "ibp_selftest_serve_admission": (_I32, [_I32, ctypes.c_uint64]),
    "ibp_selftest_index_copy": (_I32, [_I32, ctypes.c_uint64]),
    "ibp_server_stats_get": (_I32, [_P, ctypes.POINTER(ServerStats)]),
    "ibp_server_set_budget_override": (_I32, [_P, _SZ]),
    "ibp_server_request_times": (_I32, [_P, _I64, ctypes.POINTER(ctypes.c_double)]),
    "ibp_selftest_bitops": (_I32, [_I32, ctypes.c_uint64]),
    "ibp_selftest_rand": (_I32, [ctypes.c_uint32, ctypes.c_int64]),
}

STATUS = {0: "IBP_OK", 1: "IBP_ERR_INVALID", 2: "IBP_ERR_HIP", 3: "IBP_ERR_NOMEM", 4: "IBP_ERR_UNSUPPORTED",
          5: "IBP_ERR_BUSY"}

_lib = None


class IBPError(RuntimeError):
    """A libibp call failed (status != IBP_OK)."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def lib():
    """The loaded libibp.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libibp.so is not built (%s); run `python -m invariantbitpacking_amd.csrc.build`"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().ibp_last_error().decode(errors="replace")


def check(status):
    """Raises IBPError (a RuntimeError, like the reference's TORCH_CHECK) on failure."""
    if status != 0:
        raise IBPError(status, "%s: %s" % (STATUS.get(status, status), last_error()))
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))
