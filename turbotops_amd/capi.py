"""ctypes binding of libttgpu.so, the C ABI declared in include/ttgpu.h.

The library is the product's compute path; there is no CPU fallback. If it
has not been built, every entry point raises instead of computing anything.
Status codes map to the exception classes the reference's Python module
raises for the matching C++ exceptions (std::invalid_argument and
std::length_error -> ValueError, std::overflow_error -> OverflowError,
device failures -> RuntimeError).
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttgpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ttgpu.h")

TT_OK, TT_EINVAL, TT_ELENGTH, TT_EOVERFLOW, TT_EDEVICE = range(5)
METRIC_L1, METRIC_L2, METRIC_LINF, METRIC_LP = range(4)
LAYOUT_FULL, LAYOUT_LOWER = range(2)

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

# argtypes of every exported entry point (restype is tt_status unless noted)
_SIGS = {
    "tt_last_error": ([], C.c_char_p),
    "tt_version": ([], C.c_char_p),
    "tt_device_count": ([], C.c_int),
    "tt_metric_parse": ([C.c_char_p, C.POINTER(C.c_int), _f32p], None),
    "tt_dense_layout": ([C.c_uint32, C.c_uint64, C.POINTER(C.c_int)], None),
    "tt_make_dense": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_int, _vp, _vp], None),
    "tt_dist_matrix_dev": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_int, _vp, _vp, _vp, _vp],
                           None),
    "tt_dist_matrix_dev_events": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_int, _vp, _vp, _vp, _vp,
                                   _vp, _vp], None),
    "tt_radius_graph": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, _vp, _u64p, _vp, _vp],
                        None),
    "tt_knn_exact": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32, _u32p, _vp], None),
    "tt_knn_graph": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32, _vp, _vp, _u64p, _vp, _vp],
                     None),
    "tt_knn_graph_exact": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32, _vp, _u64p, _vp, _vp],
                           None),
    "tt_mst_deaths": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _vp], None),
    "tt_mst_deaths_dev": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _vp, _u32p, _u64p, _u64p, _vp], None),
    "tt_mst_deaths_dev_noalloc": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _vp, _vp], None),
    "tt_dense_edges": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, _u64p, _vp, _vp, _vp,
                        _f32p], None),
    "tt_dense_edges_dev": ([_vp, C.c_uint32, C.c_float, C.c_int, C.c_uint64, _u64p, _vp, _vp, _vp, _vp], None),
    "tt_diagram_backward": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _vp, _vp, C.c_uint32, _vp], None),
    "tt_diagram_backward_dev": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _vp, _vp, C.c_uint32, _vp,
                                 _vp], None),
    "tt_sparse_edges": ([C.c_uint32, _vp, _vp, _vp, C.c_float, _u64p, _vp, _vp, _vp], None),
    "tt_build_edges": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_int, _vp, _u64p, _vp,
                        _vp, _f32p], None),
    "tt_build_edges_witness": ([_vp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_int, _vp, _u64p,
                                _vp, _vp, _f32p], None),
    "tt_build_edges_csr": ([C.c_uint32, _vp, _vp, _vp, C.c_float, _u64p, _vp, _vp], None),
    "tt_complex_dense": ([_vp, C.c_uint32, C.c_int, C.c_int, C.c_float, C.c_uint32, C.c_int, _vp], None),
    "tt_complex_sparse": ([C.c_uint32, _vp, _vp, _vp, C.c_float, C.c_uint32, C.c_int, _vp], None),
    "tt_complex_emergent": ([_vp, C.c_uint32, _vp, _vp, C.c_uint64, _vp, _vp], None),
    "tt_complex_seed": ([_vp, C.c_uint32, _vp, _vp, C.c_uint64], None),
    "tt_complex_assemble": ([_vp, _vp, C.c_uint64, C.c_int, _u64p, _vp, _vp, _vp, _vp, _u64p], None),
    "tt_complex_destroy": ([_vp], C.c_int),
    "tt_free": ([_vp], C.c_int),
}


class TTError(RuntimeError):
    pass


_EXC = {TT_EINVAL: ValueError, TT_ELENGTH: ValueError, TT_EOVERFLOW: OverflowError, TT_EDEVICE: RuntimeError}


def declared_symbols():
    """Every tt_* function declared in include/ttgpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tt_[a-z0-9_]+)\s*\(", text)))


class Lib:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise ImportError(f"libttgpu not built: {path} is missing (run __graft_entry__.build())")
        self.path = path
        self.dll = C.CDLL(path)
        for name, (args, res) in _SIGS.items():
            fn = getattr(self.dll, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res if res is not None else C.c_int

    def has(self, name):
        return hasattr(self.dll, name)

    def check(self, rc):
        if rc != TT_OK:
            msg = self.dll.tt_last_error().decode()
            raise _EXC.get(rc, RuntimeError)(msg)

    def call(self, name, *args):
        self.check(getattr(self.dll, name)(*args))

    def last_error(self):
        return self.dll.tt_last_error().decode()


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = Lib()
    return _LIB
