"""ctypes binding of liblgmi.so (the C ABI declared in include/lgmi.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
lightgraphs.jl_amd/csrc``).  There is no CPU fallback: when the library is
missing, or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblgmi.so")

LGMI_OK = 0
LGMI_E_ARG = -1
LGMI_E_BOUNDS = -2
LGMI_E_TOO_LARGE = -3
LGMI_E_NODEV = -4
LGMI_E_ALLOC = -5
LGMI_E_HIP = -6

LGMI_POLICY_HEURISTIC = 0
LGMI_POLICY_LEARNED = 1
LGMI_POLICY_TOPDOWN = 2
LGMI_POLICY_FEATURES = 16
LGMI_POLICY_HIDDEN = 64
LGMI_POLICY_NWEIGHTS = 64 * 16 + 64 + 2 * 64 + 2 + 16 + 16
LGMI_NOALLOC_MAX_NV = 1 << 24
LGMI_NOALLOC_SCRATCH_BYTES = 17 << 20

# Every symbol include/lgmi.h and include/lgmi_gen.h declare.
EXPORTED = [
    "lgmi_graph_create", "lgmi_graph_destroy", "lgmi_gdistances", "lgmi_bfs_parents", "lgmi_bfs",
    "lgmi_bfs_run", "lgmi_bfs_fetch", "lgmi_bfs_run_many", "lgmi_bfs_run_noalloc", "lgmi_bfs_steps", "lgmi_bfs_time_steps", "lgmi_bfs_force_direction",
    "lgmi_bfs_probe", "lgmi_mlp_loss_grad", "lgmi_mlp_train", "lgmi_mlp_forward",
    "lgmi_bfs_set_policy", "lgmi_betweenness", "lgmi_betweenness_stats", "lgmi_betweenness_batch",
    "lgmi_betweenness_batch_tuning",
    "lgmi_betweenness_set_launch", "lgmi_stress", "lgmi_distance_stats",
    "lgmi_connected_components", "lgmi_has_path", "lgmi_period", "lgmi_neighborhood",
    "lgmi_graph_from_edges", "lgmi_graph_size", "lgmi_graph_csr",
    "lgmi_set_tuning", "lgmi_get_tuning",
    "lgmi_last_error", "lgmi_device_count", "lgmi_synchronize",
    "lgmi_gen_kronecker", "lgmi_gen_erdos_renyi", "lgmi_gen_barabasi_albert",
    "lgmi_gen_watts_strogatz", "lgmi_gen_grid2d", "lgmi_csr_build", "lgmi_csr_transpose",
    "lgmi_lg_parse_edges",
]


class LgmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblgmi error {code}: {msg}")
        self.code = code


class NoDeviceError(LgmiError):
    pass


class BfsStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("reached", ctypes.c_int32), ("td_steps", ctypes.c_int32),
                ("bu_steps", ctypes.c_int32), ("device_ms", ctypes.c_double),
                ("edges_reached", ctypes.c_int64), ("launches", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("steps_event_ms", ctypes.c_double)]


class BfsStep(ctypes.Structure):
    _fields_ = [("feat", ctypes.c_float * 16), ("dir", ctypes.c_int32), ("n_front", ctypes.c_int32),
                ("m_front", ctypes.c_int64), ("m_unvisited", ctypes.c_int64), ("n_next", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("scanned", ctypes.c_int64), ("cand", ctypes.c_int64),
                ("kernel_us", ctypes.c_float), ("step_us", ctypes.c_float), ("event_ms", ctypes.c_float),
                ("launch", ctypes.c_int32)]


class ProbeRec(ctypes.Structure):
    _fields_ = [("feat", ctypes.c_float * 16), ("td_us", ctypes.c_float), ("bu_us", ctypes.c_float),
                ("level", ctypes.c_int32), ("n_front", ctypes.c_int32)]


class BcStats(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int32), ("max_levels", ctypes.c_int32), ("launch", ctypes.c_int32),
                ("participants", ctypes.c_int32), ("device_ms", ctypes.c_double),
                ("forward_ms", ctypes.c_double), ("backward_ms", ctypes.c_double), ("blocks", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("wave_row", ctypes.c_int32), ("lane_rounds", ctypes.c_int32), ("range_front", ctypes.c_int32),
                ("queue_front", ctypes.c_int32), ("range_light", ctypes.c_int64), ("queue_min_edges", ctypes.c_int64),
                ("queue_max_edges", ctypes.c_int64), ("bc_seg_len", ctypes.c_int32), ("solo_front", ctypes.c_int32),
                ("solo_edges", ctypes.c_int64)]


_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

_lib = None


def load():
    """Load liblgmi.so (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C lightgraphs.jl_amd/csrc")
    lib = ctypes.CDLL(LIB_PATH)
    c = ctypes
    sig = {
        "lgmi_graph_create": ([c.c_int32, c.c_int64, _i32p, _i32p, c.c_int64, _i32p, _i32p, c.c_int,
                               c.POINTER(_vp)], c.c_int),
        "lgmi_graph_destroy": ([_vp], None),
        "lgmi_gdistances": ([_vp, _i32p, c.c_int32, _i32p], c.c_int),
        "lgmi_bfs_parents": ([_vp, c.c_int32, c.c_int, _i32p], c.c_int),
        "lgmi_bfs": ([_vp, c.c_int32, c.c_int, _i32p, _i32p], c.c_int),
        "lgmi_bfs_run": ([_vp, c.c_int32, c.c_int, c.c_int, c.POINTER(BfsStats)], c.c_int),
        "lgmi_bfs_fetch": ([_vp, _i32p, _i32p], c.c_int),
        "lgmi_bfs_run_many": ([_vp, _i32p, c.c_int32, c.c_int, _vp, _vp], c.c_int),
        "lgmi_bfs_run_noalloc": ([_vp, c.c_int32, c.c_int, _vp, _vp], c.c_int),
        "lgmi_bfs_steps": ([_vp, c.c_int32, c.POINTER(BfsStep)], c.c_int),
        "lgmi_bfs_time_steps": ([_vp, c.c_int], c.c_int),
        "lgmi_bfs_force_direction": ([_vp, c.c_int32], c.c_int),
        "lgmi_bfs_probe": ([_vp, c.c_int32, c.c_int, _i32p, c.c_int32, c.POINTER(ProbeRec)], c.c_int),
        "lgmi_mlp_loss_grad": ([_f32p, _i32p, _f32p, c.c_int32, _f32p, _f32p, _f32p], c.c_int),
        "lgmi_mlp_train": ([_f32p, _i32p, _f32p, c.c_int32, _f32p, c.c_int32, c.c_float, _f32p], c.c_int),
        "lgmi_mlp_forward": ([_f32p, c.c_int32, _f32p, _f32p], c.c_int),
        "lgmi_bfs_set_policy": ([c.c_int, c.c_float, c.c_float, _f32p, c.c_int32], c.c_int),
        "lgmi_betweenness": ([_vp, _i32p, c.c_int32, c.c_int64, c.c_int, c.c_int, c.c_int, _f64p], c.c_int),
        "lgmi_betweenness_stats": ([_vp, c.POINTER(BcStats)], c.c_int),
        "lgmi_betweenness_batch": ([c.c_int32, _i32p, _i64p, _i32p, _i64p, _i32p, _i64p, c.c_int, c.c_int,
                                    _f64p, _f64p], c.c_int),
        "lgmi_betweenness_batch_tuning": ([c.c_int32, c.c_double, c.c_int32, c.c_int32, c.c_int32], c.c_int),
        "lgmi_betweenness_set_launch": ([c.c_int], c.c_int),
        "lgmi_stress": ([_vp, _i32p, c.c_int32, _i64p], c.c_int),
        "lgmi_distance_stats": ([_vp, _i32p, c.c_int32, _i64p], c.c_int),
        "lgmi_connected_components": ([_vp, _i32p], c.c_int),
        "lgmi_has_path": ([_vp, c.c_int32, c.c_int32, _i32p, c.c_int32, c.POINTER(c.c_int)], c.c_int),
        "lgmi_period": ([_vp, _i64p], c.c_int),
        "lgmi_graph_from_edges": ([_i32p, _i32p, c.c_int64, c.c_int, c.c_int32, c.POINTER(_vp), _i64p], c.c_int),
        "lgmi_graph_size": ([_vp, _i32p, _i64p], c.c_int),
        "lgmi_graph_csr": ([_vp, _i32p, _i32p, _i32p, _i32p], c.c_int),
        "lgmi_neighborhood": ([_vp, c.c_int32, c.c_int64, c.c_int, _i32p, _i32p, _i32p], c.c_int),
        "lgmi_set_tuning": ([c.POINTER(Tuning)], c.c_int),
        "lgmi_get_tuning": ([c.POINTER(Tuning)], c.c_int),
        "lgmi_last_error": ([], c.c_char_p),
        "lgmi_device_count": ([], c.c_int),
        "lgmi_synchronize": ([], c.c_int),
        "lgmi_gen_kronecker": ([c.c_int32, c.c_int32, c.c_double, c.c_double, c.c_double, c.c_uint64,
                                _i32p, _i32p], c.c_int64),
        "lgmi_gen_erdos_renyi": ([c.c_int32, c.c_int64, c.c_uint64, _i32p, _i32p], c.c_int64),
        "lgmi_gen_barabasi_albert": ([c.c_int32, c.c_int32, c.c_uint64, _i32p, _i32p], c.c_int64),
        "lgmi_gen_watts_strogatz": ([c.c_int32, c.c_int32, c.c_double, c.c_uint64, _i32p, _i32p], c.c_int64),
        "lgmi_gen_grid2d": ([c.c_int32, c.c_int32, _i32p, _i32p], c.c_int64),
        "lgmi_csr_build": ([c.c_int32, c.c_int64, _i32p, _i32p, c.c_int, c.c_int, _i32p, _i32p, c.c_int64,
                            _i64p], c.c_int),
        "lgmi_csr_transpose": ([c.c_int32, _i32p, _i32p, _i32p, _i32p], c.c_int),
        "lgmi_lg_parse_edges": ([_vp, c.c_int64, c.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p], c.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)  # AttributeError if a declared symbol is missing
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().lgmi_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Map a status code to the exception the reference would raise."""
    if rc == LGMI_OK:
        return
    msg = last_error()
    if rc == LGMI_E_BOUNDS:
        raise IndexError(msg)          # Julia BoundsError
    if rc == LGMI_E_ARG:
        raise ValueError(msg)          # Julia ArgumentError
    if rc == LGMI_E_NODEV:
        raise NoDeviceError(rc, msg)
    raise LgmiError(rc, msg)


def i32p(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


def i64p(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)


def f32p(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32p)


def f64p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_f64p)


class DeviceBuffer:
    """Device memory from the HIP runtime for the entry points that write
    caller-owned device arrays (lgmi_bfs_run_many, lgmi_bfs_run_noalloc)."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        if self._rt().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 4))) != 0:
            raise LgmiError(LGMI_E_ALLOC, f"hipMalloc of {self.nbytes} bytes failed")

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("copy larger than the buffer")
        if self._rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes), 2) != 0:
            raise LgmiError(LGMI_E_HIP, "hipMemcpy device to host failed")
        return out

    def free(self) -> None:
        if self.ptr:
            self._rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
