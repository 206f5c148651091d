"""ORACLE / TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/liboracle_wedge.so``, the CPU restatement of
Wedge's pull engine (see ``oracle/wedge_cpu.h`` for what it restates and the
reference file:line of every part).  Only tests/, ``__graft_entry__.smoke()``
and bench.py's cpu_baseline leg may import this package, and only as the
checker or the CPU baseline -- never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_wedge.so")

WEIGHT_NONE, WEIGHT_F64, WEIGHT_U32 = 0, 1, 2

_LIB = None


class Stats(C.Structure):
    _fields_ = [
        ("running_ms", C.c_double), ("pull_ms", C.c_double), ("frontier_pull_ms", C.c_double),
        ("dense_pull_ms", C.c_double), ("wedge_ms", C.c_double), ("iterations", C.c_uint64),
        ("frontier_iterations", C.c_uint64), ("special", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        lib = C.CDLL(LIB_PATH)
        p = C.c_void_p
        u32p = C.POINTER(C.c_uint32)
        u64p = C.POINTER(C.c_uint64)
        f64p = C.POINTER(C.c_double)
        lib.wcpu_create.restype = p
        lib.wcpu_create.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, p, C.c_int, C.c_int]
        lib.wcpu_destroy.restype = None
        lib.wcpu_destroy.argtypes = [p]
        lib.wcpu_threads.restype = C.c_int
        lib.wcpu_threads.argtypes = [p]
        lib.wcpu_num_vectors.restype = C.c_uint64
        lib.wcpu_num_vectors.argtypes = [p]
        lib.wcpu_outdegrees.restype = C.c_int
        lib.wcpu_outdegrees.argtypes = [p, u64p]
        lib.wcpu_pr.restype = C.c_int
        lib.wcpu_pr.argtypes = [p, C.c_uint32, C.c_double, f64p, C.POINTER(Stats)]
        lib.wcpu_bfs.restype = C.c_int
        lib.wcpu_bfs.argtypes = [p, C.c_uint64, C.c_uint32, u64p, u32p, C.POINTER(Stats)]
        lib.wcpu_cc.restype = C.c_int
        lib.wcpu_cc.argtypes = [p, C.c_uint32, f64p, C.POINTER(Stats)]
        lib.wcpu_sssp.restype = C.c_int
        lib.wcpu_sssp.argtypes = [p, C.c_uint64, C.c_uint32, p, C.POINTER(Stats)]
        lib.wcpu_ppr.restype = C.c_int
        lib.wcpu_ppr.argtypes = [p, u64p, C.c_uint32, C.c_uint32, C.c_double, f64p]
        _LIB = lib
    return _LIB


class Oracle:
    """The reference's CPU pull engine on one graph (offsets/sources/weights arrays)."""

    def __init__(self, offsets, sources, weights=None, threads: int = 0):
        lib = _lib()
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        self.sources = np.ascontiguousarray(sources, dtype=np.uint32)
        self.weights = None if weights is None else np.ascontiguousarray(weights)
        if self.weights is None:
            wk = WEIGHT_NONE
        elif self.weights.dtype == np.float64:
            wk = WEIGHT_F64
        else:
            self.weights = self.weights.astype(np.uint32)
            wk = WEIGHT_U32
        self.weight_kind = wk
        self.V = int(self.offsets.shape[0] - 1)
        self.E = int(self.sources.shape[0])
        self._h = lib.wcpu_create(self.V, self.E, self.offsets.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  self.sources.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  self.weights.ctypes.data if self.weights is not None else None, wk, threads)
        if not self._h:
            raise RuntimeError("wcpu_create failed")

    @classmethod
    def of(cls, graph, threads: int = 0):
        return cls(graph.offsets, graph.sources, graph.weights, threads)

    def close(self):
        if getattr(self, "_h", None):
            _lib().wcpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def threads(self) -> int:
        return int(_lib().wcpu_threads(self._h))

    @property
    def num_vectors(self) -> int:
        return int(_lib().wcpu_num_vectors(self._h))

    def outdegrees(self) -> np.ndarray:
        out = np.empty(self.V, np.uint64)
        _lib().wcpu_outdegrees(self._h, out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def pr(self, iterations: int, damping: float = 0.875):
        out = np.empty(self.V, np.float64)
        st = Stats()
        rc = _lib().wcpu_pr(self._h, iterations, damping, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st))
        if rc:
            raise RuntimeError("wcpu_pr failed")
        return out, st.as_dict()

    def bfs(self, root: int, threshold_pct: int = 20):
        parent = np.empty(self.V, np.uint64)
        depth = np.empty(self.V, np.uint32)
        st = Stats()
        rc = _lib().wcpu_bfs(self._h, root, threshold_pct, parent.ctypes.data_as(C.POINTER(C.c_uint64)),
                             depth.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(st))
        if rc:
            raise RuntimeError("wcpu_bfs failed")
        return parent, depth, st.as_dict()

    def cc(self, threshold_pct: int = 20):
        out = np.empty(self.V, np.float64)
        st = Stats()
        rc = _lib().wcpu_cc(self._h, threshold_pct, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st))
        if rc:
            raise RuntimeError("wcpu_cc failed")
        return out, st.as_dict()

    def sssp(self, root: int, threshold_pct: int = 20):
        out = np.empty(self.V, np.float64 if self.weight_kind == WEIGHT_F64 else np.uint64)
        st = Stats()
        rc = _lib().wcpu_sssp(self._h, root, threshold_pct, out.ctypes.data, C.byref(st))
        if rc:
            raise RuntimeError("wcpu_sssp failed")
        return out, st.as_dict()

    def ppr(self, seeds, iterations: int, damping: float):
        """Personalised PageRank per seed: array [len(seeds)][V] (FP64)."""
        sd = np.ascontiguousarray(seeds, np.uint64)
        out = np.empty((sd.size, self.V), np.float64)
        rc = _lib().wcpu_ppr(self._h, sd.ctypes.data_as(C.POINTER(C.c_uint64)), sd.size, iterations, damping,
                             out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError("wcpu_ppr failed")
        return out
