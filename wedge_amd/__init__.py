"""Python bindings of the MI355X pull engine (ctypes over the C-ABI).

The product is native: ``lib/libwedge_gpu.so`` (HIP kernels for gfx950 behind
``include/wedge_gpu.h``) and ``lib/libwedge.so`` (C host library: ingress,
generators, execution drivers, ``include/wedge_host.h``).  This module only
marshals numpy arrays across that boundary; it never computes results itself,
and it raises if the native libraries are missing.

Mirrors the reference's application surface (stanford-mast/Wedge):
``bfs``, ``cc``, ``pr``, ``sssp`` on a destination-grouped graph, with the
phase-level entry points of include/phases.h available on :class:`Context`.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
GPU_LIB_PATH = os.path.join(LIB_DIR, "libwedge_gpu.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libwedge.so")
CLI_PATH = os.path.join(_HERE, "bin", "wedge")

# status codes (include/wedge_gpu.h)
WG_OK = 0
WG_E_INVAL = -1
WG_E_NOMEM = -2
WG_E_HIP = -3
WG_E_STATE = -4
WG_E_RANGE = -5
WG_E_IO = -6
WG_E_FORMAT = -7
WG_E_OVERFLOW = -8
WG_E_NODEV = -9

WEIGHT_NONE, WEIGHT_F64, WEIGHT_U32 = 0, 1, 2
APP_BFS, APP_CC, APP_PR, APP_SSSP = 1, 2, 3, 4
APP_MSBFS, APP_MSSSSP, APP_PPR = 5, 6, 7
MS_MAX_SOURCES = 64
NOALLOC_MAX_VERTICES = 1 << 27  # WG_NOALLOC_MAX_VERTICES
MODE_DENSE, MODE_FRONTIER = 0, 1
SCHED_PLAIN, SCHED_LEARNED, SCHED_DENSE, SCHED_FRONTIER = 0, 1, 2, 3
DEFAULT_BLOCK_LOG2 = 4
MIN_BLOCK_LOG2 = 3
MAX_BLOCK_LOG2 = 10

KERNELS = {
    "pr_tile": 0, "pr_span": 1, "pr_vertex": 2, "min_tile": 3, "min_blocks": 4,
    "bfs_dense": 5, "bfs_heavy": 6, "bfs_blocks": 7, "wedge_build": 8,
    "advance": 9, "pr_acc": 10, "loop": 11, "batch_pr": 12, "pb_push": 13, "pb_pull": 14,
}

OPT_RELABEL = 1
OPT_PR_BIN_SLICE = 2
OPT_MS_BINNED = 3
OPT_COUNT_WORK = 4
OPT_MIN_BINNED = 5
RELABEL_OFF, RELABEL_ON, RELABEL_AUTO = 0, 1, 2

UNREACHED_U64 = np.uint64(0xFFFFFFFFFFFFFFFF)
UNREACHED_U32 = np.uint32(0xFFFFFFFF)


class WedgeError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = _lib_gpu().wg_error_string(code).decode() if _GPU is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


_GPU = None
_HOST = None


def _lib_gpu():
    global _GPU
    if _GPU is None:
        if not os.path.exists(GPU_LIB_PATH):
            raise RuntimeError(f"native library missing: {GPU_LIB_PATH} (run __graft_entry__.build())")
        _GPU = C.CDLL(GPU_LIB_PATH)
        _declare_gpu(_GPU)
    return _GPU


def _lib_host():
    global _HOST
    if _HOST is None:
        _lib_gpu()
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"native library missing: {HOST_LIB_PATH} (run __graft_entry__.build())")
        _HOST = C.CDLL(HOST_LIB_PATH)
        _declare_host(_HOST)
    return _HOST


_p = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)

# learned schedule model shape (include/wedge_gpu.h)
MLP_IN, MLP_HIDDEN, MLP_ACTIONS = 12, 64, 6
MLP_GATE_LOG2 = 11  # WG_MLP_GATE_LOG2: at activity <= E >> 11 the learned schedule takes the plain frontier pull
MLP_PARAMS = 4 * MLP_IN + MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + \
    MLP_HIDDEN * MLP_ACTIONS + MLP_ACTIONS + 4  # + margins for BFS, CC, SSSP and lv_max


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]

# (name, restype, argtypes) of every C-ABI entry point in include/wedge_gpu.h
GPU_SYMBOLS = [
    ("wg_open", C.c_int, [C.POINTER(_p), C.c_int]),
    ("wg_close", None, [_p]),
    ("wg_error_string", C.c_char_p, [C.c_int]),
    ("wg_last_hip_error", C.c_int, [_p]),
    ("wg_sync", C.c_int, [_p]),
    ("wg_set_option", C.c_int, [_p, C.c_int, C.c_int64]),
    ("wg_load_csc", C.c_int, [_p, C.c_uint64, C.c_uint64, _u32p, _u32p, _p, C.c_int]),
    ("wg_load_edges", C.c_int, [_p, C.c_uint64, C.c_uint64, _u32p, _u32p, _p, C.c_int, C.c_int, _u64p]),
    ("wg_graph_info", C.c_int, [_p, _u64p, _u64p]),
    ("wg_read_outdegrees", C.c_int, [_p, _u32p]),
    ("wg_vertex_outdegree", C.c_int, [_p, C.c_uint64, _u32p]),
    ("wg_init_pr", C.c_int, [_p, C.c_double, C.c_int]),
    ("wg_init_bfs", C.c_int, [_p, C.c_uint64]),
    ("wg_init_cc", C.c_int, [_p]),
    ("wg_init_sssp", C.c_int, [_p, C.c_uint64]),
    ("wg_init_bfs_roots", C.c_int, [_p, _u64p, C.c_uint32]),
    ("wg_init_sssp_roots", C.c_int, [_p, _u64p, C.c_uint32]),
    ("wg_set_members", C.c_int, [_p, C.c_uint32, _u64p]),
    ("wg_work_counts", C.c_int, [_p, _u64p, _u64p]),
    ("wg_validate", C.c_int, [_p, _u64p]),
    ("wg_validate_values", C.c_int, [_p, C.c_int, _u64p, C.c_uint32, _p, _u64p]),
    ("wg_validate_noalloc", C.c_int, [_p, _u64p]),
    ("wg_batch_shape_at", C.c_int, [C.c_uint32, _p]),
    ("wg_batch_pr", C.c_int, [_p, C.c_uint32, C.c_double, C.c_int, _p, C.c_int]),
    ("wg_batch_autotune", C.c_int, [_p, C.c_uint32, C.c_double, C.c_int, _p, _f32p]),
    ("wg_read_batch_ranks", C.c_int, [_p, _f64p]),
    ("wg_set_pr_shape", C.c_int, [_p, _p]),
    ("wg_get_pr_shape", C.c_int, [_p, _p]),
    ("wg_pr_shape_at", C.c_int, [C.c_uint32, _p]),
    ("wg_pr_autotune", C.c_int, [_p, C.c_uint32, _p, _f32p]),
    ("wg_pr_bins_info", C.c_int, [_p, _u64p, _u64p, _u32p, _u32p]),
    ("wg_edge_pull_pr", C.c_int, [_p, _f64p]),
    ("wg_vertex_pr", C.c_int, [_p]),
    ("wg_pr_iteration", C.c_int, [_p, _f64p]),
    ("wg_read_ranks", C.c_int, [_p, _f64p]),
    ("wg_edge_pull_bfs", C.c_int, [_p, C.c_int, _u64p]),
    ("wg_edge_pull_cc", C.c_int, [_p, C.c_int, _u64p]),
    ("wg_edge_pull_sssp", C.c_int, [_p, C.c_int, _u64p]),
    ("wg_last_pull_counts", C.c_int, [_p, _u64p, _u64p, _u64p]),
    ("wg_wedge_generate_frontier", C.c_int, [_p, C.c_int]),
    ("wg_wedge_activate_single_vertex", C.c_int, [_p, C.c_uint64, C.c_int]),
    ("wg_advance_frontier", C.c_int, [_p, _u64p]),
    ("wg_read_bfs", C.c_int, [_p, _u64p]),
    ("wg_read_cc", C.c_int, [_p, _u32p]),
    ("wg_read_sssp", C.c_int, [_p, _p]),
    ("wg_reached", C.c_int, [_p, _u64p, _u64p]),
    ("wg_converge", C.c_int, [_p, _p, _u64p, _u64p]),
    ("wg_set_fused_levels", C.c_int, [_p, C.c_uint32, C.c_uint32]),
    ("wg_fused_level_count", C.c_int, [_p, _u64p]),
    ("wg_dense_pull_binned", C.c_int, [_p, C.POINTER(C.c_int)]),
    ("wg_profile_enable", C.c_int, [_p, C.c_int]),
    ("wg_profile_get", C.c_int, [_p, C.c_int, _u64p, _f64p]),
    ("wg_profile_reset", C.c_int, [_p]),
    ("wg_measure_copy", C.c_int, [_p, C.c_uint64, _f64p]),
    ("wg_msbfs", C.c_int, [_p, _u64p, C.c_uint32, _p, _u64p, _u64p]),
    ("wg_read_msbfs", C.c_int, [_p, C.c_uint32, _u32p]),
    ("wg_mssssp", C.c_int, [_p, _u64p, C.c_uint32, _p, _u64p]),
    ("wg_read_mssssp", C.c_int, [_p, C.c_uint32, _p]),
    ("wg_ppr", C.c_int, [_p, _u64p, C.c_uint32, C.c_uint32, C.c_double]),
    ("wg_read_ppr", C.c_int, [_p, C.c_uint32, _f64p]),
    ("wg_mlp_action", C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("wg_set_schedule_model", C.c_int, [_p, _f32p, C.c_uint64]),
    ("wg_mlp_infer", C.c_int, [_p, C.c_int, _f32p, _f32p, _i32p, C.c_uint64, _f32p, _i32p]),
    ("wg_mlp_train", C.c_int, [_p, _f32p, _i32p, _f32p, C.c_uint64, C.POINTER(MlpTrainOpts), _f32p, _f32p]),
]


class BatchShape(C.Structure):
    _fields_ = [("edges_per_lane", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("xcd_split", C.c_uint32)]

    def as_tuple(self):
        return int(self.edges_per_lane), int(self.blocks_per_cu), int(self.xcd_split)


BATCH_NSHAPES = 24
PR_NSHAPES = BATCH_NSHAPES + 8
PR_BINNED = 2  # xcd_split value of the binned PageRank pull


class RunOpts(C.Structure):
    _fields_ = [
        ("schedule", C.c_int), ("threshold_pct", C.c_uint32), ("block_log2", C.c_int),
        ("iterations", C.c_uint32), ("damping", C.c_double), ("contrib_fp32", C.c_int),
        ("root", C.c_uint64), ("record_trace", C.c_int), ("max_iterations", C.c_uint32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("running_ms", C.c_double), ("pull_ms", C.c_double), ("frontier_pull_ms", C.c_double),
        ("dense_pull_ms", C.c_double), ("wedge_ms", C.c_double), ("iterations", C.c_uint64),
        ("frontier_iterations", C.c_uint64), ("edges_reached", C.c_uint64),
        ("vertices_reached", C.c_uint64), ("special", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LoopParams(C.Structure):
    _fields_ = [("schedule", C.c_int), ("threshold_pct", C.c_uint32), ("block_log2", C.c_int),
                ("max_iterations", C.c_uint32)]


class TraceRow(C.Structure):
    _fields_ = [
        ("iteration", C.c_uint32), ("mode", C.c_uint32), ("block_log2", C.c_int32), ("pad", C.c_uint32),
        ("frontier_vertices", C.c_uint64), ("frontier_edges", C.c_uint64), ("unvisited", C.c_uint64),
        ("pull_ms", C.c_double), ("wedge_ms", C.c_double), ("gpu_pull_ms", C.c_double), ("gpu_build_ms", C.c_double),
    ]


class GraphStruct(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint64), ("num_edges", C.c_uint64), ("offsets", _u32p),
        ("sources", _u32p), ("weights", _p), ("weight_kind", C.c_int),
    ]


HOST_SYMBOLS = [
    ("wg_host_threads", C.c_int, []),
    ("wg_graph_rmat", C.c_int, [C.POINTER(GraphStruct), C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_uint64, C.c_int, C.c_int]),
    ("wg_graph_grid", C.c_int, [C.POINTER(GraphStruct), C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                C.c_uint32, C.c_uint64]),
    ("wg_graph_read", C.c_int, [C.POINTER(GraphStruct), C.c_char_p, C.c_int]),
    ("wg_graph_write", C.c_int, [C.POINTER(GraphStruct), C.c_char_p]),
    ("wg_graph_free", None, [C.POINTER(GraphStruct)]),
    ("wg_graph_union", C.c_int, [C.POINTER(GraphStruct), C.c_uint32, C.POINTER(GraphStruct), _u64p]),
    ("wg_graph_outdegrees", C.c_int, [C.POINTER(GraphStruct), _u32p]),
    ("wg_graph_max_outdegree_vertex", C.c_uint64, [C.POINTER(GraphStruct)]),
    ("wg_weight_f64", C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
    ("wg_rand64", C.c_uint64, [C.c_uint64, C.c_uint64]),
    ("wg_run_opts_default", None, [C.POINTER(RunOpts)]),
    ("wg_run_pr", C.c_int, [_p, C.POINTER(RunOpts), _f64p, C.POINTER(RunStats)]),
    ("wg_run_bfs", C.c_int, [_p, C.POINTER(RunOpts), _u64p, C.POINTER(RunStats)]),
    ("wg_run_cc", C.c_int, [_p, C.POINTER(RunOpts), _u32p, C.POINTER(RunStats)]),
    ("wg_run_sssp", C.c_int, [_p, C.POINTER(RunOpts), _p, C.POINTER(RunStats)]),
    ("wg_trace_count", C.c_uint64, []),
    ("wg_trace_copy", C.c_int, [C.POINTER(TraceRow), C.c_uint64]),
    ("wg_mlp_load", C.c_int, [C.c_char_p]),
    ("wg_mlp_save", C.c_int, [C.c_char_p]),
    ("wg_mlp_set", C.c_int, [_f32p, C.c_uint64]),
    ("wg_mlp_get", C.c_int, [_f32p, C.c_uint64]),
    ("wg_mlp_unload", None, []),
    ("wg_mlp_loaded", C.c_int, []),
    ("wg_mlp_feature_vector", C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                        C.c_uint64, _f32p]),
    ("wg_mlp_baseline", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, _f32p]),
    ("wg_mlp_predict", C.c_int, [C.c_int, _f32p, _f32p, C.c_int, _f32p]),
    ("wg_ctx_load_graph", C.c_int, [_p, C.POINTER(GraphStruct)]),
]


def _declare(lib, table):
    for name, res, args in table:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _declare_gpu(lib):
    _declare(lib, GPU_SYMBOLS)


def _declare_host(lib):
    _declare(lib, HOST_SYMBOLS)


def _check(rc, what=""):
    if rc != WG_OK:
        raise WedgeError(rc, what)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


# ---- graphs -----------------------------------------------------------------------


@dataclass
class Graph:
    """Destination-grouped CSC with 32-bit indices (numpy arrays owned by Python)."""

    offsets: np.ndarray  # uint32[V+1]
    sources: np.ndarray  # uint32[E], ascending within each destination
    weights: np.ndarray | None = None  # float64[E] or uint32[E]

    @property
    def num_vertices(self) -> int:
        return int(self.offsets.shape[0] - 1)

    @property
    def num_edges(self) -> int:
        return int(self.sources.shape[0])

    @property
    def weight_kind(self) -> int:
        if self.weights is None:
            return WEIGHT_NONE
        return WEIGHT_F64 if self.weights.dtype == np.float64 else WEIGHT_U32

    def destinations(self) -> np.ndarray:
        """Destination of every edge (uint32[E])."""
        deg = np.diff(self.offsets.astype(np.int64))
        return np.repeat(np.arange(self.num_vertices, dtype=np.uint32), deg)

    def outdegrees(self) -> np.ndarray:
        return np.bincount(self.sources, minlength=self.num_vertices).astype(np.uint32)

    def max_outdegree_vertex(self) -> int:
        return int(np.argmax(self.outdegrees()))

    def _struct(self) -> GraphStruct:
        s = GraphStruct()
        s.num_vertices = self.num_vertices
        s.num_edges = self.num_edges
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.uint32)
        self.sources = np.ascontiguousarray(self.sources, dtype=np.uint32)
        s.offsets = _ptr(self.offsets, C.c_uint32)
        s.sources = _ptr(self.sources, C.c_uint32)
        s.weights = self.weights.ctypes.data if self.weights is not None else None
        s.weight_kind = self.weight_kind
        return s

    def write(self, path: str) -> None:
        """Writes the reference's binary edge list (u64 V, u64 E, (src, dst[, w]) records)."""
        st = self._struct()
        _check(_lib_host().wg_graph_write(C.byref(st), os.fsencode(path)), "wg_graph_write")


def _adopt(st: GraphStruct) -> Graph:
    """Copies a C-owned graph into numpy arrays and frees the C copy."""
    V, E = int(st.num_vertices), int(st.num_edges)
    off = np.ctypeslib.as_array(st.offsets, shape=(V + 1,)).copy()
    src = np.ctypeslib.as_array(st.sources, shape=(E,)).copy() if E else np.zeros(0, np.uint32)
    w = None
    if st.weight_kind == WEIGHT_F64:
        w = np.ctypeslib.as_array(C.cast(st.weights, _f64p), shape=(E,)).copy() if E else np.zeros(0, np.float64)
    elif st.weight_kind == WEIGHT_U32:
        w = np.ctypeslib.as_array(C.cast(st.weights, _u32p), shape=(E,)).copy() if E else np.zeros(0, np.uint32)
    _lib_host().wg_graph_free(C.byref(st))
    return Graph(off, src, w)


def rmat(scale: int, edge_factor: int = 16, seed: int = 1, a=0.57, b=0.19, c=0.19, symmetrize=False,
         weights: str | None = None) -> Graph:
    """Seeded R-MAT graph, self-loops and duplicates removed, grouped by destination.

    weights='f64' attaches the FP64 weight in (0, 1] derived from (src, dst, seed)."""
    st = GraphStruct()
    wk = WEIGHT_F64 if weights == "f64" else WEIGHT_NONE
    _check(_lib_host().wg_graph_rmat(C.byref(st), scale, edge_factor, a, b, c, seed, int(symmetrize), wk),
           "wg_graph_rmat")
    return _adopt(st)


def grid(rows: int, cols: int, seed: int = 3, p_keep=0.6, p_diag=0.05, wmax=1000) -> Graph:
    """Road-like grid with integer weights in [1, wmax] (both directions stored)."""
    st = GraphStruct()
    _check(_lib_host().wg_graph_grid(C.byref(st), rows, cols, p_keep, p_diag, wmax, seed), "wg_graph_grid")
    return _adopt(st)


def read_graph(path: str, weights: str | None = None) -> Graph:
    """Reads the reference's binary edge list; weights None | 'int' | 'f64'."""
    wk = {None: WEIGHT_NONE, "int": WEIGHT_U32, "f64": WEIGHT_F64}[weights]
    st = GraphStruct()
    _check(_lib_host().wg_graph_read(C.byref(st), os.fsencode(path), wk), "wg_graph_read")
    return _adopt(st)


def from_edges(num_vertices: int, src, dst, weights=None, dedup=False) -> Graph:
    """Builds a Graph from edge arrays (sorted by (dst, src); host-side convenience)."""
    src = np.asarray(src, dtype=np.uint64)
    dst = np.asarray(dst, dtype=np.uint64)
    order = np.lexsort((src, dst))
    src, dst = src[order], dst[order]
    w = None if weights is None else np.asarray(weights)[order]
    if dedup:
        keep = np.ones(src.shape[0], bool)
        keep[1:] = (src[1:] != src[:-1]) | (dst[1:] != dst[:-1])
        keep &= src != dst
        src, dst = src[keep], dst[keep]
        if w is not None:
            w = w[keep]
    counts = np.bincount(dst.astype(np.int64), minlength=num_vertices)
    off = np.zeros(num_vertices + 1, np.uint32)
    np.cumsum(counts, out=off[1:])
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float64 if np.asarray(w).dtype.kind == "f" else np.uint32)
    return Graph(off, src.astype(np.uint32), w)


def union(graphs) -> tuple[Graph, np.ndarray]:
    """Disjoint union of member graphs (wg_graph_union): member g's vertex v
    becomes first[g] + v.  Returns the union and first (uint64[n + 1]).  All
    members must carry the same weight kind."""
    graphs = list(graphs)
    if not graphs:
        raise ValueError("union of no graphs")
    if len({g.weight_kind for g in graphs}) != 1:
        raise ValueError("members differ in weight kind")
    arr = (GraphStruct * len(graphs))()
    for i, g in enumerate(graphs):
        arr[i] = g._struct()  # keeps g's arrays contiguous and referenced by g
    first = np.zeros(len(graphs) + 1, np.uint64)
    st = GraphStruct()
    _check(_lib_host().wg_graph_union(arr, len(graphs), C.byref(st), _ptr(first, C.c_uint64)), "wg_graph_union")
    return _adopt(st), first


def batch_shape_at(i: int) -> tuple[int, int]:
    s = BatchShape()
    _check(_lib_gpu().wg_batch_shape_at(i, C.byref(s)), "wg_batch_shape_at")
    return s.as_tuple()


def pr_shape_at(i: int) -> tuple[int, int, int]:
    s = BatchShape()
    _check(_lib_gpu().wg_pr_shape_at(i, C.byref(s)), "wg_pr_shape_at")
    return s.as_tuple()


def weight_f64(src: int, dst: int, seed: int) -> float:
    return float(_lib_host().wg_weight_f64(src, dst, seed))


def rand64(key: int, counter: int) -> int:
    return int(_lib_host().wg_rand64(key, counter))


def host_threads() -> int:
    return int(_lib_host().wg_host_threads())


# ---- device context ---------------------------------------------------------------


SCHEDULES = {"plain": 0, "learned": 1, "dense": 2, "frontier": 3}


def run_opts(**kw) -> RunOpts:
    o = RunOpts()
    _lib_host().wg_run_opts_default(C.byref(o))
    sched = kw.pop("schedule", None)
    if isinstance(sched, str):
        kw["schedule"] = {"plain": SCHED_PLAIN, "learned": SCHED_LEARNED, "dense": SCHED_DENSE,
                          "frontier": SCHED_FRONTIER}[sched]
    elif sched is not None:
        kw["schedule"] = sched
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class Context:
    """One HIP device context holding one graph (include/wedge_gpu.h)."""

    def __init__(self, device: int = 0, relabel="auto"):
        self._lib = _lib_gpu()
        _lib_host()
        h = _p()
        _check(self._lib.wg_open(C.byref(h), device), "wg_open")
        self._h = h
        self.graph: Graph | None = None
        self.set_option(OPT_RELABEL, RELABEL_AUTO if relabel == "auto" else int(bool(relabel)))

    def set_option(self, option: int, value: int):
        """Context option (takes effect at the next load)."""
        self._chk(self._lib.wg_set_option(self._h, option, value), "wg_set_option")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def hip_error(self) -> int:
        return int(self._lib.wg_last_hip_error(self._h))

    def _chk(self, rc, what):
        if rc != WG_OK:
            raise WedgeError(rc, f"{what} (hip {self.hip_error()})")

    def load(self, g: Graph) -> None:
        off = np.ascontiguousarray(g.offsets, dtype=np.uint32)
        src = np.ascontiguousarray(g.sources, dtype=np.uint32)
        w = g.weights
        self._chk(self._lib.wg_load_csc(self._h, g.num_vertices, g.num_edges, _ptr(off, C.c_uint32),
                                        _ptr(src, C.c_uint32), w.ctypes.data if w is not None else None,
                                        g.weight_kind), "wg_load_csc")
        self.graph = g
        self._wkind = g.weight_kind

    def load_edges(self, num_vertices: int, src, dst, weights=None, dedup=True, drop_self_loops=True) -> int:
        """Builds the graph on the device from edges in any order (wg_load_edges);
        returns the number of edges kept.  weights: float64 or uint32 per edge."""
        s = np.ascontiguousarray(src, np.uint32)
        d = np.ascontiguousarray(dst, np.uint32)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        w, wk = None, WEIGHT_NONE
        if weights is not None:
            w = np.ascontiguousarray(weights)
            if w.dtype == np.float64:
                wk = WEIGHT_F64
            else:
                w = np.ascontiguousarray(w, np.uint32)
                wk = WEIGHT_U32
        flags = (2 if dedup else 0) | (1 if drop_self_loops else 0)
        kept = C.c_uint64()
        self._chk(self._lib.wg_load_edges(self._h, num_vertices, s.size, _ptr(s, C.c_uint32), _ptr(d, C.c_uint32),
                                          w.ctypes.data if w is not None else None, wk, flags, C.byref(kept)),
                  "wg_load_edges")
        self.graph = None
        self._wkind = wk
        return int(kept.value)

    def sync(self):
        self._chk(self._lib.wg_sync(self._h), "wg_sync")

    @property
    def V(self) -> int:
        v = C.c_uint64()
        self._lib.wg_graph_info(self._h, C.byref(v), None)
        return int(v.value)

    def outdegrees(self) -> np.ndarray:
        out = np.empty(self.V, np.uint32)
        self._chk(self._lib.wg_read_outdegrees(self._h, _ptr(out, C.c_uint32)), "wg_read_outdegrees")
        return out

    # -- phase-level API (reference include/phases.h granularity) --
    def init_pr(self, damping=0.875, contrib_fp32=False):
        self._chk(self._lib.wg_init_pr(self._h, damping, int(contrib_fp32)), "wg_init_pr")

    def init_bfs(self, root: int):
        self._chk(self._lib.wg_init_bfs(self._h, root), "wg_init_bfs")

    def init_bfs_roots(self, roots):
        r = np.ascontiguousarray(roots, np.uint64)
        self._chk(self._lib.wg_init_bfs_roots(self._h, _ptr(r, C.c_uint64), r.size), "wg_init_bfs_roots")

    def init_sssp_roots(self, roots):
        r = np.ascontiguousarray(roots, np.uint64)
        self._chk(self._lib.wg_init_sssp_roots(self._h, _ptr(r, C.c_uint64), r.size), "wg_init_sssp_roots")

    def work_counts(self) -> tuple[int, int]:
        """(edges read, destination states examined) by the pulls since the init
        (needs set_option(OPT_COUNT_WORK, 1) before the run)."""
        e, d = C.c_uint64(), C.c_uint64()
        self._chk(self._lib.wg_work_counts(self._h, C.byref(e), C.byref(d)), "wg_work_counts")
        return int(e.value), int(d.value)

    def validate(self) -> int:
        """Violations of the current BFS / CC / SSSP result (0 = correct)."""
        n = C.c_uint64()
        self._chk(self._lib.wg_validate(self._h, C.byref(n)), "wg_validate")
        return int(n.value)

    def validate_values(self, app: int, values: np.ndarray, roots=()) -> int:
        """Violations of a result computed elsewhere (wg_read_* layouts, caller's numbering)."""
        r = np.ascontiguousarray(roots, np.uint64)
        v = np.ascontiguousarray(values)
        n = C.c_uint64()
        self._chk(self._lib.wg_validate_values(self._h, app, _ptr(r, C.c_uint64) if r.size else None, r.size,
                                               v.ctypes.data, C.byref(n)), "wg_validate_values")
        return int(n.value)

    def load_batch(self, graphs) -> np.ndarray:
        """Loads member graphs as one disjoint union (relabelling off, so every
        member keeps its vertex range) and registers the members.  Returns first."""
        g, first = union(graphs)
        self.set_option(OPT_RELABEL, RELABEL_OFF)
        self.load(g)
        self.set_members(first)
        return first

    def set_members(self, first):
        f = np.ascontiguousarray(first, np.uint64)
        self._chk(self._lib.wg_set_members(self._h, f.size - 1, _ptr(f, C.c_uint64)), "wg_set_members")
        self.members = f

    def batch_pr(self, iterations=16, damping=0.85, contrib_fp32=False, shape=None, use_graph=True):
        """PageRank on every member (queued; read_batch_ranks waits).  shape:
        (edges_per_lane, blocks_per_cu[, xcd_split]) or None for the tuned one."""
        s = None if shape is None else C.byref(BatchShape(*shape))
        self._chk(self._lib.wg_batch_pr(self._h, iterations, damping, int(contrib_fp32), s, int(use_graph)),
                  "wg_batch_pr")

    def batch_autotune(self, iterations=16, damping=0.85, contrib_fp32=False):
        """Times every candidate shape; returns (best shape, {shape: ms})."""
        best = BatchShape()
        ms = np.zeros(BATCH_NSHAPES, np.float32)
        self._chk(self._lib.wg_batch_autotune(self._h, iterations, damping, int(contrib_fp32), C.byref(best),
                                              _ptr(ms, C.c_float)), "wg_batch_autotune")
        return best.as_tuple(), {batch_shape_at(i): float(ms[i]) for i in range(BATCH_NSHAPES)}

    def set_pr_shape(self, shape=None):
        """Tile shape of the single-graph PageRank: (edges_per_lane, blocks_per_cu[, xcd_split]) or None."""
        s = None if shape is None else C.byref(BatchShape(*shape))
        self._chk(self._lib.wg_set_pr_shape(self._h, s), "wg_set_pr_shape")

    def pr_shape(self):
        s = BatchShape()
        self._chk(self._lib.wg_get_pr_shape(self._h, C.byref(s)), "wg_get_pr_shape")
        return s.as_tuple()

    def pr_autotune(self, iterations=4):
        """After init_pr: picks the fastest PageRank tile shape and re-initialises;
        returns (best shape, {shape: ms per iteration})."""
        best = BatchShape()
        ms = np.zeros(PR_NSHAPES, np.float32)
        self._chk(self._lib.wg_pr_autotune(self._h, iterations, C.byref(best), _ptr(ms, C.c_float)), "wg_pr_autotune")
        return best.as_tuple(), {pr_shape_at(i): float(ms[i]) for i in range(PR_NSHAPES)}

    def pr_bins_info(self) -> dict:
        """Size of the binned PageRank layout (after init_pr with a binned shape)."""
        pairs, q = C.c_uint64(), C.c_uint64()
        push, pull = C.c_uint32(), C.c_uint32()
        self._chk(self._lib.wg_pr_bins_info(self._h, C.byref(pairs), C.byref(q), C.byref(push), C.byref(pull)),
                  "wg_pr_bins_info")
        return {"pairs": int(pairs.value), "code_entries": int(q.value), "push_groups": int(push.value),
                "pull_groups": int(pull.value)}

    def read_batch_ranks(self) -> np.ndarray:
        out = np.empty(self.V, np.float64)
        self._chk(self._lib.wg_read_batch_ranks(self._h, _ptr(out, C.c_double)), "wg_read_batch_ranks")
        return out

    def init_cc(self):
        self._chk(self._lib.wg_init_cc(self._h), "wg_init_cc")

    def init_sssp(self, root: int):
        self._chk(self._lib.wg_init_sssp(self._h, root), "wg_init_sssp")

    def pr_iteration(self) -> float:
        t = C.c_double()
        self._chk(self._lib.wg_pr_iteration(self._h, C.byref(t)), "wg_pr_iteration")
        return float(t.value)

    def edge_pull_pr(self) -> float:
        t = C.c_double()
        self._chk(self._lib.wg_edge_pull_pr(self._h, C.byref(t)), "wg_edge_pull_pr")
        return float(t.value)

    def vertex_pr(self):
        self._chk(self._lib.wg_vertex_pr(self._h), "wg_vertex_pr")

    def read_ranks(self) -> np.ndarray:
        out = np.empty(self.V, np.float64)
        self._chk(self._lib.wg_read_ranks(self._h, _ptr(out, C.c_double)), "wg_read_ranks")
        return out

    def edge_pull(self, app: int, mode: int) -> int:
        a = C.c_uint64()
        fn = {APP_BFS: self._lib.wg_edge_pull_bfs, APP_CC: self._lib.wg_edge_pull_cc,
              APP_SSSP: self._lib.wg_edge_pull_sssp}[app]
        self._chk(fn(self._h, mode, C.byref(a)), "wg_edge_pull")
        return int(a.value)

    def last_pull_counts(self):
        """(edges activated, vertices activated, vertices reached for the first time) of the last pull."""
        e, v, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._chk(self._lib.wg_last_pull_counts(self._h, C.byref(e), C.byref(v), C.byref(f)), "wg_last_pull_counts")
        return int(e.value), int(v.value), int(f.value)

    def wedge_generate_frontier(self, block_log2=DEFAULT_BLOCK_LOG2):
        self._chk(self._lib.wg_wedge_generate_frontier(self._h, block_log2), "wg_wedge_generate_frontier")

    def wedge_activate_single_vertex(self, v: int, block_log2=DEFAULT_BLOCK_LOG2):
        self._chk(self._lib.wg_wedge_activate_single_vertex(self._h, v, block_log2),
                  "wg_wedge_activate_single_vertex")

    def advance_frontier(self) -> int:
        n = C.c_uint64()
        self._chk(self._lib.wg_advance_frontier(self._h, C.byref(n)), "wg_advance_frontier")
        return int(n.value)

    def converge(self, schedule=SCHED_PLAIN, threshold_pct=20, block_log2=DEFAULT_BLOCK_LOG2, max_iterations=0):
        """Runs the initialised BFS/CC/SSSP to convergence in the device-resident loop."""
        p = LoopParams(schedule, threshold_pct, block_log2, max_iterations)
        it, fit = C.c_uint64(), C.c_uint64()
        self._chk(self._lib.wg_converge(self._h, C.byref(p), C.byref(it), C.byref(fit)), "wg_converge")
        return int(it.value), int(fit.value)

    def set_fused_levels(self, workgroups=32, apps=(APP_BFS,)):
        """Workgroups of the fused small-frontier levels of converge (0 = off)
        and the applications that use them."""
        mask = 0
        for a in apps:
            mask |= 1 << int(a)
        self._chk(self._lib.wg_set_fused_levels(self._h, int(workgroups), mask), "wg_set_fused_levels")

    def fused_level_count(self) -> int:
        """Iterations of the last converge run by fused small-frontier levels."""
        n = C.c_uint64()
        self._chk(self._lib.wg_fused_level_count(self._h, C.byref(n)), "wg_fused_level_count")
        return int(n.value)

    def dense_pull_binned(self) -> bool:
        """Whether the initialised CC / SSSP's dense pulls run on the binned layout."""
        on = C.c_int()
        self._chk(self._lib.wg_dense_pull_binned(self._h, C.byref(on)), "wg_dense_pull_binned")
        return bool(on.value)

    def read_bfs(self) -> np.ndarray:
        out = np.empty(self.V, np.uint64)
        self._chk(self._lib.wg_read_bfs(self._h, _ptr(out, C.c_uint64)), "wg_read_bfs")
        return out

    def read_cc(self) -> np.ndarray:
        out = np.empty(self.V, np.uint32)
        self._chk(self._lib.wg_read_cc(self._h, _ptr(out, C.c_uint32)), "wg_read_cc")
        return out

    def read_sssp(self) -> np.ndarray:
        f64 = getattr(self, "_wkind", WEIGHT_NONE) == WEIGHT_F64
        out = np.empty(self.V, np.float64 if f64 else np.uint32)
        self._chk(self._lib.wg_read_sssp(self._h, out.ctypes.data), "wg_read_sssp")
        return out

    def reached(self):
        v, e = C.c_uint64(), C.c_uint64()
        self._chk(self._lib.wg_reached(self._h, C.byref(v), C.byref(e)), "wg_reached")
        return int(v.value), int(e.value)

    # -- profiling --
    def profile(self, on=True):
        self._chk(self._lib.wg_profile_enable(self._h, int(on)), "wg_profile_enable")

    def measure_copy(self, bytes_=4 << 30) -> float:
        """GB/s of a device-to-device streaming copy (read + write)."""
        g = C.c_double()
        self._chk(self._lib.wg_measure_copy(self._h, bytes_, C.byref(g)), "wg_measure_copy")
        return float(g.value)

    def profile_reset(self):
        self._chk(self._lib.wg_profile_reset(self._h), "wg_profile_reset")

    def profile_get(self, kernel: str):
        n, ms = C.c_uint64(), C.c_double()
        self._chk(self._lib.wg_profile_get(self._h, KERNELS[kernel], C.byref(n), C.byref(ms)), "wg_profile_get")
        return int(n.value), float(ms.value)

    # -- whole-application drivers (reference execution_impl_*) --
    def run_pr(self, **kw):
        o = run_opts(**kw)
        out = np.empty(self.V, np.float64)
        st = RunStats()
        self._chk(_lib_host().wg_run_pr(self._h, C.byref(o), _ptr(out, C.c_double), C.byref(st)), "wg_run_pr")
        return out, st.as_dict()

    def run_bfs(self, root: int, **kw):
        o = run_opts(root=root, **kw)
        out = np.empty(self.V, np.uint64)
        st = RunStats()
        self._chk(_lib_host().wg_run_bfs(self._h, C.byref(o), _ptr(out, C.c_uint64), C.byref(st)), "wg_run_bfs")
        return out, st.as_dict()

    def run_cc(self, **kw):
        o = run_opts(**kw)
        out = np.empty(self.V, np.uint32)
        st = RunStats()
        self._chk(_lib_host().wg_run_cc(self._h, C.byref(o), _ptr(out, C.c_uint32), C.byref(st)), "wg_run_cc")
        return out, st.as_dict()

    def msbfs(self, roots, schedule="plain", threshold_pct=2, block_log2=0, max_iterations=0):
        """BFS from up to 64 roots in one traversal; returns (levels, edges reached over all roots).
        Depths per root: read_msbfs(i).  The default threshold is far below the single-source
        rule's 20 %: a batched frontier level gathers a 64-bit mask per frontier edge for every
        source at once, while a dense level streams the binned layout, so the dense level is
        cheaper from a small active share on (rmat-s24, 64 roots: 5.9 ms at 1-2 %, 6.3 ms at 5 %,
        8.3 ms at 10-40 %)."""
        r = np.ascontiguousarray(roots, np.uint64)
        p = LoopParams(schedule=SCHEDULES[schedule] if isinstance(schedule, str) else schedule,
                       threshold_pct=threshold_pct, block_log2=block_log2,
                       max_iterations=max_iterations)
        lv, er = C.c_uint64(), C.c_uint64()
        self._chk(self._lib.wg_msbfs(self._h, _ptr(r, C.c_uint64), r.size, C.byref(p), C.byref(lv), C.byref(er)),
                  "wg_msbfs")
        return int(lv.value), int(er.value)

    def mssssp(self, roots, schedule="plain", threshold_pct=20, max_iterations=0):
        """SSSP from up to 64 roots in one traversal; returns the iterations run.
        Distances per root: read_mssssp(i)."""
        r = np.ascontiguousarray(roots, np.uint64)
        p = LoopParams(schedule=SCHEDULES[schedule] if isinstance(schedule, str) else schedule,
                       threshold_pct=threshold_pct, block_log2=0, max_iterations=max_iterations)
        it = C.c_uint64()
        self._chk(self._lib.wg_mssssp(self._h, _ptr(r, C.c_uint64), r.size, C.byref(p), C.byref(it)), "wg_mssssp")
        return int(it.value)

    def read_mssssp(self, source: int) -> np.ndarray:
        f64 = getattr(self, "_wkind", WEIGHT_NONE) == WEIGHT_F64
        out = np.empty(self.V, np.float64 if f64 else np.uint32)
        self._chk(self._lib.wg_read_mssssp(self._h, source, out.ctypes.data), "wg_read_mssssp")
        return out

    def ppr(self, seeds, iterations=16, damping=0.85):
        """Personalised PageRank for up to 32 seeds; ranks per seed: read_ppr(i)."""
        s = np.ascontiguousarray(seeds, np.uint64)
        self._chk(self._lib.wg_ppr(self._h, _ptr(s, C.c_uint64), s.size, iterations, damping), "wg_ppr")

    def read_ppr(self, seed: int) -> np.ndarray:
        out = np.empty(self.V, np.float64)
        self._chk(self._lib.wg_read_ppr(self._h, seed, _ptr(out, C.c_double)), "wg_read_ppr")
        return out

    def read_msbfs(self, source: int) -> np.ndarray:
        out = np.empty(self.V, np.uint32)
        self._chk(self._lib.wg_read_msbfs(self._h, source, _ptr(out, C.c_uint32)), "wg_read_msbfs")
        return out

    def set_schedule_model(self, params: np.ndarray | None):
        """Model for the device loop's learned schedule (None clears it)."""
        if params is None:
            self._chk(self._lib.wg_set_schedule_model(self._h, None, 0), "wg_set_schedule_model")
            return
        p = np.ascontiguousarray(params, np.float32)
        self._chk(self._lib.wg_set_schedule_model(self._h, _ptr(p, C.c_float), p.size), "wg_set_schedule_model")

    def mlp_infer(self, X: np.ndarray, app: int = APP_BFS, base: np.ndarray | None = None,
                  plain: np.ndarray | None = None):
        """Device batch inference: predictions [n][A] and the actions chosen for app [n]."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
        n = X.shape[0]
        Y = np.empty((n, MLP_ACTIONS), np.float32)
        A = np.empty(n, np.int32)
        bp = None
        if base is not None:
            base = np.ascontiguousarray(base, np.float32).reshape(n, MLP_ACTIONS)
            bp = _ptr(base, C.c_float)
        pp = None
        if plain is not None:
            plain = np.ascontiguousarray(plain, np.int32)
            pp = _ptr(plain, C.c_int32)
        self._chk(self._lib.wg_mlp_infer(self._h, app, _ptr(X, C.c_float), bp, pp, n, _ptr(Y, C.c_float),
                                         _ptr(A, C.c_int32)), "wg_mlp_infer")
        return Y, A

    def mlp_train(self, X, actions, targets, params, epochs=200, batch=512, lr=1e-3, seed=1,
                  beta1=0.9, beta2=0.999, eps=1e-8) -> tuple[np.ndarray, float]:
        """Device training (hand-written Adam kernels); returns (params, mean squared error)."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
        A = np.ascontiguousarray(actions, np.int32)
        T = np.ascontiguousarray(targets, np.float32)
        P = np.ascontiguousarray(params, np.float32).copy()
        o = MlpTrainOpts(epochs=epochs, batch=batch, lr=lr, beta1=beta1, beta2=beta2, eps=eps, seed=seed)
        loss = C.c_float()
        self._chk(self._lib.wg_mlp_train(self._h, _ptr(X, C.c_float), _ptr(A, C.c_int32), _ptr(T, C.c_float),
                                         X.shape[0], C.byref(o), _ptr(P, C.c_float), C.byref(loss)), "wg_mlp_train")
        return P, float(loss.value)

    def run_sssp(self, root: int, **kw):
        o = run_opts(root=root, **kw)
        f64 = getattr(self, "_wkind", WEIGHT_NONE) == WEIGHT_F64
        out = np.empty(self.V, np.float64 if f64 else np.uint32)
        st = RunStats()
        self._chk(_lib_host().wg_run_sssp(self._h, C.byref(o), out.ctypes.data, C.byref(st)), "wg_run_sssp")
        return out, st.as_dict()


# ---- learned schedule ------------------------------------------------------------------

DEFAULT_MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "schedule_mlp.txt")


def mlp_action(a: int) -> tuple[int, int]:
    """(mode, block_log2) of schedule action a."""
    m, b = C.c_int(), C.c_int()
    _check(_lib_gpu().wg_mlp_action(a, C.byref(m), C.byref(b)), "wg_mlp_action")
    return m.value, b.value


def mlp_set(params: np.ndarray):
    p = np.ascontiguousarray(params, np.float32)
    _check(_lib_host().wg_mlp_set(_ptr(p, C.c_float), p.size), "wg_mlp_set")


def mlp_get() -> np.ndarray:
    p = np.empty(MLP_PARAMS, np.float32)
    _check(_lib_host().wg_mlp_get(_ptr(p, C.c_float), p.size), "wg_mlp_get")
    return p


def mlp_load(path: str = DEFAULT_MODEL_PATH):
    _check(_lib_host().wg_mlp_load(path.encode()), f"wg_mlp_load {path}")


def mlp_save(path: str):
    _check(_lib_host().wg_mlp_save(path.encode()), f"wg_mlp_save {path}")


def mlp_unload():
    _lib_host().wg_mlp_unload()


def mlp_loaded() -> bool:
    return bool(_lib_host().wg_mlp_loaded())


def mlp_features(app: int, iteration: int, V: int, E: int, frontier_vertices: int, frontier_edges: int,
                 unvisited: int) -> np.ndarray:
    """The schedule's feature vector of one loop state (host library)."""
    x = np.empty(MLP_IN, np.float32)
    _check(_lib_host().wg_mlp_feature_vector(app, iteration, V, E, frontier_vertices, frontier_edges, unvisited,
                                             _ptr(x, C.c_float)), "wg_mlp_feature_vector")
    return x


def mlp_base(V: int, E: int, frontier_edges: int) -> np.ndarray:
    """log2 work estimate of each action for a loop state (the unlearned part)."""
    b = np.empty(MLP_ACTIONS, np.float32)
    _check(_lib_host().wg_mlp_baseline(V, E, frontier_edges, _ptr(b, C.c_float)), "wg_mlp_baseline")
    return b


def mlp_predict(x: np.ndarray, app: int = APP_BFS, base: np.ndarray | None = None,
                plain: int = -1) -> tuple[int, np.ndarray]:
    """Host evaluation of the loaded model: (action chosen for app, settled
    against the plain rule's action plain (-1: none), predictions)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(MLP_ACTIONS, np.float32)
    bp = None
    if base is not None:
        base = np.ascontiguousarray(base, np.float32)
        bp = _ptr(base, C.c_float)
    a = _lib_host().wg_mlp_predict(app, _ptr(x, C.c_float), bp, plain, _ptr(y, C.c_float))
    if a < 0:
        raise WedgeError(a, "wg_mlp_predict")
    return a, y


def trace_rows() -> list[dict]:
    lib = _lib_host()
    n = int(lib.wg_trace_count())
    arr = (TraceRow * max(n, 1))()
    got = lib.wg_trace_copy(arr, n)
    return [{k: getattr(arr[i], k) for k, _ in TraceRow._fields_ if k != "pad"} for i in range(max(got, 0))]


def bfs_depth(packed: np.ndarray) -> np.ndarray:
    """Depth per vertex from packed BFS words (UINT32_MAX if unreached)."""
    d = (packed >> np.uint64(32)).astype(np.uint32)
    d[packed == UNREACHED_U64] = UNREACHED_U32
    return d


def bfs_parent(packed: np.ndarray) -> np.ndarray:
    """Parent per vertex from packed BFS words (UINT32_MAX if unreached)."""
    p = (packed & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    p[packed == UNREACHED_U64] = UNREACHED_U32
    return p
