"""MI355X curvelet construction: Python mirror of cpu_curvelet's in-process API.

The compute runs in libcurvelet_gpu.so (hand-written gfx950 HIP kernels behind a
C ABI, include/curvelet_gpu.h).  This module only binds that library with
ctypes and mirrors the reference's names so callers read like the reference:

    CurveletParams                 cpu_curvelet/param_settings.hpp:9-24
    read_TO_edges_from_file        cpu_curvelet/data_io.hpp:20-78
    NeighborGraph                  CPUNeighborGraph<T>, preprocess.hpp:17-31
    build_neighbor_csr_graph       preprocess.hpp:186-274
    CurveletGPU                    CurveletCPU<T>, cpu_curvelet.hpp:22-123
      .form_pairwise_bundles()     cpu_curvelet.hpp:305-361
      .build_curvelets_greedy()    cpu_curvelet.hpp:363-634
      .num_curvelets(), .chain_width(), .edge_chain_final, .curvelet_info
    run_curvelet                   cpu_curvelet/main.cpp:19-96 (timed span)

There is no CPU fallback: importing works without a GPU, but every call that
needs the device raises CurveletError when the library or the device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

__all__ = [
    "CurveletError", "CurveletParams", "NeighborGraph", "CurveletGPU", "Context", "lib_path", "load_library",
    "read_TO_edges_from_file", "build_neighbor_csr_graph", "run_curvelet", "synth_edges", "synth_preset",
    "math_eval", "write_chain_file", "write_info_file", "chain_to_info_filename", "dt_radians", "pinned_empty",
    "read_TO_edges_from_file_f32",
]

_PKG = Path(__file__).resolve().parent
_LIB_PATH = _PKG / "lib" / "libcurvelet_gpu.so"
_lib = None


class CurveletError(RuntimeError):
    def __init__(self, msg: str, status: int | None = None):
        super().__init__(msg)
        self.status = status


class _Params(C.Structure):
    _fields_ = [("nrad", C.c_double), ("gap", C.c_double), ("dx", C.c_double), ("dt_deg", C.c_double),
                ("max_k", C.c_double), ("curvelet_style", C.c_uint), ("group_max_sz", C.c_uint),
                ("out_type", C.c_uint), ("sx", C.c_double), ("st", C.c_double), ("edge_data_sz", C.c_int),
                ("max_candidates", C.c_uint)]


class _GraphInfo(C.Structure):
    _fields_ = [("num_edges", C.c_int), ("total_neighbor_pairs", C.c_int), ("max_num_of_neighbors", C.c_uint),
                ("truncated_anchors", C.c_uint), ("order_ambiguous_anchors", C.c_uint)]


class _RunStats(C.Structure):
    _fields_ = [("valid_pairs", C.c_uint), ("num_curvelets", C.c_uint), ("chain_width", C.c_uint),
                ("total_neighbor_pairs", C.c_int), ("max_num_of_neighbors", C.c_uint), ("ms_graph", C.c_double),
                ("ms_curvelets", C.c_double), ("ms_total", C.c_double), ("ms_fused_kernel", C.c_double),
                ("intersections", C.c_ulonglong), ("order_ambiguous_anchors", C.c_uint),
                ("lds_bytes_per_workgroup", C.c_uint)]


class _ServerConfig(C.Structure):
    _fields_ = [("workers", C.c_int), ("max_batch_maps", C.c_int), ("bytes_per_edge", C.c_double),
                ("hbm_budget_bytes", C.c_double)]


class _ServerResult(C.Structure):
    _fields_ = [("status", C.c_int), ("num_curvelets", C.c_uint), ("chain_width", C.c_uint),
                ("num_edges", C.c_int), ("batch_maps", C.c_int), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double)]


class _Tuning(C.Structure):
    _fields_ = [("coop_min_key", C.c_int), ("coop_waves", C.c_int), ("window_lanes", C.c_int),
                ("grid_pct", C.c_int)]


class _TuningTrial(C.Structure):
    _fields_ = [("tuning", _Tuning), ("ms", C.c_double)]


class _SynthConfig(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("num_curves", C.c_int), ("radius_min", C.c_double),
                ("radius_max", C.c_double), ("radius_log_uniform", C.c_int), ("segment_min", C.c_double),
                ("segment_max", C.c_double), ("bezier_extent", C.c_double), ("arc_max_length", C.c_double),
                ("clutter_fraction", C.c_double), ("junction_patches", C.c_int), ("curves_per_patch", C.c_int),
                ("patch_size", C.c_double), ("pos_sigma", C.c_double), ("theta_sigma_deg", C.c_double)]


_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int)
_up = _P(C.c_uint)
_ucp = _P(C.c_ubyte)

_SIGNATURES = {
    "cvg_params_default": (None, [_P(_Params)]),
    "cvg_status_string": (C.c_char_p, [C.c_int]),
    "cvg_create": (C.c_int, [C.c_int, _P(C.c_void_p)]),
    "cvg_destroy": (None, [C.c_void_p]),
    "cvg_last_error": (C.c_char_p, [C.c_void_p]),
    "cvg_read_edge_file": (C.c_int, [C.c_char_p, C.c_int, _P(_dp), _ip, C.c_char_p, C.c_size_t]),
    "cvg_free": (None, [C.c_void_p]),
    "cvg_write_chain_file": (C.c_int, [C.c_char_p, _up, C.c_uint, C.c_uint]),
    "cvg_write_info_file": (C.c_int, [C.c_char_p, _dp, C.c_uint, C.c_uint]),
    "cvg_build_neighbor_csr_graph": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, C.c_double, C.c_uint, C.c_uint]),
    "cvg_get_graph_info": (C.c_int, [C.c_void_p, _P(_GraphInfo)]),
    "cvg_get_csr": (C.c_int, [C.c_void_p, _ip, _ip, _dp]),
    "cvg_form_pairwise_bundles": (C.c_int, [C.c_void_p, _P(_Params), _up]),
    "cvg_build_curvelets_greedy": (C.c_int, [C.c_void_p, _P(_Params), _up]),
    "cvg_curvelet_shape": (C.c_int, [C.c_void_p, _up, _up]),
    "cvg_get_curvelets": (C.c_int, [C.c_void_p, _up, _dp]),
    "cvg_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _P(_Params), _P(_RunStats)]),
    "cvg_run_to_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _P(_Params), _P(C.c_uint), _dp, C.c_uint,
                                  _P(C.c_uint), _P(_RunStats)]),
    "cvg_run_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _P(_Params), _P(_RunStats)]),
    "cvg_run_f32": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _P(C.c_float), _P(_Params), _P(_RunStats)]),
    "cvg_build_neighbor_csr_graph_f32": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _P(C.c_float), C.c_double,
                                                   C.c_uint, C.c_uint]),
    "cvg_read_edge_file_f32": (C.c_int, [C.c_char_p, C.c_int, _P(_P(C.c_float)), _ip, C.c_char_p, C.c_size_t]),
    "cvg_run_device_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _P(_Params), C.c_void_p,
                                         C.c_void_p, C.c_uint, _up, _up]),
    "cvg_upload_edges": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _P(C.c_void_p)]),
    "cvg_device_curvelets": (C.c_int, [C.c_void_p, _P(C.c_void_p), _P(C.c_void_p)]),
    "cvg_get_stream": (C.c_int, [C.c_void_p, _P(C.c_void_p)]),
    "cvg_synchronize": (C.c_int, [C.c_void_p]),
    "cvg_recent_fused_ms": (C.c_int, [C.c_void_p, C.c_int, _P(C.c_float), _P(C.c_int)]),
    "cvg_set_event_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "cvg_run_batch": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, _P(_dp), _P(_Params), _P(_RunStats)]),
    "cvg_run_batch_device": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, C.c_void_p, _P(_Params), _P(_RunStats)]),
    "cvg_batch_map_shape": (C.c_int, [C.c_void_p, C.c_int, _up, _up]),
    "cvg_get_batch_map_curvelets": (C.c_int, [C.c_void_p, C.c_int, _up, _dp]),
    "cvg_get_tuning": (C.c_int, [C.c_void_p, _P(_Tuning)]),
    "cvg_set_tuning": (C.c_int, [C.c_void_p, _P(_Tuning)]),
    "cvg_autotune_batch": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, C.c_void_p, _P(_Params), C.c_int,
                                     _P(_Tuning), _dp, _P(_TuningTrial), C.c_int, _ip]),
    "cvg_server_config_default": (None, [_P(_ServerConfig)]),
    "cvg_server_create": (C.c_int, [C.c_int, _P(_Params), _P(_ServerConfig), _P(C.c_void_p)]),
    "cvg_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _P(C.c_ulonglong)]),
    "cvg_server_wait": (C.c_int, [C.c_void_p, C.c_ulonglong, _P(_ServerResult)]),
    "cvg_server_fetch": (C.c_int, [C.c_void_p, C.c_ulonglong, _up, _dp]),
    "cvg_server_view": (C.c_int, [C.c_void_p, C.c_ulonglong, _P(_up), _P(_dp)]),
    "cvg_server_release": (C.c_int, [C.c_void_p, C.c_ulonglong]),
    "cvg_host_alloc": (C.c_int, [C.c_size_t, _P(C.c_void_p)]),
    "cvg_host_free": (None, [C.c_void_p]),
    "cvg_server_counters": (C.c_int, [C.c_void_p, _P(C.c_ulonglong), _P(C.c_ulonglong), _P(C.c_ulonglong)]),
    "cvg_server_last_error": (C.c_char_p, [C.c_void_p]),
    "cvg_server_destroy": (None, [C.c_void_p]),
    "cvg_synth_preset": (C.c_int, [C.c_char_p, _P(_SynthConfig)]),
    "cvg_synth_generate": (C.c_int, [_P(_SynthConfig), C.c_ulonglong, _dp, C.c_int, _ip]),
    "cvg_debug_anchor_bundles": (C.c_int, [C.c_void_p, _P(_Params), C.c_int, _ucp, _ucp, _dp, _dp, _ip, _ip]),
    "cvg_math_eval": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]),
    "cvg_debug_cell_order": (C.c_int, [C.c_int, _dp, _dp, _ip]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def lib_path() -> Path:
    return Path(os.environ.get("CVG_LIBRARY", _LIB_PATH))


def load_library():
    """Load libcurvelet_gpu.so (built in-tree by build.py) and declare signatures."""
    global _lib
    if _lib is None:
        p = lib_path()
        if not p.exists():
            raise CurveletError(f"{p} is missing: build it with `python -m curvelet_construction_on_gpu_amd.build`")
        lib = C.CDLL(str(p))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(_P(ctype))


def dt_radians(dt_deg: float) -> float:
    """main.cpp:24: dt = T(dt_deg / 180.0) * T(M_PI)."""
    return (dt_deg / 180.0) * 3.14159265358979323846


@dataclass
class CurveletParams:
    """CurveletParams (param_settings.hpp:9-24), same defaults."""
    nrad: float = 3.5
    gap: float = 1.5
    dx: float = 0.4
    dt_deg: float = 15.0
    max_k: float = 0.3
    curvelet_style: int = 2
    group_max_sz: int = 4
    out_type: int = 0
    sx: float = 0.1
    st: float = 0.08
    edge_file: str = "eth3d_cables2.txt"
    edge_data_sz: int = 4
    max_candidates: int = 128

    def _c(self) -> _Params:
        return _Params(self.nrad, self.gap, self.dx, self.dt_deg, self.max_k, self.curvelet_style,
                       self.group_max_sz, self.out_type, self.sx, self.st, self.edge_data_sz, self.max_candidates)


@dataclass
class NeighborGraph:
    """CPUNeighborGraph<T> (preprocess.hpp:17-31)."""
    num_edges: int = 0
    total_neighbor_pairs: int = 0
    max_num_of_neighbors: int = 0
    edges: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    neighbor_offsets: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    neighbor_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    neighbor_dist2: np.ndarray = field(default_factory=lambda: np.zeros(0))
    truncated_anchors: int = 0
    order_ambiguous_anchors: int = 0
    _ctx: "Context | None" = None


class Context:
    """One HIP device context (own stream, resident device pools)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        rc = self._lib.cvg_create(device, C.byref(h))
        if rc != 0:
            raise CurveletError(f"cvg_create(device={device}) failed: "
                                f"{self._lib.cvg_status_string(rc).decode()}", rc)
        self._h = h

    @property
    def handle(self):
        return self._h

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self._lib.cvg_last_error(self._h).decode()
            raise CurveletError(f"{what}: {self._lib.cvg_status_string(rc).decode()}: {msg}", rc)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cvg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stage-wise API ----------------------------------------------------
    def build_graph(self, edges: np.ndarray, rad: float, neighbor_radius: int = 3,
                    max_candidates: int = 128) -> NeighborGraph:
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        self.check(self._lib.cvg_build_neighbor_csr_graph(self._h, n, sz, _ptr(e, C.c_double), rad,
                                                          neighbor_radius, max_candidates),
                   "build_neighbor_csr_graph")
        return self.graph(edges=e)

    def graph(self, edges: np.ndarray | None = None) -> NeighborGraph:
        gi = _GraphInfo()
        self.check(self._lib.cvg_get_graph_info(self._h, C.byref(gi)), "get_graph_info")
        off = np.zeros(gi.num_edges + 1, np.int32)
        ids = np.zeros(gi.total_neighbor_pairs, np.int32)
        d2 = np.zeros(gi.total_neighbor_pairs, np.float64)
        self.check(self._lib.cvg_get_csr(self._h, _ptr(off, C.c_int), _ptr(ids, C.c_int), _ptr(d2, C.c_double)),
                   "get_csr")
        return NeighborGraph(gi.num_edges, gi.total_neighbor_pairs, gi.max_num_of_neighbors,
                             edges if edges is not None else np.zeros((0, 4)), off, ids, d2,
                             gi.truncated_anchors, gi.order_ambiguous_anchors, self)

    def form_pairwise_bundles(self, params: CurveletParams) -> int:
        v = C.c_uint()
        p = params._c()
        self.check(self._lib.cvg_form_pairwise_bundles(self._h, C.byref(p), C.byref(v)), "form_pairwise_bundles")
        return v.value

    def build_curvelets_greedy(self, params: CurveletParams) -> int:
        v = C.c_uint()
        p = params._c()
        self.check(self._lib.cvg_build_curvelets_greedy(self._h, C.byref(p), C.byref(v)), "build_curvelets_greedy")
        return v.value

    def curvelets(self) -> tuple[np.ndarray, np.ndarray]:
        nc, w = C.c_uint(), C.c_uint()
        self.check(self._lib.cvg_curvelet_shape(self._h, C.byref(nc), C.byref(w)), "curvelet_shape")
        chains = np.zeros((nc.value, w.value), np.uint32)
        info = np.zeros((nc.value, 10), np.float64)
        self.check(self._lib.cvg_get_curvelets(self._h, _ptr(chains, C.c_uint), _ptr(info, C.c_double)),
                   "get_curvelets")
        return chains, info

    def shape(self) -> tuple[int, int]:
        """(num_curvelets(), chain_width()) of the last run."""
        nc, w = C.c_uint(), C.c_uint()
        self.check(self._lib.cvg_curvelet_shape(self._h, C.byref(nc), C.byref(w)), "curvelet_shape")
        return nc.value, w.value

    def curvelets_into(self, chains: np.ndarray, info: np.ndarray) -> int:
        """cvg_get_curvelets into caller arrays (C-contiguous, at least
        num_curvelets rows; page-locked ones from pinned_empty take one DMA
        each).  Returns num_curvelets."""
        nc, w = self.shape()
        if (chains.dtype != np.uint32 or info.dtype != np.float64 or not chains.flags.c_contiguous
                or not info.flags.c_contiguous or chains.shape[0] < nc or info.shape[0] < nc
                or chains.shape[1] != w or info.shape[1] != 10):
            raise CurveletError(f"curvelets_into: need uint32 ({nc}, {w}) and float64 ({nc}, 10) C arrays")
        self.check(self._lib.cvg_get_curvelets(self._h, _ptr(chains, C.c_uint), _ptr(info, C.c_double)),
                   "get_curvelets")
        return nc

    @property
    def stream(self) -> int:
        """The context's hipStream_t as an integer (for torch.cuda.ExternalStream)."""
        s = C.c_void_p()
        self.check(self._lib.cvg_get_stream(self._h, C.byref(s)), "get_stream")
        return s.value or 0

    # -- whole path ----------------------------------------------------------
    def run_to_host(self, edges: np.ndarray, params: CurveletParams, chains: np.ndarray, info: np.ndarray,
                    stats: bool = False) -> tuple[int, dict | None]:
        """cvg_run_to_host: host edges in, rows into the caller's chains
        (uint32, (rows, group_max_sz + 1)) and info (float64, (rows, 10))
        C arrays (page-locked ones from pinned_empty take one DMA each).
        Returns (rows, stats or None)."""
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        w = int(params.group_max_sz) + 1
        if (chains.dtype != np.uint32 or info.dtype != np.float64 or not chains.flags.c_contiguous
                or not info.flags.c_contiguous or chains.ndim != 2 or chains.shape[1] != w
                or info.shape != (chains.shape[0], 10)):
            raise CurveletError(f"run_to_host: need uint32 (rows, {w}) and float64 (rows, 10) C arrays")
        st = _RunStats()
        rows = C.c_uint()
        p = params._c()
        self.check(self._lib.cvg_run_to_host(self._h, n, sz, _ptr(e, C.c_double), C.byref(p), _ptr(chains, C.c_uint),
                                             _ptr(info, C.c_double), chains.shape[0], C.byref(rows),
                                             C.byref(st) if stats else None), "run_to_host")
        return rows.value, ({k: getattr(st, k) for k, _ in _RunStats._fields_} if stats else None)

    def run(self, edges: np.ndarray, params: CurveletParams, stats: bool = True) -> dict | None:
        """cvg_run from host memory.  stats=False returns once the work is
        queued; results are complete at their first query."""
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        st = _RunStats()
        p = params._c()
        self.check(self._lib.cvg_run(self._h, n, sz, _ptr(e, C.c_double), C.byref(p),
                                     C.byref(st) if stats else None), "run")
        return {k: getattr(st, k) for k, _ in _RunStats._fields_} if stats else None

    def run_f32(self, edges: np.ndarray, params: CurveletParams, stats: bool = True) -> dict | None:
        """run_curvelet<float> (--float): edges as float32, the reference's float arithmetic."""
        e = np.ascontiguousarray(edges, dtype=np.float32)
        n, sz = e.shape
        st = _RunStats()
        p = params._c()
        self.check(self._lib.cvg_run_f32(self._h, n, sz, _ptr(e, C.c_float), C.byref(p),
                                         C.byref(st) if stats else None), "run_f32")
        return {k: getattr(st, k) for k, _ in _RunStats._fields_} if stats else None

    def build_graph_f32(self, edges: np.ndarray, rad: float, neighbor_radius: int = 3,
                        max_candidates: int = 128) -> NeighborGraph:
        e = np.ascontiguousarray(edges, dtype=np.float32)
        n, sz = e.shape
        self.check(self._lib.cvg_build_neighbor_csr_graph_f32(self._h, n, sz, _ptr(e, C.c_float), rad,
                                                              neighbor_radius, max_candidates),
                   "build_neighbor_csr_graph_f32")
        return self.graph(edges=e)

    def upload(self, edges: np.ndarray) -> int:
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        d = C.c_void_p()
        self.check(self._lib.cvg_upload_edges(self._h, n, sz, _ptr(e, C.c_double), C.byref(d)), "upload_edges")
        return d.value

    def run_device(self, dptr: int, n: int, sz: int, params: CurveletParams, stats: bool = True) -> dict | None:
        """With stats=False the call returns once the work is queued (no wait
        for the curvelet stage); results are complete at their first query."""
        st = _RunStats()
        p = params._c()
        self.check(self._lib.cvg_run_device(self._h, n, sz, C.c_void_p(dptr), C.byref(p),
                                            C.byref(st) if stats else None), "run_device")
        return {k: getattr(st, k) for k, _ in _RunStats._fields_} if stats else None

    def run_device_noalloc(self, dptr: int, n: int, sz: int, params: CurveletParams, chains_dptr: int,
                           info_dptr: int, max_rows: int) -> tuple[int, int]:
        """cvg_run_device_noalloc: intermediate buffers in the library's static
        device scratch array, chains and info into caller device buffers of
        max_rows rows.  Returns (num_curvelets, valid_pairs)."""
        p = params._c()
        nc, vp = C.c_uint(), C.c_uint()
        self.check(self._lib.cvg_run_device_noalloc(self._h, n, sz, C.c_void_p(dptr), C.byref(p),
                                                    C.c_void_p(chains_dptr), C.c_void_p(info_dptr), max_rows,
                                                    C.byref(nc), C.byref(vp)), "run_device_noalloc")
        return nc.value, vp.value

    def recent_fused_ms(self, count: int) -> list[float]:
        """Durations (ms, HIP events) of the fused kernel in the last timed runs
        (runs with stats, or made while event timing is on), oldest first (at most 64)."""
        buf = (C.c_float * max(count, 0))()
        got = C.c_int()
        self.check(self._lib.cvg_recent_fused_ms(self._h, count, buf, C.byref(got)), "recent_fused_ms")
        return list(buf)[:got.value]

    def set_event_timing(self, on: bool) -> None:
        """Record the HIP timing events in runs without stats too (off at
        creation: each timing event holds the stream for a few microseconds)."""
        self.check(self._lib.cvg_set_event_timing(self._h, 1 if on else 0), "set_event_timing")

    def synchronize(self) -> None:
        self.check(self._lib.cvg_synchronize(self._h), "synchronize")

    # -- batched form --------------------------------------------------------
    def run_batch(self, maps: list, params: CurveletParams) -> dict:
        """Several edge maps (each N_m x sz) in one pass; results per map via batch_curvelets(m)."""
        arrs = [np.ascontiguousarray(m, dtype=np.float64).reshape(-1, np.shape(m)[1] if np.ndim(m) == 2 else 4)
                for m in maps]
        sz = arrs[0].shape[1] if arrs else 4
        if any(a.shape[1] != sz for a in arrs):
            raise CurveletError("all maps of a batch need the same values per edge")
        counts = np.array([a.shape[0] for a in arrs], dtype=np.int32)
        ptrs = (_dp * len(arrs))(*[_ptr(a, C.c_double) for a in arrs])
        st = _RunStats()
        p = params._c()
        self.check(self._lib.cvg_run_batch(self._h, len(arrs), _ptr(counts, C.c_int), sz, ptrs, C.byref(p),
                                           C.byref(st)), "run_batch")
        self._batch_keep = arrs
        return {k: getattr(st, k) for k, _ in _RunStats._fields_}

    def run_batch_device(self, dptr: int, counts, sz: int, params: CurveletParams,
                         stats: bool = True) -> dict | None:
        """stats=False: return once the work is queued (results complete at their first query)."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        st = _RunStats()
        p = params._c()
        self.check(self._lib.cvg_run_batch_device(self._h, len(c), _ptr(c, C.c_int), sz, C.c_void_p(dptr),
                                                  C.byref(p), C.byref(st) if stats else None), "run_batch_device")
        return {k: getattr(st, k) for k, _ in _RunStats._fields_} if stats else None

    @property
    def tuning(self) -> dict:
        """Launch shape (results never depend on it): coop_min_key, coop_waves, window_lanes, grid_pct."""
        t = _Tuning()
        self.check(self._lib.cvg_get_tuning(self._h, C.byref(t)), "get_tuning")
        return {k: getattr(t, k) for k, _ in _Tuning._fields_}

    @tuning.setter
    def tuning(self, value: dict) -> None:
        cur = self.tuning
        cur.update(value)  # fields not given keep their current values
        t = _Tuning(**{k: int(cur[k]) for k, _ in _Tuning._fields_})
        self.check(self._lib.cvg_set_tuning(self._h, C.byref(t)), "set_tuning")

    def autotune_batch(self, dptr: int, counts, sz: int, params: CurveletParams, reps: int = 2):
        """Times the batched path on these maps under candidate tunings, keeps the fastest.
        Returns (best tuning, its ms per run, [(tuning, ms), ...] in the order tried)."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        p = params._c()
        best, best_ms, n = _Tuning(), C.c_double(), C.c_int()
        trials = (_TuningTrial * 32)()
        self.check(self._lib.cvg_autotune_batch(self._h, len(c), _ptr(c, C.c_int), sz, C.c_void_p(dptr), C.byref(p),
                                                reps, C.byref(best), C.byref(best_ms), trials, 32, C.byref(n)),
                   "autotune_batch")
        as_dict = lambda t: {k: getattr(t, k) for k, _ in _Tuning._fields_}
        return as_dict(best), best_ms.value, [(as_dict(trials[i].tuning), trials[i].ms) for i in range(min(n.value, 32))]

    def batch_curvelets(self, m: int) -> tuple[np.ndarray, np.ndarray]:
        rows, w = C.c_uint(), C.c_uint()
        self.check(self._lib.cvg_batch_map_shape(self._h, m, C.byref(rows), C.byref(w)), "batch_map_shape")
        chains = np.zeros((rows.value, w.value), np.uint32)
        info = np.zeros((rows.value, 10))
        self.check(self._lib.cvg_get_batch_map_curvelets(self._h, m, _ptr(chains, C.c_uint), _ptr(info, C.c_double)),
                   "get_batch_map_curvelets")
        return chains, info

    def anchor_bundles(self, params: CurveletParams, anchor: int, max_deg: int):
        """Per-neighbour dir class, hypothesis flag and bundle cells of one anchor (device)."""
        K = 256
        dir_ = np.zeros(max(max_deg, 1), np.uint8)
        hyp = np.zeros(max(max_deg, 1), np.uint8)
        bmin = np.zeros(max(max_deg, 1) * K)
        bmax = np.zeros(max(max_deg, 1) * K)
        deg, cells = C.c_int(), C.c_int()
        p = params._c()
        self.check(self._lib.cvg_debug_anchor_bundles(self._h, C.byref(p), anchor, _ptr(dir_, C.c_ubyte),
                                                      _ptr(hyp, C.c_ubyte), _ptr(bmin, C.c_double),
                                                      _ptr(bmax, C.c_double), C.byref(deg), C.byref(cells)),
                   "debug_anchor_bundles")
        d, k = deg.value, cells.value
        return dir_[:d], hyp[:d], bmin[:d * k].reshape(d, k), bmax[:d * k].reshape(d, k)


class Server:
    """The served form: continuous batching over worker contexts (cvg_server_*)."""

    def __init__(self, params: CurveletParams | None = None, device: int = 0, workers: int | None = None,
                 max_batch_maps: int | None = None, hbm_budget_bytes: float | None = None,
                 bytes_per_edge: float | None = None):
        self._lib = load_library()
        cfg = _ServerConfig()
        self._lib.cvg_server_config_default(C.byref(cfg))
        if workers is not None:
            cfg.workers = workers
        if max_batch_maps is not None:
            cfg.max_batch_maps = max_batch_maps
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = hbm_budget_bytes
        if bytes_per_edge is not None:
            cfg.bytes_per_edge = bytes_per_edge
        self.config = {k: getattr(cfg, k) for k, _ in _ServerConfig._fields_}
        h = C.c_void_p()
        p = (params or CurveletParams())._c()
        rc = self._lib.cvg_server_create(device, C.byref(p), C.byref(cfg), C.byref(h))
        if rc != 0:
            raise CurveletError(f"cvg_server_create failed ({rc})", rc)
        self._h = h

    def submit(self, edges: np.ndarray) -> int:
        e = np.ascontiguousarray(edges, dtype=np.float64)
        rid = C.c_ulonglong()
        rc = self._lib.cvg_server_submit(self._h, e.shape[0], e.shape[1], _ptr(e, C.c_double), C.byref(rid))
        if rc != 0:
            raise CurveletError(f"cvg_server_submit failed ({rc})", rc)
        return rid.value

    def wait(self, rid: int) -> dict:
        res = _ServerResult()
        rc = self._lib.cvg_server_wait(self._h, rid, C.byref(res))
        if rc != 0:
            raise CurveletError(f"request {rid} failed: {self._lib.cvg_server_last_error(self._h).decode()}", rc)
        return {k: getattr(res, k) for k, _ in _ServerResult._fields_}

    def result(self, rid: int) -> tuple[np.ndarray, np.ndarray, dict]:
        r = self.wait(rid)
        chains = np.zeros((r["num_curvelets"], r["chain_width"]), np.uint32)
        info = np.zeros((r["num_curvelets"], 10))
        rc = self._lib.cvg_server_fetch(self._h, rid, _ptr(chains, C.c_uint), _ptr(info, C.c_double))
        if rc != 0:
            raise CurveletError(f"cvg_server_fetch failed ({rc})", rc)
        return chains, info, r

    def view(self, rid: int) -> tuple[np.ndarray, np.ndarray, dict]:
        """Zero-copy arrays over the server's page-locked result rows; valid until release(rid)."""
        r = self.wait(rid)
        cp, ip = _up(), _dp()
        rc = self._lib.cvg_server_view(self._h, rid, C.byref(cp), C.byref(ip))
        if rc != 0:
            raise CurveletError(f"cvg_server_view failed ({rc})", rc)
        rows, w = r["num_curvelets"], r["chain_width"]
        if rows == 0:
            return np.zeros((0, w), np.uint32), np.zeros((0, 10)), r
        return (np.ctypeslib.as_array(cp, shape=(rows, w)), np.ctypeslib.as_array(ip, shape=(rows, 10)), r)

    def release(self, rid: int) -> None:
        self._lib.cvg_server_release(self._h, rid)

    def counters(self) -> dict:
        b, m, e = C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong()
        self._lib.cvg_server_counters(self._h, C.byref(b), C.byref(m), C.byref(e))
        return {"batches": b.value, "maps": m.value, "edges": e.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.cvg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _Pinned:
    """Page-locked host block from cvg_host_alloc, freed with the last array over it."""

    def __init__(self, nbytes: int):
        self._lib = load_library()
        p = C.c_void_p()
        rc = self._lib.cvg_host_alloc(max(int(nbytes), 1), C.byref(p))
        if rc != 0:
            raise CurveletError(f"cvg_host_alloc({nbytes}) failed", rc)
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.cvg_host_free(C.c_void_p(self.ptr))
            self.ptr = None


def pinned_empty(shape, dtype=np.float64) -> np.ndarray:
    """A numpy array in page-locked host memory (DMA at full link speed for
    uploads and downloads)."""
    dt = np.dtype(dtype)
    shape = tuple(int(v) for v in np.atleast_1d(shape))
    count = int(np.prod(shape)) if shape else 1
    holder = _Pinned(count * dt.itemsize)
    buf = (C.c_char * max(count * dt.itemsize, 1)).from_address(holder.ptr)
    buf._holder = holder  # the block lives as long as the array's buffer
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


_default_ctx: Context | None = None


def _ctx() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("CVG_DEVICE", "0")))
    return _default_ctx


def build_neighbor_csr_graph(num_edges: int, sz_edge_data: int, edges: np.ndarray, rad: float,
                             neighbor_radius: int, max_candidates: int, nthreads: int,
                             graph: NeighborGraph, ctx: Context | None = None) -> bool:
    """preprocess.hpp:186-195; nthreads is accepted for parity and unused."""
    del nthreads
    e = np.ascontiguousarray(edges, dtype=np.float64).reshape(num_edges, sz_edge_data)
    c = ctx or _ctx()
    try:
        g = c.build_graph(e, rad, neighbor_radius, max_candidates)
    except CurveletError as exc:
        if exc.status in (1, 2):
            return False
        raise
    graph.__dict__.update(g.__dict__)
    return True


class CurveletGPU:
    """CurveletCPU<T> (cpu_curvelet.hpp:22-123) on the GPU; takes the graph built
    by build_neighbor_csr_graph on the same context."""

    def __init__(self, num_edges, sz_edge_data, graph: NeighborGraph, dx, dt, sx, st, max_k, group_max_sz,
                 nthreads, to_edges, nrad, dt_deg: float | None = None):
        del nthreads, to_edges
        self._ctx = graph._ctx or _ctx()
        # the C ABI takes dt in degrees (as CurveletParams); recover it when only radians are given
        if dt_deg is None:
            dt_deg = dt * 180.0 / 3.14159265358979323846
            if dt_radians(dt_deg) != dt:
                raise CurveletError("pass dt_deg: dt in radians does not round-trip to degrees exactly")
        self.params = CurveletParams(nrad=nrad, dx=dx, dt_deg=dt_deg, max_k=max_k, sx=sx, st=st,
                                     group_max_sz=group_max_sz, edge_data_sz=sz_edge_data)
        self._n = num_edges
        self.edge_chain_final = np.zeros((0, group_max_sz + 1), np.uint32)
        self.curvelet_info = np.zeros((0, 10))

    def form_pairwise_bundles(self) -> int:
        return self._ctx.form_pairwise_bundles(self.params)

    def build_curvelets_greedy(self) -> None:
        self._ctx.build_curvelets_greedy(self.params)
        self.edge_chain_final, self.curvelet_info = self._ctx.curvelets()

    def num_curvelets(self) -> int:
        return int(self.edge_chain_final.shape[0])

    def chain_width(self) -> int:
        return int(self.params.group_max_sz + 1)


def run_curvelet(edges: np.ndarray, params: CurveletParams | None = None, ctx: Context | None = None):
    """Edges (N x sz) -> (chains (C x G+1, 1-based), info (C x 10), stats) on the GPU."""
    params = params or CurveletParams()
    c = ctx or _ctx()
    stats = c.run(edges, params)
    chains, info = c.curvelets()
    return chains, info, stats


def read_TO_edges_from_file(path: str | os.PathLike, values_per_edge: int = 4) -> np.ndarray:
    """data_io.hpp:20-78 semantics on an explicit path (no ../test_files prefix)."""
    lib = load_library()
    buf = _dp()
    n = C.c_int()
    err = C.create_string_buffer(1024)
    rc = lib.cvg_read_edge_file(os.fsencode(str(path)), values_per_edge, C.byref(buf), C.byref(n), err, 1024)
    if rc != 0:
        raise CurveletError(err.value.decode(), rc)
    try:
        out = np.ctypeslib.as_array(buf, shape=(n.value * values_per_edge,)).copy()
    finally:
        lib.cvg_free(buf)
    return out.reshape(n.value, values_per_edge)


def read_TO_edges_from_file_f32(path: str | os.PathLike, values_per_edge: int = 4) -> np.ndarray:
    """data_io.hpp:20-78 on T = float: decimal text straight to float32."""
    lib = load_library()
    buf = _P(C.c_float)()
    n = C.c_int()
    err = C.create_string_buffer(1024)
    rc = lib.cvg_read_edge_file_f32(os.fsencode(str(path)), values_per_edge, C.byref(buf), C.byref(n), err, 1024)
    if rc != 0:
        raise CurveletError(err.value.decode(), rc)
    try:
        out = np.ctypeslib.as_array(buf, shape=(n.value * values_per_edge,)).copy()
    finally:
        lib.cvg_free(buf)
    return out.reshape(n.value, values_per_edge)


def write_chain_file(path, chains: np.ndarray) -> None:
    lib = load_library()
    c = np.ascontiguousarray(chains, dtype=np.uint32)
    rows, w = c.shape
    rc = lib.cvg_write_chain_file(os.fsencode(str(path)), _ptr(c, C.c_uint), rows, w)
    if rc != 0:
        raise CurveletError(f"write_chain_file({path}) failed", rc)


def write_info_file(path, info: np.ndarray) -> None:
    lib = load_library()
    d = np.ascontiguousarray(info, dtype=np.float64)
    rows, w = d.shape
    rc = lib.cvg_write_info_file(os.fsencode(str(path)), _ptr(d, C.c_double), rows, w)
    if rc != 0:
        raise CurveletError(f"write_info_file({path}) failed", rc)


def chain_to_info_filename(chain_file: str) -> str:
    """data_io.hpp:135-145."""
    pos = chain_file.find("chain")
    if pos >= 0:
        return chain_file[:pos] + "info" + chain_file[pos + 5:]
    return "info_" + chain_file


def synth_preset(name: str) -> _SynthConfig:
    lib = load_library()
    cfg = _SynthConfig()
    rc = lib.cvg_synth_preset(name.encode(), C.byref(cfg))
    if rc != 0:
        raise CurveletError(f"unknown synthetic preset {name!r}", rc)
    return cfg


def synth_edges(preset: str | _SynthConfig, seed: int, **overrides) -> np.ndarray:
    """Seeded synthetic edge map (N x 4: x, y, theta, strength)."""
    lib = load_library()
    cfg = synth_preset(preset) if isinstance(preset, str) else preset
    for k, v in overrides.items():
        setattr(cfg, k, v)
    n = C.c_int()
    rc = lib.cvg_synth_generate(C.byref(cfg), seed, None, 0, C.byref(n))
    if rc != 0:
        raise CurveletError("synth_generate failed", rc)
    out = np.zeros((n.value, 4))
    rc = lib.cvg_synth_generate(C.byref(cfg), seed, _ptr(out, C.c_double), n.value, C.byref(n))
    if rc != 0:
        raise CurveletError("synth_generate failed", rc)
    return out


MATH_OPS = {"sin": 0, "cos": 1, "atan2": 2, "sqrt": 3, "to2pi": 4, "fmod": 5, "sincos_sin": 6, "sincos_cos": 7,
            "sin_difference": 8, "sin_difference_settled": 9, "sincos_dd_sin_hi": 10, "sincos_dd_sin_lo": 11,
            "sincos_dd_cos_hi": 12, "sincos_dd_cos_lo": 13, "sinf": 14, "cosf": 15, "atan2f": 16, "to2pi_f": 17,
            "sqrtf": 18}


def math_eval(op: str, x: np.ndarray, y: np.ndarray | None = None, ctx: Context | None = None,
              device: bool = True) -> np.ndarray:
    """The library's math on the device (or, with device=False, its host build)."""
    lib = load_library()
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ya = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(xa)
    h = (ctx or _ctx()).handle if device else None
    rc = lib.cvg_math_eval(h, MATH_OPS[op], xa.size, _ptr(xa, C.c_double),
                           None if ya is None else _ptr(ya, C.c_double), _ptr(out, C.c_double))
    if rc != 0:
        raise CurveletError(f"math_eval({op}) failed", rc)
    return out


def cell_order_ranks(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Each edge's position inside its pixel cell after the reference's
    spatial-index sort (preprocess.hpp:78-92), as the library replays it
    (host code; no device needed)."""
    lib = load_library()
    xx = np.ascontiguousarray(x, dtype=np.float64)
    yy = np.ascontiguousarray(y, dtype=np.float64)
    r = np.zeros(len(xx), np.int32)
    rc = lib.cvg_debug_cell_order(len(xx), _ptr(xx, C.c_double), _ptr(yy, C.c_double), _ptr(r, C.c_int))
    if rc != 0:
        raise CurveletError("cvg_debug_cell_order failed", rc)
    return r
