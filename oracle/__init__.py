"""Test infrastructure only: ctypes bindings to the CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / the timed CPU baseline; the product
package never does.

  Oracle  - oracle/liboracle.so, the from-scratch CPU restatement of
            cpu_curvelet (curvelet_oracle.cpp).
  OracleF32 - oracle/liboracle_f32.so (glibc) and liboracle_f32_cr.so
            (correctly rounded float math), the restatement of the reference's
            float instantiation (--float, curvelet_oracle_f32.cpp).
  Ref     - oracle/_ref/libcurvelet_ref.so, the reference's own cpu_curvelet
            headers compiled in place by `make -C oracle ref` (absent where the
            reference tree is not available; callers skip then).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
ORACLE_CR_LIB = HERE / "liboracle_cr.so"
ORACLE_F32_LIB = HERE / "liboracle_f32.so"
ORACLE_F32_CR_LIB = HERE / "liboracle_f32_cr.so"
REF_LIB = HERE / "_ref" / "libcurvelet_ref.so"
REF_CLI = HERE / "_ref" / "CURVELET"
REFERENCE_SRC = Path("/root/reference/cpu_curvelet")

_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int)
_up = _P(C.c_uint)


def build(ref: bool | None = None) -> None:
    """Compile the oracle (and the in-place reference build when the sources exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if ref is None:
        ref = REFERENCE_SRC.exists()
    if ref:
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


class _OParams(C.Structure):
    _fields_ = [("nrad", C.c_double), ("dx", C.c_double), ("dt", C.c_double), ("max_k", C.c_double),
                ("sx", C.c_double), ("st", C.c_double), ("group_max_sz", C.c_uint),
                ("max_candidates", C.c_uint), ("neighbor_radius", C.c_uint)]


class _RParams(C.Structure):
    _fields_ = [("nrad", C.c_double), ("dx", C.c_double), ("dt_deg", C.c_double), ("max_k", C.c_double),
                ("sx", C.c_double), ("st", C.c_double), ("group_max_sz", C.c_uint),
                ("max_candidates", C.c_uint)]


def _f(a, t):
    return a.ctypes.data_as(_P(t))


class Oracle:
    """The CPU restatement: liboracle.so (glibc sin/cos/atan2, as the reference),
    or with cr=True liboracle_cr.so (the GPU path's correctly rounded math)."""

    def __init__(self, cr: bool = False):
        path = ORACLE_CR_LIB if cr else ORACLE_LIB
        if not path.exists():
            build(ref=False)
        lib = C.CDLL(str(path))
        sig = {
            "orc_dt_from_deg": (C.c_double, [C.c_double]),
            "orc_create": (C.c_void_p, []),
            "orc_spatial_order": (None, [C.c_int, _dp, _dp, _ip]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_build_csr": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _P(_OParams), C.c_int]),
            "orc_csr_counts": (None, [C.c_void_p, _ip, _up, _up]),
            "orc_get_csr": (None, [C.c_void_p, _ip, _ip, _dp]),
            "orc_valid_pairs": (C.c_uint, [C.c_void_p, _P(_OParams), C.c_int]),
            "orc_greedy": (C.c_uint, [C.c_void_p, _P(_OParams), C.c_int]),
            "orc_chain_width": (C.c_uint, [C.c_void_p]),
            "orc_get_curvelets": (None, [C.c_void_p, _up, _dp]),
            "orc_pair_bundle": (C.c_int, [_P(_OParams), C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, _dp, _dp]),
            "orc_bundle_cells": (C.c_int, [_P(_OParams), _up, _up]),
            "orc_dir_dot": (C.c_double, [C.c_int] + [C.c_double] * 5),
            "orc_anchor_detail": (C.c_int, [C.c_void_p, _P(_OParams), C.c_int, _P(C.c_ubyte), _P(C.c_ubyte),
                                            _dp, _dp]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(lib, k)
            fn.restype = r
            fn.argtypes = a
        self.lib = lib
        self.h = lib.orc_create()

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def spatial_order(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """The reference's sorted_edge_ids (std::sort of ids by packed cell key)."""
        xx = np.ascontiguousarray(x, dtype=np.float64)
        yy = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(len(xx), np.int32)
        self.lib.orc_spatial_order(len(xx), _f(xx, C.c_double), _f(yy, C.c_double), _f(out, C.c_int))
        return out

    def params(self, p) -> _OParams:
        """From a CurveletParams-like object (dt in degrees)."""
        return _OParams(p.nrad, p.dx, self.lib.orc_dt_from_deg(p.dt_deg), p.max_k, p.sx, p.st, p.group_max_sz,
                        p.max_candidates, 3)

    def build_csr(self, edges: np.ndarray, p, nthreads: int = 0):
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        op = self.params(p)
        rc = self.lib.orc_build_csr(self.h, n, sz, _f(e, C.c_double), C.byref(op), nthreads)
        if rc != 0:
            raise ValueError("oracle: build_csr rejected the input")
        pairs, maxdeg, trunc = C.c_int(), C.c_uint(), C.c_uint()
        self.lib.orc_csr_counts(self.h, C.byref(pairs), C.byref(maxdeg), C.byref(trunc))
        off = np.zeros(n + 1, np.int32)
        ids = np.zeros(pairs.value, np.int32)
        d2 = np.zeros(pairs.value)
        self.lib.orc_get_csr(self.h, _f(off, C.c_int), _f(ids, C.c_int), _f(d2, C.c_double))
        return {"offsets": off, "ids": ids, "dist2": d2, "pairs": pairs.value, "max_deg": maxdeg.value,
                "truncated": trunc.value}

    def valid_pairs(self, p, nthreads: int = 0) -> int:
        op = self.params(p)
        return int(self.lib.orc_valid_pairs(self.h, C.byref(op), nthreads))

    def greedy(self, p, nthreads: int = 0):
        op = self.params(p)
        nc = self.lib.orc_greedy(self.h, C.byref(op), nthreads)
        w = self.lib.orc_chain_width(self.h)
        chains = np.zeros((nc, w), np.uint32)
        info = np.zeros((nc, 10))
        self.lib.orc_get_curvelets(self.h, _f(chains, C.c_uint), _f(info, C.c_double))
        return chains, info

    def run(self, edges: np.ndarray, p, nthreads: int = 0, valid_pairs: bool = True):
        csr = self.build_csr(edges, p, nthreads)
        vp = self.valid_pairs(p, nthreads) if valid_pairs else None
        chains, info = self.greedy(p, nthreads)
        return {"csr": csr, "valid_pairs": vp, "chains": chains, "info": info}

    def pair_bundle(self, p, a, n):
        op = self.params(p)
        kx, kt = C.c_uint(), C.c_uint()
        K = self.lib.orc_bundle_cells(C.byref(op), C.byref(kx), C.byref(kt))
        lo = np.zeros(K)
        hi = np.zeros(K)
        hyp = self.lib.orc_pair_bundle(C.byref(op), a[0], a[1], a[2], n[0], n[1], n[2], _f(lo, C.c_double),
                                       _f(hi, C.c_double))
        return bool(hyp), lo, hi

    def dir_dot(self, a, n, run: int = 0) -> float:
        return self.lib.orc_dir_dot(run, a[0], a[1], a[2], n[0], n[1])

    def anchor_detail(self, p, anchor: int, deg: int):
        op = self.params(p)
        kx, kt = C.c_uint(), C.c_uint()
        K = self.lib.orc_bundle_cells(C.byref(op), C.byref(kx), C.byref(kt))
        d = max(deg, 1)
        dir_ = np.zeros(d, np.uint8)
        hyp = np.zeros(d, np.uint8)
        lo = np.zeros(d * K)
        hi = np.zeros(d * K)
        r = self.lib.orc_anchor_detail(self.h, C.byref(op), anchor, _f(dir_, C.c_ubyte), _f(hyp, C.c_ubyte),
                                       _f(lo, C.c_double), _f(hi, C.c_double))
        return dir_[:r], hyp[:r], lo[:r * K].reshape(r, K), hi[:r * K].reshape(r, K)


class Ref:
    """The reference's own cpu_curvelet, compiled in place (oracle/_ref)."""

    @staticmethod
    def available() -> bool:
        return REF_LIB.exists()

    def __init__(self):
        if not REF_LIB.exists():
            raise FileNotFoundError(f"{REF_LIB} missing (run `make -C oracle ref` where the reference exists)")
        lib = C.CDLL(str(REF_LIB))
        sig = {
            "ref_dt_from_deg": (C.c_double, [C.c_double]),
            "ref_run": (C.c_int, [C.c_int, C.c_int, _dp, _P(_RParams), C.c_int, C.c_int, C.c_int]),
            "ref_counts": (None, [_ip, _up, _up, _up, _up]),
            "ref_times": (None, [_dp]),
            "ref_get_csr": (None, [_ip, _ip, _dp]),
            "ref_get_curvelets": (None, [_up, _dp]),
            "ref_pair_bundle": (C.c_int, [_P(_RParams), C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, _dp, _dp, _ip]),
            "ref_dir_dot": (C.c_double, [C.c_double] * 5),
        }
        for k, (r, a) in sig.items():
            fn = getattr(lib, k)
            fn.restype = r
            fn.argtypes = a
        self.lib = lib

    @staticmethod
    def params(p) -> _RParams:
        return _RParams(p.nrad, p.dx, p.dt_deg, p.max_k, p.sx, p.st, p.group_max_sz, p.max_candidates)

    def run(self, edges: np.ndarray, p, nthreads: int = 1, valid_pairs: bool = True):
        e = np.ascontiguousarray(edges, dtype=np.float64)
        n, sz = e.shape
        rp = self.params(p)
        rc = self.lib.ref_run(n, sz, _f(e, C.c_double), C.byref(rp), nthreads, 1, 1 if valid_pairs else 0)
        if rc != 0:
            raise ValueError(f"reference run failed ({rc})")
        pairs, maxdeg, vp, nc, w = C.c_int(), C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
        self.lib.ref_counts(C.byref(pairs), C.byref(maxdeg), C.byref(vp), C.byref(nc), C.byref(w))
        off = np.zeros(n + 1, np.int32)
        ids = np.zeros(pairs.value, np.int32)
        d2 = np.zeros(pairs.value)
        self.lib.ref_get_csr(_f(off, C.c_int), _f(ids, C.c_int), _f(d2, C.c_double))
        chains = np.zeros((nc.value, w.value), np.uint32)
        info = np.zeros((nc.value, 10))
        self.lib.ref_get_curvelets(_f(chains, C.c_uint), _f(info, C.c_double))
        t = np.zeros(3)
        self.lib.ref_times(_f(t, C.c_double))
        return {"csr": {"offsets": off, "ids": ids, "dist2": d2, "pairs": pairs.value, "max_deg": maxdeg.value},
                "valid_pairs": vp.value if valid_pairs else None, "chains": chains, "info": info,
                "times": {"csr": t[0], "pairwise": t[1], "greedy": t[2]}}

    def run_f32(self, edges: np.ndarray, p, nthreads: int = 1, valid_pairs: bool = True):
        """run_curvelet<float> on a float32 edge array."""
        e = np.ascontiguousarray(edges, dtype=np.float32)
        n, sz = e.shape
        rp = self.params(p)
        fn = self.lib.ref_run_f32
        fn.restype = C.c_int
        fn.argtypes = [C.c_int, C.c_int, _P(C.c_float), _P(_RParams), C.c_int, C.c_int, C.c_int]
        if fn(n, sz, _f(e, C.c_float), C.byref(rp), nthreads, 1, 1 if valid_pairs else 0) != 0:
            raise ValueError("reference float run failed")
        pairs, maxdeg, vp, nc, w = C.c_int(), C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
        self.lib.ref_counts_f32.argtypes = [_ip, _up, _up, _up, _up]
        self.lib.ref_counts_f32(C.byref(pairs), C.byref(maxdeg), C.byref(vp), C.byref(nc), C.byref(w))
        off = np.zeros(n + 1, np.int32)
        ids = np.zeros(pairs.value, np.int32)
        d2 = np.zeros(pairs.value)
        self.lib.ref_get_csr_f32.argtypes = [_ip, _ip, _dp]
        self.lib.ref_get_csr_f32(_f(off, C.c_int), _f(ids, C.c_int), _f(d2, C.c_double))
        chains = np.zeros((nc.value, w.value), np.uint32)
        info = np.zeros((nc.value, 10))
        self.lib.ref_get_curvelets_f32.argtypes = [_up, _dp]
        self.lib.ref_get_curvelets_f32(_f(chains, C.c_uint), _f(info, C.c_double))
        return {"csr": {"offsets": off, "ids": ids, "dist2": d2, "pairs": pairs.value, "max_deg": maxdeg.value},
                "valid_pairs": vp.value if valid_pairs else None, "chains": chains, "info": info}

    def pair_bundle(self, p, a, n):
        rp = self.params(p)
        lo = np.zeros(256)
        hi = np.zeros(256)
        k = C.c_int()
        hyp = self.lib.ref_pair_bundle(C.byref(rp), a[0], a[1], a[2], n[0], n[1], n[2], _f(lo, C.c_double),
                                       _f(hi, C.c_double), C.byref(k))
        return bool(hyp), lo[:k.value], hi[:k.value]

    def dir_dot(self, a, n) -> float:
        return self.lib.ref_dir_dot(a[0], a[1], a[2], n[0], n[1])


class _FParams(C.Structure):
    _fields_ = [("nrad", C.c_double), ("dx", C.c_double), ("dt_deg", C.c_double), ("max_k", C.c_double),
                ("sx", C.c_double), ("st", C.c_double), ("group_max_sz", C.c_uint),
                ("max_candidates", C.c_uint), ("neighbor_radius", C.c_uint)]


class OracleF32:
    """The float instantiation's restatement; edges are float32 arrays."""

    def __init__(self, cr: bool = False):
        lib = C.CDLL(str(ORACLE_F32_CR_LIB if cr else ORACLE_F32_LIB))
        _fp = _P(C.c_float)
        sig = {
            "orcf_create": (C.c_void_p, []),
            "orcf_destroy": (None, [C.c_void_p]),
            "orcf_build_csr": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _fp, _P(_FParams), C.c_int]),
            "orcf_csr_counts": (None, [C.c_void_p, _ip, _up, _up]),
            "orcf_get_csr": (None, [C.c_void_p, _ip, _ip, _dp]),
            "orcf_valid_pairs": (C.c_uint, [C.c_void_p, _P(_FParams), C.c_int]),
            "orcf_greedy": (C.c_uint, [C.c_void_p, _P(_FParams), C.c_int]),
            "orcf_chain_width": (C.c_uint, [C.c_void_p]),
            "orcf_get_curvelets": (None, [C.c_void_p, _up, _dp]),
            "orcf_pair_bundle": (C.c_int, [_P(_FParams)] + [C.c_float] * 6 + [_fp, _fp]),
        }
        for k, (r, a) in sig.items():
            fn = getattr(lib, k)
            fn.restype = r
            fn.argtypes = a
        self.lib = lib
        self.h = lib.orcf_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orcf_destroy(self.h)
            self.h = None

    @staticmethod
    def params(p) -> _FParams:
        return _FParams(p.nrad, p.dx, p.dt_deg, p.max_k, p.sx, p.st, p.group_max_sz, p.max_candidates, 3)

    def run(self, edges: np.ndarray, p, nthreads: int = 0, valid_pairs: bool = True):
        e = np.ascontiguousarray(edges, dtype=np.float32)
        n, sz = e.shape
        fp = self.params(p)
        if self.lib.orcf_build_csr(self.h, n, sz, _f(e, C.c_float), C.byref(fp), nthreads) != 0:
            raise ValueError("orcf_build_csr failed")
        pairs, maxdeg, trunc = C.c_int(), C.c_uint(), C.c_uint()
        self.lib.orcf_csr_counts(self.h, C.byref(pairs), C.byref(maxdeg), C.byref(trunc))
        off = np.zeros(n + 1, np.int32)
        ids = np.zeros(pairs.value, np.int32)
        d2 = np.zeros(pairs.value)
        self.lib.orcf_get_csr(self.h, _f(off, C.c_int), _f(ids, C.c_int), _f(d2, C.c_double))
        vp = int(self.lib.orcf_valid_pairs(self.h, C.byref(fp), nthreads)) if valid_pairs else None
        nc = self.lib.orcf_greedy(self.h, C.byref(fp), nthreads)
        w = self.lib.orcf_chain_width(self.h)
        chains = np.zeros((nc, w), np.uint32)
        info = np.zeros((nc, 10))
        self.lib.orcf_get_curvelets(self.h, _f(chains, C.c_uint), _f(info, C.c_double))
        return {"csr": {"offsets": off, "ids": ids, "dist2": d2, "pairs": pairs.value, "max_deg": maxdeg.value,
                        "truncated": trunc.value},
                "valid_pairs": vp, "chains": chains, "info": info}

    def pair_bundle(self, p, a, n):
        fp = self.params(p)
        lo = np.zeros(256, np.float32)
        hi = np.zeros(256, np.float32)
        hyp = self.lib.orcf_pair_bundle(C.byref(fp), *[float(v) for v in a[:3]], *[float(v) for v in n[:3]],
                                        _f(lo, C.c_float), _f(hi, C.c_float))
        return bool(hyp), lo, hi


def read_edges_text(path) -> np.ndarray:
    """Plain-Python reader for fixtures (float() is correctly rounded, like strtod)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(t) for t in line.split()])
    return np.asarray(rows, dtype=np.float64)
