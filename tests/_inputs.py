"""Test inputs: the reference's shipped edge files (kept as gzipped fixtures
under tests/golden/inputs/) and seeded synthetic maps."""
from __future__ import annotations

import gzip
import io
from functools import lru_cache
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
INPUTS = GOLDEN / "inputs"

SHIPPED = ["TO_edges_digit1", "TO_edges_ABC_0006_thresh1", "TO_edges"]


def shipped_text(name: str) -> bytes:
    with gzip.open(INPUTS / f"{name}.txt.gz", "rb") as f:
        return f.read()


@lru_cache(maxsize=None)
def shipped_edges(name: str) -> np.ndarray:
    rows = [[float(t) for t in line.split()] for line in io.StringIO(shipped_text(name).decode()) if line.strip()]
    a = np.asarray(rows, dtype=np.float64)
    a.setflags(write=False)
    return a


@lru_cache(maxsize=None)
def synthetic(preset: str, seed: int) -> np.ndarray:
    from curvelet_construction_on_gpu_amd import synth_edges
    a = synth_edges(preset, seed)
    a.setflags(write=False)
    return a


def edges_for(key: str) -> np.ndarray:
    """'file:<name>' or 'synth:<preset>:<seed>'."""
    kind, _, rest = key.partition(":")
    if kind == "file":
        return shipped_edges(rest)
    preset, seed = rest.split(":")
    return synthetic(preset, int(seed))


# The 3840x2160 map of the full-size batch fixtures (the largest size class of
# the 64-map batch; curves and junction patches in proportion to area).
UHD_MAP = dict(preset="C3", seed=7001, width=3840, height=2160, num_curves=6000, junction_patches=160,
               clutter_fraction=0.3)


def uhd_edges() -> np.ndarray:
    from curvelet_construction_on_gpu_amd import synth_edges
    kw = dict(UHD_MAP)
    return synth_edges(kw.pop("preset"), kw.pop("seed"), **kw)
