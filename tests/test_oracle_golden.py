"""The oracle against the reference's own outputs (CPU).

The fixtures under tests/golden/ were written by tests/golden/make_golden.py
from the reference's cpu_curvelet code built in place (oracle/_ref): its
CURVELET binary's stdout counts and chain/info text files, and full-precision
arrays from its headers (CSR, valid pairs, chains, info, per-pair bundles and
direction values).  The glibc oracle must reproduce every one of them bit for
bit; the correctly rounded variant (the GPU path's math) must give the same
chains.  Where oracle/_ref is built (only where the reference tree exists),
the oracle is also compared live with it on inputs the fixtures do not hold.
"""
import gzip
import hashlib
import json

import numpy as np
import pytest

from curvelet_construction_on_gpu_amd import CurveletParams, write_chain_file, write_info_file
from tests._inputs import GOLDEN, edges_for

_META = {p.stem: json.loads(p.read_text()) for p in (GOLDEN / "ref_cli").glob("*.json")}
CASES = sorted(k for k, m in _META.items() if "input" in m and m.get("scalar", "double") == "double")
FLOAT_CASES = sorted(k for k, m in _META.items() if m.get("scalar") == "float")
FULL = sorted(p.stem for p in (GOLDEN / "full").glob("*.npz") if not p.stem.endswith("_rows"))
TEXT = sorted(p.name[:-len("_chain.txt.gz")] for p in (GOLDEN / "ref_cli").glob("*_chain.txt.gz")
              if not p.name.startswith(tuple(FLOAT_CASES)))


def case_meta(name):
    return json.loads((GOLDEN / "ref_cli" / f"{name}.json").read_text())


def case_edges(meta):
    key = meta["input"]
    if key.startswith("synth:"):
        _, preset, seed = key.split(":")
        f = GOLDEN / "inputs" / f"{preset}_{seed}.npy"
        if f.exists():
            return np.load(f)
    return edges_for(key)


@pytest.fixture(scope="module")
def runs(oracle):
    cache = {}

    def get(name):
        if name not in cache:
            meta = case_meta(name)
            cache[name] = (meta, oracle.run(case_edges(meta), CurveletParams(group_max_sz=meta["group_max_sz"])))
        return cache[name]
    return get


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def test_fixture_inventory():
    assert {"digit1_G4", "digit1_G7", "abc_G4", "abc_G7", "toedges_G4", "toedges_G7", "c1s1001_G4",
            "c3s1003_G7"} <= set(CASES)
    assert set(FULL) <= set(CASES) and set(TEXT) <= set(CASES)


@pytest.mark.parametrize("name", CASES)
def test_counts_and_hashes(runs, name):
    """Counts printed by the reference CLI and SHA-256 of its full-precision arrays."""
    meta, r = runs(name)
    e = case_edges(meta)
    assert len(e) == meta["edges"]
    assert r["csr"]["pairs"] == meta["pairs"]
    assert r["csr"]["max_deg"] == meta["max_deg"]
    assert r["valid_pairs"] == meta["valid_pairs"]
    assert r["chains"].shape == (meta["out_h"], meta["out_w"])
    assert sha(np.ascontiguousarray(r["chains"]).tobytes()) == meta["full_chain_sha256"]
    assert sha(np.ascontiguousarray(r["info"]).tobytes()) == meta["full_info_sha256"]
    assert np.array_equal(r["info"].sum(axis=0), np.array(meta["info_colsum"]))


@pytest.mark.parametrize("name", FULL)
def test_full_precision_arrays(runs, name):
    meta, r = runs(name)
    g = np.load(GOLDEN / "full" / f"{name}.npz")
    assert np.array_equal(r["csr"]["offsets"], g["offsets"])
    assert np.array_equal(r["csr"]["ids"], g["ids"])
    assert np.array_equal(r["csr"]["dist2"].view(np.uint64), g["dist2"].view(np.uint64))
    assert r["valid_pairs"] == int(g["valid_pairs"])
    assert np.array_equal(r["chains"], g["chains"])
    assert np.array_equal(r["info"].view(np.uint64), g["info"].view(np.uint64))


@pytest.mark.parametrize("name", CASES)
def test_output_text_identical(runs, name, tmp_path):
    """The package's writers on the oracle's arrays reproduce the reference CLI's
    chain and info files byte for byte (hash for all cases, full text where kept)."""
    meta, r = runs(name)
    chain_p = tmp_path / "chain.txt"
    info_p = tmp_path / "info.txt"
    write_chain_file(chain_p, r["chains"])
    write_info_file(info_p, r["info"])
    chain, info = chain_p.read_bytes(), info_p.read_bytes()
    assert sha(chain) == meta["chain_sha256"]
    assert sha(info) == meta["info_sha256"]
    if name in TEXT:
        with gzip.open(GOLDEN / "ref_cli" / f"{name}_chain.txt.gz") as f:
            assert chain == f.read()
        with gzip.open(GOLDEN / "ref_cli" / f"{name}_info.txt.gz") as f:
            assert info == f.read()


@pytest.mark.parametrize("name", ["abc_pairs", "c1_pairs"])
def test_pair_bundles_and_direction(oracle, name):
    """compute_curve_bundle (both transport branches, close pairs, collinear and
    mirrored geometries) and the backward-run direction value, bit for bit."""
    g = np.load(GOLDEN / "pairs" / f"{name}.npz")
    p = CurveletParams()
    for a, n, lo, hi, hyp, dot in zip(g["anchor"], g["neighbour"], g["lo"], g["hi"], g["hyp"], g["dot_sincos"]):
        h, olo, ohi = oracle.pair_bundle(p, a, n)
        assert h == bool(hyp)
        assert np.array_equal(olo.view(np.uint64), lo.view(np.uint64))
        assert np.array_equal(ohi.view(np.uint64), hi.view(np.uint64))
        assert oracle.dir_dot(a, n, run=1) == dot or (np.isnan(dot) and np.isnan(oracle.dir_dot(a, n, run=1)))


@pytest.mark.parametrize("name", ["digit1_G4", "digit1_G7", "abc_G4", "c1s1001_G4"])
def test_correctly_rounded_variant_same_chains(runs, oracle_cr, name):
    """The GPU path's math (correctly rounded sin/cos/atan2) gives the reference's
    chains on every fixture; info differs from glibc's at most in the last bits."""
    meta, r = runs(name)
    cr = oracle_cr.run(case_edges(meta), CurveletParams(group_max_sz=meta["group_max_sz"]))
    assert cr["valid_pairs"] == r["valid_pairs"]
    assert np.array_equal(cr["chains"], r["chains"])
    np.testing.assert_allclose(cr["info"], r["info"], rtol=1e-9, atol=1e-12)


def test_thread_count_independent(oracle):
    e = edges_for("file:TO_edges_ABC_0006_thresh1")
    p = CurveletParams(group_max_sz=5)
    a = oracle.run(e, p, nthreads=1)
    b = oracle.run(e, p, nthreads=8)
    assert np.array_equal(a["csr"]["ids"], b["csr"]["ids"])
    assert np.array_equal(a["chains"], b["chains"])
    assert np.array_equal(a["info"].view(np.uint64), b["info"].view(np.uint64))


# --- live comparison with the reference build (skipped where it is absent) ---

def _ref():
    from oracle import Ref
    if not Ref.available():
        pytest.skip("oracle/_ref not built (the reference tree is not present here)")
    return Ref()


def _lattice(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    return np.stack([xs.ravel(), ys.ravel(), rng.uniform(0, np.pi, xs.size), np.ones(xs.size)], axis=1)


@pytest.mark.parametrize("case", ["c2cut", "lattice_ties", "lattice_trunc", "c1_g7_wide"])
def test_live_reference(oracle, case):
    ref = _ref()
    if case == "c2cut":
        e, p = edges_for("synth:C2:1002")[:7000], CurveletParams(group_max_sz=6)
    elif case == "lattice_ties":
        e, p = _lattice(30, 24, 3), CurveletParams(group_max_sz=7)
    elif case == "lattice_trunc":
        e, p = _lattice(24, 24, 4), CurveletParams(max_candidates=9)
    else:
        e, p = edges_for("synth:C1:1004"), CurveletParams(group_max_sz=7, sx=0.05, st=0.04, max_k=0.5)
    want = ref.run(e, p, nthreads=4)
    got = oracle.run(e, p, nthreads=4)
    assert np.array_equal(got["csr"]["offsets"], want["csr"]["offsets"])
    assert np.array_equal(got["csr"]["ids"], want["csr"]["ids"])
    assert np.array_equal(got["csr"]["dist2"], want["csr"]["dist2"])
    assert got["valid_pairs"] == want["valid_pairs"]
    assert np.array_equal(got["chains"], want["chains"])
    assert np.array_equal(got["info"].view(np.uint64), want["info"].view(np.uint64))


def test_full_size_fixtures(oracle, oracle_cr):
    """The full-size fixtures the GPU tests check against: the 1080p map's
    correctly rounded oracle hashes and sampled reference rows, and the
    3840x2160 map's reference chain hash, counts and sampled rows (the glibc
    oracle reproduces the reference's rows bit for bit)."""
    from tests._inputs import uhd_edges
    meta = case_meta("c3s1003_G7")
    e = case_edges(meta)
    p = CurveletParams(group_max_sz=meta["group_max_sz"])
    w = oracle_cr.run(e, p, nthreads=0, valid_pairs=False)
    assert sha(np.ascontiguousarray(w["chains"]).tobytes()) == meta["full_chain_sha256"] == meta["cr_chain_sha256"]
    assert sha(np.ascontiguousarray(w["info"]).tobytes()) == meta["cr_info_sha256"]
    rows = np.load(GOLDEN / "full" / "c3s1003_G7_rows.npz")
    np.testing.assert_allclose(w["info"][rows["rows"]], rows["info"], rtol=1e-9, atol=1e-12)
    um = json.loads((GOLDEN / "ref_cli" / "uhd7001_G4.json").read_text())
    e = uhd_edges()
    assert sha(np.ascontiguousarray(e).tobytes()) == um["input_sha256"]
    p = CurveletParams(group_max_sz=um["group_max_sz"])
    r = oracle.run(e, p, nthreads=0)
    assert (r["csr"]["pairs"], r["csr"]["max_deg"], r["valid_pairs"]) == (um["pairs"], um["max_deg"],
                                                                          um["valid_pairs"])
    assert sha(np.ascontiguousarray(r["chains"]).tobytes()) == um["full_chain_sha256"]
    rows = np.load(GOLDEN / "full" / "uhd7001_G4_rows.npz")
    assert np.array_equal(r["chains"][rows["rows"]], rows["chains"])
    assert np.array_equal(r["info"][rows["rows"]].view(np.uint64), rows["info"].view(np.uint64))
    w = oracle_cr.run(e, p, nthreads=0, valid_pairs=False)
    assert sha(np.ascontiguousarray(w["info"]).tobytes()) == um["cr_info_sha256"]


def float_edges(meta):
    """The float edges the reference's --float reader produces from the case's text."""
    import tempfile
    from pathlib import Path
    from curvelet_construction_on_gpu_amd import read_TO_edges_from_file_f32
    from tests._inputs import shipped_text
    key = meta["input"]
    if key.startswith("file:"):
        text = shipped_text(key[5:])
    else:
        e = case_edges({"input": key})
        text = ("\n".join("\t".join(repr(float(v)) for v in row) for row in e) + "\n").encode()
    assert sha(text) == meta["input_sha256"]
    with tempfile.TemporaryDirectory() as td:
        f = Path(td) / "edges.txt"
        f.write_bytes(text)
        e32 = read_TO_edges_from_file_f32(f)
    assert sha(np.ascontiguousarray(e32).tobytes()) == meta["edges_f32_sha256"]
    return e32


@pytest.mark.parametrize("name", FLOAT_CASES)
def test_float_oracle_against_reference(name, tmp_path):
    """The float-instantiation restatement (glibc) reproduces the reference's
    --float run bit for bit (counts, full-precision chains and info, the CLI's
    text files); the correctly rounded variant (the GPU path's math) gives the
    same chains and valid pairs on these inputs; glibc's sinf and cosf are not
    correctly rounded (about 1.3% of arguments differ by an ulp), which moves
    the info of 3-7% of rows by a few float ulps (at most 3.3e-7 absolute
    here)."""
    from oracle import OracleF32
    meta = _META[name]
    e = float_edges(meta)
    p = CurveletParams(group_max_sz=meta["group_max_sz"])
    r = OracleF32().run(e, p, nthreads=0)
    assert (r["csr"]["pairs"], r["csr"]["max_deg"], r["valid_pairs"]) == (meta["pairs"], meta["max_deg"],
                                                                          meta["valid_pairs"])
    assert r["chains"].shape == (meta["out_h"], meta["out_w"])
    assert sha(np.ascontiguousarray(r["chains"]).tobytes()) == meta["full_chain_sha256"]
    assert sha(np.ascontiguousarray(r["info"]).tobytes()) == meta["full_info_sha256"]
    write_chain_file(tmp_path / "c.txt", r["chains"])
    write_info_file(tmp_path / "i.txt", r["info"])
    assert sha((tmp_path / "c.txt").read_bytes()) == meta["chain_sha256"]
    assert sha((tmp_path / "i.txt").read_bytes()) == meta["info_sha256"]
    c = OracleF32(cr=True).run(e, p, nthreads=0)
    assert c["valid_pairs"] == r["valid_pairs"]
    assert np.array_equal(c["chains"], r["chains"])
    np.testing.assert_allclose(c["info"], r["info"], rtol=1e-4, atol=1e-6)


def test_float_reader_is_strtof():
    """The --float reader converts each decimal token straight to float, as
    istream >> float does (libc strtof), not through a double."""
    import ctypes as C
    import tempfile
    from pathlib import Path
    from curvelet_construction_on_gpu_amd import read_TO_edges_from_file_f32
    libc = C.CDLL("libc.so.6")
    libc.strtof.restype = C.c_float
    libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
    rng = np.random.default_rng(8)
    # decimal strings that sit just beside float midpoints: the double nearest
    # to them rounds to the other float
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    toks = []
    for _ in range(400):
        f = np.float32(rng.uniform(0.5, 1000))
        g = np.nextafter(f, np.float32(np.inf))
        mid = (np.float64(f) + np.float64(g)) / 2
        ulp = np.nextafter(mid, np.inf) - mid
        # a quarter of a double ulp above the midpoint: the nearest double is
        # the midpoint itself, which then rounds to the even float
        toks.append(str(Decimal(float(mid)) + Decimal(float(ulp)) / 4))
        toks.append(f"{mid:.25g}")
        toks.append(repr(float(rng.uniform(-500, 500))))
    toks += ["0.1", "1e-50", "-0", "3.4028234e38", "1e-38", "7", "-2.5e-3"]
    while len(toks) % 4:
        toks.append("1")
    text = "\n".join(" ".join(toks[i:i + 4]) for i in range(0, len(toks), 4)) + "\n"
    with tempfile.TemporaryDirectory() as td:
        f = Path(td) / "e.txt"
        f.write_text(text)
        got = read_TO_edges_from_file_f32(f).ravel()
    want = np.array([libc.strtof(t.encode(), None) for t in toks], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the double route would differ on some of these
    via_double = np.array([np.float32(float(t)) for t in toks], dtype=np.float32)
    assert not np.array_equal(via_double.view(np.uint32), want.view(np.uint32))
