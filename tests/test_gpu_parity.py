"""GPU parity: the HIP path against the oracle on the same inputs.

Two checkers:
  oracle     - the restatement with glibc sin/cos/atan2, bit-exact with the
               reference build.  Bar (the reference's): neighbour lists,
               valid-pair counts and curvelet chains bit-exact; curvelet info
               within relative 1e-9 (absolute 1e-12 near zero).
  oracle_cr  - the same restatement with the device's correctly rounded
               math.  Bar: everything bit-exact (CSR, direction classes,
               hypothesis flags, every bundle cell, chains, info), which pins
               the device arithmetic (operand order, rounding points,
               compare-and-assign semantics) independently of libm.
glibc misrounds about 0.1-0.2% of sin/cos/atan2 arguments; that difference is
the only reason the first bar is a tolerance on bundles and info.
"""
import numpy as np
import pytest

from curvelet_construction_on_gpu_amd import CurveletParams, math_eval
from tests._inputs import edges_for

pytestmark = pytest.mark.gpu

INFO_RTOL = 1e-9
INFO_ATOL = 1e-12

CASES = [
    "file:TO_edges_digit1",
    "file:TO_edges_ABC_0006_thresh1",
    "file:TO_edges",
    "synth:C1:1001",
    "synth:C2:1002",
]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_device_math_matches_host_build(ctx):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-8, 8, 20000), rng.uniform(-1e4, 1e4, 2000),
                        np.ldexp(rng.uniform(-1, 1, 2000), -rng.integers(0, 60, 2000)),
                        [0.0, -0.0, 1.5707963267948966, 3.141592653589793, 6.283185307179586],
                        # Payne-Hanek range, up to the largest double
                        np.ldexp(rng.uniform(1, 2, 3000), rng.integers(19, 1024, 3000)) * rng.choice([-1.0, 1.0], 3000),
                        [6381956970095103 * 2.0 ** 797, 1.7976931348623157e308, 2.0 ** 20, 1e22]])
    y = rng.uniform(-8, 8, x.size)
    for op in ("sin", "cos", "atan2", "to2pi", "sincos_dd_sin_hi", "sincos_dd_sin_lo", "sincos_dd_cos_hi",
               "sincos_dd_cos_lo"):
        d = math_eval(op, x, y, ctx=ctx)
        h = math_eval(op, x, y, device=False)
        assert np.array_equal(bits(d), bits(h)), op
    # the transport's sine of a difference, on operands of its range
    a = rng.uniform(-0.6, 3.8, 200000)
    b = rng.uniform(-0.2, 3.4, 200000)
    a[:4000] = b[:4000] + rng.uniform(-1e-6, 1e-6, 4000)
    a[4000:4050] = b[4000:4050]
    for op in ("sin_difference", "sin_difference_settled"):
        assert np.array_equal(bits(math_eval(op, a, b, ctx=ctx)), bits(math_eval(op, a, b, device=False))), op


def test_device_sqrt_fmod_ieee(ctx):
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 50, 20000), np.ldexp(rng.uniform(0, 1, 2000), -rng.integers(0, 900, 2000))])
    assert np.array_equal(math_eval("sqrt", x, ctx=ctx), np.sqrt(x))
    a = rng.uniform(-1e6, 1e6, 20000)
    b = np.full_like(a, 2.0 * 3.141592653589793)
    assert np.array_equal(math_eval("fmod", a, b, ctx=ctx), np.fmod(a, b))


@pytest.mark.parametrize("key", CASES)
def test_csr_bit_exact(ctx, oracle, key):
    e = edges_for(key)
    p = CurveletParams()
    want = oracle.build_csr(e, p, nthreads=0)
    g = ctx.build_graph(e, p.nrad, 3, p.max_candidates)
    assert g.total_neighbor_pairs == want["pairs"]
    assert g.max_num_of_neighbors == want["max_deg"]
    assert np.array_equal(g.neighbor_offsets, want["offsets"])
    assert np.array_equal(g.neighbor_ids, want["ids"])
    assert np.array_equal(bits(g.neighbor_dist2), bits(want["dist2"]))
    assert g.truncated_anchors == want["truncated"]
    assert g.order_ambiguous_anchors == 0


@pytest.mark.parametrize("key", ["file:TO_edges_digit1", "file:TO_edges_ABC_0006_thresh1", "synth:C1:1001"])
def test_anchor_bundles(ctx, oracle, oracle_cr, key):
    e = edges_for(key)
    p = CurveletParams()
    want = oracle.build_csr(e, p)
    oracle_cr.build_csr(e, p)
    ctx.build_graph(e, p.nrad, 3, p.max_candidates)
    off = want["offsets"]
    rng = np.random.default_rng(5)
    anchors = rng.choice(len(e), size=min(len(e), 150), replace=False)
    cells = exact = 0
    for a in anchors:
        deg = int(off[a + 1] - off[a])
        if deg == 0:
            continue
        gd, gh, glo, ghi = ctx.anchor_bundles(p, int(a), deg)
        cd, ch, clo, chi = oracle_cr.anchor_detail(p, int(a), deg)
        od, oh, olo, ohi = oracle.anchor_detail(p, int(a), deg)
        # correctly rounded math on both sides: bit-exact
        assert np.array_equal(gd == 1, (cd & 1) != 0) and np.array_equal(gd == 2, (cd & 2) != 0), a
        assert np.array_equal(gh, ch), a
        assert np.array_equal(bits(glo), bits(clo)) and np.array_equal(bits(ghi), bits(chi)), a
        # glibc math: same decisions, values within tolerance
        assert np.array_equal(gd == 1, (od & 1) != 0) and np.array_equal(gd == 2, (od & 2) != 0), a
        assert np.array_equal(gh, oh), a
        np.testing.assert_allclose(glo, olo, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(ghi, ohi, rtol=1e-9, atol=1e-12)
        cells += glo.size + ghi.size
        exact += int((bits(glo) == bits(olo)).sum() + (bits(ghi) == bits(ohi)).sum())
    assert exact / cells > 0.98, exact / cells


@pytest.mark.parametrize("G", [4, 7])
@pytest.mark.parametrize("key", CASES)
def test_curvelets_parity(ctx, oracle, oracle_cr, key, G):
    e = edges_for(key)
    p = CurveletParams(group_max_sz=G)
    stats = ctx.run(e, p)
    chains, info = ctx.curvelets()
    want = oracle.run(e, p, nthreads=0)
    assert stats["total_neighbor_pairs"] == want["csr"]["pairs"]
    assert stats["valid_pairs"] == want["valid_pairs"]
    assert chains.shape == want["chains"].shape
    assert np.array_equal(chains, want["chains"])
    np.testing.assert_allclose(info, want["info"], rtol=INFO_RTOL, atol=INFO_ATOL)
    exact = oracle_cr.run(e, p, nthreads=0)
    assert np.array_equal(chains, exact["chains"])
    assert np.array_equal(bits(info), bits(exact["info"]))


def test_full_size_map_matches_reference(ctx):
    """The 1920x1080 benchmark map (207k edges, 2.04M pairs, G = 7, degrees up
    to 68, so the 2-chunk fused-kernel variants run) against the reference's
    own run on it: CSR and valid-pair counts and the SHA-256 of every chain
    row; the info bit for bit against the SHA-256 of the oracle with the
    device's correctly rounded math, and 4,096 sampled rows against the
    reference's full-precision info within 1e-9 relative."""
    import hashlib
    import json
    from tests._inputs import GOLDEN
    meta = json.loads((GOLDEN / "ref_cli" / "c3s1003_G7.json").read_text())
    e = edges_for(meta["input"])
    stats = ctx.run(e, CurveletParams(group_max_sz=meta["group_max_sz"]))
    g = ctx.graph()
    assert (g.num_edges, g.total_neighbor_pairs, g.max_num_of_neighbors) == (meta["edges"], meta["pairs"],
                                                                             meta["max_deg"])
    assert g.order_ambiguous_anchors == 0
    assert stats["valid_pairs"] == meta["valid_pairs"]
    chains, info = ctx.curvelets()
    assert chains.shape == (meta["out_h"], meta["out_w"])
    assert hashlib.sha256(np.ascontiguousarray(chains).tobytes()).hexdigest() == meta["full_chain_sha256"]
    assert hashlib.sha256(np.ascontiguousarray(info).tobytes()).hexdigest() == meta["cr_info_sha256"]
    rows = np.load(GOLDEN / "full" / "c3s1003_G7_rows.npz")
    assert np.array_equal(chains[rows["rows"]], rows["chains"])
    np.testing.assert_allclose(info[rows["rows"]], rows["info"], rtol=INFO_RTOL, atol=INFO_ATOL)


@pytest.mark.parametrize("min_cap,waves,grid", [(1, 4, 0), (1, 2, 100), (1, 1, 25), (1000, 4, 50), (8, 4, 100),
                                                (8, 4, 5)])
def test_launch_shapes_same_results(oracle_cr, monkeypatch, min_cap, waves, grid):
    """Every launch shape of the fused kernel (one anchor per wave, or one
    anchor per workgroup of 1, 2 or 4 waves; light-bucket grids from 5% to all
    of their resident capacity) gives the same bits."""
    from curvelet_construction_on_gpu_amd import Context
    monkeypatch.setenv("CVG_COOP_MIN_CAP", str(min_cap))
    monkeypatch.setenv("CVG_COOP_WAVES", str(waves))
    monkeypatch.setenv("CVG_GRID_PCT", str(grid))
    with Context(0) as c:
        for key, G in (("file:TO_edges", 7), ("synth:C2:1002", 5)):
            e = edges_for(key)
            p = CurveletParams(group_max_sz=G)
            c.run(e, p)
            chains, info = c.curvelets()
            want = oracle_cr.run(e, p, nthreads=0)
            assert np.array_equal(chains, want["chains"])
            assert np.array_equal(bits(info), bits(want["info"]))
