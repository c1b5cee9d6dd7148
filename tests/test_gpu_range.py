"""The drop-in range: inputs the reference accepts beyond the fused kernel's
and the dense cell grid's shapes.

- Coordinates spanning more pixel cells than the dense grid holds (2^28) go
  to the sorted-key cell layout; forcing that layout on ordinary maps gives
  the same bits as the dense grid (including the replayed within-cell order).
- group_max_sz above 64 and neighbour lists above 256 go to the general
  curvelet kernel; forcing that kernel on ordinary maps gives the same bits as
  the fused kernel.
- Values outside the supported domain are rejected with CVG_ERR_BAD_INPUT.
Bars as in the parity tests: chains and CSR bit-exact with the glibc oracle,
info within 1e-9 relative; everything bit-exact with the correctly rounded
oracle.
"""
import numpy as np
import pytest

from curvelet_construction_on_gpu_amd import Context, CurveletError, CurveletParams
from tests._inputs import edges_for, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def run(c, e, p):
    stats = c.run(e, p)
    chains, info = c.curvelets()
    return stats, c.graph(), chains, info


def check(c, oracle, oracle_cr, e, p):
    stats, g, chains, info = run(c, e, p)
    want = oracle.run(e, p, nthreads=0)
    assert np.array_equal(g.neighbor_offsets, want["csr"]["offsets"])
    assert np.array_equal(g.neighbor_ids, want["csr"]["ids"])
    assert stats["valid_pairs"] == want["valid_pairs"]
    assert np.array_equal(chains, want["chains"])
    np.testing.assert_allclose(info, want["info"], rtol=1e-9, atol=1e-12)
    exact = oracle_cr.run(e, p, nthreads=0)
    assert np.array_equal(chains, exact["chains"])
    assert np.array_equal(bits(info), bits(exact["info"]))
    return stats, g


def lattice(nx, ny, offset=0.0, seed=0):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx) + offset, np.arange(ny) + offset)
    th = rng.uniform(0, np.pi, xs.size)
    return np.stack([xs.ravel(), ys.ravel(), th, np.ones(xs.size)], axis=1)


def jittered(nx, ny, per_px, seed):
    rng = np.random.default_rng(seed)
    k = nx * ny * per_px
    return np.stack([rng.uniform(0, nx, k), rng.uniform(0, ny, k), rng.uniform(0, np.pi, k), np.ones(k)], axis=1)


def sparse_inputs():
    d = synthetic("C1", 1001)[:800].copy()
    return [("C2", synthetic("C2", 1002), CurveletParams(group_max_sz=7)),
            ("TO_edges", edges_for("file:TO_edges"), CurveletParams(group_max_sz=5)),
            ("half-integer lattice", lattice(25, 25, offset=0.5, seed=2), CurveletParams(group_max_sz=7)),
            ("duplicates", np.concatenate([d, d[::7]]), CurveletParams(group_max_sz=7)),
            ("cut", synthetic("C2", 1002)[:5000].copy(), CurveletParams(max_candidates=9))]


def test_sorted_key_cells_same_bits(monkeypatch, ctx, oracle, oracle_cr):
    monkeypatch.setenv("CVG_SPARSE_CELLS", "1")
    with Context(0) as c:
        for name, e, p in sparse_inputs():
            print(f"sorted-key cells: {name}", flush=True)
            stats, g = check(c, oracle, oracle_cr, e, p)
            _, gd, chains_d, info_d = run(ctx, e, p)
            assert np.array_equal(g.neighbor_ids, gd.neighbor_ids), name
            assert g.order_ambiguous_anchors == gd.order_ambiguous_anchors, name


def test_coordinates_spanning_billions_of_cells(ctx, oracle, oracle_cr):
    """Three copies of a map placed 2^31 - 10^7 apart in x and y: a dense grid
    would need ~1.8e19 cells; the sorted-key layout handles it."""
    base = synthetic("C1", 1001)[:3000].copy()
    parts = []
    for dx, dy in ((-2.137e9, -2.0e9), (0.0, 0.0), (2.137e9, 2.1e9)):
        b = base.copy()
        b[:, 0] += dx
        b[:, 1] += dy
        parts.append(b)
    e = np.concatenate(parts)
    stats, g = check(ctx, oracle, oracle_cr, e, CurveletParams(group_max_sz=6))
    assert stats["num_curvelets"] > 0


def straight_line(n=600, seed=5):
    """Edges every 0.05 px along a line, nearly collinear and aligned: most
    hypotheses keep the running bundle valid, so chains grow past 64."""
    rng = np.random.default_rng(seed)
    x = 5 + np.arange(n) * 0.05 + rng.uniform(-0.004, 0.004, n)
    y = 10 + rng.uniform(-0.004, 0.004, n)
    th = np.mod(rng.uniform(-0.002, 0.002, n), 2 * np.pi)
    return np.stack([x, y, th, np.ones(n)], axis=1)


@pytest.mark.parametrize("G", [65, 100])
def test_group_size_above_64(ctx, oracle, oracle_cr, G):
    e = jittered(30, 24, 3, seed=41)
    stats, g = check(ctx, oracle, oracle_cr, e, CurveletParams(group_max_sz=G, max_candidates=200))
    assert stats["chain_width"] == G + 1
    assert g.max_num_of_neighbors > 64
    stats, g = check(ctx, oracle, oracle_cr, straight_line(), CurveletParams(group_max_sz=G, max_candidates=200))
    chains, _ = ctx.curvelets()
    assert (chains[:, 1:] > 0).sum(axis=1).max() > 64


def test_lists_above_256(ctx, oracle, oracle_cr):
    """~10 edges per pixel: lists of up to ~400 neighbours with
    max_candidates 512."""
    e = jittered(30, 30, 10, seed=42)
    p = CurveletParams(group_max_sz=5, max_candidates=512)
    stats, g = check(ctx, oracle, oracle_cr, e, p)
    assert g.max_num_of_neighbors > 256


def test_general_kernel_same_bits_as_fused(monkeypatch, ctx):
    """The general kernel on the maps the fused kernel takes: same bits."""
    monkeypatch.setenv("CVG_GENERIC_KERNEL", "1")
    with Context(0) as c:
        for key, G in (("file:TO_edges_digit1", 4), ("file:TO_edges", 7), ("synth:C2:1002", 5),
                       ("synth:C3:1003", 7)):
            print(f"general kernel: {key} G={G}", flush=True)
            e = edges_for(key)
            p = CurveletParams(group_max_sz=G)
            s1, _, ch1, in1 = run(c, e, p)
            s2, _, ch2, in2 = run(ctx, e, p)
            assert s1["valid_pairs"] == s2["valid_pairs"], key
            assert np.array_equal(ch1, ch2), key
            assert np.array_equal(bits(in1), bits(in2)), key


def test_float_group_size_above_64(ctx):
    from oracle import OracleF32
    e = jittered(30, 24, 3, seed=43).astype(np.float32)
    p = CurveletParams(group_max_sz=80, max_candidates=200)
    stats = ctx.run_f32(e, p)
    chains, info = ctx.curvelets()
    want = OracleF32(cr=True).run(e, p, nthreads=0)
    assert stats["valid_pairs"] == want["valid_pairs"]
    assert np.array_equal(chains, want["chains"])
    assert np.array_equal(bits(info), bits(want["info"]))


def test_orientations_of_any_magnitude(ctx, oracle, oracle_cr):
    """Orientations far outside [0, 2 pi) (the reference takes any finite
    value): the trig of theta, theta +- dt and the grid angles goes through
    the large-argument reduction."""
    e = synthetic("C1", 1001)[:2500].copy()
    rng = np.random.default_rng(44)
    e[:, 2] += np.ldexp(1.0, rng.integers(10, 60, len(e))) * rng.choice([-1.0, 1.0], len(e))
    stats, g = check(ctx, oracle, oracle_cr, e, CurveletParams(group_max_sz=6))
    assert stats["num_curvelets"] > 0


@pytest.mark.parametrize("col,value", [(0, 2.2e9), (1, -2.2e9), (0, np.inf), (2, np.inf), (2, np.nan)])
def test_out_of_domain_rejected(ctx, col, value):
    e = synthetic("C1", 1001)[:200].copy()
    e[17, col] = value
    with pytest.raises(CurveletError) as ei:
        ctx.run(e, CurveletParams())
    assert ei.value.status == 2
    # the context stays usable
    ctx.run(synthetic("C1", 1001)[:100].copy(), CurveletParams())
