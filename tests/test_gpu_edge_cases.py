"""GPU edge cases against the oracle (the reference-exact CPU restatement).

Inputs the shipped files never exercise: exact distance ties (integer and
half-integer lattices, so neighbour lists of 36 with many ties go through the
std::sort replay), lrint's round-half-even cells, max_candidates cuts,
duplicate points and shared cells whose order comes from the reference's
spatial-index sort (replayed on the host), tiny and empty maps, chain sizes
below 3, bundle grids needing 2 and 4 cells per lane, lists of up to 256
neighbours, bundles in global memory, invalid input, and the stage-wise
mirror of the reference API.
"""
import numpy as np
import pytest

from curvelet_construction_on_gpu_amd import (Context, CurveletError, CurveletGPU, CurveletParams, NeighborGraph,
                                              build_neighbor_csr_graph, dt_radians)
from tests._inputs import synthetic

pytestmark = pytest.mark.gpu


def lattice(nx, ny, step=1.0, offset=0.0, seed=0):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx) * step + offset, np.arange(ny) * step + offset)
    th = rng.uniform(0, np.pi, xs.size)
    return np.stack([xs.ravel(), ys.ravel(), th, np.ones(xs.size)], axis=1)


def check_same(ctx, oracle, edges, p):
    """CSR, counts, chains bit-exact and info within 1e-9 against the oracle
    (whose spatial index and neighbour sorts are libstdc++'s std::sort, as in
    the reference, so any within-cell order the reference produces is its)."""
    want = oracle.run(edges, p)
    stats = ctx.run(edges, p)
    g = ctx.graph()
    assert np.array_equal(g.neighbor_offsets, want["csr"]["offsets"])
    assert np.array_equal(g.neighbor_ids, want["csr"]["ids"])
    assert np.array_equal(g.neighbor_dist2, want["csr"]["dist2"])
    chains, info = ctx.curvelets()
    assert stats["valid_pairs"] == want["valid_pairs"]
    assert np.array_equal(chains, want["chains"])
    np.testing.assert_allclose(info, want["info"], rtol=1e-9, atol=1e-12)
    return stats, g, want


def test_integer_lattice_ties(ctx, oracle):
    e = lattice(40, 30, seed=1)
    stats, g, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=7))
    assert g.max_num_of_neighbors == 36  # d2 <= 12.25 inside the 7x7 window
    d = g.neighbor_dist2
    assert len(np.unique(d)) < 10  # heavy ties


def test_half_integer_rows_round_half_even(ctx, oracle):
    # x = k + 0.5: lrint sends 0.5 -> 0, 1.5 -> 2, 2.5 -> 2, 3.5 -> 4, so which
    # in-radius points the +-3-cell window sees depends on round-half-even.
    # Rows 8 px apart never pair; two points sharing a cell are never at equal
    # distance from a half-integer anchor, so no anchor is order-ambiguous.
    rng = np.random.default_rng(2)
    xs = np.arange(40) + 0.5
    e = np.array([[x, y, rng.uniform(0, np.pi), 1.0] for y in (0.5, 8.5, 16.5) for x in xs])
    _, g, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=5))
    assert g.max_num_of_neighbors < 6  # the window misses in-radius points at |dx| = 3
    e2 = lattice(20, 20, step=1.5, offset=-0.5, seed=3)
    check_same(ctx, oracle, e2, CurveletParams(group_max_sz=5))


def test_half_integer_lattice_same_cell_ties_exact(ctx, oracle):
    # a 2-D half-integer lattice puts 2x2 points in one cell; anchors on their
    # bisectors see exact same-cell ties, whose order in the reference comes
    # from libstdc++'s unstable sort of the spatial index: the library counts
    # those anchors and replays that sort, so everything is bit-exact
    e = lattice(25, 25, offset=0.5, seed=2)
    _, g, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=7))
    assert g.order_ambiguous_anchors > 0


def test_lattice_with_shared_cells_long_tied_lists(ctx, oracle):
    # an integer lattice plus a second edge in every third cell: no two edges
    # of one cell tie, but lists of more than 16 have ties between cells, and
    # introsort places tied entries by the positions of all entries, including
    # the two of a shared cell: the within-cell order decides the CSR order
    rng = np.random.default_rng(9)
    base = lattice(30, 30, seed=6)
    extra = base[::3].copy()
    extra[:, 0] += 0.25
    extra[:, 1] -= 0.125
    extra[:, 2] = rng.uniform(0, np.pi, len(extra))
    e = np.concatenate([base, extra])
    e = e[rng.permutation(len(e))]
    _, g, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=7))
    assert g.order_ambiguous_anchors > 0
    # the stage-wise graph call resolves the order the same way
    p = CurveletParams()
    g2 = ctx.build_graph(e, p.nrad, 3, p.max_candidates)
    assert np.array_equal(g2.neighbor_ids, g.neighbor_ids)


def test_order_ambiguous_maps_in_a_batch(ctx, oracle):
    # per map: a 12-edge map with duplicates keeps edge-id order (its own
    # spatial-index sort is an insertion sort), the lattice map needs the replay
    rng = np.random.default_rng(10)
    small = lattice(3, 2, seed=7)
    small = np.concatenate([small, small])
    base = lattice(24, 20, seed=8)
    extra = base[::2].copy()
    extra[:, 0] += 0.3
    extra[:, 2] = rng.uniform(0, np.pi, len(extra))
    big = np.concatenate([base, extra])
    maps = [small, big, synthetic("C1", 1001)[:1500].copy()]
    p = CurveletParams(group_max_sz=6)
    ctx.run_batch(maps, p)
    for i, m in enumerate(maps):
        want = oracle.run(m, p)
        ch, inf = ctx.batch_curvelets(i)
        assert np.array_equal(ch, want["chains"]), i
        np.testing.assert_allclose(inf, want["info"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("max_cand", [1, 5, 17])
def test_max_candidates_cut(ctx, oracle, max_cand):
    e = lattice(30, 30, seed=4)
    p = CurveletParams(max_candidates=max_cand)
    _, g, want = check_same(ctx, oracle, e, p)
    assert g.truncated_anchors == want["csr"]["truncated"] > 0


def test_synthetic_cut_inside_cells(ctx, oracle):
    # subpixel data with several edges per cell: a cut can fall inside a cell,
    # where the kept set depends on the reference's within-cell order
    e = synthetic("C1", 1001)
    p = CurveletParams(max_candidates=6)
    _, g, want = check_same(ctx, oracle, e, p)
    assert g.truncated_anchors == want["csr"]["truncated"] > 0
    assert g.order_ambiguous_anchors > 0


@pytest.mark.parametrize("case", ["lattice cut", "cut inside cells", "C2", "duplicates"])
def test_counting_fill_on_repeated_maps(oracle, case):
    """A map run again on its context takes the guessed grid and buffer sizes,
    where the CSR fill counts the lists and finds the offsets itself (no
    counting pass): the same bits, pair count, largest degree and
    max_candidates count as the oracle, on both runs."""
    d = synthetic("C1", 1001)[:800].copy()
    e, p = {"lattice cut": (lattice(30, 30, seed=4), CurveletParams(max_candidates=5)),
            "cut inside cells": (synthetic("C1", 1001), CurveletParams(max_candidates=6)),
            "C2": (synthetic("C2", 1002), CurveletParams(group_max_sz=7)),
            "duplicates": (np.concatenate([d, d[::7]]), CurveletParams(group_max_sz=7))}[case]
    with Context(0) as c:
        for k in range(3):
            print(f"{case}: run {k}", flush=True)
            _, g, want = check_same(c, oracle, e, p)
            assert g.truncated_anchors == want["csr"]["truncated"], k
            assert g.max_num_of_neighbors == int(np.diff(want["csr"]["offsets"]).max()), k


def test_counting_fill_gives_up_to_the_counting_path(monkeypatch, oracle):
    """The look-backs give up after a bounded number of polls; with the limit
    forced to none, fill workgroups whose predecessors are not ready yet raise
    S_MISS and the graph is built again on the counting path, and row-offset
    scan tiles sum their predecessors' inputs directly: the same bits as the
    oracle on every run, and no wait."""
    monkeypatch.setenv("CVG_LOOKBACK_SPINS", "0")
    e, p = synthetic("C2", 1002), CurveletParams(group_max_sz=7)
    with Context(0) as c:
        for k in range(3):
            print(f"look-back limit 0: run {k}", flush=True)
            _, g, want = check_same(c, oracle, e, p)
            assert g.truncated_anchors == want["csr"]["truncated"], k


def test_duplicates_exact(ctx, oracle):
    e = synthetic("C1", 1001)[:800].copy()
    e = np.concatenate([e, e[::7]])  # exact duplicates: same cell, equal distances
    _, g, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=7))
    assert g.order_ambiguous_anchors > 0


def jittered_subpixel(nx, ny, per_px, seed):
    """per_px edges per pixel, jittered (no exact ties), random orientations."""
    rng = np.random.default_rng(seed)
    k = nx * ny * per_px
    return np.stack([rng.uniform(0, nx, k), rng.uniform(0, ny, k), rng.uniform(0, np.pi, k), np.ones(k)], axis=1)


def test_long_lists_four_chunks(ctx, oracle, oracle_cr):
    # about 5 edges per pixel and max_candidates = 256: lists of 129-256
    # neighbours take the fused kernel's 4-chunk variant
    e = jittered_subpixel(18, 18, 5, seed=12)
    p = CurveletParams(max_candidates=256, group_max_sz=5)
    stats, g, _ = check_same(ctx, oracle, e, p)
    assert g.max_num_of_neighbors > 128
    want = oracle_cr.run(e, p)
    chains, info = ctx.curvelets()
    assert np.array_equal(chains, want["chains"])
    assert np.array_equal(info.view(np.uint64), want["info"].view(np.uint64))


def test_bundles_in_global_memory(ctx, oracle_cr):
    # sx = 0.05, st = 0.04 gives 17 x 15 = 255 cells per bundle; with runs of
    # ~40 neighbours one anchor's bundles exceed the workgroup's LDS and live
    # in global memory
    e = jittered_subpixel(14, 14, 3, seed=13)
    p = CurveletParams(sx=0.05, st=0.04, group_max_sz=5)
    stats = ctx.run(e, p)
    assert stats["max_num_of_neighbors"] > 80
    want = oracle_cr.run(e, p)
    chains, info = ctx.curvelets()
    assert stats["valid_pairs"] == want["valid_pairs"]
    assert np.array_equal(chains, want["chains"])
    assert np.array_equal(info.view(np.uint64), want["info"].view(np.uint64))


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_tiny_maps(ctx, oracle, n):
    e = synthetic("C1", 1001)[:n].copy()
    check_same(ctx, oracle, e, CurveletParams())


def test_isolated_edges(ctx, oracle):
    e = lattice(10, 10, step=10.0, seed=5)
    stats, g, _ = check_same(ctx, oracle, e, CurveletParams())
    assert g.total_neighbor_pairs == 0 and stats["num_curvelets"] == 0


@pytest.mark.parametrize("G", [0, 1, 2, 3, 5, 12])
def test_group_sizes(ctx, oracle, G):
    e = synthetic("C1", 1002)
    stats, _, _ = check_same(ctx, oracle, e, CurveletParams(group_max_sz=G))
    if G < 3:
        assert stats["num_curvelets"] == 0
    assert stats["chain_width"] == G + 1


@pytest.mark.parametrize("sx,st", [(0.05, 0.08), (0.05, 0.04), (0.2, 0.3)])
def test_bundle_grid_shapes(ctx, oracle, sx, st):
    # K = 119 (2 cells per lane), 255 (4 per lane), 9 (one lane in seven)
    e = synthetic("C1", 1001)[:2500].copy()
    check_same(ctx, oracle, e, CurveletParams(sx=sx, st=st, group_max_sz=6))


@pytest.mark.parametrize("nrad,dx,max_k", [(5.0, 0.4, 0.3), (2.0, 0.25, 0.5), (3.5, 0.8, 0.1)])
def test_other_parameters(ctx, oracle, nrad, dx, max_k):
    e = synthetic("C1", 1003)
    check_same(ctx, oracle, e, CurveletParams(nrad=nrad, dx=dx, max_k=max_k))


def test_wider_rows_and_negative_coordinates(ctx, oracle):
    e = synthetic("C2", 1002)[:6000].copy()
    e[:, 0] -= 250.0
    e[:, 1] -= 100.0
    e6 = np.concatenate([e, np.full((len(e), 2), 7.0)], axis=1)  # edge_data_sz = 6
    p = CurveletParams(edge_data_sz=6)
    want = oracle.run(e6, p)
    ctx.run(e6, p)
    chains, info = ctx.curvelets()
    assert np.array_equal(chains, want["chains"])
    np.testing.assert_allclose(info, want["info"], rtol=1e-9, atol=1e-12)


def test_bad_input_rejected(ctx):
    e = synthetic("C1", 1001)[:100].copy()
    e[5, 1] = np.nan
    with pytest.raises(CurveletError) as ei:
        ctx.run(e, CurveletParams())
    assert ei.value.status == 2
    e[5, 1] = 3.0
    e[7, 0] = 1e30  # beyond the coordinate range (any finite orientation is valid)
    with pytest.raises(CurveletError):
        ctx.run(e, CurveletParams())
    with pytest.raises(CurveletError):
        ctx.run(e[:, :3], CurveletParams())  # sz_edge_data < 4 (preprocess.hpp:197)


def test_stagewise_api_matches_reference_shape(ctx, oracle):
    """build_neighbor_csr_graph -> CurveletGPU -> form_pairwise_bundles ->
    build_curvelets_greedy, the call sequence of main.cpp:47-64."""
    e = synthetic("C1", 1004)
    p = CurveletParams(group_max_sz=4)
    graph = NeighborGraph()
    assert build_neighbor_csr_graph(len(e), 4, e, p.nrad, 3, p.max_candidates, 8, graph, ctx=ctx)
    cv = CurveletGPU(len(e), 4, graph, p.dx, dt_radians(p.dt_deg), p.sx, p.st, p.max_k, p.group_max_sz, 8, e,
                     p.nrad, dt_deg=p.dt_deg)
    vp = cv.form_pairwise_bundles()
    cv.build_curvelets_greedy()
    want = oracle.run(e, p)
    assert vp == want["valid_pairs"]
    assert cv.num_curvelets() == len(want["chains"]) and cv.chain_width() == 5
    assert np.array_equal(cv.edge_chain_final, want["chains"])
    assert np.array_equal(graph.neighbor_ids, want["csr"]["ids"])
    bad = NeighborGraph()
    assert not build_neighbor_csr_graph(10, 3, e[:10, :3], p.nrad, 3, 128, 1, bad, ctx=ctx)


def test_results_independent_of_history(ctx):
    """Pools are reused across calls: a map's results must not depend on what ran before."""
    a = synthetic("C1", 1001)
    b = synthetic("C2", 1002)
    p = CurveletParams(group_max_sz=7)
    ctx.run(a, p)
    ca, ia = ctx.curvelets()
    ctx.run(b, p)
    ctx.run(a[:300], CurveletParams(group_max_sz=3))
    ctx.run(a, p)
    ca2, ia2 = ctx.curvelets()
    assert np.array_equal(ca, ca2) and np.array_equal(ia.view(np.uint64), ia2.view(np.uint64))


def test_pageable_downloads_match_page_locked(ctx):
    """Results copied into pageable memory go through page-locked staging blocks
    (several for the 1080p map's 47 MB); they must equal a direct DMA into
    page-locked memory, for the curvelets and the neighbour graph."""
    import ctypes as C
    e = synthetic("C3", 1003)
    ctx.run(e, CurveletParams(group_max_sz=7))
    chains, info = ctx.curvelets()
    lib = ctx._lib
    blocks = []

    def page_locked(shape, dtype):
        p = C.c_void_p()
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        assert lib.cvg_host_alloc(nbytes, C.byref(p)) == 0
        blocks.append(p)
        return np.frombuffer((C.c_char * nbytes).from_address(p.value), dtype=dtype).reshape(shape)

    try:
        pc = page_locked(chains.shape, np.uint32)
        pi = page_locked(info.shape, np.float64)
        ctx.check(lib.cvg_get_curvelets(ctx.handle, pc.ctypes.data_as(C.POINTER(C.c_uint)),
                                        pi.ctypes.data_as(C.POINTER(C.c_double))), "get_curvelets")
        assert info.nbytes > 3 * (8 << 20)  # four staging blocks
        assert np.array_equal(chains, pc) and np.array_equal(info.view(np.uint64), pi.view(np.uint64))
        g = ctx.graph()
        assert g.neighbor_dist2.nbytes > (8 << 20)
        pids = page_locked(g.neighbor_ids.shape, np.int32)
        pd2 = page_locked(g.neighbor_dist2.shape, np.float64)
        ctx.check(lib.cvg_get_csr(ctx.handle, None, pids.ctypes.data_as(C.POINTER(C.c_int)),
                                  pd2.ctypes.data_as(C.POINTER(C.c_double))), "get_csr")
        assert np.array_equal(g.neighbor_ids, pids)
        assert np.array_equal(g.neighbor_dist2.view(np.uint64), pd2.view(np.uint64))
    finally:
        for p in blocks:
            lib.cvg_host_free(p)


def test_runs_without_stats_return_queued_and_match(ctx):
    """A run without stats returns with its work queued; back-to-back runs
    (a different map in between, never queried) and the first query after
    them must give the results and counts of a run that waits."""
    a = synthetic("C1", 1001)
    b = synthetic("C2", 1002)
    p = CurveletParams(group_max_sz=7)
    da = ctx.upload(np.concatenate([a, b]))  # one device array: the upload buffer is reused per call
    db = da + a.size * 8
    st = ctx.run_device(da, len(a), a.shape[1], p)
    ca, ia = ctx.curvelets()
    timed = ctx.recent_fused_ms(64)  # ends with the run above (runs with stats are timed)
    for _ in range(3):
        assert ctx.run_device(db, len(b), b.shape[1], p, stats=False) is None
        assert ctx.run_device(da, len(a), a.shape[1], p, stats=False) is None
    ca2, ia2 = ctx.curvelets()
    assert ca2.shape[0] == st["num_curvelets"]
    assert np.array_equal(ca, ca2) and np.array_equal(ia.view(np.uint64), ia2.view(np.uint64))
    # runs without stats record no timing events unless asked to
    assert ctx.recent_fused_ms(64) == timed
    ctx.set_event_timing(True)
    try:
        for _ in range(3):
            assert ctx.run_device(db, len(b), b.shape[1], p, stats=False) is None
            assert ctx.run_device(da, len(a), a.shape[1], p, stats=False) is None
        ca3, ia3 = ctx.curvelets()
    finally:
        ctx.set_event_timing(False)
    assert np.array_equal(ca, ca3) and np.array_equal(ia.view(np.uint64), ia3.view(np.uint64))
    ms = ctx.recent_fused_ms(6)
    assert len(ms) == 6 and all(m > 0 for m in ms)
    assert ctx.recent_fused_ms(64)[-7] == timed[-1]


def test_downloads_into_device_memory(ctx):
    """A destination the runtime knows as device memory takes the direct copy
    (no host staging): the rows arrive intact."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    e = synthetic("C3", 1003)
    ctx.run(e, CurveletParams(group_max_sz=7))
    chains, info = ctx.curvelets()
    dc, di = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dc), C.c_size_t(chains.nbytes)) == 0
    assert hip.hipMalloc(C.byref(di), C.c_size_t(info.nbytes)) == 0
    try:
        lib = ctx._lib
        ctx.check(lib.cvg_get_curvelets(ctx.handle, C.cast(dc, C.POINTER(C.c_uint)),
                                        C.cast(di, C.POINTER(C.c_double))), "get_curvelets")
        hc, hi = np.empty_like(chains), np.empty_like(info)
        assert hip.hipMemcpy(C.c_void_p(hc.ctypes.data), dc, C.c_size_t(chains.nbytes), 2) == 0  # device to host
        assert hip.hipMemcpy(C.c_void_p(hi.ctypes.data), di, C.c_size_t(info.nbytes), 2) == 0
        assert np.array_equal(hc, chains) and np.array_equal(hi.view(np.uint64), info.view(np.uint64))
    finally:
        hip.hipFree(dc)
        hip.hipFree(di)


def _to_host(c, e, p, pinned=True, rows=None):
    from curvelet_construction_on_gpu_amd import pinned_empty
    c.run(e, p)
    want_chains, want_info = c.curvelets()
    nr = want_chains.shape[0] if rows is None else rows
    w = p.group_max_sz + 1
    if pinned:
        hc, hi = pinned_empty((nr, w), np.uint32), pinned_empty((nr, 10), np.float64)
    else:
        hc, hi = np.empty((nr, w), np.uint32), np.empty((nr, 10), np.float64)
    hc[:] = 0xFFFFFFFF
    hi[:] = np.nan
    got, _ = c.run_to_host(e, p, hc, hi)
    return got, hc, hi, want_chains, want_info


@pytest.mark.parametrize("key,G", [("synth:C3:1003", 7), ("file:TO_edges", 4), ("synth:C1:1001", 5)])
def test_run_to_host_matches_run(ctx, key, G):
    """cvg_run_to_host writes the same rows as a run followed by
    cvg_get_curvelets, into page-locked and pageable arrays."""
    from tests._inputs import edges_for
    e = edges_for(key)
    p = CurveletParams(group_max_sz=G)
    for pinned in (True, False):
        got, hc, hi, wc, wi = _to_host(ctx, e, p, pinned)
        assert got == wc.shape[0]
        assert np.array_equal(hc, wc), (key, pinned)
        assert np.array_equal(hi.view(np.uint64), wi.view(np.uint64)), (key, pinned)


def test_run_to_host_exact_order_and_general_kernel(ctx, oracle):
    """A map whose within-cell order is replayed (the stage runs again) and G
    above 64 (the general kernel) through the same call."""
    d = synthetic("C1", 1001)[:800].copy()
    e = np.concatenate([d, d[::7]])
    p = CurveletParams(group_max_sz=7)
    got, hc, hi, wc, wi = _to_host(ctx, e, p)
    want = oracle.run(e, p)
    assert ctx.graph().order_ambiguous_anchors > 0
    assert np.array_equal(hc, want["chains"]) and np.array_equal(hc, wc)
    assert np.array_equal(hi.view(np.uint64), wi.view(np.uint64))
    p = CurveletParams(group_max_sz=70)
    e = synthetic("C1", 1001)
    got, hc, hi, wc, wi = _to_host(ctx, e, p)
    assert np.array_equal(hc, wc) and np.array_equal(hi.view(np.uint64), wi.view(np.uint64))


def test_run_to_host_too_few_rows(ctx):
    from curvelet_construction_on_gpu_amd import pinned_empty
    e = synthetic("C1", 1001)
    p = CurveletParams(group_max_sz=5)
    ctx.run(e, p)
    wc, wi = ctx.curvelets()
    k = wc.shape[0] // 2
    hc, hi = pinned_empty((k, 6), np.uint32), pinned_empty((k, 10), np.float64)
    with pytest.raises(CurveletError):
        ctx.run_to_host(e, p, hc, hi)
    assert np.array_equal(hc, wc[:k])
    assert np.array_equal(hi.view(np.uint64), wi[:k].view(np.uint64))


def test_consecutive_maps_of_changing_shape(oracle):
    """One context, maps one after another whose grids, pair counts and
    longest lists grow, shift and shrink: the graph first tries the last map's
    grid and buffer sizes and must fall back whenever they do not hold."""
    from curvelet_construction_on_gpu_amd import Context
    base = synthetic("C1", 1001)
    shifted = base.copy()
    shifted[:, 0] += 500.0  # outside the last grid
    dense = np.concatenate([base, base[:, :] + np.array([0.3, 0.2, 0.0, 0.0])])  # longer lists, more pairs
    small = base[:300].copy()
    negative = base.copy()
    negative[:, :2] -= 1000.0
    p = CurveletParams(group_max_sz=6)
    with Context(0) as c:
        for name, e in (("base", base), ("base again", base), ("shifted", shifted), ("dense", dense),
                        ("small", small), ("base", base), ("negative", negative), ("dense", dense)):
            print(f"changing shapes: {name}", flush=True)
            stats = c.run(e, p)
            g = c.graph()
            want = oracle.run(e, p)
            assert np.array_equal(g.neighbor_offsets, want["csr"]["offsets"]), name
            assert np.array_equal(g.neighbor_ids, want["csr"]["ids"]), name
            chains, info = c.curvelets()
            assert stats["valid_pairs"] == want["valid_pairs"], name
            assert np.array_equal(chains, want["chains"]), name
            np.testing.assert_allclose(info, want["info"], rtol=1e-9, atol=1e-12)


def test_consecutive_maps_queued_and_to_host(oracle):
    """A sequence of maps whose grids, pair counts and longest lists change,
    through the calls that return before the counts are read (a queued run,
    then its first query) and through host rows in one call, with a bad map
    in between: every map gives the same results as the oracle, and the
    context stays usable after the error."""
    from curvelet_construction_on_gpu_amd import Context
    base = synthetic("C1", 1001)
    shifted = base.copy()
    shifted[:, 0] += 500.0
    dense = np.concatenate([base, base[:, :] + np.array([0.3, 0.2, 0.0, 0.0])])
    small = base[:300].copy()
    bad = base.copy()
    bad[7, 1] = np.inf
    p = CurveletParams(group_max_sz=6)
    with Context(0) as c:
        for k, (name, e) in enumerate((("base", base), ("base again", base), ("base third", base),
                                       ("shifted", shifted), ("dense", dense), ("dense again", dense),
                                       ("small", small), ("base", base), ("bad", bad), ("base", base),
                                       ("dense", dense))):
            print(f"no-read sequence: {name}", flush=True)
            if name == "bad":
                with pytest.raises(CurveletError) as ei:
                    c.run(e, p, stats=False)
                    c.curvelets()
                assert ei.value.status == 2
                continue
            want = oracle.run(e, p)
            if k % 2 == 0:
                c.run(e, p, stats=False)
                chains, info = c.curvelets()
                g = c.graph()
                assert np.array_equal(g.neighbor_offsets, want["csr"]["offsets"]), name
                assert np.array_equal(g.neighbor_ids, want["csr"]["ids"]), name
            else:
                rows = len(want["chains"])
                hc = np.zeros((rows + 5, 7), np.uint32)
                hi = np.zeros((rows + 5, 10), np.float64)
                got, _ = c.run_to_host(e, p, hc, hi)
                assert got == rows, name
                chains, info = hc[:rows], hi[:rows]
            assert np.array_equal(chains, want["chains"]), name
            np.testing.assert_allclose(info, want["info"], rtol=1e-9, atol=1e-12)


def test_to_host_page_locked_rows_of_changing_maps(oracle):
    """Page-locked outputs receive the last call's row count before this
    call's count is read: maps with more rows (the rest follows), fewer rows,
    rows built again in the reference's within-cell order, and a wider chain
    must all land complete and exact."""
    from curvelet_construction_on_gpu_amd import Context, pinned_empty
    base = synthetic("C1", 1001)
    dense = np.concatenate([base, base[:, :] + np.array([0.3, 0.2, 0.0, 0.0])])
    small = base[:300].copy()
    dup = base[:800].copy()
    dup = np.concatenate([dup, dup[::7]])  # exact duplicates: the within-cell order is replayed
    seq = [("base", base, 6), ("base again", base, 6), ("dense", dense, 6), ("small", small, 6),
           ("duplicates", dup, 6), ("base", base, 6), ("base, G = 9", base, 9), ("dense", dense, 6)]
    with Context(0) as c:
        for name, e, G in seq:
            print(f"page-locked rows: {name}", flush=True)
            p = CurveletParams(group_max_sz=G)
            want = oracle.run(e, p)
            rows = len(want["chains"])
            cap = 60000
            hc = pinned_empty((cap, G + 1), np.uint32)
            hi = pinned_empty((cap, 10), np.float64)
            hc[:] = 0xFFFFFFFF
            hi[:] = np.nan
            got, _ = c.run_to_host(e, p, hc, hi)
            assert got == rows, name
            assert np.array_equal(hc[:rows], want["chains"]), name
            np.testing.assert_allclose(hi[:rows], want["info"], rtol=1e-9, atol=1e-12)
