"""cvg_run_device_noalloc against cvg_run_device: the same map through the
entry point that allocates nothing (intermediate buffers in the library's
static device scratch array, outputs in caller buffers) and through the
pooled one give the same bits, on two calls in a row (the second smaller),
and sizes beyond the array or the caller's rows fail with clear errors."""
import numpy as np
import pytest

from curvelet_construction_on_gpu_amd import Context, CurveletError, CurveletParams
from curvelet_construction_on_gpu_amd.hip import DeviceArray
from tests._inputs import edges_for

pytestmark = pytest.mark.gpu

SCRATCH_MIB = 64  # CVG_NOALLOC_SCRATCH_MIB


def pooled(ctx, e, p):
    d = ctx.upload(e)
    st = ctx.run_device(d, len(e), e.shape[1], p)
    chains, info = ctx.curvelets()
    return chains, info, st["valid_pairs"]


def noalloc(ctx, e, p, rows):
    d = ctx.upload(e)
    w = p.group_max_sz + 1
    out_c = DeviceArray((rows, w), np.uint32)
    out_i = DeviceArray((rows, 10), np.float64)
    out_c.fill_bytes(0xFF)
    out_i.fill_bytes(0xFF)
    nc, vp = ctx.run_device_noalloc(d, len(e), e.shape[1], p, out_c.ptr, out_i.ptr, rows)
    return out_c.get(nc), out_i.get(nc), vp


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def nctx():
    c = Context(0)
    yield c
    c.close()


def test_noalloc_matches_pooled_two_calls(nctx):
    cases = [("synth:C1:1001", 7), ("file:TO_edges_ABC_0006_thresh1", 4)]  # second call smaller
    for key, G in cases:
        e = np.ascontiguousarray(edges_for(key))
        p = CurveletParams(group_max_sz=G)
        print(f"noalloc: {key} G={G} edges={len(e)}", flush=True)
        want_c, want_i, want_v = pooled(nctx, e, p)
        rows = len(want_c) + 17
        got_c, got_i, got_v = noalloc(nctx, e, p, rows)
        assert got_v == want_v
        assert got_c.shape == want_c.shape
        assert np.array_equal(got_c, want_c)
        assert np.array_equal(bits(got_i), bits(want_i))


def test_noalloc_rejects_sizes(nctx):
    e = np.ascontiguousarray(edges_for("synth:C1:1001"))
    p = CurveletParams(group_max_sz=4)
    want_c, _, _ = pooled(nctx, e, p)
    # too few caller rows: an error that names the row count, nothing beyond written
    rows = len(want_c) // 2
    out_c = DeviceArray((rows + 1, 5), np.uint32)
    out_i = DeviceArray((rows + 1, 10), np.float64)
    out_c.fill_bytes(0x5A)
    d = nctx.upload(e)
    with pytest.raises(CurveletError, match=str(len(want_c))):
        nctx.run_device_noalloc(d, len(e), 4, p, out_c.ptr, out_i.ptr, rows)
    assert np.all(out_c.get()[rows] == 0x5A5A5A5A)
    # a map beyond the scratch array
    big = np.ascontiguousarray(edges_for("synth:C3:1003"))
    db = nctx.upload(big)
    with pytest.raises(CurveletError, match=f"{SCRATCH_MIB} MiB"):
        nctx.run_device_noalloc(db, len(big), 4, CurveletParams(group_max_sz=7), out_c.ptr, out_i.ptr, rows)
    # the pooled path still works afterwards and matches a fresh no-alloc run
    got_c, _, _ = noalloc(nctx, e, p, len(want_c))
    assert np.array_equal(got_c, pooled(nctx, e, p)[0])
