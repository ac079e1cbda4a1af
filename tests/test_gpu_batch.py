"""Launch tiles and batches of grids (atsat_autotune, atsat_set_tile, atsat_group_*).

The tile of the RK4 stage kernel (rows per theta band, rings per workgroup) and the batching of
several grids into one launch per stage only change which workgroup computes a cell, never the
arithmetic of a cell: every result must be bit-identical to the context stepped alone with the
default tile. The contexts stepped alone are checked against the CPU oracle by
tests/test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

import atsat_amd
from atsat_amd import abi
from util import HOT_FIELDS

pytestmark = pytest.mark.gpu

# three grid shapes with different ring lengths (threads per ring 128, 64, 128) and members
GRIDS = [((21, 61, 121), 2), ((11, 31, 61), 3), ((11, 46, 91), 2)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def defaults():
    return atsat_amd.defaults("resident")


def _ctx(shape, members, prec, c, k, seed):
    g = atsat_amd.small_grid(*shape)
    ctx = atsat_amd.Context(g, c, k, members, prec)
    for m in range(members):
        st = atsat_amd.gen_v1(g, c, seed=seed, member=m, t_sigma=0.5, w_sigma=2.0 / 470.0)
        ctx.upload_state(st, member=m)
    return ctx


def _fields(ctx, members):
    return {(m, f): ctx.download(f, member=m) for m in range(members) for f in HOT_FIELDS}


def _assert_same(a, b, what):
    bad = [key for key in a if not np.array_equal(a[key], b[key], equal_nan=True)]
    assert not bad, "%s: %d fields differ, first %s" % (what, len(bad), bad[:4])


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32])
def test_tiles_bit_identical(defaults, prec):
    _, c, k = defaults
    shape, M = GRIDS[0]
    ref = None
    for band, rings in [(8, 1), (6, 3), (4, 2), (1, 1), (32, 2)]:
        say("case tile band %d rings %d, grid %s x %d, precision %d" % (band, rings, shape, M, prec))
        ctx = _ctx(shape, M, prec, c, k, 11)
        ctx.set_tile(band, rings)
        assert ctx.tile() == (band, rings)
        ctx.run(1, 4, False)
        ctx.synchronize()
        out = _fields(ctx, M)
        ctx.close()
        if ref is None:
            ref = out
        else:
            _assert_same(out, ref, "tile (%d, %d)" % (band, rings))


def test_bad_tile_rejected(defaults):
    _, c, k = defaults
    ctx = _ctx(GRIDS[0][0], 1, abi.PREC_FP64, c, k, 11)
    with pytest.raises(atsat_amd.AtsatError):
        ctx.set_tile(5, 2)                       # rings must divide band
    with pytest.raises(atsat_amd.AtsatError):
        ctx.set_tile(14, 7)                      # 7 x 128 threads > 768
    ctx.close()


def test_autotune_keeps_state(defaults):
    _, c, k = defaults
    shape, M = GRIDS[0]
    say("case autotune, grid %s x %d, FP32" % (shape, M))
    a = _ctx(shape, M, abi.PREC_FP32, c, k, 12)
    a.run(1, 1, False)
    band, rings, ms = a.autotune()
    say("  tuned tile band %d rings %d, stage 1 %.3f ms" % (band, rings, ms))
    assert ms > 0 and a.tile() == (band, rings)
    a.run(2, 3, False)
    a.synchronize()
    b = _ctx(shape, M, abi.PREC_FP32, c, k, 12)
    b.run(1, 4, False)
    b.synchronize()
    _assert_same(_fields(a, M), _fields(b, M), "after autotune")
    a.close()
    b.close()


@pytest.mark.parametrize("use_graph,layout", [(False, "tuned"), (True, "tuned"), (True, "merged"), (True, "side"),
                                              (False, "queue"), (True, "queue")])
def test_group_matches_contexts_alone(defaults, use_graph, layout):
    _, c, k = defaults
    prec = abi.PREC_FP32
    say("case group of %s, graph %s, layout %s" % ([s for s, _ in GRIDS], use_graph, layout))
    alone = []
    for gi, (shape, M) in enumerate(GRIDS):
        ctx = _ctx(shape, M, prec, c, k, 100 + gi)
        ctx.run(1, 5, False)
        ctx.synchronize()
        alone.append(_fields(ctx, M))
        ctx.close()
    ctxs = [_ctx(shape, M, prec, c, k, 100 + gi) for gi, (shape, M) in enumerate(GRIDS)]
    grp = atsat_amd.Group(ctxs)
    grp.run(1, 1, False)
    ms = grp.autotune()
    if layout != "tuned":
        grp.set_layout({"side": 0, "merged": 1, "queue": 2}[layout])
    merged = grp.merged()
    tiles = grp.tiles()
    say("  tuned %.3f ms per stage; batched launches: %s; tiles %s" % (ms, merged, tiles))
    for (shape, _), (band, rings, t) in zip(GRIDS, tiles):
        assert band % rings == 0
        assert rings * ((shape[2] - 1 + 63) // 64 * 64) <= t
        assert t == tiles[0][2] or not merged
    grp.run(2, 4, use_graph)
    grp.synchronize()
    for gi, (ctx, (shape, M)) in enumerate(zip(ctxs, GRIDS)):
        _assert_same(_fields(ctx, M), alone[gi], "grid %s in the group" % (shape,))
    grp.close()
    for ctx in ctxs:
        ctx.close()


def test_group_rejects_mixed_precision(defaults):
    _, c, k = defaults
    a = _ctx(GRIDS[1][0], 1, abi.PREC_FP32, c, k, 1)
    b = _ctx(GRIDS[1][0], 1, abi.PREC_FP64, c, k, 1)
    with pytest.raises(atsat_amd.AtsatError):
        atsat_amd.Group([a, b])
    with pytest.raises(atsat_amd.AtsatError):
        atsat_amd.Group([a, a])
    a.close()
    b.close()
