"""The budget sums on the device (atsat_budgets): the H2O, NH3 and CH4 columns in g/m2 and the
level-mean temperatures of every member, the sums of the reference's ATSAT_CWB_DIAG and
ATSAT_TB_DIAG instruments (planet/cSaturnModel.cpp:656-666, 776-787), against the oracle's
restatement of them on each member's downloaded state. The device sums in a fixed tree, the oracle
(like the reference) in loop order: the two agree to FP64 rounding, bounded at 1e-13 relative. The
oracle's sums are pinned by closed forms in test_oracle_invariants.py."""
import sys

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import stress_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(g, c, k, ctx, members, sts):
    cols, lev = ctx.budgets()
    assert cols.shape == (members, 3) and lev.shape == (members, g.im)
    O = oracle.Oracle(g, c, k)
    for m in range(members):
        st = sts[m].copy()
        ctx.download_state(st, member=m)
        oc, ol = O.budgets(st)
        say("  member %d columns %s oracle %s" % (m, cols[m], oc))
        assert np.all(np.abs(cols[m] - oc) <= 1e-13 * np.abs(oc)), (cols[m], oc)
        assert np.all(np.abs(lev[m] - ol) <= 1e-13 * np.abs(ol)), np.max(np.abs(lev[m] / ol - 1.0))


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32])
def test_budgets_match_oracle_small(defaults, prec):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    sts = [stress_state(g, c, seed=s) for s in (3, 4)]
    say("case budgets small grid, 2 members, precision", prec)
    ctx = atsat_amd.Context(g, c, k, 2, prec)
    for m, st in enumerate(sts):
        ctx.upload_state(st, member=m)
    ctx.run(1, 2, False)
    _check(g, c, k, ctx, 2, sts)
    ctx.close()


def test_budgets_match_oracle_native():
    g, c, k = atsat_amd.defaults("model")
    st = atsat_amd.gen_v1(g, c)
    say("case budgets native grid, 1 member, fp32, model physics, 2 iterations")
    ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP32)
    ctx.upload_state(st)
    ctx.run(1, 2, False)
    _check(g, c, k, ctx, 1, [st])
    cols, lev = ctx.budgets()
    assert np.all(cols[0, :2] > 0.0) and np.all(lev[0] > 0.0)
    ctx.close()
