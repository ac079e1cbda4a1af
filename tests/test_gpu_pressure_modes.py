"""Parity of the device's optional modes of the red-black pressure solver (the line pass, the
barotropic correction, both; k_pmodes.hip) against the FP64 CPU oracle, which is pinned to the
reference's own solver (tests/test_oracle_pressure_modes.py).

Tolerances: the pressure solve alone, FP64: bit-identical (every sum in the oracle's order, the
DFT tables and the sines from the host libm); FP32 context (the velocity tendencies the source is
formed from held in FP32): 1e-6 relative; whole iterations with a mode on, FP64: the bound of
tests/test_gpu_parity.py; a captured graph: bit-identical to the stream."""
import sys

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import (rel_l2, stress_state, stress_knobs, HOT_FIELDS, oracle_bound, ulp_perturbed)

pytestmark = pytest.mark.gpu
MODES = [(0, 1), (1, 0), (1, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _knobs(k, line, baro, sweeps=3):
    ks = stress_knobs(k)
    ks.press_line = line
    ks.press_baro = baro
    ks.press_sweeps = sweeps
    return ks


@pytest.mark.parametrize("line,baro", MODES)
@pytest.mark.parametrize("native", [False, True])
def test_solve_fp64_bit_identical(defaults, line, baro, native):
    g0, c, k = defaults
    g = g0 if native else atsat_amd.small_grid(21, 61, 121)
    ks = _knobs(k, line, baro)
    st = stress_state(g, c)
    say("case pressure modes line", line, "baro", baro, "native" if native else "small")
    so = st.copy()
    oracle.Oracle(g, c, ks).call("pressure_solve", so)
    ctx = atsat_amd.Context(g, c, ks, 2)
    ctx.upload_state(stress_state(g, c, seed=3), member=0)
    ctx.upload_state(st, member=1)
    ctx.call("pressure_solve")
    ctx.synchronize()
    sg = st.copy()
    ctx.download_state(sg, member=1)
    ctx.close()
    say("  rel L2", rel_l2(sg["p_dyn"], so["p_dyn"]))
    assert np.array_equal(sg["p_dyn"], so["p_dyn"])


@pytest.mark.parametrize("line,baro", MODES)
def test_solve_fp32_context(defaults, line, baro):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    ks = _knobs(k, line, baro)
    st = stress_state(g, c)
    so = st.copy()
    oracle.Oracle(g, c, ks).call("pressure_solve", so)
    ctx = atsat_amd.Context(g, c, ks, 1, abi.PREC_FP32)
    ctx.upload_state(st)
    ctx.call("pressure_solve")
    sg = st.copy()
    ctx.download_state(sg)
    ctx.close()
    e = rel_l2(sg["p_dyn"], so["p_dyn"])
    say("case fp32 pressure modes", line, baro, "rel L2 %.2e" % e)
    assert e <= 1e-6


def _run(g, c, k, st, first, n, use_graph=False):
    ctx = atsat_amd.Context(g, c, k, 1)
    ctx.upload_state(st)
    ctx.run(first, n, use_graph)
    ctx.synchronize()
    s = st.copy()
    ctx.download_state(s)
    ctx.close()
    return s


def test_steps_with_modes(defaults):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    ks = _knobs(k, 1, 1, sweeps=2)
    st = stress_state(g, c)
    say("case small stress 4 iterations, line pass and barotropic correction")
    O = oracle.Oracle(g, c, ks)
    so = st.copy()
    for it in range(1, 5):
        O.step(so, it)
    sp = ulp_perturbed(st)
    for it in range(1, 5):
        O.step(sp, it)
    sg = _run(g, c, ks, st, 1, 4)
    bad = []
    for f in HOT_FIELDS:
        e = rel_l2(sg[f], so[f])
        tol = oracle_bound(f, so, sp)
        if not (e <= tol):
            bad.append((f, e, tol))
    assert not bad, bad
    b = _run(g, c, ks, st, 1, 6, use_graph=True)
    a = _run(g, c, ks, st, 1, 6)
    for f in HOT_FIELDS:
        assert np.array_equal(a[f], b[f]), f


def test_line_pass_rejects_tall_columns(defaults):
    _, c, k = defaults
    g = atsat_amd.small_grid(70, 61, 121, dr=1.0 / 64.0)
    with pytest.raises(atsat_amd.AtsatError, match="64 levels"):
        atsat_amd.Context(g, c, _knobs(k, 1, 0), 1)
