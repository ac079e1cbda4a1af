"""Parity of the device turbulence closure (atsat_turbulence / atsat_turbulence_init, and the
step with turb_active and the closure running after restoreVar) against the FP64 CPU oracle.

Tolerances (relative L2 per field against the oracle):
  * the closure alone, FP64: <= 1e-14 (the production tensor, nu_t and the sources are the
    reference's operations in its order; the blending functions F1 / F2 use the device tanh,
    but in the free-flow form they are below 1e-10 and leave the sources unchanged);
  * whole iterations with the closure, FP64: the bound of tests/util.py oracle_bound (1e-12 for
    the prognostic fields; the oracle's own one-ulp response for the diagnostics);
  * FP32 work precision, 10 native iterations against the FP64 oracle: per field at most 3x the
    value measured on the same case (tests/util.py fp32_check).
"""
import sys

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import (fp32_check, rel_l2, stress_state, stress_knobs, HOT_FIELDS, oracle_bound,
                  ulp_perturbed)

pytestmark = pytest.mark.gpu

TURB_CMP = abi.TURB_OUT + ["vel_star"]
STEP_CMP = HOT_FIELDS + abi.TURB_OUT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small():
    return atsat_amd.small_grid(21, 61, 121)


def _knobs(k):
    kt = abi.Knobs.from_buffer_copy(k)
    kt.turb_active = 1
    kt.turb_closure = 1
    return kt


def _state(scenario, g, c):
    if scenario == "stress":
        return stress_state(g, c)
    st = atsat_amd.gen_v1(g, c)
    return st


@pytest.mark.parametrize("call", ["turbulence", "turbulence_init"])
@pytest.mark.parametrize("scenario", ["gen_v1", "stress"])
def test_closure_fp64_matches_oracle(defaults, scenario, call):
    _, c, k = defaults
    g = _small()
    kt = _knobs(k)
    st = _state(scenario, g, c)
    say("case closure", scenario, call)
    so = st.copy()
    oracle.Oracle(g, c, kt).call(call, so)
    ctx = atsat_amd.Context(g, c, kt, 1)
    ctx.upload_state(st)
    ctx.call(call)
    ctx.synchronize()
    sg = st.copy()
    ctx.download_state(sg)
    ctx.close()
    exact = [n for n in TURB_CMP if np.array_equal(sg[n], so[n])]
    say("  bit-identical:", exact)
    for n in TURB_CMP:
        e = rel_l2(sg[n], so[n])
        assert e <= 1e-14, (scenario, call, n, e)


def test_closure_needs_turb_active(defaults):
    _, c, k = defaults
    g = _small()
    ctx = atsat_amd.Context(g, c, k, 1)
    ctx.upload_state(atsat_amd.gen_v1(g, c))
    with pytest.raises(atsat_amd.AtsatError):
        ctx.turbulence()
    ctx.close()


def _run(g, c, k, st, first, n, prec=abi.PREC_FP64, use_graph=False, members=1):
    ctx = atsat_amd.Context(g, c, k, members, prec)
    for m in range(members):
        ctx.upload_state(st[m] if isinstance(st, list) else st, member=m)
    ctx.run(first, n, use_graph)
    ctx.synchronize()
    outs = []
    for m in range(members):
        s = (st[m] if isinstance(st, list) else st).copy()
        ctx.download_state(s, member=m)
        outs.append(s)
    ctx.close()
    return outs


def _oracle_run(g, c, k, st, first, n):
    so = st.copy()
    O = oracle.Oracle(g, c, k)
    for it in range(first, first + n):
        O.step(so, it)
    return so


def _check_steps(g, c, k, st, first, n, what):
    so = _oracle_run(g, c, k, st, first, n)
    sp = _oracle_run(g, c, k, ulp_perturbed(st), first, n)
    sg = _run(g, c, k, st, first, n)[0]
    bad = []
    for f in STEP_CMP:
        e = rel_l2(sg[f], so[f])
        tol = oracle_bound(f, so, sp)
        if e > 1e-14:
            say("  %-18s gpu %.2e   bound %.2e" % (f, e, tol))
        if not (e <= tol):
            bad.append((f, e, tol))
    assert not bad, (what, bad)


def test_steps_with_closure_small_stress(defaults):
    _, c, k = defaults
    g = _small()
    ks = stress_knobs(k)
    ks.turb_closure = 1
    st = stress_state(g, c)
    say("case small stress 6 iterations, closure on")
    _check_steps(g, c, ks, st, 1, 6, "small stress with the closure")


def test_steps_with_closure_native_10(defaults):
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c)
    say("case native gen_v1 10 iterations fp64, closure on")
    _check_steps(g, c, _knobs(k), st, 1, 10, "native gen_v1 with the closure")


def test_graph_replay_with_closure(defaults):
    _, c, k = defaults
    g = _small()
    kt = _knobs(k)
    st = atsat_amd.gen_v1(g, c)
    say("case graph replay small 7 iterations, closure on")
    a = _run(g, c, kt, st, 1, 7, use_graph=False)[0]
    b = _run(g, c, kt, st, 1, 7, use_graph=True)[0]
    for n in STEP_CMP:
        assert np.array_equal(a[n], b[n]), n


def test_ensemble_with_closure_matches_single(defaults):
    _, c, k = defaults
    g = _small()
    kt = _knobs(k)
    sts = [atsat_amd.gen_v1(g, c, seed=5, member=m, t_sigma=0.5, w_sigma=2.0 / 470.0) for m in range(3)]
    say("case ensemble 3 members small 4 iterations, closure on")
    ens = _run(g, c, kt, sts, 1, 4, members=3)
    for m in range(3):
        one = _run(g, c, kt, sts[m], 1, 4)[0]
        for n in STEP_CMP:
            assert np.array_equal(ens[m][n], one[n]), (m, n)


def test_fp32_step_with_closure_error_native(defaults):
    g, c, k = defaults
    kt = _knobs(k)
    st = atsat_amd.gen_v1(g, c)
    say("case native gen_v1 10 iterations fp32, closure on")
    so = _oracle_run(g, c, kt, st, 1, 10)
    sg = _run(g, c, kt, st, 1, 10, prec=abi.PREC_FP32)[0]
    errs = {n: rel_l2(sg[n], so[n]) for n in abi.PROG_NAMES + ["p_dyn", "p_stat", "nue", "tke_source"]}
    say("fp32 per-field rel L2:", {n: "%.2e" % e for n, e in errs.items()})
    # measured on this case (one MI355X); v starts at zero and its tendency is a small residual
    # of balanced O(1) terms, h2o_cloud forms from a small supersaturation: both amplify the FP32
    # rounding of the stage inputs
    MEASURED = {"t": 3.12e-09, "u": 6.96e-07, "v": 4.55e-05, "w": 2.98e-10, "h2o": 9.35e-08,
                "h2o_cloud": 9.79e-05, "h2o_ice": 2.96e-07, "ch4": 0.0, "ch4_cloud": 0.0,
                "ch4_ice": 0.0, "h2s": 0.0, "nh3": 8.93e-08, "nh3_cloud": 0.0, "nh3_ice": 3.48e-06,
                "nh4sh": 0.0, "tke": 3.47e-08, "dis": 2.80e-07, "p_dyn": 7.96e-13,
                "p_stat": 1.76e-08, "nue": 1.80e-08, "tke_source": 9.27e-06}
    bad = fp32_check(errs, MEASURED)
    assert not bad, bad


def test_drop_in_sequence_keeps_n_copies(defaults):
    """The reference's call order in drop-in mode -- restoreVar, the closure, then RungeKuttaSat
    alone -- integrates tke / dis from their copies taken before the closure."""
    _, c, k = defaults
    g = _small()
    kt = _knobs(k)
    st = stress_state(g, c)
    so = st.copy()
    O = oracle.Oracle(g, c, kt)
    for name in ("restore_var", "turbulence", "rk4_step"):
        O.call(name, so)
    ctx = atsat_amd.Context(g, c, kt, 1)
    ctx.upload_state(st)
    for name in ("restore_var", "turbulence", "rk4_step"):
        ctx.call(name)
    ctx.synchronize()
    sg = st.copy()
    ctx.download_state(sg)
    ctx.close()
    for n in ["tke", "dis", "tken", "disn", "t", "u", "v", "w"]:
        assert rel_l2(sg[n], so[n]) <= 1e-12, (n, rel_l2(sg[n], so[n]))
