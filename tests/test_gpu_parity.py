"""Parity of the HIP path (libatsat_gpu.so, through the C ABI) against the FP64 CPU oracle.

Tolerances (relative L2 per field against the oracle):
  * FP64, functions without transcendentals (density, pressure sweeps, p_dyn filter, boundary
    conditions, clamp, copies): bit-identical -- the kernels perform the reference's operations
    in the reference's order with -ffp-contract=off;
  * FP64 RK4 step: <= 1e-12 (the level mean of the buoyancy is a parallel reduction, and u
    starts at 0, so the anomaly it is driven by amplifies last-bit differences);
  * FP64 saturation adjustment, p_stat rebuild, precipitation: <= 1e-12 (device exp / log /
    pow against glibc's, last-bit differences);
  * FP64 whole iterations: <= 1e-12 for every prognostic field (u, v, w, t, the species, tke,
    dis, p_dyn, p_stat, p_metric, rho_mix). The device's exp / log / pow are correctly rounded;
    glibc's, which the reference calls, differ from them by one ulp in ~0.06% of arguments. The
    precipitation diagnostics and the last-stage tendencies amplify such a bit (a rain flux that
    evaporation cancels to a rounding residue decides the next level's branches), so they are
    bounded by max(1e-12, 10 x the oracle's own response to a one-ulp change of p_stat at the
    start) (tests/util.py oracle_bound);
  * FP32 work precision, 10 native iterations against the FP64 oracle: per field at most 3x the
    value measured on the same case (tests/util.py fp32_check).
"""
import sys

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import (fp32_check, rel_l2, stress_state, stress_knobs, cold_base_state, HOT_FIELDS,
                  oracle_bound, ulp_perturbed)

pytestmark = pytest.mark.gpu

EXACT_FUNCS = ["mixture_density", "pressure_solve", "damp_pdyn", "rk_scalar_sync",
               "boundary_conditions", "clamp_species", "restore_var"]
TRANSC_FUNCS = [("saturation_adjust", 0), ("saturation_adjust", 1), ("saturation_adjust", 2),
                ("rebuild_pstat", None), ("precipitation", None)]
COMPARE = HOT_FIELDS + ["tke", "dis", "tn", "un", "vn", "wn", "h2on", "nh3n", "aux_u", "aux_v", "aux_w"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_call(ctx, name, arg):
    if arg is None:
        ctx.call(name)
    else:
        ctx.call(name, arg)
    ctx.synchronize()


def _run_both(g, c, k, st, name, arg, prec=abi.PREC_FP64):
    so = st.copy()
    O = oracle.Oracle(g, c, k)
    if arg is None:
        O.call(name, so)
    else:
        O.call(name, so, arg)
    ctx = atsat_amd.Context(g, c, k, 1, prec)
    ctx.upload_state(st)
    _gpu_call(ctx, name, arg)
    sg = st.copy()
    ctx.download_state(sg)
    ctx.close()
    return so, sg


def _small():
    return atsat_amd.small_grid(21, 61, 121)


@pytest.mark.parametrize("scenario", ["gen_v1", "stress"])
def test_functions_fp64_small(defaults, scenario):
    _, c, k = defaults
    g = _small()
    if scenario == "stress":
        k = stress_knobs(k)
        st = stress_state(g, c)
    else:
        st = atsat_amd.gen_v1(g, c)
    for name in EXACT_FUNCS + ["rk4_step"]:
        say("case", scenario, name)
        so, sg = _run_both(g, c, k, st, name, None)
        for n in COMPARE:
            if name == "rk4_step":
                assert rel_l2(sg[n], so[n]) <= 1e-12, (name, n, rel_l2(sg[n], so[n]))
            else:
                assert np.array_equal(sg[n], so[n]), (name, n, rel_l2(sg[n], so[n]))
    for name, arg in TRANSC_FUNCS:
        say("case", scenario, name, arg)
        so, sg = _run_both(g, c, k, st, name, arg)
        for n in COMPARE:
            assert rel_l2(sg[n], so[n]) <= 1e-12, (name, arg, n, rel_l2(sg[n], so[n]))


def _steps(g, c, k, st, first, n, prec=abi.PREC_FP64, use_graph=False, members=1):
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
    sg = _steps(g, c, k, st, first, n)[0]
    bad = []
    for f in COMPARE:
        e = rel_l2(sg[f], so[f])
        tol = oracle_bound(f, so, sp)
        if e > 1e-14:
            say("  %-18s gpu %.2e   bound %.2e" % (f, e, tol))
        if not (e <= tol):
            bad.append((f, e, tol))
    assert not bad, (what, bad)


def test_steps_fp64_small_stress(defaults):
    _, c, k = defaults
    g = _small()
    k = stress_knobs(k)
    st = stress_state(g, c)
    say("case small stress 6 iterations")
    _check_steps(g, c, k, st, 1, 6, "small stress")


def test_steps_fp64_native_10(defaults):
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c)
    say("case native gen_v1 10 iterations fp64")
    _check_steps(g, c, k, st, 1, 10, "native gen_v1")


def test_graph_replay_matches_stream(defaults):
    _, c, k = defaults
    g = _small()
    st = atsat_amd.gen_v1(g, c)
    say("case graph replay small 7 iterations")
    a = _steps(g, c, k, st, 1, 7, use_graph=False)[0]
    b = _steps(g, c, k, st, 1, 7, use_graph=True)[0]
    for n in COMPARE:
        assert np.array_equal(a[n], b[n]), n


def test_fp32_run_keeps_static_pressure_terms():
    """The FP32 step's stages read pressure terms formed once per RK4 step. Inside atsat_run the
    odd iteration keeps the static ones (from p_dyn, p_stat and rho_mix) formed after the even
    iteration's physics, on the stream and in the captured graph; atsat_step called on its own forms
    them again every time. Model physics, red-black sweeps (the same work in every path): the
    three give the same bits."""
    g = _small()
    _, c, k = atsat_amd.defaults("model")
    k.press_mode = 0
    st = atsat_amd.gen_v1(g, c)
    say("case fp32 model physics small iterations 1-6: run, run with graph, step by step")
    a = _steps(g, c, k, st, 1, 6, prec=abi.PREC_FP32)[0]
    b = _steps(g, c, k, st, 1, 6, prec=abi.PREC_FP32, use_graph=True)[0]
    ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP32)
    ctx.upload_state(st)
    for it in range(1, 7):
        ctx.step(it)
    ctx.synchronize()
    s = st.copy()
    ctx.download_state(s)
    ctx.close()
    for n in COMPARE:
        assert np.array_equal(a[n], s[n]), n
        assert np.array_equal(b[n], s[n]), n


def test_ensemble_members_match_single(defaults):
    _, c, k = defaults
    g = _small()
    sts = [atsat_amd.gen_v1(g, c, seed=5, member=m, t_sigma=0.5, w_sigma=2.0 / 470.0) for m in range(3)]
    say("case ensemble 3 members small 4 iterations")
    ens = _steps(g, c, k, sts, 1, 4, members=3)
    for m in range(3):
        one = _steps(g, c, k, sts[m], 1, 4)[0]
        for n in COMPARE:
            assert np.array_equal(ens[m][n], one[n]), (m, n)


def test_fp32_step_error_native(defaults):
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c)
    say("case native gen_v1 10 iterations fp32")
    so = st.copy()
    O = oracle.Oracle(g, c, k)
    for it in range(1, 11):
        O.step(so, it)
    sg = _steps(g, c, k, st, 1, 10, prec=abi.PREC_FP32)[0]
    errs = {n: rel_l2(sg[n], so[n]) for n in abi.PROG_NAMES[:15] + ["p_dyn", "p_stat"]}
    say("fp32 per-field rel L2:", {n: "%.2e" % e for n, e in errs.items()})
    # measured on this case (one MI355X); v starts at zero and its tendency is a small residual
    # of balanced O(1) terms, h2o_cloud forms from a small supersaturation: both amplify the FP32
    # rounding of the stage inputs
    MEASURED = {"t": 7.04e-09, "u": 7.10e-07, "v": 1.72e-03, "w": 1.59e-11, "h2o": 1.72e-07,
                "h2o_cloud": 1.45e-07, "h2o_ice": 1.03e-05, "ch4": 0.0, "ch4_cloud": 0.0,
                "ch4_ice": 0.0, "h2s": 0.0, "nh3": 1.81e-07, "nh3_cloud": 0.0, "nh3_ice": 8.55e-06,
                "nh4sh": 0.0, "p_dyn": 8.75e-13, "p_stat": 1.51e-09}
    bad = fp32_check(errs, MEASURED)
    assert not bad, bad


def test_rows_api_roundtrip(defaults):
    import ctypes as C
    _, c, k = defaults
    g = _small()
    st = atsat_amd.gen_v1(g, c)
    ctx = atsat_amd.Context(g, c, k, 1)
    a = np.ascontiguousarray(st["t"])
    rows = (C.POINTER(C.c_double) * (g.im * g.jm))()
    for r in range(g.im * g.jm):
        rows[r] = C.cast(a.ctypes.data + r * g.km * 8, C.POINTER(C.c_double))
    assert atsat_amd.lib().atsat_upload_rows(ctx.h, 0, abi.FIELD["w"], rows) == 0
    assert np.array_equal(ctx.download("w"), a)
    out = np.zeros_like(a)
    for r in range(g.im * g.jm):
        rows[r] = C.cast(out.ctypes.data + r * g.km * 8, C.POINTER(C.c_double))
    assert atsat_amd.lib().atsat_download_rows(ctx.h, 0, abi.FIELD["w"], rows) == 0
    assert np.array_equal(out, a)
    ctx.close()


def test_census_counts_injected_nan(defaults):
    _, c, k = defaults
    g = _small()
    st = atsat_amd.gen_v1(g, c)
    st["t"][3, 4, 5] = np.nan
    st["h2o"][1, 2, 3] = np.inf
    ctx = atsat_amd.Context(g, c, k, 1)
    ctx.upload_state(st)
    assert ctx.nonfinite_count() == 2
    ctx.close()


def test_level_mean_skipped_cells_exact(defaults):
    """computeBuoyancyRefLevel skips cells with no positive temperature (NaN included) and cells
    whose term is not finite (RungeKutta_Sat_Turb.cpp:39-44); the device level mean must skip the
    same cells and still sum in the reference's order."""
    _, c, k = defaults
    g = _small()
    st = atsat_amd.gen_v1(g, c)
    st["t"][3, 0, 7] = 0.0
    st["t"][3, 40, 100] = -1.0
    st["t"][5, 17, 3] = np.nan
    st["p_stat"][8, 22, 9] = np.inf
    st["t"][g.im - 1, g.jm - 1, g.km - 1] = -2.0
    say("case level mean with skipped cells, small grid")
    so = st.copy()
    oracle.Oracle(g, c, k).call("buoyancy_ref_level", so)
    ctx = atsat_amd.Context(g, c, k, 1)
    ctx.upload_state(st)
    ctx.rk4_step()
    ctx.synchronize()
    sg = st.copy()
    ctx.download_state(sg)
    ctx.close()
    a, b = sg.v["buoy_ref_level"], so.v["buoy_ref_level"]
    assert np.array_equal(a, b), np.nonzero(a != b)


SRF = ["precip_srf_h2o", "precip_srf_nh3", "precip_srf_ch4", "precip_srf_nh4sh", "precip_srf_total"]


def _flip_columns(sg, so, jm, km):
    """Columns where a species' fluxes differ by more than 1e-12 of the column's largest flux, with
    for each the top differing level and whether either result holds a cancellation residue -- a
    flux 0 < F < 1e-12 of that maximum -- at that level or one above it."""
    out = []
    for names in (("P_rain", "P_snow", "P_graupel"), ("P_nh3_rain", "P_nh3_snow", "P_nh3_graupel"),
                  ("P_ch4_rain", "P_ch4_snow", "P_ch4_graupel")):
        G = np.stack([sg[n] for n in names])                  # [3][im][jm][km]
        O = np.stack([so[n] for n in names])
        mx = np.maximum(np.abs(O).max(axis=(0, 1)), 1e-300)     # [jm][km]
        rel = np.sqrt(((G - O) ** 2).sum(axis=(0, 1))) / mx
        for j, kk in zip(*np.nonzero(rel > 1e-12)):
            lev = np.nonzero(np.any(G[:, :, j, kk] != O[:, :, j, kk], axis=0))[0].max()
            win = slice(lev, min(lev + 2, G.shape[1]))
            tiny = lambda a: bool(np.any((a[:, win, j, kk] > 0.0) & (a[:, win, j, kk] < 1e-12 * mx[j, kk])))
            out.append((int(j), int(kk), int(lev), tiny(G) or tiny(O)))
    return out


@pytest.mark.parametrize("grid", ["small", "native"])
def test_precipitation_surface_maps(defaults, grid):
    """The five surface maps of Precipitation::surfaceMap (Precipitation.h:467-536) and the
    precipitation fields against the oracle, FP64, on the cold-base state (every species' flux
    reaches the base).

    Rain^(7/9) and Rain^(4/9) come from the device's correctly rounded pow, the oracle's (the
    reference's) from glibc, which differs by one ulp in ~0.06% of arguments. Where evaporation
    cancels a level's rain flux to a rounding residue (~1e-25 of 5e-9), that ulp moves the residue
    and the branches below it. Every column must therefore agree to 1e-12 of its largest flux, or
    differ only below such a residue, in at most 1e-3 of the columns; the maps of the agreeing
    columns to 1e-12, and the maps must be the sums of the device's own base-level fluxes."""
    g0, c, k = defaults
    g = _small() if grid == "small" else g0
    k = stress_knobs(k)
    st = cold_base_state(g, c)
    say("case surface maps, cold base,", grid)
    so, sg = _run_both(g, c, k, st, "precipitation", None)
    flips = _flip_columns(sg, so, g.jm, g.km)
    say("  columns differing below a cancellation residue:", len(flips))
    assert all(f[3] for f in flips), [f for f in flips if not f[3]][:5]
    assert len(flips) <= 1e-3 * g.jm * g.km, len(flips)
    keep = np.ones((g.jm, g.km), bool)
    for j, kk, _, _ in flips:
        keep[j, kk] = False
    base = {"h2o": ("P_rain", "P_snow", "P_graupel"), "nh3": ("P_nh3_rain", "P_nh3_snow", "P_nh3_graupel"),
            "ch4": ("P_ch4_rain", "P_ch4_snow", "P_ch4_graupel"), "nh4sh": ("P_nh4sh",)}
    for sp_, names in base.items():
        n = "precip_srf_" + sp_
        assert np.linalg.norm(so.v[n]) > 0.0, n
        a, b = sg.v[n].reshape(g.jm, g.km), so.v[n].reshape(g.jm, g.km)
        e = rel_l2(a[keep], b[keep])
        assert e <= 1e-12, (n, e)
        m = np.zeros((g.jm, g.km))
        for f in names:
            m = m + sg[f][0]
        assert np.array_equal(a, m), n
    for n in abi.PRECIP_OUT:
        e = rel_l2(sg[n][:, keep], so[n][:, keep])
        assert e <= 1e-12, (n, e)


def test_saturation_adjust_entry_margin(defaults):
    """The saturation adjustment's early end for a cell without condensate (SaturationAdjustment.h:
    414-448): the device decides `vapour <= its saturation value` with the library's log / exp and a
    1e-9 margin, and falls back to the correctly rounded values inside it. Vapour set at
    q_entry * (1 + d) for d in +-1e-10, +-1e-9, +-1e-8, +-1e-7 (and +1e-2, which condenses), in the
    liquid phase and (the cold upper levels, T < t_0) the ice phase: the cells the device changes are exactly the oracle's,
    and every field agrees to 1e-12 (FP64)."""
    _, c, k = defaults
    g = _small()
    st = atsat_amd.gen_v1(g, c)
    gas = abi.GAS_H2O
    T = st["t"] * c.t_ref
    rho = np.where((st["rho_mix"] > 0.0) & np.isfinite(st["rho_mix"]), st["rho_mix"], c.r_mix)
    p = st["p_stat"]
    liq = ~(T < c.t_0[gas])
    C0 = np.where(liq, c.C[gas], c.C_i[gas])
    L0 = np.where(liq, c.L0[gas], c.L0_i[gas])
    da = np.where(liq, c.del_alf[gas], c.del_alf_i[gas])
    db = np.where(liq, c.del_bet[gas], c.del_bet_i[gas])
    q_entry = rho * c.ep[gas] * np.exp(C0 + (-L0 / T + da * np.log(T) + db * T) / (1e-3 * c.R[gas])) / p
    rng = np.random.default_rng(3)
    # +1e-2: supersaturated enough for the Tao iteration to condense (its tolerance is 1e-3)
    d = rng.choice(np.array([-1e-7, -1e-8, -1e-9, -1e-10, 1e-10, 1e-9, 1e-8, 1e-7, 1e-2]), size=T.shape)
    st["h2o"][...] = q_entry * (1.0 + d)
    st["h2o_cloud"][...] = 0.0
    st["h2o_ice"][...] = 0.0
    say("case saturation adjustment at the entry margin: liquid cells %d, ice cells %d" % (liq.sum(), (~liq).sum()))
    assert liq.sum() > 0 and (~liq).sum() > 0
    so, sg = _run_both(g, c, k, st, "saturation_adjust", gas)
    for n in ("t", "h2o", "h2o_cloud", "h2o_ice"):
        ch_o = so[n] != st[n]
        ch_g = sg[n] != st[n]
        assert np.array_equal(ch_o, ch_g), (n, int((ch_o != ch_g).sum()))
        assert rel_l2(sg[n], so[n]) <= 1e-12, (n, rel_l2(sg[n], so[n]))
    assert (so["h2o"] != st["h2o"]).sum() > 0
