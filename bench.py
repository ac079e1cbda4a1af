"""Throughput of the ATSAT model time step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--members M] [--prec fp32|fp64] [--press conv|ref]
                    [--physics model|resident] [--no-graph] [--tile B R] [--other-press] [--no-cpu]
                    [--out FILE]
    python bench.py --accuracy ITERS [--members M] [--oracle]
    python bench.py --poisson
    python bench.py --batch [--press conv|ref] [--warmup W]
    python bench.py --roofline-table TRACE FETCH WRITE CAL_FETCH CAL_WRITE [--prec fp32|fp64]
    python bench.py --hbm-probe

Workload: the 16-member ensemble of the native 41 x 181 x 361 shell, each member the gen_v1 start
state with its own N(0, 0.5 K) temperature and N(0, 2 m/s) zonal-wind perturbation, advanced
together; one step is one model iteration (odd and even iterations alternate, the even ones run
the physics block, whose pressure solve is converged to a relative residual of 1e-10 by FP64
defect correction). --physics model (the default) runs the model's default physics profile: the
k-omega SST closure and the moist convective adjustment after every step and the chemistry block
on even iterations, radiation off; --physics resident switches the closure, the adjustment and
the chemistry off. The timed region is K = 100 iterations (by default) replayed from the
hipGraph of an even + odd iteration pair, captured and replayed once before the clock starts.
The headline value is member-grid-point-steps/s of the FP32 step (FP32 stencils, FP64 state and
refinement); the plain FP64 step is measured in the same run, on the same start states.

The JSON line carries:
  roofline      the fused RHS + RK4 stage kernel: algorithmic bytes per launch over its mean
                launch time, measured with HIP events on the context's stream over a profiled
                pass of two iterations launched on the stream after the tile is tuned, against the
                8 TB/s HBM peak; traffic = the HBM bytes per launch the FETCH_SIZE / WRITE_SIZE
                counters measured (profiles/step_kernels_roofline_*_ensemble16.md);
                streaming_ceiling = the HBM rate this device reaches streaming ~100 arrays at
                once, as a stage does (atsat_amd/hbm_probe, profiles/device_streams_bandwidth.md);
  cpu_baseline  the reference's own step (oracle/_ref, when it was built and shipped) or the
                FP64 CPU restatement (oracle/_build), one member, on this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def apply_physics(k, mode):
    """The physics profile of the step: 'model' = the model's defaults, the k-omega SST closure
    (turb_model k_omega_SST) and the moist convective adjustment (ATSAT_MCA 1) after every step
    and the chemistry block (chemical_reaction 1) on even iterations, radiation off; 'resident' =
    the closure, the adjustment and the chemistry off."""
    full = mode == "model"
    k.turb_active = 1 if full else 0
    k.turb_closure = 1
    k.chem_block = 1 if full else 0
    k.chemical_reaction = 1.0 if full else 0.0
    k.mca_stride = 1 if full else 0
    return k


def rk_bytes_per_cell(elem, turb=False, chem=False):
    """Algorithmic bytes of the four RK4 stage kernels per integrated cell and step, each value
    read or written once per launch. Of the live state and y_n (aliased to it; tken / disn on their
    own with the closure), t and w are FP64 and the other 13 (15 with the closure) prognostic
    fields are in the work precision `elem`, like the stage inputs X / Y, the accumulators, the
    precipitation sources, the closure and chemistry inputs, the six pressure terms formed once per
    step (k_press_terms) and rhs/aux; p_stat is FP64. Per stage: the 15 (17 with the closure)
    stencil fields, the pressure terms, p_stat and Q_precip + 9 S_precip (+ nue, tke_source, dis_source; +
    massflux_* and fluxlim_nh4sh) are read; stages 1-3 also read y_n and acc; stages 0-2 write acc
    and the next stage input, stage 3 the new state and rhs/aux u, v, w."""
    return sum(rk_stage_bytes(elem, turb, chem))


def rk_stage_bytes(elem, turb=False, chem=False):
    """The algorithmic bytes per integrated cell of each of the four RK4 stages (rk_bytes_per_cell)."""
    f = 17 if turb else 15
    state = 2 * 8 + (f - 2) * elem          # one value of every prognostic field of the state
    common = 8 + (6 + 10 + (3 if turb else 0) + (4 if chem else 0)) * elem
    s0 = state + common + 2 * f * elem
    s12 = f * elem + common + state + f * elem + 2 * f * elem
    s3 = f * elem + common + state + f * elem + state + 6 * elem
    return [s0, s12, s12, s3]


def streaming_ceiling():
    """The HBM rate this device reaches streaming ~100 arrays at once (a stage kernel's pattern):
    the lowest of the committed hbm_probe rows with 48 or more arrays, in GB/s, or None."""
    path = os.path.join(HERE, "profiles", "device_streams_bandwidth.md")
    vals = []
    try:
        with open(path) as f:
            for line in f:
                c = [x.strip() for x in line.split("|")]
                if len(c) > 7 and c[1] == "groups" and int(c[3]) >= 48:
                    vals.append(float(c[6]))
    except (OSError, ValueError):
        return None
    return min(vals) if vals else None


def measured_traffic(elem, physics):
    """HBM bytes per launch of the RK stage kernel (mean over the four stages) from the committed
    per-kernel table of the same precision (model physics only), or None."""
    if physics != "model":
        return None
    path = os.path.join(HERE, "profiles", "step_kernels_roofline_%s_ensemble16.md" % ("fp32" if elem == 4 else "fp64"))
    vals = []
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("| `k_rk_cell<"):
                    c = [x.strip() for x in line.split("|")]
                    vals.append((float(c[5]) + float(c[6])) * 1e9)
    except (OSError, ValueError, IndexError):
        return None
    return int(sum(vals) / len(vals)) if len(vals) == 4 else None


def member_states(g, c, members):
    """The ensemble's start states: gen_v1 of each member with its own N(0, 0.5 K) temperature and
    N(0, 2 m/s) zonal-wind perturbation."""
    import atsat_amd
    return [atsat_amd.gen_v1(g, c, seed=20261203, member=m, t_sigma=0.5, w_sigma=2.0 / 470.0)
            for m in range(members)]


def run_gpu(g, c, k, states, prec, steps, warmup, press_mode, autotune=True, tile=None, graph=True,
            prof_steps=2):
    """One context of len(states) members: `warmup` eager iterations, the RK4 stage tile chosen by
    the autotuner (atsat_autotune, on the warmed-up state, which it leaves unchanged), a profiled
    eager pass of `prof_steps` iterations (HIP events around every kernel family, for kernel_ms and
    the roofline), then the timed region: `steps` iterations replayed from the hipGraph of an even
    + odd iteration pair (graph=True), captured and replayed once before the clock starts, with no
    event inside; graph=False times the same iterations launched on the stream. Returns a dict."""
    import atsat_amd
    members = len(states)
    k2 = type(k).from_buffer_copy(k)
    k2.press_mode = press_mode
    t0 = time.perf_counter()
    ctx = atsat_amd.Context(g, c, k2, members, prec)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in range(members):
        ctx.upload_state(states[m], member=m)
    up = time.perf_counter() - t0
    it = 1
    ctx.run(it, warmup, False)
    it += warmup
    ctx.synchronize()
    if tile is not None:
        ctx.set_tile(tile[0], tile[1])
        tile = {"band": tile[0], "rings": tile[1], "stage_ms": None}
    elif autotune:
        band, rings, ms = ctx.autotune()
        tile = {"band": band, "rings": rings, "stage_ms": round(ms, 4)}
    ctx.profile(True)
    t0 = time.perf_counter()
    ctx.run(it, prof_steps, False)
    ctx.synchronize()
    dt_prof = time.perf_counter() - t0
    it += prof_steps
    prof = ctx.profile_read()
    ctx.profile(False)
    solve = ctx.pressure_last() if press_mode == 1 else None
    if it % 2:                                      # the graph holds an even + odd pair
        ctx.run(it, 1, False)
        it += 1
    if graph:
        ctx.run(it, 2, True)                        # capture + first replay, untimed
        it += 2
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(it, steps, graph)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    it += steps
    out = {"dt": dt, "prof": prof, "prof_steps": prof_steps, "dt_prof": dt_prof, "solve": solve,
           "setup_s": setup_s, "upload_s": up / members, "tile": tile,
           "nonfinite": ctx.nonfinite_count(), "device_bytes": ctx.device_bytes()}
    ctx.close()
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, c, k, iters):
    """The reference's own step on the host (or the restatement), one member, `iters` iterations."""
    import atsat_amd
    import oracle
    st = atsat_amd.gen_v1(g, c, seed=20261203, member=0, t_sigma=0.5, w_sigma=2.0 / 470.0)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if oracle.ref_available():
        kind = "reference"
        R = oracle.Reference(k)
        R.load(st)
        t0 = time.perf_counter()
        for it in range(1, iters + 1):
            R.call("step", it)
        dt = time.perf_counter() - t0
        R.close()
    else:
        kind = "port"
        oracle.build(ref=False)
        O = oracle.Oracle(g, c, k)
        t0 = time.perf_counter()
        for it in range(1, iters + 1):
            O.step(st, it)
        dt = time.perf_counter() - t0
    npts = g.im * g.jm * g.km
    return {"value": npts * iters / dt, "unit": "member-grid-point-steps/s", "cores": cores,
            "kind": kind, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "1 member of the native 41x181x361 shell, iterations 1..%d (odd and even, "
                      "physics block included), %.1f s" % (iters, dt)}


ACC_FIELDS = ["t", "u", "v", "w", "h2o", "h2o_cloud", "h2o_ice", "ch4", "ch4_cloud", "ch4_ice", "h2s",
              "nh3", "nh3_cloud", "nh3_ice", "nh4sh", "tke", "dis", "p_dyn", "p_stat", "Q_precip",
              "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice", "S_precip_nh3", "S_precip_nh3_cloud",
              "S_precip_nh3_ice", "P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel",
              "P_nh4sh"]


def _errs(a, b):
    d = a - b
    nb = float(np.sqrt(np.sum(b * b)))
    rel = float(np.sqrt(np.sum(d * d))) / nb if nb > 0 else float(np.sqrt(np.sum(d * d)))
    return rel, float(np.max(np.abs(d))) if d.size else 0.0


def accuracy(a):
    """Per-field error of the FP32 step against the FP64 step (both on the GPU) after a.accuracy
    iterations of a.members members, and, with --oracle, of both against the FP64 CPU
    restatement for member 0."""
    import atsat_amd
    from atsat_amd import abi
    g, c, k = atsat_amd.defaults()
    apply_physics(k, a.physics)
    sts = [atsat_amd.gen_v1(g, c, seed=20261203, member=m, t_sigma=0.5 if a.members > 1 else 0.0,
                            w_sigma=(2.0 / 470.0) if a.members > 1 else 0.0) for m in range(a.members)]
    out = {}
    res = {}
    k = type(k).from_buffer_copy(k)
    k.press_mode = 1 if a.press == "conv" else 0
    for prec, name in [(abi.PREC_FP64, "fp64"), (abi.PREC_FP32, "fp32")]:
        ctx = atsat_amd.Context(g, c, k, a.members, prec)
        for m in range(a.members):
            ctx.upload_state(sts[m], member=m)
        t0 = time.perf_counter()
        ctx.run(1, a.accuracy, True)
        ctx.synchronize()
        out[name + "_s"] = time.perf_counter() - t0
        res[name] = [{f: ctx.download(f, member=m) for f in ACC_FIELDS} for m in range(a.members)]
        ctx.close()
    per_field = {}
    # cells where the two steps put a different amount of a gas into its condensed phases (the
    # saturation adjustment's thresholds decided the other way): the condensate differs by more
    # than 1e-4 of the vapour and the vapour by the opposite amount to 1%
    flips = []
    for m in range(a.members):
        A, Bv = res["fp64"][m], res["fp32"][m]
        fl = np.zeros(A["t"].shape, bool)
        for gas in ("h2o", "nh3", "ch4"):
            dc = (Bv[gas + "_cloud"] + Bv[gas + "_ice"]) - (A[gas + "_cloud"] + A[gas + "_ice"])
            dv = Bv[gas] - A[gas]
            fl |= (np.abs(dc) > 1e-4 * np.abs(A[gas])) & (np.abs(dv + dc) < 1e-2 * np.abs(dc))
        flips.append(fl)
    for f in ACC_FIELDS:
        e = [_errs(res["fp32"][m][f], res["fp64"][m][f]) for m in range(a.members)]
        ek = [_errs(res["fp32"][m][f][~flips[m]], res["fp64"][m][f][~flips[m]])[0] for m in range(a.members)]
        per_field[f] = {"rel_l2_max_over_members": max(x[0] for x in e), "max_abs": max(x[1] for x in e),
                        "rel_l2_without_phase_flips": max(ek)}
    out["fp32_vs_fp64_gpu"] = per_field
    out["phase_flip_cells"] = [[tuple(int(x) for x in c) for c in np.argwhere(fl)] for fl in flips]
    if a.oracle and k.press_mode != 0:
        out["oracle"] = "skipped: the CPU oracle restates the reference's red-black sweeps; use --press ref"
    elif a.oracle:
        import oracle
        oracle.build(ref=False)
        so = sts[0].copy()
        O = oracle.Oracle(g, c, k)
        t0 = time.perf_counter()
        for it in range(1, a.accuracy + 1):
            O.step(so, it)
        out["oracle_s"] = time.perf_counter() - t0
        out["fp64_gpu_vs_oracle"] = {f: _errs(res["fp64"][0][f], so[f])[0] for f in ACC_FIELDS}
        out["fp32_gpu_vs_oracle"] = {f: _errs(res["fp32"][0][f], so[f])[0] for f in ACC_FIELDS}
    line = {"accuracy_iterations": a.accuracy, "members": a.members,
            "press": "converged (FP64 defect correction)" if k.press_mode == 1 else "the reference's red-black sweeps",
            "workload": "gen_v1 on the native 41x181x361 shell" + (", per-member T +N(0,0.5K), w +N(0,2m/s)" if a.members > 1 else ""),
            **out}
    s = json.dumps(line)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")


# The uneven-size batch: (grid, members). The two coarser shells keep the 500 km depth
# (dr = 1/(im - 1)) at 2 and 3 degrees.
BATCH = [((41, 181, 361), 8), ((41, 91, 181), 16), ((21, 61, 121), 32)]


LAYOUTS = {0: "each grid on its own stream, side by side", 1: "one launch per RK4 stage for all grids",
           2: "one persistent launch per RK4 stage over per-XCD work queues"}


def batch(a):
    """The uneven-size batch: every grid of BATCH advanced together (atsat_group_run: one launch
    per RK4 stage for all grids, the same as a persistent launch over a work queue, or the grids
    side by side on their own streams, whichever the batch's autotuner finds fastest, with its
    tiles; the whole iteration pair captured in a hipGraph; every layout is also timed for 10
    iterations). 50 iterations per run: a.warmup untimed, 2 for the capture, the rest
    timed. FP32 (FP64 state and refinement) and FP64 runs; per grid, the FP32 step's per-field
    error against the FP64 step after the 50 iterations. The same contexts stepped one after
    the other, each alone, give the batching's gain."""
    import atsat_amd
    from atsat_amd import abi
    g0, c, k = atsat_amd.defaults()
    apply_physics(k, a.physics)
    k = type(k).from_buffer_copy(k)
    k.press_mode = 1 if a.press == "conv" else 0
    total = 50
    timed = total - a.warmup - 2
    grids = [g0 if shape == (g0.im, g0.jm, g0.km) else atsat_amd.small_grid(*shape) for shape, _ in BATCH]
    points = sum(M * g.im * g.jm * g.km for g, (_, M) in zip(grids, BATCH))
    n_rk = sum(M * (g.im - 2) * (g.jm - 4) * (g.km - 1) for g, (_, M) in zip(grids, BATCH))
    out = {}
    res = {}
    for prec, name in [(abi.PREC_FP32, "fp32"), (abi.PREC_FP64, "fp64")]:
        elem = 4 if prec == abi.PREC_FP32 else 8
        ctxs = []
        for gi, (g, (_, M)) in enumerate(zip(grids, BATCH)):
            ctx = atsat_amd.Context(g, c, k, M, prec)
            for m in range(M):
                ctx.upload_state(atsat_amd.gen_v1(g, c, seed=20261203 + gi, member=m), member=m)
            ctxs.append(ctx)
        grp = atsat_amd.Group(ctxs)
        it = 1
        grp.run(it, a.warmup, False)
        it += a.warmup
        grp.synchronize()
        t0 = time.perf_counter()
        stage_ms = grp.autotune()
        tune_s = time.perf_counter() - t0
        tuned = grp.layout()
        grp.run(it, 2, True)                                 # capture + first replay
        it += 2
        grp.synchronize()
        t0 = time.perf_counter()
        grp.run(it, timed, True)
        grp.synchronize()
        dt = time.perf_counter() - t0
        it += timed
        # the four RK4 stages: mean algorithmic bytes per integrated cell and launch
        per_cell = rk_bytes_per_cell(elem, k.turb_active, k.chem_block) / 4.0
        out[name] = {
            "value": points * timed / dt, "ms_per_step": dt / timed * 1e3,
            "tiles": [{"grid": "%dx%dx%d" % g[0], "members": g[1], "band": t[0], "rings": t[1],
                       "threads": t[2]} for g, t in zip(BATCH, grp.tiles())],
            "layout": LAYOUTS[tuned],
            "autotune_s": round(tune_s, 3),
            "rk_stage_ms": stage_ms,
            "rk_stage_GBps": per_cell * n_rk / (stage_ms * 1e-3) / 1e9,
            "rk_stage_hbm_frac": per_cell * n_rk / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "nonfinite": sum(ctx.nonfinite_count() for ctx in ctxs),
        }
        res[name] = [[{fl: ctx.download(fl, member=m) for fl in ACC_FIELDS} for m in range(M)]
                     for ctx, (_, M) in zip(ctxs, BATCH)]
        # after the 50 iterations: every layout for 10 iterations from its own graph
        per_layout = {}
        for lay in (0, 1, 2):
            grp.set_layout(lay)
            grp.run(it, 2, True)
            it += 2
            grp.synchronize()
            t0 = time.perf_counter()
            grp.run(it, 10, True)
            grp.synchronize()
            per_layout[LAYOUTS[lay]] = round((time.perf_counter() - t0) / 10 * 1e3, 3)
            it += 10
        out[name]["ms_per_step_by_layout"] = per_layout
        grp.close()
        # the same contexts stepped one after the other, each alone with its own graph
        for ctx in ctxs:
            ctx.run(it, 2, True)
        for ctx in ctxs:
            ctx.synchronize()
        t0 = time.perf_counter()
        for ctx in ctxs:
            ctx.run(it + 2, timed, True)
            ctx.synchronize()
        out[name]["alone_one_after_another_ms_per_step"] = (time.perf_counter() - t0) / timed * 1e3
        for ctx in ctxs:
            ctx.close()
    errs = {}
    for gi, (shape, M) in enumerate(BATCH):
        per = {}
        for fl in ACC_FIELDS:
            e = [_errs(res["fp32"][gi][m][fl], res["fp64"][gi][m][fl]) for m in range(M)]
            per[fl] = max(x[0] for x in e)
        errs["%dx%dx%d" % shape] = per
    line = {"metric": "member-grid-point-steps/s, uneven-size batch", "unit": "member-grid-point-steps/s",
            "value": out["fp32"]["value"], "dtype": "fp32 work / fp64 state",
            "config": {"workload": "uneven-size batch: " + ", ".join("%d x %dx%dx%d" % ((M,) + s) for s, M in BATCH)
                       + ", gen_v1 on each grid (seed per grid), 50 iterations (%d warm-up, 2 capture, %d timed)"
                       % (a.warmup, timed), "member_grid_points": points,
                       "press": "converged" if k.press_mode == 1 else "the reference's red-black sweeps"},
            "fp32": out["fp32"], "fp64": out["fp64"],
            "fp32_over_fp64": out["fp32"]["value"] / out["fp64"]["value"],
            "fp32_vs_fp64_rel_l2_max_over_members": errs}
    s = json.dumps(line)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")


def poisson_rhs(g, seed=20261203, n_modes=32):
    """The standalone pressure-solve source: a sum of n_modes orthonormal real spherical harmonics
    Y_lm(theta, phi) (l <= 40) times sin(k pi i / (im-1)) (k <= 8), amplitudes N(0, 1) k^-2."""
    import math
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(seed)
    dthe = g.dthe_deg * math.pi / 180.0
    dphi = g.dphi_deg * math.pi / 180.0
    the = np.arange(g.jm) * dthe
    phi = np.arange(g.km) * dphi
    ii = np.arange(g.im)
    out = np.zeros((g.im, g.jm, g.km))
    for _ in range(n_modes):
        l = int(rng.integers(0, 41))
        m = int(rng.integers(-l, l + 1))
        kk = int(rng.integers(1, 9))
        amp = rng.standard_normal() * kk ** -2.0
        Y = sph_harm_y(l, abs(m), the[:, None], phi[None, :])
        Yr = Y.real if m == 0 else (math.sqrt(2.0) * (Y.real if m > 0 else Y.imag))
        out += amp * np.sin(kk * math.pi * ii / (g.im - 1))[:, None, None] * Yr[None, :, :]
    return out


def poisson(a):
    """The standalone converged pressure solve on the native grid: FP64 corrections and FP32
    corrections inside FP64 defect correction, both to a relative residual of 1e-10."""
    import atsat_amd
    from atsat_amd import abi
    g, c, k = atsat_amd.defaults()
    src = poisson_rhs(g)
    res = {}
    sols = {}
    for prec, name in [(abi.PREC_FP64, "fp64"), (abi.PREC_FP32, "fp32_in_fp64")]:
        ctx = atsat_amd.Context(g, c, k, 1, prec)
        z = np.zeros((g.im, g.jm, g.km))
        t0 = time.perf_counter()
        ctx.upload("p_dyn", z)
        ctx.poisson_solve(src, 1e-10, 12)               # set-up (eigen-decompositions) + first solve
        setup = time.perf_counter() - t0
        ctx.upload("p_dyn", z)
        t0 = time.perf_counter()
        it, rel, ok = ctx.poisson_solve(src, 1e-10, 12)
        dt = time.perf_counter() - t0
        sols[name] = ctx.download("p_dyn")
        res[name] = {"corrections": it, "relative_residual": rel, "converged": ok, "solve_ms": round(1e3 * dt, 3),
                     "first_call_with_setup_s": round(setup, 3)}
        ctx.close()
    I = (slice(1, g.im - 1), slice(1, g.jm - 1), slice(0, g.km - 1))
    w = np.ones((g.im - 2, g.jm - 2, 1))
    w[[0, -1]] *= 1.5
    w[:, [0, -1]] *= 1.5
    def demean(p):
        q = p[I]
        return q - np.sum(w * q) / (np.sum(w) * (g.km - 1))
    a64, a32 = demean(sols["fp64"]), demean(sols["fp32_in_fp64"])
    res["fp32_in_fp64"]["rel_l2_vs_fp64"] = float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64))
    res["fp32_in_fp64"]["max_abs_vs_fp64"] = float(np.max(np.abs(a32 - a64)))
    line = {"poisson": "standalone pressure solve, native 41x181x361, 32 spherical-harmonic x radial-sine "
                       "modes, null-space component removed", **res}
    s = json.dumps(line)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")


# ---- per-kernel roofline table from rocprofv3 result databases --------------------------------

def kernel_alg_bytes(name, g, members, elem, turb, stage_bytes):
    """Algorithmic bytes of one dispatch of a kernel of the model step (each value a launch reads or
    writes counted once; the state class -- t, w, the pressures, rho_mix -- FP64, the work class,
    which holds the other prognostic fields, `elem` bytes), or None where the kernel touches only
    boundary planes or scalars. stage_bytes: per-stage bytes per integrated cell of the RK4 stage
    kernel."""
    im, jm, km = g
    M = members
    N = im * jm * km * M                                # every cell
    Nrk = (im - 2) * (jm - 4) * (km - 1) * M            # RK4-integrated cells
    Ni = (im - 2) * (jm - 2) * (km - 1) * M             # Poisson / closure interior cells
    C = jm * km * M                                     # columns
    w = elem
    nint = 17 if turb else 15
    if "k_rk_cell" in name:
        st = int(name.split("k_rk_cell<")[1].split(",")[1])
        return stage_bytes[st] * Nrk
    if "k_press_terms" in name:     # all six terms: p_dyn, p_metric, p_stat, rho in; metric only: p_metric
        return (32 + 6 * w if "true>" in name else 8 + 2 * w) * Ni
    if name.startswith("Cijk_"):                        # rocBLAS GEMM of a solver transform
        e = 8 if "_DB_" in name else 4
        return 2 * e * Ni
    R = im * jm * M                                     # (member, i, j) rings of the solver
    if "k_ring_fwd" in name:                            # ring FFT: each ring read, its modes written
        e = 4 if "k_ring_fwd<float" in name else 8
        return 2 * e * R * (km - 1)
    if "k_ring_inv" in name:                            # modes read; the increment read and written
        return 8 * R * (km - 1) + (16 * Ni if "true>" in name else 8 * R * (km - 1))
    state = 2 * 8 + (nint - 2) * w                      # one value of each integrated field
    table = {
        "k_buoy_terms": 24 * N, "k_buoy_sum": 8 * N, "k_buoy_chain": 16 * N, "k_metric_pressure": (24 + w) * N,
        "k_rk_halo": (2 * jm * km + 2 * (im - 2) * km) * M * (state + 2 * nint * w),
        "k_mixture_density": 24 * N, "k_press_src": (3 * w + 16) * Ni,
        "k_damp_k": 16 * N, "k_damp_i": 16 * N,
        # t, p_stat, rho_mix and the nine species in; t and the nine species out
        "k_satadj": (32 + 18 * w) * N, "k_rebuild_pstat": 16 * N, "k_nh4sh_eq": (16 + 3 * w) * N,
        "k_precipitation": (24 + 10 * w + 20 * w) * N,
        "k_diff_mass_flux": (8 + 3 * w + 4 * w) * Ni, "k_shapiro_kj": 8 * w * N, "k_shapiro_i": 8 * w * N,
        "k_chem_massflux": (8 + 3 * w + 7 * w) * Ni,
        "k_copy_n": (2 if turb else 0) * 2 * w * N,
        # w, t, rho_mix (FP64) and u, v, k, omega and the nine species in; four outputs
        "k_turb_sources": (24 + 13 * w + 4 * w) * Ni, "k_turb_cells": 6 * w * N,
        "k_mca_smooth": (16 + 11 * w + 48 + 1) * N, "k_mca_census": (39 * 72 + 8) * C, "k_mca_mix": 8 * C,
        "k_sol_residual": (16 + w) * Ni, "k_sol_defect": 16 * Ni, "k_sol_update": 24 * Ni,
        "k_sol_add_shifted": 24 * N, "k_sol_shift": 16 * N, "k_sol_wsum": 8 * Ni, "k_sol_fill": 8 * N,
        "k_sol_radial": (w + 8) * Ni,
        "k_census": state * N,
    }
    key = name.split("<")[0].replace("void ", "").replace("atsat::", "").strip()
    return table.get(key)


def _rocpd(path, what):
    import sqlite3
    db = sqlite3.connect(path)
    cur = db.cursor()
    if what == "kernels":
        return cur.execute("select dispatch_id, name, duration from kernels order by dispatch_id").fetchall()
    if what == "copy":                                  # the runtime's 1 GiB device copy
        return max(v for n, v in cur.execute(
            "select kernel_name, sum(value) from counters_collection group by dispatch_id").fetchall()
            if "copyBuffer" in n)
    return {d: v for d, v in cur.execute(
        "select dispatch_id, sum(value) from counters_collection group by dispatch_id").fetchall()}


def trace_summary(path, top=40):
    """Markdown table of a kernel trace (rocprofv3 --kernel-trace, rocpd database): per kernel the
    launches, mean and total duration and the share of the kernel time, sorted by total; and the
    mean over every launch of the RK4 stage kernel (all four stages)."""
    import re
    kern = _rocpd(path, "kernels")
    agg = {}
    for _, n, du in kern:
        key = re.sub(r"\(.*", "", n).replace("void ", "").replace("atsat::", "")
        r = agg.setdefault(key, [0, 0.0])
        r[0] += 1
        r[1] += du
    tot = sum(v[1] for v in agg.values())
    out = ["| kernel | launches | mean ms | total ms | share of kernel time |", "|---|---|---|---|---|"]
    for key, (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append("| `%s` | %d | %.4f | %.2f | %.1f%% |" % (key[:90], n, ns / n * 1e-6, ns * 1e-6, 100.0 * ns / tot))
    out.append("")
    line = "Kernel time in the trace: %.1f ms over %d launches." % (tot * 1e-6, len(kern))
    for prec in ("float", "double"):
        rk = [du for _, n, du in kern if "k_rk_cell<%s," % prec in n]
        if rk:
            line += " RK4 stage kernel `k_rk_cell<%s, *>`: %d launches, mean %.4f ms over the four stages." % (
                prec, len(rk), sum(rk) / len(rk) * 1e-6)
    out.append(line)
    print("\n".join(out))


def roofline_table(a):
    """Markdown table of every kernel of the last timed odd + even iteration pair of a profiled
    `bench.py --steps 2` run: mean duration (kernel trace), HBM traffic from FETCH_SIZE and
    WRITE_SIZE (scaled by a plain device copy's readings), achieved GB/s and the algorithmic
    bytes' GB/s against the 8 TB/s peak."""
    import re
    tr, fe, wr, cfe, cwr = a.roofline_table
    kern = _rocpd(tr, "kernels")
    fetch, write = _rocpd(fe, "pmc"), _rocpd(wr, "pmc")
    cal_f = _rocpd(cfe, "copy")                        # the 1 GiB copy: KiB read as counted
    cal_w = _rocpd(cwr, "copy")
    rf, wf = (1 << 20) / cal_f, (1 << 20) / cal_w       # counted KiB -> true KiB
    names = [re.sub(r"\(.*", "", n) for _, n, _ in kern]
    # iterations start with the level mean (k_buoy_terms in FP64, k_buoy_chain in FP32); the window
    # is the last even iteration (the one with the saturation adjustment) and the odd one before it
    starts = [i for i, n in enumerate(names) if "k_buoy_terms<" in n or "k_buoy_chain<" in n]
    sat = [i for i, n in enumerate(names) if "k_satadj<" in n]
    ev = max(b for b in starts if b < sat[-1])
    lo = max(b for b in starts if b < ev)
    nxt = [b for b in starts if b > ev]
    cen = [i for i, n in enumerate(names) if "k_census<" in n and i > ev]
    hi = min(nxt + cen + [len(names)])
    elem = 4 if a.prec == "fp32" else 8
    turb = a.physics == "model"
    stage = rk_stage_bytes(elem, turb, turb)
    g = (41, 181, 361)
    rows = {}
    for i in range(lo, hi):
        d, n, du = kern[i]
        key = names[i].replace("void ", "").replace("atsat::", "")
        if key.startswith("Cijk_"):                     # a rocBLAS (Tensile) GEMM: its tile name
            key = key[:key.find("_SN_")] + "… (rocBLAS)" if "_SN_" in key else key[:40] + "… (rocBLAS)"
        r = rows.setdefault(key, {"n": 0, "ns": 0.0, "rd": 0.0, "wr": 0.0, "alg": 0.0, "alg_ok": True})
        r["n"] += 1
        r["ns"] += du
        r["rd"] += fetch.get(d, 0.0) * rf * 1024.0
        r["wr"] += write.get(d, 0.0) * wf * 1024.0
        ab = kernel_alg_bytes(names[i], g, a.members, elem, turb, stage)
        if ab is None:
            r["alg_ok"] = False
        else:
            r["alg"] += ab
    tot = sum(r["ns"] for r in rows.values())
    out = ["| kernel | launches per 2 iterations | ms per launch | ms per iteration | HBM read GB | HBM write GB "
           "| achieved HBM GB/s | algorithmic GB | algorithmic GB/s | fraction of 8 TB/s | traffic / algorithmic |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]
    for key, r in sorted(rows.items(), key=lambda kv: -kv[1]["ns"]):
        ms = r["ns"] / r["n"] * 1e-6
        rd, wrb = r["rd"] / r["n"], r["wr"] / r["n"]
        hbm = (rd + wrb) / (ms * 1e-3) / 1e9
        if r["alg_ok"]:
            alg = r["alg"] / r["n"]
            ag = alg / (ms * 1e-3) / 1e9
            cols = "%.3f | %.0f | %.3f | %.2f" % (alg / 1e9, ag, ag / HBM_PEAK_GBS, (rd + wrb) / alg if alg else 0.0)
        else:
            cols = "— | — | — | —"
        out.append("| `%s` | %d | %.3f | %.3f | %.3f | %.3f | %.0f | %s |" % (
            key, r["n"], ms, r["ns"] * 1e-6 / 2, rd / 1e9, wrb / 1e9, hbm, cols))
    out.append("")
    out.append("Total kernel time per iteration (mean of the odd and the even one): %.2f ms. Calibration: "
               "the 1 GiB device copy counts %.0f KiB read (FETCH_SIZE) and %.0f KiB written (WRITE_SIZE), "
               "so read bytes = FETCH_SIZE x %.3f KiB and write bytes = WRITE_SIZE x %.3f KiB."
               % (tot * 1e-6 / 2, cal_f, cal_w, rf, wf))
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--prec", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-graph", action="store_true",
                    help="time the iterations launched on the stream instead of replayed from the captured hipGraph")
    ap.add_argument("--press", default="conv", choices=["conv", "ref"],
                    help="conv: the converged pressure solve (default); ref: the reference's red-black sweeps")
    ap.add_argument("--physics", default="model", choices=["model", "resident"],
                    help="model: the closure, the adjustment and the chemistry block on (default); resident: all off")
    ap.add_argument("--other-press", action="store_true", help="also time the other pressure variant")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default RK4 stage tile")
    ap.add_argument("--tile", type=int, nargs=2, default=None, metavar=("BAND", "RINGS"),
                    help="set this RK4 stage tile instead of tuning (profiler runs: no tuning dispatches)")
    ap.add_argument("--solve-transforms", default=None, choices=["fp32", "fp64"],
                    help="FP32 step, converged solve: the corrections' forward transforms (default: the library's)")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--out", default=None)
    ap.add_argument("--accuracy", type=int, default=0, metavar="ITERS",
                    help="instead of timing, report the FP32 step's per-field error against the FP64 "
                         "step after ITERS iterations (--members members)")
    ap.add_argument("--oracle", action="store_true", help="with --accuracy: also against the CPU oracle")
    ap.add_argument("--poisson", action="store_true", help="only the standalone pressure solve")
    ap.add_argument("--batch", action="store_true",
                    help="only the uneven-size batch (8 x 41x181x361, 16 x 41x91x181, 32 x 21x61x121)")
    ap.add_argument("--trace-summary", default=None, metavar="DB",
                    help="print the per-kernel summary of a rocprofv3 --kernel-trace database and exit")
    ap.add_argument("--roofline-table", nargs=5, default=None,
                    metavar=("TRACE_DB", "FETCH_DB", "WRITE_DB", "CAL_FETCH_DB", "CAL_WRITE_DB"),
                    help="print the per-kernel roofline table of a profiled --steps 2 run from rocprofv3 result "
                         "databases (kernel trace, FETCH_SIZE, WRITE_SIZE, and the two of --copy-calibration 1)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB gibibytes (the reference point of the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only run atsat_amd/hbm_probe (the device's HBM streaming rates: a triad, many "
                         "arrays streamed at once, and a 16-field stencil stage shaped like the RK4 stage "
                         "kernel) and print its JSON lines")
    a = ap.parse_args()
    if a.hbm_probe:
        probe = os.path.join(HERE, "atsat_amd", "hbm_probe")
        if not os.path.exists(probe):
            sys.exit("bench.py: %s is not built (python -m atsat_amd.build)" % probe)
        sys.exit(subprocess.call([probe]))
    if a.trace_summary:
        trace_summary(a.trace_summary)
        return
    if a.roofline_table:
        roofline_table(a)
        return
    if a.copy_calibration > 0:
        # plain device streams of known size: the reference point of the FETCH_SIZE /
        # WRITE_SIZE counters and the HBM bandwidth the runtime's own kernels reach
        import torch
        n = int(a.copy_calibration * (1 << 30)) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        nb = n * 4
        out = {"bytes_per_array": nb}
        for name, fn, moved in [("copy", lambda: y.copy_(x), 2 * nb), ("fill", lambda: z.fill_(2.0), nb),
                                ("read_sum", lambda: x.sum(), nb),
                                ("triad", lambda: torch.add(x, y, alpha=2.0, out=z), 3 * nb)]:
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(3):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 3
            out[name + "_GBps"] = round(moved / (ms * 1e-3) / 1e9, 1)
        print(json.dumps(out))
        return
    if a.accuracy > 0:
        accuracy(a)
        return
    if a.poisson:
        poisson(a)
        return
    if a.batch:
        batch(a)
        return
    if a.gpus != 1:
        print("bench.py: this benchmark measures one GPU (ensemble members would shard across GPUs "
              "with no exchange between them); run --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.steps % 2:
        a.steps += 1                               # whole odd + even pairs
    import atsat_amd
    from atsat_amd import abi
    g, c, k = atsat_amd.defaults()
    apply_physics(k, a.physics)
    if a.solve_transforms:
        k.press_fp32_transforms = 1 if a.solve_transforms == "fp32" else 0
    npts = g.im * g.jm * g.km
    n_rk = (g.im - 2) * (g.jm - 4) * (g.km - 1)
    prec = abi.PREC_FP32 if a.prec == "fp32" else abi.PREC_FP64
    elem = 4 if prec == abi.PREC_FP32 else 8
    press_mode = 1 if a.press == "conv" else 0
    print("bench: %d members %s, %d warmup + %d timed steps, pressure %s" % (a.members, a.prec, a.warmup, a.steps, a.press),
          file=sys.stderr, flush=True)
    states = member_states(g, c, a.members)
    R = run_gpu(g, c, k, states, prec, a.steps, a.warmup, press_mode, autotune=not a.no_autotune, tile=a.tile,
                graph=not a.no_graph)
    dt, prof = R["dt"], R["prof"]
    value = a.members * npts * a.steps / dt
    fp64 = None
    if prec == abi.PREC_FP32 and not a.no_fp64:
        print("bench: FP64 comparison run", file=sys.stderr, flush=True)
        R64 = run_gpu(g, c, k, states, abi.PREC_FP64, a.steps, a.warmup, press_mode, autotune=not a.no_autotune,
                      tile=a.tile, graph=not a.no_graph)
        fp64 = {"value": a.members * npts * a.steps / R64["dt"], "ms_per_step": 1e3 * R64["dt"] / a.steps,
                "nonfinite": R64["nonfinite"], "last_pressure_solve": R64["solve"], "rk_tile": R64["tile"],
                "kernel_ms_per_step": {n: round(v[0] / R64["prof_steps"], 3) for n, v in R64["prof"].items()}}
    other = None
    if a.other_press:
        print("bench: the other pressure variant", file=sys.stderr, flush=True)
        om = 1 - press_mode
        o32 = run_gpu(g, c, k, states, abi.PREC_FP32, a.steps, a.warmup, om, graph=not a.no_graph)
        o64 = run_gpu(g, c, k, states, abi.PREC_FP64, a.steps, a.warmup, om, graph=not a.no_graph)
        other = {"press": "reference red-black sweep" if om == 0 else "converged",
                 "fp32_value": a.members * npts * a.steps / o32["dt"], "fp32_ms_per_step": 1e3 * o32["dt"] / a.steps,
                 "fp64_value": a.members * npts * a.steps / o64["dt"], "fp64_ms_per_step": 1e3 * o64["dt"] / a.steps}
    roof = None
    if prof.get("rk_stage", (0, 0))[1] > 0:
        ms, nl = prof["rk_stage"]
        per_launch_bytes = rk_bytes_per_cell(elem, k.turb_active, k.chem_block) / 4.0 * n_rk * a.members
        achieved = per_launch_bytes / (ms / nl * 1e-3) / 1e9
        traffic = measured_traffic(elem, a.physics) if a.members == 16 else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/step_kernels_roofline_%s_ensemble16.md (FETCH_SIZE and WRITE_SIZE, "
                                  "calibrated on a 1 GiB device copy)" % a.prec if traffic else None,
                "kernel": "k_rk_cell (fused RHSSat + RK4 stage)",
                "mean_launch_ms": round(ms / nl, 4), "algorithmic_bytes_per_launch": int(per_launch_bytes)}
        ceil_gbs = streaming_ceiling()
        if ceil_gbs:
            roof["streaming_ceiling"] = {"GBps": ceil_gbs, "frac_of_peak": round(ceil_gbs / HBM_PEAK_GBS, 4),
                                         "achieved_frac_of_ceiling": round(achieved / ceil_gbs, 4),
                                         "source": "profiles/device_streams_bandwidth.md (hbm_probe: 48-96 "
                                                   "arrays streamed at once, the lowest rate)"}
    cpu = None
    if not a.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(g, c, k, a.cpu_iters)
    sim_s = a.steps * k.dt * c.L_atm * 1e3 / c.u_0
    press_txt = ("converged solve to relative residual %.0e every even iteration: FP64 defect correction, "
                 "corrections with %s forward transforms" % (k.press_tol, "FP32" if prec == abi.PREC_FP32 and k.press_fp32_transforms
                                                              else "FP64")) if press_mode == 1 \
        else "the reference's red-black sweeps (%d per even iteration)" % k.press_sweeps
    line = {
        "metric": "member-grid-point-steps/s (model iteration: RK4 step + every other iteration the "
                  "pressure solve, saturation adjustment and precipitation), FP32-refined vs FP64",
        "physics": ("k-omega SST closure and moist convective adjustment every iteration, chemistry block "
                    "(NH4SH equilibrium, mass fluxes, flux limiter) every even iteration; radiation off"
                    if a.physics == "model" else "closure, chemistry, convective adjustment and radiation off"),
        "value": value, "unit": "member-grid-point-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * dt / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 stencils and RK scratch, fp64 state, reductions, physics and pressure residual" if prec == abi.PREC_FP32 else "fp64",
        "data": "synthetic (gen_v1, seeded)",
        "config": {"workload": "ensemble of %d members on the native 41x181x361 shell (gen_v1, "
                               "T +N(0,0.5K), w +N(0,2m/s) per member)" % a.members,
                   "members": a.members, "grid": [g.im, g.jm, g.km], "dt": k.dt, "press": press_txt},
        "simulated_s_per_wall_s": sim_s / dt,
        "nonfinite_after": R["nonfinite"],
        "last_pressure_solve": R["solve"],
        "device_bytes": R["device_bytes"],
        "context_setup_s": round(R["setup_s"], 2),
        "upload_ms_per_member": round(1e3 * R["upload_s"], 1),
        "fp64_same_workload": fp64,
        "timing": ("%d iterations replayed from the hipGraph of an even + odd iteration pair, no events inside "
                   "the timed region; kernel_ms_per_step and the roofline from a profiled pass of %d iterations "
                   "launched on the stream, HIP events around every kernel family"
                   % (a.steps, R["prof_steps"])) if not a.no_graph else
                  ("%d iterations launched on the stream; kernel_ms_per_step and the roofline from a profiled "
                   "pass of %d iterations" % (a.steps, R["prof_steps"])),
        "profiled_pass_ms_per_step": round(1e3 * R["dt_prof"] / R["prof_steps"], 3),
        "kernel_ms_per_step": {n: round(v[0] / R["prof_steps"], 3) for n, v in prof.items()} if prof else None,
        "rk_tile": R["tile"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if other:
        line["other_pressure_variant"] = other
    if fp64:
        line["fp32_over_fp64"] = value / fp64["value"]
    s = json.dumps(line)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")


if __name__ == "__main__":
    main()
