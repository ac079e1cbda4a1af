"""The converged pressure solve (press_mode 1) against the reference's operator restated in numpy.

The reference only relaxes the pressure equation with a fixed number of red-black sweeps
(PressureSolver.h:145-766), so the converged solve has no reference output to match; it is
pinned by what the mathematics fixes: the residual of the reference's own operator
(tests/poisson_util.py, PressureSolver.h:350-353, 476-491, boundary pass :718-740), recomputed
here in numpy from the downloaded p_dyn, and the recovery of a manufactured solution up to the
constant of the all-Neumann null space.
"""
import sys

import numpy as np
import pytest

import atsat_amd
from atsat_amd import abi
import poisson_util as pu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid():
    return atsat_amd.small_grid(21, 61, 121)


@pytest.mark.parametrize("prec,shape", [(abi.PREC_FP64, (21, 61, 121)), (abi.PREC_FP32, (21, 61, 121)),
                                        (abi.PREC_FP64, (45, 31, 61)), (abi.PREC_FP32, (45, 31, 61)),
                                        (abi.PREC_FP64, (21, 41, 57)), (abi.PREC_FP32, (21, 41, 57))])
def test_manufactured_solution_recovered(defaults, prec, shape):
    """(45, 31, 61): 43 interior levels, above the 40 of the register radial solve, so the
    corrections take the general one (per-lane arrays). Rings of 60, 90, 120, 180 and 360 points
    take the Stockham FFT in LDS; (21, 41, 57), a ring of 56 points, takes the transforms as
    rocBLAS GEMMs."""
    _, c, k = defaults
    # layers of 7.8125 km (dr = 1/64), exact in float, for the taller grid
    g = atsat_amd.small_grid(*shape) if shape[0] <= 41 else atsat_amd.small_grid(*shape, dr=1.0 / 64.0)
    rng = np.random.default_rng(11)
    shp = (2, g.im, g.jm, g.km)
    i = np.arange(g.im)[:, None, None]
    j = np.arange(g.jm)[None, :, None]
    kk = np.arange(g.km)[None, None, :]
    smooth = np.cos(2 * np.pi * kk / (g.km - 1) * 3) * np.sin(np.pi * j / (g.jm - 1)) * np.cos(np.pi * i / (g.im - 1))
    pstar = np.stack([smooth + 0.01 * rng.standard_normal(shp[1:]), 0.5 * smooth ** 2 + 0.02 * rng.standard_normal(shp[1:])])
    pstar = pu.boundary_pass(pstar)
    src = pu.apply(pstar, g, c)
    say("case manufactured solution, precision", prec, "grid", shape)
    ctx = atsat_amd.Context(g, c, k, 2, prec)
    for m in range(2):
        ctx.upload("p_dyn", np.zeros(shp[1:]), member=m)
    it, rel, ok = ctx.poisson_solve(src, tol=1e-10, max_iter=12)
    say("  corrections", it, "relative residual", rel)
    p = np.stack([ctx.download("p_dyn", member=m) for m in range(2)])
    ctx.close()
    assert ok and rel <= 1e-10
    own = pu.rel_residual(p, src, g, c)
    say("  numpy residual", own)
    assert np.all(own <= 2e-10)
    d = pu.interior(p - pstar, g)
    d -= d.mean(axis=(-3, -2, -1), keepdims=True)
    err = np.sqrt(np.sum(d * d, axis=(-3, -2, -1))) / np.sqrt(np.sum(pu.interior(pstar, g) ** 2, axis=(-3, -2, -1)))
    say("  solution error (constant removed)", err)
    assert np.all(err <= 1e-7)
    # the boundary relations hold and the gauge (weighted mean 0 at entry) is kept
    q = pu.boundary_pass(p.copy())
    assert np.allclose(q, p, rtol=0, atol=1e-12 * np.abs(p).max())
    assert np.all(np.abs(pu.wmean(p, g)) <= 1e-12 * np.abs(p).max())


def test_step_source_converges(defaults):
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    st["rhs_u"][...] = 1e-3 * np.sin(np.arange(g.km) * 0.1)[None, None, :]
    st["rhs_v"][...] = 1e-3 * np.cos(np.arange(g.jm) * 0.2)[None, :, None]
    st["rhs_w"][...] = 1e-3 * np.cos(np.arange(g.im) * 0.3)[:, None, None]
    say("case the step's own source, gen_v1 small grid")
    ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP64)
    ctx.upload_state(st)
    p0 = ctx.download("p_dyn")
    it, rel, ok = ctx.poisson_solve(None, tol=1e-10, max_iter=12)
    say("  corrections", it, "relative residual", rel)
    assert ok and rel <= 1e-10
    p = ctx.download("p_dyn")
    ctx.close()
    # the converged p_dyn satisfies Lc p = projected(clamped div(rhs) + Lc p0): the increment e = p - p0
    # satisfies Lc e = projected clamped divergence, which is bounded -- check the gauge and finiteness
    assert np.all(np.isfinite(p))
    assert abs(pu.wmean(p, g) - pu.wmean(p0, g)) <= 1e-12 * max(1.0, np.abs(p).max())


def test_step_with_converged_solve_graph_matches_stream(defaults):
    _, c, k0 = defaults
    g = _grid()
    k = abi.Knobs.from_buffer_copy(k0)        # the session's knobs stay as they are
    k.press_mode = 1
    st = atsat_amd.gen_v1(g, c)
    say("case step with press_mode 1, stream vs graph, 4 iterations")
    out = []
    for use_graph in (False, True):
        ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP64)
        ctx.upload_state(st)
        ctx.run(1, 4, use_graph)
        ctx.synchronize()
        out.append({f: ctx.download(f) for f in ("u", "v", "w", "t", "p_dyn")})
        if not use_graph:
            say("  last solve", ctx.pressure_last())
        ctx.close()
    for f in out[0]:
        a, b = out[0][f], out[1][f]
        e = np.sqrt(np.sum((a - b) ** 2)) / max(np.sqrt(np.sum(b * b)), 1e-300)
        say("  %s stream vs graph rel L2 %.2e" % (f, e))
        assert e <= 1e-8, f


def test_spherical_harmonic_source_fp32_matches_fp64(defaults):
    """The standalone-solve source (spherical harmonics x radial sines): both contexts stop at
    1e-10 or at the FP64 residual floor, agree with each other, and their residual is the one the
    numpy restatement computes."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    _, c, k = defaults
    g = _grid()
    src = bench.poisson_rhs(g)
    sols = {}
    for prec in (abi.PREC_FP64, abi.PREC_FP32):
        say("case spherical-harmonic source, precision", prec)
        ctx = atsat_amd.Context(g, c, k, 1, prec)
        ctx.upload("p_dyn", np.zeros((g.im, g.jm, g.km)))
        it, rel, ok = ctx.poisson_solve(src, tol=1e-10, max_iter=12)
        p = ctx.download("p_dyn")
        ctx.close()
        own = pu.rel_residual(p, pu.project(src, g), g, c)
        say("  corrections", it, "relative residual", rel, "numpy", own)
        assert ok and rel <= 1e-8 and abs(own - rel) <= 0.05 * rel + 1e-14
        sols[prec] = pu.interior(p, g) - pu.wmean(p, g)
    a, b = sols[abi.PREC_FP64], sols[abi.PREC_FP32]
    e = np.linalg.norm(a - b) / np.linalg.norm(a)
    say("  fp32 vs fp64 solution", e)
    assert e <= 1e-9


def test_graph_replayed_solve_reports_nonfinite(defaults):
    """A converged solve replayed from a graph does not read its residual back; the device records
    its outcome and the next synchronisation returns it: a clean replay is ATSAT_OK, a replay whose
    p_dyn holds a NaN (the source holds Lc(p_dyn), PressureSolver.h:249-282) is ATSAT_E_NONFINITE,
    as the same iteration launched on the stream is at the solve itself."""
    _, c, k0 = defaults
    g = _grid()
    k = abi.Knobs.from_buffer_copy(k0)
    k.press_mode = 1
    st = atsat_amd.gen_v1(g, c)
    say("case graph-replayed converged solve: clean, then a NaN in p_dyn")
    ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP64)
    ctx.upload_state(st)
    ctx.run(1, 5, True)                      # one eager odd iteration, the capture, replays
    ctx.synchronize()
    pd = ctx.download("p_dyn")
    pd[g.im // 2, g.jm // 2, 7] = np.nan
    ctx.upload("p_dyn", pd)
    ctx.run(6, 2, True)                      # even + odd replayed: the solve meets the NaN
    with pytest.raises(atsat_amd.AtsatError, match="non-finite"):
        ctx.synchronize()
    ctx.close()
    say("  the same on the stream")
    ctx = atsat_amd.Context(g, c, k, 1, abi.PREC_FP64)
    ctx.upload_state(st)
    ctx.run(1, 1)
    ctx.upload("p_dyn", pd)
    with pytest.raises(atsat_amd.AtsatError, match="non-finite"):
        ctx.run(2, 1)
    ctx.close()
