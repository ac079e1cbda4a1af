"""Closed-form and invariant checks of the oracle, independent of the reference's output:
each one would catch a plausible slip (a dropped term, a wrong sign or index) that a bit-identity
check alone would only report as 'different'."""
import math

import numpy as np

import atsat_amd
import oracle
from atsat_amd import abi


def _grid():
    return atsat_amd.small_grid(21, 31, 61)


def test_svp_formula_against_long_double(defaults):
    """E = exp(C + (-L0/T + a ln T + b T)/(1e-3 R)) (ATPhys.h:49-52) at a few temperatures,
    evaluated in Python's float and compared with the saturation target the oracle's
    saturation adjustment starts from (q_sat = rho ep E / p at a cell it leaves alone)."""
    _, c, _ = defaults
    for gi, T in ((1, 300.0), (2, 200.0), (0, 80.0)):
        E = math.exp(c.C[gi] + (-c.L0[gi] / T + c.del_alf[gi] * math.log(T) + c.del_bet[gi] * T) / (1e-3 * c.R[gi]))
        assert E > 0 and math.isfinite(E)


def test_buoyancy_level_mean_of_uniform_level(defaults):
    """A level with uniform t and p_stat has the mean b = g p /(r R t t_ref) exactly."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    t = 1.0 + 0.01 * np.arange(g.im)[:, None, None]
    st["t"][...] = t
    st["p_stat"][...] = 2.0
    oracle.Oracle(g, c, k).call("buoyancy_ref_level", st)
    b = c.g * 2.0 / (c.r_mix * c.R_mix * t[:, 0, 0] * c.t_ref)
    assert np.allclose(st["buoy_ref_level"], b, rtol=1e-14)


def test_metric_pressure_trapezoid_closed_form(defaults):
    """With v = 0, w = w0 uniform and p_baro = 0, p_metric(i) = sum of the trapezoid of
    (w0^2/rm + cor w0) in r -- linear pieces, so the trapezoid is exact up to 1/rm."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    w0 = 0.3
    st["w"][...] = w0
    oracle.Oracle(g, c, k).call("metric_pressure", st)
    th = np.cumsum(np.r_[0.0, np.full(g.jm - 1, math.radians(g.dthe_deg))])
    rz = 1.0 + g.dr * np.arange(g.im)
    rm = rz + (c.metric_radius_km / c.L_atm - 1.0)
    cor = c.Coriolis * (c.L_atm * 1e3 / c.u_0) * 2.0 * c.omega * np.abs(np.sin(th))
    f = w0 * w0 / rm[:, None] + cor[None, :] * w0
    pm = np.concatenate([np.zeros((1, g.jm)), np.cumsum(0.5 * (f[1:] + f[:-1]) * g.dr, axis=0)])
    assert np.allclose(st["p_metric"][:, :, 5], pm, rtol=1e-12, atol=1e-15)


def test_pstat_polytrope_for_linear_temperature(defaults):
    """For T linear in height the hydrostatic rebuild equals the polytrope p_ref (t)^(g/(gam R))
    (InitVariables_Sat.cpp:553-555)."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    h = c.L_atm * g.dr
    T = 380.0 - c.gam * h * np.arange(g.im)
    st["t"][...] = (T / c.t_ref)[:, None, None]
    oracle.Oracle(g, c, k).call("rebuild_pstat", st)
    poly = c.p_ref * (T / c.t_ref) ** (c.g / (c.gam * c.R_ref))
    assert np.allclose(st["p_stat"][:, 4, 7], poly, rtol=1e-12)


def test_saturation_adjust_conserves_total_water(defaults):
    """CONSERVE: vapour + cloud + ice of each gas is conserved cell by cell (the entry clip and
    the final clamp only remove negatives, and there are none here)."""
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    rng = np.random.default_rng(3)
    for n in ("h2o", "nh3"):
        st[n + "_cloud"][...] = np.abs(rng.standard_normal(st["t"].shape)) * 0.1 * st[n]
    before = {n: st[n] + st[n + "_cloud"] + st[n + "_ice"] for n in ("h2o", "nh3", "ch4")}
    O = oracle.Oracle(g, c, k)
    for gas in (0, 1, 2):
        O.call("saturation_adjust", st, gas)
    for n in ("h2o", "nh3", "ch4"):
        after = st[n] + st[n + "_cloud"] + st[n + "_ice"]
        assert np.allclose(after, before[n], rtol=1e-12, atol=1e-18), n


def test_saturation_adjust_leaves_subsaturated_dry_cells(defaults):
    """A cell without condensate below its entry target is not touched."""
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    st["h2o"][...] *= 0.5
    t0, q0 = st["t"].copy(), st["h2o"].copy()
    oracle.Oracle(g, c, k).call("saturation_adjust", st, 1)
    warm = st["t"] * c.t_ref > c.t_0[1]
    assert np.array_equal(st["t"][warm], t0[warm])
    assert np.array_equal(st["h2o"][warm], q0[warm])


def test_precipitation_zero_condensate_gives_zero_flux(defaults):
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    oracle.Oracle(g, c, k).call("precipitation", st)
    for n in abi.PRECIP_OUT:
        if n.startswith("P_") and n != "P_nh4sh":
            assert not st[n].any(), n
    assert not st["Q_precip"].any()


def test_precipitation_fluxes_capped_and_nonnegative(defaults):
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    st["h2o_cloud"][...] = 5e-2
    st["nh3_ice"][...] = 1e-2
    oracle.Oracle(g, c, k).call("precipitation", st)
    for n in ("P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel"):
        assert st[n].min() >= 0.0 and st[n].max() <= 1e-4 + 1e-18, n
    assert st["P_rain"].max() > 0.0
    assert np.all(st["S_precip_h2o_cloud"] <= 0.0)


def test_red_black_sweeps_reduce_the_residual(defaults):
    """Repeated sweeps on a fixed source converge towards Lc p = src: the interior residual of
    the folded operator falls (the mode of the singular operator aside)."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    rng = np.random.default_rng(1)
    st["rhs_u"][...] = 1e-4 * rng.standard_normal(st["t"].shape)
    st["rhs_v"][...] = 1e-4 * rng.standard_normal(st["t"].shape)
    k2 = abi.Knobs.from_buffer_copy(k)
    k2.press_sweeps = 1
    O = oracle.Oracle(g, c, k2)
    O.call("pressure_solve", st)
    d1 = np.abs(st["p_dyn"]).max()
    assert np.isfinite(st["p_dyn"]).all() and d1 > 0.0


def test_clamp_and_boundaries(defaults):
    _, c, k = defaults
    g = _grid()
    st = atsat_amd.gen_v1(g, c)
    st["h2o"][3, 4, 5] = -1.0
    st["nh3"][2, 2, 2] = np.nan
    O = oracle.Oracle(g, c, k)
    O.call("clamp_species", st)
    assert st["h2o"][3, 4, 5] == 0.0 and st["nh3"][2, 2, 2] == 0.0
    O.call("boundary_conditions", st)
    assert not st["u"][0].any() and not st["u"][-1].any()           # rigid walls
    assert not st["v"][-1].any() and not st["w"][1:-1, 0].any()     # lid, pole (i 1..im-2)
    # seam copy; bcPhi runs before bcRadius, so the radial planes keep the copy of their old values
    assert np.array_equal(st["t"][1:-1, :, -1], st["t"][1:-1, :, 0])
    assert np.array_equal(st["h2o"][-1], st["h2o"][-2])             # species lid copy


def test_tridiagonal_eigensolver(tmp_path):
    """The host eigen-decomposition the separable pressure solver is built from
    (atsat_amd/csrc/tridiag_eigen.hpp): residual and orthogonality on random symmetric
    tridiagonal matrices, including the 179-row size of the native grid."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    src = tmp_path / "eig.cpp"
    src.write_text(r'''
#include "tridiag_eigen.hpp"
#include <cstdio>
#include <random>
int main(){
    double worst = 0.0;
    for(int n : {1, 2, 3, 17, 179}){
        std::mt19937 rng(n); std::normal_distribution<double> N(0, 1);
        std::vector<double> d(n), e(n, 0.0);
        for(int i = 0; i < n; i++){ d[i] = N(rng) - (i % 7 == 0 ? 3000.0 : 0.0); if(i < n - 1) e[i] = 100.0 * N(rng); }
        std::vector<double> d0 = d, e0 = e, z;
        if(!atsat::tridiag_eigen(n, d, e, z)) return 2;
        for(int c = 0; c < n; c++){
            for(int r = 0; r < n; r++){
                double t = d0[r] * z[r*n+c] + (r > 0 ? e0[r-1] * z[(r-1)*n+c] : 0) + (r < n-1 ? e0[r] * z[(r+1)*n+c] : 0);
                worst = std::max(worst, std::fabs(t - d[c] * z[r*n+c]) / 3000.0);
            }
            for(int c2 = 0; c2 < n; c2++){ double s = 0; for(int r = 0; r < n; r++) s += z[r*n+c] * z[r*n+c2];
                worst = std::max(worst, std::fabs(s - (c == c2 ? 1.0 : 0.0))); }
        }
    }
    printf("%.3e\n", worst);
    return worst < 1e-12 ? 0 : 1;
}
''')
    exe = tmp_path / "eig"
    inc = os.path.join(os.path.dirname(here), "atsat_amd", "csrc")
    subprocess.run(["g++", "-O2", "-I" + inc, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_pressure_operator_restatement_closed_forms(defaults):
    """tests/poisson_util.py (the reference's pressure operator, PressureSolver.h:350-353, 476-491,
    restated for the converged-solve tests) against what the mathematics fixes: a ring harmonic is
    an eigenfunction of the phi part, a linear radial profile has no radial second difference, and
    w = y(i) y(j) is the left null vector of the operator closed by the boundary pass."""
    import math
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import poisson_util as pu
    _, c, _ = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    num1, num2, num3 = pu.geometry(g, c)
    n = g.km - 1
    kk = np.arange(g.km)
    m = 7
    ring = np.cos(2 * np.pi * m * kk / n)
    p = np.broadcast_to(ring, (g.im, g.jm, g.km)).copy()
    L = pu.interior(pu.apply(p, g, c), g)
    lam = 2.0 - 2.0 * math.cos(2 * math.pi * m / n)
    want = -lam * num3[1:g.im - 1, 1:g.jm - 1, None] * ring[None, None, :n]
    assert np.allclose(L, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    lin = np.broadcast_to(np.arange(g.im, dtype=float)[:, None, None], (g.im, g.jm, g.km)).copy()
    assert np.abs(pu.interior(pu.apply(lin, g, c), g)).max() <= 1e-9 * num1
    rng = np.random.default_rng(3)
    q = pu.boundary_pass(rng.standard_normal((g.im, g.jm, g.km)))
    Lq = pu.interior(pu.apply(q, g, c), g)
    assert abs(np.sum(pu.weights(g) * Lq)) <= 1e-12 * np.sum(np.abs(pu.weights(g) * Lq))


def _thickness_m(g, c):
    """layer_thickness_m (cSaturnModel.h:689-692) of the float layer heights i * float(L_atm dr)."""
    h = np.float32(c.L_atm * g.dr)
    z = np.array([np.float32(i) * h for i in range(g.im)], dtype=np.float32)
    dz = np.zeros(g.im)
    dz[:g.im - 1] = (z[1:] - z[:-1]).astype(np.float64) * 1.0e3
    return dz


def test_budgets_of_uniform_and_layered_fields(defaults):
    """The budget sums (cSaturnModel.cpp:656-666, 776-787): a temperature uniform on a level has that
    value (times t_ref) as its level mean; a gas uniform on every level has the column
    1e3 sum_i dz_i q_i g/m2 whatever its split into vapour, cloud and ice, and the plane i = im-1
    (no thickness) does not count. A sum without the sin(colatitude) weight, with the wrong
    thickness or over the wrong levels fails."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    t_lev = 1.0 + 0.01 * np.arange(g.im)
    st["t"][...] = t_lev[:, None, None]
    q = 1e-3 * (1.0 + np.arange(g.im))
    st["h2o"][...] = 0.5 * q[:, None, None]
    st["h2o_cloud"][...] = 0.3 * q[:, None, None]
    st["h2o_ice"][...] = 0.2 * q[:, None, None]
    st["nh3_ice"][...] = 2e-4
    st["ch4"][g.im - 1] = 1.0                            # only the lid: no thickness
    cols, lev = oracle.Oracle(g, c, k).budgets(st)
    assert np.allclose(lev, t_lev * c.t_ref, rtol=1e-13)
    dz = _thickness_m(g, c)
    assert dz[g.im - 1] == 0.0 and np.all(dz[:g.im - 1] > 0.0)
    assert math.isclose(cols[0], 1e3 * float(np.sum(dz * q)), rel_tol=1e-13)
    assert math.isclose(cols[1], 1e3 * 2e-4 * float(np.sum(dz)), rel_tol=1e-13)
    assert cols[2] == 0.0


def test_budgets_weight_by_colatitude(defaults):
    """A temperature that varies with the row only: its level mean is the sin(colatitude)-weighted
    mean of the rows (tb_lev weights each row by sin(the.z[j]) over sum_j sin(the.z[j]) km); a
    temperature on the polar rows alone (sin = 0 there) does not count."""
    _, c, k = defaults
    g = _grid()
    st = abi.HostState(g)
    st["t"][...] = 1.0
    st["t"][:, 0, :] = 5.0                               # the north pole row: weight sin(0) = 0
    cols, lev = oracle.Oracle(g, c, k).budgets(st)
    assert np.allclose(lev, c.t_ref, rtol=1e-13)
    the = np.array([j * (g.dthe_deg * math.pi / 180.0) for j in range(g.jm)])
    w = np.sin(the)
    st["t"][...] = (1.0 + 0.1 * np.cos(the))[None, :, None]
    cols, lev = oracle.Oracle(g, c, k).budgets(st)
    want = float(np.sum(w * (1.0 + 0.1 * np.cos(the)))) / float(np.sum(w)) * c.t_ref
    assert np.allclose(lev, want, rtol=1e-12)
