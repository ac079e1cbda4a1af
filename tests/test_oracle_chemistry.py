"""The CPU oracle's chemistry block (oracle/atsat_oracle.cpp: ChemistrySat::EquilibriumNH4SH,
DiffMassFluxSat, ChemMassRateSat, FluxLimiter<cSaturnModel>::nh4sh and the filters of
planet/cSaturnModel.cpp:979-994) against the reference's own golden vectors
(tests/golden/chem_summary.json, written by tests/golden/make_golden.py from the compiled
reference), against the reference harness where it is built, and against what the mathematics
fixes: conservation of N and S and the equilibrium condition of the NH4SH adjustment, the
spherical Laplacian of r^2 and of cos(theta), the thermal-diffusion term for a radial
temperature gradient, and the limiter on a linear profile."""
import json
import os

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import summary, HOT_FIELDS, stress_state, stress_knobs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold():
    with open(os.path.join(GOLD, "chem_summary.json")) as f:
        return json.load(f)


def _chem_knobs(k):
    kc = abi.Knobs.from_buffer_copy(k)
    kc.chem_block = 1
    return kc


def _check(st, ref, names, what):
    bad = [n for n in names if summary(st[n])["digest"] != ref[n]["digest"]]
    assert not bad, "%s: not bit-identical to the reference: %s" % (what, bad)


def test_chemistry_bit_identical_to_golden(defaults):
    g, c, k = defaults
    gold = _gold()["funcs"]
    O = oracle.Oracle(g, c, _chem_knobs(k))
    for scen, base in (("gen_v1", atsat_amd.gen_v1(g, c, seed=20261203)), ("stress", stress_state(g, c))):
        for call in ("nh4sh_equilibrium", "chemistry"):
            s = base.copy()
            O.call(call, s)
            _check(s, gold["%s_%s" % (scen, call)], abi.CHEM_OUT, "%s %s" % (scen, call))


def test_steps_with_chemistry_and_closure_bit_identical_to_golden(defaults):
    g, c, k = defaults
    ks = stress_knobs(k)
    ks.turb_closure = 1
    ks.chem_block = 1
    gold = _gold()["steps"]
    st = stress_state(g, c)
    O = oracle.Oracle(g, c, ks)
    names = sorted(set(HOT_FIELDS + abi.TURB_OUT + abi.CHEM_OUT))
    for it in range(1, 5):
        O.step(st, it)
        _check(st, gold[str(it)], names, "stress iteration %d, chemistry and closure" % it)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_chemistry_against_reference_harness(defaults):
    """A state the golden files do not cover: a perturbed member with seeded NH3 / H2S / NH4SH."""
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c, seed=99, member=3, t_sigma=0.5, w_sigma=2.0 / 470.0)
    rng = np.random.default_rng(5)
    fr = st["rho_mix"] / c.r_mix
    st["h2s"][...] = 0.02 * fr * np.abs(1.0 + 0.3 * rng.standard_normal(fr.shape))
    st["nh4sh"][...] = 1e-3 * fr * np.abs(rng.standard_normal(fr.shape))
    st["u"][...] = 1e-3 * rng.standard_normal(fr.shape)
    kc = _chem_knobs(k)
    for call in ("nh4sh_equilibrium", "chemistry"):
        so, sr = st.copy(), st.copy()
        oracle.Oracle(g, c, kc).call(call, so)
        R = oracle.Reference(kc)
        R.load(sr)
        R.call(call)
        R.store(sr)
        R.close()
        for n in abi.CHEM_OUT:
            assert np.array_equal(so[n], sr[n]), (call, n)


def _small(defaults):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    return g, c, _chem_knobs(k)


def test_equilibrium_conserves_and_reaches_kp(defaults):
    """Per cell: N (nh3/m_nh3 + nh4sh/m_nh4sh) and S (h2s/m_h2s + nh4sh/m_nh4sh) are conserved,
    the temperature rises by the formation heat L x m_nh4sh / (cp_mix rho), and where NH4SH
    formed the product p_NH3 p_H2S equals K_p(T) = 10^(14.82 - 4705/T) atm^2 at the new T."""
    g, c, kc = _small(defaults)
    st = atsat_amd.gen_v1(g, c)
    rng = np.random.default_rng(11)
    fr = st["rho_mix"] / c.r_mix
    st["h2s"][...] = 0.01 * fr * np.abs(1.0 + 0.3 * rng.standard_normal(fr.shape))
    st["nh4sh"][...] = 1e-4 * fr * np.abs(rng.standard_normal(fr.shape))
    s0 = st.copy()
    oracle.Oracle(g, c, kc).call("nh4sh_equilibrium", st)
    N0 = s0["nh3"] / c.m_nh3 + s0["nh4sh"] / c.m_nh4sh
    N1 = st["nh3"] / c.m_nh3 + st["nh4sh"] / c.m_nh4sh
    S0 = s0["h2s"] / c.m_h2s + s0["nh4sh"] / c.m_nh4sh
    S1 = st["h2s"] / c.m_h2s + st["nh4sh"] / c.m_nh4sh
    np.testing.assert_allclose(N1, N0, rtol=1e-13, atol=1e-25)
    np.testing.assert_allclose(S1, S0, rtol=1e-13, atol=1e-25)
    x = (st["nh4sh"] - s0["nh4sh"]) / c.m_nh4sh
    L_f = 4705.0 * np.log(10.0) * 8314.46 / 51.1114
    dT = (st["t"] - s0["t"]) * c.t_ref
    np.testing.assert_allclose(dT, L_f * c.m_nh4sh / (c.cp_mix * st["rho_mix"]) * x, rtol=1e-9, atol=1e-12)
    # the amount formed (> 0) or decomposed (< 0) is the root of
    # f(x) = (a - x)(b - x)(R T(x))^2 - K_p(T(x)), T(x) = T0 + heat x, to the bisection's
    # absolute tolerance 1e-15 max(a, b, s) (ChemistrySat.h:272-277): the root found
    # independently by Brent's method
    from scipy.optimize import brentq
    moved = np.argwhere(x != 0)
    assert len(moved) > 1000
    rng2 = np.random.default_rng(1)
    for idx in moved[rng2.choice(len(moved), 300, replace=False)]:
        i, j, k = idx
        T0 = s0["t"][i, j, k] * c.t_ref
        rho = s0["rho_mix"][i, j, k]
        a_ = max(0.0, s0["nh3"][i, j, k]) / c.m_nh3
        b_ = max(0.0, s0["h2s"][i, j, k]) / c.m_h2s
        s_ = max(0.0, s0["nh4sh"][i, j, k]) / c.m_nh4sh
        heat = L_f * c.m_nh4sh / (c.cp_mix * rho)

        def f(xx):
            TT = T0 + heat * xx
            return (a_ - xx) * (b_ - xx) * (8314.46 * TT) ** 2 - 10.0 ** (14.82 - 4705.0 / TT) * 101325.0 ** 2
        if f(0.0) > 0.0:
            root = brentq(f, 0.0, min(a_, b_), xtol=1e-300, rtol=1e-15, maxiter=500)
        elif f(-s_) <= 0.0:
            root = -s_
        else:
            root = brentq(f, -s_, 0.0, xtol=1e-300, rtol=1e-15, maxiter=500)
        assert abs(x[i, j, k] - root) <= 2e-15 * max(a_, b_, s_) + 1e-13 * abs(root), (idx, x[i, j, k], root)


def _zero_chem(st):
    for n in ("h2s", "nh3", "nh4sh", "u", "v", "w"):
        st[n][...] = 0.0
    return st


def test_laplacian_closed_forms(defaults):
    """difflux = D lap(c): c = rm^2 gives D (2 + 2/rm 2 rm) = 6 D exactly for the central
    differences; c = cos(theta) gives -2 cos(theta) / rm^2 to O(dtheta^2) where
    sin(theta) >= 0.4 (the polar floor, ChemistrySat.h:628-644)."""
    g, c, kc = _small(defaults)
    st = _zero_chem(atsat_amd.gen_v1(g, c))
    st["t"][...] = 1.0
    rm = 1.0 + np.arange(g.im) * g.dr + (c.metric_radius_km / c.L_atm - 1.0)
    the = np.arange(g.jm) * np.deg2rad(g.dthe_deg)
    st["nh3"][...] = (rm ** 2)[:, None, None]
    st["h2s"][...] = np.cos(the)[None, :, None]
    oracle.Oracle(g, c, kc).call("chemistry", st)
    # the chemistry filters difflux after forming it: a 1-2-1 pass leaves a constant unchanged
    # and moves the smooth cos(theta) profile by -dtheta^2 / 4 relative; the rows next to the
    # boundaries, which hold no difflux, are left out
    D_nh3 = c.mue_nh3 / (c.rg_nh3 * c.sc_nh3)
    D_h2s = c.mue_h2s / (c.rg_h2s * c.sc_h2s)
    d = st["difflux_nh3"][3:-3, 5:-5, :-1] / D_nh3
    np.testing.assert_allclose(d, 6.0, rtol=1e-8)
    sel = np.sin(the[5:-5]) >= 0.45
    lap_h2s = st["difflux_h2s"][3:-3, 5:-5, :-1] / D_h2s
    expect = (-2.0 * np.cos(the[5:-5])[None, :] / (rm[3:-3, None] ** 2))
    got = lap_h2s[:, sel, 7]
    np.testing.assert_allclose(got, expect[:, sel], rtol=2e-3, atol=1e-9 * np.abs(expect).max())


def test_flux_limiter_zero_on_linear_profile(defaults):
    """A linear nh4sh profile in r has equal forward and backward differences: r = 1, Superbee
    phi(1) = 1, no correction, whatever the wind."""
    g, c, kc = _small(defaults)
    st = _zero_chem(atsat_amd.gen_v1(g, c))
    st["nh4sh"][...] = (1e-3 * (1.0 + np.arange(g.im)))[:, None, None]
    rng = np.random.default_rng(2)
    st["u"][...] = 1e-2 * rng.standard_normal(st["u"].shape)
    st["v"][...] = 1e-2 * rng.standard_normal(st["v"].shape)
    st["w"][...] = 1e-2 * rng.standard_normal(st["w"].shape)
    oracle.Oracle(g, c, kc).call("chemistry", st)
    assert np.max(np.abs(st["fluxlim_nh4sh"][3:-3, 3:-3, :])) < 1e-18


def test_thermal_diffusion_closed_form(defaults):
    """Uniform species and t = 1 + beta (rm - rm_c): the thermal fluxes jT_x = D_x / t dt/dr make
    j_x = -(c_mix / r_mix) jT_x, so thermalmassflux = -(c_mix / r_mix) (beta^2 / t)
    (D_nh3 cp_nh3 + D_h2s cp_h2s + D_nh4sh cp_nh4sh) (ChemistrySat.h:401-423, 452-459); the
    filters move this smooth radial profile by far less than the tolerance."""
    g, c, kc = _small(defaults)
    st = _zero_chem(atsat_amd.gen_v1(g, c))
    rm = 1.0 + np.arange(g.im) * g.dr + (c.metric_radius_km / c.L_atm - 1.0)
    beta = 0.02
    t = 1.0 + beta * (rm - rm[g.im // 2])
    st["t"][...] = t[:, None, None]
    oracle.Oracle(g, c, kc).call("chemistry", st)
    D = {x: getattr(c, "mue_" + x) / (getattr(c, "rg_" + x) * getattr(c, "sc_" + x)) for x in ("nh3", "h2s", "nh4sh")}
    s = sum(D[x] * getattr(c, "cp_" + x) for x in D)
    expect = -(c.c_mix / c.r_mix) * beta ** 2 / t * s
    got = st["thermalmassflux"][3:-3, 5:-5, 9]
    np.testing.assert_allclose(got, np.broadcast_to(expect[3:-3, None], got.shape), rtol=1e-6)
