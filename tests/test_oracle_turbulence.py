"""The CPU oracle's turbulence closure (oracle/atsat_oracle.cpp, Turbulence<cSaturnModel>::run /
::init of planet/Turbulence.h) against the reference's own golden vectors
(tests/golden/turb_summary.json, written by tests/golden/make_golden.py from the compiled
reference), against the reference harness directly where it is built, and against closed forms
of the production term and the buoyancy sink that a sign or index slip would fail."""
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
    with open(os.path.join(GOLD, "turb_summary.json")) as f:
        return json.load(f)


def _turb_knobs(k, buoy=None):
    kt = abi.Knobs.from_buffer_copy(k)
    kt.turb_active = 1
    if buoy is not None:
        kt.turb_buoy = buoy
    return kt


def _check(st, ref, names, what):
    bad = [n for n in names if summary(st[n])["digest"] != ref[n]["digest"]]
    assert not bad, "%s: not bit-identical to the reference: %s" % (what, bad)


@pytest.fixture(scope="module")
def native(defaults):
    g, c, k = defaults
    return g, c, k, atsat_amd.gen_v1(g, c, seed=20261203)


def test_closure_bit_identical_to_golden(native):
    g, c, k, st0 = native
    gold = _gold()["funcs"]
    O = oracle.Oracle(g, c, _turb_knobs(k))
    names = abi.TURB_OUT + ["vel_star"]
    for scen, base in (("gen_v1", st0), ("stress", stress_state(g, c))):
        for call in ("turbulence", "turbulence_init"):
            s = base.copy()
            O.call(call, s)
            _check(s, gold["%s_%s" % (scen, call)], names, "%s %s" % (scen, call))


def test_steps_with_closure_bit_identical_to_golden(defaults):
    g, c, k = defaults
    ks = stress_knobs(k)
    ks.turb_closure = 1
    gold = _gold()["steps"]
    st = stress_state(g, c)
    O = oracle.Oracle(g, c, ks)
    for it in (1, 2):
        O.step(st, it)
        _check(st, gold[str(it)], HOT_FIELDS + abi.TURB_OUT, "stress iteration %d with the closure" % it)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_closure_against_reference_harness(defaults):
    """A state the golden files do not cover: a perturbed member with a seeded k / omega field."""
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c, seed=99, member=3, t_sigma=0.5, w_sigma=2.0 / 470.0)
    rng = np.random.default_rng(3)
    st["tke"][...] = np.abs(1e-4 * (1.0 + rng.standard_normal(st["tke"].shape)))
    st["dis"][...] = 1e-3 * np.exp(rng.standard_normal(st["dis"].shape))
    st["u"][...] = 1e-3 * rng.standard_normal(st["u"].shape)
    kt = _turb_knobs(k)
    so, sr = st.copy(), st.copy()
    oracle.Oracle(g, c, kt).call("turbulence", so)
    R = oracle.Reference(kt)
    R.load(sr)
    R.call("turbulence")
    R.store(sr)
    R.close()
    for n in abi.TURB_OUT + ["vel_star"]:
        assert np.array_equal(so[n], sr[n]), n


def _small_state(defaults):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    st = atsat_amd.gen_v1(g, c)
    for n in ("u", "v", "w", "tke", "dis", "nue", "prod", "tke_source", "dis_source"):
        st[n][...] = 0.0
    for sp in abi.SPECIES:
        st[sp][...] = 0.0
    return g, c, k, st


def _interior(a):
    return a[1:-1, 1:-1, :-1]


def test_uniform_state_closed_form(defaults):
    """No motion, uniform k and omega, no buoyancy term: no production, nu_t = k / omega (below
    the nu_t cap), tke_source = -beta* k omega and dis_source = -beta_2 omega^2 (F1 = 0 far from
    any wall)."""
    g, c, k, st = _small_state(defaults)
    k0, d0 = 3e-4, 2.0
    st["tke"][...] = k0
    st["dis"][...] = d0
    oracle.Oracle(g, c, _turb_knobs(k, buoy=0)).call("turbulence", st)
    assert np.all(_interior(st["prod"]) == 0.0)
    np.testing.assert_allclose(_interior(st["nue"]), k0 / d0, rtol=1e-15)
    np.testing.assert_allclose(_interior(st["tke_source"]), -0.09 * k0 * d0, rtol=1e-14)
    np.testing.assert_allclose(_interior(st["dis_source"]), -0.0368 * d0 * d0, rtol=1e-14)
    # the zero-gradient copy at i = 0 and the seam copy
    assert np.array_equal(st["nue"][0], st["nue"][1])
    assert np.array_equal(st["tke_source"][:, :, -1], st["tke_source"][:, :, 0])


def test_zonal_shear_production_closed_form(defaults):
    """w = a (rm - rm_c), u = v = 0: the rate-of-strain tensor in the spherical basis has
    S_rp = (a - w / r) / 2 and S_tp = -w cot(theta) / (2 r), so P = nu_t (2 S:S)
    = nu_t [(a - w/r)^2 + (w cot(theta) / r)^2]. A flipped curvature sign gives (a + w/r)^2."""
    g, c, k, st = _small_state(defaults)
    k0, d0, a = 1e-4, 1.0, 0.5
    st["tke"][...] = k0
    st["dis"][...] = d0
    rm = 1.0 + np.arange(g.im) * g.dr + (c.metric_radius_km / c.L_atm - 1.0)
    rc = rm[g.im // 2]
    w = a * (rm - rc)
    st["w"][...] = w[:, None, None]
    oracle.Oracle(g, c, _turb_knobs(k, buoy=0)).call("turbulence", st)
    the = (np.arange(g.jm) * np.deg2rad(g.dthe_deg))[1:-1]
    cot = np.cos(the) / np.sin(the)
    W = w[1:-1, None]
    R = rm[1:-1, None]
    expect = (k0 / d0) * ((a - W / R) ** 2 + (W * cot[None, :] / R) ** 2)
    got = st["prod"][1:-1, 1:-1, 5]
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-18)


def test_dry_buoyancy_sink_closed_form(defaults):
    """A linear T(z) with lapse G and no condensate: the buoyancy term removes
    nu_t N^2 (L/u_0)^2 / Pr_t from tke_source, N^2 = g/T (dT/dz + g/cp_mix), Pr_t = 0.9; the
    moist form (turb_buoy 3) equals the dry one without condensate, and no sink acts where the
    layer is unstable (N^2 < 0, turb_buoy 1)."""
    g, c, k, st = _small_state(defaults)
    k0, d0 = 1e-4, 10.0                           # the sink stays inside the cap 20 k omega
    st["tke"][...] = k0
    st["dis"][...] = d0
    z = np.arange(g.im) * (c.L_atm * g.dr) * 1e3                      # m
    lapse = 0.5e-3                                                   # K/m, stable (< g/cp)
    T = 400.0 - lapse * z                                            # 150 K at the lid
    st["t"][...] = (T / c.t_ref)[:, None, None]
    out = {}
    for buoy in (0, 1, 3):
        s = st.copy()
        oracle.Oracle(g, c, _turb_knobs(k, buoy=buoy)).call("turbulence", s)
        out[buoy] = s
    N2 = c.g / T * (-lapse + c.g / c.cp_mix)
    tL = c.L_atm * 1e3 / c.u_0
    sink = (k0 / d0) * N2 * tL * tL / 0.9
    d = (out[0]["tke_source"] - out[1]["tke_source"])[1:-1, 1:-1, 7]
    np.testing.assert_allclose(d, np.broadcast_to(sink[1:-1, None], d.shape), rtol=1e-6)
    assert np.array_equal(out[1]["tke_source"], out[3]["tke_source"])
    # a superadiabatic layer: no sink with turb_buoy 1
    st["t"][...] = ((1500.0 - 2.0e-3 * z) / c.t_ref)[:, None, None]
    s0, s1 = st.copy(), st.copy()
    oracle.Oracle(g, c, _turb_knobs(k, buoy=0)).call("turbulence", s0)
    oracle.Oracle(g, c, _turb_knobs(k, buoy=1)).call("turbulence", s1)
    assert np.array_equal(_interior(s0["tke_source"]), _interior(s1["tke_source"]))
