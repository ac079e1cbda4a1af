"""The CPU oracle's moist convective adjustment (oracle/atsat_oracle.cpp,
MoistConvectiveAdjustment<cSaturnModel>::run of planet/ConvectiveAdjustment.h:402-849 with the
MCA sync of planet/cSaturnModel.cpp:1147-1171) against the reference's own golden vectors
(tests/golden/mca_summary.json, written by tests/golden/make_golden.py from the compiled
reference), against the reference harness where it is built, and against what the mathematics
fixes: the column integrals of h_L and q_t are conserved, a dry superadiabatic column is mixed to
the dry adiabat (c_p T + g z uniform over each mixed segment), and a stable column is left as it
is."""
import json
import os

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import summary, HOT_FIELDS, stress_state, stress_knobs, unstable_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = abi.MCA_OUT + [n + "n" for n in abi.MCA_OUT]


def _gold():
    with open(os.path.join(GOLD, "mca_summary.json")) as f:
        return json.load(f)


def _mca_knobs(k):
    km = abi.Knobs.from_buffer_copy(k)
    km.mca_stride = 1
    return km


def _check(st, ref, names, what):
    bad = [n for n in names if summary(st[n])["digest"] != ref[n]["digest"]]
    assert not bad, "%s: not bit-identical to the reference: %s" % (what, bad)


def test_mca_bit_identical_to_golden(defaults):
    g, c, k = defaults
    gold = _gold()["funcs"]
    O = oracle.Oracle(g, c, _mca_knobs(k))
    for scen, base in (("stress", stress_state(g, c)), ("unstable", unstable_state(g, c))):
        s = base.copy()
        O.call("mca", s)
        _check(s, gold[scen], NAMES, scen)
        changed = int(np.count_nonzero(s["t"] != base["t"]))
        assert changed > 10000, (scen, changed)


def test_steps_with_mca_bit_identical_to_golden(defaults):
    g, c, k = defaults
    ks = stress_knobs(k)
    ks.turb_closure = 1
    ks.chem_block = 1
    ks.mca_stride = 1
    gold = _gold()["steps"]
    st = stress_state(g, c)
    O = oracle.Oracle(g, c, ks)
    names = sorted(set(HOT_FIELDS + abi.TURB_OUT + abi.CHEM_OUT + abi.MCA_OUT))
    for it in range(1, 4):
        O.step(st, it)
        _check(st, gold[str(it)], names, "stress iteration %d, chemistry, closure and MCA" % it)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_mca_against_reference_harness(defaults):
    """A state the golden files do not cover: another superadiabatic layer with seeded NH4SH."""
    g, c, k = defaults
    st = unstable_state(g, c, seed=23)
    rng = np.random.default_rng(4)
    fr = st["rho_mix"] / c.r_mix
    st["nh4sh"][...] = 1e-4 * fr * np.abs(rng.standard_normal(fr.shape))
    st["h2s"][...] = 0.01 * fr
    km = _mca_knobs(k)
    so, sr = st.copy(), st.copy()
    oracle.Oracle(g, c, km).call("mca", so)
    R = oracle.Reference(km)
    R.load(sr)
    R.call("mca")
    R.store(sr)
    R.close()
    for n in NAMES:
        assert np.array_equal(so[n], sr[n]), n


def _column_weights(g, c, rho):
    z = (np.arange(g.im) * np.float32(c.L_atm * g.dr)).astype(np.float32).astype(np.float64) * 1e3
    i0, i1 = 1, g.im - 2
    w = np.zeros((g.im,) + rho.shape[1:])
    for i in range(i0, i1 + 1):
        lo = 0.5 * (z[i - 1] + z[i]) if i > i0 else z[i0]
        hi = 0.5 * (z[i] + z[i + 1]) if i < i1 else z[i1]
        w[i] = rho[i] * (hi - lo)
    return z, w


def test_mca_conserves_column_integrals(defaults):
    """Per column, sum w q_t (every gas) and sum w q_tracer are unchanged, and so is sum w h_L,
    h_L = c_p T + g z - sum (lv q_liq + ls q_ice) - L_NH4SH q_NH4SH, w = rho dz, except in
    columns where an adjusted cell ends at a gas's t_0: the latent heat jumps there from ls to lv,
    the regula falsi bracket (ConvectiveAdjustment.h:447-475) closes on the jump, and the
    equilibrium keeps the residual of the jump (about 1 % of the columns here)."""
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    st = unstable_state(g, c)
    s0 = st.copy()
    oracle.Oracle(g, c, _mca_knobs(k)).call("mca", st)
    assert np.count_nonzero(st["t"] != s0["t"]) > 1000
    rho = s0["rho_mix"]
    z, w = _column_weights(g, c, rho)
    L_nh4sh = 4705.0 * np.log(10.0) * 8314.46 / 51.1114
    gases = [("h2o", 1), ("nh3", 2), ("ch4", 0)]

    def hL(s):
        h = c.cp_mix * s["t"] * c.t_ref + c.g * z[:, None, None]
        for name, q in gases:
            h = h - (c.lv[q] * np.maximum(0, s[name + "_cloud"]) + c.ls[q] * np.maximum(0, s[name + "_ice"])) / rho
        return h - L_nh4sh * s["nh4sh"] / rho
    Tn = st["t"] * c.t_ref
    at_t0 = np.zeros(Tn.shape[1:], dtype=bool)
    for q in range(3):
        at_t0 |= np.any((st["t"] != s0["t"]) & (np.abs(Tn - c.t_0[q]) < 1e-6), axis=0)
    assert at_t0.sum() < 0.02 * at_t0.size, at_t0.sum()
    a = np.sum(w * hL(s0), axis=0)[~at_t0]
    b = np.sum(w * hL(st), axis=0)[~at_t0]
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    for f in () + tuple((lambda s, n=name: (np.maximum(0, s[n]) + np.maximum(0, s[n + "_cloud"])
                                               + np.maximum(0, s[n + "_ice"])) / rho) for name, _ in gases) \
            + (lambda s: s["h2s"] / rho, lambda s: s["nh4sh"] / rho):
        a = np.sum(w * f(s0), axis=0)
        b = np.sum(w * f(st), axis=0)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_dry_superadiabatic_column_mixed_to_the_adiabat(defaults):
    """No condensables and no tracers: the Ledoux test reduces to the parcel's temperature,
    T_a - (g / c_p)(z_b - z_a) against T_b. The lowest levels at 1.5 times the dry adiabatic lapse
    are mixed into one segment of uniform h_L = c_p T + g z (the dry adiabat); a parcel from the
    segment's top is not warmer than the level above it, and the cells above it are unchanged."""
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    st = atsat_amd.gen_v1(g, c)
    for n in abi.MCA_OUT[1:]:
        st[n][...] = 0.0
    z = (np.arange(g.im) * np.float32(c.L_atm * g.dr)).astype(np.float32).astype(np.float64) * 1e3
    gam = c.g / c.cp_mix
    T = 400.0 - 0.5 * gam * z                            # stable aloft
    T[:6] = T[6] + 1.5 * gam * (z[6] - z[:6])            # superadiabatic below
    st["t"][...] = (T / c.t_ref)[:, None, None]
    s0 = st.copy()
    oracle.Oracle(g, c, _mca_knobs(k)).call("mca", st)
    for j, kk in ((3, 0), (30, 17), (45, 60), (57, 119)):
        Tn = st["t"][:, j, kk] * c.t_ref
        T0 = s0["t"][:, j, kk] * c.t_ref
        changed = np.nonzero(Tn != T0)[0]
        assert len(changed) >= 6 and changed.min() == 1, (j, kk, changed)
        a, b = changed.min(), changed.max()
        assert np.array_equal(changed, np.arange(a, b + 1)), (j, kk, changed)
        h = c.cp_mix * Tn + c.g * z
        np.testing.assert_allclose(h[a:b + 1], h[a:b + 1].mean(), rtol=1e-13)
        assert (h[b] - c.g * z[b + 1]) / c.cp_mix <= Tn[b + 1] + 1e-7
        assert np.array_equal(Tn[b + 1:], T0[b + 1:])
