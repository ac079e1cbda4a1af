"""The CPU oracle against the reference's own golden vectors (tests/golden, made by
tests/golden/make_golden.py from the reference's compiled functions).

Bit identity is checked through a blake2b digest of every hot-path field: the restatement
reproduces the reference's arithmetic operation for operation (oracle/atsat_oracle.cpp)."""
import json
import os

import numpy as np
import pytest

import atsat_amd
import oracle
from util import summary, HOT_FIELDS, stress_state, stress_knobs, cold_base_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _check(st, ref, names, what):
    bad = []
    for n in names:
        s = summary(st[n])
        if s["digest"] != ref[n]["digest"]:
            bad.append((n, s["l2"], ref[n]["l2"]))
    assert not bad, "%s: fields not bit-identical to the reference: %s" % (what, bad[:5])


@pytest.fixture(scope="module")
def native(defaults):
    g, c, k = defaults
    return g, c, k, atsat_amd.gen_v1(g, c, seed=20261203)


def test_functions_bit_identical(native):
    g, c, k, st0 = native
    gold = _load("func_summary.json")["funcs"]
    O = oracle.Oracle(g, c, k)
    for key, ref in gold.items():
        st = st0.copy()
        if key.startswith("saturation_adjust"):
            O.call("saturation_adjust", st, int(key.rsplit("_", 1)[1]))
        else:
            O.call(key, st)
        _check(st, ref, HOT_FIELDS, key)


def test_steps_bit_identical(native):
    g, c, k, st0 = native
    gold = _load("step_summary.json")
    assert gold["grid"] == [g.im, g.jm, g.km]
    O = oracle.Oracle(g, c, k)
    st = st0.copy()
    blk = np.load(os.path.join(GOLD, "step2_block.npz"))
    for it in range(1, 5):
        O.step(st, it)
        _check(st, gold["steps"][str(it)], HOT_FIELDS, "iteration %d" % it)
        if it == 2:
            for n in HOT_FIELDS:
                assert np.array_equal(st[n][:, [3, 90], :90], blk[n]), n


def test_stress_scenario_bit_identical(defaults):
    g, c, k = defaults
    ks = stress_knobs(k)
    gold = _load("stress_summary.json")["steps"]
    st = stress_state(g, c)
    O = oracle.Oracle(g, c, ks)
    for it in range(1, 3):
        O.step(st, it)
        _check(st, gold[str(it)], HOT_FIELDS + ["tke", "dis"], "stress iteration %d" % it)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_against_reference_harness_small_perturbation(defaults):
    """Directly against the reference on a state the golden files do not cover: the
    gen_v1 state of another member with T and zonal-wind perturbations."""
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c, seed=99, member=3, t_sigma=0.5, w_sigma=2.0 / 470.0)
    so, sr = st.copy(), st.copy()
    R = oracle.Reference(k)
    R.load(sr)
    O = oracle.Oracle(g, c, k)
    for it in (2, 3):
        O.step(so, it)
        R.call("step", it)
    R.store(sr)
    R.close()
    for n in HOT_FIELDS:
        assert np.array_equal(so[n], sr[n]), n


SRF = ["precip_srf_h2o", "precip_srf_nh3", "precip_srf_ch4", "precip_srf_nh4sh", "precip_srf_total"]


def _srf_check(st, ref, what):
    bad = [n for n in SRF if summary(st.v[n])["digest"] != ref[n]["digest"]]
    assert not bad, "%s: surface maps not bit-identical to the reference: %s" % (what, bad)


def test_surface_maps_bit_identical(defaults):
    """Precipitation::surfaceMap (Precipitation.h:467-536) after the precipitation alone and after
    whole iterations, against the reference's own maps (srf_summary.json). The cold-base state puts
    every species' precipitation on the base, so none of the five maps is identically zero there."""
    g, c, k = defaults
    ks = stress_knobs(k)
    gold = _load("srf_summary.json")
    for scen, mk in (("stress", stress_state), ("cold_base", cold_base_state)):
        st = mk(g, c)
        oracle.Oracle(g, c, ks).call("precipitation", st)
        ref = gold["funcs"][scen + "_precipitation"]
        _srf_check(st, ref, scen)
        _check(st, ref, [n for n in ref if n not in SRF], scen + " precipitation fields")
        if scen == "cold_base":
            assert all(ref[n]["l2"] > 0.0 for n in SRF)
            # the maps are the base level's fluxes: i_base = 0 on this grid (no topography)
            base = {"h2o": ("P_rain", "P_snow", "P_graupel"), "nh3": ("P_nh3_rain", "P_nh3_snow", "P_nh3_graupel"),
                    "ch4": ("P_ch4_rain", "P_ch4_snow", "P_ch4_graupel"), "nh4sh": ("P_nh4sh",)}
            tot = np.zeros(g.jm * g.km)
            for sp, names in base.items():
                m = np.zeros(g.jm * g.km)
                for n in names:
                    m = m + st[n][0].ravel()
                assert np.array_equal(st.v["precip_srf_" + sp], m), sp
                tot = tot + m
            assert np.allclose(st.v["precip_srf_total"], tot, rtol=1e-15, atol=0.0)
    O = oracle.Oracle(g, c, ks)
    st = cold_base_state(g, c)
    for it in (1, 2):
        O.step(st, it)
    _srf_check(st, gold["steps"]["cold_base_2"], "cold-base iteration 2")
