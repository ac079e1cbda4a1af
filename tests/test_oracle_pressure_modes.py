"""The CPU oracle's optional modes of the red-black pressure solver (oracle/atsat_oracle.cpp:
press_line_pass, press_baro_pass, press_interior_mean; PressureSolver.h:290-315, 497-708,
750-755): the line pass (ATSAT_PRESS_LINE), the barotropic correction (ATSAT_PRESS_BAROTROPIC)
and both.

The reference sums the domain mean that the line pass holds with an OpenMP reduction of no fixed
order; the oracle (and the device) sum every (i, j) row, then the rows in order. With
ATSAT_ORACLE_SEQ_MEAN=1 the oracle sums in (i, j, k) order, the reference's order on one thread,
and is then bit-identical to the reference's own solve on one thread
(tests/golden/pmodes_summary.json, and the harness where it is built); in its default order it
agrees to 1e-10 relative. The mathematics fixes: the line pass holds the interior mean of p_dyn;
the barotropic correction adds one value per column to every level, and its zonal mean has a
zero sin(theta)-weighted mean over the rows."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from atsat_amd import abi
from util import summary, stress_state, stress_knobs, rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = [(0, 1), (1, 0), (1, 1)]


def _gold():
    with open(os.path.join(GOLD, "pmodes_summary.json")) as f:
        return json.load(f)


def _knobs(k, line, baro, sweeps=3):
    ks = stress_knobs(k)
    ks.press_line = line
    ks.press_baro = baro
    ks.press_sweeps = sweeps
    return ks


class _SeqMean:
    """ATSAT_ORACLE_SEQ_MEAN=1 and one OpenMP thread for the block (the oracle's mean in the
    reference's one-thread order)."""
    def __enter__(self):
        self.gomp = ctypes.CDLL("libgomp.so.1")
        self.n = self.gomp.omp_get_max_threads()
        self.gomp.omp_set_num_threads(1)
        os.environ["ATSAT_ORACLE_SEQ_MEAN"] = "1"

    def __exit__(self, *a):
        del os.environ["ATSAT_ORACLE_SEQ_MEAN"]
        self.gomp.omp_set_num_threads(self.n)


@pytest.mark.parametrize("line,baro", MODES)
def test_modes_against_golden(defaults, line, baro):
    g, c, k = defaults
    ref = _gold()["line%d_baro%d" % (line, baro)]["p_dyn"]
    st = stress_state(g, c)
    s1 = st.copy()
    with _SeqMean():
        oracle.Oracle(g, c, _knobs(k, line, baro)).call("pressure_solve", s1)
    assert summary(s1["p_dyn"])["digest"] == ref["digest"], (line, baro)
    s2 = st.copy()
    oracle.Oracle(g, c, _knobs(k, line, baro)).call("pressure_solve", s2)
    if not line:
        assert np.array_equal(s2["p_dyn"], s1["p_dyn"])             # no mean involved
    assert rel_l2(s2["p_dyn"], s1["p_dyn"]) <= 1e-10
    np.testing.assert_allclose(summary(s2["p_dyn"])["l2"], ref["l2"], rtol=1e-10)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_modes_against_reference_harness(defaults):
    """Another state: a perturbed gen_v1 member with small velocity tendencies."""
    import atsat_amd
    g, c, k = defaults
    st = atsat_amd.gen_v1(g, c, seed=5, member=2, t_sigma=0.5, w_sigma=2.0 / 470.0)
    rng = np.random.default_rng(9)
    for n in ("rhs_u", "rhs_v", "rhs_w", "aux_u", "aux_v", "aux_w"):
        st[n][...] = 1e-3 * rng.standard_normal(st[n].shape)
    st["p_dyn"][...] = 0.1 * rng.standard_normal(st["p_dyn"].shape)
    with _SeqMean():
        for line, baro in MODES:
            ks = _knobs(k, line, baro, sweeps=2)
            so, sr = st.copy(), st.copy()
            oracle.Oracle(g, c, ks).call("pressure_solve", so)
            R = oracle.Reference(ks)
            R.load(sr)
            R.call("pressure_solve")
            R.store(sr)
            R.close()
            assert np.array_equal(so["p_dyn"], sr["p_dyn"]), (line, baro)


def _interior_mean(p):
    return p[1:-1, 1:-1, :-1].mean()


def test_line_pass_holds_the_mean(defaults):
    """p_dyn with a seeded perturbation, which the sweeps smooth: the plain sweeps move its
    interior mean, the line pass holds it (a uniform shift after every sweep)."""
    g, c, k = defaults
    st = stress_state(g, c)
    st["p_dyn"][...] += 0.05 * np.random.default_rng(2).standard_normal(st["p_dyn"].shape)
    p0 = _interior_mean(st["p_dyn"])
    tol = 1e-13 * np.abs(st["p_dyn"]).max()
    for sweeps in (1, 4):
        s = st.copy()
        oracle.Oracle(g, c, _knobs(k, 1, 0, sweeps)).call("pressure_solve", s)
        assert abs(_interior_mean(s["p_dyn"]) - p0) <= tol, sweeps
    s = st.copy()
    oracle.Oracle(g, c, _knobs(k, 0, 0, 4)).call("pressure_solve", s)
    assert abs(_interior_mean(s["p_dyn"]) - p0) > 100 * tol


def test_barotropic_correction_is_column_uniform(defaults):
    """One sweep with and without the correction: the difference is one value per column on every
    level i of the interior rows (the boundary pass keeps a column-uniform shift, c43 - c13 = 1),
    and its zonal mean has a zero sin-weighted mean over the rows (mode 0, PressureSolver.h:687-690)."""
    g, c, k = defaults
    st = stress_state(g, c)
    a, b = st.copy(), st.copy()
    oracle.Oracle(g, c, _knobs(k, 0, 1, 1)).call("pressure_solve", a)
    oracle.Oracle(g, c, _knobs(k, 0, 0, 1)).call("pressure_solve", b)
    d = (a["p_dyn"] - b["p_dyn"])[:, 1:-1, :-1]
    scale = np.abs(d).max()
    assert scale > 1e-3
    assert np.abs(d - d[20][None]).max() <= 1e-9 * scale
    q0 = d[20].mean(axis=1)                                          # the zonal mean per row
    the = np.arange(g.jm) * np.deg2rad(g.dthe_deg)
    s = np.abs(np.sin(the[1:-1]))
    assert abs(np.sum(s * q0)) <= 1e-9 * np.sum(s * np.abs(q0))
