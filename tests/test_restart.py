"""Restart files in the reference's binary format (cSaturnModel::save_state / load_state,
planet/FileIO_Sat.cpp:513-660) through the host entry points atsat_restart_write_host /
atsat_restart_read_host (no device).

Pinned against the reference's own save_state: the header and the digest of every array of the
SAT3 file it writes after two stress iterations with the chemistry, the closure and the
adjustment (tests/golden/restart_summary.json). The library writes the same state -- the
oracle's, bit-identical to the reference's over those iterations -- with the same bytes in every
array it holds; p_dynn is p_dyn (restoreVar, cSaturnModel.cpp:1566-1569); the chemistry's cell
fluxes j_* and jT_*, which the library does not hold, are zeros unless passed. Where the
reference is built, the tests compare against its file directly, with the fluxes passed the
files are byte-identical, and the reference resumed from the library's file reproduces the
reference resumed from its own bit for bit. Reading: the reference's file loads into a host
state equal to the reference's; SAT1 and SAT2 files leave the arrays they do not hold as they
were; a wrong magic, grid or a truncated file is an ATSAT_E_IO error."""
import json
import os
import shutil
import tempfile

import numpy as np
import pytest

import atsat_amd
import oracle
from atsat_amd import abi
from util import digest, stress_state, stress_knobs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FLUXES = ["j_h2s", "j_nh3", "j_nh4sh", "jT_h2s", "jT_nh3", "jT_nh4sh"]


def _gold():
    with open(os.path.join(GOLD, "restart_summary.json")) as f:
        return json.load(f)


def _knobs(k):
    ks = stress_knobs(k)
    ks.turb_closure = 1
    ks.chem_block = 1
    ks.mca_stride = 1
    return ks


@pytest.fixture(scope="module")
def stressed(defaults):
    """The stress state after iterations 1-2 with the chemistry, the closure and the adjustment
    (the oracle), and a temporary directory."""
    g, c, k = defaults
    st = stress_state(g, c)
    O = oracle.Oracle(g, c, _knobs(k))
    for it in (1, 2):
        O.step(st, it)
    d = tempfile.mkdtemp()
    yield g, c, k, st, d
    shutil.rmtree(d)


def _sections(path, g, names):
    n = g.im * g.jm * g.km
    out = {}
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(20), dtype=np.int32).tolist()
        for name in names:
            out[name] = np.frombuffer(f.read(8 * n), dtype=np.float64)
        rest = f.read()
    return hdr, out, rest


def test_sat3_file_matches_reference_golden(stressed):
    g, c, k, st, d = stressed
    gold = _gold()
    path = os.path.join(d, "ours.bin")
    atsat_amd.restart_write(path, g, 2, st)
    hdr, arr, rest = _sections(path, g, gold["order"])
    os.remove(path)
    assert hdr == gold["header"] and rest == b""
    bad = [n for n in gold["order"] if n not in FLUXES and digest(arr[n]) != gold["arrays"][n]]
    assert not bad, bad
    assert all(not arr[n].any() for n in FLUXES)


def test_sat1_layout(stressed):
    g, c, k, st, d = stressed
    gold = _gold()
    path = os.path.join(d, "sat1.bin")
    atsat_amd.restart_write(path, g, 7, st, fmt=1)
    names = gold["order"][:38]
    hdr, arr, rest = _sections(path, g, names)
    os.remove(path)
    assert hdr == [0x53415431, g.im, g.jm, g.km, 7] and rest == b""
    assert all(digest(arr[n]) == gold["arrays"][n] for n in names)


@pytest.fixture(scope="module")
def small():
    """A seeded state on a small shell (the format does not depend on the grid) and a directory."""
    g = atsat_amd.small_grid(21, 61, 121)
    st = abi.HostState(g)
    rng = np.random.default_rng(3)
    for a in st.f.values():
        a[...] = rng.standard_normal(a.shape)
    d = tempfile.mkdtemp()
    yield g, st, d
    shutil.rmtree(d)


def _fresh(g, sentinel=-7.0):
    h = abi.HostState(g)
    for a in h.f.values():
        a[...] = sentinel
    return h


def test_round_trip_sat3(small):
    g, st, d = small
    path = os.path.join(d, "rt.bin")
    x = {n: np.full(abi.grid_shape(g), 0.5 + i) for i, n in enumerate(abi.RX_NAMES)}
    atsat_amd.restart_write(path, g, 12, st, extra=x)
    h = _fresh(g)
    xr = {n: np.zeros(abi.grid_shape(g)) for n in abi.RX_NAMES}
    it, fmt = atsat_amd.restart_read(path, g, h, extra=xr)
    assert (it, fmt) == (12, 3)
    gold = _gold()
    held = {n for n in gold["order"] if n in abi.FIELD}
    for name in abi.FIELD_NAMES:
        if name in held:
            assert np.array_equal(h[name], st[name]), name
        else:
            assert (h[name] == -7.0).all(), name                    # not in the file: untouched
    for n in abi.RX_NAMES:
        assert np.array_equal(xr[n], x[n]), n


def test_sat1_and_sat2_leave_the_rest(small):
    g, st, d = small
    gold = _gold()
    n = g.im * g.jm * g.km
    p3 = os.path.join(d, "s3.bin")
    atsat_amd.restart_write(p3, g, 4, st)
    raw = open(p3, "rb").read()
    os.remove(p3)
    # SAT2: the SAT3 file without its last array (S_precip_ch4), magic "SAT2"
    p2 = os.path.join(d, "s2.bin")
    with open(p2, "wb") as f:
        f.write(np.int32(0x53415432).tobytes() + raw[4:len(raw) - 8 * n])
    h = _fresh(g)
    assert atsat_amd.restart_read(p2, g, h) == (4, 2)
    assert (h["S_precip_ch4"] == -7.0).all() and np.array_equal(h["S_precip_ch4_ice"], st["S_precip_ch4_ice"])
    os.remove(p2)
    p1 = os.path.join(d, "s1.bin")
    atsat_amd.restart_write(p1, g, 4, st, fmt=1)
    h = _fresh(g)
    assert atsat_amd.restart_read(p1, g, h) == (4, 1)
    assert np.array_equal(h["nue"], st["nue"]) and (h["rho_mix"] == -7.0).all() and (h["tke_source"] == -7.0).all()
    os.remove(p1)


def test_bad_files_are_errors(small):
    g, st, d = small
    p = os.path.join(d, "bad.bin")
    atsat_amd.restart_write(p, g, 3, st, fmt=1)
    raw = open(p, "rb").read()
    h = _fresh(g)
    with pytest.raises(atsat_amd.AtsatError, match="no file"):
        atsat_amd.restart_read(os.path.join(d, "missing.bin"), g, h)
    for name, data, msg in (("magic", b"JUP1" + raw[4:], "bad header"),
                            ("grid", raw[:4] + np.int32(40).tobytes() + raw[8:], "grid mismatch"),
                            ("short", raw[:-100], "truncated"),
                            ("header", raw[:12], "bad header")):
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(atsat_amd.AtsatError, match=msg):
            atsat_amd.restart_read(p, g, _fresh(g))
    os.remove(p)
    with pytest.raises(atsat_amd.AtsatError, match="format"):
        atsat_amd.restart_write(p, g, 3, st, fmt=2)
    small = abi.HostState(g, fields=["t"], vecs=[])
    with pytest.raises(atsat_amd.AtsatError, match="missing"):
        atsat_amd.restart_write(p, g, 3, small)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_against_reference_file_and_resume(stressed):
    """The reference writes its file; ours equals it with the fluxes passed; our reader loads
    it exactly; the reference resumed from our file (fluxes zero) reproduces the straight
    reference run over the next iteration, an odd one, which reads the arrays the last even one
    carried (FileIO_Sat.cpp:486-497)."""
    g, c, k, st, d = stressed
    ks = _knobs(k)
    R = oracle.Reference(ks)
    R.load(stress_state(g, c))
    for it in (1, 2):
        R.call("step", it)
    da, db = os.path.join(d, "a"), os.path.join(d, "b")
    os.makedirs(da)
    os.makedirs(db)
    fa = R.save_state(da, 2)
    S = st.copy()
    R.store(S)
    shp = abi.grid_shape(g)
    x = {n: R.extra(n, shp) for n in abi.RX_NAMES}
    for n in S.f:
        assert np.array_equal(S[n], st[n]), n                        # the oracle is the reference here
    fx = os.path.join(d, "with_fluxes.bin")
    atsat_amd.restart_write(fx, g, 2, S, extra=x)
    with open(fx, "rb") as f1, open(fa, "rb") as f2:
        while True:
            b1, b2 = f1.read(1 << 26), f2.read(1 << 26)
            assert b1 == b2
            if not b1:
                break
    os.remove(fx)
    h = _fresh(g)
    xr = {n: np.zeros(shp) for n in abi.RX_NAMES}
    assert atsat_amd.restart_read(fa, g, h, extra=xr) == (2, 3)
    for n in _gold()["order"]:
        if n in abi.FIELD:
            assert np.array_equal(h[n], S[n]), n
        else:
            assert np.array_equal(xr[n], x[n]), n
    del h, xr
    R.close()
    fb = os.path.join(db, "sat_restart_2.bin")
    atsat_amd.restart_write(fb, g, 2, S)
    for dd, out in ((da, S), (db, st.copy())):
        Rr = oracle.Reference(ks)
        assert Rr.load_state(dd, 2)
        Rr.call("step", 3)
        Rr.store(out)
        Rr.close()
        if dd == db:
            bad = [n for n in S.f if not np.array_equal(S[n], out[n])]
            assert not bad, bad
