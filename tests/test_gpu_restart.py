"""Restart files through the device (atsat_save_restart / atsat_load_restart): the file a context
writes for one member is byte for byte the file the host writer makes of that member's state, and
a context resumed from it continues exactly as the context that wrote it (cSaturnModel::save_state
/ load_state, planet/FileIO_Sat.cpp:572-660, and the resume at restart_from_iter + 1,
cSaturnModel.cpp:625-634)."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import atsat_amd
from atsat_amd import abi
from util import stress_state, stress_knobs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _knobs(k):
    ks = stress_knobs(k)
    ks.turb_closure = 1
    ks.chem_block = 1
    ks.mca_stride = 1
    return ks


@pytest.fixture
def tmpdir_():
    d = tempfile.mkdtemp()
    yield d
    shutil.rmtree(d)


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32])
def test_save_matches_host_writer_and_resume_is_exact(defaults, tmpdir_, prec):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    ks = _knobs(k)
    sts = [stress_state(g, c, seed=s) for s in (7, 8)]
    say("case restart small grid, 2 members, precision", prec)
    ctx = atsat_amd.Context(g, c, ks, 2, prec)
    for m, st in enumerate(sts):
        ctx.upload_state(st, member=m)
    ctx.run(1, 2, False)
    path = os.path.join(tmpdir_, "sat_restart_2.bin")
    ctx.save_restart(path, 2, member=1)
    host = sts[1].copy()
    ctx.download_state(host, member=1)
    ref = os.path.join(tmpdir_, "host.bin")
    atsat_amd.restart_write(ref, g, 2, host)
    assert open(path, "rb").read() == open(ref, "rb").read()
    # the writer continues; a fresh context with member 1's set-up state resumes from the file
    ctx.run(3, 2, False)
    cont = sts[1].copy()
    ctx.download_state(cont, member=1)
    ctx.close()
    ctx2 = atsat_amd.Context(g, c, ks, 1, prec)
    ctx2.upload_state(sts[1])
    assert ctx2.load_restart(path) == (2, 3)
    ctx2.run(3, 2, False)
    res = sts[1].copy()
    ctx2.download_state(res)
    ctx2.close()
    bad = [n for n in abi.PROG_NAMES + ["p_dyn", "p_stat", "rho_mix", "nue"] if not np.array_equal(cont[n], res[n])]
    assert not bad, bad


def test_load_errors(defaults, tmpdir_):
    _, c, k = defaults
    g = atsat_amd.small_grid(21, 61, 121)
    ctx = atsat_amd.Context(g, c, k, 1)
    with pytest.raises(atsat_amd.AtsatError, match="no file"):
        ctx.load_restart(os.path.join(tmpdir_, "none.bin"))
    g2 = atsat_amd.small_grid(21, 61, 61)
    p = os.path.join(tmpdir_, "other.bin")
    atsat_amd.restart_write(p, g2, 1, abi.HostState(g2), fmt=1)
    with pytest.raises(atsat_amd.AtsatError, match="grid mismatch"):
        ctx.load_restart(p)
    ctx.close()
