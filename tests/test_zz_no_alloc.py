"""atsat_damp_pdyn_noalloc: the p_dyn filter on caller-owned device memory, with no allocation.

The caller's memory comes from hipMalloc of the HIP runtime the library itself runs on (loaded by
its soname, so the same instance). The entry point must give, bit for bit, what the context
entry point atsat_damp_pdyn gives on the same p_dyn (AtomUtils::damp_wiggles, lib/Utils.cpp:229-325;
both run the same two kernels), on two calls in a row, the second smaller than the first, and it
must refuse a request larger than its scratch without launching anything.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import atsat_amd
from atsat_amd import abi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    atsat_amd.lib()                     # the library first: its runtime is the one found by soname
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _state(g, c, members, seed):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(members):
        st = atsat_amd.gen_v1(g, c, member=m)
        p = st["p_dyn"] + 1e-3 * rng.standard_normal(st["p_dyn"].shape)
        p[:, :, -1] = p[:, :, 0]
        out.append(np.ascontiguousarray(p))
    return np.stack(out)


def _with_context(g, c, k, p):
    ctx = atsat_amd.Context(g, c, k, p.shape[0], abi.PREC_FP64)
    for m in range(p.shape[0]):
        ctx.upload("p_dyn", p[m], member=m)
    ctx.call("damp_pdyn")
    ctx.synchronize()
    out = np.stack([ctx.download("p_dyn", member=m) for m in range(p.shape[0])])
    ctx.close()
    return out


def _noalloc(hip, g, p):
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), p.nbytes) == 0
    try:
        assert hip.hipMemcpy(d, p.ctypes.data, p.nbytes, H2D) == 0
        rc = atsat_amd.lib().atsat_damp_pdyn_noalloc(C.byref(g), p.shape[0], d, None)
        assert rc == 0, atsat_amd.lib().atsat_last_error(None)
        out = np.empty_like(p)
        assert hip.hipMemcpy(out.ctypes.data, d, p.nbytes, D2H) == 0   # null stream: after the filter
        return out
    finally:
        hip.hipFree(d)


def test_noalloc_matches_context_entry_point(defaults):
    gn, c, k = defaults
    hip = _hip()
    cases = [(gn, 1, 5), (atsat_amd.small_grid(21, 61, 121), 2, 6)]   # the second call is smaller
    for g, members, seed in cases:
        say("case damp_pdyn_noalloc %dx%dx%d x %d members" % (g.im, g.jm, g.km, members))
        p = _state(g, c, members, seed)
        want = _with_context(g, c, k, p)
        got = _noalloc(hip, g, p)
        assert not np.array_equal(want, p)
        assert np.array_equal(got, want)


def test_noalloc_refuses_oversize(defaults):
    g, _, _ = defaults
    say("case damp_pdyn_noalloc oversize request (4 native members)")
    hip = _hip()
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), 8) == 0
    try:
        rc = atsat_amd.lib().atsat_damp_pdyn_noalloc(C.byref(g), 4, d, None)
        assert rc == -1
        assert b"exceed the scratch" in atsat_amd.lib().atsat_last_error(None)
    finally:
        hip.hipFree(d)
