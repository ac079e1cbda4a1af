"""The C ABI library loads on a CPU-only host, exports every entry point include/*.h declares,
and its struct layouts agree with the ctypes mirror (no GPU call is made here)."""
import ctypes as C
import os
import re
import subprocess

import atsat_amd
from atsat_amd import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "include")


def declared():
    names = set()
    for h in os.listdir(INC):
        if h.endswith(".h"):
            src = open(os.path.join(INC, h)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(atsat_\w+)\s*\(", src, re.M))
    return names


def test_every_declared_symbol_is_exported():
    L = atsat_amd.lib()
    names = declared()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(atsat_amd.ENTRY_POINTS), names ^ set(atsat_amd.ENTRY_POINTS)


def test_struct_layouts_match_ctypes(tmp_path):
    prog = tmp_path / "sizes.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "atsat_gpu.h"\n'
                    'int main(){ printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(atsat_grid), sizeof(atsat_consts),'
                    ' sizeof(atsat_knobs), sizeof(atsat_host_state), offsetof(atsat_consts, t_0),'
                    ' offsetof(atsat_knobs, chem_inputs)); return 0; }\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I" + INC, str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = list(map(int, out))
    want = [C.sizeof(abi.Grid), C.sizeof(abi.Consts), C.sizeof(abi.Knobs), C.sizeof(abi.HostStateC),
            abi.Consts.t_0.offset, abi.Knobs.chem_inputs.offset]
    assert got == want
    assert abi.NFIELDS == len(abi.FIELD_NAMES) and abi.NPROG == 17


def test_defaults_are_the_references():
    g, c, k = atsat_amd.defaults()
    assert (g.im, g.jm, g.km, g.dr) == (41, 181, 361, 0.025)
    assert c.re == 1e5 and c.pr == 69.0 and abs(c.R_mix - 2909.2457308249) < 1e-9
    assert abs(c.cp_mix - 10058.5383502171) < 1e-9 and abs(c.r_mix - 1.382) < 1e-12
    assert k.dt == 0.01 and k.press_sweeps == 1 and k.pdyn_cap == 0.2 and k.press_mode == 0
    # turb_model k_omega_SST and chemical_reaction 1 (saturn/config_atsat.xml:17-19), ATSAT_MCA 1
    assert k.turb_active == 1 and k.turb_closure == 1 and k.chem_block == 1
    assert k.chemical_reaction == 1.0 and k.mca_stride == 1
    _, _, r = atsat_amd.defaults("resident")
    assert r.turb_active == 0 and r.chem_block == 0 and r.chemical_reaction == 0.0 and r.mca_stride == 0
    assert r.dt == k.dt and r.press_sweeps == k.press_sweeps and r.turb_buoy == k.turb_buoy


def test_generator_is_deterministic_and_finite():
    g = atsat_amd.small_grid(11, 31, 61)
    _, c, _ = atsat_amd.defaults()
    a = atsat_amd.gen_v1(g, c, seed=1, member=2, t_sigma=0.5)
    b = atsat_amd.gen_v1(g, c, seed=1, member=2, t_sigma=0.5)
    d = atsat_amd.gen_v1(g, c, seed=1, member=3, t_sigma=0.5)
    import numpy as np
    for n in a.f:
        assert np.isfinite(a[n]).all(), n
        assert np.array_equal(a[n], b[n]), n
    assert not np.array_equal(a["t"], d["t"])
    assert np.array_equal(a["t"][:, :, -1], a["t"][:, :, 0])


def test_context_errors_without_gpu_are_reported():
    """Creating a context with a bad grid fails with ATSAT_E_ARG before touching a device."""
    g = atsat_amd.small_grid(3, 31, 61)
    _, c, k = atsat_amd.defaults()
    h = C.c_void_p()
    rc = atsat_amd.lib().atsat_ctx_create(C.byref(g), C.byref(c), C.byref(k), 1, 0, C.byref(h))
    assert rc == -1
    assert b"grid" in atsat_amd.lib().atsat_last_error(None)
