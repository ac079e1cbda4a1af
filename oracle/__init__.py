"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes bindings of the FP64 CPU restatement (oracle/_build/liborc.so, built from
oracle/atsat_oracle.cpp) and, where it was built, of the reference harness
(oracle/_ref/libatsat_ref.so, the reference's own sources + oracle/ref_harness.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package; the
product (atsat_amd) never does.
"""
import ctypes as C
import os
import subprocess

from atsat_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "_build", "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libatsat_ref.so")
REF_TREE = "/root/reference"

_orc = None
_ref = None


def build(ref=True):
    """Compile the restatement (and, when the reference tree is present, the harness)."""
    targets = ["oracle"]
    if ref and os.path.isdir(REF_TREE):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, "-j8"] + targets, check=True)


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_SO):
            build(ref=False)
        lib = C.CDLL(ORC_SO)
        sig = [C.POINTER(abi.Grid), C.POINTER(abi.Consts), C.POINTER(abi.Knobs),
               C.POINTER(abi.HostStateC)]
        for n in ["orc_mixture_density", "orc_buoyancy_ref_level", "orc_metric_pressure",
                  "orc_rk4_step", "orc_prepare_projection_boundaries", "orc_pressure_solve",
                  "orc_damp_pdyn", "orc_rebuild_pstat", "orc_precipitation", "orc_rk_scalar_sync",
                  "orc_boundary_conditions", "orc_clamp_species", "orc_restore_var",
                  "orc_turbulence", "orc_turbulence_init", "orc_nh4sh_equilibrium",
                  "orc_chemistry", "orc_mca"]:
            getattr(lib, n).argtypes = sig
            getattr(lib, n).restype = C.c_int
        lib.orc_saturation_adjust.argtypes = sig + [C.c_int]
        lib.orc_budgets.argtypes = sig + [C.POINTER(C.c_double)]
        lib.orc_budgets.restype = C.c_int
        lib.orc_step.argtypes = sig + [C.c_int]
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_create.argtypes = [C.POINTER(abi.Knobs)]
        lib.ref_create.restype = C.c_void_p
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_load.argtypes = [C.c_void_p, C.POINTER(abi.HostStateC)]
        lib.ref_store.argtypes = [C.c_void_p, C.POINTER(abi.HostStateC)]
        lib.ref_call.argtypes = [C.c_void_p, C.POINTER(abi.Knobs), C.c_int, C.c_int]
        lib.ref_save_state.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        lib.ref_load_state.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        lib.ref_extra.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
        _ref = lib
    return _ref


class Oracle:
    """The restatement on one member's HostState."""

    def __init__(self, grid, consts, knobs):
        self.g, self.c, self.k = grid, consts, knobs

    def _a(self, st):
        return (C.byref(self.g), C.byref(self.c), C.byref(self.k), C.byref(st.cview()))

    def call(self, name, st, *extra):
        rc = getattr(orc(), "orc_" + name)(*self._a(st), *extra)
        if rc != 0:
            raise RuntimeError("oracle %s failed: %d" % (name, rc))

    def step(self, st, it):
        self.call("step", st, it)

    def budgets(self, st):
        """The reference's budget sums of a member: (H2O, NH3, CH4 columns in g/m2, the level-mean
        temperatures in K)."""
        import numpy as np
        out = np.zeros(3 + self.g.im)
        self.call("budgets", st, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:3].copy(), out[3:].copy()


# what-codes of ref_call (oracle/ref_harness.cpp)
REF_CALLS = {"mixture_density": 0, "pressure_solve": 1, "damp_pdyn": 2, "saturation_adjust": 3,
             "rebuild_pstat": 4, "precipitation": 5, "rk_scalar_sync": 6, "rk4_step": 7,
             "boundary_conditions": 8, "clamp_species": 9, "restore_var": 10, "step": 11,
             "buoyancy_ref_level": 12, "metric_pressure": 13,
             "prepare_projection_boundaries": 14, "turbulence": 15, "turbulence_init": 16,
             "nh4sh_equilibrium": 17, "chemistry": 18, "mca": 19}


class Reference:
    """The reference's own functions (native 41x181x361 grid only) on a HostState."""

    def __init__(self, knobs):
        self.k = knobs
        self.h = ref().ref_create(C.byref(knobs))
        if not self.h:
            raise RuntimeError("ref_create failed")

    def close(self):
        if self.h:
            ref().ref_destroy(self.h)
            self.h = None

    def load(self, st):
        if ref().ref_load(self.h, C.byref(st.cview())) != 0:
            raise RuntimeError("ref_load failed")

    def store(self, st):
        if ref().ref_store(self.h, C.byref(st.cview())) != 0:
            raise RuntimeError("ref_store failed")

    def call(self, name, arg=0):
        if ref().ref_call(self.h, C.byref(self.k), REF_CALLS[name], arg) != 0:
            raise RuntimeError("ref_call %s failed" % name)

    def save_state(self, directory, iteration):
        """cSaturnModel::save_state: writes directory/sat_restart_<iteration>.bin."""
        ref().ref_save_state(self.h, os.fsencode(directory), iteration)
        return os.path.join(directory, "sat_restart_%d.bin" % iteration)

    def load_state(self, directory, iteration):
        """cSaturnModel::load_state; True when the file was read."""
        return ref().ref_load_state(self.h, os.fsencode(directory), iteration) == 1

    def extra(self, name, shape, value=None):
        """Get (value None) or set one of the restart arrays the library does not hold."""
        import numpy as np
        a = np.zeros(shape) if value is None else np.ascontiguousarray(value, dtype=np.float64)
        if ref().ref_extra(self.h, abi.RX[name], a.ctypes.data_as(C.POINTER(C.c_double)), 0 if value is None else 1) != 0:
            raise RuntimeError("ref_extra %s failed" % name)
        return a
