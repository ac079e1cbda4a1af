"""atsat_amd -- the MI355X model time step of ATSAT behind a C ABI (include/atsat_gpu.h).

This module is a thin ctypes mirror of that ABI for Python callers (tests, bench.py). The
product is libatsat_gpu.so, built in-tree from atsat_amd/csrc by atsat_amd.build; importing the
module does not build anything, and every call fails loudly if the library is missing.

The step's entry points mirror the reference's four call sites (planet/cSaturnModel.cpp:873,
908-927, 939, 1030): Context.rk4_step() for cSaturnModel::RungeKuttaSat, pressure_solve() for
PressureSolver<cSaturnModel>::run, saturation_adjust(gas) for SaturationAdjustmentSat::run and
precipitation() for Precipitation<cSaturnModel>::run; step(iter) is one iteration of Run's loop.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import FIELD, VEC, HostState, Grid, Consts, Knobs  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libatsat_gpu.so")

_lib = None

ENTRY_POINTS = [
    "atsat_default_grid", "atsat_default_consts", "atsat_default_knobs", "atsat_resident_knobs", "atsat_gen_v1",
    "atsat_ctx_create", "atsat_ctx_destroy", "atsat_last_error", "atsat_set_knobs",
    "atsat_device_bytes", "atsat_upload", "atsat_download", "atsat_upload_vec",
    "atsat_download_vec", "atsat_upload_state", "atsat_download_state", "atsat_upload_rows",
    "atsat_download_rows", "atsat_rk4_step", "atsat_pressure_solve", "atsat_saturation_adjust",
    "atsat_precipitation", "atsat_mixture_density", "atsat_damp_pdyn", "atsat_rebuild_pstat",
    "atsat_rk_scalar_sync", "atsat_boundary_conditions", "atsat_clamp_species",
    "atsat_restore_var", "atsat_step", "atsat_run", "atsat_synchronize", "atsat_nonfinite_count",
    "atsat_profile_enable", "atsat_profile_read", "atsat_poisson_solve", "atsat_pressure_last",
    "atsat_autotune", "atsat_set_tile", "atsat_get_tile", "atsat_group_create", "atsat_group_destroy",
    "atsat_group_last_error", "atsat_group_autotune", "atsat_group_tile", "atsat_group_run",
    "atsat_group_synchronize", "atsat_group_merged", "atsat_group_set_merged",
    "atsat_turbulence", "atsat_turbulence_init", "atsat_nh4sh_equilibrium", "atsat_chemistry",
    "atsat_convective_adjustment", "atsat_restart_write_host", "atsat_restart_read_host",
    "atsat_save_restart", "atsat_load_restart", "atsat_damp_pdyn_noalloc", "atsat_budgets",
]
KERNELS = ["rk_stage", "rk_aux", "press", "damp", "satadj", "pstat", "precip", "glue", "density", "turb",
           "chem", "mca"]


class AtsatError(RuntimeError):
    pass


def lib():
    """The loaded libatsat_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AtsatError("libatsat_gpu.so is not built (run python -m atsat_amd.build): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.atsat_default_grid.argtypes = [P(Grid)]
        L.atsat_default_consts.argtypes = [P(Consts)]
        L.atsat_default_knobs.argtypes = [P(Knobs)]
        L.atsat_resident_knobs.argtypes = [P(Knobs)]
        L.atsat_gen_v1.argtypes = [P(Grid), P(Consts), C.c_ulonglong, C.c_int, C.c_double, C.c_double,
                                   P(abi.HostStateC)]
        L.atsat_ctx_create.argtypes = [P(Grid), P(Consts), P(Knobs), C.c_int, C.c_int, P(C.c_void_p)]
        L.atsat_ctx_destroy.argtypes = [C.c_void_p]
        L.atsat_last_error.argtypes = [C.c_void_p]
        L.atsat_last_error.restype = C.c_char_p
        L.atsat_set_knobs.argtypes = [C.c_void_p, P(Knobs)]
        L.atsat_device_bytes.argtypes = [C.c_void_p, P(C.c_size_t)]
        for n in ("atsat_upload", "atsat_download", "atsat_upload_vec", "atsat_download_vec"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double)]
        for n in ("atsat_upload_state", "atsat_download_state"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int, P(abi.HostStateC)]
        for n in ("atsat_upload_rows", "atsat_download_rows"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int, C.c_int, P(P(C.c_double))]
        for n in ("atsat_rk4_step", "atsat_pressure_solve", "atsat_precipitation",
                  "atsat_mixture_density", "atsat_damp_pdyn", "atsat_rebuild_pstat",
                  "atsat_rk_scalar_sync", "atsat_boundary_conditions", "atsat_clamp_species",
                  "atsat_restore_var", "atsat_synchronize", "atsat_turbulence",
                  "atsat_turbulence_init", "atsat_nh4sh_equilibrium", "atsat_chemistry",
                  "atsat_convective_adjustment"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.atsat_saturation_adjust.argtypes = [C.c_void_p, C.c_int]
        L.atsat_damp_pdyn_noalloc.argtypes = [P(Grid), C.c_int, C.c_void_p, C.c_void_p]
        L.atsat_step.argtypes = [C.c_void_p, C.c_int]
        L.atsat_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.atsat_nonfinite_count.argtypes = [C.c_void_p, P(C.c_longlong)]
        L.atsat_budgets.argtypes = [C.c_void_p, P(C.c_double)]
        L.atsat_profile_enable.argtypes = [C.c_void_p, C.c_int]
        L.atsat_profile_read.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_longlong)]
        L.atsat_poisson_solve.argtypes = [C.c_void_p, P(C.c_double), C.c_double, C.c_int, P(C.c_int), P(C.c_double)]
        L.atsat_pressure_last.argtypes = [C.c_void_p, P(C.c_int), P(C.c_double)]
        L.atsat_autotune.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_double)]
        L.atsat_set_tile.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.atsat_get_tile.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.atsat_group_create.argtypes = [P(C.c_void_p), C.c_int, P(C.c_void_p)]
        L.atsat_group_destroy.argtypes = [C.c_void_p]
        L.atsat_group_last_error.argtypes = [C.c_void_p]
        L.atsat_group_last_error.restype = C.c_char_p
        L.atsat_group_autotune.argtypes = [C.c_void_p, P(C.c_double)]
        L.atsat_group_tile.argtypes = [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]
        L.atsat_group_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.atsat_group_synchronize.argtypes = [C.c_void_p]
        L.atsat_group_merged.argtypes = [C.c_void_p, P(C.c_int)]
        L.atsat_group_set_merged.argtypes = [C.c_void_p, C.c_int]
        L.atsat_restart_write_host.argtypes = [C.c_char_p, P(Grid), C.c_int, P(abi.HostStateC),
                                               P(abi.RestartExtraC), C.c_int]
        L.atsat_restart_read_host.argtypes = [C.c_char_p, P(Grid), P(abi.HostStateC), P(abi.RestartExtraC),
                                              P(C.c_int), P(C.c_int)]
        L.atsat_save_restart.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, P(abi.RestartExtraC), C.c_int]
        L.atsat_load_restart.argtypes = [C.c_void_p, C.c_int, C.c_char_p, P(abi.RestartExtraC), P(C.c_int),
                                         P(C.c_int)]
        _lib = L
    return _lib


def restart_write(path, grid, iteration, state, extra=None, fmt=3):
    """Write one member's HostState as a restart file in the reference's format
    (cSaturnModel::save_state, planet/FileIO_Sat.cpp:572-605); extra: {name: array} of the arrays
    the library does not hold (abi.RX_NAMES), else p_dynn = p_dyn and zeros."""
    x = abi.restart_extra(extra) if extra else None
    rc = lib().atsat_restart_write_host(os.fsencode(path), C.byref(grid), iteration, C.byref(state.cview()),
                                        C.byref(x) if x is not None else None, fmt)
    if rc != 0:
        raise AtsatError("atsat_restart_write_host failed (%d): %s" % (rc, lib().atsat_last_error(None).decode()))


def restart_read(path, grid, state, extra=None):
    """Read a restart file (SAT1 / SAT2 / SAT3, cSaturnModel::load_state, FileIO_Sat.cpp:607-660)
    into a HostState (and the extras given); returns (iteration, format)."""
    it, fmt = C.c_int(), C.c_int()
    x = abi.restart_extra(extra) if extra else None
    rc = lib().atsat_restart_read_host(os.fsencode(path), C.byref(grid), C.byref(state.cview()),
                                       C.byref(x) if x is not None else None, C.byref(it), C.byref(fmt))
    if rc != 0:
        raise AtsatError("atsat_restart_read_host failed (%d): %s" % (rc, lib().atsat_last_error(None).decode()))
    return it.value, fmt.value


def defaults(profile="model"):
    """(grid, consts, knobs) with the reference's default values. profile "model": the knobs the
    reference runs by default (closure, chemistry block and convective adjustment on);
    "resident": the same with those three off (ATSAT_TURB=0, ATSAT_MCA=0, chemical_reaction 0)."""
    if profile not in ("model", "resident"):
        raise ValueError("profile must be 'model' or 'resident'")
    g, c, k = Grid(), Consts(), Knobs()
    L = lib()
    L.atsat_default_grid(C.byref(g))
    L.atsat_default_consts(C.byref(c))
    (L.atsat_default_knobs if profile == "model" else L.atsat_resident_knobs)(C.byref(k))
    return g, c, k


def small_grid(im=21, jm=61, km=121, dr=None, dthe_deg=None, dphi_deg=None):
    """A coarser shell of the same depth: dr = 1/(im-1) * 1, dtheta = 180/(jm-1) deg."""
    g = Grid()
    g.im, g.jm, g.km = im, jm, km
    g.dr = dr if dr is not None else 1.0 / (im - 1)
    g.dthe_deg = dthe_deg if dthe_deg is not None else 180.0 / (jm - 1)
    g.dphi_deg = dphi_deg if dphi_deg is not None else 360.0 / (km - 1)
    return g


def gen_v1(grid, consts, seed=20261203, member=0, t_sigma=0.0, w_sigma=0.0, state=None):
    """The seeded synthetic start state of one member (host, FP64)."""
    st = state if state is not None else HostState(grid)
    rc = lib().atsat_gen_v1(C.byref(grid), C.byref(consts), seed, member, t_sigma, w_sigma,
                            C.byref(st.cview()))
    if rc != 0:
        raise AtsatError("atsat_gen_v1 failed: %d" % rc)
    return st


class Context:
    """A device context: n_members members of one grid, FP64 or FP32 storage."""

    def __init__(self, grid, consts, knobs, n_members=1, precision=abi.PREC_FP64):
        self.grid, self.consts, self.knobs = grid, consts, knobs
        self.n_members, self.precision = n_members, precision
        h = C.c_void_p()
        rc = lib().atsat_ctx_create(C.byref(grid), C.byref(consts), C.byref(knobs), n_members,
                                    precision, C.byref(h))
        if rc != 0:
            raise AtsatError("atsat_ctx_create failed (%d): %s" % (rc, self._err(None)))
        self.h = h

    def _err(self, h):
        s = lib().atsat_last_error(h)
        return s.decode() if s else ""

    def _chk(self, rc, what):
        if rc != 0:
            raise AtsatError("%s failed (%d): %s" % (what, rc, self._err(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().atsat_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_knobs(self, knobs):
        self.knobs = knobs
        self._chk(lib().atsat_set_knobs(self.h, C.byref(knobs)), "set_knobs")

    def device_bytes(self):
        n = C.c_size_t()
        self._chk(lib().atsat_device_bytes(self.h, C.byref(n)), "device_bytes")
        return n.value

    def upload_state(self, st, member=0):
        self._chk(lib().atsat_upload_state(self.h, member, C.byref(st.cview())), "upload_state")

    def download_state(self, st, member=0):
        self._chk(lib().atsat_download_state(self.h, member, C.byref(st.cview())), "download_state")

    def upload(self, name, arr, member=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if name in FIELD:
            self._chk(lib().atsat_upload(self.h, member, FIELD[name], a.ctypes.data_as(C.POINTER(C.c_double))), "upload")
        else:
            self._chk(lib().atsat_upload_vec(self.h, member, VEC[name], a.ctypes.data_as(C.POINTER(C.c_double))), "upload_vec")

    def download(self, name, member=0):
        if name in FIELD:
            a = np.empty(abi.grid_shape(self.grid), dtype=np.float64)
            self._chk(lib().atsat_download(self.h, member, FIELD[name], a.ctypes.data_as(C.POINTER(C.c_double))), "download")
        else:
            a = np.empty(abi.vec_len(self.grid, name), dtype=np.float64)
            self._chk(lib().atsat_download_vec(self.h, member, VEC[name], a.ctypes.data_as(C.POINTER(C.c_double))), "download_vec")
        return a

    # ---- the hot path ----
    def rk4_step(self):
        self._chk(lib().atsat_rk4_step(self.h), "rk4_step")

    def pressure_solve(self):
        self._chk(lib().atsat_pressure_solve(self.h), "pressure_solve")

    def saturation_adjust(self, gas):
        self._chk(lib().atsat_saturation_adjust(self.h, gas), "saturation_adjust")

    def precipitation(self):
        self._chk(lib().atsat_precipitation(self.h), "precipitation")

    def turbulence(self):
        self._chk(lib().atsat_turbulence(self.h), "turbulence")

    def turbulence_init(self):
        self._chk(lib().atsat_turbulence_init(self.h), "turbulence_init")

    def call(self, name, *args):
        self._chk(getattr(lib(), "atsat_" + name)(self.h, *args), name)

    def save_restart(self, path, iteration, member=0, extra=None, fmt=3):
        """cSaturnModel::save_state of one member (FileIO_Sat.cpp:572-605) to path."""
        x = abi.restart_extra(extra) if extra else None
        self._chk(lib().atsat_save_restart(self.h, member, iteration, os.fsencode(path),
                                           C.byref(x) if x is not None else None, fmt), "save_restart")

    def load_restart(self, path, member=0, extra=None):
        """cSaturnModel::load_state into one member (FileIO_Sat.cpp:607-660); (iteration, format)."""
        it, fmt = C.c_int(), C.c_int()
        x = abi.restart_extra(extra) if extra else None
        self._chk(lib().atsat_load_restart(self.h, member, os.fsencode(path), C.byref(x) if x is not None else None,
                                           C.byref(it), C.byref(fmt)), "load_restart")
        return it.value, fmt.value

    def step(self, it):
        self._chk(lib().atsat_step(self.h, it), "step")

    def run(self, first_iter, n_steps, use_graph=False):
        self._chk(lib().atsat_run(self.h, first_iter, n_steps, 1 if use_graph else 0), "run")

    def synchronize(self):
        self._chk(lib().atsat_synchronize(self.h), "synchronize")

    def profile(self, on=True):
        self._chk(lib().atsat_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_read(self):
        """{kernel family: (total ms, launches)} since profile(True)."""
        out = {}
        for i, n in enumerate(KERNELS):
            ms, cnt = C.c_double(), C.c_longlong()
            self._chk(lib().atsat_profile_read(self.h, i, C.byref(ms), C.byref(cnt)), "profile_read")
            out[n] = (ms.value, cnt.value)
        return out

    def poisson_solve(self, src=None, tol=1e-10, max_iter=8):
        """Converged solve of the pressure equation for p_dyn (see atsat_poisson_solve). src: None
        (the step's own source) or an array [members, im, jm, km]. Returns (corrections, relative
        residual, converged)."""
        it, rel = C.c_int(), C.c_double()
        ptr = None
        if src is not None:
            a = np.ascontiguousarray(src, dtype=np.float64)
            assert a.size == self.n_members * self.grid.im * self.grid.jm * self.grid.km
            ptr = a.ctypes.data_as(C.POINTER(C.c_double))
        rc = lib().atsat_poisson_solve(self.h, ptr, tol, max_iter, C.byref(it), C.byref(rel))
        if rc not in (0, -6):
            self._chk(rc, "poisson_solve")
        return it.value, rel.value, rc == 0

    def pressure_last(self):
        it, rel = C.c_int(), C.c_double()
        self._chk(lib().atsat_pressure_last(self.h, C.byref(it), C.byref(rel)), "pressure_last")
        return it.value, rel.value

    def nonfinite_count(self):
        n = C.c_longlong()
        self._chk(lib().atsat_nonfinite_count(self.h, C.byref(n)), "nonfinite_count")
        return n.value

    def budgets(self):
        """The FP64 budget sums of every member (atsat_budgets): (columns [members, 3] of H2O, NH3,
        CH4 in g/m2, level-mean temperatures [members, im] in K)."""
        im = self.grid.im
        out = np.zeros(self.n_members * (3 + im))
        self._chk(lib().atsat_budgets(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "budgets")
        out = out.reshape(self.n_members, 3 + im)
        return out[:, :3].copy(), out[:, 3:].copy()

    def autotune(self):
        """Time every launch tile of the RK4 stage kernel (rounds of the four stages) on this
        context's state and keep the fastest. Returns (rows per band, rings per workgroup, mean ms
        per stage launch). The state is saved and restored."""
        band, rings, ms = C.c_int(), C.c_int(), C.c_double()
        self._chk(lib().atsat_autotune(self.h, C.byref(band), C.byref(rings), C.byref(ms)), "autotune")
        return band.value, rings.value, ms.value

    def set_tile(self, band, rings=1):
        self._chk(lib().atsat_set_tile(self.h, band, rings), "set_tile")

    def tile(self):
        band, rings = C.c_int(), C.c_int()
        self._chk(lib().atsat_get_tile(self.h, C.byref(band), C.byref(rings)), "get_tile")
        return band.value, rings.value


class Group:
    """Contexts of different grids advanced together: per iteration each context's physics on its
    own stream, the four RK4 stages of all of them as one launch each. The contexts must outlive
    the group and share precision and turb_active."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        rc = lib().atsat_group_create(arr, len(self.contexts), C.byref(h))
        if rc != 0:
            raise AtsatError("atsat_group_create failed (%d): %s" % (rc, self._err(None)))
        self.h = h

    def _err(self, h):
        s = lib().atsat_group_last_error(h)
        return s.decode() if s else ""

    def _chk(self, rc, what):
        if rc != 0:
            raise AtsatError("%s failed (%d): %s" % (what, rc, self._err(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().atsat_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        """Choose between the batched launches (with their workgroup width) and the contexts side
        by side, and each grid's tile; returns the kept layout's mean ms per stage."""
        ms = C.c_double()
        self._chk(lib().atsat_group_autotune(self.h, C.byref(ms)), "group_autotune")
        return ms.value

    def layout(self):
        """The batch's layout (the autotuner's choice): 0 the contexts step side by side on their
        own streams; 1 the RK4 stages of all grids run as one launch each; 2 the same as a
        persistent launch over a work queue."""
        m = C.c_int()
        self._chk(lib().atsat_group_merged(self.h, C.byref(m)), "group_merged")
        return m.value

    def merged(self):
        """True when the RK4 stages of all grids run in one launch each (layout 1 or 2)."""
        return self.layout() != 0

    def set_layout(self, layout):
        self._chk(lib().atsat_group_set_merged(self.h, int(layout)), "group_set_merged")

    def set_merged(self, merged):
        self.set_layout(1 if merged else 0)

    def tiles(self):
        """[(rows per band, rings per workgroup, workgroup threads)] per context."""
        out = []
        for i in range(len(self.contexts)):
            b, r, t = C.c_int(), C.c_int(), C.c_int()
            self._chk(lib().atsat_group_tile(self.h, i, C.byref(b), C.byref(r), C.byref(t)), "group_tile")
            out.append((b.value, r.value, t.value))
        return out

    def run(self, first_iter, n_steps, use_graph=False):
        self._chk(lib().atsat_group_run(self.h, first_iter, n_steps, 1 if use_graph else 0), "group_run")

    def synchronize(self):
        self._chk(lib().atsat_group_synchronize(self.h), "group_synchronize")
