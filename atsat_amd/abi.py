"""ctypes mirror of the C ABI structs in include/atsat_fields.h and include/atsat_gpu.h.

Field ids and struct layouts must match the headers exactly; tests/test_abi.py checks the
exported symbols and the struct sizes against the compiled library.
"""
import ctypes as C

import numpy as np

NPROG = 17
PROG_NAMES = ["t", "u", "v", "w", "h2o", "h2o_cloud", "h2o_ice", "ch4", "ch4_cloud", "ch4_ice",
              "h2s", "nh3", "nh3_cloud", "nh3_ice", "nh4sh", "tke", "dis"]
FIELD_NAMES = (PROG_NAMES + [n + "n" for n in PROG_NAMES] + [
    "p_dyn", "p_stat", "p_metric", "rho_mix", "p_baro",
    "rhs_u", "rhs_v", "rhs_w", "aux_u", "aux_v", "aux_w",
    "nue", "tke_source", "dis_source",
    "massflux_h2s", "massflux_nh3", "massflux_nh4sh", "fluxlim_nh4sh",
    "Q_precip",
    "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice",
    "S_precip_nh3", "S_precip_nh3_cloud", "S_precip_nh3_ice",
    "S_precip_ch4", "S_precip_ch4_cloud", "S_precip_ch4_ice",
    "P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel",
    "P_ch4_rain", "P_ch4_snow", "P_ch4_graupel", "P_nh4sh", "prod",
    "difflux_h2s", "difflux_nh3", "difflux_nh4sh", "thermalmassflux"])
NFIELDS = len(FIELD_NAMES)
FIELD = {n: i for i, n in enumerate(FIELD_NAMES)}
VEC_NAMES = ["cell_vbar0", "cell_ubar0", "buoy_ref_level",
             "precip_srf_h2o", "precip_srf_nh3", "precip_srf_ch4", "precip_srf_nh4sh",
             "precip_srf_total", "vel_star"]
NVECS = len(VEC_NAMES)
VEC = {n: i for i, n in enumerate(VEC_NAMES)}
GAS_CH4, GAS_H2O, GAS_NH3 = 0, 1, 2
PREC_FP64, PREC_FP32 = 0, 1

# The fields the chemistry block writes.
CHEM_OUT = ["t", "nh3", "h2s", "nh4sh", "difflux_h2s", "difflux_nh3", "difflux_nh4sh",
            "thermalmassflux", "massflux_h2s", "massflux_nh3", "massflux_nh4sh", "fluxlim_nh4sh"]

# The fields the moist convective adjustment writes.
MCA_OUT = ["t", "h2o", "h2o_cloud", "h2o_ice", "nh3", "nh3_cloud", "nh3_ice", "ch4", "ch4_cloud",
           "ch4_ice", "h2s", "nh4sh"]

# The fields the turbulence closure writes.
TURB_OUT = ["tke", "dis", "tken", "disn", "nue", "prod", "tke_source", "dis_source"]

# The fields the hot path writes (what parity tests compare).
SPECIES = ["h2o", "h2o_cloud", "h2o_ice", "h2s", "nh3", "nh3_cloud", "nh3_ice",
           "ch4", "ch4_cloud", "ch4_ice", "nh4sh"]
PRECIP_OUT = ["Q_precip", "S_precip_h2o", "S_precip_h2o_cloud", "S_precip_h2o_ice",
              "S_precip_nh3", "S_precip_nh3_cloud", "S_precip_nh3_ice",
              "S_precip_ch4", "S_precip_ch4_cloud", "S_precip_ch4_ice",
              "P_rain", "P_snow", "P_graupel", "P_nh3_rain", "P_nh3_snow", "P_nh3_graupel",
              "P_ch4_rain", "P_ch4_snow", "P_ch4_graupel", "P_nh4sh"]


class Grid(C.Structure):
    _fields_ = [("im", C.c_int), ("jm", C.c_int), ("km", C.c_int),
                ("dr", C.c_double), ("dthe_deg", C.c_double), ("dphi_deg", C.c_double)]


_D3 = C.c_double * 3


class Consts(C.Structure):
    _fields_ = ([(n, C.c_double) for n in
                 ["re", "pr", "g", "omega", "gam", "u_0", "t_ref", "p_ref", "R_ref", "L_atm",
                  "Coriolis", "centrifugal", "buoyancy", "metric_radius_km",
                  "r_mix", "cp_mix", "R_mix", "sc_ch4", "sc_h2o", "sc_h2s", "sc_nh3", "sc_nh4sh"]]
                + [(n, _D3) for n in ["t_0", "t_00", "ep", "lv", "ls", "C", "L0", "R",
                                      "del_alf", "del_bet", "C_i", "L0_i", "del_alf_i",
                                      "del_bet_i"]]
                + [("t_0_nh4sh", C.c_double), ("t_00_nh4sh", C.c_double)]
                + [(n, C.c_double) for n in
                   ["m_nh3", "m_h2s", "m_nh4sh", "c_mix", "mue_nh3", "mue_h2s", "mue_nh4sh",
                    "rg_nh3", "rg_h2s", "rg_nh4sh", "cp_nh3", "cp_h2s", "cp_nh4sh",
                    "k_nh3", "k_h2s", "k_nh4sh", "ep_h2s"]])


class Knobs(C.Structure):
    _fields_ = [("dt", C.c_double), ("turb_active", C.c_int), ("chemical_reaction", C.c_double),
                ("turb_coupling", C.c_double), ("precip_coupling", C.c_double),
                ("pdyn_cap", C.c_double), ("press_sweeps", C.c_int), ("press_mode", C.c_int),
                ("press_tol", C.c_double), ("precip_on", C.c_int), ("chem_inputs", C.c_int),
                ("press_max_iter", C.c_int), ("press_fp32_transforms", C.c_int),
                ("turb_closure", C.c_int), ("turb_buoy", C.c_int), ("turb_k_bg", C.c_double),
                ("turb_nue_max", C.c_double), ("chem_block", C.c_int), ("mca_stride", C.c_int),
                ("press_line", C.c_int), ("press_baro", C.c_int)]


class HostStateC(C.Structure):
    _fields_ = [("f", C.POINTER(C.c_double) * NFIELDS), ("v", C.POINTER(C.c_double) * NVECS)]


# the arrays of a restart file the library does not hold (ATSAT_RX_* of atsat_gpu.h), under the
# reference's member names
RX_NAMES = ["p_dynn", "j_h2s", "j_nh3", "j_nh4sh", "w_h2s", "w_nh3", "w_nh4sh", "BuoyancyForce",
            "PresGradForce", "Q_Sensible", "jT_h2s", "jT_nh3", "jT_nh4sh"]
RX = {n: i for i, n in enumerate(RX_NAMES)}


class RestartExtraC(C.Structure):
    _fields_ = [("a", C.POINTER(C.c_double) * len(RX_NAMES))]


def restart_extra(arrays):
    """An atsat_restart_extra over a dict {name: contiguous float64 array} (missing names: NULL)."""
    x = RestartExtraC()
    for n, a in (arrays or {}).items():
        assert a.flags.c_contiguous and a.dtype == np.float64
        x.a[RX[n]] = a.ctypes.data_as(C.POINTER(C.c_double))
    return x


def grid_shape(g):
    return (g.im, g.jm, g.km)


def vec_len(g, name):
    if name in ("cell_vbar0", "cell_ubar0"):
        return g.im * g.jm
    if name == "buoy_ref_level":
        return g.im
    return g.jm * g.km


class HostState:
    """Numpy arrays of one member, with a C view (atsat_host_state) over them."""

    def __init__(self, grid, fields=None, vecs=None):
        self.grid = grid
        shp = grid_shape(grid)
        self.f = {}
        self.v = {}
        for n in (FIELD_NAMES if fields is None else fields):
            self.f[n] = np.zeros(shp, dtype=np.float64)
        for n in (VEC_NAMES if vecs is None else vecs):
            self.v[n] = np.zeros(vec_len(grid, n), dtype=np.float64)

    def copy(self):
        h = HostState(self.grid, fields=[], vecs=[])
        h.f = {k: a.copy() for k, a in self.f.items()}
        h.v = {k: a.copy() for k, a in self.v.items()}
        return h

    def cview(self):
        s = HostStateC()
        for n, a in self.f.items():
            assert a.flags.c_contiguous and a.dtype == np.float64
            s.f[FIELD[n]] = a.ctypes.data_as(C.POINTER(C.c_double))
        for n, a in self.v.items():
            s.v[VEC[n]] = a.ctypes.data_as(C.POINTER(C.c_double))
        self._keep = s
        return s

    def __getitem__(self, name):
        return self.f[name] if name in self.f else self.v[name]
