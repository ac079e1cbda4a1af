/*
 * atsat_fields.h -- field identifiers, physical constants and frozen knobs of the ATSAT model
 * time step, shared by the GPU library (atsat_gpu.h), the CPU oracle and the test harnesses.
 *
 * Field names are the reference's Array members (planet/cSaturnModel.h:1255-1477). Note the
 * reference's naming: `u` is the RADIAL velocity, `v` the meridional (+south) and `w` the ZONAL
 * one (planet/cSaturnModel.h:1301-1304, planet/RHS_Sat_Turb.cpp:281-286).
 *
 * Every 3-D field is stored as [i][j][k] with k fastest: i = 0..im-1 (radius), j = 0..jm-1
 * (colatitude, j = 0 the north pole), k = 0..km-1 (longitude). k = 0 and k = km-1 are the same
 * meridian; with the periodic seam (ATSAT_BC_PHI_PERIODIC=1, ATPhys.h:111-139) k = 0 is
 * integrated and k = km-1 is its copy.
 */
#ifndef ATSAT_FIELDS_H
#define ATSAT_FIELDS_H

#ifdef __cplusplus
extern "C" {
#endif

/* The 17 prognostic fields the RK4 step integrates, in the order of
 * planet/RungeKutta_Sat_Turb.cpp:254-291. tke/dis are integrated only with turb_active. */
enum {
    ATSAT_T = 0, ATSAT_U, ATSAT_V, ATSAT_W,
    ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE,
    ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE,
    ATSAT_H2S,
    ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE,
    ATSAT_NH4SH,
    ATSAT_TKE, ATSAT_DIS,
    ATSAT_NPROG                                   /* = 17 */
};

/* All 3-D fields. The first ATSAT_NPROG are the live prognostic fields; the next ATSAT_NPROG
 * their start-of-step copies (tn, un, ... : planet/cSaturnModel.cpp:1551-1602). */
enum {
    ATSAT_F_N0        = ATSAT_NPROG,              /* tn = ATSAT_F_N0 + ATSAT_T, ... */
    ATSAT_P_DYN       = 2 * ATSAT_NPROG,          /* dynamic pressure (Poisson unknown) */
    ATSAT_P_STAT,                                 /* hydrostatic pressure [bar] */
    ATSAT_P_METRIC,                               /* metric-pressure column integral */
    ATSAT_RHO_MIX,                                /* local mixture density [kg/m3] */
    ATSAT_P_BARO,                                 /* frozen per-level rest (initGeoBalance) */
    ATSAT_RHS_U, ATSAT_RHS_V, ATSAT_RHS_W,        /* last-stage tendencies (Poisson source) */
    ATSAT_AUX_U, ATSAT_AUX_V, ATSAT_AUX_W,        /* tendencies without the p_dyn gradient */
    ATSAT_NUE, ATSAT_TKE_SOURCE, ATSAT_DIS_SOURCE,/* closure inputs, frozen over a step */
    ATSAT_MASSFLUX_H2S, ATSAT_MASSFLUX_NH3, ATSAT_MASSFLUX_NH4SH,
    ATSAT_FLUXLIM_NH4SH,                          /* chemistry inputs, frozen over a step */
    ATSAT_Q_PRECIP,                               /* precipitation latent heat [W/m3] */
    ATSAT_S_PRECIP_H2O, ATSAT_S_PRECIP_H2O_CLOUD, ATSAT_S_PRECIP_H2O_ICE,
    ATSAT_S_PRECIP_NH3, ATSAT_S_PRECIP_NH3_CLOUD, ATSAT_S_PRECIP_NH3_ICE,
    ATSAT_S_PRECIP_CH4, ATSAT_S_PRECIP_CH4_CLOUD, ATSAT_S_PRECIP_CH4_ICE,
    ATSAT_P_RAIN, ATSAT_P_SNOW, ATSAT_P_GRAUPEL,
    ATSAT_P_NH3_RAIN, ATSAT_P_NH3_SNOW, ATSAT_P_NH3_GRAUPEL,
    ATSAT_P_CH4_RAIN, ATSAT_P_CH4_SNOW, ATSAT_P_CH4_GRAUPEL,
    ATSAT_P_NH4SH,
    ATSAT_PROD,                                   /* shear production P_k of the closure (output) */
    ATSAT_DIFFLUX_H2S, ATSAT_DIFFLUX_NH3, ATSAT_DIFFLUX_NH4SH,   /* D * laplacian(c) (DiffMassFluxSat) */
    ATSAT_THERMALMASSFLUX,                        /* thermal diffusion heating (DiffMassFluxSat) */
    ATSAT_NFIELDS
};

/* Small arrays (not [im][jm][km]). */
enum {
    ATSAT_V_CELL_VBAR0 = 0,   /* [im*jm]  frozen zonal-mean v (initCellHold) */
    ATSAT_V_CELL_UBAR0,       /* [im*jm]  frozen zonal-mean u */
    ATSAT_V_BUOY_REF,         /* [im]     per-level buoyancy reference (computeBuoyancyRefLevel) */
    ATSAT_V_SRF_H2O,          /* [jm*km]  surface precipitation maps (Precipitation::surfaceMap) */
    ATSAT_V_SRF_NH3,
    ATSAT_V_SRF_CH4,
    ATSAT_V_SRF_NH4SH,
    ATSAT_V_SRF_TOTAL,
    ATSAT_V_VEL_STAR,         /* [jm*km]  friction velocity u_tau [m/s] (Turbulence::compute_vel_star) */
    ATSAT_NVECS
};

/* Gas selector of the saturation adjustment, in the reference's call order
 * (planet/cSaturnModel.cpp:908-927). */
enum { ATSAT_GAS_CH4 = 0, ATSAT_GAS_H2O = 1, ATSAT_GAS_NH3 = 2 };

/* Grid descriptor. The reference's grid is 41 x 181 x 361, dr = 0.025, dtheta = dphi = 1 deg
 * (planet/cSaturnModel.h:711, planet/cSaturnModel.cpp:42-52); the coordinates are built by
 * repeated addition from r0 = 1 and theta0 = phi0 = 0 (lib/Array_1D.cpp:50-61). */
typedef struct atsat_grid {
    int im, jm, km;
    double dr;          /* radial spacing (nondimensional) */
    double dthe_deg;    /* meridional spacing [deg] */
    double dphi_deg;    /* zonal spacing [deg] */
} atsat_grid;

/* Physical constants the hot path reads. Defaults are the reference's XML values
 * (saturn/config_atsat.xml) and member initialisers (planet/cSaturnModel.h:721-1056); the
 * mixture values r_mix, cp_mix, R_mix follow ThermalPropertiesSat (planet/ChemistrySat.h:475-531). */
typedef struct atsat_consts {
    double re, pr, g, omega, gam, u_0, t_ref, p_ref, R_ref, L_atm;
    double Coriolis, centrifugal, buoyancy;
    double metric_radius_km;                  /* ATSAT_METRIC_RADIUS, 58232 */
    double r_mix, cp_mix, R_mix;
    double sc_ch4, sc_h2o, sc_h2s, sc_nh3, sc_nh4sh;
    /* per gas, index ATSAT_GAS_*: liquid and ice SVP sets, latent heats, phase temperatures */
    double t_0[3], t_00[3], ep[3], lv[3], ls[3];
    double C[3], L0[3], R[3], del_alf[3], del_bet[3];
    double C_i[3], L0_i[3], del_alf_i[3], del_bet_i[3];
    double t_0_nh4sh, t_00_nh4sh;             /* NH4SH crystal band (Precipitation::sedimentNH4SH) */
    /* the chemistry block (planet/ChemistrySat.h), planet/cSaturnModel.h:818-931 */
    double m_nh3, m_h2s, m_nh4sh;             /* molar masses [kg/kmol] */
    double c_mix;                             /* mixture molar density [kmol/m3] (ThermalPropertiesSat) */
    double mue_nh3, mue_h2s, mue_nh4sh;       /* dynamic viscosities [N s/m2] */
    double rg_nh3, rg_h2s, rg_nh4sh;          /* densities of the D = mue / (rg sc) coefficients [kg/m3] */
    double cp_nh3, cp_h2s, cp_nh4sh;          /* specific heats [J/(kg K)] */
    double k_nh3, k_h2s, k_nh4sh;             /* thermal conductivities [W/(m K)] */
    double ep_h2s;                            /* H2S's ep, a tracer of the convective adjustment */
} atsat_consts;

/* Knob values of the step. Everything the reference reads from ATSAT_* environment variables
 * on the hot path is frozen at its default (see the docstrings of the entry points); these
 * are the ones a caller may set. */
typedef struct atsat_knobs {
    double dt;                  /* ATSAT_DT, 0.01 (planet/cSaturnModel.cpp:171-174, 402) */
    int    turb_active;         /* turb_model != none && ATSAT_TURB != 0 (cSaturnModel.cpp:408) */
    double chemical_reaction;   /* XML; multiplies massflux_* in the species RHS */
    double turb_coupling;       /* ATSAT_TURB_COUPLING, 1.0 */
    double precip_coupling;     /* ATSAT_PRECIP_COUPLING, 1.0 */
    double pdyn_cap;            /* ATSAT_PDYN_CAP, 0.2 */
    int    press_sweeps;        /* ATSAT_PRESS_SWEEPS, 1 (used when press_mode == 0) */
    int    press_mode;          /* 0 = the reference's red-black sweeps; 1 = converged solve */
    double press_tol;           /* press_mode 1: relative residual target, 1e-10 */
    int    precip_on;           /* ATSAT_PRECIP, 1 */
    int    chem_inputs;         /* 1: read massflux_* / fluxlim_nh4sh; 0: they are zero */
    int    press_max_iter;      /* press_mode 1: most defect corrections per solve (8) */
    int    press_fp32_transforms; /* press_mode 1, FP32 context: 1 = the corrections' forward
                                     transforms in FP32; 0 = FP64 corrections */
    /* The turbulence closure, Turbulence<cSaturnModel> (planet/Turbulence.h), k-omega SST in its
     * free-flow form (ATSAT_TURB_FREE 1, no solid cells), as the model runs it by default. */
    int    turb_closure;        /* with turb_active: 1 = atsat_step runs the closure after
                                   restoreVar (cSaturnModel.cpp:1116); 0 = nue / tke_source /
                                   dis_source stay as they are */
    int    turb_buoy;           /* ATSAT_TURB_BUOY, 3 on ATSAT (cSaturnModel.h:299): 0 none, 1 dry
                                   stable sink, 2 dry both signs, 3 moist N^2 (Turbulence.h:802-870) */
    double turb_k_bg;           /* ATSAT_TURB_K_BG [m2/s2], 1.0 (Turbulence.h:156-159) */
    double turb_nue_max;        /* ATSAT_NUE_MAX [m2/s], 1e5 (Turbulence.h:123-126) */
    /* The chemistry block of the physics iteration (planet/cSaturnModel.cpp:933, 979-994) at the
     * model's defaults: NH4SH as an equilibrium adjustment (ATSAT_NH4SH_EQ 1), difflux = D
     * laplacian(c) (ATSAT_CHEM_DIFFLUX_LAPLACIAN 1) added to massflux (ATSAT_CHEM_DIFFLUX_PLUS 1). */
    int    chem_block;          /* 1 = atsat_step runs it on even iterations (and the RK4 stages
                                   read massflux_* / fluxlim_nh4sh, as with chem_inputs) */
    /* The moist convective adjustment (planet/ConvectiveAdjustment.h:402-849) with its sync
     * (cSaturnModel.cpp:1120-1171), ATSAT's defaults (Ledoux, loading, MCA_SMOOTH 2, interior). */
    int    mca_stride;          /* ATSAT_MCA: every N iterations, 0 = off (the reference's default 1) */
    /* The optional modes of the red-black solver (press_mode 0), off by default as in the
     * reference (planet/PressureSolver.h:290-315, 497-708, 750-755). */
    int    press_line;          /* ATSAT_PRESS_LINE: after each sweep a radial Thomas solve of every
                                   column (zebra in j + k, folded Neumann ends) and the domain mean of
                                   p_dyn held */
    int    press_baro;          /* ATSAT_PRESS_BAROTROPIC: after each sweep the column-mean residual
                                   solved exactly (DFT in phi, tridiagonal in theta per mode) */
} atsat_knobs;

#ifdef __cplusplus
}
#endif
#endif
