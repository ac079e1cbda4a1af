/*
 * atsat_gpu.h -- C ABI of the MI355X model time step of ATSAT.
 *
 * The reference has no plugin or FFI boundary for this path: the step is four member calls on
 * the monolithic cSaturnModel, each mutating Array members in place. Each entry point below
 * replaces one of those calls (the reference function is cited at the declaration); the
 * binding a maintainer would add on the reference side is shown in INTEGRATION.md.
 *
 * Conventions
 *   - every function returns int: 0 = ok, < 0 = an ATSAT_E_* code; atsat_last_error() gives text;
 *   - nothing throws or exits across the ABI;
 *   - one context per host thread, not re-entrant; one HIP stream per context;
 *   - a context holds n_members independent members on one grid (an ensemble); every step entry
 *     point advances all members in the same launches;
 *   - host buffers are dense [im][jm][km] doubles, k fastest (the reference's Array::x[i][j][k]
 *     order). Device buffers are owned by the context.
 */
#ifndef ATSAT_GPU_H
#define ATSAT_GPU_H

#include <stddef.h>
#include "atsat_fields.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ATSAT_OK = 0,
    ATSAT_E_ARG = -1,        /* bad argument (null pointer, out-of-range id, bad grid) */
    ATSAT_E_HIP = -2,        /* a HIP runtime call failed */
    ATSAT_E_NOMEM = -3,      /* device allocation failed */
    ATSAT_E_UNSUPPORTED = -4,/* knob value or precision this build does not implement */
    ATSAT_E_NONFINITE = -5,  /* the census found a non-finite value */
    ATSAT_E_NOCONV = -6,     /* the converged pressure solve missed its tolerance */
    ATSAT_E_IO = -7          /* a restart file could not be opened, read or written, or its
                                header or grid does not match */
};

/* Precision of a context. FP64: every field in FP64 (the reference's arithmetic). FP32: the
 * stencils, RK4 scratch and the right-hand sides in FP32, and of the prognostic fields t and w in
 * FP64, the other 15 (u, v, the species, k, omega) stored in FP32; the pressures, rho_mix, the
 * reductions, the pressure defect correction and the physics arithmetic stay FP64. Uploads and
 * downloads are FP64 on the host side in both. */
enum { ATSAT_PREC_FP64 = 0, ATSAT_PREC_FP32 = 1 };

typedef struct atsat_ctx atsat_ctx;

/* Host-side state of one member: pointers to dense [im*jm*km] (fields) or small (vecs) arrays.
 * NULL entries are skipped by the bulk copies. */
typedef struct atsat_host_state {
    double* f[ATSAT_NFIELDS];
    double* v[ATSAT_NVECS];
} atsat_host_state;

/* ---- defaults (the reference's values) ------------------------------------------------- */
/* 41 x 181 x 361, dr 0.025, 1 deg: planet/cSaturnModel.h:711, planet/cSaturnModel.cpp:42-52 */
int atsat_default_grid(atsat_grid* g);
/* saturn/config_atsat.xml (re = 1e5 as shipped) + planet/cSaturnModel.h:721-1056 +
 * ThermalPropertiesSat planet/ChemistrySat.h:475-531 */
int atsat_default_consts(atsat_consts* c);
/* The reference's defaults (planet/cSaturnModel.cpp:93-102, 171-174, 407-408;
 * saturn/config_atsat.xml:17-19; RHS_Sat_Turb.cpp:577-578, 645-650; PressureSolver.h:168-175):
 * the k-omega SST closure, the chemistry block and the moist convective adjustment on, the
 * pressure equation relaxed by one red-black sweep per even iteration (press_mode 0). */
int atsat_default_knobs(atsat_knobs* k);
/* The reference run with ATSAT_TURB=0 ATSAT_MCA=0 and chemical_reaction 0 in the XML: the
 * closure, the chemistry block and the convective adjustment off, everything else as
 * atsat_default_knobs. No counterpart function in the reference (it is a choice of environment
 * and configuration values there). */
int atsat_resident_knobs(atsat_knobs* k);

/* ---- synthetic inputs -------------------------------------------------------------------- */
/* The seeded synthetic start state "gen_v1" of one member (host, FP64): parabolic
 * pole-to-pole T with a gam lapse rate, analytic zonal jets, vapour at 0.9 of saturation below
 * the cloud base decaying above it, k = 1e-4, omega = 1e-3, p_stat rebuilt hydrostatically
 * (InitVariables_Sat.cpp:556-578), p_dyn / p_baro from the geostrophic balance of
 * initGeoBalance (InitVariables_Sat.cpp:345-453), cell_vbar0 / ubar0 from initCellHold
 * (InitVariables_Sat.cpp:1126-1139). t_sigma [K] and w_sigma [m/s] are the member's white-noise
 * perturbations. Every field of `out` must be allocated (NULL entries are skipped). */
int atsat_gen_v1(const atsat_grid* g, const atsat_consts* c, unsigned long long seed,
                 int member, double t_sigma, double w_sigma, atsat_host_state* out);

/* ---- context ----------------------------------------------------------------------------- */
int atsat_ctx_create(const atsat_grid* g, const atsat_consts* c, const atsat_knobs* k,
                     int n_members, int precision, atsat_ctx** out);
int atsat_ctx_destroy(atsat_ctx* ctx);
const char* atsat_last_error(atsat_ctx* ctx);        /* ctx may be NULL: the global message */
int atsat_set_knobs(atsat_ctx* ctx, const atsat_knobs* k);
/* Bytes of device memory the context holds. */
int atsat_device_bytes(atsat_ctx* ctx, size_t* bytes);

/* Copy one field (ATSAT_* id) of one member between the host and the device. Fields held in
 * FP32 on the device are converted. The host buffer is dense [im][jm][km] (vecs: see
 * atsat_fields.h). These synchronise the context's stream. */
int atsat_upload(atsat_ctx* ctx, int member, int field, const double* src);
int atsat_download(atsat_ctx* ctx, int member, int field, double* dst);
int atsat_upload_vec(atsat_ctx* ctx, int member, int vec, const double* src);
int atsat_download_vec(atsat_ctx* ctx, int member, int vec, double* dst);
/* Every non-NULL entry of the state, both directions. */
int atsat_upload_state(atsat_ctx* ctx, int member, const atsat_host_state* s);
int atsat_download_state(atsat_ctx* ctx, int member, const atsat_host_state* s);
/* Drop-in mode: bind the reference's own jagged rows (Array::x[i][j], im*jm pointers to km
 * doubles, lib/Array.h:25-37) for a field; upload_rows / download_rows gather / scatter them. */
int atsat_upload_rows(atsat_ctx* ctx, int member, int field, double* const* rows);
int atsat_download_rows(atsat_ctx* ctx, int member, int field, double* const* rows);

/* ---- the four hot-path calls (all members at once) --------------------------------------- */
/* Replaces void cSaturnModel::RungeKuttaSat() (planet/cSaturnModel.h:1093;
 * planet/RungeKutta_Sat_Turb.cpp:110-387): the buoyancy reference level (:31-50), the metric
 * pressure (:60-108) and the four separated RK4 stages over RHSSat (RHS_Sat_Turb.cpp:25-1016)
 * with the final assembly and the k / omega clamps. */
int atsat_rk4_step(atsat_ctx* ctx);
/* Replaces void PressureSolver<cSaturnModel>::run() (planet/PressureSolver.h:138, 145-766)
 * including cSaturnModel::prepareProjectionBoundaries (planet/Pressure_Sat.cpp:78-194) and the
 * boundary pass. press_mode 0: knobs.press_sweeps red-black sweeps exactly as the reference;
 * press_mode 1: solve to knobs.press_tol relative residual (converged mode). */
int atsat_pressure_solve(atsat_ctx* ctx);
/* Replaces void SaturationAdjustmentSat::run(std::string gas, ...) (SaturationAdjustmentSat.h:
 * 153-161, .cpp:19-49) -> SaturationAdjustment<cSaturnModel>::run (SaturationAdjustment.h:
 * 255-746). gas is ATSAT_GAS_*; the coefficient quadruples live in atsat_consts. */
int atsat_saturation_adjust(atsat_ctx* ctx, int gas);
/* Replaces void Precipitation<cSaturnModel>::run() (planet/Precipitation.h:162, 191-270). */
int atsat_precipitation(atsat_ctx* ctx);

/* The converged solve of the pressure equation (press_mode 1 uses it inside
 * atsat_pressure_solve). No counterpart in the reference, whose PressureSolver<>::run
 * (planet/PressureSolver.h:145-766) only relaxes with a fixed number of red-black sweeps: solves
 * the reference's discrete operator (:350-353, 476-491, with the boundary pass :718-740 closing
 * it) for p_dyn to the relative residual tol, by FP64 defect correction around a separable
 * direct solver (in an FP32 context its forward transforms run in FP32, the rest in FP64). src
 * is dense [members][im][jm][km] (only the interior cells are read), or NULL for the step's own
 * source (the clamped divergence of rhs_* plus Lc(p_dyn)). The source's component along the
 * operator's left null vector is removed, and p_dyn keeps its weighted mean. The solve also
 * stops when a correction no longer halves the residual: the FP64 rounding floor of the residual
 * (about eps * den * |p| / |src|) lies above 1e-10 for sources with large-scale content, since
 * the smoothest non-constant mode of the operator has an eigenvalue of ~1e-4. Returns
 * ATSAT_E_NOCONV if neither happens in max_iter corrections. Synchronises the stream. */
int atsat_poisson_solve(atsat_ctx* ctx, const double* src, double tol, int max_iter, int* iters,
                        double* rel_residual);
/* Corrections and relative residual of the last converged solve (-1 if it ran inside a captured
 * graph, which does not check its residual on the host: there the device compares the final
 * residual with the tolerance, or with four times the last checked solve's residual where that one
 * stopped at the FP64 rounding floor, and atsat_synchronize reports the outcome). */
int atsat_pressure_last(atsat_ctx* ctx, int* iters, double* rel_residual);

/* ---- the turbulence closure ----------------------------------------------------------------- */
/* Replaces void Turbulence<cSaturnModel>::run() (planet/Turbulence.h:306-328; called at
 * planet/cSaturnModel.cpp:1116 when turb_active): u_tau, the k-omega SST sources with the
 * production tensor and the buoyancy sink (compute_sources :591-882, compute_k_omega_SST
 * :1020-1107), the deep-boundary copy (apply_wall_bc :382-424), the nu_t clamp (clamp_nue
 * :542-562), the seam copy (:512-540) and the lid corners (fill_lid_corners :492-510), for the
 * model's defaults: k_omega_SST, ATSAT_TURB_FREE 1, ATSAT_TURB_CLAMP_PASS 1, ATSAT_ABL_TAPER 0,
 * ATSAT_TURB_CURV 1, no solid cells; knobs.turb_buoy / turb_k_bg / turb_nue_max. Writes tke, dis,
 * nue, prod, tke_source, dis_source, the seam and corner cells of tken / disn, and vel_star.
 * Needs turb_active. */
int atsat_turbulence(atsat_ctx* ctx);
/* Replaces void Turbulence<cSaturnModel>::init() (planet/Turbulence.h:282-303; called at set-up,
 * planet/cSaturnModel.cpp:624): the background k and omega in every interior cell (init_fields
 * :426-490, tken / disn with them), then the pieces of run() in init's order. Needs turb_active. */
int atsat_turbulence_init(atsat_ctx* ctx);

/* ---- the chemistry block ------------------------------------------------------------------- */
/* Replaces ChemistrySat::EquilibriumNH4SH("loop") (planet/ChemistrySat.h:294-331, called at
 * planet/cSaturnModel.cpp:933 with ATSAT_NH4SH_EQ 1): NH3(g) + H2S(g) <-> NH4SH(s) brought to
 * p_NH3 p_H2S = K_p(T) in every cell, with its formation heat (equilibrateNH4SHCell :250-289).
 * Writes nh3, h2s, nh4sh, t. */
int atsat_nh4sh_equilibrium(atsat_ctx* ctx);
/* Replaces the chemistry after the RK scalar sync, planet/cSaturnModel.cpp:979-994:
 * ChemistrySat::DiffMassFluxSat (ChemistrySat.h:365-472, difflux = D laplacian(c) with
 * ATSAT_CHEM_DIFFLUX_LAPLACIAN 1, thermalmassflux), AtomUtils::damp_wiggles of difflux_* and
 * thermalmassflux, ChemMassRateSat (:46-205; with ATSAT_NH4SH_EQ 1 the rates are zero and
 * massflux_x = difflux_x, ATSAT_CHEM_DIFFLUX_PLUS 1), FluxLimiterNH4SH (FluxLimiter.h:90-190)
 * and damp_wiggles of massflux_* and fluxlim_nh4sh. */
int atsat_chemistry(atsat_ctx* ctx);

/* ---- the moist convective adjustment ----------------------------------------------------- */
/* Replaces MoistConvectiveAdjustmentSat(*this, gases, tracers).run() with the MCA sync
 * (planet/cSaturnModel.cpp:1120-1171; planet/ConvectiveAdjustment.h:402-849) at ATSAT's defaults:
 * the Ledoux criterion with condensate loading, the search on a column smoothed in latitude
 * (MCA_SMOOTH 2), the radial boundary planes left out, no relaxation, 64 sweeps at most; gases
 * H2O, NH3, CH4, tracers H2S and NH4SH (with its formation heat). Writes t, the nine condensable
 * fields, h2s and nh4sh of the cells it adjusts. Columns of at most 64 levels. Allocates its
 * scratch on the first call (not inside a graph capture) unless knobs.mca_stride was set at
 * creation. */
int atsat_convective_adjustment(atsat_ctx* ctx);

/* ---- the step glue the reference runs between those calls -------------------------------- */
/* cSaturnModel::computeMixtureDensity (planet/Pressure_Sat.cpp:49-62) */
int atsat_mixture_density(atsat_ctx* ctx);
/* AtomUtils::damp_wiggles(p_dyn, nullptr, true, false, true) (lib/Utils.cpp:229-325),
 * ATSAT_DAMP_PDYN = 3 (planet/cSaturnModel.cpp:894-897) */
int atsat_damp_pdyn(atsat_ctx* ctx);
/* No counterpart in the reference: the filter above on memory the caller owns, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops). p_dyn is a device pointer
 * to a dense [n_members][im][jm][km] double array (k fastest, row pitch km), filtered in place;
 * stream is a hipStream_t (NULL: the null stream); the call is asynchronous on it. No context,
 * no allocation: the intermediate field lives in a scratch array the library holds on each
 * device for the life of the process, at most 8,388,608 doubles (n_members x im x jm x km;
 * three members of the native grid); larger requests return ATSAT_E_ARG. Every call on one
 * device shares that scratch, so calls that overlap in time on one device (other streams,
 * other host threads) race; the caller orders them. */
int atsat_damp_pdyn_noalloc(const atsat_grid* g, int n_members, double* p_dyn, void* stream);
/* cSaturnModel::rebuildPstatHydro (planet/InitVariables_Sat.cpp:556-578) */
int atsat_rebuild_pstat(atsat_ctx* ctx);
/* The RK scalar sync after the physics block (planet/cSaturnModel.cpp:955-977) */
int atsat_rk_scalar_sync(atsat_ctx* ctx);
/* bcTheta, bcPhi, bcRadius + the species lid / wall copies (BoundaryConditions.h:151-518,
 * BC_Sat.cpp:61-103), in the loop's order (cSaturnModel.cpp:1091-1093) */
int atsat_boundary_conditions(atsat_ctx* ctx);
/* cSaturnModel::clampNegativeSpecies (planet/FileIO_Sat.cpp:716-770) */
int atsat_clamp_species(atsat_ctx* ctx);
/* cSaturnModel::restoreVar(1.0) (planet/cSaturnModel.cpp:1551-1602) */
int atsat_restore_var(atsat_ctx* ctx);

/* One iteration of the hot path, in the order of cSaturnModel::Run's loop
 * (planet/cSaturnModel.cpp:864-1171) with radiation off: on even `iter` the physics block
 * (density, pressure solve, p_dyn filter, saturation adjustment CH4 -> H2O -> NH3, p_stat rebuild,
 * with knobs.chem_block the NH4SH equilibrium, precipitation, RK scalar sync, with
 * knobs.chem_block the chemistry), then the RK4 step, the boundary conditions, the species clamp,
 * restoreVar, with turb_active and knobs.turb_closure the turbulence closure, and every
 * knobs.mca_stride iterations the moist convective adjustment. Asynchronous on the context's
 * stream. */
int atsat_step(atsat_ctx* ctx, int iter);
/* n_steps iterations starting at iteration first_iter; use_graph = 1 captures one even+odd
 * pair of iterations in a hipGraph once and replays it. The same results as atsat_step called
 * for each iteration; inside one call the FP32 step's odd iterations reuse the pressure terms
 * the even iteration before formed (atsat_step forms them every time). */
int atsat_run(atsat_ctx* ctx, int first_iter, int n_steps, int use_graph);
/* Waits for the context's stream. Returns ATSAT_E_NONFINITE (ATSAT_E_NOCONV) if a converged
 * pressure solve replayed from a graph since the last synchronisation ended with a non-finite
 * residual (a residual above its limit); the outcome is then cleared. */
int atsat_synchronize(atsat_ctx* ctx);

/* ---- launch tiles and batches of grids ----------------------------------------------------
 * No counterpart in the reference, which steps one compile-time grid with OpenMP loops
 * (planet/cSaturnModel.h:711). For ensembles of several grid sizes advanced together. */

/* The RK4 stage kernel's tile: rows per theta band and rings (rows of km - 1 longitudes) per
 * workgroup; rings must divide band and rings x (km - 1 rounded up to 64) <= 768 threads.
 * atsat_autotune times every valid tile on rounds of the four RK4 stages from the context's
 * current state, which it saves to a temporary device buffer and restores, and keeps the
 * fastest; stage_ms is that tile's mean time per stage launch. The results of a step do not
 * depend on the tile. */
int atsat_autotune(atsat_ctx* ctx, int* band, int* rings, double* stage_ms);
int atsat_set_tile(atsat_ctx* ctx, int band, int rings);
int atsat_get_tile(atsat_ctx* ctx, int* band, int* rings);

typedef struct atsat_group atsat_group;
/* A batch of 1..32 contexts -- any grids and member counts, one precision and turb_active --
 * advanced together: per iteration each context's physics block and RK4 preparation on its own
 * stream, then the four RK4 stages of all of them as one launch each, from a table of every
 * grid's shape, tile and buffers in device memory (or, where the autotuner finds it faster,
 * every context's whole iteration on its own stream, the streams side by side). The caller keeps ownership of the contexts;
 * they must outlive the group and must not be stepped on their own while a group run is in
 * flight. Each context's results equal those of atsat_run on that context alone. */
int atsat_group_create(atsat_ctx* const* ctxs, int n, atsat_group** out);
int atsat_group_destroy(atsat_group* g);
const char* atsat_group_last_error(atsat_group* g);
/* The batch's autotuner. For each workgroup width (multiples of 64 from the longest ring up to
 * 768 threads) each grid takes as many rings per workgroup as fit and its band is timed alone;
 * the width whose batched four stages are fastest is the batched candidate. The other candidate
 * steps every context on its own stream with the best tile of its own, the streams side by side.
 * The faster of the two is kept (atsat_group_merged: 1 for the batched launches), with its tiles
 * (also set on the contexts). Every context's state is saved and restored around the timing.
 * stage_ms: the kept layout's mean time per stage (all grids). */
int atsat_group_autotune(atsat_group* g, double* stage_ms);
int atsat_group_tile(atsat_group* g, int i, int* band, int* rings, int* threads);
/* The layout: 0 side by side, 1 one batched launch per stage, 2 the batched stage as a persistent
 * launch whose workgroups (as many as the CUs hold) take the batched launch's workgroup items from
 * eight per-XCD work queues in device memory, starting with their own XCD's and then draining the
 * others. The autotuner times all three and keeps the fastest. */
int atsat_group_merged(atsat_group* g, int* merged);
/* Force the layout (0, 1 or 2, as above), keeping the contexts' tiles. */
int atsat_group_set_merged(atsat_group* g, int merged);
/* n_steps iterations of every context from iteration first_iter, with atsat_run's semantics
 * (use_graph = 1 captures one even+odd pair of the whole batch once and replays it).
 * Asynchronous; atsat_group_synchronize waits for the group's and the contexts' streams. */
int atsat_group_run(atsat_group* g, int first_iter, int n_steps, int use_graph);
int atsat_group_synchronize(atsat_group* g);

/* Per-kernel timing with HIP events on the context's stream (not inside a captured graph).
 * While enabled, each launch of a kernel family is bracketed by events; read returns the summed
 * elapsed time and launch count since enabling (it synchronises the stream). No counterpart in
 * the reference, which prints std::chrono timers per routine (RungeKutta_Sat_Turb.cpp:381-383). */
enum {
    ATSAT_K_RK_STAGE = 0,   /* the four fused RHS + RK4 stage kernels */
    ATSAT_K_RK_AUX,         /* buoyancy level mean, metric pressure, stage halo copy */
    ATSAT_K_PRESS,          /* projection boundaries, source, red-black sweeps, boundary pass */
    ATSAT_K_DAMP,           /* p_dyn Shapiro filter */
    ATSAT_K_SATADJ,         /* saturation adjustment */
    ATSAT_K_PSTAT,          /* hydrostatic p_stat rebuild */
    ATSAT_K_PRECIP,         /* precipitation columns */
    ATSAT_K_GLUE,           /* boundary conditions, clamp + restore, RK scalar sync */
    ATSAT_K_DENSITY,        /* mixture density */
    ATSAT_K_TURB,           /* turbulence closure */
    ATSAT_K_CHEM,           /* chemistry block: NH4SH equilibrium, mass fluxes, flux limiter, filters */
    ATSAT_K_MCA,            /* moist convective adjustment */
    ATSAT_NKERNELS
};
int atsat_profile_enable(atsat_ctx* ctx, int on);
int atsat_profile_read(atsat_ctx* ctx, int kernel, double* ms, long long* launches);

/* Census: count non-finite values of the prognostic fields of every member
 * (FileIO_Sat.cpp:418-470 is the reference's). */
int atsat_nonfinite_count(atsat_ctx* ctx, long long* count);

/* Budgets: the sums of the reference's two budget instruments, which it only prints, as FP64
 * reductions on the device for every member. out holds n_members x (3 + im) doubles; member m's
 * are out[m (3 + im) + g], g = 0, 1, 2, the H2O, NH3 and CH4 columns (vapour + cloud + ice) in
 * g/m2, ATSAT_CWB_DIAG's cwb_col (planet/cSaturnModel.cpp:776-787), and out[m (3 + im) + 3 + i],
 * the level-mean temperature of level i in K, ATSAT_TB_DIAG's tb_lev over all rows (:656-666);
 * both weighted by sin(colatitude) over sum_j sin(colatitude_j) km. Synchronous. */
int atsat_budgets(atsat_ctx* ctx, double* out);

/* ---- restart files --------------------------------------------------------------------------
 * The reference's binary restart (cSaturnModel::save_state / load_state, planet/FileIO_Sat.cpp:
 * 513-660): a header of five int32 (magic "SAT3" 0x53415433 or "SAT1" 0x53415431, im, jm, km,
 * the iteration), then whole arrays as im*jm rows of km doubles, in the order of
 * restart_arrays() (:513-527) and, for SAT2 / SAT3, restart_arrays_carried() (:538-554). The
 * arrays the library does not hold are "extras": the previous dynamic pressure p_dynn, which
 * restoreVar sets equal to p_dyn (cSaturnModel.cpp:1566-1569), and diagnostics the step does not
 * read at the model's defaults (the chemistry's cell fluxes j_* and jT_*, the rate-law rates w_*,
 * zero with the NH4SH equilibrium, the force and sensible-heat diagnostics). */
enum {
    ATSAT_RX_P_DYNN = 0,
    ATSAT_RX_J_H2S, ATSAT_RX_J_NH3, ATSAT_RX_J_NH4SH,
    ATSAT_RX_W_H2S, ATSAT_RX_W_NH3, ATSAT_RX_W_NH4SH,
    ATSAT_RX_BUOYANCY_FORCE, ATSAT_RX_PRES_GRAD_FORCE, ATSAT_RX_Q_SENSIBLE,
    ATSAT_RX_JT_H2S, ATSAT_RX_JT_NH3, ATSAT_RX_JT_NH4SH,
    ATSAT_RX_N
};
/* Dense [im*jm*km] host arrays of the extras, or NULL entries. Writing, a NULL entry writes
 * p_dyn for p_dynn and zeros for the others; reading, a NULL entry skips the array. */
typedef struct atsat_restart_extra {
    double* a[ATSAT_RX_N];
} atsat_restart_extra;

/* Host side, no device: write / read a restart file of one member held in dense host arrays
 * (every field the format lists must be present in s when writing; NULL fields are skipped
 * when reading). format 3 writes SAT3 (the reference's default), 1 writes SAT1
 * (ATSAT_RESTART_CARRIED=0: restart_arrays() only). Reading accepts SAT1, SAT2 (no
 * S_precip_ch4) and SAT3, leaves the arrays a format does not hold as they were, as load_state
 * does, and returns the file's iteration and format; trailing data (the chromophore arrays of an
 * ATSAT_CHROM run) is ignored. x may be NULL. */
int atsat_restart_write_host(const char* path, const atsat_grid* g, int iter, const atsat_host_state* s,
                             const atsat_restart_extra* x, int format);
int atsat_restart_read_host(const char* path, const atsat_grid* g, atsat_host_state* s,
                            atsat_restart_extra* x, int* iter, int* format);
/* The same for one member of a context, through the device: save_state(iter) writes the state
 * the member holds (call it at the end of an iteration, as the reference does, cSaturnModel.cpp:
 * 1309-1333); load_state reads a file into the member (the loop resumes at iter + 1, keeping the
 * parity of the physics block, cSaturnModel.cpp:625-634). Both synchronise the stream. */
int atsat_save_restart(atsat_ctx* ctx, int member, int iter, const char* path, const atsat_restart_extra* x,
                       int format);
int atsat_load_restart(atsat_ctx* ctx, int member, const char* path, atsat_restart_extra* x, int* iter,
                       int* format);

#ifdef __cplusplus
}
#endif
#endif
