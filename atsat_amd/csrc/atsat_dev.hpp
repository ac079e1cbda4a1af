// Internal header of libatsat_gpu: device layout, kernel argument blocks and the launchers each
// .hip translation unit exports to the host layer (atsat_gpu.cpp). Not part of the C ABI.
//
// Layout in HBM. Every 3-D field of every member is one [i][j][kp] block, k fastest, with the
// row pitch kp = km rounded up to 64 elements so that every row starts on a 256-byte boundary
// and a wavefront's 64 lanes read whole cache lines. Blocks of one field are consecutive over
// members (field-major: buf[field][member][i][j][kp]), so the kernels index a member by one
// stride. The padding columns k = km..kp-1 are never read.
//
// Two storage classes. The STATE class -- the temperature t and the zonal velocity w and their
// start-of-step copies, p_dyn, p_stat, p_metric, rho_mix, p_baro and the Poisson scratch -- is
// FP64 in every precision: the quantities the step takes small differences of (the pressure
// gradients, the buoyancy anomaly, the saturation deficit, the balance of the jet's Coriolis and
// metric terms) are formed from it. The WORK class -- the other 15 prognostic fields (u, v, the
// species, k, omega) and their copies, the RK4 stage inputs and accumulators, the last-stage
// tendencies, the closure / chemistry inputs and the precipitation outputs -- is stored in the
// context's precision (FP64, or FP32 for the mixed-precision step); kernels widen it to FP64 on
// the load wherever they compute in FP64, and round once on the store (prog_state below).
#pragma once

#include <vector>

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "atsat_fields.h"

#include <type_traits>

namespace atsat {

// Internal buffers after the ATSAT_NFIELDS public ones.
enum {
    B_STAGE0 = ATSAT_NFIELDS,               // RK4 stage-input scratch X (17 fields), work class
    B_STAGE1 = B_STAGE0 + ATSAT_NPROG,      // second stage-input scratch Y (17 fields), work class
    B_ACC    = B_STAGE1 + ATSAT_NPROG,      // RK4 accumulators acc_* (RungeKutta_Sat_Turb.cpp:202)
    B_PSRC   = B_ACC + ATSAT_NPROG,         // Poisson source: clamped divergence + Lc(p_entry)
    B_TMP    = B_PSRC + 1,                  // scratch of the p_dyn filter
    B_PTERM  = B_TMP + 1,                   // the RK4 stages' pressure terms (PT_*), work class
    NBUF     = B_PTERM + 6
};

// The pressure terms of the RK4 stages' right-hand sides (RHS_Sat_Turb.cpp:154, 180, 206,
// 372-377, 814-817), each formed in FP64 from p_dyn, p_metric, p_stat and rho_mix and rounded once
// to the work class. The pressure fields and the density do not change during an RK4 step, so the
// four stages of the FP32 step read these six values per cell instead of thirteen FP64 pressure
// neighbours and the density.
enum {
    PT_DPDR = 0,        // d p_dyn / dr
    PT_DPDTHE,          // d p_dyn / dtheta
    PT_DPDPHI,          // d p_dyn / dphi
    PT_DPMDTHE,         // d p_metric / dtheta
    PT_DPMDPHI,         // d p_metric / dphi
    PT_DPSDR            // d p_stat / dr, nondimensional, times r_mix / rho_mix where rho_mix > 0
};

// The prognostic fields kept in the state class (FP64 in every precision): the temperature and
// the zonal velocity w, and their start-of-step copies. The saturation thresholds of the physics
// turn on T, and w is the O(1) jet whose Coriolis and metric terms balance the pressure gradient
// in the v and u equations, so a last-bit change of either is amplified. The other 15 (u, v, the
// eleven species, k and omega) are stored in the work class: FP32 in the FP32 step. Measured by
// rounding the live fields and their copies to FP32 after every iteration of the FP32 step, 100
// native iterations against the FP64 step: t alone moves h2o_cloud to 3e-4 and the condensed
// phase of 267 cells; w alone moves u to 9.3e-6 (from 5e-7); the other 15 together leave every
// field as it was below 1e-5 outside the cells whose phase flips.
__host__ __device__ constexpr bool prog_state(int f){ return f == ATSAT_T || f == ATSAT_W; }

__host__ __device__ constexpr bool is_state(int id){
    return (id < ATSAT_NPROG && prog_state(id)) || (id >= ATSAT_F_N0 && id < ATSAT_F_N0 + ATSAT_NPROG && prog_state(id - ATSAT_F_N0))
        || id == ATSAT_P_DYN || id == ATSAT_P_STAT || id == ATSAT_P_METRIC
        || id == ATSAT_RHO_MIX || id == ATSAT_P_BARO || id == B_PSRC || id == B_TMP;
}

// Division in a physics kernel of a context of work class T: the reference's a / b in the FP64
// step (bit-identical), a times the reciprocal rb = qrcp<T>(b), formed once, in the FP32 step,
// whose results are rounded to FP32 (an FP64 ulp of difference does not reach them). FP64
// divisions are a dozen instructions each and dominate the closure and chemistry kernels.
template<typename T>
__device__ __forceinline__ double qrcp(double b){
    if constexpr(std::is_same<T, float>::value) return 1.0 / b;
    else return 0.0;
}
template<typename T>
__device__ __forceinline__ double qdiv(double a, double b, double rb){
    if constexpr(std::is_same<T, float>::value) return a * rb;
    else return a / b;
}

// Per-(i,j) geometry, built on the host exactly as RungeKuttaSat builds CellGeometry
// (RungeKutta_Sat_Turb.cpp:129-143, 223-243) and PressureSolver its stencil weights
// (PressureSolver.h:156-162, 326-353), so device and host use bit-identical factors.
enum {
    G_RM = 0, G_COSTHE, G_COTANTHE, G_SINTHE2, G_INV_RM, G_INV_RM2, G_INV_RMSINTHE,
    G_INV_RM2SINTHE, G_INV_RM2SINTHE2, G_COSTHE_INV_RM2SINTHE, G_SINTHE_TRUE,
    G_CENT_RAD, G_CENT_THE, G_V_METRIC,
    // pressure solver (PressureSolver.h)
    G_P_NUM2, G_P_NUM3, G_P_DENOM, G_P_INV_RM, G_P_INV_RMSINTHE, G_P_LC_NUM2, G_P_LC_NUM3,
    NGEO
};

struct Dims {
    int im, jm, km, kp;          // kp: row pitch (elements)
    int M;                       // members
    long long fs;                // elements of one member's field block = im*jm*kp
    __host__ __device__ long long at(int i, int j, int k) const { return ((long long)i * jm + j) * kp + k; }
};

// Scalars every kernel may need, in double (FP32 kernels convert).
struct Scal {
    double dt, dr, dthe, dphi;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
    double re, pr, g, omega, gam, u_0, t_ref, p_ref, R_ref, L_atm;
    double Coriolis, centrifugal, buoyancy;
    double r_mix, cp_mix, R_mix;
    double sc_ch4, sc_h2o, sc_h2s, sc_nh3, sc_nh4sh;
    double nd_cor, coeff_energy_p, to_nd, precip_den, exp_pressure, pdyn_cap;
    double chemical_reaction, turb_coupling, precip_coupling;
    double tke_max_nd, dis_min_nd;
    double cor_metric;           // Coriolis * (L/u_0) * 2 * omega (metric pressure, without |sin|)
    double shift;                // metricRadius shift, R_km / L_atm - 1
    double dz_m;                 // layer thickness [m] of levels 0..im-2 (layer_thickness_m)
    double a_hydro;              // exp_pressure * gam * dz_km (rebuildPstatHydro)
    double buoy_wsum_full;       // serial sum of sin(theta_j) over every cell of a level
    double layer_h;              // the float layer spacing L_atm * dr [km] of init_layer_heights
    // the turbulence closure (Turbulence.h): caps and floors made dimensionless, the background
    // k / omega of init_fields
    double turb_nue_max_nd, turb_k_bg_nd, turb_init_tke, turb_init_dis;
    // the chemistry block (ChemistrySat.h): D_x = mue_x / (rg_x sc_x), molar masses, c_mix, the
    // species' heat capacities and conductivities, the NH4SH formation heat [J/kg]
    double D_nh3, D_h2s, D_nh4sh, m_nh3, m_h2s, m_nh4sh, c_mix;
    double cp_nh3, cp_h2s, cp_nh4sh, k_nh3, k_h2s, k_nh4sh, nh4sh_Lf;
    double ep_h2s;               // the H2S tracer of the moist convective adjustment
    int turb_active, chem_inputs, turb_buoy, chem_block;
    // per gas (CH4, H2O, NH3)
    double t_0[3], t_00[3], ep[3], lv[3], ls[3];
    double C[3], L0[3], R[3], del_alf[3], del_bet[3];
    double C_i[3], L0_i[3], del_alf_i[3], del_bet_i[3];
    double t_0_nh4sh, t_00_nh4sh, v_settle;
};

// Device pointers of every buffer (member 0; member m at +m*fs). s[id] is set for the state
// class (FP64), w[id] for the work class (T); the other entry is null.
template<typename T>
struct Bufs {
    double* s[NBUF];
    T* w[NBUF];
    double* geo;                 // [im*jm][NGEO]
    double* thez_sin;            // [jm] sin(the.z[j]) (glibc), for the level-mean weights
    double* sin_abs;             // [jm] |sin(the.z[j])| (metric pressure)
    double* rm;                  // [im] metricRadius(rad.z[i])
    double* vbar0;               // [M][im*jm]
    double* ubar0;               // [M][im*jm]
    double* buoy_ref;            // [M][im]
    double* srf;                 // [M][5][jm*km]
    double* vel_star;            // [M][jm*km] friction velocity of the closure
    double* mca;                 // scratch of the moist convective adjustment (k_mca.hip), or null
    unsigned char* mca_sf;       // [M][fs] 1 = the cell's conserved variables are valid
    double* pmodes;              // tables and scratch of the solver's optional modes (k_pmodes.hip), or null
    unsigned long long* counter; // census
    double* budget;              // [M][4][im] plane sums, then [M][3 + im] the budget sums
    double* buoy_x;              // [jm*km][buoy_ch] weighted buoyancy of every cell, level-major lanes
    long long buoy_ch;           // row pitch of buoy_x: members*im rounded up to 64
    // element n of buffer f of either class, as FP64 (ld), and a store rounded to the class (st);
    // with a compile-time f the class test folds away
    __device__ __forceinline__ double ld(int f, long long n) const { return is_state(f) ? s[f][n] : (double)w[f][n]; }
    __device__ __forceinline__ void st(int f, long long n, double x) const {
        if(is_state(f)) s[f][n] = x; else w[f][n] = (T)x;
    }
};

// ---- launchers (one per kernel family) ---------------------------------------------------
template<typename T> void launch_mixture_density(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_buoy_ref(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_metric_pressure(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_press_terms(const Dims&, const Scal&, const Bufs<T>&, bool with_static, hipStream_t);
template<typename T> void launch_rk_halo(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
struct RkTile;
template<typename T> void launch_rk_stage(const Dims&, const Scal&, const Bufs<T>&, const RkTile&, int stage, bool alias, bool clamp_species, hipStream_t);
template<typename T> struct RkGrid;
// the stage of ng grids in one launch; bit g of alias: grid g reads y_n from the live fields
template<typename T> void launch_rk_group(const RkGrid<T>* grids, int ng, int blocks, int threads, bool turb, int stage,
                                          unsigned alias, bool clamp_species, hipStream_t);
// the same stage as a persistent launch over the work queues of `heads` (kQueues unsigned
// counters in device memory) for the batched launch's `total` items
constexpr int kQueues = 8;
template<typename T> void launch_rk_queue(const RkGrid<T>* grids, int ng, int total, int threads, bool turb,
                                          int stage, unsigned alias, bool clamp_species, unsigned* heads, hipStream_t);
template<typename T> void launch_press_prepare(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_press_src(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_press_sweep(const Dims&, const Scal&, const Bufs<T>&, int colour, hipStream_t);
template<typename T> void launch_press_boundary(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_damp_pdyn(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_satadj(const Dims&, const Scal&, const Bufs<T>&, int gas_first, int gas_last, hipStream_t);
template<typename T> void launch_rebuild_pstat(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_precipitation(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_copy_n(const Dims&, const Scal&, const Bufs<T>&, unsigned mask, hipStream_t);
template<typename T> void launch_bc(const Dims&, const Scal&, const Bufs<T>&, bool clamp_species, hipStream_t);
template<typename T> void launch_clamp(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_census(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_budgets(const Dims&, const Scal&, const Bufs<T>&, double* part, double* out, hipStream_t);
// the turbulence closure: run() (init = false) or init() (init = true), k_turb.hip
template<typename T> void launch_turbulence(const Dims&, const Scal&, const Bufs<T>&, bool init, hipStream_t);
// the chemistry block, k_chem.hip: EquilibriumNH4SH("loop"), and the chemistry after the RK
// scalar sync (DiffMassFlux, ChemMassRate, FluxLimiterNH4SH and their filters)
template<typename T> void launch_nh4sh_equilibrium(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
template<typename T> void launch_chemistry(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
// the moist convective adjustment, k_mca.hip; its scratch (doubles) and the column limit of the
// kernel's 64-bit level masks
template<typename T> void launch_mca(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
size_t mca_scratch_doubles(const Dims&);
// the red-black solver's optional modes, k_pmodes.hip
template<typename T> void launch_press_line(const Dims&, const Scal&, const Bufs<T>&, int colour, hipStream_t);
template<typename T> void launch_press_baro(const Dims&, const Scal&, const Bufs<T>&, double A2, hipStream_t);
template<typename T> void launch_press_mean(const Dims&, const Bufs<T>&, int mode, hipStream_t);
size_t pmodes_doubles(const Dims&);
int pmodes_max_levels();
int pmodes_store_tables(const Dims&, const Scal&, const std::vector<double>& thez_sin, double* pm);
int mca_store_constants(const Dims&, const Scal&, double* mca);
constexpr int kMcaMaxLevels = 64;

// ---- the converged pressure solve (k_solve.hip) -------------------------------------------
constexpr int kSolveRow = 256;                  // threads per (i, j) row of the solver kernels
struct Solver;
Solver* solver_create(const Dims&, const Scal&, const double* thez_sin, const double* rm, hipStream_t, std::string* err);
void solver_destroy(Solver*);
int solver_prepare(Solver*, bool fp32);
int solver_warmup(Solver*, bool fp32);
// Lc p = src to the relative residual tol (host_check) or with max_iter corrections (no host
// synchronisation, for graph capture; the final residual is then compared on the device with
// status_limit, see solver_status). src is overwritten; scratch is one FP64 field.
int solver_run(Solver*, double* p, double* src, double* scratch, const double* geo, bool fp32, double tol,
               int max_iter, bool host_check, double status_limit);
// The worst end of the solves run without a host check since the last call (0 ok, 1 a residual
// above its limit, 2 non-finite), read and cleared; synchronous.
int solver_status(Solver*, int* status);
void solver_last(const Solver*, int* iters, double* rel);
// whether the last checked solve stopped at the FP64 rounding floor (its last correction no
// longer halved the residual) rather than at the tolerance
bool solver_last_floor(const Solver*);
const char* solver_error(const Solver*);
void solver_set_stream(Solver*, hipStream_t);
size_t solver_device_bytes(const Solver*);

// ---- device helpers ------------------------------------------------------------------------
__device__ __forceinline__ int phi_wrap(int kk, int km){      // ATPhys.h:135-139, periodic
    const int n = km - 1;
    return kk < 0 ? kk + n : (kk >= n ? kk - n : kk);
}

// The seven stencil values of an FP64 field at a cell, as loaded: the centre, r +- 1, theta +- 1,
// the east and west phi neighbours. Kernels that read several stencils load them all into these
// before using any, so that a wavefront waits for its loads once rather than field by field.
struct Raw7d { double c, rp, rm, tp, tm, pe, pw; };
template<typename E>
__device__ __forceinline__ Raw7d raw7(const E* f, long long c, long long ip, long long im, long long jp,
                                      long long jm, long long pe, long long pw){
    return Raw7d{(double)f[c], (double)f[ip], (double)f[im], (double)f[jp], (double)f[jm], (double)f[pe], (double)f[pw]};
}

// Grid mapping of the per-cell RK4 stage kernel: every workgroup holds `rings` whole longitude
// rings (consecutive rows of one band, one level, one member). Consecutive workgroups go
// round-robin to the 8 XCDs; each XCD owns a contiguous range of bands of `band` rows and walks a
// band's rows fastest, then its levels, so a ring's j neighbours were loaded by a neighbouring
// workgroup and its r neighbours `band / rings` workgroups away, inside the XCD's L2.
struct RowMap {
    int nj, nlev, M, band, nband, per;   // interior rows, levels, members, rows per band, bands
                                         // per member, bands per XCD
    int rings;                           // rings (consecutive rows of a band) per workgroup
    // Workgroup b, ring `sub` of it -> (row jr, level lev, member m); false past the grid.
    __device__ __forceinline__ bool decode(int b, int sub, int& jr, int& lev, int& m) const {
        const int l = b >> 3;
        const int wpl = band / rings;                // workgroups per band and level
        const int seq = wpl * nlev;                  // workgroups per band
        const int g = (b & 7) * per + l / seq;       // band over (members x bands)
        const int r = l % seq;
        lev = r / wpl;
        if(g >= nband * M) return false;
        m = g / nband;
        jr = (g % nband) * band + (r % wpl) * rings + sub;
        return jr < nj;
    }
    // band: rows per band, rounded up to a multiple of rings
    void set(int b, int r){
        rings = r < 1 ? 1 : r;
        band = (b < rings ? rings : b) / rings * rings;
        nband = (nj + band - 1) / band; per = (nband * M + 7) / 8;
    }
    int blocks() const { return per * (band / rings) * nlev * 8; }
};


// Workgroup -> interior (i, j) row (i 1..im-2, j 1..jm-2) of the kernels that give one longitude
// ring to a workgroup and read their neighbours through the caches (the closure's sources, the
// chemistry's fluxes): the stage kernel's XCD-contiguous band map, so that a row's theta and r
// neighbour rows are read by workgroups of the same XCD (its L2) shortly before or after it.
constexpr int kInteriorBand = 8;
inline RowMap interior_row_map(const Dims& D, int band = kInteriorBand){
    RowMap RM{D.jm - 2, D.im - 2, D.M, 0, 0, 0, 1};
    RM.set(band, 1);
    return RM;
}

// The same over every (i, j) row (the converged solve's residual and defect kernels).
inline RowMap all_row_map(const Dims& D, int band = kInteriorBand){
    RowMap RM{D.jm, D.im, D.M, 0, 0, 0, 1};
    RM.set(band, 1);
    return RM;
}

// Tile of the per-cell stage kernel: rows per band and rings per workgroup (the workgroup is
// `rings` rings of km - 1 lanes, each rounded up to a wavefront). The stencil halo is one cell
// in each direction (7-point stencils), loaded from L1 / L2, so there is no halo-depth choice.
struct RkTile { int band, rings; };
#ifndef ATSAT_ROW_THREADS
#define ATSAT_ROW_THREADS 768
#endif
constexpr int kRowThreads = ATSAT_ROW_THREADS;   // workgroup size limit of the per-cell stage kernel

// One grid of a batched stage launch (k_rk_group), read by the kernel from device memory.
template<typename T>
struct RkGrid {
    Dims D;
    Scal S;
    Bufs<T> B;
    RowMap RM;
    int rt;                      // threads per ring
    int first_block;             // first workgroup of this grid in the launch, a multiple of 8
};

int rk_ring_threads(const Dims&);
RowMap rk_row_map(const Dims&, const RkTile&);
bool rk_tile_ok(const Dims&, const RkTile&);    // rings divide band, workgroup within the limit
RkTile rk_default_tile(const Dims&);

}  // namespace atsat
