// Host layer of libatsat_gpu: the C ABI of include/atsat_gpu.h over the HIP kernels.
//
// A context owns, for n_members members on one grid, every field of the step in HBM in the
// layout described in atsat_dev.hpp, a HIP stream, the per-(i,j) geometry table and a pinned
// staging buffer for host copies. Entry points enqueue on the stream; the copies synchronise.
#include "atsat_gpu.h"
#include "atsat_dev.hpp"
#include "restart.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace atsat;

namespace {

thread_local std::string g_err;

int fail(std::string* where, int code, const std::string& msg){
    if(where) *where = msg;
    g_err = msg;
    return code;
}

#define HIPC(x) do { hipError_t e_ = (x); if(e_ != hipSuccess) \
    return fail(&err, ATSAT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while(0)

}  // namespace

void atsat::set_last_error(const std::string& msg){ g_err = msg; }

struct atsat_ctx {
    atsat_grid grid{};
    atsat_consts consts{};
    atsat_knobs knobs{};
    int precision = ATSAT_PREC_FP64;
    Dims D{};
    Scal S{};
    hipStream_t stream = nullptr;
    // a second stream for work that runs beside the main stream's (forked and joined by events,
    // also inside a captured graph): the metric pressure and the stage halo rows run there while
    // the ordered level-mean sum, which occupies few CUs, runs on the main stream; and the pressure
    // solve beside the column physics of the even iteration
    hipStream_t side = nullptr;
    hipEvent_t ev_side_fork = nullptr, ev_side_join = nullptr;
    hipEvent_t ev_phys_fork = nullptr, ev_phys_join = nullptr;
    std::string err;
    size_t elem = 8;                       // bytes per work-class element
    // The start-of-step copies equal the live state (restoreVar ran last): the RK4 stages then
    // read y_n from the live fields and the *n buffers are not maintained (see k_rk.hip).
    bool n_alias = false;
    // The iteration of the model loop whose physics the static pressure terms (B_PTERM: p_dyn,
    // p_stat, rho_mix) were formed after, or -1. Only the loop keeps them from an even iteration
    // for the odd one after it, none of whose work writes those fields; every entry point clears
    // it (check_ctx), so anything a caller changes in between is seen.
    int pterm_iter = -1;
    RkTile tile{8, 1};                     // launch tile of the RK4 stage kernel (atsat_autotune)
    void* arena = nullptr;                 // all field buffers
    size_t arena_bytes = 0;
    double* d_small = nullptr;             // geometry + tables + vecs
    size_t small_bytes = 0;
    void* mca_mem = nullptr;               // the convective adjustment's scratch (when it runs)
    size_t mca_bytes = 0;
    void* pm_mem = nullptr;                // the solver modes' tables and scratch (when asked for)
    size_t pm_bytes = 0;
    double pm_A2 = 0.0;                    // sum over i 1..im-2 of 1 / rm^2 (PressureSolver.h:579-580)
    void* pinned = nullptr;                // staging, one member field
    Bufs<double> bd{};
    Bufs<float> bf{};
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    // the converged pressure solve (press_mode 1); corrections a captured solve runs, from the
    // last solve that checked its residual
    Solver* solver = nullptr;
    int graph_corrections = 0;
    // the residual a captured solve must reach: the tolerance, or four times the last checked
    // solve's residual where that one stopped at the FP64 rounding floor above the tolerance
    double graph_rel_limit = 0.0;
    std::vector<double> h_thez_sin, h_rm;   // host copies of the tables the solver is built from
    // per-kernel-family event timing (atsat_profile_*)
    bool prof = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    double prof_ms[ATSAT_NKERNELS] = {0};
    long long prof_n[ATSAT_NKERNELS] = {0};
    void prof_begin(int kid, hipEvent_t* a, hipStream_t on = nullptr){
        if(!prof) return;
        hipEventCreate(a); hipEventRecord(*a, on ? on : stream); (void)kid;
    }
    void prof_end(int kid, hipEvent_t a, hipStream_t on = nullptr){
        if(!prof) return;
        hipEvent_t b; hipEventCreate(&b); hipEventRecord(b, on ? on : stream);
        ev.push_back({kid, {a, b}});
    }
    void prof_collect(){
        if(ev.empty()) return;
        hipStreamSynchronize(stream);
        for(auto& e : ev){
            float ms = 0.f;
            hipEventElapsedTime(&ms, e.second.first, e.second.second);
            prof_ms[e.first] += ms; prof_n[e.first]++;
            hipEventDestroy(e.second.first); hipEventDestroy(e.second.second);
        }
        ev.clear();
    }

    template<typename T> Bufs<T>& bufs();
    int check_launch(const char* what){
        hipError_t e = hipGetLastError();
        if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
        return ATSAT_OK;
    }
    ~atsat_ctx(){
        prof_collect();
        if(graph_exec) hipGraphExecDestroy(graph_exec);
        if(graph) hipGraphDestroy(graph);
        if(solver) solver_destroy(solver);
        if(arena) hipFree(arena);
        if(d_small) hipFree(d_small);
        if(mca_mem) hipFree(mca_mem);
        if(pm_mem) hipFree(pm_mem);
        if(pinned) hipHostFree(pinned);
        if(ev_side_fork) hipEventDestroy(ev_side_fork);
        if(ev_side_join) hipEventDestroy(ev_side_join);
        if(ev_phys_fork) hipEventDestroy(ev_phys_fork);
        if(ev_phys_join) hipEventDestroy(ev_phys_join);
        if(side) hipStreamDestroy(side);
        if(stream) hipStreamDestroy(stream);
    }
};
// A batch of contexts -- grids of different shapes, each with its own members -- advanced
// together: each context's physics block and RK preparation run on its own stream, the four RK4
// stages of all of them as one launch each on the group's stream (k_rk_group), joined by events.
struct atsat_group {
    std::vector<atsat_ctx*> ctx;           // not owned; they outlive the group
    int precision = ATSAT_PREC_FP64;
    bool turb = false;
    hipStream_t stream = nullptr;
    void* d_grids = nullptr;               // RkGrid<T>[n + 1]; slot n times one grid alone
    int threads = 0, blocks = 0;           // the batched launch
    std::vector<RkTile> tiles;             // the contexts' tiles the launch was laid out for
    // true: the RK4 stages of all grids as one launch each; false: every context steps on its
    // own stream with its own tile, the streams running side by side (the autotuner's choice)
    bool merged = true;
    // with merged: the batched stage as a persistent launch over a work queue (launch_rk_queue)
    bool queue = false;
    unsigned* d_heads = nullptr;           // the work queue's kQueues heads
    hipEvent_t ev_fork = nullptr, ev_rk = nullptr;
    std::vector<hipEvent_t> ev_pre, ev_post;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    std::string err;
    ~atsat_group(){
        if(stream) hipStreamSynchronize(stream);
        if(graph_exec) hipGraphExecDestroy(graph_exec);
        if(graph) hipGraphDestroy(graph);
        if(d_grids) hipFree(d_grids);
        if(d_heads) hipFree(d_heads);
        if(ev_fork) hipEventDestroy(ev_fork);
        if(ev_rk) hipEventDestroy(ev_rk);
        for(hipEvent_t e : ev_pre) if(e) hipEventDestroy(e);
        for(hipEvent_t e : ev_post) if(e) hipEventDestroy(e);
        if(stream) hipStreamDestroy(stream);
    }
};

template<> Bufs<double>& atsat_ctx::bufs<double>(){ return bd; }
template<> Bufs<float>& atsat_ctx::bufs<float>(){ return bf; }

namespace {

// ---- scalars and geometry, formed on the host in the reference's operand order ----------
void make_scal(atsat_ctx& c){
    const atsat_grid& g = c.grid; const atsat_consts& C = c.consts; const atsat_knobs& K = c.knobs;
    Scal& S = c.S;
    std::memset(&S, 0, sizeof(S));
    const double pi180 = 180.0 / M_PI;                          // cSaturnModel.cpp:40-52
    S.dt = K.dt; S.dr = g.dr; S.dthe = g.dthe_deg / pi180; S.dphi = g.dphi_deg / pi180;
    S.inv_2dr = 1.0 / (2.0 * S.dr); S.inv_2dthe = 1.0 / (2.0 * S.dthe); S.inv_2dphi = 1.0 / (2.0 * S.dphi);
    S.inv_dr2 = 1.0 / (S.dr * S.dr); S.inv_dthe2 = 1.0 / (S.dthe * S.dthe); S.inv_dphi2 = 1.0 / (S.dphi * S.dphi);
    S.re = C.re; S.pr = C.pr; S.g = C.g; S.omega = C.omega; S.gam = C.gam; S.u_0 = C.u_0;
    S.t_ref = C.t_ref; S.p_ref = C.p_ref; S.R_ref = C.R_ref; S.L_atm = C.L_atm;
    S.Coriolis = C.Coriolis; S.centrifugal = C.centrifugal; S.buoyancy = C.buoyancy;
    S.r_mix = C.r_mix; S.cp_mix = C.cp_mix; S.R_mix = C.R_mix;
    S.sc_ch4 = C.sc_ch4; S.sc_h2o = C.sc_h2o; S.sc_h2s = C.sc_h2s; S.sc_nh3 = C.sc_nh3; S.sc_nh4sh = C.sc_nh4sh;
    const double L_m = C.L_atm * 1.0e3;
    S.nd_cor = L_m / C.u_0;                                     // RHS_Sat_Turb.cpp:272-273
    S.coeff_energy_p = C.u_0 * C.u_0 / (C.cp_mix * C.t_ref);    // :352
    S.to_nd = 1.0 / (C.r_mix * C.u_0 * C.u_0 * 1.0e-5);         // :374
    S.precip_den = C.r_mix * C.cp_mix * C.u_0 * C.t_ref;        // :655-656
    S.exp_pressure = C.g / (C.gam * C.R_ref);                   // InitVariables_Sat.cpp:557
    S.pdyn_cap = K.pdyn_cap;
    S.chemical_reaction = K.chemical_reaction; S.turb_coupling = K.turb_coupling;
    S.precip_coupling = K.precip_coupling;
    S.tke_max_nd = 1000.0 / (C.u_0 * C.u_0);                    // RungeKutta_Sat_Turb.cpp:175-176
    S.dis_min_nd = 1.0e-10;
    S.cor_metric = C.Coriolis * (C.L_atm * 1.0e3 / C.u_0) * 2.0 * C.omega;   // :88
    S.shift = C.metric_radius_km / C.L_atm - 1.0;               // cSaturnModel.h:168
    const float h = C.L_atm * g.dr;                             // cSaturnModel.h:1069-1072
    const float h0 = 0 * h, h1 = 1 * h;
    S.dz_m = (double)(h1 - h0) * 1.0e3;                         // layer_thickness_m, :689-692
    S.a_hydro = S.exp_pressure * C.gam * (double)(h1 - h0);     // InitVariables_Sat.cpp:568
    S.turb_active = K.turb_active ? 1 : 0;
    S.chem_inputs = (K.chem_inputs || K.chem_block) ? 1 : 0;
    S.chem_block = K.chem_block ? 1 : 0;
    // the chemistry block, planet/ChemistrySat.h:242, 376-381
    S.D_nh3 = C.mue_nh3 / (C.rg_nh3 * C.sc_nh3);
    S.D_h2s = C.mue_h2s / (C.rg_h2s * C.sc_h2s);
    S.D_nh4sh = C.mue_nh4sh / (C.rg_nh4sh * C.sc_nh4sh);
    S.m_nh3 = C.m_nh3; S.m_h2s = C.m_h2s; S.m_nh4sh = C.m_nh4sh; S.c_mix = C.c_mix;
    S.cp_nh3 = C.cp_nh3; S.cp_h2s = C.cp_h2s; S.cp_nh4sh = C.cp_nh4sh;
    S.k_nh3 = C.k_nh3; S.k_h2s = C.k_h2s; S.k_nh4sh = C.k_nh4sh;
    S.nh4sh_Lf = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;
    S.ep_h2s = C.ep_h2s;
    S.layer_h = (double)h;
    // the turbulence closure, planet/Turbulence.h: nue_max (:566), the k floor of TURB_FREE
    // (:623-628), the background k / omega of init_fields (:431-476) with TURB_FREE's wall distance
    S.turb_buoy = K.turb_buoy;
    S.turb_nue_max_nd = K.turb_nue_max / (C.u_0 * L_m);
    S.turb_k_bg_nd = K.turb_k_bg / (C.u_0 * C.u_0);
    {
        const double C_nue = 0.028, Karman = 0.42, dis_min = 1.0e-10;
        const double inv_u02 = 1.0 / (C.u_0 * C.u_0);
        const double nd_omega = L_m / C.u_0;
        const double L_bg = L_m / double(g.im - 1);
        const double om_bg = std::pow(C_nue, -0.25) * std::sqrt(K.turb_k_bg) / (Karman * L_bg);
        const double k_phys = std::max(0.0, K.turb_k_bg);
        const double om_phys = std::pow(C_nue, -0.25) * std::sqrt(k_phys) / (Karman * 1.0e20);
        const double dis_nd = std::max(om_phys, om_bg) * nd_omega;
        S.turb_init_tke = std::max(0.0, k_phys * inv_u02);
        S.turb_init_dis = std::max(dis_min, dis_nd);
    }
    for(int q = 0; q < 3; q++){
        S.t_0[q] = C.t_0[q]; S.t_00[q] = C.t_00[q]; S.ep[q] = C.ep[q]; S.lv[q] = C.lv[q]; S.ls[q] = C.ls[q];
        S.C[q] = C.C[q]; S.L0[q] = C.L0[q]; S.R[q] = C.R[q]; S.del_alf[q] = C.del_alf[q]; S.del_bet[q] = C.del_bet[q];
        S.C_i[q] = C.C_i[q]; S.L0_i[q] = C.L0_i[q]; S.del_alf_i[q] = C.del_alf_i[q]; S.del_bet_i[q] = C.del_bet_i[q];
    }
    S.t_0_nh4sh = C.t_0_nh4sh; S.t_00_nh4sh = C.t_00_nh4sh;
    const double r_p = 1.0e-5, rho_c = 1200.0, mu = 9.0e-6;    // Precipitation.h:551-554
    S.v_settle = (2.0 / 9.0) * r_p * r_p * (rho_c - C.r_mix) * C.g / mu;
    // the weight sum of computeBuoyancyRefLevel (RungeKutta_Sat_Turb.cpp:36-47) for a level none
    // of whose cells is skipped, in its serial order
    double thez = 0.0, wsum = 0.0;
    for(int j = 0; j < g.jm; j++){
        if(j > 0) thez = thez + S.dthe;                          // lib/Array_1D.cpp:54-59
        const double wgt = std::sin(thez);
        for(int k = 0; k < g.km; k++) wsum += wgt;
    }
    S.buoy_wsum_full = wsum;
}

bool uniform_layers(const atsat_ctx& c){
    const float h = c.consts.L_atm * c.grid.dr;
    const double d0 = (double)(1 * h - 0 * h);
    for(int i = 0; i + 1 < c.grid.im; i++) if((double)((i + 1) * h - i * h) != d0) return false;
    return true;
}

// Per-(i,j) table: CellGeometry (RungeKutta_Sat_Turb.cpp:129-143, 223-243) and the solver's
// weights (PressureSolver.h:156-162, 255-259, 326-353). sin / cos come from the host libm, as
// in the reference.
void make_geo(const atsat_ctx& c, std::vector<double>& geo, std::vector<double>& sinj,
              std::vector<double>& sabs, std::vector<double>& rmv){
    const int im = c.grid.im, jm = c.grid.jm;
    const Scal& S = c.S;
    std::vector<double> radz(im), thez(jm);
    radz[0] = 1.0; for(int l = 1; l < im; l++) radz[l] = radz[l - 1] + S.dr;     // lib/Array_1D.cpp:54-59
    thez[0] = 0.0; for(int l = 1; l < jm; l++) thez[l] = thez[l - 1] + S.dthe;
    geo.assign((size_t)im * jm * NGEO, 0.0);
    sinj.resize(jm); sabs.resize(jm); rmv.resize(im);
    for(int j = 0; j < jm; j++){ sinj[j] = std::sin(thez[j]); sabs[j] = std::abs(std::sin(thez[j])); }
    for(int i = 0; i < im; i++){
        const double R_km = c.consts.metric_radius_km;
        const double rm = (R_km > 0.0) ? radz[i] + (R_km / c.consts.L_atm - 1.0) : radz[i];
        rmv[i] = rm;
        for(int j = 0; j < jm; j++){
            double* g = &geo[((size_t)i * jm + j) * NGEO];
            double s = std::sin(thez[j]);
            if(s < 0.0) s = 0.0;
            const double rm2 = rm * rm, sinthe2 = s * s, costhe = std::cos(thez[j]);
            const double inv_rm2 = 1.0 / rm2;
            g[G_RM] = rm; g[G_COSTHE] = costhe; g[G_COTANTHE] = costhe / s; g[G_SINTHE2] = sinthe2;
            g[G_INV_RM] = 1.0 / rm; g[G_INV_RM2] = inv_rm2; g[G_INV_RMSINTHE] = 1.0 / (rm * s);
            g[G_INV_RM2SINTHE] = inv_rm2 / s; g[G_INV_RM2SINTHE2] = inv_rm2 / sinthe2;
            g[G_COSTHE_INV_RM2SINTHE] = costhe * g[G_INV_RM2SINTHE];
            const double st = std::sqrt(std::max(0.0, 1.0 - costhe * costhe));       // RHS_Sat_Turb.cpp:310
            g[G_SINTHE_TRUE] = st;
            g[G_CENT_RAD] = 1.0 * c.consts.omega * c.consts.omega * rm * st * st;     // :349
            g[G_CENT_THE] = 1.0 * c.consts.omega * c.consts.omega * rm * st * costhe; // :350
            g[G_V_METRIC] = (1.0 + costhe / sinthe2) * inv_rm2;                       // :436
            double sp = std::sin(thez[j]);                                             // PressureSolver.h:158-161
            if(sp < 0.0) sp = 0.0;
            if(sp == 0.0) sp = 1.0e-5;
            const double p_inv_rm2sinthe2 = inv_rm2 / (sp * sp);
            g[G_P_NUM2] = inv_rm2 * S.inv_dthe2;                                       // :351
            g[G_P_NUM3] = p_inv_rm2sinthe2 * S.inv_dphi2;                              // :352
            const double num1 = 1.0 * S.inv_dr2;
            g[G_P_DENOM] = 2.0 * num1 + 2.0 * g[G_P_NUM2] + 2.0 * g[G_P_NUM3];        // :353
            g[G_P_INV_RM] = 1.0 / rm;
            g[G_P_INV_RMSINTHE] = 1.0 / (rm * sp);
            g[G_P_LC_NUM2] = S.inv_dthe2 / (rm * rm);                                  // :258
            g[G_P_LC_NUM3] = S.inv_dphi2 / (rm * rm * sp * sp);                        // :259
        }
    }
}

// The pressure terms are kept by the FP32 step only (the FP64 stages form them in place).
size_t buf_bytes(int b, size_t per, size_t elem){
    if(b >= B_PTERM && b < B_PTERM + 6 && elem == 8) return 0;
    return per * (is_state(b) ? 8 : elem);
}

// Scratch of the level mean (k_buoy_terms): one FP64 term per (position, member x level).
long long buoy_pitch(const Dims& D){ return ((long long)D.M * D.im + 63) / 64 * 64; }
size_t buoy_bytes(const Dims& D){ return (size_t)D.jm * D.km * buoy_pitch(D) * sizeof(double); }

template<typename T>
void assign_bufs(atsat_ctx& c){
    Bufs<T>& B = c.bufs<T>();
    std::memset(&B, 0, sizeof(B));
    char* base = static_cast<char*>(c.arena);
    const size_t per = (size_t)c.D.fs * c.D.M;
    size_t off = 0;
    for(int b = 0; b < NBUF; b++){
        if(is_state(b)) B.s[b] = reinterpret_cast<double*>(base + off);
        else B.w[b] = reinterpret_cast<T*>(base + off);
        off += buf_bytes(b, per, sizeof(T));
    }
    B.buoy_ch = buoy_pitch(c.D);
    B.buoy_x = reinterpret_cast<double*>(base + off);
    double* s = c.d_small;
    const int im = c.grid.im, jm = c.grid.jm, km = c.grid.km, M = c.D.M;
    B.geo = s; s += (size_t)im * jm * NGEO;
    B.thez_sin = s; s += jm;
    B.sin_abs = s; s += jm;
    B.rm = s; s += im;
    B.vbar0 = s; s += (size_t)M * im * jm;
    B.ubar0 = s; s += (size_t)M * im * jm;
    B.buoy_ref = s; s += (size_t)M * im;
    B.srf = s; s += (size_t)M * 5 * jm * km;
    B.vel_star = s; s += (size_t)M * jm * km;
    B.counter = reinterpret_cast<unsigned long long*>(s); s += 1;
    B.budget = s; s += (size_t)M * (5 * im + 3);
}

size_t small_elems(const atsat_grid& g, int M){
    return (size_t)g.im * g.jm * NGEO + 2 * (size_t)g.jm + g.im + 2 * (size_t)M * g.im * g.jm
         + (size_t)M * g.im + (size_t)M * 6 * g.jm * g.km + 1 + (size_t)M * (5 * g.im + 3);
}

#define PROF(c, kid, call) do { hipEvent_t a_ = nullptr; (c).prof_begin(kid, &a_); call; (c).prof_end(kid, a_); } while(0)
#define PROF_ON(c, kid, st, call) do { hipEvent_t a_ = nullptr; (c).prof_begin(kid, &a_, st); call; (c).prof_end(kid, a_, st); } while(0)

// ---- the step, templated on the storage type ---------------------------------------------
// iter: the model loop's iteration (its odd iterations keep the static pressure terms of the even
// one before), or -1 outside the loop.
template<typename T>
void do_rk_pre(atsat_ctx& c, int iter = -1){
    Bufs<T>& B = c.bufs<T>();
    const bool keep = iter >= 0 && iter % 2 == 1 && c.pterm_iter == iter - 1;
    c.pterm_iter = (iter >= 0 && iter % 2 == 0) ? iter : -1;
    // the inputs of the stages: the metric pressure, then the pressure terms (they take the
    // metric pressure's neighbours), and the halo rows on the side stream while the level mean
    // runs on the main stream
    hipEventRecord(c.ev_side_fork, c.stream);
    hipStreamWaitEvent(c.side, c.ev_side_fork, 0);
    PROF_ON(c, ATSAT_K_RK_AUX, c.side, launch_metric_pressure<T>(c.D, c.S, B, c.side));
    if(sizeof(T) == 4) PROF_ON(c, ATSAT_K_RK_AUX, c.side, launch_press_terms<T>(c.D, c.S, B, !keep, c.side));
    PROF_ON(c, ATSAT_K_RK_AUX, c.side, launch_rk_halo<T>(c.D, c.S, B, c.side));
    hipEventRecord(c.ev_side_join, c.side);
    PROF(c, ATSAT_K_RK_AUX, launch_buoy_ref<T>(c.D, c.S, B, c.stream));
    hipStreamWaitEvent(c.stream, c.ev_side_join, 0);
}

template<typename T>
int do_rk4(atsat_ctx& c, bool alias, bool clamp = false, int iter = -1){
    Bufs<T>& B = c.bufs<T>();
    do_rk_pre<T>(c, iter);
    for(int s = 0; s < 4; s++) PROF(c, ATSAT_K_RK_STAGE, launch_rk_stage<T>(c.D, c.S, B, c.tile, s, alias, clamp, c.stream));
    return c.check_launch("rk4_step");
}

// Make the *n buffers hold the start-of-step copies explicitly (they are aliased to the live
// state after restoreVar).
// With turb_active, tken / disn are never aliased: restoreVar copies tke / dis into them, and the
// closure, which runs after it, changes the live fields only.
const unsigned kTurbMask = (1u << ATSAT_TKE) | (1u << ATSAT_DIS);
bool n_split(const atsat_ctx& c, int f){ return c.S.turb_active && (f == ATSAT_TKE || f == ATSAT_DIS); }

template<typename T>
void materialize_n(atsat_ctx& c){
    if(!c.n_alias) return;
    const unsigned mask = ((1u << ATSAT_NPROG) - 1u) & ~(c.S.turb_active ? kTurbMask : 0u);
    PROF(c, ATSAT_K_GLUE, launch_copy_n<T>(c.D, c.S, c.bufs<T>(), mask, c.stream));
    c.n_alias = false;
}

template<typename T>
int do_rk4_api(atsat_ctx& c){
    // RungeKuttaSat alone leaves the *n copies as they were, so they must exist apart from L.
    materialize_n<T>(c);
    return do_rk4<T>(c, false);
}

// Whether the converged solve's corrections take their forward transforms in FP32.
bool fp32_corrections(const atsat_ctx& c){
    return c.precision == ATSAT_PREC_FP32 && c.knobs.press_fp32_transforms != 0;
}

// The converged solve of Lc p_dyn = PSRC: FP64 defect correction, the corrections' forward
// transforms in FP32 or FP64 (fp32_corrections). Outside graph capture the residual is checked on the host after every correction;
// inside, the captured solve runs the number of corrections the last checked solve needed (+ 1
// where that one ended within a factor 4 of the tolerance).
template<typename T>
int run_solver(atsat_ctx& c, double tol, int max_iter, int* iters, double* rel, hipStream_t st){
    Bufs<T>& B = c.bufs<T>();
    std::string& err = c.err;
    if(!c.solver) return fail(&c.err, ATSAT_E_UNSUPPORTED, "converged pressure solve not set up (press_mode 0 context)");
    solver_set_stream(c.solver, st);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPC(hipStreamIsCapturing(st, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    const int n = capturing ? (c.graph_corrections > 0 ? c.graph_corrections : max_iter) : max_iter;
    // before any checked solve the limit allows the FP64 floor of the worst source measured
    // (about 2.5e-9 on the native grid)
    const double limit = c.graph_rel_limit > 0.0 ? c.graph_rel_limit : std::max(tol, 1e-8);
    int rc = solver_run(c.solver, B.s[ATSAT_P_DYN], B.s[B_PSRC], B.s[B_TMP], B.geo, fp32_corrections(c), tol, n,
                        !capturing, limit);
    int it = 0; double r = -1.0;
    solver_last(c.solver, &it, &r);
    if(!capturing && rc == ATSAT_OK){
        // a captured solve runs a fixed count: the checked one's where it stopped at the rounding
        // floor (its last correction already no longer halved the residual) or at a quarter of the
        // tolerance or below (a later state whose residual is up to four times larger still
        // meets it); one more where it stopped just under the tolerance. A replay that misses
        // the tolerance anyway is reported by the status word (k_sol_status)
        const bool margin = r >= 0.0 && r <= 0.25 * tol;
        c.graph_corrections = solver_last_floor(c.solver) ? (it > 1 ? it : 1) : (margin && it > 0 ? it : it + 1);
        c.graph_rel_limit = std::max(tol, 4.0 * r);
    }
    if(iters) *iters = it;
    if(rel) *rel = r;
    if(rc != ATSAT_OK) return fail(&c.err, rc, solver_error(c.solver));
    return c.check_launch("converged pressure solve");
}

// The pressure solve on stream st (the context's own stream, or its side stream beside the column
// physics of the even iteration).
template<typename T>
int do_pressure(atsat_ctx& c, hipStream_t st){
    Bufs<T>& B = c.bufs<T>();
    PROF_ON(c, ATSAT_K_PRESS, st, launch_press_prepare<T>(c.D, c.S, B, st));
    PROF_ON(c, ATSAT_K_PRESS, st, launch_press_src<T>(c.D, c.S, B, st));
    if(c.knobs.press_mode == 1){
        int rc = ATSAT_OK;
        PROF_ON(c, ATSAT_K_PRESS, st, rc = run_solver<T>(c, c.knobs.press_tol, c.knobs.press_max_iter, nullptr, nullptr, st));
        return rc;
    }
    if(c.knobs.press_mode != 0) return fail(&c.err, ATSAT_E_UNSUPPORTED, "press_mode must be 0 or 1");
    const int n = c.knobs.press_sweeps > 0 ? c.knobs.press_sweeps : 1;
    const bool line = c.knobs.press_line != 0, baro = c.knobs.press_baro != 0;
    if((line || baro) && !B.pmodes)
        return fail(&c.err, ATSAT_E_UNSUPPORTED, "the solver modes' scratch is not allocated (press_line / press_baro set after a capture began)");
    if(line) PROF_ON(c, ATSAT_K_PRESS, st, launch_press_mean<T>(c.D, B, 0, st));       // PressureSolver.h:315
    for(int s = 0; s < n; s++){
        PROF_ON(c, ATSAT_K_PRESS, st, launch_press_sweep<T>(c.D, c.S, B, 0, st));
        PROF_ON(c, ATSAT_K_PRESS, st, launch_press_sweep<T>(c.D, c.S, B, 1, st));
        if(line){                                                                       // :502-570
            PROF_ON(c, ATSAT_K_PRESS, st, launch_press_line<T>(c.D, c.S, B, 0, st));
            PROF_ON(c, ATSAT_K_PRESS, st, launch_press_line<T>(c.D, c.S, B, 1, st));
        }
        if(baro) PROF_ON(c, ATSAT_K_PRESS, st, launch_press_baro<T>(c.D, c.S, B, c.pm_A2, st));   // :573-708
        PROF_ON(c, ATSAT_K_PRESS, st, launch_press_boundary<T>(c.D, c.S, B, st));
        if(line) PROF_ON(c, ATSAT_K_PRESS, st, launch_press_mean<T>(c.D, B, 1, st));   // :750-755
    }
    return c.check_launch("pressure_solve");
}

const unsigned kSyncMask = (1u << ATSAT_H2O) | (1u << ATSAT_H2O_CLOUD) | (1u << ATSAT_H2O_ICE)
    | (1u << ATSAT_NH3) | (1u << ATSAT_NH3_CLOUD) | (1u << ATSAT_NH3_ICE) | (1u << ATSAT_CH4)
    | (1u << ATSAT_CH4_CLOUD) | (1u << ATSAT_CH4_ICE) | (1u << ATSAT_T) | (1u << ATSAT_H2S) | (1u << ATSAT_NH4SH);

// One iteration of Run's loop. The RK scalar sync and restoreVar copy live -> *n; with the
// copies aliased to the live state both are satisfied without moving data, so after the first
// iteration the step runs aliased: y_n is read from L.
template<typename T>
int do_physics(atsat_ctx& c){
    Bufs<T>& B = c.bufs<T>();
    std::string& err = c.err;
    int rc;
    PROF(c, ATSAT_K_DENSITY, launch_mixture_density<T>(c.D, c.S, B, c.stream));
    // The pressure solve and the p_dyn filter read and write only p_dyn, rhs / aux and their own
    // scratch, none of which the saturation adjustment, the p_stat rebuild, the NH4SH
    // equilibrium, the precipitation or the chemistry read or write (and these touch nothing the
    // solve reads): the solve runs on the side stream beside them and is joined before the RK4
    // step. The column physics is enqueued first, so that a solve that checks its residual on
    // the host (outside a graph) does not hold it back.
    HIPC(hipEventRecord(c.ev_phys_fork, c.stream));
    HIPC(hipStreamWaitEvent(c.side, c.ev_phys_fork, 0));
    PROF(c, ATSAT_K_SATADJ, launch_satadj<T>(c.D, c.S, B, ATSAT_GAS_CH4, ATSAT_GAS_NH3, c.stream));
    PROF(c, ATSAT_K_PSTAT, launch_rebuild_pstat<T>(c.D, c.S, B, c.stream));
    if(c.S.chem_block) PROF(c, ATSAT_K_CHEM, launch_nh4sh_equilibrium<T>(c.D, c.S, B, c.stream));
    if(c.knobs.precip_on) PROF(c, ATSAT_K_PRECIP, launch_precipitation<T>(c.D, c.S, B, c.stream));
    if(!c.n_alias) PROF(c, ATSAT_K_GLUE, launch_copy_n<T>(c.D, c.S, B, kSyncMask, c.stream));
    if(c.S.chem_block) PROF(c, ATSAT_K_CHEM, launch_chemistry<T>(c.D, c.S, B, c.stream));
    rc = do_pressure<T>(c, c.side);
    if(rc == ATSAT_OK) PROF_ON(c, ATSAT_K_DAMP, c.side, launch_damp_pdyn<T>(c.D, c.S, B, c.side));
    HIPC(hipEventRecord(c.ev_phys_join, c.side));
    HIPC(hipStreamWaitEvent(c.stream, c.ev_phys_join, 0));
    return rc;
}

// The moist convective adjustment and its sync (cSaturnModel.cpp:1120-1171); with the *n copies
// aliased to the live state the sync moves nothing.
template<typename T>
int do_mca(atsat_ctx& c){
    if(!c.mca_mem) return fail(&c.err, ATSAT_E_UNSUPPORTED, "the convective adjustment's scratch is not allocated (mca_stride 0 at set-up)");
    PROF(c, ATSAT_K_MCA, launch_mca<T>(c.D, c.S, c.bufs<T>(), c.stream));
    if(!c.n_alias) PROF(c, ATSAT_K_GLUE, launch_copy_n<T>(c.D, c.S, c.bufs<T>(), kSyncMask, c.stream));
    return c.check_launch("convective adjustment");
}

// restoreVar(1.0): aliases the *n copies to the live state; tke / dis, with turb_active, are
// copied (their *n copies must outlive the closure's changes).
template<typename T>
void restore_var(atsat_ctx& c){
    if(c.S.turb_active) PROF(c, ATSAT_K_GLUE, launch_copy_n<T>(c.D, c.S, c.bufs<T>(), kTurbMask, c.stream));
    c.n_alias = true;
}

// After the RK4 stages: the boundary conditions (with the species clamp's boundary half),
// restoreVar(1.0), with turb_active the turbulence closure, and every mca_stride iterations the
// moist convective adjustment (cSaturnModel.cpp:1091-1171).
template<typename T>
int do_step_end(atsat_ctx& c, int iter){
    PROF(c, ATSAT_K_GLUE, launch_bc<T>(c.D, c.S, c.bufs<T>(), true, c.stream));
    restore_var<T>(c);
    if(c.S.turb_active && c.knobs.turb_closure)
        PROF(c, ATSAT_K_TURB, launch_turbulence<T>(c.D, c.S, c.bufs<T>(), false, c.stream));
    if(c.knobs.mca_stride > 0 && iter % c.knobs.mca_stride == 0){
        int rc = do_mca<T>(c);
        if(rc != ATSAT_OK) return rc;
    }
    return c.check_launch("step");
}

template<typename T>
int do_step(atsat_ctx& c, int iter){
    int rc;
    if(iter % 2 == 0 && (rc = do_physics<T>(c)) != ATSAT_OK) return rc;
    // clampNegativeSpecies is folded into the last RK stage and the radial boundary copy
    if((rc = do_rk4<T>(c, c.n_alias, true, iter)) != ATSAT_OK) return rc;
    return do_step_end<T>(c, iter);
}

// ---- launch tiles of the RK4 stage kernel --------------------------------------------------
// Candidate rows per band; a tile's band must be a multiple of its rings per workgroup.
const int kBandCands[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};

// The tuners time whole rounds of the four RK4 stages. Stage 3 writes the prognostic state and
// the last-stage tendencies rhs / aux: stage_save copies those buffers aside (a temporary device
// allocation) and stage_restore puts them back, so a tuner leaves the context as it found it.
// The other buffers the stages write (the stage inputs X, Y and the accumulators) are scratch
// that every step's stage 0 rewrites.
struct StageSave {
    void* buf = nullptr;
    std::vector<std::pair<void*, size_t>> parts;   // (device buffer, bytes)
};

template<typename T>
int stage_save(atsat_ctx& c, StageSave& sv){
    std::string& err = c.err;
    Bufs<T>& B = c.bufs<T>();
    const size_t n = (size_t)c.D.M * (size_t)c.D.fs;
    for(int f = 0; f < ATSAT_NPROG; f++)
        sv.parts.push_back(is_state(f) ? std::pair<void*, size_t>{B.s[f], n * sizeof(double)}
                                       : std::pair<void*, size_t>{B.w[f], n * sizeof(T)});
    for(int f = ATSAT_RHS_U; f <= ATSAT_AUX_W; f++) sv.parts.push_back({B.w[f], n * sizeof(T)});
    size_t total = 0;
    for(auto& p : sv.parts) total += p.second;
    if(hipMalloc(&sv.buf, total) != hipSuccess){
        sv.buf = nullptr;
        return fail(&err, ATSAT_E_NOMEM, "autotune: no device memory to save the state");
    }
    char* o = static_cast<char*>(sv.buf);
    for(auto& p : sv.parts){
        HIPC(hipMemcpyAsync(o, p.first, p.second, hipMemcpyDeviceToDevice, c.stream));
        o += p.second;
    }
    return ATSAT_OK;
}

int stage_restore(atsat_ctx& c, StageSave& sv){
    std::string& err = c.err;
    if(!sv.buf) return ATSAT_OK;
    const char* o = static_cast<const char*>(sv.buf);
    for(auto& p : sv.parts){
        HIPC(hipMemcpyAsync(p.first, o, p.second, hipMemcpyDeviceToDevice, c.stream));
        o += p.second;
    }
    HIPC(hipStreamSynchronize(c.stream));
    HIPC(hipFree(sv.buf));
    sv.buf = nullptr;
    return ATSAT_OK;
}

// Mean time per stage of whole rounds of the four stages with tile t, over reps rounds after
// one untimed round, on c.stream. The state must have been saved.
template<typename T>
int time_stages(atsat_ctx& c, const RkTile& t, int reps, double* ms){
    std::string& err = c.err;
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    HIPC(hipEventCreate(&b));
    for(int s = 0; s < 4; s++) launch_rk_stage<T>(c.D, c.S, c.bufs<T>(), t, s, c.n_alias, true, c.stream);
    hipEventRecord(a, c.stream);
    for(int r = 0; r < reps; r++)
        for(int s = 0; s < 4; s++) launch_rk_stage<T>(c.D, c.S, c.bufs<T>(), t, s, c.n_alias, true, c.stream);
    hipEventRecord(b, c.stream);
    hipError_t e = hipEventSynchronize(b);
    float f = 0.f;
    if(e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string("stage timing: ") + hipGetErrorString(e));
    *ms = f / (4.0 * reps);
    return c.check_launch("stage timing");
}

// Every valid (band, rings) tile of this grid, each timed on rounds of the four stages; the
// fastest is kept.
template<typename T>
int do_autotune(atsat_ctx& c, double* best_ms){
    StageSave sv;
    int rc = stage_save<T>(c, sv);
    if(rc == ATSAT_OK){
        do_rk_pre<T>(c);
        RkTile best = c.tile;
        double bms = 1e30;
        const int nj = c.D.jm - 4;
        for(int rings = 1; rc == ATSAT_OK && rings * rk_ring_threads(c.D) <= kRowThreads; rings++){
            for(int band : kBandCands){
                const RkTile t{band, rings};
                if(!rk_tile_ok(c.D, t) || band > nj + rings) continue;
                double ms = 0.0;
                if((rc = time_stages<T>(c, t, 2, &ms)) != ATSAT_OK) break;
                if(ms < bms){ bms = ms; best = t; }
            }
        }
        if(rc == ATSAT_OK){
            c.tile = best;
            if(best_ms) *best_ms = bms;
        }
    }
    const int rr = stage_restore(c, sv);
    return rc != ATSAT_OK ? rc : rr;
}

// ---- batches of grids ------------------------------------------------------------------------
// The device table of the batched launch from the contexts' tiles (tiles[i], or the context's own
// tile), all workgroups `threads` wide. Slot n of the table holds grid `alone` by itself when
// alone >= 0. Synchronous: not for use while capturing.
template<typename T>
int group_layout(atsat_group& g, int threads, const std::vector<RkTile>& tiles, int alone){
    std::string& err = g.err;
    const int n = (int)g.ctx.size();
    std::vector<RkGrid<T>> h(n + 1);
    int first = 0;
    for(int i = 0; i <= n; i++){
        const int src = i < n ? i : alone;
        if(src < 0) break;
        atsat_ctx& c = *g.ctx[src];
        RkGrid<T>& G = h[i];
        G.D = c.D; G.S = c.S; G.B = c.bufs<T>();
        G.RM = rk_row_map(c.D, tiles[src]);
        G.rt = rk_ring_threads(c.D);
        G.first_block = i < n ? first : 0;
        if(i < n) first += G.RM.blocks();               // a multiple of 8
    }
    HIPC(hipMemcpy(g.d_grids, h.data(), sizeof(RkGrid<T>) * (n + 1), hipMemcpyHostToDevice));
    g.threads = threads;
    g.blocks = first;
    return ATSAT_OK;
}

// Mean time per stage of rounds of the four batched stages (all grids, or slot n alone) over
// reps rounds. The contexts' states must have been saved.
template<typename T>
int group_time_stages(atsat_group& g, bool alone, int blocks, int reps, double* ms, bool queue = false){
    std::string& err = g.err;
    const int n = (int)g.ctx.size();
    const RkGrid<T>* G = static_cast<const RkGrid<T>*>(g.d_grids) + (alone ? n : 0);
    unsigned alias = 0;
    for(int i = 0; i < n; i++) if(g.ctx[i]->n_alias) alias |= 1u << i;
    if(alone) alias = ~0u;                              // timing only: y_n from the live fields
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    HIPC(hipEventCreate(&b));
    auto stage = [&](int s){
        if(queue) launch_rk_queue<T>(G, alone ? 1 : n, blocks, g.threads, g.turb, s, alias, true, g.d_heads, g.stream);
        else launch_rk_group<T>(G, alone ? 1 : n, blocks, g.threads, g.turb, s, alias, true, g.stream);
    };
    for(int s = 0; s < 4; s++) stage(s);
    hipEventRecord(a, g.stream);
    for(int r = 0; r < reps; r++)
        for(int s = 0; s < 4; s++) stage(s);
    hipEventRecord(b, g.stream);
    hipError_t e = hipEventSynchronize(b);
    float f = 0.f;
    if(e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if(e == hipSuccess) e = hipGetLastError();
    if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string("batched stage timing: ") + hipGetErrorString(e));
    *ms = f / (4.0 * reps);
    return ATSAT_OK;
}

// The autotuner of a batch. For each workgroup width (a multiple of 64 that holds the longest
// ring, up to the kernel's limit), every grid takes as many rings per workgroup as fit and
// times each band for itself; the width whose batched stages are fastest wins.
// Mean time per stage of rounds of the four stages with every context on its own stream, with
// tile tiles[i], the streams side by side (from the group stream's fork to its join).
template<typename T>
int group_time_side(atsat_group& g, const std::vector<RkTile>& tiles, int reps, double* ms){
    std::string& err = g.err;
    const int n = (int)g.ctx.size();
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    HIPC(hipEventCreate(&b));
    for(int r = 0; r <= reps; r++){
        if(r == 1) HIPC(hipEventRecord(a, g.stream));
        HIPC(hipEventRecord(g.ev_fork, g.stream));
        for(int i = 0; i < n; i++){
            atsat_ctx& c = *g.ctx[i];
            HIPC(hipStreamWaitEvent(c.stream, g.ev_fork, 0));
            for(int s = 0; s < 4; s++) launch_rk_stage<T>(c.D, c.S, c.bufs<T>(), tiles[i], s, c.n_alias, true, c.stream);
            HIPC(hipEventRecord(g.ev_post[i], c.stream));
            HIPC(hipStreamWaitEvent(g.stream, g.ev_post[i], 0));
        }
    }
    HIPC(hipEventRecord(b, g.stream));
    hipError_t e = hipEventSynchronize(b);
    float f = 0.f;
    if(e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if(e == hipSuccess) e = hipGetLastError();
    if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string("side-by-side stage timing: ") + hipGetErrorString(e));
    *ms = f / (4.0 * reps);
    return ATSAT_OK;
}

template<typename T>
int group_search(atsat_group& g, double* best_ms){
    const int n = (int)g.ctx.size();
    int rt_max = 0;
    for(atsat_ctx* c : g.ctx) rt_max = std::max(rt_max, rk_ring_threads(c->D));
    std::vector<RkTile> best_tiles(n);
    int best_threads = 0;
    double bms = 1e30;
    for(int threads = rt_max; threads <= kRowThreads; threads += 64){
        std::vector<RkTile> tiles(n);
        for(int i = 0; i < n; i++) tiles[i] = g.ctx[i]->tile;
        for(int i = 0; i < n; i++){
            const Dims& D = g.ctx[i]->D;
            const int rings = threads / rk_ring_threads(D);
            double gbest = 1e30;
            RkTile tb{rings, rings};
            for(int band : kBandCands){
                const RkTile t{band, rings};
                if(!rk_tile_ok(D, t) || band > D.jm - 4 + rings) continue;
                tiles[i] = t;
                int rc = group_layout<T>(g, threads, tiles, i);
                if(rc) return rc;
                double ms = 0.0;
                if((rc = group_time_stages<T>(g, true, rk_row_map(D, t).blocks(), 2, &ms))) return rc;
                if(ms < gbest){ gbest = ms; tb = t; }
            }
            tiles[i] = tb;
        }
        int rc = group_layout<T>(g, threads, tiles, -1);
        if(rc) return rc;
        double ms = 0.0;
        if((rc = group_time_stages<T>(g, false, g.blocks, 2, &ms))) return rc;
        if(ms < bms){ bms = ms; best_threads = threads; best_tiles = tiles; }
    }
    // the alternative: every context with the best tile of its own (any workgroup width), the
    // contexts' streams side by side
    std::vector<RkTile> own(n);
    for(int i = 0; i < n; i++){
        atsat_ctx& c = *g.ctx[i];
        double cb = 1e30;
        own[i] = best_tiles[i];
        for(int rings = 1; rings * rk_ring_threads(c.D) <= kRowThreads; rings++)
            for(int band : kBandCands){
                const RkTile t{band, rings};
                if(!rk_tile_ok(c.D, t) || band > c.D.jm - 4 + rings) continue;
                double ms = 0.0;
                int rc = time_stages<T>(c, t, 2, &ms);
                if(rc) return fail(&g.err, rc, c.err);
                if(ms < cb){ cb = ms; own[i] = t; }
            }
    }
    double side = 0.0;
    {
        int rc = group_time_side<T>(g, own, 2, &side);
        if(rc) return rc;
    }
    // the same batched layout as a persistent launch over the work queue
    double qms = 1e30;
    {
        int rc = group_layout<T>(g, best_threads, best_tiles, -1);
        if(rc) return rc;
        if((rc = group_time_stages<T>(g, false, g.blocks, 2, &qms, true))) return rc;
    }
    g.merged = std::min(bms, qms) <= side;
    g.queue = g.merged && qms < bms;
    const std::vector<RkTile>& use = g.merged ? best_tiles : own;
    for(int i = 0; i < n; i++) g.ctx[i]->tile = use[i];
    int rc = group_layout<T>(g, best_threads, best_tiles, -1);
    if(rc) return rc;
    g.tiles.clear();                       // group_refresh relays out for the contexts' tiles
    if(g.graph_exec){ hipGraphExecDestroy(g.graph_exec); g.graph_exec = nullptr; }
    if(g.graph){ hipGraphDestroy(g.graph); g.graph = nullptr; }
    if(best_ms) *best_ms = g.merged ? std::min(bms, qms) : side;
    return ATSAT_OK;
}

template<typename T>
int group_autotune(atsat_group& g, double* best_ms){
    std::string& err = g.err;
    const int n = (int)g.ctx.size();
    std::vector<StageSave> sv(n);
    int rc = ATSAT_OK;
    for(int i = 0; i < n && rc == ATSAT_OK; i++){
        if((rc = stage_save<T>(*g.ctx[i], sv[i])) != ATSAT_OK){ err = g.ctx[i]->err; break; }
        do_rk_pre<T>(*g.ctx[i]);
    }
    for(int i = 0; i < n && rc == ATSAT_OK; i++)
        if(hipStreamSynchronize(g.ctx[i]->stream) != hipSuccess) rc = fail(&err, ATSAT_E_HIP, "autotune: stream");
    if(rc == ATSAT_OK) rc = group_search<T>(g, best_ms);
    for(int i = 0; i < n; i++){
        const int rr = stage_restore(*g.ctx[i], sv[i]);
        if(rc == ATSAT_OK && rr != ATSAT_OK){ rc = rr; err = g.ctx[i]->err; }
    }
    return rc;
}

// One iteration of every context of the batch (the order of do_step per context).
template<typename T>
int group_step(atsat_group& g, int iter){
    std::string& err = g.err;
    const int n = (int)g.ctx.size();
    HIPC(hipEventRecord(g.ev_fork, g.stream));
    if(!g.merged){
        for(int i = 0; i < n; i++){
            atsat_ctx& c = *g.ctx[i];
            HIPC(hipStreamWaitEvent(c.stream, g.ev_fork, 0));
            int rc = do_step<T>(c, iter);
            if(rc != ATSAT_OK) return fail(&err, rc, c.err);
            HIPC(hipEventRecord(g.ev_post[i], c.stream));
            HIPC(hipStreamWaitEvent(g.stream, g.ev_post[i], 0));
        }
        return ATSAT_OK;
    }
    unsigned alias = 0;
    for(int i = 0; i < n; i++){
        atsat_ctx& c = *g.ctx[i];
        HIPC(hipStreamWaitEvent(c.stream, g.ev_fork, 0));
        int rc;
        if(iter % 2 == 0 && (rc = do_physics<T>(c)) != ATSAT_OK) return fail(&err, rc, c.err);
        do_rk_pre<T>(c, iter);
        if(c.n_alias) alias |= 1u << i;
        HIPC(hipEventRecord(g.ev_pre[i], c.stream));
        HIPC(hipStreamWaitEvent(g.stream, g.ev_pre[i], 0));
    }
    const RkGrid<T>* G = static_cast<const RkGrid<T>*>(g.d_grids);
    for(int s = 0; s < 4; s++){
        if(g.queue) launch_rk_queue<T>(G, n, g.blocks, g.threads, g.turb, s, alias, true, g.d_heads, g.stream);
        else launch_rk_group<T>(G, n, g.blocks, g.threads, g.turb, s, alias, true, g.stream);
    }
    HIPC(hipGetLastError());
    HIPC(hipEventRecord(g.ev_rk, g.stream));
    for(int i = 0; i < n; i++){
        atsat_ctx& c = *g.ctx[i];
        HIPC(hipStreamWaitEvent(c.stream, g.ev_rk, 0));
        int rc = do_step_end<T>(c, iter);
        if(rc != ATSAT_OK) return fail(&err, rc, c.err);
        HIPC(hipEventRecord(g.ev_post[i], c.stream));
        HIPC(hipStreamWaitEvent(g.stream, g.ev_post[i], 0));
    }
    return ATSAT_OK;
}

#define DISPATCH(c, fn, ...) ((c)->precision == ATSAT_PREC_FP64 ? fn<double>(*(c), ##__VA_ARGS__) : fn<float>(*(c), ##__VA_ARGS__))

// ---- host <-> device copies ----------------------------------------------------------------
template<typename T, typename E>
int copy_field_t(atsat_ctx& c, E* devbase, const double* src, double* dst, double* const* rows_in,
                 double* const* rows_out){
    std::string& err = c.err;
    const Dims& D = c.D;
    E* stage = static_cast<E*>(c.pinned);
    const size_t bytes = (size_t)D.fs * sizeof(E);
    if(src || rows_in){
        for(int i = 0; i < D.im; i++) for(int j = 0; j < D.jm; j++){
            const double* row = rows_in ? rows_in[(size_t)i * D.jm + j] : src + ((size_t)i * D.jm + j) * D.km;
            E* o = stage + D.at(i, j, 0);
            for(int k = 0; k < D.km; k++) o[k] = (E)row[k];
            for(int k = D.km; k < D.kp; k++) o[k] = (E)0;
        }
        HIPC(hipMemcpyAsync(devbase, stage, bytes, hipMemcpyHostToDevice, c.stream));
        HIPC(hipStreamSynchronize(c.stream));
    } else {
        HIPC(hipMemcpyAsync(stage, devbase, bytes, hipMemcpyDeviceToHost, c.stream));
        HIPC(hipStreamSynchronize(c.stream));
        for(int i = 0; i < D.im; i++) for(int j = 0; j < D.jm; j++){
            double* row = rows_out ? rows_out[(size_t)i * D.jm + j] : dst + ((size_t)i * D.jm + j) * D.km;
            const E* o = stage + D.at(i, j, 0);
            for(int k = 0; k < D.km; k++) row[k] = (double)o[k];
        }
    }
    return ATSAT_OK;
}

template<typename T>
int copy_field(atsat_ctx& c, int member, int field, const double* src, double* dst, double* const* rows_in,
               double* const* rows_out){
    const bool upload = src || rows_in;
    const bool nfield = field >= ATSAT_F_N0 && field < ATSAT_F_N0 + ATSAT_NPROG;
    int f = field;
    if(c.n_alias){
        // an upload of a live or *n field separates the two; a download of an *n field reads the
        // live state it is equal to
        if(upload) materialize_n<T>(c);
        else if(nfield && !n_split(c, field - ATSAT_F_N0)) f = field - ATSAT_F_N0;
    }
    Bufs<T>& B = c.bufs<T>();
    if(is_state(f)) return copy_field_t<T, double>(c, B.s[f] + (size_t)member * c.D.fs, src, dst, rows_in, rows_out);
    return copy_field_t<T, T>(c, B.w[f] + (size_t)member * c.D.fs, src, dst, rows_in, rows_out);
}

int vec_info(atsat_ctx& c, int member, int vec, double** dev, size_t* n){
    const atsat_grid& g = c.grid; const int M = c.D.M;
    Bufs<double>& Bd = c.bd; Bufs<float>& Bf = c.bf;
    const bool f64 = c.precision == ATSAT_PREC_FP64;
    double* vb = f64 ? Bd.vbar0 : Bf.vbar0; double* ub = f64 ? Bd.ubar0 : Bf.ubar0;
    double* br = f64 ? Bd.buoy_ref : Bf.buoy_ref; double* sr = f64 ? Bd.srf : Bf.srf;
    double* vs = f64 ? Bd.vel_star : Bf.vel_star;
    (void)M;
    switch(vec){
    case ATSAT_V_CELL_VBAR0: *n = (size_t)g.im * g.jm; *dev = vb + (size_t)member * *n; return ATSAT_OK;
    case ATSAT_V_CELL_UBAR0: *n = (size_t)g.im * g.jm; *dev = ub + (size_t)member * *n; return ATSAT_OK;
    case ATSAT_V_BUOY_REF: *n = (size_t)g.im; *dev = br + (size_t)member * *n; return ATSAT_OK;
    case ATSAT_V_VEL_STAR: *n = (size_t)g.jm * g.km; *dev = vs + (size_t)member * *n; return ATSAT_OK;
    default:
        if(vec >= ATSAT_V_SRF_H2O && vec <= ATSAT_V_SRF_TOTAL){
            *n = (size_t)g.jm * g.km;
            *dev = sr + ((size_t)member * 5 + (vec - ATSAT_V_SRF_H2O)) * *n;
            return ATSAT_OK;
        }
    }
    return fail(&c.err, ATSAT_E_ARG, "bad vec id");
}

int check_ctx(atsat_ctx* c){
    if(!c) return fail(nullptr, ATSAT_E_ARG, "null context");
    c->pterm_iter = -1;                    // the caller may have changed p_dyn, p_stat or rho_mix
    return ATSAT_OK;
}

// The convective adjustment's scratch, allocated once (outside any graph capture) when the knobs
// ask for the adjustment.
int ensure_mca(atsat_ctx& c){
    std::string& err = c.err;
    if(c.mca_mem) return ATSAT_OK;
    if(c.grid.im > kMcaMaxLevels)
        return fail(&c.err, ATSAT_E_UNSUPPORTED, "the convective adjustment handles columns of at most 64 levels");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPC(hipStreamIsCapturing(c.stream, &cap));
    if(cap != hipStreamCaptureStatusNone) return fail(&c.err, ATSAT_E_ARG, "the convective adjustment's scratch cannot be allocated during a capture");
    const size_t nd = mca_scratch_doubles(c.D), nf = (size_t)c.D.M * c.D.fs;
    c.mca_bytes = nd * sizeof(double) + nf;
    if(hipMalloc(&c.mca_mem, c.mca_bytes) != hipSuccess){
        c.mca_mem = nullptr; c.mca_bytes = 0;
        return fail(&c.err, ATSAT_E_NOMEM, "hipMalloc of the convective adjustment's scratch failed");
    }
    double* d = static_cast<double*>(c.mca_mem);
    unsigned char* f = reinterpret_cast<unsigned char*>(d + nd);
    c.bd.mca = d; c.bd.mca_sf = f; c.bf.mca = d; c.bf.mca_sf = f;
    if(mca_store_constants(c.D, c.S, d) != 0)
        return fail(&c.err, ATSAT_E_HIP, "copy of the convective adjustment's constants failed");
    return ATSAT_OK;
}

// The solver modes' tables and scratch, allocated once (outside any graph capture) when the knobs
// ask for the line pass or the barotropic correction.
int ensure_pmodes(atsat_ctx& c){
    std::string& err = c.err;
    if(c.pm_mem) return ATSAT_OK;
    if(c.knobs.press_line && c.grid.im > pmodes_max_levels())
        return fail(&c.err, ATSAT_E_UNSUPPORTED, "the line pass handles columns of at most 64 levels");
    if(c.grid.im < 4 || c.grid.jm < 4)
        return fail(&c.err, ATSAT_E_UNSUPPORTED, "the solver modes need at least 4 levels and 4 latitudes");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPC(hipStreamIsCapturing(c.stream, &cap));
    if(cap != hipStreamCaptureStatusNone) return fail(&c.err, ATSAT_E_ARG, "the solver modes' scratch cannot be allocated during a capture");
    c.pm_bytes = pmodes_doubles(c.D) * sizeof(double);
    if(hipMalloc(&c.pm_mem, c.pm_bytes) != hipSuccess){
        c.pm_mem = nullptr; c.pm_bytes = 0;
        return fail(&c.err, ATSAT_E_NOMEM, "hipMalloc of the solver modes' scratch failed");
    }
    double* d = static_cast<double*>(c.pm_mem);
    c.bd.pmodes = d; c.bf.pmodes = d;
    if(pmodes_store_tables(c.D, c.S, c.h_thez_sin, d) != 0)
        return fail(&c.err, ATSAT_E_HIP, "copy of the solver modes' tables failed");
    double A2 = 0.0;
    for(int i = 1; i < c.grid.im - 1; i++){ const double rm = c.h_rm[i]; A2 += 1.0 / (rm * rm); }
    c.pm_A2 = A2;
    return ATSAT_OK;
}

// Build the separable pressure solver once per context (the host eigen-decompositions, the
// device tables and the correction buffers of the context's precision), and run one correction
// on zeros so that every library kernel is loaded before a graph may capture the solve.
int ensure_solver(atsat_ctx& c){
    std::string& err = c.err;
    const bool fp32 = fp32_corrections(c);
    if(!c.solver){
        std::string e;
        c.solver = solver_create(c.D, c.S, c.h_thez_sin.data(), c.h_rm.data(), c.stream, &e);
        if(!c.solver) return fail(&c.err, ATSAT_E_UNSUPPORTED, e);
    }
    int rc = solver_prepare(c.solver, fp32);
    if(rc != ATSAT_OK) return fail(&c.err, rc, solver_error(c.solver));
    rc = solver_warmup(c.solver, fp32);
    if(rc != ATSAT_OK) return fail(&c.err, rc, solver_error(c.solver));
    HIPC(hipStreamSynchronize(c.stream));
    return ATSAT_OK;
}

template<typename T> static int k_mix(atsat_ctx& c){ launch_mixture_density<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("mixture_density"); }
template<typename T> static int k_damp(atsat_ctx& c){ launch_damp_pdyn<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("damp_pdyn"); }
template<typename T> static int k_pstat(atsat_ctx& c){ launch_rebuild_pstat<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("rebuild_pstat"); }
template<typename T> static int k_prec(atsat_ctx& c){ launch_precipitation<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("precipitation"); }
template<typename T> static int k_sync(atsat_ctx& c){
    if(!c.n_alias) launch_copy_n<T>(c.D, c.S, c.bufs<T>(), kSyncMask, c.stream);
    return c.check_launch("rk_scalar_sync"); }
template<typename T> static int k_bcs(atsat_ctx& c){ launch_bc<T>(c.D, c.S, c.bufs<T>(), false, c.stream); return c.check_launch("boundary_conditions"); }
template<typename T> static int k_clmp(atsat_ctx& c){ launch_clamp<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("clamp_species"); }
template<typename T> static int k_rest(atsat_ctx& c){ restore_var<T>(c); return c.check_launch("restore_var"); }
template<typename T> static int k_nh4sh(atsat_ctx& c){
    PROF(c, ATSAT_K_CHEM, launch_nh4sh_equilibrium<T>(c.D, c.S, c.bufs<T>(), c.stream));
    return c.check_launch("nh4sh_equilibrium"); }
template<typename T> static int k_chem(atsat_ctx& c){
    PROF(c, ATSAT_K_CHEM, launch_chemistry<T>(c.D, c.S, c.bufs<T>(), c.stream));
    return c.check_launch("chemistry"); }
template<typename T> static int k_turb(atsat_ctx& c, bool init){
    if(!c.S.turb_active) return fail(&c.err, ATSAT_E_UNSUPPORTED, "the turbulence closure needs turb_active");
    if(init) materialize_n<T>(c);
    PROF(c, ATSAT_K_TURB, launch_turbulence<T>(c.D, c.S, c.bufs<T>(), init, c.stream));
    return c.check_launch(init ? "turbulence_init" : "turbulence");
}
template<typename T> static int k_budg(atsat_ctx& c, double* out){
    std::string& err = c.err;
    Bufs<T>& B = c.bufs<T>();
    double* dout = B.budget + (size_t)c.D.M * 4 * c.D.im;
    launch_budgets<T>(c.D, c.S, B, B.budget, dout, c.stream);
    int rc = c.check_launch("budgets");
    if(rc) return rc;
    HIPC(hipMemcpyAsync(out, dout, (size_t)c.D.M * (3 + c.D.im) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPC(hipStreamSynchronize(c.stream));
    return ATSAT_OK;
}

template<typename T> static int k_cens(atsat_ctx& c){ launch_census<T>(c.D, c.S, c.bufs<T>(), c.stream); return c.check_launch("census"); }
template<typename T> static int k_sat(atsat_ctx& c, int gas){ launch_satadj<T>(c.D, c.S, c.bufs<T>(), gas, gas, c.stream); return c.check_launch("saturation_adjust"); }

template<typename T>
static int k_poisson(atsat_ctx& c, const double* src, double tol, int max_iter, int* iters, double* rel){
    Bufs<T>& B = c.bufs<T>();
    int rc;
    if(src){
        for(int m = 0; m < c.D.M; m++){
            rc = copy_field_t<T, double>(c, B.s[B_PSRC] + (size_t)m * c.D.fs,
                                         src + (size_t)m * c.D.im * c.D.jm * c.D.km, nullptr, nullptr, nullptr);
            if(rc) return rc;
        }
    } else {
        launch_press_prepare<T>(c.D, c.S, B, c.stream);
        launch_press_src<T>(c.D, c.S, B, c.stream);
    }
    return run_solver<T>(c, tol, max_iter, iters, rel, c.stream);
}



}  // namespace

// =============================================================================================
extern "C" {

const char* atsat_last_error(atsat_ctx* ctx){ return ctx ? ctx->err.c_str() : g_err.c_str(); }

int atsat_ctx_create(const atsat_grid* g, const atsat_consts* cs, const atsat_knobs* k,
                     int n_members, int precision, atsat_ctx** out){
    if(!g || !cs || !k || !out) return fail(nullptr, ATSAT_E_ARG, "null argument");
    *out = nullptr;
    if(g->im < 5 || g->jm < 7 || g->km < 6 || ((g->km - 1) & 1))
        return fail(nullptr, ATSAT_E_ARG, "grid too small, or km-1 odd (the red-black ring needs an even count)");
    if(g->km - 1 > 768) return fail(nullptr, ATSAT_E_UNSUPPORTED, "more than 768 longitudes");
    if(n_members < 1) return fail(nullptr, ATSAT_E_ARG, "n_members < 1");
    if(precision != ATSAT_PREC_FP64 && precision != ATSAT_PREC_FP32) return fail(nullptr, ATSAT_E_ARG, "bad precision");
    std::unique_ptr<atsat_ctx> c(new atsat_ctx());
    c->grid = *g; c->consts = *cs; c->knobs = *k; c->precision = precision;
    std::string& err = c->err;
    if(!uniform_layers(*c)) return fail(nullptr, ATSAT_E_UNSUPPORTED, "non-uniform layer thickness");
    c->elem = precision == ATSAT_PREC_FP64 ? 8 : 4;
    Dims& D = c->D;
    D.im = g->im; D.jm = g->jm; D.km = g->km; D.kp = (g->km + 63) / 64 * 64; D.M = n_members;
    D.fs = (long long)D.im * D.jm * D.kp;
    if((long long)D.fs * D.M >= (1ll << 31)) return fail(nullptr, ATSAT_E_ARG, "grid x members too large");
    make_scal(*c);
    c->tile = rk_default_tile(D);
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPC(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&c->ev_side_fork, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&c->ev_side_join, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&c->ev_phys_fork, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&c->ev_phys_join, hipEventDisableTiming));
    c->arena_bytes = 0;
    for(int b = 0; b < NBUF; b++) c->arena_bytes += buf_bytes(b, (size_t)D.fs * D.M, c->elem);
    c->arena_bytes += buoy_bytes(D);
    if(hipMalloc(&c->arena, c->arena_bytes) != hipSuccess){
        c->arena = nullptr;
        return fail(nullptr, ATSAT_E_NOMEM, "hipMalloc of the field arena failed");
    }
    HIPC(hipMemsetAsync(c->arena, 0, c->arena_bytes, c->stream));
    c->small_bytes = small_elems(*g, n_members) * sizeof(double);
    if(hipMalloc(reinterpret_cast<void**>(&c->d_small), c->small_bytes) != hipSuccess){
        c->d_small = nullptr;
        return fail(nullptr, ATSAT_E_NOMEM, "hipMalloc of the tables failed");
    }
    HIPC(hipMemsetAsync(c->d_small, 0, c->small_bytes, c->stream));
    HIPC(hipHostMalloc(&c->pinned, (size_t)D.fs * sizeof(double), hipHostMallocDefault));
    if(precision == ATSAT_PREC_FP64) assign_bufs<double>(*c); else assign_bufs<float>(*c);
    std::vector<double> geo, sinj, sabs, rmv;
    make_geo(*c, geo, sinj, sabs, rmv);
    Bufs<double>& Bd = c->bd; Bufs<float>& Bf = c->bf;
    const bool f64 = precision == ATSAT_PREC_FP64;
    HIPC(hipMemcpyAsync(f64 ? Bd.geo : Bf.geo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(f64 ? Bd.thez_sin : Bf.thez_sin, sinj.data(), sinj.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(f64 ? Bd.sin_abs : Bf.sin_abs, sabs.data(), sabs.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(f64 ? Bd.rm : Bf.rm, rmv.data(), rmv.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    c->h_thez_sin = sinj;
    c->h_rm = rmv;
    if(k->press_mode == 1){ int rc = ensure_solver(*c); if(rc) return rc; }
    if(k->mca_stride > 0){ int rc = ensure_mca(*c); if(rc) return fail(nullptr, rc, c->err); }
    if(k->press_line || k->press_baro){ int rc = ensure_pmodes(*c); if(rc) return fail(nullptr, rc, c->err); }
    *out = c.release();
    return ATSAT_OK;
}

int atsat_ctx_destroy(atsat_ctx* ctx){
    if(!ctx) return ATSAT_OK;
    if(ctx->stream) hipStreamSynchronize(ctx->stream);
    delete ctx;
    return ATSAT_OK;
}

int atsat_set_knobs(atsat_ctx* ctx, const atsat_knobs* k){
    if(check_ctx(ctx) || !k) return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "null argument");
    // the aliasing of the *n copies depends on turb_active: make them explicit under the old knobs
    if(ctx->n_alias) DISPATCH(ctx, materialize_n);
    ctx->knobs = *k;
    make_scal(*ctx);
    if(ctx->graph_exec){ hipGraphExecDestroy(ctx->graph_exec); ctx->graph_exec = nullptr; }
    if(ctx->graph){ hipGraphDestroy(ctx->graph); ctx->graph = nullptr; }
    if(k->mca_stride > 0){ int rc = ensure_mca(*ctx); if(rc) return rc; }
    if(k->press_line || k->press_baro){ int rc = ensure_pmodes(*ctx); if(rc) return rc; }
    if(k->press_mode == 1) return ensure_solver(*ctx);
    return ATSAT_OK;
}

int atsat_device_bytes(atsat_ctx* ctx, size_t* bytes){
    if(check_ctx(ctx) || !bytes) return ATSAT_E_ARG;
    *bytes = ctx->arena_bytes + ctx->small_bytes + ctx->mca_bytes + ctx->pm_bytes + solver_device_bytes(ctx->solver);
    return ATSAT_OK;
}

int atsat_upload(atsat_ctx* ctx, int member, int field, const double* src){
    if(check_ctx(ctx) || !src || member < 0 || member >= ctx->D.M || field < 0 || field >= ATSAT_NFIELDS)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad upload argument");
    return DISPATCH(ctx, copy_field, member, field, src, nullptr, nullptr, nullptr);
}
int atsat_download(atsat_ctx* ctx, int member, int field, double* dst){
    if(check_ctx(ctx) || !dst || member < 0 || member >= ctx->D.M || field < 0 || field >= ATSAT_NFIELDS)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad download argument");
    return DISPATCH(ctx, copy_field, member, field, nullptr, dst, nullptr, nullptr);
}
int atsat_upload_rows(atsat_ctx* ctx, int member, int field, double* const* rows){
    if(check_ctx(ctx) || !rows || member < 0 || member >= ctx->D.M || field < 0 || field >= ATSAT_NFIELDS)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad upload_rows argument");
    return DISPATCH(ctx, copy_field, member, field, nullptr, nullptr, rows, nullptr);
}
int atsat_download_rows(atsat_ctx* ctx, int member, int field, double* const* rows){
    if(check_ctx(ctx) || !rows || member < 0 || member >= ctx->D.M || field < 0 || field >= ATSAT_NFIELDS)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad download_rows argument");
    return DISPATCH(ctx, copy_field, member, field, nullptr, nullptr, nullptr, rows);
}
// restart files through the device, one array at a time (restart.cpp holds the format)
int atsat_save_restart(atsat_ctx* ctx, int member, int iter, const char* path, const atsat_restart_extra* x,
                       int format){
    if(check_ctx(ctx) || !path || member < 0 || member >= ctx->D.M)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad save_restart argument");
    const size_t n = (size_t)ctx->grid.im * ctx->grid.jm * ctx->grid.km;
    bool zeroed = false;
    std::string err;
    const int rc = restart_write(path, ctx->grid, iter, format,
        [&](const RestartItem& it, double* scratch, const double** a) -> int {
            if(it.extra >= 0 && x && x->a[it.extra]){ *a = x->a[it.extra]; return 0; }
            *a = scratch;
            if(it.field >= 0 || it.extra == ATSAT_RX_P_DYNN){         // restoreVar: p_dynn = p_dyn
                zeroed = false;
                return atsat_download(ctx, member, it.field >= 0 ? it.field : ATSAT_P_DYN, scratch);
            }
            if(!zeroed){ std::memset(scratch, 0, n * sizeof(double)); zeroed = true; }
            return 0;
        }, err);
    if(rc && !err.empty()) return fail(&ctx->err, rc, err);
    return rc;
}

int atsat_load_restart(atsat_ctx* ctx, int member, const char* path, atsat_restart_extra* x, int* iter,
                       int* format){
    if(check_ctx(ctx) || !path || member < 0 || member >= ctx->D.M)
        return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad load_restart argument");
    std::string err;
    const int rc = restart_read(path, ctx->grid, iter, format,
        [&](const RestartItem& it, double* scratch) -> double* {
            return it.field >= 0 ? scratch : (x ? x->a[it.extra] : nullptr);
        },
        [&](const RestartItem& it, const double* d) -> int {
            return it.field >= 0 ? atsat_upload(ctx, member, it.field, d) : 0;
        }, err);
    if(rc && !err.empty()) return fail(&ctx->err, rc, err);
    return rc;
}

int atsat_upload_vec(atsat_ctx* ctx, int member, int vec, const double* src){
    if(check_ctx(ctx) || !src || member < 0 || member >= ctx->D.M) return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad argument");
    double* dev; size_t n; int rc = vec_info(*ctx, member, vec, &dev, &n); if(rc) return rc;
    std::string& err = ctx->err;
    HIPC(hipMemcpyAsync(dev, src, n * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ATSAT_OK;
}
int atsat_download_vec(atsat_ctx* ctx, int member, int vec, double* dst){
    if(check_ctx(ctx) || !dst || member < 0 || member >= ctx->D.M) return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad argument");
    double* dev; size_t n; int rc = vec_info(*ctx, member, vec, &dev, &n); if(rc) return rc;
    std::string& err = ctx->err;
    HIPC(hipMemcpyAsync(dst, dev, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return ATSAT_OK;
}
int atsat_upload_state(atsat_ctx* ctx, int member, const atsat_host_state* s){
    if(check_ctx(ctx) || !s) return ATSAT_E_ARG;
    for(int f = 0; f < ATSAT_NFIELDS; f++) if(s->f[f]){ int rc = atsat_upload(ctx, member, f, s->f[f]); if(rc) return rc; }
    for(int v = 0; v < ATSAT_NVECS; v++) if(s->v[v]){ int rc = atsat_upload_vec(ctx, member, v, s->v[v]); if(rc) return rc; }
    return ATSAT_OK;
}
int atsat_download_state(atsat_ctx* ctx, int member, const atsat_host_state* s){
    if(check_ctx(ctx) || !s) return ATSAT_E_ARG;
    for(int f = 0; f < ATSAT_NFIELDS; f++) if(s->f[f]){ int rc = atsat_download(ctx, member, f, s->f[f]); if(rc) return rc; }
    for(int v = 0; v < ATSAT_NVECS; v++) if(s->v[v]){ int rc = atsat_download_vec(ctx, member, v, s->v[v]); if(rc) return rc; }
    return ATSAT_OK;
}

#define ENTRY(name, body) int name(atsat_ctx* ctx){ if(check_ctx(ctx)) return ATSAT_E_ARG; body; }

ENTRY(atsat_rk4_step, return DISPATCH(ctx, do_rk4_api))
ENTRY(atsat_pressure_solve, return DISPATCH(ctx, do_pressure, ctx->stream))
ENTRY(atsat_precipitation, return DISPATCH(ctx, k_prec))
ENTRY(atsat_mixture_density, return DISPATCH(ctx, k_mix))
ENTRY(atsat_damp_pdyn, return DISPATCH(ctx, k_damp))
ENTRY(atsat_rebuild_pstat, return DISPATCH(ctx, k_pstat))
ENTRY(atsat_rk_scalar_sync, return DISPATCH(ctx, k_sync))
ENTRY(atsat_boundary_conditions, return DISPATCH(ctx, k_bcs))
ENTRY(atsat_clamp_species, return DISPATCH(ctx, k_clmp))
ENTRY(atsat_restore_var, return DISPATCH(ctx, k_rest))
ENTRY(atsat_turbulence, return DISPATCH(ctx, k_turb, false))
ENTRY(atsat_turbulence_init, return DISPATCH(ctx, k_turb, true))
ENTRY(atsat_nh4sh_equilibrium, return DISPATCH(ctx, k_nh4sh))
ENTRY(atsat_chemistry, return DISPATCH(ctx, k_chem))
ENTRY(atsat_convective_adjustment, { int rc = ensure_mca(*ctx); if(rc) return rc; return DISPATCH(ctx, do_mca); })

int atsat_saturation_adjust(atsat_ctx* ctx, int gas){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    if(gas < ATSAT_GAS_CH4 || gas > ATSAT_GAS_NH3) return fail(&ctx->err, ATSAT_E_ARG, "bad gas");
    return DISPATCH(ctx, k_sat, gas);
}

int atsat_pressure_last(atsat_ctx* ctx, int* iters, double* rel_residual){
    if(check_ctx(ctx) || !iters || !rel_residual) return ATSAT_E_ARG;
    if(!ctx->solver){ *iters = 0; *rel_residual = -1.0; return ATSAT_OK; }
    solver_last(ctx->solver, iters, rel_residual);
    return ATSAT_OK;
}

int atsat_poisson_solve(atsat_ctx* ctx, const double* src, double tol, int max_iter, int* iters, double* rel_residual){
    if(check_ctx(ctx) || !iters || !rel_residual || max_iter < 0) return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "bad poisson_solve argument");
    int rc = ensure_solver(*ctx);
    if(rc) return rc;
    return DISPATCH(ctx, k_poisson, src, tol, max_iter, iters, rel_residual);
}

int atsat_step(atsat_ctx* ctx, int iter){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    return DISPATCH(ctx, do_step, iter);
}

// One iteration inside atsat_run (which checked the context): the loop's iterations follow each
// other with nothing in between, so an odd iteration keeps the even one's static pressure terms.
static int run_step(atsat_ctx* ctx, int iter){ return DISPATCH(ctx, do_step, iter); }

// The outcome of the converged solves replayed from a graph since the last synchronisation: they
// do not check their residual on the host, the device records it (solver_status).
static int graph_solve_status(atsat_ctx& c){
    if(!c.solver) return ATSAT_OK;
    int st = 0;
    if(solver_status(c.solver, &st) != ATSAT_OK) return fail(&c.err, ATSAT_E_HIP, solver_error(c.solver));
    if(st == 2) return fail(&c.err, ATSAT_E_NONFINITE, "a converged pressure solve replayed from a graph ended with a non-finite residual");
    if(st == 1) return fail(&c.err, ATSAT_E_NOCONV, "a converged pressure solve replayed from a graph ended above its residual limit");
    return ATSAT_OK;
}

int atsat_synchronize(atsat_ctx* ctx){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    std::string& err = ctx->err;
    HIPC(hipStreamSynchronize(ctx->stream));
    return graph_solve_status(*ctx);
}

int atsat_run(atsat_ctx* ctx, int first_iter, int n_steps, int use_graph){
    if(check_ctx(ctx) || n_steps < 0) return ATSAT_E_ARG;
    std::string& err = ctx->err;
    int it = first_iter, left = n_steps, rc;
    // the graph holds one even + odd pair of iterations: an adjustment stride above 2 does not
    // repeat with it, so such runs launch iteration by iteration
    if(ctx->knobs.mca_stride > 2) use_graph = 0;
    if(!use_graph){
        for(; left > 0; left--, it++) if((rc = run_step(ctx, it)) != ATSAT_OK) return rc;
        return ATSAT_OK;
    }
    if(it % 2 != 0 && left > 0){ if((rc = run_step(ctx, it)) != ATSAT_OK) return rc; it++; left--; }
    if(left >= 2 && !ctx->graph_exec){
        // capture one even + odd pair of iterations
        HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc1 = run_step(ctx, 0), rc2 = (rc1 == ATSAT_OK) ? run_step(ctx, 1) : rc1;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if(rc2 != ATSAT_OK){ if(g) hipGraphDestroy(g); return rc2; }
        if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        ctx->graph = g;
        HIPC(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
    }
    for(; left >= 2; left -= 2, it += 2) HIPC(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    if(left > 0){ if((rc = run_step(ctx, it)) != ATSAT_OK) return rc; }
    return ATSAT_OK;
}

static void drop_graph(atsat_ctx& c){
    if(c.graph_exec){ hipGraphExecDestroy(c.graph_exec); c.graph_exec = nullptr; }
    if(c.graph){ hipGraphDestroy(c.graph); c.graph = nullptr; }
}

int atsat_autotune(atsat_ctx* ctx, int* band, int* rings, double* stage_ms){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    int rc = DISPATCH(ctx, do_autotune, stage_ms);
    drop_graph(*ctx);                                       // captured with the old tile
    if(band) *band = ctx->tile.band;
    if(rings) *rings = ctx->tile.rings;
    return rc;
}

int atsat_set_tile(atsat_ctx* ctx, int band, int rings){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    const RkTile t{band, rings};
    if(!rk_tile_ok(ctx->D, t))
        return fail(&ctx->err, ATSAT_E_ARG, "bad tile: rings must divide band, and rings x (km - 1 rounded up "
                                            "to 64) must not exceed " + std::to_string(kRowThreads) + " threads");
    ctx->tile = t;
    drop_graph(*ctx);
    return ATSAT_OK;
}

int atsat_get_tile(atsat_ctx* ctx, int* band, int* rings){
    if(check_ctx(ctx) || !band || !rings) return ATSAT_E_ARG;
    *band = ctx->tile.band;
    *rings = ctx->tile.rings;
    return ATSAT_OK;
}

// ---- batches of grids ---------------------------------------------------------------------------
static int check_group(atsat_group* g){ return g ? ATSAT_OK : fail(nullptr, ATSAT_E_ARG, "null group"); }

// The batched launch's workgroup width: the tuned width, widened to hold every context's tile.
static int group_width(const atsat_group& g){
    int threads = g.threads;
    for(atsat_ctx* c : g.ctx) threads = std::max(threads, rk_ring_threads(c->D) * c->tile.rings);
    return threads;
}

static int group_refresh(atsat_group& g){
    std::vector<RkTile> tiles;
    const int threads = group_width(g);
    for(atsat_ctx* c : g.ctx) tiles.push_back(c->tile);
    const int old_threads = g.threads, old_blocks = g.blocks;
    int rc = g.precision == ATSAT_PREC_FP64 ? group_layout<double>(g, threads, tiles, -1)
                                            : group_layout<float>(g, threads, tiles, -1);
    bool same = g.threads == old_threads && g.blocks == old_blocks && tiles.size() == g.tiles.size();
    for(size_t i = 0; same && i < tiles.size(); i++)
        same = tiles[i].band == g.tiles[i].band && tiles[i].rings == g.tiles[i].rings;
    if(!same){                                              // a captured graph has the old launch shape
        if(g.graph_exec){ hipGraphExecDestroy(g.graph_exec); g.graph_exec = nullptr; }
        if(g.graph){ hipGraphDestroy(g.graph); g.graph = nullptr; }
        g.tiles = tiles;
    }
    return rc;
}

int atsat_group_create(atsat_ctx* const* ctxs, int n, atsat_group** out){
    if(!ctxs || !out || n < 1 || n > 32) return fail(nullptr, ATSAT_E_ARG, "group: 1 to 32 contexts");
    *out = nullptr;
    auto g = std::make_unique<atsat_group>();
    std::string& err = g->err;
    for(int i = 0; i < n; i++){
        if(!ctxs[i]) return fail(nullptr, ATSAT_E_ARG, "group: null context");
        for(int q = 0; q < i; q++) if(ctxs[q] == ctxs[i]) return fail(nullptr, ATSAT_E_ARG, "group: a context appears twice");
        if(ctxs[i]->precision != ctxs[0]->precision || ctxs[i]->S.turb_active != ctxs[0]->S.turb_active)
            return fail(nullptr, ATSAT_E_ARG, "group: contexts must share precision and turb_active");
        g->ctx.push_back(ctxs[i]);
    }
    g->precision = ctxs[0]->precision;
    g->turb = ctxs[0]->S.turb_active != 0;
    HIPC(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&g->ev_rk, hipEventDisableTiming));
    g->ev_pre.assign(n, nullptr);
    g->ev_post.assign(n, nullptr);
    for(int i = 0; i < n; i++){
        HIPC(hipEventCreateWithFlags(&g->ev_pre[i], hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&g->ev_post[i], hipEventDisableTiming));
    }
    const size_t entry = g->precision == ATSAT_PREC_FP64 ? sizeof(RkGrid<double>) : sizeof(RkGrid<float>);
    if(hipMalloc(&g->d_grids, entry * (n + 1)) != hipSuccess){
        g->d_grids = nullptr;
        return fail(nullptr, ATSAT_E_NOMEM, "group: device table");
    }
    if(hipMalloc(reinterpret_cast<void**>(&g->d_heads), kQueues * sizeof(unsigned)) != hipSuccess){
        g->d_heads = nullptr;
        return fail(nullptr, ATSAT_E_NOMEM, "group: work queue");
    }
    int rc = group_refresh(*g);
    if(rc) return fail(nullptr, rc, g->err);
    *out = g.release();
    return ATSAT_OK;
}

int atsat_group_destroy(atsat_group* g){
    delete g;
    return ATSAT_OK;
}

const char* atsat_group_last_error(atsat_group* g){ return g ? g->err.c_str() : g_err.c_str(); }

int atsat_group_autotune(atsat_group* g, double* stage_ms){
    if(check_group(g)) return ATSAT_E_ARG;
    std::string& err = g->err;
    HIPC(hipStreamSynchronize(g->stream));
    return g->precision == ATSAT_PREC_FP64 ? group_autotune<double>(*g, stage_ms) : group_autotune<float>(*g, stage_ms);
}

int atsat_group_tile(atsat_group* g, int i, int* band, int* rings, int* threads){
    if(check_group(g) || i < 0 || i >= (int)g->ctx.size() || !band || !rings || !threads) return ATSAT_E_ARG;
    *band = g->ctx[i]->tile.band;
    *rings = g->ctx[i]->tile.rings;
    *threads = g->merged ? group_width(*g) : rk_ring_threads(g->ctx[i]->D) * g->ctx[i]->tile.rings;
    return ATSAT_OK;
}

int atsat_group_set_merged(atsat_group* g, int merged){
    if(check_group(g) || merged < 0 || merged > 2) return ATSAT_E_ARG;
    const bool m = merged != 0, q = merged == 2;
    if(g->merged != m || g->queue != q){
        g->merged = m;
        g->queue = q;
        if(g->graph_exec){ hipGraphExecDestroy(g->graph_exec); g->graph_exec = nullptr; }
        if(g->graph){ hipGraphDestroy(g->graph); g->graph = nullptr; }
    }
    return ATSAT_OK;
}

int atsat_group_merged(atsat_group* g, int* merged){
    if(check_group(g) || !merged) return ATSAT_E_ARG;
    *merged = !g->merged ? 0 : (g->queue ? 2 : 1);
    return ATSAT_OK;
}

static int group_step_any(atsat_group* g, int iter){
    return g->precision == ATSAT_PREC_FP64 ? group_step<double>(*g, iter) : group_step<float>(*g, iter);
}

int atsat_group_run(atsat_group* g, int first_iter, int n_steps, int use_graph){
    if(check_group(g) || n_steps < 0) return ATSAT_E_ARG;
    std::string& err = g->err;
    // the contexts' scalars and tiles may have changed since the table was written
    HIPC(hipStreamSynchronize(g->stream));
    int rc = group_refresh(*g);
    if(rc) return rc;
    for(atsat_ctx* c : g->ctx) c->pterm_iter = -1;
    // an adjustment stride above 2 does not repeat with the captured even + odd pair
    for(atsat_ctx* c : g->ctx) if(c->knobs.mca_stride > 2) use_graph = 0;
    int it = first_iter, left = n_steps;
    if(!use_graph){
        for(; left > 0; left--, it++) if((rc = group_step_any(g, it)) != ATSAT_OK) return rc;
        return ATSAT_OK;
    }
    if(it % 2 != 0 && left > 0){ if((rc = group_step_any(g, it)) != ATSAT_OK) return rc; it++; left--; }
    if(left >= 2 && !g->graph_exec){
        HIPC(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        int rc1 = group_step_any(g, 0), rc2 = (rc1 == ATSAT_OK) ? group_step_any(g, 1) : rc1;
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(g->stream, &gr);
        if(rc2 != ATSAT_OK){ if(gr) hipGraphDestroy(gr); return rc2; }
        if(e != hipSuccess) return fail(&err, ATSAT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        g->graph = gr;
        HIPC(hipGraphInstantiate(&g->graph_exec, gr, nullptr, nullptr, 0));
    }
    for(; left >= 2; left -= 2, it += 2) HIPC(hipGraphLaunch(g->graph_exec, g->stream));
    if(left > 0){ if((rc = group_step_any(g, it)) != ATSAT_OK) return rc; }
    return ATSAT_OK;
}

int atsat_group_synchronize(atsat_group* g){
    if(check_group(g)) return ATSAT_E_ARG;
    std::string& err = g->err;
    HIPC(hipStreamSynchronize(g->stream));
    for(atsat_ctx* c : g->ctx) HIPC(hipStreamSynchronize(c->stream));
    for(atsat_ctx* c : g->ctx){
        const int rc = graph_solve_status(*c);
        if(rc != ATSAT_OK) return fail(&g->err, rc, c->err);
    }
    return ATSAT_OK;
}

int atsat_profile_enable(atsat_ctx* ctx, int on){
    if(check_ctx(ctx)) return ATSAT_E_ARG;
    ctx->prof_collect();
    ctx->prof = on != 0;
    for(int q = 0; q < ATSAT_NKERNELS; q++){ ctx->prof_ms[q] = 0.0; ctx->prof_n[q] = 0; }
    return ATSAT_OK;
}

int atsat_profile_read(atsat_ctx* ctx, int kernel, double* ms, long long* launches){
    if(check_ctx(ctx) || kernel < 0 || kernel >= ATSAT_NKERNELS || !ms || !launches) return ATSAT_E_ARG;
    ctx->prof_collect();
    *ms = ctx->prof_ms[kernel]; *launches = ctx->prof_n[kernel];
    return ATSAT_OK;
}

int atsat_budgets(atsat_ctx* ctx, double* out){
    if(check_ctx(ctx) || !out) return fail(ctx ? &ctx->err : nullptr, ATSAT_E_ARG, "null argument");
    return DISPATCH(ctx, k_budg, out);
}

int atsat_nonfinite_count(atsat_ctx* ctx, long long* count){
    if(check_ctx(ctx) || !count) return ATSAT_E_ARG;
    std::string& err = ctx->err;
    unsigned long long* dcount = ctx->precision == ATSAT_PREC_FP64 ? ctx->bd.counter : ctx->bf.counter;
    HIPC(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), ctx->stream));
    int rc = DISPATCH(ctx, k_cens);
    if(rc) return rc;
    unsigned long long h = 0;
    HIPC(hipMemcpyAsync(&h, dcount, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    *count = (long long)h;
    return ATSAT_OK;
}

}  // extern "C"
