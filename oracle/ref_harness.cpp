// ============================================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY. Harness around the reference's own sources.
//
// Compiled together with the reference's translation units (oracle/Makefile, outputs only in
// oracle/_ref/) into libatsat_ref.so, it drives the reference's own hot-path functions on a
// state loaded from dense host arrays, so that the CPU restatement (atsat_oracle.cpp) and the
// GPU library can be checked against the reference itself. No reference source is copied
// here: the harness only calls the reference's functions and copies fields in and out.
//
// Access to cSaturnModel's private members: the model befriends every specialisation of
// ProcessDiag (planet/cSaturnModel.h:61, 83); an explicit specialisation for a tag type of the
// harness is such a friend.
// ============================================================================================
#include "cSaturnModel.h"
#include "ChemistrySat.h"
#include "PressureSolverSat.h"
#include "PrecipitationSat.h"
#include "SaturationAdjustmentSat.h"
#include "BC_Sat.h"
#include "TurbulenceSat.h"
#include "ConvectiveAdjustmentSat.h"

#include "atsat_fields.h"
#include "atsat_gpu.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>

struct AtsatRefTag {};

template<> class ProcessDiag<AtsatRefTag> {
public:
    static Array* field(cSaturnModel& m, int id){
        switch(id){
        case ATSAT_T: return &m.t; case ATSAT_U: return &m.u; case ATSAT_V: return &m.v; case ATSAT_W: return &m.w;
        case ATSAT_H2O: return &m.h2o; case ATSAT_H2O_CLOUD: return &m.h2o_cloud; case ATSAT_H2O_ICE: return &m.h2o_ice;
        case ATSAT_CH4: return &m.ch4; case ATSAT_CH4_CLOUD: return &m.ch4_cloud; case ATSAT_CH4_ICE: return &m.ch4_ice;
        case ATSAT_H2S: return &m.h2s;
        case ATSAT_NH3: return &m.nh3; case ATSAT_NH3_CLOUD: return &m.nh3_cloud; case ATSAT_NH3_ICE: return &m.nh3_ice;
        case ATSAT_NH4SH: return &m.nh4sh; case ATSAT_TKE: return &m.tke; case ATSAT_DIS: return &m.dis;
        case ATSAT_F_N0 + ATSAT_T: return &m.tn; case ATSAT_F_N0 + ATSAT_U: return &m.un;
        case ATSAT_F_N0 + ATSAT_V: return &m.vn; case ATSAT_F_N0 + ATSAT_W: return &m.wn;
        case ATSAT_F_N0 + ATSAT_H2O: return &m.h2on; case ATSAT_F_N0 + ATSAT_H2O_CLOUD: return &m.h2o_cloudn;
        case ATSAT_F_N0 + ATSAT_H2O_ICE: return &m.h2o_icen;
        case ATSAT_F_N0 + ATSAT_CH4: return &m.ch4n; case ATSAT_F_N0 + ATSAT_CH4_CLOUD: return &m.ch4_cloudn;
        case ATSAT_F_N0 + ATSAT_CH4_ICE: return &m.ch4_icen;
        case ATSAT_F_N0 + ATSAT_H2S: return &m.h2sn;
        case ATSAT_F_N0 + ATSAT_NH3: return &m.nh3n; case ATSAT_F_N0 + ATSAT_NH3_CLOUD: return &m.nh3_cloudn;
        case ATSAT_F_N0 + ATSAT_NH3_ICE: return &m.nh3_icen;
        case ATSAT_F_N0 + ATSAT_NH4SH: return &m.nh4shn;
        case ATSAT_F_N0 + ATSAT_TKE: return &m.tken; case ATSAT_F_N0 + ATSAT_DIS: return &m.disn;
        case ATSAT_P_DYN: return &m.p_dyn; case ATSAT_P_STAT: return &m.p_stat; case ATSAT_P_METRIC: return &m.p_metric;
        case ATSAT_RHO_MIX: return &m.rho_mix;
        case ATSAT_RHS_U: return &m.rhs_u; case ATSAT_RHS_V: return &m.rhs_v; case ATSAT_RHS_W: return &m.rhs_w;
        case ATSAT_AUX_U: return &m.aux_u; case ATSAT_AUX_V: return &m.aux_v; case ATSAT_AUX_W: return &m.aux_w;
        case ATSAT_NUE: return &m.nue; case ATSAT_TKE_SOURCE: return &m.tke_source; case ATSAT_DIS_SOURCE: return &m.dis_source;
        case ATSAT_MASSFLUX_H2S: return &m.massflux_h2s; case ATSAT_MASSFLUX_NH3: return &m.massflux_nh3;
        case ATSAT_MASSFLUX_NH4SH: return &m.massflux_nh4sh; case ATSAT_FLUXLIM_NH4SH: return &m.fluxlim_nh4sh;
        case ATSAT_Q_PRECIP: return &m.Q_precip;
        case ATSAT_S_PRECIP_H2O: return &m.S_precip_h2o; case ATSAT_S_PRECIP_H2O_CLOUD: return &m.S_precip_h2o_cloud;
        case ATSAT_S_PRECIP_H2O_ICE: return &m.S_precip_h2o_ice;
        case ATSAT_S_PRECIP_NH3: return &m.S_precip_nh3; case ATSAT_S_PRECIP_NH3_CLOUD: return &m.S_precip_nh3_cloud;
        case ATSAT_S_PRECIP_NH3_ICE: return &m.S_precip_nh3_ice;
        case ATSAT_S_PRECIP_CH4: return &m.S_precip_ch4; case ATSAT_S_PRECIP_CH4_CLOUD: return &m.S_precip_ch4_cloud;
        case ATSAT_S_PRECIP_CH4_ICE: return &m.S_precip_ch4_ice;
        case ATSAT_P_RAIN: return &m.P_rain; case ATSAT_P_SNOW: return &m.P_snow; case ATSAT_P_GRAUPEL: return &m.P_graupel;
        case ATSAT_P_NH3_RAIN: return &m.P_nh3_rain; case ATSAT_P_NH3_SNOW: return &m.P_nh3_snow;
        case ATSAT_P_NH3_GRAUPEL: return &m.P_nh3_graupel;
        case ATSAT_P_CH4_RAIN: return &m.P_ch4_rain; case ATSAT_P_CH4_SNOW: return &m.P_ch4_snow;
        case ATSAT_P_CH4_GRAUPEL: return &m.P_ch4_graupel;
        case ATSAT_P_NH4SH: return &m.P_nh4sh;
        case ATSAT_PROD: return &m.prod;
        case ATSAT_DIFFLUX_H2S: return &m.difflux_h2s; case ATSAT_DIFFLUX_NH3: return &m.difflux_nh3;
        case ATSAT_DIFFLUX_NH4SH: return &m.difflux_nh4sh; case ATSAT_THERMALMASSFLUX: return &m.thermalmassflux;
        case ATSAT_P_BARO: return nullptr;           // a std::vector, handled by the caller
        default: return nullptr;
        }
    }
    static Array_2D* srf(cSaturnModel& m, int v){
        switch(v){
        case ATSAT_V_SRF_H2O: return &m.precip_srf_h2o; case ATSAT_V_SRF_NH3: return &m.precip_srf_nh3;
        case ATSAT_V_SRF_CH4: return &m.precip_srf_ch4; case ATSAT_V_SRF_NH4SH: return &m.precip_srf_nh4sh;
        case ATSAT_V_SRF_TOTAL: return &m.precip_srf_total;
        default: return nullptr;
        }
    }
    static cSaturnModel* create(const atsat_knobs& k){
        cSaturnModel* m = new cSaturnModel();
        AtomUtils::set_phi_seam_periodic(ATPhys::phi_periodic<cSaturnModel>());   // Run, cSaturnModel.cpp:383
        // saturn/config_atsat.xml as shipped (the constructor's defaults are param.py's)
        m->re = 100000.0; m->pr = 69.0; m->g = 10.0; m->omega = 0.000163; m->gam = 0.7;
        m->u_0 = 470.0; m->t_ref = 134.0; m->p_ref = 1.0; m->R_ref = 3.89; m->L_atm = 500.0;
        m->Coriolis = 1.0; m->centrifugal = 1.0; m->buoyancy = 1.0;
        m->t_pole = 380.0; m->t_equator = 400.0; m->ta = 1.0;
        m->chemical_reaction = k.chemical_reaction;
        m->init_layer_heights();
        m->resetArrays();
        ChemistrySat(*m).ThermalPropertiesSat();
        m->dt = k.dt;
        m->turb_active = k.turb_active != 0;
        const size_t N = (size_t)m->im * m->jm * m->km;
        m->p_baro.assign(N, 0.0);
        m->cell_vbar0.assign((size_t)m->im * m->jm, 0.0);
        m->cell_ubar0.assign((size_t)m->im * m->jm, 0.0);
        m->buoy_ref_level.assign(m->im, 0.0);
        return m;
    }
    static int load(cSaturnModel& m, const atsat_host_state& s){
        const int im = m.im, jm = m.jm, km = m.km;
        for(int f = 0; f < ATSAT_NFIELDS; f++){
            if(!s.f[f]) continue;
            if(f == ATSAT_P_BARO){ std::memcpy(m.p_baro.data(), s.f[f], sizeof(double) * im * jm * km); continue; }
            Array* A = field(m, f);
            if(!A) return -1;
            for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++)
                std::memcpy(A->x[i][j], s.f[f] + ((size_t)i * jm + j) * km, sizeof(double) * km);
        }
        if(s.v[ATSAT_V_CELL_VBAR0]) std::memcpy(m.cell_vbar0.data(), s.v[ATSAT_V_CELL_VBAR0], sizeof(double) * im * jm);
        if(s.v[ATSAT_V_CELL_UBAR0]) std::memcpy(m.cell_ubar0.data(), s.v[ATSAT_V_CELL_UBAR0], sizeof(double) * im * jm);
        return 0;
    }
    static int store(cSaturnModel& m, const atsat_host_state& s){
        const int im = m.im, jm = m.jm, km = m.km;
        for(int f = 0; f < ATSAT_NFIELDS; f++){
            if(!s.f[f]) continue;
            if(f == ATSAT_P_BARO){ std::memcpy(s.f[f], m.p_baro.data(), sizeof(double) * im * jm * km); continue; }
            Array* A = field(m, f);
            if(!A) return -1;
            for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++)
                std::memcpy(s.f[f] + ((size_t)i * jm + j) * km, A->x[i][j], sizeof(double) * km);
        }
        if(s.v[ATSAT_V_BUOY_REF]) for(int i = 0; i < im; i++) s.v[ATSAT_V_BUOY_REF][i] = m.buoy_ref_level[i];
        for(int v = ATSAT_V_SRF_H2O; v <= ATSAT_V_SRF_TOTAL; v++){
            if(!s.v[v]) continue;
            Array_2D* A = srf(m, v);
            for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++) s.v[v][(size_t)j * km + k] = A->y[j][k];
        }
        if(s.v[ATSAT_V_VEL_STAR])
            for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++) s.v[ATSAT_V_VEL_STAR][(size_t)j * km + k] = m.vel_star.y[j][k];
        return 0;
    }
    // The physics-block RK scalar sync is inline code of Run (cSaturnModel.cpp:955-977), so the
    // harness performs it on the model's own arrays.
    static void rk_sync(cSaturnModel& m){
        for(int i = 0; i < m.im; i++) for(int j = 0; j < m.jm; j++) for(int k = 0; k < m.km; k++){
            m.h2on.x[i][j][k] = m.h2o.x[i][j][k]; m.h2o_cloudn.x[i][j][k] = m.h2o_cloud.x[i][j][k];
            m.h2o_icen.x[i][j][k] = m.h2o_ice.x[i][j][k];
            m.nh3n.x[i][j][k] = m.nh3.x[i][j][k]; m.nh3_cloudn.x[i][j][k] = m.nh3_cloud.x[i][j][k];
            m.nh3_icen.x[i][j][k] = m.nh3_ice.x[i][j][k];
            m.ch4n.x[i][j][k] = m.ch4.x[i][j][k]; m.ch4_cloudn.x[i][j][k] = m.ch4_cloud.x[i][j][k];
            m.ch4_icen.x[i][j][k] = m.ch4_ice.x[i][j][k];
            m.tn.x[i][j][k] = m.t.x[i][j][k];
            m.h2sn.x[i][j][k] = m.h2s.x[i][j][k]; m.nh4shn.x[i][j][k] = m.nh4sh.x[i][j][k];
        }
    }
    static void satadj(cSaturnModel& m, int gas){
        if(gas == ATSAT_GAS_CH4)
            SaturationAdjustmentSat(m).run("CH4", m.coeff_ch4_A, m.coeff_ch4_B, m.coeff_ch4_A_i, m.coeff_ch4_B_i,
                m.t_0_ch4, m.t_00_ch4, m.ep_ch4, m.lv_ch4, m.ls_ch4, m.cp_ch4, m.r_ch4,
                m.C_ch4, m.L0_ch4, m.R_ch4, m.del_alf_ch4, m.del_bet_ch4, m.m_ch4, m.ch4, m.ch4_cloud, m.ch4_ice);
        else if(gas == ATSAT_GAS_H2O)
            SaturationAdjustmentSat(m).run("H2O", m.coeff_h2o_A, m.coeff_h2o_B, m.coeff_h2o_A_i, m.coeff_h2o_B_i,
                m.t_0_h2o, m.t_00_h2o, m.ep_h2o, m.lv_h2o, m.ls_h2o, m.cp_h2o, m.r_h2o,
                m.C_h2o, m.L0_h2o, m.R_h2o, m.del_alf_h2o, m.del_bet_h2o, m.m_h2o, m.h2o, m.h2o_cloud, m.h2o_ice);
        else
            SaturationAdjustmentSat(m).run("NH3", m.coeff_nh3_A, m.coeff_nh3_B, m.coeff_nh3_A_i, m.coeff_nh3_B_i,
                m.t_0_nh3, m.t_00_nh3, m.ep_nh3, m.lv_nh3, m.ls_nh3, m.cp_nh3, m.r_nh3,
                m.C_nh3, m.L0_nh3, m.R_nh3, m.del_alf_nh3, m.del_bet_nh3, m.m_nh3, m.nh3, m.nh3_cloud, m.nh3_ice);
    }
    static int turb_closure, chem_block, mca_stride;
    // the restart arrays the library does not hold, ATSAT_RX_* order (atsat_gpu.h)
    static Array* extra(cSaturnModel& m, int x){
        switch(x){
        case ATSAT_RX_P_DYNN: return &m.p_dynn;
        case ATSAT_RX_J_H2S: return &m.j_h2s; case ATSAT_RX_J_NH3: return &m.j_nh3; case ATSAT_RX_J_NH4SH: return &m.j_nh4sh;
        case ATSAT_RX_W_H2S: return &m.w_h2s; case ATSAT_RX_W_NH3: return &m.w_nh3; case ATSAT_RX_W_NH4SH: return &m.w_nh4sh;
        case ATSAT_RX_BUOYANCY_FORCE: return &m.BuoyancyForce; case ATSAT_RX_PRES_GRAD_FORCE: return &m.PresGradForce;
        case ATSAT_RX_Q_SENSIBLE: return &m.Q_Sensible;
        case ATSAT_RX_JT_H2S: return &m.jT_h2s; case ATSAT_RX_JT_NH3: return &m.jT_nh3; case ATSAT_RX_JT_NH4SH: return &m.jT_nh4sh;
        default: return nullptr;
        }
    }
    static int copy_extra(cSaturnModel& m, int x, double* a, int set){
        Array* A = extra(m, x);
        if(!A) return -1;
        for(int i = 0; i < m.im; i++) for(int j = 0; j < m.jm; j++){
            double* d = a + ((size_t)i * m.jm + j) * m.km;
            if(set) std::memcpy(A->x[i][j], d, sizeof(double) * m.km);
            else std::memcpy(d, A->x[i][j], sizeof(double) * m.km);
        }
        return 0;
    }
    // save_state / load_state write and read output_path/sat_restart_<iter>.bin
    static void save(cSaturnModel& m, const char* dir, int iter){ m.output_path = dir; m.save_state(iter); }
    static int restore(cSaturnModel& m, const char* dir, int iter){ m.output_path = dir; return m.load_state(iter) ? 1 : 0; }
    // the gases and tracers the loop hands the moist convective adjustment, cSaturnModel.cpp:1121-1139
    static void mca(cSaturnModel& m){
        typedef MoistConvectiveAdjustmentSat::Gas Gas_;
        typedef MoistConvectiveAdjustmentSat::Tracer Tr_;
        std::vector<Gas_> gases = {
            { "H2O", &m.h2o, &m.h2o_cloud, &m.h2o_ice, m.ep_h2o, m.lv_h2o, m.ls_h2o, m.t_0_h2o,
              m.C_h2o, m.L0_h2o, m.R_h2o, m.del_alf_h2o, m.del_bet_h2o, m.C_h2o_ice, m.L0_h2o_ice, m.del_alf_h2o_ice, m.del_bet_h2o_ice, m.t_00_h2o },
            { "NH3", &m.nh3, &m.nh3_cloud, &m.nh3_ice, m.ep_nh3, m.lv_nh3, m.ls_nh3, m.t_0_nh3,
              m.C_nh3, m.L0_nh3, m.R_nh3, m.del_alf_nh3, m.del_bet_nh3, m.C_nh3_ice, m.L0_nh3_ice, m.del_alf_nh3_ice, m.del_bet_nh3_ice, m.t_00_nh3 },
            { "CH4", &m.ch4, &m.ch4_cloud, &m.ch4_ice, m.ep_ch4, m.lv_ch4, m.ls_ch4, m.t_0_ch4,
              m.C_ch4, m.L0_ch4, m.R_ch4, m.del_alf_ch4, m.del_bet_ch4, m.C_ch4_ice, m.L0_ch4_ice, m.del_alf_ch4_ice, m.del_bet_ch4_ice, m.t_00_ch4 } };
        const double L_nh4sh_ = ChemistrySat::nh4sh_eq() ? ChemistrySat::nh4sh_formation_heat() : 0.0;
        std::vector<Tr_> tracers = { { "H2S", &m.h2s, m.ep_h2s, false, 0.0 }, { "NH4SH", &m.nh4sh, 0.0, true, L_nh4sh_ } };
        MoistConvectiveAdjustmentSat(m, gases, tracers).run();
        // the MCA sync, cSaturnModel.cpp:1147-1171
        for(int i = 0; i < m.im; i++) for(int j = 0; j < m.jm; j++) for(int k = 0; k < m.km; k++){
            m.tn.x[i][j][k] = m.t.x[i][j][k];
            m.h2on.x[i][j][k] = m.h2o.x[i][j][k]; m.h2o_cloudn.x[i][j][k] = m.h2o_cloud.x[i][j][k];
            m.h2o_icen.x[i][j][k] = m.h2o_ice.x[i][j][k];
            m.nh3n.x[i][j][k] = m.nh3.x[i][j][k]; m.nh3_cloudn.x[i][j][k] = m.nh3_cloud.x[i][j][k];
            m.nh3_icen.x[i][j][k] = m.nh3_ice.x[i][j][k];
            m.ch4n.x[i][j][k] = m.ch4.x[i][j][k]; m.ch4_cloudn.x[i][j][k] = m.ch4_cloud.x[i][j][k];
            m.ch4_icen.x[i][j][k] = m.ch4_ice.x[i][j][k];
            m.h2sn.x[i][j][k] = m.h2s.x[i][j][k]; m.nh4shn.x[i][j][k] = m.nh4sh.x[i][j][k];
            m.chromn.x[i][j][k] = m.chrom.x[i][j][k];
        }
    }
    static int call(cSaturnModel& m, int what, int arg){
        switch(what){
        case 0: m.computeMixtureDensity(); break;
        case 1: PressureSolverSat(m).run(); break;
        case 2: AtomUtils::damp_wiggles(m.p_dyn, nullptr, true, false, true); break;
        case 3: satadj(m, arg); break;
        case 4: m.rebuildPstatHydro("loop"); break;
        case 5: PrecipitationSat(m).run(); break;
        case 6: rk_sync(m); break;
        case 7: m.RungeKuttaSat(); break;
        case 8: BC_Sat(m).bcTheta(); BC_Sat(m).bcPhi(); BC_Sat(m).bcRadius(); break;
        case 9: m.clampNegativeSpecies(); break;
        case 10: m.restoreVar(1.0); break;
        case 11:                                   // one iteration, cSaturnModel.cpp:864-1111
            if(arg % 2 == 0){
                call(m, 0, 0); call(m, 1, 0); call(m, 2, 0);
                call(m, 3, ATSAT_GAS_CH4); call(m, 3, ATSAT_GAS_H2O); call(m, 3, ATSAT_GAS_NH3);
                call(m, 4, 0);
                if(chem_block) call(m, 17, 0);                                // :933
                call(m, 5, 0); call(m, 6, 0);
                if(chem_block) call(m, 18, 0);                                // :979-994
            }
            call(m, 7, 0); call(m, 8, 0); call(m, 9, 0); call(m, 10, 0);
            if(m.turb_active && turb_closure) call(m, 15, 0);          // cSaturnModel.cpp:1116
            if(mca_stride > 0 && arg % mca_stride == 0) call(m, 19, 0);  // :1120-1171
            break;
        case 12: m.computeBuoyancyRefLevel(); break;
        case 13: m.computeMetricPressure(); break;
        case 14: m.prepareProjectionBoundaries(true); break;
        case 15: TurbulenceSat(m).run(); break;
        case 16: TurbulenceSat(m).init(); break;
        case 17: ChemistrySat(m).EquilibriumNH4SH("loop"); break;
        case 19: mca(m); break;
        case 18:                                                           // cSaturnModel.cpp:979-994
            ChemistrySat(m).DiffMassFluxSat();
            AtomUtils::damp_wiggles(m.difflux_h2s, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.difflux_nh3, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.difflux_nh4sh, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.thermalmassflux, nullptr, true, true, true);
            ChemistrySat(m).ChemMassRateSat();
            ChemistrySat(m).FluxLimiterNH4SH();
            AtomUtils::damp_wiggles(m.massflux_h2s, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.massflux_nh3, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.massflux_nh4sh, nullptr, true, true, true);
            AtomUtils::damp_wiggles(m.fluxlim_nh4sh, nullptr, true, true, true);
            break;
        default: return -1;
        }
        return 0;
    }
};

int ProcessDiag<AtsatRefTag>::turb_closure = 0;
int ProcessDiag<AtsatRefTag>::chem_block = 0;
int ProcessDiag<AtsatRefTag>::mca_stride = 0;
typedef ProcessDiag<AtsatRefTag> H;

namespace {
// The reference prints a log line per call; the harness silences stdout while it runs.
struct Quiet {
    int saved = -1;
    Quiet(){ if(!getenv("ATSAT_REF_VERBOSE")){ fflush(stdout); saved = dup(1); int nul = open("/dev/null", O_WRONLY); dup2(nul, 1); close(nul); } }
    ~Quiet(){ if(saved >= 0){ fflush(stdout); std::cout.flush(); dup2(saved, 1); close(saved); } }
};
// The closure's knobs are read once per process (function-local statics, Turbulence.h:123-126,
// 156-159, 836-838), so the first call's values hold for the process.
void set_knob_env(const atsat_knobs& k){
    // Run sets PRESS_INCR when GEO_BALANCE is on (cSaturnModel.cpp:612); the sweep count and
    // the source cap are read on every solver call (PressureSolver.h:168-175).
    setenv("ATSAT_PRESS_INCR", "1", 0);
    char buf[64];
    snprintf(buf, sizeof buf, "%d", k.press_sweeps); setenv("ATSAT_PRESS_SWEEPS", buf, 1);
    snprintf(buf, sizeof buf, "%.17g", k.pdyn_cap); setenv("ATSAT_PDYN_CAP", buf, 1);
    snprintf(buf, sizeof buf, "%d", k.turb_buoy); setenv("ATSAT_TURB_BUOY", buf, 1);
    snprintf(buf, sizeof buf, "%.17g", k.turb_k_bg); setenv("ATSAT_TURB_K_BG", buf, 1);
    snprintf(buf, sizeof buf, "%.17g", k.turb_nue_max); setenv("ATSAT_NUE_MAX", buf, 1);
    snprintf(buf, sizeof buf, "%d", k.press_line); setenv("ATSAT_PRESS_LINE", buf, 1);
    snprintf(buf, sizeof buf, "%d", k.press_baro); setenv("ATSAT_PRESS_BAROTROPIC", buf, 1);
}
}

extern "C" {
void* ref_create(const atsat_knobs* k){
    if(!k) return nullptr;
    set_knob_env(*k);
    Quiet q;
    return H::create(*k);
}
int ref_destroy(void* h){ Quiet q; delete static_cast<cSaturnModel*>(h); return 0; }
int ref_grid(void* h, int* dims){
    cSaturnModel* m = static_cast<cSaturnModel*>(h);
    (void)m; dims[0] = 41; dims[1] = 181; dims[2] = 361; return 0;
}
int ref_load(void* h, const atsat_host_state* s){ return H::load(*static_cast<cSaturnModel*>(h), *s); }
int ref_store(void* h, atsat_host_state* s){ return H::store(*static_cast<cSaturnModel*>(h), *s); }
int ref_save_state(void* h, const char* dir, int iter){ Quiet q; H::save(*static_cast<cSaturnModel*>(h), dir, iter); return 0; }
int ref_load_state(void* h, const char* dir, int iter){ Quiet q; return H::restore(*static_cast<cSaturnModel*>(h), dir, iter); }
// copy one of the extras out of (set = 0) or into (set = 1) the model, dense [i][j][k]
int ref_extra(void* h, int x, double* a, int set){ return H::copy_extra(*static_cast<cSaturnModel*>(h), x, a, set); }
int ref_call(void* h, const atsat_knobs* k, int what, int arg){
    set_knob_env(*k);
    H::turb_closure = k->turb_closure;
    H::chem_block = k->chem_block;
    H::mca_stride = k->mca_stride;
    Quiet q;
    return H::call(*static_cast<cSaturnModel*>(h), what, arg);
}
}
