// Host-side set-up of the ATSAT step: the reference's default grid, constants and knobs, and
// the seeded synthetic start state gen_v1. No device code; compiled into libatsat_gpu.so.
#include "atsat_gpu.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

extern "C" int atsat_default_grid(atsat_grid* g){
    if(!g) return ATSAT_E_ARG;
    g->im = 41; g->jm = 181; g->km = 361;          // planet/cSaturnModel.h:711
    g->dr = 0.025;                                   // planet/cSaturnModel.cpp:50
    g->dthe_deg = 1.0; g->dphi_deg = 1.0;            // planet/cSaturnModel.cpp:42-43
    return ATSAT_OK;
}

extern "C" int atsat_default_knobs(atsat_knobs* k){
    if(!k) return ATSAT_E_ARG;
    std::memset(k, 0, sizeof(*k));
    k->dt = 0.01;                 // ATSAT_DT default, planet/cSaturnModel.cpp:171-174
    k->turb_active = 1;           // turb_model k_omega_SST (saturn/config_atsat.xml:19), cSaturnModel.cpp:407-408
    k->chemical_reaction = 1.0;   // saturn/config_atsat.xml:17
    k->turb_coupling = 1.0;       // planet/RHS_Sat_Turb.cpp:577-578
    k->precip_coupling = 1.0;     // planet/RHS_Sat_Turb.cpp:645-650
    k->pdyn_cap = 0.2;            // planet/PressureSolver.h:168
    k->press_sweeps = 1;          // planet/PressureSolver.h:174
    k->press_mode = 0;            // the reference's red-black relaxation
    k->press_tol = 1.0e-10;
    k->press_max_iter = 8;
    k->press_fp32_transforms = 1;
    k->precip_on = 1;             // planet/cSaturnModel.cpp:206-209
    k->chem_inputs = 0;           // implied by chem_block
    k->turb_closure = 1;          // cSaturnModel.cpp:1116 runs the closure whenever turb_active
    k->turb_buoy = 3;             // ATSAT_TURB_BUOY default, planet/cSaturnModel.h:299
    k->turb_k_bg = 1.0;           // ATSAT_TURB_K_BG [m2/s2], planet/Turbulence.h:156-159
    k->turb_nue_max = 1.0e5;      // ATSAT_NUE_MAX [m2/s], planet/Turbulence.h:123-126
    k->chem_block = 1;            // cSaturnModel.cpp:933, 979-994 with ATSAT_NH4SH_EQ 1
    k->mca_stride = 1;            // ATSAT_MCA default 1, planet/cSaturnModel.cpp:93-102
    k->press_line = 0;            // ATSAT_PRESS_LINE, PressureSolver.h:300
    k->press_baro = 0;            // ATSAT_PRESS_BAROTROPIC, PressureSolver.h:302
    return ATSAT_OK;
}

extern "C" int atsat_resident_knobs(atsat_knobs* k){
    const int rc = atsat_default_knobs(k);
    if(rc != ATSAT_OK) return rc;
    k->turb_active = 0;           // ATSAT_TURB=0 (cSaturnModel.cpp:250-253)
    k->chemical_reaction = 0.0;   // chemical_reaction 0 in the XML: massflux_* do not enter
    k->chem_block = 0;
    k->mca_stride = 0;            // ATSAT_MCA=0
    return ATSAT_OK;
}

extern "C" int atsat_default_consts(atsat_consts* c){
    if(!c) return ATSAT_E_ARG;
    std::memset(c, 0, sizeof(*c));
    // saturn/config_atsat.xml as shipped
    c->re = 100000.0; c->pr = 69.0; c->g = 10.0; c->omega = 0.000163; c->gam = 0.7;
    c->u_0 = 470.0; c->t_ref = 134.0; c->p_ref = 1.0; c->R_ref = 3.89; c->L_atm = 500.0;
    c->Coriolis = 1.0; c->centrifugal = 1.0; c->buoyancy = 1.0;
    c->metric_radius_km = 58232.0;                   // planet/cSaturnModel.h:164-169
    // Species constants, planet/cSaturnModel.h:815-931, and the mixture values formed as
    // ThermalPropertiesSat forms them (planet/ChemistrySat.h:484-531), same operand order.
    const double r_h2 = 0.864, r_he = 0.136, r_ch4 = 0.19, r_nh3 = 0.04, r_h2s = 0.055,
                 r_h2o = 0.09, r_nh4sh = 0.007;
    const double cp_h2 = 14.32e3, cp_he = 5.19e3, cp_ch4 = 2.232e3, cp_nh3 = 2.19e3,
                 cp_nh4sh = 2.00e3, cp_h2s = 2.24e3, cp_h2o = 1.93e3;
    const double R_h2 = 4124.2, R_he = 2077.1, R_ch4 = 518.28, R_nh3 = 488.21,
                 R_nh4sh = 261.0, R_h2s = 243.96, R_h2o = 461.52;
    c->r_mix = r_h2 + r_he + r_h2s + r_nh3 + r_nh4sh + r_h2o + r_ch4;
    c->cp_mix = (r_h2 * cp_h2 + r_h2s * cp_h2s + r_he * cp_he + r_nh3 * cp_nh3
               + r_nh4sh * cp_nh4sh + r_h2o * cp_h2o + r_ch4 * cp_ch4) / c->r_mix;
    c->R_mix = (r_h2 * R_h2 + r_he * R_he + r_h2o * R_h2o + r_h2s * R_h2s + r_nh3 * R_nh3
              + r_nh4sh * R_nh4sh + r_ch4 * R_ch4)
             / (r_h2 + r_he + r_h2o + r_h2s + r_nh3 + r_nh4sh + r_ch4);
    c->sc_ch4 = 0.99; c->sc_h2o = 0.61; c->sc_h2s = 0.94; c->sc_nh3 = 0.61; c->sc_nh4sh = 0.7;
    // Per gas: planet/cSaturnModel.h:722-746 (t_0, t_00), 934-964 (ep, lv, ls), 989-1050
    // (Sanchez-Lavega sets); the ice quadruple per gas as SaturationAdjustmentSat.cpp:37-44.
    const int CH4 = ATSAT_GAS_CH4, H2O = ATSAT_GAS_H2O, NH3 = ATSAT_GAS_NH3;
    c->t_0[CH4] = 90.69;  c->t_00[CH4] = 67.36;
    c->t_0[H2O] = 273.15; c->t_00[H2O] = 210.15;
    c->t_0[NH3] = 195.5;  c->t_00[NH3] = 140.0;
    c->ep[CH4] = 7.6752; c->ep[H2O] = 8.1253; c->ep[NH3] = 7.6752;
    c->lv[CH4] = 5.11e5; c->lv[H2O] = 2.5009e6; c->lv[NH3] = 1.3720e6;
    c->ls[CH4] = 5.11e5; c->ls[H2O] = 2.8339e6; c->ls[NH3] = 1.8320e6;
    c->C[CH4] = 1.627;  c->L0[CH4] = 553.1;  c->R[CH4] = R_ch4; c->del_alf[CH4] = 1.002;  c->del_bet[CH4] = -4.1e-3;
    c->C[H2O] = 25.096; c->L0[H2O] = 3148.2; c->R[H2O] = R_h2o; c->del_alf[H2O] = 0.0;    c->del_bet[H2O] = -8.7e-3;
    c->C[NH3] = 27.863; c->L0[NH3] = 2016.0; c->R[NH3] = R_nh3; c->del_alf[NH3] = -0.888; c->del_bet[NH3] = 0.0;
    c->C_i[CH4] = 1.627;  c->L0_i[CH4] = 553.1;  c->del_alf_i[CH4] = 1.002;  c->del_bet_i[CH4] = -4.1e-3;
    c->C_i[H2O] = 28.418; c->L0_i[H2O] = 3567.3; c->del_alf_i[H2O] = 0.0;    c->del_bet_i[H2O] = -8.7e-3;
    c->C_i[NH3] = 34.948; c->L0_i[NH3] = 2692.5; c->del_alf_i[NH3] = -0.888; c->del_bet_i[NH3] = 0.0;
    c->t_0_nh4sh = 230.0; c->t_00_nh4sh = 140.0;     // planet/cSaturnModel.h:747-748
    // The chemistry block: planet/cSaturnModel.h:818-931; c_mix as ThermalPropertiesSat forms it
    // (planet/ChemistrySat.h:486) from the member initialisers c_x = r_x / m_x (:866-872).
    const double m_h2 = 2.016, m_he = 4.0026, m_ch4 = 16.042, m_h2o = 18.01588;
    c->m_nh3 = 17.03052; c->m_h2s = 34.08088; c->m_nh4sh = 51.1114;
    c->c_mix = r_h2 / m_h2 + r_he / m_he + r_h2s / c->m_h2s + r_nh3 / c->m_nh3 + r_nh4sh / c->m_nh4sh
             + r_h2o / m_h2o + r_ch4 / m_ch4;
    c->mue_nh3 = 0.92e-5; c->mue_nh4sh = 0.99e-5; c->mue_h2s = 1.3e-5;
    c->rg_nh3 = 0.7623; c->rg_nh4sh = 1170.0; c->rg_h2s = 1.5357;
    c->cp_nh3 = cp_nh3; c->cp_nh4sh = cp_nh4sh; c->cp_h2s = cp_h2s;
    c->k_nh3 = 0.02102; c->k_nh4sh = 0.02102; c->k_h2s = 0.013;
    c->ep_h2s = 14.5192;                             // planet/cSaturnModel.h:937
    return ATSAT_OK;
}

// ============================================================================================
// gen_v1: the seeded synthetic start state.
// ============================================================================================
namespace {

// SplitMix64, keyed by (seed, member, purpose); a fresh stream per purpose so that adding a
// purpose never shifts another's draws.
struct SplitMix64 {
    uint64_t s;
    SplitMix64(uint64_t seed, uint64_t member, uint64_t purpose)
        : s(seed ^ (0x9E3779B97F4A7C15ull * (member + 1)) ^ (0xD1B54A32D192ED03ull * (purpose + 1))) {}
    uint64_t next(){
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform(){ return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }   // (0,1)
    // Box-Muller; both values of a pair are used in draw order.
    bool have = false; double spare = 0.0;
    double normal(){
        if(have){ have = false; return spare; }
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * M_PI * u2;
        spare = r * std::sin(a); have = true;
        return r * std::cos(a);
    }
};

inline double svp(double T, double C, double L0, double R, double alf, double bet){
    return std::exp(C + (-L0 / T + alf * std::log(T) + bet * T) / (1e-3 * R));
}

}  // namespace

extern "C" int atsat_gen_v1(const atsat_grid* gp, const atsat_consts* cp, unsigned long long seed,
                            int member, double t_sigma, double w_sigma, atsat_host_state* out){
    if(!gp || !cp || !out) return ATSAT_E_ARG;
    const atsat_grid& G = *gp; const atsat_consts& C = *cp;
    const int im = G.im, jm = G.jm, km = G.km;
    if(im < 5 || jm < 7 || km < 6) return ATSAT_E_ARG;
    const size_t N = (size_t)im * jm * km;
    auto at = [&](int i, int j, int k){ return ((size_t)i * jm + j) * km + k; };
    std::vector<double> t(N), w(N), ps(N), rho(N);
    const double pi180 = 180.0 / M_PI;
    const double dthe = G.dthe_deg / pi180, dphi = G.dphi_deg / pi180;
    (void)dphi;
    const double h_km = C.L_atm * G.dr;                 // layer thickness [km]
    const int jeq = (jm - 1) / 2;
    // Temperature: T0(j) = t_pole + (t_pole - t_equator)(j^2/jeq^2 - 2 j/jeq), the parabola of
    // init_temperature (InitVariables_Sat.cpp:28-37) with t_pole 380 K, t_equator 400 K;
    // minus the gam lapse; plus N(0, 0.1 K) white noise and the member's N(0, t_sigma).
    {
        SplitMix64 rng(seed, (uint64_t)member, 1);
        SplitMix64 rng_m(seed, (uint64_t)member, 2);
        const double t_pole = 380.0, t_eq = 400.0;
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++) for(int k = 0; k < km - 1; k++){
            const double dj = (double)j, dh = (double)jeq;
            const double T0 = (t_pole - t_eq) * (dj * dj / (dh * dh) - 2.0 * dj / dh) + t_pole;
            double T = T0 - C.gam * h_km * i + 0.1 * rng.normal();
            if(t_sigma > 0.0) T += t_sigma * rng_m.normal();
            t[at(i, j, k)] = T / C.t_ref;
        }
    }
    // Zonal wind w: 470 cos^8(lat) + sum_n (-1)^n 50 exp(-((|lat| - 20 n)/5)^2) m/s, over u_0.
    {
        SplitMix64 rng_m(seed, (uint64_t)member, 3);
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
            const double lat = 90.0 - j * G.dthe_deg;
            const double cl = std::cos(lat / pi180);
            double wj = 470.0 * std::pow(cl, 8.0);
            for(int n = 1; n <= 4; n++){
                const double x = (std::fabs(lat) - 20.0 * n) / 5.0;
                wj += ((n & 1) ? -50.0 : 50.0) * std::exp(-x * x);
            }
            for(int k = 0; k < km - 1; k++){
                double ww = wj;
                if(w_sigma > 0.0) ww += w_sigma * rng_m.normal();
                w[at(i, j, k)] = ww / C.u_0;
            }
        }
    }
    for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
        t[at(i, j, km - 1)] = t[at(i, j, 0)];
        w[at(i, j, km - 1)] = w[at(i, j, 0)];
    }
    // p_stat integrated up each column as rebuildPstatHydro does (InitVariables_Sat.cpp:556-578)
    const double exp_pressure = C.g / (C.gam * C.R_ref);
    for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){
        double p = C.p_ref * std::pow(t[at(0, j, k)], exp_pressure);
        ps[at(0, j, k)] = p;
        const double a = exp_pressure * C.gam * (double)(float)((float)h_km);
        for(int i = 1; i < im; i++){
            const double T0 = t[at(i - 1, j, k)] * C.t_ref, T1 = t[at(i, j, k)] * C.t_ref;
            if(T0 > 0.0 && T1 > 0.0){
                const double dT = T0 - T1;
                p *= (std::fabs(dT) > 1.0e-9 * T0) ? std::pow(T1 / T0, a / dT) : std::exp(-a / T0);
            } else p = C.p_ref * std::pow(std::max(0.0, t[at(i, j, k)]), exp_pressure);
            ps[at(i, j, k)] = p;
        }
    }
    // density, as computeMixtureDensity (Pressure_Sat.cpp:49-62)
    for(size_t n = 0; n < N; n++){
        const double T = t[n] * C.t_ref;
        if(!(T > 0.0) || !(C.R_mix > 0.0)){ rho[n] = 0.0; continue; }
        const double r = (ps[n] * 1.0e5) / (C.R_mix * T);
        rho[n] = std::isfinite(r) ? r : 0.0;
    }
    auto put = [&](int f, const std::vector<double>& src){
        if(out->f[f]) std::memcpy(out->f[f], src.data(), N * sizeof(double));
    };
    auto fill = [&](int f, double v){
        if(out->f[f]) std::fill(out->f[f], out->f[f] + N, v);
    };
    for(int f = 0; f < ATSAT_NFIELDS; f++) fill(f, 0.0);
    put(ATSAT_T, t); put(ATSAT_W, w); put(ATSAT_P_STAT, ps); put(ATSAT_RHO_MIX, rho);
    fill(ATSAT_TKE, 1.0e-4); fill(ATSAT_DIS, 1.0e-3);
    // Vapour of H2O and NH3: 0.9 of liquid saturation (rho * ep * E / p) at and below the cloud
    // base -- the highest level with T >= t_0 of the gas -- and that base value decaying with a
    // 25 km scale height above it. Capped by the deep abundance, the reference's vapour density
    // r_gas (planet/cSaturnModel.h:853-854) scaled by rho / r_mix, so that the deep, hot cells
    // where E exceeds p do not hold more vapour than the planet has.
    const int gases[2] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3 };
    const int gfield[2] = { ATSAT_H2O, ATSAT_NH3 };
    const double r_gas[2] = { 0.09, 0.04 };
    for(int q = 0; q < 2; q++){
        const int gs = gases[q];
        if(!out->f[gfield[q]]) continue;
        double* dst = out->f[gfield[q]];
        for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){
            int ib = -1;
            for(int i = 0; i < im; i++) if(t[at(i, j, k)] * C.t_ref >= C.t_0[gs]) ib = i;
            auto qsat = [&](int i){
                const double T = t[at(i, j, k)] * C.t_ref, p = ps[at(i, j, k)];
                const double r = rho[at(i, j, k)] > 0.0 ? rho[at(i, j, k)] : C.r_mix;
                return r * C.ep[gs] * svp(T, C.C[gs], C.L0[gs], C.R[gs], C.del_alf[gs], C.del_bet[gs]) / p;
            };
            auto cap = [&](int i){
                const double r = rho[at(i, j, k)] > 0.0 ? rho[at(i, j, k)] : C.r_mix;
                return r_gas[q] * r / C.r_mix;
            };
            const double qb = (ib >= 0) ? std::min(0.9 * qsat(ib), cap(ib)) : 0.0;
            for(int i = 0; i < im; i++){
                double qv = (i <= ib) ? std::min(0.9 * qsat(i), cap(i)) : qb * std::exp(-(i - ib) * h_km / 25.0);
                if(!std::isfinite(qv) || qv < 0.0) qv = 0.0;
                dst[at(i, j, k)] = qv;
            }
        }
    }
    // initGeoBalance (InitVariables_Sat.cpp:345-453): p_dyn = the barotropic balance of the
    // jet-level zonal-mean wind, p_baro = what each level's own wind needs on top of it, minus
    // the metric pressure of the start state; v = u = 0 here so cell_vbar0 = cell_ubar0 = 0.
    {
        const double nd_cor = C.L_atm * 1.0e3 / C.u_0;
        const double shift = C.metric_radius_km / C.L_atm - 1.0;
        std::vector<double> thz(jm), radz(im);
        thz[0] = 0.0; for(int j = 1; j < jm; j++) thz[j] = thz[j - 1] + dthe;
        radz[0] = 1.0; for(int i = 1; i < im; i++) radz[i] = radz[i - 1] + G.dr;
        std::vector<double> wbar((size_t)im * jm);
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
            double s = 0.0;
            for(int k = 0; k < km - 1; k++) s += w[at(i, j, k)];
            wbar[(size_t)i * jm + j] = s / (km - 1);
        }
        auto Gf = [&](int i, int j)->double{
            if(j == 0 || j == jm - 1) return 0.0;
            const double sth = std::sin(thz[j]);
            if(!(sth > 0.0)) return 0.0;
            const double cth = std::cos(thz[j]);
            const double wb = wbar[(size_t)i * jm + j];
            return (radz[i] + shift) * C.Coriolis * nd_cor * 2.0 * C.omega * cth * wb + wb * wb * cth / sth;
        };
        auto integrate = [&](const std::vector<double>& g, std::vector<double>& P){
            P[0] = 0.0;
            for(int j = 1; j < jm; j++)
                P[j] = (j >= 2) ? P[j - 2] + 2.0 * dthe * g[j - 1] : P[j - 1] + 0.5 * (g[j - 1] + g[j]) * dthe;
        };
        std::vector<double> g(jm), P(jm);
        for(int j = 0; j < jm; j++){
            int a = 0;
            for(int i = 1; i < im; i++) if(std::fabs(wbar[(size_t)i * jm + j]) > std::fabs(wbar[(size_t)a * jm + j])) a = i;
            g[j] = Gf(a, j);
        }
        integrate(g, P);
        double sw = 0.0, sp = 0.0;
        for(int j = 0; j < jm; j++){ const double s_ = std::fabs(std::sin(thz[j])); sw += s_; sp += s_ * P[j]; }
        const double p_off = sp / sw;
        for(int j = 0; j < jm; j++) P[j] -= p_off;
        std::vector<double> pdyn(N), pmet(N), pbaro(N);
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++) pdyn[at(i, j, k)] = P[j];
        // metric pressure of the start state (RungeKutta_Sat_Turb.cpp:60-108, p_baro empty, v = 0)
        for(int j = 0; j < jm; j++){
            const double cor = C.Coriolis * nd_cor * 2.0 * C.omega * std::fabs(std::sin(thz[j]));
            for(int k = 0; k < km; k++){
                auto f_ = [&](int i){
                    const double rm = radz[i] + shift, ww = w[at(i, j, k)];
                    return (ww * ww) / rm + cor * ww;
                };
                pmet[at(0, j, k)] = 0.0;
                for(int i = 1; i < im; i++) pmet[at(i, j, k)] = pmet[at(i - 1, j, k)] + 0.5 * (f_(i - 1) + f_(i)) * G.dr;
            }
        }
        std::vector<double> gi(jm), Q(jm);
        for(int i = 0; i < im; i++){
            for(int j = 0; j < jm; j++) gi[j] = Gf(i, j);
            integrate(gi, Q);
            for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++)
                pbaro[at(i, j, k)] = (Q[j] - p_off) - pdyn[at(i, j, k)] - pmet[at(i, j, k)];
        }
        put(ATSAT_P_DYN, pdyn); put(ATSAT_P_BARO, pbaro);
        std::vector<double> pm_full(N);
        for(size_t n = 0; n < N; n++) pm_full[n] = pmet[n] + pbaro[n];
        put(ATSAT_P_METRIC, pm_full);
    }
    // start-of-step copies = the live fields (restoreVar(1.0) at the end of the set-up)
    for(int f = 0; f < ATSAT_NPROG; f++)
        if(out->f[ATSAT_F_N0 + f] && out->f[f]) std::memcpy(out->f[ATSAT_F_N0 + f], out->f[f], N * sizeof(double));
    for(int vv = 0; vv < ATSAT_NVECS; vv++){
        if(!out->v[vv]) continue;
        size_t n = 0;
        if(vv == ATSAT_V_CELL_VBAR0 || vv == ATSAT_V_CELL_UBAR0) n = (size_t)im * jm;
        else if(vv == ATSAT_V_BUOY_REF) n = (size_t)im;
        else n = (size_t)jm * km;
        std::fill(out->v[vv], out->v[vv] + n, 0.0);
    }
    return ATSAT_OK;
}
