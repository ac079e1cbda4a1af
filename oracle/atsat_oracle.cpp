// ============================================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 CPU restatement (C++/OpenMP) of the reference's model time step, written from
// the reference's formulas, grid-generic, on dense [i][j][k] arrays. It is the checker of the
// GPU library and the CPU baseline of bench.py; nothing in the product links, loads or calls
// it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every function cites the reference file:line it restates. Operand order follows the
// reference expression by expression so that, compiled with -ffp-contract=off, the result is
// bit-identical to the reference build (checked against oracle/_ref where that builds).
// ============================================================================================
#include "atsat_fields.h"
#include "atsat_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Geo {
    int im, jm, km;
    double dr, dthe, dphi;
    std::vector<double> radz, thez;     // Array_1D::Coordinates, repeated addition
    std::vector<float>  hgt;            // m_layer_heights, float (cSaturnModel.h:1061-1073)
};

Geo make_geo(const atsat_grid& g, const atsat_consts& c){
    Geo G;
    G.im = g.im; G.jm = g.jm; G.km = g.km;
    const double pi180 = 180.0 / M_PI;              // cSaturnModel.cpp:40
    G.dr = g.dr;
    G.dthe = g.dthe_deg / pi180;                    // cSaturnModel.cpp:51-52
    G.dphi = g.dphi_deg / pi180;
    G.radz.resize(g.im); G.thez.resize(g.jm);
    G.radz[0] = 1.0;                                // r0, cSaturnModel.cpp:48
    for(int l = 1; l < g.im; l++) G.radz[l] = G.radz[l - 1] + G.dr;   // lib/Array_1D.cpp:54-59
    G.thez[0] = 0.0;
    for(int l = 1; l < g.jm; l++) G.thez[l] = G.thez[l - 1] + G.dthe;
    const float h = c.L_atm * G.dr;                 // init_layer_heights, cSaturnModel.h:1069
    G.hgt.resize(g.im);
    for(int i = 0; i < g.im; i++) G.hgt[i] = i * h;
    return G;
}

struct Ctx {
    Geo G;
    const atsat_consts& C;
    const atsat_knobs& K;
    atsat_host_state& S;
    Ctx(const atsat_grid& g, const atsat_consts& c, const atsat_knobs& k, atsat_host_state& s)
        : G(make_geo(g, c)), C(c), K(k), S(s) {}
    size_t at(int i, int j, int k) const { return ((size_t)i * G.jm + j) * G.km + k; }
    double* f(int id) const { return S.f[id]; }
    double metricRadius(double rm) const {            // cSaturnModel.h:164-169
        const double R_km = C.metric_radius_km;
        if(!(R_km > 0.0)) return rm;
        return rm + (R_km / C.L_atm - 1.0);
    }
    // ATPhys::phi_wrap with the periodic seam, ATPhys.h:135-139
    int wrap(int kk) const { const int n = G.km - 1; return kk < 0 ? kk + n : (kk >= n ? kk - n : kk); }
    // rho_at, cSaturnModel.h:514-518 (ATSAT_LOCAL_RHO default 1)
    double rho_at(size_t n) const {
        const double rho = S.f[ATSAT_RHO_MIX][n];
        return (rho > 0.0 && std::isfinite(rho)) ? rho : C.r_mix;
    }
    double layer_thickness_m(int i) const {            // cSaturnModel.h:689-692
        if(i < 0 || i > G.im - 2) return 0.0;
        return (double)(G.hgt[i + 1] - G.hgt[i]) * 1.0e3;
    }
};

// ATPhys::saturation_vapour_pressure, ATPhys.h:49-52
inline double svp(double T_K, double C, double L0, double R, double del_alf, double del_bet){
    return std::exp(C + (-L0 / T_K + del_alf * std::log(T_K) + del_bet * T_K) / (1e-3 * R));
}

// ---------------------------------------------------------------------------------------------
// cSaturnModel::computeMixtureDensity, planet/Pressure_Sat.cpp:49-62
void mixture_density(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    const double* t = X.f(ATSAT_T); const double* ps = X.f(ATSAT_P_STAT); double* rho = X.f(ATSAT_RHO_MIX);
    #pragma omp parallel for schedule(static)
    for(size_t n = 0; n < N; n++){
        const double T = t[n] * X.C.t_ref;
        if(!(T > 0.0) || !(X.C.R_mix > 0.0)){ rho[n] = 0.0; continue; }
        const double r = (ps[n] * 1.0e5) / (X.C.R_mix * T);
        rho[n] = std::isfinite(r) ? r : 0.0;
    }
}

// cSaturnModel::computeBuoyancyRefLevel, planet/RungeKutta_Sat_Turb.cpp:31-50
void buoyancy_ref_level(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double* t = X.f(ATSAT_T); const double* ps = X.f(ATSAT_P_STAT);
    double* out = X.S.v[ATSAT_V_BUOY_REF];
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < G.im; i++){
        double sum = 0.0, wsum = 0.0;
        for(int j = 0; j < G.jm; j++){
            const double wgt = std::sin(G.thez[j]);
            for(int k = 0; k < G.km; k++){
                const size_t n = X.at(i, j, k);
                if(!(t[n] > 0.0)) continue;
                const double b = C.g * ps[n] / (C.r_mix * C.R_mix * t[n] * C.t_ref);
                if(!std::isfinite(b)) continue;
                sum += wgt * b;
                wsum += wgt;
            }
        }
        out[i] = (wsum > 0.0) ? sum / wsum : 0.0;
    }
}

// cSaturnModel::computeMetricPressure, planet/RungeKutta_Sat_Turb.cpp:60-108
// (METRIC_SPLIT = CORIOLIS_SPLIT = 1, HYDRO_BUOY = 0, ND_COR on with GEO_BALANCE)
void metric_pressure(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double* v = X.f(ATSAT_V); const double* w = X.f(ATSAT_W);
    const double* pb = X.f(ATSAT_P_BARO); double* pm = X.f(ATSAT_P_METRIC);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < G.jm; j++){
        for(int k = 0; k < G.km; k++){
            const double cor_ = C.Coriolis * (C.L_atm * 1.0e3 / C.u_0) * 2.0 * C.omega * std::abs(std::sin(G.thez[j]));
            auto f_ = [&](int i)->double{
                const double rm_ = X.metricRadius(G.radz[i]);
                const double exp_rm_ = 1.0;
                const size_t n = X.at(i, j, k);
                return ((v[n] * v[n] + w[n] * w[n]) / rm_ + cor_ * w[n]) / exp_rm_;
            };
            pm[X.at(0, j, k)] = 0.0;
            for(int i = 1; i < G.im; i++)
                pm[X.at(i, j, k)] = pm[X.at(i - 1, j, k)] + 0.5 * (f_(i - 1) + f_(i)) * G.dr;
            if(pb) for(int i = 0; i < G.im; i++) pm[X.at(i, j, k)] += pb[X.at(i, j, k)];
        }
    }
}

// The per-(i,j) geometry RungeKuttaSat hands to RHSSat, planet/RungeKutta_Sat_Turb.cpp:129-143,
// 223-243 (cSaturnModel.h:1081-1089)
struct CellGeo {
    double rm, rm2, exp_rm, exp_2_rm, sinthe, sinthe2, costhe, cotanthe;
    double inv_rm, inv_rm2, inv_rmsinthe, inv_rm2sinthe, inv_rm2sinthe2, costhe_inv_rm2sinthe;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
};

CellGeo cell_geo(const Ctx& X, int i, int j){
    const Geo& G = X.G;
    CellGeo g;
    double s = std::sin(G.thez[j]);
    const double s_min = 0.0;                        // sinthe_min(), cSaturnModel.h:458-465
    if(s < s_min) s = s_min;
    g.rm = X.metricRadius(G.radz[i]);
    g.rm2 = g.rm * g.rm;
    g.exp_rm = 1.0; g.exp_2_rm = g.exp_rm * g.exp_rm;   // coord_stretching false
    g.sinthe = s; g.sinthe2 = g.sinthe * g.sinthe;
    g.costhe = std::cos(G.thez[j]);
    g.cotanthe = g.costhe / g.sinthe;
    g.inv_rm = 1.0 / g.rm;
    g.inv_rm2 = 1.0 / g.rm2;
    g.inv_rmsinthe = 1.0 / (g.rm * g.sinthe);
    g.inv_rm2sinthe = g.inv_rm2 / g.sinthe;
    g.inv_rm2sinthe2 = g.inv_rm2 / g.sinthe2;
    g.costhe_inv_rm2sinthe = g.costhe * g.inv_rm2sinthe;
    g.inv_2dr = 1.0 / (2.0 * G.dr);
    g.inv_2dthe = 1.0 / (2.0 * G.dthe);
    g.inv_2dphi = 1.0 / (2.0 * G.dphi);
    g.inv_dr2 = 1.0 / (G.dr * G.dr);
    g.inv_dthe2 = 1.0 / (G.dthe * G.dthe);
    g.inv_dphi2 = 1.0 / (G.dphi * G.dphi);
    return g;
}

// cSaturnModel::RHSSat, planet/RHS_Sat_Turb.cpp:25-1016, with the default knobs: METRIC_SPLIT,
// CORIOLIS_SPLIT, ND_COR (GEO_BALANCE), CELL_HOLD, ADIABATIC(_RHO), PRECIP_COUPLING 1,
// RAD_COUPLING 0, THERMAL_MASSFLUX 0, SPECIES_DIVU 0, BUOY_LEGACY 0, CHROM 0. Writes rhs[17]
// and the three aux_* of the cell.
void rhs_cell(const Ctx& X, const CellGeo& geo, int i, int j, int k, double* rhs, double* aux3){
    const Geo& G = X.G; const atsat_consts& C = X.C; const atsat_knobs& K = X.K;
    const double rm = geo.rm;
    const double costhe = geo.costhe, cotanthe = geo.cotanthe;
    const double inv_rm = geo.inv_rm, inv_rm2 = geo.inv_rm2, inv_rmsinthe = geo.inv_rmsinthe;
    const double inv_rm2sinthe = geo.inv_rm2sinthe, inv_rm2sinthe2 = geo.inv_rm2sinthe2;
    const double costhe_inv_rm2sinthe = geo.costhe_inv_rm2sinthe;
    const double inv_2dr = geo.inv_2dr, inv_2dthe = geo.inv_2dthe, inv_2dphi = geo.inv_2dphi;
    const double inv_dr2 = geo.inv_dr2, inv_dthe2 = geo.inv_dthe2, inv_dphi2 = geo.inv_dphi2;
    const double exp_rm = geo.exp_rm, exp_2_rm = geo.exp_2_rm;
    const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
    const size_t c0 = X.at(i, j, k);
    const size_t rp = X.at(i + 1, j, k), rmn = X.at(i - 1, j, k);
    const size_t tp = X.at(i, j + 1, k), tmn = X.at(i, j - 1, k);
    const size_t pe = X.at(i, j, ke), pw = X.at(i, j, kw);

    double d1r[ATSAT_NPROG], d2r[ATSAT_NPROG], d1t[ATSAT_NPROG], d2t[ATSAT_NPROG], d1p[ATSAT_NPROG], d2p[ATSAT_NPROG];
    for(int f = 0; f < ATSAT_NPROG; f++){          // COMPUTE_DR / DTHE / DPHI, :133-207
        const double* F = X.f(f);
        d1r[f] = (F[rp] - F[rmn]) * inv_2dr * exp_rm;
        d2r[f] = (F[rp] - 2.0 * F[c0] + F[rmn]) * inv_dr2 * exp_2_rm;
        d1t[f] = (F[tp] - F[tmn]) * inv_2dthe;
        d2t[f] = (F[tp] - 2.0 * F[c0] + F[tmn]) * inv_dthe2;
        d1p[f] = (F[pe] - F[pw]) * inv_2dphi;
        d2p[f] = (F[pe] - 2.0 * F[c0] + F[pw]) * inv_dphi2;
    }
    const double* pd = X.f(ATSAT_P_DYN);
    const double dpdr = (pd[rp] - pd[rmn]) * inv_2dr * exp_rm;
    const double dpdthe = (pd[tp] - pd[tmn]) * inv_2dthe;
    const double dpdphi = (pd[pe] - pd[pw]) * inv_2dphi;

    const double u_ijk = X.f(ATSAT_U)[c0], v_ijk = X.f(ATSAT_V)[c0], w_ijk = X.f(ATSAT_W)[c0];
    const double t_ijk = X.f(ATSAT_T)[c0];

    const double L_m = C.L_atm * 1.0e3;                          // :271-277
    const double nd_cor = L_m / C.u_0;
    const double nd_cent = 1.0, nd_buoy = 1.0, buoy_scale = 1.0;

    const double sinthe_true = std::sqrt(std::max(0.0, 1.0 - costhe * costhe));   // :310-321
    const double Coriolis_rad = nd_cor * -2.0 * C.omega * sinthe_true * w_ijk;
    const double Coriolis_the = nd_cor * -2.0 * C.omega * costhe * w_ijk;
    double v_cor = v_ijk, u_cor = u_ijk;
    v_cor -= X.S.v[ATSAT_V_CELL_VBAR0][(size_t)i * G.jm + j];
    u_cor -= X.S.v[ATSAT_V_CELL_UBAR0][(size_t)i * G.jm + j];
    const double Coriolis_phi = nd_cor * +2.0 * C.omega * (+ costhe * v_cor + sinthe_true * u_cor);

    const double centrifugal_rad = nd_cent * C.omega * C.omega * rm * sinthe_true * sinthe_true;   // :349-350
    const double centrifugal_the = nd_cent * C.omega * C.omega * rm * sinthe_true * costhe;
    const double coeff_energy_p = C.u_0 * C.u_0 / (C.cp_mix * C.t_ref);    // :352

    double dpstatdr = 0.0;                                        // :368-378
    {
        const double* ps = X.f(ATSAT_P_STAT);
        const double to_nd = 1.0 / (C.r_mix * C.u_0 * C.u_0 * 1.0e-5);
        dpstatdr = (ps[rp] - ps[rmn]) * inv_2dr * exp_rm * to_nd;
        const double rho_l = X.f(ATSAT_RHO_MIX)[c0];
        if(rho_l > 0.0 && std::isfinite(rho_l)) dpstatdr *= C.r_mix / rho_l;
    }
    const double pressure_t = coeff_energy_p * (u_ijk * (dpdr + dpstatdr)        // :381-383
        + v_ijk * dpdthe * inv_rm + w_ijk * dpdphi * inv_rmsinthe);

    auto transport = [&](int f){                                  // :388-431
        return u_ijk * d1r[f] + v_ijk * d1t[f] * inv_rm + w_ijk * d1p[f] * inv_rmsinthe;
    };
    const double two_inv_rm = 2.0 * inv_rm;                       // :435-491
    const double v_metric = (1.0 + costhe / geo.sinthe2) * inv_rm2;
    auto diffusion = [&](int f){
        return d2r[f] + d1r[f] * two_inv_rm + d2t[f] * inv_rm2 + d1t[f] * costhe_inv_rm2sinthe + d2p[f] * inv_rm2sinthe2;
    };
    const int U = ATSAT_U, V = ATSAT_V, W = ATSAT_W;
    const double metric_on = 0.0, cor_rad_on = 0.0;
    const double transport_t = transport(ATSAT_T);
    const double transport_u = u_ijk * d1r[U] + v_ijk * d1t[U] * inv_rm + w_ijk * d1p[U] * inv_rmsinthe
        - metric_on * (v_ijk * v_ijk + w_ijk * w_ijk) * inv_rm;
    const double transport_v = u_ijk * d1r[V] + v_ijk * d1t[V] * inv_rm + w_ijk * d1p[V] * inv_rmsinthe
        + (u_ijk * v_ijk - w_ijk * w_ijk * cotanthe) * inv_rm;
    const double transport_w = u_ijk * d1r[W] + v_ijk * d1t[W] * inv_rm + w_ijk * d1p[W] * inv_rmsinthe
        + (w_ijk * u_ijk + v_ijk * w_ijk * cotanthe) * inv_rm;
    const double diffusion_t = diffusion(ATSAT_T);
    const double diffusion_u = d2r[U] + 2.0 * u_ijk * inv_rm2 + d2t[U] * inv_rm2
        + 4.0 * d1r[U] * inv_rm + d1t[U] * costhe_inv_rm2sinthe + d2p[U] * inv_rm2sinthe2;
    const double diffusion_v = d2r[V] + d1r[V] * two_inv_rm + d2t[V] * inv_rm2
        + d1t[V] * costhe_inv_rm2sinthe - v_metric * v_ijk + d2p[V] * inv_rm2sinthe2
        + 2.0 * d1t[U] * inv_rm2 - d1p[W] * 2.0 * costhe * inv_rm2sinthe2;
    const double diffusion_w = d2r[W] + d1r[W] * two_inv_rm + d2t[W] * inv_rm2
        + d1t[W] * costhe_inv_rm2sinthe - v_metric * w_ijk + d2p[W] * inv_rm2sinthe2
        + 2.0 * d1p[U] * inv_rm2sinthe + d1p[V] * 2.0 * costhe * inv_rm2sinthe2;

    if(K.turb_active){                                            // :511-538
        const double sig_k = 0.85, sig_w = 0.5;
        const double nue_here = X.f(ATSAT_NUE)[c0];
        rhs[ATSAT_TKE] = - transport(ATSAT_TKE) + diffusion(ATSAT_TKE) * (1.0 / C.re + nue_here / sig_k)
                       + X.f(ATSAT_TKE_SOURCE)[c0];
        rhs[ATSAT_DIS] = - transport(ATSAT_DIS) + diffusion(ATSAT_DIS) * (1.0 / C.re + nue_here / sig_w)
                       + X.f(ATSAT_DIS_SOURCE)[c0];
    } else {
        rhs[ATSAT_TKE] = 0.0; rhs[ATSAT_DIS] = 0.0;
    }
    const double Pr_t = 0.9;                                      // :577-582
    double nue_t = 0.0;
    if(K.turb_active){
        const double nu = X.f(ATSAT_NUE)[c0];
        nue_t = (K.turb_coupling != 0.0 && std::isfinite(nu)) ? K.turb_coupling * std::max(0.0, nu) : 0.0;
    }
    const double nue_t_s = nue_t / Pr_t;
    const double radiation_t = 0.0;                               // :611-629, RAD_COUPLING 0
    double precip_t = 0.0;                                        // :645-665
    if(K.precip_coupling != 0.0){
        const double L_rad = C.L_atm * 1.0e3;
        if(C.r_mix > 0.0 && C.cp_mix > 0.0){
            precip_t = K.precip_coupling * X.f(ATSAT_Q_PRECIP)[c0] * L_rad / (C.r_mix * C.cp_mix * C.u_0 * C.t_ref);
            const double precip_t_max = 0.5;
            if(!std::isfinite(precip_t)) precip_t = 0.0;
            else if(precip_t > precip_t_max) precip_t = precip_t_max;
            else if(precip_t < -precip_t_max) precip_t = -precip_t_max;
        }
    }
    double buoyancy_u = 0.0;                                      // :700-714
    if(t_ijk > 0.0){
        buoyancy_u = - nd_buoy * buoy_scale * C.buoyancy
            * (C.g * X.f(ATSAT_P_STAT)[c0] / (C.r_mix * C.R_mix * t_ijk * C.t_ref) - X.S.v[ATSAT_V_BUOY_REF][i]);
    }
    const double tmf = 0.0 * K.chemical_reaction * 0.0;          // :749-756, THERMAL_MASSFLUX 0
    rhs[ATSAT_T] = + pressure_t - transport_t + diffusion_t / (C.re * C.pr) + diffusion_t * nue_t_s
                   - tmf + radiation_t + precip_t;               // :752-758
    rhs[U] = - dpdr + buoyancy_u - transport_u + diffusion_u / C.re + diffusion_u * nue_t
             - cor_rad_on * C.Coriolis * Coriolis_rad + C.centrifugal * centrifugal_rad;   // :780-786
    const double* pmet = X.f(ATSAT_P_METRIC);                    // :814-818
    const double dpmdthe = (pmet[tp] - pmet[tmn]) * inv_2dthe;
    const double dpmdphi = (pmet[pe] - pmet[pw]) * inv_2dphi;
    rhs[V] = - dpdthe * inv_rm - dpmdthe * inv_rm - transport_v + diffusion_v / C.re + diffusion_v * nue_t
             - C.Coriolis * Coriolis_the + C.centrifugal * centrifugal_the;              // :820-826
    rhs[W] = - dpdphi * inv_rmsinthe - dpmdphi * inv_rmsinthe - transport_w + diffusion_w / C.re
             + diffusion_w * nue_t - C.Coriolis * Coriolis_phi;                          // :828-833
    auto species = [&](int f, double sc, double src){            // :839-890
        return - transport(f) + diffusion(f) / (sc * C.re) + diffusion(f) * nue_t_s + src;
    };
    const double pc = K.precip_coupling, cr = K.chemical_reaction;
    const bool chem = K.chem_inputs != 0 || K.chem_block != 0;
    auto Sp = [&](int id){ return X.f(id)[c0]; };
    rhs[ATSAT_CH4]       = species(ATSAT_CH4, C.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4));
    rhs[ATSAT_CH4_CLOUD] = species(ATSAT_CH4_CLOUD, C.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_CLOUD));
    rhs[ATSAT_CH4_ICE]   = species(ATSAT_CH4_ICE, C.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_ICE));
    rhs[ATSAT_H2O]       = species(ATSAT_H2O, C.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O));
    rhs[ATSAT_H2O_CLOUD] = species(ATSAT_H2O_CLOUD, C.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_CLOUD));
    rhs[ATSAT_H2O_ICE]   = species(ATSAT_H2O_ICE, C.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_ICE));
    rhs[ATSAT_H2S]       = species(ATSAT_H2S, C.sc_h2s, cr * (chem ? Sp(ATSAT_MASSFLUX_H2S) : 0.0));
    {   // rhs_nh3 adds the chemistry term before the precipitation sink, :875-880
        const int f = ATSAT_NH3;
        rhs[f] = - transport(f) + diffusion(f) / (C.sc_nh3 * C.re) + diffusion(f) * nue_t_s
                 + cr * (chem ? Sp(ATSAT_MASSFLUX_NH3) : 0.0) + pc * Sp(ATSAT_S_PRECIP_NH3);
    }
    rhs[ATSAT_NH3_CLOUD] = species(ATSAT_NH3_CLOUD, C.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_CLOUD));
    rhs[ATSAT_NH3_ICE]   = species(ATSAT_NH3_ICE, C.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_ICE));
    {   // :932-937, the flux limiter enters before the diffusion
        const int f = ATSAT_NH4SH;
        rhs[f] = - transport(f) + (chem ? Sp(ATSAT_FLUXLIM_NH4SH) : 0.0) + diffusion(f) / (C.sc_nh4sh * C.re)
                 + diffusion(f) * nue_t_s + cr * (chem ? Sp(ATSAT_MASSFLUX_NH4SH) : 0.0);
    }
    aux3[0] = rhs[U] + dpdr;                                      // :1011-1013
    aux3[1] = rhs[V] + dpdthe * inv_rm;
    aux3[2] = rhs[W] + dpdphi * inv_rmsinthe;
}

// cSaturnModel::RungeKuttaSat, planet/RungeKutta_Sat_Turb.cpp:110-387
void rk4_step(Ctx& X){
    const Geo& G = X.G; const atsat_knobs& K = X.K;
    buoyancy_ref_level(X);                                        // :116
    metric_pressure(X);                                           // :117
    const int im = G.im, jm = G.jm, km = G.km;
    const size_t N = (size_t)im * jm * km;
    const double dt = K.dt;
    const bool turb = K.turb_active != 0;
    const double tke_max_nd = 1000.0 / (X.C.u_0 * X.C.u_0);      // :175
    const double dis_min_nd = 1.0e-10;
    std::vector<double> rhs((size_t)ATSAT_NPROG * N, 0.0);        // rhs_* of every field
    std::vector<double> acc((size_t)ATSAT_NPROG * N, 0.0);        // acc_*
    std::vector<double> aux(3 * N, 0.0);
    const int nint = turb ? ATSAT_NPROG : ATSAT_NPROG - 2;        // fields pass B integrates
    for(int stage = 0; stage < 4; stage++){                       // :208
        const double c_in = (stage == 0 || stage == 1) ? 0.5 * dt : (stage == 2 ? dt : 0.0);
        const double wgt = (stage == 0 || stage == 3) ? 1.0 : 2.0;
        #pragma omp parallel for collapse(2) schedule(static)     // pass A, :220-247
        for(int i = 1; i < im - 1; i++){
            for(int j = 2; j < jm - 2; j++){
                const CellGeo geo = cell_geo(X, i, j);
                double r[ATSAT_NPROG], a3[3];
                for(int k = 0; k < km - 1; k++){
                    rhs_cell(X, geo, i, j, k, r, a3);
                    const size_t n = X.at(i, j, k);
                    for(int f = 0; f < ATSAT_NPROG; f++) rhs[(size_t)f * N + n] = r[f];
                    for(int q = 0; q < 3; q++) aux[(size_t)q * N + n] = a3[q];
                }
            }
        }
        #pragma omp parallel for collapse(2) schedule(static)     // pass B, :250-320
        for(int i = 1; i < im - 1; i++){
            for(int j = 2; j < jm - 2; j++){
                for(int k = 0; k < km - 1; k++){
                    const size_t n = X.at(i, j, k);
                    for(int f = 0; f < nint; f++){
                        double& a = acc[(size_t)f * N + n];
                        const double r = rhs[(size_t)f * N + n];
                        a = (stage == 0) ? wgt * r : a + wgt * r;
                    }
                    if(stage < 3){
                        for(int f = 0; f < ATSAT_NPROG - 2; f++)
                            X.f(f)[n] = X.f(ATSAT_F_N0 + f)[n] + c_in * rhs[(size_t)f * N + n];
                        if(turb){
                            X.f(ATSAT_TKE)[n] = std::min(std::max(X.f(ATSAT_F_N0 + ATSAT_TKE)[n]
                                + c_in * rhs[(size_t)ATSAT_TKE * N + n], 0.0), tke_max_nd);
                            X.f(ATSAT_DIS)[n] = std::max(X.f(ATSAT_F_N0 + ATSAT_DIS)[n]
                                + c_in * rhs[(size_t)ATSAT_DIS * N + n], dis_min_nd);
                        }
                    }
                }
            }
        }
    }
    const double sixth_dt = dt / 6.0;                             // final assembly, :324-375
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 1; i < im - 1; i++){
        for(int j = 2; j < jm - 2; j++){
            for(int k = 0; k < km - 1; k++){
                const size_t n = X.at(i, j, k);
                X.f(ATSAT_T)[n] = X.f(ATSAT_F_N0 + ATSAT_T)[n] + sixth_dt * acc[(size_t)ATSAT_T * N + n];
                if(turb){
                    X.f(ATSAT_TKE)[n] = std::min(std::max(X.f(ATSAT_F_N0 + ATSAT_TKE)[n]
                        + sixth_dt * acc[(size_t)ATSAT_TKE * N + n], 0.0), tke_max_nd);
                    X.f(ATSAT_DIS)[n] = std::max(X.f(ATSAT_F_N0 + ATSAT_DIS)[n]
                        + sixth_dt * acc[(size_t)ATSAT_DIS * N + n], dis_min_nd);
                }
                for(int f = ATSAT_U; f <= ATSAT_NH4SH; f++)
                    X.f(f)[n] = X.f(ATSAT_F_N0 + f)[n] + sixth_dt * acc[(size_t)f * N + n];
            }
        }
    }
    // rhs_u/v/w and aux_u/v/w keep the last stage's values (the Poisson source of the next
    // physics iteration); only the cells the step wrote change.
    for(int i = 1; i < im - 1; i++) for(int j = 2; j < jm - 2; j++) for(int k = 0; k < km - 1; k++){
        const size_t n = X.at(i, j, k);
        X.f(ATSAT_RHS_U)[n] = rhs[(size_t)ATSAT_U * N + n];
        X.f(ATSAT_RHS_V)[n] = rhs[(size_t)ATSAT_V * N + n];
        X.f(ATSAT_RHS_W)[n] = rhs[(size_t)ATSAT_W * N + n];
        X.f(ATSAT_AUX_U)[n] = aux[n];
        X.f(ATSAT_AUX_V)[n] = aux[N + n];
        X.f(ATSAT_AUX_W)[n] = aux[2 * N + n];
    }
}

// cSaturnModel::prepareProjectionBoundaries(rigid_lid = 1), planet/Pressure_Sat.cpp:78-194
void prepare_projection_boundaries(Ctx& X){
    const Geo& G = X.G; const int im = G.im, jm = G.jm, km = G.km;
    double* au = X.f(ATSAT_AUX_U); double* av = X.f(ATSAT_AUX_V); double* aw = X.f(ATSAT_AUX_W);
    double* ru = X.f(ATSAT_RHS_U); double* rv = X.f(ATSAT_RHS_V); double* rw = X.f(ATSAT_RHS_W);
    auto cub = [&](double* F, int a, int b, int c){ return F[c] - 3.0 * F[b] + 3.0 * F[a]; };
    for(int j = 1; j < jm - 1; j++){                              // r-direction, :86-123
        for(int k = 0; k < km - 1; k++){
            const size_t n0 = X.at(0, j, k), n1 = X.at(1, j, k), n2 = X.at(2, j, k), n3 = X.at(3, j, k);
            const size_t t0 = X.at(im - 1, j, k), t1 = X.at(im - 2, j, k), t2 = X.at(im - 3, j, k), t3 = X.at(im - 4, j, k);
            au[n0] = 0.0; au[t0] = 0.0;
            double* Fs[5] = { av, aw, ru, rv, rw };
            for(double* F : Fs){ F[n0] = cub(F, n1, n2, n3); F[t0] = cub(F, t1, t2, t3); }
        }
    }
    for(int k = 0; k < km - 1; k++){                              // theta-direction, :126-148
        for(int i = 1; i < im - 1; i++){
            const size_t n0 = X.at(i, 0, k), n1 = X.at(i, 1, k), n2 = X.at(i, 2, k), n3 = X.at(i, 3, k);
            const size_t s0 = X.at(i, jm - 1, k), s1 = X.at(i, jm - 2, k), s2 = X.at(i, jm - 3, k), s3 = X.at(i, jm - 4, k);
            au[n0] = cub(au, n1, n2, n3); av[n0] = 0.0; aw[n0] = 0.0;
            au[s0] = cub(au, s1, s2, s3); av[s0] = 0.0; aw[s0] = 0.0;
            ru[n0] = cub(ru, n1, n2, n3); rv[n0] = 0.0; rw[n0] = 0.0;
            ru[s0] = cub(ru, s1, s2, s3); rv[s0] = 0.0; rw[s0] = 0.0;
        }
    }
    for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){     // phi seam copy, :150-161
        double* Fs[6] = { au, av, aw, ru, rv, rw };
        for(double* F : Fs) F[X.at(i, j, km - 1)] = F[X.at(i, j, 0)];
    }
}

// The domain mean of p_dyn over i 1..im-2, j 1..jm-2, k 0..km-2 that PRESS_LINE holds
// (PressureSolver.h:310-314). The reference's OpenMP reduction has no fixed order; here every
// (i, j) row is summed over k, then the rows in (i, j) order, the order the device keeps too.
// ATSAT_ORACLE_SEQ_MEAN=1 sums every cell in (i, j, k) order instead, the reference's order on
// one thread, so that a test can compare the restatement with the reference bit for bit.
double press_interior_mean(const Ctx& X, const double* p){
    const Geo& G = X.G;
    const char* e = std::getenv("ATSAT_ORACLE_SEQ_MEAN");
    const bool seq = e && std::atoi(e) != 0;
    double s = 0.0;
    long n = 0;
    for(int i = 1; i < G.im - 1; i++) for(int j = 1; j < G.jm - 1; j++){
        if(seq) for(int k = 0; k < G.km - 1; k++) s += p[X.at(i, j, k)];
        else {
            double r = 0.0;
            for(int k = 0; k < G.km - 1; k++) r += p[X.at(i, j, k)];
            s += r;
        }
        n += G.km - 1;
    }
    return n ? s / n : 0.0;
}

// PRESS_LINE, PressureSolver.h:502-570: per colour (j + k) & 1, each column's rows i = 1..im-2
// solved by Thomas with the theta / phi neighbours lagged and both radial ends folded
// (p0 = c43 p1 - c13 p2); a non-finite result leaves the cell as it was.
void press_line_pass(Ctx& X, double* p, const std::vector<double>& line_rhs, const std::vector<double>& sinthe_table,
                     double inv_dr2, double inv_dthe2, double inv_dphi2){
    const Geo& G = X.G;
    const int im = G.im, jm = G.jm, km = G.km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    for(int colour = 0; colour < 2; colour++){
        #pragma omp parallel
        {
            std::vector<double> aa(im), bb(im), cc(im), dd(im), cp(im), dp(im);
            #pragma omp for collapse(2) schedule(static)
            for(int j = 1; j < jm - 1; j++){
                for(int k = 0; k < km - 1; k++){
                    if(((j + k) & 1) != colour) continue;
                    const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
                    for(int i = 1; i < im - 1; i++){
                        const double rm = X.metricRadius(G.radz[i]);
                        const double num1 = 1.0 * 1.0 * inv_dr2;
                        const double num2 = inv_dthe2 / (rm * rm);
                        const double num3 = inv_dphi2 / (rm * rm * sinthe_table[j] * sinthe_table[j]);
                        const double acc_h = num2 * (p[X.at(i, j + 1, k)] + p[X.at(i, j - 1, k)])
                                           + num3 * (p[X.at(i, j, ke)] + p[X.at(i, j, kw)]);
                        aa[i] = -num1; bb[i] = 2.0 * (num1 + num2 + num3); cc[i] = -num1;
                        dd[i] = acc_h - line_rhs[X.at(i, j, k)];
                    }
                    const int lo = 1, hi = im - 2;
                    if(hi > lo){ bb[lo] += aa[lo] * c43; cc[lo] -= aa[lo] * c13; }
                    else dd[lo] -= aa[lo] * p[X.at(lo - 1, j, k)];
                    if(hi > lo){ bb[hi] += cc[hi] * c43; aa[hi] -= cc[hi] * c13; }
                    else dd[hi] -= cc[hi] * p[X.at(hi + 1, j, k)];
                    cp[lo] = cc[lo] / bb[lo];
                    dp[lo] = dd[lo] / bb[lo];
                    for(int n = lo + 1; n <= hi; n++){
                        const double dn = bb[n] - aa[n] * cp[n - 1];
                        cp[n] = cc[n] / dn;
                        dp[n] = (dd[n] - aa[n] * dp[n - 1]) / dn;
                    }
                    double x = dp[hi];
                    if(std::isfinite(x)) p[X.at(hi, j, k)] = x;
                    for(int n = hi - 1; n >= lo; n--){
                        x = dp[n] - cp[n] * x;
                        if(std::isfinite(x)) p[X.at(n, j, k)] = x;
                    }
                }
            }
        }
    }
}

// PRESS_BAROTROPIC, PressureSolver.h:573-708: the column residual R(j, k) of src - Lc(p) over
// i = 1..im-2 (end rows read as the boundary pass will set them) divided by A2 = sum 1 / rm^2; its
// DFT over the N = km-1 meridians (modes of round-off size against a zonally symmetric row
// dropped); per mode the tridiagonal in theta with folded ends, mode 0 made consistent, pinned at
// j = 1 and shifted to a zero sin-weighted mean; the inverse DFT added to every level.
void press_baro_pass(Ctx& X, double* p, const std::vector<double>& line_rhs, const std::vector<double>& sinthe_table,
                     double inv_dr2, double inv_dthe2, double inv_dphi2){
    const Geo& G = X.G;
    const int im = G.im, jm = G.jm, km = G.km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int N = km - 1, J = jm - 2, NM = N / 2;
    double A2 = 0.0;
    for(int i = 1; i < im - 1; i++){ const double rm = X.metricRadius(G.radz[i]); A2 += 1.0 / (rm * rm); }
    auto col_resid = [&](int j, int k) -> double {
        const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
        double s = 0.0;
        for(int i = 1; i < im - 1; i++){
            const double rm = X.metricRadius(G.radz[i]);
            const double num1 = 1.0 * 1.0 * inv_dr2;
            const double num2 = inv_dthe2 / (rm * rm);
            const double num3 = inv_dphi2 / (rm * rm * sinthe_table[j] * sinthe_table[j]);
            auto pj = [&](int jn) -> double {
                if(jn == 0) return c43 * p[X.at(i, 1, k)] - c13 * p[X.at(i, 2, k)];
                if(jn == jm - 1) return c43 * p[X.at(i, jm - 2, k)] - c13 * p[X.at(i, jm - 3, k)];
                return p[X.at(i, jn, k)]; };
            auto pi_ = [&](int in) -> double {
                if(in == 0) return c43 * p[X.at(1, j, k)] - c13 * p[X.at(2, j, k)];
                if(in == im - 1) return c43 * p[X.at(im - 2, j, k)] - c13 * p[X.at(im - 3, j, k)];
                return p[X.at(in, j, k)]; };
            const double p0 = p[X.at(i, j, k)];
            const double lc = num1 * (pi_(i + 1) + pi_(i - 1) - 2.0 * p0)
                            + num2 * (pj(j + 1) + pj(j - 1) - 2.0 * p0)
                            + num3 * (p[X.at(i, j, ke)] + p[X.at(i, j, kw)] - 2.0 * p0);
            s += line_rhs[X.at(i, j, k)] - lc;
        }
        return s / A2; };
    std::vector<double> R((size_t)J * N, 0.0);
    #pragma omp parallel for schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < N; k++) R[(size_t)(j - 1) * N + k] = col_resid(j, k);
    std::vector<double> ct((size_t)(NM + 1) * N), st((size_t)(NM + 1) * N);
    for(int mm = 0; mm <= NM; mm++) for(int k = 0; k < N; k++){
        const double a = 2.0 * M_PI * (double)((long)mm * k % N) / N;
        ct[(size_t)mm * N + k] = std::cos(a); st[(size_t)mm * N + k] = std::sin(a); }
    std::vector<double> Ac((size_t)(NM + 1) * J, 0.0), As((size_t)(NM + 1) * J, 0.0);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int mm = 0; mm <= NM; mm++) for(int jj = 0; jj < J; jj++){
        double sc = 0.0, ss = 0.0;
        const double* r_ = &R[(size_t)jj * N];
        for(int k = 0; k < N; k++){ sc += r_[k] * ct[(size_t)mm * N + k]; ss += r_[k] * st[(size_t)mm * N + k]; }
        if(mm > 0){
            double ra = 0.0;
            for(int k = 0; k < N; k++) ra += std::abs(r_[k]);
            const double tol = 1.0e-12 * ra;
            if(std::abs(sc) < tol) sc = 0.0;
            if(std::abs(ss) < tol) ss = 0.0;
        }
        Ac[(size_t)mm * J + jj] = sc; As[(size_t)mm * J + jj] = ss;
    }
    const double e = inv_dthe2;
    #pragma omp parallel
    {
        std::vector<double> a(J), b(J), c(J), cp(J), dp(J);
        #pragma omp for schedule(static)
        for(int mm = 0; mm <= NM; mm++){
            const double sl = std::sin(M_PI * mm / N);
            const double lam = 4.0 * sl * sl * inv_dphi2;
            for(int jj = 0; jj < J; jj++){
                const double s_ = sinthe_table[jj + 1];
                a[jj] = e; c[jj] = e; b[jj] = -2.0 * e - lam / (s_ * s_);
            }
            a[0] = 0.0;     b[0] += c43 * e;     c[0] -= c13 * e;
            c[J - 1] = 0.0; b[J - 1] += c43 * e; a[J - 1] -= c13 * e;
            for(int part = 0; part < 2; part++){
                if(mm == 0 && part == 1) continue;
                if(2 * mm == N && part == 1) continue;
                double* d = (part == 0 ? &Ac[(size_t)mm * J] : &As[(size_t)mm * J]);
                double b0 = b[0], c0 = c[0];
                if(mm == 0){
                    double yr = 0.0, ys = 0.0;
                    for(int jj = 0; jj < J; jj++){
                        const double y = (jj == 0 || jj == J - 1) ? 1.5 : 1.0;
                        const double s_ = std::abs(std::sin(G.thez[jj + 1]));
                        yr += y * d[jj]; ys += y * s_; }
                    for(int jj = 0; jj < J; jj++) d[jj] -= yr / ys * std::abs(std::sin(G.thez[jj + 1]));
                    b0 = 1.0; c0 = 0.0; d[0] = 0.0;
                }
                cp[0] = c0 / b0; dp[0] = d[0] / b0;
                for(int jj = 1; jj < J; jj++){
                    const double dn = b[jj] - a[jj] * cp[jj - 1];
                    cp[jj] = c[jj] / dn; dp[jj] = (d[jj] - a[jj] * dp[jj - 1]) / dn;
                }
                d[J - 1] = dp[J - 1];
                for(int jj = J - 2; jj >= 0; jj--) d[jj] = dp[jj] - cp[jj] * d[jj + 1];
                if(mm == 0){
                    double sq = 0.0, sw = 0.0;
                    for(int jj = 0; jj < J; jj++){ const double s_ = std::abs(std::sin(G.thez[jj + 1])); sq += s_ * d[jj]; sw += s_; }
                    for(int jj = 0; jj < J; jj++) d[jj] -= sq / sw;
                }
            }
        }
    }
    #pragma omp parallel for schedule(static)
    for(int jj = 0; jj < J; jj++){
        for(int k = 0; k < N; k++){
            double q = 0.0;
            for(int mm = 0; mm <= NM; mm++){
                const double w_ = (mm == 0 || 2 * mm == N) ? 1.0 : 2.0;
                q += w_ * (Ac[(size_t)mm * J + jj] * ct[(size_t)mm * N + k]
                         + ((mm == 0 || 2 * mm == N) ? 0.0 : As[(size_t)mm * J + jj] * st[(size_t)mm * N + k]));
            }
            q /= N;
            for(int i = 0; i < im; i++) p[X.at(i, jj + 1, k)] += q;
        }
    }
}

// PressureSolver<cSaturnModel>::run, planet/PressureSolver.h:145-766 with PRESS_SRC 1,
// PRESS_INCR 1 (set by Run with GEO_BALANCE, cSaturnModel.cpp:612), PRESS_WALL 0, rigid lid:
// knobs.press_sweeps red-black sweeps, each followed by the line pass (knobs.press_line) and the
// barotropic correction (knobs.press_baro) when asked for, the boundary pass, and with the line
// pass the domain mean of p_dyn held.
void pressure_solve_ref(Ctx& X){
    const Geo& G = X.G; const atsat_knobs& K = X.K;
    const int im = G.im, jm = G.jm, km = G.km;
    std::vector<double> sinthe_table(jm);                         // :156-162
    for(int j = 0; j < jm; j++){
        sinthe_table[j] = std::sin(G.thez[j]);
        if(sinthe_table[j] < 0.0) sinthe_table[j] = 0.0;
        if(sinthe_table[j] == 0.0) sinthe_table[j] = 1.0e-5;
    }
    const double p_dyn_cap = K.pdyn_cap;
    const int n_sweeps = K.press_sweeps > 0 ? K.press_sweeps : 1;
    prepare_projection_boundaries(X);                             // :208
    const double inv_2dr = 1.0 / (2.0 * G.dr), inv_2dthe = 1.0 / (2.0 * G.dthe), inv_2dphi = 1.0 / (2.0 * G.dphi);
    const double inv_dr2 = 1.0 / (G.dr * G.dr), inv_dthe2 = 1.0 / (G.dthe * G.dthe), inv_dphi2 = 1.0 / (G.dphi * G.dphi);
    double* p = X.f(ATSAT_P_DYN);
    const double* Su = X.f(ATSAT_RHS_U); const double* Sv = X.f(ATSAT_RHS_V); const double* Sw = X.f(ATSAT_RHS_W);
    auto S_u_at = [&](int ii, int jj, int kk)->double{            // :246-248
        if(ii == 0 || ii == im - 1) return 0.0;
        return Su[X.at(ii, jj, kk)];
    };
    std::vector<double> lc_old((size_t)im * jm * km, 0.0);        // :249-282
    #pragma omp parallel for collapse(2) schedule(static)
    for(int i = 1; i < im - 1; i++){
        for(int j = 1; j < jm - 1; j++){
            const double rm = X.metricRadius(G.radz[i]);
            const double exp_rm = 1.0;
            const double num1 = exp_rm * exp_rm * inv_dr2;
            const double num2 = inv_dthe2 / (rm * rm);
            const double num3 = inv_dphi2 / (rm * rm * sinthe_table[j] * sinthe_table[j]);
            for(int k = 0; k < km - 1; k++){
                const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
                const double p0 = p[X.at(i, j, k)];
                const double l_ = num1 * (p[X.at(i + 1, j, k)] + p[X.at(i - 1, j, k)] - 2.0 * p0)
                                + num2 * (p[X.at(i, j + 1, k)] + p[X.at(i, j - 1, k)] - 2.0 * p0)
                                + num3 * (p[X.at(i, j, ke)] + p[X.at(i, j, kw)] - 2.0 * p0);
                lc_old[X.at(i, j, k)] = l_;
            }
        }
    }
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;              // cSaturnModel.h:710
    const bool press_line = K.press_line != 0, press_baro = K.press_baro != 0;
    std::vector<double> line_rhs;                                 // :307-308, :489
    if(press_line || press_baro) line_rhs.assign((size_t)im * jm * km, 0.0);
    const double p_mean0 = press_line ? press_interior_mean(X, p) : 0.0;
    for(int sweep = 0; sweep < n_sweeps; sweep++){
        for(int colour = 0; colour < 2; colour++){
            #pragma omp parallel for collapse(2) schedule(static)  // :322-495
            for(int i = 1; i < im - 1; i++){
                for(int j = 1; j < jm - 1; j++){
                    const double rm = X.metricRadius(G.radz[i]);
                    const double rm2 = rm * rm, exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;
                    const double sinthe = sinthe_table[j], sinthe2 = sinthe * sinthe;
                    const double inv_rm = 1.0 / rm, inv_rm2 = 1.0 / rm2;
                    const double inv_rmsinthe = 1.0 / (rm * sinthe);
                    const double inv_rm2sinthe2 = inv_rm2 / sinthe2;
                    const double num1 = exp_2_rm * inv_dr2;
                    const double num2 = inv_rm2 * inv_dthe2;
                    const double num3 = inv_rm2sinthe2 * inv_dphi2;
                    const double denom = 2.0 * num1 + 2.0 * num2 + 2.0 * num3;
                    for(int k = 0; k < km - 1; k++){
                        if(((i + j + k) & 1) != colour) continue;
                        const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
                        const double du_dr = (S_u_at(i + 1, j, k) - S_u_at(i - 1, j, k)) * inv_2dr;
                        const double dv_dthe = (Sv[X.at(i, j + 1, k)] - Sv[X.at(i, j - 1, k)]) * inv_2dthe;
                        const double dw_dphi = (Sw[X.at(i, j, ke)] - Sw[X.at(i, j, kw)]) * inv_2dphi;
                        double div_src = du_dr * exp_rm + dv_dthe * inv_rm + dw_dphi * inv_rmsinthe;
                        const double acc = (p[X.at(i + 1, j, k)] + p[X.at(i - 1, j, k)]) * num1
                                         + (p[X.at(i, j + 1, k)] + p[X.at(i, j - 1, k)]) * num2
                                         + (p[X.at(i, j, ke)] + p[X.at(i, j, kw)]) * num3;
                        const double den = denom;
                        const double src_max = den * p_dyn_cap;
                        if(!std::isfinite(div_src)) div_src = 0.0;
                        else if(div_src > src_max) div_src = src_max;
                        else if(div_src < -src_max) div_src = -src_max;
                        div_src += lc_old[X.at(i, j, k)];
                        if(press_line || press_baro) line_rhs[X.at(i, j, k)] = div_src;
                        p[X.at(i, j, k)] = (acc - div_src) / den;
                    }
                }
            }
        }
        if(press_line) press_line_pass(X, p, line_rhs, sinthe_table, inv_dr2, inv_dthe2, inv_dphi2);
        if(press_baro) press_baro_pass(X, p, line_rhs, sinthe_table, inv_dr2, inv_dthe2, inv_dphi2);
        for(int k = 0; k < km; k++) for(int j = 0; j < jm; j++){  // :718-724
            p[X.at(0, j, k)] = c43 * p[X.at(1, j, k)] - c13 * p[X.at(2, j, k)];
            p[X.at(im - 1, j, k)] = c43 * p[X.at(im - 2, j, k)] - c13 * p[X.at(im - 3, j, k)];
        }
        for(int k = 0; k < km; k++) for(int i = 0; i < im; i++){  // :726-732
            p[X.at(i, 0, k)] = c43 * p[X.at(i, 1, k)] - c13 * p[X.at(i, 2, k)];
            p[X.at(i, jm - 1, k)] = c43 * p[X.at(i, jm - 2, k)] - c13 * p[X.at(i, jm - 3, k)];
        }
        for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++)  // :734-740
            p[X.at(i, j, km - 1)] = p[X.at(i, j, 0)];
        if(press_line){                                           // :750-755
            const double shift = press_interior_mean(X, p) - p_mean0;
            const size_t N = (size_t)im * jm * km;
            for(size_t n = 0; n < N; n++) p[n] -= shift;
        }
    }
}

// AtomUtils::damp_wiggles(p_dyn, nullptr, along_i = true, along_j = false, along_k = true),
// lib/Utils.cpp:229-325, seam periodic (set_phi_seam_periodic, cSaturnModel.cpp:383)
void damp_pdyn(Ctx& X){
    const Geo& G = X.G; const int im = G.im, jm = G.jm, km = G.km;
    double* F = X.f(ATSAT_P_DYN);
    const size_t N = (size_t)im * jm * km;
    std::vector<double> tmp(F, F + N);
    const double coeff = 1.0 * 0.25;
    #pragma omp parallel for collapse(2) schedule(static)          // along k, :251-275
    for(int i = 0; i < im; i++){
        for(int j = 0; j < jm; j++){
            for(int k = 0; k < km - 1; k++){
                const int km1 = (k == 0 ? km - 2 : k - 1);
                const int kp1 = (k == km - 2 ? 0 : k + 1);
                F[X.at(i, j, k)] = tmp[X.at(i, j, k)] + coeff * (tmp[X.at(i, j, km1)] - 2.0 * tmp[X.at(i, j, k)] + tmp[X.at(i, j, kp1)]);
            }
            F[X.at(i, j, km - 1)] = F[X.at(i, j, 0)];
        }
    }
    std::memcpy(tmp.data(), F, N * sizeof(double));
    #pragma omp parallel for collapse(2) schedule(static)          // along i, :301-322
    for(int j = 0; j < jm; j++){
        for(int k = 0; k < km; k++){
            for(int i = 0; i < im; i++){
                const int im1 = std::max(i - 1, 0);
                const int ip1 = std::min(i + 1, im - 1);
                F[X.at(i, j, k)] = tmp[X.at(i, j, k)] + coeff * (tmp[X.at(im1, j, k)] - 2.0 * tmp[X.at(i, j, k)] + tmp[X.at(ip1, j, k)]);
            }
        }
    }
}

// SaturationAdjustment<cSaturnModel>::run, planet/SaturationAdjustment.h:255-746, as reached
// through SaturationAdjustmentSat::run (SaturationAdjustmentSat.cpp:19-49), with the model's
// defaults: CONSERVE 1, NEWTON 1, EVAP 1, ICE_ENTRY 0, BERGERON 2, FREEZE 1, DRY_EXIT 1,
// LOCAL_RHO 1, PSTAT_HYDRO 1 (no per-cell p_stat rewrite).
void saturation_adjust(Ctx& X, int gas){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    static const int gas_fields[3][3] = {
        { ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE },
        { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE },
        { ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE } };
    double* cq = X.f(gas_fields[gas][0]); double* cloud = X.f(gas_fields[gas][1]); double* ice = X.f(gas_fields[gas][2]);
    double* t = X.f(ATSAT_T); const double* ps = X.f(ATSAT_P_STAT);
    const double t_0 = C.t_0[gas], t_00 = C.t_00[gas], ep = C.ep[gas], lv = C.lv[gas], ls = C.ls[gas];
    const double Cc = C.C[gas], L0 = C.L0[gas], R = C.R[gas], del_alf = C.del_alf[gas], del_bet = C.del_bet[gas];
    const double C_i = C.C_i[gas], L0_i = C.L0_i[gas], del_alf_i = C.del_alf_i[gas], del_bet_i = C.del_bet_i[gas];
    const int iter_prec_end = 15; const double q_diff_min = 1.0e-3;            // :250-251
    const double t_range_inv = 1.0 / (t_0 - t_00);                              // :271
    const bool conserve = true, newton = true, evap = true, dry_exit = true;
    const int bergeron = 2;
    const bool freeze_on = true && conserve && t_00 > 0.0 && t_00 < t_0;        // :300
    const double Lf = ls - lv;                                                  // :303
    const double cp_mix = C.cp_mix, t_ref = C.t_ref;
    #pragma omp parallel for collapse(2) schedule(static)
    for(int k = 0; k < G.km; k++){
        for(int j = 0; j < G.jm; j++){
            for(int i = 0; i < G.im; i++){
                const size_t n = X.at(i, j, k);
                double t_u = t[n] * t_ref;                                       // :347
                const double p_u = ps[n];
                if(!(t_u > 0.0) || !(p_u > 0.0)) continue;                       // :358
                if(t_u > t_0){                                                   // :365-385
                    if(ice[n] > 0.0){
                        double dm = ice[n];
                        if(Lf > 0.0){
                            const double hc = cp_mix * X.rho_at(n);
                            const double cap = hc * (t_u - t_0) / Lf;
                            if(dm > cap) dm = cap;
                            t_u -= Lf * dm / hc;
                            t[n] = t_u / t_ref;
                        }
                        ice[n] -= dm;
                        cloud[n] += dm;
                    }
                }
                if(freeze_on && t_u < t_00 && cloud[n] > 0.0){                  // :388-399
                    double dm = cloud[n];
                    if(Lf > 0.0){
                        const double hc = cp_mix * X.rho_at(n);
                        const double cap = hc * (t_00 - t_u) / Lf;
                        if(dm > cap) dm = cap;
                        t_u += Lf * dm / hc;
                        t[n] = t_u / t_ref;
                    }
                    cloud[n] -= dm;
                    ice[n] += dm;
                }
                if(cq[n] < 0.0) cq[n] = 0.0;                                     // :400-402
                if(cloud[n] < 0.0) cloud[n] = 0.0;
                if(ice[n] < 0.0) ice[n] = 0.0;
                const double rho_c = X.rho_at(n);                                // :414
                const double E_Rain_0 = svp(t_u, Cc, L0, R, del_alf, del_bet);    // :416-418
                const double q_Rain_0 = rho_c * ep * E_Rain_0 / p_u;
                if(q_Rain_0 <= 0.0) continue;                                    // :422
                double q_entry = q_Rain_0;
                if(bergeron >= 2 && cloud[n] <= 0.0 && ice[n] <= 0.0 && t_u < t_0)   // :433-434
                    q_entry = rho_c * ep * svp(t_u, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                if(cq[n] <= q_entry){                                            // :436-448
                    const double q_c0 = cloud[n], q_i0 = ice[n];
                    if(!evap || !conserve || (q_c0 <= 0.0 && q_i0 <= 0.0)) continue;
                    const double q_Ice_0 = rho_c * ep * svp(t_u, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                    double q_t0;
                    if(q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
                    else if(q_i0 <= 0.0) q_t0 = q_Rain_0;
                    else q_t0 = q_Ice_0;
                    if(std::fabs(cq[n] - q_t0) <= q_diff_min * q_t0) continue;
                }
                double q_v_b = cq[n], q_c_b = cloud[n], q_i_b = ice[n];         // :451-463
                double T = t_u;
                double q_v_hyp = q_v_b;
                double q_rain_now = q_Rain_0;
                for(int itr = 1; itr <= iter_prec_end; itr++){                  // :465-590
                    double CND = (T - t_00) * t_range_inv;
                    double DEP = (t_0 - T) * t_range_inv;
                    if(T <= t_00){ CND = 0.0; DEP = 1.0; }
                    if(T >= t_0){ CND = 1.0; DEP = 0.0; }
                    const bool berg = bergeron != 0 && T < t_0 && q_v_b <= q_rain_now
                                   && (q_i_b > 0.0 || (bergeron >= 2 && q_c_b <= 0.0));
                    if(berg){
                        CND = 0.0; DEP = 1.0;
                        if(q_c_b > 0.0){
                            T -= lv * q_c_b / (cp_mix * rho_c);
                            q_v_b += q_c_b;
                            q_c_b = 0.0;
                        }
                    }
                    double d_q_v = q_v_hyp - q_v_b;
                    double d_q_c = -d_q_v * CND;
                    double d_q_i = -d_q_v * DEP;
                    if(conserve){
                        if(d_q_v < -q_v_b){ d_q_v = -q_v_b; d_q_c = -d_q_v * CND; d_q_i = -d_q_v * DEP; }
                        if(q_c_b + d_q_c < 0.0){ const double ex = -(q_c_b + d_q_c); d_q_c = -q_c_b; d_q_i -= ex; }
                        if(q_i_b + d_q_i < 0.0){ const double ex = -(q_i_b + d_q_i); d_q_i = -q_i_b; d_q_c -= ex;
                                                 if(q_c_b + d_q_c < 0.0) d_q_c = -q_c_b; }
                        d_q_v = -(d_q_c + d_q_i);
                    }
                    T += (lv * d_q_c + ls * d_q_i) / (cp_mix * rho_c);
                    q_v_b += d_q_v;
                    q_c_b += d_q_c;
                    q_i_b += d_q_i;
                    if(q_v_b < 0.0) q_v_b = 0.0;
                    if(q_c_b < 0.0) q_c_b = 0.0;
                    if(q_i_b < 0.0) q_i_b = 0.0;
                    const double E_Rain = svp(T, Cc, L0, R, del_alf, del_bet);
                    const double E_Ice = svp(T, C_i, L0_i, R, del_alf_i, del_bet_i);
                    const double q_Rain = rho_c * ep * E_Rain / p_u;
                    const double q_Ice = rho_c * ep * E_Ice / p_u;
                    q_rain_now = q_Rain;
                    if(q_c_b > 0.0 && q_i_b > 0.0) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
                    else if(bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_v_hyp = q_Ice;
                    else if(q_i_b == 0.0) q_v_hyp = q_Rain;
                    else q_v_hyp = q_Ice;
                    if(T >= t_0){
                        if(q_i_b > 0.0){
                            double dm = q_i_b;
                            if(Lf > 0.0){
                                const double hc = cp_mix * rho_c;
                                const double cap = hc * (T - t_0) / Lf;
                                if(dm > cap) dm = cap;
                                T -= Lf * dm / hc;
                            }
                            q_i_b -= dm;
                            q_c_b += dm;
                        }
                    }
                    const double q_diff = std::fabs(q_v_b - q_v_hyp) / (q_v_hyp + 1e-20);
                    if(q_diff <= q_diff_min) break;
                    if(dry_exit && conserve && q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;
                    if(newton){
                        const double dT = 0.1;
                        const double q_Rain_d = rho_c * ep * svp(T + dT, Cc, L0, R, del_alf, del_bet) / p_u;
                        const double q_Ice_d = rho_c * ep * svp(T + dT, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                        double q_hyp_d;
                        if(q_c_b > 0.0 && q_i_b > 0.0) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
                        else if(bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_hyp_d = q_Ice_d;
                        else if(q_i_b == 0.0) q_hyp_d = q_Rain_d;
                        else q_hyp_d = q_Ice_d;
                        double CNDn = (T - t_00) * t_range_inv;
                        if(T <= t_00) CNDn = 0.0;
                        if(T >= t_0) CNDn = 1.0;
                        if(berg) CNDn = 0.0;
                        const double Gf = (lv * CNDn + ls * (1.0 - CNDn)) / (cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
                        double wgt = 1.0 / (1.0 + std::max(Gf, 0.0));
                        if(wgt < 0.05) wgt = 0.05;
                        q_v_hyp = q_v_b + wgt * (q_v_hyp - q_v_b);
                    } else {
                        q_v_hyp = 0.5 * (q_v_hyp + q_v_b);
                    }
                }
                cq[n] = q_v_b;                                                   // :611-614
                cloud[n] = q_c_b;
                ice[n] = q_i_b;
                t[n] = T / t_ref;
            }
        }
    }
    const size_t N = (size_t)G.im * G.jm * G.km;                                 // :673-685
    #pragma omp parallel for schedule(static)
    for(size_t n = 0; n < N; n++){
        if(cq[n] <= 0.0) cq[n] = 0.0;
        if(cloud[n] <= 0.0) cloud[n] = 0.0;
        if(ice[n] <= 0.0) ice[n] = 0.0;
    }
}

// cSaturnModel::rebuildPstatHydro, planet/InitVariables_Sat.cpp:556-578 (no CP_H2 bottom fix)
void rebuild_pstat(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double exp_pressure = C.g / (C.gam * C.R_ref);
    const double* t = X.f(ATSAT_T); double* ps = X.f(ATSAT_P_STAT);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < G.jm; j++) for(int k = 0; k < G.km; k++){
        double p = C.p_ref * std::pow(t[X.at(0, j, k)], exp_pressure);
        ps[X.at(0, j, k)] = p;
        for(int i = 1; i < G.im; i++){
            const double T0 = t[X.at(i - 1, j, k)] * C.t_ref, T1 = t[X.at(i, j, k)] * C.t_ref;
            const double a = exp_pressure * C.gam * (double)(G.hgt[i] - G.hgt[i - 1]);
            if(T0 > 0.0 && T1 > 0.0){
                const double dT = T0 - T1;
                p *= (std::fabs(dT) > 1.0e-9 * T0) ? std::pow(T1 / T0, a / dT) : std::exp(-a / T0);
            } else p = C.p_ref * std::pow(std::max(0.0, t[X.at(i, j, k)]), exp_pressure);
            ps[X.at(i, j, k)] = p;
        }
    }
}

// Precipitation<cSaturnModel>::column, planet/Precipitation.h:273-445 (coupling on)
void precip_column(Ctx& X, int gas, const int* fld, const int* Pf, const int* Sf){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const int im = G.im;
    const double c_c_au = 2.0e-8, c_i_au = 5.0e-8, q_c_crit = 5.0e-4, c_ac = 1.0e-7;   // :93-98
    const double c_rim = 1.0e-8, z_csg = 2.5e-8, a_ev = 2.76e-3;
    const double c_melt_K = 0.2, c_frz_K = 0.2, P_ref = 1.0e-4, ratio_max = 5.0, v_fall = 2.0;   // :123-139
    const double rs = 0.0045;                                     // precip_rate_scale, cSaturnModel.h:408
    const double Lf = C.ls[gas] - C.lv[gas];
    const double P_max_flux = 1.0e-4, dep_frac = 1.0, ratio_cap = ratio_max;
    const double t_ref = C.t_ref, t_frz = C.t_0[gas], t_low = C.t_00[gas];
    const double* t = X.f(ATSAT_T); const double* ps = X.f(ATSAT_P_STAT);
    const double* qv_f = X.f(fld[0]); const double* qc_f = X.f(fld[1]); const double* qi_f = X.f(fld[2]);
    double* Pr = X.f(Pf[0]); double* Ps = X.f(Pf[1]); double* Pg = X.f(Pf[2]);
    double* Sv = X.f(Sf[0]); double* Sc = X.f(Sf[1]); double* Si = X.f(Sf[2]);
    double* Qp = X.f(ATSAT_Q_PRECIP);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 0; j < G.jm; j++){
        for(int k = 0; k < G.km; k++){
            const int i_base = 0, i_top = im - 1;
            Pr[X.at(i_top, j, k)] = 0.0; Ps[X.at(i_top, j, k)] = 0.0; Pg[X.at(i_top, j, k)] = 0.0;
            for(int i = i_top - 1; i >= i_base; i--){
                const size_t n = X.at(i, j, k), nu = X.at(i + 1, j, k);
                const double T = t[n] * t_ref;
                const double p_u = ps[n];
                const double q_v = std::max(0.0, qv_f[n]);
                const double q_c = std::max(0.0, qc_f[n]);
                const double q_i = std::max(0.0, qi_f[n]);
                const double E_r = svp(T, C.C[gas], C.L0[gas], C.R[gas], C.del_alf[gas], C.del_bet[gas]);
                const double q_sat = (p_u > E_r) ? X.rho_at(n) * C.ep[gas] * E_r / p_u : q_v;
                const double dz = std::max(0.0, X.layer_thickness_m(i));
                if(dz <= 0.0){ Pr[n] = Pr[nu]; Ps[n] = Ps[nu]; Pg[n] = Pg[nu]; continue; }
                const double F_r_in = Pr[nu], F_s_in = Ps[nu], F_g_in = Pg[nu];
                const double Rain = std::min(ratio_cap, F_r_in / P_ref);
                const double Snow = std::min(ratio_cap, F_s_in / P_ref);
                const double Graupel = std::min(ratio_cap, F_g_in / P_ref);
                const double Rain_79 = (Rain > 0.0) ? std::pow(Rain, 7.0 / 9.0) : 0.0;
                const double Rain_49 = (Rain > 0.0) ? std::pow(Rain, 4.0 / 9.0) : 0.0;
                const bool warm = (T >= t_frz);
                const double S_c_au = (warm && q_c > q_c_crit) ? (rs * c_c_au) * (q_c - q_c_crit) : 0.0;
                const double S_i_au = (!warm && T >= t_low) ? (rs * c_i_au) * q_i : 0.0;
                const double S_ac = (warm) ? (rs * c_ac) * q_c * Rain_79 : 0.0;
                const double S_s_rim = (!warm) ? (rs * c_rim) * q_c * Snow : 0.0;
                const double S_g_rim = (!warm) ? (rs * c_rim) * q_c * Graupel : 0.0;
                const double S_csg = (!warm && q_c > q_c_crit) ? (rs * z_csg) * q_c * Snow : 0.0;
                const double f_melt = warm ? std::min(1.0, c_melt_K * (T - t_frz)) : 0.0;
                const double f_frz = warm ? 0.0 : std::min(1.0, c_frz_K * (t_frz - T));
                const double F_s_melt = f_melt * F_s_in, F_g_melt = f_melt * F_g_in, F_r_frz = f_frz * F_r_in;
                const double F_r_avail = std::max(0.0, F_r_in - F_r_frz + F_s_melt + F_g_melt + (S_c_au + S_ac) * dz);
                const double S_ev_raw = (warm && Rain > 0.0 && q_v < q_sat) ? a_ev * (q_sat - q_v) * Rain_49 : 0.0;
                const double F_ev_sat = std::max(0.0, q_sat - q_v) * v_fall;
                const double F_ev = std::min(std::min(S_ev_raw * dz, F_r_avail), F_ev_sat);
                const double S_ev = F_ev / dz;
                Qp[n] += Lf * (S_s_rim + S_g_rim + S_csg) + Lf * (F_r_frz - F_s_melt - F_g_melt) / dz - C.lv[gas] * F_ev / dz;
                const double tau = dz / v_fall;
                const double d_c = std::min(dep_frac * q_c, (S_c_au + S_ac + S_s_rim + S_g_rim + S_csg) * tau);
                const double d_i = std::min(dep_frac * q_i, S_i_au * tau);
                const double nd = (C.L_atm * 1.0e3) / C.u_0;
                Sc[n] = - std::max(0.0, d_c) / tau * nd;
                Si[n] = - std::max(0.0, d_i) / tau * nd;
                Sv[n] = + S_ev * nd;
                const double F_r = F_r_in - F_r_frz + F_s_melt + F_g_melt - F_ev + (S_c_au + S_ac) * dz;
                const double F_s = F_s_in - F_s_melt + (S_i_au + S_s_rim - S_csg) * dz;
                const double F_g = F_g_in - F_g_melt + F_r_frz + (S_g_rim + S_csg) * dz;
                Pr[n] = std::min(P_max_flux, std::max(0.0, F_r));
                Ps[n] = std::min(P_max_flux, std::max(0.0, F_s));
                Pg[n] = std::min(P_max_flux, std::max(0.0, F_g));
            }
        }
    }
}

// Precipitation<cSaturnModel>::run, planet/Precipitation.h:191-270, with sedimentNH4SH
// (:542-574) and surfaceMap (:467-536)
void precipitation(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const size_t N = (size_t)G.im * G.jm * G.km;
    const int zero_me[10] = { ATSAT_Q_PRECIP,
        ATSAT_S_PRECIP_H2O, ATSAT_S_PRECIP_H2O_CLOUD, ATSAT_S_PRECIP_H2O_ICE,
        ATSAT_S_PRECIP_NH3, ATSAT_S_PRECIP_NH3_CLOUD, ATSAT_S_PRECIP_NH3_ICE,
        ATSAT_S_PRECIP_CH4, ATSAT_S_PRECIP_CH4_CLOUD, ATSAT_S_PRECIP_CH4_ICE };
    for(int z : zero_me) std::fill(X.f(z), X.f(z) + N, 0.0);
    {
        const int fld[3] = { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE };
        const int Pf[3] = { ATSAT_P_RAIN, ATSAT_P_SNOW, ATSAT_P_GRAUPEL };
        const int Sf[3] = { ATSAT_S_PRECIP_H2O, ATSAT_S_PRECIP_H2O_CLOUD, ATSAT_S_PRECIP_H2O_ICE };
        precip_column(X, ATSAT_GAS_H2O, fld, Pf, Sf);
    }
    {
        const int fld[3] = { ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE };
        const int Pf[3] = { ATSAT_P_NH3_RAIN, ATSAT_P_NH3_SNOW, ATSAT_P_NH3_GRAUPEL };
        const int Sf[3] = { ATSAT_S_PRECIP_NH3, ATSAT_S_PRECIP_NH3_CLOUD, ATSAT_S_PRECIP_NH3_ICE };
        precip_column(X, ATSAT_GAS_NH3, fld, Pf, Sf);
    }
    {
        const int fld[3] = { ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE };
        const int Pf[3] = { ATSAT_P_CH4_RAIN, ATSAT_P_CH4_SNOW, ATSAT_P_CH4_GRAUPEL };
        const int Sf[3] = { ATSAT_S_PRECIP_CH4, ATSAT_S_PRECIP_CH4_CLOUD, ATSAT_S_PRECIP_CH4_ICE };
        precip_column(X, ATSAT_GAS_CH4, fld, Pf, Sf);
    }
    {   // sedimentNH4SH, :542-574
        const double rho_ref = C.r_mix, t_hi = C.t_0_nh4sh, t_lo = C.t_00_nh4sh;
        const double r_p = 1.0e-5, rho_c = 1200.0, mu = 9.0e-6;
        const double v_settle = (2.0 / 9.0) * r_p * r_p * (rho_c - rho_ref) * C.g / mu;
        const double* t = X.f(ATSAT_T); const double* q4 = X.f(ATSAT_NH4SH); double* P4 = X.f(ATSAT_P_NH4SH);
        #pragma omp parallel for schedule(static)
        for(size_t n = 0; n < N; n++){
            const double T = t[n] * C.t_ref;
            const double q = std::max(0.0, q4[n]);
            P4[n] = (T >= t_lo && T <= t_hi) ? q * v_settle : 0.0;
        }
    }
    {   // surfaceMap at i_base = 0, :467-492
        for(int j = 0; j < G.jm; j++) for(int k = 0; k < G.km; k++){
            const size_t n = X.at(0, j, k), m2 = (size_t)j * G.km + k;
            const double h2o_srf = X.f(ATSAT_P_RAIN)[n] + X.f(ATSAT_P_SNOW)[n] + X.f(ATSAT_P_GRAUPEL)[n];
            const double nh3_srf = X.f(ATSAT_P_NH3_RAIN)[n] + X.f(ATSAT_P_NH3_SNOW)[n] + X.f(ATSAT_P_NH3_GRAUPEL)[n];
            const double ch4_srf = X.f(ATSAT_P_CH4_RAIN)[n] + X.f(ATSAT_P_CH4_SNOW)[n] + X.f(ATSAT_P_CH4_GRAUPEL)[n];
            const double nh4sh_srf = X.f(ATSAT_P_NH4SH)[n];
            if(X.S.v[ATSAT_V_SRF_H2O]) X.S.v[ATSAT_V_SRF_H2O][m2] = h2o_srf;
            if(X.S.v[ATSAT_V_SRF_NH3]) X.S.v[ATSAT_V_SRF_NH3][m2] = nh3_srf;
            if(X.S.v[ATSAT_V_SRF_CH4]) X.S.v[ATSAT_V_SRF_CH4][m2] = ch4_srf;
            if(X.S.v[ATSAT_V_SRF_NH4SH]) X.S.v[ATSAT_V_SRF_NH4SH][m2] = nh4sh_srf;
            if(X.S.v[ATSAT_V_SRF_TOTAL]) X.S.v[ATSAT_V_SRF_TOTAL][m2] = h2o_srf + nh3_srf + ch4_srf + nh4sh_srf;
        }
    }
}

// RK scalar sync, planet/cSaturnModel.cpp:955-977 (ATSAT_RK_SCALAR_SYNC 2; h2s / nh4sh too,
// which is a plain copy of unchanged values when the NH4SH equilibrium is off)
void rk_scalar_sync(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    const int ids[] = { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE,
                        ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE, ATSAT_T, ATSAT_H2S, ATSAT_NH4SH };
    for(int f : ids) std::memcpy(X.f(ATSAT_F_N0 + f), X.f(f), N * sizeof(double));
}

// Species of the clamp and of the lid / wall copies (FileIO_Sat.cpp:720-725, BC_Sat.cpp:86-91;
// chrom is off).
const int kSpecies[11] = { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S,
                           ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE,
                           ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE, ATSAT_NH4SH };

// BoundaryConditions<cSaturnModel>::bcTheta / bcPhi / bcRadius (BoundaryConditions.h:151-518)
// and BC_Sat::bcRadius's species copies (BC_Sat.cpp:61-103), for the fields of the step. The
// model's defaults: bc_margin 1, POLE_COPY 1 (copy), POLE_FILL 1 (one open row), RADIUS_COPY 2
// (Neumann), BOTTOM_FORM 3 (cubic), RIGID_LID 1, TOP_TAPER 2, SPECIES_LID_COPY 1,
// SPECIES_WALL_COPY 2, periodic seam. Of the chemistry arrays of the lists, the step holds
// massflux_*, difflux_*, thermalmassflux and fluxlim_nh4sh; the fluxes j_*, jT_* and the rates w_*
// are not held (no computation here reads their boundary values), nor are the force and heat
// diagnostics.
void boundary_conditions(Ctx& X){
    const Geo& G = X.G; const int im = G.im, jm = G.jm, km = G.km;
    const bool turb = X.K.turb_active != 0;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    // bc_fields_theta_extrap (BC_Sat.cpp:132-149) among the step's fields: t, u, species
    std::vector<int> th_ext = { ATSAT_T, ATSAT_U, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE,
        ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S, ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE, ATSAT_NH4SH,
        ATSAT_MASSFLUX_H2S, ATSAT_MASSFLUX_NH3, ATSAT_MASSFLUX_NH4SH,
        ATSAT_DIFFLUX_H2S, ATSAT_DIFFLUX_NH3, ATSAT_DIFFLUX_NH4SH, ATSAT_THERMALMASSFLUX };
    // bc_fields_theta_zero (BC_Sat.cpp:154-156): v, w, fluxlim_nh4sh
    std::vector<int> th_zero = { ATSAT_V, ATSAT_W, ATSAT_FLUXLIM_NH4SH };
    std::vector<int> turbf = { ATSAT_TKE, ATSAT_DIS, ATSAT_NUE };
    const double floors[3] = { 0.0, 1.0e-10, 0.0 };                // BC_Sat.cpp:42, 178-179
    // ---- bcTheta, :292-391
    for(int k = 0; k < km - 1; k++){
        for(int i = 1; i < im - 1; i++){
            const int jf = 2;
            const double ramp = 1.0 / 2.0;
            for(int f : th_ext){ double* F = X.f(f);
                F[X.at(i, 1, k)] = F[X.at(i, jf, k)]; F[X.at(i, jm - 2, k)] = F[X.at(i, jm - 1 - jf, k)]; }
            for(int f : th_zero){ double* F = X.f(f);
                F[X.at(i, 1, k)] = ramp * F[X.at(i, jf, k)]; F[X.at(i, jm - 2, k)] = ramp * F[X.at(i, jm - 1 - jf, k)]; }
            if(turb) for(int q = 0; q < 3; q++){ double* F = X.f(turbf[q]);
                F[X.at(i, 1, k)] = std::max(floors[q], F[X.at(i, jf, k)]);
                F[X.at(i, jm - 2, k)] = std::max(floors[q], F[X.at(i, jm - 1 - jf, k)]); }
            for(int f : th_zero){ double* F = X.f(f); F[X.at(i, 0, k)] = 0.0; F[X.at(i, jm - 1, k)] = 0.0; }
            for(int f : th_ext){ double* F = X.f(f);                      // COPY
                F[X.at(i, 0, k)] = F[X.at(i, 1, k)]; F[X.at(i, jm - 1, k)] = F[X.at(i, jm - 2, k)]; }
            if(turb) for(int q = 0; q < 3; q++){ double* F = X.f(turbf[q]);
                F[X.at(i, 0, k)] = std::max(floors[q], c43 * F[X.at(i, 1, k)] - c13 * F[X.at(i, 2, k)]);
                F[X.at(i, jm - 1, k)] = std::max(floors[q], c43 * F[X.at(i, jm - 2, k)] - c13 * F[X.at(i, jm - 3, k)]); }
        }
    }
    // ---- bcPhi, periodic, :422-438 (bc_fields_phi: t, u, v, w, species)
    std::vector<int> phif = { ATSAT_T, ATSAT_U, ATSAT_V, ATSAT_W, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE,
        ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S, ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE, ATSAT_NH4SH,
        ATSAT_MASSFLUX_H2S, ATSAT_MASSFLUX_NH3, ATSAT_MASSFLUX_NH4SH,
        ATSAT_DIFFLUX_H2S, ATSAT_DIFFLUX_NH3, ATSAT_DIFFLUX_NH4SH, ATSAT_FLUXLIM_NH4SH, ATSAT_THERMALMASSFLUX };
    for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
        for(int f : phif){ double* F = X.f(f); F[X.at(i, j, km - 1)] = F[X.at(i, j, 0)]; }
        if(turb) for(int q = 0; q < 3; q++){ double* F = X.f(turbf[q]);
            const double vv = std::max(floors[q], F[X.at(i, j, 0)]);
            F[X.at(i, j, 0)] = vv; F[X.at(i, j, km - 1)] = vv; }
    }
    // ---- bcRadius, :151-286 (bc_fields_radius: t, u, v, w, species)
    std::vector<int> radf = { ATSAT_T, ATSAT_U, ATSAT_V, ATSAT_W, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE,
        ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S, ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE, ATSAT_NH4SH,
        ATSAT_MASSFLUX_H2S, ATSAT_MASSFLUX_NH3, ATSAT_MASSFLUX_NH4SH,
        ATSAT_DIFFLUX_H2S, ATSAT_DIFFLUX_NH3, ATSAT_DIFFLUX_NH4SH, ATSAT_FLUXLIM_NH4SH, ATSAT_THERMALMASSFLUX };
    for(int j = 1; j < jm - 1; j++){
        for(int k = 0; k < km - 1; k++){
            for(int f : radf){ double* F = X.f(f);
                const double a = F[X.at(1, j, k)], b = F[X.at(2, j, k)], c = F[X.at(3, j, k)];
                F[X.at(0, j, k)] = (f != ATSAT_U) ? (c - 3.0 * b + 3.0 * a) : (c43 * a - c13 * b);
                F[X.at(im - 1, j, k)] = c43 * F[X.at(im - 2, j, k)] - c13 * F[X.at(im - 3, j, k)];
            }
            if(turb) for(int q = 0; q < 3; q++){ double* F = X.f(turbf[q]);
                F[X.at(0, j, k)] = std::max(floors[q], c43 * F[X.at(1, j, k)] - c13 * F[X.at(2, j, k)]);
                F[X.at(im - 1, j, k)] = std::max(floors[q], c43 * F[X.at(im - 2, j, k)] - c13 * F[X.at(im - 3, j, k)]); }
        }
    }
    {
        double* u = X.f(ATSAT_U); double* v = X.f(ATSAT_V); double* w = X.f(ATSAT_W);
        for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){   // rigid walls, :232-240
            u[X.at(0, j, k)] = 0.0; u[X.at(im - 1, j, k)] = 0.0; }
        const int iml = im - 1;
        for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){   // top taper 2, :264-285
            v[X.at(iml, j, k)] = 0.0; w[X.at(iml, j, k)] = 0.0;
            v[X.at(iml - 1, j, k)] = (1.0 / 3.0) * v[X.at(iml - 3, j, k)];
            w[X.at(iml - 1, j, k)] = (1.0 / 3.0) * w[X.at(iml - 3, j, k)];
            v[X.at(iml - 2, j, k)] = (2.0 / 3.0) * v[X.at(iml - 3, j, k)];
            w[X.at(iml - 2, j, k)] = (2.0 / 3.0) * w[X.at(iml - 3, j, k)];
        }
    }
    for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++){       // BC_Sat.cpp:94-102
        const double fr = X.rho_at(X.at(0, j, k)) / X.rho_at(X.at(1, j, k));
        for(int f : kSpecies){ double* F = X.f(f);
            F[X.at(im - 1, j, k)] = F[X.at(im - 2, j, k)];
            F[X.at(0, j, k)] = F[X.at(1, j, k)] * fr; }
    }
}

// cSaturnModel::clampNegativeSpecies, planet/FileIO_Sat.cpp:716-770
void clamp_species(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    for(int f : kSpecies){
        double* F = X.f(f);
        #pragma omp parallel for schedule(static)
        for(size_t n = 0; n < N; n++) if(!(F[n] >= 0.0)) F[n] = 0.0;
    }
}

// cSaturnModel::restoreVar(1.0), planet/cSaturnModel.cpp:1551-1602 (p_dynn and chromn are
// not fields of this step)
void restore_var(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    for(int f = 0; f < ATSAT_NPROG; f++){
        const double* s = X.f(f); double* d = X.f(ATSAT_F_N0 + f);
        #pragma omp parallel for schedule(static)
        for(size_t n = 0; n < N; n++) d[n] = 1.0 * s[n];
    }
}

// ---------------------------------------------------------------------------------------------
// The turbulence closure, Turbulence<cSaturnModel> (planet/Turbulence.h), at the model's
// defaults: turb_model k_omega_SST (cSaturnDefaults.cpp.inc:25), TURB_FREE 1 (Turbulence.h:141-
// 146), TURB_CLAMP_PASS 1 (:174-177), ABL_TAPER 0 (:178-181), TURB_CURV 1 (:761-762), TURB_BUOY
// from the knobs (3 on ATSAT, cSaturnModel.h:299), cp_lapse = cp_mix (CpH2.h:127-128 with
// cp_lapse_in_dynamics() false, cSaturnModel.h:340), no solid cells and surface_index 0
// (cSaturnModel.h:254-255), coord_stretching false (cSaturnModel.h:1282), periodic seam,
// bc_turb_active() = turb_active (cSaturnModel.h:309). Constructor defaults (:104-107, 243-256):
// nue_gas = 1.8e-5 m2/s.
struct Turb {
    static constexpr double C_nue = 0.028, Karman = 0.42, dis_min = 1.0e-10, nue_air = 1.8e-5;
    static double clamp3(double x, double lo, double hi){ return (x < lo) ? lo : ((x > hi) ? hi : x); }
    static double blend(double inner, double outer, double F1){ return F1 * inner + (1.0 - F1) * outer; }
};

// Arrays the closure owns (fill_lid_corners :514, seam_copy :547-556)
const int kTurbArrays[8] = { ATSAT_TKE, ATSAT_DIS, ATSAT_F_N0 + ATSAT_TKE, ATSAT_F_N0 + ATSAT_DIS,
                             ATSAT_NUE, ATSAT_PROD, ATSAT_TKE_SOURCE, ATSAT_DIS_SOURCE };

double height_m(const Geo& G, int i){ return (double)G.hgt[i] * 1.0e3; }   // Turbulence.h:262-265

// compute_vel_star, Turbulence.h:335-380 (:358): TURB_FREE, u_tau = vel_star_min in every column
void turb_vel_star(Ctx& X){
    double* vs = X.S.v[ATSAT_V_VEL_STAR];
    if(!vs) return;
    const size_t n = (size_t)X.G.jm * X.G.km;
    for(size_t q = 0; q < n; q++) vs[q] = 0.05;
}

// init_fields, Turbulence.h:426-491 (TURB_FREE: the background k and omega everywhere), with
// its closing seam_copy
void turb_init_fields(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double inv_u02 = 1.0 / (C.u_0 * C.u_0);
    const double L_m = C.L_atm * 1.0e3;
    const double nd_omega = L_m / C.u_0;
    const double k_bg = X.K.turb_k_bg;
    const double L_bg = L_m / double(G.im - 1);                   // turb_l_bg, :160-163
    const double om_bg = std::pow(Turb::C_nue, -0.25) * std::sqrt(k_bg) / (Turb::Karman * L_bg);
    double* tke = X.f(ATSAT_TKE); double* dis = X.f(ATSAT_DIS);
    double* tken = X.f(ATSAT_F_N0 + ATSAT_TKE); double* disn = X.f(ATSAT_F_N0 + ATSAT_DIS);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < G.jm - 1; j++){
        for(int k = 0; k < G.km - 1; k++){
            for(int i = 1; i < G.im - 1; i++){
                const double z_safe = 1.0e20;                      // free_wall_dist, :147, 465
                const double k_phys = std::max(0.0, k_bg);
                const double om_phys = std::pow(Turb::C_nue, -0.25) * std::sqrt(k_phys) / (Turb::Karman * z_safe);
                const double dis_nd = std::max(om_phys, om_bg) * nd_omega;
                const size_t n = X.at(i, j, k);
                tke[n] = std::max(0.0, k_phys * inv_u02);
                dis[n] = std::max(Turb::dis_min, dis_nd);
                tken[n] = tke[n];
                disn[n] = dis[n];
            }
        }
    }
}

// seam_copy, Turbulence.h:542-561
void turb_seam_copy(Ctx& X){
    const Geo& G = X.G;
    for(int a = 0; a < 8; a++){
        double* F = X.f(kTurbArrays[a]);
        for(int i = 0; i < G.im; i++) for(int j = 0; j < G.jm; j++) F[X.at(i, j, G.km - 1)] = F[X.at(i, j, 0)];
    }
}

// apply_wall_bc, Turbulence.h:382-424: TURB_FREE, the zero-gradient copy at i = 0 everywhere
void turb_wall_bc(Ctx& X){
    const Geo& G = X.G;
    double* tke = X.f(ATSAT_TKE); double* dis = X.f(ATSAT_DIS); double* nue = X.f(ATSAT_NUE);
    for(int j = 0; j < G.jm; j++) for(int k = 0; k < G.km; k++){
        tke[X.at(0, j, k)] = tke[X.at(1, j, k)];
        dis[X.at(0, j, k)] = dis[X.at(1, j, k)];
        nue[X.at(0, j, k)] = nue[X.at(1, j, k)];
    }
}

// clamp_nue, Turbulence.h:564-589 (ABL_TAPER 0: f = 1)
void turb_clamp_nue(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    const double nue_max = X.K.turb_nue_max / (X.C.u_0 * (X.C.L_atm * 1.0e3));
    double* nue = X.f(ATSAT_NUE);
    for(size_t n = 0; n < N; n++) nue[n] = Turb::clamp3(nue[n], 0.0, nue_max) * 1.0;
}

// fill_lid_corners, Turbulence.h:510-540: TURB_FREE with bc_turb_active(), the corner means (:533-539)
void turb_lid_corners(Ctx& X){
    const Geo& G = X.G; const int il = G.im - 1, jm = G.jm;
    for(int k = 0; k < G.km; k++) for(int a = 0; a < 8; a++){
        double* F = X.f(kTurbArrays[a]);
        F[X.at(il, 0, k)] = 0.5 * (F[X.at(il, 1, k)] + F[X.at(il - 1, 0, k)]);
        F[X.at(il, jm - 1, k)] = 0.5 * (F[X.at(il, jm - 2, k)] + F[X.at(il - 1, jm - 1, k)]);
    }
}

// compute_sources, Turbulence.h:613-883, with compute_k_omega_SST, :1020-1107
void turb_sources(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C; const atsat_knobs& K = X.K;
    const int im = G.im, jm = G.jm, km = G.km;
    double* tke = X.f(ATSAT_TKE); double* dis = X.f(ATSAT_DIS); double* nue = X.f(ATSAT_NUE);
    double* prod = X.f(ATSAT_PROD); double* tsrc = X.f(ATSAT_TKE_SOURCE); double* dsrc = X.f(ATSAT_DIS_SOURCE);
    const double* u = X.f(ATSAT_U); const double* v = X.f(ATSAT_V); const double* w = X.f(ATSAT_W);
    const double* t = X.f(ATSAT_T);
    const double L_m = C.L_atm * 1.0e3;
    const double nue_max = K.turb_nue_max / (C.u_0 * L_m);
    const double k_bg_nd = K.turb_k_bg / (C.u_0 * C.u_0);        // k_bg_of, TURB_FREE, :623-628
    const double dis_min = Turb::dis_min;
    // the moist gases of turb_moist_gas, cSaturnModel.cpp:259-267: H2O, NH3, CH4
    const int gv[3] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
    const int gc[3] = { ATSAT_H2O_CLOUD, ATSAT_NH3_CLOUD, ATSAT_CH4_CLOUD };
    const int gi[3] = { ATSAT_H2O_ICE, ATSAT_NH3_ICE, ATSAT_CH4_ICE };
    const int gq[3] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };

    // the clamp pass before the loop, :647-652
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++)
            for(int i = 1; i < im - 1; i++){
                const size_t n = X.at(i, j, k);
                tke[n] = std::max(0.0, tke[n]);
                dis[n] = std::max(dis_min, dis[n]);
            }

    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++){
        for(int k = 0; k < km - 1; k++){
            const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
            for(int i = 1; i < im - 1; i++){
                const size_t c = X.at(i, j, k);
                const size_t ip = X.at(i + 1, j, k), imn = X.at(i - 1, j, k);
                const size_t jp = X.at(i, j + 1, k), jmn = X.at(i, j - 1, k);
                const size_t pe = X.at(i, j, ke), pw = X.at(i, j, kw);
                // geometry, :675-694
                const double rm = X.metricRadius(G.radz[i]);
                const double exp_rm = 1.0;
                double sinthe = std::sin(G.thez[j]);
                if(sinthe == 0.0) sinthe = 1.0e-5;
                const double costhe = std::cos(G.thez[j]);
                const double cotanthe = costhe / sinthe;
                const double inv_rm = 1.0 / rm;
                const double rmsinthe = rm * sinthe;
                const double inv_2dr = 1.0 / (2.0 * G.dr);
                const double inv_2dthe = 1.0 / (2.0 * G.dthe);
                const double inv_2dphi = 1.0 / (2.0 * G.dphi);
                const double inv_rm2dthe = inv_2dthe / rm;
                const double inv_rmsinthe2dphi = inv_2dphi / rmsinthe;
                // velocity gradients, :699-709
                const double dudr = (u[ip] - u[imn]) * inv_2dr * exp_rm;
                const double dvdr = (v[ip] - v[imn]) * inv_2dr * exp_rm;
                const double dwdr = (w[ip] - w[imn]) * inv_2dr * exp_rm;
                const double dudthe = (u[jp] - u[jmn]) * inv_rm2dthe;
                const double dvdthe = (v[jp] - v[jmn]) * inv_rm2dthe;
                const double dwdthe = (w[jp] - w[jmn]) * inv_rm2dthe;
                const double dudphi = (u[pe] - u[pw]) * inv_rmsinthe2dphi;
                const double dvdphi = (v[pe] - v[pw]) * inv_rmsinthe2dphi;
                const double dwdphi = (w[pe] - w[pw]) * inv_rmsinthe2dphi;
                // k / omega radial gradients (no solid cells), :711-734
                const double dtkedr = (tke[ip] - tke[imn]) * inv_2dr * exp_rm;
                const double ddisdr = (dis[ip] - dis[imn]) * inv_2dr * exp_rm;
                // nu_t from the current k / omega, :739-748
                double cnue;
                {
                    const double dis_here = std::max(dis[c], dis_min);
                    const double tke_here = std::max(tke[c], 0.0);
                    cnue = tke_here / dis_here;
                    cnue = std::min(cnue, nue_max);
                }
                // the velocity gradient tensor with its curvature terms, :763-797
                const double u_r = u[c] * inv_rm;
                const double v_r = v[c] * inv_rm;
                const double w_r = w[c] * inv_rm;
                const double v_cot_r = v_r * cotanthe;
                const double w_cot_r = w_r * cotanthe;
                const double g_rr = dudr;
                const double g_rt = dudthe - v_r;
                const double g_rp = dudphi - w_r;
                const double g_tr = dvdr;
                const double g_tt = dvdthe + u_r;
                const double g_tp = dvdphi - w_cot_r;
                const double g_pr = dwdr;
                const double g_pt = dwdthe;
                const double g_pp = dwdphi + u_r + v_cot_r;
                const double der = 0.66667 * (g_rr + g_tt + g_pp);
                const double tk = tke[c];
                prod[c] = std::max(0.0,
                      (cnue * (2.0 * g_rr - der) - 0.66667 * tk) * g_rr
                    + (cnue * (g_rt + g_tr))                     * g_rt
                    + (cnue * (g_rp + g_pr))                     * g_rp
                    + (cnue * (2.0 * g_tt - der) - 0.66667 * tk) * g_tt
                    + (cnue * (g_tr + g_rt))                     * g_tr
                    + (cnue * (g_tp + g_pt))                     * g_tp
                    + (cnue * (2.0 * g_pp - der) - 0.66667 * tk) * g_pp
                    + (cnue * (g_pr + g_rp))                     * g_pr
                    + (cnue * (g_pt + g_tp))                     * g_pt);
                const double W12 = g_rt - g_tr;
                const double W13 = g_rp - g_pr;
                const double W23 = g_tp - g_pt;
                const double Omega = std::sqrt(W12 * W12 + W13 * W13 + W23 * W23);

                // compute_k_omega_SST, :1020-1107 (TURB_FREE: y_phys = free_wall_dist)
                {
                    const double a1 = 0.31, bet_star = 0.09, bet1 = 0.0333, bet2 = 0.0368;
                    const double gam1 = 0.413, gam2 = 0.2, sig_w2 = 1.168;
                    const double nue_air_nd = Turb::nue_air / (C.u_0 * L_m);
                    const double y_phys = 1.0e20;
                    const double y_star = y_phys / L_m;
                    const double dtkedr_neu = dtkedr, ddisdr_neu = ddisdr;
                    const double dtkedthe_neu = (tke[jp] - tke[jmn]) / (rm * 2.0 * G.dthe);
                    const double ddisdthe_neu = (dis[jp] - dis[jmn]) / (rm * 2.0 * G.dthe);
                    const double dtkedphi_neu = (tke[pe] - tke[pw]) / (rmsinthe * 2.0 * G.dphi);
                    const double ddisdphi_neu = (dis[pe] - dis[pw]) / (rmsinthe * 2.0 * G.dphi);
                    const double grad_dot = dtkedr_neu * ddisdr_neu
                                          + dtkedthe_neu * ddisdthe_neu
                                          + dtkedphi_neu * ddisdphi_neu;
                    const double CD_kw = std::max(2.0 * sig_w2 / std::max(dis[c], 1.0e-20) * grad_dot, 1.0e-20);
                    const double dis_safe = std::max(dis[c], 1.0e-20);
                    const double arg1 = std::min(
                        std::max(std::sqrt(tk) / (bet_star * dis_safe * y_star),
                                 500.0 * nue_air_nd / (y_star * y_star * dis_safe)),
                        4.0 * sig_w2 * tk / (CD_kw * y_star * y_star));
                    const double arg2 = std::max(
                        2.0 * std::sqrt(tk) / (bet_star * dis_safe * y_star),
                        500.0 * nue_air_nd / (y_star * y_star * dis_safe));
                    const double F1 = std::tanh(std::pow(arg1, 4));
                    const double F2 = std::tanh(std::pow(arg2, 2));
                    nue[c] = a1 * tk / std::max(a1 * dis_safe, Omega * F2);
                    const double tke_safe = std::max(tk, 1.0e-20);
                    const double P_k = std::min(prod[c], 20.0 * bet_star * tke_safe * dis_safe);
                    const double Y_k = bet_star * tke_safe * dis_safe;
                    const double P_w = Turb::blend(gam1, gam2, F1) * P_k * dis_safe / tke_safe;
                    const double Y_w = Turb::blend(bet1, bet2, F1) * dis_safe * dis_safe;
                    const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_safe * grad_dot;
                    tsrc[c] = P_k - Y_k;
                    dsrc[c] = P_w - Y_w + D_w;
                }

                // the buoyancy sink of the k equation, :836-868
                const int buoy = K.turb_buoy;
                if(buoy != 0){
                    const double Pr_t = 0.9;
                    const double T = t[c] * C.t_ref;
                    const double dz = height_m(G, i + 1) - height_m(G, i - 1);
                    const double dTdz = (t[ip] - t[imn]) * C.t_ref / dz;
                    double Gam = C.g / C.cp_mix, fac = 1.0;
                    if(buoy == 3 && T > 0.0){
                        const double rho = X.rho_at(c);
                        double A = 0.0, B = 0.0;
                        for(int q = 0; rho > 0.0 && q < 3; q++){
                            const double qc = (std::max(0.0, X.f(gc[q])[c]) + std::max(0.0, X.f(gi[q])[c])) / rho;
                            if(!(qc > 1.0e-9)) continue;
                            const double qv = std::max(0.0, X.f(gv[q])[c]) / rho;
                            const double ep = C.ep[gq[q]], lv = C.lv[gq[q]], ls = C.ls[gq[q]], t0 = C.t_0[gq[q]];
                            const double L = T >= t0 ? lv : ls;
                            A += L * qv / (C.R_mix * T);
                            B += ep * L * L * qv / (C.cp_mix * C.R_mix * T * T);
                        }
                        if(std::isfinite(A + B)){ Gam *= (1.0 + A) / (1.0 + B); fac = 1.0 + A; }
                    }
                    const double N2 = (T > 0.0 && dz > 0.0) ? C.g / T * (dTdz + Gam) * fac : 0.0;
                    if(std::isfinite(N2) && (N2 > 0.0 || buoy == 2)){
                        const double tL = L_m / C.u_0;
                        tsrc[c] -= std::max(0.0, nue[c]) * N2 * tL * tL / Pr_t;
                    }
                }
            }
        }
    }

    // the floor and the caps after the loop, :873-881 (clamp_out :633-645)
    #pragma omp parallel for collapse(2) schedule(static)
    for(int j = 1; j < jm - 1; j++)
        for(int k = 0; k < km - 1; k++)
            for(int i = 1; i < im - 1; i++){
                const size_t n = X.at(i, j, k);
                tke[n] = std::max(k_bg_nd, tke[n]);
                dis[n] = std::max(dis_min, dis[n]);
                nue[n] = Turb::clamp3(nue[n], 0.0, nue_max);
                const double dis_src_max = 20.0 * dis[n] * dis[n];
                dsrc[n] = Turb::clamp3(dsrc[n], -dis_src_max, dis_src_max);
                const double tke_src_max = 20.0 * tke[n] * dis[n];
                tsrc[n] = Turb::clamp3(tsrc[n], -tke_src_max, tke_src_max);
            }
}

// Turbulence<>::run, Turbulence.h:306-329 (zero_land_cells does nothing without solid cells)
void turbulence_run(Ctx& X){
    turb_vel_star(X);
    turb_sources(X);
    turb_wall_bc(X);
    turb_clamp_nue(X);
    turb_seam_copy(X);
    turb_lid_corners(X);
}

// Turbulence<>::init, Turbulence.h:282-304
void turbulence_init(Ctx& X){
    turb_vel_star(X);
    turb_init_fields(X);
    turb_seam_copy(X);
    turb_wall_bc(X);
    turb_sources(X);
    turb_clamp_nue(X);
    turb_seam_copy(X);
    turb_lid_corners(X);
}

// ---------------------------------------------------------------------------------------------
// The chemistry block of the physics iteration (planet/ChemistrySat.h, planet/FluxLimiter.h), at
// the model's defaults: ATSAT_NH4SH_EQ 1 (:244-247: the equilibrium adjustment forms NH4SH and
// the rate law's w_* are zero), ATSAT_CHEM_DIFFLUX_LAPLACIAN 1 (:370-371), ATSAT_CHEM_DIFFLUX_PLUS
// 1 (:162-164), ATSAT_CHEM_ENTHALPY 0 (:41-44), no solid cells (SeaMount 0), the polar divisor
// floor 0.4 (ATPhys.h:104-108), the periodic seam.

// ChemistrySat::EquilibriumNH4SH("loop"), ChemistrySat.h:294-331, with equilibrateNH4SHCell
// :250-289: every cell in place.
void nh4sh_equilibrium(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double R_univ = 8314.46, atm = 101325.0;
    const double L_f = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;   // nh4sh_formation_heat, :242
    double* t = X.f(ATSAT_T); double* nh3 = X.f(ATSAT_NH3); double* h2s = X.f(ATSAT_H2S);
    double* nh4sh = X.f(ATSAT_NH4SH); const double* rho_mix = X.f(ATSAT_RHO_MIX);
    const size_t N = (size_t)G.im * G.jm * G.km;
    #pragma omp parallel for schedule(dynamic, 4096)
    for(size_t n = 0; n < N; n++){
        auto Kp = [&](double T){ return std::pow(10.0, 14.82 - 4705.0 / T) * atm * atm; };
        const double T0 = t[n] * C.t_ref;
        if(!(T0 > 0.0)) continue;
        double rho = rho_mix[n];
        if(!(rho > 0.0) || !std::isfinite(rho)) rho = C.r_mix;
        const double a = std::max(0.0, nh3[n]) / C.m_nh3;
        const double b = std::max(0.0, h2s[n]) / C.m_h2s;
        const double s = std::max(0.0, nh4sh[n]) / C.m_nh4sh;
        const double heat = L_f * C.m_nh4sh / (C.cp_mix * rho);
        auto f = [&](double x){
            const double T = T0 + heat * x, RT = R_univ * T;
            return (a - x) * (b - x) * RT * RT - Kp(T); };
        const double f0 = f(0.0);
        double lo, hi;
        if(f0 > 0.0){ lo = 0.0; hi = std::min(a, b); }
        else if(s > 0.0){
            if(f(-s) <= 0.0){ lo = hi = -s; }
            else { lo = -s; hi = 0.0; }
        } else continue;
        if(hi > lo) for(int it = 0; it < 80; it++){
            const double xm = 0.5 * (lo + hi);
            if(f(xm) > 0.0) lo = xm; else hi = xm;
            if(hi - lo <= 1.0e-15 * std::max(a, std::max(b, s))) break;
        }
        const double x = (f0 > 0.0) ? lo : hi;
        if(x == 0.0) continue;
        nh3[n] = std::max(0.0, nh3[n] - C.m_nh3 * x);
        h2s[n] = std::max(0.0, h2s[n] - C.m_h2s * x);
        nh4sh[n] = std::max(0.0, nh4sh[n] + C.m_nh4sh * x);
        const double dT = heat * x;
        t[n] = (T0 + dT) / C.t_ref;
    }
}

// ChemistrySat::DiffMassFluxSat, ChemistrySat.h:365-472: pass 1 (the fluxes j_*, jT_* of every
// cell, :393-425) feeds only the cell's own thermalmassflux here, since difflux_* is the
// Laplacian form (:437-442); laplacian_spherical :645-662, derivative_1_order :555-609 (no solid
// cells: the centred differences). The rate law's w_* are zero (ChemMassRateSat with
// NH4SH_EQ), so the second term of thermalmassflux is t * 0.
void diff_mass_flux(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const double D_nh3 = C.mue_nh3 / (C.rg_nh3 * C.sc_nh3);         // :376-381
    const double D_h2s = C.mue_h2s / (C.rg_h2s * C.sc_h2s);
    const double D_nh4sh = C.mue_nh4sh / (C.rg_nh4sh * C.sc_nh4sh);
    const double DT_nh3 = D_nh3, DT_h2s = D_h2s, DT_nh4sh = D_nh4sh;
    const double* t = X.f(ATSAT_T); const double* h2s = X.f(ATSAT_H2S);
    const double* nh3 = X.f(ATSAT_NH3); const double* nh4sh = X.f(ATSAT_NH4SH);
    double* df_h2s = X.f(ATSAT_DIFFLUX_H2S); double* df_nh3 = X.f(ATSAT_DIFFLUX_NH3);
    double* df_nh4sh = X.f(ATSAT_DIFFLUX_NH4SH); double* tmf = X.f(ATSAT_THERMALMASSFLUX);
    const double floor_ = 0.4;
    #pragma omp parallel for collapse(2) schedule(static)
    for(int k = 0; k < G.km - 1; k++){
        for(int j = 1; j < G.jm - 1; j++){
            const int kw = X.wrap(k - 1), ke = X.wrap(k + 1);
            for(int i = 1; i < G.im - 1; i++){
                const size_t c = X.at(i, j, k);
                const size_t ip = X.at(i + 1, j, k), imn = X.at(i - 1, j, k);
                const size_t jp = X.at(i, j + 1, k), jmn = X.at(i, j - 1, k);
                const size_t pe = X.at(i, j, ke), pw = X.at(i, j, kw);
                const double rm = X.metricRadius(G.radz[i]);
                const double sinthe = std::sin(G.thez[j]);
                const double rmsinthe = rm * sinthe;
                auto d1 = [&](const double* F, double& dr_, double& dt_, double& dp_){
                    dr_ = (F[ip] - F[imn]) / (2.0 * G.dr);
                    dt_ = (F[jp] - F[jmn]) / (2.0 * G.dthe);
                    dp_ = (F[pe] - F[pw]) / (2.0 * G.dphi);
                };
                double dtdr, dtdthe, dtdphi, a_r, a_t, a_p, b_r, b_t, b_p, s_r, s_t, s_p;
                d1(t, dtdr, dtdthe, dtdphi);
                d1(h2s, b_r, b_t, b_p);
                d1(nh3, a_r, a_t, a_p);
                d1(nh4sh, s_r, s_t, s_p);
                const double jT_nh3 = DT_nh3 / t[c] * dtdr;
                const double jT_h2s = DT_h2s / t[c] * dtdr;
                const double jT_nh4sh = DT_nh4sh / t[c] * dtdr;
                const double dnh3 = a_r + std::abs(a_t) / rm + a_p / rmsinthe;
                const double dh2s = b_r + std::abs(b_t) / rm + b_p / rmsinthe;
                const double dnh4sh = s_r + std::abs(s_t) / rm + s_p / rmsinthe;
                const double j_nh3 = C.c_mix / C.r_mix * (C.m_nh3 * D_nh3 * dnh3 - jT_nh3);
                const double j_h2s = C.c_mix / C.r_mix * (C.m_h2s * D_h2s * dh2s - jT_h2s);
                const double j_nh4sh = C.c_mix / C.r_mix * (C.m_nh4sh * D_nh4sh * dnh4sh - jT_nh4sh);
                // laplacian_spherical, :645-662 (coord_stretching false: exp_rm = 1)
                auto lap = [&](const double* F){
                    const double exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;
                    const double costhe = std::cos(G.thez[j]);
                    const double rms = rm * std::max(sinthe, floor_);
                    const double d2cdr2 = (F[ip] - 2.0 * F[c] + F[imn]) / (G.dr * G.dr) * exp_2_rm;
                    const double d2cdthe2 = (F[jp] - 2.0 * F[c] + F[jmn]) / (G.dthe * G.dthe);
                    const double d2cdphi2 = (F[pe] - 2.0 * F[c] + F[pw]) / (G.dphi * G.dphi);
                    const double dcdr = (F[ip] - F[imn]) / (2.0 * G.dr) * exp_rm;
                    const double dcdthe = (F[jp] - F[jmn]) / (2.0 * G.dthe);
                    return d2cdr2
                         + 2.0 / rm * dcdr
                         + d2cdthe2 / (rm * rm)
                         + costhe / (rm * rms) * dcdthe
                         + d2cdphi2 / (rms * rms);
                };
                df_h2s[c] = D_h2s * lap(h2s);
                df_nh3[c] = D_nh3 * lap(nh3);
                df_nh4sh[c] = D_nh4sh * lap(nh4sh);
                const double w0 = 0.0;
                tmf[c] = (j_nh3 * C.cp_nh3
                        + j_h2s * C.cp_h2s
                        + j_nh4sh * C.cp_nh4sh)
                        * (dtdr + std::abs(dtdthe) / rm + dtdphi / rmsinthe)
                        + t[c] * (w0 * C.k_nh3 + w0 * C.k_h2s + w0 * C.k_nh4sh);
            }
        }
    }
}

// AtomUtils::damp_wiggles(f, nullptr, true, true, true), lib/Utils.cpp:229-325: one 1-2-1 pass
// along k (periodic, the seam copied), then j (clamped at the poles), then i (clamped).
void damp_3d(Ctx& X, int id){
    const Geo& G = X.G; const int im = G.im, jm = G.jm, km = G.km;
    double* F = X.f(id);
    const size_t N = (size_t)im * jm * km;
    std::vector<double> tmp(N);
    const double coeff = 1.0 * 0.25;
    std::memcpy(tmp.data(), F, N * sizeof(double));                 // along k, :251-271
    for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
        for(int k = 0; k < km - 1; k++){
            const int km1 = k == 0 ? km - 2 : k - 1, kp1 = k == km - 2 ? 0 : k + 1;
            F[X.at(i, j, k)] = tmp[X.at(i, j, k)] + coeff * (tmp[X.at(i, j, km1)] - 2.0 * tmp[X.at(i, j, k)] + tmp[X.at(i, j, kp1)]);
        }
        F[X.at(i, j, km - 1)] = F[X.at(i, j, 0)];
    }
    std::memcpy(tmp.data(), F, N * sizeof(double));                 // along j, :275-293
    for(int i = 0; i < im; i++) for(int j = 0; j < jm; j++){
        const int jm1 = std::max(j - 1, 0), jp1 = std::min(j + 1, jm - 1);
        for(int k = 0; k < km; k++)
            F[X.at(i, j, k)] = tmp[X.at(i, j, k)] + coeff * (tmp[X.at(i, jm1, k)] - 2.0 * tmp[X.at(i, j, k)] + tmp[X.at(i, jp1, k)]);
    }
    std::memcpy(tmp.data(), F, N * sizeof(double));                 // along i, :297-316
    for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++)
        for(int i = 0; i < im; i++){
            const int im1 = std::max(i - 1, 0), ip1 = std::min(i + 1, im - 1);
            F[X.at(i, j, k)] = tmp[X.at(i, j, k)] + coeff * (tmp[X.at(im1, j, k)] - 2.0 * tmp[X.at(i, j, k)] + tmp[X.at(ip1, j, k)]);
        }
}

// ChemistrySat::ChemMassRateSat, ChemistrySat.h:46-205, NH4SH_EQ: w_* = 0 (:182-187) and
// massflux_x = w_x + (+1) difflux_x (:194-199); then FluxLimiter<cSaturnModel>::nh4sh,
// FluxLimiter.h:90-190 (Superbee).
void chem_mass_rate(Ctx& X){
    const Geo& G = X.G;
    const int pairs[3][2] = { { ATSAT_MASSFLUX_H2S, ATSAT_DIFFLUX_H2S }, { ATSAT_MASSFLUX_NH3, ATSAT_DIFFLUX_NH3 },
                              { ATSAT_MASSFLUX_NH4SH, ATSAT_DIFFLUX_NH4SH } };
    const double dsign = 1.0, w0 = 0.0;
    for(int k = 0; k < G.km - 1; k++) for(int j = 1; j < G.jm - 1; j++) for(int i = 1; i < G.im - 1; i++){
        const size_t c = X.at(i, j, k);
        for(int q = 0; q < 3; q++) X.f(pairs[q][0])[c] = w0 + dsign * X.f(pairs[q][1])[c];
    }
}

inline double superbee_phi(double r){                               // FluxLimiter.h:62-64
    return std::max(0.0, std::max(std::min(2.0 * r, 1.0), std::min(r, 2.0)));
}

void flux_limiter_nh4sh(Ctx& X){
    const Geo& G = X.G; const int im = G.im, jm = G.jm, km = G.km;
    const double eps = 1.0e-12, c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const double* q_ = X.f(ATSAT_NH4SH); const double* u = X.f(ATSAT_U); const double* v = X.f(ATSAT_V);
    const double* w = X.f(ATSAT_W); double* fl = X.f(ATSAT_FLUXLIM_NH4SH);
    #pragma omp parallel for collapse(2) schedule(static)
    for(int k = 0; k < km - 1; k++){
        for(int j = 1; j < jm - 1; j++){
            for(int i = 1; i < im - 1; i++){
                const size_t c = X.at(i, j, k);
                const double q = q_[c];
                const double q_rp = q_[X.at(i + 1, j, k)], q_rm = q_[X.at(i - 1, j, k)];
                const double q_tp = q_[X.at(i, j + 1, k)], q_tm = q_[X.at(i, j - 1, k)];
                const double q_pp = q_[X.at(i, j, X.wrap(k + 1))], q_pm = q_[X.at(i, j, X.wrap(k - 1))];
                const double uu = u[c], vv = v[c], ww = w[c];
                const double rm = X.metricRadius(G.radz[i]);
                const double sinthe = std::max(0.4, std::abs(std::sin(G.thez[j])));
                const double inv_rm = 1.0 / rm;
                const double inv_rmsinthe = 1.0 / (rm * sinthe);
                double corr = 0.0;
                {
                    const double df = q_rp - q, db = q - q_rm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double r = (uu >= 0.0) ? db / denom
                        : ((i + 2 < im ? q_[X.at(i + 2, j, k)] : q_rp) - q_rp) / denom;
                    corr += (1.0 - superbee_phi(r)) * std::abs(uu) * (df - db) / (2.0 * G.dr);
                }
                {
                    const double df = q_tp - q, db = q - q_tm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double q_far = (j + 2 < jm) ? q_[X.at(i, j + 2, k)] : (c43 * q_tp - c13 * q);
                    const double r = (vv >= 0.0) ? db / denom : (q_far - q_tp) / denom;
                    corr += (1.0 - superbee_phi(r)) * std::abs(vv) * inv_rm * (df - db) / (2.0 * G.dthe);
                }
                {
                    const double df = q_pp - q, db = q - q_pm;
                    const double denom = df + (df >= 0.0 ? eps : -eps);
                    const double r = (ww >= 0.0) ? db / denom
                        : (q_[X.at(i, j, X.wrap(k + 2))] - q_pp) / denom;
                    corr += (1.0 - superbee_phi(r)) * std::abs(ww) * inv_rmsinthe * (df - db) / (2.0 * G.dphi);
                }
                fl[c] = corr;
            }
        }
    }
}

// The chemistry after the RK scalar sync, cSaturnModel.cpp:979-994 (Forces and Latent_Heat,
// :996-997, are diagnostics the step does not read).
void chemistry(Ctx& X){
    diff_mass_flux(X);
    damp_3d(X, ATSAT_DIFFLUX_H2S); damp_3d(X, ATSAT_DIFFLUX_NH3); damp_3d(X, ATSAT_DIFFLUX_NH4SH);
    damp_3d(X, ATSAT_THERMALMASSFLUX);
    chem_mass_rate(X);
    flux_limiter_nh4sh(X);
    damp_3d(X, ATSAT_MASSFLUX_H2S); damp_3d(X, ATSAT_MASSFLUX_NH3); damp_3d(X, ATSAT_MASSFLUX_NH4SH);
    damp_3d(X, ATSAT_FLUXLIM_NH4SH);
}

// ---------------------------------------------------------------------------------------------
// The moist convective adjustment, MoistConvectiveAdjustment<cSaturnModel>::run
// (planet/ConvectiveAdjustment.h:402-849), called after restoreVar and the closure every
// ATSAT_MCA iterations (cSaturnModel.cpp:1120-1145), then the MCA sync (:1147-1171). ATSAT's
// defaults: MCA_LEDOUX 1, MCA_LOADING 1, MCA_RELAX 1, MCA_PASSES 64, MCA_SMOOTH 2
// (cSaturnModel.h:367), MCA_INTERIOR 1 (:360), the t_0 phase rule (MCA_PHASE needs
// SATADJ_BERGERON 0, whose default is 2, :387), cp = cp_mix (CpH2 off in the dynamics), no solid
// cells. Gases H2O, NH3, CH4 with their liquid and ice sets; tracers H2S (ep_h2s, a gas in the
// virtual temperature) and NH4SH (condensed, with its formation heat when ATSAT_NH4SH_EQ).
struct McaGas { int v, c, ice; double ep, lv, ls, t_0, C, L0, R, da, db, Ci, L0i, dai, dbi; };
struct McaTracer { int f; double ep; bool condensed; double L; };

struct McaColumn {
    // one column's search / mixing state, index i (cells) or i * ng + g
    std::vector<double> hL, qt, qr, Teq, ceq;
    std::vector<char> eq_ok, touched;
};

void mca(Ctx& X){
    const Geo& G = X.G; const atsat_consts& C = X.C;
    const int im = G.im, jm = G.jm, km = G.km;
    const int gid[3] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };
    const int gv[3] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
    McaGas Gs[3];
    for(int g = 0; g < 3; g++){
        const int q = gid[g];
        Gs[g] = { gv[g], gv[g] + 1, gv[g] + 2, C.ep[q], C.lv[q], C.ls[q], C.t_0[q], C.C[q], C.L0[q], C.R[q],
                  C.del_alf[q], C.del_bet[q], C.C_i[q], C.L0_i[q], C.del_alf_i[q], C.del_bet_i[q] };
    }
    const double L_nh4sh = 4705.0 * std::log(10.0) * 8314.46 / 51.1114;   // ChemistrySat.h:242
    const McaTracer TR[2] = { { ATSAT_H2S, C.ep_h2s, false, 0.0 }, { ATSAT_NH4SH, 0.0, true, L_nh4sh } };
    const int ng = 3, nt = 2, smooth = 2, max_pass = 64;
    const double cp = C.cp_mix, grav = C.g;
    const double tol_K = 1.0e-7, tol_T = 1.0e-10;
    const int n_iter = 60;
    auto H = [&](double TT){ return cp * TT; };
    auto qsat = [&](int g, double T, double p){
        const McaGas& s = Gs[g];
        const double E = (T >= s.t_0) ? svp(T, s.C, s.L0, s.R, s.da, s.db) : svp(T, s.Ci, s.L0i, s.R, s.dai, s.dbi);
        const double q = s.ep * E / p;
        return std::isfinite(q) ? q : 0.0;
    };
    auto lat = [&](int g, double T){ return T >= Gs[g].t_0 ? Gs[g].lv : Gs[g].ls; };
    // :447-475, the equilibrium temperature of (h_L, q_t) at (p, z) by Illinois regula falsi
    auto equil = [&](double hL, const double* qt, double p, double z, double& T, double* cond) -> bool {
        const double T_lo = (hL - grav * z) / cp;
        if(!(T_lo > 0.0) || !(p > 0.0)) return false;
        bool any = false;
        for(int g = 0; g < ng; g++) if(qt[g] > qsat(g, T_lo, p)){ any = true; break; }
        if(!any){ T = T_lo; for(int g = 0; g < ng; g++) cond[g] = 0.0; return true; }
        double span = 0.0;
        for(int g = 0; g < ng; g++) span += std::max(Gs[g].lv, Gs[g].ls) * std::max(0.0, qt[g]);
        auto fof = [&](double TT){
            double f = H(TT) + grav * z - hL;
            for(int g = 0; g < ng; g++) f -= lat(g, TT) * std::max(0.0, qt[g] - qsat(g, TT, p));
            return f; };
        double a = T_lo, b = T_lo + span / cp + 1.0e-9;
        double fa = fof(a), fb = fof(b);
        int side = 0;
        for(int it = 0; it < n_iter && b - a > tol_T; it++){
            double c = (fb - fa != 0.0) ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
            if(!(c > a && c < b)) c = 0.5 * (a + b);
            const double fc = fof(c);
            if(fc < 0.0){ a = c; fa = fc; if(side == -1) fb *= 0.5; side = -1; }
            else        { b = c; fb = fc; if(side == +1) fa *= 0.5; side = +1; }
            if(std::fabs(fc) <= cp * tol_T){ a = b = c; break; }
        }
        T = b;
        for(int g = 0; g < ng; g++) cond[g] = std::max(0.0, qt[g] - qsat(g, T, p));
        return true;
    };
    auto zh = [&](int i){ return double(G.hgt[i]) * 1.0e3; };

    // MCA_SMOOTH: the conserved variables of every cell, then the latitude 1-2-1 passes, :501-541
    const size_t nc = (size_t)im * jm * km;
    std::vector<double> SH(nc, 0.0), SQ(nc * ng, 0.0), SR(nc * nt, 0.0);
    std::vector<char> SF(nc, 0);
    #pragma omp parallel for collapse(2)
    for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++) for(int i = 0; i < im; i++){
        const size_t c = X.at(i, j, k);
        const double TT = X.f(ATSAT_T)[c] * C.t_ref, rr = X.rho_at(c), zz = zh(i);
        if(!(TT > 0.0) || !(rr > 0.0) || !std::isfinite(TT + rr)) continue;
        double h = H(TT) + grav * zz;
        for(int g = 0; g < ng; g++){
            const double v_ = std::max(0.0, X.f(Gs[g].v)[c]) / rr, c_ = std::max(0.0, X.f(Gs[g].c)[c]) / rr,
                         i_ = std::max(0.0, X.f(Gs[g].ice)[c]) / rr;
            SQ[c * ng + g] = v_ + c_ + i_;
            h -= Gs[g].lv * c_ + Gs[g].ls * i_;
        }
        for(int r = 0; r < nt; r++){ SR[c * nt + r] = X.f(TR[r].f)[c] / rr; h -= TR[r].L * SR[c * nt + r]; }
        SH[c] = h; SF[c] = 1;
    }
    for(int pass = 0; pass < smooth; pass++){
        const std::vector<double> H0(SH), Q0(SQ), R0(SR);
        #pragma omp parallel for collapse(2)
        for(int j = 0; j < jm; j++) for(int k = 0; k < km; k++) for(int i = 0; i < im; i++){
            const size_t c = X.at(i, j, k);
            if(!SF[c]) continue;
            double wn[2] = { 0.0, 0.0 }; size_t cn[2] = { c, c };
            for(int s_ = 0; s_ < 2; s_++){
                const int jj = j + (s_ ? 1 : -1);
                if(jj < 0 || jj >= jm) continue;
                const size_t c2 = X.at(i, jj, k);
                if(SF[c2]){ wn[s_] = 0.25; cn[s_] = c2; }
            }
            const double wc = 1.0 - wn[0] - wn[1];
            SH[c] = wc * H0[c] + wn[0] * H0[cn[0]] + wn[1] * H0[cn[1]];
            for(int g = 0; g < ng; g++)
                SQ[c * ng + g] = wc * Q0[c * ng + g] + wn[0] * Q0[cn[0] * ng + g] + wn[1] * Q0[cn[1] * ng + g];
            for(int r = 0; r < nt; r++)
                SR[c * nt + r] = wc * R0[c * nt + r] + wn[0] * R0[cn[0] * nt + r] + wn[1] * R0[cn[1] * nt + r];
        }
    }

    // the columns, :553-849 (MCA_INTERIOR: i 1..im-2)
    #pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for(int j = 0; j < jm; j++){
        for(int k = 0; k < km; k++){
            const int i0 = 1, i1 = im - 2;
            if(i1 - i0 < 1) continue;
            std::vector<double> p(im), z(im), rho(im), w(im), T(im);
            bool ok = true;
            for(int i = i0; i <= i1 && ok; i++){
                const size_t c = X.at(i, j, k);
                T[i] = X.f(ATSAT_T)[c] * C.t_ref;
                p[i] = X.f(ATSAT_P_STAT)[c];
                z[i] = zh(i);
                rho[i] = X.rho_at(c);
                if(!(T[i] > 0.0) || !(p[i] > 0.0) || !(rho[i] > 0.0) || !std::isfinite(T[i] + p[i] + rho[i])) ok = false;
            }
            if(!ok) continue;
            for(int i = i0; i <= i1 && ok; i++){
                const double lo = (i > i0) ? 0.5 * (z[i - 1] + z[i]) : z[i0];
                const double hi = (i < i1) ? 0.5 * (z[i] + z[i + 1]) : z[i1];
                w[i] = rho[i] * (hi - lo);
                if(!(w[i] > 0.0)) ok = false;
            }
            if(!ok) continue;
            // the column's own conserved variables (:583-597), needed after the search
            auto own = [&](McaColumn& S){
                for(int i = i0; i <= i1; i++){
                    const size_t c = X.at(i, j, k);
                    double h = H(T[i]) + grav * z[i];
                    for(int g = 0; g < ng; g++){
                        const double qv = std::max(0.0, X.f(Gs[g].v)[c]) / rho[i];
                        const double qc = std::max(0.0, X.f(Gs[g].c)[c]) / rho[i];
                        const double qi = std::max(0.0, X.f(Gs[g].ice)[c]) / rho[i];
                        S.qt[i * ng + g] = qv + qc + qi;
                        h -= Gs[g].lv * qc + Gs[g].ls * qi;
                    }
                    for(int r = 0; r < nt; r++){
                        S.qr[i * nt + r] = X.f(TR[r].f)[c] / rho[i];
                        h -= TR[r].L * S.qr[i * nt + r];
                    }
                    S.hL[i] = h;
                }
            };
            McaColumn S;
            S.hL.assign(im, 0.0); S.qt.assign(im * ng, 0.0); S.qr.assign(im * nt, 0.0);
            S.Teq.assign(im, 0.0); S.ceq.assign(im * ng, 0.0); S.eq_ok.assign(im, 0); S.touched.assign(im, 0);
            for(int i = i0; i <= i1 && ok; i++) if(!SF[X.at(i, j, k)]) ok = false;
            if(!ok) continue;
            // the search runs on the smoothed column, :605-615
            for(int i = i0; i <= i1; i++){
                const size_t c = X.at(i, j, k);
                S.hL[i] = SH[c];
                for(int g = 0; g < ng; g++) S.qt[i * ng + g] = SQ[c * ng + g];
                for(int r = 0; r < nt; r++) S.qr[i * nt + r] = SR[c * nt + r];
            }
            auto trE = [&](int q){ double e = 0.0; for(int r = 0; r < nt; r++) e += TR[r].L * S.qr[q * nt + r]; return e; };
            auto tv = [&](double TT, const double* v_, const double* cnd, int q){
                double f = 1.0;
                for(int g = 0; g < ng; g++){ f += v_[g] * (1.0 / Gs[g].ep - 1.0); f -= cnd[g]; }
                for(int r = 0; r < nt; r++){
                    const double qq = S.qr[q * nt + r];
                    if(TR[r].ep > 0.0) f += qq * (1.0 / TR[r].ep - 1.0);
                    else if(TR[r].condensed) f -= qq;
                }
                return TT * f;
            };
            double pc[3], pv[3], ec[3], ev[3];
            // :645-662, the parcel of a lifted to b against b's own equilibrium
            auto unstable = [&](int a, int b, bool* moist) -> bool {
                double Tp;
                if(!equil(S.hL[a] + trE(a), &S.qt[a * ng], p[b], z[b], Tp, pc)) return false;
                if(!S.eq_ok[b]){
                    if(!equil(S.hL[b] + trE(b), &S.qt[b * ng], p[b], z[b], S.Teq[b], &S.ceq[b * ng])) return false;
                    S.eq_ok[b] = 1;
                }
                const double Tb = S.Teq[b];
                for(int g = 0; g < ng; g++) ec[g] = S.ceq[b * ng + g];
                bool wet = false;
                for(int g = 0; g < ng; g++){ pv[g] = S.qt[a * ng + g] - pc[g]; if(pc[g] > 0.0) wet = true; ev[g] = S.qt[b * ng + g] - ec[g]; }
                const double dTv = tv(Tp, pv, pc, a) - tv(Tb, ev, ec, b);
                if(moist) *moist = wet;
                return dTv > tol_K;
            };
            // :663-677, cell q to equilibrium from its h_L / q_t
            auto settle = [&](int q) -> bool {
                double TT;
                if(!equil(S.hL[q] + trE(q), &S.qt[q * ng], p[q], z[q], TT, pc)) return false;
                S.Teq[q] = TT; S.eq_ok[q] = 1;
                for(int g = 0; g < ng; g++) S.ceq[q * ng + g] = pc[g];
                S.touched[q] = 1;
                return true;
            };
            auto mix = [&](int a, int b) -> bool {
                double sw = 0.0, sh = 0.0, sq[3] = { 0.0, 0.0, 0.0 }, sr[2] = { 0.0, 0.0 };
                for(int q = a; q <= b; q++){
                    sw += w[q]; sh += w[q] * S.hL[q];
                    for(int g = 0; g < ng; g++) sq[g] += w[q] * S.qt[q * ng + g];
                    for(int r = 0; r < nt; r++) sr[r] += w[q] * S.qr[q * nt + r];
                }
                for(int q = a; q <= b; q++){
                    S.hL[q] = sh / sw;
                    for(int g = 0; g < ng; g++) S.qt[q * ng + g] = sq[g] / sw;
                    for(int r = 0; r < nt; r++) S.qr[q * nt + r] = sr[r] / sw;
                    if(!settle(q)) return false;
                }
                return true;
            };
            // the census of the unadjusted column, :695-703
            std::vector<char> cen_u(im, 0);
            for(int i = i0; i < i1; i++) cen_u[i] = unstable(i, i + 1, nullptr);
            // the sweeps, :714-747; link[i]: cells i and i+1 were mixed in one segment
            std::vector<char> link(im, 0);
            int passes = 0; long long layers = 0; bool changed = true, fail = false;
            while(changed && passes < max_pass && !fail){
                changed = false; passes++;
                int i = i0;
                while(i < i1 && !fail){
                    const bool uns = (passes == 1 && !S.touched[i] && !S.touched[i + 1]) ? cen_u[i] != 0 : unstable(i, i + 1, nullptr);
                    if(!uns){ i++; continue; }
                    int a = i, b = i + 1;
                    for(;;){
                        if(!mix(a, b)){ fail = true; break; }
                        bool ext = false;
                        if(a > i0 && unstable(a - 1, a, nullptr)){ a--; ext = true; }
                        if(b < i1 && unstable(b, b + 1, nullptr)){ b++; ext = true; }
                        if(!ext) break;
                    }
                    if(fail) break;
                    layers += (b - a + 1);
                    for(int q = a; q < b; q++) link[q] = 1;
                    changed = true;
                    i = b;
                }
            }
            if(fail || layers == 0) continue;
            // MCA_SMOOTH: the union of the segments mixed in the column's own state, :751-783
            own(S);
            std::fill(S.touched.begin(), S.touched.end(), 0); std::fill(S.eq_ok.begin(), S.eq_ok.end(), 0);
            for(int a = i0; a < i1 && !fail; ){
                if(!link[a]){ a++; continue; }
                int b = a;
                while(b < i1 && link[b]) b++;
                if(!mix(a, b)) fail = true;
                a = b + 1;
            }
            if(fail) continue;
            // write back, :811-826
            for(int q = i0; q <= i1; q++){
                if(!S.touched[q]) continue;
                const size_t c = X.at(q, j, k);
                const double TT = S.Teq[q];
                X.f(ATSAT_T)[c] = TT / C.t_ref;
                for(int g = 0; g < ng; g++){
                    const int x = q * ng + g;
                    const double pcg = S.ceq[x];
                    const double qv = S.qt[x] - pcg;
                    double qc, qi;
                    if(TT >= Gs[g].t_0){ qc = pcg; qi = 0.0; } else { qi = pcg; qc = 0.0; }
                    X.f(Gs[g].v)[c] = qv * rho[q];
                    X.f(Gs[g].c)[c] = qc * rho[q];
                    X.f(Gs[g].ice)[c] = qi * rho[q];
                }
                for(int r = 0; r < nt; r++) X.f(TR[r].f)[c] = S.qr[q * nt + r] * rho[q];
            }
        }
    }
}

// the MCA sync, cSaturnModel.cpp:1147-1171: t, the nine condensables, h2s and nh4sh into their
// start-of-step copies
void mca_sync(Ctx& X){
    const size_t N = (size_t)X.G.im * X.G.jm * X.G.km;
    const int fl[12] = { ATSAT_T, ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_NH3, ATSAT_NH3_CLOUD,
                         ATSAT_NH3_ICE, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE, ATSAT_H2S, ATSAT_NH4SH };
    for(int f : fl){
        const double* s = X.f(f); double* d = X.f(ATSAT_F_N0 + f);
        for(size_t n = 0; n < N; n++) d[n] = s[n];
    }
}

// One iteration of cSaturnModel::Run's loop, planet/cSaturnModel.cpp:864-1171, radiation off (the
// chemistry block with knobs.chem_block, the turbulence closure with turb_active and
// knobs.turb_closure, the moist convective adjustment every knobs.mca_stride iterations).
void step(Ctx& X, int iter){
    if(iter % 2 == 0){
        mixture_density(X);                                       // :869
        pressure_solve_ref(X);                                    // :873
        damp_pdyn(X);                                             // :894-897
        saturation_adjust(X, ATSAT_GAS_CH4);                      // :908-927
        saturation_adjust(X, ATSAT_GAS_H2O);
        saturation_adjust(X, ATSAT_GAS_NH3);
        rebuild_pstat(X);                                         // :928
        if(X.K.chem_block) nh4sh_equilibrium(X);                  // :933
        if(X.K.precip_on) precipitation(X);                       // :939
        rk_scalar_sync(X);                                        // :955-977
        if(X.K.chem_block) chemistry(X);                          // :979-994
    }
    rk4_step(X);                                                  // :1030
    boundary_conditions(X);                                       // :1091-1093
    clamp_species(X);                                             // :1101
    restore_var(X);                                               // :1111
    if(X.K.turb_active && X.K.turb_closure) turbulence_run(X);    // :1116
    if(X.K.mca_stride > 0 && iter % X.K.mca_stride == 0){         // :1120-1145
        mca(X);
        mca_sync(X);                                              // :1147-1171
    }
}

// The sums of the reference's two budget instruments (print-only there, cSaturnModel.cpp:640-830),
// one value per quantity:
//   out[g], g = 0, 1, 2   the H2O, NH3 and CH4 columns, vapour + cloud + ice, in g/m2: cwb_col with
//                         i_hi = 0 (cSaturnModel.cpp:776-787), every level with a thickness
//                         (layer_thickness_m, cSaturnModel.h:689-692), weighted by sin(the.z[j]),
//                         over cw_norm = sum_j sin(the.z[j]) km (:798); the reference's OpenMP
//                         reduction has no fixed order, the loops here run k, j, i;
//   out[3 + i]            the level-mean temperature of level i in K: tb_lev (:656-666) over all
//                         rows (ATSAT_TB_ROWS unset), j then k in order, over tb_norm (:673).
void budgets(Ctx& X, double* out){
    const Geo& G = X.G;
    double norm = 0.0;
    for(int j = 0; j < G.jm; j++) norm += std::sin(G.thez[j]) * G.km;
    const int gv[3][3] = { { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE }, { ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE },
                           { ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE } };
    for(int g = 0; g < 3; g++){
        const double* a = X.f(gv[g][0]); const double* b = X.f(gv[g][1]); const double* c = X.f(gv[g][2]);
        double s_ = 0.0;
        for(int k = 0; k < G.km; k++)
            for(int j = 0; j < G.jm; j++){
                const double sj = std::sin(G.thez[j]);
                for(int i = 0; i < G.im; i++){
                    const double dz = X.layer_thickness_m(i);
                    const size_t n = X.at(i, j, k);
                    if(dz > 0.0) s_ += sj * dz * (a[n] + b[n] + c[n]);
                }
            }
        out[g] = s_ * 1.0e3 / norm;
    }
    const double* t = X.f(ATSAT_T);
    for(int i = 0; i < G.im; i++){
        double s_ = 0.0;
        for(int j = 0; j < G.jm; j++){
            const double sj = std::sin(G.thez[j]);
            for(int k = 0; k < G.km; k++) s_ += sj * t[X.at(i, j, k)];
        }
        out[3 + i] = s_ * X.C.t_ref / norm;
    }
}

}  // namespace

extern "C" int orc_budgets(const atsat_grid* g, const atsat_consts* c, const atsat_knobs* k, atsat_host_state* s,
                           double* out){
    if(!g || !c || !k || !s || !out) return -1;
    Ctx X(*g, *c, *k, *s); budgets(X, out); return 0;
}

#define ORC_ENTRY(name, body) \
    extern "C" int name(const atsat_grid* g, const atsat_consts* c, const atsat_knobs* k, atsat_host_state* s){ \
        if(!g || !c || !k || !s) return -1; Ctx X(*g, *c, *k, *s); body; return 0; }

ORC_ENTRY(orc_mixture_density, mixture_density(X))
ORC_ENTRY(orc_buoyancy_ref_level, buoyancy_ref_level(X))
ORC_ENTRY(orc_metric_pressure, metric_pressure(X))
ORC_ENTRY(orc_rk4_step, rk4_step(X))
ORC_ENTRY(orc_prepare_projection_boundaries, prepare_projection_boundaries(X))
ORC_ENTRY(orc_pressure_solve, pressure_solve_ref(X))
ORC_ENTRY(orc_damp_pdyn, damp_pdyn(X))
ORC_ENTRY(orc_rebuild_pstat, rebuild_pstat(X))
ORC_ENTRY(orc_precipitation, precipitation(X))
ORC_ENTRY(orc_rk_scalar_sync, rk_scalar_sync(X))
ORC_ENTRY(orc_boundary_conditions, boundary_conditions(X))
ORC_ENTRY(orc_clamp_species, clamp_species(X))
ORC_ENTRY(orc_restore_var, restore_var(X))
ORC_ENTRY(orc_turbulence, turbulence_run(X))
ORC_ENTRY(orc_turbulence_init, turbulence_init(X))
ORC_ENTRY(orc_nh4sh_equilibrium, nh4sh_equilibrium(X))
ORC_ENTRY(orc_chemistry, chemistry(X))
ORC_ENTRY(orc_mca, mca(X); mca_sync(X))

extern "C" int orc_saturation_adjust(const atsat_grid* g, const atsat_consts* c, const atsat_knobs* k,
                                     atsat_host_state* s, int gas){
    if(!g || !c || !k || !s || gas < 0 || gas > 2) return -1;
    Ctx X(*g, *c, *k, *s); saturation_adjust(X, gas); return 0;
}

extern "C" int orc_step(const atsat_grid* g, const atsat_consts* c, const atsat_knobs* k,
                        atsat_host_state* s, int iter){
    if(!g || !c || !k || !s) return -1;
    Ctx X(*g, *c, *k, *s); step(X, iter); return 0;
}
