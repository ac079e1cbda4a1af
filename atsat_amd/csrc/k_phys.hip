// Column and cell physics on the device: mixture density, saturation adjustment, hydrostatic
// p_stat rebuild and precipitation.
//
// References: cSaturnModel::computeMixtureDensity (planet/Pressure_Sat.cpp:49-62);
// SaturationAdjustment<cSaturnModel>::run (planet/SaturationAdjustment.h:255-746);
// cSaturnModel::rebuildPstatHydro (planet/InitVariables_Sat.cpp:556-578);
// Precipitation<cSaturnModel>::run / column / sedimentNH4SH / surfaceMap
// (planet/Precipitation.h:191-574).
//
// The saturation adjustment is independent per cell and its three gases touch a cell only
// through that cell's own temperature, so one thread runs CH4 -> H2O -> NH3 for its cell in the
// reference's order (cSaturnModel.cpp:908-927), and the final clamp of each gas (:673-685) is
// applied by the same thread after that gas. Its arithmetic is FP64 in both precisions (the
// Newton iteration near the Clausius-Clapeyron limit), on the FP64 t, p_stat and rho_mix and the
// species widened from their class (FP32 in the FP32 step), rounded back on the store. The precipitation
// arithmetic is FP64 and its outputs are stored in the work precision. exp, log and pow are the
// correctly rounded double-double versions of crmath.hpp, which agree with the host libm of the
// reference in all but a small fraction of cases. The p_stat rebuild is a prefix up each
// column and the precipitation a top-down sweep down each column: one thread per column.
#include "atsat_dev.hpp"
#include "crmath.hpp"

namespace atsat {

// ATPhys.h:49-52, with ln T_K passed in: the liquid and the ice pressure of one temperature share
// it (the same value the formula would compute twice). W: the work type of the calling kernel,
// which picks its transcendentals (crm::Phys)
template<typename W>
__device__ __forceinline__ double svp_ln(double T_K, double lnT, double C, double L0, double R, double del_alf, double del_bet){
    return crm::Phys<W>::exp(C + (-L0 / T_K + del_alf * lnT + del_bet * T_K) / (1e-3 * R));
}
template<typename W>
__device__ __forceinline__ double svp(double T_K, double C, double L0, double R, double del_alf, double del_bet){
    return svp_ln<W>(T_K, crm::Phys<W>::log(T_K), C, L0, R, del_alf, del_bet);
}

__device__ __forceinline__ double rho_at(const double* rho, long long n, double r_mix){  // cSaturnModel.h:514-518
    const double r = rho[n];
    return (r > 0.0 && isfinite(r)) ? r : r_mix;
}

// ------------------------------------------------------------------------------------------
// (both loads are issued before either is used; 128-thread blocks cover a 361-point ring in three)
template<typename T>
__global__ void __launch_bounds__(128) k_mixture_density(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = m * D.fs + D.at(i, j, k);
    const double t = B.s[ATSAT_T][n], p = B.s[ATSAT_P_STAT][n];
    const double Tk = t * S.t_ref;
    double r = 0.0;
    if((Tk > 0.0) && (S.R_mix > 0.0)){
        r = (p * 1.0e5) / (S.R_mix * Tk);
        if(!isfinite(r)) r = 0.0;
    }
    B.s[ATSAT_RHO_MIX][n] = r;
}

template<typename T>
void launch_mixture_density(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_mixture_density<T>, dim3((D.km + 127) / 128, D.jm, D.im * D.M), dim3(128), 0, st, D, S, B);
}

// ------------------------------------------------------------------------------------------
// SaturationAdjustment::run for one cell and one gas, with the model's knob values CONSERVE 1,
// NEWTON 1, EVAP 1, ICE_ENTRY 0, BERGERON 2, FREEZE 1, DRY_EXIT 1, LOCAL_RHO 1, PSTAT_HYDRO 1.
// t_u is the cell's temperature in K (updated in place); c, cl, ic the vapour / cloud / ice.
// Returns true where the reference writes t = t_u / t_ref (the entry melt :376, the entry
// freeze :395 and the write-back :614): that division need not round-trip t, so the caller
// reproduces it exactly where, and only where, the reference performs it.
template<typename W>
__device__ bool satadj_cell(const Scal& S, int gas, double rho_c, double p_u, double& t_u,
                            double& c, double& cl, double& ic){
    const double t_0 = S.t_0[gas], t_00 = S.t_00[gas], ep = S.ep[gas], lv = S.lv[gas], ls = S.ls[gas];
    const double Cc = S.C[gas], L0 = S.L0[gas], R = S.R[gas], del_alf = S.del_alf[gas], del_bet = S.del_bet[gas];
    const double C_i = S.C_i[gas], L0_i = S.L0_i[gas], del_alf_i = S.del_alf_i[gas], del_bet_i = S.del_bet_i[gas];
    const double t_range_inv = 1.0 / (t_0 - t_00);
    const double q_diff_min = 1.0e-3;
    const bool freeze_on = t_00 > 0.0 && t_00 < t_0;
    const double Lf = ls - lv;
    const double cp_mix = S.cp_mix;
    // The FP32 step's species are stored in FP32 and its transcendentals are the library's (a few
    // FP64 ulp): there the repeated divisions -- by the heat capacity of the cell, by p, by
    // 1e-3 R and by T in the saturation pressure -- are products with reciprocals formed once
    // (qrcp / qdiv, an FP64 ulp apart). The FP64 step divides, in the reference's order.
    constexpr bool F = std::is_same<W, float>::value;
    const double rhc = qrcp<W>(cp_mix * rho_c), rpu = qrcp<W>(p_u), rR3 = qrcp<W>(1e-3 * R);
    // the saturation mixing ratio rho ep E / p of the liquid (ice: false) or ice target at T, with
    // lnT = log T and, in the FP32 step, rT = 1 / T
    auto q_sat = [&](double T, double rT, double lnT, bool ice) -> double {
        const double Cx = ice ? C_i : Cc, Lx = ice ? L0_i : L0;
        const double da = ice ? del_alf_i : del_alf, db = ice ? del_bet_i : del_bet;
        if constexpr(F) return rho_c * ep * crm::Phys<W>::exp(Cx + (-Lx * rT + da * lnT + db * T) * rR3) * rpu;
        else return rho_c * ep * svp_ln<W>(T, lnT, Cx, Lx, R, da, db) / p_u;
    };
    bool wrote = false;
    if(!(t_u > 0.0) || !(p_u > 0.0)) return wrote;                            // :358
    if(t_u > t_0){                                                            // :365-385
        if(ic > 0.0){
            double dm = ic;
            if(Lf > 0.0){
                const double hc = cp_mix * rho_c;
                const double cap = hc * (t_u - t_0) / Lf;
                if(dm > cap) dm = cap;
                t_u -= Lf * dm / hc;
                wrote = true;
            }
            ic -= dm;
            cl += dm;
        }
    }
    if(freeze_on && t_u < t_00 && cl > 0.0){                                  // :388-399
        double dm = cl;
        if(Lf > 0.0){
            const double hc = cp_mix * rho_c;
            const double cap = hc * (t_00 - t_u) / Lf;
            if(dm > cap) dm = cap;
            t_u += Lf * dm / hc;
            wrote = true;
        }
        cl -= dm;
        ic += dm;
    }
    if(c < 0.0) c = 0.0;                                                      // :400-402
    if(cl < 0.0) cl = 0.0;
    if(ic < 0.0) ic = 0.0;
    // Without condensate the cell ends at :422 or :436-448 exactly when c <= q_entry, the
    // saturation value of the phase of t_u. The same expression with the library's log and exp
    // (a few ulp each) has an exponent within ex of the correctly rounded one (the log's error
    // through del_alf, one rounding of each later sum and of the quotient); when ex < 1e-10,
    // a cell whose vapour is below that value by more than 1e-9 relative takes the end. (The
    // FP32 step takes the library's functions below anyway, so there the test below decides.)
    if(!F && cl <= 0.0 && ic <= 0.0){
        const bool liq = !(t_u < t_0);
        const double a1 = -(liq ? L0 : L0_i) / t_u, a2 = (liq ? del_alf : del_alf_i) * std::log(t_u);
        const double a3 = (liq ? del_bet : del_bet_i) * t_u;
        const double x = (liq ? Cc : C_i) + (a1 + a2 + a3) / (1e-3 * R);
        const double ex = 0x1p-50 * ((fabs(a1) + 2.0 * fabs(a2) + fabs(a3)) / fabs(1e-3 * R) + fabs(x));
        if(ex < 1.0e-10){
            const double qa = rho_c * ep * std::exp(x) / p_u;
            if(isfinite(qa) && c <= qa * (1.0 - 1.0e-9)) return wrote;
        }
    }
    const double ln_tu = crm::Phys<W>::log(t_u), rtu = F ? 1.0 / t_u : 0.0;
    const double q_Rain_0 = q_sat(t_u, rtu, ln_tu, false);                    // :416-418
    if(q_Rain_0 <= 0.0) return wrote;                                         // :422
    // the ice target at t_u, needed by the entry test below in either of two places
    auto q_ice_tu = [&](){ return q_sat(t_u, rtu, ln_tu, true); };
    double q_entry = q_Rain_0;
    if(cl <= 0.0 && ic <= 0.0 && t_u < t_0) q_entry = q_ice_tu();             // :433-434
    if(c <= q_entry){                                                         // :436-448
        const double q_c0 = cl, q_i0 = ic;
        if(q_c0 <= 0.0 && q_i0 <= 0.0) return wrote;
        const double q_Ice_0 = q_ice_tu();
        double q_t0;
        if(q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
        else if(q_i0 <= 0.0) q_t0 = q_Rain_0;
        else q_t0 = q_Ice_0;
        if(fabs(c - q_t0) <= q_diff_min * q_t0) return wrote;
    }
    double q_v_b = c, q_c_b = cl, q_i_b = ic;                                 // :451-463
    double Tt = t_u;
    double q_v_hyp = q_v_b;
    double q_rain_now = q_Rain_0;
    for(int itr = 1; itr <= 15; itr++){                                       // :465-590
        double CND = (Tt - t_00) * t_range_inv;
        double DEP = (t_0 - Tt) * t_range_inv;
        if(Tt <= t_00){ CND = 0.0; DEP = 1.0; }
        if(Tt >= t_0){ CND = 1.0; DEP = 0.0; }
        const bool berg = Tt < t_0 && q_v_b <= q_rain_now && (q_i_b > 0.0 || q_c_b <= 0.0);
        if(berg){
            CND = 0.0; DEP = 1.0;
            if(q_c_b > 0.0){
                Tt -= qdiv<W>(lv * q_c_b, cp_mix * rho_c, rhc);
                q_v_b += q_c_b;
                q_c_b = 0.0;
            }
        }
        double d_q_v = q_v_hyp - q_v_b;
        double d_q_c = -d_q_v * CND;
        double d_q_i = -d_q_v * DEP;
        if(d_q_v < -q_v_b){ d_q_v = -q_v_b; d_q_c = -d_q_v * CND; d_q_i = -d_q_v * DEP; }
        if(q_c_b + d_q_c < 0.0){ const double ex = -(q_c_b + d_q_c); d_q_c = -q_c_b; d_q_i -= ex; }
        if(q_i_b + d_q_i < 0.0){ const double ex = -(q_i_b + d_q_i); d_q_i = -q_i_b; d_q_c -= ex;
                                 if(q_c_b + d_q_c < 0.0) d_q_c = -q_c_b; }
        d_q_v = -(d_q_c + d_q_i);
        Tt += qdiv<W>(lv * d_q_c + ls * d_q_i, cp_mix * rho_c, rhc);
        q_v_b += d_q_v;
        q_c_b += d_q_c;
        q_i_b += d_q_i;
        if(q_v_b < 0.0) q_v_b = 0.0;
        if(q_c_b < 0.0) q_c_b = 0.0;
        if(q_i_b < 0.0) q_i_b = 0.0;
        // the liquid target is always needed (the next pass's Bergeron test); the ice target
        // only where the hypothetical vapour uses it
        const double lnT = crm::Phys<W>::log(Tt), rT = F ? 1.0 / Tt : 0.0;
        const double q_Rain = q_sat(Tt, rT, lnT, false);
        q_rain_now = q_Rain;
        const bool mixed = q_c_b > 0.0 && q_i_b > 0.0;
        const bool dry_cold = q_c_b <= 0.0 && q_i_b <= 0.0 && Tt < t_0;
        const bool liquid_only = !mixed && !dry_cold && q_i_b == 0.0;
        const double q_Ice = liquid_only ? 0.0 : q_sat(Tt, rT, lnT, true);
        if(mixed) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
        else if(dry_cold) q_v_hyp = q_Ice;
        else if(liquid_only) q_v_hyp = q_Rain;
        else q_v_hyp = q_Ice;
        if(Tt >= t_0){
            if(q_i_b > 0.0){
                double dm = q_i_b;
                if(Lf > 0.0){
                    const double hc = cp_mix * rho_c;
                    const double cap = hc * (Tt - t_0) / Lf;
                    if(dm > cap) dm = cap;
                    Tt -= Lf * dm / hc;
                }
                q_i_b -= dm;
                q_c_b += dm;
            }
        }
        const double q_diff = fabs(q_v_b - q_v_hyp) / (q_v_hyp + 1e-20);
        if(q_diff <= q_diff_min) break;
        if(q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;          // DRY_EXIT
        const double dT = 0.1;                                                // NEWTON, :562-586
        // the same selection at Tt + dT, each target evaluated only where it is used
        const double Td = Tt + dT, lnTd = crm::Phys<W>::log(Td), rTd = F ? 1.0 / Td : 0.0;
        const bool mixed_d = q_c_b > 0.0 && q_i_b > 0.0;
        const bool dry_cold_d = q_c_b <= 0.0 && q_i_b <= 0.0 && Tt < t_0;
        const bool liquid_d = !mixed_d && !dry_cold_d && q_i_b == 0.0;
        const double q_Rain_d = (mixed_d || liquid_d) ? q_sat(Td, rTd, lnTd, false) : 0.0;
        const double q_Ice_d = liquid_d ? 0.0 : q_sat(Td, rTd, lnTd, true);
        double q_hyp_d;
        if(mixed_d) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
        else if(dry_cold_d) q_hyp_d = q_Ice_d;
        else if(liquid_d) q_hyp_d = q_Rain_d;
        else q_hyp_d = q_Ice_d;
        double CNDn = (Tt - t_00) * t_range_inv;
        if(Tt <= t_00) CNDn = 0.0;
        if(Tt >= t_0) CNDn = 1.0;
        if(berg) CNDn = 0.0;
        const double Gf = F ? (lv * CNDn + ls * (1.0 - CNDn)) * rhc * (q_hyp_d - q_v_hyp) * (1.0 / dT)
                            : (lv * CNDn + ls * (1.0 - CNDn)) / (cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
        double wgt = 1.0 / (1.0 + (Gf > 0.0 ? Gf : 0.0));
        if(wgt < 0.05) wgt = 0.05;
        q_v_hyp = q_v_b + wgt * (q_v_hyp - q_v_b);
    }
    c = q_v_b; cl = q_c_b; ic = q_i_b;                                        // :611-614
    t_u = Tt;
    return true;
}

template<typename T>
__global__ void __launch_bounds__(256) k_satadj(Dims D, Scal S, Bufs<T> B, int g0, int g1){
    // one thread per cell of an (i, member) slab, (j, k) flattened so that no lane idles at the
    // end of a 361-long row
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if(x >= D.jm * D.km) return;
    const int j = x / D.km, k = x - j * D.km;
    const int im_m = blockIdx.y;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = m * D.fs + D.at(i, j, k);
    const double rho_c = rho_at(B.s[ATSAT_RHO_MIX], n, S.r_mix);
    const double p_u = B.s[ATSAT_P_STAT][n];
    double* tp = B.s[ATSAT_T] + n;
    double t_nd = *tp;                                            // t in the model's unit
    bool t_written = false;
    const int gf[3][3] = { { ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE },
                           { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE },
                           { ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE } };
    for(int gas = g0; gas <= g1; gas++){
        T* cq = B.w[gf[gas][0]] + n; T* cl = B.w[gf[gas][1]] + n; T* ic = B.w[gf[gas][2]] + n;
        double c = (double)*cq, cld = (double)*cl, ice = (double)*ic;
        double t_u = t_nd * S.t_ref;                               // :347
        if(satadj_cell<T>(S, gas, rho_c, p_u, t_u, c, cld, ice)){ t_nd = t_u / S.t_ref; t_written = true; }
        if(c <= 0.0) c = 0.0;                                     // final clamp, :679-681
        if(cld <= 0.0) cld = 0.0;
        if(ice <= 0.0) ice = 0.0;
        *cq = (T)c; *cl = (T)cld; *ic = (T)ice;
    }
    if(t_written) *tp = t_nd;
}

template<typename T>
void launch_satadj(const Dims& D, const Scal& S, const Bufs<T>& B, int g0, int g1, hipStream_t st){
    hipLaunchKernelGGL(k_satadj<T>, dim3((D.jm * D.km + 255) / 256, D.im * D.M), dim3(256), 0, st, D, S, B, g0, g1);
}

// ------------------------------------------------------------------------------------------
// rebuildPstatHydro, InitVariables_Sat.cpp:556-578: p integrated up each column, exact for T
// linear in z between two levels. FP64 in both precisions.
template<typename T>
__global__ void __launch_bounds__(256) k_rebuild_pstat(Dims D, Scal S, Bufs<T> B){
    const int x = blockIdx.x * blockDim.x + threadIdx.x;          // the columns (j, k) flattened
    if(x >= D.jm * D.km) return;
    const int j = x / D.km, k = x - j * D.km, m = blockIdx.y;
    const long long mo = m * D.fs;
    const double* t = B.s[ATSAT_T] + mo;
    double* ps = B.s[ATSAT_P_STAT] + mo;
    const double ep_ = S.exp_pressure;
    double p = S.p_ref * crm::Phys<T>::pow(t[D.at(0, j, k)], ep_);
    ps[D.at(0, j, k)] = p;
    double T0 = t[D.at(0, j, k)] * S.t_ref;
    for(int i = 1; i < D.im; i++){
        const double t1 = t[D.at(i, j, k)];
        const double T1 = t1 * S.t_ref;
        const double a = S.a_hydro;
        if(T0 > 0.0 && T1 > 0.0){
            const double dT = T0 - T1;
            p *= (fabs(dT) > 1.0e-9 * T0) ? crm::Phys<T>::pow(T1 / T0, a / dT) : crm::Phys<T>::exp(-a / T0);
        } else p = S.p_ref * crm::Phys<T>::pow(t1 > 0.0 ? t1 : 0.0, ep_);
        ps[D.at(i, j, k)] = p;
        T0 = T1;
    }
}

template<typename T>
void launch_rebuild_pstat(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_rebuild_pstat<T>, dim3((D.jm * D.km + 255) / 256, D.M), dim3(256), 0, st, D, S, B);
}

// ------------------------------------------------------------------------------------------
// Precipitation::run: one thread per (member, j, k) column does the three species top-down
// (Precipitation.h:273-445, coupling on), the NH4SH settling flux (:542-574) and the surface
// maps (:467-492). The reference runs the species one after another down the column; they
// interact only through Q_precip, which each adds to in the reference's order (H2O, NH3, CH4)
// starting from the zeroing of :199-209. Here the thread walks the levels once and does the three
// species at each level, every value the same expression as species by species: the level's t,
// p_stat and rho_mix are loaded once for all of them, its nine species values together, Q_precip
// is summed in registers in the same order and rounding (stored once), and the settling flux rides
// along. The fluxes a species carries to the next level down are the stored (rounded) values, as
// the reference reads them back (:337-339); the surface maps take the level-0 ones.
template<typename T>
__global__ void __launch_bounds__(128) k_precipitation(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    const long long mo = m * D.fs;
    const int im = D.im;
    const double c_c_au = 2.0e-8, c_i_au = 5.0e-8, q_c_crit = 5.0e-4, c_ac = 1.0e-7;   // :93-98
    const double c_rim = 1.0e-8, z_csg = 2.5e-8, a_ev = 2.76e-3;
    const double c_melt_K = 0.2, c_frz_K = 0.2, P_ref = 1.0e-4, ratio_cap = 5.0, v_fall = 2.0;
    const double rs = 0.0045;                                     // precip_rate_scale, cSaturnModel.h:408
    const double P_max_flux = 1.0e-4, dep_frac = 1.0;
    const double dz = S.dz_m;
    const double nd = (S.L_atm * 1.0e3) / S.u_0;
    const double tau = dz / v_fall;
    // the FP32 step's products with reciprocals formed once for the divisions by P_ref, dz and tau
    // (qdiv; its outputs are FP32); the FP64 step divides, in the reference's order
    const double rPref = qrcp<T>(P_ref), rdz = qrcp<T>(dz), rtau = qrcp<T>(tau);
    const double* t = B.s[ATSAT_T] + mo;
    const double* ps = B.s[ATSAT_P_STAT] + mo;
    const double* rho = B.s[ATSAT_RHO_MIX] + mo;
    T* Qp = B.w[ATSAT_Q_PRECIP] + mo;
    const int gas_order[3] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };
    const int fld[3][3] = { { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE },
                            { ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE },
                            { ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE } };
    const int Pf[3][3] = { { ATSAT_P_RAIN, ATSAT_P_SNOW, ATSAT_P_GRAUPEL },
                           { ATSAT_P_NH3_RAIN, ATSAT_P_NH3_SNOW, ATSAT_P_NH3_GRAUPEL },
                           { ATSAT_P_CH4_RAIN, ATSAT_P_CH4_SNOW, ATSAT_P_CH4_GRAUPEL } };
    const int Sf[3][3] = { { ATSAT_S_PRECIP_H2O, ATSAT_S_PRECIP_H2O_CLOUD, ATSAT_S_PRECIP_H2O_ICE },
                           { ATSAT_S_PRECIP_NH3, ATSAT_S_PRECIP_NH3_CLOUD, ATSAT_S_PRECIP_NH3_ICE },
                           { ATSAT_S_PRECIP_CH4, ATSAT_S_PRECIP_CH4_CLOUD, ATSAT_S_PRECIP_CH4_ICE } };
    const T* q4 = B.w[ATSAT_NH4SH] + mo;
    T* P4 = B.w[ATSAT_P_NH4SH] + mo;
    // NH4SH settling (:542-574) of one cell
    auto settle = [&](long long n, double Tk){
        double q = (double)q4[n]; q = q > 0.0 ? q : 0.0;
        P4[n] = (T)((Tk >= S.t_00_nh4sh && Tk <= S.t_0_nh4sh) ? q * S.v_settle : 0.0);
    };
    {
        const long long ntop = D.at(im - 1, j, k);
#pragma unroll
        for(int q = 0; q < 3; q++){
#pragma unroll
            for(int x = 0; x < 3; x++){ B.w[Pf[q][x]][mo + ntop] = (T)0.0; B.w[Sf[q][x]][mo + ntop] = (T)0.0; }
        }
        Qp[ntop] = (T)0.0;
        settle(ntop, t[ntop] * S.t_ref);
    }
    // per species: the rain, snow and graupel fluxes into the level (held in registers, picked by
    // species with selects so that the species' code exists once)
    double Fr0 = 0.0, Fs0 = 0.0, Fg0 = 0.0, Fr1 = 0.0, Fs1 = 0.0, Fg1 = 0.0, Fr2 = 0.0, Fs2 = 0.0, Fg2 = 0.0;
    auto pick3 = [](int q, double a, double b, double c){ return q == 0 ? a : (q == 1 ? b : c); };
    for(int i = im - 2; i >= 0; i--){
        const long long n = D.at(i, j, k);
        const double Tk = t[n] * S.t_ref;
        const double p_u = ps[n];
        const double qv00 = (double)B.w[fld[0][0]][mo + n], qc00 = (double)B.w[fld[0][1]][mo + n], qi00 = (double)B.w[fld[0][2]][mo + n];
        const double qv01 = (double)B.w[fld[1][0]][mo + n], qc01 = (double)B.w[fld[1][1]][mo + n], qi01 = (double)B.w[fld[1][2]][mo + n];
        const double qv02 = (double)B.w[fld[2][0]][mo + n], qc02 = (double)B.w[fld[2][1]][mo + n], qi02 = (double)B.w[fld[2][2]][mo + n];
        T Qn = (T)0.0;
#pragma unroll 1
        for(int q = 0; q < 3; q++){
            const int gas = gas_order[q];
            const double Lf = S.ls[gas] - S.lv[gas];
            const double t_frz = S.t_0[gas], t_low = S.t_00[gas];
            const double F_r_in = pick3(q, Fr0, Fr1, Fr2), F_s_in = pick3(q, Fs0, Fs1, Fs2), F_g_in = pick3(q, Fg0, Fg1, Fg2);
            double q_v = pick3(q, qv00, qv01, qv02); q_v = q_v > 0.0 ? q_v : 0.0;
            double q_c = pick3(q, qc00, qc01, qc02); q_c = q_c > 0.0 ? q_c : 0.0;
            double q_i = pick3(q, qi00, qi01, qi02); q_i = q_i > 0.0 ? q_i : 0.0;
            const double Rain = fmin(ratio_cap, qdiv<T>(F_r_in, P_ref, rPref));
            const double Snow = fmin(ratio_cap, qdiv<T>(F_s_in, P_ref, rPref));
            const double Graupel = fmin(ratio_cap, qdiv<T>(F_g_in, P_ref, rPref));
            const bool warm = (Tk >= t_frz);
            // q_sat enters only the rain evaporation, which is zero unless a warm cell receives
            // rain (S_ev_raw = 0 there, so F_ev = min(0, F_r_avail >= 0, F_ev_sat >= 0) = 0 for
            // any q_sat); Rain^(7/9) and Rain^(4/9) enter only the warm-cell accretion and
            // evaporation. All three are formed only where they are used.
            double q_sat = q_v, Rain_79 = 0.0, Rain_49 = 0.0;
            if(warm && Rain > 0.0){
                const double E_r = svp<T>(Tk, S.C[gas], S.L0[gas], S.R[gas], S.del_alf[gas], S.del_bet[gas]);
                q_sat = (p_u > E_r) ? rho_at(rho, n, S.r_mix) * S.ep[gas] * E_r / p_u : q_v;
                crm::Phys<T>::pow2(Rain, 7.0 / 9.0, 4.0 / 9.0, Rain_79, Rain_49);
            }
            const double S_c_au = (warm && q_c > q_c_crit) ? (rs * c_c_au) * (q_c - q_c_crit) : 0.0;
            const double S_i_au = (!warm && Tk >= t_low) ? (rs * c_i_au) * q_i : 0.0;
            const double S_ac = (warm) ? (rs * c_ac) * q_c * Rain_79 : 0.0;
            const double S_s_rim = (!warm) ? (rs * c_rim) * q_c * Snow : 0.0;
            const double S_g_rim = (!warm) ? (rs * c_rim) * q_c * Graupel : 0.0;
            const double S_csg = (!warm && q_c > q_c_crit) ? (rs * z_csg) * q_c * Snow : 0.0;
            const double f_melt = warm ? fmin(1.0, c_melt_K * (Tk - t_frz)) : 0.0;
            const double f_frz = warm ? 0.0 : fmin(1.0, c_frz_K * (t_frz - Tk));
            const double F_s_melt = f_melt * F_s_in, F_g_melt = f_melt * F_g_in, F_r_frz = f_frz * F_r_in;
            const double F_r_avail = fmax(0.0, F_r_in - F_r_frz + F_s_melt + F_g_melt + (S_c_au + S_ac) * dz);
            const double S_ev_raw = (warm && Rain > 0.0 && q_v < q_sat) ? a_ev * (q_sat - q_v) * Rain_49 : 0.0;
            const double F_ev_sat = fmax(0.0, q_sat - q_v) * v_fall;
            const double F_ev = fmin(fmin(S_ev_raw * dz, F_r_avail), F_ev_sat);
            const double S_ev = qdiv<T>(F_ev, dz, rdz);
            const double dQ = Lf * (S_s_rim + S_g_rim + S_csg) + qdiv<T>(Lf * (F_r_frz - F_s_melt - F_g_melt), dz, rdz)
                            - qdiv<T>(S.lv[gas] * F_ev, dz, rdz);
            Qn = (q == 0) ? (T)(0.0 + dQ) : (T)((double)Qn + dQ);
            const double d_c = fmin(dep_frac * q_c, (S_c_au + S_ac + S_s_rim + S_g_rim + S_csg) * tau);
            const double d_i = fmin(dep_frac * q_i, S_i_au * tau);
            B.w[Sf[q][1]][mo + n] = (T)(qdiv<T>(- fmax(0.0, d_c), tau, rtau) * nd);
            B.w[Sf[q][2]][mo + n] = (T)(qdiv<T>(- fmax(0.0, d_i), tau, rtau) * nd);
            B.w[Sf[q][0]][mo + n] = (T)(+ S_ev * nd);
            const double F_r = F_r_in - F_r_frz + F_s_melt + F_g_melt - F_ev + (S_c_au + S_ac) * dz;
            const double F_s = F_s_in - F_s_melt + (S_i_au + S_s_rim - S_csg) * dz;
            const double F_g = F_g_in - F_g_melt + F_r_frz + (S_g_rim + S_csg) * dz;
            const double nr = fmin(P_max_flux, fmax(0.0, F_r));
            const double ns = fmin(P_max_flux, fmax(0.0, F_s));
            const double ng = fmin(P_max_flux, fmax(0.0, F_g));
            B.w[Pf[q][0]][mo + n] = (T)nr; B.w[Pf[q][1]][mo + n] = (T)ns; B.w[Pf[q][2]][mo + n] = (T)ng;
            // the next level down reads the stored flux (:337-339)
            const double sr_ = (double)(T)nr, ss_ = (double)(T)ns, sg_ = (double)(T)ng;
            if(q == 0){ Fr0 = sr_; Fs0 = ss_; Fg0 = sg_; }
            else if(q == 1){ Fr1 = sr_; Fs1 = ss_; Fg1 = sg_; }
            else { Fr2 = sr_; Fs2 = ss_; Fg2 = sg_; }
        }
        Qp[n] = Qn;
        settle(n, Tk);
    }
    // surface maps at i_base = 0 (:467-492), from the level-0 fluxes as stored
    {
        const long long n = D.at(0, j, k);
        const double h2o_srf = Fr0 + Fs0 + Fg0;
        const double nh3_srf = Fr1 + Fs1 + Fg1;
        const double ch4_srf = Fr2 + Fs2 + Fg2;
        const double nh4sh_srf = (double)P4[n];
        const size_t plane = (size_t)D.jm * D.km;
        double* srf = B.srf + (size_t)m * 5 * plane + (size_t)j * D.km + k;
        srf[0] = h2o_srf; srf[plane] = nh3_srf; srf[2 * plane] = ch4_srf; srf[3 * plane] = nh4sh_srf;
        srf[4 * plane] = h2o_srf + nh3_srf + ch4_srf + nh4sh_srf;
    }
}

template<typename T>
void launch_precipitation(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_precipitation<T>, dim3((D.km + 127) / 128, D.jm, D.M), dim3(128), 0, st, D, S, B);
}

#define INST(T) \
    template void launch_mixture_density<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_satadj<T>(const Dims&, const Scal&, const Bufs<T>&, int, int, hipStream_t); \
    template void launch_rebuild_pstat<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_precipitation<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
