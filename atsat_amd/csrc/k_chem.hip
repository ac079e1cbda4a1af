// The chemistry block of the physics iteration on the device (planet/ChemistrySat.h,
// planet/FluxLimiter.h; called at planet/cSaturnModel.cpp:933 and 979-994), at the model's
// defaults: NH4SH as an equilibrium adjustment (ATSAT_NH4SH_EQ 1, so the rate law's w_* are
// zero), difflux = D laplacian(c) (ATSAT_CHEM_DIFFLUX_LAPLACIAN 1) added to massflux
// (ATSAT_CHEM_DIFFLUX_PLUS 1), no reaction enthalpy term (ATSAT_CHEM_ENTHALPY 0), no solid cells,
// the periodic seam, the polar divisor floor 0.4 (ATPhys.h:104-108).
//
//   k_nh4sh_eq       EquilibriumNH4SH("loop") (:294-331, equilibrateNH4SHCell :250-289): one
//                    thread per cell; the bisection on the formed amount x with K_p(T) =
//                    10^(14.82 - 4705/T) atm^2 evaluated by the correctly rounded pow;
//   k_diff_mass_flux DiffMassFluxSat (:365-472): difflux_* = D laplacian(c) (:645-662) and
//                    thermalmassflux from the cell's fluxes j_*, jT_* (pass 1, :393-425), one
//                    thread per interior cell (the fluxes are not stored: with the Laplacian
//                    form only the cell's own thermalmassflux reads them);
//   k_chem_massflux  ChemMassRateSat (:46-205: massflux_x = 0 + difflux_x) and
//                    FluxLimiter<cSaturnModel>::nh4sh (FluxLimiter.h:90-190, Superbee);
//   k_shapiro_kj,    AtomUtils::damp_wiggles(f, nullptr, true, true, true) (lib/Utils.cpp:229-325)
//   k_shapiro_i      of four fields: the passes along k (periodic) and j (clamped) fused in
//                    registers into a scratch set, then along i (clamped) back.
//
// Every formula is evaluated in FP64 in the reference's operation order, from t, w and rho_mix
// (state class, FP64) and the species, u and v (work class, widened on the load); the species
// the equilibrium changes are rounded back to their class on the store, and difflux_*,
// thermalmassflux, massflux_* and fluxlim_nh4sh are work-class inputs of the RK4 stages, rounded
// on the store.
#include "atsat_dev.hpp"
#include "crmath.hpp"

namespace atsat {

namespace {
__device__ __forceinline__ double smax(double a, double b){ return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b){ return (b < a) ? b : a; }
constexpr double kPolarFloor = 0.4;
}

// ---- EquilibriumNH4SH ----------------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_nh4sh_eq(Dims D, Scal S, Bufs<T> B){
    const int jk = blockIdx.x * blockDim.x + threadIdx.x;         // (j, k) flattened per slab
    if(jk >= D.jm * D.km) return;
    const int j = jk / D.km, k = jk - j * D.km, im_m = blockIdx.y;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = (long long)m * D.fs + D.at(i, j, k);
    const double R_univ = 8314.46, atm = 101325.0;
    auto Kp = [&](double TK){ return crm::Phys<T>::pow(10.0, 14.82 - 4705.0 / TK) * atm * atm; };
    double* t = B.s[ATSAT_T]; T* nh3 = B.w[ATSAT_NH3]; T* h2s = B.w[ATSAT_H2S];
    T* nh4sh = B.w[ATSAT_NH4SH];
    // the cell's five values, loaded together before the first test
    const double t_n = t[n], rho_n = B.s[ATSAT_RHO_MIX][n];
    const double nh3_n = (double)nh3[n], h2s_n = (double)h2s[n], nh4sh_n = (double)nh4sh[n];
    const double T0 = t_n * S.t_ref;
    if(!(T0 > 0.0)) return;
    double rho = rho_n;
    if(!(rho > 0.0) || !isfinite(rho)) rho = S.r_mix;
    const double a = smax(0.0, nh3_n) / S.m_nh3;
    const double b = smax(0.0, h2s_n) / S.m_h2s;
    const double s = smax(0.0, nh4sh_n) / S.m_nh4sh;
    // no NH3 x H2S product and no solid: f(0) = -K_p < 0 and nothing to decompose, the
    // reference's early return, decided without K_p
    if(s == 0.0 && (a == 0.0 || b == 0.0)) return;
    const double heat = S.nh4sh_Lf * S.m_nh4sh / (S.cp_mix * rho);
    auto f = [&](double x){
        const double TK = T0 + heat * x, RT = R_univ * TK;
        return (a - x) * (b - x) * RT * RT - Kp(TK); };
    const double f0 = f(0.0);
    double lo, hi;
    if(f0 > 0.0){ lo = 0.0; hi = smin(a, b); }
    else if(s > 0.0){
        if(f(-s) <= 0.0){ lo = hi = -s; }
        else { lo = -s; hi = 0.0; }
    } else return;
    if(hi > lo) for(int it = 0; it < 80; it++){
        const double xm = 0.5 * (lo + hi);
        if(f(xm) > 0.0) lo = xm; else hi = xm;
        if(hi - lo <= 1.0e-15 * smax(a, smax(b, s))) break;
    }
    const double x = (f0 > 0.0) ? lo : hi;
    if(x == 0.0) return;
    nh3[n] = (T)smax(0.0, nh3_n - S.m_nh3 * x);
    h2s[n] = (T)smax(0.0, h2s_n - S.m_h2s * x);
    nh4sh[n] = (T)smax(0.0, nh4sh_n + S.m_nh4sh * x);
    const double dT = heat * x;
    t[n] = (T0 + dT) / S.t_ref;
}

// ---- DiffMassFluxSat -----------------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(kRowThreads) k_diff_mass_flux(Dims D, Scal S, Bufs<T> B, RowMap RM){
    const int k = threadIdx.x;
    int jr, lev, m;
    if(!RM.decode((int)blockIdx.x, 0, jr, lev, m)) return;
    const int i = 1 + lev, j = 1 + jr;
    if(k >= D.km - 1) return;
    const long long mo = (long long)m * D.fs;
    const long long c = mo + D.at(i, j, k);
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    const long long ip = c + sr, imn = c - sr, jp = c + st, jmn = c - st;
    const long long pe = mo + D.at(i, j, phi_wrap(k + 1, D.km)), pw = mo + D.at(i, j, phi_wrap(k - 1, D.km));
    const double rm = B.rm[i];
    const double sinthe = B.thez_sin[j];
    const double rmsinthe = rm * sinthe;
    const double costhe = B.geo[(long long)j * NGEO + G_COSTHE];
    const double* t = B.s[ATSAT_T];
    // the denominators of the row, and (FP32 step) their reciprocals
    const double rms = rm * smax(sinthe, kPolarFloor);
    const double dr2 = 2.0 * S.dr, dthe2 = 2.0 * S.dthe, dphi2 = 2.0 * S.dphi;
    const double drs = S.dr * S.dr, dthes = S.dthe * S.dthe, dphis = S.dphi * S.dphi;
    const double r_dr2 = qrcp<T>(dr2), r_dthe2 = qrcp<T>(dthe2), r_dphi2 = qrcp<T>(dphi2);
    const double r_drs = qrcp<T>(drs), r_dthes = qrcp<T>(dthes), r_dphis = qrcp<T>(dphis);
    const double r_rm = qrcp<T>(rm), r_rm2 = qrcp<T>(rm * rm), r_rmrms = qrcp<T>(rm * rms), r_rms2 = qrcp<T>(rms * rms);
    const double r_rmsinthe = qrcp<T>(rmsinthe);
    // the four stencils, loaded before any is used: in the FP64 step the scheduler is kept from
    // sinking them towards their uses (one wait for the wave instead of one per field: 1.90 ->
    // 1.32 ms for 16 native members); the FP32 step's shorter arithmetic keeps five waves per SIMD
    // only with the compiler's own order (0.90 against 1.04 ms)
    const Raw7d Tt = raw7(t, c, ip, imn, jp, jmn, pe, pw);
    const Raw7d H2s = raw7(B.w[ATSAT_H2S], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d Nh3 = raw7(B.w[ATSAT_NH3], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d Nh4sh = raw7(B.w[ATSAT_NH4SH], c, ip, imn, jp, jmn, pe, pw);
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_sched_barrier(0);
    // derivative_1_order, :555-609, no solid cells: the centred differences
    auto d1 = [&](const Raw7d& F, double& dr_, double& dt_, double& dp_){
        dr_ = qdiv<T>(F.rp - F.rm, dr2, r_dr2);
        dt_ = qdiv<T>(F.tp - F.tm, dthe2, r_dthe2);
        dp_ = qdiv<T>(F.pe - F.pw, dphi2, r_dphi2);
    };
    // laplacian_spherical, :645-662 (coord_stretching false: exp_rm = 1)
    auto lap = [&](const Raw7d& F){
        const double exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;
        const double d2cdr2 = qdiv<T>(F.rp - 2.0 * F.c + F.rm, drs, r_drs) * exp_2_rm;
        const double d2cdthe2 = qdiv<T>(F.tp - 2.0 * F.c + F.tm, dthes, r_dthes);
        const double d2cdphi2 = qdiv<T>(F.pe - 2.0 * F.c + F.pw, dphis, r_dphis);
        const double dcdr = qdiv<T>(F.rp - F.rm, dr2, r_dr2) * exp_rm;
        const double dcdthe = qdiv<T>(F.tp - F.tm, dthe2, r_dthe2);
        return d2cdr2
             + qdiv<T>(2.0, rm, r_rm) * dcdr
             + qdiv<T>(d2cdthe2, rm * rm, r_rm2)
             + qdiv<T>(costhe, rm * rms, r_rmrms) * dcdthe
             + qdiv<T>(d2cdphi2, rms * rms, r_rms2);
    };
    double dtdr, dtdthe, dtdphi, a_r, a_t, a_p, b_r, b_t, b_p, s_r, s_t, s_p;
    d1(Tt, dtdr, dtdthe, dtdphi);
    d1(H2s, b_r, b_t, b_p);
    d1(Nh3, a_r, a_t, a_p);
    d1(Nh4sh, s_r, s_t, s_p);
    // pass 1 at the cell, :401-423 (DT_x = D_x)
    const double tc = Tt.c, r_tc = qrcp<T>(tc);
    const double jT_nh3 = qdiv<T>(S.D_nh3, tc, r_tc) * dtdr;
    const double jT_h2s = qdiv<T>(S.D_h2s, tc, r_tc) * dtdr;
    const double jT_nh4sh = qdiv<T>(S.D_nh4sh, tc, r_tc) * dtdr;
    const double dnh3 = a_r + qdiv<T>(fabs(a_t), rm, r_rm) + qdiv<T>(a_p, rmsinthe, r_rmsinthe);
    const double dh2s = b_r + qdiv<T>(fabs(b_t), rm, r_rm) + qdiv<T>(b_p, rmsinthe, r_rmsinthe);
    const double dnh4sh = s_r + qdiv<T>(fabs(s_t), rm, r_rm) + qdiv<T>(s_p, rmsinthe, r_rmsinthe);
    const double cr = S.c_mix / S.r_mix;
    const double j_nh3 = cr * (S.m_nh3 * S.D_nh3 * dnh3 - jT_nh3);
    const double j_h2s = cr * (S.m_h2s * S.D_h2s * dh2s - jT_h2s);
    const double j_nh4sh = cr * (S.m_nh4sh * S.D_nh4sh * dnh4sh - jT_nh4sh);
    // pass 2, :437-459
    B.w[ATSAT_DIFFLUX_H2S][c] = (T)(S.D_h2s * lap(H2s));
    B.w[ATSAT_DIFFLUX_NH3][c] = (T)(S.D_nh3 * lap(Nh3));
    B.w[ATSAT_DIFFLUX_NH4SH][c] = (T)(S.D_nh4sh * lap(Nh4sh));
    const double w0 = 0.0;                  // the rates w_* (ChemMassRateSat with NH4SH_EQ)
    const double tmf = (j_nh3 * S.cp_nh3
                      + j_h2s * S.cp_h2s
                      + j_nh4sh * S.cp_nh4sh)
                      * (dtdr + qdiv<T>(fabs(dtdthe), rm, r_rm) + qdiv<T>(dtdphi, rmsinthe, r_rmsinthe))
                      + tc * (w0 * S.k_nh3 + w0 * S.k_h2s + w0 * S.k_nh4sh);
    B.w[ATSAT_THERMALMASSFLUX][c] = (T)tmf;
}

// ---- ChemMassRateSat + FluxLimiterNH4SH ----------------------------------------------------
__device__ __forceinline__ double superbee_phi(double r){               // FluxLimiter.h:62-64
    return smax(0.0, smax(smin(2.0 * r, 1.0), smin(r, 2.0)));
}

template<typename T>
__global__ void __launch_bounds__(kRowThreads) k_chem_massflux(Dims D, Scal S, Bufs<T> B, RowMap RM){
    const int k = threadIdx.x;
    int jr, lev, m;
    if(!RM.decode((int)blockIdx.x, 0, jr, lev, m)) return;
    const int i = 1 + lev, j = 1 + jr;
    if(k >= D.km - 1) return;
    const long long mo = (long long)m * D.fs;
    const long long c = mo + D.at(i, j, k);
    // every value the cell reads, loaded before the first store (the stores below could alias
    // them as far as the compiler knows, which would otherwise serialise loads and stores): the
    // diffusive fluxes, NH4SH at the cell, its six neighbours and the upwind values two cells
    // away, and the velocity
    const T* qf = B.w[ATSAT_NH4SH];
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    const double dfx_h2s = (double)B.w[ATSAT_DIFFLUX_H2S][c];
    const double dfx_nh3 = (double)B.w[ATSAT_DIFFLUX_NH3][c];
    const double dfx_nh4sh = (double)B.w[ATSAT_DIFFLUX_NH4SH][c];
    const double q = qf[c];
    const double q_rp = qf[c + sr], q_rm = qf[c - sr], q_tp = qf[c + st], q_tm = qf[c - st];
    const double q_pp = qf[mo + D.at(i, j, phi_wrap(k + 1, D.km))];
    const double q_pm = qf[mo + D.at(i, j, phi_wrap(k - 1, D.km))];
    const double q_r2 = (i + 2 < D.im) ? qf[c + 2 * sr] : q_rp;
    const double q_t2 = (j + 2 < D.jm) ? qf[c + 2 * st] : 0.0;
    const double q_p2 = qf[mo + D.at(i, j, phi_wrap(k + 2, D.km))];
    const double u = (double)B.w[ATSAT_U][c], v = (double)B.w[ATSAT_V][c], w = B.s[ATSAT_W][c];
    __builtin_amdgcn_sched_barrier(0);
    // ChemMassRateSat with NH4SH_EQ, :182-199
    const double dsign = 1.0, w0 = 0.0;
    B.w[ATSAT_MASSFLUX_H2S][c] = (T)(w0 + dsign * dfx_h2s);
    B.w[ATSAT_MASSFLUX_NH3][c] = (T)(w0 + dsign * dfx_nh3);
    B.w[ATSAT_MASSFLUX_NH4SH][c] = (T)(w0 + dsign * dfx_nh4sh);
    // FluxLimiter<cSaturnModel>::nh4sh, FluxLimiter.h:110-185
    const double eps = 1.0e-12, c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const double rm = B.rm[i];
    const double sinthe = smax(kPolarFloor, fabs(B.thez_sin[j]));
    const double inv_rm = 1.0 / rm;
    const double inv_rmsinthe = 1.0 / (rm * sinthe);
    double corr = 0.0;
    {
        const double df = q_rp - q, db = q - q_rm;
        const double denom = df + (df >= 0.0 ? eps : -eps);
        const double rr = (u >= 0.0) ? db / denom : (q_r2 - q_rp) / denom;
        corr += qdiv<T>((1.0 - superbee_phi(rr)) * fabs(u) * (df - db), 2.0 * S.dr, qrcp<T>(2.0 * S.dr));
    }
    {
        const double df = q_tp - q, db = q - q_tm;
        const double denom = df + (df >= 0.0 ? eps : -eps);
        const double q_far = (j + 2 < D.jm) ? q_t2 : (c43 * q_tp - c13 * q);
        const double rr = (v >= 0.0) ? db / denom : (q_far - q_tp) / denom;
        corr += qdiv<T>((1.0 - superbee_phi(rr)) * fabs(v) * inv_rm * (df - db), 2.0 * S.dthe, qrcp<T>(2.0 * S.dthe));
    }
    {
        const double df = q_pp - q, db = q - q_pm;
        const double denom = df + (df >= 0.0 ? eps : -eps);
        const double rr = (w >= 0.0) ? db / denom : (q_p2 - q_pp) / denom;
        corr += qdiv<T>((1.0 - superbee_phi(rr)) * fabs(w) * inv_rmsinthe * (df - db), 2.0 * S.dphi, qrcp<T>(2.0 * S.dphi));
    }
    B.w[ATSAT_FLUXLIM_NH4SH][c] = (T)corr;
}

// ---- damp_wiggles(f, nullptr, true, true, true) of four work-class fields ------------------
template<typename T>
struct Four { T* f[4]; };

// along i, clamped, from the scratch back into f (:297-316): one thread per (j, k) of one field
// marches up the column with the levels i-1, i, i+1 in registers, each value loaded once. The
// levels are loaded in groups of kShI, the next group's loads issued before the current group's
// stores (the compiler cannot move a load above a store to another buffer it may alias).
constexpr int kShI = 8;
template<typename T>
__global__ void __launch_bounds__(1024) k_shapiro_i(Dims D, Four<T> F, Four<T> tmp){
    const int k = threadIdx.x;
    const int j = blockIdx.x, m = blockIdx.y, q = blockIdx.z;
    if(k >= D.km) return;
    const long long mo = (long long)m * D.fs;
    const T* p = tmp.f[q] + mo;
    T* out = F.f[q] + mo;
    const double coeff = 1.0 * 0.25;
    const int im = D.im;
    auto lev = [&](int l) -> T { return p[D.at(l < im ? l : im - 1, j, k)]; };
    T a = lev(0), b = a;
    T cur[kShI];
#pragma unroll
    for(int g = 0; g < kShI; g++) cur[g] = lev(1 + g);           // c of levels 0 .. kShI-1
    for(int i0 = 0; i0 < im; i0 += kShI){
        T nxt[kShI];
#pragma unroll
        for(int g = 0; g < kShI; g++) nxt[g] = lev(i0 + kShI + 1 + g);
#pragma unroll
        for(int g = 0; g < kShI; g++){
            const int i = i0 + g;
            if(i >= im) break;
            const T c = cur[g];
            const double f0 = (double)b;
            out[D.at(i, j, k)] = (T)(f0 + coeff * ((double)a - 2.0 * f0 + (double)c));
            a = b; b = c;
        }
#pragma unroll
        for(int g = 0; g < kShI; g++) cur[g] = nxt[g];
    }
}

// The passes along k (periodic, the seam column k = km-1 taking the filtered k = 0,
// lib/Utils.cpp:251-271) and along j (clamped at the poles, :275-293) fused: one thread per k
// of one level and field marches down a band of kKjRows rows, holding the k-filtered values of
// rows j-1, j, j+1 in registers (each rounded to T, as the intermediate array is) and writing the
// j-filtered row j into tmp. The k-filtered row of a band edge is formed by both bands.
constexpr int kKjRows = 16;
template<typename T>
__global__ void __launch_bounds__(1024) k_shapiro_kj(Dims D, Four<T> F, Four<T> tmp){
    const int k = threadIdx.x, km = D.km;
    if(k >= km) return;
    const int j0 = blockIdx.x * kKjRows, j1 = min(j0 + kKjRows, D.jm), im_m = blockIdx.y, q = blockIdx.z;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long mo = (long long)m * D.fs;
    const T* p = F.f[q] + mo;
    T* out = tmp.f[q] + mo;
    const double coeff = 1.0 * 0.25;
    const int kk = (k == km - 1) ? 0 : k;
    const int km1 = (kk == 0 ? km - 2 : kk - 1);
    const int kp1 = (kk == km - 2 ? 0 : kk + 1);
    // the k-filtered value of row j from its three loaded values
    auto kf = [&](T w, T c, T e) -> T {
        const double f0 = (double)c;
        return (T)(f0 + coeff * ((double)w - 2.0 * f0 + (double)e));
    };
    auto row = [&](int j) -> int { return j < 0 ? 0 : (j < D.jm ? j : D.jm - 1); };
    if constexpr(sizeof(T) == 8){
        // FP64: row by row (the batched loads below take 114 VGPRs and a third of the waves)
        auto kfr = [&](int j) -> T { return kf(p[D.at(i, j, km1)], p[D.at(i, j, kk)], p[D.at(i, j, kp1)]); };
        T a = kfr(row(j0 - 1)), b = kfr(j0);
        for(int j = j0; j < j1; j++){
            const T cn = kfr(row(j + 1));
            const double f0 = (double)b;
            out[D.at(i, j, k)] = (T)(f0 + coeff * ((double)a - 2.0 * f0 + (double)cn));
            a = b; b = cn;
        }
        return;
    }
    // every row the band reads (j0 - 1 .. j1, clamped) loaded before the first store
    constexpr int NR = kKjRows + 2;
    T w[NR], c[NR], e[NR];
#pragma unroll
    for(int r = 0; r < NR; r++){
        const int jj = row(j0 - 1 + r);
        w[r] = p[D.at(i, jj, km1)]; c[r] = p[D.at(i, jj, kk)]; e[r] = p[D.at(i, jj, kp1)];
    }
    T a = kf(w[0], c[0], e[0]), b = kf(w[1], c[1], e[1]);
#pragma unroll
    for(int r = 0; r < kKjRows; r++){
        const int j = j0 + r;
        if(j >= j1) break;
        const T cn = kf(w[r + 2], c[r + 2], e[r + 2]);
        const double f0 = (double)b;
        out[D.at(i, j, k)] = (T)(f0 + coeff * ((double)a - 2.0 * f0 + (double)cn));
        a = b; b = cn;
    }
}

// the three passes: f -> t1 (k), t1 -> t2 (j), t2 -> f (i); the first two fused
template<typename T>
static void shapiro4(const Dims& D, Four<T> F, Four<T> t2, hipStream_t st){
    const dim3 row((D.km + 63) / 64 * 64);
    hipLaunchKernelGGL(k_shapiro_kj<T>, dim3((D.jm + kKjRows - 1) / kKjRows, D.im * D.M, 4), row, 0, st, D, F, t2);
    hipLaunchKernelGGL(k_shapiro_i<T>, dim3(D.jm, D.M, 4), row, 0, st, D, F, t2);
}

static int ring_threads(const Dims& D){ return (D.km - 1 + 63) / 64 * 64; }

template<typename T>
void launch_nh4sh_equilibrium(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_nh4sh_eq<T>, dim3((D.jm * D.km + 255) / 256, D.im * D.M), dim3(256), 0, st, D, S, B);
}

// cSaturnModel.cpp:979-994. The scratch of the filters is the RK4 stage-input buffer Y (four of
// its fields), which holds nothing between two steps (the stages rewrite it, its boundary cells
// are refilled before the stages).
template<typename T>
void launch_chemistry(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const RowMap RM = interior_row_map(D);
    const dim3 rows(RM.blocks());
    hipLaunchKernelGGL(k_diff_mass_flux<T>, rows, dim3(ring_threads(D)), 0, st, D, S, B, RM);
    const Four<T> t2{{B.w[B_STAGE1 + 0], B.w[B_STAGE1 + 1], B.w[B_STAGE1 + 2], B.w[B_STAGE1 + 3]}};
    shapiro4<T>(D, Four<T>{{B.w[ATSAT_DIFFLUX_H2S], B.w[ATSAT_DIFFLUX_NH3], B.w[ATSAT_DIFFLUX_NH4SH],
                            B.w[ATSAT_THERMALMASSFLUX]}}, t2, st);
    hipLaunchKernelGGL(k_chem_massflux<T>, rows, dim3(ring_threads(D)), 0, st, D, S, B, RM);
    shapiro4<T>(D, Four<T>{{B.w[ATSAT_MASSFLUX_H2S], B.w[ATSAT_MASSFLUX_NH3], B.w[ATSAT_MASSFLUX_NH4SH],
                            B.w[ATSAT_FLUXLIM_NH4SH]}}, t2, st);
}

#define INST(T) \
    template void launch_nh4sh_equilibrium<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_chemistry<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
