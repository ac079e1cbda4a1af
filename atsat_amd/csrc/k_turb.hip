// The turbulence closure on the device: Turbulence<cSaturnModel>::run and ::init
// (planet/Turbulence.h) at the model's defaults -- k-omega SST in its free-flow form
// (ATSAT_TURB_FREE 1: u_tau = 0.05 m/s everywhere, no wall law, wall distance 1e20 m), the
// clamps as separate passes (ATSAT_TURB_CLAMP_PASS 1), no boundary-layer taper, the curvature
// terms of the velocity gradient tensor (ATSAT_TURB_CURV 1), the buoyancy sink with the moist
// N^2 (ATSAT_TURB_BUOY, knob), no solid cells, the periodic seam, and the lid corner means of a
// model whose own boundary routines treat the closure's fields (bc_turb_active()).
//
// Four launches per call:
//   k_turb_sources  one thread per interior cell (i 1..im-2, j 1..jm-2, k 0..km-2): the
//                   production tensor, nu_t, the SST sources and the buoyancy sink
//                   (compute_sources :613-883 with compute_k_omega_SST :1020-1107), each output
//                   with its caps applied (the clamp pass after the loop, :873-881). The clamp
//                   pass before the loop (:647-652) is applied on the fly to every tke / dis value
//                   read, so the kernel reads the neighbours in the state the reference's loop
//                   sees and writes none of tke / dis;
//   k_turb_cells    one thread per cell: the floors of tke / dis in the interior (the two clamp
//                   passes), the zero-gradient copy at i = 0 (apply_wall_bc :382-424) and the nu_t
//                   clamp of every cell (clamp_nue :564-589);
//   k_turb_seam     the seam copy of the eight arrays (seam_copy :542-561);
//   k_turb_corners  one thread per longitude: the lid corners (fill_lid_corners :533-539) and
//                   u_tau (compute_vel_star :335-380).
// init() (:282-304) runs k_turb_init (init_fields :426-491 and the copy at i = 0 that init runs
// before the sources) in place of the i = 0 copy of k_turb_cells.
//
// Every formula is the reference's, operation for operation, in FP64 (the library is built with
// -ffp-contract=off), from t, w and rho_mix (state class, FP64) and u, v, the species, tke, dis,
// tken and disn (work class: FP64 in the FP64 step, FP32 in the FP32 step, widened on the load).
// nue, prod, tke_source and dis_source are work-class inputs of the RK4 stages, rounded once, on
// the store, to the context's precision, like the floors and copies of tke / dis. The blending functions F1 and
// F2 take std::tanh / the correctly rounded pow of the device, and are evaluated only where their
// value can reach a result: in the free-flow form their arguments are below 1e-10, so they enter
// the sources only through (1 - F1) = 1, the blends' outer constants and max(a1 omega, Omega F2) =
// a1 omega, which F1 = F2 = 0 give bit for bit.
#include "atsat_dev.hpp"
#include "crmath.hpp"

namespace atsat {

namespace {
// std::max / std::min / the reference's clamp3, with their NaN behaviour
__device__ __forceinline__ double smax(double a, double b){ return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b){ return (b < a) ? b : a; }
__device__ __forceinline__ double clamp3(double x, double lo, double hi){ return (x < lo) ? lo : ((x > hi) ? hi : x); }
__device__ __forceinline__ double blend(double inner, double outer, double F1){ return F1 * inner + (1.0 - F1) * outer; }

constexpr double kDisMin = 1.0e-10;           // Turbulence.h:195
constexpr double kNueAir = 1.8e-5;            // constructor default nue_gas, :104-107

// height_m(i) = get_layer_height(i) * 1e3 (Turbulence.h:262-265): the float layer heights
// i * h of init_layer_heights (cSaturnModel.h:1061-1073)
__device__ __forceinline__ double height_m(const Scal& S, int i){
    const float h = (float)S.layer_h;
    return (double)((float)i * h) * 1.0e3;
}
}  // namespace

template<typename T>
__global__ void __launch_bounds__(kRowThreads) k_turb_sources(Dims D, Scal S, Bufs<T> B, RowMap RM){
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
    // Every value the cell reads is loaded here, before any is used, and the scheduler is kept
    // from sinking the loads towards their uses: otherwise the wave waits for its loads in
    // several dependent rounds (the velocity stencils, then k / omega, then the buoyancy
    // sink's state), each a full memory round trip. The arithmetic below is unchanged.
    const Raw7d U7 = raw7(B.w[ATSAT_U], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d V7 = raw7(B.w[ATSAT_V], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d W7 = raw7(B.s[ATSAT_W], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d K7 = raw7(B.w[ATSAT_TKE], c, ip, imn, jp, jmn, pe, pw);
    const Raw7d O7 = raw7(B.w[ATSAT_DIS], c, ip, imn, jp, jmn, pe, pw);
    const int buoy = S.turb_buoy;
    // the buoyancy sink's state: t at the cell and its radial neighbours, rho_mix, and the
    // vapour, cloud and ice of H2O, NH3 and CH4 (turb_moist_gas, cSaturnModel.cpp:259-267)
    double t_c = 0.0, t_ip = 0.0, t_im = 0.0, rr = 0.0;
    double gas[3][3] = {};
    if(buoy != 0){
        const double* t = B.s[ATSAT_T];
        t_c = t[c]; t_ip = t[ip]; t_im = t[imn];
        rr = B.s[ATSAT_RHO_MIX][c];
        const int gv[3] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
#pragma unroll
        for(int q = 0; q < 3; q++)
#pragma unroll
            for(int p = 0; p < 3; p++) gas[q][p] = (double)B.w[gv[q] + p][c];
    }
    __builtin_amdgcn_sched_barrier(0);
    // The arithmetic type R: FP64 in the FP64 step (the reference's operations in its order,
    // bit-identical to it), FP32 in the FP32 step, whose outputs are FP32: there the differences
    // of neighbouring values are formed in FP64 from the loaded values and rounded once, the rest
    // in FP32, the divisions by the cell's omega and k and the row's denominators as products with
    // reciprocals formed once. The blending functions F1, F2 stay FP64 (reached only past the
    // thresholds below).
    using R = T;
    constexpr bool f32 = sizeof(R) == 4;
    auto rcp = [](R b) -> R { if constexpr (f32) return R(1) / b; else return R(0); };
    auto dv = [](R a, R b, R rb) -> R { if constexpr (f32) return a * rb; else return a / b; };
    auto mx = [](R a, R b) -> R { return (a < b) ? b : a; };
    auto mn = [](R a, R b) -> R { return (b < a) ? b : a; };
    auto cl3 = [](R x, R lo, R hi) -> R { return (x < lo) ? lo : ((x > hi) ? hi : x); };
    auto bl = [](R inner, R outer, R F1) -> R { return F1 * inner + (R(1) - F1) * outer; };
    // the clamp pass before the loop (:647-652) covers i 1..im-2, j 1..jm-2, k 0..km-2: the
    // neighbours at i = 0, im-1 and j = 0, jm-1 are read as they are
    const bool in_ip = i + 1 <= D.im - 2, in_im = i - 1 >= 1, in_jp = j + 1 <= D.jm - 2, in_jm = j - 1 >= 1;
    auto kc = [](double x){ return smax(0.0, x); };
    auto dc = [](double x){ return smax(kDisMin, x); };
    const double tk64 = kc(K7.c), dk64 = dc(O7.c);
    const double tk_ip = in_ip ? kc(K7.rp) : K7.rp, tk_im = in_im ? kc(K7.rm) : K7.rm;
    const double dk_ip = in_ip ? dc(O7.rp) : O7.rp, dk_im = in_im ? dc(O7.rm) : O7.rm;
    const double tk_jp = in_jp ? kc(K7.tp) : K7.tp, tk_jm = in_jm ? kc(K7.tm) : K7.tm;
    const double dk_jp = in_jp ? dc(O7.tp) : O7.tp, dk_jm = in_jm ? dc(O7.tm) : O7.tm;
    const double tk_pe = kc(K7.pe), tk_pw = kc(K7.pw), dk_pe = dc(O7.pe), dk_pw = dc(O7.pw);
    const R tk = (R)tk64, dk = (R)dk64;

    // geometry, :675-694 (coord_stretching false: exp_rm = 1)
    const R rm = (R)B.rm[i];
    const R exp_rm = R(1);
    double sinthe64 = B.thez_sin[j];
    if(sinthe64 == 0.0) sinthe64 = 1.0e-5;
    const R sinthe = (R)sinthe64;
    const R costhe = (R)B.geo[(long long)j * NGEO + G_COSTHE];
    const R cotanthe = costhe / sinthe;
    const R inv_rm = R(1) / rm;
    const R rmsinthe = rm * sinthe;
    const R inv_2dr = R(1) / (R(2) * (R)S.dr);
    const R inv_2dthe = R(1) / (R(2) * (R)S.dthe);
    const R inv_2dphi = R(1) / (R(2) * (R)S.dphi);
    const R inv_rm2dthe = inv_2dthe / rm;
    const R inv_rmsinthe2dphi = inv_2dphi / rmsinthe;
    // velocity gradients, :699-709
    const R dudr = (R)(U7.rp - U7.rm) * inv_2dr * exp_rm;
    const R dvdr = (R)(V7.rp - V7.rm) * inv_2dr * exp_rm;
    const R dwdr = (R)(W7.rp - W7.rm) * inv_2dr * exp_rm;
    const R dudthe = (R)(U7.tp - U7.tm) * inv_rm2dthe;
    const R dvdthe = (R)(V7.tp - V7.tm) * inv_rm2dthe;
    const R dwdthe = (R)(W7.tp - W7.tm) * inv_rm2dthe;
    const R dudphi = (R)(U7.pe - U7.pw) * inv_rmsinthe2dphi;
    const R dvdphi = (R)(V7.pe - V7.pw) * inv_rmsinthe2dphi;
    const R dwdphi = (R)(W7.pe - W7.pw) * inv_rmsinthe2dphi;
    // k / omega radial gradients, :729-730 (no solid cells: no Neumann faces)
    const R dtkedr = (R)(tk_ip - tk_im) * inv_2dr * exp_rm;
    const R ddisdr = (R)(dk_ip - dk_im) * inv_2dr * exp_rm;
    // nu_t from the current k / omega, capped, :739-748
    const R nue_max = (R)S.turb_nue_max_nd;
    // FP32 step: the reciprocals of the cell's omega (k and omega are floored above, so the floors
    // of 1e-10 and 1e-20 below give the same value) and of the row's denominators, formed once
    const R r_dk = rcp(mx(dk, R(1.0e-20)));
    R cnue;
    {
        const R dis_here = mx(dk, (R)kDisMin);
        const R tke_here = mx(tk, R(0));
        cnue = dv(tke_here, dis_here, r_dk);
        cnue = mn(cnue, nue_max);
    }
    // the velocity gradient tensor with its curvature terms, :763-797
    const R u_r = (R)U7.c * inv_rm;
    const R v_r = (R)V7.c * inv_rm;
    const R w_r = (R)W7.c * inv_rm;
    const R v_cot_r = v_r * cotanthe;
    const R w_cot_r = w_r * cotanthe;
    const R g_rr = dudr;
    const R g_rt = dudthe - v_r;
    const R g_rp = dudphi - w_r;
    const R g_tr = dvdr;
    const R g_tt = dvdthe + u_r;
    const R g_tp = dvdphi - w_cot_r;
    const R g_pr = dwdr;
    const R g_pt = dwdthe;
    const R g_pp = dwdphi + u_r + v_cot_r;
    const R c23 = (R)0.66667, two = R(2);
    const R der = c23 * (g_rr + g_tt + g_pp);
    const R prod = mx(R(0),
          (cnue * (two * g_rr - der) - c23 * tk) * g_rr
        + (cnue * (g_rt + g_tr))                     * g_rt
        + (cnue * (g_rp + g_pr))                     * g_rp
        + (cnue * (two * g_tt - der) - c23 * tk) * g_tt
        + (cnue * (g_tr + g_rt))                     * g_tr
        + (cnue * (g_tp + g_pt))                     * g_tp
        + (cnue * (two * g_pp - der) - c23 * tk) * g_pp
        + (cnue * (g_pr + g_rp))                     * g_pr
        + (cnue * (g_pt + g_tp))                     * g_pt);
    const R W12 = g_rt - g_tr;
    const R W13 = g_rp - g_pr;
    const R W23 = g_tp - g_pt;
    const R Omega = sqrt(W12 * W12 + W13 * W13 + W23 * W23);

    // compute_k_omega_SST, :1020-1107, with y_phys = free_wall_dist() = 1e20 m
    const double L_m = S.L_atm * 1.0e3;
    R nue, tsrc, dsrc;
    {
        const R a1 = (R)0.31, bet_star = (R)0.09, bet1 = (R)0.0333, bet2 = (R)0.0368;
        const R gam1 = (R)0.413, gam2 = (R)0.2, sig_w2 = (R)1.168;
        const R nue_air_nd = (R)(kNueAir / (S.u_0 * L_m));
        const double y_phys = 1.0e20;
        const R y_star = (R)(y_phys / L_m);
        const R dtkedr_neu = dtkedr, ddisdr_neu = ddisdr;
        const R dthe_den = rm * two * (R)S.dthe, dphi_den = rmsinthe * two * (R)S.dphi;
        const R r_dthe_den = rcp(dthe_den), r_dphi_den = rcp(dphi_den);
        const R dtkedthe_neu = dv((R)(tk_jp - tk_jm), dthe_den, r_dthe_den);
        const R ddisdthe_neu = dv((R)(dk_jp - dk_jm), dthe_den, r_dthe_den);
        const R dtkedphi_neu = dv((R)(tk_pe - tk_pw), dphi_den, r_dphi_den);
        const R ddisdphi_neu = dv((R)(dk_pe - dk_pw), dphi_den, r_dphi_den);
        const R grad_dot = dtkedr_neu * ddisdr_neu
                         + dtkedthe_neu * ddisdthe_neu
                         + dtkedphi_neu * ddisdphi_neu;
        const R dis_safe = mx(dk, R(1.0e-20));
        const R CD_kw = mx(dv(two * sig_w2, dis_safe, r_dk) * grad_dot, R(1.0e-20));
        const R sq_tk = sqrt(tk);
        const R r_y = rcp(bet_star * y_star), r_yy = rcp(y_star * y_star);
        const R low = dv(R(500) * nue_air_nd, y_star * y_star * dis_safe, r_yy * r_dk);
        const R arg1 = mn(
            mx(dv(sq_tk, bet_star * dis_safe * y_star, r_y * r_dk), low),
            R(4) * sig_w2 * tk / (CD_kw * y_star * y_star));
        const R arg2 = mx(
            dv(two * sq_tk, bet_star * dis_safe * y_star, r_y * r_dk), low);
        // F1 and F2 exactly where they can change a result: with F1 <= arg1^4 < 6.25e-18,
        // (1 - F1) rounds to 1 and F1 * gamma_1 + (1 - F1) * gamma_2 to gamma_2 (also for beta),
        // as with F1 = 0; with Omega F2 <= Omega arg2^2 below half of a1 omega, the max takes a1 omega
        // either way
        const R F1 = (arg1 < R(5.0e-5)) ? R(0) : (R)tanh(crm::Phys<T>::pow((double)arg1, 4.0));
        const R F2 = (isfinite(Omega) && Omega * (arg2 * arg2) < R(0.5) * (a1 * dis_safe))
                   ? R(0) : (R)tanh(crm::Phys<T>::pow((double)arg2, 2.0));
        nue = a1 * tk / mx(a1 * dis_safe, Omega * F2);
        const R tke_safe = mx(tk, R(1.0e-20));
        const R P_k = mn(prod, R(20) * bet_star * tke_safe * dis_safe);
        const R Y_k = bet_star * tke_safe * dis_safe;
        const R P_w = dv(bl(gam1, gam2, F1) * P_k * dis_safe, tke_safe, rcp(tke_safe));
        const R Y_w = bl(bet1, bet2, F1) * dis_safe * dis_safe;
        const R D_w = dv(two * (R(1) - F1) * sig_w2, dis_safe, r_dk) * grad_dot;
        tsrc = P_k - Y_k;
        dsrc = P_w - Y_w + D_w;
    }

    // the buoyancy sink of the k equation, :836-868; cp_lapse = cp_mix on ATSAT
    if(buoy != 0){
        const R Pr_t = (R)0.9;
        const R TK = (R)(t_c * S.t_ref);
        const R dz = (R)(height_m(S, i + 1) - height_m(S, i - 1));
        const R dTdz = dv((R)((t_ip - t_im) * S.t_ref), dz, rcp(dz));
        R Gam = (R)(S.g / S.cp_mix), fac = R(1);
        if(buoy == 3 && TK > R(0)){
            const R rho = (R)((rr > 0.0 && isfinite(rr)) ? rr : S.r_mix);      // rho_at, cSaturnModel.h:514-518
            R A = R(0), Bs = R(0);
            // turb_moist_gas, cSaturnModel.cpp:259-267: H2O, NH3, CH4
            const int gq[3] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };
            const R R_mix = (R)S.R_mix, cp_mix = (R)S.cp_mix;
            const R r_rho = rcp(rho), r_RT = rcp(R_mix * TK);
            const R r_cRTT = rcp(cp_mix * R_mix * TK * TK);
            for(int q = 0; rho > R(0) && q < 3; q++){
                const R qc = dv(mx(R(0), (R)gas[q][1]) + mx(R(0), (R)gas[q][2]), rho, r_rho);
                if(!(qc > R(1.0e-9))) continue;
                const R qv = dv(mx(R(0), (R)gas[q][0]), rho, r_rho);
                const int g = gq[q];
                const R L = TK >= (R)S.t_0[g] ? (R)S.lv[g] : (R)S.ls[g];
                A += dv(L * qv, R_mix * TK, r_RT);
                Bs += dv((R)S.ep[g] * L * L * qv, cp_mix * R_mix * TK * TK, r_cRTT);
            }
            if(isfinite(A + Bs)){ Gam *= (R(1) + A) / (R(1) + Bs); fac = R(1) + A; }
        }
        const R N2 = (TK > R(0) && dz > R(0)) ? (R)S.g / TK * (dTdz + Gam) * fac : R(0);
        if(isfinite(N2) && (N2 > R(0) || buoy == 2)){
            const R tL = (R)(L_m / S.u_0);
            tsrc -= mx(R(0), nue) * N2 * tL * tL / Pr_t;
        }
    }

    // the caps of the clamp pass after the loop, :633-645, with the cell's floored k / omega
    const R tk_out = mx((R)S.turb_k_bg_nd, tk), dk_out = mx((R)kDisMin, dk);
    nue = cl3(nue, R(0), nue_max);
    const R dis_src_max = R(20) * dk_out * dk_out;
    dsrc = cl3(dsrc, -dis_src_max, dis_src_max);
    const R tke_src_max = R(20) * tk_out * dk_out;
    tsrc = cl3(tsrc, -tke_src_max, tke_src_max);
    B.w[ATSAT_PROD][c] = (T)prod;
    B.w[ATSAT_NUE][c] = (T)nue;
    B.w[ATSAT_TKE_SOURCE][c] = (T)tsrc;
    B.w[ATSAT_DIS_SOURCE][c] = (T)dsrc;
}

// init_fields, :426-491 (TURB_FREE: k = k_bg and omega = max(omega(k, 1e20 m), om_bg) in every
// interior cell, with tken / disn), then the copy at i = 0 that init runs before the sources
// (apply_wall_bc :382-424). One thread per column (j 0..jm-1, k 0..km-2).
template<typename T>
__global__ void __launch_bounds__(256) k_turb_init(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km - 1) return;
    const long long mo = (long long)m * D.fs;
    T* tke = B.w[ATSAT_TKE] + mo; T* dis = B.w[ATSAT_DIS] + mo;
    T* tken = B.w[ATSAT_F_N0 + ATSAT_TKE] + mo; T* disn = B.w[ATSAT_F_N0 + ATSAT_DIS] + mo;
    T* nue = B.w[ATSAT_NUE] + mo;
    if(j >= 1 && j <= D.jm - 2){
        for(int i = 1; i <= D.im - 2; i++){
            const long long n = D.at(i, j, k);
            tke[n] = (T)S.turb_init_tke; dis[n] = (T)S.turb_init_dis;
            tken[n] = tke[n]; disn[n] = dis[n];
        }
    }
    const long long n0 = D.at(0, j, k), n1 = D.at(1, j, k);
    tke[n0] = tke[n1]; dis[n0] = dis[n1]; nue[n0] = nue[n1];
}

// The cell pass after the sources: the interior floors of tke / dis (the clamp pass before the
// loop, then the one after it, :630-645), the copy at i = 0 (run only, apply_wall_bc :382-424)
// and the nu_t clamp (clamp_nue :564-589), one thread per cell of the columns k 0..km-2. A thread
// at i = 0 floors the value of i = 1 it copies itself: the floors and the clamp are idempotent,
// so it does not matter whether the thread of i = 1 has already written it.
template<typename T>
__global__ void __launch_bounds__(1024) k_turb_cells(Dims D, Scal S, Bufs<T> B, int wall){
    const int k = threadIdx.x;
    const int j = blockIdx.x, im_m = blockIdx.y;
    if(k >= D.km - 1) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long mo = (long long)m * D.fs;
    T* tke = B.w[ATSAT_TKE] + mo; T* dis = B.w[ATSAT_DIS] + mo;
    T* nue = B.w[ATSAT_NUE] + mo;
    const bool jin = j >= 1 && j <= D.jm - 2;
    auto kfl = [&](double x){ return smax(S.turb_k_bg_nd, smax(0.0, x)); };
    auto dfl = [&](double x){ return smax(kDisMin, smax(kDisMin, x)); };
    auto nfl = [&](double x){ return clamp3(x, 0.0, S.turb_nue_max_nd) * 1.0; };
    const long long n = D.at(i, j, k);
    if(jin && i >= 1 && i <= D.im - 2){
        tke[n] = (T)kfl((double)tke[n]); dis[n] = (T)dfl((double)dis[n]);
    }
    if(i == 0 && wall){
        const long long n1 = D.at(1, j, k);
        const bool in1 = jin && 1 <= D.im - 2;
        tke[n] = in1 ? (T)kfl((double)tke[n1]) : tke[n1];
        dis[n] = in1 ? (T)dfl((double)dis[n1]) : dis[n1];
        nue[n] = (T)nfl((double)nue[n1]);
    } else {
        nue[n] = (T)nfl((double)nue[n]);
    }
}

// seam_copy, :542-561: the column k = km-1 of the eight arrays takes k = 0, one thread per (i, j)
template<typename T>
__global__ void __launch_bounds__(256) k_turb_seam(Dims D, Bufs<T> B){
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int im_m = blockIdx.y;
    if(j >= D.jm) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long mo = (long long)m * D.fs;
    const long long n = mo + D.at(i, j, 0), s = mo + D.at(i, j, D.km - 1);
    const int wf[8] = { ATSAT_TKE, ATSAT_DIS, ATSAT_F_N0 + ATSAT_TKE, ATSAT_F_N0 + ATSAT_DIS,
                        ATSAT_NUE, ATSAT_PROD, ATSAT_TKE_SOURCE, ATSAT_DIS_SOURCE };
    for(int q = 0; q < 8; q++) B.w[wf[q]][s] = B.w[wf[q]][n];
}

// fill_lid_corners, :533-539 (TURB_FREE, bc_turb_active()): the two lid corners (i = im-1,
// j = 0 and jm-1) of the eight arrays get the mean of their lid and pole neighbours; and u_tau,
// compute_vel_star :358 (TURB_FREE: vel_star_min). One thread per longitude k 0..km-1.
template<typename T>
__global__ void __launch_bounds__(256) k_turb_corners(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if(k >= D.km) return;
    const long long mo = (long long)m * D.fs;
    const int il = D.im - 1, jl = D.jm - 1;
    const long long a0 = mo + D.at(il, 0, k), a1 = mo + D.at(il, 1, k), a2 = mo + D.at(il - 1, 0, k);
    const long long b0 = mo + D.at(il, jl, k), b1 = mo + D.at(il, jl - 1, k), b2 = mo + D.at(il - 1, jl, k);
    const int wf[8] = { ATSAT_TKE, ATSAT_DIS, ATSAT_F_N0 + ATSAT_TKE, ATSAT_F_N0 + ATSAT_DIS,
                        ATSAT_NUE, ATSAT_PROD, ATSAT_TKE_SOURCE, ATSAT_DIS_SOURCE };
    for(int q = 0; q < 8; q++){
        T* F = B.w[wf[q]];
        F[a0] = (T)(0.5 * ((double)F[a1] + (double)F[a2]));
        F[b0] = (T)(0.5 * ((double)F[b1] + (double)F[b2]));
    }
    double* vs = B.vel_star + (long long)m * D.jm * D.km;
    for(int j = 0; j < D.jm; j++) vs[(long long)j * D.km + k] = 0.05;
}

int turb_ring_threads(const Dims& D){ return (D.km - 1 + 63) / 64 * 64; }

template<typename T>
void launch_turbulence(const Dims& D, const Scal& S, const Bufs<T>& B, bool init, hipStream_t st){
    const dim3 cols((D.km - 1 + 255) / 256, D.jm, D.M);
    if(init) hipLaunchKernelGGL(k_turb_init<T>, cols, dim3(256), 0, st, D, S, B);
    const RowMap RM = interior_row_map(D);
    hipLaunchKernelGGL(k_turb_sources<T>, dim3(RM.blocks()), dim3(turb_ring_threads(D)), 0, st, D, S, B, RM);
    hipLaunchKernelGGL(k_turb_cells<T>, dim3(D.jm, D.im * D.M), dim3(turb_ring_threads(D)), 0, st, D, S, B, init ? 0 : 1);
    hipLaunchKernelGGL(k_turb_seam<T>, dim3((D.jm + 255) / 256, D.im * D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_turb_corners<T>, dim3((D.km + 255) / 256, D.M), dim3(256), 0, st, D, S, B);
}

template void launch_turbulence<double>(const Dims&, const Scal&, const Bufs<double>&, bool, hipStream_t);
template void launch_turbulence<float>(const Dims&, const Scal&, const Bufs<float>&, bool, hipStream_t);

}  // namespace atsat
