// The RK4 step on the device: the buoyancy reference level, the metric pressure, and one fused
// kernel per RK4 stage that evaluates RHSSat for every prognostic field of a cell, folds it
// into the accumulator and writes the next stage input (or, at the last stage, the new state).
//
// Reference: cSaturnModel::RungeKuttaSat (planet/RungeKutta_Sat_Turb.cpp:110-387) and
// cSaturnModel::RHSSat (planet/RHS_Sat_Turb.cpp:25-1016). The arithmetic follows the reference
// expression by expression (the library is built with -ffp-contract=off), so the FP64
// instantiation reproduces the reference's rounding.
//
// Stage buffers. The reference keeps y_n in the *n arrays, the stage input in the live arrays
// and the running sum in acc_* (RungeKutta_Sat_Turb.cpp:201-202). Here the live state L is the
// stage-0 input, two scratch sets X and Y hold the later stage inputs, and stage 3 writes the
// new state back into L:
//     stage 0: L -> X,   stage 1: X -> Y,   stage 2: Y -> X,   stage 3: X (+ y_n, acc) -> L.
// No stage reads the buffer it writes, so a cell never sees a neighbour at another stage (the
// property RungeKutta_Sat_Turb.cpp:178-199 is about). y_n is the *n copies, or -- in the
// resident step, where restoreVar made them equal to L -- L itself ("aliased"): stages 0-2 do
// not write L, and stage 3 reads y_n of a cell before the same thread overwrites it. X's and
// Y's boundary cells, which the interior stencils read, are copied from L once per step.
//
// Precision: L and y_n are FP64 (state class); X, Y, acc and the right-hand sides are T. The
// pressure-gradient terms and the buoyancy anomaly are differences of nearly equal FP64
// numbers and are formed in FP64 before they are rounded to T.
#include "atsat_dev.hpp"

namespace atsat {

// ------------------------------------------------------------------------------------------
// computeBuoyancyRefLevel, RungeKutta_Sat_Turb.cpp:31-50. The buoyancy anomaly b - b_ref of
// the radial momentum equation is a small difference of two O(1) numbers, so a last-bit change
// of b_ref is amplified there; the level mean is therefore summed in exactly the reference's
// serial order (j, then k) to reproduce its rounding. One wavefront per (member, level): the 64
// lanes evaluate 64 consecutive cells (the FP64 division of each b) at once, then every lane
// runs the same in-order accumulation over the 64 values, broadcast lane by lane -- the serial
// chain costs two dependent adds per cell and no barrier.
template<typename T>
__global__ void __launch_bounds__(64) k_buoy_ref(Dims D, Scal S, Bufs<T> B){
    const int i = blockIdx.x, m = blockIdx.y;
    const int lane = threadIdx.x;
    const double* t = B.s[ATSAT_T] + m * D.fs;
    const double* ps = B.s[ATSAT_P_STAT] + m * D.fs;
    const double den_c = S.r_mix * S.R_mix;
    double sum = 0.0, wsum = 0.0;
    const int ncol = D.jm * D.km;
    for(int c0 = 0; c0 < ncol; c0 += 64){
        const int c = c0 + lane;
        double wb = 0.0, w = -1.0;                    // w < 0 marks a skipped cell
        if(c < ncol){
            const int j = c / D.km, k = c - j * D.km;
            const double tv = t[D.at(i, j, k)];
            if(tv > 0.0){
                const double b = S.g * ps[D.at(i, j, k)] / (den_c * tv * S.t_ref);
                if(isfinite(b)){ w = B.thez_sin[j]; wb = w * b; }
            }
        }
        #pragma unroll 16
        for(int q = 0; q < 64; q++){
            const double wq = __shfl(w, q, 64), wbq = __shfl(wb, q, 64);
            if(wq >= 0.0){ sum += wbq; wsum += wq; }
        }
    }
    if(lane == 0) B.buoy_ref[m * D.im + i] = (wsum > 0.0) ? sum / wsum : 0.0;
}

template<typename T>
void launch_buoy_ref(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_buoy_ref<T>, dim3(D.im, D.M), dim3(64), 0, st, D, S, B);
}

// ------------------------------------------------------------------------------------------
// computeMetricPressure, RungeKutta_Sat_Turb.cpp:60-108 (METRIC_SPLIT, CORIOLIS_SPLIT, no
// HYDRO_BUOY): the trapezoid of (v^2 + w^2)/rm + cor w up each column, plus the frozen p_baro.
// One thread per (member, j, k), coalesced along k.
template<typename T>
__global__ void __launch_bounds__(256) k_metric_pressure(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    const double* v = B.s[ATSAT_V] + m * D.fs;
    const double* w = B.s[ATSAT_W] + m * D.fs;
    const double* pb = B.s[ATSAT_P_BARO] + m * D.fs;
    double* pm = B.s[ATSAT_P_METRIC] + m * D.fs;
    const double cor_ = S.cor_metric * B.sin_abs[j];
    const double exp_rm_ = 1.0;
    auto f_ = [&](int i)->double{
        const long long n = D.at(i, j, k);
        const double vv = v[n], ww = w[n];
        return ((vv * vv + ww * ww) / B.rm[i] + cor_ * ww) / exp_rm_;
    };
    double acc = 0.0, fprev = f_(0);
    pm[D.at(0, j, k)] = acc + pb[D.at(0, j, k)];
    for(int i = 1; i < D.im; i++){
        const double fi = f_(i);
        acc = acc + 0.5 * (fprev + fi) * S.dr;
        fprev = fi;
        pm[D.at(i, j, k)] = acc + pb[D.at(i, j, k)];
    }
}

template<typename T>
void launch_metric_pressure(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_metric_pressure<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, S, B);
}

// ------------------------------------------------------------------------------------------
// Copy the boundary cells of the live fields into both stage scratch sets: the planes i = 0,
// im-1 and the rows j = 1, jm-2 that the interior stencils of stages 1-3 read.
template<typename T>
__global__ void __launch_bounds__(256) k_rk_halo(Dims D, Bufs<T> B, int nint){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int m = blockIdx.z;
    if(k >= D.km) return;
    for(int f = 0; f < nint; f++){
        const double* src = B.s[f] + m * D.fs;
        T* dx = B.w[B_STAGE0 + f] + m * D.fs;
        T* dy = B.w[B_STAGE1 + f] + m * D.fs;
        const long long a = D.at(0, j, k), b = D.at(D.im - 1, j, k);
        dx[a] = (T)src[a]; dy[a] = (T)src[a];
        dx[b] = (T)src[b]; dy[b] = (T)src[b];
        if(j == 1 || j == D.jm - 2)
            for(int i = 1; i < D.im - 1; i++){ const long long n = D.at(i, j, k); dx[n] = (T)src[n]; dy[n] = (T)src[n]; }
    }
}

template<typename T>
void launch_rk_halo(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const int nint = S.turb_active ? ATSAT_NPROG : ATSAT_NPROG - 2;
    hipLaunchKernelGGL(k_rk_halo<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, B, nint);
}

// ------------------------------------------------------------------------------------------
// One RK4 stage. RHSSat for the cell (RHS_Sat_Turb.cpp:25-1016), then pass B of the stage
// (RungeKutta_Sat_Turb.cpp:250-320) or, at stage 3, the final assembly (:324-375), field by
// field as each right-hand side is formed.
template<typename T>
struct Stencil {
    T d1r, d2r, d1t, d2t, d1p, d2p, c;
};

template<typename T, typename U>
__device__ __forceinline__ Stencil<T> stencil(const U* __restrict__ F, long long c0, long long sr, long long st,
                                              long long pe, long long pw, T inv_2dr, T inv_dr2,
                                              T inv_2dthe, T inv_dthe2, T inv_2dphi, T inv_dphi2){
    Stencil<T> s;
    const T fc = (T)F[c0], frp = (T)F[c0 + sr], frm = (T)F[c0 - sr], ftp = (T)F[c0 + st], ftm = (T)F[c0 - st];
    const T fpe = (T)F[pe], fpw = (T)F[pw];
    const T one = (T)1.0, two = (T)2.0;
    s.c = fc;
    s.d1r = (frp - frm) * inv_2dr * one;                        // COMPUTE_DR, :133-135
    s.d2r = (frp - two * fc + frm) * inv_dr2 * one;
    s.d1t = (ftp - ftm) * inv_2dthe;                            // COMPUTE_DTHE, :159-161
    s.d2t = (ftp - two * fc + ftm) * inv_dthe2;
    s.d1p = (fpe - fpw) * inv_2dphi;                            // COMPUTE_DPHI, :185-187
    s.d2p = (fpe - two * fc + fpw) * inv_dphi2;
    return s;
}

template<typename T, int STAGE, bool TURB>
__global__ void __launch_bounds__(256) k_rk_stage(Dims D, Scal S, Bufs<T> B, TileMap TM, int yn0){
    int tk, tj, lev, m;
    if(!TM.decode(blockIdx.x, tk, tj, lev, m)) return;
    const int k = tk * 64 + (int)(threadIdx.x & 63);
    const int j = 2 + tj * 4 + (int)(threadIdx.x >> 6);
    const int i = 1 + lev;
    if(k >= D.km - 1 || j >= D.jm - 2) return;

    const long long mo = m * D.fs;
    const long long c0 = D.at(i, j, k);
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    const long long pe = D.at(i, j, phi_wrap(k + 1, D.km)), pw = D.at(i, j, phi_wrap(k - 1, D.km));

    const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
    const T inv_rm = (T)g[G_INV_RM], inv_rm2 = (T)g[G_INV_RM2], inv_rmsinthe = (T)g[G_INV_RMSINTHE];
    const T inv_rm2sinthe = (T)g[G_INV_RM2SINTHE], inv_rm2sinthe2 = (T)g[G_INV_RM2SINTHE2];
    const T costhe_inv_rm2sinthe = (T)g[G_COSTHE_INV_RM2SINTHE];
    const T costhe = (T)g[G_COSTHE], cotanthe = (T)g[G_COTANTHE];
    const T inv_2dr = (T)S.inv_2dr, inv_dr2 = (T)S.inv_dr2, inv_2dthe = (T)S.inv_2dthe;
    const T inv_dthe2 = (T)S.inv_dthe2, inv_2dphi = (T)S.inv_2dphi, inv_dphi2 = (T)S.inv_dphi2;
    const T one = (T)1.0, two = (T)2.0, zero = (T)0.0;

    // stage input: the live state at stage 0, a scratch set after; output the other one
    const int in_w = (STAGE == 1 || STAGE == 3) ? B_STAGE0 : B_STAGE1;
    const int out_w = (STAGE == 0 || STAGE == 2) ? B_STAGE0 : B_STAGE1;
    auto SF = [&](int f) -> Stencil<T> {
        if(STAGE == 0) return stencil<T, double>(B.s[f] + mo, c0, sr, st, pe, pw, inv_2dr, inv_dr2,
                                                 inv_2dthe, inv_dthe2, inv_2dphi, inv_dphi2);
        return stencil<T, T>(B.w[in_w + f] + mo, c0, sr, st, pe, pw, inv_2dr, inv_dr2,
                             inv_2dthe, inv_dthe2, inv_2dphi, inv_dphi2);
    };

    // pass B / final assembly of field f with right-hand side r (RungeKutta_Sat_Turb.cpp:254-375)
    auto update = [&](int f, T r){
        const double yn = B.s[yn0 + f][mo + c0];
        if(STAGE < 3){
            const double c_in = (STAGE == 0 || STAGE == 1) ? 0.5 * S.dt : S.dt;
            const T wgt = (STAGE == 0) ? one : two;
            T* acc = B.w[B_ACC + f] + mo;
            acc[c0] = (STAGE == 0) ? wgt * r : acc[c0] + wgt * r;
            double x = yn + c_in * (double)r;
            if(TURB && f == ATSAT_TKE){ x = x > 0.0 ? x : 0.0; x = x < S.tke_max_nd ? x : S.tke_max_nd; }
            if(TURB && f == ATSAT_DIS){ x = x > S.dis_min_nd ? x : S.dis_min_nd; }
            B.w[out_w + f][mo + c0] = (T)x;
        } else {
            const T a = B.w[B_ACC + f][mo + c0] + one * r;
            const double sixth_dt = S.dt / 6.0;
            double x = yn + sixth_dt * (double)a;
            if(TURB && f == ATSAT_TKE){ x = x > 0.0 ? x : 0.0; x = x < S.tke_max_nd ? x : S.tke_max_nd; }
            if(TURB && f == ATSAT_DIS){ x = x > S.dis_min_nd ? x : S.dis_min_nd; }
            B.s[f][mo + c0] = x;
        }
    };

    const Stencil<T> U = SF(ATSAT_U), V = SF(ATSAT_V), W = SF(ATSAT_W);
    const T u_ijk = U.c, v_ijk = V.c, w_ijk = W.c;

    const double* pd = B.s[ATSAT_P_DYN] + mo;
    const T dpdr = (T)((pd[c0 + sr] - pd[c0 - sr]) * S.inv_2dr * 1.0);
    const T dpdthe = (T)((pd[c0 + st] - pd[c0 - st]) * S.inv_2dthe);
    const T dpdphi = (T)((pd[pe] - pd[pw]) * S.inv_2dphi);

    const T two_inv_rm = two * inv_rm;                           // :435-436
    const T v_metric = (T)g[G_V_METRIC];
    auto transport = [&](const Stencil<T>& F){                   // :388-431
        return u_ijk * F.d1r + v_ijk * F.d1t * inv_rm + w_ijk * F.d1p * inv_rmsinthe;
    };
    auto diffusion = [&](const Stencil<T>& F){                   // :438-491
        return F.d2r + F.d1r * two_inv_rm + F.d2t * inv_rm2 + F.d1t * costhe_inv_rm2sinthe + F.d2p * inv_rm2sinthe2;
    };

    // ---- closure inputs (:511-582)
    T nue_t = zero;
    if(TURB){
        const T nu = B.w[ATSAT_NUE][mo + c0];
        nue_t = (S.turb_coupling != 0.0 && isfinite((double)nu)) ? (T)S.turb_coupling * (nu > zero ? nu : zero) : zero;
    }
    const T nue_t_s = nue_t / (T)0.9;
    const T re = (T)S.re;

    // ---- momentum (:310-350, 394-404, 441-453, 700-714, 780-833)
    T rhs_u, rhs_v, rhs_w;
    {
        const T omega = (T)S.omega, nd_cor = (T)S.nd_cor;
        const T sinthe_true = (T)g[G_SINTHE_TRUE];
        const T Coriolis_rad = nd_cor * (T)-2.0 * omega * sinthe_true * w_ijk;
        const T Coriolis_the = nd_cor * (T)-2.0 * omega * costhe * w_ijk;
        T v_cor = v_ijk, u_cor = u_ijk;
        v_cor -= (T)B.vbar0[(size_t)m * D.im * D.jm + (size_t)i * D.jm + j];
        u_cor -= (T)B.ubar0[(size_t)m * D.im * D.jm + (size_t)i * D.jm + j];
        const T Coriolis_phi = nd_cor * (T)+2.0 * omega * (+ costhe * v_cor + sinthe_true * u_cor);
        const T centrifugal_rad = (T)g[G_CENT_RAD];
        const T centrifugal_the = (T)g[G_CENT_THE];
        const T metric_on = zero, cor_rad_on = zero;
        const T transport_u = u_ijk * U.d1r + v_ijk * U.d1t * inv_rm + w_ijk * U.d1p * inv_rmsinthe
            - metric_on * (v_ijk * v_ijk + w_ijk * w_ijk) * inv_rm;
        const T transport_v = u_ijk * V.d1r + v_ijk * V.d1t * inv_rm + w_ijk * V.d1p * inv_rmsinthe
            + (u_ijk * v_ijk - w_ijk * w_ijk * cotanthe) * inv_rm;
        const T transport_w = u_ijk * W.d1r + v_ijk * W.d1t * inv_rm + w_ijk * W.d1p * inv_rmsinthe
            + (w_ijk * u_ijk + v_ijk * w_ijk * cotanthe) * inv_rm;
        const T diffusion_u = U.d2r + two * u_ijk * inv_rm2 + U.d2t * inv_rm2
            + (T)4.0 * U.d1r * inv_rm + U.d1t * costhe_inv_rm2sinthe + U.d2p * inv_rm2sinthe2;
        const T diffusion_v = V.d2r + V.d1r * two_inv_rm + V.d2t * inv_rm2
            + V.d1t * costhe_inv_rm2sinthe - v_metric * v_ijk + V.d2p * inv_rm2sinthe2
            + two * U.d1t * inv_rm2 - W.d1p * two * costhe * inv_rm2sinthe2;
        const T diffusion_w = W.d2r + W.d1r * two_inv_rm + W.d2t * inv_rm2
            + W.d1t * costhe_inv_rm2sinthe - v_metric * w_ijk + W.d2p * inv_rm2sinthe2
            + two * U.d1p * inv_rm2sinthe + V.d1p * two * costhe * inv_rm2sinthe2;
        // the buoyancy anomaly against the level mean, formed in FP64 (:700-714)
        const double t_ijk = (STAGE == 0) ? B.s[ATSAT_T][mo + c0] : (double)B.w[in_w + ATSAT_T][mo + c0];
        T buoyancy_u = zero;
        if(t_ijk > 0.0){
            const double ps = B.s[ATSAT_P_STAT][mo + c0];
            buoyancy_u = (T)(- 1.0 * 1.0 * S.buoyancy
                * (S.g * ps / (S.r_mix * S.R_mix * t_ijk * S.t_ref) - B.buoy_ref[m * D.im + i]));
        }
        const T Cor = (T)S.Coriolis, Cen = (T)S.centrifugal;
        rhs_u = - dpdr + buoyancy_u - transport_u + diffusion_u / re + diffusion_u * nue_t
                - cor_rad_on * Cor * Coriolis_rad + Cen * centrifugal_rad;
        const double* pmet = B.s[ATSAT_P_METRIC] + mo;
        const T dpmdthe = (T)((pmet[c0 + st] - pmet[c0 - st]) * S.inv_2dthe);
        const T dpmdphi = (T)((pmet[pe] - pmet[pw]) * S.inv_2dphi);
        rhs_v = - dpdthe * inv_rm - dpmdthe * inv_rm - transport_v + diffusion_v / re + diffusion_v * nue_t
                - Cor * Coriolis_the + Cen * centrifugal_the;
        rhs_w = - dpdphi * inv_rmsinthe - dpmdphi * inv_rmsinthe - transport_w + diffusion_w / re
                + diffusion_w * nue_t - Cor * Coriolis_phi;
    }
    // ---- energy (:352-383, 631-665, 752-758)
    T rhs_t;
    {
        const Stencil<T> Tt = SF(ATSAT_T);
        const double* ps = B.s[ATSAT_P_STAT] + mo;
        double dpsd = (ps[c0 + sr] - ps[c0 - sr]) * S.inv_2dr * 1.0 * S.to_nd;
        const double rho_l = B.s[ATSAT_RHO_MIX][mo + c0];
        if(rho_l > 0.0 && isfinite(rho_l)) dpsd *= S.r_mix / rho_l;
        const T dpstatdr = (T)dpsd;
        const T pressure_t = (T)S.coeff_energy_p * (u_ijk * (dpdr + dpstatdr)
            + v_ijk * dpdthe * inv_rm + w_ijk * dpdphi * inv_rmsinthe);
        T precip_t = zero;
        if(S.precip_coupling != 0.0){
            precip_t = (T)S.precip_coupling * B.w[ATSAT_Q_PRECIP][mo + c0] * (T)(S.L_atm * 1.0e3) / (T)S.precip_den;
            const T pmax = (T)0.5;
            if(!isfinite((double)precip_t)) precip_t = zero;
            else if(precip_t > pmax) precip_t = pmax;
            else if(precip_t < -pmax) precip_t = -pmax;
        }
        const T diffusion_t = diffusion(Tt);
        rhs_t = + pressure_t - transport(Tt) + diffusion_t / (T)(S.re * S.pr) + diffusion_t * nue_t_s
                - zero + zero + precip_t;
    }
    update(ATSAT_T, rhs_t);
    update(ATSAT_U, rhs_u);
    update(ATSAT_V, rhs_v);
    update(ATSAT_W, rhs_w);
    if(STAGE == 3){                                               // :1011-1013
        B.w[ATSAT_RHS_U][mo + c0] = rhs_u;
        B.w[ATSAT_RHS_V][mo + c0] = rhs_v;
        B.w[ATSAT_RHS_W][mo + c0] = rhs_w;
        B.w[ATSAT_AUX_U][mo + c0] = rhs_u + dpdr;
        B.w[ATSAT_AUX_V][mo + c0] = rhs_v + dpdthe * inv_rm;
        B.w[ATSAT_AUX_W][mo + c0] = rhs_w + dpdphi * inv_rmsinthe;
    }
    // ---- species (:839-937)
    {
        const T pc = (T)S.precip_coupling, cr = (T)S.chemical_reaction;
        const bool chem = S.chem_inputs != 0;
        auto Sp = [&](int id){ return B.w[id][mo + c0]; };
        auto species = [&](int f, double sc, T src){
            const Stencil<T> F = SF(f);
            const T dif = diffusion(F);
            update(f, - transport(F) + dif / ((T)sc * re) + dif * nue_t_s + src);
        };
        species(ATSAT_CH4, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4));
        species(ATSAT_CH4_CLOUD, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_CLOUD));
        species(ATSAT_CH4_ICE, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_ICE));
        species(ATSAT_H2O, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O));
        species(ATSAT_H2O_CLOUD, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_CLOUD));
        species(ATSAT_H2O_ICE, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_ICE));
        species(ATSAT_H2S, S.sc_h2s, cr * (chem ? Sp(ATSAT_MASSFLUX_H2S) : zero));
        {   // the chemistry term before the precipitation sink, :875-880
            const Stencil<T> F = SF(ATSAT_NH3);
            const T dif = diffusion(F);
            update(ATSAT_NH3, - transport(F) + dif / ((T)S.sc_nh3 * re) + dif * nue_t_s
                              + cr * (chem ? Sp(ATSAT_MASSFLUX_NH3) : zero) + pc * Sp(ATSAT_S_PRECIP_NH3));
        }
        species(ATSAT_NH3_CLOUD, S.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_CLOUD));
        species(ATSAT_NH3_ICE, S.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_ICE));
        {   // the flux limiter before the diffusion, :932-937
            const Stencil<T> F = SF(ATSAT_NH4SH);
            const T dif = diffusion(F);
            update(ATSAT_NH4SH, - transport(F) + (chem ? Sp(ATSAT_FLUXLIM_NH4SH) : zero) + dif / ((T)S.sc_nh4sh * re)
                                + dif * nue_t_s + cr * (chem ? Sp(ATSAT_MASSFLUX_NH4SH) : zero));
        }
    }
    // ---- k / omega transport (:511-538)
    if(TURB){
        const T nue_here = B.w[ATSAT_NUE][mo + c0];
        const Stencil<T> K = SF(ATSAT_TKE), E = SF(ATSAT_DIS);
        update(ATSAT_TKE, - transport(K) + diffusion(K) * (one / re + nue_here / (T)0.85) + B.w[ATSAT_TKE_SOURCE][mo + c0]);
        update(ATSAT_DIS, - transport(E) + diffusion(E) * (one / re + nue_here / (T)0.5) + B.w[ATSAT_DIS_SOURCE][mo + c0]);
    }
}

template<typename T, int STAGE>
static void launch_stage(const Dims& D, const Scal& S, const Bufs<T>& B, int yn0, hipStream_t st){
    TileMap TM{(D.km - 1 + 63) / 64, (D.jm - 4 + 3) / 4, D.im - 2, D.M};
    if(S.turb_active) hipLaunchKernelGGL((k_rk_stage<T, STAGE, true>), dim3(TM.blocks()), dim3(256), 0, st, D, S, B, TM, yn0);
    else hipLaunchKernelGGL((k_rk_stage<T, STAGE, false>), dim3(TM.blocks()), dim3(256), 0, st, D, S, B, TM, yn0);
}

template<typename T>
void launch_rk_stage(const Dims& D, const Scal& S, const Bufs<T>& B, int stage, bool alias, hipStream_t st){
    const int yn0 = alias ? 0 : ATSAT_F_N0;
    switch(stage){
    case 0: launch_stage<T, 0>(D, S, B, yn0, st); break;
    case 1: launch_stage<T, 1>(D, S, B, yn0, st); break;
    case 2: launch_stage<T, 2>(D, S, B, yn0, st); break;
    default: launch_stage<T, 3>(D, S, B, yn0, st); break;
    }
}

#define INST(T) \
    template void launch_buoy_ref<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_metric_pressure<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_rk_halo<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_rk_stage<T>(const Dims&, const Scal&, const Bufs<T>&, int, bool, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
