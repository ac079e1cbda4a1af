// The RK4 step on the device: the buoyancy reference level, the metric pressure, and one fused
// kernel per RK4 stage that evaluates RHSSat for every prognostic field of a cell, folds it
// into the accumulator and writes the next stage input (or, at the last stage, the new state).
//
// Reference: cSaturnModel::RungeKuttaSat (planet/RungeKutta_Sat_Turb.cpp:110-387) and
// cSaturnModel::RHSSat (planet/RHS_Sat_Turb.cpp:25-1016). The arithmetic follows the reference
// expression by expression (the library is built with -ffp-contract=off), so the FP64
// instantiation reproduces the reference's rounding except where a reduction's order differs.
//
// Stage buffers. The reference keeps y_n in the *n arrays, the stage input in the live arrays
// and the running sum in acc_* (RungeKutta_Sat_Turb.cpp:201-202). Here the stage inputs
// ping-pong between the live fields (L) and one scratch set (B): stage 0 reads L and writes B,
// stage 1 reads B and writes L, stage 2 reads L and writes B, stage 3 reads B and writes the new
// state into L. No stage reads the buffer it writes, so a cell never sees a neighbour at another
// stage (the property :178-199 of the reference is about). B's boundary cells, which the
// interior stencils read, are copied from L once per step (rk_halo).
#include "atsat_dev.hpp"

namespace atsat {

// ------------------------------------------------------------------------------------------
// computeBuoyancyRefLevel, RungeKutta_Sat_Turb.cpp:31-50. One workgroup per (member, level):
// FP64 partial sums per thread in a fixed order, then an LDS tree; deterministic, and equal to
// the reference's serial sum up to the order of the additions.
template<typename T>
__global__ void __launch_bounds__(256) k_buoy_ref(Dims D, Scal S, Bufs<T> B){
    const int i = blockIdx.x, m = blockIdx.y;
    const T* t = B.b[ATSAT_T] + m * D.fs;
    const T* ps = B.b[ATSAT_P_STAT] + m * D.fs;
    const double den_c = S.r_mix * S.R_mix;
    double sum = 0.0, wsum = 0.0;
    const int ncol = D.jm * D.km;
    for(int c = threadIdx.x; c < ncol; c += blockDim.x){
        const int j = c / D.km, k = c - j * D.km;
        const long long n = D.at(i, j, k);
        const double tv = (double)t[n];
        if(!(tv > 0.0)) continue;
        const double b = S.g * (double)ps[n] / (den_c * tv * S.t_ref);
        if(!isfinite(b)) continue;
        const double wgt = B.thez_sin[j];
        sum += wgt * b;
        wsum += wgt;
    }
    __shared__ double s_sum[256], s_w[256];
    s_sum[threadIdx.x] = sum; s_w[threadIdx.x] = wsum;
    __syncthreads();
    for(int off = 128; off > 0; off >>= 1){
        if((int)threadIdx.x < off){
            s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
            s_w[threadIdx.x] += s_w[threadIdx.x + off];
        }
        __syncthreads();
    }
    if(threadIdx.x == 0) B.buoy_ref[m * D.im + i] = (s_w[0] > 0.0) ? s_sum[0] / s_w[0] : 0.0;
}

template<typename T>
void launch_buoy_ref(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_buoy_ref<T>, dim3(D.im, D.M), dim3(256), 0, st, D, S, B);
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
    const T* v = B.b[ATSAT_V] + m * D.fs;
    const T* w = B.b[ATSAT_W] + m * D.fs;
    const T* pb = B.b[ATSAT_P_BARO] + m * D.fs;
    T* pm = B.b[ATSAT_P_METRIC] + m * D.fs;
    const double cor_ = S.cor_metric * B.sin_abs[j];
    const double exp_rm_ = 1.0;
    auto f_ = [&](int i)->double{
        const long long n = D.at(i, j, k);
        const double vv = (double)v[n], ww = (double)w[n];
        return ((vv * vv + ww * ww) / B.rm[i] + cor_ * ww) / exp_rm_;
    };
    double acc = 0.0, fprev = f_(0);
    pm[D.at(0, j, k)] = (T)(acc + (double)pb[D.at(0, j, k)]);
    for(int i = 1; i < D.im; i++){
        const double fi = f_(i);
        acc = acc + 0.5 * (fprev + fi) * S.dr;
        fprev = fi;
        pm[D.at(i, j, k)] = (T)(acc + (double)pb[D.at(i, j, k)]);
    }
}

template<typename T>
void launch_metric_pressure(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_metric_pressure<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, S, B);
}

// ------------------------------------------------------------------------------------------
// Copy the boundary cells of the live fields into the stage scratch: the planes i = 0, im-1
// and the rows j = 1, jm-2 that the interior stencils of stages 1 and 3 read from B.
template<typename T>
__global__ void __launch_bounds__(256) k_rk_halo(Dims D, Bufs<T> B, int nint){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int m = blockIdx.z;
    if(k >= D.km) return;
    for(int f = 0; f < nint; f++){
        const T* src = B.b[f] + m * D.fs;
        T* dst = B.b[B_STAGE0 + f] + m * D.fs;
        dst[D.at(0, j, k)] = src[D.at(0, j, k)];
        dst[D.at(D.im - 1, j, k)] = src[D.at(D.im - 1, j, k)];
        if(j == 1 || j == D.jm - 2)
            for(int i = 1; i < D.im - 1; i++) dst[D.at(i, j, k)] = src[D.at(i, j, k)];
    }
}

template<typename T>
void launch_rk_halo(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const int nint = S.turb_active ? ATSAT_NPROG : ATSAT_NPROG - 2;
    hipLaunchKernelGGL(k_rk_halo<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, B, nint);
}

// ------------------------------------------------------------------------------------------
// One RK4 stage. RHSSat for the cell (RHS_Sat_Turb.cpp:25-1016) then pass B of the stage
// (RungeKutta_Sat_Turb.cpp:250-320) or, at stage 3, the final assembly (:324-375).
template<typename T>
struct Stencil {
    T d1r, d2r, d1t, d2t, d1p, d2p, c;
};

template<typename T>
__device__ __forceinline__ Stencil<T> stencil(const T* __restrict__ F, long long c0, long long sr, long long st,
                                              long long pe, long long pw, T inv_2dr, T inv_dr2,
                                              T inv_2dthe, T inv_dthe2, T inv_2dphi, T inv_dphi2){
    Stencil<T> s;
    const T fc = F[c0], frp = F[c0 + sr], frm = F[c0 - sr], ftp = F[c0 + st], ftm = F[c0 - st];
    const T fpe = F[pe], fpw = F[pw];
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

template<typename T, int STAGE>
__global__ void __launch_bounds__(256) k_rk_stage(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = 2 + blockIdx.y * 4 + threadIdx.y;
    const int i = 1 + (blockIdx.z % (D.im - 2));
    const int m = blockIdx.z / (D.im - 2);
    if(k >= D.km - 1 || j >= D.jm - 2) return;

    const long long mo = m * D.fs;
    const long long c0 = D.at(i, j, k);
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    const long long pe = D.at(i, j, phi_wrap(k + 1, D.km)), pw = D.at(i, j, phi_wrap(k - 1, D.km));
    // stage input: L at stages 0 and 2, B at 1 and 3; output the other one
    const int in0 = (STAGE == 0 || STAGE == 2) ? 0 : B_STAGE0;
    const int out0 = (STAGE == 0 || STAGE == 2) ? B_STAGE0 : 0;

    const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
    const T inv_rm = (T)g[G_INV_RM], inv_rm2 = (T)g[G_INV_RM2], inv_rmsinthe = (T)g[G_INV_RMSINTHE];
    const T inv_rm2sinthe = (T)g[G_INV_RM2SINTHE], inv_rm2sinthe2 = (T)g[G_INV_RM2SINTHE2];
    const T costhe_inv_rm2sinthe = (T)g[G_COSTHE_INV_RM2SINTHE];
    const T costhe = (T)g[G_COSTHE], cotanthe = (T)g[G_COTANTHE];
    const T inv_2dr = (T)S.inv_2dr, inv_dr2 = (T)S.inv_dr2, inv_2dthe = (T)S.inv_2dthe;
    const T inv_dthe2 = (T)S.inv_dthe2, inv_2dphi = (T)S.inv_2dphi, inv_dphi2 = (T)S.inv_dphi2;
    const T one = (T)1.0, two = (T)2.0, zero = (T)0.0;

    auto SF = [&](int f){ return stencil<T>(B.b[in0 + f] + mo, c0, sr, st, pe, pw, inv_2dr, inv_dr2,
                                           inv_2dthe, inv_dthe2, inv_2dphi, inv_dphi2); };
    const Stencil<T> U = SF(ATSAT_U), V = SF(ATSAT_V), W = SF(ATSAT_W);
    const T u_ijk = U.c, v_ijk = V.c, w_ijk = W.c;

    const T* pd = B.b[ATSAT_P_DYN] + mo;
    const T dpdr = (pd[c0 + sr] - pd[c0 - sr]) * inv_2dr * one;
    const T dpdthe = (pd[c0 + st] - pd[c0 - st]) * inv_2dthe;
    const T dpdphi = (pd[pe] - pd[pw]) * inv_2dphi;

    const T two_inv_rm = two * inv_rm;                           // :435-436
    const T v_metric = (T)g[G_V_METRIC];
    auto transport = [&](const Stencil<T>& F){                   // :388-431
        return u_ijk * F.d1r + v_ijk * F.d1t * inv_rm + w_ijk * F.d1p * inv_rmsinthe;
    };
    auto diffusion = [&](const Stencil<T>& F){                   // :438-491
        return F.d2r + F.d1r * two_inv_rm + F.d2t * inv_rm2 + F.d1t * costhe_inv_rm2sinthe + F.d2p * inv_rm2sinthe2;
    };

    T rhs[ATSAT_NPROG];
    // ---- closure inputs (:511-582)
    T nue_t = zero;
    const bool turb = S.turb_active != 0;
    if(turb){
        const T nu = B.b[ATSAT_NUE][mo + c0];
        nue_t = (S.turb_coupling != 0.0 && isfinite((double)nu)) ? (T)S.turb_coupling * (nu > zero ? nu : zero) : zero;
    }
    const T nue_t_s = nue_t / (T)0.9;

    // ---- momentum (:310-350, 394-404, 441-453, 780-833)
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
        // buoyancy anomaly, :700-714
        const T t_ijk = B.b[in0 + ATSAT_T][mo + c0];
        T buoyancy_u = zero;
        if(t_ijk > zero){
            const T ps = B.b[ATSAT_P_STAT][mo + c0];
            buoyancy_u = - (T)1.0 * (T)1.0 * (T)S.buoyancy
                * ((T)S.g * ps / ((T)S.r_mix * (T)S.R_mix * t_ijk * (T)S.t_ref) - (T)B.buoy_ref[m * D.im + i]);
        }
        const T re = (T)S.re;
        const T Cor = (T)S.Coriolis, Cen = (T)S.centrifugal;
        rhs[ATSAT_U] = - dpdr + buoyancy_u - transport_u + diffusion_u / re + diffusion_u * nue_t
                       - cor_rad_on * Cor * Coriolis_rad + Cen * centrifugal_rad;
        const T* pmet = B.b[ATSAT_P_METRIC] + mo;
        const T dpmdthe = (pmet[c0 + st] - pmet[c0 - st]) * inv_2dthe;
        const T dpmdphi = (pmet[pe] - pmet[pw]) * inv_2dphi;
        rhs[ATSAT_V] = - dpdthe * inv_rm - dpmdthe * inv_rm - transport_v + diffusion_v / re + diffusion_v * nue_t
                       - Cor * Coriolis_the + Cen * centrifugal_the;
        rhs[ATSAT_W] = - dpdphi * inv_rmsinthe - dpmdphi * inv_rmsinthe - transport_w + diffusion_w / re
                       + diffusion_w * nue_t - Cor * Coriolis_phi;
    }
    // ---- energy (:352-383, 631-665, 752-758)
    {
        const Stencil<T> Tt = SF(ATSAT_T);
        const T* ps = B.b[ATSAT_P_STAT] + mo;
        T dpstatdr = (ps[c0 + sr] - ps[c0 - sr]) * inv_2dr * one * (T)S.to_nd;
        const T rho_l = B.b[ATSAT_RHO_MIX][mo + c0];
        if(rho_l > zero && isfinite((double)rho_l)) dpstatdr *= (T)S.r_mix / rho_l;
        const T pressure_t = (T)S.coeff_energy_p * (u_ijk * (dpdr + dpstatdr)
            + v_ijk * dpdthe * inv_rm + w_ijk * dpdphi * inv_rmsinthe);
        T precip_t = zero;
        if(S.precip_coupling != 0.0){
            precip_t = (T)S.precip_coupling * B.b[ATSAT_Q_PRECIP][mo + c0] * (T)(S.L_atm * 1.0e3) / (T)S.precip_den;
            const T pmax = (T)0.5;
            if(!isfinite((double)precip_t)) precip_t = zero;
            else if(precip_t > pmax) precip_t = pmax;
            else if(precip_t < -pmax) precip_t = -pmax;
        }
        const T diffusion_t = diffusion(Tt);
        rhs[ATSAT_T] = + pressure_t - transport(Tt) + diffusion_t / (T)(S.re * S.pr) + diffusion_t * nue_t_s
                       - zero + zero + precip_t;
    }
    // ---- species (:839-937)
    {
        const T re = (T)S.re, pc = (T)S.precip_coupling, cr = (T)S.chemical_reaction;
        const bool chem = S.chem_inputs != 0;
        auto Sp = [&](int id){ return B.b[id][mo + c0]; };
        auto species = [&](int f, double sc, T src){
            const Stencil<T> F = SF(f);
            const T dif = diffusion(F);
            return - transport(F) + dif / ((T)sc * re) + dif * nue_t_s + src;
        };
        rhs[ATSAT_CH4]       = species(ATSAT_CH4, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4));
        rhs[ATSAT_CH4_CLOUD] = species(ATSAT_CH4_CLOUD, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_CLOUD));
        rhs[ATSAT_CH4_ICE]   = species(ATSAT_CH4_ICE, S.sc_ch4, pc * Sp(ATSAT_S_PRECIP_CH4_ICE));
        rhs[ATSAT_H2O]       = species(ATSAT_H2O, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O));
        rhs[ATSAT_H2O_CLOUD] = species(ATSAT_H2O_CLOUD, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_CLOUD));
        rhs[ATSAT_H2O_ICE]   = species(ATSAT_H2O_ICE, S.sc_h2o, pc * Sp(ATSAT_S_PRECIP_H2O_ICE));
        rhs[ATSAT_H2S]       = species(ATSAT_H2S, S.sc_h2s, cr * (chem ? Sp(ATSAT_MASSFLUX_H2S) : zero));
        {
            const Stencil<T> F = SF(ATSAT_NH3);
            const T dif = diffusion(F);
            rhs[ATSAT_NH3] = - transport(F) + dif / ((T)S.sc_nh3 * re) + dif * nue_t_s
                             + cr * (chem ? Sp(ATSAT_MASSFLUX_NH3) : zero) + pc * Sp(ATSAT_S_PRECIP_NH3);
        }
        rhs[ATSAT_NH3_CLOUD] = species(ATSAT_NH3_CLOUD, S.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_CLOUD));
        rhs[ATSAT_NH3_ICE]   = species(ATSAT_NH3_ICE, S.sc_nh3, pc * Sp(ATSAT_S_PRECIP_NH3_ICE));
        {
            const Stencil<T> F = SF(ATSAT_NH4SH);
            const T dif = diffusion(F);
            rhs[ATSAT_NH4SH] = - transport(F) + (chem ? Sp(ATSAT_FLUXLIM_NH4SH) : zero) + dif / ((T)S.sc_nh4sh * re)
                               + dif * nue_t_s + cr * (chem ? Sp(ATSAT_MASSFLUX_NH4SH) : zero);
        }
    }
    // ---- k / omega transport (:511-538)
    if(turb){
        const T nue_here = B.b[ATSAT_NUE][mo + c0];
        const Stencil<T> K = SF(ATSAT_TKE), E = SF(ATSAT_DIS);
        const T re = (T)S.re;
        rhs[ATSAT_TKE] = - transport(K) + diffusion(K) * (one / re + nue_here / (T)0.85) + B.b[ATSAT_TKE_SOURCE][mo + c0];
        rhs[ATSAT_DIS] = - transport(E) + diffusion(E) * (one / re + nue_here / (T)0.5) + B.b[ATSAT_DIS_SOURCE][mo + c0];
    }

    // ---- pass B / final assembly
    const int nint = turb ? ATSAT_NPROG : ATSAT_NPROG - 2;
    const T tke_max = (T)S.tke_max_nd, dis_min = (T)S.dis_min_nd;
    if(STAGE < 3){
        const T c_in = (STAGE == 0 || STAGE == 1) ? (T)(0.5 * S.dt) : (T)S.dt;
        const T wgt = (STAGE == 0) ? one : two;
        for(int f = 0; f < nint; f++){
            T* acc = B.b[B_ACC + f] + mo;
            acc[c0] = (STAGE == 0) ? wgt * rhs[f] : acc[c0] + wgt * rhs[f];
            const T yn = B.b[ATSAT_F_N0 + f][mo + c0];
            T x = yn + c_in * rhs[f];
            if(f == ATSAT_TKE){ x = x > zero ? x : zero; x = x < tke_max ? x : tke_max; }
            if(f == ATSAT_DIS){ x = x > dis_min ? x : dis_min; }
            B.b[out0 + f][mo + c0] = x;
        }
    } else {
        const T sixth_dt = (T)(S.dt / 6.0);
        for(int f = 0; f < nint; f++){
            const T a = B.b[B_ACC + f][mo + c0] + one * rhs[f];
            const T yn = B.b[ATSAT_F_N0 + f][mo + c0];
            T x = yn + sixth_dt * a;
            if(f == ATSAT_TKE){ x = x > zero ? x : zero; x = x < tke_max ? x : tke_max; }
            if(f == ATSAT_DIS){ x = x > dis_min ? x : dis_min; }
            B.b[out0 + f][mo + c0] = x;
        }
        // rhs_u/v/w and aux_u/v/w keep the last stage's values (:1011-1013)
        B.b[ATSAT_RHS_U][mo + c0] = rhs[ATSAT_U];
        B.b[ATSAT_RHS_V][mo + c0] = rhs[ATSAT_V];
        B.b[ATSAT_RHS_W][mo + c0] = rhs[ATSAT_W];
        B.b[ATSAT_AUX_U][mo + c0] = rhs[ATSAT_U] + dpdr;
        B.b[ATSAT_AUX_V][mo + c0] = rhs[ATSAT_V] + dpdthe * inv_rm;
        B.b[ATSAT_AUX_W][mo + c0] = rhs[ATSAT_W] + dpdphi * inv_rmsinthe;
    }
}

template<typename T>
void launch_rk_stage(const Dims& D, const Scal& S, const Bufs<T>& B, int stage, hipStream_t st){
    const dim3 grid((D.km - 1 + 63) / 64, (D.jm - 4 + 3) / 4, (D.im - 2) * D.M), block(64, 4);
    switch(stage){
    case 0: hipLaunchKernelGGL((k_rk_stage<T, 0>), grid, block, 0, st, D, S, B); break;
    case 1: hipLaunchKernelGGL((k_rk_stage<T, 1>), grid, block, 0, st, D, S, B); break;
    case 2: hipLaunchKernelGGL((k_rk_stage<T, 2>), grid, block, 0, st, D, S, B); break;
    default: hipLaunchKernelGGL((k_rk_stage<T, 3>), grid, block, 0, st, D, S, B); break;
    }
}

#define INST(T) \
    template void launch_buoy_ref<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_metric_pressure<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_rk_halo<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_rk_stage<T>(const Dims&, const Scal&, const Bufs<T>&, int, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
