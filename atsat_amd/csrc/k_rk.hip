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
// not write L, and stage 3 reads y_n of a cell before the same thread overwrites it. With
// turb_active, tke and dis always read y_n from tken / disn, which restoreVar fills: the
// turbulence closure runs after it and changes the live tke / dis only. X's and
// Y's boundary cells, which the interior stencils read, are copied from L once per step.
//
// Precision: t and w of L and y_n are FP64 (state class), the other 15 prognostic fields of L
// and y_n are T like X, Y, acc and the right-hand sides; the stage update y_n + c dt rhs is formed
// in FP64 and rounded once to its field's class. The pressure-gradient terms and the buoyancy
// anomaly are differences of nearly equal FP64 numbers and are formed in FP64 before they are
// rounded to T.
#include "atsat_dev.hpp"

#include <type_traits>

namespace atsat {

// ------------------------------------------------------------------------------------------
// computeBuoyancyRefLevel, RungeKutta_Sat_Turb.cpp:31-50. The buoyancy anomaly b - b_ref of
// the radial momentum equation is a small difference of two O(1) numbers, so a last-bit change
// of b_ref is amplified there; the level mean is therefore summed in exactly the reference's
// serial order (j, then k) to reproduce its rounding. Two passes:
//   k_buoy_terms  every cell's weighted term sin(theta_j) * b in parallel (the FP64 division is
//                 the expensive part), NaN for a cell the reference skips; written through an
//                 LDS transpose so that the terms of one position for all (member, level) chains
//                 are contiguous;
//   k_buoy_sum    one lane per (member, level) chain adds its terms in order -- the serial
//                 chain is one dependent add per cell -- fed through LDS by the whole workgroup
//                 (below). When no cell of the level was skipped, the weight
//                 sum is the same serial sum of sin(theta_j) for every level, formed once on the
//                 host; a NaN sum (some cell skipped) redoes the chain with the skips.
// The FP32 step (T = float) keeps the terms and the sums in FP64 but does not need the serial
// rounding: its stages round the anomaly to FP32, 1e8 times coarser than the last bit of b_ref.
// There k_buoy_chain sums the terms and the weights of the cells not skipped, each chain by one
// workgroup in a fixed tree: the same value up to FP64 rounding, without the 65 341-long dependent
// chain of the serial sum.
template<typename T>
__global__ void __launch_bounds__(256) k_buoy_terms(Dims D, Scal S, Bufs<T> B){
    __shared__ double tile[64][65];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npos = D.jm * D.km, nch = D.M * D.im;
    const int pos0 = blockIdx.x * 64, ch0 = blockIdx.y * 64;
    const int pos = pos0 + lane;
    int j = 0, k = 0;
    if(pos < npos){ j = pos / D.km; k = pos - j * D.km; }
    const double wgt = B.thez_sin[j];
    const double den_c = S.r_mix * S.R_mix;
    // the 32 loads of the wave's 16 cells are issued before any of them is used
    double tv[16], pv[16];
#pragma unroll
    for(int cc = 0; cc < 16; cc++){
        const int ch = ch0 + wv * 16 + cc;
        tv[cc] = 0.0; pv[cc] = 0.0;
        if(pos < npos && ch < nch){
            const int m = ch / D.im, i = ch - m * D.im;
            const long long n = (long long)m * D.fs + D.at(i, j, k);
            tv[cc] = B.s[ATSAT_T][n];
            pv[cc] = B.s[ATSAT_P_STAT][n];
        }
    }
#pragma unroll
    for(int cc = 0; cc < 16; cc++){
        const int cl = wv * 16 + cc, ch = ch0 + cl;
        double x = 0.0;
        if(pos < npos && ch < nch){
            x = __builtin_nan("");
            if(tv[cc] > 0.0){
                const double b = S.g * pv[cc] / (den_c * tv[cc] * S.t_ref);
                if(isfinite(b)) x = wgt * b;
            }
        }
        tile[cl][lane] = x;
    }
    __syncthreads();
    for(int pp = 0; pp < 16; pp++){
        const int p = wv * 16 + pp, ps = pos0 + p;
        if(ps < npos) B.buoy_x[(long long)ps * B.buoy_ch + ch0 + lane] = tile[lane][p];
    }
}

// The FP32 step's level mean: one workgroup per (member, level) chain streams its cells (the same
// terms as k_buoy_terms), eight per thread at a time with their loads issued together, and sums
// the terms and the weights of the cells not skipped; the 256 partial sums in a fixed tree.
template<typename T>
__global__ void __launch_bounds__(256) k_buoy_chain(Dims D, Scal S, Bufs<T> B){
    __shared__ double red[2][256];
    const int ch = blockIdx.x, m = ch / D.im, i = ch - m * D.im;
    const double den_c = S.r_mix * S.R_mix;
    const long long base = (long long)m * D.fs + (long long)i * D.jm * D.kp;
    const double* tf = B.s[ATSAT_T];
    const double* pf = B.s[ATSAT_P_STAT];
    double sx = 0.0, sw = 0.0;
    const int ncell = D.jm * D.km;
    constexpr int U = 8;
    for(int q0 = threadIdx.x; q0 < ncell; q0 += U * 256){
        double tv[U], pv[U], wv[U];
#pragma unroll
        for(int u = 0; u < U; u++){
            const int q = q0 + u * 256;
            tv[u] = 0.0; pv[u] = 0.0; wv[u] = 0.0;
            if(q < ncell){
                const int j = q / D.km, k = q - j * D.km;
                const long long n = base + (long long)j * D.kp + k;
                tv[u] = tf[n]; pv[u] = pf[n]; wv[u] = B.thez_sin[j];
            }
        }
#pragma unroll
        for(int u = 0; u < U; u++){
            const double b = S.g * pv[u] / (den_c * tv[u] * S.t_ref);
            if(tv[u] > 0.0 && isfinite(b)){ sx += wv[u] * b; sw += wv[u]; }
        }
    }
    const int t = threadIdx.x;
    red[0][t] = sx; red[1][t] = sw;
    __syncthreads();
    for(int h = 128; h > 0; h >>= 1){
        if(t < h){ red[0][t] += red[0][t + h]; red[1][t] += red[1][t + h]; }
        __syncthreads();
    }
    if(t == 0) B.buoy_ref[ch] = (red[1][0] > 0.0) ? red[0][0] / red[1][0] : 0.0;
}

// k_buoy_sum: 16 chains per workgroup. The lanes 0-15 of wave 0 run the 16 in-order sums out of
// an LDS chunk of kBuoyChunk positions while all 256 threads already have the next chunk's loads
// in flight in registers; between chunks the workgroup moves those registers into LDS.
constexpr int kBuoyChains = 16, kBuoyChunk = 512, kBuoyPer = kBuoyChunk * kBuoyChains / 256, kBuoyU = 16;

template<typename T>
__global__ void __launch_bounds__(256) k_buoy_sum(Dims D, Scal S, Bufs<T> B){
    __shared__ double chunk[kBuoyChunk][kBuoyChains];
    const int t = threadIdx.x;
    const int ch0 = blockIdx.x * kBuoyChains;
    const int nch = D.M * D.im, npos = D.jm * D.km;
    const long long st = B.buoy_ch;
    const int nchunk = (npos + kBuoyChunk - 1) / kBuoyChunk;
    // element e = t + 256 q of a chunk is (position e / 16, chain e % 16): 16 consecutive lanes
    // read one 128-byte row of the term matrix
    const double* xb = B.buoy_x + ch0 + (t & 15);
    double r[kBuoyPer];
    auto load = [&](int c){
        #pragma unroll
        for(int q = 0; q < kBuoyPer; q++){
            const int pos = c * kBuoyChunk + ((t + 256 * q) >> 4);
            r[q] = pos < npos ? xb[(long long)pos * st] : 0.0;
        }
    };
    auto store = [&](){
        #pragma unroll
        for(int q = 0; q < kBuoyPer; q++) chunk[(t + 256 * q) >> 4][t & 15] = r[q];
    };
    double sum = 0.0;
    load(0);
    store();
    __syncthreads();
    if(nchunk > 1) load(1);
    for(int c = 0; c < nchunk; c++){
        if(t < kBuoyChains){
            const int n = (c == nchunk - 1) ? npos - c * kBuoyChunk : kBuoyChunk;
            // in order, kBuoyU terms at a time, the next kBuoyU read from LDS while these are
            // added (enough dependent adds to cover the LDS latency)
            int p = 0;
            if(n >= kBuoyU){
                double a[kBuoyU];
                #pragma unroll
                for(int u = 0; u < kBuoyU; u++) a[u] = chunk[u][t];
                for(p = kBuoyU; p + kBuoyU <= n; p += kBuoyU){
                    double b[kBuoyU];
                    #pragma unroll
                    for(int u = 0; u < kBuoyU; u++) b[u] = chunk[p + u][t];
                    #pragma unroll
                    for(int u = 0; u < kBuoyU; u++) sum += a[u];
                    #pragma unroll
                    for(int u = 0; u < kBuoyU; u++) a[u] = b[u];
                }
                #pragma unroll
                for(int u = 0; u < kBuoyU; u++) sum += a[u];
            }
            for(; p < n; p++) sum += chunk[p][t];
        }
        __syncthreads();
        if(c + 1 < nchunk){
            store();
            __syncthreads();
            if(c + 2 < nchunk) load(c + 2);
        }
    }
    const int ch = ch0 + t;
    if(t >= kBuoyChains || ch >= nch) return;
    double wsum = S.buoy_wsum_full;
    if(isnan(sum)){                         // some cell of the level was skipped
        const double* x = B.buoy_x + ch;
        sum = 0.0; wsum = 0.0;
        for(int q = 0; q < npos; q++){
            const double xv = x[(long long)q * st];
            if(!isnan(xv)){ sum += xv; wsum += B.thez_sin[q / D.km]; }
        }
    }
    B.buoy_ref[ch] = (wsum > 0.0) ? sum / wsum : 0.0;
}

template<typename T>
void launch_buoy_ref(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const int npos = D.jm * D.km, nch = D.M * D.im;
    if(std::is_same<T, float>::value){
        hipLaunchKernelGGL(k_buoy_chain<T>, dim3(nch), dim3(256), 0, st, D, S, B);
    } else {
        hipLaunchKernelGGL(k_buoy_terms<T>, dim3((npos + 63) / 64, (unsigned)(B.buoy_ch / 64)), dim3(256), 0, st, D, S, B);
        hipLaunchKernelGGL(k_buoy_sum<T>, dim3((nch + kBuoyChains - 1) / kBuoyChains), dim3(256), 0, st, D, S, B);
    }
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
    const T* v = B.w[ATSAT_V] + m * D.fs;
    const double* w = B.s[ATSAT_W] + m * D.fs;
    const double* pb = B.s[ATSAT_P_BARO] + m * D.fs;
    double* pm = B.s[ATSAT_P_METRIC] + m * D.fs;
    const double cor_ = S.cor_metric * B.sin_abs[j];
    const double exp_rm_ = 1.0;
    auto f_ = [&](int i)->double{
        const long long n = D.at(i, j, k);
        const double vv = (double)v[n], ww = w[n];
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

// The pressure terms of the stages' right-hand sides, each formed in FP64 and rounded once to the
// work class: from p_dyn, p_stat and rho_mix (RHS_Sat_Turb.cpp:154, 180, 206, 372-377: dpdr,
// dpdthe, dpdphi, dpstatdr) and from p_metric (:814-817: dpmdthe, dpmdphi). c0 is the cell, sr /
// st the r / theta strides, pe / pw its east / west phi neighbours (the seam wrapped).
template<typename T>
__device__ __forceinline__ void pterm_static(const Scal& S, const Bufs<T>& B, long long c0, long long sr, long long st,
                                             long long pe, long long pw, T& dpdr, T& dpdthe, T& dpdphi, T& dpstatdr){
    const double* pd = B.s[ATSAT_P_DYN];
    const double* ps = B.s[ATSAT_P_STAT];
    const double pd_rp = pd[c0 + sr], pd_rm = pd[c0 - sr], pd_tp = pd[c0 + st], pd_tm = pd[c0 - st];
    const double pd_pe = pd[pe], pd_pw = pd[pw];
    const double ps_rp = ps[c0 + sr], ps_rm = ps[c0 - sr];
    const double rho_l = B.s[ATSAT_RHO_MIX][c0];
    dpdr = (T)((pd_rp - pd_rm) * S.inv_2dr * 1.0);
    dpdthe = (T)((pd_tp - pd_tm) * S.inv_2dthe);
    dpdphi = (T)((pd_pe - pd_pw) * S.inv_2dphi);
    double dpsd = (ps_rp - ps_rm) * S.inv_2dr * 1.0 * S.to_nd;
    if(rho_l > 0.0 && isfinite(rho_l)) dpsd *= S.r_mix / rho_l;
    dpstatdr = (T)dpsd;
}
template<typename T>
__device__ __forceinline__ void pterm_metric(const Scal& S, const Bufs<T>& B, long long c0, long long st,
                                             long long pe, long long pw, T& dpmdthe, T& dpmdphi){
    const double* pm = B.s[ATSAT_P_METRIC];
    const double pm_tp = pm[c0 + st], pm_tm = pm[c0 - st], pm_pe = pm[pe], pm_pw = pm[pw];
    dpmdthe = (T)((pm_tp - pm_tm) * S.inv_2dthe);
    dpmdphi = (T)((pm_pe - pm_pw) * S.inv_2dphi);
}

// The FP32 step's stages read the six terms from B_PTERM, formed once per RK4 step by
// k_press_terms (a stage is bound by its load instructions: this takes thirteen FP64 neighbour
// loads and the density off each cell of each stage). The metric terms change every iteration;
// the static ones only when p_dyn, p_stat or rho_mix do: the model loop forms them after the even
// iteration's physics and keeps them for the odd iteration (STATIC false). The FP64 step forms
// the terms inside the stage (six FP64 values would cost it more bytes than the loads they save).
// One workgroup per interior ring (i 1..im-2, j 1..jm-2) in the XCD-contiguous band map, so that
// a ring's r and theta neighbour rows are read through its XCD's L2; threads along k.
constexpr int kTermThreads = 384;
template<typename T, bool STATIC>
__global__ void __launch_bounds__(kTermThreads) k_press_terms(Dims D, Scal S, Bufs<T> B, RowMap RM){
    int jr, lev, m;
    if(!RM.decode((int)blockIdx.x, 0, jr, lev, m)) return;
    const int j = 1 + jr, i = 1 + lev;
    const long long row = (long long)m * D.fs + (long long)(i * D.jm + j) * D.kp;
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    T* const* pt = B.w + B_PTERM;
    for(int k = threadIdx.x; k < D.km - 1; k += kTermThreads){
        const long long c0 = row + k;
        const long long pe = row + phi_wrap(k + 1, D.km), pw = row + phi_wrap(k - 1, D.km);
        T a, b;
        pterm_metric<T>(S, B, c0, st, pe, pw, a, b);
        if(STATIC){
            T c, d, e, f;
            pterm_static<T>(S, B, c0, sr, st, pe, pw, c, d, e, f);
            pt[PT_DPDR][c0] = c; pt[PT_DPDTHE][c0] = d; pt[PT_DPDPHI][c0] = e; pt[PT_DPSDR][c0] = f;
        }
        pt[PT_DPMDTHE][c0] = a; pt[PT_DPMDPHI][c0] = b;
    }
}

template<typename T>
void launch_press_terms(const Dims& D, const Scal& S, const Bufs<T>& B, bool with_static, hipStream_t st){
    const RowMap RM = interior_row_map(D);
    const dim3 grid(RM.blocks());
    if(with_static) hipLaunchKernelGGL((k_press_terms<T, true>), grid, dim3(kTermThreads), 0, st, D, S, B, RM);
    else            hipLaunchKernelGGL((k_press_terms<T, false>), grid, dim3(kTermThreads), 0, st, D, S, B, RM);
}

// ------------------------------------------------------------------------------------------
// Copy the boundary cells of the live fields into both stage scratch sets: the planes i = 0,
// im-1 and the rows j = 1, jm-2 that the interior stencils of stages 1-3 read. Work item y of a
// member: y < jm the two plane cells of row j = y, then the cells (1 + r/2, 1 or jm-2) of the
// rows; a thread copies its cells of every field, loading each group of fields before storing it.
template<typename T>
__global__ void __launch_bounds__(256) k_rk_halo(Dims D, Bufs<T> B, int nint){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int m = blockIdx.z;
    if(k >= D.km) return;
    long long a, b;
    if(y < D.jm){ a = D.at(0, y, k); b = D.at(D.im - 1, y, k); }
    else {
        const int r = y - D.jm, i = 1 + (r >> 1), j = (r & 1) ? D.jm - 2 : 1;
        a = D.at(i, j, k); b = -1;
    }
    const long long mo = (long long)m * D.fs;
    constexpr int G = 6;
    for(int f0 = 0; f0 < nint; f0 += G){
        double va[G], vb[G];
#pragma unroll
        for(int g = 0; g < G; g++){
            va[g] = 0.0; vb[g] = 0.0;
            if(f0 + g < nint){ va[g] = B.ld(f0 + g, mo + a); if(b >= 0) vb[g] = B.ld(f0 + g, mo + b); }
        }
#pragma unroll
        for(int g = 0; g < G; g++){
            if(f0 + g >= nint) break;
            T* dx = B.w[B_STAGE0 + f0 + g] + mo;
            T* dy = B.w[B_STAGE1 + f0 + g] + mo;
            dx[a] = (T)va[g]; dy[a] = (T)va[g];
            if(b >= 0){ dx[b] = (T)vb[g]; dy[b] = (T)vb[g]; }
        }
    }
}

template<typename T>
void launch_rk_halo(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const int nint = S.turb_active ? ATSAT_NPROG : ATSAT_NPROG - 2;
    hipLaunchKernelGGL(k_rk_halo<T>, dim3((D.km + 255) / 256, D.jm + 2 * (D.im - 2), D.M), dim3(256), 0, st, D, B, nint);
}

// ------------------------------------------------------------------------------------------
// One RK4 stage. RHSSat for the cell (RHS_Sat_Turb.cpp:25-1016), then pass B of the stage
// (RungeKutta_Sat_Turb.cpp:250-320) or, at stage 3, the final assembly (:324-375), field by
// field as each right-hand side is formed.

template<typename T>
struct Stencil {
    T d1r, d2r, d1t, d2t, d1p, d2p, c;
};

template<typename T>
__device__ __forceinline__ Stencil<T> stencil7(T fc, T frp, T frm, T ftp, T ftm, T fpe, T fpw,
                                               T inv_2dr, T inv_dr2, T inv_2dthe, T inv_dthe2,
                                               T inv_2dphi, T inv_dphi2){
    Stencil<T> s;
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

// The pressure values a cell's right-hand side needs: the six pressure terms (pterm_static,
// pterm_metric), p_stat at the cell and the cell temperature (FP64).
template<typename T>
struct PNb {
    T dpdr, dpdthe, dpdphi, dpmdthe, dpmdphi, dpstatdr;
    double ps_c, t_c;
};

// Values that only their own cell reads in a stage (y_n, the RK4 accumulator, rhs_u/v/w) and the
// stage outputs go through non-temporal loads and stores: no other workgroup of the launch reads
// them, so they are kept from evicting the neighbour rows of the stencils from L2 (about 4%
// per stage at the native ensemble).
template<typename V>
__device__ __forceinline__ V nt_load(const V* p){ return __builtin_nontemporal_load(p); }
template<typename V>
__device__ __forceinline__ void nt_store(V v, V* p){ __builtin_nontemporal_store(v, p); }

// The seven stencil values of a field at a cell as loaded: the centre, r +- 1, theta +- 1, the
// east and west phi neighbours.
template<typename T>
struct Raw7 { T c, rp, rm, tp, tm, pe, pw; };

// A field's own values at the cell: y_n, and in stages 1-3 the RK4 accumulator.
template<typename T>
struct Own { double yn; T acc; };

// The species, k and omega in the order of their right-hand sides (RHS_Sat_Turb.cpp:511-538,
// 839-937). The species' right-hand sides differ in their sources and in the order of the terms:
// SP_PRECIP the precipitation sink; SP_CHEM the chemistry mass flux (H2S); SP_NH3 the chemistry
// term before the precipitation sink; SP_NH4SH the flux limiter before the diffusion.
enum { SP_PRECIP = 0, SP_CHEM, SP_NH3, SP_NH4SH, SP_TKE, SP_DIS };
enum { SC_CH4 = 0, SC_H2O, SC_H2S, SC_NH3, SC_NH4SH };
struct SpecDef { int f, kind, s1, s2, sc; };
__host__ __device__ constexpr SpecDef spec_def(int q){
    return q == 0 ? SpecDef{ATSAT_CH4, SP_PRECIP, ATSAT_S_PRECIP_CH4, 0, SC_CH4}
         : q == 1 ? SpecDef{ATSAT_CH4_CLOUD, SP_PRECIP, ATSAT_S_PRECIP_CH4_CLOUD, 0, SC_CH4}
         : q == 2 ? SpecDef{ATSAT_CH4_ICE, SP_PRECIP, ATSAT_S_PRECIP_CH4_ICE, 0, SC_CH4}
         : q == 3 ? SpecDef{ATSAT_H2O, SP_PRECIP, ATSAT_S_PRECIP_H2O, 0, SC_H2O}
         : q == 4 ? SpecDef{ATSAT_H2O_CLOUD, SP_PRECIP, ATSAT_S_PRECIP_H2O_CLOUD, 0, SC_H2O}
         : q == 5 ? SpecDef{ATSAT_H2O_ICE, SP_PRECIP, ATSAT_S_PRECIP_H2O_ICE, 0, SC_H2O}
         : q == 6 ? SpecDef{ATSAT_H2S, SP_CHEM, ATSAT_MASSFLUX_H2S, 0, SC_H2S}
         : q == 7 ? SpecDef{ATSAT_NH3, SP_NH3, ATSAT_MASSFLUX_NH3, ATSAT_S_PRECIP_NH3, SC_NH3}
         : q == 8 ? SpecDef{ATSAT_NH3_CLOUD, SP_PRECIP, ATSAT_S_PRECIP_NH3_CLOUD, 0, SC_NH3}
         : q == 9 ? SpecDef{ATSAT_NH3_ICE, SP_PRECIP, ATSAT_S_PRECIP_NH3_ICE, 0, SC_NH3}
         : q == 10 ? SpecDef{ATSAT_NH4SH, SP_NH4SH, ATSAT_FLUXLIM_NH4SH, ATSAT_MASSFLUX_NH4SH, SC_NH4SH}
         : q == 11 ? SpecDef{ATSAT_TKE, SP_TKE, ATSAT_TKE_SOURCE, 0, 0}
         :           SpecDef{ATSAT_DIS, SP_DIS, ATSAT_DIS_SOURCE, 0, 0};
}

// RHSSat of one cell and the stage update of every field (the body shared by both stage
// kernels). RF(f) loads field f's seven stencil values at the cell, ynv(f) its y_n; c0 is the
// cell's element offset (member included).
//
// Loads are issued ahead of the stores. The compiler cannot move a field's loads above an
// earlier field's stores (it cannot prove that the buffers do not alias), so a field-by-field
// body leaves each wavefront a dozen loads in flight and then a full memory round trip per
// field. Here every field's own values and the inputs of the first species are loaded before the
// first store, and each species' inputs two species (FP64: one) ahead of its update. No cell reads in a
// stage what any cell writes in it, so the order of loads and stores does not change a value, and
// every right-hand side is the same expression in the same order as field by field.
template<typename T, int STAGE, bool TURB, typename RFn, typename YnFn>
__device__ __forceinline__ void rk_cell(const Dims& D, const Scal& S, const Bufs<T>& B, int m, int i, int j,
                                        int c0, const RFn& RF, const YnFn& ynv, const PNb<T>& P, bool clampsp){
    const int out_w = (STAGE == 0 || STAGE == 2) ? B_STAGE0 : B_STAGE1;
    const T one = (T)1.0, two = (T)2.0, zero = (T)0.0;
    const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
    const T inv_rm = (T)g[G_INV_RM], inv_rm2 = (T)g[G_INV_RM2], inv_rmsinthe = (T)g[G_INV_RMSINTHE];
    const T inv_rm2sinthe = (T)g[G_INV_RM2SINTHE], inv_rm2sinthe2 = (T)g[G_INV_RM2SINTHE2];
    const T costhe_inv_rm2sinthe = (T)g[G_COSTHE_INV_RM2SINTHE];
    const T costhe = (T)g[G_COSTHE], cotanthe = (T)g[G_COTANTHE];
    const T inv_2dr = (T)S.inv_2dr, inv_dr2 = (T)S.inv_dr2, inv_2dthe = (T)S.inv_2dthe;
    const T inv_dthe2 = (T)S.inv_dthe2, inv_2dphi = (T)S.inv_2dphi, inv_dphi2 = (T)S.inv_dphi2;
    auto stencil = [&](const Raw7<T>& r){
        return stencil7<T>(r.c, r.rp, r.rm, r.tp, r.tm, r.pe, r.pw, inv_2dr, inv_dr2, inv_2dthe, inv_dthe2,
                           inv_2dphi, inv_dphi2);
    };
    // with the closure on, tke / dis keep their own start-of-step copies: the closure runs after
    // restoreVar and changes the live fields only (cSaturnModel.cpp:1111-1116)
    auto own = [&](int f) -> Own<T> {
        Own<T> o;
        o.yn = (TURB && f >= ATSAT_TKE) ? (double)nt_load(&B.w[ATSAT_F_N0 + f][c0]) : ynv(f);
        o.acc = (STAGE == 0) ? zero : nt_load(&B.w[B_ACC + f][c0]);
        return o;
    };
    // pass B / final assembly of field f with right-hand side r (RungeKutta_Sat_Turb.cpp:254-375)
    auto update = [&](int f, T r, const Own<T>& o){
        if(STAGE < 3){
            const double c_in = (STAGE == 0 || STAGE == 1) ? 0.5 * S.dt : S.dt;
            const T wgt = (STAGE == 0) ? one : two;
            nt_store((STAGE == 0) ? wgt * r : o.acc + wgt * r, &B.w[B_ACC + f][c0]);
            double x = o.yn + c_in * (double)r;
            if(TURB && f == ATSAT_TKE){ x = x > 0.0 ? x : 0.0; x = x < S.tke_max_nd ? x : S.tke_max_nd; }
            if(TURB && f == ATSAT_DIS){ x = x > S.dis_min_nd ? x : S.dis_min_nd; }
            nt_store((T)x, &B.w[out_w + f][c0]);
        } else {
            const T a = o.acc + one * r;
            const double sixth_dt = S.dt / 6.0;
            double x = o.yn + sixth_dt * (double)a;
            if(TURB && f == ATSAT_TKE){ x = x > 0.0 ? x : 0.0; x = x < S.tke_max_nd ? x : S.tke_max_nd; }
            if(TURB && f == ATSAT_DIS){ x = x > S.dis_min_nd ? x : S.dis_min_nd; }
            // the resident step folds clampNegativeSpecies (FileIO_Sat.cpp:746-761) in here: every
            // species value the boundary conditions then write is a copy of, or a non-negative
            // multiple of, a value written here, and the radial copy re-applies it
            if(f >= ATSAT_H2O && f <= ATSAT_NH4SH && clampsp && !(x >= 0.0)) x = 0.0;
            if(prog_state(f)) nt_store(x, &B.s[f][c0]);
            else nt_store((T)x, &B.w[f][c0]);
        }
    };

    // ---- loads: the dynamics' stencils and own values, the first two species' inputs
    const bool chem = S.chem_inputs != 0;
    struct SpIn { Raw7<T> r; T s1, s2; Own<T> o; };
    auto sp_load = [&](int q) -> SpIn {
        const SpecDef d = spec_def(q);
        SpIn x;
        x.r = RF(d.f);
        x.o = own(d.f);
        x.s1 = zero; x.s2 = zero;
        if(d.kind == SP_PRECIP || d.kind == SP_TKE || d.kind == SP_DIS || chem) x.s1 = B.w[d.s1][c0];
        if(d.kind == SP_NH3 || (d.kind == SP_NH4SH && chem)) x.s2 = B.w[d.s2][c0];
        return x;
    };
    const Raw7<T> Ur = RF(ATSAT_U), Vr = RF(ATSAT_V), Wr = RF(ATSAT_W), Tr = RF(ATSAT_T);
    const Own<T> oT = own(ATSAT_T), oU = own(ATSAT_U), oV = own(ATSAT_V), oW = own(ATSAT_W);
    const T nu = TURB ? B.w[ATSAT_NUE][c0] : zero;
    const T qprec = (S.precip_coupling != 0.0) ? B.w[ATSAT_Q_PRECIP][c0] : zero;
    // two species ahead in FP32; one in FP64, whose values take twice the registers
    constexpr int ahead = sizeof(T) == 4 ? 2 : 1;
    SpIn pa = sp_load(0), pb{};
    if(ahead == 2) pb = sp_load(1);

    const Stencil<T> Us = stencil(Ur), Vs = stencil(Vr), Ws = stencil(Wr);
    const T u_ijk = Us.c, v_ijk = Vs.c, w_ijk = Ws.c;

    const T dpdr = P.dpdr, dpdthe = P.dpdthe, dpdphi = P.dpdphi;

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
        const T transport_u = u_ijk * Us.d1r + v_ijk * Us.d1t * inv_rm + w_ijk * Us.d1p * inv_rmsinthe
            - metric_on * (v_ijk * v_ijk + w_ijk * w_ijk) * inv_rm;
        const T transport_v = u_ijk * Vs.d1r + v_ijk * Vs.d1t * inv_rm + w_ijk * Vs.d1p * inv_rmsinthe
            + (u_ijk * v_ijk - w_ijk * w_ijk * cotanthe) * inv_rm;
        const T transport_w = u_ijk * Ws.d1r + v_ijk * Ws.d1t * inv_rm + w_ijk * Ws.d1p * inv_rmsinthe
            + (w_ijk * u_ijk + v_ijk * w_ijk * cotanthe) * inv_rm;
        const T diffusion_u = Us.d2r + two * u_ijk * inv_rm2 + Us.d2t * inv_rm2
            + (T)4.0 * Us.d1r * inv_rm + Us.d1t * costhe_inv_rm2sinthe + Us.d2p * inv_rm2sinthe2;
        const T diffusion_v = Vs.d2r + Vs.d1r * two_inv_rm + Vs.d2t * inv_rm2
            + Vs.d1t * costhe_inv_rm2sinthe - v_metric * v_ijk + Vs.d2p * inv_rm2sinthe2
            + two * Us.d1t * inv_rm2 - Ws.d1p * two * costhe * inv_rm2sinthe2;
        const T diffusion_w = Ws.d2r + Ws.d1r * two_inv_rm + Ws.d2t * inv_rm2
            + Ws.d1t * costhe_inv_rm2sinthe - v_metric * w_ijk + Ws.d2p * inv_rm2sinthe2
            + two * Us.d1p * inv_rm2sinthe + Vs.d1p * two * costhe * inv_rm2sinthe2;
        // the buoyancy anomaly against the level mean, formed in FP64 (:700-714)
        const double t_ijk = P.t_c;
        T buoyancy_u = zero;
        if(t_ijk > 0.0){
            buoyancy_u = (T)(- 1.0 * 1.0 * S.buoyancy
                * (S.g * P.ps_c / (S.r_mix * S.R_mix * t_ijk * S.t_ref) - B.buoy_ref[m * D.im + i]));
        }
        const T Cor = (T)S.Coriolis, Cen = (T)S.centrifugal;
        rhs_u = - dpdr + buoyancy_u - transport_u + diffusion_u / re + diffusion_u * nue_t
                - cor_rad_on * Cor * Coriolis_rad + Cen * centrifugal_rad;
        const T dpmdthe = P.dpmdthe, dpmdphi = P.dpmdphi;
        rhs_v = - dpdthe * inv_rm - dpmdthe * inv_rm - transport_v + diffusion_v / re + diffusion_v * nue_t
                - Cor * Coriolis_the + Cen * centrifugal_the;
        rhs_w = - dpdphi * inv_rmsinthe - dpmdphi * inv_rmsinthe - transport_w + diffusion_w / re
                + diffusion_w * nue_t - Cor * Coriolis_phi;
    }
    // ---- energy (:352-383, 631-665, 752-758)
    T rhs_t;
    {
        const Stencil<T> Tt = stencil(Tr);
        const T dpstatdr = P.dpstatdr;
        const T pressure_t = (T)S.coeff_energy_p * (u_ijk * (dpdr + dpstatdr)
            + v_ijk * dpdthe * inv_rm + w_ijk * dpdphi * inv_rmsinthe);
        T precip_t = zero;
        if(S.precip_coupling != 0.0){
            precip_t = (T)S.precip_coupling * qprec * (T)(S.L_atm * 1.0e3) / (T)S.precip_den;
            const T pmax = (T)0.5;
            if(!isfinite((double)precip_t)) precip_t = zero;
            else if(precip_t > pmax) precip_t = pmax;
            else if(precip_t < -pmax) precip_t = -pmax;
        }
        const T diffusion_t = diffusion(Tt);
        rhs_t = + pressure_t - transport(Tt) + diffusion_t / (T)(S.re * S.pr) + diffusion_t * nue_t_s
                - zero + zero + precip_t;
    }
    update(ATSAT_T, rhs_t, oT);
    update(ATSAT_U, rhs_u, oU);
    update(ATSAT_V, rhs_v, oV);
    update(ATSAT_W, rhs_w, oW);
    if(STAGE == 3){                                               // :1011-1013
        nt_store(rhs_u, &B.w[ATSAT_RHS_U][c0]);
        nt_store(rhs_v, &B.w[ATSAT_RHS_V][c0]);
        nt_store(rhs_w, &B.w[ATSAT_RHS_W][c0]);
        B.w[ATSAT_AUX_U][c0] = rhs_u + dpdr;
        B.w[ATSAT_AUX_V][c0] = rhs_v + dpdthe * inv_rm;
        B.w[ATSAT_AUX_W][c0] = rhs_w + dpdphi * inv_rmsinthe;
    }
    // ---- species (:839-937), then k / omega transport (:511-538)
    const T pcp = (T)S.precip_coupling, cr = (T)S.chemical_reaction;
    auto sc_of = [&](int sc) -> double {
        return sc == SC_CH4 ? S.sc_ch4 : sc == SC_H2O ? S.sc_h2o : sc == SC_H2S ? S.sc_h2s
             : sc == SC_NH3 ? S.sc_nh3 : S.sc_nh4sh;
    };
    auto sp_do = [&](int q, const SpIn& x){
        const SpecDef d = spec_def(q);
        const Stencil<T> F = stencil(x.r);
        T rhs;
        if(d.kind == SP_TKE)
            rhs = - transport(F) + diffusion(F) * (one / re + nu / (T)0.85) + x.s1;
        else if(d.kind == SP_DIS)
            rhs = - transport(F) + diffusion(F) * (one / re + nu / (T)0.5) + x.s1;
        else {
            const T dif = diffusion(F);
            const T scre = (T)sc_of(d.sc) * re;
            if(d.kind == SP_PRECIP)   rhs = - transport(F) + dif / scre + dif * nue_t_s + pcp * x.s1;
            else if(d.kind == SP_CHEM) rhs = - transport(F) + dif / scre + dif * nue_t_s + cr * x.s1;
            else if(d.kind == SP_NH3)  rhs = - transport(F) + dif / scre + dif * nue_t_s + cr * x.s1 + pcp * x.s2;
            else                       rhs = - transport(F) + x.s1 + dif / scre + dif * nue_t_s + cr * x.s2;
        }
        update(d.f, rhs, x.o);
    };
    constexpr int nsp = TURB ? 13 : 11;
#pragma unroll
    for(int q = 0; q < nsp; q++){
        SpIn pc{};
        if(q + ahead < nsp) pc = sp_load(q + ahead);
        sp_do(q, pa);
        if(ahead == 2){ pa = pb; pb = pc; }
        else pa = pc;
    }
}

// The per-cell form of the stage: one thread per cell, every stencil value loaded from memory
// (neighbours through L1 / L2). A workgroup holds whole longitude rings (i, j): the k
// neighbours, the seam wrap included, are loaded by the workgroup itself. The row map (RowMap)
// keeps the j and r neighbours of a ring inside the L2 of its XCD; only the rows at a band edge
// miss it.
//
// One cell of the stage, for workgroup `wg` of the row map and lane `k` of its ring `sub`.
template<typename T, int STAGE, bool TURB>
__device__ __forceinline__ void rk_cell_at(const Dims& D, const Scal& S, const Bufs<T>& B, const RowMap& RM,
                                           int wg, int sub, int k, int yn0, int clampsp){
    int jr, lev, m;
    if(!RM.decode(wg, sub, jr, lev, m)) return;
    const int j = 2 + jr;
    const int i = 1 + lev;
    if(k >= D.km - 1) return;
    const int mo = m * (int)D.fs;
    const int c0 = mo + (i * D.jm + j) * D.kp + k;
    const int sr = D.jm * D.kp, st = D.kp;
    const int pe = mo + (i * D.jm + j) * D.kp + phi_wrap(k + 1, D.km);
    const int pw = mo + (i * D.jm + j) * D.kp + phi_wrap(k - 1, D.km);
    const int in_w = (STAGE == 1 || STAGE == 3) ? B_STAGE0 : B_STAGE1;
    // the stencil input: stage 0 the live state (t and w FP64, the others T), the later stages
    // the scratch set X or Y (T)
    auto RF = [&](int f) -> Raw7<T> {
        if(STAGE == 0 && prog_state(f)){
            const double* F = B.s[f];
            return Raw7<T>{(T)F[c0], (T)F[c0 + sr], (T)F[c0 - sr], (T)F[c0 + st], (T)F[c0 - st], (T)F[pe], (T)F[pw]};
        }
        const T* F = B.w[(STAGE == 0 ? 0 : in_w) + f];
        return Raw7<T>{F[c0], F[c0 + sr], F[c0 - sr], F[c0 + st], F[c0 - st], F[pe], F[pw]};
    };
    auto ynv = [&](int f) -> double {
        return prog_state(f) ? nt_load(&B.s[yn0 + f][c0]) : (double)nt_load(&B.w[yn0 + f][c0]);
    };
    const double t_c = (STAGE == 0) ? B.s[ATSAT_T][c0] : (double)B.w[in_w + ATSAT_T][c0];
    PNb<T> P;
    if constexpr(sizeof(T) == 4){
        const T* const* pt = B.w + B_PTERM;
        P = PNb<T>{pt[PT_DPDR][c0], pt[PT_DPDTHE][c0], pt[PT_DPDPHI][c0], pt[PT_DPMDTHE][c0], pt[PT_DPMDPHI][c0],
                   pt[PT_DPSDR][c0], B.s[ATSAT_P_STAT][c0], t_c};
    } else {
        pterm_static<T>(S, B, c0, sr, st, pe, pw, P.dpdr, P.dpdthe, P.dpdphi, P.dpstatdr);
        pterm_metric<T>(S, B, c0, st, pe, pw, P.dpmdthe, P.dpmdphi);
        P.ps_c = B.s[ATSAT_P_STAT][c0];
        P.t_c = t_c;
    }
    rk_cell<T, STAGE, TURB>(D, S, B, m, i, j, c0, RF, ynv, P, clampsp != 0);
}

// rt: threads per ring (km - 1 rounded up to a wavefront), a multiple of 64, so that every
// wavefront belongs to one ring.
template<typename T, int STAGE, bool TURB>
__global__ void __launch_bounds__(kRowThreads) k_rk_cell(Dims D, Scal S, Bufs<T> B, RowMap RM, int rt, int yn0, int clampsp){
    // rt is a multiple of the wavefront, so the ring is the same for every lane of a wave: said so
    // explicitly, the row decode and the ring's geometry loads stay scalar
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) / (rt >> 6);
    rk_cell_at<T, STAGE, TURB>(D, S, B, RM, (int)blockIdx.x, sub, (int)threadIdx.x - sub * rt, yn0, clampsp);
}

// The stage of several grids -- different shapes, each with its own members and tile -- in one
// launch. Grid g owns workgroups [first_block, first_block + its blocks); first_block is a
// multiple of 8 so that each grid's XCD-contiguous row map holds. All grids share the
// workgroup size; a grid with shorter rings puts several rings in one workgroup.
// A pointer read from memory has no known address space, and accesses through it compile to
// flat loads and stores. Re-forming it from its address as a global-space pointer lets the
// compiler's address-space inference turn them into global ones.
template<typename P>
__device__ __forceinline__ P* global_ptr(P* p){
    return (P*)((__attribute__((address_space(1))) P*)(uintptr_t)p);
}

// Workgroup-sized item b of the batched stage: the grid whose range of the launch holds b, then
// that grid's workgroup b - first_block.
template<typename T, int STAGE, bool TURB>
__device__ __forceinline__ void rk_group_item(const RkGrid<T>* __restrict__ G, int ng, int b, unsigned alias, int clampsp){
    // the table is read-only for the whole launch: read through the constant address space, its
    // loads are invariant and the compiler reloads values near their uses instead of holding
    // every buffer pointer in scalar registers (and spilling them) for the whole kernel
#if defined(__HIP_DEVICE_COMPILE__)
    using CGrid = const __attribute__((address_space(4))) RkGrid<T>;
#else
    using CGrid = const RkGrid<T>;                // the host pass only needs the kernel's stub
#endif
    CGrid* Gc = (CGrid*)(uintptr_t)G;
    int g = 0;
    for(int q = 1; q < ng; q++) if(b >= Gc[q].first_block) g = q;
    CGrid& R = Gc[g];
    Bufs<T> B;
    #pragma unroll
    for(int f = 0; f < NBUF; f++){ B.s[f] = global_ptr(R.B.s[f]); B.w[f] = global_ptr(R.B.w[f]); }
    B.geo = global_ptr(R.B.geo);
    B.thez_sin = global_ptr(R.B.thez_sin);
    B.sin_abs = global_ptr(R.B.sin_abs);
    B.rm = global_ptr(R.B.rm);
    B.vbar0 = global_ptr(R.B.vbar0);
    B.ubar0 = global_ptr(R.B.ubar0);
    B.buoy_ref = global_ptr(R.B.buoy_ref);
    B.srf = global_ptr(R.B.srf);
    B.counter = global_ptr(R.B.counter);
    B.buoy_x = global_ptr(R.B.buoy_x);
    B.buoy_ch = R.B.buoy_ch;
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) / (R.rt >> 6);   // wave-uniform
    if(sub >= R.RM.rings) return;
    rk_cell_at<T, STAGE, TURB>(R.D, R.S, B, R.RM, b - R.first_block, sub,
                               (int)threadIdx.x - sub * R.rt, ((alias >> g) & 1u) ? 0 : ATSAT_F_N0, clampsp);
}

template<typename T, int STAGE, bool TURB>
__global__ void __launch_bounds__(kRowThreads) k_rk_group(const RkGrid<T>* __restrict__ G, int ng, unsigned alias,
                                                          int clampsp){
    rk_group_item<T, STAGE, TURB>(G, ng, (int)blockIdx.x, alias, clampsp);
}

// The batched stage as a persistent launch over a work queue: a fixed set of workgroups takes the
// batched launch's items b = 0..total-1 from eight queues, queue q holding the items q, q + 8,
// q + 16, ... that the XCD-contiguous row maps meant for XCD q. A workgroup starts on the queue of
// the XCD it runs on (the XCC_ID hardware register) and then takes from the others, so every item
// runs exactly once and every workgroup ends when all eight queues are drained, whatever XCD the
// hardware reports. The heads are cleared by k_queue_reset before each launch.
constexpr int kQueueHwreg = (4 - 1) << 11 | 20;             // hwreg(HW_REG_XCC_ID, 0, 4)
__global__ void __launch_bounds__(64) k_queue_reset(unsigned* __restrict__ heads){
    if(threadIdx.x < kQueues) heads[threadIdx.x] = 0u;
}

template<typename T, int STAGE, bool TURB>
__global__ void __launch_bounds__(kRowThreads) k_rk_queue(const RkGrid<T>* __restrict__ G, int ng, int total,
                                                          unsigned alias, int clampsp, unsigned* __restrict__ heads){
    __shared__ int taken;
    const int xcd = (int)(__builtin_amdgcn_s_getreg(kQueueHwreg) & (kQueues - 1));
    for(int qq = 0; qq < kQueues; qq++){
        const int q = (xcd + qq) & (kQueues - 1);
        const int nq = (total - q + kQueues - 1) / kQueues;  // items q + 8 n < total
        for(;;){
            if(threadIdx.x == 0) taken = (int)atomicAdd(&heads[q], 1u);
            __syncthreads();
            const int n = taken;
            __syncthreads();
            if(n >= nq) break;
            rk_group_item<T, STAGE, TURB>(G, ng, q + kQueues * n, alias, clampsp);
        }
    }
}

template<typename T, int STAGE>
static void launch_stage(const Dims& D, const Scal& S, const Bufs<T>& B, const RkTile& tile, int yn0, int clampsp,
                         hipStream_t st){
    const RowMap RM = rk_row_map(D, tile);
    const int rt = rk_ring_threads(D);
    const dim3 grid(RM.blocks()), block(rt * RM.rings);
    if(S.turb_active) hipLaunchKernelGGL((k_rk_cell<T, STAGE, true>), grid, block, 0, st, D, S, B, RM, rt, yn0, clampsp);
    else hipLaunchKernelGGL((k_rk_cell<T, STAGE, false>), grid, block, 0, st, D, S, B, RM, rt, yn0, clampsp);
}

int rk_ring_threads(const Dims& D){ return (D.km - 1 + 63) / 64 * 64; }

RowMap rk_row_map(const Dims& D, const RkTile& tile){
    RowMap RM{D.jm - 4, D.im - 2, D.M, 0, 0, 0, 1};
    RM.set(tile.band, tile.rings);
    return RM;
}

bool rk_tile_ok(const Dims& D, const RkTile& tile){
    return tile.rings >= 1 && tile.band >= tile.rings && tile.band % tile.rings == 0
        && rk_ring_threads(D) * tile.rings <= kRowThreads;
}

// Before the autotuner has run: two rings per workgroup where they fit (they share their
// theta-neighbour rows in L1), bands of four rows; one ring, bands of eight rows otherwise.
RkTile rk_default_tile(const Dims& D){
    const RkTile two{4, 2};
    return rk_tile_ok(D, two) ? two : RkTile{8, 1};
}

template<typename T>
void launch_rk_stage(const Dims& D, const Scal& S, const Bufs<T>& B, const RkTile& tile, int stage, bool alias,
                     bool clamp, hipStream_t st){
    const int yn0 = alias ? 0 : ATSAT_F_N0;
    switch(stage){
    case 0: launch_stage<T, 0>(D, S, B, tile, yn0, 0, st); break;
    case 1: launch_stage<T, 1>(D, S, B, tile, yn0, 0, st); break;
    case 2: launch_stage<T, 2>(D, S, B, tile, yn0, 0, st); break;
    default: launch_stage<T, 3>(D, S, B, tile, yn0, clamp ? 1 : 0, st); break;
    }
}

template<typename T, int STAGE>
static void launch_group_stage(const RkGrid<T>* G, int ng, int blocks, int threads, bool turb, unsigned alias,
                               int clampsp, hipStream_t st){
    if(turb) hipLaunchKernelGGL((k_rk_group<T, STAGE, true>), dim3(blocks), dim3(threads), 0, st, G, ng, alias, clampsp);
    else hipLaunchKernelGGL((k_rk_group<T, STAGE, false>), dim3(blocks), dim3(threads), 0, st, G, ng, alias, clampsp);
}

// Workgroups of the persistent launch: as many as the CUs hold at once (the kernel's occupancy at
// this width, from the runtime), and never more than there are items.
static int queue_workgroups(const void* kernel, int threads, int total){
    int dev = 0, cus = 0, per = 0;
    if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
    const long long w = (long long)cus * per;
    return (int)(w < total ? w : total);
}

template<typename T, int STAGE>
static void launch_queue_stage(const RkGrid<T>* G, int ng, int total, int threads, bool turb, unsigned alias,
                               int clampsp, unsigned* heads, hipStream_t st){
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, heads);
    if(turb){
        const int w = queue_workgroups((const void*)k_rk_queue<T, STAGE, true>, threads, total);
        hipLaunchKernelGGL((k_rk_queue<T, STAGE, true>), dim3(w), dim3(threads), 0, st, G, ng, total, alias, clampsp, heads);
    } else {
        const int w = queue_workgroups((const void*)k_rk_queue<T, STAGE, false>, threads, total);
        hipLaunchKernelGGL((k_rk_queue<T, STAGE, false>), dim3(w), dim3(threads), 0, st, G, ng, total, alias, clampsp, heads);
    }
}

template<typename T>
void launch_rk_queue(const RkGrid<T>* G, int ng, int total, int threads, bool turb, int stage, unsigned alias,
                     bool clamp, unsigned* heads, hipStream_t st){
    switch(stage){
    case 0: launch_queue_stage<T, 0>(G, ng, total, threads, turb, alias, 0, heads, st); break;
    case 1: launch_queue_stage<T, 1>(G, ng, total, threads, turb, alias, 0, heads, st); break;
    case 2: launch_queue_stage<T, 2>(G, ng, total, threads, turb, alias, 0, heads, st); break;
    default: launch_queue_stage<T, 3>(G, ng, total, threads, turb, alias, clamp ? 1 : 0, heads, st); break;
    }
}

template<typename T>
void launch_rk_group(const RkGrid<T>* G, int ng, int blocks, int threads, bool turb, int stage, unsigned alias,
                     bool clamp, hipStream_t st){
    switch(stage){
    case 0: launch_group_stage<T, 0>(G, ng, blocks, threads, turb, alias, 0, st); break;
    case 1: launch_group_stage<T, 1>(G, ng, blocks, threads, turb, alias, 0, st); break;
    case 2: launch_group_stage<T, 2>(G, ng, blocks, threads, turb, alias, 0, st); break;
    default: launch_group_stage<T, 3>(G, ng, blocks, threads, turb, alias, clamp ? 1 : 0, st); break;
    }
}

#define INST(T) \
    template void launch_buoy_ref<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_metric_pressure<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_press_terms<T>(const Dims&, const Scal&, const Bufs<T>&, bool, hipStream_t); \
    template void launch_rk_halo<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_rk_stage<T>(const Dims&, const Scal&, const Bufs<T>&, const RkTile&, int, bool, bool, hipStream_t); \
    template void launch_rk_group<T>(const RkGrid<T>*, int, int, int, bool, int, unsigned, bool, hipStream_t); \
    template void launch_rk_queue<T>(const RkGrid<T>*, int, int, int, bool, int, unsigned, bool, unsigned*, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
