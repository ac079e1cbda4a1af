// The dynamic-pressure solve on the device (reference mode: the reference's red-black sweeps)
// and the p_dyn Shapiro filter. p_dyn and the solver's scratch are FP64 in every precision; the
// source reads the last-stage tendencies rhs_u/v/w in the work precision.
//
// Reference: PressureSolver<cSaturnModel>::run (planet/PressureSolver.h:145-766) with
// PRESS_SRC 1, PRESS_INCR 1, no obstacle, rigid lid; cSaturnModel::prepareProjectionBoundaries
// (planet/Pressure_Sat.cpp:78-194); AtomUtils::damp_wiggles (lib/Utils.cpp:229-325).
//
// The source of every cell -- the clamped divergence of rhs_* plus Lc(p_dyn at entry) -- does
// not change during a solve (rhs_* are not touched by it), so it is formed once (press_src)
// instead of once per colour pass as the reference does; the value is the same. Each colour
// pass then reads 6 neighbours and its source and writes its cell, in place, exactly the
// Gauss-Seidel order of the reference: a pass writes only cells of its colour and reads only the
// other colour.
#include "atsat_dev.hpp"
#include "atsat_gpu.h"
#include "restart.hpp"

#include <mutex>
#include <string>

namespace atsat {

// prepareProjectionBoundaries, rigid lid. Pass 1: the radial planes of every interior (j, k)
// column and the polar rows of every interior (i, k) row -- disjoint cells, no cell written here
// is read here. Pass 2: the seam copy k = km-1 <- k = 0 of every (i, j).
template<typename T>
__global__ void __launch_bounds__(256) k_press_prepare(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km - 1) return;
    const long long mo = m * D.fs;
    T* au = B.w[ATSAT_AUX_U] + mo; T* av = B.w[ATSAT_AUX_V] + mo; T* aw = B.w[ATSAT_AUX_W] + mo;
    T* ru = B.w[ATSAT_RHS_U] + mo; T* rv = B.w[ATSAT_RHS_V] + mo; T* rw = B.w[ATSAT_RHS_W] + mo;
    const T three = (T)3.0, zero = (T)0.0;
    auto cub = [&](T* F, long long a, long long b, long long c){ return F[c] - three * F[b] + three * F[a]; };
    const int im = D.im, jm = D.jm;
    if(j >= 1 && j <= jm - 2){                                    // r-direction, :86-123
        const long long n0 = D.at(0, j, k), n1 = D.at(1, j, k), n2 = D.at(2, j, k), n3 = D.at(3, j, k);
        const long long t0 = D.at(im - 1, j, k), t1 = D.at(im - 2, j, k), t2 = D.at(im - 3, j, k), t3 = D.at(im - 4, j, k);
        au[n0] = zero; au[t0] = zero;
        T* Fs[5] = { av, aw, ru, rv, rw };
        for(int q = 0; q < 5; q++){ T* F = Fs[q]; F[n0] = cub(F, n1, n2, n3); F[t0] = cub(F, t1, t2, t3); }
    }
    if(j == 0){                                                   // theta-direction, :126-148
        for(int i = 1; i < im - 1; i++){
            const long long n0 = D.at(i, 0, k), n1 = D.at(i, 1, k), n2 = D.at(i, 2, k), n3 = D.at(i, 3, k);
            const long long s0 = D.at(i, jm - 1, k), s1 = D.at(i, jm - 2, k), s2 = D.at(i, jm - 3, k), s3 = D.at(i, jm - 4, k);
            au[n0] = cub(au, n1, n2, n3); av[n0] = zero; aw[n0] = zero;
            au[s0] = cub(au, s1, s2, s3); av[s0] = zero; aw[s0] = zero;
            ru[n0] = cub(ru, n1, n2, n3); rv[n0] = zero; rw[n0] = zero;
            ru[s0] = cub(ru, s1, s2, s3); rv[s0] = zero; rw[s0] = zero;
        }
    }
}

template<typename T>
__global__ void __launch_bounds__(256) k_seam_copy6(Dims D, Bufs<T> B){
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, m = blockIdx.z;
    if(j >= D.jm) return;
    const long long mo = m * D.fs;
    const int ids[6] = { ATSAT_AUX_U, ATSAT_AUX_V, ATSAT_AUX_W, ATSAT_RHS_U, ATSAT_RHS_V, ATSAT_RHS_W };
    for(int q = 0; q < 6; q++){ T* F = B.w[ids[q]] + mo; F[D.at(i, j, D.km - 1)] = F[D.at(i, j, 0)]; }
}

template<typename T>
void launch_press_prepare(const Dims& D, const Scal&, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_press_prepare<T>, dim3((D.km - 1 + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_seam_copy6<T>, dim3((D.jm + 255) / 256, D.im, D.M), dim3(256), 0, st, D, B);
}

// The source of every relaxed cell (i 1..im-2, j 1..jm-2, k 0..km-2): the clamped divergence
// of rhs_* (PressureSolver.h:449-487, S_u zero on the radial walls :246-248) plus
// Lc(p_dyn at entry) (:249-282, :488).
template<typename T>
__global__ void __launch_bounds__(256) k_press_src(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    const int i = 1 + (blockIdx.z % (D.im - 2)), m = blockIdx.z / (D.im - 2);
    if(k >= D.km - 1) return;
    const long long mo = m * D.fs;
    const double* p = B.s[ATSAT_P_DYN] + mo;
    const T* Su = B.w[ATSAT_RHS_U] + mo; const T* Sv = B.w[ATSAT_RHS_V] + mo; const T* Sw = B.w[ATSAT_RHS_W] + mo;
    const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
    const long long c0 = D.at(i, j, k), sr = (long long)D.jm * D.kp, st = D.kp;
    const long long pe = D.at(i, j, phi_wrap(k + 1, D.km)), pw = D.at(i, j, phi_wrap(k - 1, D.km));
    // Lc(p_entry) with the weights :255-259
    const double lnum1 = 1.0 * 1.0 * S.inv_dr2;
    const double lnum2 = g[G_P_LC_NUM2], lnum3 = g[G_P_LC_NUM3];
    const double p0 = p[c0];
    const double l_ = lnum1 * (p[c0 + sr] + p[c0 - sr] - 2.0 * p0)
                    + lnum2 * (p[c0 + st] + p[c0 - st] - 2.0 * p0)
                    + lnum3 * (p[pe] + p[pw] - 2.0 * p0);
    // divergence of S = rhs_*, :449-459, the wall-normal source zero on the radial walls :246-248
    const double su_p = (i + 1 == D.im - 1) ? 0.0 : (double)Su[c0 + sr];
    const double su_m = (i - 1 == 0) ? 0.0 : (double)Su[c0 - sr];
    const double du_dr = (su_p - su_m) * S.inv_2dr;
    const double dv_dthe = ((double)Sv[c0 + st] - (double)Sv[c0 - st]) * S.inv_2dthe;
    const double dw_dphi = ((double)Sw[pe] - (double)Sw[pw]) * S.inv_2dphi;
    double div_src = du_dr * 1.0 + dv_dthe * g[G_P_INV_RM] + dw_dphi * g[G_P_INV_RMSINTHE];
    const double den = g[G_P_DENOM];
    const double src_max = den * S.pdyn_cap;                      // :484-487
    if(!isfinite(div_src)) div_src = 0.0;
    else if(div_src > src_max) div_src = src_max;
    else if(div_src < -src_max) div_src = -src_max;
    div_src += l_;
    B.s[B_PSRC][mo + c0] = div_src;
}

template<typename T>
void launch_press_src(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_press_src<T>, dim3((D.km - 1 + 255) / 256, D.jm - 2, (D.im - 2) * D.M), dim3(256), 0, st, D, S, B);
}

// One colour pass, :322-495 and :476-491: p = (sum num * p_nb - src) / den.
template<typename T>
__global__ void __launch_bounds__(256) k_press_sweep(Dims D, Scal S, Bufs<T> B, int colour){
    const int kh = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    const int i = 1 + (blockIdx.z % (D.im - 2)), m = blockIdx.z / (D.im - 2);
    const int k = 2 * kh + ((colour + i + j) & 1);
    if(k >= D.km - 1) return;
    const long long mo = m * D.fs;
    double* p = B.s[ATSAT_P_DYN] + mo;
    const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
    const long long c0 = D.at(i, j, k), sr = (long long)D.jm * D.kp, st = D.kp;
    const long long pe = D.at(i, j, phi_wrap(k + 1, D.km)), pw = D.at(i, j, phi_wrap(k - 1, D.km));
    const double num1 = 1.0 * S.inv_dr2, num2 = g[G_P_NUM2], num3 = g[G_P_NUM3];
    const double acc = (p[c0 + sr] + p[c0 - sr]) * num1 + (p[c0 + st] + p[c0 - st]) * num2 + (p[pe] + p[pw]) * num3;
    p[c0] = (acc - B.s[B_PSRC][mo + c0]) / g[G_P_DENOM];
}

template<typename T>
void launch_press_sweep(const Dims& D, const Scal& S, const Bufs<T>& B, int colour, hipStream_t st){
    const int nh = (D.km - 1 + 1) / 2;
    hipLaunchKernelGGL(k_press_sweep<T>, dim3((nh + 255) / 256, D.jm - 2, (D.im - 2) * D.M), dim3(256), 0, st, D, S, B, colour);
}

// The outer boundary pass after each sweep, :718-740: 2-point Neumann on the radial planes, then
// on the polar rows (which read the radial planes just set), then the seam copy.
template<typename T>
__global__ void __launch_bounds__(256) k_press_bnd_r(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    double* p = B.s[ATSAT_P_DYN] + m * D.fs;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    p[D.at(0, j, k)] = c43 * p[D.at(1, j, k)] - c13 * p[D.at(2, j, k)];
    p[D.at(D.im - 1, j, k)] = c43 * p[D.at(D.im - 2, j, k)] - c13 * p[D.at(D.im - 3, j, k)];
}
template<typename T>
__global__ void __launch_bounds__(256) k_press_bnd_t(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    double* p = B.s[ATSAT_P_DYN] + m * D.fs;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    p[D.at(i, 0, k)] = c43 * p[D.at(i, 1, k)] - c13 * p[D.at(i, 2, k)];
    p[D.at(i, D.jm - 1, k)] = c43 * p[D.at(i, D.jm - 2, k)] - c13 * p[D.at(i, D.jm - 3, k)];
}
template<typename T>
__global__ void __launch_bounds__(256) k_seam_copy1(Dims D, Bufs<T> B, int field){
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, m = blockIdx.z;
    if(j >= D.jm) return;
    double* p = B.s[field] + m * D.fs;
    p[D.at(i, j, D.km - 1)] = p[D.at(i, j, 0)];
}

template<typename T>
void launch_press_boundary(const Dims& D, const Scal&, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_press_bnd_r<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_press_bnd_t<T>, dim3((D.km + 255) / 256, D.im, D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_seam_copy1<T>, dim3((D.jm + 255) / 256, D.im, D.M), dim3(256), 0, st, D, B, (int)ATSAT_P_DYN);
}

// damp_wiggles(p_dyn, nullptr, along_i, !along_j, along_k): the 1-2-1 pass along the periodic
// phi ring into the scratch (the seam cell k = km-1 takes the value of k = 0), then along r with
// clamped ends back into p_dyn.
template<typename T>
__global__ void __launch_bounds__(256) k_damp_k(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    const int i = im_m % D.im, m = im_m / D.im;
    if(k >= D.km) return;
    const long long mo = m * D.fs;
    const double* p = B.s[ATSAT_P_DYN] + mo;
    double* tmp = B.s[B_TMP] + mo;
    const int kk = (k == D.km - 1) ? 0 : k;                       // :272, the seam copy
    const int km1 = (kk == 0 ? D.km - 2 : kk - 1);
    const int kp1 = (kk == D.km - 2 ? 0 : kk + 1);
    const double c = 1.0 * 0.25;
    const double f0 = p[D.at(i, j, kk)];
    tmp[D.at(i, j, k)] = f0 + c * (p[D.at(i, j, km1)] - 2.0 * f0 + p[D.at(i, j, kp1)]);
}
template<typename T>
__global__ void __launch_bounds__(256) k_damp_i(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    const int i = im_m % D.im, m = im_m / D.im;
    if(k >= D.km) return;
    const long long mo = m * D.fs;
    double* p = B.s[ATSAT_P_DYN] + mo;
    const double* tmp = B.s[B_TMP] + mo;
    const int im1 = i - 1 < 0 ? 0 : i - 1;
    const int ip1 = i + 1 > D.im - 1 ? D.im - 1 : i + 1;
    const double c = 1.0 * 0.25;
    const double f0 = tmp[D.at(i, j, k)];
    p[D.at(i, j, k)] = f0 + c * (tmp[D.at(im1, j, k)] - 2.0 * f0 + tmp[D.at(ip1, j, k)]);
}

template<typename T>
void launch_damp_pdyn(const Dims& D, const Scal&, const Bufs<T>& B, hipStream_t st){
    const dim3 grid((D.km + 255) / 256, D.jm, D.im * D.M);
    hipLaunchKernelGGL(k_damp_k<T>, grid, dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_damp_i<T>, grid, dim3(256), 0, st, D, B);
}

// ---- the p_dyn filter on caller memory, with no allocation ---------------------------------
// atsat_damp_pdyn_noalloc runs the filter's two kernels on a dense [members][im][jm][km] p_dyn the
// caller owns (row pitch km). The intermediate field lives in `scratch`, a file-scope device array
// that every GPU holds once, from the library's load until the process ends. In host code
// `scratch` names a host-side stand-in, not the device array: its device address is taken per
// device with hipGetSymbolAddress on the first call there and only that address is passed on.
constexpr size_t kDampScratchDoubles = size_t(1) << 23;     // 64 MiB: three native members
static __device__ __align__(256) double scratch[kDampScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_device(double** out, std::string* err){
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if(e != hipSuccess){ *err = std::string("hipGetDevice: ") + hipGetErrorString(e); return ATSAT_E_HIP; }
    if(dev < 0 || dev >= kMaxDevices){ *err = "device ordinal beyond the library's table"; return ATSAT_E_UNSUPPORTED; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(!g_scratch_ptr[dev]){
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p){
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return ATSAT_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ATSAT_OK;
}
}  // namespace

}  // namespace atsat

extern "C" int atsat_damp_pdyn_noalloc(const atsat_grid* g, int n_members, double* p_dyn, void* stream){
    using namespace atsat;
    if(!g || !p_dyn || n_members < 1){ set_last_error("atsat_damp_pdyn_noalloc: null argument or n_members < 1"); return ATSAT_E_ARG; }
    if(g->im < 3 || g->jm < 1 || g->km < 3){ set_last_error("atsat_damp_pdyn_noalloc: grid too small (im >= 3, km >= 3)"); return ATSAT_E_ARG; }
    const size_t n = (size_t)n_members * g->im * g->jm * g->km;
    if(n > kDampScratchDoubles){
        set_last_error("atsat_damp_pdyn_noalloc: " + std::to_string(n) + " values exceed the scratch of "
                       + std::to_string(kDampScratchDoubles) + " (members x im x jm x km)");
        return ATSAT_E_ARG;
    }
    if((long long)g->im * n_members > 65535 || g->jm > 65535){
        set_last_error("atsat_damp_pdyn_noalloc: im x members or jm above the launch grid's limit of 65535");
        return ATSAT_E_ARG;
    }
    double* tmp = nullptr;
    std::string err;
    if(int rc = scratch_on_device(&tmp, &err)){ set_last_error("atsat_damp_pdyn_noalloc: " + err); return rc; }
    Dims D{};
    D.im = g->im; D.jm = g->jm; D.km = g->km; D.kp = g->km; D.M = n_members;
    D.fs = (long long)D.im * D.jm * D.kp;
    Bufs<double> B{};
    B.s[ATSAT_P_DYN] = p_dyn;
    B.s[B_TMP] = tmp;
    launch_damp_pdyn<double>(D, Scal{}, B, static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    if(e != hipSuccess){ set_last_error(std::string("atsat_damp_pdyn_noalloc: ") + hipGetErrorString(e)); return ATSAT_E_HIP; }
    return ATSAT_OK;
}

namespace atsat {

#define INST(T) \
    template void launch_press_prepare<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_press_src<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_press_sweep<T>(const Dims&, const Scal&, const Bufs<T>&, int, hipStream_t); \
    template void launch_press_boundary<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_damp_pdyn<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
