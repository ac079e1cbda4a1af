// The optional modes of the red-black pressure solver (PressureSolver<cSaturnModel>::run), off by
// default as in the reference:
//
//   ATSAT_PRESS_LINE (PressureSolver.h:502-570, 750-755): after each sweep, per colour (j + k) & 1,
//     every column's rows i = 1..im-2 solved by Thomas with the theta / phi neighbours lagged and
//     both radial ends folded (p0 = c43 p1 - c13 p2); after the boundary pass the domain mean of
//     p_dyn is held at its value on entry;
//   ATSAT_PRESS_BAROTROPIC (:573-708): after each sweep (and the line pass) the column-mean
//     residual R(j, k) = sum_i (src - Lc p) / A2, its DFT over the N = km - 1 meridians (modes of
//     round-off size against a zonally symmetric row dropped), per mode a tridiagonal in theta with
//     folded ends (mode 0 made consistent, pinned at j = 1, shifted to a zero sin-weighted mean),
//     and the inverse DFT added to every level of the column.
//
// Kernels: k_press_line (a thread per column of one colour), k_baro_resid (a thread per column),
// k_baro_dft (a thread per mode and row), k_baro_tri (a thread per mode and cos / sin part),
// k_baro_apply (a thread per column), k_pmean_rows / k_pmean_total / k_pmean_shift (the mean:
// every (i, j) row summed over k, then the rows of a member in (i, j) order by one thread; the
// reference's OpenMP reduction has no fixed order). Every sum runs in the reference's order
// otherwise; the DFT tables, sin(pi m / N) and the guarded sin(theta) come from the host libm,
// as in the reference. FP64 throughout (p_dyn is of the state class).
#include <cmath>
#include <vector>

#include "atsat_dev.hpp"

namespace atsat {

namespace {
constexpr int kLineMax = 64;                  // rows of a column the line pass holds
constexpr double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;

struct PmLayout {
    long long N, J, NM, ct, st, lam, spg, mean0, shift, rows, R, Ac, As, cpb, dpb, total;
};
__host__ __device__ inline PmLayout pm_layout(const Dims& D){
    PmLayout L;
    L.N = D.km - 1; L.J = D.jm - 2; L.NM = L.N / 2;
    const long long M = D.M, modes = L.NM + 1;
    L.ct = 0; L.st = L.ct + modes * L.N; L.lam = L.st + modes * L.N; L.spg = L.lam + modes;
    L.mean0 = L.spg + D.jm; L.shift = L.mean0 + M; L.rows = L.shift + M;
    L.R = L.rows + M * (D.im - 2) * L.J;
    L.Ac = L.R + M * L.J * L.N; L.As = L.Ac + M * modes * L.J;
    L.cpb = L.As + M * modes * L.J; L.dpb = L.cpb + M * modes * 2 * L.J;
    L.total = L.dpb + M * modes * 2 * L.J;
    return L;
}
}  // namespace

size_t pmodes_doubles(const Dims& D){ return (size_t)pm_layout(D).total; }
int pmodes_max_levels(){ return kLineMax; }

// the host tables: cos / sin of 2 pi (m k mod N) / N, lambda_m = 4 sin^2(pi m / N) / dphi^2, and
// sin(theta_j) with the solver's guard (PressureSolver.h:156-162, 617-621, 643-644). The cos / sin
// pair of one angle is one sincos call, as the reference's GCC build evaluates it (GCC merges the
// two calls); glibc's sincos and its separate sin / cos differ in the last bit for some angles.
int pmodes_store_tables(const Dims& D, const Scal& S, const std::vector<double>& thez_sin, double* pm){
    const PmLayout L = pm_layout(D);
    std::vector<double> h((size_t)L.mean0);
    for(long long mm = 0; mm <= L.NM; mm++){
        for(long long k = 0; k < L.N; k++){
            const double a = 2.0 * M_PI * (double)((long)mm * k % L.N) / L.N;
            double sa, ca;
            ::sincos(a, &sa, &ca);
            h[L.ct + mm * L.N + k] = ca;
            h[L.st + mm * L.N + k] = sa;
        }
        const double sl = std::sin(M_PI * mm / L.N);
        h[L.lam + mm] = 4.0 * sl * sl * S.inv_dphi2;
    }
    for(int j = 0; j < D.jm; j++){
        double sp = thez_sin[j];
        if(sp < 0.0) sp = 0.0;
        if(sp == 0.0) sp = 1.0e-5;
        h[L.spg + j] = sp;
    }
    return hipMemcpy(pm, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

// ---- the line pass ------------------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_press_line(Dims D, Scal S, Bufs<T> B, int colour){
    const int kh = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y, m = blockIdx.z;
    const int k = 2 * kh + ((colour + j) & 1);                    // (j + k) & 1 == colour
    if(k >= D.km - 1) return;
    const long long mo = (long long)m * D.fs;
    double* p = B.s[ATSAT_P_DYN] + mo;
    const double* src = B.s[B_PSRC] + mo;
    const long long sr = (long long)D.jm * D.kp, st = D.kp;
    const int ke = phi_wrap(k + 1, D.km), kw = phi_wrap(k - 1, D.km);
    const int lo = 1, hi = D.im - 2;
    double cp[kLineMax], dp[kLineMax];
    const double num1 = 1.0 * 1.0 * S.inv_dr2;
    for(int i = lo; i <= hi; i++){
        const double* g = B.geo + (size_t)(i * D.jm + j) * NGEO;
        const double num2 = g[G_P_LC_NUM2], num3 = g[G_P_LC_NUM3];
        const long long c0 = D.at(i, j, 0);
        const double acc_h = num2 * (p[c0 + k + st] + p[c0 + k - st]) + num3 * (p[c0 + ke] + p[c0 + kw]);
        double a = -num1, b = 2.0 * (num1 + num2 + num3), c = -num1;
        double d = acc_h - src[c0 + k];
        if(i == lo){
            if(hi > lo){ b += a * c43; c -= a * c13; }
            else d -= a * p[c0 + k - sr];
        }
        if(i == hi){
            if(hi > lo){ b += c * c43; a -= c * c13; }
            else d -= c * p[c0 + k + sr];
        }
        if(i == lo){ cp[i] = c / b; dp[i] = d / b; }
        else {
            const double dn = b - a * cp[i - 1];
            cp[i] = c / dn;
            dp[i] = (d - a * dp[i - 1]) / dn;
        }
    }
    double x = dp[hi];
    if(isfinite(x)) p[D.at(hi, j, k)] = x;
    for(int n = hi - 1; n >= lo; n--){
        x = dp[n] - cp[n] * x;
        if(isfinite(x)) p[D.at(n, j, k)] = x;
    }
}

// ---- the barotropic correction ------------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_baro_resid(Dims D, Scal S, Bufs<T> B, double A2){
    const PmLayout L = pm_layout(D);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y, m = blockIdx.z;
    if(k >= L.N) return;
    const long long mo = (long long)m * D.fs;
    const double* p = B.s[ATSAT_P_DYN] + mo;
    const double* src = B.s[B_PSRC] + mo;
    const int ke = phi_wrap(k + 1, D.km), kw = phi_wrap(k - 1, D.km);
    const int im = D.im, jm = D.jm;
    const double num1 = 1.0 * 1.0 * S.inv_dr2;
    auto P = [&](int i, int jj, int kk){ return p[D.at(i, jj, kk)]; };
    double s = 0.0;
    for(int i = 1; i < im - 1; i++){
        const double* g = B.geo + (size_t)(i * jm + j) * NGEO;
        const double num2 = g[G_P_LC_NUM2], num3 = g[G_P_LC_NUM3];
        auto pj = [&](int jn){
            if(jn == 0) return c43 * P(i, 1, k) - c13 * P(i, 2, k);
            if(jn == jm - 1) return c43 * P(i, jm - 2, k) - c13 * P(i, jm - 3, k);
            return P(i, jn, k); };
        auto pi_ = [&](int in){
            if(in == 0) return c43 * P(1, j, k) - c13 * P(2, j, k);
            if(in == im - 1) return c43 * P(im - 2, j, k) - c13 * P(im - 3, j, k);
            return P(in, j, k); };
        const double p0 = P(i, j, k);
        const double lc = num1 * (pi_(i + 1) + pi_(i - 1) - 2.0 * p0)
                        + num2 * (pj(j + 1) + pj(j - 1) - 2.0 * p0)
                        + num3 * (P(i, j, ke) + P(i, j, kw) - 2.0 * p0);
        s += src[D.at(i, j, k)] - lc;
    }
    B.pmodes[L.R + ((long long)m * L.J + (j - 1)) * L.N + k] = s / A2;
}

template<typename T>
__global__ void __launch_bounds__(256) k_baro_dft(Dims D, Bufs<T> B){
    const PmLayout L = pm_layout(D);
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    const int mm = blockIdx.y, m = blockIdx.z;
    if(jj >= L.J) return;
    const double* pm = B.pmodes;
    const double* r_ = pm + L.R + ((long long)m * L.J + jj) * L.N;
    const double* ct = pm + L.ct + (long long)mm * L.N;
    const double* st = pm + L.st + (long long)mm * L.N;
    double sc = 0.0, ss = 0.0;
    for(int k = 0; k < L.N; k++){ sc += r_[k] * ct[k]; ss += r_[k] * st[k]; }
    if(mm > 0){
        double ra = 0.0;
        for(int k = 0; k < L.N; k++) ra += fabs(r_[k]);
        const double tol = 1.0e-12 * ra;
        if(fabs(sc) < tol) sc = 0.0;
        if(fabs(ss) < tol) ss = 0.0;
    }
    const long long o = ((long long)m * (L.NM + 1) + mm) * L.J + jj;
    B.pmodes[L.Ac + o] = sc;
    B.pmodes[L.As + o] = ss;
}

template<typename T>
__global__ void __launch_bounds__(64) k_baro_tri(Dims D, Scal S, Bufs<T> B){
    const PmLayout L = pm_layout(D);
    const int mm = blockIdx.x * blockDim.x + threadIdx.x;
    const int part = blockIdx.y, m = blockIdx.z;
    if(mm > L.NM) return;
    if(mm == 0 && part == 1) return;                              // sin of mode 0 is zero
    if(2 * mm == L.N && part == 1) return;                        // and of the Nyquist mode
    const int J = (int)L.J;
    double* pm = B.pmodes;
    const long long o = ((long long)m * (L.NM + 1) + mm) * J;
    double* d = pm + (part == 0 ? L.Ac : L.As) + o;
    double* cp = pm + L.cpb + ((long long)m * (L.NM + 1) + mm) * 2 * J + (long long)part * J;
    double* dp = pm + L.dpb + ((long long)m * (L.NM + 1) + mm) * 2 * J + (long long)part * J;
    const double e = S.inv_dthe2, lam = pm[L.lam + mm];
    const double* spg = pm + L.spg;
    const double* sabs = B.sin_abs;
    auto coef = [&](int jj, double& a, double& b, double& c){
        const double s_ = spg[jj + 1];
        a = e; c = e; b = -2.0 * e - lam / (s_ * s_);
        if(jj == 0){ a = 0.0; b += c43 * e; c -= c13 * e; }
        if(jj == J - 1){ c = 0.0; b += c43 * e; a -= c13 * e; }
    };
    double a0, b0, c0;
    coef(0, a0, b0, c0);
    if(mm == 0){
        // consistency: the part of the source the all-Neumann operator cannot reach removed
        double yr = 0.0, ys = 0.0;
        for(int jj = 0; jj < J; jj++){
            const double y = (jj == 0 || jj == J - 1) ? 1.5 : 1.0;
            yr += y * d[jj]; ys += y * sabs[jj + 1];
        }
        for(int jj = 0; jj < J; jj++) d[jj] -= yr / ys * sabs[jj + 1];
        b0 = 1.0; c0 = 0.0; d[0] = 0.0;                           // q pinned at j = 1
    }
    cp[0] = c0 / b0; dp[0] = d[0] / b0;
    for(int jj = 1; jj < J; jj++){
        double a, b, c;
        coef(jj, a, b, c);
        const double dn = b - a * cp[jj - 1];
        cp[jj] = c / dn; dp[jj] = (d[jj] - a * dp[jj - 1]) / dn;
    }
    d[J - 1] = dp[J - 1];
    for(int jj = J - 2; jj >= 0; jj--) d[jj] = dp[jj] - cp[jj] * d[jj + 1];
    if(mm == 0){                                                  // zero sin-weighted mean
        double sq = 0.0, sw = 0.0;
        for(int jj = 0; jj < J; jj++){ sq += sabs[jj + 1] * d[jj]; sw += sabs[jj + 1]; }
        for(int jj = 0; jj < J; jj++) d[jj] -= sq / sw;
    }
}

template<typename T>
__global__ void __launch_bounds__(256) k_baro_apply(Dims D, Bufs<T> B){
    const PmLayout L = pm_layout(D);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int jj = blockIdx.y, m = blockIdx.z;
    if(k >= L.N) return;
    const double* pm = B.pmodes;
    const long long o = (long long)m * (L.NM + 1) * L.J + jj;
    double q = 0.0;
    for(int mm = 0; mm <= L.NM; mm++){
        const bool edge = (mm == 0 || 2 * mm == L.N);
        const double w_ = edge ? 1.0 : 2.0;
        q += w_ * (pm[L.Ac + o + (long long)mm * L.J] * pm[L.ct + (long long)mm * L.N + k]
                 + (edge ? 0.0 : pm[L.As + o + (long long)mm * L.J] * pm[L.st + (long long)mm * L.N + k]));
    }
    q /= L.N;
    double* p = B.s[ATSAT_P_DYN] + (long long)m * D.fs;
    for(int i = 0; i < D.im; i++) p[D.at(i, jj + 1, k)] += q;
}

// ---- the mean the line pass holds -----------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_pmean_rows(Dims D, Bufs<T> B){
    const PmLayout L = pm_layout(D);
    const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 1 + blockIdx.y, m = blockIdx.z;
    if(j >= D.jm - 1) return;
    const double* p = B.s[ATSAT_P_DYN] + (long long)m * D.fs + D.at(i, j, 0);
    double r = 0.0;
    for(int k = 0; k < D.km - 1; k++) r += p[k];
    B.pmodes[L.rows + ((long long)m * (D.im - 2) + (i - 1)) * L.J + (j - 1)] = r;
}

// mode 0: the mean on entry; 1: the shift of the mean since
template<typename T>
__global__ void __launch_bounds__(64) k_pmean_total(Dims D, Bufs<T> B, int mode){
    const PmLayout L = pm_layout(D);
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if(m >= D.M) return;
    const double* rows = B.pmodes + L.rows + (long long)m * (D.im - 2) * L.J;
    double s = 0.0;
    for(long long n = 0; n < (long long)(D.im - 2) * L.J; n++) s += rows[n];
    const long nn = (long)(D.im - 2) * L.J * (D.km - 1);
    const double mean = nn ? s / nn : 0.0;
    if(mode == 0) B.pmodes[L.mean0 + m] = mean;
    else B.pmodes[L.shift + m] = mean - B.pmodes[L.mean0 + m];
}

template<typename T>
__global__ void __launch_bounds__(256) k_pmean_shift(Dims D, Bufs<T> B){
    const PmLayout L = pm_layout(D);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    B.s[ATSAT_P_DYN][(long long)m * D.fs + D.at(i, j, k)] -= B.pmodes[L.shift + m];
}

template<typename T>
void launch_press_line(const Dims& D, const Scal& S, const Bufs<T>& B, int colour, hipStream_t st){
    const int nh = (D.km - 1 + 1) / 2;
    hipLaunchKernelGGL(k_press_line<T>, dim3((nh + 255) / 256, D.jm - 2, D.M), dim3(256), 0, st, D, S, B, colour);
}

template<typename T>
void launch_press_baro(const Dims& D, const Scal& S, const Bufs<T>& B, double A2, hipStream_t st){
    const PmLayout L = pm_layout(D);
    hipLaunchKernelGGL(k_baro_resid<T>, dim3((unsigned)((L.N + 255) / 256), (unsigned)L.J, D.M), dim3(256), 0, st, D, S, B, A2);
    hipLaunchKernelGGL(k_baro_dft<T>, dim3((unsigned)((L.J + 255) / 256), (unsigned)(L.NM + 1), D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_baro_tri<T>, dim3((unsigned)((L.NM + 1 + 63) / 64), 2, D.M), dim3(64), 0, st, D, S, B);
    hipLaunchKernelGGL(k_baro_apply<T>, dim3((unsigned)((L.N + 255) / 256), (unsigned)L.J, D.M), dim3(256), 0, st, D, B);
}

template<typename T>
void launch_press_mean(const Dims& D, const Bufs<T>& B, int mode, hipStream_t st){
    hipLaunchKernelGGL(k_pmean_rows<T>, dim3((D.jm - 2 + 255) / 256, D.im - 2, D.M), dim3(256), 0, st, D, B);
    hipLaunchKernelGGL(k_pmean_total<T>, dim3((D.M + 63) / 64), dim3(64), 0, st, D, B, mode);
    if(mode == 1)
        hipLaunchKernelGGL(k_pmean_shift<T>, dim3((D.km + 255) / 256, D.jm, D.im * D.M), dim3(256), 0, st, D, B);
}

#define ATSAT_PMODES_INST(T) \
    template void launch_press_line<T>(const Dims&, const Scal&, const Bufs<T>&, int, hipStream_t); \
    template void launch_press_baro<T>(const Dims&, const Scal&, const Bufs<T>&, double, hipStream_t); \
    template void launch_press_mean<T>(const Dims&, const Bufs<T>&, int, hipStream_t);
ATSAT_PMODES_INST(double)
ATSAT_PMODES_INST(float)

}  // namespace atsat
