// The step glue on the device: boundary conditions, the negative-species clamp, the n-level
// copies and the non-finite census.
//
// References: BoundaryConditions<cSaturnModel>::bcTheta / bcPhi / bcRadius
// (planet/BoundaryConditions.h:151-518) with the model's lists (planet/BC_Sat.cpp:108-179) and
// BC_Sat::bcRadius's species lid / wall copies (planet/BC_Sat.cpp:61-103);
// clampNegativeSpecies (planet/FileIO_Sat.cpp:716-770); restoreVar (cSaturnModel.cpp:1551-1602);
// the RK scalar sync (cSaturnModel.cpp:955-977); the NaN census (FileIO_Sat.cpp:418-470).
//
// Default knob values: bc_margin 1, POLE_COPY 1 (copy), POLE_FILL 1 (one open row per pole),
// RADIUS_COPY 2 (Neumann), BOTTOM_FORM 3 (cubic, u excepted), RIGID_LID 1, TOP_TAPER 2,
// SPECIES_LID_COPY 1, SPECIES_WALL_COPY 2, periodic seam. Each BC routine is column- or
// row-local, so one thread per column / row performs the reference's passes for it in order.
#include "atsat_dev.hpp"

namespace atsat {

__constant__ int c_th_ext[13] = { ATSAT_T, ATSAT_U, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE,
    ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S, ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE, ATSAT_NH4SH };
__constant__ int c_species[11] = { ATSAT_H2O, ATSAT_H2O_CLOUD, ATSAT_H2O_ICE, ATSAT_H2S,
    ATSAT_NH3, ATSAT_NH3_CLOUD, ATSAT_NH3_ICE, ATSAT_CH4, ATSAT_CH4_CLOUD, ATSAT_CH4_ICE, ATSAT_NH4SH };
__constant__ int c_turb[3] = { ATSAT_TKE, ATSAT_DIS, ATSAT_NUE };
__constant__ double c_floor[3] = { 0.0, 1.0e-10, 0.0 };            // BC_Sat.cpp:42, 178-179
// The chemistry arrays of the lists (work class): massflux_*, difflux_*, thermalmassflux are
// extrapolated at the poles like t (bc_fields_theta_extrap, BC_Sat.cpp:132-149), fluxlim_nh4sh is
// pinned to zero there (bc_fields_theta_zero, :154-156); all eight take the seam copy and the
// radial extrapolation (bc_fields_phi, bc_fields_radius, :108-126, 158-176).
__constant__ int c_chem_bc[8] = { ATSAT_MASSFLUX_H2S, ATSAT_MASSFLUX_NH3, ATSAT_MASSFLUX_NH4SH,
    ATSAT_DIFFLUX_H2S, ATSAT_DIFFLUX_NH3, ATSAT_DIFFLUX_NH4SH, ATSAT_THERMALMASSFLUX, ATSAT_FLUXLIM_NH4SH };

// The closure fields of the BC lists (bc_turb_fields, BC_Sat.cpp:178), all three in the work
// class: tke and dis (stored in the context's precision) and the closure's output nue.
template<typename T>
struct TurbRef {
    T* tke; T* dis; T* nue;
    __device__ double get(int q, long long n) const { return (double)(q == 0 ? tke[n] : (q == 1 ? dis[n] : nue[n])); }
    __device__ void set(int q, long long n, double v) const { (q == 0 ? tke : (q == 1 ? dis : nue))[n] = (T)v; }
};

// bcTheta, BoundaryConditions.h:292-391: one thread per (i, k), i 1..im-2, k 0..km-2.
template<typename T>
__global__ void __launch_bounds__(256) k_bc_theta(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 1 + (blockIdx.y % (D.im - 2)), m = blockIdx.y / (D.im - 2);
    if(k >= D.km - 1) return;
    const long long mo = m * D.fs;
    const int jm = D.jm, jf = 2;
    const double ramp = 1.0 / 2.0;
    const long long r1 = D.at(i, 1, k), rjf = D.at(i, jf, k), s1 = D.at(i, jm - 2, k), sjf = D.at(i, jm - 1 - jf, k);
    const long long r0 = D.at(i, 0, k), s0 = D.at(i, jm - 1, k);
    for(int q = 0; q < 13; q++){ const int f = c_th_ext[q];
        B.st(f, mo + r1, B.ld(f, mo + rjf)); B.st(f, mo + s1, B.ld(f, mo + sjf)); }
    T* v = B.w[ATSAT_V] + mo; double* w = B.s[ATSAT_W] + mo;
    v[r1] = (T)(ramp * (double)v[rjf]); v[s1] = (T)(ramp * (double)v[sjf]);
    w[r1] = ramp * w[rjf]; w[s1] = ramp * w[sjf];
    for(int q = 0; q < 7; q++){ T* F = B.w[c_chem_bc[q]] + mo; F[r1] = F[rjf]; F[s1] = F[sjf]; }
    T* fl = B.w[ATSAT_FLUXLIM_NH4SH] + mo;
    fl[r1] = (T)(ramp * (double)fl[rjf]); fl[s1] = (T)(ramp * (double)fl[sjf]);
    fl[r0] = (T)0.0; fl[s0] = (T)0.0;
    for(int q = 0; q < 7; q++){ T* F = B.w[c_chem_bc[q]] + mo; F[r0] = F[r1]; F[s0] = F[s1]; }
    const TurbRef<T> TR{B.w[ATSAT_TKE] + mo, B.w[ATSAT_DIS] + mo, B.w[ATSAT_NUE] + mo};
    if(S.turb_active) for(int q = 0; q < 3; q++){ const double fl = c_floor[q];
        const double a = TR.get(q, rjf), b = TR.get(q, sjf);
        TR.set(q, r1, a > fl ? a : fl); TR.set(q, s1, b > fl ? b : fl); }
    v[r0] = (T)0.0; v[s0] = (T)0.0; w[r0] = 0.0; w[s0] = 0.0;
    for(int q = 0; q < 13; q++){ const int f = c_th_ext[q];
        B.st(f, mo + r0, B.ld(f, mo + r1)); B.st(f, mo + s0, B.ld(f, mo + s1)); }
    if(S.turb_active){
        const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
        for(int q = 0; q < 3; q++){ const double fl = c_floor[q];
            const double a = c43 * TR.get(q, r1) - c13 * TR.get(q, D.at(i, 2, k));
            const double b = c43 * TR.get(q, s1) - c13 * TR.get(q, D.at(i, jm - 3, k));
            TR.set(q, r0, a > fl ? a : fl); TR.set(q, s0, b > fl ? b : fl); }
    }
}

// bcPhi, periodic seam, :422-438: one thread per (i, j).
template<typename T>
__global__ void __launch_bounds__(256) k_bc_phi(Dims D, Scal S, Bufs<T> B){
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y % D.im, m = blockIdx.y / D.im;
    if(j >= D.jm) return;
    const long long mo = m * D.fs;
    const long long n0 = D.at(i, j, 0), n1 = D.at(i, j, D.km - 1);
    // k = 0 and k = km-1 are different cells: every value is loaded before the first store
    constexpr int NF = ATSAT_NH4SH - ATSAT_T + 1;
    double sv[NF], cv[8];
#pragma unroll
    for(int f = 0; f < NF; f++) sv[f] = B.ld(ATSAT_T + f, mo + n0);
#pragma unroll
    for(int q = 0; q < 8; q++) cv[q] = (double)B.w[c_chem_bc[q]][mo + n0];
    const TurbRef<T> TR{B.w[ATSAT_TKE] + mo, B.w[ATSAT_DIS] + mo, B.w[ATSAT_NUE] + mo};
    double tv[3];
#pragma unroll
    for(int q = 0; q < 3; q++) tv[q] = S.turb_active ? TR.get(q, n0) : 0.0;
#pragma unroll
    for(int f = 0; f < NF; f++) B.st(ATSAT_T + f, mo + n1, sv[f]);
#pragma unroll
    for(int q = 0; q < 8; q++) B.w[c_chem_bc[q]][mo + n1] = (T)cv[q];
    if(S.turb_active) for(int q = 0; q < 3; q++){ const double fl = c_floor[q];
        const double a = tv[q]; const double v = a > fl ? a : fl; TR.set(q, n0, v); TR.set(q, n1, v); }
}

// bcRadius, :151-286, then BC_Sat's species copies, BC_Sat.cpp:94-102: one thread per (j, k).
template<typename T>
__global__ void __launch_bounds__(256) k_bc_radius(Dims D, Scal S, Bufs<T> B, int clampsp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    const long long mo = m * D.fs;
    const int im = D.im;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0, three = 3.0;
    const long long n0 = D.at(0, j, k), n1 = D.at(1, j, k), n2 = D.at(2, j, k), n3 = D.at(3, j, k);
    const long long t0 = D.at(im - 1, j, k), t1 = D.at(im - 2, j, k), t2 = D.at(im - 3, j, k);
    // The cells written (i = 0, im-1) are never among those read (i = 1, 2, 3, im-2, im-3), so
    // each group of fields is loaded before any of it is stored.
    if(j >= 1 && j <= D.jm - 2 && k <= D.km - 2){
        constexpr int G = 5;
#pragma unroll
        for(int f0 = ATSAT_T; f0 <= ATSAT_NH4SH; f0 += G){
            double a[G], b[G], c[G], e1[G], e2[G];
#pragma unroll
            for(int g = 0; g < G; g++){
                const int f = f0 + g;
                a[g] = B.ld(f, mo + n1); b[g] = B.ld(f, mo + n2); c[g] = B.ld(f, mo + n3);
                e1[g] = B.ld(f, mo + t1); e2[g] = B.ld(f, mo + t2);
            }
#pragma unroll
            for(int g = 0; g < G; g++){
                const int f = f0 + g;
                B.st(f, mo + n0, (f != ATSAT_U) ? (c[g] - three * b[g] + three * a[g]) : (c43 * a[g] - c13 * b[g]));
                B.st(f, mo + t0, c43 * e1[g] - c13 * e2[g]);
            }
        }
        static_assert((ATSAT_NH4SH - ATSAT_T + 1) % 5 == 0, "bcRadius: whole groups of fields");
        for(int q0 = 0; q0 < 8; q0 += 4){
            double a[4], b[4], c[4], e1[4], e2[4];
#pragma unroll
            for(int g = 0; g < 4; g++){
                const T* F = B.w[c_chem_bc[q0 + g]] + mo;
                a[g] = F[n1]; b[g] = F[n2]; c[g] = F[n3]; e1[g] = F[t1]; e2[g] = F[t2];
            }
#pragma unroll
            for(int g = 0; g < 4; g++){
                T* F = B.w[c_chem_bc[q0 + g]] + mo;
                F[n0] = (T)(c[g] - three * b[g] + three * a[g]);
                F[t0] = (T)(c43 * e1[g] - c13 * e2[g]);
            }
        }
        const TurbRef<T> TR{B.w[ATSAT_TKE] + mo, B.w[ATSAT_DIS] + mo, B.w[ATSAT_NUE] + mo};
        if(S.turb_active) for(int q = 0; q < 3; q++){ const double fl = c_floor[q];
            const double a = c43 * TR.get(q, n1) - c13 * TR.get(q, n2), b = c43 * TR.get(q, t1) - c13 * TR.get(q, t2);
            TR.set(q, n0, a > fl ? a : fl); TR.set(q, t0, b > fl ? b : fl); }
    }
    T* u = B.w[ATSAT_U] + mo; T* v = B.w[ATSAT_V] + mo; double* w = B.s[ATSAT_W] + mo;
    u[n0] = (T)0.0; u[t0] = (T)0.0;                                      // rigid walls, :232-240
    const int iml = im - 1;                                              // top taper 2, :264-285
    const long long l0 = D.at(iml, j, k), l1 = D.at(iml - 1, j, k), l2 = D.at(iml - 2, j, k), l3 = D.at(iml - 3, j, k);
    v[l0] = (T)0.0; w[l0] = 0.0;
    v[l1] = (T)((1.0 / 3.0) * (double)v[l3]); w[l1] = (1.0 / 3.0) * w[l3];
    v[l2] = (T)((2.0 / 3.0) * (double)v[l3]); w[l2] = (2.0 / 3.0) * w[l3];
    const double* rho = B.s[ATSAT_RHO_MIX] + mo;                        // BC_Sat.cpp:97
    auto rho_at = [&](long long n){ const double r = rho[n]; return (r > 0.0 && isfinite(r)) ? r : S.r_mix; };
    const double fr = rho_at(n0) / rho_at(n1);
    double st1[11], sn1[11];
#pragma unroll
    for(int q = 0; q < 11; q++){ const T* F = B.w[c_species[q]] + mo; st1[q] = (double)F[t1]; sn1[q] = (double)F[n1]; }
#pragma unroll
    for(int q = 0; q < 11; q++){
        T* F = B.w[c_species[q]] + mo;
        F[t0] = (T)st1[q];
        double x = sn1[q] * fr;
        if(clampsp && !(x >= 0.0)) x = 0.0;                              // the folded clamp
        F[n0] = (T)x;
    }
}

template<typename T>
void launch_bc(const Dims& D, const Scal& S, const Bufs<T>& B, bool clamp, hipStream_t st){
    hipLaunchKernelGGL(k_bc_theta<T>, dim3((D.km - 1 + 255) / 256, (D.im - 2) * D.M), dim3(256), 0, st, D, S, B);
    hipLaunchKernelGGL(k_bc_phi<T>, dim3((D.jm + 255) / 256, D.im * D.M), dim3(256), 0, st, D, S, B);
    hipLaunchKernelGGL(k_bc_radius<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, S, B, clamp ? 1 : 0);
}

// clampNegativeSpecies (:746-761, "!(v >= 0)" also catches NaN).
template<typename T>
__global__ void __launch_bounds__(256) k_clamp(Dims D, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = m * D.fs + D.at(i, j, k);
    for(int q = 0; q < 11; q++){
        T* F = B.w[c_species[q]];
        const T v = F[n];
        if(!(v >= (T)0.0)) F[n] = (T)0.0;
    }
}

template<typename T>
void launch_clamp(const Dims& D, const Scal&, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_clamp<T>, dim3((D.km + 255) / 256, D.jm, D.im * D.M), dim3(256), 0, st, D, B);
}

// Copies live -> start-of-step for the fields of `mask`: restoreVar(1.0) (all of them) and the
// RK scalar sync (the nine condensables, t, h2s, nh4sh), when the copies are not aliased.
template<typename T>
__global__ void __launch_bounds__(256) k_copy_n(Dims D, Bufs<T> B, unsigned mask){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = m * D.fs + D.at(i, j, k);
    for(int f = 0; f < ATSAT_NPROG; f++) if(mask & (1u << f)){
        if(prog_state(f)) B.s[ATSAT_F_N0 + f][n] = B.s[f][n];
        else B.w[ATSAT_F_N0 + f][n] = B.w[f][n];
    }
}

template<typename T>
void launch_copy_n(const Dims& D, const Scal&, const Bufs<T>& B, unsigned mask, hipStream_t st){
    hipLaunchKernelGGL(k_copy_n<T>, dim3((D.km + 255) / 256, D.jm, D.im * D.M), dim3(256), 0, st, D, B, mask);
}

// Non-finite census over the prognostic fields of every member (FileIO_Sat.cpp:418-470 tests the
// exponent bits; isfinite is the same predicate).
template<typename T>
__global__ void __launch_bounds__(256) k_census(Dims D, Bufs<T> B, int nprog){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    unsigned long long bad = 0;
    if(k < D.km){
        const int i = im_m % D.im, m = im_m / D.im;
        const long long n = m * D.fs + D.at(i, j, k);
        for(int f = 0; f < nprog; f++) if(!isfinite(B.ld(f, n))) bad++;
    }
    // wave-level sum, then one atomic per wave
    for(int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if((threadIdx.x & 63) == 0 && bad) atomicAdd(B.counter, bad);
}

template<typename T>
void launch_census(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    hipLaunchKernelGGL(k_census<T>, dim3((D.km + 255) / 256, D.jm, D.im * D.M), dim3(256), 0, st, D, B,
                       S.turb_active ? (int)ATSAT_NPROG : (int)ATSAT_NPROG - 2);
}

// The sums of the reference's two budget instruments (print-only there), in FP64 for every member:
// the H2O, NH3 and CH4 columns, vapour + cloud + ice, in g/m2 (ATSAT_CWB_DIAG's cwb_col,
// cSaturnModel.cpp:776-787: every level with a thickness, layer_thickness_m of cSaturnModel.h:
// 689-692, weighted by sin(the.z[j])) and the level-mean temperature in K (ATSAT_TB_DIAG's tb_lev,
// :656-666, all rows); both over sum_j sin(the.z[j]) km (:673, 798). Two passes:
//   k_budget_levels  one workgroup per (level, quantity, member): the sin-weighted plane sum of t or
//                    of a gas's three phases, a fixed tree;
//   k_budget_final   one thread per (member, quantity): the temperature's level means, or the
//                    thickness-weighted sum of a gas's planes.
// The output of member m: out[m (3 + im) + g] the columns, out[m (3 + im) + 3 + i] the levels.
template<typename T>
__global__ void __launch_bounds__(256) k_budget_levels(Dims D, Bufs<T> B, double* part){
    __shared__ double red[256];
    const int i = blockIdx.x, q = blockIdx.y, m = blockIdx.z;
    const long long mo = (long long)m * D.fs;
    const int gv[3] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
    double acc = 0.0;
    for(int x = threadIdx.x; x < D.jm * D.km; x += 256){
        const int j = x / D.km, k = x - j * D.km;
        const long long n = mo + D.at(i, j, k);
        double v;
        if(q == 0) v = B.s[ATSAT_T][n];
        else {
            const int f = gv[q - 1];
            v = (double)B.w[f][n] + (double)B.w[f + 1][n] + (double)B.w[f + 2][n];
        }
        acc += B.thez_sin[j] * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for(int h = 128; h > 0; h >>= 1){
        if((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if(threadIdx.x == 0) part[((long long)m * 4 + q) * D.im + i] = red[0];
}

__global__ void __launch_bounds__(64) k_budget_final(Dims D, Scal S, const double* __restrict__ thez_sin,
                                                     const double* __restrict__ part, double* __restrict__ out){
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if(x >= D.M * 4) return;
    const int m = x / 4, q = x - m * 4;
    double norm = 0.0;
    for(int j = 0; j < D.jm; j++) norm += thez_sin[j] * D.km;
    const double* p = part + ((long long)m * 4 + q) * D.im;
    double* o = out + (long long)m * (3 + D.im);
    if(q == 0){
        for(int i = 0; i < D.im; i++) o[3 + i] = p[i] * S.t_ref / norm;
        return;
    }
    // layer_thickness_m(i): the float layer heights i h, h = float(L_atm dr) (cSaturnModel.h:1061-1073)
    const float h = (float)S.layer_h;
    double s = 0.0;
    for(int i = 0; i <= D.im - 2; i++){
        const float z0 = (float)i * h, z1 = (float)(i + 1) * h;
        const double dz = (double)(z1 - z0) * 1.0e3;
        if(dz > 0.0) s += dz * p[i];
    }
    o[q - 1] = s * 1.0e3 / norm;
}

template<typename T>
void launch_budgets(const Dims& D, const Scal& S, const Bufs<T>& B, double* part, double* out, hipStream_t st){
    hipLaunchKernelGGL(k_budget_levels<T>, dim3(D.im, 4, D.M), dim3(256), 0, st, D, B, part);
    hipLaunchKernelGGL(k_budget_final, dim3((D.M * 4 + 63) / 64), dim3(64), 0, st, D, S, B.thez_sin,
                       (const double*)part, out);
}

#define INST(T) \
    template void launch_bc<T>(const Dims&, const Scal&, const Bufs<T>&, bool, hipStream_t); \
    template void launch_clamp<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_copy_n<T>(const Dims&, const Scal&, const Bufs<T>&, unsigned, hipStream_t); \
    template void launch_census<T>(const Dims&, const Scal&, const Bufs<T>&, hipStream_t); \
    template void launch_budgets<T>(const Dims&, const Scal&, const Bufs<T>&, double*, double*, hipStream_t);
INST(double)
INST(float)

}  // namespace atsat
