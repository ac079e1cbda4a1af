// The converged dynamic-pressure solve (knobs.press_mode 1): Lc p_dyn = src to a relative
// residual tolerance, by FP64 defect correction around a separable direct solver.
//
// The operator is the reference's (PressureSolver.h:350-353, 476-491, no obstacle): at an
// interior cell (i 1..im-2, j 1..jm-2, k 0..km-2 on the periodic ring)
//     Lc p = num1 (p[i+1] + p[i-1] - 2p) + num2(i) (p[j+1] + p[j-1] - 2p) + num3(i,j) (p[k+1] + p[k-1] - 2p)
// with num1 = 1/dr^2, num2 = 1/(rm^2 dtheta^2), num3 = 1/(rm^2 sin^2(theta) dphi^2), and the
// boundary pass of :718-740 (p0 = 4/3 p1 - 1/3 p2 on the radial and polar faces, the seam copy)
// closing the system. The reference only relaxes it with a fixed number of red-black sweeps; here
// it is solved. Folding the boundary relations into the first and last interior rows gives
//     (2/3)(p2 - p1)   at i = 1 (and likewise at i = im-2, j = 1, j = jm-2),
// and multiplying the equation by rm(i)^2 separates it:
//     rm^2 num1 D_r p  +  [ dtheta^-2 D_theta + dphi^-2 S(j) D_phi ] p  =  rm^2 src.
// The bracket does not depend on i. A real DFT along the ring diagonalises D_phi (eigenvalue
// -lambda_m, lambda_m = 2 - 2 cos(2 pi m / n)); for every m the remaining tridiagonal operator in
// theta is symmetrised by a diagonal scaling and diagonalised once on the host (implicit QL). In
// those coordinates every (m, theta-mode) leaves a tridiagonal system in r, solved by Thomas. The
// ring transform is a real FFT in LDS (k_ring_fwd / k_ring_inv, rings of 60, 90, 120, 180 and 360
// points; a rocBLAS GEMM against the DFT matrix for other lengths), the theta transform a batch of
// dense products with the eigenvector matrices (rocBLAS GEMMs); both FP64 or FP32. The constant is
// the null space of the all-Neumann operator and its mode is pinned.
//
// The loop solves for the increment dp = p_new - p: after the boundary pass on p, the source of
// the increment is src - Lc p, projected onto the solvable subspace (its weighted mean with the
// left null vector w = y(i) y(j), y = 3/2 on the first and last interior row, 1 elsewhere,
// removed); then, from dp = 0,
//     r = src - Lc dp (FP64), stop if |r| / |src| <= tol for every member, or when a correction
//     no longer halves it (the FP64 rounding floor of the residual, eps * den * |dp| / |src|,
//     which the near-null modes of the operator can lift above tol),
//     e = M^-1 r (the separable solve: forward transforms FP64 or FP32, the rest FP64),
//     dp += e, boundary pass on dp,
// and finally p += dp with the weighted mean of dp removed (p keeps its gauge).
#include "atsat_dev.hpp"
#include "atsat_gpu.h"
#include "tridiag_eigen.hpp"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace atsat {

// ---- kernels -------------------------------------------------------------------------------

// One block per (i, j) row of one member. The residual of interior cells, scaled by rm(i)^2,
// into f (zero elsewhere); per-row sums of r^2 and src^2 into part[row][2].
template<typename P>
__global__ void __launch_bounds__(kSolveRow) k_sol_residual(Dims D, Scal S, const double* __restrict__ p,
                                                         const double* __restrict__ src, const double* __restrict__ geo,
                                                         const double* __restrict__ rm2, P* __restrict__ f,
                                                         double* __restrict__ part, RowMap RM){
    int j, i, m;
    if(!RM.decode((int)blockIdx.x, 0, j, i, m)) return;
    const int row = (m * D.im + i) * D.jm + j;
    const int k = threadIdx.x;
    const long long base = (long long)row * D.kp;
    const int nk = D.km - 1;
    double rr = 0.0, ss = 0.0;
    const bool inner = i >= 1 && i <= D.im - 2 && j >= 1 && j <= D.jm - 2;
    for(int kk = k; kk < D.kp; kk += blockDim.x){
        P out = (P)0;
        if(inner && kk < nk){
            const double* g = geo + (size_t)(i * D.jm + j) * NGEO;
            const long long c0 = base + kk, sr = (long long)D.jm * D.kp, st = D.kp;
            const long long pe = base + (kk + 1 < nk ? kk + 1 : 0), pw = base + (kk > 0 ? kk - 1 : nk - 1);
            const double acc = (p[c0 + sr] + p[c0 - sr]) * S.inv_dr2 + (p[c0 + st] + p[c0 - st]) * g[G_P_NUM2]
                             + (p[pe] + p[pw]) * g[G_P_NUM3];
            const double s = src[c0];
            const double r = s - (acc - g[G_P_DENOM] * p[c0]);
            rr += r * r;
            ss += s * s;
            out = (P)(rm2[i] * r);
        }
        f[base + kk] = out;
    }
    __shared__ double red[2][kSolveRow];
    red[0][k] = rr; red[1][k] = ss;
    __syncthreads();
    for(int w = blockDim.x / 2; w > 0; w >>= 1){
        if(k < w){ red[0][k] += red[0][k + w]; red[1][k] += red[1][k + w]; }
        __syncthreads();
    }
    if(k == 0){ part[2 * row] = red[0][0]; part[2 * row + 1] = red[1][0]; }
}

// Per-row sums of w * x over the interior cells (w = y(i) y(j)), for the weighted means.
__global__ void __launch_bounds__(kSolveRow) k_sol_wsum(Dims D, const double* __restrict__ x, double* __restrict__ part){
    const int row = blockIdx.x;
    const int j = row % D.jm, mi = row / D.jm, i = mi % D.im;
    const int k = threadIdx.x;
    const int nk = D.km - 1;
    double s = 0.0;
    if(i >= 1 && i <= D.im - 2 && j >= 1 && j <= D.jm - 2){
        const double w = ((i == 1 || i == D.im - 2) ? 1.5 : 1.0) * ((j == 1 || j == D.jm - 2) ? 1.5 : 1.0);
        for(int kk = k; kk < nk; kk += blockDim.x) s += w * x[(long long)row * D.kp + kk];
    }
    __shared__ double red[kSolveRow];
    red[k] = s;
    __syncthreads();
    for(int w2 = blockDim.x / 2; w2 > 0; w2 >>= 1){
        if(k < w2) red[k] += red[k + w2];
        __syncthreads();
    }
    if(k == 0){ part[2 * row] = red[0]; part[2 * row + 1] = 0.0; }
}

// One block per member: the fixed-order sums of its rows' partials -> out[2 * m + 0/1].
__global__ void __launch_bounds__(256) k_sol_reduce(int rows_per_member, const double* __restrict__ part, double* __restrict__ out){
    const int m = blockIdx.x, t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for(int r = t; r < rows_per_member; r += blockDim.x){
        a += part[2 * ((long long)m * rows_per_member + r)];
        b += part[2 * ((long long)m * rows_per_member + r) + 1];
    }
    __shared__ double red[2][256];
    red[0][t] = a; red[1][t] = b;
    __syncthreads();
    for(int w = blockDim.x / 2; w > 0; w >>= 1){
        if(t < w){ red[0][t] += red[0][t + w]; red[1][t] += red[1][t + w]; }
        __syncthreads();
    }
    if(t == 0){ out[2 * m] = red[0][0]; out[2 * m + 1] = red[1][0]; }
}

// x -= c[m] / wtot on the interior cells (c the weighted sum of x: the projection onto the
// solvable subspace).
__global__ void __launch_bounds__(256) k_sol_shift(Dims D, double* __restrict__ x, const double* __restrict__ c, double wtot){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    const int j = row % D.jm, mi = row / D.jm, i = mi % D.im, m = mi / D.im;
    if(k >= D.km - 1) return;
    if(i >= 1 && i <= D.im - 2 && j >= 1 && j <= D.jm - 2) x[(long long)row * D.kp + k] -= c[2 * m] / wtot;
}

// The radial tridiagonal of every (ring mode, theta mode, member): (rm^2 num1 D_r + mu) q = g, in
// place in the transformed buffer c2[mode][member][i][1 + e] (row pitch jm). Thomas, with the
// elimination factors of its (wavenumber, theta mode) formed once on the host (they do not depend
// on the member or the data): cf = c' and rd = 1 / (b - a c'_prev) per level, [wave][i - 1][e].
// The loops run over NI >= ni levels with those past ni (past ni - 1 for the pinned null mode)
// masked, so d' stays in registers and the column is read once and written once.
template<typename Pin, typename Pout, int NI>
__global__ void __launch_bounds__(256) k_sol_radial(Dims D, const Pin* cin, Pout* cout, const double* __restrict__ cf,
                                                    const double* __restrict__ rd, const double* __restrict__ s,
                                                    long long R, int nj, int null_e){
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int mode = blockIdx.y, m = blockIdx.z;
    if(e >= nj) return;
    const int nk = D.km - 1, half = nk / 2;
    const int wave = mode <= half ? mode : mode - half;      // cos modes 0..n/2, sin modes n/2+1..n-1
    const int ni = D.im - 2;
    const long long off = (long long)mode * R + (long long)m * D.im * D.jm + 1 + e;
    const Pin* qi = cin + off;
    Pout* q = cout + off;                                    // may be the same storage as qi
    const long long st = D.jm;
    const int n = (wave == 0 && e == null_e) ? ni - 1 : ni;  // the null mode: pin q[ni] = 0
    const long long fo = (long long)wave * ni * nj + e;
    const double two3 = 2.0 / 3.0;
    double dp[NI];
    double dprev = 0.0;
#pragma unroll
    for(int i = 1; i <= NI; i++){
        if(i <= n){
            const double a = (i == 1) ? 0.0 : ((i == ni) ? two3 * s[i] : s[i]);
            dprev = ((double)qi[i * st] - a * dprev) * rd[fo + (long long)(i - 1) * nj];
            dp[i - 1] = dprev;
        }
    }
    double x = dprev;
#pragma unroll
    for(int i = NI; i >= 1; i--){
        if(i == n) q[n * st] = (Pout)x;
        else if(i < n){
            x = dp[i - 1] - cf[fo + (long long)(i - 1) * nj] * x;
            q[i * st] = (Pout)x;
        }
    }
    if(n < ni) q[ni * st] = (Pout)0;
}

// p += e on the interior cells.
template<typename P>
__global__ void __launch_bounds__(256) k_sol_update(Dims D, double* __restrict__ p, const P* __restrict__ e){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    const int j = row % D.jm, mi = row / D.jm, i = mi % D.im;
    if(k >= D.km - 1 || i < 1 || i > D.im - 2 || j < 1 || j > D.jm - 2) return;
    p[(long long)row * D.kp + k] += (double)e[(long long)row * D.kp + k];
}

// src -= Lc p on the interior cells (the source of the increment).
__global__ void __launch_bounds__(kSolveRow) k_sol_defect(Dims D, Scal S, const double* __restrict__ p,
                                                        double* __restrict__ src, const double* __restrict__ geo,
                                                        RowMap RM){
    int j, i, m;
    if(!RM.decode((int)blockIdx.x, 0, j, i, m)) return;
    const int row = (m * D.im + i) * D.jm + j;
    if(i < 1 || i > D.im - 2 || j < 1 || j > D.jm - 2) return;
    const long long base = (long long)row * D.kp;
    const int nk = D.km - 1;
    const double* g = geo + (size_t)(i * D.jm + j) * NGEO;
    for(int kk = threadIdx.x; kk < nk; kk += blockDim.x){
        const long long c0 = base + kk, sr = (long long)D.jm * D.kp, st = D.kp;
        const long long pe = base + (kk + 1 < nk ? kk + 1 : 0), pw = base + (kk > 0 ? kk - 1 : nk - 1);
        const double acc = (p[c0 + sr] + p[c0 - sr]) * S.inv_dr2 + (p[c0 + st] + p[c0 - st]) * g[G_P_NUM2]
                         + (p[pe] + p[pw]) * g[G_P_NUM3];
        src[c0] -= acc - g[G_P_DENOM] * p[c0];
    }
}

__global__ void __launch_bounds__(256) k_sol_fill(Dims D, double* __restrict__ x, double v){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < D.kp) x[(long long)blockIdx.y * D.kp + k] = v;
}

// x += y - c[m] / wtot on every cell: the increment without its null-space constant, added to p.
__global__ void __launch_bounds__(256) k_sol_add_shifted(Dims D, double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ c, double wtot){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    const int m = row / (D.jm * D.im);
    if(k < D.km){ const long long n = (long long)row * D.kp + k; x[n] += y[n] - c[2 * m] / wtot; }
}

// The boundary pass of the solver (PressureSolver.h:718-740) on any FP64 field: the 2-point
// Neumann relation on the radial planes, then on the polar rows, then the seam copy.
__global__ void __launch_bounds__(256) k_sol_bnd_r(Dims D, double* __restrict__ x){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    double* p = x + m * D.fs;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    p[D.at(0, j, k)] = c43 * p[D.at(1, j, k)] - c13 * p[D.at(2, j, k)];
    p[D.at(D.im - 1, j, k)] = c43 * p[D.at(D.im - 2, j, k)] - c13 * p[D.at(D.im - 3, j, k)];
}
__global__ void __launch_bounds__(256) k_sol_bnd_t(Dims D, double* __restrict__ x){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    double* p = x + m * D.fs;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    p[D.at(i, 0, k)] = c43 * p[D.at(i, 1, k)] - c13 * p[D.at(i, 2, k)];
    p[D.at(i, D.jm - 1, k)] = c43 * p[D.at(i, D.jm - 2, k)] - c13 * p[D.at(i, D.jm - 3, k)];
}
__global__ void __launch_bounds__(256) k_sol_seam(Dims D, double* __restrict__ x){
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, m = blockIdx.z;
    if(j >= D.jm) return;
    double* p = x + m * D.fs;
    p[D.at(i, j, D.km - 1)] = p[D.at(i, j, 0)];
}

// ---- the ring transforms --------------------------------------------------------------------
//
// The real DFT of every (member, i, j) ring of n = 2L points. A ring x is packed into the complex
// sequence z[t] = x[2t] + i x[2t+1] of length L, which a Stockham FFT in LDS transforms (radix 2, 3,
// 4 and 5 butterflies, ping-pong buffers, the twiddles of stage s W_{Ns r}^(q k) read from a table
// of W_n^j, j = 0..n-1, formed on the host in FP64: tw[2j] = cos, tw[2j+1] = -sin); the spectrum of
// x follows from Z by the split
//     X[w] = E[w] + W_n^w O[w],  E[w] = (Z[w] + conj Z[L-w]) / 2,  O[w] = (Z[w] - conj Z[L-w]) / 2i,
// for w = 0..L (Z[L] = Z[0]). The modes are stored as the solver's matrices order them: the cosine
// coefficients a_w = Re X[w], w = 0..L, then the sine coefficients b_w = -Im X[w], w = 1..L-1, each
// as a [mode][row] plane (the layout the theta transform reads). The inverse runs the same steps
// backwards from the Hermitian spectrum X[w] = a_w - i b_w, with the 1/n of the inverse DFT. A
// workgroup transforms RB consecutive rows: they are read along the ring, the modes along the rows.
// The transforms are the preconditioner of the defect correction, whose FP64 residual alone decides
// the result, so their arithmetic is free to take fused multiply-adds.

// one radix-R butterfly in place, W = exp(-2 pi i / R) (INV: its conjugate)
template<typename P, int R, bool INV>
__device__ __forceinline__ void ring_bfly(P (&vr)[R], P (&vi)[R]){
    constexpr P sg = INV ? (P)-1 : (P)1;      // "- i sg u" = (sg u.im, -sg u.re)
    if constexpr(R == 2){
        const P ar = vr[0] + vr[1], ai = vi[0] + vi[1];
        vr[1] = vr[0] - vr[1]; vi[1] = vi[0] - vi[1];
        vr[0] = ar; vi[0] = ai;
    } else if constexpr(R == 3){
        constexpr P s3 = (P)0.86602540378443864676;                    // sin(2 pi / 3)
        const P tr = vr[1] + vr[2], ti = vi[1] + vi[2];
        const P ur = vr[1] - vr[2], ui = vi[1] - vi[2];
        const P ar = fma((P)-0.5, tr, vr[0]), ai = fma((P)-0.5, ti, vi[0]);
        vr[0] = vr[0] + tr; vi[0] = vi[0] + ti;
        vr[1] = fma(sg * s3, ui, ar);  vi[1] = fma(-sg * s3, ur, ai);
        vr[2] = fma(-sg * s3, ui, ar); vi[2] = fma(sg * s3, ur, ai);
    } else if constexpr(R == 4){
        const P ar = vr[0] + vr[2], ai = vi[0] + vi[2], br = vr[0] - vr[2], bi = vi[0] - vi[2];
        const P cr = vr[1] + vr[3], ci = vi[1] + vi[3], dr = vr[1] - vr[3], di = vi[1] - vi[3];
        vr[0] = ar + cr; vi[0] = ai + ci;
        vr[2] = ar - cr; vi[2] = ai - ci;
        vr[1] = fma(sg, di, br);  vi[1] = fma(-sg, dr, bi);
        vr[3] = fma(-sg, di, br); vi[3] = fma(sg, dr, bi);
    } else {
        static_assert(R == 5, "ring transforms: radix 2, 3, 4 or 5");
        constexpr P c1 = (P)0.30901699437494742410, c2 = (P)-0.80901699437494742410;   // cos(2 pi / 5), cos(4 pi / 5)
        constexpr P s1 = (P)0.95105651629515357212, s2 = (P)0.58778525229247312917;    // sin(2 pi / 5), sin(4 pi / 5)
        const P t1r = vr[1] + vr[4], t1i = vi[1] + vi[4], t2r = vr[2] + vr[3], t2i = vi[2] + vi[3];
        const P u1r = vr[1] - vr[4], u1i = vi[1] - vi[4], u2r = vr[2] - vr[3], u2i = vi[2] - vi[3];
        const P Ar = fma(c1, t1r, fma(c2, t2r, vr[0])), Ai = fma(c1, t1i, fma(c2, t2i, vi[0]));
        const P Cr = fma(c2, t1r, fma(c1, t2r, vr[0])), Ci = fma(c2, t1i, fma(c1, t2i, vi[0]));
        const P Br = sg * fma(s1, u1r, s2 * u2r), Bi = sg * fma(s1, u1i, s2 * u2i);    // sg (s1 u1 + s2 u2)
        const P Dr = sg * fma(s2, u1r, -s1 * u2r), Di = sg * fma(s2, u1i, -s1 * u2i);  // sg (s2 u1 - s1 u2)
        vr[0] = vr[0] + t1r + t2r; vi[0] = vi[0] + t1i + t2i;
        vr[1] = Ar + Bi; vi[1] = Ai - Br;
        vr[4] = Ar - Bi; vi[4] = Ai + Br;
        vr[2] = Cr + Di; vi[2] = Ci - Dr;
        vr[3] = Cr - Di; vi[3] = Ci + Dr;
    }
}

// One Stockham stage of radix R after stages spanning NS points, RB rows of L complex values
// (interleaved) from a to b: butterfly j (k = j mod NS) takes the points j + q L/R, each times
// W_{NS R}^(q k), and writes its outputs to (j / NS) NS R + k + q NS.
template<typename P, int L, int R, int NS, bool INV, int RB>
__device__ __forceinline__ void ring_stage(const P* a, P* b, const P* ts){
    constexpr int M = L / R;
    for(int it = threadIdx.x; it < RB * M; it += blockDim.x){
        const int r = it / M, j = it - r * M;
        const int k = j % NS;
        const P* src = a + 2 * r * L;
        P vr[R], vi[R];
#pragma unroll
        for(int q = 0; q < R; q++){
            const P xr = src[2 * (j + q * M)], xi = src[2 * (j + q * M) + 1];
            if(NS == 1 || q == 0){ vr[q] = xr; vi[q] = xi; continue; }
            const int e = 2 * q * k * (L / (NS * R));                 // W_{NS R}^(q k) = W_n^e, e < n
            const P c = ts[2 * e], s = INV ? -ts[2 * e + 1] : ts[2 * e + 1];
            vr[q] = fma(xr, c, -(xi * s));
            vi[q] = fma(xr, s, xi * c);
        }
        ring_bfly<P, R, INV>(vr, vi);
        P* dst = b + 2 * r * L;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for(int q = 0; q < R; q++){ dst[2 * (base + q * NS)] = vr[q]; dst[2 * (base + q * NS) + 1] = vi[q]; }
    }
}

// The stages R0..R3 (R3 = 1: three stages) from buffer a; returns the buffer holding the result.
template<typename P, int L, int R0, int R1, int R2, int R3, bool INV, int RB>
__device__ __forceinline__ P* ring_fft(P* a, P* b, const P* ts){
    ring_stage<P, L, R0, 1, INV, RB>(a, b, ts);
    __syncthreads();
    ring_stage<P, L, R1, R0, INV, RB>(b, a, ts);
    __syncthreads();
    ring_stage<P, L, R2, R0 * R1, INV, RB>(a, b, ts);
    __syncthreads();
    if constexpr(R3 > 1){
        ring_stage<P, L, R3, R0 * R1 * R2, INV, RB>(b, a, ts);
        __syncthreads();
        return a;
    } else {
        return b;
    }
}

template<typename P> struct Pair2;
template<> struct Pair2<float> { using T = float2; };
template<> struct Pair2<double> { using T = double2; };

constexpr int kRingThreads = 256;

template<typename P, int L, int R0, int R1, int R2, int R3, int RB>
__global__ void __launch_bounds__(kRingThreads) k_ring_fwd(const P* __restrict__ f, int kp, long long R,
                                                           const P* __restrict__ tw, P* __restrict__ c1){
    static_assert(R0 * R1 * R2 * R3 == L, "ring transforms: the radices multiply to L");
    constexpr int n = 2 * L;
    using V2 = typename Pair2<P>::T;
    __shared__ P A[RB * n], B[RB * n];
    __shared__ P ts[2 * n];
    const int tid = threadIdx.x, nt = blockDim.x;
    const long long row0 = (long long)blockIdx.x * RB;
    for(int q = tid; q < 2 * n; q += nt) ts[q] = tw[q];
    {
        // every load of the thread issued before the first LDS store
        constexpr int NQ = (RB * L + kRingThreads - 1) / kRingThreads;
        V2 v[NQ];
#pragma unroll
        for(int u = 0; u < NQ; u++){
            const int q = tid + u * kRingThreads;
            const int r = q / L, t = q - r * L;
            const long long row = row0 + r;
            if(q < RB * L && row < R) v[u] = *reinterpret_cast<const V2*>(f + row * kp + 2 * t);
            else { v[u].x = (P)0; v[u].y = (P)0; }
        }
#pragma unroll
        for(int u = 0; u < NQ; u++){
            const int q = tid + u * kRingThreads;
            if(q < RB * L){ A[2 * q] = v[u].x; A[2 * q + 1] = v[u].y; }
        }
    }
    __syncthreads();
    P* Z = ring_fft<P, L, R0, R1, R2, R3, false, RB>(A, B, ts);
    P* O = Z == A ? B : A;
    // the split into the spectrum of x, modes [mode][r]
    for(int it = tid; it < RB * (L + 1); it += nt){
        const int r = it / (L + 1), w = it - r * (L + 1);
        const P* z = Z + 2 * r * L;
        const int wa = w == L ? 0 : w, wb = w == 0 ? 0 : L - w;
        const P zr = z[2 * wa], zi = z[2 * wa + 1];
        const P yr = z[2 * wb], yi = -z[2 * wb + 1];                   // conj Z[L - w]
        const P Er = (P)0.5 * (zr + yr), Ei = (P)0.5 * (zi + yi);
        const P Or = (P)0.5 * (zi - yi), Oi = (P)-0.5 * (zr - yr);   // (Z - conj Z') / 2i
        const P c = ts[2 * w], s = ts[2 * w + 1];                     // W_n^w (w = L: -1)
        const P Xr = Er + fma(Or, c, -(Oi * s)), Xi = Ei + fma(Or, s, Oi * c);
        O[w * RB + r] = Xr;
        if(w >= 1 && w < L) O[(L + w) * RB + r] = -Xi;
    }
    __syncthreads();
    for(int q = tid; q < RB * n; q += nt){
        const int mp = q / RB, r = q - mp * RB;
        const long long row = row0 + r;
        if(row < R) c1[mp * R + row] = O[q];
    }
}

// ACC: the result is added to the interior cells of out (1 <= i <= im-2, 1 <= j <= jm-2) instead of
// stored to every row.
template<int L, int R0, int R1, int R2, int R3, int RB, bool ACC>
__global__ void __launch_bounds__(kRingThreads) k_ring_inv(const double* __restrict__ c1, long long R,
                                                           const double* __restrict__ tw, double* __restrict__ out, Dims D){
    static_assert(R0 * R1 * R2 * R3 == L, "ring transforms: the radices multiply to L");
    constexpr int n = 2 * L;
    __shared__ double A[RB * n], B[RB * n];
    __shared__ double ts[2 * n];
    const int tid = threadIdx.x, nt = blockDim.x;
    const long long row0 = (long long)blockIdx.x * RB;
    for(int q = tid; q < 2 * n; q += nt) ts[q] = tw[q];
    {
        constexpr int NQ = (RB * n + kRingThreads - 1) / kRingThreads;
        double v[NQ];
#pragma unroll
        for(int u = 0; u < NQ; u++){
            const int q = tid + u * kRingThreads;
            const int mp = q / RB, r = q - mp * RB;
            const long long row = row0 + r;
            v[u] = (q < RB * n && row < R) ? c1[mp * R + row] : 0.0;
        }
#pragma unroll
        for(int u = 0; u < NQ; u++){
            const int q = tid + u * kRingThreads;
            if(q < RB * n) A[q] = v[u];
        }
    }
    __syncthreads();
    // Z[w] = (E + i O) / L from the Hermitian spectrum, E = (X[w] + conj X[L-w]) / 2,
    // O = (X[w] - conj X[L-w]) W_n^-w / 2
    for(int it = tid; it < RB * L; it += nt){
        const int r = it / L, w = it - r * L;
        const int wb = L - w;                                        // 1..L
        const double xr = A[w * RB + r], xi = (w == 0) ? 0.0 : -A[(L + w) * RB + r];
        const double yr = A[wb * RB + r], yi = (wb == L) ? 0.0 : A[(L + wb) * RB + r];   // conj X[L - w]
        const double Er = xr + yr, Ei = xi + yi;
        const double dr = xr - yr, di = xi - yi;
        const double c = ts[2 * w], s = -ts[2 * w + 1];               // W_n^-w
        const double Or = fma(dr, c, -(di * s)), Oi = fma(dr, s, di * c);
        B[2 * (r * L + w)] = (Er - Oi) * (1.0 / n);
        B[2 * (r * L + w) + 1] = (Ei + Or) * (1.0 / n);
    }
    __syncthreads();
    double* Z = ring_fft<double, L, R0, R1, R2, R3, true, RB>(B, A, ts);
    // the rows out (interior rows only with ACC), every load of the thread before its first store
    constexpr int NQ = (RB * L + kRingThreads - 1) / kRingThreads;
    double2 u[NQ];
    bool live[NQ];
#pragma unroll
    for(int k = 0; k < NQ; k++){
        const int q = tid + k * kRingThreads;
        const int r = q / L;
        const long long row = row0 + r;
        live[k] = q < RB * L && row < R;
        if(ACC && live[k]){
            const int j = (int)(row % D.jm), i = (int)((row / D.jm) % D.im);
            live[k] = i >= 1 && i <= D.im - 2 && j >= 1 && j <= D.jm - 2;
        }
        u[k].x = 0.0; u[k].y = 0.0;
        if(ACC && live[k]) u[k] = *reinterpret_cast<const double2*>(out + row * D.kp + 2 * (q - r * L));
    }
#pragma unroll
    for(int k = 0; k < NQ; k++){
        const int q = tid + k * kRingThreads;
        const int r = q / L;
        if(!live[k]) continue;
        double2 v;
        v.x = Z[2 * q] + u[k].x; v.y = Z[2 * q + 1] + u[k].y;
        *reinterpret_cast<double2*>(out + (row0 + r) * D.kp + 2 * (q - r * L)) = v;
    }
}

// ---- host: set-up and the loop --------------------------------------------------------------

template<typename P> struct SolveMats {
    P* fwd = nullptr;    // [m'][k]   row-major: the forward real DFT, column-major (k x m') for GEMM
    P* inv = nullptr;    // [m'][k]   the inverse as a column-major (k x m') matrix
    P* vinv = nullptr;   // [m][e + nj j]  per ring wavenumber, theta modes from theta values
    P* v = nullptr;      // [m][j + nj e]  theta values from theta modes
    P* f = nullptr;      // field-shaped: the scaled residual, then the correction
    P* c1 = nullptr;     // [m'][rows]  ring modes
    P* c2 = nullptr;     // [m'][rows]  ring and theta modes
    P* tw = nullptr;     // [2 nk]  W_n^j of the ring transforms (when they are not GEMMs)
};

struct Solver {
    Dims D{};
    Scal S{};
    int nk = 0, nj = 0, ni = 0, half = 0, null_e = 0;
    long long R = 0;                        // rows = members * im * jm
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    double* d_mu = nullptr;                 // [(half+1) * nj]
    double* d_cf = nullptr;                 // [(half+1) * ni * nj]  Thomas factors c' of the radial solves
    double* d_rd = nullptr;                 // [(half+1) * ni * nj]  and 1 / den
    double* d_s = nullptr;                  // [im]  rm^2 num1
    double* d_rm2 = nullptr;                // [im]
    double* d_part = nullptr;               // [R * 2]
    double* d_red = nullptr;                // [M * 2] x 4: norms, src mean, p mean at entry, p mean now
    int* d_status = nullptr;                // worst end of a solve run without a residual check
                                            // (graph capture) since it was last read: 0 ok,
                                            // 1 above its limit, 2 non-finite
    double wtot = 0.0;                      // sum of the weights over one member's interior cells
    SolveMats<double> m64;
    SolveMats<float> m32;
    // host tables
    std::vector<double> fwd, inv, vinv, v, mu;
    std::vector<double> tw;                 // W_n^j, j = 0..nk-1 (cos, -sin)
    bool ring = false;                      // the ring transforms by k_ring_fwd / k_ring_inv
    int last_iters = 0;
    double last_rel = -1.0;
    bool last_floor = false;                // the last solve stopped at the FP64 rounding floor
    std::string err;
};

namespace {

template<typename P>
int alloc_mats(Solver& s, SolveMats<P>& M){
    if(M.f) return ATSAT_OK;
    const size_t nkk = (size_t)s.nk * s.nk, nv = (size_t)(s.half + 1) * s.nj * s.nj;
    const size_t field = (size_t)s.D.fs * s.D.M;
    std::vector<P> tmp;
    auto up = [&](P** d, const std::vector<double>& h) -> bool {
        tmp.assign(h.begin(), h.end());
        if(hipMalloc(reinterpret_cast<void**>(d), tmp.size() * sizeof(P)) != hipSuccess){ *d = nullptr; return false; }
        return hipMemcpy(*d, tmp.data(), tmp.size() * sizeof(P), hipMemcpyHostToDevice) == hipSuccess;
    };
    (void)nkk; (void)nv;
    if((!s.ring && (!up(&M.fwd, s.fwd) || !up(&M.inv, s.inv))) || !up(&M.vinv, s.vinv) || !up(&M.v, s.v)
       || (s.ring && !up(&M.tw, s.tw))){
        s.err = "solver: matrix upload failed"; return ATSAT_E_NOMEM;
    }
    for(P** b : {&M.f, &M.c1, &M.c2}){
        if(hipMalloc(reinterpret_cast<void**>(b), field * sizeof(P)) != hipSuccess){ *b = nullptr; s.err = "solver: scratch allocation failed"; return ATSAT_E_NOMEM; }
        if(hipMemset(*b, 0, field * sizeof(P)) != hipSuccess){ s.err = "solver: scratch clear failed"; return ATSAT_E_HIP; }
    }
    return ATSAT_OK;
}

template<typename P>
void free_mats(SolveMats<P>& M){
    for(P* p : {M.fwd, M.inv, M.vinv, M.v, M.f, M.c1, M.c2, M.tw}) if(p) (void)hipFree(p);
    M = SolveMats<P>();
}

rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                    const double* A, int lda, const double* B, int ldb, double* C, int ldc){
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}
rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                    const float* A, int lda, const float* B, int ldb, float* C, int ldc){
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
}
rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const double* A, int lda, long long sa,
                       const double* B, int ldb, long long sb, double* C, int ldc, long long sc, int count){
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one,
                                         A, lda, sa, B, ldb, sb, &zero, C, ldc, sc, count);
}
rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const float* A, int lda, long long sa,
                       const float* B, int ldb, long long sb, float* C, int ldc, long long sc, int count){
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one,
                                         A, lda, sa, B, ldb, sb, &zero, C, ldc, sc, count);
}

// The ring lengths with a two-pass plan (n = N1 N2); other lengths take the transforms as GEMMs.
bool ring_plan(int n){ return n == 360 || n == 180 || n == 120 || n == 90 || n == 60; }

template<typename P, int L, int R0, int R1, int R2, int R3>
void ring_fwd_launch(const Solver& s, const P* f, const P* tw, P* c1){
    constexpr int RB = sizeof(P) == 4 ? 16 : 8;
    const unsigned blocks = (unsigned)((s.R + RB - 1) / RB);
    hipLaunchKernelGGL((k_ring_fwd<P, L, R0, R1, R2, R3, RB>), dim3(blocks), dim3(kRingThreads), 0, s.stream,
                       f, s.D.kp, s.R, tw, c1);
}
template<typename P>
void ring_fwd(const Solver& s, const P* f, const P* tw, P* c1){
    switch(s.nk){
    case 360: ring_fwd_launch<P, 180, 4, 5, 3, 3>(s, f, tw, c1); break;
    case 180: ring_fwd_launch<P, 90, 2, 5, 3, 3>(s, f, tw, c1); break;
    case 120: ring_fwd_launch<P, 60, 4, 5, 3, 1>(s, f, tw, c1); break;
    case 90:  ring_fwd_launch<P, 45, 5, 3, 3, 1>(s, f, tw, c1); break;
    default:  ring_fwd_launch<P, 30, 2, 5, 3, 1>(s, f, tw, c1); break;       // 60
    }
}
template<int L, int R0, int R1, int R2, int R3>
void ring_inv_launch(const Solver& s, const double* c1, const double* tw, double* out, bool acc){
    constexpr int RB = 8;
    const unsigned blocks = (unsigned)((s.R + RB - 1) / RB);
    if(acc) hipLaunchKernelGGL((k_ring_inv<L, R0, R1, R2, R3, RB, true>), dim3(blocks), dim3(kRingThreads), 0, s.stream,
                               c1, s.R, tw, out, s.D);
    else    hipLaunchKernelGGL((k_ring_inv<L, R0, R1, R2, R3, RB, false>), dim3(blocks), dim3(kRingThreads), 0, s.stream,
                               c1, s.R, tw, out, s.D);
}
void ring_inv(const Solver& s, const double* c1, const double* tw, double* out, bool acc){
    switch(s.nk){
    case 360: ring_inv_launch<180, 4, 5, 3, 3>(s, c1, tw, out, acc); break;
    case 180: ring_inv_launch<90, 2, 5, 3, 3>(s, c1, tw, out, acc); break;
    case 120: ring_inv_launch<60, 4, 5, 3, 1>(s, c1, tw, out, acc); break;
    case 90:  ring_inv_launch<45, 5, 3, 3, 1>(s, c1, tw, out, acc); break;
    default:  ring_inv_launch<30, 2, 5, 3, 1>(s, c1, tw, out, acc); break;
    }
}

// e = M^-1 (rm^2 r): F.f holds rm^2 r on entry, Bk.f the correction on return. The forward
// transforms run in Pf, the radial solves in FP64 and the backward transforms in Pb. The
// backward half is the one whose rounding the operator amplifies (a correction rounded to
// eps relative carries white noise that Lc multiplies by its largest eigenvalue, ~1e8 times
// its smallest on the native grid), so it stays FP64 when the forward half is FP32.
// With the two-pass ring transforms and acc given, the correction is added to the interior cells
// of acc instead (*added = true).
template<typename Pf>
int separable_solve(Solver& s, SolveMats<Pf>& F, SolveMats<double>& Bk, double* acc = nullptr, bool* added = nullptr){
    const Dims& D = s.D;
    const int nk = s.nk, nj = s.nj, half = s.half;
    const long long R = s.R;
    const int cols = D.M * D.im;                            // (member, i) columns of a ring mode
    rocblas_status st;
    if(added) *added = false;
    // ring modes: c1 (R x nk, column-major) = f^T (R x nk) * fwd (nk x nk)
    if(s.ring) ring_fwd<Pf>(s, F.f, F.tw, F.c1);
    else {
        st = gemm(s.blas, rocblas_operation_transpose, rocblas_operation_none, (int)R, nk, nk, F.f, D.kp, F.fwd, nk, F.c1, (int)R);
        if(st != rocblas_status_success){ s.err = "solver: rocBLAS ring transform failed"; return ATSAT_E_HIP; }
    }
    // theta modes, cos then sin ring modes (the sin modes start at wavenumber 1)
    st = gemm_sb(s.blas, nj, cols, nj, F.vinv, nj, (long long)nj * nj, F.c1 + 1, D.jm, R, F.c2 + 1, D.jm, R, half + 1);
    if(st == rocblas_status_success && half > 1)
        st = gemm_sb(s.blas, nj, cols, nj, F.vinv + (size_t)nj * nj, nj, (long long)nj * nj,
                     F.c1 + (half + 1) * R + 1, D.jm, R, F.c2 + (half + 1) * R + 1, D.jm, R, half - 1);
    if(st != rocblas_status_success){ s.err = "solver: rocBLAS theta transform failed"; return ATSAT_E_HIP; }
    if(s.ni <= 40)
        hipLaunchKernelGGL((k_sol_radial<Pf, double, 40>), dim3((nj + 255) / 256, nk, D.M), dim3(256), 0, s.stream,
                           D, F.c2, Bk.c2, s.d_cf, s.d_rd, s.d_s, R, nj, s.null_e);
    else
        hipLaunchKernelGGL((k_sol_radial<Pf, double, 64>), dim3((nj + 255) / 256, nk, D.M), dim3(256), 0, s.stream,
                           D, F.c2, Bk.c2, s.d_cf, s.d_rd, s.d_s, R, nj, s.null_e);
    st = gemm_sb(s.blas, nj, cols, nj, Bk.v, nj, (long long)nj * nj, Bk.c2 + 1, D.jm, R, Bk.c1 + 1, D.jm, R, half + 1);
    if(st == rocblas_status_success && half > 1)
        st = gemm_sb(s.blas, nj, cols, nj, Bk.v + (size_t)nj * nj, nj, (long long)nj * nj,
                     Bk.c2 + (half + 1) * R + 1, D.jm, R, Bk.c1 + (half + 1) * R + 1, D.jm, R, half - 1);
    if(st != rocblas_status_success){ s.err = "solver: rocBLAS inverse theta transform failed"; return ATSAT_E_HIP; }
    // back to the ring: f (nk x R, column-major, ld kp) = inv (nk x nk) * c1^T
    if(s.ring){
        ring_inv(s, Bk.c1, Bk.tw, acc ? acc : Bk.f, acc != nullptr);
        if(added) *added = acc != nullptr;
        return ATSAT_OK;
    }
    st = gemm(s.blas, rocblas_operation_none, rocblas_operation_transpose, nk, (int)R, nk, Bk.inv, nk, Bk.c1, (int)R, Bk.f, D.kp);
    if(st != rocblas_status_success){ s.err = "solver: rocBLAS inverse ring transform failed"; return ATSAT_E_HIP; }
    return ATSAT_OK;
}

}  // namespace

Solver* solver_create(const Dims& D, const Scal& S, const double* thez_sin, const double* rm, hipStream_t stream,
                      std::string* err){
    Solver* s = new Solver();
    s->D = D; s->S = S; s->stream = stream;
    s->nk = D.km - 1; s->nj = D.jm - 2; s->ni = D.im - 2; s->half = s->nk / 2;
    s->R = (long long)D.M * D.im * D.jm;
    if(s->ni > 64){ *err = "converged pressure solve: more than 64 interior levels"; delete s; return nullptr; }
    const int nk = s->nk, nj = s->nj, half = s->half;
    const double pi = M_PI;
    // the real DFT of the ring: cos modes 0..n/2, sin modes 1..n/2-1 stored after them
    s->fwd.assign((size_t)nk * nk, 0.0);
    s->inv.assign((size_t)nk * nk, 0.0);
    for(int mp = 0; mp < nk; mp++){
        const bool sn = mp > half;
        const int w = sn ? mp - half : mp;
        for(int k = 0; k < nk; k++){
            const double a = 2.0 * pi * (double)((long long)w * k % nk) / nk;
            const double b = sn ? std::sin(a) : std::cos(a);
            s->fwd[(size_t)mp * nk + k] = b;
            const double scale = (w == 0 || (!sn && w == half)) ? 1.0 / nk : 2.0 / nk;
            s->inv[(size_t)mp * nk + k] = scale * b;
        }
    }
    s->ring = ring_plan(nk);
    if(s->ring){
        s->tw.assign((size_t)2 * nk, 0.0);
        for(int j = 0; j < nk; j++){
            const double a = 2.0 * pi * (double)j / nk;
            s->tw[2 * j] = std::cos(a);
            s->tw[2 * j + 1] = -std::sin(a);
        }
    }
    // theta operators: per wavenumber, dtheta^-2 D_theta (folded ends) - dphi^-2 lambda_m S(j)
    s->mu.assign((size_t)(half + 1) * nj, 0.0);
    s->vinv.assign((size_t)(half + 1) * nj * nj, 0.0);
    s->v.assign((size_t)(half + 1) * nj * nj, 0.0);
    std::vector<double> Sj(nj);
    for(int q = 0; q < nj; q++){
        double sp = thez_sin[q + 1];                              // PressureSolver.h:156-162
        if(sp == 0.0) sp = 1.0e-5;
        Sj[q] = 1.0 / (sp * sp);
    }
    const double two3 = 2.0 / 3.0;
    std::vector<double> up(nj), lo(nj), dsc(nj);
    for(int q = 0; q < nj; q++){
        up[q] = (q < nj - 1) ? S.inv_dthe2 * (q == 0 ? two3 : 1.0) : 0.0;          // L(q, q+1)
        lo[q] = (q > 0) ? S.inv_dthe2 * (q == nj - 1 ? two3 : 1.0) : 0.0;          // L(q, q-1)
    }
    dsc[0] = 1.0;
    for(int q = 0; q + 1 < nj; q++) dsc[q + 1] = dsc[q] * std::sqrt(up[q] / lo[q + 1]);
    bool ok = true;
    auto work = [&](int m0, int m1){
        std::vector<double> d(nj), e(nj, 0.0), z;
        for(int m = m0; m < m1; m++){
            const double lam = 2.0 - 2.0 * std::cos(2.0 * pi * m / nk);
            for(int q = 0; q < nj; q++){
                d[q] = S.inv_dthe2 * ((q == 0 || q == nj - 1) ? -two3 : -2.0) - S.inv_dphi2 * lam * Sj[q];
                e[q] = (q < nj - 1) ? std::sqrt(up[q] * lo[q + 1]) : 0.0;
            }
            if(!tridiag_eigen(nj, d, e, z)){ ok = false; return; }
            for(int ee = 0; ee < nj; ee++){
                s->mu[(size_t)m * nj + ee] = d[ee];
                for(int q = 0; q < nj; q++){
                    const double zq = z[(size_t)q * nj + ee];
                    s->vinv[(size_t)m * nj * nj + ee + (size_t)nj * q] = zq * dsc[q];
                    s->v[(size_t)m * nj * nj + q + (size_t)nj * ee] = zq / dsc[q];
                }
            }
        }
    };
    const int nthr = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const int per = (half + 1 + nthr - 1) / nthr;
    for(int t = 0; t < nthr; t++){
        const int a = t * per, b = std::min(half + 1, a + per);
        if(a < b) th.emplace_back(work, a, b);
    }
    for(auto& x : th) x.join();
    if(!ok){ *err = "converged pressure solve: eigen-decomposition did not converge"; delete s; return nullptr; }
    // the null mode of wavenumber 0: the theta mode of the smallest |mu|
    int ne = 0;
    for(int ee = 1; ee < nj; ee++) if(std::fabs(s->mu[ee]) < std::fabs(s->mu[ne])) ne = ee;
    s->null_e = ne;
    // radial factors and the weight total
    std::vector<double> hs(D.im), hr(D.im);
    for(int i = 0; i < D.im; i++){ hr[i] = rm[i] * rm[i]; hs[i] = hr[i] * S.inv_dr2; }
    double wi = 0.0, wj = 0.0;
    for(int i = 1; i <= D.im - 2; i++) wi += (i == 1 || i == D.im - 2) ? 1.5 : 1.0;
    for(int j = 1; j <= D.jm - 2; j++) wj += (j == 1 || j == D.jm - 2) ? 1.5 : 1.0;
    s->wtot = wi * wj * (double)nk;
    // the radial elimination factors of every (wavenumber, theta mode), as k_sol_radial applies them
    const int ni = s->ni;
    std::vector<double> cf((size_t)(half + 1) * ni * nj, 0.0), rdn((size_t)(half + 1) * ni * nj, 0.0);
    for(int w = 0; w <= half; w++){
        for(int ee = 0; ee < nj; ee++){
            const bool pin = w == 0 && ee == s->null_e;
            const double mue = pin ? 0.0 : s->mu[(size_t)w * nj + ee];
            const int n = pin ? ni - 1 : ni;
            double cprev = 0.0;
            for(int i = 1; i <= n; i++){
                const double si = hs[i];
                const double a = (i == 1) ? 0.0 : ((i == ni) ? two3 * si : si);
                const double b = ((i == 1 || i == ni) ? -two3 * si : -2.0 * si) + mue;
                const double c = (i == ni || i == n) ? 0.0 : ((i == 1) ? two3 * si : si);
                const double den = b - a * cprev;
                const size_t o = ((size_t)w * ni + (i - 1)) * nj + ee;
                cf[o] = c / den;
                rdn[o] = 1.0 / den;
                cprev = cf[o];
            }
        }
    }
    auto dmalloc = [&](double** p, size_t n) -> bool {
        return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(double)) == hipSuccess; };
    if(!dmalloc(&s->d_mu, s->mu.size()) || !dmalloc(&s->d_cf, cf.size()) || !dmalloc(&s->d_rd, rdn.size())
       || !dmalloc(&s->d_s, D.im) || !dmalloc(&s->d_rm2, D.im)
       || !dmalloc(&s->d_part, (size_t)s->R * 2) || !dmalloc(&s->d_red, (size_t)D.M * 2 * 4)){
        *err = "converged pressure solve: device allocation failed"; solver_destroy(s); return nullptr;
    }
    if(hipMalloc(reinterpret_cast<void**>(&s->d_status), sizeof(int)) != hipSuccess){
        s->d_status = nullptr;
        *err = "converged pressure solve: device allocation failed"; solver_destroy(s); return nullptr;
    }
    const int zero = 0;
    if(hipMemcpy(s->d_status, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(s->d_mu, s->mu.data(), s->mu.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(s->d_cf, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(s->d_rd, rdn.data(), rdn.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(s->d_s, hs.data(), D.im * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(s->d_rm2, hr.data(), D.im * sizeof(double), hipMemcpyHostToDevice) != hipSuccess){
        *err = "converged pressure solve: table upload failed"; solver_destroy(s); return nullptr;
    }
    if(rocblas_create_handle(&s->blas) != rocblas_status_success){
        *err = "converged pressure solve: rocblas_create_handle failed"; solver_destroy(s); return nullptr;
    }
    rocblas_set_stream(s->blas, stream);
    return s;
}

void solver_destroy(Solver* s){
    if(!s) return;
    free_mats(s->m64);
    free_mats(s->m32);
    for(double* p : {s->d_mu, s->d_cf, s->d_rd, s->d_s, s->d_rm2, s->d_part, s->d_red}) if(p) (void)hipFree(p);
    if(s->d_status) (void)hipFree(s->d_status);
    if(s->blas) rocblas_destroy_handle(s->blas);
    delete s;
}

size_t solver_device_bytes(const Solver* s){
    if(!s) return 0;
    const size_t field = (size_t)s->D.fs * s->D.M;
    const size_t mats = (s->ring ? 2 * (size_t)s->nk : 2 * (size_t)s->nk * s->nk) + 2 * (size_t)(s->half + 1) * s->nj * s->nj;
    size_t b = (s->mu.size() * (1 + 2 * (size_t)s->ni) + 2 * (size_t)s->D.im + (size_t)s->R * 2 + (size_t)s->D.M * 8) * sizeof(double);
    if(s->m64.f) b += (3 * field + mats) * sizeof(double);
    if(s->m32.f) b += (3 * field + mats) * sizeof(float);
    return b;
}

const char* solver_error(const Solver* s){ return s ? s->err.c_str() : ""; }
// the stream the solver's kernels and GEMMs go to (host state only; also during a capture)
void solver_set_stream(Solver* s, hipStream_t st){
    if(s->stream == st) return;
    s->stream = st;
    rocblas_set_stream(s->blas, st);
}
void solver_last(const Solver* s, int* iters, double* rel){ *iters = s->last_iters; *rel = s->last_rel; }
bool solver_last_floor(const Solver* s){ return s->last_floor; }

// Prepare the FP64 or FP32 correction buffers (allocations: call outside graph capture).
int solver_prepare(Solver* s, bool fp32){
    int rc = alloc_mats(*s, s->m64);
    if(rc == ATSAT_OK && fp32) rc = alloc_mats(*s, s->m32);
    return rc;
}

int solver_warmup(Solver* s, bool fp32){
    return fp32 ? separable_solve(*s, s->m32, s->m64) : separable_solve(*s, s->m64, s->m64);
}

namespace {

// The weighted means of x per member into d_red[slot] (unnormalised sums).
void wsum(Solver& s, const double* x, int slot){
    const Dims& D = s.D;
    hipLaunchKernelGGL(k_sol_wsum, dim3((unsigned)s.R), dim3(kSolveRow), 0, s.stream, D, x, s.d_part);
    hipLaunchKernelGGL(k_sol_reduce, dim3(D.M), dim3(256), 0, s.stream, D.im * D.jm, s.d_part, s.d_red + 2 * D.M * slot);
}

// The end of a solve run without a residual check on the host (inside a graph capture): per member
// the relative residual of the final iterate, formed as the host forms it, against `limit`. The
// status word keeps the worst outcome until the host reads it (solver_status): 2 non-finite,
// 1 above the limit.
__global__ void k_sol_status(int M, const double* red, double limit, int* status){
    if(threadIdx.x != 0 || blockIdx.x != 0) return;
    int st = *status;
    for(int m = 0; m < M; m++){
        const double ss = red[2 * m + 1], rr = red[2 * m];
        const double q = ss > 0.0 ? sqrt(rr / ss) : sqrt(rr);
        if(!isfinite(q)) st = 2;
        else if(q > limit && st < 1) st = 1;
    }
    *status = st;
}

template<typename P>
int run(Solver& s, SolveMats<P>& M, double* p, double* src, double* dp, const double* geo, double tol,
        int max_iter, bool host_check, double status_limit){
    const Dims& D = s.D;
    int rc;
    if((rc = alloc_mats(s, M)) != ATSAT_OK) return rc;
    if((rc = alloc_mats(s, s.m64)) != ATSAT_OK) return rc;
    const dim3 rows((unsigned)s.R);
    const dim3 cells((D.km + 255) / 256, (unsigned)s.R);
    const dim3 bgrid_r((D.km + 255) / 256, D.jm, D.M), bgrid_t((D.km + 255) / 256, D.im, D.M);
    const dim3 bgrid_s((D.jm + 255) / 256, D.im, D.M);
    auto bnd = [&](double* x){
        hipLaunchKernelGGL(k_sol_bnd_r, bgrid_r, dim3(256), 0, s.stream, D, x);
        hipLaunchKernelGGL(k_sol_bnd_t, bgrid_t, dim3(256), 0, s.stream, D, x);
        hipLaunchKernelGGL(k_sol_seam, bgrid_s, dim3(256), 0, s.stream, D, x);
    };
    // p_dyn may enter with boundary values the boundary relations do not give (the Shapiro
    // filter ran after the last boundary pass): fold them, then solve for the increment
    // dp = p_new - p: Lc dp = src - Lc p. The residual of dp is formed from small numbers, where
    // the one of p itself (a large, nearly balanced field) would cancel down to rounding.
    bnd(p);
    const RowMap RM = all_row_map(D);
    hipLaunchKernelGGL(k_sol_defect, dim3(RM.blocks()), dim3(kSolveRow), 0, s.stream, D, s.S, p, src, geo, RM);
    hipLaunchKernelGGL(k_sol_fill, cells, dim3(256), 0, s.stream, D, dp, 0.0);
    // the solvable part of the source
    wsum(s, src, 1);
    hipLaunchKernelGGL(k_sol_shift, cells, dim3(256), 0, s.stream, D, src, s.d_red + 2 * D.M * 1, s.wtot);
    std::vector<double> red(2 * D.M);
    int it = 0;
    double rel = -1.0, prev = -1.0;
    bool floor = false;
    for(;; it++){
        hipLaunchKernelGGL(k_sol_residual<P>, dim3(RM.blocks()), dim3(kSolveRow), 0, s.stream, D, s.S, dp, src, geo,
                           s.d_rm2, M.f, s.d_part, RM);
        hipLaunchKernelGGL(k_sol_reduce, dim3(D.M), dim3(256), 0, s.stream, D.im * D.jm, s.d_part, s.d_red);
        if(host_check){
            if(hipMemcpyAsync(red.data(), s.d_red, red.size() * sizeof(double), hipMemcpyDeviceToHost, s.stream) != hipSuccess
               || hipStreamSynchronize(s.stream) != hipSuccess){ s.err = "solver: residual read-back failed"; return ATSAT_E_HIP; }
            rel = 0.0;
            for(int m = 0; m < D.M; m++){
                const double ss = red[2 * m + 1], rr = red[2 * m];
                const double q = ss > 0.0 ? std::sqrt(rr / ss) : std::sqrt(rr);
                if(!std::isfinite(q)){ rel = q; break; }
                rel = std::max(rel, q);
            }
            if(!std::isfinite(rel)){ s.last_iters = it; s.last_rel = rel; s.err = "converged pressure solve: non-finite residual"; return ATSAT_E_NONFINITE; }
            if(rel <= tol || it >= max_iter) break;
            // the FP64 rounding floor: every FP64 p carries a residual of about
            // eps * den * |p| / |src|, and the smoothest modes of the operator make |p| / |src|
            // large; a correction that no longer halves the residual has reached it
            if(it >= 1 && rel > 0.5 * prev){ floor = true; break; }
            prev = rel;
        } else if(it >= max_iter) break;
        bool added = false;
        if((rc = separable_solve(s, M, s.m64, dp, &added)) != ATSAT_OK) return rc;
        if(!added) hipLaunchKernelGGL(k_sol_update<double>, cells, dim3(256), 0, s.stream, D, dp, s.m64.f);
        bnd(dp);
    }
    // without the host check, d_red holds the residual of the final iterate (the loop left after
    // forming it): its verdict goes to the status word
    if(!host_check) hipLaunchKernelGGL(k_sol_status, dim3(1), dim3(64), 0, s.stream, D.M, s.d_red, status_limit, s.d_status);
    // the increment without its null-space constant (p keeps its weighted mean), added to p
    wsum(s, dp, 3);
    hipLaunchKernelGGL(k_sol_add_shifted, cells, dim3(256), 0, s.stream, D, p, dp, s.d_red + 2 * D.M * 3, s.wtot);
    s.last_iters = it;
    s.last_rel = host_check ? rel : -1.0;
    s.last_floor = floor;
    if(host_check && rel > tol && !floor){ s.err = "converged pressure solve: tolerance not reached"; return ATSAT_E_NOCONV; }
    return ATSAT_OK;
}

}  // namespace

int solver_run(Solver* s, double* p, double* src, double* scratch, const double* geo, bool fp32, double tol,
               int max_iter, bool host_check, double status_limit){
    return fp32 ? run(*s, s->m32, p, src, scratch, geo, tol, max_iter, host_check, status_limit)
                : run(*s, s->m64, p, src, scratch, geo, tol, max_iter, host_check, status_limit);
}

int solver_status(Solver* s, int* status){
    *status = 0;
    if(!s || !s->d_status) return ATSAT_OK;
    if(hipMemcpy(status, s->d_status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess){
        s->err = "solver: status read-back failed"; return ATSAT_E_HIP;
    }
    if(*status != 0){
        const int zero = 0;
        if(hipMemcpy(s->d_status, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess){
            s->err = "solver: status reset failed"; return ATSAT_E_HIP;
        }
    }
    return ATSAT_OK;
}

}  // namespace atsat
