// The moist convective adjustment on the device: MoistConvectiveAdjustment<cSaturnModel>::run
// (planet/ConvectiveAdjustment.h:402-849), called after restoreVar and the closure every
// ATSAT_MCA iterations (cSaturnModel.cpp:1120-1145), at ATSAT's defaults: the Ledoux criterion with
// the condensate loading, no relaxation, 64 sweeps at most, the search on a column smoothed in
// latitude (MCA_SMOOTH 2), the radial boundary planes left out (MCA_INTERIOR), the t_0 phase rule,
// cp = cp_mix. Gases H2O, NH3, CH4 (liquid above t_0, ice below); tracers H2S (a gas of ep_h2s in
// the virtual temperature) and NH4SH (condensed, carrying its formation heat in h_L).
//
//   k_mca_smooth   the conserved variables h_L, q_t of every gas and the tracers' mass fractions of
//                  every cell with a valid state (:504-521) and the two latitude 1-2-1 passes over
//                  them (:522-539), one workgroup per 64 longitudes x 8 latitudes of a level: the
//                  conserved variables of 12 latitude rows and the first pass over 10 of them are
//                  held in LDS, the second pass is written once;
//   k_mca_census   one thread per adjacent pair of levels of every column: the parcel census of the
//                  smoothed column (:695-703), each pair's verdict a bit of a 64-bit mask per
//                  column (set by atomic OR into slots cleared by k_mca_cen_clear);
//   k_mca_mix      one thread per column with an unstable pair (the others end at once): the
//                  sweeps that mix unstable segments and grow them while their edges are unstable,
//                  the union of the segments mixed again in the column's own state, and the write
//                  back of every cell that was put to equilibrium (:714-826).
//
// The equilibrium temperature of (h_L, q_t) at (p, z) (:447-475) is the reference's Illinois
// regula falsi step for step. Its saturation mixing ratios are the reference's values with the work
// arranged differently and every rounding the same: log T is formed once per temperature for all
// gases; a gas with q_t = 0 takes no saturation value (it condenses nothing, so the tests q_t > q_s
// and the terms lat x max(0, q_t - q_s) are decided without it); the values at T_lo, formed for the
// saturation test, serve the bracket's lower end, and those of the bracket's current upper end
// serve the condensate of the root.
//
// Scratch (FP64, the context's MCA buffer, field-major like the arena): the smoothed variables,
// which a column that mixes then uses as its own search state, and every cell's equilibrium
// temperature and condensate. A column's bookkeeping over its levels -- which cells are in
// equilibrium, which were touched, which adjacent pairs were mixed together and the census --
// lives in 64-bit masks, so a column holds at most 64 levels.
//
// Everything is FP64 in the reference's operation order, on t and rho_mix (FP64) and the species
// widened from their class (FP32 in the FP32 step), written back rounded to it; the saturation
// vapour pressure takes the correctly rounded exp / log of crmath.hpp.
#include <algorithm>

#include "atsat_dev.hpp"
#include "crmath.hpp"

namespace atsat {

namespace {
__device__ __forceinline__ double smax(double a, double b){ return (a < b) ? b : a; }

constexpr int NG = 3, NT = 2, NV = 1 + NG + NT;
// scratch fields: the smoothed set (h_L, q_t x 3, tracers x 2), T_eq, the condensate x 3
enum { M_SH = 0, M_SQ = 1, M_SR = 4, M_TEQ = 6, M_CEQ = 7, M_NF = 10 };

// the per-gas and per-tracer constants, uniform over the launch (in the scratch buffer)
struct McaK {
    int fv[NG];                          // field of the vapour (cloud and ice follow it)
    double ep[NG], lv[NG], ls[NG], t0[NG];
    double span_l[NG];                   // max(lv, ls)
    double tv_w[NG];                     // 1 / ep - 1
    // saturation vapour pressure, liquid [0] and ice [1]: C, -L0, del_alf, del_bet, 1e-3 R
    double C[2][NG], nL0[2][NG], da[2][NG], db[2][NG], R3[2][NG];
    double rR3[2][NG];                   // 1 / R3 (the FP32 step's census)
    int ft[NT];
    int tkind[NT];                       // 1: a gas of ep in T_v (weight 1 / ep - 1), 2: condensed, 0: neither
    double tL[NT], tv_t[NT];             // formation heat in h_L; the T_v weight of kind 1
    double cp, grav, t_ref, r_mix, layer_h;
    // the census's early decisions: allowed (every ep > 0 and R > 0), and per gas whether q_s
    // does not fall where the phase changes at t_0 (the liquid exponent at t_0 not below the ice one)
    int early;
    int jump_ok[NG];
};

__device__ __forceinline__ double z_of(const McaK& K, int i){          // get_layer_height * 1e3
    const float h = (float)K.layer_h;
    return (double)((float)i * h) * 1.0e3;
}
__device__ __forceinline__ double rho_of(double r, double r_mix){       // rho_at
    return (r > 0.0 && isfinite(r)) ? r : r_mix;
}

// q_s of every gas with q_t > 0 at temperature T and pressure p (the svp of ATPhys.h:49-52 in the
// phase of T against t_0, times ep / p; a non-finite value is 0).
// DECIDE: the census of the FP32 step (W = float), which only decides whether a parcel's T_v exceeds
// its environment's by 1e-7 K, from FP32-rounded temperatures (~1e-6 K): there the saturation
// values take FP32 exp / log (1e-7 relative) and reciprocals formed once. The mixing, whose
// equilibria are written back to the state, keeps them in FP64.
template<typename W, bool DECIDE = false>
__device__ __forceinline__ void qsat_all(const McaK& K, double T, double p, const double* q, double* qs, double rp = 0.0){
    constexpr bool F = DECIDE && std::is_same<W, float>::value;
    bool any = false;
#pragma unroll
    for(int g = 0; g < NG; g++){ qs[g] = 0.0; any |= (q[g] > 0.0); }
    if(!any) return;
    const double lnT = F ? (double)logf((float)T) : crm::Phys<W>::log(T);
    const double rT = F ? 1.0 / T : 0.0;             // with DECIDE, rp = 1 / p from the caller
#pragma unroll
    for(int g = 0; g < NG; g++){
        if(!(q[g] > 0.0)) continue;
        const int s = (T >= K.t0[g]) ? 0 : 1;
        double E, v;
        if constexpr(F){
            E = (double)expf((float)(K.C[s][g] + (K.nL0[s][g] * rT + K.da[s][g] * lnT + K.db[s][g] * T) * K.rR3[s][g]));
            v = K.ep[g] * E * rp;
        } else {
            E = crm::Phys<W>::exp(K.C[s][g] + (K.nL0[s][g] / T + K.da[s][g] * lnT + K.db[s][g] * T) / K.R3[s][g]);
            v = K.ep[g] * E / p;
        }
        qs[g] = isfinite(v) ? v : 0.0;
    }
}

// cp T + g z - h - sum lat(T) max(0, q - q_s(T)), q_s given
__device__ __forceinline__ double fof(const McaK& K, double T, double z, double h, const double* q, const double* qs){
    double f = K.cp * T + K.grav * z - h;
#pragma unroll
    for(int g = 0; g < NG; g++){
        if(!(q[g] > 0.0)) continue;                  // lat x 0
        const double lat = (T >= K.t0[g]) ? K.lv[g] : K.ls[g];
        f -= lat * smax(0.0, q[g] - qs[g]);
    }
    return f;
}

// :447-475, the equilibrium temperature TT and condensate of (h, q) at (p, z). The evaluation
// points -- T_lo, the bracket's upper end, the iterates -- pass through one saturation site.
template<typename W>
__device__ __forceinline__ bool equil(const McaK& K, double h, const double* q, double p, double z, double& TT, double* cond){
    const double tol_T = 1.0e-10;
    const int n_iter = 60;
    const double T_lo = (h - K.grav * z) / K.cp;
    if(!(T_lo > 0.0) || !(p > 0.0)) return false;
    double a = T_lo, b = T_lo, fa = 0.0, fb = 0.0, x = T_lo;
    double qb[NG];
    int side = 0, it = 0, stage = 0;               // x is 0: T_lo, 1: the upper end, 2: an iterate
    for(;;){
        double qs[NG];
        qsat_all<W>(K, x, p, q, qs);
        if(stage == 0){
            bool any = false;
#pragma unroll
            for(int g = 0; g < NG; g++) any |= (q[g] > 0.0 && q[g] > qs[g]);
            if(!any){
                TT = T_lo;
#pragma unroll
                for(int g = 0; g < NG; g++) cond[g] = 0.0;
                return true;
            }
            double span = 0.0;
#pragma unroll
            for(int g = 0; g < NG; g++) span += K.span_l[g] * smax(0.0, q[g]);
            fa = fof(K, a, z, h, q, qs);
            b = T_lo + span / K.cp + 1.0e-9;
            x = b;
            stage = 1;
            continue;
        }
        if(stage == 1){
            fb = fof(K, b, z, h, q, qs);
#pragma unroll
            for(int g = 0; g < NG; g++) qb[g] = qs[g];
            stage = 2;
        } else {
            const double fc = fof(K, x, z, h, q, qs);
            const bool lower = fc < 0.0, done = fabs(fc) <= K.cp * tol_T;
            if(lower){ a = x; fa = fc; if(side == -1) fb *= 0.5; side = -1; }
            else     { b = x; fb = fc; if(side == +1) fa *= 0.5; side = +1; }
            if(done || !lower){
#pragma unroll
                for(int g = 0; g < NG; g++) qb[g] = qs[g];
            }
            it++;
            if(done){ a = b = x; break; }
        }
        if(!(it < n_iter && b - a > tol_T)) break;
        double c = (fb - fa != 0.0) ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
        if(!(c > a && c < b)) c = 0.5 * (a + b);
        x = c;
    }
    TT = b;
#pragma unroll
    for(int g = 0; g < NG; g++) cond[g] = (q[g] > 0.0) ? smax(0.0, q[g] - qb[g]) : 0.0;
    return true;
}

__device__ __forceinline__ double tr_energy(const McaK& K, const double* r){
    double e = 0.0;
#pragma unroll
    for(int t = 0; t < NT; t++) e += K.tL[t] * r[t];
    return e;
}

// the virtual temperature with the condensate loading
__device__ __forceinline__ double tv(const McaK& K, double TT, const double* v_, const double* cnd, const double* r){
    double f = 1.0;
#pragma unroll
    for(int g = 0; g < NG; g++){ f += v_[g] * K.tv_w[g]; f -= cnd[g]; }
#pragma unroll
    for(int t = 0; t < NT; t++){
        if(K.tkind[t] == 1) f += r[t] * K.tv_t[t];
        else if(K.tkind[t] == 2) f -= r[t];
    }
    return TT * f;
}
}  // namespace

// ---- the census's early decision ---------------------------------------------------------------
// The census needs only whether T_v of the parcel minus T_v of the cell exceeds 1e-7 K. Each
// equilibrium is the regula falsi above, taken one saturation evaluation at a time (EqRun); its
// brackets nest, so the temperature it ends with lies in the current bracket [a, b]. With
// F = 1 + sum (q - c) (1/ep - 1) - sum c + tracers and 0 <= c <= q, T_v = T F lies in
// [a F(c = q), b F(c = 0)]. Where every gas's exponent of the saturation pressure increases with T
// over the bracket (L0 + del_alf T + del_bet T^2 > 0 on it, and no fall of q_s at t_0), the
// condensate falls and F rises with T, and T_v lies between its values at a and b. The pair is
// decided as soon as these bounds settle the comparison by more than kDecide; otherwise both
// equilibria run to their end and the exact values decide, as in the reference. The bounds hold
// for the computed values to a few 1e-12 K; kDecide is 1e-8 K.
constexpr double kTolK = 1.0e-7, kDecide = 1.0e-8;

struct EqRun {
    double h, p, z;                      // h includes the tracers' formation energy
    double a, b, fa, fb, x;
    double tva, tvb;                     // T_v at a and at b (their own condensate)
    int side, it, stage;                 // stage: 0 x = T_lo, 1 x = b, 2 x = an iterate, 3 ended (T = b), 4 none
    bool mono;                           // T_v rises with T over [a, b]
};

__device__ __forceinline__ void eq_begin(const McaK& K, EqRun& E, double h, const double* q, double p, double z){
    E.h = h; E.p = p; E.z = z;
    E.side = 0; E.it = 0; E.fa = 0.0; E.fb = 0.0; E.mono = false; E.tva = 0.0; E.tvb = 0.0;
    const double T_lo = (h - K.grav * z) / K.cp;
    if(!(T_lo > 0.0) || !(p > 0.0)){ E.stage = 4; E.a = E.b = E.x = 0.0; return; }
    double span = 0.0;
#pragma unroll
    for(int g = 0; g < NG; g++) span += K.span_l[g] * smax(0.0, q[g]);
    E.a = T_lo; E.b = T_lo + span / K.cp + 1.0e-9; E.x = T_lo; E.stage = 0;
}

// L0 + da T + db T^2 > 0 on [Tl, Th] (T > 0), with a relative margin
__device__ __forceinline__ bool expo_rises(double L0, double da, double db, double Tl, double Th){
    const double scale = fabs(L0) + fabs(da) * Th + fabs(db) * Th * Th;
    const double m = 1.0e-9 * scale;
    if(!(L0 + da * Tl + db * Tl * Tl > m) || !(L0 + da * Th + db * Th * Th > m)) return false;
    // an interior minimum at T* = -da / (2 db) when db > 0: P(T*) = L0 - da^2 / (4 db)
    if(db > 0.0 && 2.0 * db * Tl < -da && -da < 2.0 * db * Th)
        return 4.0 * db * L0 - da * da > 4.0 * db * m;
    return true;
}

__device__ __forceinline__ bool tv_rises(const McaK& K, double a, double b, const double* q){
    if(!K.early) return false;
#pragma unroll
    for(int g = 0; g < NG; g++){
        if(!(q[g] > 0.0)) continue;
        const bool ice = a < K.t0[g], liq = b >= K.t0[g];
        if(ice && liq && !K.jump_ok[g]) return false;
        if(ice && !expo_rises(-K.nL0[1][g], K.da[1][g], K.db[1][g], a, fmin(b, K.t0[g]))) return false;
        if(liq && !expo_rises(-K.nL0[0][g], K.da[0][g], K.db[0][g], fmax(a, K.t0[g]), b)) return false;
    }
    return true;
}

// T_v at temperature T with the condensate max(0, q - qs): the reference's tv(T, q - c, c)
__device__ __forceinline__ double tv_at(const McaK& K, double T, const double* q, const double* qs, const double* r){
    double c[NG], v[NG];
#pragma unroll
    for(int g = 0; g < NG; g++){ c[g] = (q[g] > 0.0) ? smax(0.0, q[g] - qs[g]) : 0.0; v[g] = q[g] - c[g]; }
    return tv(K, T, v, c, r);
}

// one step of the regula falsi of equil(), qs = q_s at E.x
__device__ __forceinline__ void eq_advance(const McaK& K, EqRun& E, const double* q, const double* r, const double* qs){
    const double tol_T = 1.0e-10;
    const int n_iter = 60;
    const double tvx = tv_at(K, E.x, q, qs, r);
    if(E.stage == 0){
        bool any = false;
#pragma unroll
        for(int g = 0; g < NG; g++) any |= (q[g] > 0.0 && q[g] > qs[g]);
        E.tva = tvx;
        if(!any){
            // no condensate: T = T_lo, where q_s >= q makes every max(0, q - q_s) zero
            E.b = E.a; E.tvb = tvx; E.stage = 3;
            return;
        }
        E.fa = fof(K, E.a, E.z, E.h, q, qs);
        E.x = E.b;
        E.stage = 1;
        E.mono = tv_rises(K, E.a, E.b, q);
        return;
    }
    if(E.stage == 1){
        E.fb = fof(K, E.b, E.z, E.h, q, qs);
        E.tvb = tvx;
    } else {
        const double fc = fof(K, E.x, E.z, E.h, q, qs);
        const bool lower = fc < 0.0, done = fabs(fc) <= K.cp * tol_T;
        // the Illinois halving of the end that stays, as a product by 0.5 or 1 (exact): written
        // as two conditional halvings the compiler merges them into one store through a selected
        // address, which keeps fa and fb in scratch memory
        const double ha = (!lower && E.side == +1) ? 0.5 : 1.0, hb = (lower && E.side == -1) ? 0.5 : 1.0;
        if(lower){ E.a = E.x; E.fa = fc; E.tva = tvx; E.side = -1; }
        else     { E.b = E.x; E.fb = fc; E.tvb = tvx; E.side = +1; }
        E.fa *= ha; E.fb *= hb;
        E.it++;
        if(done){ E.a = E.b = E.x; E.tvb = tvx; E.stage = 3; return; }
    }
    if(!(E.it < n_iter && E.b - E.a > tol_T)){ E.stage = 3; return; }
    // the brackets nest, and what tv_rises establishes on [a, b] (each exponent above its margin,
    // no fall of q_s at t_0 inside) holds on every sub-bracket: once true it stays true
    if(!E.mono) E.mono = tv_rises(K, E.a, E.b, q);
    double c = (E.fb - E.fa != 0.0) ? E.b - E.fb * (E.b - E.a) / (E.fb - E.fa) : 0.5 * (E.a + E.b);
    if(!(c > E.a && c < E.b)) c = 0.5 * (E.a + E.b);
    E.x = c;
    E.stage = 2;
}

// bounds of the T_v the equilibrium ends with; F(c = q) = fmin, F(c = 0) = f0
__device__ __forceinline__ void eq_bounds(const McaK& K, const EqRun& E, double fmin, double f0, double& lo, double& hi){
    if(E.stage == 3){ lo = hi = E.tvb; return; }
    if(!K.early || !(fmin > 0.0)){ lo = -INFINITY; hi = INFINITY; return; }
    lo = (E.stage >= 1 && E.mono) ? E.tva : E.a * fmin;
    hi = (E.stage == 2 && E.mono) ? E.tvb : E.b * f0;
}

// ---- the smoothed conserved variables ------------------------------------------------------
constexpr int kSmK = 64, kSmJ = 8;

template<typename T>
__global__ void __launch_bounds__(kSmK * kSmJ) k_mca_smooth(Dims D, const McaK* __restrict__ Kp, Bufs<T> B){
    const McaK& K = *Kp;
    // rows: conserved rc = 0 .. kSmJ + 3 <-> latitude j0 - 2 + rc; first pass pr = 0 .. kSmJ + 1 <-> rc = pr + 1
    __shared__ double Cv[kSmJ + 4][NV][kSmK];
    __shared__ double Pv[kSmJ + 2][NV][kSmK];
    __shared__ unsigned char Fv[kSmJ + 4][kSmK];
    const int tx = threadIdx.x, ty = threadIdx.y;
    // the tiles (level plane, 64 longitudes, 8 latitudes), latitude tiles fastest; workgroup b takes
    // tile (b mod 8) per + b / 8, so that the workgroups the dispatcher hands to one XCD (b mod 8)
    // walk a contiguous run of tiles and a tile's two halo rows on either side were just read
    // through the same L2 by its neighbour tile
    const int nkb = (D.km + kSmK - 1) / kSmK, njt = (D.jm + kSmJ - 1) / kSmJ;
    const long long ntile = (long long)nkb * njt * D.im * D.M, per = (ntile + 7) / 8;
    const long long t = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if(t >= ntile) return;
    const int jt = (int)(t % njt), kb = (int)((t / njt) % nkb);
    const long long plane = t / ((long long)njt * nkb);
    const int k = kb * kSmK + tx, j0 = jt * kSmJ;
    const int i = (int)(plane % D.im), m = (int)(plane / D.im);
    const long long mo = (long long)m * D.fs;
    const long long FS = (long long)D.M * D.fs;
    const double zz = z_of(K, i);
    // the conserved variables (:504-521); rows outside the grid hold no valid state
    for(int rc = ty; rc < kSmJ + 4; rc += kSmJ){
        const int j = j0 - 2 + rc;
        unsigned char ok = 0;
        if(j >= 0 && j < D.jm && k < D.km){
            const long long n = mo + D.at(i, j, k);
            const double TT = B.s[ATSAT_T][n] * K.t_ref, rr = rho_of(B.s[ATSAT_RHO_MIX][n], K.r_mix);
            if((TT > 0.0) && (rr > 0.0) && isfinite(TT + rr)){
                double h = K.cp * TT + K.grav * zz;
                // the FP32 step (FP32 species) multiplies by 1 / rho formed once (qdiv)
                const double rrr = qrcp<T>(rr);
#pragma unroll
                for(int g = 0; g < NG; g++){
                    const double v_ = qdiv<T>(smax(0.0, (double)B.w[K.fv[g]][n]), rr, rrr),
                                 c_ = qdiv<T>(smax(0.0, (double)B.w[K.fv[g] + 1][n]), rr, rrr),
                                 i_ = qdiv<T>(smax(0.0, (double)B.w[K.fv[g] + 2][n]), rr, rrr);
                    Cv[rc][1 + g][tx] = v_ + c_ + i_;
                    h -= K.lv[g] * c_ + K.ls[g] * i_;
                }
#pragma unroll
                for(int t = 0; t < NT; t++){
                    const double q = qdiv<T>((double)B.w[K.ft[t]][n], rr, rrr);
                    Cv[rc][1 + NG + t][tx] = q;
                    h -= K.tL[t] * q;
                }
                Cv[rc][0][tx] = h;
                ok = 1;
            }
        }
        Fv[rc][tx] = ok;
    }
    __syncthreads();
    // the first 1-2-1 pass (:522-539) over rows rc = 1 .. kSmJ + 2: a valid neighbour weighs 0.25, a
    // missing one 0 (its slot taken by the cell itself); a cell without a valid state is skipped
    for(int pr = ty; pr < kSmJ + 2; pr += kSmJ){
        const int rc = pr + 1;
        if(!Fv[rc][tx]) continue;
        const bool v0 = Fv[rc - 1][tx], v1 = Fv[rc + 1][tx];
        const double w0 = v0 ? 0.25 : 0.0, w1 = v1 ? 0.25 : 0.0;
        const double wc = 1.0 - w0 - w1;
        const int n0 = v0 ? rc - 1 : rc, n1 = v1 ? rc + 1 : rc;
#pragma unroll
        for(int x = 0; x < NV; x++) Pv[pr][x][tx] = wc * Cv[rc][x][tx] + w0 * Cv[n0][x][tx] + w1 * Cv[n1][x][tx];
    }
    __syncthreads();
    // the second pass over rows rc = 2 .. kSmJ + 1 (latitudes j0 .. j0 + kSmJ - 1), written out
    const int j = j0 + ty;
    if(j >= D.jm || k >= D.km) return;
    const long long n = mo + D.at(i, j, k);
    const int rc = ty + 2, pr = rc - 1;
    B.mca_sf[n] = Fv[rc][tx];
    if(!Fv[rc][tx]) return;
    const bool v0 = Fv[rc - 1][tx], v1 = Fv[rc + 1][tx];
    const double w0 = v0 ? 0.25 : 0.0, w1 = v1 ? 0.25 : 0.0;
    const double wc = 1.0 - w0 - w1;
    const int n0 = v0 ? pr - 1 : pr, n1 = v1 ? pr + 1 : pr;
#pragma unroll
    for(int x = 0; x < NV; x++)
        B.mca[(M_SH + x) * FS + n] = wc * Pv[pr][x][tx] + w0 * Pv[n0][x][tx] + w1 * Pv[n1][x][tx];
}

// ---- the columns -----------------------------------------------------------------------------
// a column's place: one thread per (k, j) of a member, workgroups of 64 longitudes x 4 latitudes
constexpr int kColK = 64, kColJ = 4;

struct McaCol {
    int j, k, m;
    long long c0, sr, FS, ci;                 // cell i is c0 + i * sr; ci: the column's census slot
};

__device__ __forceinline__ bool mca_col(const Dims& D, McaCol& C){
    C.k = blockIdx.x * kColK + threadIdx.x;
    C.j = blockIdx.y * kColJ + threadIdx.y;
    C.m = blockIdx.z;
    if(C.k >= D.km || C.j >= D.jm) return false;
    C.sr = (long long)D.jm * D.kp;
    C.FS = (long long)D.M * D.fs;
    C.c0 = (long long)C.m * D.fs + D.at(0, C.j, C.k);
    C.ci = ((long long)C.m * D.jm + C.j) * D.kp + C.k;
    return true;
}

// The census slots start at zero (k_mca_cen_clear); k_mca_census then runs one thread per adjacent
// pair of levels of every column and sets the pair's bit where it is unstable, and the column's
// bit 63 where one of its levels holds no valid state (:573-603). Pairs are independent (each
// compares the parcel of level i lifted to i + 1 with the equilibrium of i + 1), so the census
// takes a thread per pair rather than a thread rolling up the column: a pair's tests depend only
// on its two levels, the waves of one launch are 38 times as many as the columns, and no wave
// waits for the longest of 38 dependent tests of a column. Bits are set by atomic OR; columns have
// at most 64 levels, so pairs use bits 1..61 and bit 63 is free. k_mca_mix reads a slot with bit
// 63 set as an empty census, which is what the reference's loop gives such a column.
constexpr unsigned long long kCenInvalid = 1ull << 63;

__global__ void __launch_bounds__(256) k_mca_cen_clear(unsigned long long* cen, long long n){
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(x < n) cen[x] = 0ull;
}

template<typename T>
__global__ void __launch_bounds__(kColK * kColJ) k_mca_census(Dims D, const McaK* __restrict__ Kp, Bufs<T> B, unsigned long long* cen){
    const McaK& K = *Kp;
    const int im = D.im, i0 = 1, i1 = im - 2, npair = i1 - i0;
    const int k = blockIdx.x * kColK + threadIdx.x;
    const int j = blockIdx.y * kColJ + threadIdx.y;
    const int m = blockIdx.z / npair, i = i0 + (int)(blockIdx.z % npair);
    if(k >= D.km || j >= D.jm) return;
    const long long sr = (long long)D.jm * D.kp, FS = (long long)D.M * D.fs;
    const long long c0 = (long long)m * D.fs + D.at(0, j, k);
    unsigned long long* slot = cen + ((long long)m * D.jm + j) * D.kp + k;
    const double* SC = B.mca;
    const double* tf = B.s[ATSAT_T] + c0;
    const double* pf = B.s[ATSAT_P_STAT] + c0;
    const double* rf = B.s[ATSAT_RHO_MIX] + c0;
    // a level's state must be valid (:573-603); the pair's thread reports its lower level, the
    // last pair's thread also the top one
    auto level_ok = [&](int l){
        const long long o = (long long)l * sr;
        const double Tl = tf[o] * K.t_ref, pl = pf[o], rl = rho_of(rf[o], K.r_mix);
        bool ok = (Tl > 0.0) && (pl > 0.0) && (rl > 0.0) && isfinite(Tl + pl + rl);
        const double lo = (l > i0) ? 0.5 * (z_of(K, l - 1) + z_of(K, l)) : z_of(K, i0);
        const double hi = (l < i1) ? 0.5 * (z_of(K, l) + z_of(K, l + 1)) : z_of(K, i1);
        if(!(rl * (hi - lo) > 0.0) || !B.mca_sf[c0 + o]) ok = false;
        return ok;
    };
    const bool ok_a = level_ok(i), ok_b = level_ok(i + 1);
    if(!ok_a || (i + 1 == i1 && !ok_b)) atomicOr(slot, kCenInvalid);
    if(!ok_a || !ok_b) return;
    auto load = [&](int l, double& h, double* q, double* r){
        const long long o = c0 + (long long)l * sr;
        h = SC[M_SH * FS + o];
#pragma unroll
        for(int g = 0; g < NG; g++) q[g] = SC[(M_SQ + g) * FS + o];
#pragma unroll
        for(int t = 0; t < NT; t++) r[t] = SC[(M_SR + t) * FS + o];
    };
    // the parcel of level i lifted to i + 1 against the equilibrium of i + 1 (:645-662); the two
    // equilibria advance one saturation evaluation at a time, the one with the wider bounds
    // first, until the comparison is settled
    double ha, qa[NG], ra[NT], hb, qb[NG], rb[NT];
    load(i, ha, qa, ra);
    load(i + 1, hb, qb, rb);
    const double pb = pf[(long long)(i + 1) * sr], zb = z_of(K, i + 1), rpb = 1.0 / pb;
    EqRun P, E;
    eq_begin(K, P, ha + tr_energy(K, ra), qa, pb, zb);
    eq_begin(K, E, hb + tr_energy(K, rb), qb, pb, zb);
    bool uns = false;
    if(P.stage != 4 && E.stage != 4){
        double zero[NG] = { 0.0, 0.0, 0.0 };
        const double f0P = tv(K, 1.0, qa, zero, ra), fminP = tv(K, 1.0, zero, qa, ra);
        const double f0E = tv(K, 1.0, qb, zero, rb), fminE = tv(K, 1.0, zero, qb, rb);
#pragma unroll 1
        for(;;){
            double pl, ph, el, eh;
            eq_bounds(K, P, fminP, f0P, pl, ph);
            eq_bounds(K, E, fminE, f0E, el, eh);
            if(ph - el <= kTolK - kDecide){ uns = false; break; }
            if(pl - eh > kTolK + kDecide){ uns = true; break; }
            if(P.stage == 3 && E.stage == 3){ uns = pl - eh > kTolK; break; }
            const bool advP = P.stage != 3 && (E.stage == 3 || ph - pl >= eh - el);
            double q[NG], qs[NG];
#pragma unroll
            for(int g = 0; g < NG; g++) q[g] = advP ? qa[g] : qb[g];
            qsat_all<T, true>(K, advP ? P.x : E.x, pb, q, qs, rpb);
            if(advP) eq_advance(K, P, qa, ra, qs);
            else     eq_advance(K, E, qb, rb, qs);
        }
    }
    if(uns) atomicOr(slot, 1ull << i);
}

// the columns with an unstable pair: the sweeps, the union re-mixed in the column's own state and
// the write back
template<typename T>
__global__ void __launch_bounds__(kColK * kColJ) k_mca_mix(Dims D, const McaK* __restrict__ Kp, Bufs<T> B, const unsigned long long* cen){
    const McaK& K = *Kp;
    McaCol L;
    if(!mca_col(D, L)) return;
    const unsigned long long cen_s = cen[L.ci];
    const unsigned long long cen_u = (cen_s & kCenInvalid) ? 0ull : cen_s;
    if(!cen_u) return;
    const int im = D.im, i0 = 1, i1 = im - 2;
    const long long sr = L.sr, FS = L.FS, c0 = L.c0;
    const int max_pass = 64;
    double* SC = B.mca;
    const double* tf = B.s[ATSAT_T] + c0;
    const double* pf = B.s[ATSAT_P_STAT] + c0;
    const double* rf = B.s[ATSAT_RHO_MIX] + c0;
    auto rho = [&](int i){ return rho_of(rf[(long long)i * sr], K.r_mix); };
    auto w_of = [&](int i){
        const double lo = (i > i0) ? 0.5 * (z_of(K, i - 1) + z_of(K, i)) : z_of(K, i0);
        const double hi = (i < i1) ? 0.5 * (z_of(K, i) + z_of(K, i + 1)) : z_of(K, i1);
        return rho(i) * (hi - lo);
    };

    double* hL = SC + M_SH * FS + c0;
    double* qt = SC + M_SQ * FS + c0;          // gas g: qt[g * FS + i * sr]
    double* qr = SC + M_SR * FS + c0;
    double* Teq = SC + M_TEQ * FS + c0;
    double* ceq = SC + M_CEQ * FS + c0;
    auto Q = [&](int i, int g) -> double& { return qt[g * FS + i * sr]; };
    auto R_ = [&](int i, int r) -> double& { return qr[r * FS + i * sr]; };
    auto C_ = [&](int i, int g) -> double& { return ceq[g * FS + i * sr]; };
    unsigned long long eq_ok = 0, touched = 0;
    auto bit = [](int i){ return 1ull << i; };
    // :645-662, b's equilibrium kept once formed (the census kept none: forming one again from
    // the same state gives the same values)
    auto unstable = [&](int a, int b) -> bool {
        double qa[NG], ra[NT], qb[NG], rb[NT];
#pragma unroll
        for(int g = 0; g < NG; g++){ qa[g] = Q(a, g); qb[g] = Q(b, g); }
#pragma unroll
        for(int t = 0; t < NT; t++){ ra[t] = R_(a, t); rb[t] = R_(b, t); }
        const double pb = pf[(long long)b * sr], zb = z_of(K, b);
        double Tp, pc[NG], cb[NG];
        if(!equil<T>(K, hL[(long long)a * sr] + tr_energy(K, ra), qa, pb, zb, Tp, pc)) return false;
        if(!(eq_ok & bit(b))){
            double Tb;
            if(!equil<T>(K, hL[(long long)b * sr] + tr_energy(K, rb), qb, pb, zb, Tb, cb)) return false;
            Teq[(long long)b * sr] = Tb;
#pragma unroll
            for(int g = 0; g < NG; g++) C_(b, g) = cb[g];
            eq_ok |= bit(b);
        }
        const double Tb = Teq[(long long)b * sr];
        double pv[NG], ev[NG];
#pragma unroll
        for(int g = 0; g < NG; g++){ cb[g] = C_(b, g); pv[g] = qa[g] - pc[g]; ev[g] = qb[g] - cb[g]; }
        return tv(K, Tp, pv, pc, ra) - tv(K, Tb, ev, cb, rb) > 1.0e-7;
    };
    // :663-677
    auto settle = [&](int q) -> bool {
        double TT, pc[NG], qq[NG], rq[NT];
#pragma unroll
        for(int g = 0; g < NG; g++) qq[g] = Q(q, g);
#pragma unroll
        for(int t = 0; t < NT; t++) rq[t] = R_(q, t);
        if(!equil<T>(K, hL[(long long)q * sr] + tr_energy(K, rq), qq, pf[(long long)q * sr], z_of(K, q), TT, pc)) return false;
        Teq[(long long)q * sr] = TT;
#pragma unroll
        for(int g = 0; g < NG; g++) C_(q, g) = pc[g];
        eq_ok |= bit(q); touched |= bit(q);
        return true;
    };
    // the segment a..b made uniform in h_L, q_t and the tracers (mass-weighted), every cell settled
    auto mix = [&](int a, int b) -> bool {
        double sw = 0.0, sh = 0.0, sq[NG] = { 0.0, 0.0, 0.0 }, st_[NT] = { 0.0, 0.0 };
        for(int q = a; q <= b; q++){
            const double wq = w_of(q);
            sw += wq; sh += wq * hL[(long long)q * sr];
#pragma unroll
            for(int g = 0; g < NG; g++) sq[g] += wq * Q(q, g);
#pragma unroll
            for(int r = 0; r < NT; r++) st_[r] += wq * R_(q, r);
        }
        for(int q = a; q <= b; q++){
            hL[(long long)q * sr] = sh / sw;
#pragma unroll
            for(int g = 0; g < NG; g++) Q(q, g) = sq[g] / sw;
#pragma unroll
            for(int r = 0; r < NT; r++) R_(q, r) = st_[r] / sw;
            if(!settle(q)) return false;
        }
        return true;
    };

    // the sweeps (:714-747) on the smoothed column
    unsigned long long link = 0;
    int passes = 0, layers = 0;
    bool changed = true, fail = false;
    while(changed && passes < max_pass && !fail){
        changed = false; passes++;
        int i = i0;
        while(i < i1 && !fail){
            const bool uns = (passes == 1 && !(touched & bit(i)) && !(touched & bit(i + 1)))
                           ? (cen_u & bit(i)) != 0 : unstable(i, i + 1);
            if(!uns){ i++; continue; }
            int a = i, b = i + 1;
            for(;;){
                if(!mix(a, b)){ fail = true; break; }
                bool ext = false;
                if(a > i0 && unstable(a - 1, a)){ a--; ext = true; }
                if(b < i1 && unstable(b, b + 1)){ b++; ext = true; }
                if(!ext) break;
            }
            if(fail) break;
            layers += b - a + 1;
            for(int q = a; q < b; q++) link |= bit(q);
            changed = true;
            i = b;
        }
    }
    if(fail || layers == 0) return;

    // the union of the mixed segments, mixed again in the column's own state (:751-783)
    for(int i = i0; i <= i1; i++){
        const long long o = (long long)i * sr;
        const double Ti = tf[o] * K.t_ref, ri = rho(i);
        double h = K.cp * Ti + K.grav * z_of(K, i);
#pragma unroll
        for(int g = 0; g < NG; g++){
            const double qv = smax(0.0, (double)B.w[K.fv[g]][c0 + o]) / ri;
            const double qc = smax(0.0, (double)B.w[K.fv[g] + 1][c0 + o]) / ri;
            const double qi = smax(0.0, (double)B.w[K.fv[g] + 2][c0 + o]) / ri;
            Q(i, g) = qv + qc + qi;
            h -= K.lv[g] * qc + K.ls[g] * qi;
        }
#pragma unroll
        for(int r = 0; r < NT; r++){
            R_(i, r) = (double)B.w[K.ft[r]][c0 + o] / ri;
            h -= K.tL[r] * R_(i, r);
        }
        hL[o] = h;
    }
    touched = 0; eq_ok = 0;
    for(int a = i0; a < i1 && !fail; ){
        if(!(link & bit(a))){ a++; continue; }
        int b = a;
        while(b < i1 && (link & bit(b))) b++;
        if(!mix(a, b)) fail = true;
        a = b + 1;
    }
    if(fail) return;

    // write back (:811-826)
    for(int q = i0; q <= i1; q++){
        if(!(touched & bit(q))) continue;
        const long long o = (long long)q * sr;
        const double TT = Teq[o], rq = rho(q);
        B.s[ATSAT_T][c0 + o] = TT / K.t_ref;
#pragma unroll
        for(int g = 0; g < NG; g++){
            const double pcg = C_(q, g);
            const double qv = Q(q, g) - pcg;
            double qc, qi;
            if(TT >= K.t0[g]){ qc = pcg; qi = 0.0; } else { qi = pcg; qc = 0.0; }
            B.w[K.fv[g]][c0 + o] = (T)(qv * rq);
            B.w[K.fv[g] + 1][c0 + o] = (T)(qc * rq);
            B.w[K.fv[g] + 2][c0 + o] = (T)(qi * rq);
        }
#pragma unroll
        for(int r = 0; r < NT; r++) B.w[K.ft[r]][c0 + o] = (T)(R_(q, r) * rq);
    }
}

// the scratch fields, one 8-byte census slot per column ([M][jm][kp]), then the constants (McaK)
static size_t mca_census_offset(const Dims& D){ return (size_t)M_NF * D.M * D.fs; }
static size_t mca_constants_offset(const Dims& D){ return mca_census_offset(D) + (size_t)D.M * D.jm * D.kp; }
size_t mca_scratch_doubles(const Dims& D){ return mca_constants_offset(D) + (sizeof(McaK) + 7) / 8; }

// the constants of the gases (H2O, NH3, CH4) and the tracers (H2S, NH4SH) as run() receives them
// (cSaturnModel.cpp:1123-1141)
static McaK mca_constants(const Scal& S){
    McaK K{};
    const int gq[NG] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };
    const int gv[NG] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
    for(int g = 0; g < NG; g++){
        const int q = gq[g];
        K.fv[g] = gv[g];
        K.ep[g] = S.ep[q]; K.lv[g] = S.lv[q]; K.ls[g] = S.ls[q]; K.t0[g] = S.t_0[q];
        K.span_l[g] = std::max(S.lv[q], S.ls[q]);
        K.tv_w[g] = 1.0 / S.ep[q] - 1.0;
        K.C[0][g] = S.C[q];   K.nL0[0][g] = -S.L0[q];   K.da[0][g] = S.del_alf[q];   K.db[0][g] = S.del_bet[q];
        K.C[1][g] = S.C_i[q]; K.nL0[1][g] = -S.L0_i[q]; K.da[1][g] = S.del_alf_i[q]; K.db[1][g] = S.del_bet_i[q];
        K.R3[0][g] = K.R3[1][g] = 1e-3 * S.R[q];
        K.rR3[0][g] = K.rR3[1][g] = 1.0 / K.R3[0][g];
    }
    const double tep[NT] = { S.ep_h2s, 0.0 };
    const bool tcond[NT] = { false, true };
    K.ft[0] = ATSAT_H2S;   K.tL[0] = 0.0;
    K.ft[1] = ATSAT_NH4SH; K.tL[1] = S.nh4sh_Lf;
    for(int t = 0; t < NT; t++){
        K.tkind[t] = tep[t] > 0.0 ? 1 : (tcond[t] ? 2 : 0);
        K.tv_t[t] = tep[t] > 0.0 ? 1.0 / tep[t] - 1.0 : 0.0;
    }
    K.cp = S.cp_mix; K.grav = S.g; K.t_ref = S.t_ref; K.r_mix = S.r_mix; K.layer_h = S.layer_h;
    K.early = 1;
    for(int g = 0; g < NG; g++){
        if(!(K.ep[g] > 0.0) || !(K.R3[0][g] > 0.0)) K.early = 0;
        // the exponents of the two phases at t_0 (monotone in the pressure), the ice one at most a
        // relative 1e-12 above the liquid one
        const double t0 = K.t0[g], lt = std::log(t0);
        const double xl = K.C[0][g] + (K.nL0[0][g] / t0 + K.da[0][g] * lt + K.db[0][g] * t0) / K.R3[0][g];
        const double xi = K.C[1][g] + (K.nL0[1][g] / t0 + K.da[1][g] * lt + K.db[1][g] * t0) / K.R3[1][g];
        K.jump_ok[g] = (t0 > 0.0 && xl >= xi - 1.0e-12 * std::fabs(xi)) ? 1 : 0;
    }
    if(!(S.ep_h2s > 0.0)) K.early = 0;
    return K;
}

// The constants into the scratch buffer `mca` (device memory of mca_scratch_doubles(D) doubles),
// once per context: they depend on the constants and the layer height only. Synchronous.
int mca_store_constants(const Dims& D, const Scal& S, double* mca){
    const McaK K = mca_constants(S);
    return hipMemcpy(mca + mca_constants_offset(D), &K, sizeof(K), hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

// The kernels read the constants through a pointer: passed by value they are held in scalar
// registers for the whole kernel, more than there are.
template<typename T>
void launch_mca(const Dims& D, const Scal&, const Bufs<T>& B, hipStream_t st){
    const McaK* K = reinterpret_cast<const McaK*>(B.mca + mca_constants_offset(D));
    const long long ntile = (long long)((D.km + kSmK - 1) / kSmK) * ((D.jm + kSmJ - 1) / kSmJ) * D.im * D.M;
    hipLaunchKernelGGL(k_mca_smooth<T>, dim3((unsigned)((ntile + 7) / 8 * 8)), dim3(kSmK, kSmJ), 0, st, D, K, B);
    unsigned long long* cen = reinterpret_cast<unsigned long long*>(B.mca + mca_census_offset(D));
    const dim3 cols((D.km + kColK - 1) / kColK, (D.jm + kColJ - 1) / kColJ, D.M), cb(kColK, kColJ);
    const long long nslot = (long long)D.M * D.jm * D.kp;
    hipLaunchKernelGGL(k_mca_cen_clear, dim3((unsigned)((nslot + 255) / 256)), dim3(256), 0, st, cen, nslot);
    // the rows of workgroups padded so that a level plane of a member is a multiple of 8
    // workgroups: the same column's workgroup of the next pair runs on the same XCD, where the
    // level the two pairs share was just read into L2
    const int npair = D.im - 3;
    unsigned cy = cols.y;
    while((cols.x * cy) % 8u) cy++;
    if(npair > 0)
        hipLaunchKernelGGL(k_mca_census<T>, dim3(cols.x, cy, D.M * npair), cb, 0, st, D, K, B, cen);
    hipLaunchKernelGGL(k_mca_mix<T>, cols, cb, 0, st, D, K, B, (const unsigned long long*)cen);
}

template void launch_mca<double>(const Dims&, const Scal&, const Bufs<double>&, hipStream_t);
template void launch_mca<float>(const Dims&, const Scal&, const Bufs<float>&, hipStream_t);

}  // namespace atsat
