// The moist convective adjustment on the device: MoistConvectiveAdjustment<cSaturnModel>::run
// (planet/ConvectiveAdjustment.h:402-849), called after restoreVar and the closure every
// ATSAT_MCA iterations (cSaturnModel.cpp:1120-1145), at ATSAT's defaults: the Ledoux criterion with
// the condensate loading, no relaxation, 64 sweeps at most, the search on a column smoothed in
// latitude (MCA_SMOOTH 2), the radial boundary planes left out (MCA_INTERIOR), the t_0 phase rule,
// cp = cp_mix. Gases H2O, NH3, CH4 (liquid above t_0, ice below); tracers H2S (a gas of ep_h2s in
// the virtual temperature) and NH4SH (condensed, carrying its formation heat in h_L).
//
//   k_mca_smooth_init  h_L, q_t of every gas and the tracers' mass fractions of every cell with a
//                      valid state (:504-521), one thread per cell;
//   k_mca_smooth_pass  one latitude 1-2-1 pass over them (:522-539), twice, ping-pong;
//   k_mca_columns      one thread per column (all j, all k): the parcel census, the sweeps that mix
//                      unstable segments and grow them while their edges are unstable, then the
//                      union of the segments mixed again in the column's own state, and the write
//                      back of every cell that was put to equilibrium (:553-826).
//
// Scratch (FP64, the context's MCA buffer, field-major like the arena): the smoothed variables
// (two sets for the passes), whose first set each column thread then uses as its own search
// state; every cell's equilibrium temperature and condensate. A column's bookkeeping over its
// levels -- which cells are in equilibrium, which were touched, which adjacent pairs were mixed
// together and the census -- lives in 64-bit masks, so a column holds at most 64 levels.
//
// Everything is FP64 in the reference's operation order; the saturation vapour pressure takes the
// correctly rounded exp / log of crmath.hpp.
#include "atsat_dev.hpp"
#include "crmath.hpp"

namespace atsat {

namespace {
__device__ __forceinline__ double smax(double a, double b){ return (a < b) ? b : a; }

// ATPhys.h:49-52
__device__ __forceinline__ double svp(double T_K, double C, double L0, double R, double del_alf, double del_bet){
    return crm::cr_exp(C + (-L0 / T_K + del_alf * crm::cr_log(T_K) + del_bet * T_K) / (1e-3 * R));
}

constexpr int NG = 3, NT = 2;
// scratch fields
enum { M_SH = 0, M_SQ = 1, M_SR = 4, M_SET = 6, M_TEQ = 12, M_CEQ = 13, M_NF = 16 };

struct McaC {
    // per gas (H2O, NH3, CH4): the field of the vapour (cloud and ice follow it), ep, lv, ls, t_0,
    // the liquid and the ice SVP sets
    int fv[NG];
    double ep[NG], lv[NG], ls[NG], t0[NG], C[NG], L0[NG], R[NG], da[NG], db[NG], Ci[NG], L0i[NG], dai[NG], dbi[NG];
    // tracers: H2S, NH4SH
    int ft[NT];
    double tep[NT], tL[NT];
    bool tcond[NT];
};

__device__ McaC mca_consts(const Scal& S){
    McaC K;
    const int gq[NG] = { ATSAT_GAS_H2O, ATSAT_GAS_NH3, ATSAT_GAS_CH4 };
    const int gv[NG] = { ATSAT_H2O, ATSAT_NH3, ATSAT_CH4 };
    for(int g = 0; g < NG; g++){
        const int q = gq[g];
        K.fv[g] = gv[g];
        K.ep[g] = S.ep[q]; K.lv[g] = S.lv[q]; K.ls[g] = S.ls[q]; K.t0[g] = S.t_0[q];
        K.C[g] = S.C[q]; K.L0[g] = S.L0[q]; K.R[g] = S.R[q]; K.da[g] = S.del_alf[q]; K.db[g] = S.del_bet[q];
        K.Ci[g] = S.C_i[q]; K.L0i[g] = S.L0_i[q]; K.dai[g] = S.del_alf_i[q]; K.dbi[g] = S.del_bet_i[q];
    }
    K.ft[0] = ATSAT_H2S; K.tep[0] = S.ep_h2s; K.tcond[0] = false; K.tL[0] = 0.0;
    K.ft[1] = ATSAT_NH4SH; K.tep[1] = 0.0; K.tcond[1] = true; K.tL[1] = S.nh4sh_Lf;
    return K;
}

__device__ __forceinline__ double z_of(const Scal& S, int i){          // get_layer_height * 1e3
    const float h = (float)S.layer_h;
    return (double)((float)i * h) * 1.0e3;
}
__device__ __forceinline__ double rho_of(const double* rho, long long n, double r_mix){   // rho_at
    const double r = rho[n];
    return (r > 0.0 && isfinite(r)) ? r : r_mix;
}
}  // namespace

// ---- the smoothed conserved variables ------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_mca_smooth_init(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long n = (long long)m * D.fs + D.at(i, j, k);
    const long long FS = (long long)D.M * D.fs;
    const McaC K = mca_consts(S);
    double* SC = B.mca;
    const double TT = B.s[ATSAT_T][n] * S.t_ref, rr = rho_of(B.s[ATSAT_RHO_MIX], n, S.r_mix), zz = z_of(S, i);
    if(!(TT > 0.0) || !(rr > 0.0) || !isfinite(TT + rr)){ B.mca_sf[n] = 0; return; }
    double h = S.cp_mix * TT + S.g * zz;
    for(int g = 0; g < NG; g++){
        const double v_ = smax(0.0, B.s[K.fv[g]][n]) / rr, c_ = smax(0.0, B.s[K.fv[g] + 1][n]) / rr,
                     i_ = smax(0.0, B.s[K.fv[g] + 2][n]) / rr;
        SC[(M_SQ + g) * FS + n] = v_ + c_ + i_;
        h -= K.lv[g] * c_ + K.ls[g] * i_;
    }
    for(int r = 0; r < NT; r++){
        const double q = B.s[K.ft[r]][n] / rr;
        SC[(M_SR + r) * FS + n] = q;
        h -= K.tL[r] * q;
    }
    SC[M_SH * FS + n] = h;
    B.mca_sf[n] = 1;
}

// one latitude pass, set `src` -> set `dst` (0 or M_SET), cells without a valid state skipped and
// no neighbours
template<typename T>
__global__ void __launch_bounds__(256) k_mca_smooth_pass(Dims D, Bufs<T> B, int src, int dst){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, im_m = blockIdx.z;
    if(k >= D.km) return;
    const int i = im_m % D.im, m = im_m / D.im;
    const long long mo = (long long)m * D.fs;
    const long long n = mo + D.at(i, j, k);
    const long long FS = (long long)D.M * D.fs;
    if(!B.mca_sf[n]) return;
    double wn[2] = { 0.0, 0.0 }; long long cn[2] = { n, n };
    for(int s_ = 0; s_ < 2; s_++){
        const int jj = j + (s_ ? 1 : -1);
        if(jj < 0 || jj >= D.jm) continue;
        const long long c2 = mo + D.at(i, jj, k);
        if(B.mca_sf[c2]){ wn[s_] = 0.25; cn[s_] = c2; }
    }
    const double wc = 1.0 - wn[0] - wn[1];
    const double* I = B.mca + (long long)src * FS;
    double* O = B.mca + (long long)dst * FS;
    for(int x = 0; x < 1 + NG + NT; x++){
        const double* F = I + x * FS;
        O[x * FS + n] = wc * F[n] + wn[0] * F[cn[0]] + wn[1] * F[cn[1]];
    }
}

// ---- the columns -----------------------------------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256) k_mca_columns(Dims D, Scal S, Bufs<T> B){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, m = blockIdx.z;
    if(k >= D.km) return;
    const int im = D.im, i0 = 1, i1 = im - 2;
    if(i1 - i0 < 1) return;
    const long long mo = (long long)m * D.fs;
    const long long FS = (long long)D.M * D.fs;
    const long long sr = (long long)D.jm * D.kp;
    const long long c0 = mo + D.at(0, j, k);                  // cell i is c0 + i * sr
    const McaC K = mca_consts(S);
    const double cp = S.cp_mix, grav = S.g;
    const double tol_K = 1.0e-7, tol_T = 1.0e-10;
    const int n_iter = 60, max_pass = 64;
    double* SC = B.mca;
    double* hL = SC + M_SH * FS + c0;
    double* qt = SC + M_SQ * FS + c0;          // gas g: qt[g * FS + i * sr]
    double* qr = SC + M_SR * FS + c0;
    double* Teq = SC + M_TEQ * FS + c0;
    double* ceq = SC + M_CEQ * FS + c0;
    const double* tf = B.s[ATSAT_T] + c0;
    const double* pf = B.s[ATSAT_P_STAT] + c0;
    const double* rf = B.s[ATSAT_RHO_MIX] + c0;
    auto rho = [&](int i){ return rho_of(rf, (long long)i * sr, S.r_mix); };
    auto Q = [&](int i, int g) -> double& { return qt[g * FS + i * sr]; };
    auto R_ = [&](int i, int r) -> double& { return qr[r * FS + i * sr]; };
    auto C_ = [&](int i, int g) -> double& { return ceq[g * FS + i * sr]; };

    // the column's state (:573-597)
    for(int i = i0; i <= i1; i++){
        const long long o = (long long)i * sr;
        const double Ti = tf[o] * S.t_ref, pi = pf[o], ri = rho(i);
        if(!(Ti > 0.0) || !(pi > 0.0) || !(ri > 0.0) || !isfinite(Ti + pi + ri)) return;
    }
    auto w_of = [&](int i){
        const double lo = (i > i0) ? 0.5 * (z_of(S, i - 1) + z_of(S, i)) : z_of(S, i0);
        const double hi = (i < i1) ? 0.5 * (z_of(S, i) + z_of(S, i + 1)) : z_of(S, i1);
        return rho(i) * (hi - lo);
    };
    for(int i = i0; i <= i1; i++) if(!(w_of(i) > 0.0)) return;
    for(int i = i0; i <= i1; i++) if(!B.mca_sf[c0 + (long long)i * sr]) return;

    auto qsat = [&](int g, double TT, double p){
        const double E = (TT >= K.t0[g]) ? svp(TT, K.C[g], K.L0[g], K.R[g], K.da[g], K.db[g])
                                         : svp(TT, K.Ci[g], K.L0i[g], K.R[g], K.dai[g], K.dbi[g]);
        const double q = K.ep[g] * E / p;
        return isfinite(q) ? q : 0.0;
    };
    auto lat = [&](int g, double TT){ return TT >= K.t0[g] ? K.lv[g] : K.ls[g]; };
    // :447-475
    auto equil = [&](double h, const double* q, double p, double z, double& TT, double* cond) -> bool {
        const double T_lo = (h - grav * z) / cp;
        if(!(T_lo > 0.0) || !(p > 0.0)) return false;
        bool any = false;
        for(int g = 0; g < NG; g++) if(q[g] > qsat(g, T_lo, p)){ any = true; break; }
        if(!any){ TT = T_lo; for(int g = 0; g < NG; g++) cond[g] = 0.0; return true; }
        double span = 0.0;
        for(int g = 0; g < NG; g++) span += smax(K.lv[g], K.ls[g]) * smax(0.0, q[g]);
        auto fof = [&](double X){
            double f = cp * X + grav * z - h;
            for(int g = 0; g < NG; g++) f -= lat(g, X) * smax(0.0, q[g] - qsat(g, X, p));
            return f; };
        double a = T_lo, b = T_lo + span / cp + 1.0e-9;
        double fa = fof(a), fb = fof(b);
        int side = 0;
        for(int it = 0; it < n_iter && b - a > tol_T; it++){
            double c = (fb - fa != 0.0) ? b - fb * (b - a) / (fb - fa) : 0.5 * (a + b);
            if(!(c > a && c < b)) c = 0.5 * (a + b);
            const double fc = fof(c);
            if(fc < 0.0){ a = c; fa = fc; if(side == -1) fb *= 0.5; side = -1; }
            else        { b = c; fb = fc; if(side == +1) fa *= 0.5; side = +1; }
            if(fabs(fc) <= cp * tol_T){ a = b = c; break; }
        }
        TT = b;
        for(int g = 0; g < NG; g++) cond[g] = smax(0.0, q[g] - qsat(g, TT, p));
        return true;
    };
    auto trE = [&](int q){ double e = 0.0; for(int r = 0; r < NT; r++) e += K.tL[r] * R_(q, r); return e; };
    auto tv = [&](double TT, const double* v_, const double* cnd, int q){
        double f = 1.0;
        for(int g = 0; g < NG; g++){ f += v_[g] * (1.0 / K.ep[g] - 1.0); f -= cnd[g]; }
        for(int r = 0; r < NT; r++){
            const double x = R_(q, r);
            if(K.tep[r] > 0.0) f += x * (1.0 / K.tep[r] - 1.0);
            else if(K.tcond[r]) f -= x;
        }
        return TT * f;
    };
    unsigned long long eq_ok = 0, touched = 0;
    auto bit = [](int i){ return 1ull << i; };
    // :645-662
    auto unstable = [&](int a, int b) -> bool {
        double Tp, pc[NG], qa[NG];
        for(int g = 0; g < NG; g++) qa[g] = Q(a, g);
        const double pb = pf[(long long)b * sr], zb = z_of(S, b);
        if(!equil(hL[(long long)a * sr] + trE(a), qa, pb, zb, Tp, pc)) return false;
        if(!(eq_ok & bit(b))){
            double qb[NG], cb[NG], Tb;
            for(int g = 0; g < NG; g++) qb[g] = Q(b, g);
            if(!equil(hL[(long long)b * sr] + trE(b), qb, pb, zb, Tb, cb)) return false;
            Teq[(long long)b * sr] = Tb;
            for(int g = 0; g < NG; g++) C_(b, g) = cb[g];
            eq_ok |= bit(b);
        }
        const double Tb = Teq[(long long)b * sr];
        double ec[NG], pv[NG], ev[NG];
        for(int g = 0; g < NG; g++){ ec[g] = C_(b, g); pv[g] = Q(a, g) - pc[g]; ev[g] = Q(b, g) - ec[g]; }
        const double dTv = tv(Tp, pv, pc, a) - tv(Tb, ev, ec, b);
        return dTv > tol_K;
    };
    // :663-677
    auto settle = [&](int q) -> bool {
        double TT, pc[NG], qq[NG];
        for(int g = 0; g < NG; g++) qq[g] = Q(q, g);
        if(!equil(hL[(long long)q * sr] + trE(q), qq, pf[(long long)q * sr], z_of(S, q), TT, pc)) return false;
        Teq[(long long)q * sr] = TT;
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
            for(int g = 0; g < NG; g++) sq[g] += wq * Q(q, g);
            for(int r = 0; r < NT; r++) st_[r] += wq * R_(q, r);
        }
        for(int q = a; q <= b; q++){
            hL[(long long)q * sr] = sh / sw;
            for(int g = 0; g < NG; g++) Q(q, g) = sq[g] / sw;
            for(int r = 0; r < NT; r++) R_(q, r) = st_[r] / sw;
            if(!settle(q)) return false;
        }
        return true;
    };

    // the census (:695-703) and the sweeps (:714-747) on the smoothed column
    unsigned long long cen_u = 0, link = 0;
    for(int i = i0; i < i1; i++) if(unstable(i, i + 1)) cen_u |= bit(i);
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
        const double Ti = tf[o] * S.t_ref, ri = rho(i);
        double h = cp * Ti + grav * z_of(S, i);
        for(int g = 0; g < NG; g++){
            const double qv = smax(0.0, B.s[K.fv[g]][c0 + o]) / ri;
            const double qc = smax(0.0, B.s[K.fv[g] + 1][c0 + o]) / ri;
            const double qi = smax(0.0, B.s[K.fv[g] + 2][c0 + o]) / ri;
            Q(i, g) = qv + qc + qi;
            h -= K.lv[g] * qc + K.ls[g] * qi;
        }
        for(int r = 0; r < NT; r++){
            R_(i, r) = B.s[K.ft[r]][c0 + o] / ri;
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
        B.s[ATSAT_T][c0 + o] = TT / S.t_ref;
        for(int g = 0; g < NG; g++){
            const double pcg = C_(q, g);
            const double qv = Q(q, g) - pcg;
            double qc, qi;
            if(TT >= K.t0[g]){ qc = pcg; qi = 0.0; } else { qi = pcg; qc = 0.0; }
            B.s[K.fv[g]][c0 + o] = qv * rq;
            B.s[K.fv[g] + 1][c0 + o] = qc * rq;
            B.s[K.fv[g] + 2][c0 + o] = qi * rq;
        }
        for(int r = 0; r < NT; r++) B.s[K.ft[r]][c0 + o] = R_(q, r) * rq;
    }
}

size_t mca_scratch_doubles(const Dims& D){ return (size_t)M_NF * D.M * D.fs; }

template<typename T>
void launch_mca(const Dims& D, const Scal& S, const Bufs<T>& B, hipStream_t st){
    const dim3 cells((D.km + 255) / 256, D.jm, D.im * D.M);
    hipLaunchKernelGGL(k_mca_smooth_init<T>, cells, dim3(256), 0, st, D, S, B);
    hipLaunchKernelGGL(k_mca_smooth_pass<T>, cells, dim3(256), 0, st, D, B, 0, (int)M_SET);
    hipLaunchKernelGGL(k_mca_smooth_pass<T>, cells, dim3(256), 0, st, D, B, (int)M_SET, 0);
    hipLaunchKernelGGL(k_mca_columns<T>, dim3((D.km + 255) / 256, D.jm, D.M), dim3(256), 0, st, D, S, B);
}

template void launch_mca<double>(const Dims&, const Scal&, const Bufs<double>&, hipStream_t);
template void launch_mca<float>(const Dims&, const Scal&, const Bufs<float>&, hipStream_t);

}  // namespace atsat
