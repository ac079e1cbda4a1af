// exp, log and pow, correctly rounded in all but a vanishing fraction of cases, by Ziv's method:
// a fast double-double evaluation with a proven error bound, rounded when the bound settles
// the rounding, otherwise a slower evaluation with error below 2^-95 relative, rounded once.
// The host library the reference links (glibc 2.35) documents its exp, log and pow as accurate
// to 0.51-0.52 ulp, i.e. also correctly rounded outside a narrow band near the rounding
// boundaries; the two therefore agree in all but a small fraction of cases, where the device
// library's own exp / log / pow (about 1 ulp) differ from glibc in many.
//
// Fast exp: x = (128 e + j) ln2/128 + r, ln2 split in three parts (Cody-Waite), |r| <= ln2/256;
// exp(x) = 2^e * T[j] * exp(r) with T[j] = 2^(j/128) as a double-double table and exp(r) =
// 1 + r + r^2 P(r), P of degree 4 in double; relative error below 2^-67.
// Slow exp: the same reduction without the table, r scaled by 2^-8, a degree-9 Taylor
// polynomial in double-double and eight squarings.
// log by one Newton step log x = y0 + log(x exp(-y0)) from the library value y0, with the fast
// (then the slow) exp; pow as exp(y log x). Non-finite, non-positive (for log / pow) and near
// overflow / underflow arguments go to the library functions, whose special-case results are
// the IEEE ones.
//
// Compiles for the device (HIP) and for the host (plain C++), so that it can be checked against
// the host libm and against multi-precision values on the CPU.
#pragma once

#include <cmath>

#ifdef __HIP__
#define CR_FN __device__ __forceinline__
#define CR_TAB static __constant__ const
#else
#define CR_FN inline
#define CR_TAB static const
#endif

namespace atsat {
namespace crm {

struct dd { double h, l; };

CR_FN dd two_sum(double a, double b){
    const double s = a + b, bb = s - a;
    return { s, (a - (s - bb)) + (b - bb) };
}
CR_FN dd fast_two_sum(double a, double b){               // |a| >= |b|
    const double s = a + b;
    return { s, b - (s - a) };
}
CR_FN dd two_prod(double a, double b){
    const double p = a * b;
    return { p, std::fma(a, b, -p) };
}
CR_FN dd dd_add(dd a, dd b){
    dd s = two_sum(a.h, b.h);
    s.l += a.l + b.l;
    return fast_two_sum(s.h, s.l);
}
CR_FN dd dd_mul(dd a, dd b){
    dd p = two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return fast_two_sum(p.h, p.l);
}
CR_FN dd dd_mul_d(dd a, double b){
    dd p = two_prod(a.h, b);
    p.l += a.l * b;
    return fast_two_sum(p.h, p.l);
}

// ln 2 = L1 + L2 + L3, L1 with 32 significant bits (so that k * L1 is exact for |k| < 2^20)
constexpr double kLn2_1 = 0x1.62e42fee00000p-1;
constexpr double kLn2_2 = 0x1.a39ef35793c76p-33;
constexpr double kLn2_3 = 0x1.cc01f97b57a08p-87;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;

// relative error bound of dd_exp_tab, absolute error bound of the fast log (both with margin)
constexpr double kExpErr = 0x1p-66;
constexpr double kLogErr = 0x1p-65;

// 2^(j/128), j = 0..127: the double nearest, and the double nearest the remainder
CR_TAB dd kExp2Tab[128] = {
    { 0x1.0000000000000p+0, 0x0.0p+0 }, { 0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54 },
    { 0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56 }, { 0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54 },
    { 0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55 }, { 0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55 },
    { 0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57 }, { 0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54 },
    { 0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54 }, { 0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54 },
    { 0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59 }, { 0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57 },
    { 0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54 }, { 0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58 },
    { 0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54 }, { 0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55 },
    { 0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55 }, { 0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54 },
    { 0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55 }, { 0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54 },
    { 0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54 }, { 0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54 },
    { 0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55 }, { 0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55 },
    { 0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54 }, { 0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55 },
    { 0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55 }, { 0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54 },
    { 0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55 }, { 0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59 },
    { 0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54 }, { 0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56 },
    { 0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55 }, { 0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55 },
    { 0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54 }, { 0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55 },
    { 0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54 }, { 0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54 },
    { 0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56 }, { 0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54 },
    { 0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55 }, { 0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54 },
    { 0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58 }, { 0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55 },
    { 0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59 }, { 0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54 },
    { 0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56 }, { 0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54 },
    { 0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56 }, { 0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54 },
    { 0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54 }, { 0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55 },
    { 0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55 }, { 0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55 },
    { 0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54 }, { 0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55 },
    { 0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54 }, { 0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60 },
    { 0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54 }, { 0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54 },
    { 0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54 }, { 0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55 },
    { 0x1.6623882552225p+0, -0x1.bb60987591c34p-54 }, { 0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57 },
    { 0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54 }, { 0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56 },
    { 0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57 }, { 0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55 },
    { 0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55 }, { 0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54 },
    { 0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54 }, { 0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54 },
    { 0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55 }, { 0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54 },
    { 0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56 }, { 0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54 },
    { 0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54 }, { 0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55 },
    { 0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54 }, { 0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54 },
    { 0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54 }, { 0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56 },
    { 0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55 }, { 0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54 },
    { 0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57 }, { 0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56 },
    { 0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54 }, { 0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54 },
    { 0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56 }, { 0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56 },
    { 0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54 }, { 0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54 },
    { 0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54 }, { 0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54 },
    { 0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54 }, { 0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55 },
    { 0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54 }, { 0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54 },
    { 0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57 }, { 0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54 },
    { 0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56 }, { 0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54 },
    { 0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55 }, { 0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54 },
    { 0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55 }, { 0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54 },
    { 0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54 }, { 0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56 },
    { 0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56 }, { 0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55 },
    { 0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54 }, { 0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54 },
    { 0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55 }, { 0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54 },
    { 0x1.da9e603db3285p+0, 0x1.c2300696db532p-54 }, { 0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54 },
    { 0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54 }, { 0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54 },
    { 0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55 }, { 0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54 },
    { 0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54 }, { 0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54 },
    { 0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54 }, { 0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54 },
    { 0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54 }, { 0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54 },
    { 0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55 }, { 0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57 },
};

// x - kd * ln2 / 2^s as a double-double, |kd| < 2^20
CR_FN dd reduce(dd x, double kd, int s){
    const double t1 = std::fma(-kd, std::ldexp(kLn2_1, -s), x.h);       // exact
    const dd p2 = two_prod(kd, std::ldexp(kLn2_2, -s));
    dd r = two_sum(t1, -p2.h);
    r.l += (x.l - p2.l) - kd * std::ldexp(kLn2_3, -s);
    return fast_two_sum(r.h, r.l);
}

// exp(x.h + x.l) = 2^e * (returned dd), relative error below kExpErr, |x.h| <= 708.
CR_FN dd dd_exp_tab(dd x, int& e){
    const double kd = std::nearbyint(x.h * (128.0 * kInvLn2));
    const int k = (int)kd, j = k & 127;
    e = (k - j) / 128;
    const dd r = reduce(x, kd, 7);
    const double q = r.h * r.h * (0.5 + r.h * (1.0 / 6.0 + r.h * (1.0 / 24.0 + r.h * (1.0 / 120.0 + r.h * (1.0 / 720.0)))));
    dd s = fast_two_sum(1.0, r.h);
    s.l += r.l + q;
    s = fast_two_sum(s.h, s.l);
    return dd_mul(kExp2Tab[j], s);
}

// exp(x.h + x.l) = 2^k * (returned dd), relative error below 2^-95, |x.h| <= 708.
CR_FN dd dd_exp_sq(dd x, int& k){
    const double kd = std::nearbyint(x.h * kInvLn2);
    k = (int)kd;
    dd r = reduce(x, kd, 0);
    r.h = std::ldexp(r.h, -8);
    r.l = std::ldexp(r.l, -8);
    // 1/n!, n = 0..9, as double-double
    const dd c[10] = {
        { 0x1.0000000000000p+0, 0.0 }, { 0x1.0000000000000p+0, 0.0 }, { 0x1.0000000000000p-1, 0.0 },
        { 0x1.5555555555555p-3, 0x1.5555555555555p-57 }, { 0x1.5555555555555p-5, 0x1.5555555555555p-59 },
        { 0x1.1111111111111p-7, 0x1.1111111111111p-63 }, { 0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65 },
        { 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73 }, { 0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76 },
        { 0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73 } };
    dd p = c[9];
    for(int n = 8; n >= 0; n--) p = dd_add(dd_mul(p, r), c[n]);
    for(int q = 0; q < 8; q++) p = dd_mul(p, p);
    return p;
}

// The double nearest 2^e * (r.h + r.l) if every value within err of r.h + r.l rounds to it,
// else NaN. 2^e * r.h normal.
CR_FN double settle(dd r, double err, int e){
    const double a = r.h + (r.l + err), b = r.h + (r.l - err);
    return a == b ? std::ldexp(a, e) : NAN;
}

// log x = y0 + log(x * exp(-y0)), where exp(-y0) = 2^k * (ex.h + ex.l) and y0 is within a few
// ulp of log x: t = x exp(-y0) - 1 is then tiny and log(1 + t) = t - t^2 / 2 to 2^-120.
CR_FN dd log_newton(double x, double y0, dd ex, int k){
    const double xs = std::ldexp(x, k);
    dd p = two_prod(xs, ex.h);
    p.l += xs * ex.l;
    const dd t = two_sum(p.h - 1.0, p.l);                     // p.h - 1 is exact (p.h near 1)
    dd s = two_sum(y0, t.h);
    s.l += t.l - 0.5 * t.h * t.h;
    return fast_two_sum(s.h, s.l);
}

CR_FN double cr_exp(double x){
    if(!(x > -708.0 && x < 709.0)) return std::exp(x);       // NaN, inf, overflow / subnormal range
    int e;
    const dd f = dd_exp_tab(dd{ x, 0.0 }, e);
    const double v = settle(f, kExpErr * f.h, e);
    if(v == v) return v;
    const dd s = dd_exp_sq(dd{ x, 0.0 }, e);
    return std::ldexp(s.h + s.l, e);
}

CR_FN double cr_log(double x){
    if(!(x > 0.0) || !std::isfinite(x) || x < 0x1p-1000 || x > 0x1p1000) return std::log(x);
    const double y0 = std::log(x);
    int k;
    const dd ef = dd_exp_tab(dd{ -y0, 0.0 }, k);
    const dd f = log_newton(x, y0, ef, k);
    const double v = settle(f, kLogErr, 0);
    if(v == v) return v;
    const dd es = dd_exp_sq(dd{ -y0, 0.0 }, k);
    const dd s = log_newton(x, y0, es, k);
    return s.h + s.l;
}

CR_FN double cr_pow(double x, double y){
    if(!(x > 0.0) || !std::isfinite(x) || !std::isfinite(y) || x < 0x1p-1000 || x > 0x1p1000)
        return std::pow(x, y);
    if(y == 0.0 || x == 1.0) return 1.0;
    const double y0 = std::log(x);
    int k, e;
    const dd ef = dd_exp_tab(dd{ -y0, 0.0 }, k);
    const dd lf = dd_mul_d(log_newton(x, y0, ef, k), y);
    if(!(lf.h > -708.0 && lf.h < 709.0)) return std::pow(x, y);
    const dd f = dd_exp_tab(lf, e);
    // the log's absolute error, times |y|, is a relative error of the result
    const double v = settle(f, f.h * (kExpErr + std::fabs(y) * kLogErr + 0x1p-90), e);
    if(v == v) return v;
    const dd es = dd_exp_sq(dd{ -y0, 0.0 }, k);
    const dd ls = dd_mul_d(log_newton(x, y0, es, k), y);
    const dd s = dd_exp_sq(ls, e);
    return std::ldexp(s.h + s.l, e);
}

// cr_pow(x, y1) and cr_pow(x, y2) with the logarithm of x formed once: every step is the one
// cr_pow takes, so the results are identical to two cr_pow calls.
CR_FN void cr_pow2(double x, double y1, double y2, double& r1, double& r2){
    if(!(x > 0.0) || !std::isfinite(x) || x < 0x1p-1000 || x > 0x1p1000){
        r1 = cr_pow(x, y1);
        r2 = cr_pow(x, y2);
        return;
    }
    const double y0 = std::log(x);
    int k;
    const dd ef = dd_exp_tab(dd{ -y0, 0.0 }, k);
    const dd lx = log_newton(x, y0, ef, k);
    bool slow_log = false;
    dd ls{ 0.0, 0.0 };
    auto one = [&](double y) -> double {
        if(!std::isfinite(y)) return std::pow(x, y);
        if(y == 0.0 || x == 1.0) return 1.0;
        int e;
        const dd lf = dd_mul_d(lx, y);
        if(!(lf.h > -708.0 && lf.h < 709.0)) return std::pow(x, y);
        const dd f = dd_exp_tab(lf, e);
        const double v = settle(f, f.h * (kExpErr + std::fabs(y) * kLogErr + 0x1p-90), e);
        if(v == v) return v;
        if(!slow_log){
            int k2;
            const dd es = dd_exp_sq(dd{ -y0, 0.0 }, k2);
            ls = log_newton(x, y0, es, k2);
            slow_log = true;
        }
        const dd s = dd_exp_sq(dd_mul_d(ls, y), e);
        return std::ldexp(s.h + s.l, e);
    };
    r1 = one(y1);
    r2 = one(y2);
}

// The transcendentals of a physics kernel instantiated for work type W. The FP64 step (W = double)
// takes the correctly rounded functions above, which reproduce the reference's glibc results; the
// FP32 step (W = float), whose error against the FP64 step its FP32 stencils set at ~1e-8 and
// above, takes the device library's (about one ulp), still in FP64 arithmetic.
template<typename W> struct Phys {
    static CR_FN double exp(double x){ return cr_exp(x); }
    static CR_FN double log(double x){ return cr_log(x); }
    static CR_FN double pow(double x, double y){ return cr_pow(x, y); }
    static CR_FN void pow2(double x, double y1, double y2, double& r1, double& r2){ cr_pow2(x, y1, y2, r1, r2); }
};
template<> struct Phys<float> {
    static CR_FN double exp(double x){ return std::exp(x); }
    static CR_FN double log(double x){ return std::log(x); }
    static CR_FN double pow(double x, double y){ return std::pow(x, y); }
    static CR_FN void pow2(double x, double y1, double y2, double& r1, double& r2){
        if(x > 0.0 && std::isfinite(x)){
            const double lx = std::log(x);
            r1 = std::exp(y1 * lx);
            r2 = std::exp(y2 * lx);
        } else {
            r1 = std::pow(x, y1);
            r2 = std::pow(x, y2);
        }
    }
};

}  // namespace crm
}  // namespace atsat

#undef CR_FN
#undef CR_TAB
