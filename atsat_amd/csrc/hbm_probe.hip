// hbm_probe: what HBM streaming rates this device reaches, for the roofline of the RK4 stage
// kernel. It moves synthetic data only; nothing here is part of the step.
//
//   triad      c = a + 2 b over three 1 GiB FP32 arrays, one pass (4 floats per thread per
//              array, all loads first), plain and non-temporal accesses;
//   groups     G groups of R read and W written 256 MiB arrays per thread (G (R + W) arrays in
//              flight, as a stage kernel streams ~100), non-temporal;
//   stencil    a 16-field 7-point FP32 stencil stage shaped like k_rk_cell: whole longitude
//              rings (360 cells, periodic) per workgroup, the XCD-contiguous band map, the
//              stencil input X read through the caches, y_n and the accumulator read and Y and
//              the accumulator written non-temporally -- with the full stencil and with the
//              centre value only (the same arrays streamed without neighbours).
//
// Each case runs 2 untimed and 5 timed launches (HIP events) and prints one JSON line with
// the bytes it moves (each value read or written once) per second.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#define CK(x) do{ hipError_t e_ = (x); if(e_ != hipSuccess){ fprintf(stderr, "HIP %s at line %d\n", hipGetErrorString(e_), __LINE__); exit(1);} }while(0)

namespace {

constexpr int kU = 4;                                  // floats per thread per array

template<bool NT> __device__ __forceinline__ float ld(const float* p){ return NT ? __builtin_nontemporal_load(p) : *p; }
template<bool NT> __device__ __forceinline__ void st(float* p, float v){ if(NT) __builtin_nontemporal_store(v, p); else *p = v; }

template<bool NT>
__global__ void __launch_bounds__(256) k_triad(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ c){
    const long long base = (long long)blockIdx.x * kU * 256 + threadIdx.x;
    float x[kU], y[kU];
#pragma unroll
    for(int u = 0; u < kU; u++){ x[u] = ld<NT>(a + base + u * 256); y[u] = ld<NT>(b + base + u * 256); }
#pragma unroll
    for(int u = 0; u < kU; u++) st<NT>(c + base + u * 256, x[u] + 2.0f * y[u]);
}

template<int R, int W>
__global__ void __launch_bounds__(256) k_groups(float* const* __restrict__ arr, int G){
    const long long base = (long long)blockIdx.x * kU * 256 + threadIdx.x;
#pragma unroll 1
    for(int g = 0; g < G; g++){
        float x[R][kU];
#pragma unroll
        for(int r = 0; r < R; r++)
#pragma unroll
            for(int u = 0; u < kU; u++) x[r][u] = ld<true>(arr[g * (R + W) + r] + base + u * 256);
#pragma unroll
        for(int w = 0; w < W; w++)
#pragma unroll
            for(int u = 0; u < kU; u++){
                float s = (float)w;
#pragma unroll
                for(int r = 0; r < R; r++) s += x[r][u];
                st<true>(arr[g * (R + W) + R + w] + base + u * 256, s);
            }
    }
}

// the stencil stage proxy on the native shell's shape
constexpr int kNF = 16, kIM = 41, kJM = 181, kKP = 384, kNPHI = 360;
struct Bands {
    int nj, nlev, M, band, nband, per, rings;
    __device__ __forceinline__ bool decode(int b, int sub, int& jr, int& lev, int& m) const {
        const int l = b >> 3, wpl = band / rings, seq = wpl * nlev;
        const int g = (b & 7) * per + l / seq, r = l % seq;
        lev = r / wpl;
        if(g >= nband * M) return false;
        m = g / nband;
        jr = (g % nband) * band + (r % wpl) * rings + sub;
        return jr < nj;
    }
    int blocks() const { return per * (band / rings) * nlev * 8; }
};
__device__ __forceinline__ int wrap(int k){ return k < 0 ? k + kNPHI : (k >= kNPHI ? k - kNPHI : k); }

template<bool FULL>
__global__ void __launch_bounds__(768) k_stencil(const float* __restrict__ X, const float* __restrict__ YN, const float* __restrict__ AC,
                                                 float* __restrict__ Y, float* __restrict__ AO, Bands BM, long long FS){
    const int rt = 384;                                  // threads per ring
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) / (rt >> 6);
    const int k = (int)threadIdx.x - sub * rt;
    int jr, lev, m;
    if(!BM.decode((int)blockIdx.x, sub, jr, lev, m) || k >= kNPHI) return;
    const long long row = (((long long)m * kIM + 1 + lev) * kJM + 2 + jr) * kKP;
    const long long c = row + k, sj = kKP, si = (long long)kJM * kKP;
    const long long pw = row + wrap(k - 1), pe = row + wrap(k + 1);
#pragma unroll
    for(int f = 0; f < kNF; f++){
        const float* x = X + f * FS;
        float v = 0.5f * x[c];
        if(FULL) v += 0.1f * (x[pw] + x[pe]) + 0.07f * (x[c - sj] + x[c + sj]) + 0.05f * (x[c - si] + x[c + si]);
        const float yn = ld<true>(YN + f * FS + c), a = ld<true>(AC + f * FS + c);
        st<true>(Y + f * FS + c, yn + 0.25f * v);
        st<true>(AO + f * FS + c, a + 2.0f * v);
    }
}

template<typename L>
double time_ms(L launch){
    for(int w = 0; w < 2; w++) launch();
    CK(hipGetLastError()); CK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    const int reps = 5;
    CK(hipEventRecord(e0, 0));
    for(int w = 0; w < reps; w++) launch();
    CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipEventDestroy(e0)); CK(hipEventDestroy(e1));
    return ms / reps;
}

void line(const char* kind, const char* what, double bytes, double ms, double read_share, int arrays){
    printf("{\"probe\": \"%s\", \"case\": \"%s\", \"arrays\": %d, \"read_share\": %.2f, \"ms\": %.4f, \"GBps\": %.1f}\n",
           kind, what, arrays, read_share, ms, bytes / (ms * 1e-3) / 1e9);
    fflush(stdout);
}

}  // namespace

int main(){
    // triad: three 1 GiB arrays
    {
        const long long n = 1ll << 28;
        float *a, *b, *c;
        CK(hipMalloc(&a, n * 4)); CK(hipMalloc(&b, n * 4)); CK(hipMalloc(&c, n * 4));
        CK(hipMemset(a, 0, n * 4)); CK(hipMemset(b, 0, n * 4)); CK(hipMemset(c, 0, n * 4));
        const int blocks = (int)(n / (kU * 256));
        const double bytes = 3.0 * n * 4;
        line("triad", "plain", bytes, time_ms([&]{ hipLaunchKernelGGL(k_triad<false>, dim3(blocks), dim3(256), 0, 0, a, b, c); }), 2.0 / 3, 3);
        line("triad", "non-temporal", bytes, time_ms([&]{ hipLaunchKernelGGL(k_triad<true>, dim3(blocks), dim3(256), 0, 0, a, b, c); }), 2.0 / 3, 3);
        CK(hipFree(a)); CK(hipFree(b)); CK(hipFree(c));
    }
    // groups: up to 96 arrays of 256 MiB
    {
        const long long n = 1ll << 26;
        const int narr = 96;
        std::vector<float*> h(narr);
        for(auto& p : h){ CK(hipMalloc(&p, n * 4)); CK(hipMemset(p, 0, n * 4)); }
        float** d;
        CK(hipMalloc(&d, narr * sizeof(float*)));
        CK(hipMemcpy(d, h.data(), narr * sizeof(float*), hipMemcpyHostToDevice));
        const int blocks = (int)(n / (kU * 256));
        auto grp = [&](auto tag, int G){
            constexpr int R = decltype(tag)::value / 10, W = decltype(tag)::value % 10;
            char what[64];
            snprintf(what, sizeof what, "%d groups of %d read + %d written", G, R, W);
            line("groups", what, (double)G * (R + W) * n * 4,
                 time_ms([&]{ hipLaunchKernelGGL((k_groups<R, W>), dim3(blocks), dim3(256), 0, 0, d, G); }), (double)R / (R + W), G * (R + W));
        };
        using T21 = std::integral_constant<int, 21>;          // R read, W written: 10 R + W
        using T32 = std::integral_constant<int, 32>;
        using T31 = std::integral_constant<int, 31>;
        grp(T21{}, 1); grp(T21{}, 16); grp(T21{}, 32);
        grp(T32{}, 1); grp(T32{}, 16);
        grp(T31{}, 1); grp(T31{}, 16); grp(T31{}, 24);
        for(auto p : h) CK(hipFree(p));
        CK(hipFree(d));
    }
    // the stencil stage proxy: 16 members of the native shell, band 6, 2 rings per workgroup
    {
        const int M = 16;
        const long long FS = (long long)M * kIM * kJM * kKP;
        const size_t bytes = (size_t)FS * kNF * sizeof(float);
        float *X, *YN, *AC, *Y, *AO;
        CK(hipMalloc(&X, bytes)); CK(hipMalloc(&YN, bytes)); CK(hipMalloc(&AC, bytes)); CK(hipMalloc(&Y, bytes)); CK(hipMalloc(&AO, bytes));
        CK(hipMemset(X, 0, bytes)); CK(hipMemset(YN, 0, bytes)); CK(hipMemset(AC, 0, bytes)); CK(hipMemset(Y, 0, bytes)); CK(hipMemset(AO, 0, bytes));
        Bands BM{kJM - 4, kIM - 2, M, 6, 0, 0, 2};
        BM.nband = (BM.nj + BM.band - 1) / BM.band;
        BM.per = (BM.nband * M + 7) / 8;
        const double cells = (double)M * (kIM - 2) * (kJM - 4) * kNPHI;
        const double alg = cells * kNF * 5 * 4.0;
        const dim3 grid(BM.blocks()), block(768);
        line("stencil", "16 fields, 7-point stencil", alg, time_ms([&]{ hipLaunchKernelGGL(k_stencil<true>, grid, block, 0, 0, X, YN, AC, Y, AO, BM, FS); }), 0.6, 80);
        line("stencil", "16 fields, centre only", alg, time_ms([&]{ hipLaunchKernelGGL(k_stencil<false>, grid, block, 0, 0, X, YN, AC, Y, AO, BM, FS); }), 0.6, 80);
        CK(hipFree(X)); CK(hipFree(YN)); CK(hipFree(AC)); CK(hipFree(Y)); CK(hipFree(AO));
    }
    return 0;
}
