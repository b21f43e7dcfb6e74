// ddyn_probe.hip -- measurement support for the roofline of the pair kernels.
//
//   ddyn_probe rates   issue-rate microbenchmarks on gfx950: each FP32 instruction
//                      class of the pair loop (FMA forms, mul, sub, min3, rsq),
//                      the shader clock under load, FP64 FMA and IEEE FP64
//                      sqrt + divide
//   ddyn_probe copy    a plain copy kernel of known size (256 MiB read, 256 MiB
//                      written, three launches) to calibrate the FETCH_SIZE and
//                      WRITE_SIZE counters of rocprofv3 --pmc
//
// Every kernel indexes only within its grid; buffers are sized to the grid.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CK(x)                                                                    \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                        \
        }                                                                        \
    } while (0)

// Issue rate of one FP32 instruction class, 16 independent chains per lane,
// register operands (b, c are loaded, so the compiler cannot fold literals).
enum Op { kFma3 = 0, kFmac = 1, kFmaK = 2, kMul = 3, kSub = 4, kMin3 = 5, kRsq = 6, kSqAcc = 7, kSqAdd = 8 };

template <int OP>
__global__ __launch_bounds__(256) void r_op32(float* out, const float* in, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float b = in[threadIdx.x & 7], c = in[8 + (threadIdx.x & 7)];
    float a[16], d[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        a[k] = in[16 + k] + t * 1e-7f;
        d[k] = 1e-3f * in[16 + (k ^ 1)];
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (OP == kFma3) a[k] = __builtin_fmaf(a[k], b, c);          // v_fma_f32, 3 VGPR sources
            if (OP == kFmac) a[k] = __builtin_fmaf(b, c, a[k]);          // v_fmac_f32 (VOP2)
            if (OP == kFmaK) a[k] = __builtin_fmaf(a[k], 0.999f, c);     // v_fmamk_f32, literal
            if (OP == kMul) a[k] = a[k] * b;                             // v_mul_f32
            if (OP == kSub) a[k] = b - a[k];                             // v_sub_f32
            if (OP == kMin3) a[k] = fminf(fminf(a[k], b), c);            // v_min3_f32
            if (OP == kRsq) a[k] = __builtin_amdgcn_rsqf(a[k]);          // v_rsq_f32
            if (OP == kSqAcc) a[k] = __builtin_fmaf(d[k], d[k], a[k]);   // v_fmac_f32, two distinct sources
            if (OP == kSqAdd) a[k] = __builtin_fmaf(a[k], a[k], c) * 0.5f;   // v_fma_f32 a*a+c, then v_mul
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += a[k];
    out[t] = s;
}

// Shader clock under FMA load: s_memtime (shader clocks) over s_memrealtime (100 MHz).
__global__ __launch_bounds__(256) void r_clock(unsigned long long* ticks, const float* in, float* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long c0 = __builtin_readcyclecounter(), r0 = __builtin_amdgcn_s_memrealtime();
    const float b = in[0], c = in[8];
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = in[16 + k] + t * 1e-7f;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], b, c);
    const unsigned long long c1 = __builtin_readcyclecounter(), r1 = __builtin_amdgcn_s_memrealtime();
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
    if (threadIdx.x == 0) {
        ticks[2 * blockIdx.x] = c1 - c0;
        ticks[2 * blockIdx.x + 1] = r1 - r0;
    }
}

__global__ void r_fma64(double* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a[8];
    for (int k = 0; k < 8; ++k) a[k] = t * 1e-7 + k;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = __builtin_fma(a[k], 0.999, 1e-3);
    double s = 0.0;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
}

__global__ void r_sqrtdiv64(double* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a[8];
    for (int k = 0; k < 8; ++k) a[k] = 1.0 + t * 1e-7 + k;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = 1.0 + 1e-3 / (a[k] * __builtin_sqrt(a[k]));
    double s = 0.0;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
}

__global__ void copy_kernel(const float4* __restrict__ a, float4* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

template <class F>
static float time_ms(F f, int reps)
{
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    f();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(a));
    for (int r = 0; r < reps; ++r) f();
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return ms / reps;
}

static int rates()
{
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    std::printf("device %s, %d CUs, max clock %d MHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate / 1000);
    const int blocks = 256 * 32, threads = 256, iters = 2000;
    const double n = (double)blocks * threads;
    float *o32, *in;
    double* o64;
    unsigned long long* ticks;
    CK(hipMalloc(&o32, (size_t)n * sizeof(float)));
    CK(hipMalloc(&o64, (size_t)n * sizeof(double)));
    CK(hipMalloc(&in, 64 * sizeof(float)));
    CK(hipMalloc(&ticks, 2 * (size_t)blocks * sizeof(unsigned long long)));
    float hin[64];
    for (int k = 0; k < 64; ++k) hin[k] = k < 8 ? 0.9999f : (k < 16 ? 1e-4f : 1.0f + 0.01f * k);
    CK(hipMemcpy(in, hin, sizeof(hin), hipMemcpyHostToDevice));
    const char* names[9] = {"v_fma_f32 (3 VGPR sources)", "v_fmac_f32 (3 VGPR sources)", "v_fmamk_f32 (literal)",
                            "v_mul_f32", "v_sub_f32", "v_min3_f32", "v_rsq_f32", "v_fmac_f32 d*d+a",
                            "v_fma_f32 a*a+c with v_mul"};
    double rate[9];
    float ms;
    ms = time_ms([&] { r_op32<kFma3><<<blocks, threads>>>(o32, in, iters); }, 5); rate[0] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kFmac><<<blocks, threads>>>(o32, in, iters); }, 5); rate[1] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kFmaK><<<blocks, threads>>>(o32, in, iters); }, 5); rate[2] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kMul><<<blocks, threads>>>(o32, in, iters); }, 5); rate[3] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kSub><<<blocks, threads>>>(o32, in, iters); }, 5); rate[4] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kMin3><<<blocks, threads>>>(o32, in, iters); }, 5); rate[5] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kRsq><<<blocks, threads>>>(o32, in, iters / 4); }, 5); rate[6] = n * (iters / 4) * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kSqAcc><<<blocks, threads>>>(o32, in, iters); }, 5); rate[7] = n * iters * 16 / ms * 1e-9;
    ms = time_ms([&] { r_op32<kSqAdd><<<blocks, threads>>>(o32, in, iters); }, 5); rate[8] = 2 * n * iters * 16 / ms * 1e-9;
    for (int k = 0; k < 9; ++k) std::printf("%-32s %8.2f Tlane-op/s\n", names[k], rate[k]);
    // One interaction of the FP32 pair loop as compiled: 3 v_sub; r^2 as one v_fma
    // dx*dx+eps^2 and two v_fmac d*d+r^2 (two distinct sources each); 3 v_mul;
    // 3 v_fmac accumulations (three distinct sources); 1 v_rsq; 1/2 v_min3.
    const double t_int = 3 / rate[4] + 3 / rate[7] + 3 / rate[3] + 3 / rate[1] + 1 / rate[6] + 0.5 / rate[5];
    std::printf("pair interaction, additive issue model: %.3e interactions/s (%.1f TFLOP/s at 20 FLOP, %.1f%% of 157.3)\n",
                1e12 / t_int, 20 / t_int, 20 / t_int / 157.3 * 100);
    r_clock<<<blocks, threads>>>(ticks, in, o32, 20000);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    unsigned long long* h = (unsigned long long*)std::malloc(2 * (size_t)blocks * sizeof(unsigned long long));
    CK(hipMemcpy(h, ticks, 2 * (size_t)blocks * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double f = 0.0;
    int cnt = 0;
    for (int b = 0; b < blocks; ++b)
        if (h[2 * b + 1] > 0) { f += (double)h[2 * b] / ((double)h[2 * b + 1] / 100e6); ++cnt; }
    std::free(h);
    std::printf("shader clock under FMA load: %.0f MHz (mean over %d workgroups)\n", f / cnt / 1e6, cnt);
    ms = time_ms([&] { r_fma64<<<blocks, threads>>>(o64, iters); }, 5);
    std::printf("v_fma_f64          %8.2f Gop/s  (%.1f TFLOP/s)\n", n * iters * 8 / ms * 1e-6, 2 * n * iters * 8 / ms * 1e-9);
    ms = time_ms([&] { r_sqrtdiv64<<<blocks, threads>>>(o64, iters / 10); }, 5);
    std::printf("f64 sqrt + div     %8.2f G/s\n", n * (iters / 10) * 8 / ms * 1e-6);
    CK(hipFree(o32));
    CK(hipFree(o64));
    CK(hipFree(in));
    CK(hipFree(ticks));
    return 0;
}

static int copy()
{
    const size_t bytes = (size_t)256 << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a, *b;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 1, bytes));
    CK(hipMemset(b, 0, bytes));
    for (int r = 0; r < 3; ++r) copy_kernel<<<4096, 256>>>(a, b, n);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    std::printf("copy_kernel: %zu bytes read and %zu written per launch, 3 launches\n", bytes, bytes);
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "rates") == 0) return rates();
    if (argc == 2 && std::strcmp(argv[1], "copy") == 0) return copy();
    std::fprintf(stderr, "usage: ddyn_probe rates | copy\n");
    return 1;
}
