// ddyn_probe.hip -- measurement support for the roofline of the pair kernels.
//
//   ddyn_probe rates   issue-rate microbenchmarks on gfx950: each FP32 instruction
//                      class of the pair loop (FMA forms, mul, sub, min3, rsq),
//                      the shader clock under load, FP64 FMA and IEEE FP64
//                      sqrt + divide
//   ddyn_probe mix     the FP32 pair loop's instruction sequence on fixed
//                      registers: alone, with its LDS reads, without v_min3,
//                      with v_rsq replaced by v_mul, 4 / 8 / 16 interactions
//                      per loop iteration
//   ddyn_probe rsq64   accuracy of FP64 reciprocal square roots over 2^22
//                      inputs spread over 24 decades: the v_rsq_f64 estimate
//                      alone, refined by two Newton steps, and refined by one
//                      third-order step, against long double on the host
//   ddyn_probe copy    a plain copy kernel of known size (256 MiB read, 256 MiB
//                      written, three launches) to calibrate the FETCH_SIZE and
//                      WRITE_SIZE counters of rocprofv3 --pmc
//
// Every kernel indexes only within its grid; buffers are sized to the grid.
#include <hip/hip_runtime.h>

#include <cmath>
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
// Chain k reads b[k % 4] and c[(k / 4) % 4]: consecutive instructions read
// different registers, as in the pair loop (the distance component changes
// every instruction, the force factor every third).  kFmacSame reads the same
// two registers in every instruction, a pattern that issues much slower.
enum Op { kFma3 = 0, kFmac = 1, kFmaK = 2, kMul = 3, kSub = 4, kMin3 = 5, kRsq = 6, kSqAcc = 7, kSqAdd = 8,
          kFmacSame = 9 };

template <int OP>
__global__ __launch_bounds__(256) void r_op32(float* out, const float* in, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float b[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        b[k] = in[(threadIdx.x + k) & 7];
        c[k] = in[8 + ((threadIdx.x + k) & 7)];
    }
    float a[16], d[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        a[k] = in[16 + k] + t * 1e-7f;
        d[k] = 1e-3f * in[16 + (k ^ 1)];
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float bk = b[k & 3], ck = c[(k >> 2) & 3];
            if (OP == kFma3) a[k] = __builtin_fmaf(a[k], bk, ck);        // v_fma_f32, 3 VGPR sources
            if (OP == kFmac) a[k] = __builtin_fmaf(bk, ck, a[k]);        // v_fmac_f32 (VOP2)
            if (OP == kFmaK) a[k] = __builtin_fmaf(a[k], 0.999f, ck);    // v_fmamk_f32, literal
            if (OP == kMul) a[k] = a[k] * bk;                            // v_mul_f32
            if (OP == kSub) a[k] = bk - a[k];                            // v_sub_f32
            if (OP == kMin3) a[k] = fminf(fminf(a[k], bk), ck);          // v_min3_f32
            if (OP == kRsq) a[k] = __builtin_amdgcn_rsqf(a[k]);          // v_rsq_f32
            if (OP == kSqAcc) a[k] = __builtin_fmaf(d[k], d[k], a[k]);   // v_fmac_f32, two distinct sources
            if (OP == kSqAdd) a[k] = __builtin_fmaf(a[k], a[k], ck) * 0.5f;   // v_fma_f32 a*a+c, then v_mul
            if (OP == kFmacSame) a[k] = __builtin_fmaf(b[0], c[0], a[k]);    // same two sources every time
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

// out[3i]: v_rsq_f64 estimate; out[3i+1]: two Newton steps y += y (1 - x y^2) / 2;
// out[3i+2]: one third-order step y += y e (1/2 + 3e/8), e = 1 - x y^2 (the
// form the library uses).
__global__ __launch_bounds__(256) void r_rsq64(const double* __restrict__ in, double* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in[i];
    const double y0 = __builtin_amdgcn_rsq(x);
    double y = y0;
    for (int it = 0; it < 2; ++it) {
        const double r = __builtin_fma(-x * y, y, 1.0);
        y = __builtin_fma(0.5 * y, r, y);
    }
    const double e = __builtin_fma(-x * y0, y0, 1.0);
    const double h = __builtin_fma(y0 * e, __builtin_fma(0.375, e, 0.5), y0);
    out[3 * i] = y0;
    out[3 * i + 1] = y;
    out[3 * i + 2] = h;
}

int rsq64()
{
    const int n = 1 << 22;
    double* hin = (double*)std::malloc(sizeof(double) * n);
    double* hout = (double*)std::malloc(sizeof(double) * 3 * n);
    unsigned long long s = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < n; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
        hin[i] = std::pow(10.0, -12.0 + 24.0 * u);
    }
    double *din, *dout;
    CK(hipMalloc(&din, sizeof(double) * n));
    CK(hipMalloc(&dout, sizeof(double) * 3 * n));
    CK(hipMemcpy(din, hin, sizeof(double) * n, hipMemcpyHostToDevice));
    r_rsq64<<<(n + 255) / 256, 256>>>(din, dout, n);
    CK(hipGetLastError());
    CK(hipMemcpy(hout, dout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    const char* names[3] = {"v_rsq_f64 estimate", "two Newton steps", "one third-order step"};
    for (int k = 0; k < 3; ++k) {
        long double emax = 0.0L, esum2 = 0.0L;
        for (int i = 0; i < n; ++i) {
            const long double ref = 1.0L / std::sqrt((long double)hin[i]);
            const long double e = std::fabs(((long double)hout[3 * i + k] - ref) / ref);
            emax = e > emax ? e : emax;
            esum2 += e * e;
        }
        std::printf("%-22s max rel err %.3Le (%.2Lf x 2^-53), rms %.3Le\n", names[k], emax,
                    emax / 1.1102230246251565e-16L, std::sqrt(esum2 / n));
    }
    CK(hipFree(din));
    CK(hipFree(dout));
    std::free(hin);
    std::free(hout);
    return 0;
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
    const char* names[10] = {"v_fma_f32 (3 VGPR sources)", "v_fmac_f32 (3 VGPR sources)", "v_fmamk_f32 (literal)",
                             "v_mul_f32", "v_sub_f32", "v_min3_f32", "v_rsq_f32", "v_fmac_f32 d*d+a",
                             "v_fma_f32 a*a+c with v_mul", "v_fmac_f32, same 2 sources each op"};
    double rate[10];
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
    ms = time_ms([&] { r_op32<kFmacSame><<<blocks, threads>>>(o32, in, iters); }, 5); rate[9] = n * iters * 16 / ms * 1e-9;
    for (int k = 0; k < 10; ++k) std::printf("%-36s %8.2f Tlane-op/s\n", names[k], rate[k]);
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

// ---------------------------------------------------------------------------
// The FP32 pair loop's instruction sequence on fixed registers (ddyn_probe mix)
// ---------------------------------------------------------------------------
// Four interactions as the compiler issues them in k_pairs32: v0-v15 hold four
// sources (x, y, z, m), v16-v18 the target, v19 eps^2, v20-v22 the force sums,
// v23 the close-pair minimum, v24-v51 temporaries.  The variants time the
// sequence alone, with its four LDS source reads, without v_min3, with v_rsq
// replaced by v_mul, and with two or four copies per loop iteration.  The
// values computed are irrelevant; every source register is written before use.
#define MIX4 \
    "v_sub_f32 v24, v0, v16\n" "v_sub_f32 v25, v1, v17\n" "v_sub_f32 v26, v2, v18\n" "v_sub_f32 v31, v4, v16\n" \
    "v_sub_f32 v32, v5, v17\n" "v_sub_f32 v33, v6, v18\n" "v_sub_f32 v38, v8, v16\n" "v_sub_f32 v39, v9, v17\n" \
    "v_sub_f32 v40, v10, v18\n" "v_sub_f32 v45, v12, v16\n" "v_sub_f32 v46, v13, v17\n" "v_sub_f32 v47, v14, v18\n" \
    "v_fma_f32 v27, v24, v24, v19\n" "v_fma_f32 v34, v31, v31, v19\n" "v_fma_f32 v41, v38, v38, v19\n" "v_fma_f32 v48, v45, v45, v19\n" \
    "v_fmac_f32 v27, v25, v25\n" "v_fmac_f32 v34, v32, v32\n" "v_fmac_f32 v41, v39, v39\n" "v_fmac_f32 v48, v46, v46\n" \
    "v_fmac_f32 v27, v26, v26\n" "v_fmac_f32 v34, v33, v33\n" "v_fmac_f32 v41, v40, v40\n" "v_fmac_f32 v48, v47, v47\n" \
    "v_rsq_f32 v28, v27\n" "v_rsq_f32 v35, v34\n" "v_rsq_f32 v42, v41\n" "v_rsq_f32 v49, v48\n" \
    "v_min3_f32 v23, v23, v27, v34\n" "v_min3_f32 v23, v23, v41, v48\n" "v_mul_f32 v29, v28, v28\n" "v_mul_f32 v30, v3, v28\n" \
    "v_mul_f32 v30, v30, v29\n" "v_mul_f32 v36, v35, v35\n" "v_mul_f32 v37, v7, v35\n" "v_mul_f32 v37, v37, v36\n" \
    "v_mul_f32 v43, v42, v42\n" "v_mul_f32 v44, v11, v42\n" "v_mul_f32 v44, v44, v43\n" "v_mul_f32 v50, v49, v49\n" \
    "v_mul_f32 v51, v15, v49\n" "v_mul_f32 v51, v51, v50\n" "v_fmac_f32 v20, v24, v30\n" "v_fmac_f32 v21, v25, v30\n" \
    "v_fmac_f32 v22, v26, v30\n" "v_fmac_f32 v20, v31, v37\n" "v_fmac_f32 v21, v32, v37\n" "v_fmac_f32 v22, v33, v37\n" \
    "v_fmac_f32 v20, v38, v44\n" "v_fmac_f32 v21, v39, v44\n" "v_fmac_f32 v22, v40, v44\n" "v_fmac_f32 v20, v45, v51\n" \
    "v_fmac_f32 v21, v46, v51\n" "v_fmac_f32 v22, v47, v51\n"
#define MIX4_NOMIN3 \
    "v_sub_f32 v24, v0, v16\n" "v_sub_f32 v25, v1, v17\n" "v_sub_f32 v26, v2, v18\n" "v_sub_f32 v31, v4, v16\n" \
    "v_sub_f32 v32, v5, v17\n" "v_sub_f32 v33, v6, v18\n" "v_sub_f32 v38, v8, v16\n" "v_sub_f32 v39, v9, v17\n" \
    "v_sub_f32 v40, v10, v18\n" "v_sub_f32 v45, v12, v16\n" "v_sub_f32 v46, v13, v17\n" "v_sub_f32 v47, v14, v18\n" \
    "v_fma_f32 v27, v24, v24, v19\n" "v_fma_f32 v34, v31, v31, v19\n" "v_fma_f32 v41, v38, v38, v19\n" "v_fma_f32 v48, v45, v45, v19\n" \
    "v_fmac_f32 v27, v25, v25\n" "v_fmac_f32 v34, v32, v32\n" "v_fmac_f32 v41, v39, v39\n" "v_fmac_f32 v48, v46, v46\n" \
    "v_fmac_f32 v27, v26, v26\n" "v_fmac_f32 v34, v33, v33\n" "v_fmac_f32 v41, v40, v40\n" "v_fmac_f32 v48, v47, v47\n" \
    "v_rsq_f32 v28, v27\n" "v_rsq_f32 v35, v34\n" "v_rsq_f32 v42, v41\n" "v_rsq_f32 v49, v48\n" \
    "v_mul_f32 v29, v28, v28\n" "v_mul_f32 v30, v3, v28\n" "v_mul_f32 v30, v30, v29\n" "v_mul_f32 v36, v35, v35\n" \
    "v_mul_f32 v37, v7, v35\n" "v_mul_f32 v37, v37, v36\n" "v_mul_f32 v43, v42, v42\n" "v_mul_f32 v44, v11, v42\n" \
    "v_mul_f32 v44, v44, v43\n" "v_mul_f32 v50, v49, v49\n" "v_mul_f32 v51, v15, v49\n" "v_mul_f32 v51, v51, v50\n" \
    "v_fmac_f32 v20, v24, v30\n" "v_fmac_f32 v21, v25, v30\n" "v_fmac_f32 v22, v26, v30\n" "v_fmac_f32 v20, v31, v37\n" \
    "v_fmac_f32 v21, v32, v37\n" "v_fmac_f32 v22, v33, v37\n" "v_fmac_f32 v20, v38, v44\n" "v_fmac_f32 v21, v39, v44\n" \
    "v_fmac_f32 v22, v40, v44\n" "v_fmac_f32 v20, v45, v51\n" "v_fmac_f32 v21, v46, v51\n" "v_fmac_f32 v22, v47, v51\n"
#define MIX4_RSQ_AS_MUL \
    "v_sub_f32 v24, v0, v16\n" "v_sub_f32 v25, v1, v17\n" "v_sub_f32 v26, v2, v18\n" "v_sub_f32 v31, v4, v16\n" \
    "v_sub_f32 v32, v5, v17\n" "v_sub_f32 v33, v6, v18\n" "v_sub_f32 v38, v8, v16\n" "v_sub_f32 v39, v9, v17\n" \
    "v_sub_f32 v40, v10, v18\n" "v_sub_f32 v45, v12, v16\n" "v_sub_f32 v46, v13, v17\n" "v_sub_f32 v47, v14, v18\n" \
    "v_fma_f32 v27, v24, v24, v19\n" "v_fma_f32 v34, v31, v31, v19\n" "v_fma_f32 v41, v38, v38, v19\n" "v_fma_f32 v48, v45, v45, v19\n" \
    "v_fmac_f32 v27, v25, v25\n" "v_fmac_f32 v34, v32, v32\n" "v_fmac_f32 v41, v39, v39\n" "v_fmac_f32 v48, v46, v46\n" \
    "v_fmac_f32 v27, v26, v26\n" "v_fmac_f32 v34, v33, v33\n" "v_fmac_f32 v41, v40, v40\n" "v_fmac_f32 v48, v47, v47\n" \
    "v_mul_f32 v28, v27, v27\n" "v_mul_f32 v35, v34, v34\n" "v_mul_f32 v42, v41, v41\n" "v_mul_f32 v49, v48, v48\n" \
    "v_min3_f32 v23, v23, v27, v34\n" "v_min3_f32 v23, v23, v41, v48\n" "v_mul_f32 v29, v28, v28\n" "v_mul_f32 v30, v3, v28\n" \
    "v_mul_f32 v30, v30, v29\n" "v_mul_f32 v36, v35, v35\n" "v_mul_f32 v37, v7, v35\n" "v_mul_f32 v37, v37, v36\n" \
    "v_mul_f32 v43, v42, v42\n" "v_mul_f32 v44, v11, v42\n" "v_mul_f32 v44, v44, v43\n" "v_mul_f32 v50, v49, v49\n" \
    "v_mul_f32 v51, v15, v49\n" "v_mul_f32 v51, v51, v50\n" "v_fmac_f32 v20, v24, v30\n" "v_fmac_f32 v21, v25, v30\n" \
    "v_fmac_f32 v22, v26, v30\n" "v_fmac_f32 v20, v31, v37\n" "v_fmac_f32 v21, v32, v37\n" "v_fmac_f32 v22, v33, v37\n" \
    "v_fmac_f32 v20, v38, v44\n" "v_fmac_f32 v21, v39, v44\n" "v_fmac_f32 v22, v40, v44\n" "v_fmac_f32 v20, v45, v51\n" \
    "v_fmac_f32 v21, v46, v51\n" "v_fmac_f32 v22, v47, v51\n"
#define MIXCLOB "v0","v1","v2","v3","v4","v5","v6","v7","v8","v9","v10","v11","v12","v13","v14","v15","v16","v17","v18","v19","v20","v21","v22","v23","v24","v25","v26","v27","v28","v29","v30","v31","v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51"
#define MIXINIT "v_mov_b32 v0, %0\n""v_mov_b32 v1, %0\n""v_mov_b32 v2, %0\n""v_mov_b32 v3, %0\n""v_mov_b32 v4, %0\n""v_mov_b32 v5, %0\n""v_mov_b32 v6, %0\n""v_mov_b32 v7, %0\n""v_mov_b32 v8, %0\n""v_mov_b32 v9, %0\n""v_mov_b32 v10, %0\n""v_mov_b32 v11, %0\n""v_mov_b32 v12, %0\n""v_mov_b32 v13, %0\n""v_mov_b32 v14, %0\n""v_mov_b32 v15, %0\n""v_mov_b32 v16, %0\n""v_mov_b32 v17, %0\n""v_mov_b32 v18, %0\n""v_mov_b32 v19, %0\n""v_mov_b32 v20, %0\n""v_mov_b32 v21, %0\n""v_mov_b32 v22, %0\n""v_mov_b32 v23, %0\n"

enum Mix { kMixSeq = 0, kMixLds = 1, kMixNoMin3 = 2, kMixRsqAsMul = 3 };

template <int V, int REP>
__global__ __launch_bounds__(256) void r_mix(float* out, int iters, float seed)
{
    __shared__ float4 tile[256];
    tile[threadIdx.x] = make_float4(1.0f + threadIdx.x, 2.0f, 3.0f, 1e-3f);
    __syncthreads();
    const float x = seed + threadIdx.x * 1e-7f;
    asm volatile(MIXINIT :: "v"(x) : MIXCLOB);
    for (int i = 0; i < iters; ++i) {
        if constexpr (V == kMixLds) {
            // four 16-byte reads inside the 4 KiB tile, waited for in the same block
            auto p = (__attribute__((address_space(3))) float4*)(tile + ((i * 4) & 252));
            asm volatile("ds_read_b128 v[0:3], %0\n ds_read_b128 v[4:7], %0 offset:16\n"
                         "ds_read_b128 v[8:11], %0 offset:32\n ds_read_b128 v[12:15], %0 offset:48\n"
                         "s_waitcnt lgkmcnt(0)\n" MIX4 :: "v"(p) : MIXCLOB, "memory");
        } else if constexpr (V == kMixNoMin3) {
            asm volatile(MIX4_NOMIN3 ::: MIXCLOB);
        } else if constexpr (V == kMixRsqAsMul) {
            asm volatile(MIX4_RSQ_AS_MUL ::: MIXCLOB);
        } else if constexpr (REP == 1) {
            asm volatile(MIX4 ::: MIXCLOB);
        } else if constexpr (REP == 2) {
            asm volatile(MIX4 MIX4 ::: MIXCLOB);
        } else {
            asm volatile(MIX4 MIX4 MIX4 MIX4 ::: MIXCLOB);
        }
    }
    float r;
    asm volatile("v_add_f32 %0, v20, v23" : "=v"(r) :: MIXCLOB);
    out[blockIdx.x * 256 + threadIdx.x] = r;
}

static int mix()
{
    const int blocks = 256 * 32, threads = 256, iters = 2048;
    float* o32;
    CK(hipMalloc(&o32, (size_t)blocks * threads * sizeof(float)));
    const double lanes = (double)blocks * threads;
    auto report = [&](const char* name, float ms, int per_iter, int it) {
        const double r = lanes * it * per_iter / (ms * 1e-3);
        std::printf("%-52s %.3e interactions/s (%.1f%% of 157.3 TFLOP/s at 20 FLOP)\n", name, r, 2e-11 * r / 157.3 * 100);
    };
    report("sequence of 4 interactions, registers only",
           time_ms([&] { r_mix<kMixSeq, 1><<<blocks, threads>>>(o32, iters, 1.f); }, 5), 4, iters);
    report("  + its four ds_read_b128 source reads",
           time_ms([&] { r_mix<kMixLds, 1><<<blocks, threads>>>(o32, iters, 1.f); }, 5), 4, iters);
    report("  without the two v_min3",
           time_ms([&] { r_mix<kMixNoMin3, 1><<<blocks, threads>>>(o32, iters, 1.f); }, 5), 4, iters);
    report("  with v_rsq_f32 replaced by v_mul_f32",
           time_ms([&] { r_mix<kMixRsqAsMul, 1><<<blocks, threads>>>(o32, iters, 1.f); }, 5), 4, iters);
    report("  two copies per loop iteration (8 interactions)",
           time_ms([&] { r_mix<kMixSeq, 2><<<blocks, threads>>>(o32, iters / 2, 1.f); }, 5), 8, iters / 2);
    report("  four copies per loop iteration (16 interactions)",
           time_ms([&] { r_mix<kMixSeq, 4><<<blocks, threads>>>(o32, iters / 4, 1.f); }, 5), 16, iters / 4);
    CK(hipFree(o32));
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
    if (argc == 2 && std::strcmp(argv[1], "mix") == 0) return mix();
    if (argc == 2 && std::strcmp(argv[1], "rsq64") == 0) return rsq64();
    std::fprintf(stderr, "usage: ddyn_probe rates | copy | mix | rsq64\n");
    return 1;
}
