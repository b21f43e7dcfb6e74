// ddyn_probe.hip -- measurement support for the roofline of the pair kernels.
//
//   ddyn_probe rates   issue-rate microbenchmarks on gfx950: FP32 FMA, v_rsq_f32,
//                      one pair interaction's instruction mix (12 FMA + 1 rsq),
//                      FP64 FMA, and IEEE FP64 sqrt + divide
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

__global__ void r_fma32(float* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
    for (int k = 0; k < 8; ++k) a[k] = t * 1e-7f + k;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], 0.999f, 1e-3f);
    float s = 0.f;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
}

__global__ void r_rsq32(float* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
    for (int k = 0; k < 8; ++k) a[k] = 1.0f + t * 1e-7f + k;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = __builtin_amdgcn_rsqf(a[k]);
    float s = 0.f;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[t] = s;
}

// one pair interaction's mix: 12 FMA-class ops + 1 v_rsq_f32, 4 independent chains
__global__ void r_mix32(float* out, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[4], b[4];
    for (int k = 0; k < 4; ++k) { a[k] = 1.0f + t * 1e-7f + k; b[k] = 2.0f + k; }
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float r = __builtin_amdgcn_rsqf(a[k]);
            float x = b[k];
#pragma unroll
            for (int q = 0; q < 11; ++q) x = __builtin_fmaf(x, 0.999f, 1e-4f);
            b[k] = x;
            a[k] = __builtin_fmaf(r, 1e-3f, 1.0f);
        }
    float s = 0.f;
    for (int k = 0; k < 4; ++k) s += a[k] + b[k];
    out[t] = s;
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
    const int blocks = 256 * 16, threads = 256, iters = 2000;
    const double n = (double)blocks * threads;
    float* o32;
    double* o64;
    CK(hipMalloc(&o32, (size_t)n * sizeof(float)));
    CK(hipMalloc(&o64, (size_t)n * sizeof(double)));
    float ms = time_ms([&] { r_fma32<<<blocks, threads>>>(o32, iters); }, 5);
    std::printf("v_fma_f32          %8.2f Gop/s  (%.1f TFLOP/s)\n", n * iters * 8 / ms * 1e-6, 2 * n * iters * 8 / ms * 1e-9);
    ms = time_ms([&] { r_rsq32<<<blocks, threads>>>(o32, iters); }, 5);
    std::printf("v_rsq_f32          %8.2f Gop/s\n", n * iters * 8 / ms * 1e-6);
    ms = time_ms([&] { r_mix32<<<blocks, threads>>>(o32, iters); }, 5);
    std::printf("12 FMA + 1 rsq mix %8.2f G mixes/s (%.3e pair interactions/s ceiling, %.1f TFLOP/s at 20 FLOP)\n",
                n * iters * 4 / ms * 1e-6, n * iters * 4 / ms * 1e3, 20 * n * iters * 4 / ms * 1e-9);
    ms = time_ms([&] { r_fma64<<<blocks, threads>>>(o64, iters); }, 5);
    std::printf("v_fma_f64          %8.2f Gop/s  (%.1f TFLOP/s)\n", n * iters * 8 / ms * 1e-6, 2 * n * iters * 8 / ms * 1e-9);
    ms = time_ms([&] { r_sqrtdiv64<<<blocks, threads>>>(o64, iters / 10); }, 5);
    std::printf("f64 sqrt + div     %8.2f G/s\n", n * (iters / 10) * 8 / ms * 1e-6);
    CK(hipFree(o32));
    CK(hipFree(o64));
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
