// copy_calib.hip -- the calibration kernel of the HBM-traffic figures.
//
// A plain float4 copy of known size: 256 MiB read and 256 MiB written per
// launch, three launches.  rocprofv3 --pmc FETCH_SIZE (one run) and
// --pmc WRITE_SIZE (another run) over this program give each counter's reading
// for a known number of bytes; bench.py scales the same counters of the hot
// kernels by those readings to report roofline.traffic.  Measurement only: no
// part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(call)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));               \
            return 1;                                                                     \
        }                                                                                 \
    } while (0)

__global__ void copy_kernel(const float4* __restrict__ a, float4* __restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main()
{
    const size_t bytes = (size_t)256 << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
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
