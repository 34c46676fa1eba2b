// copycal.hip — copies a known number of bytes with 8-byte and with 16-byte
// lanes, so rocprofv3's FETCH_SIZE / WRITE_SIZE (KiB per dispatch) can be
// scaled to bytes for kernels with either access width:
//     rocprofv3 --pmc FETCH_SIZE -- ./copycal 268435456
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy8(const float2* __restrict__ a, float2* __restrict__ b, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void copy16(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int main(int argc, char** argv) {
    const size_t bytes = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (1ull << 28);
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, bytes);
    const size_t n8 = bytes / 8, n16 = bytes / 16;
    hipLaunchKernelGGL(copy8, dim3((n8 + 255) / 256), dim3(256), 0, 0, (const float2*)a, (float2*)b, n8);
    hipLaunchKernelGGL(copy16, dim3((n16 + 255) / 256), dim3(256), 0, 0, (const float4*)a, (float4*)b, n16);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "copy failed\n");
        return 1;
    }
    std::printf("copied %zu bytes with 8-byte and 16-byte lanes\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
