// copycal.hip — moves a known number of bytes, so rocprofv3's FETCH_SIZE /
// WRITE_SIZE (KiB per dispatch) can be scaled to bytes:
//     rocprofv3 --pmc FETCH_SIZE -- ./copycal 268435456
// copy8 / copy16 stream the buffer with 8- and 16-byte lanes; gather8 reads
// 8-byte elements at random positions of a 1 GiB array (four times the
// 256 MB last-level cache, so nearly every gather misses it) and writes them
// out contiguously: per element 8 bytes read at random and 8 bytes written;
// gather8_local does the same with each index at most 4,096 elements from
// its own position (the locality of a renumbered slot map).
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

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void gather8(const float2* __restrict__ a, size_t n_src, float2* __restrict__ b,
                                               size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) b[i] = a[mix64(i) % n_src];
}

__global__ __launch_bounds__(256) void gather8_local(const float2* __restrict__ a, size_t n_src,
                                                     float2* __restrict__ b, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) {
        const size_t j = (i + (mix64(i) & 8191u) + n_src - 4096u) % n_src;
        b[i] = a[j];
    }
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
    // gathers: n8 elements from a 1 GiB source
    const size_t src_bytes = 1ull << 30;
    void* src = nullptr;
    if (hipMalloc(&src, src_bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(src, 1, src_bytes);
    const size_t n_src = src_bytes / 8;
    hipLaunchKernelGGL(gather8, dim3((n8 + 255) / 256), dim3(256), 0, 0, (const float2*)src, n_src, (float2*)b, n8);
    hipLaunchKernelGGL(gather8_local, dim3((n8 + 255) / 256), dim3(256), 0, 0, (const float2*)src, n_src,
                       (float2*)b, n8);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "gather failed\n");
        return 1;
    }
    std::printf("gathered %zu 8-byte elements (%zu bytes) from a %zu-byte array, random and local\n", n8, bytes,
                src_bytes);
    (void)hipFree(src);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
