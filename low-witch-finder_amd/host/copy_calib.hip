// copy_calib — a plain device-to-device copy of a known size, for scaling rocprofv3's FETCH_SIZE and WRITE_SIZE.
//
//   copy_calib [MiB]      copies MiB mebibytes (default 256) with one 16-byte load and store per lane, 5 times
//
// Each dispatch reads and writes exactly MiB MiB, so FETCH_SIZE / WRITE_SIZE of one dispatch divided by that
// size is the counter's scale on this device (the factor used to turn the scan kernels' counters into bytes).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? (size_t)std::atol(argv[1]) : 256;
    const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) { std::fprintf(stderr, "alloc\n"); return 1; }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 5; r++) hipLaunchKernelGGL(copy16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) { std::fprintf(stderr, "copy failed\n"); return 1; }
    std::printf("copy_calib: %zu MiB per dispatch (%zu bytes read, %zu bytes written), 5 dispatches\n", mib, bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
