// HBM counter calibration: a copy kernel of known traffic.  Each of its three launches
// reads 1 GiB and writes 1 GiB (16-byte vector loads, non-temporal stores), so
// rocprofv3's FETCH_SIZE and WRITE_SIZE per dispatch, divided into 1 GiB, give the
// factors that turn those counters into bytes for the other kernels.
#include <hip/hip_runtime.h>
#include <stdio.h>

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void copy4(const f4v* __restrict__ a, f4v* __restrict__ b, size_t n4) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += st) __builtin_nontemporal_store(a[i], b + i);
}

int main() {
    const size_t bytes = (size_t)1 << 30;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 3; r++)
        hipLaunchKernelGGL(copy4, dim3(8192), dim3(256), 0, 0, (const f4v*)a, (f4v*)b, bytes / 16);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    printf("copy4: 3 launches, each reads %zu and writes %zu bytes\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
