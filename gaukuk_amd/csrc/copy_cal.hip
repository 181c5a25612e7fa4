// copy_cal.hip -- counter calibration for HBM traffic figures.
//
// Copies 32 Mi doubles (256 MiB) three times.  Profiling it with rocprofv3 --pmc FETCH_SIZE and
// --pmc WRITE_SIZE gives what each counter reports for a known 256 MiB read and 256 MiB write;
// the ratio scales the same counters measured on the stage kernels into bytes.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

int main() {
    const long n = 32L << 20;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) return 1;
    if (hipMemset(a, 0, n * 8) != hipSuccess || hipMemset(b, 0, n * 8) != hipSuccess) return 1;
    for (int k = 0; k < 3; k++) hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 2;
    std::printf("copy_kernel: %ld bytes read and %ld bytes written per launch\n", n * 8, n * 8);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
