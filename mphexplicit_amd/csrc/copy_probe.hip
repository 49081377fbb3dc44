// copy_probe: a plain device-to-device copy of a known number of bytes, repeated.  Used
// only to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) against
// bytes actually moved:  copy_probe [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const double4* __restrict__ a, double4* __restrict__ b, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main(int argc, char** argv)
{
    const long mib = argc > 1 ? atol(argv[1]) : 512;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const long n = mib * 1024 * 1024 / (long)sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * sizeof(double4)) != hipSuccess || hipMalloc((void**)&b, n * sizeof(double4)) != hipSuccess) {
        std::fprintf(stderr, "allocation failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, n * sizeof(double4));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        std::printf("copy %ld MiB: read %ld + write %ld bytes in %.3f ms = %.1f GB/s\n", mib,
                    n * (long)sizeof(double4), n * (long)sizeof(double4), ms,
                    2.0 * n * sizeof(double4) / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
