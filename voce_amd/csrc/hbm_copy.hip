// hbm_copy -- calibration kernel for HBM traffic counters.
//
// Streams `n` doubles from one buffer to another, `reps` times: exactly
// 8n bytes read and 8n bytes written per launch.  Profiling it with the
// FETCH_SIZE / WRITE_SIZE counters gives the factor that turns those
// counters into bytes for the solver's kernels.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : (32L << 20);   // 256 MiB of doubles
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    double *a, *b;
    if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) return 1;
    if (hipMemset(a, 0, n * 8) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    printf("hbm_copy: %ld doubles, read %.1f MiB + write %.1f MiB per launch, best %.4f ms, %.1f GB/s\n", n,
           n * 8 / 1048576.0, n * 8 / 1048576.0, best, 2.0 * n * 8 / (best * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
