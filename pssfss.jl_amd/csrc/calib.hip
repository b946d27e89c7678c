// Kernels of known work for calibrating rocprofv3 counters (not part of the
// MoM path; built into libpssfss_calib.so, driven by bench.py --calibrate):
//   pssfss_calib_copy: a device copy of `bytes` each way, `reps` times, so
//     FETCH_SIZE / WRITE_SIZE per dispatch can be compared with the bytes moved;
//   pssfss_calib_fma: FP64 and FP32 FMA loops of 2 * blocks * 256 * iters flops,
//     for SQ_INSTS_VALU_FLOPS_FP64 / _FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ void k_calib_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) b[i] = a[i];
}

__global__ void k_calib_fill(float4 *a, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) a[i] = float4{1.f, 1.f, 1.f, 1.f};
}

template <class T>
__global__ void k_calib_fma(T *out, int iters) {
    T a = (T)threadIdx.x * (T)1e-3, b = (T)0.999, c = (T)1e-6;
    for (int i = 0; i < iters; ++i) a = a * b + c;
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = a;
}

}  // namespace

extern "C" int pssfss_calib_copy(int64_t bytes, int reps) {
    if (bytes <= 0 || bytes % 16 || reps <= 0) return -1;
    const size_t n = (size_t)bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess) return -2;
    if (hipMalloc((void **)&b, bytes) != hipSuccess) { (void)hipFree(a); return -2; }
    hipLaunchKernelGGL(k_calib_fill, dim3(8192), dim3(256), 0, 0, a, n);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_calib_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
    const int ok = hipDeviceSynchronize() == hipSuccess;
    (void)hipFree(a);
    (void)hipFree(b);
    return ok ? 0 : -3;
}

extern "C" int pssfss_calib_fma(int blocks, int iters) {
    if (blocks <= 0 || iters <= 0) return -1;
    const size_t n = (size_t)blocks * 256;
    double *d = nullptr;
    float *f = nullptr;
    if (hipMalloc((void **)&d, n * sizeof(double)) != hipSuccess) return -2;
    if (hipMalloc((void **)&f, n * sizeof(float)) != hipSuccess) { (void)hipFree(d); return -2; }
    hipLaunchKernelGGL(k_calib_fma<double>, dim3(blocks), dim3(256), 0, 0, d, iters);
    hipLaunchKernelGGL(k_calib_fma<float>, dim3(blocks), dim3(256), 0, 0, f, iters);
    const int ok = hipDeviceSynchronize() == hipSuccess;
    (void)hipFree(d);
    (void)hipFree(f);
    return ok ? 0 : -3;
}
