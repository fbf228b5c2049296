// valu_rate.hip -- VALU issue rate of wave64 FP64 and FP32 arithmetic on the
// device: the ceiling against which the stage kernels' instruction counts
// are judged (their FP64 work is VALU-issue bound, not HBM bound).
//
// Each thread runs 16 independent multiply-add chains; the grid holds w waves
// per SIMD (4 waves per 256-thread block, one block per SIMD group of the CU
// count read from the device).  Reports cycles per wave-instruction per SIMD
// at the device's peak engine clock.  Built by grhd_amd/build.py into
// grhd_amd/bin/valu_rate; `python bench.py --valu-rate` runs it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(call)                                                                   \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));         \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

constexpr int kChains = 16;

// OP 0: x = x * a + b (fused), 1: x = x + a, 2: x = x * a
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_chains(T *out, T a, T b, int iters)
{
    T x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; c++) x[c] = (T)(threadIdx.x + c);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) {
            if (OP == 0) {
                if constexpr (sizeof(T) == sizeof(double)) x[c] = __builtin_fma(x[c], a, b);
                else x[c] = __builtin_fmaf(x[c], a, b);
            }
            else if (OP == 1) x[c] = x[c] + a;
            else x[c] = x[c] * a;
        }
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < kChains; c++) s += x[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <typename T, int OP>
void run(const char *name, int cus, int waves_per_simd, double ghz)
{
    const int blocks = cus * waves_per_simd;  // a 256-thread block is one wave on each SIMD of a CU
    const int iters = 4096;
    T *out = nullptr;
    CHK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(T)));
    hipLaunchKernelGGL((k_chains<T, OP>), dim3(blocks), dim3(256), 0, 0, out, (T)1.0000001, (T)1e-7, 16);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        CHK(hipEventRecord(e0));
        hipLaunchKernelGGL((k_chains<T, OP>), dim3(blocks), dim3(256), 0, 0, out, (T)1.0000001, (T)1e-7, iters);
        CHK(hipEventRecord(e1));
        CHK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    const double winst = (double)waves_per_simd * iters * kChains;  // wave-instructions per SIMD
    const double cyc = best * 1e-3 * ghz * 1e9;
    std::printf("{\"op\": \"%s\", \"waves_per_simd\": %d, \"ms\": %.4f, \"cycles_per_wave_instruction\": %.3f}\n",
                name, waves_per_simd, best, cyc / winst);
    CHK(hipEventDestroy(e0));
    CHK(hipEventDestroy(e1));
    CHK(hipFree(out));
}

int main()
{
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const double ghz = prop.clockRate * 1e-6;  // kHz -> GHz (peak engine clock)
    std::printf("{\"device\": \"%s\", \"cus\": %d, \"peak_clock_ghz\": %.3f}\n", prop.gcnArchName, cus, ghz);
    for (int w : {1, 2, 4, 8}) {
        run<double, 0>("f64_fma", cus, w, ghz);
        run<float, 0>("f32_fma", cus, w, ghz);
    }
    run<double, 1>("f64_add", cus, 8, ghz);
    run<double, 2>("f64_mul", cus, 8, ghz);
    run<float, 1>("f32_add", cus, 8, ghz);
    return 0;
}
