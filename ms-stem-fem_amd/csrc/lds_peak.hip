// LDS rate probe for the LDS roofline of the slice passes: every lane streams
// 16-byte LDS reads (or writes), consecutive lanes on consecutive addresses
// (no bank conflicts), four 32 KiB workgroups per CU.  Prints the rate the
// hardware sustains, in TB/s and bytes per clock per CU.
//   make -C ms-stem-fem_amd/csrc lds_peak && ms-stem-fem_amd/lds_peak
#include <hip/hip_runtime.h>
#include <cstdio>

constexpr int N = 2048; // float4 per workgroup: 32 KiB

template <bool WRITE>
__global__ void __launch_bounds__(256) lds_peak(float4* out, int iters)
{
    __shared__ float4 s[N];
    const int t = threadIdx.x;
    for (int e = t; e < N; e += 256) s[e] = make_float4((float)e, 1.f, 2.f, 3.f);
    __syncthreads();
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < iters; ++k) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int a = (t + j * 256 + k * 64) & (N - 1);
            if constexpr (WRITE) {
                s[a] = make_float4(acc.x + j, acc.y, acc.z, acc.w);
            } else {
                const float4 v = s[a];
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            }
        }
        if constexpr (WRITE) acc.x += 1.f;
    }
    __syncthreads();
    if constexpr (WRITE) acc.y += s[t].x;
    out[(size_t)blockIdx.x * 256 + t] = acc;
}

template <bool W> double run(float4* out, int iters, int blocks)
{
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL(lds_peak<W>, dim3(blocks), dim3(256), 0, 0, out, 10);
    (void)hipDeviceSynchronize();
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(a, 0);
        hipLaunchKernelGGL(lds_peak<W>, dim3(blocks), dim3(256), 0, 0, out, iters);
        (void)hipEventRecord(b, 0);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best) best = ms;
    }
    const double bytes = (double)blocks * 256 * 16 * 8 * iters;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return bytes / (best * 1e-3) / 1e12;
}

int main()
{
    int dev = 0, cus = 0, clk = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev) != hipSuccess) // kHz
        return 1;
    const int blocks = cus * 4;
    float4* out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float4)) != hipSuccess) return 1;
    const int iters = 4000;
    const double rd = run<false>(out, iters, blocks);
    const double wr = run<true>(out, iters, blocks);
    printf("CUs %d clock %.0f MHz\n", cus, clk / 1e3);
    printf("LDS read  %.2f TB/s = %.1f B/clk/CU\n", rd, rd * 1e12 / (cus * clk * 1e3));
    printf("LDS write %.2f TB/s = %.1f B/clk/CU\n", wr, wr * 1e12 / (cus * clk * 1e3));
    (void)hipFree(out);
    return 0;
}
