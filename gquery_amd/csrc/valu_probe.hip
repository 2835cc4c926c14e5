// VALU issue-rate probe for the roofline of the traversal kernels (a stand-alone program, not part
// of the library): wave64 VALU instructions per second for a pure v_fma_f32 stream and for a
// v_sub_f32 / v_max_f32 stream, at 32 wavefronts per CU (4 rounds of them).  Each thread runs ITERS rounds of 8
// independent chains and stores one float to out[gid] (gid < grid size = allocation size).
#include <hip/hip_runtime.h>
#include <cstdio>
constexpr int ITERS = 4096;
__global__ void __launch_bounds__(256) k_fma(float *out, float a, float b) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float r0 = gid * 1e-7f, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
    for (int i = 0; i < ITERS; ++i) {
        r0 = fmaf(r0, a, b); r1 = fmaf(r1, a, b); r2 = fmaf(r2, a, b); r3 = fmaf(r3, a, b);
        r4 = fmaf(r4, a, b); r5 = fmaf(r5, a, b); r6 = fmaf(r6, a, b); r7 = fmaf(r7, a, b);
    }
    out[gid] = r0 + r1 + r2 + r3 + r4 + r5 + r6 + r7;
}
__global__ void __launch_bounds__(256) k_submax(float *out, float a, float b) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float r0 = gid * 1e-7f, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
    for (int i = 0; i < ITERS; ++i) {
        r0 = fmaxf(r0 - a, b); r1 = fmaxf(r1 - a, b); r2 = fmaxf(r2 - a, b); r3 = fmaxf(r3 - a, b);
        r4 = fmaxf(r4 - a, b); r5 = fmaxf(r5 - a, b); r6 = fmaxf(r6 - a, b); r7 = fmaxf(r7 - a, b);
    }
    out[gid] = r0 + r1 + r2 + r3 + r4 + r5 + r6 + r7;
}
int main() {
    int ncu = 0, khz = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, 0) != hipSuccess) return 1;
    const int blocks = ncu * 32;   // 4 rounds of 8 blocks of 4 wavefronts per CU
    float *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int kind = 0; kind < 2; ++kind) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            hipEventRecord(e0);
            if (kind == 0) hipLaunchKernelGGL(k_fma, dim3(blocks), dim3(256), 0, 0, out, 0.999f, 0.001f);
            else hipLaunchKernelGGL(k_submax, dim3(blocks), dim3(256), 0, 0, out, 0.001f, -1.0f);
            hipEventRecord(e1);
            hipEventSynchronize(e1);
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            best = ms < best ? ms : best;
        }
        const double instr = (double)blocks * 4 * ITERS * 8 * (kind == 0 ? 1 : 2);   // loop-body wave64 instructions
        const double rate = instr / (best * 1e-3);
        printf("{\"stream\": \"%s\", \"ms\": %.4f, \"wave64_instr_per_s\": %.4e, \"per_simd_per_clock\": %.3f, \"cus\": %d, \"clock_khz\": %d}\n",
               kind == 0 ? "v_fma_f32" : "v_sub_f32+v_max_f32", best, rate, rate / (ncu * 4.0 * khz * 1e3), ncu, khz);
    }
    hipFree(out);
    return 0;
}
