// Measures how VALU and MFMA work share a SIMD on this GPU: the issue costs
// behind the gridder and degridder roofline figures.  Each kernel runs a loop
// of NM independent v_mfma_f32_16x16x32_bf16 (8 accumulator chains) and NV
// independent FP32 FMAs (16 chains) per iteration, on one or two waves per
// SIMD (256-thread workgroups, one or two per CU); every thread writes one
// float.  Reported per case: wall time, shader clocks per loop iteration of
// one wave (s_memtime), and the effective clock.
//
//   make -C p9-distributed-reconstruction_amd probe && p9-distributed-reconstruction_amd/probe_mfma_valu
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kIters = 4096;

#define CHECK(call)                                                                   \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                \
            exit(1);                                                                  \
        }                                                                             \
    } while (0)

// clocks[0..1]: s_memtime and s_memrealtime deltas of workgroup 0, wave 0
template <int NM, int NV, bool INTERLEAVE>
__device__ __forceinline__ float body(int tid, unsigned long long* clocks) {
    f32x4 acc[8];
    for (int i = 0; i < 8; i++) acc[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
    bf16x8 a, b;
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16)(float)(tid + i);
        b[i] = (__bf16)(float)(i - tid);
    }
    float x[16];
    for (int i = 0; i < 16; i++) x[i] = (float)(tid * 16 + i) * 1e-3f;
    const float m = 0.999f, c = 1e-4f;
    const unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0; it < kIters; it++) {
#pragma unroll
        for (int k = 0; k < NM; k++) acc[k & 7] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k & 7], 0, 0, 0);
#pragma unroll
        for (int k = 0; k < NV; k++) x[k & 15] = __builtin_fmaf(x[k & 15], m, c);
        if constexpr (INTERLEAVE) {
            // one MFMA, then NV / NM FMAs, repeated
#pragma unroll
            for (int k = 0; k < NM; k++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, NV / NM, 0);
            }
        }
    }
    const unsigned long long t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clocks[0] = t1 - t0;
        clocks[1] = r1 - r0;
    }
    float s = 0.f;
    for (int i = 0; i < 8; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    for (int i = 0; i < 16; i++) s += x[i];
    return s;
}

template <int NM, int NV, bool INTERLEAVE, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void same_kernel(
    float* out, unsigned long long* clocks) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    out[tid] = body<NM, NV, INTERLEAVE>(tid, clocks);
}

// even waves: NM MFMAs per iteration only; odd waves: NV FMAs only
template <int NM, int NV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void split_kernel(
    float* out, unsigned long long* clocks) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    out[tid] = ((threadIdx.x >> 6) & 1) ? body<0, NV, false>(tid, clocks) : body<NM, 0, false>(tid, clocks);
}

typedef void (*Kernel)(float*, unsigned long long*);

static void run(Kernel k, int blocks, float* d_out, unsigned long long* d_clk, const char* name, int mfma, int valu) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, d_out, d_clk);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0, 0));
    constexpr int kReps = 5;
    for (int r = 0; r < kReps; r++) hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, d_out, d_clk);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long clk[2];
    CHECK(hipMemcpy(clk, d_clk, sizeof(clk), hipMemcpyDeviceToHost));
    const double cyc = (double)clk[0] / kIters;               // shader clocks per iteration, one wave
    const double ghz = clk[1] ? (double)clk[0] / (clk[1] * 10.0) : 0.0;  // s_memrealtime runs at 100 MHz
    printf("| %-44s | %2d | %3d | %7.3f | %7.1f | %5.2f |\n", name, mfma, valu, ms / kReps, cyc, ghz);
    fflush(stdout);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main() {
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, dev));
    printf("device %s (%s), %d CUs, %d iterations per wave, 256-thread workgroups\n", prop.name, prop.gcnArchName, cus,
           kIters);
    printf("| case | MFMA/it | FMA/it | ms | clocks/it (one wave) | GHz |\n|---|---|---|---|---|---|\n");
    const int one = cus, two = 2 * cus;  // one or two waves per SIMD
    float* d_out = nullptr;
    unsigned long long* d_clk = nullptr;
    CHECK(hipMalloc(&d_out, (size_t)two * 256 * sizeof(float)));
    CHECK(hipMalloc(&d_clk, 2 * sizeof(unsigned long long)));
    run(same_kernel<16, 0, false, 1>, one, d_out, d_clk, "MFMA only, 1 wave/SIMD", 16, 0);
    run(same_kernel<0, 64, false, 1>, one, d_out, d_clk, "FMA only, 1 wave/SIMD", 0, 64);
    run(same_kernel<16, 64, false, 1>, one, d_out, d_clk, "both, compiler schedule, 1 wave/SIMD", 16, 64);
    run(same_kernel<16, 64, true, 1>, one, d_out, d_clk, "both, 1 MFMA : 4 FMA interleave, 1 wave/SIMD", 16, 64);
    run(same_kernel<16, 0, false, 2>, two, d_out, d_clk, "MFMA only, 2 waves/SIMD", 16, 0);
    run(same_kernel<0, 64, false, 2>, two, d_out, d_clk, "FMA only, 2 waves/SIMD", 0, 64);
    run(same_kernel<16, 64, false, 2>, two, d_out, d_clk, "both, compiler schedule, 2 waves/SIMD", 16, 64);
    run(same_kernel<16, 64, true, 2>, two, d_out, d_clk, "both, 1 MFMA : 4 FMA interleave, 2 waves/SIMD", 16, 64);
    run(split_kernel<32, 128>, two, d_out, d_clk, "MFMA wave + FMA wave on one SIMD", 32, 128);
    CHECK(hipFree(d_out));
    CHECK(hipFree(d_clk));
    return 0;
}
