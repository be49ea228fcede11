// Measures how VALU and MFMA work share a SIMD on this GPU: the issue costs
// behind the gridder and degridder roofline figures.  Each kernel runs a loop
// of NM independent v_mfma_f32_16x16x32_bf16 (8 accumulator chains) and NV
// independent VALU instructions of one kind (16 chains) per iteration, on one
// or two waves per SIMD (256-thread workgroups, one or two per CU); every
// thread writes one float.  The VALU kinds are the gridder's: FP32 FMA,
// v_cvt_pk_bf16_f32 (operand packing), v_dot2c_f32_bf16 (the hi/lo split
// residual), FP64 FMA (phase coefficients) and v_sin_f32.  Reported per
// case: wall time, shader clocks per loop iteration of one wave (s_memtime),
// and the effective clock.
//
//   make -C p9-distributed-reconstruction_amd probe && p9-distributed-reconstruction_amd/probe_mfma_valu
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

enum Op { kFma, kCvtPk, kDot2, kFma64, kSin };

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
template <int NM, int NV, bool INTERLEAVE, int OP = kFma>
__device__ __forceinline__ float body(int tid, unsigned long long* clocks) {
    f32x4 acc[8];
    for (int i = 0; i < 8; i++) acc[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
    bf16x8 a, b;
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16)(float)(tid + i);
        b[i] = (__bf16)(float)(i - tid);
    }
    float x[16];
    double xd[16];
    for (int i = 0; i < 16; i++) x[i] = (float)(tid * 16 + i) * 1e-3f;
    for (int i = 0; i < 16; i++) xd[i] = (double)x[i];
    const float m = 0.999f, c = 1e-4f;
    const bf16x2 h2 = {(__bf16)0.5f, (__bf16)-1.0f};
    const unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0; it < kIters; it++) {
#pragma unroll
        for (int k = 0; k < NM; k++) acc[k & 7] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k & 7], 0, 0, 0);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            float& v = x[k & 15];
            if constexpr (OP == kFma) v = __builtin_fmaf(v, m, c);
            if constexpr (OP == kCvtPk) v = __builtin_bit_cast(float, __builtin_convertvector((f32x2){v, c}, bf16x2));
            if constexpr (OP == kDot2) v = __builtin_amdgcn_fdot2_f32_bf16(h2, h2, v, false);
            if constexpr (OP == kFma64) xd[k & 15] = __builtin_fma(xd[k & 15], (double)m, (double)c);
            if constexpr (OP == kSin) v = __builtin_amdgcn_sinf(v);
        }
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
    for (int i = 0; i < 16; i++) s += x[i] + (float)xd[i];
    return s;
}

template <int NM, int NV, bool INTERLEAVE, int W, int OP = kFma>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void same_kernel(
    float* out, unsigned long long* clocks) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    out[tid] = body<NM, NV, INTERLEAVE, OP>(tid, clocks);
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
    printf("| case | MFMA/it | VALU/it | ms | clocks/it (one wave) | GHz |\n|---|---|---|---|---|---|\n");
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
    // issue cost of the other VALU kinds alone
    run(same_kernel<0, 64, false, 1, kCvtPk>, one, d_out, d_clk, "v_cvt_pk_bf16_f32 only, 1 wave/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 2, kCvtPk>, two, d_out, d_clk, "v_cvt_pk_bf16_f32 only, 2 waves/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 1, kDot2>, one, d_out, d_clk, "v_dot2c_f32_bf16 only, 1 wave/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 2, kDot2>, two, d_out, d_clk, "v_dot2c_f32_bf16 only, 2 waves/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 1, kFma64>, one, d_out, d_clk, "FP64 FMA only, 1 wave/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 2, kFma64>, two, d_out, d_clk, "FP64 FMA only, 2 waves/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 1, kSin>, one, d_out, d_clk, "v_sin_f32 only, 1 wave/SIMD", 0, 64);
    run(same_kernel<0, 64, false, 2, kSin>, two, d_out, d_clk, "v_sin_f32 only, 2 waves/SIMD", 0, 64);
    CHECK(hipFree(d_out));
    CHECK(hipFree(d_clk));
    return 0;
}
