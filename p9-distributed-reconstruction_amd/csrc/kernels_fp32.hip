// FP32 vector-ALU gridder and degridder, each fused with the LDS subgrid FFT
// and the adder (gridder) or splitter (degridder).  One 256-thread workgroup
// per subgrid; workgroups are issued longest subgrid first.
//
// Gridder:   Gridder.Forward (Gridder.cs:15-87) -> SubgridFFT.Forward
//            (SubgridFFT.cs:11-52) -> Adder.Add / AdderWStack.Add
//            (Adder.cs:13-55, AdderWStack.cs:11-78) -> FFT.Shift (FFT.cs:221-264)
// Degridder: FFT.Shift -> Adder.Split / AdderWStack.Split (Adder.cs:57-95,
//            AdderWStack.cs:81-139) -> SubgridFFT.Backward (SubgridFFT.cs:54-90)
//            -> Gridder.Backwards (Gridder.cs:91-160)
// The grid's quadrant swap (FFT.Shift) is folded into the adder/splitter
// addressing, so no separate shift pass touches the grid.
#include "device_addsplit.h"

namespace p9idg {
namespace dev {

// Visibilities [k0, k1) of subgrid s (k1 < 0: all of them).
template <int S>
__device__ __forceinline__ void grid_fp32_body(const KernelArgs& a, int s, int k0, int k1, int slot) {
    constexpr int NPIX = (S * S + kThreads - 1) / kThreads;
    constexpr int CHUNK = kThreads;
    __shared__ float4 s_coef[CHUNK];
    __shared__ float2 s_vis[CHUNK];
    __shared__ float2 s_tile[2][S * S];

    const SubgridGeom g = load_geom(a, s, S);
    const int tid = threadIdx.x;

    float pl[NPIX], pm[NPIX], pn[NPIX];
    float2 acc[NPIX];
#pragma unroll
    for (int i = 0; i < NPIX; i++) {
        const int p = tid + i * kThreads;
        const int y = p / S, x = p % S;
        pl[i] = compute_l(x, S, a.image_size);
        pm[i] = compute_l(y, S, a.image_size);
        pn[i] = compute_n(pl[i], pm[i]);
        acc[i] = make_float2(0.f, 0.f);
    }

    const int C = a.C;
    const int K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    for (int j0 = k0; j0 < K; j0 += CHUNK) {
        const int j = j0 + tid;
        if (j < K) {
            const int t = g.t0 + j / C, ch = j % C;
            const float3 cf = vis_coeffs(a, g, t, ch);
            const float2 v = a.vis[vbase + j];
            s_coef[tid] = make_float4(cf.x, cf.y, cf.z, 0.f);
            s_vis[tid] = v;
        }
        __syncthreads();
        const int nv = min(CHUNK, K - j0);
        for (int v = 0; v < nv; v++) {
            const float4 cf = s_coef[v];
            const float2 vv = s_vis[v];
#pragma unroll
            for (int i = 0; i < NPIX; i++) {
                const float ph = fmaf(cf.x, pl[i], fmaf(cf.y, pm[i], cf.z * pn[i]));
                const float2 e = cis_turns(ph);
                acc[i].x = fmaf(vv.x, e.x, fmaf(-vv.y, e.y, acc[i].x));
                acc[i].y = fmaf(vv.x, e.y, fmaf(vv.y, e.x, acc[i].y));
            }
        }
        __syncthreads();
    }

    // Gridder.cs:76-79: taper (identity) and ifftshift into the tile
#pragma unroll
    for (int i = 0; i < NPIX; i++) {
        const int p = tid + i * kThreads;
        if (p < S * S) {
            const int y = p / S, x = p % S;
            s_tile[0][((y + S / 2) % S) * S + (x + S / 2) % S] = acc[i];
        }
    }
    __syncthreads();
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, +1);
    adder<S>(a, g, res, 1.0f / (float)(S * S), slot);
}

template <int S>
__device__ __forceinline__ void degrid_fp32_body(const KernelArgs& a, int s, int k0, int k1) {
    __shared__ float2 s_tile[2][S * S];
    __shared__ float s_n[S * S];
    __shared__ float s_l[S];

    const SubgridGeom g = load_geom(a, s, S);
    const int tid = threadIdx.x;

    splitter<S>(a, g, s_tile[0]);
    for (int p = tid; p < S * S; p += kThreads) {
        const int y = p / S, x = p % S;
        s_n[p] = compute_n(compute_l(x, S, a.image_size), compute_l(y, S, a.image_size));
    }
    if (tid < S) s_l[tid] = compute_l(tid, S, a.image_size);
    __syncthreads();
    float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, -1);
    float2* px = (res == s_tile[0]) ? s_tile[1] : s_tile[0];
    // Gridder.cs:101-117: pixels_copy[y][x] = taper * subgrid[ifftshift]
    for (int p = tid; p < S * S; p += kThreads) {
        const int y = p / S, x = p % S;
        px[p] = res[((y + S / 2) % S) * S + (x + S / 2) % S];
    }
    __syncthreads();

    const int C = a.C;
    const int K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    for (int j = k0 + tid; j < K; j += kThreads) {
        const int t = g.t0 + j / C, ch = j % C;
        const float3 cf = vis_coeffs(a, g, t, ch);
        float sx = 0.f, sy = 0.f;
#pragma unroll 1
        for (int y = 0; y < S; y++) {
            const float bm = cf.y * s_l[y];  // m(y) = l(y)
#pragma unroll 8
            for (int x = 0; x < S; x++) {
                const float ph = fmaf(cf.x, s_l[x], fmaf(cf.z, s_n[y * S + x], bm));
                const float2 e = cis_turns(ph);
                const float2 q = px[y * S + x];
                // pixel * exp(-i phase)  (Gridder.cs:145-147)
                sx = fmaf(q.x, e.x, fmaf(q.y, e.y, sx));
                sy = fmaf(q.y, e.x, fmaf(-q.x, e.y, sy));
            }
        }
        a.vis_out[vbase + j] = make_float2(sx * a.degrid_scale, sy * a.degrid_scale);
    }
}

template <int S>
__global__ __launch_bounds__(kThreads) void grid_fp32_kernel(KernelArgs a) {
    grid_fp32_body<S>(a, a.order[blockIdx.x], 0, -1, a.order[blockIdx.x]);
}
template <int S>
__global__ __launch_bounds__(kThreads) void grid_fp32_queue_kernel(KernelArgs a) {
    run_queue(a, [&](int s, int k0, int k1, int slot) { grid_fp32_body<S>(a, s, k0, k1, slot); });
}
template <int S>
__global__ __launch_bounds__(kThreads) void degrid_fp32_kernel(KernelArgs a) {
    degrid_fp32_body<S>(a, a.order[blockIdx.x], 0, -1);
}
template <int S>
__global__ __launch_bounds__(kThreads) void degrid_fp32_queue_kernel(KernelArgs a) {
    run_queue(a, [&](int s, int k0, int k1, int) { degrid_fp32_body<S>(a, s, k0, k1); });
}

template <int S>
__global__ __launch_bounds__(kThreads) void grid_fp32_batch_kernel(BatchArgs b) {
    run_batch_queue(b, [&](const KernelArgs& a, int s, int k0, int k1, int slot) { grid_fp32_body<S>(a, s, k0, k1, slot); });
}
template <int S>
__global__ __launch_bounds__(kThreads) void degrid_fp32_batch_kernel(BatchArgs b) {
    run_batch_queue(b, [&](const KernelArgs& a, int s, int k0, int k1, int) { degrid_fp32_body<S>(a, s, k0, k1); });
}

__global__ void reset_counter_kernel(int32_t* counter) { *counter = 0; }

}  // namespace dev

// Launches KERNEL<S> (one workgroup per subgrid) or QKERNEL<S> (persistent
// queue over the work items).
#define P9_LAUNCH_FP32(KERNEL, QKERNEL, SV)                                                   \
    do {                                                                                      \
        if (cfg.queue) {                                                                      \
            int wg = cfg.workgroups > 0 ? cfg.workgroups                                      \
                                        : persistent_workgroups((const void*)QKERNEL<SV>, dev::kThreads); \
            wg = wg < a.nitems ? wg : a.nitems;                                               \
            if (wg <= 0) return hipSuccess;                                                   \
            hipLaunchKernelGGL(dev::reset_counter_kernel, dim3(1), dim3(1), 0, st, a.counter); \
            hipLaunchKernelGGL(QKERNEL<SV>, dim3(wg), dim3(dev::kThreads), 0, st, a);         \
        } else {                                                                              \
            hipLaunchKernelGGL(KERNEL<SV>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); \
        }                                                                                     \
    } while (0)

#define P9_SWITCH_S(S, BODY_FOR)                  \
    switch (S) {                                  \
        case 8: BODY_FOR(8); break;               \
        case 16: BODY_FOR(16); break;             \
        case 24: BODY_FOR(24); break;             \
        case 32: BODY_FOR(32); break;             \
        case 48: BODY_FOR(48); break;             \
        case 64: BODY_FOR(64); break;             \
        default: return hipErrorInvalidValue;     \
    }

hipError_t launch_grid_fp32(const KernelArgs& a, int S, const LaunchCfg& cfg, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
#define P9_G(SV) P9_LAUNCH_FP32(dev::grid_fp32_kernel, dev::grid_fp32_queue_kernel, SV)
    P9_SWITCH_S(S, P9_G)
#undef P9_G
    return hipGetLastError();
}

hipError_t launch_degrid_fp32(const KernelArgs& a, int S, const LaunchCfg& cfg, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
#define P9_D(SV) P9_LAUNCH_FP32(dev::degrid_fp32_kernel, dev::degrid_fp32_queue_kernel, SV)
    P9_SWITCH_S(S, P9_D)
#undef P9_D
    return hipGetLastError();
}

hipError_t launch_fp32_batch(bool grid, const BatchArgs& b, int S, int workgroups, hipStream_t st) {
#define P9_B(SV)                                                                                   \
    do {                                                                                           \
        const void* k = grid ? (const void*)dev::grid_fp32_batch_kernel<SV>                        \
                             : (const void*)dev::degrid_fp32_batch_kernel<SV>;                     \
        int wg = workgroups > 0 ? workgroups : persistent_workgroups(k, dev::kThreads);            \
        wg = wg < b.nitems ? wg : b.nitems;                                                        \
        if (wg <= 0) return hipSuccess;                                                            \
        hipLaunchKernelGGL(dev::reset_counter_kernel, dim3(1), dim3(1), 0, st, b.counter);         \
        if (grid)                                                                                  \
            hipLaunchKernelGGL(dev::grid_fp32_batch_kernel<SV>, dim3(wg), dim3(dev::kThreads), 0, st, b); \
        else                                                                                       \
            hipLaunchKernelGGL(dev::degrid_fp32_batch_kernel<SV>, dim3(wg), dim3(dev::kThreads), 0, st, b); \
    } while (0)
    P9_SWITCH_S(S, P9_B)
#undef P9_B
    return hipGetLastError();
}

hipError_t launch_reset_counter(int32_t* counter, hipStream_t st) {
    hipLaunchKernelGGL(dev::reset_counter_kernel, dim3(1), dim3(1), 0, st, counter);
    return hipGetLastError();
}

bool subgrid_size_supported(int s) {
    return s == 8 || s == 16 || s == 24 || s == 32 || s == 48 || s == 64;
}

}  // namespace p9idg
