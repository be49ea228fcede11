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
#include "device_common.h"

namespace p9idg {
namespace dev {

// Adds the (normalised) subgrid tile onto the grid.  `tile` is the subgrid
// after the inverse FFT, indexed [y][x] in FFT order.
template <int S>
__device__ __forceinline__ void adder(const KernelArgs& a, const SubgridGeom& g,
                                      const float2* tile, float norm) {
    const int G = a.G;
    if (a.mode == P9IDG_GRID_PLAIN) {
        // Adder.cs:22-47 (range validated on the host: 0 <= U,V < G-S)
        for (int p = threadIdx.x; p < S * S; p += kThreads) {
            const int y = p / S, x = p - (p / S) * S;
            const int ys = (y + S / 2) % S, xs = (x + S / 2) % S;
            float2 v = cmulf(tile[ys * S + xs], a.adder_phasor[p]);
            v.x *= norm;
            v.y *= norm;
            float2* dst = a.grid + (size_t)shifted(g.V + y, G) * G + shifted(g.U + x, G);
            unsafeAtomicAdd(&dst->x, v.x);
            unsafeAtomicAdd(&dst->y, v.y);
        }
        return;
    }
    // AdderWStack.cs:29-74: negative w layers are mirrored and conjugated.
    const bool neg = g.W < 0;
    int X = g.U, Y = g.V, W = g.W;
    if (neg) {
        X = G - X - S + 1;
        Y = G - Y - S + 1;
        W = -W - 1;
    }
    if (!(X >= 1 && X < G - S && Y >= 1 && Y < G - S)) return;
    if (W < 0 || W >= a.L) return;  // metadata validated on the host
    float2* layer = a.grid + (a.mode == P9IDG_GRID_WSUM ? 0 : (size_t)W * G * G);
    for (int p = threadIdx.x; p < S * S; p += kThreads) {
        const int y = p / S, x = p - (p / S) * S;
        const int y_ = neg ? S - 1 - y : y;
        const int x_ = neg ? S - 1 - x : x;
        const int ys = (y_ + S / 2) % S, xs = (x_ + S / 2) % S;
        float2 v = cmulf(a.adder_phasor[y_ * S + x_], tile[ys * S + xs]);
        v.x *= norm;
        v.y *= norm;
        if (neg) v.y = -v.y;
        float2* dst = layer + (size_t)shifted(Y + y, G) * G + shifted(X + x, G);
        unsafeAtomicAdd(&dst->x, v.x);
        unsafeAtomicAdd(&dst->y, v.y);
    }
}

// Reads the subgrid from the model grid into `tile` (FFT order), zeros where
// the reference leaves the subgrid empty.
template <int S>
__device__ __forceinline__ void splitter(const KernelArgs& a, const SubgridGeom& g,
                                         float2* tile) {
    const int G = a.G;
    if (a.mode == P9IDG_DEGRID_PLAIN) {
        // Adder.cs:65-90
        const bool ok = g.U >= 0 && g.U < G - S && g.V >= 0 && g.V < G - S;
        for (int p = threadIdx.x; p < S * S; p += kThreads) {
            const int y = p / S, x = p - (p / S) * S;
            const int yd = (y + S / 2) % S, xd = (x + S / 2) % S;
            float2 v = make_float2(0.f, 0.f);
            if (ok) {
                float2 gv = a.model_grid[(size_t)shifted(g.V + y, G) * G + shifted(g.U + x, G)];
                v = cmulf(gv, conjf2(a.adder_phasor[p]));
            }
            tile[yd * S + xd] = v;
        }
        return;
    }
    // AdderWStack.cs:97-134: one grid; negative w reads the mirrored pixel
    // conjugated; the range test uses the unmirrored corner.
    const bool neg = g.W < 0;
    const int X = g.U, Y = g.V;
    const bool ok = X >= 1 && X < G - S && Y >= 1 && Y < G - S;
    for (int p = threadIdx.x; p < S * S; p += kThreads) {
        const int y = p / S, x = p - (p / S) * S;
        const int yd = (y + S / 2) % S, xd = (x + S / 2) % S;
        float2 v = make_float2(0.f, 0.f);
        if (ok) {
            const int xs = neg ? G - X - x : X + x;
            const int ys = neg ? G - Y - y : Y + y;
            float2 gv = a.model_grid[(size_t)shifted(ys, G) * G + shifted(xs, G)];
            if (neg) gv.y = -gv.y;
            v = cmulf(conjf2(a.adder_phasor[p]), gv);
        }
        tile[yd * S + xd] = v;
    }
}

template <int S>
__global__ __launch_bounds__(kThreads) void grid_fp32_kernel(KernelArgs a) {
    constexpr int NPIX = (S * S + kThreads - 1) / kThreads;
    constexpr int CHUNK = kThreads;
    __shared__ float4 s_coef[CHUNK];
    __shared__ float2 s_vis[CHUNK];
    __shared__ float2 s_tile[2][S * S];

    const int s = a.order[blockIdx.x];
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
    const int K = g.nt * C;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    for (int j0 = 0; j0 < K; j0 += CHUNK) {
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
    adder<S>(a, g, res, 1.0f / (float)(S * S));
}

template <int S>
__global__ __launch_bounds__(kThreads) void degrid_fp32_kernel(KernelArgs a) {
    __shared__ float2 s_tile[2][S * S];
    __shared__ float s_n[S * S];
    __shared__ float s_l[S];

    const int s = a.order[blockIdx.x];
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
    const int K = g.nt * C;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    for (int j = tid; j < K; j += kThreads) {
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

}  // namespace dev

#define P9_DISPATCH_S(S, KERNEL)                                                          \
    switch (S) {                                                                          \
        case 8: hipLaunchKernelGGL(KERNEL<8>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break;   \
        case 16: hipLaunchKernelGGL(KERNEL<16>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break; \
        case 24: hipLaunchKernelGGL(KERNEL<24>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break; \
        case 32: hipLaunchKernelGGL(KERNEL<32>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break; \
        case 48: hipLaunchKernelGGL(KERNEL<48>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break; \
        case 64: hipLaunchKernelGGL(KERNEL<64>, dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); break; \
        default: return hipErrorInvalidValue;                                             \
    }

hipError_t launch_grid_fp32(const KernelArgs& a, int S, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
    P9_DISPATCH_S(S, dev::grid_fp32_kernel);
    return hipGetLastError();
}

hipError_t launch_degrid_fp32(const KernelArgs& a, int S, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
    P9_DISPATCH_S(S, dev::degrid_fp32_kernel);
    return hipGetLastError();
}

bool subgrid_size_supported(int s) {
    return s == 8 || s == 16 || s == 24 || s == 32 || s == 48 || s == 64;
}

}  // namespace p9idg
