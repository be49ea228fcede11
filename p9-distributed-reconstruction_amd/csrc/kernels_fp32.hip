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

// ---------------------------------------------------------------------------
// Separable FP32 form (P9IDG_KERNEL_FP32_SEP): the phasor factorisation of the
// MFMA kernels, contracted with FP32 fused multiply-adds on the vector ALUs.
// With a_v, b_v, c_v the visibility's coefficients in turns and
// n_xy = n(l_x, 0) + n(0, m_y) + r_xy (r_xy: the per-plan table of
// ComputeN(l, m) - n(l, 0) - n(0, m), |r| < 2e-8 at the 3072 x 1.5" image),
//   A_v[x] = exp(2 pi i (a_v l_x + c_v n(l_x, 0)))
//   B_v[y] = exp(2 pi i (b_v m_y + c_v n(0, m_y)))
//   exp(2 pi i c_v r_xy) = 1 + 2 pi i c_v r_xy + O((2 pi c_v r_xy)^2)
// so a visibility needs 2 S phasors instead of S^2, and
//   gridder:   subgrid[y][x] = M0 + 2 pi i r_xy M1,
//              M0 = sum_v (V_v B_v[y]) A_v[x],  M1 = sum_v c_v (V_v B_v[y]) A_v[x]
//   degridder: vis_v = scale sum_y conj(B_v[y]) (Q0[y] - 2 pi i c_v Q1[y]),
//              Q0[y] = sum_x P[y][x] conj(A_v[x]),  Q1[y] = sum_x r_xy P[y][x] conj(A_v[x])
// Phases are FP64 (coefficients and the per-entry phase, reduced to turns)
// before the FP32 hardware sin/cos; everything else is FP32.

// exp(2 pi i ph) for ph in turns (FP64), reduced before the FP32 sin/cos.
__device__ __forceinline__ float2 cis_turns_f64(double ph) {
    const float r = (float)__builtin_amdgcn_fract(ph);
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

constexpr double kInvTwoPiSep = 1.0 / kTwoPi;
constexpr float kTwoPiSep = 6.2831853071795864769f;

// Gridder for S = 16, 32, 48, 64: thread t owns the R x R pixel block
// (t / 16, t % 16) of the subgrid (R = S / 16).  A round stages NV
// visibilities' A and V B phasors in LDS, then every thread runs the
// 8 FMAs per pixel and visibility of the two complex contractions.
template <int S>
__device__ __forceinline__ void grid_sep_body(const KernelArgs& a, int sidx, int k0, int k1, int slot) {
    constexpr int R = S / 16;
    static_assert(R * 16 == S && R <= 4, "S = 16, 32, 48 or 64");
    constexpr int NV = S <= 32 ? 64 : 32;  // visibilities per round
    __shared__ float2 s_A[NV][S];
    __shared__ float2 s_L[NV][S];
    __shared__ double s_co[NV][3];
    __shared__ float2 s_v[NV];
    __shared__ float s_c[NV];
    __shared__ double s_l[S], s_n[S];
    __shared__ float2 s_tile[2][S * S];

    const SubgridGeom g = load_geom(a, sidx, S);
    const int tid = item_tid();
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    if (tid < S) {
        const float l = compute_l(tid, S, a.image_size);
        s_l[tid] = (double)l;
        s_n[tid] = (double)compute_n(l, 0.f);
    }
    const int by = tid / 16, bx = tid % 16;
    float2 m0[R][R], m1[R][R];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int j = 0; j < R; j++) m0[i][j] = m1[i][j] = make_float2(0.f, 0.f);

    for (int j0 = k0; j0 < K; j0 += NV) {
        const int nv = min(NV, K - j0);
        __syncthreads();  // the previous round's phasors are read
        if (tid < nv) {
            const int j = j0 + tid, t = g.t0 + j / C, ch = j % C;
            const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
            const double k = a.wavenumbers[ch];
            s_co[tid][0] = (g.u_off - k * p[0]) * kInvTwoPiSep;
            s_co[tid][1] = (g.v_off - k * p[1]) * kInvTwoPiSep;
            s_co[tid][2] = (g.w_off - k * p[2]) * kInvTwoPiSep;
            s_v[tid] = a.vis[vbase + j];
        }
        __syncthreads();
        for (int e = tid; e < nv * S; e += kThreads) {
            const int v = e / S, q = e - v * S;
            const double cc = s_co[v][2];
            const double ncq = cc * s_n[q];
            s_A[v][q] = cis_turns_f64(fma(s_co[v][0], s_l[q], ncq));
            s_L[v][q] = cmulf(s_v[v], cis_turns_f64(fma(s_co[v][1], s_l[q], ncq)));
            if (q == 0) s_c[v] = (float)cc;
        }
        __syncthreads();
        for (int v = 0; v < nv; v++) {
            float2 A[R], L[R], cL[R];
            const float c = s_c[v];
#pragma unroll
            for (int i = 0; i < R; i++) {
                A[i] = s_A[v][bx * R + i];
                L[i] = s_L[v][by * R + i];
                cL[i] = make_float2(c * L[i].x, c * L[i].y);
            }
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int jx = 0; jx < R; jx++) {
                    float2& u = m0[i][jx];
                    u.x = fmaf(L[i].x, A[jx].x, fmaf(-L[i].y, A[jx].y, u.x));
                    u.y = fmaf(L[i].x, A[jx].y, fmaf(L[i].y, A[jx].x, u.y));
                    float2& w = m1[i][jx];
                    w.x = fmaf(cL[i].x, A[jx].x, fmaf(-cL[i].y, A[jx].y, w.x));
                    w.y = fmaf(cL[i].x, A[jx].y, fmaf(cL[i].y, A[jx].x, w.y));
                }
        }
    }
    // subgrid = M0 + 2 pi i r M1 into the ifftshift-ed tile (Gridder.cs:76-79)
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int jx = 0; jx < R; jx++) {
            const int y = by * R + i, x = bx * R + jx;
            const float rr = kTwoPiSep * a.rtab[y * S + x];
            s_tile[0][((y + S / 2) % S) * S + (x + S / 2) % S] =
                make_float2(m0[i][jx].x - rr * m1[i][jx].y, m0[i][jx].y + rr * m1[i][jx].x);
        }
    __syncthreads();
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, +1);
    adder<S>(a, g, res, 1.0f / (float)(S * S), slot);
}

// Degridder for S = 16 and 32: thread (v, y) = (t / S, t % S) keeps pixel
// row y and r * row y in registers; a round is kThreads / S visibilities,
// whose x-phasors are staged in LDS; the S threads of a visibility (one
// lane group of a wave) sum their rows' terms with lane shuffles.
template <int S>
__device__ __forceinline__ void degrid_sep_body(const KernelArgs& a, int sidx, int k0, int k1) {
    static_assert(S == 16 || S == 32, "a visibility's rows within one wave");
    constexpr int NV = kThreads / S;
    __shared__ float2 s_tile[2][S * S];
    __shared__ float2 s_A[NV][S];
    __shared__ double s_l[S], s_n[S];

    const SubgridGeom g = load_geom(a, sidx, S);
    const int tid = item_tid();
    const int vl = tid / S, y = tid % S;
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;

    splitter<S>(a, g, s_tile[0]);
    if (tid < S) {
        const float l = compute_l(tid, S, a.image_size);
        s_l[tid] = (double)l;
        s_n[tid] = (double)compute_n(l, 0.f);
    }
    __syncthreads();
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, -1);
    // pixels_copy[y][x] = subgrid[(y+S/2)%S][(x+S/2)%S] (Gridder.cs:101-117)
    float2 P[S], RP[S];
#pragma unroll
    for (int x = 0; x < S; x++) {
        P[x] = res[((y + S / 2) % S) * S + (x + S / 2) % S];
        const float r = a.rtab[y * S + x];
        RP[x] = make_float2(r * P[x].x, r * P[x].y);
    }
    const double ly = s_l[y], ny = s_n[y];
    for (int j0 = k0; j0 < K; j0 += NV) {
        const int j = j0 + vl;
        const int jc = j < K ? j : K - 1;
        const int t = g.t0 + jc / C, ch = jc % C;
        const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
        const double k = a.wavenumbers[ch];
        const double ca = (g.u_off - k * p[0]) * kInvTwoPiSep;
        const double cb = (g.v_off - k * p[1]) * kInvTwoPiSep;
        const double cc = (g.w_off - k * p[2]) * kInvTwoPiSep;
        __syncthreads();  // the previous round's x-phasors are read
        s_A[vl][y] = cis_turns_f64(fma(ca, s_l[y], cc * s_n[y]));  // coordinate y used as x
        const float2 B = cis_turns_f64(fma(cb, ly, cc * ny));
        __syncthreads();
        float q0r = 0.f, q0i = 0.f, q1r = 0.f, q1i = 0.f;
#pragma unroll
        for (int x = 0; x < S; x++) {
            const float2 A = s_A[vl][x];
            // P conj(A)
            q0r = fmaf(P[x].x, A.x, fmaf(P[x].y, A.y, q0r));
            q0i = fmaf(P[x].y, A.x, fmaf(-P[x].x, A.y, q0i));
            q1r = fmaf(RP[x].x, A.x, fmaf(RP[x].y, A.y, q1r));
            q1i = fmaf(RP[x].y, A.x, fmaf(-RP[x].x, A.y, q1i));
        }
        // conj(B) (Q0 - 2 pi i c Q1)
        const float c2 = kTwoPiSep * (float)cc;
        const float zr = fmaf(c2, q1i, q0r), zi = fmaf(-c2, q1r, q0i);
        float sr = fmaf(B.x, zr, B.y * zi), si = fmaf(B.x, zi, -B.y * zr);
#pragma unroll
        for (int o = S / 2; o > 0; o >>= 1) {
            sr += __shfl_xor(sr, o);
            si += __shfl_xor(si, o);
        }
        if (y == 0 && j < K) a.vis_out[vbase + j] = make_float2(sr * a.degrid_scale, si * a.degrid_scale);
    }
}

// The separable form where it exists for S, else the per-pixel one.
template <int S>
__device__ __forceinline__ void grid_fp32_any(const KernelArgs& a, int s, int k0, int k1, int slot, bool sep) {
    if constexpr (S % 16 == 0) {
        if (sep) return grid_sep_body<S>(a, s, k0, k1, slot);
    }
    grid_fp32_body<S>(a, s, k0, k1, slot);
}
template <int S>
__device__ __forceinline__ void degrid_fp32_any(const KernelArgs& a, int s, int k0, int k1, bool sep) {
    if constexpr (S == 16 || S == 32) {
        if (sep) return degrid_sep_body<S>(a, s, k0, k1);
    }
    degrid_fp32_body<S>(a, s, k0, k1);
}

template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void grid_fp32_kernel(KernelArgs a) {
    grid_fp32_any<S>(a, a.order[blockIdx.x], 0, -1, a.order[blockIdx.x], SEP);
}
template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void grid_fp32_queue_kernel(KernelArgs a) {
    run_queue(a, [&](int s, int k0, int k1, int slot) { grid_fp32_any<S>(a, s, k0, k1, slot, SEP); });
}
template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void degrid_fp32_kernel(KernelArgs a) {
    degrid_fp32_any<S>(a, a.order[blockIdx.x], 0, -1, SEP);
}
template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void degrid_fp32_queue_kernel(KernelArgs a) {
    run_queue(a, [&](int s, int k0, int k1, int) { degrid_fp32_any<S>(a, s, k0, k1, SEP); });
}

template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void grid_fp32_batch_kernel(BatchArgs b) {
    run_batch_queue(b, [&](const KernelArgs& a, int s, int k0, int k1, int slot) { grid_fp32_any<S>(a, s, k0, k1, slot, SEP); });
}
template <int S, bool SEP>
__global__ __launch_bounds__(kThreads) void degrid_fp32_batch_kernel(BatchArgs b) {
    run_batch_queue(b, [&](const KernelArgs& a, int s, int k0, int k1, int) { degrid_fp32_any<S>(a, s, k0, k1, SEP); });
}

// both queue counters (claims and exits) to zero
__global__ void reset_counter_kernel(int32_t* counter) {
    counter[0] = 0;
    counter[1] = 0;
}

}  // namespace dev

// Launches KERNEL<S> (one workgroup per subgrid) or QKERNEL<S> (persistent
// queue over the work items).
#define P9_LAUNCH_FP32_ONE(KERNEL, QKERNEL, SV, SEPV)                                          \
    do {                                                                                      \
        if (cfg.queue) {                                                                      \
            int wg = cfg.workgroups > 0 ? cfg.workgroups                                      \
                                        : persistent_workgroups((const void*)QKERNEL<SV, SEPV>, dev::kThreads); \
            wg = wg < a.nitems ? wg : a.nitems;                                               \
            if (wg <= 0) return hipSuccess;                                                   \
            hipLaunchKernelGGL((QKERNEL<SV, SEPV>), dim3(wg), dim3(dev::kThreads), 0, st, a); \
        } else {                                                                              \
            hipLaunchKernelGGL((KERNEL<SV, SEPV>), dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); \
        }                                                                                     \
    } while (0)
#define P9_LAUNCH_FP32(KERNEL, QKERNEL, SV)                                                   \
    do {                                                                                      \
        if (sep)                                                                              \
            P9_LAUNCH_FP32_ONE(KERNEL, QKERNEL, SV, true);                                    \
        else                                                                                  \
            P9_LAUNCH_FP32_ONE(KERNEL, QKERNEL, SV, false);                                   \
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

hipError_t launch_grid_fp32(const KernelArgs& a, int S, bool sep, const LaunchCfg& cfg, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
#define P9_G(SV) P9_LAUNCH_FP32(dev::grid_fp32_kernel, dev::grid_fp32_queue_kernel, SV)
    P9_SWITCH_S(S, P9_G)
#undef P9_G
    return hipGetLastError();
}

hipError_t launch_degrid_fp32(const KernelArgs& a, int S, bool sep, const LaunchCfg& cfg, hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
#define P9_D(SV) P9_LAUNCH_FP32(dev::degrid_fp32_kernel, dev::degrid_fp32_queue_kernel, SV)
    P9_SWITCH_S(S, P9_D)
#undef P9_D
    return hipGetLastError();
}

namespace {
template <int S, bool SEP>
hipError_t fp32_batch_one(bool grid, const BatchArgs& b, int workgroups, hipStream_t st) {
    const void* k = grid ? (const void*)dev::grid_fp32_batch_kernel<S, SEP>
                         : (const void*)dev::degrid_fp32_batch_kernel<S, SEP>;
    int wg = workgroups > 0 ? workgroups : persistent_workgroups(k, dev::kThreads);
    wg = wg < b.nitems ? wg : b.nitems;
    if (wg <= 0) return hipSuccess;
    if (grid)
        hipLaunchKernelGGL((dev::grid_fp32_batch_kernel<S, SEP>), dim3(wg), dim3(dev::kThreads), 0, st, b);
    else
        hipLaunchKernelGGL((dev::degrid_fp32_batch_kernel<S, SEP>), dim3(wg), dim3(dev::kThreads), 0, st, b);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_fp32_batch(bool grid, bool sep, const BatchArgs& b, int S, int workgroups, hipStream_t st) {
#define P9_B(SV) return sep ? fp32_batch_one<SV, true>(grid, b, workgroups, st) : fp32_batch_one<SV, false>(grid, b, workgroups, st)
    P9_SWITCH_S(S, P9_B)
#undef P9_B
}

hipError_t launch_reset_counter(int32_t* counter, hipStream_t st) {
    hipLaunchKernelGGL(dev::reset_counter_kernel, dim3(1), dim3(1), 0, st, counter);
    return hipGetLastError();
}

bool subgrid_size_supported(int s) {
    return s == 8 || s == 16 || s == 24 || s == 32 || s == 48 || s == 64;
}

}  // namespace p9idg
