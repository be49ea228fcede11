// Split-precision MFMA gridder and degridder (gfx950, 32x32x16 BF16/FP16
// MFMA with FP32 accumulation), each fused with the LDS subgrid FFT and the
// adder / splitter.  Same stages and semantics as kernels_fp32.hip
// (Gridder.cs:15-160, SubgridFFT.cs, Adder.cs, AdderWStack.cs).
//
// The reference's phasor for visibility v and subgrid pixel (y, x) is
//   exp(2 pi i (a_v l_x + b_v m_y + c_v n_xy))          (Gridder.cs:62-70)
// with a, b, c the per-visibility coefficients in turns (vis_coeffs) and
// n_xy = ComputeN(l_x, m_y).  ComputeN evaluates t / (t + sqrt(1 - t)),
// t = l^2 + m^2 (Gridder.cs:172-173), which is t - t^2/2 + ...; splitting it
// along the axes, n_xy = n(l_x, 0) + n(0, m_y) + r_xy with r_xy ~ -l^2 m^2
// (below 1.6e-8 for a 3072 x 1.5" image), the phasor factors into
//   A_v[x] = exp(2 pi i (a_v l_x + c_v n(l_x, 0)))   (one row of 32 per vis)
//   B_v[y] = exp(2 pi i (b_v m_y + c_v n(0, m_y)))   (one column of 32 per vis)
//   W_v[y,x] = exp(2 pi i c_v r_xy) = 1 + 2 pi i c_v r_xy + O((2 pi c_v r_xy)^2)
// and one correction term keeps the error below (2 pi |c_v| r_max)^2 / 2.
// The gridder becomes two complex GEMMs over the visibilities of a subgrid:
//   M0[y][x] = sum_v (V_v B_v[y]) A_v[x],   M1[y][x] = sum_v (c_v V_v B_v[y]) A_v[x]
//   subgrid[y][x] = M0 + 2 pi i r_xy M1
// and the degridder, with P the split subgrid pixels,
//   Q0[y][v] = sum_x P[y][x] conj(A_v[x]),  Q1[y][v] = sum_x r_xy P[y][x] conj(A_v[x])
//   vis_v = scale * sum_y conj(B_v[y]) (Q0 - 2 pi i c_v Q1)
// Phases are evaluated in FP64 (coefficients and per-entry phase), reduced to
// [0, 1) turns and passed to the hardware sin/cos.  The GEMM operands of
// the main terms are split x = hi + lo in BF16 (or FP16) and multiplied as
// hi*hi + hi*lo + lo*hi with FP32 accumulation; the correction terms use the
// hi parts only.  FP16 operands are normalised by powers of two (exact) per
// subgrid so the low parts stay in FP16's normal range.
#include "device_addsplit.h"

namespace p9idg {
namespace dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// An MFMA operand of 8 16-bit values, two per dword (element 2q in the low
// half of dword q), so pairs are converted with one v_cvt_pk instruction.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// (a, b) = hi + lo per value, both rounded to nearest in the operand type.
struct HiLo {
    uint32_t hi, lo;
};
// Run-time -1 operands of the split's residual (KernelArgs::split_k).
struct SplitK {
    uint32_t bf16_lo, bf16_hi, f16;
};
__device__ __forceinline__ SplitK split_k(const KernelArgs& a) { return {a.split_k[0], a.split_k[1], a.split_k[2]}; }

struct PrecBF16 {
    typedef __bf16 T;
    typedef bf16x8 V;
    static constexpr bool kNormalise = false;
    __device__ static __forceinline__ f32x16 mma(u32x4 a, u32x4 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(V, a), __builtin_bit_cast(V, b), c, 0, 0, 0);
    }
    __device__ static __forceinline__ f32x4 mma16(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(V, a), __builtin_bit_cast(V, b), c, 0, 0, 0);
    }
    __device__ static __forceinline__ uint32_t pack(float a, float b) {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
    }
    // a - hi.lo and b - hi.hi, exact: hi (rounded to nearest) is within a
    // factor of two of the value, so the difference is representable; each
    // is one dot2 of hi with (-1, 0) or (0, -1) accumulated onto the value
    __device__ static __forceinline__ HiLo split(float a, float b, SplitK k) {
        HiLo r;
        r.hi = pack(a, b);
        const bf16x2 h = __builtin_bit_cast(bf16x2, r.hi);
        const float la = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf16x2, k.bf16_lo), a, false);
        const float lb = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf16x2, k.bf16_hi), b, false);
        r.lo = pack(la, lb);
        return r;
    }
};
struct PrecF16 {
    typedef _Float16 T;
    typedef f16x8 V;
    static constexpr bool kNormalise = true;
    __device__ static __forceinline__ f32x16 mma(u32x4 a, u32x4 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(V, a), __builtin_bit_cast(V, b), c, 0, 0, 0);
    }
    __device__ static __forceinline__ f32x4 mma16(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(V, a), __builtin_bit_cast(V, b), c, 0, 0, 0);
    }
    __device__ static __forceinline__ uint32_t pack(float a, float b) {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, f16x2));
    }
    // a - hi.lo and b - hi.hi, exact (as for BF16), each one mixed-precision
    // FMA of the FP16 half with -1
    __device__ static __forceinline__ HiLo split(float a, float b, SplitK k) {
        HiLo r;
        r.hi = pack(a, b);
        const f16x2 h = __builtin_bit_cast(f16x2, r.hi);
        const float m1 = (float)__builtin_bit_cast(_Float16, (uint16_t)k.f16);
        const float la = __builtin_fmaf((float)h.x, m1, a);
        const float lb = __builtin_fmaf((float)h.y, m1, b);
        r.lo = pack(la, lb);
        return r;
    }
};

// Writes the split of (a, b) into the operand dwords EH (hi) and EL (lo).
#define P9_SPLIT2(P, K, a, b, EH, EL)            \
    do {                                         \
        const HiLo s_ = P::split((a), (b), (K)); \
        EH = s_.hi;                              \
        EL = s_.lo;                              \
    } while (0)
// Negates both 16-bit values of every dword (sign bits).
__device__ __forceinline__ u32x4 neg(u32x4 v) { return v ^ (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u}; }

// exp(2 pi i ph) for ph in turns (FP64), reduced to [-0.5, 0.5] before the
// FP32 hardware sin/cos.
__device__ __forceinline__ float2 cis_f64(double ph) {
    const float r = (float)__builtin_amdgcn_fract(ph);  // [0, 1) turns
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

// Separable part of ComputeN along one axis: n(l, 0) = n(0, l).
__device__ __forceinline__ double axis_n(int x, int S, float image_size) {
    return (double)compute_n(compute_l(x, S, image_size), 0.f);
}
// r_xy = ComputeN(l_x, m_y) - n(l_x, 0) - n(0, m_y) (FP64 of the float n's).
__device__ __forceinline__ double residual_n(int x, int y, int S, float image_size) {
    const float l = compute_l(x, S, image_size), m = compute_l(y, S, image_size);
    return (double)compute_n(l, m) - (double)compute_n(l, 0.f) - (double)compute_n(0.f, m);
}

// r(x, y) for every pixel of the subgrid, [y][x]; built once per plan.
__global__ void rtab_kernel(float* out, int S, float image_size) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S * S) out[i] = (float)residual_n(i % S, i / S, S, image_size);
}

// Power of two >= |x| (1 for x == 0); exact normaliser.
__device__ __forceinline__ float pow2_ceil(float x) {
    if (!(x > 0.f)) return 1.f;
    return exp2f(ceilf(log2f(x)));
}

// Block-wide max over 256 threads (4 waves).
__device__ __forceinline__ float block_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

constexpr double kInvTwoPi = 1.0 / kTwoPi;
constexpr float kTwoPiF = 6.2831853071795864769f;

// x[l] + x[l ^ 16] (rows of 16 lanes) and x[l] + x[l ^ 32] (half waves) in
// every lane l, over the row and half-wave swaps: VALU instructions, so the
// sums wait on no LDS round trip (and no LDS counter shared with the tile
// loads).  Each swap leaves x[l] in one output and x[l ^ 16] (x[l ^ 32]) in
// the other, and the sum of the two is the same in either order.
__device__ __forceinline__ float sum_xor16(float x) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float sum_xor32(float x) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// ---------------------------------------------------------------------------
// Gridder on 32x32x16 MFMA tiles, used for S = 8, 24 and 64.  One 256-thread
// workgroup per subgrid.  For S <= 32 the subgrid is one 32 x 32 MFMA tile
// and the 4 waves split the subgrid's visibilities (16 per MFMA step, 8 per
// lane group), each keeping a full set of accumulators; for S = 64 the
// subgrid is 2 x 2 tiles and wave w owns tile w over all visibilities.
// Partial subgrids are summed in LDS.
template <int S, class P>
__device__ __forceinline__ void grid_mfma_body(const KernelArgs& a, int sidx, int k0, int k1, int slot) {
    constexpr int NT = (S + 31) / 32;  // tiles per axis
    constexpr int WPK = 4 / (NT * NT); // waves sharing one tile's visibilities
    static_assert(NT * 32 >= S && NT <= 2, "subgrids up to 64");
    constexpr int PER = 2;  // visibilities staged per thread and round
    constexpr int ROUND = PER * kThreads;
    __shared__ double s_a[ROUND], s_b[ROUND], s_c[ROUND];
    __shared__ float2 s_v[ROUND];
    __shared__ float2 s_tile[2][S * S];
    __shared__ float s_red[4];

    const SubgridGeom g = load_geom(a, sidx, S);
    const SplitK sk = split_k(a);
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int li = lane & 31, grp = lane >> 5;
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    const bool chan_rec = a.chan_rec != 0;

    for (int p = tid; p < S * S; p += kThreads) s_tile[0][p] = make_float2(0.f, 0.f);
    __syncthreads();

    // FP16: scale the visibilities so |V| <= 1, and c * L by 2^-16.
    float vscale = 1.f;
    constexpr float cscale = P::kNormalise ? 1.f / 65536.f : 1.f;  // BF16: no multiply
    if (P::kNormalise) {
        float m = 0.f;
        for (int j = k0 + tid; j < K; j += kThreads) {
            const float2 v = a.vis[vbase + j];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
        vscale = 1.f / pow2_ceil(block_max(m, s_red));
    }

    // this wave's tile and the lane's row y (A operand) and column x (B operand)
    const int tile = wave / WPK, kpart = wave % WPK;
    const int ty = tile / NT, tx = tile % NT;
    const double lrow = (double)compute_l(ty * 32 + li, S, a.image_size);
    const double nrow = axis_n(ty * 32 + li, S, a.image_size);
    const double lcol = (double)compute_l(tx * 32 + li, S, a.image_size);
    const double ncol = axis_n(tx * 32 + li, S, a.image_size);

    f32x16 acc0r = {}, acc0i = {}, acc1r = {}, acc1i = {}, acc0n = {}, acc1n = {};

    // The next round's uvw, wavenumbers and visibilities are fetched into
    // registers while the current round is on the MFMA, then staged in LDS
    // as FP64 phase coefficients (a, b, c) and scaled visibilities.
    double fu[PER], fv[PER], fw[PER], fk[PER];
    float2 fvis[PER];
    auto fetch = [&](int j0) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int jc = min(j0 + i * kThreads + tid, K - 1);
            const int t = g.t0 + jc / C, ch = jc % C;
            const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
            fu[i] = p[0];
            fv[i] = p[1];
            fw[i] = p[2];
            fk[i] = a.wavenumbers[ch];
            fvis[i] = a.vis[vbase + jc];
        }
    };
    if (k0 < K) fetch(k0);
    for (int j0 = k0; j0 < K; j0 += ROUND) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = i * kThreads + tid;
            const bool in = j0 + idx < K;
            s_a[idx] = in ? (g.u_off - fk[i] * fu[i]) * kInvTwoPi : 0.0;
            s_b[idx] = in ? (g.v_off - fk[i] * fv[i]) * kInvTwoPi : 0.0;
            s_c[idx] = in ? (g.w_off - fk[i] * fw[i]) * kInvTwoPi : 0.0;
            s_v[idx] = in ? make_float2(fvis[i].x * vscale, fvis[i].y * vscale) : make_float2(0.f, 0.f);
        }
        __syncthreads();
        if (j0 + ROUND < K) fetch(j0 + ROUND);
        const int nsteps = (min(ROUND, K - j0) + 15) / 16;
        for (int st = kpart; st < nsteps; st += WPK) {
            u32x4 Lrh, Lrl, Lih, Lil, cLr, cLi, Arh, Arl, Aih, Ail;
            const int jg = st * 16 + grp * 8;  // this lane group's 8 visibilities
            if (chan_rec) {
                // one timestep's 8 channels: phasors at the first channel times
                // the channel-step phasor (FP64 phases, FP32 products)
                const double ca = s_a[jg], cb = s_b[jg], cc = s_c[jg];
                const double da = s_a[jg + 1] - ca, db = s_b[jg + 1] - cb, dc = s_c[jg + 1] - cc;
                float2 A = cis_f64(fma(ca, lcol, cc * ncol));
                float2 B = cis_f64(fma(cb, lrow, cc * nrow));
                const float2 RA = cis_f64(fma(da, lcol, dc * ncol));
                const float2 RB = cis_f64(fma(db, lrow, dc * nrow));
                const float c0 = (float)cc * cscale, d0 = (float)dc * cscale;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    float2 L[2], Aq[2];
                    float cf[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int e = 2 * q + h;
                        Aq[h] = A;
                        L[h] = cmulf(s_v[jg + e], B);
                        cf[h] = fmaf((float)e, d0, c0);
                        if (e < 7) {
                            A = cmulf(A, RA);
                            B = cmulf(B, RB);
                        }
                    }
                    cLr[q] = P::pack(cf[0] * L[0].x, cf[1] * L[1].x);
                    cLi[q] = P::pack(cf[0] * L[0].y, cf[1] * L[1].y);
                    P9_SPLIT2(P, sk, L[0].x, L[1].x, Lrh[q], Lrl[q]);
                    P9_SPLIT2(P, sk, L[0].y, L[1].y, Lih[q], Lil[q]);
                    P9_SPLIT2(P, sk, Aq[0].x, Aq[1].x, Arh[q], Arl[q]);
                    P9_SPLIT2(P, sk, Aq[0].y, Aq[1].y, Aih[q], Ail[q]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {  // elements 2q, 2q+1
                    float2 L[2], A[2];
                    float cf[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int j = jg + 2 * q + h;
                        const double ca = s_a[j], cb = s_b[j], cc = s_c[j];
                        A[h] = cis_f64(fma(ca, lcol, cc * ncol));
                        L[h] = cmulf(s_v[j], cis_f64(fma(cb, lrow, cc * nrow)));
                        cf[h] = (float)cc * cscale;
                    }
                    cLr[q] = P::pack(cf[0] * L[0].x, cf[1] * L[1].x);
                    cLi[q] = P::pack(cf[0] * L[0].y, cf[1] * L[1].y);
                    P9_SPLIT2(P, sk, L[0].x, L[1].x, Lrh[q], Lrl[q]);
                    P9_SPLIT2(P, sk, L[0].y, L[1].y, Lih[q], Lil[q]);
                    P9_SPLIT2(P, sk, A[0].x, A[1].x, Arh[q], Arl[q]);
                    P9_SPLIT2(P, sk, A[0].y, A[1].y, Aih[q], Ail[q]);
                }
            }
            // Re M0 = Lr Ar - Li Ai, Im M0 = Lr Ai + Li Ar (hi*hi + hi*lo + lo*hi);
            // the Li Ai products accumulate apart and are subtracted at the end
            acc0r = P::mma(Lrh, Arh, acc0r);
            acc0i = P::mma(Lrh, Aih, acc0i);
            acc1r = P::mma(cLr, Arh, acc1r);
            acc1i = P::mma(cLr, Aih, acc1i);
            acc0r = P::mma(Lrh, Arl, acc0r);
            acc0i = P::mma(Lrh, Ail, acc0i);
            acc1n = P::mma(cLi, Aih, acc1n);
            acc1i = P::mma(cLi, Arh, acc1i);
            acc0r = P::mma(Lrl, Arh, acc0r);
            acc0i = P::mma(Lrl, Aih, acc0i);
            acc0n = P::mma(Lih, Aih, acc0n);
            acc0i = P::mma(Lih, Arh, acc0i);
            acc0n = P::mma(Lih, Ail, acc0n);
            acc0i = P::mma(Lih, Arl, acc0i);
            acc0n = P::mma(Lil, Aih, acc0n);
            acc0i = P::mma(Lil, Arh, acc0i);
        }
        __syncthreads();
    }

    // subgrid = (M0 + 2 pi i r M1) / vscale, summed over the 4 waves into the
    // ifftshift-ed tile (Gridder.cs:76-79; identity taper)
    const float back = 1.f / vscale;
    const float back1 = back / cscale;
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = ty * 32 + 8 * (r / 4) + 4 * grp + (r % 4), x = tx * 32 + li;
                if (y < S && x < S) {
                    const float rr = kTwoPiF * a.rtab[y * S + x];
                    const float m1r = (acc1r[r] - acc1n[r]) * back1, m1i = acc1i[r] * back1;
                    const float gr = (acc0r[r] - acc0n[r]) * back - rr * m1i;
                    const float gi = acc0i[r] * back + rr * m1r;
                    float2& d = s_tile[0][((y + S / 2) % S) * S + (x + S / 2) % S];
                    d.x += gr;
                    d.y += gi;
                }
            }
        }
        __syncthreads();
    }
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, +1);
    adder<S>(a, g, res, 1.0f / (float)(S * S), slot);
}

// ---------------------------------------------------------------------------
// Gridder for S = 16, 32, 48 on 16x16x32 MFMA tiles: the subgrid is NT x NT
// tiles of 16 x 16 (S = 48 is not padded to 64).  One 256-thread workgroup
// per work item; a step is 32 visibilities, each of the 4 waves takes a
// contiguous quarter of the item's steps and accumulates all NT x NT tiles.
// Lane l = 16 g + c: c is the row (A operand) and column (B operand) within a
// tile, g picks the lane's 8 visibilities of the step.
// Per step the x-phasors of the NT column tiles are built once, then each row
// tile's V * y-phasor operands are built and multiplied into its NT tiles.
// fresh: the plan's tables (rtab, twiddles, adder phasors) are not yet in
// this workgroup's LDS (its first item, or an item of another plan).
template <int S, class P>
__device__ __forceinline__ void grid_mfma16_body(const KernelArgs& a, int sidx, int k0, int k1, int slot,
                                                 bool fresh) {
    constexpr int NT = S / 16;  // tiles per axis
    static_assert(NT * 16 == S, "subgrid size a multiple of 16");
    constexpr int PER = 2;  // visibilities staged per thread and round
    constexpr int ROUND = PER * kThreads;
    // s_buf holds a round's staged visibilities (s_a, s_b, s_c: ROUND doubles
    // each; s_v: ROUND float2) followed by the two FFT tiles; after the last
    // round the 4 waves' partial subgrids overlay it from the start, the last
    // two in the FFT tiles.
    constexpr int kStage = 4 * ROUND;  // float2 units
    constexpr int kBuf = kStage + 2 * S * S > 4 * S * S ? kStage + 2 * S * S : 4 * S * S;
    // the in-place sum of the partials into the first FFT tile needs that
    // tile to be partial 2 exactly, or clear of all four
    static_assert(kBuf - 2 * S * S == 2 * S * S || kBuf - 2 * S * S >= 4 * S * S, "partials vs FFT tiles");
    __shared__ float2 s_buf[kBuf];
    __shared__ float s_cf[ROUND];  // the staged c as FP32 (scaled for FP16)
    __shared__ float s_rt[S * S];
    __shared__ float2 s_ap[S * S], s_tw[S];
    __shared__ float s_red[4];
    double* const s_a = reinterpret_cast<double*>(s_buf);
    double* const s_b = s_a + ROUND;
    double* const s_c = s_b + ROUND;
    float2* const s_v = s_buf + 3 * ROUND;
    float2* const tile0 = s_buf + (kBuf - 2 * S * S);

    const SubgridGeom g = load_geom(a, sidx, S);
    const SplitK sk = split_k(a);
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int c = lane & 15, grp = lane >> 4;
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;
    const bool chan_rec = a.chan_rec != 0;

    if (fresh) {  // uniform over the workgroup
        for (int p = tid; p < S * S; p += kThreads) {
            s_rt[p] = a.rtab[p];
            s_ap[p] = a.adder_phasor[p];
        }
        for (int p = tid; p < S; p += kThreads) s_tw[p] = a.twiddle[p];
        __syncthreads();
    }

    // FP16: scale the visibilities so |V| <= 1, and c * L by 2^-16.
    float vscale = 1.f;
    constexpr float cscale = P::kNormalise ? 1.f / 65536.f : 1.f;  // BF16: no multiply
    if (P::kNormalise) {
        float m = 0.f;
        for (int j = k0 + tid; j < K; j += kThreads) {
            const float2 v = a.vis[vbase + j];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
        vscale = 1.f / pow2_ceil(block_max(m, s_red));
    }

    // l and the separable n of coordinate 16 t + c (rows and columns alike)
    double lt[NT], nt[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        lt[t] = (double)compute_l(16 * t + c, S, a.image_size);
        nt[t] = axis_n(16 * t + c, S, a.image_size);
    }

    f32x4 acc0r[NT][NT], acc0i[NT][NT], acc1r[NT][NT], acc1i[NT][NT];
#pragma unroll
    for (int r = 0; r < NT; r++)
#pragma unroll
        for (int q = 0; q < NT; q++) acc0r[r][q] = acc0i[r][q] = acc1r[r][q] = acc1i[r][q] = (f32x4){};

    // Each wave takes a contiguous quarter of the item's steps (32 visibilities
    // each) and stages them WROUND at a time in its own quarter of the staging
    // arrays: a wave's LDS operations run in order, so the waves need no
    // workgroup barrier until their partial tiles.
    constexpr int WROUND = ROUND / 4;
    const int nst = (K - k0 + 31) / 32;
    const int jw0 = k0 + 32 * (nst * wave / 4), jw1 = min(K, k0 + 32 * (nst * (wave + 1) / 4));
    double* const w_a = s_a + wave * WROUND;
    double* const w_b = s_b + wave * WROUND;
    double* const w_c = s_c + wave * WROUND;
    float* const w_cf = s_cf + wave * WROUND;
    float2* const w_v = s_v + wave * WROUND;

    double fu[PER], fv[PER], fw[PER], fk[PER];
    float2 fvis[PER];
    auto fetch = [&](int j0) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int jc = min(j0 + i * 64 + lane, K - 1);
            const int t = g.t0 + jc / C, ch = jc % C;
            const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
            fu[i] = p[0];
            fv[i] = p[1];
            fw[i] = p[2];
            fk[i] = a.wavenumbers[ch];
            fvis[i] = a.vis[vbase + jc];
        }
    };
    // the 8 phasors of this lane group's visibilities at coefficient pair
    // (ca, cc) for coordinate tile t: channel stepping or direct evaluation
    auto phasors = [&](int jg, const double* co, int t, float2 (&out)[8]) {
        if (chan_rec) {
            const double ca = co[jg], cc = w_c[jg];
            const double da = co[jg + 1] - ca, dc = w_c[jg + 1] - cc;
            float2 A = cis_f64(fma(ca, lt[t], cc * nt[t]));
            const float2 R = cis_f64(fma(da, lt[t], dc * nt[t]));
#pragma unroll
            for (int e = 0; e < 8; e++) {
                out[e] = A;
                if (e < 7) A = cmulf(A, R);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) out[e] = cis_f64(fma(co[jg + e], lt[t], w_c[jg + e] * nt[t]));
        }
    };

    if (jw0 < jw1) fetch(jw0);
    for (int j0 = jw0; j0 < jw1; j0 += WROUND) {  // uniform over the wave
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = i * 64 + lane;
            const bool in = j0 + idx < jw1;
            const double cc = in ? (g.w_off - fk[i] * fw[i]) * kInvTwoPi : 0.0;
            w_a[idx] = in ? (g.u_off - fk[i] * fu[i]) * kInvTwoPi : 0.0;
            w_b[idx] = in ? (g.v_off - fk[i] * fv[i]) * kInvTwoPi : 0.0;
            w_c[idx] = cc;
            w_cf[idx] = (float)cc * cscale;
            w_v[idx] = in ? make_float2(fvis[i].x * vscale, fvis[i].y * vscale) : make_float2(0.f, 0.f);
        }
        __builtin_amdgcn_wave_barrier();
        if (j0 + WROUND < jw1) fetch(j0 + WROUND);
        const int nsteps = (min(WROUND, jw1 - j0) + 31) / 32;
        for (int st = 0; st < nsteps; st++) {
            const int jg = st * 32 + grp * 8;  // this lane group's 8 visibilities
            // x-phasors of the column tiles, split; -Ai carries the products'
            // minus signs (one negation per column tile, not three per row tile)
            u32x4 Arh[NT], Arl[NT], Aih[NT], Ail[NT], nAih[NT], nAil[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                float2 A[8];
                phasors(jg, w_a, t, A);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    P9_SPLIT2(P, sk, A[2 * q].x, A[2 * q + 1].x, Arh[t][q], Arl[t][q]);
                    P9_SPLIT2(P, sk, A[2 * q].y, A[2 * q + 1].y, Aih[t][q], Ail[t][q]);
                }
                nAih[t] = neg(Aih[t]);
                nAil[t] = neg(Ail[t]);
            }
            // c of the 8 visibilities (FP32; scaled for FP16)
            float cf[8];
#pragma unroll
            for (int e = 0; e < 8; e++) cf[e] = w_cf[jg + e];
#pragma unroll
            for (int r = 0; r < NT; r++) {
                float2 B[8];
                phasors(jg, w_b, r, B);
                u32x4 Lrh, Lrl, Lih, Lil, cLr, cLi;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float2 L0 = cmulf(w_v[jg + 2 * q], B[2 * q]), L1 = cmulf(w_v[jg + 2 * q + 1], B[2 * q + 1]);
                    cLr[q] = P::pack(cf[2 * q] * L0.x, cf[2 * q + 1] * L1.x);
                    cLi[q] = P::pack(cf[2 * q] * L0.y, cf[2 * q + 1] * L1.y);
                    P9_SPLIT2(P, sk, L0.x, L1.x, Lrh[q], Lrl[q]);
                    P9_SPLIT2(P, sk, L0.y, L1.y, Lih[q], Lil[q]);
                }
#pragma unroll
                for (int q = 0; q < NT; q++) {
                    // Re M0 = Lr Ar - Li Ai, Im M0 = Lr Ai + Li Ar (hi*hi + hi*lo + lo*hi)
                    f32x4& m0r = acc0r[r][q];
                    f32x4& m0i = acc0i[r][q];
                    m0r = P::mma16(Lrh, Arh[q], m0r);
                    m0i = P::mma16(Lrh, Aih[q], m0i);
                    acc1r[r][q] = P::mma16(cLr, Arh[q], acc1r[r][q]);
                    acc1i[r][q] = P::mma16(cLr, Aih[q], acc1i[r][q]);
                    m0r = P::mma16(Lrh, Arl[q], m0r);
                    m0i = P::mma16(Lrh, Ail[q], m0i);
                    acc1r[r][q] = P::mma16(cLi, nAih[q], acc1r[r][q]);
                    acc1i[r][q] = P::mma16(cLi, Arh[q], acc1i[r][q]);
                    m0r = P::mma16(Lrl, Arh[q], m0r);
                    m0i = P::mma16(Lrl, Aih[q], m0i);
                    m0r = P::mma16(Lih, nAih[q], m0r);
                    m0i = P::mma16(Lih, Arh[q], m0i);
                    m0r = P::mma16(Lih, nAil[q], m0r);
                    m0i = P::mma16(Lih, Arl[q], m0i);
                    m0r = P::mma16(Lil, nAih[q], m0r);
                    m0i = P::mma16(Lil, Arh[q], m0i);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();  // this round's reads before the next round's stores
    }
    __syncthreads();  // every wave is done with the staging arrays the partials overlay

    // subgrid = (M0 + 2 pi i r M1) / vscale, ifftshift-ed (Gridder.cs:76-79;
    // identity taper): each wave writes its part to its own tile, then the
    // four are summed in wave order.  Lane (c, g) holds rows 4 g + e of every
    // tile at column c.
    const float back = 1.f / vscale;
    const float back1 = back / cscale;
    float2* const part = s_buf + wave * S * S;
#pragma unroll
    for (int r = 0; r < NT; r++)
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int y = 16 * r + 4 * grp + e, x = 16 * q + c;
                const float rr = kTwoPiF * s_rt[y * S + x];
                const float m1r = acc1r[r][q][e] * back1, m1i = acc1i[r][q][e] * back1;
                const float gr = acc0r[r][q][e] * back - rr * m1i;
                const float gi = acc0i[r][q][e] * back + rr * m1r;
                part[((y + S / 2) % S) * S + (x + S / 2) % S] = make_float2(gr, gi);
            }
    __syncthreads();
    for (int p = tid; p < S * S; p += kThreads) {
        const float2 p0 = s_buf[p], p1 = s_buf[S * S + p], p2 = s_buf[2 * S * S + p], p3 = s_buf[3 * S * S + p];
        tile0[p] = make_float2(((p0.x + p1.x) + p2.x) + p3.x, ((p0.y + p1.y) + p2.y) + p3.y);
    }
    __syncthreads();
    const float2* res = fft_2d<S>(tile0, tile0 + S * S, s_tw, +1);
    adder<S>(a, g, res, 1.0f / (float)(S * S), slot, s_ap);
}

// ---------------------------------------------------------------------------
// Degridder for S <= 48 on 16x16x32 MFMA tiles.  One workgroup per work item:
// splitter + forward FFT into LDS, then the subgrid pixels, split hi/lo, stay
// in registers as the MFMA A operand (row tiles of 16 y, k = x padded to 32).
// A tile is 16 visibilities (the N dimension): lane l holds the x-phasors
// of visibility l % 16 at x = 8 (l / 16) + e and accumulates output rows
// y = 16 rt + 4 (l / 16) + r, so per lane the state is 8 x-phasors and
// 4 RT y-phasors, small enough for two waves per SIMD.
template <int S, class P>
__device__ __forceinline__ void degrid_mfma_body(const KernelArgs& a, int sidx, int k0, int k1) {
    static_assert(S <= 48, "x in at most two 32-deep k steps, up to 3 row tiles");
    constexpr int RT = (S + 15) / 16;  // row tiles
    constexpr int KS = (S + 31) / 32;  // k steps of 32 along x
    constexpr int NX = 8 * KS;         // x-phasors per lane
    constexpr int NB = 4 * RT;         // y-phasors per lane
    constexpr int NR = (NX + NB) / 2;  // float4 slots of a lane's channel-step phasors
    __shared__ float2 s_tile[2][S * S];
    __shared__ double s_l[32 * KS], s_n1[32 * KS];
    __shared__ float s_red[4];
    // each lane's channel-step phasors (RA, then RB, two per float4), kept in
    // LDS rather than in 2 (8 + NB) registers: read once per channel
    __shared__ float4 s_rstep[NR][kThreads];

    const SubgridGeom g = load_geom(a, sidx, S);
    const SplitK sk = split_k(a);
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;

    splitter<S>(a, g, s_tile[0]);
    if (tid < 32 * KS) {
        s_l[tid] = (double)compute_l(tid, S, a.image_size);
        s_n1[tid] = axis_n(tid, S, a.image_size);
    }
    __syncthreads();
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, -1);

    // normalisers (powers of two): pixels to |P| <= 1 (FP16 only), r to <= 1
    float pscale = 1.f;
    if (P::kNormalise) {
        float m = 0.f;
        for (int p = tid; p < S * S; p += kThreads) m = fmaxf(m, fmaxf(fabsf(res[p].x), fabsf(res[p].y)));
        pscale = 1.f / pow2_ceil(block_max(m, s_red));
    }
    const float rscale = 1.f / pow2_ceil(fabsf(a.rtab[0]));

    // A operand: pixels_copy[y][x] = subgrid[(y+S/2)%S][(x+S/2)%S] (Gridder.cs:101-117)
    // at y = 16 rt + col, x = 32 ks + 8 grp + e; P split hi/lo and r P
    u32x4 Prh[RT][KS], Prl[RT][KS], Pih[RT][KS], Pil[RT][KS], Qr[RT][KS], Qi[RT][KS];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {  // elements 2q, 2q+1
                float2 pv[2];
                float rr[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int y = 16 * rt + col, x = 32 * ks + 8 * grp + 2 * q + u;
                    pv[u] = make_float2(0.f, 0.f);
                    rr[u] = 0.f;
                    if (y < S && x < S) {
                        pv[u] = res[((y + S / 2) % S) * S + (x + S / 2) % S];
                        pv[u].x *= pscale;
                        pv[u].y *= pscale;
                        rr[u] = a.rtab[y * S + x] * rscale;
                    }
                }
                P9_SPLIT2(P, sk, pv[0].x, pv[1].x, Prh[rt][ks][q], Prl[rt][ks][q]);
                P9_SPLIT2(P, sk, pv[0].y, pv[1].y, Pih[rt][ks][q], Pil[rt][ks][q]);
                Qr[rt][ks][q] = P::pack(rr[0] * pv[0].x, rr[1] * pv[1].x);
                Qi[rt][ks][q] = P::pack(rr[0] * pv[0].y, rr[1] * pv[1].y);
            }
        }
    }
    const float out_scale = a.degrid_scale / pscale;
    const float q1_back = kTwoPiF / rscale;
    // Up to S = 32 the products' minus signs sit on loop-invariant copies of
    // -Pr and -rPr, so no tile negates its x-phasors; S = 48 keeps the
    // registers and negates the phasors instead.
    constexpr bool kNegP = RT * KS <= 2;
    u32x4 nPrh[RT][KS], nPrl[RT][KS], nQr[RT][KS];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++)
            if constexpr (kNegP) {
                nPrh[rt][ks] = neg(Prh[rt][ks]);
                nPrl[rt][ks] = neg(Prl[rt][ks]);
                nQr[rt][ks] = neg(Qr[rt][ks]);
            }

    // One tile: Q0 = P conj(A), Q1 = (r P) conj(A) on the MFMA, then
    // vis = sum_y conj(B_v[y]) (Q0 - 2 pi i c Q1) over this lane's rows and
    // the four lane groups.  A: the NX x-phasors, B: the NB y-phasors.
    auto tile = [&](const float2 (&A)[NX], const float2 (&B)[NB], float cf) -> float2 {
        f32x4 q0r[RT], q0i[RT], q1r[RT], q1i[RT];
#pragma unroll
        for (int rt = 0; rt < RT; rt++) q0r[rt] = q0i[rt] = q1r[rt] = q1i[rt] = (f32x4){};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            u32x4 Arh, Arl, Aih, Ail;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                P9_SPLIT2(P, sk, A[8 * ks + 2 * q].x, A[8 * ks + 2 * q + 1].x, Arh[q], Arl[q]);
                P9_SPLIT2(P, sk, A[8 * ks + 2 * q].y, A[8 * ks + 2 * q + 1].y, Aih[q], Ail[q]);
            }
            // Re Q0 = Pr Ar + Pi Ai, Im Q0 = Pi Ar - Pr Ai (hi*hi + hi*lo + lo*hi)
            if constexpr (kNegP) {
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    q0r[rt] = P::mma16(Prh[rt][ks], Arh, q0r[rt]);
                    q0i[rt] = P::mma16(Pih[rt][ks], Arh, q0i[rt]);
                    q1r[rt] = P::mma16(Qr[rt][ks], Arh, q1r[rt]);
                    q1i[rt] = P::mma16(Qi[rt][ks], Arh, q1i[rt]);
                    q0r[rt] = P::mma16(Prh[rt][ks], Arl, q0r[rt]);
                    q0i[rt] = P::mma16(Pih[rt][ks], Arl, q0i[rt]);
                    q1r[rt] = P::mma16(Qi[rt][ks], Aih, q1r[rt]);
                    q1i[rt] = P::mma16(nQr[rt][ks], Aih, q1i[rt]);
                    q0r[rt] = P::mma16(Prl[rt][ks], Arh, q0r[rt]);
                    q0i[rt] = P::mma16(Pil[rt][ks], Arh, q0i[rt]);
                    q0r[rt] = P::mma16(Pih[rt][ks], Aih, q0r[rt]);
                    q0i[rt] = P::mma16(nPrh[rt][ks], Aih, q0i[rt]);
                    q0r[rt] = P::mma16(Pih[rt][ks], Ail, q0r[rt]);
                    q0i[rt] = P::mma16(nPrh[rt][ks], Ail, q0i[rt]);
                    q0r[rt] = P::mma16(Pil[rt][ks], Aih, q0r[rt]);
                    q0i[rt] = P::mma16(nPrl[rt][ks], Aih, q0i[rt]);
                }
            } else {
                const u32x4 nAih = neg(Aih), nAil = neg(Ail);
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    q0r[rt] = P::mma16(Prh[rt][ks], Arh, q0r[rt]);
                    q0i[rt] = P::mma16(Pih[rt][ks], Arh, q0i[rt]);
                    q1r[rt] = P::mma16(Qr[rt][ks], Arh, q1r[rt]);
                    q1i[rt] = P::mma16(Qi[rt][ks], Arh, q1i[rt]);
                    q0r[rt] = P::mma16(Prh[rt][ks], Arl, q0r[rt]);
                    q0i[rt] = P::mma16(Pih[rt][ks], Arl, q0i[rt]);
                    q1r[rt] = P::mma16(Qi[rt][ks], Aih, q1r[rt]);
                    q1i[rt] = P::mma16(Qr[rt][ks], nAih, q1i[rt]);
                    q0r[rt] = P::mma16(Prl[rt][ks], Arh, q0r[rt]);
                    q0i[rt] = P::mma16(Pil[rt][ks], Arh, q0i[rt]);
                    q0r[rt] = P::mma16(Pih[rt][ks], Aih, q0r[rt]);
                    q0i[rt] = P::mma16(Prh[rt][ks], nAih, q0i[rt]);
                    q0r[rt] = P::mma16(Pih[rt][ks], Ail, q0r[rt]);
                    q0i[rt] = P::mma16(Prh[rt][ks], nAil, q0i[rt]);
                    q0r[rt] = P::mma16(Pil[rt][ks], Aih, q0r[rt]);
                    q0i[rt] = P::mma16(Prl[rt][ks], nAih, q0i[rt]);
                }
            }
        }
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float2 b = B[4 * rt + r];
                const float zr = q0r[rt][r] + cf * q1i[rt][r];
                const float zi = q0i[rt][r] - cf * q1r[rt][r];
                sr = fmaf(b.x, zr, fmaf(b.y, zi, sr));
                si = fmaf(b.x, zi, fmaf(-b.y, zr, si));
            }
        }
        sr = sum_xor32(sum_xor16(sr));
        si = sum_xor32(sum_xor16(si));
        return make_float2(sr * out_scale, si * out_scale);
    };
    // x of phasor e and y of phasor i for this lane; phasors outside the
    // subgrid meet zero pixels and are kept at 0 (finite)
    auto xphasors = [&](double ca, double cc, float2 (&A)[NX]) {
#pragma unroll
        for (int e = 0; e < NX; e++) {
            const int x = 32 * (e / 8) + 8 * grp + e % 8;
            A[e] = x < S ? cis_f64(fma(ca, s_l[x], cc * s_n1[x])) : make_float2(0.f, 0.f);
        }
    };
    auto yphasors = [&](double cb, double cc, float2 (&B)[NB]) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int y = 16 * (i / 4) + 4 * grp + (i % 4);
            B[i] = y < S ? cis_f64(fma(cb, s_l[y], cc * s_n1[y])) : make_float2(0.f, 0.f);
        }
    };

    if (a.chan_step) {
        // Channel recurrence.  Column col of every tile is one timestep and
        // consecutive tiles are consecutive channels, so each lane steps its
        // phasors by one multiplication per tile.  A super-tile is 16
        // timesteps x C channels; the (super-tile, channel) sequence is cut
        // into 4 equal contiguous parts, one per wave, and the phasors are
        // evaluated directly at the start of every run and every 8 channels.
        const int tlo = k0 / C, thi = (K + C - 1) / C;
        const int nst = (thi - tlo + 15) / 16;
        const int64_t F = (int64_t)nst * C;
        int64_t f = F * wave / 4;
        const int64_t f_end = F * (wave + 1) / 4;
        while (f < f_end) {
            const int st = (int)(f / C), ch0 = (int)(f % C);
            const int run = (int)min((int64_t)(C - ch0), f_end - f);
            const int tau = tlo + st * 16 + col;  // this lane's timestep in the subgrid
            const double* p = a.uvw + ((size_t)g.bl * a.T + g.t0 + min(tau, g.nt - 1)) * 3;
            const double u = p[0], v = p[1], w = p[2];
#pragma unroll 1
            for (int cb = 0; cb < run; cb += 8) {
                const int ch = ch0 + cb;
                const double k = a.wavenumbers[ch];
                const double dk = a.wavenumbers[ch + 1 < C ? ch + 1 : ch] - k;
                const double cc = (g.w_off - k * w) * kInvTwoPi, dc = -dk * w * kInvTwoPi;
                float2 A[NX], B[NB];
                xphasors((g.u_off - k * u) * kInvTwoPi, cc, A);
                yphasors((g.v_off - k * v) * kInvTwoPi, cc, B);
                {
                    float2 RA[NX], RB[NB];
                    xphasors(-dk * u * kInvTwoPi, dc, RA);
                    yphasors(-dk * v * kInvTwoPi, dc, RB);
#pragma unroll
                    for (int i = 0; i < NX / 2; i++)
                        s_rstep[i][tid] = make_float4(RA[2 * i].x, RA[2 * i].y, RA[2 * i + 1].x, RA[2 * i + 1].y);
#pragma unroll
                    for (int i = 0; i < NB / 2; i++)
                        s_rstep[NX / 2 + i][tid] = make_float4(RB[2 * i].x, RB[2 * i].y, RB[2 * i + 1].x, RB[2 * i + 1].y);
                    // reloaded from LDS below, not forwarded in registers
                    asm volatile("" ::: "memory");
                }
                const float cf0 = (float)cc * q1_back, dcf = (float)dc * q1_back;
                const int ne = min(8, run - cb);
                // every lane holds the sums; lane group grp keeps channels
                // ch + 2 grp and ch + 2 grp + 1, so a timestep's channels are
                // written by two back-to-back stores of whole cache lines
                float2 keep0 = make_float2(0.f, 0.f), keep1 = keep0;
                // two channels per trip (a tail channel beyond ne is computed
                // and discarded), so the stepped phasors alternate between two
                // register sets instead of being copied back every channel
#pragma unroll 1
                for (int e = 0; e < ne; e += 2) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        // the next channel's x-phasors first: the split in
                        // tile() may then overwrite this channel's in place
                        float2 An[NX];
#pragma unroll
                        for (int i = 0; i < NX / 2; i++) {
                            const float4 r = s_rstep[i][tid];
                            An[2 * i] = cmulf(A[2 * i], make_float2(r.x, r.y));
                            An[2 * i + 1] = cmulf(A[2 * i + 1], make_float2(r.z, r.w));
                        }
                        const float2 vis = tile(A, B, fmaf((float)(e + h), dcf, cf0));
                        if (e + h == 2 * grp) keep0 = vis;
                        if (e + h == 2 * grp + 1) keep1 = vis;
#pragma unroll
                        for (int i = 0; i < NX; i++) A[i] = An[i];
#pragma unroll
                        for (int i = 0; i < NB / 2; i++) {
                            const float4 r = s_rstep[NX / 2 + i][tid];
                            B[2 * i] = cmulf(B[2 * i], make_float2(r.x, r.y));
                            B[2 * i + 1] = cmulf(B[2 * i + 1], make_float2(r.z, r.w));
                        }
                    }
                }
                const int j0 = tau * C + ch + 2 * grp;
                if (2 * grp < ne && tau < thi && j0 >= k0 && j0 < K) a.vis_out[vbase + j0] = keep0;
                if (2 * grp + 1 < ne && tau < thi && j0 + 1 >= k0 && j0 + 1 < K) a.vis_out[vbase + j0 + 1] = keep1;
            }
            f += run;
        }
        return;
    }

    // direct evaluation: tiles of 16 consecutive visibilities, round robin
    const int ntiles = (K - k0 + 15) / 16;
    for (int vt = wave; vt < ntiles; vt += 4) {
        const int j = k0 + vt * 16 + col;
        const int jc = j < K ? j : K - 1;
        const int t = g.t0 + jc / C, ch = jc % C;
        const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
        const double k = a.wavenumbers[ch];
        const double cc = (g.w_off - k * p[2]) * kInvTwoPi;
        float2 A[NX], B[NB];
        xphasors((g.u_off - k * p[0]) * kInvTwoPi, cc, A);
        yphasors((g.v_off - k * p[1]) * kInvTwoPi, cc, B);
        const float2 vis = tile(A, B, (float)cc * q1_back);
        if (grp == 0 && j < K) a.vis_out[vbase + j] = vis;
    }
}


// Degridder for S = 64: rows and x span two 32-wide tiles.  The split
// pixel operands (2 row tiles x 4 k-steps of 16) live in LDS, pre-arranged
// in MFMA A-operand lane order; each wave takes tiles of 32 visibilities and
// accumulates both row tiles in turn.
template <int S, class P>
__device__ __forceinline__ void degrid_mfma_wide_body(const KernelArgs& a, int sidx, int k0, int k1) {
    static_assert(S > 32 && S <= 64, "two 32-wide tiles per axis");
    constexpr int RT = 2, KS = 4, NA = 6;  // row tiles, k-steps, operand arrays
    __shared__ float2 s_tile[2][S * S];
    __shared__ u32x4 s_op[RT * KS * NA * 64];
    __shared__ double s_l[64], s_n1[64];
    __shared__ float s_red[4];

    const SubgridGeom g = load_geom(a, sidx, S);
    const SplitK sk = split_k(a);
    const int tid = item_tid(), lane = tid & 63, wave = tid >> 6;
    const int li = lane & 31, grp = lane >> 5;
    const int C = a.C, K = k1 < 0 ? g.nt * C : k1;
    const size_t vbase = ((size_t)g.bl * a.T + g.t0) * C;

    splitter<S>(a, g, s_tile[0]);
    if (tid < 64) {
        s_l[tid] = (double)compute_l(tid, S, a.image_size);
        s_n1[tid] = axis_n(tid, S, a.image_size);
    }
    __syncthreads();
    const float2* res = fft_2d<S>(s_tile[0], s_tile[1], a.twiddle, -1);

    float pscale = 1.f;
    if (P::kNormalise) {
        float m = 0.f;
        for (int p = tid; p < S * S; p += kThreads) m = fmaxf(m, fmaxf(fabsf(res[p].x), fabsf(res[p].y)));
        pscale = 1.f / pow2_ceil(block_max(m, s_red));
    }
    const float rscale = 1.f / pow2_ceil(fabsf(a.rtab[0]));

    // operand (row tile rt, k-step h) for lane ln: row y = 32 rt + ln % 32,
    // x = 16 h + 8 (ln / 32) + e; arrays Pr hi/lo, Pi hi/lo, r P re/im
    for (int q = tid; q < RT * KS * 64; q += kThreads) {
        const int ln = q & 63, h = (q >> 6) % KS, rt = q / (64 * KS);
        const int y = rt * 32 + (ln & 31);
        u32x4 prh, prl, pih, pil, qr, qi;
#pragma unroll
        for (int w = 0; w < 4; w++) {  // elements 2w, 2w+1
            float2 pv[2];
            float rr[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int x = 16 * h + 8 * (ln >> 5) + 2 * w + u;
                pv[u] = make_float2(0.f, 0.f);
                rr[u] = 0.f;
                if (y < S && x < S) {
                    pv[u] = res[((y + S / 2) % S) * S + (x + S / 2) % S];
                    pv[u].x *= pscale;
                    pv[u].y *= pscale;
                    rr[u] = a.rtab[y * S + x] * rscale;
                }
            }
            P9_SPLIT2(P, sk, pv[0].x, pv[1].x, prh[w], prl[w]);
            P9_SPLIT2(P, sk, pv[0].y, pv[1].y, pih[w], pil[w]);
            qr[w] = P::pack(rr[0] * pv[0].x, rr[1] * pv[1].x);
            qi[w] = P::pack(rr[0] * pv[0].y, rr[1] * pv[1].y);
        }
        u32x4* op = &s_op[((rt * KS + h) * NA) * 64 + ln];
        op[0 * 64] = prh;
        op[1 * 64] = prl;
        op[2 * 64] = pih;
        op[3 * 64] = pil;
        op[4 * 64] = qr;
        op[5 * 64] = qi;
    }
    __syncthreads();
    const float out_scale = a.degrid_scale / pscale;
    const float q1_back = kTwoPiF / rscale;

    const int ntiles = (K - k0 + 31) / 32;
    for (int vt = wave; vt < ntiles; vt += 4) {
        const int j = k0 + vt * 32 + li;
        const int jc = j < K ? j : K - 1;
        const int t = g.t0 + jc / C, ch = jc % C;
        const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
        const double k = a.wavenumbers[ch];
        const double ca = (g.u_off - k * p[0]) * kInvTwoPi;
        const double cb = (g.v_off - k * p[1]) * kInvTwoPi;
        const double cc = (g.w_off - k * p[2]) * kInvTwoPi;

        u32x4 Arh[KS], Arl[KS], Aih[KS], Ail[KS];
#pragma unroll
        for (int h = 0; h < KS; h++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float2 A[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int x = 16 * h + 8 * grp + 2 * q + u;
                    // padded columns (x >= S) meet zero pixels: keep them finite
                    A[u] = x < S ? cis_f64(fma(ca, s_l[x], cc * s_n1[x])) : make_float2(0.f, 0.f);
                }
                P9_SPLIT2(P, sk, A[0].x, A[1].x, Arh[h][q], Arl[h][q]);
                P9_SPLIT2(P, sk, A[0].y, A[1].y, Aih[h][q], Ail[h][q]);
            }
        }
        const float cf = (float)cc * q1_back;
        float sr = 0.f, si = 0.f;
#pragma unroll 1
        for (int rt = 0; rt < RT; rt++) {
            f32x16 q0r = {}, q0i = {}, q1r = {}, q1i = {};
#pragma unroll
            for (int h = 0; h < KS; h++) {
                const u32x4* op = &s_op[((rt * KS + h) * NA) * 64 + lane];
                const u32x4 prh = op[0 * 64], prl = op[1 * 64], pih = op[2 * 64], pil = op[3 * 64];
                const u32x4 qr = op[4 * 64], qi = op[5 * 64];
                const u32x4 nAih = neg(Aih[h]), nAil = neg(Ail[h]);
                // Re Q0 = Pr Ar + Pi Ai, Im Q0 = Pi Ar + Pr (-Ai)
                q0r = P::mma(prh, Arh[h], q0r);
                q0i = P::mma(pih, Arh[h], q0i);
                q1r = P::mma(qr, Arh[h], q1r);
                q1i = P::mma(qi, Arh[h], q1i);
                q0r = P::mma(prh, Arl[h], q0r);
                q0i = P::mma(pih, Arl[h], q0i);
                q1r = P::mma(qi, Aih[h], q1r);
                q1i = P::mma(qr, nAih, q1i);
                q0r = P::mma(prl, Arh[h], q0r);
                q0i = P::mma(pil, Arh[h], q0i);
                q0r = P::mma(pih, Aih[h], q0r);
                q0i = P::mma(prh, nAih, q0i);
                q0r = P::mma(pih, Ail[h], q0r);
                q0i = P::mma(prh, nAil, q0i);
                q0r = P::mma(pil, Aih[h], q0r);
                q0i = P::mma(prl, nAih, q0i);
            }
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int y = rt * 32 + 8 * (r / 4) + 4 * grp + (r % 4);
                if (y < S) {
                    const float2 B = cis_f64(fma(cb, s_l[y], cc * s_n1[y]));
                    const float zr = q0r[r] + cf * q1i[r];
                    const float zi = q0i[r] - cf * q1r[r];
                    sr = fmaf(B.x, zr, fmaf(B.y, zi, sr));
                    si = fmaf(B.x, zi, fmaf(-B.y, zr, si));
                }
            }
        }
        sr = sum_xor32(sr);
        si = sum_xor32(si);
        if (grp == 0 && j < K) a.vis_out[vbase + j] = make_float2(sr * out_scale, si * out_scale);
    }
}

// two waves per SIMD up to S = 48; the S = 64 gridder needs more registers
template <int S>
constexpr int kGridWaves = S <= 32 ? 2 : 1;
// multiples of 16 up to 48 on 16x16x32 tiles (as many as the subgrid holds),
// S = 8, 24 and 64 on 32x32x16
// fresh: as for grid_mfma16_body
template <int S, class P>
__device__ __forceinline__ void grid_mfma_any(const KernelArgs& a, int s, int k0, int k1, int slot, bool fresh) {
    if constexpr (S % 16 == 0 && S <= 48)
        grid_mfma16_body<S, P>(a, s, k0, k1, slot, fresh);
    else
        grid_mfma_body<S, P>(a, s, k0, k1, slot);
}
template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kGridWaves<S>))) void grid_mfma_kernel(KernelArgs a) {
    grid_mfma_any<S, P>(a, a.order[blockIdx.x], 0, -1, a.order[blockIdx.x], true);
}
template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kGridWaves<S>))) void grid_mfma_queue_kernel(KernelArgs a) {
    bool fresh = true;
    run_queue<true>(a, [&](int s, int k0, int k1, int slot) {
        grid_mfma_any<S, P>(a, s, k0, k1, slot, fresh);
        fresh = false;
    });
}

// S <= 32: 16x16x32 tiles, two waves per SIMD; S = 48: 3 row tiles and two
// k steps of 16x16x32, one wave per SIMD; S = 64: the wide form
template <int S>
constexpr int kDegridWaves = S <= 32 ? 2 : 1;
template <int S, class P>
__device__ __forceinline__ void degrid_mfma_any(const KernelArgs& a, int s, int k0, int k1) {
    if constexpr (S <= 48)
        degrid_mfma_body<S, P>(a, s, k0, k1);
    else
        degrid_mfma_wide_body<S, P>(a, s, k0, k1);
}
template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kDegridWaves<S>))) void degrid_mfma_kernel(KernelArgs a) {
    degrid_mfma_any<S, P>(a, a.order[blockIdx.x], 0, -1);
}
template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kDegridWaves<S>))) void degrid_mfma_queue_kernel(KernelArgs a) {
    run_queue(a, [&](int s, int k0, int k1, int) { degrid_mfma_any<S, P>(a, s, k0, k1); });
}

template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kGridWaves<S>))) void grid_mfma_batch_kernel(BatchArgs b) {
    // the tables are per plan: reload them when the item's plan changes
    const float* loaded = nullptr;
    run_batch_queue<true>(b, [&](const KernelArgs& a, int s, int k0, int k1, int slot) {
        grid_mfma_any<S, P>(a, s, k0, k1, slot, a.rtab != loaded);
        loaded = a.rtab;
    });
}
template <int S, class P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kDegridWaves<S>))) void degrid_mfma_batch_kernel(BatchArgs b) {
    run_batch_queue(b, [&](const KernelArgs& a, int s, int k0, int k1, int) { degrid_mfma_any<S, P>(a, s, k0, k1); });
}

}  // namespace dev

#define P9_LAUNCH_MFMA(KERNEL, QKERNEL, SV, PREC)                                                  \
    do {                                                                                           \
        if (cfg.queue) {                                                                           \
            int wg = cfg.workgroups > 0                                                            \
                         ? cfg.workgroups                                                          \
                         : persistent_workgroups((const void*)QKERNEL<SV, PREC>, dev::kThreads);   \
            wg = wg < a.nitems ? wg : a.nitems;                                                    \
            if (wg <= 0) return hipSuccess;                                                        \
            hipLaunchKernelGGL((QKERNEL<SV, PREC>), dim3(wg), dim3(dev::kThreads), 0, st, a);      \
        } else {                                                                                   \
            hipLaunchKernelGGL((KERNEL<SV, PREC>), dim3(a.nr_subgrids), dim3(dev::kThreads), 0, st, a); \
        }                                                                                          \
    } while (0)

#define P9_MFMA_SWITCH(KERNEL, QKERNEL)                                              \
    switch (S) {                                                                     \
        case 8: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 8, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 8, dev::PrecBF16); break;    \
        case 16: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 16, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 16, dev::PrecBF16); break; \
        case 24: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 24, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 24, dev::PrecBF16); break; \
        case 32: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 32, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 32, dev::PrecBF16); break; \
        case 48: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 48, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 48, dev::PrecBF16); break; \
        case 64: if (f16) P9_LAUNCH_MFMA(KERNEL, QKERNEL, 64, dev::PrecF16); else P9_LAUNCH_MFMA(KERNEL, QKERNEL, 64, dev::PrecBF16); break; \
        default: return hipErrorNotSupported;                                        \
    }

hipError_t launch_rtab(float* rtab, int S, float image_size, hipStream_t st) {
    hipLaunchKernelGGL(dev::rtab_kernel, dim3((S * S + 255) / 256), dim3(256), 0, st, rtab, S, image_size);
    return hipGetLastError();
}

hipError_t launch_grid_mfma(const KernelArgs& a, int S, int precision, const LaunchCfg& cfg,
                            hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
    const bool f16 = precision == P9IDG_KERNEL_MFMA_FP16;
    P9_MFMA_SWITCH(dev::grid_mfma_kernel, dev::grid_mfma_queue_kernel)
    return hipGetLastError();
}

namespace {
template <class K>
hipError_t launch_batch_kernel(K kernel, const BatchArgs& b, int workgroups, hipStream_t st) {
    int wg = workgroups > 0 ? workgroups : persistent_workgroups((const void*)kernel, dev::kThreads);
    wg = wg < b.nitems ? wg : b.nitems;
    if (wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(wg), dim3(dev::kThreads), 0, st, b);
    return hipGetLastError();
}

template <int S>
hipError_t batch_dispatch(bool grid, const BatchArgs& b, int kernel, int wg, hipStream_t st) {
    if (kernel == P9IDG_KERNEL_FP32 || kernel == P9IDG_KERNEL_FP32_SEP)
        return launch_fp32_batch(grid, kernel == P9IDG_KERNEL_FP32_SEP, b, S, wg, st);
    if (kernel == P9IDG_KERNEL_MFMA_FP16)
        return grid ? launch_batch_kernel(dev::grid_mfma_batch_kernel<S, dev::PrecF16>, b, wg, st)
                    : launch_batch_kernel(dev::degrid_mfma_batch_kernel<S, dev::PrecF16>, b, wg, st);
    return grid ? launch_batch_kernel(dev::grid_mfma_batch_kernel<S, dev::PrecBF16>, b, wg, st)
                : launch_batch_kernel(dev::degrid_mfma_batch_kernel<S, dev::PrecBF16>, b, wg, st);
}

hipError_t batch_switch(bool grid, const BatchArgs& b, int S, int kernel, int wg, hipStream_t st) {
    switch (S) {
        case 8: return batch_dispatch<8>(grid, b, kernel, wg, st);
        case 16: return batch_dispatch<16>(grid, b, kernel, wg, st);
        case 24: return batch_dispatch<24>(grid, b, kernel, wg, st);
        case 32: return batch_dispatch<32>(grid, b, kernel, wg, st);
        case 48: return batch_dispatch<48>(grid, b, kernel, wg, st);
        case 64: return batch_dispatch<64>(grid, b, kernel, wg, st);
        default: return hipErrorNotSupported;
    }
}
}  // namespace

hipError_t launch_grid_batch(const BatchArgs& b, int S, int kernel, int workgroups, hipStream_t st) {
    return batch_switch(true, b, S, kernel, workgroups, st);
}
hipError_t launch_degrid_batch(const BatchArgs& b, int S, int kernel, int workgroups, hipStream_t st) {
    return batch_switch(false, b, S, kernel, workgroups, st);
}

hipError_t launch_degrid_mfma(const KernelArgs& a, int S, int precision, const LaunchCfg& cfg,
                              hipStream_t st) {
    if (a.nr_subgrids <= 0) return hipSuccess;
    const bool f16 = precision == P9IDG_KERNEL_MFMA_FP16;
    P9_MFMA_SWITCH(dev::degrid_mfma_kernel, dev::degrid_mfma_queue_kernel)
    return hipGetLastError();
}

}  // namespace p9idg
