// Device-side building blocks shared by the gridder and degridder kernels:
// per-subgrid geometry, the LDS subgrid FFT, the adder (scatter onto the
// grid / w-layers) and the splitter (gather from the grid).
#pragma once

#include <hip/hip_runtime.h>

#include "p9idg_internal.h"

namespace p9idg {
namespace dev {

constexpr int kThreads = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// threadIdx.x behind an empty volatile asm: values derived from it are
// recomputed where they are used instead of being hoisted out of the
// persistent work-item loop and held (or spilled) across items.
__device__ __forceinline__ int item_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Gridder.cs:163-166, evaluated in float as the reference does.
__device__ __forceinline__ float compute_l(int x, int S, float image_size) {
    float t = (float)(x - S / 2) + 0.5f;
    t = __fmul_rn(t, image_size);
    return __fdiv_rn(t, (float)S);
}
// Gridder.cs:168-174.  The reference evaluates tmp / 1.0f + Sqrt(1.0 - tmp),
// so n = t / (t + sqrt(1 - t)) with t = l^2 + m^2 (about twice
// 1 - sqrt(1 - t) for small t); kept as written.  The sqrt is taken in double.
__device__ __forceinline__ float compute_n(float l, float m) {
    float tmp = __fadd_rn(__fmul_rn(l, l), __fmul_rn(m, m));
    double d = (double)tmp + sqrt(1.0 - (double)tmp);
    return __fdiv_rn(tmp, (float)d);
}

// Turns -> (cos, sin) of 2*pi*turns using the hardware transcendental unit
// on the argument reduced to [-0.5, 0.5].
__device__ __forceinline__ float2 cis_turns(float turns) {
    float r = turns - rintf(turns);
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

// Per-subgrid phase offsets of Gridder.cs:38-41 (FP64).
struct SubgridGeom {
    int bl, t0, nt, U, V, W;
    double u_off, v_off, w_off;
};

__device__ __forceinline__ SubgridGeom load_geom(const KernelArgs& a, int s, int S) {
    const p9idg_subgrid m = a.meta[s];
    SubgridGeom g;
    g.bl = m.baseline_idx;
    g.t0 = m.time_sample_start;
    g.nt = m.time_sample_count;
    g.U = m.u_pixel;
    g.V = m.v_pixel;
    g.W = m.w_lambda;
    g.u_off = (double)(g.U + S / 2 - a.G / 2) * a.u_scale;
    g.v_off = (double)(g.V + S / 2 - a.G / 2) * a.u_scale;
    g.w_off = a.w_offset_scale * ((double)g.W + 0.5);
    return g;
}

// Phase coefficients of one visibility, in turns per unit of l, m, n:
// phase = off - k (u l + v m + w n)  =  l (uoff - k u) + m (voff - k v) + n (woff - k w)
// (Gridder.cs:62-66).  Computed in FP64; the per-pixel evaluation is FP32.
__device__ __forceinline__ float3 vis_coeffs(const KernelArgs& a, const SubgridGeom& g, int t,
                                             int ch) {
    const double* p = a.uvw + ((size_t)g.bl * a.T + t) * 3;
    const double k = a.wavenumbers[ch];
    const double inv = 1.0 / kTwoPi;
    return make_float3((float)((g.u_off - k * p[0]) * inv), (float)((g.v_off - k * p[1]) * inv),
                       (float)((g.w_off - k * p[2]) * inv));
}

// ---------------------------------------------------------------------------
// LDS 2-D FFT of an S x S complex tile (row-major), Stockham autosort,
// radices 4, 2 and 3.  sign = +1 is FFTW's BACKWARD transform, -1 FORWARD.
// tw[m] = exp(+2 pi i m / S).  Result ends in `a` or `b`; returns it.
template <int R>
__device__ __forceinline__ void dft_small(float2* v, int sign) {
    if (R == 2) {
        float2 t0 = v[0], t1 = v[1];
        v[0] = make_float2(t0.x + t1.x, t0.y + t1.y);
        v[1] = make_float2(t0.x - t1.x, t0.y - t1.y);
    } else if (R == 4) {
        float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        // (sign * i) * d13
        float2 id13 = make_float2(-sign * d13.y, sign * d13.x);
        v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
        v[1] = make_float2(d02.x + id13.x, d02.y + id13.y);
        v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
        v[3] = make_float2(d02.x - id13.x, d02.y - id13.y);
    } else if (R == 3) {
        const float h = 0.86602540378443864676f;  // sqrt(3)/2
        float2 t = make_float2(v[1].x + v[2].x, v[1].y + v[2].y);
        float2 d = make_float2(v[1].x - v[2].x, v[1].y - v[2].y);
        float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        // (sign * i * h) * d
        float2 id = make_float2(-sign * h * d.y, sign * h * d.x);
        v[0] = make_float2(v[0].x + t.x, v[0].y + t.y);
        v[1] = make_float2(m.x + id.x, m.y + id.y);
        v[2] = make_float2(m.x - id.x, m.y - id.y);
    }
}

// One Stockham pass of radix R over all S lines of the tile.  Line b's
// element i lives at b*line_stride + i*elem_stride.  When the lines are the
// columns (line_stride 1), consecutive threads take consecutive lines so a
// wave's LDS accesses are contiguous instead of S elements apart.
template <int S, int R>
__device__ __forceinline__ void fft_pass(const float2* __restrict__ in, float2* __restrict__ out,
                                         const float2* tw, int Ns, int sign, int line_stride,
                                         int elem_stride) {
    constexpr int M = S / R;
    for (int idx = item_tid(); idx < S * M; idx += kThreads) {
        const bool cols = line_stride == 1;
        const int b = cols ? idx % S : idx / M;
        const int j = cols ? idx / S : idx - b * M;
        const int k = j % Ns;
        float2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            v[q] = in[b * line_stride + (j + q * M) * elem_stride];
            if (q > 0 && Ns > 1) {
                float2 w = tw[(q * k * (S / (Ns * R))) % S];
                if (sign < 0) w.y = -w.y;
                v[q] = cmulf(v[q], w);
            }
        }
        dft_small<R>(v, sign);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) out[b * line_stride + (base + q * Ns) * elem_stride] = v[q];
    }
}

// Radix plan per supported subgrid size.
template <int S>
struct Radices;
template <>
struct Radices<8> {
    static constexpr int n = 2;
    static constexpr int r[3] = {4, 2, 0};
};
template <>
struct Radices<16> {
    static constexpr int n = 2;
    static constexpr int r[3] = {4, 4, 0};
};
template <>
struct Radices<24> {
    static constexpr int n = 3;
    static constexpr int r[3] = {4, 2, 3};
};
template <>
struct Radices<32> {
    static constexpr int n = 3;
    static constexpr int r[3] = {4, 4, 2};
};
template <>
struct Radices<48> {
    static constexpr int n = 3;
    static constexpr int r[3] = {4, 4, 3};
};
template <>
struct Radices<64> {
    static constexpr int n = 3;
    static constexpr int r[3] = {4, 4, 4};
};

template <int S>
__device__ __forceinline__ void fft_1d_all(float2*& src, float2*& dst, const float2* tw, int sign,
                                           int line_stride, int elem_stride) {
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < Radices<S>::n; p++) {
        const int R = Radices<S>::r[p];
        if (R == 4)
            fft_pass<S, 4>(src, dst, tw, Ns, sign, line_stride, elem_stride);
        else if (R == 2)
            fft_pass<S, 2>(src, dst, tw, Ns, sign, line_stride, elem_stride);
        else
            fft_pass<S, 3>(src, dst, tw, Ns, sign, line_stride, elem_stride);
        Ns *= R;
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
    }
}

// 2-D transform of the tile in `a` (scratch `b`); returns the buffer holding
// the result.  Caller must __syncthreads() before (tile complete).
template <int S>
__device__ __forceinline__ float2* fft_2d(float2* a, float2* b, const float2* tw, int sign) {
    float2* src = a;
    float2* dst = b;
    fft_1d_all<S>(src, dst, tw, sign, S, 1);  // rows
    fft_1d_all<S>(src, dst, tw, sign, 1, S);  // columns
    return src;
}

// Persistent work queues.  counter[0] hands out item indices, counter[1]
// counts the workgroups that found the queue empty; the last of them sets both
// back to zero, so consecutive launches on the same counters need no reset
// (every claim of a workgroup returns before it counts itself out).
__device__ __forceinline__ void queue_exit(int32_t* counter) {
    if (threadIdx.x == 0 && atomicAdd(counter + 1, 1) == (int)gridDim.x - 1) {
        __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(counter + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Each workgroup takes the next WorkItem from the counter until every item is
// taken; body(sidx, k0, k1) processes one.  All threads of the workgroup
// leave together once the counter passes nitems.
// AHEAD: the next item is claimed as soon as the current one is known, so the
// counter's round trip overlaps the current item's work; the workgroup then
// holds one claimed item while it works, which can lengthen the tail (on the
// GPU the MFMA gridder gains from it and the degridder does not).
template <bool AHEAD = false, class Body>
__device__ __forceinline__ void run_queue(const KernelArgs& a, Body body) {
    __shared__ int s_item;
    if (AHEAD) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
        __syncthreads();
        int it = s_item;
        while (it < a.nitems) {  // uniform
            const WorkItem w = a.items[it];
            __syncthreads();  // every thread has read s_item
            if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
            body(w.sidx, w.k0, w.k1, w.slot);
            __syncthreads();  // this item's LDS reads are done; s_item is the next
            it = s_item;
        }
    } else {
        for (;;) {
            __syncthreads();  // previous item's LDS reads are done
            if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
            __syncthreads();
            const int it = s_item;
            if (it >= a.nitems) break;
            const WorkItem w = a.items[it];
            body(w.sidx, w.k0, w.k1, w.slot);
        }
    }
    queue_exit(a.counter);
}

// Batched queue: like run_queue, with each item's observation in args[inst].
template <bool AHEAD = false, class Body>
__device__ __forceinline__ void run_batch_queue(const BatchArgs& b, Body body) {
    __shared__ int s_item;
    if (AHEAD) {
        if (threadIdx.x == 0) s_item = atomicAdd(b.counter, 1);
        __syncthreads();
        int it = s_item;
        while (it < b.nitems) {  // uniform
            const WorkItem w = b.items[it];
            __syncthreads();  // every thread has read s_item
            if (threadIdx.x == 0) s_item = atomicAdd(b.counter, 1);
            const KernelArgs a = b.args[w.inst];
            body(a, w.sidx, w.k0, w.k1, w.slot);
            __syncthreads();  // this item's LDS reads are done; s_item is the next
            it = s_item;
        }
    } else {
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) s_item = atomicAdd(b.counter, 1);
            __syncthreads();
            const int it = s_item;
            if (it >= b.nitems) break;
            const WorkItem w = b.items[it];
            const KernelArgs a = b.args[w.inst];
            body(a, w.sidx, w.k0, w.k1, w.slot);
        }
    }
    queue_exit(b.counter);
}

// Linear index of grid element (y, x) after FFT.Shift (FFT.cs:248-264): the
// h x h blocks (h = G / 2) swap 1 <-> 3 and 2 <-> 4; for odd G the last row
// and the last column stay in place.  The map is its own inverse.
__device__ __forceinline__ size_t shift_index(int y, int x, int G) {
    const int h = G / 2;
    if ((G & 1) && (y == 2 * h || x == 2 * h)) return (size_t)y * G + x;
    return (size_t)(y >= h ? y - h : y + h) * G + (x >= h ? x - h : x + h);
}

}  // namespace dev
}  // namespace p9idg
