// Serial coordinate descent (FastSerialCD.cs) and the PSF correlation pieces
// (PaddedConvolver.cs, Common.cs:48-194) on the GPU.
//
// FastSerialCD picks, per iteration, the pixel whose elastic-net proximal
// update changes it most (first row, then first column on ties), writes it
// and adds psf2 * (old - new) to the gradients in the PSF^2 window.  Only the
// window's pixels change, so the argmax is kept as a table of per-tile bests:
// an iteration re-evaluates the tiles the window touches and reduces the
// table.  Float expressions use one rounding per operation (no contraction),
// as the reference's float code does (this file is built with
// -ffp-contract=off).
#include <climits>

#include "cd_internal.h"

namespace p9idg {
namespace cd {

namespace {

constexpr int kThreads = 256;
constexpr int kSelectThreads = 1024;
static_assert(kTile * kTile % kThreads == 0, "whole pixels per thread");

// Common.cs:27-29; Math.Max(v, 0f) gives +0 for -0
__device__ __forceinline__ float prox(float x, float lipschitz, float lambda, float alpha) {
    float num = __fsub_rn(x, __fmul_rn(lambda, alpha));
    num = num > 0.f ? num : 0.f;
    const float den = __fadd_rn(lipschitz, __fmul_rn(lambda, __fsub_rn(1.f, alpha)));
    return __fdiv_rn(num, den);
}

// Larger |diff| wins; equal nonzero |diff|: the smaller pixel index (the
// reference's row-then-column scan keeps the first maximum).
__device__ __forceinline__ bool better(const Cand& a, const Cand& b) {
    return a.diff > b.diff || (a.diff == b.diff && a.diff > 0.f && a.idx < b.idx);
}


// One step of a wave-wide best over DPP lane routes: each lane compares with
// the candidate CTRL routes to it; lanes outside ROW_MASK (rows of 16 lanes)
// keep their own.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ Cand dpp_best(const Cand& c) {
    Cand o;
    o.diff = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, c.diff),
                                                                   __builtin_bit_cast(int, c.diff), CTRL, ROW_MASK,
                                                                   0xF, false));
    o.idx = __builtin_amdgcn_update_dpp(c.idx, c.idx, CTRL, ROW_MASK, 0xF, false);
    o.xnew = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, c.xnew),
                                                                   __builtin_bit_cast(int, c.xnew), CTRL, ROW_MASK,
                                                                   0xF, false));
    o.xold = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, c.xold),
                                                                   __builtin_bit_cast(int, c.xold), CTRL, ROW_MASK,
                                                                   0xF, false));
    return better(o, c) ? o : c;
}
// The wave's best candidate, in every lane: lane pairs, quads (quad
// permutes), 8 and 16 lanes (half-row and row mirrors), then rows 0 + 1 into
// row 1 and rows 2 + 3 into row 3 (row_bcast15), and rows 0-1 into rows 2-3
// (row_bcast31), so lane 63 holds the best of all 64.  better() orders every
// candidate with a positive difference strictly, so that best does not depend
// on the order of the comparisons.
__device__ __forceinline__ Cand wave_reduce(Cand c) {
    c = dpp_best<0xB1, 0xF>(c);   // quad_perm [1, 0, 3, 2]
    c = dpp_best<0x4E, 0xF>(c);   // quad_perm [2, 3, 0, 1]
    c = dpp_best<0x141, 0xF>(c);  // row_half_mirror
    c = dpp_best<0x140, 0xF>(c);  // row_mirror
    c = dpp_best<0x142, 0xA>(c);  // row_bcast15
    c = dpp_best<0x143, 0xC>(c);  // row_bcast31
    Cand r;
    r.diff = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c.diff), 63));
    r.idx = __builtin_amdgcn_readlane(c.idx, 63);
    r.xnew = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c.xnew), 63));
    r.xold = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c.xold), 63));
    return r;
}
template <int NT>
__device__ Cand block_best(Cand c, Cand* sh) {
    c = wave_reduce(c);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; w++)
            if (better(sh[w], c)) c = sh[w];
        sh[0] = c;
    }
    __syncthreads();
    return sh[0];
}

__device__ __forceinline__ Cand none() { return Cand{0.f, INT_MAX, 0.f, 0.f}; }

// FastSerialCD.cs:150-162 for one pixel
__device__ __forceinline__ Cand eval(const float* x, const float* g, const float* amap, int idx, float lambda,
                                     float alpha) {
    const float a = amap[idx], old = x[idx];
    const float t = prox(__fadd_rn(__fmul_rn(old, a), g[idx]), a, lambda, alpha);
    return Cand{fabsf(__fsub_rn(old, t)), idx, t, old};
}

__device__ void tile_best(const float* x, const float* g, const float* amap, int H, int W, int ty, int tx,
                          float lambda, float alpha, Cand* tiles, int ntx) {
    __shared__ Cand sh[kThreads / 64];
    Cand best = none();
    for (int p = threadIdx.x; p < kTile * kTile; p += kThreads) {
        const int y = ty * kTile + p / kTile, xx = tx * kTile + p % kTile;
        if (y < H && xx < W) {
            const Cand c = eval(x, g, amap, y * W + xx, lambda, alpha);
            if (better(c, best)) best = c;
        }
    }
    best = block_best<kThreads>(best, sh);
    if (threadIdx.x == 0) tiles[ty * ntx + tx] = best;
}

__global__ __launch_bounds__(kThreads) void tiles_all_kernel(const float* x, const float* g, const float* amap,
                                                             int H, int W, float lambda, float alpha, Cand* tiles) {
    tile_best(x, g, amap, H, W, blockIdx.y, blockIdx.x, lambda, alpha, tiles, gridDim.x);
}

// FastSerialCD.cs:132-135 (GetAbsMaxDiff): the best |diff| of the tile table.
__global__ __launch_bounds__(kSelectThreads) void absmax_kernel(const Cand* tiles, int ntiles, State* st) {
    __shared__ Cand sh[kSelectThreads / 64];
    Cand best = none();
    for (int i = threadIdx.x; i < ntiles; i += kSelectThreads)
        if (better(tiles[i], best)) best = tiles[i];
    best = block_best<kSelectThreads>(best, sh);
    if (threadIdx.x == 0) st->absmax = best.diff;
}

// Common.cs:146-160: flipped PSF at (y + ypad/2 + 1, x + xpad/2 + 1), then
// FFT.Shift (quadrant swap, self-inverse for even N0 = N1)
__global__ void place_correlation_kernel(const float* psf, int ph, int pw, int ypad, int xpad, int N0, int N1,
                                         double2* buf) {
    const size_t n = (size_t)N0 * N1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(i / N1), b = (int)(i % N1);
        const int h = N0 / 2;
        // quadrant swap of the first h x h block pairs (FFT.cs:248-264)
        int p = a, q = b;
        if (a < 2 * h && b < 2 * h) {
            p = a < h ? a + h : a - h;
            q = b < h ? b + h : b - h;
        }
        const int y = p - ypad / 2 - 1, x = q - xpad / 2 - 1;
        double v = 0.0;
        if (y >= 0 && y < ph && x >= 0 && x < pw) v = (double)psf[(ph - y - 1) * pw + (pw - x - 1)];
        buf[i] = make_double2(v, 0.0);
    }
}

// Common.cs:164-172 (SetPsfInWindow): psf at (i + ph/2, j + pw/2)
__global__ void place_centered_kernel(const float* psf, int ph, int pw, int N0, int N1, double2* buf) {
    const size_t n = (size_t)N0 * N1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / N1) - ph / 2, x = (int)(i % N1) - pw / 2;
        const double v = (y >= 0 && y < ph && x >= 0 && x < pw) ? (double)psf[y * pw + x] : 0.0;
        buf[i] = make_double2(v, 0.0);
    }
}

// PaddedConvolver.cs:95-114
__global__ void insert_image_kernel(const float* img, int H, int W, int yh, int xh, int N0, int N1, double2* buf) {
    const size_t n = (size_t)N0 * N1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / N1) - yh, x = (int)(i % N1) - xh;
        const double v = (y >= 0 && y < H && x >= 0 && x < W) ? (double)img[(size_t)y * W + x] : 0.0;
        buf[i] = make_double2(v, 0.0);
    }
}

__global__ void multiply_kernel(double2* buf, const double2* k, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double2 a = buf[i], b = k[i];
        buf[i] = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
}

// PaddedConvolver.cs:58-64
__global__ void extract_kernel(const double2* buf, int N1, int yh, int xh, double norm, int H, int W, float* out) {
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i % W);
        out[i] = (float)(buf[(size_t)(y + yh) * N1 + (x + xh)].x / norm);
    }
}

// Common.cs:88-126 (QueryScan) on the prefix sums of psf^2
__global__ void amap_kernel(const double* s, int ph, int pw, int H, int W, float* amap) {
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i % W);
        const int y_over = max(0, y + (ph - ph / 2) - H), x_over = max(0, x + (pw - pw / 2) - W);
        const int y_under = -(y - ph / 2), x_under = -(x - pw / 2);
        const int yu = max(1, y_under) - 1, xu = max(1, x_under) - 1;
        auto S = [&](int a, int b) { return s[a * pw + b]; };
        double v;
        if (y_over <= 0 && x_over <= 0 && y_under <= 0 && x_under <= 0) {
            v = S(ph - 1, pw - 1);
        } else if (y_under > 0 && x_under > 0) {
            v = S(ph - 1, pw - 1) - S(ph - 1, xu) - S(yu, pw - 1) + S(yu, xu);
        } else {
            double corr = 0.0;
            if (y_under > 0) corr += S(yu, pw - x_over - 1);
            if (x_under > 0) corr += S(ph - y_over - 1, xu);
            v = S(ph - 1 - y_over, pw - 1 - x_over) - corr;
        }
        amap[i] = (float)v;
    }
}

// --- one launch per iteration ------------------------------------------------
// Best of tiles[first .. first + n) by one wave (64 lanes).
__device__ Cand wave_best(const Cand* v, int n, int lane) {
    Cand b = none();
    for (int i = lane; i < n; i += 64)
        if (better(v[i], b)) b = v[i];
    return wave_reduce(b);
}

// Rows [ty0, ty1] of the row table from the tile table, then the overall best
// (4 waves), and the selection of FastSerialCD.cs:84-97 by thread 0.
// A wave takes rows ty0 + wave, + 4, + 8, ...; four of its rows at a time have
// their tiles loaded before any is reduced, so the loads overlap (the
// comparisons run in the same order as one row at a time).
__device__ __forceinline__ void rows_and_select(const Cand* __restrict__ tiles, Cand* __restrict__ rows, int ntx, int nty, int ty0,
                                int ty1, State* st, float* x) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kWaves = kThreads / 64, kBatch = 4;
    for (int ty = ty0 + wave; ty <= ty1; ty += kWaves * kBatch) {
        Cand b[kBatch];
        if (ntx <= 128) {  // two tiles per lane at most: all loads first
            Cand v0[kBatch], v1[kBatch];
#pragma unroll
            for (int r = 0; r < kBatch; r++) {
                const int t = ty + r * kWaves;
                const Cand* row = tiles + (size_t)t * ntx;
                v0[r] = t <= ty1 && lane < ntx ? row[lane] : none();
                v1[r] = t <= ty1 && lane + 64 < ntx ? row[lane + 64] : none();
            }
#pragma unroll
            for (int r = 0; r < kBatch; r++) {
                b[r] = none();
                if (better(v0[r], b[r])) b[r] = v0[r];
                if (better(v1[r], b[r])) b[r] = v1[r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < kBatch; r++) {
                const int t = ty + r * kWaves;
                b[r] = none();
                if (t <= ty1)
                    for (int i = lane; i < ntx; i += 64)
                        if (better(tiles[(size_t)t * ntx + i], b[r])) b[r] = tiles[(size_t)t * ntx + i];
            }
        }
#pragma unroll
        for (int r = 0; r < kBatch; r++) {
            b[r] = wave_reduce(b[r]);
            const int t = ty + r * kWaves;
            if (lane == 0 && t <= ty1) rows[t] = b[r];
        }
    }
    __threadfence_block();
    __syncthreads();
    if (wave != 0) return;
    const Cand best = wave_best(rows, nty, lane);
    if (lane != 0) return;
    const float epsilon = st->epsilon;
    st->absmax = best.diff;
    if (best.diff < epsilon || best.idx == INT_MAX) {
        st->converged = best.diff < epsilon;
        st->done = 1;
        st->updated = 0;
        return;
    }
    x[best.idx] = best.xnew;
    st->last = best;
    st->updated = 1;
    st->iters += 1;
    if (st->iters >= st->max_iter) st->done = 1;
}

// All rows and the first selection (after tiles_all_kernel).
__global__ __launch_bounds__(kThreads) void first_select_kernel(const Cand* tiles, Cand* rows, int ntx, int nty,
                                                                State* st, float* x) {
    if (st->done) return;
    rows_and_select(tiles, rows, ntx, nty, 0, nty - 1, st, x);
}

// Window half of an iteration: each block owns one tile of the PSF^2 window
// of the last update, adds psf2 * (old - new) to its gradients
// (FastSerialCD.cs:219-238) and re-evaluates the tile.  The change of the
// last selection is applied even when that selection reached the iteration
// limit (the reference updates before it stops).
__global__ __launch_bounds__(kThreads) void window_kernel(float* __restrict__ x, float* __restrict__ g,
                                                          const float* __restrict__ amap,
                                                          const float* __restrict__ psf2, int p2h, int p2w, int H,
                                                          int W, Cand* __restrict__ tiles, const State* st) {
    __shared__ Cand sh[kThreads / 64];
    if (!st->updated) return;  // uniform
    const float lambda = st->lambda, alpha = st->alpha;
    const int ntx = (W + kTile - 1) / kTile;
    const int y0 = st->last.idx / W, x0 = st->last.idx % W;
    const float d = __fsub_rn(st->last.xold, st->last.xnew);
    const int yh = p2h / 2, xh = p2w / 2;
    const int r0 = max(y0 - yh, 0), r1 = min(y0 - yh + p2h, H);
    const int c0 = max(x0 - xh, 0), c1 = min(x0 - xh + p2w, W);
    const int ty = r0 / kTile + (int)blockIdx.y, tx = c0 / kTile + (int)blockIdx.x;
    if (ty > (r1 - 1) / kTile || tx > (c1 - 1) / kTile) return;
    // this thread's pixels: every load first, then the updates and
    // evaluations in pixel order (x is written only by the select kernel,
    // each gradient by its own thread)
    constexpr int kPer = kTile * kTile / kThreads;
    int idx[kPer];
    bool in[kPer], win[kPer];
    float gv[kPer], pv[kPer], xv[kPer], av[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int p = (int)threadIdx.x + k * kThreads;
        const int i = ty * kTile + p / kTile, j = tx * kTile + p % kTile;
        in[k] = i < H && j < W;
        win[k] = in[k] && i >= r0 && i < r1 && j >= c0 && j < c1;
        idx[k] = in[k] ? i * W + j : 0;
        gv[k] = in[k] ? g[idx[k]] : 0.f;
        xv[k] = in[k] ? x[idx[k]] : 0.f;
        av[k] = in[k] ? amap[idx[k]] : 0.f;
        pv[k] = win[k] ? psf2[(i - y0 + yh) * p2w + (j - x0 + xh)] : 0.f;
    }
    Cand best = none();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        if (!in[k]) continue;
        if (win[k]) {
            gv[k] = __fadd_rn(gv[k], __fmul_rn(pv[k], d));
            g[idx[k]] = gv[k];
        }
        // FastSerialCD.cs:150-162 (as eval)
        const float t = prox(__fadd_rn(__fmul_rn(xv[k], av[k]), gv[k]), av[k], lambda, alpha);
        const Cand c{fabsf(__fsub_rn(xv[k], t)), idx[k], t, xv[k]};
        if (better(c, best)) best = c;
    }
    best = block_best<kThreads>(best, sh);
    if (threadIdx.x == 0) tiles[ty * ntx + tx] = best;
}

// Selection half: refresh the rows the window spans and select the next
// update (one workgroup).
__global__ __launch_bounds__(kThreads) void row_select_kernel(const Cand* tiles, Cand* rows, int p2h, int H, int W,
                                                              State* st, float* x) {
    if (!st->updated) return;  // uniform
    if (st->done) {
        __syncthreads();
        if (threadIdx.x == 0) st->updated = 0;
        return;
    }
    const int y0 = st->last.idx / W;
    const int r0 = max(y0 - p2h / 2, 0), r1 = min(y0 - p2h / 2 + p2h, H);
    __syncthreads();  // every thread has read the state before thread 0 rewrites it
    rows_and_select(tiles, rows, (W + kTile - 1) / kTile, (H + kTile - 1) / kTile, r0 / kTile, (r1 - 1) / kTile, st, x);
}

__global__ void zero_u32_kernel(unsigned* p) { *p = 0u; }

// Residuals.GetMax (Common.cs:244-251): max(0, v) as unsigned bits
__global__ void get_max_kernel(const float* v, size_t n, unsigned* out) {
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, v[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

unsigned blocks(size_t n) {
    const size_t b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b > 4096 ? 4096 : (b == 0 ? 1 : b));
}

}  // namespace

hipError_t launch_tiles_all(const float* x, const float* g, const float* amap, int H, int W, float lambda,
                            float alpha, Cand* tiles, hipStream_t st) {
    const dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);
    hipLaunchKernelGGL(tiles_all_kernel, grid, dim3(kThreads), 0, st, x, g, amap, H, W, lambda, alpha, tiles);
    return hipGetLastError();
}

hipError_t launch_absmax(const Cand* tiles, int ntiles, State* state, hipStream_t st) {
    hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(kSelectThreads), 0, st, tiles, ntiles, state);
    return hipGetLastError();
}

hipError_t launch_place_correlation(const float* psf, int ph, int pw, int ypad, int xpad, int N0, int N1,
                                    double2* buf, hipStream_t st) {
    hipLaunchKernelGGL(place_correlation_kernel, dim3(blocks((size_t)N0 * N1)), dim3(kThreads), 0, st, psf, ph, pw,
                       ypad, xpad, N0, N1, buf);
    return hipGetLastError();
}

hipError_t launch_place_centered(const float* psf, int ph, int pw, int N0, int N1, double2* buf, hipStream_t st) {
    hipLaunchKernelGGL(place_centered_kernel, dim3(blocks((size_t)N0 * N1)), dim3(kThreads), 0, st, psf, ph, pw, N0,
                       N1, buf);
    return hipGetLastError();
}

hipError_t launch_insert_image(const float* img, int H, int W, int yh, int xh, int N0, int N1, double2* buf,
                               hipStream_t st) {
    hipLaunchKernelGGL(insert_image_kernel, dim3(blocks((size_t)N0 * N1)), dim3(kThreads), 0, st, img, H, W, yh, xh,
                       N0, N1, buf);
    return hipGetLastError();
}

hipError_t launch_multiply(double2* buf, const double2* kernel, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(multiply_kernel, dim3(blocks(n)), dim3(kThreads), 0, st, buf, kernel, n);
    return hipGetLastError();
}

hipError_t launch_extract(const double2* buf, int N1, int yh, int xh, double norm, int H, int W, float* out,
                          hipStream_t st) {
    hipLaunchKernelGGL(extract_kernel, dim3(blocks((size_t)H * W)), dim3(kThreads), 0, st, buf, N1, yh, xh, norm, H,
                       W, out);
    return hipGetLastError();
}

hipError_t launch_first_select(const Cand* tiles, Cand* rows, int H, int W, State* state, float* x, hipStream_t st) {
    hipLaunchKernelGGL(first_select_kernel, dim3(1), dim3(kThreads), 0, st, tiles, rows, (W + kTile - 1) / kTile,
                       (H + kTile - 1) / kTile, state, x);
    return hipGetLastError();
}

hipError_t launch_iteration(float* x, float* g, const float* amap, const float* psf2, int p2h, int p2w, int H, int W,
                            Cand* tiles, Cand* rows, State* state, hipStream_t st) {
    const dim3 grid((p2w + kTile - 1) / kTile + 1, (p2h + kTile - 1) / kTile + 1);
    hipLaunchKernelGGL(window_kernel, grid, dim3(kThreads), 0, st, x, g, amap, psf2, p2h, p2w, H, W, tiles, state);
    hipLaunchKernelGGL(row_select_kernel, dim3(1), dim3(kThreads), 0, st, tiles, rows, p2h, H, W, state, x);
    return hipGetLastError();
}

hipError_t launch_amap(const double* scan, int ph, int pw, int H, int W, float* amap, hipStream_t st) {
    hipLaunchKernelGGL(amap_kernel, dim3(blocks((size_t)H * W)), dim3(kThreads), 0, st, scan, ph, pw, H, W, amap);
    return hipGetLastError();
}

hipError_t launch_get_max(const float* v, size_t n, float* out, hipStream_t st) {
    unsigned* o = reinterpret_cast<unsigned*>(out);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, o);
    hipLaunchKernelGGL(get_max_kernel, dim3(blocks(n)), dim3(kThreads), 0, st, v, n, o);
    return hipGetLastError();
}

}  // namespace cd
}  // namespace p9idg
