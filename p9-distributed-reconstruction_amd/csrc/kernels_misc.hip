// Element-wise HBM-bound kernels around the gridder: buffer fills, visibility
// conversion, the PSF visibilities, and the image <-> grid steps around the
// G x G inverse / forward FFTs (FFT.cs:99-117, :159-215, :248-264).
#include <map>
#include <mutex>

#include "device_common.h"

namespace p9idg {
namespace dev {

__global__ void fill_zero_kernel(float4* p, size_t n4) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void c128_to_c64_kernel(const double2* in, float2* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const double2 v = in[i];
        out[i] = make_float2((float)v.x, (float)v.y);
    }
}

// MajorCycle.cs:36-43 / IDG.cs:54-59: 1 where unflagged, 0 where flagged.
__global__ void psf_vis_kernel(const uint8_t* flags, float2* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = make_float2((flags && flags[i]) ? 0.f : 1.f, 0.f);
}

// FFT.WStackIFFTFloat (FFT.cs:194-211) + FFT.Shift (FFT.cs:248-264): the
// layers already hold their inverse FFTs; sum (float)(Re / N) over layers in
// layer order and write the quadrant-swapped image.
__global__ void layers_to_image_kernel(const float2* layers, int L, int G, double inv_n,
                                       int shift, float* img) {
    const size_t GG = (size_t)G * G;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < GG; i += stride) {
        const int y = (int)(i / G), x = (int)(i - (size_t)y * G);
        const size_t src = shift ? shift_index(y, x, G) : i;
        float acc = 0.f;
        for (int k = 0; k < L; k++) acc += (float)((double)layers[k * GG + src].x * inv_n);
        img[i] = acc;
    }
}

// Elementwise sum of the L w-layers (any common layout).
__global__ void sum_layers_kernel(const float2* layers, int L, size_t GG, float2* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < GG; i += stride) {
        float2 acc = layers[i];
        for (int k = 1; k < L; k++) {
            const float2 v = layers[k * GG + i];
            acc.x += v.x;
            acc.y += v.y;
        }
        out[i] = acc;
    }
}

// Sum of the w-layers over the union box u of their footprints, where layer
// b.layer is nonzero only inside box b (FFT.Shift-ed storage); `out` is zero
// outside u.
__global__ void sum_layer_boxes_kernel(const float2* layers, const Box* boxes, int nboxes, int G,
                                       Box u, float2* out) {
    const int w = u.x1 - u.x0;
    const size_t n = (size_t)(u.y1 - u.y0) * (size_t)w;
    const size_t GG = (size_t)G * G;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const int y = u.y0 + (int)(i / w), x = u.x0 + (int)(i % w);
        const size_t idx = shift_index(y, x, G);
        float2 acc = make_float2(0.f, 0.f);
        for (int b = 0; b < nboxes; b++) {
            const Box bx = boxes[b];
            if (y >= bx.y0 && y < bx.y1 && x >= bx.x0 && x < bx.x1) {
                const float2 v = layers[bx.layer * GG + idx];
                acc.x += v.x;
                acc.y += v.y;
            }
        }
        out[idx] = acc;
    }
}

// FFT.Backward(Complex[,], long) (FFT.cs:159-185) + Shift, double output.
__global__ void grid_to_image_f64_kernel(const float2* grid, int G, double inv_n, int shift,
                                         double* img) {
    const size_t GG = (size_t)G * G;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < GG; i += stride) {
        const int y = (int)(i / G), x = (int)(i - (size_t)y * G);
        const size_t src = shift ? shift_index(y, x, G) : i;
        img[i] = (double)grid[src].x * inv_n;
    }
}

// MajorCycle.cs:125-126: FFT.Shift(x) then the real->complex copy of
// FFT.Forward(float[,]) (FFT.cs:105-107); the forward FFT follows.
__global__ void image_to_grid_kernel(const float* img, int G, int shift, float2* grid) {
    const size_t GG = (size_t)G * G;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < GG; i += stride) {
        const int y = (int)(i / G), x = (int)(i - (size_t)y * G);
        const size_t src = shift ? shift_index(y, x, G) : i;
        grid[i] = make_float2(img[src], 0.f);
    }
}

__global__ void scale_c64_kernel(float2* p, size_t n, double inv) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        float2 v = p[i];
        p[i] = make_float2((float)((double)v.x * inv), (float)((double)v.y * inv));
    }
}

// Visibilities.Substract (Common.cs:305-318): vis - model, 0 where flagged.
__global__ void residual_kernel(const float2* vis, const float2* model, const uint8_t* flags,
                                float2* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const float2 v = vis[i], m = model[i];
        out[i] = (flags && flags[i]) ? make_float2(0.f, 0.f) : make_float2(v.x - m.x, v.y - m.y);
    }
}

// Zeroes the (quadrant-swapped) image of each box: blockIdx.y is the box.
__global__ void zero_boxes_kernel(float2* base, const Box* boxes, int G) {
    const Box b = boxes[blockIdx.y];
    const int w = b.x1 - b.x0;
    const size_t n = (size_t)(b.y1 - b.y0) * (size_t)w;
    float2* layer = base + (size_t)b.layer * G * G;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = b.y0 + (int)(i / w), x = b.x0 + (int)(i % w);
        layer[shift_index(y, x, G)] = make_float2(0.f, 0.f);
    }
}

// Deterministic subgrid sums, pass 1: one workgroup per chunk; each of the
// tile's pixels sums the chunk's slots in slot order.  Every load reads
// inside its slot's scratch (clamped), so the loads need no branch.
constexpr int kGatherPer = kGatherTile * kGatherTile / 256;
__global__ __launch_bounds__(256) void gather_chunks_kernel(const GatherBin* bins, const GatherChunk* chunks,
                                                            const int32_t* slots, const SlotCorner* corners,
                                                            const float2* scratch, int S, float2* partial) {
    __shared__ int32_t s_slot[kGatherChunk];
    __shared__ SlotCorner s_corner[kGatherChunk];
    const GatherChunk ch = chunks[blockIdx.x];
    const GatherBin b = bins[ch.bin];
    const int n = ch.end - ch.begin;  // <= kGatherChunk
    if ((int)threadIdx.x < n) {
        const int s = slots[ch.begin + threadIdx.x];
        s_slot[threadIdx.x] = s;
        s_corner[threadIdx.x] = corners[s];
    }
    __syncthreads();
    int gy[kGatherPer], gx[kGatherPer];
    float2 acc[kGatherPer];
#pragma unroll
    for (int k = 0; k < kGatherPer; k++) {
        const int p = threadIdx.x + 256 * k;
        gy[k] = b.ty * kGatherTile + p / kGatherTile;
        gx[k] = b.tx * kGatherTile + p % kGatherTile;
        acc[k] = make_float2(0.f, 0.f);
    }
#pragma unroll 8
    for (int i = 0; i < n; i++) {
        const int s = s_slot[i];
        const SlotCorner c = s_corner[i];
        const float2* sub = scratch + (size_t)s * S * S;
#pragma unroll
        for (int k = 0; k < kGatherPer; k++) {
            const unsigned y = (unsigned)(gy[k] - c.y0), x = (unsigned)(gx[k] - c.x0);
            const bool in = y < (unsigned)S && x < (unsigned)S;
            const float2 v = sub[in ? y * S + x : 0];
            if (in) {
                acc[k].x += v.x;
                acc[k].y += v.y;
            }
        }
    }
    float2* out = partial + (size_t)blockIdx.x * kGatherTile * kGatherTile;
#pragma unroll
    for (int k = 0; k < kGatherPer; k++) out[threadIdx.x + 256 * k] = acc[k];
}

// Pass 2: one workgroup per bin; each pixel sums its tile's chunk partials in
// chunk order and is written to the (quadrant-swapped) target.
__global__ __launch_bounds__(256) void gather_bins_kernel(const GatherBin* bins, const float2* partial, int G,
                                                          float2* base) {
    const GatherBin b = bins[blockIdx.x];
    float2* layer = base + (size_t)b.layer * G * G;
    for (int p = threadIdx.x; p < kGatherTile * kGatherTile; p += blockDim.x) {
        const int gy = b.ty * kGatherTile + p / kGatherTile, gx = b.tx * kGatherTile + p % kGatherTile;
        if (gy >= G || gx >= G) continue;
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll 8
        for (int c = b.chunk0; c < b.chunk0 + b.nchunk; c++) {
            const float2 v = partial[(size_t)c * kGatherTile * kGatherTile + p];
            acc.x += v.x;
            acc.y += v.y;
        }
        layer[shift_index(gy, gx, G)] = acc;
    }
}

inline unsigned blocks_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

}  // namespace dev

int persistent_workgroups(const void* kernel, int threads) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(kernel, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 1;
    const int n = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    cache[key] = n;
    return n;
}

hipError_t launch_fill_zero(float2* p, size_t n, hipStream_t st) {
    // n complex64 values; buffers are allocated in multiples of 2 complex.
    const size_t n4 = (n + 1) / 2;
    hipLaunchKernelGGL(dev::fill_zero_kernel, dim3(dev::blocks_for(n4)), dim3(256), 0, st,
                       reinterpret_cast<float4*>(p), n4);
    return hipGetLastError();
}
hipError_t launch_c128_to_c64(const double2* in, float2* out, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(dev::c128_to_c64_kernel, dim3(dev::blocks_for(n)), dim3(256), 0, st, in,
                       out, n);
    return hipGetLastError();
}
hipError_t launch_psf_vis(const uint8_t* flags, float2* out, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(dev::psf_vis_kernel, dim3(dev::blocks_for(n)), dim3(256), 0, st, flags,
                       out, n);
    return hipGetLastError();
}
hipError_t launch_layers_to_image(const float2* layers, int L, int G, double inv_n, int shift,
                                  float* img, hipStream_t st) {
    hipLaunchKernelGGL(dev::layers_to_image_kernel, dim3(dev::blocks_for((size_t)G * G)),
                       dim3(256), 0, st, layers, L, G, inv_n, shift, img);
    return hipGetLastError();
}
hipError_t launch_sum_layers(const float2* layers, int L, size_t GG, float2* out, hipStream_t st) {
    hipLaunchKernelGGL(dev::sum_layers_kernel, dim3(dev::blocks_for(GG)), dim3(256), 0, st, layers, L,
                       GG, out);
    return hipGetLastError();
}
hipError_t launch_sum_layer_boxes(const float2* layers, const Box* boxes, int nboxes, int G, Box u,
                                  float2* out, hipStream_t st) {
    const size_t n = (size_t)(u.y1 - u.y0) * (size_t)(u.x1 - u.x0);
    hipLaunchKernelGGL(dev::sum_layer_boxes_kernel, dim3(dev::blocks_for(n)), dim3(256), 0, st,
                       layers, boxes, nboxes, G, u, out);
    return hipGetLastError();
}
hipError_t launch_grid_to_image_f64(const float2* grid, int G, double inv_n, int shift,
                                    double* img, hipStream_t st) {
    hipLaunchKernelGGL(dev::grid_to_image_f64_kernel, dim3(dev::blocks_for((size_t)G * G)),
                       dim3(256), 0, st, grid, G, inv_n, shift, img);
    return hipGetLastError();
}
hipError_t launch_image_to_grid(const float* img, int G, int shift, float2* grid,
                                hipStream_t st) {
    hipLaunchKernelGGL(dev::image_to_grid_kernel, dim3(dev::blocks_for((size_t)G * G)),
                       dim3(256), 0, st, img, G, shift, grid);
    return hipGetLastError();
}
hipError_t launch_residual(const float2* vis, const float2* model, const uint8_t* flags,
                           float2* out, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(dev::residual_kernel, dim3(dev::blocks_for(n)), dim3(256), 0, st, vis, model,
                       flags, out, n);
    return hipGetLastError();
}
hipError_t launch_zero_boxes(float2* base, const Box* boxes, int nboxes, int G, hipStream_t st) {
    if (nboxes <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::zero_boxes_kernel, dim3(64, nboxes), dim3(256), 0, st, base, boxes, G);
    return hipGetLastError();
}
hipError_t launch_gather_slots(const GatherBin* bins, int nbins, const GatherChunk* chunks, int nchunks,
                               const int32_t* slots, const SlotCorner* corners, const float2* scratch, int S, int G,
                               float2* partial, float2* base, hipStream_t st) {
    if (nbins <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::gather_chunks_kernel, dim3(nchunks), dim3(256), 0, st, bins, chunks, slots, corners,
                       scratch, S, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::gather_bins_kernel, dim3(nbins), dim3(256), 0, st, bins, partial, G, base);
    return hipGetLastError();
}
hipError_t launch_scale_c64(float2* p, size_t n, double inv_norm, hipStream_t st) {
    hipLaunchKernelGGL(dev::scale_c64_kernel, dim3(dev::blocks_for(n)), dim3(256), 0, st, p, n,
                       inv_norm);
    return hipGetLastError();
}

}  // namespace p9idg
