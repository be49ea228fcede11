// Internal declarations shared by the host sources of libp9idg.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdint>
#include <string>
#include <vector>

#include "p9idg.h"

namespace p9idg {

// Thread-local error text returned by p9idg_last_error().
void set_error(const std::string& msg);
const char* get_error();

// Messages of the reference's exceptions (Partitioner.cs:86,91,118; Adder.cs:49).
extern const char* const kMsgPlace;
extern const char* const kMsgSubgridRange;
extern const char* const kMsgInvalidGrid;

int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int fft_fail(hipfftResult r, const char* what);

#define P9_HIP(call)                                             \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::p9idg::hip_fail(e_, #call); \
    } while (0)
#define P9_FFT(call)                                             \
    do {                                                         \
        hipfftResult r_ = (call);                                \
        if (r_ != HIPFFT_SUCCESS) return ::p9idg::fft_fail(r_, #call); \
    } while (0)

// Host partitioner (partition.cpp).
int partition_host(const p9idg_constants& c, const double* uvw, int B, int T, const double* freq,
                   int C, std::vector<p9idg_subgrid>& out, std::vector<int64_t>& bl_offsets);

bool subgrid_size_supported(int s);

// One unit of gridder / degridder work: visibilities [k0, k1) of subgrid
// `sidx` (k counts the subgrid's visibilities in time-major, channel-minor
// order).  Partial subgrids of one subgrid add linearly through the subgrid
// FFT and the adder, so large subgrids can be split across items.
// Visibilities [k0, k1) of subgrid sidx; inst: the observation in a batched
// queue; slot: the item's place in (subgrid, k0) order, its scratch subgrid
// for the deterministic adder.
struct WorkItem {
    int32_t sidx, k0, k1, inst, slot;
};

// How a gridder / degridder launch is shaped.
struct LaunchCfg {
    bool queue = true;      // persistent workgroups pulling WorkItems from a counter
    int32_t workgroups = 0; // persistent workgroups (0: device CUs x occupancy)
};

// Rectangle [y0, y1) x [x0, x1) of grid layer `layer` in the grid's own
// (unshifted) coordinates; zero_boxes applies the quadrant swap.
struct Box {
    int32_t layer, y0, y1, x0, x1;
};

// Deterministic adder: a 32 x 32 tile of one layer (unshifted grid
// coordinates) whose covering scratch slots, in ascending order, are cut
// into chunks[chunk0, chunk0 + nchunk) of at most kGatherChunk slots each
// (slots[begin, end)); each slot's S x S box starts at its corner.
constexpr int kGatherTile = 32;
constexpr int kGatherChunk = 16;
struct GatherBin {
    int32_t layer, ty, tx, chunk0, nchunk;
};
struct GatherChunk {
    int32_t bin, begin, end;
};
struct SlotCorner {
    int32_t y0, x0;
};

// Kernel launch parameters shared by host and device code.
struct KernelArgs {
    const p9idg_subgrid* meta;
    const int32_t* order;  // subgrid schedule (longest first)
    int32_t nr_subgrids;
    const double* uvw;     // [B][T][3]
    const double* wavenumbers;  // [C]
    const float2* vis;     // [B][T][C]
    float2* vis_out;       // [B][T][C] (degrid)
    const float2* twiddle; // [S] exp(+2 pi i m / S)
    const float2* adder_phasor;  // [S][S] exp(i pi (x + y - S) / S)
    const float* rtab;     // [S][S] r(x, y) = n(l, m) - n(l, 0) - n(0, m) (MFMA kernels)
    float2* grid;          // grid or layer stack (gridder output)
    // non-null: the adder writes each work item's normalised subgrid to
    // scratch[slot][S][S] (subgrid order) instead of adding it onto grid
    float2* scratch;
    const float2* model_grid;  // [G][G] (degrid input, unshifted as the caller passes it)
    int32_t G, T, C, L;
    float image_size;
    double u_scale;        // 2 pi / image_size  (Gridder.cs:40)
    double w_offset_scale; // 2 pi * WStepLambda (0 in the reference)
    float degrid_scale;    // 1.0f / (S * S)  (Gridder.cs:151)
    int32_t mode;
    // nonzero: channels are uniformly spaced in wavenumber and C % 8 == 0, so
    // each aligned run of 8 visibilities is one timestep's 8 consecutive
    // channels and the MFMA gridder steps its phasors channel to channel
    int32_t chan_rec;
    // nonzero: channels are uniformly spaced in wavenumber (any C): the
    // degridder, one timestep per lane, steps its phasors channel to channel
    int32_t chan_step;
    // -1.0 as BF16 in the low half, as BF16 in the high half, and as FP16 in
    // the low half (0x0000BF80, 0xBF800000, 0x0000BC00): the MFMA kernels'
    // hi/lo split takes its residual a - hi in one dot2 (BF16) or mixed FMA
    // (FP16) against these; as compile-time constants they would be folded
    // back into a conversion and a subtraction
    uint32_t split_k[3];
    const WorkItem* items; // work queue (longest first)
    int32_t nitems;
    int32_t* counter;      // [2]: next item, finished workgroups; zero between launches
};

// Several observations in one persistent launch: item.inst indexes args.
struct BatchArgs {
    const KernelArgs* args;
    const WorkItem* items;
    int32_t nitems;
    int32_t* counter;  // [2], as in KernelArgs
};

// Launchers (kernels_*.hip); S is the subgrid size.
// sep: the separable FP32 form (P9IDG_KERNEL_FP32_SEP) where S allows it
hipError_t launch_grid_fp32(const KernelArgs& a, int S, bool sep, const LaunchCfg& cfg, hipStream_t st);
hipError_t launch_degrid_fp32(const KernelArgs& a, int S, bool sep, const LaunchCfg& cfg, hipStream_t st);
hipError_t launch_grid_mfma(const KernelArgs& a, int S, int precision, const LaunchCfg& cfg,
                            hipStream_t st);
hipError_t launch_degrid_mfma(const KernelArgs& a, int S, int precision, const LaunchCfg& cfg,
                              hipStream_t st);
// both queue counters to zero (once per allocation: the queue kernels leave
// them at zero)
hipError_t launch_reset_counter(int32_t* counter, hipStream_t st);
hipError_t launch_rtab(float* rtab, int S, float image_size, hipStream_t st);
hipError_t launch_fp32_batch(bool grid, bool sep, const BatchArgs& b, int S, int workgroups, hipStream_t st);
// kernel: p9idg_kernel; workgroups 0 = CUs x occupancy
hipError_t launch_grid_batch(const BatchArgs& b, int S, int kernel, int workgroups, hipStream_t st);
hipError_t launch_degrid_batch(const BatchArgs& b, int S, int kernel, int workgroups, hipStream_t st);
// Persistent workgroups for a kernel: CUs x resident workgroups per CU.
int persistent_workgroups(const void* kernel, int threads);
hipError_t launch_fill_zero(float2* p, size_t n, hipStream_t st);
hipError_t launch_c128_to_c64(const double2* in, float2* out, size_t n, hipStream_t st);
hipError_t launch_psf_vis(const uint8_t* flags, float2* out, size_t n, hipStream_t st);
hipError_t launch_layers_to_image(const float2* layers, int L, int G, double inv_n, int shift,
                                  float* img, hipStream_t st);
hipError_t launch_grid_to_image_f64(const float2* grid, int G, double inv_n, int shift,
                                    double* img, hipStream_t st);
hipError_t launch_image_to_grid(const float* img, int G, int shift, float2* grid,
                                hipStream_t st);
hipError_t launch_residual(const float2* vis, const float2* model, const uint8_t* flags,
                           float2* out, size_t n, hipStream_t st);
hipError_t launch_zero_boxes(float2* base, const Box* boxes, int nboxes, int G, hipStream_t st);
// Writes every pixel of each bin's tile as the sum of its chunks' partial
// sums in chunk order, each partial the sum of its slots' scratch values in
// slot order (the gridder target is zero elsewhere).  partial: one tile per
// chunk.
hipError_t launch_gather_slots(const GatherBin* bins, int nbins, const GatherChunk* chunks, int nchunks,
                               const int32_t* slots, const SlotCorner* corners, const float2* scratch, int S, int G,
                               float2* partial, float2* base, hipStream_t st);
hipError_t launch_sum_layers(const float2* layers, int L, size_t GG, float2* out, hipStream_t st);
hipError_t launch_sum_layer_boxes(const float2* layers, const Box* boxes, int nboxes, int G, Box u,
                                  float2* out, hipStream_t st);
hipError_t launch_scale_c64(float2* p, size_t n, double inv_norm, hipStream_t st);

}  // namespace p9idg
