// Device side of the serial coordinate-descent deconvolver (FastSerialCD.cs)
// and of the PSF correlation (PaddedConvolver.cs, Common.cs PSF helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace p9idg {
namespace cd {

constexpr int kTile = 32;  // edge of the argmax tiles (pixels)

// A candidate update: |x - prox|, pixel index y * W + x, new and old value.
struct Cand {
    float diff;
    int32_t idx;
    float xnew;
    float xold;
};

// Deconvolution state on the device, read by every kernel of an iteration.
// The call's parameters live here too, so a recorded graph of iterations
// serves any lambda, alpha, epsilon and iteration limit.
struct State {
    int32_t done;       // no further updates (converged or iteration limit)
    int32_t converged;  // the last selected |diff| was below epsilon
    int32_t iters;      // updates made
    int32_t updated;    // the last selection changed a pixel
    Cand last;          // that change
    float absmax;       // the last selected |diff|
    float lambda, alpha, epsilon;
    int32_t max_iter;
    int32_t pad[3];
};

// All tiles of the table for lambda / alpha.
hipError_t launch_tiles_all(const float* x, const float* g, const float* amap, int H, int W, float lambda,
                            float alpha, Cand* tiles, hipStream_t st);
// GetAbsMaxDiff: the best |diff| of the tile table into state->absmax.
hipError_t launch_absmax(const Cand* tiles, int ntiles, State* state, hipStream_t st);

// The row table (best tile per tile row) and the first selection after
// launch_tiles_all; then launch_iteration repeatedly: the window tiles take
// the gradient change and are re-evaluated, and one workgroup refreshes the
// rows the window spans and selects the next update.
// lambda, alpha, epsilon and the iteration limit come from the state.
hipError_t launch_first_select(const Cand* tiles, Cand* rows, int H, int W, State* state, float* x, hipStream_t st);
hipError_t launch_iteration(float* x, float* g, const float* amap, const float* psf2, int p2h, int p2w, int H, int W,
                            Cand* tiles, Cand* rows, State* state, hipStream_t st);

// PaddedConvolver / CalcPaddedFourierCorrelation pieces (complex double
// buffers of N0 x N1).
hipError_t launch_place_correlation(const float* psf, int ph, int pw, int ypad, int xpad, int N0, int N1,
                                    double2* buf, hipStream_t st);
hipError_t launch_place_centered(const float* psf, int ph, int pw, int N0, int N1, double2* buf,
                                 hipStream_t st);
hipError_t launch_insert_image(const float* img, int H, int W, int yh, int xh, int N0, int N1, double2* buf,
                               hipStream_t st);
hipError_t launch_multiply(double2* buf, const double2* kernel, size_t n, hipStream_t st);
// out[y][x] = (float)(Re buf[y + yh][x + xh] / norm)
hipError_t launch_extract(const double2* buf, int N1, int yh, int xh, double norm, int H, int W, float* out,
                          hipStream_t st);
// Common.cs:48-126 per pixel from the prefix sums of psf^2
hipError_t launch_amap(const double* scan, int ph, int pw, int H, int W, float* amap, hipStream_t st);
// max(0, max_i v[i]) into *out (Residuals.GetMax)
hipError_t launch_get_max(const float* v, size_t n, float* out, hipStream_t st);

}  // namespace cd
}  // namespace p9idg
