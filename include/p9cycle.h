/*
 * p9cycle.h - the serial major cycle around the image-domain gridder on the
 * GPU: PSF helpers, the padded PSF correlation, the serial coordinate-descent
 * deconvolver and MajorCycle.ReconstructSerialCD.
 *
 * Reference interfaces replaced (paths relative to Distributed-Reconstruction/):
 *   Core/Common.cs:48-235        PSF.CalcAMap / CalcMaxLipschitz / CalcPaddedFourierCorrelation /
 *                                CalcPSFSquared / CalcMaxSidelobe / Cut
 *   Core/Common.cs:244-251       Residuals.GetMax
 *   Core/PaddedConvolver.cs:19-66 PaddedConvolver(Complex[,] kernel, Rectangle) + Convolve
 *   Core/Deconvolution/FastSerialCD.cs:31-135  FastSerialCD(totalSize, psf), MaxLipschitz,
 *                                Deconvolve, GetAbsMaxDiff
 *   SingleReconstruction/MajorCycle.cs:33-134  ReconstructSerialCD
 *
 * Status codes and p9idg_last_error() are those of p9idg.h.  Host arrays are
 * row-major float[rows][cols] as the C# float[,] they replace.  Functions
 * ending in _device take device pointers of the same layout and run on the
 * object's own stream; all calls return after the work is done.
 */
#ifndef P9CYCLE_H
#define P9CYCLE_H

#include <stdint.h>

#include "p9idg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PSF helpers (Common.cs) ---------------------------------------------- */
/* PSF.Cut (Common.cs:227-237): the centred (n/factor)^2 block of an n x n PSF */
int p9psf_cut(const float* psf, int32_t n, int32_t factor, float* out);
/* PSF.CalcMaxSidelobe (Common.cs:206-218) */
int p9psf_max_sidelobe(const float* psf, int32_t n, int32_t factor, float* out);
/* PSF.CalcMaxLipschitz (Common.cs:128-135): sum of squares (float products, double sum) */
int p9psf_max_lipschitz(const float* psf, int32_t rows, int32_t cols, double* out);
/* PSF.CalcPSFSquared (Common.cs:182-194): out is float[2 rows][2 cols]; FP64 FFTs on `device` */
int p9psf_squared(int32_t device, const float* psf, int32_t rows, int32_t cols, float* out);
/* PSF.CalcAMap(psf, (0, 0, height, width)) (Common.cs:48-126): out is float[height][width] */
int p9psf_amap(int32_t device, const float* psf, int32_t rows, int32_t cols, int32_t height, int32_t width,
               float* out);

/* ---- PaddedConvolver over PSF.CalcPaddedFourierCorrelation ----------------- */
typedef struct p9conv p9conv;
/* new PaddedConvolver(PSF.CalcPaddedFourierCorrelation(psf, (0, 0, height, width)),
 *                     (0, 0, rows, cols))  (MajorCycle.cs:71); FP64 throughout */
int p9conv_create_correlation(int32_t device, const float* psf, int32_t rows, int32_t cols, int32_t height,
                              int32_t width, p9conv** out);
int p9conv_destroy(p9conv* conv);
/* PaddedConvolver.Convolve (PaddedConvolver.cs:44-66): image, out float[height][width] */
int p9conv_convolve(p9conv* conv, const float* image, float* out);
int p9conv_convolve_device(p9conv* conv, const float* d_image, float* d_out);

/* ---- FastSerialCD ----------------------------------------------------------- */
typedef struct p9cd p9cd;
/* new FastSerialCD((0, 0, height, width), psf) (FastSerialCD.cs:31-47): psf2 and the
 * Lipschitz map on the GPU.  The iterations are the reference's, with the
 * argmax kept as per-tile bests so an iteration re-reads only the PSF^2
 * window.  Ties are broken as the reference's serial scan does (first row,
 * then first column), so results are deterministic. */
int p9cd_create(int32_t device, int32_t height, int32_t width, const float* psf, int32_t rows, int32_t cols,
                p9cd** out);
int p9cd_destroy(p9cd* cd);
/* MaxLipschitz = Residuals.GetMax(psf2) */
int p9cd_max_lipschitz(const p9cd* cd, float* out);
/* GetAbsMaxDiff (FastSerialCD.cs:132-135) */
int p9cd_get_abs_max_diff(p9cd* cd, const float* x_image, const float* gradients, float lambda, float alpha,
                          float* out);
int p9cd_get_abs_max_diff_device(p9cd* cd, const float* d_x_image, const float* d_gradients, float lambda,
                                 float alpha, float* out);
/* Deconvolve (FastSerialCD.cs:77-101): x_image and gradients are updated in
 * place, as the reference updates its arrays */
int p9cd_deconvolve(p9cd* cd, float* x_image, float* gradients, float lambda, float alpha, int32_t max_iterations,
                    float epsilon, int32_t* converged, int32_t* iterations);
int p9cd_deconvolve_device(p9cd* cd, float* d_x_image, float* d_gradients, float lambda, float alpha,
                           int32_t max_iterations, float epsilon, int32_t* converged, int32_t* iterations);

/* ---- MajorCycle.ReconstructSerialCD ---------------------------------------- */
/* per-cycle figures (MajorCycle.cs:92-117) */
typedef struct p9cycle_stats {
    float max_dirty;     /* Residuals.GetMax(dirtyImage) */
    float max_gradient;  /* Residuals.GetMax(gradients) */
    float lambda;        /* currentLambda */
    float abs_max_diff;  /* GetAbsMaxDiff(xImage, gradients, lambdaTrue, alpha) */
    double objective;    /* not evaluated here (the reference only computes it): 0 */
    int32_t iterations;  /* deconvolution iterations (0 when skipped) */
    int32_t converged;
    float gradient_ms;   /* host wall time: maxima + PSF correlation of the dirty image */
    float deconv_ms;     /* GetAbsMaxDiff + Deconvolve */
    float imaging_ms;    /* ToUVW + DeGridW + residual + GridW + ToImage for the next cycle */
    int32_t pad;
} p9cycle_stats;

/* MajorCycle.ReconstructSerialCD (MajorCycle.cs:33-134) on a plan with its
 * observation set: PSF from the flags, then up to max_major_cycles cycles of
 * GridW -> ToImage -> gradients -> deconvolution -> ToUVW -> DeGridW ->
 * residual, all in HBM.  The device step ToUVW -> DeGridW -> residual ->
 * GridW -> ToImage is recorded once as the plan's cycle graph
 * (p9idg_capture_cycle, which also runs it) and replayed in every later
 * cycle (p9idg_launch_cycle).  vis (type P9IDG_C64 / P9IDG_C128, [B][T][C]) is
 * uploaded after the PSF is gridded; flags (uint8 [B][T][C], NULL = none).
 * x_image receives the model image float[G][G]; stats[max_major_cycles] (may
 * be NULL) and *ncycles the cycles run.  The deconvolver is FastSerialCD's
 * (the reference picks GPUSerialCD when a CUDA device exists: the same
 * iterations in batches of 1000 with racy ties).  With obs_name set, every
 * cycle writes <obs_name>_dirty_serial_majorCycle<k>.fits and
 * <obs_name>_model_serial_majorCycle<k>.fits as the reference does
 * (MajorCycle.cs:91,117; p9fits_write_image, float); NULL writes no files. */
int p9idg_reconstruct_serial_cd(p9idg_plan* plan, const char* obs_name, const void* vis, int32_t vis_type,
                                const uint8_t* flags, int32_t psf_cut_factor, int32_t max_major_cycles, float lambda,
                                float alpha, int32_t deconv_iterations, float deconv_epsilon, float* x_image,
                                p9cycle_stats* stats, int32_t* ncycles);

#ifdef __cplusplus
}
#endif
#endif
