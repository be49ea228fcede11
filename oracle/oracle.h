/*
 * FP64 CPU restatement of the reference image-domain gridder (IDG) with
 * w-stacking.  TEST INFRASTRUCTURE ONLY: tests, the smoke check and the
 * CPU-baseline leg of the benchmark link this library as the checker.  The
 * product library never includes, links or calls it.
 *
 * Every function follows a file of the reference C# code
 * (Distributed-Reconstruction/Core/...) line by line; the citation sits on
 * each definition in oracle.cpp.  Arrays are row-major and contiguous, the
 * same layout as the C# multi-dimensional arrays they replace:
 *   uvw        double[B][T][3]   (metres)
 *   freq       double[C]         (Hz)
 *   vis        double[B][T][C][2] (System.Numerics.Complex = {re, im})
 *   subgrids   double[N][S][S][2]
 *   grid       double[G][G][2],  layers double[L][G][G][2]
 */
#ifndef P9_ORACLE_H
#define P9_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GriddingConstants.cs:11-53 (WStepLambda is never assigned there: keep 0). */
typedef struct orc_constants {
    int64_t visibilities_count;
    int32_t grid_size;
    int32_t subgrid_size;
    int32_t kernel_size;
    int32_t max_timesteps_per_subgrid;
    float scale_arcsec; /* radians per pixel, despite the name */
    int32_t wlayer_count;
    double wstep;
    float wstep_lambda;
} orc_constants;

/* Subgrid.cs:15-24 */
typedef struct orc_subgrid {
    int32_t baseline_idx;
    int32_t time_sample_start;
    int32_t time_sample_count;
    int32_t u_pixel;
    int32_t v_pixel;
    int32_t w_lambda;
} orc_subgrid;

enum {
    ORC_OK = 0,
    ORC_ERR_PLACE = 1,        /* Partitioner.cs:86,91 */
    ORC_ERR_SUBGRID_RANGE = 2, /* Partitioner.cs:118 */
    ORC_ERR_INVALID_GRID = 3, /* Adder.cs:49 */
    ORC_ERR_CAPACITY = 4
};

void orc_set_threads(int n);
/* number of x in a pseudo-random sample where sincos(x) differs from (sin x, cos x) */
int64_t orc_sincos_mismatches(int64_t n, uint64_t seed);
int orc_get_threads(void);

/* Partitioner.cs:20-124.  Subgrids in baseline order, then time order.
 * out may be NULL (count only).  bl_offsets (B+1 entries) may be NULL. */
int orc_partition(const orc_constants* c, const double* uvw, int32_t B, int32_t T,
                  const double* freq, int32_t C, orc_subgrid* out, int64_t capacity,
                  int64_t* n_out, int64_t* bl_offsets);

/* Gridder.cs:15-87 */
void orc_gridder_forward(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                         const double* uvw, int32_t B, int32_t T, const double* vis,
                         const double* freq, int32_t C, double* subgrids_out);
/* Gridder.cs:91-160 */
void orc_gridder_backward(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                          const double* subgrids, const double* uvw, int32_t B, int32_t T,
                          const double* freq, int32_t C, double* vis_out);
/* SubgridFFT.cs:11-52 (in place) */
void orc_subgrid_fft_forward(const orc_constants* c, int64_t n, double* subgrids);
/* SubgridFFT.cs:54-90 (in place) */
void orc_subgrid_fft_backward(const orc_constants* c, int64_t n, double* subgrids);
/* Adder.cs:13-55 */
int orc_adder_add(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                  const double* subgrids, double* grid_out);
/* Adder.cs:57-95 */
void orc_adder_split(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                     const double* grid, double* subgrids_out);
/* AdderWStack.cs:11-78 */
void orc_adderw_add(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                    const double* subgrids, double* layers_out);
/* AdderWStack.cs:81-139 */
void orc_adderw_split(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                      const double* grid, double* subgrids_out);

/* FFT.cs:248-264, :221-241 (quadrant swap, in place) */
void orc_shift_c128(int32_t g, double* grid);
void orc_shift_f64(int32_t g, double* img);
void orc_shift_f32(int32_t g, float* img);
/* FFT.cs:188-215 */
void orc_wstack_ifft_float(int32_t g, int32_t L, const double* layers, int64_t vis_count,
                           float* out);
/* FFT.cs:159-185 */
void orc_fft_backward(int32_t g, const double* grid, int64_t vis_count, double* out);
/* FFT.cs:139-157 */
void orc_fft_backward_float(int32_t g, const double* grid, double norm, float* out);
/* FFT.cs:99-117 */
void orc_fft_forward_float(int32_t g, const float* img, double norm, double* out);
/* unnormalised 2-D DFT, sign -1 (FFTW forward) or +1 (FFTW backward), in place */
void orc_fft2(int32_t n0, int32_t n1, double* data, int sign);
/* Common.cs:305-318 */
void orc_substract(int64_t n, const double* vis0, const double* vis1, const uint8_t* flags,
                   double* out);

/* IDG.cs:12-20 */
int orc_grid(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
             const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
             double* grid_out);
/* IDG.cs:22-30 (grid is shifted in place, as in the reference) */
void orc_degrid(const orc_constants* c, const orc_subgrid* meta, int64_t n, double* grid,
                const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                double* vis_out);
/* IDG.cs:32-40 */
void orc_gridw(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
               const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
               double* layers_out);
/* IDG.cs:42-50 (grid is shifted in place, as in the reference) */
void orc_degridw(const orc_constants* c, const orc_subgrid* meta, int64_t n, double* grid,
                 const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                 double* vis_out);
/* IDG.cs:68-83 */
int orc_to_image(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
                 const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                 double* img_out);
/* MajorCycle.cs:88-90: GridW -> WStackIFFTFloat -> Shift */
void orc_gridw_to_image(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                        const double* vis, const double* uvw, int32_t B, int32_t T,
                        const double* freq, int32_t C, float* img_out);
/* MajorCycle.cs:125-128: Shift -> FFT.Forward(float) -> Shift, then DeGridW */
void orc_image_to_vis(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                      const float* img, const double* uvw, int32_t B, int32_t T,
                      const double* freq, int32_t C, double* vis_out);

/* ---- serial major cycle (oracle_cycle.cpp) ---------------------------- */
/* per-cycle figures of MajorCycle.cs:92-117 */
typedef struct orc_cycle_stats {
    float max_dirty;     /* Residuals.GetMax(dirtyImage) */
    float max_gradient;  /* Residuals.GetMax(gradients) */
    float lambda;        /* currentLambda */
    float abs_max_diff;  /* deconvolver.GetAbsMaxDiff(xImage, gradients, lambdaTrue, alpha) */
    double objective;    /* Residuals.CalcPenalty + ElasticNet.CalcPenalty */
    int32_t iterations;  /* deconvolution iterations (0 when skipped) */
    int32_t converged;
} orc_cycle_stats;

/* Common.cs:227-237, 206-218, 128-135, 48-126 */
void orc_psf_cut(const float* psf, int32_t n, int32_t factor, float* out);
float orc_psf_max_sidelobe(const float* psf, int32_t n, int32_t factor);
double orc_psf_max_lipschitz(const float* psf, int32_t h, int32_t w);
void orc_psf_amap(const float* psf, int32_t ph, int32_t pw, int32_t H, int32_t W, float* amap);
/* Common.cs:146-160: complex (ypad + ph) x (xpad + pw) */
void orc_psf_padded_correlation(const float* psf, int32_t ph, int32_t pw, int32_t ypad, int32_t xpad,
                                double* out);
/* PaddedConvolver.cs:44-66 */
void orc_padded_convolve(const double* kernel, int32_t kh, int32_t kw, const float* image, int32_t H,
                         int32_t W, float* out);
/* Common.cs:182-194: 2ph x 2pw */
void orc_psf_squared(const float* psf, int32_t ph, int32_t pw, float* out);
/* Common.cs:244-260, 22-40 */
float orc_residuals_get_max(const float* img, int64_t n);
double orc_residuals_penalty(const float* img, int64_t n);
double orc_elastic_penalty(const float* img, int64_t n, float lambda, float alpha);
float orc_proximal(float x, float lipschitz, float lambda, float alpha);
/* FastSerialCD.cs:132-135, 77-101 */
float orc_cd_abs_max_diff(const float* amap, const float* x, const float* g, int32_t H, int32_t W,
                          float lambda, float alpha);
int orc_cd_deconvolve(const float* amap, const float* psf2, int32_t p2h, int32_t p2w, float* x, float* g,
                      int32_t H, int32_t W, float lambda, float alpha, int32_t iterations, float epsilon,
                      int32_t* iters_out);
/* MajorCycle.cs:33-134 (FastSerialCD, no FITS output); stats[max_major] */
int orc_reconstruct_serial_cd(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                              const double* vis, const uint8_t* flags, const double* uvw, int32_t B,
                              int32_t T, const double* freq, int32_t C, int32_t psf_cut_factor,
                              int32_t max_major, float lambda, float alpha, int32_t deconv_iterations,
                              float deconv_epsilon, float* x_image, orc_cycle_stats* stats,
                              int32_t* ncycles);

#ifdef __cplusplus
}
#endif
#endif
