/*
 * p9idg -- C ABI of the MI355X (gfx950) image-domain gridder with w-stacking.
 *
 * Drop-in boundary for the gridding API of the reference C# code
 * (paths relative to Distributed-Reconstruction/):
 *
 *   GriddingConstants ctor        Core/ImageDomainGridder/GriddingConstants.cs:33
 *   Subgrid                       Core/ImageDomainGridder/Subgrid.cs:15-24
 *   Partitioner.CreatePartition   Core/ImageDomainGridder/Partitioner.cs:20
 *   IDG.Grid / DeGrid             Core/ImageDomainGridder/IDG.cs:12 / :22
 *   IDG.GridW / DeGridW           Core/ImageDomainGridder/IDG.cs:32 / :42
 *   IDG.GridPSF / ToImage / CalculatePSF   IDG.cs:52 / :68 / :85
 *   FFT.WStackIFFTFloat           Core/FFT.cs:188
 *   FFT.Forward(float[,])         Core/FFT.cs:99
 *   FFT.Backward(Complex[,],long) Core/FFT.cs:159
 *   FFT.Shift                     Core/FFT.cs:221, :248
 *   Visibilities.Substract        Core/Common.cs:305
 *
 * Conventions
 *  - Layouts are those of the C# arrays: row-major, contiguous.
 *      uvw   double[B][T][3] (metres)      freq double[C] (Hz)
 *      vis   complex[B][T][C]              grid complex[G][G]
 *      layers complex[L][G][G]             image float/double[G][G]
 *    "complex128" is {double re, double im} (System.Numerics.Complex);
 *    "complex64" is {float re, float im}.
 *  - Subgrid metadata is the flattening of List<List<Subgrid>>: all subgrids of
 *    baseline 0 in time order, then baseline 1, ...; baseline_offsets[b] is the
 *    index of the first subgrid of baseline b (B+1 entries).
 *  - The caller owns every host array; the plan owns device memory, the HIP
 *    stream and the FFT plans.  The library never frees caller memory.
 *  - Every entry point returns a p9idg_status; on failure p9idg_last_error()
 *    returns the message of the calling thread.  The three reference
 *    exceptions keep their exact texts.
 *  - Host-buffer entry points are synchronous.  Device entry points
 *    (p9idg_exec_*) are asynchronous on the plan's stream; p9idg_sync waits.
 *  - A plan is not re-entrant; separate plans may be used from separate
 *    threads concurrently.
 *  - The degrid entry points read the caller's grid without modifying it.
 *    (The reference's IDG.DeGrid/DeGridW swap the quadrants of the caller's
 *    grid in place, IDG.cs:24,44; no caller in the reference reads the grid
 *    afterwards.)
 */
#ifndef P9IDG_H
#define P9IDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P9IDG_API_VERSION 1

/* GriddingConstants.cs:11-53.  scale_arcsec is radians per pixel (the
 * reference passes cellSize in radians, RunningMethod.cs:22,45).
 * wstep_lambda must be 0: the reference never assigns WStepLambda
 * (GriddingConstants.cs:25), so its w offset is always 0. */
typedef struct p9idg_constants {
    int64_t visibilities_count; /* VisibilitiesCount (normalises the image) */
    int32_t grid_size;          /* GridSize (even) */
    int32_t subgrid_size;       /* SubgridSize: 8, 16, 24, 32, 48 or 64 */
    int32_t kernel_size;        /* KernelSize */
    int32_t max_timesteps_per_subgrid;
    float scale_arcsec;
    int32_t wlayer_count;       /* WLayerCount */
    double wstep;               /* WStep (lambda); 0 disables w binning */
    float wstep_lambda;         /* must be 0 */
} p9idg_constants;

/* Subgrid.cs:15-24 */
typedef struct p9idg_subgrid {
    int32_t baseline_idx;
    int32_t time_sample_start;
    int32_t time_sample_count;
    int32_t u_pixel;
    int32_t v_pixel;
    int32_t w_lambda; /* w-layer index, may be negative */
} p9idg_subgrid;

typedef enum p9idg_status {
    P9IDG_OK = 0,
    P9IDG_ERR_PLACE = 1,          /* Partitioner.cs:86,91 */
    P9IDG_ERR_SUBGRID_RANGE = 2,  /* Partitioner.cs:118 */
    P9IDG_ERR_INVALID_GRID = 3,   /* Adder.cs:49 */
    P9IDG_ERR_ARGUMENT = 4,
    P9IDG_ERR_CAPACITY = 5,
    P9IDG_ERR_HIP = 6,
    P9IDG_ERR_ALLOC = 7,
    P9IDG_ERR_STATE = 8,
    P9IDG_ERR_UNSUPPORTED = 9,
    P9IDG_ERR_IO = 10             /* a file that cannot be opened, read or parsed (p9fits.h) */
} p9idg_status;

typedef enum p9idg_vis_type { P9IDG_C128 = 0, P9IDG_C64 = 1 } p9idg_vis_type;

/* Phasor arithmetic of the gridder and degridder. */
typedef enum p9idg_kernel {
    P9IDG_KERNEL_FP32 = 0,      /* per-pixel phasors on the FP32 vector ALUs */
    P9IDG_KERNEL_MFMA_BF16 = 1, /* split hi/lo BF16 MFMA contraction, FP32 accumulate */
    P9IDG_KERNEL_MFMA_FP16 = 2, /* split hi/lo FP16 MFMA contraction, FP32 accumulate */
    P9IDG_KERNEL_FP32_SEP = 3   /* the MFMA kernels' separable phasors (2 S per visibility)
                                   contracted with FP32 FMAs on the vector ALUs; gridder for
                                   S = 16, 32, 48, 64 and degridder for S = 16, 32, the
                                   per-pixel FP32 form for other sizes */
} p9idg_kernel;

/* Where p9idg_exec_grid puts the subgrids. */
typedef enum p9idg_grid_mode {
    P9IDG_GRID_PLAIN = 0,  /* Adder.Add onto one grid (IDG.Grid), Adder.cs:13 */
    P9IDG_GRID_WSTACK = 1, /* AdderWStack.Add onto L layers (IDG.GridW), AdderWStack.cs:11 */
    P9IDG_GRID_WSUM = 2    /* AdderWStack.Add with every layer summed into one grid:
                              same real image after the inverse FFT, 1 FFT instead of L */
} p9idg_grid_mode;

/* Where p9idg_exec_degrid reads the subgrids from. */
typedef enum p9idg_degrid_mode {
    P9IDG_DEGRID_PLAIN = 0, /* Adder.Split (IDG.DeGrid), Adder.cs:57 */
    P9IDG_DEGRID_WSTACK = 1 /* AdderWStack.Split (IDG.DeGridW), AdderWStack.cs:81 */
} p9idg_degrid_mode;

/* Input of p9idg_exec_grid. */
typedef enum p9idg_grid_source {
    P9IDG_SRC_VIS = 0,     /* the uploaded visibilities */
    P9IDG_SRC_RESIDUAL = 1 /* the residuals of p9idg_exec_residual */
} p9idg_grid_source;

typedef struct p9idg_plan p9idg_plan;

int p9idg_api_version(void);
const char* p9idg_last_error(void);

/* ---- Partitioner (host, bit-exact with Partitioner.cs:20-124) ----------- */
/* out may be NULL to query the count.  baseline_offsets may be NULL. */
int p9idg_partition(const p9idg_constants* c, const double* uvw, int32_t nr_baselines,
                    int32_t nr_timesteps, const double* freq, int32_t nr_channels,
                    p9idg_subgrid* out, int64_t capacity, int64_t* nr_subgrids,
                    int64_t* baseline_offsets);

/* ---- Plan (handle) ------------------------------------------------------- */
/* Replaces constructing GriddingConstants; owns device state.  IDisposable
 * pattern of Core/FFT.cs:49-76. */
int p9idg_create(const p9idg_constants* c, int32_t device, p9idg_plan** plan);
/* The plan's constants and device (no counterpart in the reference: the
 * constants object is the caller's own there). */
int p9idg_get_constants(const p9idg_plan* plan, p9idg_constants* c, int32_t* device);
int p9idg_destroy(p9idg_plan* plan);
int p9idg_set_kernel(p9idg_plan* plan, int32_t kernel);
int p9idg_get_kernel(const p9idg_plan* plan, int32_t* kernel);
/* on != 0: the gridder sums subgrids onto the grid in a fixed order (each
 * work item's subgrid goes to a scratch slot, then every 32 x 32 grid tile
 * adds its covering slots in (subgrid, first visibility) order), so the
 * result is the same bit for bit from run to run, as with the serial
 * Adder.cs:13-55 / AdderWStack.cs:11-78; off (default): float atomics, whose
 * order varies at the 1e-7 level.  Costs one S^2 complex64 scratch subgrid
 * per work item and one extra pass.  Not available in batches
 * (p9idg_batch_prepare returns P9IDG_ERR_UNSUPPORTED). */
int p9idg_set_deterministic(p9idg_plan* plan, int32_t on);
/* Launch shape of the gridder / degridder: queue != 0 runs persistent
 * workgroups over a work queue whose items hold at most grid_chunk /
 * degrid_chunk visibilities of one subgrid (0 = whole subgrids), with
 * `workgroups` persistent workgroups in each direction (0 = CUs x
 * occupancy); queue == 0 launches one workgroup per subgrid.  Defaults:
 * 1, 2048, 2048, 0.  get_launch reports the gridder's workgroups (autotune
 * may pick another count for the degridder). */
int p9idg_set_launch(p9idg_plan* plan, int32_t queue, int32_t grid_chunk, int32_t degrid_chunk,
                     int32_t workgroups);
int p9idg_get_launch(const p9idg_plan* plan, int32_t* queue, int32_t* grid_chunk,
                     int32_t* degrid_chunk, int32_t* workgroups);
/* Chooses the launch shape (queue or not, visibilities per work item and
 * persistent workgroups for each direction) by timing candidates on the plan's
 * observation with HIP events (median of reps runs each), in the w-stacked
 * modes when the constants have more than one w-layer and the plain modes
 * otherwise; keeps the fastest and returns its gridder / degridder kernel
 * times.  Overwrites the grid / layer and model visibility buffers. */
int p9idg_autotune(p9idg_plan* plan, int32_t reps, float* grid_ms, float* degrid_ms);
/* Uploads uvw, frequencies and the partition to HBM (once per observation). */
int p9idg_set_observation(p9idg_plan* plan, const double* uvw, int32_t nr_baselines,
                          int32_t nr_timesteps, const double* freq, int32_t nr_channels,
                          const p9idg_subgrid* meta, int64_t nr_subgrids);

/* ---- Host-buffer drop-ins (synchronous) ---------------------------------- */
/* IDG.Grid (IDG.cs:12-20): grid_out complex128[G][G] */
int p9idg_grid(p9idg_plan* plan, const void* vis, int32_t vis_type, double* grid_out);
/* IDG.DeGrid (IDG.cs:22-30): vis_out complex128[B][T][C] */
int p9idg_degrid(p9idg_plan* plan, const double* grid, double* vis_out);
/* IDG.GridW (IDG.cs:32-40): layers_out complex128[L][G][G] */
int p9idg_gridw(p9idg_plan* plan, const void* vis, int32_t vis_type, double* layers_out);
/* IDG.DeGridW (IDG.cs:42-50): grid complex128[G][G] (one grid, as the reference) */
int p9idg_degridw(p9idg_plan* plan, const double* grid, double* vis_out);
/* IDG.GridPSF (IDG.cs:52-66): flags uint8[B][T][C] (nonzero = flagged) or NULL */
int p9idg_grid_psf(p9idg_plan* plan, const uint8_t* flags, double* grid_out);
/* IDG.ToImage (IDG.cs:68-83): img_out double[G][G] */
int p9idg_to_image(p9idg_plan* plan, const void* vis, int32_t vis_type, double* img_out);
/* IDG.CalculatePSF (IDG.cs:85-114): psf_out double[G][G] */
int p9idg_calculate_psf(p9idg_plan* plan, const uint8_t* flags, double* psf_out);
/* GridW -> FFT.WStackIFFTFloat -> FFT.Shift (MajorCycle.cs:46-48, :88-90) */
int p9idg_gridw_to_image(p9idg_plan* plan, const void* vis, int32_t vis_type, float* img_out);
/* FFT.Shift(x); FFT.Forward(x); FFT.Shift(x); IDG.DeGridW (MajorCycle.cs:125-128) */
int p9idg_image_to_vis(p9idg_plan* plan, const float* img, double* vis_out);

/* ---- FFT.cs static steps (host buffers, computed on the GPU) ------------- */
/* FFT.WStackIFFTFloat (FFT.cs:188-215): layers complex128[L][G][G] */
int p9idg_wstack_ifft_float(p9idg_plan* plan, const double* layers, int32_t nr_layers,
                            int64_t vis_count, float* out);
/* FFT.Forward(float[,], norm) (FFT.cs:99-117): grid_out complex128[G][G] */
int p9idg_fft_forward_float(p9idg_plan* plan, const float* img, double norm, double* grid_out);
/* FFT.Backward(Complex[,], long) (FFT.cs:159-185) */
int p9idg_fft_backward(p9idg_plan* plan, const double* grid, int64_t vis_count, double* out);
/* FFT.Shift (FFT.cs:248-264), host, in place; n = rows = cols */
int p9idg_shift_f32(float* a, int32_t n);
int p9idg_shift_f64(double* a, int32_t n);
int p9idg_shift_c128(double* a, int32_t n);
/* Visibilities.Substract (Common.cs:305-318), host */
int p9idg_substract(const double* vis0, const double* vis1, const uint8_t* flags, int64_t n,
                    double* out);

/* ---- Device-resident API (inputs stay in HBM) ---------------------------- */
/* Copy visibilities into the plan's complex64[B][T][C] device buffer. */
int p9idg_upload_vis(p9idg_plan* plan, const void* vis, int32_t vis_type);
/* Fill the plan's visibility buffer with 1 (0 where flagged): PSF input,
 * MajorCycle.cs:36-43. */
int p9idg_set_psf_vis(p9idg_plan* plan, const uint8_t* flags);
/* Grid the plan's visibilities (gridder -> subgrid FFT -> adder, one fused
 * launch) into its grid buffer (complex64, layers for WSTACK). */
int p9idg_exec_grid(p9idg_plan* plan, int32_t mode);
/* Inverse-FFT the grid buffer(s) into the plan's float image
 * (FFT.WStackIFFTFloat + Shift, or FFT.Backward + Shift). */
int p9idg_exec_grid_to_image(p9idg_plan* plan, int32_t mode);
/* ToUVW: image (float[G][G], device) -> model grid buffer (complex64). */
int p9idg_exec_image_to_grid(p9idg_plan* plan);
/* Degrid the model grid buffer (splitter -> subgrid FFT -> degridder, one
 * fused launch) into the plan's complex64 model-visibility buffer. */
int p9idg_exec_degrid(p9idg_plan* plan, int32_t mode);
int p9idg_sync(p9idg_plan* plan);
/* Flags uint8[B][T][C] (nonzero = flagged) used by the residuals; NULL clears. */
int p9idg_set_flags(p9idg_plan* plan, const uint8_t* flags);
/* Visibilities.Substract (Common.cs:305-318) on the device:
 * residual = vis - model_vis, 0 where flagged. */
int p9idg_exec_residual(p9idg_plan* plan);
int p9idg_set_grid_source(p9idg_plan* plan, int32_t src);
/* One device step of the major cycle (MajorCycle.cs:125-129, then :88-90):
 * ToUVW of the plan's image -> DeGridW -> residuals -> grid the residuals
 * (grid_mode) -> inverse FFT -> image (replacing the model image with the
 * dirty image of the residuals).  exec runs it eagerly; capture records it
 * into a hipGraph (after one eager run) that launch replays on the stream. */
int p9idg_exec_cycle(p9idg_plan* plan, int32_t grid_mode);
int p9idg_capture_cycle(p9idg_plan* plan, int32_t grid_mode);
int p9idg_launch_cycle(p9idg_plan* plan);
int p9idg_download_residual(p9idg_plan* plan, void* out_c64);

/* Device pointers of the plan's buffers (for zero-copy interop). */
int p9idg_device_buffers(p9idg_plan* plan, void** vis, void** grid, void** model_grid,
                         void** model_vis, void** image);
/* Host copies of the plan's buffers. */
int p9idg_download_grid(p9idg_plan* plan, int32_t mode, void* out_c64);
int p9idg_upload_model_grid(p9idg_plan* plan, const void* grid_c64);
int p9idg_download_model_grid(p9idg_plan* plan, void* out_c64);
int p9idg_download_model_vis(p9idg_plan* plan, void* out_c64);
int p9idg_upload_image(p9idg_plan* plan, const float* img);
int p9idg_download_image(p9idg_plan* plan, float* out);

/* Milliseconds of the last exec_grid / exec_degrid kernel (HIP events on the
 * plan's stream) and of the last FFT stage since the last p9idg_timing_reset;
 * -1 for a stage not run since. */
int p9idg_last_timings(p9idg_plan* plan, float* grid_ms, float* degrid_ms, float* fft_ms);
/* Collect HIP-event spans around every gridder / degridder kernel launch
 * (collect != 0) from now on, dropping earlier ones and the last timings;
 * totals are summed over the collected spans.  Adjacent spans share their
 * boundary event (one event per stream position), so a span includes the
 * launch gap before its kernel. */
int p9idg_timing_reset(p9idg_plan* plan, int32_t collect);
int p9idg_timing_totals(p9idg_plan* plan, double* grid_ms, int64_t* grid_launches,
                        double* degrid_ms, int64_t* degrid_launches);
/* Stages timed by the collected spans.  ZERO covers the zero fills of the
 * gridder target before each gridder launch (the footprint of the previous
 * gridding, or the whole buffer when its contents are unknown) and of the
 * model visibilities before a degridder launch (when needed); TO_IMAGE is
 * p9idg_exec_grid_to_image (layer sum, inverse FFT, image extraction);
 * TO_UVW is p9idg_exec_image_to_grid (shift + real-to-complex copy, forward
 * FFT).  No counterpart in the reference (its callers use stopwatches). */
typedef enum p9idg_stage {
    P9IDG_STAGE_GRIDDER = 0,
    P9IDG_STAGE_DEGRIDDER = 1,
    P9IDG_STAGE_ZERO = 2,
    P9IDG_STAGE_TO_IMAGE = 3,
    P9IDG_STAGE_TO_UVW = 4
} p9idg_stage;
/* Total milliseconds and number of the collected spans of one stage. */
int p9idg_timing_stage(p9idg_plan* plan, int32_t stage, double* ms_total, int64_t* spans);
/* Number of visibilities covered by the partition (sum of count * C). */
int p9idg_gridded_visibilities(p9idg_plan* plan, int64_t* n);

/* ---- Allocation-free gridding --------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): p9idg_exec_grid's gridder -> subgrid FFT -> adder for
 * `mode` (P9IDG_GRID_PLAIN / WSTACK / WSUM) with `kernel`, without a plan.
 * Inputs: the subgrid metadata and frequencies in host memory (read during
 * the call), uvw double[B][T][3] and visibilities complex64[B][T][C] in device
 * memory; output d_out complex64 [L][G][G] (WSTACK) or [G][G], zeroed first.
 * The work-queue counter, work items, metadata, wavenumbers and tables live in
 * a fixed 4 MiB device array per GPU (allocated with the library's code
 * object, held until the process ends): at most 65536 subgrids and 4096
 * channels (P9IDG_ERR_CAPACITY beyond).  Calls no hipMalloc / hipFree.  The
 * work runs on `stream` (a hipStream_t; NULL: the null stream); the copies of
 * the host arrays wait for the work already on the stream.  Every call on a
 * GPU shares its array: calls that overlap on one GPU (other streams, other
 * threads) race.  Atomic adder (no deterministic mode). */
int p9idg_grid_noalloc(const p9idg_constants* c, int32_t kernel, int32_t mode, const p9idg_subgrid* meta,
                       int32_t nr_subgrids, const double* d_uvw, int32_t nr_baselines, int32_t nr_timesteps,
                       const double* freq, int32_t nr_channels, const void* d_vis, void* d_out, void* stream);

/* ---- Batches: several observations in one persistent launch -------------- */
/* The plans keep their observations and buffers; a batch groups them by
 * (subgrid size, kernel) and runs each group's work items -- the K-split
 * subgrids of all its observations, longest first -- as one persistent
 * work-queued launch per direction, on the batch's own stream.  A batch run
 * (exec, capture's eager run, launch) starts after the work already queued on
 * every member plan and the plans' later work waits for it.  The plans must
 * outlive the batch; after p9idg_set_observation on a member plan the batch
 * returns P9IDG_ERR_STATE until it is prepared (and captured) again. */
typedef struct p9idg_batch p9idg_batch;
int p9idg_batch_create(p9idg_plan* const* plans, int32_t nr_plans, p9idg_batch** batch);
int p9idg_batch_destroy(p9idg_batch* batch);
/* Builds the merged work queues for the given modes (after every plan's
 * observation and kernel are set). */
int p9idg_batch_prepare(p9idg_batch* batch, int32_t grid_mode, int32_t degrid_mode);
/* what: 1 = grid (zero + gridder), 2 = degrid (zero + degridder), 3 = both */
int p9idg_batch_exec(p9idg_batch* batch, int32_t what);
/* Records p9idg_batch_exec(what) into a hipGraph; launch replays it. */
int p9idg_batch_capture(p9idg_batch* batch, int32_t what);
int p9idg_batch_launch(p9idg_batch* batch);
int p9idg_batch_sync(p9idg_batch* batch);
/* kernel-only milliseconds of the last eager p9idg_batch_exec(batch, 3) */
int p9idg_batch_timings(p9idg_batch* batch, float* grid_ms, float* degrid_ms);

#ifdef __cplusplus
}
#endif
#endif
