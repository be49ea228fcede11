/*
 * p9fits.h - FITS readers and writers for the measurement sets and images
 * the reconstruction consumes and produces (host code in libp9idg.so).
 *
 * Reference interfaces replaced (paths relative to Distributed-Reconstruction/):
 *   Core/FitsIO.cs:184-189    FitsIO.ReadFrequencies(file)            -> p9fits_read_frequencies
 *   Core/FitsIO.cs:196-219    FitsIO.ReadUVW(file)                    -> p9fits_read_uvw (0, -1)
 *   Core/FitsIO.cs:221-241    FitsIO.ReadUVW(file, fromBl, toBl)      -> p9fits_read_uvw
 *   Core/FitsIO.cs:247-255    FitsIO.CountBaselines(file)             -> p9fits_count_baselines
 *   Core/FitsIO.cs:267-294    FitsIO.ReadVisibilities(file, B, T, C, norm)
 *   Core/FitsIO.cs:296-325    FitsIO.ReadVisibilities(file, fromBl, toBl, T, C, norm)
 *                                                                     -> p9fits_read_visibilities
 *   Core/FitsIO.cs:327-357    FitsIO.ReadFlags(file, B, T, C)
 *   Core/FitsIO.cs:359-390    FitsIO.ReadFlags(file, fromBl, toBl, T, C)  -> p9fits_read_flags
 *   Core/FitsIO.cs:152-166    FitsIO.ReadImage(file)        (float)   -> p9fits_read_image F32
 *   Core/FitsIO.cs:168-182    FitsIO.ReadImageDouble(file)  (double)  -> p9fits_read_image F64
 *   Core/FitsIO.cs:90-107     FitsIO.ReadBeam(file)         (double, square)
 *   Core/FitsIO.cs:134-150    FitsIO.ReadCASAFits(file)     (float[1][1][rows][cols])
 *                                                                     -> p9fits_read_casa
 *   Core/FitsIO.cs:16-37,66-88,109-130  FitsIO.Write(Complex[,]) / Write<T>(T[,]) / WriteImag
 *                                                                     -> p9fits_write_image
 *
 * Format: the primary HDU of a FITS file (FITS Standard 4.0): 2880-byte
 * records, 80-character header cards, big-endian data with NAXIS1 varying
 * fastest.  The reference casts the primary array to nested double[] (or
 * float[]) arrays, so each reader accepts exactly the BITPIX and NAXIS those
 * casts accept and reports any other with P9IDG_ERR_IO; BSCALE and BZERO are
 * not applied (the reference reads the raw kernel).  A reader that is given
 * fewer baselines, timesteps or channels than the file holds reads the leading
 * block, as the reference's loops do; more than the file holds is an error
 * (the reference's IndexOutOfRangeException).
 *
 * Files are memory-mapped and decoded on the host threads set by
 * p9fits_set_threads (default: one per hardware thread, at most 64).  Status
 * codes and p9idg_last_error() are those of p9idg.h.
 */
#ifndef P9FITS_H
#define P9FITS_H

#include <stdint.h>

#include "p9idg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define P9FITS_MAX_AXES 8

typedef struct p9fits_info {
    int32_t bitpix;                  /* 8, 16, 32, 64, -32 or -64 */
    int32_t naxis;
    int64_t naxes[P9FITS_MAX_AXES];  /* naxes[0] = NAXIS1, the fastest-varying axis */
    int64_t data_offset;             /* byte offset of the primary array */
    int64_t data_bytes;              /* |bitpix| / 8 * prod(naxes), before padding */
} p9fits_info;

/* Element types of p9fits_read_image / p9fits_write_image */
typedef enum p9fits_type {
    P9FITS_F32 = 0,       /* float, BITPIX -32 */
    P9FITS_F64 = 1,       /* double, BITPIX -64 */
    P9FITS_C128_REAL = 2, /* write only: real parts of complex128 pixels as BITPIX -64 (Write(Complex[,])) */
    P9FITS_C128_IMAG = 3  /* write only: imaginary parts as BITPIX -64 (WriteImag) */
} p9fits_type;

/* Header of the primary HDU. */
int p9fits_read_info(const char* path, p9fits_info* out);

/* ReadFrequencies: a one-axis BITPIX -64 array.  out may be NULL to query the
 * count; otherwise it holds `capacity` doubles. */
int p9fits_read_frequencies(const char* path, double* out, int64_t capacity, int64_t* count);

/* CountBaselines, plus the timesteps per baseline, of a uvw file
 * (double[B][T][3]: NAXIS3 = B, NAXIS2 = T, NAXIS1 >= 3). */
int p9fits_count_baselines(const char* path, int32_t* baselines, int32_t* timesteps);

/* ReadUVW(file, fromBl, toBl): baselines [from_bl, to_bl) (to_bl = -1: all),
 * every timestep, into uvw double[to_bl - from_bl][T][3] with v negated. */
int p9fits_read_uvw(const char* path, int32_t from_bl, int32_t to_bl, double* uvw);

/* ReadVisibilities(file, fromBl, toBl, T, C, norm): a double[B][T][C][P][2]
 * file (P >= 4 polarizations XX, XY, YX, YY); vis[b][t][c] =
 * ((XX.re + YY.re) / norm, (XX.im + YY.im) / norm) evaluated in double, for
 * baselines [from_bl, to_bl) (to_bl = -1: all), the first `timesteps`
 * timesteps and `channels` channels.  vis_type P9IDG_C128 stores the doubles,
 * P9IDG_C64 rounds them to float. */
int p9fits_read_visibilities(const char* path, int32_t from_bl, int32_t to_bl, int32_t timesteps,
                             int32_t channels, double norm, void* vis, int32_t vis_type);

/* ReadFlags(file, fromBl, toBl, T, C): a double[B][T][C][P] file; flags[b][t][c]
 * = 1 when the P values summed in order are not 0 (NaN counts as flagged),
 * else 0.  Same ranges as p9fits_read_visibilities. */
int p9fits_read_flags(const char* path, int32_t from_bl, int32_t to_bl, int32_t timesteps,
                      int32_t channels, uint8_t* flags);

/* ReadImage (P9FITS_F32, BITPIX -32) or ReadImageDouble / ReadBeam
 * (P9FITS_F64, BITPIX -64) of a two-axis array into out[rows][cols];
 * rows = NAXIS2 and cols = NAXIS1 must match the file (read them with
 * p9fits_read_info). */
int p9fits_read_image(const char* path, int32_t type, void* out, int32_t rows, int32_t cols);

/* ReadCASAFits: plane [0][0] of a four-axis BITPIX -32 array
 * float[NAXIS4][NAXIS3][rows][cols] into out[rows][cols]. */
int p9fits_read_casa(const char* path, float* out, int32_t rows, int32_t cols);

/* Write / Write<T> / WriteImag: a primary HDU with rows = NAXIS2 and
 * cols = NAXIS1, the file created or truncated. */
int p9fits_write_image(const char* path, int32_t type, const void* img, int32_t rows, int32_t cols);

/* Host threads the readers decode on (threads <= 0: the default). */
int p9fits_set_threads(int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
