/*
 * p9obs -- seeded synthetic radio-interferometric observations (host, C ABI).
 *
 * Produces the inputs of the gridding API in the reference's layouts
 * (uvw double[B][T][3] in metres, vis complex128[B][T][C]).  The reference
 * loads these from FITS files (Core/FitsIO.cs:184-219); the measurement data
 * repository is not available, so inputs are synthesised deterministically:
 *
 *  - PRNG: SplitMix64; uniform doubles from the top 53 bits; normals from the
 *    Box-Muller transform (both written here, so the stream does not depend on
 *    a C++ standard library's distributions).
 *  - Antennas: uniform in a disk of the given DIAMETER (radius = D/2) in the
 *    local East-North plane, up = 0.
 *  - uvw: ENU -> equatorial XYZ at the site latitude, then the standard
 *    (u, v, w) rotation for hour angle H and declination dec; baseline
 *    (p, q), p < q, is antenna q minus antenna p; H steps by the sidereal
 *    rate times dt, centred on transit.
 *  - Sky: V(u,v,w) = sum_s F_s * exp(-2 pi^2 sigma_s^2 (u^2+v^2)) *
 *        exp(-2 pi i (u l_s + v m_s + w (sqrt(1 - l_s^2 - m_s^2) - 1)))
 *    with u, v, w in wavelengths; sigma_s = 0 for a point source.  A source
 *    at (l, m) = (dx, dy) * cell appears at image[G/2 + dy][G/2 + dx] of the
 *    dirty image.
 *  - Noise: real and imaginary parts each N(0, sigma^2).
 */
#ifndef P9OBS_H
#define P9OBS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct p9obs_source {
    double flux;  /* Jy (total flux for a Gaussian) */
    double l, m;  /* direction cosines */
    double sigma; /* Gaussian sigma in radians; 0 = point source */
} p9obs_source;

/* n uniform doubles in [0, 1) from SplitMix64(seed) */
void p9obs_uniform(uint64_t seed, int64_t n, double* out);
/* n standard normals (Box-Muller on the same stream) */
void p9obs_normal(uint64_t seed, int64_t n, double* out);

/* Antennas uniform in disks: n_core in a disk of diameter core_diam_m, then
 * n_outer in a disk of diameter outer_diam_m.  enu double[n_core+n_outer][3]. */
void p9obs_antennas(uint64_t seed, int32_t n_core, double core_diam_m, int32_t n_outer,
                    double outer_diam_m, double* enu);

/* uvw double[N(N-1)/2][T][3] */
void p9obs_uvw(int32_t n_ant, const double* enu, double latitude_rad, double dec_rad,
               int32_t T, double dt_s, double* uvw);

/* vis complex128[B][T][C] = sky model (overwrites) */
void p9obs_sky_vis(int32_t B, int32_t T, int32_t C, const double* uvw, const double* freq,
                   int32_t n_src, const p9obs_source* src, double* vis, int32_t threads);

/* vis += complex Gaussian noise (re, im ~ N(0, sigma^2)) */
void p9obs_add_noise(int64_t n, double sigma, uint64_t seed, double* vis);

#ifdef __cplusplus
}
#endif
#endif
