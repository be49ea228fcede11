// Seeded synthetic observations (host).  Conventions are stated in p9obs.h.
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include "p9obs.h"

#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 299792458.0;
constexpr double kSidereal = 7.2921158553e-5;  // rad/s

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" {

void p9obs_uniform(uint64_t seed, int64_t n, double* out) {
    SplitMix64 g{seed};
    for (int64_t i = 0; i < n; i++) out[i] = g.uniform();
}

void p9obs_normal(uint64_t seed, int64_t n, double* out) {
    SplitMix64 g{seed};
    for (int64_t i = 0; i < n; i += 2) {
        const double u1 = 1.0 - g.uniform();  // (0, 1]
        const double u2 = g.uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = r * std::cos(2.0 * kPi * u2);
        if (i + 1 < n) out[i + 1] = r * std::sin(2.0 * kPi * u2);
    }
}

void p9obs_antennas(uint64_t seed, int32_t n_core, double core_diam_m, int32_t n_outer,
                    double outer_diam_m, double* enu) {
    SplitMix64 g{seed};
    const int32_t n = n_core + n_outer;
    for (int32_t a = 0; a < n; a++) {
        const double radius = 0.5 * (a < n_core ? core_diam_m : outer_diam_m);
        const double r = radius * std::sqrt(g.uniform());
        const double phi = 2.0 * kPi * g.uniform();
        enu[3 * a + 0] = r * std::cos(phi);
        enu[3 * a + 1] = r * std::sin(phi);
        enu[3 * a + 2] = 0.0;
    }
}

void p9obs_uvw(int32_t n_ant, const double* enu, double lat, double dec, int32_t T, double dt,
               double* uvw) {
    const double sl = std::sin(lat), cl = std::cos(lat);
    const double sd = std::sin(dec), cd = std::cos(dec);
    int64_t b = 0;
    for (int32_t p = 0; p < n_ant; p++)
        for (int32_t q = p + 1; q < n_ant; q++, b++) {
            const double e = enu[3 * q] - enu[3 * p];
            const double nn = enu[3 * q + 1] - enu[3 * p + 1];
            const double up = enu[3 * q + 2] - enu[3 * p + 2];
            // local ENU -> equatorial XYZ (X toward H = 0, Z toward the pole)
            const double X = -sl * nn + cl * up;
            const double Y = e;
            const double Z = cl * nn + sl * up;
            for (int32_t t = 0; t < T; t++) {
                const double H = kSidereal * dt * ((double)t - 0.5 * (double)(T - 1));
                const double sh = std::sin(H), ch = std::cos(H);
                double* o = uvw + ((size_t)b * T + t) * 3;
                o[0] = sh * X + ch * Y;
                o[1] = -sd * ch * X + sd * sh * Y + cd * Z;
                o[2] = cd * ch * X - cd * sh * Y + sd * Z;
            }
        }
}

// Phases of point sources are evaluated for all sources of one (baseline,
// time) at once so the sincos calls vectorise; Gaussians (few) go one by one.
void p9obs_sky_vis(int32_t B, int32_t T, int32_t C, const double* uvw, const double* freq,
                   int32_t n_src, const p9obs_source* src, double* vis, int32_t threads) {
    std::vector<double> pl, pm, pn, pf;  // point sources: l, m, n - 1, flux
    std::vector<int32_t> gauss;
    for (int32_t s = 0; s < n_src; s++) {
        const double nm1 = std::sqrt(1.0 - src[s].l * src[s].l - src[s].m * src[s].m) - 1.0;
        if (src[s].sigma > 0.0) {
            gauss.push_back(s);
        } else {
            pl.push_back(src[s].l);
            pm.push_back(src[s].m);
            pn.push_back(nm1);
            pf.push_back(src[s].flux);
        }
    }
    const int32_t np = (int32_t)pl.size();
    // Evenly spaced channels (f_ch = f_0 + ch * df to 1e-12 relative) use the
    // rotation exp(-2 pi i d f_ch / c) = E_0 * R^ch with one sincos pair for E_0
    // and one for R; other frequency sets evaluate every channel directly.
    bool even = C > 2;
    const double df = C > 1 ? (freq[C - 1] - freq[0]) / (double)(C - 1) : 0.0;
    for (int32_t ch = 0; ch < C && even; ch++)
        even = std::fabs(freq[ch] - (freq[0] + ch * df)) <= 1e-12 * std::fabs(freq[0]);
    const int64_t BT = (int64_t)B * T;
#ifdef _OPENMP
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#else
    (void)threads;  // one thread without OpenMP
#endif
    {
        std::vector<double> er(np), ei(np), rr(np), ri(np), ang(np);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t bt = 0; bt < BT; bt++) {
            const double* p = uvw + bt * 3;
            double* out = vis + bt * C * 2;
            const double u = p[0], v = p[1], w = p[2];
            if (even) {
                const double k0 = -2.0 * kPi * (freq[0] / kC), k1 = -2.0 * kPi * (df / kC);
#pragma omp simd
                for (int32_t s = 0; s < np; s++) {
                    const double d = u * pl[s] + v * pm[s] + w * pn[s];
                    er[s] = std::cos(d * k0);
                    ei[s] = std::sin(d * k0);
                    rr[s] = std::cos(d * k1);
                    ri[s] = std::sin(d * k1);
                }
                for (int32_t ch = 0; ch < C; ch++) {
                    double sr = 0.0, si = 0.0;
#pragma omp simd reduction(+ : sr, si)
                    for (int32_t s = 0; s < np; s++) {
                        sr += pf[s] * er[s];
                        si += pf[s] * ei[s];
                        const double nr = er[s] * rr[s] - ei[s] * ri[s];
                        ei[s] = er[s] * ri[s] + ei[s] * rr[s];
                        er[s] = nr;
                    }
                    out[2 * ch] = sr;
                    out[2 * ch + 1] = si;
                }
            } else {
                for (int32_t ch = 0; ch < C; ch++) {
                    const double k = -2.0 * kPi * (freq[ch] / kC);
                    double sr = 0.0, si = 0.0;
#pragma omp simd reduction(+ : sr, si)
                    for (int32_t s = 0; s < np; s++) {
                        const double d = u * pl[s] + v * pm[s] + w * pn[s];
                        sr += pf[s] * std::cos(d * k);
                        si += pf[s] * std::sin(d * k);
                    }
                    out[2 * ch] = sr;
                    out[2 * ch + 1] = si;
                }
            }
            const double r2 = u * u + v * v;
            for (int32_t s : gauss) {
                const double nm1 = std::sqrt(1.0 - src[s].l * src[s].l - src[s].m * src[s].m) - 1.0;
                const double d = u * src[s].l + v * src[s].m + w * nm1;
                for (int32_t ch = 0; ch < C; ch++) {
                    const double k = freq[ch] / kC;
                    const double amp = src[s].flux * std::exp(-2.0 * kPi * kPi * src[s].sigma *
                                                              src[s].sigma * r2 * k * k);
                    const double ph = -2.0 * kPi * d * k;
                    out[2 * ch] += amp * std::cos(ph);
                    out[2 * ch + 1] += amp * std::sin(ph);
                }
            }
        }
    }
}

void p9obs_add_noise(int64_t n, double sigma, uint64_t seed, double* vis) {
    if (sigma == 0.0) return;
    std::vector<double> z(2 * (size_t)n);
    p9obs_normal(seed, 2 * n, z.data());
    for (int64_t i = 0; i < 2 * n; i++) vis[i] += sigma * z[i];
}

}  // extern "C"
