// FP64 CPU restatement of the reference image-domain gridder (IDG).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that tests, the smoke
// check and the CPU-baseline leg of the benchmark compare the HIP library
// against; the product library never includes, links or calls it.
//
// It follows the reference C# sources line by line (paths below are relative
// to Distributed-Reconstruction/Core/).  Arithmetic copies the C# evaluation
// order and its float/double mix (compile with -ffp-contract=off and without
// -ffast-math so no FMA contraction or reassociation changes the rounding):
//   * GriddingConstants.ImageSize = (float)scale * grid   (GriddingConstants.cs:43)
//   * ComputeL in float, ComputeN mixing float and double   (Gridder.cs:163-174)
//   * Math.Round = round-half-to-even                        (Partitioner.cs:238-239)
//   * Math.Min/Max propagate NaN; (int)double of NaN/overflow = INT32_MIN (x64 CLR)
//   * System.Numerics.Complex multiply: (ac - bd) + (bc + ad)i
// FFTW is replaced by an exact-twiddle mixed-radix FP64 FFT with FFTW's sign
// and normalisation conventions (FORWARD = e^-, BACKWARD = e^+, unnormalised).
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include "oracle.h"

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPI = 3.14159265358979323846;   // System.Math.PI
const double kSpeedOfLight = 299792458.0;    // MathFunctions.cs:11

int g_threads = 0;

struct cpx {
    double re, im;
};

// System.Numerics.Complex operator* (netcoreapp3.0)
inline cpx cmul(cpx a, cpx b) {
    cpx r;
    r.re = (a.re * b.re) - (a.im * b.im);
    r.im = (a.im * b.re) + (a.re * b.im);
    return r;
}
inline cpx cadd(cpx a, cpx b) { return cpx{a.re + b.re, a.im + b.im}; }
inline cpx cconj(cpx a) { return cpx{a.re, -a.im}; }
inline cpx cscale(cpx a, double s) { return cpx{a.re * s, a.im * s}; }

inline cpx* CX(double* p) { return reinterpret_cast<cpx*>(p); }
inline const cpx* CX(const double* p) { return reinterpret_cast<const cpx*>(p); }

// Math.Min / Math.Max (documented: NaN if either argument is NaN).
inline double cs_min(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
inline double cs_max(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
// (int)double on the x64 CLR: truncation; NaN and out-of-range give INT32_MIN.
inline int32_t cs_dtoi(double x) {
    if (std::isnan(x) || x >= 2147483648.0 || x <= -2147483649.0) return INT32_MIN;
    return static_cast<int32_t>(x);
}
// Math.Round(double): MidpointRounding.ToEven.
inline double cs_round(double x) {
    std::fesetround(FE_TONEAREST);
    return std::nearbyint(x);
}

// (cos x, sin x) of System.Math.Cos/Sin.  glibc's sincos returns the same
// values as separate cos and sin calls (orc_sincos_mismatches checks it).
inline cpx cis(double x) {
    cpx r;
    sincos(x, &r.im, &r.re);
    return r;
}

int nthreads() {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// FFT: mixed radix (any n; radices 2,3,4,5 and a plain DFT for other primes),
// twiddles exp(sign*2*pi*i*k/n) from FP64 cos/sin of the reduced angle.
struct Plan1D {
    int n = 0;
    int sign = 0;
    std::vector<int> factors;
    std::vector<cpx> tw;  // tw[k] = exp(sign * 2 pi i k / n)
};

Plan1D make_plan(int n, int sign) {
    Plan1D p;
    p.n = n;
    p.sign = sign;
    int m = n;
    for (int r : {4, 2, 3, 5}) {
        while (m % r == 0) {
            p.factors.push_back(r);
            m /= r;
        }
    }
    for (int r = 7; m > 1; r += 2) {
        while (m % r == 0) {
            p.factors.push_back(r);
            m /= r;
        }
    }
    p.tw.resize(n);
    for (int k = 0; k < n; k++) {
        // angle reduced to [0, 2pi) exactly in integers before scaling
        double a = 2.0 * kPI * (double)k / (double)n;
        p.tw[k] = cpx{std::cos(a), sign * std::sin(a)};
    }
    return p;
}

// Recursive decimation in time: out[0..n) = DFT of in[0], in[s], in[2s], ...
void fft_rec(const Plan1D& p, const cpx* in, int stride, cpx* out, int n, int fi,
             std::vector<cpx>& scratch) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int r = p.factors[fi];
    int m = n / r;
    for (int q = 0; q < r; q++) fft_rec(p, in + q * stride, stride * r, out + q * m, m, fi + 1, scratch);
    int tstep = p.n / n;  // twiddle index scale for length n
    cpx t[64];
    cpx* big = nullptr;
    std::vector<cpx> bigbuf;
    if (r > 64) {
        bigbuf.resize(r);
        big = bigbuf.data();
    }
    cpx* tt = big ? big : t;
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < r; q++) tt[q] = cmul(out[q * m + k], p.tw[(size_t)q * k * tstep % p.n]);
        for (int j = 0; j < r; j++) {
            cpx acc{0.0, 0.0};
            for (int q = 0; q < r; q++) {
                // exp(sign 2 pi i q j / r) = tw[(q*j mod r) * (N/r)]
                int idx = (int)(((long long)q * j % r) * (p.n / r));
                acc = cadd(acc, cmul(tt[q], p.tw[idx]));
            }
            scratch[j] = acc;
        }
        for (int j = 0; j < r; j++) out[j * m + k] = scratch[j];
    }
}

void fft1d(const Plan1D& p, cpx* data, int stride, std::vector<cpx>& buf_in,
           std::vector<cpx>& buf_out, std::vector<cpx>& scratch) {
    int n = p.n;
    for (int i = 0; i < n; i++) buf_in[i] = data[(size_t)i * stride];
    fft_rec(p, buf_in.data(), 1, buf_out.data(), n, 0, scratch);
    for (int i = 0; i < n; i++) data[(size_t)i * stride] = buf_out[i];
}

// In-place unnormalised 2-D DFT of a row-major n0 x n1 array.
void fft2d(int n0, int n1, cpx* data, int sign, bool parallel) {
    Plan1D prow = make_plan(n1, sign);
    Plan1D pcol = make_plan(n0, sign);
    int nmax = std::max(n0, n1);
    int nt = parallel ? nthreads() : 1;
#pragma omp parallel num_threads(nt)
    {
        std::vector<cpx> bi(nmax), bo(nmax), sc(std::max(nmax, 64));
#pragma omp for schedule(static)
        for (int y = 0; y < n0; y++) fft1d(prow, data + (size_t)y * n1, 1, bi, bo, sc);
#pragma omp for schedule(static)
        for (int x = 0; x < n1; x++) fft1d(pcol, data + x, n1, bi, bo, sc);
    }
}

// ---------------------------------------------------------------------------
// Gridder helpers (Gridder.cs:163-174)
inline float compute_l(int x, int subgrid_size, float image_size) {
    float t = (float)(x - (subgrid_size / 2)) + 0.5f;
    t = t * image_size;
    return t / (float)subgrid_size;
}
inline float compute_n(float l, float m) {
    float tmp = (l * l) + (m * m);
    double d = (double)(tmp / 1.0f) + std::sqrt(1.0 - (double)tmp);
    return tmp / (float)d;
}

// MathFunctions.cs:69-78 (the 2.0f literal is promoted to double)
std::vector<double> wavenumbers(const double* freq, int C) {
    std::vector<double> k(C);
    for (int i = 0; i < C; i++) k[i] = ((2.0 * kPI) * freq[i]) / kSpeedOfLight;
    return k;
}

inline float image_size_of(const orc_constants* c) {
    return c->scale_arcsec * (float)c->grid_size;  // GriddingConstants.cs:43
}

template <typename T>
static void shift_impl(int32_t g, T* a) {
    const int n2 = g / 2;
    for (int i = 0; i < n2; i++)
        for (int j = 0; j < n2; j++) {
            std::swap(a[(size_t)i * g + j], a[(size_t)(i + n2) * g + j + n2]);
            std::swap(a[(size_t)(i + n2) * g + j], a[(size_t)i * g + j + n2]);
        }
}
}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }

int64_t orc_sincos_mismatches(int64_t n, uint64_t seed) {
    int64_t bad = 0;
    uint64_t z = seed;
    for (int64_t i = 0; i < n; i++) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t r = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        const double x = ((double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 2.0e5;
        const cpx c = cis(x);
        if (c.re != std::cos(x) || c.im != std::sin(x)) bad++;
    }
    return bad;
}
int orc_get_threads(void) { return nthreads(); }

// ---------------------------------------------------------------------------
// Partitioner.cs:20-124 with the private Subgrid helper at :154-255
int orc_partition(const orc_constants* c, const double* uvw, int32_t B, int32_t T,
                  const double* freq, int32_t C, orc_subgrid* out, int64_t capacity,
                  int64_t* n_out, int64_t* bl_offsets) {
    const double imagesize = (double)image_size_of(c);  // :22 (float), promoted at :41-42
    std::vector<orc_subgrid> result;
    struct DP {
        double u, v, w;
    };
    std::vector<DP> dps((size_t)T * (C > 0 ? C : 1));
    int status = ORC_OK;
    if (bl_offsets) bl_offsets[0] = 0;
    for (int bl = 0; bl < B && status == ORC_OK; bl++) {
        // :30-47 datapoints[time, channel]
        for (int t = 0; t < T; t++)
            for (int ch = 0; ch < C; ch++) {
                const double* p = uvw + ((size_t)bl * T + t) * 3;
                double f = freq[ch];
                DP& d = dps[(size_t)t * C + ch];
                d.u = (p[0] * imagesize) * (f / kSpeedOfLight);  // MetersToPixels :127
                d.v = (p[1] * imagesize) * (f / kSpeedOfLight);
                d.w = p[2] * (f / kSpeedOfLight);  // MetersToLambda :129
            }
        // Subgrid state (:154-184)
        double uMin, uMax, vMin, vMax;
        int32_t wIndex;
        auto reset = [&]() {
            uMin = std::numeric_limits<double>::infinity();
            uMax = -std::numeric_limits<double>::infinity();
            vMin = std::numeric_limits<double>::infinity();
            vMax = -std::numeric_limits<double>::infinity();
            wIndex = 0;
        };
        // AddVisibility :186-221
        auto add = [&](double u, double v, double w) -> bool {
            if (std::isinf(u) || std::isinf(v)) return false;
            int32_t wi = 0;
            if (c->wstep != 0.0) wi = cs_dtoi(std::floor(w / c->wstep));
            if (!std::isinf(uMin) && wIndex != wi) return false;
            double uMin_ = cs_min(uMin, u);
            double uMax_ = cs_max(uMax, u);
            double vMin_ = cs_min(vMin, v);
            double vMax_ = cs_max(vMax, v);
            double uvWidth_ = cs_max(uMax_ - uMin_, vMax_ - vMin_);
            if ((uvWidth_ + (double)c->kernel_size) >= (double)c->subgrid_size) return false;
            uMin = uMin_;
            uMax = uMax_;
            vMin = vMin_;
            vMax = vMax_;
            wIndex = wi;
            return true;
        };
        int time = 0;
        reset();
        while (time < T) {  // :51
            reset();
            int n = 0;
            for (; time < T; time++) {  // :57-77
                const DP& d0 = dps[(size_t)time * C + 0];
                const DP& dm = dps[(size_t)time * C + (C - 1)];
                if (add(d0.u, d0.v, d0.w) && add(dm.u, dm.v, d0.w)) {  // hack: chMax with ch0's w
                    n++;
                    if (n == c->max_timesteps_per_subgrid) {
                        time++;
                        break;
                    }
                } else {
                    break;
                }
            }
            if (n == 0) {  // :80-99
                const DP& d = dps[(size_t)time * C + 0];
                if (std::isinf(d.u) || std::isinf(d.v)) {
                    status = ORC_ERR_PLACE;
                    break;
                }
                time++;
                continue;
            }
            // Finish -> ComputeCoordinates :236-248
            int32_t U = cs_dtoi(cs_round((uMax + uMin) / 2));
            int32_t V = cs_dtoi(cs_round((vMax + vMin) / 2));
            U += c->grid_size / 2;
            V += c->grid_size / 2;
            U -= c->subgrid_size / 2;
            V -= c->subgrid_size / 2;
            // InRange :223-234
            int32_t mx = std::max(U, V), mn = std::min(U, V);
            bool in_range = mn >= 1 && mx <= (c->grid_size - c->subgrid_size) &&
                            wIndex >= -c->wlayer_count && wIndex < c->wlayer_count;
            if (!in_range) {
                status = ORC_ERR_SUBGRID_RANGE;
                break;
            }
            orc_subgrid s;
            s.baseline_idx = bl;
            s.time_sample_start = time - n;
            s.time_sample_count = n;
            s.u_pixel = U;
            s.v_pixel = V;
            s.w_lambda = wIndex;
            result.push_back(s);
        }
        if (bl_offsets) bl_offsets[bl + 1] = (int64_t)result.size();
    }
    if (n_out) *n_out = (int64_t)result.size();
    if (status != ORC_OK) return status;
    if (out) {
        if ((int64_t)result.size() > capacity) return ORC_ERR_CAPACITY;
        if (!result.empty()) std::memcpy(out, result.data(), result.size() * sizeof(orc_subgrid));
    }
    return ORC_OK;
}

// ---------------------------------------------------------------------------
// Gridder.cs:15-87
void orc_gridder_forward(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                         const double* uvw, int32_t B, int32_t T, const double* vis,
                         const double* freq, int32_t C, double* subgrids_out) {
    (void)B;
    const int S = c->subgrid_size, G = c->grid_size;
    const std::vector<double> k = wavenumbers(freq, C);
    const float imagesize = image_size_of(c);
    const cpx* V = CX(vis);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
    for (int64_t s = 0; s < n; s++) {
        const orc_subgrid& m = meta[s];
        cpx* out = CX(subgrids_out) + (size_t)s * S * S;
        double uOffset = (double)(m.u_pixel + S / 2 - G / 2) * (2 * kPI / (double)imagesize);
        double vOffset = (double)(m.v_pixel + S / 2 - G / 2) * (2 * kPI / (double)imagesize);
        double tmpW = (double)c->wstep_lambda * (m.w_lambda + 0.5);
        double wOffset = 2 * kPI * tmpW;
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                cpx pixel{0.0, 0.0};
                float l = compute_l(x, S, imagesize);
                float mm = compute_l(y, S, imagesize);
                float nn = compute_n(l, mm);
                int end = m.time_sample_start + m.time_sample_count;
                for (int t = m.time_sample_start; t < end; t++) {
                    const double* p = uvw + ((size_t)m.baseline_idx * T + t) * 3;
                    double u = p[0], v = p[1], w = p[2];
                    double phaseIndex = u * l + v * mm + w * nn;
                    double phaseOffset = uOffset * l + vOffset * mm + wOffset * nn;
                    for (int ch = 0; ch < C; ch++) {
                        double phase = phaseOffset - (phaseIndex * k[ch]);
                        cpx phasor = cis(phase);
                        cpx vv = V[((size_t)m.baseline_idx * T + t) * C + ch];
                        pixel = cadd(pixel, cmul(vv, phasor));
                    }
                }
                double sph = 1.0f;  // identity spheroidal (GriddingConstants.cs:49)
                int xDest = (x + (S / 2)) % S;
                int yDest = (y + (S / 2)) % S;
                out[(size_t)yDest * S + xDest] = cscale(pixel, sph);
            }
    }
}

// Gridder.cs:91-160
void orc_gridder_backward(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                          const double* subgrids, const double* uvw, int32_t B, int32_t T,
                          const double* freq, int32_t C, double* vis_out) {
    const int S = c->subgrid_size, G = c->grid_size;
    const std::vector<double> k = wavenumbers(freq, C);
    const float imagesize = image_size_of(c);
    cpx* out = CX(vis_out);
    std::memset(vis_out, 0, sizeof(double) * 2 * (size_t)B * T * C);  // :96
    const double scale = (double)(1.0f / (float)(S * S));             // :151 float reciprocal
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
    for (int64_t s = 0; s < n; s++) {
        const orc_subgrid& m = meta[s];
        const cpx* data = CX(subgrids) + (size_t)s * S * S;
        std::vector<cpx> px((size_t)S * S);
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                double sph = 1.0f;
                int xSrc = (x + (S / 2)) % S;
                int ySrc = (y + (S / 2)) % S;
                px[(size_t)y * S + x] = cscale(data[(size_t)ySrc * S + xSrc], sph);
            }
        double uOffset = (double)(m.u_pixel + S / 2 - G / 2) * (2 * kPI / (double)imagesize);
        double vOffset = (double)(m.v_pixel + S / 2 - G / 2) * (2 * kPI / (double)imagesize);
        double tmpW = (double)c->wstep_lambda * (m.w_lambda + 0.5);
        double wOffset = 2 * kPI * tmpW;
        int end = m.time_sample_start + m.time_sample_count;
        for (int t = m.time_sample_start; t < end; t++) {
            const double* p = uvw + ((size_t)m.baseline_idx * T + t) * 3;
            double u = p[0], v = p[1], w = p[2];
            for (int ch = 0; ch < C; ch++) {
                cpx visibility{0.0, 0.0};
                for (int y = 0; y < S; y++)
                    for (int x = 0; x < S; x++) {
                        float l = compute_l(x, S, imagesize);
                        float mm = compute_l(y, S, imagesize);
                        float nn = compute_n(l, mm);
                        double phaseIndex = u * l + v * mm + w * nn;
                        double phaseOffset = uOffset * l + vOffset * mm + wOffset * nn;
                        double phase = (phaseIndex * k[ch]) - phaseOffset;
                        cpx phasor = cis(phase);
                        visibility = cadd(visibility, cmul(px[(size_t)y * S + x], phasor));
                    }
                out[((size_t)m.baseline_idx * T + t) * C + ch] = cscale(visibility, scale);
            }
        }
    }
}

// SubgridFFT.cs:11-52: FFTW BACKWARD (e^+), times 1/S^2
void orc_subgrid_fft_forward(const orc_constants* c, int64_t n, double* subgrids) {
    const int S = c->subgrid_size;
    const double norm = 1.0 / (S * S);
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t s = 0; s < n; s++) {
        cpx* d = CX(subgrids) + (size_t)s * S * S;
        fft2d(S, S, d, +1, false);
        for (int i = 0; i < S * S; i++) d[i] = cscale(d[i], norm);
    }
}

// SubgridFFT.cs:54-90: FFTW FORWARD (e^-), no normalisation
void orc_subgrid_fft_backward(const orc_constants* c, int64_t n, double* subgrids) {
    const int S = c->subgrid_size;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t s = 0; s < n; s++) fft2d(S, S, CX(subgrids) + (size_t)s * S * S, -1, false);
}

// Adder.cs:13-55
int orc_adder_add(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                  const double* subgrids, double* grid_out) {
    const int S = c->subgrid_size, G = c->grid_size;
    std::memset(grid_out, 0, sizeof(double) * 2 * (size_t)G * G);
    cpx* grid = CX(grid_out);
    for (int64_t s = 0; s < n; s++) {
        const cpx* data = CX(subgrids) + (size_t)s * S * S;
        int gridX = meta[s].u_pixel, gridY = meta[s].v_pixel;
        if (gridX >= 0 && gridX < G - S && gridY >= 0 && gridY < G - S) {
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    int xSrc = (x + (S / 2)) % S;
                    int ySrc = (y + (S / 2)) % S;
                    double phase = kPI * (double)(x + y - S) / (double)S;
                    cpx phasor{std::cos(phase), std::sin(phase)};
                    cpx& g = grid[(size_t)(gridY + y) * G + gridX + x];
                    g = cadd(g, cmul(data[(size_t)ySrc * S + xSrc], phasor));
                }
        } else {
            return ORC_ERR_INVALID_GRID;
        }
    }
    return ORC_OK;
}

// Adder.cs:57-95
void orc_adder_split(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                     const double* grid_in, double* subgrids_out) {
    const int S = c->subgrid_size, G = c->grid_size;
    const cpx* grid = CX(grid_in);
    std::memset(subgrids_out, 0, sizeof(double) * 2 * (size_t)n * S * S);
    for (int64_t s = 0; s < n; s++) {
        cpx* data = CX(subgrids_out) + (size_t)s * S * S;
        int gridX = meta[s].u_pixel, gridY = meta[s].v_pixel;
        if (gridX >= 0 && gridX < G - S && gridY >= 0 && gridY < G - S) {
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    int xDst = (x + (S / 2)) % S;
                    int yDst = (y + (S / 2)) % S;
                    double phase = -kPI * (double)(x + y - S) / (double)S;
                    cpx phasor{std::cos(phase), std::sin(phase)};
                    data[(size_t)yDst * S + xDst] =
                        cmul(grid[(size_t)(gridY + y) * G + gridX + x], phasor);
                }
        }
    }
}

// AdderWStack.cs:11-78
void orc_adderw_add(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                    const double* subgrids, double* layers_out) {
    const int S = c->subgrid_size, G = c->grid_size, L = c->wlayer_count;
    std::memset(layers_out, 0, sizeof(double) * 2 * (size_t)L * G * G);
    std::vector<cpx> pre((size_t)S * S);
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            double phase = kPI * (double)(x + y - S) / (double)S;
            pre[(size_t)y * S + x] = cpx{std::cos(phase), std::sin(phase)};
        }
    for (int64_t s = 0; s < n; s++) {
        const cpx* data = CX(subgrids) + (size_t)s * S * S;
        int X = meta[s].u_pixel, Y = meta[s].v_pixel, W = meta[s].w_lambda;
        bool neg = W < 0;
        if (neg) {
            X = G - X - S + 1;
            Y = G - Y - S + 1;
            W = -W - 1;
        }
        if (!(X >= 1 && X < G - S && Y >= 1 && Y < G - S)) continue;
        cpx* layer = CX(layers_out) + (size_t)W * G * G;
        for (int y = 0; y < S; y++) {
            int y_ = neg ? S - 1 - y : y;
            for (int x = 0; x < S; x++) {
                int x_ = neg ? S - 1 - x : x;
                int xSrc = (x_ + (S / 2)) % S;
                int ySrc = (y_ + (S / 2)) % S;
                cpx value = cmul(pre[(size_t)y_ * S + x_], data[(size_t)ySrc * S + xSrc]);
                if (neg) value = cconj(value);
                cpx& g = layer[(size_t)(Y + y) * G + X + x];
                g = cadd(g, value);
            }
        }
    }
}

// AdderWStack.cs:81-139
void orc_adderw_split(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                      const double* grid_in, double* subgrids_out) {
    const int S = c->subgrid_size, G = c->grid_size;
    const cpx* grid = CX(grid_in);
    std::vector<cpx> pre((size_t)S * S);
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            double phase = -kPI * (double)(x + y - S) / (double)S;
            pre[(size_t)y * S + x] = cpx{std::cos(phase), std::sin(phase)};
        }
    std::memset(subgrids_out, 0, sizeof(double) * 2 * (size_t)n * S * S);
    for (int64_t s = 0; s < n; s++) {
        cpx* data = CX(subgrids_out) + (size_t)s * S * S;
        int X = meta[s].u_pixel, Y = meta[s].v_pixel, W = meta[s].w_lambda;
        bool neg = W < 0;
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                int xDst = (x + (S / 2)) % S;
                int yDst = (y + (S / 2)) % S;
                int xSrc = neg ? G - X - x : X + x;
                int ySrc = neg ? G - Y - y : Y + y;
                if (X >= 1 && X < G - S && Y >= 1 && Y < G - S) {
                    cpx value = grid[(size_t)ySrc * G + xSrc];
                    if (neg) value = cconj(value);
                    data[(size_t)yDst * S + xDst] = cmul(pre[(size_t)y * S + x], value);
                }
            }
    }
}

// ---------------------------------------------------------------------------
// FFT.cs:248-264 / :221-241: swap quadrants 1<->3 and 2<->4, n2 = rows/2
void orc_shift_c128(int32_t g, double* grid) { shift_impl(g, CX(grid)); }
void orc_shift_f64(int32_t g, double* img) { shift_impl(g, img); }
void orc_shift_f32(int32_t g, float* img) { shift_impl(g, img); }

// FFT.cs:188-215
void orc_wstack_ifft_float(int32_t g, int32_t L, const double* layers, int64_t vis_count,
                           float* out) {
    std::memset(out, 0, sizeof(float) * (size_t)g * g);
    std::vector<cpx> buf((size_t)g * g);
    for (int k = 0; k < L; k++) {
        std::memcpy(buf.data(), layers + (size_t)k * g * g * 2, sizeof(cpx) * (size_t)g * g);
        fft2d(g, g, buf.data(), +1, true);
        for (size_t i = 0; i < (size_t)g * g; i++)
            out[i] += (float)(buf[i].re / (double)vis_count);
    }
}

// FFT.cs:159-185
void orc_fft_backward(int32_t g, const double* grid, int64_t vis_count, double* out) {
    std::vector<cpx> buf((size_t)g * g);
    std::memcpy(buf.data(), grid, sizeof(cpx) * (size_t)g * g);
    fft2d(g, g, buf.data(), +1, true);
    for (size_t i = 0; i < (size_t)g * g; i++) out[i] = buf[i].re / (double)vis_count;
}

// FFT.cs:139-157
void orc_fft_backward_float(int32_t g, const double* grid, double norm, float* out) {
    std::vector<cpx> buf((size_t)g * g);
    std::memcpy(buf.data(), grid, sizeof(cpx) * (size_t)g * g);
    fft2d(g, g, buf.data(), +1, true);
    for (size_t i = 0; i < (size_t)g * g; i++) out[i] = (float)(buf[i].re / norm);
}

// FFT.cs:99-117
void orc_fft_forward_float(int32_t g, const float* img, double norm, double* out) {
    cpx* o = CX(out);
    for (size_t i = 0; i < (size_t)g * g; i++) o[i] = cpx{(double)img[i], 0.0};
    fft2d(g, g, o, -1, true);
    for (size_t i = 0; i < (size_t)g * g; i++) o[i] = cpx{o[i].re / norm, o[i].im / norm};
}

void orc_fft2(int32_t n0, int32_t n1, double* data, int sign) {
    fft2d(n0, n1, CX(data), sign < 0 ? -1 : +1, true);
}

// Common.cs:305-318
void orc_substract(int64_t n, const double* vis0, const double* vis1, const uint8_t* flags,
                   double* out) {
    for (int64_t i = 0; i < n; i++) {
        if (!flags || !flags[i]) {
            out[2 * i] = vis0[2 * i] - vis1[2 * i];
            out[2 * i + 1] = vis0[2 * i + 1] - vis1[2 * i + 1];
        } else {
            out[2 * i] = 0.0;
            out[2 * i + 1] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// IDG.cs facade
int orc_grid(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
             const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
             double* grid_out) {
    const int S = c->subgrid_size;
    std::vector<double> sub((size_t)n * S * S * 2);
    orc_gridder_forward(c, meta, n, uvw, B, T, vis, freq, C, sub.data());
    orc_subgrid_fft_forward(c, n, sub.data());
    int st = orc_adder_add(c, meta, n, sub.data(), grid_out);
    if (st) return st;
    orc_shift_c128(c->grid_size, grid_out);
    return ORC_OK;
}

void orc_degrid(const orc_constants* c, const orc_subgrid* meta, int64_t n, double* grid,
                const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                double* vis_out) {
    const int S = c->subgrid_size;
    orc_shift_c128(c->grid_size, grid);
    std::vector<double> sub((size_t)n * S * S * 2);
    orc_adder_split(c, meta, n, grid, sub.data());
    orc_subgrid_fft_backward(c, n, sub.data());
    orc_gridder_backward(c, meta, n, sub.data(), uvw, B, T, freq, C, vis_out);
}

void orc_gridw(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
               const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
               double* layers_out) {
    const int S = c->subgrid_size, G = c->grid_size;
    std::vector<double> sub((size_t)n * S * S * 2);
    orc_gridder_forward(c, meta, n, uvw, B, T, vis, freq, C, sub.data());
    orc_subgrid_fft_forward(c, n, sub.data());
    orc_adderw_add(c, meta, n, sub.data(), layers_out);
    for (int k = 0; k < c->wlayer_count; k++) orc_shift_c128(G, layers_out + (size_t)k * G * G * 2);
}

void orc_degridw(const orc_constants* c, const orc_subgrid* meta, int64_t n, double* grid,
                 const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                 double* vis_out) {
    const int S = c->subgrid_size;
    orc_shift_c128(c->grid_size, grid);
    std::vector<double> sub((size_t)n * S * S * 2);
    orc_adderw_split(c, meta, n, grid, sub.data());
    orc_subgrid_fft_backward(c, n, sub.data());
    orc_gridder_backward(c, meta, n, sub.data(), uvw, B, T, freq, C, vis_out);
}

int orc_to_image(const orc_constants* c, const orc_subgrid* meta, int64_t n, const double* vis,
                 const double* uvw, int32_t B, int32_t T, const double* freq, int32_t C,
                 double* img_out) {
    const int G = c->grid_size;
    std::vector<double> grid((size_t)G * G * 2);
    int st = orc_grid(c, meta, n, vis, uvw, B, T, freq, C, grid.data());
    if (st) return st;
    orc_fft_backward(G, grid.data(), c->visibilities_count, img_out);
    orc_shift_f64(G, img_out);
    return ORC_OK;
}

void orc_gridw_to_image(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                        const double* vis, const double* uvw, int32_t B, int32_t T,
                        const double* freq, int32_t C, float* img_out) {
    const int G = c->grid_size, L = c->wlayer_count;
    std::vector<double> layers((size_t)L * G * G * 2);
    orc_gridw(c, meta, n, vis, uvw, B, T, freq, C, layers.data());
    orc_wstack_ifft_float(G, L, layers.data(), c->visibilities_count, img_out);
    orc_shift_f32(G, img_out);
}

void orc_image_to_vis(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                      const float* img, const double* uvw, int32_t B, int32_t T,
                      const double* freq, int32_t C, double* vis_out) {
    const int G = c->grid_size;
    std::vector<float> x(img, img + (size_t)G * G);
    orc_shift_f32(G, x.data());
    std::vector<double> grid((size_t)G * G * 2);
    orc_fft_forward_float(G, x.data(), 1.0, grid.data());
    orc_degridw(c, meta, n, grid.data(), uvw, B, T, freq, C, vis_out);
}

}  // extern "C"
