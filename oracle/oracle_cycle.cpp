// FP64/FP32 CPU restatement of the reference's serial major cycle and its
// deconvolver.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Follows Distributed-Reconstruction/Core/Common.cs (PSF, Residuals,
// ElasticNet), Core/PaddedConvolver.cs, Core/Deconvolution/FastSerialCD.cs and
// SingleReconstruction/MajorCycle.cs:33-134.  Float expressions of the C# code
// are evaluated in float with one rounding per operation (the library is
// built with -ffp-contract=off); double expressions in double.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle.h"

namespace {

inline size_t at(int32_t row, int32_t col, int32_t ncols) { return (size_t)row * ncols + col; }

// Math.Max(float, 0f) as .NET Core evaluates it for finite input (+0 for -0).
inline float max0(float v) { return v > 0.f ? v : 0.f; }

// Common.cs:27-29
inline float proximal(float x, float lipschitz, float lambda, float alpha) {
    const float num = max0(x - lambda * alpha);
    const float den = lipschitz + lambda * (1.f - alpha);
    return num / den;
}

// FFT.Shift (FFT.cs:248-264) on an n x n complex array.
void shift_square(std::vector<double>& a, int32_t n) {
    const int32_t n2 = n / 2;
    for (int32_t i = 0; i < n2; i++)
        for (int32_t j = 0; j < n2; j++)
            for (int k = 0; k < 2; k++) {
                std::swap(a[2 * at(i, j, n) + k], a[2 * at(i + n2, j + n2, n) + k]);
                std::swap(a[2 * at(i + n2, j, n) + k], a[2 * at(i, j + n2, n) + k]);
            }
}

struct Pixel {
    int32_t y, x;
    float diff, xnew;
};

// FastSerialCD.cs:138-172: per row the first x with the largest |x - prox|,
// then the first row with the largest of those.  The rows are scanned in
// parallel (each row's best is independent of the others) and the rows'
// bests are then compared in row order, as the reference's serial loop does.
Pixel abs_max(const float* amap, const float* x, const float* g, int32_t H, int32_t W, float lambda,
              float alpha) {
    std::vector<Pixel> rows((size_t)std::max(H, 0));
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (int32_t y = 0; y < H; y++) {
        Pixel row{-1, -1, 0.f, 0.f};
        for (int32_t xx = 0; xx < W; xx++) {
            const size_t i = at(y, xx, W);
            const float a = amap[i];
            const float old = x[i];
            const float tmp = proximal(old * a + g[i], a, lambda, alpha);
            const float diff = old - tmp;
            if (row.diff < std::fabs(diff)) row = Pixel{y, xx, std::fabs(diff), tmp};
        }
        rows[y] = row;
    }
    Pixel best{-1, -1, 0.f, 0.f};
    for (int32_t y = 0; y < H; y++)
        if (best.diff < rows[y].diff) best = rows[y];
    return best;
}

}  // namespace

extern "C" {

// Common.cs:227-237 (square PSF)
void orc_psf_cut(const float* psf, int32_t n, int32_t factor, float* out) {
    const int32_t m = n / factor;
    const int32_t off = n / 2 - m / 2;
    for (int32_t y = 0; y < m; y++)
        for (int32_t x = 0; x < m; x++) out[at(y, x, m)] = psf[at(off + y, off + x, n)];
}

// Common.cs:206-218
float orc_psf_max_sidelobe(const float* psf, int32_t n, int32_t factor) {
    const int32_t m = n / factor;
    const int32_t off = n / 2 - m / 2;
    float out = 0.f;
    for (int32_t y = 0; y < n; y++)
        for (int32_t x = 0; x < n; x++) {
            const bool in_y = y >= off && y < off + m, in_x = x >= off && x < off + m;
            if (!in_y || !in_x) out = std::fmax(out, std::fabs(psf[at(y, x, n)]));
        }
    return out;
}

// Common.cs:128-135
double orc_psf_max_lipschitz(const float* psf, int32_t h, int32_t w) {
    double sum = 0.0;
    for (int64_t i = 0; i < (int64_t)h * w; i++) sum += (double)(psf[i] * psf[i]);
    return sum;
}

// Common.cs:48-126: prefix sums of psf^2 (double) and, per pixel, the sum
// over the PSF pixels that fall inside the H x W image when the PSF is
// centred there (QueryScan's case analysis).
void orc_psf_amap(const float* psf, int32_t ph, int32_t pw, int32_t H, int32_t W, float* amap) {
    std::vector<double> s((size_t)ph * pw);
    for (int32_t i = 0; i < ph; i++)
        for (int32_t j = 0; j < pw; j++) {
            const double up = i > 0 ? s[at(i - 1, j, pw)] : 0.0;
            const double left = j > 0 ? s[at(i, j - 1, pw)] : 0.0;
            const double diag = (i > 0 && j > 0) ? s[at(i - 1, j - 1, pw)] : 0.0;
            const float cur = psf[at(i, j, pw)] * psf[at(i, j, pw)];
            s[at(i, j, pw)] = (double)cur + up + left - diag;
        }
    auto S = [&](int32_t i, int32_t j) { return s[at(i, j, pw)]; };
    for (int32_t y = 0; y < H; y++)
        for (int32_t x = 0; x < W; x++) {
            const int32_t y_over = std::max(0, y + (ph - ph / 2) - H);
            const int32_t x_over = std::max(0, x + (pw - pw / 2) - W);
            const int32_t y_under = -(y - ph / 2), x_under = -(x - pw / 2);
            const int32_t yu = std::max(1, y_under) - 1, xu = std::max(1, x_under) - 1;
            double v;
            if (y_over <= 0 && x_over <= 0 && y_under <= 0 && x_under <= 0) {
                v = S(ph - 1, pw - 1);
            } else if (y_under > 0 && x_under > 0) {
                v = S(ph - 1, pw - 1) - S(ph - 1, xu) - S(yu, pw - 1) + S(yu, xu);
            } else {
                double corr = 0.0;
                if (y_under > 0) corr += S(yu, pw - x_over - 1);
                if (x_under > 0) corr += S(ph - y_over - 1, xu);
                v = S(ph - 1 - y_over, pw - 1 - x_over) - corr;
            }
            amap[at(y, x, W)] = (float)v;
        }
}

// Common.cs:146-160: the flipped PSF in a (ypad + ph) x (xpad + pw) double
// array, FFT.Shift-ed, forward FFT; `out` is complex (re, im) row-major.
void orc_psf_padded_correlation(const float* psf, int32_t ph, int32_t pw, int32_t ypad,
                                int32_t xpad, double* out) {
    const int32_t n0 = ypad + ph, n1 = xpad + pw;
    std::vector<double> a((size_t)n0 * n1 * 2, 0.0);
    const int32_t yh = ypad / 2, xh = xpad / 2;
    for (int32_t y = 0; y < ph; y++)
        for (int32_t x = 0; x < pw; x++)
            if (y + yh + 1 < n0 && x + xh + 1 < n1)
                a[2 * at(y + yh + 1, x + xh + 1, n1)] = (double)psf[at(ph - y - 1, pw - x - 1, pw)];
    shift_square(a, n0);
    orc_fft2(n0, n1, a.data(), -1);
    std::memcpy(out, a.data(), sizeof(double) * a.size());
}

// PaddedConvolver.cs:44-66 with the kernel of orc_psf_padded_correlation
// (n0 x n1 = (H + kh) x (W + kw)) and kernelSize (0, 0, kh, kw).
void orc_padded_convolve(const double* kernel, int32_t kh, int32_t kw, const float* image, int32_t H,
                         int32_t W, float* out) {
    const int32_t n0 = H + kh, n1 = W + kw;
    const int32_t yh = kh / 2, xh = kw / 2;
    std::vector<double> b((size_t)n0 * n1 * 2, 0.0);
    for (int32_t i = yh; i < H + yh; i++)
        for (int32_t j = xh; j < W + xh; j++) b[2 * at(i, j, n1)] = (double)image[at(i - yh, j - xh, W)];
    orc_fft2(n0, n1, b.data(), -1);
    for (size_t i = 0; i < (size_t)n0 * n1; i++) {
        const double re = b[2 * i], im = b[2 * i + 1], kr = kernel[2 * i], ki = kernel[2 * i + 1];
        b[2 * i] = re * kr - im * ki;
        b[2 * i + 1] = re * ki + im * kr;
    }
    orc_fft2(n0, n1, b.data(), +1);
    const double len = (double)n0 * n1;
    for (int32_t i = 0; i < H; i++)
        for (int32_t j = 0; j < W; j++) out[at(i, j, W)] = (float)(b[2 * at(i + yh, j + xh, n1)] / len);
}

// Common.cs:182-194 (SetPsfInWindow always lands inside the 2ph x 2pw window)
void orc_psf_squared(const float* psf, int32_t ph, int32_t pw, float* out) {
    const int32_t n0 = 2 * ph, n1 = 2 * pw;
    std::vector<double> corr((size_t)n0 * n1 * 2);
    orc_psf_padded_correlation(psf, ph, pw, ph, pw, corr.data());
    std::vector<double> p((size_t)n0 * n1 * 2, 0.0);
    for (int32_t i = 0; i < ph; i++)
        for (int32_t j = 0; j < pw; j++) p[2 * at(i + ph / 2, j + pw / 2, n1)] = (double)psf[at(i, j, pw)];
    orc_fft2(n0, n1, p.data(), -1);
    for (size_t i = 0; i < (size_t)n0 * n1; i++) {
        const double re = p[2 * i], im = p[2 * i + 1], kr = corr[2 * i], ki = corr[2 * i + 1];
        p[2 * i] = re * kr - im * ki;
        p[2 * i + 1] = re * ki + im * kr;
    }
    orc_fft2(n0, n1, p.data(), +1);
    const double len = (double)n0 * n1;
    for (size_t i = 0; i < (size_t)n0 * n1; i++) out[i] = (float)(p[2 * i] / len);
}

// Common.cs:244-251
float orc_residuals_get_max(const float* img, int64_t n) {
    float m = 0.f;
    for (int64_t i = 0; i < n; i++) m = std::fmax(m, img[i]);
    return m;
}

// Common.cs:253-260
double orc_residuals_penalty(const float* img, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += (double)(img[i] * img[i]);
    return 0.5 * s;
}

// Common.cs:22-23, 31-40
double orc_elastic_penalty(const float* img, int64_t n, float lambda, float alpha) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const float v = img[i];
        const float pen = 0.5f * (1.f - alpha) * (v * v) + alpha * std::fabs(v);
        s += (double)lambda * (double)pen;
    }
    return s;
}

float orc_proximal(float x, float lipschitz, float lambda, float alpha) {
    return proximal(x, lipschitz, lambda, alpha);
}

// FastSerialCD.cs:132-135
float orc_cd_abs_max_diff(const float* amap, const float* x, const float* g, int32_t H, int32_t W,
                          float lambda, float alpha) {
    return abs_max(amap, x, g, H, W, lambda, alpha).diff;
}

// FastSerialCD.cs:77-101 with UpdateGradients (:219-238): x and g are
// updated in place; returns 1 when converged.
int orc_cd_deconvolve(const float* amap, const float* psf2, int32_t p2h, int32_t p2w, float* x, float* g,
                      int32_t H, int32_t W, float lambda, float alpha, int32_t iterations, float epsilon,
                      int32_t* iters_out) {
    bool converged = false;
    int32_t iter = 0;
    const int32_t yh = p2h / 2, xh = p2w / 2;
    while (!converged && iter < iterations) {
        const Pixel m = abs_max(amap, x, g, H, W, lambda, alpha);
        converged = m.diff < epsilon;
        if (converged) break;
        const float old = x[at(m.y, m.x, W)];
        x[at(m.y, m.x, W)] = m.xnew;
        const float d = old - m.xnew;
        const int32_t y0 = std::max(m.y - yh, 0), x0 = std::max(m.x - xh, 0);
        const int32_t y1 = std::min(m.y - yh + p2h, H), x1 = std::min(m.x - xh + p2w, W);
        for (int32_t i = y0; i < y1; i++)
            for (int32_t j = x0; j < x1; j++)
                g[at(i, j, W)] += psf2[at(i + yh - m.y, j + xh - m.x, p2w)] * d;
        iter++;
    }
    if (iters_out) *iters_out = iter;
    return converged ? 1 : 0;
}

// MajorCycle.cs:33-134 with the CPU deconvolver (FastSerialCD) and without
// the FITS output.  vis/flags as the IDG functions take them.
int orc_reconstruct_serial_cd(const orc_constants* c, const orc_subgrid* meta, int64_t n,
                              const double* vis, const uint8_t* flags, const double* uvw, int32_t B,
                              int32_t T, const double* freq, int32_t C, int32_t psf_cut_factor,
                              int32_t max_major, float lambda, float alpha, int32_t deconv_iterations,
                              float deconv_epsilon, float* x_image, orc_cycle_stats* stats,
                              int32_t* ncycles) {
    const float kMajorEpsilon = 1e-4f;
    const int32_t G = c->grid_size;
    const int64_t nv = (int64_t)B * T * C, GG = (int64_t)G * G;
    // PSF (MajorCycle.cs:36-48)
    std::vector<double> psf_vis((size_t)nv * 2, 0.0);
    for (int64_t i = 0; i < nv; i++) psf_vis[2 * i] = (flags && flags[i]) ? 0.0 : 1.0;
    std::vector<float> psf((size_t)GG);
    orc_gridw_to_image(c, meta, n, psf_vis.data(), uvw, B, T, freq, C, psf.data());
    const int32_t pc = G / psf_cut_factor;
    std::vector<float> psf_cut((size_t)pc * pc);
    orc_psf_cut(psf.data(), G, psf_cut_factor, psf_cut.data());
    const float max_sidelobe = orc_psf_max_sidelobe(psf.data(), G, psf_cut_factor);
    // FastSerialCD(totalSize, psfCut)
    std::vector<float> psf2((size_t)4 * pc * pc), amap((size_t)GG);
    orc_psf_squared(psf_cut.data(), pc, pc, psf2.data());
    orc_psf_amap(psf_cut.data(), pc, pc, G, G, amap.data());
    // gCalculator (the second convolver of the reference never computes a gradient)
    std::vector<double> kernel((size_t)(G + pc) * (G + pc) * 2);
    orc_psf_padded_correlation(psf_cut.data(), pc, pc, G, G, kernel.data());
    double max_lipschitz = orc_psf_max_lipschitz(psf_cut.data(), pc, pc);
    const float lambda_lipschitz = (float)((double)lambda * max_lipschitz);
    const float lambda_true = (float)((double)lambda * orc_psf_max_lipschitz(psf.data(), G, G));
    bool switched = false;

    std::memset(x_image, 0, sizeof(float) * (size_t)GG);
    std::vector<double> residual(vis, vis + (size_t)nv * 2), model((size_t)nv * 2);
    std::vector<float> dirty((size_t)GG), grad((size_t)GG), xs((size_t)GG);
    std::vector<double> xgrid((size_t)GG * 2);
    int32_t cycles = 0;
    for (int32_t cycle = 0; cycle < max_major; cycle++) {
        orc_gridw_to_image(c, meta, n, residual.data(), uvw, B, T, freq, C, dirty.data());
        const float max_dirty = orc_residuals_get_max(dirty.data(), GG);
        orc_padded_convolve(kernel.data(), pc, pc, dirty.data(), G, G, grad.data());
        const float max_b = orc_residuals_get_max(grad.data(), GG);
        const double correction = std::fmax((double)max_b / ((double)max_dirty * max_lipschitz), 1.0);
        const double sidelobe = (double)(max_b * max_sidelobe) * correction;
        const float current_lambda =
            (float)std::fmax(sidelobe / (double)alpha, (double)lambda_lipschitz);
        const double objective = orc_residuals_penalty(dirty.data(), GG) +
                                 orc_elastic_penalty(x_image, GG, lambda_true, alpha);
        const float abs_max_diff =
            orc_cd_abs_max_diff(amap.data(), x_image, grad.data(), G, G, lambda_true, alpha);
        int32_t iters = 0, conv = 0;
        if (abs_max_diff >= kMajorEpsilon)
            conv = orc_cd_deconvolve(amap.data(), psf2.data(), 2 * pc, 2 * pc, x_image, grad.data(), G, G,
                                     current_lambda, alpha, deconv_iterations, deconv_epsilon, &iters);
        if (stats) {
            orc_cycle_stats& s = stats[cycle];
            s.max_dirty = max_dirty;
            s.max_gradient = max_b;
            s.lambda = current_lambda;
            s.abs_max_diff = abs_max_diff;
            s.objective = objective;
            s.iterations = iters;
            s.converged = conv;
        }
        cycles = cycle + 1;
        if (lambda == current_lambda && !switched) {
            lambda = lambda_true;
            max_lipschitz = orc_psf_max_lipschitz(psf.data(), G, G);
            switched = true;
        }
        if (abs_max_diff < kMajorEpsilon) break;
        // ToUVW (MajorCycle.cs:125-129)
        std::memcpy(xs.data(), x_image, sizeof(float) * (size_t)GG);
        orc_shift_f32(G, xs.data());
        orc_fft_forward_float(G, xs.data(), 1.0, xgrid.data());
        orc_degridw(c, meta, n, xgrid.data(), uvw, B, T, freq, C, model.data());
        orc_substract(nv, vis, model.data(), flags, residual.data());
    }
    if (ncycles) *ncycles = cycles;
    return ORC_OK;
}

}  // extern "C"
