// Host side of the serial major cycle: PSF helpers, the padded PSF
// correlation (PaddedConvolver), the serial coordinate-descent deconvolver
// (FastSerialCD) on the GPU, and MajorCycle.ReconstructSerialCD over the
// plan API of p9idg.h.  Scalar float / double expressions follow the
// reference's types, one rounding per operation.
#pragma clang fp contract(off)

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cd_internal.h"
#include "p9cycle.h"
#include "p9fits.h"
#include "p9idg_internal.h"

using namespace p9idg;

namespace {

struct Guard {
    int prev = -1;
    explicit Guard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~Guard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T>
int alloc(T** p, size_t n) {
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();  // reported here, not by a later launch check
        return fail(P9IDG_ERR_ALLOC, "device allocation failed");
    }
    return P9IDG_OK;
}
template <typename T>
void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int check_dims(int32_t a, int32_t b) {
    if (a <= 0 || b <= 0) return fail(P9IDG_ERR_ARGUMENT, "dimensions must be positive");
    return P9IDG_OK;
}

// FP64 correlation spectrum of PSF.CalcPaddedFourierCorrelation (Common.cs:146-160)
// into `out` (N0 x N1 = (ypad + ph) x (xpad + pw)); `plan` is a Z2Z plan of that size.
int correlation_spectrum(const float* d_psf, int ph, int pw, int ypad, int xpad, hipfftHandle plan, double2* out,
                         hipStream_t st) {
    const int N0 = ypad + ph, N1 = xpad + pw;
    P9_HIP(cd::launch_place_correlation(d_psf, ph, pw, ypad, xpad, N0, N1, out, st));
    P9_FFT(hipfftExecZ2Z(plan, reinterpret_cast<hipfftDoubleComplex*>(out),
                         reinterpret_cast<hipfftDoubleComplex*>(out), HIPFFT_FORWARD));
    return P9IDG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// PaddedConvolver
struct p9conv {
    int device = 0;
    int H = 0, W = 0, ph = 0, pw = 0, N0 = 0, N1 = 0;
    hipStream_t stream = nullptr;
    hipfftHandle plan = 0;
    bool have_plan = false;
    double2* d_kernel = nullptr;
    double2* d_buf = nullptr;
    float* d_io = nullptr;  // [2][H][W] staging for host-array calls
};

namespace {

int conv_run(p9conv* c, const float* d_in, float* d_out) {
    const size_t n = (size_t)c->N0 * c->N1;
    P9_HIP(cd::launch_insert_image(d_in, c->H, c->W, c->ph / 2, c->pw / 2, c->N0, c->N1, c->d_buf, c->stream));
    P9_FFT(hipfftExecZ2Z(c->plan, reinterpret_cast<hipfftDoubleComplex*>(c->d_buf),
                         reinterpret_cast<hipfftDoubleComplex*>(c->d_buf), HIPFFT_FORWARD));
    P9_HIP(cd::launch_multiply(c->d_buf, c->d_kernel, n, c->stream));
    P9_FFT(hipfftExecZ2Z(c->plan, reinterpret_cast<hipfftDoubleComplex*>(c->d_buf),
                         reinterpret_cast<hipfftDoubleComplex*>(c->d_buf), HIPFFT_BACKWARD));
    P9_HIP(cd::launch_extract(c->d_buf, c->N1, c->ph / 2, c->pw / 2, (double)n, c->H, c->W, d_out, c->stream));
    P9_HIP(hipStreamSynchronize(c->stream));
    return P9IDG_OK;
}

}  // namespace

extern "C" {

int p9conv_create_correlation(int32_t device, const float* psf, int32_t rows, int32_t cols, int32_t height,
                              int32_t width, p9conv** out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (int st = check_dims(rows, cols)) return st;
    if (int st = check_dims(height, width)) return st;
    if (height + rows != width + cols)
        return fail(P9IDG_ERR_UNSUPPORTED, "the padded correlation needs a square padded size (FFT.Shift)");
    Guard g(device);
    auto* c = new p9conv();
    c->device = device;
    c->H = height;
    c->W = width;
    c->ph = rows;
    c->pw = cols;
    c->N0 = height + rows;
    c->N1 = width + cols;
    auto bail = [&](int st) {
        p9conv_destroy(c);
        return st;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(P9IDG_ERR_HIP, "stream creation failed"));
    if (hipfftPlan2d(&c->plan, c->N0, c->N1, HIPFFT_Z2Z) != HIPFFT_SUCCESS)
        return bail(fail(P9IDG_ERR_HIP, "hipFFT plan failed"));
    c->have_plan = true;
    if (hipfftSetStream(c->plan, c->stream) != HIPFFT_SUCCESS) return bail(fail(P9IDG_ERR_HIP, "hipfftSetStream"));
    const size_t n = (size_t)c->N0 * c->N1;
    int st;
    if ((st = alloc(&c->d_kernel, n)) || (st = alloc(&c->d_buf, n)) ||
        (st = alloc(&c->d_io, 2 * (size_t)height * width)))
        return bail(st);
    float* d_psf = nullptr;
    if ((st = alloc(&d_psf, (size_t)rows * cols))) return bail(st);
    hipError_t e = hipMemcpy(d_psf, psf, sizeof(float) * rows * cols, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        st = correlation_spectrum(d_psf, rows, cols, height, width, c->plan, c->d_kernel, c->stream);
        if (!st) e = hipStreamSynchronize(c->stream);
    }
    release(d_psf);
    if (e != hipSuccess) return bail(hip_fail(e, "correlation kernel"));
    if (st) return bail(st);
    *out = c;
    return P9IDG_OK;
}

int p9conv_destroy(p9conv* c) {
    if (!c) return P9IDG_OK;
    Guard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->have_plan) (void)hipfftDestroy(c->plan);
    release(c->d_kernel);
    release(c->d_buf);
    release(c->d_io);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return P9IDG_OK;
}

int p9conv_convolve_device(p9conv* c, const float* d_image, float* d_out) {
    if (!c || !d_image || !d_out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    return conv_run(c, d_image, d_out);
}

int p9conv_convolve(p9conv* c, const float* image, float* out) {
    if (!c || !image || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    const size_t n = (size_t)c->H * c->W;
    P9_HIP(hipMemcpyAsync(c->d_io, image, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (int st = conv_run(c, c->d_io, c->d_io + n)) return st;
    P9_HIP(hipMemcpyAsync(out, c->d_io + n, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    P9_HIP(hipStreamSynchronize(c->stream));
    return P9IDG_OK;
}

// ---------------------------------------------------------------------------
// PSF helpers
int p9psf_cut(const float* psf, int32_t n, int32_t factor, float* out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (n <= 0 || factor <= 0) return fail(P9IDG_ERR_ARGUMENT, "size and factor must be positive");
    const int32_t m = n / factor, off = n / 2 - m / 2;  // Common.cs:229-231
    for (int32_t y = 0; y < m; y++)
        std::memcpy(out + (size_t)y * m, psf + (size_t)(off + y) * n + off, sizeof(float) * m);
    return P9IDG_OK;
}

int p9psf_max_sidelobe(const float* psf, int32_t n, int32_t factor, float* out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (n <= 0 || factor <= 0) return fail(P9IDG_ERR_ARGUMENT, "size and factor must be positive");
    const int32_t m = n / factor, off = n / 2 - m / 2;  // Common.cs:208-209
    float v = 0.f;
    for (int32_t y = 0; y < n; y++) {
        const bool in_y = y >= off && y < off + m;
        for (int32_t x = 0; x < n; x++)
            if (!in_y || x < off || x >= off + m) v = std::fmax(v, std::fabs(psf[(size_t)y * n + x]));
    }
    *out = v;
    return P9IDG_OK;
}

int p9psf_max_lipschitz(const float* psf, int32_t rows, int32_t cols, double* out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (int st = check_dims(rows, cols)) return st;
    double s = 0.0;
    for (size_t i = 0; i < (size_t)rows * cols; i++) {
        const float sq = psf[i] * psf[i];
        s += (double)sq;
    }
    *out = s;
    return P9IDG_OK;
}

int p9psf_squared(int32_t device, const float* psf, int32_t rows, int32_t cols, float* out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (int st = check_dims(rows, cols)) return st;
    if (rows != cols) return fail(P9IDG_ERR_UNSUPPORTED, "PSF^2 needs a square PSF (FFT.Shift)");
    Guard g(device);
    const int N0 = 2 * rows, N1 = 2 * cols;
    const size_t n = (size_t)N0 * N1;
    hipStream_t st = nullptr;
    hipfftHandle plan = 0;
    bool have_plan = false;
    double2 *d_corr = nullptr, *d_p = nullptr;
    float* d_psf = nullptr;
    int status = P9IDG_OK;
    do {
        if (hipStreamCreate(&st) != hipSuccess) {
            status = fail(P9IDG_ERR_HIP, "stream creation failed");
            break;
        }
        if (hipfftPlan2d(&plan, N0, N1, HIPFFT_Z2Z) != HIPFFT_SUCCESS || hipfftSetStream(plan, st) != HIPFFT_SUCCESS) {
            status = fail(P9IDG_ERR_HIP, "hipFFT plan failed");
            break;
        }
        have_plan = true;
        if ((status = alloc(&d_corr, n)) || (status = alloc(&d_p, n)) || (status = alloc(&d_psf, (size_t)rows * cols)))
            break;
        hipError_t e = hipMemcpy(d_psf, psf, sizeof(float) * rows * cols, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            status = hip_fail(e, "psf upload");
            break;
        }
        // CalcPSFSquared: the correlation spectrum of the PSF padded by its
        // own size, times the spectrum of the centred PSF, back, / N0 N1
        if ((status = correlation_spectrum(d_psf, rows, cols, rows, cols, plan, d_corr, st))) break;
        e = cd::launch_place_centered(d_psf, rows, cols, N0, N1, d_p, st);
        if (e == hipSuccess) {
            if (hipfftExecZ2Z(plan, reinterpret_cast<hipfftDoubleComplex*>(d_p), reinterpret_cast<hipfftDoubleComplex*>(d_p),
                              HIPFFT_FORWARD) != HIPFFT_SUCCESS) {
                status = fail(P9IDG_ERR_HIP, "hipfftExecZ2Z");
                break;
            }
            e = cd::launch_multiply(d_p, d_corr, n, st);
        }
        if (e == hipSuccess &&
            hipfftExecZ2Z(plan, reinterpret_cast<hipfftDoubleComplex*>(d_p), reinterpret_cast<hipfftDoubleComplex*>(d_p),
                          HIPFFT_BACKWARD) != HIPFFT_SUCCESS) {
            status = fail(P9IDG_ERR_HIP, "hipfftExecZ2Z");
            break;
        }
        // reuse d_corr's memory as the float output (N0 x N1)
        float* d_res = reinterpret_cast<float*>(d_corr);
        if (e == hipSuccess) e = cd::launch_extract(d_p, N1, 0, 0, (double)n, N0, N1, d_res, st);
        if (e == hipSuccess) e = hipMemcpyAsync(out, d_res, sizeof(float) * n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) status = hip_fail(e, "PSF^2");
    } while (false);
    if (st) (void)hipStreamSynchronize(st);
    if (have_plan) (void)hipfftDestroy(plan);
    release(d_corr);
    release(d_p);
    release(d_psf);
    if (st) (void)hipStreamDestroy(st);
    return status;
}

}  // extern "C"

namespace {

// Prefix sums of psf^2 (Common.cs:53-64): float squares, double sums.
std::vector<double> psf_scan(const float* psf, int ph, int pw) {
    std::vector<double> s((size_t)ph * pw);
    for (int i = 0; i < ph; i++)
        for (int j = 0; j < pw; j++) {
            const double up = i > 0 ? s[(size_t)(i - 1) * pw + j] : 0.0;
            const double left = j > 0 ? s[(size_t)i * pw + j - 1] : 0.0;
            const double diag = (i > 0 && j > 0) ? s[(size_t)(i - 1) * pw + j - 1] : 0.0;
            const float cur = psf[(size_t)i * pw + j] * psf[(size_t)i * pw + j];
            s[(size_t)i * pw + j] = (double)cur + up + left - diag;
        }
    return s;
}

int amap_device(const float* psf, int ph, int pw, int H, int W, float* d_amap, hipStream_t st) {
    const std::vector<double> s = psf_scan(psf, ph, pw);
    double* d_s = nullptr;
    if (int r = alloc(&d_s, s.size())) return r;
    hipError_t e = hipMemcpyAsync(d_s, s.data(), sizeof(double) * s.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = cd::launch_amap(d_s, ph, pw, H, W, d_amap, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release(d_s);
    if (e != hipSuccess) return hip_fail(e, "Lipschitz map");
    return P9IDG_OK;
}

}  // namespace

extern "C" int p9psf_amap(int32_t device, const float* psf, int32_t rows, int32_t cols, int32_t height, int32_t width,
                          float* out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (int st = check_dims(rows, cols)) return st;
    if (int st = check_dims(height, width)) return st;
    Guard g(device);
    float* d = nullptr;
    if (int st = alloc(&d, (size_t)height * width)) return st;
    int st = amap_device(psf, rows, cols, height, width, d, nullptr);
    if (!st) {
        const hipError_t e = hipMemcpy(out, d, sizeof(float) * height * width, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_fail(e, "Lipschitz map download");
    }
    release(d);
    return st;
}

// ---------------------------------------------------------------------------
// FastSerialCD
struct p9cd {
    int device = 0;
    int H = 0, W = 0, p2h = 0, p2w = 0, ntiles = 0;
    float max_lipschitz = 0.f;
    hipStream_t stream = nullptr;
    float* d_amap = nullptr;
    float* d_psf2 = nullptr;
    cd::Cand* d_tiles = nullptr;
    cd::Cand* d_rows = nullptr;
    cd::State* d_state = nullptr;
    cd::State* h_state = nullptr;  // pinned host copy of the state (hipHostMalloc)

    float* d_io = nullptr;  // [2][H][W] staging for host-array calls

    // kBatch iterations recorded as a graph for one (x, gradients) buffer
    // pair; the parameters are read from the device state, so every call on
    // the same buffers replays it
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const float* graph_x = nullptr;
    const float* graph_g = nullptr;
};

namespace {

constexpr int kBatch = 128;  // iterations enqueued between host checks of the state

int cd_abs_max(p9cd* c, const float* d_x, const float* d_g, float lambda, float alpha, float* out) {
    P9_HIP(cd::launch_tiles_all(d_x, d_g, c->d_amap, c->H, c->W, lambda, alpha, c->d_tiles, c->stream));
    P9_HIP(cd::launch_absmax(c->d_tiles, c->ntiles, c->d_state, c->stream));
    P9_HIP(hipMemcpyAsync(c->h_state, c->d_state, sizeof(cd::State), hipMemcpyDeviceToHost, c->stream));
    P9_HIP(hipStreamSynchronize(c->stream));
    *out = c->h_state->absmax;
    return P9IDG_OK;
}

void cd_drop_graph(p9cd* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_x = c->graph_g = nullptr;
}

// Records kBatch iterations on (d_x, d_g) as a graph, unless the recorded
// one already works on these buffers.  Called with the stream idle.
int cd_graph(p9cd* c, float* d_x, float* d_g) {
    if (c->graph_exec && c->graph_x == d_x && c->graph_g == d_g) return P9IDG_OK;
    cd_drop_graph(c);
    P9_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < kBatch && e == hipSuccess; k++)
        e = cd::launch_iteration(d_x, d_g, c->d_amap, c->d_psf2, c->p2h, c->p2w, c->H, c->W, c->d_tiles, c->d_rows,
                                 c->d_state, c->stream);
    hipGraph_t graph = nullptr;
    const hipError_t e_end = hipStreamEndCapture(c->stream, &graph);
    if (e == hipSuccess) e = e_end;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(e, "deconvolution graph");
    }
    c->graph = graph;
    c->graph_exec = exec;
    c->graph_x = d_x;
    c->graph_g = d_g;
    return P9IDG_OK;
}

// The first selection, then the recorded iterations replayed until the state
// says done (launches after that return at once).  The state travels through
// pinned host memory.
int cd_run(p9cd* c, float* d_x, float* d_g, float lambda, float alpha, int32_t max_iter, float eps, int32_t* conv,
           int32_t* iters) {
    cd::State& s = *c->h_state;
    s = cd::State{};
    s.done = max_iter <= 0;
    s.lambda = lambda;
    s.alpha = alpha;
    s.epsilon = eps;
    s.max_iter = max_iter;
    P9_HIP(hipMemcpyAsync(c->d_state, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
    P9_HIP(cd::launch_tiles_all(d_x, d_g, c->d_amap, c->H, c->W, lambda, alpha, c->d_tiles, c->stream));
    P9_HIP(cd::launch_first_select(c->d_tiles, c->d_rows, c->H, c->W, c->d_state, d_x, c->stream));
    P9_HIP(hipMemcpyAsync(&s, c->d_state, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    P9_HIP(hipStreamSynchronize(c->stream));
    if (!(s.done && !s.updated)) {
        if (int st = cd_graph(c, d_x, d_g)) return st;
        while (!s.done || s.updated) {
            P9_HIP(hipGraphLaunch(c->graph_exec, c->stream));
            P9_HIP(hipMemcpyAsync(&s, c->d_state, sizeof(s), hipMemcpyDeviceToHost, c->stream));
            P9_HIP(hipStreamSynchronize(c->stream));
        }
    }
    if (conv) *conv = s.converged;
    if (iters) *iters = s.iters;
    return P9IDG_OK;
}

}  // namespace

extern "C" {

int p9cd_create(int32_t device, int32_t height, int32_t width, const float* psf, int32_t rows, int32_t cols,
                p9cd** out) {
    if (!psf || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (int st = check_dims(height, width)) return st;
    if (int st = check_dims(rows, cols)) return st;
    Guard g(device);
    auto* c = new p9cd();
    c->device = device;
    c->H = height;
    c->W = width;
    c->p2h = 2 * rows;
    c->p2w = 2 * cols;
    c->ntiles = ((height + cd::kTile - 1) / cd::kTile) * ((width + cd::kTile - 1) / cd::kTile);
    auto bail = [&](int st) {
        p9cd_destroy(c);
        return st;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(P9IDG_ERR_HIP, "stream creation failed"));
    int st;
    if ((st = alloc(&c->d_amap, (size_t)height * width)) || (st = alloc(&c->d_psf2, (size_t)c->p2h * c->p2w)) ||
        (st = alloc(&c->d_tiles, (size_t)c->ntiles)) || (st = alloc(&c->d_state, 1)) ||
        (st = alloc(&c->d_rows, (size_t)(height + cd::kTile - 1) / cd::kTile)) ||
        (st = alloc(&c->d_io, 2 * (size_t)height * width)))
        return bail(st);
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_state), sizeof(cd::State), hipHostMallocDefault) != hipSuccess) {
        c->h_state = nullptr;
        (void)hipGetLastError();
        return bail(fail(P9IDG_ERR_ALLOC, "pinned host allocation failed"));
    }
    // psf2 = PSF.CalcPSFSquared(psf), MaxLipschitz = Residuals.GetMax(psf2)
    std::vector<float> psf2((size_t)c->p2h * c->p2w);
    if ((st = p9psf_squared(device, psf, rows, cols, psf2.data()))) return bail(st);
    float m = 0.f;
    for (float v : psf2) m = std::fmax(m, v);
    c->max_lipschitz = m;
    if (hipError_t e = hipMemcpy(c->d_psf2, psf2.data(), sizeof(float) * psf2.size(), hipMemcpyHostToDevice))
        return bail(hip_fail(e, "psf2 upload"));
    if ((st = amap_device(psf, rows, cols, height, width, c->d_amap, c->stream))) return bail(st);
    *out = c;
    return P9IDG_OK;
}

int p9cd_destroy(p9cd* c) {
    if (!c) return P9IDG_OK;
    Guard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    cd_drop_graph(c);
    if (c->h_state) (void)hipHostFree(c->h_state);
    release(c->d_amap);
    release(c->d_psf2);
    release(c->d_tiles);
    release(c->d_rows);
    release(c->d_state);
    release(c->d_io);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return P9IDG_OK;
}

int p9cd_max_lipschitz(const p9cd* c, float* out) {
    if (!c || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *out = c->max_lipschitz;
    return P9IDG_OK;
}

int p9cd_get_abs_max_diff_device(p9cd* c, const float* d_x, const float* d_g, float lambda, float alpha, float* out) {
    if (!c || !d_x || !d_g || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    return cd_abs_max(c, d_x, d_g, lambda, alpha, out);
}

int p9cd_get_abs_max_diff(p9cd* c, const float* x, const float* grad, float lambda, float alpha, float* out) {
    if (!c || !x || !grad || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    const size_t n = (size_t)c->H * c->W;
    P9_HIP(hipMemcpyAsync(c->d_io, x, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    P9_HIP(hipMemcpyAsync(c->d_io + n, grad, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return cd_abs_max(c, c->d_io, c->d_io + n, lambda, alpha, out);
}

int p9cd_deconvolve_device(p9cd* c, float* d_x, float* d_g, float lambda, float alpha, int32_t max_iterations,
                           float epsilon, int32_t* converged, int32_t* iterations) {
    if (!c || !d_x || !d_g) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    return cd_run(c, d_x, d_g, lambda, alpha, max_iterations, epsilon, converged, iterations);
}

int p9cd_deconvolve(p9cd* c, float* x, float* grad, float lambda, float alpha, int32_t max_iterations, float epsilon,
                    int32_t* converged, int32_t* iterations) {
    if (!c || !x || !grad) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    Guard g(c->device);
    const size_t n = (size_t)c->H * c->W;
    P9_HIP(hipMemcpyAsync(c->d_io, x, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    P9_HIP(hipMemcpyAsync(c->d_io + n, grad, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (int st = cd_run(c, c->d_io, c->d_io + n, lambda, alpha, max_iterations, epsilon, converged, iterations))
        return st;
    P9_HIP(hipMemcpyAsync(x, c->d_io, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    P9_HIP(hipMemcpyAsync(grad, c->d_io + n, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    P9_HIP(hipStreamSynchronize(c->stream));
    return P9IDG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// MajorCycle.ReconstructSerialCD (MajorCycle.cs:33-134)
namespace {

struct CycleBuffers {
    float* x = nullptr;
    float* grad = nullptr;
    float* max = nullptr;
    ~CycleBuffers() {
        release(x);
        release(grad);
        release(max);
    }
};

int device_max(const float* d, size_t n, float* d_out, float* out) {
    P9_HIP(cd::launch_get_max(d, n, d_out, nullptr));
    P9_HIP(hipMemcpy(out, d_out, sizeof(float), hipMemcpyDeviceToHost));
    return P9IDG_OK;
}

}  // namespace

extern "C" int p9idg_reconstruct_serial_cd(p9idg_plan* plan, const char* obs_name, const void* vis, int32_t vis_type,
                                           const uint8_t* flags,
                                           int32_t psf_cut_factor, int32_t max_major_cycles, float lambda, float alpha,
                                           int32_t deconv_iterations, float deconv_epsilon, float* x_image,
                                           p9cycle_stats* stats, int32_t* ncycles) {
    const float kMajorEpsilon = 1e-4f;  // MajorCycle.cs:19
    p9idg_constants c{};
    int32_t device = 0;
    if (int st = p9idg_get_constants(plan, &c, &device)) return st;
    if (!vis || !x_image) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (psf_cut_factor <= 0 || c.grid_size / psf_cut_factor <= 0)
        return fail(P9IDG_ERR_ARGUMENT, "PSF cut factor out of range");
    const int G = c.grid_size, pc = G / psf_cut_factor;
    const size_t GG = (size_t)G * G;
    if (ncycles) *ncycles = 0;
    Guard guard(device);

    // PSF (MajorCycle.cs:36-48): ones where not flagged, GridW, ToImage, Shift
    if (int st = p9idg_set_psf_vis(plan, flags)) return st;
    if (int st = p9idg_set_grid_source(plan, P9IDG_SRC_VIS)) return st;
    if (int st = p9idg_exec_grid(plan, P9IDG_GRID_WSTACK)) return st;
    if (int st = p9idg_exec_grid_to_image(plan, P9IDG_GRID_WSTACK)) return st;
    std::vector<float> psf(GG), psf_cut((size_t)pc * pc);
    if (int st = p9idg_download_image(plan, psf.data())) return st;
    if (int st = p9idg_upload_vis(plan, vis, vis_type)) return st;
    if (int st = p9idg_set_flags(plan, flags)) return st;
    p9psf_cut(psf.data(), G, psf_cut_factor, psf_cut.data());
    float max_sidelobe = 0.f;
    p9psf_max_sidelobe(psf.data(), G, psf_cut_factor, &max_sidelobe);

    p9cd* cd = nullptr;
    p9conv* conv = nullptr;
    struct Objects {
        p9cd*& cd;
        p9conv*& conv;
        ~Objects() {
            p9cd_destroy(cd);
            p9conv_destroy(conv);
        }
    } objects{cd, conv};
    if (int st = p9cd_create(device, G, G, psf_cut.data(), pc, pc, &cd)) return st;
    // gCalculator; the reference's second convolver (full PSF) never computes a gradient
    if (int st = p9conv_create_correlation(device, psf_cut.data(), pc, pc, G, G, &conv)) return st;
    double max_lipschitz = 0.0, lipschitz_full = 0.0;
    p9psf_max_lipschitz(psf_cut.data(), pc, pc, &max_lipschitz);
    p9psf_max_lipschitz(psf.data(), G, G, &lipschitz_full);
    const float lambda_lipschitz = (float)((double)lambda * max_lipschitz);
    const float lambda_true = (float)((double)lambda * lipschitz_full);
    bool switched = false;

    CycleBuffers b;
    if (int st = alloc(&b.x, GG)) return st;
    if (int st = alloc(&b.grad, GG)) return st;
    if (int st = alloc(&b.max, 1)) return st;
    P9_HIP(hipMemsetAsync(b.x, 0, GG * sizeof(float), nullptr));
    P9_HIP(hipStreamSynchronize(nullptr));
    void* img_v = nullptr;
    if (int st = p9idg_device_buffers(plan, nullptr, nullptr, nullptr, nullptr, &img_v)) return st;
    float* d_image = static_cast<float*>(img_v);

    // first dirty image from the visibilities
    if (int st = p9idg_exec_grid(plan, P9IDG_GRID_WSTACK)) return st;
    if (int st = p9idg_exec_grid_to_image(plan, P9IDG_GRID_WSTACK)) return st;
    if (int st = p9idg_sync(plan)) return st;
    using clock = std::chrono::steady_clock;
    auto ms = [](clock::time_point a, clock::time_point b) {
        return (float)std::chrono::duration<double, std::milli>(b - a).count();
    };
    // FitsIO.Write of the dirty and model images of every cycle (MajorCycle.cs:91,117)
    std::vector<float> host;
    auto write_image = [&](const float* d_img, const char* what, int32_t cycle) -> int {
        if (!obs_name) return P9IDG_OK;
        host.resize(GG);
        P9_HIP(hipMemcpy(host.data(), d_img, GG * sizeof(float), hipMemcpyDeviceToHost));
        const std::string path = std::string(obs_name) + what + std::to_string(cycle) + ".fits";
        return p9fits_write_image(path.c_str(), P9FITS_F32, host.data(), G, G);
    };
    bool captured = false;
    for (int32_t cycle = 0; cycle < max_major_cycles; cycle++) {
        if (int st = write_image(d_image, "_dirty_serial_majorCycle", cycle)) return st;
        const auto t0 = clock::now();
        float max_dirty = 0.f, max_b = 0.f, abs_max = 0.f;
        if (int st = device_max(d_image, GG, b.max, &max_dirty)) return st;
        if (int st = p9conv_convolve_device(conv, d_image, b.grad)) return st;
        if (int st = device_max(b.grad, GG, b.max, &max_b)) return st;
        const double correction = std::fmax((double)max_b / ((double)max_dirty * max_lipschitz), 1.0);
        const float prod = max_b * max_sidelobe;
        const double sidelobe = (double)prod * correction;
        const float current_lambda = (float)std::fmax(sidelobe / (double)alpha, (double)lambda_lipschitz);
        const auto t1 = clock::now();
        if (int st = p9cd_get_abs_max_diff_device(cd, b.x, b.grad, lambda_true, alpha, &abs_max)) return st;
        int32_t it = 0, cv = 0;
        if (abs_max >= kMajorEpsilon) {
            if (int st = p9cd_deconvolve_device(cd, b.x, b.grad, current_lambda, alpha, deconv_iterations, deconv_epsilon,
                                                &cv, &it))
                return st;
        }
        const auto t2 = clock::now();
        if (int st = write_image(b.x, "_model_serial_majorCycle", cycle)) return st;
        const auto t3 = clock::now();
        if (stats) {
            p9cycle_stats& s = stats[cycle];
            s = p9cycle_stats{};
            s.max_dirty = max_dirty;
            s.max_gradient = max_b;
            s.lambda = current_lambda;
            s.abs_max_diff = abs_max;
            s.iterations = it;
            s.converged = cv;
            s.gradient_ms = ms(t0, t1);
            s.deconv_ms = ms(t1, t2);
        }
        if (ncycles) *ncycles = cycle + 1;
        if (lambda == current_lambda && !switched) {
            lambda = lambda_true;
            max_lipschitz = lipschitz_full;
            switched = true;
        }
        if (abs_max < kMajorEpsilon || cycle + 1 == max_major_cycles) break;
        // ToUVW, DeGridW, residual (MajorCycle.cs:125-129) and the next dirty image
        // the plan's stream does not wait for the null stream, and a
        // device-to-device hipMemcpy may return before the copy is done:
        // finish it before ToUVW reads the image on the plan's stream
        P9_HIP(hipMemcpyAsync(d_image, b.x, GG * sizeof(float), hipMemcpyDeviceToDevice, nullptr));
        P9_HIP(hipStreamSynchronize(nullptr));
        // the device step is recorded once as a hipGraph (capture runs it
        // eagerly first) and replayed in every later cycle
        if (!captured) {
            if (int st = p9idg_capture_cycle(plan, P9IDG_GRID_WSTACK)) return st;
            captured = true;
        } else {
            if (int st = p9idg_launch_cycle(plan)) return st;
        }
        if (int st = p9idg_sync(plan)) return st;
        if (stats) stats[cycle].imaging_ms = ms(t3, clock::now());
    }
    P9_HIP(hipMemcpy(x_image, b.x, GG * sizeof(float), hipMemcpyDeviceToHost));
    return P9IDG_OK;
}
