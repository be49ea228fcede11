// p9idg_grid_noalloc: GridW / Grid of device-resident visibilities without a
// plan and without allocating device memory, for callers that cannot call
// hipMalloc.  The gridder kernels are the plan's (p9idg_exec_grid); their
// working memory -- the work-queue counter, the work items and schedule, the
// subgrid metadata, the wavenumbers and the twiddle / adder-phasor / w-term
// tables -- lives in one file-scope device array per GPU, `scratch`.
#include <cmath>
#include <mutex>

#include "p9idg_internal.h"

namespace p9idg {
namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;  // per GPU, held until the process ends
constexpr int kMaxSubgrids = 65536;
constexpr int kMaxChannels = 4096;
constexpr int kMaxSubgridSize = 64;
constexpr int kMaxDevices = 64;

// The device array.  In host code its name is only a stand-in: every use goes
// through the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Address of `scratch` on the current device (looked up once per device).
int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(P9IDG_ERR_UNSUPPORTED, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(P9IDG_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return P9IDG_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Byte offsets of the regions of `scratch` for one call.
struct Layout {
    size_t counter, twiddle, phasor, rtab, wavenumbers, meta, order, items, end;
};
Layout layout(int S, int C, int nsub) {
    Layout l{};
    size_t o = 0;
    l.counter = o;
    o = align256(o + 2 * sizeof(int32_t));
    l.twiddle = o;
    o = align256(o + sizeof(float2) * S);
    l.phasor = o;
    o = align256(o + sizeof(float2) * S * S);
    l.rtab = o;
    o = align256(o + sizeof(float) * S * S);
    l.wavenumbers = o;
    o = align256(o + sizeof(double) * C);
    l.meta = o;
    o = align256(o + sizeof(p9idg_subgrid) * (size_t)nsub);
    l.order = o;
    o = align256(o + sizeof(int32_t) * (size_t)nsub);
    l.items = o;
    o = align256(o + sizeof(WorkItem) * (size_t)nsub);
    l.end = o;
    return l;
}

// One work item per subgrid (all its visibilities), in subgrid order.
__global__ void build_items_kernel(const p9idg_subgrid* meta, int nsub, int C, int32_t* order, WorkItem* items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsub) {
        order[i] = i;
        items[i] = WorkItem{i, 0, meta[i].time_sample_count * C, 0, i};
    }
}

}  // namespace
}  // namespace p9idg

using namespace p9idg;

extern "C" int p9idg_grid_noalloc(const p9idg_constants* c, int32_t kernel, int32_t mode, const p9idg_subgrid* meta,
                                  int32_t nsub, const double* d_uvw, int32_t B, int32_t T, const double* freq,
                                  int32_t C, const void* d_vis, void* d_out, void* stream) {
    if (!c || !freq || !d_uvw || !d_vis || !d_out || (!meta && nsub > 0))
        return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    const int G = c->grid_size, S = c->subgrid_size, L = c->wlayer_count;
    if (G <= 0) return fail(P9IDG_ERR_ARGUMENT, "grid_size must be positive");
    if (!subgrid_size_supported(S)) return fail(P9IDG_ERR_UNSUPPORTED, "subgrid_size must be one of 8, 16, 24, 32, 48, 64");
    if (S >= G) return fail(P9IDG_ERR_ARGUMENT, "subgrid_size >= grid_size");
    if (L < 1) return fail(P9IDG_ERR_ARGUMENT, "wlayer_count must be >= 1");
    if (c->visibilities_count <= 0) return fail(P9IDG_ERR_ARGUMENT, "visibilities_count must be > 0");
    if (kernel != P9IDG_KERNEL_FP32 && kernel != P9IDG_KERNEL_MFMA_BF16 && kernel != P9IDG_KERNEL_MFMA_FP16 &&
        kernel != P9IDG_KERNEL_FP32_SEP)
        return fail(P9IDG_ERR_ARGUMENT, "unknown kernel");
    if (mode < P9IDG_GRID_PLAIN || mode > P9IDG_GRID_WSUM) return fail(P9IDG_ERR_ARGUMENT, "unknown grid mode");
    if (B <= 0 || T <= 0 || C <= 0 || nsub < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid dimensions");
    if (nsub > kMaxSubgrids)
        return fail(P9IDG_ERR_CAPACITY, "p9idg_grid_noalloc takes at most " + std::to_string(kMaxSubgrids) +
                                            " subgrids, got " + std::to_string(nsub));
    if (C > kMaxChannels)
        return fail(P9IDG_ERR_CAPACITY, "p9idg_grid_noalloc takes at most " + std::to_string(kMaxChannels) +
                                            " channels, got " + std::to_string(C));
    const Layout lay = layout(S, C, nsub);
    if (lay.end > kScratchBytes) return fail(P9IDG_ERR_CAPACITY, "scratch too small for this call");
    // the metadata checks of p9idg_set_observation, and Adder.cs:30-31 for Grid
    for (int32_t i = 0; i < nsub; i++) {
        const p9idg_subgrid& m = meta[i];
        if (m.baseline_idx < 0 || m.baseline_idx >= B || m.time_sample_start < 0 || m.time_sample_count < 0 ||
            (int64_t)m.time_sample_start + m.time_sample_count > T)
            return fail(P9IDG_ERR_ARGUMENT, "subgrid " + std::to_string(i) + " indexes outside the observation");
        if (m.w_lambda < -L || m.w_lambda >= L)
            return fail(P9IDG_ERR_ARGUMENT, "subgrid " + std::to_string(i) + " has a w index outside the layers");
        if (mode == P9IDG_GRID_PLAIN && !(m.u_pixel >= 0 && m.u_pixel < G - S && m.v_pixel >= 0 && m.v_pixel < G - S))
            return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    }
    int dev = 0;
    P9_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    if (int st = scratch_address(dev, &base)) return st;
    hipStream_t st = static_cast<hipStream_t>(stream);

    // host tables on the stack (as p9idg_create and p9idg_set_observation
    // compute them), copied into the scratch regions
    float2 tw[kMaxSubgridSize], ph[kMaxSubgridSize * kMaxSubgridSize];
    for (int m = 0; m < S; m++) {
        const double a = 2.0 * M_PI * (double)m / (double)S;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            const double a = M_PI * (double)(x + y - S) / (double)S;
            ph[y * S + x] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
    double k[kMaxChannels];
    for (int i = 0; i < C; i++) k[i] = ((2.0 * M_PI) * freq[i]) / 299792458.0;  // MathFunctions.cs:69-78
    bool uniform = C >= 2;
    if (uniform) {
        const double dk = (k[C - 1] - k[0]) / (C - 1), tol = 1e-12 * std::fabs(k[C - 1]);
        for (int i = 0; i < C; i++) uniform = uniform && std::fabs(k[i] - (k[0] + i * dk)) <= tol;
    }
    // pageable host sources: each copy has read its source when it returns
    P9_HIP(hipMemcpyAsync(base + lay.twiddle, tw, sizeof(float2) * S, hipMemcpyHostToDevice, st));
    P9_HIP(hipMemcpyAsync(base + lay.phasor, ph, sizeof(float2) * S * S, hipMemcpyHostToDevice, st));
    P9_HIP(hipMemcpyAsync(base + lay.wavenumbers, k, sizeof(double) * C, hipMemcpyHostToDevice, st));
    if (nsub > 0)
        P9_HIP(hipMemcpyAsync(base + lay.meta, meta, sizeof(p9idg_subgrid) * nsub, hipMemcpyHostToDevice, st));
    const float image_size = c->scale_arcsec * (float)G;  // GriddingConstants.cs:43
    P9_HIP(launch_rtab(reinterpret_cast<float*>(base + lay.rtab), S, image_size, st));
    if (nsub > 0) {
        hipLaunchKernelGGL(build_items_kernel, dim3((nsub + 255) / 256), dim3(256), 0, st,
                           reinterpret_cast<const p9idg_subgrid*>(base + lay.meta), nsub, C,
                           reinterpret_cast<int32_t*>(base + lay.order), reinterpret_cast<WorkItem*>(base + lay.items));
        P9_HIP(hipGetLastError());
    }
    const size_t n_out = (size_t)G * G * (mode == P9IDG_GRID_WSTACK ? (size_t)L : 1);
    P9_HIP(launch_fill_zero(static_cast<float2*>(d_out), n_out, st));

    KernelArgs a{};
    a.meta = reinterpret_cast<const p9idg_subgrid*>(base + lay.meta);
    a.order = reinterpret_cast<const int32_t*>(base + lay.order);
    a.nr_subgrids = nsub;
    a.uvw = d_uvw;
    a.wavenumbers = reinterpret_cast<const double*>(base + lay.wavenumbers);
    a.vis = static_cast<const float2*>(d_vis);
    a.twiddle = reinterpret_cast<const float2*>(base + lay.twiddle);
    a.adder_phasor = reinterpret_cast<const float2*>(base + lay.phasor);
    a.rtab = reinterpret_cast<const float*>(base + lay.rtab);
    a.grid = static_cast<float2*>(d_out);
    a.G = G;
    a.T = T;
    a.C = C;
    a.L = L;
    a.image_size = image_size;
    a.u_scale = 2 * M_PI / (double)image_size;             // Gridder.cs:38
    a.w_offset_scale = 2 * M_PI * (double)c->wstep_lambda;  // Gridder.cs:40-41
    a.degrid_scale = 1.0f / (float)(S * S);
    a.mode = mode;
    a.chan_rec = uniform && C % 8 == 0;
    a.chan_step = uniform;
    a.split_k[0] = 0x0000BF80u;
    a.split_k[1] = 0xBF800000u;
    a.split_k[2] = 0x0000BC00u;
    a.items = reinterpret_cast<const WorkItem*>(base + lay.items);
    a.nitems = nsub;
    a.counter = reinterpret_cast<int32_t*>(base + lay.counter);  // zeroed below: the array keeps the last call's
    // two persistent workgroups per CU (the residency of the S <= 32
    // kernels; larger subgrids leave the extra ones to find the queue empty),
    // from the device attribute rather than the launcher's occupancy cache,
    // which would allocate on its first use
    int cus = 0;
    P9_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LaunchCfg cfg;
    cfg.queue = true;
    cfg.workgroups = 2 * (cus > 0 ? cus : 1);
    P9_HIP(launch_reset_counter(a.counter, st));
    hipError_t e;
    if (kernel == P9IDG_KERNEL_FP32 || kernel == P9IDG_KERNEL_FP32_SEP)
        e = launch_grid_fp32(a, S, kernel == P9IDG_KERNEL_FP32_SEP, cfg, st);
    else
        e = launch_grid_mfma(a, S, kernel, cfg, st);
    if (e != hipSuccess) return hip_fail(e, "gridder launch");
    return P9IDG_OK;
}
