// C ABI of libp9idg: the plan (handle) that owns device state, and the
// entry points that mirror the reference's gridding API.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "p9idg_internal.h"

namespace p9idg {

const char* const kMsgPlace =
    "could not place (all) visibilities on subgrid (too many channnels, kernel size too large)";
const char* const kMsgSubgridRange = "subgrid falls not within grid";
const char* const kMsgInvalidGrid = "invalid grid";

namespace {
thread_local std::string t_error;
}
void set_error(const std::string& msg) { t_error = msg; }
const char* get_error() { return t_error.c_str(); }

int fail(int status, const std::string& msg) {
    set_error(msg);
    return status;
}
// A reported HIP failure is also cleared from the runtime's last-error
// state, so a later launch check does not report it again.
int hip_fail(hipError_t e, const char* what) {
    (void)hipGetLastError();
    return fail(P9IDG_ERR_HIP, std::string("HIP error ") + hipGetErrorName(e) + " (" +
                                   hipGetErrorString(e) + ") in " + what);
}
int fft_fail(hipfftResult r, const char* what) {
    return fail(P9IDG_ERR_HIP, std::string("hipFFT error ") + std::to_string((int)r) + " in " + what);
}

}  // namespace p9idg

using namespace p9idg;

struct p9idg_plan {
    p9idg_constants c{};
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t kernel = P9IDG_KERNEL_FP32;

    // observation
    bool have_obs = false;
    int32_t B = 0, T = 0, C = 0;
    int64_t nsub = 0;
    // element counts the observation-shaped buffers are allocated for (0: not
    // allocated); kept apart from B, T, C and nsub, which describe the last
    // observation set successfully
    size_t alloc_uvw = 0, alloc_wn = 0, alloc_sub = 0, alloc_vis = 0;
    uint64_t obs_gen = 0;  // bumped by every p9idg_set_observation past argument checks
    bool plain_adder_ok = true;  // every subgrid passes Adder.cs:30-31
    bool chan_uniform = false;   // wavenumbers equally spaced (to 1e-12 relative)
    int64_t gridded_vis = 0;

    // device buffers
    double* d_uvw = nullptr;
    double* d_wavenumbers = nullptr;
    p9idg_subgrid* d_meta = nullptr;
    int32_t* d_order = nullptr;
    float2* d_vis = nullptr;
    float2* d_model_vis = nullptr;
    float2* d_grid = nullptr;
    float2* d_layers = nullptr;
    float2* d_model_grid = nullptr;
    float* d_image = nullptr;
    double* d_image64 = nullptr;
    float2* d_twiddle = nullptr;
    float2* d_phasor = nullptr;
    float* d_rtab = nullptr;  // [S][S] w-term residual of the MFMA phasor factorisation
    uint8_t* d_flags = nullptr;
    void* d_staging = nullptr;
    size_t staging_bytes = 0;

    // work queues (one item list per direction) and the queue counter
    std::vector<p9idg_subgrid> meta_host;
    std::vector<int32_t> order_host;
    WorkItem* d_items_grid = nullptr;
    WorkItem* d_items_degrid = nullptr;
    int32_t nitems_grid = 0, nitems_degrid = 0;
    int32_t* d_counter = nullptr;
    // launch shape (tunable): queue mode, visibilities per work item, workgroups
    // residual visibilities (vis - model_vis, 0 where flagged) and the
    // gridder's input selector
    float2* d_resid = nullptr;
    bool have_flags = false;
    int32_t grid_src = P9IDG_SRC_VIS;
    // captured major-cycle step
    bool capturing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // Zero state of the gridder targets.  A buffer is "clean" when it is zero
    // outside the footprint of the last gridding into it (dirty_fp: which
    // footprint, -1 = none); otherwise (unknown contents) it is fully zeroed.
    // The footprints (per-layer bounding boxes of the subgrids, per mode)
    // live on the device.
    bool grid_clean = false, layers_clean = false, model_vis_clean = false;
    int32_t grid_dirty_fp = -1, layers_dirty_fp = -1;
    Box* d_fp[3] = {nullptr, nullptr, nullptr};  // per grid mode
    int32_t n_fp[3] = {0, 0, 0};
    Box fp_union[3] = {};                         // bounding box of each mode's boxes
    float2* d_lsum = nullptr;                     // layer sum for the w-stacked ToImage
    int32_t use_queue = 1;
    int32_t grid_chunk = 2048;
    int32_t degrid_chunk = 2048;
    int32_t workgroups = 0;         // gridder's persistent workgroups (0: CUs x occupancy)
    int32_t workgroups_degrid = 0;  // the degridder's (its workgroups are half the size)

    // deterministic adder: each gridder work item writes its subgrid to a
    // scratch slot, then gather_slots sums the slots per grid tile in slot order
    bool deterministic = false;
    float2* d_scratch = nullptr;
    size_t alloc_scratch = 0;  // float2 elements
    // Tile lists of each grid mode, kept side by side: building one mode's
    // lists never frees another's, which a captured graph may be using.
    // They stay valid until the work items change (upload_items).
    struct Gather {
        bool valid = false;
        GatherBin* d_bins = nullptr;
        GatherChunk* d_chunks = nullptr;
        int32_t* d_bin_slots = nullptr;
        SlotCorner* d_corners = nullptr;
        float2* d_partial = nullptr;  // one 32 x 32 tile per chunk
        int32_t n_bins = 0, n_chunks = 0;
    } gather[3];

    hipfftHandle fft_one = 0;
    bool have_fft_one = false;

    hipEvent_t ev[8] = {};
    // the event marking timing slot i for p9idg_last_timings: ev[i], or the
    // span event recorded at that position while spans are collected
    hipEvent_t mark[8] = {};
    hipEvent_t last_span_event = nullptr;  // the latest event a span recorded
    bool timed_grid = false, timed_degrid = false, timed_fft = false;

    // Event pairs of every gridder / degridder launch since the last reset
    // (p9idg_timing_reset), for averages over a timed region.
    struct Span {
        hipEvent_t start, stop;
        int kind;  // p9idg_stage
    };
    std::vector<hipEvent_t> event_pool;
    std::vector<Span> spans;
    bool collect = false;
};

namespace {

size_t padded(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

template <typename T>
int dev_alloc(T** p, size_t count) {
    if (*p) return P9IDG_OK;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, padded(count * sizeof(T)));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(P9IDG_ERR_ALLOC, std::string("device allocation of ") +
                                         std::to_string(count * sizeof(T)) +
                                         " bytes failed: " + hipGetErrorString(e));
    }
    *p = static_cast<T*>(q);
    return P9IDG_OK;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

int check_plan(const p9idg_plan* p) {
    if (!p) return fail(P9IDG_ERR_ARGUMENT, "plan is NULL");
    return P9IDG_OK;
}
int check_obs(const p9idg_plan* p) {
    if (int st = check_plan(p)) return st;
    if (!p->have_obs) return fail(P9IDG_ERR_STATE, "no observation set (p9idg_set_observation)");
    return P9IDG_OK;
}

size_t grid_elems(const p9idg_plan* p) { return (size_t)p->c.grid_size * p->c.grid_size; }
size_t vis_elems(const p9idg_plan* p) { return (size_t)p->B * p->T * p->C; }

int ensure_layers(p9idg_plan* p) {
    return dev_alloc(&p->d_layers, grid_elems(p) * (size_t)p->c.wlayer_count);
}

void mark_unknown(p9idg_plan* p) {
    p->grid_clean = p->layers_clean = p->model_vis_clean = false;
}

// Per-mode footprints: per-layer bounding boxes of where the adder writes
// (Adder.cs:30-42, AdderWStack.cs:29-70).
int build_footprints(p9idg_plan* p, const p9idg_subgrid* meta, int64_t nsub) {
    const int S = p->c.subgrid_size, G = p->c.grid_size, L = p->c.wlayer_count;
    for (int mode = 0; mode < 3; mode++) {
        const int nl = mode == P9IDG_GRID_WSTACK ? L : 1;
        std::vector<Box> box(nl, Box{0, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN});
        for (int l = 0; l < nl; l++) box[l].layer = l;
        for (int64_t i = 0; i < nsub; i++) {
            int X = meta[i].u_pixel, Y = meta[i].v_pixel, W = meta[i].w_lambda, layer = 0;
            if (mode == P9IDG_GRID_PLAIN) {
                if (!(X >= 0 && X < G - S && Y >= 0 && Y < G - S)) continue;
            } else {
                if (W < 0) {
                    X = G - X - S + 1;
                    Y = G - Y - S + 1;
                    W = -W - 1;
                }
                if (!(X >= 1 && X < G - S && Y >= 1 && Y < G - S)) continue;
                if (mode == P9IDG_GRID_WSTACK) layer = W;
            }
            if (layer < 0 || layer >= nl) continue;
            Box& b = box[layer];
            b.y0 = std::min(b.y0, Y);
            b.y1 = std::max(b.y1, Y + S);
            b.x0 = std::min(b.x0, X);
            b.x1 = std::max(b.x1, X + S);
        }
        std::vector<Box> used;
        Box u{0, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
        for (const Box& b : box)
            if (b.y1 > b.y0) {
                used.push_back(b);
                u.y0 = std::min(u.y0, b.y0);
                u.y1 = std::max(u.y1, b.y1);
                u.x0 = std::min(u.x0, b.x0);
                u.x1 = std::max(u.x1, b.x1);
            }
        p->fp_union[mode] = u;
        dev_free(p->d_fp[mode]);
        p->n_fp[mode] = (int32_t)used.size();
        if (used.empty()) continue;
        if (int st = dev_alloc(&p->d_fp[mode], used.size())) return st;
        P9_HIP(hipMemcpy(p->d_fp[mode], used.data(), used.size() * sizeof(Box), hipMemcpyHostToDevice));
    }
    return P9IDG_OK;
}

// Brings the gridder target of `mode` to all-zero: the previous footprint
// when the contents are known, the whole buffer otherwise (and always while
// capturing, since a graph may be replayed after other work on the buffer).
int zero_target(p9idg_plan* p, int mode, float2* target, size_t n) {
    bool& clean = mode == P9IDG_GRID_WSTACK ? p->layers_clean : p->grid_clean;
    int32_t& fp = mode == P9IDG_GRID_WSTACK ? p->layers_dirty_fp : p->grid_dirty_fp;
    if (p->capturing || !clean) {
        P9_HIP(launch_fill_zero(target, n, p->stream));
    } else if (fp >= 0) {
        P9_HIP(launch_zero_boxes(target, p->d_fp[fp], p->n_fp[fp], p->c.grid_size, p->stream));
    }
    clean = !p->capturing;
    fp = mode;
    return P9IDG_OK;
}

int ensure_fft_one(p9idg_plan* p) {
    if (p->have_fft_one) return P9IDG_OK;
    P9_FFT(hipfftPlan2d(&p->fft_one, p->c.grid_size, p->c.grid_size, HIPFFT_C2C));
    P9_FFT(hipfftSetStream(p->fft_one, p->stream));
    p->have_fft_one = true;
    return P9IDG_OK;
}

int ensure_staging(p9idg_plan* p, size_t bytes) {
    if (p->staging_bytes >= bytes) return P9IDG_OK;
    if (p->d_staging) (void)hipFree(p->d_staging);
    p->d_staging = nullptr;
    p->staging_bytes = 0;
    hipError_t e = hipMalloc(&p->d_staging, padded(bytes));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(P9IDG_ERR_ALLOC, std::string("staging allocation failed: ") + hipGetErrorString(e));
    }
    p->staging_bytes = bytes;
    return P9IDG_OK;
}

// complex128 host array -> complex64 device buffer
int upload_c128(p9idg_plan* p, const double* src, float2* dst, size_t n) {
    if (int st = ensure_staging(p, n * 16)) return st;
    P9_HIP(hipMemcpyAsync(p->d_staging, src, n * 16, hipMemcpyHostToDevice, p->stream));
    P9_HIP(launch_c128_to_c64(static_cast<const double2*>(p->d_staging), dst, n, p->stream));
    return P9IDG_OK;
}

// complex64 device buffer -> complex128 host array
int download_c128(p9idg_plan* p, const float2* src, double* dst, size_t n) {
    std::vector<float2> tmp(n);
    P9_HIP(hipMemcpyAsync(tmp.data(), src, n * sizeof(float2), hipMemcpyDeviceToHost, p->stream));
    P9_HIP(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < n; i++) {
        dst[2 * i] = (double)tmp[i].x;
        dst[2 * i + 1] = (double)tmp[i].y;
    }
    return P9IDG_OK;
}

hipEvent_t pooled_event(p9idg_plan* p, size_t i) {
    while (p->event_pool.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        p->event_pool.push_back(e);
    }
    return p->event_pool[i];
}

// Timing slot i of p9idg_last_timings at the current stream position (not
// while capturing).  It must directly follow a span call when spans are
// collected: each event on the stream delays the next launch, so a position
// records one event, which the span and the slot share.
int record(p9idg_plan* p, int i) {
    if (p->capturing) return P9IDG_OK;
    if (p->collect && p->last_span_event) {
        p->mark[i] = p->last_span_event;
        return P9IDG_OK;
    }
    P9_HIP(hipEventRecord(p->ev[i], p->stream));
    p->mark[i] = p->ev[i];
    return P9IDG_OK;
}

// Starts a timed span on the plan's stream when collection is on.
int span_begin(p9idg_plan* p, int kind) {
    if (!p->collect || p->capturing) return P9IDG_OK;
    const size_t i = p->spans.size() * 2;
    hipEvent_t a = pooled_event(p, i), b = pooled_event(p, i + 1);
    if (!a || !b) return fail(P9IDG_ERR_HIP, "event creation failed");
    P9_HIP(hipEventRecord(a, p->stream));
    p->spans.push_back({a, b, kind});
    p->last_span_event = a;
    return P9IDG_OK;
}
int span_end(p9idg_plan* p) {
    if (!p->collect || p->capturing || p->spans.empty()) return P9IDG_OK;
    P9_HIP(hipEventRecord(p->spans.back().stop, p->stream));
    p->last_span_event = p->spans.back().stop;
    return P9IDG_OK;
}
// Ends the open span and starts the next one at the same position, with one
// event for both.
int span_next(p9idg_plan* p, int kind) {
    if (!p->collect || p->capturing) return P9IDG_OK;
    if (p->spans.empty()) return span_begin(p, kind);
    hipEvent_t b = pooled_event(p, p->spans.size() * 2 + 1);
    if (!b) return fail(P9IDG_ERR_HIP, "event creation failed");
    hipEvent_t e = p->spans.back().stop;
    P9_HIP(hipEventRecord(e, p->stream));
    p->spans.push_back({e, b, kind});
    p->last_span_event = e;
    return P9IDG_OK;
}

// Splits every subgrid's visibilities into items of at most `chunk` (rounded
// to multiples of 32; one item per subgrid for chunk 0) and orders them
// longest first.  Slots number the items in (subgrid, k0) order, so with
// chunk 0 an item's slot is its subgrid.
std::vector<WorkItem> make_items(const std::vector<p9idg_subgrid>& meta,
                                 const std::vector<int32_t>& order, int C, int chunk) {
    std::vector<std::vector<WorkItem>> per(meta.size());
    int32_t slot = 0;
    for (int32_t s = 0; s < (int32_t)meta.size(); s++) {
        const int K = meta[s].time_sample_count * C;
        if (K <= 0) {
            per[s].push_back({s, 0, 0, 0, slot++});
            continue;
        }
        const int n = chunk > 0 ? (K + chunk - 1) / chunk : 1;
        int k0 = 0;
        for (int i = 1; i <= n; i++) {
            int k1 = i == n ? K : (int)(((int64_t)K * i / n + 31) / 32 * 32);
            if (k1 > K) k1 = K;
            if (k1 > k0) per[s].push_back({s, k0, k1, 0, slot++});
            k0 = k1;
        }
    }
    std::vector<WorkItem> items;
    items.reserve(slot);
    for (int32_t s : order)
        for (const WorkItem& w : per[s]) items.push_back(w);
    std::stable_sort(items.begin(), items.end(),
                     [](const WorkItem& a, const WorkItem& b) { return a.k1 - a.k0 > b.k1 - b.k0; });
    return items;
}

void drop_graph(p9idg_plan* p);

void free_gather(p9idg_plan::Gather& g) {
    dev_free(g.d_bins);
    dev_free(g.d_chunks);
    dev_free(g.d_bin_slots);
    dev_free(g.d_corners);
    dev_free(g.d_partial);
    g.n_bins = g.n_chunks = 0;
    g.valid = false;
}

int upload_items(p9idg_plan* p) {
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));  // queued launches may still read the old arrays
    drop_graph(p);  // a captured step holds the old item arrays and tile lists
    for (auto& g : p->gather) free_gather(g);
    dev_free(p->d_items_grid);
    dev_free(p->d_items_degrid);
    p->nitems_grid = p->nitems_degrid = 0;
    if (p->nsub == 0) return P9IDG_OK;
    const auto gi = make_items(p->meta_host, p->order_host, p->C, p->grid_chunk);
    const auto di = make_items(p->meta_host, p->order_host, p->C, p->degrid_chunk);
    if (gi.size() > (size_t)INT32_MAX || di.size() > (size_t)INT32_MAX)
        return fail(P9IDG_ERR_ARGUMENT, "too many work items");
    if (int st = dev_alloc(&p->d_items_grid, gi.size())) return st;
    if (int st = dev_alloc(&p->d_items_degrid, di.size())) return st;
    P9_HIP(hipMemcpyAsync(p->d_items_grid, gi.data(), gi.size() * sizeof(WorkItem), hipMemcpyHostToDevice, p->stream));
    P9_HIP(hipMemcpyAsync(p->d_items_degrid, di.data(), di.size() * sizeof(WorkItem), hipMemcpyHostToDevice, p->stream));
    P9_HIP(hipStreamSynchronize(p->stream));
    p->nitems_grid = (int32_t)gi.size();
    p->nitems_degrid = (int32_t)di.size();
    return P9IDG_OK;
}

LaunchCfg launch_cfg(const p9idg_plan* p, bool grid) {
    LaunchCfg cfg;
    cfg.queue = p->use_queue != 0;
    cfg.workgroups = grid ? p->workgroups : p->workgroups_degrid;
    return cfg;
}

KernelArgs make_args(const p9idg_plan* p, int mode) {
    KernelArgs a{};
    a.meta = p->d_meta;
    a.order = p->d_order;
    a.nr_subgrids = (int32_t)p->nsub;
    a.uvw = p->d_uvw;
    a.wavenumbers = p->d_wavenumbers;
    a.vis = p->grid_src == P9IDG_SRC_RESIDUAL ? p->d_resid : p->d_vis;
    a.vis_out = p->d_model_vis;
    a.twiddle = p->d_twiddle;
    a.adder_phasor = p->d_phasor;
    a.rtab = p->d_rtab;
    a.model_grid = p->d_model_grid;
    a.G = p->c.grid_size;
    a.T = p->T;
    a.C = p->C;
    a.L = p->c.wlayer_count;
    a.image_size = p->c.scale_arcsec * (float)p->c.grid_size;  // GriddingConstants.cs:43
    a.u_scale = 2 * M_PI / (double)a.image_size;                // Gridder.cs:38
    a.w_offset_scale = 2 * M_PI * (double)p->c.wstep_lambda;    // Gridder.cs:40-41
    const int S = p->c.subgrid_size;
    a.degrid_scale = 1.0f / (float)(S * S);                      // Gridder.cs:151
    a.mode = mode;
    a.chan_rec = p->chan_uniform && p->C % 8 == 0;
    a.chan_step = p->chan_uniform;
    a.split_k[0] = 0x0000BF80u;
    a.split_k[1] = 0xBF800000u;
    a.split_k[2] = 0x0000BC00u;
    a.counter = p->d_counter;
    return a;
}

// Scratch and tile bins of the deterministic adder for `mode`: every gridder
// work item (the queue's items, or one per subgrid without the queue) owns
// one S x S scratch slot; each 32 x 32 tile of the target lists, in slot
// order, the slots whose box (the adder's corner after the negative-w
// mirror; the adder's skips excluded) overlaps it.  Each mode keeps its own
// lists and the scratch depends on the work items only, so preparing one
// mode frees nothing that a graph captured in another mode uses.
int prepare_gather(p9idg_plan* p, int mode) {
    p9idg_plan::Gather& gt = p->gather[mode];
    if (gt.valid) return P9IDG_OK;  // valid until the work items change (upload_items)
    const int S = p->c.subgrid_size, G = p->c.grid_size, L = p->c.wlayer_count;
    const std::vector<WorkItem> items =
        make_items(p->meta_host, p->order_host, p->C, p->use_queue ? p->grid_chunk : 0);
    const size_t need = items.size() * (size_t)S * S;
    P9_HIP(hipStreamSynchronize(p->stream));  // queued launches may still use the old buffers
    if (p->alloc_scratch != need) {
        // only a change of the work items changes the scratch size, and that
        // already dropped the graph; drop it here too so no graph can keep
        // the freed scratch
        drop_graph(p);
        dev_free(p->d_scratch);
        p->alloc_scratch = 0;
        if (int st = dev_alloc(&p->d_scratch, std::max<size_t>(need, 1))) return st;
        p->alloc_scratch = need;
    }
    const int nt = (G + kGatherTile - 1) / kGatherTile, nl = mode == P9IDG_GRID_WSTACK ? L : 1;
    std::vector<SlotCorner> corner(items.size());
    std::vector<int32_t> layer(items.size(), -1);  // -1: the adder skips the slot
    for (const WorkItem& w : items) {
        const p9idg_subgrid& m = p->meta_host[w.sidx];
        int X = m.u_pixel, Y = m.v_pixel, W = 0;
        bool ok;
        if (mode == P9IDG_GRID_PLAIN) {
            ok = X >= 0 && X < G - S && Y >= 0 && Y < G - S;
        } else {
            W = m.w_lambda;
            if (W < 0) {
                X = G - X - S + 1;
                Y = G - Y - S + 1;
                W = -W - 1;
            }
            ok = X >= 1 && X < G - S && Y >= 1 && Y < G - S && W >= 0 && W < L;
            if (mode == P9IDG_GRID_WSUM) W = 0;
        }
        corner[w.slot] = {Y, X};
        if (ok) layer[w.slot] = W;
    }
    // counting sort of (tile, slot) pairs by tile; slots ascend within a tile
    const size_t nkeys = (size_t)nl * nt * nt;
    std::vector<int32_t> start(nkeys + 1, 0);
    auto for_tiles = [&](int32_t s, auto&& f) {
        const SlotCorner c = corner[s];
        for (int ty = c.y0 / kGatherTile; ty <= (c.y0 + S - 1) / kGatherTile; ty++)
            for (int tx = c.x0 / kGatherTile; tx <= (c.x0 + S - 1) / kGatherTile; tx++)
                f(((size_t)layer[s] * nt + ty) * nt + tx);
    };
    for (int32_t s = 0; s < (int32_t)items.size(); s++)
        if (layer[s] >= 0) for_tiles(s, [&](size_t k) { start[k + 1]++; });
    for (size_t k = 0; k < nkeys; k++) start[k + 1] += start[k];
    std::vector<int32_t> slots(start[nkeys]);
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int32_t s = 0; s < (int32_t)items.size(); s++)
        if (layer[s] >= 0) for_tiles(s, [&](size_t k) { slots[fill[k]++] = s; });
    // each tile's slot list cut into chunks of at most kGatherChunk slots
    std::vector<GatherBin> bins;
    std::vector<GatherChunk> chunks;
    for (size_t k = 0; k < nkeys; k++) {
        if (start[k + 1] == start[k]) continue;
        const int32_t b = (int32_t)bins.size();
        bins.push_back({(int32_t)(k / ((size_t)nt * nt)), (int32_t)(k / nt % nt), (int32_t)(k % nt),
                        (int32_t)chunks.size(), 0});
        for (int32_t i = start[k]; i < start[k + 1]; i += kGatherChunk)
            chunks.push_back({b, i, std::min(i + kGatherChunk, start[k + 1])});
        bins.back().nchunk = (int32_t)chunks.size() - bins.back().chunk0;
    }
    free_gather(gt);  // not valid here, so no graph uses these
    if (int st = dev_alloc(&gt.d_bins, std::max<size_t>(bins.size(), 1))) return st;
    if (int st = dev_alloc(&gt.d_chunks, std::max<size_t>(chunks.size(), 1))) return st;
    if (int st = dev_alloc(&gt.d_bin_slots, std::max<size_t>(slots.size(), 1))) return st;
    if (int st = dev_alloc(&gt.d_corners, std::max<size_t>(corner.size(), 1))) return st;
    if (int st = dev_alloc(&gt.d_partial, std::max<size_t>(chunks.size(), 1) * kGatherTile * kGatherTile)) return st;
    if (!bins.empty())
        P9_HIP(hipMemcpy(gt.d_bins, bins.data(), bins.size() * sizeof(GatherBin), hipMemcpyHostToDevice));
    if (!chunks.empty())
        P9_HIP(hipMemcpy(gt.d_chunks, chunks.data(), chunks.size() * sizeof(GatherChunk), hipMemcpyHostToDevice));
    if (!slots.empty())
        P9_HIP(hipMemcpy(gt.d_bin_slots, slots.data(), slots.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!corner.empty())
        P9_HIP(hipMemcpy(gt.d_corners, corner.data(), corner.size() * sizeof(SlotCorner), hipMemcpyHostToDevice));
    gt.n_bins = (int32_t)bins.size();
    gt.n_chunks = (int32_t)chunks.size();
    gt.valid = true;
    return P9IDG_OK;
}

int do_exec_grid(p9idg_plan* p, int mode) {
    if (int st = check_obs(p)) return st;
    if (mode < P9IDG_GRID_PLAIN || mode > P9IDG_GRID_WSUM)
        return fail(P9IDG_ERR_ARGUMENT, "unknown grid mode");
    if (mode == P9IDG_GRID_PLAIN && !p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    if (p->grid_src == P9IDG_SRC_RESIDUAL && !p->d_resid)
        return fail(P9IDG_ERR_STATE, "no residual visibilities (p9idg_exec_residual)");
    DeviceGuard dg(p->device);
    float2* target = p->d_grid;
    size_t n = grid_elems(p);
    if (mode == P9IDG_GRID_WSTACK) {
        if (int st = ensure_layers(p)) return st;
        target = p->d_layers;
        n *= (size_t)p->c.wlayer_count;
    }
    KernelArgs a = make_args(p, mode);
    a.grid = target;
    a.items = p->d_items_grid;
    a.nitems = p->nitems_grid;
    if (p->deterministic) {
        if (p->capturing && !p->gather[mode].valid)
            return fail(P9IDG_ERR_STATE, "deterministic adder not prepared for this mode before capture");
        if (int st = prepare_gather(p, mode)) return st;
        a.scratch = p->d_scratch;
    }
    if (int st = span_begin(p, P9IDG_STAGE_ZERO)) return st;
    if (int st = zero_target(p, mode, target, n)) return st;
    if (int st = span_next(p, P9IDG_STAGE_GRIDDER)) return st;
    if (int st_ = record(p, 1)) return st_;
    hipError_t e;
    if (p->kernel == P9IDG_KERNEL_FP32 || p->kernel == P9IDG_KERNEL_FP32_SEP)
        e = launch_grid_fp32(a, p->c.subgrid_size, p->kernel == P9IDG_KERNEL_FP32_SEP, launch_cfg(p, true),
                             p->stream);
    else
        e = launch_grid_mfma(a, p->c.subgrid_size, p->kernel, launch_cfg(p, true), p->stream);
    if (e != hipSuccess) return hip_fail(e, "gridder launch");
    if (p->deterministic) {
        const p9idg_plan::Gather& gt = p->gather[mode];
        e = launch_gather_slots(gt.d_bins, gt.n_bins, gt.d_chunks, gt.n_chunks, gt.d_bin_slots, gt.d_corners,
                                p->d_scratch, p->c.subgrid_size, p->c.grid_size, gt.d_partial, target, p->stream);
        if (e != hipSuccess) return hip_fail(e, "ordered subgrid sum");
    }
    if (int st = span_end(p)) return st;
    if (int st_ = record(p, 2)) return st_;
    p->timed_grid = true;
    return P9IDG_OK;
}

int do_exec_degrid(p9idg_plan* p, int mode) {
    if (int st = check_obs(p)) return st;
    if (mode != P9IDG_DEGRID_PLAIN && mode != P9IDG_DEGRID_WSTACK)
        return fail(P9IDG_ERR_ARGUMENT, "unknown degrid mode");
    DeviceGuard dg(p->device);
    KernelArgs a = make_args(p, mode);
    a.items = p->d_items_degrid;
    a.nitems = p->nitems_degrid;
    // Gridder.cs:96: uncovered timesteps stay 0; only the degridder writes
    // this buffer, and always the same (covered) entries
    if (p->capturing || !p->model_vis_clean) {
        if (int st = span_begin(p, P9IDG_STAGE_ZERO)) return st;
        P9_HIP(launch_fill_zero(p->d_model_vis, vis_elems(p), p->stream));
        if (int st = span_next(p, P9IDG_STAGE_DEGRIDDER)) return st;
    } else {
        if (int st = span_begin(p, P9IDG_STAGE_DEGRIDDER)) return st;
    }
    p->model_vis_clean = !p->capturing;
    if (int st_ = record(p, 3)) return st_;
    hipError_t e;
    if (p->kernel == P9IDG_KERNEL_FP32 || p->kernel == P9IDG_KERNEL_FP32_SEP)
        e = launch_degrid_fp32(a, p->c.subgrid_size, p->kernel == P9IDG_KERNEL_FP32_SEP, launch_cfg(p, false),
                               p->stream);
    else
        e = launch_degrid_mfma(a, p->c.subgrid_size, p->kernel, launch_cfg(p, false), p->stream);
    if (e != hipSuccess) return hip_fail(e, "degridder launch");
    if (int st = span_end(p)) return st;
    if (int st_ = record(p, 4)) return st_;
    p->timed_degrid = true;
    return P9IDG_OK;
}

// FFT.WStackIFFTFloat (FFT.cs:188-215): sum over layers of
// (float)(Re(IFFT(layer)) / N).  The inverse FFT is linear and applies no
// per-layer w-screen, so it runs once on the sum of the layers (identical up
// to the per-layer float rounding).  The layers are left as they are; when
// their footprint is known only the union of the footprint boxes is summed.
int layers_image(p9idg_plan* p, double inv_n, int shift) {
    const int G = p->c.grid_size;
    const int W = P9IDG_GRID_WSTACK;
    if (int st = dev_alloc(&p->d_lsum, grid_elems(p))) return st;
    if (int st = ensure_fft_one(p)) return st;
    if (p->layers_clean && p->layers_dirty_fp == W && !p->capturing) {
        P9_HIP(launch_fill_zero(p->d_lsum, grid_elems(p), p->stream));
        if (p->n_fp[W] > 0)
            P9_HIP(launch_sum_layer_boxes(p->d_layers, p->d_fp[W], p->n_fp[W], G, p->fp_union[W], p->d_lsum,
                                          p->stream));
    } else {
        P9_HIP(launch_sum_layers(p->d_layers, p->c.wlayer_count, grid_elems(p), p->d_lsum, p->stream));
    }
    P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_lsum),
                         reinterpret_cast<hipfftComplex*>(p->d_lsum), HIPFFT_BACKWARD));
    P9_HIP(launch_layers_to_image(p->d_lsum, 1, G, inv_n, shift, p->d_image, p->stream));
    return P9IDG_OK;
}

// Inverse FFT of the grid buffer(s) followed by the image extraction.
int do_grid_to_image(p9idg_plan* p, int mode) {
    if (int st = check_obs(p)) return st;
    DeviceGuard dg(p->device);
    const int G = p->c.grid_size;
    const double inv_n = 1.0 / (double)p->c.visibilities_count;
    if (int st = span_begin(p, P9IDG_STAGE_TO_IMAGE)) return st;
    if (int st_ = record(p, 5)) return st_;
    if (mode == P9IDG_GRID_WSTACK) {
        if (!p->d_layers) return fail(P9IDG_ERR_STATE, "no layers gridded");
        if (int st = layers_image(p, inv_n, 1)) return st;
    } else {
        p->grid_clean = false;
        if (int st = ensure_fft_one(p)) return st;
        P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_grid),
                             reinterpret_cast<hipfftComplex*>(p->d_grid), HIPFFT_BACKWARD));
        P9_HIP(launch_layers_to_image(p->d_grid, 1, G, inv_n, 1, p->d_image, p->stream));
    }
    if (int st = span_end(p)) return st;
    if (int st_ = record(p, 6)) return st_;
    p->timed_fft = true;
    return P9IDG_OK;
}

int do_image_to_grid(p9idg_plan* p) {
    if (int st = check_obs(p)) return st;
    DeviceGuard dg(p->device);
    if (int st = ensure_fft_one(p)) return st;
    if (int st = span_begin(p, P9IDG_STAGE_TO_UVW)) return st;
    P9_HIP(launch_image_to_grid(p->d_image, p->c.grid_size, 1, p->d_model_grid, p->stream));
    P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_model_grid),
                         reinterpret_cast<hipfftComplex*>(p->d_model_grid), HIPFFT_FORWARD));
    return span_end(p);
}

int do_residual(p9idg_plan* p) {
    if (int st = check_obs(p)) return st;
    DeviceGuard dg(p->device);
    if (int st = dev_alloc(&p->d_resid, vis_elems(p))) return st;
    P9_HIP(launch_residual(p->d_vis, p->d_model_vis, p->have_flags ? p->d_flags : nullptr, p->d_resid,
                           vis_elems(p), p->stream));
    return P9IDG_OK;
}

// One device step of the major cycle (MajorCycle.cs:125-129 then :88-90):
// ToUVW of the model image, DeGridW, residuals, grid the residuals, image.
int cycle_step(p9idg_plan* p, int grid_mode) {
    if (int st = do_image_to_grid(p)) return st;
    if (int st = do_exec_degrid(p, P9IDG_DEGRID_WSTACK)) return st;
    if (int st = do_residual(p)) return st;
    const int32_t src = p->grid_src;
    p->grid_src = P9IDG_SRC_RESIDUAL;
    int st = do_exec_grid(p, grid_mode);
    p->grid_src = src;
    if (st) return st;
    return do_grid_to_image(p, grid_mode);
}

void drop_graph(p9idg_plan* p) {
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

int upload_vis_any(p9idg_plan* p, const void* vis, int32_t type) {
    if (int st = check_obs(p)) return st;
    if (!vis) return fail(P9IDG_ERR_ARGUMENT, "visibilities are NULL");
    DeviceGuard dg(p->device);
    const size_t n = vis_elems(p);
    if (type == P9IDG_C64) {
        P9_HIP(hipMemcpyAsync(p->d_vis, vis, n * sizeof(float2), hipMemcpyHostToDevice, p->stream));
    } else if (type == P9IDG_C128) {
        if (int st = upload_c128(p, static_cast<const double*>(vis), p->d_vis, n)) return st;
    } else {
        return fail(P9IDG_ERR_ARGUMENT, "unknown visibility type");
    }
    return P9IDG_OK;
}

// The residual kernel of a captured cycle graph holds the flag pointer it was
// recorded with (NULL without flags); switching the flags on or off, or to a
// new buffer, therefore drops the graph.  New contents in the same buffer are
// read by the next replay.
int flags_changed(p9idg_plan* p, const uint8_t* before) {
    const uint8_t* after = p->have_flags ? p->d_flags : nullptr;
    if (before != after && (p->graph || p->graph_exec)) {
        P9_HIP(hipStreamSynchronize(p->stream));
        drop_graph(p);
    }
    return P9IDG_OK;
}

int set_psf_vis(p9idg_plan* p, const uint8_t* flags) {
    if (int st = check_obs(p)) return st;
    DeviceGuard dg(p->device);
    const size_t n = vis_elems(p);
    const uint8_t* dflags = nullptr;
    if (flags) {
        const uint8_t* before = p->have_flags ? p->d_flags : nullptr;
        if (int st = dev_alloc(&p->d_flags, n)) return st;
        P9_HIP(hipMemcpyAsync(p->d_flags, flags, n, hipMemcpyHostToDevice, p->stream));
        dflags = p->d_flags;
        p->have_flags = true;
        if (int st = flags_changed(p, before)) return st;
    }
    P9_HIP(launch_psf_vis(dflags, p->d_vis, n, p->stream));
    return P9IDG_OK;
}

int upload_grid_c128(p9idg_plan* p, const double* grid) {
    if (!grid) return fail(P9IDG_ERR_ARGUMENT, "grid is NULL");
    return upload_c128(p, grid, p->d_model_grid, grid_elems(p));
}

int sync(p9idg_plan* p) {
    P9_HIP(hipStreamSynchronize(p->stream));
    return P9IDG_OK;
}

template <typename T>
void shift_host(T* a, int32_t n) {
    const int n2 = n / 2;
    for (int i = 0; i < n2; i++)
        for (int j = 0; j < n2; j++) {
            std::swap(a[(size_t)i * n + j], a[(size_t)(i + n2) * n + j + n2]);
            std::swap(a[(size_t)(i + n2) * n + j], a[(size_t)i * n + j + n2]);
        }
}
struct c128 {
    double re, im;
};

}  // namespace

extern "C" {

int p9idg_api_version(void) { return P9IDG_API_VERSION; }
const char* p9idg_last_error(void) { return get_error(); }

int p9idg_partition(const p9idg_constants* c, const double* uvw, int32_t nr_baselines,
                    int32_t nr_timesteps, const double* freq, int32_t nr_channels,
                    p9idg_subgrid* out, int64_t capacity, int64_t* nr_subgrids,
                    int64_t* baseline_offsets) {
    if (!c || !freq || (!uvw && nr_baselines > 0 && nr_timesteps > 0))
        return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (nr_baselines < 0 || nr_timesteps < 0 || nr_channels < 1)
        return fail(P9IDG_ERR_ARGUMENT, "invalid dimensions");
    std::vector<p9idg_subgrid> subs;
    std::vector<int64_t> offs;
    int st = partition_host(*c, uvw, nr_baselines, nr_timesteps, freq, nr_channels, subs, offs);
    if (nr_subgrids) *nr_subgrids = (int64_t)subs.size();
    if (st == P9IDG_ERR_PLACE) return fail(st, kMsgPlace);
    if (st == P9IDG_ERR_SUBGRID_RANGE) return fail(st, kMsgSubgridRange);
    if (st) return fail(st, "partition failed");
    // (memcpy's source must not be NULL even for 0 bytes: empty vectors are skipped)
    if (baseline_offsets && !offs.empty())
        std::memcpy(baseline_offsets, offs.data(), offs.size() * sizeof(int64_t));
    if (out) {
        if ((int64_t)subs.size() > capacity)
            return fail(P9IDG_ERR_CAPACITY, "output capacity " + std::to_string(capacity) +
                                                " < " + std::to_string(subs.size()) + " subgrids");
        if (!subs.empty()) std::memcpy(out, subs.data(), subs.size() * sizeof(p9idg_subgrid));
    }
    return P9IDG_OK;
}

int p9idg_create(const p9idg_constants* c, int32_t device, p9idg_plan** plan) {
    if (!c || !plan) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *plan = nullptr;
    if (c->grid_size <= 0) return fail(P9IDG_ERR_ARGUMENT, "grid_size must be positive");
    if (!subgrid_size_supported(c->subgrid_size))
        return fail(P9IDG_ERR_UNSUPPORTED, "subgrid_size must be one of 8, 16, 24, 32, 48, 64");
    if (c->subgrid_size >= c->grid_size) return fail(P9IDG_ERR_ARGUMENT, "subgrid_size >= grid_size");
    if (c->wlayer_count < 1) return fail(P9IDG_ERR_ARGUMENT, "wlayer_count must be >= 1");
    if (c->visibilities_count <= 0) return fail(P9IDG_ERR_ARGUMENT, "visibilities_count must be > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(P9IDG_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(P9IDG_ERR_ARGUMENT, "device index out of range");
    auto* p = new p9idg_plan();
    p->c = *c;
    p->device = device;
    DeviceGuard dg(device);
    auto bail = [&](int st) {
        p9idg_destroy(p);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    for (auto& ev : p->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
    }
    for (int i = 0; i < 8; i++) p->mark[i] = p->ev[i];
    const int S = c->subgrid_size, G = c->grid_size;
    // tables: twiddles exp(+2 pi i m / S) and the adder phasor
    // exp(i pi (x + y - S) / S) (Adder.cs:37-40, AdderWStack.cs:17-25), FP64 -> float
    std::vector<float2> tw(S), ph((size_t)S * S);
    for (int m = 0; m < S; m++) {
        double a = 2.0 * M_PI * (double)m / (double)S;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            double a = M_PI * (double)(x + y - S) / (double)S;
            ph[(size_t)y * S + x] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
    int st;
    if ((st = dev_alloc(&p->d_twiddle, (size_t)S))) return bail(st);
    if ((st = dev_alloc(&p->d_phasor, (size_t)S * S))) return bail(st);
    if ((st = dev_alloc(&p->d_grid, (size_t)G * G))) return bail(st);
    if ((st = dev_alloc(&p->d_model_grid, (size_t)G * G))) return bail(st);
    if ((st = dev_alloc(&p->d_image, (size_t)G * G))) return bail(st);
    if ((st = dev_alloc(&p->d_counter, (size_t)2))) return bail(st);
    // the queue kernels leave their counters at zero; they start there
    e = launch_reset_counter(p->d_counter, p->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "counter reset"));
    e = hipMemcpy(p->d_twiddle, tw.data(), S * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "twiddle upload"));
    e = hipMemcpy(p->d_phasor, ph.data(), ph.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "phasor upload"));
    if ((st = dev_alloc(&p->d_rtab, (size_t)S * S))) return bail(st);
    e = launch_rtab(p->d_rtab, S, p->c.scale_arcsec * (float)G, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "residual table"));
    *plan = p;
    return P9IDG_OK;
}

int p9idg_get_constants(const p9idg_plan* p, p9idg_constants* c, int32_t* device) {
    if (int st = check_plan(p)) return st;
    if (c) *c = p->c;
    if (device) *device = p->device;
    return P9IDG_OK;
}

int p9idg_destroy(p9idg_plan* p) {
    if (!p) return P9IDG_OK;
    DeviceGuard dg(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->have_fft_one) (void)hipfftDestroy(p->fft_one);
    dev_free(p->d_uvw);
    dev_free(p->d_wavenumbers);
    dev_free(p->d_meta);
    dev_free(p->d_order);
    dev_free(p->d_vis);
    dev_free(p->d_model_vis);
    dev_free(p->d_grid);
    dev_free(p->d_layers);
    dev_free(p->d_model_grid);
    dev_free(p->d_image);
    dev_free(p->d_image64);
    dev_free(p->d_lsum);
    dev_free(p->d_twiddle);
    dev_free(p->d_phasor);
    dev_free(p->d_rtab);
    dev_free(p->d_flags);
    dev_free(p->d_counter);
    dev_free(p->d_resid);
    for (auto& f : p->d_fp) dev_free(f);
    drop_graph(p);
    dev_free(p->d_items_grid);
    dev_free(p->d_items_degrid);
    dev_free(p->d_scratch);
    for (auto& g : p->gather) free_gather(g);
    if (p->d_staging) (void)hipFree(p->d_staging);
    for (auto& ev : p->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : p->event_pool) (void)hipEventDestroy(ev);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return P9IDG_OK;
}

int p9idg_set_kernel(p9idg_plan* p, int32_t kernel) {
    if (int st = check_plan(p)) return st;
    if (kernel != P9IDG_KERNEL_FP32 && kernel != P9IDG_KERNEL_MFMA_BF16 &&
        kernel != P9IDG_KERNEL_MFMA_FP16 && kernel != P9IDG_KERNEL_FP32_SEP)
        return fail(P9IDG_ERR_ARGUMENT, "unknown kernel");
    if (p->kernel != kernel) drop_graph(p);  // a captured step holds the old kernel
    p->kernel = kernel;
    return P9IDG_OK;
}

int p9idg_set_deterministic(p9idg_plan* p, int32_t on) {
    if (int st = check_plan(p)) return st;
    if ((on != 0) != p->deterministic) drop_graph(p);
    p->deterministic = on != 0;
    return P9IDG_OK;
}

int p9idg_get_kernel(const p9idg_plan* p, int32_t* kernel) {
    if (int st = check_plan(p)) return st;
    if (!kernel) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *kernel = p->kernel;
    return P9IDG_OK;
}

int p9idg_set_launch(p9idg_plan* p, int32_t queue, int32_t grid_chunk, int32_t degrid_chunk,
                     int32_t workgroups) {
    if (int st = check_plan(p)) return st;
    if (grid_chunk < 0 || degrid_chunk < 0 || workgroups < 0)
        return fail(P9IDG_ERR_ARGUMENT, "negative launch parameter");
    p->use_queue = queue != 0;
    p->grid_chunk = grid_chunk;
    p->degrid_chunk = degrid_chunk;
    p->workgroups = workgroups;
    p->workgroups_degrid = workgroups;
    if (p->have_obs) return upload_items(p);
    return P9IDG_OK;
}

int p9idg_get_launch(const p9idg_plan* p, int32_t* queue, int32_t* grid_chunk,
                     int32_t* degrid_chunk, int32_t* workgroups) {
    if (int st = check_plan(p)) return st;
    if (queue) *queue = p->use_queue;
    if (grid_chunk) *grid_chunk = p->grid_chunk;
    if (degrid_chunk) *degrid_chunk = p->degrid_chunk;
    if (workgroups) *workgroups = p->workgroups;
    return P9IDG_OK;
}

int p9idg_set_observation(p9idg_plan* p, const double* uvw, int32_t B, int32_t T,
                          const double* freq, int32_t C, const p9idg_subgrid* meta,
                          int64_t nsub) {
    if (int st = check_plan(p)) return st;
    if (B <= 0 || T <= 0 || C <= 0 || nsub < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid dimensions");
    if (!uvw || !freq || (!meta && nsub > 0)) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (nsub > INT32_MAX) return fail(P9IDG_ERR_ARGUMENT, "too many subgrids");
    const int S = p->c.subgrid_size, G = p->c.grid_size, L = p->c.wlayer_count;
    bool plain_ok = true;
    int64_t gridded = 0;
    for (int64_t i = 0; i < nsub; i++) {
        const p9idg_subgrid& m = meta[i];
        if (m.baseline_idx < 0 || m.baseline_idx >= B || m.time_sample_start < 0 ||
            m.time_sample_count < 0 || (int64_t)m.time_sample_start + m.time_sample_count > T)
            return fail(P9IDG_ERR_ARGUMENT, "subgrid " + std::to_string(i) + " indexes outside the observation");
        if (m.w_lambda < -L || m.w_lambda >= L)
            return fail(P9IDG_ERR_ARGUMENT, "subgrid " + std::to_string(i) + " has a w index outside the layers");
        if (!(m.u_pixel >= 0 && m.u_pixel < G - S && m.v_pixel >= 0 && m.v_pixel < G - S))
            plain_ok = false;
        // AdderWStack addresses only 1 <= X,Y < G-S after mirroring; the
        // splitter reads up to G - X - x.  Any other corner is skipped by
        // both, so no further check is needed for memory safety.
        gridded += (int64_t)m.time_sample_count * C;
    }
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));  // queued launches may still read the buffers freed below
    p->obs_gen++;
    // drop every buffer whose allocated size differs from the new shape (by
    // the recorded sizes, so a call that failed part-way leaves nothing
    // undersized behind)
    drop_graph(p);
    const size_t n_uvw = (size_t)B * T * 3, n_sub = (size_t)std::max<int64_t>(nsub, 1), n_vis = (size_t)B * T * C;
    if (p->alloc_vis != n_vis) {
        dev_free(p->d_vis);
        dev_free(p->d_model_vis);
        dev_free(p->d_flags);
        dev_free(p->d_resid);
        p->have_flags = false;
        p->alloc_vis = 0;
    }
    if (p->alloc_uvw != n_uvw) {
        dev_free(p->d_uvw);
        p->alloc_uvw = 0;
    }
    if (p->alloc_wn != (size_t)C) {
        dev_free(p->d_wavenumbers);
        p->alloc_wn = 0;
    }
    if (p->alloc_sub != n_sub) {
        dev_free(p->d_meta);
        dev_free(p->d_order);
        p->alloc_sub = 0;
    }
    p->have_obs = false;
    int st;
    if ((st = dev_alloc(&p->d_uvw, n_uvw))) return st;
    p->alloc_uvw = n_uvw;
    if ((st = dev_alloc(&p->d_wavenumbers, (size_t)C))) return st;
    p->alloc_wn = (size_t)C;
    if ((st = dev_alloc(&p->d_meta, n_sub)) || (st = dev_alloc(&p->d_order, n_sub))) return st;
    p->alloc_sub = n_sub;
    if ((st = dev_alloc(&p->d_vis, n_vis)) || (st = dev_alloc(&p->d_model_vis, n_vis))) return st;
    p->alloc_vis = n_vis;
    // MathFunctions.cs:69-78: k = 2 pi f / c, (2.0f promoted to double)
    std::vector<double> k(C);
    for (int i = 0; i < C; i++) k[i] = ((2.0 * M_PI) * freq[i]) / 299792458.0;
    bool uniform = C >= 2;
    if (uniform) {
        const double dk = (k[C - 1] - k[0]) / (C - 1), tol = 1e-12 * std::fabs(k[C - 1]);
        for (int i = 0; i < C; i++) uniform = uniform && std::fabs(k[i] - (k[0] + i * dk)) <= tol;
    }
    // schedule: most visibilities first
    std::vector<int32_t> order(nsub);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return meta[a].time_sample_count > meta[b].time_sample_count;
    });
    P9_HIP(hipMemcpyAsync(p->d_uvw, uvw, (size_t)B * T * 3 * sizeof(double), hipMemcpyHostToDevice, p->stream));
    P9_HIP(hipMemcpyAsync(p->d_wavenumbers, k.data(), C * sizeof(double), hipMemcpyHostToDevice, p->stream));
    if (nsub > 0) {
        P9_HIP(hipMemcpyAsync(p->d_meta, meta, nsub * sizeof(p9idg_subgrid), hipMemcpyHostToDevice, p->stream));
        P9_HIP(hipMemcpyAsync(p->d_order, order.data(), nsub * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
    }
    P9_HIP(hipStreamSynchronize(p->stream));
    p->chan_uniform = uniform;
    p->B = B;
    p->T = T;
    p->C = C;
    p->nsub = nsub;
    p->plain_adder_ok = plain_ok;
    p->gridded_vis = gridded;
    p->meta_host.assign(meta, meta + nsub);
    p->order_host = order;
    mark_unknown(p);
    if (int st = build_footprints(p, meta, nsub)) return st;
    if (int st = upload_items(p)) return st;
    p->have_obs = true;
    return P9IDG_OK;
}

// ---- host-buffer drop-ins ----------------------------------------------------
int p9idg_grid(p9idg_plan* p, const void* vis, int32_t type, double* grid_out) {
    if (int st = check_obs(p)) return st;
    if (!grid_out) return fail(P9IDG_ERR_ARGUMENT, "grid_out is NULL");
    if (!p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    if (int st = upload_vis_any(p, vis, type)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_PLAIN)) return st;
    return download_c128(p, p->d_grid, grid_out, grid_elems(p));
}

int p9idg_gridw(p9idg_plan* p, const void* vis, int32_t type, double* layers_out) {
    if (int st = check_obs(p)) return st;
    if (!layers_out) return fail(P9IDG_ERR_ARGUMENT, "layers_out is NULL");
    if (int st = upload_vis_any(p, vis, type)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_WSTACK)) return st;
    return download_c128(p, p->d_layers, layers_out, grid_elems(p) * p->c.wlayer_count);
}

static int degrid_host(p9idg_plan* p, const double* grid, double* vis_out, int mode) {
    if (int st = check_obs(p)) return st;
    if (!vis_out) return fail(P9IDG_ERR_ARGUMENT, "vis_out is NULL");
    DeviceGuard dg(p->device);
    if (int st = upload_grid_c128(p, grid)) return st;
    if (int st = do_exec_degrid(p, mode)) return st;
    return download_c128(p, p->d_model_vis, vis_out, vis_elems(p));
}

int p9idg_degrid(p9idg_plan* p, const double* grid, double* vis_out) {
    return degrid_host(p, grid, vis_out, P9IDG_DEGRID_PLAIN);
}

int p9idg_degridw(p9idg_plan* p, const double* grid, double* vis_out) {
    return degrid_host(p, grid, vis_out, P9IDG_DEGRID_WSTACK);
}

int p9idg_grid_psf(p9idg_plan* p, const uint8_t* flags, double* grid_out) {
    if (int st = check_obs(p)) return st;
    if (!grid_out) return fail(P9IDG_ERR_ARGUMENT, "grid_out is NULL");
    if (!p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    if (int st = set_psf_vis(p, flags)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_PLAIN)) return st;
    return download_c128(p, p->d_grid, grid_out, grid_elems(p));
}

static int image_f64_host(p9idg_plan* p, double* img_out) {
    DeviceGuard dg(p->device);
    if (int st = ensure_fft_one(p)) return st;
    if (int st = dev_alloc(&p->d_image64, grid_elems(p))) return st;
    p->grid_clean = false;
    P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_grid),
                         reinterpret_cast<hipfftComplex*>(p->d_grid), HIPFFT_BACKWARD));
    P9_HIP(launch_grid_to_image_f64(p->d_grid, p->c.grid_size,
                                    1.0 / (double)p->c.visibilities_count, 1, p->d_image64,
                                    p->stream));
    P9_HIP(hipMemcpyAsync(img_out, p->d_image64, grid_elems(p) * sizeof(double),
                          hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_to_image(p9idg_plan* p, const void* vis, int32_t type, double* img_out) {
    if (int st = check_obs(p)) return st;
    if (!img_out) return fail(P9IDG_ERR_ARGUMENT, "img_out is NULL");
    if (!p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    if (int st = upload_vis_any(p, vis, type)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_PLAIN)) return st;
    return image_f64_host(p, img_out);
}

int p9idg_calculate_psf(p9idg_plan* p, const uint8_t* flags, double* psf_out) {
    if (int st = check_obs(p)) return st;
    if (!psf_out) return fail(P9IDG_ERR_ARGUMENT, "psf_out is NULL");
    if (!p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    if (int st = set_psf_vis(p, flags)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_PLAIN)) return st;
    return image_f64_host(p, psf_out);
}

int p9idg_gridw_to_image(p9idg_plan* p, const void* vis, int32_t type, float* img_out) {
    if (int st = check_obs(p)) return st;
    if (!img_out) return fail(P9IDG_ERR_ARGUMENT, "img_out is NULL");
    if (int st = upload_vis_any(p, vis, type)) return st;
    if (int st = do_exec_grid(p, P9IDG_GRID_WSTACK)) return st;
    if (int st = do_grid_to_image(p, P9IDG_GRID_WSTACK)) return st;
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(img_out, p->d_image, grid_elems(p) * sizeof(float),
                          hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_image_to_vis(p9idg_plan* p, const float* img, double* vis_out) {
    if (int st = check_obs(p)) return st;
    if (!img || !vis_out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(p->d_image, img, grid_elems(p) * sizeof(float), hipMemcpyHostToDevice,
                          p->stream));
    if (int st = do_image_to_grid(p)) return st;
    if (int st = do_exec_degrid(p, P9IDG_DEGRID_WSTACK)) return st;
    return download_c128(p, p->d_model_vis, vis_out, vis_elems(p));
}

// ---- FFT.cs static steps --------------------------------------------------------
int p9idg_wstack_ifft_float(p9idg_plan* p, const double* layers, int32_t L, int64_t vis_count,
                            float* out) {
    if (int st = check_plan(p)) return st;
    if (!layers || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (L != p->c.wlayer_count) return fail(P9IDG_ERR_ARGUMENT, "layer count differs from the plan's");
    if (vis_count == 0) return fail(P9IDG_ERR_ARGUMENT, "vis_count is 0");
    DeviceGuard dg(p->device);
    if (int st = ensure_layers(p)) return st;
    const size_t n = grid_elems(p) * L;
    p->layers_clean = false;
    if (int st = upload_c128(p, layers, p->d_layers, n)) return st;
    if (int st = layers_image(p, 1.0 / (double)vis_count, 0)) return st;
    P9_HIP(hipMemcpyAsync(out, p->d_image, grid_elems(p) * sizeof(float), hipMemcpyDeviceToHost,
                          p->stream));
    return sync(p);
}

int p9idg_fft_forward_float(p9idg_plan* p, const float* img, double norm, double* grid_out) {
    if (int st = check_plan(p)) return st;
    if (!img || !grid_out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    if (int st = ensure_fft_one(p)) return st;
    const int G = p->c.grid_size;
    P9_HIP(hipMemcpyAsync(p->d_image, img, grid_elems(p) * sizeof(float), hipMemcpyHostToDevice,
                          p->stream));
    P9_HIP(launch_image_to_grid(p->d_image, G, 0, p->d_model_grid, p->stream));
    P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_model_grid),
                         reinterpret_cast<hipfftComplex*>(p->d_model_grid), HIPFFT_FORWARD));
    if (norm != 1.0) P9_HIP(launch_scale_c64(p->d_model_grid, grid_elems(p), 1.0 / norm, p->stream));
    return download_c128(p, p->d_model_grid, grid_out, grid_elems(p));
}

int p9idg_fft_backward(p9idg_plan* p, const double* grid, int64_t vis_count, double* out) {
    if (int st = check_plan(p)) return st;
    if (!grid || !out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (vis_count == 0) return fail(P9IDG_ERR_ARGUMENT, "vis_count is 0");
    DeviceGuard dg(p->device);
    if (int st = ensure_fft_one(p)) return st;
    if (int st = dev_alloc(&p->d_image64, grid_elems(p))) return st;
    p->grid_clean = false;
    if (int st = upload_c128(p, grid, p->d_grid, grid_elems(p))) return st;
    P9_FFT(hipfftExecC2C(p->fft_one, reinterpret_cast<hipfftComplex*>(p->d_grid),
                         reinterpret_cast<hipfftComplex*>(p->d_grid), HIPFFT_BACKWARD));
    P9_HIP(launch_grid_to_image_f64(p->d_grid, p->c.grid_size, 1.0 / (double)vis_count, 0,
                                    p->d_image64, p->stream));
    P9_HIP(hipMemcpyAsync(out, p->d_image64, grid_elems(p) * sizeof(double),
                          hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_shift_f32(float* a, int32_t n) {
    if (!a || n < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid argument");
    shift_host(a, n);
    return P9IDG_OK;
}
int p9idg_shift_f64(double* a, int32_t n) {
    if (!a || n < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid argument");
    shift_host(a, n);
    return P9IDG_OK;
}
int p9idg_shift_c128(double* a, int32_t n) {
    if (!a || n < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid argument");
    shift_host(reinterpret_cast<c128*>(a), n);
    return P9IDG_OK;
}

int p9idg_substract(const double* vis0, const double* vis1, const uint8_t* flags, int64_t n,
                    double* out) {
    if (!vis0 || !vis1 || !out || n < 0) return fail(P9IDG_ERR_ARGUMENT, "invalid argument");
    for (int64_t i = 0; i < n; i++) {
        if (flags && flags[i]) {
            out[2 * i] = 0.0;
            out[2 * i + 1] = 0.0;
        } else {
            out[2 * i] = vis0[2 * i] - vis1[2 * i];
            out[2 * i + 1] = vis0[2 * i + 1] - vis1[2 * i + 1];
        }
    }
    return P9IDG_OK;
}

// ---- device-resident API -----------------------------------------------------------
int p9idg_upload_vis(p9idg_plan* p, const void* vis, int32_t type) {
    if (int st = upload_vis_any(p, vis, type)) return st;
    return sync(p);
}
int p9idg_set_psf_vis(p9idg_plan* p, const uint8_t* flags) {
    if (int st = set_psf_vis(p, flags)) return st;
    return sync(p);
}
int p9idg_exec_grid(p9idg_plan* p, int32_t mode) { return do_exec_grid(p, mode); }
int p9idg_exec_grid_to_image(p9idg_plan* p, int32_t mode) { return do_grid_to_image(p, mode); }
int p9idg_exec_image_to_grid(p9idg_plan* p) { return do_image_to_grid(p); }
int p9idg_exec_degrid(p9idg_plan* p, int32_t mode) { return do_exec_degrid(p, mode); }
int p9idg_set_flags(p9idg_plan* p, const uint8_t* flags) {
    if (int st = check_obs(p)) return st;
    DeviceGuard dg(p->device);
    const uint8_t* before = p->have_flags ? p->d_flags : nullptr;
    if (!flags) {
        p->have_flags = false;
        return flags_changed(p, before);
    }
    if (int st = dev_alloc(&p->d_flags, vis_elems(p))) return st;
    P9_HIP(hipMemcpyAsync(p->d_flags, flags, vis_elems(p), hipMemcpyHostToDevice, p->stream));
    p->have_flags = true;
    if (int st = flags_changed(p, before)) return st;
    return sync(p);
}

int p9idg_exec_residual(p9idg_plan* p) { return do_residual(p); }

int p9idg_set_grid_source(p9idg_plan* p, int32_t src) {
    if (int st = check_plan(p)) return st;
    if (src != P9IDG_SRC_VIS && src != P9IDG_SRC_RESIDUAL) return fail(P9IDG_ERR_ARGUMENT, "unknown grid source");
    p->grid_src = src;
    return P9IDG_OK;
}

int p9idg_exec_cycle(p9idg_plan* p, int32_t grid_mode) {
    if (int st = check_obs(p)) return st;
    return cycle_step(p, grid_mode);
}

int p9idg_capture_cycle(p9idg_plan* p, int32_t grid_mode) {
    if (int st = check_obs(p)) return st;
    if (grid_mode == P9IDG_GRID_PLAIN && !p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
    DeviceGuard dg(p->device);
    drop_graph(p);
    // one eager step creates the FFT plans, buffers and occupancy figures
    if (int st = cycle_step(p, grid_mode)) return st;
    P9_HIP(hipStreamSynchronize(p->stream));
    P9_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    p->capturing = true;
    int st = cycle_step(p, grid_mode);
    p->capturing = false;
    mark_unknown(p);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    p->graph = g;
    e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(p);
        return hip_fail(e, "hipGraphInstantiate");
    }
    return P9IDG_OK;
}

int p9idg_launch_cycle(p9idg_plan* p) {
    if (int st = check_plan(p)) return st;
    if (!p->graph_exec) return fail(P9IDG_ERR_STATE, "no captured cycle (p9idg_capture_cycle)");
    DeviceGuard dg(p->device);
    P9_HIP(hipGraphLaunch(p->graph_exec, p->stream));
    return P9IDG_OK;
}

// Times one gridder (dir 0) or degridder (dir 1) launch with the plan's
// current launch shape; median of `reps` runs.
// Median kernel time of one direction over `reps` launches, in the mode a
// major cycle runs: w-stacked when the constants have w-layers.
static int time_direction(p9idg_plan* p, int dir, int reps, float* ms_out) {
    const bool wstack = p->c.wlayer_count > 1;
    std::vector<float> t;
    for (int r = 0; r < reps; r++) {
        int st = dir == 0 ? do_exec_grid(p, wstack ? P9IDG_GRID_WSTACK : P9IDG_GRID_PLAIN)
                          : do_exec_degrid(p, wstack ? P9IDG_DEGRID_WSTACK : P9IDG_DEGRID_PLAIN);
        if (st) return st;
        P9_HIP(hipStreamSynchronize(p->stream));
        float ms = 0.f;
        P9_HIP(hipEventElapsedTime(&ms, p->mark[dir == 0 ? 1 : 3], p->mark[dir == 0 ? 2 : 4]));
        t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return P9IDG_OK;
}

int p9idg_autotune(p9idg_plan* p, int32_t reps, float* grid_ms, float* degrid_ms) {
    if (int st = check_obs(p)) return st;
    if (reps < 1) reps = 1;
    DeviceGuard dg(p->device);
    const bool collect = p->collect;
    p->collect = false;
    int cus = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess) cus = 1;
    // candidate shapes: one workgroup per subgrid, or the queue with
    // items of `chunk` visibilities on `wg` persistent workgroups
    struct Cand {
        int queue, chunk, wg;
    };
    std::vector<Cand> cands = {{0, 0, 0}};
    for (int chunk : {512, 1024, 2048, 4096, 0})
        for (int wg : {0, cus, 2 * cus, 4 * cus}) cands.push_back({1, chunk, wg});
    float best[2] = {1e30f, 1e30f};
    Cand pick[2] = {{p->use_queue, p->grid_chunk, p->workgroups},
                    {p->use_queue, p->degrid_chunk, p->workgroups_degrid}};
    const int32_t q0 = p->use_queue, g0 = p->grid_chunk, d0 = p->degrid_chunk, w0 = p->workgroups;
    for (const Cand& c : cands) {
        p->use_queue = c.queue;
        p->grid_chunk = c.chunk;
        p->degrid_chunk = c.chunk;
        p->workgroups = c.wg;
        p->workgroups_degrid = c.wg;
        if (int st = upload_items(p)) return st;
        for (int dir = 0; dir < 2; dir++) {
            float ms = 0.f;
            if (int st = time_direction(p, dir, reps, &ms)) {
                p->collect = collect;
                return st;
            }
            if (ms < best[dir]) {
                best[dir] = ms;
                pick[dir] = c;
            }
        }
    }
    (void)q0;
    (void)g0;
    (void)d0;
    (void)w0;
    // the queue flag is shared by both directions: take it from the direction
    // with the larger time, workgroups and chunks from each (with the queue
    // off they are unused)
    const int dom = best[0] >= best[1] ? 0 : 1;
    p->use_queue = pick[dom].queue;
    p->workgroups = pick[0].wg;
    p->workgroups_degrid = pick[1].wg;
    p->grid_chunk = pick[0].chunk;
    p->degrid_chunk = pick[1].chunk;
    p->collect = collect;
    if (int st = upload_items(p)) return st;
    if (grid_ms) *grid_ms = best[0];
    if (degrid_ms) *degrid_ms = best[1];
    return P9IDG_OK;
}

int p9idg_download_residual(p9idg_plan* p, void* out) {
    if (int st = check_obs(p)) return st;
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    if (!p->d_resid) return fail(P9IDG_ERR_STATE, "no residual visibilities");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(out, p->d_resid, vis_elems(p) * sizeof(float2), hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_sync(p9idg_plan* p) {
    if (int st = check_plan(p)) return st;
    DeviceGuard dg(p->device);
    return sync(p);
}

int p9idg_device_buffers(p9idg_plan* p, void** vis, void** grid, void** model_grid,
                         void** model_vis, void** image) {
    if (int st = check_plan(p)) return st;
    if (vis) *vis = p->d_vis;
    if (grid) *grid = p->d_grid;
    mark_unknown(p);  // the caller may write the buffers
    if (model_grid) *model_grid = p->d_model_grid;
    if (model_vis) *model_vis = p->d_model_vis;
    if (image) *image = p->d_image;
    return P9IDG_OK;
}

int p9idg_download_grid(p9idg_plan* p, int32_t mode, void* out) {
    if (int st = check_obs(p)) return st;
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    const float2* src = p->d_grid;
    size_t n = grid_elems(p);
    if (mode == P9IDG_GRID_WSTACK) {
        if (!p->d_layers) return fail(P9IDG_ERR_STATE, "no layers gridded");
        src = p->d_layers;
        n *= p->c.wlayer_count;
    }
    P9_HIP(hipMemcpyAsync(out, src, n * sizeof(float2), hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_upload_model_grid(p9idg_plan* p, const void* grid) {
    if (int st = check_plan(p)) return st;
    if (!grid) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(p->d_model_grid, grid, grid_elems(p) * sizeof(float2),
                          hipMemcpyHostToDevice, p->stream));
    return sync(p);
}

int p9idg_download_model_grid(p9idg_plan* p, void* out) {
    if (int st = check_plan(p)) return st;
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(out, p->d_model_grid, grid_elems(p) * sizeof(float2),
                          hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_download_model_vis(p9idg_plan* p, void* out) {
    if (int st = check_obs(p)) return st;
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(out, p->d_model_vis, vis_elems(p) * sizeof(float2),
                          hipMemcpyDeviceToHost, p->stream));
    return sync(p);
}

int p9idg_upload_image(p9idg_plan* p, const float* img) {
    if (int st = check_plan(p)) return st;
    if (!img) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(p->d_image, img, grid_elems(p) * sizeof(float), hipMemcpyHostToDevice,
                          p->stream));
    return sync(p);
}

int p9idg_download_image(p9idg_plan* p, float* out) {
    if (int st = check_plan(p)) return st;
    if (!out) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    DeviceGuard dg(p->device);
    P9_HIP(hipMemcpyAsync(out, p->d_image, grid_elems(p) * sizeof(float), hipMemcpyDeviceToHost,
                          p->stream));
    return sync(p);
}

int p9idg_last_timings(p9idg_plan* p, float* grid_ms, float* degrid_ms, float* fft_ms) {
    if (int st = check_plan(p)) return st;
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));
    float ms = 0.f;
    if (grid_ms) {
        *grid_ms = -1.f;
        if (p->timed_grid && hipEventElapsedTime(&ms, p->mark[1], p->mark[2]) == hipSuccess) *grid_ms = ms;
    }
    if (degrid_ms) {
        *degrid_ms = -1.f;
        if (p->timed_degrid && hipEventElapsedTime(&ms, p->mark[3], p->mark[4]) == hipSuccess) *degrid_ms = ms;
    }
    if (fft_ms) {
        *fft_ms = -1.f;
        if (p->timed_fft && hipEventElapsedTime(&ms, p->mark[5], p->mark[6]) == hipSuccess) *fft_ms = ms;
    }
    return P9IDG_OK;
}

int p9idg_timing_reset(p9idg_plan* p, int32_t collect) {
    if (int st = check_plan(p)) return st;
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));
    p->spans.clear();
    p->collect = collect != 0;
    // the pooled span events are recorded afresh from here on: drop the
    // timing slots that pointed at them
    p->last_span_event = nullptr;
    for (int i = 0; i < 8; i++) p->mark[i] = p->ev[i];
    p->timed_grid = p->timed_degrid = p->timed_fft = false;
    return P9IDG_OK;
}

int p9idg_timing_stage(p9idg_plan* p, int32_t stage, double* ms_total, int64_t* spans) {
    if (int st = check_plan(p)) return st;
    if (stage < P9IDG_STAGE_GRIDDER || stage > P9IDG_STAGE_TO_UVW) return fail(P9IDG_ERR_ARGUMENT, "unknown stage");
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));
    double t = 0.0;
    int64_t n = 0;
    for (const auto& sp : p->spans) {
        if (sp.kind != stage) continue;
        float ms = 0.f;
        P9_HIP(hipEventElapsedTime(&ms, sp.start, sp.stop));
        t += ms;
        n++;
    }
    if (ms_total) *ms_total = t;
    if (spans) *spans = n;
    return P9IDG_OK;
}

int p9idg_timing_totals(p9idg_plan* p, double* grid_ms, int64_t* grid_launches,
                        double* degrid_ms, int64_t* degrid_launches) {
    if (int st = check_plan(p)) return st;
    DeviceGuard dg(p->device);
    P9_HIP(hipStreamSynchronize(p->stream));
    double t[2] = {0.0, 0.0};
    int64_t n[2] = {0, 0};
    for (const auto& sp : p->spans) {
        if (sp.kind > P9IDG_STAGE_DEGRIDDER) continue;
        float ms = 0.f;
        P9_HIP(hipEventElapsedTime(&ms, sp.start, sp.stop));
        t[sp.kind] += ms;
        n[sp.kind]++;
    }
    if (grid_ms) *grid_ms = t[0];
    if (grid_launches) *grid_launches = n[0];
    if (degrid_ms) *degrid_ms = t[1];
    if (degrid_launches) *degrid_launches = n[1];
    return P9IDG_OK;
}

}  // extern "C"

// ---- batches: several observations per persistent launch --------------------
struct p9idg_batch {
    std::vector<p9idg_plan*> plans;
    std::vector<uint64_t> obs_gen;     // each plan's observation generation at prepare
    std::vector<hipEvent_t> plan_ev;   // per plan: its stream's work before a batch run
    hipEvent_t done = nullptr;         // end of the last batch run, awaited by the plans
    int device = 0;
    hipStream_t stream = nullptr;
    struct Group {
        int S = 0, kernel = 0, workgroups = 0, workgroups_degrid = 0;
        std::vector<int> members;  // plan indices; item.inst indexes this list
        KernelArgs* d_args_grid = nullptr;
        KernelArgs* d_args_degrid = nullptr;
        WorkItem* d_items_grid = nullptr;
        WorkItem* d_items_degrid = nullptr;
        int32_t n_grid = 0, n_degrid = 0;
        int32_t* d_counter = nullptr;
    };
    std::vector<Group> groups;
    int grid_mode = P9IDG_GRID_WSTACK, degrid_mode = P9IDG_DEGRID_WSTACK;
    bool prepared = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev[5] = {};
    bool timed = false;
    // with several groups, each group's launch runs on a stream of its own
    // between a fork and a join on `stream`, so one group's launch fills the
    // CUs that another's tail leaves idle
    hipEvent_t fork = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_done;
};

namespace {

void batch_free_groups(p9idg_batch* b) {
    for (size_t i = 0; i < b->side.size(); i++) {
        if (b->side[i]) {
            (void)hipStreamSynchronize(b->side[i]);
            (void)hipStreamDestroy(b->side[i]);
        }
        if (b->side_done[i]) (void)hipEventDestroy(b->side_done[i]);
    }
    b->side.clear();
    b->side_done.clear();
    for (auto& g : b->groups) {
        dev_free(g.d_args_grid);
        dev_free(g.d_args_degrid);
        dev_free(g.d_items_grid);
        dev_free(g.d_items_degrid);
        dev_free(g.d_counter);
    }
    b->groups.clear();
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->prepared = false;
}

// The batch works on the plans' buffers on its own stream: it starts after
// the work already queued on every plan's stream, and every plan's later
// work starts after it.
int batch_check(const p9idg_batch* b) {
    if (!b->prepared) return fail(P9IDG_ERR_STATE, "batch not prepared (p9idg_batch_prepare)");
    for (size_t i = 0; i < b->plans.size(); i++) {
        if (b->plans[i]->obs_gen != b->obs_gen[i])
            return fail(P9IDG_ERR_STATE, "a plan's observation changed after p9idg_batch_prepare");
        if (b->plans[i]->deterministic)
            return fail(P9IDG_ERR_UNSUPPORTED, "the deterministic adder is not available in batches");
    }
    return P9IDG_OK;
}
int batch_order_in(p9idg_batch* b) {
    for (size_t i = 0; i < b->plans.size(); i++) {
        P9_HIP(hipEventRecord(b->plan_ev[i], b->plans[i]->stream));
        P9_HIP(hipStreamWaitEvent(b->stream, b->plan_ev[i], 0));
    }
    return P9IDG_OK;
}
int batch_order_out(p9idg_batch* b) {
    P9_HIP(hipEventRecord(b->done, b->stream));
    for (p9idg_plan* p : b->plans) P9_HIP(hipStreamWaitEvent(p->stream, b->done, 0));
    return P9IDG_OK;
}

// One persistent launch per group, on the batch stream or fanned out over the
// groups' streams.
int batch_launch_groups(p9idg_batch* b, bool grid) {
    const bool fan = b->side.size() == b->groups.size() && b->groups.size() > 1;
    if (fan) P9_HIP(hipEventRecord(b->fork, b->stream));
    for (size_t i = 0; i < b->groups.size(); i++) {
        auto& g = b->groups[i];
        hipStream_t st = fan ? b->side[i] : b->stream;
        if (fan) P9_HIP(hipStreamWaitEvent(st, b->fork, 0));
        hipError_t e;
        if (grid) {
            BatchArgs ba{g.d_args_grid, g.d_items_grid, g.n_grid, g.d_counter};
            e = launch_grid_batch(ba, g.S, g.kernel, g.workgroups, st);
        } else {
            BatchArgs ba{g.d_args_degrid, g.d_items_degrid, g.n_degrid, g.d_counter};
            e = launch_degrid_batch(ba, g.S, g.kernel, g.workgroups_degrid, st);
        }
        if (e != hipSuccess) return hip_fail(e, grid ? "batched gridder launch" : "batched degridder launch");
        if (fan) P9_HIP(hipEventRecord(b->side_done[i], st));
    }
    if (fan)
        for (size_t i = 0; i < b->groups.size(); i++) P9_HIP(hipStreamWaitEvent(b->stream, b->side_done[i], 0));
    return P9IDG_OK;
}

// Zero fills and one persistent launch per (subgrid size, kernel) group.
int batch_run(p9idg_batch* b, int what, bool capturing) {
    if (int st = batch_check(b)) return st;
    for (p9idg_plan* p : b->plans) mark_unknown(p);
    if (!capturing) P9_HIP(hipEventRecord(b->ev[0], b->stream));
    if (what & 1) {
        for (p9idg_plan* p : b->plans) {
            float2* target = b->grid_mode == P9IDG_GRID_WSTACK ? p->d_layers : p->d_grid;
            size_t n = grid_elems(p) * (b->grid_mode == P9IDG_GRID_WSTACK ? p->c.wlayer_count : 1);
            P9_HIP(launch_fill_zero(target, n, b->stream));
        }
        if (!capturing) P9_HIP(hipEventRecord(b->ev[1], b->stream));
        if (int st = batch_launch_groups(b, true)) return st;
    }
    if (!capturing) P9_HIP(hipEventRecord(b->ev[2], b->stream));
    if (what & 2) {
        for (p9idg_plan* p : b->plans) P9_HIP(launch_fill_zero(p->d_model_vis, vis_elems(p), b->stream));
        if (!capturing) P9_HIP(hipEventRecord(b->ev[3], b->stream));
        if (int st = batch_launch_groups(b, false)) return st;
    }
    if (!capturing) {
        P9_HIP(hipEventRecord(b->ev[4], b->stream));
        b->timed = (what & 1) && (what & 2);
    }
    return P9IDG_OK;
}

}  // namespace

extern "C" {

int p9idg_batch_create(p9idg_plan* const* plans, int32_t n, p9idg_batch** out) {
    if (!plans || !out || n <= 0) return fail(P9IDG_ERR_ARGUMENT, "invalid batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!plans[i]) return fail(P9IDG_ERR_ARGUMENT, "NULL plan in batch");
        if (!plans[i]->have_obs) return fail(P9IDG_ERR_STATE, "plan without observation in batch");
        if (plans[i]->device != plans[0]->device) return fail(P9IDG_ERR_ARGUMENT, "plans on different devices");
    }
    auto* b = new p9idg_batch();
    b->plans.assign(plans, plans + n);
    b->device = plans[0]->device;
    DeviceGuard dg(b->device);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipStreamCreate");
    }
    for (auto& ev : b->ev) (void)hipEventCreate(&ev);
    b->plan_ev.assign(n, nullptr);
    b->obs_gen.assign(n, 0);
    bool ok = hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (auto& ev : b->plan_ev) ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        p9idg_batch_destroy(b);
        return fail(P9IDG_ERR_HIP, "hipEventCreate failed");
    }
    *out = b;
    return P9IDG_OK;
}

int p9idg_batch_destroy(p9idg_batch* b) {
    if (!b) return P9IDG_OK;
    DeviceGuard dg(b->device);
    (void)hipStreamSynchronize(b->stream);
    batch_free_groups(b);
    for (auto& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : b->plan_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->fork) (void)hipEventDestroy(b->fork);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return P9IDG_OK;
}

int p9idg_batch_prepare(p9idg_batch* b, int32_t grid_mode, int32_t degrid_mode) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    if (grid_mode < P9IDG_GRID_PLAIN || grid_mode > P9IDG_GRID_WSUM || degrid_mode < 0 || degrid_mode > 1)
        return fail(P9IDG_ERR_ARGUMENT, "unknown mode");
    DeviceGuard dg(b->device);
    batch_free_groups(b);
    b->grid_mode = grid_mode;
    b->degrid_mode = degrid_mode;
    for (int i = 0; i < (int)b->plans.size(); i++) {
        p9idg_plan* p = b->plans[i];
        mark_unknown(p);
        if (grid_mode == P9IDG_GRID_PLAIN && !p->plain_adder_ok) return fail(P9IDG_ERR_INVALID_GRID, kMsgInvalidGrid);
        if (p->deterministic) return fail(P9IDG_ERR_UNSUPPORTED, "the deterministic adder is not available in batches");
        if (grid_mode == P9IDG_GRID_WSTACK)
            if (int st = ensure_layers(p)) return st;
        P9_HIP(hipStreamSynchronize(p->stream));
        bool found = false;
        for (auto& g : b->groups)
            if (g.S == p->c.subgrid_size && g.kernel == p->kernel) {
                g.members.push_back(i);
                found = true;
            }
        if (!found) {
            p9idg_batch::Group g;
            g.S = p->c.subgrid_size;
            g.kernel = p->kernel;
            g.workgroups = p->workgroups;
            g.workgroups_degrid = p->workgroups_degrid;
            g.members.push_back(i);
            b->groups.push_back(g);
        }
    }
    for (auto& g : b->groups) {
        std::vector<KernelArgs> ag, ad;
        std::vector<WorkItem> ig, id;
        for (int m = 0; m < (int)g.members.size(); m++) {
            p9idg_plan* p = b->plans[g.members[m]];
            KernelArgs a = make_args(p, grid_mode);
            a.vis = p->d_vis;
            a.grid = grid_mode == P9IDG_GRID_WSTACK ? p->d_layers : p->d_grid;
            ag.push_back(a);
            KernelArgs d = make_args(p, degrid_mode);
            ad.push_back(d);
            for (WorkItem w : make_items(p->meta_host, p->order_host, p->C, p->grid_chunk)) {
                w.inst = m;
                ig.push_back(w);
            }
            for (WorkItem w : make_items(p->meta_host, p->order_host, p->C, p->degrid_chunk)) {
                w.inst = m;
                id.push_back(w);
            }
        }
        auto longest = [](const WorkItem& x, const WorkItem& y) { return x.k1 - x.k0 > y.k1 - y.k0; };
        std::stable_sort(ig.begin(), ig.end(), longest);
        std::stable_sort(id.begin(), id.end(), longest);
        int st;
        if ((st = dev_alloc(&g.d_args_grid, ag.size())) || (st = dev_alloc(&g.d_args_degrid, ad.size())) ||
            (st = dev_alloc(&g.d_items_grid, std::max<size_t>(ig.size(), 1))) ||
            (st = dev_alloc(&g.d_items_degrid, std::max<size_t>(id.size(), 1))) || (st = dev_alloc(&g.d_counter, (size_t)2)))
            return st;
        P9_HIP(launch_reset_counter(g.d_counter, b->stream));  // the queue kernels leave it at zero
        P9_HIP(hipMemcpy(g.d_args_grid, ag.data(), ag.size() * sizeof(KernelArgs), hipMemcpyHostToDevice));
        P9_HIP(hipMemcpy(g.d_args_degrid, ad.data(), ad.size() * sizeof(KernelArgs), hipMemcpyHostToDevice));
        if (!ig.empty()) P9_HIP(hipMemcpy(g.d_items_grid, ig.data(), ig.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
        if (!id.empty()) P9_HIP(hipMemcpy(g.d_items_degrid, id.data(), id.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
        g.n_grid = (int32_t)ig.size();
        g.n_degrid = (int32_t)id.size();
    }
    if (b->groups.size() > 1) {
        b->side.assign(b->groups.size(), nullptr);
        b->side_done.assign(b->groups.size(), nullptr);
        for (size_t i = 0; i < b->groups.size(); i++) {
            if (hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) != hipSuccess) b->side[i] = nullptr;
            if (hipEventCreateWithFlags(&b->side_done[i], hipEventDisableTiming) != hipSuccess) b->side_done[i] = nullptr;
            if (!b->side[i] || !b->side_done[i]) {
                (void)hipGetLastError();
                batch_free_groups(b);  // also drops the streams created so far
                return fail(P9IDG_ERR_HIP, "batch stream creation failed");
            }
        }
    }
    for (size_t i = 0; i < b->plans.size(); i++) b->obs_gen[i] = b->plans[i]->obs_gen;
    b->prepared = true;
    return P9IDG_OK;
}

/* what: 1 = grid, 2 = degrid, 3 = both */
int p9idg_batch_exec(p9idg_batch* b, int32_t what) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    DeviceGuard dg(b->device);
    if (int st = batch_check(b)) return st;
    if (int st = batch_order_in(b)) return st;
    if (int st = batch_run(b, what, false)) return st;
    return batch_order_out(b);
}

int p9idg_batch_capture(p9idg_batch* b, int32_t what) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    DeviceGuard dg(b->device);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    if (int st = batch_check(b)) return st;
    if (int st = batch_order_in(b)) return st;
    if (int st = batch_run(b, what, false)) return st;  // warm: occupancy figures, first launches
    P9_HIP(hipStreamSynchronize(b->stream));
    P9_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = batch_run(b, what, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return P9IDG_OK;
}

int p9idg_batch_launch(p9idg_batch* b) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    if (!b->graph_exec) return fail(P9IDG_ERR_STATE, "no captured batch graph");
    DeviceGuard dg(b->device);
    if (int st = batch_check(b)) return st;
    if (int st = batch_order_in(b)) return st;
    P9_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    return batch_order_out(b);
}

int p9idg_batch_sync(p9idg_batch* b) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    DeviceGuard dg(b->device);
    P9_HIP(hipStreamSynchronize(b->stream));
    return P9IDG_OK;
}

/* kernel-only times of the last eager p9idg_batch_exec(b, 3) */
int p9idg_batch_timings(p9idg_batch* b, float* grid_ms, float* degrid_ms) {
    if (!b) return fail(P9IDG_ERR_ARGUMENT, "batch is NULL");
    if (!b->timed) return fail(P9IDG_ERR_STATE, "no timed batch run");
    DeviceGuard dg(b->device);
    P9_HIP(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    if (grid_ms) {
        P9_HIP(hipEventElapsedTime(&ms, b->ev[1], b->ev[2]));
        *grid_ms = ms;
    }
    if (degrid_ms) {
        P9_HIP(hipEventElapsedTime(&ms, b->ev[3], b->ev[4]));
        *degrid_ms = ms;
    }
    return P9IDG_OK;
}

int p9idg_gridded_visibilities(p9idg_plan* p, int64_t* n) {
    if (int st = check_obs(p)) return st;
    if (!n) return fail(P9IDG_ERR_ARGUMENT, "NULL argument");
    *n = p->gridded_vis;
    return P9IDG_OK;
}

}  // extern "C"
