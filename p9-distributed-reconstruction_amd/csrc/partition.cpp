// Host partitioner: cuts each baseline's time axis into subgrids.
//
// Semantics of Partitioner.CreatePartition (Distributed-Reconstruction/Core/
// ImageDomainGridder/Partitioner.cs:20-124, helper class :154-255), kept
// bit-exact: the same FP64 operations in the same order, C#'s Math.Min/Max
// NaN propagation, banker's rounding for the centre, and the x64 CLR's
// (int)double conversion.  Baselines are independent, so they are partitioned
// on a pool of host threads and concatenated in baseline order; on failure
// the error of the lowest failing baseline is reported (the reference throws
// at the first failing baseline, which is the same one).
#include <algorithm>
#include <atomic>
#include <cfenv>
#include <cmath>
#include <limits>
#include <thread>

#include "p9idg_internal.h"

namespace p9idg {
namespace {

constexpr double kC = 299792458.0;  // MathFunctions.cs:11

inline double min_nan(double a, double b) {
    if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
    if (a == b) return std::signbit(a) ? a : b;
    return b < a ? b : a;
}
inline double max_nan(double a, double b) {
    if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
    if (a == b) return std::signbit(a) ? b : a;
    return b > a ? b : a;
}
inline int32_t clr_int(double x) {
    if (!(x < 2147483648.0 && x > -2147483649.0)) return INT32_MIN;  // NaN included
    return (int32_t)x;
}

// Extent of the subgrid under construction.
struct Window {
    double u_lo, u_hi, v_lo, v_hi;
    int32_t w_index;
    void clear() {
        u_lo = v_lo = std::numeric_limits<double>::infinity();
        u_hi = v_hi = -std::numeric_limits<double>::infinity();
        w_index = 0;
    }
    // AddVisibility (Partitioner.cs:186-221)
    bool take(double u, double v, double w_lambda, const p9idg_constants& c) {
        if (std::isinf(u) || std::isinf(v)) return false;
        const int32_t wi = (c.wstep != 0.0) ? clr_int(std::floor(w_lambda / c.wstep)) : 0;
        if (!std::isinf(u_lo) && w_index != wi) return false;
        const double nu_lo = min_nan(u_lo, u), nu_hi = max_nan(u_hi, u);
        const double nv_lo = min_nan(v_lo, v), nv_hi = max_nan(v_hi, v);
        const double width = max_nan(nu_hi - nu_lo, nv_hi - nv_lo);
        if ((width + (double)c.kernel_size) >= (double)c.subgrid_size) return false;
        u_lo = nu_lo;
        u_hi = nu_hi;
        v_lo = nv_lo;
        v_hi = nv_hi;
        w_index = wi;
        return true;
    }
};

struct BaselineResult {
    std::vector<p9idg_subgrid> subgrids;
    int status = P9IDG_OK;
};

void partition_baseline(const p9idg_constants& c, const double* uvw, int bl, int T,
                        const double* freq, int C, BaselineResult& res) {
    // imagesize is a float product (Partitioner.cs:22), widened to double at
    // the MetersToPixels call (:41-42, :127).
    const double img = (double)(c.scale_arcsec * (float)c.grid_size);
    const double f0 = freq[0] / kC, fm = freq[C - 1] / kC;
    const double* track = uvw + (size_t)bl * T * 3;
    Window win;
    int t = 0;
    while (t < T) {
        win.clear();
        int taken = 0;
        double u0 = 0, v0 = 0;
        for (; t < T; t++) {
            const double* p = track + (size_t)t * 3;
            u0 = (p[0] * img) * f0;
            v0 = (p[1] * img) * f0;
            const double w0 = p[2] * f0;
            const double um = (p[0] * img) * fm;
            const double vm = (p[1] * img) * fm;
            // channel C-1 is tested with channel 0's w (Partitioner.cs:61-62)
            if (!(win.take(u0, v0, w0, c) && win.take(um, vm, w0, c))) break;
            if (++taken == c.max_timesteps_per_subgrid) {
                t++;
                break;
            }
        }
        if (taken == 0) {
            // Partitioner.cs:80-99: infinite coordinates abort, otherwise the
            // timestep that does not fit is skipped.
            if (std::isinf(u0) || std::isinf(v0)) {
                res.status = P9IDG_ERR_PLACE;
                return;
            }
            t++;
            continue;
        }
        // ComputeCoordinates (:236-248): round half to even, then shift the
        // centre to the grid's and the subgrid's top-left corner.
        const int32_t U = clr_int(std::nearbyint((win.u_hi + win.u_lo) / 2)) + c.grid_size / 2 -
                          c.subgrid_size / 2;
        const int32_t V = clr_int(std::nearbyint((win.v_hi + win.v_lo) / 2)) + c.grid_size / 2 -
                          c.subgrid_size / 2;
        // InRange (:223-234)
        if (!(std::min(U, V) >= 1 && std::max(U, V) <= c.grid_size - c.subgrid_size &&
              win.w_index >= -c.wlayer_count && win.w_index < c.wlayer_count)) {
            res.status = P9IDG_ERR_SUBGRID_RANGE;
            return;
        }
        res.subgrids.push_back(p9idg_subgrid{bl, t - taken, taken, U, V, win.w_index});
    }
}

}  // namespace

int partition_host(const p9idg_constants& c, const double* uvw, int B, int T, const double* freq,
                   int C, std::vector<p9idg_subgrid>& out, std::vector<int64_t>& bl_offsets) {
    std::fesetround(FE_TONEAREST);
    std::vector<BaselineResult> per(B);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nthreads = (int)std::min<unsigned>(std::min<unsigned>(hw, 16u), (unsigned)std::max(B, 1));
    std::atomic<int> next{0};
    auto worker = [&]() {
        std::fesetround(FE_TONEAREST);
        for (int b = next++; b < B; b = next++) partition_baseline(c, uvw, b, T, freq, C, per[b]);
    };
    if (nthreads <= 1 || (int64_t)B * T < 4096) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < nthreads; i++) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }
    out.clear();
    bl_offsets.assign((size_t)B + 1, 0);
    for (int b = 0; b < B; b++) {
        if (per[b].status != P9IDG_OK) {
            bl_offsets.resize((size_t)b + 1);
            return per[b].status;
        }
        out.insert(out.end(), per[b].subgrids.begin(), per[b].subgrids.end());
        bl_offsets[(size_t)b + 1] = (int64_t)out.size();
    }
    return P9IDG_OK;
}

}  // namespace p9idg
