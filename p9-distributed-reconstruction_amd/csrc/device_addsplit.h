// Adder (scatter of a subgrid onto the grid / w-layers) and splitter
// (gather of a subgrid from the grid), shared by the FP32 and MFMA kernels.
//   Adder.Add        Adder.cs:13-55       AdderWStack.Add    AdderWStack.cs:11-78
//   Adder.Split      Adder.cs:57-95       AdderWStack.Split  AdderWStack.cs:81-139
// The grid is addressed in its FFT.Shift-ed layout (FFT.cs:221-264).
#pragma once

#include "device_common.h"

namespace p9idg {
namespace dev {

// Adds the (normalised) subgrid tile onto the grid.  `tile` is the subgrid
// after the inverse FFT, indexed [y][x] in FFT order.  With a.scratch set the
// values that would be added at grid pixel (Y + y, X + x) are stored at
// scratch[slot][y][x] instead (Y, X: the corner after the negative-w mirror),
// for the ordered sum of gather_slots.  ap: the [S][S] adder phasors
// (a.adder_phasor or a copy of it in LDS).
template <int S>
__device__ __forceinline__ void adder(const KernelArgs& a, const SubgridGeom& g,
                                      const float2* tile, float norm, int slot, const float2* ap) {
    const int G = a.G;
    float2* const out = a.scratch ? a.scratch + (size_t)slot * S * S : nullptr;
    if (a.mode == P9IDG_GRID_PLAIN) {
        // Adder.cs:22-47 (range validated on the host: 0 <= U,V < G-S)
        for (int p = item_tid(); p < S * S; p += kThreads) {
            const int y = p / S, x = p - (p / S) * S;
            const int ys = (y + S / 2) % S, xs = (x + S / 2) % S;
            float2 v = cmulf(tile[ys * S + xs], ap[p]);
            v.x *= norm;
            v.y *= norm;
            if (out) {
                out[p] = v;
                continue;
            }
            float2* dst = a.grid + shift_index(g.V + y, g.U + x, G);
            unsafeAtomicAdd(&dst->x, v.x);
            unsafeAtomicAdd(&dst->y, v.y);
        }
        return;
    }
    // AdderWStack.cs:29-74: negative w layers are mirrored and conjugated.
    const bool neg = g.W < 0;
    int X = g.U, Y = g.V, W = g.W;
    if (neg) {
        X = G - X - S + 1;
        Y = G - Y - S + 1;
        W = -W - 1;
    }
    if (!(X >= 1 && X < G - S && Y >= 1 && Y < G - S)) return;
    if (W < 0 || W >= a.L) return;  // metadata validated on the host
    float2* layer = a.grid + (a.mode == P9IDG_GRID_WSUM ? 0 : (size_t)W * G * G);
    for (int p = item_tid(); p < S * S; p += kThreads) {
        const int y = p / S, x = p - (p / S) * S;
        const int y_ = neg ? S - 1 - y : y;
        const int x_ = neg ? S - 1 - x : x;
        const int ys = (y_ + S / 2) % S, xs = (x_ + S / 2) % S;
        float2 v = cmulf(ap[y_ * S + x_], tile[ys * S + xs]);
        v.x *= norm;
        v.y *= norm;
        if (neg) v.y = -v.y;
        if (out) {
            out[p] = v;
            continue;
        }
        float2* dst = layer + shift_index(Y + y, X + x, G);
        unsafeAtomicAdd(&dst->x, v.x);
        unsafeAtomicAdd(&dst->y, v.y);
    }
}
template <int S>
__device__ __forceinline__ void adder(const KernelArgs& a, const SubgridGeom& g,
                                      const float2* tile, float norm, int slot) {
    adder<S>(a, g, tile, norm, slot, a.adder_phasor);
}

// Reads the subgrid from the model grid into `tile` (FFT order), zeros where
// the reference leaves the subgrid empty.
template <int S>
__device__ __forceinline__ void splitter(const KernelArgs& a, const SubgridGeom& g,
                                         float2* tile) {
    const int G = a.G;
    if (a.mode == P9IDG_DEGRID_PLAIN) {
        // Adder.cs:65-90
        const bool ok = g.U >= 0 && g.U < G - S && g.V >= 0 && g.V < G - S;
        for (int p = item_tid(); p < S * S; p += kThreads) {
            const int y = p / S, x = p - (p / S) * S;
            const int yd = (y + S / 2) % S, xd = (x + S / 2) % S;
            float2 v = make_float2(0.f, 0.f);
            if (ok) {
                float2 gv = a.model_grid[shift_index(g.V + y, g.U + x, G)];
                v = cmulf(gv, conjf2(a.adder_phasor[p]));
            }
            tile[yd * S + xd] = v;
        }
        return;
    }
    // AdderWStack.cs:97-134: one grid; negative w reads the mirrored pixel
    // conjugated; the range test uses the unmirrored corner.
    const bool neg = g.W < 0;
    const int X = g.U, Y = g.V;
    const bool ok = X >= 1 && X < G - S && Y >= 1 && Y < G - S;
    for (int p = item_tid(); p < S * S; p += kThreads) {
        const int y = p / S, x = p - (p / S) * S;
        const int yd = (y + S / 2) % S, xd = (x + S / 2) % S;
        float2 v = make_float2(0.f, 0.f);
        if (ok) {
            const int xs = neg ? G - X - x : X + x;
            const int ys = neg ? G - Y - y : Y + y;
            float2 gv = a.model_grid[shift_index(ys, xs, G)];
            if (neg) gv.y = -gv.y;
            v = cmulf(conjf2(a.adder_phasor[p]), gv);
        }
        tile[yd * S + xd] = v;
    }
}

}  // namespace dev
}  // namespace p9idg
