"""Partitioner (Partitioner.cs:20-255): hand-derived cases and bit-exact
agreement between the oracle and the product's host partitioner.

Hand cases use G = 256, S = 32, K = 8, scale = 2^-10 rad (ImageSize = 0.25
exactly) and frequencies c and 2c, so MetersToPixels(u) = u/4 for channel 0
and u/2 for the last channel, exactly.  A subgrid accepts a timestep while
max(du, dv) + K < S over its channel-0 and last-channel points, i.e. spans
below 24 pixels; U = round_half_even(centre) + G/2 - S/2 = round(centre) + 112.
"""
import math

import numpy as np
import pytest

import p9idg
from p9idg import _lib

C_LIGHT = 299792458.0
SCALE = 2.0 ** -10


def consts(G=256, S=32, K=8, maxT=1024, L=1, wstep=0.0):
    return p9idg.GriddingConstants(1, G, S, K, maxT, SCALE, L, wstep)


def track(P, Q=None, W=None):
    """uvw for one baseline whose channel-0 pixel coordinates are P (u) and Q (v)."""
    P = np.asarray(P, dtype=np.float64)
    uvw = np.zeros((1, len(P), 3))
    uvw[0, :, 0] = 4.0 * P
    if Q is not None:
        uvw[0, :, 1] = 4.0 * np.asarray(Q, dtype=np.float64)
    if W is not None:
        uvw[0, :, 2] = np.asarray(W, dtype=np.float64)
    return uvw


FREQ2 = np.array([C_LIGHT, 2 * C_LIGHT])


def both(c, uvw, freq):
    """Runs the product partitioner and the oracle; returns rows or the error text."""
    import oracle as O

    res = []
    try:
        part = p9idg.Partitioner.CreatePartition(c, uvw, freq)
        res.append(("ok", [tuple(int(v) for v in r) for r in part.subgrids], list(part.offsets)))
    except p9idg.P9IDGError as e:
        res.append(("err", str(e), e.status))
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                          float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    try:
        meta, offs = O.partition(oc, uvw, freq)
        res.append(("ok", [tuple(int(v) for v in r) for r in meta], list(offs)))
    except O.OracleError as e:
        res.append(("err", str(e), e.code))
    assert res[0] == res[1], res
    return res[0]


def test_dropped_timestep():
    # t0: span 30 px -> channel-0 point taken, last channel rejected, nothing
    # committed -> timestep dropped (Partitioner.cs:80-99); t1..t2 fit.
    st, rows, _ = both(consts(), track([30, 2, 3]), FREQ2)
    assert st == "ok"
    # extent [2, 6] -> centre 4 -> U = 116; v = 0 -> V = 112
    assert rows == [(0, 1, 2, 116, 112, 0)]


def test_max_timesteps_cap():
    st, rows, _ = both(consts(maxT=3), track([2] * 7), FREQ2)
    # extent [2, 4] -> centre 3 -> U = 115
    assert rows == [(0, 0, 3, 115, 112, 0), (0, 3, 3, 115, 112, 0), (0, 6, 1, 115, 112, 0)]


def test_w_index_splits_subgrids():
    # wLambda = w (f/c = 1 at channel 0), WStep 1 -> indices 0,0,1,1,-1
    st, rows, _ = both(consts(L=2, wstep=1.0), track([1] * 5, W=[0.5, 0.7, 1.2, 1.5, -0.5]), FREQ2)
    # extent [1, 2] -> centre 1.5 -> banker's rounding 2 -> U = 114
    assert rows == [(0, 0, 2, 114, 112, 0), (0, 2, 2, 114, 112, 1), (0, 4, 1, 114, 112, -1)]


@pytest.mark.parametrize("P,Q,U,V", [(1, 0, 114, 112), (3, 0, 116, 112), (-1, 0, 110, 112), (5, 0, 120, 112),
                                     (0, 3, 112, 116), (2, -3, 115, 108)])
def test_bankers_rounding(P, Q, U, V):
    # centre (P + 2P)/2 = 1.5 P: 1.5 -> 2, 4.5 -> 4, -1.5 -> -2, 7.5 -> 8
    st, rows, _ = both(consts(), track([P], [Q]), FREQ2)
    assert rows == [(0, 0, 1, U, V, 0)]


def test_orphan_channel0_widens_extent():
    # t1's channel-0 point (12) is accepted before its last channel (24) is
    # rejected; the first subgrid keeps the widened extent [0, 12] -> U = 118.
    st, rows, _ = both(consts(), track([0, 12]), FREQ2)
    assert rows == [(0, 0, 1, 118, 112, 0), (0, 1, 1, 130, 112, 0)]


@pytest.mark.parametrize("P,ok", [(-111, True), (-112, False), (112, True), (113, False), (-115, False)])
def test_in_range_limits(P, ok):
    # single channel: U = P + 112 must lie in [1, G - S] = [1, 224]
    res = both(consts(), track([P]), np.array([C_LIGHT]))
    if ok:
        assert res[1] == [(0, 0, 1, P + 112, 112, 0)]
    else:
        assert res[0] == "err" and res[1] == "subgrid falls not within grid" and res[2] == _lib.ERR_SUBGRID_RANGE


@pytest.mark.parametrize("w,ok", [(-4.0, True), (-4.5, False), (3.9, True), (4.0, False)])
def test_w_layer_limits(w, ok):
    # -L <= floor(w / WStep) < L with L = 4
    res = both(consts(L=4, wstep=1.0), track([0], W=[w]), np.array([C_LIGHT]))
    assert (res[0] == "ok") == ok
    if not ok:
        assert res[1] == "subgrid falls not within grid"


@pytest.mark.parametrize("u,v", [(math.inf, 0.0), (0.0, -math.inf), (math.inf, math.inf)])
def test_infinite_coordinates_abort(u, v):
    uvw = np.array([[[u, v, 0.0]]])
    res = both(consts(), uvw, FREQ2)
    assert res[0] == "err" and res[2] == _lib.ERR_PLACE
    assert res[1] == "could not place (all) visibilities on subgrid (too many channnels, kernel size too large)"


def test_nan_coordinate_falls_outside_grid():
    # NaN extents are accepted by AddVisibility; (int)Math.Round(NaN) is
    # INT32_MIN on the x64 CLR, so InRange fails.
    res = both(consts(), np.array([[[math.nan, 0.0, 0.0]]]), FREQ2)
    assert res[0] == "err" and res[2] == _lib.ERR_SUBGRID_RANGE


def test_kernel_as_wide_as_subgrid_drops_everything():
    st, rows, offs = both(consts(K=32), track([0, 1, 2]), FREQ2)
    assert rows == [] and offs == [0, 0]


def test_baseline_offsets_and_order():
    uvw = np.concatenate([track([2, 2, 40, 2]), track([5, 5]).repeat(1, axis=0)[:, :2].repeat(2, axis=1)], axis=0)
    st, rows, offs = both(consts(), uvw, FREQ2)
    assert [r[0] for r in rows] == sorted(r[0] for r in rows)
    assert offs[0] == 0 and offs[-1] == len(rows)


def test_random_tracks_bit_exact():
    rng = np.random.default_rng(2024)
    for trial in range(120):
        G = int(rng.choice([64, 128, 256, 512]))
        S = int(rng.choice([8, 16, 24, 32]))
        K = int(rng.integers(1, S // 2 + 1))
        L = int(rng.integers(1, 5))
        wstep = float(rng.choice([0.0, 50.0, 300.0]))
        c = p9idg.GriddingConstants(1, G, S, K, int(rng.integers(1, 40)), float(rng.uniform(2e-6, 2e-5)), L, wstep)
        B, T, C = int(rng.integers(1, 4)), int(rng.integers(1, 60)), int(rng.integers(1, 6))
        freq = np.sort(rng.uniform(0.9e9, 1.5e9, C))
        ang = rng.uniform(0, 2 * np.pi, (B, 1))
        r = rng.uniform(0, 2000, (B, 1))
        t = np.linspace(0, rng.uniform(0.05, 1.0), T)[None, :]
        uvw = np.stack([r * np.cos(ang + t), r * np.sin(ang + t), rng.uniform(-500, 500, (B, 1)) * np.cos(t)], axis=-1)
        both(c, np.ascontiguousarray(uvw), freq)


def test_large_workload_partition_bit_exact():
    from p9idg import observation as obs

    enu = obs._antennas(2000, 45, 1000.0, 19, 8000.0)
    uvw = obs._uvw(enu, 2000, 8.0)
    freq = obs._even_freq(1.0e9, 1.4e9, 8)
    c = p9idg.GriddingConstants(uvw.shape[0] * uvw.shape[1] * 8, 3072, 32, 8, 1024, 1.5 * obs.ARCSEC, 32,
                                obs._wstep(uvw, freq, 32))
    st, rows, offs = both(c, uvw, freq)
    assert st == "ok" and len(rows) > 1000
