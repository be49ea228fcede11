"""Pins the oracle's gridder, degridder, adders and image steps with closed
forms and identities that the mathematics fixes (the reference ships no tests
or fixtures for this path).

KAT-1  one visibility at an integer uv pixel with w = 0: the dirty image is
       the closed-form fringe Re(V e^{2 pi i (p_u (x-G/2) + p_v (y-G/2)) / G}) / N.
KAT-2  visibilities V = exp(i k (u l + v m + w n)) of a point source at the
       mirror of a subgrid pixel centre (l, m), any u, v, w: at that pixel the
       gridder's phasor cancels the source phase exactly, so the subgrid value is
       N_vis * e^{i (uoff l + voff m)}; pins the signs of u, v and w and the
       w-term.  n is the reference's ComputeN, which evaluates
       tmp / 1.0f + Sqrt(1 - tmp) (Gridder.cs:172-173): n = t / (t + sqrt(1 - t))
       with t = l^2 + m^2, about twice the 1 - sqrt(1 - t) of the measurement
       equation; the test uses that n, so a kernel using the textbook n fails.
KAT-3  a noise-free point source peaks at its pixel in the dirty image.
Adjoint identities between gridder/degridder, adder/splitter and Grid/DeGrid,
and the w-stack mirror identities (a negative-w subgrid is the Hermitian
counterpart of the positive one) pin the remaining conventions.
"""
import math

import numpy as np
import pytest

from conftest import rel_err

C_LIGHT = 299792458.0


def kat_constants(O, G=256, S=32, K=8, L=1, wstep=0.0, scale=2.0 ** -10, vis_count=1):
    return O.make_constants(vis_count, G, S, K, 1024, scale, L, wstep)


def image_size(c):
    return float(np.float32(np.float32(c.scale_arcsec) * np.float32(c.grid_size)))


def compute_l(x, S, img):
    t = np.float32(np.float32(x - S // 2) + np.float32(0.5))
    t = np.float32(t * np.float32(img))
    return np.float32(t / np.float32(S))


@pytest.mark.parametrize("pu,pv", [(5, -7), (-20, 13), (0, 0), (31, 2)])
def test_kat1_single_visibility_fringe(O, pu, pv):
    G = 256
    c = kat_constants(O, G=G)
    img = image_size(c)
    freq = np.array([C_LIGHT])  # f / c = 1 exactly
    uvw = np.zeros((1, 1, 3))
    uvw[0, 0, 0] = pu / img
    uvw[0, 0, 1] = pv / img
    meta, _ = O.partition(c, uvw, freq)
    V = np.array([[[0.3 + 0.4j]]])
    out = O.to_image(c, meta, V, uvw, freq)
    y, x = np.mgrid[0:G, 0:G]
    expected = np.real(V[0, 0, 0] * np.exp(2j * np.pi * (pu * (x - G / 2) + pv * (y - G / 2)) / G))
    assert np.abs(out - expected).max() < 1e-12
    # the gridded (shifted) grid holds exactly V at the uv pixel, ~0 elsewhere
    grid = O.grid(c, meta, V, uvw, freq)
    assert abs(grid[pv % G, pu % G] - V[0, 0, 0]) < 1e-12
    grid[pv % G, pu % G] = 0
    assert np.abs(grid).max() < 1e-12


@pytest.mark.parametrize("x0,y0,wscale", [(3, 27, 0.0), (20, 9, 300.0), (31, 0, 3000.0), (16, 16, 30000.0)])
def test_kat2_gridder_cancels_source_phase(O, x0, y0, wscale):
    rng = np.random.default_rng(x0 * 100 + y0)
    G, S = 512, 32
    c = kat_constants(O, G=G, S=S, scale=1.5 / 3600 * math.pi / 180)
    img = image_size(c)
    T, C = 6, 5
    freq = 1.0e9 + 0.05e9 * np.arange(C)
    uvw = np.zeros((1, T, 3))
    uvw[0, :, 0] = rng.uniform(-3000, 3000, T)
    uvw[0, :, 1] = rng.uniform(-3000, 3000, T)
    uvw[0, :, 2] = rng.uniform(-1, 1, T) * wscale
    lf, mf = compute_l(x0, S, img), compute_l(y0, S, img)
    t = np.float32(lf * lf + mf * mf)
    n = float(np.float32(t / np.float32(float(t) + math.sqrt(1.0 - float(t)))))
    l, m = float(lf), float(mf)
    k = 2 * math.pi * freq / C_LIGHT
    d = uvw[0, :, 0, None] * l + uvw[0, :, 1, None] * m + uvw[0, :, 2, None] * n
    vis = np.exp(1j * k[None, :] * d)[None]
    U, Vp = 200, 240  # any corner; the gridder does not check spans
    meta = np.zeros(1, dtype=O.SUBGRID_DTYPE)
    meta[0] = (0, 0, T, U, Vp, 0)
    sub = O.gridder_forward(c, meta, uvw, vis, freq)[0]
    uoff = (U + S // 2 - G // 2) * (2 * math.pi / img)
    voff = (Vp + S // 2 - G // 2) * (2 * math.pi / img)
    expected = T * C * np.exp(1j * (uoff * l + voff * m))
    got = sub[(y0 + S // 2) % S, (x0 + S // 2) % S]
    assert abs(got - expected) / (T * C) < 1e-6


def test_kat3_point_source_peaks_at_its_pixel(O):
    from p9idg import observation as obs

    cell = 1.5
    ob = obs.custom(G=256, S=32, K=8, L=1, n_ant=16, T=30, C=4, f0=1.0e9, f1=1.05e9,
                    core_diam=10000.0, cell_arcsec=cell, sources=[(1.0, 17 * obs.ARCSEC * cell, -23 * obs.ARCSEC * cell, 0.0)],
                    noise=0.0, seed=3)
    c = ob.c
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                          float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    meta, _ = O.partition(oc, ob.uvw, ob.freq)
    n_grid = int(meta["timeSampleCount"].sum()) * len(ob.freq)
    assert n_grid > 0.5 * ob.vis.size
    img = O.to_image(oc, meta, ob.vis, ob.uvw, ob.freq)
    iy, ix = np.unravel_index(np.argmax(img), img.shape)
    assert (iy, ix) == (128 - 23, 128 + 17)
    # peak = gridded fraction of the flux (IDG's interpolation error is small)
    assert abs(img[iy, ix] - n_grid / ob.vis.size) < 0.05


def _random_meta(O, rng, n, G, S, B, T, wlo=0, whi=1):
    meta = np.zeros(n, dtype=O.SUBGRID_DTYPE)
    for i in range(n):
        t0 = int(rng.integers(0, T - 1))
        nt = int(rng.integers(1, T - t0 + 1))
        meta[i] = (int(rng.integers(0, B)), t0, nt, int(rng.integers(2, G - S - 1)), int(rng.integers(2, G - S - 1)),
                   int(rng.integers(wlo, whi)))
    return meta


@pytest.mark.parametrize("S", [8, 24, 32])
def test_gridder_degridder_adjoint(O, S):
    rng = np.random.default_rng(S)
    G, B, T, C = 128, 3, 7, 4
    c = kat_constants(O, G=G, S=S, K=2, scale=4e-5)
    freq = 1.2e9 + 1e7 * np.arange(C)
    uvw = rng.uniform(-800, 800, (B, T, 3))
    meta = _random_meta(O, rng, 4, G, S, B, T)
    meta["baselineIdx"] = np.arange(4) % B
    # disjoint time ranges per baseline so every visibility belongs to one subgrid
    meta["timeSampleStart"] = [0, 0, 0, 4]
    meta["timeSampleCount"] = [4, 7, 7, 3]
    meta["baselineIdx"] = [0, 1, 2, 0]
    vis = rng.standard_normal((B, T, C)) + 1j * rng.standard_normal((B, T, C))
    data = rng.standard_normal((4, S, S)) + 1j * rng.standard_normal((4, S, S))
    sub = O.gridder_forward(c, meta, uvw, vis, freq)
    back = O.gridder_backward(c, meta, data, uvw, freq)
    scale = float(np.float32(1.0) / np.float32(S * S))
    lhs = np.vdot(sub, data)
    rhs = np.vdot(vis, back) / scale
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_adder_split_adjoint_and_invalid_grid(O):
    rng = np.random.default_rng(11)
    G, S = 96, 16
    c = kat_constants(O, G=G, S=S, K=2)
    meta = _random_meta(O, rng, 5, G, S, 1, 2)
    sub = rng.standard_normal((5, S, S)) + 1j * rng.standard_normal((5, S, S))
    grid = rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))
    lhs = np.vdot(O.adder_add(c, meta, sub), grid)
    rhs = np.vdot(sub, O.adder_split(c, meta, grid))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    # Adder.cs:30-31,49: U == G - S is rejected
    bad = meta.copy()
    bad[2]["UPixel"] = G - S
    with pytest.raises(O.OracleError, match="invalid grid"):
        O.adder_add(c, bad, sub)
    # Split leaves an out-of-range subgrid empty
    assert np.abs(O.adder_split(c, bad, grid)[2]).max() == 0


def test_subgrid_fft_adjoint(O):
    rng = np.random.default_rng(12)
    S = 32
    c = kat_constants(O, G=128, S=S)
    a = rng.standard_normal((2, S, S)) + 1j * rng.standard_normal((2, S, S))
    b = rng.standard_normal((2, S, S)) + 1j * rng.standard_normal((2, S, S))
    lhs = np.vdot(O.subgrid_fft_forward(c, a), b)
    rhs = np.vdot(a, O.subgrid_fft_backward(c, b)) / (S * S)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


@pytest.mark.parametrize("S", [16, 32])
def test_grid_degrid_adjoint(O, S):
    """DeGrid = Grid^H when 1/S^2 is exact (Gridder.cs:151 uses a float)."""
    rng = np.random.default_rng(13)
    G, B, T, C = 128, 2, 6, 3
    c = kat_constants(O, G=G, S=S, K=2, scale=4e-5)
    freq = 1.0e9 + 2e7 * np.arange(C)
    uvw = rng.uniform(-900, 900, (B, T, 3))
    meta = np.zeros(3, dtype=O.SUBGRID_DTYPE)
    meta[0] = (0, 0, 4, 20, 30, 0)
    meta[1] = (0, 4, 2, 50, 41, 0)
    meta[2] = (1, 0, 6, 33, 70, 0)
    vis = rng.standard_normal((B, T, C)) + 1j * rng.standard_normal((B, T, C))
    X = rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))
    lhs = np.vdot(O.grid(c, meta, vis, uvw, freq), X)
    rhs = np.vdot(vis, O.degrid(c, meta, X, uvw, freq))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_wstack_layer_equals_adder_for_positive_w(O):
    rng = np.random.default_rng(14)
    G, S, L = 96, 16, 3
    c = kat_constants(O, G=G, S=S, K=2, L=L)
    meta = _random_meta(O, rng, 4, G, S, 1, 2, 0, L)
    sub = rng.standard_normal((4, S, S)) + 1j * rng.standard_normal((4, S, S))
    layers = O.adderw_add(c, meta, sub)
    for w in range(L):
        sel = meta["WLambda"] == w
        expected = O.adder_add(c, meta[sel], sub[sel]) if sel.any() else np.zeros((G, G), complex)
        assert np.abs(layers[w] - expected).max() < 1e-12
    grid = rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))
    assert np.abs(O.adderw_split(c, meta, grid) - O.adder_split(c, meta, grid)).max() < 1e-12


def test_wstack_negative_w_is_hermitian_mirror(O):
    """A subgrid with w-index -1-k lands mirrored and conjugated in layer k
    (AdderWStack.cs:41-70): the inverse FFT's real part is unchanged."""
    rng = np.random.default_rng(15)
    G, S, L = 128, 16, 2
    c = kat_constants(O, G=G, S=S, K=2, L=L, vis_count=1000)
    meta = _random_meta(O, rng, 3, G, S, 1, 2, 0, L)
    sub = rng.standard_normal((3, S, S)) + 1j * rng.standard_normal((3, S, S))
    neg = meta.copy()
    neg["WLambda"] = -1 - meta["WLambda"]
    img_pos, img_neg = [], []
    for m in (meta, neg):
        layers = O.adderw_add(c, m, sub)
        for k in range(L):
            O.shift(layers[k])
        (img_pos if m is meta else img_neg).append(O.wstack_ifft_float(layers, 1000))
    assert np.abs(img_pos[0] - img_neg[0]).max() <= 1e-6 * np.abs(img_pos[0]).max()
    assert np.abs(img_pos[0]).max() > 0


def test_wstack_split_negative_w_on_hermitian_grid(O):
    """For the (shifted) grid of a real image, AdderWStack.Split's mirrored,
    conjugated read equals the direct read (AdderWStack.cs:120-130)."""
    rng = np.random.default_rng(16)
    G, S = 128, 16
    c = kat_constants(O, G=G, S=S, K=2, L=2)
    meta = _random_meta(O, rng, 3, G, S, 1, 2, 0, 2)
    img = rng.standard_normal((G, G))
    grid = np.fft.fft2(img)
    O.shift(grid)  # DeGridW's in-place shift (IDG.cs:44)
    neg = meta.copy()
    neg["WLambda"] = -1 - meta["WLambda"]
    a = O.adderw_split(c, meta, grid)
    b = O.adderw_split(c, neg, grid)
    assert rel_err(b, a) < 1e-12


def test_gridw_image_equals_plain_image(O):
    """GridW -> WStackIFFTFloat -> Shift equals ToImage for one layer (float32)."""
    from p9idg import observation as obs

    ob = obs.small()
    c = ob.c
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                          float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    meta, _ = O.partition(oc, ob.uvw, ob.freq)
    a = O.gridw_to_image(oc, meta, ob.vis, ob.uvw, ob.freq)
    b = O.to_image(oc, meta, ob.vis, ob.uvw, ob.freq)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
