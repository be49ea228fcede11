"""Pins the oracle's restatement of the serial major cycle (oracle_cycle.cpp)
against independent forms: brute-force sums for PSF.CalcAMap, direct
correlation for PaddedConvolver.Convolve and PSF.CalcPSFSquared, the
closed-form solution of FastSerialCD for a one-pixel PSF, and the identity
that FastSerialCD's gradients stay b - psf2 (x) x."""
import numpy as np
import pytest

import p9idg
from p9idg import observation as obs

from conftest import oracle_constants


def corr_direct(img, psf):
    """out[y, x] = sum_rs psf[r, s] img[y - ph/2 + r, x - pw/2 + s] (zero outside)."""
    H, W = img.shape
    ph, pw = psf.shape
    out = np.zeros((H, W))
    pad = np.zeros((H + 2 * ph, W + 2 * pw))
    pad[ph:ph + H, pw:pw + W] = img
    for r in range(ph):
        for s in range(pw):
            y0, x0 = ph - ph // 2 + r, pw - pw // 2 + s
            out += psf[r, s] * pad[y0:y0 + H, x0:x0 + W]
    return out


def test_cut_and_sidelobe(O):
    rng = np.random.default_rng(1)
    psf = rng.standard_normal((16, 16)).astype(np.float32)
    cut = O.psf_cut(psf, 4)
    assert np.array_equal(cut, psf[6:10, 6:10])  # offset 16/2 - 4/2
    outside = np.abs(psf).copy()
    outside[6:10, 6:10] = 0
    assert O.psf_max_sidelobe(psf, 4) == np.float32(outside.max())
    assert O.psf_max_lipschitz(psf) == pytest.approx(float(np.sum((psf * psf).astype(np.float64))), rel=1e-15)


@pytest.mark.parametrize("n,factor", [(16, 4), (17, 2), (33, 3), (12, 1), (9, 9)])
def test_native_psf_host_helpers_match_oracle(O, n, factor):
    """PSF.Cut, CalcMaxSidelobe and CalcMaxLipschitz of libp9idg (host code,
    no GPU) against the oracle, bit for bit, odd and even sizes."""
    rng = np.random.default_rng(n * 10 + factor)
    psf = rng.standard_normal((n, n)).astype(np.float32)
    assert np.array_equal(p9idg.PSF.Cut(psf, factor), O.psf_cut(psf, factor))
    assert p9idg.PSF.CalcMaxSidelobe(psf, factor) == O.psf_max_sidelobe(psf, factor)
    assert p9idg.PSF.CalcMaxLipschitz(psf) == O.psf_max_lipschitz(psf)
    with pytest.raises(p9idg.P9IDGError) as e:
        p9idg.PSF.Cut(psf, 0)
    assert e.value.status == p9idg._lib.ERR_ARGUMENT
    with pytest.raises(p9idg.P9IDGError) as e:
        p9idg.PSF.CalcMaxLipschitz(np.zeros((0, 4), dtype=np.float32))
    assert e.value.status == p9idg._lib.ERR_ARGUMENT


@pytest.mark.parametrize("ph,pw,H,W", [(6, 6, 20, 20), (5, 5, 17, 17), (8, 4, 19, 23)])
def test_amap_is_psf_energy_inside_the_image(O, ph, pw, H, W):
    rng = np.random.default_rng(ph * 7 + W)
    psf = rng.standard_normal((ph, pw)).astype(np.float32)
    amap = O.psf_amap(psf, H, W)
    sq = (psf.astype(np.float64)) ** 2
    ref = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            i0, j0 = y - ph // 2, x - pw // 2
            ys = slice(max(0, -i0), min(ph, H - i0))
            xs = slice(max(0, -j0), min(pw, W - j0))
            ref[y, x] = sq[ys, xs].sum()
    assert np.allclose(amap, ref.astype(np.float32), rtol=2e-6, atol=0)


@pytest.mark.parametrize("H,ph", [(24, 8), (32, 16), (30, 10)])
def test_padded_correlation_convolve_is_direct_correlation(O, H, ph):
    rng = np.random.default_rng(H + ph)
    img = rng.standard_normal((H, H)).astype(np.float32)
    psf = rng.standard_normal((ph, ph)).astype(np.float32)
    k = O.psf_padded_correlation(psf, H, H)
    out = O.padded_convolve(k, ph, ph, img)
    ref = corr_direct(img.astype(np.float64), psf.astype(np.float64))
    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("ph", [4, 8, 12])
def test_psf_squared_is_autocorrelation(O, ph):
    rng = np.random.default_rng(ph)
    psf = rng.standard_normal((ph, ph)).astype(np.float32)
    p2 = O.psf_squared(psf)
    assert p2.shape == (2 * ph, 2 * ph)
    p = psf.astype(np.float64)
    ref = np.zeros((2 * ph, 2 * ph))
    for y in range(2 * ph):
        for x in range(2 * ph):
            dy, dx = y - ph, x - ph
            for r in range(ph):
                for s in range(ph):
                    if 0 <= r + dy < ph and 0 <= s + dx < ph:
                        ref[y, x] += p[r, s] * p[r + dy, s + dx]
    assert np.abs(p2 - ref).max() <= 1e-6 * np.abs(ref).max()
    assert p2[ph, ph] == np.float32((p * p).sum())


def test_proximal_operator(O):
    for x, lip, lam, a in [(3.0, 2.0, 0.5, 0.2), (-1.0, 1.0, 1.0, 0.01), (0.05, 0.7, 1.0, 0.1)]:
        f = np.float32
        num = max(f(x) - f(lam) * f(a), f(0))
        ref = f(num) / (f(lip) + f(lam) * (f(1) - f(a)))
        assert O.proximal(x, lip, lam, a) == np.float32(ref)


def test_cd_one_pixel_psf_is_closed_form(O):
    rng = np.random.default_rng(3)
    H = W = 12
    # even size: PSF.CalcPSFSquared centres the zero lag at (ph, pw), where
    # UpdateGradients reads it (half = psf2 size / 2)
    psf = np.zeros((2, 2), np.float32)
    psf[1, 1] = 1.0
    psf2 = O.psf_squared(psf)
    delta = np.zeros((4, 4), np.float32)
    delta[2, 2] = 1.0
    assert np.abs(psf2 - delta).max() < 1e-12  # FFT round-off elsewhere
    psf2 = delta
    amap = O.psf_amap(psf, H, W)
    assert np.all(amap == 1)
    g0 = rng.standard_normal((H, W)).astype(np.float32)
    lam, alpha = np.float32(0.5), np.float32(0.2)
    x, g, conv, it = O.cd_deconvolve(amap, psf2, np.zeros((H, W), np.float32), g0, lam, alpha, 10000, 1e-7)
    f = np.float32
    xs = np.maximum(g0 - lam * alpha, f(0)) / (f(1) + lam * (f(1) - alpha))
    assert conv and it == np.count_nonzero(xs)
    assert np.array_equal(x, xs)
    assert np.array_equal(g, (g0 - xs).astype(np.float32))


def test_cd_gradients_track_the_model(O):
    rng = np.random.default_rng(4)
    H = W = 40
    psf = rng.standard_normal((8, 8)).astype(np.float32) * 0.2
    psf[4, 4] = 1.0
    psf2 = O.psf_squared(psf)
    amap = O.psf_amap(psf, H, W)
    g0 = rng.standard_normal((H, W)).astype(np.float32)
    x, g, conv, it = O.cd_deconvolve(amap, psf2, np.zeros((H, W), np.float32), g0, 0.3, 0.1, 400, 1e-6)
    assert it > 0
    # g = g0 - sum over updated pixels of psf2 centred there times x
    ref = g0.astype(np.float64)
    yh = psf2.shape[0] // 2
    for y0, x0 in zip(*np.nonzero(x)):
        for i in range(max(y0 - yh, 0), min(y0 - yh + psf2.shape[0], H)):
            for j in range(max(x0 - yh, 0), min(x0 - yh + psf2.shape[1], W)):
                ref[i, j] -= psf2[i + yh - y0, j + yh - x0] * float(x[y0, x0])
    assert np.abs(g - ref).max() <= 1e-4 * np.abs(ref).max()
    assert np.all(x >= 0)
    if conv:
        assert O.cd_abs_max_diff(amap, x, g, 0.3, 0.1) < 1e-6


def test_penalties(O):
    rng = np.random.default_rng(5)
    img = rng.standard_normal((9, 7)).astype(np.float32)
    assert O.residuals_get_max(img) == max(np.float32(0), img.max())
    assert O.residuals_get_max(-np.abs(img)) == 0
    assert O.residuals_penalty(img) == pytest.approx(0.5 * float(np.sum((img * img).astype(np.float64))), rel=1e-15)
    f = np.float32
    lam, a = f(0.7), f(0.05)
    pen = f(0.5) * (f(1) - a) * (img * img) + a * np.abs(img)
    assert O.elastic_penalty(img, lam, a) == pytest.approx(float(np.sum(float(lam) * pen.astype(np.float64))), rel=1e-12)


def test_reconstruct_recovers_a_point_source(O):
    # 30" cells: the 600 m array resolves the source (a few pixels per beam)
    cell = obs.ARCSEC * 30.0
    ob = obs.custom(G=64, S=16, K=4, L=1, n_ant=8, T=24, C=4, seed=91, core_diam=600.0, cell_arcsec=30.0,
                    sources=[(1.0, 6 * cell, -4 * cell, 0.0)], noise=0.0, wstep=0.0)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    x, stats = O.reconstruct_serial_cd(oc, part.subgrids, ob.vis, None, ob.uvw, ob.freq, 4, 3, 1.0, 0.01, 2000, 1e-5)
    assert len(stats) >= 1
    dirty = O.gridw_to_image(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
    assert stats[0]["max_dirty"] == O.residuals_get_max(dirty)
    assert np.all(x >= 0)
    y, xx = np.unravel_index(np.argmax(x), x.shape)
    assert (y, xx) == (32 - 4, 32 + 6)


@pytest.mark.parametrize("lam", [1.0, 200.0])
def test_reconstruct_lambda_rule(O, lam):
    """Every cycle's lambda follows MajorCycle.cs:98-100, recomputed here from
    the cycle's own max dirty / max gradient and the PSF's cut, sidelobe and
    Lipschitz figures taken with numpy slices: lambda = max(maxB * sidelobe *
    max(maxB / (maxDirty * lip), 1) / alpha, lam * lip) (float32 products and
    casts as in the C#).  On this sky the correction factor is ~1.12, so a
    dropped correction changes lambda by 12%; lam = 200 puts every cycle on
    the Lipschitz floor."""
    cell = obs.ARCSEC * 30.0
    rng = np.random.default_rng(91)
    src = [(1.0, 6 * cell, -4 * cell, 0.0)] + [
        (float(rng.uniform(0.3, 0.8)), float(rng.integers(-3, 4) + 2) * cell, float(rng.integers(-3, 4)) * cell, 0.0)
        for _ in range(8)]
    ob = obs.custom(G=64, S=16, K=4, L=1, n_ant=8, T=24, C=4, seed=91, core_diam=600.0, cell_arcsec=30.0,
                    sources=src, noise=0.01, wstep=0.0)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    m = part.subgrids
    alpha, factor = 0.01, 4
    x, stats = O.reconstruct_serial_cd(oc, m, ob.vis, None, ob.uvw, ob.freq, factor, 4, lam, alpha, 500, 1e-5)
    # lam = 200: the first cycle's proximal steps are all zero and the cycle
    # loop stops there (abs max dx < MAJOR_EPSILON, MajorCycle.cs:104,122-123)
    assert len(stats) == (4 if lam == 1.0 else 1)
    psf = O.gridw_to_image(oc, m, np.ones_like(ob.vis), ob.uvw, ob.freq)
    G = psf.shape[0]
    pc = G // factor
    o = G // 2 - pc // 2
    cut = psf[o:o + pc, o:o + pc]
    outside = np.abs(psf).copy()
    outside[o:o + pc, o:o + pc] = 0
    sidelobe = np.float32(outside.max())
    lip = float(np.sum((cut * cut).astype(np.float64)))
    floor = float(np.float32(float(np.float32(lam)) * lip))
    f32 = np.float32
    corrections = []
    for s in stats:
        corr = max(float(s["max_gradient"]) / (float(s["max_dirty"]) * lip), 1.0)
        corrections.append(corr)
        expect = f32(max(float(f32(s["max_gradient"]) * sidelobe) * corr / float(f32(alpha)), floor))
        assert s["lambda_"] == expect
    if lam == 1.0:
        assert min(corrections) > 1.05 and all(s["lambda_"] > floor for s in stats)
    else:
        assert all(s["lambda_"] == f32(floor) for s in stats)
