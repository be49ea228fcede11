"""Pins the oracle's FFT steps (FFTW conventions) against numpy.fft.

FFTW FORWARD is exp(-2 pi i jk/n) and BACKWARD exp(+2 pi i jk/n), both
unnormalised; SubgridFFT.cs:37-43 uses BACKWARD times 1/S^2 and
SubgridFFT.cs:75 FORWARD without normalisation; FFT.cs:99-215 are the G x G
steps; FFT.Shift (FFT.cs:248-264) swaps quadrants with n2 = rows / 2.
"""
import numpy as np
import pytest

from conftest import rel_err


@pytest.mark.parametrize("shape", [(8, 8), (24, 24), (32, 32), (48, 48), (64, 64), (256, 256), (12, 15), (7, 5), (96, 40)])
def test_fft2_matches_numpy(O, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = shape[0] * shape[1]
    assert rel_err(O.fft2(a, -1), np.fft.fft2(a)) < 1e-13
    assert rel_err(O.fft2(a, +1), np.fft.ifft2(a) * n) < 1e-13


@pytest.mark.parametrize("S", [8, 16, 24, 32, 48, 64])
def test_subgrid_fft_conventions(O, S):
    rng = np.random.default_rng(S)
    c = O.make_constants(1, 4 * S, S, 4, 16, 1e-5, 1, 0.0)
    sub = rng.standard_normal((3, S, S)) + 1j * rng.standard_normal((3, S, S))
    # Forward: FFTW BACKWARD (e^+) times 1/S^2 == numpy ifft2
    assert rel_err(O.subgrid_fft_forward(c, sub), np.fft.ifft2(sub)) < 1e-13
    # Backward: FFTW FORWARD (e^-), unnormalised == numpy fft2
    assert rel_err(O.subgrid_fft_backward(c, sub), np.fft.fft2(sub)) < 1e-13


@pytest.mark.parametrize("n", [4, 8, 256])
def test_shift_even_is_fftshift(O, n):
    a = np.arange(n * n, dtype=np.float64).reshape(n, n)
    b = O.shift(a.copy())
    assert np.array_equal(b, np.fft.fftshift(a))
    c = (a + 1j * a[::-1]).astype(np.complex128)
    assert np.array_equal(O.shift(c.copy()), np.fft.fftshift(c))
    f = a.astype(np.float32)
    assert np.array_equal(O.shift(f.copy()), np.fft.fftshift(f))


def test_shift_odd_leaves_last_row_and_column(O):
    # n2 = 1 for n = 3: only [0,0]<->[1,1] and [1,0]<->[0,1] swap
    a = np.arange(9, dtype=np.float64).reshape(3, 3)
    expected = np.array([[4, 3, 2], [1, 0, 5], [6, 7, 8]], dtype=np.float64)
    assert np.array_equal(O.shift(a.copy()), expected)


def test_wstack_ifft_float_sums_layers(O):
    rng = np.random.default_rng(5)
    L, G, N = 3, 64, 12345
    layers = rng.standard_normal((L, G, G)) + 1j * rng.standard_normal((L, G, G))
    out = O.wstack_ifft_float(layers, N)
    expected = np.zeros((G, G), dtype=np.float32)
    for k in range(L):
        expected += (np.real(np.fft.ifft2(layers[k]) * G * G) / N).astype(np.float32)
    assert out.dtype == np.float32
    assert np.abs(out - expected).max() <= 4 * np.finfo(np.float32).eps * np.abs(expected).max()


def test_fft_backward_and_forward_float(O):
    rng = np.random.default_rng(6)
    G, N = 48, 777
    grid = rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))
    assert rel_err(O.fft_backward(grid, N), np.real(np.fft.ifft2(grid) * G * G) / N) < 1e-13
    img = rng.standard_normal((G, G)).astype(np.float32)
    assert rel_err(O.fft_forward_float(img, 2.0), np.fft.fft2(img.astype(np.float64)) / 2.0) < 1e-13


def test_substract(O):
    rng = np.random.default_rng(7)
    v0 = rng.standard_normal((4, 5, 3)) + 1j * rng.standard_normal((4, 5, 3))
    v1 = rng.standard_normal((4, 5, 3)) + 1j * rng.standard_normal((4, 5, 3))
    flags = rng.random((4, 5, 3)) < 0.3
    out = O.substract(v0, v1, flags)
    assert np.array_equal(out[flags], np.zeros(flags.sum(), dtype=np.complex128))
    assert np.array_equal(out[~flags], (v0 - v1)[~flags])


def test_oracle_sincos_equals_sin_and_cos(O):
    # the oracle evaluates Math.Cos/Math.Sin pairs with sincos
    assert O.sincos_mismatches(2_000_000, 7) == 0
