"""ctypes binding of the FP64 CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's
CPU-baseline leg import this module as the checker.  The product package
never imports it.

Arrays follow the reference's C# layouts (row-major, contiguous):
uvw float64[B,T,3], freq float64[C], vis complex128[B,T,C],
subgrids complex128[N,S,S], grid complex128[G,G], layers complex128[L,G,G].
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# P9IDG_ORACLE_LIB selects another build (the sanitizer build _build/asan/liboracle.so)
_LIB_PATH = os.environ.get("P9IDG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

SUBGRID_DTYPE = np.dtype(
    [
        ("baselineIdx", "<i4"),
        ("timeSampleStart", "<i4"),
        ("timeSampleCount", "<i4"),
        ("UPixel", "<i4"),
        ("VPixel", "<i4"),
        ("WLambda", "<i4"),
    ]
)

ERR_PLACE, ERR_SUBGRID_RANGE, ERR_INVALID_GRID = 1, 2, 3
MESSAGES = {
    ERR_PLACE: "could not place (all) visibilities on subgrid (too many channnels, kernel size too large)",
    ERR_SUBGRID_RANGE: "subgrid falls not within grid",
    ERR_INVALID_GRID: "invalid grid",
}


class OracleError(Exception):
    def __init__(self, code):
        super().__init__(MESSAGES.get(code, "oracle error %d" % code))
        self.code = code


class Constants(ctypes.Structure):
    _fields_ = [
        ("visibilities_count", ctypes.c_int64),
        ("grid_size", ctypes.c_int32),
        ("subgrid_size", ctypes.c_int32),
        ("kernel_size", ctypes.c_int32),
        ("max_timesteps_per_subgrid", ctypes.c_int32),
        ("scale_arcsec", ctypes.c_float),
        ("wlayer_count", ctypes.c_int32),
        ("wstep", ctypes.c_double),
        ("wstep_lambda", ctypes.c_float),
    ]


def make_constants(vis_count, grid_size, subgrid_size, kernel_size, max_timesteps, scale, wlayers, wstep):
    return Constants(int(vis_count), grid_size, subgrid_size, kernel_size, max_timesteps, scale, wlayers, wstep, 0.0)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def sincos_mismatches(n, seed=1):
    f = lib().orc_sincos_mismatches
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_int64, ctypes.c_uint64]
    return f(n, seed)


def _cc(c):
    return ctypes.byref(c)


def partition(c, uvw, freq):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    B, T = uvw.shape[0], uvw.shape[1]
    n = ctypes.c_int64(0)
    offs = np.zeros(B + 1, dtype=np.int64)
    st = lib().orc_partition(_cc(c), _p(uvw), B, T, _p(freq), len(freq), None, ctypes.c_int64(0), ctypes.byref(n), _p(offs))
    if st:
        raise OracleError(st)
    out = np.zeros(n.value, dtype=SUBGRID_DTYPE)
    st = lib().orc_partition(_cc(c), _p(uvw), B, T, _p(freq), len(freq), _p(out), ctypes.c_int64(n.value), ctypes.byref(n), _p(offs))
    if st:
        raise OracleError(st)
    return out, offs


def _dims(uvw, freq):
    return uvw.shape[0], uvw.shape[1], len(freq)


def gridder_forward(c, meta, uvw, vis, freq):
    B, T, C = _dims(uvw, freq)
    S = c.subgrid_size
    out = np.zeros((len(meta), S, S), dtype=np.complex128)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    lib().orc_gridder_forward(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(uvw), B, T, _p(vis), _p(freq), C, _p(out))
    return out


def gridder_backward(c, meta, subgrids, uvw, freq):
    B, T, C = _dims(uvw, freq)
    out = np.zeros((B, T, C), dtype=np.complex128)
    sub = np.ascontiguousarray(subgrids, dtype=np.complex128)
    lib().orc_gridder_backward(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(sub), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


def subgrid_fft_forward(c, subgrids):
    a = np.array(subgrids, dtype=np.complex128, copy=True, order="C")
    lib().orc_subgrid_fft_forward(_cc(c), ctypes.c_int64(a.shape[0]), _p(a))
    return a


def subgrid_fft_backward(c, subgrids):
    a = np.array(subgrids, dtype=np.complex128, copy=True, order="C")
    lib().orc_subgrid_fft_backward(_cc(c), ctypes.c_int64(a.shape[0]), _p(a))
    return a


def adder_add(c, meta, subgrids):
    G = c.grid_size
    grid = np.zeros((G, G), dtype=np.complex128)
    sub = np.ascontiguousarray(subgrids, dtype=np.complex128)
    st = lib().orc_adder_add(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(sub), _p(grid))
    if st:
        raise OracleError(st)
    return grid


def adder_split(c, meta, grid):
    S = c.subgrid_size
    out = np.zeros((len(meta), S, S), dtype=np.complex128)
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    lib().orc_adder_split(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(g), _p(out))
    return out


def adderw_add(c, meta, subgrids):
    G, L = c.grid_size, c.wlayer_count
    layers = np.zeros((L, G, G), dtype=np.complex128)
    sub = np.ascontiguousarray(subgrids, dtype=np.complex128)
    lib().orc_adderw_add(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(sub), _p(layers))
    return layers


def adderw_split(c, meta, grid):
    S = c.subgrid_size
    out = np.zeros((len(meta), S, S), dtype=np.complex128)
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    lib().orc_adderw_split(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(g), _p(out))
    return out


def shift(a):
    """In-place quadrant swap (FFT.cs:248-264); returns a."""
    g = a.shape[-1]
    if a.dtype == np.complex128:
        lib().orc_shift_c128(g, _p(a))
    elif a.dtype == np.float64:
        lib().orc_shift_f64(g, _p(a))
    elif a.dtype == np.float32:
        lib().orc_shift_f32(g, _p(a))
    else:
        raise TypeError(a.dtype)
    return a


def wstack_ifft_float(layers, vis_count):
    layers = np.ascontiguousarray(layers, dtype=np.complex128)
    L, G = layers.shape[0], layers.shape[1]
    out = np.zeros((G, G), dtype=np.float32)
    lib().orc_wstack_ifft_float(G, L, _p(layers), ctypes.c_int64(int(vis_count)), _p(out))
    return out


def fft_backward(grid, vis_count):
    G = grid.shape[0]
    out = np.zeros((G, G), dtype=np.float64)
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    lib().orc_fft_backward(G, _p(g), ctypes.c_int64(int(vis_count)), _p(out))
    return out


def fft_forward_float(img, norm=1.0):
    G = img.shape[0]
    im = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros((G, G), dtype=np.complex128)
    lib().orc_fft_forward_float(G, _p(im), ctypes.c_double(norm), _p(out))
    return out


def fft2(a, sign):
    a = np.array(a, dtype=np.complex128, copy=True, order="C")
    lib().orc_fft2(a.shape[0], a.shape[1], _p(a), int(sign))
    return a


def substract(vis0, vis1, flags=None):
    v0 = np.ascontiguousarray(vis0, dtype=np.complex128)
    v1 = np.ascontiguousarray(vis1, dtype=np.complex128)
    out = np.zeros_like(v0)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    lib().orc_substract(ctypes.c_int64(v0.size), _p(v0), _p(v1), None if fl is None else _p(fl), _p(out))
    return out


def grid(c, meta, vis, uvw, freq):
    B, T, C = _dims(uvw, freq)
    G = c.grid_size
    out = np.zeros((G, G), dtype=np.complex128)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    st = lib().orc_grid(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(vis), _p(uvw), B, T, _p(freq), C, _p(out))
    if st:
        raise OracleError(st)
    return out


def degrid(c, meta, grid_, uvw, freq):
    """IDG.DeGrid; unlike the reference, the caller's grid is not mutated."""
    B, T, C = _dims(uvw, freq)
    g = np.array(grid_, dtype=np.complex128, copy=True, order="C")
    out = np.zeros((B, T, C), dtype=np.complex128)
    lib().orc_degrid(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(g), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


def gridw(c, meta, vis, uvw, freq):
    B, T, C = _dims(uvw, freq)
    G, L = c.grid_size, c.wlayer_count
    out = np.zeros((L, G, G), dtype=np.complex128)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    lib().orc_gridw(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(vis), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


def degridw(c, meta, grid_, uvw, freq):
    B, T, C = _dims(uvw, freq)
    g = np.array(grid_, dtype=np.complex128, copy=True, order="C")
    out = np.zeros((B, T, C), dtype=np.complex128)
    lib().orc_degridw(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(g), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


def to_image(c, meta, vis, uvw, freq):
    B, T, C = _dims(uvw, freq)
    G = c.grid_size
    out = np.zeros((G, G), dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    st = lib().orc_to_image(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(vis), _p(uvw), B, T, _p(freq), C, _p(out))
    if st:
        raise OracleError(st)
    return out


def gridw_to_image(c, meta, vis, uvw, freq):
    B, T, C = _dims(uvw, freq)
    G = c.grid_size
    out = np.zeros((G, G), dtype=np.float32)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    lib().orc_gridw_to_image(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(vis), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


def image_to_vis(c, meta, img, uvw, freq):
    B, T, C = _dims(uvw, freq)
    im = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros((B, T, C), dtype=np.complex128)
    lib().orc_image_to_vis(_cc(c), _p(meta), ctypes.c_int64(len(meta)), _p(im), _p(uvw), B, T, _p(freq), C, _p(out))
    return out


# ---- serial major cycle (oracle_cycle.cpp) ---------------------------------
CYCLE_STATS_DTYPE = np.dtype(
    [
        ("max_dirty", "<f4"),
        ("max_gradient", "<f4"),
        ("lambda_", "<f4"),
        ("abs_max_diff", "<f4"),
        ("objective", "<f8"),
        ("iterations", "<i4"),
        ("converged", "<i4"),
    ],
    align=True,
)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def psf_cut(psf, factor):
    psf = _f32(psf)
    n = psf.shape[0]
    out = np.zeros((n // factor, n // factor), dtype=np.float32)
    lib().orc_psf_cut(_p(psf), n, int(factor), _p(out))
    return out


def psf_max_sidelobe(psf, factor):
    psf = _f32(psf)
    f = lib().orc_psf_max_sidelobe
    f.restype = ctypes.c_float
    return f(_p(psf), psf.shape[0], int(factor))


def psf_max_lipschitz(psf):
    psf = _f32(psf)
    f = lib().orc_psf_max_lipschitz
    f.restype = ctypes.c_double
    return f(_p(psf), psf.shape[0], psf.shape[1])


def psf_amap(psf, H, W):
    psf = _f32(psf)
    out = np.zeros((H, W), dtype=np.float32)
    lib().orc_psf_amap(_p(psf), psf.shape[0], psf.shape[1], int(H), int(W), _p(out))
    return out


def psf_padded_correlation(psf, ypad, xpad):
    psf = _f32(psf)
    ph, pw = psf.shape
    out = np.zeros((ypad + ph, xpad + pw), dtype=np.complex128)
    lib().orc_psf_padded_correlation(_p(psf), ph, pw, int(ypad), int(xpad), _p(out))
    return out


def padded_convolve(kernel, kh, kw, image):
    image = _f32(image)
    H, W = image.shape
    k = np.ascontiguousarray(kernel, dtype=np.complex128)
    out = np.zeros((H, W), dtype=np.float32)
    lib().orc_padded_convolve(_p(k), int(kh), int(kw), _p(image), H, W, _p(out))
    return out


def psf_squared(psf):
    psf = _f32(psf)
    ph, pw = psf.shape
    out = np.zeros((2 * ph, 2 * pw), dtype=np.float32)
    lib().orc_psf_squared(_p(psf), ph, pw, _p(out))
    return out


def residuals_get_max(img):
    img = _f32(img)
    f = lib().orc_residuals_get_max
    f.restype = ctypes.c_float
    return f(_p(img), ctypes.c_int64(img.size))


def residuals_penalty(img):
    img = _f32(img)
    f = lib().orc_residuals_penalty
    f.restype = ctypes.c_double
    return f(_p(img), ctypes.c_int64(img.size))


def elastic_penalty(img, lam, alpha):
    img = _f32(img)
    f = lib().orc_elastic_penalty
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_float]
    return f(_p(img), img.size, lam, alpha)


def proximal(x, lipschitz, lam, alpha):
    f = lib().orc_proximal
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float] * 4
    return f(x, lipschitz, lam, alpha)


def cd_abs_max_diff(amap, x, g, lam, alpha):
    amap, x, g = _f32(amap), _f32(x), _f32(g)
    f = lib().orc_cd_abs_max_diff
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float]
    return f(_p(amap), _p(x), _p(g), x.shape[0], x.shape[1], lam, alpha)


def cd_deconvolve(amap, psf2, x, g, lam, alpha, iterations, epsilon):
    """FastSerialCD.Deconvolve; returns (x, g, converged, iterations) with copies updated."""
    amap, psf2 = _f32(amap), _f32(psf2)
    x, g = np.array(x, dtype=np.float32, copy=True), np.array(g, dtype=np.float32, copy=True)
    it = ctypes.c_int32(0)
    f = lib().orc_cd_deconvolve
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_float,
                  ctypes.c_void_p]
    conv = f(_p(amap), _p(psf2), psf2.shape[0], psf2.shape[1], _p(x), _p(g), x.shape[0], x.shape[1], lam, alpha,
             int(iterations), epsilon, ctypes.addressof(it))
    return x, g, bool(conv), it.value


def reconstruct_serial_cd(c, meta, vis, flags, uvw, freq, psf_cut_factor, max_major, lam, alpha,
                          deconv_iterations, deconv_epsilon):
    """MajorCycle.ReconstructSerialCD; returns (x_image, stats[ncycles])."""
    B, T, C = _dims(uvw, freq)
    G = c.grid_size
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    x = np.zeros((G, G), dtype=np.float32)
    stats = np.zeros(max_major, dtype=CYCLE_STATS_DTYPE)
    nc = ctypes.c_int32(0)
    f = lib().orc_reconstruct_serial_cd
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    st = f(ctypes.addressof(c), _p(meta), len(meta), _p(vis), None if fl is None else _p(fl), _p(uvw), B, T,
           _p(freq), C, int(psf_cut_factor), int(max_major), lam, alpha, int(deconv_iterations), deconv_epsilon,
           _p(x), _p(stats), ctypes.addressof(nc))
    if st:
        raise OracleError(st)
    return x, stats[: nc.value]
