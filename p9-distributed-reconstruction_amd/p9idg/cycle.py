"""Python mirror of the reference's serial major cycle, running on libp9idg.

Names and argument meaning follow the reference C# code
(Distributed-Reconstruction/):

    PSF.Cut / CalcMaxSidelobe / CalcMaxLipschitz / CalcPSFSquared / CalcAMap
                                                      Core/Common.cs:48-237
    PaddedConvolver(PSF.CalcPaddedFourierCorrelation(psf, totalSize), psfSize)
        .Convolve(image)                              Core/PaddedConvolver.cs
    FastSerialCD(totalSize, psf).Deconvolve / GetAbsMaxDiff / MaxLipschitz
                                                      Core/Deconvolution/FastSerialCD.cs
    MajorCycle.ReconstructSerialCD                    SingleReconstruction/MajorCycle.cs:33-134

Images are float32[rows, cols] (C# float[,]); a total size is (height, width).
The deconvolver and the convolver run on the GPU; Deconvolve updates the
caller's x and gradient arrays in place, as the reference updates its own.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr
from .idg import _flags_arg, _vis_arg

__all__ = ["PSF", "PaddedConvolver", "FastSerialCD", "DeconvolutionResult", "MajorCycle", "CYCLE_STATS_DTYPE"]

CYCLE_STATS_DTYPE = np.dtype(
    [
        ("max_dirty", "<f4"),
        ("max_gradient", "<f4"),
        ("lambda_", "<f4"),
        ("abs_max_diff", "<f4"),
        ("objective", "<f8"),
        ("iterations", "<i4"),
        ("converged", "<i4"),
        ("gradient_ms", "<f4"),
        ("deconv_ms", "<f4"),
        ("imaging_ms", "<f4"),
        ("pad", "<i4"),
    ],
    align=True,
)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _inplace(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous and a.ndim == 2):
        raise TypeError("%s must be a C-contiguous float32[rows, cols] array (updated in place)" % name)
    return a


class PSF:
    """PSF helpers (Common.cs:48-237)."""

    @staticmethod
    def Cut(psf, factor=2):
        psf = _f32(psf)
        n = psf.shape[0]
        m = n // factor if factor > 0 else 0  # the library rejects factor <= 0
        out = np.zeros((m, m), dtype=np.float32)
        check(_lib.idg().p9psf_cut(ptr(psf), n, int(factor), ptr(out)))
        return out

    @staticmethod
    def CalcMaxSidelobe(fullPsf, cutFactor=2):
        psf = _f32(fullPsf)
        v = ctypes.c_float()
        check(_lib.idg().p9psf_max_sidelobe(ptr(psf), psf.shape[0], int(cutFactor), ctypes.byref(v)))
        return np.float32(v.value)

    @staticmethod
    def CalcMaxLipschitz(psf):
        psf = _f32(psf)
        v = ctypes.c_double()
        check(_lib.idg().p9psf_max_lipschitz(ptr(psf), psf.shape[0], psf.shape[1], ctypes.byref(v)))
        return v.value

    @staticmethod
    def CalcPSFSquared(psf, device=0):
        psf = _f32(psf)
        out = np.zeros((2 * psf.shape[0], 2 * psf.shape[1]), dtype=np.float32)
        check(_lib.idg().p9psf_squared(int(device), ptr(psf), psf.shape[0], psf.shape[1], ptr(out)))
        return out

    @staticmethod
    def CalcAMap(psf, totalSize, device=0):
        psf = _f32(psf)
        H, W = totalSize
        out = np.zeros((H, W), dtype=np.float32)
        check(_lib.idg().p9psf_amap(int(device), ptr(psf), psf.shape[0], psf.shape[1], int(H), int(W), ptr(out)))
        return out


class PaddedConvolver:
    """PaddedConvolver(PSF.CalcPaddedFourierCorrelation(psf, (0, 0, H, W)),
    (0, 0, psf rows, psf cols)): correlation of an H x W image with the PSF,
    FP64 FFTs on the GPU."""

    def __init__(self, psf, totalSize, device=0):
        psf = _f32(psf)
        self.totalSize = (int(totalSize[0]), int(totalSize[1]))
        h = ctypes.c_void_p()
        check(_lib.idg().p9conv_create_correlation(int(device), ptr(psf), psf.shape[0], psf.shape[1],
                                                   self.totalSize[0], self.totalSize[1], ctypes.byref(h)))
        self._h = h

    def Convolve(self, image):
        image = _f32(image)
        if image.shape != self.totalSize:
            raise ValueError("image must be %dx%d" % self.totalSize)
        out = np.zeros(self.totalSize, dtype=np.float32)
        check(_lib.idg().p9conv_convolve(self._h, ptr(image), ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.idg().p9conv_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeconvolutionResult:
    """Converged flag and iteration count of a deconvolution (the reference's
    DeconvolutionResult without the stopwatch)."""

    def __init__(self, converged, iterations):
        self.Converged = bool(converged)
        self.IterationCount = int(iterations)

    def __repr__(self):
        return "DeconvolutionResult(Converged=%s, IterationCount=%d)" % (self.Converged, self.IterationCount)


class FastSerialCD:
    """FastSerialCD(totalSize, psf) on the GPU (FastSerialCD.cs:31-135)."""

    def __init__(self, totalSize, psf, device=0):
        psf = _f32(psf)
        self.totalSize = (int(totalSize[0]), int(totalSize[1]))
        h = ctypes.c_void_p()
        check(_lib.idg().p9cd_create(int(device), self.totalSize[0], self.totalSize[1], ptr(psf), psf.shape[0],
                                     psf.shape[1], ctypes.byref(h)))
        self._h = h

    @property
    def MaxLipschitz(self):
        v = ctypes.c_float()
        check(_lib.idg().p9cd_max_lipschitz(self._h, ctypes.byref(v)))
        return np.float32(v.value)

    def GetAbsMaxDiff(self, xImage, gradients, lambda_, alpha):
        x, g = _f32(xImage), _f32(gradients)
        v = ctypes.c_float()
        check(_lib.idg().p9cd_get_abs_max_diff(self._h, ptr(x), ptr(g), float(lambda_), float(alpha), ctypes.byref(v)))
        return np.float32(v.value)

    def Deconvolve(self, xImage, bMap, lambda_, alpha, maxIteration, epsilon=1e-4):
        x = _inplace(xImage, "xImage")
        g = _inplace(bMap, "bMap")
        if x.shape != self.totalSize or g.shape != self.totalSize:
            raise ValueError("images must be %dx%d" % self.totalSize)
        conv, it = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.idg().p9cd_deconvolve(self._h, ptr(x), ptr(g), float(lambda_), float(alpha), int(maxIteration),
                                         float(epsilon), ctypes.byref(conv), ctypes.byref(it)))
        return DeconvolutionResult(conv.value, it.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.idg().p9cd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MajorCycle:
    @staticmethod
    def ReconstructSerialCD(plan, visibilities, flags, psfCutFactor, maxMajorCycle, lambda_, alpha,
                            deconvIterations, deconvEpsilon, obsName=None):
        """MajorCycle.ReconstructSerialCD (MajorCycle.cs:33-134) on a Plan
        whose observation is set; returns (xImage float32[G, G], stats
        structured array, one row per cycle run).  With obsName, every cycle
        writes <obsName>_dirty_serial_majorCycle<k>.fits and
        <obsName>_model_serial_majorCycle<k>.fits (MajorCycle.cs:91,117)."""
        vis, vtype = _vis_arg(visibilities, plan.shape)
        fl = _flags_arg(flags, plan.shape)
        G = plan.c.GridSize
        x = np.zeros((G, G), dtype=np.float32)
        stats = np.zeros(max(int(maxMajorCycle), 1), dtype=CYCLE_STATS_DTYPE)
        n = ctypes.c_int32()
        name = None if obsName is None else str(obsName).encode()
        check(_lib.idg().p9idg_reconstruct_serial_cd(plan._h, name, ptr(vis), vtype, ptr(fl), int(psfCutFactor),
                                                     int(maxMajorCycle), float(lambda_), float(alpha),
                                                     int(deconvIterations), float(deconvEpsilon), ptr(x), ptr(stats),
                                                     ctypes.byref(n)))
        return x, stats[: n.value]
