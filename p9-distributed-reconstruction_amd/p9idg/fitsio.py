"""Python mirror of the reference's FITS loaders, running on libp9idg.so.

Names and argument meaning follow the reference C# code
(Distributed-Reconstruction/):

    FitsIO.ReadFrequencies / ReadUVW / CountBaselines / ReadVisibilities /
    ReadFlags / ReadImage / ReadImageDouble / ReadBeam / ReadCASAFits /
    Write / WriteImag / Stitch                        Core/FitsIO.cs
    MeasurementData(vis, uvw, freq, flags), LoadLMC, LoadSimulatedPoints,
    CreateSimulatedStandardParams                     SingleReconstruction/MeasurementData.cs

The file format work (header parsing, big-endian decoding, the polarization
sum, the v flip) is native (include/p9fits.h); these wrappers allocate the
numpy results.  The readers need no GPU.  Arrays come back as the C# arrays
they replace: uvw float64[B, T, 3], visibilities complex128[B, T, C] (or
complex64 with dtype=np.complex64, the plan's compact input type), flags
bool[B, T, C], images float32 / float64[rows, cols].
"""
import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import check, ptr

__all__ = ["FitsIO", "MeasurementData"]


def _path(file):
    return os.fsencode(file)


def _vis_type(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.complex128:
        return _lib.C128
    if dtype == np.complex64:
        return _lib.C64
    raise TypeError("visibilities are read as complex128 or complex64, not %s" % dtype)


class FitsIO:
    """FitsIO (Core/FitsIO.cs): the primary HDU of a FITS file."""

    @staticmethod
    def ReadInfo(file):
        """Header of the primary HDU: (BITPIX, shape), shape in numpy order
        (NAXISn first, NAXIS1 last)."""
        info = _lib.FitsInfo()
        check(_lib.idg().p9fits_read_info(_path(file), ctypes.byref(info)))
        return info.bitpix, tuple(int(info.naxes[a]) for a in reversed(range(info.naxis)))

    @staticmethod
    def ReadFrequencies(file):
        """FitsIO.cs:184-189: double[C]."""
        n = ctypes.c_int64()
        lib = _lib.idg()
        check(lib.p9fits_read_frequencies(_path(file), None, 0, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.float64)
        check(lib.p9fits_read_frequencies(_path(file), ptr(out), out.size, ctypes.byref(n)))
        return out

    @staticmethod
    def CountBaselines(file):
        """FitsIO.cs:247-255."""
        b, t = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.idg().p9fits_count_baselines(_path(file), ctypes.byref(b), ctypes.byref(t)))
        return b.value

    @staticmethod
    def _uvw_shape(file):
        b, t = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.idg().p9fits_count_baselines(_path(file), ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    @staticmethod
    def ReadUVW(file, fromBl=None, toBl=None, out=None):
        """FitsIO.cs:196-245: double[toBl - fromBl, T, 3] with v negated.
        out, if given, is a C-contiguous float64 array of that shape to fill."""
        B, T = FitsIO._uvw_shape(file)
        f = 0 if fromBl is None else int(fromBl)
        t = B if toBl is None else int(toBl)
        if out is None:
            out = np.empty((max(t - f, 0), T, 3), dtype=np.float64)
        elif out.dtype != np.float64 or not out.flags.c_contiguous or out.shape != (t - f, T, 3):
            raise ValueError("out must be C-contiguous float64 %s" % ((t - f, T, 3),))
        check(_lib.idg().p9fits_read_uvw(_path(file), f, t, ptr(out)))
        return out

    @staticmethod
    def ReadVisibilities(file, *args, dtype=np.complex128, out=None):
        """ReadVisibilities(file, baselinesCount, timessamplesCount, channelsCount, norm)
        (FitsIO.cs:267-294) or ReadVisibilities(file, fromBl, toBl,
        timessamplesCount, channelsCount, norm) (FitsIO.cs:296-325):
        (XX + YY) / norm per baseline, timestep and channel."""
        if len(args) == 4:
            f, t = 0, int(args[0])
            T, C, norm = int(args[1]), int(args[2]), float(args[3])
        elif len(args) == 5:
            f, t = int(args[0]), int(args[1])
            T, C, norm = int(args[2]), int(args[3]), float(args[4])
        else:
            raise TypeError("ReadVisibilities(file, B, T, C, norm) or (file, fromBl, toBl, T, C, norm)")
        vt = _vis_type(dtype if out is None else out.dtype)
        shape = (max(t - f, 0), T, C)
        if out is None:
            out = np.empty(shape, dtype=dtype)
        elif not out.flags.c_contiguous or out.shape != shape:
            raise ValueError("out must be C-contiguous with shape %s" % (shape,))
        check(_lib.idg().p9fits_read_visibilities(_path(file), f, t, T, C, norm, ptr(out), vt))
        return out

    @staticmethod
    def ReadFlags(file, *args, out=None):
        """ReadFlags(file, baselinesCount, timessamplesCount, channelsCount)
        (FitsIO.cs:327-357) or ReadFlags(file, fromBl, toBl, timessamplesCount,
        channelsCount) (FitsIO.cs:359-390): True where the polarization
        values do not sum to 0."""
        if len(args) == 3:
            f, t, T, C = 0, int(args[0]), int(args[1]), int(args[2])
        elif len(args) == 4:
            f, t, T, C = int(args[0]), int(args[1]), int(args[2]), int(args[3])
        else:
            raise TypeError("ReadFlags(file, B, T, C) or (file, fromBl, toBl, T, C)")
        shape = (max(t - f, 0), T, C)
        if out is None:
            out = np.empty(shape, dtype=np.bool_)
        elif out.dtype != np.bool_ or not out.flags.c_contiguous or out.shape != shape:
            raise ValueError("out must be C-contiguous bool with shape %s" % (shape,))
        check(_lib.idg().p9fits_read_flags(_path(file), f, t, T, C, ptr(out)))
        return out

    @staticmethod
    def _read_2d(file, ftype, dtype):
        bitpix, shape = FitsIO.ReadInfo(file)
        if len(shape) != 2:
            shape = (0, 0)  # the native reader reports the mismatch
        out = np.empty(shape, dtype=dtype)
        check(_lib.idg().p9fits_read_image(_path(file), ftype, ptr(out), shape[0], shape[1]))
        return out

    @staticmethod
    def ReadImage(file):
        """FitsIO.cs:152-166: float[rows, cols]."""
        return FitsIO._read_2d(file, _lib.FITS_F32, np.float32)

    @staticmethod
    def ReadImageDouble(file):
        """FitsIO.cs:168-182: double[rows, cols]."""
        return FitsIO._read_2d(file, _lib.FITS_F64, np.float64)

    @staticmethod
    def ReadBeam(file):
        """FitsIO.cs:90-107: a square double image."""
        img = FitsIO.ReadImageDouble(file)
        if img.shape[0] != img.shape[1]:
            raise ValueError("%s: beam is %d x %d, not square" % (file, img.shape[0], img.shape[1]))
        return img

    @staticmethod
    def ReadCASAFits(file):
        """FitsIO.cs:134-150: plane [0][0] of a float[.][.][rows][cols] image."""
        bitpix, shape = FitsIO.ReadInfo(file)
        rows, cols = (shape[2], shape[3]) if len(shape) == 4 else (0, 0)
        out = np.empty((rows, cols), dtype=np.float32)
        check(_lib.idg().p9fits_read_casa(_path(file), ptr(out), rows, cols))
        return out

    @staticmethod
    def Write(img, file=None):
        """Write(Complex[,]) (real parts, double), Write<T>(T[,]) for float and
        double images, and Write(double[,,], file) (one file <file><k>.fits per
        leading index) (FitsIO.cs:16-88)."""
        img = np.asarray(img)
        if img.ndim == 3:
            base = "Outputfile" if file is None else str(file)
            for k in range(img.shape[0]):
                FitsIO.Write(np.asarray(img[k], dtype=np.float64), base + str(k) + ".fits")
            return
        if img.ndim != 2:
            raise ValueError("Write takes a 2-D image or a 3-D double cube")
        file = "Outputfile.fits" if file is None else file
        if np.iscomplexobj(img):
            a = np.ascontiguousarray(img, dtype=np.complex128)
            ftype = _lib.FITS_C128_REAL
        elif img.dtype == np.float32:
            a = np.ascontiguousarray(img)
            ftype = _lib.FITS_F32
        elif img.dtype == np.float64:
            a = np.ascontiguousarray(img)
            ftype = _lib.FITS_F64
        else:
            raise TypeError("Write takes float32, float64 or complex images, not %s" % img.dtype)
        check(_lib.idg().p9fits_write_image(_path(file), ftype, ptr(a), a.shape[0], a.shape[1]))

    @staticmethod
    def WriteImag(img, file="Outputfile_imag.fits"):
        """FitsIO.cs:109-130: imaginary parts of a complex image, double."""
        a = np.ascontiguousarray(img, dtype=np.complex128)
        if a.ndim != 2:
            raise ValueError("WriteImag takes a 2-D complex image")
        check(_lib.idg().p9fits_write_image(_path(file), _lib.FITS_C128_IMAG, ptr(a), a.shape[0], a.shape[1]))

    @staticmethod
    def Stitch(x, y):
        """FitsIO.cs:394-409: y appended to x along the baseline axis (the
        trailing axes must agree)."""
        if x.shape[1:] != y.shape[1:]:
            raise ValueError("Stitch: trailing shapes differ: %s and %s" % (x.shape[1:], y.shape[1:]))
        return np.concatenate([x, y], axis=0)

    @staticmethod
    def SetThreads(threads):
        """Host threads the native readers decode on (0: one per hardware thread)."""
        check(_lib.idg().p9fits_set_threads(int(threads)))


class MeasurementData:
    """MeasurementData (SingleReconstruction/MeasurementData.cs:11-75)."""

    def __init__(self, vis, uvw, freq, flags):
        self.Visibilities = vis
        self.UVW = uvw
        self.Frequencies = freq
        self.Flags = flags
        # number of unflagged visibilities (MeasurementData.cs:26-31)
        self.VisibilitiesCount = int(flags.size - np.count_nonzero(flags))

    @staticmethod
    def _load_parts(folder, names, with_flags, dtype):
        """Reads the (uvw, flags, vis) file triples `names` and stitches them
        along the baseline axis straight into preallocated arrays."""
        freq = FitsIO.ReadFrequencies(os.path.join(folder, "freq.fits"))
        C = freq.size
        shapes = [FitsIO._uvw_shape(os.path.join(folder, u)) for u, _, _ in names]
        T = shapes[0][1]
        for (u, _, _), (b, t) in zip(names, shapes):
            if t != T:
                raise ValueError("%s has %d timesteps, the first part %d" % (u, t, T))
        B = sum(b for b, _ in shapes)
        uvw = np.empty((B, T, 3), dtype=np.float64)
        vis = np.empty((B, T, C), dtype=dtype)
        flags = np.zeros((B, T, C), dtype=np.bool_)
        norm = 2.0
        b0 = 0
        for (u, fl, v), (b, _) in zip(names, shapes):
            sl = slice(b0, b0 + b)
            FitsIO.ReadUVW(os.path.join(folder, u), 0, b, out=uvw[sl])
            if with_flags:
                FitsIO.ReadFlags(os.path.join(folder, fl), b, T, C, out=flags[sl])
            FitsIO.ReadVisibilities(os.path.join(folder, v), b, T, C, norm, out=vis[sl])
            b0 += b
        return MeasurementData(vis, uvw, freq, flags)

    @staticmethod
    def LoadLMC(folder, dtype=np.complex128):
        """MeasurementData.cs:33-53: freq.fits and the eight parts
        uvw<i>.fits, flags<i>.fits, vis<i>.fits (norm 2), stitched in order."""
        names = [("uvw%d.fits" % i, "flags%d.fits" % i, "vis%d.fits" % i) for i in range(8)]
        return MeasurementData._load_parts(folder, names, True, dtype)

    @staticmethod
    def LoadSimulatedPoints(folder, dtype=np.complex128):
        """MeasurementData.cs:55-63: freq.fits, uvw.fits, vis.fits (norm 2),
        nothing flagged."""
        return MeasurementData._load_parts(folder, [("uvw.fits", None, "vis.fits")], False, dtype)

    @staticmethod
    def CreateSimulatedStandardParams(visibilitiesCount):
        """MeasurementData.cs:65-73: 256^2 grid, subgrid 8, kernel 4, 1024
        timesteps, 1 arcsec cells, one w-layer."""
        from .idg import GriddingConstants

        cell = np.float32(1.0 / 3600.0 * math.pi / 180.0)
        return GriddingConstants(int(visibilitiesCount), 256, 8, 4, 1024, cell, 1, 0.0)
