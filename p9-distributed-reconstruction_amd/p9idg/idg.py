"""Python mirror of the reference's gridding API, running on libp9idg.

Names, argument order and meaning follow the reference C# classes
(Distributed-Reconstruction/Core/ImageDomainGridder/*.cs and Core/FFT.cs):

    GriddingConstants(visCount, gridSize, subgridSize, kernelSize,
                      maxTimesteps, scaleArcSec, wLayerCount, wStep)
    Partitioner.CreatePartition(c, uvw, frequencies)
    IDG.Grid / DeGrid / GridW / DeGridW / GridPSF / ToImage / CalculatePSF
    FFT.WStackIFFTFloat / Forward / Backward / Shift
    Visibilities.Substract

Arrays use the C# layouts: uvw float64[B,T,3] (metres), frequencies
float64[C] (Hz), visibilities complex128[B,T,C] (complex64 accepted), grids
complex128[G,G], w-layers complex128[L,G,G], flags bool[B,T,C].

Errors raise P9IDGError carrying the reference's exception text.  Plan is
the device-resident interface: inputs are uploaded once and every stage runs
on the GPU's HBM-resident buffers.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import P9IDGError, SUBGRID_DTYPE, check, ptr

__all__ = [
    "GriddingConstants",
    "Partition",
    "Partitioner",
    "IDG",
    "FFT",
    "Visibilities",
    "Plan",
    "Batch",
    "P9IDGError",
    "SUBGRID_DTYPE",
]

_KERNELS = {"fp32": _lib.KERNEL_FP32, "fp32_sep": _lib.KERNEL_FP32_SEP, "mfma_bf16": _lib.KERNEL_MFMA_BF16,
            "mfma_fp16": _lib.KERNEL_MFMA_FP16}


class GriddingConstants:
    """Immutable gridding parameters (GriddingConstants.cs:11-53).

    ScaleArcSec is radians per pixel (the reference passes the cell size in
    radians, RunningMethod.cs:22,45) and is stored as float32, as is
    ImageSize = ScaleArcSec * GridSize.  WStepLambda is 0: the reference never
    assigns it (GriddingConstants.cs:25)."""

    def __init__(self, visCount, gridSize, subgridSize, kernelSize, maxTimesteps, scaleArcSec, wLayerCount, wStep):
        self.VisibilitiesCount = int(visCount)
        self.GridSize = int(gridSize)
        self.SubgridSize = int(subgridSize)
        self.KernelSize = int(kernelSize)
        self.MaxTimestepsPerSubgrid = int(maxTimesteps)
        self.ScaleArcSec = np.float32(scaleArcSec)
        self.ImageSize = np.float32(self.ScaleArcSec * np.float32(self.GridSize))
        self.WLayerCount = int(wLayerCount)
        self.WStep = float(wStep)
        self.WStepLambda = np.float32(0.0)

    @property
    def SubgridSpheroidal(self):
        # MathFunctions.CalcIdentitySpheroidal (MathFunctions.cs:69-78)
        return np.ones((self.SubgridSize, self.SubgridSize), dtype=np.float32)

    @property
    def GridSpheroidal(self):
        return np.ones((self.GridSize, self.GridSize), dtype=np.float32)

    def _struct(self):
        return _lib.Constants(
            self.VisibilitiesCount,
            self.GridSize,
            self.SubgridSize,
            self.KernelSize,
            self.MaxTimestepsPerSubgrid,
            float(self.ScaleArcSec),
            self.WLayerCount,
            self.WStep,
            float(self.WStepLambda),
        )


class Partition:
    """Flattened List<List<Subgrid>>: `subgrids` is a structured array in
    baseline order then time order; partition[b] is baseline b's list."""

    def __init__(self, subgrids, offsets):
        self.subgrids = subgrids
        self.offsets = offsets

    def __len__(self):
        return len(self.offsets) - 1

    @property
    def Count(self):
        return len(self)

    def __getitem__(self, b):
        return self.subgrids[self.offsets[b] : self.offsets[b + 1]]

    def __iter__(self):
        for b in range(len(self)):
            yield self[b]

    @property
    def total(self):
        return len(self.subgrids)


class Partitioner:
    @staticmethod
    def CreatePartition(c, uvw, frequencies):
        """Partitioner.cs:20-124 (bit-exact)."""
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        freq = np.ascontiguousarray(frequencies, dtype=np.float64)
        B, T = uvw.shape[0], uvw.shape[1]
        cs = c._struct()
        n = ctypes.c_int64(0)
        offs = np.zeros(B + 1, dtype=np.int64)
        check(_lib.idg().p9idg_partition(ctypes.byref(cs), ptr(uvw), B, T, ptr(freq), len(freq), None, 0, ctypes.byref(n), ptr(offs)))
        out = np.zeros(n.value, dtype=SUBGRID_DTYPE)
        check(_lib.idg().p9idg_partition(ctypes.byref(cs), ptr(uvw), B, T, ptr(freq), len(freq), ptr(out), n.value, ctypes.byref(n), ptr(offs)))
        return Partition(out, offs)

    @staticmethod
    def MetersToPixels(meters, imageSize, frequency):
        return meters * float(imageSize) * (frequency / 299792458.0)

    @staticmethod
    def MetersToLambda(meters, frequency):
        return meters * (frequency / 299792458.0)


def _as_partition(metadata):
    if isinstance(metadata, Partition):
        return metadata.subgrids
    a = np.asarray(metadata)
    if a.dtype == SUBGRID_DTYPE:
        return a
    raise TypeError("metadata must be a Partition or a SUBGRID_DTYPE array")


def _vis_arg(vis, shape=None):
    vis = np.asarray(vis)
    _expect(vis, shape, "visibilities")
    if vis.dtype == np.complex64:
        return np.ascontiguousarray(vis), _lib.C64
    return np.ascontiguousarray(vis, dtype=np.complex128), _lib.C128


def _expect(a, shape, name):
    """The C ABI reads exactly prod(shape) elements of each input: anything
    else is refused here, as the reference's shaped C# arrays would throw."""
    if shape is not None and tuple(np.shape(a)) != tuple(shape):
        raise ValueError("%s must have shape %s, got %s" % (name, tuple(shape), tuple(np.shape(a))))


def _flags_arg(flags, shape):
    if flags is None:
        return None
    _expect(flags, shape, "flags")
    return np.ascontiguousarray(flags, dtype=np.uint8)


class Plan:
    """Device-resident gridder: one HIP stream, HBM buffers for the
    observation, the visibilities, the grid / w-layers, the model grid, the
    model visibilities and the image."""

    def __init__(self, c, device=0, kernel="fp32"):
        self.c = c
        self._h = ctypes.c_void_p()
        cs = c._struct()
        check(_lib.idg().p9idg_create(ctypes.byref(cs), int(device), ctypes.byref(self._h)))
        self.set_kernel(kernel)
        self.shape = None

    def close(self):
        if self._h:
            _lib.idg().p9idg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_kernel(self, kernel):
        k = _KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        check(_lib.idg().p9idg_set_kernel(self._h, k))

    def set_deterministic(self, on=True):
        """Ordered subgrid sums: gridding results identical from run to run."""
        check(_lib.idg().p9idg_set_deterministic(self._h, 1 if on else 0))

    def set_launch(self, queue=True, grid_chunk=2048, degrid_chunk=2048, workgroups=0):
        check(_lib.idg().p9idg_set_launch(self._h, 1 if queue else 0, int(grid_chunk), int(degrid_chunk), int(workgroups)))

    def autotune(self, reps=3):
        """Times candidate launch shapes on this observation and keeps the fastest."""
        g, d = ctypes.c_float(), ctypes.c_float()
        check(_lib.idg().p9idg_autotune(self._h, int(reps), ctypes.byref(g), ctypes.byref(d)))
        out = self.get_launch()
        out.update(grid_ms=g.value, degrid_ms=d.value)
        return out

    def get_launch(self):
        v = [ctypes.c_int32() for _ in range(4)]
        check(_lib.idg().p9idg_get_launch(self._h, *[ctypes.byref(x) for x in v]))
        return {"queue": bool(v[0].value), "grid_chunk": v[1].value, "degrid_chunk": v[2].value, "workgroups": v[3].value}

    def set_observation(self, uvw, frequencies, metadata):
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        freq = np.ascontiguousarray(frequencies, dtype=np.float64)
        if uvw.ndim != 3 or uvw.shape[2] != 3:
            raise ValueError("uvw must have shape (baselines, timesteps, 3), got %s" % (uvw.shape,))
        if freq.ndim != 1:
            raise ValueError("frequencies must be one-dimensional, got shape %s" % (freq.shape,))
        meta = np.ascontiguousarray(_as_partition(metadata))
        B, T = uvw.shape[0], uvw.shape[1]
        check(_lib.idg().p9idg_set_observation(self._h, ptr(uvw), B, T, ptr(freq), len(freq), ptr(meta), len(meta)))
        self.shape = (B, T, len(freq))

    # -- host-buffer drop-ins (IDG.cs) --
    @property
    def _grid_shape(self):
        return (self.c.GridSize, self.c.GridSize)

    def grid(self, vis):
        v, t = _vis_arg(vis, self.shape)
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.complex128)
        check(_lib.idg().p9idg_grid(self._h, ptr(v), t, ptr(out)))
        return out

    def gridw(self, vis):
        v, t = _vis_arg(vis, self.shape)
        G, L = self.c.GridSize, self.c.WLayerCount
        out = np.zeros((L, G, G), dtype=np.complex128)
        check(_lib.idg().p9idg_gridw(self._h, ptr(v), t, ptr(out)))
        return out

    def degrid(self, grid):
        _expect(grid, self._grid_shape, "grid")
        g = np.ascontiguousarray(grid, dtype=np.complex128)
        out = np.zeros(self.shape, dtype=np.complex128)
        check(_lib.idg().p9idg_degrid(self._h, ptr(g), ptr(out)))
        return out

    def degridw(self, grid):
        _expect(grid, self._grid_shape, "grid")
        g = np.ascontiguousarray(grid, dtype=np.complex128)
        out = np.zeros(self.shape, dtype=np.complex128)
        check(_lib.idg().p9idg_degridw(self._h, ptr(g), ptr(out)))
        return out

    def grid_psf(self, flags=None):
        fl = _flags_arg(flags, self.shape)
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.complex128)
        check(_lib.idg().p9idg_grid_psf(self._h, ptr(fl), ptr(out)))
        return out

    def to_image(self, vis):
        v, t = _vis_arg(vis, self.shape)
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.float64)
        check(_lib.idg().p9idg_to_image(self._h, ptr(v), t, ptr(out)))
        return out

    def calculate_psf(self, flags=None):
        fl = _flags_arg(flags, self.shape)
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.float64)
        check(_lib.idg().p9idg_calculate_psf(self._h, ptr(fl), ptr(out)))
        return out

    def gridw_to_image(self, vis):
        v, t = _vis_arg(vis, self.shape)
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.float32)
        check(_lib.idg().p9idg_gridw_to_image(self._h, ptr(v), t, ptr(out)))
        return out

    def image_to_vis(self, image):
        _expect(image, self._grid_shape, "image")
        im = np.ascontiguousarray(image, dtype=np.float32)
        out = np.zeros(self.shape, dtype=np.complex128)
        check(_lib.idg().p9idg_image_to_vis(self._h, ptr(im), ptr(out)))
        return out

    # -- device-resident stages --
    def upload_vis(self, vis):
        v, t = _vis_arg(vis, self.shape)
        check(_lib.idg().p9idg_upload_vis(self._h, ptr(v), t))

    def set_psf_vis(self, flags=None):
        fl = _flags_arg(flags, self.shape)
        check(_lib.idg().p9idg_set_psf_vis(self._h, ptr(fl)))

    def exec_grid(self, mode=_lib.GRID_WSTACK):
        check(_lib.idg().p9idg_exec_grid(self._h, int(mode)))

    def exec_degrid(self, mode=_lib.DEGRID_WSTACK):
        check(_lib.idg().p9idg_exec_degrid(self._h, int(mode)))

    def exec_grid_to_image(self, mode=_lib.GRID_WSTACK):
        check(_lib.idg().p9idg_exec_grid_to_image(self._h, int(mode)))

    def exec_image_to_grid(self):
        check(_lib.idg().p9idg_exec_image_to_grid(self._h))

    def sync(self):
        check(_lib.idg().p9idg_sync(self._h))

    def set_flags(self, flags=None):
        fl = _flags_arg(flags, self.shape)
        check(_lib.idg().p9idg_set_flags(self._h, ptr(fl)))

    def exec_residual(self):
        check(_lib.idg().p9idg_exec_residual(self._h))

    def set_grid_source(self, src):
        check(_lib.idg().p9idg_set_grid_source(self._h, int(src)))

    def exec_cycle(self, mode=_lib.GRID_WSUM):
        check(_lib.idg().p9idg_exec_cycle(self._h, int(mode)))

    def capture_cycle(self, mode=_lib.GRID_WSUM):
        check(_lib.idg().p9idg_capture_cycle(self._h, int(mode)))

    def launch_cycle(self):
        check(_lib.idg().p9idg_launch_cycle(self._h))

    def download_residual(self):
        out = np.zeros(self.shape, dtype=np.complex64)
        check(_lib.idg().p9idg_download_residual(self._h, ptr(out)))
        return out

    def download_grid(self, mode=_lib.GRID_PLAIN):
        G, L = self.c.GridSize, self.c.WLayerCount
        shape = (L, G, G) if mode == _lib.GRID_WSTACK else (G, G)
        out = np.zeros(shape, dtype=np.complex64)
        check(_lib.idg().p9idg_download_grid(self._h, int(mode), ptr(out)))
        return out

    def upload_model_grid(self, grid):
        _expect(grid, self._grid_shape, "model grid")
        g = np.ascontiguousarray(grid, dtype=np.complex64)
        check(_lib.idg().p9idg_upload_model_grid(self._h, ptr(g)))

    def download_model_grid(self):
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.complex64)
        check(_lib.idg().p9idg_download_model_grid(self._h, ptr(out)))
        return out

    def download_model_vis(self):
        out = np.zeros(self.shape, dtype=np.complex64)
        check(_lib.idg().p9idg_download_model_vis(self._h, ptr(out)))
        return out

    def upload_image(self, image):
        _expect(image, self._grid_shape, "image")
        im = np.ascontiguousarray(image, dtype=np.float32)
        check(_lib.idg().p9idg_upload_image(self._h, ptr(im)))

    def download_image(self):
        G = self.c.GridSize
        out = np.zeros((G, G), dtype=np.float32)
        check(_lib.idg().p9idg_download_image(self._h, ptr(out)))
        return out

    def timings(self):
        g, d, f = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        check(_lib.idg().p9idg_last_timings(self._h, ctypes.byref(g), ctypes.byref(d), ctypes.byref(f)))
        return {"grid_ms": g.value, "degrid_ms": d.value, "fft_ms": f.value}

    def timing_reset(self, collect=True):
        check(_lib.idg().p9idg_timing_reset(self._h, 1 if collect else 0))

    def timing_totals(self):
        """(grid_ms_total, grid_launches, degrid_ms_total, degrid_launches)"""
        g, d = ctypes.c_double(), ctypes.c_double()
        ng, nd = ctypes.c_int64(), ctypes.c_int64()
        check(_lib.idg().p9idg_timing_totals(self._h, ctypes.byref(g), ctypes.byref(ng), ctypes.byref(d), ctypes.byref(nd)))
        return g.value, ng.value, d.value, nd.value

    def timing_stage(self, stage):
        """(total ms, spans) of one collected stage (p9idg_stage: gridder,
        degridder, zero fills, ToImage, ToUVW)."""
        t, n = ctypes.c_double(), ctypes.c_int64()
        check(_lib.idg().p9idg_timing_stage(self._h, int(stage), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    @property
    def gridded_visibilities(self):
        n = ctypes.c_int64()
        check(_lib.idg().p9idg_gridded_visibilities(self._h, ctypes.byref(n)))
        return n.value


class Batch:
    """Several plans' gridding / degridding as one persistent work-queued
    launch per (subgrid size, kernel) group, optionally replayed as a hipGraph."""

    def __init__(self, plans, grid_mode=_lib.GRID_WSTACK, degrid_mode=_lib.DEGRID_WSTACK):
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        self._h = ctypes.c_void_p()
        check(_lib.idg().p9idg_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(self.plans), ctypes.byref(self._h)))
        self.prepare(grid_mode, degrid_mode)

    def prepare(self, grid_mode=_lib.GRID_WSTACK, degrid_mode=_lib.DEGRID_WSTACK):
        """(Re)builds the merged work queues, e.g. after a plan's observation changed."""
        check(_lib.idg().p9idg_batch_prepare(self._h, int(grid_mode), int(degrid_mode)))

    def exec(self, what=3):
        check(_lib.idg().p9idg_batch_exec(self._h, int(what)))

    def capture(self, what=3):
        check(_lib.idg().p9idg_batch_capture(self._h, int(what)))

    def launch(self):
        check(_lib.idg().p9idg_batch_launch(self._h))

    def sync(self):
        check(_lib.idg().p9idg_batch_sync(self._h))

    def timings(self):
        g, d = ctypes.c_float(), ctypes.c_float()
        check(_lib.idg().p9idg_batch_timings(self._h, ctypes.byref(g), ctypes.byref(d)))
        return {"grid_ms": g.value, "degrid_ms": d.value}

    def close(self):
        if self._h:
            _lib.idg().p9idg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _PlanCache:
    """The static drop-ins reuse one plan per (constants, metadata, uvw,
    frequencies) so a major cycle uploads its observation once."""

    def __init__(self):
        self.key = None
        self.plan = None

    def get(self, c, metadata, uvw, frequencies, B, T):
        key = (c, metadata, uvw, frequencies)
        if self.key is not None and all(a is b for a, b in zip(self.key, key)):
            return self.plan
        if self.plan is not None:
            self.plan.close()
        plan = Plan(c)
        plan.set_observation(uvw, frequencies, metadata)
        self.key, self.plan = key, plan
        return plan


_cache = _PlanCache()


def _plan(c, metadata, uvw, frequencies):
    uvw_a = np.asarray(uvw)
    return _cache.get(c, metadata, uvw, frequencies, uvw_a.shape[0], uvw_a.shape[1])


class IDG:
    """IDG.cs:12-114 on the GPU.  DeGrid/DeGridW do not modify the caller's
    grid (the reference swaps its quadrants in place, IDG.cs:24,44)."""

    @staticmethod
    def Grid(c, metadata, visibilities, uvw, frequencies):
        return _plan(c, metadata, uvw, frequencies).grid(visibilities)

    @staticmethod
    def DeGrid(c, metadata, grid, uvw, frequencies):
        return _plan(c, metadata, uvw, frequencies).degrid(grid)

    @staticmethod
    def GridW(c, metadata, visibilities, uvw, frequencies):
        layers = _plan(c, metadata, uvw, frequencies).gridw(visibilities)
        return [layers[k] for k in range(layers.shape[0])]

    @staticmethod
    def DeGridW(c, metadata, grid, uvw, frequencies):
        return _plan(c, metadata, uvw, frequencies).degridw(grid)

    @staticmethod
    def GridPSF(c, metadata, uvw, flags, frequencies):
        return _plan(c, metadata, uvw, frequencies).grid_psf(flags)

    @staticmethod
    def ToImage(c, metadata, visibilities, uvw, frequencies):
        return _plan(c, metadata, uvw, frequencies).to_image(visibilities)

    @staticmethod
    def CalculatePSF(c, metadata, uvw, flags, frequencies):
        return _plan(c, metadata, uvw, frequencies).calculate_psf(flags)


def _square(a, name):
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("%s must be square (G, G), got shape %s" % (name, a.shape))


def _fft_plan(G, L=1):
    return Plan(GriddingConstants(1, G, 8 if G > 8 else 2, 0, 1, 1e-5, L, 0.0))


class FFT:
    """Static grid-FFT steps of Core/FFT.cs (G x G on the GPU, Shift on the host)."""

    @staticmethod
    def WStackIFFTFloat(grid, visibilityCount):
        layers = np.ascontiguousarray(np.stack(list(grid)) if isinstance(grid, (list, tuple)) else grid, dtype=np.complex128)
        if layers.ndim != 3 or layers.shape[1] != layers.shape[2]:
            raise ValueError("w-layers must have shape (L, G, G), got %s" % (layers.shape,))
        L, G = layers.shape[0], layers.shape[1]
        with _fft_plan(G, L) as p:
            out = np.zeros((G, G), dtype=np.float32)
            check(_lib.idg().p9idg_wstack_ifft_float(p._h, ptr(layers), L, int(visibilityCount), ptr(out)))
        return out

    @staticmethod
    def Forward(image, norm=1.0):
        im = np.ascontiguousarray(image, dtype=np.float32)
        _square(im, "image")
        G = im.shape[0]
        with _fft_plan(G) as p:
            out = np.zeros((G, G), dtype=np.complex128)
            check(_lib.idg().p9idg_fft_forward_float(p._h, ptr(im), float(norm), ptr(out)))
        return out

    @staticmethod
    def Backward(grid, visibilityCount):
        g = np.ascontiguousarray(grid, dtype=np.complex128)
        _square(g, "grid")
        G = g.shape[0]
        with _fft_plan(G) as p:
            out = np.zeros((G, G), dtype=np.float64)
            check(_lib.idg().p9idg_fft_backward(p._h, ptr(g), int(visibilityCount), ptr(out)))
        return out

    @staticmethod
    def Shift(a):
        """In-place quadrant swap (FFT.cs:221-264); accepts one array or a list."""
        if isinstance(a, (list, tuple)):
            for x in a:
                FFT.Shift(x)
            return a
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("Shift needs a C-contiguous array")
        _square(a, "array")
        n = a.shape[0]
        if a.dtype == np.float32:
            check(_lib.idg().p9idg_shift_f32(ptr(a), n))
        elif a.dtype == np.float64:
            check(_lib.idg().p9idg_shift_f64(ptr(a), n))
        elif a.dtype == np.complex128:
            check(_lib.idg().p9idg_shift_c128(ptr(a), n))
        else:
            raise TypeError("unsupported dtype %s" % a.dtype)
        return a


class Visibilities:
    @staticmethod
    def Substract(vis0, vis1, flag):
        """Common.cs:305-318: vis0 - vis1, 0 where flagged."""
        v0 = np.ascontiguousarray(vis0, dtype=np.complex128)
        v1 = np.ascontiguousarray(vis1, dtype=np.complex128)
        _expect(v1, v0.shape, "vis1")
        fl = _flags_arg(flag, v0.shape)
        out = np.zeros_like(v0)
        check(_lib.idg().p9idg_substract(ptr(v0), ptr(v1), ptr(fl), v0.size, ptr(out)))
        return out
