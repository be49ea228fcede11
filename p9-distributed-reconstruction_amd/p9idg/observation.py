"""Seeded synthetic observations for the benchmark and test workloads.

Each workload is fully determined by its name and seed (seed = workload
index * 1000 + instance).  Geometry and sky conventions are documented in
include/p9obs.h; in short: antennas uniform in disks given by DIAMETER,
Earth-rotation uvw centred on transit at latitude -30.7 deg, declination
-60 deg, and a point/Gaussian-source sky whose source at pixel offset
(dx, dy) appears at dirty-image pixel [G/2 + dy, G/2 + dx].

Workloads:
  "small"   256^2 image, 30 antennas (435 baselines), 60 x 10 s, 8 channels
            1.00-1.07 GHz, 2 point sources + noise 0.05, complex128, L = 1.
  "large"   3072^2 image, 64 antennas (45 in a 1 km core, 19 in an 8 km
            disk; 2016 baselines), 2000 x 8 s, 8 channels 1.0-1.4 GHz,
            500 log-uniform point sources (1 mJy-1 Jy) + 20 Gaussians
            (sigma 2-10 px, 10-100 mJy) + noise 0.01, complex64, L = 32,
            wStep = lambda(max |w|, f_max) / L (RunningMethod.cs:48-53).
  "batch"   one of 8 uneven instances: 16-64 antennas, 200-2000 timesteps,
            4-16 channels over 1.0-1.4 GHz, image 512/1024/2048, subgrid
            24/32/48, L = 8, point-source sky as "large", complex64.
Every workload has S-K margin 8 (kernel 8) and MaxTimesteps 1024
(RunningMethod.cs:21,44), cell 1.5 arcsec.
"""
import ctypes
import math
import os

import numpy as np

from . import _lib
from .idg import GriddingConstants, Partitioner

C_LIGHT = 299792458.0
ARCSEC = math.pi / 180.0 / 3600.0
LATITUDE = math.radians(-30.7)
DECLINATION = math.radians(-60.0)


def _uniform(seed, n):
    out = np.zeros(n, dtype=np.float64)
    _lib.obs().p9obs_uniform(seed, n, _lib.ptr(out))
    return out


class Observation:
    def __init__(self, name, constants, uvw, freq, vis, sources, seed, info):
        self.name = name
        self.c = constants
        self.uvw = uvw
        self.freq = freq
        self.vis = vis
        self.sources = sources
        self.seed = seed
        self.info = info
        self._partition = None

    @property
    def shape(self):
        return self.vis.shape

    def partition(self):
        if self._partition is None:
            self._partition = Partitioner.CreatePartition(self.c, self.uvw, self.freq)
        return self._partition

    def describe(self):
        d = dict(self.info)
        d.update(
            workload=self.name,
            seed=self.seed,
            baselines=int(self.uvw.shape[0]),
            timesteps=int(self.uvw.shape[1]),
            channels=int(len(self.freq)),
            grid=self.c.GridSize,
            subgrid=self.c.SubgridSize,
            kernel=self.c.KernelSize,
            wlayers=self.c.WLayerCount,
        )
        return d


def _antennas(seed, n_core, core_diam, n_outer, outer_diam):
    enu = np.zeros((n_core + n_outer, 3), dtype=np.float64)
    _lib.obs().p9obs_antennas(seed, n_core, core_diam, n_outer, outer_diam, _lib.ptr(enu))
    return enu


def _uvw(enu, T, dt):
    n = enu.shape[0]
    B = n * (n - 1) // 2
    uvw = np.zeros((B, T, 3), dtype=np.float64)
    _lib.obs().p9obs_uvw(n, _lib.ptr(enu), LATITUDE, DECLINATION, T, dt, _lib.ptr(uvw))
    return uvw


def _sky(uvw, freq, sources, threads=0):
    B, T = uvw.shape[0], uvw.shape[1]
    C = len(freq)
    src = (_lib.Source * len(sources))(*[_lib.Source(*s) for s in sources])
    vis = np.zeros((B, T, C), dtype=np.complex128)
    _lib.obs().p9obs_sky_vis(B, T, C, _lib.ptr(uvw), _lib.ptr(freq), len(sources), ctypes.cast(src, ctypes.c_void_p), _lib.ptr(vis), threads)
    return vis


def _noise(vis, sigma, seed):
    _lib.obs().p9obs_add_noise(vis.size, sigma, seed, _lib.ptr(vis))


def _even_freq(f0, f1, C):
    df = (f1 - f0) / (C - 1) if C > 1 else 0.0
    return np.array([f0 + ch * df for ch in range(C)], dtype=np.float64)


def _wstep(uvw, freq, L):
    # RunningMethod.cs:48-53
    max_w = float(np.max(np.abs(uvw[:, :, 2]))) if uvw.size else 0.0
    return Partitioner.MetersToLambda(max_w, freq[-1]) / L


def _point_sources(seed, n_points, n_blobs, G, cell):
    u = _uniform(seed, 3 * n_points + 4 * n_blobs)
    sources = []
    for i in range(n_points):
        flux = 10.0 ** (-3.0 + 3.0 * u[3 * i])
        dx = (u[3 * i + 1] - 0.5) * 0.9 * G
        dy = (u[3 * i + 2] - 0.5) * 0.9 * G
        sources.append((flux, dx * cell, dy * cell, 0.0))
    o = 3 * n_points
    for i in range(n_blobs):
        flux = 0.01 + 0.09 * u[o + 4 * i]
        sigma_px = 2.0 + 8.0 * u[o + 4 * i + 1]
        dx = (u[o + 4 * i + 2] - 0.5) * 0.8 * G
        dy = (u[o + 4 * i + 3] - 0.5) * 0.8 * G
        sources.append((flux, dx * cell, dy * cell, sigma_px * cell))
    return sources


def small(instance=0, threads=0):
    """256^2, 30 antennas, 60 timesteps, 8 channels 1.00-1.07 GHz, complex128."""
    seed = 1000 + instance
    G, S, K, L = 256, 32, 8, 1
    cell = 1.5 * ARCSEC
    enu = _antennas(seed, 30, 1000.0, 0, 0.0)
    uvw = _uvw(enu, 60, 10.0)
    freq = np.array([1.00e9 + 0.01e9 * ch for ch in range(8)], dtype=np.float64)
    sources = [(1.0, 20 * cell, -12 * cell, 0.0), (0.5, -35 * cell, 40 * cell, 0.0)]
    vis = _sky(uvw, freq, sources, threads)
    _noise(vis, 0.05, seed + 500)
    c = GriddingConstants(vis.size, G, S, K, 1024, cell, L, 0.0)
    return Observation("small", c, uvw, freq, vis, sources, seed, {"noise": 0.05, "vis_dtype": "complex128"})


def _cache_dir():
    return os.environ.get("P9IDG_CACHE", "/tmp/p9idg_cache")


def large(instance=0, threads=0, n_points=500, n_blobs=20, cache=True):
    """3072^2, 64 antennas (45 core 1 km + 19 in 8 km), 2000 x 8 s, 8 ch 1.0-1.4 GHz, L=32.

    The sky sum (32M visibilities x 520 sources) takes tens of seconds, so the
    visibilities are cached under $P9IDG_CACHE (default /tmp/p9idg_cache)."""
    seed = 2000 + instance
    G, S, K, L = 3072, 32, 8, 32
    cell = 1.5 * ARCSEC
    enu = _antennas(seed, 45, 1000.0, 19, 8000.0)
    uvw = _uvw(enu, 2000, 8.0)
    freq = _even_freq(1.0e9, 1.4e9, 8)
    sources = _point_sources(seed + 100, n_points, n_blobs, G, cell)
    path = os.path.join(_cache_dir(), "large_vis_%d_%d_%d.npy" % (seed, n_points, n_blobs))
    vis = None
    if cache and os.path.exists(path):
        try:
            vis = np.load(path, allow_pickle=False)
            if vis.shape != (uvw.shape[0], uvw.shape[1], len(freq)) or vis.dtype != np.complex64:
                vis = None
        except (OSError, ValueError):
            vis = None
    if vis is None:
        vis = _sky(uvw, freq, sources, threads)
        _noise(vis, 0.01, seed + 500)
        vis = vis.astype(np.complex64)
        if cache:
            try:
                os.makedirs(_cache_dir(), exist_ok=True)
                tmp = path + ".%d.tmp" % os.getpid()
                with open(tmp, "wb") as f:
                    np.save(f, vis, allow_pickle=False)
                os.replace(tmp, path)
            except OSError:
                pass
    wstep = _wstep(uvw, freq, L)
    c = GriddingConstants(vis.size, G, S, K, 1024, cell, L, wstep)
    return Observation("large", c, uvw, freq, vis, sources, seed, {"noise": 0.01, "vis_dtype": "complex64", "wstep": wstep})


def batch_instance(instance, threads=0):
    """One of the 8 uneven instances (parameters drawn from the seed)."""
    seed = 4000 + instance
    u = _uniform(seed, 8)
    n_ant = 16 + int(u[0] * 49)  # 16..64
    T = 200 + int(u[1] * 1801)  # 200..2000
    C = 4 + int(u[2] * 13)  # 4..16
    G = (512, 1024, 2048)[min(2, int(u[3] * 3))]
    S = (24, 32, 48)[min(2, int(u[4] * 3))]
    K, L = 8, 8
    cell = 1.5 * ARCSEC
    n_core = int(round(0.7 * n_ant))
    enu = _antennas(seed, n_core, 1000.0, n_ant - n_core, 8000.0)
    uvw = _uvw(enu, T, 8.0)
    freq = _even_freq(1.0e9, 1.4e9, C)
    sources = _point_sources(seed + 100, 50, 0, G, cell)
    vis = _sky(uvw, freq, sources, threads)
    _noise(vis, 0.01, seed + 500)
    vis = vis.astype(np.complex64)
    wstep = _wstep(uvw, freq, L)
    c = GriddingConstants(vis.size, G, S, K, 1024, cell, L, wstep)
    return Observation("batch%d" % instance, c, uvw, freq, vis, sources, seed, {"vis_dtype": "complex64", "wstep": wstep})


def custom(G=256, S=32, K=8, L=1, n_ant=12, T=20, C=4, f0=1.0e9, f1=1.1e9, cell_arcsec=1.5, core_diam=1000.0,
           sources=None, noise=0.0, seed=7, dt=10.0, wstep=None, vis_dtype=np.complex128, threads=0):
    """Small parameterised observation for tests."""
    cell = cell_arcsec * ARCSEC
    enu = _antennas(seed, n_ant, core_diam, 0, 0.0)
    uvw = _uvw(enu, T, dt)
    freq = _even_freq(f0, f1, C) if C > 1 else np.array([f0])
    if sources is None:
        sources = [(1.0, 10 * cell, -6 * cell, 0.0)]
    vis = _sky(uvw, freq, sources, threads)
    _noise(vis, noise, seed + 500)
    vis = vis.astype(vis_dtype)
    if wstep is None:
        wstep = _wstep(uvw, freq, L) if L > 1 else 0.0
    c = GriddingConstants(vis.size, G, S, K, 1024, cell, L, wstep)
    return Observation("custom", c, uvw, freq, vis, sources, seed, {})
