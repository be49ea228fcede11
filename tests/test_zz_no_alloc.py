"""p9idg_grid_noalloc through the C ABI against p9idg_exec_grid (the plan's
GridW / Grid with the same kernels) on the same inputs, two calls in a row,
the second with a smaller observation: both within the parity bar (1e-5 of
the maximum; the two differ only in the float atomics' summation order).
The test allocates its device buffers with the HIP runtime the library
uses; the entry point itself allocates nothing."""
import ctypes

import numpy as np
import pytest

import p9idg
from p9idg import _lib
from p9idg import observation as obs

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libp9idg.so itself
    loaded (found in /proc/self/maps), for the test's own device buffers."""

    H2D, D2H = 1, 2

    def __init__(self):
        _lib.idg()
        path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line)
        self.lib = ctypes.CDLL(path)
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))) == 0
        self.bufs.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), self.H2D) == 0
        return p

    def download(self, p, out):
        assert self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def noalloc(ob, part, kernel, mode):
    G, L = ob.c.GridSize, ob.c.WLayerCount
    B, T, C = ob.vis.shape
    shape = (L, G, G) if mode == p9idg.GRID_WSTACK else (G, G)
    hip = Hip()
    try:
        d_uvw = hip.upload(ob.uvw)
        d_vis = hip.upload(ob.vis.astype(np.complex64))
        d_out = hip.upload(np.full(shape, complex(1e30, -1e30), np.complex64))  # must be zeroed by the call
        meta = np.ascontiguousarray(part.subgrids)
        freq = np.ascontiguousarray(ob.freq, dtype=np.float64)
        cs = ob.c._struct()
        k = {"fp32": _lib.KERNEL_FP32, "fp32_sep": _lib.KERNEL_FP32_SEP, "mfma_bf16": _lib.KERNEL_MFMA_BF16,
             "mfma_fp16": _lib.KERNEL_MFMA_FP16}[kernel]
        _lib.check(_lib.idg().p9idg_grid_noalloc(ctypes.byref(cs), k, int(mode), _lib.ptr(meta), len(meta), d_uvw, B, T,
                                                 _lib.ptr(freq), C, d_vis, d_out, None))
        return hip.download(d_out, np.empty(shape, np.complex64))  # hipMemcpy waits for the null stream
    finally:
        hip.free()


def with_plan(ob, part, kernel, mode):
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.exec_grid(mode)
        return plan.download_grid(mode)


@pytest.mark.parametrize("kernel", ["mfma_bf16", "fp32_sep"])
def test_noalloc_matches_the_plan_twice_in_a_row(kernel):
    big = obs.custom(G=512, S=32, K=8, L=4, n_ant=16, T=80, C=8, seed=501, core_diam=2000.0)
    small = obs.custom(G=256, S=32, K=8, L=4, n_ant=8, T=30, C=6, seed=502, core_diam=1200.0)
    for ob in (big, small):
        part = ob.partition()
        assert (part.subgrids["WLambda"] < 0).any()
        for mode in (p9idg.GRID_WSTACK, p9idg.GRID_WSUM):
            got = noalloc(ob, part, kernel, mode)
            ref = with_plan(ob, part, kernel, mode)
            e = rel_err(got, ref)
            print(kernel, ob.c.GridSize, mode, "noalloc vs plan %.2e" % e)
            assert e <= TOL, (ob.c.GridSize, mode, e)


def test_noalloc_rejects_what_does_not_fit():
    ob = obs.custom(G=128, S=16, K=4, L=1, n_ant=4, T=4, C=2, seed=503)
    cs = ob.c._struct()
    meta = np.zeros(70000, dtype=p9idg.SUBGRID_DTYPE)
    freq = np.ones(5000)
    lib = _lib.idg()
    fake = 256  # never dereferenced: the checks come first
    assert lib.p9idg_grid_noalloc(ctypes.byref(cs), 1, 1, _lib.ptr(meta), 70000, fake, 4, 4, _lib.ptr(freq), 2, fake,
                                  fake, None) == _lib.ERR_CAPACITY
    assert b"65536" in lib.p9idg_last_error()
    assert lib.p9idg_grid_noalloc(ctypes.byref(cs), 1, 1, _lib.ptr(meta), 1, fake, 4, 4, _lib.ptr(freq), 5000, fake,
                                  fake, None) == _lib.ERR_CAPACITY
