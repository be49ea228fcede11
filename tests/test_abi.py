"""C ABI of libp9idg.so / libp9obs.so without a GPU: the libraries load,
export every function their headers declare, and the host-only entry points
(partitioner, FFT.Shift, Visibilities.Substract, argument checking) behave as
the reference's (error texts of Partitioner.cs:86,91,118 and Adder.cs:49)."""
import ctypes
import os
import re

import numpy as np
import pytest

import p9idg
from p9idg import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def test_idg_library_exports_every_declared_function():
    lib = _lib.idg()
    names = declared("p9idg.h", "p9idg_")
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) == set(_lib.IDG_SIGNATURES)


def test_cycle_functions_exported():
    lib = _lib.idg()
    names = declared("p9cycle.h", "p9(?:psf|conv|cd|idg)_")
    assert len(names) == 17
    assert [n for n in names if not hasattr(lib, n)] == []
    assert set(names) == set(_lib.CYCLE_SIGNATURES)


def test_fits_functions_exported():
    lib = _lib.idg()
    names = declared("p9fits.h", "p9fits_")
    assert len(names) == 10
    assert [n for n in names if not hasattr(lib, n)] == []
    assert set(names) == set(_lib.FITS_SIGNATURES)


def test_obs_library_exports_every_declared_function():
    lib = _lib.obs()
    names = declared("p9obs.h", "p9obs_")
    assert names and all(hasattr(lib, n) for n in names)
    assert set(names) == set(_lib.OBS_SIGNATURES)


def test_api_version():
    assert _lib.idg().p9idg_api_version() == 1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_idg", None)
    monkeypatch.setattr(_lib, "IDG_LIB_PATH", "/nonexistent/libp9idg.so")
    with pytest.raises(RuntimeError, match="not built"):
        _lib.idg()


def _gpu_present():
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device error path")
def test_create_without_device_reports_hip_error():
    c = p9idg.GriddingConstants(100, 256, 32, 8, 1024, 7e-6, 1, 0.0)
    with pytest.raises(p9idg.P9IDGError) as e:
        p9idg.Plan(c)
    assert e.value.status == _lib.ERR_HIP


def test_create_rejects_unsupported_sizes():
    lib = _lib.idg()
    h = ctypes.c_void_p()
    # odd grid sizes are accepted (FFT.Shift leaves their last row and column)
    for G, S in [(0, 32), (256, 20), (32, 32)]:
        cs = p9idg.GriddingConstants(100, G, S, 8, 1024, 7e-6, 1, 0.0)._struct()
        st = lib.p9idg_create(ctypes.byref(cs), 0, ctypes.byref(h))
        assert st in (_lib.ERR_UNSUPPORTED, _lib.ERR_ARGUMENT)
        assert lib.p9idg_last_error()


def test_partition_argument_and_capacity_errors():
    lib = _lib.idg()
    c = p9idg.GriddingConstants(1, 256, 32, 8, 1024, 2.0 ** -10, 1, 0.0)._struct()
    uvw = np.zeros((2, 3, 3))
    freq = np.array([1e9, 1.1e9])
    n = ctypes.c_int64()
    assert lib.p9idg_partition(None, _lib.ptr(uvw), 2, 3, _lib.ptr(freq), 2, None, 0, ctypes.byref(n), None) == _lib.ERR_ARGUMENT
    assert lib.p9idg_partition(ctypes.byref(c), _lib.ptr(uvw), 2, 3, _lib.ptr(freq), 0, None, 0, ctypes.byref(n), None) == _lib.ERR_ARGUMENT
    assert lib.p9idg_partition(ctypes.byref(c), _lib.ptr(uvw), 2, 3, _lib.ptr(freq), 2, None, 0, ctypes.byref(n), None) == _lib.OK
    assert n.value == 2
    out = np.zeros(1, dtype=_lib.SUBGRID_DTYPE)
    assert lib.p9idg_partition(ctypes.byref(c), _lib.ptr(uvw), 2, 3, _lib.ptr(freq), 2, _lib.ptr(out), 1, ctypes.byref(n), None) == _lib.ERR_CAPACITY
    assert b"capacity" in lib.p9idg_last_error()


def test_partition_errors_carry_reference_text():
    c = p9idg.GriddingConstants(1, 256, 32, 8, 1024, 2.0 ** -10, 1, 0.0)
    with pytest.raises(p9idg.P9IDGError, match="could not place"):
        p9idg.Partitioner.CreatePartition(c, np.array([[[np.inf, 0, 0]]]), np.array([1e9]))
    with pytest.raises(p9idg.P9IDGError, match="subgrid falls not within grid"):
        p9idg.Partitioner.CreatePartition(c, np.array([[[1e9, 0, 0]]]), np.array([1e9]))


def test_partition_list_of_lists_view():
    from p9idg import observation as obs

    ob = obs.small()
    part = ob.partition()
    assert part.Count == ob.uvw.shape[0] == 435
    assert sum(len(part[b]) for b in range(part.Count)) == part.total
    for b in range(part.Count):
        assert np.all(part[b]["baselineIdx"] == b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex128])
def test_shift_host(dtype):
    a = (np.arange(64 * 64).reshape(64, 64) * (1 + 1j if dtype == np.complex128 else 1)).astype(dtype)
    b = p9idg.FFT.Shift(a.copy())
    assert np.array_equal(b, np.fft.fftshift(a))


def test_substract_host():
    rng = np.random.default_rng(1)
    v0 = rng.standard_normal((3, 4, 2)) + 1j * rng.standard_normal((3, 4, 2))
    v1 = rng.standard_normal((3, 4, 2)) + 1j * rng.standard_normal((3, 4, 2))
    fl = rng.random((3, 4, 2)) < 0.5
    out = p9idg.Visibilities.Substract(v0, v1, fl)
    assert np.array_equal(out, np.where(fl, 0, v0 - v1))


def test_gridding_constants_mirror():
    c = p9idg.GriddingConstants(12345, 3072, 32, 16, 1024, 1.5 / 3600 * np.pi / 180, 24, 12.5)
    assert c.ImageSize == np.float32(np.float32(1.5 / 3600 * np.pi / 180) * np.float32(3072))
    assert c.WStepLambda == 0 and c.SubgridSpheroidal.shape == (32, 32) and np.all(c.SubgridSpheroidal == 1)
    s = c._struct()
    assert (s.grid_size, s.subgrid_size, s.kernel_size, s.wlayer_count, s.visibilities_count) == (3072, 32, 16, 24, 12345)


def handle_functions():
    """(header, name, parameter count) of every entry point whose first
    parameter is a plan, batch, convolver or deconvolver handle."""
    out = []
    for header in ("p9idg.h", "p9cycle.h"):
        text = open(os.path.join(ROOT, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\bint\s+(p9\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
            params = [p.strip() for p in m.group(2).split(",") if p.strip()]
            if params and re.match(r"(const\s+)?p9(idg_plan|idg_batch|conv|cd)\s*\*\s*\w+$", params[0]):
                out.append((header, m.group(1), len(params)))
    return out


def test_null_handles_are_rejected():
    """Every entry point taking a handle returns P9IDG_ERR_ARGUMENT for NULL
    (destroy: OK) before touching the GPU, whatever its other arguments."""
    lib = _lib.idg()
    funcs = handle_functions()
    assert len(funcs) >= 60
    for header, name, n in funcs:
        f = getattr(lib, name)
        assert len(f.argtypes) == n, name
        pointer = lambda t: t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer)
        st = f(*[None if pointer(t) else 0 for t in f.argtypes])
        if name.endswith("_destroy"):
            assert st == _lib.OK, name
        else:
            assert st == _lib.ERR_ARGUMENT, (name, st)


def test_fits_null_arguments_are_rejected():
    """The FITS readers and writers return P9IDG_ERR_ARGUMENT for NULL paths
    and buffers instead of dereferencing them."""
    lib = _lib.idg()
    pointer = lambda t: t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer)
    names = [n for n in declared("p9fits.h", "p9fits_") if n != "p9fits_set_threads"]
    assert len(names) == 9
    for name in names:
        f = getattr(lib, name)
        assert f(*[None if pointer(t) else 0 for t in f.argtypes]) == _lib.ERR_ARGUMENT, name
