"""ctypes binding of libp9idg.so (the HIP library) and libp9obs.so.

The binding refuses to run without the built HIP library: there is no CPU
fallback for any gridding entry point.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# P9IDG_LIB_DIR selects another build of the two libraries (the sanitizer
# build in _build/asan); by default the in-tree ones next to the package
LIB_DIR = os.environ.get("P9IDG_LIB_DIR") or os.path.dirname(PKG_DIR)
IDG_LIB_PATH = os.path.join(LIB_DIR, "libp9idg.so")
OBS_LIB_PATH = os.path.join(LIB_DIR, "libp9obs.so")

# p9idg_status (include/p9idg.h)
OK = 0
ERR_PLACE = 1
ERR_SUBGRID_RANGE = 2
ERR_INVALID_GRID = 3
ERR_ARGUMENT = 4
ERR_CAPACITY = 5
ERR_HIP = 6
ERR_ALLOC = 7
ERR_STATE = 8
ERR_UNSUPPORTED = 9
ERR_IO = 10

C128 = 0
C64 = 1

KERNEL_FP32 = 0
KERNEL_MFMA_BF16 = 1
KERNEL_MFMA_FP16 = 2
KERNEL_FP32_SEP = 3

STAGE_GRIDDER, STAGE_DEGRIDDER, STAGE_ZERO, STAGE_TO_IMAGE, STAGE_TO_UVW = 0, 1, 2, 3, 4
GRID_PLAIN = 0
GRID_WSTACK = 1
GRID_WSUM = 2
DEGRID_PLAIN = 0
DEGRID_WSTACK = 1
SRC_VIS = 0
SRC_RESIDUAL = 1

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


class P9IDGError(Exception):
    """Raised with the library's message; the three reference exceptions keep
    their texts (Partitioner.cs:86,91,118; Adder.cs:49)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


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


class Source(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_double), ("l", ctypes.c_double), ("m", ctypes.c_double), ("sigma", ctypes.c_double)]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_f64 = ctypes.c_double
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes; every one returns int (p9idg_status) unless listed in _RESTYPES
IDG_SIGNATURES = {
    "p9idg_api_version": [],
    "p9idg_last_error": [],
    "p9idg_partition": [_vp, _vp, _i32, _i32, _vp, _i32, _vp, _i64, _pi64, _vp],
    "p9idg_create": [_vp, _i32, ctypes.POINTER(_vp)],
    "p9idg_get_constants": [_vp, _vp, ctypes.POINTER(_i32)],
    "p9idg_destroy": [_vp],
    "p9idg_set_kernel": [_vp, _i32],
    "p9idg_set_deterministic": [_vp, _i32],
    "p9idg_get_kernel": [_vp, ctypes.POINTER(_i32)],
    "p9idg_set_observation": [_vp, _vp, _i32, _i32, _vp, _i32, _vp, _i64],
    "p9idg_set_launch": [_vp, _i32, _i32, _i32, _i32],
    "p9idg_autotune": [_vp, _i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "p9idg_get_launch": [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "p9idg_grid": [_vp, _vp, _i32, _vp],
    "p9idg_degrid": [_vp, _vp, _vp],
    "p9idg_gridw": [_vp, _vp, _i32, _vp],
    "p9idg_degridw": [_vp, _vp, _vp],
    "p9idg_grid_psf": [_vp, _vp, _vp],
    "p9idg_to_image": [_vp, _vp, _i32, _vp],
    "p9idg_calculate_psf": [_vp, _vp, _vp],
    "p9idg_gridw_to_image": [_vp, _vp, _i32, _vp],
    "p9idg_image_to_vis": [_vp, _vp, _vp],
    "p9idg_wstack_ifft_float": [_vp, _vp, _i32, _i64, _vp],
    "p9idg_fft_forward_float": [_vp, _vp, _f64, _vp],
    "p9idg_fft_backward": [_vp, _vp, _i64, _vp],
    "p9idg_shift_f32": [_vp, _i32],
    "p9idg_shift_f64": [_vp, _i32],
    "p9idg_shift_c128": [_vp, _i32],
    "p9idg_substract": [_vp, _vp, _vp, _i64, _vp],
    "p9idg_upload_vis": [_vp, _vp, _i32],
    "p9idg_set_psf_vis": [_vp, _vp],
    "p9idg_exec_grid": [_vp, _i32],
    "p9idg_exec_grid_to_image": [_vp, _i32],
    "p9idg_exec_image_to_grid": [_vp],
    "p9idg_exec_degrid": [_vp, _i32],
    "p9idg_sync": [_vp],
    "p9idg_set_flags": [_vp, _vp],
    "p9idg_exec_residual": [_vp],
    "p9idg_set_grid_source": [_vp, _i32],
    "p9idg_exec_cycle": [_vp, _i32],
    "p9idg_capture_cycle": [_vp, _i32],
    "p9idg_launch_cycle": [_vp],
    "p9idg_download_residual": [_vp, _vp],
    "p9idg_device_buffers": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "p9idg_download_grid": [_vp, _i32, _vp],
    "p9idg_upload_model_grid": [_vp, _vp],
    "p9idg_download_model_grid": [_vp, _vp],
    "p9idg_download_model_vis": [_vp, _vp],
    "p9idg_upload_image": [_vp, _vp],
    "p9idg_download_image": [_vp, _vp],
    "p9idg_last_timings": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "p9idg_timing_reset": [_vp, _i32],
    "p9idg_timing_totals": [_vp, ctypes.POINTER(_f64), _pi64, ctypes.POINTER(_f64), _pi64],
    "p9idg_timing_stage": [_vp, _i32, ctypes.POINTER(_f64), _pi64],
    "p9idg_grid_noalloc": [ctypes.POINTER(Constants), _i32, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _vp,
                           _vp],
    "p9idg_gridded_visibilities": [_vp, _pi64],
    "p9idg_batch_create": [_vp, _i32, ctypes.POINTER(_vp)],
    "p9idg_batch_destroy": [_vp],
    "p9idg_batch_prepare": [_vp, _i32, _i32],
    "p9idg_batch_exec": [_vp, _i32],
    "p9idg_batch_capture": [_vp, _i32],
    "p9idg_batch_launch": [_vp],
    "p9idg_batch_sync": [_vp],
    "p9idg_batch_timings": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
}
_f32 = ctypes.c_float
_pf32 = ctypes.POINTER(ctypes.c_float)
# p9cycle.h (same library)
CYCLE_SIGNATURES = {
    "p9psf_cut": [_vp, _i32, _i32, _vp],
    "p9psf_max_sidelobe": [_vp, _i32, _i32, _pf32],
    "p9psf_max_lipschitz": [_vp, _i32, _i32, ctypes.POINTER(_f64)],
    "p9psf_squared": [_i32, _vp, _i32, _i32, _vp],
    "p9psf_amap": [_i32, _vp, _i32, _i32, _i32, _i32, _vp],
    "p9conv_create_correlation": [_i32, _vp, _i32, _i32, _i32, _i32, ctypes.POINTER(_vp)],
    "p9conv_destroy": [_vp],
    "p9conv_convolve": [_vp, _vp, _vp],
    "p9conv_convolve_device": [_vp, _vp, _vp],
    "p9cd_create": [_i32, _i32, _i32, _vp, _i32, _i32, ctypes.POINTER(_vp)],
    "p9cd_destroy": [_vp],
    "p9cd_max_lipschitz": [_vp, _pf32],
    "p9cd_get_abs_max_diff": [_vp, _vp, _vp, _f32, _f32, _pf32],
    "p9cd_get_abs_max_diff_device": [_vp, _vp, _vp, _f32, _f32, _pf32],
    "p9cd_deconvolve": [_vp, _vp, _vp, _f32, _f32, _i32, _f32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "p9cd_deconvolve_device": [_vp, _vp, _vp, _f32, _f32, _i32, _f32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "p9idg_reconstruct_serial_cd": [_vp, ctypes.c_char_p, _vp, _i32, _vp, _i32, _i32, _f32, _f32, _i32, _f32, _vp,
                                    _vp, ctypes.POINTER(_i32)],
}

# include/p9fits.h (host code of libp9idg.so: callable without a GPU)
FITS_SIGNATURES = {
    "p9fits_set_threads": [_i32],
    "p9fits_read_info": [ctypes.c_char_p, _vp],
    "p9fits_read_frequencies": [ctypes.c_char_p, _vp, _i64, _pi64],
    "p9fits_count_baselines": [ctypes.c_char_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "p9fits_read_uvw": [ctypes.c_char_p, _i32, _i32, _vp],
    "p9fits_read_visibilities": [ctypes.c_char_p, _i32, _i32, _i32, _i32, _f64, _vp, _i32],
    "p9fits_read_flags": [ctypes.c_char_p, _i32, _i32, _i32, _i32, _vp],
    "p9fits_read_image": [ctypes.c_char_p, _i32, _vp, _i32, _i32],
    "p9fits_read_casa": [ctypes.c_char_p, _vp, _i32, _i32],
    "p9fits_write_image": [ctypes.c_char_p, _i32, _vp, _i32, _i32],
}
FITS_MAX_AXES = 8
FITS_F32 = 0
FITS_F64 = 1
FITS_C128_REAL = 2
FITS_C128_IMAG = 3


class FitsInfo(ctypes.Structure):
    _fields_ = [
        ("bitpix", ctypes.c_int32),
        ("naxis", ctypes.c_int32),
        ("naxes", ctypes.c_int64 * FITS_MAX_AXES),
        ("data_offset", ctypes.c_int64),
        ("data_bytes", ctypes.c_int64),
    ]
_RESTYPES = {"p9idg_last_error": ctypes.c_char_p}

OBS_SIGNATURES = {
    "p9obs_uniform": ([_u64, _i64, _vp], None),
    "p9obs_normal": ([_u64, _i64, _vp], None),
    "p9obs_antennas": ([_u64, _i32, _f64, _i32, _f64, _vp], None),
    "p9obs_uvw": ([_i32, _vp, _f64, _f64, _i32, _f64, _vp], None),
    "p9obs_sky_vis": ([_i32, _i32, _i32, _vp, _vp, _i32, _vp, _vp, _i32], None),
    "p9obs_add_noise": ([_i64, _f64, _u64, _vp], None),
}

_idg = None
_obs = None


def idg():
    """The HIP library; raises if it has not been built."""
    global _idg
    if _idg is None:
        if not os.path.exists(IDG_LIB_PATH):
            raise RuntimeError(
                "libp9idg.so is not built (%s); run __graft_entry__.build() or make -C %s"
                % (IDG_LIB_PATH, LIB_DIR)
            )
        lib = ctypes.CDLL(IDG_LIB_PATH)
        for table in (IDG_SIGNATURES, CYCLE_SIGNATURES, FITS_SIGNATURES):
            for name, args in table.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _idg = lib
    return _idg


def obs():
    global _obs
    if _obs is None:
        if not os.path.exists(OBS_LIB_PATH):
            raise RuntimeError("libp9obs.so is not built (%s)" % OBS_LIB_PATH)
        lib = ctypes.CDLL(OBS_LIB_PATH)
        for name, (args, res) in OBS_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _obs = lib
    return _obs


def check(status):
    if status != OK:
        msg = idg().p9idg_last_error()
        raise P9IDGError(status, msg.decode() if msg else "p9idg error %d" % status)


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
