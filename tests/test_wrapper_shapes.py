"""The Python mirror refuses inputs whose shape differs from what the C ABI
reads (the reference's C# arrays carry their shapes and throw): every such
call raises ValueError before the library reads a byte.  Runs without a GPU:
the plan object has a NULL handle, so a call that passed the checks would
only get the library's "plan is NULL" status."""
import ctypes

import numpy as np
import pytest

import p9idg
from p9idg import _lib

G, B, T, C, L = 64, 3, 5, 4, 2


def fake_plan():
    plan = object.__new__(p9idg.Plan)
    plan.c = p9idg.GriddingConstants(B * T * C, G, 16, 4, 1024, 1e-5, L, 0.0)
    plan._h = ctypes.c_void_p()
    plan.shape = (B, T, C)
    return plan


VIS_CALLS = ["grid", "gridw", "to_image", "gridw_to_image", "upload_vis"]


@pytest.mark.parametrize("name", VIS_CALLS)
@pytest.mark.parametrize("shape", [(B, T, C + 1), (B * T * C,), (B, T - 1, C)])
def test_visibilities_of_the_wrong_shape_are_refused(name, shape):
    plan = fake_plan()
    with pytest.raises(ValueError, match="visibilities"):
        getattr(plan, name)(np.zeros(shape, np.complex128))
    # the right shape reaches the library, which refuses the NULL plan
    with pytest.raises(p9idg.P9IDGError) as e:
        getattr(plan, name)(np.zeros((B, T, C), np.complex64))
    assert e.value.status == _lib.ERR_ARGUMENT


@pytest.mark.parametrize("name", ["degrid", "degridw", "upload_model_grid"])
def test_grids_of_the_wrong_shape_are_refused(name):
    plan = fake_plan()
    for shape in [(G, G - 2), (G + 2, G + 2), (L, G, G), (G * G,)]:
        with pytest.raises(ValueError, match="grid"):
            getattr(plan, name)(np.zeros(shape, np.complex128))


@pytest.mark.parametrize("name", ["image_to_vis", "upload_image"])
def test_images_of_the_wrong_shape_are_refused(name):
    plan = fake_plan()
    with pytest.raises(ValueError, match="image"):
        getattr(plan, name)(np.zeros((G, G // 2), np.float32))


@pytest.mark.parametrize("name", ["grid_psf", "calculate_psf", "set_psf_vis", "set_flags"])
def test_flags_of_the_wrong_shape_are_refused(name):
    plan = fake_plan()
    with pytest.raises(ValueError, match="flags"):
        getattr(plan, name)(np.zeros((B, T, C - 1), bool))
    with pytest.raises(p9idg.P9IDGError):  # NULL flags or the right shape pass the check
        getattr(plan, name)(None)


def test_uvw_and_frequencies_are_checked():
    plan = fake_plan()
    meta = np.zeros(0, dtype=p9idg.SUBGRID_DTYPE)
    with pytest.raises(ValueError, match="uvw"):
        plan.set_observation(np.zeros((B, T, 2)), np.ones(C), meta)
    with pytest.raises(ValueError, match="uvw"):
        plan.set_observation(np.zeros((B * T, 3)), np.ones(C), meta)
    with pytest.raises(ValueError, match="frequencies"):
        plan.set_observation(np.zeros((B, T, 3)), np.ones((C, 1)), meta)


def test_major_cycle_checks_visibilities_and_flags():
    plan = fake_plan()
    with pytest.raises(ValueError, match="visibilities"):
        p9idg.MajorCycle.ReconstructSerialCD(plan, np.zeros((B, T, C + 1), np.complex64), None, 2, 1, 1.0, 0.01, 10, 1e-5)
    with pytest.raises(ValueError, match="flags"):
        p9idg.MajorCycle.ReconstructSerialCD(plan, np.zeros((B, T, C), np.complex64), np.zeros((B, T), bool), 2, 1,
                                             1.0, 0.01, 10, 1e-5)


def test_fft_steps_need_square_arrays():
    with pytest.raises(ValueError):
        p9idg.FFT.Forward(np.zeros((8, 6), np.float32))
    with pytest.raises(ValueError):
        p9idg.FFT.Backward(np.zeros((8, 6), np.complex128), 10)
    with pytest.raises(ValueError):
        p9idg.FFT.WStackIFFTFloat(np.zeros((2, 8, 6), np.complex128), 10)
    with pytest.raises(ValueError):
        p9idg.FFT.Shift(np.zeros((8, 6), np.float32))


def test_substract_checks_shapes():
    v = np.zeros((2, 3, 4), np.complex128)
    with pytest.raises(ValueError, match="vis1"):
        p9idg.Visibilities.Substract(v, np.zeros((2, 3, 5)), None)
    with pytest.raises(ValueError, match="flags"):
        p9idg.Visibilities.Substract(v, v, np.zeros((2, 3), bool))
    assert np.array_equal(p9idg.Visibilities.Substract(v + 1, v, np.zeros(v.shape, bool)), v + 1)
