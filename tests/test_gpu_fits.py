"""A measurement set in the reference's FITS layout (MeasurementData.LoadLMC:
freq.fits and eight parts uvw<i>.fits, flags<i>.fits, vis<i>.fits, FitsIO.cs
:184-219, :267-294, :327-357) written from literal cards, loaded through the
native readers into a plan, gridded and degridded on the GPU, against the
oracle on the same files decoded by the numpy restatement (oracle_fits)."""
import os

import numpy as np
import pytest

import oracle_fits as OF
import p9idg
from p9idg import observation as obs

from conftest import oracle_constants, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def write_lmc(folder, ob, flags, rng):
    """ob's observation as eight parts along the baselines: v stored negated
    (ReadUVW flips it back), XX and YY straddling the visibility (their mean
    is what the reader keeps), XY and YX noise, and per-polarization flag
    values whose ordered sum is nonzero exactly where `flags` is set."""
    B, T, C = ob.vis.shape
    OF.write(os.path.join(folder, "freq.fits"), ob.freq, -64)
    cuts = np.linspace(0, B, 9).astype(int)
    for i in range(8):
        b0, b1 = cuts[i], cuts[i + 1]
        uvw = ob.uvw[b0:b1].copy()
        uvw[..., 1] = -uvw[..., 1]
        d = rng.standard_normal(ob.vis[b0:b1].shape) * 0.1
        pols = rng.standard_normal((b1 - b0, T, C, 4, 2))
        pols[..., 0, 0] = ob.vis[b0:b1].real + d
        pols[..., 0, 1] = ob.vis[b0:b1].imag - d
        pols[..., 3, 0] = ob.vis[b0:b1].real - d
        pols[..., 3, 1] = ob.vis[b0:b1].imag + d
        fl = np.zeros((b1 - b0, T, C, 4))
        fl[..., 2] = np.where(flags[b0:b1], 1.0, 0.0)
        OF.write(os.path.join(folder, "uvw%d.fits" % i), uvw, -64)
        OF.write(os.path.join(folder, "vis%d.fits" % i), pols, -64)
        OF.write(os.path.join(folder, "flags%d.fits" % i), fl, -64)
    return cuts


@pytest.mark.parametrize("kernel", ["mfma_bf16", "fp32"])
def test_lmc_measurement_set_gridded_and_degridded(tmp_path, O, kernel):
    ob = obs.custom(G=256, S=32, K=8, L=4, n_ant=12, T=30, C=8, seed=77, core_diam=1500.0, noise=0.01)
    B, T, C = ob.vis.shape
    rng = np.random.default_rng(17)
    flags = rng.random((B, T, C)) < 0.05
    cuts = write_lmc(str(tmp_path), ob, flags, rng)

    d = p9idg.MeasurementData.LoadLMC(str(tmp_path))
    # the numpy restatement's decode of the same files
    freq = OF.read_frequencies(tmp_path / "freq.fits")
    uvw = np.concatenate([OF.read_uvw(tmp_path / ("uvw%d.fits" % i)) for i in range(8)])
    vis = np.concatenate([OF.read_visibilities(tmp_path / ("vis%d.fits" % i), 0, cuts[i + 1] - cuts[i], T, C, 2.0)
                          for i in range(8)])
    fl = np.concatenate([OF.read_flags(tmp_path / ("flags%d.fits" % i), 0, cuts[i + 1] - cuts[i], T, C)
                         for i in range(8)])
    assert np.array_equal(d.UVW, uvw) and np.array_equal(d.Frequencies, freq)
    assert np.array_equal(d.Visibilities, vis) and np.array_equal(d.Flags, fl)
    assert np.array_equal(fl, flags) and d.VisibilitiesCount == int((~flags).sum())
    assert np.abs(uvw - ob.uvw).max() == 0.0

    c = p9idg.GriddingConstants(d.VisibilitiesCount, 256, 32, 8, 1024, ob.c.ScaleArcSec, 4, ob.c.WStep)
    part = p9idg.Partitioner.CreatePartition(c, d.UVW, d.Frequencies)
    oc = oracle_constants(c)
    m = part.subgrids
    assert (m["WLambda"] < 0).any()
    ref_l = O.gridw(oc, m, vis, uvw, freq)
    model = ref_l.sum(axis=0)
    ref_v = O.degridw(oc, m, model, uvw, freq)
    psf_vis = np.where(fl, 0.0, 1.0).astype(np.complex128)  # MajorCycle.cs:36-43
    ref_psf = O.gridw_to_image(oc, m, psf_vis, uvw, freq)
    with p9idg.Plan(c, kernel=kernel) as plan:
        plan.set_observation(d.UVW, d.Frequencies, part)
        e_l = rel_err(plan.gridw(d.Visibilities), ref_l)
        e_v = rel_err(plan.degridw(model), ref_v)
        plan.set_psf_vis(d.Flags)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        e_p = rel_err(plan.download_image(), ref_psf)
    print(kernel, "measurement set: gridw %.2e degridw %.2e psf %.2e" % (e_l, e_v, e_p))
    assert e_l <= TOL and e_v <= TOL and e_p <= TOL
