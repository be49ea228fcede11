"""The two benchmark configurations that need whole-workload checks, on the
GPU against the FP64 oracle.

* The 3072^2, 32-w-layer workload ("large": 2016 baselines x 2000 x 8
  channels, 16.2 M gridded visibilities) with the split-BF16 and split-FP16
  MFMA kernels at the autotuned launch shape: the whole dirty image
  (GridW -> FFT.WStackIFFTFloat, MajorCycle.cs:88-90), the whole PSF
  (MajorCycle.cs:36-48) and every degridded visibility (ToUVW -> DeGridW,
  MajorCycle.cs:125-128), max and RMS relative error <= 1e-5.  The oracle
  references take about a minute of 16 host threads and are computed once.
* The 8 batched uneven instances ("batch_instance" 0..7: subgrids 24 / 32 /
  48, images 512-2048, 4-16 channels), autotuned, run as one Batch eagerly
  and replayed from a hipGraph (IDG.cs:32-50 per instance): GridW layers and
  DeGridW visibilities against the oracle on a subgrid sample of every
  instance, and the full instances' degridded visibilities on the sample.

Error = max |gpu - oracle| / max |oracle| (and the RMS ratio), norm-wise."""
import time

import numpy as np
import pytest

import p9idg
from p9idg import observation as obs

from conftest import oracle_constants

pytestmark = pytest.mark.gpu

TOL = 1e-5


def errs(got, ref, mask=None):
    d = np.abs(got.astype(np.complex128) - ref)
    r = np.abs(ref)
    if mask is not None:
        d, r = d[mask], r[mask]
    return float(d.max() / r.max()), float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(r ** 2)))


def covered_mask(meta, shape):
    m = np.zeros(shape[:2], dtype=bool)
    for s in meta:
        m[s["baselineIdx"], s["timeSampleStart"]:s["timeSampleStart"] + s["timeSampleCount"]] = True
    return m


def model_image(ob):
    """The sky's point sources at their nearest pixels (a model to degrid)."""
    G = ob.c.GridSize
    cell = float(ob.c.ScaleArcSec)
    img = np.zeros((G, G), dtype=np.float32)
    for flux, l, m, sigma in ob.sources:
        x, y = G // 2 + int(round(l / cell)), G // 2 + int(round(m / cell))
        if 0 <= x < G and 0 <= y < G:
            img[y, x] += flux
    return img


@pytest.fixture(scope="module")
def large_refs(O):
    """The large workload, a model grid (ToUVW of its point sources) and the
    oracle's dirty image, PSF and degridded visibilities of all of it."""
    ob = obs.large()
    part = ob.partition()
    oc = oracle_constants(ob.c)
    G = ob.c.GridSize
    with p9idg.Plan(ob.c, kernel="fp32") as plan:  # ToUVW only (no gridder involved)
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_image(model_image(ob))
        plan.exec_image_to_grid()
        model_grid = plan.download_model_grid()
    vis128 = ob.vis.astype(np.complex128)
    t = time.perf_counter()
    dirty = O.gridw_to_image(oc, part.subgrids, vis128, ob.uvw, ob.freq)
    psf = O.gridw_to_image(oc, part.subgrids, np.ones_like(vis128), ob.uvw, ob.freq)
    del vis128
    vis = O.degridw(oc, part.subgrids, model_grid.astype(np.complex128), ob.uvw, ob.freq)
    print("oracle references of the large workload: %.1f s on %d threads" % (time.perf_counter() - t, O.get_threads()))
    assert dirty.shape == psf.shape == (G, G)
    return ob, part, model_grid, {"dirty": dirty, "psf": psf, "vis": vis}


@pytest.mark.timeout(900)
@pytest.mark.parametrize("kernel", ["mfma_bf16", "mfma_fp16"])
def test_large_workload_dirty_image_psf_and_visibilities(large_refs, kernel):
    ob, part, model_grid, ref = large_refs
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        tuned = plan.autotune(reps=2)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        dirty = plan.download_image()
        plan.set_psf_vis(None)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        psf = plan.download_image()
        plan.upload_model_grid(model_grid)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        vis = plan.download_model_vis()
    mask = covered_mask(part.subgrids, vis.shape)
    e = {"dirty": errs(dirty, ref["dirty"]), "psf": errs(psf, ref["psf"]), "vis": errs(vis, ref["vis"], mask)}
    print(kernel, "launch", tuned, {k: "max %.2e rms %.2e" % v for k, v in e.items()})
    for name, (mx, rms) in e.items():
        assert mx <= TOL and rms <= TOL, (name, mx, rms)
    # timesteps no subgrid covers stay 0 (Gridder.cs:96)
    assert np.abs(vis[~mask]).max() == 0.0


@pytest.fixture(scope="module")
def batch_obs():
    return [obs.batch_instance(i) for i in range(8)]


def sample(part, n):
    return np.ascontiguousarray(part.subgrids[:: max(1, part.total // n)])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("kernel", ["mfma_bf16", "mfma_fp16"])
def test_batched_instances_on_subgrid_samples(O, batch_obs, kernel):
    """Every instance restricted to a stride sample of its subgrids, the 8
    plans autotuned and run as one batch, eager and replayed: both directions
    against the oracle on exactly those subgrids."""
    plans, refs = [], []
    try:
        for ob in batch_obs:
            meta = sample(ob.partition(), 40)
            oc = oracle_constants(ob.c)
            ref_l = O.gridw(oc, meta, ob.vis.astype(np.complex128), ob.uvw, ob.freq)
            model = ref_l.sum(axis=0)
            ref_v = O.degridw(oc, meta, model, ob.uvw, ob.freq)
            p = p9idg.Plan(ob.c, kernel=kernel)
            p.set_observation(ob.uvw, ob.freq, meta)
            p.upload_vis(ob.vis)
            p.upload_model_grid(model.astype(np.complex64))
            p.autotune(reps=1)
            plans.append(p)
            refs.append((ref_l, ref_v, covered_mask(meta, ob.vis.shape), model.astype(np.complex64)))
        assert sorted({ob.c.SubgridSize for ob in batch_obs}) == [24, 32, 48]
        with p9idg.Batch(plans) as batch:
            batch.exec(3)
            batch.sync()
            batch.capture(3)
            for run in ("eager", "replay 1", "replay 2"):
                if run != "eager":
                    # zero both outputs with eager work first: the replay must rewrite them
                    for p, ob, (_, _, _, model) in zip(plans, batch_obs, refs):
                        p.upload_vis(np.zeros_like(ob.vis))
                        p.exec_grid(p9idg.GRID_WSTACK)
                        p.upload_model_grid(np.zeros_like(model))
                        p.exec_degrid(p9idg.DEGRID_WSTACK)
                        p.upload_vis(ob.vis)
                        p.upload_model_grid(model)
                    batch.launch()
                    batch.sync()
                for i, (p, (ref_l, ref_v, mask, _)) in enumerate(zip(plans, refs)):
                    eg = errs(p.download_grid(p9idg.GRID_WSTACK), ref_l)[0]
                    ev = errs(p.download_model_vis(), ref_v, mask)[0]
                    print(kernel, run, batch_obs[i].name, "S %d G %d C %d: gridw %.2e degridw %.2e"
                          % (p.c.SubgridSize, p.c.GridSize, batch_obs[i].vis.shape[2], eg, ev))
                    assert eg <= TOL and ev <= TOL, (run, i, eg, ev)
    finally:
        for p in plans:
            p.close()


@pytest.mark.timeout(600)
def test_batched_full_instances(O, batch_obs):
    """The 8 whole instances (55 M gridded visibilities) as one autotuned
    batch: degridded visibilities of a subgrid sample against the oracle
    (a visibility depends only on its own subgrid), and the w-layers equal a
    separate plan's GridW of the same instance up to the float atomics'
    summation order (a few 1e-7 to 1e-6 of the layers' maximum where many
    subgrids overlap), for the eager run and two graph replays."""
    plans, single, refs = [], [], []
    n_vis = 0
    try:
        for ob in batch_obs:
            part = ob.partition()
            p = p9idg.Plan(ob.c, kernel="mfma_bf16")
            p.set_observation(ob.uvw, ob.freq, part)
            p.upload_vis(ob.vis)
            p.upload_image(model_image(ob))
            p.exec_image_to_grid()
            p.autotune(reps=1)
            n_vis += p.gridded_visibilities
            model = p.download_model_grid()
            plans.append(p)
            p.exec_grid(p9idg.GRID_WSTACK)
            single.append(p.download_grid(p9idg.GRID_WSTACK))
            meta = sample(part, 30)
            oc = oracle_constants(ob.c)
            refs.append((O.degridw(oc, meta, model.astype(np.complex128), ob.uvw, ob.freq),
                         covered_mask(meta, ob.vis.shape)))
        print("batch: %d gridded visibilities" % n_vis)
        assert n_vis > 50_000_000
        with p9idg.Batch(plans) as batch:
            batch.exec(3)
            batch.sync()
            batch.capture(3)
            for run in ("eager", "replay 1", "replay 2"):
                if run != "eager":
                    batch.launch()
                    batch.sync()
                for i, (p, (ref_v, mask)) in enumerate(zip(plans, refs)):
                    ev = errs(p.download_model_vis(), ref_v, mask)[0]
                    layers = p.download_grid(p9idg.GRID_WSTACK)
                    el = float(np.abs(layers - single[i]).max() / np.abs(single[i]).max())
                    print(run, batch_obs[i].name, "degridw %.2e, layers vs single plan %.2e" % (ev, el))
                    assert ev <= TOL and el <= 5e-6, (run, i, ev, el)
    finally:
        for p in plans:
            p.close()
