"""GPU parity: every gridding entry point of libp9idg (through the C ABI)
against the FP64 oracle on the same seeded inputs, for the per-pixel and the
separable FP32 vector kernels and the split BF16 / FP16 MFMA kernels.

Tolerance: max |gpu - oracle| / max |oracle| <= 1e-5 for grids, images and
degridded visibilities (the accuracy bar of the split-precision form).  The
"large" workload is checked on a stride sample of its subgrids, and at full
size through the adjoint identity <Grid(V), X> = <V, DeGrid(X)>.
"""
import numpy as np
import pytest

import p9idg
from p9idg import observation as obs

from conftest import oracle_constants, rel_err

pytestmark = pytest.mark.gpu

KERNELS = ["fp32", "fp32_sep", "mfma_bf16", "mfma_fp16"]
TOL = 1e-5


@pytest.fixture(scope="module")
def small():
    ob = obs.small()
    return ob, ob.partition(), oracle_constants(ob.c)


@pytest.fixture(scope="module")
def small_ref(small, O):
    ob, part, oc = small
    m = part.subgrids
    ref = {}
    ref["grid"] = O.grid(oc, m, ob.vis, ob.uvw, ob.freq)
    ref["gridw"] = O.gridw(oc, m, ob.vis, ob.uvw, ob.freq)
    ref["degrid"] = O.degrid(oc, m, ref["grid"], ob.uvw, ob.freq)
    ref["degridw"] = O.degridw(oc, m, ref["grid"], ob.uvw, ob.freq)
    ref["to_image"] = O.to_image(oc, m, ob.vis, ob.uvw, ob.freq)
    ref["gridw_image"] = O.gridw_to_image(oc, m, ob.vis, ob.uvw, ob.freq)
    ref["image_to_vis"] = O.image_to_vis(oc, m, ref["gridw_image"], ob.uvw, ob.freq)
    ones = np.ones_like(ob.vis)
    ref["psf_grid"] = O.grid(oc, m, ones, ob.uvw, ob.freq)
    ref["psf"] = O.to_image(oc, m, ones, ob.uvw, ob.freq)
    return ref


@pytest.mark.parametrize("kernel", KERNELS)
def test_small_workload_every_entry_point(small, small_ref, kernel):
    ob, part, oc = small
    ref = small_ref
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        errs = {
            "grid": rel_err(plan.grid(ob.vis), ref["grid"]),
            "gridw": rel_err(plan.gridw(ob.vis), ref["gridw"]),
            "degrid": rel_err(plan.degrid(ref["grid"]), ref["degrid"]),
            "degridw": rel_err(plan.degridw(ref["grid"]), ref["degridw"]),
            "to_image": rel_err(plan.to_image(ob.vis), ref["to_image"]),
            "gridw_image": rel_err(plan.gridw_to_image(ob.vis), ref["gridw_image"]),
            "image_to_vis": rel_err(plan.image_to_vis(ref["gridw_image"]), ref["image_to_vis"]),
            "psf_grid": rel_err(plan.grid_psf(None), ref["psf_grid"]),
            "psf": rel_err(plan.calculate_psf(None), ref["psf"]),
        }
    print(kernel, {k: "%.2e" % v for k, v in errs.items()})
    assert max(errs.values()) <= TOL, errs


def test_static_drop_ins_match_oracle(small, small_ref):
    ob, part, oc = small
    c = ob.c
    assert rel_err(p9idg.IDG.Grid(c, part, ob.vis, ob.uvw, ob.freq), small_ref["grid"]) <= TOL
    layers = p9idg.IDG.GridW(c, part, ob.vis, ob.uvw, ob.freq)
    assert len(layers) == c.WLayerCount
    assert rel_err(np.stack(layers), small_ref["gridw"]) <= TOL
    grid = small_ref["grid"].copy()
    vis = p9idg.IDG.DeGridW(c, part, grid, ob.uvw, ob.freq)
    assert np.array_equal(grid, small_ref["grid"])  # the caller's grid is not modified
    assert rel_err(vis, small_ref["degridw"]) <= TOL
    assert rel_err(p9idg.IDG.ToImage(c, part, ob.vis, ob.uvw, ob.freq), small_ref["to_image"]) <= TOL
    assert rel_err(p9idg.IDG.CalculatePSF(c, part, ob.uvw, np.zeros(ob.vis.shape, bool), ob.freq), small_ref["psf"]) <= TOL


def test_psf_flags(small, O):
    ob, part, oc = small
    rng = np.random.default_rng(9)
    flags = rng.random(ob.vis.shape) < 0.2
    psf_vis = np.where(flags, 0, 1).astype(np.complex128)
    ref = O.grid(oc, part.subgrids, psf_vis, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        assert rel_err(plan.grid_psf(flags), ref) <= TOL


CASES = [
    # S, K, L, wstep-from-data, antennas, timesteps, channels
    (8, 4, 1, False, 8, 30, 5),
    (16, 4, 4, True, 10, 40, 6),
    (24, 8, 4, True, 10, 40, 6),
    (32, 8, 4, True, 12, 50, 8),
    (48, 8, 2, True, 8, 30, 4),
    (48, 8, 4, True, 10, 40, 8),  # 3 x 3 tiles of 16 with channel stepping
    (64, 8, 2, True, 8, 30, 4),
]


@pytest.mark.parametrize("S,K,L,wst,n_ant,T,C", CASES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_subgrid_sizes_and_wstack(O, S, K, L, wst, n_ant, T, C, kernel):
    ob = obs.custom(G=256, S=S, K=K, L=L, n_ant=n_ant, T=T, C=C, seed=40 + S, core_diam=1500.0,
                    sources=[(1.0, 12 * obs.ARCSEC * 1.5, -9 * obs.ARCSEC * 1.5, 0.0), (0.3, -40 * obs.ARCSEC * 1.5, 5 * obs.ARCSEC * 1.5, 0.0)],
                    noise=0.01, wstep=None if wst else 0.0)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    m = part.subgrids
    if L > 1:
        assert (m["WLambda"] < 0).any() or (m["WLambda"] > 0).any()
    ref_l = O.gridw(oc, m, ob.vis, ob.uvw, ob.freq)
    model = ref_l.sum(axis=0)
    ref_v = O.degridw(oc, m, model, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        e_g = rel_err(plan.gridw(ob.vis), ref_l)
        e_d = rel_err(plan.degridw(model), ref_v)
    print(kernel, S, "gridw %.2e degridw %.2e" % (e_g, e_d))
    assert e_g <= TOL and e_d <= TOL


@pytest.mark.parametrize("G,S", [(250, 32), (202, 24)])
@pytest.mark.parametrize("kernel", ["mfma_bf16", "fp32"])
def test_grid_size_not_a_multiple_of_32(O, G, S, kernel):
    """An even grid size that is no multiple of 32 (the deterministic sums'
    tile size): GridW / DeGridW against the oracle, atomic and ordered."""
    ob = obs.custom(G=G, S=S, K=8, L=4, n_ant=10, T=40, C=6, seed=G + S, core_diam=1500.0)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    ref_l = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
    ref_v = O.degridw(oc, part.subgrids, ref_l.sum(axis=0), ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        assert rel_err(plan.gridw(ob.vis), ref_l) <= TOL
        assert rel_err(plan.degridw(ref_l.sum(axis=0)), ref_v) <= TOL
        plan.set_deterministic(True)
        det = plan.gridw(ob.vis)
        assert rel_err(det, ref_l) <= TOL
        assert np.array_equal(plan.gridw(ob.vis), det)


@pytest.mark.parametrize("G,S", [(255, 32), (201, 24)])
@pytest.mark.parametrize("kernel", KERNELS)
def test_odd_grid_size(O, G, S, kernel):
    """An odd grid size.  FFT.Shift (FFT.cs:248-264) swaps the G/2 x G/2
    blocks and leaves the last row and column where they are; the oracle
    follows it (test_oracle_fft.py).  Grid / DeGrid / ToImage / CalculatePSF /
    ImageToVis on one layer, GridW / DeGridW on 4 w-layers with negative w,
    atomic and ordered sums, against the oracle."""
    ob = obs.custom(G=G, S=S, K=8, L=1, n_ant=10, T=40, C=6, seed=G + S, core_diam=1200.0, wstep=0.0)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    m = part.subgrids
    ref_grid = O.grid(oc, m, ob.vis, ob.uvw, ob.freq)
    ref_img = O.to_image(oc, m, ob.vis, ob.uvw, ob.freq)
    ones = np.ones_like(ob.vis)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        errs = {
            "grid": rel_err(plan.grid(ob.vis), ref_grid),
            "degrid": rel_err(plan.degrid(ref_grid), O.degrid(oc, m, ref_grid, ob.uvw, ob.freq)),
            "to_image": rel_err(plan.to_image(ob.vis), ref_img),
            "psf": rel_err(plan.calculate_psf(None), O.to_image(oc, m, ones, ob.uvw, ob.freq)),
            "image_to_vis": rel_err(plan.image_to_vis(ref_img),
                                    O.image_to_vis(oc, m, ref_img, ob.uvw, ob.freq)),
        }
    obw = obs.custom(G=G, S=S, K=8, L=4, n_ant=10, T=40, C=6, seed=G + S + 1, core_diam=1500.0)
    partw = obw.partition()
    ocw = oracle_constants(obw.c)
    mw = partw.subgrids
    assert (mw["WLambda"] < 0).any()
    ref_l = O.gridw(ocw, mw, obw.vis, obw.uvw, obw.freq)
    ref_v = O.degridw(ocw, mw, ref_l.sum(axis=0), obw.uvw, obw.freq)
    with p9idg.Plan(obw.c, kernel=kernel) as plan:
        plan.set_observation(obw.uvw, obw.freq, partw)
        errs["gridw"] = rel_err(plan.gridw(obw.vis), ref_l)
        errs["degridw"] = rel_err(plan.degridw(ref_l.sum(axis=0)), ref_v)
        plan.set_deterministic(True)
        det = plan.gridw(obw.vis)
        errs["gridw_ordered"] = rel_err(det, ref_l)
        assert np.array_equal(plan.gridw(obw.vis), det)
    print(kernel, G, S, {k: "%.2e" % v for k, v in errs.items()})
    assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("n_ant,C", [(10, 1), (2, 8), (2, 1)])
@pytest.mark.parametrize("kernel", KERNELS)
def test_single_channel_and_single_baseline(O, n_ant, C, kernel):
    """One channel (no channel stepping: C < 2), one baseline (two antennas),
    and both at once."""
    ob = obs.custom(G=256, S=32, K=8, L=2, n_ant=n_ant, T=60, C=C, seed=300 + n_ant + C, core_diam=1500.0)
    part = ob.partition()
    assert part.total > 0
    oc = oracle_constants(ob.c)
    ref_l = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
    ref_v = O.degridw(oc, part.subgrids, ref_l.sum(axis=0), ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        assert rel_err(plan.gridw(ob.vis), ref_l) <= TOL
        assert rel_err(plan.degridw(ref_l.sum(axis=0)), ref_v) <= TOL


REC_CASES = [
    # subgrid, channels, frequency jitter (Hz), work-item size (0: whole
    # subgrids): with even spacing and C % 8 == 0 the gridder steps its
    # phasors channel to channel, re-anchored every 8 visibilities and at
    # every item start (items of 96 start mid-timestep for C = 24, 11, 14); the
    # degridder steps them for any C with even spacing; uneven spacing (and,
    # in the gridder, C % 8 != 0) takes the direct evaluation
    (16, 24, 0.0, 96),
    (32, 16, 0.0, 160),
    (32, 24, 0.0, 96),
    (32, 8, 3.0e5, 64),
    (32, 9, 0.0, 160),
    (32, 11, 0.0, 96),
    (32, 13, 0.0, 64),
    (32, 5, 0.0, 64),
    (48, 16, 0.0, 160),
    (48, 24, 0.0, 96),
    (48, 8, 3.0e5, 64),
    (48, 14, 0.0, 96),
    (48, 9, 0.0, 160),
    (16, 9, 0.0, 0),
    (32, 11, 0.0, 0),
    (32, 13, 0.0, 0),
    (32, 17, 0.0, 0),
    (32, 100, 0.0, 0),
    (32, 10, 3.0e5, 0),
    (48, 9, 0.0, 0),
    (48, 15, 0.0, 0),
    (24, 11, 0.0, 0),
]


@pytest.mark.parametrize("S,C,jitter,chunk", REC_CASES)
@pytest.mark.parametrize("kernel", ["mfma_bf16", "mfma_fp16"])
def test_channel_stepping_cases(O, S, C, jitter, chunk, kernel):
    ob = obs.custom(G=256, S=S, K=8, L=4, n_ant=10, T=40, C=C, f0=1.0e9, f1=1.05e9, seed=70 + C + S,
                    core_diam=1200.0, noise=0.01)
    if jitter:
        ob.freq = ob.freq + jitter * np.sin(np.arange(C) * 1.7)
        ob._partition = None
    part = ob.partition()
    oc = oracle_constants(ob.c)
    m = part.subgrids
    ref_l = O.gridw(oc, m, ob.vis, ob.uvw, ob.freq)
    model = ref_l.sum(axis=0)
    ref_v = O.degridw(oc, m, model, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.set_launch(True, chunk, chunk, 0)
        e_g = rel_err(plan.gridw(ob.vis), ref_l)
        e_d = rel_err(plan.degridw(model), ref_v)
    print(kernel, S, C, jitter, chunk, "gridw %.2e degridw %.2e" % (e_g, e_d))
    assert e_g <= TOL and e_d <= TOL


def test_invalid_grid_error_and_wstack_skip(O):
    ob = obs.custom(G=128, S=16, K=4, L=1, n_ant=6, T=10, C=3, seed=5)
    part = ob.partition()
    meta = part.subgrids.copy()
    meta[0]["UPixel"] = ob.c.GridSize - ob.c.SubgridSize  # Adder.cs:30-31 rejects, AdderWStack skips
    oc = oracle_constants(ob.c)
    with p9idg.Plan(ob.c) as plan:
        plan.set_observation(ob.uvw, ob.freq, meta)
        with pytest.raises(p9idg.P9IDGError, match="invalid grid") as e:
            plan.grid(ob.vis)
        assert e.value.status == p9idg._lib.ERR_INVALID_GRID
        assert rel_err(plan.gridw(ob.vis), O.gridw(oc, meta, ob.vis, ob.uvw, ob.freq)) <= TOL


def test_reobserve_after_failed_set_observation(O):
    """A set_observation that fails part-way (here the visibility buffers do
    not fit in device memory) leaves no undersized buffer behind: the next,
    larger observation grids correctly."""
    ob1 = obs.custom(G=128, S=16, K=4, L=1, n_ant=5, T=8, C=2, seed=11)
    ob2 = obs.custom(G=128, S=16, K=4, L=1, n_ant=8, T=16, C=5, seed=12)
    part2 = ob2.partition()
    with p9idg.Plan(ob2.c) as plan:
        plan.set_observation(ob1.uvw, ob1.freq, ob1.partition())
        B, T, C = 1000, 1000, 50000  # 5e10 visibilities, 400 GB per FP32 buffer
        with pytest.raises(p9idg.P9IDGError) as e:
            plan.set_observation(np.zeros((B, T, 3)), np.linspace(1.0e9, 1.1e9, C), part2.subgrids[:0])
        assert e.value.status == p9idg._lib.ERR_ALLOC
        plan.set_observation(ob2.uvw, ob2.freq, part2)
        got = plan.gridw(ob2.vis)
    ref = O.gridw(oracle_constants(ob2.c), part2.subgrids, ob2.vis, ob2.uvw, ob2.freq)
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("S,kernel", [(32, "mfma_bf16"), (24, "mfma_bf16"), (48, "mfma_fp16"), (16, "fp32")])
def test_deterministic_adder(O, S, kernel):
    """Ordered subgrid sums: GridW and the w-sum grid are identical bit for
    bit across repeats and persistent-workgroup counts (same work items),
    with and without the queue, equal the oracle, and equal the atomic adder
    to float rounding; negative-w mirrors included."""
    ob = obs.custom(G=256, S=S, K=8, L=4, n_ant=10, T=40, C=6, seed=90 + S, core_diam=1500.0)
    part = ob.partition()
    assert (part.subgrids["WLambda"] < 0).any() and (part.subgrids["WLambda"] >= 0).any()
    oc = oracle_constants(ob.c)
    ref_w = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.set_launch(queue=True, grid_chunk=96, degrid_chunk=96, workgroups=0)
        atomic = plan.gridw(ob.vis)
        plan.set_deterministic(True)
        runs = []
        for wg in (0, 5, 0):
            plan.set_launch(queue=True, grid_chunk=96, degrid_chunk=96, workgroups=wg)
            runs.append(plan.gridw(ob.vis))
        for r in runs[1:]:
            assert np.array_equal(r, runs[0])
        assert rel_err(runs[0], ref_w) <= TOL
        assert rel_err(runs[0], atomic) <= 1e-6
        # one workgroup per subgrid (slot = subgrid), then whole-subgrid items
        plan.set_launch(queue=False)
        one = plan.gridw(ob.vis)
        assert rel_err(one, ref_w) <= TOL
        plan.set_launch(queue=True, grid_chunk=0, degrid_chunk=0, workgroups=3)
        assert np.array_equal(plan.gridw(ob.vis), one)
        # the w-sum grid: all layers onto one grid
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSUM)
        wsum = plan.download_grid(p9idg.GRID_WSUM)
        plan.exec_grid(p9idg.GRID_WSUM)
        assert np.array_equal(plan.download_grid(p9idg.GRID_WSUM), wsum)
        assert rel_err(wsum, ref_w.sum(axis=0)) <= TOL


def test_deterministic_adder_plain_grid_and_graph(small, small_ref, O):
    """Deterministic Grid (Adder.cs) and a replayed major-cycle graph: every
    replay equals the eager step bit for bit."""
    ob, part, oc = small
    G = ob.c.GridSize
    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_deterministic(True)
        plan.set_observation(ob.uvw, ob.freq, part)
        g1 = plan.grid(ob.vis)
        assert np.array_equal(plan.grid(ob.vis), g1)
        assert rel_err(g1, small_ref["grid"]) <= TOL
        model = np.zeros((G, G), np.float32)
        model[G // 2 - 12, G // 2 + 20] = 0.7
        plan.upload_vis(ob.vis)
        plan.upload_image(model)
        plan.exec_cycle(p9idg.GRID_WSTACK)
        eager = plan.download_image()
        plan.upload_image(model)
        plan.capture_cycle(p9idg.GRID_WSTACK)
        for rep in range(3):
            plan.upload_image(model)
            plan.exec_grid(p9idg.GRID_PLAIN)  # eager work between replays
            plan.sync()
            plan.upload_image(model)
            plan.launch_cycle()
            plan.sync()
            assert np.array_equal(plan.download_image(), eager), rep
        plan.set_launch(queue=True, grid_chunk=128, degrid_chunk=128)  # new items drop the graph
        with pytest.raises(p9idg.P9IDGError) as e:
            plan.launch_cycle()
        assert e.value.status == p9idg._lib.ERR_STATE


def test_deterministic_graph_survives_other_grid_modes():
    """A captured deterministic cycle keeps its own tile lists: gridding in
    the other modes between replays builds their lists beside it, and every
    replay still equals the eager step bit for bit (4 w-layers, negative w)."""
    ob = obs.custom(G=256, S=32, K=8, L=4, n_ant=10, T=40, C=6, seed=123, core_diam=1500.0)
    part = ob.partition()
    m = part.subgrids
    assert (m["WLambda"] < 0).any() and (m["WLambda"] > 0).any()
    G, S = ob.c.GridSize, ob.c.SubgridSize
    plain_ok = bool(((m["UPixel"] >= 0) & (m["UPixel"] < G - S) & (m["VPixel"] >= 0) & (m["VPixel"] < G - S)).all())
    others = [p9idg.GRID_WSUM, p9idg.GRID_PLAIN if plain_ok else p9idg.GRID_WSUM, p9idg.GRID_WSUM]
    model = np.zeros((G, G), np.float32)
    model[G // 2 + 7, G // 2 - 30] = 0.4
    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_deterministic(True)
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.upload_image(model)
        plan.exec_cycle(p9idg.GRID_WSTACK)
        eager = plan.download_image()
        plan.upload_image(model)
        plan.capture_cycle(p9idg.GRID_WSTACK)
        assert np.array_equal(plan.download_image(), eager)
        for rep, other in enumerate(others):
            plan.exec_grid(other)  # prepares the other mode's tile lists
            plan.sync()
            plan.upload_image(model)
            plan.launch_cycle()
            plan.sync()
            assert np.array_equal(plan.download_image(), eager), (rep, other)


def test_flags_after_capture_drop_the_graph(small, O):
    """The captured cycle holds the flag buffer it was recorded with: setting
    flags for the first time or clearing them drops the graph (launch is then
    refused); new flag contents in the same buffer are read by the replay."""
    ob, part, oc = small
    G = ob.c.GridSize
    rng = np.random.default_rng(8)
    model = np.zeros((G, G), np.float32)
    model[G // 2 - 12, G // 2 + 20] = 0.7
    fl1 = rng.random(ob.vis.shape) < 0.1
    fl2 = rng.random(ob.vis.shape) < 0.1

    def ref(flags):
        mv = O.image_to_vis(oc, part.subgrids, model, ob.uvw, ob.freq)
        return O.gridw_to_image(oc, part.subgrids, O.substract(ob.vis, mv, flags), ob.uvw, ob.freq)

    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.upload_image(model)
        plan.capture_cycle(p9idg.GRID_WSTACK)
        plan.set_flags(fl1)  # first flags: the graph recorded none
        with pytest.raises(p9idg.P9IDGError) as e:
            plan.launch_cycle()
        assert e.value.status == p9idg._lib.ERR_STATE
        plan.upload_image(model)
        plan.capture_cycle(p9idg.GRID_WSTACK)
        plan.set_flags(fl2)  # same buffer, new contents: the graph stays
        plan.upload_image(model)
        plan.launch_cycle()
        plan.sync()
        assert rel_err(plan.download_image(), ref(fl2)) <= TOL
        plan.set_flags(None)  # cleared: dropped again
        with pytest.raises(p9idg.P9IDGError) as e:
            plan.launch_cycle()
        assert e.value.status == p9idg._lib.ERR_STATE


def test_bad_metadata_rejected():
    ob = obs.custom(G=128, S=16, K=4, L=2, n_ant=6, T=10, C=3, seed=6)
    part = ob.partition()
    with p9idg.Plan(ob.c) as plan:
        for field, value in [("baselineIdx", 999), ("timeSampleStart", 9), ("WLambda", 5)]:
            meta = part.subgrids.copy()
            meta[0][field] = value
            with pytest.raises(p9idg.P9IDGError) as e:
                plan.set_observation(ob.uvw, ob.freq, meta)
            assert e.value.status == p9idg._lib.ERR_ARGUMENT


def test_empty_partition_and_single_timestep_subgrids(O):
    ob = obs.custom(G=128, S=16, K=4, L=1, n_ant=6, T=10, C=3, seed=7)
    oc = oracle_constants(ob.c)
    empty = np.zeros(0, dtype=p9idg.SUBGRID_DTYPE)
    single = ob.partition().subgrids.copy()
    single["timeSampleCount"] = 1
    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_observation(ob.uvw, ob.freq, empty)
        assert np.abs(plan.gridw(ob.vis)).max() == 0
        assert np.abs(plan.degridw(np.ones((128, 128)))).max() == 0
        plan.set_observation(ob.uvw, ob.freq, single)
        assert rel_err(plan.gridw(ob.vis), O.gridw(oc, single, ob.vis, ob.uvw, ob.freq)) <= TOL


def test_fft_steps(O):
    rng = np.random.default_rng(3)
    G, L, N = 256, 3, 4321
    layers = rng.standard_normal((L, G, G)) + 1j * rng.standard_normal((L, G, G))
    got = p9idg.FFT.WStackIFFTFloat([layers[k] for k in range(L)], N)
    ref = O.wstack_ifft_float(layers, N)
    assert rel_err(got, ref) <= TOL
    img = rng.standard_normal((G, G)).astype(np.float32)
    assert rel_err(p9idg.FFT.Forward(img, 3.0), O.fft_forward_float(img, 3.0)) <= TOL
    grid = layers[0]
    assert rel_err(p9idg.FFT.Backward(grid, N), O.fft_backward(grid, N)) <= TOL


def test_timing_spans_and_last_timings(small):
    """HIP-event timing: spans of every gridder / degridder launch while
    collecting (adjacent spans share one event per stream position), the
    last-timing slots in both modes, and a reset that clears them."""
    ob, part, _ = small
    G = ob.c.GridSize
    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.upload_model_grid(np.ones((G, G), dtype=np.complex64))
        plan.timing_reset(True)
        for _ in range(3):
            plan.exec_grid(p9idg.GRID_WSTACK)
            plan.exec_degrid(p9idg.DEGRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        plan.sync()
        g_ms, g_n, d_ms, d_n = plan.timing_totals()
        assert (g_n, d_n) == (3, 3) and g_ms > 0 and d_ms > 0
        z_ms, z_n = plan.timing_stage(p9idg.STAGE_ZERO)
        assert z_n >= 3 and z_ms >= 0
        ti_ms, ti_n = plan.timing_stage(p9idg.STAGE_TO_IMAGE)
        assert ti_n == 1 and ti_ms > 0
        t = plan.timings()
        assert t["grid_ms"] > 0 and t["degrid_ms"] > 0 and t["fft_ms"] > 0
        # the last spans and the last-timing slots are the same events
        assert t["grid_ms"] <= g_ms and t["degrid_ms"] <= d_ms
        assert abs(t["fft_ms"] - ti_ms) <= 1e-3 * max(ti_ms, 1e-3)
        plan.timing_reset(False)
        t = plan.timings()
        assert (t["grid_ms"], t["degrid_ms"], t["fft_ms"]) == (-1.0, -1.0, -1.0)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        plan.sync()
        t = plan.timings()
        assert t["grid_ms"] > 0 and t["degrid_ms"] > 0 and t["fft_ms"] == -1.0
        assert plan.timing_totals()[1] == 0  # not collecting


def test_c64_and_c128_inputs_agree(small):
    ob, part, oc = small
    with p9idg.Plan(ob.c) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        a = plan.gridw(ob.vis.astype(np.complex64))
        b = plan.gridw(ob.vis)
    assert rel_err(a, b) <= 1e-6


@pytest.fixture(scope="module")
def large():
    return obs.large()


@pytest.mark.parametrize("kernel", KERNELS)
def test_large_workload_sample(O, large, kernel):
    ob = large
    part = ob.partition()
    meta = np.ascontiguousarray(part.subgrids[::97])
    oc = oracle_constants(ob.c)
    vis128 = ob.vis.astype(np.complex128)
    ref_l = O.gridw(oc, meta, vis128, ob.uvw, ob.freq)
    model = ref_l.sum(axis=0)
    ref_v = O.degridw(oc, meta, model, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, meta)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSTACK)
        layers = plan.download_grid(p9idg.GRID_WSTACK)
        plan.upload_model_grid(model.astype(np.complex64))
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        vis = plan.download_model_vis()
    e_g = rel_err(layers, ref_l)
    covered = np.zeros(vis.shape[:2], bool)
    for s in meta:
        covered[s["baselineIdx"], s["timeSampleStart"]:s["timeSampleStart"] + s["timeSampleCount"]] = True
    e_d = float(np.abs(vis - ref_v)[covered].max() / np.abs(ref_v[covered]).max())
    print(kernel, "large sample (%d subgrids): gridw %.2e degridw %.2e" % (len(meta), e_g, e_d))
    assert e_g <= TOL and e_d <= TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_large_workload_adjoint(large, kernel):
    """<Grid(V), X> = <V, DeGrid(X)> at full size (S = 32: 1/S^2 exact)."""
    ob = large
    part = ob.partition()
    rng = np.random.default_rng(1)
    G = ob.c.GridSize
    X = (rng.standard_normal((G, G)) + 1j * rng.standard_normal((G, G))).astype(np.complex64)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_PLAIN)
        grid = plan.download_grid(p9idg.GRID_PLAIN)
        plan.upload_model_grid(X)
        plan.exec_degrid(p9idg.DEGRID_PLAIN)
        dv = plan.download_model_vis()
    lhs = np.vdot(grid.astype(np.complex128), X.astype(np.complex128))
    rhs = np.vdot(ob.vis.astype(np.complex128), dv.astype(np.complex128))
    print(kernel, "adjoint rel diff %.2e" % (abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-4


def test_wsum_mode_matches_wstack_image(large):
    """GRID_WSUM (all layers in one grid) gives the image of GRID_WSTACK."""
    ob = large
    part = ob.partition()
    with p9idg.Plan(ob.c, kernel="mfma_bf16") as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        a = plan.download_image()
        plan.exec_grid(p9idg.GRID_WSUM)
        plan.exec_grid_to_image(p9idg.GRID_WSUM)
        b = plan.download_image()
    assert rel_err(b, a) <= 1e-5


@pytest.mark.parametrize("kernel", ["fp32", "mfma_bf16"])
def test_major_cycle_step_eager_graph_and_oracle(small, O, kernel):
    """One device step of the major cycle (ToUVW -> DeGridW -> residuals ->
    GridW -> image), run eagerly and as a replayed hipGraph, against the
    oracle's MajorCycle.cs:125-129 + :88-90 sequence."""
    ob, part, oc = small
    G = ob.c.GridSize
    rng = np.random.default_rng(4)
    model = np.zeros((G, G), np.float32)
    model[G // 2 - 12, G // 2 + 20] = 0.7  # near the brighter source
    model[rng.integers(0, G, 5), rng.integers(0, G, 5)] = 0.05
    flags = rng.random(ob.vis.shape) < 0.05
    model_vis = O.image_to_vis(oc, part.subgrids, model, ob.uvw, ob.freq)
    resid = O.substract(ob.vis, model_vis, flags)
    ref = O.gridw_to_image(oc, part.subgrids, resid, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.set_flags(flags)
        plan.upload_image(model)
        plan.exec_cycle(p9idg.GRID_WSTACK)
        eager = plan.download_image()
        assert rel_err(eager, ref) <= TOL
        assert rel_err(plan.download_residual(), resid) <= TOL
        plan.upload_image(model)
        plan.capture_cycle(p9idg.GRID_WSTACK)
        for rep in range(3):
            plan.upload_image(model)  # eager work between replays
            plan.exec_grid(p9idg.GRID_PLAIN)
            plan.sync()
            plan.upload_image(model)
            plan.launch_cycle()
            plan.sync()
            img = plan.download_image()
            assert rel_err(img, eager) <= 1e-6, rep


@pytest.mark.parametrize("kernels", [("mfma_bf16",), ("fp32",), ("mfma_fp16", "fp32"), ("fp32_sep", "mfma_bf16")])
def test_batched_uneven_instances_match_oracle(O, kernels):
    """Observations with subgrid sizes 24, 32 and 48 in one batch: one
    persistent launch per (size, kernel) group, eager and replayed from a
    hipGraph; the plans' kernels taken in turn from `kernels`."""
    obsv = [
        obs.custom(G=256, S=24, K=8, L=4, n_ant=9, T=30, C=5, seed=61),
        obs.custom(G=512, S=32, K=8, L=4, n_ant=12, T=60, C=8, seed=62),
        obs.custom(G=256, S=48, K=8, L=2, n_ant=7, T=20, C=4, seed=63),
        obs.custom(G=384, S=32, K=8, L=4, n_ant=10, T=45, C=6, seed=64),
    ]
    plans, refs = [], []
    try:
        for ob in obsv:
            part = ob.partition()
            oc = oracle_constants(ob.c)
            ref_l = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
            model = ref_l.sum(axis=0)
            ref_v = O.degridw(oc, part.subgrids, model, ob.uvw, ob.freq)
            p = p9idg.Plan(ob.c, kernel=kernels[len(plans) % len(kernels)])
            p.set_launch(queue=True, grid_chunk=128, degrid_chunk=160)
            p.set_observation(ob.uvw, ob.freq, part)
            p.upload_vis(ob.vis)
            p.upload_model_grid(model.astype(np.complex64))
            plans.append(p)
            refs.append((ref_l, ref_v))
        with p9idg.Batch(plans) as batch:
            batch.exec(3)
            batch.sync()
            for p, (ref_l, ref_v) in zip(plans, refs):
                assert rel_err(p.download_grid(p9idg.GRID_WSTACK), ref_l) <= TOL
                assert rel_err(p.download_model_vis(), ref_v) <= TOL
            batch.capture(3)
            for rep in range(3):
                for p in plans:  # eager work between replays
                    p.exec_grid(p9idg.GRID_PLAIN)
                    p.sync()
                    if rep == 2:  # left queued: the batch must start after it
                        p.exec_grid(p9idg.GRID_WSTACK)
                batch.launch()
                if rep < 2:
                    batch.sync()
                # without batch.sync() the plans' downloads must wait for the batch
                for p, (ref_l, ref_v) in zip(plans, refs):
                    assert rel_err(p.download_grid(p9idg.GRID_WSTACK), ref_l) <= TOL, rep
                    assert rel_err(p.download_model_vis(), ref_v) <= TOL, rep

            # a member plan's new observation invalidates the batch until prepared again
            ob = obs.custom(G=256, S=24, K=8, L=4, n_ant=8, T=24, C=7, seed=65, wstep=obsv[0].c.WStep)
            part = ob.partition()
            oc = oracle_constants(ob.c)
            ref_l = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
            ref_v = O.degridw(oc, part.subgrids, ref_l.sum(axis=0), ob.uvw, ob.freq)
            plans[0].set_observation(ob.uvw, ob.freq, part)
            plans[0].upload_vis(ob.vis)
            plans[0].upload_model_grid(ref_l.sum(axis=0).astype(np.complex64))
            for call in (lambda: batch.exec(3), batch.launch):
                with pytest.raises(p9idg.P9IDGError) as e:
                    call()
                assert e.value.status == p9idg._lib.ERR_STATE
            batch.prepare()
            batch.exec(3)
            assert rel_err(plans[0].download_grid(p9idg.GRID_WSTACK), ref_l) <= TOL
            assert rel_err(plans[0].download_model_vis(), ref_v) <= TOL
            for p, (ref_l, ref_v) in zip(plans[1:], refs[1:]):
                assert rel_err(p.download_grid(p9idg.GRID_WSTACK), ref_l) <= TOL
            # batches keep the atomic adder: a deterministic member is refused,
            # also when it turns deterministic after the batch was prepared
            plans[1].set_deterministic(True)
            for call in (lambda: batch.exec(3), batch.prepare):
                with pytest.raises(p9idg.P9IDGError) as e:
                    call()
                assert e.value.status == p9idg._lib.ERR_UNSUPPORTED
    finally:
        for p in plans:
            p.close()


@pytest.mark.parametrize("kernel", ["mfma_bf16", "fp32_sep"])
def test_autotuned_launch_shape_matches_oracle(O, kernel):
    """p9idg_autotune picks a shape among its candidates and leaves the plan on
    it; GridW and DeGridW at that shape (atomic and ordered adder) equal the
    oracle."""
    ob = obs.custom(G=512, S=32, K=8, L=4, n_ant=20, T=100, C=8, seed=88, core_diam=2500.0, noise=0.01)
    part = ob.partition()
    oc = oracle_constants(ob.c)
    ref_l = O.gridw(oc, part.subgrids, ob.vis, ob.uvw, ob.freq)
    model = ref_l.sum(axis=0)
    ref_v = O.degridw(oc, part.subgrids, model, ob.uvw, ob.freq)
    with p9idg.Plan(ob.c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.upload_model_grid(model.astype(np.complex64))
        tuned = plan.autotune(reps=1)
        print(kernel, "autotuned", tuned)
        assert tuned["grid_ms"] > 0 and tuned["degrid_ms"] > 0
        assert tuned["grid_chunk"] in (0, 512, 1024, 2048, 4096) and tuned["degrid_chunk"] in (0, 512, 1024, 2048, 4096)
        assert plan.get_launch() == {k: tuned[k] for k in ("queue", "grid_chunk", "degrid_chunk", "workgroups")}
        plan.exec_grid(p9idg.GRID_WSTACK)
        assert rel_err(plan.download_grid(p9idg.GRID_WSTACK), ref_l) <= TOL
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        assert rel_err(plan.download_model_vis(), ref_v) <= TOL
        plan.set_deterministic(True)
        plan.exec_grid(p9idg.GRID_WSTACK)
        det = plan.download_grid(p9idg.GRID_WSTACK)
        assert rel_err(det, ref_l) <= TOL
        plan.exec_grid(p9idg.GRID_WSTACK)
        assert np.array_equal(plan.download_grid(p9idg.GRID_WSTACK), det)
