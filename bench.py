#!/usr/bin/env python3
"""Grid + degrid throughput of the image-domain gridder on one MI355X.

Workload: the "large" observation (3072^2 image, 32 w-layers, 64 antennas =
2016 baselines x 2000 timesteps x 8 channels, subgrid 32, kernel 8), inputs
resident in HBM.  One step = IDG.GridW (zero the layers' footprint, fused
gridder -> subgrid FFT -> w-stack adder) + IDG.DeGridW (fused splitter ->
subgrid FFT -> degridder) over every gridded visibility.  Visibilities per
second count the visibilities the partition places on subgrids; the total is
reported beside.

Prints one JSON line.  Beside the headline it carries:
  vs_fp32         the same step with the two FP32 vector-ALU kernels
                  (per-pixel phasors, and the separable phasors of the MFMA
                  form), each autotuned, timed in this run
  hbm             achieved GB/s of the HBM-bound stages (target zeroing,
                  ToImage, ToUVW), HIP events over repeated runs
  roofline        the slower direction's kernel against the MFMA (or FP32)
                  peak; roofline.other: the other direction's kernel
  error_vs_oracle the timed plan (its autotuned launch shape) against the FP64
                  oracle on the whole workload: dirty image, PSF, degridded
                  visibilities (max and RMS relative error)
  cpu_baseline    the FP64 oracle (a C++ restatement of the reference's
                  Parallel.For path) timed on the host: grid + degrid of the
                  whole workload on the host threads, and a 1-thread sample
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "p9-distributed-reconstruction_amd"))

METRIC = "Mvis/s grid+degrid, 3072² 32 w-layers: split-BF16 MFMA vs FP32, ≤1e-5 err"
PEAK_TFLOPS = {"fp32": 157.3, "fp32_sep": 157.3, "mfma_bf16": 2500.0, "mfma_fp16": 2500.0}
BOUND = {"fp32": "valu", "fp32_sep": "valu", "mfma_bf16": "mfma", "mfma_fp16": "mfma"}
DTYPE = {"fp32": "f32", "fp32_sep": "f32", "mfma_bf16": "bf16 hi/lo split, f32 accumulate",
         "mfma_fp16": "f16 hi/lo split, f32 accumulate"}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return max(1, min(n, 16))


def cpu_info():
    """nproc, the CPUs this process may use, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable_cpus": usable, "cpu_model": model}


def model_image(ob):
    """Model image with the sky's point sources at their nearest pixels."""
    G = ob.c.GridSize
    cell = float(ob.c.ScaleArcSec)
    img = np.zeros((G, G), dtype=np.float32)
    for flux, l, m, sigma in ob.sources:
        x = G // 2 + int(round(l / cell))
        y = G // 2 + int(round(m / cell))
        if 0 <= x < G and 0 <= y < G:
            img[y, x] += flux
    return img


def sample_subgrids(meta, C, target_vis):
    """Every k-th subgrid, k chosen so the sample holds about target_vis visibilities."""
    total = int(meta["timeSampleCount"].sum()) * C
    k = max(1, int(round(total / max(target_vis, 1))))
    return np.ascontiguousarray(meta[::k])


def rel_errs(got, ref, mask=None):
    """max |got - ref| / max |ref| and the RMS ratio (over mask)."""
    d = np.abs(got.astype(np.complex128) - ref)
    r = np.abs(ref)
    if mask is not None:
        d, r = d[mask], r[mask]
    return {"max_rel": float(d.max() / r.max()), "rms_rel": float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(r ** 2)))}


def covered_mask(meta, shape):
    m = np.zeros(shape[:2], dtype=bool)
    for s in meta:
        m[s["baselineIdx"], s["timeSampleStart"]:s["timeSampleStart"] + s["timeSampleCount"]] = True
    return m


def oracle_constants(O, c):
    return O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                            float(c.ScaleArcSec), c.WLayerCount, c.WStep)


def oracle_checks(ob, part, gpu, model_grid_c64):
    """The FP64 oracle on the whole workload: the references of the GPU's dirty
    image, PSF and degridded visibilities, and its grid + degrid time on the
    host threads (the CPU baseline), plus a 1-thread rate on a stride sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = host_threads()
    O.set_threads(threads)
    c = ob.c
    oc = oracle_constants(O, c)
    meta = part.subgrids
    C = len(ob.freq)
    vis128 = ob.vis.astype(np.complex128)
    model128 = model_grid_c64.astype(np.complex128)
    n_vis = int(meta["timeSampleCount"].sum()) * C
    t0 = time.perf_counter()
    layers = O.gridw(oc, meta, vis128, ob.uvw, ob.freq)  # IDG.GridW (IDG.cs:32-40)
    t_grid = time.perf_counter() - t0
    dirty = O.wstack_ifft_float(layers, c.VisibilitiesCount)  # FFT.WStackIFFTFloat + Shift (MajorCycle.cs:88-90)
    O.shift(dirty)
    del layers
    log("oracle: dirty image (grid %.1f s)" % t_grid)
    psf = O.gridw_to_image(oc, meta, np.ones_like(vis128), ob.uvw, ob.freq)
    log("oracle: psf")
    t0 = time.perf_counter()
    vis = O.degridw(oc, meta, model128, ob.uvw, ob.freq)  # IDG.DeGridW (IDG.cs:42-50)
    t_degrid = time.perf_counter() - t0
    log("oracle: degridded visibilities (%.1f s)" % t_degrid)
    mask = covered_mask(meta, ob.vis.shape)
    err = {
        "dirty_image": rel_errs(gpu["dirty"], dirty),
        "psf": rel_errs(gpu["psf"], psf),
        "degridded_vis": rel_errs(gpu["vis"], vis, mask),
        "uncovered_vis_all_zero": bool(np.abs(gpu["vis"][~mask]).max() == 0.0),
    }
    del dirty, psf, vis
    # the same grid + degrid on one thread, on a stride sample of the subgrids
    smp = sample_subgrids(meta, C, n_vis // 200)
    n_smp = int(smp["timeSampleCount"].sum()) * C
    O.set_threads(1)
    t0 = time.perf_counter()
    lay1 = O.gridw(oc, smp, vis128, ob.uvw, ob.freq)
    O.degridw(oc, smp, model128, ob.uvw, ob.freq)
    t_one = time.perf_counter() - t0
    del lay1
    O.set_threads(threads)
    info = cpu_info()
    cpu = {
        "value": n_vis / (t_grid + t_degrid) / 1e6,
        "unit": "Mvis/s",
        "cores": threads,
        "kind": "port",
        "sample": "the whole workload (%d gridded visibilities): FP64 GridW %.1f s + DeGridW %.1f s on %d host threads"
                  % (n_vis, t_grid, t_degrid, threads),
        "single_thread": {"value": n_smp / t_one / 1e6, "unit": "Mvis/s",
                          "sample": "every %d-th subgrid: %d visibilities, grid + degrid %.1f s"
                                    % (max(1, len(meta) // max(len(smp), 1)), n_smp, t_one)},
        "nproc": info["nproc"],
        "usable_cpus": info["usable_cpus"],
        "cpu_model": info["cpu_model"],
    }
    return cpu, err


def time_steps(plan, steps, warmup):
    """W untimed and K timed grid + degrid steps; wall time per step and the
    HIP-event averages of the gridder, degridder and zero-fill spans."""
    import p9idg

    def step():
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)

    for _ in range(warmup):
        step()
    plan.sync()
    plan.timing_reset(True)
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    plan.sync()
    t1 = time.perf_counter()
    g_ms, g_n, d_ms, d_n = plan.timing_totals()
    z_ms, _ = plan.timing_stage(p9idg.STAGE_ZERO)
    plan.timing_reset(False)
    return {"ms_per_step": (t1 - t0) * 1e3 / steps, "gridder_ms": g_ms / max(g_n, 1),
            "degridder_ms": d_ms / max(d_n, 1), "zero_ms": z_ms / steps}


def wstack_footprint(meta, G, S, L):
    """Per-layer bounding boxes of the w-stacked adder (AdderWStack.cs:29-70):
    summed area and the union box's area, in pixels."""
    X = meta["UPixel"].astype(np.int64)
    Y = meta["VPixel"].astype(np.int64)
    W = meta["WLambda"].astype(np.int64)
    neg = W < 0
    X = np.where(neg, G - X - S + 1, X)
    Y = np.where(neg, G - Y - S + 1, Y)
    W = np.where(neg, -W - 1, W)
    ok = (X >= 1) & (X < G - S) & (Y >= 1) & (Y < G - S) & (W >= 0) & (W < L)
    area = 0
    for l in range(L):
        k = ok & (W == l)
        if k.any():
            area += int((Y[k].max() + S - Y[k].min()) * (X[k].max() + S - X[k].min()))
    union = int((Y[ok].max() + S - Y[ok].min()) * (X[ok].max() + S - X[ok].min())) if ok.any() else 0
    return area, union


def comparator(ob, part, kernel, steps, warmup, tune):
    """The same step with another kernel on its own (autotuned) plan."""
    import p9idg

    with p9idg.Plan(ob.c, kernel=kernel) as p:
        p.set_observation(ob.uvw, ob.freq, part)
        p.upload_vis(ob.vis)
        p.upload_image(model_image(ob))
        p.exec_image_to_grid()
        launch = p.autotune(reps=2) if tune else p.get_launch()
        t = time_steps(p, steps, warmup)
        n_vis = p.gridded_visibilities
    return {"value": n_vis / (t["ms_per_step"] / 1e3) / 1e6, "unit": "Mvis/s", "ms_per_step": t["ms_per_step"],
            "gridder_ms": t["gridder_ms"], "degridder_ms": t["degridder_ms"], "launch": launch}


def run_cycle(args):
    """The large workload plus 10 major cycles of the serial coordinate-descent
    deconvolver (MajorCycle.ReconstructSerialCD with the reference's LMC
    settings: psf cut 16, lambda 1.0, alpha 0.01, 30000 iterations, 1e-5;
    RunningMethod.cs:41-64), everything resident in HBM."""
    import p9idg
    from p9idg import observation

    cut, cycles, lam, alpha, iters, eps = 16, args.cycles, 1.0, 0.01, 30000, 1e-5
    ob = observation.large(threads=host_threads())
    part = ob.partition()
    c = ob.c
    G = c.GridSize
    plan = p9idg.Plan(c, device=0, kernel=args.kernel)
    plan.set_deterministic(args.deterministic)
    plan.set_observation(ob.uvw, ob.freq, part)
    plan.upload_vis(ob.vis)
    if not args.no_autotune:
        log("autotuned launch shape: %s" % plan.autotune(reps=3))
    # untimed warmup: one cycle of one deconvolution iteration builds every
    # FFT plan of the run (hipFFT compiles its kernels at run time on first
    # use, about 1.5 s on a fresh box) and the graphs
    t = time.perf_counter()
    p9idg.MajorCycle.ReconstructSerialCD(plan, ob.vis, None, cut, 1, lam, alpha, 1, eps)
    log("warmup (one cycle, one iteration): %.2f s" % (time.perf_counter() - t))
    t = time.perf_counter()
    x, stats = p9idg.MajorCycle.ReconstructSerialCD(plan, ob.vis, None, cut, cycles, lam, alpha, iters, eps)
    total = time.perf_counter() - t
    n_vis = plan.gridded_visibilities
    for k, s in enumerate(stats):
        log("cycle %d: lambda %.4g, |dx|max %.3g, %d iterations (converged %d), gradients %.1f ms, "
            "deconvolution %.1f ms, imaging %.1f ms" % (k, s["lambda_"], s["abs_max_diff"], s["iterations"],
                                                       s["converged"], s["gradient_ms"], s["deconv_ms"],
                                                       s["imaging_ms"]))
    total_iters = int(stats["iterations"].sum())
    cd_ms = float(stats["deconv_ms"].sum())
    model_digest = hashlib.sha1(np.ascontiguousarray(x, dtype=np.float32).tobytes()).hexdigest()[:16]
    log("model image: sha1 %s, sum %.9g, max %.9g" % (model_digest, float(np.sum(x, dtype=np.float64)), float(x.max())))
    vs_fp64 = compare_cycle_fixture(x, stats)
    if vs_fp64:
        log("against the FP64 CPU reconstruction: model max rel err %.3g, L2 rel err %.3g"
            % (vs_fp64["model_max_rel"], vs_fp64["model_l2_rel"]))

    cpu = None
    if not args.no_oracle:
        # oracle costs on bounded samples of the same work: FastSerialCD
        # iterations on the full image (same PSF cut) and the FP64 grid +
        # degrid rate on a stride sample of subgrids
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        threads = host_threads()
        O.set_threads(threads)
        plan.set_psf_vis(None)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        psf = plan.download_image()
        psf_cut = O.psf_cut(psf, cut)
        pc = psf_cut.shape[0]
        psf2 = O.psf_squared(psf_cut)
        amap = O.psf_amap(psf_cut, G, G)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        dirty = plan.download_image()
        with p9idg.PaddedConvolver(psf_cut, (G, G)) as conv:
            g0 = conv.Convolve(dirty)
        k = 40
        t0 = time.perf_counter()
        O.cd_deconvolve(amap, psf2, np.zeros((G, G), np.float32), g0, float(stats[0]["lambda_"]), alpha, k, eps)
        t_iter = (time.perf_counter() - t0) / k
        oc = O.make_constants(c.VisibilitiesCount, G, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                              float(c.ScaleArcSec), c.WLayerCount, c.WStep)
        sample = sample_subgrids(part.subgrids, len(ob.freq), 1_000_000)
        nv = int(sample["timeSampleCount"].sum()) * len(ob.freq)
        vis128 = ob.vis.astype(np.complex128)
        t0 = time.perf_counter()
        layers = O.gridw(oc, sample, vis128, ob.uvw, ob.freq)
        O.degridw(oc, sample, layers.sum(axis=0), ob.uvw, ob.freq)
        t_vis = (time.perf_counter() - t0) / nv
        del layers
        est = total_iters * t_iter + len(stats) * n_vis * t_vis
        cpu = {
            "value": est,
            "unit": "s",
            "cores": threads,
            "kind": "port",
            "sample": "FastSerialCD: %d iterations on the 3072^2 image (%.1f ms each, argmax scan on %d threads); "
                      "grid + degrid: %d visibilities (%.2f us each, %d threads); scaled to %d iterations and %d x %d "
                      "visibilities" % (k, t_iter * 1e3, threads, nv, t_vis * 1e6, threads, total_iters, len(stats),
                                        n_vis),
        }
    plan.close()
    out = {
        "metric": "seconds for %d serial-CD major cycles, 3072^2 32 w-layers (MajorCycle.ReconstructSerialCD)" % cycles,
        "value": total,
        "unit": "s",
        "n_gpus": 1,
        "steps": len(stats),
        "warmup": 1,
        "ms_per_step": total * 1e3 / max(len(stats), 1),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.kernel] + "; deconvolution f32, PSF correlation f64",
        "data": "synthetic (seeded; see p9idg/observation.py 'large')",
        "config": {
            "workload": "large + major cycles: psf cut 16, lambda 1.0, alpha 0.01, 30000 iterations, epsilon 1e-5",
            "kernel": args.kernel,
            "deterministic_adder": bool(args.deterministic),
            "model_sha1": model_digest,
            "cycles_run": len(stats),
            "deconvolution_iterations": total_iters,
            "deconvolution_ms": cd_ms,
            "us_per_iteration": cd_ms * 1e3 / max(total_iters, 1),
            "gridded_visibilities": int(n_vis),
        },
        "cycles": [{k: (float(s[k]) if s.dtype[k].kind == "f" else int(s[k])) for k in
                    ("lambda_", "abs_max_diff", "iterations", "converged", "gradient_ms", "deconv_ms", "imaging_ms")}
                   for s in stats],
        "cpu_baseline": cpu,
        "vs_fp64_reconstruction": vs_fp64,
    }
    print(json.dumps(out), flush=True)


CYCLE_FIXTURE = os.path.join(ROOT, "tests", "golden", "cycle_large.npz")


def compare_cycle_fixture(x, stats):
    """The GPU reconstruction against the FP64 CPU one of the same workload and
    settings (tests/golden/cycle_large.npz, written by make_cycle_fixture.py
    with the oracle gridding and degridding throughout): per-cycle figures and
    the final model image's relative error.  None without the fixture."""
    try:
        f = np.load(CYCLE_FIXTURE, allow_pickle=False)
    except OSError:
        return None
    ref_stats = f["stats"]
    G = int(f["grid_size"])
    if x.shape != (G, G):
        return None
    ref = np.zeros(G * G, dtype=np.float32)
    ref[f["model_index"]] = f["model_value"]
    ref = ref.reshape(G, G)
    n = min(len(stats), len(ref_stats))
    per_cycle = []
    for k in range(n):
        a, b = stats[k], ref_stats[k]
        per_cycle.append({
            "iterations": [int(a["iterations"]), int(b["iterations"])],
            "lambda_rel": float(abs(a["lambda_"] - b["lambda_"]) / abs(b["lambda_"])) if b["lambda_"] else None,
            "abs_max_diff_rel": float(abs(a["abs_max_diff"] - b["abs_max_diff"]) / abs(b["abs_max_diff"]))
            if b["abs_max_diff"] else None,
            "max_dirty_rel": float(abs(a["max_dirty"] - b["max_dirty"]) / abs(b["max_dirty"])) if b["max_dirty"] else None,
        })
    d = x.astype(np.float64) - ref
    nz_gpu, nz_ref = x != 0, ref != 0
    return {
        "fixture": os.path.relpath(CYCLE_FIXTURE, ROOT),
        "cycles_compared": n,
        "cycles": [len(stats), len(ref_stats)],
        "per_cycle": per_cycle,
        "model_max_rel": float(np.abs(d).max() / np.abs(ref).max()),
        "model_l2_rel": float(np.linalg.norm(d) / np.linalg.norm(ref)),
        "model_sum": [float(x.sum(dtype=np.float64)), float(ref.sum(dtype=np.float64))],
        "nonzero_pixels": [int(nz_gpu.sum()), int(nz_ref.sum()), int((nz_gpu & nz_ref).sum())],
        "reference_sha1": str(f["digest"])[:16],
        "reference_seconds": float(f["seconds"]),
        "reference_threads": int(f["threads"]),
    }


def run_cd(args):
    """FastSerialCD alone on a 3072^2 image with a 192^2 PSF-like kernel (a
    Gaussian main lobe with 2% noise sidelobes) and 400 point sources:
    microseconds per deconvolution iteration."""
    import p9idg

    G, pc = 3072, 192
    rng = np.random.default_rng(0)
    y, x = np.mgrid[0:pc, 0:pc] - pc // 2
    psf = (np.exp(-(x * x + y * y) / (2.0 * 12.0 ** 2)) + 0.02 * rng.standard_normal((pc, pc))).astype(np.float32)
    psf[pc // 2, pc // 2] = 1.0
    g = np.zeros((G, G), np.float32)
    for _ in range(400):
        g[rng.integers(100, G - 100), rng.integers(100, G - 100)] = rng.uniform(1, 50)
    g += rng.standard_normal((G, G)).astype(np.float32) * 0.1
    res = {}
    with p9idg.FastSerialCD((G, G), psf) as cd:
        for it in (args.warmup * 100, args.steps * 300):
            xx = np.zeros((G, G), np.float32)
            gg = g.copy()
            t = time.perf_counter()
            r = cd.Deconvolve(xx, gg, 20.0, 0.01, it, 1e-6)
            res = {"iterations": r.IterationCount, "seconds": time.perf_counter() - t}
    us = res["seconds"] / max(res["iterations"], 1) * 1e6
    print(json.dumps({"metric": "us per FastSerialCD iteration, 3072^2 image, 192^2 PSF", "value": us,
                      "unit": "us", "n_gpus": 1, "steps": res["iterations"], "higher_is_better": False,
                      "data": "synthetic", "dtype": "f32"}), flush=True)


def run_batch(args):
    """The 8 uneven seeded observations (p9idg/observation.py batch_instance)
    gridded and degridded as one persistent launch per subgrid-size class,
    replayed from a hipGraph; per-instance error on a subgrid sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import p9idg
    from p9idg import observation

    O.set_threads(host_threads())
    plans, obs_list, parts = [], [], []
    t = time.perf_counter()
    for i in range(8):
        ob = observation.batch_instance(i, threads=host_threads())
        part = ob.partition()
        p = p9idg.Plan(ob.c, kernel=args.kernel)
        p.set_observation(ob.uvw, ob.freq, part)
        p.upload_vis(ob.vis)
        p.upload_image(model_image(ob))
        p.exec_image_to_grid()
        if not args.no_autotune:
            p.autotune(reps=2)
        plans.append(p)
        obs_list.append(ob)
        parts.append(part)
    log("batch: 8 instances ready in %.1f s" % (time.perf_counter() - t))
    n_vis = sum(p.gridded_visibilities for p in plans)
    batch = p9idg.Batch(plans)
    for _ in range(args.warmup):
        batch.exec(3)
    batch.exec(3)
    batch.sync()
    kt = batch.timings()
    batch.capture(3)
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    batch.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # per-instance error on a stride sample of its subgrids: the replayed
    # batch's degridded visibilities of those subgrids (a visibility depends
    # only on its own subgrid), and GridW of the sample on a plan with the
    # instance's tuned launch shape
    errors = []
    for ob, part, p in zip(obs_list, parts, plans):
        c = ob.c
        oc = oracle_constants(O, c)
        meta = np.ascontiguousarray(part.subgrids[:: max(1, part.total // 24)])
        vis = p.download_model_vis()
        model = p.download_model_grid().astype(np.complex128)
        ref_v = O.degridw(oc, meta, model, ob.uvw, ob.freq)
        launch = p.get_launch()
        with p9idg.Plan(c, kernel=args.kernel) as q:
            q.set_launch(launch["queue"], launch["grid_chunk"], launch["degrid_chunk"], launch["workgroups"])
            q.set_observation(ob.uvw, ob.freq, meta)
            q.upload_vis(ob.vis)
            q.exec_grid(p9idg.GRID_WSTACK)
            layers = q.download_grid(p9idg.GRID_WSTACK)
        ref = O.gridw(oc, meta, ob.vis.astype(np.complex128), ob.uvw, ob.freq)
        errors.append({"instance": ob.name, "grid": c.GridSize, "subgrid": c.SubgridSize, "channels": len(ob.freq),
                       "gridded_visibilities": int(part.subgrids["timeSampleCount"].sum()) * len(ob.freq),
                       "sample_subgrids": int(len(meta)),
                       "gridw_layers": rel_errs(layers, ref),
                       "degridw_vis": rel_errs(vis, ref_v, covered_mask(meta, vis.shape))})
    batch.close()
    for p in plans:
        p.close()
    out = {
        "metric": "Mvis/s grid+degrid, 8 batched uneven observations (one persistent launch per subgrid size, hipGraph)",
        "value": n_vis / (ms / 1e3) / 1e6,
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.kernel],
        "data": "synthetic (seeded; see p9idg/observation.py 'batch')",
        "config": {"workload": "batch: 8 seeded observations, 512-2048 px images, subgrids 24/32/48",
                   "kernel": args.kernel, "gridded_visibilities": int(n_vis)},
        "stages": {"gridder_ms": kt["grid_ms"], "degridder_ms": kt["degrid_ms"]},
        "error_vs_oracle": errors,
    }
    print(json.dumps(out), flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "idg_mfma_bf16_large_hbm_traffic.md")


def profiled_traffic(kernel, direction):
    """HBM bytes per launch of the large workload's kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated) and the launch
    shape it was measured at, or (None, None)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                line = line.strip()
                if line.startswith("hbm_bytes_per_launch:"):
                    table = json.loads(line.split(":", 1)[1])
                    v = table.get("%s/%s" % (kernel, direction))
                    if v is None:
                        return None, None
                    if isinstance(v, dict):
                        return int(v["bytes"]), v.get("launch")
                    return int(v), None
    except (OSError, ValueError, KeyError):
        return None, None
    return None, None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="mfma_bf16", choices=sorted(PEAK_TFLOPS))
    ap.add_argument("--launch", default=None, metavar="QUEUE,GRID_CHUNK,DEGRID_CHUNK,WORKGROUPS",
                    help="fixed launch shape instead of the autotuner's (e.g. 1,0,0,256)")
    ap.add_argument("--no-oracle", "--no-cpu-baseline", dest="no_oracle", action="store_true",
                    help="skip the FP64 oracle: no error_vs_oracle and no cpu_baseline (about 1.5 min of host time)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparator kernels")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--workload", default="large", choices=["large", "batch", "cycle", "cd"])
    ap.add_argument("--cycles", type=int, default=10, help="major cycles of --workload cycle")
    ap.add_argument("--deterministic", action="store_true",
                    help="ordered subgrid sums (large and cycle workloads): run-to-run identical grids")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "cycle":
        return run_cycle(args)
    if args.workload == "cd":
        return run_cd(args)

    import p9idg
    from p9idg import observation

    t = time.perf_counter()
    ob = observation.large(threads=host_threads())
    log("workload generated in %.1f s" % (time.perf_counter() - t))
    t = time.perf_counter()
    part = ob.partition()
    t_part = time.perf_counter() - t
    c = ob.c
    S, G, L = c.SubgridSize, c.GridSize, c.WLayerCount
    log("partition: %d subgrids in %.3f s" % (part.total, t_part))

    plan = p9idg.Plan(c, device=0, kernel=args.kernel)
    plan.set_deterministic(args.deterministic)
    plan.set_observation(ob.uvw, ob.freq, part)
    plan.upload_vis(ob.vis)
    model = model_image(ob)
    plan.upload_image(model)
    plan.exec_image_to_grid()  # ToUVW of the model (MajorCycle.cs:125-127)
    plan.sync()
    n_vis = plan.gridded_visibilities
    log("gridded visibilities: %d of %d" % (n_vis, ob.vis.size))
    if args.launch:
        q, gc, dc, wg = (int(v) for v in args.launch.split(","))
        plan.set_launch(q, gc, dc, wg)
    tuning = plan.get_launch()
    if not args.no_autotune and not args.launch:
        tuning = plan.autotune(reps=3)
        log("autotuned launch shape: %s" % tuning)

    tm = time_steps(plan, args.steps, args.warmup)
    ms_per_step = tm["ms_per_step"]
    value = n_vis / (ms_per_step / 1e3) / 1e6
    avg_g, avg_d = tm["gridder_ms"], tm["degridder_ms"]
    log("step %.3f ms: gridder %.3f ms, degridder %.3f ms, zero fills %.3f ms" % (ms_per_step, avg_g, avg_d,
                                                                                 tm["zero_ms"]))

    # HBM-bound stages: ToImage of the 32 layers and ToUVW, HIP events over
    # `steps` runs each (after the timed region; not part of the metric)
    plan.timing_reset(True)
    for _ in range(args.steps):
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        plan.exec_image_to_grid()
    plan.sync()
    ti_ms, ti_n = plan.timing_stage(p9idg.STAGE_TO_IMAGE)
    tu_ms, tu_n = plan.timing_stage(p9idg.STAGE_TO_UVW)
    plan.timing_reset(False)
    area, union = wstack_footprint(part.subgrids, G, S, L)
    GG = G * G
    fft_bytes = 4 * GG * 8  # an in-place G x G complex64 FFT counted as two read + write passes
    stage_bytes = {
        # the footprint boxes of the previous gridding, zeroed before each gridder launch
        "zero": area * 8,
        # zero the layer sum, sum the layer boxes, inverse FFT, real part / N into the shifted image
        "to_image": GG * 8 + area * 8 + union * 8 + fft_bytes + GG * 8 + GG * 4,
        # shifted image into a complex grid, forward FFT
        "to_uvw": GG * 4 + GG * 8 + fft_bytes,
    }
    stage_ms = {"zero": tm["zero_ms"], "to_image": ti_ms / max(ti_n, 1), "to_uvw": tu_ms / max(tu_n, 1)}
    hbm = {k: {"ms": stage_ms[k], "bytes": int(stage_bytes[k]),
               "gb_s": stage_bytes[k] / (stage_ms[k] / 1e3) / 1e9 if stage_ms[k] > 0 else None}
           for k in stage_bytes}
    for k, v in hbm.items():
        v["frac_of_peak"] = v["gb_s"] / HBM_PEAK_GBS if v["gb_s"] else None
    log("hbm stages: " + ", ".join("%s %.3f ms %.0f GB/s" % (k, v["ms"], v["gb_s"] or 0) for k, v in hbm.items()))

    # the timed plan (its launch shape and adder) on the whole workload, for the oracle
    gpu = {}
    if not args.no_oracle:
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        gpu["dirty"] = plan.download_image()
        plan.set_psf_vis(None)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        gpu["psf"] = plan.download_image()
        plan.upload_vis(ob.vis)
        plan.upload_image(model)
        plan.exec_image_to_grid()
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        gpu["vis"] = plan.download_model_vis()
    model_grid = plan.download_model_grid()
    plan.close()

    # the FP32 vector-ALU kernels on the same step, each autotuned
    vs_fp32 = None
    if args.kernel.startswith("mfma") and not args.no_fp32:
        vs_fp32 = {}
        for k in ("fp32_sep", "fp32"):
            vs_fp32[k] = comparator(ob, part, k, args.steps, args.warmup, not args.no_autotune)
            log("%s: %.3f ms per step, %.0f Mvis/s" % (k, vs_fp32[k]["ms_per_step"], vs_fp32[k]["value"]))
        best = max(vs_fp32, key=lambda k: vs_fp32[k]["value"])
        vs_fp32["faster_fp32"] = best
        vs_fp32["ratio"] = value / vs_fp32[best]["value"]

    flops_per_launch = 8.0 * S * S * n_vis  # complex MAC per pixel per visibility
    dom_name, dom_ms = ("gridder", avg_g) if avg_g >= avg_d else ("degridder", avg_d)
    achieved = flops_per_launch / (dom_ms / 1e3) / 1e12
    peak = PEAK_TFLOPS[args.kernel]
    traffic, traffic_launch = profiled_traffic(args.kernel, dom_name)
    roofline = {
        "bound": BOUND[args.kernel],
        "kernel": dom_name,
        "achieved": achieved,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": traffic,
        "traffic_launch": traffic_launch,
        "traffic_at_timed_launch": bool(traffic_launch) and all(
            traffic_launch.get(k) == tuning.get(k) for k in ("queue", "grid_chunk", "degrid_chunk", "workgroups")),
        "flops_per_launch": flops_per_launch,
        "kernel_ms": dom_ms,
    }
    if args.kernel.startswith("mfma") and S in (16, 32, 48):
        # the split product issues 16 16x16x32 MFMAs (three hi/lo terms of four
        # real products for M0, four hi-only ones for M1) per 16 x 16 tile and
        # 32 visibilities, where the complex multiply-add needs the work of 4
        roofline["mfma_issued_per_algorithmic"] = 4.0
        roofline["issued_frac"] = 4.0 * achieved / peak
    # the other direction's kernel against the same peak (same algorithmic work)
    oth_name, oth_ms = ("degridder", avg_d) if dom_name == "gridder" else ("gridder", avg_g)
    oth = flops_per_launch / (oth_ms / 1e3) / 1e12
    roofline["other"] = {"kernel": oth_name, "achieved": oth, "frac": oth / peak, "kernel_ms": oth_ms,
                         "traffic": profiled_traffic(args.kernel, oth_name)[0]}
    cpu, err = None, None
    if not args.no_oracle:
        cpu, err = oracle_checks(ob, part, gpu, model_grid)
        err["launch"] = tuning
        err["deterministic_adder"] = bool(args.deterministic)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.kernel],
        "data": "synthetic (seeded; see p9idg/observation.py 'large')",
        "config": {
            "workload": "large: 3072^2 image, 32 w-layers, 2016 baselines x 2000 timesteps x 8 channels, subgrid 32, kernel 8",
            "kernel": args.kernel,
            "deterministic_adder": bool(args.deterministic),
            "gridded_visibilities": int(n_vis),
            "total_visibilities": int(ob.vis.size),
            "subgrids": int(part.total),
            "launch": tuning,
        },
        "stages": {
            "gridder_ms": avg_g,
            "degridder_ms": avg_d,
            "zero_fill_ms": tm["zero_ms"],
            "to_image_ms": stage_ms["to_image"],
            "to_uvw_ms": stage_ms["to_uvw"],
            "partition_ms": t_part * 1e3,
        },
        "vs_fp32": vs_fp32,
        "hbm": hbm,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "error_vs_oracle": err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
