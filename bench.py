#!/usr/bin/env python3
"""Grid + degrid throughput of the image-domain gridder on one MI355X.

Workload: the "large" observation (3072^2 image, 32 w-layers, 64 antennas =
2016 baselines x 2000 timesteps x 8 channels, subgrid 32, kernel 8), inputs
resident in HBM.  One step = IDG.GridW (zero the 32 layers, fused gridder ->
subgrid FFT -> w-stack adder) + IDG.DeGridW (fused splitter -> subgrid FFT ->
degridder) over every gridded visibility.  Visibilities per second count the
visibilities the partition places on subgrids; the total is reported beside.

Prints one JSON line (see the keys below).  The CPU baseline is the FP64
oracle (a C++ restatement of the reference's Parallel.For path) timed on the
host cores on a stride sample of the subgrids; the same sample gives the
error of the GPU result against the oracle.
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
PEAK_TFLOPS = {"fp32": 157.3, "mfma_bf16": 2500.0, "mfma_fp16": 2500.0}
DTYPE = {"fp32": "f32", "mfma_bf16": "bf16 hi/lo split, f32 accumulate", "mfma_fp16": "f16 hi/lo split, f32 accumulate"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return max(1, min(n, 16))


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


def cpu_baseline_and_error(args, ob, part, model_grid_c64, kernel):
    """Times the FP64 oracle on a bounded sample and compares the GPU on it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import p9idg

    threads = host_threads()
    O.set_threads(threads)
    c = ob.c
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize,
                          c.MaxTimestepsPerSubgrid, float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    C = len(ob.freq)
    meta = part.subgrids
    vis128 = ob.vis.astype(np.complex128)
    model128 = model_grid_c64.astype(np.complex128)
    g = np.array(model128, copy=True)
    O.shift(g)  # IDG.DeGridW shifts the grid first (IDG.cs:44)

    def run(smp):
        t0 = time.perf_counter()
        sub = O.gridder_forward(oc, smp, ob.uvw, vis128, ob.freq)
        sub = O.subgrid_fft_forward(oc, sub)
        lay = O.adderw_add(oc, smp, sub)
        t1 = time.perf_counter()
        sub = O.adderw_split(oc, smp, g)
        sub = O.subgrid_fft_backward(oc, sub)
        v = O.gridder_backward(oc, smp, sub, ob.uvw, ob.freq)
        t2 = time.perf_counter()
        return lay, v, t1 - t0, t2 - t1

    # calibrate the oracle's rate on about 1/64 of the workload (whole
    # pipeline, both directions), then size the timed sample to the target
    probe = sample_subgrids(meta, C, int(meta["timeSampleCount"].sum()) * C // 64)
    _, _, pa, pb = run(probe)
    del _
    rate = int(probe["timeSampleCount"].sum()) * C / max(pa + pb, 1e-6)
    sample = sample_subgrids(meta, C, max(20000, int(rate * args.cpu_seconds)))
    n_vis = int(sample["timeSampleCount"].sum()) * C
    log("cpu baseline: %d subgrids, %d visibilities, %d threads" % (len(sample), n_vis, threads))
    layers_ref, vis_ref, tg, td = run(sample)
    del g
    for k in range(layers_ref.shape[0]):
        O.shift(layers_ref[k])  # IDG.GridW shifts each layer (IDG.cs:37)
    cpu = {
        "value": n_vis / (tg + td) / 1e6,
        "unit": "Mvis/s",
        "cores": threads,
        "kind": "port",
        "sample": "every %d-th subgrid of the workload: %d subgrids, %d visibilities; FP64 grid %.1f s + degrid %.1f s"
        % (max(1, len(meta) // max(len(sample), 1)), len(sample), n_vis, tg, td),
    }

    # GPU on the same sample
    with p9idg.Plan(c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, sample)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSTACK)
        layers = plan.download_grid(p9idg.GRID_WSTACK)
        plan.upload_model_grid(model_grid_c64)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        vis = plan.download_model_vis()
    err_grid = float(np.abs(layers - layers_ref).max() / np.abs(layers_ref).max())
    covered = np.zeros(vis.shape[:2], dtype=bool)
    for s in sample:
        covered[s["baselineIdx"], s["timeSampleStart"]:s["timeSampleStart"] + s["timeSampleCount"]] = True
    dv = np.abs(vis - vis_ref)[covered]
    err_vis = float(dv.max() / np.abs(vis_ref[covered]).max())
    rms_vis = float(np.sqrt(np.mean(dv ** 2)) / np.sqrt(np.mean(np.abs(vis_ref[covered]) ** 2)))
    err = {
        "sample_subgrids": int(len(sample)),
        "gridw_layers_max_rel": err_grid,
        "degridw_vis_max_rel": err_vis,
        "degridw_vis_rms_rel": rms_vis,
    }
    return cpu, err


def full_accuracy(ob, part, model_grid_c64, kernel):
    """Max and RMS relative error of the dirty image, the PSF and the degridded
    visibilities of the whole workload against the FP64 oracle (minutes of CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import p9idg

    O.set_threads(host_threads())
    c = ob.c
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize,
                          c.MaxTimestepsPerSubgrid, float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    meta = part.subgrids
    vis128 = ob.vis.astype(np.complex128)

    def errs(a, b, mask=None):
        d = np.abs(a.astype(np.float64) - b) if mask is None else np.abs(a - b)[mask]
        r = np.abs(b) if mask is None else np.abs(b)[mask]
        return {"max_rel": float(d.max() / r.max()), "rms_rel": float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(r ** 2)))}

    out = {}
    with p9idg.Plan(c, kernel=kernel) as plan:
        plan.set_observation(ob.uvw, ob.freq, part)
        plan.upload_vis(ob.vis)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        dirty = plan.download_image()
        plan.set_psf_vis(None)
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_grid_to_image(p9idg.GRID_WSTACK)
        psf = plan.download_image()
        plan.upload_model_grid(model_grid_c64)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)
        vis = plan.download_model_vis()
    t = time.perf_counter()
    ref = O.gridw_to_image(oc, meta, vis128, ob.uvw, ob.freq)
    out["dirty_image"] = errs(dirty, ref)
    log("full accuracy: dirty image %.1f s" % (time.perf_counter() - t))
    ref = O.gridw_to_image(oc, meta, np.ones_like(vis128), ob.uvw, ob.freq)
    out["psf"] = errs(psf, ref)
    log("full accuracy: psf done")
    g = model_grid_c64.astype(np.complex128)
    ref_v = O.degridw(oc, meta, g, ob.uvw, ob.freq)
    covered = np.zeros(vis.shape[:2], dtype=bool)
    for s in meta:
        covered[s["baselineIdx"], s["timeSampleStart"]:s["timeSampleStart"] + s["timeSampleCount"]] = True
    out["degridded_vis"] = errs(vis.astype(np.complex128), ref_v, covered)
    log("full accuracy: degrid done")
    return out


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

    cpu = None
    if not args.no_cpu_baseline:
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
            "sample": "FastSerialCD: %d iterations on the 3072^2 image (%.1f ms each, 1 thread); grid + degrid: %d "
                      "visibilities (%.2f us each, %d threads); scaled to %d iterations and %d x %d visibilities"
                      % (k, t_iter * 1e3, nv, t_vis * 1e6, threads, total_iters, len(stats), n_vis),
        }
    plan.close()
    out = {
        "metric": "seconds for %d serial-CD major cycles, 3072^2 32 w-layers (MajorCycle.ReconstructSerialCD)" % cycles,
        "value": total,
        "unit": "s",
        "n_gpus": 1,
        "steps": len(stats),
        "warmup": 0,
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
    }
    print(json.dumps(out), flush=True)


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
    errors = []
    for ob, part, p in zip(obs_list, parts, plans):
        c = ob.c
        oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize,
                              c.MaxTimestepsPerSubgrid, float(c.ScaleArcSec), c.WLayerCount, c.WStep)
        meta = np.ascontiguousarray(part.subgrids[:: max(1, part.total // 24)])
        with p9idg.Plan(c, kernel=args.kernel) as q:
            q.set_observation(ob.uvw, ob.freq, meta)
            q.upload_vis(ob.vis)
            q.exec_grid(p9idg.GRID_WSTACK)
            layers = q.download_grid(p9idg.GRID_WSTACK)
        ref = O.gridw(oc, meta, ob.vis.astype(np.complex128), ob.uvw, ob.freq)
        errors.append({"instance": ob.name, "grid": c.GridSize, "subgrid": c.SubgridSize,
                       "gridded_visibilities": int(part.subgrids["timeSampleCount"].sum()) * len(ob.freq),
                       "gridw_layers_max_rel": float(np.abs(layers - ref).max() / np.abs(ref).max())})
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
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                line = line.strip()
                if line.startswith("hbm_bytes_per_launch:"):
                    table = json.loads(line.split(":", 1)[1])
                    v = table.get("%s/%s" % (kernel, direction))
                    return None if v is None else int(v)
    except (OSError, ValueError):
        return None
    return None


def copy_probe(gib):
    """A plain copy kernel of known size, profiled beside the gridder to turn
    FETCH_SIZE / WRITE_SIZE counter readings into bytes."""
    import torch

    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy probe: %d bytes x 3 launches" % (4 * n))
    del src, dst
    torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel", default="mfma_bf16", choices=sorted(PEAK_TFLOPS))
    ap.add_argument("--launch", default=None, metavar="QUEUE,GRID_CHUNK,DEGRID_CHUNK,WORKGROUPS",
                    help="fixed launch shape instead of the autotuner's (e.g. 1,0,0,256)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="first run a plain device-to-device copy of this many GiB (3 times): the "
                         "reference kernel for scaling HBM byte counters")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--workload", default="large", choices=["large", "batch", "cycle", "cd"])
    ap.add_argument("--cycles", type=int, default=10, help="major cycles of --workload cycle")
    ap.add_argument("--deterministic", action="store_true",
                    help="ordered subgrid sums (large and cycle workloads): run-to-run identical grids")
    ap.add_argument("--full-accuracy", action="store_true",
                    help="also compare the whole workload's dirty image, PSF and degridded visibilities with the FP64 oracle")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
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
    S = c.SubgridSize
    log("partition: %d subgrids in %.3f s" % (part.total, t_part))

    plan = p9idg.Plan(c, device=0, kernel=args.kernel)
    plan.set_deterministic(args.deterministic)
    plan.set_observation(ob.uvw, ob.freq, part)
    plan.upload_vis(ob.vis)
    plan.upload_image(model_image(ob))
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

    def step():
        plan.exec_grid(p9idg.GRID_WSTACK)
        plan.exec_degrid(p9idg.DEGRID_WSTACK)

    for _ in range(args.warmup):
        step()
    plan.sync()
    log("warmup done")
    plan.timing_reset(True)
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    plan.sync()
    t1 = time.perf_counter()
    g_ms, g_n, d_ms, d_n = plan.timing_totals()
    plan.timing_reset(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n_vis / (ms_per_step / 1e3) / 1e6
    avg_g = g_ms / max(g_n, 1)
    avg_d = d_ms / max(d_n, 1)
    log("step %.3f ms: gridder %.3f ms, degridder %.3f ms" % (ms_per_step, avg_g, avg_d))

    # layer FFT + image extraction, timed once (not part of the metric)
    t = time.perf_counter()
    plan.exec_grid_to_image(p9idg.GRID_WSTACK)
    plan.sync()
    fft_wall = (time.perf_counter() - t) * 1e3

    flops_per_launch = 8.0 * S * S * n_vis  # complex MAC per pixel per visibility
    dom_name, dom_ms = ("gridder", avg_g) if avg_g >= avg_d else ("degridder", avg_d)
    achieved = flops_per_launch / (dom_ms / 1e3) / 1e12
    peak = PEAK_TFLOPS[args.kernel]
    roofline = {
        "bound": "mfma",
        "kernel": dom_name,
        "achieved": achieved,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": profiled_traffic(args.kernel, dom_name),
    }
    cpu, err, full = None, None, None
    model_grid = plan.download_model_grid()
    plan.close()
    if not args.no_cpu_baseline:
        cpu, err = cpu_baseline_and_error(args, ob, part, model_grid, args.kernel)
    if args.full_accuracy:
        full = full_accuracy(ob, part, model_grid, args.kernel)

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
            "layer_ifft_and_image_ms": fft_wall,
            "partition_ms": t_part * 1e3,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "error_vs_oracle": err,
    }
    if full is not None:
        out["error_vs_oracle_full"] = full
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
