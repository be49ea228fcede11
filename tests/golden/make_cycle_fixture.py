#!/usr/bin/env python3
"""Writes cycle_large.npz: the FP64 oracle's MajorCycle.ReconstructSerialCD
(MajorCycle.cs:33-134, restated in oracle/oracle_cycle.cpp) on the 3072^2,
32-w-layer "large" workload with the reference's LMC settings (psf cut 16,
lambda 1.0, alpha 0.01, 30000 iterations, epsilon 1e-5; RunningMethod.cs:41-64),
10 major cycles, gridding and degridding in FP64 on the CPU throughout.

Contents: the per-cycle figures (max dirty, max gradient, lambda, |dx|max,
iterations, converged, objective), the final model image as its nonzero
pixels (flat index, float32 value), and its SHA-1 digest.  bench.py
--workload cycle compares the GPU reconstruction against it.

Run from the repository root after building (python __graft_entry__.py build):
    python tests/golden/make_cycle_fixture.py --threads 8
It takes about an hour on 8 cores.

--perturb EPS multiplies every visibility by (1 + EPS z), z complex standard
normal (seed 1), before the reconstruction, and --compare FIXTURE prints the
per-cycle and final-model differences against a fixture: with EPS about the
FP32 rounding error this measures how far the deconvolution alone carries a
difference of that size, e.g.

    python tests/golden/make_cycle_fixture.py --cycles 4 --perturb 1e-7 \
        --out /tmp/cycle_perturbed.npz --compare tests/golden/cycle_large.npz"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "p9-distributed-reconstruction_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from p9idg import observation  # noqa: E402

SETTINGS = {"psf_cut": 16, "cycles": 10, "lambda": 1.0, "alpha": 0.01, "iterations": 30000, "epsilon": 1e-5}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cycle_large.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--cycles", type=int, default=SETTINGS["cycles"])
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--perturb", type=float, default=0.0, help="relative visibility perturbation")
    ap.add_argument("--compare", default=None, help="fixture to compare the result with")
    args = ap.parse_args()
    O.set_threads(args.threads)
    t = time.perf_counter()
    ob = observation.large(threads=args.threads)
    part = ob.partition()
    c = ob.c
    oc = O.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid,
                          float(c.ScaleArcSec), c.WLayerCount, c.WStep)
    print("workload: %d subgrids, %.1f s" % (part.total, time.perf_counter() - t), flush=True)
    vis = ob.vis.astype(np.complex128)
    if args.perturb:
        rng = np.random.default_rng(1)
        z = rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape)
        vis = vis * (1.0 + args.perturb * z)
    t = time.perf_counter()
    x, stats = O.reconstruct_serial_cd(oc, part.subgrids, vis, None, ob.uvw, ob.freq,
                                       SETTINGS["psf_cut"], args.cycles, SETTINGS["lambda"], SETTINGS["alpha"],
                                       SETTINGS["iterations"], SETTINGS["epsilon"])
    secs = time.perf_counter() - t
    for k, s in enumerate(stats):
        print("cycle %d: max dirty %.6g, lambda %.6g, |dx|max %.6g, %d iterations, converged %d"
              % (k, s["max_dirty"], s["lambda_"], s["abs_max_diff"], s["iterations"], s["converged"]), flush=True)
    x = np.ascontiguousarray(x, dtype=np.float32)
    idx = np.flatnonzero(x).astype(np.int32)
    digest = hashlib.sha1(x.tobytes()).hexdigest()
    np.savez_compressed(args.out, stats=stats, model_index=idx, model_value=x.ravel()[idx],
                        grid_size=np.int32(c.GridSize), digest=np.array(digest), seconds=np.float64(secs),
                        threads=np.int32(args.threads), **{"setting_" + k: np.float64(v) for k, v in SETTINGS.items()})
    print("%d cycles in %.0f s; model: %d nonzero pixels, sum %.9g, sha1 %s -> %s"
          % (len(stats), secs, idx.size, float(x.sum(dtype=np.float64)), digest[:16], args.out), flush=True)
    if args.compare:
        compare(x, stats, args.compare)


def compare(x, stats, path):
    """Prints the relative differences of the per-cycle figures and of the
    model image against the fixture at path (cycles present in both)."""
    f = np.load(path, allow_pickle=False)
    ref_stats = f["stats"]
    for k in range(min(len(stats), len(ref_stats))):
        a, b = stats[k], ref_stats[k]
        print("cycle %d vs fixture: lambda %.3g, |dx|max %.3g, max dirty %.3g (relative), iterations %d / %d"
              % (k, abs(a["lambda_"] - b["lambda_"]) / abs(b["lambda_"]),
                 abs(a["abs_max_diff"] - b["abs_max_diff"]) / abs(b["abs_max_diff"]),
                 abs(a["max_dirty"] - b["max_dirty"]) / abs(b["max_dirty"]), a["iterations"], b["iterations"]))
    if len(stats) == len(ref_stats):
        G = int(f["grid_size"])
        ref = np.zeros(G * G, np.float32)
        ref[f["model_index"]] = f["model_value"]
        d = x.astype(np.float64).ravel() - ref
        print("model vs fixture: max rel %.3g, L2 rel %.3g"
              % (np.abs(d).max() / np.abs(ref).max(), np.linalg.norm(d) / np.linalg.norm(ref)))


if __name__ == "__main__":
    main()
