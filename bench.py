#!/usr/bin/env python3
"""Benchmark of the segmentation hot path on one MI355X.

Default workload (plain method, 3D): a 256 x 1024 x 1024 uint16 nuclei stack (about
5000 randomly oriented ellipsoids, semi-axes 4-10 voxels, intensity 800-3000 on a
background of 200, Poisson shot noise + Gaussian read noise sigma 20; seed 2) resident
in HBM.  One step = gaussian_blur(sigma 2,2,2) -> float32, threshold_otsu,
connected_components_labeling_box into an int32 destination, and
statistics_of_labelled_pixels (the dict, pulled to the host).

Prints one JSON line: throughput in Mvoxels/s, the roofline of the dominant kernel
(measured with HIP events on the launch stream inside the timed region) and a CPU
baseline (the C/OpenMP restatement of the reference algorithms on a bounded slab of
the same volume).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

SHAPE = (256, 1024, 1024)
N_NUCLEI = 5000
SEED = 2
SIGMA = 2.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=256, help="depth of the slab timed on the CPU")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB float32 cle_copy first (known bytes, to scale PMC byte counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is single-device by design",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import pyclesperanto_prototype_amd as cle
    from pyclesperanto_prototype_amd import synthetic

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    t0 = time.time()
    vol_t, info = synthetic.nuclei_3d(SHAPE, N_NUCLEI, seed=SEED, device=dev)
    torch.cuda.synchronize()
    log(f"generated {SHAPE} uint16, {info['nuclei']} nuclei in {time.time() - t0:.1f}s")
    vol = cle.asarray(vol_t)
    nvox = vol.size

    blurred = cle.create(SHAPE, np.float32)
    binary = cle.create(SHAPE, np.uint8)
    labels = cle.create(SHAPE, np.int32)
    if args.copy_calibration:
        # a plain copy of known size: reads 1 GiB and writes 1 GiB
        cle.set(blurred, 1.0)
        cal = cle.create(SHAPE, np.float32)
        cle.copy(blurred, cal)
        torch.cuda.synchronize()
        del cal

    stages = ["gaussian_blur", "threshold_otsu", "connected_components_labeling_box", "statistics_of_labelled_pixels"]

    def step(ev=None):
        if ev: ev[0].record()
        cle.gaussian_blur(vol, blurred, SIGMA, SIGMA, SIGMA)
        if ev: ev[1].record()
        cle.threshold_otsu(blurred, binary)
        if ev: ev[2].record()
        cle.connected_components_labeling_box(binary, labels)
        if ev: ev[3].record()
        stats = cle.statistics_of_labelled_pixels(vol, labels)
        if ev: ev[4].record()
        return stats

    for _ in range(args.warmup):
        stats = step()
    torch.cuda.synchronize()
    n_labels = len(stats["label"])

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats = step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1e3 / args.steps
    stage_ms = {s: float(np.mean([events[k][i].elapsed_time(events[k][i + 1]) for k in range(args.steps)]))
                for i, s in enumerate(stages)}
    log("stage ms:", json.dumps({k: round(v, 3) for k, v in stage_ms.items()}))

    # dominant kernel: the blur (one fused launch) -- algorithmic bytes = read the
    # uint16 source once (2 B/voxel) + write the float32 result once (4 B/voxel)
    blur_bytes = nvox * (2 + 4)
    blur_s = stage_ms["gaussian_blur"] / 1e3
    achieved = blur_bytes / blur_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "gaussian_blur (u16 -> f32, 17 taps/axis)", "algorithmic_bytes_per_launch": blur_bytes}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(vol_t, args.cpu_slices)

    value = nvox / (ms / 1e3) / 1e6
    line = {
        "metric": "Mvoxels/s voronoi-otsu pipeline + % HBM roofline; p50/p99 serve latency",
        "value": round(value, 1),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded nuclei stack, {info['nuclei']} nuclei, {n_labels} labels found)",
        "config": {"workload": "plain method 3D: 256x1024x1024 uint16 nuclei -> gaussian_blur(2,2,2) f32 -> "
                               "threshold_otsu -> connected_components_labeling_box int32 -> "
                               "statistics_of_labelled_pixels",
                   "voxels": nvox, "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(vol_t, slices):
    """The C/OpenMP restatement of the reference algorithms (float32 separable blur,
    row histogram + Otsu, Jacobi CCL, per-row statistics) on the first `slices`
    slices of the volume, on this host's cores."""
    import numpy as np
    import torch
    from oracle import c_oracle

    sub = vol_t[:slices].contiguous().view(torch.int16).cpu().numpy().view(np.uint16).astype(np.float32)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c_oracle.lib()
    t0 = time.perf_counter()
    labels, table = c_oracle.blur_otsu_ccl_stats(sub, SIGMA, True)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {sub.shape} in {dt:.2f}s, {int(labels.max())} labels, {threads} threads")
    return {"value": round(sub.size / dt / 1e6, 2), "unit": "Mvoxels/s", "cores": threads, "kind": "port",
            "sample": f"first {slices} slices ({sub.shape[0]}x{sub.shape[1]}x{sub.shape[2]}) of the same volume, "
                      f"same pipeline, {dt:.1f}s"}


if __name__ == "__main__":
    main()
