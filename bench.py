#!/usr/bin/env python3
"""Benchmark of the segmentation hot path on one MI355X.

Headline (plain method, large): a 512 x 2048 x 2048 uint16 nuclei stack (about 100 k
ellipsoidal nuclei, semi-axes 3-8 voxels, 15 % placed as touching pairs, intensity
800-3000 on a background of 200, Poisson shot noise + Gaussian read noise sigma 20;
seed 3) resident in HBM.  One step = voronoi_otsu_labeling(spot_sigma=2.0,
outline_sigma=1.5) into int32 labels, then statistics_of_labelled_pixels (the dict
pulled to the host).  value = voxels / mean step time (Mvoxels/s).

The JSON line also holds: the roofline of the dominant kernel (the Voronoi growth's pair
launches, measured with HIP events inside the library on the launch stream over the
timed steps), a CPU baseline (the C/OpenMP restatement of the reference algorithms on a
128 x 1024 x 1024 crop of the same stack, all host threads, plus a 1-thread figure),
and secondary keys: the 512 x 512 latency case, the 2048 x 2048 random image beside the
reference's published figure, the plain 3D method on 256 x 1024 x 1024, the mixed-size
batch, the serving sweep and TIFF ingest.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--sections c3,c1,...]
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
VALU_PEAK_TFLOPS = 78.6  # MI355X float32 vector FMA without packed math

C3_SHAPE = (512, 2048, 2048)
C2_SHAPE = (256, 1024, 1024)
GROWTH_TILE_VOXELS = 1024  # a growth tile: 16 x 8 x 8 (3D), 32 x 32 (2D)
# algorithmic bytes per input voxel (inputs read once, outputs written once, on-device
# dtypes): voronoi_otsu_labeling 25, statistics (two passes over uint16 intensity +
# int32 labels) 12
C3_BYTES_PER_VOXEL = 25 + 12
SECTIONS = ("c3", "c1", "m1", "c2", "batch", "serve", "ingest", "cpu")
PUBLISHED_M1MAX_MS = 530.0  # benchmarks/apple_m1_voronoi_otsu_labeling.ipynb of the reference


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sections", default=",".join(SECTIONS), help="comma list of " + ",".join(SECTIONS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB float32 cle_copy first (known bytes, to scale PMC byte counters)")
    ap.add_argument("--batch-images", type=int, default=2000)
    ap.add_argument("--batch-only", action="store_true",
                    help="batched calls only (no per-image loop, no graph): for kernel traces of the batched path")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-qps", default="300,700,1500,6000,30000")
    ap.add_argument("--c1-calls", type=int, default=200)
    ap.add_argument("--ingest-pages", type=int, default=128)
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

    sections = set(s.strip() for s in args.sections.split(",") if s.strip())
    if args.no_cpu_baseline:
        sections.discard("cpu")
    torch.cuda.set_device(0)
    if args.copy_calibration:
        # a plain copy of known size: reads 1 GiB and writes 1 GiB
        a = cle.create(C2_SHAPE, np.float32)
        cle.set(a, 1.0)
        b = cle.create(C2_SHAPE, np.float32)
        cle.copy(a, b)
        torch.cuda.synchronize()
        del a, b

    line = {"metric": "Mvoxels/s voronoi-otsu pipeline + % HBM roofline; p50/p99 serve latency",
            "value": None, "unit": "Mvoxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": None, "config": None, "roofline": None, "cpu_baseline": None}

    c3_host_crop = None
    if "c3" in sections:
        c3, c3_host_crop = c3_section(cle, torch, np, synthetic, args.steps, args.warmup, "cpu" in sections)
        line.update(c3["line"])
    extra = {}
    if "c1" in sections:
        extra["c1_512x512"] = c1_section(cle, torch, np, synthetic, args.c1_calls)
    if "m1" in sections:
        extra["random_2048x2048"] = m1_section(cle, torch, np)
    if "c2" in sections:
        extra["plain3d"] = c2_section(cle, torch, np, synthetic)
    if "batch" in sections and args.batch_images > 0:
        extra["batched"] = batch_section(cle, torch, args.batch_images, args.batch_only)
    if "serve" in sections and args.serve_requests > 0:
        extra["serving"] = serve_section(cle, torch, args.serve_requests,
                                         [float(q) for q in args.serve_qps.split(",")])
    if "ingest" in sections and args.ingest_pages > 0:
        extra["ingest"] = ingest_section(cle, torch, args.ingest_pages)
    if "cpu" in sections and c3_host_crop is not None:
        line["cpu_baseline"] = cpu_baseline(c3_host_crop)
    line.update(extra)
    print(json.dumps(line), flush=True)


# ----------------------------------------------------------------- headline
def c3_section(cle, torch, np, synthetic, steps, warmup, want_crop):
    t0 = time.time()
    vol_t, info = synthetic.nuclei_3d(C3_SHAPE, 100_000, semi=(3.0, 8.0), seed=3, touching_fraction=0.15)
    torch.cuda.synchronize()
    log(f"c3: {C3_SHAPE} uint16, {info['nuclei']} nuclei generated in {time.time() - t0:.1f}s")
    crop = None
    if want_crop:  # the centred 128 x 1024 x 1024 crop the CPU baseline runs on
        z0, y0, x0 = (C3_SHAPE[0] - 128) // 2, (C3_SHAPE[1] - 1024) // 2, (C3_SHAPE[2] - 1024) // 2
        crop = vol_t[z0:z0 + 128, y0:y0 + 1024, x0:x0 + 1024].contiguous().view(torch.int16).cpu().numpy().view(
            np.uint16)
    vol = cle.asarray(vol_t)
    labels = cle.create(C3_SHAPE, np.int32)  # float32 intermediates, int32 labels
    nvox = vol.size
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]

    def step(e=None):
        if e: e[0].record()
        cle.voronoi_otsu_labeling(vol, labels, 2.0, 1.5)
        if e: e[1].record()
        st = cle.statistics_of_labelled_pixels(vol, labels)
        if e: e[2].record()
        return st

    # the two blurs' launch shapes (z chunking, register budget) tuned for this volume
    # ahead of the warm-up, as the plain 3D section does for its blur
    tmp = cle.create(C3_SHAPE, np.float32)
    cle.autotune_gaussian_blur(vol, tmp, 2.0, 2.0, 2.0)
    cle.autotune_gaussian_blur(vol, tmp, 1.5, 1.5, 1.5)
    del tmp
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    cle.growth_counters(reset=True)
    t0 = time.perf_counter()
    for k in range(steps):
        stats = step(ev[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gc = cle.growth_counters(reset=True)
    ms = elapsed * 1e3 / steps
    vor_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    sta_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    n_labels = len(stats["label"])
    log(f"c3: {ms:.2f} ms per step (voronoi_otsu {vor_ms:.2f}, statistics {sta_ms:.2f}), {n_labels} labels; "
        f"growth pairs: {gc}")

    # dominant kernel: the growth's pair-of-steps launches.  Per launch, its algorithmic
    # bytes are the float32 growth state of the tiles it processes read once and written
    # once (8 B per voxel of 1024-voxel tiles); the launch time is the library's event
    # time around its launches (the launch stream), summed over the timed steps.
    pair_bytes = gc["tiles"] * GROWTH_TILE_VOXELS * 8
    pair_s = gc["ms"] / 1e3
    achieved = pair_bytes / pair_s / 1e9 if pair_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_vor_pair_run3 (Voronoi growth of 3D images, two steps per launch)",
                "launches": gc["launches"] // steps if steps else 0,
                "tiles_per_step": gc["tiles"] // steps if steps else 0,
                "algorithmic_bytes_per_launch": int(pair_bytes / max(1, gc["launches"])),
                "launch_ms": round(gc["ms"] / max(1, gc["launches"]), 4)}
    traffic = measured_traffic("voronoi_growth_pairs_512x2048x2048_pmc.txt")
    if traffic:
        roofline["traffic"] = traffic["bytes"]
        roofline["traffic_source"] = traffic["source"]
    # the whole step against its algorithmic bytes
    step_frac = C3_BYTES_PER_VOXEL * nvox / (ms / 1e3) / 1e9 / HBM_PEAK_GBS
    out = {"line": {
        "value": round(nvox / (ms / 1e3) / 1e6, 1), "ms_per_step": round(ms, 3),
        "data": f"synthetic (seeded nuclei stack, {info['nuclei']} nuclei, {n_labels} labels found)",
        "config": {"workload": "voronoi_otsu_labeling(spot_sigma=2.0, outline_sigma=1.5) + "
                               "statistics_of_labelled_pixels of a 512x2048x2048 uint16 nuclei stack "
                               "(100k nuclei, semi-axes 3-8, 15% touching pairs, seed 3); float32 intermediates, "
                               "int32 labels",
                   "voxels": nvox, "stage_ms": {"voronoi_otsu_labeling": round(vor_ms, 3),
                                                "statistics_of_labelled_pixels": round(sta_ms, 3)},
                   "growth": {"pair_launches_per_step": roofline["launches"],
                              "pair_ms_per_step": round(gc["ms"] / steps, 3),
                              "tiles_per_step": roofline["tiles_per_step"]},
                   "step_hbm_frac": round(step_frac, 4), "step_bytes_per_voxel": C3_BYTES_PER_VOXEL},
        "roofline": roofline}}
    del vol, labels, vol_t
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return out, crop


def measured_traffic(name):
    """HBM bytes per launch of a kernel from a committed PMC summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy of known size)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return {"bytes": int(line.split(":")[1]), "source": str(f.relative_to(ROOT))}
    return None


# ----------------------------------------------------------------- secondary keys
def c1_section(cle, torch, np, synthetic, calls):
    """512 x 512 uint8 blobs (60 spots, seed 1): voronoi_otsu_labeling(3.5, 1.0) and
    statistics on the float32 push, one call at a time from a host-resident image (push
    included), p50 / p99 over `calls` calls."""
    img = synthetic.blobs_2d((512, 512), 60, seed=1)

    def one():
        labels = cle.voronoi_otsu_labeling(img, spot_sigma=3.5, outline_sigma=1.0)
        return cle.statistics_of_labelled_pixels(cle.push(img), labels)

    for _ in range(10):
        one()
    torch.cuda.synchronize()
    lat = []
    for _ in range(calls):
        t0 = time.perf_counter()
        st = one()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    out = {"workload": "voronoi_otsu_labeling(3.5, 1.0) + statistics_of_labelled_pixels of a 512x512 uint8 blobs "
                       "image (60 spots, seed 1), pushed from the host per call",
           "calls": calls, "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
           "labels": len(st["label"])}
    log(f"c1: {out}")
    return out


def m1_section(cle, torch, np, runs=7):
    """The reference's published case: voronoi_otsu_labeling(spot 1, outline 1) of a
    2048 x 2048 numpy random image (float64, pushed per call), 530 ms on an Apple M1
    Max CPU (benchmarks/apple_m1_voronoi_otsu_labeling.ipynb).  Timed with the push,
    as the notebook's %timeit, and device-resident."""
    img = np.random.default_rng(0).random((2048, 2048))
    for _ in range(2):
        cle.voronoi_otsu_labeling(img, spot_sigma=1, outline_sigma=1)
    torch.cuda.synchronize()
    tp = []
    for _ in range(runs):
        t0 = time.perf_counter()
        out = cle.voronoi_otsu_labeling(img, spot_sigma=1, outline_sigma=1)
        torch.cuda.synchronize()
        tp.append(time.perf_counter() - t0)
    dev = cle.push(img)
    dst = cle.create(dev.shape, np.uint32)
    cle.voronoi_otsu_labeling(dev, dst, 1, 1)
    torch.cuda.synchronize()
    td = []
    for _ in range(runs):
        t0 = time.perf_counter()
        cle.voronoi_otsu_labeling(dev, dst, 1, 1)
        torch.cuda.synchronize()
        td.append(time.perf_counter() - t0)
    ms_p, ms_d = float(np.mean(tp)) * 1e3, float(np.mean(td)) * 1e3
    res = {"workload": "voronoi_otsu_labeling(1, 1) of a 2048x2048 float64 random image",
           "with_push_ms": round(ms_p, 3), "device_resident_ms": round(ms_d, 3),
           "labels": int(cle.maximum_of_all_pixels(out)), "published_m1max_cpu_ms": PUBLISHED_M1MAX_MS,
           "speedup_vs_published": round(PUBLISHED_M1MAX_MS / ms_p, 1)}
    log(f"m1: {res}")
    return res


def c2_section(cle, torch, np, synthetic, steps=10, warmup=2, blur_reps=20):
    """Plain 3D method: 256 x 1024 x 1024 uint16 nuclei (5000, seed 2) -> gaussian_blur
    (2, 2, 2) float32 -> threshold_otsu -> connected_components_labeling_box int32 ->
    statistics, with per-stage event times and the blur's roofline."""
    vol_t, info = synthetic.nuclei_3d(C2_SHAPE, 5000, seed=2)
    vol = cle.asarray(vol_t)
    nvox = vol.size
    blurred = cle.create(C2_SHAPE, np.float32)
    binary = cle.create(C2_SHAPE, np.uint8)
    labels = cle.create(C2_SHAPE, np.int32)
    stages = ["gaussian_blur", "threshold_otsu", "connected_components_labeling_box", "statistics_of_labelled_pixels"]

    def step(ev=None):
        if ev: ev[0].record()
        cle.gaussian_blur(vol, blurred, 2, 2, 2)
        if ev: ev[1].record()
        cle.threshold_otsu(blurred, binary)
        if ev: ev[2].record()
        cle.connected_components_labeling_box(binary, labels)
        if ev: ev[3].record()
        st = cle.statistics_of_labelled_pixels(vol, labels)
        if ev: ev[4].record()
        return st

    cle.autotune_gaussian_blur(vol, blurred, 2, 2, 2)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        st = step(events[k])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    stage_ms = {s: round(float(np.mean([events[k][i].elapsed_time(events[k][i + 1]) for k in range(steps)])), 3)
                for i, s in enumerate(stages)}
    kev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    kev[0].record()
    for _ in range(blur_reps):
        cle.gaussian_blur(vol, blurred, 2, 2, 2)
    kev[1].record()
    kev[1].synchronize()
    blur_ms = kev[0].elapsed_time(kev[1]) / blur_reps
    blur_gbs = nvox * 6 / (blur_ms / 1e3) / 1e9
    stage_bytes = {"threshold_otsu": 13, "connected_components_labeling_box": 5, "statistics_of_labelled_pixels": 6}
    out = {"workload": "256x1024x1024 uint16 nuclei -> gaussian_blur(2,2,2) f32 -> threshold_otsu -> "
                       "connected_components_labeling_box int32 -> statistics_of_labelled_pixels",
           "voxels": nvox, "ms": round(ms, 3), "mvox_s": round(nvox / (ms / 1e3) / 1e6, 1), "stage_ms": stage_ms,
           "stage_hbm_frac": {k: round(b * nvox / (stage_ms[k] / 1e3) / 1e9 / HBM_PEAK_GBS, 3)
                              for k, b in stage_bytes.items()},
           "labels": len(st["label"]),
           "blur": {"launch_ms": round(blur_ms, 4), "achieved_gb_s": round(blur_gbs, 1),
                    "frac": round(blur_gbs / HBM_PEAK_GBS, 4),
                    "valu_tflops": round(nvox * 102 / (blur_ms / 1e3) / 1e12, 2)}}
    log(f"c2: {out}")
    del vol, blurred, binary, labels, vol_t
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return out


def batch_section(cle, torch, n_images, batch_only=False):
    """Mixed-size batch (70 % 2D 64^2..1024^2, 30 % 3D 32x128x128..96x512x512, per-image
    sigmas): the plain per-image loop of voronoi_otsu_labeling against one batched call."""
    import numpy as np
    from pyclesperanto_prototype_amd import synthetic

    t0 = time.time()
    imgs_t, ss, so, info = synthetic.mixed_batch(n_images, seed=4)
    torch.cuda.synchronize()
    log(f"mixed batch: {info} generated in {time.time() - t0:.1f}s")
    imgs = [cle.asarray(t) for t in imgs_t]
    plain_out = [cle.create(t.shape, np.uint32) for t in imgs_t]
    batch_out = [cle.create(t.shape, np.uint32) for t in imgs_t]
    plain = float("nan")
    if not batch_only:
        for k in range(min(20, n_images)):  # warm-up of the per-image path
            cle.voronoi_otsu_labeling(imgs[k], plain_out[k], float(ss[k]), float(so[k]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(n_images):
            cle.voronoi_otsu_labeling(imgs[k], plain_out[k], float(ss[k]), float(so[k]))
        torch.cuda.synchronize()
        plain = time.perf_counter() - t0
    cle.voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out)  # warm-up (scratch growth)
    torch.cuda.synchronize()
    batched_default = tune = float("nan")
    if not batch_only:  # (a kernel trace of the batched path keeps the default launch shapes)
        t0 = time.perf_counter()
        cle.voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out)  # default launch shapes
        torch.cuda.synchronize()
        batched_default = time.perf_counter() - t0
        # the batched blur's launch shapes tuned for this batch's groups, ahead of the timed call
        t0 = time.perf_counter()
        cle.autotune_voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out)
        torch.cuda.synchronize()
        tune = time.perf_counter() - t0
        cle.voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out)
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, steps = cle.voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out, return_steps=True)
    torch.cuda.synchronize()
    batched = time.perf_counter() - t0
    same = sum(int(torch.equal(a.tensor, b.tensor)) for a, b in zip(plain_out, batch_out))
    nv = info["voxels"]
    if batch_only:
        return {"workload": "batched calls only (kernel trace)", "images": n_images, "voxels": nv,
                "batched_ms": round(batched * 1e3, 1), "growth_steps": steps}
    log(f"batched: plain {plain * 1e3:.1f} ms, batched {batched * 1e3:.1f} ms (default launch shapes "
        f"{batched_default * 1e3:.1f} ms, tuning {tune * 1e3:.0f} ms), identical {same}/{n_images}")
    # the same batch captured once as a hipGraph and replayed
    graph_out = [cle.create(t.shape, np.uint32) for t in imgs_t]
    t0 = time.perf_counter()
    plan = cle.voronoi_otsu_labeling_batch_plan(imgs, ss, so, destinations=graph_out, graph_pairs=32)
    torch.cuda.synchronize()
    build = time.perf_counter() - t0
    plan.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.run()
    torch.cuda.synchronize()
    replay = time.perf_counter() - t0
    same_g = sum(int(torch.equal(a.tensor, b.tensor)) for a, b in zip(plain_out, graph_out))
    plan.close()
    del graph_out
    log(f"graph: build {build * 1e3:.0f} ms, replay {replay * 1e3:.1f} ms, identical {same_g}/{n_images}")
    del imgs, imgs_t, plain_out, batch_out
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return {"workload": "voronoi_otsu_labeling of a mixed-size batch (70% 2D 64^2..1024^2, 30% 3D "
                        "32x128x128..96x512x512 uint16 nuclei, density 1e-4..1e-3, spot sigma 1-5, outline sigma "
                        "0.5-2, seed 4)",
            "images": n_images, "voxels": nv, "plain_ms": round(plain * 1e3, 1),
            "plain_mvox_s": round(nv / plain / 1e6, 1), "batched_ms": round(batched * 1e3, 1),
            "batched_mvox_s": round(nv / batched / 1e6, 1), "speedup_vs_plain": round(plain / batched, 2),
            "blur_autotune": {"default_shapes_ms": round(batched_default * 1e3, 1), "tuning_ms": round(tune * 1e3)},
            "growth_steps": steps, "identical_to_plain": f"{same}/{n_images}",
            "graph": {"build_ms": round(build * 1e3, 1), "replay_ms": round(replay * 1e3, 1),
                      "replay_mvox_s": round(nv / replay / 1e6, 1), "speedup_vs_plain": round(plain / replay, 2),
                      "identical_to_plain": f"{same_g}/{n_images}"}}


def serve_section(cle, torch, n_requests, qps_list):
    """Poisson arrivals of requests drawn from the mixed-size distribution plus 5 %
    256x1024x1024 nuclei stacks, at each offered load in qps_list: the continuous-
    batching server against the plain per-image path serving the same trace first-in
    first-out (p50 / p99 latency from arrival to labels ready, sustained Mvoxels/s)."""
    import numpy as np
    from pyclesperanto_prototype_amd import serving, synthetic

    rng = np.random.default_rng(5)
    n_large = max(1, int(round(0.05 * n_requests)))
    small_t, ss, so, info = synthetic.mixed_batch(n_requests - n_large, seed=5)
    vol_c2, _ = synthetic.nuclei_3d(C2_SHAPE, 5000, seed=2)
    vol_c2 = cle.asarray(vol_c2)
    reqs = [(cle.asarray(t), float(a), float(b)) for t, a, b in zip(small_t, ss, so)]
    large = [(vol_c2, float(rng.uniform(1, 5)), float(rng.uniform(0.5, 2))) for _ in range(n_large)]
    pos = np.sort(rng.choice(n_requests, n_large, replace=False))
    for p, r in zip(pos, large):
        reqs.insert(int(p), r)
    nv = int(sum(r[0].size for r in reqs))
    # reference labels: every request alone, in arrival order
    torch.cuda.synchronize()
    points = []
    for qps in qps_list:
        arrivals = serving.poisson_trace(n_requests, qps, seed=5)
        with serving.SegmentationServer(workers=3) as srv:
            served = serving.replay(srv, reqs, arrivals)
            batch_sizes = [b[0] for b in srv.batches]
        span = max(r.finished for r in served) - min(r.arrival for r in served)
        t0 = time.perf_counter()
        plain_lat, plain_out = [], []
        for (img, a, b), at in zip(reqs, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            out = cle.voronoi_otsu_labeling(img, None, a, b)
            torch.cuda.synchronize()
            plain_lat.append(time.perf_counter() - (t0 + at))
            plain_out.append(out)
        plain_span = time.perf_counter() - t0
        same = sum(int(torch.equal(r.result().tensor, o.tensor)) for r, o in zip(served, plain_out))
        pl = np.array(plain_lat) * 1e3
        lat = serving.latency_summary(served)
        pt = {"offered_qps": qps, "served": {**lat, "mvox_s": round(nv / span / 1e6, 1),
                                             "req_s": round(n_requests / span, 1), "batches": len(batch_sizes),
                                             "mean_batch": round(float(np.mean(batch_sizes)), 1)},
              "plain": {"p50_ms": round(float(np.percentile(pl, 50)), 2),
                        "p99_ms": round(float(np.percentile(pl, 99)), 2),
                        "mvox_s": round(nv / plain_span / 1e6, 1), "req_s": round(n_requests / plain_span, 1)},
              "identical_to_plain": f"{same}/{n_requests}"}
        pt["served_over_plain_mvox"] = round(pt["served"]["mvox_s"] / pt["plain"]["mvox_s"], 2)
        log(f"serving @ {qps:g} req/s: {pt}")
        points.append(pt)
        del served, plain_out
        torch.cuda.synchronize()
    del reqs, small_t, vol_c2
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return {"workload": f"{n_requests} Poisson-arrival requests: the mixed-size distribution plus {n_large} "
                        "256x1024x1024 nuclei stacks, voronoi_otsu_labeling per request",
            "requests": n_requests, "voxels": nv, "points": points}


def ingest_section(cle, torch, pages):
    """TIFF stack -> device: imread_stack (pinned double-buffered slabs, copies on their
    own stream overlapping the decode of the next slab) against decoding every page
    first and uploading the stack with one pageable copy."""
    import tempfile
    import numpy as np
    from PIL import Image
    from pyclesperanto_prototype_amd import io as cio
    from pyclesperanto_prototype_amd import synthetic

    vol_t, _ = synthetic.nuclei_3d((pages, 1024, 1024), 600, seed=2)
    host = vol_t.contiguous().view(torch.int16).cpu().numpy().view(np.uint16)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "stack.tif")
        ims = [Image.fromarray(p) for p in host]
        ims[0].save(f, save_all=True, append_images=ims[1:])
        del ims
        cle.imread_stack(f)  # page cache, pinned allocator
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a = cle.imread_stack(f)
        torch.cuda.synchronize()
        piped = time.perf_counter() - t0
        same = bool(torch.equal(a.tensor.view(torch.int16), vol_t.view(torch.int16)))
        del a
        t0 = time.perf_counter()
        stack = np.stack(list(cio.read_tiff_pages(f)))
        b = torch.from_numpy(stack.view(np.int16)).to("cuda")
        torch.cuda.synchronize()
        plain = time.perf_counter() - t0
        del b, stack
    nbytes = host.nbytes
    log(f"ingest: {pages} pages, pipelined {piped * 1e3:.1f} ms, decode-then-copy {plain * 1e3:.1f} ms")
    return {"workload": f"uint16 TIFF stack {pages}x1024x1024 ({nbytes / 2**20:.0f} MiB) from the page cache to HBM",
            "pipelined_ms": round(piped * 1e3, 1), "pipelined_gb_s": round(nbytes / piped / 1e9, 2),
            "decode_then_copy_ms": round(plain * 1e3, 1), "decode_then_copy_gb_s": round(nbytes / plain / 1e9, 2),
            "identical": same}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(crop):
    """The C/OpenMP restatement of the reference algorithms (float32 separable blurs,
    3^d maxima, row histogram + Otsu, Jacobi CCL and Voronoi growth to a fixed point,
    per-row statistics) on the centred 128 x 1024 x 1024 crop of the headline stack,
    with every host thread, and with one thread on a 32 x 512 x 512 part of it."""
    import numpy as np
    from oracle import c_oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c_oracle.lib()

    def run(sub, nthreads):
        c_oracle.set_threads(nthreads)
        a = sub.astype(np.float32)
        t0 = time.perf_counter()
        lab = c_oracle.voronoi_otsu_labeling(a, 2.0, 1.5)
        table = c_oracle.statistics_table(a, lab)
        return time.perf_counter() - t0, int(lab.max()), table.shape[1]

    dt, nl, _ = run(crop, threads)
    small = crop[48:80, 256:768, 256:768]
    dt1, _, _ = run(small, 1)
    c_oracle.set_threads(threads)
    log(f"cpu baseline: {crop.shape} in {dt:.2f}s ({threads} threads, {nl} labels); {small.shape} 1 thread {dt1:.2f}s")
    return {"value": round(crop.size / dt / 1e6, 2), "unit": "Mvoxels/s", "cores": threads, "kind": "port",
            "sample": f"voronoi_otsu_labeling(2.0, 1.5) + statistics of the centred 128x1024x1024 crop of the same "
                      f"stack, {dt:.1f}s",
            "cpu_model": cpu_model(),
            "one_thread": {"value": round(small.size / dt1 / 1e6, 2), "unit": "Mvoxels/s",
                           "sample": f"32x512x512 part of the crop, {dt1:.1f}s"},
            "published_m1max_cpu_mvox_s": round(2048 * 2048 / (PUBLISHED_M1MAX_MS / 1e3) / 1e6, 1),
            "note": "restated reference algorithm (C/OpenMP), not reference code"}


if __name__ == "__main__":
    main()
