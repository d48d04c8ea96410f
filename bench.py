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
VALU_PEAK_TFLOPS = 78.6  # MI355X float32 vector FMA without packed math (2 flops x 64 lanes x 4 SIMD x 256 CU x 2.4 GHz)

SHAPE = (256, 1024, 1024)
N_NUCLEI = 5000
SEED = 2
SIGMA = 2.0
BLUR_REPS = 20  # back-to-back blur launches timed for the roofline


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
    ap.add_argument("--batch-images", type=int, default=2000, help="mixed-size batch (0 = skip)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batched calls only (no per-image loop, no graph): for kernel traces of the batched path")
    ap.add_argument("--serve-requests", type=int, default=400, help="serving trace length (0 = skip)")
    ap.add_argument("--serve-qps", type=float, default=300.0)
    ap.add_argument("--large", type=int, default=1,
                    help="also time voronoi_otsu_labeling + statistics of the 512x2048x2048 stack (0 = skip)")
    ap.add_argument("--ingest-pages", type=int, default=128,
                    help="TIFF ingest: pages of the 1024 x 1024 uint16 stack written and read back (0 = skip)")
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

    cle.autotune_gaussian_blur(vol, blurred, SIGMA, SIGMA, SIGMA)  # z chunking of this blur, ahead of time
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats = step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_labels = len(stats["label"])
    ms = elapsed * 1e3 / args.steps
    stage_ms = {s: float(np.mean([events[k][i].elapsed_time(events[k][i + 1]) for k in range(args.steps)]))
                for i, s in enumerate(stages)}
    log("stage ms:", json.dumps({k: round(v, 3) for k, v in stage_ms.items()}))

    # dominant kernel: the blur (one fused launch) -- algorithmic bytes = read the
    # uint16 source once (2 B/voxel) + write the float32 result once (4 B/voxel).  Its
    # launch duration: the same launch repeated back to back between two events on the
    # launch stream, so that the GPU never waits for the host (the stage time above
    # also holds the host's launch latency: the step starts on an idle GPU)
    blur_bytes = nvox * (2 + 4)
    kev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    cle.gaussian_blur(vol, blurred, SIGMA, SIGMA, SIGMA)
    kev[0].record()
    for _ in range(BLUR_REPS):
        cle.gaussian_blur(vol, blurred, SIGMA, SIGMA, SIGMA)
    kev[1].record()
    kev[1].synchronize()
    blur_launch_ms = kev[0].elapsed_time(kev[1]) / BLUR_REPS
    log(f"gaussian_blur launch: {blur_launch_ms:.3f} ms (mean of {BLUR_REPS} back-to-back launches)")
    blur_s = blur_launch_ms / 1e3
    achieved = blur_bytes / blur_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "gaussian_blur (u16 -> f32, 17 taps/axis)", "algorithmic_bytes_per_launch": blur_bytes,
                "launch_ms": round(blur_launch_ms, 4)}

    traffic = measured_traffic()
    if traffic:
        roofline["traffic"] = traffic["bytes"]
        roofline["traffic_source"] = traffic["source"]
    # the blur is VALU-bound (17 taps per axis): its place under the float32 vector roof
    flops = nvox * 3 * 17 * 2
    roofline["valu"] = {"achieved_tflops": round(flops / blur_s / 1e12, 2), "peak_tflops": VALU_PEAK_TFLOPS,
                        "frac": round(flops / blur_s / 1e12 / VALU_PEAK_TFLOPS, 4),
                        "flops_per_voxel": 3 * 17 * 2}
    # algorithmic bytes per stage: inputs read once, outputs written once at their dtypes
    # (Otsu reads the float32 image for min/max, histogram and compare: 3 x 4 + 1)
    stage_bytes = {"threshold_otsu": nvox * 13, "connected_components_labeling_box": nvox * 5,
                   "statistics_of_labelled_pixels": nvox * (4 + 2)}
    stage_roof = {k: round(b / (stage_ms[k] / 1e3) / 1e9 / HBM_PEAK_GBS, 3) for k, b in stage_bytes.items()}

    batched = batch_section(cle, torch, args.batch_images, args.batch_only) if args.batch_images > 0 else None
    serving = serve_section(cle, torch, vol, args.serve_requests, args.serve_qps) if args.serve_requests > 0 else None

    ingest = ingest_section(cle, torch, vol_t, args.ingest_pages) if args.ingest_pages > 0 else None
    large = large_section(cle, torch) if args.large else None

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
                   "voxels": nvox, "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
                   "stage_hbm_frac": stage_roof},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "ingest": ingest,
        "large": large,
    }
    print(json.dumps(line), flush=True)


def measured_traffic():
    """HBM bytes per launch of the blur from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy of known size)."""
    f = ROOT / "profiles" / "gaussian_blur_u16_256x1024x1024_pmc.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return {"bytes": int(line.split(":")[1]), "source": str(f.relative_to(ROOT))}
    return None


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
    t0 = time.perf_counter()
    _, steps = cle.voronoi_otsu_labeling_batch(imgs, ss, so, destinations=batch_out, return_steps=True)
    torch.cuda.synchronize()
    batched = time.perf_counter() - t0
    same = sum(int(torch.equal(a.tensor, b.tensor)) for a, b in zip(plain_out, batch_out))
    nv = info["voxels"]
    if not batch_only:
        log(f"batched: plain {plain * 1e3:.1f} ms, batched {batched * 1e3:.1f} ms, identical {same}/{n_images}")
    if batch_only:
        return {"workload": "batched calls only (kernel trace)", "images": n_images, "voxels": nv,
                "batched_ms": round(batched * 1e3, 1), "growth_steps": steps}
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
            "growth_steps": steps, "identical_to_plain": f"{same}/{n_images}",
            "graph": {"build_ms": round(build * 1e3, 1), "replay_ms": round(replay * 1e3, 1),
                      "replay_mvox_s": round(nv / replay / 1e6, 1), "speedup_vs_plain": round(plain / replay, 2),
                      "identical_to_plain": f"{same_g}/{n_images}"}}


def serve_section(cle, torch, vol_c2, n_requests, qps):
    """Poisson arrivals of requests drawn from the mixed-size distribution plus 5 %
    256x1024x1024 nuclei stacks: the continuous-batching server against the plain
    per-image path serving the same trace first-in first-out."""
    import numpy as np
    from pyclesperanto_prototype_amd import serving, synthetic

    rng = np.random.default_rng(5)
    n_large = max(1, int(round(0.05 * n_requests)))
    small_t, ss, so, info = synthetic.mixed_batch(n_requests - n_large, seed=5)
    reqs = [(cle.asarray(t), float(a), float(b)) for t, a, b in zip(small_t, ss, so)]
    large = [(vol_c2, float(rng.uniform(1, 5)), float(rng.uniform(0.5, 2))) for _ in range(n_large)]
    pos = np.sort(rng.choice(n_requests, n_large, replace=False))
    for p, r in zip(pos, large):
        reqs.insert(int(p), r)
    arrivals = serving.poisson_trace(n_requests, qps, seed=5)
    nv = int(sum(r[0].size for r in reqs))
    torch.cuda.synchronize()
    # continuous batching
    with serving.SegmentationServer(workers=3) as srv:
        served = serving.replay(srv, reqs, arrivals)
        batch_sizes = [b[0] for b in srv.batches]
    span = max(r.finished for r in served) - min(r.arrival for r in served)
    # the plain per-image path, first-in first-out on the same arrivals
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
    log(f"serving: {lat} batches {len(batch_sizes)}, plain p50 "
        f"{np.percentile(pl, 50):.1f} p99 {np.percentile(pl, 99):.1f} ms, identical {same}/{n_requests}")
    del reqs, served, plain_out, small_t
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return {"workload": f"Poisson arrivals at {qps:g} requests/s: mixed-size requests (as the batch) plus "
                        f"{n_large} 256x1024x1024 nuclei stacks, voronoi_otsu_labeling per request",
            "requests": n_requests, "voxels": nv, "served": {**lat,
                                                              "mvox_s": round(nv / span / 1e6, 1),
                                                              "batches": len(batch_sizes),
                                                              "mean_batch": round(float(np.mean(batch_sizes)), 1)},
            "plain": {"p50_ms": round(float(np.percentile(pl, 50)), 2),
                      "p99_ms": round(float(np.percentile(pl, 99)), 2),
                      "mvox_s": round(nv / plain_span / 1e6, 1)},
            "identical_to_plain": f"{same}/{n_requests}"}


def large_section(cle, torch, reps=3):
    """512 x 2048 x 2048 uint16 stack (about 100 k nuclei, semi-axes 3-8, 15 % touching
    pairs, seed 3) resident in HBM: voronoi_otsu_labeling(spot 2.0, outline 1.5) and
    statistics_of_labelled_pixels, timed over `reps` runs after one warm-up."""
    import numpy as np
    from pyclesperanto_prototype_amd import synthetic

    shape = (512, 2048, 2048)
    t0 = time.time()
    vol_t, info = synthetic.nuclei_3d(shape, 100_000, semi=(3.0, 8.0), seed=3, touching_fraction=0.15)
    torch.cuda.synchronize()
    log(f"large: {info['nuclei']} nuclei generated in {time.time() - t0:.1f}s")
    vol = cle.asarray(vol_t)
    labels = cle.create(shape, np.uint32)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def run(timed):
        if timed: ev[0].record()
        cle.voronoi_otsu_labeling(vol, labels, 2.0, 1.5)
        if timed: ev[1].record()
        st = cle.statistics_of_labelled_pixels(vol, labels)
        if timed: ev[2].record()
        return st

    run(False)
    torch.cuda.synchronize()
    vor, sta, tot = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        st = run(True)
        torch.cuda.synchronize()
        tot.append(time.perf_counter() - t0)
        vor.append(ev[0].elapsed_time(ev[1]))
        sta.append(ev[1].elapsed_time(ev[2]))
    nv = int(np.prod(shape))
    ms = float(np.mean(tot)) * 1e3
    log(f"large: {ms:.1f} ms per run (voronoi_otsu {np.mean(vor):.1f} ms, statistics {np.mean(sta):.1f} ms), "
        f"{len(st['label'])} labels")
    out = {"workload": "voronoi_otsu_labeling(2.0, 1.5) + statistics_of_labelled_pixels of a 512x2048x2048 uint16 "
                       f"nuclei stack ({info['nuclei']} nuclei, 15% touching pairs, seed 3)",
           "voxels": nv, "ms": round(ms, 2), "mvox_s": round(nv / (ms / 1e3) / 1e6, 1),
           "voronoi_otsu_ms": round(float(np.mean(vor)), 2), "statistics_ms": round(float(np.mean(sta)), 2),
           "labels": len(st["label"])}
    del vol, labels, vol_t
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    cle.release_pool()
    return out


def ingest_section(cle, torch, vol_t, pages):
    """TIFF stack -> device: imread_stack (pinned double-buffered slabs, copies on their
    own stream overlapping the decode of the next slab) against decoding every page
    first and uploading the stack with one pageable copy.  Both read the same file
    from the page cache (a first untimed read warms it)."""
    import tempfile
    import numpy as np
    from PIL import Image
    from pyclesperanto_prototype_amd import io as cio

    pages = min(pages, vol_t.shape[0])
    host = vol_t[:pages].contiguous().view(torch.int16).cpu().numpy().view(np.uint16)
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
        same = bool(torch.equal(a.tensor.view(torch.int16), vol_t[:pages].view(torch.int16)))
        del a
        t0 = time.perf_counter()
        stack = np.stack(list(cio.read_tiff_pages(f)))
        b = torch.from_numpy(stack.view(np.int16)).to("cuda")
        torch.cuda.synchronize()
        plain = time.perf_counter() - t0
        del b, stack
    nbytes = host.nbytes
    log(f"ingest: {pages} pages, pipelined {piped * 1e3:.1f} ms, decode-then-copy {plain * 1e3:.1f} ms")
    return {"workload": f"uint16 TIFF stack {pages}x{host.shape[1]}x{host.shape[2]} ({nbytes / 2**20:.0f} MiB) "
                        f"from the page cache to HBM",
            "pipelined_ms": round(piped * 1e3, 1), "pipelined_gb_s": round(nbytes / piped / 1e9, 2),
            "decode_then_copy_ms": round(plain * 1e3, 1), "decode_then_copy_gb_s": round(nbytes / plain / 1e9, 2),
            "identical": same}


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
