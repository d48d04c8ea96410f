"""Continuous-batching segmentation server on one GPU.

Requests (an image in device memory plus its spot / outline sigmas) arrive at any
time.  Worker threads, each with its own HIP stream, take the pending requests that
fit the memory budget (in arrival order, up to a voxel cap per batch) as one batch
and run voronoi_otsu_labeling_batch on it (a batch of one request runs the
single-image pipeline); the request's intermediates live in device memory for the
whole pipeline and only the label image is handed back.  While one batch is on the
GPU the next worker admits the requests that arrived meanwhile, so batches overlap on
the device.  Admission is by memory: a request is admitted when the bytes of the
requests in flight plus its own estimate stay within the budget (by default 80 % of
the HBM free when the server starts, less the workers' scratch reservations).  The
label images are allocated on the submitting thread's stream, and each worker's
scratch arena is sized once when it starts, so a running server makes no device
allocation of its own.

The batched pipeline gives every image the result it gets alone
(cle_voronoi_otsu_labeling_batch), so a request's labels do not depend on the
requests it is batched with.
"""
from __future__ import annotations

import ctypes
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._array import Array, asarray, create
from ._ops import voronoi_otsu_labeling, voronoi_otsu_labeling_batch

# device bytes per voxel of one request while its batch runs: the uint32 labels, the
# batch's float32 / uint8 work images (10 B), the 2x2x2-block CCL arrays and lists
BYTES_PER_VOXEL = 24
# Library scratch of a batch: at most 16 B per voxel (float32 blur / growth buffers,
# uint8 seeds and segmentation, CCL block arrays, tile lists) plus tables.  Each
# worker's stream arena is sized for a batch of max_batch_voxels when the worker starts
# and grown (once) before a larger single request, so a running server makes no driver
# allocation: a hipFree / hipMalloc while other workers' work is in flight stalled every
# worker for up to a second (HIP runtime trace of a saturated run).
SCRATCH_BYTES_PER_VOXEL = 16
SCRATCH_SLACK = 64 << 20


def scratch_bytes(voxels: int) -> int:
    return int(voxels) * SCRATCH_BYTES_PER_VOXEL + SCRATCH_SLACK


@dataclass
class Request:
    image: Array
    spot_sigma: float
    outline_sigma: float
    ready: object = None  # event recorded on the submitting thread's stream
    client_stream: object = None  # the submitting thread's stream (where the labels will be used)
    labels: Array = None  # created on the client's stream at submit time
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    error: BaseException = None
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def nbytes(self) -> int:
        return int(self.image.size) * BYTES_PER_VOXEL

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    def result(self, timeout=None) -> Array:
        if not self.done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.labels


class SegmentationServer:
    """voronoi_otsu_labeling as a service with continuous batching.

    server = SegmentationServer(); server.start()
    req = server.submit(image, spot_sigma=2, outline_sigma=1); labels = req.result()
    server.stop()
    """

    def __init__(self, workers: int = 3, memory_budget: int = None, max_batch: int = 4096,
                 max_batch_voxels: int = 1 << 28):
        import torch

        # Each worker owns a stream; the process has 4 hardware queues (HIP's default),
        # so more than 3 workers (plus the clients' stream) share queues and serialise:
        # measured on 400 mixed requests with 20 256x1024x1024 ones, 3 workers 24.1
        # Gvoxels/s, 6 workers 2.9, 8 workers 1.5.
        self.workers = workers
        self.max_batch = max_batch
        # a batch completes as a whole: capping its voxels keeps a large request from
        # holding up the small ones batched with it (a larger request runs alone).
        # Measured, same trace: 2^25 voxels 20.3 Gvoxels/s (mean batch 6.9), 2^28 24.1
        # (mean batch 10.0).
        self.max_batch_voxels = max_batch_voxels
        self.reserve_bytes = scratch_bytes(max_batch_voxels)
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info()
            memory_budget = int(free * 0.8) - workers * self.reserve_bytes
        self.budget = int(memory_budget)
        self._pending = deque()
        self._lock = threading.Lock()
        self._cv = threading.Condition(self._lock)
        self._inflight = 0
        self._stop = False
        self._threads = []
        self.batches = []  # (size, voxels, seconds) per batch run

    # ------------------------------------------------------------------ client side
    def submit(self, image, spot_sigma: float = 2.0, outline_sigma: float = 2.0) -> Request:
        import torch

        img = image if isinstance(image, Array) else asarray(image)
        client = torch.cuda.current_stream()
        req = Request(img, float(spot_sigma), float(outline_sigma), client_stream=client,
                      arrival=time.perf_counter())
        if req.nbytes > self.budget:
            req.error = MemoryError(f"request of {req.nbytes} bytes exceeds the server budget {self.budget}")
            req.finished = req.arrival
            req.done.set()
            return req
        # The labels come from the client stream's pool, where the client will use and
        # free them (memory the client freed earlier is reused here, whichever worker
        # runs the request); the worker's stream waits for the image and for the
        # client's earlier work on that memory.
        req.labels = create(img.shape, np.uint32)
        req.ready = torch.cuda.Event()
        req.ready.record(client)
        with self._cv:
            self._pending.append(req)
            self._cv.notify()
        return req

    def start(self):
        import torch

        dev = torch.cuda.current_device()
        for k in range(self.workers):
            t = threading.Thread(target=self._worker, args=(dev,), name=f"cle-serve-{k}", daemon=True)
            t.start()
            self._threads.append(t)
        return self

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._threads = []

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # ------------------------------------------------------------------ scheduling
    def _admit(self):
        """Pending requests, in arrival order, that fit the budget next to the work in
        flight (at least one when nothing is in flight, so no request starves)."""
        batch, total, vox = [], 0, 0
        while self._pending and len(batch) < self.max_batch:
            r = self._pending[0]
            if self._inflight + total + r.nbytes > self.budget and (batch or self._inflight):
                break
            if batch and vox + int(r.image.size) > self.max_batch_voxels:
                break
            self._pending.popleft()
            batch.append(r)
            total += r.nbytes
            vox += int(r.image.size)
        return batch, total

    def _worker(self, dev):
        import torch

        torch.cuda.set_device(dev)
        stream = torch.cuda.Stream()
        try:
            self._serve(stream)
        finally:
            # the library's scratch arena of this stream goes with it
            _lib.call("cle_release_stream", ctypes.c_void_p(stream.cuda_stream))

    def _serve(self, stream):
        import torch

        # best effort: without the reservation the arena grows on first use
        reserved = self.reserve_bytes
        _lib.call_rc("cle_reserve_stream", ctypes.c_void_p(stream.cuda_stream), ctypes.c_size_t(reserved))
        while True:
            with self._cv:
                while not self._stop and not self._pending:
                    self._cv.wait()
                if self._stop and not self._pending:
                    return
                batch, nbytes = self._admit()
                if not batch:
                    self._cv.wait(0.001)
                    continue
                self._inflight += nbytes
            t0 = time.perf_counter()
            need = scratch_bytes(sum(int(r.image.size) for r in batch))
            if need > reserved:  # a request above the batch cap: grow while this stream is idle
                if _lib.call_rc("cle_reserve_stream", ctypes.c_void_p(stream.cuda_stream), ctypes.c_size_t(need)) == 0:
                    reserved = need
            try:
                with torch.cuda.stream(stream):
                    outs = [r.labels for r in batch]
                    for r, o in zip(batch, outs):
                        stream.wait_event(r.ready)
                        # written on this stream: a later reuse of the block on the
                        # client's stream waits for this stream's work
                        o.tensor.record_stream(stream)
                    if len(batch) == 1:  # alone: the single-image pipeline
                        r = batch[0]
                        voronoi_otsu_labeling(r.image, outs[0], r.spot_sigma, r.outline_sigma)
                    else:
                        voronoi_otsu_labeling_batch([r.image for r in batch], [r.spot_sigma for r in batch],
                                                    [r.outline_sigma for r in batch], destinations=outs)
                    stream.synchronize()
                err = None
            except BaseException as e:  # hand the failure to every request of the batch
                outs, err = [None] * len(batch), e
            t1 = time.perf_counter()
            with self._cv:
                self._inflight -= nbytes
                self.batches.append((len(batch), int(sum(r.image.size for r in batch)), t1 - t0))
                self._cv.notify_all()
            for r, o in zip(batch, outs):
                r.admitted = t0
                r.labels = o
                r.error = err
                r.finished = t1
                r.done.set()


def poisson_trace(n: int, qps: float, seed: int = 5) -> np.ndarray:
    """Arrival offsets (seconds) of n requests with exponential inter-arrival times."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, n))


def replay(server: SegmentationServer, requests, arrivals) -> list:
    """Submit `requests` [(image, spot_sigma, outline_sigma)] at the given arrival
    offsets and wait for all of them; returns the Request objects."""
    t0 = time.perf_counter()
    out = []
    for (img, ss, so), at in zip(requests, arrivals):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        out.append(server.submit(img, ss, so))
    for r in out:
        r.done.wait()
    return out


def latency_summary(reqs) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    return {"p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_ms": round(float(lat.mean()), 2)}
