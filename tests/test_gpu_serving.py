"""Continuous batching: every request's labels are the same alone and in the mix."""
import numpy as np
import pytest

from conftest import blobs_image

pytestmark = pytest.mark.gpu


def test_server_results_equal_alone(cle):
    import torch
    from pyclesperanto_prototype_amd import serving, synthetic

    rng = np.random.default_rng(60)
    reqs = []
    for k in range(24):
        if k % 4 == 3:
            img, _ = synthetic.nuclei_overlapping((int(rng.integers(10, 30)), 96, 96), 4e-4, seed=600 + k)
            img = cle.asarray(img)
        else:
            sh = (int(rng.integers(40, 200)), int(rng.integers(40, 200)))
            img = cle.asarray((blobs_image(sh, 8, 600 + k).astype(np.float32) * 30).astype(np.uint16))
        reqs.append((img, float(rng.uniform(1, 4)), float(rng.uniform(0.5, 2))))
    alone = [cle.voronoi_otsu_labeling(im, None, a, b) for im, a, b in reqs]
    # one burst (every request within a few microseconds): requests queue behind the
    # first batch, so later ones are batched together whatever the kernels' speed
    arrivals = serving.poisson_trace(len(reqs), qps=1e6, seed=61)
    with serving.SegmentationServer(workers=2) as srv:
        done = serving.replay(srv, reqs, arrivals)
        sizes = [b[0] for b in srv.batches]
    assert max(sizes) > 1, sizes  # requests were batched together
    for r, want in zip(done, alone):
        assert torch.equal(r.result().tensor, want.tensor)
    lat = serving.latency_summary(done)
    assert lat["p99_ms"] >= lat["p50_ms"] > 0


def test_server_rejects_oversized_request(cle):
    from pyclesperanto_prototype_amd import serving

    srv = serving.SegmentationServer(workers=1, memory_budget=1000)
    r = srv.submit(cle.create((64, 64), np.float32), 2.0, 1.0)
    with pytest.raises(MemoryError):
        r.result(timeout=1)


def test_server_labels_survive_queued_client_reads(cle):
    # The client frees its labels while a read of them is still queued on its own
    # stream (behind other work), then submits again: the next batch must not be given
    # that memory before the queued read has run.
    import torch
    from pyclesperanto_prototype_amd import serving, synthetic

    a = cle.asarray(synthetic.nuclei_overlapping((256, 256), 5e-4, seed=90)[0])
    b = cle.asarray(synthetic.nuclei_overlapping((256, 256), 5e-4, seed=91)[0])
    want_a = cle.voronoi_otsu_labeling(a, None, 2.0, 1.0).tensor.clone()
    want_b = cle.voronoi_otsu_labeling(b, None, 2.0, 1.0).tensor.clone()
    assert not torch.equal(want_a, want_b)
    busy = torch.ones(1 << 28, device="cuda")
    torch.cuda.synchronize()
    client = torch.cuda.Stream()
    with serving.SegmentationServer(workers=1) as srv:
        with torch.cuda.stream(client):
            r = srv.submit(a, 2.0, 1.0)
            lab = r.result()
            r.labels = None
            for _ in range(100):
                busy.mul_(1.0)           # client work queued ahead of the read
            copy = lab.tensor.clone()    # the read of the labels, queued behind it
            del lab
            r2 = srv.submit(b, 2.0, 1.0)
            got_b = r2.result().tensor.clone()
        client.synchronize()
    assert torch.equal(copy, want_a)
    assert torch.equal(got_b, want_b)
