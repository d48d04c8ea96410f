import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def ref_tests():
    with open(GOLDEN / "reference_tests.json") as f:
        return json.load(f)


def load_npz(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def cle():
    """The package with its HIP library loaded (GPU tests only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyclesperanto_prototype_amd as cle
    cle._load_library()
    return cle


def blobs_image(shape, n, seed, sig=(3, 6), amp=(100, 200), noise=5.0, dtype=np.uint8):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    img = np.full(shape, 10.0)
    for _ in range(n):
        c = [rng.uniform(0, s) for s in shape]
        s = rng.uniform(*sig)
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        img += rng.uniform(*amp) * np.exp(-r2 / (2 * s * s))
    img += rng.normal(0, noise, shape)
    hi = 255 if dtype == np.uint8 else 65535
    return np.clip(img, 0, hi).astype(dtype)
