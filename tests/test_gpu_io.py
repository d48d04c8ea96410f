"""I/O edges on the device: TIFF stack ingest and the regionprops table."""
import numpy as np
import pytest

from conftest import load_npz
from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu


def _write_stack(path, stack):
    from PIL import Image
    pages = [Image.fromarray(p) for p in stack]
    pages[0].save(path, save_all=True, append_images=pages[1:])


@pytest.mark.parametrize("pages,slab", [(40, 16), (5, 16), (1, 4), (9, 1)])
def test_imread_stack_uint16(cle, tmp_path, pages, slab):
    import torch
    rng = np.random.default_rng(pages)
    stack = rng.integers(0, 65535, (pages, 37, 51), dtype=np.uint16)
    f = tmp_path / "s.tif"
    _write_stack(f, stack)
    a = cle.imread_stack(f, slab_pages=slab)
    want = stack[0] if pages == 1 else stack
    assert a.dtype == np.uint16 and a.shape == want.shape
    assert np.array_equal(cle.pull(a), want)
    # the device stack feeds the hot path directly (uint16 read by the blur)
    blurred = cle.gaussian_blur(a, sigma_x=1, sigma_y=1, sigma_z=1 if pages > 1 else 0)
    np.testing.assert_allclose(cle.pull(blurred), O.gaussian_blur(want.astype(np.float32), 1, 1, 1 if pages > 1 else 0),
                               atol=1e-4 * 65535)


def test_imread_float32_like_push(cle, tmp_path):
    stack = np.random.default_rng(1).integers(0, 255, (3, 20, 30), dtype=np.uint8)
    f = tmp_path / "u8.tif"
    _write_stack(f, stack)
    with pytest.warns(UserWarning):
        a = cle.imread(f)
    assert a.dtype == np.float32 and np.array_equal(cle.pull(a), stack.astype(np.float32))


def test_push_regionprops_reference_table(cle, ref_tests):
    # tests/test_statistics_of_labelled_pixels.py:5-110 of the reference: the pushed
    # regionprops equal its 36-row table at rtol 1e-4
    t = ref_tests["statistics_2d"]
    res = cle.statistics_of_labelled_pixels(np.asarray(t["intensity"], np.float32),
                                           np.asarray(t["labels"], np.float32))
    table = cle.pull(cle.push_regionprops(res, first_row_index=1))
    assert table.shape == (36, 3)
    from pyclesperanto_prototype_amd.io import REGIONPROPS_ROWS
    for i, k in enumerate(REGIONPROPS_ROWS):
        want = t["expected"]["label"] if k == "original_label" else t["expected"][k]
        assert np.allclose(table[i], want, 1e-4), k
    # the device-side table of the same statistics, in the same layout
    dev, n = cle.statistics_table(cle.push(np.asarray(t["intensity"])), cle.push(np.asarray(t["labels"])))
    assert n == 3 and np.allclose(cle.pull(dev)[1:, :3], table[1:], 1e-6)


def test_push_regionprops_column(cle):
    # tests/test_push_regionprops_column.py:4-24 of the reference
    labels = cle.push(np.asarray([[[4, 4, 2]], [[4, 1, 2]], [[3, 3, 4]]]))
    rp = cle.statistics_of_background_and_labelled_pixels(None, labels)
    areas = cle.pull(cle.push_regionprops_column(rp, "area"))
    assert np.allclose([0, 1, 2, 2, 4], areas, 0.0001)
