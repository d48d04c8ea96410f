"""Host side of the I/O edges: TIFF page decoding and the regionprops table layout."""
import numpy as np
import pytest

from pyclesperanto_prototype_amd import io as cio


def _write_stack(path, stack):
    from PIL import Image
    pages = [Image.fromarray(p) for p in stack]
    pages[0].save(path, save_all=True, append_images=pages[1:])


def test_tiff_pages_roundtrip_uint16(tmp_path):
    rng = np.random.default_rng(3)
    stack = rng.integers(0, 65535, (7, 33, 45), dtype=np.uint16)
    f = tmp_path / "s.tif"
    _write_stack(f, stack)
    assert cio.tiff_info(f) == (7, 33, 45, np.dtype(np.uint16))
    got = np.stack(list(cio.read_tiff_pages(f)))
    assert got.dtype == np.uint16 and np.array_equal(got, stack)
    assert np.array_equal(np.stack(list(cio.read_tiff_pages(f, 2, 5))), stack[2:5])


def test_regionprops_matrix_from_statistics_dict(ref_tests):
    # the reference's 2D statistics test compares push_regionprops(result) against its
    # 36-row table: rows in STATISTICS_ENTRY order, one column per label
    t = ref_tests["statistics_2d"]
    d = {k: np.asarray(v, np.float32) for k, v in t["expected"].items()}
    d["original_label"] = d["label"]
    m = cio._region_matrix(d, first_row_index=1)
    assert m.shape == (36, 3)
    assert np.array_equal(m[0], [1, 2, 3])                               # IDENTIFIER
    assert np.array_equal(m[15], t["expected"]["area"])                  # PIXEL_COUNT
    assert np.allclose(m[14], [0.47140452, 0.4714045, 0.8164966])        # STANDARD_DEVIATION_INTENSITY
    assert np.allclose(m[35], [1.1396203, 1.1396204, 1.5])               # MAX_MEAN_DISTANCE_TO_CENTROID_RATIO
    for i, k in enumerate(cio.REGIONPROPS_ROWS):
        assert np.array_equal(m[i], d[k]), k


class _Region:
    """Attributes of a 2D region as skimage.measure.regionprops reports them."""

    def __init__(self, label, bbox, area, mean, lo, hi, centroid, wcentroid):
        self.label, self.bbox, self.area = label, bbox, area
        self.mean_intensity, self.min_intensity, self.max_intensity = mean, lo, hi
        self.centroid, self.weighted_centroid = centroid, wcentroid
        self.standard_deviation_intensity = 0.5

    def __getitem__(self, k):
        return getattr(self, k)


def test_regionprops_matrix_from_region_objects():
    regions = [_Region(1, (0, 0, 2, 2), 3, 0.5, 0, 1, (0.3, 0.4), (0.5, 0.6)),
               _Region(3, (1, 2, 4, 3), 3, 3.0, 2, 4, (2.0, 2.0), (2.2, 2.0))]
    m = cio._region_matrix(regions, first_row_index=1)
    assert m.shape == (36, 3)  # labels 1..3, row of label 2 left zero
    c = {k: i for i, k in enumerate(cio.REGIONPROPS_ROWS)}
    assert list(m[c["original_label"]]) == [1, 0, 3]
    # bbox (min_row, min_col, max_row, max_col) -> x = col, y = row, end inclusive
    assert list(m[[c["bbox_min_x"], c["bbox_min_y"], c["bbox_max_x"], c["bbox_max_y"]], 2]) == [2, 1, 2, 3]
    assert m[c["bbox_width"], 2] == 1 and m[c["bbox_height"], 2] == 3 and m[c["bbox_depth"], 2] == 1
    assert m[c["sum_intensity"], 2] == 9.0
    assert m[c["centroid_x"], 0] == 0.4 and m[c["centroid_y"], 0] == 0.3
    assert m[c["mass_center_x"], 0] == 0.6 and m[c["mass_center_y"], 0] == 0.5
    col = cio._region_column(regions, "area")
    assert col.shape == (1, 4) and list(col[0]) == [0, 3, 0, 3]


def test_regionprops_column_from_dict():
    d = {"area": np.asarray([1, 2, 2, 4], np.float32)}
    assert np.array_equal(cio._region_column(d, "area"), [[1, 2, 2, 4]])
