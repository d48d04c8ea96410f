"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer / binary / label outputs must be bit-exact; float32 blur within max-abs
1e-4 relative to the image maximum; float statistics within 1e-5 relative.
"""
import numpy as np
import pytest

from conftest import blobs_image, load_npz
from oracle import cle_oracle as O

pytestmark = pytest.mark.gpu

BLUR_TOL = 1e-4    # max |gpu - oracle| <= BLUR_TOL * max(1, max |oracle|)
STATS_RTOL = 1e-5


def assert_blur_close(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert got.shape == want.shape
    err = np.abs(got - want).max() if got.size else 0.0
    assert err <= BLUR_TOL * max(1.0, np.abs(want).max()), f"max abs err {err}"


# ----------------------------------------------------------------- reference test vectors
def test_reference_vectors(cle, ref_tests):
    A = np.asarray
    t = ref_tests["detect_maxima_box"]
    out = cle.create_like(cle.push(A(t["input"])))
    cle.detect_maxima_box(A(t["input"]), out)
    assert np.array_equal(cle.pull(out), A(t["expected"]))
    for k in ("histogram_2d", "histogram_3d", "histogram_3d_2"):
        t = ref_tests[k]
        assert np.allclose(cle.pull(cle.histogram(A(t["input"]), num_bins=t["num_bins"])), t["expected"]), k
    t = ref_tests["greater_constant"]
    r = cle.create(cle.push(A(t["input"])))
    cle.greater_constant(A(t["input"]), r, t["constant"])
    assert np.array_equal(cle.pull(r), A(t["expected"]))
    t = ref_tests["binary_and"]
    assert np.array_equal(cle.pull(cle.binary_and(A(t["a"]), A(t["b"]))), A(t["expected"]))
    t = ref_tests["mask"]
    r = cle.create(cle.push(A(t["input"])))
    cle.mask(A(t["input"]), A(t["mask"]), r)
    assert np.array_equal(cle.pull(r), A(t["expected"]))
    t = ref_tests["set_nonzero_pixels_to_pixel_index"]
    assert np.array_equal(cle.pull(cle.set_non_zero_pixels_to_pixel_index(A(t["input"]))), A(t["expected"]))
    for kind in ("box", "diamond"):
        t = ref_tests["nonzero_minimum_" + kind]
        src = cle.push(A(t["input"]))
        res = cle.create(src)
        cle.set(res, 0)
        flag = cle.create((1, 1, 1))
        cle.set(flag, 0)
        getattr(cle, "nonzero_minimum_" + kind)(src, flag, res)
        assert np.allclose(cle.pull(res), A(t["expected"]), atol=1e-5)
        assert cle.pull(flag).ravel()[0] == 1
        t = ref_tests["onlyzero_overwrite_maximum_" + kind]
        flag = cle.create((1, 1, 1))
        cle.set(flag, 0)
        res = getattr(cle, "onlyzero_overwrite_maximum_" + kind)(A(t["input"]), flag)
        assert np.array_equal(cle.pull(res), A(t["expected"]))
        t = ref_tests["connected_components_labeling_" + kind]
        out = getattr(cle, "connected_components_labeling_" + kind)(A(t["input"]))
        assert out.dtype == np.uint32 and np.array_equal(cle.pull(out), A(t["expected"]))
    t = ref_tests["relabel_sequential"]
    out = cle.create_like(cle.push(A(t["input"])))
    cle.relabel_sequential(A(t["input"]), out)
    assert np.array_equal(cle.pull(out), A(t["expected"]))
    for k in ("label_spots_2d", "label_spots_3d"):
        t = ref_tests[k]
        assert np.array_equal(cle.pull(cle.label_spots(A(t["input"]))), A(t["expected"])), k
    t = ref_tests["masked_voronoi_labeling"]
    assert np.array_equal(cle.pull(cle.masked_voronoi_labeling(A(t["input"]), A(t["mask"]))), A(t["expected"]))
    t = ref_tests["voronoi_labeling"]
    assert np.array_equal(cle.pull(cle.voronoi_labeling(A(t["input"]))), A(t["expected"]))
    t = ref_tests["extend_labeling_via_voronoi"]
    out = cle.extend_labeling_via_voronoi(A(t["input"]))
    assert out.dtype == np.uint32 and np.array_equal(cle.pull(out), A(t["expected"]))
    t = ref_tests["voronoi_otsu_labeling"]
    out = cle.voronoi_otsu_labeling(A(t["input"]), spot_sigma=t["spot_sigma"], outline_sigma=t["outline_sigma"])
    assert out.dtype == np.uint32 and np.array_equal(cle.pull(out), A(t["expected"]))
    t = ref_tests["gauss_otsu_labeling"]
    out = cle.gauss_otsu_labeling(A(t["input"]), outline_sigma=t["outline_sigma"])
    assert np.array_equal(cle.pull(out), A(t["expected"]))


def test_blur_reference_bounds(cle, ref_tests):
    t = ref_tests["gaussian_blur_delta"]
    src = cle.push(np.asarray(t["input"]))
    res = cle.create(src)
    cle.gaussian_blur(src, res, *t["sigma"])
    a = cle.pull(res)
    assert a.min() > t["min_gt"] and t["max_gt"] < a.max() < t["max_lt"]
    t = ref_tests["difference_of_gaussian"]
    d = cle.pull(cle.gaussian_blur(src, None, *t["sigma1"])) - cle.pull(cle.gaussian_blur(src, None, *t["sigma2"]))
    assert t["min_gt"] < d.min() < t["min_lt"] and t["max_gt"] < d.max() < t["max_lt"]


def test_statistics_reference_tables(cle, ref_tests):
    for key in ("statistics_2d", "statistics_3d"):
        t = ref_tests[key]
        r = cle.statistics_of_labelled_pixels(np.asarray(t["intensity"]), np.asarray(t["labels"]))
        assert list(r.keys()) == O.STAT_KEYS
        for k, v in t["expected"].items():
            assert np.allclose(r[k], v, t["rtol"]), (key, k)
    t = ref_tests["statistics_std"]
    r = cle.statistics_of_labelled_pixels(np.asarray(t["intensity"]), np.asarray(t["labels"]))
    np.testing.assert_almost_equal(r["standard_deviation_intensity"][0], t["std"], decimal=t["decimal"])
    t = ref_tests["statistics_clij_csv"]
    fx = load_npz("mini_blobs.npz")
    r = cle.statistics_of_background_and_labelled_pixels(fx["image"], fx["labels"])
    for i, k in enumerate(t["columns"]):
        assert np.allclose(fx["clij_table"][:, i], r[k], t["rtol"]), k


def test_known_answers(cle, ref_tests):
    img = load_npz("blobs.npz")["image"]
    blurred = cle.gaussian_blur(img, sigma_x=1, sigma_y=1)
    binary = cle.threshold_otsu(blurred)
    labels = cle.connected_components_labeling_box(binary)
    assert cle.maximum_of_all_pixels(labels) == 62
    t = ref_tests["bbbc022_spots"]
    crop = load_npz("bbbc022_crop.npz")["image"]
    s = t["spot_sigma"]
    spots = cle.detect_maxima_box(cle.gaussian_blur(crop, sigma_x=s, sigma_y=s, sigma_z=s))
    assert cle.sum_of_all_pixels(spots) == t["detected_spots"]
    o = t["outline_sigma"]
    binary = cle.threshold_otsu(cle.gaussian_blur(crop, sigma_x=o, sigma_y=o, sigma_z=o))
    assert cle.sum_of_all_pixels(cle.binary_and(binary, spots)) == t["selected_spots"]


# ----------------------------------------------------------------- seeded parity vs the oracle
@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (5, 1), (13, 17), (64, 200), (3, 4, 5), (9, 33, 70), (40, 70, 90)])
@pytest.mark.parametrize("sig", [0.0, 0.5, 1.0, 1.5, 2.0, 3.5, 5.0])
def test_blur_parity(cle, shape, sig):
    rng = np.random.default_rng(hash((shape, sig)) % 2**32)
    a = (rng.random(shape) * 1000).astype(np.float32)
    sz = sig * 0.7
    got = cle.pull(cle.gaussian_blur(a, None, sig, sig * 1.3, sz))
    assert_blur_close(got, O.gaussian_blur(a, sig, sig * 1.3, sz))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_blur_typed_source(cle, dtype):
    rng = np.random.default_rng(3)
    a = (rng.random((6, 40, 50)) * 200).astype(dtype)
    import torch
    if dtype == np.float32:
        t = torch.from_numpy(a).cuda()
    else:
        t = torch.from_numpy(a.astype(np.int32)).cuda().to(dtype={np.uint8: torch.uint8, np.uint16: torch.uint16}[dtype])
    src = cle.asarray(t)
    assert src.dtype == np.dtype(dtype)
    dst = cle.create(src.shape, np.float32)
    cle.gaussian_blur(src, dst, 2, 2, 2)
    assert_blur_close(cle.pull(dst), O.gaussian_blur(a.astype(np.float32), 2, 2, 2))


@pytest.mark.parametrize("shape,rad", [((7, 20, 30), (1, 2, 1)), ((5, 17, 23), (2, 1, 1)), ((4, 9, 40), (0, 3, 2)),
                                       ((33, 45), (3, 0, 1)), ((19, 28), (1, 1, 0))])
def test_mean_box_parity(cle, shape, rad):
    rng = np.random.default_rng(4)
    a = (rng.random(shape) * 100).astype(np.float32)
    got = cle.pull(cle.mean_box(a, None, *rad))
    assert_blur_close(got, O.mean_box(a, *rad))


def test_mean_box_reference(cle, ref_tests):
    t = ref_tests["mean_box"]
    src = cle.push(np.asarray(t["input"]))
    res = cle.create(src)
    cle.mean_box(src, res, *t["radius"])
    assert np.allclose(cle.pull(res), np.asarray(t["expected"]))


@pytest.mark.parametrize("key", ["dilate_labels_2d", "dilate_labels_3d"])
def test_dilate_labels_reference(cle, ref_tests, key):
    # tests/test_dilate_labels.py of the reference: radius 1 is one box step, and in 3D
    # it fixes the 27-voxel neighbourhood across z
    t = ref_tests[key]
    out = cle.extend_labels_with_maximum_radius(np.asarray(t["input"]), radius=1)
    assert out.dtype == np.uint32 and np.array_equal(cle.pull(out), np.asarray(t["expected"]))
    flag = cle.create((1, 1, 1))
    cle.set(flag, 0)
    res = cle.onlyzero_overwrite_maximum_box(np.asarray(t["input"]), flag)
    assert np.array_equal(cle.pull(res), np.asarray(t["expected"])) and cle.pull(flag).ravel()[0] == 1


@pytest.mark.parametrize("shape,radius", [((30, 40), 1), ((30, 40), 4), ((9, 20, 25), 2), ((9, 20, 25), 5)])
def test_dilate_labels_parity(cle, shape, radius):
    rng = np.random.default_rng(16)
    lab = np.zeros(shape, np.uint32)
    idx = rng.choice(lab.size, max(2, lab.size // 120), replace=False)
    lab.flat[idx] = rng.integers(1, 5000, idx.size)
    got = cle.pull(cle.dilate_labels(lab, None, radius))
    assert got.dtype == np.uint32 and np.array_equal(got, O.dilate_labels(lab, radius))


@pytest.mark.parametrize("shape", [(31, 45), (6, 17, 23), (1, 9, 9), (3, 5, 300), (1, 1), (700, 3)])
def test_maxima_parity(cle, shape):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 5, shape).astype(np.float32)  # plateaus and ties
    assert np.array_equal(cle.pull(cle.detect_maxima_box(a)), O.detect_maxima_box(a))
    b = (rng.random(shape) * 10).astype(np.float32)
    assert np.array_equal(cle.pull(cle.detect_maxima_box(b, None, 1, 1, 1)), O.detect_maxima_box(b, 1, 1, 1))


@pytest.mark.parametrize("shape", [(1, 1), (50, 60), (8, 33, 47)])
def test_histogram_otsu_parity(cle, shape):
    img = blobs_image(shape, 6, 7).astype(np.float32)
    blurred = O.gaussian_blur(img, 1, 1, 1)
    d = cle.push(blurred)
    assert cle.minimum_of_all_pixels(d) == blurred.min()
    assert cle.maximum_of_all_pixels(d) == blurred.max()
    assert np.array_equal(cle.pull(cle.histogram(d, num_bins=256)), O.histogram(blurred, 256))
    assert np.array_equal(cle.pull(cle.histogram(d, num_bins=17)), O.histogram(blurred, 17))
    out, t = O.threshold_otsu(blurred, return_threshold=True)
    assert cle.otsu_threshold(d) == t
    assert np.array_equal(cle.pull(cle.threshold_otsu(d)), out)


def test_otsu_constant_image(cle):
    a = np.full((5, 6), 3.0, np.float32)
    assert cle.otsu_threshold(a) == 3.0
    assert cle.pull(cle.threshold_otsu(a)).sum() == 0


@pytest.mark.parametrize("levels", [(0.0, 1.0), (0.0, 5.0, 10.0), (2.0, 2.5, 2.5, 9.0, 9.0, 9.0)])
def test_otsu_tied_variances(cle, levels):
    # few distinct values: runs of equal between-class variances; the first maximum wins
    rng = np.random.default_rng(31)
    a = rng.choice(np.asarray(levels, np.float32), size=(40, 50)).astype(np.float32)
    out, t = O.threshold_otsu(a, return_threshold=True)
    assert cle.otsu_threshold(a) == t
    assert np.array_equal(cle.pull(cle.threshold_otsu(a)), out)


def test_saturating_stores_at_the_limits(cle):
    # float -> integer stores at and beyond the 32-bit limits (2^31 and 2^32 are the
    # limits rounded to float32), +-inf and NaN, against the oracle's convert_sat
    v = np.array([2.0 ** 31, 2.0 ** 32, 3e9, -3e9, np.inf, -np.inf, np.nan, 4294967040.0, 2147483520.0,
                  -2147483648.0, 1.5, -1.5, 255.0, 256.0, -129.0, 65535.5, 65536.0], np.float32)
    src = cle.push(v.reshape(1, -1))
    for dt in (np.uint32, np.int32, np.uint16, np.int16, np.uint8, np.int8):
        d = cle.create(src.shape, dt)
        cle.copy(src, d)
        assert np.array_equal(cle.pull(d).ravel(), O.convert_sat(v, dt)), dt
        cle.add_image_and_scalar(src, d, 0)
        assert np.array_equal(cle.pull(d).ravel(), O.convert_sat(v, dt)), dt


def test_elementwise_dtypes(cle):
    rng = np.random.default_rng(8)
    a = (rng.random((9, 10)) * 600 - 100).astype(np.float32)
    b = (rng.random((9, 10)) > 0.5).astype(np.float32)
    for dt in (np.uint8, np.uint16, np.int32, np.uint32, np.float32):
        d = cle.create(a.shape, dt)
        cle.add_image_and_scalar(a, d, 0.5)
        assert np.array_equal(cle.pull(d), O.add_image_and_scalar(a, 0.5, dt)), dt
        cle.mask(a, b, d)
        assert np.array_equal(cle.pull(d), O.mask(a, b, dt)), dt
    assert np.array_equal(cle.pull(cle.add_images_weighted(a, b, None, 2, -3)), O.add_images_weighted(a, b, 2, -3))
    assert np.array_equal(cle.pull(cle.greater_constant(a, None, 42.5)), O.greater_constant(a, 42.5))


@pytest.mark.parametrize("shape,kind,p", [((64, 64), "box", 0.45), ((120, 97), "diamond", 0.5),
                                          ((10, 30, 31), "box", 0.3), ((12, 20, 25), "diamond", 0.45),
                                          ((1, 40, 40), "box", 0.4)])
def test_ccl_parity(cle, shape, kind, p):
    rng = np.random.default_rng(9)
    b = (rng.random(shape) < p).astype(np.uint8)
    fn = cle.connected_components_labeling_box if kind == "box" else cle.connected_components_labeling_diamond
    got = cle.pull(fn(b))
    assert np.array_equal(got, O.connected_components_labeling(b, kind))
    # into a float32 destination, as the reference's masked Voronoi uses it
    d = cle.create(b.shape, np.float32)
    fn(b, d)
    assert np.array_equal(cle.pull(d), got.astype(np.float32))
    # determinism
    assert np.array_equal(cle.pull(fn(b)), got)


def test_ccl_int32_destination(cle):
    rng = np.random.default_rng(10)
    b = (rng.random((30, 40)) < 0.4).astype(np.uint8)
    d = cle.create(b.shape, np.int32)
    cle.connected_components_labeling_box(b, d)
    assert np.array_equal(cle.pull(d), O.connected_components_labeling(b).astype(np.int32))


def _ccl_expected_fp32_keys(b):
    """Composite ID rule: scipy components, min x-major key per component, rounded to
    float32, dense rank of the rounded keys (equal keys share a label)."""
    import scipy.ndimage as ndi
    sp, n = ndi.label(b != 0, structure=np.ones((3,) * b.ndim))
    keys = O.pixel_keys(b)
    mins = np.asarray(ndi.minimum(keys, sp, index=np.arange(1, n + 1)), np.int64)
    rounded = mins.astype(np.float32).astype(np.int64)
    uniq = np.unique(rounded)
    lut = np.zeros(n + 1, np.int64)
    lut[1:] = np.searchsorted(uniq, rounded) + 1
    return lut[sp].astype(np.uint32), n, len(uniq)


def test_ccl_fp32_key_merge_above_2_24(cle):
    # 8192 x 4097 pixels: keys x*H + y + 1 reach 2^25 + 8192 > 2^24.  Two single-pixel
    # components at x = 4096, y = 1002 and 1004 have keys 2^25 + 1003 and 2^25 + 1005,
    # which both round to the float32 2^25 + 1004: the reference gives them one label.
    H, W = 8192, 4097
    b = np.zeros((H, W), np.uint8)
    b[1002, W - 1] = 1
    b[1004, W - 1] = 1
    b[10:13, 5:8] = 1
    k1, k2 = (W - 1) * H + 1002 + 1, (W - 1) * H + 1004 + 1
    assert k1 > 2 ** 24 and np.float32(k1) == np.float32(k2) and k1 != k2
    want, n, n_labels = _ccl_expected_fp32_keys(b)
    assert n == 3 and n_labels == 2
    got = cle.pull(cle.connected_components_labeling_box(b))
    assert np.array_equal(got, want)
    cle.set_fp32_key_compat(False)
    try:
        exact = cle.pull(cle.connected_components_labeling_box(b))
    finally:
        cle.set_fp32_key_compat(True)
    assert exact.max() == 3


def test_relabel_and_spots_parity(cle):
    rng = np.random.default_rng(11)
    lab = rng.choice([0, 3, 7, 8, 20, 21, 1000], size=(15, 16, 17)).astype(np.float32)
    assert np.array_equal(cle.pull(cle.relabel_sequential(lab)), O.relabel_sequential(lab))
    spots = (rng.random((6, 13, 14)) < 0.05).astype(np.float32)
    assert np.array_equal(cle.pull(cle.label_spots(spots)), O.label_spots(spots))


@pytest.mark.parametrize("shape", [(50, 70), (12, 40, 45), (1, 40, 45), (9, 13, 150), (130, 70), (5, 3, 200)])
def test_masked_voronoi_parity(cle, shape):
    rng = np.random.default_rng(12)
    seeds = (rng.random(shape) < 0.01).astype(np.uint8)
    m = (rng.random(shape) < 0.8).astype(np.uint8)
    got = cle.pull(cle.masked_voronoi_labeling(seeds, m))
    assert np.array_equal(got, O.masked_voronoi_labeling(seeds, m))


@pytest.mark.parametrize("shape", [(50, 70), (12, 40, 45), (1, 40, 45), (9, 13, 150), (130, 70), (5, 3, 200)])
def test_extend_labeling_via_voronoi_parity(cle, shape):
    # sparse labels with gaps and large values, float32 destination and an int32 device source
    import torch
    rng = np.random.default_rng(14)
    lab = np.zeros(shape, np.uint32)
    idx = rng.choice(lab.size, max(2, lab.size // 150), replace=False)
    lab.flat[idx] = rng.integers(1, 100000, idx.size)
    got = cle.pull(cle.extend_labeling_via_voronoi(lab))
    assert got.dtype == np.uint32 and np.array_equal(got, O.extend_labeling_via_voronoi(lab))
    dst = cle.create(shape, np.float32)
    cle.extend_labeling_via_voronoi(cle.asarray(torch.from_numpy(lab.astype(np.int32)).cuda()), dst)
    assert np.array_equal(cle.pull(dst), O.extend_labeling_via_voronoi(lab, np.float32))


@pytest.mark.parametrize("shape", [(60, 50), (10, 30, 40)])
def test_voronoi_labeling_parity(cle, shape):
    rng = np.random.default_rng(15)
    b = (rng.random(shape) < 0.02).astype(np.uint8)
    assert np.array_equal(cle.pull(cle.voronoi_labeling(b)), O.voronoi_labeling(b))


# 3D images whose rows hold a multiple of 4 voxels take the growth that stores no state
# for tiles without a segmentation voxel (partial 16 x 8 x 8 tiles on every axis here)
@pytest.mark.parametrize("shape,ss,so", [((128, 128), 3.5, 1.0), ((200, 180), 2.0, 1.5), ((24, 80, 90), 2.0, 1.0),
                                         ((9, 64, 64), 1.0, 0.5), ((20, 44, 100), 2.0, 1.0),
                                         ((30, 72, 132), 2.0, 1.5)])
def test_voronoi_otsu_parity(cle, shape, ss, so):
    img = blobs_image(shape, 15 if len(shape) == 2 else 20, 13)
    got = cle.pull(cle.voronoi_otsu_labeling(img, spot_sigma=ss, outline_sigma=so))
    assert np.array_equal(got, O.voronoi_otsu_labeling(img, ss, so))


@pytest.mark.parametrize("shape,ss,so", [((5, 40, 48), 1.0, 1.0), ((7, 24, 36), 1.5, 1.0), ((17, 33, 52), 2.0, 0.5)])
def test_voronoi_otsu_stages_on_device_blurs(cle, shape, ss, so):
    # one partial growth tile in z (and partial tiles in y and x): the pipeline equals
    # the oracle's stages fed with the device's two blurs bit for bit (the full oracle
    # can differ here by a maxima near-tie of the blur's rounding)
    img = blobs_image(shape, 20, 13)
    got = cle.pull(cle.voronoi_otsu_labeling(img, spot_sigma=ss, outline_sigma=so))
    bs = cle.pull(cle.gaussian_blur(img, None, ss, ss, ss)).astype(np.float32)
    bo = cle.pull(cle.gaussian_blur(img, None, so, so, so)).astype(np.float32)
    want = _pipeline_from_blurs(bs, bo)[0]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(8, 32, 32), (12, 40, 64)])
def test_voronoi_otsu_no_segmentation(cle, shape):
    # a constant image: an empty segmentation, every growth tile empty, all labels 0
    img = np.full(shape, 7, np.uint8)
    got = cle.pull(cle.voronoi_otsu_labeling(img, spot_sigma=2.0, outline_sigma=1.0))
    want = O.voronoi_otsu_labeling(img, 2.0, 1.0)
    assert got.dtype == want.dtype and np.array_equal(got, want) and not got.any()


@pytest.mark.parametrize("shape", [(37, 53), (7, 31, 45), (7, 31, 48)])
def test_voronoi_otsu_destination_types(cle, shape):
    # uint32 / int32 destinations, aligned (four labels per store, then the n % 4 tail)
    # and offset by one element (element by element)
    import torch
    img = blobs_image(shape, 8, 21)
    want = O.voronoi_otsu_labeling(img, 2.0, 1.0)
    n = int(np.prod(shape))
    for tdt, ndt in ((torch.int32, np.int32), (torch.uint32, np.uint32)):
        for shift in (0, 1):
            buf = torch.zeros(n + 1, dtype=tdt, device="cuda")
            dst = cle.asarray(buf[shift:shift + n].view(shape))
            cle.voronoi_otsu_labeling(img, dst, spot_sigma=2.0, outline_sigma=1.0)
            got = cle.pull(dst)
            assert got.dtype == ndt and np.array_equal(got.astype(np.int64), want.astype(np.int64)), (ndt, shift)


def _pipeline_from_blurs(blur_s, blur_o):
    """The oracle's voronoi_otsu stages downstream of the two blurs."""
    spots = O.detect_maxima_box(blur_s, 0, 0, 0, np.float32)
    seg = O.threshold_otsu(blur_o, np.float32)
    binary = O.binary_and(spots, seg, np.uint8)
    return O.mask(O.masked_voronoi_labeling(binary, seg, np.float32), seg, np.uint32), spots, seg


def _near_tie_maxima(blur, a, b, tol):
    """Voxels where two maxima maps differ must have a 3^d neighbour within tol."""
    diff = np.argwhere(a != b)
    b3 = blur if blur.ndim == 3 else blur[None]
    for p in diff:
        p3 = tuple(p) if blur.ndim == 3 else (0,) + tuple(p)
        sl = tuple(slice(max(0, c - 1), c + 2) for c in p3)
        nb = b3[sl].ravel()
        gaps = np.abs(nb - b3[p3])
        if not np.any((gaps > 0) & (gaps <= tol)) and not np.sum(gaps == 0) > 1:
            return False
    return True


@pytest.mark.parametrize("name,ss,so", [("blobs.npz", 2.0, 1.0), ("bbbc022_crop.npz", 5.0, 1.0),
                                        ("blobs.npz", 1.0, 1.0)])
def test_voronoi_otsu_fixtures(cle, name, ss, so):
    """Real images with plateaus.  (1) The device pipeline equals, bit for bit, the
    oracle's stages fed with the device's own blurred images.  (2) Against the full
    oracle, the blurs agree to BLUR_TOL and every differing maxima / segmentation
    voxel is a near tie (within BLUR_TOL of a neighbour or of the Otsu threshold);
    the number of differing label voxels is reported."""
    img = load_npz(name)["image"]
    got = cle.pull(cle.voronoi_otsu_labeling(img, spot_sigma=ss, outline_sigma=so))
    gbs = cle.pull(cle.gaussian_blur(img, None, ss, ss, ss))
    gbo = cle.pull(cle.gaussian_blur(img, None, so, so, so))
    staged, g_spots, g_seg = _pipeline_from_blurs(gbs, gbo)
    assert np.array_equal(got, staged)
    bs = O.gaussian_blur(img, ss, ss, ss)
    bo = O.gaussian_blur(img, so, so, so)
    assert_blur_close(gbs, bs)
    assert_blur_close(gbo, bo)
    want, o_spots, o_seg = _pipeline_from_blurs(bs, bo)
    tol = BLUR_TOL * max(1.0, float(np.abs(bs).max()))
    assert _near_tie_maxima(bs, g_spots, o_spots, tol)
    _, t_o = O.threshold_otsu(bo, return_threshold=True)
    _, t_g = O.threshold_otsu(gbo, return_threshold=True)
    tol_o = BLUR_TOL * max(1.0, float(np.abs(bo).max()))
    lo, hi = min(t_o, t_g) - tol_o, max(t_o, t_g) + tol_o
    seg_diff = g_seg != o_seg
    assert np.all((bo[seg_diff] >= lo) & (bo[seg_diff] <= hi))
    n_diff = int(np.sum(got != want))
    print(f"{name}: {n_diff} label voxels differ from the full oracle "
          f"({int(np.sum(g_spots != o_spots))} near-tie maxima, {int(seg_diff.sum())} near-threshold voxels, "
          f"thresholds {t_g} / {t_o})")


def _assert_stats_close(got, want):
    assert list(got.keys()) == list(want.keys())
    for k in want:
        g, w = np.asarray(got[k], np.float64), np.asarray(want[k], np.float64)
        assert g.shape == w.shape, k
        np.testing.assert_allclose(g, w, rtol=STATS_RTOL, atol=0, equal_nan=True, err_msg=k)


@pytest.mark.parametrize("shape", [(60, 70), (8, 30, 40)])
def test_statistics_parity(cle, shape):
    img = blobs_image(shape, 8, 14).astype(np.float32)
    labels = O.voronoi_otsu_labeling(img, 2.0, 1.0)
    _assert_stats_close(cle.statistics_of_labelled_pixels(img, labels), O.statistics_of_labelled_pixels(img, labels))


def test_statistics_label_gaps(cle):
    lab = np.zeros((10, 12), np.uint32)
    lab[2:4, 3:6] = 2
    lab[7, 1:9] = 5
    img = np.arange(lab.size, dtype=np.float32).reshape(lab.shape)
    _assert_stats_close(cle.statistics_of_labelled_pixels(img, lab), O.statistics_of_labelled_pixels(img, lab))


# ----------------------------------------------------------------- tiled / vectorised paths
@pytest.mark.parametrize("shape,kind,p", [((9, 19, 150), "box", 0.3), ((9, 19, 150), "diamond", 0.5),
                                          ((70, 150), "box", 0.45), ((70, 150), "diamond", 0.55),
                                          ((13, 33, 65), "box", 0.6), ((5, 17, 129), "box", 0.2)])
def test_ccl_multi_tile(cle, shape, kind, p):
    """Shapes spanning several 4x8x64 (3D) / 32x64 (2D) tiles in every direction."""
    rng = np.random.default_rng(21)
    b = (rng.random(shape) < p).astype(np.uint8)
    fn = cle.connected_components_labeling_box if kind == "box" else cle.connected_components_labeling_diamond
    assert np.array_equal(cle.pull(fn(b)), O.connected_components_labeling(b, kind))


@pytest.mark.parametrize("shape,p", [((16, 48, 128), 0.25), ((9, 34, 68), 0.35), ((130, 256), 0.4),
                                     ((7, 19, 44), 0.3), ((2, 64, 64), 0.5)])
@pytest.mark.parametrize("dst_dtype", [np.uint32, np.int32, np.float32, np.uint16])
def test_ccl_box_uint8_device_input(cle, shape, p, dst_dtype):
    """uint8 binaries kept on the device (the 4-voxel gather for 32-bit destinations
    when rows hold a multiple of 4 voxels), across 2x2x2-block tile borders."""
    rng = np.random.default_rng(33)
    b = (rng.random(shape) < p).astype(np.uint8)
    want = O.connected_components_labeling(b, "box")
    d = cle.create(b.shape, dst_dtype)
    cle.connected_components_labeling_box(cle.asarray(b), d)
    assert np.array_equal(cle.pull(d), want.astype(dst_dtype))


def test_ccl_box_blocks_side_by_side_not_touching(cle):
    """Neighbouring 2x2x2 blocks whose foreground voxels are two apart stay separate."""
    b = np.zeros((6, 8, 12), np.uint8)
    b[0, 0, 1] = 1; b[0, 0, 4] = 1          # x = 1 and x = 4: blocks 0 and 2 (gap block 1)
    b[2, 1, 2] = 1; b[2, 1, 5] = 1          # blocks 1 and 2 along x, 3 apart
    b[3, 2, 3] = 1; b[3, 4, 3] = 1          # y = 2 and 4: gap of one row
    b[5, 6, 8] = 1; b[4, 7, 9] = 1          # diagonal across block corner: touching
    b[1, 5, 10] = 1; b[3, 6, 11] = 1        # z 1 and 3: not touching
    want = O.connected_components_labeling(b, "box")
    assert np.array_equal(cle.pull(cle.connected_components_labeling_box(cle.asarray(b))), want)
    assert np.array_equal(cle.pull(cle.connected_components_labeling_box(b)), want)


def test_ccl_diagonal_tile_crossings(cle):
    """Diagonal chains crossing tile corners (+x/-y, +x/-z, +y/-z neighbours)."""
    b = np.zeros((12, 24, 140), np.uint8)
    for t in range(12):
        b[t, 20 - t, 60 + t] = 1          # +x, -y per slice step
        b[t, 3 + t, 130 - t] = 1          # -x, +y
    for t in range(9):
        b[11 - t, 8 + t, 63 + t] = 1      # -z, +y, +x
    for kind in ("box", "diamond"):
        fn = cle.connected_components_labeling_box if kind == "box" else cle.connected_components_labeling_diamond
        assert np.array_equal(cle.pull(fn(b)), O.connected_components_labeling(b, kind)), kind


@pytest.mark.parametrize("shape,sig", [((96, 40, 70), (2.0, 2.0, 2.0)), ((90, 33, 65), (1.5, 1.0, 3.0)),
                                       ((50, 66, 130), (3.5, 3.5, 3.5))])
def test_blur_fused_z_chunks(cle, shape, sig):
    rng = np.random.default_rng(22)
    a = rng.integers(0, 4000, shape).astype(np.uint16)
    got = cle.pull(cle.gaussian_blur(a, None, *sig))
    assert_blur_close(got, O.gaussian_blur(a.astype(np.float32), *sig))


def test_histogram_many_bins(cle):
    rng = np.random.default_rng(23)
    a = (rng.random((20, 30)) * 1000).astype(np.float32)
    assert np.array_equal(cle.pull(cle.histogram(a, num_bins=1000)), O.histogram(a, 1000))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int32, np.uint32, np.float32])
def test_minmax_vector_tails(cle, dtype):
    rng = np.random.default_rng(24)
    for shape in ((7, 13, 17), (1, 3), (1030,)):
        a = rng.integers(1, 250, shape).astype(dtype)
        a.flat[-1] = 251
        a.flat[0] = 0
        d = cle.asarray(a)
        assert d.dtype == np.dtype(dtype)
        assert cle.minimum_of_all_pixels(d) == 0 and cle.maximum_of_all_pixels(d) == 251, (dtype, shape)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32, np.uint16, np.uint8])
def test_statistics_label_dtypes_and_spans(cle, dtype):
    rng = np.random.default_rng(25)
    lab = np.zeros((3, 5, 1500), np.int64)
    lab[0, 1, 10:1400] = 1
    lab[1, 2:4, 1000:1100] = 2
    lab[2, 4, 1499] = 3
    lab[:, :, 200:203] = 4
    img = (rng.random(lab.shape) * 100).astype(np.float32)
    d = cle.asarray(lab.astype(dtype))
    assert d.dtype == np.dtype(dtype)
    got = cle.statistics_of_labelled_pixels(img, d)
    _assert_stats_close(got, O.statistics_of_labelled_pixels(img, lab))


def _assert_stats_identical(got, want):
    for k in want:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        same = (g == w) | (np.isnan(g.astype(np.float64)) & np.isnan(w.astype(np.float64)))
        assert np.all(same), (k, g[~same][:5], w[~same][:5])


@pytest.mark.parametrize("shape", [(60, 70), (8, 30, 40), (3, 5, 1500)])
def test_statistics_bit_exact(cle, shape):
    """The device statistics follow the reference's float32 accumulation order (cells
    per (row y, label) in (z, x) order, then y order), so they equal the oracle's."""
    rng = np.random.default_rng(26)
    if shape == (3, 5, 1500):
        lab = np.zeros(shape, np.uint32)
        lab[0, 1, 10:1400] = 1
        lab[1, 2:4, 1000:1100] = 2
        lab[:, :, 200:203] = 4
    else:
        img0 = blobs_image(shape, 8, 27).astype(np.float32)
        lab = O.voronoi_otsu_labeling(img0, 2.0, 1.0)
    img = (rng.random(shape) * 1000).astype(np.float32)
    _assert_stats_identical(cle.statistics_of_labelled_pixels(img, lab), O.statistics_of_labelled_pixels(img, lab))


def test_ccl_large_vs_composite_rule(cle):
    """4M voxels at a percolating density: labels equal the rank of each component's
    minimum x-major key (scipy components)."""
    rng = np.random.default_rng(28)
    b = (rng.random((64, 256, 256)) < 0.12).astype(np.uint8)
    want, n, _ = _ccl_expected_fp32_keys(b)
    got = cle.pull(cle.connected_components_labeling_box(b))
    assert got.max() == n
    assert np.array_equal(got, want)


def test_statistics_label_count_derived_and_table_growth(cle):
    """num_labels derived in the run pass equals max(labels) + 1; 5000 labels exceed
    the default 4096-column table, which is grown and the call repeated; the explicit
    count gives the same table."""
    rng = np.random.default_rng(40)
    lab = np.zeros((100, 120), np.uint32)
    idx = rng.choice(lab.size, 5000, replace=False)
    lab.flat[idx] = np.arange(1, 5001, dtype=np.uint32)
    img = (rng.random(lab.shape) * 100).astype(np.float32)
    got = cle.statistics_of_labelled_pixels(img, lab)
    assert len(got["label"]) == 5000
    _assert_stats_identical(got, O.statistics_of_labelled_pixels(img, lab))
    t_explicit, n = cle.statistics_table(cle.push(img), cle.push(lab), 5001)
    t_derived, n2 = cle.statistics_table(cle.push(img), cle.push(lab))
    assert n == n2 == 5000
    assert np.array_equal(t_explicit.tensor.cpu().numpy(), t_derived.tensor[:, :n].cpu().numpy(), equal_nan=True)
    # no labels at all
    empty = cle.statistics_of_labelled_pixels(img, np.zeros_like(lab))
    assert all(len(v) == 0 for v in empty.values())


def test_statistics_run_buffer_overflow_rerun(cle):
    """A label image whose every pixel is its own run (16.7M runs) overflows the run
    buffer sized from the previous call; the pass is rerun at the exact size and the
    result equals the next call's (sized from the hint) and numpy's counts / bounds."""
    H, W = 4096, 4096
    small = np.ones((8, 8), np.uint32)
    cle.statistics_of_labelled_pixels(np.ones((8, 8), np.float32), small)  # small run-count hint
    x = np.arange(W, dtype=np.uint32)
    lab = np.broadcast_to((x % 7 + 1)[None, :], (H, W)).copy()
    lab[::2] = np.roll(lab[::2], 3, axis=1)
    img = (np.random.default_rng(41).random((H, W)) * 50).astype(np.float32)
    g_img, g_lab = cle.push(img), cle.asarray(lab)
    first = cle.statistics_of_labelled_pixels(g_img, g_lab)
    second = cle.statistics_of_labelled_pixels(g_img, g_lab)
    _assert_stats_identical(first, second)
    counts = np.bincount(lab.ravel(), minlength=8)[1:]
    assert np.array_equal(first["area"], counts.astype(np.float32))
    for k in range(7):
        ys, xs = np.nonzero(lab == k + 1)
        assert first["bbox_min_x"][k] == xs.min() and first["bbox_max_x"][k] == xs.max()
        assert first["bbox_min_y"][k] == ys.min() and first["bbox_max_y"][k] == ys.max()
        s = img[lab == k + 1].astype(np.float64).sum()
        assert abs(first["sum_intensity"][k] - s) <= 1e-3 * s


@pytest.mark.parametrize("shape", [(8, 1100, 64), (3, 40, 256)])
def test_histogram_bin_boundaries(cle, shape):
    """Values on and next to the bin boundaries (bin = floor(q + 0.5) of the float32
    quotient q = (v - min) * (bins - 1) / range): the device's reciprocal shortcut must
    hand every near-boundary value to the exact rule.  Whole-row blocks (1100 rows)
    and z-chunked blocks (40 rows) both run."""
    rng = np.random.default_rng(42)
    a = (rng.random(shape) * 255).astype(np.float32)
    flat = a.reshape(-1)
    k = rng.integers(0, 255, flat.size // 2)
    eps = rng.choice(np.array([0, 1, -1, 2, -2, 3, -3], np.float32), flat.size // 2)
    half = (k + 0.5).astype(np.float32)
    flat[: flat.size // 2] = half + eps * np.spacing(half)  # k + 0.5 and up to 3 ulp either side
    flat[0], flat[1] = 0.0, 255.0
    for bins in (256, 17, 100):
        assert np.array_equal(cle.pull(cle.histogram(a, num_bins=bins)), O.histogram(a, bins)), bins
    # a range that is not a whole number, and one whose reciprocal is not exact
    b = (a * np.float32(0.7313) + np.float32(3.1)).astype(np.float32)
    assert np.array_equal(cle.pull(cle.histogram(b, num_bins=256)), O.histogram(b, 256))
    # tiny range (subnormal): only the exact rule applies
    c = np.zeros((4, 50, 64), np.float32)
    c.flat[::3] = np.float32(1e-39)
    assert np.array_equal(cle.pull(cle.histogram(c, num_bins=256)), O.histogram(c, 256))


@pytest.mark.parametrize("shape,sig", [((64, 256, 256), 2.0), ((64, 256, 256), 1.5), ((1024, 1024), 2.6)])
def test_blur_autotuned_chunking_identical(cle, shape, sig):
    # the autotuner only changes how the volume is split into blocks and the register
    # budget the tile code runs under: the tuned and the default launch give
    # bit-identical output
    import torch
    rng = np.random.default_rng(21)
    vol = rng.integers(0, 4000, shape).astype(np.uint16)
    src = cle.asarray(torch.from_numpy(vol.view(np.int16)).cuda().view(torch.uint16))
    cle.set_autotune(False)
    outs = [cle.pull(cle.gaussian_blur(src, sigma_x=sig, sigma_y=sig, sigma_z=sig))]  # defaults
    cle.set_autotune(True)
    outs.append(cle.pull(cle.autotune_gaussian_blur(src, None, sig, sig, sig)))      # tuning run
    outs.append(cle.pull(cle.gaussian_blur(src, sigma_x=sig, sigma_y=sig, sigma_z=sig)))  # tuned choice
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert_blur_close(outs[0], O.gaussian_blur(vol.astype(np.float32), sig, sig, sig))


@pytest.mark.parametrize("c", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 20])
def test_blur_isotropic_half_widths(cle, c):
    # isotropic blurs of every register budget the half-width selects by default (3D
    # and 2D, u16 and f32 sources, tiles cut by the image edges) against the oracle
    sig = (2 * c + 0.5) / 8  # floor(8 sigma) = 2c, made odd: 2c + 1 taps
    rng = np.random.default_rng(100 + c)
    for shape in ((37, 70, 130), (150, 260)):
        a = rng.integers(0, 4000, shape).astype(np.uint16)
        want = O.gaussian_blur(a.astype(np.float32), sig, sig, sig)
        assert_blur_close(cle.pull(cle.gaussian_blur(a, None, sig, sig, sig)), want)
        assert_blur_close(cle.pull(cle.gaussian_blur(a.astype(np.float32), None, sig, sig, sig)), want)


@pytest.mark.parametrize("shape", [(70, 130), (20, 37, 70), (1, 18, 66), (17, 16, 64), (70, 128), (40, 100), (1, 4),
                                   (20, 37, 68), (5, 30, 4), (33, 29, 132)])
def test_maxima_special_values(cle, shape):
    # the closed-form visit rule against the oracle's literal visit order: plateaus
    # (ties), NaN (never takes over; a NaN centre stays a maximum), +-inf, and values
    # above 2^24 where v - 1 == v
    rng = np.random.default_rng(31)
    a = rng.integers(0, 4, shape).astype(np.float32)
    special = np.array([np.nan, np.inf, -np.inf, 3e8, 3e8 + 32, -3e8], np.float32)
    m = rng.random(shape) < 0.08
    a[m] = rng.choice(special, int(m.sum()))
    got = cle.pull(cle.detect_maxima_box(a))
    assert np.array_equal(got, O.detect_maxima_box(a))
