"""The cle functions of the segmentation hot path, drop-in for pyclesperanto_prototype.

Each function keeps the reference's name, signature, defaults, output dtype and
return value, and runs on the HIP kernels of libcle_hip.so.  Reference locations are
given per function (paths under pyclesperanto_prototype/).
"""
import ctypes
import warnings

import numpy as np

from . import _lib
from ._array import (Array, create, create_binary_like, create_labels_like, create_like, create_none, pull,
                     push, stream_handle, torch)
from ._plugin import Image, plugin_function

_S = stream_handle


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ fills / copies
@plugin_function
def set(source: Image, scalar: float = 0) -> Image:
    """_tier1/_set.py:5-33."""
    _lib.call("cle_set", source.buf(), float(scalar), _S())
    return source


@plugin_function
def copy(source: Image, destination: Image = None) -> Image:
    """_tier1/_copy.py:5-37."""
    _lib.call("cle_copy", source.buf(), destination.buf(), _S())
    return destination


# ------------------------------------------------------------------ filters
@plugin_function(categories=['filter', 'denoise', 'in assistant', 'bia-bob-suggestion'], priority=1)
def gaussian_blur(source: Image, destination: Image = None, sigma_x: float = 0, sigma_y: float = 0,
                  sigma_z: float = 0) -> Image:
    """_tier1/_gaussian_blur.py:6-58."""
    _lib.call("cle_gaussian_blur", source.buf(), destination.buf(), float(sigma_x), float(sigma_y), float(sigma_z),
              _S())
    return destination


@plugin_function(categories=['filter', 'denoise', 'in assistant'])
def mean_box(source: Image, destination: Image = None, radius_x: float = 1, radius_y: float = 1,
             radius_z: float = 1) -> Image:
    """_tier1/_mean_box.py:6-54."""
    _lib.call("cle_mean_box", source.buf(), destination.buf(), int(radius_x), int(radius_y), int(radius_z), _S())
    return destination


@plugin_function(categories=['binarize', 'in assistant'], output_creator=create_binary_like)
def detect_maxima_box(source: Image, destination: Image = None, radius_x: int = 0, radius_y: int = 0,
                      radius_z: int = 0) -> Image:
    """_tier1/_detect_maxima_box.py:5-48."""
    _lib.call("cle_detect_maxima_box", source.buf(), destination.buf(), int(radius_x), int(radius_y),
              int(radius_z), _S())
    return destination


# ------------------------------------------------------------------ reductions
def _minmax_dev(source: Array):
    res = torch.empty(2, dtype=torch.float32, device=source.tensor.device)
    _lib.call("cle_minimum_maximum", source.buf(), _ptr(res), _S())
    return res


@plugin_function
def minimum_of_all_pixels(source: Image) -> float:
    """_tier2/_minimum_of_all_pixels.py:6-59."""
    return np.float32(_minmax_dev(source)[0].item())


@plugin_function
def maximum_of_all_pixels(source: Image) -> float:
    """_tier2/_maximum_of_all_pixels.py:6-59."""
    return np.float32(_minmax_dev(source)[1].item())


@plugin_function
def sum_of_all_pixels(source: Image) -> float:
    """_tier2/_sum_of_all_pixels.py (float64 accumulation)."""
    return source.sum()


@plugin_function(output_creator=create_none)
def histogram(source: Image, destination: Image = None, num_bins: int = 256, minimum_intensity: float = None,
              maximum_intensity: float = None, determine_min_max: bool = True) -> Image:
    """_tier3/_histogram.py:25-130; returns a float32 vector of num_bins."""
    if destination is None:
        destination = create([int(num_bins)])
    if minimum_intensity is None or maximum_intensity is None or determine_min_max:
        mm = _minmax_dev(source)
        _lib.call("cle_histogram", source.buf(), destination.buf(), int(num_bins), 0.0, 0.0, _ptr(mm), _S())
    else:
        _lib.call("cle_histogram", source.buf(), destination.buf(), int(num_bins), float(minimum_intensity),
                  float(maximum_intensity), None, _S())
    return destination


@plugin_function(categories=['binarize', 'in assistant', 'bia-bob-suggestion'], priority=1,
                 output_creator=create_binary_like)
def threshold_otsu(source: Image, destination: Image = None) -> Image:
    """_tier9/_threshold_otsu.py:11-56 (histogram, Otsu search and threshold on the device)."""
    _lib.call("cle_threshold_otsu", source.buf(), destination.buf(), None, _S())
    return destination


def otsu_threshold(source) -> float:
    """The float64 threshold threshold_otsu applies (source > float32(threshold))."""
    source = push(source)
    out = create(source.shape, np.uint8)
    thr = torch.empty(1, dtype=torch.float64, device=source.tensor.device)
    _lib.call("cle_threshold_otsu", source.buf(), out.buf(), _ptr(thr), _S())
    return float(thr.item())


# ------------------------------------------------------------------ elementwise
@plugin_function(categories=['binarize', 'in assistant'], output_creator=create_binary_like)
def greater_constant(source: Image, destination: Image = None, constant: float = 0) -> Image:
    """_tier1/_greater_constant.py:5-40."""
    _lib.call("cle_greater_constant", source.buf(), destination.buf(), float(constant), _S())
    return destination


@plugin_function(categories=['combine', 'binary processing', 'in assistant', 'combine labels', 'label processing'],
                 priority=1, output_creator=create_binary_like)
def binary_and(operand1: Image, operand2: Image, destination: Image = None) -> Image:
    """_tier1/_binary_and.py:6-47."""
    _lib.call("cle_binary_and", operand1.buf(), operand2.buf(), destination.buf(), _S())
    return destination


@plugin_function(categories=[])
def mask(source: Image, mask: Image, destination: Image = None) -> Image:
    """_tier1/_mask.py:5-45."""
    _lib.call("cle_mask", source.buf(), mask.buf(), destination.buf(), _S())
    return destination


@plugin_function(categories=['filter', 'in assistant'])
def add_image_and_scalar(source: Image, destination: Image = None, scalar: float = 1) -> Image:
    """_tier1/_add_image_and_scalar.py:5-43."""
    _lib.call("cle_add_image_and_scalar", source.buf(), destination.buf(), float(scalar), _S())
    return destination


@plugin_function(categories=['combine', 'in assistant'])
def add_images_weighted(summand1: Image, summand2: Image, destination: Image = None, factor1: float = 1,
                        factor2: float = 1) -> Image:
    """_tier1/_add_images_weighted.py:5-50."""
    _lib.call("cle_add_images_weighted", summand1.buf(), summand2.buf(), destination.buf(), float(factor1),
              float(factor2), _S())
    return destination


@plugin_function(categories=['combine', 'in assistant'], priority=-1)
def add_images(summand1: Image, summand2: Image, destination: Image = None) -> Image:
    """_tier2/_add_images.py:5-29."""
    return add_images_weighted(summand1, summand2, destination, 1, 1)


# ------------------------------------------------------------------ labelling
@plugin_function
def set_non_zero_pixels_to_pixel_index(source: Image, output: Image = None, offset: float = 1) -> Image:
    """_tier1/_set_non_zero_pixels_to_pixel_index.py:5-28."""
    _lib.call("cle_set_nonzero_pixels_to_pixel_index", source.buf(), output.buf(), int(offset), _S())
    return output


@plugin_function
def nonzero_minimum_box(source: Image, flag_dst: Image, destination: Image = None) -> Image:
    """_tier1/_nonzero_minimum_box.py:5-43."""
    _lib.call("cle_nonzero_minimum_box", source.buf(), flag_dst.buf(), destination.buf(), _S())
    return destination


@plugin_function
def nonzero_minimum_diamond(source: Image, flag_dst: Image, destination: Image = None) -> Image:
    """_tier1/_nonzero_minimum_diamond.py."""
    _lib.call("cle_nonzero_minimum_diamond", source.buf(), flag_dst.buf(), destination.buf(), _S())
    return destination


@plugin_function
def onlyzero_overwrite_maximum_box(source: Image, flag_dst: Image, destination: Image = None) -> Image:
    """_tier1/_onlyzero_overwrite_maximum_box.py:5-38."""
    _lib.call("cle_onlyzero_overwrite_maximum_box", source.buf(), flag_dst.buf(), destination.buf(), _S())
    return destination


@plugin_function
def onlyzero_overwrite_maximum_diamond(source: Image, flag_dst: Image, destination: Image = None) -> Image:
    """_tier1/_onlyzero_overwrite_maximum_diamond.py:5-38."""
    _lib.call("cle_onlyzero_overwrite_maximum_diamond", source.buf(), flag_dst.buf(), destination.buf(), _S())
    return destination


_FP32_KEY_COMPAT = True


def set_fp32_key_compat(enabled: bool) -> None:
    """Labels follow the float32-rounded minimum pixel key (the reference's behaviour
    above 2^24 voxels) when enabled (default); exact keys otherwise."""
    global _FP32_KEY_COMPAT
    _FP32_KEY_COMPAT = bool(enabled)


def _ccl(binary_input: Array, labeling_destination: Array, connectivity: int) -> Array:
    opts = _lib.CLE_FP32_KEY_COMPAT if _FP32_KEY_COMPAT else 0
    _lib.call("cle_connected_components_labeling", binary_input.buf(), labeling_destination.buf(), connectivity,
              opts, None, _S())
    return labeling_destination


@plugin_function(categories=['label', 'in assistant', 'bia-bob-suggestion'], priority=1,
                 output_creator=create_labels_like)
def connected_components_labeling_box(binary_input: Image, labeling_destination: Image = None,
                                      flagged_nonzero_minimum_filter: callable = nonzero_minimum_box) -> Image:
    """_tier4/_connected_components_labeling_box.py:15-66 (union-find; the filter
    argument selects box or diamond connectivity as in the reference)."""
    conn = _lib.CLE_CONN_DIAMOND if flagged_nonzero_minimum_filter is nonzero_minimum_diamond else _lib.CLE_CONN_BOX
    return _ccl(binary_input, labeling_destination, conn)


@plugin_function(categories=['label', 'in assistant'], output_creator=create_labels_like)
def connected_components_labeling_diamond(binary_input: Image, labeling_destination: Image = None) -> Image:
    """_tier5/_connected_components_labeling_diamond.py:8-34."""
    return _ccl(binary_input, labeling_destination, _lib.CLE_CONN_DIAMOND)


@plugin_function(output_creator=create_labels_like,
                 categories=['label processing', 'in assistant', 'bia-bob-suggestion'])
def relabel_sequential(source: Image, output: Image = None, blocksize: int = 4096) -> Image:
    """_tier3/_relabel_sequential.py:12-59 (blocksize accepted for compatibility)."""
    _lib.call("cle_relabel_sequential", source.buf(), output.buf(), _S())
    return output


@plugin_function(categories=['label', 'in assistant'], output_creator=create_labels_like)
def label_spots(input_spots: Image, labelled_spots_destination: Image = None) -> Image:
    """_tier2/_label_spots.py:9-56."""
    _lib.call("cle_label_spots", input_spots.buf(), labelled_spots_destination.buf(), _S())
    return labelled_spots_destination


@plugin_function(categories=['label'], output_creator=create_labels_like)
def masked_voronoi_labeling(binary_source: Image, mask_image: Image, labeling_destination: Image = None) -> Image:
    """_tier5/_masked_voronoi_labeling.py:7-76."""
    _lib.call("cle_masked_voronoi_labeling", binary_source.buf(), mask_image.buf(), labeling_destination.buf(),
              None, _S())
    return labeling_destination


@plugin_function(categories=['label processing', 'in assistant'], output_creator=create_labels_like)
def extend_labeling_via_voronoi(labeling_source: Image, labeling_destination: Image = None) -> Image:
    """_tier4/_extend_labeling_via_voronoi.py:12-61: the labels grown into the zero
    voxels (alternating box / diamond steps, float32 buffers) until they touch."""
    _lib.call("cle_extend_labeling_via_voronoi", labeling_source.buf(), labeling_destination.buf(), None, _S())
    return labeling_destination


@plugin_function(categories=['label processing', 'in assistant', 'bia-bob-suggestion'],
                 output_creator=create_labels_like)
def dilate_labels(labeling_source: Image, labeling_destination: Image = None, radius: int = 2) -> Image:
    """_tier4/_dilate_labels.py:12-61: at most `radius` alternating box / diamond
    growth steps on float32 buffers, stopping after a step that changes nothing."""
    flip = create_like(labeling_destination)
    flop = create_like(labeling_destination)
    flag = create((1, 1, 1))
    copy(labeling_source, flip)
    it = 0
    changed = True
    while changed and it < radius:
        set(flag, 0)
        if it % 2 == 0:
            onlyzero_overwrite_maximum_box(flip, flag, flop)
        else:
            onlyzero_overwrite_maximum_diamond(flop, flag, flip)
        changed = float(pull(flag).ravel()[0]) > 0
        it += 1
    copy(flip if it % 2 == 0 else flop, labeling_destination)
    return labeling_destination


extend_labels_with_maximum_radius = dilate_labels


@plugin_function(categories=['label', 'in assistant', 'bia-bob-suggestion'], output_creator=create_labels_like)
def voronoi_labeling(binary_source: Image, labeling_destination: Image = None) -> Image:
    """_tier5/_voronoi_labeling.py:7-39: CCL box into a float32 image, then
    extend_labeling_via_voronoi."""
    flip = create_like(labeling_destination)
    connected_components_labeling_box(binary_source, flip)
    return extend_labeling_via_voronoi(flip, labeling_destination)


@plugin_function(categories=['label', 'in assistant', 'bia-bob-suggestion'], priority=1,
                 output_creator=create_labels_like)
def voronoi_otsu_labeling(source: Image, label_image_destination: Image = None, spot_sigma: float = 2,
                          outline_sigma: float = 2) -> Image:
    """_tier9/_voronoi_otsu_labeling.py:11-73, as one device pipeline."""
    _lib.call("cle_voronoi_otsu_labeling", source.buf(), label_image_destination.buf(), float(spot_sigma),
              float(outline_sigma), _S())
    return label_image_destination


def voronoi_otsu_labeling_batch(sources, spot_sigmas=2, outline_sigmas=2, destinations=None, return_steps=False):
    """voronoi_otsu_labeling of many images at once (any sizes, 2D and 3D mixed).

    `sources`: a list of images (numpy arrays are pushed as float32, as the reference
    pushes them; device Arrays keep their type, so uint16 stacks are read natively);
    `spot_sigmas` / `outline_sigmas`: one value for all or one per image.  Returns the
    list of uint32 label images (or fills `destinations`).  Each result equals
    voronoi_otsu_labeling of that image alone.
    """
    srcs = [s if isinstance(s, Array) else push(s) for s in sources]
    n = len(srcs)
    ss = np.broadcast_to(np.asarray(spot_sigmas, np.float32), (n,)).copy()
    so = np.broadcast_to(np.asarray(outline_sigmas, np.float32), (n,)).copy()
    if destinations is None:
        destinations = [create_labels_like(s) for s in srcs]
    if len(destinations) != n:
        raise ValueError(f"{len(destinations)} destinations for {n} sources")
    src_bufs = (_lib.cle_buf * max(n, 1))(*[s.buf() for s in srcs])
    dst_bufs = (_lib.cle_buf * max(n, 1))(*[d.buf() for d in destinations])
    steps = ctypes.c_int(0)
    fp = ctypes.POINTER(ctypes.c_float)
    _lib.call("cle_voronoi_otsu_labeling_batch", src_bufs, dst_bufs, ss.ctypes.data_as(fp), so.ctypes.data_as(fp),
              n, ctypes.byref(steps), _S())
    return (destinations, steps.value) if return_steps else destinations


def gaussian_blur_batch(sources, sigmas=2, destinations=None, autotune: bool = False):
    """gaussian_blur (_tier1/_gaussian_blur.py:6-58) of many 2D / 3D images at once, the
    same sigma on every axis of an image (one value for all or one per image), into
    float32 images -- the batched pipeline's blur stage.  autotune=True also times the
    launch shapes of every group (ahead of later batches of the same classes)."""
    srcs = [s if isinstance(s, Array) else push(s) for s in sources]
    n = len(srcs)
    sg = np.broadcast_to(np.asarray(sigmas, np.float32), (n,)).copy()
    if destinations is None:
        destinations = [create_like(s) for s in srcs]
    if len(destinations) != n:
        raise ValueError(f"{len(destinations)} destinations for {n} sources")
    src_bufs = (_lib.cle_buf * max(n, 1))(*[s.buf() for s in srcs])
    dst_bufs = (_lib.cle_buf * max(n, 1))(*[d.buf() for d in destinations])
    fp = ctypes.POINTER(ctypes.c_float)
    _lib.call("cle_autotune_gaussian_blur_batch" if autotune else "cle_gaussian_blur_batch", src_bufs, dst_bufs,
              sg.ctypes.data_as(fp), n, _S())
    return destinations


def autotune_voronoi_otsu_labeling_batch(sources, spot_sigmas=2, outline_sigmas=2, destinations=None):
    """Tune the batched blur's launch shapes (tile height, register budget, z chunking)
    for every blur group of this batch ahead of time; later batches of the same
    classes use the choices.  Runs the batch (synchronising) and returns its labels."""
    srcs = [s if isinstance(s, Array) else push(s) for s in sources]
    n = len(srcs)
    ss = np.broadcast_to(np.asarray(spot_sigmas, np.float32), (n,)).copy()
    so = np.broadcast_to(np.asarray(outline_sigmas, np.float32), (n,)).copy()
    if destinations is None:
        destinations = [create_labels_like(s) for s in srcs]
    if len(destinations) != n:
        raise ValueError(f"{len(destinations)} destinations for {n} sources")
    src_bufs = (_lib.cle_buf * max(n, 1))(*[s.buf() for s in srcs])
    dst_bufs = (_lib.cle_buf * max(n, 1))(*[d.buf() for d in destinations])
    fp = ctypes.POINTER(ctypes.c_float)
    _lib.call("cle_autotune_voronoi_otsu_labeling_batch", src_bufs, dst_bufs, ss.ctypes.data_as(fp),
              so.ctypes.data_as(fp), n, _S())
    return destinations


class BatchPlan:
    """A batch of voronoi_otsu_labeling captured as one hipGraph for these exact source
    and destination buffers and sigmas (cle_voronoi_otsu_labeling_batch_plan).
    Building it runs the batch (the destinations hold its labels); `run()` replays the
    graph on the sources' current contents -- write new images into the same source
    buffers and run again.  The plan owns its device memory until `close()`."""

    def __init__(self, sources, spot_sigmas=2, outline_sigmas=2, destinations=None, graph_pairs: int = 32):
        srcs = [s if isinstance(s, Array) else push(s) for s in sources]
        n = len(srcs)
        if n == 0:
            raise ValueError("a plan needs at least one image")
        ss = np.broadcast_to(np.asarray(spot_sigmas, np.float32), (n,)).copy()
        so = np.broadcast_to(np.asarray(outline_sigmas, np.float32), (n,)).copy()
        if destinations is None:
            destinations = [create_labels_like(s) for s in srcs]
        if len(destinations) != n:
            raise ValueError(f"{len(destinations)} destinations for {n} sources")
        self.sources, self.destinations = srcs, list(destinations)  # the graph holds their addresses
        src_bufs = (_lib.cle_buf * n)(*[s.buf() for s in srcs])
        dst_bufs = (_lib.cle_buf * n)(*[d.buf() for d in self.destinations])
        fp = ctypes.POINTER(ctypes.c_float)
        h = ctypes.c_void_p()
        _lib.call("cle_voronoi_otsu_labeling_batch_plan", src_bufs, dst_bufs, ss.ctypes.data_as(fp),
                  so.ctypes.data_as(fp), n, int(graph_pairs), ctypes.byref(h), _S())
        self._h = h
        self.steps = None

    def run(self):
        """Replay: the labels of the sources' current contents, in `destinations`."""
        if self._h is None:
            raise ValueError("the plan is closed")
        steps = ctypes.c_int(0)
        _lib.call("cle_plan_run", self._h, ctypes.byref(steps), _S())
        self.steps = steps.value
        return self.destinations

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.call("cle_plan_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def voronoi_otsu_labeling_batch_plan(sources, spot_sigmas=2, outline_sigmas=2, destinations=None,
                                     graph_pairs: int = 32) -> BatchPlan:
    """voronoi_otsu_labeling of a batch as a replayable hipGraph (see BatchPlan)."""
    return BatchPlan(sources, spot_sigmas, outline_sigmas, destinations, graph_pairs)


@plugin_function(categories=['label', 'in assistant'], output_creator=create_labels_like)
def gauss_otsu_labeling(source: Image, label_image_destination: Image = None, outline_sigma: float = 2) -> Image:
    """_tier9/_gauss_otsu_labeling.py:6-49."""
    _lib.call("cle_gauss_otsu_labeling", source.buf(), label_image_destination.buf(), float(outline_sigma), _S())
    return label_image_destination


# ------------------------------------------------------------------ statistics
# STATISTICS_ENTRY column order of the device table (_tier9/_statistics_entry.py:3-46)
_TABLE_COLUMNS = [
    "original_label", "bbox_min_x", "bbox_min_y", "bbox_min_z", "bbox_max_x", "bbox_max_y", "bbox_max_z",
    "bbox_width", "bbox_height", "bbox_depth", "min_intensity", "max_intensity", "mean_intensity",
    "sum_intensity", "standard_deviation_intensity", "area", "sum_intensity_times_x", "sum_intensity_times_y",
    "sum_intensity_times_z", "mass_center_x", "mass_center_y", "mass_center_z", "sum_x", "sum_y", "sum_z",
    "centroid_x", "centroid_y", "centroid_z", "sum_distance_to_mass_center", "mean_distance_to_mass_center",
    "max_distance_to_mass_center", "mean_max_distance_to_mass_center_ratio", "sum_distance_to_centroid",
    "mean_distance_to_centroid", "max_distance_to_centroid", "mean_max_distance_to_centroid_ratio",
]
# dict insertion order of the reference (_tier9/_statistics_of_labelled_pixels.py:129-291)
STATISTICS_KEYS = [
    "label", "original_label",
    "bbox_min_x", "bbox_min_y", "bbox_min_z", "bbox_max_x", "bbox_max_y", "bbox_max_z",
    "bbox_width", "bbox_height", "bbox_depth",
    "min_intensity", "max_intensity", "sum_intensity", "area", "mean_intensity",
    "sum_intensity_times_x", "mass_center_x", "sum_intensity_times_y", "mass_center_y",
    "sum_intensity_times_z", "mass_center_z",
    "sum_x", "centroid_x", "sum_y", "centroid_y", "sum_z", "centroid_z",
    "sum_distance_to_centroid", "mean_distance_to_centroid",
    "sum_distance_to_mass_center", "mean_distance_to_mass_center",
    "standard_deviation_intensity", "max_distance_to_centroid", "max_distance_to_mass_center",
    "mean_max_distance_to_centroid_ratio", "mean_max_distance_to_mass_center_ratio",
]


_STAT_COLUMNS_HINT = [4096]  # table columns allocated when the label count is derived


def statistics_table(intensity_image: Array, label_image: Array, num_labels: int = None):
    """Device float32 table (36, columns >= num_labels - 1) in STATISTICS_ENTRY column
    order, and num_labels - 1.  Without num_labels the library derives max(labels) + 1
    in its run pass (no separate maximum pass); a table found too small is grown once."""
    out_n = ctypes.c_int64(0)
    if num_labels is not None:
        n = max(int(num_labels) - 1, 0)
        table = create((_lib.CLE_STAT_COLUMNS, max(n, 1)), np.float32)
        _lib.call("cle_statistics_of_labelled_pixels", intensity_image.buf(), label_image.buf(), int(num_labels),
                  table.buf(), ctypes.byref(out_n), _S())
        return table, n
    for _ in range(2):
        cols = _STAT_COLUMNS_HINT[0]
        table = create((_lib.CLE_STAT_COLUMNS, cols), np.float32)
        rc = _lib.call_rc("cle_statistics_of_labelled_pixels", intensity_image.buf(), label_image.buf(), 0,
                          table.buf(), ctypes.byref(out_n), _S())
        need = int(out_n.value) - 1
        if rc == _lib.CLE_E_SHAPE and need > cols:
            _STAT_COLUMNS_HINT[0] = 1 << max(need - 1, 1).bit_length()
            continue
        _lib.check(rc, "cle_statistics_of_labelled_pixels")
        return table, max(need, 0)
    raise RuntimeError("statistics table could not be sized")


def _table_to_host(t) -> np.ndarray:
    """The (columns, labels) table in host memory: one copy into page-locked memory
    (torch's host cache reuses a block once no array of an earlier table holds it)
    instead of the chunked pageable path."""
    if t.numel() == 0 or not t.is_cuda:
        return t.cpu().numpy()
    pinned = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
    pinned.copy_(t, non_blocking=True)
    torch.cuda.current_stream(t.device).synchronize()
    return pinned.numpy()


@plugin_function(output_creator=create_none, categories=['bia-bob-suggestion'])
def statistics_of_labelled_pixels(intensity_image: Image = None, label_image: Image = None):
    """_tier9/_statistics_of_labelled_pixels.py:7-293: dict of numpy float32 arrays
    (label and original_label are int64), one entry per label 1..max(label)."""
    if intensity_image is None and label_image is None:
        raise ValueError("Either intensity image or labels must be set")
    if label_image is None:
        warnings.warn("No label image provided. All pixels will be analysed in one group.")
        label_image = create_like(intensity_image)
        set(label_image, 1)
    if intensity_image is None:
        intensity_image = label_image
    table, n = statistics_table(intensity_image, label_image)
    num_labels = n + 1
    host = _table_to_host(table.tensor[:, :n])
    # each column is a contiguous row of the one host table (no per-column copy; the
    # arrays keep the table alive)
    cols = {k: host[i] for i, k in enumerate(_TABLE_COLUMNS)}
    out = {}
    for k in STATISTICS_KEYS:
        if k in ("label", "original_label"):
            out[k] = np.arange(1, num_labels)
        else:
            out[k] = cols[k]
    return out


statistics_of_labeled_pixels = statistics_of_labelled_pixels


@plugin_function(output_creator=create_none)
def statistics_of_background_and_labelled_pixels(source: Image = None, labelmap: Image = None):
    """_tier9/_statistics_of_background_and_labelled_pixels.py: statistics over labels + 1."""
    temp = add_image_and_scalar(labelmap, scalar=1)
    regionprops = statistics_of_labelled_pixels(source, temp)
    regionprops['original_label'] = regionprops['label'] - 1
    return regionprops


# ------------------------------------------------------------------ aliases (reference __init__ files)
label = connected_components_labeling_box
close_index_gaps_in_label_map = relabel_sequential
set_nonzero_pixels_to_pixelindex = set_non_zero_pixels_to_pixel_index
logical_and = binary_and
binary_intersection = binary_and


def set_autotune(on: bool = True):
    """Use (on, the default) or ignore and forget the blur's tuned z chunkings; results
    never depend on them."""
    _lib.call("cle_set_autotune", 1 if on else 0)


def set_growth_method(method: str = "steps"):
    """Voronoi growth method (results are identical): "steps" (default, image-wide
    pairs of steps) or "tiles" (asynchronous exact tile solves)."""
    _lib.call("cle_set_growth_method", {"tiles": 0, "steps": 1}[method])


def growth_counters(reset: bool = False) -> dict:
    """Pair-of-steps growth launches since the last reset: launches, tiles processed
    (1024 voxels each) and their GPU milliseconds."""
    out = (ctypes.c_ulonglong * 3)()
    _lib.call("cle_growth_counters", out, 1 if reset else 0)
    return {"launches": int(out[0]), "tiles": int(out[1]), "ms": out[2] / 1e6}


def autotune_gaussian_blur(source, destination=None, sigma_x: float = 0, sigma_y: float = 0, sigma_z: float = 0):
    """Tune the isotropic blur's z chunking for this configuration ahead of time (the
    call synchronises); later gaussian_blur calls of the same class use the choice.
    Returns the blurred destination."""
    source = source if isinstance(source, Array) else push(source)
    if destination is None:
        destination = create_like(source)
    _lib.call("cle_autotune_gaussian_blur", source.buf(), destination.buf(), float(sigma_x), float(sigma_y),
              float(sigma_z), _S())
    return destination


def release_pool():
    """Free the library's cached scratch memory on the current device (synchronises it)."""
    _lib.call("cle_release_pool")


def execute(*args, **kwargs):
    """Raw OpenCL kernel execution is not part of this build."""
    raise NotImplementedError("cle.execute (raw OpenCL kernels) is not supported by the MI355X build")
