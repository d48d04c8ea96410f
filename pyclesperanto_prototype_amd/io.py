"""The data formats on either side of the segmentation path: TIFF stacks in, the
statistics table out.

Ingest.  `imread_stack` decodes a TIFF stack page by page into pinned host slabs
and copies each slab to the device on a stream of its own while the next slab is
decoded, so disk / decode time and the host-to-device copy overlap; the stack stays
in its stored element type on the device (a uint16 nuclei stack takes 2 bytes per
voxel, and the blur reads it directly).  `imread` keeps the reference's result: the
image pushed as float32 (pyclesperanto_prototype/_tier9/_imread.py:3-10, which reads
with skimage.io.imread and pushes; the float32 conversion here runs on the device).

Export.  `push_regionprops` / `push_regionprops_column` build the (36, N) table in
STATISTICS_ENTRY row order (or one column of it) from a statistics dict or from
region objects with the regionprops attributes
(pyclesperanto_prototype/_tier9/_push_regionprops.py:8-193,
_tier9/_push_regionprops_column.py:4-24).  `statistics_table` in the package returns
the same table straight from the device computation.
"""
from __future__ import annotations

import warnings

import numpy as np

from ._array import Array, _torch_dtype, device, push

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

# STATISTICS_ENTRY rows (_tier9/_statistics_entry.py:3-46) and the statistics-dict
# key each row is taken from (_tier9/_push_regionprops.py:29-102)
REGIONPROPS_ROWS = [
    "original_label", "bbox_min_x", "bbox_min_y", "bbox_min_z", "bbox_max_x", "bbox_max_y", "bbox_max_z",
    "bbox_width", "bbox_height", "bbox_depth", "min_intensity", "max_intensity", "mean_intensity",
    "sum_intensity", "standard_deviation_intensity", "area", "sum_intensity_times_x", "sum_intensity_times_y",
    "sum_intensity_times_z", "mass_center_x", "mass_center_y", "mass_center_z", "sum_x", "sum_y", "sum_z",
    "centroid_x", "centroid_y", "centroid_z", "sum_distance_to_mass_center", "mean_distance_to_mass_center",
    "max_distance_to_mass_center", "mean_max_distance_to_mass_center_ratio", "sum_distance_to_centroid",
    "mean_distance_to_centroid", "max_distance_to_centroid", "mean_max_distance_to_centroid_ratio",
]
# rows a region object may carry as extra attributes (shape descriptors)
_OPTIONAL_ROWS = REGIONPROPS_ROWS[28:]


# ----------------------------------------------------------------------------- TIFF
def tiff_info(filename):
    """(pages, height, width, numpy dtype) of a single-channel TIFF stack."""
    from PIL import Image

    with Image.open(filename) as im:
        pages = getattr(im, "n_frames", 1)
        first = _page_array(im)
    if first.ndim != 2:
        raise ValueError(f"{filename}: pages of shape {first.shape}; single-channel pages expected")
    return pages, first.shape[0], first.shape[1], first.dtype


def _page_array(im) -> np.ndarray:
    a = np.asarray(im)
    if a.dtype.byteorder not in ("=", "|"):
        a = a.astype(a.dtype.newbyteorder("="))
    return a


def read_tiff_pages(filename, start: int = 0, stop: int = None):
    """Host decode of pages [start, stop) as numpy arrays (one per page)."""
    from PIL import Image

    with Image.open(filename) as im:
        n = getattr(im, "n_frames", 1)
        stop = n if stop is None else min(stop, n)
        for p in range(start, stop):
            im.seek(p)
            yield _page_array(im)


def imread_stack(filename, slab_pages: int = 16) -> Array:
    """Device Array of a TIFF stack in its stored element type: (pages, H, W), or
    (H, W) for a single page.  Pages are decoded into two pinned host slabs in turn;
    each slab is copied on a separate stream while the next one is decoded."""
    pages, h, w, dt = tiff_info(filename)
    tdt = _torch_dtype(dt)
    # pinned staging and copies use a signed type of the same width (torch copies
    # unsigned 16/32-bit tensors only between equal types; the bits are the same)
    stage_dt = {torch.uint16: torch.int16, torch.uint32: torch.int32}.get(tdt, tdt)
    dev = device()
    out = torch.empty((pages, h, w), dtype=stage_dt, device=dev)
    slab_pages = max(1, min(int(slab_pages), pages))
    slabs = [torch.empty((slab_pages, h, w), dtype=stage_dt, pin_memory=True) for _ in range(min(2, pages))]
    done = [None] * len(slabs)
    copy_stream = torch.cuda.Stream(device=dev)
    copy_stream.wait_stream(torch.cuda.current_stream(dev))  # `out` is allocated on the current stream
    pages_iter = read_tiff_pages(filename)
    for k, s0 in enumerate(range(0, pages, slab_pages)):
        s1 = min(s0 + slab_pages, pages)
        slot = k % len(slabs)
        if done[slot] is not None:
            done[slot].synchronize()  # the slab's previous copy has left host memory
        host = slabs[slot].numpy().view(dt)
        for i in range(s1 - s0):
            host[i] = next(pages_iter)
        with torch.cuda.stream(copy_stream):
            out[s0:s1].copy_(slabs[slot][: s1 - s0], non_blocking=True)
            done[slot] = torch.cuda.Event()
            done[slot].record(copy_stream)
    torch.cuda.current_stream(dev).wait_stream(copy_stream)
    out.record_stream(copy_stream)
    for ev in done:
        if ev is not None:
            ev.synchronize()  # the pinned slabs are freed when this function returns
    t = out.view(tdt) if stage_dt != tdt else out
    return Array(t[0] if pages == 1 else t)


def imread(filename) -> Array:
    """The image as a float32 device Array (pyclesperanto_prototype/_tier9/_imread.py:3-10)."""
    warnings.warn("cle.imread is deprecated, use skimage.io.imread instead.")
    return push(imread_stack(filename).tensor)


# ----------------------------------------------------------------------------- regionprops
def _region_matrix(props, first_row_index: int = 0) -> np.ndarray:
    """Host (36, N) table in STATISTICS_ENTRY row order (the layout push_regionprops
    uploads)."""
    if isinstance(props, dict):
        return np.asarray([np.asarray(props[k], dtype=np.float64) for k in REGIONPROPS_ROWS])
    labels = [r.original_label if hasattr(r, "original_label") else r.label for r in props]
    num_rows = int(np.max(labels)) + 1 - first_row_index
    m = np.zeros((num_rows, len(REGIONPROPS_ROWS)))
    col = {k: i for i, k in enumerate(REGIONPROPS_ROWS)}
    for r, lab in zip(props, labels):
        i = int(lab)
        if i < first_row_index:
            continue
        i -= first_row_index
        row = m[i]
        row[col["original_label"]] = i + first_row_index
        bbox = r.bbox
        if len(bbox) == 4:  # (min_row, min_col, max_row, max_col), max exclusive
            row[col["bbox_min_x"]], row[col["bbox_min_y"]], row[col["bbox_min_z"]] = bbox[1], bbox[0], 0
            row[col["bbox_max_x"]], row[col["bbox_max_y"]], row[col["bbox_max_z"]] = bbox[3] - 1, bbox[2] - 1, 0
        else:
            row[col["bbox_min_x"]], row[col["bbox_min_y"]], row[col["bbox_min_z"]] = bbox[2], bbox[1], bbox[0]
            row[col["bbox_max_x"]], row[col["bbox_max_y"]], row[col["bbox_max_z"]] = bbox[5] - 1, bbox[4] - 1, bbox[3] - 1
        for a, lo, hi in (("bbox_width", "bbox_min_x", "bbox_max_x"), ("bbox_height", "bbox_min_y", "bbox_max_y"),
                          ("bbox_depth", "bbox_min_z", "bbox_max_z")):
            row[col[a]] = row[col[hi]] - row[col[lo]] + 1
        row[col["min_intensity"]] = r.min_intensity
        row[col["max_intensity"]] = r.max_intensity
        row[col["mean_intensity"]] = r.mean_intensity
        row[col["sum_intensity"]] = r.mean_intensity * r.area
        if hasattr(r, "standard_deviation_intensity"):
            row[col["standard_deviation_intensity"]] = r.standard_deviation_intensity
        else:
            warnings.warn("regionprops didn't contains stanard_deviation_intensity. Consider using "
                          "cle.statistics_of_labelled_pixels to determine the standard deviation. ")
            row[col["standard_deviation_intensity"]] = -1
        row[col["area"]] = r.area
        for name, src in (("mass_center", r.weighted_centroid), ("centroid", r.centroid)):
            if len(src) == 3:
                row[col[name + "_x"]], row[col[name + "_y"]], row[col[name + "_z"]] = src[2], src[1], src[0]
            else:
                row[col[name + "_x"]], row[col[name + "_y"]], row[col[name + "_z"]] = src[1], src[0], 0
        for k in _OPTIONAL_ROWS:
            if hasattr(r, k):
                row[col[k]] = getattr(r, k)
    return m.T


def push_regionprops(props, first_row_index: int = 0) -> Array:
    """(36, N) float32 device table from a statistics dict or region objects
    (_tier9/_push_regionprops.py:8-193)."""
    return push(_region_matrix(props, first_row_index))


def _region_column(regionprops, column: str) -> np.ndarray:
    if isinstance(regionprops, dict):
        return np.asarray([regionprops[column]])
    labels = [r.original_label if hasattr(r, "original_label") else r.label for r in regionprops]
    values = np.zeros([int(np.max(labels)) + 1])
    for r, lab in zip(regionprops, labels):
        values[int(lab)] = r[column] if hasattr(r, "__getitem__") else getattr(r, column)
    return np.asarray([values])


def push_regionprops_column(regionprops, column: str) -> Array:
    """One statistics column as a (1, N) float32 device row
    (_tier9/_push_regionprops_column.py:4-24)."""
    return push(_region_column(regionprops, column))
