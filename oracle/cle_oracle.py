"""CPU oracle for the segmentation hot path -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy, the algorithms of pyclesperanto_prototype 0.24.5 for
voronoi_otsu_labeling, connected_components_labeling_box and
statistics_of_labelled_pixels, in the reference's own order of floating-point
operations where that order is visible.  It is the checker for the HIP path: only
tests, the smoke check and the CPU-baseline leg of the benchmark may import it.
Nothing in the product package imports it.

Array convention: numpy arrays in C order (z, y, x) with x fastest; 2D arrays are
(y, x).  Every function takes and returns plain numpy arrays.

Pinning status of each function (which independent check fixes it) is recorded in
the function docstrings; the pin tests live in tests/test_oracle_pins.py.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
FLT_MAX = np.finfo(np.float32).max


# --------------------------------------------------------------------------- helpers
def _as3d(a: np.ndarray) -> np.ndarray:
    """View a 2D (y, x) array as (1, y, x)."""
    return a[np.newaxis] if a.ndim == 2 else a


def _clamp_shift(a3: np.ndarray, axis: int, v: int) -> np.ndarray:
    """a3 read at (p + v*e_axis) with clamp-to-edge.

    Reference: read_buffer3d* clamps every coordinate into [0, size-1]
    (pyclesperanto_prototype/_tier0/preamble.cl:33-49).
    """
    n = a3.shape[axis]
    idx = np.clip(np.arange(n) + v, 0, n - 1)
    return np.take(a3, idx, axis=axis)


def convert_sat(values: np.ndarray, dtype) -> np.ndarray:
    """Float -> integer store: saturate, then truncate toward zero; NaN -> 0.

    Reference: clij_convert_<t>_sat (pyclesperanto_prototype/_tier0/preamble.cl:505-591);
    pinned by tests/test_create.py:30-60 of the reference (1.5->1, -7.8->0, 2000->255).
    The reference tests `value > max` against max as a float, which rounds up to 2^31 /
    2^32 for the 32-bit types, so exactly 2^31 / 2^32 would reach an out-of-range cast
    whose result the OpenCL standard leaves to the implementation; here (and on the
    device) every value from the rounded limit on saturates.
    """
    dtype = np.dtype(dtype)
    v = np.asarray(values, dtype=np.float32)
    if dtype.kind == "f":
        return v.astype(dtype)
    info = np.iinfo(dtype)
    out = np.where(v >= np.float32(info.max), np.float64(info.max),
                   np.where(v < np.float32(info.min), np.float64(info.min), np.trunc(v.astype(np.float64))))
    out = np.nan_to_num(out, nan=0.0)
    return out.astype(dtype)


# --------------------------------------------------------------------------- A.1 blur
def sigma_to_kernel_size(sigma: float) -> int:
    """Reference: pyclesperanto_prototype/_tier0/_sigma_to_kernel_size.py:1-4."""
    n = int(float(sigma) * 8.0)
    if n % 2 == 0:
        n = n + 1
    return n


def _gauss_axis(a3: np.ndarray, axis: int, n_taps: int, sigma: float) -> np.ndarray:
    """One separable pass: normalised Gaussian of n_taps along `axis`, clamp-to-edge,
    accumulated tap by tap in float32 from v=-c to v=+c.

    Restates the clij gaussian_blur_separable kernel as launched by
    pyclesperanto_prototype/_tier1/_execute_separable_kernel.py:31,55,72 (kernel
    source absent from the reference tree).  Pinned by the closed-form centre weight
    and the difference-of-Gaussian bounds (tests/test_oracle_pins.py).
    """
    c = (n_taps - 1) // 2
    nrm = F32(-2.0) * F32(sigma) * F32(sigma)
    res = np.zeros(a3.shape, np.float32)
    hsum = F32(0.0)
    for v in range(-c, c + 1):
        h = F32(np.exp(F32(v * v) / nrm))
        res = (res + h * _clamp_shift(a3, axis, v)).astype(np.float32)
        hsum = F32(hsum + h)
    return (res / hsum).astype(np.float32)


def gaussian_blur(src: np.ndarray, sigma_x: float = 0, sigma_y: float = 0, sigma_z: float = 0) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier1/_gaussian_blur.py:6-58 and
    _execute_separable_kernel.py:7-75 (x pass, then y, then z for 3D; sigma <= 0
    copies that axis; sigma_z ignored for 2D).  Returns float32."""
    ndim = src.ndim
    a = _as3d(np.asarray(src, np.float32))
    sig = [sigma_x, sigma_y, sigma_z]
    axes = [2, 1, 0]  # x, y, z in (z, y, x) storage
    for d in range(ndim):
        if sig[d] > 0:
            a = _gauss_axis(a, axes[d], sigma_to_kernel_size(sig[d]), sig[d])
    return a[0] if ndim == 2 else a


def mean_box(src: np.ndarray, radius_x: int = 1, radius_y: int = 1, radius_z: int = 1) -> np.ndarray:
    """Separable box mean; radius 0 copies that axis.

    Reference: pyclesperanto_prototype/_tier1/_mean_box.py:6-54 (kernel size
    2r+1 per axis through execute_separable_kernel; the mean_separable kernel itself is
    absent from the reference tree).  Pinned by tests/test_mean_box.py:5-29 of the
    reference (radius 1, 1, 0) and by a float64 brute-force window mean with
    clamp-to-edge reads at several radii (tests/test_oracle_pins.py); the edge rule is
    the clamp of every separable kernel's reads (_tier0/preamble.cl:33-49).
    """
    ndim = src.ndim
    a = _as3d(np.asarray(src, np.float32))
    rad = [radius_x, radius_y, radius_z]
    axes = [2, 1, 0]
    for d in range(ndim):
        r = int(rad[d])
        if r > 0:
            n = 2 * r + 1
            s = np.zeros(a.shape, np.float32)
            for v in range(-r, r + 1):
                s = (s + _clamp_shift(a, axes[d], v)).astype(np.float32)
            a = (s / F32(n)).astype(np.float32)
    return a[0] if ndim == 2 else a


# --------------------------------------------------------------------------- A.2 maxima
def detect_maxima_box(src: np.ndarray, radius_x: int = 0, radius_y: int = 0, radius_z: int = 0,
                      out_dtype=np.uint8) -> np.ndarray:
    """Local maxima in the 3^d window with the reference's visiting order.

    Reference: pyclesperanto_prototype/_tier1/_detect_maxima_box.py:36-46 (mean_box
    first) and detect_maxima_2d_x.cl:12-34 / detect_maxima_3d_x.cl:22-48: the window
    is always radius 1; neighbours with a negative coordinate are skipped; upper
    neighbours are read clamped; the running maximum starts at v(p)-1 and a
    neighbour wins only if strictly greater; loop order dx outer, dy, dz inner.
    Pinned by the reference test tests/test_detect_maxima_box.py:4-35 and by a
    brute-force restatement on random plateaus (tests/test_oracle_pins.py).
    """
    ndim = src.ndim
    a = _as3d(mean_box(src, radius_x, radius_y, radius_z))
    D, H, W = a.shape
    local_max = (a - F32(1)).astype(np.float32)
    # position of current winner as offset from p; (0,0,0) initially
    win = np.zeros(a.shape, bool)  # True where the current winner is p itself
    win[...] = True
    zr = (-1, 0, 1) if ndim == 3 else (0,)
    zz, yy, xx = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in zr:
                qx, qy, qz = xx + dx, yy + dy, zz + dz
                valid = (qx >= 0) & (qy >= 0) & (qz >= 0)
                cx, cy, cz = np.minimum(qx, W - 1), np.minimum(qy, H - 1), np.minimum(qz, D - 1)
                val = a[cz, cy, cx]
                better = valid & (val > local_max)
                local_max = np.where(better, val, local_max)
                is_self = (dx == 0) and (dy == 0) and (dz == 0)
                win = np.where(better, is_self, win)
    out = win.astype(out_dtype)
    return out[0] if ndim == 2 else out


# --------------------------------------------------------------------------- A.3-A.5 Otsu
def minimum_of_all_pixels(src: np.ndarray) -> np.float32:
    """Reference: pyclesperanto_prototype/_tier2/_minimum_of_all_pixels.py:34-59 (exact min)."""
    return np.float32(np.min(np.asarray(src, np.float32)))


def maximum_of_all_pixels(src: np.ndarray) -> np.float32:
    """Reference: pyclesperanto_prototype/_tier2/_maximum_of_all_pixels.py (exact max)."""
    return np.float32(np.max(np.asarray(src, np.float32)))


def histogram_bin_index(values: np.ndarray, mn, mx, num_bins: int) -> np.ndarray:
    """bin = convert_uint_sat(((v - min) * (bins-1)) / (max - min) + 0.5): float32
    quotient, the literal 0.5 added in float64, then truncation with saturation.

    Restates the clij histogram kernel launched at
    pyclesperanto_prototype/_tier3/_histogram.py:117-122 (kernel source absent).
    NaN (range 0) maps to bin 0; values below min saturate to 0.  Pinned by
    tests/test_histogram.py:4-56 of the reference (integer data).
    """
    v = np.asarray(values, np.float32)
    mn = F32(mn)
    rng = F32(F32(mx) - mn)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = ((v - mn) * F32(num_bins - 1)).astype(np.float32)
        t = (t / rng).astype(np.float32)
        # the unsuffixed literal 0.5 is a double in OpenCL C on fp64 devices
        t = t.astype(np.float64) + 0.5
    t = np.nan_to_num(t, nan=0.0, posinf=np.inf, neginf=0.0)
    idx = np.floor(np.clip(t, 0, None)).astype(np.int64)
    return np.minimum(idx, num_bins - 1)


def histogram(src: np.ndarray, num_bins: int = 256, minimum_intensity=None, maximum_intensity=None,
              determine_min_max: bool = True) -> np.ndarray:
    """Per-row partial histograms (one per y, over all z and x), stored as float32,
    then summed over rows in y order in float32.

    Reference: pyclesperanto_prototype/_tier3/_histogram.py:85-128 (partials
    [H,1,bins] then sum_z_projection).  Returns float32 (num_bins,).
    """
    a = _as3d(np.asarray(src, np.float32))
    if minimum_intensity is None or maximum_intensity is None or determine_min_max:
        minimum_intensity = minimum_of_all_pixels(a)
        maximum_intensity = maximum_of_all_pixels(a)
    D, H, W = a.shape
    idx = histogram_bin_index(a, minimum_intensity, maximum_intensity, num_bins)
    partial = np.zeros((H, num_bins), np.int64)
    for y in range(H):
        partial[y] = np.bincount(idx[:, y, :].ravel(), minlength=num_bins)[:num_bins]
    total = np.zeros(num_bins, np.float32)
    for y in range(H):
        total = (total + partial[y].astype(np.float32)).astype(np.float32)
    return total


def otsu_threshold_from_histogram(hist: np.ndarray, mn, mx) -> float:
    """Otsu threshold over 256 bin centres in float64.

    Reference: pyclesperanto_prototype/_tier9/_threshold_otsu.py:38-50 (bin centres
    arange(256)*range/255 + min with range computed in float32) and the vendored
    scikit-image routine at :64-106 (cumulative sums, first argmax).  Pinned by the
    brute-force between-class-variance maximiser (tests/test_oracle_pins.py).
    """
    rng = np.float32(np.float32(mx) - np.float32(mn))
    bin_centers = np.arange(256) * np.float64(rng) / 255 + np.float64(np.float32(mn))
    counts = np.asarray(hist).astype(float)
    weight1 = np.cumsum(counts)
    weight2 = np.cumsum(counts[::-1])[::-1]
    with np.errstate(invalid="ignore", divide="ignore"):
        mean1 = np.cumsum(counts * bin_centers) / weight1
        mean2 = (np.cumsum((counts * bin_centers)[::-1]) / weight2[::-1])[::-1]
        variance12 = weight1[:-1] * weight2[1:] * (mean1[:-1] - mean2[1:]) ** 2
    idx = np.argmax(variance12)
    return float(bin_centers[:-1][idx])


def threshold_otsu(src: np.ndarray, out_dtype=np.uint8, return_threshold: bool = False):
    """Reference: pyclesperanto_prototype/_tier9/_threshold_otsu.py:11-56."""
    a = np.asarray(src, np.float32)
    mn, mx = minimum_of_all_pixels(a), maximum_of_all_pixels(a)
    hist = histogram(a, 256, mn, mx, determine_min_max=False)
    t = otsu_threshold_from_histogram(hist, mn, mx)
    out = greater_constant(a, t, out_dtype)
    return (out, t) if return_threshold else out


# --------------------------------------------------------------------------- A.6 elementwise
def greater_constant(src: np.ndarray, constant: float, out_dtype=np.uint8) -> np.ndarray:
    """src > float32(c).  Reference: pyclesperanto_prototype/_tier1/_greater_constant.py:5-40."""
    return (np.asarray(src, np.float32) > F32(constant)).astype(out_dtype)


def binary_and(a: np.ndarray, b: np.ndarray, out_dtype=np.uint8) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier1/_binary_and.py:6-47."""
    return ((np.asarray(a, np.float32) != 0) & (np.asarray(b, np.float32) != 0)).astype(out_dtype)


def mask(src: np.ndarray, mask_image: np.ndarray, out_dtype=np.float32) -> np.ndarray:
    """mask != 0 ? src : 0.  Reference: pyclesperanto_prototype/_tier1/_mask.py:5-45."""
    v = np.where(np.asarray(mask_image, np.float32) != 0, np.asarray(src, np.float32), F32(0))
    return convert_sat(v, out_dtype)


def add_image_and_scalar(src: np.ndarray, scalar: float, out_dtype=np.float32) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier1/_add_image_and_scalar.py:5-43."""
    return convert_sat((np.asarray(src, np.float32) + F32(scalar)).astype(np.float32), out_dtype)


def add_images_weighted(a: np.ndarray, b: np.ndarray, fa: float = 1, fb: float = 1, out_dtype=np.float32):
    """Reference: pyclesperanto_prototype/_tier1/_add_images_weighted.py:5-50."""
    v = (np.asarray(a, np.float32) * F32(fa) + np.asarray(b, np.float32) * F32(fb)).astype(np.float32)
    return convert_sat(v, out_dtype)


# --------------------------------------------------------------------------- A.8/A.9 CCL
def pixel_keys(src: np.ndarray, offset: int = 1) -> np.ndarray:
    """key = x*H*D + y*D + z + offset where src != 0 (int64, exact).

    Reference: pyclesperanto_prototype/_tier1/_set_non_zero_pixels_to_pixel_index.py:5-28
    (kernel absent); the 2D case is pinned by tests/test_set_nonzero_pixels_to_pixelindex.py:4-26
    of the reference.  The 3D term order (y*D + z) is inferred.
    """
    a = _as3d(np.asarray(src))
    D, H, W = a.shape
    zz, yy, xx = np.meshgrid(np.arange(D, dtype=np.int64), np.arange(H, dtype=np.int64),
                             np.arange(W, dtype=np.int64), indexing="ij")
    key = xx * H * D + yy * D + zz + offset
    key = np.where(a != 0, key, 0)
    return key[0] if src.ndim == 2 else key


def set_nonzero_pixels_to_pixel_index(src: np.ndarray, offset: int = 1) -> np.ndarray:
    """The reference stores the key into a float32 image (rounds above 2^24)."""
    return pixel_keys(src, offset).astype(np.float32)


def _neighbour_offsets(ndim: int, kind: str):
    zr = (-1, 0, 1) if ndim == 3 else (0,)
    offs = []
    for dz in zr:
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                if kind == "diamond" and abs(dx) + abs(dy) + abs(dz) != 1:
                    continue
                offs.append((dz, dy, dx))
    return offs


def _shift_clamped(a3: np.ndarray, dz: int, dy: int, dx: int) -> np.ndarray:
    out = a3
    if dz:
        out = _clamp_shift(out, 0, dz)
    if dy:
        out = _clamp_shift(out, 1, dy)
    if dx:
        out = _clamp_shift(out, 2, dx)
    return out


def nonzero_minimum(src: np.ndarray, dst_prev: np.ndarray, kind: str = "box", ndim: int = None):
    """One Jacobi step: for src(p) != 0, dst(p) = min of the non-zero values in the
    box (or diamond) neighbourhood incl. p; zero voxels keep dst_prev.  Returns
    (dst, changed).

    Reference: pyclesperanto_prototype/_tier1/_nonzero_minimum_box.py:5-43 and
    _nonzero_minimum_diamond (kernels absent); pinned by
    tests/test_nonzero_minimum_box.py:4-33 and test_nonzero_minimum_diamond.py of the reference.
    """
    ndim = src.ndim if ndim is None else ndim
    a = _as3d(np.asarray(src, np.float32))
    best = a.copy()
    for dz, dy, dx in _neighbour_offsets(ndim, kind):
        q = _shift_clamped(a, dz, dy, dx)
        best = np.where((q != 0) & (q < best), q, best)
    out = np.where(a != 0, best, _as3d(np.asarray(dst_prev, np.float32)))
    changed = bool(np.any(out != a))
    out = out.astype(np.float32)
    return (out[0] if src.ndim == 2 else out), changed


def relabel_sequential(src: np.ndarray, out_dtype=np.uint32) -> np.ndarray:
    """dst = 1-based rank of int(src) among the distinct non-zero values present.

    Reference: pyclesperanto_prototype/_tier3/_relabel_sequential.py:12-59 (flag
    vector, block sums, block enumerate, replace_intensities); pinned by
    tests/test_relabel_sequential.py:4-33 of the reference.
    """
    v = np.asarray(src, np.float32).astype(np.int64)
    present = np.unique(v[v > 0])
    lut = np.zeros(int(v.max(initial=0)) + 1, np.float32)
    lut[present] = np.arange(1, present.size + 1, dtype=np.float32)
    out = np.where(v > 0, lut[np.clip(v, 0, None)], F32(0))
    return convert_sat(out, out_dtype)


def connected_components_labeling(binary: np.ndarray, kind: str = "box", out_dtype=np.uint32,
                                  return_iterations: bool = False):
    """The reference algorithm verbatim: float32 pixel keys, Jacobi non-zero-minimum
    ping-pong until an iteration changes nothing, then relabel_sequential of the
    last written buffer.

    Reference: pyclesperanto_prototype/_tier4/_connected_components_labeling_box.py:39-64.
    Pinned by tests/test_connected_components_labeling_box.py:6-64 of the reference
    (incl. blobs.tif -> 62 labels) and by partition equality with
    scipy.ndimage.label plus the min-key order invariant (tests/test_oracle_pins.py).
    """
    ndim = binary.ndim
    t1 = set_nonzero_pixels_to_pixel_index(binary)
    t2 = np.zeros_like(t1)
    it = 0
    flag = True
    while flag:
        if it % 2 == 0:
            t2, flag = nonzero_minimum(t1, t2, kind, ndim)
        else:
            t1, flag = nonzero_minimum(t2, t1, kind, ndim)
        it += 1
    last = t1 if it % 2 == 0 else t2
    out = relabel_sequential(last, out_dtype)
    return (out, it) if return_iterations else out


# --------------------------------------------------------------------------- A.10 Voronoi
def onlyzero_overwrite_maximum(src: np.ndarray, kind: str = "box", ndim: int = None):
    """One growth step: non-zero voxels are copied; a zero voxel takes the largest
    positive value among its box (or diamond) neighbours (clamped reads), else stays 0.
    Returns (dst, changed).

    Reference: pyclesperanto_prototype/_tier1/_onlyzero_overwrite_maximum_box.py:5-38 and
    _onlyzero_overwrite_maximum_diamond.py:5-38 (kernels absent); pinned by
    tests/test_onlyzero_overwrite_maximum_box.py:4-29 and the expected array of the
    uncollected test_onlyzero_overwrite_maximum_diamond.py:14-20 of the reference.
    """
    ndim = src.ndim if ndim is None else ndim
    a = _as3d(np.asarray(src, np.float32))
    best = np.zeros_like(a)
    for dz, dy, dx in _neighbour_offsets(ndim, kind):
        q = _shift_clamped(a, dz, dy, dx)
        best = np.where(q > best, q, best)
    out = np.where(a != 0, a, best).astype(np.float32)
    changed = bool(np.any(out != a))
    return (out[0] if src.ndim == 2 else out), changed


def masked_voronoi_labeling(binary: np.ndarray, mask_image: np.ndarray, out_dtype=np.uint32,
                            return_iterations: bool = False):
    """flup = mask - 1; flop = CCL_box(binary) (float32); flip = flup + flop; then
    alternate box (even iterations) / diamond (odd) growth until an iteration
    changes nothing; finally mask the last written buffer.

    Reference: pyclesperanto_prototype/_tier5/_masked_voronoi_labeling.py:44-74.
    Pinned by tests/test_masked_voronoi_labeling.py:4-47 of the reference.
    """
    ndim = binary.ndim
    flup = (np.asarray(mask_image, np.float32) - F32(1)).astype(np.float32)
    flop = connected_components_labeling(binary, "box", np.float32)
    flip = (flup + flop).astype(np.float32)
    it = 0
    flag = True
    while flag:
        if it % 2 == 0:
            flop, flag = onlyzero_overwrite_maximum(flip, "box", ndim)
        else:
            flip, flag = onlyzero_overwrite_maximum(flop, "diamond", ndim)
        it += 1
    last = flip if it % 2 == 0 else flop
    out = mask(last, mask_image, out_dtype)
    return (out, it) if return_iterations else out


def extend_labeling_via_voronoi(labels: np.ndarray, out_dtype=np.uint32, return_iterations: bool = False):
    """flip = float32(labels); alternate box (even iterations, flip -> flop) / diamond
    (odd, flop -> flip) growth until an iteration changes nothing; the last written
    buffer is copied (saturating) into the destination type.  No mask.

    Reference: pyclesperanto_prototype/_tier4/_extend_labeling_via_voronoi.py:39-61.
    Pinned by tests/test_extend_labeling_via_voronoi.py:4-36 of the reference and by
    voronoi_labeling (_tier5/_voronoi_labeling.py:33-37 = CCL box + this function),
    whose reference test array it must reproduce.
    """
    ndim = labels.ndim
    flip = np.asarray(labels).astype(np.float32)
    flop = flip
    it = 0
    flag = True
    while flag:
        if it % 2 == 0:
            flop, flag = onlyzero_overwrite_maximum(flip, "box", ndim)
        else:
            flip, flag = onlyzero_overwrite_maximum(flop, "diamond", ndim)
        it += 1
    last = flip if it % 2 == 0 else flop
    out = convert_sat(last, out_dtype)
    return (out, it) if return_iterations else out


def dilate_labels(labels: np.ndarray, radius: int = 2, out_dtype=np.uint32) -> np.ndarray:
    """At most `radius` alternating box (even) / diamond (odd) growth steps on float32
    buffers, stopping after a step that changes nothing; the last written buffer is
    copied (saturating) into the destination type.

    Reference: pyclesperanto_prototype/_tier4/_dilate_labels.py:41-61 (exported as
    extend_labels_with_maximum_radius, _tier4/__init__.py:18).  Pinned by
    tests/test_dilate_labels.py:4-75 of the reference (2D, and 3D across z).
    """
    ndim = labels.ndim
    flip = np.asarray(labels).astype(np.float32)
    flop = flip
    it = 0
    flag = True
    while flag and it < radius:
        if it % 2 == 0:
            flop, flag = onlyzero_overwrite_maximum(flip, "box", ndim)
        else:
            flip, flag = onlyzero_overwrite_maximum(flop, "diamond", ndim)
        it += 1
    return convert_sat(flip if it % 2 == 0 else flop, out_dtype)


def voronoi_labeling(binary: np.ndarray, out_dtype=np.uint32) -> np.ndarray:
    """CCL box into float32, then extend_labeling_via_voronoi.

    Reference: pyclesperanto_prototype/_tier5/_voronoi_labeling.py:33-37."""
    return extend_labeling_via_voronoi(connected_components_labeling(binary, "box", np.float32), out_dtype)


def voronoi_otsu_labeling(src: np.ndarray, spot_sigma: float = 2, outline_sigma: float = 2,
                          out_dtype=np.uint32) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier9/_voronoi_otsu_labeling.py:57-73."""
    a = np.asarray(src, np.float32)
    temp = gaussian_blur(a, spot_sigma, spot_sigma, spot_sigma)
    spots = detect_maxima_box(temp, 0, 0, 0, np.float32)
    temp = gaussian_blur(a, outline_sigma, outline_sigma, outline_sigma)
    segmentation = threshold_otsu(temp, np.float32)
    binary = binary_and(spots, segmentation, np.uint8)
    vor = masked_voronoi_labeling(binary, segmentation, np.float32)
    return mask(vor, segmentation, out_dtype)


def gauss_otsu_labeling(src: np.ndarray, outline_sigma: float = 2, out_dtype=np.uint32) -> np.ndarray:
    """Reference: pyclesperanto_prototype/_tier9/_gauss_otsu_labeling.py:6-49 (blur, Otsu, CCL box)."""
    blurred = gaussian_blur(np.asarray(src, np.float32), outline_sigma, outline_sigma, outline_sigma)
    binary = threshold_otsu(blurred, np.uint8)
    return connected_components_labeling(binary, "box", out_dtype)


# --------------------------------------------------------------------------- A.11 spots
def label_spots(spots: np.ndarray, out_dtype=np.uint32) -> np.ndarray:
    """Non-zero voxels numbered 1.. in raster order (z, then y, then x).

    Reference: pyclesperanto_prototype/_tier2/_label_spots.py:9-56 and
    label_spots_in_x.cl:3-31 (valid while depth <= width).  Pinned by
    tests/test_label_spots.py:4-74 of the reference.
    """
    a = np.asarray(spots, np.float32)
    nz = (a != 0).ravel()
    out = np.zeros(nz.shape, np.float64)
    out[nz] = np.arange(1, int(nz.sum()) + 1)
    return convert_sat(out.reshape(a.shape).astype(np.float32), out_dtype)


# --------------------------------------------------------------------------- A.12 statistics
STAT_KEYS = [
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


def statistics_of_labelled_pixels(intensity: np.ndarray, labels: np.ndarray, sum_background: int = 0) -> dict:
    """Per-label statistics with the reference's float32 accumulation structure:
    one accumulator cell per (plane, row y, label), updated voxel by voxel in (z, x)
    order, then summed over y in y order (min/max over y for the extreme planes).

    Reference: pyclesperanto_prototype/_tier9/_statistics_of_labelled_pixels.py:43-291,
    statistics_per_label_x.cl:1-146, standard_deviation_per_label_x.cl:1-96.  Pinned
    by the full 36-column tables of tests/test_statistics_of_labelled_pixels.py:5-252
    and the CLIJ2 CSV of tests/test_statistics_of_background_and_labelled_pixels.py:221-286
    of the reference, and by float64 brute-force per-label reductions.
    """
    img = _as3d(np.asarray(intensity, np.float32))
    lab = _as3d(np.asarray(labels, np.float32)).astype(np.int64)
    D, H, W = lab.shape
    num_labels = int(lab.max()) + 1
    L = num_labels
    acc = np.zeros((16, H, L), np.float32)
    for p in (8, 10, 12, 14):
        acc[p] = FLT_MAX
    for p in (9, 11, 13, 15):
        acc[p] = -FLT_MAX
    ys = np.arange(H)
    yf = ys.astype(np.float32)
    for z in range(D):
        zf = F32(z)
        for x in range(W):
            xf = F32(x)
            l = lab[z, :, x]
            v = img[z, :, x]
            sel = (l > 0) | (sum_background != 0)
            sel &= l >= 0
            yy, ll, vv, yyf = ys[sel], l[sel], v[sel], yf[sel]
            first = acc[3, yy, ll] == 0
            acc[8, yy, ll] = np.where(first, vv, np.minimum(acc[8, yy, ll], vv))
            acc[9, yy, ll] = np.where(first, vv, np.maximum(acc[9, yy, ll], vv))
            acc[10, yy, ll] = np.where(first, xf, np.minimum(acc[10, yy, ll], xf))
            acc[11, yy, ll] = np.where(first, xf, np.maximum(acc[11, yy, ll], xf))
            acc[12, yy, ll] = np.where(first, yyf, np.minimum(acc[12, yy, ll], yyf))
            acc[13, yy, ll] = np.where(first, yyf, np.maximum(acc[13, yy, ll], yyf))
            acc[14, yy, ll] = np.where(first, zf, np.minimum(acc[14, yy, ll], zf))
            acc[15, yy, ll] = np.where(first, zf, np.maximum(acc[15, yy, ll], zf))
            acc[0, yy, ll] = acc[0, yy, ll] + xf
            acc[1, yy, ll] = acc[1, yy, ll] + yyf
            acc[2, yy, ll] = acc[2, yy, ll] + zf
            acc[3, yy, ll] = acc[3, yy, ll] + F32(1)
            acc[4, yy, ll] = acc[4, yy, ll] + (xf * vv).astype(np.float32)
            acc[5, yy, ll] = acc[5, yy, ll] + (yyf * vv).astype(np.float32)
            acc[6, yy, ll] = acc[6, yy, ll] + (zf * vv).astype(np.float32)
            acc[7, yy, ll] = acc[7, yy, ll] + vv
    sum_pl = np.zeros((16, L), np.float32)
    with np.errstate(over="ignore"):  # the +-FLT_MAX extreme planes are summed too, unused
        for y in range(H):
            sum_pl = (sum_pl + acc[:, y, :]).astype(np.float32)
    min_pl = acc.min(axis=1)
    max_pl = acc.max(axis=1)

    s = slice(1, L)
    r = {}
    r["label"] = np.arange(1, L)
    r["original_label"] = np.arange(1, L)
    bminx, bminy, bminz = min_pl[10, s], min_pl[12, s], min_pl[14, s]
    bmaxx, bmaxy, bmaxz = max_pl[11, s], max_pl[13, s], max_pl[15, s]
    r["bbox_min_x"], r["bbox_min_y"], r["bbox_min_z"] = bminx, bminy, bminz
    r["bbox_max_x"], r["bbox_max_y"], r["bbox_max_z"] = bmaxx, bmaxy, bmaxz
    with np.errstate(all="ignore"):
        r["bbox_width"] = bmaxx - bminx + 1
        r["bbox_height"] = bmaxy - bminy + 1
        r["bbox_depth"] = bmaxz - bminz + 1
        r["min_intensity"] = min_pl[8, s]
        r["max_intensity"] = max_pl[9, s]
        r["sum_intensity"] = sum_pl[7, s]
        area = sum_pl[3, s]
        r["area"] = area
        mean = (sum_pl[7, s] / area).astype(np.float32)
        r["mean_intensity"] = mean
        stat_img = np.zeros((8, L), np.float32)
        stat_img[7, s] = area
        stat_img[6, s] = mean
        names = "xyz"
        for d in range(3):
            r["sum_intensity_times_" + names[d]] = sum_pl[4 + d, s]
            mc = (sum_pl[4 + d, s] / sum_pl[7, s]).astype(np.float32)
            r["mass_center_" + names[d]] = mc
            stat_img[3 + d, s] = mc
        for d in range(3):
            r["sum_" + names[d]] = sum_pl[d, s]
            cen = (sum_pl[d, s] / area).astype(np.float32)
            r["centroid_" + names[d]] = cen
            stat_img[d, s] = cen

    # second pass
    acc2 = np.zeros((6, H, L), np.float32)
    for z in range(D):
        zf = F32(z)
        for x in range(W):
            xf = F32(x)
            l = lab[z, :, x]
            v = img[z, :, x]
            sel = ((l > 0) | (sum_background != 0)) & (l >= 0)
            yy, ll, vv, yyf = ys[sel], l[sel], v[sel], yf[sel]
            with np.errstate(all="ignore"):
                dx = (xf - stat_img[0, ll]).astype(np.float32)
                dy = (yyf - stat_img[1, ll]).astype(np.float32)
                dz = (zf - stat_img[2, ll]).astype(np.float32)
                cd = np.sqrt((dx * dx + dy * dy).astype(np.float32) + dz * dz).astype(np.float32)
                mx_ = (xf - stat_img[3, ll]).astype(np.float32)
                my_ = (yyf - stat_img[4, ll]).astype(np.float32)
                mz_ = (zf - stat_img[5, ll]).astype(np.float32)
                md = np.sqrt((mx_ * mx_ + my_ * my_).astype(np.float32) + mz_ * mz_).astype(np.float32)
                dv = (vv - stat_img[6, ll]).astype(np.float32)
                sq = ((dv * dv).astype(np.float32) / stat_img[7, ll]).astype(np.float32)
            first = acc2[3, yy, ll] == 0
            acc2[4, yy, ll] = np.where(first, cd, np.maximum(acc2[4, yy, ll], cd))
            acc2[5, yy, ll] = np.where(first, md, np.maximum(acc2[5, yy, ll], md))
            acc2[0, yy, ll] = acc2[0, yy, ll] + cd
            acc2[1, yy, ll] = acc2[1, yy, ll] + md
            acc2[2, yy, ll] = acc2[2, yy, ll] + sq
            acc2[3, yy, ll] = acc2[3, yy, ll] + F32(1)
    sum2 = np.zeros((6, L), np.float32)
    for y in range(H):
        sum2 = (sum2 + acc2[:, y, :]).astype(np.float32)
    max2 = acc2.max(axis=1)
    with np.errstate(all="ignore"):
        r["sum_distance_to_centroid"] = sum2[0, s]
        r["mean_distance_to_centroid"] = (sum2[0, s] / area).astype(np.float32)
        r["sum_distance_to_mass_center"] = sum2[1, s]
        r["mean_distance_to_mass_center"] = (sum2[1, s] / area).astype(np.float32)
        r["standard_deviation_intensity"] = np.sqrt(sum2[2, s]).astype(np.float32)
        r["max_distance_to_centroid"] = max2[4, s]
        r["max_distance_to_mass_center"] = max2[5, s]
        r["mean_max_distance_to_centroid_ratio"] = (r["max_distance_to_centroid"] / r["mean_distance_to_centroid"]).astype(np.float32)
        r["mean_max_distance_to_mass_center_ratio"] = (r["max_distance_to_mass_center"] / r["mean_distance_to_mass_center"]).astype(np.float32)
    return r


def statistics_of_background_and_labelled_pixels(intensity: np.ndarray, labels: np.ndarray) -> dict:
    """Statistics over labels + 1 (background becomes label 1), original_label = label - 1.

    Reference: pyclesperanto_prototype/_tier9/_statistics_of_background_and_labelled_pixels.py.
    """
    lab = (np.asarray(labels, np.float32) + F32(1)).astype(np.float32)
    r = statistics_of_labelled_pixels(intensity, lab)
    r["original_label"] = r["label"] - 1
    return r
