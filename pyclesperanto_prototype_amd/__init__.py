"""MI355X-native drop-in for the segmentation hot path of pyclesperanto_prototype.

    import pyclesperanto_prototype_amd as cle
    labels = cle.voronoi_otsu_labeling(image, spot_sigma=2, outline_sigma=2)
    stats = cle.statistics_of_labelled_pixels(image, labels)

The functions keep the reference's names, signatures, defaults, dtypes and returned
dict keys; the computation runs in hand-written gfx950 kernels (libcle_hip.so)
reached through a C ABI (include/cle_abi.h).
"""
from ._array import (Array, asarray, create, create_binary_like, create_labels_like, create_like, create_none,
                     create_zyx, pull, pull_zyx, push, push_zyx)
from ._plugin import Image, is_image, plugin_function
from ._ops import (STATISTICS_KEYS, add_image_and_scalar, autotune_gaussian_blur, add_images, add_images_weighted, binary_and,
                   binary_intersection, close_index_gaps_in_label_map, connected_components_labeling_box,
                   extend_labeling_via_voronoi,
                   connected_components_labeling_diamond, copy, detect_maxima_box, dilate_labels, execute,
                   extend_labels_with_maximum_radius, gauss_otsu_labeling,
                   gaussian_blur, greater_constant, growth_counters, histogram, label, label_spots, logical_and, mask,
                   masked_voronoi_labeling, maximum_of_all_pixels, mean_box, minimum_of_all_pixels, release_pool,
                   nonzero_minimum_box, nonzero_minimum_diamond, onlyzero_overwrite_maximum_box,
                   onlyzero_overwrite_maximum_diamond, otsu_threshold, relabel_sequential, set, set_autotune,
                   set_fp32_key_compat, set_growth_method, set_non_zero_pixels_to_pixel_index, set_nonzero_pixels_to_pixelindex,
                   statistics_of_background_and_labelled_pixels, statistics_of_labeled_pixels,
                   statistics_of_labelled_pixels, statistics_table, sum_of_all_pixels, threshold_otsu,
                   voronoi_labeling, voronoi_otsu_labeling, voronoi_otsu_labeling_batch,
                   voronoi_otsu_labeling_batch_plan, BatchPlan, autotune_voronoi_otsu_labeling_batch,
                   gaussian_blur_batch)
from .io import imread, imread_stack, push_regionprops, push_regionprops_column
from ._lib import load as _load_library

__version__ = "0.1.0"


def library_version() -> str:
    return _load_library().cle_version().decode()
