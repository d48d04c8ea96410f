/*
 * cle_abi.h -- C ABI of libcle_hip.so, the MI355X (gfx950) implementation of the
 * pyclesperanto_prototype segmentation hot path.
 *
 * Conventions
 *   - Every image is a cle_buf: a device pointer to a contiguous C-order array
 *     (z, y, x) with x fastest, its element type and its shape.  2D images have
 *     ndim 2 and shape {H, W}; 1D vectors ndim 1 and shape {N}.
 *   - The caller owns every cle_buf; the library never frees or keeps them after
 *     the call returns.  Scratch memory comes from a stream-ordered pool owned by
 *     the library.
 *   - Every call enqueues work on `stream` (a hipStream_t, NULL = default stream)
 *     and returns without waiting, unless its comment says it synchronises.
 *   - Return value: 0 on success, a negative CLE_E* code on error; the message of
 *     the last error on the calling thread is available from cle_last_error().
 *   - Float-to-integer stores saturate and truncate toward zero, as the reference's
 *     clij_convert_<t>_sat (pyclesperanto_prototype/_tier0/preamble.cl:505-591).
 *
 * Each entry point names the reference function it replaces (file:line under
 * pyclesperanto_prototype/).
 */
#ifndef CLE_ABI_H
#define CLE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types (numpy names) */
enum {
    CLE_U8 = 1, CLE_I8 = 2, CLE_U16 = 3, CLE_I16 = 4,
    CLE_U32 = 5, CLE_I32 = 6, CLE_F32 = 7, CLE_U64 = 8, CLE_I64 = 9, CLE_F64 = 10
};

/* error codes; the Python host maps TYPE -> TypeError, SHAPE/VALUE -> ValueError,
 * OOM -> MemoryError, HIP -> RuntimeError (reference: TypeError for unsupported
 * types, _tier0/_opencl_execute.py:158,290-294). */
enum {
    CLE_OK = 0, CLE_E_TYPE = -1, CLE_E_SHAPE = -2, CLE_E_VALUE = -3,
    CLE_E_OOM = -4, CLE_E_HIP = -5, CLE_E_INTERNAL = -6
};

typedef struct cle_buf {
    void*   data;        /* device pointer */
    int32_t dtype;       /* CLE_U8 ... */
    int32_t ndim;        /* 1, 2 or 3 */
    int64_t shape[3];    /* outermost first: (z, y, x) / (y, x) / (n) */
    int64_t strides[3];  /* element strides, outermost first; all 0 = C-contiguous.  The
                            kernels take C-contiguous images: strides that differ from the
                            C-contiguous ones (an axis of extent 1 may have any) give
                            CLE_E_SHAPE. */
} cle_buf;

/* ---- library ---------------------------------------------------------- */
const char* cle_version(void);
/* message of the last failed call on this thread ("" if none) */
const char* cle_last_error(void);
/* free the scratch arenas of the current device (synchronises the device; call it
   while no other call on that device is in flight) */
int cle_release_pool(void);
/* free the scratch arena of (current device, stream) after draining that stream, for a
   stream that is about to be destroyed (no counterpart in the reference) */
int cle_release_stream(void* stream);
/* size the scratch arena of (current device, stream) as one chunk of at least `bytes`
   (drains that stream when it has to reallocate), so that later calls on the stream
   whose scratch fits make no driver allocation; for long-lived worker streams such as
   a server's (no counterpart in the reference) */
int cle_reserve_stream(void* stream, size_t bytes);

/* ---- device memory for callers without a device-array library ----------- */
/* _tier0/_opencl_backend.py:35-38 (empty) / _tier0/_create.py: *out = `bytes` of device
 * memory on the current device (NULL for 0 bytes); free it with cle_free. */
int cle_alloc(size_t bytes, void** out);
int cle_free(void* p);
/* _tier0/_push.py / _tier0/_pull.py (from_array / get): copy `bytes` between host memory
 * and device memory in stream order; both return after the copy finished. */
int cle_upload(void* dst_device, const void* src_host, size_t bytes, void* stream);
int cle_download(void* dst_host, const void* src_device, size_t bytes, void* stream);

/* ---- fills and copies -------------------------------------------------- */
/* _tier1/_set.py:5-33: every element = value (converted to the dst type) */
int cle_set(const cle_buf* dst, float value, void* stream);
/* _tier1/_copy.py:5-37: dst = convert(src) */
int cle_copy(const cle_buf* src, const cle_buf* dst, void* stream);

/* ---- filters ------------------------------------------------------------ */
/* Autotuner of the isotropic blur's z chunking (blocks per launch) and register budget;
 * no counterpart in the reference.  cle_autotune_gaussian_blur runs the blur of these
 * arguments once per candidate (timed with events on `stream`, which it synchronises) and keeps
 * the fastest for its (half-width, dimensionality, source type, size class) for the
 * process; later cle_gaussian_blur calls of that class use it and never time anything.
 * dst holds the blur afterwards.  Results do not depend on the choice.
 * cle_set_autotune(0) (or CLE_AUTOTUNE=0) ignores and forgets the choices. */
int cle_autotune_gaussian_blur(const cle_buf* src, const cle_buf* dst,
                               float sigma_x, float sigma_y, float sigma_z, void* stream);
int cle_set_autotune(int on);
/* _tier1/_gaussian_blur.py:6-58 for many images at once (the batched pipeline's blur
 * stage): dsts[i] (float32, srcs[i]'s shape, 2D or 3D) = gaussian_blur(srcs[i],
 * sigmas[i] on every axis).  Isotropic kernels up to 41 taps run as one launch per
 * (half-width, dimensionality, source type) group, the rest image by image.  Returns
 * after the work has finished.  The autotune form also times every group's launch
 * shapes (see cle_autotune_voronoi_otsu_labeling_batch); results are the same. */
int cle_gaussian_blur_batch(const cle_buf* srcs, const cle_buf* dsts, const float* sigmas, int n_images,
                            void* stream);
int cle_autotune_gaussian_blur_batch(const cle_buf* srcs, const cle_buf* dsts, const float* sigmas, int n_images,
                                     void* stream);

/* _tier1/_gaussian_blur.py:6-58 + _tier1/_execute_separable_kernel.py:7-75:
 * separable normalised Gaussian, N(s) = floor(8 s) | 1 taps, clamp-to-edge,
 * x then y then z; sigma <= 0 leaves that axis; sigma_z ignored for 2D. */
int cle_gaussian_blur(const cle_buf* src, const cle_buf* dst,
                      float sigma_x, float sigma_y, float sigma_z, void* stream);
/* _tier1/_mean_box.py:6-54: separable box mean of radius r per axis */
int cle_mean_box(const cle_buf* src, const cle_buf* dst,
                 int radius_x, int radius_y, int radius_z, void* stream);
/* _tier1/_detect_maxima_box.py:5-48 + detect_maxima_{2d,3d}_x.cl: mean_box(r)
 * then the 3^d local-maximum flag with the reference's visiting order */
int cle_detect_maxima_box(const cle_buf* src, const cle_buf* dst,
                          int radius_x, int radius_y, int radius_z, void* stream);

/* ---- reductions, histogram, Otsu --------------------------------------- */
/* _tier2/_minimum_of_all_pixels.py:6-59 and _maximum_of_all_pixels.py:6-59:
 * out_minmax is a device float[2] = {min, max} */
int cle_minimum_maximum(const cle_buf* src, float* out_minmax, void* stream);
/* _tier2/_sum_of_all_pixels.py: device double[1] */
int cle_sum_of_all_pixels(const cle_buf* src, double* out_sum, void* stream);
/* _tier3/_histogram.py:25-130: dst is a float32 vector of num_bins.  If
 * minmax_dev is NULL the range is [mn, mx]; otherwise it is read from the device
 * float[2] (as produced by cle_minimum_maximum). */
int cle_histogram(const cle_buf* src, const cle_buf* dst, int num_bins,
                  float mn, float mx, const float* minmax_dev, void* stream);
/* _tier9/_threshold_otsu.py:11-56: dst = src > otsu(src).  If out_threshold is
 * not NULL the threshold (float64, as the reference computes it) is written to
 * that device address. */
int cle_threshold_otsu(const cle_buf* src, const cle_buf* dst, double* out_threshold, void* stream);

/* ---- elementwise -------------------------------------------------------- */
/* _tier1/_greater_constant.py:5-40 */
int cle_greater_constant(const cle_buf* src, const cle_buf* dst, float constant, void* stream);
/* _tier1/_binary_and.py:6-47 */
int cle_binary_and(const cle_buf* a, const cle_buf* b, const cle_buf* dst, void* stream);
/* _tier1/_mask.py:5-45 */
int cle_mask(const cle_buf* src, const cle_buf* mask, const cle_buf* dst, void* stream);
/* _tier1/_add_image_and_scalar.py:5-43 */
int cle_add_image_and_scalar(const cle_buf* src, const cle_buf* dst, float scalar, void* stream);
/* _tier1/_add_images_weighted.py:5-50 */
int cle_add_images_weighted(const cle_buf* a, const cle_buf* b, const cle_buf* dst,
                            float factor_a, float factor_b, void* stream);

/* ---- labelling ---------------------------------------------------------- */
/* _tier1/_set_non_zero_pixels_to_pixel_index.py:5-28: key x*H*D + y*D + z + offset */
int cle_set_nonzero_pixels_to_pixel_index(const cle_buf* src, const cle_buf* dst,
                                          int offset, void* stream);
/* _tier1/_nonzero_minimum_box.py:5-43 / _nonzero_minimum_diamond.py: one Jacobi
 * step; flag (any 1-element buffer) is set to 1 if a value changed. */
int cle_nonzero_minimum_box(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream);
int cle_nonzero_minimum_diamond(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream);
/* _tier1/_onlyzero_overwrite_maximum_box.py:5-38 / _diamond.py:5-38 */
int cle_onlyzero_overwrite_maximum_box(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream);
int cle_onlyzero_overwrite_maximum_diamond(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream);

/* connectivity for cle_connected_components_labeling */
enum { CLE_CONN_BOX = 0, CLE_CONN_DIAMOND = 1 };
/* option bits */
enum {
    /* give components whose float32-rounded minimum key is equal the same label,
     * as the reference's float32 key images do above 2^24 voxels (default on) */
    CLE_FP32_KEY_COMPAT = 1
};
/* _tier4/_connected_components_labeling_box.py:15-66 (and the diamond variant,
 * _tier5/_connected_components_labeling_diamond.py:8-34): label k is the k-th
 * component ordered by its minimum key x*H*D + y*D + z.  Union-find, no host
 * synchronisation.  If n_labels_dev is not NULL the label count (uint32) is
 * written there. */
int cle_connected_components_labeling(const cle_buf* binary, const cle_buf* dst, int connectivity,
                                      int options, uint32_t* n_labels_dev, void* stream);
/* _tier3/_relabel_sequential.py:12-59: rank of each non-zero value among the
 * values present.  Synchronises once (to size the value table, like the reference's
 * maximum_of_all_pixels). */
int cle_relabel_sequential(const cle_buf* src, const cle_buf* dst, void* stream);
/* _tier2/_label_spots.py:9-56 + label_spots_in_x.cl: non-zero voxels numbered
 * 1.. in raster (z, y, x) order */
int cle_label_spots(const cle_buf* src, const cle_buf* dst, void* stream);
/* _tier5/_masked_voronoi_labeling.py:7-76: seeds labelled by CCL box, grown inside
 * the mask by alternating box / diamond steps until no change.  Synchronises once per
 * batch of growth steps (see cle_set_growth_method).  If n_iterations is not NULL the
 * reference's iteration count is stored there (host). */
int cle_masked_voronoi_labeling(const cle_buf* binary, const cle_buf* mask, const cle_buf* dst,
                                int* n_iterations, void* stream);
/* _tier4/_extend_labeling_via_voronoi.py:12-61: the labels (converted to float32)
 * grown into the zero voxels by alternating box / diamond steps until no change,
 * then stored in dst's type.  No mask.  n_iterations as above. */
int cle_extend_labeling_via_voronoi(const cle_buf* src, const cle_buf* dst, int* n_iterations, void* stream);
/* Growth method of the Voronoi ops (no counterpart in the reference; results do not
 * depend on it): 1 (default) image-wide pairs of steps, 0 asynchronous exact tile solves
 * (which fall back to the pairs for growths of 254 steps or more; a captured batch plan
 * always records the pairs). */
int cle_set_growth_method(int method);
/* Counters of the pair-of-steps growth launches made by the calls that synchronise on
 * them (no counterpart in the reference; process-wide): out[0] launches, out[1] growth
 * tiles processed (1024 voxels each: 16 x 8 x 8 in 3D, 32 x 32 in 2D), out[2]
 * nanoseconds of GPU time of those launches (events around each batch of launches).
 * reset != 0 sets them to zero after reading. */
int cle_growth_counters(unsigned long long* out, int reset);
/* _tier9/_voronoi_otsu_labeling.py:11-73: blur(spot) -> maxima -> blur(outline) ->
 * Otsu -> AND -> masked Voronoi -> mask.  Synchronises only inside the Voronoi
 * convergence test. */
int cle_voronoi_otsu_labeling(const cle_buf* src, const cle_buf* dst,
                              float spot_sigma, float outline_sigma, void* stream);
/* cle_voronoi_otsu_labeling without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under stream capture, tight
 * serving loops, harnesses that forbid it.  Same kernels and result; its scratch is a
 * fixed device array of CLE_NOALLOC_SCRATCH_BYTES per GPU (held from the library's
 * load until the process ends), which bounds the image to about 4.5 million voxels in
 * 2D and 5.5 million in 3D (roughly 15 / 12 bytes of scratch per voxel); a larger
 * image fails with CLE_E_VALUE before anything runs.  All calls on one GPU share that array: calls that overlap on
 * one GPU (other threads, other streams) race on it.  The Voronoi growth always runs
 * as pairs of image-wide steps here (cle_set_growth_method does not apply). */
#define CLE_NOALLOC_SCRATCH_BYTES (64u << 20)
int cle_voronoi_otsu_labeling_noalloc(const cle_buf* src, const cle_buf* dst,
                                      float spot_sigma, float outline_sigma, void* stream);
/* _tier9/_gauss_otsu_labeling.py:6-49: blur -> Otsu -> CCL box */
int cle_gauss_otsu_labeling(const cle_buf* src, const cle_buf* dst, float outline_sigma, void* stream);
/* Batched voronoi_otsu_labeling (_tier9/_voronoi_otsu_labeling.py:11-73 applied to
 * every image): srcs[i] -> dsts[i] with spot_sigmas[i] / outline_sigmas[i], images of
 * any sizes and dimensionalities in one call.  Each stage is one launch over the
 * blocks of all images (the blur one launch per kernel half-width), the Otsu
 * thresholds never leave the device, and the Voronoi growth of all images advances
 * together.  Every dsts[i] equals what cle_voronoi_otsu_labeling gives for srcs[i]
 * alone.  *n_steps (if not NULL) receives the number of growth steps the slowest
 * image needed.  Returns after the work has finished (it synchronises the stream). */
int cle_voronoi_otsu_labeling_batch(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                    const float* outline_sigmas, int n_images, int* n_steps, void* stream);
/* Autotuner of the batched blur (no counterpart in the reference): runs
 * cle_voronoi_otsu_labeling_batch on these arguments, timing for every blur group
 * (images of one half-width, dimensionality and source type) the candidate launch
 * shapes -- tile height 32 or 16 rows, register budget, blocks the z chunking aims at --
 * and keeps the fastest per (half-width, dimensionality, source type, log2 voxels,
 * log2 images) for the process; later batches of that class use it and time nothing.
 * dsts hold the labels afterwards.  Results do not depend on the choices. */
int cle_autotune_voronoi_otsu_labeling_batch(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                             const float* outline_sigmas, int n_images, void* stream);

/* Replayable plan of cle_voronoi_otsu_labeling_batch for these exact buffers and
 * sigmas (no counterpart in the reference: the batched executor as a captured
 * hipGraph).  Building it runs the batch twice on a private stream (after the work
 * already queued on `stream`; dsts hold the result), then captures every stage and
 * graph_pairs pairs of growth steps into one graph; the plan owns the device memory
 * the graph uses until cle_plan_destroy.  cle_plan_run replays the graph on `stream`
 * with the sources' current contents and, if some image still grows after
 * graph_pairs pairs, finishes the growth eagerly; it returns when the labels are in
 * dsts (it synchronises the stream).  Runs of one plan must not overlap. */
typedef struct cle_plan cle_plan;
int cle_voronoi_otsu_labeling_batch_plan(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                         const float* outline_sigmas, int n_images, int graph_pairs,
                                         cle_plan** out_plan, void* stream);
int cle_plan_run(cle_plan* plan, int* n_steps, void* stream);
int cle_plan_destroy(cle_plan* plan);

/* ---- statistics --------------------------------------------------------- */
/* number of columns of the statistics table; column c holds STATISTICS_ENTRY c
 * (_tier9/_statistics_entry.py:3-46): IDENTIFIER, BOUNDING_BOX_X, ... ,
 * MAX_MEAN_DISTANCE_TO_CENTROID_RATIO */
#define CLE_STAT_COLUMNS 36
/* _tier9/_statistics_of_labelled_pixels.py:7-293 + statistics_per_label_x.cl +
 * standard_deviation_per_label_x.cl.  num_labels = max(labels) + 1 as the reference
 * computes it (_statistics_of_labelled_pixels.py:53), or 0: the library derives it
 * from the label image in the same pass that finds the label runs.  out_table is a
 * device float32 buffer of shape (CLE_STAT_COLUMNS, columns), columns >= num_labels -
 * 1; column j describes label j + 1.  *out_num_labels (if not NULL) receives
 * num_labels; with num_labels = 0 and a table that is too small the call fails with
 * CLE_E_SHAPE after setting it, so the caller can grow the table and call again.
 * Label 0 (background) and negative labels are not measured, as in the reference
 * (sum_background = 0).  Synchronises the stream once (to size the run buffers). */
int cle_statistics_of_labelled_pixels(const cle_buf* intensity, const cle_buf* labels,
                                      int64_t num_labels, const cle_buf* out_table,
                                      int64_t* out_num_labels, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CLE_ABI_H */
