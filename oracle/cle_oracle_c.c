/*
 * CPU oracle in C -- TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
 *
 * A C/OpenMP restatement of the reference algorithms of pyclesperanto_prototype
 * 0.24.5 for the segmentation hot path, in the reference's own structure (it is not
 * the reference's code, which is Python + OpenCL and cannot run here):
 *   - separable Gaussian passes through float32 temporaries, tap by tap
 *     (_tier1/_execute_separable_kernel.py:7-75, _tier0/_sigma_to_kernel_size.py:1-4)
 *   - 3^d maxima with the visiting order of _tier1/detect_maxima_{2d,3d}_x.cl
 *   - one partial histogram per row y, summed over y in float32
 *     (_tier3/_histogram.py:85-128), Otsu in float64 (_tier9/_threshold_otsu.py:64-106)
 *   - connected components by Jacobi non-zero-minimum propagation of float32 pixel
 *     keys to a fixed point, then relabel_sequential
 *     (_tier4/_connected_components_labeling_box.py:39-64)
 *   - masked Voronoi by alternating box / diamond growth to a fixed point
 *     (_tier5/_masked_voronoi_labeling.py:44-74)
 *   - statistics by per-(row, label) float32 cells, run-length accumulation along x,
 *     then float32 sums over y (_tier9/statistics_per_label_x.cl,
 *     standard_deviation_per_label_x.cl, _statistics_of_labelled_pixels.py:43-291)
 * It is pinned against the numpy oracle (oracle/cle_oracle.py) by
 * tests/test_oracle_c.py.  Only tests and the benchmark's CPU-baseline leg use it.
 *
 * Arrays are C-order (z, y, x); 2D images are D = 1, ndim = 2.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

static inline i64 clampi(i64 v, i64 hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

/* ------------------------------------------------------------------ blur */
static int kernel_size(float sigma) {
    int n = (int)(sigma * 8.0f);
    if (n % 2 == 0) n += 1;
    return n;
}

static void gauss_axis(const float* in, float* out, i64 D, i64 H, i64 W, int axis, float sigma) {
    const int n = kernel_size(sigma), c = (n - 1) / 2;
    const float nrm = -2.0f * sigma * sigma;
    float* h = (float*)malloc(sizeof(float) * (size_t)n);
    float hsum = 0.f;
    for (int v = -c; v <= c; ++v) {
        h[v + c] = expf((float)(v * v) / nrm);
        hsum += h[v + c];
    }
    const i64 na = axis == 0 ? W : (axis == 1 ? H : D);
#pragma omp parallel for collapse(2) schedule(static)
    for (i64 z = 0; z < D; ++z)
        for (i64 y = 0; y < H; ++y)
            for (i64 x = 0; x < W; ++x) {
                const i64 p = axis == 0 ? x : (axis == 1 ? y : z);
                float res = 0.f;
                for (int v = -c; v <= c; ++v) {
                    const i64 q = clampi(p + v, na - 1);
                    const i64 idx = axis == 0 ? (z * H + y) * W + q : (axis == 1 ? (z * H + q) * W + x : (q * H + y) * W + x);
                    res += h[v + c] * in[idx];
                }
                out[(z * H + y) * W + x] = res / hsum;
            }
    free(h);
}

void ora_gaussian_blur(const float* src, float* dst, i64 D, i64 H, i64 W, int ndim, float sx, float sy, float sz) {
    const i64 n = D * H * W;
    float* a = (float*)malloc(sizeof(float) * (size_t)n);
    float* b = (float*)malloc(sizeof(float) * (size_t)n);
    memcpy(a, src, sizeof(float) * (size_t)n);
    const float sig[3] = {sx, sy, sz};
    for (int d = 0; d < ndim; ++d) {
        if (sig[d] > 0) {
            gauss_axis(a, b, D, H, W, d, sig[d]);
            float* t = a; a = b; b = t;
        }
    }
    memcpy(dst, a, sizeof(float) * (size_t)n);
    free(a);
    free(b);
}

/* ------------------------------------------------------------------ maxima */
void ora_detect_maxima(const float* src, uint8_t* dst, i64 D, i64 H, i64 W, int ndim) {
    const int zr = ndim == 3 ? 1 : 0;
#pragma omp parallel for collapse(2) schedule(static)
    for (i64 z = 0; z < D; ++z)
        for (i64 y = 0; y < H; ++y)
            for (i64 x = 0; x < W; ++x) {
                float lm = src[(z * H + y) * W + x] - 1.0f;
                int self = 1;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -zr; dz <= zr; ++dz) {
                            const i64 qx = x + dx, qy = y + dy, qz = z + dz;
                            if (qx < 0 || qy < 0 || qz < 0) continue;
                            const float v = src[(clampi(qz, D - 1) * H + clampi(qy, H - 1)) * W + clampi(qx, W - 1)];
                            if (v > lm) {
                                lm = v;
                                self = (dx == 0 && dy == 0 && dz == 0);
                            }
                        }
                dst[(z * H + y) * W + x] = (uint8_t)self;
            }
}

/* ------------------------------------------------------------------ Otsu */
void ora_minmax(const float* src, i64 n, float* mn, float* mx) {
    float a = INFINITY, b = -INFINITY;
#pragma omp parallel for reduction(min : a) reduction(max : b)
    for (i64 i = 0; i < n; ++i) {
        if (src[i] < a) a = src[i];
        if (src[i] > b) b = src[i];
    }
    *mn = a;
    *mx = b;
}

static inline uint32_t bin_of(float v, float mn, float range, int bins) {
    const float q = ((v - mn) * (float)(bins - 1)) / range;
    const double t = (double)q + 0.5;
    if (!(t > 0.0)) return 0u;
    if (t >= (double)bins) return (uint32_t)(bins - 1);
    return (uint32_t)t;
}

void ora_histogram(const float* src, i64 D, i64 H, i64 W, int bins, float mn, float mx, float* hist) {
    const float range = mx - mn;
    float* partial = (float*)calloc((size_t)(H * bins), sizeof(float));
#pragma omp parallel for schedule(static)
    for (i64 y = 0; y < H; ++y) {
        uint32_t* tmp = (uint32_t*)calloc((size_t)bins, sizeof(uint32_t));
        for (i64 z = 0; z < D; ++z)
            for (i64 x = 0; x < W; ++x) tmp[bin_of(src[(z * H + y) * W + x], mn, range, bins)]++;
        for (int b = 0; b < bins; ++b) partial[y * bins + b] = (float)tmp[b];
        free(tmp);
    }
    for (int b = 0; b < bins; ++b) {
        float acc = 0.f;
        for (i64 y = 0; y < H; ++y) acc += partial[y * bins + b];
        hist[b] = acc;
    }
    free(partial);
}

double ora_otsu_threshold(const float* hist, float mn, float mx) {
    double bc[256], w1[256], w2[256], m1[256], m2[256];
    const float rng = mx - mn;
    for (int i = 0; i < 256; ++i) bc[i] = (double)i * (double)rng / 255.0 + (double)mn;
    double a = 0, s = 0;
    for (int i = 0; i < 256; ++i) {
        a += (double)hist[i];
        s += (double)hist[i] * bc[i];
        w1[i] = a;
        m1[i] = s / a;
    }
    a = 0;
    s = 0;
    for (int i = 255; i >= 0; --i) {
        a += (double)hist[i];
        s += (double)hist[i] * bc[i];
        w2[i] = a;
        m2[i] = s / a;
    }
    int best = 0;
    double bv = 0;
    for (int i = 0; i < 255; ++i) {
        const double d = m1[i] - m2[i + 1];
        const double v = (w1[i] * w2[i + 1]) * (d * d);
        if (v != v) { best = i; break; }
        if (i == 0 || v > bv) { bv = v; best = i; }
    }
    return bc[best];
}

double ora_threshold_otsu(const float* src, i64 D, i64 H, i64 W, uint8_t* dst) {
    const i64 n = D * H * W;
    float mn, mx, hist[256];
    ora_minmax(src, n, &mn, &mx);
    ora_histogram(src, D, H, W, 256, mn, mx, hist);
    const double t = ora_otsu_threshold(hist, mn, mx);
    const float tf = (float)t;
#pragma omp parallel for
    for (i64 i = 0; i < n; ++i) dst[i] = src[i] > tf ? 1 : 0;
    return t;
}

/* ------------------------------------------------------------------ CCL */
static int cmp_float(const void* a, const void* b) {
    const float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* dst = 1-based rank of each non-zero value among the distinct values present */
static void relabel_floats(const float* v, i64 n, float* out) {
    i64 m = 0;
    for (i64 i = 0; i < n; ++i) m += v[i] != 0.f;
    float* u = (float*)malloc(sizeof(float) * (size_t)(m ? m : 1));
    m = 0;
    for (i64 i = 0; i < n; ++i)
        if (v[i] != 0.f) u[m++] = v[i];
    qsort(u, (size_t)m, sizeof(float), cmp_float);
    i64 k = 0;
    for (i64 i = 0; i < m; ++i)
        if (i == 0 || u[i] != u[k - 1]) u[k++] = u[i];
#pragma omp parallel for
    for (i64 i = 0; i < n; ++i) {
        if (v[i] == 0.f) { out[i] = 0.f; continue; }
        i64 lo = 0, hi = k - 1;
        while (lo < hi) {
            const i64 mid = (lo + hi) / 2;
            if (u[mid] < v[i]) lo = mid + 1; else hi = mid;
        }
        out[i] = (float)(lo + 1);
    }
    free(u);
}

static int nonzero_min_step(const float* src, float* dst, i64 D, i64 H, i64 W, int ndim, int diamond) {
    const int zr = ndim == 3 ? 1 : 0;
    int changed = 0;
#pragma omp parallel for collapse(2) reduction(| : changed) schedule(static)
    for (i64 z = 0; z < D; ++z)
        for (i64 y = 0; y < H; ++y)
            for (i64 x = 0; x < W; ++x) {
                const i64 i = (z * H + y) * W + x;
                const float v = src[i];
                if (v == 0.f) continue;
                float best = v;
                for (int dz = -zr; dz <= zr; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (diamond && abs(dx) + abs(dy) + abs(dz) > 1) continue;
                            const float w = src[(clampi(z + dz, D - 1) * H + clampi(y + dy, H - 1)) * W + clampi(x + dx, W - 1)];
                            if (w != 0.f && w < best) best = w;
                        }
                if (best != v) changed = 1;
                dst[i] = best;
            }
    return changed;
}

/* labels (float32) of the components of `bin`; returns the iteration count */
int ora_ccl(const uint8_t* bin, i64 D, i64 H, i64 W, int ndim, int diamond, float* out) {
    const i64 n = D * H * W;
    float* t1 = (float*)malloc(sizeof(float) * (size_t)n);
    float* t2 = (float*)calloc((size_t)n, sizeof(float));
#pragma omp parallel for collapse(2)
    for (i64 z = 0; z < D; ++z)
        for (i64 y = 0; y < H; ++y)
            for (i64 x = 0; x < W; ++x) {
                const i64 i = (z * H + y) * W + x;
                t1[i] = bin[i] ? (float)(x * H * D + y * D + z + 1) : 0.f;
            }
    int it = 0, flag = 1;
    while (flag) {
        flag = (it % 2 == 0) ? nonzero_min_step(t1, t2, D, H, W, ndim, diamond)
                             : nonzero_min_step(t2, t1, D, H, W, ndim, diamond);
        ++it;
    }
    relabel_floats(it % 2 == 0 ? t1 : t2, n, out);
    free(t1);
    free(t2);
    return it;
}

/* ------------------------------------------------------------------ Voronoi */
static int onlyzero_max_step(const float* src, float* dst, i64 D, i64 H, i64 W, int ndim, int diamond) {
    const int zr = ndim == 3 ? 1 : 0;
    int changed = 0;
#pragma omp parallel for collapse(2) reduction(| : changed) schedule(static)
    for (i64 z = 0; z < D; ++z)
        for (i64 y = 0; y < H; ++y)
            for (i64 x = 0; x < W; ++x) {
                const i64 i = (z * H + y) * W + x;
                const float v = src[i];
                if (v != 0.f) { dst[i] = v; continue; }
                float best = 0.f;
                for (int dz = -zr; dz <= zr; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (dx == 0 && dy == 0 && dz == 0) continue;
                            if (diamond && abs(dx) + abs(dy) + abs(dz) > 1) continue;
                            const float w = src[(clampi(z + dz, D - 1) * H + clampi(y + dy, H - 1)) * W + clampi(x + dx, W - 1)];
                            if (w > best) best = w;
                        }
                if (best != v) changed = 1;
                dst[i] = best;
            }
    return changed;
}

int ora_masked_voronoi(const uint8_t* binary, const uint8_t* mask, i64 D, i64 H, i64 W, int ndim, uint32_t* out) {
    const i64 n = D * H * W;
    float* flip = (float*)malloc(sizeof(float) * (size_t)n);
    float* flop = (float*)malloc(sizeof(float) * (size_t)n);
    ora_ccl(binary, D, H, W, ndim, 0, flop);
#pragma omp parallel for
    for (i64 i = 0; i < n; ++i) flip[i] = ((float)mask[i] - 1.0f) + flop[i];
    int it = 0, flag = 1;
    while (flag) {
        flag = (it % 2 == 0) ? onlyzero_max_step(flip, flop, D, H, W, ndim, 0)
                             : onlyzero_max_step(flop, flip, D, H, W, ndim, 1);
        ++it;
    }
    const float* last = it % 2 == 0 ? flip : flop;
#pragma omp parallel for
    for (i64 i = 0; i < n; ++i) {
        const float v = mask[i] ? last[i] : 0.f;
        out[i] = v > 4294967295.0f ? 4294967295u : (v < 0.f ? 0u : (uint32_t)v);
    }
    free(flip);
    free(flop);
    return it;
}

void ora_voronoi_otsu_labeling(const float* src, uint32_t* out, i64 D, i64 H, i64 W, int ndim, float ss, float so) {
    const i64 n = D * H * W;
    float* temp = (float*)malloc(sizeof(float) * (size_t)n);
    uint8_t* spots = (uint8_t*)malloc((size_t)n);
    uint8_t* seg = (uint8_t*)malloc((size_t)n);
    ora_gaussian_blur(src, temp, D, H, W, ndim, ss, ss, ss);
    ora_detect_maxima(temp, spots, D, H, W, ndim);
    ora_gaussian_blur(src, temp, D, H, W, ndim, so, so, so);
    ora_threshold_otsu(temp, D, H, W, seg);
#pragma omp parallel for
    for (i64 i = 0; i < n; ++i) spots[i] = spots[i] && seg[i];
    ora_masked_voronoi(spots, seg, D, H, W, ndim, out);
    free(temp);
    free(spots);
    free(seg);
}

/* ------------------------------------------------------------------ statistics */
#define NCOL 36

/* table: float32 (36, L-1), STATISTICS_ENTRY column order */
void ora_statistics(const float* img, const uint32_t* lab, i64 D, i64 H, i64 W, i64 L, float* table) {
    const i64 cells = H * L;
    float* acc = (float*)calloc((size_t)(16 * cells), sizeof(float));
    for (int p = 8; p < 16; ++p)
        for (i64 c = 0; c < cells; ++c) acc[p * cells + c] = (p % 2 == 0) ? FLT_MAX : -FLT_MAX;
    for (i64 z = 0; z < D; ++z) {
#pragma omp parallel for schedule(static)
        for (i64 y = 0; y < H; ++y) {
            for (i64 x = 0; x < W; ++x) {
                const i64 l = (i64)(float)lab[(z * H + y) * W + x];
                if (l <= 0 || l >= L) continue;
                const float v = img[(z * H + y) * W + x];
                float* a = acc + y * L + l;
#define P(k) a[(k) * cells]
                const float xf = (float)x, yf = (float)y, zf = (float)z;
                if (P(3) == 0.f) {
                    P(8) = P(9) = v; P(10) = P(11) = xf; P(12) = P(13) = yf; P(14) = P(15) = zf;
                } else {
                    if (P(8) > v) P(8) = v;
                    if (P(9) < v) P(9) = v;
                    if (P(10) > xf) P(10) = xf;
                    if (P(11) < xf) P(11) = xf;
                    if (P(12) > yf) P(12) = yf;
                    if (P(13) < yf) P(13) = yf;
                    if (P(14) > zf) P(14) = zf;
                    if (P(15) < zf) P(15) = zf;
                }
                P(0) += xf; P(1) += yf; P(2) += zf; P(3) += 1.f;
                P(4) += xf * v; P(5) += yf * v; P(6) += zf * v; P(7) += v;
#undef P
            }
        }
    }
    float* sum = (float*)calloc((size_t)(16 * L), sizeof(float));
    float* mn = (float*)malloc(sizeof(float) * (size_t)(16 * L));
    float* mxv = (float*)malloc(sizeof(float) * (size_t)(16 * L));
    for (int p = 0; p < 16; ++p)
        for (i64 l = 0; l < L; ++l) {
            float s = 0.f, a = FLT_MAX, b = -FLT_MAX;
            for (i64 y = 0; y < H; ++y) {
                const float v = acc[p * cells + y * L + l];
                s += v;
                if (v < a) a = v;
                if (v > b) b = v;
            }
            sum[p * L + l] = s;
            mn[p * L + l] = a;
            mxv[p * L + l] = b;
        }
    /* per-label float32 results of pass 1 */
    float* st = (float*)calloc((size_t)(8 * L), sizeof(float));
    for (i64 l = 1; l < L; ++l) {
        const float area = sum[3 * L + l], sv = sum[7 * L + l];
        st[7 * L + l] = area;
        st[6 * L + l] = sv / area;
        for (int d = 0; d < 3; ++d) {
            st[(3 + d) * L + l] = sum[(4 + d) * L + l] / sv;
            st[d * L + l] = sum[d * L + l] / area;
        }
    }
    /* pass 2 */
    float* acc2 = (float*)calloc((size_t)(6 * cells), sizeof(float));
    for (i64 z = 0; z < D; ++z) {
#pragma omp parallel for schedule(static)
        for (i64 y = 0; y < H; ++y) {
            for (i64 x = 0; x < W; ++x) {
                const i64 l = (i64)(float)lab[(z * H + y) * W + x];
                if (l <= 0 || l >= L) continue;
                const float v = img[(z * H + y) * W + x];
                const float xf = (float)x, yf = (float)y, zf = (float)z;
                const float dx = xf - st[0 * L + l], dy = yf - st[1 * L + l], dz = zf - st[2 * L + l];
                const float cd = sqrtf(dx * dx + dy * dy + dz * dz);
                const float ex = xf - st[3 * L + l], ey = yf - st[4 * L + l], ez = zf - st[5 * L + l];
                const float md = sqrtf(ex * ex + ey * ey + ez * ez);
                const float dv = v - st[6 * L + l];
                const float sq = (dv * dv) / st[7 * L + l];
                float* a = acc2 + y * L + l;
#define Q(k) a[(k) * cells]
                if (Q(3) == 0.f) { Q(4) = cd; Q(5) = md; }
                else { if (Q(4) < cd) Q(4) = cd; if (Q(5) < md) Q(5) = md; }
                Q(0) += cd; Q(1) += md; Q(2) += sq; Q(3) += 1.f;
#undef Q
            }
        }
    }
    const i64 N = L - 1;
    for (i64 l = 1; l < L; ++l) {
        float s2[6] = {0, 0, 0, 0, 0, 0}, m2[6] = {-FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (i64 y = 0; y < H; ++y)
            for (int k = 0; k < 6; ++k) {
                const float v = acc2[k * cells + y * L + l];
                s2[k] += v;
                if (v > m2[k]) m2[k] = v;
            }
        const float area = sum[3 * L + l];
        float col[NCOL];
        col[0] = (float)l;
        col[1] = mn[10 * L + l]; col[2] = mn[12 * L + l]; col[3] = mn[14 * L + l];
        col[4] = mxv[11 * L + l]; col[5] = mxv[13 * L + l]; col[6] = mxv[15 * L + l];
        col[7] = col[4] - col[1] + 1.f; col[8] = col[5] - col[2] + 1.f; col[9] = col[6] - col[3] + 1.f;
        col[10] = mn[8 * L + l]; col[11] = mxv[9 * L + l];
        col[12] = st[6 * L + l]; col[13] = sum[7 * L + l];
        col[14] = sqrtf(s2[2]);
        col[15] = area;
        col[16] = sum[4 * L + l]; col[17] = sum[5 * L + l]; col[18] = sum[6 * L + l];
        col[19] = st[3 * L + l]; col[20] = st[4 * L + l]; col[21] = st[5 * L + l];
        col[22] = sum[0 * L + l]; col[23] = sum[1 * L + l]; col[24] = sum[2 * L + l];
        col[25] = st[0 * L + l]; col[26] = st[1 * L + l]; col[27] = st[2 * L + l];
        col[28] = s2[1]; col[29] = s2[1] / area; col[30] = m2[5]; col[31] = m2[5] / col[29];
        col[32] = s2[0]; col[33] = s2[0] / area; col[34] = m2[4]; col[35] = m2[4] / col[33];
        for (int c = 0; c < NCOL; ++c) table[c * N + (l - 1)] = col[c];
    }
    free(acc); free(sum); free(mn); free(mxv); free(st); free(acc2);
}

/* the 3D plain-method pipeline: blur -> Otsu -> CCL box (int32 labels) -> statistics.
 * Returns the number of labels + 1 (the statistics table width + 1). */
i64 ora_blur_otsu_ccl_stats(const float* src, i64 D, i64 H, i64 W, int ndim, float sigma, int32_t* labels_out,
                            float* table, i64 table_capacity) {
    const i64 n = D * H * W;
    float* blurred = (float*)malloc(sizeof(float) * (size_t)n);
    uint8_t* seg = (uint8_t*)malloc((size_t)n);
    float* labf = (float*)malloc(sizeof(float) * (size_t)n);
    uint32_t* lab = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    ora_gaussian_blur(src, blurred, D, H, W, ndim, sigma, sigma, sigma);
    ora_threshold_otsu(blurred, D, H, W, seg);
    ora_ccl(seg, D, H, W, ndim, 0, labf);
    i64 L = 1;
    for (i64 i = 0; i < n; ++i) {
        lab[i] = (uint32_t)labf[i];
        labels_out[i] = (int32_t)labf[i];
        if ((i64)lab[i] + 1 > L) L = (i64)lab[i] + 1;
    }
    if (table && (L - 1) * NCOL <= table_capacity) ora_statistics(src, lab, D, H, W, L, table);
    free(blurred); free(seg); free(labf); free(lab);
    return L;
}
