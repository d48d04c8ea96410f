// statistics_of_labelled_pixels as two segmented reductions.
//
// Reference semantics: pyclesperanto_prototype/_tier9/_statistics_of_labelled_pixels.py:43-291,
// statistics_per_label_x.cl:1-146 (pass 1: count, coordinate sums, intensity sums,
// intensity-weighted coordinate sums, min/max intensity, bounding box) and
// standard_deviation_per_label_x.cl:1-96 (pass 2, from the float32 centroid, mass
// centre, mean and area of pass 1: sums of Euclidean distances, their maxima, and
// the sum of (v - mean)^2 / area).  Column layout of the output table:
// _tier9/_statistics_entry.py:3-46 (STATISTICS_ENTRY).
//
// MI355X design.  Each thread walks a run of RUN consecutive voxels of one row,
// accumulating while the label stays the same and flushing a partial (one label,
// one row segment) with atomics into per-label accumulators: exact integer sums
// (uint64) for counts and coordinates, float64 for intensity sums, ordered-uint32
// atomics for min/max.  The per-voxel float32 arithmetic (x*v products, distances,
// squared deviations) is the reference's; only the summation is carried in higher
// precision, so results differ from the reference's float32 running sums by that
// rounding only.
#include <float.h>

#include "common.hpp"

namespace cle {

constexpr int RUN = 16;

struct Acc {
    unsigned long long *cnt, *sx, *sy, *sz;
    double *sv, *sxv, *syv, *szv;
    uint32_t *minv, *maxv;              // ordered-float encoding
    uint32_t *bminx, *bmaxx, *bminy, *bmaxy, *bminz, *bmaxz;
    double *sdc, *sdm, *ssq;
    uint32_t *maxdc, *maxdm;            // float bits of non-negative values
    float* img;                         // [8][L]: centroid xyz, mass centre xyz, mean, area
};

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(b);
}

// Label readers: the reference reads a label as (int) of the float it loads.
template <typename T>
struct LabT {
    const T* p;
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)(float)p[i]; }
    __device__ __forceinline__ bool nz(int64_t i) const { return p[i] != (T)0; }
};
struct LabAny {
    DBuf b;
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)ld(b, i); }
    __device__ __forceinline__ bool nz(int64_t i) const { return cle::nz(b, i); }
};

// Work split shared by both passes: a wavefront takes a span of SPAN voxels of one
// row; it first checks, with coalesced loads, whether any label in the span is
// non-zero, and skips the span if not.  Otherwise lane l walks voxels
// [x0 + l*RUN, x0 + (l+1)*RUN) of the span.
constexpr int SPAN = 64 * RUN;

template <typename LAB>
__device__ __forceinline__ bool span_has_labels(LAB lab, int64_t base, int64_t x0, int64_t x1, int lane) {
    bool any = false;
#pragma unroll 4
    for (int k = 0; k < RUN; ++k) {
        const int64_t x = x0 + lane + 64 * k;
        if (x < x1) any |= lab.nz(base + x);
    }
    return __any(any);
}

template <typename LAB>
__global__ __launch_bounds__(256) void k_stats_pass1(DBuf inten, LAB lab, Geom g, int64_t L, Acc a) {
#pragma clang fp contract(off)
    const int64_t spans = (g.W + SPAN - 1) / SPAN;
    const int64_t total = g.D * g.H * spans;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t sp = wave0; sp < total; sp += nwaves) {
        const int64_t row = sp / spans;
        const int64_t sx0 = (sp - row * spans) * SPAN;
        const int64_t sx1 = sx0 + SPAN < g.W ? sx0 + SPAN : g.W;
        const int64_t base = row * g.W;
        if (!span_has_labels(lab, base, sx0, sx1, lane)) continue;
        const int64_t x0 = sx0 + (int64_t)lane * RUN;
        const int64_t x1 = x0 + RUN < sx1 ? x0 + RUN : sx1;
        if (x0 >= x1) continue;
        const int64_t y = row % g.H, z = row / g.H;
        int cur = -1;
        unsigned long long c = 0, sx = 0;
        double sv = 0, sxv = 0, syv = 0, szv = 0;
        float mn = 0, mx = 0;
        int64_t xa = 0, xb = 0;
        const float yf = (float)y, zf = (float)z;
        for (int64_t x = x0; x <= x1; ++x) {
            const int l = x < x1 ? lab(base + x) : -1;
            if (l != cur) {
                if (cur > 0 && cur < L) {
                    atomicAdd(a.cnt + cur, c);
                    atomicAdd(a.sx + cur, sx);
                    atomicAdd(a.sy + cur, c * (unsigned long long)y);
                    atomicAdd(a.sz + cur, c * (unsigned long long)z);
                    atomicAdd(a.sv + cur, sv);
                    atomicAdd(a.sxv + cur, sxv);
                    atomicAdd(a.syv + cur, syv);
                    atomicAdd(a.szv + cur, szv);
                    atomicMin(a.minv + cur, f2o(mn));
                    atomicMax(a.maxv + cur, f2o(mx));
                    atomicMin(a.bminx + cur, (uint32_t)xa);
                    atomicMax(a.bmaxx + cur, (uint32_t)xb);
                    atomicMin(a.bminy + cur, (uint32_t)y);
                    atomicMax(a.bmaxy + cur, (uint32_t)y);
                    atomicMin(a.bminz + cur, (uint32_t)z);
                    atomicMax(a.bmaxz + cur, (uint32_t)z);
                }
                cur = l;
                c = 0; sx = 0; sv = 0; sxv = 0; syv = 0; szv = 0;
                xa = x;
                if (x < x1 && l > 0) { mn = mx = ld(inten, base + x); }
            }
            if (x == x1) break;
            if (l > 0) {
                const float v = ld(inten, base + x);
                const float xf = (float)x;
                c += 1;
                sx += (unsigned long long)x;
                sv += (double)v;
                sxv += (double)(xf * v);
                syv += (double)(yf * v);
                szv += (double)(zf * v);
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
                xb = x;
            }
        }
    }
}

// float32 centroid, mass centre, mean and area per label, in the reference's order
// of operations (divide_images of float32 sums)
__global__ void k_stats_mid(Acc a, int64_t L) {
#pragma clang fp contract(off)
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const float area = (float)a.cnt[l];
    const float sv = (float)a.sv[l];
    a.img[0 * L + l] = (float)a.sx[l] / area;
    a.img[1 * L + l] = (float)a.sy[l] / area;
    a.img[2 * L + l] = (float)a.sz[l] / area;
    a.img[3 * L + l] = (float)a.sxv[l] / sv;
    a.img[4 * L + l] = (float)a.syv[l] / sv;
    a.img[5 * L + l] = (float)a.szv[l] / sv;
    a.img[6 * L + l] = sv / area;
    a.img[7 * L + l] = area;
}

template <typename LAB>
__global__ __launch_bounds__(256) void k_stats_pass2(DBuf inten, LAB lab, Geom g, int64_t L, Acc a) {
#pragma clang fp contract(off)
    const int64_t spans = (g.W + SPAN - 1) / SPAN;
    const int64_t total = g.D * g.H * spans;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t sp = wave0; sp < total; sp += nwaves) {
        const int64_t row = sp / spans;
        const int64_t sx0 = (sp - row * spans) * SPAN;
        const int64_t sx1 = sx0 + SPAN < g.W ? sx0 + SPAN : g.W;
        const int64_t base = row * g.W;
        if (!span_has_labels(lab, base, sx0, sx1, lane)) continue;
        const int64_t x0 = sx0 + (int64_t)lane * RUN;
        const int64_t x1 = x0 + RUN < sx1 ? x0 + RUN : sx1;
        if (x0 >= x1) continue;
        const int64_t y = row % g.H, z = row / g.H;
        const float yf = (float)y, zf = (float)z;
        int cur = -1;
        double sdc = 0, sdm = 0, ssq = 0;
        float mdc = 0, mdm = 0;
        float cx = 0, cy = 0, cz = 0, mx = 0, my = 0, mz = 0, mean = 0, area = 1;
        for (int64_t x = x0; x <= x1; ++x) {
            const int l = x < x1 ? lab(base + x) : -1;
            if (l != cur) {
                if (cur > 0 && cur < L) {
                    atomicAdd(a.sdc + cur, sdc);
                    atomicAdd(a.sdm + cur, sdm);
                    atomicAdd(a.ssq + cur, ssq);
                    atomicMax(a.maxdc + cur, __float_as_uint(mdc));
                    atomicMax(a.maxdm + cur, __float_as_uint(mdm));
                }
                cur = l;
                sdc = sdm = ssq = 0;
                mdc = mdm = 0;
                if (l > 0 && l < L) {
                    cx = a.img[0 * L + l]; cy = a.img[1 * L + l]; cz = a.img[2 * L + l];
                    mx = a.img[3 * L + l]; my = a.img[4 * L + l]; mz = a.img[5 * L + l];
                    mean = a.img[6 * L + l]; area = a.img[7 * L + l];
                }
            }
            if (x == x1) break;
            if (l > 0) {
                const float v = ld(inten, base + x);
                const float xf = (float)x;
                const float dx = xf - cx, dy = yf - cy, dz = zf - cz;
                const float dc = sqrtf((dx * dx + dy * dy) + dz * dz);
                const float ex = xf - mx, ey = yf - my, ez = zf - mz;
                const float dm = sqrtf((ex * ex + ey * ey) + ez * ez);
                const float dv = v - mean;
                const float sq = (dv * dv) / area;
                sdc += (double)dc;
                sdm += (double)dm;
                ssq += (double)sq;
                mdc = fmaxf(mdc, dc);
                mdm = fmaxf(mdm, dm);
            }
        }
    }
}

template <typename LAB>
static void stats_passes(DBuf inten, LAB lab, const Geom& g, int64_t L, const Acc& a, hipStream_t s) {
    const int64_t spans = g.D * g.H * ((g.W + SPAN - 1) / SPAN);
    unsigned gr = grid_for((size_t)spans, 4);  // 4 wavefronts per block
    if (gr > 65536) gr = 65536;
    if (g.n() > 0) k_stats_pass1<<<gr, 256, 0, s>>>(inten, lab, g, L, a);
    k_stats_mid<<<grid_for((size_t)L, 256), 256, 0, s>>>(a, L);
    if (g.n() > 0) k_stats_pass2<<<gr, 256, 0, s>>>(inten, lab, g, L, a);
}

// STATISTICS_ENTRY table, one column per measured label 1..L-1
__global__ void k_stats_table(Acc a, int64_t L, float* out) {
#pragma clang fp contract(off)
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (l >= L) return;
    const int64_t N = L - 1;
    const int64_t j = l - 1;
    const bool present = a.cnt[l] > 0;
    const float minx = present ? (float)a.bminx[l] : FLT_MAX, maxx = present ? (float)a.bmaxx[l] : -FLT_MAX;
    const float miny = present ? (float)a.bminy[l] : FLT_MAX, maxy = present ? (float)a.bmaxy[l] : -FLT_MAX;
    const float minz = present ? (float)a.bminz[l] : FLT_MAX, maxz = present ? (float)a.bmaxz[l] : -FLT_MAX;
    const float area = a.img[7 * L + l];
    const float sv = (float)a.sv[l];
    const float sdm = (float)a.sdm[l], sdc = (float)a.sdc[l];
    const float meandm = sdm / area, meandc = sdc / area;
    const float maxdm = __uint_as_float(a.maxdm[l]), maxdc = __uint_as_float(a.maxdc[l]);
    float col[CLE_STAT_COLUMNS];
    col[0] = (float)l;
    col[1] = minx; col[2] = miny; col[3] = minz;
    col[4] = maxx; col[5] = maxy; col[6] = maxz;
    col[7] = maxx - minx + 1.f; col[8] = maxy - miny + 1.f; col[9] = maxz - minz + 1.f;
    col[10] = present ? o2f(a.minv[l]) : FLT_MAX;
    col[11] = present ? o2f(a.maxv[l]) : -FLT_MAX;
    col[12] = a.img[6 * L + l];
    col[13] = sv;
    col[14] = sqrtf((float)a.ssq[l]);
    col[15] = area;
    col[16] = (float)a.sxv[l]; col[17] = (float)a.syv[l]; col[18] = (float)a.szv[l];
    col[19] = a.img[3 * L + l]; col[20] = a.img[4 * L + l]; col[21] = a.img[5 * L + l];
    col[22] = (float)a.sx[l]; col[23] = (float)a.sy[l]; col[24] = (float)a.sz[l];
    col[25] = a.img[0 * L + l]; col[26] = a.img[1 * L + l]; col[27] = a.img[2 * L + l];
    col[28] = sdm; col[29] = meandm; col[30] = maxdm; col[31] = maxdm / meandm;
    col[32] = sdc; col[33] = meandc; col[34] = maxdc; col[35] = maxdc / meandc;
    for (int c = 0; c < CLE_STAT_COLUMNS; ++c) out[c * N + j] = col[c];
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_statistics_of_labelled_pixels(const cle_buf* intensity, const cle_buf* labels, int64_t num_labels,
                                      const cle_buf* out_table, void* stream) {
    Geom gi, gl, go;
    CLE_TRY(check_buf(intensity, "intensity_image", &gi));
    CLE_TRY(check_buf(labels, "label_image", &gl));
    CLE_TRY(check_same_shape(gi, gl, "intensity_image", "label_image"));
    CLE_TRY(check_buf(out_table, "table", &go));
    if (out_table->dtype != CLE_F32) return fail(CLE_E_TYPE, "statistics table must be float32");
    if (num_labels < 1) return fail(CLE_E_VALUE, "num_labels must be >= 1");
    const int64_t L = num_labels;
    if (go.n() < (int64_t)CLE_STAT_COLUMNS * (L - 1))
        return fail(CLE_E_SHAPE, "statistics table holds %lld < %lld values", (long long)go.n(),
                    (long long)CLE_STAT_COLUMNS * (L - 1));
    hipStream_t s = (hipStream_t)stream;
    // accumulators: 11 arrays of 8-byte sums, then 10 arrays of 4-byte extremes and
    // the 8-row float32 image of pass-1 results
    const size_t n8 = 11, n4 = 10 + 8;
    Scratch acc;
    CLE_TRY(acc.alloc((size_t)L * (n8 * 8 + n4 * 4), s));
    char* p = acc.get<char>();
    Acc a;
    auto take8 = [&](void) { void* r = p; p += (size_t)L * 8; return r; };
    auto take4 = [&](void) { void* r = p; p += (size_t)L * 4; return r; };
    a.cnt = (unsigned long long*)take8(); a.sx = (unsigned long long*)take8();
    a.sy = (unsigned long long*)take8(); a.sz = (unsigned long long*)take8();
    a.sv = (double*)take8(); a.sxv = (double*)take8(); a.syv = (double*)take8(); a.szv = (double*)take8();
    a.sdc = (double*)take8(); a.sdm = (double*)take8(); a.ssq = (double*)take8();
    a.minv = (uint32_t*)take4(); a.maxv = (uint32_t*)take4();
    a.bminx = (uint32_t*)take4(); a.bmaxx = (uint32_t*)take4(); a.bminy = (uint32_t*)take4();
    a.bmaxy = (uint32_t*)take4(); a.bminz = (uint32_t*)take4(); a.bmaxz = (uint32_t*)take4();
    a.maxdc = (uint32_t*)take4(); a.maxdm = (uint32_t*)take4();
    a.img = (float*)p;
    // zero the sums, set min/max sentinels
    CLE_TRY(fill_u64((unsigned long long*)acc.p, 0ull, (size_t)L * n8, s));
    CLE_TRY(fill_u32(a.minv, 0xFFFFFFFFu, (size_t)L, s));
    CLE_TRY(fill_u32(a.maxv, 0u, (size_t)L, s));
    CLE_TRY(fill_u32(a.bminx, 0xFFFFFFFFu, (size_t)L, s));
    CLE_TRY(fill_u32(a.bmaxx, 0u, (size_t)L, s));
    CLE_TRY(fill_u32(a.bminy, 0xFFFFFFFFu, (size_t)L, s));
    CLE_TRY(fill_u32(a.bmaxy, 0u, (size_t)L, s));
    CLE_TRY(fill_u32(a.bminz, 0xFFFFFFFFu, (size_t)L, s));
    CLE_TRY(fill_u32(a.bmaxz, 0u, (size_t)L, s));
    CLE_TRY(fill_u32(a.maxdc, 0u, (size_t)L, s));
    CLE_TRY(fill_u32(a.maxdm, 0u, (size_t)L, s));
    const DBuf in = dbuf(intensity);
    switch (labels->dtype) {
        case CLE_I32: stats_passes(in, LabT<int32_t>{(const int32_t*)labels->data}, gi, L, a, s); break;
        case CLE_U32: stats_passes(in, LabT<uint32_t>{(const uint32_t*)labels->data}, gi, L, a, s); break;
        case CLE_F32: stats_passes(in, LabT<float>{(const float*)labels->data}, gi, L, a, s); break;
        case CLE_U16: stats_passes(in, LabT<uint16_t>{(const uint16_t*)labels->data}, gi, L, a, s); break;
        default: stats_passes(in, LabAny{dbuf(labels)}, gi, L, a, s); break;
    }
    if (L > 1) k_stats_table<<<grid_for((size_t)(L - 1), 256), 256, 0, s>>>(a, L, (float*)out_table->data);
    CLE_LAUNCHED();
    return CLE_OK;
}

}  // extern "C"
