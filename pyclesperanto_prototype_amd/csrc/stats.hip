// statistics_of_labelled_pixels, reproducing the reference's float32 accumulation.
//
// Reference semantics: pyclesperanto_prototype/_tier9/_statistics_of_labelled_pixels.py:43-291,
// statistics_per_label_x.cl:1-146 and standard_deviation_per_label_x.cl:1-96.  The
// reference keeps one float32 cell per (plane, row y, label); the kernel for slice z
// walks row y along x and adds every voxel of label l to cell (y, l) in turn, so a
// cell holds the float32 running sum of its voxels in (z, x) order.  The per-label
// result is the float32 sum of the cells over y in y order (min / max over y for the
// extreme planes).  Pass 2 does the same with distances to the float32 centroid and
// mass centre and with (v - mean)^2 / area.  Column layout of the output table:
// _tier9/_statistics_entry.py:3-46.
//
// MI355X design.  (1) Label runs: a wavefront holds a 1024-voxel row span in registers
// (coalesced loads), finds the runs of equal measured labels with shuffles and numbers
// them with ballots, as (label, y, z, x0) keys.  A first pass counts them per span, an
// exclusive scan gives every span its offset (one synchronisation sizes the buffers),
// a second pass writes them.  (2) A
// radix sort orders the runs by key, i.e. by cell (label, y) and, inside a cell, in
// the reference's (z, x) order.  (3) One thread per cell accumulates its voxels in
// float32 with the reference's per-voxel operations (products and sums rounded
// separately).  (4) One thread per label adds its cells in y order.  Only labelled
// voxels are read after the run pass, so the cost is one read of the label image plus
// the foreground.
#include <float.h>

#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace cle {

template <typename T>
struct LabT {
    const T* p;
    // the reference reads a label as (int) of the float it loads
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)(float)p[i]; }
    __device__ __forceinline__ bool nz(int64_t i) const { return p[i] != (T)0; }
};
struct LabAny {
    DBuf b;
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)ld(b, i); }
    __device__ __forceinline__ bool nz(int64_t i) const { return cle::nz(b, i); }
};

struct KeyBits {
    int xb, zb, yb, lb;
    __host__ __device__ int total() const { return xb + zb + yb + lb; }
};

__device__ __forceinline__ uint64_t run_key(const KeyBits& kb, uint64_t l, uint64_t y, uint64_t z, uint64_t x0) {
    return (((l << kb.yb | y) << kb.zb | z) << kb.xb) | x0;
}

// Label runs of a row span, one span per wavefront, no LDS and no block barrier.
// Lane l holds the labels x = sx0 + l + 64 k (k = 0..15; coalesced loads).  A voxel
// starts a run if its label is measured (0 < l < L) and differs from its left
// neighbour inside the span, and ends one if it differs from its right neighbour;
// ballots number the starts and the ends in x order, so start j and end j bound run j.
// Runs are split at span borders, which keeps them in x order.
constexpr int RB_THREADS = 256;
constexpr int RB_K = 16;
constexpr int RB_SPAN = 64 * RB_K;

__device__ __forceinline__ uint64_t lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// PASS 0: span_count[sp] = number of runs; PASS 1: write them at span_off[sp] + rank
template <int PASS, typename LAB>
__global__ __launch_bounds__(RB_THREADS) void k_runs(LAB lab, Geom g, int64_t L, KeyBits kb, uint32_t* span_count,
                                                     const uint32_t* span_off, uint64_t* keys, uint32_t* ends) {
    const int64_t spans = (g.W + RB_SPAN - 1) / RB_SPAN;
    const int64_t total = g.D * g.H * spans;
    const int lane = threadIdx.x & 63;
    const uint64_t below = lanes_below();
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t sp = wave0; sp < total; sp += nwaves) {
        const int64_t row = sp / spans;
        const int64_t sx0 = (sp - row * spans) * RB_SPAN;
        const int n = (int)((sx0 + RB_SPAN < g.W ? sx0 + RB_SPAN : g.W) - sx0);
        const int64_t base = row * g.W + sx0;
        int v[RB_K];
        bool any = false;
#pragma unroll
        for (int k = 0; k < RB_K; ++k) {
            const int i = lane + 64 * k;
            v[k] = i < n ? lab(base + i) : 0;
            any |= v[k] != 0;
        }
        if (!__any(any)) {
            if (PASS == 0 && lane == 0) span_count[sp] = 0;
            continue;
        }
        uint32_t nstart = 0, nend = 0;
        uint64_t slot0 = 0;
        uint64_t y = 0, z = 0;
        if (PASS == 1) {
            slot0 = span_off[sp];
            y = (uint64_t)(row % g.H);
            z = (uint64_t)(row / g.H);
        }
        int prev_last = 0;  // label of voxel 64 k - 1 (lane 63 of the previous k)
#pragma unroll
        for (int k = 0; k < RB_K; ++k) {
            const int i = lane + 64 * k;
            const int l = v[k];
            int left = __shfl_up(l, 1);
            if (lane == 0) left = prev_last;
            int right = __shfl_down(l, 1);
            const int next_first = k + 1 < RB_K ? __shfl(v[k + 1 < RB_K ? k + 1 : k], 0) : 0;
            if (lane == 63) right = next_first;
            prev_last = __shfl(l, 63);
            const bool valid = i < n && l > 0 && l < L;
            const bool is_start = valid && (i == 0 || left != l);
            const bool is_end = valid && (i == n - 1 || right != l);
            const uint64_t ms = __ballot(is_start), me = __ballot(is_end);
            if (PASS == 1) {
                if (is_start)
                    keys[slot0 + nstart + __popcll(ms & below)] = run_key(kb, (uint64_t)l, y, z, (uint64_t)(sx0 + i));
                if (is_end) ends[slot0 + nend + __popcll(me & below)] = (uint32_t)(sx0 + i + 1);
            }
            nstart += (uint32_t)__popcll(ms);
            nend += (uint32_t)__popcll(me);
        }
        if (PASS == 0 && lane == 0) span_count[sp] = nstart;
    }
}

// cell (label, y) boundaries in the sorted runs
__global__ void k_cell_flags(const uint64_t* keys, int64_t n, int cell_shift, uint32_t* flags) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || (keys[i] >> cell_shift) != (keys[i - 1] >> cell_shift)) ? 1u : 0u;
}

__global__ void k_cell_scatter(const uint32_t* flags, const uint32_t* cidx, int64_t n, uint32_t* cell_start,
                               uint32_t* ncells) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (flags[i]) cell_start[cidx[i]] = (uint32_t)i;
        if (i == n - 1) {
            const uint32_t total = cidx[i] + flags[i];
            cell_start[total] = (uint32_t)n;
            *ncells = total;
        }
    }
}

// [lstart[l], lend[l]) = the cells of label l (cells are sorted by label, then y)
__global__ void k_label_ranges(const uint64_t* keys, const uint32_t* cell_start, const uint32_t* ncells_p,
                               int64_t cap, KeyBits kb, uint32_t* lstart, uint32_t* lend) {
    const uint32_t ncells = *ncells_p;
    const int lshift = kb.yb + kb.zb + kb.xb;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cap && c < ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t l = keys[cell_start[c]] >> lshift;
        if (c == 0 || (keys[cell_start[c - 1]] >> lshift) != l) lstart[l] = (uint32_t)c;
        if (c == (int64_t)ncells - 1 || (keys[cell_start[c + 1]] >> lshift) != l) lend[l] = (uint32_t)c + 1;
    }
}

struct CellGeom {
    KeyBits kb;
    int64_t H, W;
    __device__ __forceinline__ void decode(uint64_t key, int64_t& y, int64_t& z, int64_t& x0) const {
        x0 = (int64_t)(key & ((1ull << kb.xb) - 1));
        z = (int64_t)((key >> kb.xb) & ((1ull << kb.zb) - 1));
        y = (int64_t)((key >> (kb.xb + kb.zb)) & ((1ull << kb.yb) - 1));
    }
};

// pass 1, one thread per cell: 16 float32 planes in the reference's order:
// 0-2 sum x,y,z  3 count  4-6 sum x*v, y*v, z*v  7 sum v  8/9 min/max v  10-15 min/max x,y,z
__global__ __launch_bounds__(256) void k_cells_pass1(DBuf inten, CellGeom cg, const uint64_t* keys,
                                                     const uint32_t* ends, const uint32_t* cell_start,
                                                     const uint32_t* ncells_p, int64_t cap, float* acc) {
#pragma clang fp contract(off)
    const uint32_t ncells = *ncells_p;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cap && c < ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r0 = cell_start[c], r1 = cell_start[c + 1];
        float s[16];
        for (int k = 0; k < 8; ++k) s[k] = 0.f;
        for (int k = 8; k < 16; k += 2) { s[k] = FLT_MAX; s[k + 1] = -FLT_MAX; }
        for (uint32_t r = r0; r < r1; ++r) {
            int64_t y, z, x0;
            cg.decode(keys[r], y, z, x0);
            const int64_t x1 = ends[r];
            const float yf = (float)y, zf = (float)z;
            const int64_t base = (z * cg.H + y) * cg.W;
            for (int64_t x = x0; x < x1; ++x) {
                const float v = ld(inten, base + x);
                const float xf = (float)x;
                if (s[3] == 0.f) {
                    s[8] = v; s[9] = v; s[10] = xf; s[11] = xf; s[12] = yf; s[13] = yf; s[14] = zf; s[15] = zf;
                } else {
                    s[8] = s[8] > v ? v : s[8];
                    s[9] = s[9] < v ? v : s[9];
                    s[10] = s[10] > xf ? xf : s[10];
                    s[11] = s[11] < xf ? xf : s[11];
                    s[12] = s[12] > yf ? yf : s[12];
                    s[13] = s[13] < yf ? yf : s[13];
                    s[14] = s[14] > zf ? zf : s[14];
                    s[15] = s[15] < zf ? zf : s[15];
                }
                s[0] = s[0] + xf;
                s[1] = s[1] + yf;
                s[2] = s[2] + zf;
                s[3] = s[3] + 1.f;
                s[4] = s[4] + xf * v;
                s[5] = s[5] + yf * v;
                s[6] = s[6] + zf * v;
                s[7] = s[7] + v;
            }
        }
        for (int k = 0; k < 16; ++k) acc[(int64_t)k * cap + c] = s[k];
    }
}

// Per-label results of pass 1: the y-reduced planes [16][L] (sums in y order,
// min / max), and the float32 image the second pass needs [8][L]: centroid xyz,
// mass centre xyz, mean, area.
__global__ void k_labels_pass1(const float* acc, int64_t cap, const uint32_t* lstart, const uint32_t* lend,
                               int64_t L, float* lab16, float* img) {
#pragma clang fp contract(off)
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    float s[16];
    for (int k = 0; k < 8; ++k) s[k] = 0.f;
    for (int k = 8; k < 16; k += 2) { s[k] = FLT_MAX; s[k + 1] = -FLT_MAX; }
    const uint32_t c0 = lstart[l], c1 = lend[l];
    for (uint32_t c = c0; c < c1; ++c) {
        for (int k = 0; k < 8; ++k) s[k] = s[k] + acc[(int64_t)k * cap + c];
        for (int k = 8; k < 16; k += 2) {
            const float mn = acc[(int64_t)k * cap + c], mx = acc[(int64_t)(k + 1) * cap + c];
            s[k] = s[k] > mn ? mn : s[k];
            s[k + 1] = s[k + 1] < mx ? mx : s[k + 1];
        }
    }
    for (int k = 0; k < 16; ++k) lab16[(int64_t)k * L + l] = s[k];
    const float area = s[3], sv = s[7];
    img[0 * L + l] = s[0] / area;
    img[1 * L + l] = s[1] / area;
    img[2 * L + l] = s[2] / area;
    img[3 * L + l] = s[4] / sv;
    img[4 * L + l] = s[5] / sv;
    img[5 * L + l] = s[6] / sv;
    img[6 * L + l] = sv / area;
    img[7 * L + l] = area;
}

// pass 2, one thread per cell: 0 sum distance to centroid, 1 to mass centre,
// 2 sum (v - mean)^2 / area, 3 count, 4/5 max distances
__global__ __launch_bounds__(256) void k_cells_pass2(DBuf inten, CellGeom cg, const uint64_t* keys,
                                                     const uint32_t* ends, const uint32_t* cell_start,
                                                     const uint32_t* ncells_p, int64_t cap, const float* img,
                                                     int64_t L, float* acc2) {
#pragma clang fp contract(off)
    const uint32_t ncells = *ncells_p;
    const int lshift = cg.kb.yb + cg.kb.zb + cg.kb.xb;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cap && c < ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r0 = cell_start[c], r1 = cell_start[c + 1];
        const int64_t l = (int64_t)(keys[r0] >> lshift);
        const float cx = img[0 * L + l], cy = img[1 * L + l], cz = img[2 * L + l];
        const float mx = img[3 * L + l], my = img[4 * L + l], mz = img[5 * L + l];
        const float mean = img[6 * L + l], area = img[7 * L + l];
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, m4 = 0.f, m5 = 0.f;
        for (uint32_t r = r0; r < r1; ++r) {
            int64_t y, z, x0;
            cg.decode(keys[r], y, z, x0);
            const int64_t x1 = ends[r];
            const float yf = (float)y, zf = (float)z;
            const int64_t base = (z * cg.H + y) * cg.W;
            for (int64_t x = x0; x < x1; ++x) {
                const float v = ld(inten, base + x);
                const float xf = (float)x;
                const float dx = xf - cx, dy = yf - cy, dz = zf - cz;
                const float dc = sqrtf((dx * dx + dy * dy) + dz * dz);
                const float ex = xf - mx, ey = yf - my, ez = zf - mz;
                const float dm = sqrtf((ex * ex + ey * ey) + ez * ez);
                const float dv = v - mean;
                const float sq = (dv * dv) / area;
                if (s3 == 0.f) { m4 = dc; m5 = dm; }
                else { m4 = m4 < dc ? dc : m4; m5 = m5 < dm ? dm : m5; }
                s0 = s0 + dc;
                s1 = s1 + dm;
                s2 = s2 + sq;
                s3 = s3 + 1.f;
            }
        }
        acc2[0 * cap + c] = s0;
        acc2[1 * cap + c] = s1;
        acc2[2 * cap + c] = s2;
        acc2[3 * cap + c] = s3;
        acc2[4 * cap + c] = m4;
        acc2[5 * cap + c] = m5;
    }
}

// y-reduction of pass 2 and the STATISTICS_ENTRY table, one thread per label 1..L-1
__global__ void k_labels_table(const float* acc2, int64_t cap, const uint32_t* lstart, const uint32_t* lend,
                               const float* lab16, const float* img, int64_t L, float* out) {
#pragma clang fp contract(off)
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (l >= L) return;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, m4 = 0.f, m5 = 0.f;
    for (uint32_t c = lstart[l]; c < lend[l]; ++c) {
        s0 = s0 + acc2[0 * cap + c];
        s1 = s1 + acc2[1 * cap + c];
        s2 = s2 + acc2[2 * cap + c];
        // the reference's max over y also sees the zero cells of rows without the label
        const float a4 = acc2[4 * cap + c], a5 = acc2[5 * cap + c];
        m4 = m4 < a4 ? a4 : m4;
        m5 = m5 < a5 ? a5 : m5;
    }
    auto pl = [&](int k) { return lab16[(int64_t)k * L + l]; };
    const float area = pl(3), sv = pl(7);
    const int64_t N = L - 1, j = l - 1;
    float col[CLE_STAT_COLUMNS];
    col[0] = (float)l;
    col[1] = pl(10); col[2] = pl(12); col[3] = pl(14);
    col[4] = pl(11); col[5] = pl(13); col[6] = pl(15);
    col[7] = col[4] - col[1] + 1.f; col[8] = col[5] - col[2] + 1.f; col[9] = col[6] - col[3] + 1.f;
    col[10] = pl(8); col[11] = pl(9);
    col[12] = img[6 * L + l];
    col[13] = sv;
    col[14] = sqrtf(s2);
    col[15] = area;
    col[16] = pl(4); col[17] = pl(5); col[18] = pl(6);
    col[19] = img[3 * L + l]; col[20] = img[4 * L + l]; col[21] = img[5 * L + l];
    col[22] = pl(0); col[23] = pl(1); col[24] = pl(2);
    col[25] = img[0 * L + l]; col[26] = img[1 * L + l]; col[27] = img[2 * L + l];
    col[28] = s1; col[29] = s1 / area; col[30] = m5; col[31] = m5 / col[29];
    col[32] = s0; col[33] = s0 / area; col[34] = m4; col[35] = m4 / col[33];
    for (int c = 0; c < CLE_STAT_COLUMNS; ++c) out[c * N + j] = col[c];
}

__global__ void k_span_total(const uint32_t* cnt, const uint32_t* off, int64_t n, uint32_t* total) {
    *total = off[n - 1] + cnt[n - 1];
}

static int bits_for(uint64_t maxval) {
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

template <typename LAB>
static int stats_run(DBuf inten, LAB lab, const Geom& g, int64_t L, float* table, hipStream_t s) {
    KeyBits kb;
    kb.xb = bits_for((uint64_t)(g.W > 0 ? g.W - 1 : 0));
    kb.zb = bits_for((uint64_t)(g.D > 0 ? g.D - 1 : 0));
    kb.yb = bits_for((uint64_t)(g.H > 0 ? g.H - 1 : 0));
    kb.lb = bits_for((uint64_t)(L - 1));
    if (kb.total() > 64) return fail(CLE_E_SHAPE, "statistics: %d key bits exceed 64", kb.total());
    const int64_t spans = g.D * g.H * ((g.W + RB_SPAN - 1) / RB_SPAN);
    unsigned gr = grid_for((size_t)(spans > 0 ? spans : 1), RB_THREADS / 64);  // one wavefront per span
    if (gr > 32768) gr = 32768;
    // (1) count the runs per span, offsets by an exclusive scan (one synchronisation to
    //     size the buffers), then write every span's runs at its offset
    Scratch scount, soff, stot;
    const int64_t scap = spans > 0 ? spans : 1;
    CLE_TRY(scount.alloc((size_t)scap * 4, s));
    CLE_TRY(soff.alloc((size_t)scap * 4, s));
    CLE_TRY(stot.alloc(8, s));
    CLE_TRY(fill_u32(stot.get<uint32_t>(), 0u, 2, s));
    if (spans > 0) {
        k_runs<0><<<gr, RB_THREADS, 0, s>>>(lab, g, L, kb, scount.get<uint32_t>(), nullptr, nullptr, nullptr);
        size_t sb = 0;
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, scount.get<uint32_t>(), soff.get<uint32_t>(), (int)spans, s));
        Scratch st;
        CLE_TRY(st.alloc(sb, s));
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(st.p, sb, scount.get<uint32_t>(), soff.get<uint32_t>(), (int)spans, s));
        k_span_total<<<1, 1, 0, s>>>(scount.get<uint32_t>(), soff.get<uint32_t>(), spans, stot.get<uint32_t>());
    }
    CLE_LAUNCHED();
    uint32_t nruns = 0;
    CLE_HIP(hipMemcpyAsync(&nruns, stot.p, 4, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    const int64_t n = (int64_t)nruns;
    const int64_t cap = n > 0 ? n : 1;
    Scratch k0, k1, e0, e1, tmp, flags, cidx, cstart, ncell, lst, len_, acc, acc2, lab16, img;
    CLE_TRY(k0.alloc((size_t)cap * 8, s));
    CLE_TRY(k1.alloc((size_t)cap * 8, s));
    CLE_TRY(e0.alloc((size_t)cap * 4, s));
    CLE_TRY(e1.alloc((size_t)cap * 4, s));
    if (n > 0)
        k_runs<1><<<gr, RB_THREADS, 0, s>>>(lab, g, L, kb, nullptr, soff.get<uint32_t>(), k0.get<uint64_t>(),
                                            e0.get<uint32_t>());
    CLE_LAUNCHED();
    // (2) sort the runs by (label, y, z, x0)
    size_t tbytes = 0;
    CLE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, k0.get<uint64_t>(), k1.get<uint64_t>(),
                                               e0.get<uint32_t>(), e1.get<uint32_t>(), (int)cap, 0, kb.total(), s));
    CLE_TRY(tmp.alloc(tbytes, s));
    if (n > 0)
        CLE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, k0.get<uint64_t>(), k1.get<uint64_t>(),
                                                   e0.get<uint32_t>(), e1.get<uint32_t>(), (int)n, 0, kb.total(), s));
    const uint64_t* keys = k1.get<uint64_t>();
    const uint32_t* ends = e1.get<uint32_t>();
    // cells and label ranges
    CLE_TRY(flags.alloc((size_t)cap * 4, s));
    CLE_TRY(cidx.alloc((size_t)cap * 4, s));
    CLE_TRY(cstart.alloc((size_t)(cap + 1) * 4, s));
    CLE_TRY(ncell.alloc(4, s));
    CLE_TRY(fill_u32(ncell.get<uint32_t>(), 0u, 1, s));
    CLE_TRY(lst.alloc((size_t)L * 4, s));
    CLE_TRY(len_.alloc((size_t)L * 4, s));
    CLE_TRY(fill_u32(lst.get<uint32_t>(), 0u, (size_t)L, s));
    CLE_TRY(fill_u32(len_.get<uint32_t>(), 0u, (size_t)L, s));
    const unsigned gn = grid_for((size_t)cap, 256) > 65536 ? 65536 : grid_for((size_t)cap, 256);
    if (n > 0) {
        k_cell_flags<<<gn, 256, 0, s>>>(keys, n, kb.zb + kb.xb, flags.get<uint32_t>());
        size_t sbytes = 0;
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sbytes, flags.get<uint32_t>(), cidx.get<uint32_t>(), (int)n, s));
        Scratch stmp;
        CLE_TRY(stmp.alloc(sbytes, s));
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(stmp.p, sbytes, flags.get<uint32_t>(), cidx.get<uint32_t>(), (int)n, s));
        k_cell_scatter<<<gn, 256, 0, s>>>(flags.get<uint32_t>(), cidx.get<uint32_t>(), n, cstart.get<uint32_t>(),
                                          ncell.get<uint32_t>());
        k_label_ranges<<<gn, 256, 0, s>>>(keys, cstart.get<uint32_t>(), ncell.get<uint32_t>(), cap, kb,
                                          lst.get<uint32_t>(), len_.get<uint32_t>());
    }
    CLE_LAUNCHED();
    // (3) + (4): cells, then labels, twice
    CLE_TRY(acc.alloc((size_t)cap * 16 * 4, s));
    CLE_TRY(acc2.alloc((size_t)cap * 6 * 4, s));
    CLE_TRY(lab16.alloc((size_t)L * 16 * 4, s));
    CLE_TRY(img.alloc((size_t)L * 8 * 4, s));
    const CellGeom cg{kb, g.H, g.W};
    if (n > 0)
        k_cells_pass1<<<gn, 256, 0, s>>>(inten, cg, keys, ends, cstart.get<uint32_t>(), ncell.get<uint32_t>(), cap,
                                         acc.get<float>());
    k_labels_pass1<<<grid_for((size_t)L, 256), 256, 0, s>>>(acc.get<float>(), cap, lst.get<uint32_t>(),
                                                            len_.get<uint32_t>(), L, lab16.get<float>(), img.get<float>());
    if (n > 0)
        k_cells_pass2<<<gn, 256, 0, s>>>(inten, cg, keys, ends, cstart.get<uint32_t>(), ncell.get<uint32_t>(), cap,
                                         img.get<float>(), L, acc2.get<float>());
    if (L > 1)
        k_labels_table<<<grid_for((size_t)(L - 1), 256), 256, 0, s>>>(acc2.get<float>(), cap, lst.get<uint32_t>(),
                                                                      len_.get<uint32_t>(), lab16.get<float>(),
                                                                      img.get<float>(), L, table);
    CLE_LAUNCHED();
    return CLE_OK;
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
    const DBuf in = dbuf(intensity);
    float* table = (float*)out_table->data;
    switch (labels->dtype) {
        case CLE_I32: return stats_run(in, LabT<int32_t>{(const int32_t*)labels->data}, gi, L, table, s);
        case CLE_U32: return stats_run(in, LabT<uint32_t>{(const uint32_t*)labels->data}, gi, L, table, s);
        case CLE_F32: return stats_run(in, LabT<float>{(const float*)labels->data}, gi, L, table, s);
        case CLE_U16: return stats_run(in, LabT<uint16_t>{(const uint16_t*)labels->data}, gi, L, table, s);
        default: return stats_run(in, LabAny{dbuf(labels)}, gi, L, table, s);
    }
}

}  // extern "C"
