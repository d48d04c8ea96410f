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

#include <atomic>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace cle {

// raw(i) / cvt(r): the stored label and its conversion, so that a prefetch holds the
// loaded value and converts it where it is used
template <typename T>
struct LabT {
    using type = T;
    using raw_t = T;
    const T* p;
    // the reference reads a label as (int) of the float it loads
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)(float)p[i]; }
    __device__ __forceinline__ bool nz(int64_t i) const { return p[i] != (T)0; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static int cvt(raw_t r) { return (int)(float)r; }
};
struct LabAny {
    DBuf b;
    using raw_t = float;
    __device__ __forceinline__ int operator()(int64_t i) const { return (int)ld(b, i); }
    __device__ __forceinline__ bool nz(int64_t i) const { return cle::nz(b, i); }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return ld(b, i); }
    __device__ __forceinline__ static int cvt(raw_t r) { return (int)r; }
};

struct KeyBits {
    int xb, zb, yb, lb;
    __host__ __device__ int total() const { return xb + zb + yb + lb; }
};

__device__ __forceinline__ uint64_t run_key(const KeyBits& kb, uint64_t l, uint64_t y, uint64_t z, uint64_t x0) {
    return (((l << kb.yb | y) << kb.zb | z) << kb.xb) | x0;
}

// Label runs of a row span, one span of 1024 voxels per wavefront, no LDS, no atomics.
// VEC layout (rows of a multiple of 4 voxels, 16-byte aligned): lane l holds the four
// voxels x = sx0 + 256 k + 4 l + j (k, j = 0..3; one 16-byte load per k); scalar
// layout: x = sx0 + 64 k + l (k = 0..15).  A voxel starts a run if its label is
// measured (0 < label < lcap) and differs from its left neighbour inside the span,
// and ends one if it differs from its right neighbour; ballots number the starts and
// the ends in x order, so start i and end i bound run i.  Runs are split at span
// borders, which keeps them in x order.  A span writes its runs into its own
// `scap` slots (slot mode, off == nullptr; runs past scap are counted, not written)
// or at off[span] (exact mode, once the offsets are known); count[span] = its runs.
// bmax[block] = (largest span count, largest measured label) of the block.
constexpr int RB_THREADS = 256;
constexpr int RB_SPAN = 1024;

__device__ __forceinline__ uint64_t lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T> struct Vec4;
template <> struct Vec4<int32_t> { using V = int4; };
template <> struct Vec4<uint32_t> { using V = uint4; };
template <> struct Vec4<float> { using V = float4; };

// A span's labels as stored (no branch: lanes past the row end load voxel 0 of the
// span and are masked in runs_cvt), so the loads stay in flight until used.
template <bool VEC, typename LAB>
__device__ __forceinline__ void runs_fetch(const LAB& lab, int64_t base, int n, int lane,
                                           typename LAB::raw_t (&r)[RB_SPAN / 64]) {
    constexpr int E = VEC ? 4 : 1;
    constexpr int K = RB_SPAN / (64 * E);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i0 = 64 * E * k + E * lane;
        const int j = i0 < n ? i0 : 0;
        if constexpr (VEC) {  // n is a multiple of 4: all four or none
            const auto q = *(const typename Vec4<typename LAB::type>::V*)(lab.p + base + j);
            r[4 * k + 0] = q.x; r[4 * k + 1] = q.y; r[4 * k + 2] = q.z; r[4 * k + 3] = q.w;
        } else {
            r[k] = lab.raw(base + j);
        }
    }
}

template <bool VEC, typename LAB>
__device__ __forceinline__ void runs_cvt(const typename LAB::raw_t (&r)[RB_SPAN / 64], int n, int lane,
                                         int (&v)[RB_SPAN / 64]) {
    constexpr int E = VEC ? 4 : 1;
    constexpr int K = RB_SPAN / (64 * E);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i0 = 64 * E * k + E * lane;
#pragma unroll
        for (int e = 0; e < E; ++e) v[E * k + e] = i0 < n ? LAB::cvt(r[E * k + e]) : 0;
    }
}

template <bool VEC, typename LAB>
__global__ __launch_bounds__(RB_THREADS) void k_runs(LAB lab, Geom g, int64_t lcap, KeyBits kb, uint32_t scap,
                                                     const uint32_t* off, uint32_t* count, uint64_t* keys,
                                                     uint32_t* ends, uint32_t* bmax) {
    constexpr int E = VEC ? 4 : 1;         // consecutive voxels per lane
    constexpr int K = RB_SPAN / (64 * E);  // blocks of 64 * E voxels
    constexpr int NV = RB_SPAN / 64;       // voxels per lane
    const int64_t spans = (g.W + RB_SPAN - 1) / RB_SPAN;
    const int64_t total = g.D * g.H * spans;
    const int lane = threadIdx.x & 63;
    const uint64_t below = lanes_below();
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int lmax = 0;
    uint32_t cmax = 0;
    // the next span's labels are loaded while this span is processed
    typename LAB::raw_t nr[NV];
    auto span_of = [&](int64_t sp, int64_t& row, int64_t& sx0, int& n) {
        row = sp / spans;
        sx0 = (sp - row * spans) * RB_SPAN;
        n = (int)((sx0 + RB_SPAN < g.W ? sx0 + RB_SPAN : g.W) - sx0);
    };
    if (wave0 < total) {
        int64_t row, sx0;
        int n;
        span_of(wave0, row, sx0, n);
        runs_fetch<VEC>(lab, row * g.W + sx0, n, lane, nr);
    }
    for (int64_t sp = wave0; sp < total; sp += nwaves) {
        int64_t row, sx0;
        int n;
        span_of(sp, row, sx0, n);
        int v[NV];
        runs_cvt<VEC, LAB>(nr, n, lane, v);
        // both modes know where the span's runs go before they are found (the offset
        // is read before the next span's loads are issued, so waiting for it does not
        // wait for them)
        const uint64_t slot0 = off ? (uint64_t)off[sp] : (uint64_t)sp * scap;
        if (sp + nwaves < total) {
            int64_t row2, sx2;
            int n2;
            span_of(sp + nwaves, row2, sx2, n2);
            runs_fetch<VEC>(lab, row2 * g.W + sx2, n2, lane, nr);
        }
        bool any = false;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const bool valid = v[k] > 0 && v[k] < lcap;
            any |= valid;
            lmax = valid && v[k] > lmax ? v[k] : lmax;
        }
        if (!__any(any)) {
            if (lane == 0) count[sp] = 0;
            continue;
        }
        const uint32_t lim = off ? 0xFFFFFFFFu : scap;
        const uint64_t y = (uint64_t)(row % g.H), z = (uint64_t)(row / g.H);
        uint32_t ns = 0, ne = 0;
        int prev_last = 0;  // last voxel of the previous block of 64 * E
#pragma unroll
        for (int k = 0; k < K; ++k) {
            int left = __shfl_up(v[E * k + E - 1], 1);
            if (lane == 0) left = prev_last;
            int right = __shfl_down(v[E * k], 1);
            const int next_first = __shfl(v[k + 1 < K ? E * (k + 1) : E * k], 0);
            if (lane == 63) right = k + 1 < K ? next_first : 0;
            prev_last = __shfl(v[E * k + E - 1], 63);
            bool st[E], en[E];
            uint32_t ps = ns, pe = ne;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int i = 64 * E * k + E * lane + j;
                const int l = v[E * k + j];
                const int lf = j == 0 ? left : v[E * k + j - 1];
                const int rt = j == E - 1 ? right : v[E * k + j + 1];
                const bool valid = i < n && l > 0 && l < lcap;
                st[j] = valid && (i == 0 || lf != l);
                en[j] = valid && (i == n - 1 || rt != l);
                const uint64_t ms = __ballot(st[j]), me = __ballot(en[j]);
                ps += (uint32_t)__popcll(ms & below);
                pe += (uint32_t)__popcll(me & below);
                ns += (uint32_t)__popcll(ms);
                ne += (uint32_t)__popcll(me);
            }
            // ps / pe: runs of this block in earlier lanes; then this lane's own, in order
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int i = 64 * E * k + E * lane + j;
                if (st[j]) {
                    if (ps < lim) keys[slot0 + ps] = run_key(kb, (uint64_t)v[E * k + j], y, z, (uint64_t)(sx0 + i));
                    ++ps;
                }
                if (en[j]) {
                    if (pe < lim) ends[slot0 + pe] = (uint32_t)(sx0 + i + 1);
                    ++pe;
                }
            }
        }
        cmax = ns > cmax ? ns : cmax;
        if (lane == 0) count[sp] = ns;
    }
    // block maxima of (span count, label), one store per block
    __shared__ uint32_t red[2][RB_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(lmax, o);
        const uint32_t c = __shfl_xor(cmax, o);
        lmax = a > lmax ? a : lmax;
        cmax = c > cmax ? c : cmax;
    }
    if (lane == 0) {
        red[0][threadIdx.x >> 6] = cmax;
        red[1][threadIdx.x >> 6] = (uint32_t)lmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, l = 0;
        for (int w = 0; w < RB_THREADS / 64; ++w) {
            c = red[0][w] > c ? red[0][w] : c;
            l = red[1][w] > l ? red[1][w] : l;
        }
        bmax[2 * blockIdx.x] = c;
        bmax[2 * blockIdx.x + 1] = l;
    }
}

// The same runs with the 16-byte layout as 16 consecutive voxels per lane (x = sx0 + 16 l
// + e, e = 0..15; four 16-byte loads per lane, 64 bytes apart across lanes): starts and
// ends are found inside the lane and numbered with one exclusive scan of each lane's
// counts across the wavefront, instead of a ballot and two popcounts per voxel; a label
// is measured by one unsigned compare.  Same runs, same slots and order as k_runs.
// (Measured: 512x2048x2048 labels 2.65 -> 2.2 ms; looping over the set bits with a
// label picked by a select tree instead of the unrolled per-voxel stores was slower.)
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t c, int lane) {
    uint32_t x = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x - c;
}

template <typename LAB>
__global__ __launch_bounds__(RB_THREADS) void k_runs_lane(LAB lab, Geom g, int64_t lcap, KeyBits kb, uint32_t scap,
                                                          const uint32_t* off, uint32_t* count, uint64_t* keys,
                                                          uint32_t* ends, uint32_t* bmax) {
    constexpr int NV = RB_SPAN / 64;  // voxels per lane
    using V4 = typename Vec4<typename LAB::type>::V;
    const int64_t spans = (g.W + RB_SPAN - 1) / RB_SPAN;
    const int64_t total = g.D * g.H * spans;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int lmax = 0;
    uint32_t cmax = 0;
    // measured labels 1 .. lcap - 1 as one unsigned compare of l - 1
    const uint32_t ucap = (uint32_t)(lcap - 1 < (int64_t)0x80000000 ? (lcap - 1 > 0 ? lcap - 1 : 0) : 0x80000000);
    typename LAB::raw_t nr[NV];  // the next span's labels, loaded while this span is processed
    auto span_of = [&](int64_t sp, int64_t& row, int64_t& sx0, int& n) {
        row = sp / spans;
        sx0 = (sp - row * spans) * RB_SPAN;
        n = (int)((sx0 + RB_SPAN < g.W ? sx0 + RB_SPAN : g.W) - sx0);
    };
    // lanes past the row end load voxel 0 of the span (masked below); n is a multiple of 4
    auto fetch = [&](int64_t base, int n) {
#pragma unroll
        for (int k = 0; k < NV / 4; ++k) {
            const int i0 = NV * lane + 4 * k;
            const V4 q = *(const V4*)(lab.p + base + (i0 < n ? i0 : 0));
            nr[4 * k + 0] = q.x; nr[4 * k + 1] = q.y; nr[4 * k + 2] = q.z; nr[4 * k + 3] = q.w;
        }
    };
    if (wave0 < total) {
        int64_t row, sx0;
        int n;
        span_of(wave0, row, sx0, n);
        fetch(row * g.W + sx0, n);
    }
    for (int64_t sp = wave0; sp < total; sp += nwaves) {
        int64_t row, sx0;
        int n;
        span_of(sp, row, sx0, n);
        const int i0 = NV * lane;  // the lane's first voxel in the span
        int v[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) v[e] = i0 + e < n ? LAB::cvt(nr[e]) : 0;
        const uint64_t slot0 = off ? (uint64_t)off[sp] : (uint64_t)sp * scap;
        if (sp + nwaves < total) {
            int64_t row2, sx2;
            int n2;
            span_of(sp + nwaves, row2, sx2, n2);
            fetch(row2 * g.W + sx2, n2);
        }
        bool any = false;
        uint32_t S = 0, E = 0;  // bit e: voxel i0 + e starts / ends a run
        const int left = __shfl_up(v[NV - 1], 1), right = __shfl_down(v[0], 1);
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int l = v[e];
            const bool valid = (uint32_t)l - 1u < ucap;  // 0 < l < lcap; voxels past n hold 0
            any |= valid;
            lmax = valid && l > lmax ? l : lmax;
            const int lf = e > 0 ? v[e - 1] : (lane > 0 ? left : 0);
            const int rt = e < NV - 1 ? v[e + 1] : (lane < 63 ? right : 0);
            const int i = i0 + e;
            S |= (valid && (i == 0 || lf != l)) ? 1u << e : 0u;
            E |= (valid && (i == n - 1 || rt != l)) ? 1u << e : 0u;
        }
        if (!__any(any)) {
            if (lane == 0) count[sp] = 0;
            continue;
        }
        const uint32_t lim = off ? 0xFFFFFFFFu : scap;
        // run_key(l, y, z, x) = l << (yb + zb + xb) | ((y << zb | z) << xb) | x
        const uint64_t yz = (((uint64_t)(row % g.H) << kb.zb) | (uint64_t)(row / g.H)) << kb.xb;
        const int lsh = kb.yb + kb.zb + kb.xb;
        const uint32_t cs = (uint32_t)__popc(S);
        uint32_t ps = wave_excl_scan(cs, lane), pe = wave_excl_scan((uint32_t)__popc(E), lane);
        const uint32_t ns = __shfl(ps + cs, 63);
        uint64_t* kp = keys + slot0 + ps;
        uint32_t* ep = ends + slot0 + pe;
        const uint32_t x0 = (uint32_t)(sx0 + i0);
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            if (S & (1u << e)) {
                if (ps < lim) *kp = ((uint64_t)(uint32_t)v[e] << lsh) | yz | (uint64_t)(x0 + e);
                ++kp;
                ++ps;
            }
            if (E & (1u << e)) {
                if (pe < lim) *ep = x0 + e + 1;
                ++ep;
                ++pe;
            }
        }
        cmax = ns > cmax ? ns : cmax;
        if (lane == 0) count[sp] = ns;
    }
    // block maxima of (span count, label), one store per block
    __shared__ uint32_t red[2][RB_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(lmax, o);
        const uint32_t c = __shfl_xor(cmax, o);
        lmax = a > lmax ? a : lmax;
        cmax = c > cmax ? c : cmax;
    }
    if (lane == 0) {
        red[0][threadIdx.x >> 6] = cmax;
        red[1][threadIdx.x >> 6] = (uint32_t)lmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, l = 0;
        for (int w = 0; w < RB_THREADS / 64; ++w) {
            c = red[0][w] > c ? red[0][w] : c;
            l = red[1][w] > l ? red[1][w] : l;
        }
        bmax[2 * blockIdx.x] = c;
        bmax[2 * blockIdx.x + 1] = l;
    }
}

// sum[0] = runs (off[last] + count[last]), sum[1] = largest span count, sum[2] =
// largest measured label (one block)
__global__ __launch_bounds__(1024) void k_runs_summary(const uint32_t* count, const uint32_t* off, int64_t spans,
                                                      const uint32_t* bmax, int nblocks, uint32_t* sum) {
    __shared__ uint32_t sc[1024], sl[1024];
    uint32_t c = 0, l = 0;
    for (int b = threadIdx.x; b < nblocks; b += 1024) {
        c = bmax[2 * b] > c ? bmax[2 * b] : c;
        l = bmax[2 * b + 1] > l ? bmax[2 * b + 1] : l;
    }
    sc[threadIdx.x] = c;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sc[threadIdx.x] = sc[threadIdx.x] > sc[threadIdx.x + o] ? sc[threadIdx.x] : sc[threadIdx.x + o];
            sl[threadIdx.x] = sl[threadIdx.x] > sl[threadIdx.x + o] ? sl[threadIdx.x] : sl[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sum[0] = spans > 0 ? off[spans - 1] + count[spans - 1] : 0u;
        sum[1] = sc[0];
        sum[2] = sl[0];
    }
}

// slot mode -> contiguous: one wavefront per span
__global__ void k_runs_compact(const uint64_t* keys, const uint32_t* ends, const uint32_t* count,
                               const uint32_t* off, int64_t spans, uint32_t scap, uint64_t* okeys, uint32_t* oends) {
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t sp = wave0; sp < spans; sp += nwaves) {
        const uint32_t c = count[sp], o = off[sp];
        for (uint32_t i = lane; i < c; i += 64) {
            okeys[o + i] = keys[(uint64_t)sp * scap + i];
            oends[o + i] = ends[(uint64_t)sp * scap + i];
        }
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

// Voxels of a cell are visited in the reference's (z, x) order; loads go out CB at a
// time (independent), the float32 accumulation then runs over them in order.
constexpr int CB = 8;

// intensity loaders; raw(i) / cvt(r) as for the labels
struct InU16 {
    const uint16_t* p;
    using raw_t = uint32_t;
    __device__ __forceinline__ float operator()(int64_t i) const { return (float)p[i]; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static float cvt(raw_t r) { return (float)r; }
};
struct InF32 {
    const float* p;
    using raw_t = float;
    __device__ __forceinline__ float operator()(int64_t i) const { return p[i]; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static float cvt(raw_t r) { return r; }
};
struct InAny {
    DBuf b;
    using raw_t = float;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(b, i); }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return ld(b, i); }
    __device__ __forceinline__ static float cvt(raw_t r) { return r; }
};

// CB voxels of a run from x (x < xe): branch-free loads (indices past the run end
// repeat its last voxel), all in flight before the first is converted
template <typename IN>
__device__ __forceinline__ void cell_fetch(const IN& in, int64_t base, int64_t x, int64_t xe, float (&v)[CB]) {
    typename IN::raw_t r[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) r[j] = in.raw(base + (x + j < xe ? x + j : xe - 1));
    static_assert(CB == 8, "the register fence below names CB values");
    // an empty fence that reads all CB values: the compiler issues every load before
    // it (instead of sinking each one into the conditional code that uses it)
    asm volatile("" : "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]), "+v"(r[6]),
                 "+v"(r[7]));
#pragma unroll
    for (int j = 0; j < CB; ++j) v[j] = IN::cvt(r[j]);
}


// pass 1, one thread per cell: 16 float32 planes in the reference's order:
// 0-2 sum x,y,z  3 count  4-6 sum x*v, y*v, z*v  7 sum v  8/9 min/max v  10-15 min/max x,y,z
template <typename IN>
__global__ __launch_bounds__(256) void k_cells_pass1(IN in, CellGeom cg, const uint64_t* keys,
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
        uint64_t key = keys[r0];
        uint32_t x1 = ends[r0];
        for (uint32_t r = r0; r < r1; ++r) {
            int64_t y, z, x0;
            cg.decode(key, y, z, x0);
            const int64_t xe = x1;
            if (r + 1 < r1) {  // next run's bounds while this one is summed
                key = keys[r + 1];
                x1 = ends[r + 1];
            }
            const float yf = (float)y, zf = (float)z;
            const int64_t base = (z * cg.H + y) * cg.W;
            for (int64_t x = x0; x < xe; x += CB) {
                float v[CB];
                cell_fetch(in, base, x, xe, v);
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    if (x + j >= xe) break;
                    const float xf = (float)(x + j);
                    if (s[3] == 0.f) {
                        s[8] = v[j]; s[9] = v[j]; s[10] = xf; s[11] = xf; s[12] = yf; s[13] = yf; s[14] = zf; s[15] = zf;
                    } else {
                        s[8] = s[8] > v[j] ? v[j] : s[8];
                        s[9] = s[9] < v[j] ? v[j] : s[9];
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
                    s[4] = s[4] + xf * v[j];
                    s[5] = s[5] + yf * v[j];
                    s[6] = s[6] + zf * v[j];
                    s[7] = s[7] + v[j];
                }
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
template <typename IN>
__global__ __launch_bounds__(256) void k_cells_pass2(IN in, CellGeom cg, const uint64_t* keys,
                                                     const uint32_t* ends, const uint32_t* cell_start,
                                                     const uint32_t* ncells_p, int64_t cap, const float* img,
                                                     int64_t L, float* acc2) {
#pragma clang fp contract(off)
    const uint32_t ncells = *ncells_p;
    const int lshift = cg.kb.yb + cg.kb.zb + cg.kb.xb;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cap && c < ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r0 = cell_start[c], r1 = cell_start[c + 1];
        uint64_t key = keys[r0];
        uint32_t x1 = ends[r0];
        const int64_t l = (int64_t)(key >> lshift);
        const float cx = img[0 * L + l], cy = img[1 * L + l], cz = img[2 * L + l];
        const float mx = img[3 * L + l], my = img[4 * L + l], mz = img[5 * L + l];
        const float mean = img[6 * L + l], area = img[7 * L + l];
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, m4 = 0.f, m5 = 0.f;
        for (uint32_t r = r0; r < r1; ++r) {
            int64_t y, z, x0;
            cg.decode(key, y, z, x0);
            const int64_t xe = x1;
            if (r + 1 < r1) {
                key = keys[r + 1];
                x1 = ends[r + 1];
            }
            const float yf = (float)y, zf = (float)z;
            const float dy = yf - cy, dz = zf - cz, ey = yf - my, ez = zf - mz;
            const int64_t base = (z * cg.H + y) * cg.W;
            for (int64_t x = x0; x < xe; x += CB) {
                float v[CB];
                cell_fetch(in, base, x, xe, v);
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    if (x + j >= xe) break;
                    const float xf = (float)(x + j);
                    const float dx = xf - cx;
                    const float dc = sqrtf((dx * dx + dy * dy) + dz * dz);
                    const float ex = xf - mx;
                    const float dm = sqrtf((ex * ex + ey * ey) + ez * ez);
                    const float dv = v[j] - mean;
                    const float sq = (dv * dv) / area;
                    if (s3 == 0.f) { m4 = dc; m5 = dm; }
                    else { m4 = m4 < dc ? dc : m4; m5 = m5 < dm ? dm : m5; }
                    s0 = s0 + dc;
                    s1 = s1 + dm;
                    s2 = s2 + sq;
                    s3 = s3 + 1.f;
                }
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
                               const float* lab16, const float* img, int64_t L, int64_t stride, float* out) {
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
    const int64_t j = l - 1;
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
    for (int c = 0; c < CLE_STAT_COLUMNS; ++c) out[c * stride + j] = col[c];
}

static int bits_for(uint64_t maxval) {
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

// run slots per span, sized from the previous call's busiest span
static std::atomic<uint32_t> g_span_slots{32};

// 32-bit label types have the 16-byte layout; it needs rows of a multiple of 4
// labels and an aligned buffer
template <typename LAB>
constexpr bool kVecLab = std::is_same<LAB, LabT<int32_t>>::value || std::is_same<LAB, LabT<uint32_t>>::value ||
                         std::is_same<LAB, LabT<float>>::value;

// L > 0: labels 1..L-1 are measured (the caller passes max(labels) + 1).  L == 0:
// every positive label is measured and L = max(labels) + 1 comes from the run pass.
// The table holds CLE_STAT_COLUMNS rows of tcols >= L - 1 columns; a smaller table
// fails with CLE_E_SHAPE after *out_L is set, so the caller can grow it and retry.
template <typename LAB>
static int stats_run(DBuf inten, LAB lab, const Geom& g, int64_t L, float* table, int64_t tcols, int64_t* out_L,
                     hipStream_t s) {
    KeyBits kb;
    kb.xb = bits_for((uint64_t)(g.W > 0 ? g.W - 1 : 0));
    kb.zb = bits_for((uint64_t)(g.D > 0 ? g.D - 1 : 0));
    kb.yb = bits_for((uint64_t)(g.H > 0 ? g.H - 1 : 0));
    kb.lb = L > 0 ? bits_for((uint64_t)(L - 1)) : 31;  // labels are ints: at most 31 bits
    if (kb.total() > 64) return fail(CLE_E_SHAPE, "statistics: %d key bits exceed 64", kb.total());
    const int64_t spans = g.D * g.H * ((g.W + RB_SPAN - 1) / RB_SPAN);
    unsigned gr = grid_for((size_t)(spans > 0 ? spans : 1), RB_THREADS / 64);  // one wavefront per span
    if (gr > 32768) gr = 32768;
    const int64_t lcap = L > 0 ? L : (int64_t)1 << 40;
    // (1) the runs: every span writes into scap slots of its own (sized from the
    //     previous call), an exclusive scan of the span counts gives the offsets, one
    //     synchronisation reads the totals, and the slots are compacted; a span with
    //     more runs than slots reruns the pass writing at the exact offsets
    const int64_t scount = spans > 0 ? spans : 1;
    uint32_t scap = (uint32_t)g_span_slots.load(std::memory_order_relaxed);
    Scratch cnt, off, bmx, sum, k0s, e0s;
    CLE_TRY(cnt.alloc((size_t)scount * 4, s));
    CLE_TRY(off.alloc((size_t)scount * 4, s));
    CLE_TRY(bmx.alloc((size_t)gr * 8, s));
    CLE_TRY(sum.alloc(16, s));
    CLE_TRY(k0s.alloc((size_t)scount * scap * 8, s));
    CLE_TRY(e0s.alloc((size_t)scount * scap * 4, s));
    CLE_TRY(fill_u32(sum.get<uint32_t>(), 0u, 4, s));
    auto runs = [&](const uint32_t* o, uint64_t* kk, uint32_t* ee) {
        if constexpr (kVecLab<LAB>) {
            if ((g.W & 3) == 0 && ((uintptr_t)lab.p & 15) == 0) {
                k_runs_lane<<<gr, RB_THREADS, 0, s>>>(lab, g, lcap, kb, scap, o, cnt.get<uint32_t>(), kk, ee,
                                                      bmx.get<uint32_t>());
                return;
            }
        }
        k_runs<false><<<gr, RB_THREADS, 0, s>>>(lab, g, lcap, kb, scap, o, cnt.get<uint32_t>(), kk, ee,
                                                bmx.get<uint32_t>());
    };
    if (spans > 0) {
        runs(nullptr, k0s.get<uint64_t>(), e0s.get<uint32_t>());
        size_t sb = 0;
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, cnt.get<uint32_t>(), off.get<uint32_t>(), (int)spans, s));
        Scratch st;
        CLE_TRY(st.alloc(sb, s));
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(st.p, sb, cnt.get<uint32_t>(), off.get<uint32_t>(), (int)spans, s));
        k_runs_summary<<<1, 1024, 0, s>>>(cnt.get<uint32_t>(), off.get<uint32_t>(), spans, bmx.get<uint32_t>(),
                                          (int)gr, sum.get<uint32_t>());
    }
    CLE_LAUNCHED();
    uint32_t hs[3] = {0, 0, 0};
    CLE_HIP(hipMemcpyAsync(hs, sum.p, 12, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    const int64_t n = (int64_t)hs[0];
    {
        uint32_t want = 16;
        while (want < hs[1] && want < 128) want <<= 1;
        g_span_slots.store(want, std::memory_order_relaxed);
    }
    if (L <= 0) {
        L = (int64_t)hs[2] + 1;
        kb.lb = bits_for((uint64_t)(L - 1));
    }
    if (out_L) *out_L = L;
    if (tcols < L - 1)
        return fail(CLE_E_SHAPE, "statistics table holds %lld columns, %lld labels need %lld", (long long)tcols,
                    (long long)(L - 1), (long long)(L - 1));
    Scratch k0c, e0c, k1, e1, tmp, flags, cidx, cstart, ncell, lst, len_, acc, acc2, lab16, img;
    const int64_t rcap = n > 0 ? n : 1;
    CLE_TRY(k0c.alloc((size_t)rcap * 8, s));
    CLE_TRY(e0c.alloc((size_t)rcap * 4, s));
    CLE_TRY(k1.alloc((size_t)rcap * 8, s));
    CLE_TRY(e1.alloc((size_t)rcap * 4, s));
    uint64_t* k0 = k0c.get<uint64_t>();
    uint32_t* e0 = e0c.get<uint32_t>();
    if (n > 0) {
        if (hs[1] > scap) {
            runs(off.get<uint32_t>(), k0, e0);  // counts are those of the first pass
        } else {
            unsigned gc = grid_for((size_t)spans, RB_THREADS / 64);
            if (gc > 32768) gc = 32768;
            k_runs_compact<<<gc, RB_THREADS, 0, s>>>(k0s.get<uint64_t>(), e0s.get<uint32_t>(), cnt.get<uint32_t>(),
                                                     off.get<uint32_t>(), spans, scap, k0, e0);
        }
    }
    CLE_LAUNCHED();
    // (2) sort the runs by (label, y, z, x0): they are in raster (z, y, x0) order, so a
    //     stable radix sort on the (label, y) bits alone is enough
    const int bit0 = kb.zb + kb.xb;
    size_t tbytes = 0;
    CLE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, k0, k1.get<uint64_t>(), e0, e1.get<uint32_t>(),
                                               (int)rcap, bit0, kb.total(), s));
    CLE_TRY(tmp.alloc(tbytes, s));
    if (n > 0)
        CLE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, k0, k1.get<uint64_t>(), e0, e1.get<uint32_t>(),
                                                   (int)n, bit0, kb.total(), s));
    const uint64_t* keys = k1.get<uint64_t>();
    const uint32_t* ends = e1.get<uint32_t>();
    // cells and label ranges
    CLE_TRY(flags.alloc((size_t)rcap * 4, s));
    CLE_TRY(cidx.alloc((size_t)rcap * 4, s));
    CLE_TRY(cstart.alloc((size_t)(rcap + 1) * 4, s));
    CLE_TRY(ncell.alloc(4, s));
    CLE_TRY(fill_u32(ncell.get<uint32_t>(), 0u, 1, s));
    CLE_TRY(lst.alloc((size_t)L * 4, s));
    CLE_TRY(len_.alloc((size_t)L * 4, s));
    CLE_TRY(fill_u32(lst.get<uint32_t>(), 0u, (size_t)L, s));
    CLE_TRY(fill_u32(len_.get<uint32_t>(), 0u, (size_t)L, s));
    const unsigned gn = grid_for((size_t)rcap, 256) > 65536 ? 65536 : grid_for((size_t)rcap, 256);
    if (n > 0) {
        k_cell_flags<<<gn, 256, 0, s>>>(keys, n, kb.zb + kb.xb, flags.get<uint32_t>());
        size_t sbytes = 0;
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sbytes, flags.get<uint32_t>(), cidx.get<uint32_t>(), (int)n, s));
        Scratch stmp;
        CLE_TRY(stmp.alloc(sbytes, s));
        CLE_HIP(hipcub::DeviceScan::ExclusiveSum(stmp.p, sbytes, flags.get<uint32_t>(), cidx.get<uint32_t>(), (int)n, s));
        k_cell_scatter<<<gn, 256, 0, s>>>(flags.get<uint32_t>(), cidx.get<uint32_t>(), n, cstart.get<uint32_t>(),
                                          ncell.get<uint32_t>());
        k_label_ranges<<<gn, 256, 0, s>>>(keys, cstart.get<uint32_t>(), ncell.get<uint32_t>(), rcap, kb,
                                          lst.get<uint32_t>(), len_.get<uint32_t>());
    }
    CLE_LAUNCHED();
    // (3) + (4): cells, then labels, twice
    CLE_TRY(acc.alloc((size_t)rcap * 16 * 4, s));
    CLE_TRY(acc2.alloc((size_t)rcap * 6 * 4, s));
    CLE_TRY(lab16.alloc((size_t)L * 16 * 4, s));
    CLE_TRY(img.alloc((size_t)L * 8 * 4, s));
    const CellGeom cg{kb, g.H, g.W};
    auto cells = [&](auto in) {
        if (n > 0)
            k_cells_pass1<<<gn, 256, 0, s>>>(in, cg, keys, ends, cstart.get<uint32_t>(), ncell.get<uint32_t>(), rcap,
                                             acc.get<float>());
        k_labels_pass1<<<grid_for((size_t)L, 256), 256, 0, s>>>(acc.get<float>(), rcap, lst.get<uint32_t>(),
                                                                len_.get<uint32_t>(), L, lab16.get<float>(),
                                                                img.get<float>());
        if (n > 0)
            k_cells_pass2<<<gn, 256, 0, s>>>(in, cg, keys, ends, cstart.get<uint32_t>(), ncell.get<uint32_t>(), rcap,
                                             img.get<float>(), L, acc2.get<float>());
    };
    if (inten.dt == CLE_U16) cells(InU16{(const uint16_t*)inten.p});
    else if (inten.dt == CLE_F32) cells(InF32{(const float*)inten.p});
    else cells(InAny{inten});
    if (L > 1)
        k_labels_table<<<grid_for((size_t)(L - 1), 256), 256, 0, s>>>(acc2.get<float>(), rcap, lst.get<uint32_t>(),
                                                                      len_.get<uint32_t>(), lab16.get<float>(),
                                                                      img.get<float>(), L, tcols, table);
    CLE_LAUNCHED();
    return CLE_OK;
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_statistics_of_labelled_pixels(const cle_buf* intensity, const cle_buf* labels, int64_t num_labels,
                                      const cle_buf* out_table, int64_t* out_num_labels, void* stream) {
    Geom gi, gl, go;
    CLE_TRY(check_buf(intensity, "intensity_image", &gi));
    CLE_TRY(check_buf(labels, "label_image", &gl));
    CLE_TRY(check_same_shape(gi, gl, "intensity_image", "label_image"));
    CLE_TRY(check_buf(out_table, "table", &go));
    if (out_table->dtype != CLE_F32) return fail(CLE_E_TYPE, "statistics table must be float32");
    if (out_table->ndim != 2 || out_table->shape[0] != CLE_STAT_COLUMNS)
        return fail(CLE_E_SHAPE, "statistics table must have shape (%d, columns)", CLE_STAT_COLUMNS);
    if (num_labels < 0) return fail(CLE_E_VALUE, "num_labels must be >= 0");
    const int64_t tcols = out_table->shape[1];
    if (num_labels > 0 && tcols < num_labels - 1)
        return fail(CLE_E_SHAPE, "statistics table holds %lld columns < %lld labels", (long long)tcols,
                    (long long)(num_labels - 1));
    hipStream_t s = (hipStream_t)stream;
    const DBuf in = dbuf(intensity);
    float* table = (float*)out_table->data;
    int64_t* oL = out_num_labels;
    const int64_t L = num_labels;
    switch (labels->dtype) {
        case CLE_I32: return stats_run(in, LabT<int32_t>{(const int32_t*)labels->data}, gi, L, table, tcols, oL, s);
        case CLE_U32: return stats_run(in, LabT<uint32_t>{(const uint32_t*)labels->data}, gi, L, table, tcols, oL, s);
        case CLE_F32: return stats_run(in, LabT<float>{(const float*)labels->data}, gi, L, table, tcols, oL, s);
        case CLE_U16: return stats_run(in, LabT<uint16_t>{(const uint16_t*)labels->data}, gi, L, table, tcols, oL, s);
        default: return stats_run(in, LabAny{dbuf(labels)}, gi, L, table, tcols, oL, s);
    }
}

}  // extern "C"
