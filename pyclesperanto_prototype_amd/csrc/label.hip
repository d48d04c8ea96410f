// Connected-component labelling and the rank-based relabelling ops.
//
// Reference semantics (pyclesperanto_prototype/...):
//   connected_components_labeling_box  _tier4/_connected_components_labeling_box.py:15-66:
//     float32 pixel keys x*H*D + y*D + z + 1 (_tier1/_set_non_zero_pixels_to_pixel_index.py),
//     Jacobi non-zero-minimum propagation to a fixed point, then relabel_sequential.
//     The fixed point holds, in every voxel, float32(min key of its component), so
//     label k = k-th distinct float32(min key) in ascending order.
//   relabel_sequential  _tier3/_relabel_sequential.py:12-59 (rank among present values)
//   label_spots         _tier2/_label_spots.py:9-56 + label_spots_in_x.cl (raster rank)
//
// MI355X design: no iteration.  Union-find over linear indices (link to the smaller
// root with atomicMin, Playne & Hawick 2018), then each root takes the minimum
// x-major key of its component (atomicMin), the keys (optionally rounded to float32,
// reproducing the reference's merges above 2^24 voxels) are set in a bitmap over
// the key space, a popcount prefix scan gives every key its rank, and one gather
// writes the labels.  Everything stays on the device; no host synchronisation.
#include "common.hpp"

namespace cle {

constexpr uint32_t NONE = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t find_root(const uint32_t* L, uint32_t a) {
    uint32_t p = ld_relaxed(L + a);
    while (p != a) {
        a = p;
        p = ld_relaxed(L + a);
    }
    return a;
}

__device__ __forceinline__ void unite(uint32_t* L, uint32_t a, uint32_t b) {
    bool done;
    do {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a < b) {
            const uint32_t old = atomicMin(L + b, a);
            done = (old == b);
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(L + a, b);
            done = (old == a);
            a = old;
        } else {
            done = true;
        }
    } while (!done);
}

__global__ void k_ccl_init(DBuf bin, uint32_t* L, uint32_t* K, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        L[i] = nz(bin, i) ? (uint32_t)i : NONE;
        K[i] = NONE;
    }
}

// unite every foreground voxel with its foreground neighbours that come earlier in
// linear order (13 of the 26 box neighbours in 3D, 4 of 8 in 2D; diamond: 3 / 2)
__global__ void k_ccl_merge(uint32_t* L, Geom g, int diamond) {
    const int64_t n = g.n();
    const int zr = g.ndim == 3 ? 1 : 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (L[i] == NONE) continue;
        const int64_t x = i % g.W, y = (i / g.W) % g.H, z = i / (g.W * g.H);
        for (int dz = -zr; dz <= 0; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;
                    if (diamond && (abs(dx) + abs(dy) + abs(dz)) != 1) continue;
                    const int64_t qx = x + dx, qy = y + dy, qz = z + dz;
                    if (qx < 0 || qx >= g.W || qy < 0 || qy >= g.H || qz < 0) continue;
                    const int64_t q = (qz * g.H + qy) * g.W + qx;
                    if (ld_relaxed(L + q) == NONE) continue;
                    unite(L, (uint32_t)i, (uint32_t)q);
                }
    }
}

// path compression + minimum x-major key per root
__global__ void k_ccl_roots(uint32_t* L, uint32_t* K, Geom g) {
    const int64_t n = g.n();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (L[i] == NONE) continue;
        const uint32_t r = find_root(L, (uint32_t)i);
        const int64_t x = i % g.W, y = (i / g.W) % g.H, z = i / (g.W * g.H);
        const uint32_t key = (uint32_t)(x * g.H * g.D + y * g.D + z);  // 0-based
        L[i] = r;
        atomicMin(K + r, key);
    }
}

// 1-based key -> bit position; with fp32 compatibility the key is rounded to the
// nearest float32 first (the reference holds keys in float32 images)
__device__ __forceinline__ uint64_t key_bit(uint32_t key0, int compat) {
    const uint64_t k = (uint64_t)key0 + 1;
    return compat ? (uint64_t)(float)k : k;
}

__global__ void k_ccl_mark(const uint32_t* L, const uint32_t* K, int64_t n, int compat, uint32_t* bm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (L[i] != (uint32_t)i) continue;  // roots only (background holds NONE)
        const uint64_t b = key_bit(K[i], compat);
        atomicOr(bm + (b >> 5), 1u << (b & 31));
    }
}

// ----------------------------------------------------------------- bitmap ranks
constexpr int SB = 256;             // threads per scan block
constexpr int SW = 4;               // words per thread
constexpr int SWORDS = SB * SW;     // words per scan block

__global__ __launch_bounds__(SB) void k_bm_blocksum(const uint32_t* bm, int64_t nw, uint32_t* bsum) {
    const int64_t w0 = (int64_t)blockIdx.x * SWORDS;
    uint32_t c = 0;
    for (int k = 0; k < SW; ++k) {
        const int64_t w = w0 + threadIdx.x * SW + k;
        if (w < nw) c += __popc(bm[w]);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    __shared__ uint32_t sm[SB / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < SB / 64; ++k) t += sm[k];
        bsum[blockIdx.x] = t;
    }
}

// exclusive scan of the block sums (single block, chunked with a carry); the
// grand total goes to *total if requested
__global__ __launch_bounds__(SB) void k_scan_bsum(uint32_t* bsum, int64_t nb, uint32_t* total) {
    __shared__ uint32_t sm[SB];
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += SB) {
        const int64_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0;
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < SB; off <<= 1) {
            const uint32_t t = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sm[threadIdx.x] - v;
        const uint32_t chunk = sm[SB - 1];
        __syncthreads();
        carry += chunk;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// prefix[w] = number of set bits in words [0, w)
__global__ __launch_bounds__(SB) void k_bm_prefix(const uint32_t* bm, int64_t nw, const uint32_t* bsum,
                                                 uint32_t* prefix) {
    __shared__ uint32_t sm[SB];
    const int64_t w0 = (int64_t)blockIdx.x * SWORDS + threadIdx.x * SW;
    uint32_t c[SW];
    uint32_t tot = 0;
    for (int k = 0; k < SW; ++k) {
        c[k] = (w0 + k < nw) ? __popc(bm[w0 + k]) : 0;
        tot += c[k];
    }
    sm[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < SB; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = bsum[blockIdx.x] + sm[threadIdx.x] - tot;
    for (int k = 0; k < SW; ++k) {
        if (w0 + k < nw) prefix[w0 + k] = run;
        run += c[k];
    }
}

// 1-based rank of bit b (number of set bits at positions <= b)
__device__ __forceinline__ uint32_t bit_rank(const uint32_t* bm, const uint32_t* prefix, uint64_t b) {
    const uint64_t w = b >> 5;
    const uint32_t m = (2u << (uint32_t)(b & 31)) - 1u;
    return prefix[w] + __popc(bm[w] & m);
}

// A bitmap over [0, nbits) with its popcount prefix
struct RankTable {
    Scratch bm, prefix, bsum;
    int64_t nw = 0, nb = 0;
    int init(uint64_t nbits, hipStream_t s) {
        nw = (int64_t)((nbits + 31) / 32);
        if (nw < 1) nw = 1;
        nb = (nw + SWORDS - 1) / SWORDS;
        CLE_TRY(bm.alloc((size_t)nw * 4, s));
        CLE_TRY(prefix.alloc((size_t)nw * 4, s));
        CLE_TRY(bsum.alloc((size_t)nb * 4, s));
        return fill_u32(bm.get<uint32_t>(), 0u, (size_t)nw, s);
    }
    int scan(uint32_t* total, hipStream_t s) {
        k_bm_blocksum<<<(unsigned)nb, SB, 0, s>>>(bm.get<uint32_t>(), nw, bsum.get<uint32_t>());
        k_scan_bsum<<<1, SB, 0, s>>>(bsum.get<uint32_t>(), nb, total);
        k_bm_prefix<<<(unsigned)nb, SB, 0, s>>>(bm.get<uint32_t>(), nw, bsum.get<uint32_t>(), prefix.get<uint32_t>());
        CLE_LAUNCHED();
        return CLE_OK;
    }
};

__global__ void k_ccl_write(const uint32_t* L, const uint32_t* K, int64_t n, int compat, const uint32_t* bm,
                            const uint32_t* prefix, DBuf dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = L[i];
        float lab = 0.f;
        if (r != NONE) lab = (float)bit_rank(bm, prefix, key_bit(K[r], compat));
        st(dst, i, lab);
    }
}

static unsigned lg(int64_t n) {
    unsigned g = grid_for((size_t)n, 256);
    return g > 131072 ? 131072 : g;
}

int ccl_impl(DBuf binary, const Geom& g, DBuf dst, int connectivity, int options, uint32_t* n_labels_dev,
             hipStream_t s) {
    const int64_t n = g.n();
    if (n == 0) return n_labels_dev ? fill_u32(n_labels_dev, 0u, 1, s) : CLE_OK;
    const int compat = (options & CLE_FP32_KEY_COMPAT) ? 1 : 0;
    Scratch L, K;
    CLE_TRY(L.alloc((size_t)n * 4, s));
    CLE_TRY(K.alloc((size_t)n * 4, s));
    uint32_t* l = L.get<uint32_t>();
    uint32_t* k = K.get<uint32_t>();
    k_ccl_init<<<lg(n), 256, 0, s>>>(binary, l, k, n);
    k_ccl_merge<<<lg(n), 256, 0, s>>>(l, g, connectivity == CLE_CONN_DIAMOND);
    k_ccl_roots<<<lg(n), 256, 0, s>>>(l, k, g);
    CLE_LAUNCHED();
    // bit positions run to float32(n) (key n rounded up at most to the next float)
    const uint64_t maxbit = compat ? (uint64_t)(float)(uint64_t)n : (uint64_t)n;
    RankTable rt;
    CLE_TRY(rt.init(maxbit + 1, s));
    k_ccl_mark<<<lg(n), 256, 0, s>>>(l, k, n, compat, rt.bm.get<uint32_t>());
    CLE_LAUNCHED();
    CLE_TRY(rt.scan(n_labels_dev, s));
    k_ccl_write<<<lg(n), 256, 0, s>>>(l, k, n, compat, rt.bm.get<uint32_t>(), rt.prefix.get<uint32_t>(), dst);
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- relabel / spots
__global__ void k_mark_values(DBuf src, int64_t n, uint32_t* bm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = ld(src, i);
        const int64_t iv = (int64_t)v;
        if (iv > 0) atomicOr(bm + (iv >> 5), 1u << (iv & 31));
    }
}

__global__ void k_write_value_ranks(DBuf src, int64_t n, const uint32_t* bm, const uint32_t* prefix, DBuf dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iv = (int64_t)ld(src, i);
        st(dst, i, iv > 0 ? (float)bit_rank(bm, prefix, (uint64_t)iv) : 0.f);
    }
}

__global__ void k_mark_nonzero(DBuf src, int64_t n, uint32_t* bm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        // one wavefront of 64 lanes covers 2 bitmap words: build them with ballots
        const bool f = nz(src, i);
        if (f) atomicOr(bm + (i >> 5), 1u << (i & 31));
    }
}

__global__ void k_write_position_ranks(DBuf src, int64_t n, const uint32_t* bm, const uint32_t* prefix, DBuf dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst, i, nz(src, i) ? (float)bit_rank(bm, prefix, (uint64_t)i) : 0.f);
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_connected_components_labeling(const cle_buf* binary, const cle_buf* dst, int connectivity, int options,
                                      uint32_t* n_labels_dev, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(binary, "binary_input", &g));
    CLE_TRY(check_buf(dst, "labeling_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "binary_input", "labeling_destination"));
    if (connectivity != CLE_CONN_BOX && connectivity != CLE_CONN_DIAMOND)
        return fail(CLE_E_VALUE, "unknown connectivity %d", connectivity);
    return ccl_impl(dbuf(binary), g, dbuf(dst), connectivity, options, n_labels_dev, (hipStream_t)stream);
}

int cle_relabel_sequential(const cle_buf* src, const cle_buf* dst, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "output", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "output"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch mm;
    CLE_TRY(mm.alloc(8, s));
    CLE_TRY(minmax_impl(dbuf(src), n, mm.get<float>(), s));
    float hmm[2];
    CLE_HIP(hipMemcpyAsync(hmm, mm.p, 8, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    const double mx = hmm[1];
    if (!(mx < 4294967295.0)) return fail(CLE_E_VALUE, "relabel_sequential: maximum value %g out of range", mx);
    const int64_t maxv = mx > 0 ? (int64_t)mx : 0;
    RankTable rt;
    CLE_TRY(rt.init((uint64_t)maxv + 1, s));
    k_mark_values<<<lg(n), 256, 0, s>>>(dbuf(src), n, rt.bm.get<uint32_t>());
    CLE_LAUNCHED();
    CLE_TRY(rt.scan(nullptr, s));
    k_write_value_ranks<<<lg(n), 256, 0, s>>>(dbuf(src), n, rt.bm.get<uint32_t>(), rt.prefix.get<uint32_t>(),
                                              dbuf(dst));
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_label_spots(const cle_buf* src, const cle_buf* dst, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "input_spots", &g));
    CLE_TRY(check_buf(dst, "labelled_spots_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "input_spots", "labelled_spots_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    RankTable rt;
    CLE_TRY(rt.init((uint64_t)n, s));
    k_mark_nonzero<<<lg(n), 256, 0, s>>>(dbuf(src), n, rt.bm.get<uint32_t>());
    CLE_LAUNCHED();
    CLE_TRY(rt.scan(nullptr, s));
    k_write_position_ranks<<<lg(n), 256, 0, s>>>(dbuf(src), n, rt.bm.get<uint32_t>(), rt.prefix.get<uint32_t>(),
                                                 dbuf(dst));
    CLE_LAUNCHED();
    return CLE_OK;
}

}  // extern "C"
