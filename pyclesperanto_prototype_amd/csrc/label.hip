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
// MI355X design: no iteration.  Union-find (link to the smaller root with atomicMin,
// Playne & Hawick 2018), first inside 2048-voxel tiles in LDS, then across tile faces
// in global memory; every component's root carries the minimum x-major key of the
// component, the keys (optionally rounded to float32, reproducing the reference's
// merges above 2^24 voxels) are set in a bitmap over the key space, a popcount prefix
// scan gives every key its rank, and one gather writes the labels.  Everything stays
// on the device; no host synchronisation.
#include <type_traits>

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

// ----------------------------------------------------------------- tiled union-find
// Tiles of TZ x TY x TX voxels (compile-time: 8 x 16 x 32 in 3D, 1 x 64 x 64 in 2D,
// 4096 voxels, 512 threads x 8 voxels).
// Pass 1 (k_ccl_local): per tile, union-find in LDS over the tile's foreground
//   (links to the smallest local index), then L[i] = global index of i's local root
//   for every foreground voxel, K[local root] = the component's smallest 0-based
//   x-major key inside the tile, and the tile's local roots go to its root list.
// Pass 2 (k_ccl_border): foreground voxels on a tile face unite, in global memory,
//   with their foreground neighbours (13 earlier directions) in other tiles, skipping
//   neighbours whose local root repeats the previous one.
// Pass 3 (k_ccl_resolve): every local root finds its global root r, sets
//   L[local root] = r and lowers K[r] to its own key.
// Pass 4 (k_ccl_mark_roots): global roots set their key bit.
// Pass 5 (k_ccl_write): label = rank(key bit of K[L[L[i]]]).
constexpr int CT_THREADS = 512;

template <int TZ_, int TY_, int TX_>
struct TileShape {
    static constexpr int TZ = TZ_, TY = TY_, TX = TX_, N = TZ_ * TY_ * TX_;
    static constexpr int PER = N / CT_THREADS;
    static constexpr int ROOTS = N / 2;  // a tile holds at most N/2 components
};
using Tile3D = TileShape<8, 16, 32>;
using Tile2D = TileShape<1, 64, 64>;

struct TileGrid {
    int64_t nz, ny, nx;
    __host__ __device__ int64_t count() const { return nz * ny * nx; }
};

template <typename TS>
static TileGrid tile_grid(const Geom& g) {
    return TileGrid{(g.D + TS::TZ - 1) / TS::TZ, (g.H + TS::TY - 1) / TS::TY, (g.W + TS::TX - 1) / TS::TX};
}

__device__ __forceinline__ uint32_t lds_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t lfind(uint32_t* lab, uint32_t a) {
    uint32_t p = lds_load(lab + a);
    while (p != a) {
        a = p;
        p = lds_load(lab + a);
    }
    return a;
}

__device__ __forceinline__ void lunite(uint32_t* lab, uint32_t a, uint32_t b) {
    bool done;
    do {
        a = lfind(lab, a);
        b = lfind(lab, b);
        if (a < b) {
            const uint32_t old = atomicMin(lab + b, a);
            done = (old == b);
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(lab + a, b);
            done = (old == a);
            a = old;
        } else {
            done = true;
        }
    } while (!done);
}

template <typename TS, typename FG>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_local(FG fg, Geom g, TileGrid T, int diamond, uint32_t* L,
                                                         uint32_t* K, uint32_t* roots, uint32_t* nroots) {
    constexpr int N = TS::N, PER = TS::PER, TX = TS::TX, TY = TS::TY, TZ = TS::TZ;
    __shared__ uint32_t lab[N];
    __shared__ uint32_t kmin[N];
    __shared__ uint32_t cnt;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const int64_t x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int zr = g.ndim == 3 ? 1 : 0;
    if (threadIdx.x == 0) cnt = 0;
    int64_t gi[PER];
    bool f[PER];
    bool anyf = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        const int64_t x = x0 + lx, y = y0 + ly, z = z0 + lz;
        const bool in = x < g.W && y < g.H && z < g.D;
        gi[k] = in ? (z * g.H + y) * g.W + x : 0;
        f[k] = in && fg(gi[k]);
        anyf |= f[k];
        lab[li] = f[k] ? (uint32_t)li : NONE;
        kmin[li] = NONE;
    }
    if (!__syncthreads_or(anyf ? 1 : 0)) {
        if (threadIdx.x == 0) nroots[tile] = 0;
        return;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!f[k]) continue;
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        // box: when the left neighbour (same row) is foreground it has already been
        // linked to the dx = -1 and dx = 0 neighbours of the other rows, so only the
        // dx = +1 column is new
        const bool left_fg = !diamond && lx > 0 && lds_load(lab + li - 1) != NONE;
        for (int dz = -zr; dz <= 0; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;
                    if (diamond && (abs(dx) + abs(dy) + abs(dz)) != 1) continue;
                    if (left_fg && dx != 1 && !(dz == 0 && dy == 0)) continue;
                    const int qx = lx + dx, qy = ly + dy, qz = lz + dz;
                    if (qx < 0 || qx >= TX || qy < 0 || qy >= TY || qz < 0) continue;
                    const int lq = (qz * TY + qy) * TX + qx;
                    if (lds_load(lab + lq) == NONE) continue;
                    lunite(lab, (uint32_t)li, (uint32_t)lq);
                }
    }
    __syncthreads();
    uint32_t r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        r[k] = f[k] ? lfind(lab, (uint32_t)li) : NONE;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!f[k]) continue;
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        const int64_t x = x0 + lx, y = y0 + ly, z = z0 + lz;
        atomicMin(kmin + r[k], (uint32_t)(x * g.H * g.D + y * g.D + z));
        const int rx = (int)(r[k] % TX), ry = (int)((r[k] / TX) % TY), rz = (int)(r[k] / (TX * TY));
        L[gi[k]] = (uint32_t)(((z0 + rz) * g.H + (y0 + ry)) * g.W + (x0 + rx));
    }
    __syncthreads();
    uint32_t* myroots = roots + tile * TS::ROOTS;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        if (f[k] && r[k] == (uint32_t)li) {
            K[gi[k]] = kmin[li];
            const uint32_t slot = atomicAdd(&cnt, 1u);
            myroots[slot] = (uint32_t)gi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) nroots[tile] = cnt;
}

// foreground voxels on the faces of each tile unite with their foreground neighbours
// (in the 13 earlier directions) that lie in another tile
template <typename TS, typename FG>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_border(FG fg, Geom g, TileGrid T, int diamond, uint32_t* L,
                                                          const uint32_t* nroots) {
    constexpr int N = TS::N, TX = TS::TX, TY = TS::TY, TZ = TS::TZ;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    if (nroots[tile] == 0) return;  // no foreground in this tile
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const int64_t x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int zr = g.ndim == 3 ? 1 : 0;
    for (int li = threadIdx.x; li < N; li += CT_THREADS) {
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        // pairs with a neighbour in another tile: low z/y/x faces, and the high y/x
        // faces (their dz = -1 / dy = -1 neighbours reach into the next y / x tile)
        if (!((zr && lz == 0) || ly == 0 || lx == 0 || ly == TY - 1 || lx == TX - 1)) continue;
        const int64_t x = x0 + lx, y = y0 + ly, z = z0 + lz;
        if (x >= g.W || y >= g.H || z >= g.D) continue;
        const int64_t i = (z * g.H + y) * g.W + x;
        if (!fg(i)) continue;
        const uint32_t li_root = ld_relaxed(L + i);
        uint32_t last = NONE;
        // box: a foreground left neighbour (in any tile) covers the dx = -1, 0 columns
        const bool left_fg = !diamond && x > 0 && fg(i - 1);
        for (int dz = -zr; dz <= 0; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;
                    if (diamond && (abs(dx) + abs(dy) + abs(dz)) != 1) continue;
                    if (left_fg && dx != 1 && !(dz == 0 && dy == 0)) continue;
                    // only neighbours outside this tile
                    if (lz + dz >= 0 && ly + dy >= 0 && lx + dx >= 0 && lx + dx < TX && ly + dy < TY) continue;
                    const int64_t qx = x + dx, qy = y + dy, qz = z + dz;
                    if (qx < 0 || qx >= g.W || qy < 0 || qy >= g.H || qz < 0) continue;
                    const int64_t q = (qz * g.H + qy) * g.W + qx;
                    if (!fg(q)) continue;
                    const uint32_t lq = ld_relaxed(L + q);
                    if (lq == last) continue;  // same local component as the previous neighbour
                    last = lq;
                    unite(L, li_root, lq);
                }
    }
}

// Kernels over the local roots of every tile: one wavefront per tile, RT_TPB tiles per
// block (most tiles hold a few roots; a block per tile made these launches
// dispatch-bound); consecutive tiles stay on one XCD.
constexpr int RT_TPB = CT_THREADS / 64;
__device__ __forceinline__ int64_t root_tile() {
    return xcd_tile(blockIdx.x, gridDim.x) * RT_TPB + (threadIdx.x >> 6);
}
static unsigned root_grid(int64_t nt) { return (unsigned)((nt + RT_TPB - 1) / RT_TPB); }

template <typename TS>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_resolve(uint32_t* L, uint32_t* K, const uint32_t* roots,
                                                            const uint32_t* nroots, int64_t nt) {
    const int64_t tile = root_tile();
    if (tile >= nt) return;
    const uint32_t n = nroots[tile];
    for (uint32_t k = threadIdx.x & 63; k < n; k += 64) {
        const uint32_t lr = roots[tile * TS::ROOTS + k];
        const uint32_t r = find_root(L, lr);
        if (r != lr) {
            atomicMin(K + r, K[lr]);
            L[lr] = r;
        }
    }
}

template <typename TS>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_mark_roots(const uint32_t* L, const uint32_t* K,
                                                               const uint32_t* roots, const uint32_t* nroots,
                                                               int compat, uint32_t* bm, int64_t nt) {
    const int64_t tile = root_tile();
    if (tile >= nt) return;
    const uint32_t n = nroots[tile];
    for (uint32_t k = threadIdx.x & 63; k < n; k += 64) {
        const uint32_t lr = roots[tile * TS::ROOTS + k];
        if (L[lr] != lr) continue;  // not a global root
        const uint64_t b = (uint64_t)K[lr] + 1;
        const uint64_t bit = compat ? (uint64_t)(float)b : b;
        atomicOr(bm + (bit >> 5), 1u << (bit & 31));
    }
}

struct FgAny {
    DBuf b;
    __device__ __forceinline__ bool operator()(int64_t i) const { return nz(b, i); }
};
struct FgU8 {
    const uint8_t* p;
    __device__ __forceinline__ bool operator()(int64_t i) const { return p[i] != 0; }
};

// 1-based key -> bit position; with fp32 compatibility the key is rounded to the
// nearest float32 first (the reference holds keys in float32 images)
__device__ __forceinline__ uint64_t key_bit(uint32_t key0, int compat) {
    const uint64_t k = (uint64_t)key0 + 1;
    return compat ? (uint64_t)(float)k : k;
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

// exclusive scan of the block sums (single block, chunks of SB x 8 sums with a
// carry): every thread adds its 8 consecutive sums, a wavefront scans the thread sums
// with shuffles, one LDS word per wavefront joins the wavefronts; the grand total goes
// to *total if requested
__global__ __launch_bounds__(SB) void k_scan_bsum(uint32_t* bsum, int64_t nb, uint32_t* total) {
    constexpr int PT = 8;
    __shared__ uint32_t wsum[SB / 64];
    const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += (int64_t)SB * PT) {
        const int64_t i0 = base + (int64_t)threadIdx.x * PT;
        uint32_t v[PT];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < PT; ++k) {
            v[k] = i0 + k < nb ? bsum[i0 + k] : 0u;
            sum += v[k];
        }
        uint32_t x = sum;  // inclusive scan over the wavefront
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        uint32_t before = 0, chunk = 0;
#pragma unroll
        for (int w = 0; w < SB / 64; ++w) {
            const uint32_t t = wsum[w];
            before += w < wid ? t : 0u;
            chunk += t;
        }
        uint32_t run = carry + before + (x - sum);
#pragma unroll
        for (int k = 0; k < PT; ++k) {
            if (i0 + k < nb) bsum[i0 + k] = run;
            run += v[k];
        }
        carry += chunk;
        __syncthreads();  // wsum is rewritten by the next chunk
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
    static int64_t words(uint64_t nbits) {
        const int64_t w = (int64_t)((nbits + 31) / 32);
        return w < 1 ? 1 : w;
    }
    static size_t bytes(uint64_t nbits) {
        const int64_t w = words(nbits);
        return 2 * scratch_round((size_t)w * 4) + scratch_round((size_t)((w + SWORDS - 1) / SWORDS) * 4);
    }
    int init(uint64_t nbits, hipStream_t s) {
        nw = words(nbits);
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

// final label of every local root: F[lr] = rank of the key bit of its global root
template <typename TS>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_final(const uint32_t* L, const uint32_t* K, const uint32_t* roots,
                                                          const uint32_t* nroots, int compat, const uint32_t* bm,
                                                          const uint32_t* prefix, uint32_t* F, int64_t nt) {
    const int64_t tile = root_tile();
    if (tile >= nt) return;
    const uint32_t n = nroots[tile];
    for (uint32_t k = threadIdx.x & 63; k < n; k += 64) {
        const uint32_t lr = roots[tile * TS::ROOTS + k];
        F[lr] = bit_rank(bm, prefix, key_bit(K[L[lr]], compat));
    }
}

// store4 writes voxels i..i+3 with one 16-byte store (i a multiple of 4, p 16-byte aligned)
struct StLabI32 {
    static constexpr bool kCoords = false;
    int32_t* p;
    __device__ __forceinline__ static int32_t cv(float v) { return sat_cast<int32_t>(v); }
    __device__ __forceinline__ void operator()(int64_t i, float v) const { p[i] = cv(v); }
    __device__ __forceinline__ void store4(int64_t i, float a, float b, float c, float d) const {
        *(int4*)(p + i) = make_int4(cv(a), cv(b), cv(c), cv(d));
    }
};
struct StLabU32 {
    static constexpr bool kCoords = false;
    uint32_t* p;
    __device__ __forceinline__ static uint32_t cv(float v) { return sat_cast<uint32_t>(v); }
    __device__ __forceinline__ void operator()(int64_t i, float v) const { p[i] = cv(v); }
    __device__ __forceinline__ void store4(int64_t i, float a, float b, float c, float d) const {
        *(uint4*)(p + i) = make_uint4(cv(a), cv(b), cv(c), cv(d));
    }
};
struct StLabF32 {
    static constexpr bool kCoords = false;
    float* p;
    __device__ __forceinline__ void operator()(int64_t i, float v) const { p[i] = v; }
    __device__ __forceinline__ void store4(int64_t i, float a, float b, float c, float d) const {
        *(float4*)(p + i) = make_float4(a, b, c, d);
    }
};
struct StLabAny {
    static constexpr bool kCoords = false;
    DBuf b;
    __device__ __forceinline__ void operator()(int64_t i, float v) const { st(b, i, v); }
    __device__ __forceinline__ void store4(int64_t i, float a, float b_, float c, float d) const {
        st(b, i, a);
        st(b, i + 1, b_);
        st(b, i + 2, c);
        st(b, i + 3, d);
    }
};
// growth init instead of labels (GrowInitOut); put / put4 take the voxel coordinates
struct StGrowInit {
    static constexpr bool kCoords = true;
    GrowInitOut o;
    __device__ __forceinline__ int64_t tile(int64_t x, int64_t y, int64_t z) const {
        return ((z / o.tz) * o.ny + y / o.ty) * o.nx + x / o.tx;
    }
    __device__ __forceinline__ bool empty(int64_t x, int64_t y, int64_t z) const {
        return o.skip_empty && o.settled[((z >> 3) * o.ny + (y >> 3)) * o.nx + (x >> 4)] == 2;
    }
    // (k_bccl_write tests empty() once for the voxels of a block pair before put / put4)
    __device__ __forceinline__ void put(int64_t i, int64_t x, int64_t y, int64_t z, float v) const {
        const uint8_t m = o.mask[i];
        const float w = ((float)m - 1.f) + v;
        o.A[i] = w;
        o.B[i] = w;
        if (w > 0.f) o.seed[tile(x, y, z)] = -1;
        if (m) o.settled[tile(x, y, z)] = 0;  // a tile with a mask voxel is not settled
    }
    // voxels x .. x + 3 (x a multiple of 4: one growth tile, whose x extent is a
    // multiple of 4)
    __device__ __forceinline__ void put4(int64_t i, int64_t x, int64_t y, int64_t z, float a, float b, float c,
                                         float d) const {
        const uchar4 m = *(const uchar4*)(o.mask + i);
        const float4 w = make_float4(((float)m.x - 1.f) + a, ((float)m.y - 1.f) + b, ((float)m.z - 1.f) + c,
                                     ((float)m.w - 1.f) + d);
        *(float4*)(o.A + i) = w;
        *(float4*)(o.B + i) = w;
        if (w.x > 0.f || w.y > 0.f || w.z > 0.f || w.w > 0.f) o.seed[tile(x, y, z)] = -1;
        if ((m.x | m.y | m.z | m.w) != 0) o.settled[tile(x, y, z)] = 0;
    }
};

// labels are ranks (>= 0) held as float, as the reference stores them, then
// converted with saturation to the destination type; a foreground voxel's label is
// F[its local root]
template <typename FG, typename ST>
__global__ void k_ccl_write(FG fg, const uint32_t* L, const uint32_t* F, int64_t n, ST out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out(i, fg(i) ? (float)F[L[i]] : 0.f);
}

// uint8 binary, 32-bit destination: four voxels per thread (one 4-byte load; one
// 16-byte store of zeros when all four are background)
template <typename T, typename ST>
__global__ void k_ccl_write_u8x4(const uint8_t* bin, const uint32_t* L, const uint32_t* F, int64_t n, T* dst, ST out) {
    const int64_t nq = n / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b4 = ((const uint32_t*)bin)[q];
        if (b4 == 0) {
            ((uint4*)dst)[q] = make_uint4(0u, 0u, 0u, 0u);
            continue;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t i = q * 4 + k;
            out(i, ((b4 >> (8 * k)) & 0xFFu) ? (float)F[L[i]] : 0.f);
        }
    }
    for (int64_t i = nq * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out(i, bin[i] ? (float)F[L[i]] : 0.f);
}

template <typename FG>
static void launch_ccl_write(FG fg, const uint32_t* L, const uint32_t* F, int64_t n, DBuf dst, unsigned grid,
                             hipStream_t s) {
    switch (dst.dt) {
        case CLE_I32: k_ccl_write<<<grid, 256, 0, s>>>(fg, L, F, n, StLabI32{(int32_t*)dst.p}); break;
        case CLE_U32: k_ccl_write<<<grid, 256, 0, s>>>(fg, L, F, n, StLabU32{(uint32_t*)dst.p}); break;
        case CLE_F32: k_ccl_write<<<grid, 256, 0, s>>>(fg, L, F, n, StLabF32{(float*)dst.p}); break;
        default: k_ccl_write<<<grid, 256, 0, s>>>(fg, L, F, n, StLabAny{dst}); break;
    }
}

static void launch_ccl_write(FgU8 fg, const uint32_t* L, const uint32_t* F, int64_t n, DBuf dst, unsigned grid,
                             hipStream_t s) {
    const bool aligned = ((uintptr_t)fg.p & 3) == 0 && ((uintptr_t)dst.p & 15) == 0;
    if (aligned && dst.dt == CLE_I32)
        k_ccl_write_u8x4<<<grid, 256, 0, s>>>(fg.p, L, F, n, (int32_t*)dst.p, StLabI32{(int32_t*)dst.p});
    else if (aligned && dst.dt == CLE_U32)
        k_ccl_write_u8x4<<<grid, 256, 0, s>>>(fg.p, L, F, n, (uint32_t*)dst.p, StLabU32{(uint32_t*)dst.p});
    else if (aligned && dst.dt == CLE_F32)
        k_ccl_write_u8x4<<<grid, 256, 0, s>>>(fg.p, L, F, n, (float*)dst.p, StLabF32{(float*)dst.p});
    else
        launch_ccl_write<FgU8>(fg, L, F, n, dst, grid, s);
}

static unsigned lg(int64_t n) {
    unsigned g = grid_for((size_t)n, 256);
    return g > 131072 ? 131072 : g;
}

template <typename TS, typename FG>
static int ccl_run(FG fg, const Geom& g, DBuf dst, int connectivity, int options, uint32_t* n_labels_dev,
                   hipStream_t s) {
    const int64_t n = g.n();
    const int compat = (options & CLE_FP32_KEY_COMPAT) ? 1 : 0;
    const int diamond = connectivity == CLE_CONN_DIAMOND;
    const TileGrid T = tile_grid<TS>(g);
    const int64_t nt = T.count();
    Scratch Ls, Ks, Rs, Ns, Fs;
    CLE_TRY(Ls.alloc((size_t)n * 4, s));
    CLE_TRY(Ks.alloc((size_t)n * 4, s));
    CLE_TRY(Fs.alloc((size_t)n * 4, s));
    CLE_TRY(Rs.alloc((size_t)nt * TS::ROOTS * 4, s));
    CLE_TRY(Ns.alloc((size_t)nt * 4, s));
    uint32_t *l = Ls.get<uint32_t>(), *k = Ks.get<uint32_t>(), *roots = Rs.get<uint32_t>(), *nr = Ns.get<uint32_t>();
    k_ccl_local<TS><<<(unsigned)nt, CT_THREADS, 0, s>>>(fg, g, T, diamond, l, k, roots, nr);
    k_ccl_border<TS><<<(unsigned)nt, CT_THREADS, 0, s>>>(fg, g, T, diamond, l, nr);
    k_ccl_resolve<TS><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, nt);
    CLE_LAUNCHED();
    // bit positions run to float32(n) (key n rounded up at most to the next float)
    const uint64_t maxbit = compat ? (uint64_t)(float)(uint64_t)n : (uint64_t)n;
    RankTable rt;
    CLE_TRY(rt.init(maxbit + 1, s));
    k_ccl_mark_roots<TS><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, compat, rt.bm.get<uint32_t>(), nt);
    CLE_LAUNCHED();
    CLE_TRY(rt.scan(n_labels_dev, s));
    k_ccl_final<TS><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, compat, rt.bm.get<uint32_t>(),
                                                          rt.prefix.get<uint32_t>(), Fs.get<uint32_t>(), nt);
    launch_ccl_write(fg, l, Fs.get<uint32_t>(), n, dst, lg(n / 4 + 1), s);
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- 2x2x2-block union-find (box)
// With box (26-) connectivity every pair of voxels inside one 2 x 2 x 2 block (2 x 2 in
// 2D) is adjacent, so the foreground voxels of a block always form one component and
// the union-find can run over blocks: 8x fewer elements than voxels.  A block carries
// an 8-bit mask of its foreground voxels, bit (dz * 4 + dy * 2 + dx).  Two blocks at
// offset e are connected iff the block has a foreground voxel on its side facing the
// neighbour (coordinate 0 on an axis where e = -1, 1 where e = +1, any where e = 0) and
// the neighbour one on its side facing back: any two such voxels are adjacent.  The
// component key of a block is the smallest x-major key of its foreground voxels.
// Tiles of TBZ x TBY x TBX blocks run the same two-level scheme as the voxel version
// (LDS union-find per tile, then unions across tile faces in global memory); the
// final gather gives each foreground voxel the label of its block's local root.
template <int TBZ_, int TBY_, int TBX_, int BZ_>
struct BlkTile {
    static constexpr int TZ = TBZ_, TY = TBY_, TX = TBX_, BZ = BZ_, N = TBZ_ * TBY_ * TBX_;
    static constexpr int PER = N / CT_THREADS;
    static constexpr int ROOTS = N;  // blocks side by side can still be disconnected
};
using BlkTile3D = BlkTile<4, 8, 16, 2>;   // 8 x 16 x 32 voxels
using BlkTile2D = BlkTile<1, 32, 32, 1>;  // 64 x 64 pixels

struct BGrid {
    int64_t D, H, W;  // blocks along z, y, x
};

// mask bits on the side of a block facing offset e along one axis
__device__ __forceinline__ uint32_t side_bits(int e, uint32_t lo, uint32_t hi) {
    return e < 0 ? lo : (e > 0 ? hi : 0xFFu);
}
__device__ __forceinline__ uint32_t facing(int ex, int ey, int ez) {
    return side_bits(ex, 0x55u, 0xAAu) & side_bits(ey, 0x33u, 0xCCu) & side_bits(ez, 0x0Fu, 0xF0u);
}
__device__ __forceinline__ bool blocks_touch(uint32_t m, uint32_t qm, int ex, int ey, int ez) {
    return (m & facing(ex, ey, ez)) != 0 && (qm & facing(-ex, -ey, -ez)) != 0;
}

// smallest 0-based x-major key x * H * D + y * D + z among the block's voxels in m
__device__ __forceinline__ uint32_t block_min_key(uint32_t m, const Geom& g, int64_t bx, int64_t by, int64_t bz,
                                                  int BZ) {
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
                if (m & (1u << (dz * 4 + dy * 2 + dx)))
                    return (uint32_t)((2 * bx + dx) * g.H * g.D + (2 * by + dy) * g.D + (bz * BZ + dz));
    return NONE;
}

// LDS of one local-pass block
template <typename BT>
struct BcclLds {
    uint32_t lab[BT::N];
    uint32_t kmin[BT::N];
    uint8_t msk[BT::N];
    alignas(16) uint8_t vox[BT::N * 4 * BT::BZ];  // the tile's binary voxels (uint8 fast path)
    uint32_t cnt;
};

// Local pass over tile `tile` (index inside the image) of an image whose blocks start
// at global block boff; `gtile` is the tile's slot in the roots / nroots arrays.
// Block indices stored in M, L, K and roots are global (boff + local).
template <typename BT, typename FG>
__device__ __forceinline__ void bccl_local_tile(FG fg, const Geom& g, const BGrid& B, const TileGrid& T,
                                                int64_t tile, int64_t gtile, int64_t boff, uint8_t* M,
                                                uint32_t* L, uint32_t* K, uint32_t* roots, uint32_t* nroots,
                                                BcclLds<BT>& sm) {
    constexpr int PER = BT::PER, TX = BT::TX, TY = BT::TY, BZ = BT::BZ;
    uint32_t* lab = sm.lab;
    uint32_t* kmin = sm.kmin;
    uint8_t* msk = sm.msk;
    uint32_t& cnt = sm.cnt;
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const int64_t bx0 = tx * TX, by0 = ty * TY, bz0 = tz * BT::TZ;
    if (threadIdx.x == 0) cnt = 0;
    int64_t gb[PER];
    uint32_t m[PER];
    bool anyf = false;
    // uint8 binary, tile fully inside the image, rows 8-byte aligned: the tile's voxels
    // come in with one 8-byte load per thread (rows of 2 TX bytes), an empty tile
    // leaves after one block-wide test, and the block masks are read from LDS
    bool staged = false;
    if constexpr (std::is_same<FG, FgU8>::value) {
        constexpr int RB = 2 * TX;                 // bytes per voxel row of the tile
        constexpr int NROW = 2 * TY * BT::TZ * BZ;  // voxel rows of the tile
        static_assert(RB % 8 == 0 && NROW * RB == 8 * CT_THREADS, "tile must be 8 bytes per thread");
        const bool full = (bx0 + TX) * 2 <= g.W && (by0 + TY) * 2 <= g.H && (bz0 + BT::TZ) * BZ <= g.D &&
                          (g.W & 7) == 0 && ((uintptr_t)fg.p & 7) == 0;
        if (full) {
            const int r = threadIdx.x / (RB / 8), part = threadIdx.x % (RB / 8);
            const int ry = r % (2 * TY), rz = r / (2 * TY);
            const int64_t z = bz0 * BZ + rz, y = by0 * 2 + ry;
            const uint2 q = *(const uint2*)(fg.p + (z * g.H + y) * g.W + bx0 * 2 + part * 8);
            if (!__syncthreads_or((q.x | q.y) != 0)) {
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int li = threadIdx.x + k * CT_THREADS;
                    const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
                    M[boff + ((bz0 + lz) * B.H + (by0 + ly)) * B.W + (bx0 + lx)] = 0;
                }
                if (threadIdx.x == 0) nroots[gtile] = 0;
                return;
            }
            *(uint2*)(sm.vox + r * RB + part * 8) = q;
            __syncthreads();
            staged = true;
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        const int64_t bx = bx0 + lx, by = by0 + ly, bz = bz0 + lz;
        const bool in = bx < B.W && by < B.H && bz < B.D;
        uint32_t mk = 0;
        if (staged) {
#pragma unroll
            for (int dz = 0; dz < BZ; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                    const uint8_t* row = sm.vox + ((lz * BZ + dz) * 2 * TY + 2 * ly + dy) * (2 * TX) + 2 * lx;
                    if (row[0]) mk |= 1u << (dz * 4 + dy * 2);
                    if (row[1]) mk |= 1u << (dz * 4 + dy * 2 + 1);
                }
        } else if (in) {
#pragma unroll
            for (int dz = 0; dz < BZ; ++dz) {
                const int64_t z = bz * BZ + dz;
                if (z >= g.D) continue;
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                    const int64_t y = 2 * by + dy;
                    if (y >= g.H) continue;
                    const int64_t row = (z * g.H + y) * g.W;
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int64_t x = 2 * bx + dx;
                        if (x < g.W && fg(row + x)) mk |= 1u << (dz * 4 + dy * 2 + dx);
                    }
                }
            }
        }
        gb[k] = boff + (in ? (bz * B.H + by) * B.W + bx : 0);
        if (in) M[gb[k]] = (uint8_t)mk;  // neighbouring tiles read the masks in the border pass
        m[k] = mk;
        msk[li] = (uint8_t)mk;
        lab[li] = mk ? (uint32_t)li : NONE;
        kmin[li] = NONE;
        anyf |= mk != 0;
    }
    if (!__syncthreads_or(anyf ? 1 : 0)) {
        if (threadIdx.x == 0) nroots[gtile] = 0;
        return;
    }
    const int zr = BZ == 2 ? 1 : 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!m[k]) continue;
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        for (int ez = -zr; ez <= 0; ++ez)
            for (int ey = -1; ey <= 1; ++ey)
                for (int ex = -1; ex <= 1; ++ex) {
                    if (ez == 0 && (ey > 0 || (ey == 0 && ex >= 0))) continue;
                    const int qx = lx + ex, qy = ly + ey, qz = lz + ez;
                    if (qx < 0 || qx >= TX || qy < 0 || qy >= TY || qz < 0) continue;
                    const int lq = (qz * TY + qy) * TX + qx;
                    const uint32_t qm = msk[lq];
                    if (!qm || !blocks_touch(m[k], qm, ex, ey, ez)) continue;
                    lunite(lab, (uint32_t)li, (uint32_t)lq);
                }
    }
    __syncthreads();
    uint32_t r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        r[k] = m[k] ? lfind(lab, (uint32_t)li) : NONE;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!m[k]) continue;
        const int li = threadIdx.x + k * CT_THREADS;
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        atomicMin(kmin + r[k], block_min_key(m[k], g, bx0 + lx, by0 + ly, bz0 + lz, BZ));
        const int rx = (int)(r[k] % TX), ry = (int)((r[k] / TX) % TY), rz = (int)(r[k] / (TX * TY));
        L[gb[k]] = (uint32_t)(boff + ((bz0 + rz) * B.H + (by0 + ry)) * B.W + (bx0 + rx));
    }
    __syncthreads();
    uint32_t* myroots = roots + gtile * BT::ROOTS;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int li = threadIdx.x + k * CT_THREADS;
        if (m[k] && r[k] == (uint32_t)li) {
            K[gb[k]] = kmin[li];
            const uint32_t slot = atomicAdd(&cnt, 1u);
            myroots[slot] = (uint32_t)gb[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) nroots[gtile] = cnt;
}

template <typename BT, typename FG>
__global__ __launch_bounds__(CT_THREADS) void k_bccl_local(FG fg, Geom g, BGrid B, TileGrid T, uint8_t* M,
                                                          uint32_t* L, uint32_t* K, uint32_t* roots,
                                                          uint32_t* nroots) {
    __shared__ BcclLds<BT> sm;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    bccl_local_tile<BT>(fg, g, B, T, tile, tile, 0, M, L, K, roots, nroots, sm);
}

// foreground blocks on the faces of each tile unite with the touching foreground
// blocks (in the 13 earlier directions) that lie in another tile
template <typename BT>
__device__ __forceinline__ void bccl_border_tile(const BGrid& B, const TileGrid& T, int64_t tile, int64_t gtile,
                                                 int64_t boff, const uint8_t* M, uint32_t* L,
                                                 const uint32_t* nroots) {
    constexpr int N = BT::N, TX = BT::TX, TY = BT::TY;
    if (nroots[gtile] == 0) return;  // no foreground in this tile
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const int64_t bx0 = tx * TX, by0 = ty * TY, bz0 = tz * BT::TZ;
    const int zr = BT::BZ == 2 ? 1 : 0;
    for (int li = threadIdx.x; li < N; li += CT_THREADS) {
        const int lx = li % TX, ly = (li / TX) % TY, lz = li / (TX * TY);
        // low z/y/x faces, and the high y/x faces (their ez = -1 / ey = -1 neighbours
        // reach into the next y / x tile)
        if (!((zr && lz == 0) || ly == 0 || lx == 0 || ly == TY - 1 || lx == TX - 1)) continue;
        const int64_t bx = bx0 + lx, by = by0 + ly, bz = bz0 + lz;
        if (bx >= B.W || by >= B.H || bz >= B.D) continue;
        const int64_t b = boff + (bz * B.H + by) * B.W + bx;
        const uint32_t m = M[b];
        if (!m) continue;
        const uint32_t lroot = ld_relaxed(L + b);
        uint32_t last = NONE;
        for (int ez = -zr; ez <= 0; ++ez)
            for (int ey = -1; ey <= 1; ++ey)
                for (int ex = -1; ex <= 1; ++ex) {
                    if (ez == 0 && (ey > 0 || (ey == 0 && ex >= 0))) continue;
                    if (lz + ez >= 0 && ly + ey >= 0 && lx + ex >= 0 && lx + ex < TX && ly + ey < TY) continue;
                    const int64_t qx = bx + ex, qy = by + ey, qz = bz + ez;
                    if (qx < 0 || qx >= B.W || qy < 0 || qy >= B.H || qz < 0) continue;
                    const int64_t q = boff + (qz * B.H + qy) * B.W + qx;
                    const uint32_t qm = M[q];
                    if (!qm || !blocks_touch(m, qm, ex, ey, ez)) continue;
                    const uint32_t lq = ld_relaxed(L + q);
                    if (lq == last) continue;  // same local component as the previous neighbour
                    last = lq;
                    unite(L, lroot, lq);
                }
    }
}

template <typename BT>
__global__ __launch_bounds__(CT_THREADS) void k_bccl_border(BGrid B, TileGrid T, const uint8_t* M, uint32_t* L,
                                                           const uint32_t* nroots) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    bccl_border_tile<BT>(B, T, tile, tile, 0, M, L, nroots);
}

// A foreground voxel's label is F[local root of its block], held as float and
// converted with saturation like the voxel version.  The voxels are written from
// the blocks' masks (bit dz * 4 + dy * 2 + dx), so the binary image is not read
// again and each block's label is looked up once: one thread per two blocks along
// x writes their 4 x 2 x BZ voxels.  VEC: rows of a multiple of 4 voxels and a
// 16-byte aligned destination, four voxels per store.
template <int BZ, bool VEC, typename ST>
__global__ void k_bccl_write(const uint8_t* __restrict__ M, Geom g, BGrid B, const uint32_t* __restrict__ L,
                             const uint32_t* __restrict__ F, ST out) {
    const int64_t brows = B.D * B.H, np = (B.W + 1) >> 1;
    for (int64_t br = blockIdx.x; br < brows; br += gridDim.x) {
        const int64_t bz = br / B.H, by = br - bz * B.H;
        const int64_t z0 = bz * BZ, y0 = by * 2;
        const int nz = (BZ == 2 && z0 + 1 < g.D) ? 2 : 1, ny = y0 + 1 < g.H ? 2 : 1;
        for (int64_t p = threadIdx.x; p < np; p += blockDim.x) {
            const int64_t b = br * B.W + 2 * p, x = 4 * p;
            uint32_t m2;
            float f0, f1;
            uint2 l2;
            // B.W even: the pair's mask and both roots load together (a root of an empty
            // block is not used), then both labels together
            if constexpr (VEC) {
                m2 = *(const uint16_t*)(M + b);
                l2 = *(const uint2*)(L + b);
            }
            // the pair's 4 x 2 x 2 voxels lie in one growth tile (x a multiple of 4, y0
            // and z0 even): an empty one stores nothing (its flag is read alongside the
            // pair's mask and roots)
            if constexpr (ST::kCoords)
                if (out.empty(x, y0, z0)) continue;
            if constexpr (VEC) {
                const uint32_t i0 = (m2 & 0xFFu) ? l2.x : 0u, i1 = (m2 >> 8) ? l2.y : 0u;
                const uint32_t g0 = F[i0], g1 = F[i1];
                f0 = (m2 & 0xFFu) ? (float)g0 : 0.f;
                f1 = (m2 >> 8) ? (float)g1 : 0.f;
            } else {
                m2 = M[b] | (2 * p + 1 < B.W ? (uint32_t)M[b + 1] << 8 : 0u);
                f0 = (m2 & 0xFFu) ? (float)F[L[b]] : 0.f;
                f1 = (m2 >> 8) ? (float)F[L[b + 1]] : 0.f;
            }
            for (int dz = 0; dz < nz; ++dz)
                for (int dy = 0; dy < ny; ++dy) {
                    const int sh = dz * 4 + dy * 2;
                    const float v0 = (m2 >> sh) & 1u ? f0 : 0.f, v1 = (m2 >> (sh + 1)) & 1u ? f0 : 0.f;
                    const float v2 = (m2 >> (sh + 8)) & 1u ? f1 : 0.f, v3 = (m2 >> (sh + 9)) & 1u ? f1 : 0.f;
                    const int64_t i = ((z0 + dz) * g.H + y0 + dy) * g.W + x;
                    if constexpr (ST::kCoords) {
                        const int64_t yy = y0 + dy, zz = z0 + dz;
                        if constexpr (VEC) {
                            out.put4(i, x, yy, zz, v0, v1, v2, v3);
                        } else {
                            out.put(i, x, yy, zz, v0);
                            if (x + 1 < g.W) out.put(i + 1, x + 1, yy, zz, v1);
                            if (x + 2 < g.W) out.put(i + 2, x + 2, yy, zz, v2);
                            if (x + 3 < g.W) out.put(i + 3, x + 3, yy, zz, v3);
                        }
                    } else if constexpr (VEC) {
                        out.store4(i, v0, v1, v2, v3);
                    } else {
                        out(i, v0);
                        if (x + 1 < g.W) out(i + 1, v1);
                        if (x + 2 < g.W) out(i + 2, v2);
                        if (x + 3 < g.W) out(i + 3, v3);
                    }
                }
        }
    }
}

template <int BZ, bool VEC>
static void launch_bccl_write_v(const uint8_t* M, const Geom& g, const BGrid& B, const uint32_t* L, const uint32_t* F,
                                DBuf dst, unsigned grid, hipStream_t s) {
    switch (dst.dt) {
        case CLE_I32: k_bccl_write<BZ, VEC><<<grid, 256, 0, s>>>(M, g, B, L, F, StLabI32{(int32_t*)dst.p}); break;
        case CLE_U32: k_bccl_write<BZ, VEC><<<grid, 256, 0, s>>>(M, g, B, L, F, StLabU32{(uint32_t*)dst.p}); break;
        case CLE_F32: k_bccl_write<BZ, VEC><<<grid, 256, 0, s>>>(M, g, B, L, F, StLabF32{(float*)dst.p}); break;
        default: k_bccl_write<BZ, false><<<grid, 256, 0, s>>>(M, g, B, L, F, StLabAny{dst}); break;
    }
}

template <int BZ>
static void launch_bccl_write(const uint8_t* M, const Geom& g, const BGrid& B, const uint32_t* L, const uint32_t* F,
                              DBuf dst, hipStream_t s) {
    const int64_t brows = B.D * B.H;
    const unsigned grid = (unsigned)(brows < 262144 ? brows : 262144);
    if ((g.W & 3) == 0 && ((uintptr_t)dst.p & 15) == 0)
        launch_bccl_write_v<BZ, true>(M, g, B, L, F, dst, grid, s);
    else
        launch_bccl_write_v<BZ, false>(M, g, B, L, F, dst, grid, s);
}

// the block grid, tile grid and key range of the box CCL of one image, and the scratch
// bytes its buffers take
template <typename BT>
struct BcclSizes {
    BGrid B;
    TileGrid T;
    int64_t nb, nt;
    uint64_t maxbit;
    BcclSizes(const Geom& g, int compat)
        : B{(g.D + BT::BZ - 1) / BT::BZ, (g.H + 1) / 2, (g.W + 1) / 2},
          T{(B.D + BT::TZ - 1) / BT::TZ, (B.H + BT::TY - 1) / BT::TY, (B.W + BT::TX - 1) / BT::TX},
          nb(B.D * B.H * B.W), nt(T.count()),
          maxbit(compat ? (uint64_t)(float)(uint64_t)g.n() : (uint64_t)g.n()) {}
    size_t bytes() const {
        return scratch_round((size_t)nb) + 3 * scratch_round((size_t)nb * 4) +
               scratch_round((size_t)nt * BT::ROOTS * 4) + scratch_round((size_t)nt * 4) + RankTable::bytes(maxbit + 1);
    }
};

// gi: write the growth init (GrowInitOut) instead of labels into dst
template <typename BT, typename FG>
static int bccl_run(FG fg, const Geom& g, DBuf dst, int options, uint32_t* n_labels_dev, hipStream_t s,
                    const GrowInitOut* gi = nullptr) {
    const int compat = (options & CLE_FP32_KEY_COMPAT) ? 1 : 0;
    const BcclSizes<BT> Z(g, compat);
    const BGrid B = Z.B;
    const int64_t nb = Z.nb;
    const TileGrid T = Z.T;
    const int64_t nt = Z.nt;
    Scratch Ms, Ls, Ks, Rs, Ns, Fs;
    CLE_TRY(Ms.alloc((size_t)nb, s));
    CLE_TRY(Ls.alloc((size_t)nb * 4, s));
    CLE_TRY(Ks.alloc((size_t)nb * 4, s));
    CLE_TRY(Fs.alloc((size_t)nb * 4, s));
    CLE_TRY(Rs.alloc((size_t)nt * BT::ROOTS * 4, s));
    CLE_TRY(Ns.alloc((size_t)nt * 4, s));
    uint32_t *l = Ls.get<uint32_t>(), *k = Ks.get<uint32_t>(), *roots = Rs.get<uint32_t>(), *nr = Ns.get<uint32_t>();
    k_bccl_local<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(fg, g, B, T, Ms.get<uint8_t>(), l, k, roots, nr);
    k_bccl_border<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(B, T, Ms.get<uint8_t>(), l, nr);
    k_ccl_resolve<BT><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, nt);
    CLE_LAUNCHED();
    RankTable rt;
    CLE_TRY(rt.init(Z.maxbit + 1, s));
    k_ccl_mark_roots<BT><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, compat, rt.bm.get<uint32_t>(), nt);
    CLE_LAUNCHED();
    CLE_TRY(rt.scan(n_labels_dev, s));
    k_ccl_final<BT><<<root_grid(nt), CT_THREADS, 0, s>>>(l, k, roots, nr, compat, rt.bm.get<uint32_t>(),
                                                          rt.prefix.get<uint32_t>(), Fs.get<uint32_t>(), nt);
    if (gi) {
        const int64_t brows = B.D * B.H;
        const unsigned grid = (unsigned)(brows < 262144 ? brows : 262144);
        if ((g.W & 3) == 0)
            k_bccl_write<BT::BZ, true><<<grid, 256, 0, s>>>(Ms.get<uint8_t>(), g, B, l, Fs.get<uint32_t>(), StGrowInit{*gi});
        else
            k_bccl_write<BT::BZ, false><<<grid, 256, 0, s>>>(Ms.get<uint8_t>(), g, B, l, Fs.get<uint32_t>(), StGrowInit{*gi});
    } else {
        launch_bccl_write<BT::BZ>(Ms.get<uint8_t>(), g, B, l, Fs.get<uint32_t>(), dst, s);
    }
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- batched box CCL
// Every image of a batch runs the single-image block CCL on its own blocks, tiles and
// key bits; the images' 2x2x2 blocks are numbered consecutively (image j from
// boff_j), so the union-find arrays, the tile root lists and one rank bitmap serve
// the whole batch.  2D and 3D images use their own tile shapes (one launch per kind
// and pass); image j's label of key bit k is rank(kbase_j + k) - rank(kbase_j).
__host__ __device__ inline bool is3d(const Geom& g) { return g.ndim == 3 && g.D > 1; }

template <typename BT>
__host__ __device__ inline BGrid bgrid_of(const Geom& g) {
    return BGrid{(g.D + BT::BZ - 1) / BT::BZ, (g.H + 1) / 2, (g.W + 1) / 2};
}
template <typename BT>
__host__ __device__ inline TileGrid tgrid_of(const BGrid& B) {
    return TileGrid{(B.D + BT::TZ - 1) / BT::TZ, (B.H + BT::TY - 1) / BT::TY, (B.W + BT::TX - 1) / BT::TX};
}

void ccl_batch_layout(std::vector<BImg>& imgs, CclBatchSizes* sz) {
    int64_t nb = 0, bits = 0;
    for (auto& im : imgs) {
        const bool d3 = is3d(im.g);
        const BGrid B = d3 ? bgrid_of<BlkTile3D>(im.g) : bgrid_of<BlkTile2D>(im.g);
        im.boff = nb;
        nb += B.D * B.H * B.W;
        const uint64_t n = (uint64_t)im.g.n();
        const uint64_t maxbit = (uint64_t)(float)n;  // float32 keys (the voronoi pipeline's CCL)
        im.kbase = bits;
        bits += (int64_t)((maxbit + 1 + 31) / 32 * 32);
    }
    sz->nb = nb;
    sz->bits = bits;
}

template <typename BT>
__global__ __launch_bounds__(CT_THREADS) void k_bccl_local_batch(const BImg* __restrict__ imgs,
                                                                const int64_t* __restrict__ tstart, int nimg,
                                                                const uint8_t* S, uint8_t* M, uint32_t* L,
                                                                uint32_t* K, uint32_t* roots, uint32_t* nroots) {
    __shared__ BcclLds<BT> sm;
    const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
    const int j = find_seg(tstart, nimg, b);
    const BImg& im = imgs[j];
    const BGrid B = bgrid_of<BT>(im.g);
    bccl_local_tile<BT>(FgU8{S + im.off}, im.g, B, tgrid_of<BT>(B), b - tstart[j], b, im.boff, M, L, K, roots,
                        nroots, sm);
}

template <typename BT>
__global__ __launch_bounds__(CT_THREADS) void k_bccl_border_batch(const BImg* __restrict__ imgs,
                                                                 const int64_t* __restrict__ tstart, int nimg,
                                                                 const uint8_t* M, uint32_t* L,
                                                                 const uint32_t* nroots) {
    const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
    const int j = find_seg(tstart, nimg, b);
    const BImg& im = imgs[j];
    const BGrid B = bgrid_of<BT>(im.g);
    bccl_border_tile<BT>(B, tgrid_of<BT>(B), b - tstart[j], b, im.boff, M, L, nroots);
}

template <typename BT>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_mark_roots_batch(const BImg* __restrict__ imgs,
                                                                    const int64_t* __restrict__ tstart, int nimg,
                                                                    const uint32_t* L, const uint32_t* K,
                                                                    const uint32_t* roots, const uint32_t* nroots,
                                                                    uint32_t* bm) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t kbase = imgs[find_seg(tstart, nimg, tile)].kbase;
    const uint32_t n = nroots[tile];
    for (uint32_t k = threadIdx.x; k < n; k += CT_THREADS) {
        const uint32_t lr = roots[tile * BT::ROOTS + k];
        if (L[lr] != lr) continue;  // not a global root
        const uint64_t bit = (uint64_t)kbase + key_bit(K[lr], 1);
        atomicOr(bm + (bit >> 5), 1u << (bit & 31));
    }
}

template <typename BT>
__global__ __launch_bounds__(CT_THREADS) void k_ccl_final_batch(const BImg* __restrict__ imgs,
                                                               const int64_t* __restrict__ tstart, int nimg,
                                                               const uint32_t* L, const uint32_t* K,
                                                               const uint32_t* roots, const uint32_t* nroots,
                                                               const uint32_t* bm, const uint32_t* prefix,
                                                               uint32_t* F) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t kbase = imgs[find_seg(tstart, nimg, tile)].kbase;
    const uint32_t before = prefix[kbase >> 5];  // set bits of the images before this one
    const uint32_t n = nroots[tile];
    for (uint32_t k = threadIdx.x; k < n; k += CT_THREADS) {
        const uint32_t lr = roots[tile * BT::ROOTS + k];
        F[lr] = bit_rank(bm, prefix, (uint64_t)kbase + key_bit(K[L[lr]], 1)) - before;
    }
}

// float32 labels of the work image from the block masks, as k_bccl_write: one
// block per row of 2x2x2 blocks of one image.  GROW: the growth's initial state
// (G - 1) + label into out and gi.A instead, with the seed tiles stamped (see
// GrowInitBatch).
// Rows of 2x2x2 blocks a 256-thread block of the batched write takes: several short
// rows per block (small images), one long row otherwise.
__host__ __device__ inline int bw_rows_per_block(int64_t np) {
    const int64_t r = 256 / (np > 0 ? np : 1);
    return (int)(r < 1 ? 1 : (r > 64 ? 64 : r));
}

template <int BZ, bool GROW>
__global__ __launch_bounds__(256) void k_bccl_write_batch(const BImg* __restrict__ imgs,
                                                         const int64_t* __restrict__ brstart, int nimg,
                                                         const uint8_t* __restrict__ M, const uint32_t* __restrict__ L,
                                                         const uint32_t* __restrict__ F, float* out, GrowInitBatch gi) {
    const int64_t r = blockIdx.x;
    const int j = find_seg(brstart, nimg, r);
    const BImg& im = imgs[j];
    const Geom g = im.g;
    const BGrid B{(g.D + BZ - 1) / BZ, (g.H + 1) / 2, (g.W + 1) / 2};
    const int64_t np = (B.W + 1) >> 1;  // 4-voxel groups per row
    const int rpb = bw_rows_per_block(np);
    // this thread's row of blocks and first group
    int64_t br, p0, pstep;
    if (rpb > 1) {
        if (threadIdx.x >= rpb * np) return;
        br = (r - brstart[j]) * rpb + threadIdx.x / np;
        p0 = threadIdx.x % np;
        pstep = np;
    } else {
        br = r - brstart[j];
        p0 = threadIdx.x;
        pstep = blockDim.x;
    }
    if (br >= B.D * B.H) return;
    const int64_t bz = br / B.H, by = br - bz * B.H;
    const int64_t z0 = bz * BZ, y0 = by * 2;
    const int nz = (BZ == 2 && z0 + 1 < g.D) ? 2 : 1, ny = y0 + 1 < g.H ? 2 : 1;
    // growth tiles: 16 x 8 x 8 (3D), 32 x 32 (2D); a thread's four voxels lie in one
    const bool d3 = is3d(g);
    const int gtx = d3 ? 16 : 32, gty = d3 ? 8 : 32, gtz = d3 ? 8 : 1;
    const int64_t gnx = (g.W + gtx - 1) / gtx, gny = (g.H + gty - 1) / gty;
    for (int64_t p = p0; p < np; p += pstep) {
        const int64_t b = im.boff + br * B.W + 2 * p, x = 4 * p;
        const uint32_t m2 = M[b] | (2 * p + 1 < B.W ? (uint32_t)M[b + 1] << 8 : 0u);
        const float f0 = (m2 & 0xFFu) ? (float)F[L[b]] : 0.f;
        const float f1 = (m2 >> 8) ? (float)F[L[b + 1]] : 0.f;
        for (int dz = 0; dz < nz; ++dz)
            for (int dy = 0; dy < ny; ++dy) {
                const int sh = dz * 4 + dy * 2;
                const int64_t i = im.off + ((z0 + dz) * g.H + y0 + dy) * g.W + x;
                float v[4] = {(m2 >> sh) & 1u ? f0 : 0.f, (m2 >> (sh + 1)) & 1u ? f0 : 0.f,
                              (m2 >> (sh + 8)) & 1u ? f1 : 0.f, (m2 >> (sh + 9)) & 1u ? f1 : 0.f};
                const int nv = x + 4 <= g.W ? 4 : (int)(g.W - x);
                // rows of a multiple of 4 voxels: 4-voxel groups are 16-byte aligned (image
                // offsets are multiples of 64 voxels)
                const bool vec = (g.W & 3) == 0;
                if (GROW) {
                    bool pos = false, msk = false;
                    if (vec) {
                        const uchar4 m = *(const uchar4*)(gi.G + i);
                        // the reference's float32 add order: (mask - 1) + label
                        v[0] = ((float)m.x - 1.f) + v[0];
                        v[1] = ((float)m.y - 1.f) + v[1];
                        v[2] = ((float)m.z - 1.f) + v[2];
                        v[3] = ((float)m.w - 1.f) + v[3];
                        *(float4*)(gi.A + i) = make_float4(v[0], v[1], v[2], v[3]);
                        pos = v[0] > 0.f || v[1] > 0.f || v[2] > 0.f || v[3] > 0.f;
                        msk = (m.x | m.y | m.z | m.w) != 0;
                    } else {
                        for (int c = 0; c < nv; ++c) {
                            const uint8_t m = gi.G[i + c];
                            v[c] = ((float)m - 1.f) + v[c];
                            gi.A[i + c] = v[c];
                            pos |= v[c] > 0.f;
                            msk |= m != 0;
                        }
                    }
                    const int64_t t = im.tbase + (((z0 + dz) / gtz) * gny + (y0 + dy) / gty) * gnx + x / gtx;
                    if (pos) gi.seed[t] = -1;
                    if (msk) gi.settled[t] = 0;  // a tile with a mask voxel is not settled
                }
                if (vec) {
                    *(float4*)(out + i) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
                    for (int c = 0; c < nv; ++c) out[i + c] = v[c];
                }
            }
    }
}

template <typename BT>
static int plan_ccl_kind(const std::vector<BImg>& imgs, size_t imgs_off, bool want3d, const uint8_t* S, float* F,
                         uint8_t* M, uint32_t* L, uint32_t* K, uint32_t* Fl, RankTable* rt, Blob& blob,
                         std::vector<Launch>& pre, std::vector<Launch>& post, std::vector<std::shared_ptr<void>>& keep,
                         hipStream_t s, const GrowInitBatch* grow) {
    const GrowInitBatch gi = grow ? *grow : GrowInitBatch{nullptr, nullptr, nullptr, nullptr};
    const bool gw = grow != nullptr;
    const int nimg = (int)imgs.size();
    std::vector<int64_t> tst(nimg + 1), rst(nimg + 1);
    int64_t nt = 0, nr = 0;
    for (int j = 0; j < nimg; ++j) {
        tst[j] = nt;
        rst[j] = nr;
        if (is3d(imgs[j].g) != want3d || imgs[j].g.n() == 0) continue;
        const BGrid B = bgrid_of<BT>(imgs[j].g);
        nt += tgrid_of<BT>(B).count();
        const int rpb = bw_rows_per_block((B.W + 1) >> 1);
        nr += (B.D * B.H + rpb - 1) / rpb;  // write blocks: rpb rows of 2x2x2 blocks each
    }
    tst[nimg] = nt;
    rst[nimg] = nr;
    if (nt == 0) return CLE_OK;
    auto Rs = std::make_shared<Scratch>();
    auto Ns = std::make_shared<Scratch>();
    keep.push_back(Rs);
    keep.push_back(Ns);
    CLE_TRY(Rs->alloc((size_t)nt * BT::ROOTS * 4, s));
    CLE_TRY(Ns->alloc((size_t)nt * 4, s));
    uint32_t* roots = Rs->get<uint32_t>();
    uint32_t* nroots = Ns->get<uint32_t>();
    const size_t to = blob.put_vec(tst), ro = blob.put_vec(rst);
    // before the rank scan: local, border, resolve, mark
    pre.push_back([=](const char* dev) -> int {
        const BImg* im = (const BImg*)(dev + imgs_off);
        const int64_t* ts = (const int64_t*)(dev + to);
        k_bccl_local_batch<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(im, ts, nimg, S, M, L, K, roots, nroots);
        k_bccl_border_batch<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(im, ts, nimg, M, L, nroots);
        k_ccl_resolve<BT><<<root_grid(nt), CT_THREADS, 0, s>>>(L, K, roots, nroots, nt);
        k_ccl_mark_roots_batch<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(im, ts, nimg, L, K, roots, nroots,
                                                                      rt->bm.get<uint32_t>());
        CLE_LAUNCHED();
        return CLE_OK;
    });
    // after it: final labels of the local roots, then the float32 label rows
    post.push_back([=](const char* dev) -> int {
        const BImg* im = (const BImg*)(dev + imgs_off);
        const int64_t* ts = (const int64_t*)(dev + to);
        const int64_t* rs = (const int64_t*)(dev + ro);
        k_ccl_final_batch<BT><<<(unsigned)nt, CT_THREADS, 0, s>>>(im, ts, nimg, L, K, roots, nroots,
                                                                 rt->bm.get<uint32_t>(), rt->prefix.get<uint32_t>(),
                                                                 Fl);
        if (gw)
            k_bccl_write_batch<BT::BZ, true><<<(unsigned)nr, 256, 0, s>>>(im, rs, nimg, M, L, Fl, F, gi);
        else
            k_bccl_write_batch<BT::BZ, false><<<(unsigned)nr, 256, 0, s>>>(im, rs, nimg, M, L, Fl, F, gi);
        CLE_LAUNCHED();
        return CLE_OK;
    });
    return CLE_OK;
}

int plan_ccl_batch(const std::vector<BImg>& imgs, size_t imgs_off, const CclBatchSizes& sz, const uint8_t* S,
                   float* F, Blob& blob, std::vector<Launch>& out, std::vector<std::shared_ptr<void>>& keep, hipStream_t s,
                   const GrowInitBatch* grow) {
    if (sz.nb == 0) return CLE_OK;
    auto Ms = std::make_shared<Scratch>(), Ls = std::make_shared<Scratch>(), Ks = std::make_shared<Scratch>(),
         Fs = std::make_shared<Scratch>();
    keep.push_back(Ms);
    keep.push_back(Ls);
    keep.push_back(Ks);
    keep.push_back(Fs);
    CLE_TRY(Ms->alloc((size_t)sz.nb, s));
    CLE_TRY(Ls->alloc((size_t)sz.nb * 4, s));
    CLE_TRY(Ks->alloc((size_t)sz.nb * 4, s));
    CLE_TRY(Fs->alloc((size_t)sz.nb * 4, s));
    auto rts = std::make_shared<RankTable>();
    keep.push_back(rts);
    RankTable* rt = rts.get();
    std::vector<Launch> pre, post;
    CLE_TRY(rt->init((uint64_t)sz.bits, s));
    CLE_TRY(plan_ccl_kind<BlkTile3D>(imgs, imgs_off, true, S, F, Ms->get<uint8_t>(), Ls->get<uint32_t>(),
                                     Ks->get<uint32_t>(), Fs->get<uint32_t>(), rt, blob, pre, post, keep, s, grow));
    CLE_TRY(plan_ccl_kind<BlkTile2D>(imgs, imgs_off, false, S, F, Ms->get<uint8_t>(), Ls->get<uint32_t>(),
                                     Ks->get<uint32_t>(), Fs->get<uint32_t>(), rt, blob, pre, post, keep, s, grow));
    for (auto& l : pre) out.push_back(l);
    out.push_back([=](const char*) -> int { return rt->scan(nullptr, s); });
    for (auto& l : post) out.push_back(l);
    return CLE_OK;
}

int ccl_impl(DBuf binary, const Geom& g, DBuf dst, int connectivity, int options, uint32_t* n_labels_dev,
             hipStream_t s) {
    if (g.n() == 0) return n_labels_dev ? fill_u32(n_labels_dev, 0u, 1, s) : CLE_OK;
    const bool three_d = g.ndim == 3 && g.D > 1;
    if (binary.dt == CLE_U8 || binary.dt == CLE_I8) {
        FgU8 fg{(const uint8_t*)binary.p};
        if (connectivity == CLE_CONN_BOX)
            return three_d ? bccl_run<BlkTile3D>(fg, g, dst, options, n_labels_dev, s)
                           : bccl_run<BlkTile2D>(fg, g, dst, options, n_labels_dev, s);
        return three_d ? ccl_run<Tile3D>(fg, g, dst, connectivity, options, n_labels_dev, s)
                       : ccl_run<Tile2D>(fg, g, dst, connectivity, options, n_labels_dev, s);
    }
    FgAny fg{binary};
    if (connectivity == CLE_CONN_BOX)
        return three_d ? bccl_run<BlkTile3D>(fg, g, dst, options, n_labels_dev, s)
                       : bccl_run<BlkTile2D>(fg, g, dst, options, n_labels_dev, s);
    return three_d ? ccl_run<Tile3D>(fg, g, dst, connectivity, options, n_labels_dev, s)
                   : ccl_run<Tile2D>(fg, g, dst, connectivity, options, n_labels_dev, s);
}

int ccl_box_growth_init(DBuf binary, const Geom& g, const GrowInitOut& out, hipStream_t s) {
    if (g.n() == 0) return CLE_OK;
    if (((uintptr_t)out.A & 15) || ((uintptr_t)out.B & 15) || ((uintptr_t)out.mask & 3))
        return fail(CLE_E_INTERNAL, "growth init: misaligned buffers");
    const bool three_d = g.ndim == 3 && g.D > 1;
    const DBuf none{nullptr, CLE_F32};
    if (binary.dt == CLE_U8 || binary.dt == CLE_I8) {
        FgU8 fg{(const uint8_t*)binary.p};
        return three_d ? bccl_run<BlkTile3D>(fg, g, none, CLE_FP32_KEY_COMPAT, nullptr, s, &out)
                       : bccl_run<BlkTile2D>(fg, g, none, CLE_FP32_KEY_COMPAT, nullptr, s, &out);
    }
    FgAny fg{binary};
    return three_d ? bccl_run<BlkTile3D>(fg, g, none, CLE_FP32_KEY_COMPAT, nullptr, s, &out)
                   : bccl_run<BlkTile2D>(fg, g, none, CLE_FP32_KEY_COMPAT, nullptr, s, &out);
}

size_t ccl_box_scratch_bytes(const Geom& g, int options) {
    const int compat = (options & CLE_FP32_KEY_COMPAT) ? 1 : 0;
    const bool three_d = g.ndim == 3 && g.D > 1;
    return three_d ? BcclSizes<BlkTile3D>(g, compat).bytes() : BcclSizes<BlkTile2D>(g, compat).bytes();
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
