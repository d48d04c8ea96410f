// Masked Voronoi growth and the composite pipelines built on it.
//
// Reference semantics (pyclesperanto_prototype/...):
//   masked_voronoi_labeling  _tier5/_masked_voronoi_labeling.py:44-74:
//       flup = mask - 1; flop = CCL_box(binary); flip = flup + flop (float32); then
//       iteration it (even: box, flip -> flop; odd: diamond, flop -> flip) copies
//       non-zero voxels and gives a zero voxel the largest positive neighbour value
//       (_tier1/_onlyzero_overwrite_maximum_{box,diamond}.py), until an iteration
//       changes nothing; finally mask(last, mask).
//   voronoi_otsu_labeling    _tier9/_voronoi_otsu_labeling.py:57-73
//   gauss_otsu_labeling      _tier9/_gauss_otsu_labeling.py:6-49
//
// MI355X design.  The growth runs as a sequence of synchronous (Jacobi) steps on two
// float32 buffers, but a step only touches the tiles that can change: a tile T is
// processed at step t iff some tile of its 3^d tile neighbourhood changed at step
// t-1 or t-2 (at t = 0, 1 the "changes" at -1, -2 are the tiles holding a positive
// initial value).  If nothing near T changed in the last two steps, T's output at t
// equals its output at t-2 (same neighbourhood kind, same input), which is already
// in the destination buffer, so skipping T leaves both buffers exactly as the
// reference's full sweeps would.  The global "changed" flag of each step lives on
// the device; the host checks it once per batch of steps.
#include "common.hpp"

namespace cle {

int maxima_impl(DBuf src, DBuf dst, const Geom& g, hipStream_t s);

constexpr int VB = 256;  // threads per tile block
constexpr int VPT = 4;   // voxels per thread (tile = 1024 voxels)

// Tiles of 1024 voxels, close to cubes so that a growth front keeps as few voxels
// active as possible: 16 x 8 x 8 in 3D, 32 x 32 in 2D (x rows of 64 / 128 bytes)
struct Tiles {
    int tx, ty, tz;         // tile extents
    int64_t nz, ny, nx;     // tile counts
    __host__ __device__ int64_t count() const { return nz * ny * nx; }
};

__host__ __device__ inline Tiles make_tiles(const Geom& g) {
    Tiles t;
    if (g.ndim == 3 && g.D > 1) { t.tx = 16; t.ty = 8; t.tz = 8; }
    else { t.tx = 32; t.ty = 32; t.tz = 1; }
    t.nz = (g.D + t.tz - 1) / t.tz;
    t.ny = (g.H + t.ty - 1) / t.ty;
    t.nx = (g.W + t.tx - 1) / t.tx;
    return t;
}

__device__ __forceinline__ int64_t clampi(int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// A = B = (mask - 1) + label, and the tiles holding a positive value are flagged as
// "changed" at steps -1 and -2 (slots 3 and 2 of the 4-slot ring).
// tile = index inside the image, gtile = its slot in the flag arrays of ntiles slots
// A tile is "settled" once both buffers hold the same values there and none is zero:
// it can never change again, so the steps skip it (settled[gtile] = 1).
// `val(i)` is the initial value of voxel i (it may read B[i] before B[i] is written).
template <typename VF>
__device__ __forceinline__ bool vor_init_tile(VF val, float* A, float* B, const Geom& g, const Tiles& T,
                                              int64_t tile, int64_t gtile, int64_t ntiles, uint8_t* tflags,
                                              uint8_t* settled) {
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    bool pos = false, zero = false;
    for (int k = 0; k < VPT; ++k) {
        const int li = threadIdx.x + k * VB;
        const int lx = li % T.tx, ly = (li / T.tx) % T.ty, lz = li / (T.tx * T.ty);
        if (lz >= T.tz) continue;
        const int64_t x = tx * T.tx + lx, y = ty * T.ty + ly, z = tz * T.tz + lz;
        if (x >= g.W || y >= g.H || z >= g.D) continue;
        const int64_t i = (z * g.H + y) * g.W + x;
        const float v = val(i);
        A[i] = v;
        B[i] = v;
        pos |= v > 0.f;
        zero |= v == 0.f;
    }
    const int any = __syncthreads_or(pos ? 1 : 0);
    const int anyzero = __syncthreads_or(zero ? 1 : 0);
    if (threadIdx.x == 0) settled[gtile] = anyzero ? 0 : 1;
    if (tflags && threadIdx.x == 0) {
        const uint8_t f = any ? 1 : 0;
        tflags[0 * ntiles + gtile] = 0;
        tflags[1 * ntiles + gtile] = 0;
        tflags[2 * ntiles + gtile] = f;
        tflags[3 * ntiles + gtile] = f;
    }
    return any != 0;
}

// masked growth: (mask - 1) + CCL label, the label held in B (float32 add, as the
// reference's add_image_and_scalar then add_images)
struct MaskedAny {
    DBuf mask;
    const float* B;
    __device__ __forceinline__ float operator()(int64_t i) const { return (ld(mask, i) - 1.f) + B[i]; }
};
struct MaskedU8 {
    const uint8_t* mask;
    const float* B;
    __device__ __forceinline__ float operator()(int64_t i) const { return ((float)mask[i] - 1.f) + B[i]; }
};
// unmasked growth of an existing label image: the labels converted to float32 as the
// reference's copy into a float32 buffer does
struct LabelsAny {
    DBuf src;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(src, i); }
};

__global__ __launch_bounds__(VB) void k_vor_init(DBuf mask, float* A, float* B, Geom g, Tiles T, uint8_t* tflags) {
    vor_init_tile(MaskedAny{mask, B}, A, B, g, T, blockIdx.x, blockIdx.x, T.count(), tflags, tflags + 4 * T.count());
}

__global__ __launch_bounds__(VB) void k_vor_init_labels(DBuf src, float* A, float* B, Geom g, Tiles T,
                                                        uint8_t* tflags) {
    vor_init_tile(LabelsAny{src}, A, B, g, T, blockIdx.x, blockIdx.x, T.count(), tflags, tflags + 4 * T.count());
}

// one growth step t: src -> dst, box (t even) or diamond (t odd)
// One growth step on one tile.  gflag / gany: "changed" flags of this step for the
// image and for the whole launch (the same word for a single image).
// step t on one tile (src -> dst); returns, in every thread, whether the tile changed,
// and marks the tile settled when it did not change and holds no zero
__device__ __forceinline__ bool vor_tile_update(const float* __restrict__ src, float* __restrict__ dst, const Geom& g,
                                                const Tiles& T, int64_t tile, int t, uint8_t* settled_slot) {
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const bool diamond = (t & 1) != 0;
    const int zr = g.ndim == 3 ? 1 : 0;
    bool changed = false, zero = false;
    for (int k = 0; k < VPT; ++k) {
        const int li = threadIdx.x + k * VB;
        const int lx = li % T.tx, ly = (li / T.tx) % T.ty, lz = li / (T.tx * T.ty);
        if (lz >= T.tz) continue;
        const int64_t x = tx * T.tx + lx, y = ty * T.ty + ly, z = tz * T.tz + lz;
        if (x >= g.W || y >= g.H || z >= g.D) continue;
        const int64_t i = (z * g.H + y) * g.W + x;
        const float v = src[i];
        float out = v;
        if (v == 0.f) {
            float best = 0.f;
            if (diamond) {
                const int64_t xm = clampi(x - 1, g.W - 1), xp = clampi(x + 1, g.W - 1);
                const int64_t ym = clampi(y - 1, g.H - 1), yp = clampi(y + 1, g.H - 1);
                const int64_t row = (z * g.H + y) * g.W;
                best = fmaxf(best, src[row + xm]);
                best = fmaxf(best, src[row + xp]);
                best = fmaxf(best, src[(z * g.H + ym) * g.W + x]);
                best = fmaxf(best, src[(z * g.H + yp) * g.W + x]);
                if (zr) {
                    const int64_t zm = clampi(z - 1, g.D - 1), zp = clampi(z + 1, g.D - 1);
                    best = fmaxf(best, src[(zm * g.H + y) * g.W + x]);
                    best = fmaxf(best, src[(zp * g.H + y) * g.W + x]);
                }
            } else {
                for (int dz = -zr; dz <= zr; ++dz) {
                    const int64_t qz = clampi(z + dz, g.D - 1);
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int64_t row = (qz * g.H + clampi(y + dy, g.H - 1)) * g.W;
                        best = fmaxf(best, src[row + clampi(x - 1, g.W - 1)]);
                        best = fmaxf(best, src[row + x]);
                        best = fmaxf(best, src[row + clampi(x + 1, g.W - 1)]);
                    }
                }
            }
            out = best;
        }
        dst[i] = out;
        changed |= (out != v);
        zero |= out == 0.f;
    }
    const bool ch = __syncthreads_or(changed ? 1 : 0) != 0;
    const bool zr0 = __syncthreads_or(zero ? 1 : 0) != 0;
    if (!ch && !zr0 && threadIdx.x == 0) *settled_slot = 1;
    return ch;
}

__device__ __forceinline__ void vor_step_tile(const float* __restrict__ src, float* __restrict__ dst, const Geom& g,
                                              const Tiles& T, int64_t tile, int64_t tbase, int64_t ntiles, int t,
                                              uint8_t* tflags, uint32_t* gflag, uint32_t* gany) {
    const int64_t gtile = tbase + tile;
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    if (threadIdx.x == 0) tflags[((t + 1) & 3) * ntiles + gtile] = 0;
    uint8_t* settled = tflags + 4 * ntiles;
    if (settled[gtile]) return;  // no zero left and both buffers equal: nothing can change
    // activity: any tile of the 3^d tile neighbourhood changed at t-1 or t-2
    const uint8_t* f1 = tflags + ((t + 3) & 3) * ntiles;
    const uint8_t* f2 = tflags + ((t + 2) & 3) * ntiles;
    bool act = false;
    if (threadIdx.x < 27) {
        const int dx = threadIdx.x % 3 - 1, dy = (threadIdx.x / 3) % 3 - 1, dz = threadIdx.x / 9 - 1;
        const int64_t qx = tx + dx, qy = ty + dy, qz = tz + dz;
        if (qx >= 0 && qx < T.nx && qy >= 0 && qy < T.ny && qz >= 0 && qz < T.nz) {
            const int64_t q = tbase + (qz * T.ny + qy) * T.nx + qx;
            act = f1[q] | f2[q];
        }
    }
    if (!__syncthreads_or(act ? 1 : 0)) return;
    if (vor_tile_update(src, dst, g, T, tile, t, settled + gtile) && threadIdx.x == 0) {
        tflags[(t & 3) * ntiles + gtile] = 1;
        *gflag = 1;
        *gany = 1;
    }
}

__global__ __launch_bounds__(VB) void k_vor_step(const float* __restrict__ src, float* __restrict__ dst, Geom g,
                                                Tiles T, int t, uint8_t* tflags, uint32_t* gflags, int ring) {
    if (t > 0 && gflags[(t - 1) % ring] == 0) return;  // the previous step changed nothing: done
    uint32_t* f = gflags + t % ring;
    vor_step_tile(src, dst, g, T, blockIdx.x, 0, T.count(), t, tflags, f, f);
}

constexpr int VOR_BATCH = 8;
constexpr int VOR_RING = 2 * VOR_BATCH;

// Grows the labels held (identically) in A and B until convergence.  On return
// both buffers hold the converged image.
static int grow(float* A, float* B, const Geom& g, const Tiles& T, uint8_t* tflags, int* n_iterations,
                hipStream_t s) {
    Scratch gf;
    CLE_TRY(gf.alloc(VOR_RING * sizeof(uint32_t), s));
    uint32_t* gflags = gf.get<uint32_t>();
    const int64_t ntiles = T.count();
    uint32_t hflags[VOR_BATCH];
    int t = 0;
    int iterations = -1;
    while (iterations < 0) {
        CLE_TRY(fill_u32(gflags + (t % VOR_RING), 0u, VOR_BATCH, s));
        for (int j = 0; j < VOR_BATCH; ++j) {
            const int it = t + j;
            const float* src = (it & 1) ? B : A;
            float* dst = (it & 1) ? A : B;
            k_vor_step<<<(unsigned)ntiles, VB, 0, s>>>(src, dst, g, T, it, tflags, gflags, VOR_RING);
        }
        CLE_LAUNCHED();
        CLE_HIP(hipMemcpyAsync(hflags, gflags + (t % VOR_RING), sizeof(hflags), hipMemcpyDeviceToHost, s));
        CLE_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < VOR_BATCH; ++j)
            if (hflags[j] == 0) { iterations = t + j + 1; break; }
        t += VOR_BATCH;
    }
    if (n_iterations) *n_iterations = iterations;
    return CLE_OK;
}

__global__ void k_store_out(const float* A, DBuf dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst, i, A[i]);
}

__global__ void k_mask_out(const float* A, DBuf mask, DBuf dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst, i, ld(mask, i) != 0.f ? A[i] : 0.f);
}

// seg = blurred > float32(thr); spots <- spots AND seg (in place)
__global__ void k_seg_and(const float* blurred, const double* thr, uint8_t* spots, uint8_t* seg, int64_t n) {
    const float t = (float)thr[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t sg = blurred[i] > t ? 1 : 0;
        seg[i] = sg;
        spots[i] = (spots[i] != 0 && sg) ? 1 : 0;
    }
}

__global__ void k_threshold_u8(const float* blurred, const double* thr, uint8_t* seg, int64_t n) {
    const float t = (float)thr[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        seg[i] = blurred[i] > t ? 1 : 0;
}

static unsigned vg(int64_t n) {
    unsigned g = grid_for((size_t)n, 256);
    return g > 131072 ? 131072 : g;
}

// labels of `binary` grown inside `mask`, written (masked) to dst.  `B` holds
// scratch for the CCL labels (float32, n voxels).
static int masked_voronoi_core(DBuf binary, DBuf mask, DBuf dst, const Geom& g, float* A, float* B,
                               int* n_iterations, hipStream_t s) {
    const int64_t n = g.n();
    CLE_TRY(ccl_impl(binary, g, DBuf{B, CLE_F32}, CLE_CONN_BOX, CLE_FP32_KEY_COMPAT, nullptr, s));
    const Tiles T = make_tiles(g);
    Scratch tf;
    CLE_TRY(tf.alloc((size_t)T.count() * 5, s));  // 4 "changed" slots + settled
    k_vor_init<<<(unsigned)T.count(), VB, 0, s>>>(mask, A, B, g, T, tf.get<uint8_t>());
    CLE_LAUNCHED();
    CLE_TRY(grow(A, B, g, T, tf.get<uint8_t>(), n_iterations, s));
    k_mask_out<<<vg(n), 256, 0, s>>>(A, mask, dst, n);
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- batched pipeline
// voronoi_otsu_labeling of many images of any sizes: every stage is one grid over the
// blocks or tiles of all images (the blur one grid per half-width group), the Otsu
// thresholds stay on the device, and the masked-Voronoi growth steps all images
// together, each image stopping on its own convergence flag; the host checks one
// "any image changed" word per batch of steps.  Each image's result equals
// cle_voronoi_otsu_labeling of that image alone (same kernels, same arithmetic).
struct F32In {
    const float* p;
    __device__ __forceinline__ float operator()(int64_t i) const { return p[i]; }
};

// one block per row of one image (rs = row-prefix table)
__global__ __launch_bounds__(256) void k_maxima_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ rs,
                                                     int nimg, const float* F, uint8_t* S) {
    const int64_t r = blockIdx.x;
    const int j = find_seg(rs, nimg, r);
    const BImg& im = imgs[j];
    const int64_t row = r - rs[j];
    const int64_t z = row / im.g.H, y = row - z * im.g.H;
    uint8_t* o = S + im.off + row * im.g.W;
    maxima_row(F32In{F + im.off}, im.g, z, y, threadIdx.x, 256, [&](int64_t x, bool m) { o[x] = m ? 1 : 0; });
}

// G = F > float32(thr[j]); S = S AND G
__global__ __launch_bounds__(256) void k_seg_and_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ vs,
                                                      int nimg, const float* F, const double* thr, uint8_t* S,
                                                      uint8_t* G) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(vs, nimg, b);
    const BImg& im = imgs[j];
    const int64_t n = im.g.n(), i0 = (b - vs[j]) * BV, i1 = i0 + BV < n ? i0 + BV : n;
    const float t = (float)thr[j];
    for (int64_t i = im.off + i0 + threadIdx.x; i < im.off + i1; i += 256) {
        const uint8_t sg = F[i] > t ? 1 : 0;
        G[i] = sg;
        S[i] = (S[i] != 0 && sg) ? 1 : 0;
    }
}

// Work lists of the batched growth.  List t % 3 holds the tiles active at step t: the
// 3^d tile neighbourhood of a tile changed at step t-1 or t-2 (at t = 0, 1: holds a
// positive initial value).  A tile that changes at step t queues its neighbours for
// t + 1 and t + 2; stamp[t & 1][tile] = last step the tile was queued for keeps every
// list free of duplicates.  A step launch walks its list with a fixed grid, so its
// cost follows the growth front instead of the volume.
struct VorLists {
    uint32_t* items;   // 3 lists of cap entries (global tile indices)
    uint32_t* count;   // 3 counters
    int32_t* stamp;    // 2 x cap
    uint8_t* settled;  // cap
    int64_t cap;       // tiles of all images
};

__device__ __forceinline__ uint64_t lanes_below_v() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// wavefront 0 of the block queues the neighbours of `tile` (inside its image) for
// steps s1 and s2
__device__ __forceinline__ void vor_queue(const Tiles& T, int64_t tile, int64_t tbase, int s1, int s2,
                                          const VorLists& L) {
    const int lane = threadIdx.x;
    int64_t q = -1;
    if (lane < 27) {
        const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
        const int64_t qx = tx + lane % 3 - 1, qy = ty + (lane / 3) % 3 - 1, qz = tz + lane / 9 - 1;
        if (qx >= 0 && qx < T.nx && qy >= 0 && qy < T.ny && qz >= 0 && qz < T.nz)
            q = tbase + (qz * T.ny + qy) * T.nx + qx;
    }
    for (int w = 0; w < 2; ++w) {
        const int st = w ? s2 : s1;
        const bool need = q >= 0 && atomicMax(&L.stamp[(int64_t)(st & 1) * L.cap + q], st) < st;
        const uint64_t m = __ballot(need);
        uint32_t base = 0;
        if (lane == 0 && m) base = atomicAdd(&L.count[st % 3], (uint32_t)__popcll(m));
        base = __shfl(base, 0);
        if (need) L.items[(int64_t)(st % 3) * L.cap + base + __popcll(m & lanes_below_v())] = (uint32_t)q;
    }
}

// A = B = (G - 1) + labels; tiles holding a positive value queue their neighbourhood
// for steps 0 and 1
__global__ __launch_bounds__(VB) void k_vor_init_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ ts,
                                                      int nimg, const uint8_t* G, float* A, float* B, VorLists L) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(ts, nimg, b);
    const BImg& im = imgs[j];
    const Tiles T = make_tiles(im.g);
    if (vor_init_tile(MaskedU8{G + im.off, B + im.off}, A + im.off, B + im.off, im.g, T, b - ts[j], b, 0, nullptr, L.settled) &&
        threadIdx.x < 64)
        vor_queue(T, b - ts[j], ts[j], 0, 1, L);
}

__global__ __launch_bounds__(VB) void k_vor_step_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ ts,
                                                      int nimg, const float* __restrict__ src,
                                                      float* __restrict__ dst, int t, VorLists L, uint32_t* gflags,
                                                      uint32_t* gany, int ring) {
    const uint32_t n = L.count[t % 3];
    const uint32_t* items = L.items + (int64_t)(t % 3) * L.cap;
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
        const int64_t q = items[e];
        const int j = find_seg(ts, nimg, q);
        if (t > 0 && gflags[(int64_t)((t - 1) % ring) * nimg + j] == 0) continue;  // this image has converged
        if (L.settled[q]) continue;
        const BImg& im = imgs[j];
        const Tiles T = make_tiles(im.g);
        if (vor_tile_update(src + im.off, dst + im.off, im.g, T, q - ts[j], t, L.settled + q)) {
            if (threadIdx.x == 0) {
                gflags[(int64_t)(t % ring) * nimg + j] = 1;
                gany[t % ring] = 1;
            }
            if (threadIdx.x < 64) vor_queue(T, q - ts[j], ts[j], t + 1, t + 2, L);
        }
    }
}

__global__ void k_vor_list_reset(uint32_t* count, int which) { count[which] = 0; }

__global__ __launch_bounds__(256) void k_mask_out_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ vs,
                                                       int nimg, const float* A, const uint8_t* G) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(vs, nimg, b);
    const BImg& im = imgs[j];
    const int64_t n = im.g.n(), i0 = (b - vs[j]) * BV, i1 = i0 + BV < n ? i0 + BV : n;
    const DBuf d{im.dst, im.dst_dt};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) st(d, i, G[im.off + i] ? A[im.off + i] : 0.f);
}

// pinned host staging of the batch tables, released once the stream has drained
// (also when a later step fails, the copy out of it may still be in flight)
struct PinnedStage {
    void* p = nullptr;
    hipStream_t s = nullptr;
    ~PinnedStage() {
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipHostFree(p);
        }
    }
};

static int voronoi_otsu_batch(const cle_buf* srcs, const cle_buf* dsts, const float* ss, const float* so, int n,
                              int* n_steps, hipStream_t s) {
    std::vector<BImg> imgs((size_t)n);
    int64_t total = 0, ntiles = 0;
    for (int j = 0; j < n; ++j) {
        Geom g, gd;
        CLE_TRY(check_buf(&srcs[j], "source", &g));
        CLE_TRY(check_buf(&dsts[j], "label_image_destination", &gd));
        CLE_TRY(check_same_shape(g, gd, "source", "label_image_destination"));
        BImg& im = imgs[j];
        memset(&im, 0, sizeof(im));
        im.g = g;
        im.off = total;
        total += (g.n() + 63) / 64 * 64;
        im.src = srcs[j].data;
        im.src_dt = srcs[j].dtype;
        im.dst = dsts[j].data;
        im.dst_dt = dsts[j].dtype;
        im.big = g.n() >= ((int64_t)1 << 24) ? 1 : 0;
        im.tbase = ntiles;
        ntiles += g.n() > 0 ? make_tiles(g).count() : 0;
    }
    if (total == 0) return CLE_OK;
    CclBatchSizes csz;
    ccl_batch_layout(imgs, &csz);
    // packed work images: F (float32: blurs, CCL labels, growth buffer B), A (growth
    // buffer A), S (uint8 seeds), G (uint8 segmentation)
    Scratch F, A, S, G, aux, tf, gf;
    CLE_TRY(F.alloc((size_t)total * 4, s));
    CLE_TRY(A.alloc((size_t)total * 4, s));
    CLE_TRY(S.alloc((size_t)total, s));
    CLE_TRY(G.alloc((size_t)total, s));
    const size_t otsu_words = (size_t)n * (2 + 256 + 256);
    CLE_TRY(aux.alloc((size_t)n * 8 + (size_t)n * 8 + otsu_words * 4, s));
    float* mm = aux.get<float>();
    double* thr = (double*)(aux.get<char>() + (size_t)n * 8);
    uint32_t* ou = (uint32_t*)(aux.get<char>() + (size_t)n * 16);
    const int64_t tcap = ntiles > 0 ? ntiles : 1;
    CLE_TRY(tf.alloc((size_t)tcap * 3 * 4 + 16 + (size_t)tcap * 2 * 4 + (size_t)tcap, s));
    VorLists VL;
    VL.items = tf.get<uint32_t>();
    VL.count = VL.items + (size_t)tcap * 3;
    VL.stamp = (int32_t*)(VL.count + 4);
    VL.settled = (uint8_t*)(VL.stamp + (size_t)tcap * 2);
    VL.cap = tcap;
    CLE_TRY(gf.alloc(((size_t)VOR_RING * n + VOR_RING) * 4, s));
    uint32_t* gflags = gf.get<uint32_t>();
    uint32_t* gany = gflags + (size_t)VOR_RING * n;
    float* Fp = F.get<float>();
    float* Ap = A.get<float>();
    uint8_t* Sp = S.get<uint8_t>();
    uint8_t* Gp = G.get<uint8_t>();

    Blob blob;
    const size_t imgs_off = blob.put_vec(imgs);
    int64_t vblocks = 0;
    const std::vector<int64_t> vst = voxel_block_starts(imgs, &vblocks);
    const size_t vs_off = blob.put_vec(vst);
    std::vector<int64_t> tst((size_t)n + 1);
    for (int j = 0; j < n; ++j) tst[j] = imgs[j].tbase;
    tst[n] = ntiles;
    const size_t ts_off = blob.put_vec(tst);
    std::vector<int64_t> rst((size_t)n + 1);
    int64_t nrows = 0;
    for (int j = 0; j < n; ++j) {
        rst[j] = nrows;
        nrows += imgs[j].g.n() > 0 ? imgs[j].g.D * imgs[j].g.H : 0;
    }
    rst[n] = nrows;
    const size_t rs_off = blob.put_vec(rst);

    std::vector<Launch> L;
    std::vector<std::shared_ptr<void>> keep;
    std::vector<BlurReq> r1((size_t)n), r2((size_t)n);
    for (int j = 0; j < n; ++j) {
        r1[j] = BlurReq{imgs[j].src, imgs[j].src_dt, Fp + imgs[j].off, imgs[j].g, ss[j]};
        r2[j] = BlurReq{imgs[j].src, imgs[j].src_dt, Fp + imgs[j].off, imgs[j].g, so[j]};
    }
    // spots = detect_maxima_box(gaussian_blur(src, spot_sigma))
    CLE_TRY(plan_blur_batch(r1, blob, L, s));
    L.push_back([=](const char* dev) -> int {
        k_maxima_batch<<<(unsigned)nrows, 256, 0, s>>>((const BImg*)(dev + imgs_off), (const int64_t*)(dev + rs_off),
                                                       n, Fp, Sp);
        CLE_LAUNCHED();
        return CLE_OK;
    });
    // segmentation = threshold_otsu(gaussian_blur(src, outline_sigma)); seeds = spots AND segmentation
    CLE_TRY(plan_blur_batch(r2, blob, L, s));
    CLE_TRY(plan_otsu_batch(imgs, imgs_off, vs_off, vblocks, Fp, mm, thr, ou, blob, L, s));
    L.push_back([=](const char* dev) -> int {
        k_seg_and_batch<<<(unsigned)vblocks, 256, 0, s>>>((const BImg*)(dev + imgs_off),
                                                          (const int64_t*)(dev + vs_off), n, Fp, thr, Sp, Gp);
        CLE_LAUNCHED();
        return CLE_OK;
    });
    // seed labels (float32, into F), then A = F = (segmentation - 1) + labels
    CLE_TRY(plan_ccl_batch(imgs, imgs_off, csz, Sp, Fp, blob, L, keep, s));
    if (ntiles > 0)
        L.push_back([=](const char* dev) -> int {
            CLE_TRY(fill_u32(VL.count, 0u, 3, s));
            CLE_TRY(fill_u32((uint32_t*)VL.stamp, 0xFFFFFFFFu, (size_t)VL.cap * 2, s));
            k_vor_init_batch<<<(unsigned)ntiles, VB, 0, s>>>((const BImg*)(dev + imgs_off),
                                                             (const int64_t*)(dev + ts_off), n, Gp, Ap, Fp, VL);
            CLE_LAUNCHED();
            return CLE_OK;
        });

    // the tables go to the device with one copy
    Scratch dblob;
    CLE_TRY(dblob.alloc(blob.h.size(), s));
    PinnedStage pin;
    pin.s = s;
    CLE_HIP(hipHostMalloc(&pin.p, blob.h.size()));
    memcpy(pin.p, blob.h.data(), blob.h.size());
    CLE_HIP(hipMemcpyAsync(dblob.p, pin.p, blob.h.size(), hipMemcpyHostToDevice, s));
    const char* dev = dblob.get<char>();
    for (auto& l : L) CLE_TRY(l(dev));

    // growth: all images together, VOR_BATCH steps between host checks
    const BImg* dimgs = (const BImg*)(dev + imgs_off);
    const int64_t* dts = (const int64_t*)(dev + ts_off);
    int t = 0, steps = -1;
    uint32_t hany[VOR_BATCH];
    const unsigned vgrid = (unsigned)(ntiles < 8192 ? (ntiles > 0 ? ntiles : 1) : 8192);
    while (ntiles > 0 && steps < 0) {
        CLE_TRY(fill_u32(gflags + (size_t)(t % VOR_RING) * n, 0u, (size_t)VOR_BATCH * n, s));
        CLE_TRY(fill_u32(gany + t % VOR_RING, 0u, VOR_BATCH, s));
        for (int k = 0; k < VOR_BATCH; ++k) {
            const int it = t + k;
            const float* src = (it & 1) ? Fp : Ap;
            float* dst = (it & 1) ? Ap : Fp;
            k_vor_list_reset<<<1, 1, 0, s>>>(VL.count, (it + 2) % 3);  // consumed by step it - 1
            k_vor_step_batch<<<vgrid, VB, 0, s>>>(dimgs, dts, n, src, dst, it, VL, gflags, gany, VOR_RING);
        }
        CLE_LAUNCHED();
        CLE_HIP(hipMemcpyAsync(hany, gany + t % VOR_RING, sizeof(hany), hipMemcpyDeviceToHost, s));
        CLE_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < VOR_BATCH; ++k)
            if (hany[k] == 0) { steps = t + k + 1; break; }
        t += VOR_BATCH;
    }
    if (n_steps) *n_steps = steps;
    k_mask_out_batch<<<(unsigned)vblocks, 256, 0, s>>>(dimgs, (const int64_t*)(dev + vs_off), n, Ap, Gp);
    CLE_LAUNCHED();
    // the pinned tables and the scratch are released once the work has run
    CLE_HIP(hipStreamSynchronize(s));
    return CLE_OK;
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_masked_voronoi_labeling(const cle_buf* binary, const cle_buf* mask, const cle_buf* dst, int* n_iterations,
                                void* stream) {
    Geom g, gm, gd;
    CLE_TRY(check_buf(binary, "binary_source", &g));
    CLE_TRY(check_buf(mask, "mask_image", &gm));
    CLE_TRY(check_buf(dst, "labeling_destination", &gd));
    CLE_TRY(check_same_shape(g, gm, "binary_source", "mask_image"));
    CLE_TRY(check_same_shape(g, gd, "binary_source", "labeling_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch a, b;
    CLE_TRY(a.alloc((size_t)n * 4, s));
    CLE_TRY(b.alloc((size_t)n * 4, s));
    return masked_voronoi_core(dbuf(binary), dbuf(mask), dbuf(dst), g, a.get<float>(), b.get<float>(), n_iterations, s);
}

int cle_extend_labeling_via_voronoi(const cle_buf* src, const cle_buf* dst, int* n_iterations, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "labeling_source", &g));
    CLE_TRY(check_buf(dst, "labeling_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "labeling_source", "labeling_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch a, b, tf;
    CLE_TRY(a.alloc((size_t)n * 4, s));
    CLE_TRY(b.alloc((size_t)n * 4, s));
    const Tiles T = make_tiles(g);
    CLE_TRY(tf.alloc((size_t)T.count() * 5, s));  // 4 "changed" slots + settled
    k_vor_init_labels<<<(unsigned)T.count(), VB, 0, s>>>(dbuf(src), a.get<float>(), b.get<float>(), g, T,
                                                        tf.get<uint8_t>());
    CLE_LAUNCHED();
    CLE_TRY(grow(a.get<float>(), b.get<float>(), g, T, tf.get<uint8_t>(), n_iterations, s));
    k_store_out<<<vg(n), 256, 0, s>>>(a.get<float>(), dbuf(dst), n);
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_voronoi_otsu_labeling(const cle_buf* src, const cle_buf* dst, float spot_sigma, float outline_sigma,
                              void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "label_image_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "label_image_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch blurred, spots, seg, a, aux;
    CLE_TRY(blurred.alloc((size_t)n * 4, s));
    CLE_TRY(spots.alloc((size_t)n, s));
    CLE_TRY(seg.alloc((size_t)n, s));
    CLE_TRY(aux.alloc(64, s));
    float* mm = aux.get<float>();
    double* thr = (double*)(aux.get<char>() + 16);
    cle_buf bb = *src;
    bb.data = blurred.p;
    bb.dtype = CLE_F32;
    // spots = detect_maxima_box(gaussian_blur(src, spot_sigma))
    const float s1[3] = {spot_sigma, spot_sigma, spot_sigma};
    CLE_TRY(blur_impl(src, &bb, s1, s));
    CLE_TRY(maxima_impl(DBuf{blurred.p, CLE_F32}, DBuf{spots.p, CLE_U8}, g, s));
    // segmentation = threshold_otsu(gaussian_blur(src, outline_sigma)); binary = spots AND segmentation
    const float s2[3] = {outline_sigma, outline_sigma, outline_sigma};
    CLE_TRY(blur_impl(src, &bb, s2, s));
    CLE_TRY(otsu_impl(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    k_seg_and<<<vg(n), 256, 0, s>>>(blurred.get<float>(), thr, spots.get<uint8_t>(), seg.get<uint8_t>(), n);
    CLE_LAUNCHED();
    // masked Voronoi of the seeds inside the segmentation; the blurred buffer is
    // reused as the second growth buffer
    CLE_TRY(a.alloc((size_t)n * 4, s));
    return masked_voronoi_core(DBuf{spots.p, CLE_U8}, DBuf{seg.p, CLE_U8}, dbuf(dst), g, a.get<float>(),
                               blurred.get<float>(), nullptr, s);
}

int cle_voronoi_otsu_labeling_batch(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                    const float* outline_sigmas, int n_images, int* n_steps, void* stream) {
    if (n_images < 0) return fail(CLE_E_VALUE, "n_images must be >= 0");
    if (n_images == 0) return CLE_OK;
    if (!srcs || !dsts || !spot_sigmas || !outline_sigmas) return fail(CLE_E_VALUE, "null batch argument");
    return voronoi_otsu_batch(srcs, dsts, spot_sigmas, outline_sigmas, n_images, n_steps, (hipStream_t)stream);
}

int cle_gauss_otsu_labeling(const cle_buf* src, const cle_buf* dst, float outline_sigma, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "label_image_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "label_image_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch blurred, seg, aux;
    CLE_TRY(blurred.alloc((size_t)n * 4, s));
    CLE_TRY(seg.alloc((size_t)n, s));
    CLE_TRY(aux.alloc(64, s));
    float* mm = aux.get<float>();
    double* thr = (double*)(aux.get<char>() + 16);
    cle_buf bb = *src;
    bb.data = blurred.p;
    bb.dtype = CLE_F32;
    const float sg[3] = {outline_sigma, outline_sigma, outline_sigma};
    CLE_TRY(blur_impl(src, &bb, sg, s));
    CLE_TRY(otsu_impl(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    k_threshold_u8<<<vg(n), 256, 0, s>>>(blurred.get<float>(), thr, seg.get<uint8_t>(), n);
    CLE_LAUNCHED();
    return ccl_impl(DBuf{seg.p, CLE_U8}, g, dbuf(dst), CLE_CONN_BOX, CLE_FP32_KEY_COMPAT, nullptr, s);
}

}  // extern "C"
