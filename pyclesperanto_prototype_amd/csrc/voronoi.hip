// Masked Voronoi growth and the composite pipelines built on it.
//
// Reference semantics (pyclesperanto_prototype/...):
//   masked_voronoi_labeling  _tier5/_masked_voronoi_labeling.py:44-74:
//       flup = mask - 1; flop = CCL_box(binary); flip = flup + flop (float32); then
//       iteration it (even: box, flip -> flop; odd: diamond, flop -> flip) copies
//       non-zero voxels and gives a zero voxel the largest positive neighbour value
//       (_tier1/_onlyzero_overwrite_maximum_{box,diamond}.py), until an iteration
//       changes nothing; finally mask(last, mask).
//   extend_labeling_via_voronoi  _tier4/_extend_labeling_via_voronoi.py:39-61 (the same
//       growth from a label image, no mask)
//   voronoi_otsu_labeling    _tier9/_voronoi_otsu_labeling.py:57-73
//   gauss_otsu_labeling      _tier9/_gauss_otsu_labeling.py:6-49
//
// MI355X design.  One launch runs a pair of steps, box step 2b then diamond step
// 2b + 1, on every active tile: the tile and a 2-voxel halo are read from the source
// buffer (2D: staged in LDS; 3D: one z column of the region per thread, in registers),
// the box step is computed on the tile plus a 1-voxel ring, the diamond step on the
// tile, and only the tile is written to the destination buffer.  Across tiles this is
// exactly the reference's synchronous (Jacobi) steps.
//
// Stopping.  The reference stops after the first step that changes nothing.  A pair
// launch reports, per image, whether its box step and its diamond step changed a
// voxel.  If the box step changed nothing, no zero voxel has a positive box
// neighbour, so the diamond step changes nothing either and the pair's output is the
// reference's result; if only the diamond step changed nothing, the pair's output is
// the state after it, again the reference's result.  So an image stops after the
// first pair that is not "both changed", with reference iteration count
// 2b + (box changed ? 2 : 1), and its result is the destination of that pair.
//
// Work lists.  Tile T is processed in pair b iff a tile of its 3^d tile neighbourhood
// changed in pair b - 1 (pair 0: holds a positive initial value).  Otherwise T's
// neighbourhood is what it was one pair earlier, so T's output equals the output of
// the pair before, which the destination buffer (ping-pong) already holds.  A tile
// that changes queues its neighbourhood for the next pair (deduplicated by a per-tile
// stamp); a tile with no zero voxel that did not change holds the same values in both
// buffers and is never visited again ("settled").  Each launch walks its list with a
// fixed grid, so the cost follows the growth front.  Block 0 of each launch closes the
// previous pair (per-image convergence, list and flag resets; three lists and three
// flag sets rotate so that nothing it resets is in use), and a one-block launch closes
// the last pair of each chunk of pairs with one "still growing" word for the host.
// (No grid-wide "last block" ticket: the device-scope fence it needs writes back the
// L2 in every block.)
#include "common.hpp"
#include "growth.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>

namespace cle {

int maxima_impl(DBuf src, DBuf dst, const Geom& g, hipStream_t s);

constexpr int VB = 256;  // threads per tile block
constexpr int VPT = 4;   // voxels per thread (tile = 1024 voxels)
constexpr int GH = 2;    // halo of the staged region = steps per launch

// Tiles of 1024 voxels, close to cubes so that a growth front keeps as few voxels
// active as possible: 16 x 8 x 8 in 3D, 32 x 32 in 2D (x rows of 64 / 128 bytes)
struct Tiles {
    int tx, ty, tz;         // tile extents
    int64_t nz, ny, nx;     // tile counts
    __host__ __device__ int64_t count() const { return nz * ny * nx; }
};

__host__ __device__ inline bool vor_3d(const Geom& g) { return g.ndim == 3 && g.D > 1; }

__host__ __device__ inline Tiles make_tiles(const Geom& g) {
    Tiles t;
    if (vor_3d(g)) { t.tx = 16; t.ty = 8; t.tz = 8; }
    else { t.tx = 32; t.ty = 32; t.tz = 1; }
    t.nz = (g.D + t.tz - 1) / t.tz;
    t.ny = (g.H + t.ty - 1) / t.ty;
    t.nx = (g.W + t.tx - 1) / t.tx;
    return t;
}

constexpr int64_t VOR_MAX_TILES_AXIS = 65535;
constexpr int VOR_MAX_IMAGES = 65535;

// Per-tile state of the growth (one set for a single image or a whole batch; image j's
// tiles start at its tbase).
struct VorLists {
    int32_t* mark;     // 2 x cap: mark[b & 1][t] == b iff tile t changed in pair b (tiles
                       // with a seed: mark[1][t] == -1, "changed" before pair 0)
    uint8_t* settled;  // cap: the tile holds no zero and never changes again (1), or it
                       // holds no mask voxel and its state is not stored (2, empty2 only)
    int64_t cap;       // tiles of all images
    int empty2;        // 3D tiles flagged 2 are read as -1 (their buffers hold no state)
};

// Per-image state of the growth.
struct VorState {
    uint32_t* bits;    // 3 x nimg: per pair b % 3 and image, OR of the change bits (1 box, 2 diamond)
    int32_t* iters;    // per image: reference iteration count once converged, else 0
    uint32_t* word;    // [0]: an image still grows after the last pair; [1]: max iteration count
    uint32_t* ctr;     // VOR_NCTR counters of the tiles the pair launches processed, 16 words apart
};

// Tiles processed, counted per block into one of VOR_NCTR counters a cache line apart
// (block % VOR_NCTR): atomics from many blocks on one word serialise.
constexpr int VOR_NCTR = 64;
__device__ __forceinline__ void vor_count_tiles(const VorState& S, uint32_t n) {
    if (threadIdx.x == 0 && n) atomicAdd(&S.ctr[(blockIdx.x % VOR_NCTR) * 16], n);
}

// A = B = val(i) on one tile; the tile is settled when it holds no zero; returns, in
// every thread, whether the tile holds a positive value.
// `val(i)` is the initial value of voxel i (it may read B[i] before B[i] is written).
template <typename VF>
__device__ __forceinline__ bool vor_init_tile(VF val, float* A, float* B, const Geom& g, const Tiles& T,
                                              int64_t tile, uint8_t* settled) {
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    bool pos = false, zero = false;
    bool done = false;
    if constexpr (VF::kVec) {
        // rows of a multiple of 4 voxels: four consecutive x voxels per thread (tile rows
        // of 16 / 32 voxels start at multiples of 4; every buffer is 16-byte aligned)
        static_assert(VB * 4 == 1024, "one group of four voxels per thread and tile");
        if ((g.W & 3) == 0) {
            const int e = 4 * threadIdx.x;
            const int lx = e % T.tx, ly = (e / T.tx) % T.ty, lz = e / (T.tx * T.ty);
            const int64_t x = tx * T.tx + lx, y = ty * T.ty + ly, z = tz * T.tz + lz;
            if (lz < T.tz && x < g.W && y < g.H && z < g.D) {
                const int64_t i = (z * g.H + y) * g.W + x;
                float v[4];
                val.vec4(i, v);
                const float4 f = make_float4(v[0], v[1], v[2], v[3]);
                *(float4*)(A + i) = f;
                *(float4*)(B + i) = f;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    pos |= v[c] > 0.f;
                    zero |= v[c] == 0.f;
                }
            }
            done = true;
        }
    }
    for (int k = 0; k < VPT && !done; ++k) {
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
    if (threadIdx.x == 0) *settled = anyzero ? 0 : 1;
    return any != 0;
}

// masked growth: (mask - 1) + CCL label, the label held in B (float32 add, as the
// reference's add_image_and_scalar then add_images)
struct MaskedAny {
    static constexpr bool kVec = false;
    DBuf mask;
    const float* B;
    __device__ __forceinline__ float operator()(int64_t i) const { return (ld(mask, i) - 1.f) + B[i]; }
    __device__ __forceinline__ void vec4(int64_t, float*) const {}
};
struct MaskedU8 {
    static constexpr bool kVec = true;  // batch buffers: 16-byte aligned image slots
    const uint8_t* mask;
    const float* B;
    __device__ __forceinline__ float operator()(int64_t i) const { return ((float)mask[i] - 1.f) + B[i]; }
    __device__ __forceinline__ void vec4(int64_t i, float* v) const {
        const uchar4 m = *(const uchar4*)(mask + i);
        const float4 b = *(const float4*)(B + i);
        v[0] = ((float)m.x - 1.f) + b.x;
        v[1] = ((float)m.y - 1.f) + b.y;
        v[2] = ((float)m.z - 1.f) + b.z;
        v[3] = ((float)m.w - 1.f) + b.w;
    }
};
// unmasked growth of an existing label image: the labels converted to float32 as the
// reference's copy into a float32 buffer does
struct LabelsAny {
    static constexpr bool kVec = false;
    DBuf src;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(src, i); }
    __device__ __forceinline__ void vec4(int64_t, float*) const {}
};

// single image: init every tile; a tile holding a positive value counts as changed
// before pair 0 (its neighbourhood runs in pair 0).  Consecutive tiles go to one XCD.
template <typename VF>
__global__ __launch_bounds__(VB) void k_vor_init(VF val, float* A, float* B, Geom g, VorLists L) {
    const Tiles T = make_tiles(g);
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
    if (vor_init_tile(val, A, B, g, T, t, L.settled + t) && threadIdx.x == 0) L.mark[L.cap + t] = -1;
}

// The 2D form of a pair of steps: a 32 x 32 tile and a GH-voxel halo staged in LDS.
struct PairGeo2 {
    static constexpr int TX = 32, TY = 32;
    static constexpr int RX = TX + 2 * GH, RY = TY + 2 * GH;  // region extents
    static constexpr int NR = RX * RY;
    static constexpr int KV = (NR + VB - 1) / VB;  // region cells per thread (tid + k * VB)
    int64_t obase;           // offset of the region origin (may lie outside the image)
    int W;
    int xlo, xhi, ylo, yhi;  // the part inside the image, region coordinates [lo, hi)
    bool interior;           // the whole region lies inside the image
    __device__ __forceinline__ PairGeo2(const Geom& g, int tx, int ty) {
        const int64_t x0 = (int64_t)tx * TX - GH, y0 = (int64_t)ty * TY - GH;
        xlo = x0 < 0 ? (int)-x0 : 0;
        ylo = y0 < 0 ? (int)-y0 : 0;
        xhi = (int)(g.W - x0 < RX ? g.W - x0 : RX);
        yhi = (int)(g.H - y0 < RY ? g.H - y0 : RY);
        W = (int)g.W;
        obase = y0 * g.W + x0;
        interior = xlo == 0 && ylo == 0 && xhi == RX && yhi == RY;
    }
};

// issue the loads of a tile's region (voxels inside the image) into registers
__device__ __forceinline__ void vor_stage_load2(const float* __restrict__ src, const PairGeo2& R, float* v) {
    using P = PairGeo2;
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < P::KV; ++k) {
        const int c = tid + k * VB;
        const int rx = c % P::RX, ry = c / P::RX;
        const bool in = c < P::NR && (R.interior || (rx >= R.xlo && rx < R.xhi && ry >= R.ylo && ry < R.yhi));
        v[k] = in ? src[R.obase + (ry * R.W + rx)] : 0.f;
    }
}

__device__ __forceinline__ void vor_stage_store2(const float* v, float* lds) {
    using P = PairGeo2;
#pragma unroll
    for (int k = 0; k < P::KV; ++k)
        if (threadIdx.x + k * VB < P::NR) lds[threadIdx.x + k * VB] = v[k];
}

// One pair of steps (box, then diamond) on a staged tile: lds holds the region, lds2
// receives the box step's result on the tile + 1-voxel ring, the diamond step is
// written to dst.  Returns, in every thread, bit 0: a tile voxel changed in the box
// step, bit 1: in the diamond step, bit 2: a tile voxel is zero after the pair.
// Neighbour coordinates are clamped to the image, which keeps them inside the staged
// region (interior tiles need no clamping: every LDS offset is then a constant).
__device__ __forceinline__ int vor_pair_compute2(float* __restrict__ dst, const PairGeo2& R, const float* lds,
                                                 float* lds2, int* sbits) {
    using P = PairGeo2;
    constexpr int TX = P::TX, TY = P::TY, RX = P::RX;
    const int tid = threadIdx.x;
    const bool inr = R.interior;
    const int xlo = inr ? 0 : R.xlo, xhi = inr ? RX : R.xhi, ylo = inr ? 0 : R.ylo, yhi = inr ? P::RY : R.yhi;
    // a tile without a zero voxel keeps its values (non-zero voxels never change): its
    // output is its input, whatever the halo holds
    {
        bool zero = false;
#pragma unroll
        for (int k = 0; k < (TX * TY) / VB; ++k) {
            const int c = tid + k * VB;
            const int rx = c % TX + GH, ry = c / TX + GH;
            if (rx < xhi && ry < yhi) zero |= lds[ry * RX + rx] == 0.f;
        }
        if (!__syncthreads_or(zero ? 1 : 0)) {
#pragma unroll
            for (int k = 0; k < (TX * TY) / VB; ++k) {
                const int c = tid + k * VB;
                const int rx = c % TX + GH, ry = c / TX + GH;
                if (rx < xhi && ry < yhi) dst[R.obase + (ry * R.W + rx)] = lds[ry * RX + rx];
            }
            __syncthreads();  // lds is read above; the next tile's staging overwrites it
            return 0;         // no change, no zero
        }
    }
    int bits = 0;
    // box step on the tile + 1-voxel ring
    constexpr int SX = TX + 2, SY = TY + 2, NS = SX * SY;
#pragma unroll 2
    for (int c = tid; c < NS; c += VB) {
        const int rx = c % SX + 1, ry = c / SX + 1;
        if (rx < xlo || rx >= xhi || ry < ylo || ry >= yhi) continue;
        const int rc = ry * RX + rx;
        const float v = lds[rc];
        float out = v;
        if (v == 0.f) {
            const int dxm = rx > xlo ? -1 : 0, dxp = rx < xhi - 1 ? 1 : 0;
            const int dym = ry > ylo ? -RX : 0, dyp = ry < yhi - 1 ? RX : 0;
            const float* p = lds + rc;
            float best = fmaxf(0.f, fmaxf(fmaxf(p[dym + dxm], p[dym]), p[dym + dxp]));
            best = fmaxf(best, fmaxf(fmaxf(p[dxm], p[0]), p[dxp]));
            best = fmaxf(best, fmaxf(fmaxf(p[dyp + dxm], p[dyp]), p[dyp + dxp]));
            out = best;
            const bool inner = rx >= GH && rx < TX + GH && ry >= GH && ry < TY + GH;
            if (inner && out != 0.f) bits |= 1;
        }
        lds2[rc] = out;
    }
    __syncthreads();
    // diamond step on the tile
#pragma unroll
    for (int k = 0; k < (TX * TY) / VB; ++k) {
        const int c = tid + k * VB;
        const int rx = c % TX + GH, ry = c / TX + GH;
        if (rx >= xhi || ry >= yhi) continue;
        const int rc = ry * RX + rx;
        const float v = lds2[rc];
        float out = v;
        if (v == 0.f) {
            float best = fmaxf(0.f, lds2[rc + (rx > xlo ? -1 : 0)]);
            best = fmaxf(best, lds2[rc + (rx < xhi - 1 ? 1 : 0)]);
            best = fmaxf(best, lds2[rc + (ry > ylo ? -RX : 0)]);
            best = fmaxf(best, lds2[rc + (ry < yhi - 1 ? RX : 0)]);
            out = best;
            bits |= out != 0.f ? 2 : 4;
        }
        dst[R.obase + (ry * R.W + rx)] = out;
    }
    if (bits) atomicOr(sbits, bits);
    __syncthreads();
    return *sbits;
}

// Pair b over runs: a run is up to vrun (<= VOR_RUN) consecutive tiles of one tile row
// (x) of one image, 2D or 3D by D3.  Run r of a launch belongs to image find_seg(rs, r) (rs:
// run prefix over the images, an image of the other kind contributing none; null for
// one image); block b takes run xcd_tile(b), so the runs of one XCD form a contiguous
// slab of rows, processed in raster order (the halo rows a run shares with the runs
// above and below are read while in that XCD's L2).  The block gathers, with three
// loads per lane and ballots, the change stamps of the 3 x 3 rows of VOR_RUN + 2 tiles
// around its run and the settled flags of the run: tile t runs iff a neighbour
// changed in pair b - 1 and t is not settled.  It then works through the active tiles
// software-pipelined: the region of the next one is loaded into registers while the
// current one is computed from LDS.  A tile that changes stamps mark[b & 1][t] = b
// (one plain store).  The loads read src, which no block writes in this launch; only
// the block that processes a tile sets its settled flag, after it read it.  Block 0
// first closes pair b - 1: images whose pair b - 1 was not "both changed" get their
// iteration count; it also empties the change bits pair b + 1 will fill (both
// idempotent: the 2D and the 3D launch of a pair both do it).
#ifndef CLE_VOR_RUN
#define CLE_VOR_RUN 64
#endif
constexpr int VOR_RUN = CLE_VOR_RUN;  // longest run (<= 64)
#ifndef CLE_VOR_RUN_BATCH
#define CLE_VOR_RUN_BATCH 32
#endif

// Tiles per run: VOR_RUN for a single image from 4096 tiles on, CLE_VOR_RUN_BATCH in a
// batch (images enough to fill the chip); one for a smaller single image, where a few
// blocks working through runs one tile after another would leave the chip idle (its
// growth is latency-bound).  A block's start (its run's stamps, the first region) is
// not overlapped with work, so longer runs pay it less often.  Measured, growth per
// call: 512 x 2048 x 2048 stack 47 / 27.7 / 19.6 / 19.1 ms with runs of 8 / 16 / 32 / 64,
// 256 x 1024 x 1024 stack 3.7 / 2.4 / 2.25 ms with 16 / 32 / 64; mixed 2000-image batch
// 180 / 175 / 177 ms with 16 / 32 / 64 (and 438 ms instead of 387 ms with one tile per
// run below 4096 tiles).
#ifndef CLE_VOR_MIN_RUNS
#define CLE_VOR_MIN_RUNS 1024
#endif
inline int vor_run_length(const Geom& g, bool batch) {
    if (batch) return CLE_VOR_RUN_BATCH;
    const int64_t n = make_tiles(g).count();
    if (n < 4096) return 1;
    // the longest run (a power of two up to VOR_RUN) that leaves CLE_VOR_MIN_RUNS runs
    // to fill the chip
    int rl = VOR_RUN;
    while (rl > 1 && n / rl < CLE_VOR_MIN_RUNS) rl >>= 1;
    return rl;
}
// runs of an image of run length rl: pieces of rl tiles of every tile row
inline int64_t vor_run_blocks(const Geom& g, int rl) {
    if (g.n() == 0) return 0;
    const Tiles T = make_tiles(g);
    return T.nz * T.ny * ((T.nx + rl - 1) / rl);
}

// The start of a pair launch's block: block 0 closes pair b - 1 (see above); the block's
// run (image j, tile row, first tile) and the bit mask of its active tiles.  Every
// wave computes the same mask; returns false when the run has nothing to do.
struct VorRun {
    int j, xs, ty, tz, nrun;
    int64_t t0;    // first tile of the run in the per-tile arrays
    uint64_t act;  // bit k: tile xs + k runs in this pair
};
__device__ __forceinline__ bool vor_run_begin(const BImg* __restrict__ imgs, const int64_t* __restrict__ rs, int nimg,
                                              int b, const VorLists& L, const VorState& S, VorRun& R) {
    const uint32_t* prev = S.bits + (size_t)((b + 2) % 3) * nimg;  // pair b - 1
    if (blockIdx.x == 0) {
        uint32_t* next = S.bits + (size_t)((b + 1) % 3) * nimg;
        int mx = 0;
        for (int j = threadIdx.x; j < nimg; j += blockDim.x) {
            if (b > 0 && S.iters[j] == 0 && prev[j] != 3u) {
                const int it = 2 * (b - 1) + ((prev[j] & 1u) ? 2 : 1);
                S.iters[j] = it;
                mx = it > mx ? it : mx;
            }
            next[j] = 0;
        }
        if (mx) atomicMax((int*)&S.word[1], mx);
    }
    const int64_t rg = xcd_tile(blockIdx.x, gridDim.x);
    const int j = rs ? find_seg(rs, nimg, rg) : 0;
    if (S.iters[j] != 0 || (b > 0 && prev[j] != 3u)) return false;  // converged in an earlier pair
    const BImg& im = imgs[j];
    const int nx = im.vnx, ny = im.vny, nz = im.vnz;
    const int64_t tb = im.tbase;
    const int rl = im.vrun;  // tiles per run (<= VOR_RUN)
    const int64_t r = rs ? rg - rs[j] : rg;
    // runs in z-fastest order (then y, then x segment): the blocks an XCD runs at one
    // time take z-adjacent runs, whose staged regions share their z halo in that XCD's L2
    const int tz = (int)(r % nz);
    const int64_t rest = r / nz;
    const int ty = (int)(rest % ny);
    const int xs = (int)(rest / ny) * rl;
    const int64_t row = (int64_t)tz * ny + ty;
    const int nrun = nx - xs < rl ? nx - xs : rl;
    const int64_t t0 = tb + row * nx + xs;
    const int32_t* chg_prev = L.mark + (int64_t)((b - 1) & 1) * L.cap + tb;
    // active tiles of the run: bit k for tile t0 + k
    uint64_t act;
    {
        constexpr int RW = VOR_RUN + 2;  // stamps per neighbour row: x from xs - 1
        constexpr int NWB = (9 * RW + 63) / 64;  // ballot words of the 9 rows
        static_assert(VOR_RUN <= 64, "a run's active tiles fit one 64-bit mask");
        const int lane = threadIdx.x & 63;
        uint64_t w[NWB + 2];
        w[NWB] = w[NWB + 1] = 0;  // read (shifted out) past the last row
#pragma unroll
        for (int k = 0; k < NWB; ++k) {
            const int idx = lane + 64 * k;  // row idx / RW (dz major, then dy), column idx % RW
            const int rr = idx / RW, cx = idx - rr * RW;
            const int qx = xs - 1 + cx, qy = ty + rr % 3 - 1, qz = tz + rr / 3 - 1;
            bool c = false;
            if (idx < 9 * RW && cx < nrun + 2 && qx >= 0 && qx < nx && qy >= 0 && qy < ny && qz >= 0 && qz < nz)
                c = chg_prev[((int64_t)qz * ny + qy) * nx + qx] == b - 1;
            w[k] = __ballot(c);
        }
        // OR of the 9 rows: bit c of (m, mh) = some neighbour row changed at
        // x = xs - 1 + c (mh: bits 64 .. RW - 1)
        uint64_t m = 0, mh = 0;
#pragma unroll
        for (int rr = 0; rr < 9; ++rr) {
            const int o = rr * RW, wi = o >> 6, sh = o & 63;
            uint64_t lo = w[wi] >> sh, hi = 0;
            if (sh) {
                lo |= w[wi + 1] << (64 - sh);
                hi = w[wi + 1] >> sh | w[wi + 2] << (64 - sh);
            } else {
                hi = w[wi + 1];
            }
            if (RW < 64) lo &= (1ull << (RW < 64 ? RW : 0)) - 1;
            hi &= RW > 64 ? ((1ull << (RW > 64 ? RW - 64 : 0)) - 1) : 0ull;
            m |= lo;
            mh |= hi;
        }
        const bool settled = lane < nrun && L.settled[t0 + lane] != 0;
        const uint64_t st = __ballot(settled);
        // tile k runs if row bits k, k + 1 or k + 2 are set
        const uint64_t run_mask = nrun >= 64 ? ~0ull : ((1ull << nrun) - 1);
        act = ((m | (m >> 1) | (mh << 63) | (m >> 2) | (mh << 62)) & run_mask) & ~st;
    }
    R = VorRun{j, xs, ty, tz, nrun, t0, act};
    return act != 0;
}

// The 2D form: 32 x 32 tiles staged in LDS.
__global__ __launch_bounds__(VB, 4) void k_vor_pair_run2(const BImg* __restrict__ imgs, const int64_t* __restrict__ rs,
                                                       int nimg, const float* __restrict__ src, float* __restrict__ dst,
                                                       int b, VorLists L, VorState S) {
    using P = PairGeo2;
    __shared__ float lds[P::NR];
    __shared__ float lds2[P::NR];
    __shared__ int sbits[2];
    VorRun run;
    if (!vor_run_begin(imgs, rs, nimg, b, L, S, run)) return;
    uint32_t* cur = S.bits + (size_t)(b % 3) * nimg;
    const int j = run.j, xs = run.xs, ty = run.ty;
    const int64_t t0 = run.t0;
    uint64_t act = run.act;
    const BImg& im = imgs[j];
    const Geom g = im.g;
    src += im.off;
    dst += im.off;
    float v[P::KV];
    int tk = __builtin_ctzll(act);  // run index of the current tile
    act &= act - 1;
    vor_stage_load2(src, P(g, xs + tk, ty), v);
    int par = 0;
    uint32_t done = 0;
    int acc = 0;  // change bits of the block's tiles: one atomic per block
#pragma unroll 1
    for (;;) {
        ++done;
        const int tx = xs + tk;
        const P R(g, tx, ty);
        // sbits[par] was last read two tiles ago, before a barrier of the previous tile
        if (threadIdx.x == 0) sbits[par] = 0;
        vor_stage_store2(v, lds);
        __syncthreads();
        const bool more = act != 0;
        const int tn = more ? __builtin_ctzll(act) : tk;
        act &= act - 1;
        if (more) vor_stage_load2(src, P(g, xs + tn, ty), v);
        const int bits = vor_pair_compute2(dst, R, lds, lds2, &sbits[par]);
        par ^= 1;
        if (bits & 3) {
            acc |= bits & 3;
            if (threadIdx.x == 0) L.mark[(int64_t)(b & 1) * L.cap + t0 + tk] = b;
        } else if (!(bits & 4) && threadIdx.x == 0) {
            L.settled[t0 + tk] = 1;
        }
        // lds is rewritten by the next tile only after every thread passed the last
        // barrier of vor_pair_compute2 (its lds reads come before it)
        if (!more) break;
        tk = tn;
    }
    if (threadIdx.x == 0 && acc) atomicOr(&cur[j], (uint32_t)acc);
    vor_count_tiles(S, done);
}

// The 3D form on register columns.  One thread per (x, y) column of the staged region
// (the 16 x 8 x 8 tile and a 2-voxel halo: 20 x 12 columns of 12 voxels) holds the
// column in registers; region voxels outside the image hold the nearest image voxel
// (clamped loads), so the maxima over the reference's clamped neighbourhoods are plain
// maxima over the region.  The box step (tile + 1-voxel ring) is taken as separable
// maxima: over z in registers (C), over y through LDS (Y), over x through LDS; the
// diamond step (tile) reads its z neighbours from registers and its x / y neighbours
// from LDS.  Maxima are exact, so the order of the separable passes does not change a
// value.  Columns without a zero voxel skip the work whose result they do not need
// (sparse masks leave most columns of an active tile without one): the box maximum of
// a voxel is taken only where the voxel is zero, Y only next to such columns, the
// diamond only in them.  Four barriers per tile (the block's "tile holds a zero" vote
// is a flag word per tile parity set before the first, not a barrier of its own); the
// next tile's columns load while this one computes, and when it is the next tile in x
// its left halo columns come from this tile through LDS.  Every branch on the tile's
// z extent is uniform.  64 VGPRs: 7 blocks per CU (LDS bound).
#ifndef CLE_VOR3_BLOCKS
#define CLE_VOR3_BLOCKS 6  // blocks per CU the register budget must allow (<= 80 VGPRs)
#endif
// (at 8 blocks per CU registers spill; 4 and 5 blocks were measured 41 % and 21 %
// slower on the mixed batch)
constexpr int V3_RX = 20, V3_RY = 12, V3_RZ = 12;
constexpr int V3_COLS = V3_RX * V3_RY;  // 240 columns
constexpr int V3_NB = V3_RZ - 2;        // box planes: region z 1 .. 10
__global__ __launch_bounds__(VB, CLE_VOR3_BLOCKS) void k_vor_pair_run3(const BImg* __restrict__ imgs,
                                                                     const int64_t* __restrict__ rs, int nimg,
                                                                     const float* __restrict__ src,
                                                                     float* __restrict__ dst, int b, VorLists L,
                                                                     VorState S) {
    static_assert(VB >= V3_COLS, "one thread per region column");
    __shared__ float Cl[V3_NB * V3_COLS];  // C of every column, then the box result
    __shared__ float Yl[V3_NB * V3_COLS];
    __shared__ uint8_t Fl[V3_COLS];  // the column holds a zero among the box planes
    __shared__ float Hl[2][2 * V3_RY * V3_RZ];  // body columns 14, 15: the next tile's left halo
    __shared__ float Rl[2][2 * V3_RY * V3_RZ];  // body columns 0, 1: the previous tile's right halo
    __shared__ int sbits[2];
    __shared__ int zf[2];
    __shared__ uint64_t Em[9][2];  // empty2: the empty tiles of the 3 x 3 neighbour rows
    VorRun run;
    if (!vor_run_begin(imgs, rs, nimg, b, L, S, run)) return;
    uint32_t* cur = S.bits + (size_t)(b % 3) * nimg;
    const BImg& im = imgs[run.j];
    const Geom g = im.g;
    src += im.off;
    dst += im.off;
    const int tid = threadIdx.x;
    const bool col = tid < V3_COLS;
    const int rx = tid % V3_RX, ry = tid / V3_RX;
    const int64_t plane = g.H * g.W;
    const int64_t y0 = (int64_t)run.ty * 8 - 2, z0 = (int64_t)run.tz * 8 - 2;
    const int64_t yr = y0 + ry;
    const int64_t yy = yr < 0 ? 0 : (yr >= g.H ? g.H - 1 : yr);
    int64_t zoff[V3_RZ];
#pragma unroll
    for (int rz = 0; rz < V3_RZ; ++rz) {
        const int64_t z = z0 + rz;
        zoff[rz] = (z < 0 ? 0 : (z >= g.D ? g.D - 1 : z)) * plane;
    }
    // output voxel (column offset o, region plane rz inside the image)
    auto store = [&](int64_t o, int rz, float val) { dst[o + zoff[rz]] = val; };
    const bool ybox = col && ry >= 1 && ry <= V3_RY - 2;
    const bool ytile = col && ry >= 2 && ry < V3_RY - 2 && yr < g.H;
    const bool xbox = rx >= 1 && rx <= V3_RX - 2;
    // Halo columns shared by neighbouring active tiles of the run go through LDS instead
    // of being loaded again (they lie in the 64-byte lines of the neighbour's body): the
    // left halo (region x 0, 1) of tile t is tile t - 1's body columns 14, 15 (region x
    // 16, 17) when t - 1 is active, the right halo (region x 18, 19) is tile t + 1's body
    // columns 0, 1 (region x 2, 3) when t + 1 is active.  Regions load two active tiles
    // ahead, so both neighbours' bodies are in registers when a tile's halo is passed on.
    const uint64_t run_act = run.act;
    auto active = [&](int t) { return t >= 0 && t < 64 && ((run_act >> t) & 1ull) != 0; };
    // Tiles without a mask voxel (flag 2, empty2 only) store no state: their region
    // voxels read as -1.  Em[(dz + 1) * 3 + dy + 1] bit c: the tile (x: xs - 1 + c) of the
    // neighbour row (dy, dz) is empty; each thread's column lies in one neighbour row per
    // class of region planes (0, 1 / 2 .. 9 / 10, 11: z tiles of the clamped planes).
    const bool e2 = L.empty2 != 0;
    int erow[3] = {4, 4, 4};
    if (e2) {
        if (tid < 64) {
            constexpr int RW = VOR_RUN + 2;
            constexpr int NWB = (9 * RW + 63) / 64;
            const int nx = im.vnx, ny = im.vny, nz = im.vnz;
            uint64_t w[NWB + 2];
            w[NWB] = w[NWB + 1] = 0;
#pragma unroll
            for (int k = 0; k < NWB; ++k) {
                const int idx = tid + 64 * k;
                const int rr = idx / RW, cx = idx - rr * RW;
                const int qx = run.xs - 1 + cx, qy = run.ty + rr % 3 - 1, qz = run.tz + rr / 3 - 1;
                bool e = false;
                if (idx < 9 * RW && cx < run.nrun + 2 && qx >= 0 && qx < nx && qy >= 0 && qy < ny && qz >= 0 && qz < nz)
                    e = L.settled[im.tbase + ((int64_t)qz * ny + qy) * nx + qx] == 2;
                w[k] = __ballot(e);
            }
            if (tid == 0) {
#pragma unroll
                for (int rr = 0; rr < 9; ++rr) {
                    const int o = rr * RW, wi = o >> 6, sh = o & 63;
                    uint64_t lo = w[wi] >> sh, hi;
                    if (sh) {
                        lo |= w[wi + 1] << (64 - sh);
                        hi = w[wi + 1] >> sh | w[wi + 2] << (64 - sh);
                    } else {
                        hi = w[wi + 1];
                    }
                    if (RW < 64) lo &= (1ull << (RW < 64 ? RW : 0)) - 1;
                    hi &= RW > 64 ? ((1ull << (RW > 64 ? RW - 64 : 0)) - 1) : 0ull;
                    Em[rr][0] = lo;
                    Em[rr][1] = hi;
                }
            }
        }
        // neighbour rows of this thread's column: its y tile and the z tiles of the
        // clamped planes 0, 2 and 10
        const int64_t ycl = yy >> 3;
        const int dy = (int)(ycl - run.ty);
        const int64_t zc0 = z0 < 0 ? 0 : z0, zc10 = z0 + 10 >= g.D ? g.D - 1 : z0 + 10;
        erow[0] = (int)((zc0 >> 3) - run.tz + 1) * 3 + dy + 1;
        erow[1] = 3 + dy + 1;
        erow[2] = (int)((zc10 >> 3) - run.tz + 1) * 3 + dy + 1;
        __syncthreads();
    }
    const bool hsrc = col && (rx == V3_RX - 4 || rx == V3_RX - 3);
    const bool hdst = col && rx < 2;
    const bool rsrc = col && (rx == 2 || rx == 3);
    const bool rdst = col && rx >= V3_RX - 2;
    const int hidx = ((rx & 1) * V3_RY + ry) * V3_RZ;
    const int64_t xs16 = (int64_t)run.xs * 16;  // image x of the run's first tile
    // region columns of tile t (of the run), but (unless full) the halo columns that
    // come through LDS
    auto load = [&](int t, float* v, bool full) {
        if (!full && ((hdst && active(t - 1)) || (rdst && active(t + 1)))) return;
        const int64_t xr = xs16 + (int64_t)t * 16 - 2 + rx;
        const int64_t xc = xr < 0 ? 0 : (xr >= g.W ? g.W - 1 : xr);
        const float* cp = src + yy * g.W + xc;
        bool em[3] = {false, false, false};
        if (e2) {
            const int c = (int)((xc >> 4) - (run.xs - 1));  // 0 .. nrun + 1
#pragma unroll
            for (int k = 0; k < 3; ++k) em[k] = ((Em[erow[k]][c >> 6] >> (c & 63)) & 1ull) != 0;
        }
#pragma unroll
        for (int rz = 0; rz < V3_RZ; ++rz) {
            const bool e = em[rz < 2 ? 0 : (rz < V3_RZ - 2 ? 1 : 2)];
            v[rz] = !col ? 0.f : (e ? -1.f : cp[zoff[rz]]);
        }
    };
    uint64_t act = run.act;
    auto next_tile = [&]() {
        if (!act) return -1;
        const int t = __builtin_ctzll(act);
        act &= act - 1;
        return t;
    };
    int tk = next_tile();
    int tn = next_tile();
    int tn2 = next_tile();
    float v[V3_RZ], vn[V3_RZ], vnn[V3_RZ];
    load(tk, v, true);  // the first tile's region in full
    if (tn >= 0) load(tn, vn, false);
    // tile planes inside the image: q = 1 .. nzin (region plane q + 1; all 8 but in the
    // last tile plane)
    const int nzin = (int)(g.D - z0 - 2 < 8 ? g.D - z0 - 2 : 8);
    if (tid == 0) zf[0] = 0;
    __syncthreads();
    int par = 0;
    uint32_t done = 0;
    int acc = 0;
#pragma unroll 1
    for (;;) {
        ++done;
        if (tn2 >= 0) load(tn2, vnn, false);  // two tiles ahead
        const int64_t xr = xs16 + (int64_t)tk * 16 - 2 + rx;
        // tile voxel columns inside the image
        const bool tcol = ytile && rx >= 2 && rx < V3_RX - 2 && xr < g.W;
        if (tid == 0) sbits[par] = 0;  // last read two tiles ago, before a barrier
        // C: maximum over z - 1 .. z + 1 for the box planes; zero test of the tile
        float c[V3_NB];
        bool zero = false, hz = false;
#pragma unroll
        for (int q = 0; q < V3_NB; ++q) {
            c[q] = fmaxf(fmaxf(v[q], v[q + 1]), v[q + 2]);
            if (col) Cl[q * V3_COLS + tid] = c[q];
            hz |= v[q + 1] == 0.f;
            if (q >= 1 && q <= 8 && q <= nzin) zero |= v[q + 1] == 0.f;
        }
        if (col) Fl[tid] = hz ? 1 : 0;
        // the tile holds a zero voxel: one flag word per tile parity, set by the waves
        // before the barrier below and cleared for the next tile before this tile's last
        if (__ballot(zero && tcol) != 0 && (tid & 63) == 0) zf[par] = 1;
        const int64_t obase = yr * g.W + xr;  // output column offset
        __syncthreads();
        if (zf[par] == 0) {
            // no zero voxel in the tile: the output is the input
            if (tcol) {
#pragma unroll
                for (int q = 1; q <= 8; ++q)
                    if (q <= nzin) store(obase, q + 1, v[q + 1]);
            }
        } else {
            int bits = 0;
            // Y: maximum of C over y - 1 .. y + 1 (box rows), needed where this column or
            // an x neighbour holds a zero (rows 1 .. 10: the neighbours' flags lie in the
            // region)
            float y[V3_NB];
            if (ybox && (hz || Fl[tid - 1] || Fl[tid + 1])) {
#pragma unroll
                for (int q = 0; q < V3_NB; ++q) {
                    y[q] = fmaxf(fmaxf(Cl[q * V3_COLS + tid - V3_RX], c[q]), Cl[q * V3_COLS + tid + V3_RX]);
                    Yl[q * V3_COLS + tid] = y[q];
                }
            } else {
#pragma unroll
                for (int q = 0; q < V3_NB; ++q) y[q] = c[q];
            }
            __syncthreads();
            // box value of the ring + tile: maximum of Y over x - 1 .. x + 1; zero voxels
            // take max(0, box maximum), the others keep their value
            float bx[V3_NB];
#pragma unroll
            for (int q = 0; q < V3_NB; ++q) bx[q] = v[q + 1];
            if (ybox && xbox) {
                if (hz) {
#pragma unroll
                    for (int q = 0; q < V3_NB; ++q) {
                        const float m = fmaxf(fmaxf(Yl[q * V3_COLS + tid - 1], y[q]), Yl[q * V3_COLS + tid + 1]);
                        const bool z = v[q + 1] == 0.f;
                        bx[q] = z ? fmaxf(0.f, m) : v[q + 1];
                        if (q >= 1 && q <= 8 && q <= nzin) bits |= (z && bx[q] != 0.f) ? 1 : 0;
                    }
                }
#pragma unroll
                for (int q = 0; q < V3_NB; ++q) Cl[q * V3_COLS + tid] = bx[q];
            }
            if (!tcol) bits = 0;
            __syncthreads();
            // diamond on the tile: z neighbours in registers, x / y neighbours from LDS
            if (tcol && hz) {
#pragma unroll
                for (int q = 1; q <= 8; ++q) {
                    const float* p = Cl + q * V3_COLS + tid;
                    const float nb = fmaxf(fmaxf(fmaxf(p[-1], p[1]), fmaxf(p[-V3_RX], p[V3_RX])),
                                           fmaxf(bx[q - 1], bx[q + 1]));
                    const bool z = bx[q] == 0.f;
                    const float out = z ? fmaxf(0.f, nb) : bx[q];
                    if (q <= nzin) {
                        store(obase, q + 1, out);
                        bits |= z ? (out != 0.f ? 2 : 4) : 0;
                    }
                }
            } else if (tcol) {  // no zero voxel in the column: box and diamond keep it
#pragma unroll
                for (int q = 1; q <= 8; ++q)
                    if (q <= nzin) store(obase, q + 1, v[q + 1]);
            }
            if (bits) atomicOr(&sbits[par], bits);
        }
        // the next tile's halo columns from this tile and from the tile after it
        const bool nl = tn >= 0 && tn == tk + 1, nr = tn >= 0 && tn2 == tn + 1;
        if (nl && hsrc) {
#pragma unroll
            for (int rz = 0; rz < V3_RZ; ++rz) Hl[par][hidx + rz] = v[rz];
        }
        if (nr && rsrc) {
#pragma unroll
            for (int rz = 0; rz < V3_RZ; ++rz) Rl[par][hidx + rz] = vnn[rz];
        }
        if (tid == 0) zf[par ^ 1] = 0;  // last read before the previous tile's last barrier
        // every thread is past its LDS reads of this tile before the next tile's C
        // writes; the block's change bits
        __syncthreads();
        const int tb = sbits[par];
        if (tb & 3) {
            acc |= tb & 3;
            if (tid == 0) L.mark[(int64_t)(b & 1) * L.cap + run.t0 + tk] = b;
        } else if (!(tb & 4) && tid == 0) {
            L.settled[run.t0 + tk] = 1;
        }
        if (tn < 0) break;
        // on to the next tile (Hl[par] / Rl[par] are written again two tiles on, after
        // the next tile's barriers)
        if (nl && hdst) {
#pragma unroll
            for (int rz = 0; rz < V3_RZ; ++rz) v[rz] = Hl[par][hidx + rz];
        } else if (nr && rdst) {
#pragma unroll
            for (int rz = 0; rz < V3_RZ; ++rz) v[rz] = Rl[par][hidx + rz];
        } else {
#pragma unroll
            for (int rz = 0; rz < V3_RZ; ++rz) v[rz] = vn[rz];
        }
#pragma unroll
        for (int rz = 0; rz < V3_RZ; ++rz) vn[rz] = vnn[rz];
        tk = tn;
        tn = tn2;
        tn2 = next_tile();
        par ^= 1;
    }
    if (tid == 0 && acc) atomicOr(&cur[run.j], (uint32_t)acc);
    vor_count_tiles(S, done);
}

// Closes the last pair b of a chunk (one block): iteration counts of the images that
// converged in it, and the "an image still grows" word the host reads.
__global__ __launch_bounds__(VB) void k_vor_close(int nimg, int b, VorState S) {
    const uint32_t* last = S.bits + (size_t)(b % 3) * nimg;
    bool grows = false;
    int mx = 0;
    for (int j = threadIdx.x; j < nimg; j += VB) {
        if (S.iters[j] != 0) continue;
        if (last[j] == 3u) {
            grows = true;
        } else {
            const int it = 2 * b + ((last[j] & 1u) ? 2 : 1);
            S.iters[j] = it;
            mx = it > mx ? it : mx;
        }
    }
    grows = __syncthreads_or(grows ? 1 : 0) != 0;
    if (mx) atomicMax((int*)&S.word[1], mx);
    if (threadIdx.x == 0) S.word[0] = grows ? 1u : 0u;
}

constexpr int VOR_CHUNK = 8;  // pairs between host checks (the first check after 2)

// scratch of the growth over `ntiles` tiles of `nimg` images
struct VorScratch {
    Scratch mem;
    VorLists L;
    VorState S;
    static size_t bytes(int64_t ntiles, int nimg) {
        const int64_t cap = ntiles > 0 ? ntiles : 1;
        return (size_t)cap * 8 + 64 + (size_t)VOR_NCTR * 64 + (size_t)nimg * 16 + (size_t)cap + 4;
    }
    // settled_init 1: every tile starts settled (a fused growth init clears the tiles
    // holding a mask voxel); 2: every tile starts settled and empty (the segmentation
    // pass clears the tiles holding a mask voxel; the others' state is never stored)
    int alloc(int64_t ntiles, int nimg, hipStream_t s, uint32_t settled_init = 0) {
        const int64_t cap = ntiles > 0 ? ntiles : 1;
        CLE_TRY(mem.alloc(bytes(ntiles, nimg), s));
        char* p = mem.get<char>();
        L.mark = (int32_t*)p;          // 2 x cap
        p += (size_t)cap * 8;
        S.word = (uint32_t*)p;         // word[0..1], padded to 16 words
        p += 64;
        S.ctr = (uint32_t*)p;          // VOR_NCTR x 16 words
        p += (size_t)VOR_NCTR * 64;
        S.bits = (uint32_t*)p;         // 3 x nimg
        S.iters = (int32_t*)(S.bits + (size_t)3 * nimg);
        p += (size_t)nimg * 16;
        L.settled = (uint8_t*)p;
        L.cap = cap;
        L.empty2 = settled_init == 2 ? 1 : 0;
        // change stamps -2 (no pair); words, counters and per-image state 0 (contiguous
        // from word); nothing settled yet
        CLE_TRY(fill_u32((uint32_t*)L.mark, 0xFFFFFFFEu, (size_t)cap * 2, s));
        CLE_TRY(fill_u32(S.word, 0u, 16 + (size_t)VOR_NCTR * 16 + (size_t)nimg * 4, s));
        CLE_TRY(fill_u32((uint32_t*)L.settled, settled_init * 0x01010101u, ((size_t)cap + 3) / 4, s));
        return CLE_OK;
    }
};

// The runs of a growth: per kind (0: 2D images, 1: 3D images) the device run-prefix
// table over the images (null for a single image) and the number of runs.
struct VorRuns {
    const int64_t* rs[2];
    int64_t blocks[2];
};
inline VorRuns vor_runs_single(const Geom& g) {
    VorRuns R{{nullptr, nullptr}, {0, 0}};
    R.blocks[vor_3d(g) ? 1 : 0] = vor_run_blocks(g, vor_run_length(g, false));
    return R;
}

// Launches pairs [b, b + count) and the close of the last one (no host synchronisation).
static void launch_pairs(const BImg* dimgs, int nimg, float* P0, float* P1, const VorLists& L, const VorState& S,
                         const VorRuns& R, int b, int count, hipStream_t s) {
    for (int k = 0; k < count; ++k, ++b) {
        const float* src = (b & 1) ? P1 : P0;
        float* dst = (b & 1) ? P0 : P1;
        if (R.blocks[0] > 0)
            k_vor_pair_run2<<<(unsigned)R.blocks[0], VB, 0, s>>>(dimgs, R.rs[0], nimg, src, dst, b, L, S);
        if (R.blocks[1] > 0)
            k_vor_pair_run3<<<(unsigned)R.blocks[1], VB, 0, s>>>(dimgs, R.rs[1], nimg, src, dst, b, L, S);
    }
    k_vor_close<<<1, VB, 0, s>>>(nimg, b - 1, S);
}

// Runs pairs from b0 until every image has converged; P0 holds the initial state (as
// does P1); pair b reads P[b & 1] and writes P[(b + 1) & 1].  `max_iters` (host, may
// be null) receives the largest reference iteration count over the images.
// Counters of the pair launches run by grow_pairs (process-wide): launches, tiles
// processed, and their GPU time (events around each chunk of launches).
static std::atomic<unsigned long long> g_pair_launches{0}, g_pair_tiles{0}, g_pair_ns{0};

static int grow_pairs(const BImg* dimgs, int nimg, float* P0, float* P1, const VorLists& L, const VorState& S,
                      const VorRuns& R, int b0, int* max_iters, hipStream_t s) {
    int b = b0;
    uint32_t hw[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // (no events under a fixed scratch region: those calls create nothing)
    const bool timed =
        !scratch_fixed() && hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    double ms_total = 0;
    int launches = 0;
    int rc = CLE_OK;
    for (;;) {
        const int chunk = b == 0 ? 2 : VOR_CHUNK;
        if (timed) (void)hipEventRecord(ev[0], s);
        launch_pairs(dimgs, nimg, P0, P1, L, S, R, b, chunk, s);
        if (timed) (void)hipEventRecord(ev[1], s);
        b += chunk;
        launches += chunk;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hw, S.word, sizeof(hw), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(CLE_E_HIP, "Voronoi growth launch failed");
            break;
        }
        float ms = 0.f;
        if (timed && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) ms_total += ms;
        if (hw[0] == 0) break;
    }
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    CLE_TRY(rc);
    uint32_t hc[VOR_NCTR * 16];
    CLE_HIP(hipMemcpyAsync(hc, S.ctr, sizeof(hc), hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    unsigned long long tiles = 0;
    for (int k = 0; k < VOR_NCTR; ++k) tiles += hc[k * 16];
    g_pair_launches += (unsigned long long)launches;
    g_pair_tiles += tiles;
    g_pair_ns += (unsigned long long)(ms_total * 1e6);
    if (max_iters) *max_iters = (int)hw[1];
    return CLE_OK;
}

// the buffer holding the result of an image that converged with `iters` iterations
__host__ __device__ inline int vor_result_buffer(int iters) { return (((iters - 1) >> 1) + 1) & 1; }

__global__ void k_put_img(BImg* d, BImg im) { *d = im; }

// growth tiling of an image into its table entry (tile counts, run length); fails for
// more than VOR_MAX_TILES_AXIS tiles along an axis
static int vor_set_tiles(BImg& im, bool batch) {
    const Tiles T = make_tiles(im.g);
    if (T.nx > VOR_MAX_TILES_AXIS || T.ny > VOR_MAX_TILES_AXIS || T.nz > VOR_MAX_TILES_AXIS)
        return fail(CLE_E_VALUE, "image too large for the Voronoi growth tiling (%lld x %lld x %lld tiles, at most %lld "
                    "per axis)", (long long)T.nx, (long long)T.ny, (long long)T.nz, (long long)VOR_MAX_TILES_AXIS);
    im.vnx = (int)T.nx;
    im.vrun = vor_run_length(im.g, batch);
    im.vny = (int)T.ny;
    im.vnz = (int)T.nz;
    return CLE_OK;
}

// single-image growth: P0 = P1 = init state from `val`, grown; returns the buffer
// holding the result
template <typename VF>
static int grow_single(VF val, float* P0, float* P1, const Geom& g, int* n_iterations, const float** result,
                       hipStream_t s) {
    const Tiles T = make_tiles(g);
    const int64_t ntiles = T.count();
    BImg im;
    memset(&im, 0, sizeof(im));
    im.g = g;
    CLE_TRY(vor_set_tiles(im, false));
    VorScratch V;
    CLE_TRY(V.alloc(ntiles, 1, s));
    Scratch tab;
    CLE_TRY(tab.alloc(sizeof(BImg), s));
    k_put_img<<<1, 1, 0, s>>>(tab.get<BImg>(), im);
    k_vor_init<<<(unsigned)ntiles, VB, 0, s>>>(val, P0, P1, g, V.L);
    CLE_LAUNCHED();
    int iters = 0;
    CLE_TRY(grow_pairs(tab.get<BImg>(), 1, P0, P1, V.L, V.S, vor_runs_single(g), 0, &iters, s));
    if (n_iterations) *n_iterations = iters;
    *result = vor_result_buffer(iters) ? P1 : P0;
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

// mask-out of 32-bit labels with a uint8 mask: four voxels per thread (A 16-byte, mask
// 4-byte, dst 16-byte aligned), the last n % 4 voxels one by one
__global__ void k_mask_out_u8_32(const float* __restrict__ A, const uint8_t* __restrict__ mask, DBuf dst, int64_t n) {
    const int64_t nv = n / 4;
    const bool u = dst.dt == CLE_U32;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
        const float4 a = ((const float4*)A)[q];
        const uchar4 m = ((const uchar4*)mask)[q];
        const float av[4] = {m.x ? a.x : 0.f, m.y ? a.y : 0.f, m.z ? a.z : 0.f, m.w ? a.w : 0.f};
        uint32_t ov[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            ov[c] = u ? sat_cast<uint32_t>(av[c])
                      : (uint32_t)sat_cast<int32_t>(av[c]);
        ((uint4*)dst.p)[q] = make_uint4(ov[0], ov[1], ov[2], ov[3]);
    }
    for (int64_t i = nv * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst, i, mask[i] ? A[i] : 0.f);
}

// The masked result when every seed lies inside the mask (voronoi_otsu_labeling: the
// seeds are spots AND segmentation): voxels outside the mask hold exactly -1 after the
// growth and mask voxels hold a label >= 0 or 0, so mask(A, mask) == max(A, 0) and the
// mask need not be read.  Four voxels per thread as k_mask_out_u8_32.
__global__ void k_clip_out_32(const float* __restrict__ A, DBuf dst, int64_t n) {
    const int64_t nv = n / 4;
    const bool u = dst.dt == CLE_U32;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
        const float4 a = ((const float4*)A)[q];
        const float av[4] = {fmaxf(a.x, 0.f), fmaxf(a.y, 0.f), fmaxf(a.z, 0.f), fmaxf(a.w, 0.f)};
        uint32_t ov[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) ov[c] = u ? sat_cast<uint32_t>(av[c]) : (uint32_t)sat_cast<int32_t>(av[c]);
        ((uint4*)dst.p)[q] = make_uint4(ov[0], ov[1], ov[2], ov[3]);
    }
    for (int64_t i = nv * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst, i, fmaxf(A[i], 0.f));
}

// seg = blurred > float32(thr); spots <- spots AND seg (in place); four voxels per
// thread (the library's scratch buffers are 256-byte aligned), the tail one by one
__global__ void k_seg_and(const float* blurred, const double* thr, uint8_t* spots, uint8_t* seg, int64_t n) {
    const float t = (float)thr[0];
    const int64_t nv = n / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = ((const float4*)blurred)[q];
        const uchar4 sp = ((const uchar4*)spots)[q];
        const uchar4 g = make_uchar4(v.x > t ? 1 : 0, v.y > t ? 1 : 0, v.z > t ? 1 : 0, v.w > t ? 1 : 0);
        ((uchar4*)seg)[q] = g;
        ((uchar4*)spots)[q] = make_uchar4((sp.x != 0 && g.x) ? 1 : 0, (sp.y != 0 && g.y) ? 1 : 0,
                                          (sp.z != 0 && g.z) ? 1 : 0, (sp.w != 0 && g.w) ? 1 : 0);
    }
    for (int64_t i = nv * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t sg = blurred[i] > t ? 1 : 0;
        seg[i] = sg;
        spots[i] = (spots[i] != 0 && sg) ? 1 : 0;
    }
}

// k_clip_out_32 and k_seg_and over whole rows (of a multiple of 4 voxels) for a 3D
// growth whose empty tiles store no state: the segmentation pass clears the settled
// flag (2: empty) of every growth tile holding a segmentation voxel, the clip writes 0
// for the voxels of the tiles still flagged empty without reading A.  A block takes a
// range of consecutive rows (one division at its start; y and z then step along).
struct RowTiles {
    Geom g;
    int64_t tnx, tny;  // growth tiles (16 x 8 x 8) per row / column
    int64_t rpb;       // rows per block
    __host__ __device__ int64_t rows() const { return g.D * g.H; }
};
// the rows of this block: f(row, first tile of the row's tile row)
template <typename F>
__device__ __forceinline__ void for_block_rows(const RowTiles& rt, F f) {
    const int64_t r0 = (int64_t)blockIdx.x * rt.rpb;
    int64_t r1 = r0 + rt.rpb;
    if (r1 > rt.rows()) r1 = rt.rows();
    if (r0 >= r1) return;
    int64_t z = r0 / rt.g.H, y = r0 - z * rt.g.H;
    for (int64_t row = r0; row < r1; ++row) {
        f(row, ((z >> 3) * rt.tny + (y >> 3)) * rt.tnx);
        if (++y == rt.g.H) {
            y = 0;
            ++z;
        }
    }
}
__global__ __launch_bounds__(256) void k_seg_and_tiles(const float* blurred, const double* thr, uint8_t* spots,
                                                       uint8_t* seg, RowTiles rt, uint8_t* settled) {
    const float t = (float)thr[0];
    const Geom& g = rt.g;
    for_block_rows(rt, [&](int64_t row, int64_t tr) {
        for (int64_t x = 4 * threadIdx.x; x < g.W; x += 4 * 256) {
            const int64_t q = (row * g.W + x) >> 2;
            const float4 v = ((const float4*)blurred)[q];
            const uchar4 sp = ((const uchar4*)spots)[q];
            const uchar4 m = make_uchar4(v.x > t ? 1 : 0, v.y > t ? 1 : 0, v.z > t ? 1 : 0, v.w > t ? 1 : 0);
            ((uchar4*)seg)[q] = m;
            ((uchar4*)spots)[q] = make_uchar4((sp.x != 0 && m.x) ? 1 : 0, (sp.y != 0 && m.y) ? 1 : 0,
                                              (sp.z != 0 && m.z) ? 1 : 0, (sp.w != 0 && m.w) ? 1 : 0);
            // the four lanes of a 16-voxel tile column: one store, when the flag is set
            uint32_t any = (m.x | m.y | m.z | m.w) ? 1u : 0u;
            any |= __shfl_xor(any, 1);
            any |= __shfl_xor(any, 2);
            if (any && (threadIdx.x & 3) == 0 && settled[tr + (x >> 4)] != 0) settled[tr + (x >> 4)] = 0;
        }
    });
}
__global__ __launch_bounds__(256) void k_clip_out_tiles(const float* __restrict__ A, DBuf dst, RowTiles rt,
                                                        const uint8_t* __restrict__ settled) {
    const Geom& g = rt.g;
    const bool u = dst.dt == CLE_U32;
    for_block_rows(rt, [&](int64_t row, int64_t tr) {
        for (int64_t x = 4 * threadIdx.x; x < g.W; x += 4 * 256) {
            const int64_t q = (row * g.W + x) >> 2;
            uint4 o = make_uint4(0u, 0u, 0u, 0u);
            if (settled[tr + (x >> 4)] != 2) {
                const float4 a = ((const float4*)A)[q];
                const float av[4] = {fmaxf(a.x, 0.f), fmaxf(a.y, 0.f), fmaxf(a.z, 0.f), fmaxf(a.w, 0.f)};
                uint32_t ov[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) ov[c] = u ? sat_cast<uint32_t>(av[c]) : (uint32_t)sat_cast<int32_t>(av[c]);
                o = make_uint4(ov[0], ov[1], ov[2], ov[3]);
            }
            ((uint4*)dst.p)[q] = o;
        }
    });
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

// Growth method: the step-synchronous pairs of steps below (default: the faster on the
// measured workloads), or the asynchronous tile solves of growth.hip (same results; a
// growth beyond their step range falls back to the pairs).
static int g_growth_method = 1;  // 0: tile solves, 1: pairs of steps

// labels of `binary` grown inside `mask`, written (masked) to dst.  `B` holds
// scratch for the CCL labels (float32, n voxels).
// Masked growth whose initial state comes straight from the box CCL of the seeds
// (ccl_box_growth_init writes v = (mask - 1) + label into A and B and stamps the
// seed tiles): no separate init pass.  Returns the buffer holding the result.
static int grow_masked_fused(DBuf binary, const uint8_t* mask, float* A, float* B, const Geom& g,
                             int* n_iterations, const float** result, hipStream_t s, VorScratch* pre = nullptr) {
    const Tiles T = make_tiles(g);
    const int64_t ntiles = T.count();
    BImg im;
    memset(&im, 0, sizeof(im));
    im.g = g;
    CLE_TRY(vor_set_tiles(im, false));
    // pre: tiles flagged empty by the segmentation pass (their state is not stored)
    VorScratch own;
    VorScratch& V = pre ? *pre : own;
    if (!pre) CLE_TRY(V.alloc(ntiles, 1, s, 1));  // the CCL write pass clears the tiles with a mask voxel
    Scratch tab;
    CLE_TRY(tab.alloc(sizeof(BImg), s));
    k_put_img<<<1, 1, 0, s>>>(tab.get<BImg>(), im);
    CLE_LAUNCHED();
    GrowInitOut o{A, B, mask, V.L.mark + V.L.cap, V.L.settled, T.tx, T.ty, T.tz, T.nx, T.ny};
    o.skip_empty = pre != nullptr;
    CLE_TRY(ccl_box_growth_init(binary, g, o, s));
    int iters = 0;
    CLE_TRY(grow_pairs(tab.get<BImg>(), 1, A, B, V.L, V.S, vor_runs_single(g), 0, &iters, s));
    if (n_iterations) *n_iterations = iters;
    *result = vor_result_buffer(iters) ? B : A;
    return CLE_OK;
}

// masked_voronoi_labeling (_tier5/_masked_voronoi_labeling.py:44-74): label the seeds (box
// CCL), grow inside the mask, mask the result into dst.  B: a float32 work buffer of
// the image's size.
static int masked_voronoi_core(DBuf binary, DBuf mask, DBuf dst, const Geom& g, float* B, int* n_iterations,
                               hipStream_t s, bool seeds_in_mask = false, VorScratch* pre = nullptr,
                               const RowTiles* rt = nullptr) {
    const int64_t n = g.n();
    bool labelled = false;  // B holds the CCL labels
    // (the tile solves' stores are not part of the fixed-region bound: pairs there)
    if (g_growth_method == 0 && !scratch_fixed()) {
        CLE_TRY(ccl_impl(binary, g, DBuf{B, CLE_F32}, CLE_CONN_BOX, CLE_FP32_KEY_COMPAT, nullptr, s));
        labelled = true;
        bool over = false;
        CLE_TRY(grow_masked(mask, B, dst, g, n_iterations, &over, s));
        if (!over) return CLE_OK;
    }
    Scratch a;
    CLE_TRY(a.alloc((size_t)n * 4, s));
    float* A = a.get<float>();
    const float* res = nullptr;
    const bool u8mask = mask.dt == CLE_U8 && ((uintptr_t)mask.p & 3) == 0 && ((uintptr_t)A & 15) == 0 &&
                        ((uintptr_t)B & 15) == 0;
    bool empty_tiles = false;  // the growth stored no state for the tiles flagged empty
    if (u8mask && !labelled) {
        CLE_TRY(grow_masked_fused(binary, (const uint8_t*)mask.p, A, B, g, n_iterations, &res, s, pre));
        empty_tiles = pre != nullptr;
    } else {
        if (!labelled) CLE_TRY(ccl_impl(binary, g, DBuf{B, CLE_F32}, CLE_CONN_BOX, CLE_FP32_KEY_COMPAT, nullptr, s));
        // a uint8 mask and 16-byte aligned buffers take the four-voxel init
        if (u8mask)
            CLE_TRY(grow_single(MaskedU8{(const uint8_t*)mask.p, B}, A, B, g, n_iterations, &res, s));
        else
            CLE_TRY(grow_single(MaskedAny{mask, B}, A, B, g, n_iterations, &res, s));
    }
    if (empty_tiles)
        k_clip_out_tiles<<<(unsigned)((rt->rows() + rt->rpb - 1) / rt->rpb), 256, 0, s>>>(res, dst, *rt,
                                                                                          pre->L.settled);
    else if (seeds_in_mask && (dst.dt == CLE_U32 || dst.dt == CLE_I32) && ((uintptr_t)res & 15) == 0 &&
        ((uintptr_t)dst.p & 15) == 0)
        k_clip_out_32<<<vg(n / 4 + 1), 256, 0, s>>>(res, dst, n);
    else if (mask.dt == CLE_U8 && (dst.dt == CLE_U32 || dst.dt == CLE_I32) && ((uintptr_t)res & 15) == 0 &&
        ((uintptr_t)mask.p & 3) == 0 && ((uintptr_t)dst.p & 15) == 0)
        k_mask_out_u8_32<<<vg(n / 4 + 1), 256, 0, s>>>(res, (const uint8_t*)mask.p, dst, n);
    else
        k_mask_out<<<vg(n), 256, 0, s>>>(res, mask, dst, n);
    CLE_LAUNCHED();
    return CLE_OK;
}

// Scratch bound of voronoi_otsu_single: the blurred image, spots, segmentation and
// threshold words live through the call; on top of them, at most one of the blur's
// float32 temporary (its per-axis fallback), the Otsu histograms, and the growth's
// second state buffer with its tile lists together with the box CCL of the spots.
size_t voronoi_otsu_scratch_bytes(const Geom& g) {
    const size_t n = (size_t)g.n();
    const size_t held = scratch_round(n * 4) + 2 * scratch_round(n) + scratch_round(64);
    const Tiles T = make_tiles(g);
    // the growth buffers and tile lists are alive while the CCL of the seeds runs
    const size_t grow = scratch_round(n * 4) + scratch_round(VorScratch::bytes(T.count(), 1)) +
                        scratch_round(sizeof(BImg)) + ccl_box_scratch_bytes(g, CLE_FP32_KEY_COMPAT);
    size_t stage = scratch_round(n * 4);
    const size_t others[2] = {otsu_scratch_bytes(g), grow};
    for (size_t v : others) stage = v > stage ? v : stage;
    return held + stage;
}

int voronoi_otsu_single(const cle_buf* src, const cle_buf* dst, float spot_sigma, float outline_sigma,
                        hipStream_t s) {
    Geom g;
    CLE_TRY(check_buf(src, "source", &g));
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    Scratch blurred, spots, seg, aux;
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
    // (the isotropic blur folds its output's minimum and maximum into mmw on the way)
    const float s2[3] = {outline_sigma, outline_sigma, outline_sigma};
    uint32_t* mmw = (uint32_t*)(aux.get<char>() + 32);
    minmax_words_init(mmw, s);
    int folded = 0;
    CLE_TRY(blur_impl_minmax(src, &bb, s2, mmw, &folded, s));
    if (folded) {
        minmax_words_finish(mmw, mm, s);
        CLE_TRY(otsu_after_minmax(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    } else {
        CLE_TRY(otsu_impl(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    }
    // 3D images with rows of a multiple of 4 voxels, grown by pairs of steps into 32-bit
    // labels: the segmentation pass flags the growth tiles without a mask voxel, whose
    // state (-1 everywhere) is then neither stored by the growth init nor read by the
    // growth, and whose labels the final clip writes as 0 without a read
    const bool empty_tiles = vor_3d(g) && (g.W & 3) == 0 && (g_growth_method != 0 || scratch_fixed()) &&
                             (dst->dtype == CLE_U32 || dst->dtype == CLE_I32) && ((uintptr_t)dst->data & 15) == 0;
    VorScratch V;
    RowTiles rt{g, 0, 0, 1};
    if (empty_tiles) {
        const Tiles T = make_tiles(g);
        rt.tnx = T.nx;
        rt.tny = T.ny;
        // about 4096 voxels per block and at most 131072 blocks
        rt.rpb = (4096 + g.W - 1) / g.W;
        if ((rt.rows() + rt.rpb - 1) / rt.rpb > 131072) rt.rpb = (rt.rows() + 131071) / 131072;
        CLE_TRY(V.alloc(T.count(), 1, s, 2));
        k_seg_and_tiles<<<(unsigned)((rt.rows() + rt.rpb - 1) / rt.rpb), 256, 0, s>>>(
            blurred.get<float>(), thr, spots.get<uint8_t>(), seg.get<uint8_t>(), rt, V.L.settled);
    } else {
        k_seg_and<<<vg(n / 4 + 1), 256, 0, s>>>(blurred.get<float>(), thr, spots.get<uint8_t>(), seg.get<uint8_t>(),
                                                n);
    }
    CLE_LAUNCHED();
    // masked Voronoi of the seeds inside the segmentation; the seed labels go to the
    // blurred buffer
    return masked_voronoi_core(DBuf{spots.p, CLE_U8}, DBuf{seg.p, CLE_U8}, dbuf(dst), g, blurred.get<float>(), nullptr,
                               s, true, empty_tiles ? &V : nullptr, &rt);
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

// one block per maxima tile (64 x 16 x 16 voxels) of one image (ms = tile-prefix table)
__global__ __launch_bounds__(256) void k_maxima_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ ms,
                                                     int nimg, const float* F, uint8_t* S) {
    __shared__ float ring[MX_LDS];
    const int64_t b = blockIdx.x;
    const int j = find_seg(ms, nimg, b);
    const BImg& im = imgs[j];
    uint8_t* o = S + im.off;
    maxima_tile(F32In{F + im.off}, im.g, b - ms[j], ring, [&](int64_t i, bool m) { o[i] = m ? 1 : 0; });
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
    // four voxels per thread: an image starts at a multiple of 64 voxels in the batch
    // buffers (16-byte aligned float4 / 4-byte uchar4) and its slot is padded to 64, so
    // the last group may run past n inside the padding (written as 0)
    const float* f = F + im.off;
    uint8_t* sp = S + im.off;
    uint8_t* gp = G + im.off;
    for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 4 * 256) {
        const float4 v = *(const float4*)(f + i);
        const uchar4 sv = *(const uchar4*)(sp + i);
        const float vv[4] = {v.x, v.y, v.z, v.w};
        const uint8_t so[4] = {sv.x, sv.y, sv.z, sv.w};
        uint8_t go[4], sn[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            go[c] = (i + c < n && vv[c] > t) ? 1 : 0;
            sn[c] = (so[c] != 0 && go[c]) ? 1 : 0;
        }
        *(uchar4*)(gp + i) = make_uchar4(go[0], go[1], go[2], go[3]);
        *(uchar4*)(sp + i) = make_uchar4(sn[0], sn[1], sn[2], sn[3]);
    }
}

// A = B = (G - 1) + labels; tiles holding a positive value queue their neighbourhood
// for pair 0
__global__ __launch_bounds__(VB) void k_vor_init_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ ts,
                                                      int nimg, const uint8_t* G, float* A, float* B, VorLists L) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(ts, nimg, b);
    const BImg& im = imgs[j];
    const Tiles T = make_tiles(im.g);
    // a tile holding a positive value counts as changed before pair 0
    if (vor_init_tile(MaskedU8{G + im.off, B + im.off}, A + im.off, B + im.off, im.g, T, b - ts[j], L.settled + b) &&
        threadIdx.x == 0)
        L.mark[L.cap + b] = -1;
}

// dst = G ? result : 0, the result in the growth buffer of the image's last pair
__global__ __launch_bounds__(256) void k_mask_out_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ vs,
                                                       int nimg, const float* P0, const float* P1,
                                                       const int32_t* __restrict__ iters, const uint8_t* G) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(vs, nimg, b);
    const BImg& im = imgs[j];
    const int64_t n = im.g.n(), i0 = (b - vs[j]) * BV, i1 = i0 + BV < n ? i0 + BV : n;
    const DBuf d{im.dst, im.dst_dt};
    const float* A = (vor_result_buffer(iters[j]) ? P1 : P0) + im.off;
    (void)G;  // the seeds lie inside the segmentation: mask(A, G) == max(A, 0) (see k_clip_out_32)
    if ((im.dst_dt == CLE_U32 || im.dst_dt == CLE_I32) && ((uintptr_t)im.dst & 15) == 0) {
        // 32-bit labels into a 16-byte aligned destination: four voxels per thread (the
        // batch buffers are aligned as in k_seg_and_batch), the image's last partial
        // group element by element
        uint32_t* o = (uint32_t*)im.dst;
        const bool u = im.dst_dt == CLE_U32;
        for (int64_t i = i0 + 4 * threadIdx.x; i < i1; i += 4 * 256) {
            const float4 a = *(const float4*)(A + i);
            const float av[4] = {fmaxf(a.x, 0.f), fmaxf(a.y, 0.f), fmaxf(a.z, 0.f), fmaxf(a.w, 0.f)};
            if (i + 4 <= n) {
                uint32_t ov[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    ov[c] = u ? sat_cast<uint32_t>(av[c])
                              : (uint32_t)sat_cast<int32_t>(av[c]);
                *(uint4*)(o + i) = make_uint4(ov[0], ov[1], ov[2], ov[3]);
            } else {
                for (int c = 0; c < 4 && i + c < n; ++c) st(d, i + c, av[c]);
            }
        }
        return;
    }
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) st(d, i, fmaxf(A[i], 0.f));
}

// Pinned host staging of the batch tables.  Buffers are taken from a process-wide
// pool and given back when the call ends (after its stream has drained, also when a
// later step fails: the copy out of the buffer may still be in flight); they are never
// freed, so a call allocates pinned memory only when no free buffer is large enough.
struct PinnedPool {
    std::mutex m;
    std::vector<std::pair<void*, size_t>> free;
};
static PinnedPool& pinned_pool() {
    static PinnedPool* pool = new PinnedPool();  // outlives every caller (never destroyed)
    return *pool;
}
struct PinnedStage {
    void* p = nullptr;
    size_t cap = 0;
    hipStream_t s = nullptr;
    int take(size_t bytes, hipStream_t stream) {
        s = stream;
        PinnedPool& pool = pinned_pool();
        {
            std::lock_guard<std::mutex> lk(pool.m);
            for (size_t k = 0; k < pool.free.size(); ++k)
                if (pool.free[k].second >= bytes) {
                    p = pool.free[k].first;
                    cap = pool.free[k].second;
                    pool.free.erase(pool.free.begin() + (long)k);
                    return CLE_OK;
                }
        }
        cap = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
        CLE_HIP(hipHostMalloc(&p, cap));
        return CLE_OK;
    }
    // hands the buffer to a captured graph, which reads it on every replay
    void* detach() {
        void* q = p;
        p = nullptr;
        return q;
    }
    ~PinnedStage() {
        if (p) {
            (void)hipStreamSynchronize(s);
            PinnedPool& pool = pinned_pool();
            std::lock_guard<std::mutex> lk(pool.m);
            pool.free.emplace_back(p, cap);
        }
    }
};

// Capture of the batched pipeline into a graph: the growth runs a fixed number of
// pairs and copies its "still growing" word to pinned host memory instead of checking
// it; nothing synchronises.  The batch fills in what a replay needs to finish the
// growth eagerly when the fixed pairs were not enough.
struct BatchCapture {
    int graph_pairs = 0;
    uint32_t* word_host = nullptr;  // pinned uint32[2], written by the graph
    // filled in by the capture run
    void* pinned = nullptr;         // table staging the graph copies from
    const BImg* dimgs = nullptr;
    const int64_t* dvs = nullptr;
    int n = 0;
    VorRuns runs{};  // runs of the growth launches
    int64_t ntiles = 0, vblocks = 0;
    float* Ap = nullptr;
    float* Fp = nullptr;
    const uint8_t* Gp = nullptr;
    VorLists L{};
    VorState S{};
};

static int voronoi_otsu_batch(const cle_buf* srcs, const cle_buf* dsts, const float* ss, const float* so, int n,
                              int* n_steps, hipStream_t s, BatchCapture* cap = nullptr) {
    if (n > VOR_MAX_IMAGES) return fail(CLE_E_VALUE, "at most %d images per batch (got %d)", VOR_MAX_IMAGES, n);
    std::vector<BImg> imgs((size_t)n);
    int64_t total = 0, ntiles = 0, wtiles = 0;
    // the growth: tile solves, or pairs of steps (a captured graph runs a fixed number
    // of pairs without host round trips; also the fallback)
    const bool pairs = cap != nullptr || g_growth_method != 0;
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
        CLE_TRY(vor_set_tiles(im, true));
        ntiles += g.n() > 0 ? make_tiles(g).count() : 0;
        grow_tiling(im);
        if (im.wnx > 65535 || im.wny > 65535 || im.wnz > 65535)
            return fail(CLE_E_VALUE, "image %d too large for the growth tiling", j);
        im.wbase = wtiles;
        wtiles += grow_tile_count(im);
    }
    if (total == 0) return CLE_OK;
    VorRuns runs{};

    CclBatchSizes csz;
    ccl_batch_layout(imgs, &csz);
    // packed work images: F (float32: blurs, CCL labels, growth buffer B), A (growth
    // buffer A), S (uint8 seeds), G (uint8 segmentation)
    Scratch F, A, S, G, aux;
    CLE_TRY(F.alloc((size_t)total * 4, s));
    if (pairs) CLE_TRY(A.alloc((size_t)total * 4, s));
    CLE_TRY(S.alloc((size_t)total, s));
    CLE_TRY(G.alloc((size_t)total, s));
    const size_t otsu_words = (size_t)n * (2 + 256 + 256);
    CLE_TRY(aux.alloc((size_t)n * 8 + (size_t)n * 8 + otsu_words * 4, s));
    float* mm = aux.get<float>();
    double* thr = (double*)(aux.get<char>() + (size_t)n * 8);
    uint32_t* ou = (uint32_t*)(aux.get<char>() + (size_t)n * 16);
    VorScratch V;
    if (pairs) CLE_TRY(V.alloc(ntiles, n, s, 1));  // the CCL write pass clears the tiles with a mask voxel
    const VorLists VL = V.L;
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
    std::vector<int64_t> mst((size_t)n + 1);
    int64_t nmt = 0;
    for (int j = 0; j < n; ++j) {
        mst[j] = nmt;
        nmt += imgs[j].g.n() > 0 ? maxima_tiles(imgs[j].g) : 0;
    }
    mst[n] = nmt;
    const size_t ms_off = blob.put_vec(mst);
    std::vector<int64_t> wst((size_t)n + 1);
    for (int j = 0; j < n; ++j) wst[j] = imgs[j].wbase;
    wst[n] = wtiles;
    const size_t ws_off = blob.put_vec(wst);
    // run prefixes of the growth launches: 2D images, 3D images
    std::vector<int64_t> rst[2] = {std::vector<int64_t>((size_t)n + 1), std::vector<int64_t>((size_t)n + 1)};
    int64_t nrun[2] = {0, 0};
    for (int j = 0; j < n; ++j) {
        const int kind = vor_3d(imgs[j].g) ? 1 : 0;
        rst[0][j] = nrun[0];
        rst[1][j] = nrun[1];
        nrun[kind] += vor_run_blocks(imgs[j].g, imgs[j].vrun);
    }
    rst[0][n] = nrun[0];
    rst[1][n] = nrun[1];
    const size_t rs_off[2] = {blob.put_vec(rst[0]), blob.put_vec(rst[1])};

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
        k_maxima_batch<<<(unsigned)nmt, 256, 0, s>>>((const BImg*)(dev + imgs_off), (const int64_t*)(dev + ms_off),
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
    // seed labels; with the pairs of steps the CCL's write pass stores the growth's
    // initial state A = F = (segmentation - 1) + label and stamps the seed tiles, and the
    // tiles without a mask voxel are settled (no separate init pass)
    if (pairs) {
        const GrowInitBatch gi{Ap, Gp, VL.mark + VL.cap, VL.settled};
        CLE_TRY(plan_ccl_batch(imgs, imgs_off, csz, Sp, Fp, blob, L, keep, s, &gi));
    } else {
        CLE_TRY(plan_ccl_batch(imgs, imgs_off, csz, Sp, Fp, blob, L, keep, s));
    }

    // the tables go to the device with one copy
    Scratch dblob;
    CLE_TRY(dblob.alloc(blob.h.size(), s));
    PinnedStage pin;
    CLE_TRY(pin.take(blob.h.size(), s));
    memcpy(pin.p, blob.h.data(), blob.h.size());
    CLE_HIP(hipMemcpyAsync(dblob.p, pin.p, blob.h.size(), hipMemcpyHostToDevice, s));
    if (cap) cap->pinned = pin.detach();
    const char* dev = dblob.get<char>();
    for (auto& l : L) CLE_TRY(l(dev));

    // growth: all images together, each stopping on its own convergence
    const BImg* dimgs = (const BImg*)(dev + imgs_off);
    for (int k = 0; k < 2; ++k) {
        runs.rs[k] = (const int64_t*)(dev + rs_off[k]);
        runs.blocks[k] = nrun[k];
    }
    const int64_t* dvs = (const int64_t*)(dev + vs_off);
    if (cap) {
        cap->dimgs = dimgs;
        cap->dvs = dvs;
        cap->n = n;
        cap->runs = runs;
        cap->ntiles = ntiles;
        cap->vblocks = vblocks;
        cap->Ap = Ap;
        cap->Fp = Fp;
        cap->Gp = Gp;
        cap->L = V.L;
        cap->S = V.S;
        if (ntiles > 0) {
            launch_pairs(dimgs, n, Ap, Fp, V.L, V.S, runs, 0, cap->graph_pairs, s);
            CLE_LAUNCHED();
            CLE_HIP(hipMemcpyAsync(cap->word_host, V.S.word, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        }
    } else if (pairs) {
        int steps = 0;
        if (ntiles > 0) CLE_TRY(grow_pairs(dimgs, n, Ap, Fp, V.L, V.S, runs, 0, &steps, s));
        if (n_steps) *n_steps = steps;
    } else {
        // tile solves: the labels go straight to the destinations
        bool over = false;
        std::vector<int> its((size_t)n, 1);
        int kinds = 0;
        for (const BImg& im : imgs) kinds |= im.g.n() == 0 ? 0 : ((im.g.ndim == 3 && im.g.D > 1) ? 2 : 1);
        CLE_TRY(grow_batch(dimgs, (const int64_t*)(dev + ws_off), n, wtiles, kinds, Gp, Fp, its.data(), &over, s));
        if (!over) {
            if (n_steps) *n_steps = *std::max_element(its.begin(), its.end());
            CLE_HIP(hipStreamSynchronize(s));
            return CLE_OK;
        }
        // a growth beyond the tile solves' step range: pairs of steps
        CLE_TRY(A.alloc((size_t)total * 4, s));
        CLE_TRY(V.alloc(ntiles, n, s));
        Ap = A.get<float>();
        k_vor_init_batch<<<(unsigned)ntiles, VB, 0, s>>>(dimgs, (const int64_t*)(dev + ts_off), n, Gp, Ap, Fp, V.L);
        CLE_LAUNCHED();
        int steps = 0;
        CLE_TRY(grow_pairs(dimgs, n, Ap, Fp, V.L, V.S, runs, 0, &steps, s));
        if (n_steps) *n_steps = steps;
    }
    k_mask_out_batch<<<(unsigned)vblocks, 256, 0, s>>>(dimgs, dvs, n, Ap, Fp, V.S.iters, Gp);
    CLE_LAUNCHED();
    // the pinned tables and the scratch are released once the work has run (a
    // capture has recorded the work and keeps both)
    if (!cap) CLE_HIP(hipStreamSynchronize(s));
    return CLE_OK;
}

}  // namespace cle

// A replayable plan of one batch (declared opaque in the C ABI).
struct cle_plan {
    int dev = 0;
    hipStream_t ps = nullptr;  // private stream: its arena holds the plan's device memory
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* word_host = nullptr;
    cle::BatchCapture cap;
};

using namespace cle;

static void plan_free(cle_plan* P) {
    if (!P) return;
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->ps) {
        (void)arena_release_stream(P->ps);
        (void)hipStreamDestroy(P->ps);
    }
    if (P->word_host) (void)hipHostFree(P->word_host);
    if (P->cap.pinned) (void)hipHostFree(P->cap.pinned);
    delete P;
}

extern "C" {

int cle_voronoi_otsu_labeling_batch_plan(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                         const float* outline_sigmas, int n_images, int graph_pairs,
                                         cle_plan** out_plan, void* stream) {
    if (!out_plan) return fail(CLE_E_VALUE, "null plan pointer");
    *out_plan = nullptr;
    if (n_images < 1) return fail(CLE_E_VALUE, "a plan needs at least one image");
    if (graph_pairs < 1) return fail(CLE_E_VALUE, "graph_pairs must be >= 1");
    if (!srcs || !dsts || !spot_sigmas || !outline_sigmas) return fail(CLE_E_VALUE, "null batch argument");
    cle_plan* P = new cle_plan;
    int rc = CLE_OK;
    hipEvent_t ready = nullptr;
    do {
        if (hipGetDevice(&P->dev) != hipSuccess || hipStreamCreateWithFlags(&P->ps, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc((void**)&P->word_host, 4 * sizeof(uint32_t)) != hipSuccess) {
            rc = fail(CLE_E_HIP, "plan: stream / pinned allocation failed");
            break;
        }
        // work already queued on the caller's stream (the sources) comes first
        if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ready, (hipStream_t)stream) != hipSuccess || hipStreamWaitEvent(P->ps, ready, 0) != hipSuccess) {
            rc = fail(CLE_E_HIP, "plan: stream ordering failed");
            break;
        }
        // two eager runs size the private arena (the second merges its chunks), so the
        // capture allocates nothing from the driver; their results are the batch's
        for (int k = 0; k < 2 && rc == CLE_OK; ++k)
            rc = voronoi_otsu_batch(srcs, dsts, spot_sigmas, outline_sigmas, n_images, nullptr, P->ps);
        if (rc != CLE_OK) break;
        if ((rc = arena_set_owned(P->ps)) != CLE_OK) break;
        P->cap.graph_pairs = graph_pairs;
        P->cap.word_host = P->word_host;
        if (hipStreamBeginCapture(P->ps, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = fail(CLE_E_HIP, "plan: stream capture could not start");
            break;
        }
        rc = voronoi_otsu_batch(srcs, dsts, spot_sigmas, outline_sigmas, n_images, nullptr, P->ps, &P->cap);
        const hipError_t e = hipStreamEndCapture(P->ps, &P->graph);
        if (rc != CLE_OK) break;
        if (e != hipSuccess || !P->graph) {
            rc = fail(CLE_E_HIP, "plan: stream capture failed (%s)", hipGetErrorString(e));
            break;
        }
        if (hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0) != hipSuccess) {
            rc = fail(CLE_E_HIP, "plan: graph instantiation failed");
            break;
        }
    } while (false);
    if (ready) (void)hipEventDestroy(ready);
    if (rc != CLE_OK) {
        (void)hipGetLastError();
        plan_free(P);
        return rc;
    }
    *out_plan = P;
    return CLE_OK;
}

int cle_plan_run(cle_plan* P, int* n_steps, void* stream) {
    if (!P || !P->exec) return fail(CLE_E_VALUE, "invalid plan");
    int dev = 0;
    CLE_HIP(hipGetDevice(&dev));
    if (dev != P->dev) return fail(CLE_E_VALUE, "plan was built on device %d, current device is %d", P->dev, dev);
    hipStream_t s = (hipStream_t)stream;
    CLE_HIP(hipGraphLaunch(P->exec, s));
    CLE_HIP(hipStreamSynchronize(s));
    int steps = (int)P->word_host[1];
    const BatchCapture& c = P->cap;
    if (c.ntiles > 0 && P->word_host[0]) {
        // still growing after the recorded pairs: finish eagerly, then mask again
        CLE_TRY(grow_pairs(c.dimgs, c.n, c.Ap, c.Fp, c.L, c.S, c.runs, c.graph_pairs, &steps, s));
        k_mask_out_batch<<<(unsigned)c.vblocks, 256, 0, s>>>(c.dimgs, c.dvs, c.n, c.Ap, c.Fp, c.S.iters, c.Gp);
        CLE_LAUNCHED();
        CLE_HIP(hipStreamSynchronize(s));
    }
    if (n_steps) *n_steps = steps;
    return CLE_OK;
}

int cle_plan_destroy(cle_plan* P) {
    plan_free(P);
    return CLE_OK;
}

int cle_growth_counters(unsigned long long* out, int reset) {
    if (out) {
        out[0] = g_pair_launches.load();
        out[1] = g_pair_tiles.load();
        out[2] = g_pair_ns.load();
    }
    if (reset) {
        g_pair_launches = 0;
        g_pair_tiles = 0;
        g_pair_ns = 0;
    }
    return CLE_OK;
}

int cle_set_growth_method(int method) {
    if (method != 0 && method != 1) return fail(CLE_E_VALUE, "growth method must be 0 or 1 (got %d)", method);
    g_growth_method = method;
    return CLE_OK;
}

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
    Scratch b;
    CLE_TRY(b.alloc((size_t)n * 4, s));
    return masked_voronoi_core(dbuf(binary), dbuf(mask), dbuf(dst), g, b.get<float>(), n_iterations, s);
}

int cle_extend_labeling_via_voronoi(const cle_buf* src, const cle_buf* dst, int* n_iterations, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "labeling_source", &g));
    CLE_TRY(check_buf(dst, "labeling_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "labeling_source", "labeling_destination"));
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = g.n();
    if (!n) return CLE_OK;
    if (g_growth_method == 0) {
        bool over = false;
        CLE_TRY(grow_labels(dbuf(src), dbuf(dst), g, n_iterations, &over, s));
        if (!over) return CLE_OK;
    }
    Scratch a, b;
    CLE_TRY(a.alloc((size_t)n * 4, s));
    CLE_TRY(b.alloc((size_t)n * 4, s));
    const float* res = nullptr;
    CLE_TRY(grow_single(LabelsAny{dbuf(src)}, a.get<float>(), b.get<float>(), g, n_iterations, &res, s));
    k_store_out<<<vg(n), 256, 0, s>>>(res, dbuf(dst), n);
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_voronoi_otsu_labeling(const cle_buf* src, const cle_buf* dst, float spot_sigma, float outline_sigma,
                              void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "label_image_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "label_image_destination"));
    return voronoi_otsu_single(src, dst, spot_sigma, outline_sigma, (hipStream_t)stream);
}

int cle_voronoi_otsu_labeling_batch(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                    const float* outline_sigmas, int n_images, int* n_steps, void* stream) {
    if (n_images < 0) return fail(CLE_E_VALUE, "n_images must be >= 0");
    if (n_images == 0) return CLE_OK;
    if (!srcs || !dsts || !spot_sigmas || !outline_sigmas) return fail(CLE_E_VALUE, "null batch argument");
    return voronoi_otsu_batch(srcs, dsts, spot_sigmas, outline_sigmas, n_images, n_steps, (hipStream_t)stream);
}

int cle_autotune_voronoi_otsu_labeling_batch(const cle_buf* srcs, const cle_buf* dsts, const float* spot_sigmas,
                                             const float* outline_sigmas, int n_images, void* stream) {
    if (n_images < 0) return fail(CLE_E_VALUE, "n_images must be >= 0");
    if (n_images == 0) return CLE_OK;
    if (!srcs || !dsts || !spot_sigmas || !outline_sigmas) return fail(CLE_E_VALUE, "null batch argument");
    blur_tuning(true);
    const int rc = voronoi_otsu_batch(srcs, dsts, spot_sigmas, outline_sigmas, n_images, nullptr, (hipStream_t)stream);
    blur_tuning(false);
    return rc;
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
    uint32_t* mmw = (uint32_t*)(aux.get<char>() + 32);
    minmax_words_init(mmw, s);
    int folded = 0;
    CLE_TRY(blur_impl_minmax(src, &bb, sg, mmw, &folded, s));
    if (folded) {
        minmax_words_finish(mmw, mm, s);
        CLE_TRY(otsu_after_minmax(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    } else {
        CLE_TRY(otsu_impl(DBuf{blurred.p, CLE_F32}, g, mm, thr, s));
    }
    k_threshold_u8<<<vg(n), 256, 0, s>>>(blurred.get<float>(), thr, seg.get<uint8_t>(), n);
    CLE_LAUNCHED();
    return ccl_impl(DBuf{seg.p, CLE_U8}, g, dbuf(dst), CLE_CONN_BOX, CLE_FP32_KEY_COMPAT, nullptr, s);
}

}  // extern "C"
