// Shared definitions of the gfx950 kernels and the C-ABI layer: image geometry,
// typed loads/stores with the reference's saturating conversion, error reporting,
// and the stream-ordered scratch pool.
#pragma once

#include <limits>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cle_abi.h"

namespace cle {

// ----------------------------------------------------------------- geometry
// An image normalised to three dimensions: D slices of H rows of W voxels.
struct Geom {
    int64_t D, H, W;
    int ndim;
    __host__ __device__ int64_t n() const { return D * H * W; }
};

// Device view of a typed buffer: base pointer + element type code.
struct DBuf {
    void* p;
    int dt;
};

// ----------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define CLE_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::cle::fail(CLE_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                               __FILE__, __LINE__);                                    \
    } while (0)

#define CLE_TRY(expr)              \
    do {                           \
        int _rc = (expr);          \
        if (_rc != CLE_OK) return _rc; \
    } while (0)

// launch check: surfaces launch-configuration errors at the call site
#define CLE_LAUNCHED() CLE_HIP(hipGetLastError())

// ----------------------------------------------------------------- validation
int check_buf(const cle_buf* b, const char* name, Geom* g);
int check_same_shape(const Geom& a, const Geom& b, const char* na, const char* nb);
bool dtype_supported(int dt);
size_t dtype_size(int dt);
inline DBuf dbuf(const cle_buf* b) { return DBuf{b->data, b->dtype}; }

// ----------------------------------------------------------------- scratch arenas
// Stream-ordered device scratch from a grow-only arena per (device, stream); memory
// released by a call is reused by the next call on the same stream.
void* scratch_alloc(size_t bytes, hipStream_t s);
void scratch_free(void* p, hipStream_t s);
// arena of a private stream handed to a captured graph / freed with it
int arena_set_owned(hipStream_t s);
int arena_release_stream(hipStream_t s);

// Fixed scratch region: while a FixedScratch lives, the scratch allocations of the
// calling thread are bumps in the caller's device region [base, base + cap) (released
// last-in first-out) instead of arena chunks, so no driver allocation runs; a request
// beyond the region fails with CLE_E_OOM.
struct FixedScratch {
    FixedScratch(char* base, size_t cap);
    ~FixedScratch();
    FixedScratch(const FixedScratch&) = delete;
    FixedScratch& operator=(const FixedScratch&) = delete;
};
bool scratch_fixed();  // a FixedScratch is active on this thread
// bytes one scratch allocation of `bytes` takes (every allocation is 256-byte aligned)
inline size_t scratch_round(size_t bytes) { return ((bytes ? bytes : 4) + 255) / 256 * 256; }

// RAII holder for one scratch allocation on a stream
struct Scratch {
    void* p = nullptr;
    hipStream_t s = nullptr;
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch() { if (p) scratch_free(p, s); }
    template <typename T> T* get() const { return static_cast<T*>(p); }
    int alloc(size_t bytes, hipStream_t st) {
        s = st;
        p = scratch_alloc(bytes ? bytes : 4, st);
        if (!p)
            return scratch_fixed() ? fail(CLE_E_OOM, "the fixed scratch region cannot hold %zu more bytes", bytes)
                                   : fail(CLE_E_OOM, "scratch allocation of %zu bytes failed", bytes);
        return CLE_OK;
    }
};

// fill kernels (no hipMemsetAsync: callers may capture these streams in graphs)
int fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
int fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t s);
int fill_f32(float* p, float v, size_t n, hipStream_t s);
int fill_f64(double* p, double v, size_t n, hipStream_t s);

// Consecutive tiles (which share halos and cache lines) go to the same XCD: the
// hardware hands block b to XCD b % 8, each XCD has its own L2, so block b works on
// logical tile (b % 8) * per_xcd + b / 8.  A bijection on [0, nb).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t xcd = b % 8, k = b / 8;
    // the first `rem` XCDs own one extra tile
    const int64_t start = xcd * per + (xcd < rem ? xcd : rem);
    return start + k;
}

// Batches: one grid over the blocks of many independent images; image j owns the
// blocks [start[j], start[j + 1]).  Returns the largest j with start[j] <= b (images
// without blocks are skipped; start[0] = 0 <= b).  A 64-ary search by the whole
// wavefront: per round every lane tests one candidate and a ballot keeps the last one
// not above b, so 2000 images cost 2 dependent loads instead of the 11 of a binary
// search.  Every lane of the wavefront must call it together (1D blocks of a multiple
// of 64 threads, uniform control flow); the result is the same in every lane.
__device__ __forceinline__ int find_seg(const int64_t* start, int n, int64_t b) {
    const int lane = (int)(threadIdx.x & 63);
    int lo = 0, span = n;  // the answer lies in [lo, lo + span)
    while (span > 1) {
        const int step = (span + 63) >> 6;
        const int k = lo + lane * step;
        const uint64_t m = __ballot(k < lo + span && start[k] <= b);
        const int nlo = lo + (63 - __clzll((long long)m)) * step;  // lane 0 (k = lo) always holds
        span = min(step, lo + span - nlo);
        lo = nlo;
    }
    return lo;
}

inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : g);
}

// ----------------------------------------------------------------- typed access
// Saturating float -> integer conversion, then truncation toward zero
// (pyclesperanto_prototype/_tier0/preamble.cl:505-591).  The type's largest value as a
// float rounds up to 2^31 / 2^32 for the 32-bit types, so values from there on (where
// the reference's `value > max` test falls through to an out-of-range cast) saturate
// too; the limits are returned as integers, never cast from a float.  NaN -> 0.
template <typename T> __device__ __forceinline__ T sat_cast(float v) {
    constexpr T tmax = std::numeric_limits<T>::max(), tmin = std::numeric_limits<T>::lowest();
    constexpr float fhi = (float)tmax, flo = (float)tmin;
    if (v >= fhi) return tmax;
    if (v < flo) return tmin;
    if (v != v) return (T)0;
    return (T)v;  // flo <= v < fhi: inside T's range
}

__device__ __forceinline__ float ld(const DBuf b, int64_t i) {
    switch (b.dt) {
        case CLE_F32: return ((const float*)b.p)[i];
        case CLE_U8: return (float)((const uint8_t*)b.p)[i];
        case CLE_U16: return (float)((const uint16_t*)b.p)[i];
        case CLE_U32: return (float)((const uint32_t*)b.p)[i];
        case CLE_I32: return (float)((const int32_t*)b.p)[i];
        case CLE_I8: return (float)((const int8_t*)b.p)[i];
        case CLE_I16: return (float)((const int16_t*)b.p)[i];
        default: return 0.f;
    }
}

__device__ __forceinline__ void st(DBuf b, int64_t i, float v) {
    switch (b.dt) {
        case CLE_F32: ((float*)b.p)[i] = v; break;
        case CLE_U8: ((uint8_t*)b.p)[i] = sat_cast<uint8_t>(v); break;
        case CLE_U16: ((uint16_t*)b.p)[i] = sat_cast<uint16_t>(v); break;
        case CLE_U32: ((uint32_t*)b.p)[i] = sat_cast<uint32_t>(v); break;
        case CLE_I32: ((int32_t*)b.p)[i] = sat_cast<int32_t>(v); break;
        case CLE_I8: ((int8_t*)b.p)[i] = sat_cast<int8_t>(v); break;
        case CLE_I16: ((int16_t*)b.p)[i] = sat_cast<int16_t>(v); break;
        default: break;
    }
}

// Compile-time typed load for the kernels that specialise on their source type.
template <typename T> __device__ __forceinline__ float ldt(const T* p, int64_t i) { return (float)p[i]; }

// non-zero test straight on the stored bits (no float round trip)
__device__ __forceinline__ bool nz(const DBuf b, int64_t i) {
    switch (b.dt) {
        case CLE_F32: return ((const float*)b.p)[i] != 0.f;
        case CLE_U8: case CLE_I8: return ((const uint8_t*)b.p)[i] != 0;
        case CLE_U16: case CLE_I16: return ((const uint16_t*)b.p)[i] != 0;
        case CLE_U32: case CLE_I32: return ((const uint32_t*)b.p)[i] != 0;
        default: return false;
    }
}

// detect_maxima_box with radius 0: the 3^d neighbourhood of a voxel is visited x
// outer, then y, then z inner, starting from value - 1, and a strictly larger value
// takes over; the voxel is a maximum if the last take-over was its own value
// (_tier1/detect_maxima_{2d,3d}_x.cl semantics; coordinates below 0 are skipped,
// above the last index clamped).
// A block of 256 threads handles a 64 x 16 (x, y) tile of MX_ZC consecutive slices.
// Every thread owns up to MX_NPT columns (x, y) of the tile plus its 1-voxel x / y
// halo and keeps their slices z - 1, z, z + 1 (and the prefetched z + 2) in registers,
// so the z direction never touches LDS.  Per slice: C(x, y) = maximum over the visited
// z neighbours (registers) goes to LDS; Y(x, y) = max of C over y - 1 .. y + 1 (two LDS
// reads); each output reads Y at x -/+ 1 and C at y -/+ 1 (four LDS reads).  C and Y
// are double-buffered by slice parity: two barriers per slice.
// out(i, is_max) stores voxel i's flag.
constexpr int MX_TX = 64, MX_TY = 16, MX_ZC = 16;

// IEEE maxNum of two operands that are not signalling NaNs (a quiet NaN gives the other
// operand): one v_max_f32.  fmaxf compiles to the same instruction plus a quieting
// v_max_f32 x, x of every operand the compiler cannot prove quiet (values loaded from
// memory or read back from LDS); the maxima tile quiets each voxel value once instead.
__device__ __forceinline__ float qmax(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
constexpr int MX_RX = MX_TX + 2, MX_RY = MX_TY + 2, MX_RP = MX_RX * MX_RY;
constexpr int MX_NPT = (MX_RP + 255) / 256;  // columns per thread
constexpr int MX_LDS = 4 * MX_RP;            // floats of LDS per block: C and Y, two slices each

__host__ __device__ inline int64_t maxima_tiles(const Geom& g) {
    const int64_t nzc = g.ndim == 3 ? (g.D + MX_ZC - 1) / MX_ZC : 1;
    return ((g.W + MX_TX - 1) / MX_TX) * ((g.H + MX_TY - 1) / MX_TY) * nzc;
}

template <typename LD, typename OUT>
__device__ __forceinline__ void maxima_tile(const LD& in, const Geom& g, int64_t t, float* lds, const OUT& out) {
    const int tid = threadIdx.x;
    const int64_t ntx = (g.W + MX_TX - 1) / MX_TX, nty = (g.H + MX_TY - 1) / MX_TY;
    const int64_t x0 = (t % ntx) * MX_TX, y0 = ((t / ntx) % nty) * MX_TY, zc = t / (ntx * nty);
    const bool d3 = g.ndim == 3;
    const int64_t z0 = d3 ? zc * MX_ZC : 0, z1 = d3 ? (z0 + MX_ZC < g.D ? z0 + MX_ZC : g.D) : 1;
    // voxel indices in 32 bits (images hold fewer than 2^32 voxels)
    const uint32_t plane = (uint32_t)(g.H * g.W);
    const float qnan = __builtin_nanf("");
    // column k of this thread: region cell c = tid + 256 k.  Positions before the
    // image's first column / row are never visited (NaN: nothing); those past its last
    // column / row read the clamped last one, as the visit's clamped offsets do.
    // ooff: the in-plane index of the cell's output voxel (interior cells).
    uint32_t off[MX_NPT], ooff[MX_NPT];
    bool live[MX_NPT];
#pragma unroll
    for (int k = 0; k < MX_NPT; ++k) {
        const int c = tid + k * 256;
        const int64_t xx = x0 - 1 + c % MX_RX, yy = y0 - 1 + c / MX_RX;
        live[k] = c < MX_RP && xx >= 0 && yy >= 0;
        off[k] = (uint32_t)((yy < g.H ? yy : g.H - 1) * g.W + (xx < g.W ? xx : g.W - 1));
        ooff[k] = (uint32_t)(yy * g.W + xx);
    }
    float pm[MX_NPT], p0[MX_NPT], pp[MX_NPT], nx[MX_NPT];
    auto load = [&](float* r, int64_t zz) {
        const uint32_t zb = (uint32_t)zz * plane;
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) r[k] = live[k] ? in((int64_t)(zb + off[k])) : qnan;
    };
    // every value enters the maxima quieted (a signalling NaN becomes a quiet one, which
    // the visit treats alike: a NaN never takes over)
    auto quiet = [&](float* r) {
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) r[k] = __builtin_canonicalizef(r[k]);
    };
    if (d3 && z0 > 0) load(pm, z0 - 1);
    load(p0, z0);
    if (d3 && z0 + 1 < g.D) load(pp, z0 + 1);
    quiet(pm);
    quiet(p0);
    quiet(pp);
    for (int64_t z = z0; z < z1; ++z) {
        const bool below = d3 && z > 0, above = d3 && z + 1 < g.D;
        const bool pre = d3 && z + 2 < g.D && z + 1 < z1;
        if (pre) load(nx, z + 2);
        float* Cl = lds + (int)(z & 1) * (2 * MX_RP);
        float* Yl = Cl + MX_RP;
        float cm[MX_NPT];
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) {
            float m = p0[k];
            if (below) m = qmax(m, pm[k]);
            if (above) m = qmax(m, pp[k]);
            cm[k] = m;
            if (tid + k * 256 < MX_RP) Cl[tid + k * 256] = m;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) {
            const int c = tid + k * 256, cy = c / MX_RX;
            if (c < MX_RP && cy >= 1 && cy <= MX_TY) Yl[c] = qmax(qmax(Cl[c - MX_RX], cm[k]), Cl[c + MX_RX]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) {
            const int c = tid + k * 256, cx = c % MX_RX, cy = c / MX_RX;
            if (c >= MX_RP || cx < 1 || cx > MX_TX || cy < 1 || cy > MX_TY) continue;
            const int64_t xx = x0 + cx - 1, yy = y0 + cy - 1;
            if (xx >= g.W || yy >= g.H) continue;
            const float v = p0[k];
            // The visit in order (x, then y, then z offsets; start lm = v - 1, a strictly
            // larger value takes over, the flag is whether the last take-over was the
            // centre) in closed form: B / A = the largest value visited before / after
            // the centre (qmax ignores a NaN, which never takes over; NaN = nothing
            // visited).  Before the centre: the x - 1 columns, (x, y - 1), (x, y, z - 1);
            // after it: (x, y, z + 1), (x, y + 1), the x + 1 columns.  A before-value
            // takes over iff B > v - 1, leaving lm1 = max(v - 1, B); the centre takes over
            // iff v > lm1 (then lm2 = v); with no take-over before it, the flag stays set;
            // any after-value above lm2 clears it.  Exact for ties, infinities, NaN and
            // values where v - 1 == v.
            float bmax = qmax(Yl[c - 1], Cl[c - MX_RX]);
            float amax = qmax(Yl[c + 1], Cl[c + MX_RX]);
            if (below) bmax = qmax(bmax, pm[k]);
            if (d3) amax = qmax(amax, above ? pp[k] : v);
            const float lm0 = v - 1.f;
            const bool took = bmax > lm0;
            const float lm1 = took ? bmax : lm0;
            const bool centre = v > lm1;
            const float lm2 = centre ? v : lm1;
            const bool self = (centre || !took) && !(amax > lm2);
            out((int64_t)((uint32_t)z * plane + ooff[k]), self);
        }
#pragma unroll
        for (int k = 0; k < MX_NPT; ++k) {
            pm[k] = p0[k];
            p0[k] = pp[k];
            pp[k] = __builtin_canonicalizef(nx[k]);
        }
    }
}

// order-preserving map float -> uint32 (for atomicMin / atomicMax of float minima / maxima)
__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(b);
}

// ----------------------------------------------------------------- batches
// One image of a batch.  Work images of all the batch's images are packed in shared
// buffers, image j at voxel offset `off` (a multiple of 64).  Stages that walk voxels
// give each image ceil(n / BV) blocks of BV voxels.
constexpr int BV = 4096;

struct BImg {
    Geom g;
    int64_t off;        // voxel offset in the packed work buffers
    const void* src;    // caller's source image
    void* dst;          // caller's label image
    int src_dt, dst_dt;
    int big;            // n >= 2^24: Otsu runs through the single-image path
    int pad;
    int64_t kbase;      // CCL: first key bit of the image (multiple of 32)
    int64_t boff;       // CCL: first 2x2x2 block of the image
    int64_t tbase;      // voronoi: first tile of the image
    int vnx, vny, vnz;  // voronoi: tile counts along x, y, z
    int vrun;           // voronoi: tiles per run of the growth launches
    int64_t wbase;      // tile growth: first tile of the image
    int wnx, wny, wnz;  // tile growth: tile counts along x, y, z
    int pad3;
};

// Host-side staging of a batch's device tables: appended while the stages are
// planned, uploaded with one copy; a table then lives at device base + its offset.
struct Blob {
    std::vector<char> h;
    size_t put(const void* p, size_t bytes) {
        const size_t o = (h.size() + 255) & ~(size_t)255;
        h.resize(o + (bytes ? bytes : 1));
        if (bytes) memcpy(h.data() + o, p, bytes);
        return o;
    }
    template <typename T> size_t put_vec(const std::vector<T>& v) { return put(v.data(), v.size() * sizeof(T)); }
};

// A planned launch; runs with the device base of the uploaded blob.
using Launch = std::function<int(const char* dev)>;

// block-prefix table of the voxel stages: start[j] = sum over i < j of ceil(n_i / BV)
inline std::vector<int64_t> voxel_block_starts(const std::vector<BImg>& imgs, int64_t* total) {
    std::vector<int64_t> st(imgs.size() + 1);
    int64_t t = 0;
    for (size_t j = 0; j < imgs.size(); ++j) {
        st[j] = t;
        t += (imgs[j].g.n() + BV - 1) / BV;
    }
    st[imgs.size()] = t;
    *total = t;
    return st;
}

// ----------------------------------------------------------------- internal APIs
// Used across translation units by the composite entry points.
int blur_impl(const cle_buf* src, const cle_buf* dst, const float sig[3], hipStream_t s);
int minmax_impl(DBuf src, int64_t n, float* out2, hipStream_t s);
// Otsu threshold of `src`, fully on the device: writes minmax[2] and threshold
// (float64) to device memory.
int otsu_impl(DBuf src, const Geom& g, float* minmax_dev, double* thr_dev, hipStream_t s);
// the same with minmax_dev already holding src's minimum and maximum
int otsu_after_minmax(DBuf src, const Geom& g, float* minmax_dev, double* thr_dev, hipStream_t s);
// minimum / maximum words (order-preserving uint32, see f2o): reset, and to float[2]
void minmax_words_init(uint32_t* words, hipStream_t s);
void minmax_words_finish(const uint32_t* words, float* out2, hipStream_t s);
// blur_impl that also folds the output's minimum and maximum into `mm_words` (reset
// by the caller) when the blur runs as the isotropic one-launch kernel; *folded = 1
// then, 0 if the caller must take them from the output
int blur_impl_minmax(const cle_buf* src, const cle_buf* dst, const float sig[3], uint32_t* mm_words, int* folded,
                     hipStream_t s);
int ccl_impl(DBuf binary, const Geom& g, DBuf dst, int connectivity, int options,
             uint32_t* n_labels_dev, hipStream_t s);
// batched stages (defined next to their single-image kernels)
struct BlurReq {
    const void* in;
    int in_dt;
    float* out;
    Geom g;
    float sigma;
};
int plan_blur_batch(const std::vector<BlurReq>& reqs, Blob& blob, std::vector<Launch>& out, hipStream_t s);
// the calling thread's blurs (single and batched) time their candidate launch shapes
// where no choice is known yet (explicit tuning calls only)
void blur_tuning(bool on);
// per image: mm[2 j] = min, mm[2 j + 1] = max, thr[j] = Otsu threshold of the float32
// work image F (images flagged big run otsu_impl one by one)
int plan_otsu_batch(const std::vector<BImg>& imgs, size_t imgs_off, size_t vstart_off, int64_t vblocks,
                    const float* F, float* mm, double* thr, uint32_t* scratch_u32, Blob& blob,
                    std::vector<Launch>& out, hipStream_t s);
// box CCL of the uint8 work image S of every image into the float32 work image F
// (labels as the single-image path gives them, float32-key compatible).
// ccl_batch_layout sets kbase / boff of every image before the image table is staged.
struct CclBatchSizes {
    int64_t nb = 0;       // 2x2x2 blocks of all images
    int64_t bits = 0;     // key bits of all images
};
void ccl_batch_layout(std::vector<BImg>& imgs, CclBatchSizes* sz);
// Batched form of the growth init: the label write of the batched CCL stores
// v = (G - 1) + label (float32) into F and A at each image's offset, stamps
// seed[tbase + tile] = -1 for every growth tile (16 x 8 x 8 voxels in 3D, 32 x 32 in
// 2D) holding a positive v and clears settled[tbase + tile] (set beforehand) for every
// tile holding a mask voxel.
struct GrowInitBatch {
    float* A;
    const uint8_t* G;
    int32_t* seed;
    uint8_t* settled;
};
int plan_ccl_batch(const std::vector<BImg>& imgs, size_t imgs_off, const CclBatchSizes& sz, const uint8_t* S,
                   float* F, Blob& blob, std::vector<Launch>& out, std::vector<std::shared_ptr<void>>& keep,
                   hipStream_t s, const GrowInitBatch* grow = nullptr);
// Box CCL of `binary` whose write pass initialises the masked Voronoi growth instead of
// storing labels: v = (mask - 1) + label (float32, the reference's add order) into A
// and B, seed[tile] = -1 for every growth tile (tx x ty x tz voxels, nx x ny tiles per
// slice / row) holding a positive v, and settled[tile] = 0 (set beforehand) for every
// tile holding a mask voxel.  The mask is uint8; A, B 16-byte and the mask 4-byte
// aligned.
struct GrowInitOut {
    float* A;
    float* B;
    const uint8_t* mask;
    int32_t* seed;
    uint8_t* settled;
    int tx, ty, tz;
    int64_t nx, ny;
    // 3D tiles whose settled flag is 2 hold no mask voxel (flagged before this pass):
    // their state, -1 everywhere, is not written (the growth reads it as -1)
    bool skip_empty = false;
};
int ccl_box_growth_init(DBuf binary, const Geom& g, const GrowInitOut& out, hipStream_t s);
// single-image voronoi_otsu_labeling (the body of cle_voronoi_otsu_labeling)
int voronoi_otsu_single(const cle_buf* src, const cle_buf* dst, float spot_sigma, float outline_sigma,
                        hipStream_t s);
// Upper bounds of the scratch bytes (sums of scratch_round of every allocation alive at
// once) that the stages take for an image of geometry g
size_t ccl_box_scratch_bytes(const Geom& g, int options);
size_t otsu_scratch_bytes(const Geom& g);
size_t voronoi_otsu_scratch_bytes(const Geom& g);

}  // namespace cle
