// Separable stencils on gfx950: normalised Gaussian (gaussian_blur) and box mean
// (mean_box), one pass per axis with clamp-to-edge reads.
//
// Reference semantics: pyclesperanto_prototype/_tier1/_gaussian_blur.py:6-58,
// _tier1/_execute_separable_kernel.py:7-75 (x pass, then y, then z, float32
// intermediates; an axis with sigma <= 0 is copied), _tier0/_sigma_to_kernel_size.py:1-4
// (N = floor(8 sigma), +1 if even).  The kernels use symmetric, pre-normalised tap
// weights w[0..c] and pair the taps v and -v (one FMA per pair), which differs from
// the reference's tap-by-tap float32 sum by rounding only (|rel| < 1e-6).
//
// Layout: every pass streams one f32 (or typed source) image in and one out, 8 B per
// voxel for f32 -> f32.  A block stages its output tile plus the 2c halo in LDS:
//   x pass : 1 row segment of TX outputs, TX + 2c staged values
//   y/z pass: 64 consecutive x columns x TL outputs along the axis, (TL + 2c) x 64
// so every global load is a coalesced 64-wide (or wider) access and each value is
// read from HBM once per pass.
#include <math.h>
#include <vector>

#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

namespace cle {

constexpr int MAXC = 400;  // largest half-width (sigma up to 50)

struct Taps {
    int c;
    float w[MAXC + 1];
};

// ----------------------------------------------------------------- loaders/storers
// raw(i) / cvt(r): the stored value and its conversion to float, for prefetches that
// hold the loaded value in a register and convert it where it is used (a conversion
// next to the load would make the load wait at once)
struct LdF32 {
    const float* p;
    using raw_t = float;
    static constexpr int kSize = 4;
    __device__ __forceinline__ raw_t raw_b(int64_t slice, uint32_t boff) const {
        return *(const float*)((const char*)(p + slice) + boff);
    }
    __device__ __forceinline__ float operator()(int64_t i) const { return p[i]; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static float cvt(raw_t r) { return r; }
};
struct LdU16 {
    const uint16_t* p;
    using raw_t = uint32_t;
    static constexpr int kSize = 2;
    __device__ __forceinline__ raw_t raw_b(int64_t slice, uint32_t boff) const {
        return *(const uint16_t*)((const char*)(p + slice) + boff);
    }
    __device__ __forceinline__ float operator()(int64_t i) const { return (float)p[i]; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static float cvt(raw_t r) { return (float)r; }
};
struct LdU8 {
    const uint8_t* p;
    using raw_t = uint32_t;
    static constexpr int kSize = 1;
    __device__ __forceinline__ raw_t raw_b(int64_t slice, uint32_t boff) const {
        return *(const uint8_t*)((const char*)(p + slice) + boff);
    }
    __device__ __forceinline__ float operator()(int64_t i) const { return (float)p[i]; }
    __device__ __forceinline__ raw_t raw(int64_t i) const { return p[i]; }
    __device__ __forceinline__ static float cvt(raw_t r) { return (float)r; }
};
struct LdAny {
    DBuf b;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(b, i); }
};
struct StF32 {
    float* p;
    __device__ __forceinline__ void operator()(int64_t i, float v) const { p[i] = v; }
};
struct StAny {
    DBuf b;
    __device__ __forceinline__ void operator()(int64_t i, float v) const { st(b, i, v); }
};

// ----------------------------------------------------------------- x pass
constexpr int XB = 256;     // threads
constexpr int XPT = 4;      // outputs per thread
constexpr int TX = XB * XPT;

template <typename LD, typename ST>
__global__ __launch_bounds__(XB) void k_sep_x(LD in, ST out, int64_t W, int64_t rows, int segs,
                                             const Taps taps) {
    extern __shared__ float sm[];
    const int c = taps.c;
    const int64_t blk = blockIdx.x;
    const int64_t row = blk / segs;
    const int seg = (int)(blk - row * segs);
    if (row >= rows) return;
    const int64_t x0 = (int64_t)seg * TX;
    const int64_t base = row * W;
    const int n_st = TX + 2 * c;
    for (int i = threadIdx.x; i < n_st; i += XB) {
        int64_t x = x0 - c + i;
        x = x < 0 ? 0 : (x >= W ? W - 1 : x);
        sm[i] = in(base + x);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < XPT; ++k) {
        const int lx = threadIdx.x + k * XB;
        const int64_t x = x0 + lx;
        if (x < W) {
            const float* s = sm + lx + c;
            float acc = taps.w[0] * s[0];
            for (int v = 1; v <= c; ++v) acc = fmaf(taps.w[v], s[-v] + s[v], acc);
            out(base + x, acc);
        }
    }
}

// ----------------------------------------------------------------- y / z pass
constexpr int SB_X = 64;    // x columns per block (one wavefront)
constexpr int SB_Y = 4;     // wavefronts per block
constexpr int TL = 64;      // outputs along the blur axis per block
constexpr int LDS_MAX_C = 96;

// Blur along an axis of length na with element stride sa.  The image is viewed as
// (no outer lines) x (na) x (W contiguous); line o starts at o * so.
template <typename LD, typename ST>
__global__ __launch_bounds__(SB_X* SB_Y) void k_sep_strided(LD in, ST out, int64_t W, int64_t na,
                                                          int64_t sa, int64_t no, int64_t so,
                                                          int xblocks, int ablocks, const Taps taps) {
    extern __shared__ float sm[];
    const int c = taps.c;
    int64_t blk = blockIdx.x;
    const int xb = (int)(blk % xblocks);
    blk /= xblocks;
    const int ab = (int)(blk % ablocks);
    const int64_t o = blk / ablocks;
    if (o >= no) return;
    const int tx = threadIdx.x & (SB_X - 1);
    const int ty = threadIdx.x / SB_X;
    const int64_t x = (int64_t)xb * SB_X + tx;
    const bool xin = x < W;
    const int64_t a0 = (int64_t)ab * TL;
    const int64_t base = o * so + (xin ? x : 0);
    const int n_st = TL + 2 * c;
    for (int i = ty; i < n_st; i += SB_Y) {
        int64_t a = a0 - c + i;
        a = a < 0 ? 0 : (a >= na ? na - 1 : a);
        sm[i * SB_X + tx] = xin ? in(base + a * sa) : 0.f;
    }
    __syncthreads();
    if (!xin) return;
    for (int la = ty; la < TL; la += SB_Y) {
        const int64_t a = a0 + la;
        if (a >= na) break;
        const float* s = sm + (la + c) * SB_X + tx;
        float acc = taps.w[0] * s[0];
        for (int v = 1; v <= c; ++v) acc = fmaf(taps.w[v], s[-v * SB_X] + s[v * SB_X], acc);
        out(base + a * sa, acc);
    }
}

// direct-read variant for very wide kernels (halo too large for the LDS tile)
template <typename LD, typename ST>
__global__ __launch_bounds__(256) void k_sep_direct(LD in, ST out, int64_t W, int64_t na, int64_t sa,
                                                   int64_t no, int64_t so, const Taps taps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = no * na * W;
    if (i >= total) return;
    const int64_t x = i % W;
    const int64_t a = (i / W) % na;
    const int64_t o = i / (W * na);
    const int64_t base = o * so + x;
    const int c = taps.c;
    float acc = taps.w[0] * in(base + a * sa);
    for (int v = 1; v <= c; ++v) {
        int64_t am = a - v < 0 ? 0 : a - v;
        int64_t ap = a + v >= na ? na - 1 : a + v;
        acc = fmaf(taps.w[v], in(base + am * sa) + in(base + ap * sa), acc);
    }
    out(base + a * sa, acc);
}

// ----------------------------------------------------------------- fused 3D pass
// One launch for the whole separable blur of a float32 / uint16 / uint8 volume into
// float32.  A block owns a FB_TY x FB_TX column tile and a range of output slices.
// It streams the input slices through LDS: each slice tile (with its x and y halo,
// clamped reads) is blurred along x (LDS -> LDS) and then along y (LDS -> registers);
// every thread keeps the last 2*CZ+1 xy-blurred values of its two columns in a
// register ring and emits the z-blurred output slice as soon as the ring is full.
// The next slice's global loads are issued before the current slice is computed.
// HBM traffic: each input voxel read once per z-chunk (+ halo), each output written
// once -- 6 B/voxel for uint16 -> float32.
constexpr int FB_TX = 64, FB_TY = 32, FB_THREADS = 1024;
constexpr int FB_MAXC = 24;
constexpr int FB_MAXPF = ((FB_TY + 2 * FB_MAXC) * (FB_TX + 2 * FB_MAXC) + FB_THREADS - 1) / FB_THREADS;

struct FusedW {
    float wx[FB_MAXC + 1], wy[FB_MAXC + 1], wz[FB_MAXC + 1];
};

template <int CZ, typename LD>
__global__ __launch_bounds__(FB_THREADS) void k_blur_fused(LD in, float* __restrict__ out, Geom g, int cx, int cy,
                                                          int zc, int ntx, int nty, const FusedW w) {
    extern __shared__ float sm[];
    const int AW = FB_TX + 2 * cx, AH = FB_TY + 2 * cy;
    float* A = sm;
    float* B = sm + AH * AW;
    const int64_t nb = (int64_t)gridDim.x;
    const int64_t bid = xcd_tile(blockIdx.x, nb);
    const int64_t tx = bid % ntx, rest = bid / ntx;
    const int64_t ty = rest % nty, tzc = rest / nty;
    const int64_t x0 = tx * FB_TX, y0 = ty * FB_TY;
    const int64_t zb = tzc * zc;
    const int64_t ze = zb + zc < g.D ? zb + zc : g.D;
    const int row = threadIdx.x / (FB_TX / 2);
    const int col = (threadIdx.x % (FB_TX / 2)) * 2;
    const int na = AH * AW;
    float pf[FB_MAXPF];
    float r0[2 * CZ + 1], r1[2 * CZ + 1];
#pragma unroll
    for (int j = 0; j < 2 * CZ + 1; ++j) { r0[j] = 0.f; r1[j] = 0.f; }

    // in-slice offsets of this thread's staged values (a slice holds < 2^31 voxels)
    const int Hm = (int)g.H - 1, Wm = (int)g.W - 1, Wi = (int)g.W;
    int off[FB_MAXPF];
#pragma unroll
    for (int k = 0; k < FB_MAXPF; ++k) {
        const int idx = threadIdx.x + k * FB_THREADS;
        const int ay = idx / AW, ax = idx - ay * AW;
        int gy = (int)y0 - cy + ay, gx = (int)x0 - cx + ax;
        gy = gy < 0 ? 0 : (gy > Hm ? Hm : gy);
        gx = gx < 0 ? 0 : (gx > Wm ? Wm : gx);
        off[k] = gy * Wi + gx;
    }
    auto load_slice = [&](int64_t zi) {
        const int64_t zz = zi < 0 ? 0 : (zi >= g.D ? g.D - 1 : zi);
        const int64_t base = zz * g.H * g.W;
#pragma unroll
        for (int k = 0; k < FB_MAXPF; ++k)
            if (threadIdx.x + k * FB_THREADS < na) pf[k] = in(base + off[k]);
    };

    load_slice(zb - CZ);
    for (int64_t zi = zb - CZ; zi < ze + CZ; ++zi) {
#pragma unroll
        for (int k = 0; k < FB_MAXPF; ++k) {
            const int idx = threadIdx.x + k * FB_THREADS;
            if (idx < na) A[idx] = pf[k];
        }
        __syncthreads();
        if (zi + 1 < ze + CZ) load_slice(zi + 1);
        // x pass: AH rows x FB_TX columns
        for (int idx = threadIdx.x; idx < AH * FB_TX; idx += FB_THREADS) {
            const int ay = idx / FB_TX, bx = idx - ay * FB_TX;
            const float* a = A + ay * AW + bx + cx;
            float acc = w.wx[0] * a[0];
            for (int v = 1; v <= cx; ++v) acc = fmaf(w.wx[v], a[-v] + a[v], acc);
            B[idx] = acc;
        }
        __syncthreads();
        // y pass for this thread's two columns
        const float* b = B + (row + cy) * FB_TX + col;
        float v0 = w.wy[0] * b[0], v1 = w.wy[0] * b[1];
        for (int v = 1; v <= cy; ++v) {
            v0 = fmaf(w.wy[v], b[-v * FB_TX] + b[v * FB_TX], v0);
            v1 = fmaf(w.wy[v], b[-v * FB_TX + 1] + b[v * FB_TX + 1], v1);
        }
        // z ring
#pragma unroll
        for (int j = 0; j < 2 * CZ; ++j) { r0[j] = r0[j + 1]; r1[j] = r1[j + 1]; }
        r0[2 * CZ] = v0;
        r1[2 * CZ] = v1;
        if (zi >= zb + CZ) {
            float o0 = w.wz[0] * r0[CZ], o1 = w.wz[0] * r1[CZ];
#pragma unroll
            for (int k = 1; k <= CZ; ++k) {
                o0 = fmaf(w.wz[k], r0[CZ - k] + r0[CZ + k], o0);
                o1 = fmaf(w.wz[k], r1[CZ - k] + r1[CZ + k], o1);
            }
            const int64_t zo = zi - CZ, y = y0 + row, x = x0 + col;
            if (y < g.H) {
                float* o = out + (zo * g.H + y) * g.W + x;
                if (x + 1 < g.W) { o[0] = o0; o[1] = o1; }
                else if (x < g.W) o[0] = o0;
            }
        }
    }
}

// ----------------------------------------------------------------- isotropic fused pass
// The common case (sigma equal on every blurred axis): half-width C is a compile-time
// constant, so the tap loops unroll, the weights live in scalar registers and every
// LDS offset is an immediate.  Same tiling as k_blur_fused; the staged slice values
// are prefetched two slices ahead (double-buffered registers) to cover HBM latency.
template <int C>
struct IsoW {
    float w[C + 1];
};

// Isotropic tile: TY x ISO_TX outputs per slice (TY = ISO_TY, or 16 for the batched
// kernels when the autotuner picks it), two adjacent columns per thread.
// A launch aims at ISO_BLOCKS blocks (two rounds of one block per CU): fewer, longer
// z chunks re-read less of the z halo.
constexpr int ISO_TX = 64, ISO_TY = 32, ISO_THREADS = ISO_TX * ISO_TY / 2;
constexpr int ISO_BLOCKS = 512;

template <int C, int TY = ISO_TY>
struct IsoLds {
    static constexpr int AW = ISO_TX + 2 * C, AH = TY + 2 * C;
    static constexpr int AWP = AW;                // row pitch of both buffers (even: float2 reads)
    static constexpr int ASZ = AH * AWP;          // one staged slice (with x and y halo)
    static constexpr int BSZ = TY * AWP;          // one y-blurred slice (with x halo)
};

// One column tile (bid = tile index inside the image) of the isotropic blur; A and B
// hold two IsoLds<C>::ASZ and two BSZ floats (LDS).  Per input slice: the staged tile
// (x and y halo, clamped reads) is blurred along y (A -> B, over the x halo columns
// too), then along x into registers; the z pass runs over a register ring of the
// last 2 * CZ + U xy-blurred values of the thread's two columns.  Doing y before x
// blurs 1 + 2C / ISO_TX of the tile in the first pass instead of 1 + 2C / ISO_TY.
template <int C, bool Z3, typename LD, int U, int PB, int TY = ISO_TY>
__device__ __forceinline__ void blur_iso_tile(LD in, float* __restrict__ out, const Geom& g, int zc, int ntx, int nty,
                                              const IsoW<C>& w, int64_t bid, float* __restrict__ A,
                                              float* __restrict__ B, uint32_t* __restrict__ mm = nullptr) {
    constexpr int NT = ISO_TX * TY / 2;  // threads
    constexpr int AW = IsoLds<C, TY>::AW, AH = IsoLds<C, TY>::AH, NA = AW * AH;
    constexpr int AWP = IsoLds<C, TY>::AWP, ASZ = IsoLds<C, TY>::ASZ, BSZ = IsoLds<C, TY>::BSZ;
    constexpr int PF = (NA + NT - 1) / NT;
    constexpr int CZ = Z3 ? C : 0;
    static_assert(U % 2 == 0, "slices per ring shift: even");
    constexpr int RL = 2 * CZ + U;       // ring length
    constexpr int RY = 4;                // y-pass outputs per task (rows of one column)
    constexpr int YT = AW * (TY / RY);
    constexpr int YR = (YT + NT - 1) / NT;
    const int64_t tx = bid % ntx, rest = bid / ntx;
    const int64_t ty = rest % nty, tzc = rest / nty;
    const int64_t x0 = tx * ISO_TX, y0 = ty * TY;
    const int64_t zb = tzc * zc;
    const int64_t ze = zb + zc < g.D ? zb + zc : g.D;
    const int row = threadIdx.x / (ISO_TX / 2);
    const int col = (threadIdx.x % (ISO_TX / 2)) * 2;
    const int Hm = (int)g.H - 1, Wm = (int)g.W - 1, Wi = (int)g.W;
    // 64-VGPR form: global addresses as a uniform slice base plus a 32-bit byte offset
    // (no 64-bit per-lane address arithmetic); the other form indexes from the image base
    const uint32_t obyte = (uint32_t)(((int)y0 + row) * Wi + (int)x0 + col) * 4u;
    static_assert(AWP == AW, "staged slice stored at its linear index");
    int off[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int idx = threadIdx.x + k * NT;
        const int ay = idx / AW, ax = idx - ay * AW;
        int gy = (int)y0 - C + ay, gx = (int)x0 - C + ax;
        gy = gy < 0 ? 0 : (gy > Hm ? Hm : gy);
        gx = gx < 0 ? 0 : (gx > Wm ? Wm : gx);
        off[k] = PB == 1 ? (int)((uint32_t)(gy * Wi + gx) * (uint32_t)LD::kSize) : gy * Wi + gx;
    }
    using R = typename LD::raw_t;
    R pa[PF], pb[PB == 2 ? PF : 1];
    float r0[RL], r1[RL];
#pragma unroll
    for (int j = 0; j < RL; ++j) { r0[j] = 0.f; r1[j] = 0.f; }
    // minimum / maximum of this thread's outputs (folded into mm at the end when asked)
    float omn = INFINITY, omx = -INFINITY;
    const int64_t z00 = zb - CZ;           // input slice k is z00 + k
    const int64_t N = ze + CZ - z00;       // input slices

    auto load = [&](R (&pf)[PF], int64_t k) {
        if (k >= N) return;
        const int64_t zi = z00 + k;
        const int64_t zz = zi < 0 ? 0 : (zi >= g.D ? g.D - 1 : zi);
        const int64_t base = zz * g.H * g.W;
#pragma unroll
        for (int m = 0; m < PF; ++m)
            if (threadIdx.x + m * NT < NA)
                pf[m] = PB == 1 ? in.raw_b(base, (uint32_t)off[m]) : in.raw(base + off[m]);
    };
    auto store = [&](const R (&pf)[PF], float* a) {
#pragma unroll
        for (int m = 0; m < PF; ++m)
            if (threadIdx.x + m * NT < NA) a[threadIdx.x + m * NT] = LD::cvt(pf[m]);
    };
    // The x and y passes take the taps in order (w[|j - C|] * in[j], j = 0 .. 2C),
    // one fused multiply-add each, as the values stream in from LDS.
    // y pass of one staged slice: a (LDS) -> bo (LDS); a task is RY rows of one
    // column (consecutive lanes take consecutive columns)
    auto ypass = [&](const float* __restrict__ a, float* __restrict__ bo) {
#pragma unroll
        for (int k = 0; k < YR; ++k) {
            const int t = threadIdx.x + k * NT;
            if (t < YT) {
                const int q = t / AW, c = t - q * AW;
                const float* ap = a + q * RY * AWP + c;
                float acc[RY];
#pragma unroll
                for (int j = 0; j < RY; ++j) acc[j] = 0.f;
#pragma unroll
                for (int m = 0; m < RY + 2 * C; ++m) {
                    const float av = ap[m * AWP];
#pragma unroll
                    for (int j = 0; j < RY; ++j) {
                        const int d = m - j - C;  // tap offset of input row m for output j
                        if (d >= -C && d <= C) acc[j] = fmaf(w.w[d < 0 ? -d : d], av, acc[j]);
                    }
                    // 64-VGPR form: a scheduling fence every few rows keeps the compiler
                    // from hoisting all the pass's LDS reads (registers it does not have)
                    if (PB == 1 && (m % 6) == 5) __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int j = 0; j < RY; ++j) bo[(q * RY + j) * AWP + c] = acc[j];
            }
        }
    };
    // x pass of slice k (LDS -> the thread's two columns), ring slot u, z pass and
    // output when the ring covers the slice k - CZ
    auto xpass = [&](const float* __restrict__ bi, int64_t k, int u) {
        const float* bp = bi + row * AWP + col;  // B columns col .. col + 2C + 1
        float v0 = 0.f, v1 = 0.f;
#pragma unroll
        for (int m = 0; m <= C; ++m) {
            const float2 f = *(const float2*)(bp + 2 * m);
            const int d0 = 2 * m - C, d1 = 2 * m + 1 - C;  // offsets of f.x, f.y for column col
            if (d0 <= C) v0 = fmaf(w.w[d0 < 0 ? -d0 : d0], f.x, v0);
            if (d1 <= C) v0 = fmaf(w.w[d1 < 0 ? -d1 : d1], f.y, v0);
            if (d0 - 1 >= -C) v1 = fmaf(w.w[d0 - 1 < 0 ? 1 - d0 : d0 - 1], f.x, v1);
            v1 = fmaf(w.w[d1 - 1 < 0 ? 1 - d1 : d1 - 1], f.y, v1);
            if (PB == 1 && (m % 4) == 3) __builtin_amdgcn_sched_barrier(0);  // as in the y pass
        }
        r0[2 * CZ + u] = v0;
        r1[2 * CZ + u] = v1;
        if (k >= 2 * CZ) {
            float o0, o1;
            if (CZ == 0) {
                o0 = r0[u];
                o1 = r1[u];
            } else {
                o0 = w.w[0] * r0[u + CZ];
                o1 = w.w[0] * r1[u + CZ];
#pragma unroll
                for (int kk = 1; kk <= CZ; ++kk) {
                    o0 = fmaf(w.w[kk], r0[u + CZ - kk] + r0[u + CZ + kk], o0);
                    o1 = fmaf(w.w[kk], r1[u + CZ - kk] + r1[u + CZ + kk], o1);
                }
            }
            const int64_t zo = z00 + k - CZ, y = y0 + row, x = x0 + col;
            if (y < g.H) {
                float* o = PB == 1 ? (float*)((char*)(out + zo * g.H * g.W) + obyte)
                                   : out + (zo * g.H + y) * g.W + x;
                if (x + 1 < g.W) { o[0] = o0; o[1] = o1; }
                else if (x < g.W) o[0] = o0;
                omn = fminf(omn, fminf(x < g.W ? o0 : INFINITY, x + 1 < g.W ? o1 : INFINITY));
                omx = fmaxf(omx, fmaxf(x < g.W ? o0 : -INFINITY, x + 1 < g.W ? o1 : -INFINITY));
            }
        }
    };

    // Software pipeline over the input slices, one barrier per phase: phase p stages
    // slice p (into A[p & 1]), runs the y pass of slice p - 1 (A[(p-1) & 1] ->
    // B[(p-1) & 1]) and the x pass of slice p - 2 (from B[p & 1]).  The buffers a phase
    // writes were last read one phase earlier.  The loads run PB slices ahead of the
    // staging: a phase prefetches slice p + 2 (two register buffers pa / pb) or p + 1
    // (one buffer).
    if constexpr (PB == 2) {
        load(pa, 0);
        load(pb, 1);
        store(pa, A);
        load(pa, 2);
    } else {
        load(pa, 0);
        store(pa, A);
        load(pa, 1);
    }
    __syncthreads();
    if (N > 1) {
        if constexpr (PB == 2) {
            store(pb, A + ASZ);
            load(pb, 3);
        } else {
            store(pa, A + ASZ);
            load(pa, 2);
        }
    }
    ypass(A, B);
    __syncthreads();
    for (int64_t g0 = 0; g0 < N; g0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t k = g0 + u;  // slice of the x pass; p = k + 2 has parity u & 1
            if (k < N) {
                float* Ap = A + ((u & 1) ? ASZ : 0);       // A[p & 1]
                float* Aq = A + ((u & 1) ? 0 : ASZ);       // A[(p - 1) & 1]
                float* Bq = B + ((u & 1) ? 0 : BSZ);       // B[(p - 1) & 1]
                const float* Bk = B + ((u & 1) ? BSZ : 0);  // B[k & 1]
                if (k + 2 < N) {
                    if constexpr (PB == 2) {
                        if (u & 1) { store(pb, Ap); load(pb, k + 4); }
                        else { store(pa, Ap); load(pa, k + 4); }
                    } else {
                        store(pa, Ap);
                        load(pa, k + 3);
                    }
                }
                if (k + 1 < N) ypass(Aq, Bq);
                xpass(Bk, k, u);
                __syncthreads();
            }
        }
#pragma unroll
        for (int j = 0; j < 2 * CZ; ++j) { r0[j] = r0[j + U]; r1[j] = r1[j + U]; }
    }
    if (mm) {  // block minimum / maximum -> two atomics (A is free after the last barrier)
        for (int off = 32; off > 0; off >>= 1) {
            omn = fminf(omn, __shfl_xor(omn, off));
            omx = fmaxf(omx, __shfl_xor(omx, off));
        }
        constexpr int NW = NT / 64;
        if ((threadIdx.x & 63) == 0) {
            A[threadIdx.x >> 6] = omn;
            A[NW + (threadIdx.x >> 6)] = omx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < NW; ++k) {
                omn = fminf(omn, A[k]);
                omx = fmaxf(omx, A[NW + k]);
            }
            atomicMin(mm, f2o(omn));
            atomicMax(mm + 1, f2o(omx));
        }
    }
}

// Two register budgets of the same tile code.  k_blur_iso leaves the compiler its
// choice (about 80 VGPRs at C = 8: one 16-wavefront block per CU), shifts the z ring
// every U = 4 slices (2 above C = 10) and prefetches two slices ahead; k_blur_iso_o8
// holds it to 64 VGPRs (eight wavefronts per SIMD: two blocks per CU, each running
// while the other waits at a barrier) with a ring shift every 2 slices and a one-slice
// prefetch.  The 64-VGPR form exists for C <= 8 (3D) / 12 (2D); beyond that it spills.
// Both sum every voxel's taps in the same order: the output does not depend on the
// choice.
template <int C, bool Z3> constexpr bool iso_o8_ok() { return C <= (Z3 ? 8 : 12); }
template <int C> constexpr int iso_u() { return C <= 10 ? 4 : 2; }

template <int C, bool Z3, typename LD>
__global__ __launch_bounds__(ISO_THREADS) void k_blur_iso(LD in, float* __restrict__ out, Geom g, int zc, int ntx,
                                                        int nty, const IsoW<C> w, uint32_t* mm) {
    __shared__ __align__(16) float A[2 * IsoLds<C>::ASZ];
    __shared__ __align__(16) float B[2 * IsoLds<C>::BSZ];
    blur_iso_tile<C, Z3, LD, iso_u<C>(), 2>(in, out, g, zc, ntx, nty, w, xcd_tile(blockIdx.x, gridDim.x), A, B, mm);
}
template <int C, bool Z3, typename LD>
__global__ __launch_bounds__(ISO_THREADS, 8) void k_blur_iso_o8(LD in, float* __restrict__ out, Geom g, int zc,
                                                              int ntx, int nty, const IsoW<C> w, uint32_t* mm) {
    __shared__ __align__(16) float A[2 * IsoLds<C>::ASZ];
    __shared__ __align__(16) float B[2 * IsoLds<C>::BSZ];
    blur_iso_tile<C, Z3, LD, 2, 1>(in, out, g, zc, ntx, nty, w, xcd_tile(blockIdx.x, gridDim.x), A, B, mm);
}

// Batched form: one grid over the column tiles of many images that share the
// half-width C, the dimensionality and the source type; image j owns the blocks
// [bstart[j], bstart[j + 1]).
template <int C>
struct IsoJob {
    const void* in;
    float* out;
    Geom g;
    int zc, ntx, nty;
    IsoW<C> w;
};

template <typename LD> __device__ __forceinline__ LD make_ld(const void* p);
template <> __device__ __forceinline__ LdF32 make_ld<LdF32>(const void* p) { return LdF32{(const float*)p}; }
template <> __device__ __forceinline__ LdU16 make_ld<LdU16>(const void* p) { return LdU16{(const uint16_t*)p}; }
template <> __device__ __forceinline__ LdU8 make_ld<LdU8>(const void* p) { return LdU8{(const uint8_t*)p}; }

template <int C, bool Z3, typename LD, int U, int PB, int TY = ISO_TY>
__device__ __forceinline__ void blur_iso_batch_tile(const IsoJob<C>* __restrict__ jobs,
                                                    const int64_t* __restrict__ bstart, int njobs) {
    __shared__ __align__(16) float A[2 * IsoLds<C, TY>::ASZ];
    __shared__ __align__(16) float B[2 * IsoLds<C, TY>::BSZ];
    const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
    const int j = find_seg(bstart, njobs, b);
    const IsoJob<C>& jb = jobs[j];
    blur_iso_tile<C, Z3, LD, U, PB, TY>(make_ld<LD>(jb.in), jb.out, jb.g, jb.zc, jb.ntx, jb.nty, jb.w, b - bstart[j], A,
                                        B);
}
template <int C, bool Z3, typename LD, int TY = ISO_TY>
__global__ __launch_bounds__(ISO_TX * TY / 2) void k_blur_iso_batch(const IsoJob<C>* __restrict__ jobs,
                                                                  const int64_t* __restrict__ bstart, int njobs) {
    blur_iso_batch_tile<C, Z3, LD, iso_u<C>(), 2, TY>(jobs, bstart, njobs);
}
template <int C, bool Z3, typename LD>
__global__ __launch_bounds__(ISO_THREADS, 8) void k_blur_iso_batch_o8(const IsoJob<C>* __restrict__ jobs,
                                                                    const int64_t* __restrict__ bstart, int njobs) {
    blur_iso_batch_tile<C, Z3, LD, 2, 1>(jobs, bstart, njobs);
}

// ----------------------------------------------------------------- host side
static void gauss_taps(float sigma, Taps* t) {
    int n = (int)(sigma * 8.0f);  // _tier0/_sigma_to_kernel_size.py:1-4
    if (n % 2 == 0) n += 1;
    const int c = (n - 1) / 2;
    t->c = c;
    if (c > MAXC) return;  // refused by the caller; the taps would not fit
    double h[MAXC + 1];
    double sum = 0;
    const double s = (double)sigma;
    for (int v = 0; v <= c; ++v) {
        h[v] = exp(-(double)(v * v) / (2.0 * s * s));
        sum += (v == 0 ? 1.0 : 2.0) * h[v];
    }
    for (int v = 0; v <= c; ++v) t->w[v] = (float)(h[v] / sum);
}

static void box_taps(int r, Taps* t) {
    t->c = r;
    for (int v = 0; v <= r; ++v) t->w[v] = 1.0f / (float)(2 * r + 1);
}

template <int CZ, typename LD>
static void launch_fused_cz(LD in, float* out, const Geom& g, int cx, int cy, const FusedW& w, hipStream_t s) {
    const int ntx = (int)((g.W + FB_TX - 1) / FB_TX), nty = (int)((g.H + FB_TY - 1) / FB_TY);
    const int64_t nxy = (int64_t)ntx * nty;
    // enough blocks to fill the chip, but z chunks long enough to amortise the z halo
    int64_t nzc = (2048 + nxy - 1) / nxy;
    const int64_t max_nzc = g.D / (4 * CZ + 8) > 1 ? g.D / (4 * CZ + 8) : 1;
    if (nzc > max_nzc) nzc = max_nzc;
    if (nzc < 1) nzc = 1;
    const int zc = (int)((g.D + nzc - 1) / nzc);
    nzc = (g.D + zc - 1) / zc;
    const size_t lds = (size_t)((FB_TY + 2 * cy) * (FB_TX + 2 * cx) + (FB_TY + 2 * cy) * FB_TX) * sizeof(float);
    k_blur_fused<CZ, LD><<<(unsigned)(nxy * nzc), FB_THREADS, lds, s>>>(in, out, g, cx, cy, zc, ntx, nty, w);
}

template <typename LD>
static int launch_fused(LD in, float* out, const Geom& g, int cx, int cy, int cz, const FusedW& w, hipStream_t s) {
    switch (cz) {
#define CZ_CASE(n) case n: launch_fused_cz<n>(in, out, g, cx, cy, w, s); break;
        CZ_CASE(0) CZ_CASE(1) CZ_CASE(2) CZ_CASE(3) CZ_CASE(4) CZ_CASE(5) CZ_CASE(6) CZ_CASE(7) CZ_CASE(8)
        CZ_CASE(9) CZ_CASE(10) CZ_CASE(11) CZ_CASE(12) CZ_CASE(13) CZ_CASE(14) CZ_CASE(15) CZ_CASE(16)
        CZ_CASE(17) CZ_CASE(18) CZ_CASE(19) CZ_CASE(20) CZ_CASE(21) CZ_CASE(22) CZ_CASE(23) CZ_CASE(24)
#undef CZ_CASE
        default: return fail(CLE_E_INTERNAL, "fused blur: z half-width %d out of range", cz);
    }
    CLE_LAUNCHED();
    return CLE_OK;
}

// Fused path when the destination is float32, the source is f32/u16/u8 and every
// half-width fits the LDS tile; returns 1 if it ran, 0 if the caller must use the
// per-axis passes.
static int try_fused(const cle_buf* src, const cle_buf* dst, const Taps* taps[3], int* ran, hipStream_t s) {
    *ran = 0;
    if (dst->dtype != CLE_F32 || dst->data == src->data) return CLE_OK;
    if (src->dtype != CLE_F32 && src->dtype != CLE_U16 && src->dtype != CLE_U8) return CLE_OK;
    Geom g;
    CLE_TRY(check_buf(src, "source", &g));
    int c[3];
    FusedW w;
    float* ws[3] = {w.wx, w.wy, w.wz};
    for (int a = 0; a < 3; ++a) {
        c[a] = taps[a] ? taps[a]->c : 0;
        if (c[a] > FB_MAXC) return CLE_OK;
        for (int v = 0; v <= FB_MAXC; ++v) ws[a][v] = 0.f;
        if (taps[a]) for (int v = 0; v <= c[a]; ++v) ws[a][v] = taps[a]->w[v];
        else ws[a][0] = 1.f;
    }
    float* out = (float*)dst->data;
    int rc;
    if (src->dtype == CLE_F32) rc = launch_fused(LdF32{(const float*)src->data}, out, g, c[0], c[1], c[2], w, s);
    else if (src->dtype == CLE_U16) rc = launch_fused(LdU16{(const uint16_t*)src->data}, out, g, c[0], c[1], c[2], w, s);
    else rc = launch_fused(LdU8{(const uint8_t*)src->data}, out, g, c[0], c[1], c[2], w, s);
    if (rc == CLE_OK) *ran = 1;
    return rc;
}

// ---------------------------------------------------------------- autotuner
// The isotropic blur splits every column tile's z range into nzc chunks: more chunks
// fill the chip with more blocks, each chunk re-reads 2C halo slices; and it runs
// under one of two register budgets (k_blur_iso / k_blur_iso_o8).  The output is the
// same for every choice (each voxel's taps are summed in the same order).
// cle_autotune_gaussian_blur times the candidate (chunk count, budget) pairs for one
// configuration (half-width, dimensionality, source type, size class) ahead of time
// and keeps the fastest for the process; blurs look the choice up and otherwise use
// the defaults (about ISO_BLOCKS blocks, iso_o8_default), so no blur call ever pays
// for tuning.  CLE_AUTOTUNE=0 or
// cle_set_autotune(0) ignores (and forgets) the choices.
static std::atomic<int> g_autotune{-1};  // -1: read CLE_AUTOTUNE on first use

static bool autotune_on() {
    int v = g_autotune.load();
    if (v < 0) {
        const char* e = getenv("CLE_AUTOTUNE");
        v = (e && e[0] == '0') ? 0 : 1;
        g_autotune.store(v);
    }
    return v != 0;
}

struct TuneKey {
    int c, z3, src, dev, lxy, lz;
    bool operator<(const TuneKey& o) const {
        return std::tie(c, z3, src, dev, lxy, lz) < std::tie(o.c, o.z3, o.src, o.dev, o.lxy, o.lz);
    }
};
struct TuneChoice {
    int64_t nzc;
    int o8;  // register budget: 1 = k_blur_iso_o8
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune;

static int ilog2(int64_t v) {
    int l = 0;
    while (v > 1) { v >>= 1; ++l; }
    return l;
}

// Register budget of a configuration when no tuned choice is known (1: k_blur_iso_o8),
// from MI355X timings in ms, k_blur_iso / k_blur_iso_o8: 3D 256 x 1024 x 1024 u16
// C = 3 0.406 / 0.413, 4 0.605 / 0.431, 6 0.635 / 0.506, 7 0.657 / 0.551, 8 0.682 /
// 0.700 (4 VGPRs spilled); 2D 8192 x 8192 u16 C = 6 0.297 / 0.258, 8 0.324 / 0.272,
// 10 0.505 / 0.288, 12 0.487 / 0.303.
static int iso_o8_default(int c, bool z3) { return z3 ? (c >= 4 && c <= 7) : (c <= 12); }

template <int C, bool Z3, typename LD>
static void launch_iso_nzc(LD in, float* out, const Geom& g, const IsoW<C>& w, int ntx, int nty, int64_t nzc,
                           int o8, uint32_t* mm, hipStream_t s) {
    const int zc = (int)((g.D + nzc - 1) / nzc);
    nzc = (g.D + zc - 1) / zc;
    const unsigned nb = (unsigned)((int64_t)ntx * nty * nzc);
    if constexpr (iso_o8_ok<C, Z3>()) {
        if (o8) {
            k_blur_iso_o8<C, Z3, LD><<<nb, ISO_THREADS, 0, s>>>(in, out, g, zc, ntx, nty, w, mm);
            return;
        }
    }
    k_blur_iso<C, Z3, LD><<<nb, ISO_THREADS, 0, s>>>(in, out, g, zc, ntx, nty, w, mm);
}

template <typename LD> constexpr int ld_code();
template <> constexpr int ld_code<LdF32>() { return 0; }
template <> constexpr int ld_code<LdU16>() { return 1; }
template <> constexpr int ld_code<LdU8>() { return 2; }

static thread_local bool t_tuning = false;  // set by cle_autotune_gaussian_blur around its blur

template <int C, bool Z3, typename LD>
static void launch_iso_c(LD in, float* out, const Geom& g, const float* wts, uint32_t* mm, hipStream_t s) {
    const int ntx = (int)((g.W + ISO_TX - 1) / ISO_TX), nty = (int)((g.H + ISO_TY - 1) / ISO_TY);
    const int64_t nxy = (int64_t)ntx * nty;
    constexpr int CZ = Z3 ? C : 0;
    const int64_t max_nzc = g.D / (4 * CZ + 8) > 1 ? g.D / (4 * CZ + 8) : 1;
    auto nzc_for = [&](int64_t blocks) {
        int64_t v = (blocks + nxy - 1) / nxy;
        return v < 1 ? 1 : (v > max_nzc ? max_nzc : v);
    };
    IsoW<C> w;
    for (int v = 0; v <= C; ++v) w.w[v] = wts[v];
    int64_t nzc = nzc_for(ISO_BLOCKS);
    int o8 = iso_o8_default(C, Z3);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (autotune_on()) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        const TuneKey key{C, Z3 ? 1 : 0, ld_code<LD>(), dev, ilog2(nxy), ilog2(g.D)};
        bool known = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                nzc = it->second.nzc < max_nzc ? it->second.nzc : max_nzc;
                o8 = it->second.o8;
                known = true;
            }
        }
        if (!known && t_tuning && hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone) {
            std::vector<TuneChoice> cand;
            for (int v8 = 0; v8 <= (iso_o8_ok<C, Z3>() ? 1 : 0); ++v8)
                for (int64_t b : {256, 512, 1024, 2048, 4096}) {
                    const int64_t v = nzc_for(b);
                    bool dup = false;
                    for (const TuneChoice& t : cand) dup |= t.nzc == v && t.o8 == v8;
                    if (!dup) cand.push_back({v, v8});
                }
            hipEvent_t e0, e1;
            if (cand.size() > 1 && hipEventCreate(&e0) == hipSuccess) {
                if (hipEventCreate(&e1) == hipSuccess) {
                    float best = 1e30f;
                    for (const TuneChoice& t : cand) {
                        launch_iso_nzc<C, Z3>(in, out, g, w, ntx, nty, t.nzc, t.o8, mm, s);  // warm
                        for (int rep = 0; rep < 3; ++rep) {                                 // best of three
                            (void)hipEventRecord(e0, s);
                            launch_iso_nzc<C, Z3>(in, out, g, w, ntx, nty, t.nzc, t.o8, mm, s);
                            (void)hipEventRecord(e1, s);
                            float ms = 1e30f;
                            if (hipEventSynchronize(e1) == hipSuccess &&
                                hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
                                best = ms;
                                nzc = t.nzc;
                                o8 = t.o8;
                            }
                        }
                    }
                    (void)hipEventDestroy(e1);
                }
                (void)hipEventDestroy(e0);
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[key] = {nzc, o8};
        }
    }
    launch_iso_nzc<C, Z3>(in, out, g, w, ntx, nty, nzc, o8, mm, s);
}

constexpr int ISO_MAXC = 20;

// the isotropic tile code addresses a slice by 32-bit byte offsets (float32 output)
static bool iso_slice_ok(const Geom& g) { return g.H * g.W < ((int64_t)1 << 30); }

template <bool Z3, typename LD>
static int launch_iso(LD in, float* out, const Geom& g, int c, const float* wts, uint32_t* mm, hipStream_t s) {
    switch (c) {
#define ISO_CASE(n) case n: launch_iso_c<n, Z3>(in, out, g, wts, mm, s); break;
        ISO_CASE(1) ISO_CASE(2) ISO_CASE(3) ISO_CASE(4) ISO_CASE(5) ISO_CASE(6) ISO_CASE(7) ISO_CASE(8)
        ISO_CASE(9) ISO_CASE(10) ISO_CASE(11) ISO_CASE(12) ISO_CASE(13) ISO_CASE(14) ISO_CASE(15) ISO_CASE(16)
        ISO_CASE(17) ISO_CASE(18) ISO_CASE(19) ISO_CASE(20)
#undef ISO_CASE
        default: return fail(CLE_E_INTERNAL, "isotropic blur: half-width %d out of range", c);
    }
    CLE_LAUNCHED();
    return CLE_OK;
}

// Isotropic fused path: destination float32, source f32/u16/u8, equal sigma on every
// blurred axis (x, y, and z for 3D), half-width 1..ISO_MAXC.  Sets *ran.
static int try_iso(const cle_buf* src, const cle_buf* dst, const Taps* taps[3], int* ran, uint32_t* mm,
                   hipStream_t s) {
    *ran = 0;
    if (dst->dtype != CLE_F32 || dst->data == src->data) return CLE_OK;
    if (src->dtype != CLE_F32 && src->dtype != CLE_U16 && src->dtype != CLE_U8) return CLE_OK;
    Geom g;
    CLE_TRY(check_buf(src, "source", &g));
    const bool z3 = g.ndim == 3;
    if (!taps[0] || !taps[1] || (z3 && !taps[2])) return CLE_OK;
    const int c = taps[0]->c;
    if (c < 1 || c > ISO_MAXC || taps[1]->c != c || (z3 && taps[2]->c != c) || !iso_slice_ok(g)) return CLE_OK;
    for (int v = 0; v <= c; ++v)
        if (taps[1]->w[v] != taps[0]->w[v] || (z3 && taps[2]->w[v] != taps[0]->w[v])) return CLE_OK;
    float* out = (float*)dst->data;
    const float* wts = taps[0]->w;
    int rc;
    if (src->dtype == CLE_F32) {
        LdF32 in{(const float*)src->data};
        rc = z3 ? launch_iso<true>(in, out, g, c, wts, mm, s) : launch_iso<false>(in, out, g, c, wts, mm, s);
    } else if (src->dtype == CLE_U16) {
        LdU16 in{(const uint16_t*)src->data};
        rc = z3 ? launch_iso<true>(in, out, g, c, wts, mm, s) : launch_iso<false>(in, out, g, c, wts, mm, s);
    } else {
        LdU8 in{(const uint8_t*)src->data};
        rc = z3 ? launch_iso<true>(in, out, g, c, wts, mm, s) : launch_iso<false>(in, out, g, c, wts, mm, s);
    }
    if (rc == CLE_OK) *ran = 1;
    return rc;
}

template <typename LD, typename ST>
static int launch_axis(LD in, ST out, const Geom& g, int axis, const Taps& t, hipStream_t s) {
    if (axis == 0) {
        const int64_t rows = g.D * g.H;
        const int segs = (int)((g.W + TX - 1) / TX);
        const size_t lds = (size_t)(TX + 2 * t.c) * sizeof(float);
        if (lds > 64 * 1024) {
            k_sep_direct<LD, ST><<<grid_for(g.n(), 256), 256, 0, s>>>(in, out, 1, g.W, 1, rows, g.W, t);
        } else {
            k_sep_x<LD, ST><<<(unsigned)(rows * segs), XB, lds, s>>>(in, out, g.W, rows, segs, t);
        }
    } else {
        int64_t na, sa, no, so;
        if (axis == 1) { na = g.H; sa = g.W; no = g.D; so = g.H * g.W; }
        else { na = g.D; sa = g.H * g.W; no = g.H; so = g.W; }
        if (t.c > LDS_MAX_C) {
            k_sep_direct<LD, ST><<<grid_for(g.n(), 256), 256, 0, s>>>(in, out, g.W, na, sa, no, so, t);
        } else {
            const int xblocks = (int)((g.W + SB_X - 1) / SB_X);
            const int ablocks = (int)((na + TL - 1) / TL);
            const size_t lds = (size_t)(TL + 2 * t.c) * SB_X * sizeof(float);
            const int64_t nblk = (int64_t)xblocks * ablocks * no;
            k_sep_strided<LD, ST><<<(unsigned)nblk, SB_X * SB_Y, lds, s>>>(in, out, g.W, na, sa, no, so,
                                                                          xblocks, ablocks, t);
        }
    }
    CLE_LAUNCHED();
    return CLE_OK;
}

// one pass from a typed source into either a float buffer or a typed destination
static int pass_from(DBuf src, DBuf dst, const Geom& g, int axis, const Taps& t, hipStream_t s) {
    const bool dst_f32 = dst.dt == CLE_F32;
    if (src.dt == CLE_F32) {
        LdF32 in{(const float*)src.p};
        return dst_f32 ? launch_axis(in, StF32{(float*)dst.p}, g, axis, t, s)
                       : launch_axis(in, StAny{dst}, g, axis, t, s);
    }
    if (src.dt == CLE_U16) {
        LdU16 in{(const uint16_t*)src.p};
        return dst_f32 ? launch_axis(in, StF32{(float*)dst.p}, g, axis, t, s)
                       : launch_axis(in, StAny{dst}, g, axis, t, s);
    }
    if (src.dt == CLE_U8) {
        LdU8 in{(const uint8_t*)src.p};
        return dst_f32 ? launch_axis(in, StF32{(float*)dst.p}, g, axis, t, s)
                       : launch_axis(in, StAny{dst}, g, axis, t, s);
    }
    LdAny in{src};
    return dst_f32 ? launch_axis(in, StF32{(float*)dst.p}, g, axis, t, s)
                   : launch_axis(in, StAny{dst}, g, axis, t, s);
}

__global__ void k_convert(DBuf s, DBuf d, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) st(d, i, ld(s, i));
}

// Runs the separable passes given by taps[axis] (c < 0 = axis skipped).  `mm` (may be
// null): minimum / maximum words the isotropic kernel folds its output into; *folded
// tells whether it ran.
static int separable_impl(const cle_buf* src, const cle_buf* dst, const Taps* taps[3], hipStream_t s,
                          uint32_t* mm = nullptr, int* folded = nullptr) {
    if (folded) *folded = 0;
    Geom gs, gd;
    CLE_TRY(check_buf(src, "source", &gs));
    CLE_TRY(check_buf(dst, "destination", &gd));
    CLE_TRY(check_same_shape(gs, gd, "source", "destination"));
    if (gs.n() == 0) return CLE_OK;
    int axes[3], npass = 0;
    for (int a = 0; a < 3; ++a)
        if (taps[a]) axes[npass++] = a;
    DBuf in = dbuf(src), out = dbuf(dst);
    if (npass > 0 && gs.ndim >= 2) {
        int ran = 0;
        CLE_TRY(try_iso(src, dst, taps, &ran, mm, s));
        if (ran) {
            if (folded) *folded = mm != nullptr;
            return CLE_OK;
        }
        CLE_TRY(try_fused(src, dst, taps, &ran, s));
        if (ran) return CLE_OK;
    }
    if (npass == 0) {
        unsigned grid = grid_for(gs.n(), 256);
        if (grid > 65536) grid = 65536;
        k_convert<<<grid, 256, 0, s>>>(in, out, gs.n());
        CLE_LAUNCHED();
        return CLE_OK;
    }
    const size_t bytes = (size_t)gs.n() * sizeof(float);
    const bool dst_as_temp = dst->dtype == CLE_F32 && dst->data != src->data;
    Scratch t1, t2;
    DBuf tmp[2];
    if (npass >= 2) {
        if (npass == 3 && dst_as_temp) {
            CLE_TRY(t1.alloc(bytes, s));
            tmp[0] = out;  // pass 1 -> dst, pass 2 -> t1, pass 3 -> dst
            tmp[1] = DBuf{t1.p, CLE_F32};
        } else {
            CLE_TRY(t1.alloc(bytes, s));
            tmp[0] = DBuf{t1.p, CLE_F32};
            if (npass == 3) {
                CLE_TRY(t2.alloc(bytes, s));
                tmp[1] = DBuf{t2.p, CLE_F32};
            }
        }
    }
    DBuf cur = in;
    for (int i = 0; i < npass; ++i) {
        DBuf nxt = (i == npass - 1) ? out : tmp[i];
        CLE_TRY(pass_from(cur, nxt, gs, axes[i], *taps[axes[i]], s));
        cur = nxt;
    }
    return CLE_OK;
}

// ----------------------------------------------------------------- batched blur
// The isotropic requests of a batch are grouped by (half-width, dimensionality, source
// type); each group is one launch of k_blur_iso_batch over the column tiles of all
// its images.  The rest (sigma <= 0, wider kernels, other types) run blur_impl.
// A group's launch shape: the tile height (32 or 16 rows, 1024 or 512 threads), the
// register budget (TY = 32 only) and the number of blocks its z chunking aims at.  The
// default is 32 rows, the single-image budget rule and 256 blocks; the autotuner
// (cle_autotune_voronoi_otsu_labeling_batch) times the candidates of every group of
// one batch and keeps the fastest per (half-width, dimensionality, source type, log2
// voxels, log2 images) for the process.  Every shape sums each voxel's taps in the
// same order: the output does not depend on it.
struct BTuneKey {
    int c, z3, src, dev, lvox, limg;
    bool operator<(const BTuneKey& o) const {
        return std::tie(c, z3, src, dev, lvox, limg) < std::tie(o.c, o.z3, o.src, o.dev, o.lvox, o.limg);
    }
};
struct BTuneChoice {
    int ty, o8;
    int64_t target;  // blocks the z chunking aims at
};
static std::map<BTuneKey, BTuneChoice> g_btune;  // guarded by g_tune_mu

// job table and block prefix of a group for one launch shape; returns the block count
template <int C, bool Z3>
static int64_t iso_group_tables(const std::vector<BlurReq>& reqs, const std::vector<int>& idx, int ty, int64_t target,
                                std::vector<IsoJob<C>>& jobs, std::vector<int64_t>& bst) {
    constexpr int CZ = Z3 ? C : 0;
    jobs.resize(idx.size());
    bst.resize(idx.size() + 1);
    int64_t nb = 0;
    for (size_t k = 0; k < idx.size(); ++k) {
        const BlurReq& r = reqs[idx[k]];
        const Geom& g = r.g;
        IsoJob<C>& jb = jobs[k];
        jb.in = r.in;
        jb.out = r.out;
        jb.g = g;
        jb.ntx = (int)((g.W + ISO_TX - 1) / ISO_TX);
        jb.nty = (int)((g.H + ty - 1) / ty);
        const int64_t nxy = (int64_t)jb.ntx * jb.nty;
        // z chunks only for images with few column tiles
        int64_t nzc = (target + nxy - 1) / nxy;
        const int64_t max_nzc = g.D / (4 * CZ + 8) > 1 ? g.D / (4 * CZ + 8) : 1;
        if (nzc > max_nzc) nzc = max_nzc;
        if (nzc < 1) nzc = 1;
        jb.zc = (int)((g.D + nzc - 1) / nzc);
        nzc = (g.D + jb.zc - 1) / jb.zc;
        Taps t;
        gauss_taps(r.sigma, &t);
        for (int v = 0; v <= C; ++v) jb.w.w[v] = t.w[v];
        bst[k] = nb;
        nb += nxy * nzc;
    }
    bst[idx.size()] = nb;
    return nb;
}

template <int C, bool Z3, typename LD>
static void launch_iso_group(const IsoJob<C>* dj, const int64_t* db, int nj, int64_t nb, const BTuneChoice& ch,
                             hipStream_t s) {
    if (ch.ty == 16) {
        k_blur_iso_batch<C, Z3, LD, 16><<<(unsigned)nb, ISO_TX * 16 / 2, 0, s>>>(dj, db, nj);
        return;
    }
    if constexpr (iso_o8_ok<C, Z3>()) {
        if (ch.o8) {
            k_blur_iso_batch_o8<C, Z3, LD><<<(unsigned)nb, ISO_THREADS, 0, s>>>(dj, db, nj);
            return;
        }
    }
    k_blur_iso_batch<C, Z3, LD><<<(unsigned)nb, ISO_THREADS, 0, s>>>(dj, db, nj);
}

// Times the candidate shapes of one group (its tables uploaded for each) and returns
// the fastest; runs only inside an explicit tuning call.
template <int C, bool Z3, typename LD>
static BTuneChoice tune_iso_group(const std::vector<BlurReq>& reqs, const std::vector<int>& idx,
                                  const BTuneChoice& dflt, hipStream_t s) {
    BTuneChoice best = dflt;
    float best_ms = 1e30f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return best;
    }
    for (int ty : {32, 16})
        for (int o8 = 0; o8 <= (ty == 32 && iso_o8_ok<C, Z3>() ? 1 : 0); ++o8)
            for (int64_t target : {256, 1024, 4096}) {
                const BTuneChoice ch{ty, o8, target};
                std::vector<IsoJob<C>> jobs;
                std::vector<int64_t> bst;
                const int64_t nb = iso_group_tables<C, Z3>(reqs, idx, ty, target, jobs, bst);
                Scratch dt;
                const size_t jb = jobs.size() * sizeof(IsoJob<C>), bb = bst.size() * sizeof(int64_t);
                if (dt.alloc(jb + 256 + bb, s) != CLE_OK) continue;
                char* d = dt.get<char>();
                if (hipMemcpyAsync(d, jobs.data(), jb, hipMemcpyHostToDevice, s) != hipSuccess ||
                    hipMemcpyAsync(d + ((jb + 255) & ~(size_t)255), bst.data(), bb, hipMemcpyHostToDevice, s) !=
                        hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    continue;
                const IsoJob<C>* dj = (const IsoJob<C>*)d;
                const int64_t* dbs = (const int64_t*)(d + ((jb + 255) & ~(size_t)255));
                launch_iso_group<C, Z3, LD>(dj, dbs, (int)idx.size(), nb, ch, s);  // warm
                for (int rep = 0; rep < 3; ++rep) {
                    (void)hipEventRecord(e0, s);
                    launch_iso_group<C, Z3, LD>(dj, dbs, (int)idx.size(), nb, ch, s);
                    (void)hipEventRecord(e1, s);
                    float ms = 1e30f;
                    if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                        ms < best_ms) {
                        best_ms = ms;
                        best = ch;
                    }
                }
                (void)hipStreamSynchronize(s);  // the tables stay alive until the launches ran
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

template <int C, bool Z3, typename LD>
static void plan_iso_group(const std::vector<BlurReq>& reqs, const std::vector<int>& idx, Blob& blob,
                           std::vector<Launch>& out, hipStream_t s) {
    BTuneChoice ch{32, iso_o8_default(C, Z3), 256};
    bool tune = false;
    BTuneKey key{};
    if (autotune_on()) {
        int64_t vox = 0;
        for (int i : idx) vox += reqs[i].g.n();
        int dev = 0;
        (void)hipGetDevice(&dev);
        key = BTuneKey{C, Z3 ? 1 : 0, ld_code<LD>(), dev, ilog2(vox), ilog2((int64_t)idx.size())};
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_btune.find(key);
        if (it != g_btune.end()) ch = it->second;
        else tune = t_tuning;
    }
    std::vector<IsoJob<C>> jobs;
    std::vector<int64_t> bst;
    const int64_t nb = iso_group_tables<C, Z3>(reqs, idx, ch.ty, ch.target, jobs, bst);
    const size_t jo = blob.put_vec(jobs), bo = blob.put_vec(bst);
    const int nj = (int)idx.size();
    // the group's requests travel with the launch only when it tunes
    std::shared_ptr<std::vector<BlurReq>> treq;
    if (tune) {
        treq = std::make_shared<std::vector<BlurReq>>();
        for (int i : idx) treq->push_back(reqs[i]);
    }
    out.push_back([=](const char* dev) -> int {
        if (treq) {
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone) {
                std::vector<int> all(treq->size());
                for (size_t k = 0; k < all.size(); ++k) all[k] = (int)k;
                const BTuneChoice best = tune_iso_group<C, Z3, LD>(*treq, all, ch, s);
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_btune[key] = best;
            }
        }
        launch_iso_group<C, Z3, LD>((const IsoJob<C>*)(dev + jo), (const int64_t*)(dev + bo), nj, nb, ch, s);
        CLE_LAUNCHED();
        return CLE_OK;
    });
}

template <int C>
static void plan_iso_c(bool z3, int dt, const std::vector<BlurReq>& reqs, const std::vector<int>& idx, Blob& blob,
                       std::vector<Launch>& out, hipStream_t s) {
    if (dt == CLE_U16) {
        z3 ? plan_iso_group<C, true, LdU16>(reqs, idx, blob, out, s) : plan_iso_group<C, false, LdU16>(reqs, idx, blob, out, s);
    } else if (dt == CLE_U8) {
        z3 ? plan_iso_group<C, true, LdU8>(reqs, idx, blob, out, s) : plan_iso_group<C, false, LdU8>(reqs, idx, blob, out, s);
    } else {
        z3 ? plan_iso_group<C, true, LdF32>(reqs, idx, blob, out, s) : plan_iso_group<C, false, LdF32>(reqs, idx, blob, out, s);
    }
}

int plan_blur_batch(const std::vector<BlurReq>& reqs, Blob& blob, std::vector<Launch>& out, hipStream_t s) {
    std::vector<std::vector<int>> groups(2 * 3 * (ISO_MAXC + 1));
    auto key = [](int c, bool z3, int dt) { return (c * 2 + (z3 ? 1 : 0)) * 3 + (dt == CLE_U16 ? 0 : dt == CLE_U8 ? 1 : 2); };
    for (size_t i = 0; i < reqs.size(); ++i) {
        const BlurReq& r = reqs[i];
        if (r.g.n() == 0) continue;
        Taps t;
        const bool typed = r.in_dt == CLE_U16 || r.in_dt == CLE_U8 || r.in_dt == CLE_F32;
        if (r.sigma > 0 && typed) {
            gauss_taps(r.sigma, &t);
            if (t.c >= 1 && t.c <= ISO_MAXC && iso_slice_ok(r.g)) {
                groups[key(t.c, r.g.ndim == 3, r.in_dt)].push_back((int)i);
                continue;
            }
        }
        out.push_back([=](const char*) -> int {
            cle_buf sb{}, db{};
            sb.data = (void*)r.in;
            sb.dtype = r.in_dt;
            db.data = r.out;
            db.dtype = CLE_F32;
            const int nd = r.g.ndim;
            const int64_t shp[3] = {r.g.D, r.g.H, r.g.W};
            sb.ndim = db.ndim = nd;
            for (int a = 0; a < nd; ++a) sb.shape[a] = db.shape[a] = shp[3 - nd + a];
            const float sg[3] = {r.sigma, r.sigma, r.sigma};
            return blur_impl(&sb, &db, sg, s);
        });
    }
    for (int c = 1; c <= ISO_MAXC; ++c)
        for (int z = 0; z < 2; ++z)
            for (int d = 0; d < 3; ++d) {
                const std::vector<int>& idx = groups[(c * 2 + z) * 3 + d];
                if (idx.empty()) continue;
                const int dt = d == 0 ? CLE_U16 : d == 1 ? CLE_U8 : CLE_F32;
                switch (c) {
#define ISO_PLAN(n) case n: plan_iso_c<n>(z == 1, dt, reqs, idx, blob, out, s); break;
                    ISO_PLAN(1) ISO_PLAN(2) ISO_PLAN(3) ISO_PLAN(4) ISO_PLAN(5) ISO_PLAN(6) ISO_PLAN(7)
                    ISO_PLAN(8) ISO_PLAN(9) ISO_PLAN(10) ISO_PLAN(11) ISO_PLAN(12) ISO_PLAN(13) ISO_PLAN(14)
                    ISO_PLAN(15) ISO_PLAN(16) ISO_PLAN(17) ISO_PLAN(18) ISO_PLAN(19) ISO_PLAN(20)
#undef ISO_PLAN
                    default: break;
                }
            }
    return CLE_OK;
}

int blur_impl(const cle_buf* src, const cle_buf* dst, const float sig[3], hipStream_t s) {
    return blur_impl_minmax(src, dst, sig, nullptr, nullptr, s);
}

int blur_impl_minmax(const cle_buf* src, const cle_buf* dst, const float sig[3], uint32_t* mm_words, int* folded,
                     hipStream_t s) {
    if (folded) *folded = 0;
    Taps tx, ty, tz;
    const Taps* taps[3] = {nullptr, nullptr, nullptr};
    Taps* store[3] = {&tx, &ty, &tz};
    const int nd = src->ndim;
    for (int a = 0; a < 3; ++a) {
        if (a >= nd) break;  // sigma_z ignored for 2D, y and z for 1D
        if (sig[a] > 0) {
            gauss_taps(sig[a], store[a]);
            if (store[a]->c > MAXC) return fail(CLE_E_VALUE, "sigma %g too large (max %d taps)", sig[a], 2 * MAXC + 1);
            taps[a] = store[a];
        }
    }
    return separable_impl(src, dst, taps, s, mm_words, folded);
}

void blur_tuning(bool on) { t_tuning = on; }

// Isotropic blurs of many images into float32 destinations, planned as in the batched
// pipeline (one launch per group); synchronises the stream.
static int blur_batch(const cle_buf* srcs, const cle_buf* dsts, const float* sig, int n, bool tune, hipStream_t s) {
    if (n < 0) return fail(CLE_E_VALUE, "n_images must be >= 0");
    if (n == 0) return CLE_OK;
    if (!srcs || !dsts || !sig) return fail(CLE_E_VALUE, "null batch argument");
    std::vector<BlurReq> reqs((size_t)n);
    for (int j = 0; j < n; ++j) {
        Geom g, gd;
        CLE_TRY(check_buf(&srcs[j], "source", &g));
        CLE_TRY(check_buf(&dsts[j], "destination", &gd));
        CLE_TRY(check_same_shape(g, gd, "source", "destination"));
        if (dsts[j].dtype != CLE_F32) return fail(CLE_E_TYPE, "destination %d: batched blurs write float32", j);
        if (!(sig[j] >= 0.f)) return fail(CLE_E_VALUE, "sigma of image %d must be >= 0", j);
        if (g.ndim == 1) return fail(CLE_E_SHAPE, "image %d: batched blurs take 2D and 3D images", j);
        reqs[j] = BlurReq{srcs[j].data, srcs[j].dtype, (float*)dsts[j].data, g, sig[j]};
    }
    Blob blob;
    std::vector<Launch> L;
    CLE_TRY(plan_blur_batch(reqs, blob, L, s));
    Scratch dev;
    CLE_TRY(dev.alloc(blob.h.size(), s));
    CLE_HIP(hipMemcpyAsync(dev.p, blob.h.data(), blob.h.size(), hipMemcpyHostToDevice, s));
    t_tuning = tune;
    int rc = CLE_OK;
    for (auto& l : L)
        if (rc == CLE_OK) rc = l(dev.get<char>());
    t_tuning = false;
    CLE_TRY(rc);
    CLE_HIP(hipStreamSynchronize(s));  // the host tables and the scratch outlive the work
    return CLE_OK;
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_set_autotune(int on) {
    g_autotune.store(on ? 1 : 0);
    if (!on) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune.clear();
        g_btune.clear();
    }
    return CLE_OK;
}

int cle_gaussian_blur_batch(const cle_buf* srcs, const cle_buf* dsts, const float* sigmas, int n_images,
                            void* stream) {
    return blur_batch(srcs, dsts, sigmas, n_images, false, (hipStream_t)stream);
}

int cle_autotune_gaussian_blur_batch(const cle_buf* srcs, const cle_buf* dsts, const float* sigmas, int n_images,
                                     void* stream) {
    if (!autotune_on()) return blur_batch(srcs, dsts, sigmas, n_images, false, (hipStream_t)stream);
    return blur_batch(srcs, dsts, sigmas, n_images, true, (hipStream_t)stream);
}

int cle_autotune_gaussian_blur(const cle_buf* src, const cle_buf* dst, float sigma_x, float sigma_y, float sigma_z,
                               void* stream) {
    if (!autotune_on()) return CLE_OK;
    const float sig[3] = {sigma_x, sigma_y, sigma_z};
    t_tuning = true;
    const int rc = blur_impl(src, dst, sig, (hipStream_t)stream);
    t_tuning = false;
    return rc;
}

int cle_gaussian_blur(const cle_buf* src, const cle_buf* dst, float sigma_x, float sigma_y, float sigma_z,
                      void* stream) {
    const float sig[3] = {sigma_x, sigma_y, sigma_z};
    return blur_impl(src, dst, sig, (hipStream_t)stream);
}

int cle_mean_box(const cle_buf* src, const cle_buf* dst, int radius_x, int radius_y, int radius_z, void* stream) {
    Taps tx, ty, tz;
    const Taps* taps[3] = {nullptr, nullptr, nullptr};
    Taps* store[3] = {&tx, &ty, &tz};
    const int r[3] = {radius_x, radius_y, radius_z};
    for (int a = 0; a < 3 && a < src->ndim; ++a) {
        if (r[a] > 0) {
            if (r[a] > MAXC) return fail(CLE_E_VALUE, "radius %d too large", r[a]);
            box_taps(r[a], store[a]);
            taps[a] = store[a];
        }
    }
    return separable_impl(src, dst, taps, (hipStream_t)stream);
}

}  // extern "C"
