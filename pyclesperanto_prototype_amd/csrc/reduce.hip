// Reductions of the Otsu stage: global min/max, sum, the 256-bin histogram and the
// Otsu threshold search, all on the device (no host round trip between them).
//
// Reference semantics (pyclesperanto_prototype/...):
//   minimum/maximum_of_all_pixels  _tier2/_minimum_of_all_pixels.py:6-59 (exact)
//   histogram                      _tier3/_histogram.py:25-130: one partial histogram
//       per row y over all z and x, bin = convert_uint_sat(((v-min)*(bins-1))/(max-min)+0.5)
//       in float32, partials stored as float32 and summed over y in y order
//   threshold_otsu                 _tier9/_threshold_otsu.py:38-54 + the vendored
//       scikit-image routine :64-106 (float64 cumulative sums, first argmax)
//
// Histogram layout: counts are exact uint32 per (row y, bin) -- hist2d[H][bins] --
// built from per-wavefront LDS sub-histograms (block = one row y over a chunk of z
// slices), then each bin is summed over y in float32 in y order, which reproduces
// the reference's float32 totals bit for bit (also above 2^24 per bin).
#include <float.h>

#include "common.hpp"

namespace cle {

constexpr int RB = 256;

template <typename LD>
__global__ __launch_bounds__(RB) void k_minmax(LD in, int64_t n, uint32_t* out) {
    float mn = INFINITY, mx = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
        const float v = in(i);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off));
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    __shared__ float smn[RB / 64], smx[RB / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { smn[w] = mn; smx[w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < RB / 64; ++k) { mn = fminf(mn, smn[k]); mx = fmaxf(mx, smx[k]); }
        atomicMin(out, f2o(mn));
        atomicMax(out + 1, f2o(mx));
    }
}

// Element k of a 16-byte vector holding elements of type T, converted to float by
// explicit shifts / bit casts (no type punning through pointers).
template <typename T> __device__ __forceinline__ float vec_elem(const uint4& q, int k);
template <> __device__ __forceinline__ float vec_elem<uint8_t>(const uint4& q, int k) {
    const uint32_t w = k < 4 ? q.x : (k < 8 ? q.y : (k < 12 ? q.z : q.w));
    return (float)((w >> (8 * (k & 3))) & 0xFFu);
}
template <> __device__ __forceinline__ float vec_elem<uint16_t>(const uint4& q, int k) {
    const uint32_t w = k < 2 ? q.x : (k < 4 ? q.y : (k < 6 ? q.z : q.w));
    return (float)((w >> (16 * (k & 1))) & 0xFFFFu);
}
template <> __device__ __forceinline__ float vec_elem<uint32_t>(const uint4& q, int k) {
    const uint32_t w = k == 0 ? q.x : (k == 1 ? q.y : (k == 2 ? q.z : q.w));
    return (float)w;
}
template <> __device__ __forceinline__ float vec_elem<int32_t>(const uint4& q, int k) {
    const uint32_t w = k == 0 ? q.x : (k == 1 ? q.y : (k == 2 ? q.z : q.w));
    return (float)(int32_t)w;
}
template <> __device__ __forceinline__ float vec_elem<float>(const uint4& q, int k) {
    const uint32_t w = k == 0 ? q.x : (k == 1 ? q.y : (k == 2 ? q.z : q.w));
    return __uint_as_float(w);
}

// min/max with 16-byte vector loads (T = element type)
template <typename T>
__global__ __launch_bounds__(RB) void k_minmax_vec(const T* __restrict__ p, int64_t n, uint32_t* out) {
    constexpr int E = 16 / sizeof(T);
    float mn = INFINITY, mx = -INFINITY;
    const int64_t nv = n / E;
    const uint4* pv = (const uint4*)p;
    const int64_t step = (int64_t)gridDim.x * RB;
    int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    // four independent 16-byte loads in flight per thread
    for (; i + 3 * step < nv; i += 4 * step) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = pv[i + u * step];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const float v = vec_elem<T>(q[u], k);
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
            }
    }
    for (; i < nv; i += step) {
        const uint4 q = pv[i];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const float v = vec_elem<T>(q, k);
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
    }
    // tail
    for (int64_t i = nv * E + (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
        const float v = (float)p[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off));
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    __shared__ float smn[RB / 64], smx[RB / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { smn[w] = mn; smx[w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < RB / 64; ++k) { mn = fminf(mn, smn[k]); mx = fmaxf(mx, smx[k]); }
        atomicMin(out, f2o(mn));
        atomicMax(out + 1, f2o(mx));
    }
}

__global__ void k_minmax_init(uint32_t* o) {
    o[0] = 0xFFFFFFFFu;
    o[1] = 0u;
}
__global__ void k_minmax_fin(const uint32_t* o, float* out2) {
    out2[0] = o2f(o[0]);
    out2[1] = o2f(o[1]);
}

struct LdF32R {
    const float* p;
    __device__ __forceinline__ float operator()(int64_t i) const { return p[i]; }
};
struct LdAnyR {
    DBuf b;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(b, i); }
};

int minmax_impl(DBuf src, int64_t n, float* out2, hipStream_t s) {
    Scratch tmp;
    CLE_TRY(tmp.alloc(8, s));
    k_minmax_init<<<1, 1, 0, s>>>(tmp.get<uint32_t>());
    unsigned g = grid_for((size_t)n, RB * 16);
    if (g > 2048) g = 2048;
    const bool aligned = ((uintptr_t)src.p & 15) == 0;
    if (n > 0) {
        uint32_t* o = tmp.get<uint32_t>();
        if (aligned && src.dt == CLE_F32) k_minmax_vec<float><<<g, RB, 0, s>>>((const float*)src.p, n, o);
        else if (aligned && src.dt == CLE_I32) k_minmax_vec<int32_t><<<g, RB, 0, s>>>((const int32_t*)src.p, n, o);
        else if (aligned && src.dt == CLE_U32) k_minmax_vec<uint32_t><<<g, RB, 0, s>>>((const uint32_t*)src.p, n, o);
        else if (aligned && src.dt == CLE_U16) k_minmax_vec<uint16_t><<<g, RB, 0, s>>>((const uint16_t*)src.p, n, o);
        else if (aligned && src.dt == CLE_U8) k_minmax_vec<uint8_t><<<g, RB, 0, s>>>((const uint8_t*)src.p, n, o);
        else k_minmax<<<g, RB, 0, s>>>(LdAnyR{src}, n, o);
    }
    k_minmax_fin<<<1, 1, 0, s>>>(tmp.get<uint32_t>(), out2);
    CLE_LAUNCHED();
    return CLE_OK;
}

__global__ __launch_bounds__(RB) void k_sum(DBuf in, int64_t n, double* out) {
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) acc += ld(in, i);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double sm[RB / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < RB / 64; ++k) acc += sm[k];
        atomicAdd(out, acc);
    }
}

// ----------------------------------------------------------------- histogram
constexpr int HB = 256;                // threads per block
constexpr int HWAVES = HB / 64;
constexpr int HIST_LDS_MAX_BINS = 480;  // LDS replicas: 32 x (bins | 1) x 4 B <= 64 KB

// The float32 quotient gets the literal 0.5 added in float64 (an unsuffixed OpenCL
// literal is double on fp64 devices), then convert_uint_sat truncates.
__device__ __forceinline__ uint32_t bin_of(float v, float mn, float range, int bins) {
#pragma clang fp contract(off)
    const float q = ((v - mn) * (float)(bins - 1)) / range;
    const double t = (double)q + 0.5;
    // convert_uint_sat: NaN -> 0, negative -> 0, truncation toward zero
    if (!(t > 0.0)) return 0u;
    if (t >= (double)bins) return (uint32_t)(bins - 1);
    return (uint32_t)t;
}

// Fast bin: bin_of needs floor(q + 0.5) of the correctly rounded float32 quotient q.
// qa = a * (1 / range) is within 2^-23 relative of q (two roundings); t = qa + 0.5f
// adds at most half an ulp of t.  For t < 2048 the total error stays below 4e-4, so
// when t lies more than 2^-10 from an integer floor(t) = floor(q + 0.5) and the
// division and float64 add are skipped; otherwise (near a bin boundary, NaN, huge or
// tiny ranges) the exact bin_of decides.
struct BinFast {
    float mn, range, rinv, fb1;
    int bins;
    bool ok;  // range is a normal float with a finite reciprocal
};

__device__ __forceinline__ BinFast bin_fast_init(float mn, float range, int bins) {
    BinFast c;
    c.mn = mn;
    c.range = range;
    c.rinv = 1.0f / range;
    c.fb1 = (float)(bins - 1);
    c.bins = bins;
    c.ok = range >= FLT_MIN && c.rinv <= FLT_MAX && bins + 2 <= 2048;
    return c;
}

__device__ __forceinline__ uint32_t bin_fast(float v, const BinFast& c) {
#pragma clang fp contract(off)
    if (c.ok) {
        const float a = (v - c.mn) * c.fb1;
        const float t = a * c.rinv + 0.5f;
        const float k = floorf(t);
        const float d = t - k;
        if (d > 0.0009765625f && d < 0.9990234375f && t < 2048.f)  // NaN fails these tests
            return k <= 0.f ? 0u : (k >= c.fb1 ? (uint32_t)(c.bins - 1) : (uint32_t)k);
        if (t >= 2048.f) return (uint32_t)(c.bins - 1);  // far above the last bin
    }
    return bin_of(v, c.mn, c.range, c.bins);
}

// One block = one row y over slices [z0, z0 + zc) (all slices when the image has
// enough rows to fill the chip: then the block owns its row's counts and stores them;
// otherwise the z chunks of a row add theirs atomically).  Thread t takes the float4
// column x = 4t (+ 4*256 k) of every slice -- coalesced across the block -- eight slices
// per iteration with the loads issued together, and counts runs of equal bins in
// registers across all its voxels (blurred backgrounds fall into one or two bins),
// flushing a run into one of HREP LDS replicas per wavefront (lane % HREP; replicas
// padded to an odd stride so equal bins of different replicas sit in different LDS
// banks).  Rows whose length is not a multiple of 4 or whose start is not 16-byte
// aligned take the scalar loads.
constexpr int HREP = 4;

template <typename LD>
__global__ __launch_bounds__(HB) void k_hist_rows(LD in, const float* f4base, Geom g, int bins, int zc, int zchunks,
                                                 float mn_arg, float mx_arg, const float* mm, uint32_t* hist2d) {
    extern __shared__ uint32_t sh[];
    const int stride = bins | 1;
    const int64_t y = blockIdx.x / zchunks;
    const int64_t z0 = (int64_t)(blockIdx.x % zchunks) * zc;
    const float mn = mm ? mm[0] : mn_arg;
    const float mx = mm ? mm[1] : mx_arg;
    const BinFast bc = bin_fast_init(mn, mx - mn, bins);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nsub = HWAVES * HREP;
    for (int i = threadIdx.x; i < stride * nsub; i += HB) sh[i] = 0;
    __syncthreads();
    uint32_t* mine = sh + (wave * HREP + (lane % HREP)) * stride;
    const int64_t z1 = (z0 + zc < g.D) ? z0 + zc : g.D;
    uint32_t cur = 0xFFFFFFFFu, run = 0;
    auto add = [&](float v) {
        const uint32_t b = bin_fast(v, bc);
        if (b == cur) {
            ++run;
        } else {
            if (run) atomicAdd(&mine[cur], run);
            cur = b;
            run = 1;
        }
    };
    const bool vec = f4base != nullptr && (g.W & 3) == 0;
    const int64_t sstride = g.H * g.W;
    if (vec) {
        const int64_t nq = g.W / 4;
        for (int64_t q = threadIdx.x; q < nq; q += HB) {
            const float4* col = (const float4*)(f4base + (z0 * g.H + y) * g.W) + q;
            const int64_t s4 = sstride / 4;
            int64_t z = z0;
            for (; z + 8 <= z1; z += 8) {
                float4 v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = col[j * s4];
                col += 8 * s4;
#pragma unroll
                for (int j = 0; j < 8; ++j) { add(v[j].x); add(v[j].y); add(v[j].z); add(v[j].w); }
            }
            for (; z < z1; ++z) {
                const float4 v = col[0];
                col += s4;
                add(v.x); add(v.y); add(v.z); add(v.w);
            }
        }
    } else {
        for (int64_t z = z0; z < z1; ++z) {
            const int64_t base = (z * g.H + y) * g.W;
            for (int64_t x = threadIdx.x; x < g.W; x += HB) add(in(base + x));
        }
    }
    if (run) atomicAdd(&mine[cur], run);
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += HB) {
        uint32_t c = 0;
        for (int w = 0; w < nsub; ++w) c += sh[w * stride + b];
        if (zchunks == 1) hist2d[y * bins + b] = c;
        else if (c) atomicAdd(&hist2d[y * bins + b], c);
    }
}

// fallback for very many bins: global atomics
template <typename LD>
__global__ void k_hist_global(LD in, Geom g, int bins, float mn_arg, float mx_arg, const float* mm, uint32_t* hist2d) {
    const float mn = mm ? mm[0] : mn_arg;
    const float mx = mm ? mm[1] : mx_arg;
    const float range = mx - mn;
    const int64_t n = g.n();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = (i / g.W) % g.H;
        atomicAdd(&hist2d[y * bins + bin_of(in(i), mn, range, bins)], 1u);
    }
}

// total[b] = sum over y (in y order, float32) of float(hist2d[y][b]).  One wavefront
// per bin: the lanes stage the bin's column in LDS in chunks, lane 0 adds it in order.
constexpr int RS_CHUNK = 2048;
__global__ __launch_bounds__(64) void k_hist_rowsum(const uint32_t* hist2d, int64_t H, int bins, DBuf out) {
    __shared__ float col[RS_CHUNK];
    const int b = blockIdx.x;
    float acc = 0.f;
    for (int64_t y0 = 0; y0 < H; y0 += RS_CHUNK) {
        const int n = (int)(H - y0 < RS_CHUNK ? H - y0 : RS_CHUNK);
        for (int k = threadIdx.x; k < n; k += 64) col[k] = (float)hist2d[(y0 + k) * bins + b];
        __syncthreads();
        if (threadIdx.x == 0) {
            int k = 0;
            for (; k + 8 <= n; k += 8) {
                const float4 a = *(const float4*)(col + k);
                const float4 c = *(const float4*)(col + k + 4);
                acc += a.x; acc += a.y; acc += a.z; acc += a.w;
                acc += c.x; acc += c.y; acc += c.z; acc += c.w;
            }
            for (; k < n; ++k) acc += col[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) st(out, b, acc);
}

template <typename LD>
static int hist_launch(LD in, const float* f4base, const Geom& g, int bins, float mn, float mx, const float* mm,
                       uint32_t* h2, hipStream_t s) {
    if (bins <= HIST_LDS_MAX_BINS) {
        // rows split into z chunks so that the grid holds at least 4096 blocks (the
        // chunks of a row add their exact integer counts atomically; short images:
        // chunks of about 16 K voxels)
        int zc = (int)g.D;
        if (g.H >= 1024) {
            const int64_t chunks = (4096 + g.H - 1) / g.H;
            zc = (int)((g.D + chunks - 1) / chunks);
            if (zc < 8 && g.D >= 8) zc = 8;
        } else {
            const int64_t per_block = 16384;
            zc = (int)(per_block / (g.W > 0 ? g.W : 1));
            if (zc < 1) zc = 1;
            if (zc > g.D) zc = (int)g.D;
        }
        const int zchunks = (int)((g.D + zc - 1) / zc);
        const size_t lds = (size_t)(bins | 1) * HWAVES * HREP * sizeof(uint32_t);
        k_hist_rows<<<(unsigned)(g.H * zchunks), HB, lds, s>>>(in, f4base, g, bins, zc, zchunks, mn, mx, mm, h2);
    } else {
        unsigned gr = grid_for((size_t)g.n(), 256);
        if (gr > 65536) gr = 65536;
        k_hist_global<<<gr, 256, 0, s>>>(in, g, bins, mn, mx, mm, h2);
    }
    CLE_LAUNCHED();
    return CLE_OK;
}

static int histogram_impl(DBuf src, const Geom& g, int bins, float mn, float mx, const float* mm, DBuf out,
                          hipStream_t s) {
    Scratch h2;
    const size_t cells = (size_t)g.H * bins;
    CLE_TRY(h2.alloc(cells * sizeof(uint32_t), s));
    CLE_TRY(fill_u32(h2.get<uint32_t>(), 0u, cells, s));
    if (g.n() > 0) {
        const bool f4 = src.dt == CLE_F32 && ((uintptr_t)src.p & 15) == 0;
        if (src.dt == CLE_F32)
            CLE_TRY(hist_launch(LdF32R{(const float*)src.p}, f4 ? (const float*)src.p : nullptr, g, bins, mn, mx, mm,
                                h2.get<uint32_t>(), s));
        else CLE_TRY(hist_launch(LdAnyR{src}, nullptr, g, bins, mn, mx, mm, h2.get<uint32_t>(), s));
    }
    k_hist_rowsum<<<(unsigned)bins, 64, 0, s>>>(h2.get<uint32_t>(), g.H, bins, out);
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- Otsu
// Single block of 256 threads.  The histogram is staged in LDS; thread 0 runs the
// cumulative sums in the exact order numpy's cumsum uses (sequential, float64, no
// contraction), thread 64 the reversed ones; the class means (sum / weight, the same
// operands as a division inside the running loop), the 255 between-class variances and
// the first maximum (NaN counting as maximum, as numpy.argmax does: the first NaN if
// there is one, else the first index of the largest value) are formed in parallel.
struct OtsuLds {
    double bc[256], hs[256], w1[256], w2[256], m1[256], m2[256];
    double wv[4];
    int wi[4];
    int first_nan;
};

__device__ __forceinline__ void otsu_body(const float* hist, const float* mm, double* thr, OtsuLds& sm) {
#pragma clang fp contract(off)
    double* bc = sm.bc;
    double* hs = sm.hs;
    double* w1 = sm.w1;
    double* w2 = sm.w2;
    double* m1 = sm.m1;
    double* m2 = sm.m2;
    const int t = threadIdx.x;
    const float mnf = mm[0], mxf = mm[1];
    const float rng = mxf - mnf;  // float32 subtraction, as in the reference
    bc[t] = (double)t * (double)rng / 255.0 + (double)mnf;
    hs[t] = (double)hist[t];
    if (t == 0) sm.first_nan = 256;
    __syncthreads();
    if (t == 0) {
        double a = 0, b = 0;
        for (int i = 0; i < 256; ++i) {
            const double c = hs[i];
            a = a + c;
            b = b + c * bc[i];
            w1[i] = a;
            m1[i] = b;  // running sum; divided below
        }
    } else if (t == 64) {
        double a = 0, b = 0;
        for (int i = 255; i >= 0; --i) {
            const double c = hs[i];
            a = a + c;
            b = b + c * bc[i];
            w2[i] = a;
            m2[i] = b;
        }
    }
    __syncthreads();
    m1[t] = m1[t] / w1[t];
    m2[t] = m2[t] / w2[t];
    __syncthreads();
    double v = -1.0;  // never the maximum: every variance is >= 0 or NaN
    if (t < 255) {
        const double d = m1[t] - m2[t + 1];
        const double vt = (w1[t] * w2[t + 1]) * (d * d);
        if (vt != vt) atomicMin(&sm.first_nan, t);
        else v = vt;
    }
    // largest value, smallest index among equals: within each wavefront, then across
    int vi = t;
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oi = __shfl_xor(vi, o);
        if (ov > v || (ov == v && oi < vi)) { v = ov; vi = oi; }
    }
    if ((t & 63) == 0) { sm.wv[t >> 6] = v; sm.wi[t >> 6] = vi; }
    __syncthreads();
    if (t == 0) {
        int best = sm.wi[0];
        double bv = sm.wv[0];
        for (int w = 1; w < 4; ++w)
            if (sm.wv[w] > bv) { bv = sm.wv[w]; best = sm.wi[w]; }  // waves in index order: ties keep the first
        if (sm.first_nan < 255) best = sm.first_nan;
        thr[0] = bc[best];
    }
}

__global__ __launch_bounds__(256) void k_otsu(const float* hist, const float* mm, double* thr) {
    __shared__ OtsuLds sm;
    otsu_body(hist, mm, thr, sm);
}

// dst = src > float32(threshold); float32 -> uint8 with 16-byte loads, 4-byte stores
__global__ __launch_bounds__(256) void k_threshold_f32_u8(const float* __restrict__ s, uint8_t* __restrict__ d,
                                                          const double* thr, int64_t n) {
    const float t = (float)thr[0];
    const int64_t nv = n / 4;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = ((const float4*)s)[i];
        const uint32_t o = (v.x > t ? 1u : 0u) | (v.y > t ? 1u << 8 : 0u) | (v.z > t ? 1u << 16 : 0u) |
                           (v.w > t ? 1u << 24 : 0u);
        ((uint32_t*)d)[i] = o;
    }
    for (int64_t i = nv * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i] > t ? 1 : 0;
}

__global__ void k_threshold(DBuf s, DBuf d, const double* thr, int64_t n) {
    const float t = (float)thr[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(d, i, ld(s, i) > t ? 1.f : 0.f);
}

// minmax_impl's 8 bytes are released before the histograms are taken
size_t otsu_scratch_bytes(const Geom& g) {
    return scratch_round(256 * sizeof(float)) + scratch_round((size_t)g.H * 256 * sizeof(uint32_t));
}

int otsu_impl(DBuf src, const Geom& g, float* minmax_dev, double* thr_dev, hipStream_t s) {
    CLE_TRY(minmax_impl(src, g.n(), minmax_dev, s));
    return otsu_after_minmax(src, g, minmax_dev, thr_dev, s);
}

void minmax_words_init(uint32_t* words, hipStream_t s) { k_minmax_init<<<1, 1, 0, s>>>(words); }
void minmax_words_finish(const uint32_t* words, float* out2, hipStream_t s) {
    k_minmax_fin<<<1, 1, 0, s>>>(words, out2);
}

int otsu_after_minmax(DBuf src, const Geom& g, float* minmax_dev, double* thr_dev, hipStream_t s) {
    Scratch hist;
    CLE_TRY(hist.alloc(256 * sizeof(float), s));
    CLE_TRY(histogram_impl(src, g, 256, 0.f, 0.f, minmax_dev, DBuf{hist.p, CLE_F32}, s));
    k_otsu<<<1, 256, 0, s>>>(hist.get<float>(), minmax_dev, thr_dev);
    CLE_LAUNCHED();
    return CLE_OK;
}

// ----------------------------------------------------------------- batched Otsu
// Per image of a batch: min/max, the 256-bin histogram and the threshold, each stage
// one launch over the BV-voxel blocks of all images.  An image below 2^24 voxels has
// every partial histogram sum below 2^24, so its float32 totals equal the exact
// integer counts whatever the summation order: the counts are accumulated as uint32
// (LDS per block, then atomics).  Larger images take otsu_impl (per-row partials in
// y order) one by one.
__global__ __launch_bounds__(256) void k_minmax_batch(const BImg* __restrict__ imgs,
                                                     const int64_t* __restrict__ vstart, int nimg,
                                                     const float* __restrict__ F, uint32_t* mmo) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(vstart, nimg, b);
    const BImg& im = imgs[j];
    if (im.big) return;
    const int64_t n = im.g.n();
    const int64_t i0 = (b - vstart[j]) * BV;
    const int64_t i1 = i0 + BV < n ? i0 + BV : n;
    const float* p = F + im.off;
    float mn = INFINITY, mx = -INFINITY;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const float v = p[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off));
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    __shared__ float smn[4], smx[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { smn[w] = mn; smx[w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) { mn = fminf(mn, smn[k]); mx = fmaxf(mx, smx[k]); }
        atomicMin(mmo + 2 * j, f2o(mn));
        atomicMax(mmo + 2 * j + 1, f2o(mx));
    }
}

__global__ void k_mm_init_batch(uint32_t* mmo, int nimg) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nimg; j += gridDim.x * blockDim.x) {
        mmo[2 * j] = 0xFFFFFFFFu;
        mmo[2 * j + 1] = 0u;
    }
}

__global__ void k_mm_fin_batch(const BImg* __restrict__ imgs, const uint32_t* mmo, float* mm, int nimg) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nimg; j += gridDim.x * blockDim.x) {
        if (imgs[j].big) continue;
        mm[2 * j] = o2f(mmo[2 * j]);
        mm[2 * j + 1] = o2f(mmo[2 * j + 1]);
    }
}

__global__ __launch_bounds__(256) void k_hist_batch(const BImg* __restrict__ imgs, const int64_t* __restrict__ vstart,
                                                   int nimg, const float* __restrict__ F, const float* mm,
                                                   uint32_t* cnt) {
    __shared__ uint32_t h[4 * 257];
    const int64_t b = blockIdx.x;
    const int j = find_seg(vstart, nimg, b);
    const BImg& im = imgs[j];
    if (im.big) return;
    for (int i = threadIdx.x; i < 4 * 257; i += 256) h[i] = 0;
    __syncthreads();
    const BinFast bc = bin_fast_init(mm[2 * j], mm[2 * j + 1] - mm[2 * j], 256);
    const int64_t n = im.g.n();
    const int64_t i0 = (b - vstart[j]) * BV;
    const int64_t i1 = i0 + BV < n ? i0 + BV : n;
    const float* p = F + im.off;
    uint32_t* mine = h + (threadIdx.x >> 6) * 257;
    uint32_t cur = 0xFFFFFFFFu, run = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const uint32_t bin = bin_fast(p[i], bc);
        if (bin == cur) {
            ++run;
        } else {
            if (run) atomicAdd(&mine[cur], run);
            cur = bin;
            run = 1;
        }
    }
    if (run) atomicAdd(&mine[cur], run);
    __syncthreads();
    const int t = threadIdx.x;
    const uint32_t c = h[t] + h[257 + t] + h[514 + t] + h[771 + t];
    if (c) atomicAdd(&cnt[256 * j + t], c);
}

__global__ __launch_bounds__(256) void k_otsu_batch(const BImg* __restrict__ imgs, const uint32_t* cnt,
                                                   float* histf, const float* mm, double* thr) {
    __shared__ OtsuLds sm;
    const int j = blockIdx.x;
    if (imgs[j].big) return;
    histf[256 * j + threadIdx.x] = (float)cnt[256 * j + threadIdx.x];
    __syncthreads();
    otsu_body(histf + 256 * j, mm + 2 * j, thr + j, sm);
}

int plan_otsu_batch(const std::vector<BImg>& imgs, size_t imgs_off, size_t vstart_off, int64_t vblocks,
                    const float* F, float* mm, double* thr, uint32_t* scratch_u32, Blob& blob,
                    std::vector<Launch>& out, hipStream_t s) {
    (void)blob;
    const int nimg = (int)imgs.size();
    // scratch_u32: 2 * nimg ordered min/max, then 256 * nimg counts, then 256 * nimg floats
    uint32_t* mmo = scratch_u32;
    uint32_t* cnt = scratch_u32 + 2 * (size_t)nimg;
    float* histf = (float*)(cnt + 256 * (size_t)nimg);
    out.push_back([=](const char* dev) -> int {
        const BImg* im = (const BImg*)(dev + imgs_off);
        const int64_t* vs = (const int64_t*)(dev + vstart_off);
        const unsigned gi = grid_for((size_t)nimg, 256);
        k_mm_init_batch<<<gi, 256, 0, s>>>(mmo, nimg);
        CLE_TRY(fill_u32(cnt, 0u, 256 * (size_t)nimg, s));
        k_minmax_batch<<<(unsigned)vblocks, 256, 0, s>>>(im, vs, nimg, F, mmo);
        k_mm_fin_batch<<<gi, 256, 0, s>>>(im, mmo, mm, nimg);
        k_hist_batch<<<(unsigned)vblocks, 256, 0, s>>>(im, vs, nimg, F, mm, cnt);
        k_otsu_batch<<<(unsigned)nimg, 256, 0, s>>>(im, cnt, histf, mm, thr);
        CLE_LAUNCHED();
        return CLE_OK;
    });
    for (int j = 0; j < nimg; ++j) {
        if (!imgs[j].big) continue;
        const BImg im = imgs[j];
        out.push_back([=](const char*) -> int {
            return otsu_impl(DBuf{(void*)(F + im.off), CLE_F32}, im.g, mm + 2 * j, thr + j, s);
        });
    }
    return CLE_OK;
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_minimum_maximum(const cle_buf* src, float* out_minmax, void* stream) {
    Geom g;
    CLE_TRY(check_buf(src, "source", &g));
    if (!out_minmax) return fail(CLE_E_VALUE, "null output");
    return minmax_impl(dbuf(src), g.n(), out_minmax, (hipStream_t)stream);
}

int cle_sum_of_all_pixels(const cle_buf* src, double* out_sum, void* stream) {
    Geom g;
    CLE_TRY(check_buf(src, "source", &g));
    hipStream_t s = (hipStream_t)stream;
    CLE_TRY(fill_f64(out_sum, 0.0, 1, s));
    if (!g.n()) return CLE_OK;
    unsigned gr = grid_for((size_t)g.n(), RB * 8);
    if (gr > 2048) gr = 2048;
    k_sum<<<gr, RB, 0, s>>>(dbuf(src), g.n(), out_sum);
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_histogram(const cle_buf* src, const cle_buf* dst, int num_bins, float mn, float mx,
                  const float* minmax_dev, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "destination", &gd));
    if (num_bins < 1) return fail(CLE_E_VALUE, "num_bins must be positive");
    if (gd.n() < num_bins) return fail(CLE_E_SHAPE, "destination holds %lld < %d bins", (long long)gd.n(), num_bins);
    return histogram_impl(dbuf(src), g, num_bins, mn, mx, minmax_dev, dbuf(dst), (hipStream_t)stream);
}

int cle_threshold_otsu(const cle_buf* src, const cle_buf* dst, double* out_threshold, void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "destination"));
    hipStream_t s = (hipStream_t)stream;
    if (!g.n()) return CLE_OK;
    Scratch mm, thr;
    CLE_TRY(mm.alloc(2 * sizeof(float), s));
    double* t = out_threshold;
    if (!t) {
        CLE_TRY(thr.alloc(sizeof(double), s));
        t = thr.get<double>();
    }
    CLE_TRY(otsu_impl(dbuf(src), g, mm.get<float>(), t, s));
    unsigned gr = grid_for((size_t)g.n(), 256 * 4);
    if (gr > 131072) gr = 131072;
    const bool vec = src->dtype == CLE_F32 && dst->dtype == CLE_U8 && ((uintptr_t)src->data & 15) == 0 &&
                     ((uintptr_t)dst->data & 3) == 0;
    if (vec) k_threshold_f32_u8<<<gr, 256, 0, s>>>((const float*)src->data, (uint8_t*)dst->data, t, g.n());
    else k_threshold<<<gr, 256, 0, s>>>(dbuf(src), dbuf(dst), t, g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

}  // extern "C"
