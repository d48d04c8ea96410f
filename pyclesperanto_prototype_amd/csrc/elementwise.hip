// Elementwise ops and single-step neighbourhood filters of the hot path.
//
// Reference semantics (pyclesperanto_prototype/...):
//   greater_constant       _tier1/_greater_constant.py:5-40      src > float32(c)
//   binary_and             _tier1/_binary_and.py:6-47            (a != 0) && (b != 0)
//   mask                   _tier1/_mask.py:5-45                  mask != 0 ? src : 0
//   add_image_and_scalar   _tier1/_add_image_and_scalar.py:5-43  src + s
//   add_images_weighted    _tier1/_add_images_weighted.py:5-50   a*fa + b*fb
//   set_non_zero_pixels_to_pixel_index  _tier1/_set_non_zero_pixels_to_pixel_index.py:5-28
//   nonzero_minimum_{box,diamond}       _tier1/_nonzero_minimum_box.py:5-43
//   onlyzero_overwrite_maximum_{box,diamond}  _tier1/_onlyzero_overwrite_maximum_box.py:5-38
//   detect_maxima_box      _tier1/_detect_maxima_box.py:5-48 + detect_maxima_{2d,3d}_x.cl
// All reads of neighbours clamp to the edge (_tier0/preamble.cl:33-49).
#include "common.hpp"

namespace cle {

#define GRID_STRIDE(i, n)                                                       \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (int64_t)gridDim.x * blockDim.x)

static unsigned ew_grid(int64_t n) {
    unsigned g = grid_for((size_t)n, 256);
    return g > 131072 ? 131072 : g;
}

__global__ void k_greater_constant(DBuf s, DBuf d, float c, int64_t n) {
    GRID_STRIDE(i, n) st(d, i, ld(s, i) > c ? 1.f : 0.f);
}

__global__ void k_binary_and(DBuf a, DBuf b, DBuf d, int64_t n) {
    GRID_STRIDE(i, n) st(d, i, (ld(a, i) != 0.f && ld(b, i) != 0.f) ? 1.f : 0.f);
}

__global__ void k_mask(DBuf s, DBuf m, DBuf d, int64_t n) {
    GRID_STRIDE(i, n) st(d, i, ld(m, i) != 0.f ? ld(s, i) : 0.f);
}

__global__ void k_add_scalar(DBuf s, DBuf d, float v, int64_t n) {
    GRID_STRIDE(i, n) st(d, i, ld(s, i) + v);
}

__global__ void k_add_weighted(DBuf a, DBuf b, DBuf d, float fa, float fb, int64_t n) {
    GRID_STRIDE(i, n) st(d, i, __fadd_rn(__fmul_rn(ld(a, i), fa), __fmul_rn(ld(b, i), fb)));
}

__global__ void k_pixel_index(DBuf s, DBuf d, Geom g, int offset) {
    const int64_t n = g.n();
    GRID_STRIDE(i, n) {
        const int64_t x = i % g.W, y = (i / g.W) % g.H, z = i / (g.W * g.H);
        const int64_t key = x * g.H * g.D + y * g.D + z + offset;
        st(d, i, ld(s, i) != 0.f ? (float)key : 0.f);
    }
}

__device__ __forceinline__ int64_t clampi(int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// One Jacobi step of the non-zero minimum filter; zero voxels are not written.
__global__ void k_nonzero_min(DBuf s, DBuf flag, DBuf d, Geom g, int diamond) {
    const int64_t n = g.n();
    const int zr = g.ndim == 3 ? 1 : 0;
    GRID_STRIDE(i, n) {
        const float v = ld(s, i);
        if (v == 0.f) continue;
        const int64_t x = i % g.W, y = (i / g.W) % g.H, z = i / (g.W * g.H);
        float best = v;
        for (int dz = -zr; dz <= zr; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (diamond && (abs(dx) + abs(dy) + abs(dz)) > 1) continue;
                    const int64_t q = (clampi(z + dz, g.D - 1) * g.H + clampi(y + dy, g.H - 1)) * g.W +
                                      clampi(x + dx, g.W - 1);
                    const float w = ld(s, q);
                    if (w != 0.f && w < best) best = w;
                }
        if (best != v) st(flag, 0, 1.f);
        st(d, i, best);
    }
}

// One growth step: non-zero voxels copied, zero voxels take the largest positive
// neighbour value (or stay 0).
__global__ void k_onlyzero_max(DBuf s, DBuf flag, DBuf d, Geom g, int diamond) {
    const int64_t n = g.n();
    const int zr = g.ndim == 3 ? 1 : 0;
    GRID_STRIDE(i, n) {
        const float v = ld(s, i);
        if (v != 0.f) {
            st(d, i, v);
            continue;
        }
        const int64_t x = i % g.W, y = (i / g.W) % g.H, z = i / (g.W * g.H);
        float best = 0.f;
        for (int dz = -zr; dz <= zr; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    if (diamond && (abs(dx) + abs(dy) + abs(dz)) > 1) continue;
                    const int64_t q = (clampi(z + dz, g.D - 1) * g.H + clampi(y + dy, g.H - 1)) * g.W +
                                      clampi(x + dx, g.W - 1);
                    const float w = ld(s, q);
                    if (w > best) best = w;
                }
        if (best != v) st(flag, 0, 1.f);
        st(d, i, best);
    }
}

// 3^d local maximum with the reference's visiting order: dx outer, dy, dz inner;
// neighbours with a negative coordinate are skipped, upper ones read clamped; the
// running maximum starts at v - 1 and a neighbour wins only if strictly greater.
template <typename LD>
__global__ __launch_bounds__(256) void k_maxima_t(LD in, DBuf d, Geom g) {
    __shared__ float ring[MX_LDS];
    // consecutive tiles (which share halo lines) on one XCD
    maxima_tile(in, g, xcd_tile(blockIdx.x, gridDim.x), ring, [&](int64_t i, bool m) { st(d, i, m ? 1.f : 0.f); });
}
// uint8 flags (the pipelines' spots)
template <typename LD>
__global__ __launch_bounds__(256) void k_maxima_u8(LD in, uint8_t* __restrict__ d, Geom g) {
    __shared__ float ring[MX_LDS];
    maxima_tile(in, g, xcd_tile(blockIdx.x, gridDim.x), ring, [&](int64_t i, bool m) { d[i] = m ? 1 : 0; });
}

// float32 -> uint8 flags for rows of a multiple of 4 voxels (16-byte aligned source,
// 4-byte aligned destination): four x-consecutive voxels per thread, the same closed
// form as maxima_tile.  A block's 256 threads form 16 rows of 16 quads: rows 1 .. 14
// are the tile's output rows (64 x 14 voxels), rows 0 and 15 its y halo; threads 0 ..
// 31 also carry one x-halo cell (column x0 - 1 for rows 0 .. 15, then column x0 + 64).
// A quad keeps its z neighbours in registers; C (z maximum) and Y (C over y - 1 .. y
// + 1) go through LDS once per slice as float4 rows, and a thread reads back only the
// Y of its left and right neighbour voxels.  Two barriers per slice; one buffer each:
// the C writes of a slice follow the previous slice's second barrier, the Y writes its
// own first one.  The first cell of a quad past the last column stands for column W,
// which the visit reads clamped to column W - 1.
constexpr int MQ_TY = 14, MQ_RP = 72;  // C / Y row pitch: x0 - 1 at 3, x0 at 4, x0 + 64 at 68

__host__ __device__ inline int64_t maxima_q4_tiles(const Geom& g) {
    const int64_t nzc = g.ndim == 3 ? (g.D + MX_ZC - 1) / MX_ZC : 1;
    return ((g.W + 63) / 64) * ((g.H + MQ_TY - 1) / MQ_TY) * nzc;
}

__device__ __forceinline__ float4 qmax4(float4 a, float4 b) {
    return make_float4(qmax(a.x, b.x), qmax(a.y, b.y), qmax(a.z, b.z), qmax(a.w, b.w));
}
__device__ __forceinline__ float4 quiet4(float4 a) {
    return make_float4(__builtin_canonicalizef(a.x), __builtin_canonicalizef(a.y), __builtin_canonicalizef(a.z),
                       __builtin_canonicalizef(a.w));
}
__device__ __forceinline__ float f4at(const float4& a, int i) { return i == 0 ? a.x : i == 1 ? a.y : i == 2 ? a.z : a.w; }

__global__ __launch_bounds__(256) void k_maxima_q4(const float* __restrict__ in, uint8_t* __restrict__ d, Geom g) {
    __shared__ __align__(16) float Cl[16 * MQ_RP];
    __shared__ __align__(16) float Yl[16 * MQ_RP];
    const int tid = threadIdx.x, r = tid >> 4, q = tid & 15;
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t ntx = (g.W + 63) / 64, nty = (g.H + MQ_TY - 1) / MQ_TY;
    const int64_t x0 = (t % ntx) * 64, y0 = ((t / ntx) % nty) * MQ_TY, zc = t / (ntx * nty);
    const bool d3 = g.ndim == 3;
    const int64_t z0 = d3 ? zc * MX_ZC : 0, z1 = d3 ? (z0 + MX_ZC < g.D ? z0 + MX_ZC : g.D) : 1;
    const uint32_t plane = (uint32_t)(g.H * g.W);
    const float qnan = __builtin_nanf("");
    const float4 qnan4 = make_float4(qnan, qnan, qnan, qnan);
    // own quad: row y0 - 1 + r, columns xq .. xq + 3 (rows before the first are never
    // visited: NaN; rows past the last read the last one)
    const int64_t yy = y0 - 1 + r, xq = x0 + 4 * q;
    const bool rlive = yy >= 0, past = xq >= g.W;
    const int64_t yc = yy < 0 ? 0 : (yy < g.H ? yy : g.H - 1);
    const uint32_t qoff = (uint32_t)(yc * g.W + (past ? g.W - 4 : xq));
    // x-halo cell
    const bool hx = tid < 32;
    const int hr = tid & 15, hcol = tid < 16 ? 3 : 68;
    const int64_t hy = y0 - 1 + hr, hxx = tid < 16 ? x0 - 1 : x0 + 64;
    const bool hlive = hx && hy >= 0 && hxx >= 0;
    const int64_t hyc = hy < 0 ? 0 : (hy < g.H ? hy : g.H - 1);
    const uint32_t hoff = (uint32_t)(hyc * g.W + (hxx < 0 ? 0 : (hxx < g.W ? hxx : g.W - 1)));
    // Plain loads (every address is in bounds), so that the compiler waits for each
    // only where it is used; cells that are never visited become NaN, and a quad past
    // the last column takes column W - 1 in its first cell (the x + 1 neighbour of
    // column W - 1, the only one of its cells that is read), when C is stored.
    auto load = [&](float4& v, float& h, int64_t zz) {
        const uint32_t zb = (uint32_t)zz * plane;
        v = *(const float4*)(in + (zb + qoff));
        h = in[zb + hoff];  // used by the x-halo threads only
    };
    // Register ring of five slices: slot (z - z0 + 1) % 5 holds slice z.  Slice z runs
    // with z - 1, z, z + 1 (quieted when z + 1 enters use) and issues the load of z + 3
    // into the slot of z - 2: two slices of work hide each load, and with the loop
    // unrolled over the five slots no loaded register is ever moved.
    float4 R[5];
    float Hh[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        R[j] = qnan4;
        Hh[j] = qnan;
    }
    if (d3 && z0 > 0) load(R[0], Hh[0], z0 - 1);
    load(R[1], Hh[1], z0);
    if (d3 && z0 + 1 < g.D) load(R[2], Hh[2], z0 + 1);
    if (d3 && z0 + 2 < g.D && z0 + 1 < z1) load(R[3], Hh[3], z0 + 2);
    R[0] = quiet4(R[0]);
    R[1] = quiet4(R[1]);
    Hh[0] = __builtin_canonicalizef(Hh[0]);
    Hh[1] = __builtin_canonicalizef(Hh[1]);
    const bool orow = r >= 1 && r <= MQ_TY;
    float* const Crow = Cl + r * MQ_RP + 4 + 4 * q;
    float* const Yrow = Yl + r * MQ_RP + 4 + 4 * q;
    const int64_t zmax = z1 < g.D ? z1 : g.D - 1;
    for (int64_t zb = z0; zb < z1; zb += 5) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int64_t z = zb + k;
            if (z < z1) {
                const int sm = k % 5, s0 = (k + 1) % 5, sp = (k + 2) % 5, sn = (k + 4) % 5;
                const bool below = d3 && z > 0, above = d3 && z + 1 < g.D;
                // slice z + 3, clamped to the last slice the chunk reads (a clamped load
                // re-reads that slice and is never used)
                load(R[sn], Hh[sn], z + 3 < zmax ? z + 3 : zmax);
                R[sp] = quiet4(R[sp]);
                Hh[sp] = __builtin_canonicalizef(Hh[sp]);
                const float4 pm = R[sm], p0 = R[s0], pp = R[sp];
                float4 c = qmax4(p0, p0);  // (a copy the compiler cannot fold into the loads)
                if (below) c = qmax4(c, pm);
                if (above) c = qmax4(c, pp);
                if (past) c.x = c.w;
                if (!rlive) c = qnan4;
                *(float4*)Crow = c;
                if (hx) {
                    float ch = qmax(Hh[s0], Hh[s0]);
                    if (below) ch = qmax(ch, Hh[sm]);
                    if (above) ch = qmax(ch, Hh[sp]);
                    Cl[hr * MQ_RP + hcol] = hlive ? ch : qnan;
                }
                __syncthreads();
                float4 cu = qnan4, cd = qnan4, y = qnan4;
                if (orow) {
                    cu = *(const float4*)(Crow - MQ_RP);
                    cd = *(const float4*)(Crow + MQ_RP);
                    y = qmax4(qmax4(cu, c), cd);
                    *(float4*)Yrow = y;
                }
                if (hx && hr >= 1 && hr <= MQ_TY) {
                    const float* cp = Cl + hr * MQ_RP + hcol;
                    Yl[hr * MQ_RP + hcol] = qmax(qmax(cp[-MQ_RP], cp[0]), cp[MQ_RP]);
                }
                __syncthreads();
                if (orow && !past && yy < g.H) {
                    const float yl = Yrow[-1], yr = Yrow[4];  // Y at x - 1 and at x + 4
                    uint32_t w = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float v = f4at(p0, i);
                        // the visit in closed form, as maxima_tile: B / A = the largest
                        // value visited before / after the centre
                        float bmax = qmax(i == 0 ? yl : f4at(y, i - 1), f4at(cu, i));
                        float amax = qmax(i == 3 ? yr : f4at(y, i + 1), f4at(cd, i));
                        if (below) bmax = qmax(bmax, f4at(pm, i));
                        if (d3) amax = qmax(amax, above ? f4at(pp, i) : v);
                        const float lm0 = v - 1.f;
                        const bool took = bmax > lm0;
                        const float lm1 = took ? bmax : lm0;
                        const bool centre = v > lm1;
                        const float lm2 = centre ? v : lm1;
                        const bool self = (centre || !took) && !(amax > lm2);
                        w |= (uint32_t)self << (8 * i);
                    }
                    *(uint32_t*)(d + ((uint32_t)z * plane + (uint32_t)(yy * g.W + xq))) = w;
                }
            }
        }
    }
}

struct LdAnyE {
    DBuf b;
    __device__ __forceinline__ float operator()(int64_t i) const { return ld(b, i); }
};
struct LdF32E {
    const float* p;
    __device__ __forceinline__ float operator()(int64_t i) const { return p[i]; }
};

int maxima_impl(DBuf src, DBuf dst, const Geom& g, hipStream_t s) {
    if (g.n() == 0) return CLE_OK;
    const unsigned grid = (unsigned)maxima_tiles(g);
    if (src.dt == CLE_F32 && dst.dt == CLE_U8 && (g.W & 3) == 0 && ((uintptr_t)src.p & 15) == 0 &&
        ((uintptr_t)dst.p & 3) == 0)
        k_maxima_q4<<<(unsigned)maxima_q4_tiles(g), 256, 0, s>>>((const float*)src.p, (uint8_t*)dst.p, g);
    else if (src.dt == CLE_F32 && dst.dt == CLE_U8)
        k_maxima_u8<<<grid, 256, 0, s>>>(LdF32E{(const float*)src.p}, (uint8_t*)dst.p, g);
    else if (src.dt == CLE_F32)
        k_maxima_t<<<grid, 256, 0, s>>>(LdF32E{(const float*)src.p}, dst, g);
    else
        k_maxima_t<<<grid, 256, 0, s>>>(LdAnyE{src}, dst, g);
    CLE_LAUNCHED();
    return CLE_OK;
}

static int unary_check(const cle_buf* a, const cle_buf* d, Geom* g) {
    Geom gd;
    CLE_TRY(check_buf(a, "source", g));
    CLE_TRY(check_buf(d, "destination", &gd));
    return check_same_shape(*g, gd, "source", "destination");
}

static int binary_check(const cle_buf* a, const cle_buf* b, const cle_buf* d, Geom* g) {
    Geom gb, gd;
    CLE_TRY(check_buf(a, "source1", g));
    CLE_TRY(check_buf(b, "source2", &gb));
    CLE_TRY(check_buf(d, "destination", &gd));
    CLE_TRY(check_same_shape(*g, gb, "source1", "source2"));
    return check_same_shape(*g, gd, "source1", "destination");
}

static int flag_check(const cle_buf* f) {
    Geom gf;
    CLE_TRY(check_buf(f, "flag", &gf));
    if (gf.n() < 1) return fail(CLE_E_SHAPE, "flag buffer is empty");
    return CLE_OK;
}

}  // namespace cle

using namespace cle;

extern "C" {

int cle_greater_constant(const cle_buf* src, const cle_buf* dst, float c, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    if (!g.n()) return CLE_OK;
    k_greater_constant<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(dst), c, g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_binary_and(const cle_buf* a, const cle_buf* b, const cle_buf* dst, void* stream) {
    Geom g;
    CLE_TRY(binary_check(a, b, dst, &g));
    if (!g.n()) return CLE_OK;
    k_binary_and<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(a), dbuf(b), dbuf(dst), g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_mask(const cle_buf* src, const cle_buf* m, const cle_buf* dst, void* stream) {
    Geom g;
    CLE_TRY(binary_check(src, m, dst, &g));
    if (!g.n()) return CLE_OK;
    k_mask<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(m), dbuf(dst), g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_add_image_and_scalar(const cle_buf* src, const cle_buf* dst, float scalar, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    if (!g.n()) return CLE_OK;
    k_add_scalar<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(dst), scalar, g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_add_images_weighted(const cle_buf* a, const cle_buf* b, const cle_buf* dst, float fa, float fb,
                            void* stream) {
    Geom g;
    CLE_TRY(binary_check(a, b, dst, &g));
    if (!g.n()) return CLE_OK;
    k_add_weighted<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(a), dbuf(b), dbuf(dst), fa, fb, g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_set_nonzero_pixels_to_pixel_index(const cle_buf* src, const cle_buf* dst, int offset, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    if (!g.n()) return CLE_OK;
    k_pixel_index<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(dst), g, offset);
    CLE_LAUNCHED();
    return CLE_OK;
}

static int nzmin(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, int diamond, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    CLE_TRY(flag_check(flag));
    if (src->data == dst->data) return fail(CLE_E_VALUE, "source and destination must differ");
    if (!g.n()) return CLE_OK;
    k_nonzero_min<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(flag), dbuf(dst), g, diamond);
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_nonzero_minimum_box(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream) {
    return nzmin(src, flag, dst, 0, stream);
}
int cle_nonzero_minimum_diamond(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream) {
    return nzmin(src, flag, dst, 1, stream);
}

static int ozmax(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, int diamond, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    CLE_TRY(flag_check(flag));
    if (src->data == dst->data) return fail(CLE_E_VALUE, "source and destination must differ");
    if (!g.n()) return CLE_OK;
    k_onlyzero_max<<<ew_grid(g.n()), 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(flag), dbuf(dst), g, diamond);
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_onlyzero_overwrite_maximum_box(const cle_buf* src, const cle_buf* flag, const cle_buf* dst, void* stream) {
    return ozmax(src, flag, dst, 0, stream);
}
int cle_onlyzero_overwrite_maximum_diamond(const cle_buf* src, const cle_buf* flag, const cle_buf* dst,
                                           void* stream) {
    return ozmax(src, flag, dst, 1, stream);
}

int cle_detect_maxima_box(const cle_buf* src, const cle_buf* dst, int rx, int ry, int rz, void* stream) {
    Geom g;
    CLE_TRY(unary_check(src, dst, &g));
    hipStream_t s = (hipStream_t)stream;
    if (!g.n()) return CLE_OK;
    if (rx > 0 || ry > 0 || rz > 0) {
        Scratch t;
        CLE_TRY(t.alloc((size_t)g.n() * 4, s));
        cle_buf tb = *src;
        tb.data = t.p;
        tb.dtype = CLE_F32;
        CLE_TRY(cle_mean_box(src, &tb, rx, ry, rz, stream));
        return maxima_impl(DBuf{t.p, CLE_F32}, dbuf(dst), g, s);
    }
    return maxima_impl(dbuf(src), dbuf(dst), g, s);
}

}  // extern "C"
