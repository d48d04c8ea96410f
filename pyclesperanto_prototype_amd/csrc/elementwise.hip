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
    if (src.dt == CLE_F32 && dst.dt == CLE_U8)
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
