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
constexpr int VTX = 64;  // tile extent along x
constexpr int VPT = 4;   // voxels per thread (tile = 1024 voxels)

struct Tiles {
    int tz, ty;             // tile extents in z and y (tx = VTX)
    int64_t nz, ny, nx;     // tile counts
    __host__ __device__ int64_t count() const { return nz * ny * nx; }
};

static Tiles make_tiles(const Geom& g) {
    Tiles t;
    if (g.ndim == 3 && g.D > 1) { t.tz = 4; t.ty = 4; }
    else { t.tz = 1; t.ty = 16; }
    t.nz = (g.D + t.tz - 1) / t.tz;
    t.ny = (g.H + t.ty - 1) / t.ty;
    t.nx = (g.W + VTX - 1) / VTX;
    return t;
}

__device__ __forceinline__ int64_t clampi(int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// A = B = (mask - 1) + label, and the tiles holding a positive value are flagged as
// "changed" at steps -1 and -2 (slots 3 and 2 of the 4-slot ring).
__global__ __launch_bounds__(VB) void k_vor_init(DBuf mask, float* A, float* B, Geom g, Tiles T, uint8_t* tflags) {
    const int64_t tile = blockIdx.x;
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    const int64_t ntiles = T.count();
    bool pos = false;
    for (int k = 0; k < VPT; ++k) {
        const int li = threadIdx.x + k * VB;
        const int lx = li % VTX, ly = (li / VTX) % T.ty, lz = li / (VTX * T.ty);
        if (lz >= T.tz) continue;
        const int64_t x = tx * VTX + lx, y = ty * T.ty + ly, z = tz * T.tz + lz;
        if (x >= g.W || y >= g.H || z >= g.D) continue;
        const int64_t i = (z * g.H + y) * g.W + x;
        const float v = (ld(mask, i) - 1.f) + B[i];
        A[i] = v;
        B[i] = v;
        pos |= v > 0.f;
    }
    const int any = __syncthreads_or(pos ? 1 : 0);
    if (threadIdx.x == 0) {
        const uint8_t f = any ? 1 : 0;
        tflags[0 * ntiles + tile] = 0;
        tflags[1 * ntiles + tile] = 0;
        tflags[2 * ntiles + tile] = f;
        tflags[3 * ntiles + tile] = f;
    }
}

// one growth step t: src -> dst, box (t even) or diamond (t odd)
__global__ __launch_bounds__(VB) void k_vor_step(const float* __restrict__ src, float* __restrict__ dst, Geom g,
                                                Tiles T, int t, uint8_t* tflags, uint32_t* gflags, int ring) {
    if (t > 0 && gflags[(t - 1) % ring] == 0) return;  // the previous step changed nothing: done
    const int64_t tile = blockIdx.x;
    const int64_t ntiles = T.count();
    const int64_t tx = tile % T.nx, ty = (tile / T.nx) % T.ny, tz = tile / (T.nx * T.ny);
    if (threadIdx.x == 0) tflags[((t + 1) & 3) * ntiles + tile] = 0;
    // activity: any tile of the 3^d tile neighbourhood changed at t-1 or t-2
    const uint8_t* f1 = tflags + ((t + 3) & 3) * ntiles;
    const uint8_t* f2 = tflags + ((t + 2) & 3) * ntiles;
    bool act = false;
    if (threadIdx.x < 27) {
        const int dx = threadIdx.x % 3 - 1, dy = (threadIdx.x / 3) % 3 - 1, dz = threadIdx.x / 9 - 1;
        const int64_t qx = tx + dx, qy = ty + dy, qz = tz + dz;
        if (qx >= 0 && qx < T.nx && qy >= 0 && qy < T.ny && qz >= 0 && qz < T.nz) {
            const int64_t q = (qz * T.ny + qy) * T.nx + qx;
            act = f1[q] | f2[q];
        }
    }
    if (!__syncthreads_or(act ? 1 : 0)) return;

    const bool diamond = (t & 1) != 0;
    const int zr = g.ndim == 3 ? 1 : 0;
    bool changed = false;
    for (int k = 0; k < VPT; ++k) {
        const int li = threadIdx.x + k * VB;
        const int lx = li % VTX, ly = (li / VTX) % T.ty, lz = li / (VTX * T.ty);
        if (lz >= T.tz) continue;
        const int64_t x = tx * VTX + lx, y = ty * T.ty + ly, z = tz * T.tz + lz;
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
    }
    if (__syncthreads_or(changed ? 1 : 0) && threadIdx.x == 0) {
        tflags[(t & 3) * ntiles + tile] = 1;
        gflags[t % ring] = 1;
    }
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
    CLE_TRY(tf.alloc((size_t)T.count() * 4, s));
    k_vor_init<<<(unsigned)T.count(), VB, 0, s>>>(mask, A, B, g, T, tf.get<uint8_t>());
    CLE_LAUNCHED();
    CLE_TRY(grow(A, B, g, T, tf.get<uint8_t>(), n_iterations, s));
    k_mask_out<<<vg(n), 256, 0, s>>>(A, mask, dst, n);
    CLE_LAUNCHED();
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
