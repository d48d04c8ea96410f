// Library runtime: error reporting, argument validation, the scratch pool, fills,
// and the set / copy entry points.
#include <stdarg.h>
#include <stdio.h>
#include <mutex>

#include "common.hpp"

namespace cle {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

bool dtype_supported(int dt) {
    switch (dt) {
        case CLE_U8: case CLE_I8: case CLE_U16: case CLE_I16:
        case CLE_U32: case CLE_I32: case CLE_F32: return true;
        default: return false;
    }
}

size_t dtype_size(int dt) {
    switch (dt) {
        case CLE_U8: case CLE_I8: return 1;
        case CLE_U16: case CLE_I16: return 2;
        case CLE_U32: case CLE_I32: case CLE_F32: return 4;
        case CLE_U64: case CLE_I64: case CLE_F64: return 8;
        default: return 0;
    }
}

int check_buf(const cle_buf* b, const char* name, Geom* g) {
    if (!b) return fail(CLE_E_VALUE, "%s: null buffer descriptor", name);
    if (!dtype_supported(b->dtype))
        return fail(CLE_E_TYPE, "%s: unsupported element type code %d", name, b->dtype);
    if (b->ndim < 1 || b->ndim > 3) return fail(CLE_E_SHAPE, "%s: ndim %d not in 1..3", name, b->ndim);
    int64_t s[3] = {1, 1, 1};
    for (int i = 0; i < b->ndim; ++i) {
        if (b->shape[i] < 0) return fail(CLE_E_SHAPE, "%s: negative extent", name);
        s[3 - b->ndim + i] = b->shape[i];
    }
    g->D = s[0]; g->H = s[1]; g->W = s[2];
    g->ndim = b->ndim;
    if (g->n() > 0 && !b->data) return fail(CLE_E_VALUE, "%s: null data pointer", name);
    if (g->n() >= (int64_t)0xFFFFFFFFLL)
        return fail(CLE_E_SHAPE, "%s: %lld voxels exceed the 2^32-1 limit", name, (long long)g->n());
    return CLE_OK;
}

int check_same_shape(const Geom& a, const Geom& b, const char* na, const char* nb) {
    if (a.D != b.D || a.H != b.H || a.W != b.W)
        return fail(CLE_E_SHAPE, "%s (%lld,%lld,%lld) and %s (%lld,%lld,%lld) differ in shape", na,
                    (long long)a.D, (long long)a.H, (long long)a.W, nb, (long long)b.D,
                    (long long)b.H, (long long)b.W);
    return CLE_OK;
}

// ----------------------------------------------------------------- pool
static std::once_flag g_pool_once;

static void pool_init() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = ~0ull;  // keep freed blocks cached in the pool
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
}

void* scratch_alloc(size_t bytes, hipStream_t s) {
    std::call_once(g_pool_once, pool_init);
    void* p = nullptr;
    if (hipMallocAsync(&p, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void scratch_free(void* p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
}

// ----------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static int fill_t(T* p, T v, size_t n, hipStream_t s) {
    if (n == 0) return CLE_OK;
    unsigned g = grid_for(n, 256);
    if (g > 65536) g = 65536;
    k_fill<T><<<g, 256, 0, s>>>(p, v, n);
    CLE_LAUNCHED();
    return CLE_OK;
}

int fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) { return fill_t(p, v, n, s); }
int fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t s) { return fill_t(p, v, n, s); }
int fill_f32(float* p, float v, size_t n, hipStream_t s) { return fill_t(p, v, n, s); }
int fill_f64(double* p, double v, size_t n, hipStream_t s) { return fill_t(p, v, n, s); }

__global__ void k_set(DBuf d, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) st(d, i, v);
}

__global__ void k_copy(DBuf s, DBuf d, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) st(d, i, ld(s, i));
}

}  // namespace cle

using namespace cle;

extern "C" {

const char* cle_version(void) { return "cle-mi355x 0.1.0 (gfx950)"; }

const char* cle_last_error(void) { return g_err; }

int cle_release_pool(void) {
    int dev = 0;
    CLE_HIP(hipGetDevice(&dev));
    CLE_HIP(hipDeviceSynchronize());
    hipMemPool_t pool;
    CLE_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    CLE_HIP(hipMemPoolTrimTo(pool, 0));
    return CLE_OK;
}

int cle_set(const cle_buf* dst, float value, void* stream) {
    Geom g;
    CLE_TRY(check_buf(dst, "dst", &g));
    if (g.n() == 0) return CLE_OK;
    unsigned grid = grid_for(g.n(), 256);
    if (grid > 65536) grid = 65536;
    k_set<<<grid, 256, 0, (hipStream_t)stream>>>(dbuf(dst), value, g.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

int cle_copy(const cle_buf* src, const cle_buf* dst, void* stream) {
    Geom gs, gd;
    CLE_TRY(check_buf(src, "src", &gs));
    CLE_TRY(check_buf(dst, "dst", &gd));
    if (gs.n() != gd.n()) return fail(CLE_E_SHAPE, "copy: element counts differ");
    if (gs.n() == 0) return CLE_OK;
    unsigned grid = grid_for(gs.n(), 256);
    if (grid > 65536) grid = 65536;
    k_copy<<<grid, 256, 0, (hipStream_t)stream>>>(dbuf(src), dbuf(dst), gs.n());
    CLE_LAUNCHED();
    return CLE_OK;
}

}  // extern "C"
