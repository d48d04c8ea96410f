// Shared definitions of the gfx950 kernels and the C-ABI layer: image geometry,
// typed loads/stores with the reference's saturating conversion, error reporting,
// and the stream-ordered scratch pool.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

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
        if (!p) return fail(CLE_E_OOM, "scratch allocation of %zu bytes failed", bytes);
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

inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : g);
}

// ----------------------------------------------------------------- typed access
// Saturating float -> integer conversion, then truncation toward zero
// (pyclesperanto_prototype/_tier0/preamble.cl:505-591).
template <typename T> __device__ __forceinline__ T sat_cast(float v, float lo, float hi) {
    if (v > hi) return (T)hi;
    if (v < lo) return (T)lo;
    if (v != v) return (T)0;
    return (T)v;
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
        case CLE_U8: ((uint8_t*)b.p)[i] = sat_cast<uint8_t>(v, 0.f, 255.f); break;
        case CLE_U16: ((uint16_t*)b.p)[i] = sat_cast<uint16_t>(v, 0.f, 65535.f); break;
        case CLE_U32: ((uint32_t*)b.p)[i] = sat_cast<uint32_t>(v, 0.f, 4294967295.f); break;
        case CLE_I32: ((int32_t*)b.p)[i] = sat_cast<int32_t>(v, -2147483648.f, 2147483647.f); break;
        case CLE_I8: ((int8_t*)b.p)[i] = sat_cast<int8_t>(v, -128.f, 127.f); break;
        case CLE_I16: ((int16_t*)b.p)[i] = sat_cast<int16_t>(v, -32768.f, 32767.f); break;
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

// ----------------------------------------------------------------- internal APIs
// Used across translation units by the composite entry points.
int blur_impl(const cle_buf* src, const cle_buf* dst, const float sig[3], hipStream_t s);
int minmax_impl(DBuf src, int64_t n, float* out2, hipStream_t s);
// Otsu threshold of `src`, fully on the device: writes minmax[2] and threshold
// (float64) to device memory.
int otsu_impl(DBuf src, const Geom& g, float* minmax_dev, double* thr_dev, hipStream_t s);
int ccl_impl(DBuf binary, const Geom& g, DBuf dst, int connectivity, int options,
             uint32_t* n_labels_dev, hipStream_t s);
int voronoi_grow_impl(float* a, float* b, const Geom& g, int* n_iterations, int* final_in_a,
                      hipStream_t s);

}  // namespace cle
