// Library runtime: error reporting, argument validation, the scratch arenas, fills,
// and the set / copy entry points.
#include <stdarg.h>
#include <stdio.h>
#include <map>
#include <mutex>
#include <vector>

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
    if (b->strides[0] || b->strides[1] || b->strides[2]) {
        int64_t expect = 1;  // C-contiguous element stride of axis i
        for (int i = b->ndim - 1; i >= 0; --i) {
            if (b->shape[i] > 1 && b->strides[i] != expect)
                return fail(CLE_E_SHAPE, "%s: strides (%lld, %lld, %lld) are not C-contiguous; the kernels take "
                            "contiguous images", name, (long long)b->strides[0], (long long)b->strides[1],
                            (long long)b->strides[2]);
            expect *= b->shape[i];
        }
    }
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

// ----------------------------------------------------------------- scratch arenas
// One grow-only arena per (device, stream).  Calls on one stream run in stream order,
// so memory a call released on the host can be handed to the next call on the same
// stream at once: allocation is a bump of an offset, release pops a LIFO stack, and no
// driver allocator (whose free can block the host until the device drains) runs on the
// steady-state path.  An arena grows by adding a chunk; when it is idle again and holds
// more than one chunk, its stream is drained and the chunks are merged into one.
namespace {
constexpr size_t ARENA_ALIGN = 256;
constexpr size_t ARENA_MIN_CHUNK = 64ull << 20;

struct ArenaBlock {
    int chunk;
    size_t off, size;
    bool freed;
};

struct Arena {
    std::mutex mu;  // the arena's own state; g_arena_mu guards only the map
    hipStream_t s = nullptr;
    std::vector<std::pair<char*, size_t>> chunks;
    std::vector<ArenaBlock> stack;
    int cur = 0;      // chunk the bump offset points into
    size_t top = 0;   // bump offset in chunks[cur]
    bool owned = false;  // held by a captured graph: cle_release_pool leaves it alone
    size_t capacity() const {
        size_t t = 0;
        for (auto& c : chunks) t += c.second;
        return t;
    }
};

std::mutex g_arena_mu;
std::map<std::pair<int, hipStream_t>, Arena*> g_arenas;

// The arena of (current device, s), created on first use.  Arenas are deleted only by
// cle_release_pool / arena_release_stream, which the caller runs while no call on that
// stream is in flight, so the pointer stays valid for the calling thread.
Arena* arena_for(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_arena_mu);
    auto& a = g_arenas[{dev, s}];
    if (!a) {
        a = new Arena;
        a->s = s;
    }
    return a;
}

// One chunk of at least `want` bytes in place of the arena's chunks (caller holds a->mu,
// the arena is idle).  Drains the arena's stream only: other streams' arenas and calls
// are not held up.
bool arena_merge(Arena* a, size_t want) {
    size_t total = a->capacity();
    if (total < want) total = want;
    if (hipStreamSynchronize(a->s) != hipSuccess) return false;
    for (auto& c : a->chunks) (void)hipFree(c.first);
    a->chunks.clear();
    a->cur = 0;
    a->top = 0;
    void* p = nullptr;
    if (hipMalloc(&p, total) != hipSuccess) return false;
    a->chunks.push_back({(char*)p, total});
    return true;
}
}  // namespace

// ----------------------------------------------------------------- fixed region
namespace {
struct FixedRegion {
    char* base = nullptr;
    size_t cap = 0, top = 0;
    int n = 0;  // live allocations
    struct Blk {
        size_t off;
        bool freed;
    } blk[64];
};
thread_local FixedRegion t_fixed;
}  // namespace

FixedScratch::FixedScratch(char* base, size_t cap) {
    t_fixed.base = base;
    t_fixed.cap = cap;
    t_fixed.top = 0;
    t_fixed.n = 0;
}

FixedScratch::~FixedScratch() {
    t_fixed.base = nullptr;
    t_fixed.cap = t_fixed.top = 0;
    t_fixed.n = 0;
}

bool scratch_fixed() { return t_fixed.base != nullptr; }

static void* fixed_alloc(size_t bytes) {
    FixedRegion& f = t_fixed;
    const size_t need = scratch_round(bytes);
    if (f.n == (int)(sizeof(f.blk) / sizeof(f.blk[0])) || need > f.cap - f.top) return nullptr;
    f.blk[f.n++] = {f.top, false};
    char* p = f.base + f.top;
    f.top += need;
    return p;
}

static void fixed_free(void* p) {
    FixedRegion& f = t_fixed;
    for (int i = f.n; i-- > 0;)
        if (f.base + f.blk[i].off == (char*)p) {
            f.blk[i].freed = true;
            break;
        }
    while (f.n > 0 && f.blk[f.n - 1].freed) f.top = f.blk[--f.n].off;
}

void* scratch_alloc(size_t bytes, hipStream_t s) {
    if (t_fixed.base) return fixed_alloc(bytes);
    Arena* a = arena_for(s);
    if (!a) return nullptr;
    std::lock_guard<std::mutex> lk(a->mu);
    const size_t need = (bytes + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
    if (a->stack.empty()) {
        if (a->chunks.size() > 1 && !arena_merge(a, 0)) {
            (void)hipGetLastError();
            return nullptr;
        }
        a->cur = 0;
        a->top = 0;
    }
    while (a->cur < (int)a->chunks.size() && a->top + need > a->chunks[a->cur].second) {
        ++a->cur;
        a->top = 0;
    }
    if (a->cur == (int)a->chunks.size()) {
        size_t sz = a->capacity();
        if (sz < ARENA_MIN_CHUNK) sz = ARENA_MIN_CHUNK;
        if (sz < need) sz = need;
        void* p = nullptr;
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        a->chunks.push_back({(char*)p, sz});
        a->top = 0;
    }
    char* p = a->chunks[a->cur].first + a->top;
    a->stack.push_back({a->cur, a->top, need, false});
    a->top += need;
    return p;
}

// The arena of (current device, s) becomes owned by a captured graph: its memory
// stays where the graph's kernels expect it until arena_release_stream(s).
int arena_set_owned(hipStream_t s) {
    Arena* a = arena_for(s);
    if (!a) return fail(CLE_E_HIP, "no arena for the stream");
    std::lock_guard<std::mutex> lk(a->mu);
    a->owned = true;
    return CLE_OK;
}

// Frees the arena of (current device, s) after draining s.
int arena_release_stream(hipStream_t s) {
    int dev = 0;
    CLE_HIP(hipGetDevice(&dev));
    CLE_HIP(hipStreamSynchronize(s));
    Arena* a = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_arena_mu);
        auto it = g_arenas.find({dev, s});
        if (it == g_arenas.end()) return CLE_OK;
        a = it->second;
        g_arenas.erase(it);
    }
    // out of the map: the frees hold up no other stream's lookups
    int rc = CLE_OK;
    for (auto& c : a->chunks)
        if (hipFree(c.first) != hipSuccess && rc == CLE_OK) rc = fail(CLE_E_HIP, "hipFree of a scratch chunk failed");
    delete a;
    return rc;
}

void scratch_free(void* p, hipStream_t s) {
    if (!p) return;
    if (t_fixed.base && (char*)p >= t_fixed.base && (char*)p < t_fixed.base + t_fixed.cap) {
        fixed_free(p);
        return;
    }
    Arena* a = arena_for(s);
    if (!a) return;
    std::lock_guard<std::mutex> lk(a->mu);
    for (size_t i = a->stack.size(); i-- > 0;) {
        ArenaBlock& b = a->stack[i];
        if (a->chunks[b.chunk].first + b.off == (char*)p) {
            b.freed = true;
            break;
        }
    }
    while (!a->stack.empty() && a->stack.back().freed) {
        a->cur = a->stack.back().chunk;
        a->top = a->stack.back().off;
        a->stack.pop_back();
    }
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
    std::lock_guard<std::mutex> lk(g_arena_mu);
    for (auto it = g_arenas.begin(); it != g_arenas.end();) {
        Arena* a = it->second;
        std::unique_lock<std::mutex> al(a->mu);
        if (it->first.first != dev || !a->stack.empty() || a->owned) {
            ++it;
            continue;
        }
        for (auto& c : a->chunks) CLE_HIP(hipFree(c.first));
        al.unlock();
        delete a;
        it = g_arenas.erase(it);
    }
    return CLE_OK;
}

int cle_release_stream(void* stream) { return arena_release_stream((hipStream_t)stream); }

int cle_reserve_stream(void* stream, size_t bytes) {
    Arena* a = arena_for((hipStream_t)stream);
    if (!a) return fail(CLE_E_HIP, "cle_reserve_stream: no current device");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->stack.empty()) return fail(CLE_E_VALUE, "cle_reserve_stream: a call on this stream holds scratch");
    if (a->chunks.size() == 1 && a->chunks[0].second >= bytes) return CLE_OK;
    if (!arena_merge(a, bytes)) {
        const hipError_t e = hipGetLastError();
        return fail(CLE_E_OOM, "cle_reserve_stream of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
    }
    return CLE_OK;
}

int cle_alloc(size_t bytes, void** out) {
    if (!out) return fail(CLE_E_VALUE, "cle_alloc: null output pointer");
    *out = nullptr;
    if (bytes == 0) return CLE_OK;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(CLE_E_OOM, "cle_alloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
    }
    *out = p;
    return CLE_OK;
}

int cle_free(void* p) {
    if (p) CLE_HIP(hipFree(p));
    return CLE_OK;
}

int cle_upload(void* dst_device, const void* src_host, size_t bytes, void* stream) {
    if (!bytes) return CLE_OK;
    if (!dst_device || !src_host) return fail(CLE_E_VALUE, "cle_upload: null pointer");
    const hipStream_t s = (hipStream_t)stream;
    CLE_HIP(hipMemcpyAsync(dst_device, src_host, bytes, hipMemcpyHostToDevice, s));
    CLE_HIP(hipStreamSynchronize(s));
    return CLE_OK;
}

int cle_download(void* dst_host, const void* src_device, size_t bytes, void* stream) {
    if (!bytes) return CLE_OK;
    if (!dst_host || !src_device) return fail(CLE_E_VALUE, "cle_download: null pointer");
    const hipStream_t s = (hipStream_t)stream;
    CLE_HIP(hipMemcpyAsync(dst_host, src_device, bytes, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
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
