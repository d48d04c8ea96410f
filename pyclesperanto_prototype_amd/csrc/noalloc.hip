// voronoi_otsu_labeling without any allocation: for callers that cannot allocate
// device memory (code under stream capture, tight serving loops).  Its scratch is one
// file-scope device array per GPU, reached through the address the runtime reports for
// the symbol on that GPU (the array's name in host code is a host-side stand-in, not a
// device address).  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU (two host threads, two streams) race on it.  Nothing in the array has to
// start at zero: each stage writes what it reads.
#include <mutex>

#include "common.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)CLE_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[CLE_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` on each GPU

// the device address of this GPU's array, looked up on the first call on it
int scratch_address(void** out) {
    int dev = 0;
    CLE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return cle::fail(CLE_E_VALUE, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cle::fail(CLE_E_HIP, "hipGetSymbolAddress of the fixed scratch array failed on device %d (%s)", dev,
                             hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return CLE_OK;
}

}  // namespace

using namespace cle;

extern "C" {

int cle_voronoi_otsu_labeling_noalloc(const cle_buf* src, const cle_buf* dst, float spot_sigma, float outline_sigma,
                                      void* stream) {
    Geom g, gd;
    CLE_TRY(check_buf(src, "source", &g));
    CLE_TRY(check_buf(dst, "label_image_destination", &gd));
    CLE_TRY(check_same_shape(g, gd, "source", "label_image_destination"));
    if (!g.n()) return CLE_OK;
    const size_t need = voronoi_otsu_scratch_bytes(g);
    if (need > kScratchBytes)
        return fail(CLE_E_VALUE,
                    "voronoi_otsu_labeling_noalloc: a (%lld, %lld, %lld) image needs %zu bytes of scratch, more than "
                    "the fixed %zu (roughly 12 to 15 bytes per voxel)",
                    (long long)g.D, (long long)g.H, (long long)g.W, need, kScratchBytes);
    void* base = nullptr;
    CLE_TRY(scratch_address(&base));
    FixedScratch fixed((char*)base, kScratchBytes);
    return voronoi_otsu_single(src, dst, spot_sigma, outline_sigma, (hipStream_t)stream);
}

}  // extern "C"
