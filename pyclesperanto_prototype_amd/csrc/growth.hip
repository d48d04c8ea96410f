// Masked / unmasked Voronoi growth by asynchronous exact tile solves.
//
// Reference semantics (pyclesperanto_prototype/_tier5/_masked_voronoi_labeling.py:44-74,
// _tier4/_extend_labeling_via_voronoi.py:39-61): starting from v0 (float32), step T = 0,
// 1, 2, ... gives every zero voxel the largest positive value among its box (T even)
// or diamond (T odd) neighbours, synchronously, and the process stops after the first
// step that changes nothing anywhere in the image.
//
// The same result, computed without image-wide steps.  Let t(v) be the step at which a
// zero voxel is assigned in the process run without the stop (seeds: t = -1).  Then
// t(v) = min over neighbours u with t(u) finite of: t(u) + 1 if u is a face neighbour,
// the smallest even number above t(u) otherwise; and v's value is the largest value of
// the neighbours u in the step's neighbourhood with t(u) < t(v).  Every dependency goes
// to a strictly smaller t, so these equations have exactly one solution, and it is
// the reference's process: by induction over T, both assign the same voxels with the
// same values before every step T.  The reference's stop at step S (the first step with
// no assignment) keeps exactly the voxels with t < S.
//
// So the image is cut into tiles (62 x 6 x 6 voxels in 3D, 62 x 62 in 2D: with a one-
// voxel halo, region rows of 64 voxels, 64 rows), and one wavefront solves a tile
// exactly from its halo: the halo voxels' (t, value) are replayed, each turning "on"
// at its step, and the tile runs the reference's steps on 64-bit masks of its rows.
// With at most GKMAX distinct values in the region, the on-set is kept per value (one
// mask per value and row, a lane per row, neighbour rows by cross-lane shuffles) and a
// step dilates the masks from the largest value down, so the first value that reaches
// a zero voxel is its maximum; regions with more values take a per-voxel maximum in
// LDS (a second kernel).  A solve writes the tile's output and publishes the (t, value)
// of its face voxels; when they change, the neighbouring tiles that read them are
// queued and solved again.  The t of every solve are upper bounds that only decrease,
// and once they are final the values settle in order of t, so the rounds end, at the
// unique solution, with every tile's output from its last solve.  Each solve also
// records the steps at which it assigned voxels; a gap in an image's steps below their
// maximum is the reference's stop, and that image's tiles are solved once more with
// the steps capped there.
//
// Steps are held in bytes (code c: 0 seed, 1 .. 254 assigned at step c - 1, on from
// step c; 255 never); a tile that would need a step >= 254 reports an overflow and the
// caller runs the step-synchronous growth instead.
#include <type_traits>

#include "common.hpp"
#include "growth.hpp"

namespace cle {

constexpr int GW = 64;                 // threads per tile (one wavefront)
constexpr int GKMAX = 16;              // distinct values a register-resident solve holds
constexpr uint8_t GNONE = 0xFF;        // never on (wall, unassigned, outside)
constexpr uint64_t GROW_INT = 0x7FFFFFFFFFFFFFFEull;  // interior bits of a region row
constexpr uint32_t GFREE = 0xFFFFFFFFu;

__host__ __device__ inline bool g_is3d(const Geom& g) { return g.ndim == 3 && g.D > 1; }

void grow_tiling(BImg& im) {
    const Geom& g = im.g;
    im.wnx = (int)((g.W + GX - 1) / GX);
    if (g_is3d(g)) {
        im.wny = (int)((g.H + GT3 - 1) / GT3);
        im.wnz = (int)((g.D + GT3 - 1) / GT3);
    } else {
        im.wny = (int)((g.H + GT2 - 1) / GT2);
        im.wnz = 1;
    }
    if (g.n() == 0) im.wnx = im.wny = im.wnz = 0;
}

int64_t grow_tile_count(const BImg& im) { return (int64_t)im.wnx * im.wny * im.wnz; }

struct GItem {
    uint64_t c;  // image << 48 | tz << 32 | ty << 16 | tx
    int64_t gt;  // the tile's index over all images
};
__host__ __device__ inline uint64_t g_pack(int j, int tx, int ty, int tz) {
    return ((uint64_t)j << 48) | ((uint64_t)tz << 32) | ((uint64_t)ty << 16) | (uint64_t)tx;
}

// store entry of a tile-face voxel (tile-local lx, ly, lz) on the face of the tile
// that holds it, read from the side of the neighbour at offset (dx, dy, dz): its x
// face if dx != 0, else its y face, else its z face (side 0 low, side 1 high)
template <bool D3>
__device__ __forceinline__ int g_face_entry(int dx, int dy, int dz, int lx, int ly, int lz) {
    if (D3) {
        if (dx) return (dx < 0 ? GF3_X1 : GF3_X0) + lz * GT3 + ly;
        if (dy) return (dy < 0 ? GF3_Y1 : GF3_Y0) + lz * GX + lx;
        return (dz < 0 ? GF3_Z1 : GF3_Z0) + ly * GX + lx;
    } else {
        if (dx) return (dx < 0 ? GF2_X1 : GF2_X0) + ly;
        return (dy < 0 ? GF2_Y1 : GF2_Y0) + lx;
    }
}

// ----------------------------------------------------------------- sources / outputs
// kFast sources read v0 as ((float)u8[i] - 1) + f32[i] from two packed buffers (the
// loads are issued for a whole region at once), and their output is 0 for every voxel
// that is not on (kZeroOff); the others call v0() per voxel.
//
// voronoi_otsu pipeline and batches: v0 = (segmentation - 1) + seed label (float32, as
// the reference's add_image_and_scalar then add_images), output masked by the
// segmentation into the image's destination
struct GSrcSegLab {
    static constexpr bool kFast = true, kZeroOff = true;
    const uint8_t* seg;
    const float* lab;
    __device__ __forceinline__ const uint8_t* u8p(const BImg& im) const { return seg + im.off; }
    __device__ __forceinline__ const float* f32p(const BImg& im) const { return lab + im.off; }
    __device__ __forceinline__ float v0(const BImg& im, int64_t i) const {
        return ((float)seg[im.off + i] - 1.f) + lab[im.off + i];
    }
    __device__ __forceinline__ void out(const BImg& im, int64_t i, float v) const {
        st(DBuf{im.dst, im.dst_dt}, i, seg[im.off + i] ? v : 0.f);
    }
};
// masked_voronoi_labeling with a mask of any type (one image)
struct GSrcMaskLab {
    static constexpr bool kFast = false, kZeroOff = false;
    DBuf mask;
    const float* lab;
    DBuf dst;
    __device__ __forceinline__ const uint8_t* u8p(const BImg&) const { return nullptr; }
    __device__ __forceinline__ const float* f32p(const BImg&) const { return nullptr; }
    __device__ __forceinline__ float v0(const BImg&, int64_t i) const { return (ld(mask, i) - 1.f) + lab[i]; }
    __device__ __forceinline__ void out(const BImg&, int64_t i, float v) const {
        st(dst, i, ld(mask, i) != 0.f ? v : 0.f);
    }
};
// extend_labeling_via_voronoi: the labels converted to float32, no mask (one image)
struct GSrcLabels {
    static constexpr bool kFast = false, kZeroOff = false;
    DBuf src;
    DBuf dst;
    __device__ __forceinline__ const uint8_t* u8p(const BImg&) const { return nullptr; }
    __device__ __forceinline__ const float* f32p(const BImg&) const { return nullptr; }
    __device__ __forceinline__ float v0(const BImg&, int64_t i) const { return ld(src, i); }
    __device__ __forceinline__ void out(const BImg&, int64_t i, float v) const { st(dst, i, v); }
};

// ----------------------------------------------------------------- LDS of the solves
// value-sliced solve: per region voxel (row * 64 + x) its code and the hash slot of
// its value (voxels that are or turn on); the hash set of the region's values, their
// ranks (ascending value) and the slot of each rank
struct GLdsB {
    uint8_t code[4096];
    uint8_t slot[4096];
    uint32_t hkey[64];   // 0: empty (values that turn on are positive)
    uint32_t hrank[64];
    uint32_t rslot[GKMAX];
    uint32_t nslot[27];
    uint32_t nkeys;
};
// per-voxel solve: values and codes of the region, the on-set of the rows, the voxels
// assigned in a step and their prefix over the rows
struct GLdsV {
    uint32_t L[4096];
    uint8_t code[4096];
    uint64_t on[64];
    uint64_t nw[64];
    int pf[65];
    uint32_t nslot[27];
};

__device__ __forceinline__ int g_hash(uint32_t k) { return (int)((k * 2654435761u) >> 26); }

// the hash slot of value k (inserted if new; any slot once the set holds more than
// GKMAX values: the tile then goes to the per-voxel solve)
__device__ __forceinline__ int g_insert(GLdsB& s, uint32_t k) {
    int h = g_hash(k);
    for (int probe = 0; probe < 64; ++probe) {
        const uint32_t prev = atomicCAS(&s.hkey[h], 0u, k);
        if (prev == 0u) {
            atomicAdd(&s.nkeys, 1u);
            return h;
        }
        if (prev == k) return h;
        if (s.nkeys > (uint32_t)GKMAX) return 0;
        h = (h + 1) & 63;
    }
    return 0;
}

struct GSinkB {
    GLdsB& s;
    uint32_t last;
    int lslot;
    __device__ __forceinline__ void put(int ri, uint32_t c, uint32_t v) {
        s.code[ri] = (uint8_t)c;
        if (c != GNONE) {
            if (v != last) {  // neighbouring voxels mostly repeat a value
                last = v;
                lslot = g_insert(s, v);
            }
            s.slot[ri] = (uint8_t)lslot;
        }
    }
};
struct GSinkV {
    GLdsV& s;
    __device__ __forceinline__ void put(int ri, uint32_t c, uint32_t v) {
        s.code[ri] = (uint8_t)c;
        s.L[ri] = v;
    }
};

// ----------------------------------------------------------------- region load
// The 27 neighbour tiles' store slots (GFREE: none, or not active) into nslot.
__device__ __forceinline__ void g_neighbour_slots(const BImg& im, int tx, int ty, int tz,
                                                  const uint32_t* __restrict__ map, uint32_t* nslot) {
    const int k = threadIdx.x;
    if (k < 27) {
        const int qx = tx + k % 3 - 1, qy = ty + (k / 3) % 3 - 1, qz = tz + k / 9 - 1;
        uint32_t sl = GFREE;
        if (qx >= 0 && qx < im.wnx && qy >= 0 && qy < im.wny && qz >= 0 && qz < im.wnz && k != 13)
            sl = map[im.wbase + ((int64_t)qz * im.wny + qy) * im.wnx + qx];
        nslot[k] = sl;
    }
    __syncthreads();
}

constexpr int GCHUNK = 16;  // region rows whose loads are in flight together

// The region of tile (tx, ty, tz), lane = region x: every voxel's code (0: on from the
// start, 1..254: on from that step, 255: never) and value go to `sink`; lane r gets
// row r's interior zero voxels (zb) and the voxels that are or turn on (onp).  Halo
// voxels come from the neighbours' stores when `nslot` is given and the neighbour is
// active, else from v0; only interior voxels are zero (growable).
template <bool D3, typename SRC, typename SINK>
__device__ __forceinline__ void g_load(const SRC& src, const BImg& im, int tx, int ty, int tz, const uint32_t* nslot,
                                       const uint8_t* __restrict__ scode, const uint32_t* __restrict__ sL, SINK& sink,
                                       uint64_t& zb, uint64_t& onp) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const Geom& g = im.g;
    const int lane = threadIdx.x;
    const int64_t X = (int64_t)tx * GX - 1 + lane;
    const bool xin = X >= 0 && X < g.W;
    const int64_t Y0 = (int64_t)ty * TY - 1, Z0 = D3 ? (int64_t)tz * TZ - 1 : 0;
    const int dx = lane == 0 ? -1 : (lane == GW - 1 ? 1 : 0);
    zb = 0;
    onp = 0;
#pragma unroll 1
    for (int r0 = 0; r0 < 64; r0 += GCHUNK) {
        uint32_t vv[GCHUNK], cc[GCHUNK];
        uint32_t fromstore = 0, inside = 0, zero = 0;
#pragma unroll
        for (int k = 0; k < GCHUNK; ++k) {
            const int r = r0 + k;
            const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
            const int64_t Y = Y0 + ry, Z = D3 ? Z0 + rz : 0;
            const bool in = xin && Y >= 0 && Y < g.H && Z >= 0 && Z < g.D;
            const int dy = ry == 0 ? -1 : (ry == TY + 1 ? 1 : 0);
            const int dz = D3 ? (rz == 0 ? -1 : (rz == TZ + 1 ? 1 : 0)) : 0;
            const int64_t i = (Z * g.H + Y) * g.W + X;
            int64_t a = -1;
            if (in && nslot && (dx | dy | dz)) {
                const uint32_t sl = nslot[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                if (sl != GFREE) {
                    const int lx = dx < 0 ? GX - 1 : (dx > 0 ? 0 : lane - 1);
                    const int ly = dy < 0 ? TY - 1 : (dy > 0 ? 0 : ry - 1);
                    const int lz = D3 ? (dz < 0 ? TZ - 1 : (dz > 0 ? 0 : rz - 1)) : 0;
                    a = (int64_t)sl * GSLOT + g_face_entry<D3>(dx, dy, dz, lx, ly, lz);
                }
            }
            fromstore |= (a >= 0 ? 1u : 0u) << k;
            inside |= (in ? 1u : 0u) << k;
            if constexpr (SRC::kFast) {
                // two loads per voxel, from the sources or a neighbour's store, addresses
                // chosen without branches so that the chunk's loads overlap
                const uint8_t* p8 = a >= 0 ? scode + a : src.u8p(im) + (in ? i : 0);
                const uint32_t* p32 = a >= 0 ? sL + a : (const uint32_t*)src.f32p(im) + (in ? i : 0);
                cc[k] = *p8;
                vv[k] = *p32;
            } else {
                if (a >= 0) {
                    cc[k] = scode[a];
                    vv[k] = sL[a];
                } else if (in) {
                    const float v = src.v0(im, i);
                    vv[k] = __float_as_uint(v);
                    cc[k] = v > 0.f ? 0u : GNONE;
                    zero |= (v == 0.f && !(dx | dy | dz) ? 1u : 0u) << k;
                } else {
                    cc[k] = GNONE;
                    vv[k] = 0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < GCHUNK; ++k) {
            const int r = r0 + k;
            const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
            const bool interior = lane >= 1 && lane <= GX && ry >= 1 && ry <= TY && rz >= 1 && rz <= TZ;
            uint32_t c = cc[k], v = vv[k];
            bool z = (zero >> k) & 1u;
            if constexpr (SRC::kFast) {
                if (!((inside >> k) & 1u)) {
                    c = GNONE;
                    v = 0;
                } else if (!((fromstore >> k) & 1u)) {
                    const float f = ((float)c - 1.f) + __uint_as_float(v);
                    v = __float_as_uint(f);
                    c = f > 0.f ? 0u : GNONE;
                    z = f == 0.f && interior;
                }
            }
            sink.put(r * 64 + lane, c, v);
            const uint64_t zr = __ballot(z), orow = __ballot(c != GNONE);
            if (lane == r) {
                zb = zr;
                onp = orow;
            }
        }
    }
    const int ry = D3 ? (lane & 7) : lane, rz = D3 ? (lane >> 3) : 1;
    if (!(ry >= 1 && ry <= TY && rz >= 1 && rz <= TZ)) zb = 0;
    zb &= GROW_INT;
    __syncthreads();
}

// the m-th set bit (from 0) of w
__device__ __forceinline__ int g_nth_bit(uint64_t w, int m) {
    int base = 0;
    uint32_t x = (uint32_t)w;
    int c = __popc(x);
    if (m >= c) {
        m -= c;
        x = (uint32_t)(w >> 32);
        base = 32;
    }
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        const uint32_t lo = x & ((1u << half) - 1u);
        c = __popc(lo);
        if (m >= c) {
            m -= c;
            x >>= half;
            base += half;
        }
    }
    return base;
}

__device__ __forceinline__ uint64_t g_shfl64(uint64_t v, int src) { return (uint64_t)__shfl((long long)v, src); }

struct GSteps {
    uint64_t w[4];
    __device__ __forceinline__ void clear() { w[0] = w[1] = w[2] = w[3] = 0; }
    __device__ __forceinline__ void set(int T) {
        const uint64_t b = 1ull << (T & 63);
        if (T < 64) w[0] |= b;
        else if (T < 128) w[1] |= b;
        else if (T < 192) w[2] |= b;
        else w[3] |= b;
    }
};

// lane r: the smallest code of the voxels in `pend` of row r
__device__ __forceinline__ int g_first_code(const uint8_t* code, int r, uint64_t pend) {
    int m = 255;
    for (uint64_t p = pend; p; p &= p - 1) {
        const int c = code[r * 64 + __ffsll((unsigned long long)p) - 1];
        m = c < m ? c : m;
    }
    return m;
}

// ----------------------------------------------------------------- value-sliced solve
// The region's values ranked (ascending; positive floats order as their bits).
// Returns their count K (> GKMAX: too many for this solve).
__device__ __forceinline__ int g_rank_values(GLdsB& s) {
    const int lane = threadIdx.x;
    const int K = (int)s.nkeys;
    if (K <= GKMAX) {
        const uint32_t key = s.hkey[lane];
        int rank = 0;
        for (int j = 0; j < 64; ++j) {
            const uint32_t o = (uint32_t)__shfl((int)key, j);
            rank += (o != 0u && o < key) ? 1 : 0;
        }
        s.hrank[lane] = (uint32_t)rank;
        if (key) s.rslot[rank] = (uint32_t)lane;
    }
    __syncthreads();
    return K;
}

// The reference's steps on the tile, lane = region row, value-sliced (K <= GKMAX): one
// on-mask per value; a step dilates them from the largest value down.  Steps T >= cap
// are not run when cap < GMAXSTEP.  st: the steps with an assignment; *over when a
// step >= GMAXSTEP would assign a voxel.
template <bool D3>
__device__ __forceinline__ void g_solve_bits(GLdsB& s, uint64_t zb, uint64_t onp, int K, int cap, GSteps& st,
                                             bool* over) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const int r = threadIdx.x;
    const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
    const bool irow = ry >= 1 && ry <= TY && rz >= 1 && rz <= TZ;
    uint64_t on[GKMAX];
#pragma unroll
    for (int k = 0; k < GKMAX; ++k) on[k] = 0;
    uint64_t zero = zb, pend = onp;  // seeds turn on at step 0, halo voxels at their step
    int nact = g_first_code(s.code, r, pend);
    st.clear();
    int T = 0;
    for (;;) {
        if (T >= cap && cap < GMAXSTEP) break;
        if (pend && nact <= T) {
            int na = 255;
            for (uint64_t p = pend; p; p &= p - 1) {
                const int b = __ffsll((unsigned long long)p) - 1;
                const int c = s.code[r * 64 + b];
                if (c <= T) {
                    const int rk = (int)s.hrank[s.slot[r * 64 + b]];
#pragma unroll
                    for (int k = 0; k < GKMAX; ++k)
                        if (k == rk) on[k] |= 1ull << b;
                    pend &= ~(1ull << b);
                } else {
                    na = c < na ? c : na;
                }
            }
            nact = na;
        }
        uint64_t taken = 0;
        const bool box = (T & 1) == 0;
#pragma unroll
        for (int k = GKMAX - 1; k >= 0; --k) {
            if (k < K) {
                const uint64_t o = on[k];
                uint64_t d;
                if (box) {
                    const uint64_t h = o | (o << 1) | (o >> 1);
                    const uint64_t a = h | g_shfl64(h, r - 1) | g_shfl64(h, r + 1);
                    d = D3 ? (a | g_shfl64(a, r - 8) | g_shfl64(a, r + 8)) : a;
                } else {
                    d = (o << 1) | (o >> 1) | g_shfl64(o, r - 1) | g_shfl64(o, r + 1);
                    if (D3) d |= g_shfl64(o, r - 8) | g_shfl64(o, r + 8);
                }
                // a value's own dilation never reads the masks of the smaller values,
                // so its mask takes its new voxels at once
                const uint64_t hit = irow ? (zero & d & ~taken) : 0ull;
                if (hit) {
                    on[k] = o | hit;
                    const uint8_t sl = (uint8_t)s.rslot[k];
                    for (uint64_t w = hit; w; w &= w - 1) {
                        const int b = __ffsll((unsigned long long)w) - 1;
                        s.code[r * 64 + b] = (uint8_t)(T + 1);
                        s.slot[r * 64 + b] = sl;
                    }
                    taken |= hit;
                }
            }
        }
        if (__ballot(taken != 0)) {
            if (T >= GMAXSTEP) {
                *over = true;
                break;
            }
            zero &= ~taken;
            st.set(T);
            T += 1;
        } else if (T & 1) {
            T += 1;  // a diamond step without change can be followed by a box step that changes
        } else {
            if (!__ballot(pend != 0)) break;  // nothing changes any more
            int m = pend ? nact : 255;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const int q = __shfl_xor(m, o);
                m = q < m ? q : m;
            }
            T = m > T + 1 ? m : T + 1;  // skip to the next step a halo voxel turns on
        }
    }
    __syncthreads();
}

// ----------------------------------------------------------------- per-voxel solve
// The same steps for regions with more than GKMAX values: the on-set per row, the
// maximum per assigned voxel over its neighbours' values in LDS.
template <bool D3>
__device__ __forceinline__ void g_solve_vox(GLdsV& s, uint64_t zb, uint64_t onp, int cap, GSteps& st, bool* over) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const int r = threadIdx.x;
    const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
    const bool irow = ry >= 1 && ry <= TY && rz >= 1 && rz <= TZ;
    uint64_t on = 0, zero = zb, pend = onp;
    int nact = g_first_code(s.code, r, pend);
    st.clear();
    int T = 0;
    for (;;) {
        if (T >= cap && cap < GMAXSTEP) break;
        if (pend && nact <= T) {
            int na = 255;
            for (uint64_t p = pend; p; p &= p - 1) {
                const int b = __ffsll((unsigned long long)p) - 1;
                const int c = s.code[r * 64 + b];
                if (c <= T) {
                    on |= 1ull << b;
                    pend &= ~(1ull << b);
                } else {
                    na = c < na ? c : na;
                }
            }
            nact = na;
        }
        __syncthreads();  // the previous step's reads of s.on are done
        s.on[r] = on;
        __syncthreads();
        uint64_t dil = 0;
        if (irow) {
            if ((T & 1) == 0) {
#pragma unroll
                for (int dz = (D3 ? -1 : 0); dz <= (D3 ? 1 : 0); ++dz)
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy) {
                        const uint64_t w = s.on[r + dy + 8 * dz];
                        dil |= w | (w << 1) | (w >> 1);
                    }
            } else {
                dil = (on << 1) | (on >> 1) | s.on[r - 1] | s.on[r + 1];
                if (D3) dil |= s.on[r - 8] | s.on[r + 8];
            }
        }
        const uint64_t nw = zero & dil;
        if (__ballot(nw != 0)) {
            if (T >= GMAXSTEP) {
                *over = true;
                break;
            }
            const int cnt = __popcll(nw);
            int inc = cnt;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(inc, d);
                if (r >= d) inc += o;
            }
            s.nw[r] = nw;
            s.pf[r] = inc - cnt;
            if (r == 63) s.pf[64] = inc;
            __syncthreads();
            const int total = s.pf[64];
            for (int k = r; k < total; k += 64) {
                int q = 0;
#pragma unroll
                for (int stp = 32; stp >= 1; stp >>= 1)
                    if (s.pf[q + stp] <= k) q += stp;
                const int b = g_nth_bit(s.nw[q], k - s.pf[q]);
                uint32_t best = 0;
                if ((T & 1) == 0) {
#pragma unroll
                    for (int dz = (D3 ? -1 : 0); dz <= (D3 ? 1 : 0); ++dz)
#pragma unroll
                        for (int dy = -1; dy <= 1; ++dy) {
                            const int qq = q + dy + 8 * dz;
                            const uint64_t w = s.on[qq] >> (b - 1);
#pragma unroll
                            for (int dx = 0; dx < 3; ++dx)
                                if (((w >> dx) & 1ull) && !(dz == 0 && dy == 0 && dx == 1)) {
                                    const uint32_t l = s.L[qq * 64 + b - 1 + dx];
                                    best = l > best ? l : best;
                                }
                        }
                } else {
                    const int nb[6] = {q * 64 + b - 1, q * 64 + b + 1, (q - 1) * 64 + b, (q + 1) * 64 + b,
                                       (D3 ? q - 8 : q) * 64 + b, (D3 ? q + 8 : q) * 64 + b};
                    const bool onb[6] = {((s.on[q] >> (b - 1)) & 1ull) != 0, ((s.on[q] >> (b + 1)) & 1ull) != 0,
                                         ((s.on[q - 1] >> b) & 1ull) != 0, ((s.on[q + 1] >> b) & 1ull) != 0,
                                         D3 && ((s.on[D3 ? q - 8 : q] >> b) & 1ull) != 0,
                                         D3 && ((s.on[D3 ? q + 8 : q] >> b) & 1ull) != 0};
#pragma unroll
                    for (int k2 = 0; k2 < (D3 ? 6 : 4); ++k2)
                        if (onb[k2]) {
                            const uint32_t l = s.L[nb[k2]];
                            best = l > best ? l : best;
                        }
                }
                s.L[q * 64 + b] = best;
                s.code[q * 64 + b] = (uint8_t)(T + 1);
            }
            on |= nw;
            zero &= ~nw;
            st.set(T);
            T += 1;
        } else if (T & 1) {
            T += 1;
        } else {
            if (!__ballot(pend != 0)) break;
            int m = pend ? nact : 255;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const int q = __shfl_xor(m, o);
                m = q < m ? q : m;
            }
            T = m > T + 1 ? m : T + 1;
        }
    }
    __syncthreads();
}

// ----------------------------------------------------------------- publish / output
// 27-bit mask of the neighbour offsets whose halo holds tile voxel (lx, ly, lz)
template <bool D3>
__device__ __forceinline__ uint32_t g_nbr_mask(int lx, int ly, int lz) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const uint32_t xm = 2u | (lx == 0 ? 1u : 0u) | (lx == GX - 1 ? 4u : 0u);
    const uint32_t ym = 2u | (ly == 0 ? 1u : 0u) | (ly == TY - 1 ? 4u : 0u);
    const uint32_t zm = D3 ? (2u | (lz == 0 ? 1u : 0u) | (lz == TZ - 1 ? 4u : 0u)) : 2u;
    uint32_t m = 0;
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y)
            if (((zm >> z) & 1u) && ((ym >> y) & 1u)) m |= xm << (3 * y + 9 * z);
    return m;
}

// value bits and code of region voxel ri after a solve
struct GViewB {
    const GLdsB& s;
    __device__ __forceinline__ uint8_t code(int ri) const { return s.code[ri]; }
    __device__ __forceinline__ uint32_t val(int ri) const { return s.hkey[s.slot[ri]]; }
};
struct GViewV {
    const GLdsV& s;
    __device__ __forceinline__ uint8_t code(int ri) const { return s.code[ri]; }
    __device__ __forceinline__ uint32_t val(int ri) const { return s.L[ri]; }
};

template <bool D3, typename VIEW>
__device__ __forceinline__ uint32_t g_pub(const VIEW& v, const Geom& g, int64_t X0, int64_t Y0, int64_t Z0, int lx,
                                          int ly, int lz, int64_t a, uint8_t* scode, uint32_t* sL, bool first) {
    if (X0 + lx >= g.W || Y0 + ly >= g.H || Z0 + lz >= g.D) return 0;
    const int ri = D3 ? ((lz + 1) * 8 + ly + 1) * 64 + lx + 1 : (ly + 1) * 64 + lx + 1;
    const uint8_t c = v.code(ri);
    const uint32_t l = c != GNONE ? v.val(ri) : 0u;
    bool ch;
    if (first) {
        ch = c != 0 && c != GNONE;
    } else {
        const uint8_t oc = scode[a];
        ch = oc != c || (c != GNONE && sL[a] != l);
    }
    scode[a] = c;
    sL[a] = l;
    return ch ? g_nbr_mask<D3>(lx, ly, lz) : 0u;
}

// Publish the tile's faces (lane = position along a face); returns the OR over the
// wavefront of the neighbour offsets that must solve again.
template <bool D3, typename VIEW>
__device__ __forceinline__ uint32_t g_publish(const VIEW& v, const BImg& im, int tx, int ty, int tz, uint32_t slot,
                                              uint8_t* scode, uint32_t* sL, bool first) {
    constexpr int TY = D3 ? GT3 : GT2;
    const Geom& g = im.g;
    const int64_t X0 = (int64_t)tx * GX, Y0 = (int64_t)ty * TY, Z0 = D3 ? (int64_t)tz * GT3 : 0;
    const int64_t base = (int64_t)slot * GSLOT;
    const int lane = threadIdx.x;
    uint32_t m = 0;
    if (D3) {
        for (int e = lane; e < GT3 * GT3; e += GW) {
            const int ly = e % GT3, lz = e / GT3;
            m |= g_pub<true>(v, g, X0, Y0, Z0, 0, ly, lz, base + GF3_X0 + e, scode, sL, first);
            m |= g_pub<true>(v, g, X0, Y0, Z0, GX - 1, ly, lz, base + GF3_X1 + e, scode, sL, first);
        }
        for (int e = lane; e < GX * GT3; e += GW) {
            const int lx = e % GX, o = e / GX;
            m |= g_pub<true>(v, g, X0, Y0, Z0, lx, 0, o, base + GF3_Y0 + e, scode, sL, first);
            m |= g_pub<true>(v, g, X0, Y0, Z0, lx, GT3 - 1, o, base + GF3_Y1 + e, scode, sL, first);
            m |= g_pub<true>(v, g, X0, Y0, Z0, lx, o, 0, base + GF3_Z0 + e, scode, sL, first);
            m |= g_pub<true>(v, g, X0, Y0, Z0, lx, o, GT3 - 1, base + GF3_Z1 + e, scode, sL, first);
        }
    } else {
        for (int e = lane; e < GX; e += GW) {
            m |= g_pub<false>(v, g, X0, Y0, Z0, 0, e, 0, base + GF2_X0 + e, scode, sL, first);
            m |= g_pub<false>(v, g, X0, Y0, Z0, GX - 1, e, 0, base + GF2_X1 + e, scode, sL, first);
            m |= g_pub<false>(v, g, X0, Y0, Z0, e, 0, 0, base + GF2_Y0 + e, scode, sL, first);
            m |= g_pub<false>(v, g, X0, Y0, Z0, e, GT2 - 1, 0, base + GF2_Y1 + e, scode, sL, first);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m |= (uint32_t)__shfl_xor((int)m, o);
    return m & ~(1u << 13);
}

// the tile's output, lane = region x: grown and seed values; the other voxels their
// v0 (kZeroOff sources: 0)
template <bool D3, typename SRC, typename VIEW>
__device__ __forceinline__ void g_output(const SRC& src, const VIEW& v, const BImg& im, int tx, int ty, int tz) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const Geom& g = im.g;
    const int lane = threadIdx.x;
    const int64_t X = (int64_t)tx * GX - 1 + lane;
    if (lane < 1 || lane > GX || X >= g.W) return;
    for (int lz = 0; lz < TZ; ++lz) {
        const int64_t Z = D3 ? (int64_t)tz * TZ + lz : 0;
        if (Z >= g.D) break;
        for (int ly = 0; ly < TY; ++ly) {
            const int64_t Y = (int64_t)ty * TY + ly;
            if (Y >= g.H) break;
            const int ri = D3 ? ((lz + 1) * 8 + ly + 1) * 64 + lane : (ly + 1) * 64 + lane;
            const int64_t i = (Z * g.H + Y) * g.W + X;
            float val;
            if (v.code(ri) != GNONE) val = __uint_as_float(v.val(ri));
            else val = SRC::kZeroOff ? 0.f : src.v0(im, i);
            src.out(im, i, val);
        }
    }
}

// ----------------------------------------------------------------- kernels
// Every tile: the output of tiles without a zero voxel (final: nothing grows there),
// and a store slot for the others.  One wavefront per tile, lane = x.
template <typename SRC, bool D3>
__device__ __forceinline__ void g_classify_tile(const SRC& src, const BImg& im, int j, int tx, int ty, int tz,
                                                int64_t b, uint32_t* map, GItem* slots, uint32_t* nslots) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1, NR = TY * TZ;
    const Geom& g = im.g;
    const int lane = threadIdx.x;
    const int64_t X = (int64_t)tx * GX + lane;
    const bool xin = lane < GX && X < g.W;
    float v[NR];
    bool zero = false;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int ly = k % TY, lz = k / TY;
        const int64_t Y = (int64_t)ty * TY + ly, Z = D3 ? (int64_t)tz * TZ + lz : 0;
        const bool in = xin && Y < g.H && Z < g.D;
        v[k] = in ? src.v0(im, (Z * g.H + Y) * g.W + X) : 1.f;
        zero |= in && v[k] == 0.f;
    }
    const bool active = __ballot(zero) != 0;
    if (!active) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int ly = k % TY, lz = k / TY;
            const int64_t Y = (int64_t)ty * TY + ly, Z = D3 ? (int64_t)tz * TZ + lz : 0;
            if (xin && Y < g.H && Z < g.D) {
                const float o = SRC::kZeroOff ? (v[k] > 0.f ? v[k] : 0.f) : v[k];
                src.out(im, (Z * g.H + Y) * g.W + X, o);
            }
        }
    }
    if (lane == 0) {
        if (active) {
            const uint32_t slot = atomicAdd(nslots, 1u);
            map[b] = slot;
            slots[slot] = GItem{g_pack(j, tx, ty, tz), b};
        } else {
            map[b] = GFREE;
        }
    }
}

template <typename SRC>
__global__ __launch_bounds__(GW) void k_g_classify(SRC src, const BImg* __restrict__ imgs,
                                                   const int64_t* __restrict__ ts, int nimg, int64_t ntiles,
                                                   uint32_t* map, GItem* slots, uint32_t* nslots) {
    for (int64_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        const int j = find_seg(ts, nimg, b);
        const BImg& im = imgs[j];
        const int64_t t = b - ts[j];
        const int tx = (int)(t % im.wnx), ty = (int)((t / im.wnx) % im.wny),
                  tz = (int)(t / ((int64_t)im.wnx * im.wny));
        if (g_is3d(im.g)) g_classify_tile<SRC, true>(src, im, j, tx, ty, tz, b, map, slots, nslots);
        else g_classify_tile<SRC, false>(src, im, j, tx, ty, tz, b, map, slots, nslots);
    }
}

enum { G_FIRST = 0, G_ROUND = 1, G_CAPPED = 2 };

// Arguments shared by the solve kernels.
struct GArgs {
    const BImg* imgs;
    const uint32_t* map;
    uint8_t* scode;
    uint32_t* sL;
    int32_t* stamp;
    uint64_t* rec;        // per slot: the steps with an assignment of its last solve
    const int* caps;      // G_CAPPED: per image step cap, < 0: skip
    uint32_t* overflow;
    GItem* many;          // tiles with more than GKMAX values (for the per-voxel kernel)
    uint32_t* many_count;
};

// after a solve: queue the active neighbours in `m` for round + 1, record the steps
__device__ __forceinline__ void g_after(const GArgs& A, const BImg& im, int j, int tx, int ty, int tz, uint32_t slot,
                                        uint32_t m, int round, GItem* out_items, uint32_t* out_count,
                                        const GSteps& st, bool over) {
    const int lane = threadIdx.x;
    if (over && lane == 0) atomicOr(A.overflow, 1u);
    if (lane < 4) A.rec[(size_t)slot * 4 + lane] = lane == 0 ? st.w[0] : lane == 1 ? st.w[1] : lane == 2 ? st.w[2]
                                                                                                          : st.w[3];
    if (m && lane < 27 && ((m >> lane) & 1u)) {
        const int qx = tx + lane % 3 - 1, qy = ty + (lane / 3) % 3 - 1, qz = tz + lane / 9 - 1;
        if (qx >= 0 && qx < im.wnx && qy >= 0 && qy < im.wny && qz >= 0 && qz < im.wnz) {
            const int64_t q = im.wbase + ((int64_t)qz * im.wny + qy) * im.wnx + qx;
            if (A.map[q] != GFREE && atomicMax(A.stamp + q, round + 1) < round + 1) {
                const uint32_t pos = atomicAdd(out_count, 1u);
                out_items[pos] = GItem{g_pack(j, qx, qy, qz), q};
            }
        }
    }
}

template <bool D3, typename SRC, int MODE>
__device__ __forceinline__ void g_tile_bits(const SRC& src, const GArgs& A, GLdsB& s, const GItem& it, int round,
                                            GItem* out_items, uint32_t* out_count) {
    const int j = (int)(it.c >> 48);
    const int tz = (int)((it.c >> 32) & 0xFFFF), ty = (int)((it.c >> 16) & 0xFFFF), tx = (int)(it.c & 0xFFFF);
    const uint32_t slot = A.map[it.gt];
    const BImg& im = A.imgs[j];
    const int cap = MODE == G_CAPPED ? A.caps[j] : GMAXSTEP;
    const int lane = threadIdx.x;
    s.hkey[lane] = 0;
    if (lane == 0) s.nkeys = 0;
    if (MODE != G_FIRST) g_neighbour_slots(im, tx, ty, tz, A.map, s.nslot);
    else __syncthreads();
    GSinkB sink{s, 0u, 0};
    uint64_t zb, onp;
    g_load<D3>(src, im, tx, ty, tz, MODE != G_FIRST ? s.nslot : nullptr, A.scode, A.sL, sink, zb, onp);
    const int K = g_rank_values(s);
    if (K > GKMAX) {  // the per-voxel kernel takes this tile
        if (lane == 0) A.many[atomicAdd(A.many_count, 1u)] = it;
        return;
    }
    GSteps st;
    bool over = false;
    g_solve_bits<D3>(s, zb, onp, K, cap, st, &over);
    const GViewB v{s};
    g_output<D3>(src, v, im, tx, ty, tz);
    uint32_t m = 0;
    if (MODE != G_CAPPED) m = g_publish<D3>(v, im, tx, ty, tz, slot, A.scode, A.sL, MODE == G_FIRST);
    g_after(A, im, j, tx, ty, tz, slot, m, round, out_items, out_count, st, over);
}

template <bool D3, typename SRC, int MODE>
__device__ __forceinline__ void g_tile_vox(const SRC& src, const GArgs& A, GLdsV& s, const GItem& it, int round,
                                           GItem* out_items, uint32_t* out_count) {
    const int j = (int)(it.c >> 48);
    const int tz = (int)((it.c >> 32) & 0xFFFF), ty = (int)((it.c >> 16) & 0xFFFF), tx = (int)(it.c & 0xFFFF);
    const uint32_t slot = A.map[it.gt];
    const BImg& im = A.imgs[j];
    const int cap = MODE == G_CAPPED ? A.caps[j] : GMAXSTEP;
    if (MODE != G_FIRST) g_neighbour_slots(im, tx, ty, tz, A.map, s.nslot);
    GSinkV sink{s};
    uint64_t zb, onp;
    g_load<D3>(src, im, tx, ty, tz, MODE != G_FIRST ? s.nslot : nullptr, A.scode, A.sL, sink, zb, onp);
    GSteps st;
    bool over = false;
    g_solve_vox<D3>(s, zb, onp, cap, st, &over);
    const GViewV v{s};
    g_output<D3>(src, v, im, tx, ty, tz);
    uint32_t m = 0;
    if (MODE != G_CAPPED) m = g_publish<D3>(v, im, tx, ty, tz, slot, A.scode, A.sL, MODE == G_FIRST);
    g_after(A, im, j, tx, ty, tz, slot, m, round, out_items, out_count, st, over);
}

// Solves the tiles of a list that are of its kind (D3: 3D images, else 2D).
// G_FIRST: every active tile, halos from v0; G_ROUND: the queued tiles, halos from the
// stores; both write the output, publish the faces and queue the active neighbours
// whose halo changed (stamp = round + 1).  G_CAPPED: the tiles of images with caps[j]
// >= 0, steps capped there, output only.  Block 0 empties `reset` (the list two rounds
// back).  VOX: the per-voxel solve (the tiles that the value-sliced solve of the same
// pass found with more than GKMAX values).
template <typename SRC, int MODE, bool VOX, bool D3>
__global__ __launch_bounds__(GW) void k_g_solve(SRC src, GArgs A, const GItem* __restrict__ items,
                                                const uint32_t* __restrict__ count, int round, GItem* out_items,
                                                uint32_t* out_count, uint32_t* reset) {
    using Lds = typename std::conditional<VOX, GLdsV, GLdsB>::type;
    __shared__ Lds s;
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
    const uint32_t n = *count;
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
        const GItem it = items[e];
        const int j = (int)(it.c >> 48);
        if (A.map[it.gt] == GFREE || g_is3d(A.imgs[j].g) != D3) continue;
        if (MODE == G_CAPPED && A.caps[j] < 0) continue;
        __syncthreads();  // the previous tile's LDS reads are done
        if constexpr (VOX) g_tile_vox<D3, SRC, MODE>(src, A, s, it, round, out_items, out_count);
        else g_tile_bits<D3, SRC, MODE>(src, A, s, it, round, out_items, out_count);
    }
}

// the steps recorded by every active tile's last solve, OR-ed per image (a wavefront
// merges the slots of one image before its atomics)
__global__ __launch_bounds__(256) void k_g_steps(const GItem* __restrict__ slots, const uint32_t* __restrict__ nslots,
                                                 const uint64_t* __restrict__ rec, uint32_t* stepbm) {
    const uint32_t n = *nslots;
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const uint32_t e = base + threadIdx.x;
        const bool ok = e < n;
        const int j = ok ? (int)(slots[e].c >> 48) : -1;
        uint64_t w[4] = {0, 0, 0, 0};
        if (ok)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = rec[(size_t)e * 4 + k];
        bool todo = ok;
        while (__ballot(todo)) {
            const int lead = __ffsll((unsigned long long)__ballot(todo)) - 1;
            const int jl = __shfl(j, lead);
            const bool mine = todo && j == jl;
            uint64_t acc[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc[k] = mine ? w[k] : 0ull;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) acc[k] |= g_shfl64(acc[k], lane ^ o);
            }
            if (lane == lead)
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t word = (uint32_t)(acc[k >> 1] >> ((k & 1) * 32));
                    if (word) atomicOr(stepbm + (size_t)jl * 8 + k, word);
                }
            todo = todo && !mine;
        }
    }
}

__global__ void k_g_zero(uint32_t* p) { *p = 0; }

// ----------------------------------------------------------------- host driver
// one pass over a list: the value-sliced solve of the 2D and / or 3D tiles (`kinds`
// bit 0 / bit 1), then the per-voxel solve of the tiles it left
template <typename SRC, int MODE>
static void launch_pass(const SRC& src, const GArgs& A, int kinds, const GItem* items, const uint32_t* count,
                        unsigned grid, int round, GItem* out_items, uint32_t* out_count, uint32_t* reset,
                        hipStream_t s) {
    if (kinds & 1)
        k_g_solve<SRC, MODE, false, false><<<grid, GW, 0, s>>>(src, A, items, count, round, out_items, out_count,
                                                               reset);
    if (kinds & 2)
        k_g_solve<SRC, MODE, false, true><<<grid, GW, 0, s>>>(src, A, items, count, round, out_items, out_count,
                                                              reset);
    if (kinds & 1)
        k_g_solve<SRC, MODE, true, false><<<256, GW, 0, s>>>(src, A, A.many, A.many_count, round, out_items,
                                                             out_count, nullptr);
    if (kinds & 2)
        k_g_solve<SRC, MODE, true, true><<<256, GW, 0, s>>>(src, A, A.many, A.many_count, round, out_items,
                                                            out_count, nullptr);
    k_g_zero<<<1, 1, 0, s>>>(A.many_count);
}

template <typename SRC>
static int grow_run(const SRC& src, const BImg* dimgs, const int64_t* dts, int nimg, int64_t ntiles, int kinds,
                    int* iters, bool* overflow, hipStream_t s) {
    *overflow = false;
    if (ntiles <= 0) {
        if (iters)
            for (int j = 0; j < nimg; ++j) iters[j] = 1;
        return CLE_OK;
    }
    const size_t nt = (size_t)ntiles;
    // slots, 3 round lists, the many-values list (16 B entries), records (32 B per
    // slot), map, stamps, counters, per-image step bitmaps and caps
    Scratch mem, store;
    const size_t bytes = 5 * nt * 16 + nt * 32 + 2 * nt * 4 + 64 + (size_t)nimg * 36 + 256;
    CLE_TRY(mem.alloc(bytes, s));
    char* p = mem.get<char>();
    GItem* slots = (GItem*)p;
    GItem* lists = slots + nt;  // 3 x nt
    GItem* many = lists + 3 * nt;
    p += 5 * nt * 16;
    uint64_t* rec = (uint64_t*)p;
    p += nt * 32;
    uint32_t* map = (uint32_t*)p;
    p += nt * 4;
    int32_t* stamp = (int32_t*)p;
    p += nt * 4;
    uint32_t* cnt = (uint32_t*)p;  // [0..2] round lists, [3] slots, [4] overflow, [5] many
    p += 64;
    uint32_t* stepbm = (uint32_t*)p;
    p += (size_t)nimg * 32;
    int* caps = (int*)p;
    // the face store for every tile that can be active (no host round trip to size it)
    CLE_TRY(store.alloc(nt * GSLOT * 5, s));
    uint32_t* sL = store.get<uint32_t>();
    uint8_t* scode = (uint8_t*)(sL + nt * GSLOT);
    CLE_TRY(fill_u32((uint32_t*)stamp, 0u, nt, s));
    CLE_TRY(fill_u32(cnt, 0u, 16 + (size_t)nimg * 8, s));  // counters, then the step bitmaps
    const GArgs A{dimgs, map, scode, sL, stamp, rec, caps, cnt + 4, many, cnt + 5};

    k_g_classify<SRC><<<(unsigned)(ntiles < 65536 ? ntiles : 65536), GW, 0, s>>>(src, dimgs, dts, nimg, ntiles, map,
                                                                               slots, cnt + 3);
    CLE_LAUNCHED();
    const unsigned grid = (unsigned)(ntiles < 65536 ? ntiles : 65536);
    auto list = [&](int k) { return lists + (size_t)(k % 3) * nt; };
    auto count = [&](int k) { return cnt + (k % 3); };
    // round 1: every active tile from v0 halos; queues into list 2 (round 2)
    launch_pass<SRC, G_FIRST>(src, A, kinds, slots, cnt + 3, grid, 1, list(2), count(2), count(0), s);
    CLE_LAUNCHED();
    // rounds k >= 2 read list k % 3, queue into list (k + 1) % 3, empty list (k + 2) % 3;
    // two rounds per check of the queue
    for (int k = 2;; k += 2) {
        // every round lowers a step or settles a value; a bound far above any real
        // image guards the host loop
        if (k > (1 << 16)) return fail(CLE_E_HIP, "Voronoi tile growth did not settle after %d rounds", k);
        uint32_t c = 0;
        CLE_HIP(hipMemcpyAsync(&c, count(k), 4, hipMemcpyDeviceToHost, s));
        CLE_HIP(hipStreamSynchronize(s));
        if (c == 0) break;
        launch_pass<SRC, G_ROUND>(src, A, kinds, list(k), count(k), c < grid ? c : grid, k, list(k + 1), count(k + 1),
                                  count(k + 2), s);
        launch_pass<SRC, G_ROUND>(src, A, kinds, list(k + 1), count(k + 1), grid, k + 1, list(k + 2), count(k + 2),
                                  count(k + 3), s);
        CLE_LAUNCHED();
    }
    // the steps of every tile's last solve, per image
    const size_t sb = (nt + 255) / 256;
    k_g_steps<<<(unsigned)(sb < 4096 ? sb : 4096), 256, 0, s>>>(slots, cnt + 3, rec, stepbm);
    CLE_LAUNCHED();
    std::vector<uint32_t> h((size_t)nimg * 8 + 8);
    CLE_HIP(hipMemcpyAsync(h.data(), cnt + 4, 4, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipMemcpyAsync(h.data() + 8, stepbm, (size_t)nimg * 32, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    if (h[0]) {
        *overflow = true;
        return CLE_OK;
    }
    // the reference stops at the first step without an assignment
    std::vector<int> hc((size_t)nimg, -1);
    bool redo = false;
    for (int j = 0; j < nimg; ++j) {
        const uint32_t* w = h.data() + 8 + (size_t)j * 8;
        int S = 0;
        while (S < 256 && ((w[S >> 5] >> (S & 31)) & 1u)) ++S;
        bool later = false;
        for (int t = S; t < 256; ++t) later |= ((w[t >> 5] >> (t & 31)) & 1u) != 0;
        if (later) {
            hc[j] = S;
            redo = true;
        }
        if (iters) iters[j] = S + 1;
    }
    if (redo) {
        CLE_HIP(hipMemcpyAsync(caps, hc.data(), (size_t)nimg * 4, hipMemcpyHostToDevice, s));
        launch_pass<SRC, G_CAPPED>(src, A, kinds, slots, cnt + 3, grid, 0, nullptr, nullptr, nullptr, s);
        CLE_LAUNCHED();
        CLE_HIP(hipStreamSynchronize(s));  // `hc` (host) is read by the copy above
    }
    return CLE_OK;
}

__global__ void k_g_put(BImg* d, int64_t* ts, BImg im, int64_t n) {
    *d = im;
    ts[0] = 0;
    ts[1] = n;
}

// one image: its table entry (with the destination `dst`) and tile prefix on the device
static int grow_single_table(const Geom& g, DBuf dst, Scratch& tab, int64_t* ntiles, hipStream_t s) {
    BImg im;
    memset(&im, 0, sizeof(BImg));
    im.g = g;
    im.dst = dst.p;
    im.dst_dt = dst.dt;
    grow_tiling(im);
    if (im.wnx > 65535 || im.wny > 65535 || im.wnz > 65535)
        return fail(CLE_E_VALUE, "image too large for the growth tiling (%d x %d x %d tiles)", im.wnx, im.wny, im.wnz);
    *ntiles = grow_tile_count(im);
    CLE_TRY(tab.alloc(sizeof(BImg) + 32, s));
    k_g_put<<<1, 1, 0, s>>>(tab.get<BImg>(), (int64_t*)(tab.get<char>() + sizeof(BImg)), im, *ntiles);
    CLE_LAUNCHED();
    return CLE_OK;
}

int grow_masked(DBuf mask, const float* lab, DBuf dst, const Geom& g, int* iters, bool* overflow, hipStream_t s) {
    Scratch tab;
    int64_t nt = 0;
    CLE_TRY(grow_single_table(g, dst, tab, &nt, s));
    const BImg* d = tab.get<BImg>();
    const int64_t* ts = (const int64_t*)(tab.get<char>() + sizeof(BImg));
    if (mask.dt == CLE_U8) return grow_run(GSrcSegLab{(const uint8_t*)mask.p, lab}, d, ts, 1, nt, g_is3d(g) ? 2 : 1, iters, overflow, s);
    return grow_run(GSrcMaskLab{mask, lab, dst}, d, ts, 1, nt, g_is3d(g) ? 2 : 1, iters, overflow, s);
}

int grow_labels(DBuf src, DBuf dst, const Geom& g, int* iters, bool* overflow, hipStream_t s) {
    Scratch tab;
    int64_t nt = 0;
    CLE_TRY(grow_single_table(g, dst, tab, &nt, s));
    return grow_run(GSrcLabels{src, dst}, tab.get<BImg>(), (const int64_t*)(tab.get<char>() + sizeof(BImg)), 1, nt,
                    g_is3d(g) ? 2 : 1, iters, overflow, s);
}

int grow_batch(const BImg* dimgs, const int64_t* dts, int nimg, int64_t ntiles, int kinds, const uint8_t* seg,
               const float* lab, int* iters, bool* overflow, hipStream_t s) {
    return grow_run(GSrcSegLab{seg, lab}, dimgs, dts, nimg, ntiles, kinds, iters, overflow, s);
}

}  // namespace cle
