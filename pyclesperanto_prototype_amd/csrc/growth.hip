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
// exactly in LDS from its halo: the halo voxels' (t, value) are replayed, each turning
// "on" at its step, and the tile runs the reference's steps on bit masks of its rows
// (dilating the on-set with shifts and ORs of 64-bit rows), taking the label maximum
// only for the voxels assigned in a step.  A tile publishes the (t, value) of its face
// voxels; when they change, the neighbouring tiles that read them are queued and
// solved again.  The (t) of every solve are upper bounds that only decrease, and once
// they are final the values settle in order of t, so the rounds end, at the unique
// solution.  A final pass solves every active tile once more from the final halos,
// writes the output and records the steps at which voxels were assigned; a gap in
// those steps below their maximum is the reference's stop, and the tiles of that image
// are solved once more with the steps capped there.
//
// Steps are held in bytes (code = t + 2 of the step it turns on: 0 seed, 1 .. 254
// assigned at step code - 1, 255 never); a tile that would need a step >= 254 reports
// an overflow and the caller runs the step-synchronous growth instead.
#include "common.hpp"
#include "growth.hpp"

namespace cle {

constexpr int GW = 64;                 // threads per tile (one wavefront)
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

struct GLds {
    uint32_t L[4096];    // value (float32 bits) of every region voxel
    uint8_t code[4096];  // when the voxel is on: 0 from the start, c from step c, 255 never
    uint64_t on[64];     // on-set of every region row before the current step
    uint64_t zb[64];     // interior zero voxels (load only)
    uint64_t pend[64];   // halo voxels that turn on at a later step (load only)
    uint64_t nw[64];     // voxels assigned in the current step
    int pf[65];          // exclusive prefix of their counts over the rows
};

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

// region of tile (tx, ty, tz) into LDS; lane = region x.  Halo voxels come from the
// neighbours' stores (when `store_halo` and the neighbour is an active tile) or from
// v0; only interior voxels can be zero (growable).
template <bool D3, typename SRC>
__device__ __forceinline__ void g_load(const SRC& src, const BImg& im, int tx, int ty, int tz, bool store_halo,
                                       const uint32_t* __restrict__ map, const uint8_t* __restrict__ scode,
                                       const uint32_t* __restrict__ sL, GLds& s) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const Geom& g = im.g;
    const int lane = threadIdx.x;
    const int64_t X = (int64_t)tx * GX - 1 + lane;
    const bool xin = X >= 0 && X < g.W;
    const int64_t Y0 = (int64_t)ty * TY - 1, Z0 = D3 ? (int64_t)tz * TZ - 1 : 0;
    const int dx = lane == 0 ? -1 : (lane == GW - 1 ? 1 : 0);
#pragma unroll 4
    for (int r = 0; r < 64; ++r) {
        const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
        const int64_t Y = Y0 + ry, Z = D3 ? Z0 + rz : 0;
        const bool in = xin && Y >= 0 && Y < g.H && Z >= 0 && Z < g.D;
        const int dy = ry == 0 ? -1 : (ry == TY + 1 ? 1 : 0);
        const int dz = D3 ? (rz == 0 ? -1 : (rz == TZ + 1 ? 1 : 0)) : 0;
        uint32_t lv = 0;
        uint8_t c = GNONE;
        bool zero = false;
        if (in) {
            bool from_v0 = true;
            if (store_halo && (dx | dy | dz)) {
                const int64_t nt = im.wbase + ((int64_t)(tz + dz) * im.wny + (ty + dy)) * im.wnx + (tx + dx);
                const uint32_t slot = map[nt];
                if (slot != GFREE) {
                    from_v0 = false;
                    const int lx = dx < 0 ? GX - 1 : (dx > 0 ? 0 : lane - 1);
                    const int ly = dy < 0 ? TY - 1 : (dy > 0 ? 0 : ry - 1);
                    const int lz = D3 ? (dz < 0 ? TZ - 1 : (dz > 0 ? 0 : rz - 1)) : 0;
                    const int64_t a = (int64_t)slot * GSLOT + g_face_entry<D3>(dx, dy, dz, lx, ly, lz);
                    c = scode[a];
                    lv = sL[a];
                }
            }
            if (from_v0) {
                const float v = src.v0(im, (Z * g.H + Y) * g.W + X);
                lv = __float_as_uint(v);
                if (v > 0.f) c = 0;
                zero = v == 0.f && !(dx | dy | dz);
            }
        }
        s.L[r * 64 + lane] = lv;
        s.code[r * 64 + lane] = c;
        const uint64_t zb = __ballot(zero), ob = __ballot(c == 0), pb = __ballot(c != 0 && c != GNONE);
        if (lane == 0) {
            s.zb[r] = zb;
            s.on[r] = ob;
            s.pend[r] = pb;
        }
    }
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

// The reference's steps on the tile (lane = region row).  Steps T >= cap are not run
// (cap < GMAXSTEP: a capped solve).  Returns the steps with an assignment in st[0..3]
// (same in every lane); *over is set when a step >= GMAXSTEP would assign a voxel.
template <bool D3>
__device__ __forceinline__ void g_solve(GLds& s, int cap, uint64_t* st, bool* over) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const int r = threadIdx.x;
    const int ry = D3 ? (r & 7) : r, rz = D3 ? (r >> 3) : 1;
    const bool irow = ry >= 1 && ry <= TY && rz >= 1 && rz <= TZ;
    uint64_t on = s.on[r], zero = irow ? (s.zb[r] & GROW_INT) : 0ull, pend = s.pend[r];
    int nact = 255;
    for (uint64_t p = pend; p; p &= p - 1) {
        const int c = s.code[r * 64 + __ffsll((unsigned long long)p) - 1];
        nact = c < nact ? c : nact;
    }
    st[0] = st[1] = st[2] = st[3] = 0;
    int T = 0;
    for (;;) {
        if (T >= cap && cap < GMAXSTEP) break;
        if (pend && nact <= T) {  // halo voxels turning on at this step
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
                                       (q - 8) * 64 + b, (q + 8) * 64 + b};
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
            st[T >> 6] |= 1ull << (T & 63);
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

// one face voxel: publish (code, value), report a change
template <bool D3>
__device__ __forceinline__ uint32_t g_pub(const GLds& s, const Geom& g, int64_t X0, int64_t Y0, int64_t Z0, int lx,
                                          int ly, int lz, int64_t a, uint8_t* scode, uint32_t* sL, bool first) {
    if (X0 + lx >= g.W || Y0 + ly >= g.H || Z0 + lz >= g.D) return 0;
    const int ri = D3 ? ((lz + 1) * 8 + ly + 1) * 64 + lx + 1 : (ly + 1) * 64 + lx + 1;
    const uint8_t c = s.code[ri];
    const uint32_t l = s.L[ri];
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
template <bool D3>
__device__ __forceinline__ uint32_t g_publish(const GLds& s, const BImg& im, int tx, int ty, int tz, uint32_t slot,
                                              uint8_t* scode, uint32_t* sL, bool first) {
    constexpr int TY = D3 ? GT3 : GT2, TZ = D3 ? GT3 : 1;
    const Geom& g = im.g;
    const int64_t X0 = (int64_t)tx * GX, Y0 = (int64_t)ty * TY, Z0 = D3 ? (int64_t)tz * TZ : 0;
    const int64_t base = (int64_t)slot * GSLOT;
    const int lane = threadIdx.x;
    uint32_t m = 0;
    if (D3) {
        for (int e = lane; e < GT3 * GT3; e += GW) {
            const int ly = e % GT3, lz = e / GT3;
            m |= g_pub<true>(s, g, X0, Y0, Z0, 0, ly, lz, base + GF3_X0 + e, scode, sL, first);
            m |= g_pub<true>(s, g, X0, Y0, Z0, GX - 1, ly, lz, base + GF3_X1 + e, scode, sL, first);
        }
        for (int e = lane; e < GX * GT3; e += GW) {
            const int lx = e % GX, o = e / GX;
            m |= g_pub<true>(s, g, X0, Y0, Z0, lx, 0, o, base + GF3_Y0 + e, scode, sL, first);
            m |= g_pub<true>(s, g, X0, Y0, Z0, lx, GT3 - 1, o, base + GF3_Y1 + e, scode, sL, first);
            m |= g_pub<true>(s, g, X0, Y0, Z0, lx, o, 0, base + GF3_Z0 + e, scode, sL, first);
            m |= g_pub<true>(s, g, X0, Y0, Z0, lx, o, GT3 - 1, base + GF3_Z1 + e, scode, sL, first);
        }
    } else {
        for (int e = lane; e < GX; e += GW) {
            m |= g_pub<false>(s, g, X0, Y0, Z0, 0, e, 0, base + GF2_X0 + e, scode, sL, first);
            m |= g_pub<false>(s, g, X0, Y0, Z0, GX - 1, e, 0, base + GF2_X1 + e, scode, sL, first);
            m |= g_pub<false>(s, g, X0, Y0, Z0, e, 0, 0, base + GF2_Y0 + e, scode, sL, first);
            m |= g_pub<false>(s, g, X0, Y0, Z0, e, GT2 - 1, 0, base + GF2_Y1 + e, scode, sL, first);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m |= (uint32_t)__shfl_xor((int)m, o);
    return m & ~(1u << 13);
}

// the tile's output, lane = region x
template <bool D3, typename SRC>
__device__ __forceinline__ void g_output(const SRC& src, const GLds& s, const BImg& im, int tx, int ty, int tz) {
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
            src.out(im, (Z * g.H + Y) * g.W + X, __uint_as_float(s.L[ri]));
        }
    }
}

// Every tile: its voxels' output from v0 (final for tiles without a zero voxel; the
// final pass overwrites the others), and a slot in the store for tiles holding a zero
// voxel.  One wavefront per tile, lane = x.
template <typename SRC>
__global__ __launch_bounds__(GW) void k_g_classify(SRC src, const BImg* __restrict__ imgs,
                                                   const int64_t* __restrict__ ts, int nimg, uint32_t* map,
                                                   GItem* slots, uint32_t* nslots) {
    const int64_t b = blockIdx.x;
    const int j = find_seg(ts, nimg, b);
    const BImg& im = imgs[j];
    const Geom& g = im.g;
    const int64_t t = b - ts[j];
    const int tx = (int)(t % im.wnx), ty = (int)((t / im.wnx) % im.wny), tz = (int)(t / ((int64_t)im.wnx * im.wny));
    const bool d3 = g_is3d(g);
    const int TY = d3 ? GT3 : GT2, TZ = d3 ? GT3 : 1;
    const int lane = threadIdx.x;
    const int64_t X = (int64_t)tx * GX + lane;
    bool zero = false;
    if (lane < GX && X < g.W) {
        for (int lz = 0; lz < TZ; ++lz) {
            const int64_t Z = d3 ? (int64_t)tz * TZ + lz : 0;
            if (Z >= g.D) break;
            for (int ly = 0; ly < TY; ++ly) {
                const int64_t Y = (int64_t)ty * TY + ly;
                if (Y >= g.H) break;
                const int64_t i = (Z * g.H + Y) * g.W + X;
                const float v = src.v0(im, i);
                zero |= v == 0.f;
                src.out(im, i, v);
            }
        }
    }
    const bool active = __ballot(zero) != 0;
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

enum { G_FIRST = 0, G_ROUND = 1, G_FINAL = 2 };

// Solves the tiles of a list.  G_FIRST: every active tile, halos from v0; G_ROUND: the
// queued tiles, halos from the stores; both publish their faces and queue the active
// neighbours whose halo changed (stamp = round + 1).  G_FINAL: every active tile of an
// image with caps[j] >= 0, steps capped at caps[j], output written, steps recorded.
// Block 0 empties `reset` (the list two rounds back).
template <typename SRC, int MODE>
__global__ __launch_bounds__(GW) void k_g_solve(SRC src, const BImg* __restrict__ imgs, const GItem* __restrict__ items,
                                                const uint32_t* __restrict__ count, const uint32_t* __restrict__ map,
                                                uint8_t* scode, uint32_t* sL, int round, int32_t* stamp,
                                                GItem* out_items, uint32_t* out_count, uint32_t* reset,
                                                const int* __restrict__ caps, uint32_t* stepbm, uint32_t* overflow) {
    __shared__ GLds s;
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
    const uint32_t n = *count;
    const int lane = threadIdx.x;
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
        const GItem it = items[e];
        const int j = (int)(it.c >> 48);
        const int tz = (int)((it.c >> 32) & 0xFFFF), ty = (int)((it.c >> 16) & 0xFFFF), tx = (int)(it.c & 0xFFFF);
        const uint32_t slot = map[it.gt];
        if (slot == GFREE) continue;
        const int cap = MODE == G_FINAL ? caps[j] : GMAXSTEP;
        if (cap < 0) continue;
        const BImg& im = imgs[j];
        const bool d3 = g_is3d(im.g);
        uint64_t st[4];
        bool over = false;
        uint32_t m = 0;
        __syncthreads();  // the previous tile's LDS reads are done
        if (d3) {
            g_load<true>(src, im, tx, ty, tz, MODE != G_FIRST, map, scode, sL, s);
            g_solve<true>(s, cap, st, &over);
            if (MODE == G_FINAL) g_output<true>(src, s, im, tx, ty, tz);
            else m = g_publish<true>(s, im, tx, ty, tz, slot, scode, sL, MODE == G_FIRST);
        } else {
            g_load<false>(src, im, tx, ty, tz, MODE != G_FIRST, map, scode, sL, s);
            g_solve<false>(s, cap, st, &over);
            if (MODE == G_FINAL) g_output<false>(src, s, im, tx, ty, tz);
            else m = g_publish<false>(s, im, tx, ty, tz, slot, scode, sL, MODE == G_FIRST);
        }
        if (over && lane == 0) atomicOr(overflow, 1u);
        if (MODE == G_FINAL) {
            if (lane < 8) {
                const uint32_t w = (uint32_t)(st[lane >> 1] >> ((lane & 1) * 32));
                if (w) atomicOr(stepbm + (size_t)j * 8 + lane, w);
            }
        } else if (m && lane < 27 && ((m >> lane) & 1u)) {
            const int qx = tx + lane % 3 - 1, qy = ty + (lane / 3) % 3 - 1, qz = tz + lane / 9 - 1;
            if (qx >= 0 && qx < im.wnx && qy >= 0 && qy < im.wny && qz >= 0 && qz < im.wnz) {
                const int64_t q = im.wbase + ((int64_t)qz * im.wny + qy) * im.wnx + qx;
                if (map[q] != GFREE && atomicMax(stamp + q, round + 1) < round + 1) {
                    const uint32_t pos = atomicAdd(out_count, 1u);
                    out_items[pos] = GItem{g_pack(j, qx, qy, qz), q};
                }
            }
        }
    }
}

// ----------------------------------------------------------------- sources / outputs
// voronoi_otsu pipeline and batches: v0 = (segmentation - 1) + seed label (float32, as
// the reference's add_image_and_scalar then add_images), output masked by the
// segmentation into the image's destination
struct GSrcSegLab {
    const uint8_t* seg;
    const float* lab;
    __device__ __forceinline__ float v0(const BImg& im, int64_t i) const {
        return ((float)seg[im.off + i] - 1.f) + lab[im.off + i];
    }
    __device__ __forceinline__ void out(const BImg& im, int64_t i, float v) const {
        st(DBuf{im.dst, im.dst_dt}, i, seg[im.off + i] ? v : 0.f);
    }
};
// masked_voronoi_labeling with a mask of any type (one image)
struct GSrcMaskLab {
    DBuf mask;
    const float* lab;
    DBuf dst;
    __device__ __forceinline__ float v0(const BImg&, int64_t i) const { return (ld(mask, i) - 1.f) + lab[i]; }
    __device__ __forceinline__ void out(const BImg&, int64_t i, float v) const {
        st(dst, i, ld(mask, i) != 0.f ? v : 0.f);
    }
};
// extend_labeling_via_voronoi: the labels converted to float32, no mask (one image)
struct GSrcLabels {
    DBuf src;
    DBuf dst;
    __device__ __forceinline__ float v0(const BImg&, int64_t i) const { return ld(src, i); }
    __device__ __forceinline__ void out(const BImg&, int64_t i, float v) const { st(dst, i, v); }
};

// ----------------------------------------------------------------- host driver
struct GScratch {
    Scratch mem, store;
    uint32_t* map = nullptr;
    int32_t* stamp = nullptr;
    GItem* slots = nullptr;
    GItem* lists = nullptr;  // 3 x ntiles
    uint32_t* cnt = nullptr;  // [0..2] list counts, [3] slots, [4] overflow
    int* caps = nullptr;
    uint32_t* stepbm = nullptr;
    uint8_t* scode = nullptr;
    uint32_t* sL = nullptr;
};

template <typename SRC>
static int grow_run(const SRC& src, const BImg* dimgs, const int64_t* dts, int nimg, int64_t ntiles, int* iters,
                    bool* overflow, hipStream_t s) {
    *overflow = false;
    if (ntiles <= 0) {
        if (iters)
            for (int j = 0; j < nimg; ++j) iters[j] = 1;
        return CLE_OK;
    }
    GScratch G;
    const size_t nt = (size_t)ntiles;
    const size_t bytes = nt * 4 + nt * 4 + nt * 16 + 3 * nt * 16 + 64 + (size_t)nimg * 4 + (size_t)nimg * 32 + 256;
    CLE_TRY(G.mem.alloc(bytes, s));
    char* p = G.mem.get<char>();
    G.slots = (GItem*)p;
    p += nt * 16;
    G.lists = (GItem*)p;
    p += 3 * nt * 16;
    G.map = (uint32_t*)p;
    p += nt * 4;
    G.stamp = (int32_t*)p;
    p += nt * 4;
    G.cnt = (uint32_t*)p;  // 16 words
    p += 64;
    G.stepbm = (uint32_t*)p;
    p += (size_t)nimg * 32;
    G.caps = (int*)p;
    // the store for every tile that can be active (no host round trip to size it)
    CLE_TRY(G.store.alloc(nt * GSLOT * 5, s));
    G.sL = G.store.get<uint32_t>();
    G.scode = (uint8_t*)(G.sL + nt * GSLOT);
    CLE_TRY(fill_u32((uint32_t*)G.stamp, 0u, nt, s));
    CLE_TRY(fill_u32(G.cnt, 0u, 16 + (size_t)nimg * 8, s));  // counts, then the step bitmaps
    CLE_TRY(fill_u32((uint32_t*)G.caps, (uint32_t)GMAXSTEP, (size_t)nimg, s));

    k_g_classify<SRC><<<(unsigned)ntiles, GW, 0, s>>>(src, dimgs, dts, nimg, G.map, G.slots, G.cnt + 3);
    CLE_LAUNCHED();
    const unsigned grid = (unsigned)(ntiles < 16384 ? ntiles : 16384);
    auto list = [&](int k) { return G.lists + (size_t)(k % 3) * nt; };
    auto count = [&](int k) { return G.cnt + (k % 3); };
    // round 1: every active tile from v0 halos; queues into list 2 (round 2)
    k_g_solve<SRC, G_FIRST><<<grid, GW, 0, s>>>(src, dimgs, G.slots, G.cnt + 3, G.map, G.scode, G.sL, 1, G.stamp,
                                                list(2), count(2), count(0), G.caps, G.stepbm, G.cnt + 4);
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
        for (int kk = k; kk < k + 2; ++kk) {
            const unsigned gr = (unsigned)((kk == k ? c : ntiles) < grid ? (kk == k ? c : ntiles) : grid);
            k_g_solve<SRC, G_ROUND><<<gr, GW, 0, s>>>(src, dimgs, list(kk), count(kk), G.map, G.scode, G.sL, kk,
                                                      G.stamp, list(kk + 1), count(kk + 1), count(kk + 2), G.caps,
                                                      G.stepbm, G.cnt + 4);
        }
        CLE_LAUNCHED();
    }
    // final pass: output of every active tile, steps with an assignment per image
    k_g_solve<SRC, G_FINAL><<<grid, GW, 0, s>>>(src, dimgs, G.slots, G.cnt + 3, G.map, G.scode, G.sL, 0, G.stamp,
                                                nullptr, nullptr, nullptr, G.caps, G.stepbm, G.cnt + 4);
    CLE_LAUNCHED();
    std::vector<uint32_t> h((size_t)nimg * 8 + 8);
    CLE_HIP(hipMemcpyAsync(h.data(), G.cnt + 4, 4, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipMemcpyAsync(h.data() + 8, G.stepbm, (size_t)nimg * 32, hipMemcpyDeviceToHost, s));
    CLE_HIP(hipStreamSynchronize(s));
    if (h[0]) {
        *overflow = true;
        return CLE_OK;
    }
    // the reference stops at the first step without an assignment
    std::vector<int> caps((size_t)nimg, -1);
    bool redo = false;
    for (int j = 0; j < nimg; ++j) {
        const uint32_t* w = h.data() + 8 + (size_t)j * 8;
        int S = 0;
        while (S < 256 && ((w[S >> 5] >> (S & 31)) & 1u)) ++S;
        bool later = false;
        for (int t = S; t < 256; ++t) later |= ((w[t >> 5] >> (t & 31)) & 1u) != 0;
        if (later) {
            caps[j] = S;
            redo = true;
        }
        if (iters) iters[j] = S + 1;
    }
    if (redo) {
        Scratch ct;
        CLE_TRY(ct.alloc((size_t)nimg * 4, s));
        CLE_HIP(hipMemcpyAsync(ct.p, caps.data(), (size_t)nimg * 4, hipMemcpyHostToDevice, s));
        k_g_solve<SRC, G_FINAL><<<grid, GW, 0, s>>>(src, dimgs, G.slots, G.cnt + 3, G.map, G.scode, G.sL, 0, G.stamp,
                                                    nullptr, nullptr, nullptr, ct.get<int>(), G.stepbm, G.cnt + 4);
        CLE_LAUNCHED();
        CLE_HIP(hipStreamSynchronize(s));  // `caps` (host) is read by the copy above
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
    if (mask.dt == CLE_U8) return grow_run(GSrcSegLab{(const uint8_t*)mask.p, lab}, d, ts, 1, nt, iters, overflow, s);
    return grow_run(GSrcMaskLab{mask, lab, dst}, d, ts, 1, nt, iters, overflow, s);
}

int grow_labels(DBuf src, DBuf dst, const Geom& g, int* iters, bool* overflow, hipStream_t s) {
    Scratch tab;
    int64_t nt = 0;
    CLE_TRY(grow_single_table(g, dst, tab, &nt, s));
    return grow_run(GSrcLabels{src, dst}, tab.get<BImg>(), (const int64_t*)(tab.get<char>() + sizeof(BImg)), 1, nt,
                    iters, overflow, s);
}

int grow_batch(const BImg* dimgs, const int64_t* dts, int nimg, int64_t ntiles, const uint8_t* seg, const float* lab,
               int* iters, bool* overflow, hipStream_t s) {
    return grow_run(GSrcSegLab{seg, lab}, dimgs, dts, nimg, ntiles, iters, overflow, s);
}

}  // namespace cle
