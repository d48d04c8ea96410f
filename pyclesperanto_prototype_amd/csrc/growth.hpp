// Exact Voronoi growth by asynchronous tile solves (growth.hip): tile shape, the layout
// of a tile's published faces, and the entry points used by the pipelines.
#pragma once
#include "common.hpp"

namespace cle {

constexpr int GX = 62;        // tile extent along x (region rows of 64 voxels with the halo)
constexpr int GT3 = 6;        // 3D tiles: extent along y and z (8 x 8 region rows)
constexpr int GT2 = 62;       // 2D tiles: extent along y (64 region rows)
constexpr int GMAXSTEP = 254; // steps 0 .. 253 fit the byte codes

// published faces of a tile: (code, value) per face voxel, at these entry offsets
constexpr int GF3_X0 = 0, GF3_X1 = GT3 * GT3, GF3_Y0 = 2 * GT3 * GT3, GF3_Y1 = GF3_Y0 + GX * GT3,
              GF3_Z0 = GF3_Y1 + GX * GT3, GF3_Z1 = GF3_Z0 + GX * GT3;
constexpr int GF2_X0 = 0, GF2_X1 = GT2, GF2_Y0 = 2 * GT2, GF2_Y1 = 2 * GT2 + GX;
constexpr int GSLOT = 1568;   // entries per tile (3D: 1560, 2D: 248)
static_assert(GF3_Z1 + GX * GT3 <= GSLOT && GF2_Y1 + GX <= GSLOT, "face store");

void grow_tiling(BImg& im);
int64_t grow_tile_count(const BImg& im);

// masked growth of one image: v0 = (mask - 1) + lab (float32), output = mask != 0 ?
// grown : 0 into dst.  *overflow: the growth needs more than GMAXSTEP steps (nothing
// reliable was written; run the step-synchronous growth).  iters: the reference's
// iteration count.
int grow_masked(DBuf mask, const float* lab, DBuf dst, const Geom& g, int* iters, bool* overflow, hipStream_t s);
// unmasked growth of a label image (extend_labeling_via_voronoi) into dst
int grow_labels(DBuf src, DBuf dst, const Geom& g, int* iters, bool* overflow, hipStream_t s);
// batch: images of `dimgs` (device table with wbase / wn* set, packed at .off, output
// to .dst), tile prefix `dts` (nimg + 1 entries), kinds: bit 0 some image is 2D, bit 1
// some image is 3D; v0 = (seg - 1) + lab
int grow_batch(const BImg* dimgs, const int64_t* dts, int nimg, int64_t ntiles, int kinds, const uint8_t* seg,
               const float* lab, int* iters, bool* overflow, hipStream_t s);

}  // namespace cle
