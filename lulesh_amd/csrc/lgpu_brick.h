// lgpu_brick.h -- host-side set-up of the structured-brick path: recognising
// the reference's brick and boundary set-up in a mesh description, and the
// brick's tiling (workgroup tiles and z-chunks of the fused kernels).
#pragma once

#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <vector>

#include "lgpu_kernels.h"
#include "lulesh_gpu.h"

namespace lgpu {

inline FastDiv make_fastdiv(uint32_t d)
{
   FastDiv f;
   f.d = d;
   f.m = UINT64_MAX / d + 1;
   return f;
}

// True when the description is the reference's structured brick: lattice
// element-to-node map (lulesh-init.cc:245-266) and node-to-corner lists in
// ascending corner order (lulesh-init.cc:271-319).
inline bool is_reference_brick(const lgpu_mesh_desc *d)
{
   const int64_t s = d->sizeX;
   if (s < 1 || d->sizeY != s || d->sizeZ != s) return false;
   const int64_t en = s + 1;
   if (d->numElem != s * s * s || d->numNode != en * en * en) return false;
   int64_t e = 0;
   for (int64_t c = 0; c < s; ++c)
      for (int64_t b = 0; b < s; ++b)
         for (int64_t a = 0; a < s; ++a, ++e) {
            const int64_t n0 = (c * en + b) * en + a;
            const int64_t want[8] = {n0, n0 + 1, n0 + en + 1, n0 + en, n0 + en * en,
                                     n0 + en * en + 1, n0 + en * en + en + 1, n0 + en * en + en};
            for (int j = 0; j < 8; ++j)
               if (d->nodelist[8 * e + j] != want[j]) return false;
         }
   // Corner lists: every node's corners in ascending order.
   const int64_t N = d->numNode;
   std::vector<int32_t> cnt((size_t)N, 0);
   for (int64_t c = 0; c < 8 * (int64_t)d->numElem; ++c) cnt[(size_t)d->nodelist[c]]++;
   int64_t start = 0;
   for (int64_t n = 0; n < N; ++n) {
      if (d->nodeElemStart[n] != start) return false;
      start += cnt[(size_t)n];
   }
   if (d->nodeElemStart[N] != start) return false;
   std::fill(cnt.begin(), cnt.end(), 0);
   for (int64_t c = 0; c < 8 * (int64_t)d->numElem; ++c) {
      const int32_t n = d->nodelist[c];
      if (d->nodeElemCornerList[d->nodeElemStart[n] + cnt[(size_t)n]++] != (int32_t)c) return false;
   }
   return true;
}

// True when the face neighbours and boundary masks are those of the brick
// with symmetry planes on the low faces and free surfaces on the high faces
// (the reference's single-domain set-up, lulesh-init.cc:340-398, 497-573), so
// that k_eos can compute them (brick_faces in lgpu_step.inc).  A neighbour
// index has to match only where the mask lets k_eos read it (bits 0 or COMM).
inline bool brick_faces_computable(const lgpu_mesh_desc *d)
{
   const int64_t s = d->sizeX, s2 = s * s;
   const int32_t *nb[6] = {d->lxim, d->lxip, d->letam, d->letap, d->lzetam, d->lzetap};
   const int32_t field[6] = {0x00007, 0x00038, 0x001c0, 0x00e00, 0x07000, 0x38000};
   const int32_t comm[6] = {0x00004, 0x00020, 0x00100, 0x00800, 0x04000, 0x20000};
   int64_t e = 0;
   for (int64_t c = 0; c < s; ++c)
      for (int64_t b = 0; b < s; ++b)
         for (int64_t a = 0; a < s; ++a, ++e) {
            const int32_t want = (a == 0 ? 0x00001 : 0) | (a == s - 1 ? 0x00010 : 0) |
                                 (b == 0 ? 0x00040 : 0) | (b == s - 1 ? 0x00400 : 0) |
                                 (c == 0 ? 0x01000 : 0) | (c == s - 1 ? 0x10000 : 0);
            if (d->elemBC[e] != want) return false;
            const int64_t wantN[6] = {e - 1, e + 1, e - s, e + s, e - s2, e + s2};
            for (int f = 0; f < 6; ++f) {
               const int32_t bits = want & field[f];
               if ((bits == 0 || bits == comm[f]) && nb[f][e] != wantN[f]) return false;
            }
         }
   return true;
}

inline void set_brick_chunk(Brick &g, int chunk)
{
   g.chunkZ = std::max(1, std::min(chunk, g.en));
   g.chunks = (g.en + g.chunkZ - 1) / g.chunkZ;
}

// Planes per workgroup of a column-walking kernel: a launch of `columns`
// column tiles, each cut into z-chunks of `chunk` planes out of `planes`,
// runs in rounds of `slots` resident workgroups, and a workgroup walks one
// plane more than its chunk.  Picks the chunk that minimises rounds x
// (chunk + 1): redundant planes against a partly filled last round.
inline int pick_chunk(int64_t columns, int planes, int slots, int fallback)
{
   if (slots < 1) return fallback;
   int best = fallback;
   double bestCost = 1e300;
   for (int ch = 4; ch <= planes; ++ch) {
      const int64_t blocks = columns * ((planes + ch - 1) / ch);
      const int64_t rounds = (blocks + slots - 1) / slots;
      const double cost = (double)rounds * (ch + 1);
      if (cost < bestCost) {
         bestCost = cost;
         best = ch;
      }
   }
   return best;
}

// Default node planes per workgroup of the fused nodal kernel.  s=192 on
// 256 CUs: 65 for FP64 (2 workgroups per CU: all 507 resident at once, 2%
// faster than 22), 22 for the mixed form (3 per CU).
inline int pick_brick_chunk(const Brick &g, int slots)
{
   return pick_chunk((int64_t)g.tilesX * g.tilesY, g.en, slots, 8);
}

inline void set_kin_chunk(Brick &g, int chunk)
{
   g.kinChunk = std::max(1, std::min(chunk, g.s));
   g.kinChunks = (g.s + g.kinChunk - 1) / g.kinChunk;
}

// Default element planes per workgroup of the kinematics kernel.  s=192 on
// 256 CUs: 12 for FP64 (3 workgroups per CU), 28 for the mixed form (4 per
// CU: 1,008 workgroups, all resident at once; 3% faster than 12).
inline int pick_kin_chunk(const Brick &g, int slots)
{
   return pick_chunk((int64_t)g.kinTiles * g.kinTiles, g.s, slots, 12);
}

// Fused nodal kernel tiles of w threads by h rows: (w-1) x (h-1) nodes each.
inline void set_brick_tile(Brick &g, int w, int h)
{
   g.tileW = w;
   g.rows = h;
   g.tilesX = (g.en + w - 2) / (w - 1);
   g.tilesY = (g.en + h - 2) / (h - 1);
}

// The tile of an ensemble member's nodal workgroups (256 threads): the
// width w (rows 256 / w) that needs the fewest tiles per plane, so that few
// lanes fall outside a small mesh (s=16: 18 x 14 threads, 2 tiles instead of
// 4 of 16 x 16); ties go to the width nearest 16.
inline void pick_member_tile(Brick &g)
{
   int bestW = 16;
   int64_t best = INT64_MAX;
   for (int w = 4; w <= 64; ++w) {
      const int h = 256 / w;
      if (h < 3) break;
      const int64_t tiles = (int64_t)((g.en + w - 2) / (w - 1)) * ((g.en + h - 2) / (h - 1));
      if (tiles < best || (tiles == best && std::abs(w - 16) < std::abs(bestW - 16))) {
         best = tiles;
         bestW = w;
      }
   }
   set_brick_tile(g, bestW, 256 / bestW);
}

// Kinematics tiles of w x h elements (one per thread of a 256-thread
// workgroup), kept within the LDS of the 16 x 16 tile: (h+1) node rows of
// 2 ceil((w+1)/2) values, at most 17 x 18.
inline bool kin_tile_fits(int w, int h)
{
   return w >= 1 && h >= 1 && w * h <= 256 && (h + 1) * 2 * ((w + 2) / 2) <= 17 * 18;
}

inline void set_kin_tile(Brick &g, int w, int h)
{
   g.kinW = w;
   g.kinH = h;
   g.kinTilesX = (g.s + w - 1) / w;
   g.kinTilesY = (g.s + h - 1) / h;
}

// An ensemble member's kinematics tile: the width (rows 256 / width) with
// the fewest tiles per plane that fits the LDS; ties go to the width
// nearest 16.
inline void pick_member_kin_tile(Brick &g)
{
   int bestW = 16;
   int64_t best = INT64_MAX;
   for (int w = 4; w <= 64; ++w) {
      const int h = 256 / w;
      if (!kin_tile_fits(w, h)) continue;
      const int64_t tiles = (int64_t)((g.s + w - 1) / w) * ((g.s + h - 1) / h);
      if (tiles < best || (tiles == best && std::abs(w - 16) < std::abs(bestW - 16))) {
         best = tiles;
         bestW = w;
      }
   }
   set_kin_tile(g, bestW, 256 / bestW);
}

// Tiling of an s^3 brick with the default chunks (the caller may pick
// others): 15x15-node tiles of the fused nodal kernel, 16x16-element tiles
// of the kinematics kernel.
inline void init_brick(Brick &g, int s)
{
   g.s = s;
   g.en = s + 1;
   set_brick_tile(g, 16, 16);
   g.divS = make_fastdiv((uint32_t)g.s);
   g.divS2 = make_fastdiv((uint32_t)(g.s * g.s));
   g.kinTiles = (g.s + 15) / 16;
   set_kin_tile(g, 16, 16);
   set_brick_chunk(g, 8);
   set_kin_chunk(g, 12);
}

// Workgroups of one launch of each brick kernel.
inline int64_t nodal_blocks(const Brick &g) { return (int64_t)g.tilesX * g.tilesY * g.chunks; }
inline int64_t kin_blocks(const Brick &g) { return (int64_t)g.kinTilesX * g.kinTilesY * g.kinChunks; }

} // namespace lgpu
