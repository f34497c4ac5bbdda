// Halo (temporal) blocking decomposition of one axis.
//
// Restates the reference's blocking core (src/blocking_core/blocking_alogrithm.hpp:8-41)
// and the tile placement of its simulator (simulator_expt_platform.cu:120-198, 523-540,
// 792-795).  An axis of length N is cut into tiles of length B that each advance I
// steps; a tile's outer I-1 cells on each interior side are halo, recomputed by the
// neighbouring tile, so only the "valid" range [vb(k), vb(k+1)) of tile k is written.
//   n tiles:  1 if N <= B; 2 if N <= 2(B+1-I); else 2 + ceil((N - 2(B+1-I)) / (B+2-2I))
//   vb(k):    0 for k = 0, N for k = n, else (B+1-I) + (k-1)(B+2-2I)
//   loaded:   [max(vb(k) - (I-1), 0), min(vb(k+1) + (I-1), N))  (the tile origin is the start)
#pragma once
#include <algorithm>
#include <cstdint>

namespace culbm {

// validBlkAxisConfig: the reference's own acceptance test, kept for its error
// behaviour on the command line.
inline bool blk_axis_valid_ref(uint32_t dom, uint32_t blk, uint32_t iter) {
  if (blk < dom && blk + 1 <= iter) return false;
  if (2 * (blk + 1 - iter) < dom && blk + 2 <= iter) return false;
  return true;
}

// Stricter test under which the decomposition is well defined: every tile
// advances at least one valid cell (B+1-I > 0 when tiled) and interior tiles
// move by a positive stride B+2-2I when more than two tiles are needed.
inline bool blk_axis_valid(int64_t dom, int64_t blk, int64_t iter) {
  if (dom < 1 || blk < 1 || iter < 1) return false;
  if (dom <= blk) return true;
  if (blk + 1 - iter <= 0) return false;
  if (dom <= 2 * (blk + 1 - iter)) return true;
  return blk + 2 - 2 * iter > 0;
}

inline int64_t blk_num(int64_t dom, int64_t blk, int64_t iter) {
  if (dom <= blk) return 1;
  if (dom <= 2 * (blk + 1 - iter)) return 2;
  const int64_t den = blk + 2 - 2 * iter;
  return 2 + (dom - 2 * (blk + 1 - iter) + den - 1) / den;
}

inline int64_t blk_valid_start(int64_t k, int64_t blk, int64_t num, int64_t iter, int64_t dom) {
  if (k == num) return dom;
  if (k == 0) return 0;
  return (blk + 1 - iter) + (k - 1) * (blk + 2 - 2 * iter);
}

inline int64_t blk_load_begin(int64_t k, int64_t blk, int64_t num, int64_t iter, int64_t dom) {
  return std::max<int64_t>(blk_valid_start(k, blk, num, iter, dom) - (iter - 1), 0);
}

inline int64_t blk_load_end(int64_t k, int64_t blk, int64_t num, int64_t iter, int64_t dom) {
  return std::min<int64_t>(blk_valid_start(k + 1, blk, num, iter, dom) + (iter - 1), dom);
}

}  // namespace culbm
