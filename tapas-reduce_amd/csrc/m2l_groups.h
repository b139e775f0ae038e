// m2l_groups.h -- the direct-mapped table of M2L offset groups, shared by
// the traversal (which gives every M2L pair its slot as it emits the pair)
// and the batched M2L (which turns the used slots into groups).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tfmm {
namespace m2lg {

// Slot of a (target level, level difference, offset) in the direct-mapped
// table: offsets are in units of half the finer width; same-level offsets are
// even with |o| <= 22, offsets across levels odd with |o| <= 23.  An M2L
// pair's parents were not well separated, so a same-level or adjacent-level
// offset is at most about 2 (2 / theta + 1) half widths (the table covers
// theta >= 0.2 there); pairs 2 or 3 levels apart, which a leaf next to a much
// finer region makes, reach |o| = 15 and 23 at theta = 0.4.
constexpr int SAME_HALF = 11;                // |o/2| <= 11
constexpr int SAME_N = 2 * SAME_HALF + 1;    // 23
constexpr int CROSS_MAX = 23;                // |o| <= 23, odd
constexpr int CROSS_N = CROSS_MAX + 1;       // (o + 23) / 2 in [0, 23]
constexpr int DL_MAX = 3;                    // level differences -3 .. 3
constexpr int SLOTS_SAME = SAME_N * SAME_N * SAME_N;
constexpr int SLOTS_CROSS = CROSS_N * CROSS_N * CROSS_N;
constexpr int SLOTS_PER_LEVEL = SLOTS_SAME + 2 * DL_MAX * SLOTS_CROSS;
__device__ inline int32_t group_slot(int lt, int ls, const int o[3]) {
  const int dl = ls - lt;
  int idx;
  if (dl == 0) {
    int a[3];
    for (int d = 0; d < 3; d++) {
      if (o[d] & 1) return -1;
      a[d] = o[d] / 2 + SAME_HALF;
      if (a[d] < 0 || a[d] >= SAME_N) return -1;
    }
    idx = (a[0] * SAME_N + a[1]) * SAME_N + a[2];
  } else if (dl >= -DL_MAX && dl <= DL_MAX) {
    int a[3];
    for (int d = 0; d < 3; d++) {
      if (!(o[d] & 1)) return -1;
      a[d] = (o[d] + CROSS_MAX) / 2;
      if (a[d] < 0 || a[d] >= CROSS_N) return -1;
    }
    const int blk = 2 * ((dl > 0 ? dl : -dl) - 1) + (dl < 0 ? 1 : 0);
    idx = SLOTS_SAME + blk * SLOTS_CROSS + (a[0] * CROSS_N + a[1]) * CROSS_N + a[2];
  } else {
    return -1;
  }
  return lt * SLOTS_PER_LEVEL + idx;
}

// slot of an M2L pair (target t, source s) from the cells' depths, widths
// (max over the dimensions) and centres: the offset in half widths of the
// finer cell; -1 outside the table
__device__ inline int32_t pair_slot(int lt, int ls, double wt, double ws, const double *ct,
                                    const double *cs) {
  const double half = 0.5 * fmin(wt, ws);
  int o[3];
  for (int d = 0; d < 3; d++) {
    const double v = (ct[d] - cs[d]) / half;
    if (!(fabs(v) < 64.0)) return -1;
    o[d] = (int)rint(v);
  }
  return group_slot(lt, ls, o);
}

}  // namespace m2lg
}  // namespace tfmm
