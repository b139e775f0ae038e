// m2l_batched.hip -- M2L regrouped by relative cell offset and run as
// block-sparse batched contractions on the gfx950 matrix cores (MFMA).
//
// Math (reference: LaplaceSphericalCPU_tapas.h:208-273, evalLocal :102-135):
// the M2L of one pair is a linear map of the source's 55 complex multipole
// coefficients onto the target's 55 local coefficients, whose entries are the
// singular harmonics Y_{j+n}^{m-k}(dX) with signs.  Written on real and
// imaginary parts it is a 112 x 112 real matrix T(dX) (110 + 2 padding rows
// and columns), non-zero only where j + n <= P - 1; in 16 x 16 blocks only 18
// of the 49 blocks are non-zero.
//
// Cells of one level sit on a lattice, so dX of a pair is an integer
// multiple of half the finer cell's width: pairs are grouped by
// (target level, source level, offset).  Every pair of a group shares T, so a
// group's pairs become the columns of a matrix product T * [M_s1 ... M_s16].
// The harmonics of each group are formed once, in FP64.
//
// Scaling: with sigma the width of each cell, the kernel works on
//   M~_n = M_n / sigma_s^n,  L~_j = L_j sigma_t^(j+1),  Y~_p = Y_p sigma_t^(p+1)
// so that L~_j = sum_n 2^(n (l_t - l_s)) Y~_(j+n) M~_n with all quantities
// O(1): FP32 neither under- nor overflows at P = 10 on deep trees, and the
// power-of-two factors are exact.  Y~ depends only on the offset in units of
// the cells' widths, so one group serves every tree of a batch.
//
// Work decomposition: targets are tiled by 128 consecutive cells (breadth
// first, so one tile is one Morton range of one level); the pairs of a tile
// are sorted by group, and each run of up to 16 pairs of one group is a work
// item = 18 blocks x 4 MFMA 16x16x4 steps.  One workgroup (4 waves) owns a
// tile and accumulates its local expansions in LDS; each wave owns fixed row
// blocks of T, so the waves never write the same LDS word, walk the tile's
// items independently, and the accumulation order is fixed (deterministic).
// The item loop is software pipelined (the next items' operands are in
// flight while an item is multiplied), and tiles are dealt to the XCDs in
// runs of consecutive tiles so that one XCD's L2 holds the sources its tiles
// share.  Pairs whose offset falls outside the group table (level
// difference above 3, or an offset beyond the table's range) go to the
// per-pair kernel of m2l.hip.
//
// Per-group factors: the level factor 2^(n dl), n the source degree and
// dl = l_t - l_s, is split as 2^(p dl) 2^(-j dl) with p = j + n, the first
// folded into the group's harmonics (index p) and the second into the rows
// (degree j) of the group's A operands, which are formed once per group
// before the tile kernel; both are exact powers of two.
#include <algorithm>
#include <cmath>
#include <vector>

#include "m2l_codes.h"
#include "m2l_groups.h"
#include "tfmm_device.h"
#include "tfmm_kernels.h"
#include "tfmm_internal.h"

namespace tfmm {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

using m2lc::BLK_CB;
using m2lc::BLK_RB;
using m2lc::NBLK;
using m2lc::NR;
constexpr int NCODE = NBLK * 4 * 64;
constexpr int SRC_BITS = 25;  // source cell bits of a grouped pair's value
constexpr int TILE = 128;
constexpr int TILE_WAVES = 4;
constexpr int WAVE_SYNC = 64;  // items between block barriers in the tile loop

// ---------------------------------------------------------------------------
// group table
// ---------------------------------------------------------------------------
// group table (m2l_groups.h)
using m2lg::CROSS_MAX;
using m2lg::CROSS_N;
using m2lg::DL_MAX;
using m2lg::SAME_HALF;
using m2lg::SAME_N;
using m2lg::SLOTS_CROSS;
using m2lg::SLOTS_PER_LEVEL;
using m2lg::SLOTS_SAME;
// column word of the tile kernel: target slot (7 bits) | 128 | (l_t - l_s +
// DL_MAX) << COL_DL | group << COL_G
constexpr int COL_DL = 8;
constexpr int COL_G = 11;

__device__ inline void slot_decode(int32_t slot, int &lt, int &ls, int o[3]) {
  lt = slot / SLOTS_PER_LEVEL;
  int idx = slot - lt * SLOTS_PER_LEVEL;
  if (idx < SLOTS_SAME) {
    ls = lt;
    o[2] = (idx % SAME_N - SAME_HALF) * 2;
    idx /= SAME_N;
    o[1] = (idx % SAME_N - SAME_HALF) * 2;
    o[0] = (idx / SAME_N - SAME_HALF) * 2;
  } else {
    idx -= SLOTS_SAME;
    const int blk = idx / SLOTS_CROSS;
    ls = lt + ((blk & 1) ? -(blk / 2 + 1) : (blk / 2 + 1));
    idx %= SLOTS_CROSS;
    o[2] = (idx % CROSS_N) * 2 - CROSS_MAX;
    idx /= CROSS_N;
    o[1] = (idx % CROSS_N) * 2 - CROSS_MAX;
    o[0] = (idx / CROSS_N) * 2 - CROSS_MAX;
  }
}


// Split the executed M2L pairs into grouped ones (sort key: tile << gbits |
// group, value: source cell | slot in tile << SRC_BITS) and the rest (key:
// target | source).  Within one (tile, group) segment every target occurs at
// most once, so the order inside a segment does not affect any sum and the
// key needs no slot bits.
// Grouped keys stay at the pair's index; a pair outside the groups leaves
// the key NO_GROUP there, which sorts after every grouped key (the tile
// field of a key is at most 2^tbits - 2), so a sort of all n keys puts the
// grouped pairs first.  The pairs outside the groups are appended compactly
// (counters[1], one atomic per block that has any); k_grouped_count then
// sets counters[0] to the grouped count.
// With dn the pair count is *dn (n its capacity), and group indices are
// clamped below gcap (a capacity the caller checks after the launch).
constexpr uint32_t NO_GROUP = 0xffffffffu;

__global__ __launch_bounds__(1024) void k_classify(const int2 *pairs, const int32_t *slot_of,
                                                  uint32_t n, const uint32_t *gid_of_slot, int gbits,
                                                  int sbits, uint32_t *gkeys, uint32_t *gvals,
                                                  uint64_t *fkeys, uint32_t *fvals,
                                                  uint32_t *counters, const uint32_t *dn,
                                                  uint32_t gcap) {
  critical_path_priority();
  __shared__ uint32_t scratch[17];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dn) n = min(n, *dn);
  const bool valid = i < n;
  int2 p = make_int2(0, 0);
  int32_t slot = -1;
  if (valid) {
    p = pairs[i];
    slot = slot_of[i];
  }
  const bool grouped = valid && slot >= 0;
  const bool fallback = valid && slot < 0;
  if (valid) {
    const uint64_t tile = (uint32_t)p.x / TILE, in_tile = (uint32_t)p.x % TILE;
    gkeys[i] = grouped ? (uint32_t)(tile << gbits) | min(gid_of_slot[slot], gcap - 1u) : NO_GROUP;
    gvals[i] = grouped ? ((uint32_t)p.y & CELL_MASK) | (uint32_t)(in_tile << SRC_BITS) : 0u;
  }
  if (!__syncthreads_or(fallback)) return;  // the usual block: nothing outside the groups
  const uint32_t af = block_append<16>(&counters[1], fallback, 1, scratch);
  if (fallback) {
    fkeys[af] = ((uint64_t)(uint32_t)p.x << sbits) | (uint64_t)(uint32_t)p.y;
    fvals[af] = 0;
  }
}

// counters[0] = the grouped pairs: the pair count less those outside
__global__ void k_grouped_count(uint32_t n, const uint32_t *dn, uint32_t *counters) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (dn) n = min(n, *dn);
  counters[0] = n - min(n, counters[1]);
}

// k_classify when every pair is in the table: keys and values in pair order
__global__ void k_classify_all(const int2 *pairs, const int32_t *slot_of, uint32_t n,
                               const uint32_t *gid_of_slot, int gbits, uint32_t *gkeys,
                               uint32_t *gvals) {
  critical_path_priority();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 p = pairs[i];
  const uint64_t tile = (uint32_t)p.x / TILE, in_tile = (uint32_t)p.x % TILE;
  gkeys[i] = (uint32_t)(tile << gbits) | gid_of_slot[slot_of[i]];
  gvals[i] = ((uint32_t)p.y & CELL_MASK) | (uint32_t)(in_tile << SRC_BITS);
}

// Harmonics of every used group, FP64, scaled: Y~_p^q = Y_p^q sigma_t^(p+1)
// 2^(p dl), dl = l_t - l_s, i.e. the harmonics of the offset measured in
// target widths.  Stored as NR reals (re/im interleaved, index 2 tri(p, q)),
// in FP64 and FP32.
__global__ void k_group_harmonics(const uint32_t *used, const uint32_t *gid_of_slot,
                                  int32_t nslots, double *Y64, float *Y32, int8_t *gdl,
                                  int8_t *glt, uint32_t gcap) {
  critical_path_priority();
  const int32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots || !used[slot]) return;
  const uint32_t g = gid_of_slot[slot];
  if (g >= gcap) return;  // more groups than the buffers hold (checked by the caller)
  int lt, ls, o[3];
  slot_decode(slot, lt, ls, o);
  const int lf = lt > ls ? lt : ls;
  const double half = ldexp(0.5, lt - lf);  // half the finer width, target width = 1
  const double dx = o[0] * half, dy = o[1] * half, dz = o[2] * half;
  double rho, alpha, beta;
  cart2sph(dx, dy, dz, rho, alpha, beta);
  const double sig = 1.0;
  const double x = cos(alpha), y = sin(alpha);
  const cd ei = mk(cos(beta), sin(beta));
  double *yo = Y64 + (int64_t)g * NR;
  float *yf = Y32 + (int64_t)g * NR;
  for (int m = 0; m < P; m++) {
    local_column(m, rho, x, y, ei, [&](int n, int mm, cd v) {
      double sc = ldexp(sig, n * (lt - ls));  // sigma^(n+1) 2^(n dl)
      for (int i = 0; i < n; i++) sc *= sig;
      const int t = tri(n, mm);
      yo[2 * t] = v.re * sc;
      yo[2 * t + 1] = v.im * sc;
      yf[2 * t] = (float)(v.re * sc);
      yf[2 * t + 1] = (float)(v.im * sc);
    });
  }
  yo[110] = yo[111] = 0;
  yf[110] = yf[111] = 0;
  gdl[g] = (int8_t)(lt - ls);
  glt[g] = (int8_t)lt;
}

// M~ of every cell, in the quad-interleaved layout the MFMA B operand reads:
// Mq[cell][q][t] = M~[cell][4 t + q], row r = 2 tri(n, m) + (0 re | 1 im).
template <class T>
__global__ void k_scale_m(Cells c, int32_t ncells, const double *M, T *Mq) {
  critical_path_priority();
  const int32_t ci = blockIdx.x;
  if (ci >= ncells) return;
  const int r = threadIdx.x;  // 0..111
  if (r >= NR) return;
  const int q = r & 3, t = r >> 2;
  double v = 0;
  if (r < 110) {
    const int tt = r >> 1;
    int n = 0;
    while (tri(n + 1, 0) <= tt) n++;
    const double w = c.width[ci];
    const double isig = w > 0 ? 1.0 / w : 0.0;  // zero width: only degree 0 is non-zero
    double sc = 1;
    for (int i = 0; i < n; i++) sc *= isig;
    v = M[2 * ((int64_t)ci * NTERM + tt) + (r & 1)] * sc;
  }
  Mq[(int64_t)ci * NR + q * 28 + t] = (T)v;
}

// ---------------------------------------------------------------------------
// work items
// ---------------------------------------------------------------------------
// Position of every grouped pair in its (tile, group) segment, from the
// sorted keys alone.  A target occurs at most once per segment, so a segment
// holds at most TILE = 128 pairs: the block reads the 128 keys before its
// own and 64 after, marks the segment starts (ballot per row of 64 keys),
// and every pair finds its segment's start among the marks of its row and
// the two rows before, and the next start among those of its row and the
// row after.  Items are runs of 16 pairs of one segment: a pair starts one
// when its position in the segment is a multiple of 16; ipos = position in
// the segment | (pairs of the item) << 8 for an item start, and bcount[b]
// the item starts among block b's pairs.
constexpr int IP_WAVES = 16;
constexpr int IP_ROWS = IP_WAVES + 3;  // 2 rows before the block's, 1 after

__global__ __launch_bounds__(64 * IP_WAVES) void k_item_pos(const uint32_t *keys, uint32_t n,
                                                            uint32_t *ipos, uint32_t *bcount,
                                                            const uint32_t *dn) {
  critical_path_priority();
  __shared__ uint32_t sk[IP_ROWS * 64];
  __shared__ uint64_t mark[IP_ROWS];
  __shared__ uint32_t wcount[IP_WAVES];
  if (dn) n = min(n, *dn);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t w0 = (int64_t)blockIdx.x * (64 * IP_WAVES) - 128;  // key of window slot 0
  for (int t = threadIdx.x; t < IP_ROWS * 64; t += 64 * IP_WAVES) {
    const int64_t j = w0 + t;
    // outside [0, n): a value no key takes (the tile field of a key is at
    // most 2^tbits - 2), so the last segment ends at n
    sk[t] = (j >= 0 && j < (int64_t)n) ? keys[j] : 0xffffffffu;
  }
  __syncthreads();
  for (int r = wid; r < IP_ROWS; r += IP_WAVES) {
    const int t = 64 * r + lane;
    const uint64_t m = __ballot(t == 0 || sk[t] != sk[t - 1]);
    if (lane == 0) mark[r] = m;
  }
  __syncthreads();
  const int R = wid + 2;
  const int64_t i = w0 + 64 * R + lane;
  bool first = false;
  if (i < (int64_t)n) {
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // bits <= lane
    int start;  // window slot of the segment's first pair
    if (mark[R] & upto) start = 64 * R + 63 - __clzll(mark[R] & upto);
    else if (mark[R - 1]) start = 64 * (R - 1) + 63 - __clzll(mark[R - 1]);
    else start = 64 * (R - 2) + 63 - __clzll(mark[R - 2]);  // mark[R - 2] has slot 0 at least
    const int d = min(64 * R + lane - start, TILE - 1);
    int next;  // window slot of the next segment's first pair
    if (mark[R] & ~upto) next = 64 * R + __ffsll((unsigned long long)(mark[R] & ~upto)) - 1;
    else if (mark[R + 1]) next = 64 * (R + 1) + __ffsll((unsigned long long)mark[R + 1]) - 1;
    else next = 64 * (R + 2);
    const int cnt = min(16, next - (64 * R + lane));
    first = (d & 15) == 0;
    ipos[i] = (uint32_t)d | (first ? (uint32_t)cnt << 8 : 0u);
  }
  const uint64_t fm = __ballot(first);
  if (lane == 0) wcount[wid] = (uint32_t)__popcll(fm);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < IP_WAVES; w++) t += wcount[w];
    bcount[blockIdx.x] = t;
  }
}

// Item descriptors (first pair, pairs, group, tile) and the item's 16 column
// descriptors: cols[16 item + j] = (source cell, target slot in tile
// | 128 if column j holds a pair | (level difference + DL_MAX) << COL_DL
// | group << COL_G).
// Every column carries the group, so that every lane of the tile kernel can
// fetch the group's data.
// Blocks as in k_item_pos; bstart: exclusive prefix of its block counts.
// With dn the pair count is *dn (n its capacity); items at or past
// items_cap are not written (a capacity the caller checks).
__global__ __launch_bounds__(64 * IP_WAVES) void k_items(const uint32_t *keys,
                                                         const uint32_t *vals,
                                                         const uint32_t *ipos,
                                                         const uint32_t *bstart, uint32_t n,
                                                         int gbits, const int8_t *gdl,
                                                         int4 *items, uint2 *cols,
                                                         const uint32_t *dn, uint32_t items_cap) {
  critical_path_priority();
  __shared__ uint32_t wcount[IP_WAVES];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dn) n = min(n, *dn);
  const bool valid = i < n;
  const uint32_t ip = valid ? ipos[i] : 1u;
  const bool first = valid && (ip & 15u) == 0;
  // item of pair i = item starts at or before i, less one
  const uint64_t fm = __ballot(first);
  if (lane == 0) wcount[wid] = (uint32_t)__popcll(fm);
  __syncthreads();
  if (!valid) return;
  uint32_t before = bstart[blockIdx.x];
  for (int w = 0; w < wid; w++) before += wcount[w];
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const uint32_t item = before + (uint32_t)__popcll(fm & upto) - 1;
  const uint32_t jj = ip & 15u;
  const uint32_t k = keys[i], v = vals[i];
  const uint32_t g = k & ((1u << gbits) - 1u);
  const uint32_t gw = (uint32_t)(gdl[g] + DL_MAX) << COL_DL | (g << COL_G);
  if (item >= items_cap) return;
  cols[(int64_t)item * 16 + jj] =
      make_uint2(v & ((1u << SRC_BITS) - 1u), (v >> SRC_BITS) | 128u | gw);
  if (first) {  // first pair of the item: descriptor and padding columns
    const uint32_t cnt = ip >> 8;  // pairs of the segment from i on, at most 16
    int4 it;
    it.x = (int)i;                             // first pair
    it.y = (int)cnt;                           // pairs in the item
    it.z = (int)g;                             // group
    it.w = (int)(k >> gbits);                  // tile
    items[item] = it;
    for (uint32_t j2 = cnt; j2 < 16; j2++) cols[(int64_t)item * 16 + j2] = make_uint2(0u, gw);
  }
}

__global__ void k_tile_offsets(const int4 *items, uint32_t nitems, int32_t ntiles,
                               int32_t *off, const uint32_t *dnitems) {
  critical_path_priority();
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  if (dnitems) nitems = min(nitems, *dnitems);
  uint32_t lo = 0, hi = nitems;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (items[mid].w < t) lo = mid + 1;
    else hi = mid;
  }
  off[t] = (int32_t)lo;
}

// ---------------------------------------------------------------------------
// the tile kernel
// ---------------------------------------------------------------------------
template <class T>
struct Mfma;
// Operands (both types): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16].  Result i of lane l is D[row(l / 16, i)][l % 16], with
// the row order differing between the two instructions (checked on gfx950).
template <>
struct Mfma<float> {
  typedef f4 V;
  static __device__ inline V step(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int q, int i) { return 4 * q + i; }
};
template <>
struct Mfma<double> {
  typedef d4 V;
  static __device__ inline V step(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int q, int i) { return q + 4 * i; }
};

// Work split of one item over the 4 waves of a tile: each wave owns fixed
// (row block, column block) pairs of T, so the waves write disjoint rows of
// the tile's expansions and need no atomics and no barriers between items.
// Row block 0 has 7 column blocks and is split between waves 0 and 1 (wave 1
// accumulates its part in a separate array, added at the end).
//   wave 0: (0,0) (0,1) (0,2) (0,3) (6,0)
//   wave 1: (0,4) (0,5) (0,6) (4,0) (5,0)
//   wave 2: (1,0) (1,1) (1,2) (1,3)
//   wave 3: (2,0) (2,1) (3,0) (3,1)
// Blocks of one row block accumulate into one MFMA accumulator (WAVE_F).
constexpr int WAVE_NBLK[TILE_WAVES] = {5, 5, 4, 4};
constexpr int WAVE_BLK[TILE_WAVES][5] = {
    {0, 1, 2, 3, 17}, {4, 5, 6, 15, 16}, {7, 8, 9, 10, -1}, {11, 12, 13, 14, -1}};
constexpr int WAVE_NRB[TILE_WAVES] = {2, 3, 1, 2};
constexpr int WAVE_RBS[TILE_WAVES][3] = {{0, 6, -1}, {0, 4, 5}, {1, -1, -1}, {2, 3, -1}};
constexpr int WAVE_F[TILE_WAVES][5] = {
    {0, 0, 0, 0, 1}, {0, 0, 0, 1, 2}, {0, 0, 0, 0, -1}, {0, 0, 1, 1, -1}};

__host__ __device__ constexpr bool wave_needs_cb(int w, int cb) {
  for (int e = 0; e < WAVE_NBLK[w]; e++)
    if (BLK_CB[WAVE_BLK[w][e]] == cb) return true;
  return false;
}

__device__ inline int row_degree(int r) {
  if (r >= 110) return 0;
  const int t = r >> 1;
  int n = 0;
  while (tri(n + 1, 0) <= t) n++;
  return n;
}

// B fragments (4 k-steps) of the column blocks wave W multiplies, for the
// source whose M~ starts at mq (already offset by the lane's k-row quad)
template <class T, int W>
__device__ inline void load_b(T (&b)[7][4], const T *mq) {
#pragma unroll
  for (int cb = 0; cb < 7; cb++) {
    if (wave_needs_cb(W, cb)) {
#pragma unroll
      for (int kk = 0; kk < 4; kk++) b[cb][kk] = mq[4 * cb + kk];
    }
  }
}

// A operands of every group, in the order the tile kernel's lanes read them:
// Af[g][AF_SLOT(w, e)][lane][kk] is the entry of block WAVE_BLK[w][e] at
// k-step kk that lane supplies (a 16-byte vector per lane and block in FP32,
// 32 in FP64), so a wave reads its operands of a new group with one
// contiguous, conflict-free load per block.  Slots number the (wave, block)
// pairs wave by wave: slot = sum of WAVE_NBLK[0 .. w-1] + e.
constexpr int AF_PER_GROUP = NBLK * 64 * 4;  // elements per group
__host__ __device__ constexpr int af_slot(int w, int e) {
  int s = e;
  for (int i = 0; i < w; i++) s += WAVE_NBLK[i];
  return s;
}

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef f4 V;
};
template <>
struct Vec4<double> {
  typedef d4 V;
};

// A operands of wave W for the group whose fragments start at af
template <class T, int W>
__device__ inline void load_a(T (&a)[WAVE_NBLK[W] * 4], const T *af, int lane) {
  typedef typename Vec4<T>::V V;
#pragma unroll
  for (int e = 0; e < WAVE_NBLK[W]; e++) {
    const V v = *reinterpret_cast<const V *>(af + (af_slot(W, e) * 64 + lane) * 4);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) a[e * 4 + kk] = v[kk];
  }
}

// Pipeline state of one wave at item it: everything was loaded during item
// it - 1, so nothing the item consumes is still in flight when it starts;
// acc holds the products of item it - 1, stored during item it.
template <class T, int W>
struct WaveState {
  uint2 c_cur;  // columns of it (sources of its B operands were c_b of it - 1)
  uint2 c_b;    // columns of it + 1 (sources and group of the next operands)
  uint2 c_st;   // columns of it - 1 (target slots and level difference of acc)
  T b[7][4];    // B operands of it
  typename Mfma<T>::V acc[WAVE_NRB[W]];  // products of it - 1
};

// add the products of one item (columns c) into the tile's expansions
template <class T, int W>
__device__ inline void store_item(const typename Mfma<T>::V (&acc)[WAVE_NRB[W]], uint2 c,
                                  T (*Lt)[NR + 1], T (*Lt0b)[17], int q) {
  if (!(c.y & 128u)) return;
  const uint32_t ts = c.y & 127u;
#pragma unroll
  for (int f = 0; f < WAVE_NRB[W]; f++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = Mfma<T>::row(q, i);
      const T v = acc[f][i];
      if (W == 1 && WAVE_RBS[W][f] == 0) {
        Lt0b[ts][r] += v;
      } else {
        const int row = WAVE_RBS[W][f] * 16 + r;
        if (row < 110) Lt[ts][row] += v;
      }
    }
  }
}

// One item of wave W with the state of it in cur; fills nxt for it + 1.
// Order: the item's MFMAs are issued first; the loads of the next item's
// operands and the store of the previous item's products are independent of
// them and overlap their execution.  Loads issued here are consumed by the
// next call.  The state is double buffered by unrolling the caller's loop by
// two (a register rotation would wait for the loads just issued).
template <class T, int W>
__device__ inline void wave_item(WaveState<T, W> &cur, WaveState<T, W> &nxt,
                                 T (&a)[WAVE_NBLK[W] * 4], int32_t it, int32_t ib, int32_t last,
                                 const uint2 *cols, const T *Af, const T *Mq, T (*Lt)[NR + 1],
                                 T (*Lt0b)[17], int lane) {
  constexpr int NB = WAVE_NBLK[W], NF = WAVE_NRB[W];
  const int q = lane >> 4, j = lane & 15;
  // the matrix work of it
#pragma unroll
  for (int f = 0; f < NF; f++) nxt.acc[f] = (typename Mfma<T>::V){0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
#pragma unroll
    for (int e = 0; e < NB; e++)
      nxt.acc[WAVE_F[W][e]] = Mfma<T>::step(a[e * 4 + kk], cur.b[BLK_CB[WAVE_BLK[W][e]]][kk],
                                            nxt.acc[WAVE_F[W][e]]);
  }
  nxt.c_st = cur.c_cur;
  // loads for the next items (indices clamped: loads stay in range)
  nxt.c_cur = cur.c_b;
  nxt.c_b = cols[(int64_t)min(it + 2, last) * 16 + j];
  load_b<T, W>(nxt.b, Mq + (int64_t)cur.c_b.x * NR + q * 28);
  // A operands of it + 1, loaded over those of it (the MFMAs above have
  // read them) when the group changes: consecutive items of one group share
  // them, and every column of an item carries its group, so the branch is
  // uniform
  const uint32_t gn = cur.c_b.y >> COL_G;
  if (gn != (cur.c_cur.y >> COL_G)) load_a<T, W>(a, Af + (int64_t)gn * AF_PER_GROUP, lane);
  // the products of it - 1
  store_item<T, W>(cur.acc, cur.c_st, Lt, Lt0b, q);
  // keep the 4 waves within a few items of each other: they read the same
  // source multipoles, which then come from L2 once instead of 4 times
  if (((it - ib) & (WAVE_SYNC - 1)) == WAVE_SYNC - 1) __syncthreads();
}

// The item loop of wave W over the tile's items [ib, ie).
template <class T, int W>
__device__ inline void run_wave(const uint2 *cols, int32_t ib, int32_t ie, const T *Af,
                                const T *Mq, T (*Lt)[NR + 1], T (*Lt0b)[17], int lane) {
  constexpr int NF = WAVE_NRB[W];
  if (ib >= ie) return;
  const int q = lane >> 4, j = lane & 15;
  const int32_t last = ie - 1;

  WaveState<T, W> s0, s1;
  s0.c_cur = cols[(int64_t)ib * 16 + j];
  s0.c_b = cols[(int64_t)min(ib + 1, last) * 16 + j];
  s0.c_st = make_uint2(0u, 0u);  // no previous item
#pragma unroll
  for (int f = 0; f < NF; f++) s0.acc[f] = (typename Mfma<T>::V){0, 0, 0, 0};
  load_b<T, W>(s0.b, Mq + (int64_t)s0.c_cur.x * NR + q * 28);
  T a[WAVE_NBLK[W] * 4];  // A operands of the current item
  load_a<T, W>(a, Af + (int64_t)(s0.c_cur.y >> COL_G) * AF_PER_GROUP, lane);
  for (int32_t it = ib; it < ie; it += 2) {
    wave_item<T, W>(s0, s1, a, it, ib, last, cols, Af, Mq, Lt, Lt0b, lane);
    if (it + 1 < ie) wave_item<T, W>(s1, s0, a, it + 1, ib, last, cols, Af, Mq, Lt, Lt0b, lane);
  }
  // the products of the last item
  if ((ie - ib) & 1) store_item<T, W>(s1.acc, s1.c_st, Lt, Lt0b, q);
  else store_item<T, W>(s0.acc, s0.c_st, Lt, Lt0b, q);
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b to XCD
// b % 8) and every XCD has its own L2.  Consecutive tiles are neighbouring
// target ranges of one level and share most of their source cells, so each
// run of 256 tiles is dealt out as 8 blocks of 32 consecutive tiles, one per
// XCD: the tiles an XCD runs at the same time read the same sources from
// its L2.  The last partial run keeps the identity order (a bijection).
// The run length is a tuning knob (32 by default).
constexpr int NXCD = 8;
__device__ inline int tile_of_block(int b, int ntiles, int run) {
  const int S = NXCD * run;
  if (run <= 1 || b >= (ntiles / S) * S) return b;
  const int local = b % S;
  return (b - local) + (local % NXCD) * run + local / NXCD;
}

// Work units: a tile's items are cut into runs of about `chunk` items, one
// workgroup each, so that a tile with many sparse groups (a coarse leaf next
// to a much finer region) does not hold the whole launch back.  Unit k of a
// tile starts at the first segment (group) boundary at or after item
// k chunk: a segment never straddles two units, and since a target occurs
// at most once per segment, which of a target's terms a unit sums does not
// depend on the order of the items inside a segment (which the traversal's
// atomics leave open).  The first unit adds its products to L directly, the
// others leave theirs in part[] (one 128 x 110 slab per extra unit, FP64),
// which k_unit_combine adds in unit order after the launch, so the sums do
// not depend on timing.
__global__ void k_unit_counts(const int32_t *tile_off, int32_t ntiles, int32_t chunk,
                              uint32_t *ucnt, const int32_t *dchunk) {
  critical_path_priority();
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  if (dchunk) chunk = *dchunk;
  const int32_t n = tile_off[t + 1] - tile_off[t];
  ucnt[t] = n > chunk ? (uint32_t)((n + chunk - 1) / chunk) : 1u;
}

__global__ void k_unit_map(const uint32_t *ucnt, const uint32_t *uoff, int32_t ntiles,
                           int2 *units, uint32_t units_cap) {
  critical_path_priority();
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  for (uint32_t k = 0; k < ucnt[t] && uoff[t] + k < units_cap; k++)
    units[uoff[t] + k] = make_int2(t, (int)k);
}

// slab of extra unit k >= 1 of tile t: every tile has at least one unit
__device__ inline int64_t extra_slab(const uint32_t *uoff, int32_t t, int k) {
  return (int64_t)uoff[t] - t + k - 1;
}

// nslab: slabs the part buffer holds (units past them were not run; the
// caller checks the unit count against its capacity)
__global__ void k_unit_combine(const uint32_t *ucnt, const uint32_t *uoff, int32_t ntiles,
                               const double *part, int32_t ncells, double *L, int64_t nslab) {
  critical_path_priority();
  const int32_t t = blockIdx.x;
  int n = (int)ucnt[t];
  while (n > 1 && extra_slab(uoff, t, n - 1) >= nslab) n--;
  if (n <= 1) return;
  const int32_t first = t * TILE;
  for (int e = threadIdx.x; e < TILE * 110; e += blockDim.x) {
    const int sl = e / 110, row = e - sl * 110;
    const int32_t ci = first + sl;
    if (ci >= ncells) continue;
    double *dst = &L[2 * ((int64_t)ci * NTERM + (row >> 1)) + (row & 1)];
    double v = *dst;
    for (int k = 1; k < n; k++) v += part[extra_slab(uoff, t, k) * (TILE * 110) + e];
    *dst = v;
  }
}

// A fragments of every group from its harmonics: the entry that lane
// supplies at (block, k-step) is the sum of two table elements addressed by
// the operand code (byte offsets into the harmonics at multiples +1, -1, +2,
// -2); the multiples are exact, so each entry is formed with one rounding.
template <class T>
__device__ inline T table_elem(const T *y, uint32_t byte_off) {
  const uint32_t i = byte_off / (uint32_t)sizeof(T);
  const uint32_t m = i / NR;
  const T v = y[i - m * NR];
  return m == 0 ? v : m == 1 ? -v : m == 2 ? v + v : -(v + v);
}

template <class T>
__global__ void k_group_afrag(const T *Yg, uint32_t ng, const uint32_t *codes, const int8_t *gdl,
                              T *Af, const uint32_t *dng) {
  critical_path_priority();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (dng) ng = min(ng, *dng);
  if (t >= (int64_t)ng * (NBLK * 64)) return;
  const int64_t g = t / (NBLK * 64);
  const int r = (int)(t - g * (NBLK * 64));
  const int slot = r >> 6, lane = r & 63;
  int w = 0, e = slot;
  while (w < TILE_WAVES - 1 && e >= WAVE_NBLK[w]) {
    e -= WAVE_NBLK[w];
    w++;
  }
  const int b = WAVE_BLK[w][e];
  const T *y = Yg + g * NR;
  // row factor 2^(-j dl) of the entry's row (degree j)
  const int rexp = -row_degree(BLK_RB[b] * 16 + (lane & 15)) * (int)gdl[g];
  typename Vec4<T>::V out;
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const uint32_t cw = codes[(b * 4 + kk) * 64 + lane];
    out[kk] = ldexp(table_elem(y, cw & 0xffffu) + table_elem(y, cw >> 16), rexp);
  }
  *reinterpret_cast<typename Vec4<T>::V *>(Af + g * AF_PER_GROUP + (slot * 64 + lane) * 4) = out;
}

template <class T>
__global__ __launch_bounds__(64 * TILE_WAVES) void k_m2l_tiles(
    const int32_t *tile_off, const int2 *units, const uint32_t *ucnt, const uint32_t *uoff,
    int32_t ntiles, int32_t chunk, int xcd_run, const uint2 *cols, const T *Af, const T *Mq,
    Cells c, int32_t ncells, double *L, double *part, const int32_t *dchunk, int64_t nslab) {
  critical_path_priority();
  __shared__ T Lt[TILE][NR + 1];
  __shared__ T Lt0b[TILE][17];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int u = tile_of_block(blockIdx.x, gridDim.x, xcd_run);
  if ((uint32_t)u >= uoff[ntiles - 1] + ucnt[ntiles - 1]) return;  // grid is an upper bound
  const int2 tu = units[u];
  const int tile = tu.x, k = tu.y;
  if (k > 0 && extra_slab(uoff, tile, k) >= nslab) return;  // no slab (capacity exceeded)
  if (dchunk) chunk = *dchunk;
  for (int i = threadIdx.x; i < TILE * (NR + 1); i += 64 * TILE_WAVES) (&Lt[0][0])[i] = 0;
  for (int i = threadIdx.x; i < TILE * 17; i += 64 * TILE_WAVES) (&Lt0b[0][0])[i] = 0;
  __syncthreads();
  const int32_t tb = tile_off[tile], te = tile_off[tile + 1];
  // first segment start at or after item j of the tile (a segment holds at
  // most TILE / 16 items, so the walk is short)
  auto seg_start = [&](int64_t j) -> int32_t {
    if (j <= tb) return tb;
    if (j >= te) return te;
    int32_t i = (int32_t)j;
    while (i < te && (cols[(int64_t)i * 16].y >> COL_G) == (cols[(int64_t)(i - 1) * 16].y >> COL_G)) i++;
    return i;
  };
  const int32_t ib = seg_start((int64_t)tb + (int64_t)k * chunk);
  const int32_t ie = seg_start((int64_t)tb + (int64_t)(k + 1) * chunk);
  // every wave walks all items of the tile, in order, for its own blocks
  switch (wave) {
    case 0: run_wave<T, 0>(cols, ib, ie, Af, Mq, Lt, Lt0b, lane); break;
    case 1: run_wave<T, 1>(cols, ib, ie, Af, Mq, Lt, Lt0b, lane); break;
    case 2: run_wave<T, 2>(cols, ib, ie, Af, Mq, Lt, Lt0b, lane); break;
    default: run_wave<T, 3>(cols, ib, ie, Af, Mq, Lt, Lt0b, lane); break;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TILE * 16; e += 64 * TILE_WAVES) {
    const int sl = e >> 4, r = e & 15;
    Lt[sl][r] += Lt0b[sl][r];
  }
  __syncthreads();
  // L_j = L~_j / sigma_t^(j+1): added to the target's FP64 local expansion
  // by the tile's first unit, kept in the unit's slab by the others
  const int32_t first = tile * TILE;
  double *slab = k ? part + extra_slab(uoff, tile, k) * (TILE * 110) : nullptr;
  for (int e = threadIdx.x; e < TILE * 110; e += 64 * TILE_WAVES) {
    const int sl = e / 110, row = e - sl * 110;
    const int32_t ci = first + sl;
    if (ci >= ncells) continue;
    const int jk = row >> 1;
    int jj = 0;
    while (tri(jj + 1, 0) <= jk) jj++;
    const double v = (double)Lt[sl][row];
    double out = 0;
    if (v != 0) {  // (nothing arrived; a zero-width tree must not form 0 * inf)
      const double isig = 1.0 / c.width[ci];
      double sc = isig;
      for (int i = 0; i < jj; i++) sc *= isig;
      out = v * sc;
    }
    if (slab) slab[e] = out;
    else if (v != 0) L[2 * ((int64_t)ci * NTERM + jk) + (row & 1)] += out;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Operand codes of both element types, FP32 first (NCODE words each).
hipError_t m2l_upload_codes(uint32_t *dev_codes, hipStream_t s) {
  std::vector<uint32_t> c32, c64;
  if (!m2lc::build_codes(c32, 4) || !m2lc::build_codes(c64, 8)) return hipErrorInvalidValue;
  if (c32.size() != (size_t)NCODE || c64.size() != (size_t)NCODE) return hipErrorInvalidValue;
  c32.insert(c32.end(), c64.begin(), c64.end());
  hipError_t e = hipMemcpyAsync(dev_codes, c32.data(), c32.size() * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

size_t m2l_codes_bytes() { return (size_t)2 * NCODE * 4; }

int32_t m2l_group_slots(int max_depth) { return (max_depth + 1) * SLOTS_PER_LEVEL; }
int m2l_group_bits() { return 32 - COL_G; }

hipError_t launch_m2l_classify_all(const int2 *pairs, const int32_t *slot_of, uint32_t n,
                                   const uint32_t *gid_of_slot, int gbits, uint32_t *gkeys,
                                   uint32_t *gvals, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_classify_all<<<(n + 255) / 256, 256, 0, s>>>(pairs, slot_of, n, gid_of_slot, gbits, gkeys,
                                                 gvals);
  return hipGetLastError();
}

hipError_t launch_m2l_classify(const int2 *pairs, const int32_t *slot_of, uint32_t n,
                               const uint32_t *gid_of_slot, int gbits, int sbits,
                               uint32_t *gkeys, uint32_t *gvals, uint64_t *fkeys,
                               uint32_t *fvals, uint32_t *counters, hipStream_t s,
                               const uint32_t *dn, uint32_t gcap) {
  if (n == 0) return hipSuccess;
  k_classify<<<(n + 1023) / 1024, 1024, 0, s>>>(pairs, slot_of, n, gid_of_slot, gbits, sbits,
                                                  gkeys, gvals, fkeys, fvals, counters, dn,
                                                  gcap);
  k_grouped_count<<<1, 64, 0, s>>>(n, dn, counters);
  return hipGetLastError();
}

hipError_t launch_m2l_group_harmonics(const uint32_t *used, const uint32_t *gid_of_slot,
                                      int32_t nslots, double *Y64, float *Y32, int8_t *gdl,
                                      int8_t *glt, hipStream_t s, uint32_t gcap) {
  if (nslots <= 0) return hipSuccess;
  k_group_harmonics<<<(nslots + 127) / 128, 128, 0, s>>>(used, gid_of_slot, nslots, Y64, Y32,
                                                         gdl, glt, gcap);
  return hipGetLastError();
}

hipError_t launch_m2l_scale_m(Cells c, int32_t ncells, const double *M, void *Mq, bool fp32,
                              hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (fp32) k_scale_m<float><<<ncells, NR, 0, s>>>(c, ncells, M, (float *)Mq);
  else k_scale_m<double><<<ncells, NR, 0, s>>>(c, ncells, M, (double *)Mq);
  return hipGetLastError();
}

hipError_t launch_m2l_count_items(const uint32_t *keys, uint32_t n, uint32_t *ipos,
                                  uint32_t *bcount, uint32_t *bstart, void *scan_tmp,
                                  uint32_t *h_pinned, uint32_t *nitems_out, hipStream_t s) {
  *nitems_out = 0;
  if (n == 0) return hipSuccess;
  const unsigned per = 64 * IP_WAVES, nb = (n + per - 1) / per;
  k_item_pos<<<nb, per, 0, s>>>(keys, n, ipos, bcount, nullptr);
  hipError_t e = scan_exclusive_u32(bcount, bstart, nb, scan_tmp, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned, bstart + (nb - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned + 1, bcount + (nb - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  *nitems_out = h_pinned[0] + h_pinned[1];
  return hipGetLastError();
}

// the item count on the device: the block counts of k_item_pos summed
__global__ void k_items_total(const uint32_t *bcount, const uint32_t *bstart, uint32_t nb,
                              uint32_t *nitems) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *nitems = bstart[nb - 1] + bcount[nb - 1];
}

hipError_t launch_m2l_count_items_dev(const uint32_t *keys, uint32_t n, const uint32_t *dn,
                                      uint32_t *ipos, uint32_t *bcount, uint32_t *bstart,
                                      void *scan_tmp, uint32_t *nitems, hipStream_t s) {
  const unsigned per = 64 * IP_WAVES, nb = (std::max<uint32_t>(n, 1) + per - 1) / per;
  k_item_pos<<<nb, per, 0, s>>>(keys, n, ipos, bcount, dn);
  hipError_t e = scan_exclusive_u32(bcount, bstart, nb, scan_tmp, s);
  if (e != hipSuccess) return e;
  k_items_total<<<1, 64, 0, s>>>(bcount, bstart, nb, nitems);
  return hipGetLastError();
}

hipError_t launch_m2l_write_items(const uint32_t *keys, const uint32_t *vals, uint32_t n,
                                  int gbits, const uint32_t *ipos, const uint32_t *bstart,
                                  const int8_t *gdl, int4 *items, uint2 *cols, hipStream_t s,
                                  const uint32_t *dn, uint32_t items_cap) {
  if (n == 0) return hipSuccess;
  const unsigned per = 64 * IP_WAVES;
  k_items<<<(n + per - 1) / per, per, 0, s>>>(keys, vals, ipos, bstart, n, gbits, gdl, items,
                                             cols, dn, items_cap);
  return hipGetLastError();
}

hipError_t launch_m2l_tile_offsets(const int4 *items, uint32_t nitems, int32_t ntiles,
                                   int32_t *off, hipStream_t s, const uint32_t *dnitems) {
  k_tile_offsets<<<(ntiles + 1 + 255) / 256, 256, 0, s>>>(items, nitems, ntiles, off, dnitems);
  return hipGetLastError();
}

// m2l_unit_chunk on the device, from the item and tile counts there
__global__ void k_unit_chunk(const uint32_t *nitems, const uint32_t *ntiles, double factor,
                             int32_t *chunk) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (factor <= 0) {
    *chunk = INT32_MAX / 2;
    return;
  }
  const double mean = *ntiles > 0 ? (double)*nitems / (double)*ntiles : 1.0;
  *chunk = (int32_t)fmin((double)(INT32_MAX / 2), fmax(64.0, ceil(mean * factor)));
}

hipError_t launch_m2l_unit_chunk(const uint32_t *nitems, const uint32_t *ntiles, double factor,
                                 int32_t *chunk, hipStream_t s) {
  k_unit_chunk<<<1, 64, 0, s>>>(nitems, ntiles, factor, chunk);
  return hipGetLastError();
}

int32_t m2l_unit_chunk(uint32_t nitems, int32_t ntiles, double factor) {
  if (factor <= 0) return INT32_MAX / 2;  // one unit per tile
  const double mean = ntiles > 0 ? (double)nitems / ntiles : 1.0;
  return (int32_t)std::min<double>(INT32_MAX / 2, std::max(64.0, std::ceil(mean * factor)));
}

uint32_t m2l_units_bound(uint32_t nitems, int32_t ntiles, int32_t chunk) {
  return (uint32_t)ntiles + nitems / (uint32_t)chunk;
}

size_t m2l_unit_part_bytes(uint32_t units_bound, int32_t ntiles) {
  const size_t extra = units_bound > (uint32_t)ntiles ? units_bound - (uint32_t)ntiles : 0;
  return std::max<size_t>(extra, 1) * TILE * 110 * 8;
}

hipError_t launch_m2l_units(const int32_t *tile_off, int32_t ntiles, int32_t chunk,
                            uint32_t *ucnt, uint32_t *uoff, int2 *units, void *scan_tmp,
                            hipStream_t s, uint32_t units_cap, const int32_t *dchunk) {
  if (ntiles <= 0) return hipSuccess;
  k_unit_counts<<<(ntiles + 255) / 256, 256, 0, s>>>(tile_off, ntiles, chunk, ucnt, dchunk);
  hipError_t e = scan_exclusive_u32(ucnt, uoff, ntiles, scan_tmp, s);
  if (e != hipSuccess) return e;
  k_unit_map<<<(ntiles + 255) / 256, 256, 0, s>>>(ucnt, uoff, ntiles, units, units_cap);
  return hipGetLastError();
}

size_t m2l_afrag_bytes(uint32_t ngroups, bool fp32) {
  return (size_t)std::max<uint32_t>(ngroups, 1) * AF_PER_GROUP * (fp32 ? 4 : 8);
}

hipError_t launch_m2l_group_afrag(const void *Yg, uint32_t ngroups, const uint32_t *codes,
                                  const int8_t *gdl, void *Af, bool fp32, hipStream_t s,
                                  const uint32_t *dng) {
  if (ngroups == 0) return hipSuccess;
  const int64_t n = (int64_t)ngroups * NBLK * 64;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (fp32)
    k_group_afrag<float><<<nb, 256, 0, s>>>((const float *)Yg, ngroups, codes, gdl, (float *)Af,
                                            dng);
  else
    k_group_afrag<double><<<nb, 256, 0, s>>>((const double *)Yg, ngroups, codes + NCODE, gdl,
                                             (double *)Af, dng);
  return hipGetLastError();
}

hipError_t launch_m2l_tiles(const int32_t *tile_off, const int2 *units, const uint32_t *ucnt,
                            const uint32_t *uoff, int32_t ntiles, int32_t chunk, int xcd_run,
                            uint32_t units_bound, const uint2 *cols, const void *Af,
                            const void *Mq, Cells c, int32_t ncells, double *L, double *part,
                            bool fp32, hipStream_t s, const int32_t *dchunk) {
  if (ntiles <= 0) return hipSuccess;
  // slabs of the part buffer (m2l_unit_part_bytes)
  const int64_t nslab = units_bound > (uint32_t)ntiles ? (int64_t)units_bound - ntiles : 0;
  if (fp32)
    k_m2l_tiles<float><<<units_bound, 64 * TILE_WAVES, 0, s>>>(
        tile_off, units, ucnt, uoff, ntiles, chunk, xcd_run, cols, (const float *)Af,
        (const float *)Mq, c, ncells, L, part, dchunk, nslab);
  else
    k_m2l_tiles<double><<<units_bound, 64 * TILE_WAVES, 0, s>>>(
        tile_off, units, ucnt, uoff, ntiles, chunk, xcd_run, cols, (const double *)Af,
        (const double *)Mq, c, ncells, L, part, dchunk, nslab);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_unit_combine<<<ntiles, 256, 0, s>>>(ucnt, uoff, ntiles, part, ncells, L, nslab);
  return hipGetLastError();
}

int m2l_tile_size() { return TILE; }
int m2l_src_bits() { return SRC_BITS; }

}  // namespace tfmm
