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
// Scaling: with sigma_l = w_0 2^-l (w_0 the root width), the kernel works on
//   M~_n = M_n / sigma_s^n,  L~_j = L_j sigma_t^(j+1),  Y~_p = Y_p sigma_t^(p+1)
// so that L~_j = sum_n 2^(n (l_t - l_s)) Y~_(j+n) M~_n with all quantities
// O(1): FP32 neither under- nor overflows at P = 10 on deep trees, and the
// power-of-two factors are exact.
//
// Work decomposition: targets are tiled by 128 consecutive cells (breadth
// first, so one tile is one Morton range of one level); the pairs of a tile
// are sorted by group, and each run of up to 16 pairs of one group is a work
// item = 18 blocks x 4 MFMA 16x16x4 steps.  One workgroup (4 waves) owns a
// tile: it accumulates the tile's local expansions in LDS (LDS atomics, since
// waves work on different items) and writes each target once, without
// global atomics.  Pairs whose offset falls outside the group table (level
// difference >= 2) go to the per-pair kernel of m2l.hip.
#include <algorithm>
#include <vector>

#include "tfmm_device.h"
#include "tfmm_kernels.h"
#include "tfmm_internal.h"

namespace tfmm {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NR = 112;  // padded real rows/columns
constexpr int NBLK = 18;
constexpr int NSLOT = NBLK * 4;
constexpr int TILE = 128;
constexpr int TILE_WAVES = 4;

// non-zero 16x16 blocks (row block, column block) of T
constexpr int8_t BLK_RB[NBLK] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 5, 6};
constexpr int8_t BLK_CB[NBLK] = {0, 1, 2, 3, 4, 5, 6, 0, 1, 2, 3, 0, 1, 0, 1, 0, 0, 0};

// ---------------------------------------------------------------------------
// group table
// ---------------------------------------------------------------------------
// Slot of a (target level, level difference, offset) in the direct-mapped
// table: offsets are in units of half the finer width; same-level offsets are
// even with |o| <= 14, adjacent-level offsets odd with |o| <= 15.
constexpr int SAME_HALF = 7;                 // |o/2| <= 7
constexpr int SAME_N = 2 * SAME_HALF + 1;    // 15
constexpr int CROSS_N = 16;                  // (o + 15) / 2 in [0, 15]
constexpr int SLOTS_SAME = SAME_N * SAME_N * SAME_N;
constexpr int SLOTS_CROSS = CROSS_N * CROSS_N * CROSS_N;
constexpr int SLOTS_PER_LEVEL = SLOTS_SAME + 2 * SLOTS_CROSS;

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
  } else if (dl == 1 || dl == -1) {
    int a[3];
    for (int d = 0; d < 3; d++) {
      if (!(o[d] & 1)) return -1;
      a[d] = (o[d] + 15) / 2;
      if (a[d] < 0 || a[d] >= CROSS_N) return -1;
    }
    idx = SLOTS_SAME + (dl == 1 ? 0 : SLOTS_CROSS) + (a[0] * CROSS_N + a[1]) * CROSS_N + a[2];
  } else {
    return -1;
  }
  return lt * SLOTS_PER_LEVEL + idx;
}

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
    ls = idx < SLOTS_CROSS ? lt + 1 : lt - 1;
    idx %= SLOTS_CROSS;
    o[2] = (idx % CROSS_N) * 2 - 15;
    idx /= CROSS_N;
    o[1] = (idx % CROSS_N) * 2 - 15;
    o[0] = (idx / CROSS_N) * 2 - 15;
  }
}

// offset of an M2L pair in half widths of the finer cell; slot or -1
__device__ inline int32_t pair_slot(const Cells &c, int32_t t, int32_t s) {
  const int lt = c.depth[t], ls = c.depth[s];
  const double half = 0.5 * fmin(c.width[t], c.width[s]);
  int o[3];
  for (int d = 0; d < 3; d++) {
    const double v = (c.center[3 * (int64_t)t + d] - c.center[3 * (int64_t)s + d]) / half;
    if (!(fabs(v) < 64.0)) return -1;
    o[d] = (int)rint(v);
  }
  return group_slot(lt, ls, o);
}

__global__ void k_mark_slots(Cells c, const int2 *pairs, uint32_t n, uint32_t *used) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 p = pairs[i];
  const int32_t s = pair_slot(c, p.x, p.y);
  if (s >= 0) used[s] = 1u;
}

// Split the executed M2L pairs into grouped ones (sort key: tile | group |
// slot-in-tile, value: source cell) and the rest (key: target | source).
__global__ __launch_bounds__(1024) void k_classify(Cells c, const int2 *pairs, uint32_t n,
                                                  const uint32_t *gid_of_slot, int gbits,
                                                  int sbits, uint64_t *gkeys, uint32_t *gvals,
                                                  uint64_t *fkeys, uint32_t *fvals,
                                                  uint32_t *counters) {
  __shared__ uint32_t scratch[17];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int2 p = make_int2(0, 0);
  int32_t slot = -1;
  if (valid) {
    p = pairs[i];
    slot = pair_slot(c, p.x, p.y);
  }
  const bool grouped = valid && slot >= 0;
  const bool fallback = valid && slot < 0;
  const uint32_t at = block_append<16>(&counters[0], grouped, 1, scratch);
  if (grouped) {
    const uint64_t tile = (uint32_t)p.x / TILE, in_tile = (uint32_t)p.x % TILE;
    gkeys[at] = (tile << (gbits + 7)) | ((uint64_t)gid_of_slot[slot] << 7) | in_tile;
    gvals[at] = (uint32_t)p.y;
  }
  const uint32_t af = block_append<16>(&counters[1], fallback, 1, scratch);
  if (fallback) {
    fkeys[af] = ((uint64_t)(uint32_t)p.x << sbits) | (uint64_t)(uint32_t)p.y;
    fvals[af] = 0;
  }
}

// Harmonics of every used group, FP64, scaled: Y~_p^q = Y_p^q sigma_t^(p+1).
// Stored as NR reals (re/im interleaved, index 2 tri(p, q)), in FP64 and FP32.
__global__ void k_group_harmonics(const uint32_t *used, const uint32_t *gid_of_slot,
                                  int32_t nslots, double w0, double *Y64, float *Y32,
                                  int8_t *gdl, int8_t *glt) {
  const int32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots || !used[slot]) return;
  const uint32_t g = gid_of_slot[slot];
  int lt, ls, o[3];
  slot_decode(slot, lt, ls, o);
  const int lf = lt > ls ? lt : ls;
  const double half = ldexp(w0, -lf) * 0.5;
  const double dx = o[0] * half, dy = o[1] * half, dz = o[2] * half;
  double rho, alpha, beta;
  cart2sph(dx, dy, dz, rho, alpha, beta);
  const double sig = ldexp(w0, -lt);
  const double x = cos(alpha), y = sin(alpha);
  const cd ei = mk(cos(beta), sin(beta));
  double *yo = Y64 + (int64_t)g * NR;
  float *yf = Y32 + (int64_t)g * NR;
  for (int m = 0; m < P; m++) {
    local_column(m, rho, x, y, ei, [&](int n, int mm, cd v) {
      double sc = sig;  // sigma^(n+1)
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
__global__ void k_scale_m(Cells c, int32_t ncells, const double *M, double w0, T *Mq) {
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
    const double isig = 1.0 / ldexp(w0, -c.depth[ci]);
    double sc = 1;
    for (int i = 0; i < n; i++) sc *= isig;
    v = M[2 * ((int64_t)ci * NTERM + tt) + (r & 1)] * sc;
  }
  Mq[(int64_t)ci * NR + q * 28 + t] = (T)v;
}

// ---------------------------------------------------------------------------
// work items
// ---------------------------------------------------------------------------
__global__ void k_seg_flags(const uint64_t *keys, uint32_t n, uint32_t *flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || (keys[i] >> 7) != (keys[i - 1] >> 7)) ? 1u : 0u;
}

__global__ void k_seg_starts(const uint32_t *flag, const uint32_t *excl, uint32_t n,
                             uint32_t *segstart) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  segstart[excl[i]] = i;
}

__global__ void k_item_flags(const uint32_t *flag, const uint32_t *excl,
                             const uint32_t *segstart, uint32_t n, uint32_t *iflag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t seg = excl[i] + flag[i] - 1;
  iflag[i] = ((i - segstart[seg]) % 16u == 0u) ? 1u : 0u;
}

__global__ void k_items(const uint64_t *keys, const uint32_t *flag, const uint32_t *excl,
                        const uint32_t *segstart, uint32_t nseg, const uint32_t *iflag,
                        const uint32_t *iexcl, uint32_t n, int gbits, int4 *items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !iflag[i]) return;
  const uint32_t seg = excl[i] + flag[i] - 1;
  const uint32_t segend = (seg + 1 < nseg) ? segstart[seg + 1] : n;
  const uint32_t cnt = min(16u, segend - i);
  const uint64_t k = keys[i];
  int4 it;
  it.x = (int)i;                                         // first pair
  it.y = (int)cnt;                                       // pairs in the item
  it.z = (int)((k >> 7) & ((1ull << gbits) - 1ull));     // group
  it.w = (int)(k >> (gbits + 7));                        // tile
  items[iexcl[i]] = it;
}

__global__ void k_tile_offsets(const int4 *items, uint32_t nitems, int32_t ntiles,
                               int32_t *off) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
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

// A-operand code of one (block, k-step, lane): T[r][c] = s1 Y~[i1] + s2 Y~[i2],
// times 2^(n dl) with n the column's degree.  s1, s2 in {-2..2} (the +m and
// -m terms of a k = 0 row can hit the same harmonic).  Bits: i1 0-6,
// s1 + 2 7-9, i2 10-16, s2 + 2 17-19, n 20-23.
__device__ inline float mult_of(uint32_t f) { return (float)f - 2.f; }

// A value of code cw from the group's harmonics in LDS
template <class T>
__device__ inline T a_value(const T *ys, uint32_t cw, int dl) {
  const T y1 = ys[cw & 127u] * (T)mult_of((cw >> 7) & 7u);
  const T y2 = ys[(cw >> 10) & 127u] * (T)mult_of((cw >> 17) & 7u);
  return ldexp(y1 + y2, (int)((cw >> 20) & 15u) * dl);
}

// Work split of one item over the 4 waves of a tile: each wave owns fixed
// (row block, column block) pairs of T, so the waves write disjoint rows of
// the tile's expansions and need no atomics and no barriers between items.
// Row block 0 has 7 column blocks and is split between waves 0 and 1 (wave 1
// accumulates its part in a separate array, added at the end).
//   wave 0: (0,0) (0,1) (0,2) (0,3) (6,0)
//   wave 1: (0,4) (0,5) (0,6) (4,0) (5,0)
//   wave 2: (1,0) (1,1) (1,2) (1,3)
//   wave 3: (2,0) (2,1) (3,0) (3,1)
constexpr int WAVE_NBLK[TILE_WAVES] = {5, 5, 4, 4};
constexpr int WAVE_BLK[TILE_WAVES][5] = {
    {0, 1, 2, 3, 17}, {4, 5, 6, 15, 16}, {7, 8, 9, 10, -1}, {11, 12, 13, 14, -1}};

template <class T, int W>
__device__ inline void wave_item(typename Mfma<T>::V *acc, const uint32_t *code, const T *ys,
                                 const T *mq, bool valid, int dl) {
  typedef typename Mfma<T>::V V;
  // B fragments of the column blocks this wave touches (4 k-steps each)
  T b[7][4];
#pragma unroll
  for (int cb = 0; cb < 7; cb++) {
    bool need = false;
#pragma unroll
    for (int e = 0; e < WAVE_NBLK[W]; e++) need |= BLK_CB[WAVE_BLK[W][e]] == cb;
    if (need) {
#pragma unroll
      for (int kk = 0; kk < 4; kk++) b[cb][kk] = valid ? mq[4 * cb + kk] : (T)0;
    }
  }
#pragma unroll
  for (int e = 0; e < WAVE_NBLK[W]; e++) acc[e] = (V){0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
#pragma unroll
    for (int e = 0; e < WAVE_NBLK[W]; e++) {
      const int blk = WAVE_BLK[W][e];
      const T a = a_value(ys, code[e * 4 + kk], dl);
      acc[e] = Mfma<T>::step(a, b[BLK_CB[blk]][kk], acc[e]);
    }
  }
}

// accumulate the wave's results of one item into the tile (rows of the
// wave's row blocks; the second half of row block 0 goes to Lt0b)
template <class T, int W>
__device__ inline void wave_store(const typename Mfma<T>::V *acc, T (*Lt)[NR + 1],
                                  T (*Lt0b)[17], uint32_t tslot, int q) {
  // blocks of one row block are summed first
  typename Mfma<T>::V sum[5];
  int rbs[5];
  int nrb = 0;
#pragma unroll
  for (int e = 0; e < WAVE_NBLK[W]; e++) {
    const int rb = BLK_RB[WAVE_BLK[W][e]];
    bool found = false;
#pragma unroll
    for (int f = 0; f < 5; f++) {
      if (f < nrb && rbs[f] == rb) {
        sum[f] += acc[e];
        found = true;
      }
    }
    if (!found) {
      rbs[nrb] = rb;
      sum[nrb] = acc[e];
      nrb++;
    }
  }
#pragma unroll
  for (int f = 0; f < 5; f++) {
    if (f >= nrb) break;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = Mfma<T>::row(q, i);
      if (W == 1 && rbs[f] == 0) {
        Lt0b[tslot][r] += sum[f][i];
      } else {
        const int row = rbs[f] * 16 + r;
        if (row < 110) Lt[tslot][row] += sum[f][i];
      }
    }
  }
}

template <class T, int W>
__device__ inline void run_wave(const int4 *items, int32_t ib, int32_t ie, const uint64_t *keys,
                                const uint32_t *src, const T *Yg, const int8_t *gdl,
                                const T *Mq, const uint32_t *codes, T (*Lt)[NR + 1],
                                T (*Lt0b)[17], T *ys, int lane) {
  typedef typename Mfma<T>::V V;
  uint32_t code[5 * 4];
#pragma unroll
  for (int e = 0; e < WAVE_NBLK[W]; e++)
#pragma unroll
    for (int kk = 0; kk < 4; kk++) code[e * 4 + kk] = codes[(WAVE_BLK[W][e] * 4 + kk) * 64 + lane];
  const int q = lane >> 4, j = lane & 15;
  for (int32_t it = ib; it < ie; it++) {
    const int4 item = items[it];
    const int g = item.z;
    const T *yg = Yg + (int64_t)g * NR;
    __builtin_amdgcn_wave_barrier();
    ys[lane] = yg[lane];
    if (lane < NR - 64) ys[64 + lane] = yg[64 + lane];
    const int dl = gdl[g];
    const bool valid = j < item.y;
    uint32_t s = 0, tslot = 0;
    if (valid) {
      s = src[item.x + j];
      tslot = (uint32_t)(keys[item.x + j] & 127u);
    }
    __builtin_amdgcn_wave_barrier();
    V acc[5];
    wave_item<T, W>(acc, code, ys, Mq + (int64_t)s * NR + q * 28, valid, dl);
    if (valid) wave_store<T, W>(acc, Lt, Lt0b, tslot, q);
  }
}

template <class T>
__global__ __launch_bounds__(64 * TILE_WAVES) void k_m2l_tiles(
    const int4 *items, const int32_t *tile_off, const uint64_t *keys, const uint32_t *src,
    const T *Yg, const int8_t *gdl, const T *Mq, const uint32_t *codes, Cells c, int32_t ncells,
    double w0, double *L) {
  __shared__ T Lt[TILE][NR + 1];
  __shared__ T Lt0b[TILE][17];
  __shared__ T Ys[TILE_WAVES][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = blockIdx.x;
  for (int i = threadIdx.x; i < TILE * (NR + 1); i += 64 * TILE_WAVES) (&Lt[0][0])[i] = 0;
  for (int i = threadIdx.x; i < TILE * 17; i += 64 * TILE_WAVES) (&Lt0b[0][0])[i] = 0;
  __syncthreads();
  const int32_t ib = tile_off[tile], ie = tile_off[tile + 1];
  // every wave walks all items of the tile, in order, for its own blocks
  switch (wave) {
    case 0: run_wave<T, 0>(items, ib, ie, keys, src, Yg, gdl, Mq, codes, Lt, Lt0b, Ys[0], lane); break;
    case 1: run_wave<T, 1>(items, ib, ie, keys, src, Yg, gdl, Mq, codes, Lt, Lt0b, Ys[1], lane); break;
    case 2: run_wave<T, 2>(items, ib, ie, keys, src, Yg, gdl, Mq, codes, Lt, Lt0b, Ys[2], lane); break;
    default: run_wave<T, 3>(items, ib, ie, keys, src, Yg, gdl, Mq, codes, Lt, Lt0b, Ys[3], lane); break;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TILE * 16; e += 64 * TILE_WAVES) {
    const int sl = e >> 4, r = e & 15;
    Lt[sl][r] += Lt0b[sl][r];
  }
  __syncthreads();
  // L_j = L~_j / sigma_t^(j+1), added to the target's FP64 local expansion
  const int32_t first = tile * TILE;
  for (int e = threadIdx.x; e < TILE * 110; e += 64 * TILE_WAVES) {
    const int sl = e / 110, row = e - sl * 110;
    const int32_t ci = first + sl;
    if (ci >= ncells) continue;
    const int jk = row >> 1;
    int jj = 0;
    while (tri(jj + 1, 0) <= jk) jj++;
    const double v = (double)Lt[sl][row];
    if (v == 0) continue;  // nothing arrived (and a zero-width tree must not form 0 * inf)
    const double isig = 1.0 / ldexp(w0, -c.depth[ci]);
    double sc = isig;
    for (int i = 0; i < jj; i++) sc *= isig;
    L[2 * ((int64_t)ci * NTERM + jk) + (row & 1)] += v * sc;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// A-operand codes: for block b, k-step kk and lane l, the entry
// T[r][c] with r = 16 rb + l % 16, c = 16 cb + 4 kk + l / 16.
static bool build_codes(std::vector<uint32_t> &codes) {
  // T entries as up to two (flat index, sign) terms
  struct Ent { int i1 = -1, s1 = 0, i2 = -1, s2 = 0; };
  std::vector<Ent> T((size_t)NR * NR);
  bool ok = true;
  auto add = [&](int r, int cc, int yi, int sg) {
    Ent &e = T[(size_t)r * NR + cc];
    if (e.i1 < 0) { e.i1 = yi; e.s1 = sg; }
    else if (e.i1 == yi) { e.s1 += sg; }
    else if (e.i2 < 0) { e.i2 = yi; e.s2 = sg; }
    else if (e.i2 == yi) { e.s2 += sg; }
    else { ok = false; }  // more than two harmonics in one entry
  };
  auto oddeven = [](int n) { return (n & 1) ? -1 : 1; };
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      const int jk = tri(j, k);
      const int Cnm = oddeven(j);
      for (int n = 0; n < P - j; n++) {
        for (int m = -n; m <= n; m++) {
          const int p = j + n, qq = m - k;
          // Y_p^qq = cy + i dy, stored for qq >= 0; conj for qq < 0
          const int base = 2 * tri(p, qq >= 0 ? qq : -qq);
          const int cy = base, dy = base + 1, dys = qq >= 0 ? 1 : -1;
          const int mm = m >= 0 ? m : -m;
          const int cn = 2 * tri(n, mm);
          int s = (m < 0) ? Cnm : Cnm * oddeven((k - m) * (k < m) + m);
          if (m >= 0) {
            add(2 * jk, cn, cy, s);
            add(2 * jk, cn + 1, dy, -s * dys);
            add(2 * jk + 1, cn, dy, s * dys);
            add(2 * jk + 1, cn + 1, cy, s);
          } else {
            add(2 * jk, cn, cy, s);
            add(2 * jk, cn + 1, dy, s * dys);
            add(2 * jk + 1, cn, dy, s * dys);
            add(2 * jk + 1, cn + 1, cy, -s);
          }
        }
      }
    }
  }
  auto enc = [&](int s) -> uint32_t {
    if (s < -2 || s > 2) ok = false;
    return (uint32_t)(s + 2);
  };
  auto degree_of_col = [](int cc) {
    if (cc >= 110) return 0;
    const int t = cc >> 1;
    int n = 0;
    while (tri(n + 1, 0) <= t) n++;
    return n;
  };
  codes.assign((size_t)NSLOT * 64, 0u);
  for (int b = 0; b < NBLK; b++) {
    for (int kk = 0; kk < 4; kk++) {
      for (int l = 0; l < 64; l++) {
        const int r = 16 * BLK_RB[b] + (l % 16);
        const int cc = 16 * BLK_CB[b] + 4 * kk + l / 16;
        const Ent e = T[(size_t)r * NR + cc];
        uint32_t w = 0;
        w |= (e.i1 >= 0 ? (uint32_t)e.i1 : 0u) | (enc(e.i1 >= 0 ? e.s1 : 0) << 7);
        w |= ((e.i2 >= 0 ? (uint32_t)e.i2 : 0u) << 10) | (enc(e.i2 >= 0 ? e.s2 : 0) << 17);
        w |= (uint32_t)degree_of_col(cc) << 20;
        codes[(size_t)(b * 4 + kk) * 64 + l] = w;
      }
    }
  }
  // every non-zero entry of T must lie in one of the listed blocks
  for (int r = 0; r < NR; r++) {
    for (int cc = 0; cc < NR; cc++) {
      const Ent &e = T[(size_t)r * NR + cc];
      if ((e.i1 < 0 || e.s1 == 0) && (e.i2 < 0 || e.s2 == 0)) continue;
      bool in = false;
      for (int b = 0; b < NBLK; b++)
        in |= (r / 16 == BLK_RB[b] && cc / 16 == BLK_CB[b]);
      ok &= in;
    }
  }
  return ok;
}

hipError_t m2l_upload_codes(uint32_t *dev_codes, hipStream_t s) {
  std::vector<uint32_t> codes;
  if (!build_codes(codes)) return hipErrorInvalidValue;
  hipError_t e = hipMemcpyAsync(dev_codes, codes.data(), codes.size() * 4,
                                hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

int32_t m2l_group_slots(int max_depth) { return (max_depth + 1) * SLOTS_PER_LEVEL; }

hipError_t launch_m2l_mark_slots(Cells c, const int2 *pairs, uint32_t n, uint32_t *used,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_mark_slots<<<(n + 255) / 256, 256, 0, s>>>(c, pairs, n, used);
  return hipGetLastError();
}

hipError_t launch_m2l_classify(Cells c, const int2 *pairs, uint32_t n,
                               const uint32_t *gid_of_slot, int gbits, int sbits,
                               uint64_t *gkeys, uint32_t *gvals, uint64_t *fkeys,
                               uint32_t *fvals, uint32_t *counters, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_classify<<<(n + 1023) / 1024, 1024, 0, s>>>(c, pairs, n, gid_of_slot, gbits, sbits, gkeys,
                                                  gvals, fkeys, fvals, counters);
  return hipGetLastError();
}

hipError_t launch_m2l_group_harmonics(const uint32_t *used, const uint32_t *gid_of_slot,
                                      int32_t nslots, double w0, double *Y64, float *Y32,
                                      int8_t *gdl, int8_t *glt, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  k_group_harmonics<<<(nslots + 127) / 128, 128, 0, s>>>(used, gid_of_slot, nslots, w0, Y64,
                                                         Y32, gdl, glt);
  return hipGetLastError();
}

hipError_t launch_m2l_scale_m(Cells c, int32_t ncells, const double *M, double w0, void *Mq,
                              bool fp32, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (fp32) k_scale_m<float><<<ncells, NR, 0, s>>>(c, ncells, M, w0, (float *)Mq);
  else k_scale_m<double><<<ncells, NR, 0, s>>>(c, ncells, M, w0, (double *)Mq);
  return hipGetLastError();
}

hipError_t launch_m2l_items(const uint64_t *keys, uint32_t n, int gbits, uint32_t *flag,
                            uint32_t *excl, uint32_t *segstart, uint32_t *iflag,
                            uint32_t *iexcl, void *scan_tmp, uint32_t *h_pinned, int4 *items,
                            size_t items_cap, uint32_t *nitems_out, hipStream_t s) {
  *nitems_out = 0;
  if (n == 0) return hipSuccess;
  const unsigned nb = (n + 255) / 256;
  k_seg_flags<<<nb, 256, 0, s>>>(keys, n, flag);
  hipError_t e = scan_exclusive_u32(flag, excl, n, scan_tmp, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned, excl + (n - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned + 1, flag + (n - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const uint32_t nseg = h_pinned[0] + h_pinned[1];
  k_seg_starts<<<nb, 256, 0, s>>>(flag, excl, n, segstart);
  k_item_flags<<<nb, 256, 0, s>>>(flag, excl, segstart, n, iflag);
  e = scan_exclusive_u32(iflag, iexcl, n, scan_tmp, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned, iexcl + (n - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(h_pinned + 1, iflag + (n - 1), 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const uint32_t nitems = h_pinned[0] + h_pinned[1];
  if (nitems > items_cap) {
    *nitems_out = nitems;
    return hipErrorOutOfMemory;
  }
  k_items<<<nb, 256, 0, s>>>(keys, flag, excl, segstart, nseg, iflag, iexcl, n, gbits, items);
  *nitems_out = nitems;
  return hipGetLastError();
}

hipError_t launch_m2l_tile_offsets(const int4 *items, uint32_t nitems, int32_t ntiles,
                                   int32_t *off, hipStream_t s) {
  k_tile_offsets<<<(ntiles + 1 + 255) / 256, 256, 0, s>>>(items, nitems, ntiles, off);
  return hipGetLastError();
}

hipError_t launch_m2l_tiles(const int4 *items, const int32_t *tile_off, int32_t ntiles,
                            const uint64_t *keys, const uint32_t *src, const void *Yg,
                            const int8_t *gdl, const void *Mq, const uint32_t *codes, Cells c,
                            int32_t ncells, double w0, double *L, bool fp32, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (fp32)
    k_m2l_tiles<float><<<ntiles, 64 * TILE_WAVES, 0, s>>>(items, tile_off, keys, src,
                                                          (const float *)Yg, gdl,
                                                          (const float *)Mq, codes, c, ncells,
                                                          w0, L);
  else
    k_m2l_tiles<double><<<ntiles, 64 * TILE_WAVES, 0, s>>>(items, tile_off, keys, src,
                                                           (const double *)Yg, gdl,
                                                           (const double *)Mq, codes, c, ncells,
                                                           w0, L);
  return hipGetLastError();
}

int m2l_tile_size() { return TILE; }

}  // namespace tfmm
