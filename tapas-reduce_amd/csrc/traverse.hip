// traverse.hip -- dual tree traversal on the GPU.
//
// The reference recursion FMM_DTT / tapas_splitCell
// (sample/exafmm-dev-13274dd4ac68/examples/tapas_exafmm.cxx:209-286) visits a
// deterministic set of (target, source) cell pairs:
//   R2 = |centre_i - centre_j|^2  (TP/geometry.h:80-83, TP/vec.h:52-58)
//   R  = (max_d width_d / 2 * 1.00001) / theta
//   R2 > (Ri + Rj)^2          -> M2L(Ci, Cj)
//   both leaves               -> P2P(Ci, Cj)   (skipped if either is empty)
//   Cj leaf                   -> (children of Ci) x Cj
//   Ci leaf                   -> Ci x (children of Cj)
//   mutual and Ci == Cj       -> (children a) x (children b), a <= b
//   Ri >= Rj                  -> (children of Ci) x Cj
//   else                      -> Ci x (children of Cj)
// With mutual (the reference's --mutual 1) every pair is an unordered pair
// whose M2L and P2P act on both cells (LaplaceSphericalCPU_tapas.h:208-273,
// LaplaceP2PCPU_tapas.h:66-71; the a <= b product is TP/hot/mapper.h:29-121):
// such a pair is emitted into the lists in both directions, (i <- j) and
// (j <- i), so the target-major M2L and P2P kernels serve it unchanged.
// Here the recursion is expanded breadth first: one launch per frontier,
// each pair either emits list entries or its 8 (36) successors.
// Counters: *nout the next frontier; cnt (uint32 words) 1 M2L list entries
// executed, 2 P2P list entries, 3 M2L calls (numM2L: a mutual pair once,
// empty cells included), 4 P2P leaf-pair calls, words 8-11 two uint64 sums:
// body pairs the reference's P2P functor visits (numP2P: i x j, or the
// j <= i pairs of a leaf with itself under mutual) and the r = 0 pairs of a
// body with itself among them.  The input size is read on the device
// (*nin_dev, at most nin_max), so consecutive steps need no host round trip.
// The MAC is evaluated in FP64 without contraction, in the reference's
// operation order, so every decision is bit-identical to the reference's.
//
// The step reads each cell through one 32-byte record (centre, MAC radius R)
// and one 8-byte record (first child, body count), built once per traversal,
// instead of six separate arrays: the traversal gathers cells at random, so
// the number of cache lines a pair touches sets its cost.
#include "m2l_groups.h"
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

constexpr int DTT_THREADS = 1024;
constexpr int DTT_PPT = 4;  // frontier pairs per thread
constexpr int DTT_WAVES = DTT_THREADS / 64;

// per-cell traversal records: (centre, R) with R = (width / 2 * 1.00001) / theta
// (tapas_exafmm.cxx:222-226), and (first child, body count)
__global__ void k_dtt_cells(Cells c, int32_t ncells, double theta, double4 *geo, int2 *topo) {
#pragma clang fp contract(off)
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells) return;
  double R = c.width[i];
  R = (R / 2 * 1.00001) / theta;
  geo[i] = make_double4(c.center[3 * (int64_t)i], c.center[3 * (int64_t)i + 1],
                        c.center[3 * (int64_t)i + 2], R);
  topo[i] = make_int2(c.child[i], c.nb[i]);
}

// Kind of a frontier pair after the MAC: what the step emits for it.
enum : int { DK_NONE = 0, DK_M2L, DK_FAR_EMPTY, DK_P2P, DK_SPLIT_I, DK_SPLIT_J, DK_SPLIT_BOTH };

__global__ __launch_bounds__(DTT_THREADS) void k_dtt_step(const int2 *in, const uint32_t *nin_dev,
                                                  uint32_t nin_max, int2 *out, uint32_t out_cap,
                                                  uint32_t *nout, int2 *m2l, uint32_t m2l_cap,
                                                  int2 *p2p, uint32_t p2p_cap, uint32_t *cnt,
                                                  const double4 *geo, const int2 *topo,
                                                  const int32_t *depth, const double *width,
                                                  int32_t *m2l_slot, uint32_t *used,
                                                  uint32_t *n_outside, int mutual, double cycle) {
#pragma clang fp contract(off)
  const uint32_t nin = nin_dev ? *nin_dev : nin_max;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  // Each thread takes DTT_PPT pairs (rows r: pair blockIdx.x * DTT_THREADS *
  // DTT_PPT + r * DTT_THREADS + threadIdx.x).  Pass 1 classifies them and
  // counts the block's appends to the four lists (0 next frontier, 8 entries
  // per pair; 1 executed M2L; 2 P2P; 3 all M2L, counted only) per row and
  // wave; one global atomic per list per block reserves the block's ranges
  // (same-address atomics are what bound this kernel); pass 2 writes.
  __shared__ uint32_t wc[DTT_PPT][DTT_WAVES][5];
  __shared__ unsigned long long ws[DTT_WAVES][2];
  int32_t pi[DTT_PPT], pj[DTT_PPT], kind[DTT_PPT], fch[DTT_PPT], grp[DTT_PPT], grp2[DTT_PPT];
  bool dbl[DTT_PPT];
  unsigned long long v = 0, self = 0;
#pragma unroll
  for (int r = 0; r < DTT_PPT; r++) {
    const uint32_t t = (blockIdx.x * DTT_PPT + r) * DTT_THREADS + threadIdx.x;
    const bool valid = t < nin && t < nin_max;
    int32_t i = 0, j = 0, k = DK_NONE, f = 0, g = -1, g2 = -1;
    bool two = false;  // the pair's entries go into the lists in both directions
    if (valid) {
      const int2 pr = in[t];
      i = pr.x;
      j = pr.y;  // source cell and image (periodic images)
      const int32_t jc = (int32_t)((uint32_t)j & CELL_MASK);
      const uint32_t img = (uint32_t)j >> IMG_SHIFT;
      double sx, sy, sz;
      image_shift(img, cycle, sx, sy, sz);
      const bool same = i == jc && (cycle == 0 || img == IMG_CENTRE);  // a cell with itself
      const double4 gi = geo[i], gj = geo[jc];
      const int2 ti = topo[i], tj = topo[jc];
      // dX = Xi - Xj - Xperiodic (EX/include/traversal.h:37)
      const double d0 = gi.x - gj.x - sx;
      const double d1 = gi.y - gj.y - sy;
      const double d2 = gi.z - gj.z - sz;
      double R2 = 0;
      R2 += d0 * d0;
      R2 += d1 * d1;
      R2 += d2 * d2;
      const double Ri = gi.w, Rj = gj.w;
      const bool leaf_i = ti.x < 0, leaf_j = tj.x < 0;
      const bool nonempty = ti.y > 0 && tj.y > 0;
      if (R2 > (Ri + Rj) * (Ri + Rj)) {
        k = nonempty ? DK_M2L : DK_FAR_EMPTY;
        two = mutual != 0;
        if (nonempty) {
          // the pair's M2L offset group (same arithmetic on the same centres
          // as a separate pass over the list would do)
          const double ci[3] = {gi.x, gi.y, gi.z}, cj[3] = {gj.x + sx, gj.y + sy, gj.z + sz};
          g = m2lg::pair_slot(depth[i], depth[jc], width[i], width[jc], ci, cj);
          if (two) g2 = m2lg::pair_slot(depth[jc], depth[i], width[jc], width[i], cj, ci);
        }
      } else if (leaf_i && leaf_j) {
        if (nonempty) {
          k = DK_P2P;
          two = mutual != 0 && !same;
          if (mutual != 0 && same) {
            v += (unsigned long long)ti.y * (unsigned long long)(ti.y + 1) / 2;
          } else {
            v += (unsigned long long)ti.y * (unsigned long long)tj.y;
          }
          if (same) self += (unsigned long long)ti.y;
        }
      } else if (leaf_j) {
        k = DK_SPLIT_I;
        f = ti.x;
      } else if (leaf_i) {
        k = DK_SPLIT_J;
        f = tj.x | (int32_t)(img << IMG_SHIFT);
      } else if (mutual != 0 && same) {
        k = DK_SPLIT_BOTH;
        f = ti.x;
      } else if (Ri >= Rj) {
        k = DK_SPLIT_I;
        f = ti.x;
      } else {
        k = DK_SPLIT_J;
        f = tj.x | (int32_t)(img << IMG_SHIFT);
      }
    }
    pi[r] = i;
    pj[r] = j;
    kind[r] = k;
    fch[r] = f;
    grp[r] = g;
    grp2[r] = g2;
    dbl[r] = two;
    const uint64_t m0 = __ballot(k == DK_SPLIT_I || k == DK_SPLIT_J);
    const uint64_t m0b = __ballot(k == DK_SPLIT_BOTH);
    const uint64_t m1 = __ballot(k == DK_M2L), m1d = __ballot(k == DK_M2L && two);
    const uint64_t m2 = __ballot(k == DK_P2P), m2d = __ballot(k == DK_P2P && two);
    const uint64_t m3 = __ballot(k == DK_M2L || k == DK_FAR_EMPTY);
    if (lane == 0) {
      wc[r][wid][0] = (uint32_t)__popcll(m0) * 8u + (uint32_t)__popcll(m0b) * 36u;
      wc[r][wid][1] = (uint32_t)(__popcll(m1) + __popcll(m1d));
      wc[r][wid][2] = (uint32_t)(__popcll(m2) + __popcll(m2d));
      wc[r][wid][3] = (uint32_t)__popcll(m3);
      wc[r][wid][4] = (uint32_t)__popcll(m2);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, 64);
    self += __shfl_xor(self, o, 64);
  }
  if (lane == 0) {
    ws[wid][0] = v;
    ws[wid][1] = self;
  }
  __syncthreads();
  // one thread per list: the block's total, its global range, and the
  // (row, wave) counts turned into absolute starts
  if (threadIdx.x < 5) {
    const int q = threadIdx.x;
    uint32_t tot = 0;
    for (int r = 0; r < DTT_PPT; r++)
      for (int w = 0; w < DTT_WAVES; w++) tot += wc[r][w][q];
    uint32_t run = tot ? atomicAdd(q == 0 ? nout : &cnt[q], tot) : 0u;
    for (int r = 0; r < DTT_PPT; r++)
      for (int w = 0; w < DTT_WAVES; w++) {
        const uint32_t c = wc[r][w][q];
        wc[r][w][q] = run;
        run += c;
      }
  } else if (threadIdx.x == 64) {
    unsigned long long tv = 0, ts = 0;
    for (int w = 0; w < DTT_WAVES; w++) {
      tv += ws[w][0];
      ts += ws[w][1];
    }
    unsigned long long *pc = reinterpret_cast<unsigned long long *>(cnt + 8);
    if (tv) atomicAdd(&pc[0], tv);
    if (ts) atomicAdd(&pc[1], ts);
  }
  __syncthreads();
  uint32_t outside = 0;
#pragma unroll
  for (int r = 0; r < DTT_PPT; r++) {
    const int k = kind[r];
    const int32_t i = pi[r], j = pj[r];
    const bool two = dbl[r];
    const bool split = k == DK_SPLIT_I || k == DK_SPLIT_J;
    const uint64_t m0 = __ballot(split), m0b = __ballot(k == DK_SPLIT_BOTH);
    const uint64_t m1 = __ballot(k == DK_M2L), m1d = __ballot(k == DK_M2L && two);
    const uint64_t m2 = __ballot(k == DK_P2P), m2d = __ballot(k == DK_P2P && two);
    if (k == DK_M2L) {
      const uint32_t sl =
          wc[r][wid][1] + (uint32_t)(__popcll(m1 & lt) + __popcll(m1d & lt));
      for (int e = 0; e < (two ? 2 : 1); e++) {
        const int32_t ge = e ? grp2[r] : grp[r];
        if (sl + e < m2l_cap) {
          m2l[sl + e] = e ? make_int2(j, i) : make_int2(i, j);
          m2l_slot[sl + e] = ge;
          if (ge >= 0) used[ge] = 1u;
          else outside++;
        }
      }
    } else if (k == DK_P2P) {
      const uint32_t sl =
          wc[r][wid][2] + (uint32_t)(__popcll(m2 & lt) + __popcll(m2d & lt));
      if (sl < p2p_cap) p2p[sl] = make_int2(i, j);
      if (two && sl + 1 < p2p_cap) p2p[sl + 1] = make_int2(j, i);
    } else if (split || k == DK_SPLIT_BOTH) {
      const uint32_t sl =
          wc[r][wid][0] + (uint32_t)__popcll(m0 & lt) * 8u + (uint32_t)__popcll(m0b & lt) * 36u;
      const int32_t f = fch[r];
      if (k == DK_SPLIT_BOTH) {
        if (sl + 36 <= out_cap) {
          uint32_t o = sl;
          for (int a = 0; a < 8; a++)
            for (int b = a; b < 8; b++) out[o++] = make_int2(f + a, f + b);
        }
      } else if (sl + 8 <= out_cap) {
        if (k == DK_SPLIT_I) {
          for (int c = 0; c < 8; c++) out[sl + c] = make_int2(f + c, j);
        } else {
          for (int c = 0; c < 8; c++) out[sl + c] = make_int2(i, f + c);
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) outside += __shfl_xor(outside, o, 64);
  if (lane == 0 && outside) atomicAdd(n_outside, outside);
}

// Group-table slots of M2L pairs that come from outside the traversal (the
// host recursion of the reference's own FMM_DTT functor): the same slot, used
// flag and outside count a traversal step gives the pairs it emits.
__global__ void k_pair_slots(const int2 *pairs, uint32_t n, const double4 *geo,
                             const int32_t *depth, const double *width, int32_t *slot,
                             uint32_t *used, uint32_t *n_outside) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int2 pr = pairs[t];
  const double4 gi = geo[pr.x], gj = geo[pr.y];
  const double ci[3] = {gi.x, gi.y, gi.z}, cj[3] = {gj.x, gj.y, gj.z};
  const int32_t g = m2lg::pair_slot(depth[pr.x], depth[pr.y], width[pr.x], width[pr.y], ci, cj);
  slot[t] = g;
  if (g >= 0) used[g] = 1u;
  else atomicAdd(n_outside, 1u);
}

__global__ void k_unpack_pairs(const uint64_t *keys, uint32_t n, int sbits, int32_t *src,
                               const uint32_t *dn) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || (dn && t >= *dn)) return;
  src[t] = (int32_t)(keys[t] & ((1ull << sbits) - 1ull));
}

// off[t] = first entry whose target is >= t (t = 0..ncells)
__global__ void k_csr_offsets(const uint64_t *keys, uint32_t n, int sbits, int32_t ncells,
                              int32_t *off, const uint32_t *dn) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ncells) return;
  if (dn) n = min(n, *dn);
  const uint64_t v = (uint64_t)(uint32_t)t << sbits;
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  off[t] = (int32_t)lo;
}

// Leaf-pair list in CSR form, in two steps: a stable radix sort of the
// pairs by target only, then every target's sources sorted by one wave
// (rank = number of smaller sources, ties by position).  The result is the
// (target, source) order of a sort of the whole pair, with half the radix
// passes.
__global__ void k_split_pairs(const int2 *pairs, uint32_t n, uint32_t *tgt, uint32_t *src,
                              const uint32_t *dn) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || (dn && t >= *dn)) return;
  const int2 p = pairs[t];
  tgt[t] = (uint32_t)p.x;
  src[t] = (uint32_t)p.y;
}

// off[t] = first entry whose target is >= t (t = 0..ncells)
__global__ void k_csr_offsets32(const uint32_t *tgt, uint32_t n, int32_t ncells, int32_t *off,
                                const uint32_t *dn) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ncells) return;
  if (dn) n = min(n, *dn);
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (tgt[mid] < (uint32_t)t) lo = mid + 1;
    else hi = mid;
  }
  off[t] = (int32_t)lo;
}

constexpr int SEG_CAP = 2048;  // sources of one target staged in LDS

template <class P>
__device__ inline void rank_sources(P v, int32_t n, int lane, int32_t *out) {
  for (int32_t i = lane; i < n; i += 64) {
    const uint32_t x = v[i];
    int32_t r = 0;
    for (int32_t k = 0; k < n; k++) {
      const uint32_t y = v[k];
      r += (y < x || (y == x && k < i)) ? 1 : 0;
    }
    out[r] = (int32_t)x;
  }
}

// one wave per target: its sources in ascending order (up to 64 sorted in
// registers, up to SEG_CAP ranked from LDS); a longer list is queued in
// longq (word 0 the count, then the targets) for k_seg_sort_long
__global__ __launch_bounds__(64) void k_seg_sort(const int32_t *off, const uint32_t *in,
                                                 int32_t *out, uint32_t *longq) {
  __shared__ uint32_t v[SEG_CAP];
  const int lane = threadIdx.x;
  const int32_t b = off[blockIdx.x], n = off[blockIdx.x + 1] - b;
  if (n <= 0) return;
  if (n > SEG_CAP) {
    if (lane == 0) longq[1 + atomicAdd(longq, 1u)] = blockIdx.x;
    return;
  }
  if (n <= 64) {
    // one source per lane: bitonic sort across the wave, ascending (lanes
    // past n hold a value above every cell index)
    uint32_t x = lane < n ? in[b + lane] : 0xffffffffu;
    for (int k = 2; k <= 64; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        const uint32_t y = __shfl_xor(x, j, 64);
        const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
        x = keep_min ? min(x, y) : max(x, y);
      }
    if (lane < n) out[b + lane] = (int32_t)x;
  } else {
    for (int32_t i = lane; i < n; i += 64) v[i] = in[b + i];
    __syncthreads();
    rank_sources(static_cast<const uint32_t *>(v), n, lane, out + b);
  }
}

// The queued long lists, one workgroup of 1024 threads at a time: up to
// LONG_CAP sources bitonic-sorted in LDS (padded with the largest word),
// longer ones ranked by the whole workgroup from global memory.  Sources of
// one target are distinct words, so the order is unique.
constexpr int LONG_CAP = 16384;
constexpr int LONG_THREADS = 1024;

__global__ __launch_bounds__(LONG_THREADS) void k_seg_sort_long(const int32_t *off,
                                                                const uint32_t *in, int32_t *out,
                                                                const uint32_t *longq) {
  __shared__ uint32_t v[LONG_CAP];
  const uint32_t nq = longq[0];
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const int32_t t = (int32_t)longq[1 + q];
    const int32_t b = off[t], n = off[t + 1] - b;
    if (n <= LONG_CAP) {
      int32_t m = 1;
      while (m < n) m <<= 1;
      for (int32_t i = threadIdx.x; i < m; i += LONG_THREADS) v[i] = i < n ? in[b + i] : ~0u;
      __syncthreads();
      for (int32_t k = 2; k <= m; k <<= 1)
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
          for (int32_t i = threadIdx.x; i < m; i += LONG_THREADS) {
            const int32_t p = i ^ j;
            if (p > i) {
              const uint32_t a = v[i], c = v[p];
              if ((a > c) == ((i & k) == 0)) {
                v[i] = c;
                v[p] = a;
              }
            }
          }
          __syncthreads();
        }
      for (int32_t i = threadIdx.x; i < n; i += LONG_THREADS) out[b + i] = (int32_t)v[i];
    } else {
      for (int32_t i = threadIdx.x; i < n; i += LONG_THREADS) {
        const uint32_t x = in[b + i];
        int32_t r = 0;
        for (int32_t k = 0; k < n; k++) {
          const uint32_t y = in[b + k];
          r += (y < x || (y == x && k < i)) ? 1 : 0;
        }
        out[b + r] = (int32_t)x;
      }
    }
    __syncthreads();  // v is reused by the next list
  }
}

}  // namespace

hipError_t launch_split_pairs(const int2 *pairs, uint32_t n, uint32_t *tgt, uint32_t *src,
                              hipStream_t s, const uint32_t *dn) {
  if (n == 0) return hipSuccess;
  k_split_pairs<<<(n + 255) / 256, 256, 0, s>>>(pairs, n, tgt, src, dn);
  return hipGetLastError();
}

// off (ncells + 1) from the target-sorted list; out: each target's sources
// (in, in the target-sorted order) in ascending order; longq: ncells + 1
// words of scratch (the queue of lists longer than one wave sorts)
hipError_t launch_csr_sorted_sources(const uint32_t *tgt, const uint32_t *in, uint32_t n,
                                     int32_t ncells, int32_t *off, int32_t *out, uint32_t *longq,
                                     hipStream_t s, const uint32_t *dn) {
  k_csr_offsets32<<<(ncells + 1 + 255) / 256, 256, 0, s>>>(tgt, n, ncells, off, dn);
  if (n > 0 && ncells > 0) {
    hipError_t e = launch_fill_zero(longq, 4, s);
    if (e != hipSuccess) return e;
    k_seg_sort<<<ncells, 64, 0, s>>>(off, in, out, longq);
    k_seg_sort_long<<<64, LONG_THREADS, 0, s>>>(off, in, out, longq);
  }
  return hipGetLastError();
}

hipError_t launch_dtt_cells(Cells c, int32_t ncells, double theta, double4 *geo, int2 *topo,
                            hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_dtt_cells<<<(ncells + 255) / 256, 256, 0, s>>>(c, ncells, theta, geo, topo);
  return hipGetLastError();
}

hipError_t launch_dtt_step(const int2 *in, const uint32_t *nin_dev, uint32_t nin_max, int2 *out,
                           uint32_t out_cap, uint32_t *nout, int2 *m2l, uint32_t m2l_cap,
                           int2 *p2p, uint32_t p2p_cap, uint32_t *cnt, const double4 *geo,
                           const int2 *topo, const int32_t *depth, const double *width,
                           int32_t *m2l_slot, uint32_t *used, uint32_t *n_outside, int mutual,
                           hipStream_t s, double cycle) {
  if (nin_max == 0) return hipSuccess;
  const uint32_t per = DTT_THREADS * DTT_PPT;
  k_dtt_step<<<(nin_max + per - 1) / per, DTT_THREADS, 0, s>>>(
      in, nin_dev, nin_max, out, out_cap, nout, m2l, m2l_cap, p2p, p2p_cap, cnt, geo, topo, depth,
      width, m2l_slot, used, n_outside, mutual, cycle);
  return hipGetLastError();
}

hipError_t launch_pair_slots(const int2 *pairs, uint32_t n, const double4 *geo,
                             const int32_t *depth, const double *width, int32_t *slot,
                             uint32_t *used, uint32_t *n_outside, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_pair_slots<<<(n + 255) / 256, 256, 0, s>>>(pairs, n, geo, depth, width, slot, used, n_outside);
  return hipGetLastError();
}

hipError_t launch_unpack_pairs(const uint64_t *keys, uint32_t n, int sbits, int32_t *src,
                               hipStream_t s, const uint32_t *dn) {
  if (n == 0) return hipSuccess;
  k_unpack_pairs<<<(n + 255) / 256, 256, 0, s>>>(keys, n, sbits, src, dn);
  return hipGetLastError();
}

hipError_t launch_csr_offsets(const uint64_t *sorted_keys, uint32_t n, int sbits,
                              int32_t ncells, int32_t *off, hipStream_t s, const uint32_t *dn) {
  k_csr_offsets<<<(ncells + 1 + 255) / 256, 256, 0, s>>>(sorted_keys, n, sbits, ncells, off, dn);
  return hipGetLastError();
}

}  // namespace tfmm
