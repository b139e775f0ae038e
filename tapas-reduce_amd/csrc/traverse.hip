// traverse.hip -- dual tree traversal on the GPU.
//
// The reference recursion FMM_DTT / tapas_splitCell
// (sample/exafmm-dev-13274dd4ac68/examples/tapas_exafmm.cxx:209-286, with
// mutual = 0) visits a deterministic set of (target, source) cell pairs:
//   R2 = |centre_i - centre_j|^2  (TP/geometry.h:80-83, TP/vec.h:52-58)
//   R  = (max_d width_d / 2 * 1.00001) / theta
//   R2 > (Ri + Rj)^2          -> M2L(Ci, Cj)
//   both leaves               -> P2P(Ci, Cj)   (skipped if either is empty)
//   Cj leaf                   -> (children of Ci) x Cj
//   Ci leaf                   -> Ci x (children of Cj)
//   Ri >= Rj                  -> (children of Ci) x Cj
//   else                      -> Ci x (children of Cj)
// Here that recursion is expanded breadth first: one launch per frontier,
// each pair either emits an M2L entry, a P2P entry or its 8 successors.
// Counters: *nout the next frontier; cnt (uint32 words) 1 executed M2L pairs,
// 2 P2P pairs, 3 all M2L pairs, words 8-11 two uint64 sums: body-pair
// interactions of the P2P pairs and those of the self pairs.  The input size
// is read on the device (*nin_dev, at most nin_max), so consecutive steps
// need no host round trip.  The
// MAC is evaluated in FP64 without contraction, in the reference's operation
// order, so every decision is bit-identical to the reference's.
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

__global__ __launch_bounds__(DTT_THREADS) void k_dtt_step(const int2 *in, const uint32_t *nin_dev,
                                                  uint32_t nin_max, int2 *out, uint32_t out_cap,
                                                  uint32_t *nout, int2 *m2l, uint32_t m2l_cap,
                                                  int2 *p2p, uint32_t p2p_cap, uint32_t *cnt,
                                                  const double4 *geo, const int2 *topo,
                                                  const int32_t *depth, const double *width,
                                                  int32_t *m2l_slot, uint32_t *used,
                                                  uint32_t *n_outside) {
#pragma clang fp contract(off)
  const uint32_t nin = nin_dev ? *nin_dev : nin_max;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < nin && t < nin_max;
  int32_t i = 0, j = 0;
  int2 ti = make_int2(0, 0), tj = make_int2(0, 0);
  bool far = false, p2p_pair = false, split_i = false, split_j = false, nonempty = false;
  double4 gi = make_double4(0, 0, 0, 0), gj = gi;
  if (valid) {
    const int2 pr = in[t];
    i = pr.x;
    j = pr.y;
    gi = geo[i];
    gj = geo[j];
    ti = topo[i];
    tj = topo[j];
    const double d0 = gi.x - gj.x;
    const double d1 = gi.y - gj.y;
    const double d2 = gi.z - gj.z;
    double R2 = 0;
    R2 += d0 * d0;
    R2 += d1 * d1;
    R2 += d2 * d2;
    const double Ri = gi.w, Rj = gj.w;
    const bool leaf_i = ti.x < 0, leaf_j = tj.x < 0;
    nonempty = ti.y > 0 && tj.y > 0;
    if (R2 > (Ri + Rj) * (Ri + Rj)) {
      far = true;
    } else if (leaf_i && leaf_j) {
      p2p_pair = nonempty;
    } else if (leaf_j) {
      split_i = true;
    } else if (leaf_i) {
      split_j = true;
    } else if (Ri >= Rj) {
      split_i = true;
    } else {
      split_j = true;
    }
  }
  // Appends of the block to the four lists (0 next frontier, 8 entries per
  // pair; 1 executed M2L; 2 P2P; 3 all M2L, counted only) and the body-pair
  // sums: one global atomic per list per block, two block barriers.
  const bool split = split_i || split_j;
  const bool pred[4] = {split, far && nonempty, p2p_pair, far};
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t m[4];
#pragma unroll
  for (int q = 0; q < 4; q++) m[q] = __ballot(pred[q]);
  unsigned long long v = 0, self = 0;
  if (p2p_pair) {
    v = (unsigned long long)ti.y * (unsigned long long)tj.y;
    if (i == j) self = (unsigned long long)ti.y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, 64);
    self += __shfl_xor(self, o, 64);
  }
  __shared__ uint32_t wc[DTT_WAVES][4];
  __shared__ unsigned long long ws[DTT_WAVES][2];
  __shared__ uint32_t base[4];
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) wc[wid][q] = (uint32_t)__popcll(m[q]) * (q == 0 ? 8u : 1u);
    ws[wid][0] = v;
    ws[wid][1] = self;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < DTT_WAVES; w++) tot += wc[w][q];
    base[q] = tot ? atomicAdd(q == 0 ? nout : &cnt[q], tot) : 0u;
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
  uint32_t slot[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    uint32_t pre = base[q];
    for (int w = 0; w < wid; w++) pre += wc[w][q];
    slot[q] = pre + (uint32_t)__popcll(m[q] & lt) * (q == 0 ? 8u : 1u);
  }
  bool outside = false;
  if (far && nonempty && slot[1] < m2l_cap) {
    m2l[slot[1]] = make_int2(i, j);
    // the pair's M2L offset group (same arithmetic on the same centres as
    // a separate pass over the list would do)
    const double ci[3] = {gi.x, gi.y, gi.z}, cj[3] = {gj.x, gj.y, gj.z};
    const int32_t g = m2lg::pair_slot(depth[i], depth[j], width[i], width[j], ci, cj);
    m2l_slot[slot[1]] = g;
    if (g >= 0) used[g] = 1u;
    outside = g < 0;
  }
  {
    const uint64_t mo = __ballot(outside);
    if (lane == 0 && mo) atomicAdd(n_outside, (uint32_t)__popcll(mo));
  }
  if (p2p_pair && slot[2] < p2p_cap) p2p[slot[2]] = make_int2(i, j);
  if (split && slot[0] + 8 <= out_cap) {
    if (split_i) {
      const int32_t f = ti.x;
      for (int k = 0; k < 8; k++) out[slot[0] + k] = make_int2(f + k, j);
    } else {
      const int32_t f = tj.x;
      for (int k = 0; k < 8; k++) out[slot[0] + k] = make_int2(i, f + k);
    }
  }
}

__global__ void k_unpack_pairs(const uint64_t *keys, uint32_t n, int sbits, int32_t *src) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  src[t] = (int32_t)(keys[t] & ((1ull << sbits) - 1ull));
}

// off[t] = first entry whose target is >= t (t = 0..ncells)
__global__ void k_csr_offsets(const uint64_t *keys, uint32_t n, int sbits, int32_t ncells,
                              int32_t *off) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ncells) return;
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
__global__ void k_split_pairs(const int2 *pairs, uint32_t n, uint32_t *tgt, uint32_t *src) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int2 p = pairs[t];
  tgt[t] = (uint32_t)p.x;
  src[t] = (uint32_t)p.y;
}

// off[t] = first entry whose target is >= t (t = 0..ncells)
__global__ void k_csr_offsets32(const uint32_t *tgt, uint32_t n, int32_t ncells, int32_t *off) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ncells) return;
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

// one wave per target: its sources in ascending order
__global__ __launch_bounds__(64) void k_seg_sort(const int32_t *off, const uint32_t *in,
                                                 int32_t *out) {
  __shared__ uint32_t v[SEG_CAP];
  const int lane = threadIdx.x;
  const int32_t b = off[blockIdx.x], n = off[blockIdx.x + 1] - b;
  if (n <= 0) return;
  if (n <= SEG_CAP) {
    for (int32_t i = lane; i < n; i += 64) v[i] = in[b + i];
    __syncthreads();
    rank_sources(static_cast<const uint32_t *>(v), n, lane, out + b);
  } else {
    rank_sources(in + b, n, lane, out + b);  // long lists rank from global memory
  }
}

}  // namespace

hipError_t launch_split_pairs(const int2 *pairs, uint32_t n, uint32_t *tgt, uint32_t *src,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_split_pairs<<<(n + 255) / 256, 256, 0, s>>>(pairs, n, tgt, src);
  return hipGetLastError();
}

// off (ncells + 1) from the target-sorted list; out: each target's sources
// (in, in the target-sorted order) in ascending order
hipError_t launch_csr_sorted_sources(const uint32_t *tgt, const uint32_t *in, uint32_t n,
                                     int32_t ncells, int32_t *off, int32_t *out, hipStream_t s) {
  k_csr_offsets32<<<(ncells + 1 + 255) / 256, 256, 0, s>>>(tgt, n, ncells, off);
  if (n > 0 && ncells > 0) k_seg_sort<<<ncells, 64, 0, s>>>(off, in, out);
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
                           int32_t *m2l_slot, uint32_t *used, uint32_t *n_outside,
                           hipStream_t s) {
  if (nin_max == 0) return hipSuccess;
  k_dtt_step<<<(nin_max + DTT_THREADS - 1) / DTT_THREADS, DTT_THREADS, 0, s>>>(
      in, nin_dev, nin_max, out, out_cap, nout, m2l, m2l_cap, p2p, p2p_cap, cnt, geo, topo, depth,
      width, m2l_slot, used, n_outside);
  return hipGetLastError();
}

hipError_t launch_unpack_pairs(const uint64_t *keys, uint32_t n, int sbits, int32_t *src,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_unpack_pairs<<<(n + 255) / 256, 256, 0, s>>>(keys, n, sbits, src);
  return hipGetLastError();
}

hipError_t launch_csr_offsets(const uint64_t *sorted_keys, uint32_t n, int sbits,
                              int32_t ncells, int32_t *off, hipStream_t s) {
  k_csr_offsets<<<(ncells + 1 + 255) / 256, 256, 0, s>>>(sorted_keys, n, sbits, ncells, off);
  return hipGetLastError();
}

}  // namespace tfmm
