// tree.hip -- octree construction on the GPU: scalar-cube bounding box,
// Morton keys, body permutation and the level-synchronous cell build.
//
// Reference semantics (cpp/include/tapas/):
//   bounding box    hot/buildtree.h:72-87 (a Vec assigned from a scalar inside a
//                   per-dimension loop: min/max over every coordinate of every
//                   body except body 0's x and y) and vec.h:39-43
//   keys            hot/buildtree.h:504-539, sfc_morton.h:387-402
//   cells           hot/buildtree.h:388-459: a cell is a leaf iff nb <= ncrit;
//                   every internal cell gets all 8 children, empty ones too;
//                   region by halving (sfc_morton.h:536-557), centre (max+min)/2
//                   (hot.h:232-238)
// The cell set is built breadth first: level d+1 is produced from level d by
// one scan over the "split" flags, so cells of one level are contiguous and
// the 8 children of a cell are adjacent, which the level-synchronous upward
// and downward passes rely on.
#include <algorithm>

#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

constexpr int BB_THREADS = 256;

__global__ __launch_bounds__(BB_THREADS) void k_bbox_partial(const double *x, const double *y,
                                                             const double *z, int64_t n,
                                                             double *pmin, double *pmax) {
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * BB_THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BB_THREADS) {
    const double vz = z[i];
    lo = fmin(lo, vz);
    hi = fmax(hi, vz);
    if (i != 0) {
      const double vx = x[i], vy = y[i];
      lo = fmin(lo, fmin(vx, vy));
      hi = fmax(hi, fmax(vx, vy));
    }
  }
  __shared__ double slo[BB_THREADS], shi[BB_THREADS];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int o = BB_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + o]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmin[blockIdx.x] = slo[0];
    pmax[blockIdx.x] = shi[0];
  }
}

__global__ __launch_bounds__(BB_THREADS) void k_bbox_final(const double *pmin,
                                                           const double *pmax, int nparts,
                                                           double *region) {
  __shared__ double slo[BB_THREADS], shi[BB_THREADS];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < nparts; i += BB_THREADS) {
    lo = fmin(lo, pmin[i]);
    hi = fmax(hi, pmax[i]);
  }
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int o = BB_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + o]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    region[threadIdx.x] = slo[0];
    region[3 + threadIdx.x] = shi[0];
  }
}

// Problem of body i of a forest: the last p with off[p] <= i.
__device__ inline int problem_of(const int64_t *off, int nprob, int64_t i) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Keys against the region of the body's own problem (off == nullptr: one
// problem); the problem index is kept per body for the forest sort.
// Also packs each body into one 32-byte record (x, y, z, q): the permutation
// gathers in random order, and one record per body is one cache line where
// four arrays are four.
__global__ void k_keys(const double *x, const double *y, const double *z, const double *q,
                       int64_t n, const double *regions, const int64_t *off, int nprob,
                       uint64_t *keys, uint32_t *idx, uint32_t *bprob, double4 *rec) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pb = off ? problem_of(off, nprob, i) : 0;
  if (bprob) bprob[i] = (uint32_t)pb;
  const double *region = regions + 6 * pb;
  const int nfin = 1 << MAXDEPTH;
  const double p[3] = {x[i], y[i], z[i]};
  int a[3];
  for (int d = 0; d < 3; d++) {
    const double pitch = (region[3 + d] - region[d]) / (double)nfin;
    double o = p[d] - region[d];
    o = o / pitch;
    int v = (int)o;
    if (v == nfin) v--;
    a[d] = v;
  }
  keys[i] = finest_key(a[0], a[1], a[2]);
  idx[i] = (uint32_t)i;
  rec[i] = make_double4(p[0], p[1], p[2], q[i]);
}

// Bounding box of every problem of a forest, one block per problem, with
// the same scalar-cube rule as the single-problem box: problem p is its own
// Partition, so its first body's x and y are the ones left out.
__global__ __launch_bounds__(BB_THREADS) void k_bbox_forest(const double *x, const double *y,
                                                            const double *z, const int64_t *off,
                                                            double *regions) {
  const int pb = blockIdx.x;
  const int64_t b = off[pb], e = off[pb + 1];
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t i = b + threadIdx.x; i < e; i += BB_THREADS) {
    const double vz = z[i];
    lo = fmin(lo, vz);
    hi = fmax(hi, vz);
    if (i != b) {
      const double vx = x[i], vy = y[i];
      lo = fmin(lo, fmin(vx, vy));
      hi = fmax(hi, fmax(vx, vy));
    }
  }
  __shared__ double slo[BB_THREADS], shi[BB_THREADS];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int o = BB_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + o]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    regions[6 * pb + threadIdx.x] = slo[0];
    regions[6 * pb + 3 + threadIdx.x] = shi[0];
  }
}

// Forest order, step 1: after the bodies are sorted by key, the problem of
// each sorted body becomes the key of a second, stable sort.
__global__ void k_problem_keys(const uint32_t *order, const uint32_t *bprob, int64_t n,
                               uint32_t *pkey, uint32_t *pos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pkey[i] = bprob[order[i]];
  pos[i] = (uint32_t)i;
}

// Forest order, step 2: keys and body order by (problem, key).
__global__ void k_problem_gather(const uint32_t *pos, const uint64_t *keys1,
                                 const uint32_t *order1, int64_t n, uint64_t *keys,
                                 uint32_t *order) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = pos[i];
  keys[i] = keys1[j];
  order[i] = order1[j];
}

// Gather the caller's bodies into sorted order (SoA FP64) and keep the
// permutation.  perm[sorted] = original index.
__global__ void k_permute(const uint32_t *order, const double4 *rec, int64_t n, double *sx,
                          double *sy, double *sz, double *sq, int32_t *perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = order[i];
  const double4 r = rec[o];
  sx[i] = r.x;
  sy[i] = r.y;
  sz[i] = r.z;
  sq[i] = r.w;
  perm[i] = (int32_t)o;
}

// Root cells: one per problem (off == nullptr: the single root 0 of n bodies).
__global__ void k_roots(Cells c, int64_t n, const int64_t *off, int nprob,
                        const double *regions, int32_t *lev) {
  const int pb = blockIdx.x * blockDim.x + threadIdx.x;
  if (pb == 0 && lev) {
    lev[0] = 0;
    lev[1] = nprob;
  }
  if (pb >= nprob) return;
  const int64_t b = off ? off[pb] : 0, e = off ? off[pb + 1] : n;
  c.key[pb] = 0;
  c.beg[pb] = (int32_t)b;
  c.nb[pb] = (int32_t)(e - b);
  c.parent[pb] = -1;
  c.child[pb] = -1;
  c.depth[pb] = 0;
  const double *region = regions + 6 * pb;
  double w = 0;
  for (int d = 0; d < 3; d++) {
    c.center[3 * (int64_t)pb + d] = (region[3 + d] + region[d]) / 2;
    const double wd = region[3 + d] - region[d];
    w = (w < wd) ? wd : w;
  }
  c.width[pb] = w;
}

// Traversal start: the pair (root, root) of every problem.
// The traversal's first frontier: every root with itself, or with periodic
// images the one root with each of its 27 images (EX/include/traversal.h:
// 256-266).
__global__ void k_root_pairs(int2 *front, int nroots, uint32_t *count, int periodic) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = periodic ? 27 : nroots;
  if (i < n) front[i] = periodic ? make_int2(0, (int32_t)((uint32_t)i << IMG_SHIFT)) : make_int2(i, i);
  if (i == 0) *count = (uint32_t)n;
}

__global__ void k_split_flags(Cells c, int32_t lbeg, int32_t lcnt, int ncrit, uint32_t *flag) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= lcnt) return;
  const int32_t ci = lbeg + i;
  const bool split = c.nb[ci] > ncrit && c.depth[ci] < MAXDEPTH;
  flag[i] = split ? 1u : 0u;
  if (!split) c.child[ci] = -1;
}

__device__ inline int32_t lower_bound_keys(const uint64_t *k, int32_t lo, int32_t hi,
                                           uint64_t v) {
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (k[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// child ch of split parent p, at cell index first + ch
__device__ inline void make_child(Cells c, int32_t p, int ch, int32_t first, const uint64_t *keys,
                                  const double *regions, const uint32_t *sprob) {
  const int32_t ci = first + ch;
  const uint64_t pk = c.key[p];
  const uint64_t ck = key_child(pk, ch);
  const int32_t pb = c.beg[p], pe = c.beg[p] + c.nb[p];
  const int32_t b = lower_bound_keys(keys, pb, pe, ck);
  const int32_t e = (ch == 7) ? pe : lower_bound_keys(keys, pb, pe, key_child(pk, ch + 1));
  c.key[ci] = ck;
  c.beg[ci] = b;
  c.nb[ci] = e - b;
  c.parent[ci] = p;
  c.child[ci] = -1;
  c.depth[ci] = c.depth[p] + 1;
  // a split cell holds bodies, so its first body names its problem
  const double *region = regions + (sprob ? 6 * (int64_t)sprob[pb] : 0);
  double rmin[3] = {region[0], region[1], region[2]};
  double rmax[3] = {region[3], region[4], region[5]};
  double mn[3], mx[3];
  cell_region(ck, rmin, rmax, mn, mx);
  double w = 0;
  for (int d = 0; d < 3; d++) {
    c.center[3 * (int64_t)ci + d] = (mx[d] + mn[d]) / 2;
    const double wd = mx[d] - mn[d];
    w = (w < wd) ? wd : w;
  }
  c.width[ci] = w;
  if (ch == 0) c.child[p] = first;
}

__global__ void k_children(Cells c, int32_t lbeg, int32_t lcnt, const uint32_t *flag,
                           const uint32_t *slot, int32_t next_beg, const uint64_t *keys,
                           const double *regions, const uint32_t *sprob) {
  // one thread per (parent, child) -> 8 threads per parent
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)lcnt * 8) return;
  const int32_t i = (int32_t)(t >> 3);
  const int ch = (int)(t & 7);
  if (!flag[i]) return;
  make_child(c, lbeg + i, ch, next_beg + 8 * (int32_t)slot[i], keys, regions, sprob);
}

// ---------------------------------------------------------------------------
// The same level step with every size on the device (no host round trip per
// level): lev[2 d], lev[2 d + 1] = first cell and cell count of level d.  The
// grids come from the sizes of an earlier tree; the host checks afterwards
// that every level fitted its grid and the cell capacity (status bits), and
// builds again the synchronous way if not.
constexpr int LV_BLOCK = 256;

// split flags, the exclusive prefix of the flags inside each 256-cell block,
// and the block totals
__global__ __launch_bounds__(LV_BLOCK) void k_level_flags(Cells c, const int32_t *lev, int d,
                                                          int ncrit, uint32_t *flag,
                                                          uint32_t *loc, uint32_t *bsum) {
  const int32_t lbeg = lev[2 * d], lcnt = lev[2 * d + 1];
  const int32_t i = blockIdx.x * LV_BLOCK + threadIdx.x;
  bool split = false;
  if (i < lcnt) {
    const int32_t ci = lbeg + i;
    split = c.nb[ci] > ncrit && c.depth[ci] < MAXDEPTH;
    if (!split) c.child[ci] = -1;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t m = __ballot(split);
  __shared__ uint32_t wsum[LV_BLOCK / 64];
  if (lane == 0) wsum[wid] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int w = 0; w < LV_BLOCK / 64; w++) {
    pre += w < wid ? wsum[w] : 0u;
    tot += wsum[w];
  }
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (i < lcnt) {
    flag[i] = split ? 1u : 0u;
    loc[i] = pre + (uint32_t)__popcll(m & lt);
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// one block: exclusive scan of the block totals, the next level's extent;
// status bit 1: cell capacity exceeded, bit 2: level larger than its grid,
// bit 4: splits on the last planned level
__global__ __launch_bounds__(1024) void k_level_scan(int32_t *lev, int d, uint32_t *bsum,
                                                     int32_t nblk, int32_t cap,
                                                     uint32_t *status, int last) {
  const int32_t lbeg = lev[2 * d], lcnt = lev[2 * d + 1];
  const int32_t nb = min(nblk, (lcnt + LV_BLOCK - 1) / LV_BLOCK);
  __shared__ uint32_t part[1024];
  // each thread scans a contiguous run of the block totals
  const int per = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  uint32_t s = 0;
  for (int b = b0; b < b1; b++) s += bsum[b];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (int b = b0; b < b1; b++) {
    const uint32_t v = bsum[b];
    bsum[b] = run;
    run += v;
  }
  if (threadIdx.x == 0) {
    const uint32_t total = part[1023];
    const int64_t next_beg = (int64_t)lbeg + lcnt;
    int64_t next_cnt = 8 * (int64_t)total;
    uint32_t st = 0;
    if (lcnt > nblk * LV_BLOCK) st |= 2u;
    if (next_beg + next_cnt > cap) st |= 1u;
    if (last && next_cnt > 0) st |= 4u;
    if (st) {
      atomicOr(status, st);
      next_cnt = 0;  // stop: this tree is built again on the host's path
    }
    lev[2 * (d + 1)] = (int32_t)next_beg;
    lev[2 * (d + 1) + 1] = (int32_t)next_cnt;
  }
}

__global__ void k_level_children(Cells c, const int32_t *lev, int d, const uint32_t *flag,
                                 const uint32_t *loc, const uint32_t *bsum,
                                 const uint64_t *keys, const double *regions,
                                 const uint32_t *sprob) {
  const int32_t lbeg = lev[2 * d], lcnt = lev[2 * d + 1];
  const int32_t next_beg = lev[2 * (d + 1)], next_cnt = lev[2 * (d + 1) + 1];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (next_cnt == 0 || t >= (int64_t)lcnt * 8) return;
  const int32_t i = (int32_t)(t >> 3);
  const int ch = (int)(t & 7);
  if (!flag[i]) return;
  const uint32_t slot = bsum[i / LV_BLOCK] + loc[i];
  make_child(c, lbeg + i, ch, next_beg + 8 * (int32_t)slot, keys, regions, sprob);
}

// flag = 1 when a coordinate or charge of the n bodies is not finite (the
// host path rejects such bodies while unpacking them; device arrays are
// checked here)
__global__ void k_check_finite(const double *x, const double *y, const double *z,
                               const double *q, int64_t n, uint32_t *flag) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    bad |= !(isfinite(x[i]) && isfinite(y[i]) && isfinite(z[i]) && isfinite(q[i]));
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// Periodic images: the root cell is the periodic cell, a cube of side cycle
// centred on the bodies' scalar bounding cube (region: min x, y, z, max x, y,
// z); flag = 1 when the bodies span more than one period.
__global__ void k_periodic_region(double *region, double cycle, uint32_t *flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double lo = region[0], hi = region[3];
  if (!(hi - lo <= cycle)) *flag = 1u;
  const double c = 0.5 * (lo + hi), h = 0.5 * cycle;
  for (int d = 0; d < 3; d++) {
    region[d] = c - h;
    region[3 + d] = c + h;
  }
}

}  // namespace

hipError_t launch_periodic_region(double *region, double cycle, uint32_t *flag, hipStream_t s) {
  k_periodic_region<<<1, 64, 0, s>>>(region, cycle, flag);
  return hipGetLastError();
}

hipError_t launch_check_finite(const double *x, const double *y, const double *z,
                               const double *q, int64_t n, uint32_t *flag, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>(2048, (n + 255) / 256);
  k_check_finite<<<(unsigned)blocks, 256, 0, s>>>(x, y, z, q, n, flag);
  return hipGetLastError();
}

hipError_t launch_bbox(const double *x, const double *y, const double *z, int64_t n,
                       double *partial, int nparts, double *region, hipStream_t s) {
  k_bbox_partial<<<nparts, BB_THREADS, 0, s>>>(x, y, z, n, partial, partial + nparts);
  k_bbox_final<<<1, BB_THREADS, 0, s>>>(partial, partial + nparts, nparts, region);
  return hipGetLastError();
}

hipError_t launch_keys(const double *x, const double *y, const double *z, const double *q,
                       int64_t n, const double *regions, const int64_t *off, int nprob,
                       uint64_t *keys, uint32_t *idx, uint32_t *bprob, double4 *rec,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_keys<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, y, z, q, n, regions, off, nprob, keys,
                                                      idx, bprob, rec);
  return hipGetLastError();
}

hipError_t launch_bbox_forest(const double *x, const double *y, const double *z,
                              const int64_t *off, int nprob, double *regions, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  k_bbox_forest<<<nprob, BB_THREADS, 0, s>>>(x, y, z, off, regions);
  return hipGetLastError();
}

hipError_t launch_problem_keys(const uint32_t *order, const uint32_t *bprob, int64_t n,
                               uint32_t *pkey, uint32_t *pos, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_problem_keys<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(order, bprob, n, pkey, pos);
  return hipGetLastError();
}

hipError_t launch_problem_gather(const uint32_t *pos, const uint64_t *keys1,
                                 const uint32_t *order1, int64_t n, uint64_t *keys,
                                 uint32_t *order, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_problem_gather<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos, keys1, order1, n, keys,
                                                                order);
  return hipGetLastError();
}

hipError_t launch_root_pairs(int2 *front, int nroots, uint32_t *count, hipStream_t s,
                             bool periodic) {
  const int n = periodic ? 27 : nroots;
  k_root_pairs<<<(n + 255) / 256, 256, 0, s>>>(front, nroots, count, periodic ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_permute(const uint32_t *order, const double4 *rec, int64_t n, double *sx,
                          double *sy, double *sz, double *sq, int32_t *perm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_permute<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(order, rec, n, sx, sy, sz, sq, perm);
  return hipGetLastError();
}

hipError_t launch_roots(Cells c, int64_t n, const int64_t *off, int nprob,
                        const double *regions, int32_t *lev, hipStream_t s) {
  k_roots<<<(nprob + 255) / 256, 256, 0, s>>>(c, n, off, nprob, regions, lev);
  return hipGetLastError();
}

int32_t level_blocks(int32_t cells) { return (cells + LV_BLOCK - 1) / LV_BLOCK; }

hipError_t launch_level_step(Cells c, int32_t *lev, int d, int32_t grid_cells, int ncrit,
                             int32_t cap, uint32_t *flag, uint32_t *loc, uint32_t *bsum,
                             uint32_t *status, const uint64_t *keys, const double *regions,
                             const uint32_t *sprob, hipStream_t s, bool last) {
  const int32_t nblk = level_blocks(grid_cells);
  if (nblk <= 0) return hipSuccess;
  k_level_flags<<<nblk, LV_BLOCK, 0, s>>>(c, lev, d, ncrit, flag, loc, bsum);
  k_level_scan<<<1, 1024, 0, s>>>(lev, d, bsum, nblk, cap, status, last ? 1 : 0);
  const int64_t nt = (int64_t)nblk * LV_BLOCK * 8;
  k_level_children<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(c, lev, d, flag, loc, bsum, keys,
                                                                 regions, sprob);
  return hipGetLastError();
}

hipError_t launch_split_flags(Cells c, int32_t lbeg, int32_t lcnt, int ncrit, uint32_t *flag,
                              hipStream_t s) {
  if (lcnt <= 0) return hipSuccess;
  k_split_flags<<<(lcnt + 255) / 256, 256, 0, s>>>(c, lbeg, lcnt, ncrit, flag);
  return hipGetLastError();
}

hipError_t launch_children(Cells c, int32_t lbeg, int32_t lcnt, const uint32_t *flag,
                           const uint32_t *slot, int32_t next_beg, const uint64_t *keys,
                           const double *regions, const uint32_t *sprob, hipStream_t s) {
  if (lcnt <= 0) return hipSuccess;
  const int64_t nt = (int64_t)lcnt * 8;
  k_children<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(c, lbeg, lcnt, flag, slot, next_beg,
                                                           keys, regions, sprob);
  return hipGetLastError();
}

}  // namespace tfmm
