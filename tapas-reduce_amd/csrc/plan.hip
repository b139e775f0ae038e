// plan.hip -- the device-side sizes of an evaluation recorded as one hipGraph.
//
// In that mode every launch of the evaluation is sized by a plan (capacities
// taken from an earlier evaluation of the context, with headroom) and reads
// the actual sizes of this evaluation from device memory: the tree's extent
// from the level table, list lengths from the traversal's counters, the
// group, item and unit counts from the words written here.  No size is read
// back by the host before the end.  Every kernel stays inside its
// capacities whatever the sizes are; the last kernel of the graph records
// which capacity an evaluation exceeded (status bits), and the host then
// evaluates again the ordinary way and makes a new plan.
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

// The tree's extent after the level steps (lev: first cell and count of each
// of the L planned levels, then the level after them).  A tree that did not
// fit its plan (status bits of the level steps) is replaced by empty roots:
// every cell becomes an empty leaf (k_pad_cells with ncells = 0), so the
// rest of the graph has nothing to do, and the host evaluates again; the
// level counts are cleared too, so the per-level kernels find no parents.
__global__ void k_tree_final(int32_t *lev, int L, const uint32_t *lev_status, int32_t tile,
                             uint32_t *sz) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t st = *lev_status;
  if (lev[2 * L + 1] != 0) st |= ST_DEPTH;
  int64_t ncells = 0, nleaves = 0;
  int depth = -1;
  for (int d = 0; d < L && lev[2 * d + 1] > 0; d++) {
    ncells = (int64_t)lev[2 * d] + lev[2 * d + 1];
    nleaves += lev[2 * d + 1] - lev[2 * d + 3] / 8;
    depth = d;
  }
  if (st) {
    sz[SZ_STATUS] |= st;
    ncells = nleaves = 0;
    for (int d = 0; d <= L; d++) lev[2 * d + 1] = 0;
  }
  sz[SZ_NCELLS] = (uint32_t)ncells;
  sz[SZ_NLEAVES] = (uint32_t)nleaves;
  sz[SZ_DEPTH] = (uint32_t)depth;
  sz[SZ_NTILES] = (uint32_t)((ncells + tile - 1) / tile);
}

// cells [ncells, cap): empty leaves (no bodies, no children) at the origin
__global__ void k_pad_cells(Cells c, const uint32_t *sz, int32_t cap) {
  const int32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= cap || ci < (int32_t)sz[SZ_NCELLS]) return;
  c.key[ci] = 0;
  c.beg[ci] = 0;
  c.nb[ci] = 0;
  c.parent[ci] = -1;
  c.child[ci] = -1;
  c.depth[ci] = 0;
  c.center[3 * ci] = c.center[3 * ci + 1] = c.center[3 * ci + 2] = 0;
  c.width[ci] = 0;
}

// groups in use: the exclusive scan of the slot flags plus the last flag
__global__ void k_groups_count(const uint32_t *gid, const uint32_t *used, int32_t nslots,
                               uint32_t gcap, uint32_t *sz) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t ng = gid[nslots - 1] + used[nslots - 1];
  sz[SZ_NG] = ng;
  if (ng > gcap) sz[SZ_STATUS] |= ST_GROUPS;
}

// The capacities of the rest of the evaluation against what it produced.
__global__ void k_plan_check(const uint32_t *cnt, int nsteps, const uint32_t *step_grid,
                             PlanCaps caps, const uint32_t *gcnt, const uint32_t *ucnt,
                             const uint32_t *uoff, int32_t ntiles, uint32_t *sz) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t st = 0;
  // traversal: lists, frontiers (each step within its grid), termination
  if (cnt[1] > caps.m2l || cnt[2] > caps.p2p || cnt[16 + nsteps] != 0) st |= ST_TRAVERSAL;
  for (int i = 0; i < nsteps; i++) {
    const uint32_t f = cnt[16 + i];
    if (f > caps.front || f > step_grid[i]) st |= ST_TRAVERSAL;
  }
  if (cnt[16 + nsteps] > caps.front) st |= ST_TRAVERSAL;
  if (gcnt[1] > caps.fallback) st |= ST_FALLBACK;
  if (sz[SZ_NITEMS] > caps.items) st |= ST_ITEMS;
  const uint32_t units = ntiles > 0 ? uoff[ntiles - 1] + ucnt[ntiles - 1] : 0u;
  sz[SZ_NUNITS] = units;
  if (units > caps.units) st |= ST_UNITS;
  sz[SZ_STATUS] |= st;
}

}  // namespace

hipError_t launch_tree_final(int32_t *lev, int L, const uint32_t *lev_status, int32_t tile,
                             uint32_t *sz, hipStream_t s) {
  k_tree_final<<<1, 64, 0, s>>>(lev, L, lev_status, tile, sz);
  return hipGetLastError();
}

hipError_t launch_pad_cells(Cells c, const uint32_t *sz, int32_t cap, hipStream_t s) {
  if (cap <= 0) return hipSuccess;
  k_pad_cells<<<(cap + 255) / 256, 256, 0, s>>>(c, sz, cap);
  return hipGetLastError();
}

hipError_t launch_groups_count(const uint32_t *gid, const uint32_t *used, int32_t nslots,
                               uint32_t gcap, uint32_t *sz, hipStream_t s) {
  k_groups_count<<<1, 64, 0, s>>>(gid, used, nslots, gcap, sz);
  return hipGetLastError();
}

hipError_t launch_plan_check(const uint32_t *cnt, int nsteps, const uint32_t *step_grid,
                             const PlanCaps &caps, const uint32_t *gcnt, const uint32_t *ucnt,
                             const uint32_t *uoff, int32_t ntiles, uint32_t *sz, hipStream_t s) {
  k_plan_check<<<1, 64, 0, s>>>(cnt, nsteps, step_grid, caps, gcnt, ucnt, uoff, ntiles, sz);
  return hipGetLastError();
}

}  // namespace tfmm
