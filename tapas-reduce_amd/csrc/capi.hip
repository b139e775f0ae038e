// capi.hip -- the C ABI (include/tfmm.h): context, device buffers and the
// phase orchestration of one FMM evaluation on one GPU.
//
// Phases follow the reference driver (sample/exafmm-dev-13274dd4ac68/examples/
// tapas_exafmm.cxx:557-641): Partition (tree), Map(FMM_Upward),
// Map(FMM_DTT, Product(root, root)), Map(FMM_Downward), CopyBackResult.
// All work is queued on the context's stream; the host only waits where a
// size decides the next launch (cells per level, frontier per traversal step).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tfmm.h"
#include "tfmm_device.h"
#include "tfmm_internal.h"
#include "tfmm_kernels.h"

using namespace tfmm;

namespace {

enum Ev {
  EV_BEGIN,
  EV_TREE,
  EV_UP,
  EV_DTT0,
  EV_DTT1,
  EV_M2L0,
  EV_M2LT,
  EV_M2L1,
  EV_P2PREADY,
  EV_P2P0,
  EV_P2P1,
  EV_P2PJOIN,  // the P2P stream's end, for the join only (never timed)
  EV_TRAV,
  EV_DOWN,
  EV_COUNT
};

// Capacities of an evaluation recorded as one graph (TFMM_TUNE_GRAPH = 2),
// taken from an earlier evaluation of the context with headroom, and the
// executable graph recorded with them.
struct Plan {
  bool valid = false;
  // what the plan was made for
  int64_t n = 0;
  int nprob = 0;
  tfmm_params prm{};
  double unit_factor = 0;
  int xcd_run = 0, p2p_schedule = 0, p2p_waves = 0;
  // capacities
  int L = 0;                        // tree levels
  std::vector<int32_t> grid;        // cells per level
  int64_t cells = 0;                // cells in all
  int max_depth = 0, nsteps = 0, cwords = 0;
  int32_t nslots = 0, ntiles = 0;
  std::vector<uint32_t> step_grid;  // pairs per traversal step
  PlanCaps caps{};
  uint32_t groups = 0;
  int gbits = 0, sbits = 0, tbits = 0;
  // sort parities (fixed by the bit ranges)
  bool keys_alt = false, gkeys_alt = false;
  // the graph and the buffer addresses it was recorded with
  hipGraphExec_t exec = nullptr;
  uint64_t signature = 0;
};

}  // namespace

struct tfmm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // P2P runs here, overlapping the M2L work
  std::string err;

  // bodies (caller order); a batch of nprob problems is one forest, problem p
  // holding bodies [prob_off[p], prob_off[p+1])
  int64_t n = 0;
  int nprob = 1;
  std::vector<int64_t> prob_off;
  DevBuf d_prob_off, bprob, pkey, pkey_alt, ppos, ppos_alt;
  const double *bx = nullptr, *by = nullptr, *bz = nullptr, *bq = nullptr;
  DevBuf own_x, own_y, own_z, own_q;

  tfmm_params prm{};
  int stage = 0;  // 0 none, 1 tree, 2 upward, 3 traverse, 4 downward

  // tree
  DevBuf keys, keys_alt, idx, idx_alt, sort_tmp, sx, sy, sz, sq, perm, region, bbox_part, brec;
  DevBuf c_key, c_beg, c_nb, c_parent, c_child, c_depth, c_center, c_width;
  size_t cell_cap = 0;
  DevBuf flag, slot, scan_tmp, lev, lev_bsum, lev_status;
  std::vector<int32_t> lvl;
  std::vector<int32_t> tree_hint;  // cells per level of the last tree
  int64_t ncells = 0, nleaves = 0;
  bool keys_in_alt = false;
  double h_region[6] = {0, 0, 0, 0, 0, 0};
  bool tops_ready = false;  // unit translation operators built

  // expansions
  DevBuf M, L;

  // traversal
  DevBuf front_a, front_b, m2l_pairs, p2p_pairs, cnt, lkeys, lkeys_alt, lidx, lidx_alt;
  DevBuf dtt_geo, dtt_topo;
  DevBuf m2l_off, m2l_src, p2p_off, p2p_src;
  DevBuf seg_long;  // queue of the per-target lists too long for one wave's sort
  uint32_t m2l_cap = 0, p2p_cap = 0, front_cap = 0;
  std::vector<uint32_t> dtt_hint;  // frontier sizes of the last traversal
  int64_t dtt_hint_cells = -1;
  uint32_t n_m2l_exec = 0, n_p2p = 0;
  uint64_t n_m2l_all = 0, n_p2p_calls = 0, n_p2p_self = 0;
  uint64_t n_p2p_calls_leaf = 0;  // P2P leaf-pair calls (a mutual pair once)
  uint32_t dtt_steps = 0;

  // batched M2L (offset groups)
  DevBuf g_used, g_gid, g_y64, g_y32, g_af, g_dl, g_lt, g_cnt;
  DevBuf gkeys, gvals, gkeys_alt, gvals_alt, fkeys, fvals, fkeys_alt, fvals_alt;
  // it_flag: position in segment per grouped pair; it_iflag/it_iexcl: item starts per block
  DevBuf it_flag, it_iflag, it_iexcl, items, cols, tile_off, mq, codes;
  DevBuf fb_off, fb_src, fb_part, relpos, leaf_of, tops, slot_of;
  DevBuf u_cnt, u_off, units, u_part;
  double m2l_unit_factor = 0.0;  // tile work-unit size in mean items per tile (0: off)
  int m2l_xcd_run = 32;          // consecutive work units dealt to one XCD
  int p2p_schedule = 0;          // where the P2P starts (TFMM_TUNE_P2P_SCHEDULE)
  int use_graph = 0;             // tfmm_evaluate's numeric phase as one hipGraph
  int p2p_waves = 0;             // resident P2P waves per SIMD (0: one workgroup per cell)
  int n_cu = 0;                  // compute units of the device
  DevBuf p2p_next;               // the resident P2P's work counter
  int ext_events = -1;           // external event records in a capture: -1 not probed
  bool fused = false;            // last evaluation ran the fused order
  bool capturing = false;        // inside the numeric phase's stream capture
  bool p2p_forked = false;       // this evaluation's P2P already started in phase A
  bool inner_timed = true;       // the last evaluation recorded its per-phase events
  hipGraphExec_t gexec = nullptr;
  Plan plan;                      // the whole-evaluation graph (use_graph == 2)
  DevBuf gsz, gstep;              // its device sizes and traversal step grids
  uint32_t *h_plan = nullptr;     // pinned read-back of the whole graph (PLAN_HOST_BYTES)
  bool one_graph = false;         // the last tfmm_evaluate was one launch of it
  // periodic images (tfmm_set_periodic): image sublevels, 0 = none, and the
  // period; the root cell is then the periodic cell
  int images = 0;
  double cycle = 0;
  DevBuf pflag;  // bodies spanning more than one period
  // expansion basis (tfmm_set_basis): TFMM_BASIS_SPHERICAL or _CARTESIAN,
  // and the Cartesian M2L's schedule tables (built once per context)
  int basis = TFMM_BASIS_SPHERICAL;
  DevBuf cart_tab;
  bool cart_ready = false;
  // Cartesian expansions (CART_TERMS doubles per cell); M and L then hold
  // the spherical images the tile kernel works on (TFMM_TUNE_CART_M2L 0)
  DevBuf Mc, Lc;
  int cart_direct = 0;  // TFMM_TUNE_CART_M2L: 1 = the direct Cartesian M2L kernel
  int32_t unit_chunk = 0;
  uint32_t units_bound = 0;
  bool codes_ready = false;
  uint32_t n_groups = 0, n_grouped = 0, n_fallback = 0, n_items = 0;
  bool grouped_keys_in_alt = false;

  // picker
  DevBuf feat_acc, feat_x, feat_out;

  // fields
  DevBuf acc, trg, out;

  // this device's copy of the static scratch of tfmm_direct_noalloc
  void *noalloc = nullptr;
  DevBuf finite_flag;  // the finiteness check of device-resident bodies

  // host staging (pinned)
  uint32_t *h_u32 = nullptr;  // pinned staging, H_U32_WORDS words
  unsigned long long *h_u64 = nullptr;

  hipEvent_t ev[EV_COUNT] = {};
  float ms_dtt = 0, ms_m2l = 0, ms_p2p = 0;
  bool timed_phases = false;
};

namespace {

int fail(tfmm_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(tfmm_ctx *c, hipError_t e, const char *where) {
  return fail(c, e == hipErrorOutOfMemory ? TFMM_ENOMEM : TFMM_EHIP,
              std::string(where) + ": " + hipGetErrorString(e));
}

#define TRY(expr)                                     \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hipfail(c, _e, #expr); \
  } while (0)

hipError_t grow_copy(DevBuf &b, size_t bytes, size_t used, hipStream_t s) {
  if (bytes <= b.cap) return hipSuccess;
  size_t want = b.cap ? b.cap : 256;
  while (want < bytes) want *= 2;
  if (!dev_fits(0, want)) return hipErrorOutOfMemory;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) return e;
  dev_account(0, want);
  if (b.ptr && used) {
    e = hipMemcpyAsync(p, b.ptr, used, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      if (hipFree(p) == hipSuccess) dev_account(want, 0);
      return e;
    }
  }
  if (b.ptr && hipFree(b.ptr) == hipSuccess) dev_account(b.cap, 0);
  b.ptr = p;
  b.cap = want;
  return hipSuccess;
}

Cells cells_of(tfmm_ctx *c) {
  Cells k;
  k.key = dp<uint64_t>(c->c_key);
  k.beg = dp<int32_t>(c->c_beg);
  k.nb = dp<int32_t>(c->c_nb);
  k.parent = dp<int32_t>(c->c_parent);
  k.child = dp<int32_t>(c->c_child);
  k.depth = dp<int32_t>(c->c_depth);
  k.center = dp<double>(c->c_center);
  k.width = dp<double>(c->c_width);
  return k;
}

Bodies sorted_bodies(tfmm_ctx *c) {
  Bodies b;
  b.x = dp<double>(c->sx);
  b.y = dp<double>(c->sy);
  b.z = dp<double>(c->sz);
  b.q = dp<double>(c->sq);
  b.n = c->n;
  return b;
}

// every device buffer of the context
std::vector<DevBuf *> all_buffers(tfmm_ctx *c) {
  return {&c->own_x, &c->own_y, &c->own_z, &c->own_q, &c->keys, &c->keys_alt, &c->idx,
          &c->idx_alt, &c->sort_tmp, &c->sx, &c->sy, &c->sz, &c->sq, &c->perm,
          &c->region, &c->bbox_part, &c->c_key, &c->c_beg, &c->c_nb, &c->c_parent,
          &c->c_child, &c->c_depth, &c->c_center, &c->c_width, &c->flag, &c->slot,
          &c->scan_tmp, &c->M, &c->L, &c->front_a, &c->front_b, &c->m2l_pairs,
          &c->p2p_pairs, &c->cnt, &c->lkeys, &c->lkeys_alt, &c->lidx, &c->lidx_alt,
          &c->m2l_off, &c->m2l_src, &c->p2p_off, &c->p2p_src, &c->acc,
          &c->trg, &c->out, &c->g_used, &c->g_gid, &c->g_y64, &c->g_y32, &c->g_af, &c->g_dl,
          &c->g_lt, &c->g_cnt, &c->gkeys, &c->gvals, &c->gkeys_alt, &c->gvals_alt,
          &c->fkeys, &c->fvals, &c->fkeys_alt, &c->fvals_alt, &c->it_flag,
          &c->it_iflag, &c->it_iexcl, &c->items,
          &c->tile_off, &c->cols, &c->mq, &c->codes, &c->fb_off, &c->fb_src, &c->fb_part,
          &c->relpos, &c->leaf_of, &c->tops, &c->slot_of,
          &c->feat_acc, &c->feat_x, &c->feat_out, &c->d_prob_off, &c->bprob,
          &c->pkey, &c->pkey_alt, &c->ppos, &c->ppos_alt, &c->brec, &c->lev,
          &c->lev_bsum, &c->lev_status, &c->dtt_geo, &c->dtt_topo, &c->u_cnt,
          &c->u_off, &c->units, &c->u_part, &c->p2p_next, &c->finite_flag, &c->gsz, &c->gstep,
          &c->pflag, &c->cart_tab, &c->Mc, &c->Lc, &c->seg_long};
}

hipError_t reserve_cells(tfmm_ctx *c, size_t need, size_t used, hipStream_t s) {
  if (need <= c->cell_cap) return hipSuccess;
  size_t cap = c->cell_cap ? c->cell_cap : 4096;
  while (cap < need) cap *= 2;
  hipError_t e;
  if ((e = grow_copy(c->c_key, cap * 8, used * 8, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_beg, cap * 4, used * 4, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_nb, cap * 4, used * 4, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_parent, cap * 4, used * 4, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_child, cap * 4, used * 4, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_depth, cap * 4, used * 4, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_center, cap * 24, used * 24, s)) != hipSuccess) return e;
  if ((e = grow_copy(c->c_width, cap * 8, used * 8, s)) != hipSuccess) return e;
  c->cell_cap = cap;
  return hipSuccess;
}

constexpr int H_U32_WORDS = 256;

// the period the kernels displace images by (0 without periodic images)
double period(const tfmm_ctx *c) { return c->images > 0 ? c->cycle : 0.0; }

bool cartesian(const tfmm_ctx *c) { return c->basis == TFMM_BASIS_CARTESIAN; }
// the Cartesian basis with its own M2L kernel (no spherical images)
bool cart_direct(const tfmm_ctx *c) { return cartesian(c) && c->cart_direct; }
// the tile kernel in FP32 (mixed precision, spherical basis); the Cartesian
// basis keeps its expansions and their M2L in FP64
bool m2l_fp32(const tfmm_ctx *c) { return c->prm.precision == TFMM_MIXED && !cartesian(c); }

// bytes of a spherical expansion array (M or L) of every cell
size_t sph_bytes(int64_t cells) { return (size_t)cells * NTERM * 16; }
// bytes of a Cartesian expansion array (Mc or Lc) of every cell
size_t cart_bytes(int64_t cells) { return (size_t)cells * CART_TERMS * 8; }

// the expansion arrays of the basis and route
hipError_t reserve_expansions(tfmm_ctx *c) {
  hipError_t e;
  if (cartesian(c)) {
    if ((e = dev_reserve(c->Mc, cart_bytes(c->ncells))) != hipSuccess) return e;
    if ((e = dev_reserve(c->Lc, cart_bytes(c->ncells))) != hipSuccess) return e;
    if (c->cart_direct) return hipSuccess;
  }
  if ((e = dev_reserve(c->M, sph_bytes(c->ncells))) != hipSuccess) return e;
  return dev_reserve(c->L, sph_bytes(c->ncells));
}

// the Cartesian M2L's tables, on the device once per context
int cart_tables(tfmm_ctx *c) {
  if (c->cart_ready) return TFMM_OK;
  TRY(dev_reserve(c->cart_tab, cart_tables_bytes()));
  hipError_t e = cart_upload_tables(c->cart_tab.ptr);
  if (e == hipErrorInvalidValue) return fail(c, TFMM_EHIP, "Cartesian M2L schedule is inconsistent");
  TRY(e);
  c->cart_ready = true;
  return TFMM_OK;
}

// the direct sum's image range for the image sublevels (the reference's
// prange, tapas_exafmm.cxx:314-317)
int image_range(const tfmm_ctx *c) {
  int r = 0, p3 = 1;
  for (int i = 0; i < c->images; i++, p3 *= 3) r += p3;
  return r;
}
// source field of the fallback M2L keys (target << FB_SRC_BITS | source word):
// the whole 32-bit source word, cell and periodic image
constexpr int FB_SRC_BITS = 32;

int bits_for(int64_t v) {
  int b = 1;
  while (b < 62 && (((int64_t)1 << b) <= v)) b++;
  return b;
}

int check_params(tfmm_ctx *c, const tfmm_params *p) {
  if (!p) return fail(c, TFMM_EINVAL, "params is null");
  if (p->ncrit < 1) return fail(c, TFMM_EINVAL, "ncrit must be >= 1");
  if (!(p->theta > 0) || !(p->theta < 1e300))
    return fail(c, TFMM_EINVAL, "theta must be a positive finite number");
  if (p->mutual != 0 && p->mutual != 1) return fail(c, TFMM_EINVAL, "mutual must be 0 or 1");
  if (p->precision != TFMM_FP64 && p->precision != TFMM_MIXED)
    return fail(c, TFMM_EINVAL, "precision must be TFMM_FP64 or TFMM_MIXED");
  return TFMM_OK;
}

// The cell build with all level sizes on the device, grids sized from the
// last tree's levels (with headroom), one host round trip at the end.
// Returns 0 when the tree is complete, 1 when it must be built the
// synchronous way (a level outgrew its grid, the cell capacity or the
// hinted depth), or an error.
int build_levels_hinted(tfmm_ctx *c, int64_t n, const int64_t *off, int np,
                        const uint64_t *skeys, const uint32_t *sprob) {
  hipStream_t s = c->stream;
  const int L = (int)c->tree_hint.size();
  if (2 * (L + 1) + 1 > H_U32_WORDS) return 1;
  std::vector<int32_t> grid(L);
  int64_t total = 0, gmax = 0;
  for (int d = 0; d < L; d++) {
    const int64_t h = c->tree_hint[d];
    grid[d] = (int32_t)std::min<int64_t>(INT32_MAX / 16, h + h / 8 + 256);
    if (d == 0) grid[d] = std::max<int32_t>(grid[d], np);
    total += grid[d];
    gmax = std::max<int64_t>(gmax, grid[d]);
  }
  if (total + 4096 > INT32_MAX / 2) return 1;
  TRY(reserve_cells(c, (size_t)total + 4096, 0, s));
  const int32_t cap = (int32_t)std::min<size_t>(c->cell_cap, INT32_MAX / 2);
  TRY(dev_reserve(c->lev, (size_t)2 * (L + 1) * 4));
  TRY(dev_reserve(c->lev_status, 4));
  TRY(dev_reserve(c->flag, (size_t)gmax * 4 + 1024));
  TRY(dev_reserve(c->slot, (size_t)gmax * 4 + 1024));
  TRY(dev_reserve(c->lev_bsum, (size_t)level_blocks((int32_t)gmax) * 4 + 4));
  TRY(hipMemsetAsync(c->lev.ptr, 0, (size_t)2 * (L + 1) * 4, s));
  TRY(hipMemsetAsync(c->lev_status.ptr, 0, 4, s));
  const Cells k = cells_of(c);
  int32_t *lev = dp<int32_t>(c->lev);
  TRY(launch_roots(k, n, off, np, dp<double>(c->region), lev, s));
  for (int d = 0; d < L; d++)
    TRY(launch_level_step(k, lev, d, grid[d], c->prm.ncrit, cap, dp<uint32_t>(c->flag),
                          dp<uint32_t>(c->slot), dp<uint32_t>(c->lev_bsum),
                          dp<uint32_t>(c->lev_status), skeys, dp<double>(c->region), sprob, s));
  TRY(hipMemcpyAsync(c->h_u32, lev, (size_t)2 * (L + 1) * 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(c->h_u32 + 2 * (L + 1), c->lev_status.ptr, 4, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  const int32_t *hl = reinterpret_cast<const int32_t *>(c->h_u32);
  if (c->h_u32[2 * (L + 1)] != 0 || hl[2 * L + 1] != 0) return 1;
  c->lvl.assign(1, 0);
  int64_t nleaves = 0;
  for (int d = 0; d < L && hl[2 * d + 1] > 0; d++) {
    c->lvl.push_back(hl[2 * d] + hl[2 * d + 1]);
    nleaves += hl[2 * d + 1] - hl[2 * (d + 1) + 1] / 8;
  }
  c->ncells = c->lvl.back();
  c->nleaves = nleaves;
  c->tree_hint.clear();
  for (size_t d = 0; d + 1 < c->lvl.size(); d++) c->tree_hint.push_back(c->lvl[d + 1] - c->lvl[d]);
  return 0;
}

int do_build_tree(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const int64_t n = c->n;
  TRY(dev_reserve(c->keys, n * 8));
  TRY(dev_reserve(c->keys_alt, n * 8));
  TRY(dev_reserve(c->idx, n * 4));
  TRY(dev_reserve(c->idx_alt, n * 4));
  TRY(dev_reserve(c->sort_tmp, radix_tmp_bytes(n)));
  TRY(dev_reserve(c->sx, n * 8));
  TRY(dev_reserve(c->sy, n * 8));
  TRY(dev_reserve(c->sz, n * 8));
  TRY(dev_reserve(c->sq, n * 8));
  TRY(dev_reserve(c->perm, n * 4));
  const int np = c->nprob;
  const bool forest = np > 1;
  TRY(dev_reserve(c->region, 6 * 8 * (size_t)np));
  const int64_t *off = forest ? dp<int64_t>(c->d_prob_off) : nullptr;
  if (forest) {
    TRY(dev_reserve(c->bprob, n * 4));
    TRY(launch_bbox_forest(c->bx, c->by, c->bz, off, np, dp<double>(c->region), s));
  } else {
    const int nparts = (int)std::min<int64_t>(1024, (n + 255) / 256);
    TRY(dev_reserve(c->bbox_part, 2 * 8 * (size_t)nparts));
    TRY(launch_bbox(c->bx, c->by, c->bz, n, dp<double>(c->bbox_part), nparts,
                    dp<double>(c->region), s));
    if (c->images > 0) {
      TRY(dev_reserve(c->pflag, 4));
      TRY(launch_fill_zero(c->pflag.ptr, 4, s));
      TRY(launch_periodic_region(dp<double>(c->region), c->cycle, dp<uint32_t>(c->pflag), s));
    }
  }
  TRY(dev_reserve(c->brec, n * 32));
  TRY(launch_keys(c->bx, c->by, c->bz, c->bq, n, dp<double>(c->region), off, np,
                  dp<uint64_t>(c->keys), dp<uint32_t>(c->idx),
                  forest ? dp<uint32_t>(c->bprob) : nullptr, dp<double4>(c->brec), s));
  bool alt = false;
  TRY(radix_sort_pairs(dp<uint64_t>(c->keys), dp<uint32_t>(c->idx), dp<uint64_t>(c->keys_alt),
                       dp<uint32_t>(c->idx_alt), n, DEPTHBITS, DEPTHBITS + 3 * MAXDEPTH,
                       c->sort_tmp.ptr, s, &alt));
  const uint32_t *sprob = nullptr;
  if (forest) {
    // (problem, key) order: a stable sort by problem of the key-sorted bodies
    TRY(dev_reserve(c->pkey, n * 4));
    TRY(dev_reserve(c->pkey_alt, n * 4));
    TRY(dev_reserve(c->ppos, n * 4));
    TRY(dev_reserve(c->ppos_alt, n * 4));
    const uint64_t *k1 = dp<uint64_t>(alt ? c->keys_alt : c->keys);
    const uint32_t *o1 = dp<uint32_t>(alt ? c->idx_alt : c->idx);
    TRY(launch_problem_keys(o1, dp<uint32_t>(c->bprob), n, dp<uint32_t>(c->pkey),
                            dp<uint32_t>(c->ppos), s));
    bool palt = false;
    TRY(radix_sort_pairs32(dp<uint32_t>(c->pkey), dp<uint32_t>(c->ppos), dp<uint32_t>(c->pkey_alt),
                           dp<uint32_t>(c->ppos_alt), n, 0, bits_for(np - 1), c->sort_tmp.ptr, s,
                           &palt));
    sprob = dp<uint32_t>(palt ? c->pkey_alt : c->pkey);
    TRY(launch_problem_gather(dp<uint32_t>(palt ? c->ppos_alt : c->ppos), k1, o1, n,
                              dp<uint64_t>(alt ? c->keys : c->keys_alt),
                              dp<uint32_t>(alt ? c->idx : c->idx_alt), s));
    alt = !alt;
  }
  c->keys_in_alt = alt;
  const uint64_t *skeys = dp<uint64_t>(alt ? c->keys_alt : c->keys);
  const uint32_t *order = dp<uint32_t>(alt ? c->idx_alt : c->idx);
  TRY(launch_permute(order, dp<double4>(c->brec), n, dp<double>(c->sx), dp<double>(c->sy),
                     dp<double>(c->sz), dp<double>(c->sq), dp<int32_t>(c->perm), s));

  // level-synchronous cell build: without a host round trip per level when
  // an earlier tree gives the level sizes to size the grids by
  if (!c->tree_hint.empty()) {
    int rc = build_levels_hinted(c, n, off, np, skeys, sprob);
    if (rc < 0) return rc;
    if (rc == 0) {
      TRY(hipMemcpyAsync(c->h_region, c->region.ptr, 48, hipMemcpyDeviceToHost, s));
      return TFMM_OK;
    }
  }
  TRY(reserve_cells(c, std::max<size_t>(4096, 8 * (size_t)np), 0, s));
  TRY(launch_roots(cells_of(c), n, off, np, dp<double>(c->region), nullptr, s));
  c->lvl.assign(1, 0);
  int32_t lbeg = 0, lcnt = np;
  int64_t nleaves = 0;
  while (true) {
    TRY(dev_reserve(c->flag, (size_t)lcnt * 4));
    TRY(dev_reserve(c->slot, (size_t)lcnt * 4));
    TRY(dev_reserve(c->scan_tmp, scan_tmp_bytes(lcnt)));
    TRY(launch_split_flags(cells_of(c), lbeg, lcnt, c->prm.ncrit, dp<uint32_t>(c->flag), s));
    TRY(scan_exclusive_u32(dp<uint32_t>(c->flag), dp<uint32_t>(c->slot), lcnt,
                           c->scan_tmp.ptr, s));
    TRY(hipMemcpyAsync(c->h_u32, dp<uint32_t>(c->flag) + (lcnt - 1), 4, hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(c->h_u32 + 1, dp<uint32_t>(c->slot) + (lcnt - 1), 4,
                       hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    const int64_t splits = (int64_t)c->h_u32[0] + c->h_u32[1];
    nleaves += lcnt - splits;
    c->lvl.push_back(lbeg + lcnt);
    if (splits == 0) break;
    const int64_t next_beg = (int64_t)lbeg + lcnt;
    const int64_t need = next_beg + 8 * splits;
    if (need > INT32_MAX) return fail(c, TFMM_EINVAL, "too many cells");
    TRY(reserve_cells(c, (size_t)need, (size_t)next_beg, s));
    TRY(launch_children(cells_of(c), lbeg, lcnt, dp<uint32_t>(c->flag), dp<uint32_t>(c->slot),
                        (int32_t)next_beg, skeys, dp<double>(c->region), sprob, s));
    lbeg = (int32_t)next_beg;
    lcnt = (int32_t)(8 * splits);
  }
  c->ncells = c->lvl.back();
  c->nleaves = nleaves;
  c->tree_hint.clear();
  for (size_t d = 0; d + 1 < c->lvl.size(); d++) c->tree_hint.push_back(c->lvl[d + 1] - c->lvl[d]);
  TRY(hipMemcpyAsync(c->h_region, c->region.ptr, 48, hipMemcpyDeviceToHost, s));
  return TFMM_OK;
}

hipError_t rec(tfmm_ctx *c, int ev, hipStream_t st);
int sort_list(tfmm_ctx *c, DevBuf &pairs, uint32_t npairs, DevBuf &off, DevBuf &src);

// The Cartesian basis's upward pass: P2M, then M2M level by level (Mc);
// for the tile kernel's route also the spherical images of the multipoles (M).
int cart_upward(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  TRY(launch_fill_zero(c->Mc.ptr, cart_bytes(c->ncells), s));
  TRY(launch_cart_p2m(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->Mc), s));
  const int nl = (int)c->lvl.size() - 1;
  for (int d = nl - 2; d >= 0; d--)
    TRY(launch_cart_m2m(k, c->lvl[d + 1], (c->lvl[d + 2] - c->lvl[d + 1]) / 8, dp<double>(c->Mc),
                        s));
  if (!c->cart_direct)
    TRY(launch_cart_to_sph((int32_t)c->ncells, dp<double>(c->Mc), dp<double>(c->M),
                           c->cart_tab.ptr, s));
  return TFMM_OK;
}

// The direct Cartesian M2L: every pair of the traversal from one per-target
// list (sorted by sort_list into m2l_off / m2l_src), then the far field of
// periodic images.
int cart_m2l(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  TRY(launch_fill_zero(c->Lc.ptr, cart_bytes(c->ncells), s));
  TRY(rec(c, EV_M2L0, s));
  Csr ml{dp<int32_t>(c->m2l_off), dp<int32_t>(c->m2l_src)};
  TRY(launch_cart_m2l(k, ml, (int32_t)c->ncells, dp<double>(c->Mc), dp<double>(c->Lc),
                      c->cart_tab.ptr, period(c), s));
  TRY(rec(c, EV_M2LT, s));
  TRY(launch_cart_periodic_far(dp<double>(c->Mc), dp<double>(c->Lc), c->cycle, c->images,
                               c->cart_tab.ptr, s));
  return TFMM_OK;
}

// The list of the direct Cartesian M2L (no offset groups, no MFMA batches).
int cart_lists(tfmm_ctx *c) {
  int rc = sort_list(c, c->m2l_pairs, c->n_m2l_exec, c->m2l_off, c->m2l_src);
  if (rc) return rc;
  c->n_groups = c->n_grouped = c->n_fallback = c->n_items = 0;
  return cart_tables(c);
}

// The tile kernel's route: its spherical locals (L) into the Cartesian ones.
hipError_t cart_from_tiles(tfmm_ctx *c) {
  if (!cartesian(c) || c->cart_direct) return hipSuccess;
  return launch_sph_to_cart((int32_t)c->ncells, dp<double>(c->L), dp<double>(c->Lc),
                            c->cart_tab.ptr, c->stream);
}

int cart_downward(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  const int nl = (int)c->lvl.size() - 1;
  for (int d = 0; d + 1 < nl; d++)
    TRY(launch_cart_l2l(k, c->lvl[d + 1], c->lvl[d + 2] - c->lvl[d + 1], dp<double>(c->Lc), s));
  TRY(launch_cart_l2p(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->Lc),
                      dp<double>(c->acc), dp<int32_t>(c->perm), dp<double>(c->trg),
                      dp<double>(c->out), s));
  return TFMM_OK;
}

int do_upward(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  TRY(reserve_expansions(c));
  if (cartesian(c)) {
    int rc = cart_tables(c);
    if (rc) return rc;
    return cart_upward(c);
  }
  const size_t bytes = sph_bytes(c->ncells);
  TRY(hipMemsetAsync(c->M.ptr, 0, bytes, s));
  const Cells k = cells_of(c);
  TRY(launch_p2m(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->M), s));
  const int nl = (int)c->lvl.size() - 1;
  // the unit translation operators, shared by every level (built once)
  if (!c->tops_ready) {
    TRY(dev_reserve(c->tops, translation_ops_doubles() * 8));
    TRY(launch_translation_ops(dp<double>(c->tops), s));
    c->tops_ready = true;
  }
  for (int d = nl - 2; d >= 0; d--)
    TRY(launch_m2m(k, c->lvl[d + 1], (c->lvl[d + 2] - c->lvl[d + 1]) / 8, dp<double>(c->tops),
                   dp<double>(c->M), s));
  return TFMM_OK;
}

int sort_list(tfmm_ctx *c, DevBuf &pairs, uint32_t npairs, DevBuf &off, DevBuf &src) {
  hipStream_t s = c->stream;
  const int sbits = bits_for(c->ncells);
  // (target, source) order: targets by a stable radix sort, then each
  // target's sources sorted in place of the sort's low passes
  TRY(dev_reserve(c->lkeys, (size_t)npairs * 4 + 4));
  TRY(dev_reserve(c->lkeys_alt, (size_t)npairs * 4 + 4));
  TRY(dev_reserve(c->lidx, (size_t)npairs * 4 + 4));
  TRY(dev_reserve(c->lidx_alt, (size_t)npairs * 4 + 4));
  TRY(dev_reserve(c->sort_tmp, radix_tmp_bytes(npairs)));
  TRY(dev_reserve(off, (size_t)(c->ncells + 1) * 4));
  TRY(dev_reserve(src, (size_t)npairs * 4 + 4));
  TRY(launch_split_pairs(dp<int2>(pairs), npairs, dp<uint32_t>(c->lkeys), dp<uint32_t>(c->lidx),
                         s));
  bool alt = false;
  TRY(radix_sort_pairs32(dp<uint32_t>(c->lkeys), dp<uint32_t>(c->lidx), dp<uint32_t>(c->lkeys_alt),
                         dp<uint32_t>(c->lidx_alt), npairs, 0, sbits, c->sort_tmp.ptr, s, &alt));
  TRY(dev_reserve(c->seg_long, (size_t)(c->ncells + 1) * 4));
  TRY(launch_csr_sorted_sources(dp<uint32_t>(alt ? c->lkeys_alt : c->lkeys),
                                dp<uint32_t>(alt ? c->lidx_alt : c->lidx), npairs,
                                (int32_t)c->ncells, dp<int32_t>(off), dp<int32_t>(src),
                                dp<uint32_t>(c->seg_long), s));
  return TFMM_OK;
}

// The traversal runs 2 max_depth + 2 breadth-first steps (every step refines
// one side of a pair by one level, so no pair survives longer), each reading
// its input size on the device: one host round trip for the whole traversal.
// The frontier buffers and the lists have capacities; if any count exceeds
// its capacity the traversal is run again with the observed sizes.
int do_dtt(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  const int max_depth = (int)c->lvl.size() - 2;
  const int nsteps = 2 * max_depth + 2;
  const int cwords = 16 + nsteps + 1;  // list counters, then frontier sizes 0..nsteps
  TRY(dev_reserve(c->cnt, (size_t)cwords * 4));
  if (c->m2l_cap == 0) c->m2l_cap = (uint32_t)std::min<int64_t>(UINT32_MAX / 2, std::max<int64_t>(1 << 16, 200 * c->ncells));
  if (c->p2p_cap == 0) c->p2p_cap = (uint32_t)std::min<int64_t>(UINT32_MAX / 2, std::max<int64_t>(1 << 16, 40 * c->nleaves));
  if (c->front_cap == 0) c->front_cap = (uint32_t)std::min<int64_t>(UINT32_MAX / 2, std::max<int64_t>(1 << 20, 256 * c->ncells));
  if (cwords > H_U32_WORDS) return fail(c, TFMM_EINVAL, "tree too deep");
  TRY(dev_reserve(c->dtt_geo, (size_t)c->ncells * 32));
  TRY(dev_reserve(c->dtt_topo, (size_t)c->ncells * 8));
  TRY(launch_dtt_cells(k, (int32_t)c->ncells, c->prm.theta, dp<double4>(c->dtt_geo),
                       dp<int2>(c->dtt_topo), s));
  // A frontier that overflowed truncates every later step, so the counts
  // seen after an overflow are lower bounds: capacities at least double.
  // M2L group-table flags (one per slot), then the count of M2L pairs
  // outside the table, filled by the steps beside the lists
  const int32_t nslots = m2l_group_slots(max_depth);
  TRY(dev_reserve(c->g_used, (size_t)(nslots + 1) * 4));
  for (int attempt = 0; attempt < 12; attempt++) {
    TRY(dev_reserve(c->m2l_pairs, (size_t)c->m2l_cap * 8));
    TRY(dev_reserve(c->slot_of, (size_t)c->m2l_cap * 4));
    TRY(hipMemsetAsync(c->g_used.ptr, 0, (size_t)(nslots + 1) * 4, s));
    TRY(dev_reserve(c->p2p_pairs, (size_t)c->p2p_cap * 8));
    TRY(dev_reserve(c->front_a, (size_t)c->front_cap * 8));
    TRY(dev_reserve(c->front_b, (size_t)c->front_cap * 8));
    uint32_t *cnt = dp<uint32_t>(c->cnt);
    TRY(hipMemsetAsync(cnt, 0, (size_t)cwords * 4, s));
    TRY(launch_root_pairs(dp<int2>(c->front_a), c->nprob, cnt + 16, s, c->images > 0));
    // grid of step st: the frontier size of the previous traversal of the same
    // tree shape when there is one (checked below), else the capacity
    const bool hint = (int)c->dtt_hint.size() == nsteps + 1 && c->dtt_hint_cells == c->ncells;
    std::vector<uint32_t> grid(nsteps);
    for (int st = 0; st < nsteps; st++) {
      grid[st] = hint ? std::min<uint32_t>(c->front_cap, c->dtt_hint[st]) : c->front_cap;
      TRY(launch_dtt_step(dp<int2>(c->front_a), cnt + 16 + st, grid[st], dp<int2>(c->front_b),
                          c->front_cap, cnt + 16 + st + 1, dp<int2>(c->m2l_pairs), c->m2l_cap,
                          dp<int2>(c->p2p_pairs), c->p2p_cap, cnt, dp<double4>(c->dtt_geo),
                          dp<int2>(c->dtt_topo), k.depth, k.width, dp<int32_t>(c->slot_of),
                          dp<uint32_t>(c->g_used), dp<uint32_t>(c->g_used) + nslots,
                          c->prm.mutual, s, period(c)));
      std::swap(c->front_a, c->front_b);
    }
    TRY(hipMemcpyAsync(c->h_u32, cnt, (size_t)cwords * 4, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    uint32_t fmax = 0;
    c->dtt_steps = 0;
    for (int st = 0; st <= nsteps; st++) {
      fmax = std::max(fmax, c->h_u32[16 + st]);
      if (c->h_u32[16 + st]) c->dtt_steps = st + 1;
    }
    bool grid_short = false;
    for (int st = 0; st < nsteps; st++)
      grid_short |= grid[st] < c->front_cap && c->h_u32[16 + st] > grid[st];
    if (grid_short) {  // a step saw more pairs than its hinted grid: run again unhinted
      c->dtt_hint.clear();
      attempt--;
      continue;
    }
    if (c->h_u32[16 + nsteps] != 0) return fail(c, TFMM_EHIP, "traversal did not terminate");
    const uint32_t nm = c->h_u32[1], np = c->h_u32[2];
    if (nm <= c->m2l_cap && np <= c->p2p_cap && fmax <= c->front_cap) {
      c->dtt_hint.assign(c->h_u32 + 16, c->h_u32 + 16 + nsteps + 1);
      c->dtt_hint_cells = c->ncells;
      c->n_m2l_exec = nm;
      c->n_p2p = np;
      c->n_m2l_all = c->h_u32[3];
      c->n_p2p_calls_leaf = c->h_u32[4];
      return TFMM_OK;
    }
    auto grow = [](uint32_t cap, uint32_t seen) {
      const uint64_t want = std::max<uint64_t>(2 * (uint64_t)cap, (uint64_t)seen + seen / 4 + 1024);
      return (uint32_t)std::min<uint64_t>(UINT32_MAX / 2, want);
    };
    const bool front_over = fmax > c->front_cap;
    if (front_over || nm > c->m2l_cap) c->m2l_cap = grow(c->m2l_cap, nm);
    if (front_over || np > c->p2p_cap) c->p2p_cap = grow(c->p2p_cap, np);
    if (front_over) c->front_cap = grow(c->front_cap, fmax);
  }
  return fail(c, TFMM_EHIP, "traversal lists did not converge");
}

// Group the executed M2L pairs by (target level, source level, offset), build
// the MFMA work items and the per-target list of the pairs outside the groups.
int build_m2l_batches(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  const uint32_t n = c->n_m2l_exec;
  const int max_depth = (int)c->lvl.size() - 2;
  const int32_t nslots = m2l_group_slots(max_depth);
  // g_used (nslots flags, then the count of pairs outside the table) and
  // slot_of (every pair's slot) come from the traversal
  TRY(dev_reserve(c->g_gid, (size_t)nslots * 4));
  TRY(dev_reserve(c->scan_tmp, scan_tmp_bytes(std::max<int64_t>(nslots, n))));
  TRY(scan_exclusive_u32(dp<uint32_t>(c->g_used), dp<uint32_t>(c->g_gid), nslots,
                         c->scan_tmp.ptr, s));
  TRY(hipMemcpyAsync(c->h_u32, dp<uint32_t>(c->g_gid) + nslots - 1, 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(c->h_u32 + 1, dp<uint32_t>(c->g_used) + nslots - 1, 8,
                     hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  const uint32_t ng = c->h_u32[0] + c->h_u32[1];
  const uint32_t n_outside = c->h_u32[2];
  c->n_groups = ng;
  const size_t ngr = std::max<uint32_t>(ng, 1);
  TRY(dev_reserve(c->g_y64, ngr * 112 * 8));
  TRY(dev_reserve(c->g_y32, ngr * 112 * 4));
  TRY(dev_reserve(c->g_dl, ngr));
  TRY(dev_reserve(c->g_lt, ngr));
  TRY(launch_m2l_group_harmonics(dp<uint32_t>(c->g_used), dp<uint32_t>(c->g_gid), nslots,
                                 dp<double>(c->g_y64), dp<float>(c->g_y32), dp<int8_t>(c->g_dl),
                                 dp<int8_t>(c->g_lt), s));
  // split grouped / fallback pairs
  const size_t nn = (size_t)std::max<uint32_t>(n, 1);
  TRY(dev_reserve(c->gkeys, nn * 4));
  TRY(dev_reserve(c->gvals, nn * 4));
  TRY(dev_reserve(c->gkeys_alt, nn * 4));
  TRY(dev_reserve(c->gvals_alt, nn * 4));
  TRY(dev_reserve(c->fkeys, nn * 8));
  TRY(dev_reserve(c->fvals, nn * 4));
  TRY(dev_reserve(c->fkeys_alt, nn * 8));
  TRY(dev_reserve(c->fvals_alt, nn * 4));
  TRY(dev_reserve(c->g_cnt, 16));
  TRY(hipMemsetAsync(c->g_cnt.ptr, 0, 16, s));
  const int gbits = bits_for(ngr);
  const int sbits = bits_for(c->ncells);
  {
    const int64_t nt = (c->ncells + m2l_tile_size() - 1) / m2l_tile_size();
    if (bits_for(nt) + gbits > 32 || gbits > m2l_group_bits() ||
        c->ncells >= ((int64_t)1 << m2l_src_bits()))
      return fail(c, TFMM_EHIP, "tree too large for the batched M2L keys");
  }
  uint32_t ngp = n, nfb = 0;
  if (n_outside == 0) {  // the usual case: no compaction, no read-back
    TRY(launch_m2l_classify_all(dp<int2>(c->m2l_pairs), dp<int32_t>(c->slot_of), n,
                                dp<uint32_t>(c->g_gid), gbits, dp<uint32_t>(c->gkeys),
                                dp<uint32_t>(c->gvals), s));
  } else {
    TRY(launch_m2l_classify(dp<int2>(c->m2l_pairs), dp<int32_t>(c->slot_of), n, dp<uint32_t>(c->g_gid), gbits, FB_SRC_BITS,
                            dp<uint32_t>(c->gkeys), dp<uint32_t>(c->gvals), dp<uint64_t>(c->fkeys),
                            dp<uint32_t>(c->fvals), dp<uint32_t>(c->g_cnt), s));
    TRY(hipMemcpyAsync(c->h_u32, c->g_cnt.ptr, 8, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    ngp = c->h_u32[0];
    nfb = c->h_u32[1];
    if (nfb != n_outside || ngp + nfb != n) return fail(c, TFMM_EHIP, "M2L pair split is inconsistent");
  }
  c->n_grouped = ngp;
  c->n_fallback = nfb;
  const int32_t tile = m2l_tile_size();
  const int32_t ntiles = (int32_t)((c->ncells + tile - 1) / tile);
  const int tbits = bits_for(ntiles);
  TRY(dev_reserve(c->sort_tmp, radix_tmp_bytes(n)));
  bool alt = false;
  // all n keys: the grouped ones sort to the first ngp places
  TRY(radix_sort_pairs32(dp<uint32_t>(c->gkeys), dp<uint32_t>(c->gvals), dp<uint32_t>(c->gkeys_alt),
                         dp<uint32_t>(c->gvals_alt), n, 0, tbits + gbits, c->sort_tmp.ptr, s,
                         &alt));
  c->grouped_keys_in_alt = alt;
  const uint32_t *gk = dp<uint32_t>(alt ? c->gkeys_alt : c->gkeys);
  // work items: runs of <= 16 pairs of one group within one tile
  const size_t ng1 = (size_t)std::max<uint32_t>(ngp, 1);
  TRY(dev_reserve(c->it_flag, ng1 * 4));
  const size_t nblk = (ng1 + 1023) / 1024;
  TRY(dev_reserve(c->it_iflag, nblk * 4));
  TRY(dev_reserve(c->it_iexcl, nblk * 4));
  uint32_t nitems = 0;
  TRY(launch_m2l_count_items(gk, ngp, dp<uint32_t>(c->it_flag), dp<uint32_t>(c->it_iflag),
                             dp<uint32_t>(c->it_iexcl), c->scan_tmp.ptr, c->h_u32, &nitems, s));
  const size_t ni1 = std::max<uint32_t>(nitems, 1);
  TRY(dev_reserve(c->items, ni1 * 16));
  TRY(dev_reserve(c->cols, ni1 * 16 * 8));
  const uint32_t *gvs = dp<uint32_t>(alt ? c->gvals_alt : c->gvals);
  TRY(launch_m2l_write_items(gk, gvs, ngp, gbits, dp<uint32_t>(c->it_flag),
                             dp<uint32_t>(c->it_iexcl), dp<int8_t>(c->g_dl), dp<int4>(c->items),
                             dp<uint2>(c->cols), s));
  c->n_items = nitems;
  TRY(dev_reserve(c->tile_off, (size_t)(ntiles + 1) * 4));
  TRY(launch_m2l_tile_offsets(dp<int4>(c->items), nitems, ntiles, dp<int32_t>(c->tile_off), s));
  c->unit_chunk = m2l_unit_chunk(nitems, ntiles, c->m2l_unit_factor);
  c->units_bound = m2l_units_bound(nitems, ntiles, c->unit_chunk);
  TRY(dev_reserve(c->u_cnt, (size_t)ntiles * 4));
  TRY(dev_reserve(c->u_off, (size_t)ntiles * 4));
  TRY(dev_reserve(c->units, (size_t)c->units_bound * 8));
  TRY(dev_reserve(c->u_part, m2l_unit_part_bytes(c->units_bound, ntiles)));
  TRY(dev_reserve(c->scan_tmp, scan_tmp_bytes(std::max<int64_t>(ntiles, std::max<int64_t>(nslots, n)))));
  TRY(launch_m2l_units(dp<int32_t>(c->tile_off), ntiles, c->unit_chunk, dp<uint32_t>(c->u_cnt),
                       dp<uint32_t>(c->u_off), dp<int2>(c->units), c->scan_tmp.ptr, s));
  // fallback pairs: per-target list sorted by (target, source)
  bool falt = false;
  TRY(radix_sort_pairs(dp<uint64_t>(c->fkeys), dp<uint32_t>(c->fvals), dp<uint64_t>(c->fkeys_alt),
                       dp<uint32_t>(c->fvals_alt), nfb, 0, FB_SRC_BITS + sbits, c->sort_tmp.ptr, s,
                       &falt));
  const uint64_t *fk = dp<uint64_t>(falt ? c->fkeys_alt : c->fkeys);
  TRY(dev_reserve(c->fb_off, (size_t)(c->ncells + 1) * 4));
  TRY(dev_reserve(c->fb_src, (size_t)std::max<uint32_t>(nfb, 1) * 4));
  TRY(launch_csr_offsets(fk, nfb, FB_SRC_BITS, (int32_t)c->ncells, dp<int32_t>(c->fb_off), s));
  TRY(launch_unpack_pairs(fk, nfb, FB_SRC_BITS, dp<int32_t>(c->fb_src), s));
  if (!c->codes_ready) {
    TRY(dev_reserve(c->codes, m2l_codes_bytes()));
    hipError_t e = m2l_upload_codes(dp<uint32_t>(c->codes), s);
    if (e == hipErrorInvalidValue) return fail(c, TFMM_EHIP, "M2L operator table is inconsistent");
    TRY(e);
    c->codes_ready = true;
  }
  // every group's A operands in the tile kernel's order
  const bool mixed = m2l_fp32(c);
  TRY(dev_reserve(c->g_af, m2l_afrag_bytes(ng, mixed)));
  TRY(launch_m2l_group_afrag(mixed ? (const void *)c->g_y32.ptr : (const void *)c->g_y64.ptr,
                             ng, dp<uint32_t>(c->codes), dp<int8_t>(c->g_dl), c->g_af.ptr,
                             mixed, s));
  return TFMM_OK;
}

// resident P2P workgroups (one wave each) for the knob's waves per SIMD
int p2p_resident(tfmm_ctx *c) { return c->p2p_waves > 0 ? c->p2p_waves * 4 * c->n_cu : 0; }

// The M2L and P2P pairs of the traversal (c->m2l_pairs with their group
// slots, c->p2p_pairs, the counters) into sorted lists, M2L batches and the
// M2L and P2P kernels.
int run_lists(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  const bool mixed = c->prm.precision == TFMM_MIXED;
  int rc = sort_list(c, c->p2p_pairs, c->n_p2p, c->p2p_off, c->p2p_src);
  if (rc) return rc;
  if (mixed) {
    TRY(dev_reserve(c->relpos, (size_t)c->n * 16));
    TRY(launch_leaf_relpos(k, (int32_t)c->ncells, sorted_bodies(c), dp<float4>(c->relpos), s));
  }
  // P2P on the side stream: it needs only the P2P list and the bodies.
  // Schedule 0 starts it as soon as its list is sorted, so it covers the host
  // round trips of the M2L batch building (which then gets few CU slots);
  // schedule 1 starts it together with the tile kernel, so the P2P (FP32
  // VALU) and the tiles (MFMA, LDS) share the CUs.  Joined before the
  // downward pass.
  TRY(dev_reserve(c->acc, (size_t)c->n * 32));
  auto launch_p2p = [&]() -> int {
    TRY(hipEventRecord(c->ev[EV_P2PREADY], s));
    TRY(hipStreamWaitEvent(c->side, c->ev[EV_P2PREADY], 0));
    TRY(hipEventRecord(c->ev[EV_P2P0], c->side));
    Csr pl{dp<int32_t>(c->p2p_off), dp<int32_t>(c->p2p_src)};
    if (mixed)
      TRY(launch_p2p_fp32(k, (int32_t)c->ncells, pl, sorted_bodies(c), dp<float4>(c->relpos),
                          dp<double>(c->acc), p2p_resident(c), dp<uint32_t>(c->p2p_next),
                          c->side, period(c)));
    else
      TRY(launch_p2p_fp64(k, (int32_t)c->ncells, pl, sorted_bodies(c), dp<double>(c->acc),
                          p2p_resident(c), dp<uint32_t>(c->p2p_next), c->side, period(c)));
    TRY(hipEventRecord(c->ev[EV_P2P1], c->side));
    return TFMM_OK;
  };
  const bool p2p_early = c->p2p_schedule != 1;
  if (p2p_early) {
    rc = launch_p2p();
    if (rc) return rc;
  }
  if (cart_direct(c)) {
    rc = cart_lists(c);
    if (!rc) {
      TRY(hipEventRecord(c->ev[EV_DTT1], s));
      if (!p2p_early) rc = launch_p2p();
    }
    if (!rc) rc = cart_m2l(c);
    if (!rc) TRY(hipEventRecord(c->ev[EV_M2L1], s));
    TRY(hipStreamWaitEvent(s, c->ev[EV_P2P1], 0));
    if (rc) return rc;
    TRY(hipMemcpyAsync(c->h_u64, static_cast<char *>(c->cnt.ptr) + 32, 16,
                       hipMemcpyDeviceToHost, s));
    return TFMM_OK;
  }
  rc = build_m2l_batches(c);
  if (rc) {
    if (p2p_early) (void)hipStreamWaitEvent(s, c->ev[EV_P2P1], 0);
    return rc;
  }
  // M~ in the MFMA operand layout
  const bool tf32 = m2l_fp32(c);
  TRY(dev_reserve(c->mq, (size_t)c->ncells * 112 * (tf32 ? 4 : 8)));
  TRY(launch_m2l_scale_m(k, (int32_t)c->ncells, dp<double>(c->M), c->mq.ptr, tf32, s));
  TRY(hipEventRecord(c->ev[EV_DTT1], s));

  const size_t lbytes = sph_bytes(c->ncells);
  TRY(hipMemsetAsync(c->L.ptr, 0, lbytes, s));
  TRY(hipEventRecord(c->ev[EV_M2L0], s));
  if (!p2p_early) {
    rc = launch_p2p();
    if (rc) return rc;
  }
  const int32_t tile = m2l_tile_size();
  const int32_t ntiles = (int32_t)((c->ncells + tile - 1) / tile);
  TRY(launch_m2l_tiles(dp<int32_t>(c->tile_off), dp<int2>(c->units), dp<uint32_t>(c->u_cnt),
                       dp<uint32_t>(c->u_off), ntiles, c->unit_chunk, c->m2l_xcd_run,
                       c->units_bound,
                       dp<uint2>(c->cols),
                       c->g_af.ptr, c->mq.ptr, k, (int32_t)c->ncells,
                       dp<double>(c->L), dp<double>(c->u_part), tf32, s));
  TRY(hipEventRecord(c->ev[EV_M2LT], s));
  if (c->n_fallback) {
    Csr fl{dp<int32_t>(c->fb_off), dp<int32_t>(c->fb_src)};
    TRY(dev_reserve(c->fb_part, m2l_pairs_scratch_bytes((int32_t)c->n_fallback)));
    TRY(launch_m2l_fp64(k, (int32_t)c->ncells, fl, (int32_t)c->n_fallback, dp<double>(c->M),
                        dp<double>(c->L), c->fb_part.ptr, s, nullptr, period(c)));
  }
  TRY(launch_periodic_far(dp<double>(c->M), dp<double>(c->L), c->cycle, c->images, s));
  TRY(cart_from_tiles(c));
  TRY(hipEventRecord(c->ev[EV_M2L1], s));

  TRY(hipStreamWaitEvent(s, c->ev[EV_P2P1], 0));
  TRY(hipMemcpyAsync(c->h_u64, static_cast<char *>(c->cnt.ptr) + 32, 16, hipMemcpyDeviceToHost, s));
  return TFMM_OK;
}

int do_traverse(tfmm_ctx *c) {
  TRY(hipEventRecord(c->ev[EV_DTT0], c->stream));
  int rc = do_dtt(c);
  if (rc) return rc;
  return run_lists(c);
}

// Interaction lists recorded outside the library (the reference's FMM_DTT
// functor run on host cells, TP/hot/mapper.h:452-690 records its P2P products
// the same way) in place of the GPU traversal: M2L pairs (target, source) as
// the functor called M2L, empty cells included and counted like the
// reference's numM2L; P2P pairs of leaves.  Pairs with an empty cell are
// counted but not executed, as in the traversal.
int traverse_given(tfmm_ctx *c, const int32_t *m2l, int64_t nm, const int32_t *p2p, int64_t np) {
  if (c->images > 0)
    return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: lists of periodic images are not taken");
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  if (nm < 0 || np < 0 || (nm && !m2l) || (np && !p2p))
    return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: bad list");
  if (nm > (int64_t)(UINT32_MAX / 2) || np > (int64_t)(UINT32_MAX / 2))
    return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: list too long");
  const int64_t nc = c->ncells;
  std::vector<int32_t> nb(nc), child(nc);
  TRY(hipMemcpyAsync(nb.data(), k.nb, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(child.data(), k.child, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  // Counters as the reference counts calls: with mutual the lists hold each
  // unordered pair in both directions, one call of the reference's functors
  // (the entry with i <= j counts; a leaf with itself visits the j <= i body
  // pairs).
  const bool mutual = c->prm.mutual != 0;
  std::vector<int2> me, pe;
  me.reserve((size_t)nm);
  pe.reserve((size_t)np);
  uint64_t m2l_calls = 0, p2p_calls = 0;
  for (int64_t e = 0; e < nm; e++) {
    const int32_t i = m2l[2 * e], j = m2l[2 * e + 1];
    if (i < 0 || i >= nc || j < 0 || j >= nc)
      return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: M2L cell index out of range");
    if (!mutual || i <= j) m2l_calls++;
    if (nb[i] > 0 && nb[j] > 0) me.push_back(make_int2(i, j));
  }
  unsigned long long pairs = 0, self = 0;
  for (int64_t e = 0; e < np; e++) {
    const int32_t i = p2p[2 * e], j = p2p[2 * e + 1];
    if (i < 0 || i >= nc || j < 0 || j >= nc)
      return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: P2P cell index out of range");
    if (child[i] >= 0 || child[j] >= 0)
      return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: P2P pair of a cell that is not a leaf");
    if (nb[i] == 0 || nb[j] == 0) continue;
    pe.push_back(make_int2(i, j));
    if (mutual && i > j) continue;
    p2p_calls++;
    if (mutual && i == j)
      pairs += (unsigned long long)nb[i] * (unsigned long long)(nb[i] + 1) / 2;
    else
      pairs += (unsigned long long)nb[i] * (unsigned long long)nb[j];
    if (i == j) self += (unsigned long long)nb[i];
  }
  TRY(hipEventRecord(c->ev[EV_DTT0], s));
  const int max_depth = (int)c->lvl.size() - 2;
  const int32_t nslots = m2l_group_slots(max_depth);
  const size_t nme = std::max<size_t>(me.size(), 1), npe = std::max<size_t>(pe.size(), 1);
  TRY(dev_reserve(c->cnt, 16 * 4));
  TRY(dev_reserve(c->g_used, (size_t)(nslots + 1) * 4));
  TRY(dev_reserve(c->m2l_pairs, nme * 8));
  TRY(dev_reserve(c->slot_of, nme * 4));
  TRY(dev_reserve(c->p2p_pairs, npe * 8));
  TRY(dev_reserve(c->dtt_geo, (size_t)nc * 32));
  TRY(dev_reserve(c->dtt_topo, (size_t)nc * 8));
  TRY(hipMemsetAsync(c->g_used.ptr, 0, (size_t)(nslots + 1) * 4, s));
  c->h_u64[0] = pairs;
  c->h_u64[1] = self;
  TRY(hipMemcpyAsync(static_cast<char *>(c->cnt.ptr) + 32, c->h_u64, 16, hipMemcpyHostToDevice, s));
  if (!me.empty())
    TRY(hipMemcpyAsync(c->m2l_pairs.ptr, me.data(), me.size() * 8, hipMemcpyHostToDevice, s));
  if (!pe.empty())
    TRY(hipMemcpyAsync(c->p2p_pairs.ptr, pe.data(), pe.size() * 8, hipMemcpyHostToDevice, s));
  TRY(launch_dtt_cells(k, (int32_t)nc, c->prm.theta, dp<double4>(c->dtt_geo),
                       dp<int2>(c->dtt_topo), s));
  TRY(launch_pair_slots(dp<int2>(c->m2l_pairs), (uint32_t)me.size(), dp<double4>(c->dtt_geo),
                        k.depth, k.width, dp<int32_t>(c->slot_of), dp<uint32_t>(c->g_used),
                        dp<uint32_t>(c->g_used) + nslots, s));
  // the uploads read host vectors that end with this call
  TRY(hipStreamSynchronize(s));
  c->n_m2l_exec = (uint32_t)me.size();
  c->n_m2l_all = m2l_calls;
  c->n_p2p = (uint32_t)pe.size();
  c->n_p2p_calls_leaf = p2p_calls;
  return run_lists(c);
}

int do_downward(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  if (cartesian(c)) {
    TRY(dev_reserve(c->trg, (size_t)c->n * 32));
    TRY(dev_reserve(c->out, (size_t)c->n * 32));
    return cart_downward(c);
  }
  const int nl = (int)c->lvl.size() - 1;
  for (int d = 0; d + 1 < nl; d++)
    TRY(launch_l2l(k, c->lvl[d + 1], (c->lvl[d + 2] - c->lvl[d + 1]) / 8, dp<double>(c->tops),
                   dp<double>(c->L), s));
  TRY(dev_reserve(c->trg, (size_t)c->n * 32));
  TRY(dev_reserve(c->out, (size_t)c->n * 32));
  TRY(dev_reserve(c->leaf_of, (size_t)c->n * 4));
  TRY(launch_l2p(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->L), dp<double>(c->acc),
                 dp<int32_t>(c->perm), dp<int32_t>(c->leaf_of), dp<double>(c->trg),
                 dp<double>(c->out), s));
  return TFMM_OK;
}

// ---------------------------------------------------------------------------
// tfmm_evaluate's order.  Phase A holds everything whose sizes the host reads
// back: the traversal, the list sorts and the M2L batch building, then every
// buffer the numeric phase writes is reserved.  Phase B, the numeric work
// (P2M, M2M, P2P, M2L, L2L, L2P), has all its sizes fixed by then: it
// allocates nothing, reads nothing back and clears memory with a kernel, so
// it can be recorded into one hipGraph and launched as one unit.  P2P forks
// to the second stream at the start of phase B and overlaps the upward pass
// and the M2L kernels.
// A timing event: inside a capture it must become an event node of the graph
// (an external record), or the host could not read its time afterwards.  A
// HIP runtime without external records in captures (found by a probe) gets
// none: the phase times inside the graph then read 0.
hipError_t rec(tfmm_ctx *c, int ev, hipStream_t st) {
  if (!c->capturing) return hipEventRecord(c->ev[ev], st);
  if (c->ext_events != 1) return hipSuccess;
  return hipEventRecordWithFlags(c->ev[ev], st, hipEventRecordExternal);
}

// Whether this HIP runtime records external events inside a stream capture
// (an empty capture with one such record).
void probe_external_events(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  hipGraph_t g = nullptr;
  bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    const bool rec_ok =
        hipEventRecordWithFlags(c->ev[EV_DOWN], s, hipEventRecordExternal) == hipSuccess;
    ok = hipStreamEndCapture(s, &g) == hipSuccess && rec_ok;
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  c->ext_events = ok ? 1 : 0;
}

// The P2P on the second stream (forked from and later joined to the context
// stream by events).
// Schedule 3: the resident FP32 P2P starts early with its few waves, and a
// full-occupancy launch of the same kernel on the context stream takes the
// target cells still left on its counter once the tile kernel is done.
bool p2p_drains(const tfmm_ctx *c) {
  return c->p2p_schedule == 3 && c->p2p_waves > 0 && c->prm.precision == TFMM_MIXED;
}

int drain_p2p(tfmm_ctx *c, int64_t ncells) {
  if (!p2p_drains(c)) return TFMM_OK;
  const Cells k = cells_of(c);
  Csr pl{dp<int32_t>(c->p2p_off), dp<int32_t>(c->p2p_src)};
  TRY(launch_p2p_fp32_drain(k, (int32_t)ncells, pl, sorted_bodies(c), dp<float4>(c->relpos),
                            dp<double>(c->acc), 8 * 4 * c->n_cu, dp<uint32_t>(c->p2p_next),
                            c->stream, period(c)));
  return TFMM_OK;
}

int fork_p2p(tfmm_ctx *c, int64_t ncells = -1) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  if (ncells < 0) ncells = c->ncells;
  const bool mixed = c->prm.precision == TFMM_MIXED;
  // the drained schedule clears the shared counter on this stream, before
  // either launch can take from it
  if (p2p_drains(c)) TRY(launch_fill_zero(c->p2p_next.ptr, 4, s));
  // schedule 4: on the context stream, alone (the kernel's own time)
  hipStream_t ps = c->p2p_schedule == 4 ? s : c->side;
  if (ps != s) {
    TRY(hipEventRecord(c->ev[EV_P2PREADY], s));
    TRY(hipStreamWaitEvent(ps, c->ev[EV_P2PREADY], 0));
  }
  TRY(rec(c, EV_P2P0, ps));
  Csr pl{dp<int32_t>(c->p2p_off), dp<int32_t>(c->p2p_src)};
  if (p2p_drains(c))
    TRY(launch_p2p_fp32_drain(k, (int32_t)ncells, pl, sorted_bodies(c), dp<float4>(c->relpos),
                              dp<double>(c->acc), p2p_resident(c), dp<uint32_t>(c->p2p_next),
                              ps, period(c)));
  else if (mixed)
    TRY(launch_p2p_fp32(k, (int32_t)ncells, pl, sorted_bodies(c), dp<float4>(c->relpos),
                        dp<double>(c->acc), p2p_resident(c), dp<uint32_t>(c->p2p_next), ps,
                        period(c)));
  else
    TRY(launch_p2p_fp64(k, (int32_t)ncells, pl, sorted_bodies(c), dp<double>(c->acc),
                        p2p_resident(c), dp<uint32_t>(c->p2p_next), ps,
                        period(c)));
  TRY(rec(c, EV_P2P1, ps));
  TRY(hipEventRecord(c->ev[EV_P2PJOIN], ps));
  return TFMM_OK;
}

int fused_prepare(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  const bool mixed = c->prm.precision == TFMM_MIXED;
  TRY(hipEventRecord(c->ev[EV_DTT0], s));
  int rc = do_dtt(c);
  if (rc) return rc;
  rc = sort_list(c, c->p2p_pairs, c->n_p2p, c->p2p_off, c->p2p_src);
  if (rc) return rc;
  if (mixed) {
    TRY(dev_reserve(c->relpos, (size_t)c->n * 16));
    TRY(launch_leaf_relpos(k, (int32_t)c->ncells, sorted_bodies(c), dp<float4>(c->relpos), s));
  }
  TRY(dev_reserve(c->acc, (size_t)c->n * 32));
  c->p2p_forked = false;
  if (c->p2p_schedule >= 2 && !c->use_graph) {
    // the P2P list is complete: start the P2P now, beside the M2L batch
    // building (with resident waves it leaves that work CU slots)
    rc = fork_p2p(c);
    if (rc) return rc;
    c->p2p_forked = true;
  }
  rc = cart_direct(c) ? cart_lists(c) : build_m2l_batches(c);
  if (!rc && cartesian(c)) rc = cart_tables(c);
  if (rc) {
    if (c->p2p_forked) (void)hipStreamWaitEvent(s, c->ev[EV_P2PJOIN], 0);
    return rc;
  }
  TRY(hipMemcpyAsync(c->h_u64, static_cast<char *>(c->cnt.ptr) + 32, 16, hipMemcpyDeviceToHost, s));
  TRY(reserve_expansions(c));
  if (cart_direct(c)) {
    TRY(dev_reserve(c->trg, (size_t)c->n * 32));
    TRY(dev_reserve(c->out, (size_t)c->n * 32));
    TRY(hipEventRecord(c->ev[EV_DTT1], s));
    return TFMM_OK;
  }
  if (!c->tops_ready) {
    TRY(dev_reserve(c->tops, translation_ops_doubles() * 8));
    TRY(launch_translation_ops(dp<double>(c->tops), s));
    c->tops_ready = true;
  }
  TRY(dev_reserve(c->mq, (size_t)c->ncells * 112 * (m2l_fp32(c) ? 4 : 8)));
  TRY(dev_reserve(c->trg, (size_t)c->n * 32));
  TRY(dev_reserve(c->out, (size_t)c->n * 32));
  TRY(dev_reserve(c->leaf_of, (size_t)c->n * 4));
  if (c->n_fallback)
    TRY(dev_reserve(c->fb_part, m2l_pairs_scratch_bytes((int32_t)c->n_fallback)));
  TRY(hipEventRecord(c->ev[EV_DTT1], s));
  return TFMM_OK;
}

int fused_numeric(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  const Cells k = cells_of(c);
  // P2P on the second stream, joined before the downward pass; forked at the
  // start of the phase (schedule 0), at the tile kernel (1), or already in
  // phase A (2)
  int rc;
  if (!c->p2p_forked && c->p2p_schedule != 1 && (rc = fork_p2p(c))) return rc;
  if (cart_direct(c)) {
    if ((rc = cart_upward(c))) return rc;
    TRY(rec(c, EV_UP, s));
    if (!c->p2p_forked && c->p2p_schedule == 1 && (rc = fork_p2p(c))) return rc;
    if ((rc = cart_m2l(c))) return rc;
    TRY(rec(c, EV_M2L1, s));
    if ((rc = drain_p2p(c, c->ncells))) return rc;
    TRY(hipStreamWaitEvent(s, c->ev[EV_P2PJOIN], 0));
    TRY(rec(c, EV_TRAV, s));
    if ((rc = cart_downward(c))) return rc;
    TRY(rec(c, EV_DOWN, s));
    return TFMM_OK;
  }
  // upward (the Cartesian basis: in its own arrays, then the spherical images)
  const size_t ebytes = sph_bytes(c->ncells);
  const int nl = (int)c->lvl.size() - 1;
  if (cartesian(c)) {
    if ((rc = cart_upward(c))) return rc;
  } else {
    TRY(launch_fill_zero(c->M.ptr, ebytes, s));
    TRY(launch_p2m(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->M), s));
    for (int d = nl - 2; d >= 0; d--)
      TRY(launch_m2m(k, c->lvl[d + 1], (c->lvl[d + 2] - c->lvl[d + 1]) / 8,
                     dp<double>(c->tops), dp<double>(c->M), s));
  }
  TRY(rec(c, EV_UP, s));
  // M2L
  const bool mixed = m2l_fp32(c);
  TRY(launch_m2l_scale_m(k, (int32_t)c->ncells, dp<double>(c->M), c->mq.ptr, mixed, s));
  TRY(launch_fill_zero(c->L.ptr, ebytes, s));
  TRY(rec(c, EV_M2L0, s));
  if (!c->p2p_forked && c->p2p_schedule == 1 && (rc = fork_p2p(c))) return rc;
  const int32_t tile = m2l_tile_size();
  const int32_t ntiles = (int32_t)((c->ncells + tile - 1) / tile);
  TRY(launch_m2l_tiles(dp<int32_t>(c->tile_off), dp<int2>(c->units), dp<uint32_t>(c->u_cnt),
                       dp<uint32_t>(c->u_off), ntiles, c->unit_chunk, c->m2l_xcd_run,
                       c->units_bound, dp<uint2>(c->cols),
                       c->g_af.ptr, c->mq.ptr, k, (int32_t)c->ncells,
                       dp<double>(c->L), dp<double>(c->u_part), mixed, s));
  TRY(rec(c, EV_M2LT, s));
  if (c->n_fallback) {
    Csr fl{dp<int32_t>(c->fb_off), dp<int32_t>(c->fb_src)};
    TRY(launch_m2l_fp64(k, (int32_t)c->ncells, fl, (int32_t)c->n_fallback, dp<double>(c->M),
                        dp<double>(c->L), c->fb_part.ptr, s, nullptr, period(c)));
  }
  TRY(launch_periodic_far(dp<double>(c->M), dp<double>(c->L), c->cycle, c->images, s));
  TRY(cart_from_tiles(c));
  TRY(rec(c, EV_M2L1, s));
  if ((rc = drain_p2p(c, c->ncells))) return rc;
  TRY(hipStreamWaitEvent(s, c->ev[EV_P2PJOIN], 0));
  TRY(rec(c, EV_TRAV, s));
  // downward
  if (cartesian(c)) {
    if ((rc = cart_downward(c))) return rc;
    TRY(rec(c, EV_DOWN, s));
    return TFMM_OK;
  }
  for (int d = 0; d + 1 < nl; d++)
    TRY(launch_l2l(k, c->lvl[d + 1], (c->lvl[d + 2] - c->lvl[d + 1]) / 8, dp<double>(c->tops),
                   dp<double>(c->L), s));
  TRY(launch_l2p(k, (int32_t)c->ncells, sorted_bodies(c), dp<double>(c->L), dp<double>(c->acc),
                 dp<int32_t>(c->perm), dp<int32_t>(c->leaf_of), dp<double>(c->trg),
                 dp<double>(c->out), s));
  TRY(rec(c, EV_DOWN, s));
  return TFMM_OK;
}

// Phase B recorded into a graph and launched; the executable graph of the
// previous evaluation is updated in place when only parameters changed.
int fused_numeric_graph(tfmm_ctx *c) {
  hipStream_t s = c->stream;
  if (c->ext_events < 0) probe_external_events(c);
  TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  c->capturing = true;
  const int rc = fused_numeric(c);
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
  if (c->gexec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(c->gexec, g, &bad, &res) != hipSuccess ||
        res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
  }
  hipError_t e2 = hipSuccess;
  if (!c->gexec) e2 = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    c->gexec = nullptr;
    return hipfail(c, e2, "hipGraphInstantiate");
  }
  TRY(hipGraphLaunch(c->gexec, s));
  if (c->ext_events != 1) TRY(hipEventRecord(c->ev[EV_DOWN], s));  // the evaluation's end
  return TFMM_OK;
}

// ---------------------------------------------------------------------------
// The whole evaluation as one hipGraph (TFMM_TUNE_GRAPH = 2).  The first
// evaluation of a context runs the ordinary way and leaves a plan:
// capacities from its sizes with headroom.  Later evaluations with the same
// bodies' count, parameters and knobs launch one graph recorded with those
// capacities: every kernel reads the evaluation's actual sizes on the device
// (plan.hip), nothing is read back before the end, and the graph's last
// kernel checks the sizes against the capacities.  An evaluation that
// exceeded one runs again the ordinary way and makes a new plan; the graph
// is recorded again when a buffer moved.
// Pinned read-back of the graph: sizes, group counters, traversal counters,
// level table, pair counts and region.
constexpr int PH_SZ = 0, PH_GCNT = 16, PH_CNT = 32, PH_LEV = 96, PH_WORDS = 160;
constexpr size_t PH_U64_BYTE = PH_WORDS * 4, PH_REGION_BYTE = PH_U64_BYTE + 16;
constexpr size_t PLAN_HOST_BYTES = PH_REGION_BYTE + 48;
constexpr int PLAN_MAX_LEVELS = 30;  // 2 (L + 1) words of PH_LEV
constexpr int PLAN_MAX_CWORDS = 64;  // 16 + steps + 1 words of PH_CNT

uint64_t buffers_signature(tfmm_ctx *c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  for (DevBuf *b : all_buffers(c)) mix((uint64_t)(uintptr_t)b->ptr);
  mix((uint64_t)(uintptr_t)c->bx);
  mix((uint64_t)(uintptr_t)c->by);
  mix((uint64_t)(uintptr_t)c->bz);
  mix((uint64_t)(uintptr_t)c->bq);
  return h;
}

bool plan_fits(tfmm_ctx *c, const tfmm_params *p) {
  const Plan &pl = c->plan;
  return pl.valid && pl.n == c->n && pl.nprob == c->nprob && pl.prm.ncrit == p->ncrit &&
         pl.prm.theta == p->theta && pl.prm.precision == p->precision &&
         pl.prm.mutual == p->mutual && pl.unit_factor == c->m2l_unit_factor &&
         pl.xcd_run == c->m2l_xcd_run && pl.p2p_schedule == c->p2p_schedule &&
         pl.p2p_waves == c->p2p_waves;
}

void plan_drop(tfmm_ctx *c) {
  if (c->plan.exec) (void)hipGraphExecDestroy(c->plan.exec);
  c->plan = Plan();
}

// Capacities from the sizes of the evaluation that just ran (its tree, its
// traversal, its M2L batches): one more level than it had, an eighth more
// of everything else.  No plan when the sizes leave no room in the keys.
void plan_make(tfmm_ctx *c) {
  plan_drop(c);
  Plan pl;
  const int L0 = (int)c->tree_hint.size();
  if (L0 < 1 || L0 + 1 > PLAN_MAX_LEVELS || c->dtt_hint.empty()) return;
  auto more = [](int64_t v, int64_t extra) { return v + v / 8 + extra; };
  pl.n = c->n;
  pl.nprob = c->nprob;
  pl.prm = c->prm;
  pl.unit_factor = c->m2l_unit_factor;
  pl.xcd_run = c->m2l_xcd_run;
  pl.p2p_schedule = c->p2p_schedule;
  pl.p2p_waves = c->p2p_waves;
  pl.L = L0 + 1;
  pl.grid.resize(pl.L);
  pl.cells = 4096;
  for (int d = 0; d < pl.L; d++) {
    const int64_t h = d < L0 ? c->tree_hint[d] : std::max<int64_t>(c->tree_hint[L0 - 1], 1024);
    int64_t g = more(h, 256);
    if (d == 0) g = std::max<int64_t>(g, c->nprob);
    pl.grid[d] = (int32_t)std::min<int64_t>(g, INT32_MAX / 16);
    pl.cells += pl.grid[d];
  }
  if (pl.cells > INT32_MAX / 4) return;
  pl.max_depth = pl.L - 1;
  pl.nsteps = 2 * pl.max_depth + 2;
  pl.cwords = 16 + pl.nsteps + 1;
  if (pl.cwords > PLAN_MAX_CWORDS) return;
  pl.nslots = m2l_group_slots(pl.max_depth);
  int64_t fmax = 0;
  pl.step_grid.resize(pl.nsteps);
  for (int st = 0; st < pl.nsteps; st++) {
    const int64_t h = st < (int)c->dtt_hint.size() ? c->dtt_hint[st] : 0;
    pl.step_grid[st] = (uint32_t)std::min<int64_t>(UINT32_MAX / 4, more(h, 4096));
    fmax = std::max<int64_t>(fmax, pl.step_grid[st]);
  }
  auto cap32 = [](int64_t v) { return (uint32_t)std::min<int64_t>(UINT32_MAX / 4, v); };
  pl.caps.front = cap32(fmax);
  pl.caps.m2l = cap32(more(c->n_m2l_exec, 65536));
  pl.caps.p2p = cap32(more(c->n_p2p, 65536));
  pl.caps.items = cap32(more(c->n_items, 4096));
  pl.caps.fallback = cap32(more(c->n_fallback, 4096));
  pl.groups = cap32(more(c->n_groups, 256));
  pl.ntiles = (int32_t)((pl.cells + m2l_tile_size() - 1) / m2l_tile_size());
  const double f = c->m2l_unit_factor;
  pl.caps.units =
      cap32(f > 0 ? pl.ntiles + (int64_t)std::ceil(pl.ntiles / f) + 1 : (int64_t)pl.ntiles);
  pl.gbits = bits_for(pl.groups);
  pl.sbits = bits_for(pl.cells);
  pl.tbits = bits_for(pl.ntiles);
  if (pl.tbits + pl.gbits > 32 || pl.gbits > m2l_group_bits() ||
      pl.cells >= ((int64_t)1 << m2l_src_bits()) || FB_SRC_BITS + pl.sbits > 64 ||
      (int64_t)pl.caps.fallback > (1 << 22))
    return;
  pl.valid = true;
  c->plan = pl;
}

// every buffer of the graph, at the plan's capacities
int plan_reserve(tfmm_ctx *c) {
  const Plan &pl = c->plan;
  hipStream_t s = c->stream;
  const int64_t n = c->n;
  const bool mixed = pl.prm.precision == TFMM_MIXED;
  const size_t cells = (size_t)pl.cells;
  const size_t nm = pl.caps.m2l, npp = pl.caps.p2p;
  // tree
  TRY(dev_reserve(c->keys, n * 8));
  TRY(dev_reserve(c->keys_alt, n * 8));
  TRY(dev_reserve(c->idx, n * 4));
  TRY(dev_reserve(c->idx_alt, n * 4));
  const size_t sort_n = std::max<size_t>({(size_t)n, nm, npp});
  TRY(dev_reserve(c->sort_tmp, radix_tmp_bytes((int64_t)sort_n)));
  TRY(dev_reserve(c->sx, n * 8));
  TRY(dev_reserve(c->sy, n * 8));
  TRY(dev_reserve(c->sz, n * 8));
  TRY(dev_reserve(c->sq, n * 8));
  TRY(dev_reserve(c->perm, n * 4));
  TRY(dev_reserve(c->region, 6 * 8 * (size_t)c->nprob));
  TRY(dev_reserve(c->brec, n * 32));
  if (c->nprob > 1) {
    TRY(dev_reserve(c->bprob, n * 4));
    TRY(dev_reserve(c->pkey, n * 4));
    TRY(dev_reserve(c->pkey_alt, n * 4));
    TRY(dev_reserve(c->ppos, n * 4));
    TRY(dev_reserve(c->ppos_alt, n * 4));
  } else {
    const int nparts = (int)std::min<int64_t>(1024, (n + 255) / 256);
    TRY(dev_reserve(c->bbox_part, 2 * 8 * (size_t)nparts));
  }
  TRY(reserve_cells(c, cells, 0, s));
  int32_t gmax = 0;
  for (int32_t g : pl.grid) gmax = std::max(gmax, g);
  TRY(dev_reserve(c->lev, (size_t)2 * (pl.L + 1) * 4));
  TRY(dev_reserve(c->lev_status, 4));
  TRY(dev_reserve(c->flag, (size_t)gmax * 4 + 1024));
  TRY(dev_reserve(c->slot, (size_t)gmax * 4 + 1024));
  TRY(dev_reserve(c->lev_bsum, (size_t)level_blocks(gmax) * 4 + 4));
  // traversal and P2P list
  TRY(dev_reserve(c->cnt, (size_t)pl.cwords * 4));
  TRY(dev_reserve(c->dtt_geo, cells * 32));
  TRY(dev_reserve(c->dtt_topo, cells * 8));
  TRY(dev_reserve(c->g_used, (size_t)(pl.nslots + 1) * 4));
  TRY(dev_reserve(c->m2l_pairs, nm * 8));
  TRY(dev_reserve(c->slot_of, nm * 4));
  TRY(dev_reserve(c->p2p_pairs, npp * 8));
  TRY(dev_reserve(c->front_a, (size_t)pl.caps.front * 8));
  TRY(dev_reserve(c->front_b, (size_t)pl.caps.front * 8));
  TRY(dev_reserve(c->lkeys, npp * 4 + 4));
  TRY(dev_reserve(c->lkeys_alt, npp * 4 + 4));
  TRY(dev_reserve(c->lidx, npp * 4 + 4));
  TRY(dev_reserve(c->lidx_alt, npp * 4 + 4));
  TRY(dev_reserve(c->p2p_off, (cells + 1) * 4));
  TRY(dev_reserve(c->seg_long, (cells + 1) * 4));
  TRY(dev_reserve(c->p2p_src, npp * 4 + 4));
  if (mixed) TRY(dev_reserve(c->relpos, (size_t)n * 16));
  TRY(dev_reserve(c->acc, (size_t)n * 32));
  // M2L batches
  const size_t ng = pl.groups;
  const size_t item_blocks = (nm + 1023) / 1024;
  TRY(dev_reserve(c->g_gid, (size_t)pl.nslots * 4));
  TRY(dev_reserve(c->scan_tmp,
                  scan_tmp_bytes(std::max<int64_t>({(int64_t)pl.nslots, (int64_t)pl.ntiles,
                                                    (int64_t)item_blocks}))));
  TRY(dev_reserve(c->g_y64, ng * 112 * 8));
  TRY(dev_reserve(c->g_y32, ng * 112 * 4));
  TRY(dev_reserve(c->g_dl, ng));
  TRY(dev_reserve(c->g_lt, ng));
  TRY(dev_reserve(c->g_af, m2l_afrag_bytes((uint32_t)ng, mixed)));
  TRY(dev_reserve(c->gkeys, nm * 4));
  TRY(dev_reserve(c->gvals, nm * 4));
  TRY(dev_reserve(c->gkeys_alt, nm * 4));
  TRY(dev_reserve(c->gvals_alt, nm * 4));
  TRY(dev_reserve(c->fkeys, nm * 8));
  TRY(dev_reserve(c->fvals, nm * 4));
  TRY(dev_reserve(c->fkeys_alt, nm * 8));
  TRY(dev_reserve(c->fvals_alt, nm * 4));
  TRY(dev_reserve(c->g_cnt, 16));
  TRY(dev_reserve(c->it_flag, nm * 4));
  TRY(dev_reserve(c->it_iflag, item_blocks * 4));
  TRY(dev_reserve(c->it_iexcl, item_blocks * 4));
  TRY(dev_reserve(c->items, (size_t)pl.caps.items * 16));
  TRY(dev_reserve(c->cols, (size_t)pl.caps.items * 16 * 8));
  TRY(dev_reserve(c->tile_off, (size_t)(pl.ntiles + 1) * 4));
  TRY(dev_reserve(c->u_cnt, (size_t)pl.ntiles * 4));
  TRY(dev_reserve(c->u_off, (size_t)pl.ntiles * 4));
  TRY(dev_reserve(c->units, (size_t)pl.caps.units * 8));
  TRY(dev_reserve(c->u_part, m2l_unit_part_bytes(pl.caps.units, pl.ntiles)));
  TRY(dev_reserve(c->fb_off, (cells + 1) * 4));
  TRY(dev_reserve(c->fb_src, (size_t)pl.caps.fallback * 4 + 4));
  TRY(dev_reserve(c->fb_part, m2l_pairs_scratch_bytes((int32_t)pl.caps.fallback)));
  if (!c->codes_ready) {
    TRY(dev_reserve(c->codes, m2l_codes_bytes()));
    const hipError_t e = m2l_upload_codes(dp<uint32_t>(c->codes), s);
    if (e == hipErrorInvalidValue) return fail(c, TFMM_EHIP, "M2L operator table is inconsistent");
    TRY(e);
    c->codes_ready = true;
  }
  // expansions and fields
  const size_t ebytes = cells * NTERM * 16;
  TRY(dev_reserve(c->M, ebytes));
  TRY(dev_reserve(c->L, ebytes));
  if (!c->tops_ready) {
    TRY(dev_reserve(c->tops, translation_ops_doubles() * 8));
    TRY(launch_translation_ops(dp<double>(c->tops), s));
    c->tops_ready = true;
  }
  TRY(dev_reserve(c->mq, cells * 112 * (mixed ? 4 : 8)));
  TRY(dev_reserve(c->trg, (size_t)n * 32));
  TRY(dev_reserve(c->out, (size_t)n * 32));
  TRY(dev_reserve(c->leaf_of, (size_t)n * 4));
  // the graph's own words
  TRY(dev_reserve(c->gsz, SZ_WORDS * 4));
  TRY(dev_reserve(c->gstep, (size_t)pl.nsteps * 4));
  TRY(hipMemcpyAsync(c->gstep.ptr, pl.step_grid.data(), (size_t)pl.nsteps * 4,
                     hipMemcpyHostToDevice, s));
  if (!c->h_plan) TRY(hipHostMalloc((void **)&c->h_plan, PLAN_HOST_BYTES, 0));
  TRY(hipStreamSynchronize(s));
  return TFMM_OK;
}

// The evaluation with the plan's capacities and the device's sizes, enqueued
// on the context stream (recorded into the graph by the caller).
int whole_record(tfmm_ctx *c) {
  Plan &pl = c->plan;
  hipStream_t s = c->stream;
  const int64_t n = c->n;
  const int np = c->nprob;
  const bool forest = np > 1;
  const bool mixed = pl.prm.precision == TFMM_MIXED;
  const Cells k = cells_of(c);
  const int32_t cells = (int32_t)pl.cells;
  uint32_t *sz = dp<uint32_t>(c->gsz);
  TRY(launch_fill_zero(c->gsz.ptr, SZ_WORDS * 4, s));
  // bodies to sorted keys
  double *region = dp<double>(c->region);
  const int64_t *off = forest ? dp<int64_t>(c->d_prob_off) : nullptr;
  if (forest) {
    TRY(launch_bbox_forest(c->bx, c->by, c->bz, off, np, region, s));
  } else {
    const int nparts = (int)std::min<int64_t>(1024, (n + 255) / 256);
    TRY(launch_bbox(c->bx, c->by, c->bz, n, dp<double>(c->bbox_part), nparts, region, s));
  }
  TRY(launch_keys(c->bx, c->by, c->bz, c->bq, n, region, off, np, dp<uint64_t>(c->keys),
                  dp<uint32_t>(c->idx), forest ? dp<uint32_t>(c->bprob) : nullptr,
                  dp<double4>(c->brec), s));
  bool alt = false;
  TRY(radix_sort_pairs(dp<uint64_t>(c->keys), dp<uint32_t>(c->idx), dp<uint64_t>(c->keys_alt),
                       dp<uint32_t>(c->idx_alt), n, DEPTHBITS, DEPTHBITS + 3 * MAXDEPTH,
                       c->sort_tmp.ptr, s, &alt));
  const uint32_t *sprob = nullptr;
  if (forest) {
    const uint64_t *k1 = dp<uint64_t>(alt ? c->keys_alt : c->keys);
    const uint32_t *o1 = dp<uint32_t>(alt ? c->idx_alt : c->idx);
    TRY(launch_problem_keys(o1, dp<uint32_t>(c->bprob), n, dp<uint32_t>(c->pkey),
                            dp<uint32_t>(c->ppos), s));
    bool palt = false;
    TRY(radix_sort_pairs32(dp<uint32_t>(c->pkey), dp<uint32_t>(c->ppos), dp<uint32_t>(c->pkey_alt),
                           dp<uint32_t>(c->ppos_alt), n, 0, bits_for(np - 1), c->sort_tmp.ptr, s,
                           &palt));
    sprob = dp<uint32_t>(palt ? c->pkey_alt : c->pkey);
    TRY(launch_problem_gather(dp<uint32_t>(palt ? c->ppos_alt : c->ppos), k1, o1, n,
                              dp<uint64_t>(alt ? c->keys : c->keys_alt),
                              dp<uint32_t>(alt ? c->idx : c->idx_alt), s));
    alt = !alt;
  }
  pl.keys_alt = alt;
  const uint64_t *skeys = dp<uint64_t>(alt ? c->keys_alt : c->keys);
  TRY(launch_permute(dp<uint32_t>(alt ? c->idx_alt : c->idx), dp<double4>(c->brec), n,
                     dp<double>(c->sx), dp<double>(c->sy), dp<double>(c->sz), dp<double>(c->sq),
                     dp<int32_t>(c->perm), s));
  // cells, level by level, every size on the device
  int32_t *lev = dp<int32_t>(c->lev);
  TRY(launch_fill_zero(c->lev.ptr, (size_t)2 * (pl.L + 1) * 4, s));
  TRY(launch_fill_zero(c->lev_status.ptr, 4, s));
  TRY(launch_roots(k, n, off, np, region, lev, s));
  for (int d = 0; d < pl.L; d++)
    TRY(launch_level_step(k, lev, d, pl.grid[d], pl.prm.ncrit, cells, dp<uint32_t>(c->flag),
                          dp<uint32_t>(c->slot), dp<uint32_t>(c->lev_bsum),
                          dp<uint32_t>(c->lev_status), skeys, region, sprob, s, d == pl.L - 1));
  TRY(launch_tree_final(lev, pl.L, dp<uint32_t>(c->lev_status), m2l_tile_size(), sz, s));
  TRY(launch_pad_cells(k, sz, cells, s));
  TRY(rec(c, EV_TREE, s));
  TRY(rec(c, EV_DTT0, s));
  // traversal
  uint32_t *cnt = dp<uint32_t>(c->cnt);
  uint32_t *used = dp<uint32_t>(c->g_used);
  TRY(launch_dtt_cells(k, cells, pl.prm.theta, dp<double4>(c->dtt_geo), dp<int2>(c->dtt_topo),
                       s));
  TRY(launch_fill_zero(used, (size_t)(pl.nslots + 1) * 4, s));
  TRY(launch_fill_zero(cnt, (size_t)pl.cwords * 4, s));
  TRY(launch_root_pairs(dp<int2>(c->front_a), np, cnt + 16, s));
  DevBuf *fa = &c->front_a, *fb = &c->front_b;
  for (int st = 0; st < pl.nsteps; st++) {
    TRY(launch_dtt_step(dp<int2>(*fa), cnt + 16 + st, pl.step_grid[st], dp<int2>(*fb),
                        pl.caps.front, cnt + 16 + st + 1, dp<int2>(c->m2l_pairs), pl.caps.m2l,
                        dp<int2>(c->p2p_pairs), pl.caps.p2p, cnt, dp<double4>(c->dtt_geo),
                        dp<int2>(c->dtt_topo), k.depth, k.width, dp<int32_t>(c->slot_of), used,
                        used + pl.nslots, pl.prm.mutual, s));
    std::swap(fa, fb);
  }
  // P2P list: targets sorted, each target's sources sorted
  {
    const uint32_t npp = pl.caps.p2p;
    TRY(launch_split_pairs(dp<int2>(c->p2p_pairs), npp, dp<uint32_t>(c->lkeys),
                           dp<uint32_t>(c->lidx), s, cnt + 2));
    bool lalt = false;
    TRY(radix_sort_pairs32(dp<uint32_t>(c->lkeys), dp<uint32_t>(c->lidx),
                           dp<uint32_t>(c->lkeys_alt), dp<uint32_t>(c->lidx_alt), npp, 0,
                           pl.sbits, c->sort_tmp.ptr, s, &lalt, cnt + 2));
    TRY(launch_csr_sorted_sources(dp<uint32_t>(lalt ? c->lkeys_alt : c->lkeys),
                                  dp<uint32_t>(lalt ? c->lidx_alt : c->lidx), npp, cells,
                                  dp<int32_t>(c->p2p_off), dp<int32_t>(c->p2p_src),
                                  dp<uint32_t>(c->seg_long), s, cnt + 2));
  }
  if (mixed)
    TRY(launch_leaf_relpos(k, cells, sorted_bodies(c), dp<float4>(c->relpos), s));
  // schedules 2 and 3: the P2P branches off here, beside the M2L batching
  int rc;
  const bool early = c->p2p_schedule >= 2;
  if (early && (rc = fork_p2p(c, cells))) return rc;
  // M2L batches
  {
    const uint32_t nm = pl.caps.m2l;
    uint32_t *gcnt = dp<uint32_t>(c->g_cnt);
    TRY(scan_exclusive_u32(used, dp<uint32_t>(c->g_gid), pl.nslots, c->scan_tmp.ptr, s));
    TRY(launch_groups_count(dp<uint32_t>(c->g_gid), used, pl.nslots, pl.groups, sz, s));
    TRY(launch_m2l_group_harmonics(used, dp<uint32_t>(c->g_gid), pl.nslots,
                                   dp<double>(c->g_y64), dp<float>(c->g_y32),
                                   dp<int8_t>(c->g_dl), dp<int8_t>(c->g_lt), s, pl.groups));
    TRY(launch_fill_zero(gcnt, 16, s));
    TRY(launch_m2l_classify(dp<int2>(c->m2l_pairs), dp<int32_t>(c->slot_of), nm,
                            dp<uint32_t>(c->g_gid), pl.gbits, FB_SRC_BITS, dp<uint32_t>(c->gkeys),
                            dp<uint32_t>(c->gvals), dp<uint64_t>(c->fkeys),
                            dp<uint32_t>(c->fvals), gcnt, s, cnt + 1, pl.groups));
    bool galt = false;
    // every pair's key (the grouped ones sort first, gcnt[0] of them)
    TRY(radix_sort_pairs32(dp<uint32_t>(c->gkeys), dp<uint32_t>(c->gvals),
                           dp<uint32_t>(c->gkeys_alt), dp<uint32_t>(c->gvals_alt), nm, 0,
                           pl.tbits + pl.gbits, c->sort_tmp.ptr, s, &galt, cnt + 1));
    pl.gkeys_alt = galt;
    const uint32_t *gk = dp<uint32_t>(galt ? c->gkeys_alt : c->gkeys);
    const uint32_t *gv = dp<uint32_t>(galt ? c->gvals_alt : c->gvals);
    TRY(launch_m2l_count_items_dev(gk, nm, gcnt, dp<uint32_t>(c->it_flag),
                                   dp<uint32_t>(c->it_iflag), dp<uint32_t>(c->it_iexcl),
                                   c->scan_tmp.ptr, sz + SZ_NITEMS, s));
    TRY(launch_m2l_write_items(gk, gv, nm, pl.gbits, dp<uint32_t>(c->it_flag),
                               dp<uint32_t>(c->it_iexcl), dp<int8_t>(c->g_dl), dp<int4>(c->items),
                               dp<uint2>(c->cols), s, gcnt, pl.caps.items));
    TRY(launch_m2l_tile_offsets(dp<int4>(c->items), pl.caps.items, pl.ntiles,
                                dp<int32_t>(c->tile_off), s, sz + SZ_NITEMS));
    int32_t *chunk = reinterpret_cast<int32_t *>(sz + SZ_CHUNK);
    TRY(launch_m2l_unit_chunk(sz + SZ_NITEMS, sz + SZ_NTILES, c->m2l_unit_factor, chunk, s));
    TRY(launch_m2l_units(dp<int32_t>(c->tile_off), pl.ntiles, 0, dp<uint32_t>(c->u_cnt),
                         dp<uint32_t>(c->u_off), dp<int2>(c->units), c->scan_tmp.ptr, s,
                         pl.caps.units, chunk));
    // the pairs outside the groups (appended compactly): at most the
    // fallback capacity of them are sorted
    const uint32_t nfb = std::min<uint32_t>(nm, pl.caps.fallback);
    bool falt = false;
    TRY(radix_sort_pairs(dp<uint64_t>(c->fkeys), dp<uint32_t>(c->fvals), dp<uint64_t>(c->fkeys_alt),
                         dp<uint32_t>(c->fvals_alt), nfb, 0, FB_SRC_BITS + pl.sbits, c->sort_tmp.ptr, s,
                         &falt, gcnt + 1));
    const uint64_t *fk = dp<uint64_t>(falt ? c->fkeys_alt : c->fkeys);
    TRY(launch_csr_offsets(fk, nfb, FB_SRC_BITS, cells, dp<int32_t>(c->fb_off), s, gcnt + 1));
    TRY(launch_unpack_pairs(fk, nfb, FB_SRC_BITS, dp<int32_t>(c->fb_src), s, gcnt + 1));
    TRY(launch_m2l_group_afrag(mixed ? (const void *)c->g_y32.ptr : (const void *)c->g_y64.ptr,
                               pl.groups, dp<uint32_t>(c->codes), dp<int8_t>(c->g_dl),
                               c->g_af.ptr, mixed, s, sz + SZ_NG));
  }
  TRY(rec(c, EV_DTT1, s));
  // the numeric phase
  if (!early && c->p2p_schedule != 1 && (rc = fork_p2p(c, cells))) return rc;
  const size_t ebytes = (size_t)cells * NTERM * 16;
  TRY(launch_fill_zero(c->M.ptr, ebytes, s));
  TRY(launch_p2m(k, cells, sorted_bodies(c), dp<double>(c->M), s));
  for (int d = pl.L - 2; d >= 0; d--)
    TRY(launch_m2m(k, 0, pl.grid[d + 1] / 8 + 1, dp<double>(c->tops), dp<double>(c->M), s, lev,
                   d + 1));
  TRY(rec(c, EV_UP, s));
  TRY(launch_m2l_scale_m(k, cells, dp<double>(c->M), c->mq.ptr, mixed, s));
  TRY(launch_fill_zero(c->L.ptr, ebytes, s));
  TRY(rec(c, EV_M2L0, s));
  if (c->p2p_schedule == 1 && (rc = fork_p2p(c, cells))) return rc;
  TRY(launch_m2l_tiles(dp<int32_t>(c->tile_off), dp<int2>(c->units), dp<uint32_t>(c->u_cnt),
                       dp<uint32_t>(c->u_off), pl.ntiles, 0, c->m2l_xcd_run, pl.caps.units,
                       dp<uint2>(c->cols), c->g_af.ptr, c->mq.ptr, k, cells, dp<double>(c->L),
                       dp<double>(c->u_part), mixed, s,
                       reinterpret_cast<const int32_t *>(sz + SZ_CHUNK)));
  TRY(rec(c, EV_M2LT, s));
  if (pl.caps.fallback > 0) {
    Csr fl{dp<int32_t>(c->fb_off), dp<int32_t>(c->fb_src)};
    TRY(launch_m2l_fp64(k, cells, fl, (int32_t)pl.caps.fallback, dp<double>(c->M),
                        dp<double>(c->L), c->fb_part.ptr, s, dp<uint32_t>(c->g_cnt) + 1));
  }
  TRY(rec(c, EV_M2L1, s));
  if ((rc = drain_p2p(c, cells))) return rc;
  TRY(hipStreamWaitEvent(s, c->ev[EV_P2PJOIN], 0));
  TRY(rec(c, EV_TRAV, s));
  for (int d = 0; d + 1 < pl.L; d++)
    TRY(launch_l2l(k, 0, pl.grid[d + 1] / 8 + 1, dp<double>(c->tops), dp<double>(c->L), s, lev,
                   d + 1));
  TRY(launch_l2p(k, cells, sorted_bodies(c), dp<double>(c->L), dp<double>(c->acc),
                 dp<int32_t>(c->perm), dp<int32_t>(c->leaf_of), dp<double>(c->trg),
                 dp<double>(c->out), s));
  // capacities, then the read-back
  TRY(launch_plan_check(cnt, pl.nsteps, dp<uint32_t>(c->gstep), pl.caps, dp<uint32_t>(c->g_cnt),
                        dp<uint32_t>(c->u_cnt), dp<uint32_t>(c->u_off), pl.ntiles, sz, s));
  char *h = reinterpret_cast<char *>(c->h_plan);
  TRY(hipMemcpyAsync(h + PH_SZ * 4, sz, SZ_WORDS * 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h + PH_GCNT * 4, c->g_cnt.ptr, 8, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h + PH_CNT * 4, cnt, (size_t)pl.cwords * 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h + PH_LEV * 4, lev, (size_t)2 * (pl.L + 1) * 4, hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h + PH_U64_BYTE, static_cast<char *>(c->cnt.ptr) + 32, 16,
                     hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h + PH_REGION_BYTE, region, 48, hipMemcpyDeviceToHost, s));
  return TFMM_OK;
}

// reserve, record and instantiate the plan's graph
int whole_capture(tfmm_ctx *c) {
  Plan &pl = c->plan;
  if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
  pl.exec = nullptr;
  int rc = plan_reserve(c);
  if (rc) return rc;
  if (c->ext_events < 0) probe_external_events(c);
  hipStream_t s = c->stream;
  TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  c->capturing = true;
  rc = whole_record(c);
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
  const hipError_t e2 = hipGraphInstantiate(&pl.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    pl.exec = nullptr;
    return hipfail(c, e2, "hipGraphInstantiate");
  }
  pl.signature = buffers_signature(c);
  return TFMM_OK;
}

// the host's view of the evaluation the graph just ran
void whole_finish(tfmm_ctx *c) {
  const Plan &pl = c->plan;
  const uint32_t *h = c->h_plan;
  const uint32_t *hs = h + PH_SZ, *gc = h + PH_GCNT, *cn = h + PH_CNT;
  const int32_t *hl = reinterpret_cast<const int32_t *>(h + PH_LEV);
  c->ncells = hs[SZ_NCELLS];
  c->nleaves = hs[SZ_NLEAVES];
  c->lvl.assign(1, 0);
  c->tree_hint.clear();
  for (int d = 0; d < pl.L && hl[2 * d + 1] > 0; d++) {
    c->lvl.push_back(hl[2 * d] + hl[2 * d + 1]);
    c->tree_hint.push_back(hl[2 * d + 1]);
  }
  c->keys_in_alt = pl.keys_alt;
  c->grouped_keys_in_alt = pl.gkeys_alt;
  c->n_m2l_exec = cn[1];
  c->n_p2p = cn[2];
  c->n_m2l_all = cn[3];
  c->n_p2p_calls_leaf = cn[4];
  c->dtt_steps = 0;
  for (int st = 0; st <= pl.nsteps; st++)
    if (cn[16 + st]) c->dtt_steps = st + 1;
  memcpy(c->h_u64, reinterpret_cast<const char *>(h) + PH_U64_BYTE, 16);
  memcpy(c->h_region, reinterpret_cast<const char *>(h) + PH_REGION_BYTE, 48);
  c->n_groups = hs[SZ_NG];
  c->n_grouped = gc[0];
  c->n_fallback = gc[1];
  c->n_items = hs[SZ_NITEMS];
  c->unit_chunk = (int32_t)hs[SZ_CHUNK];
  c->units_bound = pl.caps.units;
}

// One launch of the plan's graph (recorded again when a buffer moved).
// 0: done; 1: a capacity was exceeded (the caller evaluates the ordinary
// way); < 0: error.
int whole_run(tfmm_ctx *c, const tfmm_params *p) {
  TRY(hipSetDevice(c->device));
  c->prm = *p;
  Plan &pl = c->plan;
  if (!pl.exec || pl.signature != buffers_signature(c)) {
    const int rc = whole_capture(c);
    if (rc) return rc;
  }
  hipStream_t s = c->stream;
  c->stage = 0;
  TRY(hipEventRecord(c->ev[EV_BEGIN], s));
  if (c->ext_events != 1) TRY(hipEventRecord(c->ev[EV_TREE], s));  // no inner times
  TRY(hipGraphLaunch(pl.exec, s));
  TRY(hipEventRecord(c->ev[EV_DOWN], s));
  TRY(hipStreamSynchronize(s));
  if (c->h_plan[PH_SZ + SZ_STATUS] != 0) {
    plan_drop(c);
    return 1;
  }
  whole_finish(c);
  c->one_graph = true;
  c->fused = true;
  c->inner_timed = c->ext_events == 1;
  c->stage = 4;
  return TFMM_OK;
}

int need_stage(tfmm_ctx *c, int st, const char *what) {
  if (!c) return TFMM_EINVAL;
  if (c->stage < st) return fail(c, TFMM_ESTATE, std::string(what) + " called out of order");
  return TFMM_OK;
}

}  // namespace

extern "C" {

namespace {

// Scratch of tfmm_direct_noalloc: the target indices (8 bytes each), then the
// fields (32 bytes each) of up to TFMM_DIRECT_NOALLOC_MAX targets.  Host code
// uses it only through the per-device address hipGetSymbolAddress gives.
constexpr size_t NOALLOC_IDX_BYTES = (size_t)TFMM_DIRECT_NOALLOC_MAX * 8;
constexpr size_t NOALLOC_BYTES = (size_t)TFMM_DIRECT_NOALLOC_MAX * 40;
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES] = {};

// the scratch's address on the current device (looked up once per device)
hipError_t scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

}  // namespace

int tfmm_create(int hip_device, tfmm_ctx **out) {
  if (!out) return TFMM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TFMM_EHIP;
  if (hip_device < 0 || hip_device >= ndev) return TFMM_EINVAL;
  if (hipSetDevice(hip_device) != hipSuccess) return TFMM_EHIP;
  tfmm_ctx *c = new tfmm_ctx();
  c->device = hip_device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TFMM_EHIP;
  }
  if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, hip_device) !=
          hipSuccess ||
      dev_reserve(c->p2p_next, 256) != hipSuccess) {
    delete c;
    return TFMM_EHIP;
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
    tfmm_destroy(c);
    return TFMM_EHIP;
  }
  for (int i = 0; i < EV_COUNT; i++) {
    if (hipEventCreate(&c->ev[i]) != hipSuccess) {
      tfmm_destroy(c);
      return TFMM_EHIP;
    }
  }
  if (hipHostMalloc((void **)&c->h_u32, H_U32_WORDS * 4, 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_u64, 64, 0) != hipSuccess) {
    tfmm_destroy(c);
    return TFMM_ENOMEM;
  }
  if (scratch_address(hip_device, &c->noalloc) != hipSuccess) {
    tfmm_destroy(c);
    return TFMM_EHIP;
  }
  *out = c;
  return TFMM_OK;
}

void tfmm_destroy(tfmm_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (DevBuf *b : all_buffers(c)) dev_release(*b);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->plan.exec) (void)hipGraphExecDestroy(c->plan.exec);
  if (c->h_plan) (void)hipHostFree(c->h_plan);
  for (int i = 0; i < EV_COUNT; i++)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->h_u32) (void)hipHostFree(c->h_u32);
  if (c->h_u64) (void)hipHostFree(c->h_u64);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *tfmm_last_error(const tfmm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int tfmm_set_periodic(tfmm_ctx *c, int images, double cycle) {
  if (!c) return TFMM_EINVAL;
  if (images < 0 || images > TFMM_PERIODIC_MAX_IMAGES)
    return fail(c, TFMM_EINVAL, "images must be in [0, TFMM_PERIODIC_MAX_IMAGES]");
  if (images > 0 && !(cycle > 0 && cycle < 1e300))
    return fail(c, TFMM_EINVAL, "the period must be a positive finite number");
  c->images = images;
  c->cycle = images > 0 ? cycle : 0.0;
  c->stage = 0;  // a new tree and traversal
  plan_drop(c);
  return TFMM_OK;
}

int tfmm_set_basis(tfmm_ctx *c, int basis) {
  if (!c) return TFMM_EINVAL;
  if (basis != TFMM_BASIS_SPHERICAL && basis != TFMM_BASIS_CARTESIAN)
    return fail(c, TFMM_EINVAL, "basis must be TFMM_BASIS_SPHERICAL or TFMM_BASIS_CARTESIAN");
  if (basis != c->basis) {
    c->basis = basis;
    if (c->stage > 1) c->stage = 1;  // the tree stands; expansions are made again
    plan_drop(c);
  }
  return TFMM_OK;
}

int tfmm_get_basis(tfmm_ctx *c, int *basis) {
  if (!c || !basis) return TFMM_EINVAL;
  *basis = c->basis;
  return TFMM_OK;
}

int tfmm_set_tuning(tfmm_ctx *c, int knob, double value) {
  if (!c) return TFMM_EINVAL;
  switch (knob) {
    case TFMM_TUNE_M2L_UNIT:
      if (!(value >= 0) || !(value <= 1e6)) return fail(c, TFMM_EINVAL, "unit factor out of range");
      c->m2l_unit_factor = value;
      return TFMM_OK;
    case TFMM_TUNE_M2L_XCD_RUN:
      if (!(value >= 1) || !(value <= 4096) || value != std::floor(value))
        return fail(c, TFMM_EINVAL, "XCD run must be an integer in [1, 4096]");
      c->m2l_xcd_run = (int)value;
      return TFMM_OK;
    case TFMM_TUNE_P2P_SCHEDULE:
      if (value != 0 && value != 1 && value != 2 && value != 3 && value != 4)
        return fail(c, TFMM_EINVAL, "P2P schedule must be 0 ... 4");
      c->p2p_schedule = (int)value;
      return TFMM_OK;
    case TFMM_TUNE_GRAPH:
      if (value != 0 && value != 1 && value != 2)
        return fail(c, TFMM_EINVAL, "graph must be 0, 1 or 2");
      c->use_graph = (int)value;
      return TFMM_OK;
    case TFMM_TUNE_P2P_WAVES:
      if (!(value >= 0) || !(value <= 8) || value != std::floor(value))
        return fail(c, TFMM_EINVAL, "P2P waves per SIMD must be an integer in [0, 8]");
      c->p2p_waves = (int)value;
      return TFMM_OK;
    case TFMM_TUNE_DEVICE_LIMIT:
      if (!(value >= 0) || !(value < 9.0e18) || value != std::floor(value))
        return fail(c, TFMM_EINVAL, "device limit must be a whole number of bytes >= 0");
      dev_set_limit((size_t)value);
      return TFMM_OK;
    case TFMM_TUNE_CART_M2L:
      if (value != 0 && value != 1) return fail(c, TFMM_EINVAL, "Cartesian M2L route must be 0 or 1");
      if ((int)value != c->cart_direct) {
        c->cart_direct = (int)value;
        if (cartesian(c) && c->stage > 1) c->stage = 1;  // the expansions are made again
      }
      return TFMM_OK;
    default:
      return fail(c, TFMM_EINVAL, "unknown tuning knob");
  }
}

// Candidates of the tile-kernel knobs, tried jointly; the evaluation's time
// (HIP events around it) decides, median over the repetitions.  The P2P
// schedule stays as set: running the P2P beside the tiles measured within
// noise of running it beside the batch building, and it would fold the
// P2P's time into the tile kernel's.
// Candidates of the kernel knobs, timed in two stages: the tile kernel's
// (unit factor x XCD run) at the current P2P setting, then where the P2P
// starts and how many resident waves it holds, at the best tile setting
// (running the P2P from phase A with a few resident waves lets the upward
// pass and the tiles have CU slots beside it: on the 4M Plummer 12.96 ms
// against 13.82 with the P2P taking every slot).
int tfmm_autotune(tfmm_ctx *c, const tfmm_params *p, int reps, double *chosen) {
  if (!c) return TFMM_EINVAL;
  if (reps < 1 || reps > 64) return fail(c, TFMM_EINVAL, "reps must be in [1, 64]");
  if (!p) return fail(c, TFMM_EINVAL, "params is null");
  static const double kUnit[] = {0.0, 1.5, 2.0, 3.0};
  static const int kRun[] = {8, 16, 32, 64};
  const double unit0 = c->m2l_unit_factor;
  const int run0 = c->m2l_xcd_run, sched0 = c->p2p_schedule, waves0 = c->p2p_waves;
  struct Cand { double unit; int run, sched, waves; std::vector<double> ms; };
  auto restore = [&] {
    c->m2l_unit_factor = unit0;
    c->m2l_xcd_run = run0;
    c->p2p_schedule = sched0;
    c->p2p_waves = waves0;
  };
  auto apply = [&](const Cand &k) {
    c->m2l_unit_factor = k.unit;
    c->m2l_xcd_run = k.run;
    c->p2p_schedule = k.sched;
    c->p2p_waves = k.waves;
  };
  // Each candidate is evaluated as a block: AT_SETTLE evaluations that are
  // not counted, then reps counted ones.  The time of a setting depends on
  // the few evaluations before it (measured: the P2P schedule 2 with 3
  // resident waves took 12.5, 11.9, then 11.7 ms after a switch, and 13.0
  // when the candidates were interleaved), so interleaving would rank the
  // settings by their neighbours.
  constexpr int AT_SETTLE = 2;
  auto run = [&](std::vector<Cand> &cand, size_t &best, double &best_ms) -> int {
    for (Cand &k : cand) {
      apply(k);
      for (int rep = -AT_SETTLE; rep < reps; rep++) {
        int rc = tfmm_evaluate(c, p);
        if (rc) return rc;
        if (rep < 0) continue;
        tfmm_stats_t st;
        rc = tfmm_get_stats(c, &st);
        if (rc) return rc;
        k.ms.push_back(st.ms_total);
      }
    }
    for (size_t i = 0; i < cand.size(); i++) {
      std::vector<double> v = cand[i].ms;
      std::sort(v.begin(), v.end());
      const double med = v[v.size() / 2];
      if (med < best_ms) {
        best_ms = med;
        best = i;
      }
    }
    return TFMM_OK;
  };
  int rc = tfmm_evaluate(c, p);  // sizes the buffers
  if (rc) {
    restore();
    return rc;
  }
  std::vector<Cand> tiles;
  for (double u : kUnit)
    for (int r : kRun) tiles.push_back({u, r, sched0, waves0, {}});
  size_t bt = 0;
  double bt_ms = 1e300;
  if ((rc = run(tiles, bt, bt_ms))) {
    restore();
    return rc;
  }
  // the P2P's start and resident waves (schedule 2 to evaluations without
  // a graph)
  std::vector<Cand> p2p;
  // schedules 0-2 with 0, 3 or 4 resident waves (schedule 3 with 2 waves
  // measured level with schedule 2 with 3 on the 4M Plummer, 11.7 ms, and
  // 5 % slower on the 8M two-cluster problem, where the autotuner took it
  // once on noise: it stays a knob, not a candidate)
  struct SW { int sched, waves; };
  static const SW kP2P[] = {{0, 0}, {0, 3}, {0, 4}, {1, 0}, {1, 3},
                            {1, 4}, {2, 0}, {2, 3}, {2, 4}};
  for (const SW &k : kP2P) {
    if (k.sched >= 2 && c->use_graph == 1) continue;  // the early start needs phase A
    if (k.sched == sched0 && k.waves == waves0) continue;  // timed in the first stage
    p2p.push_back({tiles[bt].unit, tiles[bt].run, k.sched, k.waves, {}});
  }
  size_t bp = 0;
  double bp_ms = bt_ms;
  bool p2p_better = false;
  {
    double m = bt_ms;
    if ((rc = run(p2p, bp, m))) {
      restore();
      return rc;
    }
    p2p_better = m < bt_ms;
    bp_ms = m;
  }
  const Cand &w = p2p_better ? p2p[bp] : tiles[bt];
  apply(w);
  if (chosen) {
    chosen[0] = w.unit;
    chosen[1] = w.run;
    chosen[2] = w.sched;
    chosen[3] = p2p_better ? bp_ms : bt_ms;
  }
  return TFMM_OK;
}

int tfmm_get_tuning(tfmm_ctx *c, int knob, double *value) {
  if (!c || !value) return TFMM_EINVAL;
  switch (knob) {
    case TFMM_TUNE_M2L_UNIT: *value = c->m2l_unit_factor; return TFMM_OK;
    case TFMM_TUNE_M2L_XCD_RUN: *value = c->m2l_xcd_run; return TFMM_OK;
    case TFMM_TUNE_P2P_SCHEDULE: *value = c->p2p_schedule; return TFMM_OK;
    case TFMM_TUNE_GRAPH: *value = c->use_graph; return TFMM_OK;
    case TFMM_TUNE_P2P_WAVES: *value = c->p2p_waves; return TFMM_OK;
    case TFMM_TUNE_DEVICE_LIMIT: *value = (double)dev_limit(); return TFMM_OK;
    case TFMM_TUNE_CART_M2L: *value = c->cart_direct; return TFMM_OK;
    default: return fail(c, TFMM_EINVAL, "unknown tuning knob");
  }
}

void *tfmm_stream(tfmm_ctx *c) { return c ? (void *)c->stream : nullptr; }

int tfmm_sync(tfmm_ctx *c) {
  if (!c) return TFMM_EINVAL;
  TRY(hipStreamSynchronize(c->stream));
  return TFMM_OK;
}

namespace {

int check_desc(tfmm_ctx *c, const void *host, const tfmm_body_desc *d) {
  if (!d || !host) return fail(c, TFMM_EINVAL, "null bodies or descriptor");
  if (d->n < 1 || d->n > INT32_MAX) return fail(c, TFMM_EINVAL, "n must be in [1, 2^31)");
  if (d->stride < 24 || d->off_x > d->stride - 24 || d->off_q > d->stride - 8)
    return fail(c, TFMM_EINVAL, "body descriptor offsets exceed the stride");
  return TFMM_OK;
}

// the caller's array-of-structs bodies into SoA host arrays, from index at
int unpack_bodies(tfmm_ctx *c, const void *host, const tfmm_body_desc *d, int64_t at,
                  std::vector<double> &x, std::vector<double> &y, std::vector<double> &z,
                  std::vector<double> &q) {
  const char *base = reinterpret_cast<const char *>(host);
  for (int64_t i = 0; i < d->n; i++) {
    const char *b = base + (size_t)i * d->stride;
    double X[3], Q;
    memcpy(X, b + d->off_x, 24);
    memcpy(&Q, b + d->off_q, 8);
    if (!std::isfinite(X[0]) || !std::isfinite(X[1]) || !std::isfinite(X[2]) || !std::isfinite(Q))
      return fail(c, TFMM_EINVAL, "non-finite body coordinate or charge");
    x[at + i] = X[0];
    y[at + i] = X[1];
    z[at + i] = X[2];
    q[at + i] = Q;
  }
  return TFMM_OK;
}

int upload_bodies(tfmm_ctx *c, int64_t n, const std::vector<double> &x,
                  const std::vector<double> &y, const std::vector<double> &z,
                  const std::vector<double> &q) {
  TRY(dev_reserve(c->own_x, n * 8));
  TRY(dev_reserve(c->own_y, n * 8));
  TRY(dev_reserve(c->own_z, n * 8));
  TRY(dev_reserve(c->own_q, n * 8));
  TRY(hipMemcpyAsync(c->own_x.ptr, x.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  TRY(hipMemcpyAsync(c->own_y.ptr, y.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  TRY(hipMemcpyAsync(c->own_z.ptr, z.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  TRY(hipMemcpyAsync(c->own_q.ptr, q.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  c->n = n;
  c->bx = dp<double>(c->own_x);
  c->by = dp<double>(c->own_y);
  c->bz = dp<double>(c->own_z);
  c->bq = dp<double>(c->own_q);
  c->stage = 0;
  return TFMM_OK;
}

// problem offsets of a batch onto the device
int set_problems(tfmm_ctx *c, int nprob, const int64_t *off) {
  c->nprob = nprob;
  c->prob_off.assign(off, off + nprob + 1);
  if (nprob > 1) {
    TRY(dev_reserve(c->d_prob_off, (size_t)(nprob + 1) * 8));
    TRY(hipMemcpyAsync(c->d_prob_off.ptr, c->prob_off.data(), (size_t)(nprob + 1) * 8,
                       hipMemcpyHostToDevice, c->stream));
    TRY(hipStreamSynchronize(c->stream));
  }
  return TFMM_OK;
}

}  // namespace

int tfmm_set_bodies(tfmm_ctx *c, const void *host, const tfmm_body_desc *d) {
  if (!c) return TFMM_EINVAL;
  int rc = check_desc(c, host, d);
  if (rc) return rc;
  TRY(hipSetDevice(c->device));
  const int64_t n = d->n;
  std::vector<double> x(n), y(n), z(n), q(n);
  rc = unpack_bodies(c, host, d, 0, x, y, z, q);
  if (rc) return rc;
  rc = upload_bodies(c, n, x, y, z, q);
  if (rc) return rc;
  const int64_t off[2] = {0, n};
  return set_problems(c, 1, off);
}

int tfmm_set_batch(tfmm_ctx *c, int nprob, const tfmm_body_desc *descs,
                   const void *const *bodies) {
  if (!c) return TFMM_EINVAL;
  if (nprob < 1 || nprob > TFMM_MAX_PROBLEMS)
    return fail(c, TFMM_EINVAL, "nprob must be in [1, TFMM_MAX_PROBLEMS]");
  if (!descs || !bodies) return fail(c, TFMM_EINVAL, "null descriptors or bodies");
  std::vector<int64_t> off(nprob + 1, 0);
  for (int p = 0; p < nprob; p++) {
    int rc = check_desc(c, bodies[p], &descs[p]);
    if (rc) return rc;
    off[p + 1] = off[p] + descs[p].n;
  }
  if (off[nprob] > INT32_MAX) return fail(c, TFMM_EINVAL, "total body count must be < 2^31");
  TRY(hipSetDevice(c->device));
  const int64_t n = off[nprob];
  std::vector<double> x(n), y(n), z(n), q(n);
  for (int p = 0; p < nprob; p++) {
    int rc = unpack_bodies(c, bodies[p], &descs[p], off[p], x, y, z, q);
    if (rc) return rc;
  }
  int rc = upload_bodies(c, n, x, y, z, q);
  if (rc) return rc;
  return set_problems(c, nprob, off.data());
}

// device-resident bodies: every coordinate and charge must be finite, as the
// host path requires (one small reduction and a read-back per set call)
int check_device_bodies(tfmm_ctx *c, int64_t n, const double *x, const double *y,
                        const double *z, const double *q) {
  TRY(dev_reserve(c->finite_flag, 4));
  TRY(launch_fill_zero(c->finite_flag.ptr, 16, c->stream));
  TRY(launch_check_finite(x, y, z, q, n, dp<uint32_t>(c->finite_flag), c->stream));
  TRY(hipMemcpyAsync(c->h_u32, c->finite_flag.ptr, 4, hipMemcpyDeviceToHost, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  if (c->h_u32[0]) return fail(c, TFMM_EINVAL, "a body coordinate or charge is not finite");
  return TFMM_OK;
}

int tfmm_set_batch_device(tfmm_ctx *c, int nprob, const int64_t *offsets, const double *x,
                          const double *y, const double *z, const double *q) {
  if (!c) return TFMM_EINVAL;
  if (nprob < 1 || nprob > TFMM_MAX_PROBLEMS)
    return fail(c, TFMM_EINVAL, "nprob must be in [1, TFMM_MAX_PROBLEMS]");
  if (!offsets) return fail(c, TFMM_EINVAL, "null offsets");
  if (!x || !y || !z || !q) return fail(c, TFMM_EINVAL, "null body array");
  if (offsets[0] != 0) return fail(c, TFMM_EINVAL, "offsets[0] must be 0");
  for (int p = 0; p < nprob; p++)
    if (offsets[p + 1] <= offsets[p]) return fail(c, TFMM_EINVAL, "every problem needs a body");
  if (offsets[nprob] > INT32_MAX) return fail(c, TFMM_EINVAL, "total body count must be < 2^31");
  TRY(hipSetDevice(c->device));
  int rc = check_device_bodies(c, offsets[nprob], x, y, z, q);
  if (rc) return rc;
  c->n = offsets[nprob];
  c->bx = x;
  c->by = y;
  c->bz = z;
  c->bq = q;
  c->stage = 0;
  return set_problems(c, nprob, offsets);
}

int tfmm_batch_evaluate(tfmm_ctx *c, int nprob, const tfmm_body_desc *descs,
                        const void *const *bodies, const tfmm_params *p) {
  int rc = tfmm_set_batch(c, nprob, descs, bodies);
  if (rc) return rc;
  return tfmm_evaluate(c, p);
}

int tfmm_num_problems(const tfmm_ctx *c, int64_t *offsets) {
  if (!c) return TFMM_EINVAL;
  if (offsets) memcpy(offsets, c->prob_off.data(), c->prob_off.size() * 8);
  return c->nprob;
}

int tfmm_set_bodies_device(tfmm_ctx *c, int64_t n, const double *x, const double *y,
                           const double *z, const double *q) {
  if (!c) return TFMM_EINVAL;
  if (n < 1 || n > INT32_MAX) return fail(c, TFMM_EINVAL, "n must be in [1, 2^31)");
  if (!x || !y || !z || !q) return fail(c, TFMM_EINVAL, "null body array");
  TRY(hipSetDevice(c->device));
  int rc = check_device_bodies(c, n, x, y, z, q);
  if (rc) return rc;
  c->n = n;
  c->bx = x;
  c->by = y;
  c->bz = z;
  c->bq = q;
  c->stage = 0;
  const int64_t off[2] = {0, n};
  return set_problems(c, 1, off);
}

int tfmm_build_tree(tfmm_ctx *c, const tfmm_params *p) {
  if (!c) return TFMM_EINVAL;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (c->n < 1) return fail(c, TFMM_ESTATE, "no bodies set");
  if (c->images > 0 && c->nprob > 1)
    return fail(c, TFMM_EINVAL, "periodic images apply to one problem, not a batch");
  if (c->images > 0 && p->mutual)
    return fail(c, TFMM_EINVAL, "periodic images are traversed with mutual = 0");
  TRY(hipSetDevice(c->device));
  c->prm = *p;
  c->stage = 0;
  c->one_graph = false;
  TRY(hipEventRecord(c->ev[EV_BEGIN], c->stream));
  rc = do_build_tree(c);
  if (rc) return rc;
  if (c->images > 0) {
    TRY(hipMemcpyAsync(c->h_u32, c->pflag.ptr, 4, hipMemcpyDeviceToHost, c->stream));
    TRY(hipStreamSynchronize(c->stream));
    if (c->h_u32[0]) return fail(c, TFMM_EINVAL, "the bodies span more than one period");
  }
  TRY(hipEventRecord(c->ev[EV_TREE], c->stream));
  c->stage = 1;
  return TFMM_OK;
}

int tfmm_set_params(tfmm_ctx *c, const tfmm_params *p) {
  if (!c) return TFMM_EINVAL;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (c->stage >= 1 && p->ncrit != c->prm.ncrit)
    return fail(c, TFMM_EINVAL, "ncrit changes the tree: call tfmm_build_tree");
  c->prm = *p;
  if (c->stage > 2) c->stage = 2;  // the traversal has to be redone
  return TFMM_OK;
}

int tfmm_upward(tfmm_ctx *c) {
  int rc = need_stage(c, 1, "tfmm_upward");
  if (rc) return rc;
  c->fused = false;
  c->inner_timed = true;
  rc = do_upward(c);
  if (rc) return rc;
  TRY(hipEventRecord(c->ev[EV_UP], c->stream));
  c->stage = 2;
  return TFMM_OK;
}

int tfmm_traverse(tfmm_ctx *c) {
  int rc = need_stage(c, 2, "tfmm_traverse");
  if (rc) return rc;
  rc = do_traverse(c);
  if (rc) return rc;
  TRY(hipEventRecord(c->ev[EV_TRAV], c->stream));
  c->stage = 3;
  return TFMM_OK;
}

int tfmm_traverse_lists(tfmm_ctx *c, const int32_t *m2l, int64_t n_m2l, const int32_t *p2p,
                        int64_t n_p2p) {
  int rc = need_stage(c, 2, "tfmm_traverse_lists");
  if (rc) return rc;
  if (c->nprob > 1) return fail(c, TFMM_EINVAL, "tfmm_traverse_lists: one problem, not a batch");
  rc = traverse_given(c, m2l, n_m2l, p2p, n_p2p);
  if (rc) return rc;
  TRY(hipEventRecord(c->ev[EV_TRAV], c->stream));
  c->stage = 3;
  return TFMM_OK;
}

int tfmm_downward(tfmm_ctx *c) {
  int rc = need_stage(c, 3, "tfmm_downward");
  if (rc) return rc;
  rc = do_downward(c);
  if (rc) return rc;
  TRY(hipEventRecord(c->ev[EV_DOWN], c->stream));
  c->stage = 4;
  return TFMM_OK;
}

int tfmm_evaluate(tfmm_ctx *c, const tfmm_params *p) {
  const bool whole = c && c->use_graph == 2 && c->images == 0 && !cartesian(c);
  if (whole) {
    int rc = check_params(c, p);
    if (rc) return rc;
    if (c->n < 1) return fail(c, TFMM_ESTATE, "no bodies set");
    if (plan_fits(c, p)) {
      rc = whole_run(c, p);
      if (rc <= 0) return rc;
    }
  }
  int rc = tfmm_build_tree(c, p);
  if (rc) return rc;
  c->fused = true;
  c->one_graph = false;
  rc = fused_prepare(c);
  if (rc) return rc;
  const bool graph = c->use_graph == 1;
  rc = graph ? fused_numeric_graph(c) : fused_numeric(c);
  if (rc) return rc;
  c->inner_timed = !graph || c->ext_events == 1;
  c->stage = 4;
  if (whole) plan_make(c);  // from this evaluation's sizes (already on the host)
  return TFMM_OK;
}

int tfmm_get_fields(tfmm_ctx *c, double *dst, int order, int on_device) {
  int rc = need_stage(c, 4, "tfmm_get_fields");
  if (rc) return rc;
  if (!dst) return fail(c, TFMM_EINVAL, "null destination");
  const void *src = order == 0 ? c->out.ptr : c->trg.ptr;
  TRY(hipMemcpyAsync(dst, src, (size_t)c->n * 32,
                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return TFMM_OK;
}

int tfmm_get_stats(tfmm_ctx *c, tfmm_stats_t *o) {
  if (!c || !o) return TFMM_EINVAL;
  TRY(hipStreamSynchronize(c->stream));
  memset(o, 0, sizeof(*o));
  o->n = c->n;
  o->ncells = c->ncells;
  o->nleaves = c->nleaves;
  o->nlevels = c->stage >= 1 ? (int32_t)c->lvl.size() - 1 : 0;
  o->max_depth = o->nlevels - 1;
  o->n_problems = c->nprob;
  o->one_graph = c->one_graph && c->stage >= 4 ? 1 : 0;
  for (int d = 0; d < 3; d++) {
    o->region_min[d] = c->h_region[d];
    o->region_max[d] = c->h_region[3 + d];
  }
  if (c->stage >= 3) {
    o->n_m2l = c->n_m2l_all;
    o->n_m2l_exec = c->n_m2l_exec;
    o->n_p2p_leafpairs = c->n_p2p_calls_leaf;
    o->n_p2p_calls = c->h_u64[0];
    o->n_p2p_self = c->h_u64[1];
    o->n_m2l_groups = c->n_groups;
    o->n_m2l_items = c->n_items;
    o->n_m2l_fallback = c->n_fallback;
  }
  auto el = [&](int a, int b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) {
      (void)hipGetLastError();  // not timed: leave no error for the next call to find
      return 0.0;
    }
    return (double)ms;
  };
  if (c->stage >= 1) o->ms_tree = el(EV_BEGIN, EV_TREE);
  if (!c->inner_timed) {
    // a graph without per-phase events: the totals only
    if (c->stage >= 3) o->ms_dtt = el(EV_DTT0, EV_DTT1);
    if (c->stage >= 4) o->ms_total = el(EV_BEGIN, EV_DOWN);
    return TFMM_OK;
  }
  if (c->fused && c->stage >= 4) {
    // fused order: traversal and list building, then upward, M2L/P2P, downward
    o->ms_upward = el(EV_DTT1, EV_UP);
    o->ms_traverse = el(EV_DTT0, EV_DTT1) + el(EV_UP, EV_TRAV);
  } else {
    if (c->stage >= 2) o->ms_upward = el(EV_TREE, EV_UP);
    if (c->stage >= 3) o->ms_traverse = el(EV_UP, EV_TRAV);
  }
  if (c->stage >= 3) {
    o->ms_dtt = el(EV_DTT0, EV_DTT1);
    o->ms_m2l = el(EV_M2L0, EV_M2L1);
    o->ms_m2l_tiles = el(EV_M2L0, EV_M2LT);
    o->ms_p2p = el(EV_P2P0, EV_P2P1);
  }
  if (c->stage >= 4) {
    o->ms_downward = el(EV_TRAV, EV_DOWN);
    o->ms_total = el(EV_BEGIN, EV_DOWN);
  }
  return TFMM_OK;
}

int tfmm_get_sorted(tfmm_ctx *c, uint64_t *keys, int32_t *perm) {
  int rc = need_stage(c, 1, "tfmm_get_sorted");
  if (rc) return rc;
  const void *k = c->keys_in_alt ? c->keys_alt.ptr : c->keys.ptr;
  if (keys) TRY(hipMemcpyAsync(keys, k, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
  if (perm) TRY(hipMemcpyAsync(perm, c->perm.ptr, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
  TRY(hipStreamSynchronize(c->stream));
  return TFMM_OK;
}

int tfmm_get_cells(tfmm_ctx *c, uint64_t *key, int32_t *beg, int32_t *nb, int32_t *leaf,
                   double *center, double *width) {
  int rc = need_stage(c, 1, "tfmm_get_cells");
  if (rc) return rc;
  const size_t nc = (size_t)c->ncells;
  hipStream_t s = c->stream;
  if (key) TRY(hipMemcpyAsync(key, c->c_key.ptr, nc * 8, hipMemcpyDeviceToHost, s));
  if (beg) TRY(hipMemcpyAsync(beg, c->c_beg.ptr, nc * 4, hipMemcpyDeviceToHost, s));
  if (nb) TRY(hipMemcpyAsync(nb, c->c_nb.ptr, nc * 4, hipMemcpyDeviceToHost, s));
  if (center) TRY(hipMemcpyAsync(center, c->c_center.ptr, nc * 24, hipMemcpyDeviceToHost, s));
  if (width) TRY(hipMemcpyAsync(width, c->c_width.ptr, nc * 8, hipMemcpyDeviceToHost, s));
  if (leaf) {
    std::vector<int32_t> ch(nc);
    TRY(hipMemcpyAsync(ch.data(), c->c_child.ptr, nc * 4, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < nc; i++) leaf[i] = ch[i] < 0 ? 1 : 0;
  }
  TRY(hipStreamSynchronize(s));
  return TFMM_OK;
}

int tfmm_get_expansions(tfmm_ctx *c, double *M, double *L) {
  int rc = need_stage(c, 2, "tfmm_get_expansions");
  if (rc) return rc;
  const bool cart = cartesian(c);
  const size_t bytes = cart ? cart_bytes(c->ncells) : sph_bytes(c->ncells);
  if (M) TRY(hipMemcpyAsync(M, (cart ? c->Mc : c->M).ptr, bytes, hipMemcpyDeviceToHost, c->stream));
  if (L) {
    if (c->stage < 3) return fail(c, TFMM_ESTATE, "L is formed by tfmm_traverse");
    TRY(hipMemcpyAsync(L, (cart ? c->Lc : c->L).ptr, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  TRY(hipStreamSynchronize(c->stream));
  return TFMM_OK;
}

int tfmm_get_lists(tfmm_ctx *c, int32_t *m2l_off, int32_t *m2l_src, int32_t *p2p_off,
                   int32_t *p2p_src) {
  int rc = need_stage(c, 3, "tfmm_get_lists");
  if (rc) return rc;
  hipStream_t s = c->stream;
  const size_t no = (size_t)(c->ncells + 1) * 4;
  if (m2l_off || m2l_src) {
    std::vector<int2> pr(c->n_m2l_exec);
    if (c->n_m2l_exec)
      TRY(hipMemcpyAsync(pr.data(), c->m2l_pairs.ptr, pr.size() * 8, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    std::sort(pr.begin(), pr.end(), [](const int2 &a, const int2 &b) {
      return a.x != b.x ? a.x < b.x : a.y < b.y;
    });
    if (m2l_off) {
      size_t e = 0;
      for (int64_t t = 0; t <= c->ncells; t++) {
        while (e < pr.size() && pr[e].x < t) e++;
        m2l_off[t] = (int32_t)e;
      }
    }
    if (m2l_src)
      for (size_t i = 0; i < pr.size(); i++) m2l_src[i] = pr[i].y;
  }
  if (p2p_off) TRY(hipMemcpyAsync(p2p_off, c->p2p_off.ptr, no, hipMemcpyDeviceToHost, s));
  if (p2p_src && c->n_p2p)
    TRY(hipMemcpyAsync(p2p_src, c->p2p_src.ptr, (size_t)c->n_p2p * 4, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  return TFMM_OK;
}

int tfmm_direct(tfmm_ctx *c, const int64_t *targets, int64_t nt, double *out) {
  if (!c) return TFMM_EINVAL;
  if (c->n < 1) return fail(c, TFMM_ESTATE, "no bodies set");
  if (nt < 0 || (nt > 0 && (!targets || !out))) return fail(c, TFMM_EINVAL, "bad targets");
  if (nt == 0) return TFMM_OK;
  for (int64_t i = 0; i < nt; i++)
    if (targets[i] < 0 || targets[i] >= c->n) return fail(c, TFMM_EINVAL, "target index out of range");
  TRY(hipSetDevice(c->device));
  // targets grouped by problem; each group sums over its own problem's bodies
  const int np = c->nprob;
  std::vector<int64_t> rel(nt), where(nt), cnt(np + 1, 0);
  auto prob_of = [&](int64_t i) {
    return (int)(std::upper_bound(c->prob_off.begin(), c->prob_off.begin() + np, i) -
                 c->prob_off.begin()) - 1;
  };
  for (int64_t i = 0; i < nt; i++) cnt[prob_of(targets[i]) + 1]++;
  for (int p = 0; p < np; p++) cnt[p + 1] += cnt[p];
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < nt; i++) {
      const int p = prob_of(targets[i]);
      const int64_t at = fill[p]++;
      rel[at] = targets[i] - c->prob_off[p];
      where[at] = i;
    }
  }
  DevBuf t, o;
  std::vector<double> grouped((size_t)nt * 4);
  hipError_t e = dev_reserve(t, (size_t)nt * 8);
  if (e == hipSuccess) e = dev_reserve(o, (size_t)nt * 32);
  if (e == hipSuccess) e = hipMemcpyAsync(t.ptr, rel.data(), (size_t)nt * 8, hipMemcpyHostToDevice, c->stream);
  for (int p = 0; p < np && e == hipSuccess; p++) {
    const int64_t b = c->prob_off[p], m = c->prob_off[p + 1] - b;
    if (cnt[p + 1] > cnt[p])
      e = launch_direct_fp64(c->bx + b, c->by + b, c->bz + b, c->bq + b, m,
                             dp<int64_t>(t) + cnt[p], 0, cnt[p + 1] - cnt[p],
                             dp<double>(o) + 4 * cnt[p], c->stream, image_range(c), period(c));
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(grouped.data(), o.ptr, (size_t)nt * 32, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_release(t);
  dev_release(o);
  if (e != hipSuccess) return hipfail(c, e, "tfmm_direct");
  for (int64_t a = 0; a < nt; a++) memcpy(out + 4 * where[a], &grouped[4 * a], 32);
  return TFMM_OK;
}

int tfmm_direct_noalloc(tfmm_ctx *c, const int64_t *targets, int64_t nt, double *out) {
  if (!c) return TFMM_EINVAL;
  if (c->n < 1) return fail(c, TFMM_ESTATE, "no bodies set");
  if (nt < 0 || (nt > 0 && (!targets || !out))) return fail(c, TFMM_EINVAL, "bad targets");
  if (nt > TFMM_DIRECT_NOALLOC_MAX)
    return fail(c, TFMM_EINVAL, "tfmm_direct_noalloc: more targets than TFMM_DIRECT_NOALLOC_MAX");
  if (nt == 0) return TFMM_OK;
  for (int64_t i = 0; i < nt; i++)
    if (targets[i] < 0 || targets[i] >= c->n) return fail(c, TFMM_EINVAL, "target index out of range");
  if (!c->noalloc) return fail(c, TFMM_ESTATE, "tfmm_direct_noalloc: no scratch address");
  TRY(hipSetDevice(c->device));
  int64_t *d_idx = static_cast<int64_t *>(c->noalloc);
  double *d_out = reinterpret_cast<double *>(static_cast<unsigned char *>(c->noalloc) +
                                             NOALLOC_IDX_BYTES);
  hipError_t e = hipMemcpyAsync(d_idx, targets, (size_t)nt * 8, hipMemcpyHostToDevice, c->stream);
  // one launch per run of adjacent targets of one problem; the kernel turns
  // the caller's (forest) indices into the problem's own by its base
  const int np = c->nprob;
  for (int64_t i = 0; i < nt && e == hipSuccess;) {
    const int p = (int)(std::upper_bound(c->prob_off.begin(), c->prob_off.begin() + np,
                                         targets[i]) - c->prob_off.begin()) - 1;
    const int64_t b = c->prob_off[p], end = c->prob_off[p + 1];
    int64_t j = i + 1;
    while (j < nt && targets[j] >= b && targets[j] < end) j++;
    e = launch_direct_fp64(c->bx + b, c->by + b, c->bz + b, c->bq + b, end - b, d_idx + i, b,
                           j - i, d_out + 4 * i, c->stream, image_range(c), period(c));
    i = j;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_out, (size_t)nt * 32, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "tfmm_direct_noalloc");
  return TFMM_OK;
}

int tfmm_tree_features(tfmm_ctx *c, float *feat) {
  int rc = need_stage(c, 1, "tfmm_tree_features");
  if (rc) return rc;
  if (!feat) return fail(c, TFMM_EINVAL, "null feature array");
  if (c->nprob > 1) return fail(c, TFMM_EINVAL, "tree features describe a single problem, not a batch");
  hipStream_t s = c->stream;
  TRY(dev_reserve(c->feat_acc, 16 * 8));
  TRY(hipMemsetAsync(c->feat_acc.ptr, 0, 16 * 8, s));
  TRY(launch_tree_stats(cells_of(c), (int32_t)c->ncells, c->prm.ncrit, dp<double>(c->feat_acc), s));
  double a[16];
  TRY(hipMemcpyAsync(a, c->feat_acc.ptr, 16 * 8, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  const double nl = a[8] > 0 ? a[8] : 1;
  const double mean = a[10] / nl;
  const double var = std::max(0.0, a[11] / nl - mean * mean);
  feat[0] = (float)(std::log2((double)c->n) / 23.0);
  feat[1] = (float)(((int)c->lvl.size() - 2) / 20.0);
  for (int b = 0; b < 8; b++) feat[2 + b] = (float)(a[b] / nl);
  feat[10] = (float)mean;
  feat[11] = (float)std::sqrt(var);
  feat[12] = (float)(a[12] / nl);
  feat[13] = (float)(a[9] / std::max(1.0, a[8] + a[9]));
  feat[14] = (float)(a[13] / nl / 20.0);
  feat[15] = (float)(a[14] / (double)c->n);
  return TFMM_OK;
}

int tfmm_pick_params(tfmm_ctx *c, tfmm_picker *picker, tfmm_params *io) {
  if (!c) return TFMM_EINVAL;
  if (!io) return fail(c, TFMM_EINVAL, "null params");
  static const int32_t kNcrit[5] = {16, 32, 64, 128, 256};
  io->ncrit = 64;  // fallback
  io->theta = 0.4;
  io->mutual = 0;
  if (!picker) return TFMM_OK;
  if (c->n < 1) return fail(c, TFMM_ESTATE, "no bodies set");
  if (c->nprob > 1) return fail(c, TFMM_EINVAL, "the picker chooses parameters for a single problem, not a batch");
  if (c->stage < 1 || c->prm.ncrit != 64) {
    tfmm_params trial = *io;
    int rc = tfmm_build_tree(c, &trial);
    if (rc) return rc;
  }
  float feat[16];
  int rc = tfmm_tree_features(c, feat);
  if (rc) return rc;
  hipStream_t s = c->stream;
  TRY(dev_reserve(c->feat_x, 16 * 4));
  TRY(dev_reserve(c->feat_out, 8 * 4));
  TRY(hipMemcpyAsync(c->feat_x.ptr, feat, 64, hipMemcpyHostToDevice, s));
  TRY(launch_mlp_forward(picker_weights(picker), dp<float>(c->feat_x), 1, dp<float>(c->feat_out), s));
  float o[6];
  TRY(hipMemcpyAsync(o, c->feat_out.ptr, 24, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  int best = -1;
  for (int i = 0; i < 5; i++) {
    if (!std::isfinite(o[i])) return TFMM_OK;  // fallback stays
    if (best < 0 || o[i] > o[best]) best = i;
  }
  if (!std::isfinite(o[5])) return TFMM_OK;
  io->ncrit = kNcrit[best];
  io->theta = std::min(0.2 + 0.4 / (1.0 + std::exp(-(double)o[5])), (double)picker_theta_max(picker));
  return TFMM_OK;
}

}  // extern "C"
