// tfmm_kernels.h -- device data structures and the launch wrappers of the
// TapasFMM kernels (one wrapper per kernel family, defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfmm {

// Cells in breadth-first order (level-contiguous, siblings adjacent).
struct Cells {
  uint64_t *key;    // Morton key with depth bits (TP/sfc_morton.h)
  int32_t *beg;     // first sorted body (valid for every cell)
  int32_t *nb;      // number of bodies
  int32_t *parent;  // -1 for the root
  int32_t *child;   // index of the first of 8 children, -1 for leaves
  int32_t *depth;
  double *center;   // 3 per cell
  double *width;    // max_d (max_d - min_d), the quantity the MAC radius uses
};

// Sorted bodies, SoA.
struct Bodies {
  const double *x, *y, *z, *q;
  int64_t n;
};

// A CSR interaction list: entries [off[t], off[t+1]) of src are the sources
// of target cell t.
struct Csr {
  const int32_t *off;  // ncells + 1
  const int32_t *src;
};

// Sizes of an evaluation recorded as one graph (plan.hip): words of the
// device sizes buffer, and the status bits of the capacities it exceeded.
enum : int {
  SZ_STATUS = 0,
  SZ_NCELLS,
  SZ_NLEAVES,
  SZ_DEPTH,
  SZ_NTILES,
  SZ_NG,
  SZ_NITEMS,
  SZ_CHUNK,
  SZ_NUNITS,
  SZ_WORDS = 16
};
enum : uint32_t {
  ST_CELLS = 1,       // cell capacity (level steps)
  ST_LEVEL_GRID = 2,  // a level larger than its grid (level steps)
  ST_DEPTH = 4,       // a tree deeper than the planned levels
  ST_TRAVERSAL = 8,   // lists, frontiers or step grids of the traversal
  ST_GROUPS = 16,
  ST_ITEMS = 32,
  ST_UNITS = 64,
  ST_FALLBACK = 128
};
struct PlanCaps {
  uint32_t m2l, p2p, front, items, units, fallback;
};

// tree.hip
// *flag |= 1 when any x, y, z or q of the n bodies is not finite
hipError_t launch_check_finite(const double *x, const double *y, const double *z,
                               const double *q, int64_t n, uint32_t *flag, hipStream_t s);
hipError_t launch_bbox(const double *x, const double *y, const double *z, int64_t n,
                       double *partial, int nparts, double *region, hipStream_t s);
// keys, and the bodies packed as (x, y, z, q) records for the permutation
hipError_t launch_keys(const double *x, const double *y, const double *z, const double *q,
                       int64_t n, const double *regions, const int64_t *off, int nprob,
                       uint64_t *keys, uint32_t *idx, uint32_t *bprob, double4 *rec,
                       hipStream_t s);
// forest of nprob problems, problem p = bodies [off[p], off[p+1])
hipError_t launch_bbox_forest(const double *x, const double *y, const double *z,
                              const int64_t *off, int nprob, double *regions, hipStream_t s);
hipError_t launch_problem_keys(const uint32_t *order, const uint32_t *bprob, int64_t n,
                               uint32_t *pkey, uint32_t *pos, hipStream_t s);
hipError_t launch_problem_gather(const uint32_t *pos, const uint64_t *keys1,
                                 const uint32_t *order1, int64_t n, uint64_t *keys,
                                 uint32_t *order, hipStream_t s);
// the first traversal frontier: each root with itself, or (periodic) the one
// root with its 27 images
hipError_t launch_root_pairs(int2 *front, int nroots, uint32_t *count, hipStream_t s,
                             bool periodic = false);
hipError_t launch_permute(const uint32_t *order, const double4 *rec, int64_t n, double *sx,
                          double *sy, double *sz, double *sq, int32_t *perm, hipStream_t s);
hipError_t launch_roots(Cells c, int64_t n, const int64_t *off, int nprob,
                        const double *regions, int32_t *lev, hipStream_t s);
// one level of the cell build with sizes on the device (lev: first cell and
// count per level); grid_cells bounds the level's size; status bit 1: cell
// capacity cap exceeded, bit 2: level larger than grid_cells
int32_t level_blocks(int32_t cells);
// periodic images: the region (one problem) becomes the periodic cell of
// side cycle around the bodies' bounding cube; *flag = 1 if they span more
hipError_t launch_periodic_region(double *region, double cycle, uint32_t *flag, hipStream_t s);
// last: the planned tree has no level after d (splits there set ST_DEPTH)
hipError_t launch_level_step(Cells c, int32_t *lev, int d, int32_t grid_cells, int ncrit,
                             int32_t cap, uint32_t *flag, uint32_t *loc, uint32_t *bsum,
                             uint32_t *status, const uint64_t *keys, const double *regions,
                             const uint32_t *sprob, hipStream_t s, bool last = false);
// plan.hip
hipError_t launch_tree_final(int32_t *lev, int L, const uint32_t *lev_status, int32_t tile,
                             uint32_t *sz, hipStream_t s);
hipError_t launch_pad_cells(Cells c, const uint32_t *sz, int32_t cap, hipStream_t s);
hipError_t launch_groups_count(const uint32_t *gid, const uint32_t *used, int32_t nslots,
                               uint32_t gcap, uint32_t *sz, hipStream_t s);
hipError_t launch_plan_check(const uint32_t *cnt, int nsteps, const uint32_t *step_grid,
                             const PlanCaps &caps, const uint32_t *gcnt, const uint32_t *ucnt,
                             const uint32_t *uoff, int32_t ntiles, uint32_t *sz, hipStream_t s);
hipError_t launch_split_flags(Cells c, int32_t lbeg, int32_t lcnt, int ncrit, uint32_t *flag,
                              hipStream_t s);
hipError_t launch_children(Cells c, int32_t lbeg, int32_t lcnt, const uint32_t *flag,
                           const uint32_t *slot, int32_t next_beg, const uint64_t *keys,
                           const double *regions, const uint32_t *sprob, hipStream_t s);

// expansions.hip (FP64)
// zero fill that can be captured into a hipGraph (ptr 16-byte aligned)
hipError_t launch_fill_zero(void *ptr, size_t bytes, hipStream_t s);
hipError_t launch_p2m(Cells c, int32_t ncells, Bodies b, double *M, hipStream_t s);
// translate.hip (FP64 M2M / L2L on MFMA)
size_t translation_ops_doubles();
hipError_t launch_translation_ops(double *ops, hipStream_t s);
// periodic images beyond the first shell (images > 1): their far field into
// the root's local expansion, from the root's multipole
hipError_t launch_periodic_far(const double *Mroot, double *Lroot, double cycle, int images,
                               hipStream_t s);
// one level of parents (npar, children from first_child); with lev the
// children's level dl is read on the device and npar is the capacity
hipError_t launch_m2m(Cells c, int32_t first_child, int32_t npar, const double *ops, double *M,
                      hipStream_t s, const int32_t *lev = nullptr, int dl = 0);
hipError_t launch_l2l(Cells c, int32_t first_child, int32_t npar, const double *ops, double *L,
                      hipStream_t s, const int32_t *lev = nullptr, int dl = 0);
hipError_t launch_l2p(Cells c, int32_t ncells, Bodies b, const double *L, const double *p2p,
                      const int32_t *perm, int32_t *leaf_of, double *trg, double *out,
                      hipStream_t s);

// cartesian.hip: the Cartesian Taylor basis, CART_TERMS doubles per
// expansion (FP64); the M2L and the periodic far field read tables built
// by cart_upload_tables (cart_tables_bytes of device memory, synchronous)
constexpr int CART_TERMS = 220;  // P (P+1) (P+2) / 6 for P = 10
constexpr int CART_BINS = 6;
size_t cart_tables_bytes();
hipError_t cart_upload_tables(void *dst);
hipError_t launch_cart_p2m(Cells c, int32_t ncells, Bodies b, double *M, hipStream_t s);
// one level: npar groups of 8 siblings from first_child into their parents
hipError_t launch_cart_m2m(Cells c, int32_t first_child, int32_t npar, double *M, hipStream_t s);
// list: per-target CSR of source words (cell | image << IMG_SHIFT)
hipError_t launch_cart_m2l(Cells c, Csr list, int32_t ncells, const double *M, double *L,
                           const void *tables, double cycle, hipStream_t s);
// one level: nchild cells from first_child, from their parents
hipError_t launch_cart_l2l(Cells c, int32_t first_child, int32_t nchild, double *L,
                           hipStream_t s);
hipError_t launch_cart_l2p(Cells c, int32_t ncells, Bodies b, const double *L, const double *p2p,
                           const int32_t *perm, double *trg, double *out, hipStream_t s);
hipError_t launch_cart_periodic_far(const double *Mroot, double *Lroot, double cycle, int images,
                                    const void *tables, hipStream_t s);
// basis change of every cell: Cartesian multipoles into spherical ones (the
// library's NTERM complex layout) and spherical locals into Cartesian ones
hipError_t launch_cart_to_sph(int32_t ncells, const double *Mc, double *Ms, const void *tables,
                              hipStream_t s);
hipError_t launch_sph_to_cart(int32_t ncells, const double *Ls, double *Lc, const void *tables,
                              hipStream_t s);
// traverse.hip
struct DttCounters {
  // [0] next-frontier size, [1] executed M2L pairs, [2] executed P2P leaf pairs,
  // [3] all M2L pairs (also those with an empty cell)
  uint32_t *cnt;
};
// per-cell records the traversal steps read (geo: centre + MAC radius)
hipError_t launch_dtt_cells(Cells c, int32_t ncells, double theta, double4 *geo, int2 *topo,
                            hipStream_t s);
// One breadth-first traversal step.  Every M2L pair stored also gets its
// group-table slot (m2l_slot, -1 outside the table), marks used[slot] and
// counts the pairs outside the table in *n_outside.
hipError_t launch_dtt_step(const int2 *in, const uint32_t *nin_dev, uint32_t nin_max, int2 *out,
                           uint32_t out_cap, uint32_t *nout, int2 *m2l, uint32_t m2l_cap,
                           int2 *p2p, uint32_t p2p_cap, uint32_t *cnt, const double4 *geo,
                           const int2 *topo, const int32_t *depth, const double *width,
                           int32_t *m2l_slot, uint32_t *used, uint32_t *n_outside, int mutual,
                           hipStream_t s, double cycle = 0);
// group-table slots of M2L pairs given from outside the traversal
hipError_t launch_pair_slots(const int2 *pairs, uint32_t n, const double4 *geo,
                             const int32_t *depth, const double *width, int32_t *slot,
                             uint32_t *used, uint32_t *n_outside, hipStream_t s);
// leaf-pair list: targets and sources of the pairs; then, from the
// target-sorted list, the CSR offsets and every target's sources ascending.
// With dn the pair count is *dn on the device (n its capacity).
hipError_t launch_split_pairs(const int2 *pairs, uint32_t n, uint32_t *tgt, uint32_t *src,
                              hipStream_t s, const uint32_t *dn = nullptr);
// longq: ncells + 1 words of scratch (lists longer than one wave sorts)
hipError_t launch_csr_sorted_sources(const uint32_t *tgt, const uint32_t *in, uint32_t n,
                                     int32_t ncells, int32_t *off, int32_t *out, uint32_t *longq,
                                     hipStream_t s, const uint32_t *dn = nullptr);
hipError_t launch_unpack_pairs(const uint64_t *keys, uint32_t n, int sbits, int32_t *src,
                               hipStream_t s, const uint32_t *dn = nullptr);
hipError_t launch_csr_offsets(const uint64_t *sorted_keys, uint32_t n, int sbits,
                              int32_t ncells, int32_t *off, hipStream_t s,
                              const uint32_t *dn = nullptr);
// m2l.hip
size_t m2l_pairs_scratch_bytes(int32_t npairs);
// with dn the pair count is *dn on the device (npairs its capacity)
// (list sources are source words: cell and periodic image, cycle the period)
hipError_t launch_m2l_fp64(Cells c, int32_t ncells, Csr list, int32_t npairs, const double *M,
                           double *L, void *scratch, hipStream_t s, const uint32_t *dn = nullptr,
                           double cycle = 0);

// p2p.hip
// (list sources are source words: leaf and periodic image, cycle the period)
// resident > 0: that many resident waves take the target cells from *next
// (cleared first), as launch_p2p_fp32
hipError_t launch_p2p_fp64(Cells c, int32_t ncells, Csr list, Bodies b, double *acc,
                           int resident, uint32_t *next, hipStream_t s, double cycle = 0);

// direct.hip
// target t sums over the n bodies at x, y, z, q: index targets[t] - tbase
// with prange > 0, also over the images of every body at offsets (ix, iy, iz)
// cycle, |ix|, |iy|, |iz| <= prange
hipError_t launch_direct_fp64(const double *x, const double *y, const double *z,
                              const double *q, int64_t n, const int64_t *targets, int64_t tbase,
                              int64_t nt, double *out, hipStream_t s, int prange = 0,
                              double cycle = 0);

}  // namespace tfmm

namespace tfmm {
// m2l_batched.hip
int m2l_group_bits();  // bits of a group index in the tile kernel's column word
int32_t m2l_group_slots(int max_depth);
int m2l_tile_size();
int m2l_src_bits();
hipError_t m2l_upload_codes(uint32_t *dev_codes, hipStream_t s);
size_t m2l_codes_bytes();
// every pair in the group table: grouped keys and values in pair order
hipError_t launch_m2l_classify_all(const int2 *pairs, const int32_t *slot_of, uint32_t n,
                                   const uint32_t *gid_of_slot, int gbits, uint32_t *gkeys,
                                   uint32_t *gvals, hipStream_t s);
// Grouped keys and values at the pair's index, the key 0xffffffff (after
// every grouped key) for a pair outside the groups, which is appended to
// fkeys/fvals; counters[1] = pairs outside, counters[0] = grouped pairs.  A
// sort of all n keys puts the grouped pairs first.
// With dn the pair count is *dn on the device (n its capacity); group
// indices are clamped below gcap, groups at or past gcap get no harmonics.
hipError_t launch_m2l_classify(const int2 *pairs, const int32_t *slot_of, uint32_t n,
                               const uint32_t *gid_of_slot, int gbits, int sbits,
                               uint32_t *gkeys, uint32_t *gvals, uint64_t *fkeys,
                               uint32_t *fvals, uint32_t *counters, hipStream_t s,
                               const uint32_t *dn = nullptr, uint32_t gcap = UINT32_MAX);
hipError_t launch_m2l_group_harmonics(const uint32_t *used, const uint32_t *gid_of_slot,
                                      int32_t nslots, double *Y64, float *Y32, int8_t *gdl,
                                      int8_t *glt, hipStream_t s, uint32_t gcap = UINT32_MAX);
hipError_t launch_m2l_scale_m(Cells c, int32_t ncells, const double *M, void *Mq, bool fp32,
                              hipStream_t s);
// items of the (tile, group)-sorted grouped pairs: every pair's position in
// its segment and the item starts (runs of 16) per block of 1024 pairs,
// their count read back; bcount/bstart: one word per block
hipError_t launch_m2l_count_items(const uint32_t *keys, uint32_t n, uint32_t *ipos,
                                  uint32_t *bcount, uint32_t *bstart, void *scan_tmp,
                                  uint32_t *h_pinned, uint32_t *nitems_out, hipStream_t s);
// the same with the pair count *dn on the device (n its capacity) and the
// item count left on the device in *nitems
hipError_t launch_m2l_count_items_dev(const uint32_t *keys, uint32_t n, const uint32_t *dn,
                                      uint32_t *ipos, uint32_t *bcount, uint32_t *bstart,
                                      void *scan_tmp, uint32_t *nitems, hipStream_t s);
// items at or past items_cap are not written
hipError_t launch_m2l_write_items(const uint32_t *keys, const uint32_t *vals, uint32_t n,
                                  int gbits, const uint32_t *ipos, const uint32_t *bstart,
                                  const int8_t *gdl, int4 *items, uint2 *cols, hipStream_t s,
                                  const uint32_t *dn = nullptr, uint32_t items_cap = UINT32_MAX);
hipError_t launch_m2l_tile_offsets(const int4 *items, uint32_t nitems, int32_t ntiles,
                                   int32_t *off, hipStream_t s,
                                   const uint32_t *dnitems = nullptr);
// m2l_unit_chunk of the device's item and tile counts, into *chunk
hipError_t launch_m2l_unit_chunk(const uint32_t *nitems, const uint32_t *ntiles, double factor,
                                 int32_t *chunk, hipStream_t s);
// work units of the tile kernel: runs of at most `chunk` items of one tile
int32_t m2l_unit_chunk(uint32_t nitems, int32_t ntiles, double factor);
uint32_t m2l_units_bound(uint32_t nitems, int32_t ntiles, int32_t chunk);
size_t m2l_unit_part_bytes(uint32_t units_bound, int32_t ntiles);
// units at or past units_cap are not written; with dchunk the chunk is read
// on the device
hipError_t launch_m2l_units(const int32_t *tile_off, int32_t ntiles, int32_t chunk,
                            uint32_t *ucnt, uint32_t *uoff, int2 *units, void *scan_tmp,
                            hipStream_t s, uint32_t units_cap = UINT32_MAX,
                            const int32_t *dchunk = nullptr);
// A fragments of every group (the tile kernel's operand order, row factors
// 2^(-j dl) included) from the group harmonics (FP32 or FP64), the operand
// codes of both types and the groups' level differences
size_t m2l_afrag_bytes(uint32_t ngroups, bool fp32);
hipError_t launch_m2l_group_afrag(const void *Yg, uint32_t ngroups, const uint32_t *codes,
                                  const int8_t *gdl, void *Af, bool fp32, hipStream_t s,
                                  const uint32_t *dng = nullptr);
hipError_t launch_m2l_tiles(const int32_t *tile_off, const int2 *units, const uint32_t *ucnt,
                            const uint32_t *uoff, int32_t ntiles, int32_t chunk, int xcd_run,
                            uint32_t units_bound, const uint2 *cols, const void *Af,
                            const void *Mq, Cells c, int32_t ncells, double *L, double *part,
                            bool fp32, hipStream_t s, const int32_t *dchunk = nullptr);

// p2p.hip (mixed precision)
hipError_t launch_leaf_relpos(Cells c, int32_t ncells, Bodies b, float4 *rel, hipStream_t s);
// resident > 0: that many resident waves take target cells from the counter
// next (4 bytes, zeroed here); 0: one workgroup per cell
// the resident FP32 P2P on a counter cleared by the caller (waves
// workgroups of one wave each); launches sharing the counter split the cells
hipError_t launch_p2p_fp32_drain(Cells c, int32_t ncells, Csr list, Bodies b, const float4 *rel,
                                 double *acc, int waves, uint32_t *next, hipStream_t s,
                                 double cycle);
hipError_t launch_p2p_fp32(Cells c, int32_t ncells, Csr list, Bodies b, const float4 *rel,
                           double *acc, int resident, uint32_t *next, hipStream_t s,
                           double cycle = 0);
}  // namespace tfmm

namespace tfmm {
// mlp.hip
int mlp_nparam();
hipError_t launch_mlp_grad(const float *W, const float *X, const int32_t *cls,
                           const float *theta, int n, float *G, float *loss, hipStream_t s);
hipError_t launch_mlp_adam(float *W, const float *G, int n, float *m, float *v, float lr,
                           int step, float *gmean, hipStream_t s);
hipError_t launch_mlp_forward(const float *W, const float *X, int n, float *out, hipStream_t s);
hipError_t launch_tree_stats(Cells c, int32_t ncells, int ncrit, double *acc, hipStream_t s);
}  // namespace tfmm
