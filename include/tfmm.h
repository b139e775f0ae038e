/*
 * tfmm.h -- C ABI of the MI355X (gfx950) TapasFMM library (libtfmm.so).
 *
 * Plain C types only: pointers, sizes and POD structs; no exceptions, no
 * aborts across the boundary.  Every call returns TFMM_OK (0) or a negative
 * status; tfmm_last_error() gives the message of the last failure.
 *
 * Each entry point replaces a piece of the reference's path (paths relative to
 * the reference repository; TP = cpp/include/tapas, EX =
 * sample/exafmm-dev-13274dd4ac68):
 *
 *   tfmm_set_bodies*     the body copy Partition makes of the caller's array
 *                        (TP/hot/buildtree.h:72-73, TP/hot.h:1349-1352)
 *   tfmm_build_tree      TapasFMM::Partition -> SamplingOctree::Build
 *                        (TP/hot.h:1240-1284, TP/hot/buildtree.h:329-459)
 *   tfmm_upward          TapasFMM::Map(FMM_Upward(), *root, theta)
 *                        (EX/examples/tapas_exafmm.cxx:586, :87-122)
 *   tfmm_traverse        TapasFMM::Map(FMM_DTT(), Product(*root, *root), mutual,
 *                        nspawn, theta) (tapas_exafmm.cxx:606, :209-286); on the
 *                        reference's GPU build the GPUMapper records the P2P
 *                        products and launches them at the end
 *                        (TP/hot/mapper.h:452-690, TP/vectormap_cuda.h:911-1004)
 *   tfmm_downward        TapasFMM::Map(FMM_Downward(), *root)
 *                        (tapas_exafmm.cxx:627, :124-140)
 *   tfmm_get_fields      CopyBackResult: root->local_body_attr(i)
 *                        (tapas_exafmm.cxx:388-398, TP/hot.h:1037-1050)
 *   tfmm_get_sorted /    root->data().local_body_keys_, ht_ (TP/hot/shared_data.h:31-68)
 *   tfmm_get_cells
 *   tfmm_stats           RankCSV columns total/upward/traverse/downward/tree and
 *                        the -DCOUNT numP2P/numM2L counters
 *                        (tapas_exafmm.cxx:651-677, :39-50)
 *   tfmm_direct          CheckResult's direct sum (tapas_exafmm.cxx:288-383)
 *   tfmm_set_batch*,     no single counterpart: what the reference does by
 *   tfmm_batch_evaluate  running Partition + Map(...) once per problem
 *                        (tapas_exafmm.cxx:560-627 in a loop), done as one
 *                        forest in one set of launches
 */
#ifndef TFMM_H_
#define TFMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFMM_OK 0
#define TFMM_EINVAL (-22) /* invalid argument or parameters */
#define TFMM_ENOMEM (-12) /* device allocation failed */
#define TFMM_EHIP (-5)    /* HIP runtime error */
#define TFMM_ESTATE (-77) /* call out of order (e.g. upward before build_tree) */

#define TFMM_P 10      /* expansion order (EXPANSION=10) */
#define TFMM_NTERM 55  /* P*(P+1)/2 coefficients per cell */

#define TFMM_MAX_PROBLEMS 65536 /* problems in one batch */

#define TFMM_FP64 0  /* FP64 throughout */
#define TFMM_MIXED 1 /* FP32 P2P with compensated sums, FP32 batched M2L; FP64 tree/L2L/L2P */

typedef struct tfmm_ctx tfmm_ctx;

/* FMM parameters (the reference's -c/--ncrit, -t/--theta, -m/--mutual of
 * EX/include/args.h:101-158). */
typedef struct {
  int32_t ncrit;     /* a cell is a leaf iff it holds <= ncrit bodies */
  int32_t mutual;    /* 0 or 1: 1 visits every unordered cell pair once and applies
                        its M2L and P2P to both cells (the reference's default
                        --mutual 1, EX/include/args.h:102; its CUDA path refused
                        it, tapas_exafmm.cxx:507-515) */
  int32_t precision; /* TFMM_FP64 or TFMM_MIXED */
  int32_t reserved;
  double theta;      /* multipole acceptance parameter, 0 < theta */
} tfmm_params;

/* Layout of the caller's array-of-structs bodies (the reference's Body is
 * 96 B with X at 0 and SRC at 24, EX/include/types.h:62-74). */
typedef struct {
  int64_t n;
  size_t stride;    /* bytes between consecutive bodies */
  size_t off_x;     /* byte offset of double X[3] */
  size_t off_q;     /* byte offset of double SRC */
} tfmm_body_desc;

typedef struct {
  int64_t n;
  int64_t ncells;
  int64_t nleaves;
  int32_t nlevels;    /* number of non-empty levels (max depth + 1) */
  int32_t max_depth;
  double region_min[3];
  double region_max[3];
  uint64_t n_m2l;            /* M2L calls of the traversal (numM2L; a mutual pair once) */
  uint64_t n_m2l_exec;       /* M2L list entries executed (both cells non-empty; a
                                mutual pair is two entries) */
  uint64_t n_p2p_leafpairs;  /* P2P leaf-pair calls (both non-empty; a mutual pair once) */
  uint64_t n_p2p_calls;      /* body pairs the reference's P2P functor visits (numP2P:
                                includes r == 0 self pairs; under mutual the pairs
                                j <= i of a leaf with itself) */
  uint64_t n_p2p_self;       /* r == 0 self pairs among them */
  /* milliseconds of the last evaluation, measured with HIP events on the
   * context stream; kernel entries are summed over their launches */
  double ms_tree, ms_upward, ms_traverse, ms_downward, ms_total;
  double ms_p2p, ms_m2l, ms_dtt;
  /* M2L batching: offset groups, MFMA work items (<= 16 pairs of one group
   * and one 128-cell target tile each) and pairs outside the groups */
  uint64_t n_m2l_groups, n_m2l_items, n_m2l_fallback;
  /* the batched M2L tile kernel alone (its one launch, inside ms_m2l) */
  double ms_m2l_tiles;
  /* problems in the forest (1 unless the bodies were set as a batch);
   * region_min/max above are those of problem 0 */
  int64_t n_problems;
  /* 1 when the last tfmm_evaluate ran as one graph launch (TFMM_TUNE_GRAPH
   * = 2), 0 when it ran the ordinary way */
  int64_t one_graph;
} tfmm_stats_t;

int tfmm_create(int hip_device, tfmm_ctx **out);
void tfmm_destroy(tfmm_ctx *ctx);
const char *tfmm_last_error(const tfmm_ctx *ctx);
/* the hipStream_t all work of the context is queued on */
void *tfmm_stream(tfmm_ctx *ctx);
int tfmm_sync(tfmm_ctx *ctx);

/* Copy n host bodies (AoS as described) into the context. */
int tfmm_set_bodies(tfmm_ctx *ctx, const void *host_bodies, const tfmm_body_desc *desc);
/* Use n device-resident bodies (SoA FP64 arrays in HBM) without a copy; the
 * arrays must stay valid until the next set_bodies call or destroy. */
int tfmm_set_bodies_device(tfmm_ctx *ctx, int64_t n, const double *x, const double *y,
                           const double *z, const double *q);

/* A batch of nprob independent problems evaluated as one forest: problem p
 * has its own bounding box (with the single-problem rules), its own tree
 * and no interaction with any other problem; every phase runs over all
 * problems in the same launches.  Bodies are concatenated in problem order:
 * fields, sorted keys, cells and lists are those of the concatenation, and
 * the cells of problem p descend from root cell p.  tfmm_set_bodies*
 * returns the context to a single problem. */
/* copies: problem p = descs[p].n bodies at bodies[p] (AoS as described) */
int tfmm_set_batch(tfmm_ctx *ctx, int nprob, const tfmm_body_desc *descs,
                   const void *const *bodies);
/* device-resident SoA arrays of offsets[nprob] bodies; problem p holds
 * [offsets[p], offsets[p+1]) (host array, offsets[0] = 0, no empty problem) */
int tfmm_set_batch_device(tfmm_ctx *ctx, int nprob, const int64_t *offsets, const double *x,
                          const double *y, const double *z, const double *q);
/* tfmm_set_batch + tfmm_evaluate */
int tfmm_batch_evaluate(tfmm_ctx *ctx, int nprob, const tfmm_body_desc *descs,
                        const void *const *bodies, const tfmm_params *params);
/* number of problems; offsets (nprob + 1 entries) when not null */
int tfmm_num_problems(const tfmm_ctx *ctx, int64_t *offsets);

/* Periodic boundary conditions (ExaFMM's images, EX/include/traversal.h:
 * 126-185 and 247-268; Tapas takes the Xperiodic argument of M2L and P2P,
 * EX/examples/LaplaceSphericalCPU_tapas.h:209-220, LaplaceP2PCPU_tapas.h:50-54,
 * and checks against the periodic direct sum, tapas_exafmm.cxx:288-345, but
 * its traversal passes Xperiodic = 0, tapas_exafmm.cxx:224).  images = 0
 * (default): none.  images >= 1: the root cell becomes the periodic cell, a
 * cube of side cycle centred on the bodies' bounding cube (the bodies must
 * span at most one period); the traversal starts from the root and each of
 * its 27 images at (ix, iy, iz) cycle, ix, iy, iz in {-1, 0, 1}, with
 * mutual = 0; for images >= 2 the copies farther out, up to the
 * (3^images)^3 copies of the periodic cell around it, add their far field
 * through multipoles of ever larger blocks of copies into the root.
 * tfmm_direct then sums over the same (3^images)^3 copies (the reference's
 * prange): 27^images times the work of the plain direct sum.  One problem
 * only (not a batch), no mutual interaction, no
 * recorded lists (tfmm_traverse_lists); TFMM_TUNE_GRAPH 2 evaluates the
 * ordinary way.  Takes effect at the next tree build. */
#define TFMM_PERIODIC_MAX_IMAGES 6
int tfmm_set_periodic(tfmm_ctx *ctx, int images, double cycle);

/* Expansion basis (the reference's compile-time BASIS, EX/Makefile.include:13
 * and EX/include/types.h:39-46, as a run-time choice).
 * TFMM_BASIS_SPHERICAL (default): spherical harmonics, the Tapas kernels
 * (EX/examples/LaplaceSphericalCPU_tapas.h), 55 complex terms per expansion.
 * TFMM_BASIS_CARTESIAN: Cartesian Taylor expansions of ExaFMM's
 * LaplaceCartesianCPU.cxx, TFMM_CART_TERMS real terms per expansion (terms
 * of degree 0 .. 9; term (x, y, z) of degree d = x + y + z at
 * d (d+1) (d+2) / 6 + (y + z) (y + z + 1) / 2 + z).  Both bases truncate
 * the same series at total degree 9, so their fields agree to rounding.
 * TFMM_TUNE_CART_M2L picks the route of the Cartesian M2L (by default the
 * FP64 tile kernel between exact basis changes; the stats' group, item and
 * fallback counts are then the tile kernel's, and 0 for the direct kernel).
 * The P2P follows the precision parameter; the Cartesian expansions and
 * their M2L are FP64 in either precision.  Mutual interaction, batches and
 * periodic images work in either basis; TFMM_TUNE_GRAPH 2 evaluates the
 * ordinary way.  A change keeps the tree: the next tfmm_upward (or
 * tfmm_evaluate) forms the expansions in the new basis. */
#define TFMM_BASIS_SPHERICAL 0
#define TFMM_BASIS_CARTESIAN 1
#define TFMM_CART_TERMS 220
int tfmm_set_basis(tfmm_ctx *ctx, int basis);
int tfmm_get_basis(tfmm_ctx *ctx, int *basis);

int tfmm_build_tree(tfmm_ctx *ctx, const tfmm_params *params);
/* Change theta or the precision of the following phases without rebuilding
 * the tree (the reference passes theta to Map(FMM_DTT, ...), not to
 * Partition).  A different ncrit is rejected: it changes the tree. */
int tfmm_set_params(tfmm_ctx *ctx, const tfmm_params *params);
int tfmm_upward(tfmm_ctx *ctx);
int tfmm_traverse(tfmm_ctx *ctx);
/* tfmm_traverse with interaction lists recorded by the caller instead of the
 * GPU traversal: the reference's own FMM_DTT functor run on host cells, its
 * M2L(Ci, Cj, ...) calls and Map(P2P(), Product(Ci.bodies(), Cj.bodies()), ...)
 * products recorded, as the reference's GPUMapper records P2P products
 * (TP/hot/mapper.h:452-690), then executed by the library's M2L and P2P
 * kernels.  m2l: n_m2l (target, source) cell pairs, 2 int32 each, in the
 * breadth-first cell order of tfmm_get_cells, empty cells included (counted
 * like the reference's numM2L, not executed); p2p: n_p2p leaf pairs (pairs with
 * an empty leaf are skipped).  With the mutual parameter set (tfmm_set_params)
 * the lists hold every unordered pair in both directions (a leaf with itself
 * once) and the counters count each pair once, as the reference's functors
 * are called.  One problem only (not a batch). */
int tfmm_traverse_lists(tfmm_ctx *ctx, const int32_t *m2l, int64_t n_m2l, const int32_t *p2p,
                        int64_t n_p2p);
int tfmm_downward(tfmm_ctx *ctx);
/* build_tree + upward + traverse + downward: the traversal and the list
 * building first, then the numeric kernels (as one hipGraph, see
 * TFMM_TUNE_GRAPH); results as from the four calls. */
int tfmm_evaluate(tfmm_ctx *ctx, const tfmm_params *params);

/* Fields (pot, ax, ay, az) per body, 4 doubles each.  order 0: the caller's
 * body order; 1: sorted (Morton) order.  dst is a host pointer, or a device
 * pointer when on_device != 0. */
int tfmm_get_fields(tfmm_ctx *ctx, double *dst, int order, int on_device);

int tfmm_get_stats(tfmm_ctx *ctx, tfmm_stats_t *out);
/* Sorted body keys and, per sorted body, its index in the caller's order. */
int tfmm_get_sorted(tfmm_ctx *ctx, uint64_t *keys, int32_t *perm);
/* Per cell (breadth-first order): key, first sorted body, body count,
 * leaf flag, centre (3 doubles) and width (max over the dimensions).  Any
 * pointer may be null. */
int tfmm_get_cells(tfmm_ctx *ctx, uint64_t *key, int32_t *beg, int32_t *nb, int32_t *leaf,
                   double *center, double *width);
/* M and L of every cell: ncells * 55 complex (re, im) doubles each, or with
 * the Cartesian basis ncells * TFMM_CART_TERMS doubles each. */
int tfmm_get_expansions(tfmm_ctx *ctx, double *M, double *L);
/* Interaction lists in CSR form over target cells: off has ncells + 1
 * entries; src the source cell of each entry (sorted by target, then source).
 * With periodic images a source word is cell | image << 27 (bits of a
 * uint32), image 9 (ix + 1) + 3 (iy + 1) + (iz + 1) for the source seen at
 * (ix, iy, iz) * cycle. */
int tfmm_get_lists(tfmm_ctx *ctx, int32_t *m2l_off, int32_t *m2l_src, int32_t *p2p_off,
                   int32_t *p2p_src);

/* Tuning knobs of the kernels (no counterpart in the reference; results do
 * not depend on them beyond the summation order).
 * TFMM_TUNE_M2L_UNIT: the batched M2L kernel runs each 128-cell target tile
 * as work units of at most factor x (mean items per tile) items, so that a
 * few heavy tiles do not hold the launch back; 0 = one unit per tile. */
#define TFMM_TUNE_M2L_UNIT 1
/* TFMM_TUNE_M2L_XCD_RUN: consecutive work units dealt to one of the 8 XCDs
 * (each has its own L2; neighbouring tiles share source expansions). */
#define TFMM_TUNE_M2L_XCD_RUN 2
/* TFMM_TUNE_P2P_SCHEDULE: where the P2P kernel (second stream) starts.  In
 * tfmm_evaluate, 0 starts it with the numeric phase, beside the upward pass
 * and the M2L kernels; 1 starts it with the M2L tile kernel; 2 as soon as
 * its list is sorted, beside the M2L batch building (also inside the
 * one-launch graph of TFMM_TUNE_GRAPH 2; with TFMM_TUNE_GRAPH 1 as 0); 3 as 2 with the resident FP32 P2P (TFMM_TUNE_P2P_WAVES > 0,
 * mixed precision) holding its few waves until the tile kernel is done, when
 * a full-occupancy launch of the same kernel on the context stream takes the
 * target cells still left on its counter (each cell is computed by one
 * wave either way: the fields do not change); 4 runs it on the context
 * stream with nothing beside it (from phase A, serially: for timing the
 * kernel alone).  In the separate
 * tfmm_traverse, 1 starts it with the tile kernel, any other value as soon
 * as its list is sorted. */
#define TFMM_TUNE_P2P_SCHEDULE 3
/* TFMM_TUNE_GRAPH: 1 records tfmm_evaluate's numeric phase (P2M, M2M, P2P,
 * M2L, L2L, L2P, whose sizes are fixed once the traversal and the M2L
 * batches are built) into one hipGraph and launches it, updating the
 * executable graph in place between evaluations; 0 (default) launches the
 * same kernels one by one (measured equal in time: the GPU is never short of
 * queued work).  2 makes every tfmm_evaluate one hipGraph launch of the
 * whole evaluation (tree build, traversal, list sorts, M2L batching and the
 * numeric phase): the first evaluation with given bodies' count, parameters
 * and knobs runs the ordinary way and sets capacities (its sizes plus an
 * eighth, one tree level more); later ones are one launch whose kernels read
 * the evaluation's sizes on the device, with no host read-back before the
 * end, where tfmm_evaluate waits for the launch and checks the capacities.
 * An evaluation that exceeded one runs again the ordinary way and sets new
 * capacities; the fields are those of the ordinary evaluation bit for bit
 * either way (tfmm_stats_t.one_graph tells which ran).  With a HIP runtime
 * that cannot record timing events inside a capture, the per-phase times
 * inside a graph read 0.  tfmm_upward / _traverse / _downward called
 * separately always launch one by one. */
#define TFMM_TUNE_GRAPH 4
/* TFMM_TUNE_P2P_WAVES: 0 (default) runs the P2P (FP32 or FP64) as one
 * workgroup per cell; k in 1..8 runs it as k resident waves per SIMD that
 * take target leaves from a counter, so it holds a bounded share of the CUs
 * and the kernels beside it keep slots (see TFMM_TUNE_P2P_SCHEDULE 2 and 3;
 * the drain of schedule 3 is the FP32 P2P's). */
#define TFMM_TUNE_P2P_WAVES 5
/* TFMM_TUNE_DEVICE_LIMIT: a limit in bytes (0 = none, the default) on the
 * device memory the library's buffers hold, over all contexts of the
 * process; a call whose buffers would exceed it fails with TFMM_ENOMEM,
 * leaving the context usable for a smaller problem or after the limit is
 * raised.  For deployments that share the GPU, and to test the allocation
 * failure path. */
#define TFMM_TUNE_DEVICE_LIMIT 6
/* TFMM_TUNE_CART_M2L: the M2L of the Cartesian basis (tfmm_set_basis).
 * 0 (default): through the spherical tile kernel, exactly -- the Cartesian
 * multipoles turned into spherical ones per degree (T1), the FP64 tile
 * kernel (and its per-pair FP64 fallback, and the far field of periodic
 * images), the spherical locals turned back (T2); both bases truncate the
 * same series and a contraction with derivatives of 1/r sees only the
 * traceless part of a moment tensor, so this is the Cartesian M2L up to
 * rounding.  1: the direct Cartesian kernel, the reference's formula
 * (5005 multiply-adds per pair) on every pair of the list.  A change makes
 * the expansions again (as tfmm_set_basis). */
#define TFMM_TUNE_CART_M2L 7
int tfmm_set_tuning(tfmm_ctx *ctx, int knob, double value);
int tfmm_get_tuning(tfmm_ctx *ctx, int knob, double *value);
/* Autotuner: evaluates the current bodies with params, reps times per
 * candidate (after 2 uncounted evaluations of the candidate, which let the
 * device settle to the setting), in two stages -- the tile-kernel knobs (unit factor {0, 1.5,
 * 2, 3} x XCD run {8, 16, 32, 64}) at the current P2P setting, then the
 * P2P's schedule {0, 1, 2} x resident waves {0, 3, 4} (waves for mixed
 * precision only, schedule 2 not with TFMM_TUNE_GRAPH 1) at the best tile
 * setting --
 * and keeps the setting whose evaluation was fastest (median).  chosen (4
 * doubles, may be null): unit factor, XCD run, P2P schedule, milliseconds
 * per evaluation; the P2P waves chosen stay set (tfmm_get_tuning). */
int tfmm_autotune(tfmm_ctx *ctx, const tfmm_params *params, int reps, double *chosen);

/* FP64 direct sum on the current bodies for the given target indices (caller
 * order), normalised by the target charge; out: nt * 4 doubles (host).  In
 * a batch each target sums over the bodies of its own problem; with periodic
 * images over the (3^images)^3 copies of the bodies (tfmm_set_periodic). */
int tfmm_direct(tfmm_ctx *ctx, const int64_t *targets, int64_t nt, double *out);

/* tfmm_direct without any device allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The target indices and the fields are staged in a
 * static device array of the library sized for TFMM_DIRECT_NOALLOC_MAX
 * targets (40 bytes each, 2.5 MiB; every GPU the library runs on holds its
 * copy until the process ends); more targets return TFMM_EINVAL.  Same
 * results as tfmm_direct, bit for bit.  All calls on one GPU share that
 * GPU's array, which keeps the last call's contents: calls that overlap on
 * one GPU (from several contexts or host threads) race.  Targets of one
 * problem that are adjacent in `targets` are summed by one launch. */
#define TFMM_DIRECT_NOALLOC_MAX 65536
int tfmm_direct_noalloc(tfmm_ctx *ctx, const int64_t *targets, int64_t nt, double *out);

/* ---------------------------------------------------------------------------
 * Learned parameter picker.  No counterpart in the reference (its ncrit and
 * theta are fixed command-line values, EX/include/args.h:101-158): an FP32
 * MLP, 16 tree statistics -> 64 -> 64 (ReLU) -> 5 logits over ncrit in
 * {16, 32, 64, 128, 256} and theta = 0.2 + 0.4 sigmoid(o5), trained and
 * evaluated by the library's own GPU kernels.
 * ------------------------------------------------------------------------- */
typedef struct tfmm_picker tfmm_picker;

int64_t tfmm_picker_nparam(void);
/* Glorot-uniform initial weights from the seed */
int tfmm_picker_create(int hip_device, uint64_t seed, tfmm_picker **out);
void tfmm_picker_destroy(tfmm_picker *p);
const char *tfmm_picker_last_error(const tfmm_picker *p);
int tfmm_picker_set_weights(tfmm_picker *p, const float *w, int64_t n);
int tfmm_picker_get_weights(tfmm_picker *p, float *w, int64_t n);
/* Upper bound on the theta tfmm_pick_params returns (default 0.6): the
 * largest theta among the training labels, so that the regression head does
 * not extrapolate past the configurations it was trained on. */
int tfmm_picker_set_theta_max(tfmm_picker *p, float theta_max);
/* Full-batch Adam on n samples (16 features each, ncrit class 0..4, theta in
 * [0.2, 0.6]); loss_per_step (optional) receives the mean loss of each step. */
int tfmm_picker_train(tfmm_picker *p, const float *X, const int32_t *cls, const float *theta,
                      int n, int steps, float lr, float *loss_per_step);
/* Mean gradient of the loss over the batch (tfmm_picker_nparam() floats) and
 * per-sample losses; the weights are not changed. */
int tfmm_picker_gradient(tfmm_picker *p, const float *X, const int32_t *cls, const float *theta,
                         int n, float *grad, float *loss);
/* Raw network outputs (6 per sample) */
int tfmm_picker_predict(tfmm_picker *p, const float *X, int n, float *out);

/* The 16 tree statistics of the built tree: log2 N / 23, max depth / 20,
 * fraction of non-empty leaves per depth bin (<=3, 4..9, >=10), mean, spread
 * and third moment of leaf occupancy nb / ncrit, fraction of empty leaves,
 * mean leaf depth / 20, fraction of bodies in leaves at depth >= 8. */
int tfmm_tree_features(tfmm_ctx *ctx, float *feat16);
/* Pick ncrit and theta for the current bodies: builds a trial tree at
 * ncrit = 64, reads its statistics and runs the picker's network; with a null
 * picker (or a non-finite output) the fallback ncrit = 64, theta = 0.4 is
 * returned.  theta is capped by tfmm_picker_set_theta_max.  The precision field of *io is kept. */
int tfmm_pick_params(tfmm_ctx *ctx, tfmm_picker *picker, tfmm_params *io);

#ifdef __cplusplus
}
#endif
#endif /* TFMM_H_ */
