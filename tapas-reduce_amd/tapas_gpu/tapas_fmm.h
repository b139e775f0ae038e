// tapas_fmm.h -- the Tapas FMM user surface (Partition / Map / Product, the
// Cell and Body layout and the FMM functor names) on top of the MI355X
// library's C ABI (include/tfmm.h).  Header-only C++11; link libtfmm.so.
//
// Reference surface it mirrors (TP = cpp/include/tapas,
// EX = sample/exafmm-dev-13274dd4ac68):
//   Body / Source layout            EX/include/types.h:62-74 (96 bytes, X at 0,
//                                   SRC at 24, IBODY at 32, TRG at 56)
//   BodyAttr = kvec4, CellAttr      EX/examples/tapas_exafmm.h:10-17
//   TapasFMM::Partition             TP/hot.h:1349-1352
//   TapasFMM::Map overloads         TP/hot.h:1354-1396
//   tapas::Product                  TP/iterator.h:407-434
//   FMM_Upward / FMM_DTT /          EX/examples/tapas_exafmm.cxx:87-140, 209-286
//   FMM_Downward
//   root->local_body(i) /           TP/hot.h:949-1050, used by CopyBackResult
//   local_body_attr(i) / local_nb() (tapas_exafmm.cxx:388-398)
//
// The reference's mapper runs the user functor on host cells (or records P2P
// products for its CUDA vectormap, TP/hot/mapper.h:452-690).  Here the
// mapper recognises the three FMM functor types and runs the corresponding
// phase of the GPU library instead: the functor objects are tags, their
// bodies are the HIP kernels.  Errors of the library are thrown as
// tapas_gpu::Error (the reference aborts).
#ifndef TAPAS_GPU_TAPAS_FMM_H_
#define TAPAS_GPU_TAPAS_FMM_H_

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tfmm.h"

namespace tapas_gpu {

typedef double real_t;
typedef std::complex<real_t> complex_t;
typedef int64_t index_t;

// The reference's Body (types.h:62-74, built without -mavx).
struct Body {
  real_t X[3];
  real_t SRC;
  int IBODY;
  int IRANK;
  uint64_t ICELL;
  real_t WEIGHT;
  real_t TRG[4];
} __attribute__((aligned(16)));
static_assert(sizeof(Body) == 96, "Body must keep the reference's 96-byte layout");
static_assert(offsetof(Body, SRC) == 24 && offsetof(Body, TRG) == 56, "Body field offsets");

typedef std::array<real_t, 4> BodyAttr;  // kvec4: (potential, ax, ay, az)

struct CellAttr {
  real_t R;
  complex_t M[TFMM_NTERM];
  complex_t L[TFMM_NTERM];
};

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// FMM functors: the user-visible names of the reference application.
struct FMM_Upward {};
struct FMM_Downward {};
struct FMM_DTT {};

enum class Precision { FP64 = TFMM_FP64, Mixed = TFMM_MIXED };

namespace detail {

struct Tree {
  tfmm_ctx *ctx = nullptr;
  tfmm_params prm{};
  index_t n = 0;
  std::vector<Body> bodies;        // sorted order, as root->local_body(i)
  std::vector<BodyAttr> attrs;     // sorted order, as root->local_body_attr(i)
  std::vector<uint64_t> key;
  std::vector<int32_t> beg, nb, leaf;
  std::vector<double> center, width;
  std::vector<int32_t> first_child;  // -1 for leaves
  std::vector<int32_t> parent;
  std::vector<CellAttr> cell_attr;   // filled on demand
  bool attr_valid = false;
  int stage = 0;

  ~Tree() {
    if (ctx) tfmm_destroy(ctx);
  }
  void check(int rc, const char *what) {
    if (rc != TFMM_OK)
      throw Error(std::string(what) + ": " + (ctx ? tfmm_last_error(ctx) : "no context"));
  }
};

}  // namespace detail

class Cell {
 public:
  bool IsLeaf() const { return t_->first_child[idx_] < 0; }
  bool IsRoot() const { return idx_ == 0; }
  index_t nb() const { return t_->nb[idx_]; }
  index_t local_nb() const { return t_->nb[idx_]; }
  int nsubcells() const { return IsLeaf() ? 0 : 8; }
  uint64_t key() const { return t_->key[idx_]; }
  int depth() const { return (int)(key() & 31u); }
  std::array<real_t, 3> center() const {
    return {{t_->center[3 * idx_], t_->center[3 * idx_ + 1], t_->center[3 * idx_ + 2]}};
  }
  // max over the dimensions of the region width (the MAC radius uses it)
  real_t width(int) const { return t_->width[idx_]; }
  Cell subcell(int i) const {
    if (IsLeaf()) throw Error("Cell::subcell called for a leaf cell");
    return Cell(t_, t_->first_child[idx_] + i);
  }
  Cell parent() const {
    if (IsRoot()) throw Error("Cell::parent called for the root");
    return Cell(t_, t_->parent[idx_]);
  }
  bool operator==(const Cell &o) const { return t_ == o.t_ && idx_ == o.idx_; }

  // body i of a leaf (sorted order)
  Body &body(index_t i) { return t_->bodies[t_->beg[idx_] + i]; }
  BodyAttr &body_attr(index_t i) { return t_->attrs[t_->beg[idx_] + i]; }
  // bodies of the whole (root) cell, as used by CopyBackResult
  Body &local_body(index_t i) { return t_->bodies[i]; }
  BodyAttr &local_body_attr(index_t i) { return t_->attrs[i]; }

  // multipole / local expansion (downloaded from the device on first use)
  const CellAttr &attr() const {
    if (!t_->attr_valid) {
      const size_t nc = t_->key.size();
      std::vector<double> M(nc * TFMM_NTERM * 2), L(nc * TFMM_NTERM * 2);
      const bool have_l = t_->stage >= 3;
      t_->check(tfmm_get_expansions(t_->ctx, M.data(), have_l ? L.data() : nullptr),
                "tfmm_get_expansions");
      t_->cell_attr.assign(nc, CellAttr());
      for (size_t c = 0; c < nc; c++) {
        for (int i = 0; i < TFMM_NTERM; i++) {
          const size_t o = 2 * (c * TFMM_NTERM + i);
          t_->cell_attr[c].M[i] = complex_t(M[o], M[o + 1]);
          t_->cell_attr[c].L[i] = complex_t(L[o], L[o + 1]);
        }
        t_->cell_attr[c].R = 0;
      }
      t_->attr_valid = true;
    }
    return t_->cell_attr[idx_];
  }

  std::shared_ptr<detail::Tree> tree() const { return t_; }
  index_t index() const { return idx_; }

  Cell(std::shared_ptr<detail::Tree> t, index_t i) : t_(std::move(t)), idx_(i) {}

 private:
  std::shared_ptr<detail::Tree> t_;
  index_t idx_;
};

// tapas::Product of two cells (the only product the FMM driver maps over)
struct CellProduct {
  Cell a, b;
};
inline CellProduct Product(Cell &a, Cell &b) { return CellProduct{a, b}; }

struct TapasFMM {
  typedef tapas_gpu::Cell Cell;
  typedef tapas_gpu::Body Body;

  // Copies the bodies and builds the octree on the GPU.  theta and the
  // precision are taken at Map(FMM_DTT, ...); the tree does not depend on them.
  static Cell *Partition(Body *b, index_t nb, int ncrit, int device = 0,
                         Precision precision = Precision::FP64) {
    auto t = std::make_shared<detail::Tree>();
    t->check(tfmm_create(device, &t->ctx), "tfmm_create");
    tfmm_body_desc d;
    d.n = nb;
    d.stride = sizeof(Body);
    d.off_x = offsetof(Body, X);
    d.off_q = offsetof(Body, SRC);
    t->check(tfmm_set_bodies(t->ctx, b, &d), "tfmm_set_bodies");
    t->prm.ncrit = ncrit;
    t->prm.mutual = 0;
    t->prm.precision = (int32_t)precision;
    t->prm.theta = 0.4;
    t->check(tfmm_build_tree(t->ctx, &t->prm), "tfmm_build_tree");
    t->n = nb;
    tfmm_stats_t st;
    t->check(tfmm_get_stats(t->ctx, &st), "tfmm_get_stats");
    const size_t nc = (size_t)st.ncells;
    t->key.resize(nc);
    t->beg.resize(nc);
    t->nb.resize(nc);
    t->leaf.resize(nc);
    t->center.resize(3 * nc);
    t->width.resize(nc);
    t->check(tfmm_get_cells(t->ctx, t->key.data(), t->beg.data(), t->nb.data(), t->leaf.data(),
                            t->center.data(), t->width.data()),
             "tfmm_get_cells");
    // children of a cell are adjacent and levels are contiguous: the first
    // child of the i-th internal cell of the breadth-first order is 1 + 8 i
    t->first_child.assign(nc, -1);
    t->parent.assign(nc, -1);
    int32_t next = 1;
    for (size_t c = 0; c < nc; c++) {
      if (!t->leaf[c]) {
        t->first_child[c] = next;
        for (int k = 0; k < 8; k++) t->parent[next + k] = (int32_t)c;
        next += 8;
      }
    }
    std::vector<uint64_t> keys(nb);
    std::vector<int32_t> perm(nb);
    t->check(tfmm_get_sorted(t->ctx, keys.data(), perm.data()), "tfmm_get_sorted");
    t->bodies.resize(nb);
    for (index_t i = 0; i < nb; i++) t->bodies[i] = b[perm[i]];
    t->attrs.assign(nb, BodyAttr{{0, 0, 0, 0}});
    t->stage = 1;
    return new Cell(t, 0);
  }

  // Map(FMM_Upward(), *root, theta): P2M at the leaves, M2M upward
  static void Map(FMM_Upward, Cell &root, real_t /*theta*/) {
    auto t = root.tree();
    if (!root.IsRoot()) throw Error("the GPU mapper runs FMM_Upward from the root");
    t->check(tfmm_upward(t->ctx), "tfmm_upward");
    t->attr_valid = false;
    t->stage = 2;
  }

  // Map(FMM_DTT(), Product(*root, *root), mutual, nspawn, theta): the dual
  // tree traversal with its M2L and P2P kernels
  static void Map(FMM_DTT, CellProduct prod, int mutual, int /*nspawn*/, real_t theta) {
    if (!(prod.a.IsRoot() && prod.b.IsRoot()))
      throw Error("the GPU mapper runs FMM_DTT from Product(root, root)");
    if (mutual) throw Error("mutual interaction is not supported (mutual must be 0)");
    auto t = prod.a.tree();
    t->prm.theta = theta;
    t->check(tfmm_set_params(t->ctx, &t->prm), "tfmm_set_params");
    t->check(tfmm_traverse(t->ctx), "tfmm_traverse");
    t->attr_valid = false;
    t->stage = 3;
  }

  // Map(FMM_Downward(), *root): L2L downward, L2P at the leaves; the fields
  // become visible through local_body_attr(i)
  static void Map(FMM_Downward, Cell &root) {
    auto t = root.tree();
    if (!root.IsRoot()) throw Error("the GPU mapper runs FMM_Downward from the root");
    t->check(tfmm_downward(t->ctx), "tfmm_downward");
    std::vector<double> f(4 * (size_t)t->n);
    t->check(tfmm_get_fields(t->ctx, f.data(), 1, 0), "tfmm_get_fields");
    for (index_t i = 0; i < t->n; i++)
      for (int d = 0; d < 4; d++) t->attrs[i][d] = f[4 * i + d];
    t->attr_valid = false;
    t->stage = 4;
  }

  static tfmm_stats_t Stats(Cell &root) {
    tfmm_stats_t st;
    root.tree()->check(tfmm_get_stats(root.tree()->ctx, &st), "tfmm_get_stats");
    return st;
  }
};

}  // namespace tapas_gpu

#endif  // TAPAS_GPU_TAPAS_FMM_H_
