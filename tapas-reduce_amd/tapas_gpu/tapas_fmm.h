// tapas_fmm.h -- the Tapas FMM user surface on top of the MI355X library's C
// ABI (include/tfmm.h): Partition, the Map overloads over cells, subcells,
// bodies and their products, tapas::Product, the Cell and Body layout, and the
// FMM kernel functions the application's functors call.  Header-only C++11;
// link libtfmm.so.
//
// Reference surface it mirrors (TP = cpp/include/tapas,
// EX = sample/exafmm-dev-13274dd4ac68):
//   Body / Source layout            EX/include/types.h:62-74 (96 bytes, X at 0,
//                                   SRC at 24, IBODY at 32, TRG at 56)
//   BodyAttr = kvec4, CellAttr      EX/examples/tapas_exafmm.h:10-17
//   TapasFMM::Partition             TP/hot.h:1349-1352
//   TapasFMM::Map overloads         TP/hot.h:1354-1396 (products, subcells,
//                                   bodies, one cell)
//   tapas::Product                  TP/iterator.h:407-434
//   Cell::Distance(c, tapas::Center) TP/hot.h:427-429, TP/geometry.h:80-83
//   kernel calls of the functors    EX/examples/LaplaceSphericalCPU_tapas.h
//                                   (P2M :138, M2M :165, M2L :208, L2P :275,
//                                   L2L :309), LaplaceP2PCPU_tapas.h:48-74 (P2P)
//   root->local_body(i) /           TP/hot.h:949-1050, used by CopyBackResult
//   local_body_attr(i) / local_nb() (tapas_exafmm.cxx:388-398)
//
// Two ways to drive the GPU:
//  * The application's own functors (FMM_Upward, FMM_DTT, FMM_Downward as
//    written in EX/examples/tapas_exafmm.cxx:87-286, any struct with the
//    same operator() signatures).  A Map from the root runs the functor on
//    host cells, as the reference's mapper does; the kernel functions below
//    only record what is asked of them, and when the root-level Map returns
//    the recorded work is launched on the GPU (the reference's GPUMapper
//    records P2P products and launches them the same way,
//    TP/hot/mapper.h:452-690).  The recorded kernels also tell the mapper
//    which pass the functor is (P2M/M2M: upward, L2L/L2P: downward), as the
//    reference inspects a functor's direction by running it on proxy cells
//    (TP/hot/exact_let.h:322-371).  The traversal's M2L and P2P lists are
//    executed exactly as recorded (tfmm_traverse_lists), so a user-written
//    traversal decides the interaction set.
//  * The tags tapas_gpu::FMM_Upward / FMM_DTT / FMM_Downward, which run the
//    library's own GPU traversal and passes directly.
// Host-side values (Cell::attr(), body attributes) are copies downloaded from
// the device after a pass; writing them does not change the GPU's data.
// Library errors and unsupported calls throw tapas_gpu::Error (the reference
// aborts).
#ifndef TAPAS_GPU_TAPAS_FMM_H_
#define TAPAS_GPU_TAPAS_FMM_H_

#include <unistd.h>  // gethostname

#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/tfmm.h"

namespace tapas_gpu {

typedef double real_t;
typedef std::complex<real_t> complex_t;
typedef int64_t index_t;

// Fixed-size vector with the operations the FMM application uses on
// positions, body attributes and expansion coefficients (EX/include/vec.h).
template <int N, class T>
struct vec {
  T a[N];
  vec() {
    for (int i = 0; i < N; i++) a[i] = T(0);
  }
  vec(const T &s) {  // NOLINT: broadcast, as the reference's vec(const T&)
    for (int i = 0; i < N; i++) a[i] = s;
  }
  vec &operator=(const T &s) {
    for (int i = 0; i < N; i++) a[i] = s;
    return *this;
  }
  T &operator[](int i) { return a[i]; }
  const T &operator[](int i) const { return a[i]; }
  vec &operator+=(const vec &o) {
    for (int i = 0; i < N; i++) a[i] += o.a[i];
    return *this;
  }
  vec &operator-=(const vec &o) {
    for (int i = 0; i < N; i++) a[i] -= o.a[i];
    return *this;
  }
  vec &operator*=(const T &s) {
    for (int i = 0; i < N; i++) a[i] *= s;
    return *this;
  }
  vec operator+(const vec &o) const { return vec(*this) += o; }
  vec operator-(const vec &o) const { return vec(*this) -= o; }
  vec operator*(const T &s) const { return vec(*this) *= s; }
  bool operator==(const vec &o) const {
    for (int i = 0; i < N; i++)
      if (!(a[i] == o.a[i])) return false;
    return true;
  }
  bool operator!=(const vec &o) const { return !(*this == o); }
};
template <int N, class T>
inline T norm(const vec<N, T> &v) {  // squared length, as EX/include/vec.h's norm
  T s = 0;
  for (int i = 0; i < N; i++) s += v[i] * v[i];
  return s;
}

typedef vec<3, real_t> vec3;
typedef vec<4, real_t> kvec4;
typedef vec<TFMM_NTERM, complex_t> vecP;

// The reference's Body (types.h:62-74, built without -mavx).
struct Body {
  vec3 X;
  real_t SRC;
  int IBODY;
  int IRANK;
  uint64_t ICELL;
  real_t WEIGHT;
  kvec4 TRG;
} __attribute__((aligned(16)));
static_assert(sizeof(Body) == 96, "Body must keep the reference's 96-byte layout");
static_assert(offsetof(Body, SRC) == 24 && offsetof(Body, TRG) == 56, "Body field offsets");

typedef kvec4 BodyAttr;  // (potential, ax, ay, az)

struct CellAttr {
  real_t R;
  vecP M;
  vecP L;
};

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// Tags that run the library's own GPU passes (see the file comment).
struct FMM_Upward {};
struct FMM_Downward {};
struct FMM_DTT {};

enum class Precision { FP64 = TFMM_FP64, Mixed = TFMM_MIXED };
// expansion basis (the reference's compile-time BASIS, EX/Makefile.include:13)
enum class Basis { Spherical = TFMM_BASIS_SPHERICAL, Cartesian = TFMM_BASIS_CARTESIAN };

// tapas::Center: Distance between cell centres
struct CenterTag {};

namespace detail {

// What a root-level Map recorded while the application's functor ran on
// host cells.
struct Recorder {
  bool active = false;
  std::vector<int32_t> m2l, p2p;        // (target, source) pairs
  std::vector<uint8_t> p2m, m2m, l2l, l2p;  // per cell
  void start(size_t nc) {
    active = true;
    m2l.clear();
    p2p.clear();
    p2m.assign(nc, 0);
    m2m.assign(nc, 0);
    l2l.assign(nc, 0);
    l2p.assign(nc, 0);
  }
  static bool any(const std::vector<uint8_t> &v) {
    for (uint8_t x : v)
      if (x) return true;
    return false;
  }
};

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
  std::vector<CellAttr> cell_attr;   // host copy, filled on demand
  bool attr_valid = false;
  int stage = 0;
  bool opt_mutual = false;  // root->SetOptMutual (the reference's --mutual)
  Recorder rec;

  ~Tree() {
    if (ctx) tfmm_destroy(ctx);
  }
  void check(int rc, const char *what) {
    if (rc != TFMM_OK)
      throw Error(std::string(what) + ": " + (ctx ? tfmm_last_error(ctx) : "no context"));
  }
  size_t ncells() const { return key.size(); }
  void load_attrs() {
    const size_t nc = ncells();
    cell_attr.assign(nc, CellAttr());
    if (stage >= 2) {
      std::vector<double> M(nc * TFMM_NTERM * 2), L(nc * TFMM_NTERM * 2);
      check(tfmm_get_expansions(ctx, M.data(), stage >= 3 ? L.data() : nullptr),
            "tfmm_get_expansions");
      for (size_t c = 0; c < nc; c++)
        for (int i = 0; i < TFMM_NTERM; i++) {
          const size_t o = 2 * (c * TFMM_NTERM + i);
          cell_attr[c].M[i] = complex_t(M[o], M[o + 1]);
          cell_attr[c].L[i] = complex_t(L[o], L[o + 1]);
        }
    }
    attr_valid = true;
  }
};

}  // namespace detail

// Report files in the reference's RankCSV format (TP/util.h:114-317): one
// header line, then one row per rank, every column right-aligned in 14
// characters (a value that fills its column gets one space after it), doubles
// in scientific notation.  One GPU is one rank.
namespace util {
class RankCSV {
 public:
  explicit RankCSV(std::vector<std::string> cols) : cols_() {
    cols_.push_back("Rank");
    for (auto &c : cols) cols_.push_back(c);
    vals_.assign(cols_.size(), std::string("N/A"));
    At("Rank", 0);
  }
  void At(const std::string &col, double v) {
    std::ostringstream ss;
    ss << std::scientific << v;
    set(col, ss.str());
  }
  void At(const std::string &col, int64_t v) { set(col, std::to_string(v)); }
  void At(const std::string &col, uint64_t v) { set(col, std::to_string(v)); }
  void At(const std::string &col, int v) { set(col, std::to_string(v)); }
  void At(const std::string &col, const std::string &v) { set(col, v); }
  std::string Text() const {
    std::string out;
    for (auto &c : cols_) out += pad(c);
    out += "\n";
    for (size_t i = 0; i < vals_.size(); i++) {
      out += pad(vals_[i]);
      if (i + 1 < vals_.size() && vals_[i].size() >= kWidth) out += " ";
    }
    out += "\n";
    return out;
  }
  void Dump(const std::string &fname) const {
    std::ofstream f(fname);
    if (!f) throw Error("cannot write " + fname);
    f << Text();
  }

 private:
  static const size_t kWidth = 14;
  static std::string pad(const std::string &v) {
    return v.size() < kWidth ? std::string(kWidth - v.size(), ' ') + v : v;
  }
  void set(const std::string &col, const std::string &v) {
    for (size_t i = 0; i < cols_.size(); i++)
      if (cols_[i] == col) {
        vals_[i] = v;
        return;
      }
    throw Error("unknown report column " + col);
  }
  std::vector<std::string> cols_, vals_;
};

// TAPAS_REPORT_PREFIX + name + TAPAS_REPORT_SUFFIX + ".csv", as the
// reference names its report files (TP/hot/report.h:48-58)
inline std::string ReportPath(const std::string &name) {
  const char *p = getenv("TAPAS_REPORT_PREFIX"), *s = getenv("TAPAS_REPORT_SUFFIX");
  return std::string(p ? p : "") + name + std::string(s ? s : "") + ".csv";
}
}  // namespace util

class SubCellRange;
class BodyRange;

class Cell {
 public:
  bool IsLeaf() const { return t_->first_child[idx_] < 0; }
  bool IsRoot() const { return idx_ == 0; }
  index_t nb() const { return t_->nb[idx_]; }
  index_t local_nb() const { return t_->nb[idx_]; }
  int nsubcells() const { return IsLeaf() ? 0 : 8; }
  uint64_t key() const { return t_->key[idx_]; }
  int depth() const { return (int)(key() & 31u); }
  vec3 center() const {
    vec3 c;
    for (int d = 0; d < 3; d++) c[d] = t_->center[3 * idx_ + d];
    return c;
  }
  // region width along dimension d (cells are cubes: the bounding region is a
  // scalar cube, TP/hot/buildtree.h:77-86)
  real_t width(int) const { return t_->width[idx_]; }
  // squared distance of the centres, ((dx^2 + dy^2) + dz^2) as
  // TP/geometry.h:80-83 sums it
  real_t Distance(const Cell &o, CenterTag) const {
    real_t r2 = 0;
    for (int d = 0; d < 3; d++) {
      const real_t dd = t_->center[3 * idx_ + d] - o.t_->center[3 * o.idx_ + d];
      r2 += dd * dd;
    }
    return r2;
  }
  Cell subcell(int i) const {
    if (IsLeaf()) throw Error("Cell::subcell(int) is called for a leaf cell");
    if (i < 0 || i >= 8) throw Error("Cell::subcell(int): index out of range");
    return Cell(t_, t_->first_child[idx_] + i);
  }
  Cell parent() const {
    if (IsRoot()) throw Error("Cell::parent() is called for the root");
    return Cell(t_, t_->parent[idx_]);
  }
  inline SubCellRange subcells() const;
  inline BodyRange bodies() const;
  bool operator==(const Cell &o) const { return t_ == o.t_ && idx_ == o.idx_; }
  bool operator!=(const Cell &o) const { return !(*this == o); }
  // mutual interaction of the whole tree (TP/hot.h:407-409): a product of a
  // cell's subcells with themselves then visits the pairs a <= b only
  bool GetOptMutual() const { return t_->opt_mutual; }
  bool SetOptMutual(bool b) {
    const bool old = t_->opt_mutual;
    t_->opt_mutual = b;
    return old;
  }

  // body i of a leaf (sorted order)
  Body &body(index_t i) { return t_->bodies[t_->beg[idx_] + i]; }
  BodyAttr &body_attr(index_t i) { return t_->attrs[t_->beg[idx_] + i]; }
  // bodies of the whole (root) cell, as used by CopyBackResult
  Body &local_body(index_t i) { return t_->bodies[i]; }
  BodyAttr &local_body_attr(index_t i) { return t_->attrs[i]; }

  // multipole / local expansion: a host copy, downloaded from the device
  // after a pass (zero before the upward pass)
  CellAttr &attr() const {
    if (!t_->attr_valid) t_->load_attrs();
    return t_->cell_attr[idx_];
  }

  std::shared_ptr<detail::Tree> tree() const { return t_; }
  index_t index() const { return idx_; }

  // root->Report(): the reference's report files (TP/hot/report.h:39-152)
  // for the last evaluation.  One GPU is one rank: no sampling, body
  // exchange, global tree or LET, whose columns are 0; the tree build (key
  // sort included) is the local growth; map2's net traversal and device call
  // are the GPU traversal with its M2L and P2P kernels.
  inline void Report() const;

  Cell(std::shared_ptr<detail::Tree> t, index_t i) : t_(std::move(t)), idx_(i) {}

 private:
  std::shared_ptr<detail::Tree> t_;
  index_t idx_;
};

inline void Cell::Report() const {
  tfmm_stats_t st;
  t_->check(tfmm_get_stats(t_->ctx, &st), "tfmm_get_stats");
  char host[200] = {0};
  gethostname(host, sizeof(host) - 1);
  util::RankCSV hn({"Hostname"});
  hn.At("Hostname", std::string(host));
  hn.Dump(util::ReportPath("hostnames"));
  {  // misc (rank 0): no sampling on one GPU
    util::RankCSV misc({"SamplingRate", "NB"});
    misc.At("SamplingRate", 0.0);
    misc.At("NB", (int64_t)t_->n);
    // the reference's misc.csv is a CSV, not a RankCSV: no Rank column
    std::string txt = misc.Text();
    const size_t cut = std::string("          Rank").size();
    std::string out;
    std::istringstream lines(txt);
    for (std::string l; std::getline(lines, l);) out += l.substr(cut) + "\n";
    std::ofstream f(util::ReportPath("misc"));
    if (!f) throw Error("cannot write misc report");
    f << out;
  }
  util::RankCSV lb({"NumBodies1", "NumBodies2", "NumLeaves", "NumCells"});
  lb.At("NumBodies1", (int64_t)t_->n);
  lb.At("NumBodies2", (int64_t)t_->n);
  lb.At("NumLeaves", (int64_t)st.nleaves);
  lb.At("NumCells", (int64_t)st.ncells);
  lb.Dump(util::ReportPath("load_balancing"));
  util::RankCSV tc({"all", "sample", "exchange", "grow-local", "grow-global"});
  tc.At("all", st.ms_tree * 1e-3);
  tc.At("sample", 0.0);
  tc.At("exchange", 0.0);
  tc.At("grow-local", st.ms_tree * 1e-3);
  tc.At("grow-global", 0.0);
  tc.Dump(util::ReportPath("tree_construction"));
  util::RankCSV let({"LET-All", "LET-Trav-main", "LET-Trav-sub", "LET-Req-all", "LET-Req-comm",
                     "LET-Res-all", "LET-Res-acomm", "LET-Res-bcomm", "LET-Reg"});
  for (const char *c : {"LET-All", "LET-Trav-main", "LET-Trav-sub", "LET-Req-all", "LET-Req-comm",
                        "LET-Res-all", "LET-Res-acomm", "LET-Res-bcomm", "LET-Reg"})
    let.At(c, 0.0);
  let.Dump(util::ReportPath("let_construction"));
  util::RankCSV m2({"all", "let", "net_traverse", "device_call"});
  m2.At("all", st.ms_traverse * 1e-3);
  m2.At("let", 0.0);
  m2.At("net_traverse", st.ms_traverse * 1e-3);
  m2.At("device_call", (st.ms_m2l + st.ms_p2p) * 1e-3);
  m2.Dump(util::ReportPath("map2"));
}

// Cell::subcells(): the 8 children of an internal cell
class SubCellRange {
 public:
  explicit SubCellRange(const Cell &c) : cell_(c) {}
  int size() const { return cell_.nsubcells(); }
  Cell operator[](int i) const { return cell_.subcell(i); }
  const Cell &cell() const { return cell_; }

 private:
  Cell cell_;
};

// Cell::bodies(): the bodies of a leaf with their attributes
class BodyRange {
 public:
  explicit BodyRange(const Cell &c) : cell_(c) {}
  index_t size() const { return cell_.nb(); }
  Body &operator[](index_t i) { return cell_.body(i); }
  BodyAttr &attr(index_t i) { return cell_.body_attr(i); }
  const Cell &cell() const { return cell_; }

 private:
  Cell cell_;
};

inline SubCellRange Cell::subcells() const { return SubCellRange(*this); }
inline BodyRange Cell::bodies() const {
  if (!IsLeaf()) throw Error("Cell::bodies() is called for a cell that is not a leaf");
  return BodyRange(*this);
}

// tapas::Product of two ranges (TP/iterator.h:407-434)
template <class A, class B>
struct ProductRange {
  A t1;
  B t2;
};
inline ProductRange<Cell, Cell> Product(const Cell &a, const Cell &b) { return {a, b}; }
inline ProductRange<SubCellRange, Cell> Product(const SubCellRange &a, const Cell &b) {
  return {a, b};
}
inline ProductRange<Cell, SubCellRange> Product(const Cell &a, const SubCellRange &b) {
  return {a, b};
}
inline ProductRange<SubCellRange, SubCellRange> Product(const SubCellRange &a,
                                                        const SubCellRange &b) {
  return {a, b};
}
inline ProductRange<BodyRange, BodyRange> Product(const BodyRange &a, const BodyRange &b) {
  return {a, b};
}
// the CellProduct of the tag interface
typedef ProductRange<Cell, Cell> CellProduct;

// ---------------------------------------------------------------------------
// Kernel functions called by the application's functors (the names and
// signatures of LaplaceSphericalCPU_tapas.h / LaplaceP2PCPU_tapas.h).  Inside
// a root-level Map they record the call; the GPU runs the recorded work.
// ---------------------------------------------------------------------------
namespace detail {
inline Recorder &recorder(const Cell &c, const char *what) {
  Recorder &r = c.tree()->rec;
  if (!r.active)
    throw Error(std::string(what) + " is called outside a TapasFMM::Map from the root");
  return r;
}
inline void no_periodic(const vec3 &xp) {
  if (xp[0] != 0 || xp[1] != 0 || xp[2] != 0)
    throw Error("periodic images are not supported (Xperiodic must be 0)");
}
}  // namespace detail

inline void P2M(Cell &c) {
  if (!c.IsLeaf()) throw Error("P2M is called for a cell that is not a leaf");
  detail::recorder(c, "P2M").p2m[c.index()] = 1;
}
inline void M2M(Cell &c) {
  if (c.IsLeaf()) throw Error("M2M is called for a leaf");
  detail::recorder(c, "M2M").m2m[c.index()] = 1;
}
inline void L2L(Cell &c) {
  if (c.IsRoot()) throw Error("L2L is called for the root");
  detail::recorder(c, "L2L").l2l[c.index()] = 1;
}
// with mutual the call also gives Cj the M2L of Ci's multipole
// (LaplaceSphericalCPU_tapas.h:225): recorded as the reversed pair too
inline void M2L(Cell &Ci, Cell &Cj, vec3 Xperiodic, bool mutual) {
  detail::no_periodic(Xperiodic);
  detail::Recorder &r = detail::recorder(Ci, "M2L");
  r.m2l.push_back((int32_t)Ci.index());
  r.m2l.push_back((int32_t)Cj.index());
  if (mutual) {
    r.m2l.push_back((int32_t)Cj.index());
    r.m2l.push_back((int32_t)Ci.index());
  }
}
// L2P of one body (the reference passes the leaf as a pointer); mapped over a
// leaf's bodies it records the leaf's L2P once
inline void L2P(Body &, BodyAttr &, Cell *c) {
  if (!c) throw Error("L2P needs its leaf");
  detail::recorder(*c, "L2P").l2p[c->index()] = 1;
}
struct P2P {
  template <class B, class A>
  void operator()(B &, A &, B &, A &, vec3, int) const {
    throw Error("P2P runs on the GPU: map it over Product(Ci.bodies(), Cj.bodies())");
  }
};

namespace detail {
template <class F>
struct is_l2p : std::false_type {};
template <>
struct is_l2p<void (*)(Body &, BodyAttr &, Cell *)> : std::true_type {};
template <class F>
struct is_p2p : std::is_same<typename std::decay<F>::type, P2P> {};
}  // namespace detail

struct TapasFMM {
  typedef tapas_gpu::Cell Cell;
  typedef tapas_gpu::Body Body;
  typedef tapas_gpu::BodyAttr BodyAttr;
  typedef tapas_gpu::CellAttr CellAttr;

  // Copies the bodies and builds the octree on the GPU.  theta and the
  // precision are taken at Map(FMM_DTT, ...); the tree does not depend on them.
  // images > 0: periodic images of period cycle (tfmm_set_periodic; the
  // reference's --images, EX/include/args.h, with its cycle 2 pi).  basis:
  // the expansions' basis (tfmm_set_basis).
  static Cell *Partition(Body *b, index_t nb, int ncrit, int device = 0,
                         Precision precision = Precision::FP64, int images = 0,
                         double cycle = 0, Basis basis = Basis::Spherical) {
    auto t = std::make_shared<detail::Tree>();
    t->check(tfmm_create(device, &t->ctx), "tfmm_create");
    if (images > 0) t->check(tfmm_set_periodic(t->ctx, images, cycle), "tfmm_set_periodic");
    t->check(tfmm_set_basis(t->ctx, (int)basis), "tfmm_set_basis");
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
    t->attrs.assign(nb, BodyAttr(0.0));
    t->stage = 1;
    return new Cell(t, 0);
  }

  // ---- the library's own passes (tags) ----------------------------------

  // Map(FMM_Upward(), *root, theta): P2M at the leaves, M2M upward
  static void Map(FMM_Upward, Cell &root, real_t /*theta*/) {
    if (!root.IsRoot()) throw Error("the GPU mapper runs FMM_Upward from the root");
    upward(root.tree());
  }
  // Map(FMM_DTT(), Product(*root, *root), mutual, nspawn, theta): the GPU
  // dual tree traversal with its M2L and P2P kernels
  static void Map(FMM_DTT, CellProduct prod, int mutual, int /*nspawn*/, real_t theta) {
    if (!(prod.t1.IsRoot() && prod.t2.IsRoot()))
      throw Error("the GPU mapper runs FMM_DTT from Product(root, root)");
    auto t = prod.t1.tree();
    t->prm.mutual = mutual ? 1 : 0;
    set_theta(t, theta);
    t->check(tfmm_traverse(t->ctx), "tfmm_traverse");
    t->attr_valid = false;
    t->stage = 3;
  }
  // Map(FMM_Downward(), *root): L2L downward, L2P at the leaves; the fields
  // become visible through local_body_attr(i)
  static void Map(FMM_Downward, Cell &root) {
    if (!root.IsRoot()) throw Error("the GPU mapper runs FMM_Downward from the root");
    downward(root.tree());
  }

  // ---- the application's functors (TP/hot.h:1354-1396) ------------------

  // One cell.  From the root (outside any other Map) the functor runs on the
  // host cells with its kernel calls recorded, and the recorded pass runs on
  // the GPU when it returns; inside a Map the functor is simply called.
  template <class F, class... Args>
  static void Map(F f, Cell &c, Args... args) {
    auto t = c.tree();
    if (t->rec.active) {
      f(c, args...);
      return;
    }
    if (!c.IsRoot()) throw Error("the GPU mapper starts a pass from the root");
    t->rec.start(t->ncells());
    try {
      f(c, args...);
    } catch (...) {
      t->rec.active = false;
      throw;
    }
    t->rec.active = false;
    run_recorded_pass(t);
  }

  // Subcells of a cell (inside a Map)
  template <class F, class... Args>
  static void Map(F f, SubCellRange r, Args... args) {
    for (int i = 0; i < r.size(); i++) {
      Cell ci = r[i];
      Map(f, ci, args...);
    }
  }

  // Bodies of a leaf: L2P is recorded for the leaf; any other body functor
  // runs on the host copies of the bodies
  template <class F, class... Args>
  static void Map(F f, BodyRange r, Args... args) {
    map_bodies(f, r, typename detail::is_l2p<F>::type(), args...);
  }

  // Cell products: from Product(root, root) outside any Map, the functor is
  // a traversal whose M2L calls and P2P products are recorded, then run by
  // the GPU kernels; inside a Map every pair of the product is visited.
  template <class F, class... Args>
  static void Map(F f, ProductRange<Cell, Cell> p, Args... args) {
    auto t = p.t1.tree();
    if (t->rec.active) {
      f(p.t1, p.t2, args...);
      return;
    }
    if (!(p.t1.IsRoot() && p.t2.IsRoot()))
      throw Error("the GPU mapper starts a traversal from Product(root, root)");
    t->rec.start(t->ncells());
    try {
      f(p.t1, p.t2, args...);
    } catch (...) {
      t->rec.active = false;
      throw;
    }
    t->rec.active = false;
    run_recorded_pass(t);
  }
  template <class F, class... Args>
  static void Map(F f, ProductRange<SubCellRange, Cell> p, Args... args) {
    for (int i = 0; i < p.t1.size(); i++) {
      Cell ci = p.t1[i], cj = p.t2;
      f(ci, cj, args...);
    }
  }
  template <class F, class... Args>
  static void Map(F f, ProductRange<Cell, SubCellRange> p, Args... args) {
    for (int j = 0; j < p.t2.size(); j++) {
      Cell ci = p.t1, cj = p.t2[j];
      f(ci, cj, args...);
    }
  }
  // with the tree's mutual option, the subcells of one cell with themselves
  // give the pairs a <= b only (TP/hot/mapper.h:29-121)
  template <class F, class... Args>
  static void Map(F f, ProductRange<SubCellRange, SubCellRange> p, Args... args) {
    const bool tri = p.t1.cell().GetOptMutual() && p.t1.cell() == p.t2.cell();
    for (int i = 0; i < p.t1.size(); i++)
      for (int j = tri ? i : 0; j < p.t2.size(); j++) {
        Cell ci = p.t1[i], cj = p.t2[j];
        f(ci, cj, args...);
      }
  }

  // Body products (the P2P of two leaves): recorded for the GPU P2P kernel;
  // an empty product is skipped, as TP/hot.h:1367 skips it.  With mutual the
  // source leaf gets the pair's interaction too (LaplaceP2PCPU_tapas.h:66-71):
  // recorded as the reversed leaf pair (a leaf with itself needs no second
  // entry: the kernel gives every body of it all its pairs).
  template <class... Args>
  static void Map(P2P, ProductRange<BodyRange, BodyRange> p, vec3 Xperiodic, int mutual,
                  Args...) {
    detail::no_periodic(Xperiodic);
    if (p.t1.size() == 0 || p.t2.size() == 0) return;
    detail::Recorder &r = detail::recorder(p.t1.cell(), "Map(P2P, Product(bodies, bodies))");
    const int32_t i = (int32_t)p.t1.cell().index(), j = (int32_t)p.t2.cell().index();
    r.p2p.push_back(i);
    r.p2p.push_back(j);
    if (mutual && i != j) {
      r.p2p.push_back(j);
      r.p2p.push_back(i);
    }
  }

  static tfmm_stats_t Stats(Cell &root) {
    tfmm_stats_t st;
    root.tree()->check(tfmm_get_stats(root.tree()->ctx, &st), "tfmm_get_stats");
    return st;
  }

 private:
  static void set_theta(const std::shared_ptr<detail::Tree> &t, real_t theta) {
    t->prm.theta = theta;
    t->check(tfmm_set_params(t->ctx, &t->prm), "tfmm_set_params");
  }
  static void upward(const std::shared_ptr<detail::Tree> &t) {
    t->check(tfmm_upward(t->ctx), "tfmm_upward");
    t->attr_valid = false;
    t->stage = 2;
  }
  static void downward(const std::shared_ptr<detail::Tree> &t) {
    t->check(tfmm_downward(t->ctx), "tfmm_downward");
    std::vector<double> f(4 * (size_t)t->n);
    t->check(tfmm_get_fields(t->ctx, f.data(), 1, 0), "tfmm_get_fields");
    for (index_t i = 0; i < t->n; i++)
      for (int d = 0; d < 4; d++) t->attrs[i][d] = f[4 * i + d];
    t->attr_valid = false;
    t->stage = 4;
  }

  template <class F, class... Args>
  static void map_bodies(F, BodyRange r, std::true_type, Args...) {
    if (r.size() == 0) return;
    detail::recorder(r.cell(), "Map(L2P, bodies)").l2p[r.cell().index()] = 1;
  }
  template <class F, class... Args>
  static void map_bodies(F f, BodyRange r, std::false_type, Args... args) {
    for (index_t i = 0; i < r.size(); i++) f(r[i], r.attr(i), args...);
  }

  // The pass the recorded kernels make up, launched on the GPU.  A pass must
  // be whole (the GPU kernels treat every cell): the upward pass has P2M at
  // every leaf and M2M at every internal cell, the downward pass L2L at every
  // cell but the root and L2P at every non-empty leaf.
  static void run_recorded_pass(const std::shared_ptr<detail::Tree> &t) {
    detail::Recorder &r = t->rec;
    const bool up = r.any(r.p2m) || r.any(r.m2m);
    const bool down = r.any(r.l2l) || r.any(r.l2p);
    const bool trav = !r.m2l.empty() || !r.p2p.empty();
    if ((int)up + (int)down + (int)trav > 1)
      throw Error("a functor mixes kernels of different passes (upward, traversal, downward)");
    const size_t nc = t->ncells();
    if (up) {
      for (size_t c = 0; c < nc; c++) {
        const bool leaf = t->first_child[c] < 0;
        if (leaf ? !r.p2m[c] : !r.m2m[c])
          throw Error("the GPU mapper runs whole upward passes: a cell got no P2M/M2M");
      }
      upward(t);
    } else if (down) {
      if (t->stage < 3) throw Error("downward pass before the traversal");
      for (size_t c = 1; c < nc; c++) {
        const bool leaf = t->first_child[c] < 0;
        if (!r.l2l[c] || (leaf && t->nb[c] > 0 && !r.l2p[c]))
          throw Error("the GPU mapper runs whole downward passes: a cell got no L2L/L2P");
      }
      downward(t);
    } else if (trav) {
      if (t->stage < 2) throw Error("traversal before the upward pass");
      // the counters of a mutual traversal count each recorded pair once
      t->prm.mutual = t->opt_mutual ? 1 : 0;
      t->check(tfmm_set_params(t->ctx, &t->prm), "tfmm_set_params");
      t->check(tfmm_traverse_lists(t->ctx, r.m2l.data(), (int64_t)(r.m2l.size() / 2),
                                   r.p2p.data(), (int64_t)(r.p2p.size() / 2)),
               "tfmm_traverse_lists");
      t->attr_valid = false;
      t->stage = 3;
    }
  }
};

}  // namespace tapas_gpu

// The reference's tapas::Product and tapas::Center spellings.
namespace tapas {
using tapas_gpu::Product;
constexpr tapas_gpu::CenterTag Center{};
}  // namespace tapas

#endif  // TAPAS_GPU_TAPAS_FMM_H_
