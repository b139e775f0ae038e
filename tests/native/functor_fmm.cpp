// functor_fmm.cpp -- FMM functors in the reference application's shape
// (sample/exafmm-dev-13274dd4ac68/examples/tapas_exafmm.cxx:87-286: the same
// operator() signatures, kernel calls and TapasFMM::Map calls) run on the
// MI355X library through tapas_gpu/tapas_exafmm_gpu.h, against the library's
// own GPU traversal on the same bodies.
//
//   functor_fmm <in> <out> <ncrit> <theta> [mutual]
// in: int64 n, then n doubles each of x, y, z, q.  out: n * 4 doubles, the
// fields of the functor run in input order.  Prints the counts of both runs;
// exit status 0 when both runs have the same interaction lists and fields
// within relative 1e-12, 1 otherwise, 2 on an error.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../tapas-reduce_amd/tapas_gpu/tapas_exafmm_gpu.h"

// Upward pass: reset, then P2M at a leaf or the children first and M2M.
struct FMM_Upward {
  template <class Cell>
  inline void operator()(Cell &c, real_t theta) {
    auto attr = c.attr();
    attr.R = 0;
    attr.M = 0;
    attr.L = 0;
    c.attr() = attr;
    if (c.IsLeaf()) {
      P2M(c);
    } else {
      TapasFMM::Map(*this, c.subcells(), theta);
      M2M(c);
    }
  }
};

// Downward pass: L2L below the root, L2P over a non-empty leaf's bodies.
struct FMM_Downward {
  template <class Cell>
  inline void operator()(Cell &c) {
    if (!c.IsRoot()) L2L(c);
    if (c.IsLeaf()) {
      if (c.nb() > 0) TapasFMM::Map(L2P, c.bodies(), &c);
    } else {
      TapasFMM::Map(*this, c.subcells());
    }
  }
};

// Dual tree traversal: the multipole acceptance test on the centres' squared
// distance against the two radii (width / 2 * 1.00001 / theta), M2L when it
// passes, P2P of two leaves, otherwise a split of the larger cell.
struct FMM_DTT {
  template <class Cell>
  inline void operator()(Cell &Ci, Cell &Cj, int mutual, int nspawn, real_t theta) {
    const real_t R2 = Ci.Distance(Cj, tapas::Center);
    vec3 Xperiodic = 0;
    real_t Ri = 0, Rj = 0;
    for (int d = 0; d < 3; d++) {
      Ri = std::max(Ri, Ci.width(d));
      Rj = std::max(Rj, Cj.width(d));
    }
    Ri = (Ri / 2 * 1.00001) / theta;
    Rj = (Rj / 2 * 1.00001) / theta;
    if (R2 > (Ri + Rj) * (Ri + Rj)) {
      M2L(Ci, Cj, Xperiodic, mutual);
    } else if (Ci.IsLeaf() && Cj.IsLeaf()) {
      TapasFMM::Map(P2P(), tapas::Product(Ci.bodies(), Cj.bodies()), Xperiodic, mutual);
    } else {
      split(Ci, Cj, Ri, Rj, mutual, nspawn, theta);
    }
  }
  template <class Cell>
  inline void split(Cell &Ci, Cell &Cj, real_t Ri, real_t Rj, int mutual, int nspawn,
                    real_t theta) {
    if (Cj.IsLeaf()) {
      TapasFMM::Map(*this, tapas::Product(Ci.subcells(), Cj), mutual, nspawn, theta);
    } else if (Ci.IsLeaf()) {
      TapasFMM::Map(*this, tapas::Product(Ci, Cj.subcells()), mutual, nspawn, theta);
    } else if (mutual && Ci == Cj) {
      TapasFMM::Map(*this, tapas::Product(Ci.subcells(), Cj.subcells()), mutual, nspawn, theta);
    } else if (Ri >= Rj) {
      TapasFMM::Map(*this, tapas::Product(Ci.subcells(), Cj), mutual, nspawn, theta);
    } else {
      TapasFMM::Map(*this, tapas::Product(Ci, Cj.subcells()), mutual, nspawn, theta);
    }
  }
};

struct Run {
  std::vector<double> fields;  // input order
  std::vector<int32_t> m2l_off, m2l_src, p2p_off, p2p_src;
  tfmm_stats_t st;
};

static void collect(TapasFMM::Cell *root, int64_t n, Run &r) {
  auto t = root->tree();
  r.st = TapasFMM::Stats(*root);
  const size_t nc = (size_t)r.st.ncells;
  r.m2l_off.resize(nc + 1);
  r.p2p_off.resize(nc + 1);
  t->check(tfmm_get_lists(t->ctx, r.m2l_off.data(), nullptr, r.p2p_off.data(), nullptr),
           "tfmm_get_lists");
  r.m2l_src.resize((size_t)r.m2l_off[nc] + 1);
  r.p2p_src.resize((size_t)r.p2p_off[nc] + 1);
  t->check(tfmm_get_lists(t->ctx, r.m2l_off.data(), r.m2l_src.data(), r.p2p_off.data(),
                          r.p2p_src.data()),
           "tfmm_get_lists");
  r.fields.assign(4 * (size_t)n, 0);
  for (int64_t i = 0; i < n; i++) {
    const Body &b = root->local_body(i);
    for (int d = 0; d < 4; d++) r.fields[4 * (size_t)b.IBODY + d] = root->local_body_attr(i)[d];
  }
}

int main(int argc, char **argv) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "usage: functor_fmm <in> <out> <ncrit> <theta> [mutual]\n");
    return 2;
  }
  const int ncrit = atoi(argv[3]);
  const double theta = atof(argv[4]);
  const int mutual = argc > 5 ? atoi(argv[5]) : 0;
  FILE *fi = fopen(argv[1], "rb");
  if (!fi) return 2;
  int64_t n = 0;
  if (fread(&n, 8, 1, fi) != 1 || n <= 0) return 2;
  std::vector<double> a(4 * (size_t)n);
  if (fread(a.data(), 8, a.size(), fi) != a.size()) return 2;
  fclose(fi);
  std::vector<Body> bodies((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) bodies[i].X[d] = a[(size_t)d * n + i];
    bodies[i].SRC = a[3 * (size_t)n + i];
    bodies[i].IBODY = (int)i;
  }
  Run gpu, fun;
  try {
    // the library's own traversal
    TapasFMM::Cell *r0 = TapasFMM::Partition(bodies.data(), n, ncrit);
    r0->SetOptMutual(mutual);
    TapasFMM::Map(tapas_gpu::FMM_Upward(), *r0, theta);
    TapasFMM::Map(tapas_gpu::FMM_DTT(), tapas::Product(*r0, *r0), mutual, 1000, theta);
    TapasFMM::Map(tapas_gpu::FMM_Downward(), *r0);
    collect(r0, n, gpu);
    delete r0;
    // the functors above, recorded on host cells and run by the GPU kernels
    TapasFMM::Cell *r1 = TapasFMM::Partition(bodies.data(), n, ncrit);
    r1->SetOptMutual(mutual);
    TapasFMM::Map(FMM_Upward(), *r1, theta);
    TapasFMM::Map(FMM_DTT(), tapas::Product(*r1, *r1), mutual, 1000, theta);
    TapasFMM::Map(FMM_Downward(), *r1);
    collect(r1, n, fun);
    delete r1;
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  printf("gpu traversal: cells %lld m2l %llu p2p leaf pairs %llu body pairs %llu\n",
         (long long)gpu.st.ncells, (unsigned long long)gpu.st.n_m2l,
         (unsigned long long)gpu.st.n_p2p_leafpairs, (unsigned long long)gpu.st.n_p2p_calls);
  printf("functors:      cells %lld m2l %llu p2p leaf pairs %llu body pairs %llu\n",
         (long long)fun.st.ncells, (unsigned long long)fun.st.n_m2l,
         (unsigned long long)fun.st.n_p2p_leafpairs, (unsigned long long)fun.st.n_p2p_calls);
  const bool lists = gpu.m2l_off == fun.m2l_off && gpu.m2l_src == fun.m2l_src &&
                     gpu.p2p_off == fun.p2p_off && gpu.p2p_src == fun.p2p_src &&
                     gpu.st.n_m2l == fun.st.n_m2l && gpu.st.n_p2p_calls == fun.st.n_p2p_calls;
  double num = 0, den = 0;
  for (size_t i = 0; i < gpu.fields.size(); i++) {
    num += (gpu.fields[i] - fun.fields[i]) * (gpu.fields[i] - fun.fields[i]);
    den += gpu.fields[i] * gpu.fields[i];
  }
  const double rel = std::sqrt(num / (den > 0 ? den : 1));
  printf("lists equal: %d, fields rel-L2 difference %.3e\n", (int)lists, rel);
  FILE *fo = fopen(argv[2], "wb");
  if (!fo || fwrite(fun.fields.data(), 8, fun.fields.size(), fo) != fun.fields.size()) return 2;
  fclose(fo);
  return (lists && rel <= 1e-12) ? 0 : 1;
}
