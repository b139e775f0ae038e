// ref_driver.cxx -- runs the reference's own TapasFMM pipeline on bodies read
// from a file and dumps what the oracle and the HIP library are checked
// against.  TEST INFRASTRUCTURE ONLY; built by oracle/Makefile into
// oracle/_ref/ from the reference sources where they lie.
//
// The reference application file is compiled unchanged; only its main() is
// renamed so that this driver can call the same functors
// (FMM_Upward / FMM_DTT / FMM_Downward, EX/examples/tapas_exafmm.cxx:87-286)
// in the same sequence as its main (tapas_exafmm.cxx:557-641), with the
// mutual flag of the command line (default 0; the reference's main sets it
// with root->SetOptMutual(args.mutual), tapas_exafmm.cxx:570).
//
// Input file (little endian):  int64 n, int32 ncrit, double theta,
//                              double x[n], y[n], z[n], q[n]
// Output file:                 see the end of main() below.
//
// Under mpiexec with P > 1 ranks it only times the pipeline: rank r takes
// bodies [r n / P, (r + 1) n / P) (the reference's Partition then samples
// and redistributes them, TP/hot.h:1240-1284), every phase is bracketed by
// barriers, and rank 0 writes int64 n, int32 P and the four phase times
// (the maxima over the ranks) instead of the dump.
#define main tapas_reference_app_main
#include "tapas_exafmm.cxx"
#undef main

#include <algorithm>
#include <cstdio>
#include <vector>

namespace {

template <class T>
void put(FILE *f, const T &v) { fwrite(&v, sizeof(T), 1, f); }
template <class T>
void put_n(FILE *f, const T *v, size_t n) { fwrite(v, sizeof(T), n, f); }

struct CellRec {
  uint64_t key;
  int64_t nb;
  int64_t bid;
  int32_t leaf;
  const TapasFMM::Cell *cell;
};

}  // namespace

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  if (argc < 3) {
    fprintf(stderr, "usage: %s in.bin out.bin [dump_expansions=0|1 [mutual=0|1]]\n", argv[0]);
    return 2;
  }
  const bool dump_exp = argc > 3 && atoi(argv[3]) != 0;
  const int mutual = argc > 4 ? atoi(argv[4]) : 0;
  FILE *fi = fopen(argv[1], "rb");
  if (!fi) { perror("input"); return 2; }
  int64_t n = 0; int32_t ncrit = 0; double theta = 0;
  if (fread(&n, 8, 1, fi) != 1 || fread(&ncrit, 4, 1, fi) != 1 || fread(&theta, 8, 1, fi) != 1) return 2;
  std::vector<double> x(n), y(n), z(n), q(n);
  if (fread(x.data(), 8, n, fi) != (size_t)n || fread(y.data(), 8, n, fi) != (size_t)n ||
      fread(z.data(), 8, n, fi) != (size_t)n || fread(q.data(), 8, n, fi) != (size_t)n) return 2;
  fclose(fi);
  int rank = 0, nranks = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &nranks);
  const int64_t lo = n * rank / nranks, hi = n * (rank + 1) / nranks;

  Bodies bodies(hi - lo);
  for (int64_t i = lo; i < hi; i++) {
    Body &b = bodies[i - lo];
    b.X[0] = x[i]; b.X[1] = y[i]; b.X[2] = z[i];
    b.SRC = q[i];
    b.IBODY = (int)i;
    b.IRANK = rank;
    b.ICELL = 0;
    b.WEIGHT = 1;
    b.TRG = 0;
  }

  if (nranks > 1) {  // timing only
    double t[5];
    MPI_Barrier(MPI_COMM_WORLD);
    t[0] = GetTime();
    TapasFMM::Cell *root = TapasFMM::Partition(bodies.data(), bodies.size(), ncrit);
    MPI_Barrier(MPI_COMM_WORLD);
    t[1] = GetTime();
    root->SetOptMutual(0);
    TapasFMM::Map(FMM_Upward(), *root, (real_t)theta);
    MPI_Barrier(MPI_COMM_WORLD);
    t[2] = GetTime();
    TapasFMM::Map(FMM_DTT(), tapas::Product(*root, *root), 0, 1000, (real_t)theta);
    MPI_Barrier(MPI_COMM_WORLD);
    t[3] = GetTime();
    TapasFMM::Map(FMM_Downward(), *root);
    MPI_Barrier(MPI_COMM_WORLD);
    t[4] = GetTime();
    double ph[4] = {t[1] - t[0], t[2] - t[1], t[3] - t[2], t[4] - t[3]}, mx[4];
    MPI_Reduce(ph, mx, 4, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    if (rank == 0) {
      FILE *fo = fopen(argv[2], "wb");
      if (!fo) { perror("output"); return 2; }
      put(fo, (int64_t)n);
      put(fo, (int32_t)nranks);
      for (int d = 0; d < 4; d++) put(fo, mx[d]);
      fclose(fo);
      printf("ref: n=%lld ranks=%d tree=%.3f up=%.3f dtt=%.3f down=%.3f\n", (long long)n,
             nranks, mx[0], mx[1], mx[2], mx[3]);
    }
    MPI_Finalize();
    return 0;
  }

  double t0 = GetTime();
  TapasFMM::Cell *root = TapasFMM::Partition(bodies.data(), bodies.size(), ncrit);
  double t1 = GetTime();
  root->SetOptMutual(mutual);
  TapasFMM::Map(FMM_Upward(), *root, (real_t)theta);
  double t2 = GetTime();
  ResetCount();
  TapasFMM::Map(FMM_DTT(), tapas::Product(*root, *root), mutual, 1000, (real_t)theta);
  uint64_t c_m2l = 0, c_p2p = 0;
#ifdef COUNT
  c_m2l = numM2L; c_p2p = numP2P;
#endif
  double t3 = GetTime();
  TapasFMM::Map(FMM_Downward(), *root);
  double t4 = GetTime();

  auto &data = root->data();
  const int64_t nloc = (int64_t)root->local_nb();
  Body *b0 = &root->local_body(0);

  std::vector<CellRec> cells;
  cells.reserve(data.ht_.size());
  for (auto &kv : data.ht_) {
    const TapasFMM::Cell *c = kv.second;
    CellRec r;
    r.key = kv.first;
    r.leaf = c->IsLeaf() ? 1 : 0;
    r.nb = (int64_t)c->local_nb();
    r.bid = -1;
    if (r.leaf && r.nb > 0) {
      r.bid = (int64_t)(&const_cast<TapasFMM::Cell *>(c)->body(0) - b0);
    }
    r.cell = c;
    cells.push_back(r);
  }
  std::sort(cells.begin(), cells.end(),
            [](const CellRec &a, const CellRec &b) { return a.key < b.key; });

  FILE *fo = fopen(argv[2], "wb");
  if (!fo) { perror("output"); return 2; }
  put(fo, (int64_t)nloc);
  for (int d = 0; d < 3; d++) put(fo, (double)data.region_.min()[d]);
  for (int d = 0; d < 3; d++) put(fo, (double)data.region_.max()[d]);
  put(fo, (int32_t)data.max_depth_);
  // sorted keys and original indices of sorted bodies
  put_n(fo, data.local_body_keys_.data(), nloc);
  for (int64_t i = 0; i < nloc; i++) put(fo, (int64_t)b0[i].IBODY);
  // cells sorted by key
  put(fo, (int64_t)cells.size());
  for (auto &c : cells) {
    put(fo, c.key); put(fo, c.nb); put(fo, c.bid); put(fo, c.leaf);
    for (int d = 0; d < 3; d++) put(fo, (double)c.cell->center()[d]);
    for (int d = 0; d < 3; d++) put(fo, (double)c.cell->width(d));
  }
  // fields (pot, ax, ay, az) in original body order
  std::vector<double> trg(4 * nloc);
  for (int64_t i = 0; i < nloc; i++) {
    int64_t o = b0[i].IBODY;
    const auto &a = root->local_body_attr(i);
    for (int d = 0; d < 4; d++) trg[4 * o + d] = a[d];
  }
  put_n(fo, trg.data(), trg.size());
  put(fo, c_m2l); put(fo, c_p2p);
  put(fo, t1 - t0); put(fo, t2 - t1); put(fo, t3 - t2); put(fo, t4 - t3);
  put(fo, (int32_t)(dump_exp ? 1 : 0));
  if (dump_exp) {
    for (auto &c : cells) {
      const auto &a = c.cell->attr();
      for (int i = 0; i < NTERM; i++) { put(fo, a.M[i].real()); put(fo, a.M[i].imag()); }
      for (int i = 0; i < NTERM; i++) { put(fo, a.L[i].real()); put(fo, a.L[i].imag()); }
    }
  }
  fclose(fo);
  printf("ref: n=%lld cells=%zu depth=%d m2l=%llu p2p=%llu tree=%.3f up=%.3f dtt=%.3f down=%.3f\n",
         (long long)nloc, cells.size(), data.max_depth_, (unsigned long long)c_m2l,
         (unsigned long long)c_p2p, t1 - t0, t2 - t1, t3 - t2, t4 - t3);
  MPI_Finalize();
  return 0;
}
