// tapas_exafmm_gpu.cxx -- the TapasFMM driver on the MI355X library, written
// against the same user surface as the reference application
// (sample/exafmm-dev-13274dd4ac68/examples/tapas_exafmm.cxx:448-712):
// Partition, Map(FMM_Upward), Map(FMM_DTT, Product(root, root)),
// Map(FMM_Downward), CopyBackResult, the main.csv timing report, the
// 100-target direct-sum check and root->Report()'s CSV files.
//
// Options (as the reference's EX/include/args.h): -n bodies, -c ncrit,
// -t theta, -m mutual (0 or 1), -d cube|plummer|sphere, -r repeat,
// -k check (0/1), -i images (periodic images of period 2 pi, which the
// reference's args take but its Tapas traversal leaves at 0; with -m 0),
// plus -p fp64|mixed and -b spherical|cartesian (the reference's
// compile-time BASIS, EX/Makefile.include:13).  Bodies follow the reference's
// dataset semantics (EX/include/dataset.h:131-271): cube U[-pi,pi)^3 or the
// Plummer sphere scaled by 3pi/16, charges drand48() - 0.5, srand48(0).
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "tapas_fmm.h"

using namespace tapas_gpu;

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static std::vector<Body> init_bodies(int n, const std::string &dist) {
  std::vector<Body> b(n);
  srand48(0);
  if (dist == "plummer") {
    for (int i = 0; i < n;) {
      const double X1 = drand48(), X2 = drand48(), X3 = drand48();
      const double R = 1.0 / std::sqrt(std::pow(X1, -2.0 / 3.0) - 1.0);
      if (R >= 100.0) continue;
      const double Z = (1.0 - 2.0 * X2) * R;
      const double s = std::sqrt(R * R - Z * Z);
      const double scale = 3.0 * M_PI / 16.0;
      b[i].X[0] = s * std::cos(2.0 * M_PI * X3) * scale;
      b[i].X[1] = s * std::sin(2.0 * M_PI * X3) * scale;
      b[i].X[2] = Z * scale;
      i++;
    }
  } else if (dist == "sphere") {
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) b[i].X[d] = drand48() * 2 - 1;
      const double r = std::sqrt(b[i].X[0] * b[i].X[0] + b[i].X[1] * b[i].X[1] +
                                 b[i].X[2] * b[i].X[2]);
      for (int d = 0; d < 3; d++) b[i].X[d] /= r * 1.1;
    }
  } else {
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) b[i].X[d] = drand48() * 2 * M_PI - M_PI;
  }
  srand48(0);
  for (int i = 0; i < n; i++) {
    b[i].SRC = drand48() - .5;
    b[i].IBODY = i;
    b[i].IRANK = 0;
    b[i].ICELL = 0;
    b[i].WEIGHT = 1;
    for (int d = 0; d < 4; d++) b[i].TRG[d] = 0;
  }
  return b;
}

int main(int argc, char **argv) {
  int n = 1000000, ncrit = 16, mutual = 1, repeat = 1, check = 1, images = 0;
  double theta = 0.4;
  std::string dist = "cube", prec = "fp64";
  int opt;
  std::string basis_name = "spherical";
  while ((opt = getopt(argc, argv, "n:c:t:m:d:r:k:p:i:b:")) != -1) {
    switch (opt) {
      case 'n': n = atoi(optarg); break;
      case 'c': ncrit = atoi(optarg); break;
      case 't': theta = atof(optarg); break;
      case 'm': mutual = atoi(optarg); break;
      case 'd': dist = optarg; break;
      case 'r': repeat = atoi(optarg); break;
      case 'k': check = atoi(optarg); break;
      case 'p': prec = optarg; break;
      case 'i': images = atoi(optarg); break;
      case 'b': basis_name = optarg; break;
      default: fprintf(stderr, "unknown option\n"); return 2;
    }
  }
  if (mutual != 0 && mutual != 1) {
    fprintf(stderr, "TapasFMM: [Error] mutual must be 0 or 1.\n");
    return 1;
  }
  if (images > 0 && mutual) {
    fprintf(stderr, "TapasFMM: [Error] periodic images are traversed with -m 0.\n");
    return 1;
  }
  const Precision precision = prec == "mixed" ? Precision::Mixed : Precision::FP64;
  if (basis_name != "spherical" && basis_name != "cartesian") {
    fprintf(stderr, "TapasFMM: [Error] basis must be spherical or cartesian.\n");
    return 1;
  }
  const Basis basis = basis_name == "cartesian" ? Basis::Cartesian : Basis::Spherical;
  const double cycle = 2 * M_PI;  // the reference's period (tapas_exafmm.cxx:537)
  std::vector<Body> bodies = init_bodies(n, dist);
  for (int t = 0; t < repeat; t++) {
    const double t0 = now();
    Cell *root = TapasFMM::Partition(bodies.data(), (index_t)bodies.size(), ncrit, 0, precision,
                                     images, cycle, basis);
    root->SetOptMutual(mutual);
    const double t1 = now();
    TapasFMM::Map(FMM_Upward(), *root, theta);
    const double t2 = now();
    TapasFMM::Map(FMM_DTT(), Product(*root, *root), mutual, 1000, theta);
    const double t3 = now();
    TapasFMM::Map(FMM_Downward(), *root);
    const double t4 = now();
    // CopyBackResult
    std::vector<Body> out(root->local_nb());
    for (index_t i = 0; i < root->local_nb(); i++) {
      out[i] = root->local_body(i);
      for (int d = 0; d < 4; d++) out[i].TRG[d] = root->local_body_attr(i)[d];
    }
    const tfmm_stats_t st = TapasFMM::Stats(*root);
    // the reference's main.csv (tapas_exafmm.cxx:651-677, built with COUNT)
    tapas_gpu::util::RankCSV csv({"total", "upward", "traverse", "downward", "tree", "numP2P",
                                  "numM2L"});
    csv.At("total", t4 - t0);
    csv.At("upward", t2 - t1);
    csv.At("traverse", t3 - t2);
    csv.At("downward", t4 - t3);
    csv.At("tree", t1 - t0);
    csv.At("numP2P", (uint64_t)st.n_p2p_calls);
    csv.At("numM2L", (uint64_t)st.n_m2l);
    csv.Dump(tapas_gpu::util::ReportPath("main"));
    printf("bodies %d cells %lld depth %d  tree %.4f upward %.4f traverse %.4f downward %.4f"
           " total %.4f s (device: tree %.2f up %.2f trav %.2f down %.2f ms)\n",
           n, (long long)st.ncells, st.max_depth, t1 - t0, t2 - t1, t3 - t2, t4 - t3, t4 - t0,
           st.ms_tree, st.ms_upward, st.ms_traverse, st.ms_downward);
    if (check) {
      // 100 stride-sampled targets (sorted order) against the direct sum (with
      // periodic images over the same copies, CheckResult's prange loops)
      const int ns = std::min<int>(100, (int)out.size());
      const int stride = (int)out.size() / ns;
      std::vector<int64_t> tg(ns);
      for (int i = 0; i < ns; i++) tg[i] = out[(size_t)i * stride].IBODY;
      std::vector<double> ref(4 * (size_t)ns);
      if (tfmm_direct(root->tree()->ctx, tg.data(), ns, ref.data()) != TFMM_OK) return 1;
      double pd = 0, pn = 0, ad = 0, an = 0;
      for (int i = 0; i < ns; i++) {
        const Body &b = out[(size_t)i * stride];
        pd += (b.TRG[0] - ref[4 * i]) * (b.TRG[0] - ref[4 * i]);
        pn += ref[4 * i] * ref[4 * i];
        for (int d = 1; d < 4; d++) {
          ad += (b.TRG[d] - ref[4 * i + d]) * (b.TRG[d] - ref[4 * i + d]);
          an += ref[4 * i + d] * ref[4 * i + d];
        }
      }
      printf("Rel. L2 Error (pot)            : %e\n", std::sqrt(pd / pn));
      printf("Rel. L2 Error (acc)            : %e\n", std::sqrt(ad / an));
    }
    root->Report();  // the reference's report files (tapas_exafmm.cxx:700)
    delete root;
  }
  return 0;
}
