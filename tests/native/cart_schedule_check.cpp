// CPU check of the Cartesian M2L as the kernel (cartesian.hip, k_cart_m2l)
// computes it: the schedule tables of cart_sched.h, the derivative
// recurrence driven by the tables' index and factorial entries, the lanes'
// bin accumulators over several sources, the zero slot and zero padding,
// and the fixed-order sum of each term's pieces -- against the oracle's
// per-pair M2L (orc_cart_m2l, the restatement of
// sample/exafmm-dev-13274dd4ac68/kernels/LaplaceCartesianCPU.cxx:659-673).
// Also checks that every product (n, k), |n| + |k| <= ORDER - 1, is
// scheduled exactly once.  Prints "max_rel <value>"; exit 1 on a broken
// table.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../tapas-reduce_amd/csrc/cart_sched.h"

extern "C" {
void orc_cart_m2l(double *Li, const double *Mj, const double *dX);
}

using namespace tfmm::cart;

static CartTables tb;

// the kernel's wave_derivs on one pair
static void derivs(double *C, const double dX[3]) {
  const double r2 = dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2];
  const double invR2 = 1.0 / r2;
  C[0] = std::sqrt(invR2);
  for (int d = 1; d < ORDER; d++) {
    const int nt = (d + 1) * (d + 2) / 2;
    for (int lane = 0; lane < nt; lane++) {
      const int t = cdeg_off(d) + lane;
      const uint64_t r = tb.rec[t];
      const double s1 = dX[0] * C[r & 0xff] + dX[1] * C[(r >> 8) & 0xff] +
                        dX[2] * C[(r >> 16) & 0xff];
      const double s2 = C[(r >> 24) & 0xff] + C[(r >> 32) & 0xff] + C[(r >> 40) & 0xff];
      C[t] = ((double)(1 - 2 * d) / d * s1 + (double)(1 - d) / d * s2) * invR2;
    }
  }
  for (int t = 0; t < CN; t++) C[t] *= tb.fact[t];
}

int main(int argc, char **argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 40;
  if (!build_tables(tb)) {
    std::printf("tables inconsistent\n");
    return 1;
  }
  // coverage: every (n, k) once
  std::vector<int> seen(CN * CN, 0);
  for (int lane = 0; lane < 64; lane++) {
    int j = 0;
    for (int b = 0; b < NBINS; b++)
      for (int i = 0; i < BL[b]; i++, j++) {
        const uint32_t w = tb.ent[j >> 1][lane];
        const int off = (int)((j & 1) ? (w >> 16) : (w & 0xffffu));
        if (off % 8 || off / 8 > ZS) {
          std::printf("bad offset\n");
          return 1;
        }
        const int c = off / 8, k = (int)tb.kb[b][lane] / 8 + i;
        if (c == ZS) continue;
        if (k >= CN) {
          std::printf("product past the terms\n");
          return 1;
        }
        // n = (n + k) - k
        const int n = cidx(term_part(c, 0) - term_part(k, 0), term_part(c, 1) - term_part(k, 1),
                           term_part(c, 2) - term_part(k, 2));
        seen[n * CN + k]++;
      }
  }
  for (int n = 0; n < CN; n++)
    for (int k = 0; k < CN; k++) {
      const int dn = term_part(n, 0) + term_part(n, 1) + term_part(n, 2);
      const int dk = term_part(k, 0) + term_part(k, 1) + term_part(k, 2);
      if (seen[n * CN + k] != (dn + dk <= ORDER - 1 ? 1 : 0)) {
        std::printf("product (%d, %d) scheduled %d times\n", n, k, seen[n * CN + k]);
        return 1;
      }
    }
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  double worst = 0;
  for (int trial = 0; trial < trials; trial++) {
    const int nsrc = 1 + trial % 4;
    std::vector<double> acc(NSLOT, 0.0), Lref(CN, 0.0);
    for (int s = 0; s < nsrc; s++) {
      double Ms[MS_LEN] = {0}, C[CN + 1];
      for (int t = 0; t < CN; t++) Ms[t] = u(rng) / (1 + t % 7);
      double dX[3];
      const double dist = 1.5 + 3.0 * (0.5 + 0.5 * u(rng));
      double nn = 0;
      for (int d = 0; d < 3; d++) {
        dX[d] = u(rng);
        nn += dX[d] * dX[d];
      }
      for (int d = 0; d < 3; d++) dX[d] *= dist / std::sqrt(nn);
      C[ZS] = 0;
      derivs(C, dX);
      for (int lane = 0; lane < 64; lane++) {
        int j = 0;
        for (int b = 0; b < NBINS; b++) {
          const double *m = Ms + tb.kb[b][lane] / 8;
          double a = acc[b * 64 + lane];
          for (int i = 0; i < BL[b]; i++, j++) {
            const uint32_t w = tb.ent[j >> 1][lane];
            const uint32_t off = (j & 1) ? (w >> 16) : (w & 0xffffu);
            a = std::fma(m[i], C[off / 8], a);
          }
          acc[b * 64 + lane] = a;
        }
      }
      orc_cart_m2l(Lref.data(), Ms, dX);
    }
    double num = 0, den = 0;
    for (int t = 0; t < CN; t++) {
      const uint32_t pc = tb.piece[t];
      double s = 0;
      for (uint32_t p = 0; p < (pc >> 16); p++) s += acc[tb.slot[(pc & 0xffffu) + p]];
      num = std::fmax(num, std::fabs(s - Lref[t]));
      den = std::fmax(den, std::fabs(Lref[t]));
    }
    worst = std::fmax(worst, num / den);
  }
  std::printf("max_rel %.3e\n", worst);
  return 0;
}
