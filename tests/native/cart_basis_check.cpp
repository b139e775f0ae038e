// CPU check of the basis change between the Cartesian and the spherical
// expansions (cart_sched.h, build_basis_tables), on the oracle's kernels:
//   t1:   T1 (Cartesian P2M) = spherical P2M of the same bodies
//         (orc_cart_p2m, orc_p2m);
//   t2:   the fields of a spherical local expansion (orc_l2p) = the fields of
//         its T2 image in the Cartesian basis (orc_cart_l2p);
//   m2l:  T2 (spherical M2L (T1 M)) = Cartesian M2L (M) for well separated
//         pairs (orc_m2l, orc_cart_m2l: LaplaceSphericalCPU_tapas.h:208-273,
//         LaplaceCartesianCPU.cxx:659-673).
// Prints "t1 <max rel> t2 <max rel> m2l <max rel>"; exit 1 on a broken table.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../tapas-reduce_amd/csrc/cart_sched.h"

extern "C" {
void orc_p2m(double *M, const double *cen, const double *x, const double *y, const double *z,
             const double *q, int64_t nb);
void orc_m2l(double *Li, const double *Mj, const double *dX);
void orc_l2p(double *trg, const double *bx, double q, const double *L, const double *cen);
void orc_cart_p2m(double *M, const double *cen, const double *x, const double *y,
                  const double *z, const double *q, int64_t nb);
void orc_cart_m2l(double *Li, const double *Mj, const double *dX);
void orc_cart_l2p(double *trg, const double *bx, double q, const double *L, const double *cen);
}

using namespace tfmm::cart;
typedef std::complex<double> cd;

static BasisTables bt;

static void to_sph(const double *Mc, cd *Ms) {
  for (int t = 0; t < SPH_TERMS; t++) {
    cd s = 0;
    for (uint32_t e = bt.t1_off[t]; e < bt.t1_off[t + 1]; e++)
      s += cd(bt.t1_re[e], bt.t1_im[e]) * Mc[bt.t1_k[e]];
    Ms[t] = s;
  }
}

static void to_cart(const cd *Ls, double *Lc) {
  for (int k = 0; k < CN; k++) {
    double s = 0;
    for (uint32_t e = bt.t2_off[k]; e < bt.t2_off[k + 1]; e++)
      s += bt.t2_a[e] * Ls[bt.t2_t[e]].real() + bt.t2_b[e] * Ls[bt.t2_t[e]].imag();
    Lc[k] = s;
  }
}

static double rel(const double *a, const double *b, int n) {
  double num = 0, den = 0;
  for (int i = 0; i < n; i++) {
    num = std::fmax(num, std::fabs(a[i] - b[i]));
    den = std::fmax(den, std::fabs(b[i]));
  }
  return num / den;
}

int main(int argc, char **argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 20;
  if (!build_basis_tables(bt)) {
    std::printf("basis tables inconsistent\n");
    return 1;
  }
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  double e1 = 0, e2 = 0, e3 = 0;
  for (int trial = 0; trial < trials; trial++) {
    const int nb = 25;
    std::vector<double> x(nb), y(nb), z(nb), q(nb);
    const double cen[3] = {0.1 * u(rng), 0.1 * u(rng), 0.1 * u(rng)};
    for (int i = 0; i < nb; i++) {
      x[i] = cen[0] + 0.3 * u(rng);
      y[i] = cen[1] + 0.3 * u(rng);
      z[i] = cen[2] + 0.3 * u(rng);
      q[i] = u(rng);
    }
    // t1
    std::vector<double> Mc(CN, 0.0);
    std::vector<cd> Ms(SPH_TERMS, 0.0), Mt(SPH_TERMS);
    orc_cart_p2m(Mc.data(), cen, x.data(), y.data(), z.data(), q.data(), nb);
    orc_p2m(reinterpret_cast<double *>(Ms.data()), cen, x.data(), y.data(), z.data(), q.data(),
            nb);
    to_sph(Mc.data(), Mt.data());
    e1 = std::fmax(e1, rel(reinterpret_cast<double *>(Mt.data()),
                           reinterpret_cast<double *>(Ms.data()), 2 * SPH_TERMS));
    // t2: a random spherical local expansion (Im of m = 0 has no effect)
    std::vector<cd> Ls(SPH_TERMS);
    for (int n = 0; n < ORDER; n++)
      for (int m = 0; m <= n; m++) Ls[n * (n + 1) / 2 + m] = cd(u(rng), m ? u(rng) : 0.0);
    std::vector<double> Lc(CN);
    to_cart(Ls.data(), Lc.data());
    for (int p = 0; p < 5; p++) {
      const double b[3] = {cen[0] + 0.4 * u(rng), cen[1] + 0.4 * u(rng), cen[2] + 0.4 * u(rng)};
      double fs[4] = {0, 0, 0, 0}, fc[4] = {0, 0, 0, 0};
      orc_l2p(fs, b, 1.0, reinterpret_cast<const double *>(Ls.data()), cen);
      orc_cart_l2p(fc, b, 1.0, Lc.data(), cen);
      e2 = std::fmax(e2, rel(fc, fs, 4));
    }
    // m2l through the spherical basis
    double dX[3], nn = 0;
    for (int d = 0; d < 3; d++) {
      dX[d] = u(rng);
      nn += dX[d] * dX[d];
    }
    const double dist = 1.5 + 2.0 * (0.5 + 0.5 * u(rng));
    for (int d = 0; d < 3; d++) dX[d] *= dist / std::sqrt(nn);
    std::vector<cd> Lsph(SPH_TERMS, 0.0);
    orc_m2l(reinterpret_cast<double *>(Lsph.data()), reinterpret_cast<const double *>(Mt.data()),
            dX);
    std::vector<double> L1(CN), L2(CN, 0.0);
    to_cart(Lsph.data(), L1.data());
    orc_cart_m2l(L2.data(), Mc.data(), dX);
    e3 = std::fmax(e3, rel(L1.data(), L2.data(), CN));
  }
  std::printf("t1 %.3e t2 %.3e m2l %.3e\n", e1, e2, e3);
  return 0;
}
