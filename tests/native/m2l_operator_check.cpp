// CPU check of the batched M2L formulation used by the MFMA tile kernel:
// operator codes (m2l_codes.h), per-group harmonic scaling with the level
// factor folded into the harmonics (2^(p dl)) and into the rows of the A
// operands (2^(-j dl)), and the M~/L~ scalings, applied in FP64 exactly as the
// kernel applies them (A fragment by A fragment: block, k-step, lane), against
// the oracle's per-pair M2L
// (orc_m2l, the restatement of LaplaceSphericalCPU_tapas.h:208-273).
// Prints "max_rel <value>" and exits 0 when the codes build.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../tapas-reduce_amd/csrc/m2l_codes.h"

extern "C" {
void orc_eval_local(double rho, double alpha, double beta, double *ynm);
void orc_m2l(double *Li, const double *Mj, const double *dX);
}

using namespace tfmm::m2lc;

int main(int argc, char **argv) {
  const int trials = argc > 1 ? atoi(argv[1]) : 60;
  std::vector<uint32_t> c8, c4;
  if (!build_codes(c8, 8) || !build_codes(c4, 4)) {
    printf("codes inconsistent\n");
    return 1;
  }
  for (size_t i = 0; i < c8.size(); i++) {
    if ((c8[i] & 0xffffu) != 2 * (c4[i] & 0xffffu) || (c8[i] >> 16) != 2 * (c4[i] >> 16)) {
      printf("fp32/fp64 codes disagree at %zu\n", i);
      return 1;
    }
  }
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(-1, 1);
  double worst = 0;
  for (int t = 0; t < trials; t++) {
    const int dl = (int)(t % 7) - 3;             // l_t - l_s in -3 .. 3
    const double sig_t = std::ldexp(1.0, -(t % 7)) * (1.0 + 0.5 * (u(rng) + 1));
    const double sig_s = std::ldexp(sig_t, dl);  // sigma_l = w0 2^-l
    double dX[3];
    double r = 0;
    do {
      for (int d = 0; d < 3; d++) dX[d] = 5 * sig_t * u(rng);
      r = std::sqrt(dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
    } while (r < 2 * sig_t);
    if (t % 11 == 0) dX[0] = dX[1] = 0;  // on the z axis
    r = std::sqrt(dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
    std::vector<double> Y(2 * P * P);
    orc_eval_local(r, std::acos(dX[2] / r), std::atan2(dX[1], dX[0]), Y.data());
    // group harmonics: Y'_p^q = Y_p^q sigma_t^(p+1) 2^(p dl), q >= 0
    double yv[NR] = {0};
    for (int p = 0; p < P; p++)
      for (int q = 0; q <= p; q++)
        for (int c = 0; c < 2; c++)
          yv[2 * tri(p, q) + c] =
              Y[2 * (p * p + p + q) + c] * std::pow(sig_t, p + 1) * std::ldexp(1.0, p * dl);
    std::vector<double> tbl(NMULT * NR);
    for (int i = 0; i < NR; i++) {
      tbl[i] = yv[i];
      tbl[NR + i] = -yv[i];
      tbl[2 * NR + i] = 2 * yv[i];
      tbl[3 * NR + i] = -2 * yv[i];
    }
    // source multipole, M~_n = M_n / sigma_s^n
    double M[110], Mt[NR] = {0};
    for (int i = 0; i < 110; i++) M[i] = u(rng);
    for (int n = 0; n < P; n++) M[2 * tri(n, 0) + 1] = 0;  // m = 0 terms are real
    for (int i = 0; i < 110; i++) Mt[i] = M[i] / std::pow(sig_s, degree_of(i));
    // the kernel's A fragments (row factor 2^(-j dl) included), its
    // block-sparse product, then unscaling
    double Lt[NR] = {0};
    for (int b = 0; b < NBLK; b++)
      for (int kk = 0; kk < 4; kk++)
        for (int l = 0; l < 64; l++) {
          const uint32_t cw = c8[(size_t)(b * 4 + kk) * 64 + l];
          const int row = 16 * BLK_RB[b] + l % 16, col = 16 * BLK_CB[b] + 4 * kk + l / 16;
          const double a = std::ldexp(tbl[(cw & 0xffffu) / 8] + tbl[(cw >> 16) / 8],
                                      -degree_of(row) * dl);
          Lt[row] += a * Mt[col];
        }
    double L[110];
    for (int i = 0; i < 110; i++) L[i] = Lt[i] / std::pow(sig_t, degree_of(i) + 1);
    double Lref[110] = {0};
    orc_m2l(Lref, M, dX);
    double num = 0, den = 0;
    for (int i = 0; i < 110; i++) {
      num += (L[i] - Lref[i]) * (L[i] - Lref[i]);
      den += Lref[i] * Lref[i];
    }
    const double rel = std::sqrt(num / den);
    if (!(rel <= worst)) worst = rel;
  }
  printf("max_rel %.3e\n", worst);
  return 0;
}
