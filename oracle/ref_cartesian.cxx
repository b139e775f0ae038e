// ref_cartesian.cxx -- runs the reference's own Cartesian-basis kernels
// (sample/exafmm-dev-13274dd4ac68/kernels/LaplaceCartesianCPU.cxx, compiled
// from where it lies with the BASIS = Cartesian flags of the reference's
// Makefile.include: -DCartesian -DEXPANSION=10 -DFP64) on fixed cells and
// bodies, and writes inputs and outputs to one binary file.
//
// TEST INFRASTRUCTURE ONLY: the source of the golden vectors of the oracle's
// Cartesian operators (tests/golden/make_golden.py).  Nothing from the
// reference is copied here; this file only calls kernel::P2M / M2M / M2L /
// L2L / L2P through the reference's kernel.h.
//
// Cells (indices):
//   0        parent of 1..8 (M2M of their multipoles)
//   1..8     leaves with bodies (P2M)
//   9, 10    target and source leaves (P2M; M2L 9 <- 10 with mutual = 1)
//   11       target at the centre of 9: M2L 11 <- 10 in a periodic image
//   12       parent of 13..20, its L = L of 9 after the mutual M2L
//   13..20   leaves with bodies: L2L from 12, then L2P (targets preset)
//
// Output (all little-endian doubles): ncells, nbodies, NTERM, then per cell
// X[3], M[NTERM], L[NTERM], IBODY, NBODY, ICHILD, NCHILD, IPARENT, then per
// body X[3], SRC, TRG before the kernels [4], TRG after [4], then the two
// Xperiodic vectors used (0 and the image's).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kernel.h"

namespace {

struct Lcg {
  uint64_t s;
  double uniform() {  // [0, 1)
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  double range(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

void put(FILE *f, double v) { fwrite(&v, sizeof(v), 1, f); }

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s out.bin [seed]\n", argv[0]);
    return 2;
  }
  Lcg rng{argc > 2 ? (uint64_t)std::atoll(argv[2]) : 12345u};
  const int ncells = 21;
  // centres and half widths
  vec3 cen[ncells];
  double half[ncells];
  const double ax[3] = {0.1, -0.2, 0.3}, tx[3] = {0.0, 0.0, 0.0}, sx[3] = {1.3, 0.4, -0.9},
               ux[3] = {-0.5, 0.6, 0.2};
  for (int d = 0; d < 3; d++) {
    cen[0][d] = ax[d];
    cen[9][d] = tx[d];
    cen[10][d] = sx[d];
    cen[11][d] = tx[d];
    cen[12][d] = ux[d];
  }
  half[0] = 0.5;
  half[9] = half[10] = half[11] = 0.25;
  half[12] = 0.25;
  for (int o = 0; o < 8; o++) {
    for (int d = 0; d < 3; d++) {
      const double s = ((o >> d) & 1) ? 1.0 : -1.0;
      cen[1 + o][d] = cen[0][d] + s * 0.25;
      cen[13 + o][d] = cen[12][d] + s * 0.125;
    }
    half[1 + o] = 0.25;
    half[13 + o] = 0.125;
  }
  // bodies of the leaves
  std::vector<int> nb(ncells, 0);
  for (int o = 0; o < 8; o++) {
    nb[1 + o] = 4 + o;
    nb[13 + o] = 3 + o;
  }
  nb[9] = 12;
  nb[10] = 9;
  std::vector<int> first(ncells, 0);
  int nbodies = 0;
  for (int c = 0; c < ncells; c++) {
    first[c] = nbodies;
    nbodies += nb[c];
  }
  Bodies bodies(nbodies);
  for (int c = 0; c < ncells; c++)
    for (int i = first[c]; i < first[c] + nb[c]; i++) {
      for (int d = 0; d < 3; d++) bodies[i].X[d] = cen[c][d] + rng.range(-half[c], half[c]);
      bodies[i].SRC = rng.range(0.2, 1.2);
      for (int d = 0; d < 4; d++) bodies[i].TRG[d] = c >= 13 ? rng.range(-1.0, 1.0) : 0.0;
    }
  std::vector<double> trg_in(4 * (size_t)nbodies);
  for (int i = 0; i < nbodies; i++)
    for (int d = 0; d < 4; d++) trg_in[4 * i + d] = bodies[i].TRG[d];
  Cells cells(ncells);
  for (int c = 0; c < ncells; c++) {
    Cell &C = cells[c];
    C.X = cen[c];
    C.R = half[c];
    C.M = 0;
    C.L = 0;
    C.IBODY = first[c];
    C.NBODY = nb[c];
    C.BODY = bodies.begin() + first[c];
    C.ICHILD = 0;
    C.NCHILD = 0;
    C.IPARENT = 0;
  }
  cells[0].ICHILD = 1;
  cells[0].NCHILD = 8;
  cells[12].ICHILD = 13;
  cells[12].NCHILD = 8;
  for (int o = 0; o < 8; o++) {
    cells[1 + o].IPARENT = 0;
    cells[13 + o].IPARENT = 12;
  }
  const C_iter C0 = cells.begin();
  for (int c = 0; c < ncells; c++)
    if (nb[c] > 0) kernel::P2M(C0 + c);
  kernel::M2M(C0, C0);
  vec3 none, image;
  none = 0.0;
  image[0] = 2.0;
  image[1] = -2.0;
  image[2] = 2.0;
  kernel::M2L(C0 + 9, C0 + 10, none, true);
  kernel::M2L(C0 + 11, C0 + 10, image, false);
  cells[12].L = cells[9].L;
  for (int o = 0; o < 8; o++) kernel::L2L(C0 + 13 + o, C0);
  for (int o = 0; o < 8; o++) kernel::L2P(C0 + 13 + o);

  FILE *f = fopen(argv[1], "wb");
  if (!f) return 1;
  put(f, ncells);
  put(f, nbodies);
  put(f, NTERM);
  for (int c = 0; c < ncells; c++) {
    const Cell &C = cells[c];
    for (int d = 0; d < 3; d++) put(f, C.X[d]);
    for (int t = 0; t < NTERM; t++) put(f, C.M[t]);
    for (int t = 0; t < NTERM; t++) put(f, C.L[t]);
    put(f, C.IBODY);
    put(f, C.NBODY);
    put(f, C.ICHILD);
    put(f, C.NCHILD);
    put(f, C.IPARENT);
  }
  for (int i = 0; i < nbodies; i++) {
    for (int d = 0; d < 3; d++) put(f, bodies[i].X[d]);
    put(f, bodies[i].SRC);
    for (int d = 0; d < 4; d++) put(f, trg_in[4 * i + d]);
    for (int d = 0; d < 4; d++) put(f, bodies[i].TRG[d]);
  }
  for (int d = 0; d < 3; d++) put(f, none[d]);
  for (int d = 0; d < 3; d++) put(f, image[d]);
  fclose(f);
  return 0;
}
