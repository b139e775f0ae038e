/*
 * fmm_oracle.c -- CPU restatement of the Tapas FMM path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls it.
 *
 * It restates, in plain C, the algorithm of the reference path
 * (paths relative to the reference repository):
 *   TP = cpp/include/tapas/, EX = sample/exafmm-dev-13274dd4ac68/
 *
 *   bounding box (scalar cube)     TP/hot/buildtree.h:72-87, TP/vec.h:39-43
 *   body keys                      TP/hot/buildtree.h:504-539,
 *                                  TP/sfc_morton.h:387-402, :208-211
 *   key sort                       TP/common.h:248-273 (ties: any order)
 *   cell generation                TP/hot/buildtree.h:388-459
 *   cell region / centre           TP/sfc_morton.h:536-557, TP/hot.h:232-238
 *   upward pass (P2M, M2M)         EX/examples/tapas_exafmm.cxx:87-122,
 *                                  EX/examples/LaplaceSphericalCPU_tapas.h:138-206
 *   dual tree traversal            EX/examples/tapas_exafmm.cxx:209-286,
 *                                  TP/hot/mapper.h:29-169, TP/geometry.h:80-83
 *   M2L                            EX/examples/LaplaceSphericalCPU_tapas.h:208-273
 *   P2P                            EX/examples/LaplaceP2PCPU_tapas.h:48-74
 *   downward pass (L2L, L2P)       EX/examples/tapas_exafmm.cxx:124-140,
 *                                  EX/examples/LaplaceSphericalCPU_tapas.h:275-345
 *   solid harmonics                EX/examples/LaplaceSphericalCPU_tapas.h:43-135
 *   Cartesian basis (P2M, M2M, M2L, EX/kernels/LaplaceCartesianCPU.cxx:3-704
 *     L2L, L2P)                    (the reference's BASIS = Cartesian build)
 *   periodic images                EX/include/traversal.h:126-185, :247-268
 *   direct sum + error metric      EX/examples/tapas_exafmm.cxx:288-383,
 *                                  EX/include/verify.h:18-92
 *
 * Arithmetic is written in the reference's operation order and must be
 * compiled with -ffp-contract=off (no FMA), so that the key, region and
 * multipole-acceptance decisions are bit-identical to the reference's
 * x86-64 build.  Loop orders (tree build, traversal, accumulation) follow
 * the reference too, so field sums round the same way.
 *
 * Parity pinning of each function is listed in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_oracle.h"

#define P ORC_P
#define NTERM ORC_NTERM
#define MAXDEPTH 19
#define DEPTHBITS 5

typedef struct { double re, im; } cx;

static inline cx cx_make(double re, double im) { cx r; r.re = re; r.im = im; return r; }
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
/* complex * complex: (ac - bd, ad + bc), no fused terms */
static inline cx cx_mul(cx a, cx b) {
  double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
  return cx_make(ac - bd, ad + bc);
}
static inline cx cx_scale(cx a, double s) { return cx_make(a.re * s, a.im * s); }
static inline cx cx_rscale(double s, cx a) { return cx_make(s * a.re, s * a.im); }
static inline cx cx_conj(cx a) { return cx_make(a.re, -a.im); }

static inline int oddeven(int n) { return (n & 1) ? -1 : 1; }
static inline int ipow2n(int n) { return (n >= 0) ? 1 : oddeven(n); }

/* ------------------------------------------------------------------ */
/* Geometry and Morton keys                                             */
/* ------------------------------------------------------------------ */

void orc_bbox(const double *x, const double *y, const double *z, int64_t n,
              double *out_min, double *out_max) {
  /* The reference assigns a scalar to a Vec inside a per-dimension loop;
   * the broadcast constructor makes every component equal.  Body 0 thus
   * contributes only its z, every later body all three coordinates. */
  double mx = 0, mn = 0;
  for (int64_t i = 0; i < n; i++) {
    const double pos[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; d++) {
      if (i == 0) {
        mx = pos[d];
        mn = pos[d];
      } else {
        /* std::max(a, b) returns a if !(a < b); std::min(a, b) returns a if !(b < a) */
        mx = (pos[d] < mx) ? mx : pos[d];
        mn = (mn < pos[d]) ? mn : pos[d];
      }
    }
  }
  for (int d = 0; d < 3; d++) { out_min[d] = mn; out_max[d] = mx; }
}

uint64_t orc_finest_key(const int anchor[3]) {
  uint64_t k = 0;
  int mask = 1 << (MAXDEPTH - 1);
  for (int i = 0; i < MAXDEPTH; i++) {
    for (int d = 2; d >= 0; d--) {
      k = (k << 1) | (uint64_t)((anchor[d] & mask) >> (MAXDEPTH - i - 1));
    }
    mask >>= 1;
  }
  return (k << DEPTHBITS) | MAXDEPTH;
}

void orc_keys(const double *x, const double *y, const double *z, int64_t n,
              const double *rmin, const double *rmax, uint64_t *keys) {
  const int nfin = 1 << MAXDEPTH;
  double pitch[3];
  for (int d = 0; d < 3; d++) pitch[d] = (rmax[d] - rmin[d]) / (double)nfin;
  for (int64_t i = 0; i < n; i++) {
    double ofst[3] = {x[i], y[i], z[i]};
    int anchor[3];
    for (int d = 0; d < 3; d++) {
      ofst[d] -= rmin[d];
      ofst[d] /= pitch[d];
    }
    for (int d = 0; d < 3; d++) {
      anchor[d] = (int)ofst[d];
      if (anchor[d] == nfin) anchor[d]--;
    }
    keys[i] = orc_finest_key(anchor);
  }
}

static inline int key_depth(uint64_t k) { return (int)(k & 31u); }
static inline uint64_t key_next(uint64_t k) {
  int d = key_depth(k);
  return k + ((uint64_t)1 << (3 * (MAXDEPTH - d) + DEPTHBITS));
}
static inline uint64_t key_child(uint64_t k, int idx) {
  int d = key_depth(k) + 1;
  uint64_t kk = ((k >> DEPTHBITS) << DEPTHBITS) | (uint64_t)d;
  return kk | ((uint64_t)idx << ((MAXDEPTH - d) * 3 + DEPTHBITS));
}
static inline int key_dir(uint64_t k, int dim, int depth) {
  uint64_t dir = (k >> DEPTHBITS) >> ((MAXDEPTH - depth) * 3);
  return (int)((dir >> dim) & 1);
}

void orc_cell_region(uint64_t key, const double *rmin, const double *rmax,
                     double *cmin, double *cmax) {
  double mx[3], mn[3];
  for (int d = 0; d < 3; d++) { mx[d] = rmax[d]; mn[d] = rmin[d]; }
  int depth = key_depth(key);
  for (int dep = 1; dep <= depth; dep++) {
    for (int dim = 0; dim < 3; dim++) {
      double c = (mx[dim] + mn[dim]) / 2;
      if (key_dir(key, dim, dep) == 1) mn[dim] = c;
      else mx[dim] = c;
    }
  }
  for (int d = 0; d < 3; d++) { cmin[d] = mn[d]; cmax[d] = mx[d]; }
}

/* ------------------------------------------------------------------ */
/* Solid harmonics                                                      */
/* ------------------------------------------------------------------ */

static void cart2sph(const double dX[3], double *r, double *theta, double *phi) {
  double n2 = 0;
  for (int i = 0; i < 3; i++) n2 += dX[i] * dX[i];
  *r = sqrt(n2);
  *theta = (*r == 0) ? 0 : acos(dX[2] / *r);
  *phi = atan2(dX[1], dX[0]);
}

/* r^n Y_n^m and its theta derivative; Ynm has P*P entries */
static void eval_multipole(double rho, double alpha, double beta, cx *Ynm, cx *YnmTheta) {
  double x = cos(alpha), y = sin(alpha);
  double fact = 1, pn = 1, rhom = 1;
  cx ei = cx_make(cos(beta), sin(beta));
  cx eim = cx_make(1.0, 0.0);
  for (int m = 0; m < P; m++) {
    double p = pn;
    int npn = m * m + 2 * m, nmn = m * m;
    Ynm[npn] = cx_rscale(rhom * p, eim);
    Ynm[nmn] = cx_conj(Ynm[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    YnmTheta[npn] = cx_rscale(rhom * (p - (m + 1) * x * p1) / y, eim);
    rhom *= rho;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      rhon /= -(n + m);
      Ynm[npm] = cx_rscale(rhon * p, eim);
      Ynm[nmm] = cx_conj(Ynm[npm]);
      double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      YnmTheta[npm] = cx_rscale(rhon * ((n - m + 1) * p - (n + 1) * x * p1) / y, eim);
      rhon *= rho;
    }
    rhom /= -(2 * m + 2) * (2 * m + 1);
    pn = -pn * fact * y;
    fact += 2;
    eim = cx_mul(eim, ei);
  }
}

/* r^{-n-1} Y_n^m */
static void eval_local(double rho, double alpha, double beta, cx *Ynm) {
  double x = cos(alpha), y = sin(alpha);
  double fact = 1, pn = 1;
  double invR = -1.0 / rho;
  double rhom = -invR;
  cx ei = cx_make(cos(beta), sin(beta));
  cx eim = cx_make(1.0, 0.0);
  for (int m = 0; m < P; m++) {
    double p = pn;
    int npn = m * m + 2 * m, nmn = m * m;
    Ynm[npn] = cx_rscale(rhom * p, eim);
    Ynm[nmn] = cx_conj(Ynm[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    rhom *= invR;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      Ynm[npm] = cx_rscale(rhon * p, eim);
      Ynm[nmm] = cx_conj(Ynm[npm]);
      double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhon *= invR * (n - m + 1);
    }
    pn = -pn * fact * y;
    fact += 2;
    eim = cx_mul(eim, ei);
  }
}

void orc_eval_multipole(double rho, double alpha, double beta, double *ynm, double *ynm_theta) {
  cx Y[P * P], T[P * P];
  memset(T, 0, sizeof(T));
  eval_multipole(rho, alpha, beta, Y, T);
  memcpy(ynm, Y, sizeof(Y));
  if (ynm_theta) memcpy(ynm_theta, T, sizeof(T));
}

void orc_eval_local(double rho, double alpha, double beta, double *ynm) {
  cx Y[P * P];
  eval_local(rho, alpha, beta, Y);
  memcpy(ynm, Y, sizeof(Y));
}

/* ------------------------------------------------------------------ */
/* Expansion kernels on raw coefficient arrays                          */
/* ------------------------------------------------------------------ */

/* M[nms] += q * Y[n^2+n-m] over the bodies (x,y,z,q) of a leaf */
static void p2m(cx *M, const double *cen, const double *x, const double *y, const double *z,
                const double *q, int64_t nb) {
  cx Ynm[P * P], YnmTheta[P * P];
  for (int64_t i = 0; i < nb; i++) {
    double dX[3] = {x[i] - cen[0], y[i] - cen[1], z[i] - cen[2]};
    double rho, alpha, beta;
    cart2sph(dX, &rho, &alpha, &beta);
    eval_multipole(rho, alpha, beta, Ynm, YnmTheta);
    for (int n = 0; n < P; n++) {
      for (int m = 0; m <= n; m++) {
        int nm = n * n + n - m;
        int nms = n * (n + 1) / 2 + m;
        M[nms] = cx_add(M[nms], cx_rscale(q[i], Ynm[nm]));
      }
    }
  }
}

/* parent M += shifted child M; dX = parent centre - child centre */
static void m2m(cx *Mp, const cx *Mc, const double dX[3]) {
  cx Ynm[P * P], YnmTheta[P * P];
  double rho, alpha, beta;
  cart2sph(dX, &rho, &alpha, &beta);
  eval_multipole(rho, alpha, beta, Ynm, YnmTheta);
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      int jks = j * (j + 1) / 2 + k;
      cx M = cx_make(0, 0);
      for (int n = 0; n <= j; n++) {
        int mlo = (-n > -j + k + n) ? -n : (-j + k + n);
        int mhi = (k - 1 < n) ? (k - 1) : n;
        for (int m = mlo; m <= mhi; m++) {
          int jnkms = (j - n) * (j - n + 1) / 2 + k - m;
          int nm = n * n + n - m;
          M = cx_add(M, cx_scale(cx_mul(Mc[jnkms], Ynm[nm]), (double)(ipow2n(m) * oddeven(n))));
        }
        int mhi2 = (n < j + k - n) ? n : (j + k - n);
        for (int m = k; m <= mhi2; m++) {
          int jnkms = (j - n) * (j - n + 1) / 2 - k + m;
          int nm = n * n + n - m;
          M = cx_add(M, cx_scale(cx_mul(cx_conj(Mc[jnkms]), Ynm[nm]), (double)oddeven(k + n + m)));
        }
      }
      Mp[jks] = cx_add(Mp[jks], M);
    }
  }
}

/* target L += M2L(source M) with the singular harmonics Ynmi of the offset */
static void m2l_y(cx *Li, const cx *Mj, const cx *Ynmi) {
  for (int j = 0; j < P; j++) {
    double Cnm = oddeven(j);
    for (int k = 0; k <= j; k++) {
      int jks = j * (j + 1) / 2 + k;
      cx L = cx_make(0, 0);
      for (int n = 0; n < P - j; n++) {
        for (int m = -n; m < 0; m++) {
          int nms = n * (n + 1) / 2 - m;
          int jnkm = (j + n) * (j + n) + j + n + m - k;
          L = cx_add(L, cx_mul(cx_scale(cx_conj(Mj[nms]), Cnm), Ynmi[jnkm]));
        }
        for (int m = 0; m <= n; m++) {
          int nms = n * (n + 1) / 2 + m;
          int jnkm = (j + n) * (j + n) + j + n + m - k;
          double Cnm2 = Cnm * oddeven((k - m) * (k < m) + m);
          L = cx_add(L, cx_mul(cx_scale(Mj[nms], Cnm2), Ynmi[jnkm]));
        }
      }
      Li[jks] = cx_add(Li[jks], L);
    }
  }
}

/* target L += M2L(source M); dX = target centre - source centre */
static void m2l(cx *Li, const cx *Mj, const double dX[3]) {
  cx Ynmi[P * P];
  double rho, alpha, beta;
  cart2sph(dX, &rho, &alpha, &beta);
  eval_local(rho, alpha, beta, Ynmi);
  m2l_y(Li, Mj, Ynmi);
}

/* Mutual M2L (LaplaceSphericalCPU_tapas.h:208-273 with mutual = 1): L_i +=
 * M2L(M_j) as above, and L_j += M2L(M_i) with the harmonics evaluated at
 * (rho, alpha + pi, beta), the reference's form of the reversed offset (:225) */
static void m2l_mutual(cx *Li, cx *Lj, const cx *Mi, const cx *Mj, const double dX[3]) {
  cx Ynmi[P * P], Ynmj[P * P];
  double rho, alpha, beta;
  cart2sph(dX, &rho, &alpha, &beta);
  eval_local(rho, alpha, beta, Ynmi);
  eval_local(rho, alpha + 3.14159265358979323846 /* M_PI */, beta, Ynmj);
  m2l_y(Li, Mj, Ynmi);
  m2l_y(Lj, Mi, Ynmj);
}

/* child L += shifted parent L; dX = child centre - parent centre */
static void l2l(cx *Lc, const cx *Lp, const double dX[3]) {
  cx Ynm[P * P], YnmTheta[P * P];
  double rho, alpha, beta;
  cart2sph(dX, &rho, &alpha, &beta);
  eval_multipole(rho, alpha, beta, Ynm, YnmTheta);
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      int jks = j * (j + 1) / 2 + k;
      cx L = cx_make(0, 0);
      for (int n = j; n < P; n++) {
        for (int m = j + k - n; m < 0; m++) {
          int jnkm = (n - j) * (n - j) + n - j + m - k;
          int nms = n * (n + 1) / 2 - m;
          L = cx_add(L, cx_scale(cx_mul(cx_conj(Lp[nms]), Ynm[jnkm]), (double)oddeven(k)));
        }
        for (int m = 0; m <= n; m++) {
          if (n - j >= abs(m - k)) {
            int jnkm = (n - j) * (n - j) + n - j + m - k;
            int nms = n * (n + 1) / 2 + m;
            L = cx_add(L, cx_scale(cx_mul(Lp[nms], Ynm[jnkm]), (double)oddeven((m - k) * (m < k))));
          }
        }
      }
      Lc[jks] = cx_add(Lc[jks], L);
    }
  }
}

/* trg (pot, ax, ay, az) of one body: divide P2P sum by q, add local field */
static void l2p(double *trg, const double bx[3], double q, const cx *L, const double cen[3]) {
  cx Ynm[P * P], YnmTheta[P * P];
  double dX[3] = {bx[0] - cen[0], bx[1] - cen[1], bx[2] - cen[2]};
  double sph[3] = {0, 0, 0}, cart[3];
  double r, theta, phi;
  cart2sph(dX, &r, &theta, &phi);
  eval_multipole(r, theta, phi, Ynm, YnmTheta);
  for (int d = 0; d < 4; d++) trg[d] /= q;
  for (int n = 0; n < P; n++) {
    int nm = n * n + n;
    int nms = n * (n + 1) / 2;
    double re = cx_mul(L[nms], Ynm[nm]).re;
    trg[0] += re;
    sph[0] += re / r * n;
    sph[1] += cx_mul(L[nms], YnmTheta[nm]).re;
    for (int m = 1; m <= n; m++) {
      nm = n * n + n + m;
      nms = n * (n + 1) / 2 + m;
      cx ly = cx_mul(L[nms], Ynm[nm]);
      trg[0] += 2 * ly.re;
      sph[0] += 2 * ly.re / r * n;
      sph[1] += 2 * cx_mul(L[nms], YnmTheta[nm]).re;
      sph[2] += 2 * cx_mul(ly, cx_make(0, 1)).re * m;
    }
  }
  cart[0] = sin(theta) * cos(phi) * sph[0] + cos(theta) * cos(phi) / r * sph[1]
          - sin(phi) / r / sin(theta) * sph[2];
  cart[1] = sin(theta) * sin(phi) * sph[0] + cos(theta) * sin(phi) / r * sph[1]
          + cos(phi) / r / sin(theta) * sph[2];
  cart[2] = cos(theta) * sph[0] - sin(theta) / r * sph[1];
  trg[1] += cart[0];
  trg[2] += cart[1];
  trg[3] += cart[2];
}

/* Raw-array kernel entry points (for per-kernel parity tests). */
void orc_p2m(double *M, const double *cen, const double *x, const double *y, const double *z,
             const double *q, int64_t nb) {
  p2m((cx *)M, cen, x, y, z, q, nb);
}
void orc_m2m(double *Mp, const double *Mc, const double *dX) { m2m((cx *)Mp, (const cx *)Mc, dX); }
void orc_m2l(double *Li, const double *Mj, const double *dX) { m2l((cx *)Li, (const cx *)Mj, dX); }
void orc_l2l(double *Lc, const double *Lp, const double *dX) { l2l((cx *)Lc, (const cx *)Lp, dX); }
void orc_l2p(double *trg, const double *bx, double q, const double *L, const double *cen) {
  l2p(trg, bx, q, (const cx *)L, cen);
}

/* ------------------------------------------------------------------ */
/* Cartesian Taylor basis (EX/kernels/LaplaceCartesianCPU.cxx, the    */
/* reference's BASIS = Cartesian build, EX/include/types.h:39-41)      */
/* ------------------------------------------------------------------ */
/* Terms of degree 0 .. P-1 in three variables: CN = P (P+1) (P+2) / 6.
 * Term k = (kx, ky, kz) of degree d sits at
 *   index(k) = d (d+1) (d+2) / 6 + (d-kx) (d-kx+1) / 2 + kz,
 * i.e. degree by degree, kx falling, then ky falling (the order of the
 * reference's Index<nx,ny,nz>::I, LaplaceCartesianCPU.cxx:3-25).
 *   multipole  M_k = sum_b q_b (X - x_b)^k / k!             (P2M, :636-644)
 *   local      phi(X + e) = sum_n L_n e^n / n!
 *   M2L        L_n += sum_{|k| <= P-1-|n|} D^{n+k}(1/r)(R) M_k,
 *              R = Xi - Xj - Xperiodic                        (:659-673)
 *   M2M        M_n += sum_{k <= n} (Xp - Xc)^{n-k} / (n-k)! M_k  (:646-657)
 *   L2L        L_n += sum_{|k| <= P-1-|n|} (Xc - Xp)^k / k! L_{n+k}  (:675-689)
 *   L2P        trg /= q; pot += sum_k e^k/k! L_k, grad_d += sum_{|k| <= P-2}
 *              e^k/k! L_{k+e_d}, e = x - X                    (:691-704)
 * D^k(1/r) comes from the Taylor coefficients a_k = D^k(1/r) / k! by the
 * recurrence of 1/r (:45-104, :312-324):
 *   |k| r^2 a_k = (1 - 2|k|) sum_i x_i a_{k-e_i} + (1 - |k|) sum_i a_{k-2e_i}. */
#define CN (P * (P + 1) * (P + 2) / 6)

static int cart_index(int x, int y, int z) {
  const int d = x + y + z;
  return d * (d + 1) * (d + 2) / 6 + (d - x) * (d - x + 1) / 2 + z;
}

/* multi-index of every term, in index order */
static void cart_terms(int mi[CN][3]) {
  int t = 0;
  for (int d = 0; d < P; d++)
    for (int x = d; x >= 0; x--)
      for (int y = d - x; y >= 0; y--) {
        mi[t][0] = x;
        mi[t][1] = y;
        mi[t][2] = d - x - y;
        t++;
      }
}

/* C_k = c0 dX^k / k!, each term from its predecessor along z, else y, else x
 * (the reference's power, :167-171, :198-201, :229-232) */
static void cart_powers(double *C, const double dX[3], double c0) {
  int mi[CN][3];
  cart_terms(mi);
  C[0] = c0;
  for (int t = 1; t < CN; t++) {
    const int x = mi[t][0], y = mi[t][1], z = mi[t][2];
    if (z > 0) C[t] = C[cart_index(x, y, z - 1)] * dX[2] / z;
    else if (y > 0) C[t] = C[cart_index(x, y - 1, 0)] * dX[1] / y;
    else C[t] = C[cart_index(x - 1, 0, 0)] * dX[0] / x;
  }
}

/* C_k = D^k (1/r) at dX */
static void cart_derivs(double *C, const double dX[3]) {
  int mi[CN][3];
  cart_terms(mi);
  const double r2 = dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2];
  const double invR2 = 1.0 / r2;
  double a[CN];
  a[0] = sqrt(invR2);
  for (int t = 1; t < CN; t++) {
    const int k[3] = {mi[t][0], mi[t][1], mi[t][2]};
    const int n = k[0] + k[1] + k[2];
    double s1 = 0, s2 = 0;
    for (int i = 0; i < 3; i++) {
      int km[3] = {k[0], k[1], k[2]};
      if (k[i] >= 1) {
        km[i] -= 1;
        s1 += dX[i] * a[cart_index(km[0], km[1], km[2])];
      }
      if (k[i] >= 2) {
        km[i] -= 1;
        s2 += a[cart_index(km[0], km[1], km[2])];
      }
    }
    a[t] = ((1 - 2 * n) * s1 + (1 - n) * s2) / n * invR2;
  }
  for (int t = 0; t < CN; t++) {
    double f = 1;
    for (int i = 0; i < 3; i++)
      for (int j = 2; j <= mi[t][i]; j++) f *= j;
    C[t] = a[t] * f;
  }
}

static void cart_p2m(double *M, const double *cen, const double *x, const double *y,
                     const double *z, const double *q, int64_t nb) {
  double C[CN];
  for (int64_t b = 0; b < nb; b++) {
    const double dX[3] = {cen[0] - x[b], cen[1] - y[b], cen[2] - z[b]};
    cart_powers(C, dX, q[b]);
    for (int t = 0; t < CN; t++) M[t] += C[t];
  }
}

/* dX = parent centre - child centre */
static void cart_m2m(double *Mp, const double *Mc, const double dX[3]) {
  int mi[CN][3];
  cart_terms(mi);
  double C[CN];
  cart_powers(C, dX, 1.0);
  for (int t = 0; t < CN; t++) {
    const int nx = mi[t][0], ny = mi[t][1], nz = mi[t][2];
    double s = 0;
    for (int kx = 0; kx <= nx; kx++)
      for (int ky = 0; ky <= ny; ky++)
        for (int kz = 0; kz <= nz; kz++)
          s += C[cart_index(nx - kx, ny - ky, nz - kz)] * Mc[cart_index(kx, ky, kz)];
    Mp[t] += s;
  }
}

/* L_n += sum_{|k| <= P-1-|n|} C_{n+k} M_k */
static void cart_contract(double *L, const double *C, const double *M) {
  int mi[CN][3];
  cart_terms(mi);
  for (int t = 0; t < CN; t++) {
    const int nx = mi[t][0], ny = mi[t][1], nz = mi[t][2];
    const int left = P - 1 - (nx + ny + nz);
    double s = 0;
    for (int u = 0; u < CN; u++) {
      if (mi[u][0] + mi[u][1] + mi[u][2] > left) break;
      s += C[cart_index(nx + mi[u][0], ny + mi[u][1], nz + mi[u][2])] * M[u];
    }
    L[t] += s;
  }
}

/* dX = target centre - source centre (- Xperiodic) */
static void cart_m2l(double *Li, const double *Mj, const double dX[3]) {
  double C[CN];
  cart_derivs(C, dX);
  cart_contract(Li, C, Mj);
}

/* both directions of one pair: D^k(1/r) at -dX is (-1)^|k| D^k(1/r) at dX
 * (the reference's Coefs::negate, :612-633) */
static void cart_m2l_mutual(double *Li, double *Lj, const double *Mi, const double *Mj,
                            const double dX[3]) {
  int mi[CN][3];
  cart_terms(mi);
  double C[CN];
  cart_derivs(C, dX);
  cart_contract(Li, C, Mj);
  for (int t = 0; t < CN; t++)
    if ((mi[t][0] + mi[t][1] + mi[t][2]) & 1) C[t] = -C[t];
  cart_contract(Lj, C, Mi);
}

/* dX = child centre - parent centre */
static void cart_l2l(double *Lc, const double *Lp, const double dX[3]) {
  double C[CN];
  cart_powers(C, dX, 1.0);
  cart_contract(Lc, Lp, C);
}

static void cart_l2p(double *trg, const double bx[3], double q, const double *L,
                     const double cen[3]) {
  int mi[CN][3];
  cart_terms(mi);
  const double dX[3] = {bx[0] - cen[0], bx[1] - cen[1], bx[2] - cen[2]};
  double C[CN];
  cart_powers(C, dX, 1.0);
  for (int d = 0; d < 4; d++) trg[d] /= q;
  double s[4] = {0, 0, 0, 0};
  for (int t = 0; t < CN; t++) {
    const int x = mi[t][0], y = mi[t][1], z = mi[t][2];
    s[0] += C[t] * L[t];
    if (x + y + z <= P - 2) {
      s[1] += C[t] * L[cart_index(x + 1, y, z)];
      s[2] += C[t] * L[cart_index(x, y + 1, z)];
      s[3] += C[t] * L[cart_index(x, y, z + 1)];
    }
  }
  for (int d = 0; d < 4; d++) trg[d] += s[d];
}

/* Raw-array Cartesian entry points (CN doubles per expansion). */
int orc_cart_terms(void) { return CN; }
int orc_cart_index(int x, int y, int z) { return cart_index(x, y, z); }
void orc_cart_derivs(double *C, const double *dX) { cart_derivs(C, dX); }
void orc_cart_powers(double *C, const double *dX, double c0) { cart_powers(C, dX, c0); }
void orc_cart_p2m(double *M, const double *cen, const double *x, const double *y,
                  const double *z, const double *q, int64_t nb) {
  cart_p2m(M, cen, x, y, z, q, nb);
}
void orc_cart_m2m(double *Mp, const double *Mc, const double *dX) { cart_m2m(Mp, Mc, dX); }
void orc_cart_m2l(double *Li, const double *Mj, const double *dX) { cart_m2l(Li, Mj, dX); }
void orc_cart_m2l_mutual(double *Li, double *Lj, const double *Mi, const double *Mj,
                         const double *dX) {
  cart_m2l_mutual(Li, Lj, Mi, Mj, dX);
}
void orc_cart_l2l(double *Lc, const double *Lp, const double *dX) { cart_l2l(Lc, Lp, dX); }
void orc_cart_l2p(double *trg, const double *bx, double q, const double *L, const double *cen) {
  cart_l2p(trg, bx, q, L, cen);
}

/* ------------------------------------------------------------------ */
/* Tree, traversal and full evaluation                                  */
/* ------------------------------------------------------------------ */

typedef struct {
  uint64_t key;
  int64_t nb, bid;     /* bid = first sorted body of a leaf; 0 for internal cells */
  int32_t leaf, depth;
  int32_t child[8];    /* -1 for leaves */
  int32_t parent;
  double cmin[3], cmax[3], center[3];
} ocell;

struct orc_ctx {
  int64_t n;
  int ncrit;
  double theta;
  double rmin[3], rmax[3];
  /* sorted bodies */
  uint64_t *keys;
  int64_t *perm;        /* sorted position -> original index */
  double *x, *y, *z, *q;
  double *trg;          /* 4 per sorted body */
  /* cells in creation (depth-first) order */
  ocell *cells;
  int64_t ncells, cap;
  cx *M, *L;            /* NTERM per cell */
  int max_depth;
  /* counters */
  uint64_t n_m2l, n_p2p_leafpairs, n_p2p_bodypairs, n_p2p_selfpairs;
  uint64_t n_m2l_rec, n_p2p_rec; /* entries of the recorded (directed) lists */
  /* mutual interaction (the reference's --mutual 1) */
  int mutual;
  /* optional interaction lists (cell indices) */
  int record;
  int64_t *m2l_list, n_m2l_cap;
  int64_t *p2p_list, n_p2p_cap;
  /* periodic images: the image the traversal is in (xp = its Xperiodic,
   * img = its code 9 (ix+1) + 3 (iy+1) + (iz+1), recorded in the top bits
   * of list sources as the library's lists hold it) */
  double xp[3];
  int64_t img;
  /* expansion basis: 0 spherical (M, L: NTERM complex per cell), 1 Cartesian
   * (Mc, Lc: CN doubles per cell) */
  int basis;
  double *Mc, *Lc;
};

static int64_t lower_bound_u64(const uint64_t *a, int64_t lo, int64_t hi, uint64_t v) {
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static int64_t new_cell(orc_ctx *c) {
  if (c->ncells == c->cap) {
    c->cap = c->cap ? c->cap * 2 : 1024;
    c->cells = (ocell *)realloc(c->cells, sizeof(ocell) * c->cap);
  }
  return c->ncells++;
}

static int64_t generate_cell(orc_ctx *c, uint64_t k, int64_t pbeg, int64_t pend, int32_t parent) {
  uint64_t k2 = key_next(k);
  int64_t rb = lower_bound_u64(c->keys, pbeg, pend, k);
  int64_t re = lower_bound_u64(c->keys, pbeg, pend, k2);
  int64_t nb = re - rb;
  int depth = key_depth(k);
  /* A cell at the finest depth cannot be split further; the reference would
   * form invalid child keys there.  Such a cell stays a leaf. */
  int is_leaf = (nb <= c->ncrit) || depth == MAXDEPTH;
  int64_t idx = new_cell(c);
  ocell *cell = &c->cells[idx];
  cell->key = k;
  cell->nb = nb;
  cell->bid = is_leaf ? rb : 0;
  cell->leaf = is_leaf;
  cell->depth = depth;
  cell->parent = parent;
  orc_cell_region(k, c->rmin, c->rmax, cell->cmin, cell->cmax);
  for (int d = 0; d < 3; d++) cell->center[d] = (cell->cmax[d] + cell->cmin[d]) / 2;
  for (int i = 0; i < 8; i++) cell->child[i] = -1;
  if (depth > c->max_depth) c->max_depth = depth;
  if (!is_leaf) {
    for (int i = 0; i < 8; i++) {
      int64_t ch = generate_cell(c, key_child(k, i), rb, re, (int32_t)idx);
      c->cells[idx].child[i] = (int32_t)ch;
    }
  }
  return idx;
}

typedef struct { uint64_t key; int64_t idx; } kidx;
static int cmp_kidx(const void *a, const void *b) {
  const kidx *x = (const kidx *)a, *y = (const kidx *)b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

static double cell_width(const ocell *c, int d) { return c->cmax[d] - c->cmin[d]; }

static void upward_cart(orc_ctx *c, int64_t ci) {
  ocell *cell = &c->cells[ci];
  double *M = c->Mc + ci * CN;
  for (int i = 0; i < CN; i++) M[i] = c->Lc[ci * CN + i] = 0;
  if (cell->leaf) {
    int64_t b = cell->bid;
    cart_p2m(M, cell->center, c->x + b, c->y + b, c->z + b, c->q + b, cell->nb);
  } else {
    for (int i = 0; i < 8; i++) upward_cart(c, cell->child[i]);
    for (int i = 0; i < 8; i++) {
      const ocell *ch = &c->cells[cell->child[i]];
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = cell->center[d] - ch->center[d];
      cart_m2m(M, c->Mc + (int64_t)cell->child[i] * CN, dX);
    }
  }
}

static void upward(orc_ctx *c, int64_t ci) {
  if (c->basis) {
    upward_cart(c, ci);
    return;
  }
  ocell *cell = &c->cells[ci];
  cx *M = c->M + ci * NTERM;
  for (int i = 0; i < NTERM; i++) { M[i] = cx_make(0, 0); c->L[ci * NTERM + i] = cx_make(0, 0); }
  if (cell->leaf) {
    int64_t b = cell->bid;
    p2m(M, cell->center, c->x + b, c->y + b, c->z + b, c->q + b, cell->nb);
  } else {
    for (int i = 0; i < 8; i++) upward(c, cell->child[i]);
    for (int i = 0; i < 8; i++) {
      const ocell *ch = &c->cells[cell->child[i]];
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = cell->center[d] - ch->center[d];
      m2m(M, c->M + (int64_t)cell->child[i] * NTERM, dX);
    }
  }
}

static void push_pair(int64_t **list, int64_t *cap, uint64_t cnt, int64_t a, int64_t b) {
  if ((int64_t)(2 * cnt + 2) > *cap) {
    *cap = *cap ? *cap * 2 : 4096;
    *list = (int64_t *)realloc(*list, sizeof(int64_t) * (*cap));
  }
  (*list)[2 * cnt] = a;
  (*list)[2 * cnt + 1] = b;
}

/* P2P functor on one body pair (LaplaceP2PCPU_tapas.h:48-74): target i, and
 * with mutual the source j too when the positions differ; dX = Xi - Xj -
 * Xperiodic (:54) */
static void p2p_pair(orc_ctx *c, int64_t i, int64_t j, int mutual) {
  double dX[3] = {c->x[i] - c->x[j] - c->xp[0], c->y[i] - c->y[j] - c->xp[1],
                  c->z[i] - c->z[j] - c->xp[2]};
  double R2 = 0;
  for (int d = 0; d < 3; d++) R2 += dX[d] * dX[d];
  R2 += 0.0; /* softening EPS2 = 0 */
  if (R2 != 0) {
    double invR2 = 1.0 / R2;
    double invR = c->q[i] * c->q[j] * sqrt(invR2);
    for (int d = 0; d < 3; d++) dX[d] *= invR2 * invR;
    c->trg[4 * i + 0] += invR;
    c->trg[4 * i + 1] -= dX[0];
    c->trg[4 * i + 2] -= dX[1];
    c->trg[4 * i + 3] -= dX[2];
    if (mutual && (c->x[i] != c->x[j] || c->y[i] != c->y[j] || c->z[i] != c->z[j])) {
      c->trg[4 * j + 0] += invR;
      c->trg[4 * j + 1] += dX[0];
      c->trg[4 * j + 2] += dX[1];
      c->trg[4 * j + 3] += dX[2];
    }
    c->n_p2p_bodypairs++;
  } else {
    c->n_p2p_selfpairs++;
  }
}

/* Map(P2P, Product(Ci.bodies(), Cj.bodies())): every pair, or with mutual
 * on one leaf the pairs j <= i only (TP/hot/mapper.h:126-169) */
static void p2p_leaves(orc_ctx *c, const ocell *ci, const ocell *cj) {
  const int same = c->mutual && ci == cj;
  for (int64_t i = ci->bid; i < ci->bid + ci->nb; i++) {
    const int64_t jend = same ? i + 1 : cj->bid + cj->nb;
    for (int64_t j = cj->bid; j < jend; j++) p2p_pair(c, i, j, c->mutual);
  }
}

/* FMM_DTT (tapas_exafmm.cxx:209-286), in the periodic image c->xp
 * (EX/include/traversal.h:36-62: dX = Xi - Xj - Xperiodic) */
static void dtt(orc_ctx *c, int64_t ii, int64_t jj) {
  const ocell *Ci = &c->cells[ii], *Cj = &c->cells[jj];
  double d0 = Ci->center[0] - Cj->center[0] - c->xp[0];
  double d1 = Ci->center[1] - Cj->center[1] - c->xp[1];
  double d2 = Ci->center[2] - Cj->center[2] - c->xp[2];
  double R2 = 0;
  R2 += d0 * d0;
  R2 += d1 * d1;
  R2 += d2 * d2;
  double Ri = 0, Rj = 0;
  for (int d = 0; d < 3; d++) {
    double wi = cell_width(Ci, d), wj = cell_width(Cj, d);
    Ri = (Ri < wi) ? wi : Ri;
    Rj = (Rj < wj) ? wj : Rj;
  }
  Ri = (Ri / 2 * 1.00001) / c->theta;
  Rj = (Rj / 2 * 1.00001) / c->theta;
  if (R2 > (Ri + Rj) * (Ri + Rj)) {
    double dX[3] = {d0, d1, d2};
    if (c->basis && c->mutual)
      cart_m2l_mutual(c->Lc + ii * CN, c->Lc + jj * CN, c->Mc + ii * CN, c->Mc + jj * CN, dX);
    else if (c->basis)
      cart_m2l(c->Lc + ii * CN, c->Mc + jj * CN, dX);
    else if (c->mutual)
      m2l_mutual(c->L + ii * NTERM, c->L + jj * NTERM, c->M + ii * NTERM, c->M + jj * NTERM, dX);
    else
      m2l(c->L + ii * NTERM, c->M + jj * NTERM, dX);
    /* lists hold directed (target, source) pairs: a mutual call is two */
    if (c->record) push_pair(&c->m2l_list, &c->n_m2l_cap, c->n_m2l_rec++, ii, jj | (c->img << 27));
    if (c->record && c->mutual) push_pair(&c->m2l_list, &c->n_m2l_cap, c->n_m2l_rec++, jj, ii);
    c->n_m2l++;
  } else if (Ci->leaf && Cj->leaf) {
    if (Ci->nb > 0 && Cj->nb > 0) {
      if (c->record) push_pair(&c->p2p_list, &c->n_p2p_cap, c->n_p2p_rec++, ii, jj | (c->img << 27));
      if (c->record && c->mutual && ii != jj)
        push_pair(&c->p2p_list, &c->n_p2p_cap, c->n_p2p_rec++, jj, ii);
      c->n_p2p_leafpairs++;
      p2p_leaves(c, Ci, Cj);
    }
  } else if (Cj->leaf) {
    for (int i = 0; i < 8; i++) dtt(c, c->cells[ii].child[i], jj);
  } else if (Ci->leaf) {
    for (int j = 0; j < 8; j++) dtt(c, ii, c->cells[jj].child[j]);
  } else if (c->mutual && ii == jj) {
    /* both sides split, each unordered pair of children once
     * (tapas_exafmm.cxx:270-271, TP/hot/mapper.h:29-121 with i <= j) */
    for (int i = 0; i < 8; i++)
      for (int j = i; j < 8; j++) dtt(c, c->cells[ii].child[i], c->cells[jj].child[j]);
  } else if (Ri >= Rj) {
    for (int i = 0; i < 8; i++) dtt(c, c->cells[ii].child[i], jj);
  } else {
    for (int j = 0; j < 8; j++) dtt(c, ii, c->cells[jj].child[j]);
  }
}

static void downward(orc_ctx *c, int64_t ci) {
  ocell *cell = &c->cells[ci];
  if (cell->parent >= 0) {
    const ocell *par = &c->cells[cell->parent];
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = cell->center[d] - par->center[d];
    if (c->basis)
      cart_l2l(c->Lc + ci * CN, c->Lc + (int64_t)cell->parent * CN, dX);
    else
      l2l(c->L + ci * NTERM, c->L + (int64_t)cell->parent * NTERM, dX);
  }
  if (cell->leaf) {
    for (int64_t b = cell->bid; b < cell->bid + cell->nb; b++) {
      double bx[3] = {c->x[b], c->y[b], c->z[b]};
      if (c->basis)
        cart_l2p(c->trg + 4 * b, bx, c->q[b], c->Lc + ci * CN, cell->center);
      else
        l2p(c->trg + 4 * b, bx, c->q[b], c->L + ci * NTERM, cell->center);
    }
  } else {
    for (int i = 0; i < 8; i++) downward(c, cell->child[i]);
  }
}

orc_ctx *orc_run(const double *x, const double *y, const double *z, const double *q, int64_t n,
                 int ncrit, double theta, int phases, int record_lists) {
  return orc_run_mutual(x, y, z, q, n, ncrit, theta, phases, record_lists, 0);
}

/* The tree over the bodies (SamplingOctree::Build on one rank); with
 * cycle > 0 the root region is the periodic cell: a cube of side cycle
 * centred on the bodies' scalar bounding cube (NULL if they span more). */
static orc_ctx *setup(const double *x, const double *y, const double *z, const double *q,
                      int64_t n, int ncrit, double theta, int record_lists, int mutual,
                      double cycle, int basis) {
  if (n <= 0 || ncrit <= 0 || !(theta > 0) || basis < 0 || basis > 1) return NULL;
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->mutual = mutual ? 1 : 0;
  c->basis = basis;
  c->n = n;
  c->ncrit = ncrit;
  c->theta = theta;
  c->record = record_lists;
  orc_bbox(x, y, z, n, c->rmin, c->rmax);
  if (cycle > 0) {
    const double lo = c->rmin[0], hi = c->rmax[0];
    if (!(hi - lo <= cycle)) {
      free(c);
      return NULL;
    }
    const double mid = 0.5 * (lo + hi), h = 0.5 * cycle;
    for (int d = 0; d < 3; d++) {
      c->rmin[d] = mid - h;
      c->rmax[d] = mid + h;
    }
  }
  uint64_t *k0 = (uint64_t *)malloc(sizeof(uint64_t) * n);
  orc_keys(x, y, z, n, c->rmin, c->rmax, k0);
  kidx *ki = (kidx *)malloc(sizeof(kidx) * n);
  for (int64_t i = 0; i < n; i++) { ki[i].key = k0[i]; ki[i].idx = i; }
  qsort(ki, (size_t)n, sizeof(kidx), cmp_kidx);
  c->keys = k0;
  c->perm = (int64_t *)malloc(sizeof(int64_t) * n);
  c->x = (double *)malloc(sizeof(double) * n);
  c->y = (double *)malloc(sizeof(double) * n);
  c->z = (double *)malloc(sizeof(double) * n);
  c->q = (double *)malloc(sizeof(double) * n);
  c->trg = (double *)calloc((size_t)n * 4, sizeof(double));
  for (int64_t i = 0; i < n; i++) {
    int64_t o = ki[i].idx;
    c->keys[i] = ki[i].key;
    c->perm[i] = o;
    c->x[i] = x[o]; c->y[i] = y[o]; c->z[i] = z[o]; c->q[i] = q[o];
  }
  free(ki);
  generate_cell(c, 0, 0, n, -1);
  if (basis) {
    c->Mc = (double *)calloc((size_t)c->ncells * CN, sizeof(double));
    c->Lc = (double *)calloc((size_t)c->ncells * CN, sizeof(double));
  } else {
    c->M = (cx *)calloc((size_t)c->ncells * NTERM, sizeof(cx));
    c->L = (cx *)calloc((size_t)c->ncells * NTERM, sizeof(cx));
  }
  return c;
}

orc_ctx *orc_run_mutual(const double *x, const double *y, const double *z, const double *q,
                        int64_t n, int ncrit, double theta, int phases, int record_lists,
                        int mutual) {
  return orc_run_basis(x, y, z, q, n, ncrit, theta, phases, record_lists, mutual, 0);
}

orc_ctx *orc_run_basis(const double *x, const double *y, const double *z, const double *q,
                       int64_t n, int ncrit, double theta, int phases, int record_lists,
                       int mutual, int basis) {
  orc_ctx *c = setup(x, y, z, q, n, ncrit, theta, record_lists, mutual, 0.0, basis);
  if (!c) return NULL;
  if (phases >= 1) upward(c, 0);
  if (phases >= 2) dtt(c, 0, 0);
  if (phases >= 3) downward(c, 0);
  return c;
}

/* The far field of the periodic images beyond the first shell
 * (EX/include/traversal.h:126-185): Ms, the multipole of a block of copies
 * (about the root's centre), starts as the root's; per sublevel the 26 x 27
 * copies at (3 i + c) cycle (i in {-1,0,1}^3 without 0, c in {-1,0,1}^3) add
 * their M2L into the root's local expansion (dX = Xroot - Xroot - Xperiodic),
 * then the 27 copies i cycle shift into the block three times larger
 * (M2M, dX = parent - child centre; the reference's loop resets the parent
 * before this M2M, leaving the centre copy out: the sum over all
 * (3^images)^3 copies, which its direct check computes, needs it). */
static void periodic_far_cart(orc_ctx *c, int images, double cycle) {
  double Ms[CN], Mn[CN];
  memcpy(Ms, c->Mc, sizeof(Ms));
  for (int level = 0; level < images - 1; level++) {
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          if (ix == 0 && iy == 0 && iz == 0) continue;
          for (int cx_ = -1; cx_ <= 1; cx_++)
            for (int cy = -1; cy <= 1; cy++)
              for (int cz = -1; cz <= 1; cz++) {
                const double dX[3] = {-((ix * 3 + cx_) * cycle), -((iy * 3 + cy) * cycle),
                                      -((iz * 3 + cz) * cycle)};
                cart_m2l(c->Lc, Ms, dX);
              }
        }
    memset(Mn, 0, sizeof(Mn));
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          const double dX[3] = {-(ix * cycle), -(iy * cycle), -(iz * cycle)};
          cart_m2m(Mn, Ms, dX);
        }
    memcpy(Ms, Mn, sizeof(Ms));
    cycle *= 3;
  }
}

static void periodic_far(orc_ctx *c, int images, double cycle) {
  if (c->basis) {
    periodic_far_cart(c, images, cycle);
    return;
  }
  cx Ms[NTERM], Mn[NTERM];
  memcpy(Ms, c->M, sizeof(Ms));
  for (int level = 0; level < images - 1; level++) {
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          if (ix == 0 && iy == 0 && iz == 0) continue;
          for (int cx_ = -1; cx_ <= 1; cx_++)
            for (int cy = -1; cy <= 1; cy++)
              for (int cz = -1; cz <= 1; cz++) {
                const double dX[3] = {-((ix * 3 + cx_) * cycle), -((iy * 3 + cy) * cycle),
                                      -((iz * 3 + cz) * cycle)};
                m2l(c->L, Ms, dX);
              }
        }
    memset(Mn, 0, sizeof(Mn));
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          const double dX[3] = {-(ix * cycle), -(iy * cycle), -(iz * cycle)};
          m2m(Mn, Ms, dX);
        }
    memcpy(Ms, Mn, sizeof(Ms));
    cycle *= 3;
  }
}

/* The FMM with periodic images (EX/include/traversal.h:247-268, images >= 1):
 * the traversal from the root and each of its 27 images (Xperiodic =
 * (ix, iy, iz) cycle, the reference's loop order, mutual = 0), then the
 * farther copies (images >= 2) into the root's local expansion. */
orc_ctx *orc_run_periodic(const double *x, const double *y, const double *z, const double *q,
                          int64_t n, int ncrit, double theta, int images, double cycle,
                          int phases, int record_lists) {
  return orc_run_periodic_basis(x, y, z, q, n, ncrit, theta, images, cycle, phases,
                                record_lists, 0);
}

orc_ctx *orc_run_periodic_basis(const double *x, const double *y, const double *z,
                                const double *q, int64_t n, int ncrit, double theta, int images,
                                double cycle, int phases, int record_lists, int basis) {
  if (images < 1 || !(cycle > 0)) return NULL;
  orc_ctx *c = setup(x, y, z, q, n, ncrit, theta, record_lists, 0, cycle, basis);
  if (!c) return NULL;
  if (phases >= 1) upward(c, 0);
  if (phases >= 2) {
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          c->xp[0] = ix * cycle;
          c->xp[1] = iy * cycle;
          c->xp[2] = iz * cycle;
          c->img = 9 * (ix + 1) + 3 * (iy + 1) + (iz + 1);
          dtt(c, 0, 0);
        }
    c->xp[0] = c->xp[1] = c->xp[2] = 0;
    c->img = 0;
    periodic_far(c, images, cycle);
  }
  if (phases >= 3) downward(c, 0);
  return c;
}

void orc_free(orc_ctx *c) {
  if (!c) return;
  free(c->keys); free(c->perm); free(c->x); free(c->y); free(c->z); free(c->q); free(c->trg);
  free(c->cells); free(c->M); free(c->L); free(c->Mc); free(c->Lc);
  free(c->m2l_list); free(c->p2p_list);
  free(c);
}

void orc_region(const orc_ctx *c, double *rmin, double *rmax) {
  for (int d = 0; d < 3; d++) { rmin[d] = c->rmin[d]; rmax[d] = c->rmax[d]; }
}
int64_t orc_ncells(const orc_ctx *c) { return c->ncells; }
int orc_max_depth(const orc_ctx *c) { return c->max_depth; }

void orc_sorted(const orc_ctx *c, uint64_t *keys, int64_t *perm) {
  if (keys) memcpy(keys, c->keys, sizeof(uint64_t) * c->n);
  if (perm) memcpy(perm, c->perm, sizeof(int64_t) * c->n);
}

void orc_cells(const orc_ctx *c, uint64_t *keys, int64_t *nb, int64_t *bid, int32_t *leaf,
               double *center, double *width) {
  for (int64_t i = 0; i < c->ncells; i++) {
    const ocell *e = &c->cells[i];
    if (keys) keys[i] = e->key;
    if (nb) nb[i] = e->nb;
    if (bid) bid[i] = e->bid;
    if (leaf) leaf[i] = e->leaf;
    for (int d = 0; d < 3; d++) {
      if (center) center[3 * i + d] = e->center[d];
      if (width) width[3 * i + d] = cell_width(e, d);
    }
  }
}

/* NTERM complex (spherical) or CN doubles (Cartesian) per cell */
void orc_expansions(const orc_ctx *c, double *M, double *L) {
  const size_t bytes = c->basis ? sizeof(double) * (size_t)c->ncells * CN
                                : sizeof(cx) * (size_t)c->ncells * NTERM;
  if (M) memcpy(M, c->basis ? (const void *)c->Mc : (const void *)c->M, bytes);
  if (L) memcpy(L, c->basis ? (const void *)c->Lc : (const void *)c->L, bytes);
}

/* fields in the caller's original body order */
void orc_fields(const orc_ctx *c, double *trg4) {
  for (int64_t i = 0; i < c->n; i++) {
    int64_t o = c->perm[i];
    for (int d = 0; d < 4; d++) trg4[4 * o + d] = c->trg[4 * i + d];
  }
}

void orc_counts(const orc_ctx *c, uint64_t *out4) {
  out4[0] = c->n_m2l;
  out4[1] = c->n_p2p_leafpairs;
  out4[2] = c->n_p2p_bodypairs;
  out4[3] = c->n_p2p_selfpairs;
}

/* recorded lists as (target cell index, source cell index) pairs */
void orc_lists(const orc_ctx *c, int64_t *m2l, int64_t *p2p) {
  if (m2l && c->m2l_list) memcpy(m2l, c->m2l_list, sizeof(int64_t) * 2 * c->n_m2l_rec);
  if (p2p && c->p2p_list) memcpy(p2p, c->p2p_list, sizeof(int64_t) * 2 * c->n_p2p_rec);
}

/* entries of the recorded lists (directed pairs: a mutual call is two) */
void orc_list_sizes(const orc_ctx *c, int64_t *out2) {
  out2[0] = (int64_t)c->n_m2l_rec;
  out2[1] = (int64_t)c->n_p2p_rec;
}

/* FP64 direct sum on a target subset, normalised by the target charge as the
 * reference's CheckResult does (tapas_exafmm.cxx:338-367). */
void orc_direct(const double *x, const double *y, const double *z, const double *q, int64_t n,
                const int64_t *targets, int64_t nt, double *trg4) {
  for (int64_t t = 0; t < nt; t++) {
    int64_t i = targets[t];
    double acc[4] = {0, 0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      double dX[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
      double R2 = 0;
      for (int d = 0; d < 3; d++) R2 += dX[d] * dX[d];
      if (R2 != 0) {
        double invR2 = 1.0 / R2;
        double invR = q[i] * q[j] * sqrt(invR2);
        for (int d = 0; d < 3; d++) dX[d] *= invR2 * invR;
        acc[0] += invR;
        acc[1] -= dX[0];
        acc[2] -= dX[1];
        acc[3] -= dX[2];
      }
    }
    for (int d = 0; d < 4; d++) trg4[4 * t + d] = acc[d] / q[i];
  }
}

/* The direct sum over the copies of the bodies at (ix, iy, iz) cycle, |ix|,
 * |iy|, |iz| <= prange, prange = sum_{i < images} 3^i, in the reference's
 * CheckResult order (tapas_exafmm.cxx:314-345), normalised by the target
 * charge. */
void orc_direct_periodic(const double *x, const double *y, const double *z, const double *q,
                         int64_t n, const int64_t *targets, int64_t nt, int images, double cycle,
                         double *trg4) {
  int prange = 0, p3 = 1;
  for (int i = 0; i < images; i++, p3 *= 3) prange += p3;
  for (int64_t t = 0; t < nt; t++) {
    int64_t i = targets[t];
    double acc[4] = {0, 0, 0, 0};
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          const double xp[3] = {ix * cycle, iy * cycle, iz * cycle};
          for (int64_t j = 0; j < n; j++) {
            double dX[3] = {x[i] - x[j] - xp[0], y[i] - y[j] - xp[1], z[i] - z[j] - xp[2]};
            double R2 = 0;
            for (int d = 0; d < 3; d++) R2 += dX[d] * dX[d];
            if (R2 != 0) {
              double invR2 = 1.0 / R2;
              double invR = q[i] * q[j] * sqrt(invR2);
              for (int d = 0; d < 3; d++) dX[d] *= invR2 * invR;
              acc[0] += invR;
              acc[1] -= dX[0];
              acc[2] -= dX[1];
              acc[3] -= dX[2];
            }
          }
        }
    for (int d = 0; d < 4; d++) trg4[4 * t + d] = acc[d] / q[i];
  }
}

/* rel-L2 error of potential and acceleration (verify.h:18-92) */
void orc_rel_l2(const double *trg, const double *ref, int64_t n, double *err2) {
  double pd = 0, pn = 0, ad = 0, an = 0;
  for (int64_t i = 0; i < n; i++) {
    double dp = trg[4 * i] - ref[4 * i];
    pd += dp * dp;
    pn += ref[4 * i] * ref[4 * i];
    double dx = trg[4 * i + 1] - ref[4 * i + 1];
    double dy = trg[4 * i + 2] - ref[4 * i + 2];
    double dz = trg[4 * i + 3] - ref[4 * i + 3];
    ad += dx * dx + dy * dy + dz * dz;
    an += ref[4 * i + 1] * ref[4 * i + 1] + ref[4 * i + 2] * ref[4 * i + 2] +
          ref[4 * i + 3] * ref[4 * i + 3];
  }
  err2[0] = sqrt(pd / pn);
  err2[1] = sqrt(ad / an);
}

/* key algebra entry points (for the known-answer tests of the key functions) */
uint64_t orc_key_child(uint64_t k, int idx) { return key_child(k, idx); }
uint64_t orc_key_next(uint64_t k) { return key_next(k); }
