// translate.hip -- FP64 M2M and L2L as per-level matrix products on the
// gfx950 matrix cores.
//
// Reference (sample/exafmm-dev-13274dd4ac68/examples/):
//   M2M  LaplaceSphericalCPU_tapas.h:165-206  (dX = parent - child centre)
//   L2L  LaplaceSphericalCPU_tapas.h:309-345  (dX = child - parent centre)
//   called per cell by FMM_Upward / FMM_Downward, tapas_exafmm.cxx:87-140.
// Both are real-linear maps of 55 complex (110 real) coefficients whose
// entries are regular harmonics of dX: the (j, n) entry of M2M scales as
// |dX|^(j-n), that of L2L as |dX|^(n-j).  The 8 children of a parent sit at
// dX = sigma_p u_o (sigma_p the parent's width, u_o = (+-1/4, +-1/4, +-1/4) for
// octant o), so with operators T_o = T(u_o), S_o = S(u_o) built once:
//   M2M:  M_p[j]   += sigma_p^j    sum_o sum_n T_o[j][n] sigma_p^-n M_c(o)[n]
//   L2L:  L_c(o)[j] += sigma_p^-j  sum_n S_o[j][n] sigma_p^n L_p[n]
// i.e. one product per level, [T_0 .. T_7] (110 x 880) against the stacked
// children coefficients (the 8 children of a parent are adjacent cells, one
// contiguous column of 880 doubles), with per-parent column and row factors.
// Parents of any width (different levels, or different problems of a batch)
// share the operators.  The operators are formed by applying the reference's
// formulas to unit coefficient vectors; the products run on
// v_mfma_f64_16x16x4 in FP64.  (dX of the reference comes from the cell
// centres, which equal sigma_p u_o up to rounding.)
#include <algorithm>

#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NV = 2 * NTERM;  // 110 reals per cell
constexpr int KM = 8 * NV;     // 880: stacked children
constexpr int KS = (NV + 3) / 4;  // 28 k-steps of 4
// The operators again in the order the MFMA A operand is read: for each
// (row block, k-step, octant) of M2M and (octant, row block, k-step) of L2L,
// 64 consecutive doubles, lane l holding A[16 rb + l % 16][4 ks + l / 16]
// (zero outside the 110 x 110 blocks).  One k-step of one wave then reads
// 512 contiguous bytes per operand instead of 16 scattered rows, so a block
// streams each operator once through its CU instead of four times.
constexpr int FRAG_M = 7 * KS * 8 * 64;
constexpr int FRAG_L = 8 * 7 * KS * 64;

__device__ inline int sgn(int n) { return (n & 1) ? -1 : 1; }
__device__ inline int ipow2n(int n) { return (n >= 0) ? 1 : sgn(n); }

__device__ inline void untri(int t, int &j, int &k) {
  j = 0;
  while (tri(j + 1, 0) <= t) j++;
  k = t - tri(j, 0);
}

// M2M output coefficient (j, k) of one child with coefficients Mc and
// harmonics Y of dX (LaplaceSphericalCPU_tapas.h:178-200)
__device__ inline cd m2m_entry(int j, int k, const cd *Mc, const cd *Y) {
  cd acc = mk(0, 0);
  for (int n = 0; n <= j; n++) {
    const int mlo = max(-n, -j + k + n), mhi = min(k - 1, n);
    for (int m = mlo; m <= mhi; m++) {
      const cd t = cmul(Mc[tri(j - n, k - m)], yfull(Y, n, -m));
      acc = cadd(acc, cscale(t, (double)(ipow2n(m) * sgn(n))));
    }
    const int mhi2 = min(n, j + k - n);
    for (int m = k; m <= mhi2; m++) {
      const cd t = cmul(cconj(Mc[tri(j - n, m - k)]), yfull(Y, n, -m));
      acc = cadd(acc, cscale(t, (double)sgn(k + n + m)));
    }
  }
  return acc;
}

// L2L output coefficient (j, k) from the parent's Lp and harmonics Y of dX
// (LaplaceSphericalCPU_tapas.h:322-341)
__device__ inline cd l2l_entry(int j, int k, const cd *Lp, const cd *Y) {
  cd acc = mk(0, 0);
  for (int n = j; n < P; n++) {
    for (int m = j + k - n; m < 0; m++) {
      const cd t = cmul(cconj(Lp[tri(n, -m)]), yfull(Y, n - j, m - k));
      acc = cadd(acc, cscale(t, (double)sgn(k)));
    }
    for (int m = 0; m <= n; m++) {
      if (n - j >= abs(m - k)) {
        const cd t = cmul(Lp[tri(n, m)], yfull(Y, n - j, m - k));
        acc = cadd(acc, cscale(t, (double)sgn((m - k) * (m < k))));
      }
    }
  }
  return acc;
}

__host__ __device__ constexpr int deg_of_coef(int t) {
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= t) n++;
  return n;
}
__host__ __device__ constexpr int m2m_kend(int rb) {
  const int last = (16 * rb + 15 < NV - 1 ? 16 * rb + 15 : NV - 1) >> 1;
  const int d = deg_of_coef(last);
  return (d + 1) * (d + 2) < NV ? (d + 1) * (d + 2) : NV;
}
__host__ __device__ constexpr int l2l_kbeg(int rb) {
  const int d = deg_of_coef((16 * rb) >> 1);
  return (d * (d + 1)) & ~3;
}

// The unit operators.  Block (column, octant, kind): the image of unit input
// vector `column` under M2M (kind 0) or L2L (kind 1) of the octant.
//   Tm: [110 rows][880 = octant * 110 + child column]
//   Tl: [octant][110 rows][110 parent columns]
__global__ __launch_bounds__(64) void k_build_ops(double *Tm, double *Tl) {
  const int col = blockIdx.x, oct = blockIdx.y, kind = blockIdx.z;
  const int lane = threadIdx.x;
  __shared__ cd Y[NTERM];
  __shared__ cd U[NTERM];
  if (lane == 0) {
    // child centre - parent centre for a parent of width 1
    const double u[3] = {(oct & 1) ? 0.25 : -0.25, (oct & 2) ? 0.25 : -0.25,
                         (oct & 4) ? 0.25 : -0.25};
    const double sg = kind == 0 ? -1.0 : 1.0;  // M2M: parent - child; L2L: child - parent
    double rho, ca, sa;
    cd ei;
    sph_frame(sg * u[0], sg * u[1], sg * u[2], rho, ca, sa, ei);
    multipole_all<false>(rho, ca, sa, ei, [&](int n, int m, cd v, cd) { Y[tri(n, m)] = v; });
  }
  if (lane < NTERM) U[lane] = mk(lane == (col >> 1) && !(col & 1) ? 1.0 : 0.0,
                                 lane == (col >> 1) && (col & 1) ? 1.0 : 0.0);
  __syncthreads();
  if (lane >= NTERM) return;
  int j, k;
  untri(lane, j, k);
  if (kind == 0) {
    const cd e = m2m_entry(j, k, U, Y);
    Tm[(int64_t)(2 * lane) * KM + oct * NV + col] = e.re;
    Tm[(int64_t)(2 * lane + 1) * KM + oct * NV + col] = e.im;
  } else {
    const cd e = l2l_entry(j, k, U, Y);
    Tl[((int64_t)oct * NV + 2 * lane) * NV + col] = e.re;
    Tl[((int64_t)oct * NV + 2 * lane + 1) * NV + col] = e.im;
  }
}

__global__ void k_frag_ops(const double *Tm, const double *Tl, double *Fm, double *Fl) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < FRAG_M) {
    const int lane = e & 63, o = (e >> 6) & 7, ks = (e >> 9) % KS, rb = (e >> 9) / KS;
    const int row = rb * 16 + (lane & 15), kc = 4 * ks + (lane >> 4);
    Fm[e] = (row < NV && kc < NV) ? Tm[(int64_t)row * KM + o * NV + kc] : 0.0;
  } else if (e < FRAG_M + FRAG_L) {
    const int f = e - FRAG_M;
    const int lane = f & 63, ks = (f >> 6) % KS, rb = ((f >> 6) / KS) % 7, oct = (f >> 6) / (KS * 7);
    const int row = rb * 16 + (lane & 15), kc = 4 * ks + (lane >> 4);
    Fl[f] = (row < NV && kc < NV) ? Tl[((int64_t)oct * NV + row) * NV + kc] : 0.0;
  }
}

// per-parent factors sigma^(e * degree(k)) for k = 0..111 of the 16 parents
// of a block (e = +1 or -1), into LDS.  A parent of zero width (all its
// bodies at one point) keeps only the degree-0 terms: its children sit at
// dX = 0, where the shifts reduce to the identity on coefficients that are
// zero above degree 0 anyway.
__device__ inline void parent_powers(double (*f)[NV + 2], const double *sig, int e) {
  for (int t = threadIdx.x; t < 16 * (NV + 2); t += blockDim.x) {
    const int colp = t / (NV + 2), kk = t % (NV + 2);
    const int d = kk < NV ? deg_of_coef(kk >> 1) : 0;
    const double sg = sig[colp];
    double v = 1;
    if (sg == 0) {
      v = d == 0 ? 1.0 : 0.0;
    } else {
      const double s = e > 0 ? sg : 1.0 / sg;
      for (int i = 0; i < d; i++) v *= s;
    }
    f[colp][kk] = v;
  }
}

// MFMA f64 16x16x4 operands: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; result i of lane l is D[l / 16 + 4 i][l % 16].
//
// Degree bounds of the 7 row blocks of 16 real coefficients: M2M row block
// rb only takes child columns below m2m_kend(rb) (child degree <= parent
// degree), L2L row block rb only parent columns from l2l_kbeg(rb) on
// (parent degree >= child degree; rounded down to the k-step of 4).
// With lev, the children's level dl is read on the device (first cell
// lev[2 dl], lev[2 dl + 1] cells, 8 per parent) and the launch's npar is only
// the capacity its grid was sized by.  Both variants of a level are then
// launched and the level's size picks the one that runs, as the host picks
// it from a known size: sel > 0 runs only for fewer than sel blocks of 16
// parents, sel < 0 only for at least -sel.
__device__ inline bool level_extent(const int32_t *lev, int dl, int sel, int32_t &first_child,
                                    int32_t &npar) {
  if (!lev) return true;
  first_child = lev[2 * dl];
  const int32_t act = lev[2 * dl + 1] >> 3;
  const int32_t nb = (act + 15) / 16;
  if ((sel > 0 && nb >= sel) || (sel < 0 && nb < -sel)) return false;
  npar = min(npar, act);
  return true;
}

// M2M of one level: columns are the level's parents with children, in the
// order of their children (the t-th such parent has children at
// first_child + 8 t); block = 16 parents, wave rb = row block rb.
__global__ __launch_bounds__(7 * 64) void k_m2m_mfma(Cells c, int32_t first_child, int32_t npar,
                                                     const double *Fm, double *M,
                                                     const int32_t *lev, int dl, int sel) {
  critical_path_priority();
  if (!level_extent(lev, dl, sel, first_child, npar)) return;
  if ((int32_t)blockIdx.x * 16 >= npar) return;
  const int lane = threadIdx.x & 63, rb = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  __shared__ double sig[16];
  __shared__ double cf[16][NV + 2];  // sigma_p^-n of the B rows (child coefficients)
  __shared__ double rf[16][NV + 2];  // sigma_p^j of the output rows
  if (threadIdx.x < 16) {
    const int32_t tt = (int32_t)blockIdx.x * 16 + threadIdx.x;
    const int32_t cc = first_child + 8 * (tt < npar ? tt : 0);
    sig[threadIdx.x] = c.width[c.parent[cc]];
  }
  __syncthreads();
  parent_powers(cf, sig, -1);
  parent_powers(rf, sig, +1);
  __syncthreads();
  const int32_t t = (int32_t)blockIdx.x * 16 + col;
  const bool valid = t < npar;
  const int32_t c0 = first_child + 8 * (valid ? t : 0);
  const double *B = M + (int64_t)c0 * NV;
  const double *A = Fm + (int64_t)rb * KS * 8 * 64 + lane;
  int kend = NV;
  switch (rb) {  // constants per wave
    case 0: kend = m2m_kend(0); break;
    case 1: kend = m2m_kend(1); break;
    case 2: kend = m2m_kend(2); break;
    case 3: kend = m2m_kend(3); break;
    case 4: kend = m2m_kend(4); break;
    case 5: kend = m2m_kend(5); break;
    default: kend = m2m_kend(6); break;
  }
  // Two operand sets in flight: the loads of k-step k0 + 4 are issued before
  // step k0 is multiplied, into the other register set (no register moves,
  // which would wait for the loads).  Loads are unconditional from clamped
  // indices and masked afterwards: loads under a branch make every k-step
  // wait for all loads in flight.  On the small upper levels a launch is one
  // block, and its time is the load latency of its k-steps.
  d4 acc = {0, 0, 0, 0};
  double a0[8], b0[8], a1[8], b1[8];
  auto load = [&](int k0, double (&ao)[8], double (&bo)[8]) {
    const int kc = k0 + kq;
    const bool in = kc < kend;
    const int kk = in ? kc : 0;
    const int ks = min(k0 >> 2, KS - 1);
    // the mask as a factor (the operands are finite), so no load sits under a branch
    const double f = in ? cf[col][kk] : 0.0;
#pragma unroll
    for (int o = 0; o < 8; o++) {
      ao[o] = A[(ks * 8 + o) * 64];
      bo[o] = B[o * NV + kk] * f;
    }
  };
  auto mma = [&](const double (&ao)[8], const double (&bo)[8]) {
#pragma unroll
    for (int o = 0; o < 8; o++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ao[o], bo[o], acc, 0, 0, 0);
  };
  load(0, a0, b0);
  for (int k0 = 0; k0 < kend; k0 += 8) {
    load(k0 + 4, a1, b1);
    mma(a0, b0);
    load(k0 + 8, a0, b0);
    mma(a1, b1);
  }
  if (!valid) return;
  const int32_t p = c.parent[c0];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = rb * 16 + kq + 4 * i;
    if (r < NV) M[(int64_t)p * NV + r] += acc[i] * rf[col][r];
  }
}

// L2L of one level: the same parent columns; wave o = child octant o, all
// 7 row blocks of the child's coefficients at once (one B fragment per
// k-step feeds 7 independent accumulators).
__global__ __launch_bounds__(8 * 64) void k_l2l_mfma(Cells c, int32_t first_child, int32_t npar,
                                                     const double *Fl, double *L,
                                                     const int32_t *lev, int dl, int sel) {
  if (!level_extent(lev, dl, sel, first_child, npar)) return;
  if ((int32_t)blockIdx.x * 16 >= npar) return;
  const int lane = threadIdx.x & 63, oct = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  __shared__ double sig[16];
  __shared__ double cf[16][NV + 2];  // sigma_p^n of the B rows (parent coefficients)
  __shared__ double rf[16][NV + 2];  // sigma_p^-j of the output rows
  if (threadIdx.x < 16) {
    const int32_t tt = (int32_t)blockIdx.x * 16 + threadIdx.x;
    const int32_t cc = first_child + 8 * (tt < npar ? tt : 0);
    sig[threadIdx.x] = c.width[c.parent[cc]];
  }
  __syncthreads();
  parent_powers(cf, sig, +1);
  parent_powers(rf, sig, -1);
  __syncthreads();
  const int32_t t = (int32_t)blockIdx.x * 16 + col;
  const bool valid = t < npar;
  const int32_t c0 = first_child + 8 * (valid ? t : 0);
  const int32_t p = c.parent[c0];
  const double *B = L + (int64_t)p * NV;
  const double *S = Fl + (int64_t)oct * 7 * KS * 64 + lane;
  d4 acc[7];
#pragma unroll
  for (int rb = 0; rb < 7; rb++) acc[rb] = (d4){0, 0, 0, 0};
  // two operand sets in flight, unconditional clamped loads (as in M2M)
  double a0[7], a1[7], b0, b1;
  auto load = [&](int k0, double (&ao)[7], double &bo) {
    const int kc = k0 + kq;
    const bool in = kc < NV;
    const int kk = in ? kc : 0;
    const int ks = min(k0 >> 2, KS - 1);
    bo = B[kk] * (in ? cf[col][kk] : 0.0);
#pragma unroll
    for (int rb = 0; rb < 7; rb++) ao[rb] = S[(rb * KS + ks) * 64];
  };
  auto mma = [&](int k0, const double (&ao)[7], double bo) {
#pragma unroll
    for (int rb = 0; rb < 7; rb++)
      if (k0 >= l2l_kbeg(rb)) acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ao[rb], bo, acc[rb], 0, 0, 0);
  };
  load(0, a0, b0);
  for (int k0 = 0; k0 < NV; k0 += 8) {
    load(k0 + 4, a1, b1);
    mma(k0, a0, b0);
    load(k0 + 8, a0, b0);
    mma(k0 + 4, a1, b1);
  }
  if (!valid) return;
  const int64_t cidx = c0 + oct;
#pragma unroll
  for (int rb = 0; rb < 7; rb++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = rb * 16 + kq + 4 * i;
      if (r < NV) L[cidx * NV + r] += acc[rb][i] * rf[col][r];
    }
  }
}

// Levels with few parents: one block per (16 parents, row block) and one
// wave per two octants (M2M) or per quarter of the k-steps (L2L), every
// wave's operands of an octant loaded at once; the partial products meet in
// LDS.  A single wave walking all 28 k-steps of a row block waits out one
// load latency per step, which on the upper levels (a handful of blocks) is
// the whole launch.  The M2M block is 4 waves (one per SIMD) so that it
// fits beside the resident P2P's waves (184 VGPRs a wave: an 8-wave block,
// two waves a SIMD, waited for the P2P to end).
constexpr int M2M_SMALL_WAVES = 4;

__global__ __launch_bounds__(M2M_SMALL_WAVES * 64) void k_m2m_small(
    Cells c, int32_t first_child, int32_t npar, const double *Fm, double *M, const int32_t *lev,
    int dl, int sel) {
  critical_path_priority();
  if (!level_extent(lev, dl, sel, first_child, npar)) return;
  if ((int32_t)blockIdx.x * 16 >= npar) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, rb = blockIdx.y;
  const int col = lane & 15, kq = lane >> 4;
  __shared__ double sig[16];
  __shared__ double cf[16][NV + 2];
  __shared__ double rf[16][NV + 2];
  __shared__ double red[8][64][4];
  if (threadIdx.x < 16) {
    const int32_t tt = (int32_t)blockIdx.x * 16 + threadIdx.x;
    const int32_t cc = first_child + 8 * (tt < npar ? tt : 0);
    sig[threadIdx.x] = c.width[c.parent[cc]];
  }
  __syncthreads();
  parent_powers(cf, sig, -1);
  parent_powers(rf, sig, +1);
  __syncthreads();
  const int32_t t = (int32_t)blockIdx.x * 16 + col;
  const bool valid = t < npar;
  const int32_t c0 = first_child + 8 * (valid ? t : 0);
  const int kend = m2m_kend(rb);
  // per octant, the k-steps in two batches of KH, each batch's operands
  // loaded at once (one batch of every operand would not fit beside the P2P)
  constexpr int KH = KS / 2;
  static_assert(2 * KH == KS, "two equal batches of k-steps");
#pragma unroll 1
  for (int o = w; o < 8; o += M2M_SMALL_WAVES) {
    const double *B = M + (int64_t)c0 * NV + o * NV;
    const double *A = Fm + (int64_t)rb * KS * 8 * 64 + o * 64 + lane;
    d4 acc = {0, 0, 0, 0};
#pragma unroll 1
    for (int k0 = 0; k0 < KS; k0 += KH) {
      double a[KH], b[KH];
#pragma unroll
      for (int j = 0; j < KH; j++) {
        const int kc = 4 * (k0 + j) + kq;
        const bool in = kc < kend;
        const int kk = in ? kc : 0;
        a[j] = A[(k0 + j) * 8 * 64];
        b[j] = B[kk] * (in ? cf[col][kk] : 0.0);
      }
#pragma unroll
      for (int j = 0; j < KH; j++)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[j], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) red[o][lane][i] = acc[i];
  }
  __syncthreads();
  if (w != 0 || !valid) return;
  const int32_t p = c.parent[c0];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double v = 0;
#pragma unroll
    for (int w = 0; w < 8; w++) v += red[w][lane][i];
    const int r = rb * 16 + kq + 4 * i;
    if (r < NV) M[(int64_t)p * NV + r] += v * rf[col][r];
  }
}

constexpr int L2L_SPLIT = 4;                         // waves per (parents, octant, row block)
constexpr int L2L_KPW = (KS + L2L_SPLIT - 1) / L2L_SPLIT;  // k-steps per wave

__global__ __launch_bounds__(L2L_SPLIT * 64) void k_l2l_small(Cells c, int32_t first_child,
                                                              int32_t npar, const double *Fl,
                                                              double *L, const int32_t *lev,
                                                              int dl, int sel) {
  if (!level_extent(lev, dl, sel, first_child, npar)) return;
  if ((int32_t)blockIdx.x * 16 >= npar) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int oct = blockIdx.y, rb = blockIdx.z;
  const int col = lane & 15, kq = lane >> 4;
  __shared__ double sig[16];
  __shared__ double cf[16][NV + 2];
  __shared__ double rf[16][NV + 2];
  __shared__ double red[L2L_SPLIT][64][4];
  if (threadIdx.x < 16) {
    const int32_t tt = (int32_t)blockIdx.x * 16 + threadIdx.x;
    const int32_t cc = first_child + 8 * (tt < npar ? tt : 0);
    sig[threadIdx.x] = c.width[c.parent[cc]];
  }
  __syncthreads();
  parent_powers(cf, sig, +1);
  parent_powers(rf, sig, -1);
  __syncthreads();
  const int32_t t = (int32_t)blockIdx.x * 16 + col;
  const bool valid = t < npar;
  const int32_t c0 = first_child + 8 * (valid ? t : 0);
  const int32_t p = c.parent[c0];
  const double *B = L + (int64_t)p * NV;
  const double *S = Fl + ((int64_t)oct * 7 + rb) * KS * 64 + lane;
  const int kbeg = l2l_kbeg(rb);
  double a[L2L_KPW], b[L2L_KPW];
#pragma unroll
  for (int j = 0; j < L2L_KPW; j++) {
    const int ks = w * L2L_KPW + j;
    const int kc = 4 * ks + kq;
    const bool in = ks < KS && kc < NV && 4 * ks >= kbeg;
    const int kk = in ? kc : 0;
    a[j] = S[(ks < KS ? ks : 0) * 64];
    b[j] = B[kk] * (in ? cf[col][kk] : 0.0);
  }
  d4 acc = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < L2L_KPW; j++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[j], acc, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; i++) red[w][lane][i] = acc[i];
  __syncthreads();
  if (w != 0 || !valid) return;
  const int64_t cidx = c0 + oct;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double v = 0;
#pragma unroll
    for (int q = 0; q < L2L_SPLIT; q++) v += red[q][lane][i];
    const int r = rb * 16 + kq + 4 * i;
    if (r < NV) L[cidx * NV + r] += v * rf[col][r];
  }
}

// levels below this many 16-parent blocks take the split kernels (measured
// crossover: the split L2L does 7 x 8 blocks per 16 parents)
constexpr int SMALL_BLOCKS_M2M = 128;
constexpr int SMALL_BLOCKS_L2L = 24;

}  // namespace

size_t translation_ops_doubles() {
  return (size_t)NV * KM + (size_t)8 * NV * NV + FRAG_M + FRAG_L;
}

namespace {

// M2L output coefficient (j, k) from source coefficients Ms and the singular
// harmonics Ys of dX (LaplaceSphericalCPU_tapas.h:227-250, mutual = 0)
__device__ inline cd m2l_entry(int j, int k, const cd *Ms, const cd *Ys) {
  const double Cnm = sgn(j);
  cd acc = mk(0, 0);
  for (int n = 0; n < P - j; n++) {
    for (int m = -n; m < 0; m++)
      acc = cadd(acc, cmul(cscale(cconj(Ms[tri(n, -m)]), Cnm), yfull(Ys, j + n, m - k)));
    for (int m = 0; m <= n; m++) {
      const double Cnm2 = Cnm * sgn((k - m) * (k < m) + m);
      acc = cadd(acc, cmul(cscale(Ms[tri(n, m)], Cnm2), yfull(Ys, j + n, m - k)));
    }
  }
  return acc;
}

// image (ix, iy, iz) in {-1, 0, 1}^3 of index a in the reference's loop order
__device__ inline void image_of(int a, int &ix, int &iy, int &iz) {
  ix = a / 9 - 1;
  iy = a / 3 % 3 - 1;
  iz = a % 3 - 1;
}

// The far field of the periodic images beyond the first shell
// (EX/include/traversal.h:126-185, images > 1): at sublevel l the copies of
// the periodic cell form cells of side cycle_l = cycle 3^l whose multipole
// (about the root's centre) is M_l, M_0 the root's.  The 26 x 27 copies
// (3 i + c) cycle_l, i in {-1,0,1}^3 without 0, c in {-1,0,1}^3, add their M2L
// into the root's local expansion (dX = -Xperiodic: the root with an image of
// itself), and M_{l+1} is the M2M of the 27 copies i cycle_l, i in {-1,0,1}^3,
// into the cell three times larger.  (The copy at the centre is part of that
// cell: the reference's loop resets the parent before its M2M, which would
// leave it out; this follows the sum over all (3^images)^3 copies that its
// direct check computes.)  One workgroup: 4 waves share the M2L offsets, each
// in a fixed order, their partial sums added in wave order.
__global__ __launch_bounds__(256) void k_periodic_far(const double *Mroot, double *Lroot,
                                                      double cycle, int images) {
  critical_path_priority();
  __shared__ cd Ms[NTERM], Mn[NTERM], Yl[NTERM];
  __shared__ cd Ys[4][NTERM];
  __shared__ cd part[4][NTERM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int j = 0, k = 0;
  if (lane < NTERM) untri(lane, j, k);
  const cd *M0 = reinterpret_cast<const cd *>(Mroot);
  if (threadIdx.x < NTERM) Ms[threadIdx.x] = M0[threadIdx.x];
  cd acc = mk(0, 0);
  for (int level = 0; level + 1 < images; level++) {
    __syncthreads();
    for (int o = w; o < 26 * 27; o += 4) {
      const int a = o / 27, ac = a < 13 ? a : a + 1;  // the 26 images other than the centre
      int ix, iy, iz, cx, cy, cz;
      image_of(ac, ix, iy, iz);
      image_of(o % 27, cx, cy, cz);
      __builtin_amdgcn_wave_barrier();
      if (lane < P) {
        const double dx = -((ix * 3 + cx) * cycle), dy = -((iy * 3 + cy) * cycle),
                     dz = -((iz * 3 + cz) * cycle);
        double rho, ca, sa;
        cd ei;
        sph_frame(dx, dy, dz, rho, ca, sa, ei);
        local_column(lane, rho, ca, sa, ei, [&](int n, int m, cd Y) { Ys[w][tri(n, m)] = Y; });
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      if (lane < NTERM) acc = cadd(acc, m2l_entry(j, k, Ms, Ys[w]));
    }
    __syncthreads();
    if (w == 0) {
      cd mn = mk(0, 0);
      for (int a = 0; a < 27; a++) {
        int ix, iy, iz;
        image_of(a, ix, iy, iz);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {  // dX = parent - child centre = -(i cycle_l)
          double rho, ca, sa;
          cd ei;
          sph_frame(-(ix * cycle), -(iy * cycle), -(iz * cycle), rho, ca, sa, ei);
          multipole_all<false>(rho, ca, sa, ei, [&](int n, int m, cd v, cd) { Yl[tri(n, m)] = v; });
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (lane < NTERM) mn = cadd(mn, m2m_entry(j, k, Ms, Yl));
      }
      if (lane < NTERM) Mn[lane] = mn;
    }
    __syncthreads();
    if (threadIdx.x < NTERM) Ms[threadIdx.x] = Mn[threadIdx.x];
    cycle *= 3;
  }
  if (lane < NTERM) part[w][lane] = acc;
  __syncthreads();
  if (threadIdx.x < NTERM) {
    cd s = part[0][threadIdx.x];
    for (int u = 1; u < 4; u++) s = cadd(s, part[u][threadIdx.x]);
    cd *L0 = reinterpret_cast<cd *>(Lroot);
    L0[threadIdx.x] = cadd(L0[threadIdx.x], s);
  }
}

}  // namespace

hipError_t launch_periodic_far(const double *Mroot, double *Lroot, double cycle, int images,
                               hipStream_t s) {
  if (images < 2) return hipSuccess;
  k_periodic_far<<<1, 256, 0, s>>>(Mroot, Lroot, cycle, images);
  return hipGetLastError();
}

// The unit operators (built once per context); ops: translation_ops_doubles()
// doubles: the M2M operator, the 8 L2L operators, then both in fragment order.
hipError_t launch_translation_ops(double *ops, hipStream_t s) {
  double *Tm = ops, *Tl = ops + (size_t)NV * KM, *Fm = Tl + (size_t)8 * NV * NV, *Fl = Fm + FRAG_M;
  k_build_ops<<<dim3(NV, 8, 2), 64, 0, s>>>(Tm, Tl);
  k_frag_ops<<<(FRAG_M + FRAG_L + 255) / 256, 256, 0, s>>>(Tm, Tl, Fm, Fl);
  return hipGetLastError();
}

// parents of one level: those whose children are the cells
// [first_child, first_child + 8 npar) of the next level
hipError_t launch_m2m(Cells c, int32_t first_child, int32_t npar, const double *ops, double *M,
                      hipStream_t s, const int32_t *lev, int dl) {
  if (npar <= 0) return hipSuccess;
  const double *Fm = ops + (size_t)NV * KM + (size_t)8 * NV * NV;
  const int nb = (npar + 15) / 16;
  if (lev) {  // the level's size on the device picks the variant
    k_m2m_small<<<dim3(std::min(nb, SMALL_BLOCKS_M2M - 1), 7), M2M_SMALL_WAVES * 64, 0, s>>>(
        c, first_child, npar, Fm, M, lev, dl, SMALL_BLOCKS_M2M);
    if (nb >= SMALL_BLOCKS_M2M)
      k_m2m_mfma<<<nb, 7 * 64, 0, s>>>(c, first_child, npar, Fm, M, lev, dl, -SMALL_BLOCKS_M2M);
  } else if (nb < SMALL_BLOCKS_M2M) {
    k_m2m_small<<<dim3(nb, 7), M2M_SMALL_WAVES * 64, 0, s>>>(c, first_child, npar, Fm, M, lev,
                                                             dl, 0);
  } else {
    k_m2m_mfma<<<nb, 7 * 64, 0, s>>>(c, first_child, npar, Fm, M, lev, dl, 0);
  }
  return hipGetLastError();
}

hipError_t launch_l2l(Cells c, int32_t first_child, int32_t npar, const double *ops, double *L,
                      hipStream_t s, const int32_t *lev, int dl) {
  if (npar <= 0) return hipSuccess;
  const double *Fl = ops + (size_t)NV * KM + (size_t)8 * NV * NV + FRAG_M;
  const int nb = (npar + 15) / 16;
  if (lev) {  // the level's size on the device picks the variant
    k_l2l_small<<<dim3(std::min(nb, SMALL_BLOCKS_L2L - 1), 8, 7), L2L_SPLIT * 64, 0, s>>>(
        c, first_child, npar, Fl, L, lev, dl, SMALL_BLOCKS_L2L);
    if (nb >= SMALL_BLOCKS_L2L)
      k_l2l_mfma<<<nb, 8 * 64, 0, s>>>(c, first_child, npar, Fl, L, lev, dl, -SMALL_BLOCKS_L2L);
  } else if (nb < SMALL_BLOCKS_L2L) {
    k_l2l_small<<<dim3(nb, 8, 7), L2L_SPLIT * 64, 0, s>>>(c, first_child, npar, Fl, L, lev, dl,
                                                          0);
  } else {
    k_l2l_mfma<<<nb, 8 * 64, 0, s>>>(c, first_child, npar, Fl, L, lev, dl, 0);
  }
  return hipGetLastError();
}

}  // namespace tfmm
