// expansions.hip -- FP64 upward (P2M, M2M) and downward (L2L, L2P) kernels.
//
// Reference (sample/exafmm-dev-13274dd4ac68/examples/):
//   FMM_Upward / FMM_Downward   tapas_exafmm.cxx:87-140
//   P2M  LaplaceSphericalCPU_tapas.h:138-163     M2M  :165-206
//   L2L  LaplaceSphericalCPU_tapas.h:309-345     L2P  :275-307
// The reference recurses over cells on the host; here every level of the
// breadth-first cell array is one launch (children before parents upward,
// parents before children downward), one wave64 per cell.
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

__device__ inline int sgn(int n) { return (n & 1) ? -1 : 1; }
__device__ inline int ipow2n(int n) { return (n >= 0) ? 1 : sgn(n); }

// (j, k) of coefficient index t = tri(j, k)
__device__ inline void untri(int t, int &j, int &k) {
  j = 0;
  while (tri(j + 1, 0) <= t) j++;
  k = t - tri(j, 0);
}

// P2M: M[n,m] += q * conj(Y_n^m(body - centre)), summed over the leaf's bodies.
__global__ __launch_bounds__(64) void k_p2m(Cells c, Bodies b, double *M) {
  const int32_t ci = blockIdx.x;
  if (c.child[ci] >= 0 || c.nb[ci] == 0) return;
  const int lane = threadIdx.x;
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  double are = 0, aim = 0;  // lane t < 55 owns M[t]
  for (int32_t base = beg; base < end; base += 64) {
    const int32_t i = base + lane;
    const bool valid = i < end;
    double dx = 1, dy = 0, dz = 0, q = 0;
    if (valid) {
      dx = b.x[i] - cx;
      dy = b.y[i] - cy;
      dz = b.z[i] - cz;
      q = b.q[i];
    }
    double rho, ca, sa;
    cd ei;
    sph_frame(dx, dy, dz, rho, ca, sa, ei);
    multipole_all<false>(rho, ca, sa, ei, [&](int n, int m, cd Y, cd) {
      const double re = wave_sum(q * Y.re);
      const double im = wave_sum(-q * Y.im);
      if (lane == tri(n, m)) {
        are += re;
        aim += im;
      }
    });
  }
  if (lane < NTERM) {
    M[2 * (NTERM * (int64_t)ci + lane)] += are;
    M[2 * (NTERM * (int64_t)ci + lane) + 1] += aim;
  }
}

// M2M for the internal cells of one level: parent M += shift(child M) for the
// 8 children in order; dX = parent centre - child centre.
__global__ __launch_bounds__(64) void k_m2m(Cells c, int32_t lbeg, double *M) {
  const int32_t p = lbeg + blockIdx.x;
  const int32_t first = c.child[p];
  if (first < 0) return;
  __shared__ cd Ys[8][NTERM];
  const int lane = threadIdx.x;
  if (lane < 8) {
    const int32_t ch = first + lane;
    const double dx = c.center[3 * (int64_t)p] - c.center[3 * (int64_t)ch];
    const double dy = c.center[3 * (int64_t)p + 1] - c.center[3 * (int64_t)ch + 1];
    const double dz = c.center[3 * (int64_t)p + 2] - c.center[3 * (int64_t)ch + 2];
    double rho, ca, sa;
    cd ei;
    sph_frame(dx, dy, dz, rho, ca, sa, ei);
    multipole_all<false>(rho, ca, sa, ei,
                         [&](int n, int m, cd Y, cd) { Ys[lane][tri(n, m)] = Y; });
  }
  __syncthreads();
  if (lane >= NTERM) return;
  int j, k;
  untri(lane, j, k);
  const cd *Mall = reinterpret_cast<const cd *>(M);
  cd out = mk(0, 0);
  for (int chi = 0; chi < 8; chi++) {
    const cd *Mc = Mall + (int64_t)(first + chi) * NTERM;
    const cd *Y = Ys[chi];
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
    out = cadd(out, acc);
  }
  cd *Mp = reinterpret_cast<cd *>(M) + (int64_t)p * NTERM;
  Mp[lane] = cadd(Mp[lane], out);
}

// L2L for the children of the internal cells of one (parent) level:
// child L += shift(parent L); dX = child centre - parent centre.
__global__ __launch_bounds__(64) void k_l2l(Cells c, int32_t lbeg, double *L) {
  const int32_t p = lbeg + blockIdx.x;
  const int32_t first = c.child[p];
  if (first < 0) return;
  __shared__ cd Ys[8][NTERM];
  __shared__ cd Lp[NTERM];
  const int lane = threadIdx.x;
  cd *Lall = reinterpret_cast<cd *>(L);
  if (lane < NTERM) Lp[lane] = Lall[(int64_t)p * NTERM + lane];
  if (lane < 8) {
    const int32_t ch = first + lane;
    const double dx = c.center[3 * (int64_t)ch] - c.center[3 * (int64_t)p];
    const double dy = c.center[3 * (int64_t)ch + 1] - c.center[3 * (int64_t)p + 1];
    const double dz = c.center[3 * (int64_t)ch + 2] - c.center[3 * (int64_t)p + 2];
    double rho, ca, sa;
    cd ei;
    sph_frame(dx, dy, dz, rho, ca, sa, ei);
    multipole_all<false>(rho, ca, sa, ei,
                         [&](int n, int m, cd Y, cd) { Ys[lane][tri(n, m)] = Y; });
  }
  __syncthreads();
  if (lane >= NTERM) return;
  int j, k;
  untri(lane, j, k);
  for (int chi = 0; chi < 8; chi++) {
    const cd *Y = Ys[chi];
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
    cd *Lc = Lall + (int64_t)(first + chi) * NTERM;
    Lc[lane] = cadd(Lc[lane], acc);
  }
}

// L2P: trg = p2p / q + local field of the leaf's expansion.
__global__ __launch_bounds__(64) void k_l2p(Cells c, Bodies b, const double *L,
                                            const double *p2p, double *trg) {
  const int32_t ci = blockIdx.x;
  if (c.child[ci] >= 0 || c.nb[ci] == 0) return;
  __shared__ cd Ls[NTERM];
  const int lane = threadIdx.x;
  if (lane < NTERM) Ls[lane] = reinterpret_cast<const cd *>(L)[(int64_t)ci * NTERM + lane];
  __syncthreads();
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  for (int32_t i = beg + lane; i < end; i += 64) {
    const double dx = b.x[i] - cx, dy = b.y[i] - cy, dz = b.z[i] - cz;
    const double q = b.q[i];
    double r, ct, st;
    cd ei;
    sph_frame(dx, dy, dz, r, ct, st, ei);
    double t0 = p2p[4 * (int64_t)i] / q, t1 = p2p[4 * (int64_t)i + 1] / q,
           t2 = p2p[4 * (int64_t)i + 2] / q, t3 = p2p[4 * (int64_t)i + 3] / q;
    double s0 = 0, s1 = 0, s2 = 0;
    multipole_all<true>(r, ct, st, ei, [&](int n, int m, cd Y, cd YT) {
      const cd l = Ls[tri(n, m)];
      const cd ly = cmul(l, Y);
      const double lyt = cmul(l, YT).re;
      if (m == 0) {
        t0 += ly.re;
        s0 += ly.re / r * n;
        s1 += lyt;
      } else {
        t0 += 2 * ly.re;
        s0 += 2 * ly.re / r * n;
        s1 += 2 * lyt;
        s2 += 2 * (-ly.im) * m;
      }
    });
    const double sp = ei.im, cp = ei.re;
    t1 += st * cp * s0 + ct * cp / r * s1 - sp / r / st * s2;
    t2 += st * sp * s0 + ct * sp / r * s1 + cp / r / st * s2;
    t3 += ct * s0 - st / r * s1;
    trg[4 * (int64_t)i] = t0;
    trg[4 * (int64_t)i + 1] = t1;
    trg[4 * (int64_t)i + 2] = t2;
    trg[4 * (int64_t)i + 3] = t3;
  }
}

__global__ void k_unpermute(const int32_t *perm, const double *src, int64_t n, double *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = perm[i];
  for (int d = 0; d < 4; d++) dst[4 * o + d] = src[4 * i + d];
}

}  // namespace

hipError_t launch_p2m(Cells c, int32_t ncells, Bodies b, double *M, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_p2m<<<ncells, 64, 0, s>>>(c, b, M);
  return hipGetLastError();
}

hipError_t launch_m2m(Cells c, int32_t lbeg, int32_t lcnt, double *M, hipStream_t s) {
  if (lcnt <= 0) return hipSuccess;
  k_m2m<<<lcnt, 64, 0, s>>>(c, lbeg, M);
  return hipGetLastError();
}

hipError_t launch_l2l(Cells c, int32_t lbeg, int32_t lcnt, double *L, hipStream_t s) {
  if (lcnt <= 0) return hipSuccess;
  k_l2l<<<lcnt, 64, 0, s>>>(c, lbeg, L);
  return hipGetLastError();
}

hipError_t launch_l2p(Cells c, int32_t ncells, Bodies b, const double *L, const double *p2p,
                      double *trg, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_l2p<<<ncells, 64, 0, s>>>(c, b, L, p2p, trg);
  return hipGetLastError();
}

hipError_t launch_unpermute(const int32_t *perm, const double *trg_sorted, int64_t n,
                            double *trg_orig, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_unpermute<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(perm, trg_sorted, n, trg_orig);
  return hipGetLastError();
}

}  // namespace tfmm
