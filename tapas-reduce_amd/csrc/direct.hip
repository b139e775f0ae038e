// direct.hip -- FP64 O(N * targets) direct sum, the error judge of the FMM.
//
// Same interaction as the P2P functor and the same normalisation as the
// reference's CheckResult (sample/exafmm-dev-13274dd4ac68/examples/
// tapas_exafmm.cxx:338-367): each target accumulates q_i q_j / r and
// -q_i q_j dX / r^3 over all bodies j (r = 0 skipped), then divides by q_i.
// With periodic images (prange > 0) every body j is also seen at
// x_j + (ix, iy, iz) cycle for |ix|, |iy|, |iz| <= prange, dX = x_i - x_j -
// Xperiodic, in the reference's loop order (tapas_exafmm.cxx:328-345).
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

constexpr int DT = 256;

__global__ __launch_bounds__(DT) void k_direct(const double *x, const double *y,
                                               const double *z, const double *q, int64_t n,
                                               const int64_t *targets, int64_t tbase,
                                               int64_t nt, double *out, int prange,
                                               double cycle) {
  __shared__ double sx[DT], sy[DT], sz[DT], sq[DT];
  const int64_t t = (int64_t)blockIdx.x * DT + threadIdx.x;
  const bool valid = t < nt;
  double xi = 0, yi = 0, zi = 0, qi = 1;
  if (valid) {
    const int64_t i = targets[t] - tbase;
    xi = x[i];
    yi = y[i];
    zi = z[i];
    qi = q[i];
  }
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int64_t base = 0; base < n; base += DT) {
    __syncthreads();
    const int64_t j = base + threadIdx.x;
    if (j < n) {
      sx[threadIdx.x] = x[j];
      sy[threadIdx.x] = y[j];
      sz[threadIdx.x] = z[j];
      sq[threadIdx.x] = q[j];
    }
    __syncthreads();
    const int cnt = (int)((n - base) < DT ? (n - base) : DT);
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          const double px = ix * cycle, py = iy * cycle, pz = iz * cycle;
          for (int jj = 0; jj < cnt; jj++) {
            const double dx = xi - sx[jj] - px, dy = yi - sy[jj] - py, dz = zi - sz[jj] - pz;
            const double R2 = dx * dx + dy * dy + dz * dz;
            if (R2 != 0) {
              const double invR2 = 1.0 / R2;
              const double invR = qi * sq[jj] * sqrt(invR2);
              const double f = invR2 * invR;
              a0 += invR;
              a1 -= dx * f;
              a2 -= dy * f;
              a3 -= dz * f;
            }
          }
        }
  }
  if (valid) {
    out[4 * t] = a0 / qi;
    out[4 * t + 1] = a1 / qi;
    out[4 * t + 2] = a2 / qi;
    out[4 * t + 3] = a3 / qi;
  }
}

}  // namespace

hipError_t launch_direct_fp64(const double *x, const double *y, const double *z,
                              const double *q, int64_t n, const int64_t *targets, int64_t tbase,
                              int64_t nt, double *out, hipStream_t s, int prange, double cycle) {
  if (nt <= 0) return hipSuccess;
  k_direct<<<(unsigned)((nt + DT - 1) / DT), DT, 0, s>>>(x, y, z, q, n, targets, tbase, nt,
                                                         out, prange, cycle);
  return hipGetLastError();
}

}  // namespace tfmm
