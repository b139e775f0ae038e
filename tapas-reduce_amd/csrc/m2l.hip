// m2l.hip -- FP64 M2L over a per-target CSR list.
//
// Reference: M2L (sample/exafmm-dev-13274dd4ac68/examples/
// LaplaceSphericalCPU_tapas.h:208-273) with evalLocal (:102-135), mutual = 0:
//   L_i[j,k] += sum_{n < P-j} sum_{|m| <= n} s(j,k,n,m) M_j[n,m] Y_{j+n}^{m-k}(dX)
// dX = centre_i - centre_j.  One wave64 per target cell owns the target's L
// for the whole source list (no atomics); per source, lanes 0..9 build one
// column of the singular harmonics each into LDS, then lane t < 55 forms
// output coefficient t.
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

__device__ inline int sgn(int n) { return (n & 1) ? -1 : 1; }

__global__ __launch_bounds__(64) void k_m2l_fp64(Cells c, Csr list, const double *M,
                                                 double *L) {
  const int32_t ti = blockIdx.x;
  const int32_t b = list.off[ti], e = list.off[ti + 1];
  if (b == e) return;
  __shared__ cd Ys[NTERM];
  __shared__ cd Ms[NTERM];
  const int lane = threadIdx.x;
  int j = 0, k = 0;
  if (lane < NTERM) {
    while (tri(j + 1, 0) <= lane) j++;
    k = lane - tri(j, 0);
  }
  const double cx = c.center[3 * (int64_t)ti], cy = c.center[3 * (int64_t)ti + 1],
               cz = c.center[3 * (int64_t)ti + 2];
  const cd *Mall = reinterpret_cast<const cd *>(M);
  const double Cnm = sgn(j);
  cd acc = mk(0, 0);
  for (int32_t s = b; s < e; s++) {
    const int32_t sj = list.src[s];
    if (lane < NTERM) Ms[lane] = Mall[(int64_t)sj * NTERM + lane];
    if (lane < P) {
      const double dx = cx - c.center[3 * (int64_t)sj];
      const double dy = cy - c.center[3 * (int64_t)sj + 1];
      const double dz = cz - c.center[3 * (int64_t)sj + 2];
      double rho, ca, sa;
      cd ei;
      sph_frame(dx, dy, dz, rho, ca, sa, ei);
      local_column(lane, rho, ca, sa, ei,
                   [&](int n, int m, cd Y) { Ys[tri(n, m)] = Y; });
    }
    __syncthreads();
    if (lane < NTERM) {
      for (int n = 0; n < P - j; n++) {
        for (int m = -n; m < 0; m++) {
          const cd t = cmul(cscale(cconj(Ms[tri(n, -m)]), Cnm), yfull(Ys, j + n, m - k));
          acc = cadd(acc, t);
        }
        for (int m = 0; m <= n; m++) {
          const double Cnm2 = Cnm * sgn((k - m) * (k < m) + m);
          const cd t = cmul(cscale(Ms[tri(n, m)], Cnm2), yfull(Ys, j + n, m - k));
          acc = cadd(acc, t);
        }
      }
    }
    __syncthreads();
  }
  if (lane < NTERM) {
    cd *Li = reinterpret_cast<cd *>(L) + (int64_t)ti * NTERM;
    Li[lane] = cadd(Li[lane], acc);
  }
}

}  // namespace

hipError_t launch_m2l_fp64(Cells c, int32_t ncells, Csr list, const double *M, double *L,
                           hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_m2l_fp64<<<ncells, 64, 0, s>>>(c, list, M, L);
  return hipGetLastError();
}

}  // namespace tfmm
