// m2l.hip -- FP64 M2L of individual pairs (those outside the batched
// groups: level difference >= 2).
//
// Reference: M2L (sample/exafmm-dev-13274dd4ac68/examples/
// LaplaceSphericalCPU_tapas.h:208-273) with evalLocal (:102-135), mutual = 0:
//   L_i[j,k] += sum_{n < P-j} sum_{|m| <= n} s(j,k,n,m) M_j[n,m] Y_{j+n}^{m-k}(dX)
// dX = centre_i - centre_j.  One wave64 per pair: lanes 0..9 build one column
// of the singular harmonics each into LDS, then lane t < 55 forms output
// coefficient t into a per-pair partial.  A second kernel adds the partials
// of each target in list order (deterministic, no atomics).
#include <algorithm>

#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

__device__ inline int sgn(int n) { return (n & 1) ? -1 : 1; }

constexpr int PAIR_WAVES = 4;

__global__ __launch_bounds__(64 * PAIR_WAVES) void k_m2l_pairs_fp64(Cells c, Csr list,
                                                                   int32_t p0, int32_t npairs,
                                                                   const int32_t *tgt,
                                                                   const double *M, cd *part,
                                                                   const uint32_t *dn,
                                                                   double cycle) {
  critical_path_priority();
  __shared__ cd Ys[PAIR_WAVES][NTERM];
  __shared__ cd Ms[PAIR_WAVES][NTERM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t pl = blockIdx.x * PAIR_WAVES + w;  // pair within the chunk
  if (dn) npairs = min(npairs, (int32_t)min(*dn, (uint32_t)INT32_MAX) - p0);
  if (pl >= npairs) return;  // whole wave leaves together; no block barrier follows
  const int32_t pi = p0 + pl;
  const int32_t ti = tgt[pi];
  const uint32_t sw = (uint32_t)list.src[pi];  // source cell and periodic image
  const int32_t sj = (int32_t)(sw & CELL_MASK);
  int j = 0, k = 0;
  if (lane < NTERM) {
    while (tri(j + 1, 0) <= lane) j++;
    k = lane - tri(j, 0);
  }
  const cd *Mall = reinterpret_cast<const cd *>(M);
  if (lane < NTERM) Ms[w][lane] = Mall[(int64_t)sj * NTERM + lane];
  if (lane < P) {
    double sx, sy, sz;
    image_shift(sw >> IMG_SHIFT, cycle, sx, sy, sz);
    // dX = Xi - Xj - Xperiodic (LaplaceSphericalCPU_tapas.h:218-220)
    const double dx = c.center[3 * (int64_t)ti] - c.center[3 * (int64_t)sj] - sx;
    const double dy = c.center[3 * (int64_t)ti + 1] - c.center[3 * (int64_t)sj + 1] - sy;
    const double dz = c.center[3 * (int64_t)ti + 2] - c.center[3 * (int64_t)sj + 2] - sz;
    double rho, ca, sa;
    cd ei;
    sph_frame(dx, dy, dz, rho, ca, sa, ei);
    local_column(lane, rho, ca, sa, ei, [&](int n, int m, cd Y) { Ys[w][tri(n, m)] = Y; });
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  if (lane >= NTERM) return;
  const double Cnm = sgn(j);
  cd acc = mk(0, 0);
  for (int n = 0; n < P - j; n++) {
    for (int m = -n; m < 0; m++) {
      const cd t = cmul(cscale(cconj(Ms[w][tri(n, -m)]), Cnm), yfull(Ys[w], j + n, m - k));
      acc = cadd(acc, t);
    }
    for (int m = 0; m <= n; m++) {
      const double Cnm2 = Cnm * sgn((k - m) * (k < m) + m);
      const cd t = cmul(cscale(Ms[w][tri(n, m)], Cnm2), yfull(Ys[w], j + n, m - k));
      acc = cadd(acc, t);
    }
  }
  part[(int64_t)pl * NTERM + lane] = acc;
}

// target of every list entry (of the first npairs)
__global__ void k_list_targets(const int32_t *off, int32_t ncells, int32_t npairs, int32_t *tgt) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) return;
  const int32_t e1 = min(off[t + 1], npairs);
  for (int32_t e = off[t]; e < e1; e++) tgt[e] = t;
}

// L_t += partials of t's pairs within the chunk [p0, p1), in list order; one
// wave per target
__global__ __launch_bounds__(256) void k_m2l_reduce(Csr list, int32_t ncells, int32_t p0,
                                                    int32_t p1, const cd *part, double *L) {
  critical_path_priority();
  const int lane = threadIdx.x & 63;
  const int32_t ti = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (ti >= ncells || lane >= NTERM) return;
  const int32_t b = max(list.off[ti], p0), e = min(list.off[ti + 1], p1);
  if (b >= e) return;
  cd acc = mk(0, 0);
  for (int32_t p = b; p < e; p++) acc = cadd(acc, part[(int64_t)(p - p0) * NTERM + lane]);
  cd *Li = reinterpret_cast<cd *>(L) + (int64_t)ti * NTERM;
  Li[lane] = cadd(Li[lane], acc);
}

}  // namespace

constexpr int32_t PAIR_CHUNK = 1 << 22;  // pairs per partials buffer (3.7 GB)

size_t m2l_pairs_scratch_bytes(int32_t npairs) {
  const size_t chunk = (size_t)std::min(npairs, PAIR_CHUNK);
  return chunk * NTERM * sizeof(cd) + (size_t)npairs * sizeof(int32_t);
}

// scratch: m2l_pairs_scratch_bytes(npairs) bytes (partials of one chunk of
// pairs, then the target of every pair); chunks are added in list order
hipError_t launch_m2l_fp64(Cells c, int32_t ncells, Csr list, int32_t npairs, const double *M,
                           double *L, void *scratch, hipStream_t s, const uint32_t *dn,
                           double cycle) {
  if (ncells <= 0 || npairs <= 0) return hipSuccess;
  const int32_t chunk = std::min(npairs, PAIR_CHUNK);
  cd *part = static_cast<cd *>(scratch);
  int32_t *tgt = reinterpret_cast<int32_t *>(part + (size_t)chunk * NTERM);
  k_list_targets<<<(ncells + 255) / 256, 256, 0, s>>>(list.off, ncells, npairs, tgt);
  for (int32_t p0 = 0; p0 < npairs; p0 += chunk) {
    const int32_t n = std::min(chunk, npairs - p0);
    k_m2l_pairs_fp64<<<(n + PAIR_WAVES - 1) / PAIR_WAVES, 64 * PAIR_WAVES, 0, s>>>(
        c, list, p0, n, tgt, M, part, dn, cycle);
    k_m2l_reduce<<<(ncells + 3) / 4, 256, 0, s>>>(list, ncells, p0, p0 + n, part, L);
  }
  return hipGetLastError();
}

}  // namespace tfmm
