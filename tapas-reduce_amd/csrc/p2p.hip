// p2p.hip -- near-field P2P over a per-target-leaf CSR list of source leaves.
//
// Reference: P2P functor (sample/exafmm-dev-13274dd4ac68/examples/
// LaplaceP2PCPU_tapas.h:48-74) applied by ProductMapImpl over leaf body
// products (cpp/include/tapas/hot/mapper.h:126-169), mutual = 0:
//   dX = x_i - x_j; R2 = |dX|^2; if R2 != 0:
//     invR = q_i q_j / |dX|;  acc_i += (invR, -dX invR / R2)
// (the reference's CUDA path accumulates with an atomicCAS loop per term,
// TP/map.h:159-206; here one wave owns a target leaf for its whole source
// list and every target body is written once).
//
// Wave mapping: a target leaf of nt <= 64 bodies is spread over g = 64 / nt
// lanes per body (lane = body * g + sub); the g lanes of one body split the
// source bodies and their partial sums are added in lane order through LDS
// at the end, so small leaves still keep (almost) all 64 lanes busy.
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

constexpr int P2P_CHUNK = 64;

__global__ __launch_bounds__(64) void k_p2p_fp64(Cells c, Csr list, Bodies b, double *acc) {
  const int32_t ti = blockIdx.x;
  if (c.child[ti] >= 0) return;
  const int32_t nt = c.nb[ti];
  if (nt == 0) return;
  const int32_t tb = c.beg[ti];
  const int lane = threadIdx.x;
  __shared__ double sx[P2P_CHUNK], sy[P2P_CHUNK], sz[P2P_CHUNK], sq[P2P_CHUNK];
  const int32_t lb = list.off[ti], le = list.off[ti + 1];
  __shared__ double red[4][64];
  // lanes per target body
  const int g = nt >= 64 ? 1 : 64 / nt;
  const int tpp = 64 / g;  // target bodies per pass
  for (int32_t t0 = 0; t0 < nt; t0 += tpp) {
    const int32_t tl = t0 + lane / g;
    const int sub = lane % g;
    const bool tvalid = lane < g * tpp && tl < nt;
    double xi = 0, yi = 0, zi = 0, qi = 0;
    if (tvalid) {
      xi = b.x[tb + tl];
      yi = b.y[tb + tl];
      zi = b.z[tb + tl];
      qi = b.q[tb + tl];
    }
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int32_t s = lb; s < le; s++) {
      const int32_t sj = list.src[s];
      const int32_t sb = c.beg[sj], ns = c.nb[sj];
      for (int32_t c0 = 0; c0 < ns; c0 += P2P_CHUNK) {
        const int32_t cn = min(P2P_CHUNK, ns - c0);
        __syncthreads();
        if (lane < cn) {
          sx[lane] = b.x[sb + c0 + lane];
          sy[lane] = b.y[sb + c0 + lane];
          sz[lane] = b.z[sb + c0 + lane];
          sq[lane] = b.q[sb + c0 + lane];
        }
        __syncthreads();
        if (tvalid) {
          for (int jj = sub; jj < cn; jj += g) {
            const double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
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
    }
    __syncthreads();
    red[0][lane] = a0;
    red[1][lane] = a1;
    red[2][lane] = a2;
    red[3][lane] = a3;
    __syncthreads();
    if (tvalid && sub == 0) {
      for (int u = 1; u < g; u++) {
        a0 += red[0][lane + u];
        a1 += red[1][lane + u];
        a2 += red[2][lane + u];
        a3 += red[3][lane + u];
      }
      double *o = acc + 4 * (int64_t)(tb + tl);
      o[0] = a0;
      o[1] = a1;
      o[2] = a2;
      o[3] = a3;
    }
  }
}

}  // namespace

hipError_t launch_p2p_fp64(Cells c, int32_t ncells, Csr list, Bodies b, double *acc,
                           hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_p2p_fp64<<<ncells, 64, 0, s>>>(c, list, b, acc);
  return hipGetLastError();
}

}  // namespace tfmm

// ---------------------------------------------------------------------------
// Mixed precision: FP32 interactions, compensated accumulation.
// ---------------------------------------------------------------------------
namespace tfmm {
namespace {

// body position relative to its leaf's centre (FP32) and charge
__global__ __launch_bounds__(64) void k_leaf_relpos(Cells c, Bodies b, float4 *rel) {
  const int32_t ci = blockIdx.x;
  if (c.child[ci] >= 0) return;
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  for (int32_t i = beg + threadIdx.x; i < end; i += 64)
    rel[i] = make_float4((float)(b.x[i] - cx), (float)(b.y[i] - cy), (float)(b.z[i] - cz),
                         (float)b.q[i]);
}

// Kahan step: (s, c) += v
__device__ inline void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Target leaf per wave; positions are relative to the own leaf centre, so a
// source leaf is shifted by the FP64 centre difference (rounded once) and
// dX = x_i - x_j keeps the FP32 precision of the leaf size, not of the
// coordinates.  Each source leaf's contribution is summed in FP32 and then
// added to the running per-target sums with Kahan compensation.
//
// Latency: the metadata of 64 source leaves at a time (first body, count,
// centre offset) is loaded by the 64 lanes together and read back with
// readlane; the bodies of the next source chunk (<= 64 bodies) are loaded
// into registers while the current chunk is computed from LDS.  The block is
// one wave, whose LDS operations complete in issue order, so one staging
// buffer suffices.
// The work of one target leaf (one wave); returns at once for other cells.
__device__ inline void p2p_fp32_leaf(int32_t ti, Cells c, Csr list, Bodies b, const float4 *rel,
                                     double *acc, float4 *sb, double (*red)[64]) {
  if (c.child[ti] >= 0) return;
  const int32_t nt = c.nb[ti];
  if (nt == 0) return;
  const int32_t tb = c.beg[ti];
  const int lane = threadIdx.x;
  const int32_t lb = list.off[ti], le = list.off[ti + 1];
  const double tcx = c.center[3 * (int64_t)ti], tcy = c.center[3 * (int64_t)ti + 1],
               tcz = c.center[3 * (int64_t)ti + 2];
  const int g = nt >= 64 ? 1 : 64 / nt;  // lanes per target body
  const int tpp = 64 / g;                // target bodies per pass
  for (int32_t t0 = 0; t0 < nt; t0 += tpp) {
    const int32_t tl = t0 + lane / g;
    const int sub = lane % g;
    const bool tvalid = lane < g * tpp && tl < nt;
    float4 xi = make_float4(0, 0, 0, 0);
    if (tvalid) xi = rel[tb + tl];
    float s0 = 0, s1 = 0, s2 = 0, s3 = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (int32_t base = lb; base < le; base += 64) {
      // metadata of source leaves base .. base + 63, one per lane
      const int nsl = min(64, le - base);
      int32_t mbeg = 0, mnb = 0, msrc = -1;
      float mox = 0, moy = 0, moz = 0;
      if (lane < nsl) {
        const int32_t sj = list.src[base + lane];
        msrc = sj;
        mbeg = c.beg[sj];
        mnb = c.nb[sj];
        mox = (float)(tcx - c.center[3 * (int64_t)sj]);
        moy = (float)(tcy - c.center[3 * (int64_t)sj + 1]);
        moz = (float)(tcz - c.center[3 * (int64_t)sj + 2]);
      }
      // chunk (k, off): bodies off .. off + 63 of source leaf k
      int k = 0, off = 0;
      int32_t kb = __builtin_amdgcn_readlane(mbeg, 0), kn = __builtin_amdgcn_readlane(mnb, 0);
      int32_t ks = __builtin_amdgcn_readlane(msrc, 0);
      float kox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mox), 0));
      float koy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moy), 0));
      float koz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moz), 0));
      float4 v = make_float4(0, 0, 0, 0);
      if (lane < kn) v = rel[kb + lane];
      while (true) {
        const int cn = min(P2P_CHUNK, kn - off);
        const bool self = ks == ti;  // only the own leaf can hold r = 0 pairs
        // stage the current chunk (shifted into the target leaf's frame)
        __builtin_amdgcn_wave_barrier();
        sb[lane] = make_float4(v.x - kox, v.y - koy, v.z - koz, v.w);
        __builtin_amdgcn_wave_barrier();
        // advance to the next chunk and start its load
        off += P2P_CHUNK;
        bool more = true;
        if (off >= kn) {
          off = 0;
          k++;
          if (k >= nsl) {
            more = false;
          } else {
            kb = __builtin_amdgcn_readlane(mbeg, k);
            kn = __builtin_amdgcn_readlane(mnb, k);
            ks = __builtin_amdgcn_readlane(msrc, k);
            kox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mox), k));
            koy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moy), k));
            koz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moz), k));
          }
        }
        if (more && off + lane < kn) v = rel[kb + off + lane];
        // the current chunk
        float p0 = 0, p1 = 0, p2 = 0, p3 = 0;
        if (tvalid) {
          if (self) {
            for (int jj = sub; jj < cn; jj += g) {
              const float4 xj = sb[jj];
              const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
              const float r2 = dx * dx + dy * dy + dz * dz;
              const float rinv = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
              const float qr = xj.w * rinv;
              const float f = qr * rinv * rinv;
              p0 += qr;
              p1 -= dx * f;
              p2 -= dy * f;
              p3 -= dz * f;
            }
          } else {
            // bodies of different leaves sit at different positions (equal
            // positions give equal keys), so r > 0 unless they are closer than
            // the FP32 resolution of the leaf-relative coordinates
            auto pair = [&](const float4 xj) {
              const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
              const float r2 = dx * dx + dy * dy + dz * dz;
              const float rinv = __builtin_amdgcn_rsqf(r2);
              const float qr = xj.w * rinv;
              const float f = qr * rinv * rinv;
              p0 += qr;
              p1 -= dx * f;
              p2 -= dy * f;
              p3 -= dz * f;
            };
            if (g == 1) {
              // one lane per target (leaves of >= 33 bodies, most of the
              // pairs): the source loop is uniform across the wave, four
              // sources per trip at immediate LDS offsets, so the counter,
              // bound test and address step (three of the ~19 VALU slots of
              // a pair) are paid once per four pairs; same order of the sums
              int jj = 0;
              for (; jj + 3 < cn; jj += 4) {
                const float4 xa = sb[jj], xb = sb[jj + 1], xc = sb[jj + 2], xd = sb[jj + 3];
                pair(xa);
                pair(xb);
                pair(xc);
                pair(xd);
              }
              for (; jj < cn; jj++) pair(sb[jj]);
            } else {
              for (int jj = sub; jj < cn; jj += g) pair(sb[jj]);
            }
          }
        }
        kahan_add(s0, c0, p0);
        kahan_add(s1, c1, p1);
        kahan_add(s2, c2, p2);
        kahan_add(s3, c3, p3);
        if (!more) break;
      }
    }
    double d0 = (double)s0 - c0, d1 = (double)s1 - c1, d2 = (double)s2 - c2, d3 = (double)s3 - c3;
    __syncthreads();
    red[0][lane] = d0;
    red[1][lane] = d1;
    red[2][lane] = d2;
    red[3][lane] = d3;
    __syncthreads();
    if (tvalid && sub == 0) {
      for (int u = 1; u < g; u++) {
        d0 += red[0][lane + u];
        d1 += red[1][lane + u];
        d2 += red[2][lane + u];
        d3 += red[3][lane + u];
      }
      // the FP64 path stores q_i * sum (the reference's P2P sum, divided by
      // q_i in L2P); keep that convention
      const double qi = b.q[tb + tl];
      double *o = acc + 4 * (int64_t)(tb + tl);
      o[0] = d0 * qi;
      o[1] = d1 * qi;
      o[2] = d2 * qi;
      o[3] = d3 * qi;
    }
  }
}


__global__ __launch_bounds__(64) void k_p2p_fp32(Cells c, Csr list, Bodies b,
                                                const float4 *rel, double *acc) {
  __shared__ float4 sb[P2P_CHUNK];
  __shared__ double red[4][64];
  p2p_fp32_leaf(blockIdx.x, c, list, b, rel, acc, sb, red);
}

// A fixed number of resident waves that take target cells from a counter
// (zeroed before the launch) until none are left: the P2P then holds a
// bounded share of the CUs, and kernels on the other stream still find
// slots.  Each target is computed by one wave exactly as above, so the order
// in which the waves take them does not change any result.
__global__ __launch_bounds__(64) void k_p2p_fp32_resident(Cells c, int32_t ncells, Csr list,
                                                         Bodies b, const float4 *rel,
                                                         double *acc, uint32_t *next) {
  __shared__ float4 sb[P2P_CHUNK];
  __shared__ double red[4][64];
  __shared__ uint32_t take;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) take = atomicAdd(next, 1u);
    __syncthreads();
    const uint32_t ti = take;
    if (ti >= (uint32_t)ncells) return;
    p2p_fp32_leaf((int32_t)ti, c, list, b, rel, acc, sb, red);
  }
}
}  // namespace

hipError_t launch_leaf_relpos(Cells c, int32_t ncells, Bodies b, float4 *rel, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_leaf_relpos<<<ncells, 64, 0, s>>>(c, b, rel);
  return hipGetLastError();
}

hipError_t launch_p2p_fp32(Cells c, int32_t ncells, Csr list, Bodies b, const float4 *rel,
                           double *acc, int resident, uint32_t *next, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  if (resident > 0 && next) {
    hipError_t e = launch_fill_zero(next, 4, s);
    if (e != hipSuccess) return e;
    const int grid = resident < ncells ? resident : ncells;
    k_p2p_fp32_resident<<<grid, 64, 0, s>>>(c, ncells, list, b, rel, acc, next);
  } else {
    k_p2p_fp32<<<ncells, 64, 0, s>>>(c, list, b, rel, acc);
  }
  return hipGetLastError();
}

}  // namespace tfmm
