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
// One wave per target leaf.  The bodies of all the leaf's source leaves form
// one source stream, read in list order through an LDS ring and cut into
// chunks of 16 G consecutive bodies that cross source-leaf boundaries.  A
// target body is served by G lanes; lane s of a target takes segment s (16
// consecutive bodies) of every chunk, so all lanes run the same uniform
// 16-pair loop whatever the source leaves' sizes.  Each lane holds two
// targets (t and t + 64/G) and uses every source body it reads from LDS for
// both, so a pass over the stream serves 2 * (64 / G) targets.  G is chosen
// per leaf from its size to fill the lanes: a leaf of 40 bodies takes G = 3
// (one pass: 21 lanes x 2 targets, 120 of 126 pair slots busy) where one
// lane per target leaves 24 of 64 lanes idle.  The G partial sums of a target
// are added in lane order at the end of its pass, so results do not depend
// on timing.
//
// The R2 != 0 test costs one instruction: R2 is formed as dX.dX - tiny
// (tiny far below the smallest distance the coordinates resolve), so a zero
// distance gives a negative argument, rsqrt gives NaN, and max(NaN, 0) = 0
// drops the pair exactly.
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

constexpr int SEG = 16;  // bodies per lane and chunk

// Bodies of the source stream live in the ring at slot(p); one pad slot per
// 16 keeps the segments of one chunk (16 bodies apart) in different banks.
template <int RING>
__device__ inline int ring_slot(int32_t p) {
  const int r = p & (RING - 1);
  return r + (r >> 4);
}

// lanes per target body for a leaf of nt bodies: fewest pair slots over the
// passes (2 * 64 / g targets each; a last pass with at most 64 / g targets
// runs one target per lane), counting one pass over the source stream
// (staging) as ~1/30 of a 16-body segment
template <int GMAX>
__device__ inline int lanes_per_target(int32_t nt) {
  int g_best = 1;
  float best = 3.0e38f;
  for (int g = 1; g <= GMAX; g++) {
    const int tpp = 64 / g;
    const int full = nt / (2 * tpp), rem = nt - full * 2 * tpp;
    const float one = (60.f + 2.f * (float)g) / (float)g, two = (120.f + 2.f * (float)g) / (float)g;
    const float cost = (float)full * two + (rem == 0 ? 0.f : rem <= tpp ? one : two);
    if (cost < best * 0.999f) {
      best = cost;
      g_best = g;
    }
  }
  return g_best;
}

__device__ inline float rsq_guarded(float r2) {
  return __builtin_fmaxf(__builtin_amdgcn_rsqf(r2), 0.f);
}

// 1/sqrt(x) in FP64: the hardware estimate (relative error 5.2e-8 at most,
// about 24 bits) and one Newton step y <- y + y/2 (1 - x y^2), which squares
// the error: 4.1e-15 at most, a few ulp (both measured on gfx950 over 2^20
// arguments spread over [1e-12, 1e12]; a second step gives 1.4e-16 and costs
// 4 of the pair's 23 VALU instructions).  A negative x gives NaN, and
// max(NaN, 0) = 0.
__device__ inline double rsq_guarded(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  return __builtin_fmax(y, 0.0);
}

__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

// One pair: target xi, source xj (w = charge); p0 += q/r, p1..3 -= q dX / r^3
template <typename T, typename V4>
__device__ inline void pair(const V4 &xi, const V4 &xj, T tiny, T &p0, T &p1, T &p2, T &p3) {
  const T dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
  T r2 = fma_t(dx, dx, -tiny);
  r2 = fma_t(dy, dy, r2);
  r2 = fma_t(dz, dz, r2);
  const T rinv = rsq_guarded(r2);
  const T qr = xj.w * rinv;
  const T f = qr * (rinv * rinv);
  p0 += qr;
  p1 = fma_t(-dx, f, p1);
  p2 = fma_t(-dy, f, p2);
  p3 = fma_t(-dz, f, p3);
}

// Kahan step: (s, c) += v
__device__ inline void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Running sums of one lane: FP32 partials over at most 64 pairs, folded into
// Kahan-compensated FP32 sums (mixed); plain FP64 sums (FP64).
template <typename T> struct Acc;
template <> struct Acc<float> {
  float p0 = 0, p1 = 0, p2 = 0, p3 = 0;
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  __device__ void flush() {
    kahan_add(s0, c0, p0);
    kahan_add(s1, c1, p1);
    kahan_add(s2, c2, p2);
    kahan_add(s3, c3, p3);
    p0 = p1 = p2 = p3 = 0;
  }
  __device__ double4 total() {
    return make_double4((double)s0 - c0, (double)s1 - c1, (double)s2 - c2, (double)s3 - c3);
  }
};
template <> struct Acc<double> {
  double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
  __device__ void flush() {}
  __device__ double4 total() { return make_double4(p0, p1, p2, p3); }
};

template <typename T>
__device__ inline typename Vec4<T>::type load_body(const Bodies &b, const float4 *rel,
                                                   int32_t i);
template <>
__device__ inline float4 load_body<float>(const Bodies &, const float4 *rel, int32_t i) {
  return rel[i];
}
template <>
__device__ inline double4 load_body<double>(const Bodies &b, const float4 *, int32_t i) {
  return make_double4(b.x[i], b.y[i], b.z[i], b.q[i]);
}

// One pass over the source stream for the targets t (and t2 when TWO) of
// every lane; see p2p_leaf.
template <typename T, int RING, bool TWO>
__device__ inline void p2p_pass(const Cells &c, const Csr &list, const Bodies &b,
                                const float4 *rel, double *acc, typename Vec4<T>::type *ring,
                                int32_t tb, int32_t lb, int32_t le, double tcx, double tcy,
                                double tcz, double cycle, T tiny, int G, int32_t t, int32_t t2,
                                bool tvalid, bool tvalid2) {
  using V4 = typename Vec4<T>::type;
  constexpr bool MIX = sizeof(T) == 4;
  const int lane = threadIdx.x;
  const int B = SEG * G;
  const int seg = lane % G;
  double4 *red = reinterpret_cast<double4 *>(ring);
  V4 xi{}, xi2{};
  if (tvalid) xi = load_body<T>(b, rel, tb + t);
  if (TWO && tvalid2) xi2 = load_body<T>(b, rel, tb + t2);
  Acc<T> a, a2;
  int32_t wpos = 0, cpos = 0;
  int since = 0;
  // chunk [cpos, cpos + B) of the stream: segment seg of this lane
  auto compute = [&]() {
    if (tvalid) {
      const V4 *sp = ring + ring_slot<RING>(cpos + SEG * seg);
#pragma unroll
      for (int k = 0; k < SEG; k++) {
        const V4 xj = sp[k];
        pair<T>(xi, xj, tiny, a.p0, a.p1, a.p2, a.p3);
        if (TWO) pair<T>(xi2, xj, tiny, a2.p0, a2.p1, a2.p2, a2.p3);
      }
    }
    cpos += B;
    since += SEG;
    if (since >= 64) {
      a.flush();
      if (TWO) a2.flush();
      since = 0;
    }
  };
  for (int32_t base = lb; base < le; base += 64) {
    // metadata of source leaves base .. base + 63, one per lane
    const int nsl = min(64, le - base);
    int32_t mbeg = 0, mnb = 0, mimg = 0;
    float mox = 0, moy = 0, moz = 0;
    if (lane < nsl) {
      // source word: leaf and periodic image (the leaf's bodies displaced by
      // the image's Xperiodic, LaplaceP2PCPU_tapas.h:54)
      const uint32_t sw = (uint32_t)list.src[base + lane];
      const int32_t sj = (int32_t)(sw & CELL_MASK);
      mimg = (int32_t)(sw >> IMG_SHIFT);
      mbeg = c.beg[sj];
      mnb = c.nb[sj];
      if constexpr (MIX) {
        double sx, sy, sz;
        image_shift((uint32_t)mimg, cycle, sx, sy, sz);
        mox = (float)(tcx - c.center[3 * (int64_t)sj] - sx);
        moy = (float)(tcy - c.center[3 * (int64_t)sj + 1] - sy);
        moz = (float)(tcz - c.center[3 * (int64_t)sj + 2] - sz);
      }
    }
    // FP64: the image displacement of the current source leaf
    double ksx = 0, ksy = 0, ksz = 0;
    if constexpr (!MIX)
      image_shift((uint32_t)__builtin_amdgcn_readlane(mimg, 0), cycle, ksx, ksy, ksz);
    // piece (k, off): bodies off .. off + 63 of source leaf k
    int k = 0, off = 0;
    int32_t kb = __builtin_amdgcn_readlane(mbeg, 0), kn = __builtin_amdgcn_readlane(mnb, 0);
    float kox = 0, koy = 0, koz = 0;
    if constexpr (MIX) {
      kox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mox), 0));
      koy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moy), 0));
      koz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moz), 0));
    }
    V4 v{};
    if (lane < kn) v = load_body<T>(b, rel, kb + lane);
    while (true) {
      const int n = min(64, kn - off);
      // stage the piece (shifted into the target leaf's frame)
      __builtin_amdgcn_wave_barrier();
      if (lane < n) {
        if constexpr (MIX)
          ring[ring_slot<RING>(wpos + lane)] = make_float4(v.x - kox, v.y - koy, v.z - koz, v.w);
        else if (cycle != 0)
          ring[ring_slot<RING>(wpos + lane)] = make_double4(v.x + ksx, v.y + ksy, v.z + ksz, v.w);
        else
          ring[ring_slot<RING>(wpos + lane)] = v;
      }
      __builtin_amdgcn_wave_barrier();
      wpos += n;
      // advance to the next piece and start its load
      off += 64;
      bool more = true;
      if (off >= kn) {
        off = 0;
        k++;
        if (k >= nsl) {
          more = false;
        } else {
          kb = __builtin_amdgcn_readlane(mbeg, k);
          kn = __builtin_amdgcn_readlane(mnb, k);
          if constexpr (MIX) {
            kox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mox), k));
            koy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moy), k));
            koz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(moz), k));
          } else {
            image_shift((uint32_t)__builtin_amdgcn_readlane(mimg, k), cycle, ksx, ksy, ksz);
          }
        }
      }
      if (more && off + lane < kn) v = load_body<T>(b, rel, kb + off + lane);
      // the chunks the ring now holds in full
      while (wpos - cpos >= B) compute();
      if (!more) break;
    }
  }
  // last, partial chunk: the rest of the chunk holds zero-charge bodies at
  // the target frame's origin (their pairs add exactly zero)
  if (wpos > cpos) {
    const int pad = cpos + B - wpos;
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < pad; i += 64) ring[ring_slot<RING>(wpos + i)] = V4{};
    __builtin_amdgcn_wave_barrier();
    compute();
  }
  a.flush();
  if (TWO) a2.flush();
  // the G partials of a target, added in lane order
  for (int h = 0; h < (TWO ? 2 : 1); h++) {
    const double4 d = h ? a2.total() : a.total();
    const int32_t th = h ? t2 : t;
    __builtin_amdgcn_wave_barrier();
    red[lane] = d;
    __builtin_amdgcn_wave_barrier();
    if ((h ? tvalid2 : tvalid) && seg == 0) {
      double d0 = d.x, d1 = d.y, d2 = d.z, d3 = d.w;
      for (int u = 1; u < G; u++) {
        const double4 e = red[lane + u];
        d0 += e.x;
        d1 += e.y;
        d2 += e.z;
        d3 += e.w;
      }
      // stored as q_i * sum, the reference's P2P sum (divided by q_i in L2P)
      const double qi = b.q[tb + th];
      double *o = acc + 4 * (int64_t)(tb + th);
      o[0] = d0 * qi;
      o[1] = d1 * qi;
      o[2] = d2 * qi;
      o[3] = d3 * qi;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// The work of one target leaf (one wave); returns at once for other cells.
// ring: RING + RING / 16 entries of LDS; it also holds the lane partials at
// the end of every pass.
template <typename T, int RING>
__device__ inline void p2p_leaf(int32_t ti, const Cells &c, const Csr &list, const Bodies &b,
                                const float4 *rel, double *acc, typename Vec4<T>::type *ring,
                                double cycle) {
  constexpr bool MIX = sizeof(T) == 4;
  constexpr int GMAX = RING / (2 * SEG);  // a chunk (16 G) plus a 64-body piece fit the ring
  if (c.child[ti] >= 0) return;
  const int32_t nt = c.nb[ti];
  if (nt == 0) return;
  const int32_t tb = c.beg[ti];
  const int lane = threadIdx.x;
  const int32_t lb = list.off[ti], le = list.off[ti + 1];
  const double tcx = c.center[3 * (int64_t)ti], tcy = c.center[3 * (int64_t)ti + 1],
               tcz = c.center[3 * (int64_t)ti + 2];
  // tiny: far below the square of the smallest resolved distance (FP32:
  // 2^-24 of the leaf width; FP64: any normal number)
  T tiny;
  if constexpr (MIX) {
    const float w = (float)c.width[ti];
    tiny = __builtin_fmaxf(w * w * 0x1p-80f, 0x1p-126f);
  } else {
    tiny = 0x1p-1022;
  }
  const int G = lanes_per_target<GMAX>(nt);
  const int tpp = 64 / G;  // targets per lane set
  const int tloc = lane / G;
  for (int32_t t0 = 0; t0 < nt; t0 += 2 * tpp) {
    // two targets per lane, t and t + tpp, while the pass has both
    const int32_t t = t0 + tloc, t2 = t + tpp;
    const bool tvalid = tloc < tpp && t < nt, tvalid2 = tloc < tpp && t2 < nt;
    if (t0 + tpp < nt)
      p2p_pass<T, RING, true>(c, list, b, rel, acc, ring, tb, lb, le, tcx, tcy, tcz, cycle, tiny,
                              G, t, t2, tvalid, tvalid2);
    else
      p2p_pass<T, RING, false>(c, list, b, rel, acc, ring, tb, lb, le, tcx, tcy, tcz, cycle, tiny,
                               G, t, t2, tvalid, false);
  }
}

constexpr int RING32 = 256;  // float4 entries: 4.25 KB with the pads
constexpr int RING64 = 128;  // double4 entries: 4.25 KB with the pads
constexpr int ring_entries(int ring) { return ring + ring / 16; }

__global__ __launch_bounds__(64) void k_p2p_fp64(Cells c, Csr list, Bodies b, double *acc,
                                                 double cycle) {
  __shared__ double4 ring[ring_entries(RING64)];
  p2p_leaf<double, RING64>(blockIdx.x, c, list, b, nullptr, acc, ring, cycle);
}

// The FP64 P2P as resident waves taking target cells from a counter, as
// k_p2p_fp32_resident below (each target computed by one wave exactly as by
// k_p2p_fp64, so the fields do not change).
__global__ __launch_bounds__(64) void k_p2p_fp64_resident(Cells c, int32_t ncells, Csr list,
                                                         Bodies b, double *acc, uint32_t *next,
                                                         double cycle) {
  __shared__ double4 ring[ring_entries(RING64)];
  __shared__ uint32_t take;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) take = atomicAdd(next, 1u);
    __syncthreads();
    const uint32_t ti = take;
    if (ti >= (uint32_t)ncells) return;
    p2p_leaf<double, RING64>((int32_t)ti, c, list, b, nullptr, acc, ring, cycle);
  }
}

}  // namespace

hipError_t launch_p2p_fp64(Cells c, int32_t ncells, Csr list, Bodies b, double *acc,
                           int resident, uint32_t *next, hipStream_t s, double cycle) {
  if (ncells <= 0) return hipSuccess;
  if (resident > 0 && next) {
    hipError_t e = launch_fill_zero(next, 4, s);
    if (e != hipSuccess) return e;
    const int grid = resident < ncells ? resident : ncells;
    k_p2p_fp64_resident<<<grid, 64, 0, s>>>(c, ncells, list, b, acc, next, cycle);
  } else {
    k_p2p_fp64<<<ncells, 64, 0, s>>>(c, list, b, acc, cycle);
  }
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

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_p2p_fp32(Cells c, Csr list, Bodies b, const float4 *rel, double *acc, double cycle) {
  __shared__ float4 ring[ring_entries(RING32)];
  p2p_leaf<float, RING32>(blockIdx.x, c, list, b, rel, acc, ring, cycle);
}

// A fixed number of resident waves that take target cells from a counter
// (zeroed before the launch) until none are left: the P2P then holds a
// bounded share of the CUs, and kernels on the other stream still find
// slots.  Each target is computed by one wave exactly as above, so the order
// in which the waves take them does not change any result.
__global__ __launch_bounds__(64) void k_p2p_fp32_resident(Cells c, int32_t ncells, Csr list,
                                                         Bodies b, const float4 *rel,
                                                         double *acc, uint32_t *next,
                                                         double cycle) {
  __shared__ float4 ring[ring_entries(RING32)];
  __shared__ uint32_t take;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) take = atomicAdd(next, 1u);
    __syncthreads();
    const uint32_t ti = take;
    if (ti >= (uint32_t)ncells) return;
    p2p_leaf<float, RING32>((int32_t)ti, c, list, b, rel, acc, ring, cycle);
  }
}
}  // namespace

hipError_t launch_leaf_relpos(Cells c, int32_t ncells, Bodies b, float4 *rel, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_leaf_relpos<<<ncells, 64, 0, s>>>(c, b, rel);
  return hipGetLastError();
}

// resident waves taking target cells from *next without clearing it first:
// several such launches (on different streams) share the cells of one
// cleared counter, each cell computed once by one wave
hipError_t launch_p2p_fp32_drain(Cells c, int32_t ncells, Csr list, Bodies b, const float4 *rel,
                                 double *acc, int waves, uint32_t *next, hipStream_t s,
                                 double cycle) {
  if (ncells <= 0 || waves <= 0) return hipSuccess;
  const int grid = waves < ncells ? waves : ncells;
  k_p2p_fp32_resident<<<grid, 64, 0, s>>>(c, ncells, list, b, rel, acc, next, cycle);
  return hipGetLastError();
}

hipError_t launch_p2p_fp32(Cells c, int32_t ncells, Csr list, Bodies b, const float4 *rel,
                           double *acc, int resident, uint32_t *next, hipStream_t s,
                           double cycle) {
  if (ncells <= 0) return hipSuccess;
  if (resident > 0 && next) {
    hipError_t e = launch_fill_zero(next, 4, s);
    if (e != hipSuccess) return e;
    const int grid = resident < ncells ? resident : ncells;
    k_p2p_fp32_resident<<<grid, 64, 0, s>>>(c, ncells, list, b, rel, acc, next, cycle);
  } else {
    k_p2p_fp32<<<ncells, 64, 0, s>>>(c, list, b, rel, acc, cycle);
  }
  return hipGetLastError();
}

}  // namespace tfmm
