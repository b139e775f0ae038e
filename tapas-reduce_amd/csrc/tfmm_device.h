// tfmm_device.h -- device-side building blocks shared by the TapasFMM HIP
// kernels: Morton-key algebra, FP64 complex arithmetic and the solid-harmonic
// recurrences of the Laplace spherical FMM at expansion order P = 10.
//
// Layout conventions (gfx950, wave64):
//  * coefficients of one cell: 55 complex doubles, index n*(n+1)/2 + m
//    (0 <= m <= n < P), as CellAttr::M/L of the reference
//    (sample/exafmm-dev-13274dd4ac68/examples/tapas_exafmm.h:10-17);
//  * harmonics: only the m >= 0 half is stored (55 entries, same index);
//    Y_n^{-m} = conj(Y_n^m), the conjugate relation the reference uses at
//    LaplaceSphericalCPU_tapas.h:76,87,116,125.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfmm {

constexpr int P = 10;
constexpr int NTERM = P * (P + 1) / 2;  // 55
constexpr int MAXDEPTH = 19;            // 64-bit keys, 3-D (sfc_morton.h:154-180)
constexpr int DEPTHBITS = 5;
constexpr int WAVE = 64;

// Periodic images (EX/include/traversal.h:247-268): the source side of a
// traversal pair, list entry or P2P/M2L source carries the image it is seen
// through in its top bits: word = cell | v << IMG_SHIFT, image (ix, iy, iz)
// in {-1, 0, 1}^3 as v = 9 (ix + 1) + 3 (iy + 1) + (iz + 1) (the reference's
// loop order), the source displaced by (ix, iy, iz) * cycle.  Without
// periodic images cycle = 0 and v = 0: no displacement.
constexpr int IMG_SHIFT = 27;
constexpr uint32_t CELL_MASK = (1u << IMG_SHIFT) - 1u;
constexpr uint32_t IMG_CENTRE = 13;  // (0, 0, 0)
__host__ __device__ inline void image_shift(uint32_t v, double cycle, double &sx, double &sy,
                                            double &sz) {
  if (cycle == 0) {  // +0: subtracting it changes no value (nor the sign of a zero)
    sx = sy = sz = 0;
    return;
  }
  sx = cycle * (double)((int)(v / 9) - 1);
  sy = cycle * (double)((int)(v / 3 % 3) - 1);
  sz = cycle * (double)((int)(v % 3) - 1);
}

__host__ __device__ inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// ---------------------------------------------------------------------------
// Morton keys (cpp/include/tapas/sfc_morton.h:187-211, 288-295, 387-402, 536-557)
// ---------------------------------------------------------------------------
__host__ __device__ inline int key_depth(uint64_t k) { return (int)(k & 31u); }

__host__ __device__ inline uint64_t key_next(uint64_t k) {
  return k + ((uint64_t)1 << (3 * (MAXDEPTH - key_depth(k)) + DEPTHBITS));
}

__host__ __device__ inline uint64_t key_child(uint64_t k, int idx) {
  const int d = key_depth(k) + 1;
  const uint64_t kk = ((k >> DEPTHBITS) << DEPTHBITS) | (uint64_t)d;
  return kk | ((uint64_t)idx << ((MAXDEPTH - d) * 3 + DEPTHBITS));
}

// spread the low 19 bits of v to every third bit (bit b -> bit 3b)
__host__ __device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x7ffffu;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Finest-level key of an integer anchor: per level the bits are ordered
// z, y, x (x least significant), most significant level first; the anchor's
// low 19 bits are used, as the reference masks them (sfc_morton.h:387-402).
__host__ __device__ inline uint64_t finest_key(int ax, int ay, int az) {
  const uint64_t bits = spread3((uint32_t)ax) | (spread3((uint32_t)ay) << 1) |
                        (spread3((uint32_t)az) << 2);
  return (bits << DEPTHBITS) | (uint64_t)MAXDEPTH;
}

// Region of a cell by repeated halving from the root region.  The halving
// order (depth outer, dimension inner) and the midpoint expression match
// sfc_morton.h:536-557 so that widths and centres are bit-identical.
__host__ __device__ inline void cell_region(uint64_t key, const double rmin[3],
                                            const double rmax[3], double cmin[3],
                                            double cmax[3]) {
#pragma clang fp contract(off)
  double mx[3] = {rmax[0], rmax[1], rmax[2]};
  double mn[3] = {rmin[0], rmin[1], rmin[2]};
  const int depth = key_depth(key);
  const uint64_t body = key >> DEPTHBITS;
  for (int dep = 1; dep <= depth; dep++) {
    const uint32_t dir = (uint32_t)(body >> ((MAXDEPTH - dep) * 3));
    for (int dim = 0; dim < 3; dim++) {
      const double c = (mx[dim] + mn[dim]) / 2;
      if ((dir >> dim) & 1) mn[dim] = c;
      else mx[dim] = c;
    }
  }
  for (int d = 0; d < 3; d++) { cmin[d] = mn[d]; cmax[d] = mx[d]; }
}

// ---------------------------------------------------------------------------
// FP64 complex
// ---------------------------------------------------------------------------
struct cd {
  double re, im;
};
__host__ __device__ inline cd mk(double re, double im) { cd r; r.re = re; r.im = im; return r; }
__host__ __device__ inline cd cadd(cd a, cd b) { return mk(a.re + b.re, a.im + b.im); }
__host__ __device__ inline cd cmul(cd a, cd b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ inline cd cscale(cd a, double s) { return mk(a.re * s, a.im * s); }
__host__ __device__ inline cd cconj(cd a) { return mk(a.re, -a.im); }

// ---------------------------------------------------------------------------
// Spherical coordinates (LaplaceSphericalCPU_tapas.h:43-47)
// ---------------------------------------------------------------------------
__device__ inline void cart2sph(double dx, double dy, double dz, double &r, double &theta,
                                double &phi) {
  r = sqrt(dx * dx + dy * dy + dz * dz);
  theta = (r == 0) ? 0.0 : acos(dz / r);
  phi = atan2(dy, dx);
}

// What the harmonic recurrences actually consume -- cos(theta), sin(theta)
// and e^{i phi} -- formed without inverse trigonometry:
//   cos(theta) = z / r,  sin(theta) = r_xy / r,  e^{i phi} = (x + i y) / r_xy.
// The reference forms theta = acos(z/r), phi = atan2(y, x) and then takes
// cos/sin of them again; the values agree to rounding.  Degenerate cases
// follow the reference: r = 0 gives theta = 0 (cos 1, sin 0) and r_xy = 0
// gives phi = atan2(0, x) (e^{i phi} = +-1).
__device__ inline void sph_frame(double dx, double dy, double dz, double &r, double &ct,
                                 double &st, cd &eip) {
  const double rxy2 = dx * dx + dy * dy;
  r = sqrt(rxy2 + dz * dz);
  if (r == 0) {
    ct = 1;
    st = 0;
    eip = mk(1, 0);
    return;
  }
  ct = dz / r;
  const double rxy = sqrt(rxy2);
  st = rxy / r;
  if (rxy > 0) eip = mk(dx / rxy, dy / rxy);
  else eip = mk(dx < 0 ? -1.0 : 1.0, 0.0);
}

// Column m of the singular solid harmonics r^{-n-1} Y_n^m (evalLocal,
// LaplaceSphericalCPU_tapas.h:102-135), restarted at column m.
template <class Store>
__device__ inline void local_column(int m, double rho, double x, double y, cd ei, Store store) {
  double fact = 1, pn = 1;
  const double invR = -1.0 / rho;
  double rhom = -invR;
  cd eim = mk(1.0, 0.0);
  for (int i = 0; i < m; i++) {
    rhom *= invR;
    pn = -pn * fact * y;
    fact += 2;
    eim = cmul(eim, ei);
  }
  double p = pn;
  store(m, m, cscale(eim, rhom * p));
  double p1 = p;
  p = x * (2 * m + 1) * p1;
  double rhon = rhom * invR;
  for (int n = m + 1; n < P; n++) {
    store(n, m, cscale(eim, rhon * p));
    const double p2 = p1;
    p1 = p;
    p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
    rhon *= invR * (n - m + 1);
  }
}

// Reciprocals of the integer divisors of the recurrences: the divisions of
// the reference's recurrences become multiplications (a difference of at
// most one rounding per step).
__host__ __device__ constexpr double recip_int(int k) { return 1.0 / (double)k; }

// All columns of the regular harmonics in the reference's order (m outer,
// n inner), one serial recurrence per lane (evalMultipole,
// LaplaceSphericalCPU_tapas.h:63-99).  store(n, m, Y, dY/dtheta) is called
// once per (n, m >= 0), 55 calls in a fixed order; the theta derivative is
// only formed if WithTheta (inv_y = 1 / sin(theta) is then required).
template <bool WithTheta, class Store>
__device__ inline void multipole_all(double rho, double x, double y, cd ei, Store store,
                                     double inv_y = 0.0) {
  double fact = 1, pn = 1, rhom = 1;
  cd eim = mk(1.0, 0.0);
#pragma unroll
  for (int m = 0; m < P; m++) {
    double p = pn;
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    store(m, m, cscale(eim, rhom * p1),
          WithTheta ? cscale(eim, rhom * (p - (m + 1) * x * p1) * inv_y) : mk(0, 0));
    rhom *= rho;
    double rhon = rhom;
#pragma unroll
    for (int n = m + 1; n < P; n++) {
      rhon *= -recip_int(n + m);
      const cd yv = cscale(eim, rhon * p);
      const double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) * recip_int(n - m + 1);
      store(n, m, yv,
            WithTheta ? cscale(eim, rhon * ((n - m + 1) * p - (n + 1) * x * p1) * inv_y)
                      : mk(0, 0));
      rhon *= rho;
    }
    rhom *= -recip_int((2 * m + 2) * (2 * m + 1));
    pn = -pn * fact * y;
    fact += 2;
    eim = cmul(eim, ei);
  }
}

// Y_n^{mm} for any sign of mm from the stored m >= 0 half.
__device__ inline cd yfull(const cd *Y, int n, int mm) {
  return mm >= 0 ? Y[tri(n, mm)] : cconj(Y[tri(n, -mm)]);
}

// Append from every thread of a block with pred set: one global atomic per
// block (same-address atomics per wave serialise in L2 and dominated the
// traversal).  Returns the slot of this thread (valid where pred is set);
// every thread of the block must call it.  scratch: >= NW + 1 uints of LDS.
template <int NW>
__device__ inline uint32_t block_append(uint32_t *counter, bool pred, uint32_t per_lane,
                                        uint32_t *scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t m = __ballot(pred);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  if (lane == 0) scratch[wid] = (uint32_t)__popcll(m) * per_lane;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < NW; w++) tot += scratch[w];
    scratch[NW] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  uint32_t pre = scratch[NW];
  for (int w = 0; w < wid; w++) pre += scratch[w];
  const uint32_t slot = pre + (uint32_t)__popcll(m & lt) * per_lane;
  __syncthreads();
  return slot;
}

// Issue priority of the calling wave among the waves of its SIMD.  The
// kernels that run beside the P2P before it is joined (the M2L batch
// building and its sorts, the upward pass, the M2L kernels) raise theirs:
// the P2P's resident waves, at the default priority, then take the VALU
// issue slots those leave instead of competing for every one, and the
// path to the downward pass shortens.  (Measured on the 4M Plummer with
// the P2P started early as 3 resident waves per SIMD: 11.8 against 12.9
// ms per evaluation; on the 256-problem forest 14.4 against 16.5.)
__device__ inline void critical_path_priority() { __builtin_amdgcn_s_setprio(2); }

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

}  // namespace tfmm
