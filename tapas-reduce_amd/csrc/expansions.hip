// expansions.hip -- FP64 upward (P2M, M2M) and downward (L2L, L2P) kernels.
//
// Reference (sample/exafmm-dev-13274dd4ac68/examples/):
//   FMM_Upward / FMM_Downward   tapas_exafmm.cxx:87-140
//   P2M  LaplaceSphericalCPU_tapas.h:138-163     L2P  :275-307
// (M2M and L2L are in translate.hip.)
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

// P2M: M[n,m] += q * conj(Y_n^m(body - centre)), summed over the leaf's
// bodies.  One wave per leaf, one body per lane (chunks of 64 bodies).  The
// harmonics come out of the recurrence in a fixed visit order; every 16 real
// values (8 terms) are staged through LDS as a 64 x 16 block whose columns
// the wave sums (16 rows per quarter-wave, then across quarters), so the
// body sum costs 16 LDS writes + 16 LDS reads per 16 values instead of a
// full wave reduction per value.  Summation order is fixed.
constexpr int P2M_WAVES = 4;
constexpr int NVAL = 2 * NTERM;          // 110 real values per cell
constexpr int NCHUNK = (NVAL + 15) / 16;  // 7

// (n, m) of the v-th term of multipole_all's visit order (m outer, n inner)
__device__ inline int visit_term(int v) {
  int m = 0;
  while (v >= P - m) {
    v -= P - m;
    m++;
  }
  return tri(m + v, m);
}

__global__ __launch_bounds__(64 * P2M_WAVES) void k_p2m(Cells c, int32_t ncells, Bodies b,
                                                        double *M) {
  critical_path_priority();
  __shared__ double stage[P2M_WAVES][64][17];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ci = blockIdx.x * P2M_WAVES + w;
  if (ci >= ncells || c.child[ci] >= 0 || c.nb[ci] == 0) return;  // per wave; no block barrier
  double (*st)[17] = stage[w];
  const int col = lane & 15, rg = lane >> 4;
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  double own0 = 0, own1 = 0;  // this lane's sums of chunks rg and rg + 4
  for (int32_t base = beg; base < end; base += 64) {
    const int32_t i = base + lane;
    const int rows = min(64, end - base);
    double dx = 1, dy = 0, dz = 0, q = 0;
    if (i < end) {
      dx = b.x[i] - cx;
      dy = b.y[i] - cy;
      dz = b.z[i] - cz;
      q = b.q[i];
    }
    double rho, ca, sa;
    cd ei;
    sph_frame(dx, dy, dz, rho, ca, sa, ei);
    double buf[16];
    int v = 0;  // visit counter (a constant after unrolling)
    multipole_all<false>(rho, ca, sa, ei, [&](int, int, cd Y, cd) {
      buf[(2 * v) & 15] = q * Y.re;
      buf[(2 * v + 1) & 15] = -q * Y.im;
      const int filled = 2 * v + 2;
      if ((filled & 15) == 0 || filled == NVAL) {
        const int ch = (filled - 1) >> 4;
        const int nval = filled - 16 * ch;
#pragma unroll
        for (int k = 0; k < 16; k++)
          if (k < nval) st[lane][k] = buf[k];
        __builtin_amdgcn_wave_barrier();
        double sum = 0;
        if (col < nval) {
          const int r1 = min(rows, 16 * rg + 16);
          for (int r = 16 * rg; r < r1; r++) sum += st[r][col];
        }
        sum += __shfl_xor(sum, 16, 64);
        sum += __shfl_xor(sum, 32, 64);
        if (ch == rg) own0 += sum;
        if (ch == rg + 4) own1 += sum;
        __builtin_amdgcn_wave_barrier();
      }
      v++;
    });
  }
  double *Mc = M + 2 * (int64_t)NTERM * ci;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int ch = rg + 4 * h;
    const int f = 16 * ch + col;  // value index in visit order
    if (ch < NCHUNK && f < NVAL) Mc[2 * visit_term(f >> 1) + (f & 1)] += h ? own1 : own0;
  }
}

// leaf of every (sorted) body
__global__ void k_body_leaf(Cells c, int32_t ncells, int32_t *leaf_of) {
  const int32_t ci = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (ci >= ncells || c.child[ci] >= 0) return;
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  for (int32_t i = beg + (threadIdx.x & 63); i < end; i += 64) leaf_of[i] = ci;
}

// L2P: trg = p2p / q + local field of the leaf's expansion; one thread per
// body, written in sorted order (trg) and in input order (out[perm]).
// The block's bodies are consecutive in sorted order, so they belong to a few
// runs of one leaf each: the expansions of the first L2P_RUNS runs are
// staged in LDS and read from there (bodies of later runs read global
// memory).
constexpr int L2P_RUNS = 16;

__global__ __launch_bounds__(256) void k_l2p(Cells c, Bodies b, const int32_t *leaf_of,
                                             const double *L, const double *p2p,
                                             const int32_t *perm, double *trg, double *out) {
  __shared__ cd Lrun[L2P_RUNS][NTERM];
  __shared__ int32_t run_leaf[L2P_RUNS];
  __shared__ uint32_t wave_runs[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  const bool valid = i < b.n;
  const int32_t ci = valid ? leaf_of[i] : -1;
  // run index of this body: run starts before it in the block
  const bool start = valid && (i == i0 || leaf_of[i - 1] != ci);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t m = __ballot(start);
  if (lane == 0) wave_runs[wid] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wid; w++) before += wave_runs[w];
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int run = (int)(before + (uint32_t)__popcll(m & upto)) - 1;
  if (start && run < L2P_RUNS) run_leaf[run] = ci;
  __syncthreads();
  const int nruns = (int)(wave_runs[0] + wave_runs[1] + wave_runs[2] + wave_runs[3]);
  const int staged = nruns < L2P_RUNS ? nruns : L2P_RUNS;
  for (int e = threadIdx.x; e < staged * NTERM; e += blockDim.x) {
    const int r = e / NTERM, t = e - r * NTERM;
    Lrun[r][t] = reinterpret_cast<const cd *>(L)[(int64_t)run_leaf[r] * NTERM + t];
  }
  __syncthreads();
  if (!valid) return;
  const cd *Ls = run < L2P_RUNS ? Lrun[run] : reinterpret_cast<const cd *>(L) + (int64_t)ci * NTERM;
  const double dx = b.x[i] - c.center[3 * (int64_t)ci], dy = b.y[i] - c.center[3 * (int64_t)ci + 1],
               dz = b.z[i] - c.center[3 * (int64_t)ci + 2];
  const double q = b.q[i];
  double r, ct, st;
  cd ei;
  sph_frame(dx, dy, dz, r, ct, st, ei);
  const double iq = 1.0 / q, ir = 1.0 / r;
  double t0 = p2p[4 * i] * iq, t1 = p2p[4 * i + 1] * iq, t2 = p2p[4 * i + 2] * iq,
         t3 = p2p[4 * i + 3] * iq;
  double s0 = 0, s1 = 0, s2 = 0;
  multipole_all<true>(
      r, ct, st, ei,
      [&](int n, int m, cd Y, cd YT) {
        const cd l = Ls[tri(n, m)];
        const cd ly = cmul(l, Y);
        const double lyt = cmul(l, YT).re;
        if (m == 0) {
          t0 += ly.re;
          s0 += ly.re * ir * n;
          s1 += lyt;
        } else {
          t0 += 2 * ly.re;
          s0 += 2 * ly.re * ir * n;
          s1 += 2 * lyt;
          s2 += 2 * (-ly.im) * m;
        }
      },
      1.0 / st);
  const double sp = ei.im, cp = ei.re;
  t1 += st * cp * s0 + ct * cp * ir * s1 - sp * ir / st * s2;
  t2 += st * sp * s0 + ct * sp * ir * s1 + cp * ir / st * s2;
  t3 += ct * s0 - st * ir * s1;
  trg[4 * i] = t0;
  trg[4 * i + 1] = t1;
  trg[4 * i + 2] = t2;
  trg[4 * i + 3] = t3;
  const int64_t o = perm[i];
  out[4 * o] = t0;
  out[4 * o + 1] = t1;
  out[4 * o + 2] = t2;
  out[4 * o + 3] = t3;
}

// Zero fill (in place of hipMemsetAsync, which must not be recorded into a
// captured graph): 16-byte stores, the tail by single bytes.
__global__ void k_fill_zero(uint4 *p, size_t n16, unsigned char *tail, int ntail) {
  critical_path_priority();
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) p[i] = make_uint4(0, 0, 0, 0);
  if (i < (size_t)ntail) tail[i] = 0;
}

}  // namespace

hipError_t launch_fill_zero(void *ptr, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const size_t n16 = bytes / 16;
  const int ntail = (int)(bytes - n16 * 16);
  const size_t nt = n16 > (size_t)ntail ? n16 : (size_t)ntail;
  k_fill_zero<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(
      static_cast<uint4 *>(ptr), n16, static_cast<unsigned char *>(ptr) + n16 * 16, ntail);
  return hipGetLastError();
}

hipError_t launch_p2m(Cells c, int32_t ncells, Bodies b, double *M, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_p2m<<<(ncells + P2M_WAVES - 1) / P2M_WAVES, 64 * P2M_WAVES, 0, s>>>(c, ncells, b, M);
  return hipGetLastError();
}

// leaf_of: scratch of b.n int32
hipError_t launch_l2p(Cells c, int32_t ncells, Bodies b, const double *L, const double *p2p,
                      const int32_t *perm, int32_t *leaf_of, double *trg, double *out,
                      hipStream_t s) {
  if (ncells <= 0 || b.n <= 0) return hipSuccess;
  k_body_leaf<<<(ncells + 3) / 4, 256, 0, s>>>(c, ncells, leaf_of);
  k_l2p<<<(unsigned)((b.n + 255) / 256), 256, 0, s>>>(c, b, leaf_of, L, p2p, perm, trg, out);
  return hipGetLastError();
}

}  // namespace tfmm
