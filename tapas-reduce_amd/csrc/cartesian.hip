// cartesian.hip -- the Cartesian Taylor basis (FP64): P2M, M2M, M2L, L2L,
// L2P and the far field of periodic images.
//
// Reference: sample/exafmm-dev-13274dd4ac68/kernels/LaplaceCartesianCPU.cxx
// (the reference's BASIS = Cartesian build, EX/include/types.h:39-41).
// Terms of degree 0 .. P-1 in three variables, CART_TERMS = P (P+1) (P+2) / 6
// doubles per expansion; term k = (kx, ky, kz) of degree d at
//   index(k) = d (d+1) (d+2) / 6 + (ky + kz) (ky + kz + 1) / 2 + kz
// (degree by degree, kx falling, then ky falling: Index<nx,ny,nz>::I, :3-25).
//   P2M  M_k += q (X - x)^k / k!                                   :636-644
//   M2M  M_n += sum_{k <= n} (Xp - Xc)^{n-k} / (n-k)! M_k             :646-657
//   M2L  L_n += sum_{|k| <= P-1-|n|} D^{n+k}(1/r)(Xi - Xj) M_k        :659-673
//   L2L  L_n += sum_{|k| <= P-1-|n|} (Xc - Xp)^k / k! L_{n+k}         :675-689
//   L2P  trg /= q; pot += sum_k e^k/k! L_k, grad += sum e^k/k! L_{k+e_d}  :691-704
// D^k(1/r) = k! a_k with the Taylor coefficients of 1/r from the recurrence
//   |k| r^2 a_k = (1 - 2|k|) sum_i x_i a_{k-e_i} + (1 - |k|) sum_i a_{k-2e_i}
// (the reference's DerivativeSum, :45-104, and scale, :177-181).
//
// The M2L is the one kernel with real work: 5005 products per pair, split
// very unevenly over the 220 output terms.  One wave handles one target
// cell; the products of a pair are dealt to its lanes by the fixed schedule
// of cart_sched.h (every lane issues 84 products per pair against the 78.2
// of a perfect split).  A lane keeps one FP64 accumulator per bin over all
// the target's sources; the pieces of a term are summed in a fixed order at
// the end.  The schedule (the C offsets of a lane's products, in registers)
// and the recurrence's index table are built on the host once per context.
#include <utility>

#include "cart_sched.h"
#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {
namespace {

using cart::BasisTables;
using cart::BL;
using cart::CartTables;
using cart::SPH_TERMS;
using cart::cdeg_off;
using cart::cidx;
using cart::CN;
using cart::CW;
using cart::MS_LEN;
using cart::NSLOT;
using cart::term_part;
using cart::ZS;
static_assert(cart::ORDER == P && CN == CART_TERMS && cart::NBINS == CART_BINS,
              "the schedule's sizes are the library's");
// multi-index of every term, for the kernels' run-time lookups
struct TermTab {
  uint8_t x[CN], y[CN], z[CN];
};
constexpr TermTab make_terms() {
  TermTab tt{};
  for (int t = 0; t < CN; t++) {
    tt.x[t] = (uint8_t)term_part(t, 0);
    tt.y[t] = (uint8_t)term_part(t, 1);
    tt.z[t] = (uint8_t)term_part(t, 2);
  }
  return tt;
}
__constant__ TermTab c_terms = make_terms();

// 1 / e for the power chains x^e / e! (a product in place of a division)
struct Inv {
  double v[P];
};
constexpr Inv make_inv() {
  Inv r{};
  for (int e = 1; e < P; e++) r.v[e] = 1.0 / e;
  return r;
}
__constant__ Inv c_inv = make_inv();
#define kInv c_inv.v

// M2M products: for each parent term n, the pairs (k, n - k), k <= n
// componentwise (k | (n - k) << 8), term by term from m2m_off[n]
constexpr int M2M_ENTRIES = 5005;
struct M2mTab {
  uint16_t off[CN + 1];
  uint16_t ent[M2M_ENTRIES];
};
constexpr M2mTab make_m2m() {
  M2mTab r{};
  int e = 0;
  for (int n = 0; n < CN; n++) {
    r.off[n] = (uint16_t)e;
    const int nx = term_part(n, 0), ny = term_part(n, 1), nz = term_part(n, 2);
    for (int kx = 0; kx <= nx; kx++)
      for (int ky = 0; ky <= ny; ky++)
        for (int kz = 0; kz <= nz; kz++)
          r.ent[e++] = (uint16_t)(cidx(kx, ky, kz) | cidx(nx - kx, ny - ky, nz - kz) << 8);
  }
  r.off[CN] = (uint16_t)e;
  return r;
}
__constant__ M2mTab c_m2m = make_m2m();

// ---------------------------------------------------------------------------
// P2M: one wave per leaf, lane l owns terms l, l+64, l+128, l+192.  Every
// body of a chunk of 64 writes its powers (X - x)^a / a! per axis (the
// charge folded into x's) to LDS; each lane then sums its terms over the
// chunk's bodies in body order.
constexpr int PW = 2;
constexpr int PWR = 3 * P;

__global__ __launch_bounds__(64 * PW) void k_cart_p2m(Cells c, int32_t ncells, Bodies b,
                                                      double *M) {
  critical_path_priority();
  __shared__ double tab[PW][64][PWR + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ci = blockIdx.x * PW + w;
  if (ci >= ncells || c.child[ci] >= 0 || c.nb[ci] == 0) return;  // per wave; no block barrier
  double (*tb)[PWR + 1] = tab[w];
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  int ox[4], oy[4], oz[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = min(lane + 64 * r, CN - 1);
    ox[r] = c_terms.x[t];
    oy[r] = P + c_terms.y[t];
    oz[r] = 2 * P + c_terms.z[t];
  }
  double acc[4] = {0, 0, 0, 0};
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  for (int32_t base = beg; base < end; base += 64) {
    const int32_t i = base + lane;
    const int rows = min(64, end - base);
    if (i < end) {
      const double d[3] = {cx - b.x[i], cy - b.y[i], cz - b.z[i]};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        double p = a == 0 ? b.q[i] : 1.0;
        tb[lane][a * P] = p;
#pragma unroll
        for (int e = 1; e < P; e++) {
          p = p * d[a] * kInv[e];
          tb[lane][a * P + e] = p;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    for (int r = 0; r < rows; r++) {
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] += tb[r][ox[k]] * tb[r][oy[k]] * tb[r][oz[k]];
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = lane + 64 * r;
    if (t < CN) M[(int64_t)ci * CN + t] += acc[r];
  }
}

// ---------------------------------------------------------------------------
// M2M: one block per group of 8 siblings (children [c0, c0 + 8) of one
// parent); thread t owns term t of the parent.  Each child's multipole and
// its shift powers (Xp - Xc)^m / m! (all terms) are staged in LDS; the
// products of a term come from the (k, n - k) table.
__global__ __launch_bounds__(256) void k_cart_m2m(Cells c, int32_t first_child, int32_t npar,
                                                  double *M) {
  critical_path_priority();
  __shared__ double Mc[8][CN];
  __shared__ double Pw[8][CN];
  __shared__ double pw[8][PWR];
  const int32_t g = blockIdx.x;
  if (g >= npar) return;
  const int32_t c0 = first_child + 8 * g;
  const int32_t par = c.parent[c0];
  for (int e = threadIdx.x; e < 8 * CN; e += 256) Mc[e / CN][e % CN] = M[(int64_t)c0 * CN + e];
  if (threadIdx.x < 24) {
    const int j = threadIdx.x / 3, a = threadIdx.x % 3;
    const double d = c.center[3 * (int64_t)par + a] - c.center[3 * (int64_t)(c0 + j) + a];
    double p = 1;
    pw[j][a * P] = p;
    for (int e = 1; e < P; e++) {
      p = p * d * kInv[e];
      pw[j][a * P + e] = p;
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < CN) {
    const int x = c_terms.x[t], y = c_terms.y[t], z = c_terms.z[t];
    for (int j = 0; j < 8; j++) Pw[j][t] = pw[j][x] * pw[j][P + y] * pw[j][2 * P + z];
  }
  __syncthreads();
  if (t >= CN) return;
  double s = 0;
  const int e1 = c_m2m.off[t + 1];
  for (int e = c_m2m.off[t]; e < e1; e++) {
    const int en = c_m2m.ent[e], k = en & 0xff, m = en >> 8;
#pragma unroll
    for (int j = 0; j < 8; j++) s += Pw[j][m] * Mc[j][k];
  }
  M[(int64_t)par * CN + t] += s;
}

// ---------------------------------------------------------------------------
// The M2L wave: state of one wave that sums the M2L of many (source, offset)
// pairs into one local expansion.
struct CartWave {
  uint32_t ent[CW / 2];
  uint32_t kb[CART_BINS];
  double acc[CART_BINS];
};

// Block-shared tables (loaded once per block, then __syncthreads).
struct CartShared {
  uint64_t rec[CN];
  double fact[CN];
};

__device__ inline void load_shared(CartShared &sh, const CartTables *tb) {
  for (int t = threadIdx.x; t < CN; t += blockDim.x) {
    sh.rec[t] = tb->rec[t];
    sh.fact[t] = tb->fact[t];
  }
}

// per wave: its lane's schedule, zero accumulators, the zero pads
__device__ inline void wave_init(CartWave &wv, const CartTables *tb, double *Ms, double *Cs,
                                 int lane) {
#pragma unroll
  for (int j = 0; j < CW / 2; j++) wv.ent[j] = tb->ent[j][lane];
#pragma unroll
  for (int b = 0; b < CART_BINS; b++) {
    wv.kb[b] = tb->kb[b][lane];
    wv.acc[b] = 0;
  }
  for (int t = CN + lane; t < MS_LEN; t += 64) Ms[t] = 0;
  if (lane == 0) Cs[ZS] = 0;
  __builtin_amdgcn_wave_barrier();
}

// D^k(1/r) at d (all CN terms) into Cs: a_0, then degree by degree (the
// terms of one degree, at most 55, one per lane), then the k! scaling
__device__ inline void wave_derivs(const CartShared &sh, double *Cs, double dx, double dy,
                                   double dz, int lane) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double invR2 = 1.0 / r2;
  if (lane == 0) Cs[0] = sqrt(invR2);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int d = 1; d < P; d++) {
    const int nt = (d + 1) * (d + 2) / 2;
    if (lane < nt) {
      const int t = cdeg_off(d) + lane;
      const uint64_t r = sh.rec[t];
      const double s1 = dx * Cs[r & 0xff] + dy * Cs[(r >> 8) & 0xff] + dz * Cs[(r >> 16) & 0xff];
      const double s2 = Cs[(r >> 24) & 0xff] + Cs[(r >> 32) & 0xff] + Cs[(r >> 40) & 0xff];
      Cs[t] = ((double)(1 - 2 * d) / d * s1 + (double)(1 - d) / d * s2) * invR2;
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = lane + 64 * r;
    if (t < CN) Cs[t] *= sh.fact[t];
  }
  __builtin_amdgcn_wave_barrier();
}

// acc[b] += sum over the lane's products of bin b: Ms[kb + i] * Cs[off]
__device__ inline void wave_contract(CartWave &wv, const double *Ms, const double *Cs) {
  const char *mb = reinterpret_cast<const char *>(Ms);
  const char *cb = reinterpret_cast<const char *>(Cs);
  int j = 0;
#pragma unroll
  for (int b = 0; b < CART_BINS; b++) {
    const double *m = reinterpret_cast<const double *>(mb + wv.kb[b]);
    double a = wv.acc[b];
#pragma unroll
    for (int i = 0; i < BL[b]; i++, j++) {
      const uint32_t w = wv.ent[j >> 1];
      const uint32_t off = (j & 1) ? (w >> 16) : (w & 0xffffu);
      a = fma(m[i], *reinterpret_cast<const double *>(cb + off), a);
    }
    wv.acc[b] = a;
  }
}

// the pieces' sums into L (out: CN terms), in the fixed piece order
__device__ inline void wave_finish(const CartWave &wv, const CartTables *tb, double *part,
                                   double *out, int lane) {
#pragma unroll
  for (int b = 0; b < CART_BINS; b++) part[b * 64 + lane] = wv.acc[b];
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = lane + 64 * r;
    if (t >= CN) continue;
    const uint32_t pc = tb->piece[t];
    const int first = (int)(pc & 0xffffu), cnt = (int)(pc >> 16);
    double s = 0;
    for (int p = 0; p < cnt; p++) s += part[tb->slot[first + p]];
    out[t] += s;
  }
}

constexpr int MW = 4;  // M2L waves (target cells) per block

// M2L over a per-target CSR list (source words: cell | image << IMG_SHIFT)
__global__ __launch_bounds__(64 * MW) void k_cart_m2l(Cells c, Csr list, int32_t ncells,
                                                      const double *M, double *L,
                                                      const CartTables *tb, double cycle) {
  critical_path_priority();
  __shared__ CartShared sh;
  __shared__ double sMs[MW][MS_LEN];
  __shared__ double sCs[MW][CN + 4];
  __shared__ double spart[MW][NSLOT];
  load_shared(sh, tb);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ti = blockIdx.x * MW + w;
  if (ti >= ncells) return;  // per wave from here on
  const int32_t beg = list.off[ti], end = list.off[ti + 1];
  if (beg >= end) return;
  double *Ms = sMs[w], *Cs = sCs[w];
  CartWave wv;
  wave_init(wv, tb, Ms, Cs, lane);
  const double tx = c.center[3 * (int64_t)ti], ty = c.center[3 * (int64_t)ti + 1],
               tz = c.center[3 * (int64_t)ti + 2];
  for (int32_t e = beg; e < end; e++) {
    const uint32_t sw = (uint32_t)list.src[e];
    const int32_t sj = (int32_t)(sw & CELL_MASK);
    double shx, shy, shz;
    image_shift(sw >> IMG_SHIFT, cycle, shx, shy, shz);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = lane + 64 * r;
      if (t < CN) Ms[t] = M[(int64_t)sj * CN + t];
    }
    const double dx = tx - c.center[3 * (int64_t)sj] - shx;
    const double dy = ty - c.center[3 * (int64_t)sj + 1] - shy;
    const double dz = tz - c.center[3 * (int64_t)sj + 2] - shz;
    wave_derivs(sh, Cs, dx, dy, dz, lane);
    wave_contract(wv, Ms, Cs);
    __builtin_amdgcn_wave_barrier();
  }
  wave_finish(wv, tb, spart[w], L + (int64_t)ti * CN, lane);
}

// ---------------------------------------------------------------------------
// Basis change (cart_sched.h): the Cartesian M2L computed exactly as
// T2 (spherical M2L (T1 M)), the spherical M2L being the library's MFMA
// tile kernel.  One wave per cell; the source expansion staged in LDS.
constexpr int BW = 4;

// spherical multipoles (NTERM complex per cell) from Cartesian ones
__global__ __launch_bounds__(64 * BW) void k_cart_to_sph(int32_t ncells, const double *Mc,
                                                         double *Ms, const BasisTables *bt) {
  critical_path_priority();
  __shared__ double sm[BW][CN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ci = blockIdx.x * BW + w;
  if (ci >= ncells) return;
  for (int t = lane; t < CN; t += 64) sm[w][t] = Mc[(int64_t)ci * CN + t];
  __builtin_amdgcn_wave_barrier();
  if (lane >= SPH_TERMS) return;
  double re = 0, im = 0;
  for (uint32_t e = bt->t1_off[lane]; e < bt->t1_off[lane + 1]; e++) {
    const double v = sm[w][bt->t1_k[e]];
    re += bt->t1_re[e] * v;
    im += bt->t1_im[e] * v;
  }
  Ms[((int64_t)ci * SPH_TERMS + lane) * 2] = re;
  Ms[((int64_t)ci * SPH_TERMS + lane) * 2 + 1] = im;
}

// Cartesian locals from spherical ones
__global__ __launch_bounds__(64 * BW) void k_sph_to_cart(int32_t ncells, const double *Ls,
                                                         double *Lc, const BasisTables *bt) {
  critical_path_priority();
  __shared__ double sl[BW][2 * SPH_TERMS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ci = blockIdx.x * BW + w;
  if (ci >= ncells) return;
  for (int t = lane; t < 2 * SPH_TERMS; t += 64) sl[w][t] = Ls[(int64_t)ci * 2 * SPH_TERMS + t];
  __builtin_amdgcn_wave_barrier();
  for (int k = lane; k < CN; k += 64) {
    double s = 0;
    for (uint32_t e = bt->t2_off[k]; e < bt->t2_off[k + 1]; e++) {
      const int t = (int)bt->t2_t[e];
      s += bt->t2_a[e] * sl[w][2 * t] + bt->t2_b[e] * sl[w][2 * t + 1];
    }
    Lc[(int64_t)ci * CN + k] = s;
  }
}

// ---------------------------------------------------------------------------
// L2L: one block per child cell; thread t owns term t of the child.
__global__ __launch_bounds__(256) void k_cart_l2l(Cells c, int32_t first_child, int32_t nchild,
                                                  double *L) {
  __shared__ double Lp[CN];
  __shared__ double pw[PWR];
  const int32_t ch = first_child + (int32_t)blockIdx.x;
  if ((int32_t)blockIdx.x >= nchild) return;
  const int32_t par = c.parent[ch];
  for (int t = threadIdx.x; t < CN; t += 256) Lp[t] = L[(int64_t)par * CN + t];
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    const double d = c.center[3 * (int64_t)ch + a] - c.center[3 * (int64_t)par + a];
    double p = 1;
    pw[a * P] = p;
    for (int e = 1; e < P; e++) {
      p = p * d * kInv[e];
      pw[a * P + e] = p;
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= CN) return;
  const int nx = c_terms.x[t], ny = c_terms.y[t], nz = c_terms.z[t];
  const int left = P - 1 - (nx + ny + nz);
  double s = 0;
  for (int kx = 0; kx <= left; kx++)
    for (int ky = 0; ky <= left - kx; ky++)
      for (int kz = 0; kz <= left - kx - ky; kz++)
        s += pw[kx] * pw[P + ky] * pw[2 * P + kz] * Lp[cidx(nx + kx, ny + ky, nz + kz)];
  L[(int64_t)ch * CN + t] += s;
}

// ---------------------------------------------------------------------------
// L2P: one wave per leaf, one body per lane (chunks of 64); the leaf's
// expansion staged in LDS (every lane reads the same term at once), every
// term unrolled at compile time.
template <int T>
__device__ inline void l2p_term(const double *Lc, const double (&px)[P], const double (&py)[P],
                                const double (&pz)[P], double (&s)[4]) {
  constexpr int x = term_part(T, 0), y = term_part(T, 1), z = term_part(T, 2);
  const double cc = px[x] * py[y] * pz[z];
  s[0] += cc * Lc[T];
  if constexpr (x + y + z <= P - 2) {
    s[1] += cc * Lc[cidx(x + 1, y, z)];
    s[2] += cc * Lc[cidx(x, y + 1, z)];
    s[3] += cc * Lc[cidx(x, y, z + 1)];
  }
}

template <int... T>
__device__ inline void l2p_all(std::integer_sequence<int, T...>, const double *Lc,
                               const double (&px)[P], const double (&py)[P], const double (&pz)[P],
                               double (&s)[4]) {
  (l2p_term<T>(Lc, px, py, pz, s), ...);
}

constexpr int LW = 4;  // L2P waves (leaves) per block

__global__ __launch_bounds__(64 * LW) void k_cart_l2p(Cells c, int32_t ncells, Bodies b,
                                                      const double *L, const double *p2p,
                                                      const int32_t *perm, double *trg,
                                                      double *out) {
  __shared__ double sL[LW][CN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t ci = blockIdx.x * LW + w;
  if (ci >= ncells || c.child[ci] >= 0 || c.nb[ci] == 0) return;  // per wave; no block barrier
  double *Ls = sL[w];
  for (int t = lane; t < CN; t += 64) Ls[t] = L[(int64_t)ci * CN + t];
  __builtin_amdgcn_wave_barrier();
  const double cx = c.center[3 * (int64_t)ci], cy = c.center[3 * (int64_t)ci + 1],
               cz = c.center[3 * (int64_t)ci + 2];
  const int32_t beg = c.beg[ci], end = beg + c.nb[ci];
  for (int32_t i = beg + lane; i - lane < end; i += 64) {
    // the expansion's LDS reads are the same in every chunk: without this
    // barrier they are hoisted out of the loop, hundreds of values held live
    asm volatile("" ::: "memory");
    const bool on = i < end;
    const int32_t k = on ? i : beg;
    const double d[3] = {b.x[k] - cx, b.y[k] - cy, b.z[k] - cz};
    double px[P], py[P], pz[P];
    px[0] = py[0] = pz[0] = 1;
#pragma unroll
    for (int e = 1; e < P; e++) {
      px[e] = px[e - 1] * d[0] * kInv[e];
      py[e] = py[e - 1] * d[1] * kInv[e];
      pz[e] = pz[e - 1] * d[2] * kInv[e];
    }
    double s[4] = {0, 0, 0, 0};
    l2p_all(std::make_integer_sequence<int, CN>{}, Ls, px, py, pz, s);
    if (!on) continue;
    const double q = b.q[i];
    double t[4];
#pragma unroll
    for (int e = 0; e < 4; e++) t[e] = p2p[4 * (int64_t)i + e] / q + s[e];
#pragma unroll
    for (int e = 0; e < 4; e++) trg[4 * (int64_t)i + e] = t[e];
    const int64_t o = perm[i];
#pragma unroll
    for (int e = 0; e < 4; e++) out[4 * o + e] = t[e];
  }
}

// ---------------------------------------------------------------------------
// The far field of the periodic images beyond the first shell into the
// root's local expansion (EX/include/traversal.h:126-185, in this basis):
// per sublevel the M2L of the block multipole from the 26 x 27 copies at
// (3 i + c) cycle, then the M2M of its 27 copies into the block three times
// larger (centre copy included).  One block: four waves take the copies in
// turn, their pieces summed wave by wave at the end.
__global__ __launch_bounds__(256) void k_cart_periodic_far(const double *Mroot, double *Lroot,
                                                           double cycle, int images,
                                                           const CartTables *tb) {
  __shared__ CartShared sh;
  __shared__ double sMs[MS_LEN];
  __shared__ double sCs[4][CN + 4];
  __shared__ double spart[4][NSLOT];
  __shared__ double Mn[CN];
  __shared__ double Lsum[4][CN];
  __shared__ double pw[PWR];
  load_shared(sh, tb);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < MS_LEN; t += 256) sMs[t] = t < CN ? Mroot[t] : 0.0;
  for (int t = threadIdx.x; t < 4 * CN; t += 256) Lsum[t / CN][t % CN] = 0;
  __syncthreads();
  CartWave wv;
  for (int level = 0; level < images - 1; level++) {
    wave_init(wv, tb, sMs, sCs[w], lane);
    __syncthreads();  // wave_init zeroes the pads of the shared multipole
    for (int k = w; k < 26 * 27; k += 4) {
      int i = k / 27, cc = k % 27;
      if (i >= 13) i++;  // skip the centre block (0, 0, 0)
      const int ix = i / 9 - 1, iy = (i / 3) % 3 - 1, iz = i % 3 - 1;
      const int cx = cc / 9 - 1, cy = (cc / 3) % 3 - 1, cz = cc % 3 - 1;
      const double dx = -((ix * 3 + cx) * cycle), dy = -((iy * 3 + cy) * cycle),
                   dz = -((iz * 3 + cz) * cycle);
      wave_derivs(sh, sCs[w], dx, dy, dz, lane);
      wave_contract(wv, sMs, sCs[w]);
      __builtin_amdgcn_wave_barrier();
    }
    wave_finish(wv, tb, spart[w], Lsum[w], lane);
    __syncthreads();
    // M2M of the 27 copies (dX = -(i cycle)) into the next block multipole
    const int t = threadIdx.x;
    double s = 0;
    int nx = 0, ny = 0, nz = 0;
    if (t < CN) {
      nx = c_terms.x[t];
      ny = c_terms.y[t];
      nz = c_terms.z[t];
    }
    for (int i = 0; i < 27; i++) {
      const double dd[3] = {-((i / 9 - 1) * cycle), -(((i / 3) % 3 - 1) * cycle),
                            -((i % 3 - 1) * cycle)};
      __syncthreads();
      if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        double p = 1;
        pw[a * P] = p;
        for (int e = 1; e < P; e++) {
          p = p * dd[a] * kInv[e];
          pw[a * P + e] = p;
        }
      }
      __syncthreads();
      if (t < CN)
        for (int kx = 0; kx <= nx; kx++)
          for (int ky = 0; ky <= ny; ky++)
            for (int kz = 0; kz <= nz; kz++)
              s += pw[nx - kx] * pw[P + ny - ky] * pw[2 * P + nz - kz] * sMs[cidx(kx, ky, kz)];
    }
    if (t < CN) Mn[t] = s;
    __syncthreads();
    if (t < CN) sMs[t] = Mn[t];
    __syncthreads();
    cycle *= 3;
  }
  // the waves' sums in wave order
  for (int t = threadIdx.x; t < CN; t += 256)
    Lroot[t] += ((Lsum[0][t] + Lsum[1][t]) + Lsum[2][t]) + Lsum[3][t];
}

// the device copy of both table sets
struct DeviceTables {
  CartTables sched;
  BasisTables basis;
};

const CartTables *sched_of(const void *p) { return &static_cast<const DeviceTables *>(p)->sched; }
const BasisTables *basis_of(const void *p) { return &static_cast<const DeviceTables *>(p)->basis; }

}  // namespace

size_t cart_tables_bytes() { return sizeof(DeviceTables); }

hipError_t cart_upload_tables(void *dst) {
  static DeviceTables host;
  static int state = 0;  // 0 not built, 1 built, -1 failed
  if (state == 0)
    state = cart::build_tables(host.sched) && cart::build_basis_tables(host.basis) ? 1 : -1;
  if (state < 0) return hipErrorInvalidValue;
  return hipMemcpy(dst, &host, sizeof(host), hipMemcpyHostToDevice);
}

hipError_t launch_cart_to_sph(int32_t ncells, const double *Mc, double *Ms, const void *tables,
                              hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_cart_to_sph<<<(ncells + BW - 1) / BW, 64 * BW, 0, s>>>(ncells, Mc, Ms, basis_of(tables));
  return hipGetLastError();
}

hipError_t launch_sph_to_cart(int32_t ncells, const double *Ls, double *Lc, const void *tables,
                              hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_sph_to_cart<<<(ncells + BW - 1) / BW, 64 * BW, 0, s>>>(ncells, Ls, Lc, basis_of(tables));
  return hipGetLastError();
}

hipError_t launch_cart_p2m(Cells c, int32_t ncells, Bodies b, double *M, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_cart_p2m<<<(ncells + PW - 1) / PW, 64 * PW, 0, s>>>(c, ncells, b, M);
  return hipGetLastError();
}

hipError_t launch_cart_m2m(Cells c, int32_t first_child, int32_t npar, double *M, hipStream_t s) {
  if (npar <= 0) return hipSuccess;
  k_cart_m2m<<<npar, 256, 0, s>>>(c, first_child, npar, M);
  return hipGetLastError();
}

hipError_t launch_cart_m2l(Cells c, Csr list, int32_t ncells, const double *M, double *L,
                           const void *tables, double cycle, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_cart_m2l<<<(ncells + MW - 1) / MW, 64 * MW, 0, s>>>(c, list, ncells, M, L,
                                                       sched_of(tables), cycle);
  return hipGetLastError();
}

hipError_t launch_cart_l2l(Cells c, int32_t first_child, int32_t nchild, double *L,
                           hipStream_t s) {
  if (nchild <= 0) return hipSuccess;
  k_cart_l2l<<<nchild, 256, 0, s>>>(c, first_child, nchild, L);
  return hipGetLastError();
}

hipError_t launch_cart_l2p(Cells c, int32_t ncells, Bodies b, const double *L, const double *p2p,
                           const int32_t *perm, double *trg, double *out, hipStream_t s) {
  if (ncells <= 0 || b.n <= 0) return hipSuccess;
  k_cart_l2p<<<(ncells + LW - 1) / LW, 64 * LW, 0, s>>>(c, ncells, b, L, p2p, perm, trg, out);
  return hipGetLastError();
}

hipError_t launch_cart_periodic_far(const double *Mroot, double *Lroot, double cycle, int images,
                                    const void *tables, hipStream_t s) {
  if (images < 2 || !(cycle > 0)) return hipSuccess;
  k_cart_periodic_far<<<1, 256, 0, s>>>(Mroot, Lroot, cycle, images, sched_of(tables));
  return hipGetLastError();
}

}  // namespace tfmm
