// sort.hip -- device-wide exclusive scan and stable LSD radix sort of
// (uint64 key, uint32 value) pairs for gfx950 (wave64).
//
// These replace the host std::sort of the reference tree build
// (cpp/include/tapas/common.h:248-273, TP/hot/buildtree.h:299-324); the
// reference sort is unstable, so only the key order is specified and equal
// keys may come out in any order (this sort keeps input order for them).
#include <atomic>

#include "tfmm_internal.h"

namespace tfmm {
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ROWS = 16;  // keys per lane
constexpr int RS_TILE = RS_THREADS * RS_ROWS;
constexpr int RS_MAXBITS = 9;
constexpr int RS_BINS = 1 << RS_MAXBITS;

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v, lane);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int w = 0; w < SCAN_THREADS / 64; w++) {
    if (w < wid) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t *in, int64_t n,
                                                              uint32_t *sums) {
  __builtin_amdgcn_s_setprio(2);  // as the kernels beside the early P2P (critical_path_priority)
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    const int64_t j = base + i;
    if (j < n) s += in[j];
  }
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t *in, int64_t n,
                                                            const uint32_t *block_off,
                                                            uint32_t *out) {
  __builtin_amdgcn_s_setprio(2);  // as the kernels beside the early P2P (critical_path_priority)
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    const int64_t j = base + i;
    v[i] = (j < n) ? in[j] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot) + (block_off ? block_off[blockIdx.x] : 0u);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    const int64_t j = base + i;
    if (j < n) out[j] = run;
    run += v[i];
  }
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
// LSD passes of up to RS_MAXBITS bits; a pass = per-tile digit histogram,
// device-wide exclusive scan of the (digit, tile) counts, and a stable
// scatter that ranks each tile in LDS (wave ballots), stages it in digit
// order and writes every digit run to consecutive positions.

template <class K>
__device__ inline uint32_t digit_of(K k, int shift, uint32_t mask) {
  return (uint32_t)(k >> shift) & mask;
}

template <class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const K *__restrict__ keys, int64_t n,
                                                           int shift,
                                                           uint32_t mask, uint32_t *counts,
                                                           int nblocks, const uint32_t *dn) {
  __builtin_amdgcn_s_setprio(2);  // as the kernels beside the early P2P (critical_path_priority)
  __shared__ uint32_t hist[RS_BINS];
  if (dn) n = min(n, (int64_t)*dn);  // the count on the device, at most the capacity n
  for (int i = threadIdx.x; i < RS_BINS; i += RS_THREADS) hist[i] = 0;
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * RS_TILE;
  for (int i = threadIdx.x; i < RS_TILE; i += RS_THREADS) {
    const int64_t j = tile + i;
    if (j < n) atomicAdd(&hist[digit_of(keys[j], shift, mask)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d <= (int)mask; d += RS_THREADS)
    counts[(int64_t)d * nblocks + blockIdx.x] = hist[d];
}

template <class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const K *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, K *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, int64_t n, int shift, int nbits,
    const uint32_t *__restrict__ offsets, int nblocks, const uint32_t *dn) {
  __builtin_amdgcn_s_setprio(2);  // as the kernels beside the early P2P (critical_path_priority)
  // The wave histograms are dead once every key knows its place in the tile,
  // so the tile's staging arrays reuse their LDS (more blocks per CU).
  constexpr size_t HIST_BYTES = sizeof(uint32_t) * RS_WAVES * RS_BINS;
  constexpr size_t STAGE_BYTES = (sizeof(K) + sizeof(uint32_t)) * RS_TILE;
  __shared__ __align__(16) unsigned char pool[HIST_BYTES > STAGE_BYTES ? HIST_BYTES : STAGE_BYTES];
  uint32_t (*whist)[RS_BINS] = reinterpret_cast<uint32_t (*)[RS_BINS]>(pool);
  K *sk = reinterpret_cast<K *>(pool);
  uint32_t *sv = reinterpret_cast<uint32_t *>(pool + sizeof(K) * RS_TILE);
  __shared__ uint32_t dstart[RS_BINS];
  __shared__ uint32_t gshift[RS_BINS];  // global position of the tile's run of digit d, less dstart[d]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t mask = (1u << nbits) - 1u;
  if (dn) n = min(n, (int64_t)*dn);
  for (int i = lane; i < RS_BINS; i += 64) whist[wid][i] = 0;
  __syncthreads();

  const int64_t wbase = (int64_t)blockIdx.x * RS_TILE + (int64_t)wid * (64 * RS_ROWS);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  K k[RS_ROWS];
  uint32_t v[RS_ROWS];
  uint32_t rank[RS_ROWS];
#pragma unroll
  for (int r = 0; r < RS_ROWS; r++) {
    const int64_t j = wbase + r * 64 + lane;
    const bool valid = j < n;
    k[r] = valid ? keys_in[j] : (K)0;
    v[r] = valid ? vals_in[j] : 0u;
    const uint32_t d = digit_of(k[r], shift, mask);
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; b++) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t before = (uint32_t)__popcll(m & lt);
    uint32_t base = 0;
    if (valid) base = whist[wid][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) whist[wid][d] = base + (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
    rank[r] = base + before;
  }
  __syncthreads();
  // per digit: offsets of the waves (exclusive, in wave order) and the tile's
  // count; then the tile's digit starts (exclusive scan over digits).  Each
  // thread owns RS_BINS / RS_THREADS consecutive digits.
  constexpr int DPT = RS_BINS / RS_THREADS;
  uint32_t cnt[DPT];
  uint32_t mine = 0;
#pragma unroll
  for (int u = 0; u < DPT; u++) {
    const int d = threadIdx.x * DPT + u;
    cnt[u] = 0;
    for (int w = 0; w < RS_WAVES; w++) {
      const uint32_t t = whist[w][d];
      whist[w][d] = cnt[u];
      cnt[u] += t;
    }
    mine += cnt[u];
  }
  uint32_t unused;
  uint32_t start = block_excl_scan(mine, &unused);
#pragma unroll
  for (int u = 0; u < DPT; u++) {
    const int d = threadIdx.x * DPT + u;
    dstart[d] = start;
    // the digit's global offsets, read once here so the write loop below
    // waits on no global load
    if (d <= (int)mask) gshift[d] = offsets[(int64_t)d * nblocks + blockIdx.x] - start;
    start += cnt[u];
  }
  __syncthreads();
  // each key's place in the tile (digit order, stable); then, with the wave
  // histograms no longer read, stage the tile in LDS in that order and write
  // each digit's run to consecutive global positions
#pragma unroll
  for (int r = 0; r < RS_ROWS; r++) {
    const uint32_t d = digit_of(k[r], shift, mask);
    rank[r] += dstart[d] + whist[wid][d];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ROWS; r++) {
    const int64_t j = wbase + r * 64 + lane;
    if (j < n) {
      sk[rank[r]] = k[r];
      sv[rank[r]] = v[r];
    }
  }
  __syncthreads();
  const int64_t tbase = (int64_t)blockIdx.x * RS_TILE;
  const int tn = (int)((n - tbase) < RS_TILE ? (n - tbase) : RS_TILE);  // <= 0 past the count
  for (int i = threadIdx.x; i < tn; i += RS_THREADS) {
    const K key = sk[i];
    const uint32_t d = digit_of(key, shift, mask);
    const uint32_t pos = gshift[d] + (uint32_t)i;
    keys_out[pos] = key;
    vals_out[pos] = sv[i];
  }
}

}  // namespace

// Device bytes held by the library's buffers (all contexts), and the limit
// TFMM_TUNE_DEVICE_LIMIT sets on them (0: none).
std::atomic<size_t> g_held{0}, g_limit{0};

bool dev_fits(size_t release, size_t want) {
  const size_t lim = g_limit.load();
  return lim == 0 || g_held.load() - release + want <= lim;
}

void dev_account(size_t freed, size_t allocated) {
  g_held -= freed;
  g_held += allocated;
}

void dev_set_limit(size_t bytes) { g_limit = bytes; }
size_t dev_limit() { return g_limit.load(); }

hipError_t dev_reserve(DevBuf &b, size_t bytes) {
  if (bytes <= b.cap) return hipSuccess;
  size_t want = b.cap ? b.cap : 256;
  while (want < bytes) want *= 2;
  if (!dev_fits(b.cap, want)) return hipErrorOutOfMemory;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    if (e != hipSuccess) return e;
    dev_account(b.cap, 0);
    b.ptr = nullptr;
    b.cap = 0;
  }
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e == hipSuccess) {
    b.cap = want;
    dev_account(0, want);
  }
  return e;
}

void dev_release(DevBuf &b) {
  if (b.ptr && hipFree(b.ptr) == hipSuccess) dev_account(b.cap, 0);
  b.ptr = nullptr;
  b.cap = 0;
}

size_t scan_tmp_bytes(int64_t n) {
  size_t total = 0;
  int64_t m = n;
  while (m > SCAN_TILE) {
    m = ceil_div(m, SCAN_TILE);
    total += (size_t)m * 2 * sizeof(uint32_t) + 256;
  }
  return total + 256;
}

hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, void *tmp,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = ceil_div(n, SCAN_TILE);
  if (nb == 1) {
    k_scan_down<<<1, SCAN_THREADS, 0, s>>>(in, n, nullptr, out);
    return hipGetLastError();
  }
  uint32_t *sums = reinterpret_cast<uint32_t *>(tmp);
  uint32_t *sums_scan = sums + nb;
  char *next = reinterpret_cast<char *>(sums_scan + nb);
  next = reinterpret_cast<char *>(((uintptr_t)next + 255) & ~(uintptr_t)255);
  k_scan_reduce<<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, sums);
  hipError_t e = scan_exclusive_u32(sums, sums_scan, nb, next, s);
  if (e != hipSuccess) return e;
  k_scan_down<<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, sums_scan, out);
  return hipGetLastError();
}

size_t radix_tmp_bytes(int64_t n) {
  const int64_t nb = ceil_div(n > 0 ? n : 1, RS_TILE);
  const int64_t ncount = nb * RS_BINS;
  return 2 * (size_t)ncount * sizeof(uint32_t) + scan_tmp_bytes(ncount) + 512;
}

// Sorts the bits [bit_lo, bit_hi) of the keys with as few passes of at most
// RS_MAXBITS bits as possible, digit widths balanced.
// With dn the number of pairs is read on the device (*dn, at most the
// capacity n): grids and passes depend on n and the bit range only.
template <class K>
static hipError_t radix_sort_impl(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt,
                                  int64_t n, int bit_lo, int bit_hi, void *tmp, hipStream_t s,
                                  bool *result_in_alt, const uint32_t *dn) {
  *result_in_alt = false;
  if ((n <= 1 && !dn) || n <= 0 || bit_hi <= bit_lo) return hipSuccess;
  const int nb = (int)ceil_div(n, RS_TILE);
  const int64_t ncount = (int64_t)nb * RS_BINS;
  uint32_t *counts = reinterpret_cast<uint32_t *>(tmp);
  uint32_t *offs = counts + ncount;
  char *stmp = reinterpret_cast<char *>(offs + ncount);
  stmp = reinterpret_cast<char *>(((uintptr_t)stmp + 255) & ~(uintptr_t)255);
  K *kin = keys, *kout = keys_alt;
  uint32_t *vin = vals, *vout = vals_alt;
  const int bits = bit_hi - bit_lo;
  const int passes = (bits + RS_MAXBITS - 1) / RS_MAXBITS;
  int lo = bit_lo;
  for (int ps = 0; ps < passes; ps++) {
    const int nbits = (bit_hi - lo + (passes - ps) - 1) / (passes - ps);  // balanced widths
    const uint32_t mask = (1u << nbits) - 1u;
    const int64_t nc = (int64_t)nb * (mask + 1);
    k_radix_hist<K><<<nb, RS_THREADS, 0, s>>>(kin, n, lo, mask, counts, nb, dn);
    hipError_t e = scan_exclusive_u32(counts, offs, nc, stmp, s);
    if (e != hipSuccess) return e;
    k_radix_scatter<K><<<nb, RS_THREADS, 0, s>>>(kin, vin, kout, vout, n, lo, nbits, offs, nb,
                                                  dn);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    K *tk = kin; kin = kout; kout = tk;
    uint32_t *tv = vin; vin = vout; vout = tv;
    *result_in_alt = !*result_in_alt;
    lo += nbits;
  }
  return hipSuccess;
}

hipError_t radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                            uint32_t *vals_alt, int64_t n, int bit_lo, int bit_hi, void *tmp,
                            hipStream_t s, bool *result_in_alt, const uint32_t *dn) {
  return radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, bit_lo, bit_hi, tmp, s,
                                   result_in_alt, dn);
}

hipError_t radix_sort_pairs32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                              uint32_t *vals_alt, int64_t n, int bit_lo, int bit_hi, void *tmp,
                              hipStream_t s, bool *result_in_alt, const uint32_t *dn) {
  return radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, bit_lo, bit_hi, tmp, s,
                                   result_in_alt, dn);
}

}  // namespace tfmm
