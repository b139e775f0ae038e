// cart_sched.h -- the Cartesian basis's term order and the fixed schedule of
// its M2L contraction (cartesian.hip), host and device alike: the kernels
// take the tables built here; tests/native/cart_schedule_check.cpp applies
// them on the CPU as the kernel does.
//
// Term k = (kx, ky, kz) of degree d = kx + ky + kz sits at
//   cidx(k) = d (d+1) (d+2) / 6 + (ky + kz) (ky + kz + 1) / 2 + kz,
// the order of the reference's Index<nx,ny,nz>::I
// (sample/exafmm-dev-13274dd4ac68/kernels/LaplaceCartesianCPU.cxx:3-25).
//
// The M2L contraction L_n += sum_{|k| <= P-1-|n|} C_{n+k} M_k has 5005
// products per pair, from 220 for n = 0 down to one for each n of degree
// P-1.  They are dealt to the 64 lanes of a wave in NBINS bins of fixed
// lengths BL, the same for every lane; (lane, bin) holds a run of
// consecutive k of one output term (a piece: a long term is cut into
// several) or nothing.  Placement: terms by falling product count; a term's
// rest goes to the smallest free (lane, bin) that holds it, else the
// largest free one takes as much as it holds.  Product i of a piece reads
// M at (its first k) + i, and C at the offset stored for it (a padding
// product reads the zero slot ZS of C).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <tuple>
#include <utility>
#include <vector>

#ifdef __HIPCC__
#define CART_HD __host__ __device__
#else
#define CART_HD
#endif

namespace tfmm {
namespace cart {

constexpr int ORDER = 10;  // terms of degree < ORDER
constexpr int CN = ORDER * (ORDER + 1) * (ORDER + 2) / 6;
constexpr int NBINS = 6;
constexpr int ZS = CN;  // index of the zero slot after the terms
constexpr int BL[NBINS] = {29, 21, 15, 10, 7, 2};
constexpr int cart_width() {
  int w = 0;
  for (int b = 0; b < NBINS; b++) w += BL[b];
  return w;
}
constexpr int CW = cart_width();  // products per lane per pair
static_assert(CW % 2 == 0, "two C offsets per word");
constexpr int NSLOT = NBINS * 64;
constexpr int MS_LEN = 256;  // source expansion in LDS: terms, then zeros
static_assert(CN + 29 <= MS_LEN, "a padded piece reads past the terms into zeros");

CART_HD constexpr int cdeg_off(int d) { return d * (d + 1) * (d + 2) / 6; }
CART_HD constexpr int cidx(int x, int y, int z) {
  return cdeg_off(x + y + z) + (y + z) * (y + z + 1) / 2 + z;
}
// multi-index of term t (x, y, z)
CART_HD constexpr int term_part(int t, int which) {
  for (int d = 0; d < ORDER; d++)
    for (int x = d; x >= 0; x--)
      for (int y = d - x; y >= 0; y--)
        if (cidx(x, y, d - x - y) == t) return which == 0 ? x : which == 1 ? y : d - x - y;
  return -1;
}

// The tables the kernels read (one device copy per context).
struct CartTables {
  uint32_t ent[CW / 2][64];  // C byte offsets of a lane's products, two per word
  uint32_t kb[NBINS][64];  // byte offset of the first M term of each (bin, lane)
  uint32_t piece[CN];          // first slot-list entry | piece count << 16
  uint32_t slot[NSLOT];        // pieces of term 0, then of term 1, ...: bin * 64 + lane
  uint64_t rec[CN];            // recurrence: indices of k-e_x,y,z then k-2e_x,y,z (ZS: none)
  double fact[CN];             // kx! ky! kz!
};

// products of output term t in a pair
inline int cart_count(int t) {
  const int left = ORDER - 1 - (term_part(t, 0) + term_part(t, 1) + term_part(t, 2));
  return (left + 1) * (left + 2) * (left + 3) / 6;
}

inline bool build_tables(CartTables &tb) {
  struct Slot {
    int cap, lane, bin;
    bool operator<(const Slot &o) const {
      return std::tie(cap, lane, bin) < std::tie(o.cap, o.lane, o.bin);
    }
  };
  std::vector<Slot> free_slots;
  for (int l = 0; l < 64; l++)
    for (int b = 0; b < NBINS; b++) free_slots.push_back({BL[b], l, b});
  std::vector<int> order(CN);
  for (int t = 0; t < CN; t++) order[t] = t;
  std::stable_sort(order.begin(), order.end(),
                   [](int a, int b) { return cart_count(a) > cart_count(b); });
  // assigned piece of every (bin, lane): term, first k, length (term -1: none)
  std::vector<int> pt(NSLOT, -1), pk(NSLOT, 0), pl(NSLOT, 0);
  for (int t : order) {
    int rem = cart_count(t), kb = 0;
    while (rem > 0) {
      if (free_slots.empty()) return false;
      auto best = free_slots.end();
      for (auto it = free_slots.begin(); it != free_slots.end(); ++it)
        if (it->cap >= rem && (best == free_slots.end() || *it < *best)) best = it;
      if (best == free_slots.end()) best = std::max_element(free_slots.begin(), free_slots.end());
      const Slot s = *best;
      free_slots.erase(best);
      const int take = std::min(rem, s.cap);
      const int id = s.bin * 64 + s.lane;
      pt[id] = t;
      pk[id] = kb;
      pl[id] = take;
      kb += take;
      rem -= take;
    }
  }
  // products of every lane in bin order
  for (int l = 0; l < 64; l++) {
    int j = 0;
    uint16_t off[CW];
    for (int b = 0; b < NBINS; b++) {
      const int id = b * 64 + l;
      tb.kb[b][l] = pt[id] >= 0 ? (uint32_t)pk[id] * 8u : 0u;
      for (int i = 0; i < BL[b]; i++, j++) {
        int ci = ZS;
        if (pt[id] >= 0 && i < pl[id]) {
          const int t = pt[id], k = pk[id] + i;
          ci = cidx(term_part(t, 0) + term_part(k, 0), term_part(t, 1) + term_part(k, 1),
                    term_part(t, 2) + term_part(k, 2));
          if (ci >= CN) return false;
        }
        off[j] = (uint16_t)(ci * 8);
      }
    }
    for (int q = 0; q < CW / 2; q++) tb.ent[q][l] = (uint32_t)off[2 * q] | (uint32_t)off[2 * q + 1] << 16;
  }
  // slot lists: each term's pieces in k order
  int n = 0;
  for (int t = 0; t < CN; t++) {
    std::vector<std::pair<int, int>> ps;
    int covered = 0;
    for (int id = 0; id < NSLOT; id++)
      if (pt[id] == t) {
        ps.push_back({pk[id], id});
        covered += pl[id];
      }
    if (covered != cart_count(t)) return false;
    std::sort(ps.begin(), ps.end());
    tb.piece[t] = (uint32_t)n | (uint32_t)ps.size() << 16;
    for (auto &p : ps) tb.slot[n++] = (uint32_t)p.second;
  }
  for (; n < NSLOT; n++) tb.slot[n] = 0;
  // the recurrence's indices and the factorials
  for (int t = 0; t < CN; t++) {
    const int k[3] = {term_part(t, 0), term_part(t, 1), term_part(t, 2)};
    uint64_t r = 0;
    for (int a = 0; a < 3; a++) {
      int m1 = ZS, m2 = ZS;
      int km[3] = {k[0], k[1], k[2]};
      if (k[a] >= 1) {
        km[a] -= 1;
        m1 = cidx(km[0], km[1], km[2]);
      }
      if (k[a] >= 2) {
        km[a] -= 1;
        m2 = cidx(km[0], km[1], km[2]);
      }
      r |= (uint64_t)m1 << (8 * a);
      r |= (uint64_t)m2 << (8 * (a + 3));
    }
    tb.rec[t] = r;
    double f = 1;
    for (int a = 0; a < 3; a++)
      for (int e = 2; e <= k[a]; e++) f *= e;
    tb.fact[t] = f;
  }
  return true;
}


}  // namespace cart
}  // namespace tfmm
