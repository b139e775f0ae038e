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


// ---------------------------------------------------------------------------
// Basis change between the Cartesian expansions and the spherical ones of
// the library (55 complex terms, (n, m) at n (n+1) / 2 + m, m >= 0).
// Both bases truncate the same double series at total degree ORDER - 1, and
// a contraction with derivatives of 1/r sees only the traceless part of a
// moment tensor, so the Cartesian M2L is exactly
//   L_cart = T2 (M2L_spherical (T1 M_cart))
// with T1, T2 block diagonal by degree:
//   M_sph[n,m] = sum_{|k|=n} conj(y_{nm,k}) (-1)^n k! M_cart[k]
//     (P2M stores q conj(Y_n^m(x - X)); M_cart[k] = sum q (X - x)^k / k!)
//   L_cart[k]  = k! sum_{m=0..n} w_m Re(L_sph[n,m] y_{nm,k}),  w_0 = 1, w_m = 2
//     (L2P sums w_m Re(L Y_n^m(e)); the Cartesian local is sum L e^k / k!)
// where y_{nm,k} are the coefficients of the regular solid harmonic
//   Y_n^m(e) = (-1)^n / (n+m)! rho^n P_n^m(cos alpha) e^{i m beta}
// (the normalisation of the library's harmonics recurrence,
// LaplaceSphericalCPU_tapas.h:43-100) as a polynomial in e:
//   R_m^m = (-1)^m (2m-1)!! (x + i y)^m,  R_{m+1}^m = (2m+1) z R_m^m,
//   R_n^m = ((2n-1) z R_{n-1}^m - (n+m-1) rho^2 R_{n-2}^m) / (n-m).
constexpr int SPH_TERMS = ORDER * (ORDER + 1) / 2;  // 55
constexpr int BASIS_ENTRIES = 1705;  // sum_n (n+1) (n+1) (n+2) / 2, either map

struct BasisTables {
  // T1: spherical term t from Cartesian terms t1_k[t1_off[t] .. t1_off[t+1])
  uint32_t t1_off[SPH_TERMS + 1];
  uint32_t t1_k[BASIS_ENTRIES];
  double t1_re[BASIS_ENTRIES], t1_im[BASIS_ENTRIES];
  // T2: Cartesian term k from spherical terms t2_t[t2_off[k] .. t2_off[k+1]]:
  //   L_cart[k] = sum a Re L_sph[t] + b Im L_sph[t]
  uint32_t t2_off[CN + 1];
  uint32_t t2_t[BASIS_ENTRIES];
  double t2_a[BASIS_ENTRIES], t2_b[BASIS_ENTRIES];
};

inline double factorial(int n) {
  double f = 1;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

// coefficients of Y_n^m over all Cartesian terms (zero off degree n)
inline void solid_harmonics(std::vector<std::vector<double>> &re,
                            std::vector<std::vector<double>> &im) {
  re.assign(SPH_TERMS, std::vector<double>(CN, 0.0));
  im.assign(SPH_TERMS, std::vector<double>(CN, 0.0));
  using Poly = std::pair<std::vector<double>, std::vector<double>>;  // re, im over terms
  auto zero = [] { return Poly(std::vector<double>(CN, 0.0), std::vector<double>(CN, 0.0)); };
  auto shift = [&](const Poly &p, int ax, int times) {  // p * e_ax^times
    Poly q = zero();
    for (int t = 0; t < CN; t++) {
      if (p.first[t] == 0 && p.second[t] == 0) continue;
      int e[3] = {term_part(t, 0), term_part(t, 1), term_part(t, 2)};
      e[ax] += times;
      if (e[0] + e[1] + e[2] >= ORDER) continue;
      const int u = cidx(e[0], e[1], e[2]);
      q.first[u] += p.first[t];
      q.second[u] += p.second[t];
    }
    return q;
  };
  for (int m = 0; m < ORDER; m++) {
    // R_m^m = (-1)^m (2m-1)!! (x + i y)^m
    Poly r = zero();
    r.first[0] = 1;
    for (int j = 0; j < m; j++) {  // multiply by (x + i y)
      const Poly px = shift(r, 0, 1), py = shift(r, 1, 1);
      for (int t = 0; t < CN; t++) {
        r.first[t] = px.first[t] - py.second[t];
        r.second[t] = px.second[t] + py.first[t];
      }
    }
    double df = 1;
    for (int j = 1; j <= 2 * m - 1; j += 2) df *= j;
    const double s = (m & 1 ? -1.0 : 1.0) * df;
    for (int t = 0; t < CN; t++) {
      r.first[t] *= s;
      r.second[t] *= s;
    }
    Poly prev = zero(), cur = r;  // R_{n-1}^m, R_n^m
    for (int n = m; n < ORDER; n++) {
      if (n > m) {
        const Poly zc = shift(cur, 2, 1);
        Poly nx = zero();
        const Poly rx = shift(prev, 0, 2), ry = shift(prev, 1, 2), rz = shift(prev, 2, 2);
        for (int t = 0; t < CN; t++) {
          const double r2re = rx.first[t] + ry.first[t] + rz.first[t];
          const double r2im = rx.second[t] + ry.second[t] + rz.second[t];
          nx.first[t] = ((2 * n - 1) * zc.first[t] - (n + m - 1) * r2re) / (n - m);
          nx.second[t] = ((2 * n - 1) * zc.second[t] - (n + m - 1) * r2im) / (n - m);
        }
        prev = cur;
        cur = nx;
      }
      const int st = n * (n + 1) / 2 + m;
      const double f = (n & 1 ? -1.0 : 1.0) / factorial(n + m);
      for (int t = 0; t < CN; t++) {
        re[st][t] = cur.first[t] * f;
        im[st][t] = cur.second[t] * f;
      }
    }
  }
}

inline bool build_basis_tables(BasisTables &bt) {
  std::vector<std::vector<double>> yre, yim;
  solid_harmonics(yre, yim);
  auto kfact = [](int t) {
    return factorial(term_part(t, 0)) * factorial(term_part(t, 1)) * factorial(term_part(t, 2));
  };
  auto degree = [](int t) { return term_part(t, 0) + term_part(t, 1) + term_part(t, 2); };
  int e = 0;
  for (int n = 0; n < ORDER; n++)
    for (int m = 0; m <= n; m++) {
      const int st = n * (n + 1) / 2 + m;
      bt.t1_off[st] = (uint32_t)e;
      for (int k = 0; k < CN; k++) {
        if (degree(k) != n) continue;
        if (e >= BASIS_ENTRIES) return false;
        const double s = (n & 1 ? -1.0 : 1.0) * kfact(k);
        bt.t1_k[e] = (uint32_t)k;
        bt.t1_re[e] = yre[st][k] * s;   // conj(y) (-1)^n k!
        bt.t1_im[e] = -yim[st][k] * s;
        e++;
      }
    }
  bt.t1_off[SPH_TERMS] = (uint32_t)e;
  if (e != BASIS_ENTRIES) return false;
  e = 0;
  for (int k = 0; k < CN; k++) {
    bt.t2_off[k] = (uint32_t)e;
    const int n = degree(k);
    for (int m = 0; m <= n; m++) {
      const int st = n * (n + 1) / 2 + m;
      if (e >= BASIS_ENTRIES) return false;
      const double w = (m == 0 ? 1.0 : 2.0) * kfact(k);
      bt.t2_t[e] = (uint32_t)st;
      bt.t2_a[e] = w * yre[st][k];    // Re(L y) = Re L Re y - Im L Im y
      bt.t2_b[e] = -w * yim[st][k];
      e++;
    }
  }
  bt.t2_off[CN] = (uint32_t)e;
  return e == BASIS_ENTRIES;
}

}  // namespace cart
}  // namespace tfmm
