// m2l_codes.h -- the sparsity pattern of the M2L operator as a real matrix,
// and the operand codes the batched M2L kernel reads it through.  Host-only
// C++ (no HIP), so that it can be checked on the CPU.
//
// The M2L of one cell pair (reference: LaplaceSphericalCPU_tapas.h:208-273,
// mutual = 0) maps the source's 55 complex multipole coefficients M_n^m onto
// the target's 55 local coefficients L_j^k:
//   L_j^k += sum_{n < P-j} sum_{|m| <= n} C(j,k,n,m) M_n^m Y_{j+n}^{m-k}
// with C = (-1)^j for m < 0 and (-1)^j (-1)^{(k-m)[k<m] + m} otherwise, the
// m < 0 coefficients being conj(M_n^{-m}) and Y_p^{-q} = conj(Y_p^q).
// On real and imaginary parts this is a 112 x 112 real matrix T (110 rows
// and columns plus 2 of padding), row r = 2 tri(j, k) + (0 re | 1 im),
// column c = 2 tri(n, |m|) + (0 re | 1 im).  Every entry is a combination of
// at most two stored harmonic components (flat index 2 tri(p, q) + re/im),
// each with a coefficient in {-2, ..., 2}; only 18 of the 49 16 x 16 blocks
// are non-zero.
//
// Operand codes: with a group's harmonics laid out at four multiples
// {+1, -1, +2, -2} (element m * 112 + i), the A-operand entry that lane l
// supplies at k-step kk of block b is the sum of two table elements whose
// byte offsets are packed into one 32-bit code (low 16 bits, high 16 bits).
// Absent terms point at element 110, a zero.  The batched M2L forms every
// group's A operands from these codes once, before the tile kernel.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace tfmm {
namespace m2lc {

constexpr int P = 10;
constexpr int NR = 112;
constexpr int NBLK = 18;
constexpr int NMULT = 4;  // table multiples +1, -1, +2, -2
constexpr int ZERO_ELEM = 110;

// non-zero 16 x 16 blocks (row block, column block) of T
constexpr int8_t BLK_RB[NBLK] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 5, 6};
constexpr int8_t BLK_CB[NBLK] = {0, 1, 2, 3, 4, 5, 6, 0, 1, 2, 3, 0, 1, 0, 1, 0, 0, 0};

inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// degree j of real row (or column) r; rows >= 110 are padding (degree 0)
inline int degree_of(int r) {
  if (r >= 110) return 0;
  const int t = r >> 1;
  int n = 0;
  while (tri(n + 1, 0) <= t) n++;
  return n;
}

// T entry: up to two (harmonic component, coefficient) terms
struct Ent {
  int i1 = -1, s1 = 0, i2 = -1, s2 = 0;
};

// T as entries; false if an entry needs more than two harmonic components or
// a non-zero entry falls outside the listed blocks.
inline bool build_entries(std::vector<Ent> &T) {
  T.assign((size_t)NR * NR, Ent());
  bool ok = true;
  auto add = [&](int r, int cc, int yi, int sg) {
    Ent &e = T[(size_t)r * NR + cc];
    if (e.i1 < 0) {
      e.i1 = yi;
      e.s1 = sg;
    } else if (e.i1 == yi) {
      e.s1 += sg;
    } else if (e.i2 < 0) {
      e.i2 = yi;
      e.s2 = sg;
    } else if (e.i2 == yi) {
      e.s2 += sg;
    } else {
      ok = false;
    }
  };
  auto oddeven = [](int n) { return (n & 1) ? -1 : 1; };
  for (int j = 0; j < P; j++) {
    for (int k = 0; k <= j; k++) {
      const int jk = tri(j, k);
      const int Cnm = oddeven(j);
      for (int n = 0; n < P - j; n++) {
        for (int m = -n; m <= n; m++) {
          const int p = j + n, qq = m - k;
          const int base = 2 * tri(p, qq >= 0 ? qq : -qq);
          const int cy = base, dy = base + 1, dys = qq >= 0 ? 1 : -1;
          const int cn = 2 * tri(n, m >= 0 ? m : -m);
          const int s = (m < 0) ? Cnm : Cnm * oddeven((k - m) * (k < m) + m);
          // (a + i b)(c + i d) with M = a + i b (conj for m < 0), Y = c + i d
          if (m >= 0) {
            add(2 * jk, cn, cy, s);
            add(2 * jk, cn + 1, dy, -s * dys);
            add(2 * jk + 1, cn, dy, s * dys);
            add(2 * jk + 1, cn + 1, cy, s);
          } else {
            add(2 * jk, cn, cy, s);
            add(2 * jk, cn + 1, dy, s * dys);
            add(2 * jk + 1, cn, dy, s * dys);
            add(2 * jk + 1, cn + 1, cy, -s);
          }
        }
      }
    }
  }
  for (int r = 0; r < NR; r++) {
    for (int cc = 0; cc < NR; cc++) {
      const Ent &e = T[(size_t)r * NR + cc];
      if ((e.i1 < 0 || e.s1 == 0) && (e.i2 < 0 || e.s2 == 0)) continue;
      bool in = false;
      for (int b = 0; b < NBLK; b++) in |= (r / 16 == BLK_RB[b] && cc / 16 == BLK_CB[b]);
      ok &= in;
    }
  }
  return ok;
}

// table element of coefficient s times harmonic component i
inline int table_elem(int i, int s, bool &ok) {
  switch (s) {
    case 0: return ZERO_ELEM;
    case 1: return i;
    case -1: return NR + i;
    case 2: return 2 * NR + i;
    case -2: return 3 * NR + i;
    default: ok = false; return ZERO_ELEM;
  }
}

// Codes for block b, k-step kk, lane l at index (b * 4 + kk) * 64 + l: the
// entry T[16 rb + l % 16][16 cb + 4 kk + l / 16] as two byte offsets into the
// table of elem_bytes-sized elements.
inline bool build_codes(std::vector<uint32_t> &codes, int elem_bytes) {
  std::vector<Ent> T;
  bool ok = build_entries(T);
  codes.assign((size_t)NBLK * 4 * 64, 0u);
  for (int b = 0; b < NBLK; b++) {
    for (int kk = 0; kk < 4; kk++) {
      for (int l = 0; l < 64; l++) {
        const int r = 16 * BLK_RB[b] + (l % 16);
        const int cc = 16 * BLK_CB[b] + 4 * kk + l / 16;
        const Ent &e = T[(size_t)r * NR + cc];
        const int e1 = e.i1 >= 0 ? table_elem(e.i1, e.s1, ok) : ZERO_ELEM;
        const int e2 = e.i2 >= 0 ? table_elem(e.i2, e.s2, ok) : ZERO_ELEM;
        const uint32_t o1 = (uint32_t)(e1 * elem_bytes), o2 = (uint32_t)(e2 * elem_bytes);
        if (o1 > 0xffffu || o2 > 0xffffu) ok = false;
        codes[(size_t)(b * 4 + kk) * 64 + l] = o1 | (o2 << 16);
      }
    }
  }
  return ok;
}

}  // namespace m2lc
}  // namespace tfmm
