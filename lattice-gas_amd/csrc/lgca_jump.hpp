// Host-side jump-ahead tables for the exact CPU random streams.
//
// The reference draws one xorshift64* value per fluid cell, sequentially per
// row band and per step (src/lattice.rs:232-256), and 4 (3) per fluid inflow
// cell from one serial stream (src/lattice.rs:270-285). A device lane that
// owns a 64-cell word must start at the right place in its stream. xorshift's
// state update is linear over GF(2)^64, s -> T s, so
//     T^n s = c(T) s,  c(z) = z^n mod P(z),
// where P is T's characteristic polynomial (Cayley-Hamilton). P is found once
// by Berlekamp-Massey on one output bit of the state sequence; the jump is
// then a 64-term XOR over the stream's first 64 states. This is the method of
// Haramoto, Matsumoto, Nishimura, Panneton and L'Ecuyer, "Efficient Jump Ahead
// for F2-Linear Random Number Generators", INFORMS J. Computing 20(3), 2008.
//
// The polynomials depend only on the obstacle layout (which cells draw), so
// they are built once per load and reused by every step.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace lgca {

inline uint64_t xs_step(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

// Low 64 coefficients of P(z) = z^64 + sum_{i<64} p_i z^i.
inline uint64_t charpoly_find() {
  // Bit 0 of the state sequence from an arbitrary non-zero start.
  const int N = 256;
  uint8_t a[N];
  uint64_t s = 0x0123456789ABCDEFull;
  for (int i = 0; i < N; ++i) {
    a[i] = (uint8_t)(s & 1u);
    s = xs_step(s);
  }
  // Berlekamp-Massey over GF(2): connection polynomial C(z) = 1 + c_1 z + ...
  std::vector<uint8_t> C(N + 1, 0), B(N + 1, 0), T;
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < N; ++n) {
    uint8_t d = a[n];
    for (int i = 1; i <= L; ++i) d ^= (uint8_t)(C[i] & a[n - i]);
    if (d == 0) {
      ++m;
    } else if (2 * L <= n) {
      T = C;
      for (int i = 0; i + m <= N; ++i) C[i + m] ^= B[i];
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      for (int i = 0; i + m <= N; ++i) C[i + m] ^= B[i];
      ++m;
    }
  }
  // P(z) = z^L C(1/z): coefficient of z^(L-j) is c_j.
  uint64_t p = 0;
  if (L == 64)
    for (int j = 1; j <= 64; ++j)
      if (C[j]) p |= 1ull << (64 - j);
  return p;
}
inline uint64_t charpoly_low() {
  static const uint64_t p = charpoly_find();  // thread-safe one-time initialisation
  return p;
}

// c(z) * z^f mod P(z), one bit at a time
inline uint64_t poly_mulz_slow(uint64_t c, uint32_t f, uint64_t p) {
  for (uint32_t i = 0; i < f; ++i) {
    uint64_t carry = c >> 63;
    c <<= 1;
    if (carry) c ^= p;
  }
  return c;
}

// c(z) * z^f mod P(z) for f <= 32: the f bits shifted past z^63 are folded
// back with four byte tables, R[i][b] = b(z) z^(64 + 8i) mod P(z).
struct MulzTables {
  uint64_t p = 0;
  uint64_t R[4][256];
  explicit MulzTables(uint64_t poly) : p(poly) {
    for (int i = 0; i < 4; ++i)
      for (uint32_t b = 0; b < 256; ++b) {
        uint64_t acc = 0;
        for (int k = 0; k < 8; ++k)
          if ((b >> k) & 1u) acc ^= poly_mulz_slow(1, 64 + 8 * i + k, poly);
        R[i][b] = acc;
      }
  }
  uint64_t mulz(uint64_t c, uint32_t f) const {
    if (f == 0) return c;
    const uint64_t over = c >> (64 - f);  // coefficients of z^64 .. z^(63+f)
    return (c << f) ^ R[0][over & 255u] ^ R[1][(over >> 8) & 255u] ^ R[2][(over >> 16) & 255u] ^
           R[3][(over >> 24) & 255u];
  }
};

inline uint64_t poly_mulz(uint64_t c, uint32_t f, uint64_t p) {
  static const MulzTables tabs(p);
  if (p != tabs.p) return poly_mulz_slow(c, f, p);
  while (f > 32) {
    c = tabs.mulz(c, 32);
    f -= 32;
  }
  return tabs.mulz(c, f);
}

// a(z) * b(z) mod P(z)
inline uint64_t poly_mulmod(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t r = 0;
  for (int i = 63; i >= 0; --i) {
    uint64_t carry = r >> 63;
    r <<= 1;
    if (carry) r ^= p;
    if ((b >> i) & 1u) r ^= a;
  }
  return r;
}

// z^n mod P(z)
inline uint64_t poly_pow(uint64_t n, uint64_t p) {
  uint64_t r = 1, base = 2;  // z
  while (n) {
    if (n & 1u) r = poly_mulmod(r, base, p);
    base = poly_mulmod(base, base, p);
    n >>= 1;
  }
  return r;
}

// T^n s, evaluated through the polynomial (for tests and the host tables).
inline uint64_t jump_state(uint64_t s, uint64_t n) {
  uint64_t c = poly_pow(n, charpoly_low());
  uint64_t st = 0, b = s;
  for (int i = 0; i < 64; ++i) {
    if ((c >> i) & 1u) st ^= b;
    b = xs_step(b);
  }
  return st;
}

// Band geometry of src/lattice.rs:232-240.
inline uint32_t band_rows(uint32_t h, uint32_t threads) {
  uint32_t nt = threads < 1 ? 1 : threads;
  if (nt > h) nt = h;
  return (h + nt - 1) / nt;
}
inline uint32_t n_bands(uint32_t h, uint32_t threads) {
  uint32_t b = band_rows(h, threads);
  return (h + b - 1) / b;
}

// Static per-load tables.
//   cjump[2 * (y * wpr + j) + h]: z^n mod P, n = draws of the row band's
//                       stream before the first cell of word (y, j) (h = 0)
//                       or before its cell 32 (h = 1); a draw is one fluid
//                       cell in row-major order.
//   row_band[y]:        the band (collision stream) of row y.
//   ijump[2 * (y * iw + j) + h]: the same for the inflow stream (4 draws per
//                       fluid inflow cell, 3 without rest).
//   order[k]:           the k-th word a step visits, packed (y << 16) | j:
//                       first the inflow words, then the words holding a
//                       solid site or the row's ragged end, then the rest in
//                       row-major order, so that waves of ordinary words never
//                       diverge into the slower paths.
// `solid` is the solid plane, wpr words per row; tail bits must be zero.
inline void build_tables(uint32_t w, uint32_t h, uint32_t bands, uint32_t inlet_cols, bool rest,
                         const std::vector<uint64_t>& solid, std::vector<uint64_t>& cjump,
                         std::vector<uint32_t>& row_band, std::vector<uint64_t>& ijump,
                         std::vector<uint32_t>& order) {
  const uint64_t p = charpoly_low();
  const uint32_t wpr = (w + 63) / 64;
  const uint32_t band = band_rows(h, bands);
  const uint64_t LO = 0xFFFFFFFFull;
  cjump.assign((size_t)2 * wpr * h, 0);
  row_band.assign(h, 0);
  uint64_t c = 1;
  for (uint32_t y = 0; y < h; ++y) {
    if (y % band == 0) c = 1;
    row_band[y] = y / band;
    for (uint32_t j = 0; j < wpr; ++j) {
      size_t word = (size_t)y * wpr + j;
      uint32_t ncell = (j + 1) * 64 <= w ? 64 : w - j * 64;
      uint64_t cells = ncell == 64 ? ~0ull : ((1ull << ncell) - 1);
      uint64_t fl = cells & ~solid[word];
      cjump[2 * word] = c;
      c = poly_mulz(c, (uint32_t)__builtin_popcountll(fl & LO), p);
      cjump[2 * word + 1] = c;
      c = poly_mulz(c, (uint32_t)__builtin_popcountll(fl & ~LO), p);
    }
  }
  const uint32_t cols = inlet_cols < w ? inlet_cols : w;
  const uint32_t iw = (cols + 63) / 64;
  const uint32_t per = rest ? 4 : 3;
  ijump.assign((size_t)2 * iw * h, 0);
  c = 1;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = 0; j < iw; ++j) {
      uint32_t span = cols - j * 64;
      uint64_t m = span >= 64 ? ~0ull : ((1ull << span) - 1);
      uint64_t fl = m & ~solid[(size_t)y * wpr + j];
      ijump[2 * ((size_t)y * iw + j)] = c;
      c = poly_mulz(c, (uint32_t)__builtin_popcountll(fl & LO) * per, p);
      ijump[2 * ((size_t)y * iw + j) + 1] = c;
      c = poly_mulz(c, (uint32_t)__builtin_popcountll(fl & ~LO) * per, p);
    }
  const bool ragged = (w % 64) != 0;
  order.clear();
  order.reserve((size_t)wpr * h);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = 0; j < iw; ++j) order.push_back((y << 16) | j);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = iw; j < wpr; ++j)
      if (solid[(size_t)y * wpr + j] || (ragged && j == wpr - 1)) order.push_back((y << 16) | j);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = iw; j < wpr; ++j)
      if (!(solid[(size_t)y * wpr + j] || (ragged && j == wpr - 1))) order.push_back((y << 16) | j);
}

// Per-cell inflow jumps for inflow widths of at most 64 cells (one word):
// icell[y * cols + x] = z^n mod P, n = draws of the inflow stream before
// cell (y, x) (4 per fluid inflow cell, 3 without rest, row-major), 0 for a
// solid cell, which takes no draw (src/lattice.rs:270-285). The same
// positions ijump gives per word, one cell at a time.
inline void build_inflow_cells(uint32_t w, uint32_t h, uint32_t inlet_cols, bool rest,
                               const std::vector<uint64_t>& solid, std::vector<uint64_t>& icell) {
  const uint64_t p = charpoly_low();
  const uint32_t wpr = (w + 63) / 64;
  const uint32_t cols = inlet_cols < w ? inlet_cols : w;
  const uint32_t per = rest ? 4 : 3;
  icell.assign((size_t)cols * h, 0);
  uint64_t c = 1;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < cols; ++x) {
      if ((solid[(size_t)y * wpr + x / 64] >> (x % 64)) & 1u) continue;
      icell[(size_t)y * cols + x] = c;
      c = poly_mulz(c, per, p);
    }
}

// Jumps to a point before each word's first draw, for the resident bands'
// plain-chain draw of a half whose cells without a draw form one run: lead[w]
// = cjump[2 w] * z^(-len) mod P, len the count of the word's low-half cells
// that take no draw when those form one run (else len = 0), so the chain of
// 32 draws ends on the half's last draw (lg_run_fix_lead). z^(-len) =
// z^(2^64 - 1 - len): P is primitive, so z has order 2^64 - 1. Returns
// whether every half of every word has at most one such run (the layouts
// the single-run draw covers: plates, boxes, discs, a row's tail).
inline bool build_lead_jumps(uint32_t w, uint32_t h, const std::vector<uint64_t>& solid,
                             const std::vector<uint64_t>& cjump, std::vector<uint64_t>& lead) {
  const uint64_t p = charpoly_low();
  const uint32_t wpr = (w + 63) / 64;
  struct Inv {
    uint64_t z[33];
    Inv() {
      for (uint32_t k = 0; k <= 32; ++k) z[k] = k ? poly_pow(~0ull - k, charpoly_low()) : 1ull;
    }
  };
  static const Inv inv;  // z^-k for k <= 32, built once per process (thread-safe)
  const uint64_t* zinv = inv.z;
  auto one_run = [](uint32_t fh) {
    const uint32_t z = ~fh;
    return ((z + (z & (0u - z))) & z) == 0u;
  };
  bool all = true;
  lead.assign((size_t)wpr * h, 0);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = 0; j < wpr; ++j) {
      const size_t word = (size_t)y * wpr + j;
      const uint32_t ncell = (j + 1) * 64 <= w ? 64 : w - j * 64;
      const uint64_t cells = ncell == 64 ? ~0ull : ((1ull << ncell) - 1);
      const uint64_t fl = cells & ~solid[word];
      const uint32_t f0 = (uint32_t)fl, f1 = (uint32_t)(fl >> 32);
      all = all && one_run(f0) && one_run(f1);
      const uint32_t len = one_run(f0) ? (uint32_t)__builtin_popcount(~f0) : 0u;
      lead[word] = len ? poly_mulmod(cjump[2 * word], zinv[len], p) : cjump[2 * word];
    }
  return all;
}

// Window tables of the maps T^(4g), g = 0..15 (T^0 the identity): T^(4g) is
// linear over GF(2)^64 and the same for every stream and step, so
//     T^(4g) s = XOR over w = 0..15 of pw[g * 256 + w * 16 + ((s >> 4w) & 15)],
//     pw[g * 256 + w * 16 + v] = T^(4g) (v << 4w).
// A lane then reaches state 4g of a freshly seeded stream with 16 lookups
// instead of walking 4g steps (lg_xs_pow4).
inline void xs_pow4_windows(std::vector<uint64_t>& pw) {
  pw.assign((size_t)16 * 256, 0);
  for (uint32_t g = 0; g < 16; ++g) {
    uint64_t col[64];  // T^(4g) of each basis vector
    for (uint32_t i = 0; i < 64; ++i) {
      uint64_t x = 1ull << i;
      for (uint32_t k = 0; k < 4 * g; ++k) x = xs_step(x);
      col[i] = x;
    }
    for (uint32_t w = 0; w < 16; ++w)
      for (uint32_t v = 0; v < 16; ++v) {
        uint64_t e = 0;
        for (uint32_t b = 0; b < 4; ++b)
          if ((v >> b) & 1u) e ^= col[4 * w + b];
        pw[(size_t)g * 256 + w * 16 + v] = e;
      }
  }
}

}  // namespace lgca
