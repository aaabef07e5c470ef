// Test harness: the HIP engine's per-word step (lgca_device.h), compiled for
// the host and driven over a whole lattice, so CPU tests can compare the
// exact device algorithm with the oracle without a GPU. Not part of the
// product: nothing outside tests/ loads it.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../lattice-gas_amd/csrc/lgca_device.h"
#include "../../lattice-gas_amd/csrc/lgca_jump.hpp"

namespace {
struct Sim {
  LgGeom G;
  uint64_t seed, steps = 0;
  uint32_t bands;
  std::vector<uint64_t> a, b, solid, cjump, ijump, basis;
  std::vector<uint32_t> row_band, order;
  uint32_t units = 1;
  uint32_t passes = 1;
  uint32_t draw_units = 2;
  std::vector<uint32_t> kbuf, ibuf;
};
}  // namespace

extern "C" {

void* hs_new(uint32_t w, uint32_t h, int rest, uint64_t seed, uint32_t bands, uint32_t inlet_cols,
             const uint32_t* thresh7) {
  auto* S = new Sim();
  LgGeom& G = S->G;
  std::memset(&G, 0, sizeof(G));
  G.w = w;
  G.h = h;
  G.wpr = (w + 63) / 64;
  G.total = G.wpr * h;
  G.lastword = (w - 1) / 64;
  G.lastbit = (w - 1) % 64;
  G.tailmask = G.lastbit == 63 ? ~0ull : ((1ull << (G.lastbit + 1)) - 1);
  G.inlet_cols = inlet_cols < w ? inlet_cols : w;
  G.iw = (G.inlet_cols + 63) / 64;
  G.rest = rest ? 1u : 0u;
  G.nbands = lgca::n_bands(h, bands);
  for (int d = 0; d < 7; ++d) G.thresh[d] = thresh7 ? thresh7[d] : 0;
  S->seed = seed;
  S->bands = bands;
  S->a.assign((size_t)7 * G.total, 0);
  S->b.assign((size_t)7 * G.total, 0);
  S->solid.assign(G.total, 0);
  S->basis.assign((size_t)(G.nbands + 1) * LG_WTAB, 0);
  return S;
}

void hs_free(void* p) { delete static_cast<Sim*>(p); }
void hs_set_units(void* p, uint32_t units) { static_cast<Sim*>(p)->units = units; }
void hs_set_passes(void* p, uint32_t passes) { static_cast<Sim*>(p)->passes = passes; }
// lg_step2_role for blocks 0..cblocks+dblocks-1 into draw[], rb[]
void hs_step2_roles(uint32_t cblocks, uint32_t dblocks, uint32_t* draw, uint32_t* rb) {
  for (uint32_t b = 0; b < cblocks + dblocks; ++b) lg_step2_role(b, cblocks, dblocks, &draw[b], &rb[b]);
}
void hs_set_draw_units(void* p, uint32_t units) { static_cast<Sim*>(p)->draw_units = units; }

void hs_load(void* p, const uint8_t* cells) {
  auto* S = static_cast<Sim*>(p);
  LgGeom& G = S->G;
  std::fill(S->a.begin(), S->a.end(), 0);
  std::fill(S->solid.begin(), S->solid.end(), 0);
  for (uint32_t y = 0; y < G.h; ++y)
    for (uint32_t x = 0; x < G.w; ++x) {
      uint8_t c = cells[(size_t)y * G.w + x];
      size_t word = (size_t)y * G.wpr + x / 64;
      uint64_t bit = 1ull << (x % 64);
      for (int d = 0; d < 7; ++d)
        if (c & (1u << d)) S->a[(size_t)d * G.total + word] |= bit;
      if (c & 0x80u) S->solid[word] |= bit;
    }
  lgca::build_tables(G.w, G.h, S->bands, G.inlet_cols, G.rest != 0, S->solid, S->cjump,
                     S->row_band, S->ijump, S->order);
}

void hs_store(void* p, uint8_t* cells) {
  auto* S = static_cast<Sim*>(p);
  LgGeom& G = S->G;
  for (uint32_t y = 0; y < G.h; ++y)
    for (uint32_t x = 0; x < G.w; ++x) {
      size_t word = (size_t)y * G.wpr + x / 64;
      uint32_t k = x % 64;
      uint8_t c = 0;
      for (int d = 0; d < 7; ++d) c |= (uint8_t)(((S->a[(size_t)d * G.total + word] >> k) & 1u) << d);
      c |= (uint8_t)(((S->solid[word] >> k) & 1u) << 7);
      cells[(size_t)y * G.w + x] = c;
    }
}

void hs_advance(void* p, uint64_t n) {
  auto* S = static_cast<Sim*>(p);
  LgGeom& G = S->G;
  for (uint64_t it = 0; it < n; ++it) {
    uint64_t t = S->steps;
    for (uint32_t s = 0; s <= G.nbands; ++s) {
      uint64_t st = lg_rng_init(s < G.nbands ? lg_band_seed(S->seed, t, s)
                                             : lg_inlet_seed(S->seed, t + 1));
      lg_build_tab(st, S->basis.data() + (size_t)s * LG_WTAB);
    }
    if (S->passes == 2) {
      // the two-pass step: every half's draws, then every unit's collide
      const size_t nhalf = (size_t)2 * G.total, nih = (size_t)2 * G.iw * G.h;
      S->kbuf.assign(LG_KDIGITS * nhalf, 0xDEADBEEFu);
      S->ibuf.assign(std::max<size_t>(7 * nih, 1), 0xDEADBEEFu);
      for (size_t g = 0; g < (size_t)G.total * S->draw_units; ++g) {
        if (S->draw_units == 1) {
          uint32_t e = S->order[g];
          lg_draw_unit<1>(G, S->solid.data(), S->cjump.data(), S->row_band.data(), S->ijump.data(),
                          S->basis.data(), e >> 16, e & 0xFFFFu, 0, S->kbuf.data(), nhalf,
                          S->ibuf.data(), nih);
        } else {
          uint32_t e = S->order[g >> 1];
          lg_draw_unit<2>(G, S->solid.data(), S->cjump.data(), S->row_band.data(), S->ijump.data(),
                          S->basis.data(), e >> 16, e & 0xFFFFu, (uint32_t)(g & 1u), S->kbuf.data(),
                          nhalf, S->ibuf.data(), nih);
        }
      }
      for (size_t u = 0; u < (size_t)G.total * S->units; ++u) {
        uint32_t word = (uint32_t)(S->units == 1 ? u : u >> 1);
        uint32_t y = word / G.wpr, j = word - y * G.wpr;
        if (S->units == 1)
          lg_collide_unit<1>(G, S->a.data(), S->b.data(), S->solid.data(), S->kbuf.data(), nhalf,
                             S->ibuf.data(), nih, y, j, 0);
        else
          lg_collide_unit<2>(G, S->a.data(), S->b.data(), S->solid.data(), S->kbuf.data(), nhalf,
                             S->ibuf.data(), nih, y, j, (uint32_t)(u & 1u));
      }
      std::swap(S->a, S->b);
      S->steps += 1;
      continue;
    }
    // visit the words in the device's order, each as S->units units
    for (uint32_t e : S->order)
      for (uint32_t hh = 0; hh < S->units; ++hh) {
        if (S->units == 1)
          lg_step_unit<1>(G, S->a.data(), S->b.data(), S->solid.data(), S->cjump.data(),
                          S->row_band.data(), S->ijump.data(), S->basis.data(), e >> 16, e & 0xFFFFu, 0);
        else
          lg_step_unit<2>(G, S->a.data(), S->b.data(), S->solid.data(), S->cjump.data(),
                          S->row_band.data(), S->ijump.data(), S->basis.data(), e >> 16, e & 0xFFFFu, hh);
      }
    std::swap(S->a, S->b);
    S->steps += 1;
  }
}

// The temporally blocked collide pass (lg_collide_k_kernel) emulated on the
// host: steps are drawn as the draw pass draws them, then every tile of L =
// NW x M rows (R = L - 2 ks valid) is stepped ks steps with zero rows beyond
// its edges and its words' neighbours taken across the row (periodic), and
// only its valid rows are stored. Groups of ks steps until n are done.
void hs_advance_tiled(void* p, uint64_t n, uint32_t ks, uint32_t L) {
  auto* S = static_cast<Sim*>(p);
  LgGeom& G = S->G;
  const size_t nhalf = (size_t)2 * G.total, nih = (size_t)2 * G.iw * G.h;
  const size_t T = G.total;
  uint64_t done = 0;
  while (done < n) {
    const uint32_t k = (uint32_t)std::min<uint64_t>(ks, n - done);
    const uint32_t R = L - 2 * k;
    // draws of the group's k steps
    std::vector<uint32_t> kb((size_t)k * LG_KDIGITS * nhalf, 0xDEADBEEFu);
    std::vector<uint32_t> ib(std::max<size_t>((size_t)k * 7 * nih, 1), 0xDEADBEEFu);
    for (uint32_t s = 0; s < k; ++s) {
      const uint64_t t = S->steps + s;
      for (uint32_t q = 0; q <= G.nbands; ++q) {
        uint64_t st = lg_rng_init(q < G.nbands ? lg_band_seed(S->seed, t, q)
                                               : lg_inlet_seed(S->seed, t + 1));
        lg_build_tab(st, S->basis.data() + (size_t)q * LG_WTAB);
      }
      for (uint32_t e : S->order)
        lg_draw_unit<1>(G, S->solid.data(), S->cjump.data(), S->row_band.data(), S->ijump.data(),
                        S->basis.data(), e >> 16, e & 0xFFFFu, 0,
                        kb.data() + (size_t)s * LG_KDIGITS * nhalf, nhalf,
                        ib.data() + (size_t)s * 7 * nih, nih);
    }
    const uint32_t ntiles = (G.h + R - 1) / R;
    std::vector<uint64_t> P((size_t)L * 7 * G.wpr), Q(P.size());
    std::vector<uint32_t> Y(L);
    for (uint32_t t = 0; t < ntiles; ++t) {
      for (uint32_t i = 0; i < L; ++i) {
        Y[i] = lg_tile_row(t, i, R, k, G.h);
        for (int d = 0; d < 7; ++d)
          for (uint32_t j = 0; j < G.wpr; ++j)
            P[((size_t)i * 7 + d) * G.wpr + j] = S->a[(size_t)d * T + (size_t)Y[i] * G.wpr + j];
      }
      auto at = [&](uint32_t i, int d, uint32_t j) -> uint64_t {  // zero beyond the tile
        return i < L ? P[((size_t)i * 7 + d) * G.wpr + j] : 0ull;
      };
      for (uint32_t s = 0; s < k; ++s) {
        const uint32_t* kbs = kb.data() + (size_t)s * LG_KDIGITS * nhalf;
        const uint32_t* ibs = ib.data() + (size_t)s * 7 * nih;
        for (uint32_t i = 0; i < L; ++i) {
          const uint32_t y = Y[i], odd = y & 1u, up = i - 1, dn = i + 1;  // up wraps past L at i = 0
          for (uint32_t j = 0; j < G.wpr; ++j) {
            const uint32_t jl = (j == 0) ? G.lastword : j - 1, jr = (j == G.lastword) ? 0 : j + 1;
            const uint32_t cl = (j == 0) ? G.lastbit : 63u, cr = (j == G.lastword) ? G.lastbit : 63u;
            const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
            const size_t word = (size_t)y * G.wpr + j;
            uint64_t nn[7];
            lg_gather(odd, cl, cr, at(i, 0, j), at(i, 0, jl), at(up, 1, j), at(up, 1, jl), at(up, 2, j),
                      at(up, 2, jr), at(i, 3, j), at(i, 3, jr), at(dn, 4, j), at(dn, 4, jr),
                      at(dn, 5, j), at(dn, 5, jl), at(i, 6, j), nn);
            uint32_t Kl[LG_KDIGITS] = {0, 0, 0, 0, 0}, Kh[LG_KDIGITS] = {0, 0, 0, 0, 0};
            const int nk = G.rest ? LG_COLLIDE_REST_KBITS : LG_COLLIDE_NOREST_KBITS;
            for (int b = 0; b < nk; ++b) {
              Kl[b] = kbs[(size_t)b * nhalf + 2 * word];
              Kh[b] = kbs[(size_t)b * nhalf + 2 * word + 1];
            }
            uint64_t res[7];
            lg_collide_word(G, nn, S->solid[word], Kl, Kh, res);
            if (j < G.iw) {
              const size_t iw2 = 2 * ((size_t)y * G.iw + j);
              uint64_t bits[7];
              for (int d = 0; d < 7; ++d)
                bits[d] = ((uint64_t)ibs[(size_t)d * nih + iw2 + 1] << 32) | ibs[(size_t)d * nih + iw2];
              lg_inflow_apply(G, j, S->solid[word], tail, bits, res);
            }
            for (int d = 0; d < 7; ++d) Q[((size_t)i * 7 + d) * G.wpr + j] = res[d] & tail;
          }
        }
        std::swap(P, Q);
      }
      for (uint32_t i = k; i < L - k; ++i) {
        if ((uint64_t)t * R + (i - k) >= G.h) continue;
        for (int d = 0; d < 7; ++d)
          for (uint32_t j = 0; j < G.wpr; ++j)
            S->b[(size_t)d * T + (size_t)Y[i] * G.wpr + j] = P[((size_t)i * 7 + d) * G.wpr + j];
      }
    }
    std::swap(S->a, S->b);
    S->steps += k;
    done += k;
  }
}

// Exhaustive check of the generated circuit: for every state and every k,
// the collided state (stride 30 with rest, 6 without) into out[s * stride + k].
void hs_collide_table(int rest, uint8_t* out) {
  const int ns = rest ? 128 : 64;
  const int stride = rest ? LG_COLLIDE_REST_STRIDE : LG_COLLIDE_NOREST_STRIDE;
  for (int s = 0; s < ns; ++s)
    for (int k = 0; k < stride; ++k) {
      uint64_t n[7], o[7];
      for (int d = 0; d < 7; ++d) n[d] = ((s >> d) & 1) ? ~0ull : 0ull;
      uint8_t r = 0;
      if (rest) {
        uint64_t K[LG_COLLIDE_REST_KBITS];
        for (int b = 0; b < LG_COLLIDE_REST_KBITS; ++b) K[b] = ((k >> b) & 1) ? ~0ull : 0ull;
        lg_collide_rest<uint64_t>(n, K, o);
      } else {
        uint64_t K[LG_COLLIDE_NOREST_KBITS];
        for (int b = 0; b < LG_COLLIDE_NOREST_KBITS; ++b) K[b] = ((k >> b) & 1) ? ~0ull : 0ull;
        lg_collide_norest<uint64_t>(n, K, o);
      }
      int bad = 0;
      for (int d = 0; d < 7; ++d) {
        if (o[d] == ~0ull) r |= (uint8_t)(1u << d);
        else if (o[d] != 0) bad = 1;
      }
      out[s * stride + k] = bad ? 0xFF : r;
    }
}

// The branch-free masked half (lg_draw_half_sel) against the branching one
// (lg_draw_masked32) and the plain one for full masks, on n pseudo-random
// (state, mask) pairs: digits, and the state after the drawing cells.
// Returns the number of mismatches.
int hs_draw_sel_check(uint64_t seed, int n) {
  uint64_t x = seed | 1u;
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    x = lg_xs(x);
    const uint64_t st = lg_rng_init(x);
    x = lg_xs(x);
    uint32_t fh = (uint32_t)(x >> 17);
    if (i % 4 == 0) fh = ~0u;
    if (i % 4 == 1) fh &= (uint32_t)(x >> 40);  // sparser
    uint32_t A[5], B[5], C[5];
    const uint64_t end = lg_draw_half_sel<5>(st, fh, 30, A);
    lg_draw_masked32<5>(st, fh, 30, B);
    uint64_t want = st;
    for (int c = 0; c < __builtin_popcount(fh); ++c) want = lg_xs(want);
    for (int b = 0; b < 5; ++b) bad += A[b] != B[b];
    bad += end != want;
    if (fh == ~0u) {
      const uint64_t e2 = lg_draw_half<5>(st, 30, C);
      for (int b = 0; b < 5; ++b) bad += A[b] != C[b];
      bad += e2 != end;
    }
    uint32_t A3[3], B3[3];
    lg_draw_half_sel<3>(st, fh, 6, A3);
    lg_draw_masked32<3>(st, fh, 6, B3);
    for (int b = 0; b < 3; ++b) bad += A3[b] != B3[b];
  }
  return bad;
}

// The single-run draw (lg_draw_half_run) against the branching masked draw
// (lg_draw_masked32) on n pseudo-random states and masks whose non-drawing
// cells form one run (anywhere, of any length, including none and all 32),
// and lg_one_run against a cell-by-cell count of runs on arbitrary masks.
// Returns the number of mismatches.
int hs_draw_run_check(uint64_t seed, int n) {
  uint64_t x = seed | 1u;
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    x = lg_xs(x);
    const uint64_t st = lg_rng_init(x);
    x = lg_xs(x);
    const uint32_t a = (uint32_t)(x % 33), len = (uint32_t)((x >> 8) % (33 - a));
    const uint64_t run = len ? (((len >= 32 ? ~0ull : (1ull << len) - 1ull)) << a) : 0ull;
    const uint32_t fh = ~(uint32_t)run;
    uint32_t A[5], B[5], A3[3], B3[3];
    lg_draw_half_run<5>(st, fh, 30, A);
    lg_draw_masked32<5>(st, fh, 30, B);
    for (int b = 0; b < 5; ++b) bad += A[b] != B[b];
    lg_draw_half_run<3>(st, fh, 6, A3);
    lg_draw_masked32<3>(st, fh, 6, B3);
    for (int b = 0; b < 3; ++b) bad += A3[b] != B3[b];
    bad += !lg_one_run(fh);
    // arbitrary masks: at most one run of zero bits?
    x = lg_xs(x);
    const uint32_t m = (uint32_t)(x >> 13) | (i % 3 == 0 ? (uint32_t)(x >> 29) : 0u);
    int runs = 0;
    for (int c = 0; c < 32; ++c)
      if (!((m >> c) & 1u) && (c == 0 || ((m >> (c - 1)) & 1u))) ++runs;
    bad += lg_one_run(m) != (runs <= 1);
  }
  return bad;
}

// The lead jumps (lgca::build_lead_jumps) on a w x h layout of boxes (one
// run per half where a box edge falls inside a half, plus the row's tail):
// for every word, the plain chain from its lead jump with lg_run_fix_lead
// gives the low half's masked digits (lg_draw_masked32 from the half's own
// jump) and ends on the state the high half starts from (its own jump); the
// high half's plain chain with lg_run_fix gives its masked digits. Returns
// the number of mismatches, or -1 if build_lead_jumps did not report the
// layout as single-run.
int hs_lead_check(uint32_t w, uint32_t h, uint64_t seed) {
  const uint32_t wpr = (w + 63) / 64;
  std::vector<uint64_t> solid((size_t)wpr * h, 0), cj, ij, lead;
  std::vector<uint32_t> rb, ord;
  uint64_t x = seed | 1u;
  // one box row-chord per row, some rows none; clear of the half that holds
  // a ragged row's tail, so that no half has two runs
  const uint32_t room = (w % 64) ? 64 * (wpr - 1) + ((w % 64) > 32 ? 32 : 0) : w;
  for (uint32_t y = 0; y < h && room > 0; ++y) {
    x = lg_xs(x);
    if (x % 3 == 0) continue;
    const uint32_t x0 = (uint32_t)((x >> 8) % room), len = 1 + (uint32_t)((x >> 24) % 9);
    for (uint32_t c = x0; c < std::min(room, x0 + len); ++c)
      solid[(size_t)y * wpr + c / 64] |= 1ull << (c % 64);
  }
  lgca::build_tables(w, h, 3, 0, true, solid, cj, rb, ij, ord);
  if (!lgca::build_lead_jumps(w, h, solid, cj, lead)) return -1;
  {  // two runs in one half: not covered
    std::vector<uint64_t> s2 = solid, l2;
    s2[0] |= (1ull << 2) | (1ull << 5);
    std::vector<uint64_t> cj2;
    lgca::build_tables(w, h, 3, 0, true, s2, cj2, rb, ij, ord);
    if (lgca::build_lead_jumps(w, h, s2, cj2, l2)) return -2;
  }
  std::vector<uint64_t> tab(LG_WTAB);
  lg_build_tab(lg_rng_init(seed ^ 0x1EAD), tab.data());
  int bad = 0;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t j = 0; j < wpr; ++j) {
      const size_t word = (size_t)y * wpr + j;
      const uint32_t ncell = (j + 1) * 64 <= w ? 64 : w - j * 64;
      const uint64_t cells = ncell == 64 ? ~0ull : ((1ull << ncell) - 1);
      const uint64_t fm = cells & ~solid[word];
      uint32_t A[5], B[5];
      const uint64_t end = lg_draw_half<5>(lg_jump_tab(lead[word], tab.data()), 30, A);
      lg_run_fix_lead<5>((uint32_t)fm, A);
      lg_draw_masked32<5>(lg_jump_tab(cj[2 * word], tab.data()), (uint32_t)fm, 30, B);
      for (int b = 0; b < 5; ++b) bad += A[b] != B[b];
      const uint64_t hi = lg_jump_tab(cj[2 * word + 1], tab.data());
      bad += end != hi;
      lg_draw_half<5>(end, 30, A);
      lg_run_fix<5>((uint32_t)(fm >> 32), A);
      lg_draw_masked32<5>(hi, (uint32_t)(fm >> 32), 30, B);
      for (int b = 0; b < 5; ++b) bad += A[b] != B[b];
    }
  return bad;
}

// The T^(4g) windows (lgca::xs_pow4_windows, lg_xs_pow4) against 4g
// sequential steps, and a stream's window tables assembled group by group
// (lg_tab_group) against lg_build_tab, on n pseudo-random states. Returns the
// number of mismatches.
int hs_pow4_check(uint64_t seed, int n) {
  std::vector<uint64_t> pw, ref(LG_WTAB), tab(LG_WTAB);
  lgca::xs_pow4_windows(pw);
  uint64_t x = seed | 1u;
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    x = lg_xs(x);
    const uint64_t s = lg_rng_init(x);
    uint64_t t = s;
    for (uint32_t g = 0; g < 16; ++g) {
      bad += lg_xs_pow4(s, g, pw.data()) != t;
      lg_tab_group(lg_xs_pow4(s, g, pw.data()), tab.data() + g * 16);
      for (int k = 0; k < 4; ++k) t = lg_xs(t);
    }
    lg_build_tab(s, ref.data());
    for (int e = 0; e < LG_WTAB; ++e) bad += tab[e] != ref[e];
  }
  return bad;
}

// Per-cell inflow jumps (build_inflow_cells) against the per-word ones
// (build_tables' ijump): for a random solid layout and inflow width <= 64,
// every fluid inflow cell sampled from its own jump gives the bits the
// serial per-word chain gives. Returns the number of mismatching words.
int hs_inflow_cells_check(uint32_t w, uint32_t h, uint32_t cols, int rest, uint64_t seed) {
  const uint32_t wpr = (w + 63) / 64;
  std::vector<uint64_t> solid((size_t)wpr * h, 0), cj, ij, ic;
  std::vector<uint32_t> rb, ord;
  uint64_t x = seed | 1u;
  for (auto& v : solid) {
    x = lg_xs(x);
    v = x & lg_xs(x) & lg_xs(lg_xs(x));  // about one cell in eight
  }
  for (uint32_t y = 0; y < h; ++y)
    if (w % 64) solid[(size_t)y * wpr + wpr - 1] &= (1ull << (w % 64)) - 1;
  lgca::build_tables(w, h, 1, cols, rest != 0, solid, cj, rb, ij, ord);
  lgca::build_inflow_cells(w, h, cols, rest != 0, solid, ic);
  LgGeom G;
  std::memset(&G, 0, sizeof(G));
  G.w = w;
  G.rest = rest ? 1u : 0u;
  G.inlet_cols = cols < w ? cols : w;
  G.iw = (G.inlet_cols + 63) / 64;
  const uint32_t th[7] = {5000000, 9000000, 3000000, 12000000, 7000000, 1000000, 8000000};
  for (int d = 0; d < 7; ++d) G.thresh[d] = th[d];
  std::vector<uint64_t> tab(LG_WTAB);
  lg_build_tab(lg_rng_init(seed ^ 0x77), tab.data());
  int bad = 0;
  for (uint32_t y = 0; y < h; ++y) {
    const uint64_t sm = solid[(size_t)y * wpr];
    const uint64_t apply = lg_inlet_apply(G, 0, ~sm);
    uint64_t want[7];
    lg_inlet_bits(G, lg_jump_tab(ij[2 * (size_t)y * G.iw], tab.data()), apply, want);
    uint64_t got[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c = 0; c < G.inlet_cols; ++c) {
      if ((sm >> c) & 1u) continue;
      uint64_t st = lg_jump_tab(ic[(size_t)y * G.inlet_cols + c], tab.data());
      const uint32_t bits = lg_eq_cell(G, st);
      for (int d = 0; d < 7; ++d) got[d] |= (uint64_t)((bits >> d) & 1u) << c;
    }
    for (int d = 0; d < 7; ++d)
      if (got[d] != want[d]) {
        ++bad;
        break;
      }
  }
  return bad;
}

uint64_t hs_charpoly(void) { return lgca::charpoly_low(); }

// The device source's generator: Rng::new(seed), then n next_u64 outputs.
uint64_t hs_rng_state(uint64_t seed) { return lg_rng_init(seed); }
void hs_rng_u64(uint64_t seed, size_t n, uint64_t* out) {
  uint64_t s = lg_rng_init(seed);
  for (size_t i = 0; i < n; ++i) {
    s = lg_xs(s);
    out[i] = s * LG_RNG_MUL;
  }
}
// The device source's stream seeds (band b of step t; inflow after step t).
uint64_t hs_band_seed(uint64_t seed, uint64_t t, uint32_t b) { return lg_band_seed(seed, t, b); }
uint64_t hs_inlet_seed(uint64_t seed, uint64_t t) { return lg_inlet_seed(seed, t + 1); }

// Table-driven multiplication by z^f mod P against the bit-serial one, on n
// pseudo-random (c, f) pairs with f up to 199; returns the mismatch count.
int hs_mulz_check(uint64_t seed, int n) {
  const uint64_t p = lgca::charpoly_low();
  uint64_t x = seed | 1u;
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    x = lg_xs(x);
    uint64_t c = x * 0x9E3779B97F4A7C15ull;
    uint32_t f = (uint32_t)(x >> 40) % 200u;
    bad += lgca::poly_mulz(c, f, p) != lgca::poly_mulz_slow(c, f, p);
  }
  return bad;
}
uint64_t hs_jump_state(uint64_t s, uint64_t n) { return lgca::jump_state(s, n); }

}  // extern "C"
