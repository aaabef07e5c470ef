// Device-side core of the MI355X FHP engine: the per-word step on u64
// bitplanes, reproducing the reference CPU path bit for bit.
//
// The code here is __host__ __device__: the HIP kernels call it on gfx950,
// and a host build of the same source is what the CPU-side kernel tests run.
//
// Layout (one bit per cell, 64 cells per word, little-bit order):
//   plane p in 0..6 (six directions, then rest), word (y, j) at
//   p * total + y * wpr + j, cell x of row y is bit x % 64 of word x / 64
//   (the reference's plane layout, src/gpu.rs:29-32,586-602, widened to u64).
//   The solid plane is separate and never changes (src/gpu.rs:453-458).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LG_HD __host__ __device__
#else
#define LG_HD
#endif

typedef uint64_t lg_word;

// ---------------------------------------------------------------------------
// Three-input lookup: result bit = tt[(a << 2) | (b << 1) | c]. On gfx950 this
// is v_bitop3_b32 (src0 <-> 0xF0, src1 <-> 0xCC, src2 <-> 0xAA), once per half.
// ---------------------------------------------------------------------------
template <uint32_t TT>
LG_HD inline uint64_t lg_op3(uint64_t a, uint64_t b, uint64_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, TT);
  uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32),
                                            (uint32_t)(c >> 32), TT);
  return ((uint64_t)hi << 32) | lo;
#else
  uint64_t r = 0;
  for (uint32_t idx = 0; idx < 8; ++idx)
    if ((TT >> idx) & 1u)
      r |= ((idx & 4) ? a : ~a) & ((idx & 2) ? b : ~b) & ((idx & 1) ? c : ~c);
  return r;
#endif
}
#define LG_OP3(a, b, c, tt) lg_op3<(tt)>((a), (b), (c))

#include "lgca_collide.gen.h"

// ---------------------------------------------------------------------------
// Geometry constants (src/hex.rs:36-41)
// ---------------------------------------------------------------------------
#define LG_OPP(d) (((d) + 3) % 6)
// Source column offset of the gather for direction d at a destination row of
// the given parity: DX_{EVEN,ODD}[OPP[d]] (src/lattice.rs:333-342).
LG_HD inline int lg_src_dx(int d, uint32_t odd) {
  const int DXE[6] = {1, 0, -1, -1, -1, 0};
  const int DXO[6] = {1, 1, 0, -1, 0, 1};
  return odd ? DXO[LG_OPP(d)] : DXE[LG_OPP(d)];
}
// Source row offset DY[OPP[d]]
LG_HD inline int lg_src_dy(int d) {
  const int DY[6] = {0, 1, 1, 0, -1, -1};
  return DY[LG_OPP(d)];
}

// ---------------------------------------------------------------------------
// xorshift64* (src/rng.rs:8-29)
// ---------------------------------------------------------------------------
#define LG_RNG_MUL 0x2545F4914F6CDD1Dull
LG_HD inline uint64_t lg_rng_init(uint64_t seed) {
  uint64_t x = seed ^ 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return (x ^ (x >> 31)) | 1ull;
}
LG_HD inline uint64_t lg_xs(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}
// next_u32 of the state just produced: the high half of x * M.
LG_HD inline uint32_t lg_out32(uint64_t x) { return (uint32_t)((x * LG_RNG_MUL) >> 32); }

// Per-step stream seeds (src/lattice.rs:233,242,276).
#define LG_STEP_MIX 0x9E3779B97F4A7C15ull
#define LG_INLET_MIX 0xD1B54A32D192ED03ull
LG_HD inline uint64_t lg_band_seed(uint64_t seed, uint64_t t, uint32_t band) {
  return (seed ^ (t * LG_STEP_MIX)) ^ ((uint64_t)(band + 1) << 32);
}
LG_HD inline uint64_t lg_inlet_seed(uint64_t seed, uint64_t t_after) {
  return seed ^ (t_after * LG_INLET_MIX);
}

// Jump-ahead: the state n draws into a stream whose start is s is
// T^n s = sum_i c_i T^i s with c(z) = z^n mod P(z), P the characteristic
// polynomial of the xorshift matrix T. B[i] = T^i s is the stream's basis,
// computed once per stream per step; c is static per word.
LG_HD inline uint64_t lg_jump(uint64_t c, const uint64_t* B) {
  uint64_t st = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int i = 0; i < 64; ++i) st ^= B[i] & (0ull - ((c >> i) & 1ull));
  return st;
}

// ---------------------------------------------------------------------------
// Step parameters
// ---------------------------------------------------------------------------
struct LgGeom {
  uint32_t w, h, wpr, total;  // total = wpr * h words per plane
  uint32_t lastword;          // (w - 1) / 64
  uint32_t lastbit;           // (w - 1) % 64
  uint64_t tailmask;          // valid bits of the last word of a row
  uint32_t iw;                // inlet words per row: ceil(inlet_cols / 64)
  uint32_t inlet_cols;        // min(cols, w)
  uint32_t rest;              // rest particles present (stride 30) or not (6)
  uint32_t nbands;            // collision streams per step
  uint32_t thresh[7];         // inflow thresholds (src/lattice.rs:87-92)
};

// Gather plane d for destination word (y, j): bit x takes the bit of the
// source site (x + dx, y + dy), periodic in x and y (src/lattice.rs:345-356).
LG_HD inline uint64_t lg_gather(const LgGeom& G, const uint64_t* __restrict plane, uint32_t y,
                                uint32_t j, int dx, int dy) {
  int sy = (int)y + dy;
  if (sy < 0) sy += (int)G.h;
  if (sy >= (int)G.h) sy -= (int)G.h;
  const uint64_t* row = plane + (size_t)sy * G.wpr;
  uint64_t w0 = row[j];
  if (dx == 0) return w0;
  if (dx > 0) {  // bit x takes bit x + 1
    uint64_t hi = (j + 1 < G.wpr) ? row[j + 1] : 0ull;
    uint64_t v = (w0 >> 1) | (hi << 63);
    if (j == G.lastword)  // the last cell takes cell 0 of the row
      v = (v & ~(1ull << G.lastbit)) | ((row[0] & 1ull) << G.lastbit);
    return v;
  }
  // bit x takes bit x - 1
  uint64_t lo = (j > 0) ? row[j - 1] : 0ull;
  uint64_t v = (w0 << 1) | (lo >> 63);
  if (j == 0) v = (v & ~1ull) | ((row[G.lastword] >> G.lastbit) & 1ull);
  return v;
}

// Deposit the binary digits of k (k < 32) for cell x into the digit planes:
// the digit word is rotated so digit b lands on bit (x + b) % 32 of a 32-bit
// accumulator per digit; lg_k_finish rotates each accumulator back by b.
template <int NK>
LG_HD inline void lg_k_deposit(uint32_t acc[NK], uint32_t k, uint32_t x) {
  uint32_t xr = x & 31u;
  uint32_t rk = (k << xr) | (xr ? (k >> (32u - xr)) : 0u);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int b = 0; b < NK; ++b) {
    uint32_t m = 1u << ((xr + (uint32_t)b) & 31u);
    acc[b] = (rk & m) | (acc[b] & ~m);
  }
}
template <int NK>
LG_HD inline uint32_t lg_k_finish_one(uint32_t a, int b) {
  return b ? ((a >> b) | (a << (32 - b))) : a;
}

// Collision draws for one word: K[b] holds digit b of k for every cell that
// draws (fluid cells inside the row), zeros elsewhere. `st` is the stream
// state before the word's first draw; fm the cells that draw, in order.
template <int NK>
LG_HD inline uint64_t lg_draw_word(uint64_t st, uint64_t fm, uint32_t stride, uint64_t K[NK]) {
  uint32_t lo[NK], hi[NK];
  for (int b = 0; b < NK; ++b) lo[b] = hi[b] = 0;
  if (fm == ~0ull) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (uint32_t x = 0; x < 32; ++x) {
      st = lg_xs(st);
      uint32_t k = (uint32_t)(((uint64_t)lg_out32(st) * stride) >> 32);
      lg_k_deposit<NK>(lo, k, x);
    }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (uint32_t x = 32; x < 64; ++x) {
      st = lg_xs(st);
      uint32_t k = (uint32_t)(((uint64_t)lg_out32(st) * stride) >> 32);
      lg_k_deposit<NK>(hi, k, x);
    }
  } else {
    for (uint32_t x = 0; x < 64; ++x) {
      uint32_t take = (uint32_t)((fm >> x) & 1ull);
      uint64_t nx = lg_xs(st);
      st = take ? nx : st;
      uint32_t k = take ? (uint32_t)(((uint64_t)lg_out32(st) * stride) >> 32) : 0u;
      if (x < 32)
        lg_k_deposit<NK>(lo, k, x);
      else
        lg_k_deposit<NK>(hi, k, x);
    }
  }
  for (int b = 0; b < NK; ++b)
    K[b] = ((uint64_t)lg_k_finish_one<NK>(hi[b], b) << 32) | lg_k_finish_one<NK>(lo[b], b);
  return st;
}

// Inflow re-seed for the inlet cells `apply` of one word (src/lattice.rs:102-115,
// 270-285): each applied cell takes 4 draws (3 without rest) in order, and
// its seven bits replace res[] at that cell.
LG_HD inline void lg_inlet_word(const LgGeom& G, uint64_t st, uint64_t apply, uint64_t res[7]) {
  uint64_t bits[7] = {0, 0, 0, 0, 0, 0, 0};
  for (uint32_t x = 0; x < 64; ++x) {
    if (!((apply >> x) & 1ull)) continue;
    uint64_t s = 0;
    for (int d = 0; d < 6; d += 2) {
      st = lg_xs(st);
      uint64_t r = st * LG_RNG_MUL;
      s |= (uint64_t)((uint32_t)(r >> 40) < G.thresh[d]) << d;
      s |= (uint64_t)(((uint32_t)(r >> 16) & 0xFFFFFFu) < G.thresh[d + 1]) << (d + 1);
    }
    if (G.rest) {
      st = lg_xs(st);
      uint64_t r = st * LG_RNG_MUL;
      s |= (uint64_t)((uint32_t)(r >> 40) < G.thresh[6]) << 6;
    }
    for (int d = 0; d < 7; ++d) bits[d] |= ((s >> d) & 1ull) << x;
  }
  for (int d = 0; d < 7; ++d) res[d] = (res[d] & ~apply) | bits[d];
}

// One word of the fused step: gather, exact-stream collision, bounce-back on
// solid sites, inflow re-seed, tail mask (src/lattice.rs:228-285).
//   src/dst:  7 planes each (ping-pong)
//   solid:    the solid plane
//   cjump:    per-word collision jump polynomial
//   row_band: collision stream of each row
//   ijump:    per inlet word (y * iw + j) inlet jump polynomial
//   basis:    (nbands + 1) x 64 stream bases for this step; the last is the
//             inflow stream of the step's end
LG_HD inline void lg_step_word(const LgGeom& G, const uint64_t* __restrict src,
                               uint64_t* __restrict dst, const uint64_t* __restrict solid,
                               const uint64_t* __restrict cjump, const uint32_t* __restrict row_band,
                               const uint64_t* __restrict ijump, const uint64_t* __restrict basis,
                               uint32_t y, uint32_t j) {
  const size_t word = (size_t)y * G.wpr + j;
  const uint32_t odd = y & 1u;
  uint64_t n[7];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int d = 0; d < 6; ++d)
    n[d] = lg_gather(G, src + (size_t)d * G.total, y, j, lg_src_dx(d, odd), lg_src_dy(d));
  n[6] = src[(size_t)6 * G.total + word];

  const uint64_t sm = solid[word];
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  const uint64_t fm = ~sm & tail;

  uint64_t st = lg_jump(cjump[word], basis + (size_t)row_band[y] * 64);
  uint64_t o[7];
  if (G.rest) {
    uint64_t K[LG_COLLIDE_REST_KBITS];
    lg_draw_word<LG_COLLIDE_REST_KBITS>(st, fm, LG_COLLIDE_REST_STRIDE, K);
    lg_collide_rest(n, K, o);
  } else {
    uint64_t K[LG_COLLIDE_NOREST_KBITS];
    lg_draw_word<LG_COLLIDE_NOREST_KBITS>(st, fm, LG_COLLIDE_NOREST_STRIDE, K);
    lg_collide_norest(n, K, o);
  }

  // A wall sends every particle back the way it came and never collides;
  // the rest particle stays (src/hex.rs:45-48, src/lattice.rs:358-359).
  uint64_t res[7];
  for (int d = 0; d < 6; ++d) res[d] = (sm & n[LG_OPP(d)]) | (~sm & o[d]);
  res[6] = (sm & n[6]) | (~sm & o[6]);

  if (j < G.iw) {
    uint32_t x0 = j * 64u;
    uint32_t span = G.inlet_cols - x0;  // > 0 since j < iw
    uint64_t imask = span >= 64 ? ~0ull : ((1ull << span) - 1ull);
    uint64_t apply = imask & fm;
    if (apply) {
      uint64_t ist = lg_jump(ijump[(size_t)y * G.iw + j], basis + (size_t)G.nbands * 64);
      lg_inlet_word(G, ist, apply, res);
    }
  }
  for (int d = 0; d < 7; ++d) dst[(size_t)d * G.total + word] = res[d] & tail;
}
