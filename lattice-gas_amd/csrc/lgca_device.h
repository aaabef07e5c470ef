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

#if defined(__HIP_DEVICE_COMPILE__)
#define LG_UNROLL _Pragma("unroll")
#else
#define LG_UNROLL
#endif

// ---------------------------------------------------------------------------
// Three-input lookup: result bit = TT[(a << 2) | (b << 1) | c]. On gfx950 this
// is v_bitop3_b32 (src0 <-> 0xF0, src1 <-> 0xCC, src2 <-> 0xAA).
// ---------------------------------------------------------------------------
template <uint32_t TT>
LG_HD inline uint32_t lg_op3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
#else
  uint32_t r = 0;
  for (uint32_t idx = 0; idx < 8; ++idx)
    if ((TT >> idx) & 1u)
      r |= ((idx & 4) ? a : ~a) & ((idx & 2) ? b : ~b) & ((idx & 1) ? c : ~c);
  return r;
#endif
}
template <uint32_t TT>
LG_HD inline uint64_t lg_op3(uint64_t a, uint64_t b, uint64_t c) {
  return ((uint64_t)lg_op3<TT>((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32)) << 32) |
         lg_op3<TT>((uint32_t)a, (uint32_t)b, (uint32_t)c);
}

#include "lgca_collide.gen.h"

// ---------------------------------------------------------------------------
// xorshift64* (src/rng.rs:8-29) and the per-step stream seeds
// (src/lattice.rs:233,242,276)
// ---------------------------------------------------------------------------
#define LG_RNG_MUL 0x2545F4914F6CDD1Dull
#define LG_STEP_MIX 0x9E3779B97F4A7C15ull
#define LG_INLET_MIX 0xD1B54A32D192ED03ull

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
// The collision table's column k = (r * stride) >> 32 (src/collision.rs:115-119).
LG_HD inline uint32_t lg_kcol(uint32_t r, uint32_t stride) {
  return (uint32_t)(((uint64_t)r * stride) >> 32);
}
LG_HD inline uint64_t lg_band_seed(uint64_t seed, uint64_t t, uint32_t band) {
  return (seed ^ (t * LG_STEP_MIX)) ^ ((uint64_t)(band + 1) << 32);
}
LG_HD inline uint64_t lg_inlet_seed(uint64_t seed, uint64_t t_after) {
  return seed ^ (t_after * LG_INLET_MIX);
}

// Jump-ahead: the state n draws into a stream whose start is s is
// T^n s = sum_i c_i T^i s with c(z) = z^n mod P(z), P the characteristic
// polynomial of the xorshift matrix T; c is static per word.
//
// lg_jump: against the 64 states B[i] = T^i s.
LG_HD inline uint64_t lg_jump(uint64_t c, const uint64_t* B) {
  uint64_t st = 0;
  for (int i = 0; i < 64; ++i) st ^= B[i] & (0ull - ((c >> i) & 1ull));
  return st;
}
// lg_jump_tab: against 4-bit window tables, tab[g * 16 + v] = XOR of
// T^(4g + b) s over the set bits b of v: 16 lookups instead of 64 terms.
#define LG_WTAB 256  // entries per stream per step
LG_HD inline uint64_t lg_jump_tab(uint64_t c, const uint64_t* __restrict tab) {
  uint64_t e[16];
  LG_UNROLL
  for (int g = 0; g < 16; ++g) e[g] = tab[g * 16 + (uint32_t)((c >> (4 * g)) & 15u)];
  // XOR of the 16 entries as a tree of three-input XORs (0x96: a ^ b ^ c)
  const uint64_t x0 = lg_op3<0x96>(e[0], e[1], e[2]), x1 = lg_op3<0x96>(e[3], e[4], e[5]);
  const uint64_t x2 = lg_op3<0x96>(e[6], e[7], e[8]), x3 = lg_op3<0x96>(e[9], e[10], e[11]);
  const uint64_t x4 = lg_op3<0x96>(e[12], e[13], e[14]);
  return lg_op3<0x96>(x0, x1, x2) ^ lg_op3<0x96>(x3, x4, e[15]);
}
// State 4g of a stream from its start s (g < 16): T^(4g) s through 4-bit
// windows of s against pw = lgca::xs_pow4_windows (16 lookups).
LG_HD inline uint64_t lg_xs_pow4(uint64_t s, uint32_t g, const uint64_t* __restrict pw) {
  const uint64_t* __restrict t = pw + (size_t)g * 256;
  uint64_t e[16];
  LG_UNROLL
  for (int w = 0; w < 16; ++w) e[w] = t[w * 16 + (uint32_t)((s >> (4 * w)) & 15u)];
  const uint64_t x0 = lg_op3<0x96>(e[0], e[1], e[2]), x1 = lg_op3<0x96>(e[3], e[4], e[5]);
  const uint64_t x2 = lg_op3<0x96>(e[6], e[7], e[8]), x3 = lg_op3<0x96>(e[9], e[10], e[11]);
  const uint64_t x4 = lg_op3<0x96>(e[12], e[13], e[14]);
  return lg_op3<0x96>(x0, x1, x2) ^ lg_op3<0x96>(x3, x4, e[15]);
}
// Window group g (entries g * 16 .. g * 16 + 15) of a stream's tables from
// its state 4g: entry v is the XOR of states 4g + b over the set bits b of v.
LG_HD inline void lg_tab_group(uint64_t b0, uint64_t* __restrict o) {
  const uint64_t b1 = lg_xs(b0), b2 = lg_xs(b1), b3 = lg_xs(b2);
  const uint64_t b01 = b0 ^ b1, b23 = b2 ^ b3;
  o[0] = 0;
  o[1] = b0;
  o[2] = b1;
  o[3] = b01;
  o[4] = b2;
  o[5] = b0 ^ b2;
  o[6] = b1 ^ b2;
  o[7] = b01 ^ b2;
  o[8] = b3;
  o[9] = b0 ^ b3;
  o[10] = b1 ^ b3;
  o[11] = b01 ^ b3;
  o[12] = b23;
  o[13] = b0 ^ b23;
  o[14] = b1 ^ b23;
  o[15] = b01 ^ b23;
}
// Fill one stream's window tables from its seeded start state.
LG_HD inline void lg_build_tab(uint64_t s, uint64_t* __restrict tab) {
  uint64_t B[64];
  for (int i = 0; i < 64; ++i) {
    B[i] = s;
    s = lg_xs(s);
  }
  for (int g = 0; g < 16; ++g)
    for (uint32_t v = 0; v < 16; ++v) {
      uint64_t e = 0;
      for (int b = 0; b < 4; ++b)
        if ((v >> b) & 1u) e ^= B[4 * g + b];
      tab[g * 16 + v] = e;
    }
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

// ---------------------------------------------------------------------------
// Coarse-graining of block g of a bw-wide grid of bx x by blocks: exact
// integer sums over fluid cells (src/moments.rs:37-65 FLUID,
// src/render.rs:93-106), then the f32 running average of src/render.rs:107-115
// with every operation rounded on its own (no fused multiply-add), as the CPU
// computes it. sums[4 g ..] receives (mass, px2, py2, nsolid).
#if defined(__HIPCC__)
// Exact partial sums over the (row, word) pieces item0, item0 + istep, ...
// of block g (by rows x the words the block's bx columns touch).
__device__ inline void lg_block_sums(const LgGeom& G, const uint64_t* __restrict planes,
                                     const uint64_t* __restrict solid, uint32_t bx, uint32_t by,
                                     uint32_t bw, uint32_t g, uint32_t item0, uint32_t istep,
                                     int64_t s[4]) {
  const uint32_t jb = g / bw, ib = g - jb * bw;
  const int CX2[6] = {2, 1, -1, -2, -1, 1};
  const int CY2[6] = {0, 1, 1, 0, -1, -1};
  const uint32_t x0 = ib * bx, x1 = x0 + bx;
  const uint32_t w0 = x0 >> 6, nwb = ((x1 - 1) >> 6) - w0 + 1;
  int64_t mass = 0, px = 0, py = 0, ns = 0;
  for (uint32_t it = item0; it < by * nwb; it += istep) {
    const uint32_t r = it / nwb, wi = w0 + (it - r * nwb);
    const size_t base = (size_t)(jb * by + r) * G.wpr;
    const uint32_t lo = max(x0, wi << 6) - (wi << 6);
    const uint32_t hi = min(x1, (wi + 1) << 6) - (wi << 6);
    const uint64_t m = (hi - lo >= 64) ? ~0ull : (((1ull << (hi - lo)) - 1ull) << lo);
    const uint64_t sm = solid[base + wi] & m;
    ns += __popcll(sm);
    const uint64_t fm = m & ~sm;
    for (int d = 0; d < 6; ++d) {
      const int64_t c = __popcll(planes[(size_t)d * G.total + base + wi] & fm);
      mass += c;
      px += CX2[d] * c;
      py += CY2[d] * c;
    }
    mass += __popcll(planes[(size_t)6 * G.total + base + wi] & fm);
  }
  s[0] = mass;
  s[1] = px;
  s[2] = py;
  s[3] = ns;
}

// Block g's exact sums into sums[4 g ..] and the f32 running average of
// src/render.rs:107-115, every operation rounded on its own as the CPU does.
__device__ inline void lg_block_field(uint32_t bx, uint32_t by, float alpha, const int64_t s[4],
                                      float* __restrict ux, float* __restrict uy,
                                      float* __restrict rho, float* __restrict sol,
                                      int64_t* __restrict sums, uint32_t g) {
  const int64_t mass = s[0], px = s[1], py = s[2], ns = s[3];
  sums[(size_t)g * 4 + 0] = mass;
  sums[(size_t)g * 4 + 1] = px;
  sums[(size_t)g * 4 + 2] = py;
  sums[(size_t)g * 4 + 3] = ns;
  const float SQRT3_2 = 0.8660254f;
  float n = (float)(bx * by);
  float vx = 0.0f, vy = 0.0f;
  if (mass != 0) {
    float mf = (float)mass;
    vx = __fdiv_rn(__fmul_rn((float)px, 0.5f), mf);
    vy = __fdiv_rn(__fmul_rn((float)py, SQRT3_2), mf);
  }
  ux[g] = __fadd_rn(ux[g], __fmul_rn(alpha, __fsub_rn(vx, ux[g])));
  uy[g] = __fadd_rn(uy[g], __fmul_rn(alpha, __fsub_rn(vy, uy[g])));
  rho[g] = __fadd_rn(rho[g], __fmul_rn(alpha, __fsub_rn(__fdiv_rn((float)mass, n), rho[g])));
  sol[g] = __fdiv_rn((float)ns, n);
}

// One wavefront per block: each lane takes every 64th (row, word) piece, the
// partial sums meet by a butterfly of lane exchanges, lane 0 writes.
__device__ inline void lg_sample_block_wave(const LgGeom& G, const uint64_t* __restrict planes,
                                            const uint64_t* __restrict solid, uint32_t bx,
                                            uint32_t by, uint32_t bw, float alpha,
                                            float* __restrict ux, float* __restrict uy,
                                            float* __restrict rho, float* __restrict sol,
                                            int64_t* __restrict sums, uint32_t g, uint32_t lane) {
  int64_t s[4];
  lg_block_sums(G, planes, solid, bx, by, bw, g, lane, 64, s);
  for (int off = 32; off > 0; off >>= 1)
    for (int i = 0; i < 4; ++i) s[i] += __shfl_xor(s[i], off, 64);
  if (lane == 0) lg_block_field(bx, by, alpha, s, ux, uy, rho, sol, sums, g);
}
#endif

// Collision draws for one word: the binary digits of k for the 64 cells.
// ---------------------------------------------------------------------------
// Fast path (every cell draws). Cell x's digits are OR-ed into accumulator
// x % 8 at bits x..x+4 (rotated past bit 31), one shift-or per cell; the
// digit planes are then gathered back with a rotate and a mask per
// accumulator. Cells sharing an accumulator are 8 apart, so their 5-bit
// windows never overlap.
template <int NK>
LG_HD inline uint64_t lg_draw_half(uint64_t st, uint32_t stride, uint32_t P[NK]) {
  uint32_t R[8];
  LG_UNROLL
  for (int m = 0; m < 8; ++m) R[m] = 0;
  LG_UNROLL
  for (uint32_t x = 0; x < 32; ++x) {
    st = lg_xs(st);
    uint32_t k = lg_kcol(lg_out32(st), stride);
    if (x + NK <= 32) {
#if defined(__HIP_DEVICE_COMPILE__)
      // one shift-or per cell (the compiler otherwise emits a funnel shift
      // and a masked or, not knowing k < 2^NK)
      uint32_t t;
      asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(t) : "v"(k), "i"(x), "v"(R[x & 7]));
      R[x & 7] = t;
#else
      R[x & 7] |= k << x;
#endif
    } else {
      R[x & 7] |= (k << x) | (k >> (32 - x));
    }
  }
  // Digit b of cell x sits at bit (x + b) mod 32 of R[x % 8]: select those
  // bits of every accumulator, then one rotate right by b lines them up.
  LG_UNROLL
  for (int b = 0; b < NK; ++b) {
    uint32_t acc = 0;
    LG_UNROLL
    for (int m = 0; m < 8; ++m) {
      const uint32_t sel = 0x01010101u << m;
      // (R & mask) | acc as one three-input op (0xEA)
      acc = lg_op3<0xEA>(R[m], b ? ((sel << b) | (sel >> (32 - b))) : sel, acc);
    }
    P[b] = b ? ((acc >> b) | (acc << (32 - b))) : acc;
  }
  return st;
}

// Whether the cells of a half that take no draw (the zero bits of fh) form
// at most one run of consecutive cells (a plate, a disc's chord, a row's
// tail end).
LG_HD inline bool lg_one_run(uint32_t fh) {
  const uint32_t z = ~fh;
  return ((z + (z & (0u - z))) & z) == 0u;
}

// The digits of a half whose non-drawing cells form one run [a, a + len)
// (lg_one_run), from the plain chain's: cells before the run take the
// chain's draws 0 .. a - 1, the run takes none (digits zero, as
// lg_draw_half_sel gives), and a cell x after it takes draw x - len, which
// the plain chain put at cell x - len: one shift and two masks per digit
// plane instead of the masked chain.
template <int NK>
LG_HD inline void lg_run_fix(uint32_t fh, uint32_t K[NK]) {
  const uint32_t z = ~fh;
  const uint32_t len = (uint32_t)__builtin_popcount(z);
  const uint32_t a = z ? (uint32_t)__builtin_ctz(z) : 32u, b = a + len;  // b <= 32
  const uint32_t before = a >= 32 ? ~0u : (1u << a) - 1u;
  const uint32_t after = b >= 32 ? 0u : ~((1u << b) - 1u);
  LG_UNROLL
  for (int q = 0; q < NK; ++q) K[q] = (K[q] & before) | ((K[q] << (len & 31u)) & after);
}
// The same when the chain started len draws before the half's first draw
// (its jump is z^(n - len), lgca::build_lead_jumps): chain cell i holds draw
// i - len, so a cell x after the run finds its draw in place and a cell x
// before it at chain cell x + len; and the chain's end state is that of the
// half's last draw, from which the next half continues.
template <int NK>
LG_HD inline void lg_run_fix_lead(uint32_t fh, uint32_t K[NK]) {
  const uint32_t z = ~fh;
  const uint32_t len = (uint32_t)__builtin_popcount(z);
  const uint32_t a = z ? (uint32_t)__builtin_ctz(z) : 32u, b = a + len;
  const uint32_t before = a >= 32 ? ~0u : (1u << a) - 1u;
  const uint32_t after = b >= 32 ? 0u : ~((1u << b) - 1u);
  LG_UNROLL
  for (int q = 0; q < NK; ++q) K[q] = ((K[q] >> (len & 31u)) & before) | (K[q] & after);
}
// The digits of such a half by the plain chain from the half's first draw;
// the chain's end state is that of 32 draws, not the half's 32 - len: a
// caller that continues the stream jumps to the next half's start instead.
template <int NK>
LG_HD inline void lg_draw_half_run(uint64_t st, uint32_t fh, uint32_t stride, uint32_t K[NK]) {
  lg_draw_half<NK>(st, stride, K);
  lg_run_fix<NK>(fh, K);
}

// The same 32 cells when only the cells in fh draw (solid sites and cells
// past the row's end take no draw, src/lattice.rs:358-362), without
// branches: every cell advances a candidate state and keeps it only if it
// draws, and a non-drawing cell's digits are zero (what lg_draw_masked32
// gives). It returns the stream state after the drawing cells, so the next
// half can continue the chain. About a fifth more work than lg_draw_half,
// the same for every lane: a wave where some lane's half is partial runs it
// for all its lanes instead of both paths one after the other.
template <int NK>
LG_HD inline uint64_t lg_draw_half_sel(uint64_t st, uint32_t fh, uint32_t stride, uint32_t P[NK]) {
  uint32_t R[8];
  LG_UNROLL
  for (int m = 0; m < 8; ++m) R[m] = 0;
  LG_UNROLL
  for (uint32_t x = 0; x < 32; ++x) {
    const uint64_t nx = lg_xs(st);
    const bool take = (fh >> x) & 1u;
    st = take ? nx : st;
    uint32_t k = lg_kcol(lg_out32(st), stride);
    k = take ? k : 0u;
    if (x + NK <= 32)
      R[x & 7] |= k << x;
    else
      R[x & 7] |= (k << x) | (k >> (32 - x));
  }
  LG_UNROLL
  for (int b = 0; b < NK; ++b) {
    uint32_t acc = 0;
    LG_UNROLL
    for (int m = 0; m < 8; ++m) {
      const uint32_t sel = 0x01010101u << m;
      acc = lg_op3<0xEA>(R[m], b ? ((sel << b) | (sel >> (32 - b))) : sel, acc);
    }
    P[b] = b ? ((acc >> b) | (acc << (32 - b))) : acc;
  }
  return st;
}

// General path: only the cells in `fm` draw, in order (solid sites and the
// cells past the row's end take no draw, src/lattice.rs:358-362).
template <int NK>
LG_HD inline void lg_draw_masked(uint64_t st, uint64_t fm, uint32_t stride, uint32_t lo[NK],
                                 uint32_t hi[NK]) {
  for (int b = 0; b < NK; ++b) lo[b] = hi[b] = 0;
  for (uint32_t x = 0; x < 64; ++x) {
    if (!((fm >> x) & 1ull)) continue;
    st = lg_xs(st);
    uint32_t k = lg_kcol(lg_out32(st), stride);
    for (int b = 0; b < NK; ++b) {
      uint32_t bit = (k >> b) & 1u;
      if (x < 32)
        lo[b] |= bit << x;
      else
        hi[b] |= bit << (x - 32);
    }
  }
}

template <int NK>
LG_HD inline void lg_draw_masked32(uint64_t st, uint32_t fm, uint32_t stride, uint32_t K[NK]) {
  for (int b = 0; b < NK; ++b) K[b] = 0;
  for (uint32_t x = 0; x < 32; ++x) {
    if (!((fm >> x) & 1u)) continue;
    st = lg_xs(st);
    uint32_t k = lg_kcol(lg_out32(st), stride);
    for (int b = 0; b < NK; ++b) K[b] |= ((k >> b) & 1u) << x;
  }
}

template <int NK>
LG_HD inline void lg_draw_word(uint64_t st, uint64_t fm, uint32_t stride, uint32_t lo[NK],
                               uint32_t hi[NK]) {
  if (fm == ~0ull) {
    st = lg_draw_half<NK>(st, stride, lo);
    lg_draw_half<NK>(st, stride, hi);
  } else {
    lg_draw_masked<NK>(st, fm, stride, lo, hi);
  }
}

// Inflow re-seed / equilibrium sampling for the cells `apply` of one word
// (src/lattice.rs:102-115, 197-208, 270-285): each applied cell takes 4 draws
// (3 without rest) in order; bits[d] receives the seven sampled planes.
// Equilibrium::sample of one cell (src/lattice.rs:102-115) from the stream
// state st (advanced past its 4 draws, 3 without rest): bits 0-6.
LG_HD inline uint32_t lg_eq_cell(const LgGeom& G, uint64_t& st) {
  uint32_t s = 0;
  for (int d = 0; d < 6; d += 2) {
    st = lg_xs(st);
    uint64_t r = st * LG_RNG_MUL;
    s |= (uint32_t)((uint32_t)(r >> 40) < G.thresh[d]) << d;
    s |= (uint32_t)(((uint32_t)(r >> 16) & 0xFFFFFFu) < G.thresh[d + 1]) << (d + 1);
  }
  if (G.rest) {
    st = lg_xs(st);
    uint64_t r = st * LG_RNG_MUL;
    s |= (uint32_t)((uint32_t)(r >> 40) < G.thresh[6]) << 6;
  }
  return s;
}

LG_HD inline void lg_inlet_bits(const LgGeom& G, uint64_t st, uint64_t apply, uint64_t bits[7]) {
  for (int d = 0; d < 7; ++d) bits[d] = 0;
  for (uint32_t x = 0; x < 64; ++x) {
    if (!((apply >> x) & 1ull)) continue;
    const uint32_t s = lg_eq_cell(G, st);
    for (int d = 0; d < 7; ++d) bits[d] |= (uint64_t)((s >> d) & 1u) << x;
  }
}
LG_HD inline void lg_inlet_word(const LgGeom& G, uint64_t st, uint64_t apply, uint64_t res[7]) {
  uint64_t bits[7];
  lg_inlet_bits(G, st, apply, bits);
  for (int d = 0; d < 7; ++d) res[d] = (res[d] & ~apply) | bits[d];
}

// ---------------------------------------------------------------------------
// Propagation. Bit d of site (x, y) comes from the site (x + dx, y + dy)
// with (dx, dy) = (DX_{EVEN,ODD}, DY)[OPP[d]] for the destination row's
// parity, periodic in x and y (src/lattice.rs:333-356):
//   d:        0   1   2   3   4   5
//   dy:       0  -1  -1   0  +1  +1
//   dx even: -1  -1   0  +1   0  -1
//   dx odd:  -1   0  +1  +1  +1   0
// dx = -1: bit x takes bit x-1, with the carry from the word on the left
// (the row's last cell for j = 0); dx = +1: bit x takes bit x+1, with the
// carry from the word on the right (cell 0 for the last word).
// ---------------------------------------------------------------------------
LG_HD inline uint64_t lg_shl_in(uint64_t w0, uint64_t wl, uint32_t csh) {
  return (w0 << 1) | ((wl >> csh) & 1ull);
}
LG_HD inline uint64_t lg_shr_in(uint64_t w0, uint64_t wr, uint32_t csh) {
  return (w0 >> 1) | ((wr & 1ull) << csh);
}

// One work unit of the fused step: a whole word (NH = 1) or one 32-bit half
// of it (NH = 2, half h): gather, exact-stream collision, bounce-back on
// solid sites, inflow re-seed, tail mask (src/lattice.rs:228-285).
//   src/dst:  7 planes each (ping-pong)
//   solid:    the solid plane
//   cjump:    2 per word: jump polynomial to the word's first draw and to
//             the first draw of its upper half
//   row_band: collision stream of each row
//   ijump:    2 per inlet word (y * iw + j), the same for the inflow stream
//   tabs:     (nbands + 1) window tables for this step; the last is the
//             inflow stream of the step's end
template <int NH>
LG_HD inline void lg_step_unit(const LgGeom& G, const uint64_t* __restrict src,
                               uint64_t* __restrict dst, const uint64_t* __restrict solid,
                               const uint64_t* __restrict cjump, const uint32_t* __restrict row_band,
                               const uint64_t* __restrict ijump, const uint64_t* __restrict tabs,
                               uint32_t y, uint32_t j, uint32_t h) {
  const size_t word = (size_t)y * G.wpr + j;
  const uint32_t odd = y & 1u;
  const size_t T = G.total;
  const uint32_t ym = (y == 0) ? G.h - 1 : y - 1;
  const uint32_t yp = (y + 1 == G.h) ? 0 : y + 1;
  const uint32_t jl = (j == 0) ? G.lastword : j - 1;
  const uint32_t jr = (j == G.lastword) ? 0 : j + 1;
  const uint32_t cl = (j == 0) ? G.lastbit : 63u;         // carry bit taken from the left word
  const uint32_t cr = (j == G.lastword) ? G.lastbit : 63u;  // where the right carry lands
  const size_t r0 = (size_t)y * G.wpr, rm = (size_t)ym * G.wpr, rp = (size_t)yp * G.wpr;

  const uint64_t c = cjump[2 * word + h];
  const uint32_t band = row_band[y];
  const uint64_t sm = solid[word];
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  const uint64_t fm = ~sm & tail;
  const uint64_t st = lg_jump_tab(c, tabs + (size_t)band * LG_WTAB);

  // Draw first and gather afterwards, so the draw loop does not hold the
  // gathered words in registers (more waves fit a SIMD). A whole-word unit
  // draws both halves, the upper half's chain continuing the lower one's when
  // every cell of the lower half draws (a second, interleaved chain started
  // by its own jump measured slower); a half-word unit draws its own.
  uint32_t Kh[LG_COLLIDE_REST_KBITS], Kl[LG_COLLIDE_REST_KBITS];
  if (NH == 2) {
    const uint32_t fh = (uint32_t)(fm >> (32 * h));
    if (G.rest) {
      if (fh == ~0u)
        lg_draw_half<LG_COLLIDE_REST_KBITS>(st, LG_COLLIDE_REST_STRIDE, Kh);
      else
        lg_draw_masked32<LG_COLLIDE_REST_KBITS>(st, fh, LG_COLLIDE_REST_STRIDE, Kh);
    } else {
      if (fh == ~0u)
        lg_draw_half<LG_COLLIDE_NOREST_KBITS>(st, LG_COLLIDE_NOREST_STRIDE, Kh);
      else
        lg_draw_masked32<LG_COLLIDE_NOREST_KBITS>(st, fh, LG_COLLIDE_NOREST_STRIDE, Kh);
    }
  } else {
    const uint32_t f0 = (uint32_t)fm, f1 = (uint32_t)(fm >> 32);
    uint64_t st1;
    if (G.rest) {
      if (f0 == ~0u) {
        st1 = lg_draw_half<LG_COLLIDE_REST_KBITS>(st, LG_COLLIDE_REST_STRIDE, Kl);
      } else {
        lg_draw_masked32<LG_COLLIDE_REST_KBITS>(st, f0, LG_COLLIDE_REST_STRIDE, Kl);
        st1 = lg_jump_tab(cjump[2 * word + 1], tabs + (size_t)band * LG_WTAB);
      }
      if (f1 == ~0u)
        lg_draw_half<LG_COLLIDE_REST_KBITS>(st1, LG_COLLIDE_REST_STRIDE, Kh);
      else
        lg_draw_masked32<LG_COLLIDE_REST_KBITS>(st1, f1, LG_COLLIDE_REST_STRIDE, Kh);
    } else {
      if (f0 == ~0u) {
        st1 = lg_draw_half<LG_COLLIDE_NOREST_KBITS>(st, LG_COLLIDE_NOREST_STRIDE, Kl);
      } else {
        lg_draw_masked32<LG_COLLIDE_NOREST_KBITS>(st, f0, LG_COLLIDE_NOREST_STRIDE, Kl);
        st1 = lg_jump_tab(cjump[2 * word + 1], tabs + (size_t)band * LG_WTAB);
      }
      if (f1 == ~0u)
        lg_draw_half<LG_COLLIDE_NOREST_KBITS>(st1, LG_COLLIDE_NOREST_STRIDE, Kh);
      else
        lg_draw_masked32<LG_COLLIDE_NOREST_KBITS>(st1, f1, LG_COLLIDE_NOREST_STRIDE, Kh);
    }
  }
#if defined(__HIP_DEVICE_COMPILE__)
  // keep the plane loads below the draw loop
  asm volatile("" : "+v"(Kh[0]) :: "memory");
#endif

  const uint64_t a0 = src[0 * T + r0 + j], a0l = src[0 * T + r0 + jl];
  const uint64_t a1 = src[1 * T + rm + j], a1l = src[1 * T + rm + jl];
  const uint64_t a2 = src[2 * T + rm + j], a2r = src[2 * T + rm + jr];
  const uint64_t a3 = src[3 * T + r0 + j], a3r = src[3 * T + r0 + jr];
  const uint64_t a4 = src[4 * T + rp + j], a4r = src[4 * T + rp + jr];
  const uint64_t a5 = src[5 * T + rp + j], a5l = src[5 * T + rp + jl];
  const uint64_t a6 = src[6 * T + word];

  uint64_t n[7];
  n[0] = lg_shl_in(a0, a0l, cl);
  n[1] = odd ? a1 : lg_shl_in(a1, a1l, cl);
  n[2] = odd ? lg_shr_in(a2, a2r, cr) : a2;
  n[3] = lg_shr_in(a3, a3r, cr);
  n[4] = odd ? lg_shr_in(a4, a4r, cr) : a4;
  n[5] = odd ? a5 : lg_shl_in(a5, a5l, cl);
  n[6] = a6;

  uint64_t o[7];
  {
    uint32_t nh[7], oh[7];
    if (NH == 1) {
      LG_UNROLL
      for (int d = 0; d < 7; ++d) nh[d] = (uint32_t)n[d];
      if (G.rest)
        lg_collide_rest<uint32_t>(nh, Kl, oh);
      else
        lg_collide_norest<uint32_t>(nh, Kl, oh);
      LG_UNROLL
      for (int d = 0; d < 7; ++d) o[d] = oh[d];
    }
    // the upper half (NH = 1) or this lane's half (NH = 2)
    const uint32_t sh = NH == 1 ? 32u : 32u * h;
    LG_UNROLL
    for (int d = 0; d < 7; ++d) nh[d] = (uint32_t)(n[d] >> sh);
    if (G.rest)
      lg_collide_rest<uint32_t>(nh, Kh, oh);
    else
      lg_collide_norest<uint32_t>(nh, Kh, oh);
    LG_UNROLL
    for (int d = 0; d < 7; ++d) o[d] = (NH == 1 ? o[d] : 0ull) | ((uint64_t)oh[d] << sh);
  }

  // A wall sends every particle back the way it came and never collides;
  // the rest particle stays (src/hex.rs:45-48, src/lattice.rs:358-359).
  uint64_t res[7];
  LG_UNROLL
  for (int d = 0; d < 6; ++d) res[d] = (sm & n[(d + 3) % 6]) | (~sm & o[d]);
  res[6] = (sm & n[6]) | (~sm & o[6]);

  const uint64_t mine = NH == 1 ? ~0ull : (0xFFFFFFFFull << (32 * h));
  if (j < G.iw) {
    uint32_t span = G.inlet_cols - j * 64u;  // > 0 since j < iw
    uint64_t imask = span >= 64 ? ~0ull : ((1ull << span) - 1ull);
    uint64_t apply = imask & fm & mine;
    if (apply) {
      uint64_t ist = lg_jump_tab(ijump[2 * ((size_t)y * G.iw + j) + (NH == 1 ? 0 : h)],
                                 tabs + (size_t)G.nbands * LG_WTAB);
      lg_inlet_word(G, ist, apply, res);
    }
  }
  if (NH == 1) {
    LG_UNROLL
    for (int d = 0; d < 7; ++d) dst[(size_t)d * T + word] = res[d] & tail;
  } else {
    uint32_t* dst32 = reinterpret_cast<uint32_t*>(dst);
    LG_UNROLL
    for (int d = 0; d < 7; ++d)
      dst32[2 * ((size_t)d * T + word) + h] = (uint32_t)((res[d] & tail) >> (32 * h));
  }
}

// ---------------------------------------------------------------------------
// Two-pass step. The random draws of a step depend only on the seed, the step
// number and which cells are fluid, never on the lattice state, so they can
// be made ahead of the state update:
//   draw pass:    for each 32-cell half, the collision column digits of its
//                 cells (exact band stream), and for an inflow half the seven
//                 re-seeded planes of its inflow cells (exact inflow stream);
//   collide pass: gather, collide with those digits, bounce back, overwrite
//                 the inflow cells, store.
// The draw pass needs few registers, so many waves hide its serial chains;
// the collide pass carries no random state. The draw pass visits the halves
// in order-table sequence (its masked draws diverge), but files them by
// position: half h of word w has its digits at index 2 w + h of each digit
// plane (kstride halves apart) and, in an inflow word (y, j), its inflow
// planes at index 2 (y iw + j) + h of each inflow plane (istride apart). The
// collide pass, with no divergent paths, runs in plain word order and finds
// them without a table lookup.
// ---------------------------------------------------------------------------
#define LG_KDIGITS 5  // digit planes reserved per step (3 used without rest)

// Role (draw = 1: draw work, 0: collide work) and role-local block index rb of
// block b in a launch of cblocks collide and dblocks draw blocks: the roles
// alternate block by block while both have blocks left, then the longer
// role's remaining blocks follow in order. A bijection onto {collide
// 0..cblocks-1} + {draw 0..dblocks-1}. Consecutive blocks share CUs, so
// single-block alternation puts both roles on every CU, where the collide
// work's memory waits overlap the draw work's arithmetic: alternating in runs
// of 2, 4, 8, 16 or 64 blocks measured 3-10% slower at 4096 x 2304.
LG_HD inline void lg_step2_role(uint32_t b, uint32_t cblocks, uint32_t dblocks, uint32_t* draw,
                                uint32_t* rb) {
  const uint32_t m = cblocks < dblocks ? cblocks : dblocks;
  if (b < 2 * m) {
    *draw = b & 1u;
    *rb = b >> 1;
  } else {
    *draw = cblocks > dblocks ? 0u : 1u;
    *rb = b - m;
  }
}

LG_HD inline uint64_t lg_inlet_apply(const LgGeom& G, uint32_t j, uint64_t fm) {
  uint32_t span = G.inlet_cols - j * 64u;  // > 0 for j < iw
  uint64_t imask = span >= 64 ? ~0ull : ((1ull << span) - 1ull);
  return imask & fm;
}

// The collision digits of one 32-cell half (fluid mask fh) from state st.
template <int NK>
LG_HD inline uint64_t lg_draw_digits(uint64_t st, uint32_t fh, uint32_t stride, uint32_t K[NK]) {
  if (fh == ~0u) return lg_draw_half<NK>(st, stride, K);
  lg_draw_masked32<NK>(st, fh, stride, K);
  return 0;  // the chain's end is not needed past a masked half
}

// Draw pass for word (y, j): both halves (DNH = 1; the upper half's chain
// continues the lower one's when every cell of the lower half draws) or
// half h alone (DNH = 2). tabs are this step's window tables,
// kplanes/iplanes the step's digit and inflow planes.
template <int NK, int DNH>
LG_HD inline void lg_draw_unit_k(const LgGeom& G, uint64_t fm, const uint64_t* __restrict cj,
                                 const uint64_t* __restrict tab, uint32_t h, uint32_t stride,
                                 uint32_t* __restrict kout, size_t kstride) {
  uint32_t K[NK];
  const uint32_t h0 = DNH == 1 ? 0u : h;
  const uint32_t f0 = (uint32_t)(fm >> (32 * h0));
  uint64_t st = lg_draw_digits<NK>(lg_jump_tab(cj[h0], tab), f0, stride, K);
  LG_UNROLL
  for (int b = 0; b < NK; ++b) kout[b * kstride + h0] = K[b];
  if (DNH == 1) {
    if (f0 != ~0u) st = lg_jump_tab(cj[1], tab);
    lg_draw_digits<NK>(st, (uint32_t)(fm >> 32), stride, K);
    LG_UNROLL
    for (int b = 0; b < NK; ++b) kout[b * kstride + 1] = K[b];
  }
}
template <int DNH>
LG_HD inline void lg_draw_unit(const LgGeom& G, const uint64_t* __restrict solid,
                               const uint64_t* __restrict cjump, const uint32_t* __restrict row_band,
                               const uint64_t* __restrict ijump, const uint64_t* __restrict tabs,
                               uint32_t y, uint32_t j, uint32_t h, uint32_t* __restrict kplanes,
                               size_t kstride, uint32_t* __restrict iplanes, size_t istride) {
  const size_t word = (size_t)y * G.wpr + j;
  uint32_t* __restrict kout = kplanes + 2 * word;
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  const uint64_t fm = ~solid[word] & tail;
  const uint64_t* __restrict tab = tabs + (size_t)row_band[y] * LG_WTAB;
  if (G.rest)
    lg_draw_unit_k<LG_COLLIDE_REST_KBITS, DNH>(G, fm, cjump + 2 * word, tab, h,
                                               LG_COLLIDE_REST_STRIDE, kout, kstride);
  else
    lg_draw_unit_k<LG_COLLIDE_NOREST_KBITS, DNH>(G, fm, cjump + 2 * word, tab, h,
                                                 LG_COLLIDE_NOREST_STRIDE, kout, kstride);
  if (j < G.iw) {
    const uint64_t mine = DNH == 1 ? ~0ull : (0xFFFFFFFFull << (32 * h));
    const uint64_t apply = lg_inlet_apply(G, j, fm) & mine;
    if (apply) {
      const size_t iword = (size_t)y * G.iw + j;
      uint64_t ist = lg_jump_tab(ijump[2 * iword + (DNH == 1 ? 0 : h)],
                                 tabs + (size_t)G.nbands * LG_WTAB);
      uint64_t bits[7];
      lg_inlet_bits(G, ist, apply, bits);
      uint32_t* __restrict iout = iplanes + 2 * iword;
      LG_UNROLL
      for (int d = 0; d < 7; ++d) {
        if (DNH == 1) {
          iout[d * istride] = (uint32_t)bits[d];
          iout[d * istride + 1] = (uint32_t)(bits[d] >> 32);
        } else {
          iout[d * istride + h] = (uint32_t)(bits[d] >> (32 * h));
        }
      }
    }
  }
}

// Gather of destination word (y, j) from its source words (parity `odd` of
// row y): aD the plane-D word straight above/below/beside it as the table
// of src/lattice.rs:333-356 says (planes 1, 2 from row y - 1; 4, 5 from row
// y + 1; 0, 3, 6 from row y), aDl / aDr the same plane's word to the left /
// right; cl = the bit taken from a left word (63, or lastbit for j = 0), cr =
// where a right word's bit 0 lands (63, or lastbit for the last word).
LG_HD inline void lg_gather(uint32_t odd, uint32_t cl, uint32_t cr, uint64_t a0, uint64_t a0l,
                            uint64_t a1, uint64_t a1l, uint64_t a2, uint64_t a2r, uint64_t a3,
                            uint64_t a3r, uint64_t a4, uint64_t a4r, uint64_t a5, uint64_t a5l,
                            uint64_t a6, uint64_t n[7]) {
  n[0] = lg_shl_in(a0, a0l, cl);
  n[1] = odd ? a1 : lg_shl_in(a1, a1l, cl);
  n[2] = odd ? lg_shr_in(a2, a2r, cr) : a2;
  n[3] = lg_shr_in(a3, a3r, cr);
  n[4] = odd ? lg_shr_in(a4, a4r, cr) : a4;
  n[5] = odd ? a5 : lg_shl_in(a5, a5l, cl);
  n[6] = a6;
}

// Collision of a whole gathered word with the digits of its two halves,
// then bounce-back on the solid sites sm (src/hex.rs:45-48,
// src/lattice.rs:358-359): res = the word's new planes before inflow and tail.
LG_HD inline void lg_collide_word(const LgGeom& G, const uint64_t n[7], uint64_t sm,
                                  const uint32_t Kl[], const uint32_t Kh[], uint64_t res[7]) {
  uint64_t o[7];
  uint32_t nh[7], oh[7];
  LG_UNROLL
  for (int d = 0; d < 7; ++d) nh[d] = (uint32_t)n[d];
  if (G.rest)
    lg_collide_rest<uint32_t>(nh, Kl, oh);
  else
    lg_collide_norest<uint32_t>(nh, Kl, oh);
  LG_UNROLL
  for (int d = 0; d < 7; ++d) o[d] = oh[d];
  LG_UNROLL
  for (int d = 0; d < 7; ++d) nh[d] = (uint32_t)(n[d] >> 32);
  if (G.rest)
    lg_collide_rest<uint32_t>(nh, Kh, oh);
  else
    lg_collide_norest<uint32_t>(nh, Kh, oh);
  LG_UNROLL
  for (int d = 0; d < 7; ++d) o[d] |= (uint64_t)oh[d] << 32;
  LG_UNROLL
  for (int d = 0; d < 6; ++d) res[d] = (sm & n[(d + 3) % 6]) | (~sm & o[d]);
  res[6] = (sm & n[6]) | (~sm & o[6]);
}

// Collide pass for one unit of word (y, j): the whole word (NH = 1) or its
// half h (NH = 2); kplanes/iplanes are the step's digit and inflow planes.
template <int NH>
LG_HD inline void lg_collide_unit(const LgGeom& G, const uint64_t* __restrict src,
                                  uint64_t* __restrict dst, const uint64_t* __restrict solid,
                                  const uint32_t* __restrict kplanes, size_t kstride,
                                  const uint32_t* __restrict iplanes, size_t istride, uint32_t y,
                                  uint32_t j, uint32_t h) {
  const size_t word = (size_t)y * G.wpr + j;
  const uint32_t* __restrict kin = kplanes + 2 * word + (NH == 1 ? 0 : h);
  const uint32_t odd = y & 1u;
  const size_t T = G.total;
  const uint32_t ym = (y == 0) ? G.h - 1 : y - 1;
  const uint32_t yp = (y + 1 == G.h) ? 0 : y + 1;
  const uint32_t jl = (j == 0) ? G.lastword : j - 1;
  const uint32_t jr = (j == G.lastword) ? 0 : j + 1;
  const uint32_t cl = (j == 0) ? G.lastbit : 63u;
  const uint32_t cr = (j == G.lastword) ? G.lastbit : 63u;
  const size_t r0 = (size_t)y * G.wpr, rm = (size_t)ym * G.wpr, rp = (size_t)yp * G.wpr;
  const uint64_t sm = solid[word];
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;

  uint32_t Kl[LG_KDIGITS], Kh[LG_KDIGITS];
  const int nk = G.rest ? LG_COLLIDE_REST_KBITS : LG_COLLIDE_NOREST_KBITS;
  LG_UNROLL
  for (int b = 0; b < LG_KDIGITS; ++b) {
    Kl[b] = b < nk ? kin[b * kstride] : 0u;
    Kh[b] = (NH == 1 && b < nk) ? kin[b * kstride + 1] : Kl[b];
  }

  const uint64_t a0 = src[0 * T + r0 + j], a0l = src[0 * T + r0 + jl];
  const uint64_t a1 = src[1 * T + rm + j], a1l = src[1 * T + rm + jl];
  const uint64_t a2 = src[2 * T + rm + j], a2r = src[2 * T + rm + jr];
  const uint64_t a3 = src[3 * T + r0 + j], a3r = src[3 * T + r0 + jr];
  const uint64_t a4 = src[4 * T + rp + j], a4r = src[4 * T + rp + jr];
  const uint64_t a5 = src[5 * T + rp + j], a5l = src[5 * T + rp + jl];
  const uint64_t a6 = src[6 * T + word];

  uint64_t n[7];
  lg_gather(odd, cl, cr, a0, a0l, a1, a1l, a2, a2r, a3, a3r, a4, a4r, a5, a5l, a6, n);

  uint64_t res[7];
  if (NH == 1) {
    lg_collide_word(G, n, sm, Kl, Kh, res);
  } else {
    uint64_t o[7];
    uint32_t nh[7], oh[7];
    const uint32_t sh = 32u * h;
    LG_UNROLL
    for (int d = 0; d < 7; ++d) nh[d] = (uint32_t)(n[d] >> sh);
    if (G.rest)
      lg_collide_rest<uint32_t>(nh, Kh, oh);
    else
      lg_collide_norest<uint32_t>(nh, Kh, oh);
    LG_UNROLL
    for (int d = 0; d < 7; ++d) o[d] = (uint64_t)oh[d] << sh;
    // walls: bounce back, no collision (src/hex.rs:45-48, src/lattice.rs:358-359)
    LG_UNROLL
    for (int d = 0; d < 6; ++d) res[d] = (sm & n[(d + 3) % 6]) | (~sm & o[d]);
    res[6] = (sm & n[6]) | (~sm & o[6]);
  }

  if (j < G.iw) {
    const uint64_t mine = NH == 1 ? ~0ull : (0xFFFFFFFFull << (32 * h));
    const uint64_t apply = lg_inlet_apply(G, j, ~sm & tail) & mine;
    if (apply) {
      const uint32_t* __restrict iin = iplanes + 2 * ((size_t)y * G.iw + j) + (NH == 1 ? 0 : h);
      LG_UNROLL
      for (int d = 0; d < 7; ++d) {
        uint64_t bits = NH == 1 ? ((uint64_t)iin[d * istride + 1] << 32) | iin[d * istride]
                                : (uint64_t)iin[d * istride] << (32 * h);
        res[d] = (res[d] & ~apply) | (bits & apply);
      }
    }
  }
  if (NH == 1) {
    LG_UNROLL
    for (int d = 0; d < 7; ++d) dst[(size_t)d * T + word] = res[d] & tail;
  } else {
    uint32_t* dst32 = reinterpret_cast<uint32_t*>(dst);
    LG_UNROLL
    for (int d = 0; d < 7; ++d)
      dst32[2 * ((size_t)d * T + word) + h] = (uint32_t)((res[d] & tail) >> (32 * h));
  }
}

// ---------------------------------------------------------------------------
// Temporal blocking of the collide pass. A tile is a band of L rows, full
// width, stepped `ks` steps in registers before one store: tile t holds the
// rows t R - ks .. t R + R + ks - 1 (mod h), R = L - 2 ks valid rows. A row
// next to the band's edge sees zero planes beyond it and goes wrong after the
// first step, its neighbour after the second: after ks steps the middle R
// rows are exact, and only those are stored. The random digits of every step
// come from the draw pass (they depend on the step, not on the state), so
// the edge rows repeat only the gather and the circuit.
// ---------------------------------------------------------------------------

// Global row of tile-local row i of tile t (R valid rows, ks halo rows).
LG_HD inline uint32_t lg_tile_row(uint32_t t, uint32_t i, uint32_t R, uint32_t ks, uint32_t h) {
  const uint32_t base = (uint32_t)(((uint64_t)t * R + i) % h);
  const uint32_t back = ks % h;
  return base >= back ? base - back : base + h - back;
}

// Inflow re-seed of word (y, j < iw) from its sampled planes bits[7]: the
// fluid cells of the inflow columns take them (src/lattice.rs:270-285).
LG_HD inline void lg_inflow_apply(const LgGeom& G, uint32_t j, uint64_t sm, uint64_t tail,
                                  const uint64_t bits[7], uint64_t res[7]) {
  const uint64_t apply = lg_inlet_apply(G, j, ~sm & tail);
  LG_UNROLL
  for (int d = 0; d < 7; ++d) res[d] = (res[d] & ~apply) | (bits[d] & apply);
}

