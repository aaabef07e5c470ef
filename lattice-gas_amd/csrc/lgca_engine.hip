// MI355X FHP lattice-gas engine: HIP kernels and the C ABI of include/lgca.h.
//
// Hot path per step (default: one launch of lg_step2_kernel): draw work --
// jump each word's position in its CPU band stream, draw the collision
// columns of its cells and, in inflow words, the re-seeded states, for the
// next step -- and collide work -- gather six direction planes with in-word
// shifts and cross-word carries, evaluate the generated collision circuit with
// the digits drawn one launch earlier, bounce back on solid sites, overwrite
// the inflow cells, mask the ragged tail. The alternatives (one fused kernel
// per step; draws on a second stream) give the same cells. The per-step
// stream jump tables come from lg_tables_kernel, launched for the next
// basis_steps steps whenever a call's steps leave the ones it built last.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lgca_internal.h"
#include "lgca_jump.hpp"

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Stream tables for `nsteps` consecutive steps from t0: for step t0 + k and
// stream s (s < nbands: collision band s of src/lattice.rs:242; s == nbands:
// the inflow stream of src/lattice.rs:276, seeded with the incremented step)
// write the 4-bit window tables of the freshly seeded generator. One thread
// per (step, stream, window group g): it walks to T^(4g) s0 and fills the
// 16 entries of its group.
__global__ void __launch_bounds__(256) lg_tables_kernel(uint64_t seed, uint64_t t0, uint32_t nsteps,
                                                        uint32_t nbands, uint64_t* __restrict tabs) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t per = nbands + 1;
  if (g >= nsteps * per * 16) return;
  uint32_t grp = g & 15u, ks = g >> 4;
  uint32_t k = ks / per, s = ks - k * per;
  uint64_t t = t0 + k;
  uint64_t st = lg_rng_init(s < nbands ? lg_band_seed(seed, t, s) : lg_inlet_seed(seed, t + 1));
  for (uint32_t i = 0; i < 4 * grp; ++i) st = lg_xs(st);
  uint64_t b0 = st, b1 = lg_xs(b0), b2 = lg_xs(b1), b3 = lg_xs(b2);
  uint64_t* out = tabs + (size_t)ks * LG_WTAB + grp * 16;
  for (uint32_t v = 0; v < 16; ++v) {
    uint64_t e = 0;
    if (v & 1u) e ^= b0;
    if (v & 2u) e ^= b1;
    if (v & 4u) e ^= b2;
    if (v & 8u) e ^= b3;
    out[v] = e;
  }
}

// One step. Thread g steps unit g % NH of word order[g / NH]: the static
// order puts the inflow words first (as the reference does,
// src/gpu.rs:212-228), then the words holding a solid site or a ragged row
// end, then the rest, so waves of ordinary words never diverge.
template <int NH>
__device__ __forceinline__ void lg_step_body(const LgGeom& G, const uint64_t* __restrict src,
                                             uint64_t* __restrict dst,
                                             const uint64_t* __restrict solid,
                                             const uint64_t* __restrict cjump,
                                             const uint32_t* __restrict row_band,
                                             const uint64_t* __restrict ijump,
                                             const uint32_t* __restrict order,
                                             const uint64_t* __restrict tabs) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total * NH) return;
  const uint32_t e = order[NH == 1 ? g : g >> 1];
  lg_step_unit<NH>(G, src, dst, solid, cjump, row_band, ijump, tabs, e >> 16, e & 0xFFFFu,
                   NH == 1 ? 0u : (g & 1u));
}

// Both unit modes run at 4 waves per SIMD (the register budget that fits);
// 5 for half-word units measured slower on the 4096 x 2304 plate workload.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
lg_step_kernel_w(LgGeom G, const uint64_t* __restrict src, uint64_t* __restrict dst,
                 const uint64_t* __restrict solid, const uint64_t* __restrict cjump,
                 const uint32_t* __restrict row_band, const uint64_t* __restrict ijump,
                 const uint32_t* __restrict order, const uint64_t* __restrict tabs) {
  lg_step_body<1>(G, src, dst, solid, cjump, row_band, ijump, order, tabs);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
lg_step_kernel_h(LgGeom G, const uint64_t* __restrict src, uint64_t* __restrict dst,
                 const uint64_t* __restrict solid, const uint64_t* __restrict cjump,
                 const uint32_t* __restrict row_band, const uint64_t* __restrict ijump,
                 const uint32_t* __restrict order, const uint64_t* __restrict tabs) {
  lg_step_body<2>(G, src, dst, solid, cjump, row_band, ijump, order, tabs);
}

// Two-pass step, one launch per step: collide blocks update the lattice for
// step t with the digits drawn earlier (kcur/icur), draw blocks make step
// t + 1's digits (knext/inext, from tabs_next) -- the draws never read the
// lattice, so the two kinds of work share a launch and the draw blocks keep
// the VALUs busy while collide blocks wait on memory. Roles alternate by
// block index while both remain, so every CU gets a mix; a launch may carry
// only one of them (cblocks or dblocks zero). Collide unit u is word u
// (NH = 1) or half u % 2 of word u / 2 (NH = 2) in row-major order; draw unit
// v is word order[v] (DNH = 1) or half v % 2 of word order[v / 2] (DNH = 2).
template <int NH, int DNH>
__global__ void __launch_bounds__(256)
lg_step2_kernel(LgGeom G, const uint64_t* __restrict src, uint64_t* __restrict dst,
                const uint64_t* __restrict solid, const uint64_t* __restrict cjump,
                const uint32_t* __restrict row_band, const uint64_t* __restrict ijump,
                const uint32_t* __restrict order, const uint64_t* __restrict tabs_next,
                const uint32_t* __restrict kcur, const uint32_t* __restrict icur,
                uint32_t* __restrict knext, uint32_t* __restrict inext, uint32_t cblocks,
                uint32_t dblocks) {
  uint32_t b, draw;
  lg_step2_role(blockIdx.x, cblocks, dblocks, &draw, &b);
  const size_t kstride = 2 * (size_t)G.total, istride = 2 * (size_t)G.iw * G.h;
  const uint32_t u = b * blockDim.x + threadIdx.x;
  if (draw) {
    if (u >= G.total * DNH) return;
    const uint32_t e = order[DNH == 1 ? u : u >> 1];
    lg_draw_unit<DNH>(G, solid, cjump, row_band, ijump, tabs_next, e >> 16, e & 0xFFFFu,
                      DNH == 1 ? 0u : (u & 1u), knext, kstride, inext, istride);
  } else {
    if (u >= G.total * NH) return;
    const uint32_t word = NH == 1 ? u : u >> 1;
    const uint32_t y = word / G.wpr;
    lg_collide_unit<NH>(G, src, dst, solid, kcur, kstride, icur, istride, y, word - y * G.wpr,
                        NH == 1 ? 0u : (u & 1u));
  }
}
// Draw-ahead pipeline, draw side: step blockIdx.y of this launch for draw
// unit u (word order[u], DNH = 1, or half u % 2 of word order[u / 2]);
// tabs0, kb0, ib0: the launch's first step's tables and slots.
template <int DNH>
__global__ void __launch_bounds__(256)
lg_draw_kernel(LgGeom G, const uint64_t* __restrict solid, const uint64_t* __restrict cjump,
               const uint32_t* __restrict row_band, const uint64_t* __restrict ijump,
               const uint32_t* __restrict order, const uint64_t* __restrict tabs0, uint32_t per,
               uint32_t* __restrict kb0, uint32_t* __restrict ib0) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= G.total * DNH) return;
  const size_t kstride = 2 * (size_t)G.total, istride = 2 * (size_t)G.iw * G.h;
  const uint32_t s = blockIdx.y;
  const uint32_t e = order[DNH == 1 ? u : u >> 1];
  lg_draw_unit<DNH>(G, solid, cjump, row_band, ijump, tabs0 + (size_t)s * per * LG_WTAB, e >> 16,
                    e & 0xFFFFu, DNH == 1 ? 0u : (u & 1u), kb0 + (size_t)s * LG_KDIGITS * kstride,
                    kstride, ib0 + (size_t)s * 7 * istride, istride);
}

// Draw-ahead pipeline, collide side: one step from its digit and inflow slot.
template <int NH>
__global__ void __launch_bounds__(256)
lg_collide_kernel(LgGeom G, const uint64_t* __restrict src, uint64_t* __restrict dst,
                  const uint64_t* __restrict solid, const uint32_t* __restrict kc,
                  const uint32_t* __restrict ic) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= G.total * NH) return;
  const uint32_t word = NH == 1 ? u : u >> 1;
  const uint32_t y = word / G.wpr;
  lg_collide_unit<NH>(G, src, dst, solid, kc, 2 * (size_t)G.total, ic, 2 * (size_t)G.iw * G.h, y,
                      word - y * G.wpr, NH == 1 ? 0u : (u & 1u));
}

// Word of the lane to the left (j - 1) or right (j + 1) of this one in the
// row, periodic in x. WAVEROW: the row is exactly the 64 lanes of the wave
// (64 words), so a whole-wave DPP rotation by one lane is the neighbour
// (wave_ror:1: lane n reads lane n - 1, lane 0 reads lane 63; wave_rol:1 the
// other way). Otherwise ds_bpermute from the byte address of the neighbour
// lane (addr), which wraps at the row's own length.
template <bool WAVEROW>
__device__ __forceinline__ uint64_t lg_lane_left(uint64_t x, int addr) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (WAVEROW) {
    lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x13C, 0xF, 0xF, false);
    hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x13C, 0xF, 0xF, false);
  } else {
    lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)lo);
    hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)hi);
  }
  return ((uint64_t)hi << 32) | lo;
}
template <bool WAVEROW>
__device__ __forceinline__ uint64_t lg_lane_right(uint64_t x, int addr) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (WAVEROW) {
    lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x134, 0xF, 0xF, false);
    hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x134, 0xF, 0xF, false);
  } else {
    lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)lo);
    hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)hi);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Temporally blocked collide pass (rows <= 64 words): workgroup t steps tile
// t of the lattice (see lg_tile_row) ks steps with the digit and inflow
// planes of those steps (kb0 / ib0: the first step's slot, the next steps'
// slots following), then stores its R valid rows. Lane j holds word j of its
// rows (lanes past the row's end idle along), wave w holds tile rows w M ..
// w M + M - 1 in registers: rows inside a wave meet in the lane, the wave's
// first and last rows meet the neighbouring waves' through LDS (one barrier
// per step, two exchange slots alternating), words meet their left and right
// neighbours across lanes. Every global access is to a row y < h (mod h) and
// word j < wpr; the launch has ceil(h / R) workgroups of L / M waves and
// L / M * 4 KB of dynamic LDS.
template <int M, bool WAVEROW>
__global__ void __launch_bounds__(1024)
lg_collide_k_kernel(LgGeom G, const uint64_t* __restrict src, uint64_t* __restrict dst,
                    const uint64_t* __restrict solid, const uint32_t* __restrict kb0,
                    const uint32_t* __restrict ib0, uint32_t ks, uint32_t R) {
  extern __shared__ uint64_t xch[];  // [2 slots][NW waves][4 planes][64 lanes]
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const uint32_t L = NW * M;
  const size_t T = G.total, kstride = 2 * T, istride = 2 * (size_t)G.iw * G.h;
  const bool live = lane < G.wpr;
  const uint32_t j = live ? lane : 0u;
  const uint32_t cl = (j == 0) ? G.lastbit : 63u, cr = (j == G.lastword) ? G.lastbit : 63u;
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  const int addr_l = (int)(4u * (live ? (j == 0 ? G.wpr - 1 : j - 1) : lane));
  const int addr_r = (int)(4u * (live ? (j + 1 == G.wpr ? 0u : j + 1) : lane));
  const int nk = G.rest ? LG_COLLIDE_REST_KBITS : LG_COLLIDE_NOREST_KBITS;

  uint64_t P[M][7], S[M];
  uint32_t Y[M];
  LG_UNROLL
  for (int m = 0; m < M; ++m) {
    Y[m] = lg_tile_row(blockIdx.x, w * M + m, R, ks, G.h);
    const size_t word = (size_t)Y[m] * G.wpr + j;
    LG_UNROLL
    for (int d = 0; d < 7; ++d) P[m][d] = src[(size_t)d * T + word];
    S[m] = solid[word];
  }
  for (uint32_t s = 0; s < ks; ++s) {
    const uint32_t* __restrict kb = kb0 + (size_t)s * LG_KDIGITS * kstride;
    const uint32_t* __restrict ib = ib0 + (size_t)s * 7 * istride;
    // this step's digits of every row (in flight across the exchange)
    uint64_t K[M][LG_KDIGITS];
    LG_UNROLL
    for (int m = 0; m < M; ++m) {
      const size_t word = (size_t)Y[m] * G.wpr + j;
      LG_UNROLL
      for (int b = 0; b < LG_KDIGITS; ++b)
        K[m][b] = b < nk ? reinterpret_cast<const uint64_t*>(kb + (size_t)b * kstride)[word] : 0ull;
    }
    uint64_t* slot = xch + (size_t)(s & 1u) * NW * 256;
    slot[(w * 4 + 0) * 64 + lane] = P[0][4];
    slot[(w * 4 + 1) * 64 + lane] = P[0][5];
    slot[(w * 4 + 2) * 64 + lane] = P[M - 1][1];
    slot[(w * 4 + 3) * 64 + lane] = P[M - 1][2];
    __syncthreads();
    // planes 1, 2 of the row above this wave's first row, 4, 5 of the row
    // below its last (zero beyond the tile's edges)
    uint64_t a1 = 0, a2 = 0, dn4 = 0, dn5 = 0;
    if (w > 0) {
      a1 = slot[((w - 1) * 4 + 2) * 64 + lane];
      a2 = slot[((w - 1) * 4 + 3) * 64 + lane];
    }
    if (w + 1 < NW) {
      dn4 = slot[((w + 1) * 4 + 0) * 64 + lane];
      dn5 = slot[((w + 1) * 4 + 1) * 64 + lane];
    }
    LG_UNROLL
    for (int m = 0; m < M; ++m) {
      const uint64_t a4 = (m + 1 < M) ? P[m + 1][4] : dn4;
      const uint64_t a5 = (m + 1 < M) ? P[m + 1][5] : dn5;
      const uint32_t odd = Y[m] & 1u;  // the same for the whole wave
      const uint64_t a0l = lg_lane_left<WAVEROW>(P[m][0], addr_l);
      const uint64_t a3r = lg_lane_right<WAVEROW>(P[m][3], addr_r);
      uint64_t a1l = 0, a5l = 0, a2r = 0, a4r = 0;
      if (odd) {
        a2r = lg_lane_right<WAVEROW>(a2, addr_r);
        a4r = lg_lane_right<WAVEROW>(a4, addr_r);
      } else {
        a1l = lg_lane_left<WAVEROW>(a1, addr_l);
        a5l = lg_lane_left<WAVEROW>(a5, addr_l);
      }
      uint64_t n[7];
      lg_gather(odd, cl, cr, P[m][0], a0l, a1, a1l, a2, a2r, P[m][3], a3r, a4, a4r, a5, a5l,
                P[m][6], n);
      const uint64_t o1 = P[m][1], o2 = P[m][2];  // row m's old planes 1, 2, for row m + 1
      uint32_t Kl[LG_KDIGITS], Kh[LG_KDIGITS];
      LG_UNROLL
      for (int b = 0; b < LG_KDIGITS; ++b) {
        Kl[b] = (uint32_t)K[m][b];
        Kh[b] = (uint32_t)(K[m][b] >> 32);
      }
      uint64_t res[7];
      lg_collide_word(G, n, S[m], Kl, Kh, res);
      if (j < G.iw) {
        const uint64_t* __restrict iin =
            reinterpret_cast<const uint64_t*>(ib) + (size_t)Y[m] * G.iw + j;
        uint64_t bits[7];
        LG_UNROLL
        for (int d = 0; d < 7; ++d) bits[d] = iin[(size_t)d * (istride / 2)];
        lg_inflow_apply(G, j, S[m], tail, bits, res);
      }
      LG_UNROLL
      for (int d = 0; d < 7; ++d) P[m][d] = res[d] & tail;
      a1 = o1;
      a2 = o2;
    }
  }
  if (!live) return;
  LG_UNROLL
  for (int m = 0; m < M; ++m) {
    const uint32_t i = w * M + m;
    // valid rows only, and each global row once (the last tile may run past h)
    if (i < ks || i >= L - ks || (uint64_t)blockIdx.x * R + (i - ks) >= G.h) continue;
    const size_t word = (size_t)Y[m] * G.wpr + j;
    LG_UNROLL
    for (int d = 0; d < 7; ++d) dst[(size_t)d * T + word] = P[m][d];
  }
}

// Bytes -> planes (GpuLattice::load, src/gpu.rs:572-610), one thread per word.
__global__ void __launch_bounds__(256) lg_pack_kernel(LgGeom G, const uint8_t* __restrict cells,
                                                      uint64_t* __restrict planes,
                                                      uint64_t* __restrict solid) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total) return;
  uint32_t y = g / G.wpr, j = g - y * G.wpr;
  uint32_t x0 = j * 64, n = min(64u, G.w - x0);
  const uint8_t* row = cells + (size_t)y * G.w + x0;
  uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t x = 0; x < n; ++x) {
    uint32_t c = row[x];
#pragma unroll
    for (int d = 0; d < 8; ++d) p[d] |= (uint64_t)((c >> d) & 1u) << x;
  }
#pragma unroll
  for (int d = 0; d < 7; ++d) planes[(size_t)d * G.total + g] = p[d];
  solid[g] = p[7];
}

// Planes -> bytes (GpuLattice::store, src/gpu.rs:630-652).
__global__ void __launch_bounds__(256) lg_unpack_kernel(LgGeom G, const uint64_t* __restrict planes,
                                                        const uint64_t* __restrict solid,
                                                        uint8_t* __restrict cells) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total) return;
  uint32_t y = g / G.wpr, j = g - y * G.wpr;
  uint32_t x0 = j * 64, n = min(64u, G.w - x0);
  uint64_t p[8];
#pragma unroll
  for (int d = 0; d < 7; ++d) p[d] = planes[(size_t)d * G.total + g];
  p[7] = solid[g];
  uint8_t* row = cells + (size_t)y * G.w + x0;
  for (uint32_t x = 0; x < n; ++x) {
    uint32_t c = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) c |= (uint32_t)((p[d] >> x) & 1u) << d;
    row[x] = (uint8_t)c;
  }
}

// Coarse-graining, one thread per block of the field grid: exact integer sums
// over fluid cells (src/moments.rs:37-65 FLUID, src/render.rs:93-106), then
// the f32 running average of src/render.rs:107-115 with every operation
// rounded on its own (no fused multiply-add), as the CPU computes it.
// One wavefront per coarse-grain block (four per workgroup).
__global__ void __launch_bounds__(256) lg_sample_kernel(
    LgGeom G, const uint64_t* __restrict planes, const uint64_t* __restrict solid, uint32_t bx,
    uint32_t by, uint32_t bw, uint32_t bh, float alpha, float* __restrict ux, float* __restrict uy,
    float* __restrict rho, float* __restrict sol, int64_t* __restrict sums) {
  const uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= bw * bh) return;  // whole wavefronts leave together
  lg_sample_block_wave(G, planes, solid, bx, by, bw, alpha, ux, uy, rho, sol, sums, g,
                       threadIdx.x & 63u);
}

// Obstacles (src/lattice.rs:210-220): every site whose f32 position
// (src/hex.rs:52-55) satisfies the predicate becomes SOLID_BIT with no
// particles. kind 0: box x0 <= px <= x1 && y0 <= py <= y1 (src/main.rs:584-588,
// tests/golden.rs:78-83); kind 1: disc (px-a)^2 + (py-b)^2 <= r^2
// (src/main.rs:589-592) with a = x0, b = x1, r = y0.
__global__ void __launch_bounds__(256) lg_solid_kernel(LgGeom G, int kind, float x0, float x1,
                                                       float y0, float y1,
                                                       uint64_t* __restrict planes,
                                                       uint64_t* __restrict solid) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total) return;
  uint32_t y = g / G.wpr, j = g - y * G.wpr;
  const float SQRT3_2 = 0.8660254f;
  float py = __fmul_rn((float)y, SQRT3_2);
  uint64_t m = 0;
  uint32_t n = min(64u, G.w - j * 64);
  for (uint32_t b = 0; b < n; ++b) {
    float px = __fadd_rn((float)(j * 64 + b), (y & 1u) ? 0.5f : 0.0f);
    bool in;
    if (kind == 0) {
      in = px >= x0 && px <= x1 && py >= y0 && py <= y1;
    } else {
      float dx = __fsub_rn(px, x0), dy = __fsub_rn(py, x1);
      in = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)) <= __fmul_rn(y0, y0);
    }
    m |= (uint64_t)in << b;
  }
  if (!m) return;
  solid[g] |= m;
#pragma unroll
  for (int d = 0; d < 7; ++d) planes[(size_t)d * G.total + g] &= ~m;
}

// Lattice::init_equilibrium (src/lattice.rs:197-208): every fluid cell in
// index order takes Equilibrium::sample from the stream seeded with
// seed ^ 0xA5A51234; the word's start in that stream is given by `cinit`.
__global__ void __launch_bounds__(256) lg_init_kernel(LgGeom G, const uint64_t* __restrict cinit,
                                                      const uint64_t* __restrict basis,
                                                      uint64_t* __restrict planes,
                                                      const uint64_t* __restrict solid) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total) return;
  uint32_t j = g % G.wpr;
  uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  uint64_t fm = ~solid[g] & tail;
  uint64_t res[7] = {0, 0, 0, 0, 0, 0, 0};
  if (fm) lg_inlet_word(G, lg_jump(cinit[g], basis), fm, res);
#pragma unroll
  for (int d = 0; d < 7; ++d) planes[(size_t)d * G.total + g] = res[d];
}

// Lattice::init_equilibrium with a profile: cell (x, y) is drawn from the
// equilibrium of column x (axis 0) or of row y (axis 1), thresholds thr[i * 7 +
// d]; same stream and order as lg_init_kernel (the column form is how
// src/transport.rs:238-251 builds its start states).
__global__ void __launch_bounds__(256) lg_init_profile_kernel(LgGeom G, const uint64_t* __restrict cinit,
                                                              const uint64_t* __restrict basis,
                                                              const uint32_t* __restrict thr, int axis,
                                                              uint64_t* __restrict planes,
                                                              const uint64_t* __restrict solid) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.total) return;
  uint32_t y = g / G.wpr, j = g - y * G.wpr;
  uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  uint64_t fm = ~solid[g] & tail;
  uint64_t bits[7] = {0, 0, 0, 0, 0, 0, 0};
  uint64_t st = fm ? lg_jump(cinit[g], basis) : 0;
  for (uint32_t x = 0; x < 64; ++x) {
    if (!((fm >> x) & 1ull)) continue;
    const uint32_t* t = thr + (size_t)(axis == 0 ? j * 64 + x : y) * 7;
    uint64_t s = 0;
    for (int d = 0; d < 6; d += 2) {
      st = lg_xs(st);
      uint64_t r = st * LG_RNG_MUL;
      s |= (uint64_t)((uint32_t)(r >> 40) < t[d]) << d;
      s |= (uint64_t)(((uint32_t)(r >> 16) & 0xFFFFFFu) < t[d + 1]) << (d + 1);
    }
    if (G.rest) {
      st = lg_xs(st);
      uint64_t r = st * LG_RNG_MUL;
      s |= (uint64_t)((uint32_t)(r >> 40) < t[6]) << 6;
    }
    for (int d = 0; d < 7; ++d) bits[d] |= ((s >> d) & 1ull) << x;
  }
#pragma unroll
  for (int d = 0; d < 7; ++d) planes[(size_t)d * G.total + g] = bits[d];
}

// Per-row and per-column moment sums over every cell, obstacle sites
// included (moments::WITH_WALLS, src/moments.rs:74): row y -> (mass, px2),
// column x -> (mass, py2). Exact integers; transport::row_ux / column_uy
// (src/transport.rs:51-91) derive their f32 velocities from these.
__global__ void __launch_bounds__(256) lg_rows_kernel(LgGeom G, const uint64_t* __restrict planes,
                                                      int64_t* __restrict out) {
  uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= G.h) return;
  const int CX2[6] = {2, 1, -1, -2, -1, 1};
  int64_t mass = 0, px = 0;
  for (uint32_t j = 0; j < G.wpr; ++j) {
    size_t word = (size_t)y * G.wpr + j;
    for (int d = 0; d < 6; ++d) {
      int64_t c = __popcll(planes[(size_t)d * G.total + word]);
      mass += c;
      px += CX2[d] * c;
    }
    mass += __popcll(planes[(size_t)6 * G.total + word]);
  }
  out[2 * y] = mass;
  out[2 * y + 1] = px;
}

__global__ void __launch_bounds__(256) lg_cols_kernel(LgGeom G, const uint64_t* __restrict planes,
                                                      int64_t* __restrict out) {
  uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= G.w) return;
  const int CY2[6] = {0, 1, 1, 0, -1, -1};
  uint32_t j = x >> 6, b = x & 63u;
  int64_t mass = 0, py = 0;
  for (uint32_t y = 0; y < G.h; ++y) {
    size_t word = (size_t)y * G.wpr + j;
    for (int d = 0; d < 7; ++d) {
      int64_t c = (int64_t)((planes[(size_t)d * G.total + word] >> b) & 1ull);
      mass += c;
      if (d < 6) py += CY2[d] * c;
    }
  }
  out[2 * x] = mass;
  out[2 * x + 1] = py;
}

// Total particles, obstacle sites included (src/lattice.rs:299-312): one
// thread per word, a 64-bit atomic per wave.
__global__ void __launch_bounds__(256) lg_count_kernel(LgGeom G, const uint64_t* __restrict planes,
                                                       unsigned long long* __restrict out) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  if (g < G.total) {
#pragma unroll
    for (int d = 0; d < 7; ++d) c += __popcll(planes[(size_t)d * G.total + g]);
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// Exhaustive check of the generated circuit as the device evaluates it: one
// thread per (state, k), k < stride; lanes must agree (out 0xFF otherwise).
__global__ void __launch_bounds__(256) lg_verify_kernel(int rest, uint8_t* __restrict out) {
  const int ns = rest ? 128 : 64;
  const int stride = rest ? LG_COLLIDE_REST_STRIDE : LG_COLLIDE_NOREST_STRIDE;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns * stride) return;
  int s = i / stride, k = i - s * stride;
  uint64_t n[7], o[7];
  for (int d = 0; d < 7; ++d) n[d] = ((s >> d) & 1) ? ~0ull : 0ull;
  if (rest) {
    uint64_t K[LG_COLLIDE_REST_KBITS];
    for (int b = 0; b < LG_COLLIDE_REST_KBITS; ++b) K[b] = ((k >> b) & 1) ? ~0ull : 0ull;
    lg_collide_rest<uint64_t>(n, K, o);
  } else {
    uint64_t K[LG_COLLIDE_NOREST_KBITS];
    for (int b = 0; b < LG_COLLIDE_NOREST_KBITS; ++b) K[b] = ((k >> b) & 1) ? ~0ull : 0ull;
    lg_collide_norest<uint64_t>(n, K, o);
  }
  int r = 0, bad = 0;
  for (int d = 0; d < 7; ++d) {
    if (o[d] == ~0ull) r |= 1 << d;
    else if (o[d] != 0) bad = 1;
  }
  out[i] = bad ? 0xFF : (uint8_t)r;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace lgca_internal {

thread_local std::string g_err;

namespace {

struct MemCache {
  std::mutex mu;
  struct Block {
    int dev;       // -1: page-locked host memory
    size_t bytes;  // the size class
  };
  std::unordered_map<void*, Block> live;
  std::map<std::pair<int, size_t>, std::vector<void*>> free;  // (device, class) -> blocks
  size_t cached = 0;                                          // bytes on the free lists
};
MemCache& mem_cache() {
  static MemCache* c = new MemCache();  // never destroyed: handles may outlive static teardown
  return *c;
}
// Cached bytes above which a released block is freed at once instead.
constexpr size_t kCacheMax = 16ull << 30;

size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes < (1u << 20)) {
    size_t c = 512;
    while (c < bytes) c <<= 1;
    return c;
  }
  const size_t m = 2u << 20;
  return (bytes + m - 1) / m * m;
}

void raw_free(int dev, void* p) {
  if (dev < 0)
    (void)hipHostFree(p);
  else
    (void)hipFree(p);
}

// Empty the free lists of `dev` (every device and the host when dev is -2).
size_t trim_locked(MemCache& c, int dev) {
  size_t got = 0;
  for (auto it = c.free.begin(); it != c.free.end();) {
    if (dev != -2 && it->first.first != dev) {
      ++it;
      continue;
    }
    for (void* p : it->second) {
      raw_free(it->first.first, p);
      got += it->first.second;
    }
    c.cached -= it->first.second * it->second.size();
    it = c.free.erase(it);
  }
  return got;
}

hipError_t cached_alloc(void** p, size_t bytes, bool host) {
  int dev = -1;
  if (!host) {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
  }
  const size_t cls = size_class(bytes);
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.free.find({dev, cls});
  if (it != c.free.end() && !it->second.empty()) {
    *p = it->second.back();
    it->second.pop_back();
    c.cached -= cls;
    c.live[*p] = {dev, cls};
    return hipSuccess;
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    hipError_t e = host ? hipHostMalloc(p, cls, hipHostMallocDefault) : hipMalloc(p, cls);
    if (e == hipSuccess) {
      c.live[*p] = {dev, cls};
      return hipSuccess;
    }
    (void)hipGetLastError();
    if (attempt == 0 && trim_locked(c, dev) == 0) return e;  // nothing cached to give back
  }
  return hipErrorOutOfMemory;
}

void cached_release(void* p) {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) return;
  const MemCache::Block b = it->second;
  c.live.erase(it);
  if (c.cached + b.bytes > kCacheMax) {
    raw_free(b.dev, p);
    return;
  }
  c.free[{b.dev, b.bytes}].push_back(p);
  c.cached += b.bytes;
}

}  // namespace

hipError_t dev_alloc(void** p, size_t bytes) { return cached_alloc(p, bytes, false); }
void dev_release(void* p) { cached_release(p); }
hipError_t host_alloc(void** p, size_t bytes) { return cached_alloc(p, bytes, true); }
void host_release(void* p) { cached_release(p); }
size_t trim_caches() {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  return trim_locked(c, -2);
}

// Equilibrium thresholds exactly as src/lattice.rs:59-82 computes them: each
// f32 operation rounded on its own, in source order.
uint32_t threshold_f32(float p) {
  const float ONE = 16777216.0f;
  volatile float prod = p * ONE;
  float t = std::ceil((float)prod);
  if (t >= ONE) return (uint32_t)ONE;
  if (t > 0.0f) return (uint32_t)t;
  return 0;
}

void equilibrium_thresholds(float density, float ux, float uy, bool rest, uint32_t out[7]) {
  const float SQRT3_2 = 0.8660254f;
  const float CXF[6] = {1.0f, 0.5f, -0.5f, -1.0f, -0.5f, 0.5f};
  const float CYF[6] = {0.0f, SQRT3_2, SQRT3_2, 0.0f, -SQRT3_2, -SQRT3_2};
  const float bias = rest ? (2.0f * 7.0f / 6.0f) : 2.0f;
  for (int d = 0; d < 6; ++d) {
    volatile float a = CXF[d] * ux;
    volatile float b = CYF[d] * uy;
    volatile float dot = a + b;
    volatile float sc = bias * dot;
    volatile float one_plus = 1.0f + sc;
    volatile float p = density * one_plus;
    out[d] = threshold_f32(p);
  }
  out[6] = rest ? threshold_f32(density) : 0;
}


void destroy_lattice(lg_lattice* g) {
  (void)hipStreamSynchronize(g->stream);
  if (g->dstream) (void)hipStreamSynchronize(g->dstream);
  band_forget_stream(g);
  g->a.release();
  g->b.release();
  g->solid.release();
  g->cjump.release();
  g->ijump.release();
  g->icell.release();
  g->basis.release();
  g->row_band.release();
  g->order.release();
  g->bytes.release();
  g->band_halo.release();
  g->band_flags.release();
  g->band_err_dev.release();
  g->band_part.release();
  g->band_bcnt.release();
  g->band_field.release();
  if (g->band_err_host) host_release(g->band_err_host);
  g->band_err_host = nullptr;
  g->fux.release();
  g->fuy.release();
  g->frho.release();
  g->fsol.release();
  g->fsums.release();
  g->counter.release();
  for (hipEvent_t e : g->tev) (void)hipEventDestroy(e);
  for (int q = 0; q < 2; ++q) {
    if (g->ev_drawn[q]) (void)hipEventDestroy(g->ev_drawn[q]);
    if (g->ev_used[q]) (void)hipEventDestroy(g->ev_used[q]);
  }
  if (g->ev_start) (void)hipEventDestroy(g->ev_start);
  if (g->dstream) (void)hipStreamDestroy(g->dstream);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // namespace lgca_internal

using namespace lgca_internal;


namespace {


int rebuild_tables(lg_t* g, const std::vector<uint64_t>& solid_host) {
  std::vector<uint64_t> cj, ij;
  std::vector<uint32_t> rb, ord;
  lgca::build_tables(g->G.w, g->G.h, g->bands, g->G.inlet_cols, g->G.rest != 0, solid_host, cj, rb,
                     ij, ord);
  if (lgca::charpoly_low() == 0) return fail(LG_EDEVICE, "jump polynomial not found");
  HIP_TRY(g->cjump.alloc(cj.size()));
  HIP_TRY(g->row_band.alloc(rb.size()));
  HIP_TRY(g->ijump.alloc(std::max<size_t>(ij.size(), 1)));
  HIP_TRY(hipMemcpyAsync(g->cjump.p, cj.data(), cj.size() * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(g->row_band.p, rb.data(), rb.size() * 4, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(g->order.alloc(ord.size()));
  HIP_TRY(hipMemcpyAsync(g->order.p, ord.data(), ord.size() * 4, hipMemcpyHostToDevice, g->stream));
  if (!ij.empty())
    HIP_TRY(hipMemcpyAsync(g->ijump.p, ij.data(), ij.size() * 8, hipMemcpyHostToDevice, g->stream));
  std::vector<uint64_t> ic;
  if (g->G.inlet_cols > 0 && g->G.inlet_cols <= 64) {  // resident bands' per-cell inflow jumps
    lgca::build_inflow_cells(g->G.w, g->G.h, g->G.inlet_cols, g->G.rest != 0, solid_host, ic);
    HIP_TRY(g->icell.alloc(ic.size()));
    HIP_TRY(hipMemcpyAsync(g->icell.p, ic.data(), ic.size() * 8, hipMemcpyHostToDevice, g->stream));
  }
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->tables_dirty = false;
  g->tables_version += 1;  // the resident bands' lead jumps follow on their next launch
  return LG_OK;
}

// The solid plane of a byte lattice, for the jump tables (host side).
std::vector<uint64_t> solid_plane(const LgGeom& G, const uint8_t* cells) {
  std::vector<uint64_t> s((size_t)G.total, 0);
  for (uint32_t y = 0; y < G.h; ++y) {
    const uint8_t* row = cells + (size_t)y * G.w;
    for (uint32_t x = 0; x < G.w; ++x)
      if (row[x] & 0x80u) s[(size_t)y * G.wpr + x / 64] |= 1ull << (x % 64);
  }
  return s;
}

}  // namespace

int lgca_internal::download_solid(lg_t* g, std::vector<uint64_t>& s) {
  s.resize((size_t)g->G.total);
  HIP_TRY(hipMemcpyAsync(s.data(), g->solid.p, s.size() * 8, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return LG_OK;
}

namespace {

int add_solid(lg_t* g, int kind, float a, float b, float c, float d) {
  hipLaunchKernelGGL(lg_solid_kernel, dim3(blocks_for(g->G.total)), dim3(256), 0, g->stream, g->G,
                     kind, a, b, c, d, g->a.p, g->solid.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->loaded = true;
  g->tables_dirty = true;
  return LG_OK;
}

int sample_now(lg_t* g, float alpha) {
  // a lattice smaller than one block has an empty field (src/render.rs:73-75)
  if (!g->has_field || g->bw == 0 || g->bh == 0) return LG_OK;
  hipLaunchKernelGGL(lg_sample_kernel, dim3(blocks_for((size_t)g->bw * g->bh, 4)), dim3(256), 0,
                     g->stream, g->G, g->a.p, g->solid.p, g->bx, g->by, g->bw, g->bh, alpha,
                     g->fux.p, g->fuy.p, g->frho.p, g->fsol.p, g->fsums.p);
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

}  // namespace

// Digit and inflow planes of the two-pass step: two slots of one step each
// (the step being collided, the step being drawn), sized for the current
// inflow width.
// With draw-ahead: two groups of `ahead` steps each.
int lgca_internal::ensure_draw_buffers(lg_t* g) {
  const size_t nhalf = (size_t)2 * g->G.total, nih = (size_t)2 * g->G.iw * g->G.h;
  const size_t slots = 2 * (size_t)std::max<uint32_t>(1, g->ahead);
  if (g->kbuf.n >= slots * LG_KDIGITS * nhalf && g->ibuf.n >= std::max<size_t>(slots * 7 * nih, 1))
    return LG_OK;
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(g->kbuf.alloc(slots * LG_KDIGITS * nhalf));
  HIP_TRY(g->ibuf.alloc(std::max<size_t>(slots * 7 * nih, 1)));
  return LG_OK;
}

namespace {

// One two-pass launch: collide the current step from digit slot `cur` if
// `collide`, draw into slot cur ^ 1 from tabs if `draw`.
void launch_step2(lg_t* g, uint32_t cur, bool collide, bool draw, const uint64_t* tabs) {
  const size_t nhalf = (size_t)2 * g->G.total, nih = (size_t)2 * g->G.iw * g->G.h;
  const uint32_t cblocks = collide ? blocks_for((size_t)g->G.total * g->units) : 0;
  const uint32_t dblocks = draw ? blocks_for((size_t)g->G.total * g->draw_units) : 0;
  const uint32_t* kc = g->kbuf.p + (size_t)cur * LG_KDIGITS * nhalf;
  const uint32_t* ic = g->ibuf.p + (size_t)cur * 7 * nih;
  uint32_t* kn = g->kbuf.p + (size_t)(cur ^ 1u) * LG_KDIGITS * nhalf;
  uint32_t* in = g->ibuf.p + (size_t)(cur ^ 1u) * 7 * nih;
  auto k = g->units == 2 ? (g->draw_units == 2 ? lg_step2_kernel<2, 2> : lg_step2_kernel<2, 1>)
                         : (g->draw_units == 2 ? lg_step2_kernel<1, 2> : lg_step2_kernel<1, 1>);
  hipLaunchKernelGGL(k, dim3(cblocks + dblocks), dim3(256), 0, g->stream, g->G, g->a.p, g->b.p,
                     g->solid.p, g->cjump.p, g->row_band.p, g->ijump.p, g->order.p, tabs, kc, ic,
                     kn, in, cblocks, dblocks);
}

// Steps per collide launch: the temporal depth where it applies (draw-ahead
// pipeline, rows of at most 64 words), else 1.
uint32_t collide_depth(const lg_t* g) {
  if (g->tdepth < 2 || g->ahead == 0 || g->G.wpr > 64) return 1;
  return g->tdepth;
}

// Tile shape of a temporally blocked collide of ks steps: rows per wave M,
// waves per workgroup NW (<= 16), valid rows per tile R. By default about one
// tile per CU (R = ceil(h / CUs)), grown to fill whole waves.
struct TileShape {
  uint32_t M, NW, R, ntiles;
};
TileShape tile_shape(const lg_t* g, uint32_t ks) {
  TileShape t;
  t.M = 2;
  uint32_t R = g->trows ? g->trows : (g->G.h + (uint32_t)g->ncu - 1) / (uint32_t)g->ncu;
  R = std::max<uint32_t>(R, 1);
  uint32_t L = R + 2 * ks;
  t.NW = (L + t.M - 1) / t.M;
  if (t.NW > 16) t.NW = 16;
  L = t.NW * t.M;
  t.R = L - 2 * ks;  // >= 1: ks <= LG_MAX_TDEPTH keeps 2 ks < 32
  t.ntiles = (g->G.h + t.R - 1) / t.R;
  return t;
}

// One temporally blocked collide launch of ks steps from the digit and
// inflow slots kc / ic (steps following each other) on the handle's stream.
void launch_collide_k(lg_t* g, const uint32_t* kc, const uint32_t* ic, uint32_t ks) {
  const TileShape t = tile_shape(g, ks);
  const dim3 grid(t.ntiles), block(64 * t.NW);
  const size_t lds = (size_t)t.NW * 4096;
  if (g->G.wpr == 64)
    hipLaunchKernelGGL((lg_collide_k_kernel<2, true>), grid, block, lds, g->stream, g->G, g->a.p,
                       g->b.p, g->solid.p, kc, ic, ks, t.R);
  else
    hipLaunchKernelGGL((lg_collide_k_kernel<2, false>), grid, block, lds, g->stream, g->G, g->a.p,
                       g->b.p, g->solid.p, kc, ic, ks, t.R);
}

// The second stream and the events of the draw-ahead pipeline (once).
int ensure_ahead_stream(lg_t* g) {
  if (g->dstream) return LG_OK;
  HIP_TRY(hipStreamCreateWithFlags(&g->dstream, hipStreamNonBlocking));
  for (int q = 0; q < 2; ++q) {
    HIP_TRY(hipEventCreateWithFlags(&g->ev_drawn[q], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&g->ev_used[q], hipEventDisableTiming));
    HIP_TRY(hipEventRecord(g->ev_used[q], g->stream));  // both groups start free
  }
  HIP_TRY(hipEventCreate(&g->ev_start));
  return LG_OK;
}

// Draw-ahead: launch the draws of steps [s0, s0 + nd) of the chunk whose
// tables are in basis into the next slot group on dstream, then their
// collides on the main stream.
int enqueue_ahead_group(lg_t* g, const uint64_t* tabs, uint32_t s0, uint32_t nd) {
  const uint32_t per = g->G.nbands + 1;
  const size_t nhalf = (size_t)2 * g->G.total, nih = (size_t)2 * g->G.iw * g->G.h;
  const uint32_t q = (uint32_t)(g->dgroup++ & 1u);
  uint32_t* kb = g->kbuf.p + (size_t)q * g->ahead * LG_KDIGITS * nhalf;
  uint32_t* ib = g->ibuf.p + (size_t)q * g->ahead * 7 * nih;
  HIP_TRY(hipStreamWaitEvent(g->dstream, g->ev_used[q], 0));
  const dim3 dg(blocks_for((size_t)g->G.total * g->draw_units), nd);
  const uint64_t* tabs0 = tabs + (size_t)s0 * per * LG_WTAB;
  if (g->draw_units == 2)
    hipLaunchKernelGGL(lg_draw_kernel<2>, dg, dim3(256), 0, g->dstream, g->G, g->solid.p,
                       g->cjump.p, g->row_band.p, g->ijump.p, g->order.p, tabs0, per, kb, ib);
  else
    hipLaunchKernelGGL(lg_draw_kernel<1>, dg, dim3(256), 0, g->dstream, g->G, g->solid.p,
                       g->cjump.p, g->row_band.p, g->ijump.p, g->order.p, tabs0, per, kb, ib);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(g->ev_drawn[q], g->dstream));
  HIP_TRY(hipStreamWaitEvent(g->stream, g->ev_drawn[q], 0));
  const uint32_t kd = collide_depth(g);
  for (uint32_t k = 0; k < nd;) {
    const uint32_t* kc = kb + (size_t)k * LG_KDIGITS * nhalf;
    const uint32_t* ic = ib + (size_t)k * 7 * nih;
    const uint32_t ks = std::min(kd, nd - k);
    if (ks >= 2) {
      launch_collide_k(g, kc, ic, ks);
      std::swap(g->a.p, g->b.p);
      g->steps += ks;
      k += ks;
      continue;
    }
    if (g->units == 2)
      hipLaunchKernelGGL(lg_collide_kernel<2>, dim3(blocks_for(nhalf)), dim3(256), 0, g->stream,
                         g->G, g->a.p, g->b.p, g->solid.p, kc, ic);
    else
      hipLaunchKernelGGL(lg_collide_kernel<1>, dim3(blocks_for(g->G.total)), dim3(256), 0,
                         g->stream, g->G, g->a.p, g->b.p, g->solid.p, kc, ic);
    std::swap(g->a.p, g->b.p);
    g->steps += 1;
    k += 1;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(g->ev_used[q], g->stream));
  return LG_OK;
}

// At least `count` timing events for lg_advance_timed. They release at
// device scope: recorded around launches they then add no write-back of the
// caches to system scope (the call's results reach the host through copies
// that synchronise by themselves).
int timing_events(lg_t* g, size_t count) {
  while (g->tev.size() < count) {
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
    g->tev.push_back(e);
  }
  return LG_OK;
}

// The stream tables of steps [t, t + nsteps) (nsteps <= basis_steps): from
// the ones `basis` holds when they cover those steps, else the tables of the
// next basis_steps steps from t, launched on `stream` first. Tables depend
// only on the seed and the step, so they stay good across calls; work that
// reads them is ordered before a later refill by the streams' order (a refill
// on the draw stream waits for the main stream's earlier work, and the main
// stream's later work waits for the draws).
const uint64_t* stream_tables(lg_t* g, uint64_t t, uint32_t nsteps, hipStream_t stream) {
  const uint32_t per = g->G.nbands + 1;
  if (g->tab_n == 0 || t < g->tab_t0 || t + nsteps > g->tab_t0 + g->tab_n) {
    hipLaunchKernelGGL(lg_tables_kernel, dim3(blocks_for((size_t)g->basis_steps * per * 16)),
                       dim3(256), 0, stream, g->seed, t, g->basis_steps, g->G.nbands, g->basis.p);
    g->tab_t0 = t;
    g->tab_n = g->basis_steps;
  }
  return g->basis.p + (size_t)(t - g->tab_t0) * per * LG_WTAB;
}

// Enqueue `n` steps (and samples every `every`), no synchronisation.
int enqueue_steps(lg_t* g, uint64_t n, uint64_t every, float alpha, bool timed) {
  int trc = ensure_tables(g);
  if (trc) return trc;
  resolve_pipeline(g);
  uint64_t left = n, since = 0;
  const uint32_t per = g->G.nbands + 1;
  if (timed) g->tev_used = 0;
  uint32_t passes = step_passes(g);
  uint32_t brows = 0, bnb = 0;
  // a call whose first chunk is too short to repay the band launch takes
  // the other pipelines unless the caller asked for resident bands
  const uint64_t first = std::min<uint64_t>(n, every > 0 ? every : n);
  const bool resident = resident_active(g, &brows, &bnb);
  uint32_t units = g->units;  // cells per lane of the fused step
  if (g->resident != 1 && first <= 2 && g->G.inlet_cols == 0 && g->passes == 0 &&
      g->tdepth < 2 && g->auto_ahead && g->G.total < LG_AUTO_FUSED_WORDS) {
    // one or two steps without inflow: the fused step's call is the
    // shortest, with 32-cell lanes unless the caller set the lane width
    // (2048 x 1280, one step: 17.7 us against 18.7 with 64-cell lanes, 20.3
    // two-pass, 29 resident; two steps: 24.6, 26.5, 28.2, 33.0)
    passes = 1;
    if (g->auto_units) units = 2;
  } else if (resident &&
             (g->resident == 1 || (double)first * band_gain_us(g, brows) >= LG_BAND_LAUNCH_US)) {
    // resident bands: one launch per chunk (it makes its own stream tables);
    // if the runtime declines the launch, the rest of the call takes the
    // other pipelines below
    while (left > 0) {
      uint64_t chunk = std::min<uint64_t>(left, LG_BAND_MAX_CHUNK);
      if (every > 0) chunk = std::min<uint64_t>(chunk, every - since);
      if (timed) {
        if (int rc = timing_events(g, 2 * (g->tev_used + 1))) return rc;
        HIP_TRY(hipEventRecord(g->tev[2 * g->tev_used], g->stream));
      }
      // a chunk followed by a sample coarse-grains its final state in the
      // same launch where the field's blocks fit
      const bool sample_after = every > 0 && (since + chunk >= every || left == chunk);
      const bool fused = sample_after && band_sample_fits(g);
      const int brc = launch_band(g, (uint32_t)chunk, fused, alpha);
      if (brc == kBandRefused) break;
      if (brc) return brc;
      if (timed) HIP_TRY(hipEventRecord(g->tev[2 * g->tev_used++ + 1], g->stream));
      std::swap(g->a.p, g->b.p);
      g->steps += chunk;
      left -= chunk;
      since += chunk;
      if (every > 0 && (since >= every || left == 0)) {
        since = 0;
        if (!fused) {
          int rc = sample_now(g, alpha);
          if (rc) return rc;
        }
      }
    }
    if (left == 0) return LG_OK;
  }
  const bool ahead = passes == 2 && g->ahead > 0;
  if (ahead) {
    int rc = ensure_ahead_stream(g);
    if (!rc) rc = ensure_draw_buffers(g);
    if (rc) return rc;
    // the draw stream starts with this call's work on the main stream
    HIP_TRY(hipEventRecord(g->ev_start, g->stream));
    HIP_TRY(hipStreamWaitEvent(g->dstream, g->ev_start, 0));
  }
  while (left > 0) {
    const uint32_t extra = (passes == 2 && !ahead) ? 1u : 0u;  // see the two-pass branch
    uint64_t chunk = std::min<uint64_t>(left, g->basis_steps - extra);
    if (every > 0) chunk = std::min<uint64_t>(chunk, every - since);
    const uint64_t* tabs =
        stream_tables(g, g->steps, (uint32_t)chunk + extra, ahead ? g->dstream : g->stream);
    HIP_TRY(hipGetLastError());
    if (timed && ahead) {
      // one interval for the whole call: the draw stream runs ahead of the
      // main stream's step boundaries
      if (g->tev_used == 0) {
        if (int rc = timing_events(g, 2)) return rc;
        HIP_TRY(hipEventRecord(g->tev[0], g->stream));
        g->tev_used = 1;
      }
    } else if (timed) {
      if (int rc = timing_events(g, 2 * (g->tev_used + 1))) return rc;
      HIP_TRY(hipEventRecord(g->tev[2 * g->tev_used], g->stream));
    }
    if (ahead) {
      for (uint64_t s0 = 0; s0 < chunk; s0 += g->ahead) {
        int rc = enqueue_ahead_group(g, tabs, (uint32_t)s0,
                                     (uint32_t)std::min<uint64_t>(g->ahead, chunk - s0));
        if (rc) return rc;
      }
    } else if (passes == 1) {
      for (uint64_t k = 0; k < chunk; ++k) {
        if (units == 2)
          hipLaunchKernelGGL(lg_step_kernel_h, dim3(blocks_for((size_t)g->G.total * 2)), dim3(256),
                             0, g->stream, g->G, g->a.p, g->b.p, g->solid.p, g->cjump.p,
                             g->row_band.p, g->ijump.p, g->order.p,
                             tabs + (size_t)k * per * LG_WTAB);
        else
          hipLaunchKernelGGL(lg_step_kernel_w, dim3(blocks_for(g->G.total)), dim3(256), 0,
                             g->stream, g->G, g->a.p, g->b.p, g->solid.p, g->cjump.p,
                             g->row_band.p, g->ijump.p, g->order.p,
                             tabs + (size_t)k * per * LG_WTAB);
        std::swap(g->a.p, g->b.p);
        g->steps += 1;
      }
    } else {
      // The tables of this chunk cover one step more than it takes, so its
      // last launch can draw the next chunk's first step.
      int rc = ensure_draw_buffers(g);
      if (rc) return rc;
      if (left == n) {  // the call's first step: draw it on its own
        launch_step2(g, 1, false, true, tabs);
        g->kslot = 0;
      }
      for (uint64_t k = 0; k < chunk; ++k) {
        const bool more = k + 1 < chunk || left > chunk;
        launch_step2(g, g->kslot, true, more, tabs + (size_t)(k + 1) * per * LG_WTAB);
        g->kslot ^= 1u;
        std::swap(g->a.p, g->b.p);
        g->steps += 1;
      }
    }
    HIP_TRY(hipGetLastError());
    if (timed && ahead) {
      if (left == chunk) HIP_TRY(hipEventRecord(g->tev[1], g->stream));  // the call's last chunk
    } else if (timed) {
      HIP_TRY(hipEventRecord(g->tev[2 * g->tev_used++ + 1], g->stream));
    }
    left -= chunk;
    since += chunk;
    if (every > 0 && (since >= every || left == 0)) {
      since = 0;
      int rc = sample_now(g, alpha);
      if (rc) return rc;
    }
  }
  return LG_OK;
}

}  // namespace

bool lgca_internal::resident_active(const lg_t* g, uint32_t* rows, uint32_t* nb) {
  if (g->resident == 0 || g->band_refused) return false;
  if (!band_eligible(g, rows, nb)) return false;
  if (g->resident == 1) return true;
  // by the size rule: unless the caller chose another pipeline, and where
  // the bands' step beats the two-pass step (calls also need a first chunk
  // long enough to repay the band launch: enqueue_steps)
  if (g->passes != 0 || g->tdepth >= 2 || !g->auto_ahead) return false;
  return band_gain_us(g, *rows) > 0.0;
}

double lgca_internal::band_gain_us(const lg_t* g, uint32_t rows) {
  const double resident = 4.0 + 0.47 * rows;
  const double two_pass = 4.8 + 0.055e-3 * (double)g->G.total + (g->G.inlet_cols > 0 ? 1.3 : 0.0);
  return two_pass - resident;
}

uint32_t lgca_internal::step_passes(const lg_t* g) {
  if (g->passes) return g->passes;
  if (g->tdepth >= 2 && g->G.wpr <= 64) return 2u;  // temporal blocking is a two-pass pipeline
  return g->G.total >= LG_AUTO_FUSED_WORDS ? 1u : 2u;
}

void lgca_internal::resolve_pipeline(lg_t* g) {
  const bool small = g->G.total < LG_SMALL_WORDS, inflow = g->G.inlet_cols > 0;
  if (g->auto_draw_units) g->draw_units = small && !inflow ? 2u : 1u;
  if (g->auto_ahead) g->ahead = small && inflow ? 8u : 0u;
}

int lgca_internal::ensure_tables(lg_t* g) {
  if (!g->tables_dirty) return LG_OK;
  std::vector<uint64_t> s;
  int rc = download_solid(g, s);
  if (rc) return rc;
  return rebuild_tables(g, s);
}

extern "C" {

const char* lg_last_error(void) { return g_err.c_str(); }

int lg_trim_memory(size_t* freed) {
  const size_t got = trim_caches();
  if (freed) *freed = got;
  return LG_OK;
}

lg_t* lg_create(size_t w, size_t h, int rest, uint64_t seed, size_t bands) {
  if (h % 2 != 0) {
    fail(LG_EINVAL, "row count must be even for the lattice to wrap cleanly in y");
    return nullptr;
  }
  if (w <= 4 || h <= 4) {
    fail(LG_EINVAL, "lattice is too small");
    return nullptr;
  }
  if (h >= 65536 || (w + 63) / 64 >= 65536 || ((w + 63) / 64) * h * 2 >= (1ull << 32)) {
    fail(LG_EINVAL, "lattice is too large (h and words per row must stay below 65536)");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    fail(LG_EDEVICE, "no HIP device on this machine");
    return nullptr;
  }
  auto* g = new lg_lattice();
  auto bail = [&](const std::string& m) -> lg_t* {
    fail(LG_EDEVICE, m);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return nullptr;
  };
  if (hipGetDevice(&g->device) != hipSuccess) return bail("hipGetDevice failed");
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, g->device) == hipSuccess) g->ncu = prop.multiProcessorCount;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return bail("hipStreamCreate failed");
  LgGeom& G = g->G;
  G.w = (uint32_t)w;
  G.h = (uint32_t)h;
  G.wpr = (uint32_t)((w + 63) / 64);
  G.total = G.wpr * G.h;
  G.lastword = (uint32_t)((w - 1) / 64);
  G.lastbit = (uint32_t)((w - 1) % 64);
  G.tailmask = G.lastbit == 63 ? ~0ull : ((1ull << (G.lastbit + 1)) - 1);
  G.iw = 0;
  G.inlet_cols = 0;
  G.rest = rest ? 1u : 0u;
  g->bands = (uint32_t)std::max<size_t>(1, std::min<size_t>(bands, h));
  G.nbands = lgca::n_bands(G.h, g->bands);
  g->seed = seed;
  const size_t per = G.nbands + 1;
  // At least 3 steps: a batch chunk takes at least 2 steps and the two-pass
  // step keeps one table spare for the next chunk's first draw.
  g->basis_steps = (uint32_t)std::max<size_t>(
      3, std::min<size_t>(256, (32u << 20) / (per * LG_WTAB * 8)));
  if (g->a.alloc((size_t)7 * G.total) || g->b.alloc((size_t)7 * G.total) ||
      g->solid.alloc(G.total) || g->basis.alloc((size_t)g->basis_steps * per * LG_WTAB) ||
      g->counter.alloc(1))
    return bail("hipMalloc failed for the lattice planes");
  if (hipMemsetAsync(g->a.p, 0, g->a.n * 8, g->stream) ||
      hipMemsetAsync(g->b.p, 0, g->b.n * 8, g->stream) ||
      hipMemsetAsync(g->solid.p, 0, g->solid.n * 8, g->stream) ||
      hipStreamSynchronize(g->stream))
    return bail("hipMemset failed");
  if (lgca::charpoly_low() == 0) return bail("jump polynomial not found");
  g->tables_dirty = true;  // built before the first step
  return g;
}

void lg_destroy(lg_t* g) {
  if (!g) return;
  if (g->batch_refs) {  // still a batch member: the last lg_batch_destroy frees it
    g->destroyed = true;
    return;
  }
  destroy_lattice(g);
}

int lg_device_name(lg_t* g, char* buf, size_t len) {
  if (!g || !buf || len == 0) return fail(LG_EINVAL, "null argument");
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, g->device));
  // Some drivers leave the marketing name empty; the architecture is always set.
  if (p.name[0])
    std::snprintf(buf, len, "%s (%s)", p.name, p.gcnArchName);
  else
    std::snprintf(buf, len, "AMD GPU %s, %d CUs", p.gcnArchName, p.multiProcessorCount);
  return LG_OK;
}

int lg_set_inlet(lg_t* g, size_t cols, const uint32_t thresh[7]) {
  if (!g || (!thresh && cols)) return fail(LG_EINVAL, "null argument");
  for (int d = 0; d < 7; ++d) {
    uint32_t t = thresh ? thresh[d] : 0;
    if (t > (1u << 24)) return fail(LG_EINVAL, "threshold above 2^24");
    g->G.thresh[d] = t;
  }
  uint32_t c = (uint32_t)std::min<size_t>(cols, g->G.w);
  g->G.inlet_cols = c;
  g->G.iw = (c + 63) / 64;
  g->tables_dirty = true;  // the inflow tables and word order change
  return LG_OK;
}

int lg_set_inlet_flow(lg_t* g, size_t cols, float density, float ux, float uy) {
  if (!g) return fail(LG_EINVAL, "null handle");
  uint32_t t[7];
  equilibrium_thresholds(density, ux, uy, g->G.rest != 0, t);
  return lg_set_inlet(g, cols, t);
}

int lg_attach_field(lg_t* g, size_t bx, size_t by) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  if (bx == 0 || by == 0) return fail(LG_EINVAL, "block size must be positive");
  g->bx = (uint32_t)bx;
  g->by = (uint32_t)by;
  g->bw = g->G.w / g->bx;
  g->bh = g->G.h / g->by;
  size_t nb = std::max<size_t>((size_t)g->bw * g->bh, 1);
  HIP_TRY(g->fux.alloc(nb));
  HIP_TRY(g->fuy.alloc(nb));
  HIP_TRY(g->frho.alloc(nb));
  HIP_TRY(g->fsol.alloc(nb));
  HIP_TRY(g->fsums.alloc(nb * 4));
  HIP_TRY(hipMemsetAsync(g->fux.p, 0, nb * 4, g->stream));
  HIP_TRY(hipMemsetAsync(g->fuy.p, 0, nb * 4, g->stream));
  HIP_TRY(hipMemsetAsync(g->frho.p, 0, nb * 4, g->stream));
  HIP_TRY(hipMemsetAsync(g->fsol.p, 0, nb * 4, g->stream));
  HIP_TRY(hipMemsetAsync(g->fsums.p, 0, nb * 32, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->has_field = true;
  return LG_OK;
}

int lg_load(lg_t* g, const uint8_t* cells, size_t n) {
  if (!g || !cells) return fail(LG_EINVAL, "null argument");
  g->poisoned = false;  // a reload sets the whole state
  if (int rc = settle(g)) return rc;
  if (n != (size_t)g->G.w * g->G.h) return fail(LG_EINVAL, "wrong number of cells");
  if (g->bytes.n < n) HIP_TRY(g->bytes.alloc(n));
  HIP_TRY(hipMemcpyAsync(g->bytes.p, cells, n, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(lg_pack_kernel, dim3(blocks_for(g->G.total)), dim3(256), 0, g->stream, g->G,
                     g->bytes.p, g->a.p, g->solid.p);
  HIP_TRY(hipGetLastError());
  int rc = rebuild_tables(g, solid_plane(g->G, cells));
  if (rc) return rc;
  g->loaded = true;
  return LG_OK;
}

int lg_restart(lg_t* g, const uint8_t* cells, size_t n, uint64_t seed) {
  int rc = lg_load(g, cells, n);
  if (rc) return rc;
  g->seed = seed;
  g->steps = 0;
  g->tab_n = 0;  // tables of the old seed
  return LG_OK;
}

int lg_store(lg_t* g, uint8_t* cells, size_t n) {
  if (!g || !cells) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  if (n != (size_t)g->G.w * g->G.h) return fail(LG_EINVAL, "wrong number of cells");
  if (g->bytes.n < n) HIP_TRY(g->bytes.alloc(n));
  hipLaunchKernelGGL(lg_unpack_kernel, dim3(blocks_for(g->G.total)), dim3(256), 0, g->stream,
                     g->G, g->a.p, g->solid.p, g->bytes.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(cells, g->bytes.p, n, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return LG_OK;
}

int lg_advance(lg_t* g, uint64_t n) { return lg_advance_sampling(g, n, 0, 0.0f); }

int lg_advance_sampling(lg_t* g, uint64_t n, uint64_t every, float alpha) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  int rc = enqueue_steps(g, n, every, alpha, false);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  return check_band(g);
}

int lg_advance_async(lg_t* g, uint64_t n) { return lg_advance_sampling_async(g, n, 0, 0.0f); }

int lg_advance_sampling_async(lg_t* g, uint64_t n, uint64_t every, float alpha) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  return enqueue_steps(g, n, every, alpha, false);
}

// Work queued on the handle's stream is all of it: the draw-ahead stream's
// launches are each waited for by a later launch on the handle's stream.
int lg_poll(lg_t* g) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (g->after) {  // an async batch advance stepped it last
    hipError_t q = hipEventQuery(g->after);
    if (q == hipErrorNotReady) return 0;
    if (q != hipSuccess) return fail(LG_EDEVICE, std::string("hipEventQuery: ") + hipGetErrorString(q));
    g->after = nullptr;
  }
  hipError_t e = hipStreamQuery(g->stream);
  if (e == hipSuccess) {
    if (int rc = check_band(g)) return rc;
    return 1;
  }
  if (e == hipErrorNotReady) return 0;
  return fail(LG_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int lg_wait(lg_t* g) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  return check_band(g);
}

int lg_advance_sampling_timed(lg_t* g, uint64_t n, uint64_t every, float alpha, float* ms) {
  if (!g || !ms) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  if (n == 0) {
    *ms = 0.0f;
    return LG_OK;
  }
  int rc = enqueue_steps(g, n, every, alpha, true);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  if ((rc = check_band(g))) return rc;
  float total = 0.0f;
  for (size_t i = 0; i < g->tev_used; ++i) {
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, g->tev[2 * i], g->tev[2 * i + 1]));
    total += t;
  }
  *ms = total;
  return LG_OK;
}

int lg_advance_timed(lg_t* g, uint64_t n, float* ms) {
  return lg_advance_sampling_timed(g, n, 0, 0.0f, ms);
}

int lg_sample_field_async(lg_t* g, float alpha) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  return sample_now(g, alpha);
}

int lg_sample_field(lg_t* g, float alpha) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  int rc = sample_now(g, alpha);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  return LG_OK;
}

int lg_field_dims(lg_t* g, size_t* bw, size_t* bh, size_t* bx, size_t* by) {
  if (!g || !bw || !bh || !bx || !by) return fail(LG_EINVAL, "null argument");
  if (!g->has_field) return fail(LG_ESTATE, "no field attached to this lattice");
  *bw = g->bw, *bh = g->bh, *bx = g->bx, *by = g->by;
  return LG_OK;
}

int lg_read_field(lg_t* g, float* ux, float* uy, float* rho, float* solid, size_t nblocks) {
  if (!g || !ux || !uy || !rho || !solid) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  if (!g->has_field) return fail(LG_ESTATE, "no field attached to this lattice");
  if (nblocks != (size_t)g->bw * g->bh)
    return fail(LG_EINVAL, "the field on the device has a different block grid");
  if (nblocks == 0) return LG_OK;
  HIP_TRY(hipMemcpyAsync(ux, g->fux.p, nblocks * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(uy, g->fuy.p, nblocks * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(rho, g->frho.p, nblocks * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(solid, g->fsol.p, nblocks * 4, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return LG_OK;
}

int lg_read_block_sums(lg_t* g, int64_t* mass, int64_t* px2, int64_t* py2, int64_t* nsolid,
                       size_t nblocks) {
  if (!g || !mass || !px2 || !py2 || !nsolid) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  if (!g->has_field) return fail(LG_ESTATE, "no field attached to this lattice");
  if (nblocks != (size_t)g->bw * g->bh)
    return fail(LG_EINVAL, "the field on the device has a different block grid");
  std::vector<int64_t> s(nblocks * 4);
  if (nblocks) {
    HIP_TRY(hipMemcpyAsync(s.data(), g->fsums.p, s.size() * 8, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  for (size_t i = 0; i < nblocks; ++i) {
    mass[i] = s[i * 4 + 0];
    px2[i] = s[i * 4 + 1];
    py2[i] = s[i * 4 + 2];
    nsolid[i] = s[i * 4 + 3];
  }
  return LG_OK;
}

int lg_total_particles(lg_t* g, uint64_t* out) {
  if (!g || !out) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  HIP_TRY(hipMemsetAsync(g->counter.p, 0, 8, g->stream));
  hipLaunchKernelGGL(lg_count_kernel, dim3(blocks_for(g->G.total)), dim3(256), 0, g->stream,
                     g->G, g->a.p, g->counter.p);
  HIP_TRY(hipGetLastError());
  unsigned long long v = 0;
  HIP_TRY(hipMemcpyAsync(&v, g->counter.p, 8, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  *out = v;
  return LG_OK;
}

int lg_dims(lg_t* g, size_t* w, size_t* h) {
  if (!g || !w || !h) return fail(LG_EINVAL, "null argument");
  *w = g->G.w;
  *h = g->G.h;
  return LG_OK;
}

uint64_t lg_steps(lg_t* g) { return g ? g->steps : 0; }

int lg_set_work_unit(lg_t* g, uint32_t cells_per_lane) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (cells_per_lane != 64 && cells_per_lane != 32)
    return fail(LG_EINVAL, "cells per lane must be 64 or 32");
  g->units = cells_per_lane == 64 ? 1u : 2u;
  g->auto_units = false;
  return LG_OK;
}

int lg_set_draw_ahead(lg_t* g, uint32_t steps) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (steps > 64) return fail(LG_EINVAL, "draw-ahead is at most 64 steps");
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->ahead = steps;
  g->auto_ahead = false;
  return LG_OK;
}

int lg_set_temporal_depth(lg_t* g, uint32_t steps, uint32_t rows_per_tile) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (steps < 1 || steps > LG_MAX_TDEPTH)
    return fail(LG_EINVAL, "temporal depth must be 1.." + std::to_string(LG_MAX_TDEPTH));
  if (rows_per_tile > 32 - 2 * steps)
    return fail(LG_EINVAL, "a tile holds at most 32 rows with its halo (2 x depth rows)");
  if (steps >= 2 && g->G.wpr > 64)
    return fail(LG_EINVAL, "temporal blocking needs rows of at most 64 words (4096 cells)");
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (g->dstream) HIP_TRY(hipStreamSynchronize(g->dstream));
  g->tdepth = steps;
  g->trows = rows_per_tile;
  if (steps >= 2 && g->ahead < steps) {  // blocks of draws at least as deep
    g->ahead = steps;
    g->auto_ahead = false;
  }
  return LG_OK;
}

int lg_set_resident(lg_t* g, uint32_t on) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (on > 2) return fail(LG_EINVAL, "resident must be 0 (off), 1 (on) or 2 (by the size rule)");
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->resident = on;
  g->band_refused = false;  // a declined cooperative launch is tried again
  return LG_OK;
}

int lg_get_temporal_depth(lg_t* g, uint32_t* steps, uint32_t* rows_per_tile) {
  if (!g || !steps || !rows_per_tile) return fail(LG_EINVAL, "null argument");
  *steps = g->tdepth;
  *rows_per_tile = g->trows;
  return LG_OK;
}

int lg_autotune(lg_t* g, uint64_t budget_steps, uint64_t* used_steps) {
  if (!g || !used_steps) return fail(LG_EINVAL, "null argument");
  *used_steps = 0;
  if (int rc = settle(g)) return rc;
  struct Cand {
    uint32_t resident, passes;
    int ahead;   // -1: by the size rule
    uint32_t depth;
  };
  // the size rule first, so that a budget that runs out keeps it
  std::vector<Cand> cands = {{2, 0, -1, 1}, {0, 2, 0, 1}, {0, 2, 16, 1}, {0, 1, -1, 1}};
  uint32_t rows = 0, nb = 0;
  if (band_eligible(g, &rows, &nb)) cands.insert(cands.begin() + 1, Cand{1, 0, -1, 1});
  if (g->G.wpr <= 64)
    for (uint32_t k : {2u, 4u, 8u}) cands.push_back(Cand{0, 2, 16, k});
  auto apply = [&](const Cand& c) -> int {
    int rc = lg_set_temporal_depth(g, 1, 0);
    if (!rc) rc = lg_set_passes(g, c.passes);  // 0 returns the draw unit and draw-ahead to the rule
    if (!rc && c.ahead >= 0) rc = lg_set_draw_ahead(g, (uint32_t)c.ahead);
    if (!rc && c.depth > 1) rc = lg_set_temporal_depth(g, c.depth, 0);
    if (!rc) rc = lg_set_resident(g, c.resident);
    return rc;
  };
  const uint64_t trial = 32;
  constexpr int kTimed = 3;
  double best = 1e300;
  size_t bi = 0;
  uint64_t used = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    if (budget_steps && used + (1 + kTimed) * trial > budget_steps) break;
    if (int rc = apply(cands[i])) return rc;
    if (int rc = lg_advance(g, trial)) return rc;  // warm: streams, buffers, code
    double dt = 1e300;
    for (int r = 0; r < kTimed; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      if (int rc = lg_advance(g, trial)) return rc;
      dt = std::min(dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    used += (1 + kTimed) * trial;
    if (dt < best) {
      best = dt;
      bi = i;
    }
  }
  *used_steps = used;
  return apply(cands[bi]);
}

int lg_get_resident(lg_t* g, uint32_t* rows_per_band, uint32_t* bands) {
  if (!g || !rows_per_band || !bands) return fail(LG_EINVAL, "null argument");
  uint32_t r = 0, n = 0;
  const bool ok = resident_active(g, &r, &n);
  *rows_per_band = ok ? r : 0;
  *bands = ok ? n : 0;
  return ok ? 1 : 0;
}

int lg_set_draw_unit(lg_t* g, uint32_t cells_per_lane) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (cells_per_lane != 64 && cells_per_lane != 32 && cells_per_lane != 0)
    return fail(LG_EINVAL, "cells per lane must be 64, 32 or 0 (by size)");
  g->auto_draw_units = cells_per_lane == 0;
  if (cells_per_lane) g->draw_units = cells_per_lane == 64 ? 1u : 2u;
  return LG_OK;
}

int lg_set_passes(lg_t* g, uint32_t passes) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (passes > 2) return fail(LG_EINVAL, "passes per step must be 0 (by size), 1 or 2");
  g->passes = passes;
  if (passes == 0) g->auto_draw_units = g->auto_ahead = true;  // the whole pipeline by size
  return LG_OK;
}

int lg_get_pipeline(lg_t* g, uint32_t* passes, uint32_t* cells_per_lane,
                    uint32_t* draw_cells_per_lane, uint32_t* draw_ahead) {
  if (!g || !passes || !cells_per_lane || !draw_cells_per_lane || !draw_ahead)
    return fail(LG_EINVAL, "null argument");
  resolve_pipeline(g);
  *passes = step_passes(g);
  *cells_per_lane = g->units == 1 ? 64u : 32u;
  *draw_cells_per_lane = g->draw_units == 1 ? 64u : 32u;
  *draw_ahead = *passes == 2 ? g->ahead : 0u;
  return LG_OK;
}

int lg_set_steps(lg_t* g, uint64_t steps) {
  if (!g) return fail(LG_EINVAL, "null handle");
  g->steps = steps;
  return LG_OK;
}

int lg_add_solid_box(lg_t* g, float x0, float x1, float y0, float y1) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  return add_solid(g, 0, x0, x1, y0, y1);
}

int lg_add_solid_disc(lg_t* g, float cx, float cy, float r) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  return add_solid(g, 1, cx, cy, r, 0.0f);
}

}  // extern "C"

namespace {

// Start of every word in the init stream (seed ^ 0xA5A51234, 4 draws per
// fluid cell, 3 without rest, src/lattice.rs:197-208) and the stream basis.
int init_stream(lg_t* g, std::vector<uint64_t>& cinit, uint64_t basis[64]) {
  std::vector<uint64_t> s;
  int rc = download_solid(g, s);
  if (rc) return rc;
  const LgGeom& G0 = g->G;
  const uint64_t p = lgca::charpoly_low();
  const uint32_t per = G0.rest ? 4 : 3;
  cinit.assign((size_t)G0.total, 0);
  uint64_t c = 1;
  for (uint32_t y = 0; y < G0.h; ++y)
    for (uint32_t j = 0; j < G0.wpr; ++j) {
      size_t word = (size_t)y * G0.wpr + j;
      cinit[word] = c;
      uint64_t tail = (j == G0.lastword) ? G0.tailmask : ~0ull;
      c = lgca::poly_mulz(c, (uint32_t)__builtin_popcountll(~s[word] & tail) * per, p);
    }
  uint64_t st = lg_rng_init(g->seed ^ 0xA5A51234ull);
  for (int i = 0; i < 64; ++i) {
    basis[i] = st;
    st = lg_xs(st);
  }
  return LG_OK;
}

}  // namespace

extern "C" {

int lg_init_equilibrium(lg_t* g, float density, float ux, float uy) {
  if (!g) return fail(LG_EINVAL, "null handle");
  if (int rc = settle(g)) return rc;
  std::vector<uint64_t> cinit;
  uint64_t basis[64];
  int rc = init_stream(g, cinit, basis);
  if (rc) return rc;
  LgGeom G = g->G;
  equilibrium_thresholds(density, ux, uy, G.rest != 0, G.thresh);
  DevBuf<uint64_t> dc, db;
  HIP_TRY(dc.alloc(cinit.size()));
  HIP_TRY(db.alloc(64));
  HIP_TRY(hipMemcpyAsync(dc.p, cinit.data(), cinit.size() * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(db.p, basis, 64 * 8, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(lg_init_kernel, dim3(blocks_for(G.total)), dim3(256), 0, g->stream, G, dc.p,
                     db.p, g->a.p, g->solid.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->loaded = true;
  return LG_OK;
}

int lg_init_profile(lg_t* g, int axis, float density, const float* ux, const float* uy, size_t n) {
  if (!g || !ux || !uy) return fail(LG_EINVAL, "null argument");
  if (int rc = settle(g)) return rc;
  if (axis != 0 && axis != 1) return fail(LG_EINVAL, "axis must be 0 (columns) or 1 (rows)");
  if (n != (axis == 0 ? g->G.w : g->G.h)) return fail(LG_EINVAL, "profile length differs from the lattice");
  std::vector<uint32_t> thr(n * 7);
  for (size_t i = 0; i < n; ++i)
    equilibrium_thresholds(density, ux[i], uy[i], g->G.rest != 0, &thr[i * 7]);
  std::vector<uint64_t> cinit;
  uint64_t basis[64];
  int rc = init_stream(g, cinit, basis);
  if (rc) return rc;
  DevBuf<uint64_t> dc, db;
  DevBuf<uint32_t> dt;
  HIP_TRY(dc.alloc(cinit.size()));
  HIP_TRY(db.alloc(64));
  HIP_TRY(dt.alloc(thr.size()));
  HIP_TRY(hipMemcpyAsync(dc.p, cinit.data(), cinit.size() * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(db.p, basis, 64 * 8, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(dt.p, thr.data(), thr.size() * 4, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(lg_init_profile_kernel, dim3(blocks_for(g->G.total)), dim3(256), 0, g->stream,
                     g->G, dc.p, db.p, dt.p, axis, g->a.p, g->solid.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->loaded = true;
  return LG_OK;
}

// ---------------------------------------------------------------------------
// transport::measure on the engine (src/transport.rs:24-319)
// ---------------------------------------------------------------------------

}  // extern "C"

namespace {

// Equilibrium::sample (src/lattice.rs:102-115) on the host, for the initial
// states of the measurement's realisations.
uint8_t eq_sample(const uint32_t t[7], bool rest, uint64_t& st) {
  uint8_t s = 0;
  for (int d = 0; d < 6; d += 2) {
    st = lg_xs(st);
    uint64_t r = st * LG_RNG_MUL;
    s |= (uint8_t)(((uint32_t)(r >> 40) < t[d]) ? 1u : 0u) << d;
    s |= (uint8_t)((((uint32_t)(r >> 16) & 0xFFFFFFu) < t[d + 1]) ? 1u : 0u) << (d + 1);
  }
  if (rest) {
    st = lg_xs(st);
    uint64_t r = st * LG_RNG_MUL;
    if ((uint32_t)(r >> 40) < t[6]) s |= 0x40u;
  }
  return s;
}

struct Realisation {
  std::vector<uint8_t> cells;
  uint64_t seed;
};

const uint64_t TR_TRANSIENT = 200, TR_MIN_INTERVAL = 4, TR_REPS = 4, TR_MAX_WINDOW = 30000;
const size_t TR_SAMPLES = 20;
const float TR_SQRT3_2 = 0.8660254f;
const float TR_PI_F = 3.14159265358979323846f;

double fit_slope(const std::vector<std::pair<double, double>>& pts) {
  double n = (double)pts.size(), sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (auto& p : pts) sx += p.first;
  for (auto& p : pts) sy += p.second;
  for (auto& p : pts) sxx += p.first * p.first;
  for (auto& p : pts) sxy += p.first * p.second;
  return (n * sxy - sx * sy) / (n * sxx - sx * sx);
}

// The measurement's projection: exact row or column sums from the device,
// then the reference's f32 velocity and f32 cos/sin projection in f64.
struct Projector {
  lg_t* g;
  bool rows;
  float k;
  DevBuf<int64_t> buf;
  std::vector<int64_t> host;
  int init() {
    size_t n = rows ? g->G.h : g->G.w;
    host.resize(2 * n);
    HIP_TRY(buf.alloc(2 * n));
    return LG_OK;
  }
  int project(std::pair<double, double>* out) {
    size_t n = rows ? g->G.h : g->G.w;
    if (rows)
      hipLaunchKernelGGL(lg_rows_kernel, dim3(blocks_for(n)), dim3(256), 0, g->stream, g->G, g->a.p,
                         buf.p);
    else
      hipLaunchKernelGGL(lg_cols_kernel, dim3(blocks_for(n)), dim3(256), 0, g->stream, g->G, g->a.p,
                         buf.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host.data(), buf.p, host.size() * 8, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    double a = 0.0, b = 0.0;
    for (size_t i = 0; i < n; ++i) {
      int64_t m = host[2 * i], p = host[2 * i + 1];
      float u, ph;
      if (rows) {  // Moments::velocity().0 (src/moments.rs:91-97)
        u = m == 0 ? 0.0f : (float)p * 0.5f / (float)m;
        ph = k * (float)i * TR_SQRT3_2;
      } else {  // column_uy (src/transport.rs:75-83)
        u = m > 0 ? (float)p * TR_SQRT3_2 / (float)m : 0.0f;
        ph = k * (float)i;
      }
      a += (double)(u * std::cos(ph));
      b += (double)(u * std::sin(ph));
    }
    out->first = 2.0 * a / (double)n;
    out->second = 2.0 * b / (double)n;
    return LG_OK;
  }
};

struct Handle {
  lg_t* g = nullptr;
  ~Handle() { lg_destroy(g); }
};

struct BatchHandle {
  lg_batch_t* b = nullptr;
  ~BatchHandle() { lg_batch_destroy(b); }
};

// averaged() of src/transport.rs:131-171 over the engine. decay_window runs
// realisation 0 alone, as a batch of one (its chunks of steps replay as a
// graph: the 256 x 256 box is far too small for launch latency not to
// dominate); the REPS trajectories then run together as one batch, each
// member exactly as it would run alone. The samples are summed in the
// reference's order (realisation by realisation), and the advance after the
// last sample, which no result reads, is left out.
template <class StartFn>
int measure_mode(lg_t* const* hs, Projector& P, const StartFn& start, std::vector<double>& times,
                 std::vector<std::pair<double, double>>& acc) {
  int rc;
  std::pair<double, double> ab;
  uint64_t t = 0;
  {
    Realisation r0 = start(0);
    if ((rc = lg_restart(hs[0], r0.cells.data(), r0.cells.size(), r0.seed))) return rc;
    BatchHandle B;
    B.b = lg_batch_create(hs, 1);
    if (!B.b) return LG_EDEVICE;
    if ((rc = lg_batch_advance(B.b, TR_TRANSIENT))) return rc;
    P.g = hs[0];
    if ((rc = P.project(&ab))) return rc;
    double a0 = std::max(std::sqrt(ab.first * ab.first + ab.second * ab.second), 1e-12);
    while (t < TR_MAX_WINDOW) {
      if ((rc = lg_batch_advance(B.b, 25))) return rc;
      t += 25;
      if ((rc = P.project(&ab))) return rc;
      if (std::sqrt(ab.first * ab.first + ab.second * ab.second) < a0 / 3.0) break;
    }
  }
  uint64_t window = std::min(std::max(t, (uint64_t)TR_SAMPLES * TR_MIN_INTERVAL), TR_MAX_WINDOW);
  uint64_t interval = std::max(window / (uint64_t)TR_SAMPLES, TR_MIN_INTERVAL);
  for (uint64_t rep = 0; rep < TR_REPS; ++rep) {
    Realisation r = start(rep);
    if ((rc = lg_restart(hs[rep], r.cells.data(), r.cells.size(), r.seed))) return rc;
  }
  BatchHandle B;
  B.b = lg_batch_create(hs, TR_REPS);
  if (!B.b) return LG_EDEVICE;
  if ((rc = lg_batch_advance(B.b, TR_TRANSIENT))) return rc;
  std::vector<std::pair<double, double>> traj(TR_REPS * TR_SAMPLES);
  for (size_t i = 0; i < TR_SAMPLES; ++i) {
    for (uint64_t rep = 0; rep < TR_REPS; ++rep) {
      P.g = hs[rep];
      if ((rc = P.project(&traj[rep * TR_SAMPLES + i]))) return rc;
    }
    if (i + 1 < TR_SAMPLES && (rc = lg_batch_advance(B.b, interval))) return rc;
  }
  acc.assign(TR_SAMPLES, {0.0, 0.0});
  for (uint64_t rep = 0; rep < TR_REPS; ++rep)
    for (size_t i = 0; i < TR_SAMPLES; ++i) {
      acc[i].first += traj[rep * TR_SAMPLES + i].first / (double)TR_REPS;
      acc[i].second += traj[rep * TR_SAMPLES + i].second / (double)TR_REPS;
    }
  times.resize(TR_SAMPLES);
  for (size_t i = 0; i < TR_SAMPLES; ++i) times[i] = (double)(i * interval);
  return LG_OK;
}

// REPS handles for one measurement: quiescent periodic boxes (box_of,
// src/transport.rs:176-179).
struct Boxes {
  Handle h[TR_REPS];
  lg_t* g[TR_REPS] = {};
  int create(size_t size, int rest, size_t threads) {
    for (uint64_t i = 0; i < TR_REPS; ++i) {
      h[i].g = g[i] = lg_create(size, size, rest, 1, threads);
      if (!g[i]) return LG_EDEVICE;
    }
    return LG_OK;
  }
};

}  // namespace

extern "C" {

// transport::viscosity (src/transport.rs:183-221)
int lg_measure_viscosity(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* nu) {
  if (!nu) return fail(LG_EINVAL, "null argument");
  const size_t w = size, h = size;
  const float k = 2.0f * TR_PI_F / ((float)h * TR_SQRT3_2);
  const float amp = 0.05f;
  auto start = [&](uint64_t rep) {
    uint64_t s = seed ^ ((rep * 0x9E3779B9ull) << 20);
    uint64_t st = lg_rng_init(s ^ 0xBEEF);
    Realisation r{std::vector<uint8_t>(w * h), s};
    uint32_t t[7];
    for (size_t y = 0; y < h; ++y) {
      float ux = amp * std::sin(k * (float)y * TR_SQRT3_2);
      equilibrium_thresholds(density, ux, 0.0f, rest != 0, t);
      for (size_t x = 0; x < w; ++x) r.cells[y * w + x] = eq_sample(t, rest != 0, st);
    }
    return r;
  };
  Boxes X;
  int rc = X.create(size, rest, threads);
  if (rc) return rc;
  Projector P{X.g[0], true, k, {}, {}};
  if ((rc = P.init())) return rc;
  std::vector<double> times;
  std::vector<std::pair<double, double>> mode;
  if ((rc = measure_mode(X.g, P, start, times, mode))) return rc;
  std::vector<std::pair<double, double>> pts;
  for (size_t i = 0; i < times.size(); ++i) {
    double a = mode[i].first, b = mode[i].second;
    pts.push_back({times[i], std::log(std::max(std::sqrt(a * a + b * b), 1e-12))});
  }
  *nu = (float)(-fit_slope(pts) / (double)(k * k));
  return LG_OK;
}

// transport::advection_factor (src/transport.rs:225-289)
int lg_measure_advection(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* g_out) {
  if (!g_out) return fail(LG_EINVAL, "null argument");
  const size_t w = size, h = size;
  const float u0 = 0.25f;
  const float k = 2.0f * TR_PI_F / (float)w;
  const float amp = 0.05f;
  std::vector<uint32_t> cols((size_t)w * 7);
  for (size_t x = 0; x < w; ++x)
    equilibrium_thresholds(density, u0, amp * std::sin(k * (float)x), rest != 0, &cols[x * 7]);
  auto start = [&](uint64_t rep) {
    uint64_t s = seed ^ ((rep * 0x9E3779B9ull) << 20) ^ 0x1357;
    uint64_t st = lg_rng_init(s ^ 0xF00D);
    Realisation r{std::vector<uint8_t>(w * h), s};
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) r.cells[y * w + x] = eq_sample(&cols[x * 7], rest != 0, st);
    return r;
  };
  // The realised mean speed (src/transport.rs:264-265): WITH_WALLS sums of
  // the first realisation, Moments::velocity().0.
  float u_actual;
  {
    Realisation r0 = start(0);
    int64_t mass = 0, px = 0;
    const int CX2[6] = {2, 1, -1, -2, -1, 1};
    for (uint8_t c : r0.cells) {
      for (int d = 0; d < 6; ++d)
        if (c & (1u << d)) {
          mass += 1;
          px += CX2[d];
        }
      if (c & 0x40u) mass += 1;
    }
    u_actual = mass == 0 ? 0.0f : (float)px * 0.5f / (float)mass;
  }
  Boxes X;
  int rc = X.create(size, rest, threads);
  if (rc) return rc;
  Projector P{X.g[0], false, k, {}, {}};
  if ((rc = P.init())) return rc;
  std::vector<double> times;
  std::vector<std::pair<double, double>> mode;
  if ((rc = measure_mode(X.g, P, start, times, mode))) return rc;
  const double PI = 3.14159265358979323846;
  double unwrapped = 0.0, prev = std::atan2(mode[0].first, mode[0].second);
  std::vector<std::pair<double, double>> pts;
  for (size_t i = 0; i < mode.size(); ++i) {
    double ph = std::atan2(mode[i].first, mode[i].second);
    double delta = ph - prev;
    while (delta > PI) delta -= 2.0 * PI;
    while (delta < -PI) delta += 2.0 * PI;
    unwrapped += delta;
    prev = ph;
    pts.push_back({times[i], unwrapped});
  }
  *g_out = (float)(-fit_slope(pts) / (double)(k * u_actual));
  return LG_OK;
}

// transport::measure (src/transport.rs:291-319): both measurements, each on
// its own handle and stream, run concurrently from two host threads as the
// reference does on its GPU path.
int lg_measure_transport(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* nu, float* g_out) {
  if (!nu || !g_out) return fail(LG_EINVAL, "null argument");
  int rc_nu = 0;
  std::string err_nu;
  std::thread t([&] {
    rc_nu = lg_measure_viscosity(density, rest, threads, seed, size, nu);
    if (rc_nu) err_nu = g_err;
  });
  int rc_g = lg_measure_advection(density, rest, threads, seed, size, g_out);
  t.join();
  if (rc_nu) return fail(rc_nu, err_nu);
  return rc_g;
}

// Test hook: the generated circuit evaluated on the device for every state
// and every k < stride (out: n_states * stride bytes, 0xFF where lanes of a
// word disagreed). Returns the stride.
int lg_debug_collide_table(int rest, uint8_t* out, size_t n) {
  const size_t ns = rest ? 128 : 64;
  const size_t stride = rest ? LG_COLLIDE_REST_STRIDE : LG_COLLIDE_NOREST_STRIDE;
  if (!out || n < ns * stride) return fail(LG_EINVAL, "output too small");
  uint8_t* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, ns * stride));
  hipLaunchKernelGGL(lg_verify_kernel, dim3(blocks_for(ns * stride)), dim3(256), 0, 0, rest, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, ns * stride, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(LG_EDEVICE, hipGetErrorString(e));
  return (int)stride;
}

}  // extern "C"
