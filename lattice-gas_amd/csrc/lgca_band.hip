// Resident-band pipeline: one persistent workgroup per band of rows steps
// its rows through a whole chunk of steps without leaving registers.
//
// Rows of at most 64 words only: lane j holds word j of its row, wave w row
// y0 + w of the band, for the whole launch. Per step each row gathers its six
// neighbour words -- across lanes for x (DPP rotation, or ds_bpermute when
// the row is shorter than the wave), through LDS from the neighbouring waves
// for y, and, for a band's first and last row, from the neighbouring bands'
// halo rows published in global memory -- collides with the digits it drew
// for this step, publishes its new halo rows, draws the next step's digits
// (the expensive, state-independent work, which hides the neighbours'
// latency) and meets the other waves at one barrier.
//
// Bands meet through per-band step counters ("flags") and halo words, all
// stored and loaded as agent-scope atomics (coherent across the GPU's XCDs;
// see lg_band_wait for the ordering). Every poll is bounded: a wait that runs out records the launch's
// epoch in an error word and every later wait of that launch returns at
// once, so the kernel always ends and the host reports the failure. Every
// band's workgroup must be resident at once: the process's first band launch
// on a device is cooperative, the later ones plain and ordered one after the
// other (BandOrder below).
//
// Flags count steps from a per-launch epoch that only grows (2^21 apart),
// so a launch never reads an older launch's counts as its own; halo rows sit
// in two slots by step parity, which the counters keep apart: a band writes
// slot (s + 1) & 1 only after both neighbours have read slot (s - 1) & 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "lgca_internal.h"
#include "lgca_jump.hpp"

namespace {

// Spin until *f >= v. On running out, record `epoch` in *err; return false
// at once if *err already holds this launch's epoch.
//
// No L2 write-back or invalidate is needed around the counters: everything a
// band hands to another -- halo words and counters -- moves as agent-scope
// atomics (written through to, and read from, the device's coherence point,
// past the per-XCD L2s), a publisher waits for its halo stores to complete
// before it stores the counter (lg_band_publish), and a reader issues its
// halo loads only after the counter load has returned the awaited value.
// The compiler fences keep the compiler from moving the halo accesses across
// the counter accesses; the agent-scope acquire/release fences the compiler
// would emit instead (buffer_inv / buffer_wbl2 sc1) cost about 4.5 us per
// exchange on MI355X and protect only plain loads and stores, which the
// exchange does not use.
__device__ __forceinline__ bool lg_band_wait(uint64_t* f, uint64_t v, uint64_t* err,
                                             uint64_t* err_host, uint64_t epoch) {
  for (uint32_t it = 0; it < LG_BAND_SPIN_LIMIT; ++it) {
    if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= v) {
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      return true;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(err, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(err_host, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  return false;
}

// Store a band's step counter once its halo stores have completed.
__device__ __forceinline__ void lg_band_publish(uint64_t* f, uint64_t v) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0): the halo stores are done
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lg_ld_agent(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lg_st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Word of the lane to the left / right in the row (periodic); see
// lg_collide_k_kernel for the two forms.
template <bool WAVEROW>
__device__ __forceinline__ uint64_t lg_band_left(uint64_t x, int addr) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (WAVEROW) {
    lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x13C, 0xF, 0xF, false);  // wave_ror:1
    hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x13C, 0xF, 0xF, false);
  } else {
    lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)lo);
    hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)hi);
  }
  return ((uint64_t)hi << 32) | lo;
}
template <bool WAVEROW>
__device__ __forceinline__ uint64_t lg_band_right(uint64_t x, int addr) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (WAVEROW) {
    lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x134, 0xF, 0xF, false);  // wave_rol:1
    hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x134, 0xF, 0xF, false);
  } else {
    lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)lo);
    hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)hi);
  }
  return ((uint64_t)hi << 32) | lo;
}

// The collision digits of word (y, j) for one step (bt: its row band's
// stream table of that step) for fluid mask fm: both halves -- lg_draw_unit<1>'s
// digits -- into the row's digit slot in LDS (kout[q * 64 + lane] the low
// half's digit q, kout[(LG_KDIGITS + q) * 64 + lane] the high half's), where
// the row's collide of the step reads them (in LDS rather than registers
// across the barrier, which leaves the step loop's registers to the rest).
// LEAD (every half of the lattice has at most one run of cells without a
// draw): cj0 is the word's lead jump (lgca::build_lead_jumps), len draws
// before its first draw where the low half has such a run; both halves take
// the plain chain, and the digits are moved past the runs (lg_run_fix_lead,
// lg_run_fix) in the waves where some lane needs it. A wave whose lanes all draw every cell of a half takes the
// plain chain for it. Where some lane's half has cells without a draw, and
// in every lane those form at most one run (a plate, a row's tail), the wave
// takes the plain chain and moves the digits past the run
// (lg_draw_half_run); the upper half then starts from its own jump (cjhi:
// the word's cjump[2 * word + 1]) rather than where the lower one ended.
// Otherwise the whole wave takes the branch-free masked chain, which
// continues through the masked cells, so the upper half starts where the
// lower one ended without a second jump.
template <int NK, bool OPAQUE_FM>
__device__ __forceinline__ void lg_band_digits(uint64_t fm, uint64_t cj0,
                                               const uint64_t* __restrict cjhi,
                                               const uint64_t* __restrict bt, uint32_t stride,
                                               uint32_t lane, uint32_t* __restrict kout,
                                               bool lead) {
  uint32_t Kl[NK], Kh[NK];
#if defined(__HIP_DEVICE_COMPILE__)
  // OPAQUE_FM: the fluid mask is the same every step, and left to itself the
  // compiler hoists the masked chains' per-cell lane masks out of the step
  // loop, 64 scalar register pairs that then spill and push other values out
  // too. Measured per instance: in the sampling instances hiding the mask
  // from the optimiser shortens the step; in the others it lengthens it.
  if (OPAQUE_FM) asm volatile("" : "+v"(fm));
#endif
  uint64_t st = lg_jump_tab(cj0, bt);
  const uint32_t f0 = (uint32_t)fm, f1 = (uint32_t)(fm >> 32);
  if (lead) {
    st = lg_draw_half<NK>(st, stride, Kl);
    if (__ballot(f0 != ~0u)) lg_run_fix_lead<NK>(f0, Kl);
    LG_UNROLL
    for (int q = 0; q < LG_KDIGITS; ++q) kout[q * 64 + lane] = q < NK ? Kl[q] : 0u;
    lg_draw_half<NK>(st, stride, Kh);
    if (__ballot(f1 != ~0u)) lg_run_fix<NK>(f1, Kh);
    LG_UNROLL
    for (int q = 0; q < LG_KDIGITS; ++q) kout[(LG_KDIGITS + q) * 64 + lane] = q < NK ? Kh[q] : 0u;
    return;
  }
  if (__ballot(f0 != ~0u)) {
    if (__ballot(!lg_one_run(f0))) {
      st = lg_draw_half_sel<NK>(st, f0, stride, Kl);
    } else {
      const uint64_t cj1 = *cjhi;
      lg_draw_half_run<NK>(st, f0, stride, Kl);
      st = lg_jump_tab(cj1, bt);
    }
  } else {
    st = lg_draw_half<NK>(st, stride, Kl);
  }
  LG_UNROLL
  for (int q = 0; q < LG_KDIGITS; ++q) kout[q * 64 + lane] = q < NK ? Kl[q] : 0u;
  if (__ballot(f1 != ~0u)) {
    if (__ballot(!lg_one_run(f1)))
      lg_draw_half_sel<NK>(st, f1, stride, Kh);
    else
      lg_draw_half_run<NK>(st, f1, stride, Kh);
  } else {
    lg_draw_half<NK>(st, stride, Kh);
  }
  LG_UNROLL
  for (int q = 0; q < LG_KDIGITS; ++q) kout[(LG_KDIGITS + q) * 64 + lane] = q < NK ? Kh[q] : 0u;
}

}  // namespace

// The band row wave w holds. Waves go to the CU's four SIMDs in turn (wave
// w to SIMD w % 4), so a band of B rows puts one wave more on SIMDs 0 ..
// B % 4 - 1. The band's first and last rows wait for and publish to the
// neighbouring bands; they go to the first waves on the other SIMDs, where
// fewer waves share the issue slots around those waits, and the rows in
// between fill the remaining waves in order.
LG_HD inline uint32_t lg_band_row_of_wave(uint32_t w, uint32_t B) {
  const uint32_t heavy = B % 4;
  if (heavy == 0 || B < 6) return w;
  const uint32_t wt = heavy, wb = heavy + 1 < 4 ? heavy + 1 : heavy + 4;  // light waves
  if (w == wt) return 0;
  if (w == wb) return B - 1;
  // the other waves in order take rows 1 .. B - 2
  return 1 + w - (w > wt ? 1u : 0u) - (w > wb ? 1u : 0u);
}

// The band row wave w holds when the band's first row has its draws split
// between two helper waves (bands of B = 4m + 1 rows): waves 0 .. B - 2 hold
// rows 1 .. B - 1 (the last row, which waits for and publishes to the band
// below, on wave 1, which shares its SIMD with the table wave), wave B - 1
// holds row 0 and only collides it; the table wave (B) and the helpers (B +
// 1, B + 2) fill the other SIMDs' third slots, so every SIMD holds 2m rows
// and a part of the remaining row's work.
LG_HD inline uint32_t lg_band_row_split(uint32_t w, uint32_t B) {
  if (w == B - 1) return 0;
  if (B >= 3 && w <= 1) return w == 0 ? 1u : B - 1;
  return w;
}

// The stream tables one band needs for step t into LDS (out: nstreams x
// LG_WTAB entries): its row bands' collision streams rb0, rb0 + 1, ... and,
// last, the inflow stream of the step's end (src/lattice.rs:233,242,276);
// the same 4-bit window tables lg_tables_kernel builds. One lane per
// (stream, window group): it reaches the group's first state 4g from the
// seeded start through the T^(4g) windows in LDS (pw) and writes the group's
// 16 entries; the lanes of one wave take the tasks in turn.
__device__ __forceinline__ void lg_band_tables(const lgca_internal::LgBandArgs& A, uint32_t rb0,
                                               uint32_t nsb, uint64_t t, uint32_t lane,
                                               const uint64_t* __restrict pw,
                                               uint64_t* __restrict out) {
  const uint32_t tasks = (nsb + 1) * 16;
  for (uint32_t q = lane; q < tasks; q += 64) {
    const uint32_t sidx = q >> 4, grp = q & 15u;
    const uint64_t st = lg_rng_init(sidx < nsb ? lg_band_seed(A.seed, t, rb0 + sidx)
                                               : lg_inlet_seed(A.seed, t + 1));
    lg_tab_group(lg_xs_pow4(st, grp, pw), out + (size_t)sidx * LG_WTAB + grp * 16);
  }
}

// Step t's inflow cells of the band's rows, drawn by the table wave (inflow
// of at most 64 cells, lanes per row `cpr` a power of two): lane q of chunk c
// takes cell q % cpr of band row c * (64 / cpr) + q / cpr from its own jump
// (icl[c * 64 + q], the cell's icell entry kept in LDS for the launch; zero
// for a solid cell, which takes no draw) against the inflow stream's table
// itab; each row's seven state planes of cells 0 .. cols - 1
// go to out[row * 8 + d], which the row's collide of step t applies.
__device__ __forceinline__ void lg_band_inflow(const LgGeom& G, const uint64_t* __restrict icl,
                                               uint32_t Bb, uint32_t cpr, uint32_t lane,
                                               const uint64_t* __restrict itab,
                                               uint64_t* __restrict out) {
  const uint32_t cols = G.inlet_cols, rpc = 64u / cpr;
  const uint64_t fieldmask = cpr >= 64 ? ~0ull : (1ull << cpr) - 1ull;
  for (uint32_t r0 = 0; r0 < Bb; r0 += rpc) {
    const uint32_t r = r0 + lane / cpr, x = lane % cpr;
    uint32_t s = 0;
    if (r < Bb && x < cols) {
      const uint64_t c = icl[(r0 / rpc) * 64 + lane];
      if (c) {
        uint64_t st = lg_jump_tab(c, itab);
        s = lg_eq_cell(G, st);
      }
    }
    uint64_t m[7];
    LG_UNROLL
    for (int d = 0; d < 7; ++d) m[d] = __ballot((s >> d) & 1u);
    if (lane < rpc && r0 + lane < Bb) {
      uint64_t* o = out + (size_t)(r0 + lane) * 8;
      LG_UNROLL
      for (int d = 0; d < 7; ++d) o[d] = (m[d] >> ((lane * cpr) & 63u)) & fieldmask;
    }
  }
}

// The coarse-graining of the launch's final state (src/render.rs:93-115, the
// sums of lg_block_sums and the f32 update of lg_block_field), done by the
// bands at the launch's end instead of by a second kernel. Each lane adds its
// word's fluid-cell counts per block into the band's block sums in LDS (the
// exchange slots, free by now); the band stores them for its block rows and
// counts itself in at each block row; the band that arrives last at a block
// row adds up that row's sums over the bands it spans and updates the field.
// The sums are integers, so the result does not depend on the order of
// arrival. The stores and counts move as agent-scope atomics like the halo
// exchange, so the last band reads what the others stored.
__device__ __forceinline__ void lg_band_sample(const lgca_internal::LgBandField& A, float alpha,
                                               uint32_t rows, int32_t* acc, const uint64_t P[7],
                                               uint64_t sm, bool own, uint32_t y, uint32_t j,
                                               uint32_t b, uint32_t y0, uint32_t Bb) {
  const uint32_t tid = threadIdx.x, nthr = blockDim.x;
  const uint32_t per_band = A.nbr * A.bw * 4;
  for (uint32_t i = tid; i < per_band; i += nthr) acc[i] = 0;
  __syncthreads();
  const uint32_t jb0 = y0 / A.by, xend = A.bw * A.bx;  // columns past the last whole block: none
  if (own && y < A.bh * A.by && 64 * j < xend) {
    const int CX2[6] = {2, 1, -1, -2, -1, 1};
    const int CY2[6] = {0, 1, 1, 0, -1, -1};
    const uint32_t k = y / A.by - jb0, c0 = 64 * j, c1 = min(c0 + 64, xend);
    for (uint32_t ib = c0 / A.bx; ib * A.bx < c1; ++ib) {
      const uint32_t lo = max(ib * A.bx, c0) - c0, hi = min((ib + 1) * A.bx, c1) - c0;
      const uint64_t m = (hi - lo >= 64) ? ~0ull : (((1ull << (hi - lo)) - 1ull) << lo);
      const uint64_t fm = m & ~sm;
      int32_t mass = 0, px = 0, py = 0;
      LG_UNROLL
      for (int d = 0; d < 6; ++d) {
        const int32_t c = __popcll(P[d] & fm);
        mass += c;
        px += CX2[d] * c;
        py += CY2[d] * c;
      }
      mass += __popcll(P[6] & fm);
      int32_t* a = acc + ((size_t)k * A.bw + ib) * 4;
      atomicAdd(a + 0, mass);
      atomicAdd(a + 1, px);
      atomicAdd(a + 2, py);
      atomicAdd(a + 3, (int32_t)__popcll(sm & m));
    }
  }
  __syncthreads();
  // block rows [jb0, jb0 + nk) hold rows of this band
  const uint32_t jb1 = min(A.bh, (y0 + Bb - 1) / A.by + 1);
  const uint32_t nk = jb0 < jb1 ? jb1 - jb0 : 0u;
  int32_t* mine = A.part + (size_t)b * per_band;
  for (uint32_t i = tid; i < nk * A.bw * 4; i += nthr)
    __hip_atomic_store(mine + i, acc[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // this thread's stores are done
  __syncthreads();                // ... and so are the whole band's
  uint32_t fin = 0;
  if (tid < nk) {
    const uint32_t jb = jb0 + tid;
    const uint32_t want = (jb * A.by + A.by - 1) / rows - (jb * A.by) / rows + 1;
    const uint32_t old =
        __hip_atomic_fetch_add(A.bcnt + jb, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    fin = old + 1 == want;
    // the last band resets the count for the next launch: every band of
    // this block row has counted itself in
    if (fin) __hip_atomic_store(A.bcnt + jb, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (tid < nk) acc[tid] = (int32_t)fin;
  __syncthreads();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  for (uint32_t q = 0; q < nk; ++q) {
    if (!acc[q]) continue;
    const uint32_t jb = jb0 + q;
    const uint32_t bfirst = (jb * A.by) / rows, blast = (jb * A.by + A.by - 1) / rows;
    for (uint32_t ib = tid; ib < A.bw; ib += nthr) {
      int64_t s4[4] = {0, 0, 0, 0};
      for (uint32_t bb = bfirst; bb <= blast; ++bb) {
        const uint32_t kk = jb - (bb * rows) / A.by;
        const int32_t* pp = A.part + (size_t)bb * per_band + ((size_t)kk * A.bw + ib) * 4;
        LG_UNROLL
        for (int c = 0; c < 4; ++c)
          s4[c] += __hip_atomic_load(const_cast<int32_t*>(pp + c), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      }
      lg_block_field(A.bx, A.by, alpha, s4, A.ux, A.uy, A.rho, A.sol, A.sums, jb * A.bw + ib);
    }
  }
}

// One band per workgroup (blockIdx.x), one row per wave, one word per lane,
// plus one wave that builds the next step's stream tables; A.n steps from
// A.src into A.dst. LDS: two exchange slots of the rows' planes 1, 2, 4, 5
// (rows x 4 KB), two slots of stream tables (A.nstreams x 4 KB), the helper
// waves' digits (split bands), the T^(4g) windows (32 KB), two slots of
// the table wave's inflow planes (rows x 64 B), each row's digits for its
// next collide (rows x 2.5 KB) and (INFLOW 0) the inflow cells' jumps
// (1.5 KB): at most 159 KB (11 rows, 12 streams).
// SAMPLE: the launch ends with the fused coarse-graining (an instance of its
// own, so launches without it keep the step loop's register allocation).
// INFLOW (how the inflow cells draw): 0, the table wave draws every inflow
// cell of the band (inflow of at most 64 cells, few cells per band); 1, in
// each row's wave lane x draws cell x (at most 64 cells); 2, the serial
// per-word chain in the lanes of the inflow words (wider inflow). Each choice
// is an instance of its own, so the others' registers are not held across
// the step loop.
template <bool WAVEROW, bool SAMPLE, int INFLOW>
__global__ void __launch_bounds__(64 * (LG_BAND_MAX_ROWS + 1))
    lg_band_kernel(lgca_internal::LgBandArgs A) {
  extern __shared__ uint64_t xch[];  // [2 slots][rows][4 planes][64 lanes]
  const LgGeom& G = A.G;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, NW = A.rows;
  uint64_t* const tabs_lds = xch + (size_t)2 * NW * 256;  // [2 slots][nstreams][LG_WTAB]
  const bool table_wave = w == NW;
  // split draws: row 0's digits are drawn by two helper waves, half a word
  // per lane each, into LDS ([2 slots][2 halves][LG_KDIGITS][64 lanes])
  const bool split = A.split != 0;
  const bool helper = split && w > NW;
  const uint32_t hh = helper ? w - NW - 1 : 0u;  // the half a helper draws
  const bool drawless = split && w + 1 == NW;    // row 0's wave: digits from the helpers
  uint32_t* const kx = reinterpret_cast<uint32_t*>(tabs_lds + (size_t)2 * A.nstreams * LG_WTAB);
  // [16 groups][16 windows][16 values], then [2 slots][rows][8] inflow planes
  uint64_t* const pw_lds =
      tabs_lds + (size_t)2 * A.nstreams * LG_WTAB + (split ? 2 * LG_KDIGITS * 64 : 0);
  uint64_t* const ilds = pw_lds + 16 * 256;
  // [rows][2 halves][LG_KDIGITS][64 lanes]: each row's digits for its next collide
  uint32_t* const kr = reinterpret_cast<uint32_t*>(ilds + (size_t)2 * NW * 8);
  // INFLOW 0: the table wave's inflow cells' jumps ([LG_BAND_ICHUNKS][64])
  uint64_t* const icl = reinterpret_cast<uint64_t*>(kr + (size_t)NW * 2 * LG_KDIGITS * 64);
  const uint32_t b = blockIdx.x, y0 = b * A.rows;
  const uint32_t Bb = min(A.rows, G.h - y0);  // rows of this band
  // the band row it holds (helpers: none, but y below falls on row 0, whose
  // words they draw)
  const uint32_t r = (table_wave || helper) ? NW
                     : split                ? lg_band_row_split(w, NW)
                                            : lg_band_row_of_wave(w, NW);
  const bool row_live = r < Bb;
  // the table wave's work (the next step's stream tables, the step's inflow
  // cells) is one wave's serial chain; left at the default priority it
  // issues after the older row waves of its SIMD and is the last to reach
  // each pass's barrier, so it issues first
  if (table_wave) __builtin_amdgcn_s_setprio(3);
  const uint32_t y = y0 + (row_live ? r : 0u);
  const bool live = lane < G.wpr;
  const uint32_t j = live ? lane : 0u;
  const uint32_t cl = (j == 0) ? G.lastbit : 63u, cr = (j == G.lastword) ? G.lastbit : 63u;
  const uint64_t tail = (j == G.lastword) ? G.tailmask : ~0ull;
  const int addr_l = (int)(4u * (live ? (j == 0 ? G.wpr - 1 : j - 1) : lane));
  const int addr_r = (int)(4u * (live ? (j + 1 == G.wpr ? 0u : j + 1) : lane));
  const uint32_t bu = (b + A.nb - 1) % A.nb, bd = (b + 1) % A.nb;
  const bool top = r == 0, bot = r + 1 == Bb;
  const size_t T = G.total, word = (size_t)y * G.wpr + j;
  // the row bands (collision streams) of this band's rows: rb0 .. rb0 + nsb - 1
  const uint32_t rb0 = A.row_band[y0], nsb = A.row_band[y0 + Bb - 1] - rb0 + 1;
  const size_t hstride = (size_t)A.nb * 128;  // one halo kind of one slot: nb x 2 planes x 64

  uint64_t P[7];
  LG_UNROLL
  for (int d = 0; d < 7; ++d) P[d] = A.src[(size_t)d * T + word];
  const uint64_t sm = A.solid[word];
  const uint64_t fm = live ? ~sm & tail : 0ull;
  // a helper's half starts at its own jump; a row's word at its lead jump
  const uint64_t cj0 = (A.lead_ok && !helper) ? A.lead[word] : A.cjump[2 * word + hh];
  const uint32_t stream = A.row_band[y];
  // inflow: INFLOW 1, lane x samples cell x of the row's first word from its
  // own jump (icell); INFLOW 2, the serial per-word chain (ijump) in the
  // lanes of the inflow words; INFLOW 0, the table wave draws them all
  constexpr bool lane_inflow = INFLOW == 1;
  const uint64_t sm0 = A.solid[(size_t)y * G.wpr];
  const bool icell_live = lane_inflow && lane < G.inlet_cols && !((sm0 >> lane) & 1ull);
  const uint64_t ic = icell_live ? A.icell[(size_t)y * G.inlet_cols + lane] : 0ull;
  const uint64_t ij = (INFLOW == 2 && j < G.iw) ? A.ijump[2 * ((size_t)y * G.iw + j)] : 0ull;
  // the step's inflow: wide, the word's seven planes (I); narrow, the lane's
  // cell (ibits), gathered across the wave where it is applied
  uint64_t I[7];
  uint32_t ibits = 0;
  LG_UNROLL
  for (int d = 0; d < 7; ++d) I[d] = 0;
  for (uint32_t q = threadIdx.x; q < 16 * 256; q += blockDim.x) pw_lds[q] = A.pw[q];
  if (INFLOW == 0 && table_wave && G.inlet_cols > 0) {
    // the inflow cells' jumps, read once (the table wave alone reads them,
    // each lane its own entries)
    const uint32_t cols = G.inlet_cols, rpc = 64u / A.icpr;
    for (uint32_t c = 0; c * rpc < Bb; ++c) {
      const uint32_t rr = c * rpc + lane / A.icpr, x = lane % A.icpr;
      icl[c * 64 + lane] = (rr < Bb && x < cols) ? A.icell[(size_t)(y0 + rr) * cols + x] : 0ull;
    }
  }
  __syncthreads();
  if (table_wave) lg_band_tables(A, rb0, nsb, A.t0, lane, pw_lds, tabs_lds);
  if (row_live) {
    uint64_t* s0 = xch + (size_t)r * 256;
    s0[0 * 64 + lane] = P[1];
    s0[1 * 64 + lane] = P[2];
    s0[2 * 64 + lane] = P[4];
    s0[3 * 64 + lane] = P[5];
  }
  __syncthreads();

  // Pass i collides step i - 1 (i > 0) and draws step i (i < n): one copy
  // of each in the code.
  for (uint32_t i = 0; i <= A.n; ++i) {
    const uint32_t s = i - 1;  // the step collided in this pass
    if (row_live && i > 0) {
      const uint64_t* slot = xch + (size_t)(s & 1u) * NW * 256;
      uint64_t a1, a2, a4, a5;
      // planes 1, 2 of the row above
      if (!top) {
        a1 = slot[((r - 1) * 4 + 0) * 64 + lane];
        a2 = slot[((r - 1) * 4 + 1) * 64 + lane];
      } else if (s == 0) {
        const size_t up = (size_t)((y0 + G.h - 1) % G.h) * G.wpr + j;
        a1 = A.src[1 * T + up];
        a2 = A.src[2 * T + up];
      } else {
        lg_band_wait(&A.flags[A.nb + bu], A.epoch + s, A.err, A.err_host, A.epoch);
        const uint64_t* hb = A.halo + (size_t)(s & 1u) * 2 * hstride + hstride + (size_t)bu * 128;
        a1 = lg_ld_agent(hb + lane);
        a2 = lg_ld_agent(hb + 64 + lane);
      }
      // planes 4, 5 of the row below
      if (!bot) {
        a4 = slot[((r + 1) * 4 + 2) * 64 + lane];
        a5 = slot[((r + 1) * 4 + 3) * 64 + lane];
      } else if (s == 0) {
        const size_t dn = (size_t)((y0 + Bb) % G.h) * G.wpr + j;
        a4 = A.src[4 * T + dn];
        a5 = A.src[5 * T + dn];
      } else {
        lg_band_wait(&A.flags[bd], A.epoch + s, A.err, A.err_host, A.epoch);
        const uint64_t* ht = A.halo + (size_t)(s & 1u) * 2 * hstride + (size_t)bd * 128;
        a4 = lg_ld_agent(ht + lane);
        a5 = lg_ld_agent(ht + 64 + lane);
      }
      const uint32_t odd = y & 1u;
      const uint64_t a0l = lg_band_left<WAVEROW>(P[0], addr_l);
      const uint64_t a3r = lg_band_right<WAVEROW>(P[3], addr_r);
      uint64_t a1l = 0, a5l = 0, a2r = 0, a4r = 0;
      if (odd) {
        a2r = lg_band_right<WAVEROW>(a2, addr_r);
        a4r = lg_band_right<WAVEROW>(a4, addr_r);
      } else {
        a1l = lg_band_left<WAVEROW>(a1, addr_l);
        a5l = lg_band_left<WAVEROW>(a5, addr_l);
      }
      uint64_t n[7];
      lg_gather(odd, cl, cr, P[0], a0l, a1, a1l, a2, a2r, P[3], a3r, a4, a4r, a5, a5l, P[6], n);
      uint32_t Kl[LG_KDIGITS], Kh[LG_KDIGITS];
      if (drawless) {
        const uint32_t* kin = kx + (size_t)(s & 1u) * 2 * LG_KDIGITS * 64;
        LG_UNROLL
        for (int q = 0; q < LG_KDIGITS; ++q) {
          Kl[q] = kin[q * 64 + lane];
          Kh[q] = kin[(LG_KDIGITS + q) * 64 + lane];
        }
      } else {
        const uint32_t* kin = kr + (size_t)r * 2 * LG_KDIGITS * 64;
        LG_UNROLL
        for (int q = 0; q < LG_KDIGITS; ++q) {
          Kl[q] = kin[q * 64 + lane];
          Kh[q] = kin[(LG_KDIGITS + q) * 64 + lane];
        }
      }
      uint64_t res[7];
      lg_collide_word(G, n, sm, Kl, Kh, res);
      if constexpr (INFLOW == 1) {  // cells 0 .. inlet_cols - 1 of word 0
        LG_UNROLL
        for (int d = 0; d < 7; ++d) I[d] = __ballot((ibits >> d) & 1u);
      } else if constexpr (INFLOW == 0) {  // drawn by the table wave
        const uint64_t* iin = ilds + ((size_t)(s & 1u) * NW + r) * 8;
        LG_UNROLL
        for (int d = 0; d < 7; ++d) I[d] = iin[d];
      }
      if (j < G.iw) lg_inflow_apply(G, j, sm, tail, I, res);
      LG_UNROLL
      for (int d = 0; d < 7; ++d) P[d] = res[d] & tail;

      // publish the new rows: to the neighbouring waves, and the band's edge
      // rows to the neighbouring bands
      const uint32_t ns = (s + 1) & 1u;
      uint64_t* out = xch + ((size_t)ns * NW + r) * 256;
      out[0 * 64 + lane] = P[1];
      out[1 * 64 + lane] = P[2];
      out[2 * 64 + lane] = P[4];
      out[3 * 64 + lane] = P[5];
      if (top) {  // planes 4, 5 of the band's first row, for the band above
        uint64_t* ht = A.halo + (size_t)ns * 2 * hstride + (size_t)b * 128;
        lg_st_agent(ht + lane, P[4]);
        lg_st_agent(ht + 64 + lane, P[5]);
        lg_band_publish(&A.flags[b], A.epoch + s + 1);
      }
      if (bot) {  // planes 1, 2 of the band's last row, for the band below
        uint64_t* hb = A.halo + (size_t)ns * 2 * hstride + hstride + (size_t)b * 128;
        lg_st_agent(hb + lane, P[1]);
        lg_st_agent(hb + 64 + lane, P[2]);
        lg_band_publish(&A.flags[A.nb + b], A.epoch + s + 1);
      }
    }
    // the next step's tables, beside the rows' work, and (INFLOW 0) this
    // step's inflow cells
    if (table_wave && i + 1 < A.n)
      lg_band_tables(A, rb0, nsb, A.t0 + i + 1, lane, pw_lds,
                     tabs_lds + (size_t)((i + 1) & 1u) * A.nstreams * LG_WTAB);
    if (INFLOW == 0 && table_wave && i < A.n && G.inlet_cols > 0)
      lg_band_inflow(G, icl, Bb, A.icpr, lane,
                     tabs_lds + (size_t)(i & 1u) * A.nstreams * LG_WTAB + (size_t)nsb * LG_WTAB,
                     ilds + (size_t)(i & 1u) * NW * 8);
    // step i's draws
    if (row_live && i < A.n) {
      const uint64_t* tab = tabs_lds + (size_t)(i & 1u) * A.nstreams * LG_WTAB;
      const uint64_t* itab = tab + (size_t)nsb * LG_WTAB;
      uint64_t c0 = cj0, ci = ic, cw = ij;
#if defined(__HIP_DEVICE_COMPILE__)
      // the jumps' window offsets are the same every step; left to itself
      // the compiler hoists them out of the loop and spills them
      asm volatile("" : "+v"(c0), "+v"(ci), "+v"(cw));
#endif
      if (live && !drawless) {
        const uint64_t* cjhi = A.cjump + 2 * word + 1;
        if (G.rest)
          lg_band_digits<LG_COLLIDE_REST_KBITS, SAMPLE>(fm, c0, cjhi,
                                                tab + (size_t)(stream - rb0) * LG_WTAB,
                                                LG_COLLIDE_REST_STRIDE, lane,
                                                kr + (size_t)r * 2 * LG_KDIGITS * 64,
                                                A.lead_ok != 0);
        else
          lg_band_digits<LG_COLLIDE_NOREST_KBITS, SAMPLE>(fm, c0, cjhi,
                                                  tab + (size_t)(stream - rb0) * LG_WTAB,
                                                  LG_COLLIDE_NOREST_STRIDE, lane,
                                                  kr + (size_t)r * 2 * LG_KDIGITS * 64,
                                                  A.lead_ok != 0);
      }
      if constexpr (lane_inflow) {
        // cell `lane` of word 0 (zero where it takes no inflow)
        ibits = 0;
        if (icell_live) {
          uint64_t st = lg_jump_tab(ci, itab);
          ibits = lg_eq_cell(G, st);
        }
      } else if (INFLOW == 2 && live && j < G.iw) {
        const uint64_t apply = lg_inlet_apply(G, j, fm);
        if (apply) lg_inlet_bits(G, lg_jump_tab(cw, itab), apply, I);
      }
    }
    // step i's draws of row 0, one half per helper
    if (helper && i < A.n) {
      const uint64_t* tab = tabs_lds + (size_t)(i & 1u) * A.nstreams * LG_WTAB +
                            (size_t)(stream - rb0) * LG_WTAB;
      uint64_t c0 = cj0;
#if defined(__HIP_DEVICE_COMPILE__)
      asm volatile("" : "+v"(c0));
#endif
      uint32_t Kx[LG_KDIGITS];
      LG_UNROLL
      for (int q = 0; q < LG_KDIGITS; ++q) Kx[q] = 0;
      if (live) {
        uint32_t fh = hh ? (uint32_t)(fm >> 32) : (uint32_t)fm;
#if defined(__HIP_DEVICE_COMPILE__)
        if (SAMPLE) asm volatile("" : "+v"(fh));  // as in lg_band_digits
#endif
        const uint64_t st = lg_jump_tab(c0, tab);
        // each half starts at its own jump: a half with one run of cells
        // without a draw needs no end state (lg_draw_half_run)
        const bool masked = __ballot(fh != ~0u) != 0;
        const bool runs = masked && __ballot(!lg_one_run(fh)) == 0;
        if (G.rest) {
          if (runs)
            lg_draw_half_run<LG_COLLIDE_REST_KBITS>(st, fh, LG_COLLIDE_REST_STRIDE, Kx);
          else if (masked)
            lg_draw_half_sel<LG_COLLIDE_REST_KBITS>(st, fh, LG_COLLIDE_REST_STRIDE, Kx);
          else
            lg_draw_half<LG_COLLIDE_REST_KBITS>(st, LG_COLLIDE_REST_STRIDE, Kx);
        } else {
          if (runs)
            lg_draw_half_run<LG_COLLIDE_NOREST_KBITS>(st, fh, LG_COLLIDE_NOREST_STRIDE, Kx);
          else if (masked)
            lg_draw_half_sel<LG_COLLIDE_NOREST_KBITS>(st, fh, LG_COLLIDE_NOREST_STRIDE, Kx);
          else
            lg_draw_half<LG_COLLIDE_NOREST_KBITS>(st, LG_COLLIDE_NOREST_STRIDE, Kx);
        }
      }
      uint32_t* o = kx + ((size_t)(i & 1u) * 2 + hh) * LG_KDIGITS * 64;
      LG_UNROLL
      for (int q = 0; q < LG_KDIGITS; ++q) o[q * 64 + lane] = Kx[q];
    }
    __syncthreads();
  }
  // the coarse-graining first: its hand-offs between bands then need not
  // wait for the final state's stores to complete
  if (SAMPLE) {
    // the field's layout is read from device memory here, after the step
    // loop: as kernel arguments its loads would be hoisted to the kernel's
    // start and held in registers the loop needs
    const lgca_internal::LgBandField* fs = A.fs;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(fs));
#endif
    lg_band_sample(*fs, A.alpha, A.rows, reinterpret_cast<int32_t*>(xch), P, sm, row_live && live,
                   y, j, b, y0, Bb);
  }
  if (row_live && live) {
    LG_UNROLL
    for (int d = 0; d < 7; ++d) A.dst[(size_t)d * T + word] = P[d];
  }
}

namespace lgca_internal {

// Band launches of this process, per device. The grid needs every band's
// workgroup resident at once. Before the device's first band launch the
// process makes a cooperative one (of no steps), which the runtime declines
// while another process that has made one is alive: the process that gets it
// runs band launches from then on and the other takes the other pipelines.
// The band launches themselves are plain (the cooperative launch's runtime
// work costs ~10-15 us a call) and are ordered one after the other on the
// device: a launch from another stream than the last one waits
// for that stream's work, so no two band grids of the process compete for
// CUs. Other kernels that share the device end on their own; band
// workgroups that wait for CUs meanwhile only delay their neighbours' waits,
// which are bounded.
namespace {
struct BandOrder {
  std::mutex mu;
  bool claimed[LG_MAX_DEVICES] = {};
  hipStream_t last[LG_MAX_DEVICES] = {};  // stream of the latest band launch
  hipEvent_t mark[LG_MAX_DEVICES] = {};   // recorded there when the stream changes
};
BandOrder& band_order() {
  static BandOrder o;
  return o;
}
}  // namespace

void band_forget_stream(const lg_t* g) {
  if (g->device < 0 || g->device >= LG_MAX_DEVICES) return;
  BandOrder& o = band_order();
  std::lock_guard<std::mutex> lk(o.mu);
  if (o.last[g->device] == g->stream) o.last[g->device] = nullptr;
}

bool band_eligible(const lg_t* g, uint32_t* rows, uint32_t* nb) {
  const uint32_t ncu = (uint32_t)std::max(g->ncu, 1);
  if (g->G.wpr > 64) return false;
  const uint32_t r = (g->G.h + ncu - 1) / ncu;
  if (r > LG_BAND_MAX_ROWS) return false;
  *rows = std::max<uint32_t>(r, 1);
  *nb = (g->G.h + *rows - 1) / *rows;
  return true;
}

int ensure_band_buffers(lg_t* g, uint32_t nb) {
  if (g->band_nb >= nb && g->band_halo.p && g->lead_version == g->tables_version) return LG_OK;
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(g->band_halo.alloc((size_t)2 * 2 * nb * 128));
  HIP_TRY(g->band_flags.alloc((size_t)2 * nb));
  HIP_TRY(hipMemsetAsync(g->band_flags.p, 0, (size_t)2 * nb * 8, g->stream));
  if (!g->band_err_dev.p) {
    HIP_TRY(g->band_err_dev.alloc(1));
    HIP_TRY(hipMemsetAsync(g->band_err_dev.p, 0, 8, g->stream));
  }
  if (!g->band_err_host) {
    HIP_TRY(host_alloc((void**)&g->band_err_host, 8));
    *g->band_err_host = 0;
  }
  if (g->lead_version != g->tables_version) {
    std::vector<uint64_t> solid, cj(g->cjump.n), lead;
    HIP_TRY(hipMemcpy(cj.data(), g->cjump.p, cj.size() * 8, hipMemcpyDeviceToHost));
    if (int rc = download_solid(g, solid)) return rc;
    g->lead_ok = lgca::build_lead_jumps(g->G.w, g->G.h, solid, cj, lead);
    HIP_TRY(g->lead.alloc(lead.size()));
    HIP_TRY(hipMemcpy(g->lead.p, lead.data(), lead.size() * 8, hipMemcpyHostToDevice));
    g->lead_version = g->tables_version;
  }
  if (!g->band_pw.p) {
    std::vector<uint64_t> pw;
    lgca::xs_pow4_windows(pw);
    HIP_TRY(g->band_pw.alloc(pw.size()));
    HIP_TRY(hipMemcpy(g->band_pw.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->band_nb = nb;  // the epoch keeps growing: cleared flags lie below any target
  return LG_OK;
}

// Block rows one band of `rows` rows can touch, for blocks `by` rows tall.
static uint32_t band_block_rows(uint32_t rows, uint32_t by) { return (rows - 1) / by + 2; }

bool band_sample_fits(const lg_t* g) {
  uint32_t rows = 0, nb = 0;
  if (!g->has_field || g->bw == 0 || g->bh == 0 || !band_eligible(g, &rows, &nb)) return false;
  // the block sums of one band live in the exchange slots (rows x 4 KB)
  return (size_t)band_block_rows(rows, g->by) * g->bw * 16 <= (size_t)rows * 4096;
}

// Launch n steps of the resident-band kernel on the handle's stream, planes
// a -> b; with `sample`, the launch also coarse-grains its final state into
// the attached field with weight alpha.
int launch_band(lg_t* g, uint32_t n, bool sample, float alpha) {
  uint32_t rows = 0, nb = 0;
  if (!band_eligible(g, &rows, &nb)) return fail(LG_EINVAL, "lattice not eligible for resident bands");
  if (int rc = ensure_band_buffers(g, nb)) return rc;
  if (sample && !band_sample_fits(g)) return fail(LG_EINVAL, "field blocks do not fit a band's LDS");
  const uint32_t nbr = sample ? band_block_rows(rows, g->by) : 0;
  if (sample) {
    const size_t np = (size_t)nb * nbr * g->bw * 4;
    if (g->band_part.n < np) HIP_TRY(g->band_part.alloc(np));
    if (g->band_bcnt.n < g->bh) {
      HIP_TRY(g->band_bcnt.alloc(g->bh));
      HIP_TRY(hipMemsetAsync(g->band_bcnt.p, 0, (size_t)g->bh * 4, g->stream));
    }
    if (!g->band_field.p) HIP_TRY(g->band_field.alloc(1));
    LgBandField f;
    std::memset(&f, 0, sizeof f);
    f.bx = g->bx;
    f.by = g->by;
    f.bw = g->bw;
    f.bh = g->bh;
    f.nbr = nbr;
    f.part = g->band_part.p;
    f.bcnt = g->band_bcnt.p;
    f.ux = g->fux.p;
    f.uy = g->fuy.p;
    f.rho = g->frho.p;
    f.sol = g->fsol.p;
    f.sums = g->fsums.p;
    if (std::memcmp(&f, &g->band_field_host, sizeof f) != 0) {  // attached or resized since
      HIP_TRY(hipStreamSynchronize(g->stream));  // no launch in flight reads the old one
      HIP_TRY(hipMemcpy(g->band_field.p, &f, sizeof f, hipMemcpyHostToDevice));
      g->band_field_host = f;
    }
  }
  LgBandArgs A;
  A.G = g->G;
  A.src = g->a.p;
  A.dst = g->b.p;
  A.solid = g->solid.p;
  A.cjump = g->cjump.p;
  A.row_band = g->row_band.p;
  A.ijump = g->ijump.p;
  A.icell = g->icell.p;
  A.halo = g->band_halo.p;
  A.flags = g->band_flags.p;
  A.err = g->band_err_dev.p;
  A.err_host = g->band_err_host;
  g->band_epoch += 1ull << 21;
  A.epoch = g->band_epoch;
  A.n = n;
  A.rows = rows;
  A.nb = nb;
  A.seed = g->seed;
  A.t0 = g->steps;
  // stream tables per band: its row bands (at most `rows` of them) and the
  // inflow stream
  const uint32_t cpu_band = lgca::band_rows(g->G.h, g->bands);
  A.nstreams = std::min<uint32_t>(rows, (rows + cpu_band - 2) / cpu_band + 1) + 1;
  // bands of 4m + 1 rows split their first row's draws between two helper
  // waves, where the workgroup has room for them
  A.split = (rows % 4 == 1 && rows + 3 <= LG_BAND_MAX_ROWS + 1) ? 1u : 0u;
  A.alpha = alpha;
  A.fs = g->band_field.p;
  A.pw = g->band_pw.p;
  A.lead = g->lead.p;
  A.lead_ok = g->lead_ok ? 1u : 0u;
  // inflow: the table wave draws every inflow cell of the band where that
  // takes it at most three passes over its lanes (64 / icpr rows a pass),
  // else each row's lanes draw their own (at most 64 cells) or the serial
  // per-word chain runs (wider)
  const uint32_t cols = g->G.inlet_cols;
  uint32_t cpr = 1;
  while (cpr < cols && cpr < 64) cpr <<= 1;
  A.icpr = cpr;
  const int inflow =
      cols > 64 ? 2 : ((rows + 64 / cpr - 1) / (64 / cpr) <= LG_BAND_ICHUNKS ? 0 : 1);
  void* args[] = {&A};
  const dim3 grid(nb), block(64 * (rows + 1 + 2 * A.split));
  // [rows][stream tables][helpers' digits][T^(4g) windows][inflow planes]
  // [rows' digits]
  const size_t lds = (size_t)(rows + A.nstreams) * 4096 +
                     (A.split ? 2 * 2 * LG_KDIGITS * 64 * 4 : 0) + 16 * 256 * 8 +
                     (size_t)2 * rows * 8 * 8 + (size_t)rows * 2 * LG_KDIGITS * 64 * 4 +
                     (inflow == 0 ? LG_BAND_ICHUNKS * 64 * 8 : 0);
  // [rows of 64 words][sampling][inflow]
  const void* const kernels[2][2][3] = {
      {{(const void*)lg_band_kernel<false, false, 0>, (const void*)lg_band_kernel<false, false, 1>,
        (const void*)lg_band_kernel<false, false, 2>},
       {(const void*)lg_band_kernel<false, true, 0>, (const void*)lg_band_kernel<false, true, 1>,
        (const void*)lg_band_kernel<false, true, 2>}},
      {{(const void*)lg_band_kernel<true, false, 0>, (const void*)lg_band_kernel<true, false, 1>,
        (const void*)lg_band_kernel<true, false, 2>},
       {(const void*)lg_band_kernel<true, true, 0>, (const void*)lg_band_kernel<true, true, 1>,
        (const void*)lg_band_kernel<true, true, 2>}}};
  if (lds > 64 * 1024 && !g->band_lds_raised) {
    for (int q = 0; q < 12; ++q)
      HIP_TRY(hipFuncSetAttribute(kernels[q / 6][(q / 3) & 1][q % 3],
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    g->band_lds_raised = true;
  }
  const void* kf = kernels[g->G.wpr == 64][sample][inflow];
  if (g->device < 0 || g->device >= LG_MAX_DEVICES) return fail(LG_EINVAL, "device index out of range");
  BandOrder& o = band_order();
  std::lock_guard<std::mutex> lk(o.mu);
  const int dev = g->device;
  hipError_t e = hipSuccess;
  if (!o.claimed[dev]) {
    // the claim: a cooperative launch of no steps (it only copies the state
    // to the next plane set, which this launch's steps then overwrite), so
    // that the call's own launch is already a plain one: the first plain
    // launch on a queue pays the runtime's set-up of its scratch memory
    LgBandArgs C = A;
    C.n = 0;
    void* cargs[] = {&C};
    e = hipLaunchCooperativeKernel(kernels[g->G.wpr == 64][0][inflow], grid, block, cargs, lds,
                                   g->stream);
    if (e == hipSuccess) o.claimed[dev] = true;
  } else if (o.last[dev] && o.last[dev] != g->stream) {
    if (!o.mark[dev]) HIP_TRY(hipEventCreateWithFlags(&o.mark[dev], hipEventDisableTiming));
    HIP_TRY(hipEventRecord(o.mark[dev], o.last[dev]));
    HIP_TRY(hipStreamWaitEvent(g->stream, o.mark[dev], 0));
  }
  if (e == hipSuccess) e = hipLaunchKernel(kf, grid, block, args, lds, g->stream);
  if (e != hipSuccess) {
    // Refused (the runtime can decline a cooperative launch, e.g. while
    // another process holds the device's cooperative-launch resources):
    // nothing ran; this lattice takes the other pipelines from here on.
    (void)hipGetLastError();
    g->band_refused = true;
    return kBandRefused;
  }
  o.last[dev] = g->stream;
  g->band_check = true;
  return LG_OK;
}

// After the handle's stream has drained: did a launch's wait run out?
int check_band(lg_t* g) {
  if (!g->band_check) return LG_OK;
  g->band_check = false;
  // epochs only grow: a failure since the last check carries a larger one
  const uint64_t e = __atomic_load_n(g->band_err_host, __ATOMIC_ACQUIRE);
  const bool failed = e > g->band_epoch_checked;
  g->band_epoch_checked = g->band_epoch;
  if (failed) {
    g->poisoned = true;
    return fail(LG_EDEVICE, "a resident-band wait timed out: the lattice's state is unknown");
  }
  return LG_OK;
}

}  // namespace lgca_internal
