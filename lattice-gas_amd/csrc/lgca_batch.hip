// Batched stepping: many independent lattices of uneven sizes advanced by one
// step kernel per step, driven by a tile table, with the step loop captured in
// a hipGraph and replayed.
//
// The reference never batched lattices ("would need replicas inside the step
// kernel", docs/plan.md:482-485). Here each member keeps its own handle
// (lg_t) and device buffers; a batch only adds
//   - a descriptor per member (geometry, plane pointers, stream tables, seed,
//     step counter and which plane set is current), in device memory;
//   - a tile table: tile i = (member, first word of a run of TILE words);
//   - per chunk of K steps: one tables launch for every member's streams, K
//     step launches over all tiles (grid-stride, so a fixed persistent grid
//     can walk any number of tiles), one launch that bumps the counters.
// With the two-pass step (the default) the tile table holds collide tiles
// and draw tiles, alternating: launch k collides step k of every member and
// draws its step k + 1 into the member's other digit slot, and a call begins
// with one draw-only launch for its first step.
// Every member computes exactly what its own lg_advance would: the per-unit
// work is the same device code with the member's own geometry and streams.
#include <chrono>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lgca_internal.h"

using namespace lgca_internal;

struct LgDesc {
  LgGeom G;
  uint64_t* p0;  // plane set 0 (7 planes)
  uint64_t* p1;  // plane set 1
  const uint64_t* solid;
  const uint64_t* cjump;
  const uint32_t* row_band;
  const uint64_t* ijump;
  const uint32_t* order;
  uint64_t* tabs;  // >= (K + 1) * (nbands + 1) * LG_WTAB
  uint32_t* kb;    // two digit slots (two-pass step)
  uint32_t* ib;    // two inflow slots
  uint64_t seed;
  uint64_t steps;      // step index of the current state
  uint64_t remaining;  // steps this member still takes in the current call
  uint32_t cur;        // plane set holding the current state
  uint32_t kpar;       // digit slot holding the current step's digits
  uint32_t kfresh;     // 1 until the call's first chunk: its first step is not drawn yet
  uint32_t pad;
  // coarse field (src/render.rs:21-48); fux == nullptr when none is attached
  float *fux, *fuy, *frho, *fsol;
  int64_t* fsums;
  uint32_t bx, by, bw, bh;
  uint64_t snap_off;   // offset of this member's snapshot in the staging buffer
};

// Stream tables for members blockIdx.y, steps steps..steps+nsteps-1.
__global__ void __launch_bounds__(256) lg_tables_batch_kernel(const LgDesc* __restrict D,
                                                              uint32_t nsteps) {
  const LgDesc& d = D[blockIdx.y];
  uint32_t per = d.G.nbands + 1;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nsteps * per * 16) return;
  uint32_t grp = g & 15u, ks = g >> 4;
  uint32_t k = ks / per, s = ks - k * per;
  if (k >= d.remaining) return;
  uint64_t t = d.steps + k;
  uint64_t st = lg_rng_init(s < d.G.nbands ? lg_band_seed(d.seed, t, s) : lg_inlet_seed(d.seed, t + 1));
  for (uint32_t i = 0; i < 4 * grp; ++i) st = lg_xs(st);
  uint64_t b0 = st, b1 = lg_xs(b0), b2 = lg_xs(b1), b3 = lg_xs(b2);
  uint64_t* out = d.tabs + (size_t)ks * LG_WTAB + grp * 16;
  for (uint32_t v = 0; v < 16; ++v) {
    uint64_t e = 0;
    if (v & 1u) e ^= b0;
    if (v & 2u) e ^= b1;
    if (v & 4u) e ^= b2;
    if (v & 8u) e ^= b3;
    out[v] = e;
  }
}

// The batch kernels are flattened: with a dozen kernels sharing the per-unit
// device functions the compiler would otherwise reach them through calls,
// paying register saves and stack on every unit (measured 1.8x slower).
#define LG_FLAT __attribute__((flatten))

// Step k of the chunk for every tile. A tile is TILE consecutive thread
// indices of one member, mapped to work units exactly as lg_step_kernel_{w,h} map
// them (the member's static word order, NH units per word). Register target
// WPE waves per SIMD: three, the default, leaves the tile loop room to keep
// everything in registers (134 VGPRs, no scratch); four spills 32 bytes a
// lane but hides more of the dependency waits, which pays on some mixes;
// two and four are there for lg_batch_set_waves and the autotuner.
template <int TILE, int NH, int WPE>
__global__ void __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(WPE))) LG_FLAT lg_step_batch_kernel(const LgDesc* __restrict D,
                                                             const uint2* __restrict tiles,
                                                             uint32_t ntiles, uint32_t k) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint2 tl = tiles[t];
    const LgDesc& d = D[tl.x];
    const uint32_t g = tl.y + threadIdx.x;
    if (k >= d.remaining || g >= d.G.total * NH) continue;
    const uint32_t e = d.order[NH == 1 ? g : g >> 1];
    const uint32_t par = (d.cur + k) & 1u;
    const uint64_t* src = par ? d.p1 : d.p0;
    uint64_t* dst = par ? d.p0 : d.p1;
    const uint64_t* tabs = d.tabs + (size_t)k * (d.G.nbands + 1) * LG_WTAB;
    lg_step_unit<NH>(d.G, src, dst, d.solid, d.cjump, d.row_band, d.ijump, tabs, e >> 16,
                     e & 0xFFFFu, NH == 1 ? 0u : (g & 1u));
  }
}

// Two-pass batch step. A tile is (member, first unit | role << 31): a
// collide tile runs units of step k from digit slot kpar + k, a draw tile
// draws units of step k + 1 into the other slot, both exactly as
// lg_step2_kernel maps them for a single lattice.
// Register target WPE waves per SIMD: by default five, as the single-lattice
// kernel's (about 100 VGPRs); left to itself the compiler unrolls the tile
// loop into 225 VGPRs for whole-word units, 2 waves per SIMD. Whole-word
// collide and draw lanes also come at four and six.
template <int TILE, int NH, int DNH, int WPE>
__global__ void __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(WPE))) LG_FLAT
    lg_step2_batch_kernel(const LgDesc* __restrict D, const uint2* __restrict tiles,
                          uint32_t ntiles, uint32_t k) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint2 tl = tiles[t];
    const LgDesc& d = D[tl.x];
    const uint32_t u = (tl.y & 0x7FFFFFFFu) + threadIdx.x;
    const size_t kstride = 2 * (size_t)d.G.total, istride = 2 * (size_t)d.G.iw * d.G.h;
    const uint32_t slot = (d.kpar + k) & 1u;
    if (tl.y >> 31) {
      if (k + 1 >= d.remaining || u >= d.G.total * DNH) continue;
      const uint32_t e = d.order[DNH == 1 ? u : u >> 1];
      const uint64_t* tabs = d.tabs + (size_t)(k + 1) * (d.G.nbands + 1) * LG_WTAB;
      lg_draw_unit<DNH>(d.G, d.solid, d.cjump, d.row_band, d.ijump, tabs, e >> 16, e & 0xFFFFu,
                        DNH == 1 ? 0u : (u & 1u), d.kb + (size_t)(slot ^ 1u) * LG_KDIGITS * kstride,
                        kstride, d.ib + (size_t)(slot ^ 1u) * 7 * istride, istride);
    } else {
      if (k >= d.remaining || u >= d.G.total * NH) continue;
      const uint32_t par = (d.cur + k) & 1u;
      const uint64_t* src = par ? d.p1 : d.p0;
      uint64_t* dst = par ? d.p0 : d.p1;
      const uint32_t word = NH == 1 ? u : u >> 1;
      const uint32_t y = word / d.G.wpr;
      lg_collide_unit<NH>(d.G, src, dst, d.solid, d.kb + (size_t)slot * LG_KDIGITS * kstride,
                          kstride, d.ib + (size_t)slot * 7 * istride, istride, y,
                          word - y * d.G.wpr, NH == 1 ? 0u : (u & 1u));
    }
  }
}

// The call's first step for members still fresh: its draw tiles only.
template <int TILE, int DNH>
__global__ void __launch_bounds__(TILE) LG_FLAT
    lg_draw0_batch_kernel(const LgDesc* __restrict D, const uint2* __restrict tiles,
                          uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint2 tl = tiles[t];
    const LgDesc& d = D[tl.x];
    const uint32_t u = (tl.y & 0x7FFFFFFFu) + threadIdx.x;
    if (!(tl.y >> 31) || !d.kfresh || d.remaining == 0 || u >= d.G.total * DNH) continue;
    const size_t kstride = 2 * (size_t)d.G.total, istride = 2 * (size_t)d.G.iw * d.G.h;
    const uint32_t e = d.order[DNH == 1 ? u : u >> 1];
    lg_draw_unit<DNH>(d.G, d.solid, d.cjump, d.row_band, d.ijump, d.tabs, e >> 16, e & 0xFFFFu,
                      DNH == 1 ? 0u : (u & 1u), d.kb + (size_t)d.kpar * LG_KDIGITS * kstride,
                      kstride, d.ib + (size_t)d.kpar * 7 * istride, istride);
  }
}

// After a chunk: advance each member's counters; fresh_after = 1 when the
// chunk ran the fused step (no digits were drawn ahead), so that a following
// two-pass chunk draws its own first step.
__global__ void lg_bump_kernel(LgDesc* __restrict D, uint32_t n, uint32_t count,
                               uint32_t fresh_after) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t adv = D[i].remaining < count ? D[i].remaining : count;
  D[i].steps += adv;
  D[i].cur ^= (uint32_t)(adv & 1u);
  D[i].remaining -= adv;
  D[i].kpar ^= count & 1u;
  D[i].kfresh = fresh_after;
}

// Coarse-grain every member with a field (blockIdx.y = member), exactly as
// lg_sample_kernel does, and copy the result (ux, uy, rho, solid arrays of
// the member, back to back) into the staging buffer for one read-back.
__global__ void __launch_bounds__(256) lg_snapshot_kernel(const LgDesc* __restrict D, float alpha,
                                                          float* __restrict staging) {
  const LgDesc& d = D[blockIdx.y];
  if (!d.fux) return;
  const uint32_t nb = d.bw * d.bh;
  const uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6);  // one wavefront per block
  if (g >= nb) return;
  const uint64_t* planes = d.cur ? d.p1 : d.p0;
  lg_sample_block_wave(d.G, planes, d.solid, d.bx, d.by, d.bw, alpha, d.fux, d.fuy, d.frho,
                       d.fsol, d.fsums, g, threadIdx.x & 63u);
  if ((threadIdx.x & 63u) == 0) {
    float* out = staging + d.snap_off;
    out[g] = d.fux[g];
    out[nb + g] = d.fuy[g];
    out[2 * nb + g] = d.frho[g];
    out[3 * nb + g] = d.fsol[g];
  }
}

struct lg_batch {
  std::vector<lg_t*> members;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;  // end of the last async advance; members' own calls wait for it
  DevBuf<LgDesc> desc;
  DevBuf<uint2> tiles128, tiles256;
  uint32_t ntiles128 = 0, ntiles256 = 0;
  uint32_t tiles_key = 0;   // pipeline the tile tables were built for (0: none)
  uint32_t max_tab_blocks_per_step = 0;  // tables launch width per step of chunk
  uint32_t kmax = 0;                     // largest chunk the members' tables hold
  // tuned launch shape
  uint32_t tile = 256;   // words per tile (128 or 256)
  uint32_t grid = 0;     // step-kernel workgroups; 0 = one per tile
  uint32_t chunk = 32;   // steps per graph replay (even)
  uint32_t units = 1;    // collide (or fused) lanes: 1 = 64 cells, 2 = 32 cells
  uint32_t passes = 2;   // 2: two-pass step, 1: fused step (set by size at creation)
  uint32_t draw_units = 1;  // draw lanes of the two-pass step
  uint32_t waves = 0;       // step kernel's register target, waves per SIMD (0: default)
  // graph for `chunk` steps with the current shape
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_chunk = 0, graph_tile = 0, graph_grid = 0, graph_key = 0, graph_waves = 0;
  int ncu = 256;
  // Descriptor uploads from two page-locked host slots in turn: a slot is
  // filled again only once the copy from it has run (its event), so an
  // upload never waits for the advance queued before it.
  LgDesc* hdesc[2] = {nullptr, nullptr};
  hipEvent_t desc_ev[2] = {nullptr, nullptr};
  bool desc_rec[2] = {false, false};
  uint32_t dslot = 0;                // slot of the latest upload
  std::vector<uint64_t> advanced;    // per member, steps taken by the last call
  // Snapshots of sampling advances: two slots (device staging, page-locked
  // copy, event after the copy), read back oldest first.
  DevBuf<float> staging[2];
  float* pin[2] = {nullptr, nullptr};
  size_t pin_bytes[2] = {0, 0};
  hipEvent_t snap_ev[2] = {nullptr, nullptr};
  uint32_t stage_head = 0, stage_count = 0;
  size_t staging_floats = 0;
  hipEvent_t own_ev = nullptr;       // marks a member's own stream for the batch to wait on
  uint32_t max_field_blocks = 0;
  uint32_t calls = 0;              // advance calls on this membership
  std::vector<uint2> host_tiles128, host_tiles256;
};

namespace {

void drop_graph(lg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_chunk = 0;
}

// The launch pipeline as one key (tile tables and graphs depend on it).
uint32_t pipeline_key(const lg_batch* b) {
  return b->passes * 100 + b->units * 10 + (b->passes == 2 ? b->draw_units : 0);
}

// Fill host descriptor slot h (no device work, no allocation: the slots are
// page-locked buffers made at batch creation).
void fill_desc(lg_batch* b, LgDesc* h, const uint64_t* limits, uint64_t n) {
  size_t off = 0;
  for (size_t i = 0; i < b->members.size(); ++i) {
    lg_t* g = b->members[i];
    LgDesc& d = h[i];
    std::memset(&d, 0, sizeof(d));
    d.G = g->G;
    d.p0 = g->a.p;
    d.p1 = g->b.p;
    d.solid = g->solid.p;
    d.cjump = g->cjump.p;
    d.row_band = g->row_band.p;
    d.ijump = g->ijump.p;
    d.order = g->order.p;
    d.tabs = g->basis.p;
    g->tab_n = 0;  // the batch's tables kernel overwrites them
    d.kb = g->kbuf.p;
    d.ib = g->ibuf.p;
    d.seed = g->seed;
    d.steps = g->steps;
    d.remaining = limits ? std::min<uint64_t>(limits[i], n) : n;
    d.cur = 0;
    d.kpar = 0;
    d.kfresh = 1;
    if (g->has_field) {
      d.fux = g->fux.p;
      d.fuy = g->fuy.p;
      d.frho = g->frho.p;
      d.fsol = g->fsol.p;
      d.fsums = g->fsums.p;
      d.bx = g->bx;
      d.by = g->by;
      d.bw = g->bw;
      d.bh = g->bh;
      d.snap_off = off;
      off += (size_t)4 * g->bw * g->bh;
    }
  }
  b->staging_floats = off;
}

// Queue the descriptors of a call on the batch stream, from the host slot
// not used by the latest upload (waiting only for the copy from that slot
// two uploads ago). Allocates nothing.
int queue_desc(lg_batch* b, const uint64_t* limits, uint64_t n) {
  const uint32_t slot = b->dslot ^ 1u;
  if (b->desc_rec[slot]) HIP_TRY(hipEventSynchronize(b->desc_ev[slot]));
  fill_desc(b, b->hdesc[slot], limits, n);
  HIP_TRY(hipMemcpyAsync(b->desc.p, b->hdesc[slot], b->members.size() * sizeof(LgDesc),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipEventRecord(b->desc_ev[slot], b->stream));
  b->desc_rec[slot] = true;
  b->dslot = slot;
  return LG_OK;
}

// The members' stream-jump tables, built where their obstacles or inflow
// changed since the last build: on several host threads when there are many
// (each build is host work of about a millisecond at 2048 x 2048, and a new
// batch of hundreds of members would otherwise build them one after another
// inside its first call). Each member's build uses only its own stream.
int ensure_member_tables(lg_batch* b) {
  std::vector<lg_t*> todo;
  for (lg_t* g : b->members)
    if (g->tables_dirty) todo.push_back(g);
  const size_t nthreads = std::min<size_t>(
      todo.size() / 4, std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency())));
  if (nthreads < 2) {
    for (lg_t* g : todo)
      if (int rc = ensure_tables(g)) return rc;
    return LG_OK;
  }
  std::atomic<size_t> next{0};
  std::vector<int> rcs(nthreads, LG_OK);
  std::vector<std::string> msgs(nthreads);
  std::vector<std::thread> pool;
  for (size_t t = 0; t < nthreads; ++t)
    pool.emplace_back([&, t] {
      for (size_t i; (i = next.fetch_add(1)) < todo.size();) {
        // the member's device (a thread's current device starts at 0)
        int rc = hipSetDevice(todo[i]->device) == hipSuccess
                     ? ensure_tables(todo[i])
                     : fail(LG_EDEVICE, "hipSetDevice failed for a batch member's device");
        if (rc) {
          rcs[t] = rc;
          msgs[t] = g_err;  // the worker's own error text
          return;
        }
      }
    });
  for (std::thread& th : pool) th.join();
  for (size_t t = 0; t < nthreads; ++t)
    if (rcs[t]) return fail(rcs[t], msgs[t]);
  return LG_OK;
}

int upload_desc(lg_batch* b, const uint64_t* limits, uint64_t n) {
  if (int rc = ensure_member_tables(b)) return rc;
  for (lg_t* g : b->members)
    if (int rc = ensure_draw_buffers(g)) return rc;  // the pipeline may change within a call
  return queue_desc(b, limits, n);
}

// Tiles of TILE units: for the fused step, every member's units; for the
// two-pass step, each member's collide tiles and draw tiles (role bit 31)
// alternating. The list, member after member, is then cut into eight equal
// contiguous segments laid out so that segment x holds the positions equal to
// x mod 8: blocks go to the eight XCDs in turn, so each XCD's L2 serves about
// an eighth of the members, and with them an eighth of the per-member stream
// tables every unit's jump reads (17 x 2 KB per member and step). Measured
// on the 256-lattice ensemble: 806 GCUPS against 776-780 with the plain
// member order; served batches unchanged. (Placing whole members on one XCD
// instead confines a large member to an eighth of the chip: 499 against 579
// GCUPS served.) Positions past the list's end within the last row of eight
// hold a no-op tile (member 0, unit 0x7FFFFFFF), which every kernel skips as
// beyond its member's units.
template <int TILE>
int build_tiles(lg_batch* b, DevBuf<uint2>& buf, uint32_t& n, std::vector<uint2>& t) {
  std::vector<uint2> all, c, dr;
  for (size_t i = 0; i < b->members.size(); ++i) {
    c.clear();
    dr.clear();
    uint32_t total = b->members[i]->G.total * b->units;
    for (uint32_t g0 = 0; g0 < total; g0 += TILE) c.push_back(make_uint2((uint32_t)i, g0));
    if (b->passes == 2) {
      uint32_t dtotal = b->members[i]->G.total * b->draw_units;
      for (uint32_t g0 = 0; g0 < dtotal; g0 += TILE)
        dr.push_back(make_uint2((uint32_t)i, g0 | 0x80000000u));
    }
    size_t a = 0, q = 0;
    while (a < c.size() || q < dr.size()) {
      if (a < c.size()) all.push_back(c[a++]);
      if (q < dr.size()) all.push_back(dr[q++]);
    }
  }
  const size_t N = all.size(), L = (N + 7) / 8;
  t.clear();
  for (size_t r = 0; r < L; ++r)
    for (size_t x = 0; x < 8; ++x) {
      const size_t i = x * L + r;
      t.push_back(i < N ? all[i] : make_uint2(0u, 0x7FFFFFFFu));
    }
  while (!t.empty() && t.back().y == 0x7FFFFFFFu) t.pop_back();
  n = (uint32_t)t.size();
  HIP_TRY(buf.alloc(std::max<size_t>(t.size(), 1)));
  if (!t.empty())
    HIP_TRY(hipMemcpyAsync(buf.p, t.data(), t.size() * sizeof(uint2), hipMemcpyHostToDevice,
                           b->stream));
  return LG_OK;
}

// The register target (waves per SIMD) the batch's step kernel runs at: the
// requested one where that pipeline has an instance for it, else the
// pipeline's default.
uint32_t effective_waves(const lg_batch* b) {
  if (b->passes == 1) {
    // whole-word lanes: three by default. Which is faster depends on the
    // mix (GCUPS at three / four waves per SIMD: the 256-lattice ensemble
    // 951 / 1020, 36 of its members 876 / 910; batches of 10-40 served
    // requests, up to 4096 x 2304 with inflow, 692 / 683 to 962 / 929): the
    // autotuner decides per batch.
    if (b->units == 1) return (b->waves >= 2 && b->waves <= 4) ? b->waves : 3;
    return 3;
  }
  if (b->units == 1 && b->draw_units == 1 && (b->waves == 4 || b->waves == 6)) return b->waves;
  return 5;
}

// Enqueue `count` steps (count <= kmax) on the batch stream.
template <int TILE>
void launch_steps(lg_batch* b, uint32_t grid, const uint2* tiles, uint32_t ntiles, uint32_t k,
                  bool first) {
  const dim3 gd(grid), bd(TILE);
  const uint32_t wpe = effective_waves(b);
  if (b->passes == 1) {
    if (b->units == 2)
      hipLaunchKernelGGL((lg_step_batch_kernel<TILE, 2, 3>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles, k);
    else if (wpe == 2)
      hipLaunchKernelGGL((lg_step_batch_kernel<TILE, 1, 2>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles, k);
    else if (wpe == 4)
      hipLaunchKernelGGL((lg_step_batch_kernel<TILE, 1, 4>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles, k);
    else
      hipLaunchKernelGGL((lg_step_batch_kernel<TILE, 1, 3>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles, k);
    return;
  }
  if (first) {
    if (b->draw_units == 1)
      hipLaunchKernelGGL((lg_draw0_batch_kernel<TILE, 1>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles);
    else
      hipLaunchKernelGGL((lg_draw0_batch_kernel<TILE, 2>), gd, bd, 0, b->stream, b->desc.p, tiles,
                         ntiles);
  }
  if (b->units == 1 && b->draw_units == 1 && wpe == 4)
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 1, 1, 4>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
  else if (b->units == 1 && b->draw_units == 1 && wpe == 6)
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 1, 1, 6>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
  else if (b->units == 1 && b->draw_units == 1)
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 1, 1, 5>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
  else if (b->units == 1)
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 1, 2, 5>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
  else if (b->draw_units == 1)
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 2, 1, 5>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
  else
    hipLaunchKernelGGL((lg_step2_batch_kernel<TILE, 2, 2, 5>), gd, bd, 0, b->stream, b->desc.p,
                       tiles, ntiles, k);
}

int enqueue_chunk(lg_batch* b, uint32_t count) {
  const uint32_t nm = (uint32_t)b->members.size();
  // two-pass: one table more, for the draws of the next chunk's first step
  const uint32_t ntab = count + (b->passes == 2 ? 1u : 0u);
  dim3 tgrid(b->max_tab_blocks_per_step * ntab, nm);
  hipLaunchKernelGGL(lg_tables_batch_kernel, tgrid, dim3(256), 0, b->stream, b->desc.p, ntab);
  HIP_TRY(hipGetLastError());
  const uint32_t ntiles = b->tile == 128 ? b->ntiles128 : b->ntiles256;
  const uint2* tiles = b->tile == 128 ? b->tiles128.p : b->tiles256.p;
  uint32_t grid = b->grid ? std::min(b->grid, ntiles) : ntiles;
  for (uint32_t k = 0; k < count; ++k) {
    if (b->tile == 128)
      launch_steps<128>(b, grid, tiles, ntiles, k, k == 0);
    else
      launch_steps<256>(b, grid, tiles, ntiles, k, k == 0);
  }
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(lg_bump_kernel, dim3(blocks_for(nm)), dim3(256), 0, b->stream, b->desc.p, nm,
                     count, b->passes == 1 ? 1u : 0u);
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

int ensure_tiles(lg_batch* b) {
  if (b->tiles_key == pipeline_key(b)) return LG_OK;
  HIP_TRY(hipStreamSynchronize(b->stream));  // the old tables may still be read
  int rc = build_tiles<128>(b, b->tiles128, b->ntiles128, b->host_tiles128);
  if (rc) return rc;
  rc = build_tiles<256>(b, b->tiles256, b->ntiles256, b->host_tiles256);
  if (rc) return rc;
  b->tiles_key = pipeline_key(b);
  return LG_OK;
}

int ensure_graph(lg_batch* b) {
  if (b->exec && b->graph_chunk == b->chunk && b->graph_tile == b->tile &&
      b->graph_grid == b->grid && b->graph_key == pipeline_key(b) &&
      b->graph_waves == effective_waves(b))
    return LG_OK;
  drop_graph(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_chunk(b, b->chunk);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return fail(LG_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  b->graph = gr;
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_chunk = b->chunk;
  b->graph_tile = b->tile;
  b->graph_grid = b->grid;
  b->graph_key = pipeline_key(b);
  b->graph_waves = effective_waves(b);
  return LG_OK;
}

// Advance every member by n steps; the descriptors must be current.
int run_steps(lg_batch* b, uint64_t n, bool use_graph) {
  int trc = ensure_tiles(b);
  if (trc) return trc;
  uint64_t left = n;
  if (use_graph && left >= b->chunk) {
    int rc = ensure_graph(b);
    if (rc) return rc;
    while (left >= b->chunk) {
      HIP_TRY(hipGraphLaunch(b->exec, b->stream));
      left -= b->chunk;
    }
  }
  while (left > 0) {
    uint32_t c = (uint32_t)std::min<uint64_t>(left, b->kmax);
    int rc = enqueue_chunk(b, c);
    if (rc) return rc;
    left -= c;
  }
  return LG_OK;
}

void commit_members(lg_batch* b, const uint64_t* limits, uint64_t n) {
  b->advanced.assign(b->members.size(), 0);
  for (size_t i = 0; i < b->members.size(); ++i) {
    lg_t* g = b->members[i];
    uint64_t adv = limits ? std::min<uint64_t>(limits[i], n) : n;
    g->steps += adv;
    if (adv & 1u) std::swap(g->a.p, g->b.p);
    b->advanced[i] = adv;
  }
}

}  // namespace

// A batch call on a member that lg_destroy has released to the batch would
// run on a handle its owner has given up: refuse it.
static int members_alive(const lg_batch* b) {
  for (const lg_t* g : b->members) {
    if (g->destroyed) return fail(LG_EINVAL, "a member of the batch was destroyed; destroy the batch");
    if (g->poisoned)
      return fail(LG_ESTATE, "a member's state is unknown after a failed batch call; reload it");
  }
  return LG_OK;
}

// A batch call that fails after some of its steps were queued: the members'
// host counters and plane pointers no longer say where their device state
// is, so wait for what was queued and retire the members (their calls fail
// with LG_ESTATE from now on; they can only be destroyed, or reloaded).
static int abandon_queued(lg_batch* b, int rc) {
  const std::string msg = g_err;
  (void)hipStreamSynchronize(b->stream);
  for (lg_t* g : b->members) g->poisoned = true;
  return fail(rc, msg + " (batch members left in an unknown state)");
}

// Wait for everything queued for the members: by their own calls and by
// async batch advances (this batch's or another's).
static int settle_members(lg_batch* b) {
  for (lg_t* g : b->members) {
    if (int rc = settle(g)) return rc;
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  return LG_OK;
}

// The same order kept on the device instead, for a call that queues work and
// returns: the batch stream waits for the async batch advance that last
// stepped each member (one wait per distinct advance: usually the previous
// batch's) and for work still queued on a member's own stream. The host does
// not wait, so a serving loop can queue the next quantum, on a batch of
// another membership, before the last one has ended.
static int order_members(lg_batch* b) {
  hipEvent_t waited[8];
  int nw = 0;
  for (lg_t* g : b->members) {
    if (g->poisoned)
      return fail(LG_ESTATE, "the lattice's state is unknown after a failed batch call; reload it");
    if (g->after && g->after != b->done) {
      bool seen = false;
      for (int i = 0; i < nw; ++i) seen = seen || waited[i] == g->after;
      if (!seen) {
        HIP_TRY(hipStreamWaitEvent(b->stream, g->after, 0));
        if (nw < 8) waited[nw++] = g->after;
      }
    }
    const hipError_t q = hipStreamQuery(g->stream);
    if (q == hipErrorNotReady) {
      HIP_TRY(hipEventRecord(b->own_ev, g->stream));
      HIP_TRY(hipStreamWaitEvent(b->stream, b->own_ev, 0));
    } else if (q != hipSuccess) {
      return fail(LG_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(q));
    }
  }
  return LG_OK;
}

// The batch's page-locked slots and events (its stream has drained).
static void release_host_side(lg_batch* b) {
  for (int q = 0; q < 2; ++q) {
    if (b->hdesc[q]) host_release(b->hdesc[q]);
    if (b->pin[q]) host_release(b->pin[q]);
    if (b->desc_ev[q]) (void)hipEventDestroy(b->desc_ev[q]);
    if (b->snap_ev[q]) (void)hipEventDestroy(b->snap_ev[q]);
    b->hdesc[q] = nullptr;
    b->pin[q] = nullptr;
    b->desc_ev[q] = b->snap_ev[q] = nullptr;
  }
  if (b->own_ev) (void)hipEventDestroy(b->own_ev);
  b->own_ev = nullptr;
}

extern "C" {

lg_batch_t* lg_batch_create(lg_t* const* members, size_t n) {
  if (!members || n == 0) {
    fail(LG_EINVAL, "a batch needs at least one lattice");
    return nullptr;
  }
  auto* b = new lg_batch();
  auto bail = [&](const std::string& m, int code = LG_EDEVICE) -> lg_batch_t* {
    fail(code, m);
    release_host_side(b);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return nullptr;
  };
  b->members.assign(members, members + n);
  for (size_t i = 0; i < n; ++i) {
    if (!members[i]) return bail("null member", LG_EINVAL);
    if (members[i]->destroyed) return bail("a member was destroyed", LG_EINVAL);
    // a tile's unit index (word x cells-per-lane split) must stay below bit 31,
    // which marks draw tiles, and below the no-op tile's 0x7FFFFFFF
    if (members[i]->G.total >= (1u << 30))
      return bail("a batch member must have fewer than 2^30 words per plane", LG_EINVAL);
    for (size_t k = 0; k < i; ++k)
      if (members[k] == members[i]) return bail("a lattice appears twice in the batch", LG_EINVAL);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail("hipStreamCreate failed");
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    b->done = nullptr;
    return bail("hipEventCreate failed");
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, members[0]->device) == hipSuccess) b->ncu = p.multiProcessorCount;
  b->kmax = ~0u;
  for (lg_t* g : b->members) {
    b->kmax = std::min(b->kmax, g->basis_steps - 1);  // one table spare for the two-pass step
    uint32_t per = g->G.nbands + 1;
    b->max_tab_blocks_per_step = std::max(b->max_tab_blocks_per_step, blocks_for((size_t)per * 16));
  }
  // every member's tables hold at least 3 steps (lg_create), so a chunk of
  // 2 always fits beside the two-pass step's spare table
  if (b->kmax < 2) return bail("a member's stream tables hold fewer than 3 steps", LG_EINVAL);
  b->chunk = std::min<uint32_t>(32, b->kmax & ~1u);
  // the same size rule as one lattice's default step, over all members' words
  size_t words = 0;
  for (lg_t* g : b->members) words += g->G.total;
  b->passes = words >= LG_AUTO_FUSED_WORDS ? 1u : 2u;
  // The fused step runs best on 128-word tiles (the 256-lattice ensemble: 958
  // against 920 GCUPS; served batches: 591-597 against 572-595).
  b->tile = b->passes == 1 ? 128u : 256u;
  if (b->desc.alloc(n)) return bail("hipMalloc failed for descriptors");
  for (int q = 0; q < 2; ++q) {
    if (host_alloc((void**)&b->hdesc[q], n * sizeof(LgDesc)) != hipSuccess) {
      b->hdesc[q] = nullptr;
      return bail("page-locked allocation failed for descriptors");
    }
    if (hipEventCreateWithFlags(&b->desc_ev[q], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->snap_ev[q], hipEventDisableTiming) != hipSuccess)
      return bail("hipEventCreate failed");
  }
  if (hipEventCreateWithFlags(&b->own_ev, hipEventDisableTiming) != hipSuccess)
    return bail("hipEventCreate failed");
  if (ensure_tiles(b)) return bail(g_err);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return bail("hipStreamSynchronize failed");
  for (lg_t* g : b->members) g->batch_refs += 1;
  return b;
}

void lg_batch_destroy(lg_batch_t* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  b->desc.release();
  b->tiles128.release();
  b->tiles256.release();
  b->staging[0].release();
  b->staging[1].release();
  release_host_side(b);
  for (lg_t* g : b->members)
    if (g->after == b->done) g->after = nullptr;  // ended: the stream was synchronised
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (lg_t* g : b->members)
    if (--g->batch_refs == 0 && g->destroyed) destroy_lattice(g);
  delete b;
}

int lg_batch_advance_limited(lg_batch_t* b, uint64_t n, const uint64_t* limits) {
  if (!b) return fail(LG_EINVAL, "null batch");
  b->stage_count = 0;  // unread snapshots are discarded
  if (int rc = members_alive(b)) return rc;
  if (n == 0) return LG_OK;
  if (int rc = settle_members(b)) return rc;
  int rc = upload_desc(b, limits, n);
  if (rc) return rc;
  // a fresh membership runs its first call eagerly; from the second call on
  // the chunk is captured once and replayed
  rc = run_steps(b, n, b->calls > 0);
  if (rc) return abandon_queued(b, rc);
  HIP_TRY(hipStreamSynchronize(b->stream));
  commit_members(b, limits, n);
  b->calls += 1;
  return LG_OK;
}

int lg_batch_advance(lg_batch_t* b, uint64_t n) { return lg_batch_advance_limited(b, n, nullptr); }

// Queue an advance, and with `sample` one sample of every member's field at
// its end into the next snapshot slot (the descriptors the advance leaves on
// the device name each member's current planes), copied on to the slot's
// page-locked buffer on the batch stream. Nothing waits on the host: the
// batch stream waits on the device for the members' earlier work
// (order_members), and the descriptors go up from the host slot the last
// call did not use, so a second advance can be queued behind this one.
static int advance_async(lg_batch* b, uint64_t n, const uint64_t* limits, bool sample,
                         float alpha) {
  if (int rc = members_alive(b)) return rc;
  if (n == 0 && !sample) return LG_OK;
  if (sample && b->stage_count == 2)
    return fail(LG_ESTATE, "two snapshots wait to be read (lg_batch_read_snapshot) first");
  if (int rc = order_members(b)) return rc;
  int rc = upload_desc(b, limits, n);
  if (rc) return rc;
  // everything that can fail without device work queued comes first: the
  // snapshot slot is sized by upload_desc and allocated before any step
  const uint32_t slot = (b->stage_head + b->stage_count) & 1u;
  const size_t bytes = b->staging_floats * 4;
  if (sample && bytes) {
    if (b->staging[slot].n < b->staging_floats) HIP_TRY(b->staging[slot].alloc(b->staging_floats));
    if (b->pin_bytes[slot] < bytes) {
      if (b->pin[slot]) host_release(b->pin[slot]);
      b->pin[slot] = nullptr;
      b->pin_bytes[slot] = 0;
      HIP_TRY(host_alloc((void**)&b->pin[slot], bytes));
      b->pin_bytes[slot] = bytes;
    }
  }
  if (n) {
    rc = run_steps(b, n, b->calls > 0);
    if (rc) return abandon_queued(b, rc);
  }
  if (sample && bytes) {
    uint32_t maxnb = 0;
    for (lg_t* g : b->members)
      if (g->has_field) maxnb = std::max(maxnb, g->bw * g->bh);
    dim3 grid(blocks_for(maxnb, 4), (uint32_t)b->members.size());
    hipLaunchKernelGGL(lg_snapshot_kernel, grid, dim3(256), 0, b->stream, b->desc.p, alpha,
                       b->staging[slot].p);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->pin[slot], b->staging[slot].p, bytes, hipMemcpyDeviceToHost, b->stream);
    if (e != hipSuccess)
      return abandon_queued(b, fail(LG_EDEVICE, std::string("snapshot: ") + hipGetErrorString(e)));
  }
  if (sample) {
    hipError_t e = hipEventRecord(b->snap_ev[slot], b->stream);
    if (e != hipSuccess)
      return abandon_queued(b, fail(LG_EDEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e)));
    b->stage_count += 1;
  }
  // a member's own calls wait for this advance first (settle), and a later
  // batch's stream on the device (order_members): no wait is queued on the
  // members' streams, which would hold up unrelated work sharing their
  // hardware queues
  hipError_t e = hipEventRecord(b->done, b->stream);
  if (e != hipSuccess)
    return abandon_queued(b, fail(LG_EDEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e)));
  for (lg_t* g : b->members) g->after = b->done;
  commit_members(b, limits, n);
  if (n) b->calls += 1;
  return LG_OK;
}

int lg_batch_advance_async(lg_batch_t* b, uint64_t n, const uint64_t* limits) {
  if (!b) return fail(LG_EINVAL, "null batch");
  return advance_async(b, n, limits, false, 0.0f);
}

int lg_batch_advance_sampling_async(lg_batch_t* b, uint64_t n, const uint64_t* limits,
                                    float alpha) {
  if (!b) return fail(LG_EINVAL, "null batch");
  return advance_async(b, n, limits, true, alpha);
}

int lg_batch_read_snapshot(lg_batch_t* b, float* out, size_t nfloats) {
  if (!b || (!out && nfloats)) return fail(LG_EINVAL, "null argument");
  if (!b->stage_count) return fail(LG_ESTATE, "no unread snapshot of a sampling advance");
  if (nfloats != b->staging_floats)
    return fail(LG_EINVAL, "output size differs from the members' fields (4 x blocks each)");
  const uint32_t slot = b->stage_head;
  HIP_TRY(hipEventSynchronize(b->snap_ev[slot]));
  if (nfloats) std::memcpy(out, b->pin[slot], nfloats * 4);
  b->stage_head ^= 1u;
  b->stage_count -= 1;
  return LG_OK;
}

int lg_batch_poll(lg_batch_t* b) {
  if (!b) return fail(LG_EINVAL, "null batch");
  hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(LG_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int lg_batch_wait(lg_batch_t* b) {
  if (!b) return fail(LG_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return LG_OK;
}

}  // extern "C"

// Sample every member's field into `staging` (device, >= staging_floats
// floats) and copy it out; the descriptors on the device must be current.
static int snapshot_into(lg_batch* b, float alpha, float* staging, float* out, size_t nfloats) {
  uint32_t maxnb = 0;
  for (lg_t* g : b->members)
    if (g->has_field) maxnb = std::max(maxnb, g->bw * g->bh);
  dim3 grid(blocks_for(maxnb, 4), (uint32_t)b->members.size());
  hipLaunchKernelGGL(lg_snapshot_kernel, grid, dim3(256), 0, b->stream, b->desc.p, alpha, staging);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, staging, nfloats * 4, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return LG_OK;
}

extern "C" {

int lg_batch_snapshot(lg_batch_t* b, float alpha, float* out, size_t nfloats) {
  if (!b || !out) return fail(LG_EINVAL, "null argument");
  b->stage_count = 0;  // unread snapshots are discarded
  if (int rc = members_alive(b)) return rc;
  if (int rc = settle_members(b)) return rc;
  int rc = upload_desc(b, nullptr, 0);
  if (rc) return rc;
  if (nfloats != b->staging_floats)
    return fail(LG_EINVAL, "output size differs from the members' fields (4 x blocks each)");
  if (nfloats == 0) return LG_OK;
  if (b->staging[0].n < nfloats) HIP_TRY(b->staging[0].alloc(nfloats));
  return snapshot_into(b, alpha, b->staging[0].p, out, nfloats);
}

size_t lg_batch_snapshot_size(lg_batch_t* b) {
  if (!b) return 0;
  size_t n = 0;
  for (lg_t* g : b->members)
    if (g->has_field) n += (size_t)4 * g->bw * g->bh;
  return n;
}

int lg_batch_set_units(lg_batch_t* b, uint32_t cells_per_lane) {
  if (!b) return fail(LG_EINVAL, "null batch");
  if (cells_per_lane != 64 && cells_per_lane != 32)
    return fail(LG_EINVAL, "cells per lane must be 64 or 32");
  b->units = cells_per_lane == 64 ? 1u : 2u;
  return LG_OK;
}

int lg_batch_set_passes(lg_batch_t* b, uint32_t passes) {
  if (!b) return fail(LG_EINVAL, "null batch");
  if (passes != 1 && passes != 2) return fail(LG_EINVAL, "passes per step must be 1 or 2");
  b->passes = passes;
  return LG_OK;
}

int lg_batch_set_draw_unit(lg_batch_t* b, uint32_t cells_per_lane) {
  if (!b) return fail(LG_EINVAL, "null batch");
  if (cells_per_lane != 64 && cells_per_lane != 32)
    return fail(LG_EINVAL, "cells per lane must be 64 or 32");
  b->draw_units = cells_per_lane == 64 ? 1u : 2u;
  return LG_OK;
}

int lg_batch_get_pipeline(lg_batch_t* b, uint32_t* passes, uint32_t* cells_per_lane,
                          uint32_t* draw_cells_per_lane) {
  if (!b || !passes || !cells_per_lane || !draw_cells_per_lane)
    return fail(LG_EINVAL, "null argument");
  *passes = b->passes;
  *cells_per_lane = b->units == 1 ? 64u : 32u;
  *draw_cells_per_lane = b->draw_units == 1 ? 64u : 32u;
  return LG_OK;
}

int lg_batch_set_shape(lg_batch_t* b, uint32_t tile, uint32_t grid, uint32_t chunk) {
  if (!b) return fail(LG_EINVAL, "null batch");
  if (tile != 128 && tile != 256) return fail(LG_EINVAL, "tile must be 128 or 256 words");
  if (chunk == 0 || (chunk & 1u) || chunk > b->kmax)
    return fail(LG_EINVAL, "chunk must be even and at most the members' table capacity");
  b->tile = tile;
  b->grid = grid;
  b->chunk = chunk;
  return LG_OK;
}

int lg_batch_set_waves(lg_batch_t* b, uint32_t waves_per_simd) {
  if (!b) return fail(LG_EINVAL, "null batch");
  if (waves_per_simd != 0 && (waves_per_simd < 2 || waves_per_simd > 6))
    return fail(LG_EINVAL, "waves per SIMD must be 0 (default) or 2..6");
  b->waves = waves_per_simd;
  return LG_OK;
}

int lg_batch_get_waves(lg_batch_t* b, uint32_t* waves_per_simd) {
  if (!b || !waves_per_simd) return fail(LG_EINVAL, "null argument");
  *waves_per_simd = effective_waves(b);
  return LG_OK;
}

int lg_batch_get_units(lg_batch_t* b, uint32_t* cells_per_lane) {
  if (!b || !cells_per_lane) return fail(LG_EINVAL, "null argument");
  *cells_per_lane = b->units == 1 ? 64u : 32u;
  return LG_OK;
}

int lg_batch_get_shape(lg_batch_t* b, uint32_t* tile, uint32_t* grid, uint32_t* chunk) {
  if (!b || !tile || !grid || !chunk) return fail(LG_EINVAL, "null argument");
  *tile = b->tile;
  *grid = b->grid;
  *chunk = b->chunk;
  return LG_OK;
}

int lg_batch_autotune(lg_batch_t* b, uint64_t budget_steps, uint64_t* used_steps) {
  if (!b || !used_steps) return fail(LG_EINVAL, "null argument");
  *used_steps = 0;
  b->stage_count = 0;  // unread snapshots are discarded
  if (int rc = members_alive(b)) return rc;
  if (int rc = settle_members(b)) return rc;
  int rc = upload_desc(b, nullptr, ~0ull);
  if (rc) return rc;
  struct Cand {
    uint32_t passes, units, draw_units, tile, grid, waves;
  };
  // Most promising first, so that a budget that runs out keeps the best so
  // far: the pipeline the size rule picks leads (at 256 lattices of 256..2048
  // a side the fused step's best shape, 128-word tiles, runs 958 GCUPS
  // against 920 for 256-word tiles; a single lattice's size favours the
  // two-pass step), then its register targets (waves per SIMD), then the
  // other tile sizes, grids and lane widths.
  const uint32_t g16 = 16u * (uint32_t)b->ncu, g32 = 32u * (uint32_t)b->ncu;
  const Cand two_pass[] = {
      {2, 1, 1, 256, 0, 0},   {2, 1, 1, 256, 0, 4}, {2, 1, 1, 256, 0, 6}, {2, 1, 1, 128, 0, 0},
      {2, 1, 2, 256, 0, 0},   {2, 2, 2, 256, 0, 0}, {2, 1, 1, 256, g16, 0}, {2, 2, 1, 256, 0, 0},
  };
  const Cand fused[] = {
      {1, 1, 1, 128, 0, 0},   {1, 1, 1, 128, 0, 4}, {1, 1, 1, 128, 0, 2}, {1, 1, 1, 256, 0, 0},
      {1, 1, 1, 256, g32, 0}, {1, 2, 1, 256, 0, 0}, {1, 1, 1, 256, g16, 0},
  };
  size_t words = 0;
  for (lg_t* g : b->members) words += g->G.total;
  std::vector<Cand> cands;
  if (words >= LG_AUTO_FUSED_WORDS) {
    cands.assign(std::begin(fused), std::end(fused));
    cands.insert(cands.end(), std::begin(two_pass), std::end(two_pass));
  } else {
    cands.assign(std::begin(two_pass), std::end(two_pass));
    cands.insert(cands.end(), std::begin(fused), std::end(fused));
  }
  const uint32_t chunk = std::min<uint32_t>(32, b->kmax & ~1u);  // >= 2 (lg_batch_create)
  double best = 1e300;
  Cand bestc = {b->passes, b->units, b->draw_units, b->tile, b->grid, b->waves};
  uint64_t used = 0;
  // one warm replay (it captures the graph), then the best of three timed ones
  constexpr int kTimed = 3;
  for (const Cand& c : cands) {
    const uint64_t trial = chunk;
    if (budget_steps && used + (1 + kTimed) * trial > budget_steps) break;
    b->passes = c.passes;
    b->units = c.units;
    b->draw_units = c.draw_units;
    b->tile = c.tile;
    b->grid = c.grid;
    b->waves = c.waves;
    b->chunk = chunk;
    if ((rc = run_steps(b, trial, true))) return abandon_queued(b, rc);
    HIP_TRY(hipStreamSynchronize(b->stream));
    double dt = 1e300;
    for (int r = 0; r < kTimed; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      if ((rc = run_steps(b, trial, true))) return abandon_queued(b, rc);
      HIP_TRY(hipStreamSynchronize(b->stream));
      dt = std::min(dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    used += (1 + kTimed) * trial;
    if (dt < best) {
      best = dt;
      bestc = c;
    }
  }
  commit_members(b, nullptr, used);
  b->passes = bestc.passes;
  b->units = bestc.units;
  b->draw_units = bestc.draw_units;
  b->tile = bestc.tile;
  b->grid = bestc.grid;
  b->waves = bestc.waves;
  b->chunk = chunk;
  *used_steps = used;
  return LG_OK;
}

// Fold every member that has a field attached into it (src/render.rs:61-120),
// on the members' own streams, and wait.
int lg_batch_sample(lg_batch_t* b, float alpha) {
  if (!b) return fail(LG_EINVAL, "null batch");
  b->stage_count = 0;  // unread snapshots are discarded
  if (int rc = members_alive(b)) return rc;
  for (lg_t* g : b->members) {
    int rc = lg_sample_field_async(g, alpha);
    if (rc) return rc;
  }
  for (lg_t* g : b->members) HIP_TRY(hipStreamSynchronize(g->stream));
  return LG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// lg_batch_snapshot without any allocation, for callers that cannot allocate
// device memory once set up. The staging buffer is this file-scope device
// array; each GPU holds its own copy from the first use until the process
// ends. In host code `scratch` names a host-side stand-in, never the device
// copy: kernels and copies get the address hipGetSymbolAddress returns for
// the current device, fetched once per device and kept in g_scratch_ptr.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (from several threads, or with another batch's call of this entry
// point) race on it. The snapshot kernel writes every float that is copied
// out, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
#define LG_NOALLOC_SCRATCH_FLOATS (2u << 20)  // 8 MiB
static __device__ __align__(256) float scratch[LG_NOALLOC_SCRATCH_FLOATS];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_ptr[LG_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev`, looked up once per device
// (with that device current) and cached.
int scratch_on(int dev, float** ptr) {
  if (dev < 0 || dev >= LG_MAX_DEVICES) return fail(LG_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(LG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return LG_OK;
}

}  // namespace

extern "C" {

size_t lg_batch_snapshot_noalloc_max(void) { return LG_NOALLOC_SCRATCH_FLOATS; }

int lg_batch_snapshot_noalloc(lg_batch_t* b, float alpha, float* out, size_t nfloats) {
  if (!b || !out) return fail(LG_EINVAL, "null argument");
  b->stage_count = 0;  // unread snapshots are discarded
  if (int rc = members_alive(b)) return rc;
  float* staging = nullptr;
  if (int rc = scratch_on(b->members[0]->device, &staging)) return rc;
  if (int rc = settle_members(b)) return rc;
  if (int rc = queue_desc(b, nullptr, 0)) return rc;
  if (nfloats != b->staging_floats)
    return fail(LG_EINVAL, "output size differs from the members' fields (4 x blocks each)");
  if (nfloats > LG_NOALLOC_SCRATCH_FLOATS)
    return fail(LG_EINVAL, "the members' fields exceed the allocation-free snapshot's " +
                               std::to_string(LG_NOALLOC_SCRATCH_FLOATS) + " floats");
  if (nfloats == 0) return LG_OK;
  return snapshot_into(b, alpha, staging, out, nfloats);
}

}  // extern "C"
