// Internal host-side definitions shared by the engine's translation units:
// error plumbing, device buffers, and the handle behind lg_t.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lgca.h"
#include "lgca_device.h"
#include "lgca_error.h"

struct lg_lattice;

namespace lgca_internal {

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return lgca_internal::fail(LG_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device memory for the library's buffers goes through a per-device cache:
// a released block joins a free list of its size class (powers of two up to
// 1 MiB, then multiples of 2 MiB) and the next request of that class takes
// it back. hipFree waits for the whole device, so handles that come and go
// while other work runs (a serving loop's lattices and batches) would
// otherwise stall it. A block is released only after every stream that used
// it has drained (a handle's or batch's own stream; a lattice a batch holds
// is freed by the batch's destroy), so reuse never overlaps its last use.
// When hipMalloc fails the cache is emptied and the request tried again;
// lg_trim_memory empties it on request.
hipError_t dev_alloc(void** p, size_t bytes);
void dev_release(void* p);
// Page-locked host memory, cached the same way (hipHostFree waits too).
hipError_t host_alloc(void** p, size_t bytes);
void host_release(void* p);
// Free every cached block now; returns the bytes given back.
size_t trim_caches();

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = dev_alloc((void**)&p, count * sizeof(T));
    if (e == hipSuccess)
      n = count;
    else
      p = nullptr;
    return e;
  }
  void release() {
    if (p) dev_release(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

inline uint32_t blocks_for(size_t n, uint32_t t = 256) { return (uint32_t)((n + t - 1) / t); }

// Equilibrium thresholds exactly as src/lattice.rs:59-82 computes them.
void equilibrium_thresholds(float density, float ux, float uy, bool rest, uint32_t out[7]);

// Words per plane from which the default step is the fused one (measured on
// MI355X: at 147,456 words, 4096 x 2304, the two-pass step is faster -- 647 vs
// ~500 GCUPS with a plate and inflow, 686 vs 673 without; from 331,776 words,
// 6144 x 3456, the fused step is, up to 1,134 vs 991 GCUPS at 8.4M words).
#define LG_AUTO_FUSED_WORDS (1u << 18)

// Kernels per step actually used: lg_set_passes's choice, or for 0 (the
// default) the fused step from LG_AUTO_FUSED_WORDS words up -- where one
// launch per step holds enough waves to fill the chip -- and the two-pass
// step below it.
uint32_t step_passes(const lg_lattice* g);

// Words per plane below which the two-pass step's draw work is split finer by
// default: into 32-cell draw lanes, or, with inflow columns, into draws run 8
// steps ahead on the second stream (so that an inflow word's long serial
// draws overlap other steps). Measured on MI355X with the reference's own
// benchmark lattices (density 0.22, drift 0.4): 256 x 256, 1,024 words, 5.2
// -> 5.0 us per step; 2048 x 1280, 40,960 words, 7.9 -> 6.8 us, with inflow
// 11.1 -> 8.8 us; 1024 x 1024 with inflow 10.7 -> 7.7 us; at 131,072 words
// (4096 x 2048) 32-cell draw lanes are 17% slower and draw-ahead 10% slower.
#define LG_SMALL_WORDS (1u << 16)

// Largest temporal depth of the collide pass (steps per collide launch): a
// tile of at most 32 rows keeps 2 x depth halo rows and at least 16 valid.
#define LG_MAX_TDEPTH 8u

// Settle the draw unit and draw-ahead that follow the size rule (those the
// caller has not set): see LG_SMALL_WORDS.
void resolve_pipeline(lg_lattice* g);

// Resident bands: rows per band at most (one wave per row), and the polls a
// neighbour wait makes before it gives up (each sleeps ~64 clocks: ~0.1 s).
#define LG_BAND_MAX_ROWS 11u
// Most passes of the table wave over its lanes for a band's inflow cells
// (INFLOW 0): 64 / cpr rows a pass, cpr the inflow width rounded up to a
// power of two
#define LG_BAND_ICHUNKS 3u
// The size rule's resident bands, from measured step costs (wall-clock per
// step in calls of 10 and 100 steps on MI355X): a step of B-row bands costs
// about 4.0 + 0.47 B us whatever the row width (at most: 4-row to 9-row
// bands measured 5.6 to 8.3 us), the two-pass step about 4.8 us + 0.055 us
// per 1000 words, 1.3 us more with inflow; besides its steps a band call
// costs about 23 us, a two-pass call about 15 (2048 x 1280, its stream
// tables already built), so a call takes resident bands where its first
// chunk's steps save more than LG_BAND_LAUNCH_US (band_gain_us; the model
// puts the gain at 2048 x 1280 near 0.7 us a step, measured 0.9-1.6, and
// the calls cross over at 5-6 steps). (Measured per step, resident against
// two-pass: 2048 x 1280 5.6 against 6.5, with inflow 5.8 against 7.9;
// 1024 x 1024 with inflow 5.9 against 7.2; 256 x 256 4.5 against 5.5;
// 4096 x 2304 with inflow 8.3 against 14.0; 1024 x 2304 with inflow, 9-row
// bands of 16 words, 8.2 against 8.2; 512 x 1280 5.6 against 5.4.)
#define LG_BAND_LAUNCH_US 4.0
// Step time saved by resident bands of `rows` rows (us; negative: lost).
double band_gain_us(const lg_lattice* g, uint32_t rows);
#define LG_BAND_SPIN_LIMIT (1u << 20)
// Steps per resident launch at most (its step counters run below 2^21).
#define LG_BAND_MAX_CHUNK (1u << 16)
// Devices the process-wide band launch order keeps track of.
#define LG_MAX_DEVICES 64

// Arguments of one resident-band launch (lgca_band.hip).
struct LgBandArgs {
  LgGeom G;
  const uint64_t* src;
  uint64_t* dst;
  const uint64_t* solid;
  const uint64_t* cjump;
  const uint32_t* row_band;
  const uint64_t* ijump;
  const uint64_t* icell;  // per-cell inflow jumps (inflow widths <= 64)
  uint64_t* halo;        // [2 slots][2 kinds: top, bottom][nb][2 planes][64]
  uint64_t* flags;       // [2 kinds][nb]: steps published, from epoch
  uint64_t* err;         // epoch of a launch whose wait ran out (device)
  uint64_t* err_host;    // the same, in page-locked host memory
  uint64_t epoch;
  uint64_t seed, t0;  // the run's seed, the launch's first step
  uint32_t n, rows, nb;
  uint32_t nstreams;  // stream tables per band and step in LDS (band streams + inflow)
  uint32_t split;     // row 0's draws by two helper waves (bands of 4m + 1 rows)
  // window tables of T^(4g) (lgca::xs_pow4_windows), copied into LDS at the
  // launch's start: the table wave reaches each window group's first state
  // with 16 lookups
  const uint64_t* pw;
  // inflow drawn by the table wave: lanes per band row (a power of two, at
  // least the inflow width)
  uint32_t icpr;
  // lead jumps (lgca::build_lead_jumps) and whether every half of the
  // lattice has at most one run of cells without a draw (then the rows'
  // draws take the plain chain from the lead jumps)
  const uint64_t* lead;
  uint32_t lead_ok;
  // Coarse-graining of the launch's final state with weight alpha (the
  // kernel instance that does it); the field's layout in device memory at
  // `fs`, read only after the step loop.
  float alpha;
  const struct LgBandField* fs;
};

// The attached field as the resident-band launch's fused coarse-graining
// sees it: bx x by blocks, bw x bh of them, nbr block rows at most per band.
struct LgBandField {
  uint32_t bx, by, bw, bh, nbr;
  int32_t* part;    // [nb][nbr][bw][4]: each band's block sums over its rows
  uint32_t* bcnt;   // [bh]: bands arrived per block row (0 between launches)
  float *ux, *uy, *rho, *sol;
  int64_t* sums;
};

// Free a handle and everything it owns (after its streams drain).
void destroy_lattice(lg_lattice* g);

// Rebuild the static jump and order tables of a lattice if its obstacle
// layout or inflow width changed since they were last built.
int ensure_tables(lg_lattice* g);
// The solid plane on the host (waits for the handle's stream).
int download_solid(lg_lattice* g, std::vector<uint64_t>& s);

// Allocate the two digit/inflow slots of the two-pass step if needed.
int ensure_draw_buffers(lg_lattice* g);

// Resident bands (lgca_band.hip): whether the lattice can run them (rows of
// at most 64 words, at most LG_BAND_MAX_ROWS rows per band with one band per
// CU) and with which band height and count; one launch of n steps from the
// tables at tabs, planes a -> b; and, once the stream has drained, whether a
// launch since the last check ran out of a wait.
bool band_eligible(const lg_lattice* g, uint32_t* rows, uint32_t* nb);
bool resident_active(const lg_lattice* g, uint32_t* rows, uint32_t* nb);
// Whether a band launch can take the field's coarse-graining into its end
// (the block sums of one band fit the LDS it no longer needs by then).
bool band_sample_fits(const lg_lattice* g);
// n steps; with `sample`, the coarse-graining of the final state with
// weight alpha (what sample_now would do next) is done by the same launch.
// Returns kBandRefused (and nothing ran) when the runtime declined the
// cooperative launch; the lattice then runs the other pipelines.
constexpr int kBandRefused = 1;
int launch_band(lg_lattice* g, uint32_t n, bool sample = false, float alpha = 0.0f);
int check_band(lg_lattice* g);
// A handle's stream is going away (its work has ended): band launches of
// other handles no longer order themselves after it.
void band_forget_stream(const lg_lattice* g);

}  // namespace lgca_internal

// The handle behind lg_t: owns every device allocation of one lattice.
struct lg_lattice {
  template <typename T>
  using DevBuf = lgca_internal::DevBuf<T>;

  int device = 0;
  hipStream_t stream = nullptr;
  // Timed advance: one (start, end) event pair around each chunk's step
  // launches, so table and sampling kernels stay out of the step time.
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  LgGeom G{};
  uint64_t seed = 0, steps = 0;
  uint32_t bands = 1;
  bool loaded = false;
  bool tables_dirty = true;  // cjump/ijump/row_band/order out of date
  // planes: a (current), b (next); solid; static tables
  DevBuf<uint64_t> a, b, solid, cjump, ijump, basis;
  DevBuf<uint64_t> icell;  // per-cell inflow jumps, inflow widths <= 64 (resident bands)
  // resident bands: per-word lead jumps, and whether every half of the
  // lattice has at most one run of cells without a draw (lgca::build_lead_jumps),
  // built at the first band launch after the tables (tables_version) change
  DevBuf<uint64_t> lead;
  bool lead_ok = false;
  uint64_t tables_version = 0, lead_version = ~0ull;
  DevBuf<uint32_t> row_band, order;
  uint32_t units = 1;  // 1: a lane steps a 64-cell word; 2: a 32-cell half
  bool auto_units = true;  // calls of one or two steps may take 32-cell lanes
  // passes per step: 1 = one fused kernel per step; 2 = one launch per step
  // that collides step t and draws step t + 1's random digits into the other
  // slot of kbuf/ibuf (kslot: the slot holding the current step's digits);
  // 0 = chosen by size (step_passes)
  uint32_t passes = 0;
  uint32_t draw_units = 1;  // draw lanes of the two-pass step: 1 = 64 cells, 2 = 32
  // draw_units and ahead follow the size rule until the caller sets them
  bool auto_draw_units = true, auto_ahead = true;
  // Batches that hold this handle. lg_destroy of a held handle only marks it
  // destroyed; the last lg_batch_destroy frees it, and the batches' calls
  // fail until then.
  uint32_t batch_refs = 0;
  bool destroyed = false;
  // A batch call failed after queueing steps for this handle: its host-side
  // step counter and plane pointers are not trusted until lg_load/lg_restart.
  bool poisoned = false;
  // End of the async batch advance that last stepped this handle, until a
  // call of its own has waited for it (cleared by that batch's destroy).
  hipEvent_t after = nullptr;
  uint32_t kslot = 0;
  // Draw-ahead (ahead > 0): the draw work runs on `dstream` in launches of up
  // to `ahead` steps into two groups of slots; `stream` runs one collide
  // launch per step. ev_drawn[q] / ev_used[q]: group q drawn / collided.
  uint32_t ahead = 0;
  // Temporal blocking of the collide pass (draw-ahead only): tdepth steps per
  // collide launch (1: one launch per step), tiles of trows valid rows (0:
  // chosen by lattice height and CU count). Needs rows of at most 64 words.
  uint32_t tdepth = 1, trows = 0;
  int ncu = 256;
  // Resident bands (lgca_band.hip): 0 = off, 1 = on where eligible, 2 (the
  // default) = where eligible unless another pipeline was chosen.
  uint32_t resident = 2;
  DevBuf<uint64_t> band_halo, band_flags, band_err_dev;
  DevBuf<uint64_t> band_pw;  // window tables of T^(4g) (lgca::xs_pow4_windows)
  uint64_t* band_err_host = nullptr;  // page-locked
  uint32_t band_nb = 0;
  uint64_t band_epoch = 0, band_epoch_checked = 0;
  bool band_check = false;
  bool band_lds_raised = false;
  bool band_refused = false;  // the runtime declined a cooperative launch
  DevBuf<int32_t> band_part;   // per-band block sums of the fused coarse-graining
  DevBuf<uint32_t> band_bcnt;  // bands arrived per block row
  DevBuf<lgca_internal::LgBandField> band_field;  // what the launch reads of the field
  lgca_internal::LgBandField band_field_host{};   // ... as last copied there
  hipStream_t dstream = nullptr;
  hipEvent_t ev_drawn[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};
  hipEvent_t ev_start = nullptr;
  uint64_t dgroup = 0;
  DevBuf<uint32_t> kbuf, ibuf;
  DevBuf<uint8_t> bytes;
  uint32_t basis_steps = 0;  // capacity of `basis` in steps
  // `basis` holds the stream tables of steps [tab_t0, tab_t0 + tab_n): calls
  // whose steps fall inside launch no tables kernel (stream_tables); a batch
  // that writes its members' tables, and a restart, clear tab_n
  uint64_t tab_t0 = 0;
  uint32_t tab_n = 0;
  // field
  bool has_field = false;
  uint32_t bx = 0, by = 0, bw = 0, bh = 0;
  DevBuf<float> fux, fuy, frho, fsol;
  DevBuf<int64_t> fsums;
  DevBuf<unsigned long long> counter;
};

namespace lgca_internal {

// A handle that an async batch advance has stepped: its planes belong to the
// batch's stream until that advance ends. Calls of the handle's own that
// touch the device wait for it here (on the host) before they queue work.
inline int settle(lg_lattice* g) {
  if (g->poisoned)
    return fail(LG_ESTATE, "the lattice's state is unknown after a failed batch call; reload it");
  if (!g->after) return 0;
  hipError_t e = hipEventSynchronize(g->after);
  if (e != hipSuccess) return fail(LG_EDEVICE, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
  g->after = nullptr;
  return 0;
}

}  // namespace lgca_internal
