/* lgca.h -- C ABI of the MI355X FHP lattice-gas engine.
 *
 * A drop-in for the reference's GPU backend: every entry point maps to one
 * method of `GpuLattice` (src/gpu.rs:453-755), which `Sim` (src/sim.rs:19-175)
 * drives, so a Rust `Backend::Gpu` can call these through `extern "C"` instead
 * of Metal. Unlike the reference's Metal kernels, this engine reproduces the
 * reference CPU path (src/lattice.rs:228-285) bit for bit, including its
 * random streams, which depend on the CPU thread count: that count is the
 * `bands` argument of lg_create.
 *
 * Conventions
 *   - Every call is synchronous (returns after the device work finished), as
 *     the reference's methods are (src/gpu.rs:705,720,752), except the
 *     *_async calls, which return once their work is queued (lg_poll and
 *     lg_wait, lg_batch_poll and lg_batch_wait follow it).
 *   - Functions returning int give 0 on success and a negative code on error;
 *     the message is then available from lg_last_error() (thread-local). The
 *     library never aborts on bad input; where the reference asserts
 *     (src/gpu.rs:480-481,573,631,727-731) it returns LG_EINVAL.
 *   - A handle owns all its device memory and its own HIP stream (a second one
 *     for draw-ahead, lg_set_draw_ahead); use one handle from one thread at a
 *     time. Distinct handles may run concurrently.
 *   - Host arrays are borrowed for the duration of the call only.
 *   - Cells are one byte per site in the reference's layout (src/hex.rs:7-23):
 *     bits 0-5 moving directions, bit 6 rest, bit 7 solid; row-major, y * w + x.
 */
#ifndef LGCA_H
#define LGCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LG_OK 0
#define LG_EINVAL (-1)   /* bad argument (the reference would assert) */
#define LG_EDEVICE (-2)  /* HIP error, or no device */
#define LG_ESTATE (-3)   /* call out of order, e.g. read_field before attach_field */
#define LG_EIO (-4)      /* a frame file could not be written */

typedef struct lg_lattice lg_t;

/* GpuLattice::new (src/gpu.rs:479-515). h must be even and w, h > 4
 * (src/lattice.rs:137-138). `bands` is the CPU thread count the run is to be
 * bit-identical with (src/lattice.rs:232-242); it is new here because the
 * reference's GPU path did not reproduce the CPU streams. Returns NULL on
 * error (see lg_last_error). */
lg_t* lg_create(size_t w, size_t h, int rest, uint64_t seed, size_t bands);

/* Drop for GpuLattice. NULL is accepted. A handle that a batch still holds
 * is freed when the last such batch is destroyed (see lg_batch_create). */
void lg_destroy(lg_t* g);

/* Message of the last failed call on this thread ("" if none). */
const char* lg_last_error(void);

/* New: give back to the runtime the device and page-locked memory the
 * library keeps for reuse. Buffers of destroyed lattices and batches are
 * kept, by size class, and handed to later lattices and batches, because
 * hipFree waits for the whole device (a serving loop that frees while other
 * runs step would stall). The cache is also emptied when an allocation
 * fails, before it is tried again. *freed (may be NULL) receives the bytes
 * given back. Call it with no other call of the library running. */
int lg_trim_memory(size_t* freed);

/* GpuLattice::device_name (src/gpu.rs:517-519). */
int lg_device_name(lg_t* g, char* buf, size_t len);

/* GpuLattice::set_inlet (src/gpu.rs:525-528) given the seven 24-bit
 * thresholds of Equilibrium::thresholds (src/lattice.rs:87-92), rest last. */
int lg_set_inlet(lg_t* g, size_t cols, const uint32_t thresh[7]);

/* GpuLattice::set_inlet with the equilibrium computed here exactly as
 * Equilibrium::new (src/lattice.rs:59-82) computes it, in f32. */
int lg_set_inlet_flow(lg_t* g, size_t cols, float density, float ux, float uy);

/* GpuLattice::attach_field (src/gpu.rs:532-545): a device-resident coarse
 * field of (w / bx) x (h / by) blocks, zero-initialised. */
int lg_attach_field(lg_t* g, size_t bx, size_t by);

/* GpuLattice::load (src/gpu.rs:572-610). n must be w * h. */
int lg_load(lg_t* g, const uint8_t* cells, size_t n);

/* GpuLattice::restart (src/gpu.rs:616-620): load, set the seed, steps = 0. */
int lg_restart(lg_t* g, const uint8_t* cells, size_t n, uint64_t seed);

/* GpuLattice::store (src/gpu.rs:630-652). n must be w * h. */
int lg_store(lg_t* g, uint8_t* cells, size_t n);

/* GpuLattice::advance (src/gpu.rs:657-659). */
int lg_advance(lg_t* g, uint64_t n);

/* GpuLattice::advance_sampling (src/gpu.rs:668-706): fold the lattice into
 * the attached field every `every` steps and after a ragged last chunk;
 * every == 0 skips sampling. */
int lg_advance_sampling(lg_t* g, uint64_t n, uint64_t every, float alpha);

/* New: lg_advance and lg_advance_sampling without the wait, for serving loops
 * that overlap host work with the device (the reference's calls always wait,
 * src/gpu.rs:657-706). They return once the steps are queued on the handle's
 * stream; lg_steps already counts them. Every later call on the handle queues
 * behind them, and the calls that return data (lg_store, lg_read_field, ...)
 * wait for them; lg_poll reports whether the handle's queued work has ended
 * (1) or not (0), lg_wait waits for it. */
int lg_advance_async(lg_t* g, uint64_t n);
int lg_advance_sampling_async(lg_t* g, uint64_t n, uint64_t every, float alpha);
int lg_poll(lg_t* g);
int lg_wait(lg_t* g);

/* GpuLattice::sample_field (src/gpu.rs:710-721). No-op without a field. */
int lg_sample_field(lg_t* g, float alpha);

/* New: the block grid of the attached field (Field's bw, bh, bx, by,
 * src/render.rs:21-34), so that callers can size the arrays of lg_read_field. */
int lg_field_dims(lg_t* g, size_t* bw, size_t* bh, size_t* bx, size_t* by);

/* GpuLattice::read_field (src/gpu.rs:725-736): copy the four f32 arrays out.
 * nblocks must be (w / bx) * (h / by). */
int lg_read_field(lg_t* g, float* ux, float* uy, float* rho, float* solid, size_t nblocks);

/* New: the exact integer block sums behind the last sample (mass, px2, py2 in
 * the units of src/hex.rs:29-30, and solid count), the quantities
 * Field::sample derives its f32 values from (src/render.rs:93-116). */
int lg_read_block_sums(lg_t* g, int64_t* mass, int64_t* px2, int64_t* py2, int64_t* nsolid,
                       size_t nblocks);

/* GpuLattice::total_particles (src/gpu.rs:740-754), as a 64-bit count. */
int lg_total_particles(lg_t* g, uint64_t* out);

/* Public fields w, h, steps of GpuLattice (src/gpu.rs:467-471). */
int lg_dims(lg_t* g, size_t* w, size_t* h);
uint64_t lg_steps(lg_t* g);

/* New: set the step counter, so that cells + seed + steps + bands resume a
 * run exactly (the reference can only restart at step 0). */
int lg_set_steps(lg_t* g, uint64_t steps);

/* Launch shape of the step: each lane steps 64 cells (a whole u64 word, the
 * default) or 32 (half a word, twice the lanes). The results are identical
 * either way. Until set, calls of one or two steps without inflow columns on
 * lattices below 2^18 words per plane take the fused step with 32-cell lanes
 * (the shortest call), unless lg_set_passes or lg_set_resident(g, 1) chose
 * the pipeline. */
int lg_set_work_unit(lg_t* g, uint32_t cells_per_lane);

/* Work per step launch: 2 splits the step into a draw pass (the random
 * collision columns and inflow samples, which depend only on the seed, the
 * step and the obstacle layout) and a gather-collide-store pass; each launch
 * collides step t while it draws step t + 1, so memory waits of the one
 * overlap the arithmetic of the other. 1 fuses both passes of one step into
 * each lane. 0 (the default) picks by size: the two-pass step for lattices
 * whose one-launch step holds too few waves to keep every CU busy to its end,
 * the fused step for larger ones. The results are identical either way. */
int lg_set_passes(lg_t* g, uint32_t passes);
/* (lg_set_passes(g, 0) also returns the draw unit and the draw-ahead below to
 * their size rule.) */

/* The pipeline the steps of a long call will use when it does not run
 * resident bands (lg_get_resident): kernels per step (1 or 2), cells per lane
 * of the collide (or fused) and draw work, draw-ahead steps. */
int lg_get_pipeline(lg_t* g, uint32_t* passes, uint32_t* cells_per_lane,
                    uint32_t* draw_cells_per_lane, uint32_t* draw_ahead);

/* Draw-ahead for the two-pass step: 0 runs each step's draw work in the
 * launch that collides the step before it; N > 0 (at most 64) runs the draws
 * of up to N steps per launch on a second HIP stream, double-buffered ahead of
 * the one-collide-per-step launches on the handle's stream, so the draws fill
 * the gaps of the serial collide chain. Results are identical either way.
 * Until set, it follows the size rule: 8 for lattices below 2^16 words per
 * plane that have inflow columns (their inflow words' serial draws then
 * overlap other steps), else 0. */
int lg_set_draw_ahead(lg_t* g, uint32_t steps);

/* Cells per lane (64 or 32) of the two-pass step's draw work, as
 * lg_set_work_unit sets them for its collide work; 0, the default, follows the
 * size rule: 32 for lattices below 2^16 words per plane without inflow
 * columns, else 64. Results are identical. */
int lg_set_draw_unit(lg_t* g, uint32_t cells_per_lane);

/* New. Temporal blocking of the collide pass: each collide launch steps a
 * band of rows `steps` steps (1..8) in registers, its rows' neighbours
 * exchanged across lanes and through LDS, and stores once; the band's
 * 2 x steps edge rows repeat the collision, never the draws (the digits come
 * from the draw pass, which runs `steps` or more steps ahead on the second
 * stream: lg_set_draw_ahead is raised to `steps` if lower). rows_per_tile =
 * valid rows per band (0: about one band per CU; at most 32 - 2 x steps).
 * Rows of at most 64 words (4096 cells) only; steps = 1 turns it off.
 * Results are identical for every choice. */
int lg_set_temporal_depth(lg_t* g, uint32_t steps, uint32_t rows_per_tile);

/* New. Resident bands (on = 1; 2, the default, = on unless lg_set_passes,
 * lg_set_draw_ahead or lg_set_temporal_depth chose another pipeline; 0 =
 * off): each call's steps run in one launch per chunk (a chunk ends at a sample or at 65536 steps; a chunk that
 * ends at a sample takes it in the same launch where a band's block sums fit
 * its LDS), in which one
 * workgroup per band of rows keeps its rows in registers for the whole
 * chunk, draws its own random digits one step ahead and exchanges only its
 * edge rows with the neighbouring bands through global memory and per-band
 * step counters, and a further wave builds each step's stream tables in LDS.
 * For rows of at most 64 words (4096 cells) and at most 11 rows per band at
 * one band per CU; other lattices keep the other pipelines. The size rule
 * (2) takes it where, by measured step costs, bands of B rows step faster
 * than the two-pass step, for calls whose first chunk (the call, or up to
 * its first sample) saves more step time than a band call costs over a
 * two-pass call (about 4 us). Results are identical. lg_get_resident returns 1 (and the band
 * height and count) when resident bands are in use for calls of enough
 * steps, else 0. The process's first band launch on a device is
 * cooperative; if the runtime declines it (it does while another process
 * that has made one is alive), the lattice runs the other pipelines from
 * then on (until lg_set_resident is called again). Later band launches are
 * plain and the library runs them one after another on the device (a launch
 * from another stream waits for the previous one's stream). A wait
 * between bands that runs out (it cannot hang the device: every wait is
 * bounded) makes the call fail with LG_EDEVICE and the lattice must be
 * reloaded. */
int lg_set_resident(lg_t* g, uint32_t on);
int lg_get_resident(lg_t* g, uint32_t* rows_per_band, uint32_t* bands);
/* New: the temporal depth and tile rows set by lg_set_temporal_depth. */
int lg_get_temporal_depth(lg_t* g, uint32_t* steps, uint32_t* rows_per_tile);
/* New: time the step pipelines this lattice can run -- the size rule,
 * resident bands on and off, the two-pass step in one launch, draw-ahead by
 * 16 steps, the fused step, and the temporally blocked collide at depths 2,
 * 4 and 8 (rows of at most 64 words) -- on the lattice itself, and keep the
 * fastest. Each candidate runs one warm call of up to 32 steps and is timed
 * as the best of three more; the steps are real steps of the lattice
 * (results are identical in every pipeline), reported in *used_steps;
 * budget_steps = 0: no limit. The choice is read back with lg_get_pipeline,
 * lg_get_resident and lg_get_temporal_depth. */
int lg_autotune(lg_t* g, uint64_t budget_steps, uint64_t* used_steps);

/* New: advance n steps and report the device time of the step kernels in ms:
 * the sum over HIP event pairs recorded on the handle's stream around each
 * run of back-to-back step launches (the random-table and sampling kernels
 * between those runs are not counted). */
int lg_advance_timed(lg_t* g, uint64_t n, float* ms);

/* New: lg_advance_sampling with the same device-time report. */
int lg_advance_sampling_timed(lg_t* g, uint64_t n, uint64_t every, float alpha, float* ms);

/* Lattice::add_solid (src/lattice.rs:210-220) on the device, for the two
 * obstacle shapes the reference builds: the plate box
 * x0 <= px <= x1 && y0 <= py <= y1 (src/main.rs:584-588, tests/golden.rs:78-83)
 * and the disc (px-cx)^2 + (py-cy)^2 <= r^2 (src/main.rs:589-592), over the
 * f32 site positions of src/hex.rs:52-55. Marked sites lose their particles. */
int lg_add_solid_box(lg_t* g, float x0, float x1, float y0, float y1);
int lg_add_solid_disc(lg_t* g, float cx, float cy, float r);

/* Lattice::init_equilibrium (src/lattice.rs:197-208) on the device: every
 * fluid site drawn from Equilibrium(density, ux, uy) in index order from the
 * stream seeded with seed ^ 0xA5A51234, bit-identical with the CPU. */
int lg_init_equilibrium(lg_t* g, float density, float ux, float uy);

/* Lattice::init_equilibrium with a profile: site (x, y) is drawn from
 * Equilibrium(density, ux[i], uy[i]) with i = x (axis 0, n = w) or i = y
 * (axis 1, n = h), in the same stream and order as lg_init_equilibrium; the
 * column form is how src/transport.rs:238-251 builds its start states. */
int lg_init_profile(lg_t* g, int axis, float density, const float* ux, const float* uy, size_t n);

/* transport::measure (src/transport.rs:183-319) on the engine: the viscosity
 * from a decaying shear wave and the advection factor from a drifting wave in
 * size x size periodic boxes, with the reference's protocol (transient, adaptive
 * window, 4 realisations, 20 samples). `threads` is the CPU thread count whose
 * random streams the runs reproduce; with the same value the results equal the
 * CPU path's bit for bit. lg_measure_transport runs both concurrently. */
int lg_measure_viscosity(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* nu);
int lg_measure_advection(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* g);
int lg_measure_transport(float density, int rest, size_t threads, uint64_t seed, size_t size,
                         float* nu, float* g);

/* Batches (no counterpart in the reference, which never batched lattices,
 * docs/plan.md:482-485). A batch is a fixed set of distinct handles that are
 * advanced together: one step kernel per step covers every member through a
 * tile table, and the step loop is captured in a hipGraph and replayed. Each
 * member ends exactly where its own lg_advance would have left it. While a
 * batch exists its members may still be used on their own between batch calls
 * (lg_batch_advance re-reads their state each call). lg_destroy of a member
 * that a batch holds is deferred: the handle is freed by the last
 * lg_batch_destroy of the batches holding it, and until then those batches'
 * advance, snapshot, sample and autotune calls fail with LG_EINVAL. Creating,
 * destroying and calling batches that share a member must not overlap in
 * time (one thread, or the caller's own lock). Members must have fewer than
 * 2^30 words per plane (68.7G cells; one such lattice fills the device). */
typedef struct lg_batch lg_batch_t;
lg_batch_t* lg_batch_create(lg_t* const* members, size_t n);
void lg_batch_destroy(lg_batch_t* b);
int lg_batch_advance(lg_batch_t* b, uint64_t n);
/* Advance member i by min(n, limits[i]) steps in the same launches (members
 * with fewer steps left stop early); limits has one entry per member. */
int lg_batch_advance_limited(lg_batch_t* b, uint64_t n, const uint64_t* limits);
/* lg_batch_advance_limited without the wait (limits may be NULL: n for
 * all). It queues the steps behind the work queued earlier for the members
 * (their own calls, async advances of this or other batches) -- the batch's
 * queue waits for that on the device, the host does not -- and returns; the
 * members' step counters already count them. A serving loop can so queue a
 * quantum on a batch of another membership before the last one has ended. A
 * member's own calls that touch the device wait for them first (on the
 * host), and other batches' calls on the device; lg_poll of a member
 * includes them. lg_batch_poll reports whether the batch's queued work has
 * ended (1) or not (0), lg_batch_wait waits for it. */
int lg_batch_advance_async(lg_batch_t* b, uint64_t n, const uint64_t* limits);
int lg_batch_poll(lg_batch_t* b);
int lg_batch_wait(lg_batch_t* b);
/* lg_batch_advance_async followed, in the same queue, by lg_batch_snapshot's
 * sample of every member's field (n = 0: the sample alone), copied on to
 * page-locked memory in the same queue. Up to two such snapshots wait to be
 * read (a third sampling advance fails with LG_ESTATE): lg_batch_read_snapshot
 * waits for the oldest one and copies it out in lg_batch_snapshot's layout
 * (nfloats as there), so a loop can queue the next quantum before it reads
 * the last one. A synchronous advance, snapshot, sample or autotune call of
 * the batch discards unread snapshots; lg_batch_advance_async, lg_batch_poll
 * and lg_batch_wait keep them. */
int lg_batch_advance_sampling_async(lg_batch_t* b, uint64_t n, const uint64_t* limits,
                                    float alpha);
int lg_batch_read_snapshot(lg_batch_t* b, float* out, size_t nfloats);
/* Sample every member's attached field (lg_sample_field with alpha) in one
 * launch and copy all of them out: for each member with a field, in member
 * order, its ux, uy, rho, solid arrays back to back. nfloats must equal
 * lg_batch_snapshot_size(b). */
int lg_batch_snapshot(lg_batch_t* b, float alpha, float* out, size_t nfloats);
size_t lg_batch_snapshot_size(lg_batch_t* b);
/* No counterpart in the reference. lg_batch_snapshot for callers that cannot
 * allocate device memory once set up (serving loops and harnesses that
 * forbid hipMalloc): the same sample and output, staged through a static
 * device array of lg_batch_snapshot_noalloc_max() floats (8 MiB, held per
 * GPU until the process ends) instead of a buffer the batch allocates.
 * nfloats above that maximum is LG_EINVAL. Calls on one GPU share the array:
 * calls that overlap on one GPU (several threads, several batches) race. */
int lg_batch_snapshot_noalloc(lg_batch_t* b, float alpha, float* out, size_t nfloats);
size_t lg_batch_snapshot_noalloc_max(void);
/* Launch shape: tile = words per workgroup (128 or 256; by default 128 for
 * the fused step, 256 for the two-pass step), grid = step-kernel workgroups
 * (0, the default: one per tile; otherwise a persistent grid walking the
 * tiles), chunk = steps per graph replay (even, at most 256). */
int lg_batch_set_shape(lg_batch_t* b, uint32_t tile, uint32_t grid, uint32_t chunk);
int lg_batch_get_shape(lg_batch_t* b, uint32_t* tile, uint32_t* grid, uint32_t* chunk);
/* New: register target of the batch's step kernel, in waves per SIMD (0: the
 * pipeline's default, 3 for the fused step and 5 for the two-pass step).
 * Whole-word lanes also run at 2 or 4 (fused) and 4 or 6 (two-pass); other
 * values (2..6) fall back to the default, and lg_batch_get_waves reports the
 * target actually used. Results are identical for every choice. */
int lg_batch_set_waves(lg_batch_t* b, uint32_t waves_per_simd);
int lg_batch_get_waves(lg_batch_t* b, uint32_t* waves_per_simd);
/* Cells per lane (64 or 32) for the batch's step kernel, as lg_set_work_unit. */
int lg_batch_set_units(lg_batch_t* b, uint32_t cells_per_lane);
int lg_batch_get_units(lg_batch_t* b, uint32_t* cells_per_lane);
/* Kernels per step (2: collide while drawing the next step; 1: fused; the
 * default follows lg_set_passes's size rule applied to all members' words
 * together) and the draw work's cells per lane, as lg_set_passes /
 * lg_set_draw_unit for one lattice. */
int lg_batch_set_passes(lg_batch_t* b, uint32_t passes);
int lg_batch_set_draw_unit(lg_batch_t* b, uint32_t cells_per_lane);
int lg_batch_get_pipeline(lg_batch_t* b, uint32_t* passes, uint32_t* cells_per_lane,
                          uint32_t* draw_cells_per_lane);
/* lg_sample_field for every member with a field, issued together. */
int lg_batch_sample(lg_batch_t* b, float alpha);
/* lg_sample_field without the wait (the sample is queued on the handle's
 * stream; the next synchronous call on the handle sees it). */
int lg_sample_field_async(lg_t* g, float alpha);
/* Time candidate launch pipelines and shapes (kernels per step, cells per
 * lane for the collide and draw work, tile, grid, waves per SIMD) on the
 * batch itself, most promising first, and keep the fastest: each candidate
 * runs one warm graph replay of up to 32 steps and is timed as the best of
 * three more. The timed steps are real steps of every member, reported in
 * *used_steps; budget_steps = 0: no limit. */
int lg_batch_autotune(lg_batch_t* b, uint64_t budget_steps, uint64_t* used_steps);

/* ---- Frame output (host code; no GPU except lg_frames_write_lattice) ------
 *
 * The reference's frame writers, byte-identical with its files for the same
 * field. A field here is bw x bh blocks of bx x by cells, as lg_read_field
 * returns it: row-major f32 arrays ux, uy, solid of bw * bh values. */

/* Field::vorticity (src/render.rs:122-139): centred differences, one-sided
 * at the x ends, periodic in y; out holds bw * bh values. */
int lg_field_vorticity(const float* ux, const float* uy, size_t bw, size_t bh, size_t bx,
                       size_t by, float* out);

/* Field::mean_velocity (src/render.rs:142-157): mean of ux and uy over the
 * blocks whose solid fraction is below 0.5, summed in f32 in block order;
 * (0, 0) when there are none. */
int lg_field_mean_velocity(const float* ux, const float* uy, const float* solid, size_t n,
                           float* mx, float* my);

/* The colour scale of write_frame (src/main.rs:219-229): element n' * 98 /
 * 100 of the ascending |vort| over the n' blocks whose solid fraction is below
 * 0.12, at least 1e-5, and 1e-3 when n' is 0. LG_EINVAL on a NaN, where the
 * reference's sort would panic. */
int lg_vorticity_scale(const float* vort, const float* solid, size_t n, float* wmax);

/* png::deflate (src/png.rs:103-162): one fixed-Huffman block with greedy
 * LZ77. *len receives the compressed size; the bytes are copied to out when
 * out is not NULL (LG_EINVAL if cap is smaller than *len). */
int lg_png_deflate(const uint8_t* data, size_t n, uint8_t* out, size_t cap, size_t* len);

/* png::write_rgb (src/png.rs:174-197): an 8-bit RGB PNG of w x h pixels,
 * rows top to bottom. LG_EIO names the path when the file cannot be written. */
int lg_write_png_rgb(const char* path, size_t w, size_t h, const uint8_t* rgb);

/* render::write_vorticity (src/render.rs:176-204): each block is scale
 * pixels wide and round(scale * sqrt(3)/2 * by / bx) high, coloured by
 * vorticity / wmax on a blue-white-red map, solid blocks dark. */
int lg_write_vorticity(const char* path, const float* ux, const float* uy, const float* solid,
                       size_t bw, size_t bh, size_t bx, size_t by, size_t scale, float wmax);

/* render::write_arrows (src/render.rs:209-267): an SVG arrow per fluid block
 * of the velocity minus (frame_x, frame_y), bx * scale lattice units per unit
 * of speed; arrows shorter than 0.35 are left out. */
int lg_write_arrows(const char* path, const float* ux, const float* uy, const float* solid,
                    size_t bw, size_t bh, size_t bx, size_t by, float scale, float frame_x,
                    float frame_y);

/* Frames (src/main.rs:214-281): writer threads behind a queue that holds at
 * most `workers` frames (1..256), so that a run outpacing its writers waits
 * for them. Each frame writes the vorticity PNG, scaled by the 98th
 * percentile above, then the arrow SVG (write_frame, src/main.rs:220-234). */
typedef struct lg_frames lg_frames_t;
lg_frames_t* lg_frames_create(size_t workers);

/* Frames::write (src/main.rs:262-268): copy the field and queue the frame,
 * waiting while the queue is full. Returns LG_EIO with "<png path>: <error>"
 * of the first frame that failed, if any has; that failure surfaces a frame
 * or two late, as in the reference. */
int lg_frames_write(lg_frames_t* fr, const float* ux, const float* uy, const float* solid,
                    size_t bw, size_t bh, size_t bx, size_t by, const char* png, const char* svg,
                    size_t scale, float arrow_scale, float mean_x, float mean_y);

/* New: the run loop's frame (src/main.rs:676-688) from a lattice's attached
 * field: read it (as lg_read_field), take its mean velocity as the arrows'
 * frame, and queue it as lg_frames_write does. */
int lg_frames_write_lattice(lg_frames_t* fr, lg_t* g, const char* png, const char* svg,
                            size_t scale, float arrow_scale);

/* Frames::finish (src/main.rs:271-280): write everything queued, stop the
 * writers and free the pool; returns the first failure as lg_frames_write. */
int lg_frames_finish(lg_frames_t* fr);

/* Diagnostics. The generated collision circuit as the device evaluates it,
 * for every state and every draw index k < stride, into out[state * stride +
 * k] (0xFF where the lanes of a word disagreed). Returns the stride (30 with
 * rest particles, 6 without) or a negative error. The same exhaustive check
 * as the reference's lgca_verify kernel (src/gpu.rs:399-417, 775-820). */
int lg_debug_collide_table(int rest, uint8_t* out, size_t n);

/* Diagnostics, no counterpart in the reference: the VALU issue-rate probe.
 * Runs op (0..14: lg_probe_op_name gives its instruction; LG_PROBE_STREAM:
 * the draw work's exact random stream, one cell after another) on every CU
 * at waves_per_simd waves per SIMD (1..8), with `chains` independent
 * dependency chains per lane (1 or 8; ignored for the stream), iters loop
 * iterations of 16 instructions per chain (32 cells for the stream). Gives
 * wave-instructions (cells for the stream) issued per CU per shader clock,
 * from the median workgroup's s_memtime interval, and that clock in GHz
 * (s_memtime against the 100 MHz s_memrealtime). */
#define LG_PROBE_STREAM 100
const char* lg_probe_op_name(int op);
int lg_probe_valu(int op, uint32_t waves_per_simd, uint32_t chains, uint32_t iters,
                  double* per_cu_cycle, double* clock_ghz);

#ifdef __cplusplus
}
#endif
#endif /* LGCA_H */
