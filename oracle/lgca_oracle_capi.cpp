// C ABI over the CPU oracle, for the Python tests and the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: what results are compared with, never the thing
// measured or shipped.
// Each entry point forwards to the restatement in lgca_oracle.hpp (the
// lattice), lgca_transport.hpp (transport measurement) or lgca_render.hpp
// (frame output), which cite the reference file:line they follow.
#include "lgca_oracle.hpp"
#include "lgca_render.hpp"
#include "lgca_transport.hpp"

using namespace oracle;

extern "C" {

// Lattice::new (src/lattice.rs:127-152). Returns NULL where the reference
// asserts: h odd, or w or h not above 4 (src/lattice.rs:137-138).
void* orc_new(size_t w, size_t h, float density, float ux, float uy, size_t inlet_cols, int rest,
              size_t threads, uint64_t seed) {
  if (h % 2 != 0 || w <= 4 || h <= 4) return nullptr;
  return new Lattice(w, h, density, ux, uy, inlet_cols, rest != 0, threads, seed);
}
void orc_free(void* p) { delete static_cast<Lattice*>(p); }
void orc_set_parallel(void* p, int on) { static_cast<Lattice*>(p)->parallel = on != 0; }
void orc_add_solid_box(void* p, float x0, float x1, float y0, float y1) {
  static_cast<Lattice*>(p)->add_solid_box(x0, x1, y0, y1);
}
void orc_add_solid_disc(void* p, float cx, float cy, float r) {
  static_cast<Lattice*>(p)->add_solid_disc(cx, cy, r);
}
// The public `inflow` field of Lattice (src/lattice.rs:18), for workloads
// that initialise at one velocity and drive the inflow at another.
void orc_set_inflow(void* p, float ux, float uy) {
  static_cast<Lattice*>(p)->inflow_x = ux;
  static_cast<Lattice*>(p)->inflow_y = uy;
}
void orc_init_equilibrium(void* p) { static_cast<Lattice*>(p)->init_equilibrium(); }
void orc_init_profile(void* p, int axis, const float* ux, const float* uy) {
  static_cast<Lattice*>(p)->init_profile(axis, ux, uy);
}
int orc_restart(void* p, const uint8_t* cells, size_t n, uint64_t seed) {
  auto* L = static_cast<Lattice*>(p);
  if (n != L->cells.size()) return -1;
  L->restart(cells, seed);
  return 0;
}
void orc_advance(void* p, uint64_t n) { static_cast<Lattice*>(p)->advance(n); }
uint64_t orc_steps(void* p) { return static_cast<Lattice*>(p)->steps; }
size_t orc_n_bands(void* p) { return static_cast<Lattice*>(p)->n_bands(); }
int orc_get_cells(void* p, uint8_t* out, size_t n) {
  auto* L = static_cast<Lattice*>(p);
  if (n != L->cells.size()) return -1;
  std::memcpy(out, L->cells.data(), n);
  return 0;
}
int orc_set_cells(void* p, const uint8_t* in, size_t n) {
  auto* L = static_cast<Lattice*>(p);
  if (n != L->cells.size()) return -1;
  std::memcpy(L->cells.data(), in, n);
  return 0;
}
uint64_t orc_total_particles(void* p) { return static_cast<Lattice*>(p)->total_particles(); }
void orc_mean_velocity(void* p, float* ux, float* uy) {
  static_cast<Lattice*>(p)->mean_velocity(ux, uy);
}
size_t orc_active_states(void* p) { return static_cast<Lattice*>(p)->table.active_states(); }

// Field (src/render.rs:35-120)
void* orc_field_new(size_t w, size_t h, size_t bx, size_t by) {
  if (bx == 0 || by == 0) return nullptr;
  return new Field(w, h, bx, by);
}
void orc_field_free(void* f) { delete static_cast<Field*>(f); }
void orc_field_dims(void* f, size_t* bw, size_t* bh) {
  *bw = static_cast<Field*>(f)->bw;
  *bh = static_cast<Field*>(f)->bh;
}
int orc_field_sample(void* f, void* lat, float alpha) {
  return static_cast<Field*>(f)->sample(*static_cast<Lattice*>(lat), alpha) ? 0 : -1;
}
void orc_field_read(void* f, float* ux, float* uy, float* rho, float* solid) {
  auto* F = static_cast<Field*>(f);
  size_t n = F->bw * F->bh;
  std::memcpy(ux, F->ux.data(), n * 4);
  std::memcpy(uy, F->uy.data(), n * 4);
  std::memcpy(rho, F->rho.data(), n * 4);
  std::memcpy(solid, F->solid.data(), n * 4);
}
uint64_t orc_field_checksum(void* f) { return field_checksum(*static_cast<Field*>(f)); }
// Exact per-block integer sums (mass, px2, py2, nsolid), block-row-major.
void orc_block_sums(void* f, void* lat, int64_t* mass, int64_t* px2, int64_t* py2,
                    int64_t* nsolid) {
  auto* F = static_cast<Field*>(f);
  auto* L = static_cast<Lattice*>(lat);
  for (size_t jb = 0; jb < F->bh; ++jb)
    for (size_t ib = 0; ib < F->bw; ++ib) {
      Moments m = F->block_moments(*L, ib, jb);
      size_t k = jb * F->bw + ib;
      mass[k] = m.mass;
      px2[k] = m.px2;
      py2[k] = m.py2;
      nsolid[k] = m.solid;
    }
}

// tests/golden.rs:40-47
uint64_t orc_fnv1a(const uint8_t* bytes, size_t n) { return fnv1a(bytes, n); }

// Equilibrium::thresholds (src/lattice.rs:72-92)
void orc_thresholds(float density, float ux, float uy, int rest, uint32_t* out7) {
  Equilibrium(density, ux, uy, rest != 0).thresholds(out7);
}
// Equilibrium::sample repeated n times from Rng::new(seed) (src/lattice.rs:102-115)
void orc_eq_samples(float density, float ux, float uy, int rest, uint64_t seed, size_t n,
                    uint8_t* out) {
  Equilibrium eq(density, ux, uy, rest != 0);
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) out[i] = eq.sample(rng);
}
// Rng (src/rng.rs:8-29)
uint64_t orc_rng_new(uint64_t seed) { return Rng(seed).s; }
void orc_rng_u64(uint64_t seed, size_t n, uint64_t* out) {
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) out[i] = rng.next_u64();
}
// CollisionTable (src/collision.rs:72-98). Writes n_states*stride bytes when
// `out` is large enough; returns the stride.
size_t orc_table(int rest, uint8_t* out, size_t cap) {
  CollisionTable t(rest != 0);
  if (out && cap >= t.data.size()) std::memcpy(out, t.data.data(), t.data.size());
  return t.stride;
}
// classes (src/collision.rs:62-69): members concatenated into `members`,
// class sizes into `sizes`; returns the number of classes.
size_t orc_classes(int rest, uint8_t* members, uint32_t* sizes) {
  auto cl = classes(rest != 0);
  size_t off = 0;
  for (size_t c = 0; c < cl.size(); ++c) {
    sizes[c] = (uint32_t)cl[c].size();
    for (uint8_t m : cl[c]) members[off++] = m;
  }
  return cl.size();
}
// moments tables (src/moments.rs:70-74)
void orc_moment_tables(uint64_t* fluid, uint64_t* with_walls) {
  std::memcpy(fluid, tables().fluid, sizeof(tables().fluid));
  std::memcpy(with_walls, tables().with_walls, sizeof(tables().with_walls));
}
// moments::sum + Moments::velocity over a byte slice (src/moments.rs:91-138)
void orc_moments_sum(int walls, const uint8_t* cells, size_t n, int64_t* out4, float* v2) {
  Moments m = sum(walls ? tables().with_walls : tables().fluid, cells, n);
  out4[0] = m.mass;
  out4[1] = m.px2;
  out4[2] = m.py2;
  out4[3] = m.solid;
  m.velocity(&v2[0], &v2[1]);
}

// transport::viscosity / advection_factor (src/transport.rs:183-289)
float orc_viscosity(float density, int rest, size_t threads, uint64_t seed, size_t size) {
  return viscosity(density, rest != 0, threads, seed, size);
}
float orc_advection_factor(float density, int rest, size_t threads, uint64_t seed, size_t size) {
  return advection_factor(density, rest != 0, threads, seed, size);
}
// Per-row x velocity and per-column y velocity of a byte lattice
// (src/transport.rs:51-91), for checking the engine's projections.
void orc_row_ux(const uint8_t* cells, size_t w, size_t h, float* out) {
  auto u = row_ux(cells, w, h);
  std::memcpy(out, u.data(), h * 4);
}
void orc_column_uy(const uint8_t* cells, size_t w, size_t h, float* out) {
  auto u = column_uy(cells, w, h);
  std::memcpy(out, u.data(), w * 4);
}

// ---- frame output (lgca_render.hpp) --------------------------------------
// The byte outputs are returned by size: the call copies them to `out` when
// `cap` is large enough and always returns the full size.
static size_t copy_out(const void* src, size_t n, void* out, size_t cap) {
  if (out && cap >= n) std::memcpy(out, src, n);
  return n;
}

// Field::vorticity (src/render.rs:122-139)
void orc_vorticity(const float* ux, const float* uy, size_t bw, size_t bh, size_t bx, size_t by,
                   float* out) {
  auto w = oracle_render::vorticity(bw, bh, bx, by, ux, uy);
  std::memcpy(out, w.data(), w.size() * 4);
}
// write_frame's colour scale (src/main.rs:219-229)
float orc_vorticity_scale(const float* vort, const float* solid, size_t n) {
  return oracle_render::vorticity_scale(std::vector<float>(vort, vort + n), solid);
}
// Field::mean_velocity (src/render.rs:142-157)
void orc_field_mean_velocity(const float* ux, const float* uy, const float* solid, size_t n,
                             float* mx, float* my) {
  oracle_render::mean_velocity(ux, uy, solid, n, mx, my);
}
// png::deflate (src/png.rs:103-162)
size_t orc_deflate(const uint8_t* data, size_t n, uint8_t* out, size_t cap) {
  auto z = oracle_render::deflate(data, n);
  return copy_out(z.data(), z.size(), out, cap);
}
// png::crc32, png::adler32 (src/png.rs:4-22)
uint32_t orc_crc32(const uint8_t* d, size_t n) { return oracle_render::crc32(d, n); }
uint32_t orc_adler32(const uint8_t* d, size_t n) { return oracle_render::adler32(d, n); }
// png::write_rgb (src/png.rs:174-197), the file's bytes
size_t orc_png_rgb(size_t w, size_t h, const uint8_t* rgb, uint8_t* out, size_t cap) {
  auto p = oracle_render::png_rgb(w, h, rgb);
  return copy_out(p.data(), p.size(), out, cap);
}
// render::write_vorticity (src/render.rs:176-204), the file's bytes
size_t orc_vorticity_png(const float* ux, const float* uy, const float* solid, size_t bw,
                         size_t bh, size_t bx, size_t by, size_t scale, float wmax, uint8_t* out,
                         size_t cap) {
  auto p = oracle_render::vorticity_png(bw, bh, bx, by, ux, uy, solid, scale, wmax);
  return copy_out(p.data(), p.size(), out, cap);
}
// render::write_arrows (src/render.rs:209-267), the file's bytes
size_t orc_arrows_svg(const float* ux, const float* uy, const float* solid, size_t bw, size_t bh,
                      size_t bx, size_t by, float scale, float fx, float fy, char* out,
                      size_t cap) {
  auto s = oracle_render::arrows_svg(bw, bh, bx, by, ux, uy, solid, scale, fx, fy);
  return copy_out(s.data(), s.size(), out, cap);
}
// Rust's `{:.1}` and `{}` of an f32
size_t orc_fmt1(float x, char* out, size_t cap) {
  auto s = oracle_render::fmt1(x);
  return copy_out(s.c_str(), s.size() + 1, out, cap);
}
size_t orc_fmt_short(float x, char* out, size_t cap) {
  auto s = oracle_render::fmt_short(x);
  return copy_out(s.c_str(), s.size() + 1, out, cap);
}

}  // extern "C"
