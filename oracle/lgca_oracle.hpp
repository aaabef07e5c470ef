// CPU oracle for the FHP lattice-gas step: a line-by-line restatement of the
// reference's byte-per-cell CPU path.
//
// TEST INFRASTRUCTURE ONLY. This file is what the HIP engine's results are
// compared against; it is never part of the product. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Every function cites the reference file:line it restates. All f32
// arithmetic is plain IEEE single precision evaluated in source order; the
// library is compiled with -ffp-contract=off so that no multiply-add is fused,
// matching Rust, which never contracts.
//
// Pins (what checks each part, beyond restating it):
//   Rng, Lattice::step/apply_inlet/init_equilibrium, Field::sample, fnv1a:
//       the five-row GOLDEN table of tests/golden.rs:141-147 (bit-exact cell
//       and field checksums, exact mass/momentum) -- tests/test_oracle_golden.py
//   CollisionTable: conservation + closure (src/collision.rs:126-150) and the
//       class census 80/53 recomputed from the enumeration -- tests/test_oracle_rules.py
//   moments: brute force over all 256 bytes (src/moments.rs:147-175)
//   propagation/walls: straight flight and wall reversal closed forms
//       (src/lattice.rs:408-465)
//   Equilibrium: rate/correlation/velocity statistics (src/lattice.rs:471-559)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

namespace oracle {

// ---- src/hex.rs:7-55 -------------------------------------------------------
constexpr int NDIR = 6;
constexpr uint8_t REST_BIT = 1u << 6;
constexpr uint8_t SOLID_BIT = 1u << 7;
constexpr uint8_t STATE_MASK = 0x7F;
constexpr uint8_t MOVING_MASK = 0x3F;
constexpr float SQRT3_2 = 0.8660254f;  // f32 0x3f5db3d7 (src/hex.rs:25)
constexpr int CX2[NDIR] = {2, 1, -1, -2, -1, 1};
constexpr int CY2[NDIR] = {0, 1, 1, 0, -1, -1};
constexpr float CXF[NDIR] = {1.0f, 0.5f, -0.5f, -1.0f, -0.5f, 0.5f};
constexpr float CYF[NDIR] = {0.0f, SQRT3_2, SQRT3_2, 0.0f, -SQRT3_2, -SQRT3_2};
constexpr int OPP[NDIR] = {3, 4, 5, 0, 1, 2};
constexpr int DX_EVEN[NDIR] = {1, 0, -1, -1, -1, 0};
constexpr int DX_ODD[NDIR] = {1, 1, 0, -1, 0, 1};
constexpr int DY[NDIR] = {0, 1, 1, 0, -1, -1};

// src/hex.rs:45-48
inline uint8_t reverse(uint8_t state) {
  uint8_t m = state & MOVING_MASK;
  return (uint8_t)((state & REST_BIT) | ((uint8_t)((m << 3) | (m >> 3)) & MOVING_MASK));
}

// src/hex.rs:52-55
inline void site_position(size_t x, size_t y, float* px, float* py) {
  *px = (float)x + ((y & 1) == 1 ? 0.5f : 0.0f);
  *py = (float)y * SQRT3_2;
}

// ---- src/rng.rs:4-35 -------------------------------------------------------
struct Rng {
  uint64_t s;
  // src/rng.rs:8-14
  explicit Rng(uint64_t seed) {
    uint64_t x = seed ^ 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    s = (x ^ (x >> 31)) | 1;
  }
  // src/rng.rs:17-24
  inline uint64_t next_u64() {
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return x * 0x2545F4914F6CDD1Dull;
  }
  // src/rng.rs:27-29
  inline uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
};

// ---- src/lattice.rs:41-116 -------------------------------------------------
// src/lattice.rs:59-69
inline uint32_t threshold(float p) {
  const float ONE = 16777216.0f;
  float t = std::ceil(p * ONE);
  if (t >= ONE) return (uint32_t)ONE;
  if (t > 0.0f) return (uint32_t)t;
  return 0;  // also NaN
}

struct Equilibrium {
  uint32_t thresh[NDIR];
  bool has_rest;
  uint32_t rest;
  // src/lattice.rs:72-82
  Equilibrium(float density, float ux, float uy, bool rest_particles) {
    const float bias = rest_particles ? (2.0f * 7.0f / 6.0f) : 2.0f;
    for (int d = 0; d < NDIR; ++d) {
      float dot = CXF[d] * ux + CYF[d] * uy;
      float p = density * (1.0f + bias * dot);
      thresh[d] = threshold(p);
    }
    has_rest = rest_particles;
    rest = rest_particles ? threshold(density) : 0;
  }
  // src/lattice.rs:87-92 (rest particle last, 0 without)
  void thresholds(uint32_t out[7]) const {
    for (int d = 0; d < NDIR; ++d) out[d] = thresh[d];
    out[NDIR] = has_rest ? rest : 0;
  }
  // src/lattice.rs:102-115
  inline uint8_t sample(Rng& rng) const {
    uint8_t s = 0;
    for (int d = 0; d < NDIR; d += 2) {
      uint64_t r = rng.next_u64();
      s |= (uint8_t)(((uint32_t)(r >> 40) < thresh[d]) ? 1 : 0) << d;
      s |= (uint8_t)((((uint32_t)(r >> 16) & 0xFFFFFFu) < thresh[d + 1]) ? 1 : 0) << (d + 1);
    }
    if (has_rest) {
      if ((uint32_t)(rng.next_u64() >> 40) < rest) s |= REST_BIT;
    }
    return s;
  }
};

// ---- src/collision.rs:37-120 -----------------------------------------------
// src/collision.rs:37-52
inline std::tuple<uint32_t, int, int> invariants(uint8_t state) {
  uint32_t mass = 0;
  int px = 0, py = 0;
  for (int d = 0; d < NDIR; ++d)
    if (state & (1u << d)) {
      mass += 1;
      px += CX2[d];
      py += CY2[d];
    }
  if (state & REST_BIT) mass += 1;
  return {mass, px, py};
}

// src/collision.rs:62-69: classes in (mass, px, py) key order (BTreeMap
// ordering = lexicographic tuple order), members ascending.
inline std::vector<std::vector<uint8_t>> classes(bool rest_particles) {
  int n_states = rest_particles ? 128 : 64;
  std::map<std::tuple<uint32_t, int, int>, std::vector<uint8_t>> groups;
  for (int s = 0; s < n_states; ++s) groups[invariants((uint8_t)s)].push_back((uint8_t)s);
  std::vector<std::vector<uint8_t>> out;
  for (auto& kv : groups) out.push_back(kv.second);
  return out;
}

inline size_t gcd_(size_t a, size_t b) { return b == 0 ? a : gcd_(b, a % b); }
inline size_t lcm_(size_t a, size_t b) { return a / gcd_(a, b) * b; }

struct CollisionTable {
  size_t stride = 1;
  std::vector<uint8_t> data;
  size_t n_states = 0;
  bool rest_particles = false;
  // src/collision.rs:72-98
  explicit CollisionTable(bool rest) : rest_particles(rest) {
    n_states = rest ? 128 : 64;
    std::map<std::tuple<uint32_t, int, int>, std::vector<uint8_t>> groups;
    for (size_t s = 0; s < n_states; ++s) groups[invariants((uint8_t)s)].push_back((uint8_t)s);
    stride = 1;
    for (auto& kv : groups) stride = lcm_(stride, kv.second.size());
    data.assign(n_states * stride, 0);
    for (size_t s = 0; s < n_states; ++s) {
      const auto& g = groups[invariants((uint8_t)s)];
      for (size_t k = 0; k < stride; ++k) data[s * stride + k] = g[k % g.size()];
    }
  }
  // src/collision.rs:101-108
  size_t active_states() const {
    size_t n = 0;
    for (size_t s = 0; s < n_states; ++s) {
      bool any = false;
      for (size_t k = 0; k < stride; ++k) any |= data[s * stride + k] != s;
      n += any;
    }
    return n;
  }
  // src/collision.rs:115-119
  inline uint8_t apply(uint8_t state, uint32_t rand) const {
    size_t k = (size_t)(((uint64_t)rand * (uint64_t)stride) >> 32);
    return data[(size_t)state * stride + k];
  }
};

// ---- src/moments.rs:16-138 -------------------------------------------------
constexpr uint32_t M_MASS = 0, M_PX = 16, M_PY = 32, M_SOLID = 48;
constexpr int64_t PX_BIAS = 4, PY_BIAS = 2;
constexpr size_t MAX_BLOCK = 65535 / 8;  // src/moments.rs:32 (8191)

// src/moments.rs:37-65
inline void moment_table(bool walls, uint64_t t[256]) {
  for (int s = 0; s < 256; ++s) {
    uint8_t c = (uint8_t)s;
    bool solid = (c & SOLID_BIT) != 0;
    int64_t mass = 0, px = 0, py = 0;
    if (walls || !solid) {
      for (int d = 0; d < NDIR; ++d)
        if (c & (1u << d)) {
          mass += 1;
          px += CX2[d];
          py += CY2[d];
        }
      if (c & REST_BIT) mass += 1;
    }
    t[s] = ((uint64_t)mass << M_MASS) | ((uint64_t)(px + PX_BIAS) << M_PX) |
           ((uint64_t)(py + PY_BIAS) << M_PY) | ((uint64_t)(solid ? 1 : 0) << M_SOLID);
  }
}

struct Tables {
  uint64_t fluid[256];
  uint64_t with_walls[256];
  Tables() {
    moment_table(false, fluid);
    moment_table(true, with_walls);
  }
};
inline const Tables& tables() {
  static Tables t;
  return t;
}

struct Moments {
  int64_t mass = 0, px2 = 0, py2 = 0, solid = 0;
  // src/moments.rs:91-97
  void velocity(float* vx, float* vy) const {
    if (mass == 0) {
      *vx = 0.0f;
      *vy = 0.0f;
      return;
    }
    float m = (float)mass;
    *vx = (float)px2 * 0.5f / m;
    *vy = (float)py2 * SQRT3_2 / m;
  }
};

// src/moments.rs:103-110
inline Moments unpack(uint64_t acc, size_t n) {
  Moments m;
  m.mass = (int64_t)((acc >> M_MASS) & 0xFFFF);
  m.px2 = (int64_t)((acc >> M_PX) & 0xFFFF) - PX_BIAS * (int64_t)n;
  m.py2 = (int64_t)((acc >> M_PY) & 0xFFFF) - PY_BIAS * (int64_t)n;
  m.solid = (int64_t)((acc >> M_SOLID) & 0xFFFF);
  return m;
}

// src/moments.rs:126-138
inline Moments sum(const uint64_t table[256], const uint8_t* cells, size_t n) {
  Moments total;
  for (size_t c0 = 0; c0 < n; c0 += MAX_BLOCK) {
    size_t len = std::min(MAX_BLOCK, n - c0);
    uint64_t acc = 0;
    for (size_t i = 0; i < len; ++i) acc += table[cells[c0 + i]];
    Moments m = unpack(acc, len);
    total.mass += m.mass;
    total.px2 += m.px2;
    total.py2 += m.py2;
    total.solid += m.solid;
  }
  return total;
}

// ---- src/lattice.rs:8-364 --------------------------------------------------
struct Lattice {
  size_t w, h;
  std::vector<uint8_t> cells, scratch;
  CollisionTable table;
  float density;
  float inflow_x, inflow_y;
  size_t inlet_cols;
  uint64_t steps = 0;
  size_t threads;
  uint64_t seed;
  // When true, each band of a step runs on its own std::thread, spawned per
  // step as src/lattice.rs:239-259 does. The result does not depend on it.
  bool parallel = false;

  // src/lattice.rs:127-152 (the asserts are the caller's job here; the C API
  // checks them and returns an error instead of aborting)
  Lattice(size_t w_, size_t h_, float density_, float ux, float uy, size_t inlet_cols_,
          bool rest, size_t threads_, uint64_t seed_)
      : w(w_), h(h_), cells(w_ * h_, 0), scratch(w_ * h_, 0), table(rest), density(density_),
        inflow_x(ux), inflow_y(uy), inlet_cols(inlet_cols_), threads(std::max<size_t>(threads_, 1)),
        seed(seed_) {}

  // src/lattice.rs:188-193
  void restart(const uint8_t* c, uint64_t seed_) {
    std::memcpy(cells.data(), c, cells.size());
    seed = seed_;
    steps = 0;
  }

  // src/lattice.rs:197-208
  void init_equilibrium() {
    Rng rng(seed ^ 0xA5A51234ull);
    Equilibrium eq(density, inflow_x, inflow_y, table.rest_particles);
    for (size_t i = 0; i < cells.size(); ++i)
      cells[i] = (cells[i] & SOLID_BIT) ? SOLID_BIT : eq.sample(rng);
  }

  // init_equilibrium (src/lattice.rs:197-208) with one equilibrium per column
  // (axis 0) or per row (axis 1), the column form being how
  // src/transport.rs:238-251 builds its start states.
  void init_profile(int axis, const float* ux, const float* uy) {
    Rng rng(seed ^ 0xA5A51234ull);
    size_t n = axis == 0 ? w : h;
    std::vector<Equilibrium> eq;
    for (size_t i = 0; i < n; ++i) eq.emplace_back(density, ux[i], uy[i], table.rest_particles);
    for (size_t i = 0; i < cells.size(); ++i) {
      size_t k = axis == 0 ? i % w : i / w;
      cells[i] = (cells[i] & SOLID_BIT) ? SOLID_BIT : eq[k].sample(rng);
    }
  }

  // src/lattice.rs:210-220 with the plate predicate of tests/golden.rs:78-83
  // and src/main.rs:584-588: x0 <= px <= x1 && y0 <= py <= y1.
  void add_solid_box(float x0, float x1, float y0, float y1) {
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        float px, py;
        site_position(x, y, &px, &py);
        if (px >= x0 && px <= x1 && py >= y0 && py <= y1) cells[y * w + x] = SOLID_BIT;
      }
  }
  // src/main.rs:589-592: (x-cx)^2 + (y-yc)^2 <= r*r, with powi(2) == x*x.
  void add_solid_disc(float cx, float cy, float r) {
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        float px, py;
        site_position(x, y, &px, &py);
        float dx = px - cx, dy = py - cy;
        if (dx * dx + dy * dy <= r * r) cells[y * w + x] = SOLID_BIT;
      }
  }

  // src/lattice.rs:324-364
  static void stream_and_collide_row(const uint8_t* old, uint8_t* out, const CollisionTable& table,
                                     Rng& rng, size_t y, size_t w, size_t h) {
    const int* dxtab = (y & 1) ? DX_ODD : DX_EVEN;
    int sdx[NDIR];
    size_t srow[NDIR];
    for (int d = 0; d < NDIR; ++d) {
      int back = OPP[d];
      sdx[d] = dxtab[back];
      long sy = ((long)y + DY[back]) % (long)h;
      if (sy < 0) sy += (long)h;
      srow[d] = (size_t)sy * w;
    }
    size_t row = y * w;
    for (size_t x = 0; x < w; ++x) {
      uint8_t here = old[row + x];
      uint8_t state = here & REST_BIT;
      for (int d = 0; d < NDIR; ++d) {
        long sx = (long)x + sdx[d];
        if (sx < 0) sx += (long)w;
        if (sx >= (long)w) sx -= (long)w;
        state |= old[srow[d] + (size_t)sx] & (uint8_t)(1u << d);
      }
      out[x] = (here & SOLID_BIT) ? (uint8_t)(SOLID_BIT | reverse(state))
                                  : table.apply(state, rng.next_u32());
    }
  }

  // Band geometry of src/lattice.rs:232-240: the rows are cut into
  // ceil(h / band) chunks of `band` rows, band = ceil(h / min(threads, h)).
  size_t band_rows() const {
    size_t nthreads = std::min(threads, h);
    return (h + nthreads - 1) / nthreads;
  }
  size_t n_bands() const {
    size_t band = band_rows();
    return (h + band - 1) / band;
  }

  // src/lattice.rs:228-264
  void step() {
    const uint8_t* old = cells.data();
    uint64_t base_seed = seed ^ (steps * 0x9E3779B97F4A7C15ull);
    size_t band = band_rows();
    size_t nb = n_bands();
    auto run_band = [&](size_t b) {
      size_t y0 = b * band;
      size_t y1 = std::min(h, y0 + band);
      Rng rng(base_seed ^ ((uint64_t)(b + 1) << 32));
      for (size_t y = y0; y < y1; ++y)
        stream_and_collide_row(old, scratch.data() + y * w, table, rng, y, w, h);
    };
    if (parallel && nb > 1) {
      std::vector<std::thread> pool;
      pool.reserve(nb);
      for (size_t b = 0; b < nb; ++b) pool.emplace_back(run_band, b);
      for (auto& t : pool) t.join();
    } else {
      for (size_t b = 0; b < nb; ++b) run_band(b);
    }
    std::swap(cells, scratch);
    steps += 1;
    apply_inlet();
  }

  void advance(uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) step();
  }

  // src/lattice.rs:270-285
  void apply_inlet() {
    if (inlet_cols == 0) return;
    Equilibrium eq(density, inflow_x, inflow_y, table.rest_particles);
    Rng rng(seed ^ (steps * 0xD1B54A32D192ED03ull));
    size_t cols = std::min(inlet_cols, w);
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < cols; ++x) {
        size_t i = y * w + x;
        if ((cells[i] & SOLID_BIT) == 0) cells[i] = eq.sample(rng);
      }
  }

  // src/lattice.rs:292-294
  void mean_velocity(float* ux, float* uy) const {
    sum(tables().with_walls, cells.data(), cells.size()).velocity(ux, uy);
  }

  // src/lattice.rs:299-312
  uint64_t total_particles() const {
    uint64_t n = 0;
    for (uint8_t c : cells) n += (uint64_t)__builtin_popcount(c & STATE_MASK);
    return n;
  }
};

// ---- src/render.rs:21-120 --------------------------------------------------
struct Field {
  size_t bw, bh, bx, by;
  std::vector<float> ux, uy, rho, solid;
  // src/render.rs:35-48
  Field(size_t w, size_t h, size_t bx_, size_t by_) : bw(w / bx_), bh(h / by_), bx(bx_), by(by_) {
    ux.assign(bw * bh, 0.0f);
    uy.assign(bw * bh, 0.0f);
    rho.assign(bw * bh, 0.0f);
    solid.assign(bw * bh, 0.0f);
  }

  // Exact per-block integer moments over fluid cells (the sums the f32 field
  // is computed from; src/render.rs:93-106).
  Moments block_moments(const Lattice& lat, size_t ib, size_t jb) const {
    uint64_t acc = 0;
    const auto& T = tables().fluid;
    for (size_t y = jb * by; y < (jb + 1) * by; ++y) {
      const uint8_t* row = lat.cells.data() + y * lat.w + ib * bx;
      for (size_t x = 0; x < bx; ++x) acc += T[row[x]];
    }
    return unpack(acc, bx * by);
  }

  // src/render.rs:61-120. Each block is independent of the thread split
  // (src/render.rs:53-56), so the blocks are visited in order here.
  // Returns false when the block exceeds MAX_BLOCK (the reference asserts).
  bool sample(const Lattice& lat, float alpha) {
    if (bx * by > MAX_BLOCK) return false;
    if (bw == 0 || bh == 0) return true;
    float n = (float)(bx * by);
    for (size_t jb = 0; jb < bh; ++jb)
      for (size_t ib = 0; ib < bw; ++ib) {
        Moments m = block_moments(lat, ib, jb);
        float vx, vy;
        m.velocity(&vx, &vy);
        size_t k = jb * bw + ib;
        ux[k] += alpha * (vx - ux[k]);
        uy[k] += alpha * (vy - uy[k]);
        rho[k] += alpha * ((float)m.mass / n - rho[k]);
        solid[k] = (float)m.solid / n;
      }
    return true;
  }
};

// ---- tests/golden.rs:40-47 -------------------------------------------------
// The multiplier is the reference's literal 0x1000_0000_01b3 (2^44 + 0x1b3),
// which is not the textbook 64-bit FNV prime (2^40 + 0x1b3); the GOLDEN
// checksums were recorded with it.
inline uint64_t fnv1a(const uint8_t* bytes, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
  for (size_t i = 0; i < n; ++i) {
    h ^= bytes[i];
    h *= 0x1000000001b3ull;
  }
  return h;
}

// tests/golden.rs:124-127: ux ‖ uy ‖ rho ‖ solid as little-endian f32 bits.
inline uint64_t field_checksum(const Field& f) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (const auto* v : {&f.ux, &f.uy, &f.rho, &f.solid})
    h = fnv1a(reinterpret_cast<const uint8_t*>(v->data()), v->size() * 4, h);
  return h;
}

}  // namespace oracle
