// CPU oracle for the transport measurement (src/transport.rs): viscosity from
// the decay of a transverse shear wave, advection factor from the phase drift
// of a wave carried by a uniform stream.
//
// TEST INFRASTRUCTURE ONLY: what results are compared with. f32 expressions
// are evaluated in source order with no contraction; f32 sin/cos are the C library's sinf/cosf
// (the reference calls f32::sin/cos, i.e. the platform libm); f64 as in Rust.
#pragma once
#include <cmath>
#include <vector>

#include "lgca_oracle.hpp"

namespace oracle {

// src/transport.rs:24-29
constexpr uint64_t TRANSIENT = 200;
constexpr uint64_t MIN_INTERVAL = 4;
constexpr size_t SAMPLES = 20;
constexpr uint64_t REPS = 4;
constexpr uint64_t MAX_WINDOW = 30000;

// src/transport.rs:51-84
inline std::vector<float> column_uy(const uint8_t* cells, size_t w, size_t h) {
  std::vector<int64_t> mass(w, 0), py2(w, 0);
  std::vector<uint64_t> acc(w, 0);
  const auto& T = tables().with_walls;
  size_t y0 = 0;
  while (y0 < h) {
    size_t y1 = std::min(y0 + MAX_BLOCK, h);
    std::fill(acc.begin(), acc.end(), 0);
    for (size_t y = y0; y < y1; ++y)
      for (size_t x = 0; x < w; ++x) acc[x] += T[cells[y * w + x]];
    for (size_t x = 0; x < w; ++x) {
      Moments m = unpack(acc[x], y1 - y0);
      mass[x] += m.mass;
      py2[x] += m.py2;
    }
    y0 = y1;
  }
  std::vector<float> u(w);
  for (size_t x = 0; x < w; ++x)
    u[x] = mass[x] > 0 ? (float)py2[x] * SQRT3_2 / (float)mass[x] : 0.0f;
  return u;
}

// src/transport.rs:87-91
inline std::vector<float> row_ux(const uint8_t* cells, size_t w, size_t h) {
  std::vector<float> u(h);
  for (size_t y = 0; y < h; ++y) {
    float vx, vy;
    sum(tables().with_walls, cells + y * w, w).velocity(&vx, &vy);
    u[y] = vx;
  }
  return u;
}

// src/transport.rs:93-100
inline double fit_slope(const std::vector<std::pair<double, double>>& pts) {
  double n = (double)pts.size(), sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (auto& p : pts) sx += p.first;
  for (auto& p : pts) sy += p.second;
  for (auto& p : pts) sxx += p.first * p.first;
  for (auto& p : pts) sxy += p.first * p.second;
  return (n * sxy - sx * sy) / (n * sxx - sx * sx);
}

struct Start {
  std::vector<uint8_t> cells;
  uint64_t seed;
};

// A quiescent periodic box (src/transport.rs:176-179).
inline Lattice box_of(size_t size, float density, bool rest, size_t threads) {
  Lattice lat(size, size, density, 0.0f, 0.0f, 0, rest, threads, 1);
  return lat;
}

template <class Project>
std::vector<std::pair<double, double>> trajectory(Lattice& sim, const Start& s, const Project& project,
                                                  uint64_t interval) {
  // src/transport.rs:109-127
  sim.restart(s.cells.data(), s.seed);
  sim.advance(TRANSIENT);
  std::vector<std::pair<double, double>> out;
  for (size_t i = 0; i < SAMPLES; ++i) {
    out.push_back(project(sim.cells.data()));
    sim.advance(interval);
  }
  return out;
}

template <class Project>
uint64_t decay_window(Lattice& sim, const Start& s, const Project& project) {
  // src/transport.rs:131-150
  sim.restart(s.cells.data(), s.seed);
  sim.advance(TRANSIENT);
  auto ab = project(sim.cells.data());
  double a0 = std::max(std::sqrt(ab.first * ab.first + ab.second * ab.second), 1e-12);
  const uint64_t block = 25;
  uint64_t t = 0;
  while (t < MAX_WINDOW) {
    sim.advance(block);
    t += block;
    auto c = project(sim.cells.data());
    if (std::sqrt(c.first * c.first + c.second * c.second) < a0 / 3.0) break;
  }
  return std::min(std::max(t, (uint64_t)SAMPLES * MIN_INTERVAL), MAX_WINDOW);
}

template <class StartFn, class Project>
void averaged(Lattice& sim, const StartFn& start, const Project& project, std::vector<double>& times,
              std::vector<std::pair<double, double>>& acc) {
  // src/transport.rs:153-171
  uint64_t window = decay_window(sim, start(0), project);
  uint64_t interval = std::max(window / (uint64_t)SAMPLES, MIN_INTERVAL);
  acc.assign(SAMPLES, {0.0, 0.0});
  for (uint64_t rep = 0; rep < REPS; ++rep) {
    auto tr = trajectory(sim, start(rep), project, interval);
    for (size_t i = 0; i < SAMPLES; ++i) {
      acc[i].first += tr[i].first / (double)REPS;
      acc[i].second += tr[i].second / (double)REPS;
    }
  }
  times.resize(SAMPLES);
  for (size_t i = 0; i < SAMPLES; ++i) times[i] = (double)(i * interval);
}

// src/transport.rs:183-221
inline float viscosity(float density, bool rest, size_t threads, uint64_t seed, size_t size) {
  const size_t w = size, h = size;
  const float PI_F = 3.14159265358979323846f;
  const float k = 2.0f * PI_F / ((float)h * SQRT3_2);
  const float amp = 0.05f;
  auto start = [&](uint64_t rep) {
    uint64_t s = seed ^ ((rep * 0x9E3779B9ull) << 20);
    Rng rng(s ^ 0xBEEF);
    Start st{std::vector<uint8_t>(w * h), s};
    for (size_t y = 0; y < h; ++y) {
      float ux = amp * std::sin(k * (float)y * SQRT3_2);
      Equilibrium eq(density, ux, 0.0f, rest);
      for (size_t x = 0; x < w; ++x) st.cells[y * w + x] = eq.sample(rng);
    }
    return st;
  };
  auto project = [&](const uint8_t* cells) {
    auto u = row_ux(cells, w, h);
    double a = 0, b = 0;
    for (size_t y = 0; y < h; ++y) {
      float p = k * (float)y * SQRT3_2;
      a += (double)(u[y] * std::cos(p));
      b += (double)(u[y] * std::sin(p));
    }
    return std::make_pair(2.0 * a / (double)h, 2.0 * b / (double)h);
  };
  Lattice sim = box_of(size, density, rest, threads);
  std::vector<double> times;
  std::vector<std::pair<double, double>> mode;
  averaged(sim, start, project, times, mode);
  std::vector<std::pair<double, double>> pts;
  for (size_t i = 0; i < times.size(); ++i) {
    double a = mode[i].first, b = mode[i].second;
    pts.push_back({times[i], std::log(std::max(std::sqrt(a * a + b * b), 1e-12))});
  }
  return (float)(-fit_slope(pts) / (double)(k * k));
}

// src/transport.rs:225-289
inline float advection_factor(float density, bool rest, size_t threads, uint64_t seed, size_t size) {
  const size_t w = size, h = size;
  const float PI_F = 3.14159265358979323846f;
  const float u0 = 0.25f;
  const float k = 2.0f * PI_F / (float)w;
  const float amp = 0.05f;
  auto start = [&](uint64_t rep) {
    uint64_t s = seed ^ ((rep * 0x9E3779B9ull) << 20) ^ 0x1357;
    Rng rng(s ^ 0xF00D);
    std::vector<Equilibrium> cols;
    for (size_t x = 0; x < w; ++x)
      cols.emplace_back(density, u0, amp * std::sin(k * (float)x), rest);
    Start st{std::vector<uint8_t>(w * h), s};
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) st.cells[y * w + x] = cols[x].sample(rng);
    return st;
  };
  auto project = [&](const uint8_t* cells) {
    auto u = column_uy(cells, w, h);
    double a = 0, b = 0;
    for (size_t x = 0; x < w; ++x) {
      float p = k * (float)x;
      a += (double)(u[x] * std::cos(p));
      b += (double)(u[x] * std::sin(p));
    }
    return std::make_pair(2.0 * a / (double)w, 2.0 * b / (double)w);
  };
  float ua, ub;
  {
    Start s0 = start(0);
    sum(tables().with_walls, s0.cells.data(), s0.cells.size()).velocity(&ua, &ub);
  }
  const float u_actual = ua;
  Lattice sim = box_of(size, density, rest, threads);
  std::vector<double> times;
  std::vector<std::pair<double, double>> mode;
  averaged(sim, start, project, times, mode);
  const double PI = 3.14159265358979323846;
  double unwrapped = 0.0;
  double prev = std::atan2(mode[0].first, mode[0].second);
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
  return (float)(-fit_slope(pts) / (double)(k * u_actual));
}

}  // namespace oracle
