// A compiled consumer of include/lgca.h: plays the reference's Sim
// (src/sim.rs:35-174) over the C ABI for the golden case
// channel/plate/inlet/t4 (tests/golden.rs:49-86) and checks the reference's
// own checksums (tests/golden.rs:141-147). Built by the test setup with the
// host compiler against liblgca_hip.so, so every call below is checked
// against the header's declarations at compile time; run by a gpu test.
// Exit status 0 and a last line "harness: ok" on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/lgca.h"

namespace {

// tests/golden.rs:40-47, with the reference's multiplier as written there:
// 0x1000_0000_01b3 = 2^44 + 0x1b3, not the textbook 2^40 + 0x1b3
uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x1000000001b3ull;
  }
  return h;
}

int check(int rc, const char* what) {
  if (rc < 0) std::fprintf(stderr, "harness: %s failed: %s\n", what, lg_last_error());
  return rc < 0;
}

}  // namespace

int main() {
  const size_t w = 128, h = 64, threads = 4, steps = 200, block = 16;
  const uint64_t seed = 0x5EED;
  const float density = 0.22f, speed = 0.3f, SQRT3_2 = 0.8660254f;
  // Sim::new(lat, true): GpuLattice::new, set_inlet, load of the prepared
  // lattice -- here the obstacle and the equilibrium are made on the device
  lg_t* g = lg_create(w, h, 1, seed, threads);
  if (!g) {
    std::fprintf(stderr, "harness: lg_create failed: %s\n", lg_last_error());
    return 1;
  }
  char name[256];
  int bad = check(lg_device_name(g, name, sizeof name), "lg_device_name");
  if (!bad) std::printf("harness: %s\n", name);
  bad |= check(lg_set_inlet_flow(g, 8, density, speed, 0.0f), "lg_set_inlet_flow");
  const float yc = (float)h * SQRT3_2 / 2.0f, x0 = (float)w / 5.0f;  // tests/golden.rs:78-83
  bad |= check(lg_add_solid_box(g, x0, x0 + 3.0f, yc - 12.0f, yc + 12.0f), "lg_add_solid_box");
  bad |= check(lg_init_equilibrium(g, density, speed, 0.0f), "lg_init_equilibrium");
  bad |= check(lg_attach_field(g, block, block), "lg_attach_field");
  size_t dw = 0, dh = 0;
  bad |= check(lg_dims(g, &dw, &dh), "lg_dims");
  if (dw != w || dh != h) bad = 1;
  // a wrong cell count is refused with a message, not an abort
  std::vector<uint8_t> cells(w * h);
  if (lg_store(g, cells.data(), cells.size() - 1) != LG_EINVAL || !*lg_last_error()) {
    std::fprintf(stderr, "harness: a short buffer was not refused\n");
    bad = 1;
  }
  // Sim::advance, then Sim::sample (src/sim.rs:148-165)
  uint32_t rows = 0, nbands = 0;
  const int resident_before = lg_get_resident(g, &rows, &nbands);
  bad |= check(lg_advance(g, steps), "lg_advance");
  // the engine may run resident bands; if the runtime declines their
  // cooperative launch (another process holding the device's cooperative
  // launch resources), the same steps run in the other pipelines
  std::printf("harness: resident bands %s\n",
              resident_before != 1                    ? "not used"
              : lg_get_resident(g, &rows, &nbands) == 1 ? "used"
                                                        : "declined by the runtime, other pipeline used");
  if (lg_steps(g) != steps) bad = 1;
  bad |= check(lg_sample_field(g, 1.0f), "lg_sample_field");
  size_t bw = 0, bh = 0, bx = 0, by = 0;
  bad |= check(lg_field_dims(g, &bw, &bh, &bx, &by), "lg_field_dims");
  const size_t nb = bw * bh;
  std::vector<float> ux(nb), uy(nb), rho(nb), sol(nb);
  bad |= check(lg_read_field(g, ux.data(), uy.data(), rho.data(), sol.data(), nb), "lg_read_field");
  std::vector<int64_t> mass(nb), px(nb), py(nb), ns(nb);
  bad |= check(lg_read_block_sums(g, mass.data(), px.data(), py.data(), ns.data(), nb),
               "lg_read_block_sums");
  // Sim::cells, Sim::total_particles
  bad |= check(lg_store(g, cells.data(), cells.size()), "lg_store");
  uint64_t total = 0;
  bad |= check(lg_total_particles(g, &total), "lg_total_particles");
  lg_destroy(g);
  if (bad) return 1;

  // tests/golden.rs:146: cells, field (ux, uy, rho, solid as little-endian
  // f32, one array after another), exact mass
  std::vector<uint8_t> fb;
  for (const std::vector<float>* a : {&ux, &uy, &rho, &sol}) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(a->data());
    fb.insert(fb.end(), p, p + a->size() * 4);
  }
  const uint64_t hc = fnv1a(cells.data(), cells.size()), hf = fnv1a(fb.data(), fb.size());
  int64_t msum = 0;
  for (int64_t m : mass) msum += m;
  std::printf("harness: cells %016llx field %016llx particles %llu block mass %lld\n",
              (unsigned long long)hc, (unsigned long long)hf, (unsigned long long)total,
              (long long)msum);
  if (hc != 0x151d9c62d8f873cfull || hf != 0x4f0354593ef1135dull || total != 11685) return 1;
  std::printf("harness: ok\n");
  return 0;
}
