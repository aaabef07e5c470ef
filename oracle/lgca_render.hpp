// CPU oracle for the frame output of the reference run loop: the vorticity map
// (src/render.rs:122-204), the arrow plot (src/render.rs:205-267), the frame's
// colour scale (src/main.rs:219-229) and the PNG writer (src/png.rs).
//
// TEST INFRASTRUCTURE ONLY: what results are compared with. Only tests/ and
// bench.py's cpu_baseline leg load it. f32 expressions are evaluated in source
// order with no contraction (built -ffp-contract=off). Rust's `as u8` from f32
// saturates (NaN to 0), `{:.1}` rounds the exact binary value half to even
// and `{}` of an f32 prints the shortest digits that read back to it.
//
// Pins: the reference's docs/vortex-street.png re-encoded from its pixels
// byte for byte, and docs/velocity-arrows.svg rebuilt from its arrows; the
// PNG container and zlib stream against Python's zlib (inflate, CRC-32,
// Adler-32); the reference's deflate tests (src/png.rs:262-313); the
// vorticity of linear flows against its closed form; `{:.1}` against
// Python's formatting and `{}` against numpy's shortest digits.
#pragma once
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace oracle_render {

constexpr float SQRT3_2 = 0.8660254f;  // src/hex.rs:25
constexpr float SOLID = 0.12f;         // src/render.rs:17

// src/render.rs:122-139
inline std::vector<float> vorticity(size_t bw, size_t bh, size_t bx, size_t by, const float* ux,
                                    const float* uy) {
  const float dx = (float)bx;
  const float dy = (float)by * SQRT3_2;
  std::vector<float> w(bw * bh, 0.0f);
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i) {
      size_t ip = std::min(i + 1, bw - 1);
      size_t im = i == 0 ? 0 : i - 1;
      size_t jp = j + 1 < bh ? j + 1 : 0;
      size_t jm = j == 0 ? bh - 1 : j - 1;
      float dvdx = (uy[j * bw + ip] - uy[j * bw + im]) / ((float)(ip - im) * dx);
      float dudy = (ux[jp * bw + i] - ux[jm * bw + i]) / (2.0f * dy);
      w[j * bw + i] = dvdx - dudy;
    }
  return w;
}

// src/render.rs:142-157
inline void mean_velocity(const float* ux, const float* uy, const float* solid, size_t n,
                          float* mx, float* my) {
  float sx = 0.0f, sy = 0.0f, c = 0.0f;
  for (size_t k = 0; k < n; ++k)
    if (solid[k] < 0.5f) {
      sx += ux[k];
      sy += uy[k];
      c += 1.0f;
    }
  *mx = c == 0.0f ? 0.0f : sx / c;
  *my = c == 0.0f ? 0.0f : sy / c;
}

// src/main.rs:219-229: the 98th percentile of |vorticity| over fluid blocks.
inline float vorticity_scale(const std::vector<float>& vort, const float* solid) {
  std::vector<float> v;
  for (size_t k = 0; k < vort.size(); ++k)
    if (solid[k] < 0.12f) v.push_back(std::fabs(vort[k]));
  std::sort(v.begin(), v.end());
  size_t idx = v.size() * 98 / 100;
  float x = idx < v.size() ? v[idx] : 1e-3f;
  return std::fmax(x, 1e-5f);
}

// Rust's saturating f32 -> u8 cast.
inline uint8_t as_u8(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

// src/render.rs:159-173
inline void diverging(float t, uint8_t rgb[3]) {
  if (!std::isnan(t)) t = std::min(std::max(t, -1.0f), 1.0f);
  float r, g, b;
  if (t >= 0.0f) {
    r = 1.0f;
    g = 1.0f - 0.75f * t;
    b = 1.0f - 0.9f * t;
  } else {
    r = 1.0f + 0.9f * t;
    g = 1.0f + 0.6f * t;
    b = 1.0f;
  }
  rgb[0] = as_u8(r * 255.0f);
  rgb[1] = as_u8(g * 255.0f);
  rgb[2] = as_u8(b * 255.0f);
}

// Rust's saturating f32 -> i32 cast.
inline long long as_i32(float x) {
  if (std::isnan(x)) return 0;
  return (long long)std::fmax(std::fmin((double)x, 2147483647.0), -2147483648.0);
}

// Rust's f32::round (half away from zero), then the saturating usize cast.
inline size_t round_usize(float x) {
  float r = std::round(x);
  if (std::isnan(r) || r <= 0.0f) return 0;
  return (size_t)r;
}

// ---- src/png.rs -------------------------------------------------------------

inline uint32_t crc32(const uint8_t* d, size_t n) {  // src/png.rs:4-13
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; ++i) {
    c ^= d[i];
    for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
  }
  return c ^ 0xFFFFFFFFu;
}

inline uint32_t adler32(const uint8_t* d, size_t n) {  // src/png.rs:15-22
  uint32_t a = 1, b = 0;
  for (size_t i = 0; i < n; ++i) {
    a = (a + d[i]) % 65521u;
    b = (b + a) % 65521u;
  }
  return (b << 16) | a;
}

struct BitWriter {  // src/png.rs:27-62
  std::vector<uint8_t> out;
  uint32_t buf = 0, n = 0;
  void bits(uint32_t value, uint32_t count) {
    buf |= value << n;
    n += count;
    while (n >= 8) {
      out.push_back((uint8_t)buf);
      buf >>= 8;
      n -= 8;
    }
  }
  void code(uint32_t value, uint32_t count) {
    uint32_t rev = 0;
    for (uint32_t i = 0; i < count; ++i) rev |= ((value >> i) & 1u) << (count - 1 - i);
    bits(rev, count);
  }
  std::vector<uint8_t> finish() {
    if (n > 0) out.push_back((uint8_t)buf);
    return out;
  }
};

// src/png.rs:64-78
constexpr uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint32_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                    2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t DIST_BASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                    33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                    1025, 1537, 2049, 3073, 4097, 6145,  8193,  12289, 16385, 24577};
constexpr uint32_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                     6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

inline void fixed_symbol(uint32_t sym, uint32_t* code, uint32_t* bits) {  // src/png.rs:81-88
  if (sym <= 143) {
    *code = 0x30u + sym;
    *bits = 8;
  } else if (sym <= 255) {
    *code = 0x190u + (sym - 144);
    *bits = 9;
  } else if (sym <= 279) {
    *code = sym - 256;
    *bits = 7;
  } else {
    *code = 0xC0u + (sym - 280);
    *bits = 8;
  }
}

constexpr size_t WINDOW = 32768, MIN_MATCH = 3, MAX_MATCH = 258, MAX_CHAIN = 32;  // :90-96
constexpr size_t NONE = ~(size_t)0;

inline size_t hash3(const uint8_t* d) {  // src/png.rs:98-100
  return (((size_t)d[0] << 10) ^ ((size_t)d[1] << 5) ^ (size_t)d[2]) & 0x7FFF;
}

// partition_point(|b| b <= x) - 1 over an ascending table
template <size_t N>
inline size_t last_le(const uint16_t (&t)[N], size_t x) {
  size_t k = 0;
  while (k < N && t[k] <= x) ++k;
  return k - 1;
}

inline std::vector<uint8_t> deflate(const uint8_t* data, size_t n) {  // src/png.rs:103-162
  BitWriter bw;
  bw.bits(1, 1);
  bw.bits(1, 2);
  std::vector<size_t> head(1u << 15, NONE), prev(std::max<size_t>(n, 1), NONE);
  size_t i = 0;
  while (i < n) {
    size_t best_len = 0, best_dist = 0;
    if (i + MIN_MATCH <= n) {
      size_t limit = std::min(n - i, MAX_MATCH);
      size_t cand = head[hash3(data + i)];
      size_t chain = 0;
      while (cand != NONE && chain < MAX_CHAIN && i - cand <= WINDOW) {
        size_t l = 0;
        while (l < limit && data[cand + l] == data[i + l]) ++l;
        if (l > best_len) {
          best_len = l;
          best_dist = i - cand;
          if (l == limit) break;
        }
        cand = prev[cand];
        ++chain;
      }
    }
    size_t advance;
    uint32_t c, b;
    if (best_len >= MIN_MATCH) {
      size_t li = last_le(LEN_BASE, best_len);
      fixed_symbol(257 + (uint32_t)li, &c, &b);
      bw.code(c, b);
      bw.bits((uint32_t)(best_len - LEN_BASE[li]), LEN_EXTRA[li]);
      size_t di = last_le(DIST_BASE, best_dist);
      bw.code((uint32_t)di, 5);
      bw.bits((uint32_t)(best_dist - DIST_BASE[di]), DIST_EXTRA[di]);
      advance = best_len;
    } else {
      fixed_symbol(data[i], &c, &b);
      bw.code(c, b);
      advance = 1;
    }
    for (size_t k = i; k < i + advance; ++k)
      if (k + MIN_MATCH <= n) {
        size_t h = hash3(data + k);
        prev[k] = head[h];
        head[h] = k;
      }
    i += advance;
  }
  uint32_t c, b;
  fixed_symbol(256, &c, &b);
  bw.code(c, b);
  return bw.finish();
}

inline void put_be32(std::vector<uint8_t>& o, uint32_t v) {
  o.push_back((uint8_t)(v >> 24));
  o.push_back((uint8_t)(v >> 16));
  o.push_back((uint8_t)(v >> 8));
  o.push_back((uint8_t)v);
}

inline void chunk(std::vector<uint8_t>& out, const char kind[4], const std::vector<uint8_t>& body) {
  put_be32(out, (uint32_t)body.size());  // src/png.rs:164-171
  std::vector<uint8_t> framed(kind, kind + 4);
  framed.insert(framed.end(), body.begin(), body.end());
  out.insert(out.end(), framed.begin(), framed.end());
  put_be32(out, crc32(framed.data(), framed.size()));
}

inline std::vector<uint8_t> png_rgb(size_t w, size_t h, const uint8_t* rgb) {  // :174-197
  std::vector<uint8_t> raw;
  raw.reserve(h * (1 + w * 3));
  for (size_t y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb + y * w * 3, rgb + (y + 1) * w * 3);
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  std::vector<uint8_t> d = deflate(raw.data(), raw.size());
  z.insert(z.end(), d.begin(), d.end());
  put_be32(z, adler32(raw.data(), raw.size()));
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, (uint32_t)w);
  put_be32(ihdr, (uint32_t)h);
  for (uint8_t v : {8, 2, 0, 0, 0}) ihdr.push_back(v);
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", {});
  return png;
}

// ---- frames ------------------------------------------------------------------

// src/render.rs:175-200
inline std::vector<uint8_t> vorticity_png(size_t bw, size_t bh, size_t bx, size_t by,
                                          const float* ux, const float* uy, const float* solid,
                                          size_t scale, float wmax) {
  std::vector<float> vort = vorticity(bw, bh, bx, by, ux, uy);
  size_t px = std::max<size_t>(scale, 1);
  size_t py = std::max<size_t>(round_usize((float)px * SQRT3_2 * (float)by / (float)bx), 1);
  size_t iw = bw * px, ih = bh * py;
  std::vector<uint8_t> img(iw * ih * 3, 0);
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i) {
      size_t k = j * bw + i;
      uint8_t c[3] = {30, 30, 35};
      if (!(solid[k] > SOLID)) diverging(vort[k] / wmax, c);
      for (size_t dy = 0; dy < py; ++dy) {
        size_t iy = ih - 1 - (j * py + dy);
        for (size_t dx = 0; dx < px; ++dx) {
          size_t o = (iy * iw + i * px + dx) * 3;
          img[o] = c[0];
          img[o + 1] = c[1];
          img[o + 2] = c[2];
        }
      }
    }
  return png_rgb(iw, ih, img.data());
}

// Rust prints non-finite f32 as "NaN", "inf", "-inf".
inline bool rust_nonfinite(float x, std::string& out) {
  if (std::isnan(x)) return out = "NaN", true;
  if (std::isinf(x)) return out = x < 0 ? "-inf" : "inf", true;
  return false;
}

// `{:.1}`: the exact binary value rounded half to even at one decimal, signed
// whenever the sign bit is set. glibc's printf rounds exactly that way.
inline std::string fmt1(float x) {
  std::string s;
  if (rust_nonfinite(x, s)) return s;
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.1f", (double)x);
  return buf;
}

// Digits d1 d2 ... and exponent e of x = 0.d1d2... x 10^e, shortest that reads
// back to the same f32: the first precision of "%.*e" that round-trips.
inline void shortest_digits(float x, std::string& digits, int& e10) {
  char buf[64];
  for (int prec = 0; prec < 12; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec, (double)std::fabs(x));
    if (std::strtof(buf, nullptr) == std::fabs(x)) break;
  }
  char* e = std::strchr(buf, 'e');
  digits.clear();
  for (char* q = buf; q < e; ++q)
    if (*q != '.') digits += *q;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  e10 = std::atoi(e + 1) + 1;
}

// `{}`: the shortest digits that read back to the same f32, written out in
// plain decimal (zeros fill out to the point; no exponent).
inline std::string fmt_short(float x) {
  std::string s;
  if (rust_nonfinite(x, s)) return s;
  if (std::signbit(x)) s = "-";
  if (x == 0.0f) return s + "0";
  std::string d;
  int pos;
  shortest_digits(x, d, pos);
  if (pos <= 0) return s + "0." + std::string((size_t)-pos, '0') + d;
  if ((size_t)pos >= d.size()) return s + d + std::string((size_t)pos - d.size(), '0');
  return s + d.substr(0, (size_t)pos) + "." + d.substr((size_t)pos);
}

// src/render.rs:205-266
inline std::string arrows_svg(size_t bw, size_t bh, size_t bx, size_t by, const float* ux,
                              const float* uy, const float* solid, float scale, float fx,
                              float fy) {
  const float w = (float)bw * (float)bx;
  const float h = (float)bh * (float)by * SQRT3_2;
  std::string s;
  s += "<svg xmlns=\"http://www.w3.org/2000/svg\" viewBox=\"0 0 " + fmt_short(w) + " " +
       fmt_short(h) + "\" width=\"" + std::to_string(as_i32(std::round(w))) + "\" height=\"" +
       std::to_string(as_i32(std::round(h))) +
       "\">\n<rect width=\"100%\" height=\"100%\" fill=\"#ffffff\"/>\n"
       "<g stroke=\"#1a1a1a\" stroke-width=\"1.1\" fill=\"none\" stroke-linecap=\"round\">\n";
  const float arrow = (float)bx * scale;
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i) {
      size_t k = j * bw + i;
      if (solid[k] > SOLID) continue;
      float cx = ((float)i + 0.5f) * (float)bx;
      float cy = h - ((float)j + 0.5f) * (float)by * SQRT3_2;
      float vx = (ux[k] - fx) * arrow;
      float vy = -(uy[k] - fy) * arrow;
      float len = std::sqrt(vx * vx + vy * vy);
      if (len < 0.35f) continue;
      float x0 = cx - vx * 0.5f, y0 = cy - vy * 0.5f;
      float x1 = cx + vx * 0.5f, y1 = cy + vy * 0.5f;
      float hx = vx / len, hy = vy / len;
      float head = std::fmin(len * 0.35f, (float)bx * 0.4f);  // f32::min
      s += "<path d=\"M" + fmt1(x0) + " " + fmt1(y0) + "L" + fmt1(x1) + " " + fmt1(y1) + "m" +
           fmt1(-head * (hx + hy * 0.5f)) + " " + fmt1(-head * (hy - hx * 0.5f)) + "L" +
           fmt1(x1) + " " + fmt1(y1) + "l" + fmt1(-head * (hx - hy * 0.5f)) + " " +
           fmt1(-head * (hy + hx * 0.5f)) + "\"/>\n";
    }
  s += "</g>\n<g fill=\"#1a1a1a\">\n";
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i)
      if (solid[j * bw + i] > SOLID) {
        float x = (float)i * (float)bx;
        float y = h - (float)(j + 1) * (float)by * SQRT3_2;
        s += "<rect x=\"" + fmt1(x) + "\" y=\"" + fmt1(y) + "\" width=\"" + fmt1((float)bx) +
             "\" height=\"" + fmt1((float)by * SQRT3_2) + "\"/>\n";
      }
  s += "</g>\n</svg>\n";
  return s;
}

}  // namespace oracle_render
