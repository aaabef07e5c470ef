// Frame output: the vorticity map and the velocity-arrow plot of a coarse
// field, the PNG writer under the map, and a pool of writer threads that
// keeps encoding off the stepping loop.
//
// Follows src/render.rs:121-266 (Field::vorticity, Field::mean_velocity,
// diverging, write_vorticity, write_arrows), src/png.rs (crc32, adler32,
// fixed-Huffman deflate with greedy LZ77, write_rgb) and src/main.rs:196-281
// (Frame, Frames, write_frame). For the same field every file is
// byte-identical with the reference's: the deflate makes the same greedy
// parse over the same hash chains, f32 expressions are evaluated in the
// reference's order without contraction, and numbers are printed as Rust
// prints them. Speed comes from elsewhere: a 64-bit bit buffer, 8-byte match
// compares, an early reject on the one byte a longer match needs, table
// lookups for the length and distance codes, scanlines built in place with
// repeated rows copied, and an O(n) selection for the colour scale.
//
// Pure host code: nothing here touches the GPU except lg_frames_write_lattice,
// which reads a lattice's field through lg_read_field.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "lgca_error.h"

using lgca_internal::fail;

namespace {

constexpr float kSqrt3_2 = 0.8660254f;  // src/hex.rs:25
constexpr float kSolid = 0.12f;         // src/render.rs:17

// ---- deflate tables (RFC 1951 3.2.5-3.2.6) -----------------------------------

struct Codes {
  uint16_t lit[288];  // fixed Huffman codes, bit-reversed for LSB-first output
  uint8_t lit_bits[288];
  uint16_t lbase[29];
  uint8_t lextra[29];
  uint16_t dbase[30];
  uint8_t dextra[30];
  uint8_t dcode[30];       // 5-bit distance codes, bit-reversed
  uint8_t len_idx[259];    // match length 3..258 -> length code index
  uint8_t dist_idx[32769]; // distance 1..32768 -> distance code index
  uint32_t crc[4][256];    // CRC-32 (reflected 0xEDB88320), slicing by four

  static uint32_t reverse(uint32_t v, unsigned n) {
    uint32_t r = 0;
    for (unsigned i = 0; i < n; ++i) r |= ((v >> i) & 1u) << (n - 1 - i);
    return r;
  }

  Codes() {
    for (unsigned s = 0; s < 288; ++s) {
      unsigned c, b;
      if (s < 144) c = 0x30 + s, b = 8;
      else if (s < 256) c = 0x190 + (s - 144), b = 9;
      else if (s < 280) c = s - 256, b = 7;
      else c = 0xC0 + (s - 280), b = 8;
      lit[s] = (uint16_t)reverse(c, b);
      lit_bits[s] = (uint8_t)b;
    }
    // Lengths: eight codes without extra bits from 3, four codes for each
    // extra-bit count 1..5, and a last code for 258 alone.
    unsigned k = 0, base = 3;
    for (; k < 8; ++k) lbase[k] = (uint16_t)base++, lextra[k] = 0;
    for (unsigned e = 1; e <= 5; ++e)
      for (unsigned r = 0; r < 4; ++r, ++k) lbase[k] = (uint16_t)base, lextra[k] = (uint8_t)e, base += 1u << e;
    lbase[28] = 258, lextra[28] = 0;
    // Distances: four codes without extra bits from 1, then two codes for
    // each extra-bit count 1..13.
    k = 0, base = 1;
    for (; k < 4; ++k) dbase[k] = (uint16_t)base++, dextra[k] = 0;
    for (unsigned e = 1; e <= 13; ++e)
      for (unsigned r = 0; r < 2; ++r, ++k) dbase[k] = (uint16_t)base, dextra[k] = (uint8_t)e, base += 1u << e;
    for (unsigned i = 0; i < 30; ++i) dcode[i] = (uint8_t)reverse(i, 5);
    // The code of a value is the last one whose base does not exceed it.
    for (unsigned l = 3, i = 0; l <= 258; ++l) {
      while (i + 1 < 29 && lbase[i + 1] <= l) ++i;
      len_idx[l] = (uint8_t)i;
    }
    for (unsigned d = 1, i = 0; d <= 32768; ++d) {
      while (i + 1 < 30 && dbase[i + 1] <= d) ++i;
      dist_idx[d] = (uint8_t)i;
    }
    for (uint32_t n = 0; n < 256; ++n) {
      uint32_t c = n;
      for (int j = 0; j < 8; ++j) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      crc[0][n] = c;
    }
    for (uint32_t n = 0; n < 256; ++n)
      for (int t = 1; t < 4; ++t) crc[t][n] = (crc[t - 1][n] >> 8) ^ crc[0][crc[t - 1][n] & 0xFF];
  }
};

const Codes& codes() {
  static const Codes c;
  return c;
}

// CRC-32 as png::crc32 (src/png.rs:4-13), four bytes per table round.
uint32_t crc32(const uint8_t* d, size_t n) {
  const Codes& T = codes();
  uint32_t c = 0xFFFFFFFFu;
  for (; n >= 4; d += 4, n -= 4) {
    uint32_t w;
    std::memcpy(&w, d, 4);
    c ^= w;
    c = T.crc[3][c & 0xFF] ^ T.crc[2][(c >> 8) & 0xFF] ^ T.crc[1][(c >> 16) & 0xFF] ^ T.crc[0][c >> 24];
  }
  for (; n; --n) c = T.crc[0][(c ^ *d++) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

// Adler-32 as png::adler32 (src/png.rs:15-22), reducing mod 65521 once per
// 5552 bytes, the longest run whose sums cannot overflow 32 bits.
uint32_t adler32(const uint8_t* d, size_t n) {
  uint32_t a = 1, b = 0;
  while (n) {
    size_t m = std::min<size_t>(n, 5552);
    n -= m;
    for (; m; --m) a += *d++, b += a;
    a %= 65521u;
    b %= 65521u;
  }
  return (b << 16) | a;
}

struct BitOut {
  uint8_t* p;
  uint64_t acc = 0;
  unsigned n = 0;
  void put(uint32_t v, unsigned count) {
    acc |= (uint64_t)v << n;
    n += count;
    if (n >= 32) {
      uint32_t w = (uint32_t)acc;
      std::memcpy(p, &w, 4);
      p += 4;
      acc >>= 32;
      n -= 32;
    }
  }
  void finish() {
    while (n) {
      *p++ = (uint8_t)acc;
      acc >>= 8;
      n = n > 8 ? n - 8 : 0;
    }
  }
};

inline uint32_t hash3(const uint8_t* d) {  // src/png.rs:98-100
  return (((uint32_t)d[0] << 10) ^ ((uint32_t)d[1] << 5) ^ (uint32_t)d[2]) & 0x7FFF;
}

inline size_t common_prefix(const uint8_t* a, const uint8_t* b, size_t limit) {
  size_t l = 0;
  for (; l + 8 <= limit; l += 8) {
    uint64_t x, y;
    std::memcpy(&x, a + l, 8);
    std::memcpy(&y, b + l, 8);
    if (uint64_t z = x ^ y) return l + ((unsigned)__builtin_ctzll(z) >> 3);
  }
  while (l < limit && a[l] == b[l]) ++l;
  return l;
}

// Buffers a thread reuses from frame to frame, so that encoding a frame does
// not allocate (and fault in) megabytes each time. The writer threads, and
// any caller thread, keep theirs until they exit.
struct Scratch {
  std::vector<int32_t> head = std::vector<int32_t>(1u << 15);
  std::vector<int32_t> prev;
  std::vector<uint8_t> raw, file;
  std::string svg;
};

Scratch& scratch() {
  thread_local Scratch s;
  return s;
}

// png::deflate (src/png.rs:103-162): one final fixed-Huffman block; at each
// position the longest of up to 32 earlier occurrences of the same 3-byte hash
// within 32 KiB (the first one found among equals), taken when 3 bytes or more,
// else a literal; every position that starts three bytes is hashed in.
// Appends to `out`.
void deflate(const uint8_t* d, size_t n, std::vector<uint8_t>& out) {
  if (n > 0x7FFFFFFFu)  // chain positions are 31-bit
    throw std::length_error("deflate input over 2 GiB");
  const Codes& T = codes();
  Scratch& S = scratch();
  if (S.prev.size() < n) S.prev.resize(n);
  int32_t* head = S.head.data();
  int32_t* prev = S.prev.data();
  std::fill(S.head.begin(), S.head.end(), -1);
  size_t at = out.size();
  out.resize(at + n + n / 8 + 16);  // fixed codes spend at most 9 bits a byte
  BitOut b{out.data() + at};
  b.put(3, 3);  // BFINAL, then BTYPE 01
  const size_t last_start = n >= 3 ? n - 2 : 0;  // positions k with k + 3 <= n
  size_t i = 0;
  while (i < n) {
    size_t best_len = 0, best_dist = 0;
    if (i < last_start) {
      const size_t limit = std::min<size_t>(n - i, 258);
      int32_t cand = head[hash3(d + i)];
      for (unsigned chain = 0; cand >= 0 && chain < 32 && i - (size_t)cand <= 32768; ++chain) {
        const uint8_t* a = d + cand;
        // A candidate can only win by matching one byte past the best so far.
        if (a[best_len] == d[i + best_len]) {
          size_t l = common_prefix(a, d + i, limit);
          if (l > best_len) {
            best_len = l;
            best_dist = i - (size_t)cand;
            if (l == limit) break;
          }
        }
        cand = prev[cand];
      }
    }
    size_t adv;
    if (best_len >= 3) {
      unsigned li = T.len_idx[best_len];
      b.put(T.lit[257 + li], T.lit_bits[257 + li]);
      b.put((uint32_t)(best_len - T.lbase[li]), T.lextra[li]);
      unsigned di = T.dist_idx[best_dist];
      b.put(T.dcode[di], 5);
      b.put((uint32_t)(best_dist - T.dbase[di]), T.dextra[di]);
      adv = best_len;
    } else {
      b.put(T.lit[d[i]], T.lit_bits[d[i]]);
      adv = 1;
    }
    for (size_t k = i, e = std::min(i + adv, last_start); k < e; ++k) {
      const uint32_t h = hash3(d + k);
      prev[k] = head[h];
      head[h] = (int32_t)k;
    }
    i += adv;
  }
  b.put(T.lit[256], T.lit_bits[256]);
  b.finish();
  out.resize((size_t)(b.p - out.data()));
}

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
  uint8_t q[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
  o.insert(o.end(), q, q + 4);
}

// Length, type, body, CRC of type + body (src/png.rs:164-171). The body is
// out[from..] on entry, preceded by eight bytes reserved for length and type.
void seal_chunk(std::vector<uint8_t>& out, size_t from, const char type[4]) {
  uint32_t len = (uint32_t)(out.size() - from);
  uint8_t* h = out.data() + from - 8;
  h[0] = (uint8_t)(len >> 24), h[1] = (uint8_t)(len >> 16), h[2] = (uint8_t)(len >> 8), h[3] = (uint8_t)len;
  std::memcpy(h + 4, type, 4);
  put_be32(out, crc32(h + 4, len + 4));
}

// png::write_rgb (src/png.rs:174-197) given the filtered scanlines (a zero
// filter byte before each row of w * 3 bytes); the file's bytes go to `png`.
void png_from_scanlines(size_t w, size_t h, const std::vector<uint8_t>& raw,
                        std::vector<uint8_t>& png) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  png.assign(sig, sig + 8);
  png.resize(png.size() + 8);
  size_t body = png.size();
  put_be32(png, (uint32_t)w);
  put_be32(png, (uint32_t)h);
  for (uint8_t v : {8, 2, 0, 0, 0}) png.push_back(v);  // 8-bit RGB, no interlace
  seal_chunk(png, body, "IHDR");
  png.resize(png.size() + 8);
  body = png.size();
  png.push_back(0x78);  // zlib: deflate, 32 KiB window, no dictionary
  png.push_back(0x01);
  deflate(raw.data(), raw.size(), png);
  put_be32(png, adler32(raw.data(), raw.size()));
  seal_chunk(png, body, "IDAT");
  png.resize(png.size() + 8);
  seal_chunk(png, png.size(), "IEND");
}

int write_file(const char* path, const void* data, size_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(LG_EIO, std::string(path) + ": " + std::strerror(errno));
  size_t put = n ? std::fwrite(data, 1, n, f) : 0;
  int err = put == n ? 0 : errno;
  if (std::fclose(f) != 0 && !err) err = errno;
  if (err || put != n) return fail(LG_EIO, std::string(path) + ": " + std::strerror(err ? err : EIO));
  return LG_OK;
}

// ---- field products ----------------------------------------------------------

void vorticity(const float* ux, const float* uy, size_t bw, size_t bh, size_t bx, size_t by,
               float* w) {
  const float dx = (float)bx;
  const float dy = (float)by * kSqrt3_2;
  const float two_dy = 2.0f * dy;
  for (size_t j = 0; j < bh; ++j) {
    const size_t jp = j + 1 < bh ? j + 1 : 0;
    const size_t jm = j == 0 ? bh - 1 : j - 1;
    for (size_t i = 0; i < bw; ++i) {
      const size_t ip = std::min(i + 1, bw - 1);
      const size_t im = i == 0 ? 0 : i - 1;
      const float dvdx = (uy[j * bw + ip] - uy[j * bw + im]) / ((float)(ip - im) * dx);
      const float dudy = (ux[jp * bw + i] - ux[jm * bw + i]) / two_dy;
      w[j * bw + i] = dvdx - dudy;
    }
  }
}

// write_frame's colour scale (src/main.rs:219-229): element len * 98 / 100 of
// the ascending |w| over blocks below the solid threshold, at least 1e-5, and
// 1e-3 when there are none. nth_element finds the same element as the sort.
int vorticity_scale(const float* w, const float* solid, size_t n, float* out) {
  std::vector<float> v;
  v.reserve(n);
  for (size_t k = 0; k < n; ++k)
    if (solid[k] < 0.12f) {
      float a = std::fabs(w[k]);
      if (std::isnan(a))
        return fail(LG_EINVAL, "vorticity is NaN (a field one block wide has no x difference)");
      v.push_back(a);
    }
  size_t idx = v.size() * 98 / 100;
  float x = 1e-3f;
  if (idx < v.size()) {
    std::nth_element(v.begin(), v.begin() + (ptrdiff_t)idx, v.end());
    x = v[idx];
  }
  *out = std::fmax(x, 1e-5f);
  return LG_OK;
}

// Rust's `as u8` from f32: saturating, NaN to 0.
inline uint8_t sat_u8(float x) {
  if (!(x > 0.0f)) return 0;
  return x >= 255.0f ? 255 : (uint8_t)x;
}

// diverging (src/render.rs:159-173). f32::clamp keeps NaN, which then takes
// the negative branch.
inline void diverging(float t, uint8_t* c) {
  if (t < -1.0f) t = -1.0f;
  if (t > 1.0f) t = 1.0f;
  float r, g, b;
  if (t >= 0.0f) r = 1.0f, g = 1.0f - 0.75f * t, b = 1.0f - 0.9f * t;
  else r = 1.0f + 0.9f * t, g = 1.0f + 0.6f * t, b = 1.0f;
  c[0] = sat_u8(r * 255.0f);
  c[1] = sat_u8(g * 255.0f);
  c[2] = sat_u8(b * 255.0f);
}

// Pixel rows per block row: (px * SQRT3_2 * by / bx).round() as usize, at least 1.
size_t rows_per_block(size_t px, size_t bx, size_t by) {
  float r = std::round((float)px * kSqrt3_2 * (float)by / (float)bx);
  size_t v = (std::isnan(r) || r <= 0.0f) ? 0 : (size_t)r;
  return std::max<size_t>(v, 1);
}

// render::write_vorticity (src/render.rs:176-204) given the vorticity.
int vorticity_png(const char* path, const float* vort, const float* solid, size_t bw, size_t bh,
                  size_t bx, size_t by, size_t scale, float wmax) {
  const size_t px = std::max<size_t>(scale, 1);
  const size_t py = rows_per_block(px, bx, by);
  const size_t iw = bw * px, ih = bh * py;
  const size_t stride = 1 + iw * 3;
  Scratch& S = scratch();
  std::vector<uint8_t>& raw = S.raw;
  raw.resize(ih * stride);
  // Image rows run downward and lattice rows upward: block row j fills image
  // rows ih - 1 - (j * py + dy).
  for (size_t j = 0; j < bh; ++j) {
    uint8_t* row = raw.data() + (ih - 1 - j * py) * stride;
    row[0] = 0;
    for (size_t i = 0; i < bw; ++i) {
      const size_t k = j * bw + i;
      uint8_t c[3] = {30, 30, 35};
      if (!(solid[k] > kSolid)) diverging(vort[k] / wmax, c);
      uint8_t* p = row + 1 + i * px * 3;
      for (size_t x = 0; x < px; ++x, p += 3) p[0] = c[0], p[1] = c[1], p[2] = c[2];
    }
    for (size_t dy = 1; dy < py; ++dy) std::memcpy(row - dy * stride, row, stride);
  }
  png_from_scanlines(iw, ih, raw, S.file);
  return write_file(path, S.file.data(), S.file.size());
}

// Rust spells the non-finite f32 values "NaN", "inf" and "-inf".
char* put_nonfinite(char* p, float x) {
  const char* t = std::isnan(x) ? "NaN" : x < 0 ? "-inf" : "inf";
  size_t n = std::strlen(t);
  std::memcpy(p, t, n);
  return p + n;
}

// `{:.1}` of an f32: the exact binary value rounded to one decimal, ties to
// even, with a sign whenever the sign bit is set ("-0.0").
char* put_tenths(char* p, float x) {
  if (!std::isfinite(x)) return put_nonfinite(p, x);
  const double v = (double)x * 10.0;  // exact: 24 + 4 significant bits
  if (!(std::fabs(v) < 9.0e15)) return p + std::sprintf(p, "%.1f", (double)x);
  const double fl = std::floor(v);
  const double frac = v - fl;  // exact
  int64_t r = (int64_t)fl;
  if (frac > 0.5 || (frac == 0.5 && (r & 1))) ++r;
  if (std::signbit(x)) *p++ = '-';
  const uint64_t a = (uint64_t)(r < 0 ? -r : r);
  p = std::to_chars(p, p + 24, a / 10).ptr;
  *p++ = '.';
  *p++ = (char)('0' + a % 10);
  return p;
}

// `{}` of an f32: the shortest digits that read back to it, written out in
// plain decimal, zeros filling out to the point, never an exponent.
char* put_shortest(char* p, float x) {
  if (!std::isfinite(x)) return put_nonfinite(p, x);
  if (std::signbit(x)) *p++ = '-';
  if (x == 0.0f) return *p++ = '0', p;
  char sci[48];
  char* end = std::to_chars(sci, sci + sizeof sci, std::fabs(x), std::chars_format::scientific).ptr;
  *end = 0;
  char* e = std::strchr(sci, 'e');
  char digits[16];
  int nd = 0;
  for (char* q = sci; q < e; ++q)
    if (*q != '.') digits[nd++] = *q;
  const int point = std::atoi(e + 1) + 1;  // digits before the decimal point
  if (point <= 0) {
    *p++ = '0', *p++ = '.';
    for (int z = 0; z < -point; ++z) *p++ = '0';
    std::memcpy(p, digits, (size_t)nd);
    return p + nd;
  }
  for (int k = 0; k < std::max(point, nd); ++k) {
    if (k == point) *p++ = '.';
    *p++ = k < nd ? digits[k] : '0';
  }
  return p;
}

// `as i32` from f32: saturating, NaN to 0.
int32_t sat_i32(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}

void append(std::string& s, const char* a) { s.append(a); }

// render::write_arrows (src/render.rs:209-267), into s.
void arrows_svg(std::string& s, const float* ux, const float* uy, const float* solid, size_t bw,
                size_t bh, size_t bx, size_t by, float scale, float fx, float fy) {
  const float w = (float)bw * (float)bx;
  const float h = (float)bh * (float)by * kSqrt3_2;
  s.clear();
  s.reserve(bw * bh * 72 + 512);
  char buf[1024];  // ten numbers of at most 42 characters and the markup
  char* p = buf;
  p += std::sprintf(p, "<svg xmlns=\"http://www.w3.org/2000/svg\" viewBox=\"0 0 ");
  p = put_shortest(p, w);
  *p++ = ' ';
  p = put_shortest(p, h);
  p += std::sprintf(p, "\" width=\"%d\" height=\"%d\">\n", (int)sat_i32(std::round(w)),
                    (int)sat_i32(std::round(h)));
  s.append(buf, (size_t)(p - buf));
  append(s, "<rect width=\"100%\" height=\"100%\" fill=\"#ffffff\"/>\n"
            "<g stroke=\"#1a1a1a\" stroke-width=\"1.1\" fill=\"none\" stroke-linecap=\"round\">\n");
  const float arrow = (float)bx * scale;
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i) {
      const size_t k = j * bw + i;
      if (solid[k] > kSolid) continue;
      const float cx = ((float)i + 0.5f) * (float)bx;
      const float cy = h - ((float)j + 0.5f) * (float)by * kSqrt3_2;
      const float vx = (ux[k] - fx) * arrow;
      const float vy = -(uy[k] - fy) * arrow;
      const float len = std::sqrt(vx * vx + vy * vy);
      if (len < 0.35f) continue;
      const float x0 = cx - vx * 0.5f, y0 = cy - vy * 0.5f;
      const float x1 = cx + vx * 0.5f, y1 = cy + vy * 0.5f;
      const float hx = vx / len, hy = vy / len;
      const float head = std::fmin(len * 0.35f, (float)bx * 0.4f);  // f32::min keeps the number
      p = buf;
      std::memcpy(p, "<path d=\"M", 10), p += 10;
      p = put_tenths(p, x0), *p++ = ' ', p = put_tenths(p, y0), *p++ = 'L';
      p = put_tenths(p, x1), *p++ = ' ', p = put_tenths(p, y1), *p++ = 'm';
      p = put_tenths(p, -head * (hx + hy * 0.5f)), *p++ = ' ';
      p = put_tenths(p, -head * (hy - hx * 0.5f)), *p++ = 'L';
      p = put_tenths(p, x1), *p++ = ' ', p = put_tenths(p, y1), *p++ = 'l';
      p = put_tenths(p, -head * (hx - hy * 0.5f)), *p++ = ' ';
      p = put_tenths(p, -head * (hy + hx * 0.5f));
      std::memcpy(p, "\"/>\n", 4), p += 4;
      s.append(buf, (size_t)(p - buf));
    }
  append(s, "</g>\n<g fill=\"#1a1a1a\">\n");
  for (size_t j = 0; j < bh; ++j)
    for (size_t i = 0; i < bw; ++i)
      if (solid[j * bw + i] > kSolid) {
        const float x = (float)i * (float)bx;
        const float y = h - (float)(j + 1) * (float)by * kSqrt3_2;
        p = buf;
        std::memcpy(p, "<rect x=\"", 9), p += 9;
        p = put_tenths(p, x);
        std::memcpy(p, "\" y=\"", 5), p += 5;
        p = put_tenths(p, y);
        std::memcpy(p, "\" width=\"", 9), p += 9;
        p = put_tenths(p, (float)bx);
        std::memcpy(p, "\" height=\"", 10), p += 10;
        p = put_tenths(p, (float)by * kSqrt3_2);
        std::memcpy(p, "\"/>\n", 4), p += 4;
        s.append(buf, (size_t)(p - buf));
      }
  append(s, "</g>\n</svg>\n");
}

int check_field(const float* ux, const float* uy, const float* solid, size_t bw, size_t bh,
                size_t bx, size_t by) {
  if ((!ux || !uy || !solid) && bw && bh) return fail(LG_EINVAL, "null field array");
  if (bx == 0 || by == 0) return fail(LG_EINVAL, "block sides must be positive");
  return LG_OK;
}

// Runs f, turning a C++ exception (out of host memory, a thread that cannot
// start) into an error return: nothing thrown crosses the C ABI.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(LG_EINVAL, "out of host memory");
  } catch (const std::exception& e) {
    return fail(LG_EINVAL, e.what());
  }
}

}  // namespace

// ---- the writer pool (src/main.rs:196-281) ------------------------------------

struct lg_frames {
  struct Job {
    std::vector<float> ux, uy, solid;
    size_t bw = 0, bh = 0, bx = 0, by = 0, scale = 0;
    float arrow_scale = 0, mean_x = 0, mean_y = 0;
    std::string png, svg;
  };
  std::mutex m;
  std::condition_variable room, work;
  std::deque<std::unique_ptr<Job>> queue;
  size_t bound = 1;
  bool closing = false;
  std::string failure;  // the first failed frame, "" while none has
  std::vector<std::thread> workers;

  // write_frame (src/main.rs:220-234); the vorticity is computed once for
  // both the colour scale and the map.
  static int write_frame(const Job& f) {
    const size_t n = f.bw * f.bh;
    std::vector<float> vort(n);
    vorticity(f.ux.data(), f.uy.data(), f.bw, f.bh, f.bx, f.by, vort.data());
    float wmax;
    int rc = vorticity_scale(vort.data(), f.solid.data(), n, &wmax);
    if (rc == LG_OK)
      rc = vorticity_png(f.png.c_str(), vort.data(), f.solid.data(), f.bw, f.bh, f.bx, f.by, f.scale, wmax);
    if (rc == LG_OK) {
      std::string& s = scratch().svg;
      arrows_svg(s, f.ux.data(), f.uy.data(), f.solid.data(), f.bw, f.bh, f.bx, f.by,
                 f.arrow_scale, f.mean_x, f.mean_y);
      rc = write_file(f.svg.c_str(), s.data(), s.size());
    }
    return rc;
  }

  void run() {
    for (;;) {
      std::unique_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lk(m);
        work.wait(lk, [&] { return closing || !queue.empty(); });
        if (queue.empty()) return;
        job = std::move(queue.front());
        queue.pop_front();
      }
      room.notify_one();
      if (guarded([&] { return write_frame(*job); }) != LG_OK) {
        // Strip the path write_file put in front: the message names the PNG,
        // as the reference's does, whichever of the two files failed.
        std::string msg = lgca_internal::g_err;
        for (const std::string* pth : {&job->png, &job->svg})
          if (msg.compare(0, pth->size() + 2, *pth + ": ") == 0) msg.erase(0, pth->size() + 2);
        std::lock_guard<std::mutex> lk(m);
        if (failure.empty()) failure = job->png + ": " + msg;
      }
    }
  }

  int status() {
    std::lock_guard<std::mutex> lk(m);
    return failure.empty() ? LG_OK : fail(LG_EIO, failure);
  }

  int submit(std::unique_ptr<Job> job) {
    {
      std::unique_lock<std::mutex> lk(m);
      room.wait(lk, [&] { return queue.size() < bound; });
      queue.push_back(std::move(job));
    }
    work.notify_one();
    return status();
  }
};

extern "C" {

int lg_field_vorticity(const float* ux, const float* uy, size_t bw, size_t bh, size_t bx,
                       size_t by, float* out) {
  static const float zero = 0.0f;
  if (int rc = check_field(ux, uy, &zero, bw, bh, bx, by)) return rc;
  if (!out && bw && bh) return fail(LG_EINVAL, "null output");
  vorticity(ux, uy, bw, bh, bx, by, out);
  return LG_OK;
}

int lg_field_mean_velocity(const float* ux, const float* uy, const float* solid, size_t n,
                           float* mx, float* my) {
  if ((!ux || !uy || !solid) && n) return fail(LG_EINVAL, "null field array");
  if (!mx || !my) return fail(LG_EINVAL, "null output");
  float sx = 0.0f, sy = 0.0f, c = 0.0f;
  for (size_t k = 0; k < n; ++k)
    if (solid[k] < 0.5f) sx += ux[k], sy += uy[k], c += 1.0f;
  *mx = c == 0.0f ? 0.0f : sx / c;
  *my = c == 0.0f ? 0.0f : sy / c;
  return LG_OK;
}

int lg_vorticity_scale(const float* vort, const float* solid, size_t n, float* wmax) {
  if ((!vort || !solid) && n) return fail(LG_EINVAL, "null array");
  if (!wmax) return fail(LG_EINVAL, "null output");
  return guarded([&] { return vorticity_scale(vort, solid, n, wmax); });
}

int lg_png_deflate(const uint8_t* data, size_t n, uint8_t* out, size_t cap, size_t* len) {
  if ((!data && n) || !len) return fail(LG_EINVAL, "null argument");
  return guarded([&] {
    std::vector<uint8_t> z;
    deflate(data, n, z);
    *len = z.size();
    if (!out) return LG_OK;
    if (cap < z.size())
      return fail(LG_EINVAL, "output buffer too small (" + std::to_string(z.size()) + " bytes needed)");
    std::memcpy(out, z.data(), z.size());
    return LG_OK;
  });
}

int lg_write_png_rgb(const char* path, size_t w, size_t h, const uint8_t* rgb) {
  if (!path || (!rgb && w && h)) return fail(LG_EINVAL, "null argument");
  if (w > 0x7FFFFFFFu || h > 0x7FFFFFFFu) return fail(LG_EINVAL, "image too large for PNG");
  return guarded([&] {
    const size_t stride = 1 + w * 3;
    Scratch& S = scratch();
    S.raw.resize(h * stride);
    for (size_t y = 0; y < h; ++y) {
      S.raw[y * stride] = 0;
      if (w) std::memcpy(S.raw.data() + y * stride + 1, rgb + y * w * 3, w * 3);
    }
    png_from_scanlines(w, h, S.raw, S.file);
    return write_file(path, S.file.data(), S.file.size());
  });
}

int lg_write_vorticity(const char* path, const float* ux, const float* uy, const float* solid,
                       size_t bw, size_t bh, size_t bx, size_t by, size_t scale, float wmax) {
  if (!path) return fail(LG_EINVAL, "null path");
  if (int rc = check_field(ux, uy, solid, bw, bh, bx, by)) return rc;
  return guarded([&] {
    std::vector<float> vort(bw * bh);
    vorticity(ux, uy, bw, bh, bx, by, vort.data());
    return vorticity_png(path, vort.data(), solid, bw, bh, bx, by, scale, wmax);
  });
}

int lg_write_arrows(const char* path, const float* ux, const float* uy, const float* solid,
                    size_t bw, size_t bh, size_t bx, size_t by, float scale, float frame_x,
                    float frame_y) {
  if (!path) return fail(LG_EINVAL, "null path");
  if (int rc = check_field(ux, uy, solid, bw, bh, bx, by)) return rc;
  return guarded([&] {
    std::string& s = scratch().svg;
    arrows_svg(s, ux, uy, solid, bw, bh, bx, by, scale, frame_x, frame_y);
    return write_file(path, s.data(), s.size());
  });
}

lg_frames_t* lg_frames_create(size_t workers) {
  if (workers == 0 || workers > 256) {
    fail(LG_EINVAL, "writer count must be 1..256");
    return nullptr;
  }
  lg_frames* fr = new (std::nothrow) lg_frames;
  if (!fr) {
    fail(LG_EINVAL, "out of host memory");
    return nullptr;
  }
  fr->bound = workers;
  try {
    for (size_t i = 0; i < workers; ++i) fr->workers.emplace_back([fr] { fr->run(); });
  } catch (const std::exception& e) {
    {
      std::lock_guard<std::mutex> lk(fr->m);
      fr->closing = true;
    }
    fr->work.notify_all();
    for (auto& t : fr->workers) t.join();
    delete fr;
    fail(LG_EINVAL, std::string("cannot start writer threads: ") + e.what());
    return nullptr;
  }
  return fr;
}

int lg_frames_write(lg_frames_t* fr, const float* ux, const float* uy, const float* solid,
                    size_t bw, size_t bh, size_t bx, size_t by, const char* png, const char* svg,
                    size_t scale, float arrow_scale, float mean_x, float mean_y) {
  if (!fr || !png || !svg) return fail(LG_EINVAL, "null argument");
  if (int rc = check_field(ux, uy, solid, bw, bh, bx, by)) return rc;
  return guarded([&] {
    auto job = std::make_unique<lg_frames::Job>();
    const size_t n = bw * bh;
    job->ux.assign(ux, ux + n);
    job->uy.assign(uy, uy + n);
    job->solid.assign(solid, solid + n);
    job->bw = bw, job->bh = bh, job->bx = bx, job->by = by, job->scale = scale;
    job->arrow_scale = arrow_scale, job->mean_x = mean_x, job->mean_y = mean_y;
    job->png = png, job->svg = svg;
    return fr->submit(std::move(job));
  });
}

int lg_frames_write_lattice(lg_frames_t* fr, lg_t* g, const char* png, const char* svg,
                            size_t scale, float arrow_scale) {
  if (!fr || !g || !png || !svg) return fail(LG_EINVAL, "null argument");
  size_t bw, bh, bx, by;
  if (int rc = lg_field_dims(g, &bw, &bh, &bx, &by)) return rc;
  return guarded([&] {
    auto job = std::make_unique<lg_frames::Job>();
    const size_t n = bw * bh;
    job->ux.resize(n), job->uy.resize(n), job->solid.resize(n);
    std::vector<float> rho(n);
    if (int rc = lg_read_field(g, job->ux.data(), job->uy.data(), rho.data(), job->solid.data(), n))
      return rc;
    lg_field_mean_velocity(job->ux.data(), job->uy.data(), job->solid.data(), n, &job->mean_x,
                           &job->mean_y);
    job->bw = bw, job->bh = bh, job->bx = bx, job->by = by, job->scale = scale;
    job->arrow_scale = arrow_scale;
    job->png = png, job->svg = svg;
    return fr->submit(std::move(job));
  });
}

int lg_frames_finish(lg_frames_t* fr) {
  if (!fr) return fail(LG_EINVAL, "null frames");
  {
    std::lock_guard<std::mutex> lk(fr->m);
    fr->closing = true;
  }
  fr->work.notify_all();
  for (auto& t : fr->workers) t.join();
  int rc = fr->status();
  delete fr;
  return rc;
}

}  // extern "C"
