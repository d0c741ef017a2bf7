// Native glTF composite import for gbr_add_file().  See gbr_import.h for the
// conventions; float arithmetic is float32 in the same order as the
// renderer's Python importer (habitat_sim_amd/assets.py), so both produce
// identical file descriptions.
#include "gbr_import.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <utility>

#pragma STDC FP_CONTRACT OFF

namespace gbr {
namespace imp {

namespace {

// ---------------------------------------------------------------------------
// JSON (RFC 8259), enough for glTF: objects keep member order
// ---------------------------------------------------------------------------

struct Json {
  enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<Json> a;
  std::vector<std::pair<std::string, Json>> o;

  const Json* get(const char* key) const {
    if (type != Obj) return nullptr;
    for (const auto& kv : o)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  bool has(const char* key) const { return get(key) != nullptr; }
  size_t size() const { return type == Arr ? a.size() : 0; }
  const Json& operator[](size_t i) const { return a[i]; }
};

struct JsonParser {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool fail(const char* what) {
    if (err.empty()) err = std::string("JSON: ") + what;
    return false;
  }
  static void utf8(uint32_t c, std::string& out) {
    if (c < 0x80) {
      out += (char)c;
    } else if (c < 0x800) {
      out += (char)(0xC0 | (c >> 6));
      out += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      out += (char)(0xE0 | (c >> 12));
      out += (char)(0x80 | ((c >> 6) & 0x3F));
      out += (char)(0x80 | (c & 0x3F));
    } else {
      out += (char)(0xF0 | (c >> 18));
      out += (char)(0x80 | ((c >> 12) & 0x3F));
      out += (char)(0x80 | ((c >> 6) & 0x3F));
      out += (char)(0x80 | (c & 0x3F));
    }
  }
  bool hex4(uint32_t& v) {
    if (end - p < 4) return fail("short \\u escape");
    v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string& out) {
    if (p >= end || *p != '"') return fail("expected a string");
    ++p;
    while (p < end && *p != '"') {
      if (*p != '\\') {
        out += *p++;
        continue;
      }
      if (++p >= end) return fail("unterminated escape");
      const char c = *p++;
      switch (c) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t v;
          if (!hex4(v)) return false;
          if (v >= 0xD800 && v < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
            p += 2;
            uint32_t lo;
            if (!hex4(lo)) return false;
            v = 0x10000 + ((v - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(v, out);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(Json& v, int depth) {
    if (depth > 256) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end");
    const char c = *p;
    if (c == '{') {
      v.type = Json::Obj;
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      while (true) {
        ws();
        std::string key;
        if (!string(key)) return false;
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        ++p;
        v.o.emplace_back(std::move(key), Json());
        if (!value(v.o.back().second, depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.type = Json::Arr;
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      while (true) {
        v.a.emplace_back();
        if (!value(v.a.back(), depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = Json::Str;
      return string(v.s);
    }
    if (end - p >= 4 && !std::strncmp(p, "true", 4)) {
      v.type = Json::Bool;
      v.b = true;
      p += 4;
      return true;
    }
    if (end - p >= 5 && !std::strncmp(p, "false", 5)) {
      v.type = Json::Bool;
      p += 5;
      return true;
    }
    if (end - p >= 4 && !std::strncmp(p, "null", 4)) {
      p += 4;
      return true;
    }
    // number
    const char* s = p;
    if (p < end && (*p == '-' || *p == '+')) ++p;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                       *p == '-' || *p == '+'))
      ++p;
    if (p == s) return fail("unexpected character");
    std::string num(s, p);
    char* e = nullptr;
    v.type = Json::Num;
    v.n = std::strtod(num.c_str(), &e);
    if (!e || *e) return fail("bad number");
    return true;
  }
};

double num(const Json* j, double dflt) { return j && j->type == Json::Num ? j->n : dflt; }
int64_t inum(const Json* j, int64_t dflt) {
  return j && j->type == Json::Num ? (int64_t)j->n : dflt;
}

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg(0, std::ios::end);
  const std::streamoff n = f.tellg();
  if (n < 0) return false;
  f.seekg(0, std::ios::beg);
  out.resize((size_t)n);
  if (n > 0) f.read(reinterpret_cast<char*>(out.data()), n);
  return (bool)f || f.eof();
}

uint32_t rd32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }
uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

bool base64(const std::string& in, std::vector<uint8_t>& out) {
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
  };
  uint32_t acc = 0;
  int bits = 0;
  for (char c : in) {
    if (c == '=' || c == '\n' || c == '\r' || c == ' ') continue;
    const int v = val(c);
    if (v < 0) return false;
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((uint8_t)((acc >> bits) & 0xFF));
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Images
// ---------------------------------------------------------------------------

// (..., C) -> RGBA with alpha 255 where missing; 2 channels = luminance-alpha
void expand_rgba(const uint8_t* src, size_t n, int comps, uint8_t* dst) {
  for (size_t i = 0; i < n; ++i) {
    const uint8_t* s = src + i * comps;
    uint8_t* d = dst + i * 4;
    switch (comps) {
      case 1: d[0] = d[1] = d[2] = s[0]; d[3] = 255; break;
      case 2: d[0] = d[1] = d[2] = s[0]; d[3] = s[1]; break;
      case 3: d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = 255; break;
      default: std::memcpy(d, s, 4); break;
    }
  }
}

void flip_rows(std::vector<uint8_t>& img, uint32_t w, uint32_t h, uint32_t layers) {
  const size_t row = (size_t)w * 4;
  std::vector<uint8_t> tmp(row);
  for (uint32_t l = 0; l < layers; ++l) {
    uint8_t* base = img.data() + (size_t)l * row * h;
    for (uint32_t y = 0; y < h / 2; ++y) {
      uint8_t* a = base + (size_t)y * row;
      uint8_t* b = base + (size_t)(h - 1 - y) * row;
      std::memcpy(tmp.data(), a, row);
      std::memcpy(a, b, row);
      std::memcpy(b, tmp.data(), row);
    }
  }
}

void rgb565(uint32_t c, int* out) {
  const int r = (c >> 11) & 31, g = (c >> 5) & 63, b = c & 31;
  out[0] = (r << 3) | (r >> 2);
  out[1] = (g << 2) | (g >> 4);
  out[2] = (b << 3) | (b >> 2);
}

// BC1 blocks -> RGBA, block row 0 = image row 0
void decode_bc1(const uint8_t* data, uint32_t w, uint32_t h, bool alpha, uint8_t* out) {
  const uint32_t bw = (w + 3) / 4, bh = (h + 3) / 4;
  for (uint32_t by = 0; by < bh; ++by)
    for (uint32_t bx = 0; bx < bw; ++bx) {
      const uint8_t* blk = data + ((size_t)by * bw + bx) * 8;
      const uint32_t c0 = blk[0] | (blk[1] << 8), c1 = blk[2] | (blk[3] << 8);
      const uint32_t bits = rd32(blk + 4);
      int p[4][4];
      rgb565(c0, p[0]);
      rgb565(c1, p[1]);
      p[0][3] = p[1][3] = 255;
      for (int k = 0; k < 3; ++k) {
        if (c0 > c1) {
          p[2][k] = (2 * p[0][k] + p[1][k] + 1) / 3;
          p[3][k] = (p[0][k] + 2 * p[1][k] + 1) / 3;
        } else {
          p[2][k] = (p[0][k] + p[1][k]) / 2;
          p[3][k] = 0;
        }
      }
      p[2][3] = 255;
      p[3][3] = (c0 > c1 || !alpha) ? 255 : 0;
      for (uint32_t py = 0; py < 4; ++py)
        for (uint32_t px = 0; px < 4; ++px) {
          const uint32_t x = bx * 4 + px, y = by * 4 + py;
          if (x >= w || y >= h) continue;
          const int idx = (bits >> (2 * (py * 4 + px))) & 3;
          uint8_t* d = out + ((size_t)y * w + x) * 4;
          for (int k = 0; k < 4; ++k) d[k] = (uint8_t)p[idx][k];
        }
    }
}

const uint8_t kKtx2Magic[12] = {0xAB, 'K', 'T', 'X', ' ', '2', '0', 0xBB, '\r', '\n', 0x1A, '\n'};

bool decode_ktx2(const uint8_t* d, size_t n, Texture& t, uint32_t& layers_field, std::string& err) {
  if (n < 80 || std::memcmp(d, kKtx2Magic, 12)) return (err = "not a KTX2 file"), false;
  const uint32_t vk = rd32(d + 12), w = rd32(d + 20), h = rd32(d + 24), depth = rd32(d + 28);
  const uint32_t layers = rd32(d + 32), faces = rd32(d + 36), levels = rd32(d + 40);
  const uint32_t sc = rd32(d + 44);
  if (sc != 0) return (err = "supercompressed KTX2 is not supported"), false;
  if (depth > 1 || faces != 1) return (err = "only 2D / 2D-array KTX2 images are supported"), false;
  int comps = 0;
  bool bc1 = false, bc1a = false;
  switch (vk) {
    case 9: comps = 1; break;
    case 16: comps = 2; break;
    case 23: comps = 3; break;
    case 37: comps = 4; break;
    case 131: case 132: bc1 = true; break;
    case 133: case 134: bc1 = bc1a = true; break;
    default: return (err = "unsupported KTX2 vkFormat " + std::to_string(vk)), false;
  }
  const uint32_t kvd_off = rd32(d + 56), kvd_len = rd32(d + 60);
  std::string orientation = "ru";
  if ((size_t)kvd_off + kvd_len <= n) {
    size_t q = 0;
    const uint8_t* kv = d + kvd_off;
    while (q + 4 <= kvd_len) {
      const uint32_t len = rd32(kv + q);
      if (q + 4 + len > kvd_len) break;
      const char* e = reinterpret_cast<const char*>(kv + q + 4);
      const size_t klen = strnlen(e, len);
      if (std::string(e, klen) == "KTXorientation" && klen + 1 < len) {
        const char* v = e + klen + 1;
        orientation = std::string(v, strnlen(v, len - klen - 1));
      }
      q += 4 + ((len + 3) & ~3u);
    }
  }
  const uint32_t nl = layers ? layers : 1;
  const uint32_t nlev = levels ? levels : 1;
  if (80 + 24 * (size_t)nlev > n) return (err = "truncated KTX2 level index"), false;
  t.width = w;
  t.height = h;
  t.layers = nl;
  t.compressed = bc1;
  t.levels.clear();
  for (uint32_t lv = 0; lv < nlev; ++lv) {
    const uint64_t off = rd64(d + 80 + 24 * lv), len = rd64(d + 88 + 24 * lv);
    if (off + len > n) return (err = "truncated KTX2 level"), false;
    const uint32_t lw = std::max(w >> lv, 1u), lh = std::max(h >> lv, 1u);
    std::vector<uint8_t> img((size_t)nl * lw * lh * 4);
    if (!bc1) {
      if (len < (uint64_t)nl * lw * lh * comps) return (err = "short KTX2 level"), false;
      expand_rgba(d + off, (size_t)nl * lw * lh, comps, img.data());
      if (orientation.size() > 1 && orientation[1] == 'd') flip_rows(img, lw, lh, nl);
    } else {
      const size_t bsz = (size_t)((lw + 3) / 4) * ((lh + 3) / 4) * 8;
      if (len < bsz * nl) return (err = "short KTX2 level"), false;
      for (uint32_t l = 0; l < nl; ++l)
        decode_bc1(d + off + l * bsz, lw, lh, bc1a, img.data() + (size_t)l * lw * lh * 4);
    }
    t.levels.push_back(std::move(img));
  }
  layers_field = layers;
  return true;
}

}  // namespace

// PNG: IHDR/PLTE/tRNS/IDAT/IEND, zlib stream, the five filters, Adam7;
// greyscale / RGB / palette / grey+alpha / RGBA at 1-16 bits per sample
// (16-bit samples keep their high byte, low-depth grey scales to 0..255)
bool decode_png(const uint8_t* d, size_t n, uint32_t& w, uint32_t& h, std::vector<uint8_t>& rgba,
                std::string& err) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  if (n < 8 || std::memcmp(d, sig, 8)) return (err = "not a PNG file"), false;
  size_t p = 8;
  uint8_t depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  w = h = 0;
  while (p + 12 <= n) {
    const uint32_t len = be32(d + p);
    const uint8_t* type = d + p + 4;
    const uint8_t* body = d + p + 8;
    if (p + 12 + (size_t)len > n) return (err = "truncated PNG chunk"), false;
    if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
      w = be32(body);
      h = be32(body + 4);
      depth = body[8];
      ctype = body[9];
      interlace = body[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(body, body + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(body, body + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), body, body + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    p += 12 + len;
  }
  int ch;
  switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: return (err = "bad PNG colour type"), false;
  }
  if (!w || !h || (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16))
    return (err = "bad PNG header"), false;
  if (ctype == 3 && plte.empty()) return (err = "PNG palette missing"), false;
  const uint32_t bpp_bits = (uint32_t)ch * depth;
  const uint32_t bpp = std::max(1u, bpp_bits / 8);  // filter byte distance
  // pass geometry (Adam7 or one pass)
  static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
  const int passes = interlace ? 7 : 1;
  size_t raw_size = 0;
  for (int ps = 0; ps < passes; ++ps) {
    const uint32_t pw = interlace ? (w + adx[ps] - 1 - ax0[ps]) / adx[ps] : w;
    const uint32_t ph = interlace ? (h + ady[ps] - 1 - ay0[ps]) / ady[ps] : h;
    if (pw && ph) raw_size += (size_t)ph * (1 + ((size_t)pw * bpp_bits + 7) / 8);
  }
  std::vector<uint8_t> raw(raw_size);
  uLongf out_len = (uLongf)raw_size;
  if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK ||
      out_len != raw_size)
    return (err = "PNG data does not inflate to the image size"), false;
  rgba.assign((size_t)w * h * 4, 0);
  size_t q = 0;
  for (int ps = 0; ps < passes; ++ps) {
    const uint32_t pw = interlace ? (w + adx[ps] - 1 - ax0[ps]) / adx[ps] : w;
    const uint32_t ph = interlace ? (h + ady[ps] - 1 - ay0[ps]) / ady[ps] : h;
    if (!pw || !ph) continue;
    const size_t stride = ((size_t)pw * bpp_bits + 7) / 8;
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    for (uint32_t y = 0; y < ph; ++y) {
      const uint8_t f = raw[q++];
      for (size_t i = 0; i < stride; ++i) {
        const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
        int x = raw[q + i];
        switch (f) {
          case 0: break;
          case 1: x += a; break;
          case 2: x += b; break;
          case 3: x += (a + b) / 2; break;
          case 4: {
            const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
            x += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
            break;
          }
          default: return (err = "bad PNG filter"), false;
        }
        cur[i] = (uint8_t)x;
      }
      q += stride;
      for (uint32_t x = 0; x < pw; ++x) {
        uint32_t s[4] = {0, 0, 0, 0};
        for (int k = 0; k < ch; ++k) {
          const size_t bit = ((size_t)x * ch + k) * depth;
          uint32_t v;
          if (depth == 16) v = cur[bit / 8];  // high byte
          else if (depth == 8) v = cur[bit / 8];
          else v = (cur[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
          s[k] = v;
        }
        uint8_t* o;
        const uint32_t ix = interlace ? ax0[ps] + x * adx[ps] : x;
        const uint32_t iy = interlace ? ay0[ps] + y * ady[ps] : y;
        o = rgba.data() + ((size_t)iy * w + ix) * 4;
        if (ctype == 3) {
          const uint32_t i = s[0];
          if ((size_t)i * 3 + 2 >= plte.size()) return (err = "PNG palette index out of range"), false;
          o[0] = plte[i * 3];
          o[1] = plte[i * 3 + 1];
          o[2] = plte[i * 3 + 2];
          o[3] = i < trns.size() ? trns[i] : 255;
          continue;
        }
        if (depth < 8) s[0] = s[0] * 255 / ((1u << depth) - 1);
        switch (ch) {
          case 1: o[0] = o[1] = o[2] = (uint8_t)s[0]; o[3] = 255; break;
          case 2: o[0] = o[1] = o[2] = (uint8_t)s[0]; o[3] = (uint8_t)s[1]; break;
          case 3: o[0] = (uint8_t)s[0]; o[1] = (uint8_t)s[1]; o[2] = (uint8_t)s[2]; o[3] = 255; break;
          default: for (int k = 0; k < 4; ++k) o[k] = (uint8_t)s[k]; break;
        }
      }
      std::swap(prev, cur);
    }
  }
  flip_rows(rgba, w, h, 1);  // first row at the bottom
  return true;
}

namespace {

// ---------------------------------------------------------------------------
// glTF
// ---------------------------------------------------------------------------

struct Gltf {
  std::string dir;
  Json j;
  std::vector<uint8_t> bin0;
  bool has_bin0 = false;
  std::map<int64_t, std::vector<uint8_t>> buffers;
  std::string err;

  bool open(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    std::vector<uint8_t> raw;
    if (!read_file(path, raw)) return (err = "can't read " + path), false;
    const uint8_t* jsrc = raw.data();
    size_t jlen = raw.size();
    if (raw.size() >= 12 && !std::memcmp(raw.data(), "glTF", 4)) {
      const uint32_t length = std::min<uint32_t>(rd32(raw.data() + 8), (uint32_t)raw.size());
      size_t p = 12;
      jsrc = nullptr;
      while (p + 8 <= length) {
        const uint32_t clen = rd32(raw.data() + p), ctype = rd32(raw.data() + p + 4);
        if (p + 8 + (size_t)clen > length) return (err = "truncated glb chunk"), false;
        if (ctype == 0x4E4F534Au) {
          jsrc = raw.data() + p + 8;
          jlen = clen;
        } else if (ctype == 0x004E4942u) {
          bin0.assign(raw.data() + p + 8, raw.data() + p + 8 + clen);
          has_bin0 = true;
        }
        p += 8 + clen;
      }
      if (!jsrc) return (err = "glb without a JSON chunk"), false;
    }
    JsonParser jp{reinterpret_cast<const char*>(jsrc), reinterpret_cast<const char*>(jsrc) + jlen, {}};
    if (!jp.value(j, 0)) return (err = jp.err), false;
    if (j.type != Json::Obj) return (err = "glTF root is not an object"), false;
    return true;
  }

  const Json* arr(const char* key) const {
    const Json* a = j.get(key);
    return a && a->type == Json::Arr ? a : nullptr;
  }

  const std::vector<uint8_t>* buffer(int64_t i) {
    auto it = buffers.find(i);
    if (it != buffers.end()) return &it->second;
    const Json* bs = arr("buffers");
    if (!bs || i < 0 || (size_t)i >= bs->size()) return (err = "bad buffer index"), nullptr;
    const Json* uri = (*bs)[(size_t)i].get("uri");
    std::vector<uint8_t> data;
    if (!uri) {
      if (!has_bin0) return (err = "buffer without uri and no GLB chunk"), nullptr;
      data = bin0;
    } else if (uri->s.compare(0, 5, "data:") == 0) {
      const size_t comma = uri->s.find(',');
      if (comma == std::string::npos || !base64(uri->s.substr(comma + 1), data))
        return (err = "bad data URI"), nullptr;
    } else if (!read_file(dir + "/" + uri->s, data)) {
      return (err = "can't read buffer " + uri->s), nullptr;
    }
    return &(buffers[i] = std::move(data));
  }

  // bytes of a buffer view: pointer, length, stride
  bool view(int64_t vi, const uint8_t*& ptr, size_t& len, size_t& stride) {
    const Json* vs = arr("bufferViews");
    if (!vs || vi < 0 || (size_t)vi >= vs->size()) return (err = "bad buffer view"), false;
    const Json& v = (*vs)[(size_t)vi];
    const std::vector<uint8_t>* b = buffer(inum(v.get("buffer"), -1));
    if (!b) return false;
    const size_t off = (size_t)inum(v.get("byteOffset"), 0);
    len = (size_t)inum(v.get("byteLength"), 0);
    stride = (size_t)inum(v.get("byteStride"), 0);
    if (off + len > b->size()) return (err = "buffer view out of range"), false;
    ptr = b->data() + off;
    return true;
  }

  // accessor as float32 [count][comps]: floats as is, integers converted,
  // divided by the type's maximum when `normalize` (glTF normalized ints)
  bool accessor_f(int64_t ai, std::vector<float>& out, int& comps, bool normalize,
                  int* csize = nullptr) {
    const Json* as = arr("accessors");
    if (!as || ai < 0 || (size_t)ai >= as->size()) return (err = "bad accessor"), false;
    const Json& a = (*as)[(size_t)ai];
    const int ct = (int)inum(a.get("componentType"), 0);
    const Json* ty = a.get("type");
    const std::string t = ty ? ty->s : "";
    comps = t == "SCALAR" ? 1 : t == "VEC2" ? 2 : t == "VEC3" ? 3 : t == "VEC4" ? 4 : t == "MAT4" ? 16 : 0;
    int esz;
    switch (ct) {
      case 5120: case 5121: esz = 1; break;
      case 5122: case 5123: esz = 2; break;
      case 5125: case 5126: esz = 4; break;
      default: esz = 0;
    }
    if (!comps || !esz) return (err = "unsupported accessor type"), false;
    if (csize) *csize = esz;
    const size_t count = (size_t)inum(a.get("count"), 0);
    const uint8_t* base;
    size_t len, stride;
    if (!view(inum(a.get("bufferView"), -1), base, len, stride)) return false;
    const size_t off = (size_t)inum(a.get("byteOffset"), 0);
    const size_t el = (size_t)esz * comps;
    if (stride == 0) stride = el;
    if (count && off + (count - 1) * stride + el > len) return (err = "accessor out of range"), false;
    out.resize(count * comps);
    for (size_t i = 0; i < count; ++i) {
      const uint8_t* e = base + off + i * stride;
      for (int k = 0; k < comps; ++k) {
        const uint8_t* c = e + (size_t)k * esz;
        float v;
        switch (ct) {
          case 5120: v = (float)(int8_t)c[0]; if (normalize) v = v / 127.0f; break;
          case 5121: v = (float)c[0]; if (normalize) v = v / 255.0f; break;
          case 5122: { int16_t x; std::memcpy(&x, c, 2); v = (float)x; if (normalize) v = v / 32767.0f; break; }
          case 5123: { uint16_t x; std::memcpy(&x, c, 2); v = (float)x; if (normalize) v = v / 65535.0f; break; }
          case 5125: { uint32_t x; std::memcpy(&x, c, 4); v = (float)x; break; }
          default: std::memcpy(&v, c, 4); break;
        }
        out[i * comps + k] = v;
      }
    }
    return true;
  }

  bool accessor_u32(int64_t ai, std::vector<uint32_t>& out, int& csize) {
    const Json* as = arr("accessors");
    if (!as || ai < 0 || (size_t)ai >= as->size()) return (err = "bad accessor"), false;
    const Json& a = (*as)[(size_t)ai];
    const int ct = (int)inum(a.get("componentType"), 0);
    csize = ct == 5121 ? 1 : ct == 5123 ? 2 : ct == 5125 ? 4 : 0;
    if (!csize) return (err = "unsupported index type"), false;
    const size_t count = (size_t)inum(a.get("count"), 0);
    const uint8_t* base;
    size_t len, stride;
    if (!view(inum(a.get("bufferView"), -1), base, len, stride)) return false;
    const size_t off = (size_t)inum(a.get("byteOffset"), 0);
    if (stride == 0) stride = (size_t)csize;
    if (count && off + (count - 1) * stride + csize > len) return (err = "index accessor out of range"), false;
    out.resize(count);
    for (size_t i = 0; i < count; ++i) {
      const uint8_t* e = base + off + i * stride;
      if (csize == 1) out[i] = e[0];
      else if (csize == 2) out[i] = (uint32_t)e[0] | ((uint32_t)e[1] << 8);
      else out[i] = rd32(e);
    }
    return true;
  }

  bool image(int64_t ii, std::vector<uint8_t>& data, std::string& mime) {
    const Json* ims = arr("images");
    if (!ims || ii < 0 || (size_t)ii >= ims->size()) return (err = "bad image index"), false;
    const Json& im = (*ims)[(size_t)ii];
    const Json* mt = im.get("mimeType");
    mime = mt ? mt->s : "";
    if (im.has("bufferView")) {
      const uint8_t* p;
      size_t len, stride;
      if (!view(inum(im.get("bufferView"), -1), p, len, stride)) return false;
      data.assign(p, p + len);
      return true;
    }
    const Json* uri = im.get("uri");
    if (!uri) return (err = "image without data"), false;
    if (uri->s.compare(0, 5, "data:") == 0) {
      const size_t comma = uri->s.find(',');
      if (comma == std::string::npos || !base64(uri->s.substr(comma + 1), data))
        return (err = "bad image data URI"), false;
      return true;
    }
    if (!read_file(dir + "/" + uri->s, data)) return (err = "can't read image " + uri->s), false;
    if (mime.empty()) {
      const std::string& u = uri->s;
      mime = u.size() >= 5 && u.compare(u.size() - 5, 5, ".ktx2") == 0 ? "image/ktx2" : "image/png";
    }
    return true;
  }
};

// 3x3 product on column-major arrays, accumulated from +0 (Magnum order)
void mat3_mul(const float* a, const float* b, float* out) {
  float t[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      float s = 0.0f;
      for (int k = 0; k < 3; ++k) {
        const float p = a[k * 3 + r] * b[c * 3 + k];
        s = s + p;
      }
      t[c * 3 + r] = s;
    }
  std::memcpy(out, t, sizeof t);
}

void node_transform(const Json& n, float* m) {
  const Json* mat = n.get("matrix");
  if (mat && mat->size() == 16) {
    for (int i = 0; i < 16; ++i) m[i] = (float)(*mat)[(size_t)i].n;
    return;
  }
  float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  if (const Json* v = n.get("translation"))
    for (size_t i = 0; i < 3 && i < v->size(); ++i) t[i] = (float)(*v)[i].n;
  if (const Json* v = n.get("rotation"))
    for (size_t i = 0; i < 4 && i < v->size(); ++i) q[i] = (float)(*v)[i].n;
  if (const Json* v = n.get("scale"))
    for (size_t i = 0; i < 3 && i < v->size(); ++i) s[i] = (float)(*v)[i].n;
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  // rotation columns of the unit quaternion
  float rot[3][3];
  {
    float a = y * y, b = z * z, c = a + b; rot[0][0] = 1.0f - 2.0f * c;
    a = x * y; b = z * w; c = a + b; rot[0][1] = 2.0f * c;
    a = x * z; b = y * w; c = a - b; rot[0][2] = 2.0f * c;
    a = x * y; b = z * w; c = a - b; rot[1][0] = 2.0f * c;
    a = x * x; b = z * z; c = a + b; rot[1][1] = 1.0f - 2.0f * c;
    a = y * z; b = x * w; c = a + b; rot[1][2] = 2.0f * c;
    a = x * z; b = y * w; c = a + b; rot[2][0] = 2.0f * c;
    a = y * z; b = x * w; c = a - b; rot[2][1] = 2.0f * c;
    a = x * x; b = y * y; c = a + b; rot[2][2] = 1.0f - 2.0f * c;
  }
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[c * 4 + r] = rot[c][r] * s[c];
  for (int r = 0; r < 3; ++r) m[12 + r] = t[r];
}

bool triangulate(std::vector<uint32_t>& idx, int mode, std::string& err) {
  if (mode == 4) {
    idx.resize(idx.size() / 3 * 3);
    return true;
  }
  std::vector<uint32_t> out;
  const size_t n = idx.size();
  if (mode == 6) {
    for (size_t i = 0; i + 2 < n; ++i) {
      out.push_back(idx[0]);
      out.push_back(idx[i + 1]);
      out.push_back(idx[i + 2]);
    }
  } else if (mode == 5) {
    for (size_t i = 0; i + 2 < n; ++i) {
      if (i % 2 == 0) {
        out.push_back(idx[i]);
        out.push_back(idx[i + 1]);
      } else {
        out.push_back(idx[i + 1]);
        out.push_back(idx[i]);
      }
      out.push_back(idx[i + 2]);
    }
  } else {
    err = "unsupported primitive mode " + std::to_string(mode);
    return false;
  }
  idx.swap(out);
  return true;
}

}  // namespace

bool load_gltf(const std::string& path, File& out, std::string& err) {
  Gltf g;
  if (!g.open(path)) {
    err = g.err;
    return false;
  }
  out = File();

  // textures: KHR_texture_ktx layers of one (image, sampler) pair share a texture
  std::map<int64_t, int32_t> tex_map;
  std::map<int64_t, uint32_t> tex_layer;
  std::map<std::pair<int64_t, int64_t>, int32_t> dedup;
  struct Cached {
    Texture t;
  };
  std::map<int64_t, std::shared_ptr<Cached>> images;
  const Json* samplers = g.arr("samplers");
  if (const Json* texs = g.arr("textures")) {
    for (size_t ti = 0; ti < texs->size(); ++ti) {
      const Json& t = (*texs)[ti];
      const Json* ext = nullptr;
      if (const Json* e = t.get("extensions")) ext = e->get("KHR_texture_ktx");
      const int64_t src = ext ? inum(ext->get("source"), -1) : inum(t.get("source"), -1);
      const uint32_t layer = ext ? (uint32_t)inum(ext->get("layer"), 0) : 0;
      if (src < 0) return (err = "texture " + std::to_string(ti) + " has no image"), false;
      const int64_t sidx = inum(t.get("sampler"), -1);
      const Json* smp = (samplers && sidx >= 0 && (size_t)sidx < samplers->size()) ? &(*samplers)[(size_t)sidx] : nullptr;
      const auto key = std::make_pair(src, sidx);
      if (ext && dedup.count(key)) {
        tex_map[(int64_t)ti] = dedup[key];
        tex_layer[(int64_t)ti] = layer;
        continue;
      }
      if (!images.count(src)) {
        std::vector<uint8_t> data;
        std::string mime;
        if (!g.image(src, data, mime)) return (err = g.err), false;
        auto c = std::make_shared<Cached>();
        const bool ktx = mime.find("ktx2") != std::string::npos ||
                         (data.size() >= 12 && !std::memcmp(data.data(), kKtx2Magic, 12));
        if (ktx) {
          uint32_t lf = 0;
          if (!decode_ktx2(data.data(), data.size(), c->t, lf, err)) return false;
        } else {
          uint32_t w, h;
          std::vector<uint8_t> rgba;
          if (!decode_png(data.data(), data.size(), w, h, rgba, err)) return false;
          c->t.width = w;
          c->t.height = h;
          c->t.layers = 1;
          c->t.levels.push_back(std::move(rgba));
        }
        images[src] = c;
      }
      Texture td = images[src]->t;
      if (smp) {
        td.min_filter = (uint32_t)inum(smp->get("minFilter"), 9987);
        td.mag_filter = (uint32_t)inum(smp->get("magFilter"), 9729);
        td.wrap_s = (uint32_t)inum(smp->get("wrapS"), 10497);
        td.wrap_t = (uint32_t)inum(smp->get("wrapT"), 10497);
      }
      tex_map[(int64_t)ti] = (int32_t)out.textures.size();
      tex_layer[(int64_t)ti] = layer;
      if (ext) dedup[key] = tex_map[(int64_t)ti];
      out.textures.push_back(std::move(td));
    }
  }

  // materials
  const float flip[9] = {1, 0, 0, 0, -1, 0, 0, 1, 1};  // v -> 1 - v, column-major
  if (const Json* mats = g.arr("materials")) {
    for (size_t mi = 0; mi < mats->size(); ++mi) {
      const Json& m = (*mats)[mi];
      Material md;
      if (const Json* e = m.get("extensions")) md.flat = e->has("KHR_materials_unlit");
      const Json* pbr = m.get("pbrMetallicRoughness");
      if (pbr) {
        if (const Json* c = pbr->get("baseColorFactor"))
          for (size_t k = 0; k < 4 && k < c->size(); ++k) md.base_color[k] = (float)(*c)[k].n;
        if (const Json* bct = pbr->get("baseColorTexture")) {
          const int64_t ti = inum(bct->get("index"), -1);
          if (!tex_map.count(ti)) return (err = "material references a missing texture"), false;
          md.texture = tex_map[ti];
          md.layer = tex_layer[ti];
          const Json* tt = nullptr;
          if (const Json* e = bct->get("extensions")) tt = e->get("KHR_texture_transform");
          if (tt) {
            float off[2] = {0, 0}, sc[2] = {1, 1};
            if (const Json* v = tt->get("offset"))
              for (size_t k = 0; k < 2 && k < v->size(); ++k) off[k] = (float)(*v)[k].n;
            if (const Json* v = tt->get("scale"))
              for (size_t k = 0; k < 2 && k < v->size(); ++k) sc[k] = (float)(*v)[k].n;
            const double rr = num(tt->get("rotation"), 0.0);
            const float cs = (float)std::cos(rr), sn = (float)std::sin(rr);
            float gm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 1};  // glTF T * R * S
            gm[0] = cs * sc[0];
            gm[1] = -sn * sc[0];
            gm[3] = sn * sc[1];
            gm[4] = cs * sc[1];
            gm[6] = off[0];
            gm[7] = off[1];
            float tmp[9];
            mat3_mul(flip, gm, tmp);
            mat3_mul(tmp, flip, md.texture_matrix);
          }
        }
      }
      out.materials.push_back(md);
    }
  }

  // meshes: one per primitive
  std::vector<std::vector<std::pair<uint32_t, int32_t>>> prim_ids;
  if (const Json* meshes = g.arr("meshes")) {
    for (size_t mi = 0; mi < meshes->size(); ++mi) {
      std::vector<std::pair<uint32_t, int32_t>> ids;
      const Json* prims = (*meshes)[mi].get("primitives");
      for (size_t pi = 0; prims && pi < prims->size(); ++pi) {
        const Json& prim = (*prims)[pi];
        const Json* at = prim.get("attributes");
        if (!at || !at->has("POSITION")) return (err = "primitive without POSITION"), false;
        Mesh mesh;
        int comps;
        if (!g.accessor_f(inum(at->get("POSITION"), -1), mesh.positions, comps, false)) return (err = g.err), false;
        if (comps != 3) return (err = "POSITION is not VEC3"), false;
        const size_t nv = mesh.positions.size() / 3;
        const int mode = (int)inum(prim.get("mode"), 4);
        if (prim.has("indices")) {
          int isz;
          if (!g.accessor_u32(inum(prim.get("indices"), -1), mesh.indices, isz)) return (err = g.err), false;
          mesh.index_size = (uint32_t)isz;
        } else {
          mesh.indices.resize(nv);
          for (size_t i = 0; i < nv; ++i) mesh.indices[i] = (uint32_t)i;
          mesh.index_size = 4;
        }
        if (!triangulate(mesh.indices, mode, err)) return false;
        for (uint32_t ix : mesh.indices)
          if (ix >= nv) return (err = "index out of range"), false;
        if (at->has("TEXCOORD_0")) {
          if (!g.accessor_f(inum(at->get("TEXCOORD_0"), -1), mesh.texcoords, comps, true)) return (err = g.err), false;
          if (comps != 2) return (err = "TEXCOORD_0 is not VEC2"), false;
          for (size_t i = 0; i < nv; ++i) mesh.texcoords[2 * i + 1] = 1.0f - mesh.texcoords[2 * i + 1];
        }
        if (at->has("NORMAL")) {
          if (!g.accessor_f(inum(at->get("NORMAL"), -1), mesh.normals, comps, false)) return (err = g.err), false;
          if (comps != 3) return (err = "NORMAL is not VEC3"), false;
        }
        if (at->has("COLOR_0")) {
          std::vector<float> c;
          if (!g.accessor_f(inum(at->get("COLOR_0"), -1), c, comps, true)) return (err = g.err), false;
          if (comps != 3 && comps != 4) return (err = "COLOR_0 is not VEC3/VEC4"), false;
          mesh.colors.resize(nv * 4);
          for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 4; ++k) mesh.colors[4 * i + k] = k < comps ? c[i * comps + k] : 1.0f;
        }
        ids.emplace_back((uint32_t)out.meshes.size(), (int32_t)inum(prim.get("material"), -1));
        out.meshes.push_back(std::move(mesh));
      }
      prim_ids.push_back(std::move(ids));
    }
  }

  // scene: nodes in index order, parents from children lists
  const Json* scenes = g.arr("scenes");
  if (!scenes || scenes->size() == 0) {
    out.has_scene = false;
    return true;
  }
  if (scenes->size() != 1) return (err = "expected exactly one scene, got " + std::to_string(scenes->size())), false;
  const Json* nodes = g.arr("nodes");
  const size_t nn = nodes ? nodes->size() : 0;
  std::vector<int32_t> parent(nn, -1);
  for (size_t i = 0; i < nn; ++i)
    if (const Json* ch = (*nodes)[i].get("children"))
      for (size_t k = 0; k < ch->size(); ++k) {
        const int64_t c = (int64_t)(*ch)[k].n;
        if (c < 0 || (size_t)c >= nn) return (err = "bad child index"), false;
        parent[(size_t)c] = (int32_t)i;
      }
  for (size_t i = 0; i < nn; ++i) {
    const Json& n = (*nodes)[i];
    Node nd;
    nd.parent = parent[i];
    node_transform(n, nd.transform);
    if (const Json* nm = n.get("name")) {
      nd.has_name = nm->type == Json::Str;
      nd.name = nm->s;
    }
    out.nodes.push_back(nd);
    if (!n.has("mesh")) continue;
    const int64_t mesh = inum(n.get("mesh"), -1);
    if (mesh < 0 || (size_t)mesh >= prim_ids.size()) return (err = "bad mesh index"), false;
    Entry base;
    base.node = (uint32_t)i;
    if (const Json* ex = n.get("extras")) {
      if (ex->has("meshViewIndexOffset")) {
        if (!ex->has("meshViewIndexCount") || !ex->has("meshViewMaterial"))
          return (err = "incomplete mesh view fields"), false;
        base.has_view = true;
        base.view_offset = (uint32_t)inum(ex->get("meshViewIndexOffset"), 0);
        base.view_count = (uint32_t)inum(ex->get("meshViewIndexCount"), 0);
        base.view_material = (int32_t)inum(ex->get("meshViewMaterial"), -1);
      }
    }
    for (const auto& id : prim_ids[(size_t)mesh]) {
      Entry e = base;
      e.mesh = id.first;
      e.material = id.second;
      out.entries.push_back(e);
    }
  }
  return true;
}

gbr_file_desc File::desc() {
  mesh_desc.assign(meshes.size(), gbr_mesh_desc{});
  for (size_t i = 0; i < meshes.size(); ++i) {
    const Mesh& m = meshes[i];
    gbr_mesh_desc& d = mesh_desc[i];
    d.vertex_count = (uint32_t)(m.positions.size() / 3);
    d.positions = m.positions.data();
    d.texcoords = m.texcoords.empty() ? nullptr : m.texcoords.data();
    d.colors = m.colors.empty() ? nullptr : m.colors.data();
    d.colors_u8 = 0;
    d.texcoords_half = 0;
    d.index_count = (uint32_t)m.indices.size();
    d.indices = m.indices.data();
    d.source_index_size = m.index_size;
    d.normals = m.normals.empty() ? nullptr : m.normals.data();
  }
  texel_blobs.assign(textures.size(), {});
  texture_desc.assign(textures.size(), gbr_texture_desc{});
  for (size_t i = 0; i < textures.size(); ++i) {
    const Texture& t = textures[i];
    for (const auto& lv : t.levels) texel_blobs[i].insert(texel_blobs[i].end(), lv.begin(), lv.end());
    gbr_texture_desc& d = texture_desc[i];
    d.width = t.width;
    d.height = t.height;
    d.layers = t.layers;
    d.levels = (uint32_t)t.levels.size();
    d.texels = texel_blobs[i].data();
    d.min_filter = t.min_filter;
    d.mag_filter = t.mag_filter;
    d.wrap_s = t.wrap_s;
    d.wrap_t = t.wrap_t;
    d.compressed = t.compressed ? 1 : 0;
  }
  material_desc.assign(materials.size(), gbr_material_desc{});
  for (size_t i = 0; i < materials.size(); ++i) {
    const Material& m = materials[i];
    gbr_material_desc& d = material_desc[i];
    d.flat = m.flat ? 1 : 0;
    std::memcpy(d.base_color, m.base_color, sizeof d.base_color);
    d.texture = m.texture;
    d.layer = m.layer;
    std::memcpy(d.texture_matrix, m.texture_matrix, sizeof d.texture_matrix);
  }
  node_desc.assign(nodes.size(), gbr_node_desc{});
  for (size_t i = 0; i < nodes.size(); ++i) {
    node_desc[i].parent = nodes[i].parent;
    std::memcpy(node_desc[i].transform, nodes[i].transform, sizeof node_desc[i].transform);
    node_desc[i].name = nodes[i].has_name ? nodes[i].name.c_str() : nullptr;
  }
  entry_desc.assign(entries.size(), gbr_mesh_entry_desc{});
  for (size_t i = 0; i < entries.size(); ++i) {
    const Entry& e = entries[i];
    gbr_mesh_entry_desc& d = entry_desc[i];
    d.node = e.node;
    d.mesh = e.mesh;
    d.material = e.material;
    d.has_view = e.has_view ? 1 : 0;
    d.view_index_offset_bytes = e.view_offset;
    d.view_index_count = e.view_count;
    d.view_material = e.view_material;
  }
  gbr_file_desc f{};
  f.mesh_count = (uint32_t)meshes.size();
  f.meshes = mesh_desc.data();
  f.texture_count = (uint32_t)textures.size();
  f.textures = texture_desc.data();
  f.material_count = (uint32_t)materials.size();
  f.materials = material_desc.data();
  f.node_count = (uint32_t)nodes.size();
  f.nodes = node_desc.data();
  f.mesh_entry_count = (uint32_t)entries.size();
  f.mesh_entries = entry_desc.data();
  f.has_scene = has_scene ? 1 : 0;
  return f;
}

}  // namespace imp
}  // namespace gbr
