// .strands cache loader: parse + validate + pack, behind the C ABI.
//
// Restates the reference's asset loader (src/strand_cache.rs):
//   * binary layout (parse_binary_cache, :110-158): magic "SSRSTRD\0",
//     u32 version (= 1), f32 scale, u64 vertex / strand / index counts, then
//     vertex_count * 3 f32, strand_count * (u32 count, u32 offset),
//     index_count * u32, all little-endian, and nothing after the payload;
//   * JSON layout (JsonStrandCacheFile, :32-40): {version, source, scale
//     (default 1), vertices: [[x, y, z]...], strands: [[index...]...]}, each
//     strand's indices appended in order (TryFrom, :46-67);
//   * format choice by the magic prefix (parse_cache, :165-171);
//   * validate_cache (:174-209);
// and the packing of set_cached_strand_geometry + pack_indexed_strands
// (src/plugin.rs:1468-1490, 1357-1393): vertices scaled by `scale`, stored
// (x, y, z, 1); StrandMeta{count, offset, material 0}; one material slab
// entry; AABB over every vertex; max_segments_in_strand = max vertex count.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/strand_rt.h"

namespace {

constexpr char kMagic[8] = {'S', 'S', 'R', 'S', 'T', 'R', 'D', '\0'};
constexpr uint32_t kVersion = 1;

struct Cache {
  uint32_t version = 0;
  float scale = 1.0f;
  std::vector<float> xyz;            // 3 per vertex
  std::vector<uint32_t> counts, offsets;
  std::vector<uint32_t> indices;
};

struct Error {
  std::string msg;
  bool set(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    msg = buf;
    return false;
  }
};

template <class T>
T load_le(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));  // the build targets are little-endian
  return v;
}

// parse_binary_cache: sequential reads; a read past the end is
// "unexpected end of strand cache".
bool parse_binary(const uint8_t* b, uint64_t n, Cache& c, Error& e) {
  uint64_t at = 0;
  auto need = [&](uint64_t k) { return k <= n && at <= n - k; };
  if (!need(8)) return e.set("unexpected end of strand cache");
  if (std::memcmp(b, kMagic, 8) != 0) return e.set("invalid strand cache magic");
  at = 8;
  if (!need(4)) return e.set("unexpected end of strand cache");
  c.version = load_le<uint32_t>(b + at);
  at += 4;
  if (c.version != kVersion) return e.set("unsupported strand cache version %u", c.version);
  if (!need(4 + 24)) return e.set("unexpected end of strand cache");
  c.scale = load_le<float>(b + at);
  const uint64_t vc = load_le<uint64_t>(b + at + 4), sc = load_le<uint64_t>(b + at + 12),
                 ic = load_le<uint64_t>(b + at + 20);
  at += 28;
  // every count must fit the bytes that remain (guards the allocations too)
  const uint64_t rem = n - at;
  if (vc > rem / 12u) return e.set("unexpected end of strand cache");
  if (sc > (rem - vc * 12u) / 8u) return e.set("unexpected end of strand cache");
  if (ic > (rem - vc * 12u - sc * 8u) / 4u) return e.set("unexpected end of strand cache");
  const uint64_t payload = vc * 12u + sc * 8u + ic * 4u;
  if (rem != payload) return e.set("strand cache has %llu trailing bytes after payload",
                                   (unsigned long long)(rem - payload));
  c.xyz.resize(vc * 3u);
  c.counts.resize(sc);
  c.offsets.resize(sc);
  c.indices.resize(ic);
  std::memcpy(c.xyz.data(), b + at, vc * 12u);
  at += vc * 12u;
  for (uint64_t i = 0; i < sc; ++i) {
    c.counts[i] = load_le<uint32_t>(b + at + 8u * i);
    c.offsets[i] = load_le<uint32_t>(b + at + 8u * i + 4u);
  }
  at += sc * 8u;
  std::memcpy(c.indices.data(), b + at, ic * 4u);
  return true;
}

// ---------------------------------------------------------------- JSON
// A strict reader for the one schema above; unknown keys are skipped like
// serde's default, "source" may be a string or null.
struct Json {
  const char* p;
  const char* end;
  Error& e;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool lit(char ch) {
    ws();
    if (p < end && *p == ch) { ++p; return true; }
    return false;
  }
  bool expect(char ch) { return lit(ch) || e.set("invalid JSON strand cache: expected '%c' at byte %ld", ch, (long)(p - start)); }
  const char* start;

  bool number(double& v) {
    ws();
    const char* q = p;
    if (q < end && (*q == '-' || *q == '+')) ++q;
    bool digits = false;
    while (q < end && ((*q >= '0' && *q <= '9') || *q == '.' || *q == 'e' || *q == 'E' || *q == '-' || *q == '+')) {
      digits = digits || (*q >= '0' && *q <= '9');
      ++q;
    }
    if (!digits) return e.set("invalid JSON strand cache: expected a number at byte %ld", (long)(p - start));
    std::string s(p, q);
    char* stop = nullptr;
    v = std::strtod(s.c_str(), &stop);
    if (!stop || *stop) return e.set("invalid JSON strand cache: bad number '%s'", s.c_str());
    p = q;
    return true;
  }
  bool u32(uint32_t& v) {
    double d;
    if (!number(d)) return false;
    if (!(d >= 0.0 && d <= 4294967295.0) || std::floor(d) != d)
      return e.set("invalid JSON strand cache: %g is not a u32", d);
    v = (uint32_t)d;
    return true;
  }
  bool str(std::string* out) {
    ws();
    if (p >= end || *p != '"') return e.set("invalid JSON strand cache: expected a string at byte %ld", (long)(p - start));
    ++p;
    std::string s;
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) break;
        switch (*p) {
          case 'n': s += '\n'; break;
          case 't': s += '\t'; break;
          case 'r': s += '\r'; break;
          case 'b': s += '\b'; break;
          case 'f': s += '\f'; break;
          case 'u': s += '?'; p += (end - p > 4) ? 4 : (end - p - 1); break;
          default: s += *p;
        }
        ++p;
      } else {
        s += *p++;
      }
    }
    if (p >= end) return e.set("invalid JSON strand cache: unterminated string");
    ++p;
    if (out) *out = s;
    return true;
  }
  bool word(const char* w) {
    ws();
    size_t k = std::strlen(w);
    if ((size_t)(end - p) >= k && std::memcmp(p, w, k) == 0) { p += k; return true; }
    return false;
  }
  bool skip_value(int depth = 0) {
    ws();
    if (depth > 64) return e.set("invalid JSON strand cache: nesting too deep");
    if (p >= end) return e.set("invalid JSON strand cache: unexpected end");
    if (*p == '"') return str(nullptr);
    if (word("null") || word("true") || word("false")) return true;
    if (*p == '[') {
      ++p;
      if (lit(']')) return true;
      do { if (!skip_value(depth + 1)) return false; } while (lit(','));
      return expect(']');
    }
    if (*p == '{') {
      ++p;
      if (lit('}')) return true;
      do {
        if (!str(nullptr) || !expect(':') || !skip_value(depth + 1)) return false;
      } while (lit(','));
      return expect('}');
    }
    double d;
    return number(d);
  }
};

bool parse_json(const uint8_t* b, uint64_t n, Cache& c, Error& e) {
  Json j{reinterpret_cast<const char*>(b), reinterpret_cast<const char*>(b) + n, e,
         reinterpret_cast<const char*>(b)};
  bool have_version = false, have_vertices = false, have_strands = false;
  if (!j.expect('{')) return false;
  if (!j.lit('}')) {
    do {
      std::string key;
      if (!j.str(&key) || !j.expect(':')) return false;
      if (key == "version") {
        if (!j.u32(c.version)) return false;
        have_version = true;
      } else if (key == "scale") {
        double d;
        if (!j.number(d)) return false;
        c.scale = (float)d;
      } else if (key == "source") {
        if (!j.word("null") && !j.str(nullptr)) return false;
      } else if (key == "vertices") {
        if (!j.expect('[')) return false;
        c.xyz.clear();
        if (!j.lit(']')) {
          do {
            if (!j.expect('[')) return false;
            for (int k = 0; k < 3; ++k) {
              double d;
              if ((k && !j.expect(',')) || !j.number(d)) return false;
              c.xyz.push_back((float)d);
            }
            if (!j.expect(']')) return false;
          } while (j.lit(','));
          if (!j.expect(']')) return false;
        }
        have_vertices = true;
      } else if (key == "strands") {
        if (!j.expect('[')) return false;
        c.counts.clear(); c.offsets.clear(); c.indices.clear();
        if (!j.lit(']')) {
          do {
            if (!j.expect('[')) return false;
            const uint32_t off = (uint32_t)c.indices.size();
            if (!j.lit(']')) {
              do {
                uint32_t v;
                if (!j.u32(v)) return false;
                c.indices.push_back(v);
              } while (j.lit(','));
              if (!j.expect(']')) return false;
            }
            c.counts.push_back((uint32_t)c.indices.size() - off);
            c.offsets.push_back(off);
          } while (j.lit(','));
          if (!j.expect(']')) return false;
        }
        have_strands = true;
      } else if (!j.skip_value()) {
        return false;
      }
    } while (j.lit(','));
    if (!j.expect('}')) return false;
  }
  j.ws();
  if (j.p != j.end) return e.set("invalid JSON strand cache: trailing characters at byte %ld", (long)(j.p - j.start));
  if (!have_version) return e.set("invalid JSON strand cache: missing field `version`");
  if (!have_vertices) return e.set("invalid JSON strand cache: missing field `vertices`");
  if (!have_strands) return e.set("invalid JSON strand cache: missing field `strands`");
  return true;
}

// validate_cache (src/strand_cache.rs:174-209)
bool validate(const Cache& c, Error& e) {
  if (c.version != kVersion) return e.set("unsupported strand cache version %u", c.version);
  const uint64_t nv = c.xyz.size() / 3u;
  for (size_t s = 0; s < c.counts.size(); ++s) {
    if (c.counts[s] < 2u) return e.set("strand %zu has fewer than two vertices", s);
    const uint64_t end = (uint64_t)c.offsets[s] + c.counts[s];
    if (end > 0xFFFFFFFFull) return e.set("strand %zu index range overflows", s);
    if (end > c.indices.size())
      return e.set("strand %zu range %u..%llu exceeds %zu cache indices", s, c.offsets[s], (unsigned long long)end,
                   c.indices.size());
    for (uint64_t i = c.offsets[s]; i < end; ++i)
      if (c.indices[i] >= nv)
        return e.set("strand %zu references vertex %u, but cache has %llu vertices", s, c.indices[i],
                     (unsigned long long)nv);
  }
  return true;
}

}  // namespace

extern "C" int sr_strand_cache_parse(const void* bytes, uint64_t size, sr_strand_cache_info* info, float* vertices,
                                     sr_strand_meta* meta, uint32_t* indices, char* error, uint64_t error_capacity) {
  Error e;
  auto fail = [&](int code) {
    if (error && error_capacity) {
      std::snprintf(error, error_capacity, "%s", e.msg.c_str());
    }
    return code;
  };
  if (!info || (!bytes && size)) {
    e.msg = "sr_strand_cache_parse: info and bytes are required";
    return fail(SR_ERR_INVALID_ARGUMENT);
  }
  const uint8_t* b = static_cast<const uint8_t*>(bytes);
  Cache c;
  const bool binary = size >= 8 && std::memcmp(b, kMagic, 8) == 0;
  if (binary) {
    if (!parse_binary(b, size, c, e)) return fail(SR_ERR_INVALID_ARGUMENT);
  } else {
    if (!parse_json(b, size, c, e)) return fail(SR_ERR_INVALID_ARGUMENT);
  }
  if (!validate(c, e)) return fail(SR_ERR_INVALID_ARGUMENT);
  const uint64_t nv = c.xyz.size() / 3u;
  if (nv == 0 || c.counts.empty()) {  // pack_indexed_strands returns None
    e.msg = "no valid cached strands";
    return fail(SR_ERR_INVALID_ARGUMENT);
  }
  std::memset(info, 0, sizeof(*info));
  info->version = c.version;
  info->scale = c.scale;
  info->format = binary ? 0u : 1u;
  info->vertex_count = nv;
  info->strand_count = c.counts.size();
  info->index_count = c.indices.size();
  uint32_t max_count = 0;
  for (uint32_t k : c.counts) max_count = k > max_count ? k : max_count;
  info->max_segments_in_strand = max_count;
  // Vec3 * scale, then Aabb3d::from_point_cloud over every vertex
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t v = 0; v < nv; ++v) {
    for (int k = 0; k < 3; ++k) {
      const float x = c.xyz[3 * v + k] * c.scale;
      mn[k] = x < mn[k] ? x : mn[k];
      mx[k] = x > mx[k] ? x : mx[k];
      if (vertices) vertices[4 * v + k] = x;
    }
    if (vertices) vertices[4 * v + 3] = 1.0f;
  }
  for (int k = 0; k < 3; ++k) {
    info->aabb.min[k] = mn[k];
    info->aabb.max[k] = mx[k];
  }
  if (meta)
    for (size_t s = 0; s < c.counts.size(); ++s) meta[s] = sr_strand_meta{c.counts[s], c.offsets[s], 0u, 0u};
  if (indices && !c.indices.empty()) std::memcpy(indices, c.indices.data(), c.indices.size() * 4u);
  if (error && error_capacity) error[0] = '\0';
  return SR_OK;
}
