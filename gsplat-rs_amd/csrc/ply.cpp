// ply.cpp -- bounded ASCII / binary PLY loader producing reference SceneBuffers.
//
// Behaviour (accepted inputs, limits, error kinds and texts, RDF->RUF
// conversion) follows crates/gsplat-io-ply/src/lib.rs; each function cites
// the part it restates.

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "common.h"

namespace gs {
namespace {

// PlyLoadError (io-ply/src/lib.rs:128-159) with its code mapping (:161-176).
[[noreturn]] void fail(Code c, const std::string &m) { throw Error(c, m); }
[[noreturn]] void err_io() { fail(Code::NotFound, "I/O error while reading PLY"); }
[[noreturn]] void err_unsupported_format() {
  fail(Code::Unsupported,
       "unsupported PLY format; supported: ascii 1.0, binary_little_endian 1.0, binary_big_endian 1.0");
}
[[noreturn]] void err_malformed() { fail(Code::ParseFailed, "malformed PLY header"); }
[[noreturn]] void err_missing(const char *f) {
  fail(Code::ParseFailed, std::string("missing required field `") + f + "`");
}
[[noreturn]] void err_field_count() {
  fail(Code::ParseFailed, "vertex row does not match declared property count");
}
[[noreturn]] void err_parse_number() { fail(Code::ParseFailed, "failed to parse numeric value"); }
[[noreturn]] void err_count_mismatch() {
  fail(Code::ParseFailed, "vertex rows do not match declared vertex count");
}
[[noreturn]] void err_invalid_scene() { fail(Code::ParseFailed, "parsed scene buffers are inconsistent"); }
[[noreturn]] void err_overflow(const char *what) {
  fail(Code::Unsupported, std::string("PLY resource size overflow while computing ") + what);
}
void ensure_limit(const char *resource, uint64_t requested, uint64_t limit) {
  if (requested > limit)
    fail(Code::Unsupported, std::string("PLY resource limit exceeded for ") + resource + ": requested " +
                                std::to_string(requested) + ", limit " + std::to_string(limit));
}

const char *const kRequired[14] = {"x",     "y",     "z",     "opacity", "scale_0", "scale_1", "scale_2",
                                   "rot_0", "rot_1", "rot_2", "rot_3",   "f_dc_0",  "f_dc_1",  "f_dc_2"};

// SH_FLIP_RDF_TO_RUF (io-ply/src/lib.rs:54-56).
const float kShFlip[15] = {-1.0f, 1.0f, 1.0f, -1.0f, -1.0f, 1.0f, 1.0f, 1.0f,
                           -1.0f, -1.0f, -1.0f, 1.0f, 1.0f, 1.0f, 1.0f};

// ---------------------------------------------------------------- UTF-8 text

// Decodes one UTF-8 code point at s[i]; returns its length or 0 if invalid.
size_t utf8_next(const uint8_t *s, size_t n, size_t i, uint32_t *cp) {
  uint8_t c = s[i];
  if (c < 0x80) {
    *cp = c;
    return 1;
  }
  size_t len;
  uint32_t v, min;
  if ((c & 0xE0) == 0xC0) {
    len = 2;
    v = c & 0x1F;
    min = 0x80;
  } else if ((c & 0xF0) == 0xE0) {
    len = 3;
    v = c & 0x0F;
    min = 0x800;
  } else if ((c & 0xF8) == 0xF0) {
    len = 4;
    v = c & 0x07;
    min = 0x10000;
  } else {
    return 0;
  }
  if (i + len > n) return 0;
  for (size_t k = 1; k < len; ++k) {
    if ((s[i + k] & 0xC0) != 0x80) return 0;
    v = (v << 6) | (s[i + k] & 0x3F);
  }
  if (v < min || v > 0x10FFFF || (v >= 0xD800 && v <= 0xDFFF)) return 0;
  *cp = v;
  return len;
}

bool valid_utf8(const uint8_t *s, size_t n) {
  size_t i = 0;
  uint32_t cp;
  while (i < n) {
    size_t l = utf8_next(s, n, i, &cp);
    if (l == 0) return false;
    i += l;
  }
  return true;
}

// char::is_whitespace (Unicode White_Space).
bool is_ws(uint32_t c) {
  return (c >= 0x09 && c <= 0x0D) || c == 0x20 || c == 0x85 || c == 0xA0 || c == 0x1680 ||
         (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F ||
         c == 0x3000;
}

// A view of valid UTF-8 text.
struct Str {
  const char *p = nullptr;
  size_t n = 0;
  std::string s() const { return std::string(p, n); }
  bool eq(const char *t) const { return n == std::strlen(t) && std::memcmp(p, t, n) == 0; }
  bool starts_with(const char *t) const {
    size_t l = std::strlen(t);
    return n >= l && std::memcmp(p, t, l) == 0;
  }
};

// str::trim
Str trim(Str t) {
  const uint8_t *u = reinterpret_cast<const uint8_t *>(t.p);
  size_t b = 0, e = t.n;
  uint32_t cp;
  while (b < e) {
    size_t l = utf8_next(u, e, b, &cp);
    if (!is_ws(cp)) break;
    b += l;
  }
  while (e > b) {
    size_t s = e - 1;
    while (s > b && (u[s] & 0xC0) == 0x80) --s;
    utf8_next(u, e, s, &cp);
    if (!is_ws(cp)) break;
    e = s;
  }
  return Str{t.p + b, e - b};
}

// str::split_whitespace
std::vector<Str> split_ws(Str t) {
  std::vector<Str> out;
  const uint8_t *u = reinterpret_cast<const uint8_t *>(t.p);
  size_t i = 0, start = 0;
  bool in_tok = false;
  uint32_t cp;
  while (i < t.n) {
    size_t l = utf8_next(u, t.n, i, &cp);
    if (is_ws(cp)) {
      if (in_tok) out.push_back(Str{t.p + start, i - start});
      in_tok = false;
    } else if (!in_tok) {
      in_tok = true;
      start = i;
    }
    i += l;
  }
  if (in_tok) out.push_back(Str{t.p + start, t.n - start});
  return out;
}

// str::lines: split on '\n', strip one trailing '\r', no final empty line.
std::vector<Str> lines(Str t) {
  std::vector<Str> out;
  size_t start = 0;
  for (size_t i = 0; i < t.n; ++i) {
    if (t.p[i] == '\n') {
      size_t e = i;
      if (e > start && t.p[e - 1] == '\r') --e;
      out.push_back(Str{t.p + start, e - start});
      start = i + 1;
    }
  }
  if (start < t.n) {
    size_t e = t.n;
    if (e > start && t.p[e - 1] == '\r') --e;
    out.push_back(Str{t.p + start, e - start});
  }
  return out;
}

// <usize as FromStr>: optional '+', decimal digits, no overflow.
bool parse_usize(Str t, uint64_t *out) {
  size_t i = 0;
  if (t.n > 0 && t.p[0] == '+') i = 1;
  if (i >= t.n) return false;
  uint64_t v = 0;
  for (; i < t.n; ++i) {
    char c = t.p[i];
    if (c < '0' || c > '9') return false;
    uint64_t d = (uint64_t)(c - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

bool ieq(const char *a, size_t n, const char *lit) {
  size_t l = std::strlen(lit);
  if (n != l) return false;
  for (size_t i = 0; i < n; ++i) {
    char c = a[i];
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    if (c != lit[i]) return false;
  }
  return true;
}

// <f32 as FromStr>: [+-]? (inf | infinity | nan | digits[.digits] | .digits)
// ([eE][+-]?digits)?, case-insensitive words, correctly rounded.
bool parse_f32(Str t, float *out) {
  const char *p = t.p;
  size_t n = t.n, i = 0;
  bool neg = false;
  if (i < n && (p[i] == '+' || p[i] == '-')) {
    neg = p[i] == '-';
    ++i;
  }
  const char *w = p + i;
  size_t wn = n - i;
  if (ieq(w, wn, "inf") || ieq(w, wn, "infinity")) {
    *out = neg ? -INFINITY : INFINITY;
    return true;
  }
  if (ieq(w, wn, "nan")) {
    *out = NAN;
    return true;
  }
  size_t digits = 0, frac = 0;
  while (i < n && p[i] >= '0' && p[i] <= '9') ++i, ++digits;
  if (i < n && p[i] == '.') {
    ++i;
    while (i < n && p[i] >= '0' && p[i] <= '9') ++i, ++frac;
  }
  if (digits + frac == 0) return false;
  if (i < n && (p[i] == 'e' || p[i] == 'E')) {
    ++i;
    if (i < n && (p[i] == '+' || p[i] == '-')) ++i;
    size_t ed = 0;
    while (i < n && p[i] >= '0' && p[i] <= '9') ++i, ++ed;
    if (ed == 0) return false;
  }
  if (i != n) return false;
  std::string s(p, n);
  errno = 0;
  *out = std::strtof(s.c_str(), nullptr);
  return true;
}

// normalize_opacity_logit (io-ply/src/lib.rs:993-1003).
float normalize_opacity(float v) {
  if (std::isfinite(v)) return v;
  if (v == INFINITY) return 16.0f;
  if (v == -INFINITY) return -16.0f;
  err_parse_number();
}

// ------------------------------------------------------------------ header

enum class Format { Ascii, BinLE, BinBE };
enum class Ty { I8, U8, I16, U16, I32, U32, F32, F64 };
size_t ty_size(Ty t) {
  switch (t) {
    case Ty::I8:
    case Ty::U8: return 1;
    case Ty::I16:
    case Ty::U16: return 2;
    case Ty::F64: return 8;
    default: return 4;
  }
}
bool parse_ty(Str t, Ty *out) {
  if (t.eq("char") || t.eq("int8")) *out = Ty::I8;
  else if (t.eq("uchar") || t.eq("uint8")) *out = Ty::U8;
  else if (t.eq("short") || t.eq("int16")) *out = Ty::I16;
  else if (t.eq("ushort") || t.eq("uint16")) *out = Ty::U16;
  else if (t.eq("int") || t.eq("int32")) *out = Ty::I32;
  else if (t.eq("uint") || t.eq("uint32")) *out = Ty::U32;
  else if (t.eq("float") || t.eq("float32")) *out = Ty::F32;
  else if (t.eq("double") || t.eq("float64")) *out = Ty::F64;
  else return false;
  return true;
}

struct Prop {
  std::string name;
  Ty ty;
};
struct Header {
  Format format;
  uint64_t vertex_count;
  std::vector<Prop> props;
};

// parse_header_text (io-ply/src/lib.rs:516-598).
Header parse_header_text(Str text, const PlyLimits &lim) {
  std::vector<Str> ls = lines(text);
  if (ls.empty() || !trim(ls[0]).eq("ply")) err_malformed();
  bool have_format = false, saw_end = false, have_count = false, in_vertex = false;
  Format format = Format::Ascii;
  uint64_t count = 0;
  std::vector<Prop> props;
  for (size_t li = 1; li < ls.size(); ++li) {
    Str t = trim(ls[li]);
    if (t.n == 0 || t.starts_with("comment")) continue;
    if (t.eq("end_header")) {
      saw_end = true;
      break;
    }
    std::vector<Str> parts = split_ws(t);
    if (parts.empty()) continue;
    if (parts[0].eq("format")) {
      have_format = false;
      if (parts.size() >= 3 && parts[2].eq("1.0")) {
        if (parts[1].eq("ascii")) have_format = true, format = Format::Ascii;
        else if (parts[1].eq("binary_little_endian")) have_format = true, format = Format::BinLE;
        else if (parts[1].eq("binary_big_endian")) have_format = true, format = Format::BinBE;
      }
    } else if (parts[0].eq("element")) {
      in_vertex = parts.size() >= 2 && parts[1].eq("vertex");
      if (in_vertex) {
        uint64_t c;
        if (parts.size() < 3 || !parse_usize(parts[2], &c)) err_malformed();
        ensure_limit("vertices", c, lim.max_vertices);
        count = c;
        have_count = true;
      }
    } else if (parts[0].eq("property") && in_vertex) {
      if (parts.size() < 2) err_malformed();
      if (parts[1].eq("list")) err_unsupported_format();
      if (parts.size() < 3) err_malformed();
      Ty ty;
      if (!parse_ty(parts[1], &ty)) err_malformed();
      ensure_limit("vertex properties", (uint64_t)props.size() + 1, lim.max_vertex_properties);
      props.push_back(Prop{parts[2].s(), ty});
    }
  }
  if (!have_format) err_unsupported_format();
  if (!saw_end) err_malformed();
  if (!have_count) err_malformed();
  return Header{format, count, props};
}

// split_header_body (io-ply/src/lib.rs:473-514): returns header and body offset.
Header split_header_body(const uint8_t *in, size_t len, const PlyLimits &lim, size_t *body_off) {
  size_t cursor = 0;
  bool found = false;
  size_t header_end = 0;
  while (cursor < len) {
    uint64_t scan_end64 = std::min<uint64_t>(len, lim.max_header_bytes == UINT64_MAX
                                                      ? UINT64_MAX
                                                      : lim.max_header_bytes + 1);
    size_t scan_end = (size_t)scan_end64;
    if (cursor >= scan_end)
      fail(Code::Unsupported, "PLY resource limit exceeded for header bytes: requested " +
                                  std::to_string((uint64_t)cursor + 1) + ", limit " +
                                  std::to_string(lim.max_header_bytes));
    size_t line_end = scan_end;
    const void *nl = std::memchr(in + cursor, '\n', scan_end - cursor);
    if (nl) line_end = (size_t)((const uint8_t *)nl - in);
    size_t next_cursor = line_end < len ? line_end + 1 : line_end;
    ensure_limit("header bytes", next_cursor, lim.max_header_bytes);
    if (!valid_utf8(in + cursor, line_end - cursor)) err_malformed();
    Str line{(const char *)in + cursor, line_end - cursor};
    while (line.n > 0 && line.p[line.n - 1] == '\r') --line.n;  // trim_end_matches('\r')
    if (trim(line).eq("end_header")) {
      found = true;
      header_end = next_cursor;
      break;
    }
    cursor = line_end + 1;
  }
  if (!found) err_malformed();
  if (!valid_utf8(in, header_end)) err_malformed();
  *body_off = header_end;
  return parse_header_text(Str{(const char *)in, header_end}, lim);
}

// infer_sh_rest_layout / infer_sh_degree (io-ply/src/lib.rs:614-688).
uint32_t infer_sh(const Header &h, std::vector<size_t> *rest_props) {
  std::vector<std::pair<uint64_t, size_t>> pairs;
  for (size_t i = 0; i < h.props.size(); ++i) {
    const std::string &n = h.props[i].name;
    if (n.compare(0, 7, "f_rest_") == 0) {
      uint64_t idx;
      if (parse_usize(Str{n.data() + 7, n.size() - 7}, &idx)) pairs.emplace_back(idx, i);
    }
  }
  if (pairs.empty()) return 0;
  std::stable_sort(pairs.begin(), pairs.end(),
                   [](const std::pair<uint64_t, size_t> &a, const std::pair<uint64_t, size_t> &b) {
                     return a.first < b.first;
                   });
  uint64_t max_idx = pairs.back().first;
  if (pairs.size() != max_idx + 1) return 0;
  for (size_t e = 0; e < pairs.size(); ++e)
    if (pairs[e].first != e) return 0;
  uint64_t total = pairs.size();
  if (total % 3 != 0) return 0;
  uint64_t coeff_total = total / 3 + 1;
  uint64_t root = (uint64_t)std::sqrt((double)coeff_total);
  while (root * root > coeff_total) --root;
  while ((root + 1) * (root + 1) <= coeff_total) ++root;
  if (root * root != coeff_total) return 0;
  uint64_t degree = root - 1;
  if (degree == 0 || degree > 4) return 0;
  rest_props->clear();
  for (auto &p : pairs) rest_props->push_back(p.second);
  return (uint32_t)degree;
}

bool mul_ovf(uint64_t a, uint64_t b, uint64_t *out) { return __builtin_mul_overflow(a, b, out); }

float read_scalar(const uint8_t *rec, size_t off, Ty ty, bool big) {
  uint8_t b[8];
  size_t sz = ty_size(ty);
  for (size_t i = 0; i < sz; ++i) b[i] = rec[off + (big ? sz - 1 - i : i)];
  switch (ty) {
    case Ty::I8: return (float)(int8_t)b[0];
    case Ty::U8: return (float)b[0];
    case Ty::I16: { int16_t v; std::memcpy(&v, b, 2); return (float)v; }
    case Ty::U16: { uint16_t v; std::memcpy(&v, b, 2); return (float)v; }
    case Ty::I32: { int32_t v; std::memcpy(&v, b, 4); return (float)v; }
    case Ty::U32: { uint32_t v; std::memcpy(&v, b, 4); return (float)v; }
    case Ty::F32: { float v; std::memcpy(&v, b, 4); return v; }
    case Ty::F64: { double v; std::memcpy(&v, b, 8); return (float)v; }
  }
  return 0.0f;
}

bool rot_layout_xyzw() {
  const char *e = std::getenv("GSPLAT_ROT_LAYOUT");
  return e && (std::strcmp(e, "xyzw") == 0 || std::strcmp(e, "XYZW") == 0);
}

}  // namespace

// parse_ply_bytes_with_limits (io-ply/src/lib.rs:228-294).
HostScene ply_parse(const uint8_t *in, size_t len, const PlyLimits &lim, bool text_only) {
  if (text_only) {  // parse_ply_text_with_limits (:214-225)
    size_t off;
    Header h = split_header_body(in, len, lim, &off);
    if (h.format != Format::Ascii) err_unsupported_format();
  }
  ensure_limit("input bytes", len, lim.max_input_bytes);
  size_t body_off = 0;
  Header h = split_header_body(in, len, lim, &body_off);
  const uint8_t *body = in + body_off;
  size_t body_len = len - body_off;

  std::map<std::string, size_t> idx;
  for (size_t i = 0; i < h.props.size(); ++i) idx[h.props[i].name] = i;  // last one wins
  for (const char *r : kRequired)
    if (!idx.count(r)) err_missing(r);

  std::vector<size_t> rest_props;
  uint32_t degree = infer_sh(h, &rest_props);
  bool has_rest = degree > 0;
  uint64_t rest_stride = has_rest ? 3ull * sh_per_channel(degree) : 0;

  // validate_body_before_allocation (:361-390) and the binary layout (:863-873).
  std::vector<size_t> offsets;
  uint64_t stride = 0;
  for (const Prop &p : h.props) {
    offsets.push_back((size_t)stride);
    if (__builtin_add_overflow(stride, (uint64_t)ty_size(p.ty), &stride)) err_malformed();
  }
  if (h.format == Format::Ascii) {
    if (!valid_utf8(body, body_len)) err_unsupported_format();
    uint64_t rows = 0;
    for (Str l : lines(Str{(const char *)body, body_len})) {
      if (rows >= h.vertex_count) break;
      Str t = trim(l);
      if (t.n == 0 || t.starts_with("comment")) continue;
      ++rows;
    }
    if (rows < h.vertex_count) err_count_mismatch();
  } else {
    uint64_t required;
    if (mul_ovf(h.vertex_count, stride, &required)) err_overflow("binary body bytes");
    if (body_len < required) err_count_mismatch();
  }

  // allocate_scene budgets (:307-359): 56 B per splat + 4 B per SH float.
  uint64_t base_bytes, rest_cap, rest_bytes, scene_bytes;
  if (mul_ovf(h.vertex_count, 56, &base_bytes)) err_overflow("base scene bytes");
  rest_cap = 0;
  if (has_rest && mul_ovf(h.vertex_count, rest_stride, &rest_cap)) err_overflow("SH coefficient capacity");
  if (mul_ovf(rest_cap, 4, &rest_bytes)) err_overflow("SH coefficient bytes");
  if (__builtin_add_overflow(base_bytes, rest_bytes, &scene_bytes)) err_overflow("total scene bytes");
  ensure_limit("decoded scene bytes", scene_bytes, lim.max_scene_bytes);

  HostScene s;
  s.count = (uint32_t)h.vertex_count;
  s.sh_degree = degree;
  s.has_sh_rest = has_rest;
  try {
    s.positions.reserve(3 * h.vertex_count);
    s.opacity.reserve(h.vertex_count);
    s.scale_xyz.reserve(3 * h.vertex_count);
    s.rotation_xyzw.reserve(4 * h.vertex_count);
    s.color_dc.reserve(3 * h.vertex_count);
    if (has_rest) s.sh_rest.reserve(rest_cap);
  } catch (const std::bad_alloc &) {
    fail(Code::Internal, "failed to reserve memory for PLY scene buffers");
  }
  const bool xyzw = rot_layout_xyzw();
  const size_t ix = idx["x"], iy = idx["y"], iz = idx["z"], io = idx["opacity"];
  size_t isc[3] = {idx["scale_0"], idx["scale_1"], idx["scale_2"]};
  size_t ir[4] = {idx["rot_0"], idx["rot_1"], idx["rot_2"], idx["rot_3"]};
  size_t idc[3] = {idx["f_dc_0"], idx["f_dc_1"], idx["f_dc_2"]};

  auto push_rot = [&](float r0, float r1, float r2, float r3) {
    // rotation_input_to_xyzw (:63-74): wxyz -> xyzw unless GSPLAT_ROT_LAYOUT=xyzw.
    if (xyzw) {
      s.rotation_xyzw.insert(s.rotation_xyzw.end(), {r0, r1, r2, r3});
    } else {
      s.rotation_xyzw.insert(s.rotation_xyzw.end(), {r1, r2, r3, r0});
    }
  };

  if (h.format == Format::Ascii) {
    // parse_ascii_body (:690-753).
    std::vector<Str> ls = lines(Str{(const char *)body, body_len});
    size_t li = 0;
    for (uint64_t v = 0; v < h.vertex_count; ++v) {
      Str row;
      bool got = false;
      while (li < ls.size()) {
        Str t = trim(ls[li++]);
        if (t.n == 0 || t.starts_with("comment")) continue;
        row = t;
        got = true;
        break;
      }
      if (!got) err_count_mismatch();
      std::vector<Str> vals = split_ws(row);
      if (vals.size() < h.props.size()) err_field_count();
      auto field = [&](size_t i) {
        float f;
        if (i >= vals.size()) err_field_count();
        if (!parse_f32(vals[i], &f) || !std::isfinite(f)) err_parse_number();
        return f;
      };
      float x = field(ix), y = field(iy), z = field(iz);
      s.positions.insert(s.positions.end(), {x, y, z});
      float o;
      if (!parse_f32(vals[io], &o)) err_parse_number();
      s.opacity.push_back(normalize_opacity(o));
      float s0 = field(isc[0]), s1 = field(isc[1]), s2 = field(isc[2]);
      s.scale_xyz.insert(s.scale_xyz.end(), {s0, s1, s2});
      float r0 = field(ir[0]), r1 = field(ir[1]), r2 = field(ir[2]), r3 = field(ir[3]);
      push_rot(r0, r1, r2, r3);
      float d0 = field(idc[0]), d1 = field(idc[1]), d2 = field(idc[2]);
      s.color_dc.insert(s.color_dc.end(), {d0, d1, d2});
      if (has_rest)
        for (size_t p : rest_props) s.sh_rest.push_back(field(p));
    }
  } else {
    // parse_binary_body (:800-861).
    const bool big = h.format == Format::BinBE;
    for (uint64_t v = 0; v < h.vertex_count; ++v) {
      const uint8_t *rec = body + v * stride;
      auto field = [&](size_t i) {
        float f = read_scalar(rec, offsets[i], h.props[i].ty, big);
        if (!std::isfinite(f)) err_parse_number();
        return f;
      };
      float x = field(ix), y = field(iy), z = field(iz);
      s.positions.insert(s.positions.end(), {x, y, z});
      s.opacity.push_back(normalize_opacity(read_scalar(rec, offsets[io], h.props[io].ty, big)));
      float s0 = field(isc[0]), s1 = field(isc[1]), s2 = field(isc[2]);
      s.scale_xyz.insert(s.scale_xyz.end(), {s0, s1, s2});
      float r0 = field(ir[0]), r1 = field(ir[1]), r2 = field(ir[2]), r3 = field(ir[3]);
      push_rot(r0, r1, r2, r3);
      float d0 = field(idc[0]), d1 = field(idc[1]), d2 = field(idc[2]);
      s.color_dc.insert(s.color_dc.end(), {d0, d1, d2});
      if (has_rest)
        for (size_t p : rest_props) s.sh_rest.push_back(field(p));
    }
  }

  // convert_scene_rdf_to_ruf (:76-114).
  for (uint32_t i = 0; i < s.count; ++i) {
    s.positions[3 * i + 1] = -s.positions[3 * i + 1];
    s.rotation_xyzw[4 * i + 0] = -s.rotation_xyzw[4 * i + 0];
    s.rotation_xyzw[4 * i + 2] = -s.rotation_xyzw[4 * i + 2];
  }
  if (has_rest) {
    uint32_t pc = sh_per_channel(degree);
    for (uint32_t i = 0; i < s.count; ++i)
      for (uint32_t ch = 0; ch < 3; ++ch)
        for (uint32_t k = 0; k < pc; ++k) s.sh_rest[(size_t)i * 3 * pc + ch * pc + k] *= (k < 15 ? kShFlip[k] : 1.0f);
  }
  try {
    validate_scene(s);
  } catch (const Error &) {
    err_invalid_scene();
  }
  return s;
}

// load_ply_with_limits + read_ply_bytes_with_limits (io-ply/src/lib.rs:183-305).
HostScene ply_load(const std::string &path, const PlyLimits &lim) {
  struct stat st;
  if (::stat(path.c_str(), &st) != 0 || S_ISDIR(st.st_mode)) err_io();
  ensure_limit("input bytes", (uint64_t)st.st_size, lim.max_input_bytes);
  std::ifstream f(path, std::ios::binary);
  if (!f) err_io();
  std::vector<uint8_t> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (f.bad()) err_io();
  ensure_limit("input bytes", raw.size(), lim.max_input_bytes);
  return ply_parse(raw.data(), raw.size(), lim, false);
}

}  // namespace gs
