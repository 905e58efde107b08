// STL ingestion: binary / ASCII STL bytes to a triangle soup, and the centre-and-rescale normalisation the
// reference applies to uploaded meshes (src/geometry/stlLoader.ts:9-127).
//
// Format decision (stlLoader.ts:9-38): under 84 bytes the bytes are ASCII; a face count at offset 80 whose
// 84 + 50*count equals the length selects binary; otherwise a header that starts with "solid" (after leading
// whitespace, case-insensitive, within the first 256 bytes) and a text holding both "facet" and "vertex"
// selects ASCII; otherwise a face count that fits and is non-zero selects binary; otherwise the bytes are
// rejected.
//
// ASCII vertices (stlLoader.ts:78-95) are the matches of the regular expression
//   vertex\s+N\s+N\s+N,  N = [+-]?\d*\.?\d+(?:[eE][+-]?\d+)?
// with ECMAScript's \s over the UTF-8 text. The matcher below works on the bytes: ASCII bytes and complete
// UTF-8 whitespace sequences decode to themselves whatever surrounds them, so matching on bytes finds the
// same matches as matching on the decoded text. Each number is parsed to the nearest double and then
// rounded to float, as parseFloat followed by a Float32Array store does.
#include <clocale>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "geometry.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "binary STL fields are read as little-endian");

namespace {

typedef const uint8_t* Cur;

// Length of the ECMAScript whitespace / line terminator at p (0 if none): \t \n \v \f \r space, U+00A0,
// U+1680, U+2000..U+200A, U+2028, U+2029, U+202F, U+205F, U+3000, U+FEFF.
int ws_len(Cur p, Cur end) {
  const uint8_t c = *p;
  if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
  const ptrdiff_t n = end - p;
  if (c == 0xC2) return (n >= 2 && p[1] == 0xA0) ? 2 : 0;
  if (n < 3) return 0;
  const uint8_t a = p[1], b = p[2];
  switch (c) {
    case 0xE1: return (a == 0x9A && b == 0x80) ? 3 : 0;
    case 0xE2:
      if (a == 0x80 && ((b >= 0x80 && b <= 0x8A) || b == 0xA8 || b == 0xA9 || b == 0xAF)) return 3;
      return (a == 0x81 && b == 0x9F) ? 3 : 0;
    case 0xE3: return (a == 0x80 && b == 0x80) ? 3 : 0;
    case 0xEF: return (a == 0xBB && b == 0xBF) ? 3 : 0;
    default: return 0;
  }
}

// \s+ at p: the end of the whitespace run, or nullptr when p is not whitespace.
Cur skip_ws1(Cur p, Cur end) {
  if (p >= end) return nullptr;
  int l = ws_len(p, end);
  if (!l) return nullptr;
  while (p < end && (l = ws_len(p, end)) > 0) p += l;
  return p;
}

inline bool digit(Cur p, Cur end) { return p < end && *p >= '0' && *p <= '9'; }
inline Cur digits(Cur p, Cur end) {
  while (digit(p, end)) ++p;
  return p;
}

// The match of N at p that a backtracking matcher reports first when nothing has to follow it: the sign if
// present; then all leading digits, a point and the digits after it when at least one follows, otherwise
// the leading digits alone (at least one); then an exponent when [eE][+-]?digit+ follows. nullptr: no match.
Cur number_first(Cur p, Cur end) {
  if (p < end && (*p == '+' || *p == '-')) ++p;
  Cur d = digits(p, end);
  Cur q;
  if (d < end && *d == '.' && digit(d + 1, end)) q = digits(d + 1, end);
  else if (d > p) q = d;
  else return nullptr;
  if (q < end && (*q == 'e' || *q == 'E')) {
    Cur e = q + 1;
    if (e < end && (*e == '+' || *e == '-')) ++e;
    if (digit(e, end)) q = digits(e, end);
  }
  return q;
}

// N followed by \s+: the match must end at the first whitespace after p, so the whole token [p, ws) must
// be a number. Returns the token's end, or nullptr.
Cur number_before_ws(Cur p, Cur end) {
  Cur t = p;
  while (t < end && !ws_len(t, end)) ++t;
  if (t == end || t == p) return nullptr;  // \s+ must follow
  Cur a = p;
  if (*a == '+' || *a == '-') ++a;
  Cur d = digits(a, end);
  Cur m;
  if (d < t && *d == '.') {
    m = digits(d + 1, end);
    if (m == d + 1) return nullptr;  // a point needs a digit after it
  } else {
    if (d == a) return nullptr;
    m = d;
  }
  if (m < t && (*m == 'e' || *m == 'E')) {
    Cur e = m + 1;
    if (e < t && (*e == '+' || *e == '-')) ++e;
    Cur ed = digits(e, end);
    if (ed == e) return nullptr;
    m = ed;
  }
  return m == t ? t : nullptr;
}

locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

float to_float(Cur a, Cur b) {
  char buf[64];
  std::string big;
  const size_t n = (size_t)(b - a);
  char* s = buf;
  if (n >= sizeof(buf)) {
    big.assign((const char*)a, n);
    s = &big[0];
  } else {
    std::memcpy(buf, a, n);
    buf[n] = 0;
  }
  const double v = strtod_l(s, nullptr, c_locale());  // nearest double, +-inf on overflow, subnormals kept
  return (float)v;
}

bool contains(Cur p, size_t n, const char* w) {
  const size_t m = std::strlen(w);
  if (n < m) return false;
  for (size_t i = 0; i + m <= n; ++i)
    if (std::memcmp(p + i, w, m) == 0) return true;
  return false;
}

// parseAsciiSTL (stlLoader.ts:78-95): every non-overlapping match of the vertex pattern, left to right.
int parse_ascii(Cur p, size_t n, std::vector<float>& pos, std::string& err) {
  Cur end = p + n;
  Cur i = p;
  while (end - i >= 6) {
    if (std::memcmp(i, "vertex", 6) != 0) {
      ++i;
      continue;
    }
    Cur q = skip_ws1(i + 6, end);
    Cur n1 = q ? number_before_ws(q, end) : nullptr;
    Cur q2 = n1 ? skip_ws1(n1, end) : nullptr;
    Cur n2 = q2 ? number_before_ws(q2, end) : nullptr;
    Cur q3 = n2 ? skip_ws1(n2, end) : nullptr;
    Cur n3 = q3 ? number_first(q3, end) : nullptr;
    if (!n3) {
      ++i;
      continue;
    }
    pos.push_back(to_float(q, n1));
    pos.push_back(to_float(q2, n2));
    pos.push_back(to_float(q3, n3));
    i = n3;
  }
  if (pos.size() < 9 || pos.size() % 9 != 0) {
    err = "ASCII STL file contains no valid triangles.";
    return AERO_E_FORMAT;
  }
  return AERO_OK;
}

uint32_t u32le(Cur p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

// Binary layout: 80-byte header, u32 face count, then per face 12 B normal, 3 x 12 B vertices, u16.
void parse_binary(Cur p, uint64_t faces, float* out) {
  for (uint64_t f = 0; f < faces; ++f) std::memcpy(out + f * 9, p + 84 + f * 50 + 12, 36);
}

bool header_says_solid(Cur p, size_t n) {
  const size_t lim = n < 256 ? n : 256;
  Cur end = p + lim;
  Cur i = p;
  int l;
  while (i < end && (l = ws_len(i, end)) > 0) i += l;
  if (end - i < 5) return false;
  const char* w = "solid";
  for (int k = 0; k < 5; ++k) {
    uint8_t c = i[k];
    if (c >= 'A' && c <= 'Z') c = (uint8_t)(c - 'A' + 'a');
    if (c != (uint8_t)w[k]) return false;
  }
  return true;
}

enum Kind { ASCII, BINARY, INVALID };

Kind classify(Cur p, size_t n, uint64_t* faces) {
  if (n < 84) return ASCII;
  const uint64_t fc = u32le(p + 80);
  const uint64_t expected = 84 + fc * 50;
  *faces = fc;
  if (expected == n) return BINARY;
  if (header_says_solid(p, n) && contains(p, n, "facet") && contains(p, n, "vertex")) return ASCII;
  if (expected <= n && fc > 0) return BINARY;
  return INVALID;
}

}  // namespace

extern "C" {

int aero_stl_parse(const uint8_t* bytes, int64_t len, float* out, int64_t cap, int64_t* ntri) {
  if (!ntri || len < 0 || (len > 0 && !bytes)) {
    aero::set_host_error("aero_stl_parse: null argument or negative length");
    return AERO_E_ARG;
  }
  static const uint8_t empty = 0;
  Cur p = len ? bytes : &empty;
  uint64_t faces = 0;
  const Kind k = classify(p, (size_t)len, &faces);
  if (k == INVALID) {
    aero::set_host_error("Invalid or corrupted STL file.");
    return AERO_E_FORMAT;
  }
  if (k == BINARY) {
    *ntri = (int64_t)faces;
    if (!out) return AERO_OK;
    if (cap < (int64_t)faces) {
      aero::set_host_error("aero_stl_parse: output capacity below the triangle count");
      return AERO_E_ARG;
    }
    parse_binary(p, faces, out);
    return AERO_OK;
  }
  std::vector<float> pos;
  std::string err;
  const int rc = parse_ascii(p, (size_t)len, pos, err);
  if (rc != AERO_OK) {
    aero::set_host_error(err);
    return rc;
  }
  *ntri = (int64_t)(pos.size() / 9);
  if (!out) return AERO_OK;
  if (cap < *ntri) {
    aero::set_host_error("aero_stl_parse: output capacity below the triangle count");
    return AERO_E_ARG;
  }
  std::memcpy(out, pos.data(), pos.size() * sizeof(float));
  return AERO_OK;
}

// normalizeAndCenterGeometry (stlLoader.ts:101-127) on a triangle soup, in place. The bounding box is taken
// on the float positions in double; the translation and the scale are THREE.Matrix4 transforms applied by
// Vector3.applyMatrix4 (the full row dot product plus the projective w, in double) and stored back to
// float. The size that decides the rescale is measured on the translated positions, because the reference's
// box object is recomputed by the translation.
int aero_stl_normalize(float* tri, int64_t ntri, int32_t* scaled_from_mm) {
  if (ntri < 0 || (ntri > 0 && !tri)) {
    aero::set_host_error("aero_stl_normalize: null argument or negative count");
    return AERO_E_ARG;
  }
  const size_t nv = (size_t)ntri * 3;
  auto box = [&](double* mn, double* mx) {
    for (int k = 0; k < 3; ++k) {
      mn[k] = INFINITY;
      mx[k] = -INFINITY;
    }
    for (size_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) {
        const double v = tri[i * 3 + k];
        mn[k] = aero::js_min(mn[k], v);
        mx[k] = aero::js_max(mx[k], v);
      }
  };
  auto empty = [](const double* mn, const double* mx) { return mx[0] < mn[0] || mx[1] < mn[1] || mx[2] < mn[2]; };
  // column-major 4x4, as THREE.Matrix4.elements
  auto apply = [&](const double* e) {
    for (size_t i = 0; i < nv; ++i) {
      const double x = tri[i * 3], y = tri[i * 3 + 1], z = tri[i * 3 + 2];
      const double w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
      tri[i * 3] = (float)((e[0] * x + e[4] * y + e[8] * z + e[12]) * w);
      tri[i * 3 + 1] = (float)((e[1] * x + e[5] * y + e[9] * z + e[13]) * w);
      tri[i * 3 + 2] = (float)((e[2] * x + e[6] * y + e[10] * z + e[14]) * w);
    }
  };
  double mn[3], mx[3];
  box(mn, mx);
  double c[3] = {0, 0, 0};
  if (!empty(mn, mx))
    for (int k = 0; k < 3; ++k) c[k] = (mn[k] + mx[k]) * 0.5;
  const double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -c[0], -c[1], -c[2], 1};
  apply(T);
  box(mn, mx);
  double size[3] = {0, 0, 0};
  if (!empty(mn, mx))
    for (int k = 0; k < 3; ++k) size[k] = mx[k] - mn[k];
  const double maxDim = aero::js_max(aero::js_max(size[0], size[1]), size[2]);
  int scaled = 0;
  if (maxDim > 10.0) {
    const double s = maxDim > 50 ? 0.001 : 1.0 / maxDim;
    const double S[16] = {s, 0, 0, 0, 0, s, 0, 0, 0, 0, s, 0, 0, 0, 0, 1};
    apply(S);
    scaled = 1;
  }
  if (scaled_from_mm) *scaled_from_mm = scaled;
  return AERO_OK;
}

}  // extern "C"
