// cn_safetensors.cpp -- the safetensors container on the host
// (LeanSafetensors: Header.lean, Reader.lean, Writer.lean).
//
// Layout: u64 little-endian header length, a UTF-8 JSON header, then the
// tensor bytes with offsets relative to the end of the header.  The header
// maps tensor names to {dtype, shape, data_offsets} plus an optional
// "__metadata__" object of strings.
//
// Reading follows SafeTensors.parse + validate (Reader.lean:33-61): every
// tensor's slice must lie in the body and hold dtype size x element count
// bytes.  The header is decoded the way the reference decodes it: its JSON
// objects become key-sorted maps (a repeated key keeps its last value), so
// tensors and metadata come back in key order; a non-string metadata value
// is kept as its JSON text (Header.lean:168-177); shape and offset entries
// must be non-negative integers (Header.lean:133-152).
//
// Writing follows Writer.layout + encode (Writer.lean:20-58): tensors laid
// out contiguously in the given order, "__metadata__" first when present,
// compact JSON (LeanGltf JSON.render, with its string escapes), and the
// header padded with spaces so the body starts 8-byte aligned.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "curvenet_la.h"

namespace cn {
void set_error(const char *fmt, ...);
}

namespace {

// ---- dtypes (Header.lean:36-66), in the reference's declaration order ----
struct DtypeDesc {
  const char *tag;
  int size;
};
const DtypeDesc kDtypes[] = {{"F32", 4}, {"F64", 8}, {"F16", 2}, {"BF16", 2}, {"U8", 1},
                             {"U16", 2}, {"U32", 4}, {"U64", 8}, {"I8", 1},   {"I16", 2},
                             {"I32", 4}, {"I64", 8}, {"BOOL", 1}};
constexpr int kNumDtypes = sizeof(kDtypes) / sizeof(kDtypes[0]);

int dtype_of(const std::string &s) {
  for (int i = 0; i < kNumDtypes; ++i)
    if (s == kDtypes[i].tag) return i;
  return -1;
}

// ---- a small JSON reader ---------------------------------------------------
struct Json {
  enum Kind { NUL, BOOL, INT, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  int64_t i = 0;
  std::string text;                   // STR: the string; NUM / INT: the source text
  std::vector<Json> arr;
  std::map<std::string, Json> obj;    // key-sorted, last duplicate wins
};

struct Parser {
  const char *p, *e;
  std::string err;

  bool fail(const char *m) {
    if (err.empty()) err = m;
    return false;
  }
  void ws() {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  static void put_utf8(std::string &s, uint32_t c) {
    if (c < 0x80) {
      s.push_back((char)c);
    } else if (c < 0x800) {
      s.push_back((char)(0xC0 | (c >> 6)));
      s.push_back((char)(0x80 | (c & 0x3F)));
    } else if (c < 0x10000) {
      s.push_back((char)(0xE0 | (c >> 12)));
      s.push_back((char)(0x80 | ((c >> 6) & 0x3F)));
      s.push_back((char)(0x80 | (c & 0x3F)));
    } else {
      s.push_back((char)(0xF0 | (c >> 18)));
      s.push_back((char)(0x80 | ((c >> 12) & 0x3F)));
      s.push_back((char)(0x80 | ((c >> 6) & 0x3F)));
      s.push_back((char)(0x80 | (c & 0x3F)));
    }
  }
  bool hex4(uint32_t &v) {
    if (e - p < 4) return fail("truncated \\u escape");
    v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string &s) {
    if (p >= e || *p != '"') return fail("expected a string");
    ++p;
    while (p < e && *p != '"') {
      const unsigned char c = (unsigned char)*p++;
      if (c < 0x20) return fail("control character in a string");
      if (c != '\\') {
        s.push_back((char)c);
        continue;
      }
      if (p >= e) return fail("truncated escape");
      const char x = *p++;
      switch (x) {
        case '"': s.push_back('"'); break;
        case '\\': s.push_back('\\'); break;
        case '/': s.push_back('/'); break;
        case 'b': s.push_back('\b'); break;
        case 'f': s.push_back('\f'); break;
        case 'n': s.push_back('\n'); break;
        case 'r': s.push_back('\r'); break;
        case 't': s.push_back('\t'); break;
        case 'u': {
          uint32_t v;
          if (!hex4(v)) return false;
          if (v >= 0xD800 && v < 0xDC00) {  // surrogate pair
            uint32_t lo;
            if (e - p < 2 || p[0] != '\\' || p[1] != 'u') return fail("lone surrogate");
            p += 2;
            if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return fail("bad surrogate pair");
            v = 0x10000 + ((v - 0xD800) << 10) + (lo - 0xDC00);
          } else if (v >= 0xDC00 && v < 0xE000) {
            return fail("lone surrogate");
          }
          put_utf8(s, v);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p >= e) return fail("unterminated string");
    ++p;
    return true;
  }
  bool number(Json &j) {
    const char *s = p;
    if (p < e && *p == '-') ++p;
    if (p >= e || !(*p >= '0' && *p <= '9')) return fail("bad number");
    if (*p == '0') ++p;
    else while (p < e && *p >= '0' && *p <= '9') ++p;
    bool integer = true;
    if (p < e && *p == '.') {
      integer = false;
      ++p;
      if (p >= e || !(*p >= '0' && *p <= '9')) return fail("bad number");
      while (p < e && *p >= '0' && *p <= '9') ++p;
    }
    if (p < e && (*p == 'e' || *p == 'E')) {
      integer = false;
      ++p;
      if (p < e && (*p == '+' || *p == '-')) ++p;
      if (p >= e || !(*p >= '0' && *p <= '9')) return fail("bad number");
      while (p < e && *p >= '0' && *p <= '9') ++p;
    }
    j.text.assign(s, p);
    j.kind = Json::NUM;
    if (integer) {
      // an integer that fits int64 is an INT; longer ones stay NUM text
      errno = 0;
      char *end = nullptr;
      const long long v = strtoll(j.text.c_str(), &end, 10);
      if (errno == 0 && end && *end == 0) {
        j.kind = Json::INT;
        j.i = v;
      }
    }
    return true;
  }
  bool lit(const char *w) {
    const size_t n = strlen(w);
    if ((size_t)(e - p) < n || memcmp(p, w, n) != 0) return fail("bad literal");
    p += n;
    return true;
  }
  bool value(Json &j, int depth) {
    if (depth > 64) return fail("nesting too deep");
    ws();
    if (p >= e) return fail("unexpected end of header");
    const char c = *p;
    if (c == '{') {
      ++p;
      j.kind = Json::OBJ;
      ws();
      if (p < e && *p == '}') {
        ++p;
        return true;
      }
      while (true) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (p >= e || *p != ':') return fail("expected ':'");
        ++p;
        Json v;
        if (!value(v, depth + 1)) return false;
        j.obj[k] = std::move(v);
        ws();
        if (p < e && *p == ',') {
          ++p;
          continue;
        }
        if (p < e && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      j.kind = Json::ARR;
      ws();
      if (p < e && *p == ']') {
        ++p;
        return true;
      }
      while (true) {
        Json v;
        if (!value(v, depth + 1)) return false;
        j.arr.push_back(std::move(v));
        ws();
        if (p < e && *p == ',') {
          ++p;
          continue;
        }
        if (p < e && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      j.kind = Json::STR;
      return string(j.text);
    }
    if (c == 't') {
      j.kind = Json::BOOL;
      j.b = true;
      return lit("true");
    }
    if (c == 'f') {
      j.kind = Json::BOOL;
      return lit("false");
    }
    if (c == 'n') {
      j.kind = Json::NUL;
      return lit("null");
    }
    return number(j);
  }
};

// ---- JSON writer (LeanGltf JSON.render / escapeString) ---------------------
void escape(std::string &o, const std::string &s) {
  static const char *hex = "0123456789abcdef";
  o.push_back('"');
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case 0x08: o += "\\b"; break;
      case 0x0c: o += "\\f"; break;
      default:
        if (c < 0x20) {
          o += "\\u00";
          o.push_back(hex[c >> 4]);
          o.push_back(hex[c & 15]);
        } else {
          o.push_back((char)c);
        }
    }
  }
  o.push_back('"');
}

void render(std::string &o, const Json &j) {
  switch (j.kind) {
    case Json::NUL: o += "null"; break;
    case Json::BOOL: o += j.b ? "true" : "false"; break;
    case Json::INT: o += std::to_string(j.i); break;
    case Json::NUM: o += j.text; break;
    case Json::STR: escape(o, j.text); break;
    case Json::ARR:
      o.push_back('[');
      for (size_t k = 0; k < j.arr.size(); ++k) {
        if (k) o.push_back(',');
        render(o, j.arr[k]);
      }
      o.push_back(']');
      break;
    case Json::OBJ: {
      o.push_back('{');
      bool first = true;
      for (const auto &kv : j.obj) {
        if (!first) o.push_back(',');
        first = false;
        escape(o, kv.first);
        o.push_back(':');
        render(o, kv.second);
      }
      o.push_back('}');
      break;
    }
  }
}

}  // namespace

struct cn_st_t {
  std::vector<unsigned char> body;
  std::vector<std::string> names, mkeys, mvals;
  std::vector<int32_t> dtype;
  std::vector<std::vector<int64_t>> shape;
  std::vector<int64_t> lo, hi;
};

namespace {

bool decode_uint(const Json &j, int64_t &v) {
  if (j.kind != Json::INT || j.i < 0) return false;
  v = j.i;
  return true;
}

cn_status parse_into(const unsigned char *b, int64_t size, cn_st_t &st) {
  if (size < 8) {
    cn::set_error("safetensors: file < 8 bytes");
    return CN_BAD_ARG;
  }
  uint64_t hlen = 0;
  for (int k = 0; k < 8; ++k) hlen |= (uint64_t)b[k] << (8 * k);
  if (hlen > (uint64_t)size - 8) {
    cn::set_error("safetensors: header length %llu exceeds file size %lld",
                  (unsigned long long)hlen, (long long)size);
    return CN_BAD_ARG;
  }
  Parser ps{(const char *)b + 8, (const char *)b + 8 + hlen, {}};
  Json root;
  if (!ps.value(root, 0)) {
    cn::set_error("safetensors: header JSON: %s", ps.err.c_str());
    return CN_BAD_ARG;
  }
  ps.ws();
  if (ps.p != ps.e) {
    cn::set_error("safetensors: header JSON: trailing characters");
    return CN_BAD_ARG;
  }
  if (root.kind != Json::OBJ) {
    cn::set_error("safetensors: header root is not an object");
    return CN_BAD_ARG;
  }
  for (const auto &kv : root.obj) {
    const std::string &k = kv.first;
    const Json &v = kv.second;
    if (k == "__metadata__") {
      if (v.kind != Json::OBJ) {
        cn::set_error("safetensors: __metadata__ is not an object");
        return CN_BAD_ARG;
      }
      for (const auto &m : v.obj) {
        st.mkeys.push_back(m.first);
        if (m.second.kind == Json::STR) {
          st.mvals.push_back(m.second.text);
        } else {
          std::string t;
          render(t, m.second);
          st.mvals.push_back(t);
        }
      }
      continue;
    }
    if (v.kind != Json::OBJ) {
      cn::set_error("safetensors: tensor entry is not an object");
      return CN_BAD_ARG;
    }
    auto dt = v.obj.find("dtype");
    if (dt == v.obj.end() || dt->second.kind != Json::STR) {
      cn::set_error("safetensors: tensor entry missing/invalid dtype");
      return CN_BAD_ARG;
    }
    const int d = dtype_of(dt->second.text);
    if (d < 0) {
      cn::set_error("safetensors: unknown dtype \"%s\"", dt->second.text.c_str());
      return CN_BAD_ARG;
    }
    auto sh = v.obj.find("shape");
    if (sh == v.obj.end() || sh->second.kind != Json::ARR) {
      cn::set_error("safetensors: tensor entry missing/invalid shape");
      return CN_BAD_ARG;
    }
    std::vector<int64_t> shape;
    for (const Json &x : sh->second.arr) {
      int64_t s;
      if (!decode_uint(x, s)) {
        cn::set_error("safetensors: negative or non-integer shape entry");
        return CN_BAD_ARG;
      }
      shape.push_back(s);
    }
    auto of = v.obj.find("data_offsets");
    if (of == v.obj.end() || of->second.kind != Json::ARR) {
      cn::set_error("safetensors: tensor entry missing/invalid data_offsets");
      return CN_BAD_ARG;
    }
    if (of->second.arr.size() != 2) {
      cn::set_error("safetensors: data_offsets must have 2 entries");
      return CN_BAD_ARG;
    }
    int64_t lo, hi;
    if (!decode_uint(of->second.arr[0], lo) || !decode_uint(of->second.arr[1], hi)) {
      cn::set_error("safetensors: negative or non-integer data_offsets entry");
      return CN_BAD_ARG;
    }
    st.names.push_back(k);
    st.dtype.push_back(d);
    st.shape.push_back(shape);
    st.lo.push_back(lo);
    st.hi.push_back(hi);
  }
  const int64_t bsize = size - 8 - (int64_t)hlen;
  // SafeTensors.validate (Reader.lean:33-45)
  for (size_t t = 0; t < st.names.size(); ++t) {
    if (st.lo[t] > st.hi[t]) {
      cn::set_error("safetensors: tensor %s has lo > hi", st.names[t].c_str());
      return CN_BAD_ARG;
    }
    if (st.hi[t] > bsize) {
      cn::set_error("safetensors: tensor %s extends beyond body (%lld > %lld)",
                    st.names[t].c_str(), (long long)st.hi[t], (long long)bsize);
      return CN_BAD_ARG;
    }
    // dtype size x element count, without overflow
    __int128 expect = kDtypes[st.dtype[t]].size;
    for (int64_t s : st.shape[t]) {
      expect *= s;
      if (expect > ((__int128)1 << 62)) break;
    }
    if ((__int128)(st.hi[t] - st.lo[t]) != expect) {
      cn::set_error("safetensors: tensor %s size %lld != dtype.shape product",
                    st.names[t].c_str(), (long long)(st.hi[t] - st.lo[t]));
      return CN_BAD_ARG;
    }
  }
  st.body.assign(b + 8 + hlen, b + size);
  return CN_OK;
}

}  // namespace

// host allocation failures become a status, not an exception through C
#define ST_TRY try {
#define ST_CATCH                                                              \
  }                                                                           \
  catch (...) {                                                               \
    cn::set_error("safetensors: host allocation failed");                     \
    return CN_HIP_ERROR;                                                      \
  }

extern "C" {

cn_status cn_st_parse(const void *bytes, int64_t size, cn_st *out) {
  if (!out || (size > 0 && !bytes) || size < 0) {
    cn::set_error("cn_st_parse: bad arguments");
    return CN_BAD_ARG;
  }
  ST_TRY
  auto st = std::make_unique<cn_st_t>();
  const cn_status r = parse_into((const unsigned char *)bytes, size, *st);
  if (r != CN_OK) return r;
  *out = st.release();
  return CN_OK;
  ST_CATCH
}

cn_status cn_st_read_file(const char *path, cn_st *out) {
  if (!path || !out) {
    cn::set_error("cn_st_read_file: bad arguments");
    return CN_BAD_ARG;
  }
  ST_TRY
  FILE *f = fopen(path, "rb");
  if (!f) {
    cn::set_error("cn_st_read_file: cannot open %s", path);
    return CN_BAD_ARG;
  }
  std::vector<unsigned char> buf;
  unsigned char tmp[1 << 16];
  size_t n;
  while ((n = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  const bool bad = ferror(f) != 0;
  fclose(f);
  if (bad) {
    cn::set_error("cn_st_read_file: read error on %s", path);
    return CN_BAD_ARG;
  }
  return cn_st_parse(buf.data(), (int64_t)buf.size(), out);
  ST_CATCH
}

int32_t cn_st_count(cn_st st) { return st ? (int32_t)st->names.size() : 0; }

cn_status cn_st_tensor_at(cn_st st, int32_t i, cn_st_tensor *out) {
  if (!st || !out || i < 0 || i >= (int32_t)st->names.size()) {
    cn::set_error("cn_st_tensor_at: bad arguments");
    return CN_BAD_ARG;
  }
  out->name = st->names[i].c_str();
  out->dtype = st->dtype[i];
  out->ndim = (int32_t)st->shape[i].size();
  out->shape = st->shape[i].data();
  out->data = st->body.data() + st->lo[i];
  out->nbytes = st->hi[i] - st->lo[i];
  return CN_OK;
}

int32_t cn_st_find(cn_st st, const char *name) {
  if (!st || !name) return -1;
  for (size_t t = 0; t < st->names.size(); ++t)
    if (st->names[t] == name) return (int32_t)t;
  return -1;
}

int32_t cn_st_metadata_count(cn_st st) { return st ? (int32_t)st->mkeys.size() : 0; }

cn_status cn_st_metadata_at(cn_st st, int32_t i, const char **key, const char **value) {
  if (!st || !key || !value || i < 0 || i >= (int32_t)st->mkeys.size()) {
    cn::set_error("cn_st_metadata_at: bad arguments");
    return CN_BAD_ARG;
  }
  *key = st->mkeys[i].c_str();
  *value = st->mvals[i].c_str();
  return CN_OK;
}

const char *cn_st_metadata_get(cn_st st, const char *key) {
  if (!st || !key) return nullptr;
  for (size_t m = 0; m < st->mkeys.size(); ++m)
    if (st->mkeys[m] == key) return st->mvals[m].c_str();
  return nullptr;
}

void cn_st_free(cn_st st) { delete st; }

cn_status cn_st_encode(int32_t n, const cn_st_tensor *tensors, int32_t nmeta,
                       const char *const *keys, const char *const *values, void *buf,
                       int64_t cap, int64_t *size) {
  if (n < 0 || nmeta < 0 || (n > 0 && !tensors) || (nmeta > 0 && (!keys || !values)) ||
      !size) {
    cn::set_error("cn_st_encode: bad arguments");
    return CN_BAD_ARG;
  }
  ST_TRY
  // Writer.layout: contiguous offsets in input order, sizes checked
  std::string h = "{";
  bool first = true;
  if (nmeta > 0) {
    h += "\"__metadata__\":{";
    for (int32_t m = 0; m < nmeta; ++m) {
      if (!keys[m] || !values[m]) {
        cn::set_error("cn_st_encode: null metadata entry");
        return CN_BAD_ARG;
      }
      if (m) h.push_back(',');
      escape(h, keys[m]);
      h.push_back(':');
      escape(h, values[m]);
    }
    h += "}";
    first = false;
  }
  int64_t off = 0;
  for (int32_t t = 0; t < n; ++t) {
    const cn_st_tensor &x = tensors[t];
    if (!x.name || x.dtype < 0 || x.dtype >= kNumDtypes || x.ndim < 0 ||
        (x.ndim > 0 && !x.shape) || x.nbytes < 0 || (x.nbytes > 0 && !x.data)) {
      cn::set_error("cn_st_encode: bad tensor %d", t);
      return CN_BAD_ARG;
    }
    int64_t expect = kDtypes[x.dtype].size;
    for (int32_t d = 0; d < x.ndim; ++d) {
      if (x.shape[d] < 0) {
        cn::set_error("cn_st_encode: tensor %s has a negative dimension", x.name);
        return CN_BAD_ARG;
      }
      expect *= x.shape[d];
    }
    if (expect != x.nbytes) {
      cn::set_error("safetensors: tensor %s bytes %lld != dtype.shape product %lld", x.name,
                    (long long)x.nbytes, (long long)expect);
      return CN_BAD_ARG;
    }
    if (!first) h.push_back(',');
    first = false;
    escape(h, x.name);
    h += ":{\"dtype\":\"";
    h += kDtypes[x.dtype].tag;
    h += "\",\"shape\":[";
    for (int32_t d = 0; d < x.ndim; ++d) {
      if (d) h.push_back(',');
      h += std::to_string(x.shape[d]);
    }
    h += "],\"data_offsets\":[" + std::to_string(off) + "," + std::to_string(off + x.nbytes) +
         "]}";
    off += x.nbytes;
  }
  h += "}";
  // Writer.encode: pad with spaces so the body starts on an 8-byte boundary
  while ((8 + h.size()) % 8 != 0) h.push_back(' ');
  const int64_t total = 8 + (int64_t)h.size() + off;
  *size = total;
  if (!buf) return CN_OK;
  if (cap < total) {
    cn::set_error("cn_st_encode: buffer of %lld bytes, %lld needed", (long long)cap,
                  (long long)total);
    return CN_BAD_ARG;
  }
  unsigned char *o = (unsigned char *)buf;
  const uint64_t hl = h.size();
  for (int k = 0; k < 8; ++k) o[k] = (unsigned char)(hl >> (8 * k));
  memcpy(o + 8, h.data(), h.size());
  int64_t pos = 8 + (int64_t)h.size();
  for (int32_t t = 0; t < n; ++t) {
    if (tensors[t].nbytes) memcpy(o + pos, tensors[t].data, (size_t)tensors[t].nbytes);
    pos += tensors[t].nbytes;
  }
  return CN_OK;
  ST_CATCH
}

cn_status cn_st_write_file(const char *path, int32_t n, const cn_st_tensor *tensors,
                           int32_t nmeta, const char *const *keys, const char *const *values) {
  if (!path) {
    cn::set_error("cn_st_write_file: bad arguments");
    return CN_BAD_ARG;
  }
  ST_TRY
  int64_t size = 0;
  cn_status r = cn_st_encode(n, tensors, nmeta, keys, values, nullptr, 0, &size);
  if (r != CN_OK) return r;
  std::vector<unsigned char> buf((size_t)size);
  r = cn_st_encode(n, tensors, nmeta, keys, values, buf.data(), size, &size);
  if (r != CN_OK) return r;
  FILE *f = fopen(path, "wb");
  if (!f) {
    cn::set_error("cn_st_write_file: cannot open %s", path);
    return CN_BAD_ARG;
  }
  const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  const bool closed = fclose(f) == 0;
  if (!ok || !closed) {
    cn::set_error("cn_st_write_file: write error on %s", path);
    return CN_BAD_ARG;
  }
  return CN_OK;
  ST_CATCH
}

}  // extern "C"
