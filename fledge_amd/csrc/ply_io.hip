// Mesh ingestion: MakeTriangleMesh (src/plymesh.cpp:19-86) without the
// resource allocator.  The reference reads PLY files through miniply (a
// third-party reader fetched at configure time, not part of its tree); this
// reader follows the PLY format as Greg Turk's 1994 description defines it
// and takes from the file what MakeTriangleMesh takes:
//   * the "vertex" element's x, y, z as float (find_pos), nx, ny, nz when all
//     three are present (find_normal), and one texture-coordinate pair (u/v,
//     s/t, texture_u/texture_v or texture_s/texture_t; find_texcoord);
//   * the "face" element's "vertex_indices" lists as int, three per face
//     (convert_list_to_fixed_size with 3); a face of another size is an error
//     here;
//   * other elements and properties are skipped.
// Formats ascii 1.0, binary_little_endian 1.0 and binary_big_endian 1.0;
// scalar types char/int8, uchar/uint8, short/int16, ushort/uint16,
// int/int32, uint/uint32, float/float32, double/float64, converted to the
// destination type as C does (float from double rounds to nearest).  Indices
// outside [0, vertex count) are an error (the reference asserts them in debug
// builds, src/shape.cpp:111-113).  Host code only.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace flg {

namespace {

enum PType { kI8, kU8, kI16, kU16, kI32, kU32, kF32, kF64, kBad };

PType parse_type(const std::string& s) {
  if (s == "char" || s == "int8") return kI8;
  if (s == "uchar" || s == "uint8") return kU8;
  if (s == "short" || s == "int16") return kI16;
  if (s == "ushort" || s == "uint16") return kU16;
  if (s == "int" || s == "int32") return kI32;
  if (s == "uint" || s == "uint32") return kU32;
  if (s == "float" || s == "float32") return kF32;
  if (s == "double" || s == "float64") return kF64;
  return kBad;
}

int type_size(PType t) {
  switch (t) {
    case kI8: case kU8: return 1;
    case kI16: case kU16: return 2;
    case kI32: case kU32: case kF32: return 4;
    case kF64: return 8;
    default: return 0;
  }
}

struct Prop {
  std::string name;
  PType type = kBad;
  bool list = false;
  PType count_type = kBad;
};

struct Element {
  std::string name;
  int64_t count = 0;
  std::vector<Prop> props;
};

enum Format { kAscii, kLE, kBE };

// A value of any scalar type, kept exactly until converted.
struct Value {
  bool is_float = false;
  double f = 0;
  int64_t i = 0;
  float as_float() const { return is_float ? (float)f : (float)i; }
  bool as_int(int64_t* out) const {
    if (is_float) {
      // within int64's range (NaN fails) before the cast, then integral
      if (!(f >= -9.0e18 && f <= 9.0e18) || f != (double)(int64_t)f) return false;
      *out = (int64_t)f;
      return true;
    }
    *out = i;
    return true;
  }
};

struct Reader {
  const char* p;
  const char* end;
  Format fmt;
  bool ok = true;

  // ascii: the next whitespace-separated token
  bool token(std::string& t) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
    const char* s = p;
    while (p < end && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
    t.assign(s, p);
    return !t.empty();
  }

  Value read(PType t) {
    Value v;
    if (fmt == kAscii) {
      std::string tok;
      if (!token(tok)) {
        ok = false;
        return v;
      }
      char* e = nullptr;
      errno = 0;
      if (t == kF32) {
        v.is_float = true;
        v.f = (double)std::strtof(tok.c_str(), &e);  // the float the text names
      } else if (t == kF64) {
        v.is_float = true;
        v.f = std::strtod(tok.c_str(), &e);
      } else {
        v.i = std::strtoll(tok.c_str(), &e, 10);
      }
      // ERANGE on a float is underflow (a subnormal or zero result, which
      // the text legitimately names) unless the value overflowed to inf
      const bool range = errno == ERANGE && (!v.is_float || std::isinf(v.f));
      if (!e || *e != '\0' || range) ok = false;
      return v;
    }
    const int n = type_size(t);
    if (end - p < n) {
      ok = false;
      return v;
    }
    unsigned char b[8];
    for (int k = 0; k < n; ++k) b[k] = (unsigned char)p[fmt == kLE ? k : n - 1 - k];
    p += n;
    switch (t) {
      case kI8: { int8_t x; std::memcpy(&x, b, 1); v.i = x; break; }
      case kU8: { uint8_t x; std::memcpy(&x, b, 1); v.i = x; break; }
      case kI16: { int16_t x; std::memcpy(&x, b, 2); v.i = x; break; }
      case kU16: { uint16_t x; std::memcpy(&x, b, 2); v.i = x; break; }
      case kI32: { int32_t x; std::memcpy(&x, b, 4); v.i = x; break; }
      case kU32: { uint32_t x; std::memcpy(&x, b, 4); v.i = x; break; }
      case kF32: { float x; std::memcpy(&x, b, 4); v.is_float = true; v.f = x; break; }
      case kF64: { double x; std::memcpy(&x, b, 8); v.is_float = true; v.f = x; break; }
      default: ok = false;
    }
    return v;
  }
};

int find(const Element& e, const char* name) {
  for (size_t k = 0; k < e.props.size(); ++k)
    if (!e.props[k].list && e.props[k].name == name) return (int)k;
  return -1;
}

int load(const char* path, flg_ply_mesh* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return FLG_ERR_INVALID;
  std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  // header
  size_t pos = 0;
  auto line = [&](std::string& l) {
    if (pos >= data.size()) return false;
    size_t e = data.find('\n', pos);
    if (e == std::string::npos) e = data.size();
    l = data.substr(pos, e - pos);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    pos = e + 1;
    return true;
  };
  std::string l;
  if (!line(l) || l != "ply") return FLG_ERR_INVALID;
  Format fmt = kAscii;
  bool have_format = false, have_end = false;
  std::vector<Element> elems;
  while (line(l)) {
    std::istringstream ss(l);
    std::string kw;
    ss >> kw;
    if (kw.empty() || kw == "comment" || kw == "obj_info") continue;
    if (kw == "format") {
      std::string f1, ver;
      ss >> f1 >> ver;
      if (ver != "1.0") return FLG_ERR_INVALID;
      if (f1 == "ascii") fmt = kAscii;
      else if (f1 == "binary_little_endian") fmt = kLE;
      else if (f1 == "binary_big_endian") fmt = kBE;
      else return FLG_ERR_INVALID;
      have_format = true;
    } else if (kw == "element") {
      Element e;
      ss >> e.name >> e.count;
      if (ss.fail() || e.count < 0) return FLG_ERR_INVALID;
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) return FLG_ERR_INVALID;
      Prop pr;
      std::string t;
      ss >> t;
      if (t == "list") {
        std::string ct, it;
        ss >> ct >> it >> pr.name;
        pr.list = true;
        pr.count_type = parse_type(ct);
        pr.type = parse_type(it);
        if (pr.count_type == kBad || pr.count_type == kF32 || pr.count_type == kF64) return FLG_ERR_INVALID;
      } else {
        pr.type = parse_type(t);
        ss >> pr.name;
      }
      if (pr.type == kBad || pr.name.empty()) return FLG_ERR_INVALID;
      elems.back().props.push_back(pr);
    } else if (kw == "end_header") {
      have_end = true;
      break;
    } else {
      return FLG_ERR_INVALID;
    }
  }
  if (!have_format || !have_end) return FLG_ERR_INVALID;
  const Element* ve = nullptr;
  const Element* fe = nullptr;
  for (const Element& e : elems) {
    if (e.name == "vertex" && !ve) ve = &e;
    if (e.name == "face" && !fe) fe = &e;
  }
  if (!ve || !fe) return FLG_ERR_INVALID;
  const int ix = find(*ve, "x"), iy = find(*ve, "y"), iz = find(*ve, "z");
  if (ix < 0 || iy < 0 || iz < 0) return FLG_ERR_INVALID;
  int inx = find(*ve, "nx"), iny = find(*ve, "ny"), inz = find(*ve, "nz");
  const bool normals = inx >= 0 && iny >= 0 && inz >= 0;
  int iu = -1, iv = -1;
  const char* uv_names[4][2] = {{"u", "v"}, {"s", "t"}, {"texture_u", "texture_v"}, {"texture_s", "texture_t"}};
  for (auto& nm : uv_names) {
    if (find(*ve, nm[0]) >= 0 && find(*ve, nm[1]) >= 0) {
      iu = find(*ve, nm[0]);
      iv = find(*ve, nm[1]);
      break;
    }
  }
  int ifl = -1;
  for (size_t k = 0; k < fe->props.size(); ++k)
    if (fe->props[k].list && fe->props[k].name == "vertex_indices") ifl = (int)k;
  if (ifl < 0) return FLG_ERR_INVALID;
  if (ve->count >= ((int64_t)1 << 31) || fe->count >= ((int64_t)1 << 30)) return FLG_ERR_RANGE;
  // A row takes at least one byte per value in ascii (a token) and the
  // values' sizes in binary; counts the body cannot hold are an error before
  // any buffer is sized from them.
  auto min_row = [&](const Element& e) {
    int64_t b = 0;
    for (const Prop& pr : e.props)
      b += fmt == kAscii ? 1 : type_size(pr.list ? pr.count_type : pr.type);
    return b;
  };
  const int64_t body = (int64_t)(data.size() - std::min(pos, data.size()));
  for (const Element* e : {ve, fe})
    if (e->count * min_row(*e) > body) return FLG_ERR_INVALID;
  std::vector<float> verts((size_t)ve->count * 3), nrm(normals ? (size_t)ve->count * 3 : 0),
      uv(iu >= 0 ? (size_t)ve->count * 2 : 0);
  std::vector<int32_t> ind((size_t)fe->count * 3);
  Reader R{data.data() + pos, data.data() + data.size(), fmt};
  std::vector<Value> row;
  for (const Element& e : elems) {
    const bool is_v = &e == ve, is_f = &e == fe;
    for (int64_t r = 0; r < e.count; ++r) {
      row.assign(e.props.size(), Value{});
      for (size_t k = 0; k < e.props.size(); ++k) {
        const Prop& pr = e.props[k];
        if (!pr.list) {
          row[k] = R.read(pr.type);
          continue;
        }
        int64_t cnt = 0;
        if (!R.read(pr.count_type).as_int(&cnt) || cnt < 0) return FLG_ERR_INVALID;
        if (is_f && (int)k == ifl && cnt != 3) return FLG_ERR_INVALID;
        // a list item takes at least one byte in ascii and its type's size
        // in binary: a count the rest of the body cannot hold is an error
        // before the items are read
        const int64_t item = fmt == kAscii ? 1 : type_size(pr.type);
        if (cnt > (int64_t)(R.end - R.p) / item) return FLG_ERR_INVALID;
        for (int64_t c = 0; c < cnt; ++c) {
          const Value v = R.read(pr.type);
          if (!R.ok) return FLG_ERR_INVALID;
          if (is_f && (int)k == ifl) {
            int64_t idx = 0;
            if (!v.as_int(&idx) || idx < 0 || idx >= ve->count) return FLG_ERR_INVALID;
            ind[(size_t)r * 3 + (size_t)c] = (int32_t)idx;
          }
        }
        if (!R.ok) return FLG_ERR_INVALID;
      }
      if (!R.ok) return FLG_ERR_INVALID;
      if (is_v) {
        verts[(size_t)r * 3] = row[ix].as_float();
        verts[(size_t)r * 3 + 1] = row[iy].as_float();
        verts[(size_t)r * 3 + 2] = row[iz].as_float();
        if (normals) {
          nrm[(size_t)r * 3] = row[inx].as_float();
          nrm[(size_t)r * 3 + 1] = row[iny].as_float();
          nrm[(size_t)r * 3 + 2] = row[inz].as_float();
        }
        if (iu >= 0) {
          uv[(size_t)r * 2] = row[iu].as_float();
          uv[(size_t)r * 2 + 1] = row[iv].as_float();
        }
      }
    }
  }
  auto copy_out = [](const auto& v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    if (v.empty()) return (T*)nullptr;
    T* m = (T*)std::malloc(v.size() * sizeof(T));
    if (m) std::memcpy(m, v.data(), v.size() * sizeof(T));
    return m;
  };
  out->nverts = (int32_t)ve->count;
  out->ntris = (int32_t)fe->count;
  out->verts = copy_out(verts);
  out->ind = copy_out(ind);
  out->normals = copy_out(nrm);
  out->uv = copy_out(uv);
  if ((!verts.empty() && !out->verts) || (!ind.empty() && !out->ind) || (!nrm.empty() && !out->normals) ||
      (!uv.empty() && !out->uv)) {
    flg_ply_free(out);
    return FLG_ERR_OOM;
  }
  return FLG_OK;
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_ply_load(const char* path, flg_ply_mesh* out) {
  return host_guard(nullptr, "flg_ply_load", [&]() -> int {
    if (!path || !out) return FLG_ERR_INVALID;
    *out = flg_ply_mesh{0, 0, nullptr, nullptr, nullptr, nullptr};
    try {
      return load(path, out);
    } catch (...) {  // allocation failures of the parser's own buffers
      flg_ply_free(out);
      return FLG_ERR_OOM;
    }
  });
}

void flg_ply_free(flg_ply_mesh* m) {
  if (!m) return;
  std::free(m->verts);
  std::free(m->ind);
  std::free(m->normals);
  std::free(m->uv);
  *m = flg_ply_mesh{0, 0, nullptr, nullptr, nullptr, nullptr};
}

}  // extern "C"
