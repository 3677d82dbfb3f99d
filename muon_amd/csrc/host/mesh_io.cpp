// OBJ / PLY readers for the .muon `load` command; the import rules are listed
// in mesh_io.h.
#include "mesh_io.h"

#include "numbers.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <unordered_map>
#include <sstream>

namespace muonx {
namespace host {
namespace {

// Builds one mesh: corners with the same attribute key share a vertex. Keys
// are kept in one flat array and found through a hash of their words.
class MeshBuilder {
 public:
  explicit MeshBuilder(std::vector<LoadedMesh> &out) : out_(out) {}
  // key: attribute bits identifying the corner; pos: its position
  int Corner(const std::vector<uint64_t> &key, const std::array<float, 3> &pos) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the key words
    for (uint64_t w : key) h = (h ^ w) * 1099511628211ull;
    auto it = head_.find(h);
    for (int v = it == head_.end() ? -1 : it->second; v >= 0; v = next_[size_t(v)]) {
      const size_t o = off_[size_t(v)], len = off_[size_t(v) + 1] - o;
      if (len == key.size() && std::equal(key.begin(), key.end(), keys_.begin() + long(o))) return v;
    }
    const int id = int(cur_.pos.size());
    cur_.pos.push_back(pos);
    keys_.insert(keys_.end(), key.begin(), key.end());
    off_.push_back(keys_.size());
    next_.push_back(it == head_.end() ? -1 : it->second);
    head_[h] = id;
    return id;
  }
  // Adds a polygon as a fan from its first corner, dropping triangles with
  // two equal corner positions.
  void Polygon(const std::vector<int> &c) {
    for (size_t k = 1; k + 1 < c.size(); ++k) {
      const std::array<int, 3> t{c[0], c[k], c[k + 1]};
      const auto &a = cur_.pos[size_t(t[0])], &b = cur_.pos[size_t(t[1])], &d = cur_.pos[size_t(t[2])];
      if (a == b || b == d || a == d) continue;
      cur_.tris.push_back(t);
    }
  }
  void Finish() {
    if (!cur_.tris.empty()) out_.push_back(std::move(cur_));
    cur_ = LoadedMesh();
    head_.clear();
    keys_.clear();
    off_.assign(1, 0);
    next_.clear();
  }

 private:
  std::vector<LoadedMesh> &out_;
  LoadedMesh cur_;
  std::unordered_map<uint64_t, int> head_;  // key hash -> newest vertex with it
  std::vector<uint64_t> keys_;              // key words of every vertex
  std::vector<size_t> off_{0};              // vertex v's key is keys_[off_[v], off_[v + 1])
  std::vector<int> next_;                   // older vertex with the same hash, or -1
};

uint64_t Bits(float f) {
  if (f == 0.0f) f = 0.0f;  // +0 and -0 are the same value
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  return u;
}

// ------------------------------------------------------------------ OBJ

// Resolves a 1-based (or negative, counted from the end) OBJ index.
bool ObjResolve(long v, size_t count, long &out) {
  if (v == 0) return false;
  const long idx = v > 0 ? v - 1 : long(count) + v;
  if (idx < 0 || size_t(idx) >= count) return false;
  out = idx;
  return true;
}

// Reads an optionally negative decimal integer of up to 9 digits at s.
bool ObjInt(const char *&s, const char *e, long &out) {
  bool neg = false;
  if (s < e && *s == '-') {
    neg = true;
    ++s;
  }
  const char *b = s;
  long v = 0;
  while (s < e && *s >= '0' && *s <= '9' && s - b < 9) v = v * 10 + (*s++ - '0');
  if (s == b || (s < e && *s >= '0' && *s <= '9')) return false;
  out = neg ? -v : v;
  return true;
}

// Splits the corners of an "f" line ("i", "i/j", "i//k", "i/j/k"; 0 marks
// an absent index); false for anything else, which the stream path reads.
bool ObjFaceFast(const char *s, const char *e, std::vector<std::array<long, 3>> &out) {
  out.clear();
  while (true) {
    while (s < e && IsSpace(*s)) ++s;
    if (s == e) return true;
    std::array<long, 3> c{0, 0, 0};
    if (!ObjInt(s, e, c[0])) return false;
    if (s < e && *s == '/') {
      ++s;
      if (s < e && *s != '/' && !ObjInt(s, e, c[1])) return false;
      if (s < e && *s == '/') {
        ++s;
        if (!ObjInt(s, e, c[2])) return false;
      }
    }
    if (s < e && !IsSpace(*s)) return false;
    out.push_back(c);
  }
}

bool ParseObj(const std::string &data, std::vector<LoadedMesh> &out, std::string &err) {
  std::vector<std::array<float, 3>> v, vn;
  std::vector<std::array<float, 3>> vt;
  MeshBuilder mb(out);
  size_t lineno = 0;
  std::vector<int> corners;
  std::vector<std::array<long, 3>> raw;
  std::vector<uint64_t> key;
  // one face corner from its raw OBJ indices (0: absent)
  auto corner = [&](long rv, long rt, long rn) {
    long iv, it = -1, in_ = -1;
    if (!ObjResolve(rv, v.size(), iv) || (rt != 0 && !ObjResolve(rt, vt.size(), it)) ||
        (rn != 0 && !ObjResolve(rn, vn.size(), in_)))
      return false;
    key.clear();
    for (float c : v[size_t(iv)]) key.push_back(Bits(c));
    key.push_back(it >= 0 ? 1u : 0u);
    if (it >= 0)
      for (float c : vt[size_t(it)]) key.push_back(Bits(c));
    key.push_back(in_ >= 0 ? 1u : 0u);
    if (in_ >= 0)
      for (float c : vn[size_t(in_)]) key.push_back(Bits(c));
    corners.push_back(mb.Corner(key, v[size_t(iv)]));
    return true;
  };
  auto face = [&]() {
    if (corners.size() < 3) {
      err = "obj line " + std::to_string(lineno) + ": face with fewer than 3 corners";
      return false;
    }
    mb.Polygon(corners);
    return true;
  };
  for (size_t next = 0; next < data.size();) {
    size_t lb = next, le = data.find('\n', next);
    if (le == std::string::npos) le = data.size();
    next = le + 1;
    ++lineno;
    const char *p = data.data() + lb, *pe = data.data() + le;
    if (const void *h = std::memchr(p, '#', size_t(pe - p))) pe = static_cast<const char *>(h);
    while (p < pe && IsSpace(*p)) ++p;
    if (p == pe) continue;
    // plain "v x y z" and "f ..." lines without a stream
    if (pe - p > 2 && p[0] == 'v' && IsSpace(p[1])) {
      float x[3];
      if (FastNumbers(p + 2, pe, x, 3)) {
        v.push_back({x[0], x[1], x[2]});
        continue;
      }
    } else if (pe - p > 2 && p[0] == 'f' && IsSpace(p[1]) && ObjFaceFast(p + 2, pe, raw)) {
      corners.clear();
      for (const auto &c : raw)
        if (!corner(c[0], c[1], c[2])) {
          err = "obj line " + std::to_string(lineno) + ": bad face index";
          return false;
        }
      if (!face()) return false;
      continue;
    }
    std::istringstream ls(std::string(p, pe));
    std::string cmd;
    if (!(ls >> cmd)) continue;
    if (cmd == "v" || cmd == "vn") {
      std::array<float, 3> q{};
      if (!(ls >> q[0] >> q[1] >> q[2])) {
        err = "obj line " + std::to_string(lineno) + ": bad " + cmd;
        return false;
      }
      (cmd == "v" ? v : vn).push_back(q);
    } else if (cmd == "vt") {
      std::array<float, 3> q{};
      if (!(ls >> q[0])) {
        err = "obj line " + std::to_string(lineno) + ": bad vt";
        return false;
      }
      ls >> q[1] >> q[2];  // optional components
      vt.push_back(q);
    } else if (cmd == "f") {
      corners.clear();
      std::string tok;
      while (ls >> tok) {
        std::string part[3];
        size_t s0 = tok.find('/');
        part[0] = tok.substr(0, s0);
        if (s0 != std::string::npos) {
          const size_t s1 = tok.find('/', s0 + 1);
          part[1] = tok.substr(s0 + 1, s1 == std::string::npos ? std::string::npos : s1 - s0 - 1);
          if (s1 != std::string::npos) part[2] = tok.substr(s1 + 1);
        }
        long iv, it = 0, in_ = 0;
        char *end = nullptr;
        bool ok = !part[0].empty();
        if (ok) iv = std::strtol(part[0].c_str(), &end, 10), ok = *end == '\0';
        if (ok && !part[1].empty()) it = std::strtol(part[1].c_str(), &end, 10), ok = *end == '\0' && it != 0;
        if (ok && !part[2].empty()) in_ = std::strtol(part[2].c_str(), &end, 10), ok = *end == '\0' && in_ != 0;
        if (!ok || !corner(iv, it, in_)) {
          err = "obj line " + std::to_string(lineno) + ": bad face index '" + tok + "'";
          return false;
        }
      }
      if (!face()) return false;
    } else if (cmd == "o" || cmd == "g" || cmd == "usemtl") {
      mb.Finish();
    }
    // other statements (mtllib, s, l, p, ...) carry nothing muon uses
  }
  mb.Finish();
  return true;
}

// ------------------------------------------------------------------ PLY

enum class PlyType { kI8, kU8, kI16, kU16, kI32, kU32, kF32, kF64, kBad };

PlyType PlyTypeOf(const std::string &s) {
  if (s == "char" || s == "int8") return PlyType::kI8;
  if (s == "uchar" || s == "uint8") return PlyType::kU8;
  if (s == "short" || s == "int16") return PlyType::kI16;
  if (s == "ushort" || s == "uint16") return PlyType::kU16;
  if (s == "int" || s == "int32") return PlyType::kI32;
  if (s == "uint" || s == "uint32") return PlyType::kU32;
  if (s == "float" || s == "float32") return PlyType::kF32;
  if (s == "double" || s == "float64") return PlyType::kF64;
  return PlyType::kBad;
}

size_t PlySize(PlyType t) {
  switch (t) {
    case PlyType::kI8: case PlyType::kU8: return 1;
    case PlyType::kI16: case PlyType::kU16: return 2;
    case PlyType::kI32: case PlyType::kU32: case PlyType::kF32: return 4;
    case PlyType::kF64: return 8;
    default: return 0;
  }
}

struct PlyProp {
  std::string name;
  bool list = false;
  PlyType count_type = PlyType::kBad, type = PlyType::kBad;
};

struct PlyElement {
  std::string name;
  size_t count = 0;
  std::vector<PlyProp> props;
};

// Reads values of a PLY body: ASCII tokens or binary in either byte order.
class PlyReader {
 public:
  PlyReader(const std::string &data, size_t pos, int format) : d_(data), p_(pos), fmt_(format) {}
  bool Read(PlyType t, double &out) {
    if (fmt_ == 0) {
      while (p_ < d_.size() && std::isspace(static_cast<unsigned char>(d_[p_]))) ++p_;
      const size_t b = p_;
      while (p_ < d_.size() && !std::isspace(static_cast<unsigned char>(d_[p_]))) ++p_;
      if (b == p_) return false;
      const std::string tok = d_.substr(b, p_ - b);
      char *end = nullptr;
      out = std::strtod(tok.c_str(), &end);
      return *end == '\0';
    }
    const size_t n = PlySize(t);
    if (n == 0 || p_ + n > d_.size()) return false;
    unsigned char b[8];
    std::memcpy(b, d_.data() + p_, n);
    p_ += n;
    if (fmt_ == 2) std::reverse(b, b + n);  // big endian -> host (little endian)
    switch (t) {
      case PlyType::kI8: { int8_t v; std::memcpy(&v, b, 1); out = v; break; }
      case PlyType::kU8: { uint8_t v; std::memcpy(&v, b, 1); out = v; break; }
      case PlyType::kI16: { int16_t v; std::memcpy(&v, b, 2); out = v; break; }
      case PlyType::kU16: { uint16_t v; std::memcpy(&v, b, 2); out = v; break; }
      case PlyType::kI32: { int32_t v; std::memcpy(&v, b, 4); out = v; break; }
      case PlyType::kU32: { uint32_t v; std::memcpy(&v, b, 4); out = v; break; }
      case PlyType::kF32: { float v; std::memcpy(&v, b, 4); out = v; break; }
      case PlyType::kF64: { double v; std::memcpy(&v, b, 8); out = v; break; }
      default: return false;
    }
    return true;
  }
  // Bytes left in the body: every value takes at least one (a binary
  // scalar, or an ASCII token with its separator), so a count above this
  // cannot be read and marks a truncated or corrupt file.
  size_t Remaining() const { return p_ < d_.size() ? d_.size() - p_ : 0; }

 private:
  const std::string &d_;
  size_t p_;
  int fmt_;  // 0 ascii, 1 little endian, 2 big endian
};

// A list count or vertex index read as a double: accepted when it is finite,
// integral and within [0, limit].
bool PlyIndex(double v, double limit, long &out) {
  if (!(v >= 0.0 && v <= limit) || v != std::floor(v)) return false;
  out = long(v);
  return true;
}

bool ParsePly(const std::string &data, std::vector<LoadedMesh> &out, std::string &err) {
  std::istringstream hs(data);
  std::string line;
  if (!std::getline(hs, line) || line.substr(0, 3) != "ply") {
    err = "ply: missing magic";
    return false;
  }
  int fmt = -1;
  std::vector<PlyElement> elems;
  bool ended = false;
  while (std::getline(hs, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string kw;
    if (!(ls >> kw)) continue;
    if (kw == "format") {
      std::string f;
      ls >> f;
      fmt = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
    } else if (kw == "element") {
      PlyElement e;
      if (!(ls >> e.name >> e.count)) {
        err = "ply: bad element line";
        return false;
      }
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) {
        err = "ply: property before element";
        return false;
      }
      PlyProp p;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> p.name;
        p.list = true;
        p.count_type = PlyTypeOf(ct);
        p.type = PlyTypeOf(it);
        if (p.count_type == PlyType::kBad || p.type == PlyType::kBad || p.count_type == PlyType::kF32 ||
            p.count_type == PlyType::kF64) {
          err = "ply: bad list property types";
          return false;
        }
      } else {
        p.type = PlyTypeOf(t);
        ls >> p.name;
        if (p.type == PlyType::kBad) {
          err = "ply: bad property type '" + t + "'";
          return false;
        }
      }
      elems.back().props.push_back(p);
    } else if (kw == "end_header") {
      ended = true;
      break;
    }
    // comment / obj_info lines are skipped
  }
  if (!ended || fmt < 0) {
    err = "ply: incomplete header or unknown format";
    return false;
  }
  const std::streamoff body = hs.tellg();
  if (body < 0) {
    err = "ply: no body";
    return false;
  }
  PlyReader rd(data, size_t(body), fmt);
  // vertex keys (the bits of every scalar property, `vstride` per vertex)
  // and face corner lists, flat
  std::vector<uint64_t> vkey;
  size_t vstride = 0;
  std::vector<std::array<float, 3>> vpos;
  std::vector<long> fidx;
  std::vector<size_t> foff{0};
  for (const PlyElement &e : elems) {
    const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
    int ix = -1, iy = -1, iz = -1, ifl = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const PlyProp &p = e.props[k];
      if (p.list) {
        if (p.name == "vertex_indices" || p.name == "vertex_index") ifl = int(k);
      } else if (p.name == "x") ix = int(k);
      else if (p.name == "y") iy = int(k);
      else if (p.name == "z") iz = int(k);
    }
    if (is_vertex && (ix < 0 || iy < 0 || iz < 0)) {
      err = "ply: vertex element without x, y, z";
      return false;
    }
    if (is_face && ifl < 0) {
      err = "ply: face element without vertex_indices";
      return false;
    }
    if (is_vertex) {
      vstride = 0;
      for (const PlyProp &p : e.props) vstride += p.list ? 0 : 1;
    }
    if (e.props.empty()) continue;  // nothing to read, whatever the declared count
    if (e.count > rd.Remaining()) {
      err = "ply: element '" + e.name + "' declares more rows than the body holds";
      return false;
    }
    for (size_t r = 0; r < e.count; ++r) {
      std::array<float, 3> pos{};
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PlyProp &p = e.props[k];
        double v;
        if (p.list) {
          double cnt;
          long n = 0;
          if (!rd.Read(p.count_type, cnt) || !PlyIndex(cnt, double(rd.Remaining()), n)) {
            err = "ply: truncated list or bad list count";
            return false;
          }
          for (long j = 0; j < n; ++j) {
            if (!rd.Read(p.type, v)) {
              err = "ply: truncated list";
              return false;
            }
            if (is_face && int(k) == ifl) {
              long idx;
              if (!PlyIndex(v, 2147483647.0, idx)) {
                err = "ply: face index is not a non-negative integer";
                return false;
              }
              fidx.push_back(idx);
            }
          }
          continue;
        }
        if (!rd.Read(p.type, v)) {
          err = "ply: truncated element '" + e.name + "'";
          return false;
        }
        if (is_vertex) {
          const float f = float(v);
          vkey.push_back(Bits(f));
          if (int(k) == ix) pos[0] = f;
          else if (int(k) == iy) pos[1] = f;
          else if (int(k) == iz) pos[2] = f;
        }
      }
      if (is_vertex) vpos.push_back(pos);
      else if (is_face) foff.push_back(fidx.size());
    }
  }
  MeshBuilder mb(out);
  std::vector<int> joined(vpos.size());
  std::vector<uint64_t> key(vstride);
  for (size_t i = 0; i < vpos.size(); ++i) {
    std::copy(vkey.begin() + long(i * vstride), vkey.begin() + long((i + 1) * vstride), key.begin());
    joined[i] = mb.Corner(key, vpos[i]);
  }
  std::vector<int> c;
  for (size_t fi = 0; fi + 1 < foff.size(); ++fi) {
    if (foff[fi + 1] - foff[fi] < 3) {
      err = "ply: face with fewer than 3 corners";
      return false;
    }
    c.clear();
    for (size_t j = foff[fi]; j < foff[fi + 1]; ++j) {
      const long i = fidx[j];
      if (i < 0 || size_t(i) >= vpos.size()) {
        err = "ply: face index out of range";
        return false;
      }
      c.push_back(joined[size_t(i)]);
    }
    mb.Polygon(c);
  }
  mb.Finish();
  return true;
}

std::string Lower(std::string s) {
  for (char &c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

}  // namespace

bool LoadMeshData(const std::string &data, const std::string &ext, std::vector<LoadedMesh> &out,
                  std::string &err) {
  std::vector<LoadedMesh> meshes;
  const std::string e = Lower(ext);
  bool ok;
  if (e == "obj") ok = ParseObj(data, meshes, err);
  else if (e == "ply") ok = ParsePly(data, meshes, err);
  else {
    err = "unsupported mesh format '" + ext + "' (obj and ply are read)";
    return false;
  }
  if (ok) out = std::move(meshes);
  return ok;
}

bool LoadMeshFile(const std::string &path, std::vector<LoadedMesh> &out, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open mesh file: " + path;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  const size_t dot = path.find_last_of('.');
  const std::string ext = dot == std::string::npos ? std::string() : path.substr(dot + 1);
  return LoadMeshData(ss.str(), ext, out, err);
}

}  // namespace host
}  // namespace muonx
