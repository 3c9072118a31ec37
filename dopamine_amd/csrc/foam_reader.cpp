// OpenFOAM polyMesh reader (include/dopamine_mesh.h): points, faces, owner, neighbour and boundary
// of a constant/polyMesh directory, ASCII or binary, as the reference's foam_read takes them
// (src/mesh/mod_foam_reader.f90:16-97, file parsers :139-415, binary forms :617-1104).
//   - FoamFile header: format ascii|binary, arch "LSB|MSB;label=32|64;scalar=32|64" (:617-666)
//   - binary lists: count, '(', raw values, ')'; faces either as a compact pair of lists (offsets
//     starting at 0, then the concatenated point labels) or as (n, labels...) records, told apart
//     by the first label, 0 only for the compact form (:697-790, :842-876)
//   - boundary: per patch its name, type, nFaces, startFace (:346-415); type -> bc id (:1144-1167)
// The checks the reference makes on the files are made here too (negative cell index,
// owner >= neighbour on an internal face, a boundary face outside every patch, patch ranges,
// an odd number of cyclic patches); the mesh builder (dm_mesh_build_poly) repeats the topology
// checks for callers that bring their own arrays.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dopamine_mesh.h"

namespace {

thread_local std::string g_foam_err;

struct FoamMesh {
  std::vector<double> points;
  std::vector<int64_t> face_offset{0};
  std::vector<int32_t> face_nodes, owner, neighbour;
  std::vector<std::string> patch_name, patch_type;
  std::vector<int64_t> patch_start, patch_nfaces;
  std::vector<int32_t> patch_bc;
  int64_t n_cells = 0;
};

// Whole file plus a cursor; the header's format and arch fields.
struct Src {
  std::string b;
  size_t pos = 0;
  bool binary = false, little = true;
  int label_bytes = 4, scalar_bytes = 8;
  std::string name, cls;

  [[noreturn]] void fail(const std::string& what) const { throw std::runtime_error(name + ": " + what); }

  bool eof() const { return pos >= b.size(); }
  char peek() const { return eof() ? '\0' : b[pos]; }

  // whitespace, // line comments and /* */ block comments
  void skip_ws() {
    while (!eof()) {
      const char c = b[pos];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v') {
        ++pos;
      } else if (c == '/' && pos + 1 < b.size() && b[pos + 1] == '/') {
        while (!eof() && b[pos] != '\n') ++pos;
      } else if (c == '/' && pos + 1 < b.size() && b[pos + 1] == '*') {
        const size_t e = b.find("*/", pos + 2);
        pos = e == std::string::npos ? b.size() : e + 2;
      } else {
        break;
      }
    }
  }

  void expect(char c) {
    skip_ws();
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++pos;
  }

  int64_t integer() {
    skip_ws();
    const char* s = b.c_str() + pos;
    char* e = nullptr;
    const long long v = std::strtoll(s, &e, 10);
    if (e == s) fail("expected an integer");
    pos += size_t(e - s);
    return int64_t(v);
  }

  double real() {
    skip_ws();
    const char* s = b.c_str() + pos;
    char* e = nullptr;
    const double v = std::strtod(s, &e);
    if (e == s) fail("expected a number");
    pos += size_t(e - s);
    return v;
  }

  // a word or a quoted string up to whitespace, ';', '{', '}', '(' or ')'
  std::string word() {
    skip_ws();
    if (peek() == '"') {
      const size_t e = b.find('"', pos + 1);
      if (e == std::string::npos) fail("unterminated string");
      std::string w = b.substr(pos + 1, e - pos - 1);
      pos = e + 1;
      return w;
    }
    const size_t s = pos;
    while (!eof() && !std::strchr(" \t\r\n;{}()", b[pos])) ++pos;
    return b.substr(s, pos - s);
  }

  uint64_t raw(int nbytes) {
    if (pos + size_t(nbytes) > b.size()) fail("binary list runs past the end of the file");
    uint64_t v = 0;
    for (int i = 0; i < nbytes; ++i) {
      const uint64_t byte = uint8_t(b[pos + (little ? size_t(nbytes - 1 - i) : size_t(i))]);
      v = (v << 8) | byte;
    }
    pos += size_t(nbytes);
    return v;
  }

  int64_t label() {
    if (!binary) return integer();
    const uint64_t v = raw(label_bytes);
    return label_bytes == 4 ? int64_t(int32_t(uint32_t(v))) : int64_t(v);
  }

  double scalar() {
    if (!binary) return real();
    const uint64_t v = raw(scalar_bytes);
    if (scalar_bytes == 4) {
      const uint32_t u = uint32_t(v);
      float f;
      std::memcpy(&f, &u, 4);
      return double(f);
    }
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }

  // list size and the opening parenthesis; binary data starts right after it (no skipping:
  // the first value's leading byte may look like whitespace)
  int64_t open_list() {
    const int64_t n = integer();
    if (n < 0) fail("negative list size");
    expect('(');
    return n;
  }
  void close_list() { expect(')'); }
};

// Reads the file and its FoamFile header dictionary (mod_foam_reader.f90:617-666).
Src load(const std::string& path) {
  Src s;
  s.name = path;
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::ostringstream ss;
  ss << in.rdbuf();
  s.b = ss.str();
  const size_t h = s.b.find("FoamFile");
  if (h == std::string::npos) s.fail("no FoamFile header");
  s.pos = h + 8;
  s.expect('{');
  for (;;) {
    s.skip_ws();
    if (s.peek() == '}') {
      ++s.pos;
      break;
    }
    if (s.eof()) s.fail("unterminated FoamFile header");
    const std::string key = s.word();
    if (key.empty()) s.fail("malformed FoamFile header");
    s.skip_ws();
    // a quoted value may hold ';' (arch "LSB;label=32;scalar=64")
    const size_t q = s.peek() == '"' ? s.b.find('"', s.pos + 1) : s.pos;
    if (q == std::string::npos) s.fail("unterminated string in FoamFile header");
    const size_t e = s.b.find(';', q);
    if (e == std::string::npos) s.fail("malformed FoamFile header");
    std::string val = s.b.substr(s.pos, e - s.pos);
    s.pos = e + 1;
    while (!val.empty() && (val.back() == ' ' || val.back() == '\t')) val.pop_back();
    if (val.size() >= 2 && val.front() == '"' && val.back() == '"') val = val.substr(1, val.size() - 2);
    if (key == "format") {
      if (val == "binary") s.binary = true;
      else if (val == "ascii") s.binary = false;
      else s.fail("unknown format " + val);
    } else if (key == "arch") {
      if (val.find("MSB") != std::string::npos) s.little = false;
      if (val.find("label=64") != std::string::npos) s.label_bytes = 8;
      if (val.find("scalar=32") != std::string::npos) s.scalar_bytes = 4;
    } else if (key == "class") {
      s.cls = val;
    }
  }
  return s;
}

void read_points(const std::string& path, FoamMesh& m) {
  Src s = load(path);
  const int64_t n = s.open_list();
  m.points.resize(3 * size_t(n));
  for (int64_t i = 0; i < n; ++i) {
    if (!s.binary) s.expect('(');
    for (int k = 0; k < 3; ++k) m.points[3 * size_t(i) + k] = s.scalar();
    if (!s.binary) s.expect(')');
  }
  s.close_list();
}

void read_faces(const std::string& path, FoamMesh& m) {
  Src s = load(path);
  const int64_t n = s.open_list();
  const size_t data = s.pos;
  const bool compact = s.cls == "faceCompactList" || (s.binary && n > 0 && s.label() == 0);
  s.pos = data;
  m.face_offset.assign(1, 0);
  m.face_nodes.clear();
  if (compact) {
    // offsets (n entries, from 0, non-decreasing), then the concatenated labels
    if (n < 2) s.fail("compact face offset table too small");
    std::vector<int64_t> off(n);
    for (int64_t i = 0; i < n; ++i) off[i] = s.label();
    s.close_list();
    if (off[0] != 0) s.fail("compact face offsets must start at 0");
    for (int64_t i = 1; i < n; ++i)
      if (off[i] < off[i - 1]) s.fail("compact face offsets are not monotonic");
    const int64_t np = s.open_list();
    if (np != off[n - 1]) s.fail("compact face offsets do not match the label list");
    m.face_nodes.resize(np);
    for (int64_t i = 0; i < np; ++i) m.face_nodes[i] = int32_t(s.label());
    s.close_list();
    m.face_offset.assign(off.begin(), off.end());
  } else {
    for (int64_t i = 0; i < n; ++i) {
      const int64_t k = s.label();
      if (k < 0) s.fail("negative node count in a face");
      if (!s.binary) s.expect('(');
      for (int64_t j = 0; j < k; ++j) m.face_nodes.push_back(int32_t(s.label()));
      if (!s.binary) s.expect(')');
      m.face_offset.push_back(int64_t(m.face_nodes.size()));
    }
    s.close_list();
  }
}

void read_labels(const std::string& path, std::vector<int32_t>& out) {
  Src s = load(path);
  const int64_t n = s.open_list();
  out.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t v = s.label();
    if (v < INT32_MIN || v > INT32_MAX) s.fail("label out of the 32-bit range");
    out[i] = int32_t(v);
  }
  s.close_list();
}

// bc_type_to_id (mod_foam_reader.f90:1144-1167)
int32_t bc_of_type(const std::string& t) {
  if (t == "inlet" || t == "fixedValue" || t == "inletOutlet") return DM_BC_INLET;
  if (t == "outlet" || t == "inletOutlet_out") return DM_BC_OUTLET;
  if (t == "zeroGradient") return DM_BC_ZEROGRADIENT;
  if (t == "wall") return DM_BC_WALL_MODEL;
  if (t == "fixedTemperature") return DM_BC_FIXEDTEMP;
  if (t == "movingWall" || t == "wallVelocity") return DM_BC_WALL_VELOCITY;
  if (t == "symmetry" || t == "symmetryPlane") return DM_BC_SYMMETRY;
  if (t == "empty") return DM_BC_SLIP;
  if (t == "cyclic" || t == "cyclicAMI") return DM_BC_PERIODIC;
  return DM_BC_ZEROGRADIENT;
}

// name { key value; ... } per patch; values may hold parenthesised lists (inGroups 1(wall);)
void read_boundary(const std::string& path, FoamMesh& m) {
  Src s = load(path);
  const int64_t n = s.open_list();
  for (int64_t p = 0; p < n; ++p) {
    const std::string name = s.word();
    if (name.empty()) s.fail("malformed patch entry");
    s.expect('{');
    std::string type = "patch";
    int64_t nfaces = 0, start = 0;
    for (;;) {
      s.skip_ws();
      if (s.peek() == '}') {
        ++s.pos;
        break;
      }
      if (s.eof()) s.fail("unterminated patch entry " + name);
      const std::string key = s.word();
      if (key.empty()) s.fail("malformed entry in patch " + name);
      std::string val;
      int depth = 0;
      s.skip_ws();
      while (!s.eof() && !(depth == 0 && s.peek() == ';')) {
        const char c = s.b[s.pos++];
        if (c == '(') ++depth;
        if (c == ')') --depth;
        val.push_back(c);
      }
      if (s.eof()) s.fail("unterminated entry " + key + " in patch " + name);
      ++s.pos;  // ';'
      while (!val.empty() && (val.back() == ' ' || val.back() == '\t')) val.pop_back();
      if (key == "type") {
        type = val;
      } else if (key == "nFaces" || key == "startFace") {
        char* e = nullptr;
        const long long v = std::strtoll(val.c_str(), &e, 10);
        if (e == val.c_str() || *e != '\0') s.fail("bad " + key + " in patch " + name);
        (key == "nFaces" ? nfaces : start) = int64_t(v);
      }
    }
    m.patch_name.push_back(name);
    m.patch_type.push_back(type);
    m.patch_start.push_back(start);
    m.patch_nfaces.push_back(nfaces);
    m.patch_bc.push_back(bc_of_type(type));
  }
  s.close_list();
}

FoamMesh read(const std::string& dir) {
  FoamMesh m;
  read_points(dir + "/points", m);
  read_faces(dir + "/faces", m);
  read_labels(dir + "/owner", m.owner);
  read_labels(dir + "/neighbour", m.neighbour);
  read_boundary(dir + "/boundary", m);
  const int64_t nf = int64_t(m.face_offset.size()) - 1, nint = int64_t(m.neighbour.size());
  const int64_t npts = int64_t(m.points.size() / 3);
  if (int64_t(m.owner.size()) != nf) throw std::runtime_error("owner list length differs from the face count");
  if (nint > nf) throw std::runtime_error("more neighbours than faces");
  int64_t nc = 0;
  for (int32_t o : m.owner) {
    if (o < 0) throw std::runtime_error("negative cell index in owner/neighbour");
    nc = std::max<int64_t>(nc, int64_t(o) + 1);
  }
  for (int64_t f = 0; f < nint; ++f) {
    if (m.neighbour[f] < 0) throw std::runtime_error("negative cell index in owner/neighbour");
    nc = std::max<int64_t>(nc, int64_t(m.neighbour[f]) + 1);
    if (m.owner[f] >= m.neighbour[f])
      throw std::runtime_error("face " + std::to_string(f) + " violates owner < neighbour");
  }
  for (int32_t v : m.face_nodes)
    if (v < 0 || v >= npts) throw std::runtime_error("face point label out of range");
  m.n_cells = nc;
  // every boundary face in exactly the range of a patch (mod_foam_reader.f90:99-137)
  std::vector<char> hit(size_t(nf - nint), 0);
  int n_cyclic = 0;
  for (size_t p = 0; p < m.patch_name.size(); ++p) {
    const int64_t s = m.patch_start[p], n = m.patch_nfaces[p];
    if (n < 0 || s < nint || s + n > nf)
      throw std::runtime_error("patch " + m.patch_name[p] + ": face range outside the boundary faces");
    for (int64_t f = s; f < s + n; ++f) hit[size_t(f - nint)] = 1;
    if (m.patch_bc[p] == DM_BC_PERIODIC) ++n_cyclic;
  }
  for (char h : hit)
    if (!h) throw std::runtime_error("boundary faces not mapped to a patch");
  if (n_cyclic % 2) throw std::runtime_error("odd number of periodic patches");
  return m;
}

}  // namespace

extern "C" {

const char* dm_foam_last_error(void) { return g_foam_err.c_str(); }

void* dm_foam_read(const char* dir) {
  try {
    g_foam_err.clear();
    if (!dir) throw std::runtime_error("null path");
    return new FoamMesh(read(dir));
  } catch (const std::exception& e) {
    g_foam_err = e.what();
    return nullptr;
  }
}

int dm_foam_get_view(const void* h, dm_poly_input* v) {
  const auto* m = static_cast<const FoamMesh*>(h);
  if (!m || !v) return -1;
  v->n_points = int64_t(m->points.size() / 3);
  v->points = m->points.data();
  v->n_faces = int64_t(m->face_offset.size()) - 1;
  v->face_offset = m->face_offset.data();
  v->face_nodes = m->face_nodes.data();
  v->n_int_faces = int64_t(m->neighbour.size());
  v->owner = m->owner.data();
  v->neighbour = m->neighbour.data();
  v->n_patches = int32_t(m->patch_name.size());
  v->patch_start = m->patch_start.data();
  v->patch_nfaces = m->patch_nfaces.data();
  v->patch_bc = m->patch_bc.data();
  return 0;
}

const char* dm_foam_patch_name(const void* h, int p) {
  const auto* m = static_cast<const FoamMesh*>(h);
  if (!m || p < 0 || p >= int(m->patch_name.size())) return nullptr;
  return m->patch_name[p].c_str();
}

const char* dm_foam_patch_type(const void* h, int p) {
  const auto* m = static_cast<const FoamMesh*>(h);
  if (!m || p < 0 || p >= int(m->patch_type.size())) return nullptr;
  return m->patch_type[p].c_str();
}

int32_t dm_foam_bc_of_type(const char* type) { return type ? bc_of_type(type) : DM_BC_ZEROGRADIENT; }

void dm_foam_free(void* h) { delete static_cast<FoamMesh*>(h); }

}  // extern "C"
