// Mesh setup: Gmsh-style element lists, or face-based (OpenFOAM polyMesh) input -> the
// cell-centred FV mesh every hot-path kernel reads.
//
// Host-side setup code (runs once per mesh). Each step names the reference routine whose result
// it reproduces, so a mesh built here carries the same face numbering, geometry, periodic pairs and
// least-squares weights as the reference's gmsh_read + build_connectivity + compute_geometry +
// build_periodic_pairs + build_ls_stencil on the same element lists.
#include "dopamine_mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

constexpr double SMALL = 1.0e-30;   // src/mod_parameters.f90:25
constexpr double LS_DET_TOL = 1.0e-12;  // src/mesh/mod_mesh.f90:20
constexpr double LS_REG_FRAC = 1.0e-6;  // src/mesh/mod_mesh.f90:21

thread_local std::string g_err;

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sumsq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Mesh {
  int nc = 0, nf = 0, nint = 0, nbnd = 0, npatch = 0, nvert = 0, max_cell_faces = 0, ls_degenerate = 0;
  std::vector<double> cell_centroid, cell_volume;
  std::vector<int32_t> cell_face_offset, cell_faces;
  std::vector<int32_t> face_owner, face_neighbour, face_patch;
  std::vector<double> face_area, face_centre, face_normal, face_sf, face_dcf, face_gc, face_skew, face_eps;
  std::vector<int32_t> patch_bc, patch_start, patch_nfaces;
  std::vector<int32_t> periodic_pair;
  std::vector<double> periodic_pair_dcf, periodic_gc;
  std::vector<int32_t> ls_offset, ls_cell, ls_face;
  std::vector<double> ls_weight;
  std::vector<double> vertex_coords;
  std::vector<int64_t> cell_vertex_offset;
  std::vector<int32_t> cell_vertices;
  std::vector<int64_t> face_node_offset;  // per face, right-hand normal along face_normal
  std::vector<int32_t> face_nodes;

  V3 centroid(int c) const { return {cell_centroid[3 * c], cell_centroid[3 * c + 1], cell_centroid[3 * c + 2]}; }
  V3 fcentre(int f) const { return {face_centre[3 * f], face_centre[3 * f + 1], face_centre[3 * f + 2]}; }
  V3 fnormal(int f) const { return {face_normal[3 * f], face_normal[3 * f + 1], face_normal[3 * f + 2]}; }
};

// Exact tet centroid and volume (mod_gmsh_reader.f90:773-781).
void tet_geometry(V3 p1, V3 p2, V3 p3, V3 p4, V3& c, double& vol) {
  V3 a = sub(p2, p1), b = sub(p3, p1), d = sub(p4, p1);
  vol = std::fabs(a.x * (b.y * d.z - b.z * d.y) - a.y * (b.x * d.z - b.z * d.x) + a.z * (b.x * d.y - b.y * d.x)) / 6.0;
  V3 s = add(add(add(p1, p2), p3), p4);
  c = {s.x / 4.0, s.y / 4.0, s.z / 4.0};
}

// Sum of tet moments over a fixed decomposition (hex: 5-tet Freudenthal partition, prism 3, pyramid 2;
// mod_gmsh_reader.f90:784-841).
void decomposed_geometry(const V3* p, const int (*conn)[4], int ntet, V3& centroid, double& volume) {
  volume = 0.0;
  centroid = {0.0, 0.0, 0.0};
  for (int i = 0; i < ntet; ++i) {
    V3 tc;
    double tv;
    tet_geometry(p[conn[i][0]], p[conn[i][1]], p[conn[i][2]], p[conn[i][3]], tc, tv);
    volume = volume + tv;
    centroid = add(centroid, scl(tv, tc));
  }
  if (volume > SMALL) centroid = {centroid.x / volume, centroid.y / volume, centroid.z / volume};
}

void cell_geometry(int type, const V3* p, V3& c, double& vol) {
  static const int hex[5][4] = {{0, 1, 3, 4}, {1, 2, 3, 6}, {3, 4, 6, 7}, {1, 4, 5, 6}, {1, 3, 4, 6}};
  static const int prism[3][4] = {{0, 1, 2, 3}, {1, 3, 4, 2}, {2, 3, 4, 5}};
  static const int pyr[2][4] = {{0, 1, 3, 4}, {1, 2, 3, 4}};
  switch (type) {
    case DM_GMSH_TET4: tet_geometry(p[0], p[1], p[2], p[3], c, vol); break;
    case DM_GMSH_HEX8: decomposed_geometry(p, hex, 5, c, vol); break;
    case DM_GMSH_PRISM6: decomposed_geometry(p, prism, 3, c, vol); break;
    case DM_GMSH_PYR5: decomposed_geometry(p, pyr, 2, c, vol); break;
    default: throw std::runtime_error("unsupported volume element type");
  }
}

int nodes_per_type(int type) {
  switch (type) {
    case DM_GMSH_TRI3: return 3;
    case DM_GMSH_QUAD4: return 4;
    case DM_GMSH_TET4: return 4;
    case DM_GMSH_HEX8: return 8;
    case DM_GMSH_PRISM6: return 6;
    case DM_GMSH_PYR5: return 5;
    default: return -1;
  }
}

// Local faces of a volume element in the reference's order (mod_gmsh_reader.f90:887-932).
int element_faces(int type, const int32_t* n, int fn[6][4], int fcount[6]) {
  auto set = [&](int i, int a, int b, int c, int d) {
    fn[i][0] = a; fn[i][1] = b; fn[i][2] = c; fn[i][3] = d;
    fcount[i] = (d < 0) ? 3 : 4;
  };
  switch (type) {
    case DM_GMSH_TET4:
      set(0, n[0], n[1], n[2], -1); set(1, n[0], n[1], n[3], -1);
      set(2, n[1], n[2], n[3], -1); set(3, n[0], n[2], n[3], -1);
      return 4;
    case DM_GMSH_HEX8:
      set(0, n[0], n[1], n[2], n[3]); set(1, n[4], n[5], n[6], n[7]);
      set(2, n[0], n[1], n[5], n[4]); set(3, n[2], n[3], n[7], n[6]);
      set(4, n[0], n[3], n[7], n[4]); set(5, n[1], n[2], n[6], n[5]);
      return 6;
    case DM_GMSH_PRISM6:
      set(0, n[0], n[1], n[2], -1); set(1, n[3], n[4], n[5], -1);
      set(2, n[0], n[1], n[4], n[3]); set(3, n[1], n[2], n[5], n[4]);
      set(4, n[0], n[2], n[5], n[3]);
      return 5;
    case DM_GMSH_PYR5:
      set(0, n[0], n[1], n[2], n[3]); set(1, n[0], n[1], n[4], -1);
      set(2, n[1], n[2], n[4], -1); set(3, n[2], n[3], n[4], -1);
      set(4, n[0], n[3], n[4], -1);
      return 5;
  }
  return 0;
}

// Vertex-mean centre and fan area vector about it (mod_gmsh_reader.f90:843-869).
void face_geometry(const int* fnodes, int nfn, const double* coords, V3& centre, V3& normal, double& area) {
  V3 s = {0.0, 0.0, 0.0};
  for (int i = 0; i < nfn; ++i) s = add(s, V3{coords[3 * fnodes[i]], coords[3 * fnodes[i] + 1], coords[3 * fnodes[i] + 2]});
  centre = {s.x / double(nfn), s.y / double(nfn), s.z / double(nfn)};
  V3 nacc = {0.0, 0.0, 0.0};
  for (int i = 0; i < nfn; ++i) {
    const int a = fnodes[i], b = fnodes[(i + 1) % nfn];
    V3 pi = {coords[3 * a], coords[3 * a + 1], coords[3 * a + 2]};
    V3 pj = {coords[3 * b], coords[3 * b + 1], coords[3 * b + 2]};
    nacc = add(nacc, scl(0.5, cross(sub(pi, centre), sub(pj, centre))));
  }
  area = std::sqrt(sumsq(nacc));
  normal = {nacc.x / (area + SMALL), nacc.y / (area + SMALL), nacc.z / (area + SMALL)};
}

// Right-hand-rule area vector of a point list (fan about its first point).
V3 fan_area(const int* nodes, int n, const double* coords) {
  auto P = [&](int i) { return V3{coords[3 * nodes[i]], coords[3 * nodes[i] + 1], coords[3 * nodes[i] + 2]}; };
  V3 acc = {0.0, 0.0, 0.0};
  for (int i = 1; i + 1 < n; ++i) acc = add(acc, cross(sub(P(i), P(0)), sub(P(i + 1), P(0))));
  return acc;
}

// Appends a face's points so that the right-hand normal follows `normal`.
void push_face_nodes(std::vector<int32_t>& out, const int* nodes, int n, const double* coords, V3 normal) {
  if (dot(fan_area(nodes, n, coords), normal) < 0.0) {
    out.push_back(nodes[0]);
    for (int i = n - 1; i >= 1; --i) out.push_back(nodes[i]);
  } else {
    out.insert(out.end(), nodes, nodes + n);
  }
}

using FaceKey = std::array<int32_t, 4>;
struct FaceKeyHash {
  size_t operator()(const FaceKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 4; ++i) h = (h ^ uint64_t(uint32_t(k[i]))) * 1099511628211ull;
    return size_t(h);
  }
};
FaceKey canonical(const int* n, int cnt) {
  FaceKey k = {-1, -1, -1, -1};
  for (int i = 0; i < cnt; ++i) k[i] = n[i];
  std::sort(k.begin(), k.begin() + cnt);
  return k;
}

// Symmetric 3x3 inverse by cofactors with a sign-preserving determinant guard (mod_mesh.f90:1259-1283).
void invert3x3(const double A[3][3], double Ai[3][3], double& det) {
  const double c11 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c12 = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]);
  const double c13 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double c21 = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]);
  const double c22 = A[0][0] * A[2][2] - A[0][2] * A[2][0];
  const double c23 = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]);
  const double c31 = A[0][1] * A[1][2] - A[0][2] * A[1][1];
  const double c32 = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]);
  const double c33 = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  det = A[0][0] * c11 + A[0][1] * c12 + A[0][2] * c13;
  const double g = std::max(std::fabs(det), SMALL);
  const double idet = 1.0 / (det < 0.0 ? -g : g);
  Ai[0][0] = c11 * idet; Ai[0][1] = c21 * idet; Ai[0][2] = c31 * idet;
  Ai[1][0] = c12 * idet; Ai[1][1] = c22 * idet; Ai[1][2] = c32 * idet;
  Ai[2][0] = c13 * idet; Ai[2][1] = c23 * idet; Ai[2][2] = c33 * idet;
}

void build_faces(Mesh& m, const dm_mesh_input& in) {
  const int nc = int(in.n_vol);
  m.nc = nc;
  m.cell_centroid.assign(3 * size_t(nc), 0.0);
  m.cell_volume.assign(nc, 0.0);
  m.nvert = int(in.n_nodes);
  m.vertex_coords.assign(in.coords, in.coords + 3 * in.n_nodes);
  m.cell_vertex_offset.assign(in.vol_offset, in.vol_offset + nc + 1);
  m.cell_vertices.assign(in.vol_nodes, in.vol_nodes + in.vol_offset[nc]);

  for (int c = 0; c < nc; ++c) {
    const int t = in.vol_type[c];
    const int64_t o = in.vol_offset[c];
    if (nodes_per_type(t) < 0 || in.vol_offset[c + 1] - o != nodes_per_type(t))
      throw std::runtime_error("volume element " + std::to_string(c) + ": bad type or node count");
    V3 p[8];
    for (int i = 0; i < nodes_per_type(t); ++i) {
      const int32_t v = in.vol_nodes[o + i];
      if (v < 0 || v >= in.n_nodes) throw std::runtime_error("volume element node out of range");
      p[i] = {in.coords[3 * v], in.coords[3 * v + 1], in.coords[3 * v + 2]};
    }
    V3 cen;
    double vol;
    cell_geometry(t, p, cen, vol);
    m.cell_centroid[3 * c] = cen.x;
    m.cell_centroid[3 * c + 1] = cen.y;
    m.cell_centroid[3 * c + 2] = cen.z;
    m.cell_volume[c] = vol;
  }

  // Interior faces are numbered when the second cell meets a face; the first cell owns it
  // (mod_gmsh_reader.f90:426-460). state: 1 open, 2 interior, 3 closed by a boundary element.
  struct Slot {
    int32_t owner;
    int32_t state;
  };
  std::unordered_map<FaceKey, Slot, FaceKeyHash> table;
  table.reserve(size_t(nc) * 4);
  std::vector<int32_t> fown, fnb, fpatch, fnodes;
  std::vector<int64_t> foff{0};
  std::vector<double> fcen, fnrm, farea;
  fown.reserve(size_t(nc) * 3);
  for (int c = 0; c < nc; ++c) {
    int fn[6][4], fcnt[6];
    const int nfc = element_faces(in.vol_type[c], in.vol_nodes + in.vol_offset[c], fn, fcnt);
    for (int fi = 0; fi < nfc; ++fi) {
      FaceKey k = canonical(fn[fi], fcnt[fi]);
      auto it = table.find(k);
      if (it == table.end()) {
        table.emplace(k, Slot{c, 1});
      } else if (it->second.state == 1) {
        const int owner = it->second.owner;
        V3 cen, nrm;
        double area;
        face_geometry(fn[fi], fcnt[fi], in.coords, cen, nrm, area);
        const V3 d = sub(m.centroid(c), m.centroid(owner));
        if (dot(nrm, d) < 0.0) nrm = scl(-1.0, nrm);  // orient_normal
        push_face_nodes(fnodes, fn[fi], fcnt[fi], in.coords, nrm);
        foff.push_back(int64_t(fnodes.size()));
        fown.push_back(owner);
        fnb.push_back(c);
        fpatch.push_back(-1);
        fcen.insert(fcen.end(), {cen.x, cen.y, cen.z});
        fnrm.insert(fnrm.end(), {nrm.x, nrm.y, nrm.z});
        farea.push_back(area);
        it->second.state = 2;
      }
    }
  }
  const int nint = int(fown.size());

  // Boundary faces from the 2-D elements, owner found through the same table (:470-504).
  const int nb = int(in.n_bnd);
  std::vector<int32_t> bown(nb), bpatch(nb), bnodes(4 * size_t(nb));
  std::vector<double> bcen(3 * size_t(nb)), bnrm(3 * size_t(nb)), barea(nb);
  for (int e = 0; e < nb; ++e) {
    const int64_t o = in.bnd_offset[e];
    const int cnt = int(in.bnd_offset[e + 1] - o);
    if (cnt != 3 && cnt != 4) throw std::runtime_error("boundary element with unsupported node count");
    int nodes[4];
    for (int i = 0; i < cnt; ++i) nodes[i] = in.bnd_nodes[o + i];
    auto it = table.find(canonical(nodes, cnt));
    if (it == table.end()) throw std::runtime_error("boundary owner not found for boundary element " + std::to_string(e));
    bown[e] = it->second.owner;
    if (it->second.state == 1) it->second.state = 3;
    V3 cen, nrm;
    double area;
    face_geometry(nodes, cnt, in.coords, cen, nrm, area);
    if (dot(nrm, sub(cen, m.centroid(bown[e]))) < 0.0) nrm = scl(-1.0, nrm);  // orient_outward_normal
    {
      std::vector<int32_t> o;
      push_face_nodes(o, nodes, cnt, in.coords, nrm);
      for (int i = 0; i < 4; ++i) bnodes[4 * size_t(e) + i] = i < cnt ? o[i] : -1;
    }
    bcen[3 * e] = cen.x; bcen[3 * e + 1] = cen.y; bcen[3 * e + 2] = cen.z;
    bnrm[3 * e] = nrm.x; bnrm[3 * e + 1] = nrm.y; bnrm[3 * e + 2] = nrm.z;
    barea[e] = area;
    const int g = in.bnd_group[e];
    if (g < 0 || g >= in.n_groups) throw std::runtime_error("boundary element without a 2-D physical group");
    bpatch[e] = g;
  }
  for (const auto& kv : table)
    if (kv.second.state == 1) throw std::runtime_error("unlabelled boundary faces (open cells)");

  // Stable reorder of boundary faces by patch; empty groups get no patch (:526-581, 618-649).
  std::vector<int> count(in.n_groups, 0), pmap(in.n_groups, -1);
  for (int e = 0; e < nb; ++e) count[bpatch[e]]++;
  int np = 0;
  for (int g = 0; g < in.n_groups; ++g)
    if (count[g] > 0) {
      pmap[g] = np++;
      m.patch_bc.push_back(in.group_bc[g]);
      m.patch_nfaces.push_back(count[g]);
    }
  m.patch_start.assign(np, 0);
  for (int p = 1; p < np; ++p) m.patch_start[p] = m.patch_start[p - 1] + m.patch_nfaces[p - 1];
  std::vector<int> cursor = m.patch_start;
  std::vector<int> order(nb);
  for (int e = 0; e < nb; ++e) order[cursor[pmap[bpatch[e]]]++] = e;

  m.nint = nint;
  m.nbnd = nb;
  m.nf = nint + nb;
  m.npatch = np;
  m.face_owner = fown;
  m.face_neighbour = fnb;
  m.face_patch = fpatch;
  m.face_centre = fcen;
  m.face_normal = fnrm;
  m.face_area = farea;
  m.face_nodes = fnodes;
  m.face_node_offset = foff;
  for (int i = 0; i < nb; ++i) {
    const int e = order[i];
    for (int k = 0; k < 4 && bnodes[4 * size_t(e) + k] >= 0; ++k) m.face_nodes.push_back(bnodes[4 * size_t(e) + k]);
    m.face_node_offset.push_back(int64_t(m.face_nodes.size()));
    m.face_owner.push_back(bown[e]);
    m.face_neighbour.push_back(-1);
    m.face_patch.push_back(pmap[bpatch[e]]);
    m.face_centre.insert(m.face_centre.end(), {bcen[3 * e], bcen[3 * e + 1], bcen[3 * e + 2]});
    m.face_normal.insert(m.face_normal.end(), {bnrm[3 * e], bnrm[3 * e + 1], bnrm[3 * e + 2]});
    m.face_area.push_back(barea[e]);
  }
}

// Face-based input (OpenFOAM polyMesh layout): faces in input order, geometry by the reference's
// polyhedral formulas (mod_mesh.f90:1336-1461), cell vertex lists in first-appearance order over
// the cell's owner faces then its neighbour faces (build_cell_topology_metadata,
// mod_foam_reader.f90:417-579), patches as given, empty ones included (:346-415).
void build_faces_poly(Mesh& m, const dm_poly_input& in) {
  const int64_t nf = in.n_faces, nint = in.n_int_faces, npts = in.n_points;
  if (nf <= 0 || nint < 0 || nint > nf) throw std::runtime_error("bad face counts");
  if (nf > INT32_MAX || npts > INT32_MAX) throw std::runtime_error("mesh too large for 32-bit indices");
  if (in.face_offset[0] != 0) throw std::runtime_error("face offsets must start at 0");
  for (int64_t f = 0; f < nf; ++f)
    if (in.face_offset[f + 1] < in.face_offset[f]) throw std::runtime_error("face offsets are not monotonic");
  for (int64_t k = 0; k < in.face_offset[nf]; ++k)
    if (in.face_nodes[k] < 0 || in.face_nodes[k] >= npts) throw std::runtime_error("face point label out of range");
  int64_t nc = 0;
  for (int64_t f = 0; f < nf; ++f) {
    if (in.owner[f] < 0) throw std::runtime_error("negative cell index in owner/neighbour");
    nc = std::max<int64_t>(nc, int64_t(in.owner[f]) + 1);
  }
  for (int64_t f = 0; f < nint; ++f) {
    if (in.neighbour[f] < 0) throw std::runtime_error("negative cell index in owner/neighbour");
    nc = std::max<int64_t>(nc, int64_t(in.neighbour[f]) + 1);
    if (in.owner[f] >= in.neighbour[f])
      throw std::runtime_error("face " + std::to_string(f) + " violates owner < neighbour");
  }
  m.nc = int(nc);
  m.nf = int(nf);
  m.nint = int(nint);
  m.nbnd = int(nf - nint);
  m.nvert = int(npts);
  m.vertex_coords.assign(in.points, in.points + 3 * npts);
  m.face_node_offset.assign(in.face_offset, in.face_offset + nf + 1);
  m.face_nodes.assign(in.face_nodes, in.face_nodes + in.face_offset[nf]);
  m.face_owner.assign(in.owner, in.owner + nf);
  m.face_neighbour.assign(nf, -1);
  for (int64_t f = 0; f < nint; ++f) m.face_neighbour[f] = in.neighbour[f];

  // patches: ranges inside the boundary faces, covering each boundary face once
  m.npatch = in.n_patches;
  m.face_patch.assign(nf, -1);
  for (int p = 0; p < in.n_patches; ++p) {
    const int64_t s = in.patch_start[p], n = in.patch_nfaces[p];
    if (n < 0 || s < nint || s + n > nf)
      throw std::runtime_error("patch " + std::to_string(p) + ": face range outside the boundary faces");
    for (int64_t f = s; f < s + n; ++f) {
      if (m.face_patch[f] >= 0) throw std::runtime_error("boundary face in two patches");
      m.face_patch[f] = p;
    }
    m.patch_bc.push_back(in.patch_bc[p]);
    m.patch_start.push_back(int32_t(s - nint));
    m.patch_nfaces.push_back(int32_t(n));
  }
  for (int64_t f = nint; f < nf; ++f)
    if (m.face_patch[f] < 0) throw std::runtime_error("boundary faces not mapped to a patch");

  auto P = [&](int64_t k) {
    const int v = in.face_nodes[k];
    return V3{in.points[3 * size_t(v)], in.points[3 * size_t(v) + 1], in.points[3 * size_t(v) + 2]};
  };
  // compute_polygon_face_geometry (mod_mesh.f90:1415-1461)
  m.face_centre.assign(3 * size_t(nf), 0.0);
  m.face_normal.assign(3 * size_t(nf), 0.0);
  m.face_area.assign(nf, 0.0);
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t b = in.face_offset[f], e = in.face_offset[f + 1];
    if (e - b < 3) {
      m.face_area[f] = SMALL;
      continue;
    }
    const V3 pref = P(b);
    V3 nacc = {0.0, 0.0, 0.0};
    for (int64_t i = b + 1; i < e - 1; ++i) nacc = add(nacc, scl(0.5, cross(sub(P(i), pref), sub(P(i + 1), pref))));
    const double area = std::sqrt(nacc.x * nacc.x + nacc.y * nacc.y + nacc.z * nacc.z) + SMALL;
    const V3 nrm = {nacc.x / area, nacc.y / area, nacc.z / area};
    V3 ctr = {0.0, 0.0, 0.0};
    double asum = 0.0;
    for (int64_t i = b + 1; i < e - 1; ++i) {
      const V3 pi = P(i), pj = P(i + 1);
      const double ta = dot(scl(0.5, cross(sub(pi, pref), sub(pj, pref))), nrm);
      const V3 t = add(add(pref, pi), pj);
      ctr = add(ctr, V3{ta * t.x / 3.0, ta * t.y / 3.0, ta * t.z / 3.0});
      asum += ta;
    }
    if (std::fabs(asum) > SMALL) ctr = {ctr.x / asum, ctr.y / asum, ctr.z / asum};
    else ctr = pref;
    m.face_area[f] = area;
    m.face_normal[3 * f] = nrm.x;
    m.face_normal[3 * f + 1] = nrm.y;
    m.face_normal[3 * f + 2] = nrm.z;
    m.face_centre[3 * f] = ctr.x;
    m.face_centre[3 * f + 1] = ctr.y;
    m.face_centre[3 * f + 2] = ctr.z;
  }

  // compute_polyhedron_cell_moments (mod_mesh.f90:1336-1412): signed tets about a per-cell
  // reference point, the first point of the first face (of >= 3 points) that touches the cell
  std::vector<V3> ref(nc), mom(nc, V3{0.0, 0.0, 0.0});
  std::vector<char> ref_set(nc, 0);
  std::vector<double> vol(nc, 0.0);
  for (int64_t f = 0; f < nf; ++f) {
    if (in.face_offset[f + 1] - in.face_offset[f] < 3) continue;
    for (int side = 0; side < 2; ++side) {
      const int c = side == 0 ? in.owner[f] : (f < nint ? in.neighbour[f] : -1);
      if (c < 0 || ref_set[c]) continue;
      ref[c] = P(in.face_offset[f]);
      ref_set[c] = 1;
    }
  }
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t b = in.face_offset[f], e = in.face_offset[f + 1];
    if (e - b < 3) continue;
    for (int side = 0; side < 2; ++side) {
      const int c = side == 0 ? in.owner[f] : (f < nint ? in.neighbour[f] : -1);
      if (c < 0) break;
      const V3 r = ref[c], v0 = sub(P(b), r);
      for (int64_t i = b + 1; i < e - 1; ++i) {
        const V3 v1 = sub(P(i), r), v2 = sub(P(i + 1), r);
        const double sv = dot(v0, cross(v1, v2)) / 6.0;
        const V3 t = add(add(v0, v1), v2);
        const V3 sm = {sv * t.x / 4.0, sv * t.y / 4.0, sv * t.z / 4.0};
        if (side == 0) {
          vol[c] += sv;
          mom[c] = add(mom[c], sm);
        } else {
          vol[c] -= sv;
          mom[c] = sub(mom[c], sm);
        }
      }
    }
  }
  m.cell_centroid.assign(3 * size_t(nc), 0.0);
  m.cell_volume.assign(nc, 0.0);
  for (int64_t c = 0; c < nc; ++c) {
    if (std::fabs(vol[c]) > SMALL) {
      m.cell_centroid[3 * c] = ref[c].x + mom[c].x / vol[c];
      m.cell_centroid[3 * c + 1] = ref[c].y + mom[c].y / vol[c];
      m.cell_centroid[3 * c + 2] = ref[c].z + mom[c].z / vol[c];
    }
    m.cell_volume[c] = std::fabs(vol[c]);
  }

  // cell vertex lists: owner faces in face order, then neighbour faces, first appearance
  std::vector<int64_t> cfo(nc + 1, 0);
  for (int64_t f = 0; f < nf; ++f) cfo[in.owner[f] + 1]++;
  for (int64_t f = 0; f < nint; ++f) cfo[in.neighbour[f] + 1]++;
  for (int64_t c = 0; c < nc; ++c) cfo[c + 1] += cfo[c];
  std::vector<int32_t> cf(cfo[nc]);
  {
    std::vector<int64_t> fill(cfo.begin(), cfo.end() - 1);
    for (int64_t f = 0; f < nf; ++f) cf[fill[in.owner[f]]++] = int32_t(f);
    for (int64_t f = 0; f < nint; ++f) cf[fill[in.neighbour[f]]++] = int32_t(f);
  }
  std::vector<int64_t> marker(npts, -1);
  m.cell_vertex_offset.assign(1, 0);
  m.cell_vertices.clear();
  for (int64_t c = 0; c < nc; ++c) {
    for (int64_t k = cfo[c]; k < cfo[c + 1]; ++k) {
      const int f = cf[k];
      for (int64_t q = in.face_offset[f]; q < in.face_offset[f + 1]; ++q) {
        const int v = in.face_nodes[q];
        if (marker[v] != c) {
          marker[v] = c;
          m.cell_vertices.push_back(v);
        }
      }
    }
    m.cell_vertex_offset.push_back(int64_t(m.cell_vertices.size()));
  }
}

// Cell->face CSR in ascending face order (mod_mesh.f90:215-259).
void build_connectivity(Mesh& m) {
  std::vector<int32_t> deg(m.nc, 0);
  for (int f = 0; f < m.nf; ++f) {
    deg[m.face_owner[f]]++;
    if (m.face_neighbour[f] >= 0) deg[m.face_neighbour[f]]++;
  }
  m.cell_face_offset.assign(m.nc + 1, 0);
  for (int c = 0; c < m.nc; ++c) m.cell_face_offset[c + 1] = m.cell_face_offset[c] + deg[c];
  m.cell_faces.assign(m.cell_face_offset[m.nc], 0);
  std::vector<int32_t> fill(m.nc, 0);
  m.max_cell_faces = 0;
  for (int c = 0; c < m.nc; ++c) m.max_cell_faces = std::max(m.max_cell_faces, deg[c]);
  for (int f = 0; f < m.nf; ++f) {
    const int o = m.face_owner[f];
    m.cell_faces[m.cell_face_offset[o] + fill[o]++] = f;
    const int n = m.face_neighbour[f];
    if (n >= 0) m.cell_faces[m.cell_face_offset[n] + fill[n]++] = f;
  }
}

// Face area vector, distances, interpolation weight, skewness offset (mod_mesh.f90:437-489).
void compute_geometry(Mesh& m) {
  m.face_sf.assign(3 * size_t(m.nf), 0.0);
  m.face_dcf.assign(m.nf, 0.0);
  m.face_gc.assign(m.nf, 0.0);
  m.face_skew.assign(m.nf, 0.0);
  m.face_eps.assign(3 * size_t(m.nf), 0.0);
  for (int f = 0; f < m.nf; ++f) {
    const double a = m.face_area[f];
    const V3 n = m.fnormal(f), ctr = m.fcentre(f);
    m.face_sf[3 * f] = a * n.x;
    m.face_sf[3 * f + 1] = a * n.y;
    m.face_sf[3 * f + 2] = a * n.z;
    const int o = m.face_owner[f], nb = m.face_neighbour[f];
    if (nb >= 0) {
      const V3 d_on = sub(m.centroid(nb), m.centroid(o));
      const double d_on2 = sumsq(d_on);
      m.face_dcf[f] = std::max(std::sqrt(d_on2), 1.0e-30);
      const V3 d_oc = sub(ctr, m.centroid(o));
      const double w = dot(d_oc, d_on) / (d_on2 + SMALL);
      const double gc = std::max(0.0, std::min(1.0, 1.0 - w));
      m.face_gc[f] = gc;
      const V3 e = sub(ctr, add(m.centroid(o), scl(1.0 - gc, d_on)));
      m.face_eps[3 * f] = e.x;
      m.face_eps[3 * f + 1] = e.y;
      m.face_eps[3 * f + 2] = e.z;
      // mesh_tag_refinement_faces (mod_mesh.f90:491-509)
      const V3 mid = sub(ctr, scl(0.5, add(m.centroid(o), m.centroid(nb))));
      m.face_skew[f] = std::sqrt(sumsq(mid)) / std::max(m.face_dcf[f], 1.0e-30);
    } else {
      const V3 d_cn = sub(ctr, m.centroid(o));
      m.face_dcf[f] = std::max(std::sqrt(sumsq(d_cn)), 1.0e-30);
      m.face_gc[f] = 1.0;
    }
  }
}

// Geometric pairing of periodic patches by translated face centres (mod_bc_setup.f90:150-330).
void build_periodic_pairs(Mesh& m) {
  m.periodic_pair.assign(m.nbnd, -1);
  m.periodic_pair_dcf.assign(m.nbnd, 0.0);
  std::vector<int> per;
  for (int p = 0; p < m.npatch; ++p)
    if (m.patch_bc[p] == DM_BC_PERIODIC) per.push_back(p);
  std::vector<char> used(m.npatch, 0);
  const double KC1 = 3.14159265358979, KC2 = 2.71828182845904, KC3 = 1.41421356237310;
  for (size_t i = 0; i < per.size(); ++i) {
    const int p = per[i];
    if (used[p]) continue;
    int best_q = -1;
    double best_err = 1e300;
    std::vector<int> best_map;
    for (size_t j = i + 1; j < per.size(); ++j) {
      const int q = per[j];
      if (used[q]) continue;
      const int n = m.patch_nfaces[p];
      if (n == 0 || n != m.patch_nfaces[q]) continue;
      std::vector<V3> cp(n), cq(n);
      V3 sp = {0, 0, 0}, sq = {0, 0, 0}, snp = {0, 0, 0}, snq = {0, 0, 0};
      double scale = 1.0;
      for (int k = 0; k < n; ++k) {
        const int fp = m.nint + m.patch_start[p] + k, fq = m.nint + m.patch_start[q] + k;
        cp[k] = m.fcentre(fp);
        cq[k] = m.fcentre(fq);
        sp = add(sp, cp[k]);
        sq = add(sq, cq[k]);
        snp = add(snp, m.fnormal(fp));
        snq = add(snq, m.fnormal(fq));
        scale = std::max({scale, std::fabs(cp[k].x), std::fabs(cp[k].y), std::fabs(cp[k].z),
                          std::fabs(cq[k].x), std::fabs(cq[k].y), std::fabs(cq[k].z)});
      }
      const V3 shift = sub(scl(1.0 / 1.0, V3{sq.x / n, sq.y / n, sq.z / n}), V3{sp.x / n, sp.y / n, sp.z / n});
      const double avg_dot = dot(snp, snq);
      const double tol = 1.0e-5 * scale;
      std::vector<double> kp(n), kq(n);
      std::vector<int> op(n), oq(n);
      for (int k = 0; k < n; ++k) {
        kp[k] = KC1 * (cp[k].x + shift.x) + KC2 * (cp[k].y + shift.y) + KC3 * (cp[k].z + shift.z);
        kq[k] = KC1 * cq[k].x + KC2 * cq[k].y + KC3 * cq[k].z;
        op[k] = oq[k] = k;
      }
      std::stable_sort(op.begin(), op.end(), [&](int a, int b) { return kp[a] < kp[b]; });
      std::stable_sort(oq.begin(), oq.end(), [&](int a, int b) { return kq[a] < kq[b]; });
      std::vector<int> map(n);
      double max_err = 0.0;
      for (int k = 0; k < n; ++k) {
        map[op[k]] = oq[k];
        max_err = std::max(max_err, std::sqrt(sumsq(sub(cq[oq[k]], add(cp[op[k]], shift)))));
      }
      if (max_err <= tol && avg_dot < 0.0 && max_err < best_err) {
        best_err = max_err;
        best_q = q;
        best_map = map;
      }
    }
    if (best_q < 0) throw std::runtime_error("no periodic partner for patch " + std::to_string(p));
    for (int k = 0; k < m.patch_nfaces[p]; ++k) {
      const int fbp = m.patch_start[p] + k, fbq = m.patch_start[best_q] + best_map[k];
      m.periodic_pair[fbp] = fbq;
      m.periodic_pair[fbq] = fbp;
      m.periodic_pair_dcf[fbp] = m.face_dcf[m.nint + fbq];
      m.periodic_pair_dcf[fbq] = m.face_dcf[m.nint + fbp];
    }
    used[p] = used[best_q] = 1;
  }
  // finalize_periodic_partner_metrics: unpaired faces fall back to their own d_cf.
  for (int b = 0; b < m.nbnd; ++b)
    if (m.periodic_pair_dcf[b] <= 0.0) m.periodic_pair_dcf[b] = m.face_dcf[m.nint + b];
  // Owner weight of the periodic-pair interpolation (mod_mesh.f90:419-425).
  m.periodic_gc.assign(m.nbnd, 0.5);
  for (int b = 0; b < m.nbnd; ++b)
    if (m.periodic_pair_dcf[b] > 0.0)
      m.periodic_gc[b] = m.periodic_pair_dcf[b] / (m.face_dcf[m.nint + b] + m.periodic_pair_dcf[b] + SMALL);
}

// Least-squares stencil: face neighbours (periodic partners through the pair), boundary
// pseudo-neighbours at face centres, then vertex-adjacent cells in vertex order; weights from
// inverse-distance normal equations with Tikhonov fallback (mod_mesh.f90:513-754).
void build_ls_stencil(Mesh& m) {
  const int nc = m.nc;
  const int max_stencil = std::max(32, 4 * std::max(1, m.max_cell_faces));
  std::vector<int32_t> cnt(nc, 0), tnb(size_t(nc) * max_stencil), tface(size_t(nc) * max_stencil, -1);
  std::vector<char> tper(size_t(nc) * max_stencil, 0);
  for (int i = 0; i < nc; ++i) {
    for (int k = m.cell_face_offset[i]; k < m.cell_face_offset[i + 1]; ++k) {
      const int f = m.cell_faces[k];
      int nb = m.face_owner[f];
      if (nb == i) nb = m.face_neighbour[f];
      bool is_per = false;
      if (nb < 0) {
        const int fb = f - m.nint, pid = m.face_patch[f];
        if (pid >= 0 && m.patch_bc[pid] == DM_BC_PERIODIC) {
          const int pair = m.periodic_pair[fb];
          if (pair >= 0) nb = m.face_owner[m.nint + pair];
          if (nb >= 0) is_per = true;
        }
        if (nb < 0) nb = -1;
      }
      const size_t s = size_t(i) * max_stencil + cnt[i];
      tnb[s] = nb;
      tper[s] = is_per;
      if (nb == -1) tface[s] = f;
      cnt[i]++;
    }
  }
  // vertex -> cells CSR, cells ascending
  std::vector<int64_t> voff(m.nvert + 1, 0);
  for (int c = 0; c < nc; ++c)
    for (int64_t k = m.cell_vertex_offset[c]; k < m.cell_vertex_offset[c + 1]; ++k) voff[m.cell_vertices[k] + 1]++;
  for (int v = 0; v < m.nvert; ++v) voff[v + 1] += voff[v];
  std::vector<int32_t> vcell(voff[m.nvert]);
  {
    std::vector<int64_t> ptr(voff.begin(), voff.end() - 1);
    for (int c = 0; c < nc; ++c)
      for (int64_t k = m.cell_vertex_offset[c]; k < m.cell_vertex_offset[c + 1]; ++k)
        vcell[ptr[m.cell_vertices[k]]++] = c;
  }
  for (int i = 0; i < nc; ++i) {
    int32_t* row = &tnb[size_t(i) * max_stencil];
    for (int64_t k = m.cell_vertex_offset[i]; k < m.cell_vertex_offset[i + 1]; ++k) {
      const int v = m.cell_vertices[k];
      for (int64_t kk = voff[v]; kk < voff[v + 1]; ++kk) {
        const int cnb = vcell[kk];
        if (cnb == i) continue;
        bool found = false;
        for (int jj = 0; jj < cnt[i]; ++jj)
          if (row[jj] == cnb) { found = true; break; }
        if (found) continue;
        if (cnt[i] < max_stencil) {
          row[cnt[i]] = cnb;
          tper[size_t(i) * max_stencil + cnt[i]] = 0;
          tface[size_t(i) * max_stencil + cnt[i]] = -1;
          cnt[i]++;
        }
      }
    }
  }
  m.ls_offset.assign(nc + 1, 0);
  for (int i = 0; i < nc; ++i) m.ls_offset[i + 1] = m.ls_offset[i] + cnt[i];
  const int64_t n_ls = m.ls_offset[nc];
  m.ls_cell.assign(n_ls, 0);
  m.ls_face.assign(n_ls, -1);
  m.ls_weight.assign(3 * n_ls, 0.0);
  m.ls_degenerate = 0;
  std::vector<V3> dv(max_stencil);
  for (int i = 0; i < nc; ++i) {
    const int off = m.ls_offset[i];
    double AtA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int idx = 0; idx < cnt[i]; ++idx) {
      const size_t s = size_t(i) * max_stencil + idx;
      const int nb = tnb[s];
      m.ls_cell[off + idx] = nb;
      m.ls_face[off + idx] = tface[s];
      V3 d;
      if (nb >= 0 && tper[s]) {
        // periodic neighbour: (d_cf_own + d_cf_partner) along the face normal
        const int f = m.cell_faces[m.cell_face_offset[i] + idx];
        const double d_o = m.face_dcf[f];
        double d_nb = d_o;
        const int fb = f - m.nint;
        if (fb >= 0 && fb < m.nbnd && m.periodic_pair_dcf[fb] > 0.0) d_nb = m.periodic_pair_dcf[fb];
        const V3 n = m.fnormal(f);
        d = {(d_o + d_nb) * n.x, (d_o + d_nb) * n.y, (d_o + d_nb) * n.z};
      } else if (nb >= 0) {
        d = sub(m.centroid(nb), m.centroid(i));
      } else {
        const int f = m.cell_faces[m.cell_face_offset[i] + idx];
        d = sub(m.fcentre(f), m.centroid(i));
      }
      dv[idx] = d;
      const double w = 1.0 / (std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z) + SMALL);
      const double ww = w * w;
      AtA[0][0] += ww * d.x * d.x; AtA[0][1] += ww * d.x * d.y;
      AtA[0][2] += ww * d.x * d.z; AtA[1][1] += ww * d.y * d.y;
      AtA[1][2] += ww * d.y * d.z; AtA[2][2] += ww * d.z * d.z;
    }
    AtA[1][0] = AtA[0][1]; AtA[2][0] = AtA[0][2]; AtA[2][1] = AtA[1][2];
    const double scale = (AtA[0][0] + AtA[1][1] + AtA[2][2]) / 3.0;
    double Ai[3][3], det;
    invert3x3(AtA, Ai, det);
    if (scale <= 0.0 || std::fabs(det) < LS_DET_TOL * scale * scale * scale) {
      for (int k = 0; k < 3; ++k) AtA[k][k] += LS_REG_FRAC * std::max(scale, SMALL);
      invert3x3(AtA, Ai, det);
      m.ls_degenerate++;
    }
    for (int idx = 0; idx < cnt[i]; ++idx) {
      const V3 d = dv[idx];
      const double w = 1.0 / (std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z) + SMALL);
      const double r[3] = {w * w * d.x, w * w * d.y, w * w * d.z};
      for (int a = 0; a < 3; ++a) m.ls_weight[3 * size_t(off + idx) + a] = Ai[a][0] * r[0] + Ai[a][1] * r[1] + Ai[a][2] * r[2];
    }
  }
}

}  // namespace

extern "C" {

const char* dm_last_error(void) { return g_err.c_str(); }

void* dm_mesh_build(const dm_mesh_input* in) {
  g_err.clear();
  try {
    if (!in || in->n_vol <= 0) throw std::runtime_error("empty mesh input");
    auto* m = new Mesh();
    try {
      build_faces(*m, *in);
      build_connectivity(*m);
      compute_geometry(*m);
      build_periodic_pairs(*m);
      build_ls_stencil(*m);
    } catch (...) {
      delete m;
      throw;
    }
    return m;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void* dm_mesh_build_poly(const dm_poly_input* in) {
  g_err.clear();
  try {
    if (!in || in->n_faces <= 0) throw std::runtime_error("empty mesh input");
    auto* m = new Mesh();
    try {
      build_faces_poly(*m, *in);
      build_connectivity(*m);
      compute_geometry(*m);
      build_periodic_pairs(*m);
      build_ls_stencil(*m);
    } catch (...) {
      delete m;
      throw;
    }
    return m;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// Reverse Cuthill-McKee order of the cells over the face graph (interior faces and periodic
// pairs): breadth-first from a pseudo-peripheral cell (the last cell reached from the lowest-
// degree one), neighbours taken by increasing degree then index, components one after another,
// the whole order reversed. order[new] = old.
int dm_mesh_rcm_order(const void* mesh, int32_t* order) {
  g_err.clear();
  try {
    const Mesh& m = *static_cast<const Mesh*>(mesh);
    const int nc = m.nc;
    std::vector<std::vector<int>> adj(nc);
    for (int f = 0; f < m.nint; ++f) {
      adj[m.face_owner[f]].push_back(m.face_neighbour[f]);
      adj[m.face_neighbour[f]].push_back(m.face_owner[f]);
    }
    for (int b = 0; b < m.nbnd; ++b) {
      const int q = m.periodic_pair[b];
      if (q > b) {
        const int a = m.face_owner[m.nint + b], c = m.face_owner[m.nint + q];
        if (a != c) {
          adj[a].push_back(c);
          adj[c].push_back(a);
        }
      }
    }
    for (auto& l : adj) {
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
    }
    for (auto& l : adj)
      std::stable_sort(l.begin(), l.end(), [&](int a, int c) { return adj[a].size() < adj[c].size(); });
    std::vector<char> seen(nc, 0);
    std::vector<int> out;
    out.reserve(nc);
    auto bfs = [&](int start, std::vector<int>& visit, std::vector<char>& mark) {
      size_t head = visit.size();
      visit.push_back(start);
      mark[start] = 1;
      while (head < visit.size()) {
        const int c = visit[head++];
        for (int nb : adj[c])
          if (!mark[nb]) {
            mark[nb] = 1;
            visit.push_back(nb);
          }
      }
    };
    for (int c0 = 0; c0 < nc; ++c0) {
      if (seen[c0]) continue;
      // lowest-degree cell of this component, then the last cell a search from it reaches
      std::vector<int> comp;
      std::vector<char> mark(seen);
      bfs(c0, comp, mark);
      int lo = comp[0];
      for (int c : comp)
        if (adj[c].size() < adj[lo].size()) lo = c;
      std::vector<int> probe;
      std::vector<char> mark2(seen);
      bfs(lo, probe, mark2);
      bfs(probe.back(), out, seen);
    }
    for (int i = 0; i < nc; ++i) order[i] = out[nc - 1 - i];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int dm_mesh_get_view(const void* mesh, dm_mesh_view* v) {
  const Mesh& m = *static_cast<const Mesh*>(mesh);
  std::memset(v, 0, sizeof(*v));
  v->n_cells = m.nc;
  v->n_faces = m.nf;
  v->n_int_faces = m.nint;
  v->n_bnd_faces = m.nbnd;
  v->n_patches = m.npatch;
  v->n_vertices = m.nvert;
  v->n_ls = int64_t(m.ls_cell.size());
  v->max_cell_faces = m.max_cell_faces;
  v->ls_degenerate = m.ls_degenerate;
  v->cell_centroid = m.cell_centroid.data();
  v->cell_volume = m.cell_volume.data();
  v->cell_face_offset = m.cell_face_offset.data();
  v->cell_faces = m.cell_faces.data();
  v->face_owner = m.face_owner.data();
  v->face_neighbour = m.face_neighbour.data();
  v->face_patch = m.face_patch.data();
  v->face_area = m.face_area.data();
  v->face_centre = m.face_centre.data();
  v->face_normal = m.face_normal.data();
  v->face_sf = m.face_sf.data();
  v->face_dcf = m.face_dcf.data();
  v->face_gc = m.face_gc.data();
  v->face_skew = m.face_skew.data();
  v->face_eps = m.face_eps.data();
  v->patch_bc = m.patch_bc.data();
  v->patch_start = m.patch_start.data();
  v->patch_nfaces = m.patch_nfaces.data();
  v->periodic_pair = m.periodic_pair.data();
  v->periodic_pair_dcf = m.periodic_pair_dcf.data();
  v->periodic_gc = m.periodic_gc.data();
  v->ls_offset = m.ls_offset.data();
  v->ls_cell = m.ls_cell.data();
  v->ls_face = m.ls_face.data();
  v->ls_weight = m.ls_weight.data();
  v->vertex_coords = m.vertex_coords.data();
  v->cell_vertex_offset = m.cell_vertex_offset.data();
  v->cell_vertices = m.cell_vertices.data();
  v->face_node_offset = m.face_node_offset.data();
  v->face_nodes = m.face_nodes.data();
  return 0;
}

void dm_mesh_free(void* mesh) { delete static_cast<Mesh*>(mesh); }

}  // extern "C"
