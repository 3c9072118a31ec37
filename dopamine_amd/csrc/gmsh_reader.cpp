// Gmsh v4 ASCII reader (include/dopamine_mesh.h): the sections the reference's gmsh_read parses
// (src/mesh/mod_gmsh_reader.f90:45-236) into the element lists the mesh builder takes.
//   $PhysicalNames  dim, tag, "name"                               (read_physical_names, :118-135)
//   $Entities       entity tag -> first physical tag, per dimension (read_entities_v4, :165-197;
//                   parse_entity_line_for_phys, :651-676)
//   $Nodes          blocks of tags then coordinates; a node's index is its position in the file
//                   (tag_to_local of build_mesh_from_gmsh, :255-262)
//   $Elements       blocks by entity; 3-D elements become cells in file order, 2-D elements
//                   boundary elements tagged with their entity's physical group (:199-236)
// Boundary groups are numbered among the 2-D physical groups in $PhysicalNames order, which is
// the order the reference gives its patches (build_patches, :618-649).
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dopamine_mesh.h"

namespace {

thread_local std::string g_gmsh_err;

struct GmshFile {
  std::vector<double> coords;
  std::vector<int32_t> vol_type, vol_nodes, bnd_type, bnd_nodes, bnd_group;
  std::vector<int64_t> vol_offset{0}, bnd_offset{0};
  std::vector<std::string> group_names;  // 2-D physical groups
};

// nodes per first-order element type (gmsh_nodes_per_element); 0 = unsupported
int nodes_of(int t) {
  switch (t) {
    case 15: return 1;  // point
    case 1: return 2;   // line
    case DM_GMSH_TRI3: return 3;
    case DM_GMSH_QUAD4: return 4;
    case DM_GMSH_TET4: return 4;
    case DM_GMSH_HEX8: return 8;
    case DM_GMSH_PRISM6: return 6;
    case DM_GMSH_PYR5: return 5;
    default: return 0;
  }
}

std::string strip_quotes(std::string s) {
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') s = s.substr(1, s.size() - 2);
  return s;
}

GmshFile read(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  GmshFile g;
  struct Phys {
    int dim, tag;
    std::string name;
  };
  std::vector<Phys> phys;
  std::map<std::pair<int, int>, int> entity_phys;  // (dim, entity tag) -> physical tag
  std::vector<int64_t> node_tags;
  bool got_nodes = false, got_elements = false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream first(line);
    std::string key;
    first >> key;
    if (key == "$MeshFormat") {
      double version = 0.0;
      int file_type = 0;
      in >> version >> file_type;
      if (version < 4.0 || version >= 5.0) throw std::runtime_error("only Gmsh v4 meshes are read");
      if (file_type != 0) throw std::runtime_error("only ASCII Gmsh files are read");
    } else if (key == "$PhysicalNames") {
      int n = 0;
      in >> n;
      std::getline(in, line);
      for (int i = 0; i < n; ++i) {
        std::getline(in, line);
        std::istringstream ls(line);
        Phys p;
        ls >> p.dim >> p.tag;
        std::string rest;
        std::getline(ls, rest);
        const size_t a = rest.find_first_not_of(" \t");
        p.name = strip_quotes(a == std::string::npos ? std::string() : rest.substr(a, rest.find_last_not_of(" \t\r") - a + 1));
        phys.push_back(p);
      }
    } else if (key == "$Entities") {
      int64_t counts[4] = {0, 0, 0, 0};
      in >> counts[0] >> counts[1] >> counts[2] >> counts[3];
      std::getline(in, line);
      for (int dim = 0; dim < 4; ++dim)
        for (int64_t i = 0; i < counts[dim]; ++i) {
          std::getline(in, line);
          std::istringstream ls(line);
          std::vector<std::string> tok;
          for (std::string t; ls >> t;) tok.push_back(t);
          // points: tag x y z nphys tags...; others: tag minx miny minz maxx maxy maxz nphys tags...
          const size_t np_at = dim == 0 ? 4 : 7;
          if (tok.size() <= np_at) continue;
          const int nphys = std::stoi(tok[np_at]);
          if (nphys <= 0 || tok.size() <= np_at + 1) continue;
          entity_phys[{dim, std::stoi(tok[0])}] = std::stoi(tok[np_at + 1]);  // the first group only
        }
    } else if (key == "$Nodes") {
      int64_t nblocks = 0, nnodes = 0, mn = 0, mx = 0;
      in >> nblocks >> nnodes >> mn >> mx;
      node_tags.reserve(size_t(nnodes));
      g.coords.reserve(3 * size_t(nnodes));
      for (int64_t b = 0; b < nblocks; ++b) {
        int edim = 0, etag = 0, param = 0;
        int64_t n = 0;
        in >> edim >> etag >> param >> n;
        if (param != 0) throw std::runtime_error("parametric node blocks are not read");
        for (int64_t i = 0; i < n; ++i) {
          int64_t t = 0;
          in >> t;
          node_tags.push_back(t);
        }
        for (int64_t i = 0; i < 3 * n; ++i) {
          double v = 0.0;
          in >> v;
          g.coords.push_back(v);
        }
      }
      if (int64_t(node_tags.size()) != nnodes || !in) throw std::runtime_error("malformed $Nodes section");
      got_nodes = true;
    } else if (key == "$Elements") {
      if (!got_nodes) throw std::runtime_error("$Elements before $Nodes");
      std::map<int64_t, int32_t> t2l;
      for (size_t i = 0; i < node_tags.size(); ++i) t2l[node_tags[i]] = int32_t(i);
      // 2-D groups: index among the 2-D physical groups, by physical tag
      std::map<int, int> group_of_tag;
      for (const Phys& p : phys)
        if (p.dim == 2) {
          group_of_tag[p.tag] = int(g.group_names.size());
          g.group_names.push_back(p.name);
        }
      int64_t nblocks = 0, nelem = 0, mn = 0, mx = 0;
      in >> nblocks >> nelem >> mn >> mx;
      for (int64_t b = 0; b < nblocks; ++b) {
        int edim = 0, etag = 0, etype = 0;
        int64_t n = 0;
        in >> edim >> etag >> etype >> n;
        const int nn = nodes_of(etype);
        if (nn < 1)
          throw std::runtime_error("unsupported Gmsh element type " + std::to_string(etype) +
                                   " (export first-order elements)");
        const auto ep = entity_phys.find({edim, etag});
        const int phys_tag = ep == entity_phys.end() ? -1 : ep->second;
        for (int64_t i = 0; i < n; ++i) {
          int64_t tag = 0;
          in >> tag;
          std::vector<int32_t> nodes(nn);
          for (int k = 0; k < nn; ++k) {
            int64_t t = 0;
            in >> t;
            const auto it = t2l.find(t);
            if (it == t2l.end()) throw std::runtime_error("element node tag not in $Nodes");
            nodes[k] = it->second;
          }
          if (etype == DM_GMSH_TET4 || etype == DM_GMSH_HEX8 || etype == DM_GMSH_PRISM6 || etype == DM_GMSH_PYR5) {
            g.vol_type.push_back(etype);
            g.vol_nodes.insert(g.vol_nodes.end(), nodes.begin(), nodes.end());
            g.vol_offset.push_back(int64_t(g.vol_nodes.size()));
          } else if (etype == DM_GMSH_TRI3 || etype == DM_GMSH_QUAD4) {
            const auto gi = group_of_tag.find(phys_tag);
            if (gi == group_of_tag.end())
              throw std::runtime_error("boundary element with physical tag " + std::to_string(phys_tag) +
                                       " belongs to no 2-D physical group");
            g.bnd_type.push_back(etype);
            g.bnd_nodes.insert(g.bnd_nodes.end(), nodes.begin(), nodes.end());
            g.bnd_offset.push_back(int64_t(g.bnd_nodes.size()));
            g.bnd_group.push_back(gi->second);
          }
        }
      }
      if (!in) throw std::runtime_error("malformed $Elements section");
      got_elements = true;
    }
  }
  if (!got_nodes) throw std::runtime_error("missing $Nodes section");
  if (!got_elements) throw std::runtime_error("missing $Elements section");
  if (g.vol_type.empty()) throw std::runtime_error("no volume elements");
  return g;
}

}  // namespace

extern "C" {

const char* dm_gmsh_last_error(void) { return g_gmsh_err.c_str(); }

void* dm_gmsh_read(const char* path) {
  try {
    g_gmsh_err.clear();
    if (!path) throw std::runtime_error("null path");
    return new GmshFile(read(path));
  } catch (const std::exception& e) {
    g_gmsh_err = e.what();
    return nullptr;
  }
}

int dm_gmsh_get_view(const void* h, dm_gmsh_view* v) {
  const auto* g = static_cast<const GmshFile*>(h);
  if (!g || !v) return -1;
  v->n_nodes = int64_t(g->coords.size() / 3);
  v->coords = g->coords.data();
  v->n_vol = int64_t(g->vol_type.size());
  v->vol_type = g->vol_type.data();
  v->vol_offset = g->vol_offset.data();
  v->vol_nodes = g->vol_nodes.data();
  v->n_bnd = int64_t(g->bnd_type.size());
  v->bnd_type = g->bnd_type.data();
  v->bnd_offset = g->bnd_offset.data();
  v->bnd_nodes = g->bnd_nodes.data();
  v->bnd_group = g->bnd_group.data();
  v->n_groups = int32_t(g->group_names.size());
  return 0;
}

const char* dm_gmsh_group_name(const void* h, int group) {
  const auto* g = static_cast<const GmshFile*>(h);
  if (!g || group < 0 || group >= int(g->group_names.size())) return nullptr;
  return g->group_names[group].c_str();
}

void dm_gmsh_free(void* h) { delete static_cast<GmshFile*>(h); }

}  // extern "C"
