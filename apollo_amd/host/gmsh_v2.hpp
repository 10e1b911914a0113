// gmsh_v2.hpp -- Gmsh v2.2 ASCII meshes for the C++ host, in the numbering
// Apollo's subsolver sees after ApSolver::createMesh
// (src/solvers/apsolver.cc:296-359) on one rank:
//  - DMPlexCreateGmsh: cells are the triangles (element type 2) in file
//    order, vertices the $Nodes in file order, each cone the node list as
//    written; "Face Sets" is the physical tag (first tag) of the line
//    element (type 1) on the same vertex pair;
//  - DMPlexDistribute on one rank keeps that numbering;
//  - DMPlexConstructGhostCells appends ghost cells after the interior
//    cells, which the RHS never reads, so they are not materialised here.
// Other element types (points, quads) are skipped.
#pragma once

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>

#include "wx_nodal_dg2d_hip.hpp"

namespace apl_host {

inline NodalDG2dMesh read_gmsh_v2(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw WxExcept(APL_ERR_ARG, "cannot open mesh file " + path);
  std::string line;
  auto expect_section = [&](const char *name) {
    while (std::getline(in, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      if (line == name) return;
    }
    throw WxExcept(APL_ERR_ARG, std::string("mesh file has no ") + name + " section");
  };
  expect_section("$MeshFormat");
  std::getline(in, line);
  if (line.empty() || line[0] != '2')
    throw WxExcept(APL_ERR_ARG, "only Gmsh format 2.x is supported");
  NodalDG2dMesh m;
  expect_section("$Nodes");
  long nn = 0;
  std::getline(in, line);
  nn = std::stol(line);
  std::map<long, int> ids;
  m.VX.resize(nn);
  m.VY.resize(nn);
  for (long j = 0; j < nn; ++j) {
    std::getline(in, line);
    std::istringstream ss(line);
    long id;
    double x, y, z;
    ss >> id >> x >> y >> z;
    ids[id] = (int)j;
    m.VX[j] = x;
    m.VY[j] = y;
  }
  expect_section("$Elements");
  std::getline(in, line);
  const long ne = std::stol(line);
  std::map<std::pair<int, int>, int> edge_tag;
  for (long j = 0; j < ne; ++j) {
    std::getline(in, line);
    std::istringstream ss(line);
    long id;
    int type, ntags;
    ss >> id >> type >> ntags;
    int phys = 0;
    for (int t = 0; t < ntags; ++t) {
      int tag;
      ss >> tag;
      if (t == 0) phys = tag;
    }
    const int nv = type == 2 ? 3 : (type == 1 ? 2 : 0);
    int v[3] = {0, 0, 0};
    for (int k = 0; k < nv; ++k) {
      long nid;
      ss >> nid;
      auto it = ids.find(nid);
      if (it == ids.end()) throw WxExcept(APL_ERR_ARG, "element refers to an unknown node");
      v[k] = it->second;
    }
    if (type == 2) {
      m.cones.insert(m.cones.end(), {v[0], v[1], v[2]});
    } else if (type == 1) {
      edge_tag[{std::min(v[0], v[1]), std::max(v[0], v[1])}] = phys;
    }
  }
  const size_t K = m.cones.size() / 3;
  m.faceSets.assign(3 * K, 0);
  for (size_t k = 0; k < K; ++k)
    for (int f = 0; f < 3; ++f) {
      const int a = m.cones[3 * k + f], b = m.cones[3 * k + (f + 1) % 3];
      auto it = edge_tag.find({std::min(a, b), std::max(a, b)});
      if (it != edge_tag.end()) m.faceSets[3 * k + f] = it->second;
    }
  return m;
}

}  // namespace apl_host
