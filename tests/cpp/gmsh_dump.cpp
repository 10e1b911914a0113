// Reads a Gmsh v2.2 mesh with the C++ host's reader (apollo_amd/host/gmsh_v2.hpp)
// and writes K, Nv, VX, VY, cones and face tags for the Python test to compare.
//   gmsh_dump <mesh.msh> <out.bin>
#include <cstdio>
#include <vector>

#include "../../apollo_amd/host/gmsh_v2.hpp"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  try {
    apl_host::NodalDG2dMesh m = apl_host::read_gmsh_v2(argv[1]);
    std::FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 3;
    const double hdr[2] = {(double)(m.cones.size() / 3), (double)m.VX.size()};
    std::fwrite(hdr, sizeof(double), 2, f);
    std::fwrite(m.VX.data(), sizeof(double), m.VX.size(), f);
    std::fwrite(m.VY.data(), sizeof(double), m.VY.size(), f);
    std::vector<double> c(m.cones.begin(), m.cones.end()), t(m.faceSets.begin(), m.faceSets.end());
    std::fwrite(c.data(), sizeof(double), c.size(), f);
    std::fwrite(t.data(), sizeof(double), t.size(), f);
    std::fclose(f);
  } catch (const apl_host::WxExcept &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 4;
  }
  return 0;
}
