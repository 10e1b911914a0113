// A C++ caller of the Apollo-shaped interface: sets up the Maxwell TM
// cavity on a jittered walled mesh, evaluates one RHS through step(), runs
// LSERK4 steps through advance() and two output frames of Apollo's own
// stepping through solve(), and writes the results for the Python test to
// compare with the oracle.  Exit status 0 on success.
//   subsolver_demo <n> <N> <nsteps> <out.bin>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../apollo_amd/host/wx_nodal_dg2d_hip.hpp"

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s n N nsteps out.bin\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]), N = std::atoi(argv[2]), nsteps = std::atoi(argv[3]);
  using namespace apl_host;
  NodalDG2dMesh mesh;
  const int K = 2 * n * n, Nv = (n + 1) * (n + 1);
  mesh.VX.resize(Nv);
  mesh.VY.resize(Nv);
  mesh.cones.resize(3 * K);
  mesh.faceSets.resize(3 * K);
  mesh.vertexId.resize(Nv);
  if (apl_mesh_structured(n, 0.0, 1.0, 0.0, 1.0, 0.1, 1, 0, mesh.VX.data(), mesh.VY.data(),
                          mesh.cones.data(), mesh.vertexId.data(), mesh.faceSets.data()))
    return 3;
  NodalDG2dConfig cfg;
  cfg.polynomialOrder = N;
  cfg.equation = "maxwellTM";
  cfg.cfl = 0.5 / (2 * N + 1);
  apl_bc tm{};
  tm.kind = APL_BC_MAXWELL_TM;
  cfg.boundaryConditions.push_back(tm);
  try {
    WxNodalDG2dMethodHip dg;
    dg.setup(cfg, mesh);
    const size_t len = (size_t)dg.numElements() * dg.NpElem() * dg.totalEqns();
    std::vector<double> q(len), rhs(len);
    const double pi = 3.14159265358979323846;
    const double dt0 = dg.init(0.0, [&](double, double x, double y, double *qq) {
      qq[0] = std::sin(3 * pi * x) * std::sin(4 * pi * y);
      qq[1] = 0.0;
      qq[2] = 0.0;
    }, q.data());
    std::vector<double> q0 = q;
    WxStepperStatus<double> st = dg.step(0.0, dt0, q.data(), rhs.data());
    const double dt = st.getSuggestedDt();
    dg.advance(0.0, dt, nsteps, q.data());
    // two output frames of Apollo's own stepping from q0, each restarted at dt0
    std::vector<double> qf = q0;
    int frames = 0;
    dg.solve(2 * nsteps * dt, 2, dt0, qf.data(), [&](int, double, const double *) { ++frames; });
    if (frames != 2) return 5;
    std::FILE *f = std::fopen(argv[4], "wb");
    if (!f) return 4;
    const double hdr[3] = {dt0, dt, (double)len};
    std::fwrite(hdr, sizeof(double), 3, f);
    std::fwrite(q0.data(), sizeof(double), len, f);
    std::fwrite(rhs.data(), sizeof(double), len, f);
    std::fwrite(q.data(), sizeof(double), len, f);
    std::fwrite(qf.data(), sizeof(double), len, f);
    std::fclose(f);
    std::printf("subsolver_demo: K=%u Np=%u m=%u dt0=%.6e dt=%.6e\n", dg.numElements(), dg.NpElem(),
                dg.totalEqns(), dt0, dt);
  } catch (const WxExcept &e) {
    std::fprintf(stderr, "WxExcept: %s\n", e.what());
    return 10 + e.status();
  }
  return 0;
}
