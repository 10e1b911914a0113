// apl_host.h -- host-side (CPU) part of the engine: operators, mesh,
// connectivity and geometry, all in FP64.  Pure C++; no HIP types.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace apl {

// Nodal, cubature and face-Gauss operators of one order, row-major FP64.
// Built the way WxCubature2d (src/lib/wxcubature2d.cc:9-169) and
// wxNodalDGgeometry2D (src/lib/wxnodaldggeometry2d.cc:10-107) build theirs.
struct Operators {
  int N = 0, Np = 0, Nc = 0, Ng = 0, Nfp = 0;
  double rmin = 0.0;
  std::vector<double> r, s;         // [Np]
  std::vector<int> fmask;           // [3*Nfp]
  std::vector<double> V, invV, VVT; // [Np*Np]
  std::vector<double> cr, cs, cw;   // [Nc]
  std::vector<double> Vc, Drc, Dsc; // [Nc*Np]
  std::vector<double> gz, gw;       // [Ng]
  std::vector<double> If;           // [3Ng*Np]
};

// Returns false if N is outside 1..8.
bool build_operators(int N, Operators &ops);

// Orthonormal Jacobi polynomial and its derivative (src/lib/wxmath.h:303-368).
double jacobi_p(double x, double alpha, double beta, int n);
double grad_jacobi_p(double x, double alpha, double beta, int n);
// n-point Gauss-Legendre rule on [-1,1].
void gauss_legendre(int n, double *z, double *w);
// Dense inverse; false if singular.
bool invert(int n, const double *A, double *Ainv);

// Connectivity with FacePair2d semantics (src/lib/wxnodaldggeometry2d.cc:
// 235-367).  etetf[K*6].  Returns 0, or APL_ERR_MESH_TOPOLOGY if a face is
// shared by more than two elements.
int face_pair(int K, int Nv, const int *EToV, const int *vert_id,
              const int *face_tag, std::vector<int> &etetf);

// Per-element geometry (src/lib/wxnodaldggeometry2d.cc:179-233,382-442).
struct Geometry {
  std::vector<double> x, y;   // [K*Np] node coordinates
  std::vector<double> geom;   // [K*14] rx,sx,ry,sy,J, (nx,ny,L/2) x 3
  std::vector<double> vert;   // [K*6] x1,y1,x2,y2,x3,y3 (cone order)
  double dtscale = 0.0;
};
// Returns 0, APL_ERR_MESH_JACOBIAN or APL_ERR_MESH_EDGE; `bad` receives the
// offending element.
int build_geometry(const Operators &ops, int K, const double *VX,
                   const double *VY, const int *EToV, Geometry &g, int &bad);

// Structured jittered mesh (see apl_mesh_structured in apollo_dg.h).
void structured_mesh(int n, double x0, double x1, double y0, double y1,
                     double jitter, uint64_t seed, bool periodic, double *VX,
                     double *VY, int *EToV, int *vert_id, int *face_tag);

}  // namespace apl
