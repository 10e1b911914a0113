// apl_host.cpp -- operators, connectivity, geometry and synthetic meshes.
//
// Host-side setup of the engine, in FP64.  The formulas follow Apollo's
// setup code so that the operators agree with Apollo's to rounding:
//   nodes / Fmask     src/lib/wxpnodaldgfunctions.h:112-278 (tables)
//   V, V^-1, V V^T    src/lib/wxnodaldggeometry2d.cc:49-76
//   basis             src/lib/wxNodalDGMatrices.h:7-168, src/lib/wxmath.h:303-368
//   cubature ops      src/lib/wxcubature2d.cc:18-133
//   connectivity      src/lib/wxnodaldggeometry2d.cc:235-367
//   geometry          src/lib/wxnodaldggeometry2d.cc:179-233,382-442
#include "apl_host.h"

#include <cmath>
#include <cstring>
#include <random>
#include <unordered_map>

#include "../../include/apollo_dg.h"
#include "apl_tables.h"

namespace apl {

// ---------------------------------------------------------------------------
// polynomial basis
// ---------------------------------------------------------------------------

// Orthonormal Jacobi polynomial P_n^(alpha,beta) by the three-term
// recurrence of Hesthaven & Warburton (2008), eq. (A.1)-(A.4).
double jacobi_p(double x, double alpha, double beta, int n) {
  const double ab = alpha + beta, ab1 = ab + 1.0, a1 = alpha + 1.0, b1 = beta + 1.0;
  const double g0 = std::pow(2.0, ab1) / ab1 * std::tgamma(a1) * std::tgamma(b1) / std::tgamma(ab1);
  double p_prev = 1.0 / std::sqrt(g0);
  if (n == 0) return p_prev;
  const double g1 = a1 * b1 / (ab + 3.0) * g0;
  double p_cur = ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / std::sqrt(g1);
  if (n == 1) return p_cur;
  double a_old = 2.0 / (2.0 + ab) * std::sqrt(a1 * b1 / (ab + 3.0));
  for (int i = 1; i < n; ++i) {
    const double h = 2.0 * i + ab;
    const double a_new = 2.0 / (h + 2.0) *
        std::sqrt((i + 1) * (i + ab1) * (i + a1) * (i + b1) / (h + 1.0) / (h + 3.0));
    const double b_new = -(alpha * alpha - beta * beta) / h / (h + 2.0);
    const double p_next = 1.0 / a_new * (-a_old * p_prev + (x - b_new) * p_cur);
    a_old = a_new;
    p_prev = p_cur;
    p_cur = p_next;
  }
  return p_cur;
}

double grad_jacobi_p(double x, double alpha, double beta, int n) {
  if (n == 0) return 0.0;
  return std::sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1);
}

// Collapsed coordinates (a, b) of (r, s).
static inline void collapse(double r, double s, double &a, double &b) {
  b = s;
  a = (s == 1.0) ? -1.0 : 2.0 * (1.0 + r) / (1.0 - s) - 1.0;
}

// Mode (i,j) of the orthonormal PKDO basis on the reference triangle.
static double pkdo(double a, double b, int i, int j) {
  return std::sqrt(2.0) * jacobi_p(a, 0.0, 0.0, i) * jacobi_p(b, 2.0 * i + 1, 0.0, j) *
         std::pow(1.0 - b, (double)i);
}

static void pkdo_grad(double a, double b, int i, int j, double &dr, double &ds) {
  const double fa = jacobi_p(a, 0.0, 0.0, i), dfa = grad_jacobi_p(a, 0.0, 0.0, i);
  const double gb = jacobi_p(b, 2.0 * i + 1, 0.0, j), dgb = grad_jacobi_p(b, 2.0 * i + 1, 0.0, j);
  const double half1mb = 0.5 * (1.0 - b);
  dr = dfa * gb;
  ds = dfa * gb * 0.5 * (1.0 + a);
  if (i > 0) {
    dr *= std::pow(half1mb, i - 1.0);
    ds *= std::pow(half1mb, i - 1.0);
  }
  double t = dgb * std::pow(half1mb, (double)i);
  if (i > 0) t -= 0.5 * i * gb * std::pow(half1mb, i - 1.0);
  ds += fa * t;
  const double norm = std::pow(2.0, i + 0.5);
  dr *= norm;
  ds *= norm;
}

// V[k*Np + mode] = phi_mode(r_k, s_k), modes ordered i outer, j inner.
static void vandermonde(int N, int n, const double *r, const double *s, double *V) {
  const int Np = (N + 1) * (N + 2) / 2;
  for (int k = 0; k < n; ++k) {
    double a, b;
    collapse(r[k], s[k], a, b);
    int mode = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j) V[k * Np + mode++] = pkdo(a, b, i, j);
  }
}

static void grad_vandermonde(int N, int n, const double *r, const double *s, double *Vr, double *Vs) {
  const int Np = (N + 1) * (N + 2) / 2;
  for (int k = 0; k < n; ++k) {
    double a, b;
    collapse(r[k], s[k], a, b);
    int mode = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j, ++mode) pkdo_grad(a, b, i, j, Vr[k * Np + mode], Vs[k * Np + mode]);
  }
}

// Gauss-Legendre nodes and weights by Newton's method on P_n (the
// Numerical Recipes `gauleg` scheme Apollo uses, src/lib/wxmath.h:171-206):
// iterate until |dz| <= 3e-13 and take the weight from P_n' at the last
// iterate before the update, which reproduces Apollo's weights bit for bit.
void gauss_legendre(int n, double *z, double *w) {
  const double pi = 3.14159265358979323846;
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double x = std::cos(pi * (i - 0.25) / (n + 0.5)), x_prev, dp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int k = 1; k <= n; ++k) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * k - 1.0) * x * p2 - (k - 1.0) * p3) / k;
      }
      dp = n * (x * p1 - p2) / (x * x - 1.0);
      x_prev = x;
      x = x_prev - p1 / dp;
    } while (std::fabs(x - x_prev) > 3.0e-13);
    z[i - 1] = -x;
    z[n - i] = x;
    w[i - 1] = w[n - i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

bool invert(int n, const double *A, double *Ainv) {
  std::vector<double> M(A, A + (size_t)n * n), I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(M[(size_t)i * n + c]) > std::fabs(M[(size_t)p * n + c])) p = i;
    if (M[(size_t)p * n + c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < n; ++j) {
        std::swap(M[(size_t)p * n + j], M[(size_t)c * n + j]);
        std::swap(I[(size_t)p * n + j], I[(size_t)c * n + j]);
      }
    for (int i = c + 1; i < n; ++i) {
      const double l = M[(size_t)i * n + c] / M[(size_t)c * n + c];
      if (l == 0.0) continue;
      for (int j = c; j < n; ++j) M[(size_t)i * n + j] -= l * M[(size_t)c * n + j];
      for (int j = 0; j < n; ++j) I[(size_t)i * n + j] -= l * I[(size_t)c * n + j];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    const double d = M[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) I[(size_t)c * n + j] /= d;
    for (int i = 0; i < c; ++i) {
      const double l = M[(size_t)i * n + c];
      if (l == 0.0) continue;
      for (int j = 0; j < n; ++j) I[(size_t)i * n + j] -= l * I[(size_t)c * n + j];
    }
  }
  std::memcpy(Ainv, I.data(), sizeof(double) * (size_t)n * n);
  return true;
}

static void gemm(int M, int K, int N, const double *A, const double *B, double *C) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double acc = 0.0;
      for (int k = 0; k < K; ++k) acc += A[(size_t)i * K + k] * B[(size_t)k * N + j];
      C[(size_t)i * N + j] = acc;
    }
}

bool build_operators(int N, Operators &o) {
  if (N < 1 || N > APL_TAB_MAX_ORDER) return false;
  o = Operators();
  o.N = N;
  o.Np = (N + 1) * (N + 2) / 2;
  o.Nfp = N + 1;
  o.Ng = 2 * (N + 1);                            // wxcubature2d.cc:23
  const int cub_order = 3 * (N + 1);             // wxcubature2d.cc:20
  o.Nc = apl_tab_cub_npts[N - 1];                // wxcubature2d.cc:26-28 (rule cub_order)
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng;
  o.r.assign(apl_tab_node_r + apl_tab_node_off[N - 1], apl_tab_node_r + apl_tab_node_off[N]);
  o.s.assign(apl_tab_node_s + apl_tab_node_off[N - 1], apl_tab_node_s + apl_tab_node_off[N]);
  o.fmask.assign(apl_tab_fmask + apl_tab_fmask_off[N - 1], apl_tab_fmask + apl_tab_fmask_off[N]);
  o.rmin = std::fabs(o.r[1] - o.r[0]);           // wxnodaldggeometry2d.cc:91

  o.V.resize((size_t)Np * Np);
  o.invV.resize((size_t)Np * Np);
  o.VVT.resize((size_t)Np * Np);
  vandermonde(N, Np, o.r.data(), o.s.data(), o.V.data());
  if (!invert(Np, o.V.data(), o.invV.data())) return false;
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      double acc = 0.0;
      for (int k = 0; k < Np; ++k) acc += o.V[(size_t)i * Np + k] * o.V[(size_t)j * Np + k];
      o.VVT[(size_t)i * Np + j] = acc;
    }

  o.cr.resize(Nc);
  o.cs.resize(Nc);
  o.cw.resize(Nc);
  const double *rule = apl_tab_cub_rsw + apl_tab_cub_off[N - 1];   // cub2D_<cub_order>
  for (int k = 0; k < Nc; ++k) {
    o.cr[k] = rule[3 * k];
    o.cs[k] = rule[3 * k + 1];
    o.cw[k] = rule[3 * k + 2];
  }
  std::vector<double> Vq((size_t)Nc * Np), Vr((size_t)Nc * Np), Vs((size_t)Nc * Np);
  vandermonde(N, Nc, o.cr.data(), o.cs.data(), Vq.data());
  grad_vandermonde(N, Nc, o.cr.data(), o.cs.data(), Vr.data(), Vs.data());
  o.Vc.resize((size_t)Nc * Np);
  o.Drc.resize((size_t)Nc * Np);
  o.Dsc.resize((size_t)Nc * Np);
  gemm(Nc, Np, Np, Vq.data(), o.invV.data(), o.Vc.data());
  gemm(Nc, Np, Np, Vr.data(), o.invV.data(), o.Drc.data());
  gemm(Nc, Np, Np, Vs.data(), o.invV.data(), o.Dsc.data());

  o.gz.resize(Ng);
  o.gw.resize(Ng);
  gauss_legendre(Ng, o.gz.data(), o.gw.data());
  // face parameterisation (wxcubature2d.cc:118-124)
  std::vector<double> fr(3 * Ng), fs(3 * Ng), gV((size_t)3 * Ng * Np);
  for (int k = 0; k < Ng; ++k) {
    fr[k] = o.gz[k];            fs[k] = -1.0;
    fr[Ng + k] = -o.gz[k];      fs[Ng + k] = o.gz[k];
    fr[2 * Ng + k] = -1.0;      fs[2 * Ng + k] = -o.gz[k];
  }
  vandermonde(N, 3 * Ng, fr.data(), fs.data(), gV.data());
  o.If.resize((size_t)3 * Ng * Np);
  gemm(3 * Ng, Np, Np, gV.data(), o.invV.data(), o.If.data());
  return true;
}

// ---------------------------------------------------------------------------
// connectivity
// ---------------------------------------------------------------------------

int face_pair(int K, int Nv, const int *EToV, const int *vert_id, const int *face_tag,
              std::vector<int> &etetf) {
  static const int vn[3][2] = {{0, 1}, {1, 2}, {2, 0}};   // wxnodaldggeometry2d.cc:263
  const int F = 3 * K;
  std::vector<int> partner(F, -1);
  std::unordered_map<uint64_t, int> open;
  open.reserve((size_t)F);
  int err = 0;
  for (int fid = 0; fid < F; ++fid) {
    const int e = fid / 3, f = fid % 3;
    int a = EToV[3 * e + vn[f][0]], b = EToV[3 * e + vn[f][1]];
    if (vert_id) { a = vert_id[a]; b = vert_id[b]; }
    const uint64_t key = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
    auto it = open.find(key);
    if (it == open.end()) {
      open.emplace(key, fid);
    } else if (it->second >= 0) {
      partner[fid] = it->second;
      partner[it->second] = fid;
      it->second = -1;           // closed: a third face on this edge is an error
    } else {
      err = APL_ERR_MESH_TOPOLOGY;
    }
  }
  (void)Nv;
  etetf.assign((size_t)6 * K, -1);                          // :336-340
  if (face_tag) {
    // "Face Sets" labels (:348-356): every entry of the face's first
    // supporting cell, faces visited in creation order (first appearance
    // over cells, local faces in cone order).
    for (int fid = 0; fid < F; ++fid) {
      const int p = partner[fid];
      if (p >= 0 && p < fid) continue;       // created earlier by cell p/3
      int tag = face_tag[fid];
      if (p >= 0 && tag == 0) tag = face_tag[p];
      if (tag != 0)
        for (int k = 0; k < 6; ++k) etetf[(size_t)6 * (fid / 3) + k] = -tag;
    }
  }
  for (int fid = 0; fid < F; ++fid) {                       // :359-363
    const int p = partner[fid];
    if (p < 0) continue;
    etetf[(size_t)6 * (fid / 3) + 2 * (fid % 3)] = p / 3;
    etetf[(size_t)6 * (fid / 3) + 2 * (fid % 3) + 1] = p % 3;
  }
  return err;
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

int build_geometry(const Operators &o, int K, const double *VX, const double *VY,
                   const int *EToV, Geometry &g, int &bad) {
  const int Np = o.Np, Nfp = o.Nfp;
  g.x.resize((size_t)K * Np);
  g.y.resize((size_t)K * Np);
  g.geom.resize((size_t)K * 14);
  g.vert.resize((size_t)K * 6);
  double dts = 1.0e6;
  bad = -1;
  for (int k = 0; k < K; ++k) {
    const double p1x = VX[EToV[3 * k]], p1y = VY[EToV[3 * k]];
    const double p2x = VX[EToV[3 * k + 1]], p2y = VY[EToV[3 * k + 1]];
    const double p3x = VX[EToV[3 * k + 2]], p3y = VY[EToV[3 * k + 2]];
    double *vk = &g.vert[(size_t)6 * k];
    vk[0] = p1x; vk[1] = p1y; vk[2] = p2x; vk[3] = p2y; vk[4] = p3x; vk[5] = p3y;
    double *xk = &g.x[(size_t)k * Np], *yk = &g.y[(size_t)k * Np];
    for (int n = 0; n < Np; ++n) {          // affine map, :211-218
      const double r = o.r[n], s = o.s[n];
      xk[n] = 0.5 * (-p1x * (r + s) + p2x * (1. + r) + p3x * (1. + s));
      yk[n] = 0.5 * (-p1y * (r + s) + p2y * (1. + r) + p3y * (1. + s));
    }
    // inscribed radius (:220-227)
    const double l1 = std::sqrt(std::pow(p1x - p2x, 2) + std::pow(p1y - p2y, 2));
    const double l2 = std::sqrt(std::pow(p2x - p3x, 2) + std::pow(p2y - p3y, 2));
    const double l3 = std::sqrt(std::pow(p3x - p1x, 2) + std::pow(p3y - p1y, 2));
    const double sp = 0.5 * (l1 + l2 + l3);
    const double area = std::sqrt(std::fabs(sp * (sp - l1) * (sp - l2) * (sp - l3)));
    dts = std::min(dts, area / sp);
    // geometric factors from the three vertex nodes (:412-442)
    const double x1 = xk[o.fmask[0]], y1 = yk[o.fmask[0]];
    const double x2 = xk[o.fmask[Nfp]], y2 = yk[o.fmask[Nfp]];
    const double x3 = xk[o.fmask[2 * Nfp]], y3 = yk[o.fmask[2 * Nfp]];
    const double xr = (x2 - x1) / 2, xs = (x3 - x1) / 2, yr = (y2 - y1) / 2, ys = (y3 - y1) / 2;
    const double J = -xs * yr + xr * ys;
    if (J <= 0) { bad = k; return APL_ERR_MESH_JACOBIAN; }
    double *gk = &g.geom[(size_t)14 * k];
    gk[0] = ys / J;
    gk[1] = -yr / J;
    gk[2] = -xs / J;
    gk[3] = xr / J;
    gk[4] = J;
    // outward unit normals and half edge lengths (:382-410)
    const double ex[3] = {x2 - x1, x3 - x2, x1 - x3}, ey[3] = {y2 - y1, y3 - y2, y1 - y3};
    for (int f = 0; f < 3; ++f) {
      const double nx = ey[f], ny = -ex[f];
      const double len = std::sqrt(nx * nx + ny * ny);
      if (len <= 0) { bad = k; return APL_ERR_MESH_EDGE; }
      gk[5 + 3 * f] = nx / len;
      gk[6 + 3 * f] = ny / len;
      gk[7 + 3 * f] = len / 2.;
    }
  }
  g.dtscale = dts;
  return 0;
}

// ---------------------------------------------------------------------------
// synthetic meshes
// ---------------------------------------------------------------------------

void structured_mesh(int n, double x0, double x1, double y0, double y1, double jitter,
                     uint64_t seed, bool periodic, double *VX, double *VY, int *EToV,
                     int *vert_id, int *face_tag) {
  const double hx = (x1 - x0) / n, hy = (y1 - y0) / n;
  std::mt19937_64 rng(seed);
  auto uniform = [&]() { return (double)(rng() >> 11) * (1.0 / 9007199254740992.0); };
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) {
      const int v = j * (n + 1) + i;
      double x = x0 + i * hx, y = y0 + j * hy;
      if (i == n) x = x1;
      if (j == n) y = y1;
      if (i > 0 && i < n && j > 0 && j < n && jitter != 0.0) {
        x += (2.0 * uniform() - 1.0) * jitter * hx;
        y += (2.0 * uniform() - 1.0) * jitter * hy;
      }
      VX[v] = x;
      VY[v] = y;
      if (vert_id) vert_id[v] = periodic ? (j % n) * (n + 1) + (i % n) : v;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const int q = j * n + i;
      const int v00 = j * (n + 1) + i, v10 = v00 + 1, v01 = v00 + (n + 1), v11 = v01 + 1;
      int *t0 = EToV + 6 * q, *t1 = t0 + 3;
      t0[0] = v00; t0[1] = v10; t0[2] = v11;
      t1[0] = v00; t1[1] = v11; t1[2] = v01;
      if (face_tag) {
        int *f0 = face_tag + 6 * q, *f1 = f0 + 3;
        const int tag = periodic ? 0 : 1;
        f0[0] = (j == 0) ? tag : 0;       // bottom edge (v00,v10)
        f0[1] = (i == n - 1) ? tag : 0;   // right edge (v10,v11)
        f0[2] = 0;                        // diagonal
        f1[0] = 0;                        // diagonal
        f1[1] = (j == n - 1) ? tag : 0;   // top edge (v11,v01)
        f1[2] = (i == 0) ? tag : 0;       // left edge (v01,v00)
      }
    }
}

}  // namespace apl
