// apl_device.h -- device code of the nodal-DG RHS engine (gfx950).
//
// One 64-lane wave owns a batch of 16 elements.  Columns of every GEMM are
// (variable, element) pairs, one 16-column MFMA tile per variable, so that a
// lane always sees element e = lane % 16.  The RHS of Apollo's nodalDG2d
// (src/subsolvers/wxnodaldg2dmethod.cc:320-510) is evaluated as
//
//   Qc  = V_c Q                         GEMM1   (Nc x Np)(Np x 16)   per var
//   Phi_r = J (rx F(Qc) + ry G(Qc)),  Phi_s = J (sx F + sy G)   pointwise
//   R  += A_r Phi_r + A_s Phi_s         GEMM2   (Np x Nc)(Nc x 16)   per var
//   S   = F*(QM, QP) * L/2              face Gauss points, pointwise
//   R  -= L S                           GEMM3   (Np x 3Ng)(3Ng x 16) per var
//   rhs = R / J
//
// with A_r = V V^T Dr_c^T diag(w), A_s = V V^T Ds_c^T diag(w) and
// L = V V^T I_f^T diag(gw) folded on the host.  Cubature points are processed
// in chunks of 16: GEMM1's accumulator of a chunk holds, in lane l, the
// cubature points acc_row(l/16, r) (r = 0..3) of element l%16, which is
// exactly the B operand GEMM2 needs if the k index of its r-th MFMA is taken
// as acc_row(l/16, r); the operator fragments A_r/A_s are stored with their
// columns permuted accordingly.  Nothing but the state and the neighbour face nodes
// is staged through LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/apollo_dg.h"

namespace apl {

// ---------------------------------------------------------------------------
// compile-time sizes
// ---------------------------------------------------------------------------

__host__ __device__ constexpr int cub_points(int N) {
  // cubPoints[3(N+1)-1] (src/lib/wxcubature2d.cc:26-28)
  return N == 1 ? 12 : N == 2 ? 19 : N == 3 ? 36 : N == 4 ? 54 : N == 5 ? 73
       : N == 6 ? 93 : N == 7 ? 118 : 145;
}

template <int N_>
struct Order {
  static constexpr int N = N_;
  static constexpr int Np = (N + 1) * (N + 2) / 2;
  static constexpr int Nfp = N + 1;
  static constexpr int Ng = 2 * (N + 1);
  static constexpr int Nc = cub_points(N);
  static constexpr int NKS1 = (Np + 3) / 4;      // k-steps of GEMM1 (K = 4)
  static constexpr int NCH = (Nc + 15) / 16;     // cubature chunks
  static constexpr int NRT = (Np + 15) / 16;     // output row tiles
  static constexpr int NKF = (3 * Ng + 3) / 4;   // k-steps of the face GEMM
};

template <int EQN> struct Vars;
template <> struct Vars<APL_EQN_EULER5> { static constexpr int M = 5; };
template <> struct Vars<APL_EQN_EULER4> { static constexpr int M = 4; };
template <> struct Vars<APL_EQN_PHMAXWELL8> { static constexpr int M = 8; };
template <> struct Vars<APL_EQN_MAXWELL_TM3> { static constexpr int M = 3; };

// per-element FP64 geometry words
enum {
  G_JRX = 0, G_JRY = 1, G_JSX = 2, G_JSY = 3, G_INVJ = 4, G_J = 5,
  G_NX = 6, G_NY = 9, G_SJ = 12, G_WORDS = 16
};

// error flag bits (device word)
enum { EF_NAN = 1u, EF_DENSITY = 2u, EF_PRESSURE = 4u };

struct PhysParams {
  double gamma, min_dens, min_pres, c0, chi, gph;
};

struct BCTable {
  int n;
  int kind[APL_MAX_BC];
  double p[APL_MAX_BC][8];
};

struct RhsArgs {
  int K;
  int mode;                // 0: write rhs to `out`; 1: LSERK4 stage update
  const double *q;         // [K][Np][M] stage input
  double *out;             // rhs (mode 0) or the stage output state (mode 1)
  double *res;             // LSERK4 residual (mode 1)
  const double *geom;      // [K][G_WORDS]
  const double *vert;      // [K][6]
  const int *fcode;        // [K][3]: nbr*4+nf (interior) or -(bc index + 1)
  const void *A1;          // V_c fragments        [NCH][NKS1][64]
  const void *A2;          // A_r|A_s fragments    [NRT][NCH][64][8]
  const void *A3;          // -L fragments         [NRT][NKF][64]
  const double *I1D;       // [3][Ng][Nfp] face-node -> Gauss interpolation
  const int *fmask;        // [3][Nfp]
  const double *tptr;      // device time (graph replay) or null
  double t, tofs;          // time = (tptr ? *tptr : t) + tofs
  double rk_a, rk_b, dt;   // stage coefficients (mode 1)
  double *speed_part;      // [nbatch] max LF speed per batch, or null
  unsigned *err;           // error flags
  PhysParams ph;
  BCTable bc;
};

// ---------------------------------------------------------------------------
// MFMA wrappers, 16x16x4: A[i][k] in lane i+16k, B[k][j] in lane j+16k.
// The accumulator layout differs by type (measured on gfx950):
//   f32 (and f32 += bf16 16x16x32): D[i][j] in lane j+16*(i/4), element i%4
//   f64:                            D[i][j] in lane j+16*(i%4), element i/4
// acc_row<T>(g, r) is the row held in element r of lane group g = lane/16.
// ---------------------------------------------------------------------------

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };

template <typename T>
__host__ __device__ constexpr int acc_row(int g, int r) {
  return sizeof(T) == 8 ? g + 4 * r : 4 * g + r;
}

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// physics (pointwise, scalar type S)
// ---------------------------------------------------------------------------

// Physical fluxes F (x) and G (y) of WxEulerEqn::flux (wxeulereqn.cc:806-871)
// and WxPHMaxwellEqn::flux (wxphmaxwelleqn.cc:447-498).  EULER4 is EULER5
// with rho w == 0 and that component dropped; TM3 is the (Ez, Bx, By) subset
// of phMaxwell with chi == gamma == 0.
template <int EQN, typename S>
__device__ __forceinline__ unsigned phys_flux(const PhysParams &P, const S *q, S *F, S *G) {
  unsigned ef = 0;
  if constexpr (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) {
    constexpr bool W = (EQN == APL_EQN_EULER5);
    constexpr int IE = W ? 4 : 3;
    S rho = q[0];
    if (P.min_dens == 0.0 && !(rho > S(0))) ef |= EF_DENSITY;
    if (rho < S(P.min_dens)) rho = S(P.min_dens);
    const S u = q[1] / rho, v = q[2] / rho;
    const S w = W ? q[3] / rho : S(0);
    const S E = q[IE];
    S p = S(P.gamma - 1.0) * (E - S(0.5) * rho * (u * u + v * v + w * w));
    if (!(p > S(0))) {
      if (P.min_pres == 0.0) ef |= EF_PRESSURE;
      else p = S(P.min_pres);
    }
    const S ru = rho * u, rv = rho * v;
    F[0] = ru;        G[0] = rv;
    F[1] = ru * u + p; G[1] = rv * u;
    F[2] = ru * v;    G[2] = rv * v + p;
    if constexpr (W) { F[3] = ru * w; G[3] = rv * w; }
    F[IE] = (E + p) * u;
    G[IE] = (E + p) * v;
  } else if constexpr (EQN == APL_EQN_PHMAXWELL8) {
    const S c2 = S(P.c0 * P.c0), chi = S(P.chi), g = S(P.gph);
    F[0] = chi * c2 * q[6]; F[1] = c2 * q[5];  F[2] = -c2 * q[4]; F[3] = g * q[7];
    F[4] = -q[2];           F[5] = q[1];       F[6] = chi * q[0]; F[7] = g * c2 * q[3];
    G[1] = chi * c2 * q[6]; G[2] = c2 * q[3];  G[0] = -c2 * q[5]; G[4] = g * q[7];
    G[5] = -q[0];           G[3] = q[2];       G[6] = chi * q[1]; G[7] = g * c2 * q[4];
  } else {
    const S c2 = S(P.c0 * P.c0);
    F[0] = -c2 * q[2]; F[1] = S(0); F[2] = -q[0];
    G[0] = c2 * q[1];  G[1] = q[0]; G[2] = S(0);
  }
  return ef;
}

// LF wave speed |u| + c (wxeulereqn.cc:1106-1110, primitive variables
// :1324-1341).
template <int EQN>
__device__ __forceinline__ double euler_speed(const PhysParams &P, const double *q) {
  constexpr bool W = (EQN == APL_EQN_EULER5);
  constexpr int IE = W ? 4 : 3;
  const double r = q[0];
  const double mw = W ? q[3] : 0.0;
  const double p = (P.gamma - 1.) * (q[IE] - 0.5 * (q[1] * q[1] + q[2] * q[2] + mw * mw) / r);
  const double u = q[1] / r, v = q[2] / r, w = mw / r;
  return sqrt(u * u + v * v + w * w) + sqrt(P.gamma * p / r);
}

// Numerical flux F*(n, qM, qP): LF (wxeulereqn.cc:1088-1117,
// wxphmaxwelleqn.cc:500-522) or, for TM3, the exact upwind flux (LF minus
// the dissipation of the normal-B jump).  Returns flags; *lam = LF speed.
template <int EQN, int FLUX>
__device__ __forceinline__ unsigned num_flux(const PhysParams &P, double nx, double ny,
                                             const double *qM, const double *qP, double *fs,
                                             double *lam) {
  constexpr int M = Vars<EQN>::M;
  double FM[M], GM[M], FP[M], GP[M];
  unsigned ef = phys_flux<EQN, double>(P, qM, FM, GM);
  ef |= phys_flux<EQN, double>(P, qP, FP, GP);
  double l;
  if constexpr (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) {
    const double cM = euler_speed<EQN>(P, qM), cP = euler_speed<EQN>(P, qP);
    l = cM > cP ? cM : cP;
  } else {
    const double a = P.chi * P.c0, b = P.gph * P.c0;
    const double m1 = a > b ? a : b;
    l = m1 > P.c0 ? m1 : P.c0;
  }
#pragma unroll
  for (int c = 0; c < M; ++c)
    fs[c] = 0.5 * (nx * (FM[c] + FP[c]) + ny * (GM[c] + GP[c]) + l * (qM[c] - qP[c]));
  if constexpr (FLUX == APL_FLUX_UPWIND && EQN == APL_EQN_MAXWELL_TM3) {
    const double jb = nx * (qM[1] - qP[1]) + ny * (qM[2] - qP[2]);
    fs[1] -= 0.5 * l * nx * jb;
    fs[2] -= 0.5 * l * ny * jb;
  }
  *lam = l;
  return ef;
}

// Exterior state at a boundary Gauss point (wxgridbc.cc:21-26 and
// wxisentropicvortexbc.cc:20-43, wxmaxwelltranversemagneticbc.cc:13-25,
// wxphmaxwellconductingwallbc.cc:12-41).
template <int EQN>
__device__ __forceinline__ void boundary_state(const BCTable &B, int idx, double t, double x,
                                               double y, double nx, double ny, const double *qM,
                                               double *qP) {
  constexpr int M = Vars<EQN>::M;
  const int kind = (idx >= 0 && idx < B.n) ? B.kind[idx] : APL_BC_COPY;
  const double *p = B.p[(idx >= 0 && idx < B.n) ? idx : 0];
#pragma unroll
  for (int c = 0; c < M; ++c) qP[c] = qM[c];
  if constexpr (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) {
    if (kind == APL_BC_ISENTROPIC_VORTEX) {
      const double uo = p[0], vo = p[1], xo = p[2], yo = p[3], gamma = p[4], beta = p[5];
      const double pi = 3.141592653589793;
      const double xmut = x - uo * t, ymvt = y - vo * t;
      const double r = sqrt((xmut - xo) * (xmut - xo) + (ymvt - yo) * (ymvt - yo));
      const double ex = exp(1. - r * r);
      const double u = uo - beta * ex * (ymvt - yo) / (2. * pi);
      const double v = vo + beta * ex * (xmut - xo) / (2. * pi);
      const double base = 1. - ((gamma - 1.) * beta * beta * exp(2. * (1. - r * r)) / (16. * gamma * pi * pi));
      const double rho1 = pow(base, 1. / (gamma - 1));
      const double p1 = pow(rho1, gamma);
      qP[0] = rho1;
      qP[1] = rho1 * u;
      qP[2] = rho1 * v;
      if constexpr (EQN == APL_EQN_EULER5) {
        qP[3] = 0.0;
        qP[4] = p1 / (gamma - 1) + 0.5 * rho1 * (u * u + v * v);
      } else {
        qP[3] = p1 / (gamma - 1) + 0.5 * rho1 * (u * u + v * v);
      }
    }
  } else if constexpr (EQN == APL_EQN_PHMAXWELL8) {
    if (kind == APL_BC_MAXWELL_TM) {
      qP[2] = -qM[2];
    } else if (kind == APL_BC_PEC) {
      const double ex = qM[0], ey = qM[1], bx = qM[3], by = qM[4];
      const double en = ex * nx + ey * ny, et = ex * ny - ey * nx;
      qP[0] = en * nx - et * ny;
      qP[1] = en * ny + et * nx;
      qP[2] = -qM[2];
      const double bn = -bx * nx - by * ny, bt = -bx * ny + by * nx;
      qP[3] = bn * nx - bt * ny;
      qP[4] = bn * ny + bt * nx;
      qP[6] = -qM[6];
    }
  } else {  // TM3: (Ez, Bx, By)
    if (kind == APL_BC_MAXWELL_TM) {
      qP[0] = -qM[0];
    } else if (kind == APL_BC_PEC) {
      const double bx = qM[1], by = qM[2];
      const double bn = -bx * nx - by * ny, bt = -bx * ny + by * nx;
      qP[0] = -qM[0];
      qP[1] = bn * nx - bt * ny;
      qP[2] = bn * ny + bt * nx;
    }
  }
}

// ---------------------------------------------------------------------------
// the fused RHS kernel
// ---------------------------------------------------------------------------

template <int N, int EQN>
struct Lds {
  static constexpr int M = Vars<EQN>::M;
  static constexpr int X = 16 * Order<N>::Np * M;          // own state
  static constexpr int QP = 16 * 3 * Order<N>::Nfp * M;    // neighbour face nodes
  static constexpr int doubles = X + QP;
};

template <int N, int EQN, int FLUX, typename T>
__global__ void __launch_bounds__(64) rhs_kernel(RhsArgs a) {
  using O = Order<N>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Nc = O::Nc, Ng = O::Ng, Nfp = O::Nfp;
  constexpr int NKS1 = O::NKS1, NCH = O::NCH, NRT = O::NRT, NKF = O::NKF;
  constexpr int ES = Np * M;  // doubles per element
  typedef typename Acc<T>::type V4;

  extern __shared__ double lds[];
  double *X = lds;                       // [16][Np][M]
  double *QPn = lds + Lds<N, EQN>::X;    // [16][3][Nfp][M]

  const int lane = threadIdx.x;
  const int el = lane & 15, lg = lane >> 4;
  const long e0 = (long)blockIdx.x * 16;
  const int cnt = (int)min((long)16, (long)a.K - e0);

  // 1. stage the batch's state (coalesced; missing elements repeat element 0)
  const double *qb = a.q + e0 * ES;
  for (int i = lane; i < 16 * ES; i += 64) X[i] = (i < cnt * ES) ? qb[i] : qb[i % ES];

  // 2. gather the neighbours' face nodes
  for (int it = lane; it < 16 * 3 * Nfp; it += 64) {
    const int e = it / (3 * Nfp), rem = it % (3 * Nfp), f = rem / Nfp, i = rem % Nfp;
    const long ge = e0 + (e < cnt ? e : 0);
    const int code = a.fcode[ge * 3 + f];
    double *dst = QPn + ((e * 3 + f) * Nfp + i) * M;
    if (code >= 0) {
      const long nb = code >> 2;
      const int nf = code & 3;
      const double *src = a.q + (nb * Np + a.fmask[nf * Nfp + i]) * M;
#pragma unroll
      for (int v = 0; v < M; ++v) dst[v] = src[v];
    }
  }
  __syncthreads();

  // 3. geometry of this lane's element
  const long ge = e0 + (el < cnt ? el : 0);
  const double *gk = a.geom + ge * G_WORDS;
  const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];

  V4 acc[NRT][M];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int v = 0; v < M; ++v) acc[rt][v] = V4{0, 0, 0, 0};

  unsigned ef = 0;
  const T *A1 = (const T *)a.A1;
  const T *A2 = (const T *)a.A2;
  const T *A3 = (const T *)a.A3;
  const double *Xe = X + el * ES;

  // 4. volume: chunks of 16 cubature points
#pragma unroll 1
  for (int cc = 0; cc < NCH; ++cc) {
    V4 qc[M];
#pragma unroll
    for (int v = 0; v < M; ++v) qc[v] = V4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < NKS1; ++ks) {
      const T av = A1[(cc * NKS1 + ks) * 64 + lane];
      const int node = 4 * ks + lg;
#pragma unroll
      for (int v = 0; v < M; ++v) {
        const T bv = (node < Np) ? T(Xe[node * M + v]) : T(0);
        qc[v] = mma(av, bv, qc[v]);
      }
    }
    T phr[M][4], phs[M][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = 16 * cc + acc_row<T>(lg, r);
      if (c < Nc) {
        T qv[M], F[M], G[M];
#pragma unroll
        for (int v = 0; v < M; ++v) qv[v] = qc[v][r];
        ef |= phys_flux<EQN, T>(a.ph, qv, F, G);
#pragma unroll
        for (int v = 0; v < M; ++v) {
          phr[v][r] = T(jrx * double(F[v]) + jry * double(G[v]));
          phs[v][r] = T(jsx * double(F[v]) + jsy * double(G[v]));
        }
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v) phr[v][r] = phs[v][r] = T(0);
      }
    }
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
      const T *frag = A2 + ((size_t)(rt * NCH + cc) * 64 + lane) * 8;
      T a2[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) a2[j] = frag[j];
#pragma unroll
      for (int v = 0; v < M; ++v)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          acc[rt][v] = mma(a2[r], phr[v][r], acc[rt][v]);
          acc[rt][v] = mma(a2[4 + r], phs[v][r], acc[rt][v]);
        }
    }
  }

  // 5. surface: face Gauss point kf = 4 ks + lg of element el
  const double tnow = (a.tptr ? *a.tptr : a.t) + a.tofs;
  double lmax = 0.0;
#pragma unroll 1
  for (int ks = 0; ks < NKF; ++ks) {
    const int kf = 4 * ks + lg;
    T fsT[M];
    if (kf < 3 * Ng) {
      const int f = kf / Ng, g = kf % Ng;
      double qM[M], qP[M], fs[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] = 0.0;
      const double *I = a.I1D + (f * Ng + g) * Nfp;
#pragma unroll
      for (int i = 0; i < Nfp; ++i) {
        const double w = I[i];
        const double *src = Xe + a.fmask[f * Nfp + i] * M;
#pragma unroll
        for (int v = 0; v < M; ++v) qM[v] += w * src[v];
      }
      const int code = a.fcode[ge * 3 + f];
      const double nx = gk[G_NX + f], ny = gk[G_NY + f];
      if (code >= 0) {
        const int nf = code & 3;
        const double *In = a.I1D + (nf * Ng + (Ng - 1 - g)) * Nfp;
        const double *src = QPn + (el * 3 + f) * Nfp * M;
#pragma unroll
        for (int v = 0; v < M; ++v) qP[v] = 0.0;
#pragma unroll
        for (int i = 0; i < Nfp; ++i) {
          const double w = In[i];
#pragma unroll
          for (int v = 0; v < M; ++v) qP[v] += w * src[i * M + v];
        }
      } else {
        // Gauss point coordinates on the straight edge
        const double *vk = a.vert + ge * 6;
        const int va = f, vb = (f + 1) % 3;
        // I1D rows carry the 1D abscissa implicitly; recover z from the
        // Gauss rule stored after I1D
        const double z = a.I1D[3 * Ng * Nfp + g];
        const double s = 0.5 * (1.0 + z);
        const double xg = vk[2 * va] + s * (vk[2 * vb] - vk[2 * va]);
        const double yg = vk[2 * va + 1] + s * (vk[2 * vb + 1] - vk[2 * va + 1]);
        boundary_state<EQN>(a.bc, -code - 1, tnow, xg, yg, nx, ny, qM, qP);
      }
      double lam;
      ef |= num_flux<EQN, FLUX>(a.ph, nx, ny, qM, qP, fs, &lam);
      lmax = lam > lmax ? lam : lmax;
      const double sj = gk[G_SJ + f];
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = T(fs[v] * sj);
    } else {
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = T(0);
    }
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
      const T a3 = A3[(rt * NKF + ks) * 64 + lane];
#pragma unroll
      for (int v = 0; v < M; ++v) acc[rt][v] = mma(a3, fsT[v], acc[rt][v]);
    }
  }

  // 6. epilogue: / J, stage through LDS, coalesced write
  __syncthreads();
  const double Jel = gk[G_J];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * rt + acc_row<T>(lg, r);
      if (n < Np) {
#pragma unroll
        for (int v = 0; v < M; ++v) X[el * ES + n * M + v] = double(acc[rt][v][r]) / Jel;
      }
    }
  __syncthreads();
  const long base = e0 * ES;
  if (a.mode == 0) {
    for (int i = lane; i < cnt * ES; i += 64) {
      const double val = X[i];
      if (val != val) ef |= EF_NAN;
      a.out[base + i] = val;
    }
  } else {
    for (int i = lane; i < cnt * ES; i += 64) {
      const double val = X[i];
      if (val != val) ef |= EF_NAN;
      const double r = (a.rk_a == 0.0) ? a.dt * val : a.rk_a * a.res[base + i] + a.dt * val;
      a.res[base + i] = r;
      a.out[base + i] = a.q[base + i] + a.rk_b * r;
    }
  }
  // wave-wide max speed and flags
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off);
    lmax = o > lmax ? o : lmax;
    ef |= __shfl_xor(ef, off);
  }
  if (lane == 0) {
    if (a.speed_part) a.speed_part[blockIdx.x] = lmax;
    if (ef) atomicOr(a.err, ef);
  }
}

// Advance the device clock by dt at the end of a captured step.
int launch_advance_time(hipStream_t s, double *t, double dt);

// Launch entry of one order (defined in apl_kernels_N<N>.hip).
template <int N>
int launch_rhs(int eqn, int flux, int prec, hipStream_t s, const RhsArgs &a, int nblocks);

}  // namespace apl
