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

#include <type_traits>

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

// Row stride (doubles) of the transposed batch state in LDS: 16 elements plus
// one pad column, so that the staging writes (consecutive lanes two rows
// apart) spread over the LDS banks instead of all hitting one.
constexpr int XS = 17;

// rows of the transposed batch state staged in LDS: the K extent of the
// first contraction, 4-row steps for FP32/FP64 MFMA and 32-row steps for
// BF16 (rows >= Np are zero)
template <int N, typename T>
__host__ __device__ constexpr int xt_rows() {
  return sizeof(T) == 1 ? (Order<N>::Np + 31) / 32 * 32 : Order<N>::NKS1 * 4;
}

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
  int mode;                // 0: write rhs to `out`; 1: LSERK4 stage update;
                           // 2: out = ca q + cb rhs (+ cz z)
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
  const double *z;         // mode 2: third vector (own element only), or null
  double ca, cb, cz;       // mode 2 coefficients
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
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Split-BF16x3 contractions: each operand x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even), and a.b ~ ah.bh + ah.bl + al.bh
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Tag type of the
// kernels' operand-type parameter; pointwise work runs in FP32.
struct Bf16x3 {};

// one lane's A fragment of a 16x16x32 step: hi and lo halves, 32 bytes
struct __attribute__((aligned(16))) Bf8x2 {
  bf8 hi, lo;
};

template <typename T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };
template <> struct Acc<Bf16x3> { typedef f4 type; };

// scalar type of the pointwise flux work that feeds a contraction
template <typename T> struct Scal { typedef T type; };
template <> struct Scal<Bf16x3> { typedef float type; };

template <typename T> struct IsF64 { static constexpr bool value = false; };
template <> struct IsF64<double> { static constexpr bool value = true; };

template <typename T>
__host__ __device__ constexpr int acc_row(int g, int r) {
  return IsF64<T>::value ? g + 4 * r : 4 * g + r;
}

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// c += (ah + al)(bh + bl) without the al.bl term; the K index of a
// 16x16x32 step is k = 8 (lane/16) + item in both operands
__device__ __forceinline__ f4 mma3(const Bf8x2 &a, const bf8 &bh, const bf8 &bl, f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, bh, c, 0, 0, 0);
  return c;
}

__device__ __forceinline__ void split_bf16(float x, bf8 &h, bf8 &l, int it) {
  const __bf16 hi = (__bf16)x;
  h[it] = hi;
  l[it] = (__bf16)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// physics (pointwise, scalar type S)
// ---------------------------------------------------------------------------

// 1/x from the hardware reciprocal estimate refined by Newton-Raphson
// iterations r <- r + r (1 - x r) (two for FP64, one for FP32): within an
// ulp or so of the IEEE quotient for the positive, normal densities it is
// used on, in about half the instructions of the IEEE-exact division
// sequence (scale, reciprocal, refinement, fix-up).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ float recip(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

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
    const S ir = recip(rho);   // one reciprocal for the three velocities
    const S u = q[1] * ir, v = q[2] * ir;
    const S w = W ? q[3] * ir : S(0);
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

// Euler flux of one trace and its LF wave speed |u| + c
// (wxeulereqn.cc:806-871, 1106-1110) from one reciprocal of the density and
// one pressure; the density and pressure floors act as in phys_flux.
// Square roots of the LF wave speed: IEEE for FP64; the hardware's for FP32
// (within an ulp), in the FP32 face flux of the FP32 and split-BF16x3 kernels.
__device__ __forceinline__ double speed_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float speed_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int EQN, typename S = double>
__device__ __forceinline__ unsigned euler_flux_speed(const PhysParams &P, const S *q, S *F, S *G,
                                                     S *speed) {
  constexpr bool W = (EQN == APL_EQN_EULER5);
  constexpr int IE = W ? 4 : 3;
  unsigned ef = 0;
  S rho = q[0];
  if (P.min_dens == 0.0 && !(rho > S(0))) ef |= EF_DENSITY;
  if (rho < S(P.min_dens)) rho = S(P.min_dens);
  const S ir = recip(rho);
  const S u = q[1] * ir, v = q[2] * ir, w = W ? q[3] * ir : S(0);
  const S E = q[IE];
  const S k2 = u * u + v * v + w * w;
  S p = S(P.gamma - 1.0) * (E - S(0.5) * rho * k2);
  if (!(p > S(0))) {
    if (P.min_pres == 0.0) ef |= EF_PRESSURE;
    else p = S(P.min_pres);
  }
  const S ru = rho * u, rv = rho * v;
  F[0] = ru;         G[0] = rv;
  F[1] = ru * u + p; G[1] = rv * u;
  F[2] = ru * v;     G[2] = rv * v + p;
  if constexpr (W) { F[3] = ru * w; G[3] = rv * w; }
  F[IE] = (E + p) * u;
  G[IE] = (E + p) * v;
  *speed = speed_sqrt(k2) + speed_sqrt(S(P.gamma) * p * ir);
  return ef;
}

// The Euler LF flux in FP32 (wxeulereqn.cc:1088-1117): the face flux of the
// FP32 and split-BF16x3 kernels, whose fluxes enter FP32 products anyway.
template <int EQN>
__device__ __forceinline__ unsigned euler_lf_f32(const PhysParams &P, float nx, float ny,
                                                 const float *qM, const float *qP, float *fs,
                                                 float *lam) {
  constexpr int M = Vars<EQN>::M;
  float FM[M], GM[M], FP[M], GP[M], cM, cP;
  unsigned ef = euler_flux_speed<EQN, float>(P, qM, FM, GM, &cM);
  ef |= euler_flux_speed<EQN, float>(P, qP, FP, GP, &cP);
  const float l = cM > cP ? cM : cP;
#pragma unroll
  for (int c = 0; c < M; ++c)
    fs[c] = 0.5f * (nx * (FM[c] + FP[c]) + ny * (GM[c] + GP[c]) + l * (qM[c] - qP[c]));
  *lam = l;
  return ef;
}

// HLL (wxeulereqn.cc:1119-1202) and Roe (:1204-1268) fluxes of the Euler
// equations in the face's (normal, tangential) frame, in Apollo's order of
// steps.  Roe is reproduced as Apollo computes it: the back-rotation assigns
// the x-momentum slot twice, so it holds the y-momentum flux, and the
// y-momentum slot keeps the 0 its buffer starts with.  Both zero the z
// momentum flux.  *lam = max(|u| + c) of the two traces.
template <int EQN, int FLUX>
__device__ __forceinline__ unsigned euler_riemann_flux(const PhysParams &P, double nx, double ny,
                                                       const double *qMi, const double *qPi,
                                                       double *fs, double *lam) {
  constexpr int M = Vars<EQN>::M;
  constexpr bool W = (EQN == APL_EQN_EULER5);
  constexpr int IE = W ? 4 : 3;
  const double g = P.gamma;
  double qM[M], qP[M];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    qM[c] = qMi[c];
    qP[c] = qPi[c];
  }
  qM[1] = nx * qMi[1] + ny * qMi[2];
  qM[2] = -ny * qMi[1] + nx * qMi[2];
  qP[1] = nx * qPi[1] + ny * qPi[2];
  qP[2] = -ny * qPi[1] + nx * qPi[2];
  double FM[M], GM[M], FP[M], GP[M];
  unsigned ef = phys_flux<EQN, double>(P, qM, FM, GM);
  ef |= phys_flux<EQN, double>(P, qP, FP, GP);
  // rotated primitives: rho, u (normal), v, w, p
  const double iM = 1.0 / qM[0], iP = 1.0 / qP[0];
  const double uM = qM[1] * iM, vM = qM[2] * iM, wM = W ? qM[3] * iM : 0.0;
  const double uP = qP[1] * iP, vP = qP[2] * iP, wP = W ? qP[3] * iP : 0.0;
  const double mwM = W ? qM[3] : 0.0, mwP = W ? qP[3] : 0.0;
  const double pM = (g - 1.) * (qM[IE] - 0.5 * (qM[1] * qM[1] + qM[2] * qM[2] + mwM * mwM) * iM);
  const double pP = (g - 1.) * (qP[IE] - 0.5 * (qP[1] * qP[1] + qP[2] * qP[2] + mwP * mwP) * iP);
  const double cM = sqrt(g * pM * iM), cP = sqrt(g * pP * iP);
  // Roe averages
  const double HM = (qM[IE] + pM) * iM, HP = (qP[IE] + pP) * iP;
  const double rM = sqrt(qM[0]), rP = sqrt(qP[0]), irs = 1.0 / (rM + rP);
  const double u = (rM * uM + rP * uP) * irs, v = (rM * vM + rP * vP) * irs;
  const double w = (rM * wM + rP * wP) * irs, H = (rM * HM + rP * HP) * irs;
  const double c2 = (g - 1.) * (H - 0.5 * (u * u + v * v + w * w)), c = sqrt(c2);
  double fx[M];
  if constexpr (FLUX == APL_FLUX_HLL) {
    const double SL = fmin(uM - cM, u - c), SR = fmax(uP + cP, u + c);
    const double t1 = (fmin(SR, 0.0) - fmin(0.0, SL)) / (SR - SL);
    const double t2 = 1. - t1;
    const double t3 = (SR * fabs(SL) - SL * fabs(SR)) / (2. * (SR - SL));
#pragma unroll
    for (int k = 0; k < M; ++k) fx[k] = t1 * FP[k] + t2 * FM[k] - t3 * (qP[k] - qM[k]);
    fs[1] = nx * fx[1] - ny * fx[2];
    fs[2] = ny * fx[1] + nx * fx[2];
  } else {
    const double rho = rM * rP;
    const double dp = pP - pM;
    double dw1 = -0.5 * rho * (uP - uM) / c + 0.5 * dp / c2;
    double dw2 = (qP[0] - qM[0]) - dp / c2;
    double dw3 = rho * (vP - vM);
    double dw4 = 0.5 * rho * (uP - uM) / c + 0.5 * dp / c2;
    dw1 = fabs(u - c) * dw1;
    dw2 = fabs(u) * dw2;
    dw3 = fabs(u) * dw3;
    dw4 = fabs(u + c) * dw4;
#pragma unroll
    for (int k = 0; k < M; ++k) fx[k] = 0.5 * (FM[k] + FP[k]);
    fx[0] -= 0.5 * (dw1 + dw2 + dw4);
    fx[1] -= 0.5 * (dw1 * (u - c) + dw2 * u + dw4 * (u + c));
    fx[2] -= 0.5 * (dw1 * v + dw2 * v + dw3 + dw4 * v);
    fx[IE] -= 0.5 * (dw1 * (H - u * c) + dw2 * (u * u + v * v) / 2. + dw3 * v + dw4 * (H + u * c));
    fs[1] = ny * fx[1] + nx * fx[2];
    fs[2] = 0.0;
  }
  fs[0] = fx[0];
  if constexpr (W) fs[3] = 0.0;
  fs[IE] = fx[IE];
  const double lM = sqrt(uM * uM + vM * vM + wM * wM) + cM;
  const double lP = sqrt(uP * uP + vP * vP + wP * wP) + cP;
  *lam = lM > lP ? lM : lP;
  return ef;
}

// Numerical flux F*(n, qM, qP): LF (wxeulereqn.cc:1088-1117,
// wxphmaxwelleqn.cc:500-522), Euler HLL / Roe / Wave, or, for TM3, the exact
// upwind flux (LF minus the dissipation of the normal-B jump).  Returns
// flags; *lam = the fastest wave speed.
template <int EQN, int FLUX>
__device__ __forceinline__ unsigned num_flux(const PhysParams &P, double nx, double ny,
                                             const double *qM, const double *qP, double *fs,
                                             double *lam) {
  constexpr int M = Vars<EQN>::M;
  if constexpr ((EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) &&
                (FLUX == APL_FLUX_HLL || FLUX == APL_FLUX_ROE))
    return euler_riemann_flux<EQN, FLUX>(P, nx, ny, qM, qP, fs, lam);
  if constexpr ((EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) && FLUX == APL_FLUX_WAVE) {
    // Apollo's "Wave": momentum of both traces rotated to the face frame,
    // then LF with the Cartesian normal.
    double rM[M], rP[M];
#pragma unroll
    for (int c = 0; c < M; ++c) {
      rM[c] = qM[c];
      rP[c] = qP[c];
    }
    rM[1] = nx * qM[1] + ny * qM[2];
    rM[2] = -ny * qM[1] + nx * qM[2];
    rP[1] = nx * qP[1] + ny * qP[2];
    rP[2] = -ny * qP[1] + nx * qP[2];
    return num_flux<EQN, APL_FLUX_LF>(P, nx, ny, rM, rP, fs, lam);
  }
  double FM[M], GM[M], FP[M], GP[M];
  unsigned ef = 0;
  double l;
  if constexpr (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) {
    // flux and |u| + c of each trace from one reciprocal and one pressure
    double cM, cP;
    ef |= euler_flux_speed<EQN>(P, qM, FM, GM, &cM);
    ef |= euler_flux_speed<EQN>(P, qP, FP, GP, &cP);
    l = cM > cP ? cM : cP;
  } else {
    ef |= phys_flux<EQN, double>(P, qM, FM, GM);
    ef |= phys_flux<EQN, double>(P, qP, FP, GP);
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

// Apollo's 5-variable Euler state (rho, rho u, rho v, rho w, E) from the
// engine's layout: EULER4 (rho, rho u, rho v, E) embeds with rho w = 0.
template <int EQN>
__device__ __forceinline__ void load_euler5(const double *q, double *q5) {
  if constexpr (EQN == APL_EQN_EULER5) {
#pragma unroll
    for (int c = 0; c < 5; ++c) q5[c] = q[c];
  } else {
    q5[0] = q[0];
    q5[1] = q[1];
    q5[2] = q[2];
    q5[3] = 0.0;
    q5[4] = q[3];
  }
}

template <int EQN>
__device__ __forceinline__ void store_euler5(const double *q5, double *q) {
  if constexpr (EQN == APL_EQN_EULER5) {
#pragma unroll
    for (int c = 0; c < 5; ++c) q[c] = q5[c];
  } else {
    q[0] = q5[0];
    q[1] = q5[1];
    q[2] = q5[2];
    q[3] = q5[4];
  }
}

// Euler exterior states on the 5-variable state: isentropic vortex
// (wxisentropicvortexbc.cc:20-43), wall (wxeulerwallbc.cc:18-25), inflow
// (wxeulerinflowbc.cc:21-36), Dirichlet (wxeulerdirichletbc.cc:22-30); any
// other kind copies (zero gradient, wxeulerzerogradientbc.cc:18-25).
__device__ __forceinline__ void euler_bc5(int kind, const double *p, double t, double x, double y,
                                          double nx, double ny, const double *q, double *o) {
#pragma unroll
  for (int c = 0; c < 5; ++c) o[c] = q[c];
  const double pi = 3.141592653589793;
  if (kind == APL_BC_ISENTROPIC_VORTEX) {
    const double uo = p[0], vo = p[1], xo = p[2], yo = p[3], gamma = p[4], beta = p[5];
    const double xmut = x - uo * t, ymvt = y - vo * t;
    const double r = sqrt((xmut - xo) * (xmut - xo) + (ymvt - yo) * (ymvt - yo));
    const double ex = exp(1. - r * r);
    const double u = uo - beta * ex * (ymvt - yo) / (2. * pi);
    const double v = vo + beta * ex * (xmut - xo) / (2. * pi);
    const double base = 1. - ((gamma - 1.) * beta * beta * exp(2. * (1. - r * r)) / (16. * gamma * pi * pi));
    const double rho1 = pow(base, 1. / (gamma - 1));
    const double p1 = pow(rho1, gamma);
    o[0] = rho1;
    o[1] = rho1 * u;
    o[2] = rho1 * v;
    o[3] = 0.0;
    o[4] = p1 / (gamma - 1) + 0.5 * rho1 * (u * u + v * v);
  } else if (kind == APL_BC_EULER_WALL) {
    const double qn = nx * q[1] + ny * q[2];
    o[1] = q[1] - 2. * qn * nx;
    o[2] = q[2] - 2. * qn * ny;
    o[3] = 0.0;
  } else if (kind == APL_BC_EULER_INFLOW) {
    const double mach = p[0], angle = p[1], gamma = p[2], pres = p[3], rho = p[4];
    const double sound = sqrt(gamma * pres / rho);
    const double u = mach * sound * cos(angle * pi / 180.);
    const double v = mach * sound * sin(angle * pi / 180.);
    o[0] = rho;
    o[1] = rho * u;
    o[2] = rho * v;
    o[3] = 0.0;
    o[4] = pres / (gamma - 1) + 0.5 * rho * (u * u + v * v);
  } else if (kind == APL_BC_EULER_DIRICHLET) {
    const double rho = p[0], u = p[1], v = p[2], w = p[3], gamma = p[4], pres = p[5];
    o[0] = rho;
    o[1] = rho * u;
    o[2] = rho * v;
    o[3] = rho * w;
    o[4] = pres / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w);
  }
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
    if (kind != APL_BC_COPY && kind != APL_BC_NONE) {
      double q5[5], o5[5];
      load_euler5<EQN>(qM, q5);
      euler_bc5(kind, p, t, x, y, nx, ny, q5, o5);
      store_euler5<EQN>(o5, qP);
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
// batch staging helpers (one wave, 16 elements of ES doubles each)
// ---------------------------------------------------------------------------

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin, and each
// XCD has its own L2.  Batch order follows mesh order, so the face
// neighbours of batch b sit in nearby batches: the batch a workgroup takes
// is remapped so that each XCD walks one contiguous range of batches (a
// bijection of [0, nb): XCD x = b % 8 takes start(x) + b / 8).
__device__ __forceinline__ int xcd_batch(int b, int nb) {
  constexpr int NX = 8;
  const int q = nb / NX, r = nb - q * NX;
  const int x = b % NX, i = b / NX;
  return x * q + (x < r ? x : r) + i;
}

// Waves per workgroup.  Each wave owns one 16-element batch and its own LDS
// slice; workgroups of 4 such waves stream HBM about twice as fast as
// single-wave workgroups (measured on the fused-stage traffic pattern).  A
// group size W keeps eff(W) = min(waves whose LDS fits a CU's 160 KB,
// target) waves resident, where target is what the registers allow; the
// largest W with eff(W) >= 3/4 of the best eff is taken.
__host__ __device__ constexpr int lds_resident_waves(int w, int lds_bytes_per_wave) {
  return w * lds_bytes_per_wave > 163840 ? 0 : (163840 / (w * lds_bytes_per_wave)) * w;
}

__host__ __device__ constexpr int eff_waves(int w, int lds_doubles_per_wave, int target) {
  return lds_resident_waves(w, 8 * lds_doubles_per_wave) < target
             ? lds_resident_waves(w, 8 * lds_doubles_per_wave)
             : target;
}

__host__ __device__ constexpr int group_waves(int lds_doubles_per_wave, int target_waves_per_cu) {
  // best of the three sizes, then the largest size within 3/4 of it
  return 4 * eff_waves(4, lds_doubles_per_wave, target_waves_per_cu) >=
                 3 * (eff_waves(2, lds_doubles_per_wave, target_waves_per_cu) >
                              eff_waves(1, lds_doubles_per_wave, target_waves_per_cu)
                          ? eff_waves(2, lds_doubles_per_wave, target_waves_per_cu)
                          : eff_waves(1, lds_doubles_per_wave, target_waves_per_cu)) &&
                 eff_waves(4, lds_doubles_per_wave, target_waves_per_cu) > 0
             ? 4
             : (4 * eff_waves(2, lds_doubles_per_wave, target_waves_per_cu) >=
                        3 * eff_waves(1, lds_doubles_per_wave, target_waves_per_cu) &&
                    eff_waves(2, lds_doubles_per_wave, target_waves_per_cu) > 0
                    ? 2
                    : 1);
}

// Orders one wave's LDS accesses across its lanes.  A wave only touches its
// own LDS slice, so no workgroup barrier is needed: a workgroup-wide barrier
// would make the group's waves wait for each other's gathers.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The batch of this wave: workgroup g (XCD-remapped) holds batches g W ..
// g W + W - 1.  A wave past the last batch returns at once.
struct WaveBatch {
  int lane, wave, b, cnt;
  long e0;
  bool active;
};

template <int W>
__device__ __forceinline__ WaveBatch wave_batch(int K) {
  WaveBatch w;
  // lane id from mbcnt: the compiler knows its range [0, 64)
  w.lane = W == 1 ? (int)threadIdx.x : (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  w.wave = W == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform in the wave
  const int nb = (K + 15) / 16;
  const int bq = xcd_batch(blockIdx.x, gridDim.x) * W + w.wave;
  w.active = bq < nb;
  w.b = w.active ? bq : nb - 1;
  w.e0 = (long)w.b * 16;
  w.cnt = (int)min((long)16, (long)K - w.e0);
  return w;
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// XT[w*RS + e] = q[(e0 + e)*ES + w]: all loads of a full batch are issued
// before any LDS write (16-byte vector loads; the batch base is 128-byte
// aligned because 16*ES*8 is a multiple of 128).
template <int ES, int RS = XS, typename XE = double>
__device__ __forceinline__ void stage_transposed(XE *XT, const double *qb, int cnt, int lane) {
  constexpr int NV2 = 16 * ES / 2;
  constexpr int NIT = (NV2 + 63) / 64;
  if (cnt == 16) {
    dbl2 buf[NIT];
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) buf[j] = ((const dbl2 *)qb)[idx];
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
          XT[w * RS + e] = XE(buf[j][h]);
        }
      }
    }
  } else {
    for (int i = lane; i < 16 * ES; i += 64) {
      const int e = i / ES, w = i - e * ES;
      XT[w * RS + e] = XE((e < cnt) ? qb[i] : qb[w]);
    }
  }
}

// Write the batch result OUT[e*ES + w] (LDS): rhs (mode 0), the LSERK4
// stage update res = a res + dt rhs, q_out = q_in + b res (mode 1), or the
// linear combination out = ca q + cb rhs (+ cz z) of an SSP stage (mode 2;
// z may be `out` itself: each element reads its own z before writing).
template <int ES, typename OE = double>
__device__ __forceinline__ unsigned write_batch(const RhsArgs &a, const OE *OUT, long e0,
                                                int cnt, int lane) {
  constexpr int NV2 = 16 * ES / 2;
  constexpr int NIT = (NV2 + 63) / 64;
  const long base = e0 * ES;
  unsigned ef = 0;
  if (a.mode == 2) {
    if (cnt == 16) {
      dbl2 qv[NIT], zv[NIT];
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) {
          qv[j] = ((const dbl2 *)(a.q + base))[idx];
          if (a.z) zv[j] = ((const dbl2 *)(a.z + base))[idx];
        }
      }
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) {
          const dbl2 v = dbl2{OUT[2 * idx], OUT[2 * idx + 1]};
          if (v[0] != v[0] || v[1] != v[1]) ef |= EF_NAN;
          dbl2 o = a.ca * qv[j] + a.cb * v;
          if (a.z) o = o + a.cz * zv[j];
          ((dbl2 *)(a.out + base))[idx] = o;
        }
      }
    } else {
      for (int i = lane; i < cnt * ES; i += 64) {
        const double val = OUT[i];
        if (val != val) ef |= EF_NAN;
        double o = a.ca * a.q[base + i] + a.cb * val;
        if (a.z) o = o + a.cz * a.z[base + i];
        a.out[base + i] = o;
      }
    }
    return ef;
  }
  if (cnt == 16) {
    if (a.mode == 0) {
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) {
          const dbl2 v = dbl2{OUT[2 * idx], OUT[2 * idx + 1]};
          if (v[0] != v[0] || v[1] != v[1]) ef |= EF_NAN;
          ((dbl2 *)(a.out + base))[idx] = v;
        }
      }
    } else {
      dbl2 rb[NIT], qv[NIT];
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) {
          if (a.rk_a != 0.0) rb[j] = ((const dbl2 *)(a.res + base))[idx];
          qv[j] = ((const dbl2 *)(a.q + base))[idx];
        }
      }
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) {
          dbl2 r;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const double val = OUT[2 * idx + h];
            if (val != val) ef |= EF_NAN;
            r[h] = (a.rk_a == 0.0) ? a.dt * val : a.rk_a * rb[j][h] + a.dt * val;
          }
          ((dbl2 *)(a.res + base))[idx] = r;
          ((dbl2 *)(a.out + base))[idx] = qv[j] + a.rk_b * r;
        }
      }
    }
  } else {
    for (int i = lane; i < cnt * ES; i += 64) {
      const double val = OUT[i];
      if (val != val) ef |= EF_NAN;
      if (a.mode == 0) {
        a.out[base + i] = val;
      } else {
        const double r = (a.rk_a == 0.0) ? a.dt * val : a.rk_a * a.res[base + i] + a.dt * val;
        a.res[base + i] = r;
        a.out[base + i] = a.q[base + i] + a.rk_b * r;
      }
    }
  }
  return ef;
}

// ---------------------------------------------------------------------------
// the fused RHS kernel
// ---------------------------------------------------------------------------

// LDS of one wave (doubles).  The batch state is staged transposed,
// XT[(node*M + v)*XS + e], so that B-operand and face-node reads
// of the 16 lanes of a lane group hit consecutive addresses (the own face
// nodes are read from it through Fmask); the neighbours' face nodes are
// staged as PT[((f*Nfp + i)*M + v)*16 + e] in the own face's node order
// (neighbour node Fmask[nf][Nfp-1-i] for own node i).  The small tables
// I1D[3][Ng][Nfp] and Fmask[3][Nfp] follow.
template <int N, int EQN, typename T>
struct Lds {
  static constexpr int M = Vars<EQN>::M;
  static constexpr int Np = Order<N>::Np, Nfp = Order<N>::Nfp, Ng = Order<N>::Ng;
  // the staged state and neighbour face nodes: FP32 for the FP32 and
  // BF16x3 kernels (half the LDS, so more waves fit a CU), FP64 for FP64
  using XE = typename std::conditional<sizeof(T) == 8, double, float>::type;
  static constexpr int XT = xt_rows<N, T>() * M * XS;   // elements; rows >= Np are zero
  static constexpr int PT = 3 * Nfp * M * 16;           // elements
  static constexpr int XT_D = (XT * (int)sizeof(XE) + 7) / 8;   // in doubles
  static constexpr int PT_D = (PT * (int)sizeof(XE) + 7) / 8;
  static constexpr int I1D = 3 * Ng * Nfp;
  static constexpr int FM = (3 * Nfp + 1) / 2;   // ints packed into doubles
  static constexpr int TABLES = XT_D + PT_D + I1D + FM;
  static constexpr int OUT = 16 * Np * M;        // FP64 result staging, over the same space
  static constexpr int doubles = TABLES > OUT ? TABLES : OUT;
};

// Per-lane state of the cubature kernel's surface pass (XE: the staged
// state's element type).
template <typename XE>
struct FaceCtx {
  int lg, el;
  long ge;
  int code[3];
  double nx[3], ny[3], sj[3];
  const XE *XT, *PT;
  const double *I1;
  const int *FM;
  double tnow, lmax;
  unsigned ef;
};

// Scaled numerical flux fsT = F* L/2 at face Gauss point kf = 4 ks + lg of
// element el (zero past the last point): traces from the staged own and
// neighbour face nodes, the exterior state on boundary faces.
template <int N, int EQN, int FLUX, typename S, typename XE>
__device__ __forceinline__ void face_flux_cub(const RhsArgs &a, int ks, S *fsT, FaceCtx<XE> &F) {
  constexpr int M = Vars<EQN>::M;
  constexpr int Ng = Order<N>::Ng, Nfp = Order<N>::Nfp;
  const int kf = 4 * ks + F.lg;
  // FP32 face flux for the FP32 and split-BF16x3 Euler LF kernels up to
  // N = 5 (2D Euler vortex, K = 524288, N = 4: 444 -> 402 us per stage; at
  // N = 8 it takes the single-RHS error past 2e-5, and there the volume
  // dominates)
  constexpr bool F32 = sizeof(S) == 4 && FLUX == APL_FLUX_LF && N <= 5 &&
                       (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4);
  if constexpr (F32) {
    // FP32 traces and flux; the boundary state stays FP64
    if (kf < 3 * Ng) {
      const int f = kf / Ng, g = kf - f * Ng;
      const int c0 = F.code[0], c1 = F.code[1], c2 = F.code[2];
      const double x0 = F.nx[0], x1 = F.nx[1], x2 = F.nx[2];
      const double y0 = F.ny[0], y1 = F.ny[1], y2 = F.ny[2];
      const double s0 = F.sj[0], s1 = F.sj[1], s2 = F.sj[2];
      const int code = f == 0 ? c0 : (f == 1 ? c1 : c2);
      const double nx = f == 0 ? x0 : (f == 1 ? x1 : x2);
      const double ny = f == 0 ? y0 : (f == 1 ? y1 : y2);
      const double sj = f == 0 ? s0 : (f == 1 ? s1 : s2);
      float qM[M], qP[M], fs[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] = qP[v] = 0.0f;
      const double *Io = F.I1 + (f * Ng + g) * Nfp;
#pragma unroll
      for (int i = 0; i < Nfp; ++i) {
        const float w = float(Io[i]);
        const XE *xo = F.XT + F.FM[f * Nfp + i] * M * XS + F.el;
#pragma unroll
        for (int v = 0; v < M; ++v) qM[v] += w * float(xo[v * XS]);
      }
      if (code >= 0) {
        const int nf = code & 3;
        const double *In = F.I1 + (nf * Ng + (Ng - 1 - g)) * Nfp;
#pragma unroll
        for (int i = 0; i < Nfp; ++i) {
          const float w = float(In[Nfp - 1 - i]);
#pragma unroll
          for (int v = 0; v < M; ++v) qP[v] += w * float(F.PT[((f * Nfp + i) * M + v) * 16 + F.el]);
        }
      } else {
        const double *vk = a.vert + F.ge * 6;
        const int va = f, vb = (f + 1) % 3;
        const double z = a.I1D[3 * Ng * Nfp + g];
        const double sz = 0.5 * (1.0 + z);
        const double xg = vk[2 * va] + sz * (vk[2 * vb] - vk[2 * va]);
        const double yg = vk[2 * va + 1] + sz * (vk[2 * vb + 1] - vk[2 * va + 1]);
        double dM[M], dP[M];
#pragma unroll
        for (int v = 0; v < M; ++v) dM[v] = qM[v];
        boundary_state<EQN>(a.bc, -code - 1, F.tnow, xg, yg, nx, ny, dM, dP);
#pragma unroll
        for (int v = 0; v < M; ++v) qP[v] = float(dP[v]);
      }
      float lam;
      F.ef |= euler_lf_f32<EQN>(a.ph, float(nx), float(ny), qM, qP, fs, &lam);
      F.lmax = double(lam) > F.lmax ? double(lam) : F.lmax;
      const float sjf = float(sj);
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = S(fs[v] * sjf);
    } else {
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = S(0);
    }
    return;
  }
  if (kf < 3 * Ng) {
    const int f = kf / Ng, g = kf - f * Ng;
    // selects of loaded values (a select of addresses would keep F in memory)
    const int c0 = F.code[0], c1 = F.code[1], c2 = F.code[2];
    const double x0 = F.nx[0], x1 = F.nx[1], x2 = F.nx[2];
    const double y0 = F.ny[0], y1 = F.ny[1], y2 = F.ny[2];
    const double s0 = F.sj[0], s1 = F.sj[1], s2 = F.sj[2];
    const int code = f == 0 ? c0 : (f == 1 ? c1 : c2);
    const double nx = f == 0 ? x0 : (f == 1 ? x1 : x2);
    const double ny = f == 0 ? y0 : (f == 1 ? y1 : y2);
    const double sj = f == 0 ? s0 : (f == 1 ? s1 : s2);
    double qM[M], qP[M], fs[M];
#pragma unroll
    for (int v = 0; v < M; ++v) qM[v] = qP[v] = 0.0;
    const double *Io = F.I1 + (f * Ng + g) * Nfp;
#pragma unroll
    for (int i = 0; i < Nfp; ++i) {
      const double w = Io[i];
      const XE *xo = F.XT + F.FM[f * Nfp + i] * M * XS + F.el;
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] += w * double(xo[v * XS]);
    }
    if (code >= 0) {
      // own node i <-> neighbour node Fmask[nf][Nfp-1-i]; the neighbour's
      // trace at its Gauss point Ng-1-g is sum_j I1D[nf][Ng-1-g][j] q[Fmask[nf][j]]
      const int nf = code & 3;
      const double *In = F.I1 + (nf * Ng + (Ng - 1 - g)) * Nfp;
#pragma unroll
      for (int i = 0; i < Nfp; ++i) {
        const double w = In[Nfp - 1 - i];
#pragma unroll
        for (int v = 0; v < M; ++v) qP[v] += w * double(F.PT[((f * Nfp + i) * M + v) * 16 + F.el]);
      }
    } else {
      const double *vk = a.vert + F.ge * 6;
      const int va = f, vb = (f + 1) % 3;
      const double z = a.I1D[3 * Ng * Nfp + g];
      const double s = 0.5 * (1.0 + z);
      const double xg = vk[2 * va] + s * (vk[2 * vb] - vk[2 * va]);
      const double yg = vk[2 * va + 1] + s * (vk[2 * vb + 1] - vk[2 * va + 1]);
      boundary_state<EQN>(a.bc, -code - 1, F.tnow, xg, yg, nx, ny, qM, qP);
    }
    double lam;
    F.ef |= num_flux<EQN, FLUX>(a.ph, nx, ny, qM, qP, fs, &lam);
    F.lmax = lam > F.lmax ? lam : F.lmax;
#pragma unroll
    for (int v = 0; v < M; ++v) fsT[v] = S(fs[v] * sj);
  } else {
#pragma unroll
    for (int v = 0; v < M; ++v) fsT[v] = S(0);
  }
}

// BF16x3 B operands of a K = Np product: row 32 kk + 4 it + lg of the staged
// state (element el) split into hi/lo item it of step kk.
template <int NK8, int M, typename XE>
__device__ __forceinline__ void split_state(const XE *XT, int lg, int el, bf8 (&qh)[NK8][M],
                                            bf8 (&ql)[NK8][M]) {
#pragma unroll
  for (int kk = 0; kk < NK8; ++kk)
#pragma unroll
    for (int it = 0; it < 8; ++it)
#pragma unroll
      for (int v = 0; v < M; ++v)
        split_bf16(float(XT[((32 * kk + 4 * it + lg) * M + v) * XS + el]), qh[kk][v], ql[kk][v], it);
}

// Waves per SIMD the register allocator is asked to allow, by measurement
// (K = 131072, stage time one wave -> two):
//  - two for N <= 4, and for FP32 at every order (with the FP32-staged
//    state: Euler N = 7 515 -> 383 us, N = 8 616 -> 594);
//  - BF16x3: two for Maxwell at every order (N = 7 315 -> 234, N = 8
//    371 -> 276), one for Euler above N = 4 (N = 8 565 -> 739 with two);
//  - FP64: two for Maxwell up to N = 6 (N = 6 303 -> 293), one for Euler
//    above N = 4 and for N >= 7, where two waves spill (Euler N = 5
//    346 -> 373, N = 7 637 -> 1380).
//  - FP32 2D Euler (4 variables) at N = 4: four (127 VGPRs with the FP32
//    face flux; K = 524288: 402 -> 380 us; at N = 3 four measured slower,
//    K = 8192: 17.2 -> 18.6 us; the 5-variable kernel would spill at four).
template <int N, int EQN, typename T>
struct MinWaves {
  static constexpr bool TM = EQN == APL_EQN_MAXWELL_TM3;
  static constexpr int value =
      EQN == APL_EQN_PHMAXWELL8 ? 1
      : (sizeof(T) == 4 && N == 4 && EQN == APL_EQN_EULER4) ? 4
      : (N <= 4 || sizeof(T) == 4) ? 2
      : sizeof(T) == 1         ? (TM ? 2 : 1)
      : N <= 6                 ? (TM ? 2 : 1)
                               : 1;
};

// Split-BF16x3 cubature kernels stay at one wave per group: with the barriers
// of a multi-wave group their schedule no longer fits 256 VGPRs (measured).
template <int N, int EQN, typename T>
struct RhsGroup {
  static constexpr int W =
      sizeof(T) == 1 ? 1 : group_waves(Lds<N, EQN, T>::doubles, 4 * MinWaves<N, EQN, T>::value);
};

template <int N, int EQN, int FLUX, typename T>
__global__ void __launch_bounds__((64 * RhsGroup<N, EQN, T>::W), (MinWaves<N, EQN, T>::value))
    rhs_kernel(RhsArgs a) {
  using O = Order<N>;
  using L = Lds<N, EQN, T>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Nc = O::Nc, Ng = O::Ng, Nfp = O::Nfp;
  constexpr int NKS1 = O::NKS1, NCH = O::NCH, NRT = O::NRT, NKF = O::NKF;
  constexpr int ES = Np * M;  // doubles per element
  typedef typename Acc<T>::type V4;
  typedef typename Scal<T>::type S;          // FP32 for split-BF16x3
  constexpr bool BF = sizeof(T) == 1;
  // 16x16x32 k-steps of GEMM1 (K = Np) and the face GEMM (K = 3 Ng): the
  // k index 8 (lane/16) + item stands for the FP32 k-step 8 kk + item
  constexpr int NK8 = (NKS1 + 7) / 8, NKF8 = (NKF + 7) / 8;

  constexpr int W = RhsGroup<N, EQN, T>::W;
  const WaveBatch wb = wave_batch<W>(a.K);
  if (!wb.active) return;
  extern __shared__ double lds_all[];
  double *lds = lds_all + wb.wave * L::doubles;
  using XE = typename L::XE;
  XE *XT = (XE *)lds;
  XE *PT = (XE *)(lds + L::XT_D);
  double *I1 = lds + L::XT_D + L::PT_D;
  int *FM = (int *)(I1 + L::I1D);

  const int lane = wb.lane;
  const int el = lane & 15, lg = lane >> 4;
  const long e0 = wb.e0;
  const int cnt = wb.cnt;

  // element data of this lane (element el; missing elements repeat element 0)
  const long ge = e0 + (el < cnt ? el : 0);
  const double *gk = a.geom + ge * G_WORDS;
  const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];
  const S jrx_s = S(jrx), jry_s = S(jry), jsx_s = S(jsx), jsy_s = S(jsy);
  // 1/J as stored (one FP64 division per element on the host, as the
  // vector-ALU kernels use it): multiplying by it instead of dividing each
  // residual entry by J changes the result by at most an ulp and drops ~12
  // FP64 instructions per entry
  const double invJ = gk[G_INVJ];
  const double nx0 = gk[G_NX], nx1 = gk[G_NX + 1], nx2 = gk[G_NX + 2];
  const double ny0 = gk[G_NY], ny1 = gk[G_NY + 1], ny2 = gk[G_NY + 2];
  const double sj0 = gk[G_SJ], sj1 = gk[G_SJ + 1], sj2 = gk[G_SJ + 2];
  const int code0 = a.fcode[ge * 3], code1 = a.fcode[ge * 3 + 1], code2 = a.fcode[ge * 3 + 2];

  // 1. small tables
  for (int i = lane; i < L::I1D; i += 64) I1[i] = a.I1D[i];
  for (int i = lane; i < 3 * Nfp; i += 64) FM[i] = a.fmask[i];

  // 2. stage the batch state, transposed (coalesced global reads)
  stage_transposed<ES, XS, XE>(XT, a.q + e0 * ES, cnt, lane);
  for (int i = ES * XS + lane; i < L::XT; i += 64) XT[i] = XE(0);
  wave_sync();

  // 3. gather the neighbours' face nodes; item (k = f*Nfp + i, e), e fastest
  constexpr int NGI = (16 * 3 * Nfp + 63) / 64;
#pragma unroll
  for (int j = 0; j < NGI; ++j) {
    const int it = lane + 64 * j;
    if (it < 16 * 3 * Nfp) {
      const int e = it & 15, k = it >> 4, f = k / Nfp, i = k - f * Nfp;
      const long gee = e0 + (e < cnt ? e : 0);
      const int code = a.fcode[gee * 3 + f];
      double v8[M];
      if (code >= 0) {
        const long nb = code >> 2;
        const int nf = code & 3;
        const double *src = a.q + (nb * Np + FM[nf * Nfp + (Nfp - 1 - i)]) * M;
#pragma unroll
        for (int v = 0; v < M; ++v) v8[v] = src[v];
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v) v8[v] = 0.0;
      }
#pragma unroll
      for (int v = 0; v < M; ++v) PT[(k * M + v) * 16 + e] = XE(v8[v]);
    }
  }
  wave_sync();

  V4 acc[NRT][M];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int v = 0; v < M; ++v) acc[rt][v] = V4{0, 0, 0, 0};

  unsigned ef = 0;
  const S *A1 = (const S *)a.A1;
  const S *A2 = (const S *)a.A2;
  const S *A3 = (const S *)a.A3;
  const Bf8x2 *B1 = (const Bf8x2 *)a.A1;   // BF16x3: [NCH][NK8][64]
  const Bf8x2 *B2 = (const Bf8x2 *)a.A2;   // BF16x3: [NRT][NCH][64]
  const Bf8x2 *B3 = (const Bf8x2 *)a.A3;   // BF16x3: [NRT][NKF8][64]

  // BF16x3: B operands of GEMM1 (the state split into hi/lo), built once
  // when they fit a handful of registers, else per chunk
  constexpr bool HOIST = NK8 * M <= 8;
  bf8 qh[NK8][M], ql[NK8][M];
  if constexpr (BF && HOIST) split_state<NK8, M>(XT, lg, el, qh, ql);

  // 4. volume: chunks of 16 cubature points.  The A_r|A_s fragments of the
  // next chunk are loaded while the current one computes; GEMM1's operands
  // are loaded in one batch at the top of each chunk.
  S a2n[BF ? 1 : NRT][8];
  Bf8x2 f2n[BF ? NRT : 1];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) {
    if constexpr (BF) {
      f2n[rt] = B2[(size_t)(rt * NCH) * 64 + lane];
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) a2n[rt][j] = A2[((size_t)(rt * NCH) * 64 + lane) * 8 + j];
    }
  }

#pragma unroll 2
  for (int cc = 0; cc < NCH; ++cc) {
    V4 qc[M];
#pragma unroll
    for (int v = 0; v < M; ++v) qc[v] = V4{0, 0, 0, 0};
    S a2[BF ? 1 : NRT][8];
    Bf8x2 f2[BF ? NRT : 1];
    if constexpr (BF) {
      if constexpr (!HOIST) split_state<NK8, M>(XT, lg, el, qh, ql);
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) f2[rt] = f2n[rt];
      if (cc + 1 < NCH) {
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) f2n[rt] = B2[(size_t)(rt * NCH + cc + 1) * 64 + lane];
      }
#pragma unroll
      for (int kk = 0; kk < NK8; ++kk) {
        const Bf8x2 f1 = B1[(cc * NK8 + kk) * 64 + lane];
#pragma unroll
        for (int v = 0; v < M; ++v) qc[v] = mma3(f1, qh[kk][v], ql[kk][v], qc[v]);
      }
    } else {
      S a1[NKS1], b1[NKS1][M];
#pragma unroll
      for (int ks = 0; ks < NKS1; ++ks) {
        a1[ks] = A1[(cc * NKS1 + ks) * 64 + lane];
#pragma unroll
        for (int v = 0; v < M; ++v) b1[ks][v] = S(XT[((4 * ks + lg) * M + v) * XS + el]);
      }
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int j = 0; j < 8; ++j) a2[rt][j] = a2n[rt][j];
      if (cc + 1 < NCH) {
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
          for (int j = 0; j < 8; ++j)
            a2n[rt][j] = A2[((size_t)(rt * NCH + cc + 1) * 64 + lane) * 8 + j];
      }
#pragma unroll
      for (int ks = 0; ks < NKS1; ++ks)
#pragma unroll
        for (int v = 0; v < M; ++v) qc[v] = mma(a1[ks], b1[ks][v], qc[v]);
    }
    // cubature point r of this lane: flux, geometric scaling, and its two
    // rank-4 updates of the output tiles (BF16x3: items r and 4 + r of one
    // 16x16x32 step)
    bf8 gh[M], gl[M];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = 16 * cc + acc_row<T>(lg, r);
      S phr[M], phs[M];
      if (c < Nc) {
        S qv[M], F[M], G[M];
#pragma unroll
        for (int v = 0; v < M; ++v) qv[v] = qc[v][r];
        ef |= phys_flux<EQN, S>(a.ph, qv, F, G);
        if constexpr (sizeof(S) == 4) {
          // FP32 kernels: the scaled fluxes enter FP32 products, so they are
          // formed in FP32
#pragma unroll
          for (int v = 0; v < M; ++v) {
            phr[v] = jrx_s * F[v] + jry_s * G[v];
            phs[v] = jsx_s * F[v] + jsy_s * G[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < M; ++v) {
            phr[v] = S(jrx * double(F[v]) + jry * double(G[v]));
            phs[v] = S(jsx * double(F[v]) + jsy * double(G[v]));
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v) phr[v] = phs[v] = S(0);
      }
      if constexpr (BF) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
          split_bf16(phr[v], gh[v], gl[v], r);
          split_bf16(phs[v], gh[v], gl[v], 4 + r);
        }
      } else if (!IsF64<T>::value || 16 * cc + 4 * r < Nc) {
        // FP64 holds points 4r..4r+3 of the chunk in step r: a step of the
        // last chunk with no point left adds only zeros and is skipped
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
          for (int v = 0; v < M; ++v) {
            acc[rt][v] = mma(a2[rt][r], phr[v], acc[rt][v]);
            acc[rt][v] = mma(a2[rt][4 + r], phs[v], acc[rt][v]);
          }
      }
    }
    if constexpr (BF) {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int v = 0; v < M; ++v) acc[rt][v] = mma3(f2[rt], gh[v], gl[v], acc[rt][v]);
    }
  }

  // 5. surface: face Gauss point kf = 4 ks + lg of element el
  FaceCtx<XE> F;
  F.lg = lg;
  F.el = el;
  F.ge = ge;
  F.code[0] = code0; F.code[1] = code1; F.code[2] = code2;
  F.nx[0] = nx0; F.nx[1] = nx1; F.nx[2] = nx2;
  F.ny[0] = ny0; F.ny[1] = ny1; F.ny[2] = ny2;
  F.sj[0] = sj0; F.sj[1] = sj1; F.sj[2] = sj2;
  F.XT = XT;
  F.PT = PT;
  F.I1 = I1;
  F.FM = FM;
  F.tnow = (a.tptr ? *a.tptr : a.t) + a.tofs;
  F.lmax = 0.0;
  F.ef = 0;
  if constexpr (BF) {
#pragma unroll 1
    for (int kk = 0; kk < NKF8; ++kk) {
      bf8 fh[M], fl[M];
#pragma unroll
      for (int it = 0; it < 8; ++it) {
        S fsT[M];
        face_flux_cub<N, EQN, FLUX, S>(a, 8 * kk + it, fsT, F);
#pragma unroll
        for (int v = 0; v < M; ++v) split_bf16(fsT[v], fh[v], fl[v], it);
      }
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) {
        const Bf8x2 f3 = B3[(rt * NKF8 + kk) * 64 + lane];
#pragma unroll
        for (int v = 0; v < M; ++v) acc[rt][v] = mma3(f3, fh[v], fl[v], acc[rt][v]);
      }
    }
  } else {
#pragma unroll 2
    for (int ks = 0; ks < NKF; ++ks) {
      S a3[NRT];
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) a3[rt] = A3[(rt * NKF + ks) * 64 + lane];
      S fsT[M];
      face_flux_cub<N, EQN, FLUX, S>(a, ks, fsT, F);
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int v = 0; v < M; ++v) acc[rt][v] = mma(a3[rt], fsT[v], acc[rt][v]);
    }
  }

  double lmax = F.lmax;
  ef |= F.ef;

  // 6. epilogue: x 1/J, stage through LDS as [e][n][v], coalesced write
  wave_sync();
  double *OUT = lds;
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * rt + acc_row<T>(lg, r);
      if (n < Np) {
#pragma unroll
        for (int v = 0; v < M; ++v) OUT[el * ES + n * M + v] = double(acc[rt][v][r]) * invJ;
      }
    }
  wave_sync();
  if (wb.active) ef |= write_batch<ES>(a, OUT, e0, cnt, lane);
  // wave-wide max speed and flags
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off);
    lmax = o > lmax ? o : lmax;
    ef |= __shfl_xor(ef, off);
  }
  if (lane == 0 && wb.active) {
    if (a.speed_part) a.speed_part[wb.b] = lmax;
    if (ef) atomicOr(a.err, ef);
  }
}

// ---------------------------------------------------------------------------
// the nodal (linear-flux) RHS kernel
// ---------------------------------------------------------------------------
//
// For a linear flux F(q) = Ax q, G(q) = Ay q the cubature volume term is
// A_r Phi_r + A_s Phi_s with Phi_r = (jrx Ax + jry Ay)(V_c Q), which equals
// Dt_r Q (jrx Ax + jry Ay)^T with Dt_r = A_r V_c (Np x Np), and the face term
// L S(Gauss points) equals Lt S(face nodes) with Lt = L blkdiag(I1D) because
// F* is linear in (q-, q+) and the boundary states are linear reflections.
// Same result to rounding, 4-5x fewer MFMAs:
//   P_r = Dt_r Q, P_s = Dt_s Q        GEMM   (Np x Np)(Np x 16) per var
//   R   = jrx F(P_r) + jry G(P_r) + jsx F(P_s) + jsy G(P_s)   at the nodes
//   R  -= Lt S,  S = F*(q-, q+) L/2 at the 3(N+1) face nodes
//   rhs = R / J

// LDS of the nodal kernel: only the transposed batch state (the own face
// nodes are read from it through Fmask; the neighbours' face nodes go straight
// to registers), so that LDS does not bound the number of resident waves.
template <int N, int EQN, typename T>
struct LinLds {
  static constexpr int M = Vars<EQN>::M;
  // The staged state, which feeds the FP32 B operands and the face traces
  // whose fluxes enter the FP32 products: FP32 (half the LDS) for BF16x3
  // and for FP32 at N >= 6, where it measured faster (K = 294912, N = 8:
  // FP32 371 -> 356 us per stage, BF16x3 385 -> 350; N = 5 BF16x3 68.4 ->
  // 66.1); FP64 for the FP64 kernel and for FP32 at N <= 5, where the
  // conversions cost registers at four waves per SIMD (64.7 -> 70.4 us).
  using XE = typename std::conditional<sizeof(T) == 8 || (sizeof(T) == 4 && N <= 5), double,
                                       float>::type;
  static constexpr int XT = xt_rows<N, T>() * M * XS;          // elements of XE
  static constexpr int doubles = (XT * (int)sizeof(XE) + 7) / 8;
};

template <int N>
struct LinOrder {
  static constexpr int NKP = (3 * Order<N>::Nfp + 3) / 4;   // k-steps of the face GEMM
};

// The neighbour's values at face node kf = f Nfp + i of an element (its
// node Fmask[nf][Nfp-1-i], given as fnb[nf]); zero on boundary faces and
// past the last face node.
template <int Np, int Nfp, int M>
__device__ __forceinline__ void gather_face_node(const double *q, int kf, int code0, int code1,
                                                 int code2, const int (&fnb)[3], double *dst) {
  const int f = kf / Nfp;
  const int code = f == 0 ? code0 : (f == 1 ? code1 : code2);
  if (kf < 3 * Nfp && code >= 0) {
    const int nf = code & 3;
    const int node = nf == 0 ? fnb[0] : (nf == 1 ? fnb[1] : fnb[2]);
    const double *src = q + ((long)(code >> 2) * Np + node) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) dst[v] = src[v];
  } else {
#pragma unroll
    for (int v = 0; v < M; ++v) dst[v] = 0.0;
  }
}

// Waves per SIMD requested from the register allocator (a lower bound: the
// FP64 kernel fits three at N = 5 and four or more below N = 5).  FP64
// requests two at every order: at N = 7, 8 that costs a few spilled VGPRs
// (1 and 5) but beats one wave with 380 registers (N = 8, K = 294912: 446
// against 511 us per stage; N = 7: 394 against 419).  Requesting four at
// N = 5 spills 14 VGPRs and runs 107 instead of 71 us per stage (K =
// 131072).  FP32 and BF16x3 at three waves for N >= 7 measured no faster
// than two.
template <int N, int EQN, typename T>
struct LinWaves {
  static constexpr int value = EQN == APL_EQN_PHMAXWELL8 ? 1
                              : sizeof(T) == 8 ? 2
                                               : (N <= 5 ? 4 : (N == 6 ? 3 : 2));   // FP32, BF16x3
};

// The nodal kernel keeps one wave per group: it is register-bound at the
// same waves per CU either way, and a multi-wave group holds its LDS until
// its slowest wave ends (measured at N=5: 65.5 us per stage with one wave,
// 69.0 us with four).
template <int N, int EQN, typename T>
struct LinGroup {
  static constexpr int W = 1;
};

template <int N, int EQN, int FLUX, typename T>
__global__ void __launch_bounds__((64 * LinGroup<N, EQN, T>::W), (LinWaves<N, EQN, T>::value))
    rhs_linear_kernel(RhsArgs a) {
  using O = Order<N>;
  using L = LinLds<N, EQN, T>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Nfp = O::Nfp;
  constexpr int NKS1 = O::NKS1, NRT = O::NRT, NKP = LinOrder<N>::NKP;
  constexpr int ES = Np * M;
  typedef typename Acc<T>::type V4;
  typedef typename Scal<T>::type S;          // FP32 for split-BF16x3
  constexpr bool BF = sizeof(T) == 1;
  // 16x16x32 k-steps of the volume (K = Np) and face (K = 3 Nfp) products;
  // k = 8 (lane/16) + item stands for the FP32 k-step 8 kk + item, so each
  // lane's 8 items are the values it holds for eight 4-row steps
  constexpr int NK8 = (NKS1 + 7) / 8, NKP8 = (NKP + 7) / 8;

  constexpr int W = LinGroup<N, EQN, T>::W;
  const WaveBatch wb = wave_batch<W>(a.K);
  if (!wb.active) return;
  extern __shared__ double lds_all[];
  using XE = typename L::XE;
  XE *XT = (XE *)(lds_all + wb.wave * L::doubles);

  const int lane = wb.lane;
  const int el = lane & 15, lg = lane >> 4;
  const long e0 = wb.e0;
  const int cnt = wb.cnt;

  const long ge = e0 + (el < cnt ? el : 0);
  const double *gk = a.geom + ge * G_WORDS;
  const int code0 = a.fcode[ge * 3], code1 = a.fcode[ge * 3 + 1], code2 = a.fcode[ge * 3 + 2];

  // face node kf = 4 j + lg of element el: own node Fmask[f][i] and the
  // neighbour's node Fmask[nf][Nfp-1-i] for each possible neighbour face nf
  int fown[NKP], fnb[NKP][3];
#pragma unroll
  for (int j = 0; j < NKP; ++j) {
    const int kf = 4 * j + lg;
    const int f = kf < 3 * Nfp ? kf / Nfp : 0, i = kf < 3 * Nfp ? kf - f * Nfp : 0;
    fown[j] = a.fmask[f * Nfp + i];
#pragma unroll
    for (int nf = 0; nf < 3; ++nf) fnb[j][nf] = a.fmask[nf * Nfp + (Nfp - 1 - i)];
  }

  // Neighbour face-node values (zero on boundary faces).  FP32 and BF16x3 up
  // to 5 variables load them all into registers before the batch staging,
  // so their latency overlaps it; FP64 and phMaxwell load them per k-step
  // (FP64, measured: K = 131072, N = 5 80.7 -> 77.5 us per stage;
  // K = 294912, N = 8 455.7 -> 441.9 us; FP32 without the early gather:
  // N = 5 66.3 -> 68.2 us, N = 8 364 -> 388 us).
  const double *qg = a.q;
  constexpr bool PRE = M <= 5 && sizeof(T) != 8;
  double qP[PRE ? NKP : 1][M];
  if constexpr (PRE) {
#pragma unroll
    for (int j = 0; j < NKP; ++j)
      gather_face_node<Np, Nfp, M>(qg, 4 * j + lg, code0, code1, code2, fnb[j], qP[j]);
  }

  stage_transposed<ES, XS, XE>(XT, a.q + e0 * ES, cnt, lane);
  for (int i = ES * XS + lane; i < L::XT; i += 64) XT[i] = XE(0);
  wave_sync();

  unsigned ef = 0;
  const S *A1 = (const S *)a.A1;   // [tile][NKS1][64]: Dt_r, Dt_s (tiles below)
  const S *A3 = (const S *)a.A3;   // [NRT][NKP][64]: -Lt
  const Bf8x2 *B1 = (const Bf8x2 *)a.A1;   // BF16x3: [2][NRT][NK8][64]
  const Bf8x2 *B3 = (const Bf8x2 *)a.A3;   // BF16x3: [NRT][NKP8][64]
  bf8 fh[M], fl[M];                          // BF16x3: face fluxes of one k-step

  V4 acc[NRT][M];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int v = 0; v < M; ++v) acc[rt][v] = V4{0, 0, 0, 0};

  // surface at the face nodes kf = 4 ks + lg: R = -Lt S
  double lmax = 0.0;
#pragma unroll
  for (int ks = 0; ks < NKP; ++ks) {
    const int kf = 4 * ks + lg;
    S a3[NRT];
    if constexpr (!BF) {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) a3[rt] = A3[(rt * NKP + ks) * 64 + lane];
    }
    S fsT[M];
    if (kf < 3 * Nfp) {
      const int f = kf / Nfp;
      const int code = f == 0 ? code0 : (f == 1 ? code1 : code2);
      const double nx = gk[G_NX + f], ny = gk[G_NY + f], sj = gk[G_SJ + f];
      double qM[M], fs[M];
      double *qp = qP[PRE ? ks : 0];
      if constexpr (!PRE) gather_face_node<Np, Nfp, M>(qg, kf, code0, code1, code2, fnb[ks], qp);
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] = double(XT[(fown[ks] * M + v) * XS + el]);
      if (code < 0) boundary_state<EQN>(a.bc, -code - 1, 0.0, 0.0, 0.0, nx, ny, qM, qp);
      double lam;
      ef |= num_flux<EQN, FLUX>(a.ph, nx, ny, qM, qp, fs, &lam);
      lmax = lam > lmax ? lam : lmax;
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = S(fs[v] * sj);
    } else {
#pragma unroll
      for (int v = 0; v < M; ++v) fsT[v] = S(0);
    }
    if constexpr (BF) {
      const int it = ks & 7;
#pragma unroll
      for (int v = 0; v < M; ++v) split_bf16(fsT[v], fh[v], fl[v], it);
      if (it == 7 || ks == NKP - 1) {
#pragma unroll
        for (int j = it + 1; j < 8; ++j)
#pragma unroll
          for (int v = 0; v < M; ++v) fh[v][j] = fl[v][j] = (__bf16)0.0f;
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
          const Bf8x2 f3 = B3[(rt * NKP8 + ks / 8) * 64 + lane];
#pragma unroll
          for (int v = 0; v < M; ++v) acc[rt][v] = mma3(f3, fh[v], fl[v], acc[rt][v]);
        }
      }
    } else {
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int v = 0; v < M; ++v) acc[rt][v] = mma(a3[rt], fsT[v], acc[rt][v]);
    }
  }

  // volume at the nodes: P_r = Dt_r Q then P_s = Dt_s Q, each mixed by the
  // flux matrices and the geometric factors straight into the accumulator
  const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];
  // B operands: held in registers for FP32; re-read from LDS per k-step for
  // FP64, whose doubled register footprint would otherwise cost a wave/SIMD
  constexpr bool KEEP_B = sizeof(T) == 4;
  S b1[KEEP_B ? NKS1 : 1][M];
  if constexpr (KEEP_B) {
#pragma unroll
    for (int ks = 0; ks < NKS1; ++ks)
#pragma unroll
      for (int v = 0; v < M; ++v) b1[ks][v] = S(XT[((4 * ks + lg) * M + v) * XS + el]);
  }
  // BF16x3: the state split once into hi/lo B operands
  bf8 bh[BF ? NK8 : 1][M], bl[BF ? NK8 : 1][M];
  if constexpr (BF) {
#pragma unroll
    for (int kk = 0; kk < NK8; ++kk)
#pragma unroll
      for (int it = 0; it < 8; ++it)
#pragma unroll
        for (int v = 0; v < M; ++v)
          split_bf16(float(XT[((32 * kk + 4 * it + lg) * M + v) * XS + el]), bh[kk][v], bl[kk][v], it);
  }

  // Operand tiles of the volume products.  FP32 and BF16x3: [d][rt], tile t
  // = the derivative t & 1 of output rows 16 (t >> 1) ... .  FP64: Dt_r and
  // Dt_s stacked in one operand, Dt_s from row OFF (Np rounded up to a
  // multiple of 4), cut into 16-row tiles that may straddle the two.  An
  // FP64 accumulator holds rows g + 4 r of its tile in lane group g; a shift
  // by a multiple of 4 keeps each row's lane group, so element r of a
  // stacked tile is, in every lane, element r2 of the accumulator tile rt2
  // of one derivative d.  Fewer tiles where 2 Np is short of 32 NRT (N = 5:
  // 3 instead of 4, 54 MFMAs instead of 72); the same products in the same
  // order.
  constexpr bool STK = IsF64<T>::value;
  constexpr int OFF = (Np + 3) / 4 * 4;
  constexpr int NVT = STK ? (OFF + Np + 15) / 16 : 2 * NRT;
#pragma unroll
  for (int t = 0; t < NVT; ++t) {
    const int tile = STK ? t : ((t & 1) * NRT + (t >> 1));
    V4 p[M];
#pragma unroll
    for (int v = 0; v < M; ++v) p[v] = V4{0, 0, 0, 0};
    if constexpr (BF) {
#pragma unroll
      for (int kk = 0; kk < NK8; ++kk) {
        const Bf8x2 f1 = B1[(tile * NK8 + kk) * 64 + lane];
#pragma unroll
        for (int v = 0; v < M; ++v) p[v] = mma3(f1, bh[kk][v], bl[kk][v], p[v]);
      }
    } else {
      S ad[NKS1];
#pragma unroll
      for (int ks = 0; ks < NKS1; ++ks) ad[ks] = A1[(tile * NKS1 + ks) * 64 + lane];
#pragma unroll
      for (int ks = 0; ks < NKS1; ++ks)
#pragma unroll
        for (int v = 0; v < M; ++v) {
          const S b = KEEP_B ? b1[KEEP_B ? ks : 0][v] : S(XT[((4 * ks + lg) * M + v) * XS + el]);
          p[v] = mma(ad[ks], b, p[v]);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // stacked rows 16 t + 4 r + g are nodes n0 + g of derivative d
      const int s0 = 16 * t + 4 * r;
      const int d = STK ? (s0 >= OFF ? 1 : 0) : (t & 1);
      const int n0 = d == 1 && STK ? s0 - OFF : s0;
      const int rt2 = STK ? n0 / 16 : (t >> 1), r2 = STK ? (n0 % 16) / 4 : r;
      if (rt2 >= NRT) continue;   // padding rows only
      const double cx = d == 0 ? jrx : jsx, cy = d == 0 ? jry : jsy;
      S pv[M], F[M], G[M];
#pragma unroll
      for (int v = 0; v < M; ++v) pv[v] = p[v][r];
      phys_flux<EQN, S>(a.ph, pv, F, G);
      if constexpr (sizeof(S) == 8) {
        // two multiply-adds per entry
#pragma unroll
        for (int v = 0; v < M; ++v) {
          acc[rt2][v][r2] += cx * F[v];
          acc[rt2][v][r2] += cy * G[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v)
          acc[rt2][v][r2] += S(cx * double(F[v]) + cy * double(G[v]));
      }
    }
  }

  // epilogue: x 1/J (the stored reciprocal, as in rhs_kernel), stage through
  // LDS as [e][n][v], coalesced write
  const double invJ = gk[G_INVJ];
  wave_sync();
  XE *OUT = XT;
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * rt + acc_row<T>(lg, r);
      if (n < Np) {
#pragma unroll
        for (int v = 0; v < M; ++v) OUT[el * ES + n * M + v] = XE(double(acc[rt][v][r]) * invJ);
      }
    }
  wave_sync();
  if (wb.active) ef |= write_batch<ES>(a, OUT, e0, cnt, lane);
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off);
    lmax = o > lmax ? o : lmax;
    ef |= __shfl_xor(ef, off);
  }
  if (lane == 0 && wb.active) {
    if (a.speed_part) a.speed_part[wb.b] = lmax;
    if (ef) atomicOr(a.err, ef);
  }
}

// Advance the device clock by dt at the end of a captured step.
int launch_advance_time(hipStream_t s, double *t, double dt);

// out[0] = max(in[0 .. n)), one workgroup.
int launch_max_reduce(hipStream_t s, const double *in, int n, double *out);

// Launch entry of one order (defined in apl_kernels_N<N>.hip); `linear`
// selects the nodal kernel (Maxwell only).
template <int N>
int launch_rhs(int eqn, int flux, int prec, int linear, hipStream_t s, const RhsArgs &a,
               int nblocks);

}  // namespace apl
