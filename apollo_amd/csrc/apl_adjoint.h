// apl_adjoint.h -- discrete adjoint of the cubature-form RHS (FP64, gfx950).
//
// No counterpart in the reference.  For rhs = (1/J)[A_r Phi_r + A_s Phi_s -
// L S] (apl_device.h), the transpose applied to lam is, per element,
//   lh     = lam / J
//   Phib_r = A_r^T lh, Phib_s = A_s^T lh                   (Nc x m)
//   Fb = jrx Phib_r + jsx Phib_s,  Gb = jry Phib_r + jsy Phib_s
//   out   += V_c^T [ (dF/dq)^T Fb + (dG/dq)^T Gb ](V_c Q)
//   Sb     = -L^T lh                                      (3Ng x m)
//   out   += I_f^T qbM,  qbM = (dF*/dq-)^T Sb sJ  (own flux)
//                            + (dF*_nbr/dq+)^T Sb_nbr sJ_nbr (neighbour's flux,
//                              whose exterior state is this element's trace)
//                            [+ (dqBC/dq-)^T (dF*/dq+)^T Sb sJ on a boundary]
// Kernel A writes Sb for every element; kernel B gathers the neighbours' Sb,
// so every element writes only its own output (no atomics).  In a reverse
// LSERK4 stage lam = rbar + b qbar, and kernel B applies the updates
// qbar += dt out, rbar = a lam; it then forms the next reverse stage's
// lam' = rbar + b' qbar and writes that stage's Sb into a second buffer
// (its own element only), so that each reverse stage after the first is one
// launch of kernel B: the neighbours' Sb it reads were written by the
// previous launch.
#pragma once

#include "apl_device.h"

namespace apl {

struct AdjArgs {
  int K;
  int mode;                 // 0: out = J^T lam; 1: reverse LSERK4 stage
  const double *q;          // forward state of the stage
  const double *lam;        // mode 0: lam; mode 1: rbar (updated in place)
  double *qbar;             // mode 1: qbar (kernel A reads, kernel B updates)
  double *out;              // mode 0 output
  double *sbar;             // [K][3Ng][M] face weights -L^T lh (kernel A -> B)
  const double *geom;
  const double *vert;
  const int *fcode;
  const double *I1D;        // [3][Ng][Nfp] + Gauss abscissae
  const int *fmask;
  const double *Avc;        // V_c fragments        [NCH][NKS1][64]
  const double *Aart;       // A_r^T fragments      [NCH][NKS1][64]
  const double *Aast;       // A_s^T fragments      [NCH][NKS1][64]
  const double *Avct;       // V_c^T fragments      [NRT][NCH][64][4] (acc_row permuted)
  const double *Alt;        // -L^T fragments       [NFT][NKS1][64]
  const double *Aift;       // I_f^T fragments      [NRT][NKF][64]
  double *sbar_out;         // mode 1: the next stage's Sb (kernel B), or null
  int face_pass;            // 1: kernel A computes Sb before kernel B runs
  double t;                 // time of the stage (boundary states)
  double rk_a, rk_b, dt;    // reverse stage coefficients (mode 1)
  double rk_b_next;         // mode 1: b of the next reverse stage
  unsigned *err;
  PhysParams ph;
  BCTable bc;
};

template <int N>
struct AdjOrder {
  static constexpr int NFT = (3 * Order<N>::Ng + 15) / 16;   // face-point tiles
};

// ---------------------------------------------------------------------------
// pointwise vector-Jacobian products
// ---------------------------------------------------------------------------

// qb += (dF_d/dq)^T fb for one direction of the Euler flux; `mu` is the
// momentum normal to the direction, t1/t2 the tangential ones (t2 < 0 when
// absent, EULER4).  Smooth branch (no density/pressure floors).
template <int EQN>
__device__ __forceinline__ void euler_dir_vjp(const PhysParams &P, const double *q, int mu, int t1,
                                              int t2, const double *fb, double *qb) {
  constexpr int IE = (EQN == APL_EQN_EULER5) ? 4 : 3;
  const double g1 = P.gamma - 1.0;
  const double rho = q[0], u = q[mu] / rho, v = q[t1] / rho, w = t2 >= 0 ? q[t2] / rho : 0.0;
  const double E = q[IE];
  const double k2 = u * u + v * v + w * w;
  const double p = g1 * (E - 0.5 * rho * k2);
  const double pr = 0.5 * g1 * k2, pu = -g1 * u, pv = -g1 * v, pw = -g1 * w, pE = g1;
  const double b0 = fb[0], bu = fb[mu], bv = fb[t1], bw = t2 >= 0 ? fb[t2] : 0.0, b4 = fb[IE];
  qb[0] += bu * (-u * u + pr) + bv * (-u * v) + bw * (-u * w) + b4 * (pr * u - (E + p) * u / rho);
  qb[mu] += b0 + bu * (2 * u + pu) + bv * v + bw * w + b4 * (pu * u + (E + p) / rho);
  qb[t1] += bu * pv + bv * u + b4 * pv * u;
  if (t2 >= 0) qb[t2] += bu * pw + bw * u + b4 * pw * u;
  qb[IE] += bu * pE + b4 * (1 + pE) * u;
}

// qb += (dF/dq)^T Fb + (dG/dq)^T Gb
template <int EQN>
__device__ __forceinline__ void flux_vjp(const PhysParams &P, const double *q, const double *Fb,
                                         const double *Gb, double *qb) {
  if constexpr (EQN == APL_EQN_EULER5) {
    euler_dir_vjp<EQN>(P, q, 1, 2, 3, Fb, qb);
    euler_dir_vjp<EQN>(P, q, 2, 3, 1, Gb, qb);
  } else if constexpr (EQN == APL_EQN_EULER4) {
    euler_dir_vjp<EQN>(P, q, 1, 2, -1, Fb, qb);
    euler_dir_vjp<EQN>(P, q, 2, 1, -1, Gb, qb);
  } else if constexpr (EQN == APL_EQN_PHMAXWELL8) {
    const double c2 = P.c0 * P.c0, chi = P.chi, g = P.gph;
    // transpose of phys_flux's F rows
    qb[6] += chi * c2 * Fb[0]; qb[5] += c2 * Fb[1]; qb[4] += -c2 * Fb[2]; qb[7] += g * Fb[3];
    qb[2] += -Fb[4];           qb[1] += Fb[5];      qb[0] += chi * Fb[6]; qb[3] += g * c2 * Fb[7];
    // and of its G rows
    qb[6] += chi * c2 * Gb[1]; qb[3] += c2 * Gb[2]; qb[5] += -c2 * Gb[0]; qb[7] += g * Gb[4];
    qb[0] += -Gb[5];           qb[2] += Gb[3];      qb[1] += chi * Gb[6]; qb[4] += g * c2 * Gb[7];
  } else {
    const double c2 = P.c0 * P.c0;
    qb[2] += -c2 * Fb[0]; qb[0] += -Fb[2];
    qb[1] += c2 * Gb[0];  qb[0] += Gb[1];
  }
}

// d c / d q for the LF speed c = |u| + sqrt(gamma p / rho); returns c.
template <int EQN>
__device__ __forceinline__ double euler_speed_grad(const PhysParams &P, const double *q, double *dc) {
  constexpr bool W = (EQN == APL_EQN_EULER5);
  constexpr int IE = W ? 4 : 3;
  const double g = P.gamma, g1 = g - 1.0;
  const double rho = q[0];
  const double mw = W ? q[3] : 0.0;
  const double u[3] = {q[1] / rho, q[2] / rho, mw / rho};
  const double k2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double p = g1 * (q[IE] - 0.5 * (q[1] * q[1] + q[2] * q[2] + mw * mw) / rho);
  const double sp = sqrt(k2), a = sqrt(g * p / rho);
  const double da_dp = g / (2 * a * rho);
  dc[0] = da_dp * 0.5 * g1 * k2 - a / (2 * rho);
  const int nmom = W ? 3 : 2;
  for (int i = 0; i < nmom; ++i) {
    const double du = sp > 0 ? u[i] / sp : 0.0;
    dc[0] += -du * u[i] / rho;
    dc[1 + i] = du / rho + da_dp * (-g1 * u[i]);
  }
  dc[IE] = da_dp * g1;
  return sp + a;
}

// qMb += (dF*/dqM)^T nb and qPb += (dF*/dqP)^T nb for the numerical flux.
template <int EQN, int FLUX>
__device__ __forceinline__ void num_flux_vjp(const PhysParams &P, double nx, double ny,
                                             const double *qM, const double *qP, const double *nb,
                                             double *qMb, double *qPb) {
  constexpr int M = Vars<EQN>::M;
  double fx[M], fy[M];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    fx[c] = 0.5 * nx * nb[c];
    fy[c] = 0.5 * ny * nb[c];
  }
  flux_vjp<EQN>(P, qM, fx, fy, qMb);
  flux_vjp<EQN>(P, qP, fx, fy, qPb);
  double l;
  if constexpr (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4) {
    double dcM[M], dcP[M];
    const double cM = euler_speed_grad<EQN>(P, qM, dcM), cP = euler_speed_grad<EQN>(P, qP, dcP);
    l = cM > cP ? cM : cP;
    double jn = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) jn += (qM[c] - qP[c]) * nb[c];
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (cM > cP) qMb[c] += 0.5 * jn * dcM[c];
      else qPb[c] += 0.5 * jn * dcP[c];
    }
  } else {
    const double a = P.chi * P.c0, b = P.gph * P.c0;
    const double m1 = a > b ? a : b;
    l = m1 > P.c0 ? m1 : P.c0;
  }
#pragma unroll
  for (int c = 0; c < M; ++c) {
    qMb[c] += 0.5 * l * nb[c];
    qPb[c] -= 0.5 * l * nb[c];
  }
  if constexpr (FLUX == APL_FLUX_UPWIND && EQN == APL_EQN_MAXWELL_TM3) {
    const double sdot = -0.5 * l * (nx * nb[1] + ny * nb[2]);
    qMb[1] += sdot * nx; qMb[2] += sdot * ny;
    qPb[1] -= sdot * nx; qPb[2] -= sdot * ny;
  }
}

// qMb += (dqBC/dqM)^T qBCb
template <int EQN>
__device__ __forceinline__ void boundary_vjp(const BCTable &B, int idx, double nx, double ny,
                                             const double *qBCb, double *qMb) {
  constexpr int M = Vars<EQN>::M;
  const int kind = (idx >= 0 && idx < B.n) ? B.kind[idx] : APL_BC_COPY;
  if (kind == APL_BC_ISENTROPIC_VORTEX) return;   // exterior state independent of q-
  double t[M];
#pragma unroll
  for (int c = 0; c < M; ++c) t[c] = qBCb[c];
  if constexpr (EQN == APL_EQN_PHMAXWELL8) {
    if (kind == APL_BC_MAXWELL_TM) {
      t[2] = -qBCb[2];
    } else if (kind == APL_BC_PEC) {
      const double En = qBCb[0] * nx + qBCb[1] * ny, Et = -qBCb[0] * ny + qBCb[1] * nx;
      t[0] = En * nx + Et * ny;
      t[1] = En * ny - Et * nx;
      t[2] = -qBCb[2];
      const double Bn = qBCb[3] * nx + qBCb[4] * ny, Bt = -qBCb[3] * ny + qBCb[4] * nx;
      t[3] = -Bn * nx - Bt * ny;
      t[4] = -Bn * ny + Bt * nx;
      t[6] = -qBCb[6];
    }
  } else if constexpr (EQN == APL_EQN_MAXWELL_TM3) {
    if (kind == APL_BC_MAXWELL_TM) {
      t[0] = -qBCb[0];
    } else if (kind == APL_BC_PEC) {
      t[0] = -qBCb[0];
      const double Bn = qBCb[1] * nx + qBCb[2] * ny, Bt = -qBCb[1] * ny + qBCb[2] * nx;
      t[1] = -Bn * nx - Bt * ny;
      t[2] = -Bn * ny + Bt * nx;
    }
  }
#pragma unroll
  for (int c = 0; c < M; ++c) qMb[c] += t[c];
}

// y[i] + b x[i] over one batch of 16 elements of ES doubles, staged
// transposed as stage_transposed does (XT[w*16 + e]; zero past the batch).
template <int ES>
__device__ __forceinline__ void axpy_stage_transposed(double *XT, const double *y, const double *x,
                                                      double b, int cnt, int lane) {
  constexpr int NV2 = 16 * ES / 2, NIT = (NV2 + 63) / 64;
  if (cnt == 16) {
    dbl2 yv[NIT], xv[NIT];
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) {
        yv[j] = ((const dbl2 *)y)[idx];
        xv[j] = ((const dbl2 *)x)[idx];
      }
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) {
        const dbl2 v = yv[j] + b * xv[j];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
          XT[w * 16 + e] = v[h];
        }
      }
    }
  } else {
    for (int i = lane; i < 16 * ES; i += 64) {
      const int e = i / ES, w = i - e * ES;
      if (e < cnt) {
        XT[w * 16 + e] = y[i] + b * x[i];
      } else {
        XT[w * 16 + e] = 0.0;
      }
    }
  }
}

// dst[i] = src[i] (LDS to global) for cnt elements of ES doubles each.
template <int ES>
__device__ __forceinline__ void copy_batch(double *dst, const double *src, int cnt, int lane) {
  constexpr int NV2 = 16 * ES / 2, NIT = (NV2 + 63) / 64;
  if (cnt == 16) {
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) ((dbl2 *)dst)[idx] = dbl2{src[2 * idx], src[2 * idx + 1]};
    }
  } else {
    for (int i = lane; i < cnt * ES; i += 64) dst[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// kernel A: lh = lam / J (mode 1: lam = rbar + b qbar), Sb = -L^T lh
// ---------------------------------------------------------------------------

// Sb = -L^T (lam / J) for one batch: LT holds lam transposed (rows >= Np
// are zeroed here), sc is 1/J of lane el's element (0 past the batch end);
// SO (LDS, 16 FS doubles) stages the result, which is written to dst.
template <int N, int EQN>
__device__ __forceinline__ void face_weights(const AdjArgs &a, double *LT, double *SO, double sc,
                                             double *dst, int cnt, int lane) {
  using O = Order<N>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Ng = O::Ng, NKS1 = O::NKS1, NFT = AdjOrder<N>::NFT;
  constexpr int ES = Np * M, FS = 3 * Ng * M;
  typedef typename Acc<double>::type V4;
  const int el = lane & 15, lg = lane >> 4;
  for (int i = 16 * ES + lane; i < NKS1 * 4 * M * 16; i += 64) LT[i] = 0.0;
  wave_sync();
  double b[NKS1][M];
#pragma unroll
  for (int ks = 0; ks < NKS1; ++ks)
#pragma unroll
    for (int v = 0; v < M; ++v) b[ks][v] = LT[((4 * ks + lg) * M + v) * 16 + el] * sc;
  wave_sync();   // SO may overlay LT's space in the caller
#pragma unroll
  for (int ft = 0; ft < NFT; ++ft) {
    V4 acc[M];
#pragma unroll
    for (int v = 0; v < M; ++v) acc[v] = V4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < NKS1; ++ks) {
      const double av = a.Alt[(ft * NKS1 + ks) * 64 + lane];
#pragma unroll
      for (int v = 0; v < M; ++v) acc[v] = mma(av, b[ks][v], acc[v]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kf = 16 * ft + acc_row<double>(lg, r);
      if (kf < 3 * Ng) {
#pragma unroll
        for (int v = 0; v < M; ++v) SO[el * FS + kf * M + v] = acc[v][r];
      }
    }
  }
  wave_sync();
  copy_batch<FS>(dst, SO, cnt, lane);
}

template <int N, int EQN>
struct AdjFaceLds {
  static constexpr int M = Vars<EQN>::M;
  static constexpr int LT = Order<N>::NKS1 * 4 * M * 16;   // lh transposed, rows >= Np zero
  static constexpr int SO = 16 * 3 * Order<N>::Ng * M;     // Sb, [e][kf][v]
  static constexpr int doubles = LT + SO;
};

template <int N, int EQN>
struct AdjFaceGroup {
  static constexpr int W = group_waves(AdjFaceLds<N, EQN>::doubles, 8);
};

template <int N, int EQN>
__global__ void __launch_bounds__((64 * AdjFaceGroup<N, EQN>::W)) adj_face_kernel(AdjArgs a) {
  using O = Order<N>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Ng = O::Ng, NKS1 = O::NKS1, NFT = AdjOrder<N>::NFT;
  constexpr int ES = Np * M, FS = 3 * Ng * M;
  constexpr int W = AdjFaceGroup<N, EQN>::W;
  const WaveBatch wb = wave_batch<W>(a.K);
  if (!wb.active) return;
  extern __shared__ double lds_all[];
  double *lds = lds_all + wb.wave * AdjFaceLds<N, EQN>::doubles;
  double *LT = lds;
  double *SO = lds + AdjFaceLds<N, EQN>::LT;

  const int lane = wb.lane, el = lane & 15, lg = lane >> 4;
  const long e0 = wb.e0;
  const int cnt = wb.cnt;
  // mode 1: LT = (rbar + b qbar) transposed; mode 0: lam transposed.  1/J
  // is applied to the B operands (lane el's own element; 0 past the batch
  // end).
  const long ge = e0 + (el < cnt ? el : 0);
  const double sc = el < cnt ? a.geom[ge * G_WORDS + G_INVJ] : 0.0;
  if (a.mode == 1)
    axpy_stage_transposed<ES>(LT, a.lam + e0 * ES, a.qbar + e0 * ES, a.rk_b, cnt, lane);
  else
    stage_transposed<ES, 16>(LT, a.lam + e0 * ES, cnt, lane);
  face_weights<N, EQN>(a, LT, SO, sc, a.sbar + e0 * FS, cnt, lane);
}

// ---------------------------------------------------------------------------
// kernel B: the transposed RHS
// ---------------------------------------------------------------------------

// Write the batch's transposed RHS OUT[e*ES + w] (LDS): to `out` (mode 0),
// or the reverse LSERK4 update qbar += dt OUT, rbar = a lam (mode 1), with
// lam = rbar + b qbar as staged transposed in LT (LT[w*16 + e]).  With LN
// (LDS), mode 1 also stages the next reverse stage's lam' = rbar + b' qbar
// of the updated values, transposed the same way (zero past the batch end).
template <int ES, bool STAGE>
__device__ __forceinline__ unsigned adj_write_batch(const AdjArgs &a, const double *OUT,
                                                    const double *LT, double *LN, long e0, int cnt,
                                                    int lane) {
  constexpr int NV2 = 16 * ES / 2, NIT = (NV2 + 63) / 64;
  const long base = e0 * ES;
  double *rb = (double *)a.lam;
  unsigned ef = 0;
  if (cnt == 16) {
    dbl2 qv[NIT];
    if constexpr (STAGE) {
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = lane + 64 * j;
        if (idx < NV2) qv[j] = ((const dbl2 *)(a.qbar + base))[idx];
      }
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int idx = lane + 64 * j;
      if (idx < NV2) {
        const dbl2 v = dbl2{OUT[2 * idx], OUT[2 * idx + 1]};
        if (v[0] != v[0] || v[1] != v[1]) ef |= EF_NAN;
        if constexpr (!STAGE) {
          ((dbl2 *)(a.out + base))[idx] = v;
        } else {
          dbl2 lam;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
            lam[h] = LT[w * 16 + e];
          }
          const dbl2 qn = qv[j] + a.dt * v, rn = a.rk_a * lam;
          ((dbl2 *)(a.qbar + base))[idx] = qn;
          ((dbl2 *)(rb + base))[idx] = rn;
          if (LN) {
            const dbl2 ln = rn + a.rk_b_next * qn;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
              LN[w * 16 + e] = ln[h];
            }
          }
        }
      }
    }
  } else {
    for (int i = lane; i < 16 * ES; i += 64) {
      const int e = i / ES, w = i - e * ES;
      if (i >= cnt * ES) {
        if (LN) LN[w * 16 + e] = 0.0;
        continue;
      }
      const double val = OUT[i];
      if (val != val) ef |= EF_NAN;
      if constexpr (!STAGE) {
        a.out[base + i] = val;
      } else {
        const double qn = a.qbar[base + i] + a.dt * val, rn = a.rk_a * LT[w * 16 + e];
        a.qbar[base + i] = qn;
        rb[base + i] = rn;
        if (LN) LN[w * 16 + e] = rn + a.rk_b_next * qn;
      }
    }
  }
  return ef;
}

template <int N, int EQN>
struct AdjLds {
  static constexpr int M = Vars<EQN>::M;
  static constexpr int Np = Order<N>::Np, Nfp = Order<N>::Nfp, Ng = Order<N>::Ng;
  static constexpr int XT = Order<N>::NKS1 * 4 * M * 16;
  static constexpr int PT = 3 * Nfp * M * 16;
  static constexpr int I1D = 3 * Ng * Nfp;
  static constexpr int MAIN = 2 * XT + PT + I1D + 2 * Nfp;
  // epilogue: OUT (16 ES, inside XT's space), then SO (16 FS) over it; the
  // staged lam (LT) is read while the next lam is staged past both
  static constexpr int SO = 16 * 3 * Ng * M > 2 * XT ? 16 * 3 * Ng * M : 2 * XT;
  static constexpr int EPI_LN = SO;
  static constexpr int EPI = EPI_LN + XT;
  static constexpr int doubles = MAIN > EPI ? MAIN : EPI;
};

// Waves per SIMD requested from the register allocator: two where the
// kernel fits 256 VGPRs without spilling (measured with
// -Rpass-analysis=kernel-resource-usage), one otherwise.
template <int N, int EQN>
struct AdjWaves {
  static constexpr int value =
      (EQN == APL_EQN_MAXWELL_TM3 || (EQN == APL_EQN_EULER4 && N <= 4)) ? 2 : 1;
};

template <int N, int EQN>
struct AdjGroup {
  static constexpr int W = group_waves(AdjLds<N, EQN>::doubles, 4 * AdjWaves<N, EQN>::value);
};

// STAGE: the reverse-LSERK4 stage kernel (mode 1: lam = rbar + b qbar, the
// stage update and the next stage's face weights); otherwise out = J^T lam
// (mode 0).
template <int N, int EQN, int FLUX, bool STAGE>
__global__ void __launch_bounds__((64 * AdjGroup<N, EQN>::W), (AdjWaves<N, EQN>::value))
    adj_rhs_kernel(AdjArgs a) {
  using O = Order<N>;
  using L = AdjLds<N, EQN>;
  constexpr int M = Vars<EQN>::M;
  constexpr int Np = O::Np, Nc = O::Nc, Ng = O::Ng, Nfp = O::Nfp;
  constexpr int NKS1 = O::NKS1, NCH = O::NCH, NRT = O::NRT, NKF = O::NKF;
  constexpr int ES = Np * M, FS = 3 * Ng * M;
  typedef typename Acc<double>::type V4;

  constexpr int W = AdjGroup<N, EQN>::W;
  const WaveBatch wb = wave_batch<W>(a.K);
  if (!wb.active) return;
  extern __shared__ double lds_all[];
  double *lds = lds_all + wb.wave * L::doubles;
  double *XT = lds;                 // state, transposed
  double *LT = XT + L::XT;          // lam / J, transposed
  double *PT = LT + L::XT;          // neighbour face nodes (own order)
  double *I1 = PT + L::PT;
  int *FM = (int *)(I1 + L::I1D);

  const int lane = wb.lane, el = lane & 15, lg = lane >> 4;
  const long e0 = wb.e0;
  const int cnt = wb.cnt;
  const long ge = e0 + (el < cnt ? el : 0);
  const double *gk = a.geom + ge * G_WORDS;
  const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];

  // 1/J of lane el's element (0 past the batch end) scales the lam operands
  const double sc = el < cnt ? gk[G_INVJ] : 0.0;
  for (int i = lane; i < L::I1D; i += 64) I1[i] = a.I1D[i];
  for (int i = lane; i < 3 * Nfp; i += 64) FM[i] = a.fmask[i];
  stage_transposed<ES, 16>(XT, a.q + e0 * ES, cnt, lane);
  if constexpr (STAGE)
    axpy_stage_transposed<ES>(LT, a.lam + e0 * ES, a.qbar + e0 * ES, a.rk_b, cnt, lane);
  else
    stage_transposed<ES, 16>(LT, a.lam + e0 * ES, cnt, lane);
  for (int i = 16 * ES + lane; i < L::XT; i += 64) XT[i] = LT[i] = 0.0;
  wave_sync();
  for (int it = lane; it < 16 * 3 * Nfp; it += 64) {
    const int e = it & 15, k = it >> 4, f = k / Nfp, i = k - f * Nfp;
    const long gee = e0 + (e < cnt ? e : 0);
    const int code = a.fcode[gee * 3 + f];
#pragma unroll
    for (int v = 0; v < M; ++v) {
      double pv = 0.0;
      if (code >= 0)
        pv = a.q[((long)(code >> 2) * Np + FM[(code & 3) * Nfp + (Nfp - 1 - i)]) * M + v];
      PT[(k * M + v) * 16 + e] = pv;
    }
  }
  wave_sync();

  V4 acc[NRT][M];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int v = 0; v < M; ++v) acc[rt][v] = V4{0, 0, 0, 0};

  // volume
#pragma unroll 1
  for (int cc = 0; cc < NCH; ++cc) {
    V4 qc[M], fr[M], fs[M];
#pragma unroll
    for (int v = 0; v < M; ++v) qc[v] = fr[v] = fs[v] = V4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < NKS1; ++ks) {
      const double a1 = a.Avc[(cc * NKS1 + ks) * 64 + lane];
      const double ar = a.Aart[(cc * NKS1 + ks) * 64 + lane];
      const double as = a.Aast[(cc * NKS1 + ks) * 64 + lane];
#pragma unroll
      for (int v = 0; v < M; ++v) {
        const double bx = XT[((4 * ks + lg) * M + v) * 16 + el];
        const double bl = LT[((4 * ks + lg) * M + v) * 16 + el] * sc;
        qc[v] = mma(a1, bx, qc[v]);
        fr[v] = mma(ar, bl, fr[v]);
        fs[v] = mma(as, bl, fs[v]);
      }
    }
    double qcb[M][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = 16 * cc + acc_row<double>(lg, r);
#pragma unroll
      for (int v = 0; v < M; ++v) qcb[v][r] = 0.0;
      if (c < Nc) {
        double qv[M], Fb[M], Gb[M], qb[M];
#pragma unroll
        for (int v = 0; v < M; ++v) {
          qv[v] = qc[v][r];
          Fb[v] = jrx * fr[v][r] + jsx * fs[v][r];
          Gb[v] = jry * fr[v][r] + jsy * fs[v][r];
          qb[v] = 0.0;
        }
        flux_vjp<EQN>(a.ph, qv, Fb, Gb, qb);
#pragma unroll
        for (int v = 0; v < M; ++v) qcb[v][r] = qb[v];
      }
    }
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
      const double *frag = a.Avct + ((size_t)(rt * NCH + cc) * 64 + lane) * 4;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double av = frag[r];
#pragma unroll
        for (int v = 0; v < M; ++v) acc[rt][v] = mma(av, qcb[v][r], acc[rt][v]);
      }
    }
  }

  // surface
  const int code0 = a.fcode[ge * 3], code1 = a.fcode[ge * 3 + 1], code2 = a.fcode[ge * 3 + 2];
  unsigned ef = 0;
#pragma unroll 1
  for (int ks = 0; ks < NKF; ++ks) {
    const int kf = 4 * ks + lg;
    double qMb[M];
#pragma unroll
    for (int v = 0; v < M; ++v) qMb[v] = 0.0;
    if (kf < 3 * Ng) {
      const int f = kf / Ng, g = kf - f * Ng;
      const int code = f == 0 ? code0 : (f == 1 ? code1 : code2);
      const double nx = gk[G_NX + f], ny = gk[G_NY + f], sj = gk[G_SJ + f];
      double qM[M], qP[M], nb[M], qPb[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] = qP[v] = qPb[v] = 0.0;
      const double *Io = I1 + (f * Ng + g) * Nfp;
#pragma unroll
      for (int i = 0; i < Nfp; ++i) {
        const double *xo = XT + FM[f * Nfp + i] * M * 16 + el;
#pragma unroll
        for (int v = 0; v < M; ++v) qM[v] += Io[i] * xo[v * 16];
      }
      const double *so = a.sbar + ge * FS + kf * M;
#pragma unroll
      for (int v = 0; v < M; ++v) nb[v] = so[v] * sj;
      if (code >= 0) {
        const long nbr = code >> 2;
        const int nf = code & 3;
        const double *In = I1 + (nf * Ng + (Ng - 1 - g)) * Nfp;
#pragma unroll
        for (int i = 0; i < Nfp; ++i)
#pragma unroll
          for (int v = 0; v < M; ++v) qP[v] += In[Nfp - 1 - i] * PT[((f * Nfp + i) * M + v) * 16 + el];
        // The neighbour's flux at its Gauss point Ng-1-g has this trace qM as
        // its exterior state and the opposite normal; the fluxes the adjoint
        // covers (LF, TM upwind) are conservative, F*(-n, qP, qM) =
        // -F*(n, qM, qP), so its part is -(dF*/dqM)^T of its weights, and
        // one product with the combined weights gives both.
        const double snj = a.geom[nbr * G_WORDS + G_SJ + nf];
        const double *sn = a.sbar + nbr * FS + (nf * Ng + (Ng - 1 - g)) * M;
#pragma unroll
        for (int v = 0; v < M; ++v) nb[v] -= sn[v] * snj;
        num_flux_vjp<EQN, FLUX>(a.ph, nx, ny, qM, qP, nb, qMb, qPb);
      } else {
        const double *vk = a.vert + ge * 6;
        const int va = f, vb = (f + 1) % 3;
        const double z = a.I1D[3 * Ng * Nfp + g], s = 0.5 * (1.0 + z);
        const double xg = vk[2 * va] + s * (vk[2 * vb] - vk[2 * va]);
        const double yg = vk[2 * va + 1] + s * (vk[2 * vb + 1] - vk[2 * va + 1]);
        boundary_state<EQN>(a.bc, -code - 1, a.t, xg, yg, nx, ny, qM, qP);
        num_flux_vjp<EQN, FLUX>(a.ph, nx, ny, qM, qP, nb, qMb, qPb);
        boundary_vjp<EQN>(a.bc, -code - 1, nx, ny, qPb, qMb);
      }
    }
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
      const double av = a.Aift[(rt * NKF + ks) * 64 + lane];
#pragma unroll
      for (int v = 0; v < M; ++v) acc[rt][v] = mma(av, qMb[v], acc[rt][v]);
    }
  }

  // epilogue
  wave_sync();
  double *OUT = lds;
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * rt + acc_row<double>(lg, r);
      if (n < Np) {
#pragma unroll
        for (int v = 0; v < M; ++v) OUT[el * ES + n * M + v] = acc[rt][v][r];
      }
    }
  wave_sync();
  if constexpr (STAGE) {
    const bool next = a.sbar_out != nullptr;
    double *LN = lds + L::EPI_LN;
    ef |= adj_write_batch<ES, true>(a, OUT, LT, next ? LN : nullptr, e0, cnt, lane);
    // the next reverse stage's Sb of this batch (SO over OUT)
    if (next) face_weights<N, EQN>(a, LN, lds, sc, a.sbar_out + e0 * FS, cnt, lane);
  } else {
    ef |= adj_write_batch<ES, false>(a, OUT, LT, nullptr, e0, cnt, lane);
  }
  for (int off = 32; off > 0; off >>= 1) ef |= __shfl_xor(ef, off);
  if (lane == 0 && ef) atomicOr(a.err, ef);
}

// ---------------------------------------------------------------------------
// objective J = sum_k sum_c w_c J_k q_var(x_c)^2 [|x_c - x0| < R] and dJ/dq
// ---------------------------------------------------------------------------

struct ObjArgs {
  int K, var;
  const double *q;
  double *grad;         // [K][Np][M] or null
  double *part;         // [nblocks] partial sums of J
  const double *geom, *vert;
  const double *Vc;     // [Nc][Np] row-major
  const double *cub;    // [Nc][3] r, s, w
  double xc, yc, R;
};

template <int N, int M>
__global__ void __launch_bounds__(64) objective_kernel(ObjArgs a) {
  using O = Order<N>;
  constexpr int Np = O::Np, Nc = O::Nc;
  const long k = (long)blockIdx.x * 64 + threadIdx.x;
  double Jk = 0.0;
  if (k < a.K) {
    const double *q = a.q + k * Np * M;
    const double *vk = a.vert + k * 6;
    const double J = a.geom[k * G_WORDS + G_J];
    double gq[Np];
#pragma unroll
    for (int j = 0; j < Np; ++j) gq[j] = 0.0;
    for (int c = 0; c < Nc; ++c) {
      const double r = a.cub[3 * c], s = a.cub[3 * c + 1], w = a.cub[3 * c + 2];
      const double x = 0.5 * (-vk[0] * (r + s) + vk[2] * (1. + r) + vk[4] * (1. + s));
      const double y = 0.5 * (-vk[1] * (r + s) + vk[3] * (1. + r) + vk[5] * (1. + s));
      if ((x - a.xc) * (x - a.xc) + (y - a.yc) * (y - a.yc) < a.R * a.R) {
        double qc = 0.0;
#pragma unroll
        for (int j = 0; j < Np; ++j) qc += a.Vc[c * Np + j] * q[j * M + a.var];
        Jk += w * J * qc * qc;
#pragma unroll
        for (int j = 0; j < Np; ++j) gq[j] += 2.0 * w * J * qc * a.Vc[c * Np + j];
      }
    }
    if (a.grad) {
      double *g = a.grad + k * Np * M;
#pragma unroll
      for (int j = 0; j < Np; ++j)
#pragma unroll
        for (int v = 0; v < M; ++v) g[j * M + v] = (v == a.var) ? gq[j] : 0.0;
    }
  }
  // fixed-order reduction per block
  for (int off = 32; off > 0; off >>= 1) Jk += __shfl_xor(Jk, off);
  if (threadIdx.x == 0) a.part[blockIdx.x] = Jk;
}

// launch entries (apl_kernels.hip)
template <int N>
int launch_adjoint(int eqn, int flux, hipStream_t s, const AdjArgs &a, int nblocks);
template <int N>
int launch_objective(int m, hipStream_t s, const ObjArgs &a, int nblocks);

}  // namespace apl
