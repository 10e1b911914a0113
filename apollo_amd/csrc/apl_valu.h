// apl_valu.h -- the cubature-form RHS on the vector ALU for N <= 3 (gfx950).
//
// The same RHS as rhs_kernel (apl_device.h), Apollo's over-integrated
// cubature form (src/subsolvers/wxnodaldg2dmethod.cc:320-510,
// src/lib/wxcubature2d.cc:288-382):
//
//   rhs = (1/J) [ sum_c (A_r[:,c] Phi_r(c) + A_s[:,c] Phi_s(c))
//                 - sum_k L[:,k] F*(k) ]
//   Phi_r(c) = J (rx F + ry G)(V_c[c,:] Q),  Phi_s likewise,
//
// evaluated without matrix cores.  At N <= 3 an element has Np = 3, 6, 10
// nodes: a 16x16x4 MFMA tile pads the K = Np and M = Np extents of every
// product by 25-80%, and the products are too short to keep the matrix pipe
// fed.  Here one lane owns one element and every product is a chain of
// FMAs whose operator entry is wave-uniform, so it comes from a scalar load
// (constant address space) and enters the FMA as a scalar operand; the
// operators are never staged per lane.
//
// A workgroup of W = 4 or 8 waves owns 64 elements.  Their states are
// staged through LDS once (coalesced 16-byte loads; element stride ES + pad,
// odd, so the per-lane reads spread over the banks).  The waves split the
// points of the RHS: the first 3W/4 waves take the Gauss points of one face
// (W = 4: all Ng; W = 8: half of them each) and NVF cubature points, the
// last W/4 waves the remaining cubature points.  Each wave keeps a partial
// residual of its 64 elements in registers; the partials are summed through
// LDS pairwise in a fixed order (W = 4: (R0 + R2) + (R1 + R3)), scaled by 1/J
// in FP64, and written by the first four waves as four 16-element batches
// with the same epilogue (rhs, LSERK4 stage, SSP combination) as the MFMA
// kernels.  Eight waves split a group's work twice as finely; they serve
// meshes too small to give every CU two four-wave groups.
#pragma once

#include "apl_device.h"

namespace apl {

template <int N, int EQN, int W>
struct ValuCfg {
  using O = Order<N>;
  static constexpr int M = Vars<EQN>::M;
  static constexpr int Np = O::Np, Nc = O::Nc, Ng = O::Ng, Nfp = O::Nfp;
  static constexpr int ES = Np * M;     // doubles per element
  static constexpr int ESP = ES | 1;    // LDS element stride (odd)
  static constexpr int VW = 3 * Np;     // point row: V_c[c,:], A_r[:,c], A_s[:,c]
  static constexpr int FW = 3 * W / 4;  // face waves
  static constexpr int NFG = Ng * 4 / W;   // Gauss points per face wave (Ng is even)
  // VALU instructions per point (estimates): a cubature point interpolates
  // (Np M), updates the residual (2 Np M) and evaluates the flux (~14 M);
  // a face point forms two traces (2 Nfp M), the numerical flux (~30 M) and
  // the lift (Np M).  Each face wave takes as many cubature points as
  // balance the waves; the others share the rest.
  static constexpr int CV = 3 * Np * M + 14 * M;
  static constexpr int CF = 2 * Nfp * M + 30 * M + Np * M;
  static constexpr int TOT = Nc * CV + 3 * Ng * CF;
  static constexpr int NVF0 = (TOT / W - NFG * CF + CV / 2) / CV;
  static constexpr int NVF = NVF0 < 0 ? 0 : (FW * NVF0 > Nc ? Nc / FW : NVF0);
  static constexpr int NVR = Nc - FW * NVF;   // points of the volume-only waves
  // cubature points [cb, ce) of wave w
  static __device__ __forceinline__ int cb(int w) {
    return w < FW ? w * NVF : FW * NVF + (w - FW) * NVR / (W - FW);
  }
  static __device__ __forceinline__ int ce(int w) {
    return w < FW ? (w + 1) * NVF : FW * NVF + (w - FW + 1) * NVR / (W - FW);
  }
};

// LDS of one workgroup, in doubles: the staged states [64][ESP] (element
// type S), reused for W/2 partial residuals [W/2][64][ESP] and then for the
// FP64 result [64][ES]; then I1D[3][Ng][Nfp], Fmask[3][Nfp], and the
// per-wave speed maxima and flags.
template <int N, int EQN, typename S, int W>
struct ValuLds {
  using C = ValuCfg<N, EQN, W>;
  static constexpr int RED_D = ((W / 2) * 64 * C::ESP * (int)sizeof(S) + 15) / 16 * 2;
  static constexpr int OUT_D = 64 * C::ES;
  static constexpr int BASE = RED_D > OUT_D ? RED_D : OUT_D;
  static constexpr int I1D = 3 * C::Ng * C::Nfp;
  static constexpr int FM = (3 * C::Nfp + 1) / 2;
  static constexpr int doubles = BASE + I1D + FM + W + (W + 1) / 2;
};

// Meshes with fewer 64-element groups than this run eight-wave groups
// where their LDS fits a CU.
constexpr int VALU_SMALL_GROUPS = 512;

template <int N, int EQN, typename S>
struct ValuWaves {
  static constexpr int value = sizeof(S) == 8 ? 2 : 3;
};

template <typename S>
using cptr_t = const __attribute__((address_space(4))) S *;

template <int N, int EQN, int FLUX, typename S, int W>
__global__ void __launch_bounds__(64 * W, (ValuWaves<N, EQN, S>::value))
    rhs_valu_kernel(RhsArgs a) {
  using C = ValuCfg<N, EQN, W>;
  using L = ValuLds<N, EQN, S, W>;
  constexpr int M = C::M, Np = C::Np, Ng = C::Ng, Nfp = C::Nfp;
  constexpr int ES = C::ES, ESP = C::ESP, VW = C::VW, FW = C::FW, NFG = C::NFG;
  constexpr int NT = 64 * W;   // threads
  // FP32 face traces and flux where the FP32 MFMA kernel uses them
  constexpr bool F32 = sizeof(S) == 4 && FLUX == APL_FLUX_LF && N <= 5 &&
                       (EQN == APL_EQN_EULER5 || EQN == APL_EQN_EULER4);
  typedef typename std::conditional<F32, float, double>::type TR;   // trace type

  // operator tables, read with wave-uniform indices (scalar loads)
  const cptr_t<S> OPV = (cptr_t<S>)a.A1;          // [Nc][VW]
  const cptr_t<S> OPL = (cptr_t<S>)a.A3;          // [3 Ng][Np]: -L[:, k]
  const cptr_t<double> I1c = (cptr_t<double>)a.I1D;
  const cptr_t<int> FMc = (cptr_t<int>)a.fmask;

  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = (int)threadIdx.x;
  const int grp = xcd_batch(blockIdx.x, gridDim.x);
  const long e0 = (long)grp * 64;
  const int cnt = (int)min((long)64, (long)a.K - e0);
  const long ge = e0 + (lane < cnt ? lane : 0);

  extern __shared__ double lds_all[];
  S *QS = (S *)lds_all;                 // [64][ESP] staged states
  S *RED = (S *)lds_all;                // [W/2][64][ESP] partial residuals
  double *OUT = lds_all;                // [64][ES] result
  double *I1 = lds_all + L::BASE;       // [3][Ng][Nfp]
  int *FM = (int *)(I1 + L::I1D);       // [3][Nfp]
  double *SPD = I1 + L::I1D + L::FM;    // [W] speed maxima
  unsigned *EFW = (unsigned *)(SPD + W);   // [W] flags

  // 1. tables and the group's states (missing elements repeat element 0)
  for (int i = tid; i < L::I1D; i += NT) I1[i] = a.I1D[i];
  for (int i = tid; i < 3 * Nfp; i += NT) FM[i] = a.fmask[i];
  {
    const double *qb = a.q + e0 * ES;
    if (cnt == 64) {
      constexpr int NV2 = 64 * ES / 2, NIT = (NV2 + NT - 1) / NT;
      dbl2 buf[NIT];
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = tid + NT * j;
        if (idx < NV2) buf[j] = ((const dbl2 *)qb)[idx];
      }
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = tid + NT * j;
        if (idx < NV2) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
            QS[e * ESP + w] = S(buf[j][h]);
          }
        }
      }
    } else {
      for (int i = tid; i < 64 * ES; i += NT) {
        const int e = i / ES, w = i - e * ES;
        QS[e * ESP + w] = S(e < cnt ? qb[i] : qb[w]);
      }
    }
  }
  __syncthreads();

  const double *gk = a.geom + ge * G_WORDS;
  const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];
  const double invJ = gk[G_INVJ];
  const S jrx_s = S(jrx), jry_s = S(jry), jsx_s = S(jsx), jsy_s = S(jsy);

  S R[Np][M];
#pragma unroll
  for (int n = 0; n < Np; ++n)
#pragma unroll
    for (int v = 0; v < M; ++v) R[n][v] = S(0);
  double lmax = 0.0;
  unsigned ef = 0;

  // 2. surface: face wave w takes Gauss points [g0, g0 + NFG) of face f
  if (wave < FW) {
    const int f = W == 4 ? wave : wave >> 1;
    const int g0 = W == 4 ? 0 : (wave & 1) * NFG;
    const int code = a.fcode[ge * 3 + f];
    const double nx = gk[G_NX + f], ny = gk[G_NY + f], sj = gk[G_SJ + f];
    const double tnow = (a.tptr ? *a.tptr : a.t) + a.tofs;
    // own face nodes (Fmask[f][i]) and the neighbour's face nodes in its
    // own order (Fmask[nf][j]); its Gauss point Ng-1-g meets our point g
    TR PM[Nfp][M], PP[Nfp][M];
#pragma unroll
    for (int i = 0; i < Nfp; ++i) {
      const int fm = FMc[f * Nfp + i];
#pragma unroll
      for (int v = 0; v < M; ++v) PM[i][v] = TR(QS[lane * ESP + fm * M + v]);
    }
    const int nf = code >= 0 ? (code & 3) : 0;
    if (code >= 0) {
      const long nb = code >> 2;
#pragma unroll
      for (int j = 0; j < Nfp; ++j) {
        const double *src = a.q + (nb * Np + FM[nf * Nfp + j]) * M;
#pragma unroll
        for (int v = 0; v < M; ++v) PP[j][v] = TR(src[v]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < Nfp; ++j)
#pragma unroll
        for (int v = 0; v < M; ++v) PP[j][v] = TR(0);
    }
#pragma unroll 1
    for (int g = g0; g < g0 + NFG; ++g) {
      TR qM[M], qP[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qM[v] = qP[v] = TR(0);
#pragma unroll
      for (int i = 0; i < Nfp; ++i) {
        const TR w = TR(I1c[(f * Ng + g) * Nfp + i]);
#pragma unroll
        for (int v = 0; v < M; ++v) qM[v] += w * PM[i][v];
      }
      S fsT[M];
      if constexpr (F32) {
        if (code >= 0) {
          const double *In = I1 + (nf * Ng + (Ng - 1 - g)) * Nfp;
#pragma unroll
          for (int j = 0; j < Nfp; ++j) {
            const float w = float(In[j]);
#pragma unroll
            for (int v = 0; v < M; ++v) qP[v] += w * PP[j][v];
          }
        } else {
          const double *vk = a.vert + ge * 6;
          const int va = f, vb = f == 2 ? 0 : f + 1;
          const double sz = 0.5 * (1.0 + I1c[3 * Ng * Nfp + g]);
          const double xg = vk[2 * va] + sz * (vk[2 * vb] - vk[2 * va]);
          const double yg = vk[2 * va + 1] + sz * (vk[2 * vb + 1] - vk[2 * va + 1]);
          double dM[M], dP[M];
#pragma unroll
          for (int v = 0; v < M; ++v) dM[v] = qM[v];
          boundary_state<EQN>(a.bc, -code - 1, tnow, xg, yg, nx, ny, dM, dP);
#pragma unroll
          for (int v = 0; v < M; ++v) qP[v] = float(dP[v]);
        }
        float fs[M], lam;
        ef |= euler_lf_f32<EQN>(a.ph, float(nx), float(ny), qM, qP, fs, &lam);
        lmax = double(lam) > lmax ? double(lam) : lmax;
        const float sjf = float(sj);
#pragma unroll
        for (int v = 0; v < M; ++v) fsT[v] = S(fs[v] * sjf);
      } else {
        if (code >= 0) {
          const double *In = I1 + (nf * Ng + (Ng - 1 - g)) * Nfp;
#pragma unroll
          for (int j = 0; j < Nfp; ++j) {
            const double w = In[j];
#pragma unroll
            for (int v = 0; v < M; ++v) qP[v] += w * PP[j][v];
          }
        } else {
          const double *vk = a.vert + ge * 6;
          const int va = f, vb = f == 2 ? 0 : f + 1;
          const double sz = 0.5 * (1.0 + I1c[3 * Ng * Nfp + g]);
          const double xg = vk[2 * va] + sz * (vk[2 * vb] - vk[2 * va]);
          const double yg = vk[2 * va + 1] + sz * (vk[2 * vb + 1] - vk[2 * va + 1]);
          boundary_state<EQN>(a.bc, -code - 1, tnow, xg, yg, nx, ny, qM, qP);
        }
        double fs[M], lam;
        ef |= num_flux<EQN, FLUX>(a.ph, nx, ny, qM, qP, fs, &lam);
        lmax = lam > lmax ? lam : lmax;
#pragma unroll
        for (int v = 0; v < M; ++v) fsT[v] = S(fs[v] * sj);
      }
      const int k = f * Ng + g;
#pragma unroll
      for (int n = 0; n < Np; ++n) {
        const S l = OPL[k * Np + n];
#pragma unroll
        for (int v = 0; v < M; ++v) R[n][v] += l * fsT[v];
      }
    }
  }

  // 3. volume: cubature points [cb, ce) of this wave
  {
    S Q[Np][M];
#pragma unroll
    for (int n = 0; n < Np; ++n)
#pragma unroll
      for (int v = 0; v < M; ++v) Q[n][v] = QS[lane * ESP + n * M + v];
    const int cb = C::cb(wave), ce = C::ce(wave);
#pragma unroll 1
    for (int c = cb; c < ce; ++c) {
      const cptr_t<S> vr = OPV + c * VW;
      S qc[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qc[v] = S(0);
#pragma unroll
      for (int j = 0; j < Np; ++j) {
        const S w = vr[j];
#pragma unroll
        for (int v = 0; v < M; ++v) qc[v] += w * Q[j][v];
      }
      S F[M], G[M], phr[M], phs[M];
      ef |= phys_flux<EQN, S>(a.ph, qc, F, G);
      if constexpr (sizeof(S) == 4) {
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
#pragma unroll
      for (int n = 0; n < Np; ++n) {
        const S ar = vr[Np + n], as = vr[2 * Np + n];
        // two separate multiply-adds per entry (one expression would cost a
        // multiply, a multiply-add and an add)
#pragma unroll
        for (int v = 0; v < M; ++v) {
          R[n][v] += ar * phr[v];
          R[n][v] += as * phs[v];
        }
      }
    }
  }

  // 4. sum the partial residuals pairwise: wave w + h adds into wave w
  // (h = W/2, ..., 1), so W = 4 sums (R0 + R2) + (R1 + R3)
  __syncthreads();   // every wave has read its state from QS
#pragma unroll
  for (int h = W / 2; h >= 1; h >>= 1) {
    if (wave >= h && wave < 2 * h) {
      S *dst = RED + (wave - h) * 64 * ESP + lane * ESP;
#pragma unroll
      for (int n = 0; n < Np; ++n)
#pragma unroll
        for (int v = 0; v < M; ++v) dst[n * M + v] = R[n][v];
    }
    __syncthreads();
    if (wave < h) {
      const S *src = RED + wave * 64 * ESP + lane * ESP;
#pragma unroll
      for (int n = 0; n < Np; ++n)
#pragma unroll
        for (int v = 0; v < M; ++v) R[n][v] += src[n * M + v];
    }
    if (h > 1) __syncthreads();
  }
  if (wave == 0) {
    // wave 0's reads of RED are all issued before its first write of OUT,
    // which overlaps them; one wave's LDS operations complete in order
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int n = 0; n < Np; ++n)
#pragma unroll
      for (int v = 0; v < M; ++v) OUT[lane * ES + n * M + v] = double(R[n][v]) * invJ;
  }
  __syncthreads();

  // 5. wave w < 4 writes elements [16 w, 16 w + 16) of the group
  const long eb = e0 + 16 * wave;
  if (wave < 4 && eb < a.K) {
    const int cw = (int)min((long)16, (long)a.K - eb);
    ef |= write_batch<ES>(a, OUT + 16 * wave * ES, eb, cw, lane);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off);
    lmax = o > lmax ? o : lmax;
    ef |= __shfl_xor(ef, off);
  }
  if (lane == 0) {
    SPD[wave] = lmax;
    EFW[wave] = ef;
  }
  __syncthreads();
  if (lane == 0 && wave < 4 && eb < a.K) {
    double m = SPD[0];
    unsigned e = EFW[0];
#pragma unroll
    for (int w = 1; w < W; ++w) {
      m = SPD[w] > m ? SPD[w] : m;
      e |= EFW[w];
    }
    if (a.speed_part) a.speed_part[grp * 4 + wave] = m;
    if (wave == 0 && e) atomicOr(a.err, e);
  }
}


// ---------------------------------------------------------------------------
// the nodal (linear-flux) form on the vector ALU
// ---------------------------------------------------------------------------
//
// The nodal form of rhs_linear_kernel (apl_device.h), one lane per element,
// 64 elements per workgroup of W waves:
//   surface  S(kf) = F*(q-, q+) L/2 at the 3 Nfp face nodes: wave w takes
//            face nodes [w NFW, (w+1) NFW) and stages S in LDS;
//   lift     R(n) = -Lt[n, :] S for its rows n in [w RPW, (w+1) RPW);
//   volume   R(n) += jrx F(P_r) + jry G(P_r) + jsx F(P_s) + jsy G(P_s),
//            P_r = Dt_r[n, :] Q, P_s = Dt_s[n, :] Q, for the same rows;
// then R / J through LDS to the shared epilogue.  Operator entries are
// wave-uniform (scalar loads); the state is read from its LDS copy once per
// node for all of a wave's rows.  Each wave holds only its rows' residual,
// so a wave needs few registers and many stay resident: the MFMA nodal
// kernel's batch of 16 elements per wave is latency-bound on memory when
// registers allow two FP64 waves per SIMD.

template <int N, int EQN>
struct ValuLinCfg {
  using O = Order<N>;
  static constexpr int M = Vars<EQN>::M;
  static constexpr int Np = O::Np, Nfp = O::Nfp;
  static constexpr int ES = Np * M, ESP = ES | 1;
  static constexpr int NF = 3 * Nfp;                       // face nodes
  static constexpr int W = Np >= 16 ? 8 : (Np >= 6 ? 4 : 2);   // waves per group
  static constexpr int RPW = (Np + W - 1) / W;              // rows per wave
  static constexpr int NFW = (NF + W - 1) / W;              // face nodes per wave
};

// LDS, in doubles: the staged states [64][ESP] (element type S), then the
// FP64 result [64][ES] over them; the face fluxes [NF][M][64] (S); speed
// maxima and flags.
template <int N, int EQN, typename S>
struct ValuLinLds {
  using C = ValuLinCfg<N, EQN>;
  static constexpr int QS_D = (64 * C::ESP * (int)sizeof(S) + 15) / 16 * 2;
  static constexpr int OUT_D = 64 * C::ES;
  static constexpr int BASE = QS_D > OUT_D ? QS_D : OUT_D;
  static constexpr int SF_D = (C::NF * C::M * 64 * (int)sizeof(S) + 15) / 16 * 2;
  static constexpr int doubles = BASE + SF_D + C::W + (C::W + 1) / 2;
};

// Waves per SIMD the register allocator is asked to allow: a wave holds
// its rows' residual and their two nodal derivatives (3 RPW M values).
template <int N, int EQN, typename S>
struct ValuLinWaves {
  static constexpr int value = sizeof(S) == 8 ? (N <= 5 ? 4 : (N == 6 ? 3 : 2))
                                              : (N <= 5 ? 5 : (N == 6 ? 4 : 3));
};

template <int N, int EQN, int FLUX, typename S>
__global__ void __launch_bounds__((64 * ValuLinCfg<N, EQN>::W), (ValuLinWaves<N, EQN, S>::value))
    rhs_valu_linear_kernel(RhsArgs a) {
  using C = ValuLinCfg<N, EQN>;
  using L = ValuLinLds<N, EQN, S>;
  constexpr int M = C::M, Np = C::Np, Nfp = C::Nfp, ES = C::ES, ESP = C::ESP, NF = C::NF;
  constexpr int W = C::W, RPW = C::RPW, NFW = C::NFW, NT = 64 * W;

  const cptr_t<S> DT = (cptr_t<S>)a.A1;        // [2][Np][Np]: Dt_r, Dt_s
  const cptr_t<S> LT = (cptr_t<S>)a.A3;        // [Np][NF]: -Lt
  const cptr_t<int> FMc = (cptr_t<int>)a.fmask;

  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = (int)threadIdx.x;
  const int grp = xcd_batch(blockIdx.x, gridDim.x);
  const long e0 = (long)grp * 64;
  const int cnt = (int)min((long)64, (long)a.K - e0);
  const long ge = e0 + (lane < cnt ? lane : 0);

  extern __shared__ double lds_all[];
  S *QS = (S *)lds_all;                    // [64][ESP] staged states
  double *OUT = lds_all;                   // [64][ES] result
  S *SF = (S *)(lds_all + L::BASE);        // [NF][M][64] face fluxes
  double *SPD = lds_all + L::BASE + L::SF_D;   // [W]
  unsigned *EFW = (unsigned *)(SPD + W);       // [W]

  // 1. the group's states (missing elements repeat element 0)
  {
    const double *qb = a.q + e0 * ES;
    if (cnt == 64) {
      constexpr int NV2 = 64 * ES / 2, NIT = (NV2 + NT - 1) / NT;
      dbl2 buf[NIT];
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = tid + NT * j;
        if (idx < NV2) buf[j] = ((const dbl2 *)qb)[idx];
      }
#pragma unroll
      for (int j = 0; j < NIT; ++j) {
        const int idx = tid + NT * j;
        if (idx < NV2) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int i = 2 * idx + h, e = i / ES, w = i - e * ES;
            QS[e * ESP + w] = S(buf[j][h]);
          }
        }
      }
    } else {
      for (int i = tid; i < 64 * ES; i += NT) {
        const int e = i / ES, w = i - e * ES;
        QS[e * ESP + w] = S(e < cnt ? qb[i] : qb[w]);
      }
    }
  }
  __syncthreads();

  const double *gk = a.geom + ge * G_WORDS;
  double lmax = 0.0;
  unsigned ef = 0;

  // 2. surface: face nodes [wave NFW, wave NFW + NFW) of this lane's element
#pragma unroll 1
  for (int kf = wave * NFW; kf < (wave + 1) * NFW && kf < NF; ++kf) {
    const int f = kf / Nfp, i = kf - f * Nfp;
    const int code = a.fcode[ge * 3 + f];
    const double nx = gk[G_NX + f], ny = gk[G_NY + f], sj = gk[G_SJ + f];
    double qM[M], qP[M], fs[M];
    const int fm = FMc[f * Nfp + i];
#pragma unroll
    for (int v = 0; v < M; ++v) qM[v] = double(QS[lane * ESP + fm * M + v]);
    if (code >= 0) {
      // own node i <-> the neighbour's node Fmask[nf][Nfp-1-i]
      const int nf = code & 3;
      const double *src = a.q + ((long)(code >> 2) * Np + a.fmask[nf * Nfp + (Nfp - 1 - i)]) * M;
#pragma unroll
      for (int v = 0; v < M; ++v) qP[v] = src[v];
    } else {
      boundary_state<EQN>(a.bc, -code - 1, 0.0, 0.0, 0.0, nx, ny, qM, qP);
    }
    double lam;
    ef |= num_flux<EQN, FLUX>(a.ph, nx, ny, qM, qP, fs, &lam);
    lmax = lam > lmax ? lam : lmax;
#pragma unroll
    for (int v = 0; v < M; ++v) SF[(kf * M + v) * 64 + lane] = S(fs[v] * sj);
  }
  __syncthreads();

  // 3. lift and volume for rows [n0, n0 + RPW)
  const int n0 = wave * RPW;
  S R[RPW][M];
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int v = 0; v < M; ++v) R[r][v] = S(0);
#pragma unroll 2
  for (int kf = 0; kf < NF; ++kf) {
    S sv[M];
#pragma unroll
    for (int v = 0; v < M; ++v) sv[v] = SF[(kf * M + v) * 64 + lane];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      if (n0 + r < Np) {
        const S l = LT[(n0 + r) * NF + kf];
#pragma unroll
        for (int v = 0; v < M; ++v) R[r][v] += l * sv[v];
      }
    }
  }
  {
    S pr[RPW][M], ps[RPW][M];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int v = 0; v < M; ++v) pr[r][v] = ps[r][v] = S(0);
#pragma unroll 2
    for (int k = 0; k < Np; ++k) {
      S qk[M];
#pragma unroll
      for (int v = 0; v < M; ++v) qk[v] = QS[lane * ESP + k * M + v];
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        if (n0 + r < Np) {
          const S dr = DT[(n0 + r) * Np + k], ds = DT[Np * Np + (n0 + r) * Np + k];
#pragma unroll
          for (int v = 0; v < M; ++v) {
            pr[r][v] += dr * qk[v];
            ps[r][v] += ds * qk[v];
          }
        }
      }
    }
    const double jrx = gk[G_JRX], jry = gk[G_JRY], jsx = gk[G_JSX], jsy = gk[G_JSY];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      S F[M], G[M];
      phys_flux<EQN, S>(a.ph, pr[r], F, G);
#pragma unroll
      for (int v = 0; v < M; ++v) {
        R[r][v] += S(jrx) * F[v];
        R[r][v] += S(jry) * G[v];
      }
      phys_flux<EQN, S>(a.ph, ps[r], F, G);
#pragma unroll
      for (int v = 0; v < M; ++v) {
        R[r][v] += S(jsx) * F[v];
        R[r][v] += S(jsy) * G[v];
      }
    }
  }

  // 4. R / J into LDS over the staged states, then the shared epilogue: wave
  // w writes the 16-element batches w, w + W, ...
  __syncthreads();   // every wave has read the states
  const double invJ = gk[G_INVJ];
#pragma unroll
  for (int r = 0; r < RPW; ++r)
    if (n0 + r < Np) {
#pragma unroll
      for (int v = 0; v < M; ++v) OUT[lane * ES + (n0 + r) * M + v] = double(R[r][v]) * invJ;
    }
  __syncthreads();
#pragma unroll
  for (int sb = wave; sb < 4; sb += W) {
    const long eb = e0 + 16 * sb;
    if (eb < a.K) {
      const int cw = (int)min((long)16, (long)a.K - eb);
      ef |= write_batch<ES>(a, OUT + 16 * sb * ES, eb, cw, lane);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off);
    lmax = o > lmax ? o : lmax;
    ef |= __shfl_xor(ef, off);
  }
  if (lane == 0) {
    SPD[wave] = lmax;
    EFW[wave] = ef;
  }
  __syncthreads();
  if (lane == 0) {
    double m = SPD[0];
    unsigned e = EFW[0];
#pragma unroll
    for (int w = 1; w < W; ++w) {
      m = SPD[w] > m ? SPD[w] : m;
      e |= EFW[w];
    }
#pragma unroll
    for (int sb = wave; sb < 4; sb += W)
      if (e0 + 16 * sb < a.K && a.speed_part) a.speed_part[grp * 4 + sb] = m;
    if (wave == 0 && e) atomicOr(a.err, e);
  }
}

}  // namespace apl
