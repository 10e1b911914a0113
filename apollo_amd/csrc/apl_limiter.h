// apl_limiter.h -- Apollo's Hesthaven-Warburton Euler limiter
// (src/hyperapps/euler/wxhesthavenwarburtoneulerlimiter.cc:61-484) on the
// GPU, as three element-parallel passes over HBM-resident state:
//
//   lim_average_kernel      cell averages of the conserved variables
//                           (computeConservedAndPrimitiveAVEVariables,
//                           :487-522), one FP64 5-vector per element;
//   lim_gradient_kernel     the four-element patch of each element: patch
//                           centres and areas, ghost averages through the
//                           boundary states, face averages at the six face
//                           ends, the three edge gradients and their area
//                           average (:97-314);
//   lim_reconstruct_kernel  the limited gradient from the neighbours' cell
//                           gradients (own edge gradients on boundary faces)
//                           and the reconstruction with its density and
//                           pressure floors (:317-467).
//
// Each pass is bound by HBM traffic.  The average and reconstruction
// passes move the state of 32 consecutive elements per wave between HBM and
// LDS with coalesced 16-byte accesses, one lane per element on the LDS copy;
// the gradient pass, a gather of a few hundred bytes per element, runs one
// thread per element.  The
// reference's reading order and quirks are kept (see the oracle's
// orc_limit): AVE[0..2] with the vertex nodes for the patch centres, the
// boundary-face node states read from the last neighbour visited, the face
// ends indexed by element-node id into the face-ordered averages, and the
// (yv3 - yv2), (xv3 - xv2) factors of the third edge gradient.  Orders
// N <= 4 only: for N > 4 the reference reads outside its face array.
#pragma once

#include "apl_device.h"

namespace apl {

struct LimArgs {
  int K;
  const double *qin;       // [K][Np][M]
  double *qout;            // [K][Np][M], may equal qin
  double *avg;             // [K][8]   conserved cell averages, then the
                           //          patch centre xc, yc and area share A0
  double *grad;            // [K][8][5] C0x,C0y,E1x,E1y,E2x,E2y,E3x,E3y; the
                           //          edge rows only for elements with a
                           //          boundary face (the only readers)
  const double *geom;      // [K][G_WORDS]
  const double *vert;      // [K][6]
  const int *fcode;        // [K][3]
  const int *fmask;        // [3][Nfp]
  const double *tab;       // ave[Np], r[Np], s[Np], drop[Np][Np]
  double gamma, min_dens, min_pres;
  unsigned *err;
  BCTable bc;
};

// Limiter primitives (wxhesthavenwarburtoneulerlimiter.cc:535-543): no floors.
__device__ __forceinline__ void lim_prim(double g, const double *c, double *p) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = c[3] / c[0];
  p[4] = (g - 1.) * (c[4] - 0.5 * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) / c[0]);
}

// The state of LE consecutive elements moves between HBM and LDS with
// coalesced 16-byte accesses by all 64 lanes of a wave; lane e < LE then
// works on element e from LDS.
constexpr int LE = 32;

template <int ES>
__device__ __forceinline__ void lim_stage_in(const double *src, double *lds, int cnt, int lane) {
  const int n = cnt * ES;
  if (cnt == LE) {
    // every load of the wave in flight before the first LDS write
    constexpr int NV2 = LE * ES / 2, NIT = (NV2 + 63) / 64;
    dbl2 buf[NIT];
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int i = lane + 64 * j;
      if (i < NV2) buf[j] = ((const dbl2 *)src)[i];
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int i = lane + 64 * j;
      if (i < NV2) ((dbl2 *)lds)[i] = buf[j];
    }
  } else {
    for (int i = lane; i < n; i += 64) lds[i] = src[i];
  }
}

template <int ES>
__device__ __forceinline__ void lim_stage_out(const double *lds, double *dst, int cnt, int lane) {
  const int n = cnt * ES;
  if (cnt == LE) {
    constexpr int NV2 = LE * ES / 2;
#pragma unroll 4
    for (int i = lane; i < NV2; i += 64) ((dbl2 *)dst)[i] = ((const dbl2 *)lds)[i];
  } else {
    for (int i = lane; i < n; i += 64) dst[i] = lds[i];
  }
}

template <int N, int EQN>
__global__ void __launch_bounds__(64) lim_average_kernel(LimArgs a) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, ES = Np * M;
  __shared__ __align__(16) double st[LE * ES];
  const int lane = threadIdx.x;
  const long e0 = (long)blockIdx.x * LE;
  const int cnt = (int)((a.K - e0) < LE ? (a.K - e0) : LE);
  lim_stage_in<ES>(a.qin + e0 * ES, st, cnt, lane);
  __syncthreads();
  if (lane >= cnt) return;
  const double *ave = a.tab;
  const double *q = st + lane * ES;
  double c5[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) c5[c] = 0.0;
#pragma unroll
  for (int n = 0; n < Np; ++n) {
    double q5[5];
    load_euler5<EQN>(q + n * M, q5);
#pragma unroll
    for (int c = 0; c < 5; ++c) c5[c] += ave[n] * q5[c];
  }
  // the element's patch centre and area share (:121-126), which its
  // neighbours' patches reuse (:136-144): same arithmetic either way
  const long e = e0 + lane;
  const double *vk = a.vert + e * 6;
  // J from the vertices as the geometry computes it (wxnodaldggeometry2d.cc:412-426)
  const double xr = (vk[2] - vk[0]) / 2, xs = (vk[4] - vk[0]) / 2;
  const double yr = (vk[3] - vk[1]) / 2, ys = (vk[5] - vk[1]) / 2;
  const double J = -xs * yr + xr * ys;
  double xc = 0.0, yc = 0.0, A0 = 0.0;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    xc += ave[kk] * vk[2 * kk];
    yc += ave[kk] * vk[2 * kk + 1];
    A0 += 2. / 3. * ave[kk] * J;
  }
  double *o = a.avg + e * 8;
#pragma unroll
  for (int c = 0; c < 5; ++c) o[c] = c5[c];
  o[5] = xc;
  o[6] = yc;
  o[7] = A0;
}

template <int N, int EQN>
__global__ void __launch_bounds__(256) lim_gradient_kernel(LimArgs a) {
  constexpr int Np = Order<N>::Np, Nfp = Order<N>::Nfp, M = Vars<EQN>::M;
  static_assert(N <= 4, "the limiter's face-end reads stay inside 3 Nfp entries for N <= 4");
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.K) return;
  const int *F = a.fmask;
  const double *gk = a.geom + (long)e * G_WORDS;
  const double *vk = a.vert + (long)e * 6;
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];

  // patch centres and areas (:121-147) and patch averages (:149-233);
  // src[f] is the element whose nodes the reference reads for face f's
  // boundary states: the last neighbour visited before it, or the element
  // itself
  double xc[4], yc[4], A0[4], pr[4][5], own[5];
  {
    const double *o = a.avg + (long)e * 8;
#pragma unroll
    for (int c = 0; c < 5; ++c) own[c] = o[c];
    xc[0] = o[5];
    yc[0] = o[6];
    A0[0] = o[7];
  }
  lim_prim(a.gamma, own, pr[0]);
  long src[3];
  long cur = e;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    double cf[5];
    if (code[f] < 0) {
      const double H1 = 2.0 * A0[0] / gk[G_SJ + f];
      xc[f + 1] = xc[0] + 2.0 * gk[G_NX + f] * H1;
      yc[f + 1] = yc[0] + 2.0 * gk[G_NY + f] * H1;
      A0[f + 1] = 0.0;
      const int bi = -code[f] - 1;
      const int kind = (bi >= 0 && bi < a.bc.n) ? a.bc.kind[bi] : APL_BC_COPY;
      const double *bp = a.bc.p[(bi >= 0 && bi < a.bc.n) ? bi : 0];
      euler_bc5(kind, bp, 0.0, xc[f + 1], yc[f + 1], gk[G_NX + f], gk[G_NY + f], own, cf);
    } else {
      cur = code[f] >> 2;
      const double *o = a.avg + cur * 8;
#pragma unroll
      for (int c = 0; c < 5; ++c) cf[c] = o[c];
      xc[f + 1] = o[5];
      yc[f + 1] = o[6];
      A0[f + 1] = o[7];
    }
    src[f] = cur;
    lim_prim(a.gamma, cf, pr[f + 1]);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) A0[f + 1] += A0[0];

  // face averages at the six face ends, read at face-node position
  // p = element-node id of the end (:241-265).  The ends are the vertex
  // nodes (v1, v2, v2, v3, v3, v1): each face's node list starts where the
  // previous one ends (checked on the host), so three evaluations serve.
  double fp[6][5];
  const int ends[3] = {F[0], F[Nfp - 1], F[2 * Nfp - 1]};
  const double *qe = a.qin + (long)e * Np * M;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int p = ends[i];
    const int f = p / Nfp, j = p - f * Nfp;
    double qm[5], qp[5];
    load_euler5<EQN>(qe + F[p] * M, qm);
    const int cf = f == 0 ? code[0] : (f == 1 ? code[1] : code[2]);
    if (cf >= 0) {
      const long nb = cf >> 2;
      const int nf = cf & 3;
      load_euler5<EQN>(a.qin + (nb * Np + F[nf * Nfp + Nfp - 1 - j]) * M, qp);
    } else {
      const long sf = f == 0 ? src[0] : (f == 1 ? src[1] : src[2]);
      double qs[5];
      load_euler5<EQN>(a.qin + (sf * Np + F[p]) * M, qs);
      const int bi = -cf - 1;
      const int kind = (bi >= 0 && bi < a.bc.n) ? a.bc.kind[bi] : APL_BC_COPY;
      const double *bp = a.bc.p[(bi >= 0 && bi < a.bc.n) ? bi : 0];
      euler_bc5(kind, bp, 0.0, 0.0, 0.0, gk[G_NX + f], gk[G_NY + f], qs, qp);
    }
    double fc[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) fc[c] = 0.5 * (qp[c] + qm[c]);
    lim_prim(a.gamma, fc, fp[i == 0 ? 0 : 2 * i - 1]);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    fp[2][c] = fp[1][c];
    fp[4][c] = fp[3][c];
    fp[5][c] = fp[0][c];
  }

  // edge and cell gradients of the primitives (:269-314)
  const double xv1 = vk[0], yv1 = vk[1], xv2 = vk[2], yv2 = vk[3], xv3 = vk[4], yv3 = vk[5];
  double *gr = a.grad + (long)e * 40;
  const bool edge_rows = code[0] < 0 || code[1] < 0 || code[2] < 0;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double e1x = 0.5 * ((pr[1][c] - pr[0][c]) * (yv2 - yv1) + (fp[0][c] - fp[1][c]) * (yc[1] - yc[0])) / A0[1];
    const double e1y = -0.5 * ((pr[1][c] - pr[0][c]) * (xv2 - xv1) + (fp[0][c] - fp[1][c]) * (xc[1] - xc[0])) / A0[1];
    const double e2x = 0.5 * ((pr[2][c] - pr[0][c]) * (yv3 - yv2) + (fp[2][c] - fp[3][c]) * (yc[2] - yc[0])) / A0[2];
    const double e2y = -0.5 * ((pr[2][c] - pr[0][c]) * (xv3 - xv2) + (fp[2][c] - fp[3][c]) * (xc[2] - xc[0])) / A0[2];
    const double e3x = 0.5 * ((pr[3][c] - pr[0][c]) * (yv3 - yv2) + (fp[4][c] - fp[5][c]) * (yc[3] - yc[0])) / A0[3];
    const double e3y = -0.5 * ((pr[3][c] - pr[0][c]) * (xv3 - xv2) + (fp[4][c] - fp[5][c]) * (xc[3] - xc[0])) / A0[3];
    const double sa = A0[1] + A0[2] + A0[3];
    gr[0 * 5 + c] = (A0[1] * e1x + A0[2] * e2x + A0[3] * e3x) / sa;
    gr[1 * 5 + c] = (A0[1] * e1y + A0[2] * e2y + A0[3] * e3y) / sa;
    if (edge_rows) {
      gr[2 * 5 + c] = e1x;
      gr[3 * 5 + c] = e1y;
      gr[4 * 5 + c] = e2x;
      gr[5 * 5 + c] = e2y;
      gr[6 * 5 + c] = e3x;
      gr[7 * 5 + c] = e3y;
    }
  }
}

// One element of the reconstruction pass; writes its Np x M result to out.
template <int N, int EQN>
__device__ __forceinline__ void lim_reconstruct_element(const LimArgs &a, long e, double *out) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M;
  const double *rn = a.tab + Np, *sn = a.tab + 2 * Np, *drop = a.tab + 3 * Np;
  const double *vk = a.vert + (long)e * 6;
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];

  // limited gradients (:331-389)
  double Lx[5], Ly[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double gx[3], gy[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const long row = code[f] < 0 ? (long)e * 8 + 2 + 2 * f : (long)(code[f] >> 2) * 8;
      gx[f] = a.grad[row * 5 + c];
      gy[f] = a.grad[(row + 1) * 5 + c];
    }
    const double g1 = gx[0] * gx[0] + gy[0] * gy[0];
    const double g2 = gx[1] * gx[1] + gy[1] * gy[1];
    const double g3 = gx[2] * gx[2] + gy[2] * gy[2];
    const double fac = g1 * g1 + g2 * g2 + g3 * g3;
    const double epse = 1.e-14;
    const double fac3e = fac + 3.0 * epse;
    const double w1 = (g2 * g3 + epse) / fac3e;
    const double w2 = (g1 * g3 + epse) / fac3e;
    const double w3 = (g1 * g2 + epse) / fac3e;
    Lx[c] = w1 * gx[0] + w2 * gx[1] + w3 * gx[2];
    Ly[c] = w1 * gy[0] + w2 * gy[1] + w3 * gy[2];
  }

  // node coordinates, as the geometry builds them (wxnodaldggeometry2d.cc:211-218)
  double xn[Np], yn[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    const double r = rn[j], s = sn[j];
    xn[j] = 0.5 * (-vk[0] * (r + s) + vk[2] * (1. + r) + vk[4] * (1. + s));
    yn[j] = 0.5 * (-vk[1] * (r + s) + vk[3] * (1. + r) + vk[5] * (1. + s));
  }

  double AC[5], AP[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) AC[c] = a.avg[(long)e * 8 + c];
  lim_prim(a.gamma, AC, AP);
  const double g = a.gamma;
  const double averho = AP[0], aveu = AP[1], avev = AP[2], avew = AP[3], avep = AP[4];
  unsigned ef = 0;
#pragma unroll 1
  for (int n = 0; n < Np; ++n) {
    double ldx = 0.0, ldy = 0.0;
#pragma unroll
    for (int j = 0; j < Np; ++j) {
      ldx += drop[n * Np + j] * xn[j];
      ldy += drop[n * Np + j] * yn[j];
    }
    double drho = ldx * Lx[0] + ldy * Ly[0];
    const double du = ldx * Lx[1] + ldy * Ly[1];
    const double dv = ldx * Lx[2] + ldy * Ly[2];
    const double dw = ldx * Lx[3] + ldy * Ly[3];
    double dp = ldx * Lx[4] + ldy * Ly[4];

    // reconstruction with the density and pressure floors (:394-466)
    double Lrho = averho + drho;
    if (Lrho < a.min_dens) {
      for (int it = 0; it < 4; ++it) {
        drho *= 0.5;
        Lrho = averho + drho;
        if (Lrho > a.min_dens) break;
        else if (it == 3) Lrho = AC[0];
      }
    }
    const double Lrhou = AC[1] + averho * du + drho * aveu;
    const double Lrhov = AC[2] + averho * dv + drho * avev;
    const double Lrhow = AC[3] + averho * dw + drho * avew;
    const double ke = 0.5 * (Lrhou * Lrhou + Lrhov * Lrhov + Lrhow * Lrhow);
    double dEner = dp / (g - 1.) + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
                   averho * (aveu * du + avev * dv + avew * dw);
    double LEner = AC[4] + dEner;
    double Lp = (g - 1.) * (LEner - ke / Lrho);
    if (Lp < a.min_pres) {
      for (int it = 0; it < 4; ++it) {
        dp *= 0.5;
        dEner = dp / (g - 1.) + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
                averho * (aveu * du + avev * dv + avew * dw);
        LEner = AC[4] + dEner;
        Lp = (g - 1.) * (LEner - ke / Lrho);
        if (Lp > a.min_pres) break;
        else if (it == 3) LEner = avep / (g - 1.) + ke / Lrho;
      }
    }
    const double o5[5] = {Lrho, Lrhou, Lrhov, Lrhow, LEner};
#pragma unroll
    for (int c = 0; c < 5; ++c)
      if (!__builtin_isfinite(o5[c])) ef |= EF_NAN;   // wxnodaldg2dmethod.cc:246
    store_euler5<EQN>(o5, out + n * M);
  }
  if (ef) atomicOr(a.err, ef);
}

template <int N, int EQN>
__global__ void __launch_bounds__(64) lim_reconstruct_kernel(LimArgs a) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, ES = Np * M;
  __shared__ __align__(16) double st[LE * ES];
  const int lane = threadIdx.x;
  const long e0 = (long)blockIdx.x * LE;
  const int cnt = (int)((a.K - e0) < LE ? (a.K - e0) : LE);
  if (lane < cnt) lim_reconstruct_element<N, EQN>(a, e0 + lane, st + lane * ES);
  __syncthreads();
  lim_stage_out<ES>(st, a.qout + e0 * ES, cnt, lane);
}

// Launch entry of one order (defined in apl_kernels_N<N>.hip for N <= 4).
template <int N>
int launch_limiter(int eqn, hipStream_t s, const LimArgs &a);

}  // namespace apl
