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
// The average and reconstruction passes move the state of 32 consecutive
// elements per wave between HBM and LDS with coalesced 16-byte accesses;
// the average pass works one lane per element on the LDS copy, the
// reconstruction two lanes per element (half of its nodes each).  The
// gradient pass, a gather of a few hundred bytes per element, runs one
// thread per element and issues all of its reads at once.  The
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
  double *grad;            // [K][2][5] cell gradients C0x, C0y, then
                           //          [K][6][5] edge gradients E1x, E1y,
                           //          E2x, E2y, E3x, E3y, written only for
                           //          elements with a boundary face (their
                           //          only readers)
  const double *geom;      // [K][G_WORDS]
  const double *vert;      // [K][6]
  const int *fcode;        // [K][3]
  const int *fmask;        // [3][Nfp]
  const double *tab;       // ave[Np], r[Np], s[Np]
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
  // the element's vertices, read together with the state
  const long e = e0 + lane;
  double vk[6];
  if (lane < cnt) {
#pragma unroll
    for (int i = 0; i < 6; ++i) vk[i] = a.vert[e * 6 + i];
  }
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
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];

  // src[f] is the element whose averages an interior face f reads, and
  // whose nodes a boundary face f's states read: the last neighbour visited
  // up to face f, or the element itself (:136-233).  Every address below
  // follows from the face codes and the (uniform) Fmask entries, so all the
  // reads of the pass are issued together, before any arithmetic.
  long src[3];
  {
    long cur = e;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (code[f] >= 0) cur = code[f] >> 2;
      src[f] = cur;
    }
  }
  // face averages at the six face ends are read at face-node position
  // p = element-node id of the end (:241-265).  The ends are the vertex
  // nodes (v1, v2, v2, v3, v3, v1): each face's node list starts where the
  // previous one ends (checked on the host), so three evaluations serve.
  // End i lies on face fe[i] at position je[i]; the neighbour across it
  // reads its face node Nfp-1-je[i] of whichever face nf it shares.
  constexpr int pe[3] = {0, Nfp - 1, 2 * Nfp - 1};   // positions in F of the ends
  int nodeq[3], fe[3], nbr[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int p = F[pe[i]];
    fe[i] = p / Nfp;
    const int j = p - fe[i] * Nfp;
    nodeq[i] = F[p];
#pragma unroll
    for (int nf = 0; nf < 3; ++nf) nbr[i][nf] = F[nf * Nfp + Nfp - 1 - j];
  }

  double av[4][8];   // own row, then the rows of src[0..2]
#pragma unroll
  for (int w = 0; w < 8; ++w) av[0][w] = a.avg[(long)e * 8 + w];
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int w = 0; w < 8; ++w) av[f + 1][w] = a.avg[src[f] * 8 + w];
  double qm[3][5], qo[3][5];
  const double *qe = a.qin + (long)e * Np * M;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int f = fe[i];
    const int cf = f == 0 ? code[0] : (f == 1 ? code[1] : code[2]);
    const long sf = f == 0 ? src[0] : (f == 1 ? src[1] : src[2]);
    const int nf = cf & 3;
    const int nn = nf == 0 ? nbr[i][0] : (nf == 1 ? nbr[i][1] : nbr[i][2]);
    const long row = cf >= 0 ? (long)(cf >> 2) * Np + nn : sf * Np + nodeq[i];
    load_euler5<EQN>(qe + nodeq[i] * M, qm[i]);
    load_euler5<EQN>(a.qin + row * M, qo[i]);
  }
  const double *vk = a.vert + (long)e * 6;
  const double xv1 = vk[0], yv1 = vk[1], xv2 = vk[2], yv2 = vk[3], xv3 = vk[4], yv3 = vk[5];

  // patch centres and areas (:121-147) and patch averages (:149-233)
  const bool bnd = code[0] < 0 || code[1] < 0 || code[2] < 0;
  const double *gk = a.geom + (long)e * G_WORDS;
  double xc[4], yc[4], A0[4], pr[4][5], own[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) own[c] = av[0][c];
  xc[0] = av[0][5];
  yc[0] = av[0][6];
  A0[0] = av[0][7];
  lim_prim(a.gamma, own, pr[0]);
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
#pragma unroll
      for (int c = 0; c < 5; ++c) cf[c] = av[f + 1][c];
      xc[f + 1] = av[f + 1][5];
      yc[f + 1] = av[f + 1][6];
      A0[f + 1] = av[f + 1][7];
    }
    lim_prim(a.gamma, cf, pr[f + 1]);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) A0[f + 1] += A0[0];

  double fp[6][5];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int f = fe[i];
    const int cf = f == 0 ? code[0] : (f == 1 ? code[1] : code[2]);
    double qp[5];
    if (cf >= 0) {
#pragma unroll
      for (int c = 0; c < 5; ++c) qp[c] = qo[i][c];
    } else {
      const int bi = -cf - 1;
      const int kind = (bi >= 0 && bi < a.bc.n) ? a.bc.kind[bi] : APL_BC_COPY;
      const double *bp = a.bc.p[(bi >= 0 && bi < a.bc.n) ? bi : 0];
      euler_bc5(kind, bp, 0.0, 0.0, 0.0, gk[G_NX + f], gk[G_NY + f], qo[i], qp);
    }
    double fc[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) fc[c] = 0.5 * (qp[c] + qm[i][c]);
    lim_prim(a.gamma, fc, fp[i == 0 ? 0 : 2 * i - 1]);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    fp[2][c] = fp[1][c];
    fp[4][c] = fp[3][c];
    fp[5][c] = fp[0][c];
  }

  // edge and cell gradients of the primitives (:269-314); the cell
  // gradient C0 goes to the dense [K][2][5] block, the edge gradients of an
  // element with a boundary face to the [K][6][5] block after it
  double *g0 = a.grad + (long)e * 10;
  double *ge = a.grad + (long)a.K * 10 + (long)e * 30;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double e1x = 0.5 * ((pr[1][c] - pr[0][c]) * (yv2 - yv1) + (fp[0][c] - fp[1][c]) * (yc[1] - yc[0])) / A0[1];
    const double e1y = -0.5 * ((pr[1][c] - pr[0][c]) * (xv2 - xv1) + (fp[0][c] - fp[1][c]) * (xc[1] - xc[0])) / A0[1];
    const double e2x = 0.5 * ((pr[2][c] - pr[0][c]) * (yv3 - yv2) + (fp[2][c] - fp[3][c]) * (yc[2] - yc[0])) / A0[2];
    const double e2y = -0.5 * ((pr[2][c] - pr[0][c]) * (xv3 - xv2) + (fp[2][c] - fp[3][c]) * (xc[2] - xc[0])) / A0[2];
    const double e3x = 0.5 * ((pr[3][c] - pr[0][c]) * (yv3 - yv2) + (fp[4][c] - fp[5][c]) * (yc[3] - yc[0])) / A0[3];
    const double e3y = -0.5 * ((pr[3][c] - pr[0][c]) * (xv3 - xv2) + (fp[4][c] - fp[5][c]) * (xc[3] - xc[0])) / A0[3];
    const double sa = A0[1] + A0[2] + A0[3];
    g0[c] = (A0[1] * e1x + A0[2] * e2x + A0[3] * e3x) / sa;
    g0[5 + c] = (A0[1] * e1y + A0[2] * e2y + A0[3] * e3y) / sa;
    if (bnd) {
      ge[0 * 5 + c] = e1x;
      ge[1 * 5 + c] = e1y;
      ge[2 * 5 + c] = e2x;
      ge[3 * 5 + c] = e2y;
      ge[4 * 5 + c] = e3x;
      ge[5 * 5 + c] = e3y;
    }
  }
}

// One element of the reconstruction pass, shared by two lanes: both form
// the limited gradients, lane `half` reconstructs nodes [half*NH, half*NH+NH)
// and writes them to out (the element's Np x M block).
template <int N, int EQN>
__device__ __forceinline__ void lim_reconstruct_element(const LimArgs &a, long e, int half,
                                                        double *out) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, NH = (Np + 1) / 2;
  const double *ave = a.tab, *rn = a.tab + Np, *sn = a.tab + 2 * Np;
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];
  const double *vk = a.vert + (long)e * 6;
  const double v0 = vk[0], v1 = vk[1], v2 = vk[2], v3 = vk[3], v4 = vk[4], v5 = vk[5];
  double AC[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) AC[c] = a.avg[(long)e * 8 + c];

  // limited gradients (:331-389): a neighbour's cell gradient across an
  // interior face, the element's own edge gradient across a boundary face
  const double *gf[3];
#pragma unroll
  for (int f = 0; f < 3; ++f)
    gf[f] = code[f] < 0 ? a.grad + (long)a.K * 10 + (long)e * 30 + 10 * f
                        : a.grad + (long)(code[f] >> 2) * 10;
  double gx[3][5], gy[3][5];
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      gx[f][c] = gf[f][c];
      gy[f][c] = gf[f][5 + c];
    }
  double Lx[5], Ly[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double g1 = gx[0][c] * gx[0][c] + gy[0][c] * gy[0][c];
    const double g2 = gx[1][c] * gx[1][c] + gy[1][c] * gy[1][c];
    const double g3 = gx[2][c] * gx[2][c] + gy[2][c] * gy[2][c];
    const double fac = g1 * g1 + g2 * g2 + g3 * g3;
    const double epse = 1.e-14;
    const double fac3e = fac + 3.0 * epse;
    const double w1 = (g2 * g3 + epse) / fac3e;
    const double w2 = (g1 * g3 + epse) / fac3e;
    const double w3 = (g1 * g2 + epse) / fac3e;
    Lx[c] = w1 * gx[0][c] + w2 * gx[1][c] + w3 * gx[2][c];
    Ly[c] = w1 * gy[0][c] + w2 * gy[1][c] + w3 * gy[2][c];
  }

  // node coordinates, as the geometry builds them (wxnodaldggeometry2d.cc:211-218),
  // and their offsets from the element's node average.  The reference
  // applies dropAVE = I - 1 ave^T as a matrix (:331-340); x_n - sum_j ave_j x_j
  // is the same product, rounded differently.
  double xb = 0.0, yb = 0.0;
#pragma unroll
  for (int j = 0; j < Np; ++j) {
    const double r = rn[j], s = sn[j];
    xb += ave[j] * (0.5 * (-v0 * (r + s) + v2 * (1. + r) + v4 * (1. + s)));
    yb += ave[j] * (0.5 * (-v1 * (r + s) + v3 * (1. + r) + v5 * (1. + s)));
  }

  double AP[5];
  lim_prim(a.gamma, AC, AP);
  const double g = a.gamma;
  const double averho = AP[0], aveu = AP[1], avev = AP[2], avew = AP[3], avep = AP[4];
  unsigned ef = 0;
  const int nb = half * NH;
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int n = nb + k;
    if (n >= Np) break;
    const double r = rn[n], s = sn[n];
    const double ldx = 0.5 * (-v0 * (r + s) + v2 * (1. + r) + v4 * (1. + s)) - xb;
    const double ldy = 0.5 * (-v1 * (r + s) + v3 * (1. + r) + v5 * (1. + s)) - yb;
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

// LE elements per 64-lane wave, lanes l and l + LE on element l.
template <int N, int EQN>
__global__ void __launch_bounds__(64) lim_reconstruct_kernel(LimArgs a) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, ES = Np * M;
  static_assert(2 * LE == 64, "two lanes per element");
  __shared__ __align__(16) double st[LE * ES];
  const int lane = threadIdx.x, el = lane % LE, half = lane / LE;
  const long e0 = (long)blockIdx.x * LE;
  const int cnt = (int)((a.K - e0) < LE ? (a.K - e0) : LE);
  if (el < cnt) lim_reconstruct_element<N, EQN>(a, e0 + el, half, st + el * ES);
  __syncthreads();
  lim_stage_out<ES>(st, a.qout + e0 * ES, cnt, lane);
}

// Launch entry of one order (defined in apl_kernels_N<N>.hip for N <= 4).
template <int N>
int launch_limiter(int eqn, hipStream_t s, const LimArgs &a);

}  // namespace apl
