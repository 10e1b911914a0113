// apl_limiter.h -- Apollo's Hesthaven-Warburton Euler limiter
// (src/hyperapps/euler/wxhesthavenwarburtoneulerlimiter.cc:61-484) on the
// GPU, as three element-parallel passes over HBM-resident state:
//
//   lim_patch_kernel        (once per handle) each element's patch centre
//                           and area share, mesh constants;
//   lim_average_kernel      cell averages of the conserved variables
//                           (computeConservedAndPrimitiveAVEVariables,
//                           :487-522), one FP64 5-vector per element, and
//                           the states at the few nodes the gradient pass
//                           reads (LimTrace), so that it gathers from a
//                           compact array instead of the full state (at
//                           N = 1 those nodes are the whole state, which it
//                           then reads directly);
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
// The average and reconstruction passes move the state of consecutive
// elements between HBM and LDS with coalesced 16-byte accesses: the average
// pass 32 elements per wave, one lane per element on the LDS copy; the
// reconstruction 16 elements per wave, four lanes per element (a quarter
// of its nodes each).  The gradient pass, a gather of a few hundred bytes
// per element, runs one thread per element and issues all of its reads at
// once.  The
// reference's reading order and quirks are kept (see the oracle's
// orc_limit): AVE[0..2] with the vertex nodes for the patch centres, the
// boundary-face node states read from the last neighbour visited, the face
// ends indexed by element-node id into the face-ordered averages, and the
// (yv3 - yv2), (xv3 - xv2) factors of the third edge gradient.  Orders
// N <= 4 only: for N > 4 the reference reads outside its face array.
#pragma once

#include "apl_device.h"

namespace apl {

// The element nodes whose states the gradient pass reads, as slots of a
// per-element trace the average pass writes.  End i of the three face ends
// lies on face `face[i]`; its own state is slot `own[i]` of the element (or,
// on a boundary face, of the last neighbour visited), and the neighbour
// sharing its face nf contributes slot `nbr[i][nf]`.  Built on the host from
// Fmask: the three vertices for N = 1, 2, 4, seven nodes for N = 3.
constexpr int LIM_MAX_TR = 12;
struct LimTrace {
  int s;
  int node[LIM_MAX_TR];
  int face[3];
  int own[3];
  int nbr[3][3];
};

struct LimArgs {
  int K;
  const double *qin;       // [K][Np][M]
  double *qout;            // [K][Np][M], may equal qin
  double *avg;             // [K][8]   conserved cell averages (every call),
                           //          then the patch centre xc, yc and
                           //          area share A0 (once per handle)
  double *grad;            // [K][2][5] cell gradients C0x, C0y, then
                           //          [K][6][5] edge gradients E1x, E1y,
                           //          E2x, E2y, E3x, E3y, written only for
                           //          elements with a boundary face (their
                           //          only readers)
  const double *geom;      // [K][G_WORDS]
  const double *vert;      // [K][6]
  const int *fcode;        // [K][3]
  double *trc;             // [K][tr.s][5] states at the trace nodes, written by
                           //          the average pass when trc_write
  const double *trc_rd;    // the trace the gradient pass reads: trc, or the
                           //          state itself when the slots are nodes
                           //          0 .. Np-1 of a 5-variable state (N = 1)
  int trc_write;
  int clear_err;           // the average pass zeroes *err first (the passes
                           //          before the reconstruction set no flags)
  LimTrace tr;
  const double *tab;       // ave[Np], then r[Np] and s[Np] centred on
                           //          their ave-weighted means
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

// The state of LB consecutive elements moves between HBM and LDS with
// coalesced 16-byte accesses by all 64 lanes of a wave.  The average pass
// takes LE elements per wave, one lane each; the reconstruction LER
// elements per wave, 64 / LER lanes each.
constexpr int LE = 32;
constexpr int LER = 16;

template <int ES, int LB>
__device__ __forceinline__ void lim_stage_in(const double *src, double *lds, int cnt, int lane) {
  const int n = cnt * ES;
  if (cnt == LB) {
    // every load of the wave in flight before the first LDS write
    constexpr int NV2 = LB * ES / 2, NIT = (NV2 + 63) / 64;
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

template <int ES, int LB>
__device__ __forceinline__ void lim_stage_out(const double *lds, double *dst, int cnt, int lane) {
  const int n = cnt * ES;
  if (cnt == LB) {
    constexpr int NV2 = LB * ES / 2;
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
  const long e = e0 + lane;
  if (a.clear_err && blockIdx.x == 0 && lane == 0) *a.err = 0u;
  lim_stage_in<ES, LE>(a.qin + e0 * ES, st, cnt, lane);
  __syncthreads();
  if (lane < cnt) {
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
    // words 5..7 of the row hold the patch geometry (lim_patch_kernel)
    double *o = a.avg + e * 8;
#pragma unroll
    for (int c = 0; c < 5; ++c) o[c] = c5[c];
  }
  // the trace of the wave's elements, one contiguous [cnt][S][5] block,
  // gathered from the LDS copy with consecutive lanes on consecutive words
  if (!a.trc_write) return;
  const int S5 = a.tr.s * 5;
  double *tb = a.trc + e0 * S5;
  for (int t = lane; t < cnt * S5; t += 64) {
    const int l = t / S5, rem = t - l * S5, sl = rem / 5, c = rem - sl * 5;
    int node = a.tr.node[0];
#pragma unroll
    for (int k = 1; k < LIM_MAX_TR; ++k) node = sl == k ? a.tr.node[k] : node;
    const double *q = st + l * ES + node * M;
    double v;
    if constexpr (M == 5) v = q[c];
    else v = c == 3 ? 0.0 : q[c == 4 ? 3 : c];   // EULER4: rho w == 0, E last
    tb[t] = v;
  }
}

template <int N, int EQN>
__global__ void __launch_bounds__(256) lim_gradient_kernel(LimArgs a) {
  static_assert(N <= 4, "the reference's face-end reads leave its face array for N > 4");
  // at N = 1 each XCD takes one contiguous range of elements, so that most
  // neighbour rows are in its own L2 (K = 524288: 55.8 -> 50.3 us); at N = 4
  // the plain order measured 3 us faster
  const int blk = N == 1 ? xcd_batch(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int e = blk * blockDim.x + threadIdx.x;
  if (e >= a.K) return;
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];

  // src[f] is the element whose averages an interior face f reads, and
  // whose nodes a boundary face f's states read: the last neighbour visited
  // up to face f, or the element itself (:136-233).  Every address below
  // follows from the face codes and the (uniform) trace slots, so all the
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
  // Their node states come from the trace slots (LimTrace).
  double av[4][8];   // own row, then the rows of src[0..2]
#pragma unroll
  for (int w = 0; w < 8; ++w) av[0][w] = a.avg[(long)e * 8 + w];
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int w = 0; w < 8; ++w) av[f + 1][w] = a.avg[src[f] * 8 + w];
  const long S = a.tr.s;
  double qm[3][5], qo[3][5];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int f = a.tr.face[i];
    const int cf = f == 0 ? code[0] : (f == 1 ? code[1] : code[2]);
    const long sf = f == 0 ? src[0] : (f == 1 ? src[1] : src[2]);
    const int nf = cf & 3;
    const int ns = nf == 0 ? a.tr.nbr[i][0] : (nf == 1 ? a.tr.nbr[i][1] : a.tr.nbr[i][2]);
    const long row = cf >= 0 ? (long)(cf >> 2) * S + ns : sf * S + a.tr.own[i];
    const double *pm = a.trc_rd + ((long)e * S + a.tr.own[i]) * 5, *po = a.trc_rd + row * 5;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      qm[i][c] = pm[c];
      qo[i][c] = po[c];
    }
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
    const int f = a.tr.face[i];
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

// One element of the reconstruction pass, shared by P = 64 / LER lanes:
// each forms the limited gradients, lane `part` reconstructs nodes
// [part*NH, part*NH + NH) and writes them to out (the element's Np x M
// block).
template <int N, int EQN>
__device__ __forceinline__ void lim_reconstruct_element(const LimArgs &a, long e, int part,
                                                        double *out) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, P = 64 / LER, NH = (Np + P - 1) / P;
  const double *cr = a.tab + Np, *cs = a.tab + 2 * Np;
  int code[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) code[f] = a.fcode[(long)e * 3 + f];
  const double *vk = a.vert + (long)e * 6;
  const double dx1 = vk[2] - vk[0], dy1 = vk[3] - vk[1], dx2 = vk[4] - vk[0], dy2 = vk[5] - vk[1];
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
    // the three weights share one reciprocal of fac + 3 eps (>= 3e-14)
    const double ifac = 1.0 / (fac + 3.0 * epse);
    const double w1 = (g2 * g3 + epse) * ifac;
    const double w2 = (g1 * g3 + epse) * ifac;
    const double w3 = (g1 * g2 + epse) * ifac;
    Lx[c] = w1 * gx[0][c] + w2 * gx[1][c] + w3 * gx[2][c];
    Ly[c] = w1 * gy[0][c] + w2 * gy[1][c] + w3 * gy[2][c];
  }

  double AP[5];
  lim_prim(a.gamma, AC, AP);
  const double g = a.gamma, igm1 = 1.0 / (g - 1.);
  const double averho = AP[0], aveu = AP[1], avev = AP[2], avew = AP[3], avep = AP[4];
  unsigned ef = 0;
  const int nb = part * NH;
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int n = nb + k;
    if (n >= Np) break;
    // node n's offset from the element's node average.  The reference
    // forms the node coordinates (wxnodaldggeometry2d.cc:211-218) and
    // applies dropAVE = I - 1 ave^T (:331-340); with x_n = x_1 +
    // ((x_2 - x_1)(1 + r_n) + (x_3 - x_1)(1 + s_n)) / 2 that product is
    // ((x_2 - x_1) cr_n + (x_3 - x_1) cs_n) / 2 with the centred reference
    // coordinates cr_n = r_n - sum_j ave_j r_j (and cs_n likewise), free of
    // the cancellation of the absolute coordinates.
    const double ldx = 0.5 * (dx1 * cr[n] + dx2 * cs[n]);
    const double ldy = 0.5 * (dy1 * cr[n] + dy2 * cs[n]);
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
    double dEner = dp * igm1 + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
                   averho * (aveu * du + avev * dv + avew * dw);
    double LEner = AC[4] + dEner;
    double Lp = (g - 1.) * (LEner - ke / Lrho);
    if (Lp < a.min_pres) {
      for (int it = 0; it < 4; ++it) {
        dp *= 0.5;
        dEner = dp * igm1 + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
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

// LER elements per 64-lane wave; lanes l, l + LER, l + 2 LER, ... on element l.
template <int N, int EQN>
__global__ void __launch_bounds__(64) lim_reconstruct_kernel(LimArgs a) {
  constexpr int Np = Order<N>::Np, M = Vars<EQN>::M, ES = Np * M;
  static_assert(64 % LER == 0, "whole lanes per element");
  __shared__ __align__(16) double st[LER * ES];
  const int lane = threadIdx.x, el = lane % LER, part = lane / LER;
  const long e0 = (long)blockIdx.x * LER;
  const int cnt = (int)((a.K - e0) < LER ? (a.K - e0) : LER);
  if (el < cnt) lim_reconstruct_element<N, EQN>(a, e0 + el, part, st + el * ES);
  __syncthreads();
  lim_stage_out<ES, LER>(st, a.qout + e0 * ES, cnt, lane);
}

// Launch entry of one order (defined in apl_kernels_N<N>.hip for N <= 4).
template <int N>
int launch_limiter(int eqn, hipStream_t s, const LimArgs &a);

// The mesh constants of the averages rows, once per handle: each element's
// patch centre xc, yc and area share A0 (:121-126), which its neighbours'
// patches reuse (:136-144), into words 5..7 of its row.  Needs a.K, a.vert,
// a.tab (the weights of the vertex nodes) and a.avg.
int launch_limiter_patch(hipStream_t s, const LimArgs &a);

}  // namespace apl
