/*
 * apollo_oracle.c -- CPU restatement of Apollo's nodal-DG RHS.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the MI355X engine and
 * is never linked into it.  Each function follows the reference file:line
 * it cites (paths relative to the Apollo source tree).  Arithmetic is kept
 * in the reference's order where that is cheap, so that differences to
 * Apollo itself come only from PETSc's dense LU / MatMatMult rounding.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "apollo_oracle.h"
#include "../apollo_amd/csrc/apl_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ALLOC(T, n) ((T *)calloc((size_t)(n) > 0 ? (size_t)(n) : 1, sizeof(T)))

/* ------------------------------------------------------------------------ */
/* Polynomials: src/lib/wxmath.h:303-368                                    */
/* ------------------------------------------------------------------------ */

/* JacobiP, orthonormal Jacobi polynomial (wxmath.h:303-348). */
double orc_jacobiP(double x, double alpha, double beta, int N) {
  double aold = 0.0, anew = 0.0, bnew = 0.0, h1 = 0.0;
  double gamma0 = 0.0, gamma1 = 0.0;
  double ab = alpha + beta, ab1 = alpha + beta + 1.0, a1 = alpha + 1.0, b1 = beta + 1.0;
  double PL0, PL1, PL2 = 0.0, prow, x_bnew;
  gamma0 = pow(2.0, ab1) / (ab1)*tgamma(a1) * tgamma(b1) / tgamma(ab1);
  if (0 == N) return 1.0 / sqrt(gamma0);
  PL0 = 1.0 / sqrt(gamma0);
  gamma1 = (a1) * (b1) / (ab + 3.0) * gamma0;
  prow = ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / sqrt(gamma1);
  if (1 == N) return prow;
  PL1 = prow;
  aold = 2.0 / (2.0 + ab) * sqrt((a1) * (b1) / (ab + 3.0));
  for (int i = 1; i <= (N - 1); ++i) {
    h1 = 2.0 * i + ab;
    anew = 2.0 / (h1 + 2.0) * sqrt((i + 1) * (i + ab1) * (i + a1) * (i + b1) / (h1 + 1.0) / (h1 + 3.0));
    bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    x_bnew = x - bnew;
    PL2 = 1.0 / anew * (-aold * PL0 + x_bnew * PL1);
    aold = anew;
    PL0 = PL1;
    PL1 = PL2;
  }
  return PL2;
}

/* GradJacobiP (wxmath.h:350-368). */
double orc_grad_jacobiP(double z, double alpha, double beta, int N) {
  if (0 == N) return 0.0;
  return sqrt(N * (N + alpha + beta + 1)) * orc_jacobiP(z, alpha + 1, beta + 1, N - 1);
}

/* Simplex2DP (wxNodalDGMatrices.h:7-20). */
static double simplex2DP(double a, double b, int i, int j) {
  double number = 2. * i + 1., zero = 0;
  double h1 = orc_jacobiP(a, zero, zero, i), h2 = orc_jacobiP(b, number, zero, j);
  double tv1 = sqrt(2.) * h1 * h2, tv2 = pow(1.0 - b, (double)i);
  return tv1 * tv2;
}

/* GradSimplex2DP (wxNodalDGMatrices.h:67-105). */
static void grad_simplex2DP(double a, double b, int id, int jd, double *dr, double *ds) {
  double fa, dfa, gb, dgb, tmp, dmodedr, dmodeds;
  fa = orc_jacobiP(a, 0.0, 0.0, id);
  dfa = orc_grad_jacobiP(a, 0.0, 0.0, id);
  gb = orc_jacobiP(b, 2.0 * id + 1, 0.0, jd);
  dgb = orc_grad_jacobiP(b, 2.0 * id + 1, 0.0, jd);
  dmodedr = dfa * gb;
  if (id > 0) dmodedr *= pow(0.5 * (1.0 - b), (id - 1.0));
  dmodeds = dfa * gb * 0.5 * (1.0 + a);
  if (id > 0) dmodeds *= pow(0.5 * (1.0 - b), (id - 1.0));
  tmp = dgb * pow(0.5 * (1.0 - b), (double)id);
  if (id > 0) tmp -= (0.5 * id) * gb * pow(0.5 * (1.0 - b), (id - 1.0));
  dmodeds += fa * tmp;
  dmodedr *= pow(2.0, (id + 0.5));
  dmodeds *= pow(2.0, (id + 0.5));
  *dr = dmodedr;
  *ds = dmodeds;
}

static void rs_to_ab(double r, double s, double *a, double *b) {
  *b = s;
  if (s == 1.0) *a = -1.;
  else *a = 2.0 * (1. + r) / (1.0 - s) - 1.0;
}

/* Vandermonde2D (wxNodalDGMatrices.h:35-65): V[k][mode], row-major Nx x Np. */
static void vandermonde2D(int N, int Nx, const double *rr, const double *ss, double *V) {
  int Np = (N + 1) * (N + 2) / 2, sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      for (int k = 0; k < Nx; k++) {
        double a, b;
        rs_to_ab(rr[k], ss[k], &a, &b);
        V[k * Np + sk] = simplex2DP(a, b, i, j);
      }
      ++sk;
    }
}

/* GradVandermonde2D (wxNodalDGMatrices.h:107-142). */
static void grad_vandermonde2D(int N, int Nx, const double *rr, const double *ss,
                               double *Vr, double *Vs) {
  int Np = (N + 1) * (N + 2) / 2, sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      for (int k = 0; k < Nx; k++) {
        double a, b, dr, ds;
        rs_to_ab(rr[k], ss[k], &a, &b);
        grad_simplex2DP(a, b, i, j, &dr, &ds);
        Vr[k * Np + sk] = dr;
        Vs[k * Np + sk] = ds;
      }
      ++sk;
    }
}

/* gauleg (wxmath.h:171-206), with 0-based output arrays. */
void orc_gauleg(double x1, double x2, double *x0, double *w0, int n) {
  double *x = x0 - 1, *w = w0 - 1;
  int m, j, i;
  double z1, z, xm, xl, pp, p3, p2, p1;
  double EPS = 3.0e-13;
  double PI = 3.14159265358979323846;
  m = (n + 1) / 2;
  xm = 0.5 * (x2 + x1);
  xl = 0.5 * (x2 - x1);
  for (i = 1; i <= m; i++) {
    z = cos(PI * (i - 0.25) / (n + 0.5));
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > EPS);
    x[i] = xm - xl * z;
    x[n + 1 - i] = xm + xl * z;
    w[i] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n + 1 - i] = w[i];
  }
}

/* Dense inverse by LU with partial pivoting.  Stands in for PETSc's dense
 * MatLUFactor + MatMatSolve (wxnodaldggeometry2d.cc:147-177); last-bit
 * rounding relative to PETSc: parity unpinned. */
int orc_invert(int n, const double *A, double *Ainv) {
  double *LU = ALLOC(double, n * n);
  int *piv = ALLOC(int, n);
  memcpy(LU, A, sizeof(double) * n * n);
  for (int i = 0; i < n; i++) piv[i] = i;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(LU[k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(LU[i * n + k]) > best) { best = fabs(LU[i * n + k]); p = i; }
    if (best == 0.0) { free(LU); free(piv); return 1; }
    if (p != k) {
      for (int j = 0; j < n; j++) { double t = LU[k * n + j]; LU[k * n + j] = LU[p * n + j]; LU[p * n + j] = t; }
      int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
    }
    for (int i = k + 1; i < n; i++) {
      double l = LU[i * n + k] / LU[k * n + k];
      LU[i * n + k] = l;
      for (int j = k + 1; j < n; j++) LU[i * n + j] -= l * LU[k * n + j];
    }
  }
  /* solve LU X = P I column by column */
  for (int c = 0; c < n; c++) {
    double *x = ALLOC(double, n);
    for (int i = 0; i < n; i++) x[i] = (piv[i] == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) x[i] -= LU[i * n + j] * x[j];
    for (int i = n - 1; i >= 0; i--) {
      for (int j = i + 1; j < n; j++) x[i] -= LU[i * n + j] * x[j];
      x[i] /= LU[i * n + i];
    }
    for (int i = 0; i < n; i++) Ainv[i * n + c] = x[i];
    free(x);
  }
  free(LU);
  free(piv);
  return 0;
}

static void matmul(int M, int Kd, int Nn, const double *A, const double *B, double *C) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < Nn; j++) {
      double acc = 0.0;
      for (int k = 0; k < Kd; k++) acc += A[i * Kd + k] * B[k * Nn + j];
      C[i * Nn + j] = acc;
    }
}

static void transpose(int M, int Nn, const double *A, double *AT) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < Nn; j++) AT[j * M + i] = A[i * Nn + j];
}

/* Operators: wxNodalDGgeometry2D ctor (wxnodaldggeometry2d.cc:34-98) and
 * WxCubature2d ctor (wxcubature2d.cc:18-156). */
int orc_build_ops(int N, orc_ops *o) {
  memset(o, 0, sizeof(*o));
  if (N < 1 || N > 8) return 2;
  int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  int cubOrd = (int)floor(3.0 * (N + 1));          /* wxcubature2d.cc:20 */
  int Ng = (int)floor((N + 1) * 2);                 /* wxcubature2d.cc:23 */
  int Nc = apl_tab_cub_npts[N - 1];                 /* wxcubature2d.cc:26-28, rule cubOrd */
  o->N = N; o->Np = Np; o->Nfp = Nfp; o->Nc = Nc; o->Ng = Ng;
  o->r = ALLOC(double, Np); o->s = ALLOC(double, Np);
  o->fmask = ALLOC(int, 3 * Nfp);
  /* nodalNaturalCoordinates (wxpnodaldgfunctions.h:258-278) */
  for (int i = 0; i < Np; i++) {
    o->r[i] = apl_tab_node_r[apl_tab_node_off[N - 1] + i];
    o->s[i] = apl_tab_node_s[apl_tab_node_off[N - 1] + i];
  }
  for (int i = 0; i < 3 * Nfp; i++) o->fmask[i] = apl_tab_fmask[apl_tab_fmask_off[N - 1] + i];
  o->rmin = fabs(o->r[1] - o->r[0]);               /* wxnodaldggeometry2d.cc:91 */

  o->V = ALLOC(double, Np * Np); o->invV = ALLOC(double, Np * Np);
  o->VVT = ALLOC(double, Np * Np);
  o->Dr = ALLOC(double, Np * Np); o->Ds = ALLOC(double, Np * Np);
  vandermonde2D(N, Np, o->r, o->s, o->V);
  if (orc_invert(Np, o->V, o->invV)) return 1;
  {
    double *Vr = ALLOC(double, Np * Np), *Vs = ALLOC(double, Np * Np), *VT = ALLOC(double, Np * Np);
    grad_vandermonde2D(N, Np, o->r, o->s, Vr, Vs);
    matmul(Np, Np, Np, Vr, o->invV, o->Dr);
    matmul(Np, Np, Np, Vs, o->invV, o->Ds);
    transpose(Np, Np, o->V, VT);
    matmul(Np, Np, Np, o->V, VT, o->VVT);           /* wxnodaldggeometry2d.cc:72-74 */
    free(Vr); free(Vs); free(VT);
  }
  /* cubature (wxcubature2d.cc:29-75) */
  o->cr = ALLOC(double, Nc); o->cs = ALLOC(double, Nc); o->cw = ALLOC(double, Nc);
  for (int k = 0; k < Nc; k++) {
    const double *t = apl_tab_cub_rsw + apl_tab_cub_off[N - 1] + 3 * k;   /* cub2D_<cubOrd> */
    o->cr[k] = t[0]; o->cs[k] = t[1]; o->cw[k] = t[2];
  }
  o->Vc = ALLOC(double, Nc * Np); o->Drc = ALLOC(double, Nc * Np); o->Dsc = ALLOC(double, Nc * Np);
  o->VcT = ALLOC(double, Nc * Np); o->DrcT = ALLOC(double, Nc * Np); o->DscT = ALLOC(double, Nc * Np);
  {
    double *Vout = ALLOC(double, Nc * Np), *Vr = ALLOC(double, Nc * Np), *Vs = ALLOC(double, Nc * Np);
    vandermonde2D(N, Nc, o->cr, o->cs, Vout);        /* :80-82 */
    matmul(Nc, Np, Np, Vout, o->invV, o->Vc);
    grad_vandermonde2D(N, Nc, o->cr, o->cs, Vr, Vs); /* :93 DifferentiationMatrices2D */
    matmul(Nc, Np, Np, Vr, o->invV, o->Drc);
    matmul(Nc, Np, Np, Vs, o->invV, o->Dsc);
    transpose(Nc, Np, o->Vc, o->VcT);
    transpose(Nc, Np, o->Drc, o->DrcT);
    transpose(Nc, Np, o->Dsc, o->DscT);
    free(Vout); free(Vr); free(Vs);
  }
  /* face Gauss points (wxcubature2d.cc:110-133) */
  o->gz = ALLOC(double, Ng); o->gw = ALLOC(double, Ng);
  orc_gauleg(-1.0, 1.0, o->gz, o->gw, Ng);
  {
    double *fr = ALLOC(double, 3 * Ng), *fs = ALLOC(double, 3 * Ng), *gV = ALLOC(double, 3 * Ng * Np);
    for (int k = 0; k < Ng; k++) {
      fr[0 * Ng + k] = o->gz[k];  fs[0 * Ng + k] = -1.0;
      fr[1 * Ng + k] = -o->gz[k]; fs[1 * Ng + k] = o->gz[k];
      fr[2 * Ng + k] = -1.0;      fs[2 * Ng + k] = -o->gz[k];
    }
    vandermonde2D(N, 3 * Ng, fr, fs, gV);
    o->If = ALLOC(double, 3 * Ng * Np); o->IfT = ALLOC(double, 3 * Ng * Np);
    matmul(3 * Ng, Np, Np, gV, o->invV, o->If);
    transpose(3 * Ng, Np, o->If, o->IfT);
    free(fr); free(fs); free(gV);
  }
  return 0;
}

void orc_free_ops(orc_ops *o) {
  free(o->r); free(o->s); free(o->fmask); free(o->V); free(o->invV); free(o->VVT);
  free(o->Dr); free(o->Ds); free(o->cr); free(o->cs); free(o->cw);
  free(o->Vc); free(o->Drc); free(o->Dsc); free(o->VcT); free(o->DrcT); free(o->DscT);
  free(o->gz); free(o->gw); free(o->If); free(o->IfT);
  memset(o, 0, sizeof(*o));
}

/* ------------------------------------------------------------------------ */
/* Connectivity: FacePair2d (wxnodaldggeometry2d.cc:235-367)                */
/* ------------------------------------------------------------------------ */

typedef struct { long long key; int face; } face_key;

static int cmp_face_key(const void *a, const void *b) {
  const face_key *x = (const face_key *)a, *y = (const face_key *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->face - y->face;
}

/* Faces vn = {{0,1},{1,2},{2,0}} (:263).  FtoF = FtoV FtoV^T - 2I has a 2
 * exactly where two faces share both vertices (:279-323); here the same
 * pairs come from sorting faces by their (sorted) vertex-id pair.  Boundary
 * labels (:348-356) are applied first to all six entries of the cell that
 * supports the face, in face-creation order (DMPlexInterpolate numbers faces
 * in order of first appearance over cells, local faces in cone order -- an
 * assumption about PETSc, parity unpinned), then interior pairs overwrite
 * their two entries (:359-363). */
int orc_facepair(int K, int Nv, const int *EToV, const int *vert_id,
                 const int *face_tag, int *etetf) {
  static const int vn[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  int F = 3 * K;
  face_key *keys = ALLOC(face_key, F);
  for (int e = 0; e < K; e++)
    for (int f = 0; f < 3; f++) {
      int a = EToV[3 * e + vn[f][0]], b = EToV[3 * e + vn[f][1]];
      if (vert_id) { a = vert_id[a]; b = vert_id[b]; }
      int lo = a < b ? a : b, hi = a < b ? b : a;
      keys[3 * e + f].key = (long long)lo * (long long)(Nv + 1) + hi;
      keys[3 * e + f].face = 3 * e + f;
    }
  qsort(keys, F, sizeof(face_key), cmp_face_key);
  int *partner = ALLOC(int, F);
  for (int i = 0; i < F; i++) partner[i] = -1;
  int err = 0;
  for (int i = 0; i < F;) {
    int j = i;
    while (j < F && keys[j].key == keys[i].key) j++;
    if (j - i == 2) {
      partner[keys[i].face] = keys[i + 1].face;
      partner[keys[i + 1].face] = keys[i].face;
    } else if (j - i > 2) {
      err = 5;
    }
    i = j;
  }
  for (int i = 0; i < 6 * K; i++) etetf[i] = -1;            /* :336-340 */
  if (face_tag) {
    for (int e = 0; e < K; e++)
      for (int f = 0; f < 3; f++) {
        int fid = 3 * e + f, p = partner[fid];
        /* only the first appearance (lowest cell) creates the face; its
         * support's first cell is that lowest cell */
        if (p >= 0 && p < fid) continue;
        int tag = face_tag[fid];
        if (p >= 0 && tag == 0) tag = face_tag[p];
        if (tag != 0)
          for (int k = 0; k < 6; k++) etetf[6 * e + k] = -tag;
      }
  }
  for (int fid = 0; fid < F; fid++) {                        /* :359-363 */
    int p = partner[fid];
    if (p < 0) continue;
    etetf[6 * (fid / 3) + 2 * (fid % 3)] = p / 3;
    etetf[6 * (fid / 3) + 2 * (fid % 3) + 1] = p % 3;
  }
  free(keys);
  free(partner);
  return err;
}

/* CalculateNodeCoordinates2d (wxnodaldggeometry2d.cc:179-233). */
void orc_node_coords(const orc_ops *o, int K, const double *VX, const double *VY,
                     const int *EToV, double *x, double *y, double *dtscale) {
  double dts = 1.0e6;
  int Np = o->Np;
  for (int k = 0; k < K; k++) {
    double p1x = VX[EToV[3 * k]], p1y = VY[EToV[3 * k]];
    double p2x = VX[EToV[3 * k + 1]], p2y = VY[EToV[3 * k + 1]];
    double p3x = VX[EToV[3 * k + 2]], p3y = VY[EToV[3 * k + 2]];
    for (int n = 0; n < Np; n++) {
      double r = o->r[n], s = o->s[n];
      x[k * Np + n] = 0.5 * (-p1x * (r + s) + p2x * (1. + r) + p3x * (1. + s));
      y[k * Np + n] = 0.5 * (-p1y * (r + s) + p2y * (1. + r) + p3y * (1. + s));
    }
    double len1 = sqrt(pow(p1x - p2x, 2) + pow(p1y - p2y, 2));
    double len2 = sqrt(pow(p2x - p3x, 2) + pow(p2y - p3y, 2));
    double len3 = sqrt(pow(p3x - p1x, 2) + pow(p3y - p1y, 2));
    double sper = 0.5 * (len1 + len2 + len3);
    double Area = sqrt(fabs(sper * (sper - len1) * (sper - len2) * (sper - len3)));
    double v = Area / sper;
    dts = dts < v ? dts : v;                                   /* dmin */
  }
  *dtscale = dts;
}

/* GeometricFactors2d (:412-442) and Normals2d (:382-410). */
int orc_elem_geom(const orc_ops *o, const double *x, const double *y, double *g) {
  int Nfp = o->Nfp;
  const int *Fm = o->fmask;
  double x1 = x[Fm[0 * Nfp]], y1 = y[Fm[0 * Nfp]];
  double x2 = x[Fm[1 * Nfp]], y2 = y[Fm[1 * Nfp]];
  double x3 = x[Fm[2 * Nfp]], y3 = y[Fm[2 * Nfp]];
  double dxdr = (x2 - x1) / 2, dxds = (x3 - x1) / 2;
  double dydr = (y2 - y1) / 2, dyds = (y3 - y1) / 2;
  double J = -dxds * dydr + dxdr * dyds;
  if (J <= 0) return 3;
  g[0] = dyds / J;
  g[1] = -dydr / J;
  g[2] = -dxds / J;
  g[3] = dxdr / J;
  g[4] = J;
  double nr[9];
  nr[0] = (y2 - y1); nr[1] = -(x2 - x1);
  nr[3] = (y3 - y2); nr[4] = -(x3 - x2);
  nr[6] = (y1 - y3); nr[7] = -(x1 - x3);
  for (int f = 0; f < 3; ++f) {
    double sJ = sqrt(nr[3 * f] * nr[3 * f] + nr[3 * f + 1] * nr[3 * f + 1]);
    if (sJ <= 0) return 4;
    nr[3 * f] /= sJ;
    nr[3 * f + 1] /= sJ;
    nr[3 * f + 2] = sJ / 2.;
  }
  memcpy(g + 5, nr, sizeof(nr));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Physics                                                                   */
/* ------------------------------------------------------------------------ */

int orc_meqn(const orc_phys *ph) { return ph->eqn == ORC_EULER ? 5 : 8; }

static void euler_dirs(int d, int *mu, int *mv, int *mw) {
  if (d == 0) { *mu = 1; *mv = 2; *mw = 3; }
  else if (d == 1) { *mu = 2; *mv = 3; *mw = 1; }
  else { *mu = 3; *mv = 1; *mw = 2; }
}

static void maxwell_dirs(int d, int *ix) {
  /* iex, iey, iez, ibx, iby, ibz (wxphmaxwelleqn.cc:454-485) */
  static const int t[3][6] = {{0, 1, 2, 3, 4, 5}, {1, 2, 0, 4, 5, 3}, {2, 0, 1, 5, 3, 4}};
  for (int i = 0; i < 6; i++) ix[i] = t[d][i];
}

/* WxEulerEqn::flux (wxeulereqn.cc:806-871) / WxPHMaxwellEqn::flux
 * (wxphmaxwelleqn.cc:447-498). */
int orc_flux(const orc_phys *ph, int d, const double *q, double *f) {
  if (ph->eqn == ORC_EULER) {
    double rho, u, v, w, E, p;
    double gas_gamma1 = ph->gamma - 1.;
    int mu, mv, mw;
    euler_dirs(d, &mu, &mv, &mw);
    rho = q[0];
    if (ph->min_dens == 0.0 && rho <= 0) return 8;
    if (rho < ph->min_dens) rho = ph->min_dens;
    u = q[mu] / rho;
    v = q[mv] / rho;
    w = q[mw] / rho;
    E = q[4];
    p = gas_gamma1 * (E - 0.5 * rho * (u * u + v * v + w * w));
    if (p <= 0) {
      if (ph->min_pres == 0.0) return 9;
      p = ph->min_pres;
    }
    f[0] = rho * u;
    f[mu] = rho * u * u + p;
    f[mv] = rho * u * v;
    f[mw] = rho * u * w;
    f[4] = (E + p) * u;
    return 0;
  }
  int ix[6];
  maxwell_dirs(d, ix);
  double c0 = ph->c0, chi = ph->chi, g = ph->gph;
  f[ix[0]] = chi * c0 * c0 * q[6];
  f[ix[1]] = c0 * c0 * q[ix[5]];
  f[ix[2]] = -c0 * c0 * q[ix[4]];
  f[ix[3]] = g * q[7];
  f[ix[4]] = -q[ix[2]];
  f[ix[5]] = q[ix[1]];
  f[6] = chi * q[ix[0]];
  f[7] = g * c0 * c0 * q[ix[3]];
  return 0;
}

/* primitiveVariables (wxeulereqn.cc:1324-1341), including its in-place
 * floor of the conserved state. */
static void euler_primitive(const orc_phys *ph, double *qc, double *qp) {
  qp[0] = qc[0];
  if (qp[0] <= 0.) { qc[0] = ph->min_dens; qp[0] = ph->min_dens; }
  qp[1] = qc[1] / qc[0];
  qp[2] = qc[2] / qc[0];
  qp[3] = qc[3] / qc[0];
  qp[4] = (ph->gamma - 1.) * (qc[4] - 0.5 * (qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]) / qc[0]);
  if (qp[4] <= 0.) {
    qp[4] = ph->min_pres;
    qc[4] = ph->min_pres / (ph->gamma - 1.) + 0.5 * qp[0] * (qp[1] * qp[1] + qp[2] * qp[2] + qp[3] * qp[3]);
  }
}

static double dmax2(double a, double b) { return a > b ? a : b; }          /* wxmath.h:15 */
static double dmax3(double a, double b, double c) { double m1 = a > b ? a : b; return m1 > c ? m1 : c; }

static double dmin2(double a, double b) { return a < b ? a : b; }          /* wxmath.h:28 */

/* Momentum rotated in place to (normal, tangential) components, as
 * applyHLLFluxes / applyRoeFluxes do to their arguments (wxeulereqn.cc:
 * 1136-1143, 1214-1221); the RHS passes copies (wxnodaldg2dmethod.cc:450-452). */
static void rotate_momentum(const double *n, double *q) {
  double ru = q[1], rv = q[2];
  q[1] = n[0] * ru + n[1] * rv;
  q[2] = -n[1] * ru + n[0] * rv;
}

/* Roe averages of the rotated primitive states (wxeulereqn.cc:1160-1170,
 * 1232-1243): out = u, v, w, H, c2, rho. */
static void roe_average(const orc_phys *ph, const double *qM, const double *pM, const double *qP,
                        const double *pP, double *out) {
  double HM = (qM[4] + pM[4]) / pM[0], HP = (qP[4] + pP[4]) / pP[0];
  double rhoMs = sqrt(pM[0]), rhoPs = sqrt(pP[0]);
  double u = (rhoMs * pM[1] + rhoPs * pP[1]) / (rhoMs + rhoPs);
  double v = (rhoMs * pM[2] + rhoPs * pP[2]) / (rhoMs + rhoPs);
  double w = (rhoMs * pM[3] + rhoPs * pP[3]) / (rhoMs + rhoPs);
  double H = (rhoMs * HM + rhoPs * HP) / (rhoMs + rhoPs);
  out[0] = u;
  out[1] = v;
  out[2] = w;
  out[3] = H;
  out[4] = (ph->gamma - 1.) * (H - 0.5 * (u * u + v * v + w * w));
  out[5] = rhoMs * rhoPs;
}

/* WxEulerEqn::applyHLLFluxes (wxeulereqn.cc:1119-1202).  `abs(SR)` at :1183
 * is taken as the floating-point absolute value. */
static int euler_hll(const orc_phys *ph, const double *n, double *qM, double *qP, double *nflux,
                     double *lambda) {
  int rc;
  double fM[5], fP[5], gM[5], gP[5], pM[5], pP[5], avg[6], fx[5];
  rotate_momentum(n, qM);
  rotate_momentum(n, qP);
  if ((rc = orc_flux(ph, 0, qM, fM)) || (rc = orc_flux(ph, 0, qP, fP)) ||
      (rc = orc_flux(ph, 1, qM, gM)) || (rc = orc_flux(ph, 1, qP, gP)))
    return rc;
  euler_primitive(ph, qM, pM);
  euler_primitive(ph, qP, pP);
  double cM = sqrt(ph->gamma * pM[4] / pM[0]), cP = sqrt(ph->gamma * pP[4] / pP[0]);
  roe_average(ph, qM, pM, qP, pP, avg);
  double u = avg[0], c = sqrt(avg[4]);
  double SL = dmin2(pM[1] - cM, u - c), SR = dmax2(pP[1] + cP, u + c);
  double t1 = (dmin2(SR, 0.0) - dmin2(0.0, SL)) / (SR - SL);
  double t2 = 1. - t1;
  double t3 = (SR * fabs(SL) - SL * fabs(SR)) / (2. * (SR - SL));
  for (int c5 = 0; c5 < 5; c5++) fx[c5] = t1 * fP[c5] + t2 * fM[c5] - t3 * (qP[c5] - qM[c5]);
  nflux[0] = fx[0];
  nflux[1] = n[0] * fx[1] - n[1] * fx[2];
  nflux[2] = n[1] * fx[1] + n[0] * fx[2];
  nflux[3] = 0.0;
  nflux[4] = fx[4];
  double c0M = sqrt(pM[1] * pM[1] + pM[2] * pM[2] + pM[3] * pM[3]) + cM;
  double c0P = sqrt(pP[1] * pP[1] + pP[2] * pP[2] + pP[3] * pP[3]) + cP;
  *lambda = dmax2(c0M, c0P);
  return 0;
}

/* WxEulerEqn::applyRoeFluxes (wxeulereqn.cc:1204-1268), as written: the
 * back-rotation assigns nflux[1] twice (:1259-1260), so the x-momentum slot
 * receives the y-momentum flux n_y fx1 + n_x fx2, and nflux[2] is never
 * written.  Apollo passes a buffer allocated zero-filled at setup
 * (wxnodaldg2dmethod.cc:98, wxmath.h:49-55) that only this call writes, so
 * the y-momentum slot stays 0. */
static int euler_roe(const orc_phys *ph, const double *n, double *qM, double *qP, double *nflux,
                     double *lambda) {
  int rc;
  double fM[5], fP[5], pM[5], pP[5], avg[6], fx[5];
  rotate_momentum(n, qM);
  rotate_momentum(n, qP);
  euler_primitive(ph, qM, pM);
  euler_primitive(ph, qP, pP);
  roe_average(ph, qM, pM, qP, pP, avg);
  double u = avg[0], v = avg[1], H = avg[3], c2 = avg[4], c = sqrt(c2), rho = avg[5];
  double dw1 = -0.5 * rho * (pP[1] - pM[1]) / c + 0.5 * (pP[4] - pM[4]) / c2;
  double dw2 = (pP[0] - pM[0]) - (pP[4] - pM[4]) / c2;
  double dw3 = rho * (pP[2] - pM[2]);
  double dw4 = 0.5 * rho * (pP[1] - pM[1]) / c + 0.5 * (pP[4] - pM[4]) / c2;
  dw1 = fabs(u - c) * dw1;
  dw2 = fabs(u) * dw2;
  dw3 = fabs(u) * dw3;
  dw4 = fabs(u + c) * dw4;
  if ((rc = orc_flux(ph, 0, qM, fM)) || (rc = orc_flux(ph, 0, qP, fP))) return rc;
  for (int c5 = 0; c5 < 5; c5++) fx[c5] = 0.5 * (fM[c5] + fP[c5]);
  fx[0] -= 0.5 * (dw1 * 1. + dw2 * 1. + dw3 * 0. + dw4 * 1.);
  fx[1] -= 0.5 * (dw1 * (u - c) + dw2 * u + dw3 * 0. + dw4 * (u + c));
  fx[2] -= 0.5 * (dw1 * v + dw2 * v + dw3 * 1. + dw4 * v);
  fx[3] = 0.0;
  fx[4] -= 0.5 * (dw1 * (H - u * c) + dw2 * (u * u + v * v) / 2. + dw3 * v + dw4 * (H + u * c));
  nflux[0] = fx[0];
  nflux[1] = n[1] * fx[1] + n[0] * fx[2];
  nflux[2] = 0.0;
  nflux[3] = 0.0;
  nflux[4] = fx[4];
  double c0M = sqrt(pM[1] * pM[1] + pM[2] * pM[2] + pM[3] * pM[3]) + sqrt(ph->gamma * pM[4] / pM[0]);
  double c0P = sqrt(pP[1] * pP[1] + pP[2] * pP[2] + pP[3] * pP[3]) + sqrt(ph->gamma * pP[4] / pP[0]);
  *lambda = dmax2(c0M, c0P);
  return 0;
}

int orc_numflux(const orc_phys *ph, const double *n, double *qM, double *qP,
                double *nflux, double *lambda_out) {
  int m = orc_meqn(ph), rc;
  if (ph->eqn == ORC_EULER && ph->flux == ORC_FLUX_HLL) return euler_hll(ph, n, qM, qP, nflux, lambda_out);
  if (ph->eqn == ORC_EULER && ph->flux == ORC_FLUX_ROE) return euler_roe(ph, n, qM, qP, nflux, lambda_out);
  if (ph->eqn == ORC_EULER && ph->flux == ORC_FLUX_WAVE) {
    /* "Wave": applyWavePropagationFluxes rotates both traces' momentum in
     * place (wxeulereqn.cc:1292-1300); its result is then overwritten, since
     * case 4 of DGnumericalFlux has no break and falls through to LF
     * (:888-892), which runs on the rotated traces. */
    rotate_momentum(n, qM);
    rotate_momentum(n, qP);
  }
  double fM[8], fP[8], gM[8], gP[8];
  /* flux evaluations first, then primitives (wxeulereqn.cc:1097-1104) */
  if ((rc = orc_flux(ph, 0, qM, fM))) return rc;
  if ((rc = orc_flux(ph, 0, qP, fP))) return rc;
  if ((rc = orc_flux(ph, 1, qM, gM))) return rc;
  if ((rc = orc_flux(ph, 1, qP, gP))) return rc;
  double lambda;
  if (ph->eqn == ORC_EULER) {
    double pM[5], pP[5];
    euler_primitive(ph, qM, pM);
    euler_primitive(ph, qP, pP);
    double c0M = sqrt(pM[1] * pM[1] + pM[2] * pM[2] + pM[3] * pM[3]) + sqrt(ph->gamma * pM[4] / pM[0]);
    double c0P = sqrt(pP[1] * pP[1] + pP[2] * pP[2] + pP[3] * pP[3]) + sqrt(ph->gamma * pP[4] / pP[0]);
    lambda = dmax2(c0M, c0P);
  } else {
    lambda = dmax3(ph->chi * ph->c0, ph->gph * ph->c0, ph->c0);   /* :513 */
  }
  for (int c = 0; c < m; c++)
    nflux[c] = 0.5 * (n[0] * (fM[c] + fP[c]) + n[1] * (gM[c] + gP[c]) + lambda * (qM[c] - qP[c]));
  if (ph->eqn == ORC_PHMAXWELL && ph->flux == ORC_FLUX_UPWIND) {
    /* build extension: exact upwind flux of the TM fields (chi = gamma = 0):
     * LF minus the dissipation of the normal B jump. */
    double jb = n[0] * (qM[3] - qP[3]) + n[1] * (qM[4] - qP[4]);
    nflux[3] -= 0.5 * lambda * n[0] * jb;
    nflux[4] -= 0.5 * lambda * n[1] * jb;
  }
  *lambda_out = lambda;
  return 0;
}

/* Boundary states: WxIsentropicVortexBC::applyBC (wxisentropicvortexbc.cc:20-43),
 * WxMaxwellTransverseMagnetic::applyBC (wxmaxwelltranversemagneticbc.cc:13-25),
 * WxPHMaxwellConductingWallBC::applyBC (wxphmaxwellconductingwallbc.cc:12-41). */
void orc_bc(const orc_phys *ph, int bcnum, const double *xc, const double *nx,
            const double *q, double *qBC) {
  int m = orc_meqn(ph);
  int idx = bcnum - 1;
  int kind = (idx >= 0 && idx < ph->nbc) ? ph->bc_kind[idx] : ORC_BC_NONE;
  const double *P = (idx >= 0 && idx < ph->nbc) ? ph->bc_p[idx] : NULL;
  switch (kind) {
    case ORC_BC_VORTEX: {
      double uo = P[0], vo = P[1], xo = P[2], yo = P[3], gamma = P[4], beta = P[5];
      double t = xc[0], x = xc[1], y = xc[2];
      double pi = 3.141592653589793;
      double xmut = x - uo * t;
      double ymvt = y - vo * t;
      double r = sqrt((xmut - xo) * (xmut - xo) + (ymvt - yo) * (ymvt - yo));
      double u = uo - beta * exp(1. - r * r) * (ymvt - yo) / (2. * pi);
      double v = vo + beta * exp(1. - r * r) * (xmut - xo) / (2. * pi);
      double base = 1. - ((gamma - 1.) * beta * beta * exp(2. * (1. - r * r)) / (16. * gamma * pi * pi));
      double rho1 = pow(base, 1. / (gamma - 1));
      double p1 = pow(rho1, gamma);
      qBC[0] = rho1;
      qBC[1] = rho1 * u;
      qBC[2] = rho1 * v;
      qBC[3] = 0.0;
      qBC[4] = p1 / (gamma - 1) + 0.5 * rho1 * (u * u + v * v);
      break;
    }
    case ORC_BC_TM:
      for (int c = 0; c < m; c++) qBC[c] = q[c];
      qBC[2] = -q[2];
      break;
    case ORC_BC_PEC: {
      double ex = q[0], ey = q[1], ez = q[2], bx = q[3], by = q[4], bz = q[5], phi = q[6], psi = q[7];
      double enorm = ex * nx[0] + ey * nx[1];
      double etang = ex * nx[1] - ey * nx[0];
      qBC[0] = enorm * nx[0] - etang * nx[1];
      qBC[1] = enorm * nx[1] + etang * nx[0];
      qBC[2] = -ez;
      double bnorm = -bx * nx[0] - by * nx[1];
      double btang = -bx * nx[1] + by * nx[0];
      qBC[3] = bnorm * nx[0] - btang * nx[1];
      qBC[4] = bnorm * nx[1] + btang * nx[0];
      qBC[5] = bz;
      qBC[6] = -phi;
      qBC[7] = psi;
      break;
    }
    case ORC_BC_EULER_WALL: {                 /* wxeulerwallbc.cc:18-25 */
      double qn = nx[0] * q[1] + nx[1] * q[2];
      qBC[0] = q[0];
      qBC[1] = q[1] - 2. * qn * nx[0];
      qBC[2] = q[2] - 2. * qn * nx[1];
      qBC[3] = 0.0;
      qBC[4] = q[4];
      break;
    }
    case ORC_BC_EULER_INFLOW: {               /* wxeulerinflowbc.cc:21-36 */
      double mach = P[0], angle = P[1], gamma = P[2], pres = P[3], rho = P[4];
      double pi = 3.141592653589793;
      double sound = sqrt(gamma * pres / rho);
      double u = mach * sound * cos(angle * pi / 180.);
      double v = mach * sound * sin(angle * pi / 180.);
      qBC[0] = rho;
      qBC[1] = rho * u;
      qBC[2] = rho * v;
      qBC[3] = 0.0;
      qBC[4] = pres / (gamma - 1) + 0.5 * rho * (u * u + v * v);
      break;
    }
    case ORC_BC_EULER_DIRICHLET: {            /* wxeulerdirichletbc.cc:22-30 */
      double rho = P[0], u = P[1], v = P[2], w = P[3], gamma = P[4], pres = P[5];
      qBC[0] = rho;
      qBC[1] = rho * u;
      qBC[2] = rho * v;
      qBC[3] = rho * w;
      qBC[4] = pres / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w);
      break;
    }
    default: /* copy: eulerZeroGradientBC (wxeulerzerogradientbc.cc:18-25) */
      for (int c = 0; c < m; c++) qBC[c] = q[c];
  }
}

/* Geometric factors at cubature points: Apollo recomputes them from Dr_c x
 * and Ds_c x (wxcubature2d.cc:305,327-356), which for straight-sided
 * triangles equals the element's affine factors up to rounding.  With
 * exact_geom set the affine factors are used instead (a diagnostic that
 * isolates that rounding; off by default). */
static int g_exact_geom = 0;
void orc_set_exact_geometry(int on) { g_exact_geom = on; }

static void cub_geom(const orc_ops *o, const double *xcoord, const double *ycoord,
                     const double *geo, double *xr, double *xs, double *yr, double *ys);

/* With reverse_sums set, every matrix-vector product sums its terms in the
 * reverse order: another FP64 evaluation of the same algorithm, which
 * measures the RHS's rounding floor (a diagnostic; off by default). */
static int g_reverse_sums = 0;
void orc_set_reverse_sums(int on) { g_reverse_sums = on; }

/* MatrixVectorMult (wxcubature2d.cc:431-442). */
static void mvm(int rows, int cols, int meqn, const double *A, const double *x, double *y) {
  for (int k = 0; k < rows * meqn; k++) y[k] = 0.0;
  if (g_reverse_sums) {
    for (int kx = 0; kx < rows; kx++)
      for (int ky = cols - 1; ky >= 0; ky--)
        for (int kz = 0; kz < meqn; kz++) y[kx * meqn + kz] += A[kx * cols + ky] * x[ky * meqn + kz];
    return;
  }
  for (int kx = 0; kx < rows; kx++)
    for (int ky = 0; ky < cols; ky++)
      for (int kz = 0; kz < meqn; kz++) y[kx * meqn + kz] += A[kx * cols + ky] * x[ky * meqn + kz];
}

static void cub_geom(const orc_ops *o, const double *xcoord, const double *ycoord,
                     const double *geo, double *xr, double *xs, double *yr, double *ys) {
  const int Nc = o->Nc, Np = o->Np;
  if (!g_exact_geom) {
    mvm(Nc, Np, 1, o->Drc, xcoord, xr);
    mvm(Nc, Np, 1, o->Drc, ycoord, yr);
    mvm(Nc, Np, 1, o->Dsc, xcoord, xs);
    mvm(Nc, Np, 1, o->Dsc, ycoord, ys);
    return;
  }
  /* invert rx = ys/J, sx = -yr/J, ry = -xs/J, sy = xr/J */
  const double J = geo[4];
  for (int c = 0; c < Nc; c++) {
    xr[c] = geo[3] * J;
    xs[c] = -geo[2] * J;
    yr[c] = -geo[1] * J;
    ys[c] = geo[0] * J;
  }
}

/* ------------------------------------------------------------------------ */
/* RHS: WxNodalDG2dMethod::step (wxnodaldg2dmethod.cc:235-542)              */
/* ------------------------------------------------------------------------ */

static int rhs_element(const orc_ops *o, const orc_phys *ph, int kelem,
                       const double *xall, const double *yall, const int *etetf,
                       double t, const double *u, double *out, double *maxSpeed) {
  const int Np = o->Np, Nc = o->Nc, Ng = o->Ng, m = orc_meqn(ph);
  double xc[5];
  xc[0] = t;
  double nx[2];
  const double *xcoord = xall + (size_t)kelem * Np, *ycoord = yall + (size_t)kelem * Np;
  double q_vol[45 * 8], vec_rhs[45 * 8], volInt[45 * 8], q_surf[45 * 8], qtemp[45 * 8];
  double Iq_vol[145 * 8], If_vol[145 * 8], Ig_vol[145 * 8];
  double QP[54 * 8], QM[54 * 8], numFlux[54 * 8], Xcrd[54], Ycrd[54], qgtemp[54 * 8];
  double Fflux[8], Gflux[8], Qvar[8], qM[8], qP[8], nfl[8];
  int rc;
  memcpy(q_vol, u + (size_t)kelem * Np * m, sizeof(double) * Np * m);   /* :330-332 */
  double geo[14];
  if ((rc = orc_elem_geom(o, xcoord, ycoord, geo))) return rc;         /* :336-337 */
  const double *normals = geo + 5;
  /* volume (:347-377) */
  mvm(Nc, Np, m, o->Vc, q_vol, Iq_vol);
  for (int pt = 0; pt < Nc; pt++) {
    for (int c = 0; c < m; c++) Qvar[c] = Iq_vol[pt * m + c];
    if ((rc = orc_flux(ph, 0, Qvar, Fflux))) return rc;
    if ((rc = orc_flux(ph, 1, Qvar, Gflux))) return rc;
    for (int c = 0; c < m; c++) {
      If_vol[pt * m + c] = Fflux[c];
      Ig_vol[pt * m + c] = Gflux[c];
    }
  }
  {
    /* evaluatedVolumeIntegrals + geometricFactors2D (wxcubature2d.cc:295-356);
     * the source term is identically zero for these configurations and its
     * V_c^T (wJ * 0) product is skipped. */
    double xr[145], xs[145], yr[145], ys[145];
    double RR[145 * 8], SS[145 * 8], ddx[45 * 8], ddy[45 * 8];
    cub_geom(o, xcoord, ycoord, geo, xr, xs, yr, ys);
    for (int kk = 0; kk < Nc; kk++) {
      double J = xr[kk] * ys[kk] - xs[kk] * yr[kk];
      if (J <= 0) return 3;
      double rx = ys[kk] / J, sx = -yr[kk] / J, ry = -xs[kk] / J, sy = xr[kk] / J;
      for (int mm = 0; mm < m; mm++) {
        RR[kk * m + mm] = o->cw[kk] * J * (rx * If_vol[kk * m + mm] + ry * Ig_vol[kk * m + mm]);
        SS[kk * m + mm] = o->cw[kk] * J * (sx * If_vol[kk * m + mm] + sy * Ig_vol[kk * m + mm]);
      }
    }
    mvm(Np, Nc, m, o->DrcT, RR, ddx);
    mvm(Np, Nc, m, o->DscT, SS, ddy);
    for (int kk = 0; kk < Np * m; kk++) volInt[kk] = ddx[kk] + ddy[kk] + 0.0;
  }
  /* surface (:390-472) */
  mvm(3 * Ng, Np, m, o->If, q_vol, QM);
  mvm(3 * Ng, Np, 1, o->If, xcoord, Xcrd);
  mvm(3 * Ng, Np, 1, o->If, ycoord, Ycrd);
  const int *connect = etetf + 6 * (size_t)kelem;
  for (int edge = 0; edge < 3; edge++) {
    int edgeNum = connect[2 * edge + 1];
    if (edgeNum < 0) {
      for (int g = 0; g < Ng; g++) {
        xc[1] = Xcrd[edge * Ng + g];
        xc[2] = Ycrd[edge * Ng + g];
        nx[0] = normals[3 * edge];
        nx[1] = normals[3 * edge + 1];
        for (int c = 0; c < m; c++) qM[c] = QM[(edge * Ng + g) * m + c];
        orc_bc(ph, abs(edgeNum), xc, nx, qM, qP);
        for (int c = 0; c < m; c++) QP[(edge * Ng + g) * m + c] = qP[c];
      }
    } else {
      int plusElem = connect[2 * edge];
      memcpy(qtemp, u + (size_t)plusElem * Np * m, sizeof(double) * Np * m);
      mvm(3 * Ng, Np, m, o->If, qtemp, qgtemp);                     /* q7 */
      for (int g = 0; g < Ng; g++)
        for (int c = 0; c < m; c++)
          QP[(edge * Ng + g) * m + c] = qgtemp[(edgeNum * Ng + Ng - 1 - g) * m + c];
    }
  }
  for (int kk = 0; kk < Ng * 3; kk++) {
    int curEdge = kk / Ng;
    for (int c = 0; c < m; c++) { qM[c] = QM[kk * m + c]; qP[c] = QP[kk * m + c]; }
    double lambda;
    nx[0] = normals[3 * curEdge + 0];
    nx[1] = normals[3 * curEdge + 1];
    if ((rc = orc_numflux(ph, nx, qM, qP, nfl, &lambda))) return rc;
    lambda = sqrt(lambda * lambda);
    *maxSpeed = dmax2(*maxSpeed, lambda);
    for (int c = 0; c < m; c++) numFlux[kk * m + c] = nfl[c] * normals[3 * curEdge + 2];
  }
  /* calculateSurfaceIntegral (wxcubature2d.cc:365-382) */
  for (int kk = 0; kk < 3 * Ng; kk++)
    for (int mm = 0; mm < m; mm++) numFlux[kk * m + mm] *= o->gw[kk % Ng];
  mvm(Np, 3 * Ng, m, o->IfT, numFlux, q_surf);
  for (int kk = 0; kk < Np * m; kk++) {                             /* :475-489 */
    volInt[kk] -= q_surf[kk];
    if (q_surf[kk] != q_surf[kk]) return 7;
    if (volInt[kk] != volInt[kk]) return 7;
  }
  mvm(Np, Np, m, o->VVT, volInt, vec_rhs);                          /* :492 */
  double *rhs = out + (size_t)kelem * Np * m;
  for (int kne = 0; kne < Np * m; kne++) {
    rhs[kne] = vec_rhs[kne] / geo[4];
    if (vec_rhs[kne] != vec_rhs[kne]) return 7;
  }
  return 0;
}

int orc_rhs(const orc_ops *o, const orc_phys *ph, int K, const double *x,
            const double *y, const int *etetf, double t, const double *q,
            double *rhs, double *maxSpeed, int nthreads) {
  double ms = 0.0;
  int err = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(max : ms) reduction(max : err)
#endif
  for (int k = 0; k < K; k++) {
    double local = 0.0;
    int rc = rhs_element(o, ph, k, x, y, etetf, t, q, rhs, &local);
    if (rc > err) err = rc;
    if (local > ms) ms = local;
  }
  (void)nthreads;
  if (maxSpeed) *maxSpeed = ms;
  return err;
}

/* Initial dt (wxnodaldg2dmethod.cc:201-228): Roe speeds of the node state
 * against itself (wxeulereqn.cc:95-234). */
double orc_initial_dt(const orc_ops *o, const orc_phys *ph, int K, const double *q,
                      double cfl, double dtscale) {
  int m = orc_meqn(ph), Np = o->Np;
  double maxSpeed = 0.0;
  for (int k = 0; k < K; k++)
    for (int n = 0; n < Np; n++) {
      const double *d = q + ((size_t)k * Np + n) * m;
      double lam = 0.0;
      if (ph->eqn == ORC_EULER) {
        for (int dir = 0; dir < 2; dir++) {
          int mu, mv, mw;
          euler_dirs(dir, &mu, &mv, &mw);
          double g1 = ph->gamma - 1;
          double rs = sqrt(d[0]);
          double p = g1 * (d[4] - 0.5 * (d[mu] * d[mu] + d[mv] * d[mv] + d[mw] * d[mw]) / d[0]);
          double rs2 = rs + rs;
          double u = (d[mu] / rs + d[mu] / rs) / rs2;
          double v = (d[mv] / rs + d[mv] / rs) / rs2;
          double w = (d[mw] / rs + d[mw] / rs) / rs2;
          double u2 = u * u + v * v + w * w;
          double enth = ((d[4] + p) / rs + (d[4] + p) / rs) / rs2;
          double a = sqrt(g1 * (enth - 0.5 * u2));
          double s[3] = {u - a, u, u + a};
          for (int i = 0; i < 3; i++) lam = dmax2(lam, s[i] * s[i]);
        }
      } else {
        double c = dmax3(ph->chi * ph->c0, ph->gph * ph->c0, ph->c0);
        lam = c * c;
      }
      lam = sqrt(lam);
      maxSpeed = dmax2(maxSpeed, lam);
    }
  return 2. / 3. * cfl * dtscale * (o->rmin / maxSpeed);
}

/* ------------------------------------------------------------------------ */
/* LSERK4 (build extension; Carpenter & Kennedy 1994 (5,4) 2N-storage)      */
/* ------------------------------------------------------------------------ */

const double orc_rk4a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                            -2404267990393.0 / 2016746695238.0,
                            -3550918686646.0 / 2091501179385.0,
                            -1275806237668.0 / 842570457699.0};
const double orc_rk4b[5] = {1432997174477.0 / 9575080441755.0,
                            5161836677717.0 / 13612068292357.0,
                            1720146321549.0 / 2090206949498.0,
                            3134564353537.0 / 4481467310338.0,
                            2277821191437.0 / 14882151754819.0};
const double orc_rk4c[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                            2526269341429.0 / 6820363962896.0,
                            2006345519317.0 / 3224310063776.0,
                            2802321613138.0 / 2924317926251.0};

int orc_lserk4(const orc_ops *o, const orc_phys *ph, int K, const double *x,
               const double *y, const int *etetf, double t0, double dt, int nsteps,
               double *q, int nthreads) {
  size_t n = (size_t)K * o->Np * orc_meqn(ph);
  double *res = ALLOC(double, n), *rhs = ALLOC(double, n);
  int rc = 0;
  for (int step = 0; step < nsteps && !rc; step++) {
    double t = t0 + step * dt;
    for (int s = 0; s < 5 && !rc; s++) {
      double ms;
      if (ph->limiter && (rc = orc_limit(o, ph, K, x, y, etetf, q, q))) break;
      rc = orc_rhs(o, ph, K, x, y, etetf, t + orc_rk4c[s] * dt, q, rhs, &ms, nthreads);
      for (size_t i = 0; i < n; i++) {
        res[i] = orc_rk4a[s] * res[i] + dt * rhs[i];
        q[i] = q[i] + orc_rk4b[s] * res[i];
      }
    }
  }
  free(res);
  free(rhs);
  return rc;
}

/* One output frame as ApSolver::solve advances it (src/solvers/apsolver.cc:
 * 172-216): PETSc TSSSP (src/subsolvers/wxpetsctimestepping.h:33) of type
 * rks2, the s-stage low-storage second-order SSP method (Ketcheson 2008;
 * PETSc's default type, 5 stages), started by TSSetInitialTimeStep(t0, dt0)
 * with exact final time MATCHSTEP (wxpetsctimestepping.h:92-96).  Per step:
 * MonitorVTK clamps dt to the frame end (apsolver.cc:421-440), MATCHSTEP
 * shortens a step that would pass t_end, and ComputeRHSforTS sets the next
 * dt to min(suggested, this step's dt) after every RHS (apsolver.cc:
 * 442-480), so the last stage's value stands.  The suggested dt is
 * (2/3) cfl dtscale rmin / maxSpeed (wxnodaldg2dmethod.cc:523). */
int orc_ssp_advance(const orc_ops *o, const orc_phys *ph, int K, const double *x,
                    const double *y, const int *etetf, double t0, double t_end, double dt0,
                    double cfl, double dtscale, int stages, double *q, int max_steps,
                    int *nsteps, double *dts, int nthreads) {
  size_t n = (size_t)K * o->Np * orc_meqn(ph);
  double *w = ALLOC(double, n), *F = ALLOC(double, n);
  double t = t0, dt_temp = dt0;
  int rc = 0, steps = 0;
  for (;;) {
    if (fabs(t_end - t) < dt_temp) dt_temp = fabs(t_end - t);
    if (t >= t_end) break;
    double dt = dt_temp;
    if (t + dt > t_end) dt = t_end - t;
    if (steps >= max_steps) { rc = 1; break; }
    double ms = 0.0;
    memcpy(w, q, n * sizeof(double));
    /* with a limiter, step() limits the vector TSSSP hands it in place
     * (wxnodaldg2dmethod.cc:243-247): the stage vector w, never q */
    for (int i = 0; i < stages - 1 && !rc; i++) {
      if (ph->limiter && (rc = orc_limit(o, ph, K, x, y, etetf, w, w))) break;
      rc = orc_rhs(o, ph, K, x, y, etetf, t + dt * (i / (stages - 1.)), w, F, &ms, nthreads);
      for (size_t k = 0; k < n; k++) w[k] = w[k] + dt / (stages - 1.) * F[k];
    }
    if (rc) break;
    if (ph->limiter && (rc = orc_limit(o, ph, K, x, y, etetf, w, w))) break;
    rc = orc_rhs(o, ph, K, x, y, etetf, t + dt, w, F, &ms, nthreads);
    if (rc) break;
    for (size_t k = 0; k < n; k++)
      q[k] = (stages - 1.) / stages * w[k] + dt / stages * F[k] + 1. / stages * q[k];
    double suggested = 2. / 3. * cfl * dtscale * (o->rmin / ms);
    dt_temp = fmin(suggested, dt);
    t = t + dt;
    if (dts) dts[steps] = dt;
    steps++;
  }
  if (nsteps) *nsteps = steps;
  free(w);
  free(F);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* The Hesthaven-Warburton Euler limiter                                      */
/* ------------------------------------------------------------------------ */

/* Limiter's primitive variables (wxhesthavenwarburtoneulerlimiter.cc:534-543):
 * no floors. */
static void lim_primitive(double g, const double *qc, double *qp) {
  qp[0] = qc[0];
  qp[1] = qc[1] / qc[0];
  qp[2] = qc[2] / qc[0];
  qp[3] = qc[3] / qc[0];
  qp[4] = (g - 1.) * (qc[4] - 0.5 * (qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]) / qc[0]);
}

/* Cell-average weights (wxnodaldggeometry2d.h:84-98): AVE[k] = 0.5 sum_m
 * Mass[m][k] with Mass = (V V^T)^-1 (wxnodaldggeometry2d.cc:71-76). */
int orc_limiter_weights(const orc_ops *o, double *ave) {
  int Np = o->Np;
  double *mass = ALLOC(double, Np * Np);
  int rc = orc_invert(Np, o->VVT, mass);
  for (int k = 0; k < Np; k++) {
    ave[k] = 0.0;
    for (int m = 0; m < Np; m++) ave[k] += 0.5 * mass[k + m * Np];
  }
  free(mass);
  return rc ? 2 : 0;
}

/* WxHestavenWarburtonEulerLimiter::applyLimiter
 * (wxhesthavenwarburtoneulerlimiter.cc:61-484), restated as written:
 *  - the patch centres and areas use AVE[0..2] with the vertex nodes
 *    Fmask[0], Fmask[Nfp], Fmask[2 Nfp] (:119-144);
 *  - the face-node states of a boundary face are read from the last
 *    neighbour read earlier in the face loop, or from the element itself
 *    when there was none (:200-214 read qIn, which :221 re-points);
 *  - the face-averaged primitives at the six face ends are read from the
 *    face-ordered array at the element-node ids of the ends (:241-265);
 *  - the third edge gradient uses (yv3 - yv2), (xv3 - xv2) (:301-305);
 *  - every node is overwritten by the reconstruction (:394-466).
 * The end-node ids index an array of 3 Nfp entries, so orders N > 4 read
 * outside it in the reference; they are refused (11), as is a boundary
 * face with the isentropic-vortex state, which reads three coordinates
 * from a two-entry array there (:186-192, wxisentropicvortexbc.cc:23-25).
 * qout may equal qin. */
int orc_limit(const orc_ops *o, const orc_phys *ph, int K, const double *x,
              const double *y, const int *etetf, const double *qin, double *qout) {
  enum { M = 5 };
  const int Np = o->Np, NpF = o->Nfp;
  const int *F = o->fmask;
  const double g = ph->lim_gamma;
  if (ph->eqn != ORC_EULER) return 1;
  if (o->N > 4) return 11;
  double *ave = ALLOC(double, Np), *drop = ALLOC(double, Np * Np);
  double *avc = ALLOC(double, (size_t)K * M), *avp = ALLOC(double, (size_t)K * M);
  double *grad = ALLOC(double, (size_t)K * 8 * M);   /* E1x,E1y,E2x,E2y,E3x,E3y,C0x,C0y */
  double *qo = ALLOC(double, (size_t)K * Np * M);
  double *QM = ALLOC(double, 3 * NpF * M), *QP = ALLOC(double, 3 * NpF * M);
  double *fc = ALLOC(double, 3 * NpF * M);
  int rc = orc_limiter_weights(o, ave);
  for (int i = 0; i < Np; i++)
    for (int j = 0; j < Np; j++) drop[i * Np + j] = -ave[j] + (i == j ? 1. : 0.);
  const int endIDs[6] = {F[0], F[NpF - 1], F[NpF], F[2 * NpF - 1], F[2 * NpF], F[3 * NpF - 1]};

  for (int e = 0; e < K && !rc; e++) {
    const int *cn = etetf + 6 * e;
    const double *xe = x + (size_t)e * Np, *ye = y + (size_t)e * Np;
    double geo[14], gn[14];
    if ((rc = orc_elem_geom(o, xe, ye, geo))) break;
    double xc[4] = {0, 0, 0, 0}, yc[4] = {0, 0, 0, 0}, A0[4] = {0, 0, 0, 0};
    for (int kk = 0; kk < 3; kk++) {                         /* :121-126 */
      xc[0] += ave[kk] * xe[F[kk * NpF]];
      yc[0] += ave[kk] * ye[F[kk * NpF]];
      A0[0] += 2. / 3. * ave[kk] * geo[4];
    }
    for (int f = 0; f < 3; f++) {                            /* :128-144 */
      const double *nrm = geo + 5 + 3 * f;
      if (cn[2 * f] < 0) {
        double H1 = 2.0 * A0[0] / nrm[2];
        xc[f + 1] = xc[0] + 2.0 * nrm[0] * H1;
        yc[f + 1] = yc[0] + 2.0 * nrm[1] * H1;
      } else {
        const int nb = cn[2 * f];
        const double *xn = x + (size_t)nb * Np, *yn = y + (size_t)nb * Np;
        if ((rc = orc_elem_geom(o, xn, yn, gn))) break;
        for (int node = 0; node < 3; node++) {
          xc[f + 1] += ave[node] * xn[F[node * NpF]];
          yc[f + 1] += ave[node] * yn[F[node * NpF]];
          A0[f + 1] += 2. / 3. * ave[node] * gn[4];
        }
      }
    }
    if (rc) break;
    for (int f = 0; f < 3; f++) A0[f + 1] += A0[0];

    /* step 2: patch averages (:149-233) */
    double cons[M][4], prim[M][4], qc[M], qp[M], qb[M];
    const double *qIn = qin + (size_t)e * Np * M;
    for (int c = 0; c < M; c++) {
      qc[c] = 0.0;
      for (int n = 0; n < Np; n++) qc[c] += ave[n] * qIn[n * M + c];
    }
    lim_primitive(g, qc, qp);
    for (int c = 0; c < M; c++) {
      cons[c][0] = qc[c];
      prim[c][0] = qp[c];
      avc[e * M + c] = qc[c];
      avp[e * M + c] = qp[c];
    }
    for (int f = 0; f < 3; f++)
      for (int n = 0; n < NpF; n++)
        for (int c = 0; c < M; c++) QM[(f * NpF + n) * M + c] = qIn[F[f * NpF + n] * M + c];
    for (int f = 0; f < 3 && !rc; f++) {
      const double *nrm = geo + 5 + 3 * f;
      if (cn[2 * f] < 0) {
        const int bcnum = abs(cn[2 * f + 1]);
        const int bi = bcnum - 1;
        if (bi >= 0 && bi < ph->nbc && ph->bc_kind[bi] == ORC_BC_VORTEX) { rc = 11; break; }
        double NX[2] = {nrm[0], nrm[1]}, XC[3] = {0, xc[f + 1], yc[f + 1]};
        for (int c = 0; c < M; c++) qc[c] = cons[c][0];
        orc_bc(ph, bcnum, XC, NX, qc, qb);
        lim_primitive(g, qb, qp);
        for (int c = 0; c < M; c++) {
          cons[c][f + 1] = qb[c];
          prim[c][f + 1] = qp[c];
        }
        for (int n = 0; n < NpF; n++) {
          const int node = F[f * NpF + n];
          double XN[3] = {0, xe[node], ye[node]};
          for (int c = 0; c < M; c++) qc[c] = qIn[node * M + c];
          orc_bc(ph, bcnum, XN, NX, qc, qb);
          for (int c = 0; c < M; c++) QP[(f * NpF + n) * M + c] = qb[c];
        }
      } else {
        qIn = qin + (size_t)cn[2 * f] * Np * M;
        for (int c = 0; c < M; c++) {
          qc[c] = 0.0;
          for (int n = 0; n < Np; n++) qc[c] += ave[n] * qIn[n * M + c];
        }
        lim_primitive(g, qc, qp);
        for (int c = 0; c < M; c++) {
          cons[c][f + 1] = qc[c];
          prim[c][f + 1] = qp[c];
        }
        const int fn = cn[2 * f + 1];
        for (int n = 0; n < NpF; n++)
          for (int c = 0; c < M; c++)
            QP[(f * NpF + n) * M + c] = qIn[F[fn * NpF + NpF - 1 - n] * M + c];
      }
    }
    if (rc) break;

    /* step 3: face averages and their primitives at the face ends (:236-266) */
    for (int k = 0; k < 3 * NpF * M; k++) fc[k] = 0.5 * (QP[k] + QM[k]);
    double fp[6][M];
    for (int n = 0; n < 6; n++) {
      const double *a = fc + endIDs[n] * M;
      fp[n][0] = a[0];
      fp[n][1] = a[1] / a[0];
      fp[n][2] = a[2] / a[0];
      fp[n][3] = a[3] / a[0];
      fp[n][4] = (g - 1.) * (a[4] - 0.5 * (a[1] * a[1] + a[2] * a[2] + a[3] * a[3]) / a[0]);
    }

    /* step 4: edge and cell gradients of the primitives (:269-314) */
    const double xv1 = xe[endIDs[0]], yv1 = ye[endIDs[0]];
    const double xv2 = xe[endIDs[2]], yv2 = ye[endIDs[2]];
    const double xv3 = xe[endIDs[4]], yv3 = ye[endIDs[4]];
    double *gr = grad + (size_t)e * 8 * M;
    for (int c = 0; c < M; c++) {
      double e1x = 0.5 * ((prim[c][1] - prim[c][0]) * (yv2 - yv1) + (fp[0][c] - fp[1][c]) * (yc[1] - yc[0])) / A0[1];
      double e1y = -0.5 * ((prim[c][1] - prim[c][0]) * (xv2 - xv1) + (fp[0][c] - fp[1][c]) * (xc[1] - xc[0])) / A0[1];
      double e2x = 0.5 * ((prim[c][2] - prim[c][0]) * (yv3 - yv2) + (fp[2][c] - fp[3][c]) * (yc[2] - yc[0])) / A0[2];
      double e2y = -0.5 * ((prim[c][2] - prim[c][0]) * (xv3 - xv2) + (fp[2][c] - fp[3][c]) * (xc[2] - xc[0])) / A0[2];
      double e3x = 0.5 * ((prim[c][3] - prim[c][0]) * (yv3 - yv2) + (fp[4][c] - fp[5][c]) * (yc[3] - yc[0])) / A0[3];
      double e3y = -0.5 * ((prim[c][3] - prim[c][0]) * (xv3 - xv2) + (fp[4][c] - fp[5][c]) * (xc[3] - xc[0])) / A0[3];
      gr[0 * M + c] = e1x;
      gr[1 * M + c] = e1y;
      gr[2 * M + c] = e2x;
      gr[3 * M + c] = e2y;
      gr[4 * M + c] = e3x;
      gr[5 * M + c] = e3y;
      gr[6 * M + c] = (A0[1] * e1x + A0[2] * e2x + A0[3] * e3x) / (A0[1] + A0[2] + A0[3]);
      gr[7 * M + c] = (A0[1] * e1y + A0[2] * e2y + A0[3] * e3y) / (A0[1] + A0[2] + A0[3]);
    }
  }

  /* second pass: limited gradients and reconstruction (:317-467) */
  for (int e = 0; e < K && !rc; e++) {
    const int *cn = etetf + 6 * e;
    const double *xe = x + (size_t)e * Np, *ye = y + (size_t)e * Np;
    double ldx[16], ldy[16], qdv[16 * M];
    for (int i = 0; i < Np; i++) {
      ldx[i] = 0.0;
      ldy[i] = 0.0;
    }
    for (int i = 0; i < Np; i++)
      for (int j = 0; j < Np; j++) {
        ldx[i] += drop[i * Np + j] * xe[j];
        ldy[i] += drop[i * Np + j] * ye[j];
      }
    for (int c = 0; c < M; c++) {
      double gx[3], gy[3];
      for (int f = 0; f < 3; f++) {
        if (cn[2 * f] < 0) {
          gx[f] = grad[((size_t)e * 8 + 2 * f) * M + c];
          gy[f] = grad[((size_t)e * 8 + 2 * f + 1) * M + c];
        } else {
          gx[f] = grad[((size_t)cn[2 * f] * 8 + 6) * M + c];
          gy[f] = grad[((size_t)cn[2 * f] * 8 + 7) * M + c];
        }
      }
      double g1 = gx[0] * gx[0] + gy[0] * gy[0];
      double g2 = gx[1] * gx[1] + gy[1] * gy[1];
      double g3 = gx[2] * gx[2] + gy[2] * gy[2];
      double fac = g1 * g1 + g2 * g2 + g3 * g3;
      double epse = 1.e-14;
      double fac3e = fac + 3.0 * epse;
      double w1 = (g2 * g3 + epse) / fac3e;
      double w2 = (g1 * g3 + epse) / fac3e;
      double w3 = (g1 * g2 + epse) / fac3e;
      double Lx = w1 * gx[0] + w2 * gx[1] + w3 * gx[2];
      double Ly = w1 * gy[0] + w2 * gy[1] + w3 * gy[2];
      for (int n = 0; n < Np; n++) qdv[n * M + c] = ldx[n] * Lx + ldy[n] * Ly;
    }
    const double *AP = avp + (size_t)e * M, *AC = avc + (size_t)e * M;
    double *out = qo + (size_t)e * Np * M;
    for (int n = 0; n < Np; n++) {
      double averho = AP[0], aveu = AP[1], avev = AP[2], avew = AP[3], avep = AP[4];
      double drho = qdv[n * M + 0], du = qdv[n * M + 1], dv = qdv[n * M + 2];
      double dw = qdv[n * M + 3], dp = qdv[n * M + 4];
      double Lrho = averho + drho, Lrhou, Lrhov, Lrhow, LEner, Lp;
      if (Lrho < ph->lim_min_dens) {
        for (int it = 0; it < 4; it++) {
          drho *= 0.5;
          Lrho = averho + drho;
          if (Lrho > ph->lim_min_dens) break;
          else if (it == 3) Lrho = AC[0];
        }
      }
      Lrhou = AC[1] + averho * du + drho * aveu;
      Lrhov = AC[2] + averho * dv + drho * avev;
      Lrhow = AC[3] + averho * dw + drho * avew;
      double dEner = dp / (g - 1.) + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
                     averho * (aveu * du + avev * dv + avew * dw);
      LEner = AC[4] + dEner;
      Lp = (g - 1.) * (LEner - 0.5 * (Lrhou * Lrhou + Lrhov * Lrhov + Lrhow * Lrhow) / Lrho);
      if (Lp < ph->lim_min_pres) {
        for (int it = 0; it < 4; it++) {
          dp *= 0.5;
          dEner = dp / (g - 1.) + 0.5 * drho * (aveu * aveu + avev * avev + avew * avew) +
                  averho * (aveu * du + avev * dv + avew * dw);
          LEner = AC[4] + dEner;
          Lp = (g - 1.) * (LEner - 0.5 * (Lrhou * Lrhou + Lrhov * Lrhov + Lrhow * Lrhow) / Lrho);
          if (Lp > ph->lim_min_pres) break;
          else if (it == 3)
            LEner = avep / (g - 1.) + 0.5 * (Lrhou * Lrhou + Lrhov * Lrhov + Lrhow * Lrhow) / Lrho;
        }
      }
      out[n * M + 0] = Lrho;
      out[n * M + 1] = Lrhou;
      out[n * M + 2] = Lrhov;
      out[n * M + 3] = Lrhow;
      out[n * M + 4] = LEner;
    }
  }
  if (!rc) {
    for (size_t i = 0; i < (size_t)K * Np * M; i++)
      if (isnan(qo[i]) || isinf(qo[i])) { rc = 7; break; }   /* wxnodaldg2dmethod.cc:246 */
    if (!rc) memcpy(qout, qo, sizeof(double) * (size_t)K * Np * M);
  }
  free(ave); free(drop); free(avc); free(avp); free(grad); free(qo); free(QM); free(QP); free(fc);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (build extension)                                        */
/* ------------------------------------------------------------------------ */

/* v^T (d f_d / d q) for the Euler flux with Apollo's index rotation; no
 * density/pressure floors (the smooth branch). */
static void euler_flux_vjp(const orc_phys *ph, int d, const double *q, const double *fb, double *qb) {
  int mu, mv, mw;
  euler_dirs(d, &mu, &mv, &mw);
  double g1 = ph->gamma - 1.0;
  double rho = q[0], u = q[mu] / rho, v = q[mv] / rho, w = q[mw] / rho, E = q[4];
  double p = g1 * (E - 0.5 * rho * (u * u + v * v + w * w));
  double pr = 0.5 * g1 * (u * u + v * v + w * w), pu = -g1 * u, pv = -g1 * v, pw = -g1 * w, pE = g1;
  double b0 = fb[0], bu = fb[mu], bv = fb[mv], bw = fb[mw], b4 = fb[4];
  qb[0] += bu * (-u * u + pr) + bv * (-u * v) + bw * (-u * w) + b4 * (pr * u - (E + p) * u / rho);
  qb[mu] += b0 + bu * (2 * u + pu) + bv * v + bw * w + b4 * (pu * u + (E + p) / rho);
  qb[mv] += bu * pv + bv * u + b4 * pv * u;
  qb[mw] += bu * pw + bw * u + b4 * pw * u;
  qb[4] += bu * pE + b4 * (1 + pE) * u;
}

static void maxwell_flux_vjp(const orc_phys *ph, int d, const double *fb, double *qb) {
  int ix[6];
  maxwell_dirs(d, ix);
  double c0 = ph->c0, chi = ph->chi, g = ph->gph;
  qb[6] += chi * c0 * c0 * fb[ix[0]];
  qb[ix[5]] += c0 * c0 * fb[ix[1]];
  qb[ix[4]] += -c0 * c0 * fb[ix[2]];
  qb[7] += g * fb[ix[3]];
  qb[ix[2]] += -fb[ix[4]];
  qb[ix[1]] += fb[ix[5]];
  qb[ix[0]] += chi * fb[6];
  qb[ix[3]] += g * c0 * c0 * fb[7];
}

static void flux_vjp(const orc_phys *ph, int d, const double *q, const double *fb, double *qb) {
  if (ph->eqn == ORC_EULER) euler_flux_vjp(ph, d, q, fb, qb);
  else maxwell_flux_vjp(ph, d, fb, qb);
}

/* d c / d q for c = |u| + sqrt(gamma p / rho) (LF wave speed). */
static double euler_speed_grad(const orc_phys *ph, const double *q, double *dc) {
  double g = ph->gamma, g1 = g - 1.0;
  double rho = q[0], u[3] = {q[1] / rho, q[2] / rho, q[3] / rho};
  double p = g1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]);
  double sp = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double a = sqrt(g * p / rho);
  double pr = 0.5 * g1 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double da_dp = g / (2 * a * rho);
  dc[0] = da_dp * pr - a / (2 * rho);
  for (int i = 0; i < 3; i++) {
    double du = sp > 0 ? u[i] / sp : 0.0;
    dc[0] += -du * u[i] / rho;
    dc[1 + i] = du / rho + da_dp * (-g1 * u[i]);
  }
  dc[4] = da_dp * g1;
  return sp + a;
}

static void numflux_vjp(const orc_phys *ph, const double *n, const double *qM,
                        const double *qP, const double *nb, double *qMb, double *qPb) {
  int m = orc_meqn(ph);
  double t[8];
  for (int c = 0; c < m; c++) t[c] = 0.5 * n[0] * nb[c];
  flux_vjp(ph, 0, qM, t, qMb);
  flux_vjp(ph, 0, qP, t, qPb);
  for (int c = 0; c < m; c++) t[c] = 0.5 * n[1] * nb[c];
  flux_vjp(ph, 1, qM, t, qMb);
  flux_vjp(ph, 1, qP, t, qPb);
  double lambda;
  if (ph->eqn == ORC_EULER) {
    double dcM[5], dcP[5];
    double cM = euler_speed_grad(ph, qM, dcM), cP = euler_speed_grad(ph, qP, dcP);
    lambda = dmax2(cM, cP);
    double jn = 0.0;
    for (int c = 0; c < m; c++) jn += (qM[c] - qP[c]) * nb[c];
    for (int c = 0; c < m; c++) {
      if (cM > cP) qMb[c] += 0.5 * jn * dcM[c];
      else qPb[c] += 0.5 * jn * dcP[c];
    }
  } else {
    lambda = dmax3(ph->chi * ph->c0, ph->gph * ph->c0, ph->c0);
  }
  for (int c = 0; c < m; c++) {
    qMb[c] += 0.5 * lambda * nb[c];
    qPb[c] -= 0.5 * lambda * nb[c];
  }
  if (ph->eqn == ORC_PHMAXWELL && ph->flux == ORC_FLUX_UPWIND) {
    double s = -0.5 * lambda * (n[0] * nb[3] + n[1] * nb[4]);
    qMb[3] += s * n[0]; qMb[4] += s * n[1];
    qPb[3] -= s * n[0]; qPb[4] -= s * n[1];
  }
}

/* qMb += (d qBC / d qM)^T qBCb */
static void bc_vjp(const orc_phys *ph, int bcnum, const double *nx, const double *qBCb, double *qMb) {
  int m = orc_meqn(ph);
  int idx = bcnum - 1;
  int kind = (idx >= 0 && idx < ph->nbc) ? ph->bc_kind[idx] : ORC_BC_NONE;
  switch (kind) {
    case ORC_BC_VORTEX: break;
    case ORC_BC_TM:
      for (int c = 0; c < m; c++) qMb[c] += qBCb[c];
      qMb[2] -= 2.0 * qBCb[2];
      break;
    case ORC_BC_PEC: {
      double En = qBCb[0] * nx[0] + qBCb[1] * nx[1];
      double Et = -qBCb[0] * nx[1] + qBCb[1] * nx[0];
      qMb[0] += En * nx[0] + Et * nx[1];
      qMb[1] += En * nx[1] - Et * nx[0];
      qMb[2] += -qBCb[2];
      double Bn = qBCb[3] * nx[0] + qBCb[4] * nx[1];
      double Bt = -qBCb[3] * nx[1] + qBCb[4] * nx[0];
      qMb[3] += -Bn * nx[0] - Bt * nx[1];
      qMb[4] += -Bn * nx[1] + Bt * nx[0];
      qMb[5] += qBCb[5];
      qMb[6] += -qBCb[6];
      qMb[7] += qBCb[7];
      break;
    }
    default:
      for (int c = 0; c < m; c++) qMb[c] += qBCb[c];
  }
}

/* Transpose of rhs_element's linearisation.  Each element's own terms are
 * computed in parallel; the adjoint traces it hands to its neighbours (qPb
 * per face point) are kept and scattered afterwards in element order, so
 * the result does not depend on the thread count. */
int orc_adjoint_rhs(const orc_ops *o, const orc_phys *ph, int K, const double *xall,
                    const double *yall, const int *etetf, double t, const double *u,
                    const double *lam, double *out) {
  const int Np = o->Np, Nc = o->Nc, Ng = o->Ng, m = orc_meqn(ph);
  size_t nst = (size_t)K * Np * m;
  memset(out, 0, sizeof(double) * nst);
  double *pbuf = ALLOC(double, (size_t)K * 3 * Ng * m);
  if (!pbuf) return 10;
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int k = 0; k < K; k++) {
    const double *xcoord = xall + (size_t)k * Np, *ycoord = yall + (size_t)k * Np;
    const double *q = u + (size_t)k * Np * m;
    double geo[14];
    int rc;
    if ((rc = orc_elem_geom(o, xcoord, ycoord, geo))) { err |= rc; continue; }
    const double *normals = geo + 5;
    double a[45 * 8], lt[45 * 8];
    for (int i = 0; i < Np * m; i++) lt[i] = lam[(size_t)k * Np * m + i] / geo[4];
    /* a = VVT^T lt */
    for (int j = 0; j < Np; j++)
      for (int c = 0; c < m; c++) {
        double s = 0.0;
        for (int i = 0; i < Np; i++) s += o->VVT[i * Np + j] * lt[i * m + c];
        a[j * m + c] = s;
      }
    double *ok = out + (size_t)k * Np * m;
    /* volume */
    double Qc[145 * 8], RRb[145 * 8], SSb[145 * 8], Qcb[145 * 8];
    mvm(Nc, Np, m, o->Vc, q, Qc);
    mvm(Nc, Np, m, o->Drc, a, RRb);
    mvm(Nc, Np, m, o->Dsc, a, SSb);
    double xr[145], xs[145], yr[145], ys[145];
    cub_geom(o, xcoord, ycoord, geo, xr, xs, yr, ys);
    for (int c = 0; c < Nc; c++) {
      double J = xr[c] * ys[c] - xs[c] * yr[c];
      double rx = ys[c] / J, sx = -yr[c] / J, ry = -xs[c] / J, sy = xr[c] / J;
      double Fb[8], Gb[8], qb[8] = {0};
      for (int v = 0; v < m; v++) {
        Fb[v] = o->cw[c] * J * (rx * RRb[c * m + v] + sx * SSb[c * m + v]);
        Gb[v] = o->cw[c] * J * (ry * RRb[c * m + v] + sy * SSb[c * m + v]);
      }
      flux_vjp(ph, 0, Qc + c * m, Fb, qb);
      flux_vjp(ph, 1, Qc + c * m, Gb, qb);
      for (int v = 0; v < m; v++) Qcb[c * m + v] = qb[v];
    }
    for (int j = 0; j < Np; j++)
      for (int v = 0; v < m; v++) {
        double s = 0.0;
        for (int c = 0; c < Nc; c++) s += o->Vc[c * Np + j] * Qcb[c * m + v];
        ok[j * m + v] += s;
      }
    /* surface */
    double QM[54 * 8], QP[54 * 8], Sb[54 * 8], Xcrd[54], Ycrd[54], qgtemp[54 * 8];
    mvm(3 * Ng, Np, m, o->If, q, QM);
    mvm(3 * Ng, Np, 1, o->If, xcoord, Xcrd);
    mvm(3 * Ng, Np, 1, o->If, ycoord, Ycrd);
    mvm(3 * Ng, Np, m, o->If, a, Sb);
    const int *connect = etetf + 6 * (size_t)k;
    for (int edge = 0; edge < 3; edge++) {
      int edgeNum = connect[2 * edge + 1];
      if (edgeNum < 0) {
        for (int g = 0; g < Ng; g++) {
          double xc[3] = {t, Xcrd[edge * Ng + g], Ycrd[edge * Ng + g]};
          orc_bc(ph, abs(edgeNum), xc, normals + 3 * edge, QM + (edge * Ng + g) * m, QP + (edge * Ng + g) * m);
        }
      } else {
        mvm(3 * Ng, Np, m, o->If, u + (size_t)connect[2 * edge] * Np * m, qgtemp);
        for (int g = 0; g < Ng; g++)
          for (int c = 0; c < m; c++)
            QP[(edge * Ng + g) * m + c] = qgtemp[(edgeNum * Ng + Ng - 1 - g) * m + c];
      }
    }
    double QMb[54 * 8];
    memset(QMb, 0, sizeof(QMb));
    for (int kk = 0; kk < 3 * Ng; kk++) {
      int edge = kk / Ng, g = kk % Ng;
      double nb[8], qMb[8] = {0}, qPb[8] = {0};
      double scale = -normals[3 * edge + 2] * o->gw[g];
      for (int c = 0; c < m; c++) nb[c] = scale * Sb[kk * m + c];
      numflux_vjp(ph, normals + 3 * edge, QM + kk * m, QP + kk * m, nb, qMb, qPb);
      int edgeNum = connect[2 * edge + 1];
      if (edgeNum < 0) {
        bc_vjp(ph, abs(edgeNum), normals + 3 * edge, qPb, qMb);
      } else {
        /* the neighbour's share, scattered below */
        for (int c = 0; c < m; c++) pbuf[((size_t)k * 3 * Ng + kk) * m + c] = qPb[c];
      }
      for (int c = 0; c < m; c++) QMb[kk * m + c] = qMb[c];
    }
    for (int j = 0; j < Np; j++)
      for (int c = 0; c < m; c++) {
        double s = 0.0;
        for (int kk = 0; kk < 3 * Ng; kk++) s += o->If[kk * Np + j] * QMb[kk * m + c];
        ok[j * m + c] += s;
      }
  }
  /* neighbour terms: row edgeNum*Ng + Ng-1-g of If applied to q_nbr */
  for (int k = 0; k < K && !err; k++) {
    const int *connect = etetf + 6 * (size_t)k;
    for (int kk = 0; kk < 3 * Ng; kk++) {
      int edge = kk / Ng, g = kk % Ng, edgeNum = connect[2 * edge + 1];
      if (edgeNum < 0) continue;
      int nb_e = connect[2 * edge], row = edgeNum * Ng + Ng - 1 - g;
      const double *qPb = pbuf + ((size_t)k * 3 * Ng + kk) * m;
      double *on = out + (size_t)nb_e * Np * m;
      for (int j = 0; j < Np; j++)
        for (int c = 0; c < m; c++) on[j * m + c] += o->If[row * Np + j] * qPb[c];
    }
  }
  free(pbuf);
  return err;
}

int orc_lserk4_adjoint(const orc_ops *o, const orc_phys *ph, int K, const double *x,
                       const double *y, const int *etetf, double t0, double dt,
                       int nsteps, const double *q0, const double *qbarT, double *dJdq0) {
  size_t n = (size_t)K * o->Np * orc_meqn(ph);
  double *traj = ALLOC(double, n * (size_t)(nsteps + 1));
  double *stg = ALLOC(double, n * 5), *res = ALLOC(double, n), *rhs = ALLOC(double, n);
  double *qb = ALLOC(double, n), *rb = ALLOC(double, n), *tmp = ALLOC(double, n);
  int rc = 0;
  memcpy(traj, q0, sizeof(double) * n);
  for (int step = 0; step < nsteps && !rc; step++) {
    memcpy(traj + n * (step + 1), traj + n * step, sizeof(double) * n);
    rc = orc_lserk4(o, ph, K, x, y, etetf, t0 + step * dt, dt, 1, traj + n * (step + 1), 0);
  }
  memcpy(qb, qbarT, sizeof(double) * n);
  for (int step = nsteps - 1; step >= 0 && !rc; step--) {
    double t = t0 + step * dt;
    /* recompute the stage inputs q_{s-1} of this step */
    double *q = tmp;
    memcpy(q, traj + n * step, sizeof(double) * n);
    memset(res, 0, sizeof(double) * n);
    for (int s = 0; s < 5; s++) {
      double ms;
      memcpy(stg + n * s, q, sizeof(double) * n);
      rc |= orc_rhs(o, ph, K, x, y, etetf, t + orc_rk4c[s] * dt, q, rhs, &ms, 0);
      for (size_t i = 0; i < n; i++) {
        res[i] = orc_rk4a[s] * res[i] + dt * rhs[i];
        q[i] += orc_rk4b[s] * res[i];
      }
    }
    memset(rb, 0, sizeof(double) * n);
    for (int s = 4; s >= 0; s--) {
      for (size_t i = 0; i < n; i++) rb[i] += orc_rk4b[s] * qb[i];
      rc |= orc_adjoint_rhs(o, ph, K, x, y, etetf, t + orc_rk4c[s] * dt, stg + n * s, rb, rhs);
      for (size_t i = 0; i < n; i++) {
        qb[i] += dt * rhs[i];
        rb[i] *= orc_rk4a[s];
      }
    }
  }
  memcpy(dJdq0, qb, sizeof(double) * n);
  free(traj); free(stg); free(res); free(rhs); free(qb); free(rb); free(tmp);
  return rc;
}

double orc_objective_disc_sq(const orc_ops *o, int K, const double *x, const double *y,
                             int m, const double *q, int var, double xc, double yc,
                             double radius, double *dJdq) {
  int Np = o->Np, Nc = o->Nc;
  double J = 0.0;
  if (dJdq) memset(dJdq, 0, sizeof(double) * (size_t)K * Np * m);
  for (int k = 0; k < K; k++) {
    double geo[14];
    orc_elem_geom(o, x + (size_t)k * Np, y + (size_t)k * Np, geo);
    for (int c = 0; c < Nc; c++) {
      double qc = 0.0, px = 0.0, py = 0.0;
      for (int j = 0; j < Np; j++) {
        qc += o->Vc[c * Np + j] * q[((size_t)k * Np + j) * m + var];
        px += o->Vc[c * Np + j] * x[(size_t)k * Np + j];
        py += o->Vc[c * Np + j] * y[(size_t)k * Np + j];
      }
      double r2 = (px - xc) * (px - xc) + (py - yc) * (py - yc);
      if (r2 < radius * radius) {
        J += o->cw[c] * geo[4] * qc * qc;
        if (dJdq)
          for (int j = 0; j < Np; j++)
            dJdq[((size_t)k * Np + j) * m + var] += 2.0 * o->cw[c] * geo[4] * qc * o->Vc[c * Np + j];
      }
    }
  }
  return J;
}
