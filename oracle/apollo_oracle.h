/*
 * apollo_oracle.h -- CPU restatement of Apollo's nodal-DG RHS (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the MI355X engine.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it; the
 * engine never links or calls it.
 *
 * Every function restates the reference's algorithm in plain C and cites
 * the reference file:line it follows (paths relative to the Apollo tree).
 * Apollo's own variable counts are used: 5 for `eulerEqn`, 8 for
 * `phMaxwellEqn`.
 *
 * Parity status of each part (what pins it) is listed next to each
 * declaration: "pinned" = checked in tests against reference-held tables,
 * closed forms, invariants or brute force; "parity unpinned" = no such
 * check exists.
 */
#ifndef APOLLO_ORACLE_H
#define APOLLO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Operators of one polynomial order, row-major FP64. */
typedef struct {
  int N, Np, Nc, Ng, Nfp;
  double rmin;
  double *r, *s;          /* [Np] nodes (wxpnodaldgfunctions.h:112-256)      */
  int *fmask;             /* [3*Nfp]                                          */
  double *V, *invV, *VVT; /* [Np*Np] (wxnodaldggeometry2d.cc:49-76)            */
  double *Dr, *Ds;        /* [Np*Np] nodal differentiation                    */
  double *cr, *cs, *cw;   /* [Nc] cubature rule (wxcubature2d.cc:20-75)       */
  double *Vc, *Drc, *Dsc; /* [Nc*Np] (wxcubature2d.cc:77-100)                 */
  double *VcT, *DrcT, *DscT; /* [Np*Nc]                                       */
  double *gz, *gw;        /* [Ng] Gauss-Legendre (wxcubature2d.cc:110-116)    */
  double *If, *IfT;       /* [3Ng*Np], [Np*3Ng] (wxcubature2d.cc:118-133)     */
} orc_ops;

/* pinned: JacobiP/GradJacobiP closed forms, NUDG tables V, invV, Dr, Ds
 * (src/lib/dataN*.h), cubature exactness, gauleg vs numpy. */
int orc_build_ops(int N, orc_ops *ops);
void orc_free_ops(orc_ops *ops);

/* Individual building blocks (exported for tests). */
double orc_jacobiP(double x, double alpha, double beta, int n);
double orc_grad_jacobiP(double x, double alpha, double beta, int n);
void orc_gauleg(double x1, double x2, double *x, double *w, int n);
int orc_invert(int n, const double *A, double *Ainv);

/* Connectivity, FacePair2d semantics (wxnodaldggeometry2d.cc:235-367),
 * with faces matched on (vert_id[a], vert_id[b]) pairs.  etetf[K*6].
 * pinned: brute-force O(F^2) matcher on small meshes, Gmsh fixtures. */
int orc_facepair(int K, int Nv, const int *EToV, const int *vert_id,
                 const int *face_tag, int *etetf);

/* Node coordinates and dtscale (wxnodaldggeometry2d.cc:179-233).
 * x, y [K*Np]. */
void orc_node_coords(const orc_ops *ops, int K, const double *VX,
                     const double *VY, const int *EToV, double *x, double *y,
                     double *dtscale);

/* Per-element geometric factors and normals (wxnodaldggeometry2d.cc:382-442):
 * geom[14] = rx, sx, ry, sy, J, nx0, ny0, sJ0/2, nx1, ny1, sJ1/2, nx2, ny2, sJ2/2.
 * Returns 0, or 3 (J <= 0) / 4 (zero edge). */
int orc_elem_geom(const orc_ops *ops, const double *x, const double *y, double *geom);

/* Physics configuration of the RHS. */
enum { ORC_EULER = 0, ORC_PHMAXWELL = 1 };
/* Euler numerical fluxes follow WxEulerEqn::DGnumericalFlux
 * (wxeulereqn.cc:873-894): LF, HLL, Roe, and "Wave", which falls through to
 * LF of the face-frame traces.
 * Maxwell: LF, or the build's exact upwind flux (UPWIND). */
enum { ORC_FLUX_LF = 0, ORC_FLUX_UPWIND = 1, ORC_FLUX_HLL = 2, ORC_FLUX_ROE = 3, ORC_FLUX_WAVE = 4 };
enum { ORC_BC_NONE = 0, ORC_BC_VORTEX = 1, ORC_BC_TM = 2, ORC_BC_PEC = 3, ORC_BC_COPY = 4,
       ORC_BC_EULER_WALL = 5, ORC_BC_EULER_INFLOW = 6, ORC_BC_EULER_DIRICHLET = 7 };
typedef struct {
  int eqn;              /* ORC_EULER (m = 5) or ORC_PHMAXWELL (m = 8)          */
  int flux;             /* ORC_FLUX_*: UPWIND for Maxwell (build ext), HLL/ROE/WAVE Euler */
  double gamma, min_dens, min_pres;
  double c0, chi, gph;
  int nbc;
  int bc_kind[8];
  double bc_p[8][8];
  /* eulerLimiterHW block: gas_gamma, min_density, min_pressure */
  double lim_gamma, lim_min_dens, lim_min_pres;
  int limiter;          /* 1: limit each RHS input first (wxnodaldg2dmethod.cc:243-247) */
} orc_phys;

int orc_meqn(const orc_phys *ph);

/* Physical flux (wxeulereqn.cc:806-871, wxphmaxwelleqn.cc:447-498).
 * Returns 0 or 8 / 9 (negative density / pressure). */
int orc_flux(const orc_phys *ph, int d, const double *q, double *f);
/* Numerical flux (wxhyperboliceqnset.cc:184-203 -> wxeulereqn.cc:1088-1117,
 * wxphmaxwelleqn.cc:500-522).  qM, qP may be modified (wxeulereqn.cc:1329-1340). */
int orc_numflux(const orc_phys *ph, const double *n, double *qM, double *qP,
                double *nflux, double *lambda);
/* Boundary state (wxgridbc.cc:21-26 and the applyBC files): isentropic
 * vortex, Maxwell TM, PEC, and the Euler wall, inflow (p = Mach, angle in
 * degrees, gas_gamma, pressure, density), Dirichlet (p = density, u, v, w,
 * gas_gamma, pressure) and zero-gradient (copy) states.
 * pinned: closed forms and reflection invariants in the tests. */
void orc_bc(const orc_phys *ph, int bcnum, const double *xc, const double *n,
            const double *qM, double *qP);

/* The RHS, WxNodalDG2dMethod::step (wxnodaldg2dmethod.cc:235-542):
 * rhs[K*Np*m]; *maxSpeed receives max lambda.  nthreads <= 0 uses OpenMP's
 * default.  Returns 0 or an error code (7 NaN, 8/9 flux, 3/4 geometry).
 * parity unpinned by reference data (no reference output exists); pinned
 * by free-stream preservation, conservation, the analytic TM cavity mode
 * and the closed-form flux checks. */
int orc_rhs(const orc_ops *ops, const orc_phys *ph, int K, const double *x,
            const double *y, const int *etetf, double t, const double *q,
            double *rhs, double *maxSpeed, int nthreads);

/* Diagnostic switch: use each element's affine geometric factors at the
 * cubature points instead of Apollo's Dr_c x / Ds_c x recomputation
 * (wxcubature2d.cc:305,327-356).  Off (0) by default; process-global. */
void orc_set_exact_geometry(int on);
/* Diagnostic: every matrix-vector product of the RHS sums in reverse order. */
void orc_set_reverse_sums(int on);

/* Initial dt (wxnodaldg2dmethod.cc:201-228) using Roe speeds of a state
 * against itself (wxeulereqn.cc:95-200) / c0 max(chi,gamma,1) for Maxwell. */
double orc_initial_dt(const orc_ops *ops, const orc_phys *ph, int K,
                      const double *q, double cfl, double dtscale);

/* Cell-average weights of the limiter, ave[Np] (wxnodaldggeometry2d.h:84-98). */
int orc_limiter_weights(const orc_ops *ops, double *ave);
/* Hesthaven-Warburton Euler limiter, applyLimiter as written
 * (wxhesthavenwarburtoneulerlimiter.cc:61-484); Euler, N <= 4, no vortex
 * boundary (11 otherwise); 7 on a NaN/Inf result.  qout may equal qin.
 * pinned: preserves cell averages and linear states away from the patch
 * quirks, reproduces by hand-computed patches, see tests. */
int orc_limit(const orc_ops *ops, const orc_phys *ph, int K, const double *x,
              const double *y, const int *etetf, const double *qin, double *qout);

/* LSERK4 (Carpenter & Kennedy 1994 five-stage, as tabulated in Hesthaven &
 * Warburton 2008): no reference counterpart. */
extern const double orc_rk4a[5], orc_rk4b[5], orc_rk4c[5];
int orc_lserk4(const orc_ops *ops, const orc_phys *ph, int K, const double *x,
               const double *y, const int *etetf, double t0, double dt,
               int nsteps, double *q, int nthreads);

/* Discrete adjoint of orc_rhs: out = (d rhs/d q)^T lam.  No reference
 * counterpart; pinned by the dot-product test and finite differences. */
int orc_adjoint_rhs(const orc_ops *ops, const orc_phys *ph, int K,
                    const double *x, const double *y, const int *etetf,
                    double t, const double *q, const double *lam, double *out);

/* Reverse LSERK4 with full recomputation from q0 (no checkpoint tricks):
 * dJdq0 = (d q_T / d q0)^T qbarT. */
/* One output frame of Apollo's stepping: PETSc TSSSP rks2 (s stages) under
 * ApSolver's dt control (apsolver.cc:172-216, 421-480).  Returns 0, an RHS
 * error code, or 1 when more than max_steps steps would be needed.
 * parity unpinned against PETSc (its TSSSP source is not in the reference
 * tree): pinned by the method's stability polynomial, second-order
 * convergence and the dt-control invariants in the tests. */
int orc_ssp_advance(const orc_ops *ops, const orc_phys *ph, int K, const double *x,
                    const double *y, const int *etetf, double t0, double t_end, double dt0,
                    double cfl, double dtscale, int stages, double *q, int max_steps,
                    int *nsteps, double *dts, int nthreads);

int orc_lserk4_adjoint(const orc_ops *ops, const orc_phys *ph, int K,
                       const double *x, const double *y, const int *etetf,
                       double t0, double dt, int nsteps, const double *q0,
                       const double *qbarT, double *dJdq0);

/* J = sum_k sum_c w_c J_k q_var(x_c)^2 [r_c < radius], and dJ/dq. */
double orc_objective_disc_sq(const orc_ops *ops, int K, const double *x,
                             const double *y, int m, const double *q, int var,
                             double xc, double yc, double radius, double *dJdq);

#ifdef __cplusplus
}
#endif
#endif
