/*
 * apollo_dg.h -- C-ABI of the MI355X nodal-DG right-hand-side engine.
 *
 * The engine replaces the hot path of Apollo's `nodalDG2d` subsolver: the
 * over-integrated (cubature + edge-Gauss) weak-form RHS of
 * `WxNodalDG2dMethod<double>::step` and the time stepper that calls it.
 * Every entry point below names the reference interface it stands in for.
 *
 * Conventions
 *  - Every function returns an `int` status (APL_OK == 0).  No function
 *    calls exit(), throws, or prints; numerical faults that Apollo turns into
 *    exit(1) come back as status codes.
 *  - The state layout is Apollo's: element-major, then node, then variable,
 *    FP64, i.e. q[(k*Np + n)*m + v] for k in [0,K), n in [0,Np), v in [0,m)
 *    (src/subsolvers/wxnodaldg2dmethod.cc:194-199,329-332,494-497).
 *  - Pointers to state are host pointers when `where == APL_HOST` and device
 *    (HIP) pointers when `where == APL_DEVICE`.
 *  - A handle is not thread-safe; the caller serialises calls on it.
 *  - Plain C types only: no HIP, PETSc or torch types cross this boundary.
 */
#ifndef APOLLO_DG_H
#define APOLLO_DG_H

#ifdef __cplusplus
extern "C" {
#endif

#define APL_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum apl_status {
  APL_OK = 0,
  APL_ERR_ARG = 1,          /* bad argument / unsupported combination        */
  APL_ERR_ORDER = 2,        /* N outside 1..8 (wxpnodaldgfunctions.h:271-276)*/
  APL_ERR_MESH_JACOBIAN = 3,/* J <= 0 (wxnodaldggeometry2d.cc:428-434)       */
  APL_ERR_MESH_EDGE = 4,    /* zero edge length (wxnodaldggeometry2d.cc:399) */
  APL_ERR_MESH_TOPOLOGY = 5,/* a face shared by more than two elements, or a
                               boundary face with no boundary condition      */
  APL_ERR_HIP = 6,          /* a HIP runtime call failed                      */
  APL_ERR_NAN = 7,          /* NaN in the RHS (wxnodaldg2dmethod.cc:478-503)  */
  APL_ERR_NEG_DENSITY = 8,  /* rho <= 0 in the Euler flux (wxeulereqn.cc:838) */
  APL_ERR_NEG_PRESSURE = 9, /* p <= 0 in the Euler flux (wxeulereqn.cc:851)   */
  APL_ERR_ALLOC = 10,       /* host or device allocation failed               */
  APL_ERR_UNSUPPORTED = 11  /* valid request the library does not provide     */
};
/* The kernels raise flags and the host reads them once per call (once per
 * step in the steppers).  When one call raises several, the status is the
 * first of NEG_DENSITY, NEG_PRESSURE, NAN.  Apollo exits at the first failing
 * node instead, so where a state has both a negative density and a
 * negative pressure the two may name different ones. */

/* ---- equation systems --------------------------------------------------- */
enum apl_eqn {
  APL_EQN_EULER5 = 0,      /* Apollo `eulerEqn`: (rho, rho u, rho v, rho w, E)
                              src/hyperapps/euler/wxeulereqn.h:28-30         */
  APL_EQN_EULER4 = 1,      /* 2D restriction rho w == 0: (rho, rho u, rho v, E)*/
  APL_EQN_PHMAXWELL8 = 2,  /* Apollo `phMaxwellEqn`: (Ex,Ey,Ez,Bx,By,Bz,phi,psi)
                              src/hyperapps/maxwell/wxphmaxwelleqn.cc:447-498 */
  APL_EQN_MAXWELL_TM3 = 3  /* TM subset (Ez, Bx, By) with chi == gamma == 0    */
};

enum apl_flux {
  APL_FLUX_LF = 0,         /* Lax-Friedrichs (Rusanov): wxeulereqn.cc:1088-1117,
                              wxphmaxwelleqn.cc:500-522                      */
  APL_FLUX_UPWIND = 1,     /* exact upwind flux, APL_EQN_MAXWELL_TM3 only (no
                              counterpart in the reference)                  */
  APL_FLUX_HLL = 2,        /* Euler HLL, wxeulereqn.cc:1119-1202             */
  APL_FLUX_ROE = 3,        /* Euler Roe as Apollo computes it,
                              wxeulereqn.cc:1204-1268: the x-momentum slot
                              holds the y-momentum flux and the y-momentum
                              flux is 0 (:1259-1260 assign nflux[1] twice)   */
  APL_FLUX_WAVE = 4        /* Euler "Wave" as Apollo computes it: the wave-
                              propagation routine rotates both traces'
                              momentum to the face frame in place
                              (wxeulereqn.cc:1292-1300) and the switch falls
                              through to LF (:888-892), so the flux is LF of
                              the rotated traces                             */
};

enum apl_precision {
  APL_PREC_FP64 = 0,       /* FP64 everywhere                                 */
  APL_PREC_FP32_MFMA = 1,  /* contractions on FP32 MFMA, FP32 accumulation;
                              geometry, 1/J, dt and stage update in FP64      */
  APL_PREC_BF16X3 = 2      /* contractions as split-BF16 x3 MFMA products with
                              FP32 accumulation; FP64 as above               */
};

enum apl_algorithm {
  APL_ALGO_AUTO = 0,       /* cubature form for nonlinear fluxes (Euler); for
                              linear fluxes (Maxwell) the algebraically
                              identical nodal form V V^T Dr_c^T W V_c Q ... on
                              straight-sided elements (no reference
                              counterpart; same result to rounding).  The
                              cubature form runs on the vector-ALU kernel at
                              N <= 3 in FP64 and FP32, on MFMA otherwise     */
  APL_ALGO_CUBATURE = 1,   /* always Apollo's over-integrated cubature form
                              (wxcubature2d.cc:288-382), on the kernels
                              APL_ALGO_AUTO picks for it                     */
  APL_ALGO_CUBATURE_MFMA = 2, /* the cubature form on the MFMA kernel at every
                              order (padded 16x16x4 tiles at N <= 3)         */
  APL_ALGO_CUBATURE_VALU = 3, /* the cubature form on the vector-ALU kernel
                              (one lane per element, operators as scalar
                              operands); N <= 3, FP64 or FP32 only, else
                              APL_ERR_UNSUPPORTED                            */
  APL_ALGO_NODAL_MFMA = 4, /* the nodal form of the linear fluxes on the MFMA
                              kernel (Maxwell only, else APL_ERR_ARG)        */
  APL_ALGO_NODAL_VALU = 5  /* the nodal form on the vector-ALU kernel (one
                              lane per element, rows split over the waves of
                              a workgroup); MAXWELL_TM3 in FP64 or FP32 only,
                              else APL_ERR_UNSUPPORTED                       */
};

enum apl_bc_kind {
  APL_BC_NONE = 0,
  APL_BC_ISENTROPIC_VORTEX = 1, /* exact vortex state, wxisentropicvortexbc.cc:20-43
                                   p[0..5] = u0, v0, x0, y0, gas_gamma, beta  */
  APL_BC_MAXWELL_TM = 2,        /* Ez -> -Ez, rest copied,
                                   wxmaxwelltranversemagneticbc.cc:13-25     */
  APL_BC_PEC = 3,               /* conducting wall, wxphmaxwellconductingwallbc.cc:12-41 */
  APL_BC_COPY = 4,              /* q+ = q- (zero-gradient outflow,
                                   wxeulerzerogradientbc.cc:18-25)           */
  APL_BC_EULER_WALL = 5,        /* reflect the normal momentum, rho w = 0,
                                   wxeulerwallbc.cc:18-25                    */
  APL_BC_EULER_INFLOW = 6,      /* fixed supersonic inflow, wxeulerinflowbc.cc:21-36
                                   p[0..4] = Mach_number, Flow_angle (deg),
                                   gas_gamma, Pressure, density              */
  APL_BC_EULER_DIRICHLET = 7    /* fixed state, wxeulerdirichletbc.cc:22-30
                                   p[0..5] = density, xvelocity, yvelocity,
                                   zvelocity, gas_gamma, pressure            */
};

enum apl_where { APL_HOST = 0, APL_DEVICE = 1 };

#define APL_MAX_BC 8

typedef struct {
  int kind;        /* apl_bc_kind */
  double p[8];     /* kind-specific parameters, see apl_bc_kind */
} apl_bc;

/* Mesh as DMPlex hands it to Apollo after DMPlexCreateGmsh + Distribute on
 * one rank: cells [0,K) in file order, each with a counter-clockwise cone of
 * three vertices; faces vn = {{0,1},{1,2},{2,0}}
 * (src/lib/wxnodaldggeometry2d.cc:243-263). */
typedef struct {
  int K;                 /* number of triangles                                */
  int Nv;                /* number of vertices                                 */
  const double *VX, *VY; /* [Nv] vertex coordinates                            */
  const int *EToV;       /* [K*3] 0-based vertex ids in cone order              */
  const int *vert_id;    /* [Nv] topological id used to match faces, or NULL
                            (identity).  Periodic meshes map each image vertex
                            to the id of its partner.                          */
  const int *face_tag;   /* [K*3] Gmsh physical tag of face f of element k
                            ("Face Sets" label), >= 0, 0 = none; NULL = no
                            tags.  A negative tag is APL_ERR_ARG.             */
} apl_mesh;

typedef struct {
  int order;             /* polynomial order N, 1..8                           */
  int eqn;               /* apl_eqn                                            */
  int flux;              /* apl_flux                                           */
  int precision;         /* apl_precision                                      */
  int algorithm;         /* apl_algorithm                                      */
  double gas_gamma;      /* Euler `gas_gamma`                                  */
  double min_density;    /* Euler `minDensity`  (0: rho<=0 is an error)        */
  double min_pressure;   /* Euler `minPressure` (0: p<=0 is an error)          */
  double c0, chi, gamma_ph; /* Maxwell `c0`, `chi`, `gamma`                    */
  int n_bc;              /* boundary-condition list; a face with tag T uses
                            bc[T-1] (src/subsolvers/wxnodaldg2dmethod.cc:569)  */
  apl_bc bc[APL_MAX_BC];
  int limiter;           /* 1: Apollo's `Limiter = [eulerLimiterHW]`
                            (wxnodaldg2dmethod.cc:126-134): the steppers limit
                            every RHS input first (:243-247); Euler, N <= 4,
                            no isentropic-vortex boundary condition           */
  double lim_gas_gamma, lim_min_density, lim_min_pressure;
                         /* the limiter block's gas_gamma, min_density,
                            min_pressure (wxhesthavenwarburtoneulerlimiter.cc:
                            21-23)                                            */
} apl_params;

typedef struct {
  int K, N, Np, Nc, Ng, Nfp, m;
  double dtscale;        /* min inscribed radius (wxnodaldggeometry2d.cc:220-227) */
  double rmin;           /* |r1 - r0|          (wxnodaldggeometry2d.cc:91)         */
  long long device_bytes;/* device memory held by the handle                     */
  int device;            /* HIP device the handle was created on                 */
} apl_info;

typedef struct apl_handle_s apl_handle;

/* ---- lifetime ----------------------------------------------------------- */

/* Replaces WxNodalDG2dMethod<double>::setup (wxnodaldg2dmethod.cc:50-152):
 * builds the nodal, cubature and face-Gauss operators, the geometric factors
 * and the connectivity on the host in FP64 and uploads them to the current
 * HIP device. */
int apl_create(const apl_mesh *mesh, const apl_params *params, apl_handle **out);

/* Replaces ~WxNodalDG2dMethod (wxnodaldg2dmethod.cc:18-45). */
int apl_destroy(apl_handle *h);

int apl_get_info(const apl_handle *h, apl_info *out);

/* Work on a caller-provided hipStream_t (NULL = the null stream). */
int apl_set_stream(apl_handle *h, void *hip_stream);

/* ---- index maps and geometry (bit-exact checks) ------------------------- */

/* etetf[K*6]: wxNodalDGgeometry2D::ElementTOElementANDFace
 *   (wxnodaldggeometry2d.h:78-81, built by FacePair2d :235-367): neighbour
 *   element and face per face, or -(tag) twice on a boundary face.
 * fmask[3*(N+1)]: face node lists (wxpnodaldgfunctions.h:112-256).
 * vmapP[K*3*(N+1)]: for face f of element k, the global node index
 *   (element*Np + node) of the neighbour node that coincides with own face
 *   node i, or -1 on a boundary face.  Any pointer may be NULL. */
int apl_get_index_maps(const apl_handle *h, int *etetf, int *fmask, int *vmapP);

/* geom[K*14] per element: rx, sx, ry, sy, J (GeometricFactors2d :412-442),
 * then nx, ny, L/2 for faces 0,1,2 (Normals2d :382-410).
 * x, y [K*Np]: node coordinates (CalculateNodeCoordinates2d :179-233).
 * Any pointer may be NULL. */
int apl_get_geometry(const apl_handle *h, double *geom, double *x, double *y);

/* Operators in row-major FP64 (the values the kernels are built from):
 * which = 0: V_c [Nc x Np], 1: Dr_c [Nc x Np], 2: Ds_c [Nc x Np],
 *         3: I_f [3Ng x Np], 4: V V^T [Np x Np], 5: cubature w [Nc],
 *         6: Gauss z [Ng], 7: Gauss w [Ng], 8: node r [Np], 9: node s [Np]. */
int apl_get_operator(const apl_handle *h, int which, double *out, int capacity);

/* ---- the RHS and the stepper -------------------------------------------- */

/* Replaces WxNodalDG2dMethod<double>::step (wxnodaldg2dmethod.cc:235-542):
 * rhs = M^-1 [ volume - surface ] / J for every element, overwriting `rhs`.
 * `suggested_dt` (may be NULL) receives (2/3) cfl dtscale rmin / maxSpeed
 * (wxnodaldg2dmethod.cc:523) with `cfl` as given.  With APL_DEVICE, `rhs`
 * must not overlap `q` (APL_ERR_ARG): the kernels gather neighbour nodes of
 * q while other elements' results are stored. */
int apl_rhs(apl_handle *h, double t, const double *q, double *rhs,
            double cfl, double *suggested_dt, int where);

/* apl_rhs on device pointers that allocates nothing and calls no
 * allocating routine, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same kernels and results
 * as apl_rhs(..., APL_DEVICE).  Its per-batch speed maxima and error flags
 * live in a static device array of the library (8 MB per GPU, held until the
 * process ends) instead of the handle's buffers; K above 16 777 216 elements
 * is APL_ERR_ARG.  Calls that overlap on one GPU share that array and race:
 * the caller serialises them.  Synchronises the handle's stream once to read
 * the flags and the speed maximum back. */
int apl_rhs_noalloc(apl_handle *h, double t, const double *q, double *rhs, double cfl,
                    double *suggested_dt);

/* Replaces WxNodalDG2dMethod::applyLimiter -> WxHestavenWarburtonEulerLimiter::
 * applyLimiter (wxnodaldg2dmethod.cc:575-582, src/hyperapps/euler/
 * wxhesthavenwarburtoneulerlimiter.cc:61-484): every element is replaced
 * by the reconstruction from its cell average and its limited gradient.
 * q_out may equal q_in.  Needs a handle created with limiter = 1.  apl_rhs
 * does not limit; Apollo's step() limits its input before the RHS, which is
 * what apl_lserk4 and apl_ssp_advance do on such handles.  A NaN or Inf in
 * the result is APL_ERR_NAN (wxnodaldg2dmethod.cc:246). */
int apl_limit(apl_handle *h, const double *q_in, double *q_out, int where);

/* Initial-condition dt of WxNodalDG2dMethod::init (wxnodaldg2dmethod.cc:
 * 201-228): Roe wave speeds at every node, (2/3) cfl dtscale rmin / maxSpeed. */
int apl_initial_dt(apl_handle *h, const double *q, double cfl, double *dt, int where);

/* Low-storage RK4(5) (Carpenter & Kennedy 1994) over `nsteps` steps of a
 * fixed dt, starting at time t0; q is advanced in place.  Each step runs as
 * one captured hipGraph of five fused RHS + stage-update kernels.  Stands in
 * for the PETSc TS loop around ApSolver::ComputeRHSforTS
 * (src/solvers/apsolver.cc:442-480, src/subsolvers/wxpetsctimestepping.h).
 * On a handle with the limiter, each stage limits its input state in place
 * before the RHS. */
int apl_lserk4(apl_handle *h, double t0, double dt, int nsteps, double *q, int where);

/* Apollo's own stepping of one output frame [t0, t_end], in place of the
 * PETSc TSSolve that ApSolver::solve runs per frame (src/solvers/
 * apsolver.cc:172-216): TSSSP (src/subsolvers/wxpetsctimestepping.h:33) of
 * type rks2 with `stages` stages (PETSc's default is 5), TSSetInitialTimeStep
 * (t0, dt0), exact final time MATCHSTEP, the MonitorVTK clamp to the frame
 * end (apsolver.cc:421-440) and ComputeRHSforTS's dt = min(suggested, dt)
 * after each RHS (apsolver.cc:442-480), with the suggested dt of
 * WxNodalDG2dMethod::step at `cfl`.  q is advanced in place; *nsteps
 * receives the step count and dts[0 .. *nsteps) (optional, max_steps long)
 * each step's dt.  More than max_steps steps is APL_ERR_ARG.  On a handle
 * with the limiter, each RHS input (the TSSSP stage vector, never q itself)
 * is limited in place first, as step() does (wxnodaldg2dmethod.cc:243-247). */
int apl_ssp_advance(apl_handle *h, double t0, double t_end, double dt0, double cfl, int stages,
                    double *q, int where, int max_steps, int *nsteps, double *dts);

/* ---- discrete adjoint (no counterpart in the reference) ----------------- */

/* out = (d rhs / d q)^T lam at state q and time t (FP64).  With APL_DEVICE,
 * `out` must not overlap `q` or `lam` (APL_ERR_ARG). */
int apl_adjoint_rhs(apl_handle *h, double t, const double *q, const double *lam,
                    double *out, int where);

/* Reverse-mode LSERK4: given q0 and the terminal adjoint qbarT = dJ/dq(T)
 * after nsteps fixed-dt steps from t0, returns dJdq0 = dJ/dq(0).  Forward
 * states are checkpointed every `ckpt_every` steps in device memory; each
 * segment is recomputed once from its checkpoint, keeping the five stage
 * inputs of each of its steps (device memory for nsteps / ckpt_every + 1 +
 * 5 ckpt_every states).  Optionally returns q(T). */
int apl_lserk4_adjoint(apl_handle *h, double t0, double dt, int nsteps,
                       int ckpt_every, const double *q0, const double *qbarT,
                       double *dJdq0, double *qT, int where);

/* Objective J = sum_k sum_c w_c J_k q_var(x_c)^2 [ |x_c - (xc,yc)| < radius ]
 * over cubature points, and its gradient dJ/dq (same layout as q). */
int apl_objective_disc_sq(apl_handle *h, const double *q, int var, double xc,
                          double yc, double radius, double *J, double *dJdq,
                          int where);

/* ---- synthetic meshes --------------------------------------------------- */

/* Structured n x n quads on [x0,x1]x[y0,y1], each split along the
 * (i,j)->(i+1,j+1) diagonal into T0 = (v_ij, v_i+1j, v_i+1j+1) and
 * T1 = (v_ij, v_i+1j+1, v_ij+1); elements in row-major quad order, T0 then
 * T1.  Interior vertices are jittered by independent U(-a h, a h) in x and
 * y from std::mt19937_64(seed) mapped as (x >> 11) * 2^-53.  With
 * periodic != 0 vertices on i = n / j = n carry the topological id of
 * i = 0 / j = 0; otherwise every boundary face gets tag 1.
 * Sizes: K = 2 n^2, Nv = (n+1)^2.  Output arrays are caller-allocated. */
int apl_mesh_structured(int n, double x0, double x1, double y0, double y1,
                        double jitter, unsigned long long seed, int periodic,
                        double *VX, double *VY, int *EToV, int *vert_id,
                        int *face_tag);

/* ---- host-only utilities (no device needed) ----------------------------- */

/* The connectivity apl_create would build for `mesh` (FacePair2d,
 * src/lib/wxnodaldggeometry2d.cc:235-367): etetf[K*6] as in
 * apl_get_index_maps. */
int apl_host_connectivity(const apl_mesh *mesh, int *etetf);

/* The operators apl_create would build for order N (see apl_get_operator
 * for `which`); `capacity` is the size of `out` in doubles. */
int apl_host_operator(int order, int which, double *out, int capacity);

/* ---- errors -------------------------------------------------------------- */
const char *apl_status_string(int status);
/* Detail of the last failure on this handle (or of the last failed
 * apl_create when h is NULL). */
const char *apl_last_error(const apl_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* APOLLO_DG_H */
