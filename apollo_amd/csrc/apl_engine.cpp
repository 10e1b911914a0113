// apl_engine.cpp -- the C-ABI of the engine (include/apollo_dg.h).
//
// Host side: builds operators, geometry and connectivity (apl_host.cpp),
// folds and permutes the operators into MFMA fragments, uploads everything
// to the current HIP device and launches the kernels of apl_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/apollo_dg.h"
#include "apl_adjoint.h"
#include "apl_device.h"
#include "apl_limiter.h"
#include "apl_host.h"

namespace apl {
template <int N>
int launch_rhs(int eqn, int flux, int prec, int linear, hipStream_t s, const RhsArgs &a,
               int nblocks);
template <int N>
int launch_adjoint(int eqn, int flux, hipStream_t s, const AdjArgs &a, int nblocks);
template <int N>
int launch_objective(int m, hipStream_t s, const ObjArgs &a, int nblocks);
}  // namespace apl

using namespace apl;

static thread_local std::string g_create_error;

struct apl_handle_s {
  apl_params params;
  int K = 0, M = 0, Nv = 0;
  bool linear = false;   // nodal linear-flux kernel (Maxwell, APL_ALGO_AUTO)
  bool valu = false;     // cubature form on the VALU kernel (N <= 3)
  Operators ops;
  Geometry geo;
  std::vector<int> etetf, fcode, vmapP;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  long long dev_bytes = 0;
  std::string last_error;
  // device buffers
  double *d_geom = nullptr, *d_vert = nullptr, *d_I1D = nullptr;
  int *d_fcode = nullptr, *d_fmask = nullptr;
  void *d_A1 = nullptr, *d_A2 = nullptr, *d_A3 = nullptr;
  unsigned *d_err = nullptr;
  double *d_speed = nullptr, *d_t = nullptr;
  double *d_qa = nullptr, *d_qb = nullptr, *d_res = nullptr, *d_out = nullptr;
  // captured LSERK4 steps (even: A->B, odd: B->A)
  hipGraphExec_t step_exec[2] = {nullptr, nullptr};
  hipGraph_t step_graph[2] = {nullptr, nullptr};
  const double *graph_qa = nullptr;
  double graph_dt = 0.0;
  // adjoint: FP64 operator fragments, work buffers, checkpoints
  double *d_adjops[6] = {nullptr};   // Avc, Aart, Aast, Avct, Alt, Aift
  double *d_Vc = nullptr, *d_cub = nullptr;
  double *d_sbar = nullptr, *d_qbar = nullptr, *d_rbar = nullptr;
  double *d_sbar2 = nullptr;         // the reverse stepper's second face-weight buffer
  double *d_stg = nullptr;           // 2 stage buffers of forward_step
  double *d_seg = nullptr;           // stage inputs of one checkpoint segment
  long long seg_slots = 0;
  double *d_ckpt = nullptr;
  long long ckpt_slots = 0;
  double *d_opart = nullptr;
  // limiter: cell averages, gradients, weight and node tables, and the
  // states at the trace nodes
  double *d_lavg = nullptr, *d_lgrad = nullptr, *d_ltab = nullptr, *d_ltrc = nullptr;
  LimTrace ltr{};
  bool ltr_is_state = false;   // trace slots = nodes 0..Np-1 of a 5-variable state
  // pinned host words the flag and speed read-backs land in (two doubles)
  double *h_pin = nullptr;
  // this device's copy of apl_rhs_noalloc's scratch array (not owned)
  double *scratch = nullptr;
  std::string scratch_error;
};

// ---------------------------------------------------------------------------
// the device scratch of apl_rhs_noalloc
// ---------------------------------------------------------------------------
// apl_rhs_noalloc uses this file-scope array instead of the handle's
// buffers, so that it allocates nothing.  Layout in doubles: [0] the reduced
// speed maximum, [1] the error flags (its low 32 bits), [2, 32) unused,
// [32, 32 + nb) the per-batch speed maxima of one RHS over nb = ceil(K/16)
// batches.  Every GPU the library runs on holds its own copy (8 MB) until
// the process ends; calls that overlap on one GPU share it and race.
constexpr int NOALLOC_HEAD = 32;
constexpr long NOALLOC_MAX_BATCHES = 1L << 20;            // K <= 16 777 216
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_HEAD + NOALLOC_MAX_BATCHES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of this device's copy of `scratch` (the array's name is
// a host-side stand-in, not a device address), looked up once per device
// with that device current and kept for later handles.  A failure is kept
// on the handle and reported by apl_rhs_noalloc; nothing else needs it.
static void scratch_setup(apl_handle *h) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (h->device < 0 || h->device >= NOALLOC_MAX_DEVICES) {
    h->scratch_error = "device id " + std::to_string(h->device) + " beyond the scratch table";
    return;
  }
  if (!g_scratch_ptr[h->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      h->scratch_error = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
      return;
    }
    g_scratch_ptr[h->device] = (double *)p;
  }
  h->scratch = g_scratch_ptr[h->device];
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

static int fail(apl_handle *h, int code, const std::string &msg) {
  if (h) h->last_error = msg;
  else g_create_error = msg;
  return code;
}

#define HIPCHK(h, call)                                                            \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(h, APL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dev_upload(apl_handle *h, T **dst, const T *src, size_t n) {
  HIPCHK(h, hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(n, 1)));
  h->dev_bytes += (long long)(sizeof(T) * n);
  if (n) HIPCHK(h, hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
  return APL_OK;
}

static int dev_alloc(apl_handle *h, void **dst, size_t bytes) {
  HIPCHK(h, hipMalloc(dst, std::max<size_t>(bytes, 1)));
  h->dev_bytes += (long long)bytes;
  return APL_OK;
}

// The cubature form at N <= 3 (FP64 and FP32) runs on the VALU kernel
// unless APL_ALGO_CUBATURE_MFMA asks for the MFMA kernel: measured faster
// on the 2D Euler vortex at N = 3 (us per fused stage, VALU vs MFMA):
// K = 8192 FP64 18.4 vs 20.4, FP32 15.4 vs 16.8; K = 131072 FP64 96.2 vs
// 106.4, FP32 67.7 vs 75.2.
constexpr bool VALU_BY_DEFAULT = true;
// The nodal form of MAXWELL_TM3 runs on the VALU kernel under
// APL_ALGO_AUTO (FP64 and FP32).
constexpr bool NODAL_VALU_BY_DEFAULT = false;

static int nvars(int eqn) {
  switch (eqn) {
    case APL_EQN_EULER5: return 5;
    case APL_EQN_EULER4: return 4;
    case APL_EQN_PHMAXWELL8: return 8;
    case APL_EQN_MAXWELL_TM3: return 3;
  }
  return 0;
}

static int dispatch_rhs(apl_handle *h, const RhsArgs &a) {
  const int nb = (h->K + 15) / 16;
  const apl_params &p = h->params;
  // kernel: 0 cubature form on MFMA, 1 nodal form on MFMA, 2 cubature form
  // on VALU, 3 nodal form on VALU
  const int lin = h->linear ? (h->valu ? 3 : 1) : (h->valu ? 2 : 0);
  int rc;
  switch (p.order) {
    case 1: rc = launch_rhs<1>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 2: rc = launch_rhs<2>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 3: rc = launch_rhs<3>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 4: rc = launch_rhs<4>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 5: rc = launch_rhs<5>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 6: rc = launch_rhs<6>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 7: rc = launch_rhs<7>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    case 8: rc = launch_rhs<8>(p.eqn, p.flux, p.precision, lin, h->stream, a, nb); break;
    default: rc = APL_ERR_ORDER;
  }
  if (rc) return fail(h, rc, "RHS kernel launch failed");
  return APL_OK;
}

static RhsArgs base_args(apl_handle *h) {
  RhsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.K = h->K;
  a.geom = h->d_geom;
  a.vert = h->d_vert;
  a.fcode = h->d_fcode;
  a.A1 = h->d_A1;
  a.A2 = h->d_A2;
  a.A3 = h->d_A3;
  a.I1D = h->d_I1D;
  a.fmask = h->d_fmask;
  a.err = h->d_err;
  const apl_params &p = h->params;
  a.ph.gamma = p.gas_gamma;
  a.ph.min_dens = p.min_density;
  a.ph.min_pres = p.min_pressure;
  a.ph.c0 = p.c0;
  a.ph.chi = p.chi;
  a.ph.gph = p.gamma_ph;
  a.bc.n = p.n_bc;
  for (int i = 0; i < APL_MAX_BC; ++i) {
    a.bc.kind[i] = p.bc[i].kind;
    for (int j = 0; j < 8; ++j) a.bc.p[i][j] = p.bc[i].p[j];
  }
  return a;
}

// Status of a set of kernel error flags.
static int flags_status(apl_handle *h, unsigned ef) {
  if (ef & EF_DENSITY) return fail(h, APL_ERR_NEG_DENSITY, "negative density in the Euler flux");
  if (ef & EF_PRESSURE) return fail(h, APL_ERR_NEG_PRESSURE, "negative pressure in the Euler flux");
  if (ef & EF_NAN) return fail(h, APL_ERR_NAN, "NaN in the RHS");
  return APL_OK;
}

// Read the error flags after everything enqueued so far (one stream
// synchronisation).
static int check_flags(apl_handle *h) {
  unsigned *pf = (unsigned *)h->h_pin;
  HIPCHK(h, hipMemcpyAsync(pf, h->d_err, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return flags_status(h, *pf);
}

// The folded operators of the cubature form: A_r = V V^T Dr_c^T diag(w)
// and A_s likewise [Np x Nc], L = V V^T I_f^T diag(gw) [Np x 3Ng].
static void fold_operators(const Operators &o, std::vector<double> &Ar, std::vector<double> &As,
                           std::vector<double> &L) {
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng;
  Ar.assign((size_t)Np * Nc, 0.0);
  As.assign((size_t)Np * Nc, 0.0);
  L.assign((size_t)Np * 3 * Ng, 0.0);
  for (int n = 0; n < Np; ++n) {
    for (int c = 0; c < Nc; ++c) {
      double sr = 0.0, ss = 0.0;
      for (int j = 0; j < Np; ++j) {
        sr += o.VVT[(size_t)n * Np + j] * o.Drc[(size_t)c * Np + j];
        ss += o.VVT[(size_t)n * Np + j] * o.Dsc[(size_t)c * Np + j];
      }
      Ar[(size_t)n * Nc + c] = sr * o.cw[c];
      As[(size_t)n * Nc + c] = ss * o.cw[c];
    }
    for (int k = 0; k < 3 * Ng; ++k) {
      double s = 0.0;
      for (int j = 0; j < Np; ++j) s += o.VVT[(size_t)n * Np + j] * o.If[(size_t)k * Np + j];
      L[(size_t)n * 3 * Ng + k] = s * o.gw[k % Ng];
    }
  }
}

// Fold and permute the operators into MFMA fragments (see apl_device.h).
// LT is the operand type whose accumulator layout the fragments follow.
template <typename T, typename LT = T>
static void make_fragments(const Operators &o, std::vector<T> &A1, std::vector<T> &A2,
                           std::vector<T> &A3) {
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng;
  const int NKS1 = (Np + 3) / 4, NCH = (Nc + 15) / 16, NRT = (Np + 15) / 16, NKF = (3 * Ng + 3) / 4;
  std::vector<double> Ar, As, L;
  fold_operators(o, Ar, As, L);
  A1.assign((size_t)NCH * NKS1 * 64, T(0));
  for (int cc = 0; cc < NCH; ++cc)
    for (int ks = 0; ks < NKS1; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int c = 16 * cc + (l & 15), j = 4 * ks + (l >> 4);
        if (c < Nc && j < Np) A1[((size_t)cc * NKS1 + ks) * 64 + l] = T(o.Vc[(size_t)c * Np + j]);
      }
  A2.assign((size_t)NRT * NCH * 64 * 8, T(0));
  for (int rt = 0; rt < NRT; ++rt)
    for (int cc = 0; cc < NCH; ++cc)
      for (int l = 0; l < 64; ++l)
        for (int r = 0; r < 4; ++r) {
          const int n = 16 * rt + (l & 15), c = 16 * cc + acc_row<LT>(l >> 4, r);
          if (n < Np && c < Nc) {
            const size_t b = (((size_t)rt * NCH + cc) * 64 + l) * 8;
            A2[b + r] = T(Ar[(size_t)n * Nc + c]);
            A2[b + 4 + r] = T(As[(size_t)n * Nc + c]);
          }
        }
  A3.assign((size_t)NRT * NKF * 64, T(0));
  for (int rt = 0; rt < NRT; ++rt)
    for (int ks = 0; ks < NKF; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int n = 16 * rt + (l & 15), k = 4 * ks + (l >> 4);
        if (n < Np && k < 3 * Ng) A3[((size_t)rt * NKF + ks) * 64 + l] = T(-L[(size_t)n * 3 * Ng + k]);
      }
}

// Point tables of the VALU cubature kernel (apl_valu.h): row c of A1 holds
// V_c[c, 0..Np), A_r[0..Np, c), A_s[0..Np, c); row k of A3 holds -L[0..Np, k).
template <typename T>
static void make_valu_tables(const Operators &o, std::vector<T> &A1, std::vector<T> &A3) {
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng;
  std::vector<double> Ar, As, L;
  fold_operators(o, Ar, As, L);
  A1.assign((size_t)Nc * 3 * Np, T(0));
  for (int c = 0; c < Nc; ++c)
    for (int n = 0; n < Np; ++n) {
      A1[(size_t)c * 3 * Np + n] = T(o.Vc[(size_t)c * Np + n]);
      A1[(size_t)c * 3 * Np + Np + n] = T(Ar[(size_t)n * Nc + c]);
      A1[(size_t)c * 3 * Np + 2 * Np + n] = T(As[(size_t)n * Nc + c]);
    }
  A3.assign((size_t)3 * Ng * Np, T(0));
  for (int k = 0; k < 3 * Ng; ++k)
    for (int n = 0; n < Np; ++n) A3[(size_t)k * Np + n] = T(-L[(size_t)n * 3 * Ng + k]);
}

// The folded operators of the nodal linear-flux form: D0 = Dt_r = A_r V_c,
// D1 = Dt_s [Np x Np] and Lt = L blkdiag(I1D_f) [Np x 3Nfp] (see apl_device.h).
static void fold_linear_operators(const Operators &o, std::vector<double> (&D)[2],
                                  std::vector<double> &Lt) {
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng, Nfp = o.Nfp;
  D[0].assign((size_t)Np * Np, 0.0);
  D[1].assign((size_t)Np * Np, 0.0);
  Lt.assign((size_t)Np * 3 * Nfp, 0.0);
  for (int n = 0; n < Np; ++n) {
    // rows of A_r, A_s (V V^T Dr_c^T diag(w)) and L (V V^T I_f^T diag(gw))
    std::vector<double> ar(Nc), as(Nc), lr(3 * Ng);
    for (int c = 0; c < Nc; ++c) {
      double sr = 0.0, ss = 0.0;
      for (int j = 0; j < Np; ++j) {
        sr += o.VVT[(size_t)n * Np + j] * o.Drc[(size_t)c * Np + j];
        ss += o.VVT[(size_t)n * Np + j] * o.Dsc[(size_t)c * Np + j];
      }
      ar[c] = sr * o.cw[c];
      as[c] = ss * o.cw[c];
    }
    for (int k = 0; k < 3 * Ng; ++k) {
      double sl = 0.0;
      for (int j = 0; j < Np; ++j) sl += o.VVT[(size_t)n * Np + j] * o.If[(size_t)k * Np + j];
      lr[k] = sl * o.gw[k % Ng];
    }
    for (int j = 0; j < Np; ++j) {
      double sr = 0.0, ss = 0.0;
      for (int c = 0; c < Nc; ++c) {
        sr += ar[c] * o.Vc[(size_t)c * Np + j];
        ss += as[c] * o.Vc[(size_t)c * Np + j];
      }
      D[0][(size_t)n * Np + j] = sr;
      D[1][(size_t)n * Np + j] = ss;
    }
    for (int f = 0; f < 3; ++f)
      for (int i = 0; i < Nfp; ++i) {
        double sl = 0.0;
        for (int g = 0; g < Ng; ++g)
          sl += lr[f * Ng + g] * o.If[((size_t)f * Ng + g) * Np + o.fmask[f * Nfp + i]];
        Lt[(size_t)n * 3 * Nfp + f * Nfp + i] = sl;
      }
  }
}

// Tables of the nodal VALU kernel (apl_valu.h): A1 = [Dt_r; Dt_s] and
// A3 = -Lt, row-major.
template <typename T>
static void make_valu_linear_tables(const Operators &o, std::vector<T> &A1, std::vector<T> &A3) {
  std::vector<double> D[2], Lt;
  fold_linear_operators(o, D, Lt);
  A1.resize(2 * D[0].size());
  for (size_t i = 0; i < D[0].size(); ++i) {
    A1[i] = T(D[0][i]);
    A1[D[0].size() + i] = T(D[1][i]);
  }
  A3.resize(Lt.size());
  for (size_t i = 0; i < Lt.size(); ++i) A3[i] = T(-Lt[i]);
}

// Fragments of the nodal linear-flux kernel: A1 = [Dt_r; Dt_s] with
// Dt_h = A_h V_c, A3 = -Lt with Lt = L blkdiag(I1D_f) (see apl_device.h).
// A1 holds 16-row tiles [h][rt] of each derivative, or, `stacked` (the FP64
// kernel), of one operand with Dt_r in rows 0..Np-1 and Dt_s from row Np
// rounded up to a multiple of 4.
template <typename T>
static void make_linear_fragments(const Operators &o, std::vector<T> &A1, std::vector<T> &A3,
                                  bool stacked) {
  const int Np = o.Np, Nfp = o.Nfp;
  const int NKS1 = (Np + 3) / 4, NRT = (Np + 15) / 16, NKP = (3 * Nfp + 3) / 4;
  const int OFF = (Np + 3) / 4 * 4;
  const int NT = stacked ? (OFF + Np + 15) / 16 : 2 * NRT;
  std::vector<double> D[2], Lt;
  fold_linear_operators(o, D, Lt);
  A1.assign((size_t)NT * NKS1 * 64, T(0));
  for (int t = 0; t < NT; ++t)
    for (int ks = 0; ks < NKS1; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int row = 16 * t + (l & 15), j = 4 * ks + (l >> 4);
        const int hh = stacked ? (row >= OFF ? 1 : 0) : t / NRT;
        const int n = stacked ? row - hh * OFF : 16 * (t % NRT) + (l & 15);
        if (n < Np && j < Np) A1[((size_t)t * NKS1 + ks) * 64 + l] = T(D[hh][(size_t)n * Np + j]);
      }
  A3.assign((size_t)NRT * NKP * 64, T(0));
  for (int rt = 0; rt < NRT; ++rt)
    for (int ks = 0; ks < NKP; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int n = 16 * rt + (l & 15), k = 4 * ks + (l >> 4);
        if (n < Np && k < 3 * Nfp) A3[((size_t)rt * NKP + ks) * 64 + l] = T(-Lt[(size_t)n * 3 * Nfp + k]);
      }
}

template <typename T>
static int upload_fragments(apl_handle *h) {
  std::vector<T> A1, A2, A3;
  if (h->linear && h->valu) {
    make_valu_linear_tables<T>(h->ops, A1, A3);
    A2.assign(1, T(0));
  } else if (h->linear) {
    make_linear_fragments<T>(h->ops, A1, A3, std::is_same<T, double>::value);
    A2.assign(1, T(0));
  } else if (h->valu) {
    make_valu_tables<T>(h->ops, A1, A3);
    A2.assign(1, T(0));
  } else {
    make_fragments<T>(h->ops, A1, A2, A3);
  }
  int rc;
  if ((rc = dev_upload<T>(h, (T **)&h->d_A1, A1.data(), A1.size()))) return rc;
  if ((rc = dev_upload<T>(h, (T **)&h->d_A2, A2.data(), A2.size()))) return rc;
  if ((rc = dev_upload<T>(h, (T **)&h->d_A3, A3.data(), A3.size()))) return rc;
  return APL_OK;
}

// BF16 (bfloat16) bits of x, rounded to nearest even; NaN and Inf pass.
static uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// One lane's split-BF16 fragment of a 16x16x32 step: 8 hi then 8 lo values
// (the layout of Bf8x2 in apl_device.h).
struct Bf16Pair {
  uint16_t v[16];
};

static_assert(sizeof(Bf16Pair) == sizeof(apl::Bf8x2), "one lane's hi/lo fragment is 32 bytes");

static void split_into(double x, Bf16Pair &p, int item) {
  const float xf = (float)x;
  const uint16_t hi = bf16_rne(xf);
  p.v[item] = hi;
  p.v[8 + item] = bf16_rne(xf - bf16_value(hi));
}

// FP32-layout fragments [outer][nks][64] (one value per lane per 4-row
// k-step) regrouped into 16x16x32 steps [outer][(nks+7)/8][64]: item i of
// step kk is the value of k-step 8 kk + i (zero past nks).
static std::vector<Bf16Pair> regroup_bf16(const std::vector<double> &F, int outer, int nks) {
  const int nk8 = (nks + 7) / 8;
  std::vector<Bf16Pair> out((size_t)outer * nk8 * 64);
  for (auto &p : out) std::memset(p.v, 0, sizeof(p.v));
  for (int o = 0; o < outer; ++o)
    for (int ks = 0; ks < nks; ++ks)
      for (int l = 0; l < 64; ++l)
        split_into(F[((size_t)o * nks + ks) * 64 + l], out[((size_t)o * nk8 + ks / 8) * 64 + l], ks % 8);
  return out;
}

// Split-BF16x3 fragments: the FP32 fragments (FP32 accumulator layout) in
// FP64, regrouped per 16x16x32 step and split into hi/lo.
static int upload_fragments_bf16x3(apl_handle *h) {
  const Operators &o = h->ops;
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng, Nfp = o.Nfp;
  const int NKS1 = (Np + 3) / 4, NCH = (Nc + 15) / 16, NRT = (Np + 15) / 16;
  std::vector<double> A1, A2, A3;
  std::vector<Bf16Pair> B1, B2, B3;
  if (h->linear) {
    make_linear_fragments<double>(o, A1, A3, false);
    B1 = regroup_bf16(A1, 2 * NRT, NKS1);
    B3 = regroup_bf16(A3, NRT, (3 * Nfp + 3) / 4);
    B2.resize(1);
    std::memset(B2[0].v, 0, sizeof(B2[0].v));
  } else {
    make_fragments<double, float>(o, A1, A2, A3);
    B1 = regroup_bf16(A1, NCH, NKS1);
    // A_r|A_s fragments already hold 8 values per lane: items 0-3 (A_r) and
    // 4-7 (A_s) of one step
    B2.resize((size_t)NRT * NCH * 64);
    for (size_t i = 0; i < B2.size(); ++i)
      for (int it = 0; it < 8; ++it) split_into(A2[i * 8 + it], B2[i], it);
    B3 = regroup_bf16(A3, NRT, (3 * Ng + 3) / 4);
  }
  int rc;
  if ((rc = dev_upload<Bf16Pair>(h, (Bf16Pair **)&h->d_A1, B1.data(), B1.size()))) return rc;
  if ((rc = dev_upload<Bf16Pair>(h, (Bf16Pair **)&h->d_A2, B2.data(), B2.size()))) return rc;
  if ((rc = dev_upload<Bf16Pair>(h, (Bf16Pair **)&h->d_A3, B3.data(), B3.size()))) return rc;
  return APL_OK;
}

static void free_graphs(apl_handle *h) {
  for (int i = 0; i < 2; ++i) {
    if (h->step_exec[i]) (void)hipGraphExecDestroy(h->step_exec[i]);
    if (h->step_graph[i]) (void)hipGraphDestroy(h->step_graph[i]);
    h->step_exec[i] = nullptr;
    h->step_graph[i] = nullptr;
  }
  h->graph_qa = nullptr;
}

static void free_handle(apl_handle *h) {
  if (!h) return;
  free_graphs(h);
  void *bufs[] = {h->d_geom, h->d_vert, h->d_I1D, h->d_fcode, h->d_fmask, h->d_A1, h->d_A2,
                  h->d_A3, h->d_speed, h->d_t, h->d_qa, h->d_qb, h->d_res, h->d_out,
                  h->d_ckpt, h->d_Vc, h->d_cub, h->d_sbar, h->d_sbar2, h->d_qbar, h->d_rbar, h->d_stg,
                  h->d_seg, h->d_opart, h->d_lavg, h->d_lgrad, h->d_ltab, h->d_ltrc};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->h_pin) (void)hipHostFree(h->h_pin);
  for (double *b : h->d_adjops)
    if (b) (void)hipFree(b);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static size_t state_len(const apl_handle *h) { return (size_t)h->K * h->ops.Np * h->M; }

// Whether two state-sized device ranges overlap.  The RHS and adjoint kernels
// gather neighbour face nodes from their input while other waves store their
// output, so an output that overlaps an input gives wrong results silently.
static bool states_overlap(const apl_handle *h, const void *a, const void *b) {
  const size_t bytes = sizeof(double) * state_len(h);
  const char *pa = (const char *)a, *pb = (const char *)b;
  return pa < pb + bytes && pb < pa + bytes;
}

static int ensure_work(apl_handle *h) {
  const size_t bytes = sizeof(double) * state_len(h);
  int rc;
  if (!h->d_qa && (rc = dev_alloc(h, (void **)&h->d_qa, bytes))) return rc;
  if (!h->d_qb && (rc = dev_alloc(h, (void **)&h->d_qb, bytes))) return rc;
  if (!h->d_res && (rc = dev_alloc(h, (void **)&h->d_res, bytes))) return rc;
  if (!h->d_out && (rc = dev_alloc(h, (void **)&h->d_out, bytes))) return rc;
  return APL_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *apl_status_string(int s) {
  switch (s) {
    case APL_OK: return "ok";
    case APL_ERR_ARG: return "invalid argument";
    case APL_ERR_ORDER: return "polynomial order outside 1..8";
    case APL_ERR_MESH_JACOBIAN: return "element with J <= 0";
    case APL_ERR_MESH_EDGE: return "element with a zero-length edge";
    case APL_ERR_MESH_TOPOLOGY: return "face shared by more than two elements";
    case APL_ERR_HIP: return "HIP runtime error";
    case APL_ERR_NAN: return "NaN in the RHS";
    case APL_ERR_NEG_DENSITY: return "negative density";
    case APL_ERR_NEG_PRESSURE: return "negative pressure";
    case APL_ERR_ALLOC: return "allocation failed";
    case APL_ERR_UNSUPPORTED: return "unsupported configuration";
  }
  return "unknown status";
}

const char *apl_last_error(const apl_handle *h) {
  return h ? h->last_error.c_str() : g_create_error.c_str();
}

int apl_mesh_structured(int n, double x0, double x1, double y0, double y1, double jitter,
                        unsigned long long seed, int periodic, double *VX, double *VY,
                        int *EToV, int *vert_id, int *face_tag) {
  if (n < 1 || !VX || !VY || !EToV || (periodic && n < 3) || !(x1 > x0) || !(y1 > y0) ||
      jitter < 0.0 || jitter >= 0.5)
    return APL_ERR_ARG;
  structured_mesh(n, x0, x1, y0, y1, jitter, seed, periodic != 0, VX, VY, EToV, vert_id, face_tag);
  return APL_OK;
}

// The integer arrays of a mesh: EToV entries in [0, Nv) and face tags >= 0.
// A negative tag would be stored as a positive "neighbour" (|tag|, |tag|) in
// ETETF (FacePair2d writes -tag on boundary faces) and read as an interior
// face.  Returns the reason, or NULL when the arrays are valid.
static const char *bad_mesh_ids(const apl_mesh *mesh) {
  for (long i = 0; i < 3L * mesh->K; ++i)
    if (mesh->EToV[i] < 0 || mesh->EToV[i] >= mesh->Nv) return "EToV entry out of range";
  if (mesh->face_tag)
    for (long i = 0; i < 3L * mesh->K; ++i)
      if (mesh->face_tag[i] < 0) return "negative face tag";
  return nullptr;
}

int apl_create(const apl_mesh *mesh, const apl_params *params, apl_handle **out) {
  g_create_error.clear();
  if (!mesh || !params || !out) return fail(nullptr, APL_ERR_ARG, "null argument");
  *out = nullptr;
  if (params->order < 1 || params->order > 8) return fail(nullptr, APL_ERR_ORDER, "order outside 1..8");
  const int M = nvars(params->eqn);
  if (!M) return fail(nullptr, APL_ERR_ARG, "unknown equation kind");
  const bool euler = params->eqn == APL_EQN_EULER5 || params->eqn == APL_EQN_EULER4;
  if (params->flux != APL_FLUX_LF &&
      !(params->flux == APL_FLUX_UPWIND && params->eqn == APL_EQN_MAXWELL_TM3) &&
      !((params->flux == APL_FLUX_HLL || params->flux == APL_FLUX_ROE ||
         params->flux == APL_FLUX_WAVE) && euler))
    return fail(nullptr, APL_ERR_ARG,
                "flux: upwind is for MAXWELL_TM3 only, HLL, Roe and Wave for the Euler equations only");
  if (params->precision != APL_PREC_FP64 && params->precision != APL_PREC_FP32_MFMA &&
      params->precision != APL_PREC_BF16X3)
    return fail(nullptr, APL_ERR_ARG, "unknown precision");
  if (params->algorithm < APL_ALGO_AUTO || params->algorithm > APL_ALGO_NODAL_VALU)
    return fail(nullptr, APL_ERR_ARG, "unknown algorithm");
  const bool maxwell = params->eqn == APL_EQN_MAXWELL_TM3 || params->eqn == APL_EQN_PHMAXWELL8;
  if ((params->algorithm == APL_ALGO_NODAL_MFMA || params->algorithm == APL_ALGO_NODAL_VALU) &&
      !maxwell)
    return fail(nullptr, APL_ERR_ARG, "the nodal form needs a linear flux (Maxwell)");
  if (params->algorithm == APL_ALGO_NODAL_VALU &&
      (params->eqn != APL_EQN_MAXWELL_TM3 || params->precision == APL_PREC_BF16X3))
    return fail(nullptr, APL_ERR_UNSUPPORTED,
                "the VALU nodal kernel covers MAXWELL_TM3 in FP64 and FP32");
  if (params->algorithm == APL_ALGO_CUBATURE_VALU &&
      (params->order > 3 || params->precision == APL_PREC_BF16X3))
    return fail(nullptr, APL_ERR_UNSUPPORTED,
                "the VALU cubature kernel covers N <= 3 in FP64 and FP32");
  if (params->eqn == APL_EQN_MAXWELL_TM3 && (params->chi != 0.0 || params->gamma_ph != 0.0))
    return fail(nullptr, APL_ERR_ARG, "MAXWELL_TM3 requires chi == gamma == 0");
  if (params->n_bc < 0 || params->n_bc > APL_MAX_BC) return fail(nullptr, APL_ERR_ARG, "n_bc");
  for (int i = 0; i < params->n_bc; ++i)
    if (params->eqn == APL_EQN_EULER4 && params->bc[i].kind == APL_BC_EULER_DIRICHLET &&
        params->bc[i].p[3] != 0.0)
      return fail(nullptr, APL_ERR_ARG, "EULER4 has no z momentum: Dirichlet zvelocity must be 0");
  if (params->limiter) {
    if (params->limiter != 1 || !euler)
      return fail(nullptr, APL_ERR_ARG, "limiter: the Hesthaven-Warburton limiter is for the Euler equations");
    if (params->order > 4)
      return fail(nullptr, APL_ERR_UNSUPPORTED,
                  "limiter: orders above 4 read outside the face-average array in the reference");
    for (int i = 0; i < params->n_bc; ++i)
      if (params->bc[i].kind == APL_BC_ISENTROPIC_VORTEX)
        return fail(nullptr, APL_ERR_UNSUPPORTED,
                    "limiter: the isentropic-vortex state reads three coordinates from the "
                    "limiter's two-entry array in the reference");
  }
  if (mesh->K < 1 || mesh->Nv < 3 || !mesh->VX || !mesh->VY || !mesh->EToV)
    return fail(nullptr, APL_ERR_ARG, "empty or incomplete mesh");
  if (const char *why = bad_mesh_ids(mesh)) return fail(nullptr, APL_ERR_ARG, why);

  apl_handle *h = new apl_handle();
  h->params = *params;
  h->linear = maxwell && (params->algorithm == APL_ALGO_AUTO ||
                          params->algorithm == APL_ALGO_NODAL_MFMA ||
                          params->algorithm == APL_ALGO_NODAL_VALU);
  if (h->linear)
    h->valu = params->algorithm == APL_ALGO_NODAL_VALU ||
              (params->algorithm == APL_ALGO_AUTO && NODAL_VALU_BY_DEFAULT &&
               params->eqn == APL_EQN_MAXWELL_TM3 && params->precision != APL_PREC_BF16X3);
  else
    h->valu = params->order <= 3 && params->precision != APL_PREC_BF16X3 &&
              (params->algorithm == APL_ALGO_CUBATURE_VALU ||
               (params->algorithm != APL_ALGO_CUBATURE_MFMA && VALU_BY_DEFAULT));
  h->K = mesh->K;
  h->M = M;
  h->Nv = mesh->Nv;
  auto bail = [&](int rc) {
    g_create_error = h->last_error;
    free_handle(h);
    return rc;
  };
  if (!build_operators(params->order, h->ops)) {
    h->last_error = "operator construction failed";
    return bail(APL_ERR_ORDER);
  }
  const Operators &o = h->ops;
  int bad = -1;
  int rc = build_geometry(o, h->K, mesh->VX, mesh->VY, mesh->EToV, h->geo, bad);
  if (rc) {
    h->last_error = "bad element " + std::to_string(bad);
    return bail(rc);
  }
  rc = face_pair(h->K, mesh->Nv, mesh->EToV, mesh->vert_id, mesh->face_tag, h->etetf);
  if (rc) {
    h->last_error = "a face is shared by more than two elements";
    return bail(rc);
  }
  // face codes and the vmapP index map
  const int Np = o.Np, Nfp = o.Nfp;
  h->fcode.resize((size_t)3 * h->K);
  h->vmapP.assign((size_t)3 * h->K * Nfp, -1);
  for (int k = 0; k < h->K; ++k)
    for (int f = 0; f < 3; ++f) {
      const int e2 = h->etetf[(size_t)6 * k + 2 * f], f2 = h->etetf[(size_t)6 * k + 2 * f + 1];
      if (f2 >= 0) {
        h->fcode[(size_t)3 * k + f] = e2 * 4 + f2;
        for (int i = 0; i < Nfp; ++i)
          h->vmapP[((size_t)3 * k + f) * Nfp + i] = e2 * Np + o.fmask[f2 * Nfp + (Nfp - 1 - i)];
      } else {
        // boundary condition index |tag| - 1 (wxnodaldg2dmethod.cc:418,569)
        const int idx = std::abs(f2) - 1;
        if (idx >= params->n_bc || params->bc[idx].kind == APL_BC_NONE) {
          h->last_error = "boundary face of element " + std::to_string(k) +
                          " has no boundary condition (tag " + std::to_string(std::abs(f2)) + ")";
          return bail(APL_ERR_MESH_TOPOLOGY);
        }
        h->fcode[(size_t)3 * k + f] = -(idx + 1);
      }
    }
  if (h->K > (1 << 29)) {
    h->last_error = "K too large for 32-bit face codes";
    return bail(APL_ERR_ARG);
  }

  // device side
  hipError_t he = hipGetDevice(&h->device);
  if (he != hipSuccess) {
    h->last_error = std::string("hipGetDevice: ") + hipGetErrorString(he);
    return bail(APL_ERR_HIP);
  }
  he = hipStreamCreate(&h->stream);
  if (he != hipSuccess) {
    h->last_error = std::string("hipStreamCreate: ") + hipGetErrorString(he);
    return bail(APL_ERR_HIP);
  }
  h->own_stream = true;
  scratch_setup(h);
  std::vector<double> g((size_t)h->K * G_WORDS, 0.0);
  for (int k = 0; k < h->K; ++k) {
    const double *s = &h->geo.geom[(size_t)14 * k];
    double *d = &g[(size_t)G_WORDS * k];
    const double J = s[4];
    d[G_JRX] = J * s[0];
    d[G_JRY] = J * s[2];
    d[G_JSX] = J * s[1];
    d[G_JSY] = J * s[3];
    d[G_INVJ] = 1.0 / J;
    d[G_J] = J;
    for (int f = 0; f < 3; ++f) {
      d[G_NX + f] = s[5 + 3 * f];
      d[G_NY + f] = s[6 + 3 * f];
      d[G_SJ + f] = s[7 + 3 * f];
    }
  }
  // face interpolation: I1D[f][g][i] = I_f[f Ng + g][Fmask[f][i]], then the
  // Gauss abscissae
  const int Ng = o.Ng;
  std::vector<double> I1D((size_t)3 * Ng * Nfp + Ng);
  for (int f = 0; f < 3; ++f)
    for (int gq = 0; gq < Ng; ++gq)
      for (int i = 0; i < Nfp; ++i)
        I1D[((size_t)f * Ng + gq) * Nfp + i] = o.If[((size_t)f * Ng + gq) * Np + o.fmask[f * Nfp + i]];
  for (int gq = 0; gq < Ng; ++gq) I1D[(size_t)3 * Ng * Nfp + gq] = o.gz[gq];

  if ((rc = dev_upload<double>(h, &h->d_geom, g.data(), g.size()))) return bail(rc);
  if ((rc = dev_upload<double>(h, &h->d_vert, h->geo.vert.data(), h->geo.vert.size()))) return bail(rc);
  if ((rc = dev_upload<int>(h, &h->d_fcode, h->fcode.data(), h->fcode.size()))) return bail(rc);
  if ((rc = dev_upload<int>(h, &h->d_fmask, o.fmask.data(), o.fmask.size()))) return bail(rc);
  if ((rc = dev_upload<double>(h, &h->d_I1D, I1D.data(), I1D.size()))) return bail(rc);
  if (params->precision == APL_PREC_FP64) rc = upload_fragments<double>(h);
  else if (params->precision == APL_PREC_FP32_MFMA) rc = upload_fragments<float>(h);
  else rc = upload_fragments_bf16x3(h);
  if (rc) return bail(rc);
  if (hipHostMalloc((void **)&h->h_pin, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    h->h_pin = nullptr;
    h->last_error = "pinned host allocation failed";
    return bail(APL_ERR_HIP);
  }
  if ((rc = dev_alloc(h, (void **)&h->d_speed, sizeof(double) * ((h->K + 15) / 16)))) return bail(rc);
  // d_t: [0] device clock, [1] a step's speed maximum, [2] the error flags
  // (an unsigned), so that one copy reads back both of a step's results
  if ((rc = dev_alloc(h, (void **)&h->d_t, sizeof(double) * 4))) return bail(rc);
  h->d_err = (unsigned *)(h->d_t + 2);
  if (params->limiter) {
    // the limiter's face ends are the vertex nodes shared by consecutive
    // faces: Fmask[f][Nfp-1] == Fmask[f+1][0]
    const int *F = o.fmask.data();
    if (F[Nfp - 1] != F[Nfp] || F[2 * Nfp - 1] != F[2 * Nfp] || F[3 * Nfp - 1] != F[0]) {
      h->last_error = "limiter: face node lists do not chain at the vertices";
      return bail(APL_ERR_UNSUPPORTED);
    }
    // trace slots: the nodes the gradient pass reads at the three face ends
    // p = Fmask[0][0], Fmask[0][Nfp-1], Fmask[1][Nfp-1], each taken as the
    // face-node position (face p / Nfp, position j = p % Nfp): its own node
    // Fmask[p], and node Nfp-1-j of whichever face nf a neighbour shares
    LimTrace &tr = h->ltr;
    tr.s = 0;
    auto slot = [&tr](int node) {
      for (int q = 0; q < tr.s; ++q)
        if (tr.node[q] == node) return q;
      tr.node[tr.s] = node;
      return tr.s++;
    };
    const int pe[3] = {0, Nfp - 1, 2 * Nfp - 1};
    for (int i = 0; i < 3; ++i) {
      const int p = F[pe[i]];
      if (p >= 3 * Nfp) {
        h->last_error = "limiter: a face end falls outside the face node lists";
        return bail(APL_ERR_UNSUPPORTED);
      }
      const int fe = p / Nfp, j = p - fe * Nfp;
      tr.face[i] = fe;
      tr.own[i] = slot(F[p]);
      for (int nf = 0; nf < 3; ++nf) tr.nbr[i][nf] = slot(F[nf * Nfp + Nfp - 1 - j]);
    }
    // ave[Np] = 0.5 sum_m Mass[m][k], Mass = (V V^T)^-1 (wxnodaldggeometry2d.h:84-98),
    // then the centred reference coordinates r_k - sum_j ave_j r_j and
    // s_k - sum_j ave_j s_j
    std::vector<double> mass((size_t)Np * Np), tab((size_t)3 * Np);
    if (!invert(Np, o.VVT.data(), mass.data())) {
      h->last_error = "mass matrix inversion failed";
      return bail(APL_ERR_ORDER);
    }
    double rbar = 0.0, sbar = 0.0;
    for (int k = 0; k < Np; ++k) {
      double a = 0.0;
      for (int m = 0; m < Np; ++m) a += 0.5 * mass[(size_t)k + (size_t)m * Np];
      tab[k] = a;
      rbar += a * o.r[k];
      sbar += a * o.s[k];
    }
    for (int k = 0; k < Np; ++k) {
      tab[Np + k] = o.r[k] - rbar;
      tab[2 * Np + k] = o.s[k] - sbar;
    }
    if ((rc = dev_upload<double>(h, &h->d_ltab, tab.data(), tab.size()))) return bail(rc);
    if ((rc = dev_alloc(h, (void **)&h->d_lavg, sizeof(double) * 8 * (size_t)h->K))) return bail(rc);
    {
      LimArgs pa{};
      pa.K = h->K;
      pa.vert = h->d_vert;
      pa.tab = h->d_ltab;
      pa.avg = h->d_lavg;
      if (launch_limiter_patch(h->stream, pa) != APL_OK || hipStreamSynchronize(h->stream) != hipSuccess) {
        h->last_error = "limiter patch geometry kernel failed";
        return bail(APL_ERR_HIP);
      }
    }
    if ((rc = dev_alloc(h, (void **)&h->d_lgrad, sizeof(double) * 40 * (size_t)h->K))) return bail(rc);
    h->ltr_is_state = tr.s == Np && params->eqn == APL_EQN_EULER5;
    for (int q = 0; q < tr.s && h->ltr_is_state; ++q) h->ltr_is_state = tr.node[q] == q;
    if (!h->ltr_is_state &&
        (rc = dev_alloc(h, (void **)&h->d_ltrc, sizeof(double) * 5 * (size_t)tr.s * h->K)))
      return bail(rc);
  }
  *out = h;
  return APL_OK;
}

// The limiter's three passes on the handle's stream (qout may equal qin).
// clear_err: the first pass zeroes the error flags (apl_limit's own call;
// the steppers keep the flags of the step's earlier kernels).
static int enqueue_limit(apl_handle *h, const double *qin, double *qout, bool clear_err = false) {
  LimArgs a;
  a.K = h->K;
  a.qin = qin;
  a.qout = qout;
  a.avg = h->d_lavg;
  a.grad = h->d_lgrad;
  a.geom = h->d_geom;
  a.vert = h->d_vert;
  a.fcode = h->d_fcode;
  a.trc = h->d_ltrc;
  a.trc_rd = h->ltr_is_state ? qin : h->d_ltrc;
  a.trc_write = h->ltr_is_state ? 0 : 1;
  a.clear_err = clear_err ? 1 : 0;
  a.tr = h->ltr;
  a.tab = h->d_ltab;
  a.gamma = h->params.lim_gas_gamma;
  a.min_dens = h->params.lim_min_density;
  a.min_pres = h->params.lim_min_pressure;
  a.err = h->d_err;
  a.bc = base_args(h).bc;
  int rc;
  switch (h->params.order) {
    case 1: rc = launch_limiter<1>(h->params.eqn, h->stream, a); break;
    case 2: rc = launch_limiter<2>(h->params.eqn, h->stream, a); break;
    case 3: rc = launch_limiter<3>(h->params.eqn, h->stream, a); break;
    case 4: rc = launch_limiter<4>(h->params.eqn, h->stream, a); break;
    default: rc = APL_ERR_UNSUPPORTED;
  }
  if (rc) return fail(h, rc, "limiter launch");
  return APL_OK;
}

int apl_limit(apl_handle *h, const double *q_in, double *q_out, int where) {
  if (!h || !q_in || !q_out) return APL_ERR_ARG;
  if (!h->params.limiter) return fail(h, APL_ERR_ARG, "apl_limit: handle created without the limiter");
  int rc;
  if ((rc = ensure_work(h))) return rc;
  if (where == APL_HOST) {
    const size_t bytes = sizeof(double) * state_len(h);
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q_in, bytes, hipMemcpyHostToDevice, h->stream));
    if ((rc = enqueue_limit(h, h->d_qa, h->d_qa, true))) return rc;
    HIPCHK(h, hipMemcpyAsync(q_out, h->d_qa, bytes, hipMemcpyDeviceToHost, h->stream));
  } else {
    if ((rc = enqueue_limit(h, q_in, q_out, true))) return rc;
  }
  return check_flags(h);
}

int apl_destroy(apl_handle *h) {
  if (!h) return APL_ERR_ARG;
  (void)hipStreamSynchronize(h->stream);
  free_handle(h);
  return APL_OK;
}

int apl_get_info(const apl_handle *h, apl_info *o) {
  if (!h || !o) return APL_ERR_ARG;
  o->K = h->K;
  o->N = h->ops.N;
  o->Np = h->ops.Np;
  o->Nc = h->ops.Nc;
  o->Ng = h->ops.Ng;
  o->Nfp = h->ops.Nfp;
  o->m = h->M;
  o->dtscale = h->geo.dtscale;
  o->rmin = h->ops.rmin;
  o->device_bytes = h->dev_bytes;
  o->device = h->device;
  return APL_OK;
}

int apl_set_stream(apl_handle *h, void *s) {
  if (!h) return APL_ERR_ARG;
  (void)hipStreamSynchronize(h->stream);
  free_graphs(h);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  h->own_stream = false;
  h->stream = (hipStream_t)s;
  return APL_OK;
}

int apl_get_index_maps(const apl_handle *h, int *etetf, int *fmask, int *vmapP) {
  if (!h) return APL_ERR_ARG;
  if (etetf) std::memcpy(etetf, h->etetf.data(), sizeof(int) * h->etetf.size());
  if (fmask) std::memcpy(fmask, h->ops.fmask.data(), sizeof(int) * h->ops.fmask.size());
  if (vmapP) std::memcpy(vmapP, h->vmapP.data(), sizeof(int) * h->vmapP.size());
  return APL_OK;
}

int apl_get_geometry(const apl_handle *h, double *geom, double *x, double *y) {
  if (!h) return APL_ERR_ARG;
  if (geom) std::memcpy(geom, h->geo.geom.data(), sizeof(double) * h->geo.geom.size());
  if (x) std::memcpy(x, h->geo.x.data(), sizeof(double) * h->geo.x.size());
  if (y) std::memcpy(y, h->geo.y.data(), sizeof(double) * h->geo.y.size());
  return APL_OK;
}

static int copy_operator(const Operators &o, int which, double *out, int capacity) {
  const std::vector<double> *v = nullptr;
  switch (which) {
    case 0: v = &o.Vc; break;
    case 1: v = &o.Drc; break;
    case 2: v = &o.Dsc; break;
    case 3: v = &o.If; break;
    case 4: v = &o.VVT; break;
    case 5: v = &o.cw; break;
    case 6: v = &o.gz; break;
    case 7: v = &o.gw; break;
    case 8: v = &o.r; break;
    case 9: v = &o.s; break;
    default: return APL_ERR_ARG;
  }
  if ((size_t)capacity < v->size()) return APL_ERR_ARG;
  std::memcpy(out, v->data(), sizeof(double) * v->size());
  return APL_OK;
}

int apl_get_operator(const apl_handle *h, int which, double *out, int capacity) {
  if (!h || !out) return APL_ERR_ARG;
  return copy_operator(h->ops, which, out, capacity);
}

int apl_host_operator(int order, int which, double *out, int capacity) {
  if (!out) return APL_ERR_ARG;
  Operators o;
  if (!build_operators(order, o)) return APL_ERR_ORDER;
  return copy_operator(o, which, out, capacity);
}

int apl_host_connectivity(const apl_mesh *mesh, int *etetf) {
  if (!mesh || !etetf || mesh->K < 1 || !mesh->EToV) return APL_ERR_ARG;
  if (bad_mesh_ids(mesh)) return APL_ERR_ARG;
  std::vector<int> et;
  const int rc = face_pair(mesh->K, mesh->Nv, mesh->EToV, mesh->vert_id, mesh->face_tag, et);
  std::memcpy(etetf, et.data(), sizeof(int) * et.size());
  return rc;
}

int apl_rhs(apl_handle *h, double t, const double *q, double *rhs, double cfl,
            double *suggested_dt, int where) {
  if (!h || !q || !rhs) return APL_ERR_ARG;
  if (where == APL_DEVICE && states_overlap(h, q, rhs))
    return fail(h, APL_ERR_ARG, "rhs: the output overlaps the input state on the device");
  const size_t bytes = sizeof(double) * state_len(h);
  int rc;
  const double *dq = q;
  double *drhs = rhs;
  if (where == APL_HOST) {
    if ((rc = ensure_work(h))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q, bytes, hipMemcpyHostToDevice, h->stream));
    dq = h->d_qa;
    drhs = h->d_out;
  }
  HIPCHK(h, hipMemsetAsync(h->d_err, 0, sizeof(unsigned), h->stream));
  RhsArgs a = base_args(h);
  a.mode = 0;
  a.q = dq;
  a.out = drhs;
  a.t = t;
  a.speed_part = h->d_speed;
  if ((rc = dispatch_rhs(h, a))) return rc;
  if (where == APL_HOST)
    HIPCHK(h, hipMemcpyAsync(rhs, drhs, bytes, hipMemcpyDeviceToHost, h->stream));
  if ((rc = check_flags(h))) return rc;
  if (suggested_dt) {
    const int nb = (h->K + 15) / 16;
    std::vector<double> sp(nb);
    HIPCHK(h, hipMemcpy(sp.data(), h->d_speed, sizeof(double) * nb, hipMemcpyDeviceToHost));
    double ms = 0.0;
    for (double s : sp) ms = s > ms ? s : ms;
    // wxnodaldg2dmethod.cc:523
    *suggested_dt = 2. / 3. * cfl * h->geo.dtscale * (h->ops.rmin / ms);
  }
  return APL_OK;
}

int apl_rhs_noalloc(apl_handle *h, double t, const double *q, double *rhs, double cfl,
                    double *suggested_dt) {
  if (!h || !q || !rhs) return APL_ERR_ARG;
  if (!h->scratch)
    return fail(h, APL_ERR_HIP, "rhs_noalloc: no device scratch array (" + h->scratch_error + ")");
  const long nb = ((long)h->K + 15) / 16;
  if (nb > NOALLOC_MAX_BATCHES)
    return fail(h, APL_ERR_ARG, "rhs_noalloc: K = " + std::to_string(h->K) +
                                    " exceeds the scratch capacity of " +
                                    std::to_string(16 * NOALLOC_MAX_BATCHES) + " elements");
  if (states_overlap(h, q, rhs))
    return fail(h, APL_ERR_ARG, "rhs_noalloc: the output overlaps the input state");
  // the flags and the reduced maximum start at zero (the array keeps the
  // last call's contents)
  HIPCHK(h, hipMemsetAsync(h->scratch, 0, sizeof(double) * NOALLOC_HEAD, h->stream));
  RhsArgs a = base_args(h);
  a.mode = 0;
  a.q = q;
  a.out = rhs;
  a.t = t;
  a.err = (unsigned *)(h->scratch + 1);
  a.speed_part = h->scratch + NOALLOC_HEAD;
  int rc;
  if ((rc = dispatch_rhs(h, a))) return rc;
  if (suggested_dt && launch_max_reduce(h->stream, h->scratch + NOALLOC_HEAD, (int)nb, h->scratch))
    return fail(h, APL_ERR_HIP, "rhs_noalloc: max kernel");
  double back[2];
  HIPCHK(h, hipMemcpyAsync(back, h->scratch, sizeof(back), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  unsigned ef;
  std::memcpy(&ef, &back[1], sizeof(ef));
  if ((rc = flags_status(h, ef))) return rc;
  // wxnodaldg2dmethod.cc:523
  if (suggested_dt) *suggested_dt = 2. / 3. * cfl * h->geo.dtscale * (h->ops.rmin / back[0]);
  return APL_OK;
}

int apl_initial_dt(apl_handle *h, const double *q, double cfl, double *dt, int where) {
  if (!h || !q || !dt) return APL_ERR_ARG;
  const size_t n = state_len(h);
  std::vector<double> hq;
  const double *qq = q;
  if (where == APL_DEVICE) {
    hq.resize(n);
    HIPCHK(h, hipMemcpy(hq.data(), q, sizeof(double) * n, hipMemcpyDeviceToHost));
    qq = hq.data();
  }
  // Roe speeds of the node state against itself (wxnodaldg2dmethod.cc:201-213,
  // wxeulereqn.cc:95-234); phMaxwell's fastest wave is c0 max(chi, gamma, 1).
  const apl_params &p = h->params;
  const int M = h->M;
  double maxSpeed = 0.0;
  const bool euler = (p.eqn == APL_EQN_EULER5 || p.eqn == APL_EQN_EULER4);
  for (size_t i = 0; i < n / M; ++i) {
    const double *d = qq + i * M;
    double lam = 0.0;
    if (euler) {
      const double rho = d[0], mu = d[1], mv = d[2];
      const double mw = (p.eqn == APL_EQN_EULER5) ? d[3] : 0.0;
      const double E = (p.eqn == APL_EQN_EULER5) ? d[4] : d[3];
      const double g1 = p.gas_gamma - 1.0;
      const double pr = g1 * (E - 0.5 * (mu * mu + mv * mv + mw * mw) / rho);
      const double u = mu / rho, v = mv / rho, w = mw / rho;
      const double enth = (E + pr) / rho;
      const double a = std::sqrt(g1 * (enth - 0.5 * (u * u + v * v + w * w)));
      for (double s : {u - a, u, u + a, v - a, v, v + a}) lam = std::max(lam, s * s);
      lam = std::sqrt(lam);
    } else {
      lam = std::max(std::max(p.chi * p.c0, p.gamma_ph * p.c0), p.c0);
    }
    maxSpeed = std::max(maxSpeed, lam);
  }
  *dt = 2. / 3. * cfl * h->geo.dtscale * (h->ops.rmin / maxSpeed);
  return APL_OK;
}

// LSERK4 coefficients (Carpenter & Kennedy 1994, five-stage fourth order).
static const double rk4a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                               -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0,
                               -1275806237668.0 / 842570457699.0};
static const double rk4b[5] = {1432997174477.0 / 9575080441755.0,
                               5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0,
                               3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double rk4c[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                               2526269341429.0 / 6820363962896.0,
                               2006345519317.0 / 3224310063776.0,
                               2802321613138.0 / 2924317926251.0};

// Enqueue the five stages of one step on h->stream; the state starts in
// `src` and ends in `dst` (the other buffer).
static int enqueue_step(apl_handle *h, double *src, double *dst, double dt, bool advance) {
  RhsArgs a = base_args(h);
  a.mode = 1;
  a.res = h->d_res;
  a.dt = dt;
  a.tptr = h->d_t;
  a.speed_part = nullptr;
  double *bufs[2] = {src, dst};
  for (int s = 0; s < 5; ++s) {
    a.q = bufs[s & 1];
    a.out = bufs[(s + 1) & 1];
    a.rk_a = rk4a[s];
    a.rk_b = rk4b[s];
    a.tofs = rk4c[s] * dt;
    int rc;
    if (h->params.limiter && (rc = enqueue_limit(h, bufs[s & 1], bufs[s & 1]))) return rc;
    rc = dispatch_rhs(h, a);
    if (rc) return rc;
  }
  if (advance) {
    if (launch_advance_time(h->stream, h->d_t, dt)) return fail(h, APL_ERR_HIP, "time kernel");
  }
  return APL_OK;
}

// Whether a kernel reads the time: only the isentropic-vortex boundary
// state depends on it.  Otherwise the captured step skips the clock kernel.
static bool time_dependent(const apl_handle *h) {
  for (int i = 0; i < h->params.n_bc; ++i)
    if (h->params.bc[i].kind == APL_BC_ISENTROPIC_VORTEX) return true;
  return false;
}

static int build_graphs(apl_handle *h, double *qa, double *qb, double dt) {
  free_graphs(h);
  const bool adv = time_dependent(h);
  for (int i = 0; i < 2; ++i) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = (i == 0) ? enqueue_step(h, qa, qb, dt, adv) : enqueue_step(h, qb, qa, dt, adv);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(h, APL_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    h->step_graph[i] = g;
    HIPCHK(h, hipGraphInstantiate(&h->step_exec[i], g, nullptr, nullptr, 0));
  }
  h->graph_qa = qa;
  h->graph_dt = dt;
  return APL_OK;
}

// Runs nsteps LSERK4 steps on device buffers qa (in/out) and qb (scratch).
static int run_lserk4(apl_handle *h, double t0, double dt, int nsteps, double *qa, double *qb) {
  int rc;
  if (h->graph_qa != qa || h->graph_dt != dt || !h->step_exec[0])
    if ((rc = build_graphs(h, qa, qb, dt))) return rc;
  HIPCHK(h, hipMemcpyAsync(h->d_t, &t0, sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(h->d_err, 0, sizeof(unsigned), h->stream));
  for (int s = 0; s < nsteps; ++s) HIPCHK(h, hipGraphLaunch(h->step_exec[s & 1], h->stream));
  if (nsteps & 1)
    HIPCHK(h, hipMemcpyAsync(qa, qb, sizeof(double) * state_len(h), hipMemcpyDeviceToDevice, h->stream));
  return check_flags(h);
}

int apl_lserk4(apl_handle *h, double t0, double dt, int nsteps, double *q, int where) {
  if (!h || !q || nsteps < 0) return APL_ERR_ARG;
  if (nsteps == 0) return APL_OK;
  int rc;
  if ((rc = ensure_work(h))) return rc;
  const size_t bytes = sizeof(double) * state_len(h);
  if (where == APL_HOST) {
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q, bytes, hipMemcpyHostToDevice, h->stream));
    if ((rc = run_lserk4(h, t0, dt, nsteps, h->d_qa, h->d_qb))) return rc;
    HIPCHK(h, hipMemcpy(q, h->d_qa, bytes, hipMemcpyDeviceToHost));
    return APL_OK;
  }
  return run_lserk4(h, t0, dt, nsteps, q, h->d_qb);
}

// ---------------------------------------------------------------------------
// Apollo's stepper within one output frame: PETSc TSSSP (type rks2, s stages)
// under ApSolver's dt control
// ---------------------------------------------------------------------------

// One frame as ApSolver::solve runs it (src/solvers/apsolver.cc:172-216):
// TSSetInitialTimeStep(t0, dt0), exact final time MATCHSTEP, and per step
//  - MonitorVTK (apsolver.cc:421-440): dt = |t_end - t| if that is smaller;
//  - MATCHSTEP: dt = t_end - t if t + dt would pass t_end;
//  - TSSSP rks2 with s stages, the low-storage second-order SSP scheme
//    (Ketcheson 2008): w = u; for i < s-1: w += dt/(s-1) L(t + dt i/(s-1), w);
//    u = (s-1)/s w + dt/s L(t + dt, w) + u/s;
//  - ComputeRHSforTS (apsolver.cc:442-480): after every RHS the next dt is
//    min(suggested, dt of this step), so the last stage's value stands.
// The stages run as fused kernels (epilogue mode 2); the last stage's speed
// maximum is reduced on the device and read back once per step.
int apl_ssp_advance(apl_handle *h, double t0, double t_end, double dt0, double cfl, int stages,
                    double *q, int where, int max_steps, int *nsteps, double *dts) {
  if (!h || !q || !(dt0 > 0.0) || stages < 2 || max_steps < 0 || !(t_end >= t0))
    return APL_ERR_ARG;
  int rc;
  if ((rc = ensure_work(h))) return rc;
  const size_t bytes = sizeof(double) * state_len(h);
  double *sol = q;
  if (where == APL_HOST) {
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q, bytes, hipMemcpyHostToDevice, h->stream));
    sol = h->d_qa;
  }
  double *bufs[2] = {h->d_qb, h->d_res};
  const int nb = (h->K + 15) / 16;
  double *d_smax = h->d_t + 1;
  HIPCHK(h, hipMemsetAsync(h->d_err, 0, sizeof(unsigned), h->stream));
  RhsArgs a = base_args(h);
  a.mode = 2;
  double t = t0, dt_temp = dt0;
  int n = 0;
  for (;;) {
    if (std::fabs(t_end - t) < dt_temp) dt_temp = std::fabs(t_end - t);   // MonitorVTK
    if (t >= t_end) break;
    double dt = dt_temp;
    if (t + dt > t_end) dt = t_end - t;                                      // MATCHSTEP
    if (n >= max_steps) return fail(h, APL_ERR_ARG, "apl_ssp_advance: more than max_steps steps");
    // TSSSP works on a copy w of sol; with the limiter, step() limits w in
    // place before each RHS (wxnodaldg2dmethod.cc:243-247), so the first
    // stage limits sol into the free buffer and later stages in place
    const double *src = sol;
    for (int i = 0; i < stages - 1; ++i) {
      if (h->params.limiter) {
        double *lim = (i == 0) ? bufs[1] : bufs[(i - 1) & 1];
        if ((rc = enqueue_limit(h, src, lim))) return rc;
        src = lim;
      }
      a.q = src;
      a.out = bufs[i & 1];
      a.z = nullptr;
      a.ca = 1.0;
      a.cb = dt / (stages - 1.);
      a.cz = 0.0;
      a.t = t + dt * (i / (stages - 1.));
      a.speed_part = nullptr;
      if ((rc = dispatch_rhs(h, a))) return rc;
      src = bufs[i & 1];
    }
    if (h->params.limiter) {
      // with two stages the first limited copy sits in bufs[1] and the
      // stage result in bufs[0]; src is always a work buffer here
      if ((rc = enqueue_limit(h, src, (double *)src))) return rc;
    }
    a.q = src;
    a.out = sol;
    a.z = sol;
    a.ca = (stages - 1.) / stages;
    a.cb = dt / stages;
    a.cz = 1. / stages;
    a.t = t + dt;
    a.speed_part = h->d_speed;
    if ((rc = dispatch_rhs(h, a))) return rc;
    if (launch_max_reduce(h->stream, h->d_speed, nb, d_smax)) return fail(h, APL_ERR_HIP, "max kernel");
    // the speed maximum and the flags word after it, in one copy
    HIPCHK(h, hipMemcpyAsync(h->h_pin, d_smax, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double ms = h->h_pin[0];
    if (*(const unsigned *)(h->h_pin + 1)) return check_flags(h);
    // wxnodaldg2dmethod.cc:523, then apsolver.cc:474
    const double suggested = 2. / 3. * cfl * h->geo.dtscale * (h->ops.rmin / ms);
    dt_temp = std::fmin(suggested, dt);
    t = t + dt;
    if (dts) dts[n] = dt;
    ++n;
  }
  if (nsteps) *nsteps = n;
  if (where == APL_HOST) HIPCHK(h, hipMemcpy(q, sol, bytes, hipMemcpyDeviceToHost));
  return check_flags(h);
}

// ---------------------------------------------------------------------------
// discrete adjoint
// ---------------------------------------------------------------------------

// FP64 fragments of the transposed operators (see apl_adjoint.h).
static int ensure_adjoint_ops(apl_handle *h) {
  if (h->d_adjops[0]) return APL_OK;
  const Operators &o = h->ops;
  const int Np = o.Np, Nc = o.Nc, Ng = o.Ng;
  const int NKS1 = (Np + 3) / 4, NCH = (Nc + 15) / 16, NRT = (Np + 15) / 16, NKF = (3 * Ng + 3) / 4;
  const int NFT = (3 * Ng + 15) / 16;
  std::vector<double> Ar((size_t)Np * Nc), As((size_t)Np * Nc), L((size_t)Np * 3 * Ng);
  for (int n = 0; n < Np; ++n) {
    for (int c = 0; c < Nc; ++c) {
      double sr = 0.0, ss = 0.0;
      for (int j = 0; j < Np; ++j) {
        sr += o.VVT[(size_t)n * Np + j] * o.Drc[(size_t)c * Np + j];
        ss += o.VVT[(size_t)n * Np + j] * o.Dsc[(size_t)c * Np + j];
      }
      Ar[(size_t)n * Nc + c] = sr * o.cw[c];
      As[(size_t)n * Nc + c] = ss * o.cw[c];
    }
    for (int k = 0; k < 3 * Ng; ++k) {
      double sl = 0.0;
      for (int j = 0; j < Np; ++j) sl += o.VVT[(size_t)n * Np + j] * o.If[(size_t)k * Np + j];
      L[(size_t)n * 3 * Ng + k] = sl * o.gw[k % Ng];
    }
  }
  std::vector<double> vc((size_t)NCH * NKS1 * 64, 0.0), ar(vc.size(), 0.0), as(vc.size(), 0.0);
  for (int cc = 0; cc < NCH; ++cc)
    for (int ks = 0; ks < NKS1; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int c = 16 * cc + (l & 15), j = 4 * ks + (l >> 4);
        if (c < Nc && j < Np) {
          const size_t i = ((size_t)cc * NKS1 + ks) * 64 + l;
          vc[i] = o.Vc[(size_t)c * Np + j];
          ar[i] = Ar[(size_t)j * Nc + c];
          as[i] = As[(size_t)j * Nc + c];
        }
      }
  std::vector<double> vct((size_t)NRT * NCH * 64 * 4, 0.0);
  for (int rt = 0; rt < NRT; ++rt)
    for (int cc = 0; cc < NCH; ++cc)
      for (int l = 0; l < 64; ++l)
        for (int r = 0; r < 4; ++r) {
          const int n = 16 * rt + (l & 15), c = 16 * cc + acc_row<double>(l >> 4, r);
          if (n < Np && c < Nc) vct[(((size_t)rt * NCH + cc) * 64 + l) * 4 + r] = o.Vc[(size_t)c * Np + n];
        }
  std::vector<double> lt((size_t)NFT * NKS1 * 64, 0.0);
  for (int ft = 0; ft < NFT; ++ft)
    for (int ks = 0; ks < NKS1; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int kf = 16 * ft + (l & 15), n = 4 * ks + (l >> 4);
        if (kf < 3 * Ng && n < Np) lt[((size_t)ft * NKS1 + ks) * 64 + l] = -L[(size_t)n * 3 * Ng + kf];
      }
  std::vector<double> ift((size_t)NRT * NKF * 64, 0.0);
  for (int rt = 0; rt < NRT; ++rt)
    for (int ks = 0; ks < NKF; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int n = 16 * rt + (l & 15), kf = 4 * ks + (l >> 4);
        if (n < Np && kf < 3 * Ng) ift[((size_t)rt * NKF + ks) * 64 + l] = o.If[(size_t)kf * Np + n];
      }
  const std::vector<double> *src[6] = {&vc, &ar, &as, &vct, &lt, &ift};
  int rc;
  for (int i = 0; i < 6; ++i)
    if ((rc = dev_upload<double>(h, &h->d_adjops[i], src[i]->data(), src[i]->size()))) return rc;
  std::vector<double> cub((size_t)3 * Nc);
  for (int c = 0; c < Nc; ++c) {
    cub[3 * c] = o.cr[c];
    cub[3 * c + 1] = o.cs[c];
    cub[3 * c + 2] = o.cw[c];
  }
  if ((rc = dev_upload<double>(h, &h->d_Vc, o.Vc.data(), o.Vc.size()))) return rc;
  if ((rc = dev_upload<double>(h, &h->d_cub, cub.data(), cub.size()))) return rc;
  const size_t n = state_len(h);
  if ((rc = dev_alloc(h, (void **)&h->d_sbar, sizeof(double) * (size_t)h->K * 3 * Ng * h->M))) return rc;
  if ((rc = dev_alloc(h, (void **)&h->d_sbar2, sizeof(double) * (size_t)h->K * 3 * Ng * h->M))) return rc;
  if ((rc = dev_alloc(h, (void **)&h->d_qbar, sizeof(double) * n))) return rc;
  if ((rc = dev_alloc(h, (void **)&h->d_rbar, sizeof(double) * n))) return rc;
  if ((rc = dev_alloc(h, (void **)&h->d_stg, sizeof(double) * n * 2))) return rc;
  return APL_OK;
}

static AdjArgs adj_args(apl_handle *h) {
  AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  RhsArgs r = base_args(h);
  a.K = h->K;
  a.sbar = h->d_sbar;
  a.geom = h->d_geom;
  a.vert = h->d_vert;
  a.fcode = h->d_fcode;
  a.I1D = h->d_I1D;
  a.fmask = h->d_fmask;
  a.Avc = h->d_adjops[0];
  a.Aart = h->d_adjops[1];
  a.Aast = h->d_adjops[2];
  a.Avct = h->d_adjops[3];
  a.Alt = h->d_adjops[4];
  a.Aift = h->d_adjops[5];
  a.err = h->d_err;
  a.ph = r.ph;
  a.bc = r.bc;
  return a;
}

static int dispatch_adjoint(apl_handle *h, const AdjArgs &a) {
  const int nb = (h->K + 15) / 16;
  const apl_params &p = h->params;
  if (p.flux == APL_FLUX_HLL || p.flux == APL_FLUX_ROE || p.flux == APL_FLUX_WAVE)
    return fail(h, APL_ERR_UNSUPPORTED, "the adjoint covers the Lax-Friedrichs and upwind fluxes");
  int rc;
  switch (p.order) {
    case 1: rc = launch_adjoint<1>(p.eqn, p.flux, h->stream, a, nb); break;
    case 2: rc = launch_adjoint<2>(p.eqn, p.flux, h->stream, a, nb); break;
    case 3: rc = launch_adjoint<3>(p.eqn, p.flux, h->stream, a, nb); break;
    case 4: rc = launch_adjoint<4>(p.eqn, p.flux, h->stream, a, nb); break;
    case 5: rc = launch_adjoint<5>(p.eqn, p.flux, h->stream, a, nb); break;
    case 6: rc = launch_adjoint<6>(p.eqn, p.flux, h->stream, a, nb); break;
    case 7: rc = launch_adjoint<7>(p.eqn, p.flux, h->stream, a, nb); break;
    case 8: rc = launch_adjoint<8>(p.eqn, p.flux, h->stream, a, nb); break;
    default: rc = APL_ERR_ORDER;
  }
  if (rc) return fail(h, rc, "adjoint kernel launch failed");
  return APL_OK;
}

int apl_adjoint_rhs(apl_handle *h, double t, const double *q, const double *lam, double *out,
                    int where) {
  if (!h || !q || !lam || !out) return APL_ERR_ARG;
  if (h->params.limiter) return fail(h, APL_ERR_UNSUPPORTED, "the adjoint does not cover the limiter");
  if (where == APL_DEVICE && (states_overlap(h, q, out) || states_overlap(h, lam, out)))
    return fail(h, APL_ERR_ARG, "adjoint_rhs: the output overlaps an input on the device");
  int rc;
  if ((rc = ensure_adjoint_ops(h)) || (rc = ensure_work(h))) return rc;
  const size_t bytes = sizeof(double) * state_len(h);
  const double *dq = q, *dl = lam;
  double *dout = out;
  if (where == APL_HOST) {
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_qb, lam, bytes, hipMemcpyHostToDevice, h->stream));
    dq = h->d_qa;
    dl = h->d_qb;
    dout = h->d_out;
  }
  HIPCHK(h, hipMemsetAsync(h->d_err, 0, sizeof(unsigned), h->stream));
  AdjArgs a = adj_args(h);
  a.mode = 0;
  a.face_pass = 1;
  a.q = dq;
  a.lam = dl;
  a.out = dout;
  a.t = t;
  if ((rc = dispatch_adjoint(h, a))) return rc;
  if (where == APL_HOST) HIPCHK(h, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, h->stream));
  return check_flags(h);
}

// One forward LSERK4 step from `in` to `result`, its stage inputs
// alternating between two buffers.
static int forward_step(apl_handle *h, double t, double dt, const double *in, double *result) {
  const size_t n = state_len(h);
  double *stg = h->d_stg;
  RhsArgs a = base_args(h);
  a.mode = 1;
  a.res = h->d_res;
  a.dt = dt;
  a.t = t;
  // stage 0 reads `in` (which may be `result`: stage 4 writes it last)
  for (int s = 0; s < 5; ++s) {
    a.q = s == 0 ? in : stg + (s & 1) * n;
    a.out = (s == 4) ? result : stg + ((s + 1) & 1) * n;
    a.rk_a = rk4a[s];
    a.rk_b = rk4b[s];
    a.tofs = rk4c[s] * dt;
    int rc = dispatch_rhs(h, a);
    if (rc) return rc;
  }
  return APL_OK;
}

// One LSERK4 step whose stage s reads stg[s n] and writes stage s + 1's
// input to stg[(s + 1) n] (s < 4) or the new state to `result`; stg[0] holds
// the step's input state.
static int forward_stages(apl_handle *h, double t, double dt, double *stg, double *result) {
  const size_t n = state_len(h);
  RhsArgs a = base_args(h);
  a.mode = 1;
  a.res = h->d_res;
  a.dt = dt;
  a.t = t;
  for (int s = 0; s < 5; ++s) {
    a.q = stg + (size_t)s * n;
    a.out = (s == 4) ? result : stg + (size_t)(s + 1) * n;
    a.rk_a = rk4a[s];
    a.rk_b = rk4b[s];
    a.tofs = rk4c[s] * dt;
    int rc = dispatch_rhs(h, a);
    if (rc) return rc;
  }
  return APL_OK;
}

int apl_lserk4_adjoint(apl_handle *h, double t0, double dt, int nsteps, int ckpt_every,
                       const double *q0, const double *qbarT, double *dJdq0, double *qT,
                       int where) {
  if (!h || !q0 || !qbarT || !dJdq0 || nsteps < 0 || ckpt_every < 1) return APL_ERR_ARG;
  if (h->params.precision != APL_PREC_FP64)
    return fail(h, APL_ERR_UNSUPPORTED, "the adjoint runs on FP64 handles");
  if (h->params.limiter) return fail(h, APL_ERR_UNSUPPORTED, "the adjoint does not cover the limiter");
  int rc;
  if ((rc = ensure_adjoint_ops(h)) || (rc = ensure_work(h))) return rc;
  const size_t n = state_len(h), bytes = sizeof(double) * n;
  const int C = ckpt_every;
  const long long nck = nsteps / C + 1;
  if (h->ckpt_slots < nck) {
    if (h->d_ckpt) (void)hipFree(h->d_ckpt);
    h->dev_bytes -= (long long)bytes * h->ckpt_slots;
    h->d_ckpt = nullptr;
    h->ckpt_slots = 0;
    if ((rc = dev_alloc(h, (void **)&h->d_ckpt, bytes * (size_t)nck))) return rc;
    h->ckpt_slots = nck;
  }
  // the five stage inputs of every step of one segment
  const long long nseg = 5LL * C;
  if (h->seg_slots < nseg) {
    if (h->d_seg) (void)hipFree(h->d_seg);
    h->dev_bytes -= (long long)bytes * h->seg_slots;
    h->d_seg = nullptr;
    h->seg_slots = 0;
    if ((rc = dev_alloc(h, (void **)&h->d_seg, bytes * (size_t)nseg))) return rc;
    h->seg_slots = nseg;
  }
  const hipMemcpyKind in_kind = where == APL_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  HIPCHK(h, hipMemsetAsync(h->d_err, 0, sizeof(unsigned), h->stream));
  // forward sweep with checkpoints; the steps of the last segment keep their
  // stage inputs in d_seg at once, so that segment needs no recompute
  const long long last = nsteps > 0 ? (nsteps - 1) / C : -1;
  const int sl = (int)(last * C);
  double *cur = h->d_ckpt;
  HIPCHK(h, hipMemcpyAsync(cur, q0, bytes, in_kind, h->stream));
  for (int step = 0; step < nsteps; ++step) {
    if (step >= sl) {
      double *slot = h->d_seg + (size_t)5 * (step - sl) * n;
      if (step == sl) HIPCHK(h, hipMemcpyAsync(slot, cur, bytes, hipMemcpyDeviceToDevice, h->stream));
      double *next = step + 1 < nsteps ? slot + 5 * n : h->d_qa;
      if ((rc = forward_stages(h, t0 + step * dt, dt, slot, next))) return rc;
      cur = next;
    } else {
      double *dst = ((step + 1) % C == 0) ? h->d_ckpt + (size_t)((step + 1) / C) * n : h->d_qa;
      if ((rc = forward_step(h, t0 + step * dt, dt, cur, dst))) return rc;
      cur = dst;
    }
  }
  if (qT) {
    const hipMemcpyKind out_kind = where == APL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPCHK(h, hipMemcpyAsync(qT, cur, bytes, out_kind, h->stream));
  }
  // reverse sweep
  HIPCHK(h, hipMemcpyAsync(h->d_qbar, qbarT, bytes, in_kind, h->stream));
  HIPCHK(h, hipMemsetAsync(h->d_rbar, 0, bytes, h->stream));
  AdjArgs a = adj_args(h);
  a.mode = 1;
  a.lam = h->d_rbar;
  a.qbar = h->d_qbar;
  a.dt = dt;
  // The reverse stages run in the order (step i, stage s) = (n-1, 4), ...,
  // (0, 0).  The first launch computes its face weights Sb in a face pass;
  // every launch writes the next stage's Sb into the other buffer.
  double *sb[2] = {h->d_sbar, h->d_sbar2};
  int sbi = 0;
  bool first = true;
  for (long long seg = (nsteps - 1) / C; seg >= 0 && nsteps > 0; --seg) {
    const int s0 = (int)(seg * C), s1 = std::min(nsteps, s0 + C);
    // the segment recomputed from its checkpoint (the last one was kept by
    // the forward sweep), keeping the stage inputs of step i in slots
    // 5 (i - s0) .. 5 (i - s0) + 4 (slot 5 (i - s0) is q_i), so that the
    // reverse stages need no further forward work
    if (seg != last) {
      HIPCHK(h, hipMemcpyAsync(h->d_seg, h->d_ckpt + (size_t)seg * n, bytes,
                               hipMemcpyDeviceToDevice, h->stream));
      for (int i = s0; i < s1; ++i) {
        double *slot = h->d_seg + (size_t)5 * (i - s0) * n;
        double *next = i + 1 < s1 ? slot + 5 * n : h->d_qa;
        if ((rc = forward_stages(h, t0 + i * dt, dt, slot, next))) return rc;
      }
    }
    for (int i = s1 - 1; i >= s0; --i) {
      const double t = t0 + i * dt;
      for (int s = 4; s >= 0; --s) {
        a.q = h->d_seg + (size_t)(5 * (i - s0) + s) * n;
        a.t = t + rk4c[s] * dt;
        a.rk_a = rk4a[s];
        a.rk_b = rk4b[s];
        a.rk_b_next = s > 0 ? rk4b[s - 1] : rk4b[4];
        a.face_pass = first ? 1 : 0;
        a.sbar = sb[sbi];
        a.sbar_out = (i == 0 && s == 0) ? nullptr : sb[sbi ^ 1];
        if ((rc = dispatch_adjoint(h, a))) return rc;
        sbi ^= 1;
        first = false;
      }
    }
  }
  const hipMemcpyKind out_kind = where == APL_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPCHK(h, hipMemcpyAsync(dJdq0, h->d_qbar, bytes, out_kind, h->stream));
  return check_flags(h);
}

int apl_objective_disc_sq(apl_handle *h, const double *q, int var, double xc, double yc,
                          double radius, double *J, double *dJdq, int where) {
  if (!h || !q || !J || var < 0 || var >= h->M) return APL_ERR_ARG;
  int rc;
  if ((rc = ensure_adjoint_ops(h)) || (rc = ensure_work(h))) return rc;
  const size_t bytes = sizeof(double) * state_len(h);
  const int nb = (h->K + 63) / 64;
  if (!h->d_opart && (rc = dev_alloc(h, (void **)&h->d_opart, sizeof(double) * nb))) return rc;
  const double *dq = q;
  if (where == APL_HOST) {
    HIPCHK(h, hipMemcpyAsync(h->d_qa, q, bytes, hipMemcpyHostToDevice, h->stream));
    dq = h->d_qa;
  }
  ObjArgs a;
  a.K = h->K;
  a.var = var;
  a.q = dq;
  a.grad = dJdq ? (where == APL_HOST ? h->d_out : dJdq) : nullptr;
  a.part = h->d_opart;
  a.geom = h->d_geom;
  a.vert = h->d_vert;
  a.Vc = h->d_Vc;
  a.cub = h->d_cub;
  a.xc = xc;
  a.yc = yc;
  a.R = radius;
  switch (h->params.order) {
    case 1: rc = launch_objective<1>(h->M, h->stream, a, nb); break;
    case 2: rc = launch_objective<2>(h->M, h->stream, a, nb); break;
    case 3: rc = launch_objective<3>(h->M, h->stream, a, nb); break;
    case 4: rc = launch_objective<4>(h->M, h->stream, a, nb); break;
    case 5: rc = launch_objective<5>(h->M, h->stream, a, nb); break;
    case 6: rc = launch_objective<6>(h->M, h->stream, a, nb); break;
    case 7: rc = launch_objective<7>(h->M, h->stream, a, nb); break;
    case 8: rc = launch_objective<8>(h->M, h->stream, a, nb); break;
    default: rc = APL_ERR_ORDER;
  }
  if (rc) return fail(h, rc, "objective kernel launch failed");
  std::vector<double> part(nb);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->d_opart, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  if (dJdq && where == APL_HOST)
    HIPCHK(h, hipMemcpyAsync(dJdq, h->d_out, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double sum = 0.0;
  for (double v : part) sum += v;     // fixed order: deterministic
  *J = sum;
  return APL_OK;
}

}  // extern "C"
