// wx_nodal_dg2d_hip.hpp -- Apollo-shaped C++ host interface over the C-ABI.
//
// Mirrors the plugin interface of Apollo's `nodalDG2d` subsolver so that a
// caller written against it reads the same:
//   ApSubSolver<REAL>::setup / init / step    src/subsolvers/apsubsolver.h:46
//   WxNodalDG2dMethod<REAL>::setup            src/subsolvers/wxnodaldg2dmethod.cc:50-152
//   WxNodalDG2dMethod<REAL>::init             src/subsolvers/wxnodaldg2dmethod.cc:154-233
//   WxNodalDG2dMethod<REAL>::step             src/subsolvers/wxnodaldg2dmethod.cc:235-542
//   WxStepperStatus<REAL>                     src/solvers/wxstepperstatus.h:6-66
//   ApSolver::ComputeRHSforTS                 src/solvers/apsolver.cc:442-480
// PETSc is not part of this interface: the mesh arrives as the arrays
// DMPlex would hand over (cells, cones, vertex coordinates, "Face Sets"
// labels) and vectors as plain FP64 arrays in Apollo's [cell][node][var]
// layout.  Where Apollo calls exit(1) or throws WxExcept this header throws
// apl_host::WxExcept carrying the engine's status.  Header-only; link with
// -lapollo_dg.
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/apollo_dg.h"

namespace apl_host {

class WxExcept : public std::runtime_error {
 public:
  WxExcept(int status, const std::string &what)
      : std::runtime_error(what + ": " + apl_status_string(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Return value of step(): success flag and the suggested time step.
template <typename REAL>
class WxStepperStatus {
 public:
  WxStepperStatus() : status_(true), dt_(0) {}
  WxStepperStatus(bool s, REAL dt) : status_(s), dt_(dt) {}
  bool getStatus() const { return status_; }
  REAL getSuggestedDt() const { return dt_; }
  void setStatus(bool s) { status_ = s; }
  void setSuggestedDt(REAL dt) { dt_ = dt; }

 private:
  bool status_;
  REAL dt_;
};

// The configuration keys WxNodalDG2dMethod::setup reads from its
// WxCryptSet block (wxnodaldg2dmethod.cc:58-144), with the equation and
// boundary-condition blocks flattened.
struct NodalDG2dConfig {
  int polynomialOrder = 3;
  double cfl = 0.5;
  std::string equation = "eulerEqn";     // "eulerEqn" | "phMaxwellEqn" | "euler2d" | "maxwellTM"
  // Euler "Numerical_Flux" (wxeulereqn.cc:44-84): "LF" | "HLL" | "Roe" |
  // "Wave"; any other name, or none, selects LF as Apollo does.  "upwind"
  // selects the exact TM upwind flux (maxwellTM only).
  std::string numericalFlux = "LF";
  double gas_gamma = 1.4, minDensity = 0.0, minPressure = 0.0;
  double c0 = 1.0, chi = 0.0, gamma = 0.0;
  int precision = APL_PREC_FP64;
  int algorithm = APL_ALGO_AUTO;
  // `Limiter = [<eulerLimiterHW block>]` (wxnodaldg2dmethod.cc:126-134) and
  // that block's gas_gamma, min_density, min_pressure
  // (wxhesthavenwarburtoneulerlimiter.cc:21-23)
  bool limiter = false;
  double limiterGasGamma = 1.4, limiterMinDensity = 0.0, limiterMinPressure = 0.0;
  std::vector<apl_bc> boundaryConditions;   // entry T-1 serves "Face Sets" tag T
};

// The mesh as DMPlex presents it: cells in order, counter-clockwise cones,
// vertex coordinates, optional face labels and periodic vertex identities.
struct NodalDG2dMesh {
  std::vector<double> VX, VY;
  std::vector<int> cones;      // [K*3]
  std::vector<int> faceSets;   // [K*3], 0 = unlabelled; empty = none
  std::vector<int> vertexId;   // [Nv], empty = identity
};

class WxNodalDG2dMethodHip {
 public:
  WxNodalDG2dMethodHip() = default;
  WxNodalDG2dMethodHip(const WxNodalDG2dMethodHip &) = delete;
  WxNodalDG2dMethodHip &operator=(const WxNodalDG2dMethodHip &) = delete;
  ~WxNodalDG2dMethodHip() {
    if (h_) apl_destroy(h_);
  }

  // Operators, connectivity and geometry (wxnodaldg2dmethod.cc:50-152).
  void setup(const NodalDG2dConfig &cfg, const NodalDG2dMesh &mesh) {
    cfg_ = cfg;
    apl_mesh m;
    m.K = (int)(mesh.cones.size() / 3);
    m.Nv = (int)mesh.VX.size();
    m.VX = mesh.VX.data();
    m.VY = mesh.VY.data();
    m.EToV = mesh.cones.data();
    m.vert_id = mesh.vertexId.empty() ? nullptr : mesh.vertexId.data();
    m.face_tag = mesh.faceSets.empty() ? nullptr : mesh.faceSets.data();
    apl_params p{};
    p.order = cfg.polynomialOrder;
    p.eqn = equation_kind(cfg.equation);
    p.flux = flux_kind(cfg.numericalFlux);
    p.precision = cfg.precision;
    p.algorithm = cfg.algorithm;
    p.gas_gamma = cfg.gas_gamma;
    p.min_density = cfg.minDensity;
    p.min_pressure = cfg.minPressure;
    p.c0 = cfg.c0;
    p.chi = cfg.chi;
    p.gamma_ph = cfg.gamma;
    p.limiter = cfg.limiter ? 1 : 0;
    p.lim_gas_gamma = cfg.limiterGasGamma;
    p.lim_min_density = cfg.limiterMinDensity;
    p.lim_min_pressure = cfg.limiterMinPressure;
    p.n_bc = (int)cfg.boundaryConditions.size();
    if (p.n_bc > APL_MAX_BC) throw WxExcept(APL_ERR_ARG, "too many boundary conditions");
    for (int i = 0; i < p.n_bc; ++i) p.bc[i] = cfg.boundaryConditions[i];
    if (h_) apl_destroy(h_);
    h_ = nullptr;
    const int rc = apl_create(&m, &p, &h_);
    if (rc) throw WxExcept(rc, std::string("setup: ") + apl_last_error(nullptr));
    apl_get_info(h_, &info_);
  }

  // Nodal initial condition and the suggested initial dt (Roe speeds,
  // wxnodaldg2dmethod.cc:183-228).  `ic(t, x, y, q)` fills one node.
  double init(double t, const std::function<void(double, double, double, double *)> &ic,
              double *out) {
    std::vector<double> x((size_t)info_.K * info_.Np), y(x.size());
    check(apl_get_geometry(h_, nullptr, x.data(), y.data()), "init");
    for (size_t i = 0; i < x.size(); ++i) ic(t, x[i], y[i], out + i * info_.m);
    double dt = 0.0;
    check(apl_initial_dt(h_, out, cfg_.cfl, &dt, APL_HOST), "init");
    dt_ = dt;
    return dt;
  }

  // The RHS of one stage (wxnodaldg2dmethod.cc:235-542): with a limiter,
  // `in` is limited in place first (:243-247), as Apollo does to the vector
  // it is handed; then `out` is overwritten with the RHS.  The status
  // carries (2/3) cfl dtscale rmin / maxSpeed.
  WxStepperStatus<double> step(double t, double /*dt*/, double *in, double *out,
                               int where = APL_HOST) {
    if (cfg_.limiter) applyLimiter(in, in, where);
    double sdt = 0.0;
    check(apl_rhs(h_, t, in, out, cfg_.cfl, &sdt, where), "step");
    return WxStepperStatus<double>(true, sdt);
  }

  // WxNodalDG2dMethod::applyLimiter (wxnodaldg2dmethod.cc:575-582).
  void applyLimiter(const double *qin, double *qlimited, int where = APL_HOST) {
    check(apl_limit(h_, qin, qlimited, where), "applyLimiter");
  }

  // nsteps fixed-dt LSERK4 steps in place (one hipGraph per step).
  void advance(double t0, double dt, int nsteps, double *q, int where = APL_HOST) {
    check(apl_lserk4(h_, t0, dt, nsteps, q, where), "advance");
  }

  // One output frame [t0, t_end] as ApSolver::solve runs it
  // (apsolver.cc:172-216): TSSSP rks2 with `stages` stages from dt0, the
  // MonitorVTK clamp and ComputeRHSforTS's dt = min(suggested, dt).
  // Returns the number of steps taken.
  int advanceFrame(double t0, double t_end, double dt0, double *q, int stages = 5,
                   int where = APL_HOST) {
    int n = 0;
    check(apl_ssp_advance(h_, t0, t_end, dt0, cfg_.cfl, stages, q, where, 100000000, &n, nullptr),
          "advanceFrame");
    return n;
  }

  // ApSolver::solve (apsolver.cc:172-216): nout frames of length t_end/nout,
  // each restarted from dt, with `frame(i, t, q)` called after each frame.
  void solve(double t_end, int nout, double dt, double *q,
             const std::function<void(int, double, const double *)> &frame) {
    const double tsize = t_end / nout;
    for (int i = 0; i < nout; ++i) {
      advanceFrame(i * tsize, (i + 1) * tsize, dt, q);
      if (frame) frame(i, (i + 1) * tsize, q);
    }
  }

  unsigned totalEqns() const { return (unsigned)info_.m; }
  unsigned NpElem() const { return (unsigned)info_.Np; }
  unsigned numElements() const { return (unsigned)info_.K; }
  double getDt() const { return dt_; }
  apl_handle *handle() { return h_; }

  // ApSolver::ComputeRHSforTS (apsolver.cc:442-480): zero `out`, add every
  // per-step subsolver's RHS, keep the smallest suggested dt.
  static WxStepperStatus<double> computeRHSforTS(
      const std::vector<WxNodalDG2dMethodHip *> &subsolvers, double t, double dt,
      double *in, double *out, size_t n) {
    std::vector<double> x(n);
    for (size_t i = 0; i < n; ++i) out[i] = 0.0;
    double dt_temp = dt;
    for (WxNodalDG2dMethodHip *s : subsolvers) {
      WxStepperStatus<double> st = s->step(t, dt_temp, in, x.data());
      for (size_t i = 0; i < n; ++i) out[i] += x[i];
      dt_temp = st.getSuggestedDt() < dt_temp ? st.getSuggestedDt() : dt_temp;
    }
    return WxStepperStatus<double>(true, dt_temp);
  }

 private:
  static int flux_kind(const std::string &f) {
    if (f == "HLL") return APL_FLUX_HLL;
    if (f == "Roe") return APL_FLUX_ROE;
    if (f == "Wave") return APL_FLUX_WAVE;
    if (f == "upwind") return APL_FLUX_UPWIND;
    return APL_FLUX_LF;
  }
  static int equation_kind(const std::string &e) {
    if (e == "eulerEqn") return APL_EQN_EULER5;
    if (e == "phMaxwellEqn") return APL_EQN_PHMAXWELL8;
    if (e == "euler2d") return APL_EQN_EULER4;
    if (e == "maxwellTM") return APL_EQN_MAXWELL_TM3;
    throw WxExcept(APL_ERR_ARG, "unknown equation " + e);
  }
  void check(int rc, const char *what) {
    if (rc) throw WxExcept(rc, std::string(what) + ": " + apl_last_error(h_));
  }

  NodalDG2dConfig cfg_;
  apl_handle *h_ = nullptr;
  apl_info info_{};
  double dt_ = 0.0;
};

}  // namespace apl_host
