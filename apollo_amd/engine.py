"""ctypes binding of ``libapollo_dg.so`` (include/apollo_dg.h).

This is the Python face of the engine used by the test suite and the
benchmark.  It mirrors the C-ABI one to one; device pointers come from torch
tensors (``tensor.data_ptr()``) when the state lives on the GPU.

The library must have been built (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C apollo_amd/csrc``); importing this module on a
machine without it raises immediately -- there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the engine library: the in-tree build, or another build of the same
# sources named by APOLLO_DG_LIB (for side-by-side timing of two builds)
LIB_PATH = os.environ.get("APOLLO_DG_LIB") or os.path.join(_HERE, "libapollo_dg.so")

# enums (include/apollo_dg.h)
APL_OK = 0
EQN_EULER5, EQN_EULER4, EQN_PHMAXWELL8, EQN_MAXWELL_TM3 = 0, 1, 2, 3
FLUX_LF, FLUX_UPWIND, FLUX_HLL, FLUX_ROE, FLUX_WAVE = 0, 1, 2, 3, 4
PREC_FP64, PREC_FP32_MFMA, PREC_BF16X3 = 0, 1, 2
ALGO_AUTO, ALGO_CUBATURE, ALGO_CUBATURE_MFMA, ALGO_CUBATURE_VALU, ALGO_NODAL_MFMA, ALGO_NODAL_VALU = (
    0, 1, 2, 3, 4, 5)
BC_NONE, BC_ISENTROPIC_VORTEX, BC_MAXWELL_TM, BC_PEC, BC_COPY = 0, 1, 2, 3, 4
BC_EULER_WALL, BC_EULER_INFLOW, BC_EULER_DIRICHLET = 5, 6, 7
HOST, DEVICE = 0, 1
MAX_BC = 8
NVARS = {EQN_EULER5: 5, EQN_EULER4: 4, EQN_PHMAXWELL8: 8, EQN_MAXWELL_TM3: 3}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ApolloError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d): %s" % (status_string(status), status, msg))
        self.status = status


class _BC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 8)]


class _Mesh(ctypes.Structure):
    _fields_ = [("K", ctypes.c_int), ("Nv", ctypes.c_int), ("VX", _dp), ("VY", _dp),
                ("EToV", _ip), ("vert_id", _ip), ("face_tag", _ip)]


class _Params(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("eqn", ctypes.c_int), ("flux", ctypes.c_int),
                ("precision", ctypes.c_int), ("algorithm", ctypes.c_int),
                ("gas_gamma", ctypes.c_double),
                ("min_density", ctypes.c_double), ("min_pressure", ctypes.c_double),
                ("c0", ctypes.c_double), ("chi", ctypes.c_double), ("gamma_ph", ctypes.c_double),
                ("n_bc", ctypes.c_int), ("bc", _BC * MAX_BC),
                ("limiter", ctypes.c_int), ("lim_gas_gamma", ctypes.c_double),
                ("lim_min_density", ctypes.c_double), ("lim_min_pressure", ctypes.c_double)]


class _Info(ctypes.Structure):
    _fields_ = [("K", ctypes.c_int), ("N", ctypes.c_int), ("Np", ctypes.c_int),
                ("Nc", ctypes.c_int), ("Ng", ctypes.c_int), ("Nfp", ctypes.c_int),
                ("m", ctypes.c_int), ("dtscale", ctypes.c_double), ("rmin", ctypes.c_double),
                ("device_bytes", ctypes.c_longlong), ("device", ctypes.c_int)]


EXPORTS = [
    "apl_create", "apl_destroy", "apl_get_info", "apl_set_stream", "apl_get_index_maps",
    "apl_get_geometry", "apl_get_operator", "apl_rhs", "apl_initial_dt", "apl_lserk4",
    "apl_ssp_advance", "apl_limit", "apl_adjoint_rhs", "apl_lserk4_adjoint", "apl_objective_disc_sq",
    "apl_mesh_structured", "apl_rhs_noalloc",
    "apl_status_string", "apl_last_error", "apl_host_connectivity", "apl_host_operator",
]

_lib = None


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libapollo_dg.so not built: run __graft_entry__.build() "
                           "or `make -C apollo_amd/csrc`")
    # One HIP runtime per process: torch's HIP libraries load their bundled
    # libamdhip64 by file name, so load torch first (when present) and let
    # the engine's libamdhip64.so.7 dependency resolve to that same runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.apl_create.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), ctypes.POINTER(vp)]
    L.apl_destroy.argtypes = [vp]
    L.apl_get_info.argtypes = [vp, ctypes.POINTER(_Info)]
    L.apl_set_stream.argtypes = [vp, vp]
    L.apl_get_index_maps.argtypes = [vp, _ip, _ip, _ip]
    L.apl_get_geometry.argtypes = [vp, _dp, _dp, _dp]
    L.apl_get_operator.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
    L.apl_rhs.argtypes = [vp, ctypes.c_double, vp, vp, ctypes.c_double, _dp, ctypes.c_int]
    L.apl_rhs_noalloc.argtypes = [vp, ctypes.c_double, vp, vp, ctypes.c_double, _dp]
    L.apl_initial_dt.argtypes = [vp, vp, ctypes.c_double, _dp, ctypes.c_int]
    L.apl_lserk4.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, ctypes.c_int]
    L.apl_ssp_advance.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), _dp]
    L.apl_limit.argtypes = [vp, vp, vp, ctypes.c_int]
    L.apl_adjoint_rhs.argtypes = [vp, ctypes.c_double, vp, vp, vp, ctypes.c_int]
    L.apl_lserk4_adjoint.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, vp, vp, vp, vp, ctypes.c_int]
    L.apl_objective_disc_sq.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _dp, vp, ctypes.c_int]
    L.apl_mesh_structured.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_ulonglong, ctypes.c_int, _dp, _dp, _ip, _ip, _ip]
    L.apl_host_connectivity.argtypes = [ctypes.POINTER(_Mesh), _ip]
    L.apl_host_operator.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
    L.apl_status_string.restype = ctypes.c_char_p
    L.apl_status_string.argtypes = [ctypes.c_int]
    L.apl_last_error.restype = ctypes.c_char_p
    L.apl_last_error.argtypes = [vp]
    _lib = L
    return L


def status_string(s):
    return lib().apl_status_string(s).decode()


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------

class MeshData:
    """Triangle mesh in the engine's input form (apl_mesh)."""

    def __init__(self, VX, VY, EToV, vert_id=None, face_tag=None):
        self.VX = np.ascontiguousarray(VX, dtype=np.float64)
        self.VY = np.ascontiguousarray(VY, dtype=np.float64)
        self.EToV = np.ascontiguousarray(EToV, dtype=np.int32).reshape(-1, 3)
        self.vert_id = None if vert_id is None else np.ascontiguousarray(vert_id, dtype=np.int32)
        self.face_tag = None if face_tag is None else np.ascontiguousarray(
            face_tag, dtype=np.int32).reshape(-1, 3)
        self.K = self.EToV.shape[0]
        self.Nv = self.VX.shape[0]

    def c_struct(self):
        m = _Mesh()
        m.K, m.Nv = self.K, self.Nv
        m.VX, m.VY = _d(self.VX), _d(self.VY)
        m.EToV = _i(self.EToV)
        m.vert_id = _i(self.vert_id)
        m.face_tag = _i(self.face_tag)
        return m


def structured_mesh(n, x0, x1, y0, y1, jitter=0.0, seed=0, periodic=False):
    """apl_mesh_structured: n x n quads split into 2 counter-clockwise triangles."""
    K, Nv = 2 * n * n, (n + 1) * (n + 1)
    VX, VY = np.zeros(Nv), np.zeros(Nv)
    EToV = np.zeros((K, 3), dtype=np.int32)
    vid = np.zeros(Nv, dtype=np.int32)
    tags = np.zeros((K, 3), dtype=np.int32)
    rc = lib().apl_mesh_structured(n, x0, x1, y0, y1, jitter, seed, 1 if periodic else 0,
                                   _d(VX), _d(VY), _i(EToV), _i(vid), _i(tags))
    if rc:
        raise ApolloError(rc, "apl_mesh_structured")
    return MeshData(VX, VY, EToV, vid, tags)


def host_connectivity(mesh):
    """ETETF the engine builds for `mesh` (no device needed)."""
    et = np.zeros((mesh.K, 6), dtype=np.int32)
    ms = mesh.c_struct()
    rc = lib().apl_host_connectivity(ctypes.byref(ms), _i(et))
    if rc:
        raise ApolloError(rc, "apl_host_connectivity")
    return et


def host_operator(order, which):
    """Operator `which` (see apl_get_operator) of order N, built on the host."""
    Np = (order + 1) * (order + 2) // 2
    cap = 400 * Np
    out = np.zeros(cap)
    rc = lib().apl_host_operator(order, which, _d(out), cap)
    if rc:
        raise ApolloError(rc, "apl_host_operator")
    Nc = [12, 19, 36, 54, 73, 93, 118, 145][order - 1]
    Ng = 2 * (order + 1)
    sizes = {0: (Nc, Np), 1: (Nc, Np), 2: (Nc, Np), 3: (3 * Ng, Np), 4: (Np, Np), 5: (Nc,),
             6: (Ng,), 7: (Ng,), 8: (Np,), 9: (Np,)}
    n = int(np.prod(sizes[which]))
    return out[:n].reshape(sizes[which]).copy()


# ---------------------------------------------------------------------------
# handle
# ---------------------------------------------------------------------------

class Params:
    """apl_params with keyword defaults."""

    def __init__(self, order, eqn, flux=FLUX_LF, precision=PREC_FP64, gas_gamma=1.4,
                 min_density=0.0, min_pressure=0.0, c0=1.0, chi=0.0, gamma_ph=0.0, bcs=(),
                 algorithm=ALGO_AUTO, limiter=None):
        """limiter: None, or (gas_gamma, min_density, min_pressure) of an
        eulerLimiterHW block."""
        self.order, self.eqn, self.flux, self.precision = order, eqn, flux, precision
        self.limiter = None if limiter is None else tuple(limiter)
        self.algorithm = algorithm
        self.gas_gamma, self.min_density, self.min_pressure = gas_gamma, min_density, min_pressure
        self.c0, self.chi, self.gamma_ph = c0, chi, gamma_ph
        self.bcs = list(bcs)

    def c_struct(self):
        p = _Params()
        p.order, p.eqn, p.flux, p.precision = self.order, self.eqn, self.flux, self.precision
        p.algorithm = self.algorithm
        p.gas_gamma, p.min_density, p.min_pressure = self.gas_gamma, self.min_density, self.min_pressure
        p.c0, p.chi, p.gamma_ph = self.c0, self.chi, self.gamma_ph
        if self.limiter is not None:
            p.limiter = 1
            p.lim_gas_gamma, p.lim_min_density, p.lim_min_pressure = self.limiter
        p.n_bc = len(self.bcs)
        for i, (kind, vals) in enumerate(self.bcs):
            p.bc[i].kind = kind
            for j, v in enumerate(vals):
                p.bc[i].p[j] = v
        return p

    @property
    def m(self):
        return NVARS[self.eqn]


class Engine:
    """One apl_handle: a nodal-DG RHS engine on one mesh."""

    def __init__(self, mesh, params):
        L = lib()
        self.mesh, self.params = mesh, params
        h = ctypes.c_void_p()
        ms, ps = mesh.c_struct(), params.c_struct()
        rc = L.apl_create(ctypes.byref(ms), ctypes.byref(ps), ctypes.byref(h))
        if rc:
            raise ApolloError(rc, L.apl_last_error(None).decode())
        self._h = h
        info = _Info()
        L.apl_get_info(h, ctypes.byref(info))
        self.K, self.N, self.Np, self.Nc, self.Ng, self.Nfp, self.m = (
            info.K, info.N, info.Np, info.Nc, info.Ng, info.Nfp, info.m)
        self.dtscale, self.rmin = info.dtscale, info.rmin
        self.device_bytes = info.device_bytes
        self.device = info.device

    def close(self):
        if getattr(self, "_h", None):
            lib().apl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise ApolloError(rc, "%s: %s" % (what, lib().apl_last_error(self._h).decode()))

    @property
    def shape(self):
        return (self.K, self.Np, self.m)

    # -- maps and geometry -------------------------------------------------
    def index_maps(self):
        et = np.zeros((self.K, 6), dtype=np.int32)
        fm = np.zeros((3, self.Nfp), dtype=np.int32)
        vp = np.zeros((self.K, 3, self.Nfp), dtype=np.int32)
        self._check(lib().apl_get_index_maps(self._h, _i(et), _i(fm), _i(vp)), "index maps")
        return et, fm, vp

    def geometry(self):
        g = np.zeros((self.K, 14))
        x = np.zeros((self.K, self.Np))
        y = np.zeros((self.K, self.Np))
        self._check(lib().apl_get_geometry(self._h, _d(g), _d(x), _d(y)), "geometry")
        return g, x, y

    def operator(self, which):
        sizes = {0: (self.Nc, self.Np), 1: (self.Nc, self.Np), 2: (self.Nc, self.Np),
                 3: (3 * self.Ng, self.Np), 4: (self.Np, self.Np), 5: (self.Nc,),
                 6: (self.Ng,), 7: (self.Ng,), 8: (self.Np,), 9: (self.Np,)}
        out = np.zeros(sizes[which])
        self._check(lib().apl_get_operator(self._h, which, _d(out), out.size), "operator")
        return out

    # -- compute -------------------------------------------------------------
    def _ptr(self, a, name="state"):
        """(pointer, where) of a state-sized numpy array or torch tensor.

        The C-ABI reads and writes K*Np*m doubles at the pointer, so anything
        else is refused here: a dtype other than float64, another element
        count, a non-contiguous layout, or a device tensor on a GPU other
        than the handle's."""
        n = self.K * self.Np * self.m
        if isinstance(a, np.ndarray):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("%s: need a C-contiguous float64 array, got %s" % (name, a.dtype))
            if a.size != n:
                raise ValueError("%s: need K*Np*m = %d values, got %d" % (name, n, a.size))
            return a.ctypes.data, HOST
        if str(getattr(a, "dtype", "")) != "torch.float64" or not a.is_contiguous():
            raise ValueError("%s: need a contiguous torch.float64 tensor, got %s"
                             % (name, getattr(a, "dtype", type(a))))
        if a.numel() != n:
            raise ValueError("%s: need K*Np*m = %d values, got %d" % (name, n, a.numel()))
        if a.is_cuda:
            if a.device.index != self.device:
                raise ValueError("%s: tensor on cuda:%d, handle on device %d"
                                 % (name, a.device.index, self.device))
            return a.data_ptr(), DEVICE
        return a.data_ptr(), HOST

    def _pair(self, a, b, na, nb):
        """Pointers of two state arguments that must live on the same side."""
        pa, wa = self._ptr(a, na)
        pb, wb = self._ptr(b, nb)
        if wa != wb:
            raise ValueError("%s and %s must both be host or both be device memory" % (na, nb))
        return pa, pb, wa

    def rhs(self, t, q, out, cfl=1.0):
        """Apollo step(): out = RHS(q).  Returns the suggested dt."""
        qp, op, where = self._pair(q, out, "q", "out")
        dt = ctypes.c_double(0.0)
        self._check(lib().apl_rhs(self._h, t, qp, op, cfl, ctypes.byref(dt), where), "rhs")
        return dt.value

    def initial_dt(self, q, cfl):
        qp, where = self._ptr(q, "q")
        dt = ctypes.c_double(0.0)
        self._check(lib().apl_initial_dt(self._h, qp, cfl, ctypes.byref(dt), where), "initial_dt")
        return dt.value

    def lserk4(self, t0, dt, nsteps, q):
        qp, where = self._ptr(q, "q")
        self._check(lib().apl_lserk4(self._h, t0, dt, nsteps, qp, where), "lserk4")

    def ssp_advance(self, t0, t_end, dt0, cfl, q, stages=5, max_steps=1000000):
        """Apollo's stepping of one output frame (TSSSP rks2 under ApSolver's
        dt control), in place.  Returns the per-step dt values."""
        qp, where = self._ptr(q)
        dts = np.zeros(max(max_steps, 1))
        n = ctypes.c_int(0)
        self._check(lib().apl_ssp_advance(self._h, t0, t_end, dt0, cfl, stages, qp, where,
                                          max_steps, ctypes.byref(n), _d(dts)), "ssp_advance")
        return dts[:n.value].copy()

    def limit(self, q_in, q_out):
        """Apollo's Hesthaven-Warburton Euler limiter: q_out = limited q_in
        (q_out may be q_in)."""
        ip, op, where = self._pair(q_in, q_out, "q_in", "q_out")
        self._check(lib().apl_limit(self._h, ip, op, where), "limit")

    def adjoint_rhs(self, t, q, lam, out):
        qp, lp, where = self._pair(q, lam, "q", "lam")
        op, _, _ = self._pair(out, q, "out", "q")
        self._check(lib().apl_adjoint_rhs(self._h, t, qp, lp, op, where), "adjoint_rhs")

    def lserk4_adjoint(self, t0, dt, nsteps, ckpt_every, q0, qbarT, dJdq0, qT=None):
        q0p, qbp, where = self._pair(q0, qbarT, "q0", "qbarT")
        gp, _, _ = self._pair(dJdq0, q0, "dJdq0", "q0")
        qtp = self._pair(qT, q0, "qT", "q0")[0] if qT is not None else None
        self._check(lib().apl_lserk4_adjoint(self._h, t0, dt, nsteps, ckpt_every, q0p, qbp, gp,
                                             qtp, where), "lserk4_adjoint")

    def objective_disc_sq(self, q, var, xc, yc, radius, dJdq=None):
        qp, where = self._ptr(q, "q")
        gp = self._pair(dJdq, q, "dJdq", "q")[0] if dJdq is not None else None
        J = ctypes.c_double(0.0)
        self._check(lib().apl_objective_disc_sq(self._h, qp, var, xc, yc, radius, ctypes.byref(J),
                                                gp, where), "objective")
        return J.value

    def suggested_dt(self, cfl, max_speed):
        return 2.0 / 3.0 * cfl * self.dtscale * (self.rmin / max_speed)

    def solve(self, t_end, nout, dt, cfl, q, run_name=None, mesh=None, stages=5):
        """ApSolver::solve (src/solvers/apsolver.cc:172-216): write frame 0,
        then nout frames of length t_end/nout, each TSSolve restarted from
        dt, writing `<run_name>_<i>.vtu` after each when run_name is given
        (needs `mesh`).  q (host) is advanced in place; returns the number
        of steps of each frame."""
        from . import vtu
        if run_name is not None:
            vtu.write_frame(vtu.frame_name(run_name, 0), mesh, q)
        tsize = t_end / nout
        steps = []
        for i in range(nout):
            steps.append(len(self.ssp_advance(i * tsize, (i + 1) * tsize, dt, cfl, q, stages=stages)))
            if run_name is not None:
                vtu.write_frame(vtu.frame_name(run_name, i + 1), mesh, q)
        return steps


# ---------------------------------------------------------------------------
# initial conditions of the benchmark configurations (closed forms)
# ---------------------------------------------------------------------------

def vortex_state(x, y, t=0.0, beta=5.0, gamma=1.4, x0=0.0, y0=0.0, u0=1.0, v0=0.0, m=4):
    """Isentropic vortex (examples/unstructuredDG/euler/isentropicVortex/isentropicVortex.pin:47-66),
    translated by (u0, v0) t.  m = 4: (rho, rho u, rho v, E); m = 5 adds rho w = 0."""
    xm, ym = x - u0 * t - x0, y - v0 * t - y0
    r2 = xm * xm + ym * ym
    pi = 3.141592653589793
    u = u0 - beta * np.exp(1.0 - r2) * ym / (2.0 * pi)
    v = v0 + beta * np.exp(1.0 - r2) * xm / (2.0 * pi)
    base = 1.0 - ((gamma - 1.0) * beta * beta * np.exp(2.0 * (1.0 - r2)) / (16.0 * gamma * pi * pi))
    rho = base ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    comps = [rho, rho * u, rho * v] + ([np.zeros_like(rho)] if m == 5 else []) + [E]
    return np.stack(comps, axis=-1)


def periodic_vortex_state(x, y, t, L, beta=5.0, gamma=1.4, x0=0.0, y0=0.0, u0=1.0, v0=0.0, m=4):
    """The vortex at time t on a periodic square of side L: evaluated at the
    periodic image of (x - u0 t, y - v0 t) nearest to the centre (x0, y0)."""
    xs = (x - u0 * t - x0 + L / 2) % L - L / 2
    ys = (y - v0 * t - y0 + L / 2) % L - L / 2
    return vortex_state(xs, ys, 0.0, beta=beta, gamma=gamma, x0=0.0, y0=0.0, u0=u0, v0=v0, m=m)


def tm_cavity_mode(x, y, t=0.0, p=3, q=4, c0=1.0, m=3):
    """PEC cavity mode on the unit square: Ez = sin(p pi x) sin(q pi y) cos(w t),
    Bx = -(q/k) sin(p pi x) cos(q pi y) sin(w t), By = (p/k) cos(p pi x) sin(q pi y) sin(w t),
    k = sqrt(p^2+q^2), w = c0 pi k.  m = 8 embeds (Ez, Bx, By) in phMaxwell."""
    k = np.sqrt(p * p + q * q)
    w = c0 * np.pi * k
    Ez = np.sin(p * np.pi * x) * np.sin(q * np.pi * y) * np.cos(w * t)
    Bx = -(q / k) * np.sin(p * np.pi * x) * np.cos(q * np.pi * y) * np.sin(w * t)
    By = (p / k) * np.cos(p * np.pi * x) * np.sin(q * np.pi * y) * np.sin(w * t)
    if m == 3:
        return np.stack([Ez, Bx, By], axis=-1)
    z = np.zeros_like(Ez)
    return np.stack([z, z, Ez, Bx, By, z, z, z], axis=-1)


def random_fourier_ez(x, y, nmodes=16, seed=4, m=3):
    """Ez = sum over the nmodes smallest (p,q) >= 1 (by p^2+q^2, then p) of
    a_pq sin(p pi x) sin(q pi y), a_pq ~ N(0,1)/(p^2+q^2); H = 0."""
    cand = sorted(((p * p + q * q, p, q) for p in range(1, 9) for q in range(1, 9)))[:nmodes]
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(len(cand))
    Ez = np.zeros_like(x)
    for ai, (k2, p, q) in zip(a, cand):
        Ez = Ez + ai / k2 * np.sin(p * np.pi * x) * np.sin(q * np.pi * y)
    z = np.zeros_like(Ez)
    if m == 3:
        return np.stack([Ez, z, z], axis=-1)
    return np.stack([z, z, Ez, z, z, z, z, z], axis=-1)
