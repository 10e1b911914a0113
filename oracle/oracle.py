"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  The engine never does.

The oracle restates Apollo's nodal-DG RHS in C (``apollo_oracle.c``); see
the header of that file for the reference file:line each function follows.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

ORC_EULER, ORC_PHMAXWELL = 0, 1
ORC_FLUX_LF, ORC_FLUX_UPWIND, ORC_FLUX_HLL, ORC_FLUX_ROE, ORC_FLUX_WAVE = 0, 1, 2, 3, 4
ORC_BC_NONE, ORC_BC_VORTEX, ORC_BC_TM, ORC_BC_PEC, ORC_BC_COPY = 0, 1, 2, 3, 4
ORC_BC_EULER_WALL, ORC_BC_EULER_INFLOW, ORC_BC_EULER_DIRICHLET = 5, 6, 7

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Ops(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Np", ctypes.c_int), ("Nc", ctypes.c_int),
                ("Ng", ctypes.c_int), ("Nfp", ctypes.c_int), ("rmin", ctypes.c_double),
                ("r", _dp), ("s", _dp), ("fmask", _ip),
                ("V", _dp), ("invV", _dp), ("VVT", _dp), ("Dr", _dp), ("Ds", _dp),
                ("cr", _dp), ("cs", _dp), ("cw", _dp),
                ("Vc", _dp), ("Drc", _dp), ("Dsc", _dp),
                ("VcT", _dp), ("DrcT", _dp), ("DscT", _dp),
                ("gz", _dp), ("gw", _dp), ("If", _dp), ("IfT", _dp)]


class _Phys(ctypes.Structure):
    _fields_ = [("eqn", ctypes.c_int), ("flux", ctypes.c_int),
                ("gamma", ctypes.c_double), ("min_dens", ctypes.c_double),
                ("min_pres", ctypes.c_double), ("c0", ctypes.c_double),
                ("chi", ctypes.c_double), ("gph", ctypes.c_double),
                ("nbc", ctypes.c_int), ("bc_kind", ctypes.c_int * 8),
                ("bc_p", (ctypes.c_double * 8) * 8),
                ("lim_gamma", ctypes.c_double), ("lim_min_dens", ctypes.c_double),
                ("lim_min_pres", ctypes.c_double), ("limiter", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_build_ops.argtypes = [ctypes.c_int, ctypes.POINTER(_Ops)]
        L.orc_free_ops.argtypes = [ctypes.POINTER(_Ops)]
        L.orc_jacobiP.restype = ctypes.c_double
        L.orc_jacobiP.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int]
        L.orc_grad_jacobiP.restype = ctypes.c_double
        L.orc_grad_jacobiP.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int]
        L.orc_gauleg.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_facepair.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip]
        L.orc_node_coords.argtypes = [ctypes.POINTER(_Ops), ctypes.c_int, _dp, _dp, _ip,
                                      _dp, _dp, _dp]
        L.orc_elem_geom.argtypes = [ctypes.POINTER(_Ops), _dp, _dp, _dp]
        L.orc_flux.argtypes = [ctypes.POINTER(_Phys), ctypes.c_int, _dp, _dp]
        L.orc_numflux.argtypes = [ctypes.POINTER(_Phys), _dp, _dp, _dp, _dp, _dp]
        L.orc_bc.argtypes = [ctypes.POINTER(_Phys), ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                              _dp, _dp, _ip, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int]
        L.orc_initial_dt.restype = ctypes.c_double
        L.orc_initial_dt.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                                     _dp, ctypes.c_double, ctypes.c_double]
        L.orc_lserk4.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                                 _dp, _dp, _ip, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, _dp, ctypes.c_int]
        L.orc_adjoint_rhs.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                                      _dp, _dp, _ip, ctypes.c_double, _dp, _dp, _dp]
        L.orc_limit.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                                _dp, _dp, _ip, _dp, _dp]
        L.orc_limiter_weights.argtypes = [ctypes.POINTER(_Ops), _dp]
        L.orc_ssp_advance.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys), ctypes.c_int,
                                      _dp, _dp, _ip, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), _dp, ctypes.c_int]
        L.orc_lserk4_adjoint.argtypes = [ctypes.POINTER(_Ops), ctypes.POINTER(_Phys),
                                         ctypes.c_int, _dp, _dp, _ip, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.orc_set_exact_geometry.argtypes = [ctypes.c_int]
        L.orc_set_reverse_sums.argtypes = [ctypes.c_int]
        L.orc_objective_disc_sq.restype = ctypes.c_double
        L.orc_objective_disc_sq.argtypes = [ctypes.POINTER(_Ops), ctypes.c_int, _dp, _dp,
                                            ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def jacobiP(x, a, b, n):
    return lib().orc_jacobiP(x, a, b, n)


def grad_jacobiP(x, a, b, n):
    return lib().orc_grad_jacobiP(x, a, b, n)


def gauleg(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauleg(-1.0, 1.0, _d(x), _d(w), n)
    return x, w


class Ops:
    """Operators of one order, copied into numpy arrays."""

    def __init__(self, N):
        L = lib()
        o = _Ops()
        rc = L.orc_build_ops(N, ctypes.byref(o))
        if rc:
            raise ValueError("orc_build_ops(%d) failed: %d" % (N, rc))
        self._o = o
        self.N, self.Np, self.Nc, self.Ng, self.Nfp = o.N, o.Np, o.Nc, o.Ng, o.Nfp
        self.rmin = o.rmin
        Np, Nc, Ng = o.Np, o.Nc, o.Ng

        def arr(p, *shape):
            n = int(np.prod(shape))
            return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(shape)

        self.r, self.s = arr(o.r, Np), arr(o.s, Np)
        self.fmask = arr(o.fmask, 3, o.Nfp)
        self.V, self.invV, self.VVT = arr(o.V, Np, Np), arr(o.invV, Np, Np), arr(o.VVT, Np, Np)
        self.Dr, self.Ds = arr(o.Dr, Np, Np), arr(o.Ds, Np, Np)
        self.cr, self.cs, self.cw = arr(o.cr, Nc), arr(o.cs, Nc), arr(o.cw, Nc)
        self.Vc, self.Drc, self.Dsc = arr(o.Vc, Nc, Np), arr(o.Drc, Nc, Np), arr(o.Dsc, Nc, Np)
        self.gz, self.gw = arr(o.gz, Ng), arr(o.gw, Ng)
        self.If = arr(o.If, 3 * Ng, Np)

    def __del__(self):
        try:
            lib().orc_free_ops(ctypes.byref(self._o))
        except Exception:
            pass


def phys(eqn, flux=ORC_FLUX_LF, gamma=1.4, min_dens=0.0, min_pres=0.0, c0=1.0, chi=0.0,
         gph=0.0, bcs=(), limiter=None):
    """bcs: sequence of (kind, params) indexed by tag-1.  limiter: None, or
    (gas_gamma, min_density, min_pressure) of an eulerLimiterHW block."""
    p = _Phys()
    if limiter is not None:
        p.limiter = 1
        p.lim_gamma, p.lim_min_dens, p.lim_min_pres = limiter
    p.eqn, p.flux = eqn, flux
    p.gamma, p.min_dens, p.min_pres = gamma, min_dens, min_pres
    p.c0, p.chi, p.gph = c0, chi, gph
    p.nbc = len(bcs)
    for i, (kind, params) in enumerate(bcs):
        p.bc_kind[i] = kind
        for j, v in enumerate(params):
            p.bc_p[i][j] = v
    return p


def meqn(ph):
    return 5 if ph.eqn == ORC_EULER else 8


def facepair(EToV, Nv, vert_id=None, face_tag=None):
    EToV = np.ascontiguousarray(EToV, dtype=np.int32)
    K = EToV.shape[0]
    et = np.zeros((K, 6), dtype=np.int32)
    vid = None if vert_id is None else np.ascontiguousarray(vert_id, dtype=np.int32)
    ft = None if face_tag is None else np.ascontiguousarray(face_tag, dtype=np.int32)
    rc = lib().orc_facepair(K, Nv, _i(EToV), _i(vid), _i(ft), _i(et))
    if rc:
        raise ValueError("orc_facepair: %d" % rc)
    return et


def node_coords(ops, VX, VY, EToV):
    EToV = np.ascontiguousarray(EToV, dtype=np.int32)
    VX = np.ascontiguousarray(VX, dtype=np.float64)
    VY = np.ascontiguousarray(VY, dtype=np.float64)
    K = EToV.shape[0]
    x = np.zeros((K, ops.Np))
    y = np.zeros((K, ops.Np))
    dts = np.zeros(1)
    lib().orc_node_coords(ctypes.byref(ops._o), K, _d(VX), _d(VY), _i(EToV), _d(x), _d(y), _d(dts))
    return x, y, float(dts[0])


def elem_geom(ops, x, y):
    g = np.zeros(14)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    rc = lib().orc_elem_geom(ctypes.byref(ops._o), _d(x), _d(y), _d(g))
    if rc:
        raise ValueError("orc_elem_geom: %d" % rc)
    return g


def flux(ph, d, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    f = np.zeros(meqn(ph))
    rc = lib().orc_flux(ctypes.byref(ph), d, _d(q), _d(f))
    if rc:
        raise ValueError("orc_flux: %d" % rc)
    return f


def numflux(ph, n, qM, qP):
    n = np.ascontiguousarray(n, dtype=np.float64)
    qM = np.array(qM, dtype=np.float64)
    qP = np.array(qP, dtype=np.float64)
    out = np.zeros(meqn(ph))
    lam = np.zeros(1)
    rc = lib().orc_numflux(ctypes.byref(ph), _d(n), _d(qM), _d(qP), _d(out), _d(lam))
    if rc:
        raise ValueError("orc_numflux: %d" % rc)
    return out, float(lam[0])


def bc(ph, bcnum, xc, n, qM):
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    qM = np.ascontiguousarray(qM, dtype=np.float64)
    qP = np.zeros(meqn(ph))
    lib().orc_bc(ctypes.byref(ph), bcnum, _d(xc), _d(n), _d(qM), _d(qP))
    return qP


class Mesh:
    """Mesh + connectivity + node coordinates as the oracle sees them."""

    def __init__(self, ops, VX, VY, EToV, vert_id=None, face_tag=None):
        self.ops = ops
        self.VX = np.ascontiguousarray(VX, dtype=np.float64)
        self.VY = np.ascontiguousarray(VY, dtype=np.float64)
        self.EToV = np.ascontiguousarray(EToV, dtype=np.int32)
        self.K = self.EToV.shape[0]
        self.etetf = facepair(self.EToV, len(self.VX), vert_id, face_tag)
        self.x, self.y, self.dtscale = node_coords(ops, self.VX, self.VY, self.EToV)


class exact_geometry:
    """Context manager: oracle RHS with affine geometric factors at the
    cubature points (see orc_set_exact_geometry)."""

    def __enter__(self):
        lib().orc_set_exact_geometry(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_exact_geometry(0)
        return False


class reverse_sums:
    """Context manager: every matrix-vector product of the oracle RHS sums
    its terms in reverse order (see orc_set_reverse_sums)."""

    def __enter__(self):
        lib().orc_set_reverse_sums(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_reverse_sums(0)
        return False


def rhs(mesh, ph, t, q, nthreads=0):
    ops = mesh.ops
    m = meqn(ph)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(mesh.K, ops.Np, m)
    out = np.zeros_like(q)
    ms = np.zeros(1)
    rc = lib().orc_rhs(ctypes.byref(ops._o), ctypes.byref(ph), mesh.K, _d(mesh.x), _d(mesh.y),
                       _i(mesh.etetf), t, _d(q), _d(out), _d(ms), nthreads)
    if rc:
        raise ValueError("orc_rhs: %d" % rc)
    return out, float(ms[0])


def suggested_dt(mesh, cfl, max_speed):
    """wxnodaldg2dmethod.cc:523."""
    return 2. / 3. * cfl * mesh.dtscale * (mesh.ops.rmin / max_speed)


def initial_dt(mesh, ph, q, cfl):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().orc_initial_dt(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K, _d(q),
                                cfl, mesh.dtscale)


def lserk4(mesh, ph, t0, dt, nsteps, q, nthreads=0):
    q = np.array(q, dtype=np.float64, copy=True)
    rc = lib().orc_lserk4(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K, _d(mesh.x),
                          _d(mesh.y), _i(mesh.etetf), t0, dt, nsteps, _d(q), nthreads)
    if rc:
        raise ValueError("orc_lserk4: %d" % rc)
    return q


def ssp_advance(mesh, ph, t0, t_end, dt0, cfl, q, stages=5, max_steps=100000, nthreads=0):
    """One output frame of Apollo's stepping (TSSSP rks2 under ApSolver's
    dt control).  Returns (q, dts)."""
    q = np.array(q, dtype=np.float64, copy=True)
    dts = np.zeros(max(max_steps, 1))
    n = ctypes.c_int(0)
    rc = lib().orc_ssp_advance(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K, _d(mesh.x),
                               _d(mesh.y), _i(mesh.etetf), t0, t_end, dt0, cfl, mesh.dtscale,
                               stages, _d(q), max_steps, ctypes.byref(n), _d(dts), nthreads)
    if rc:
        raise ValueError("orc_ssp_advance: %d" % rc)
    return q, dts[:n.value].copy()


def limiter_weights(ops):
    ave = np.zeros(ops.Np)
    rc = lib().orc_limiter_weights(ctypes.byref(ops._o), _d(ave))
    if rc:
        raise ValueError("orc_limiter_weights: %d" % rc)
    return ave


def limit(mesh, ph, q):
    """Apollo's Hesthaven-Warburton Euler limiter applied to q (a copy)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros_like(q)
    rc = lib().orc_limit(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K, _d(mesh.x),
                         _d(mesh.y), _i(mesh.etetf), _d(q), _d(out))
    if rc:
        raise ValueError("orc_limit: %d" % rc)
    return out


def adjoint_rhs(mesh, ph, t, q, lam):
    q = np.ascontiguousarray(q, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.zeros_like(q)
    rc = lib().orc_adjoint_rhs(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K, _d(mesh.x),
                               _d(mesh.y), _i(mesh.etetf), t, _d(q), _d(lam), _d(out))
    if rc:
        raise ValueError("orc_adjoint_rhs: %d" % rc)
    return out


def lserk4_adjoint(mesh, ph, t0, dt, nsteps, q0, qbarT):
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    qbarT = np.ascontiguousarray(qbarT, dtype=np.float64)
    out = np.zeros_like(q0)
    rc = lib().orc_lserk4_adjoint(ctypes.byref(mesh.ops._o), ctypes.byref(ph), mesh.K,
                                  _d(mesh.x), _d(mesh.y), _i(mesh.etetf), t0, dt, nsteps,
                                  _d(q0), _d(qbarT), _d(out))
    if rc:
        raise ValueError("orc_lserk4_adjoint: %d" % rc)
    return out


def objective_disc_sq(mesh, q, var, xc, yc, radius, m):
    q = np.ascontiguousarray(q, dtype=np.float64)
    g = np.zeros_like(q)
    J = lib().orc_objective_disc_sq(ctypes.byref(mesh.ops._o), mesh.K, _d(mesh.x), _d(mesh.y), m,
                                    _d(q), var, xc, yc, radius, _d(g))
    return J, g
