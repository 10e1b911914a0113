"""The five benchmark workloads: mesh, parameters, initial state, dt rule and
step count of each, shared by the benchmark and the full-size parity tests.

Meshes are n x n quads split into two counter-clockwise triangles with the
interior vertices jittered (apl_mesh_structured).  dt is Apollo's formula
(src/subsolvers/wxnodaldg2dmethod.cc:523, (2/3) cfl dtscale rmin / maxSpeed)
evaluated once from the initial state and then held fixed.

  vortex_64_N3         Euler isentropic vortex (strength 5, centre (0,0),
                       drift (1,0)), [-5,5]^2 periodic, 64x64x2 (K=8192),
                       jitter 0.1h seed 0, N=3, LF, CFL 0.5, 100 steps
  tm_cavity_256_N5     Maxwell TM (Ez,Bx,By) in the unit square, PEC walls
                       (Apollo's TM boundary state), 256x256x2 (K=131072),
                       jitter 0.1h seed 1, N=5, upwind flux,
                       Ez = sin(3 pi x) sin(4 pi y), H = 0, 500 steps
  vortex_512_N4        Euler isentropic vortex, [-10,10]^2 periodic,
                       512x512x2 (K=524288), jitter 0.15h seed 2, N=4, LF,
                       200 steps
  adjoint_vortex_256_N3  Euler isentropic vortex, [-5,5]^2 periodic,
                       256x256x2 (K=131072), jitter 0.1h seed 3, N=3, LF;
                       J = int_{r<1} rho^2 after 200 steps, checkpoint
                       every 20 steps
  tm_fourier_384_N8    Maxwell TM, unit square, PEC, 384x384x2 (K=294912),
                       jitter 0.1h seed 4, N=8, upwind, Ez from 16 Fourier
                       modes with N(0,1)/k^2 amplitudes, 300 steps

The Maxwell decks pass Apollo's CFL already divided by 2N+1
(examples/unstructuredDG/euler/isentropicVortex/isentropicVortex.pin:38),
so the TM workloads use 0.5/(2N+1); the Euler workloads use 0.5.
"""
from . import engine as E

NAMES = ("vortex_64_N3", "tm_cavity_256_N5", "vortex_512_N4", "adjoint_vortex_256_N3",
         "tm_fourier_384_N8")


class Workload:
    def __init__(self, name, n, box, jitter, seed, periodic, order, eqn, flux, bcs, cfl, steps,
                 state, desc, ckpt=None):
        self.name, self.n, self.box, self.jitter, self.seed = name, n, box, jitter, seed
        self.periodic, self.order, self.eqn, self.flux, self.bcs = periodic, order, eqn, flux, bcs
        self.cfl, self.steps, self.state_kind, self.desc, self.ckpt = cfl, steps, state, desc, ckpt
        self._mesh = None

    @property
    def m(self):
        return E.NVARS[self.eqn]

    @property
    def K(self):
        return 2 * self.n * self.n

    @property
    def L(self):
        return self.box[1] - self.box[0]

    @property
    def linear(self):
        return self.eqn in (E.EQN_MAXWELL_TM3, E.EQN_PHMAXWELL8)

    def mesh(self):
        if self._mesh is None:
            x0, x1 = self.box
            self._mesh = E.structured_mesh(self.n, x0, x1, x0, x1, self.jitter, self.seed,
                                           self.periodic)
        return self._mesh

    def params(self, precision=E.PREC_FP64, algorithm=E.ALGO_AUTO):
        return E.Params(self.order, self.eqn, flux=self.flux, precision=precision,
                        algorithm=algorithm, bcs=self.bcs)

    def initial_state(self, x, y):
        """[K][Np][m] state at node coordinates x, y [K][Np]."""
        if self.state_kind == "cavity":
            return E.tm_cavity_mode(x, y, 0.0, m=self.m)
        if self.state_kind == "fourier":
            return E.random_fourier_ez(x, y, 16, 4, m=self.m)
        return E.periodic_vortex_state(x, y, 0.0, self.L, m=self.m)

    def exact_state(self, x, y, t):
        """The analytic solution at time t, or None when there is none."""
        if self.state_kind == "cavity":
            return E.tm_cavity_mode(x, y, t, m=self.m)
        if self.state_kind == "vortex":
            return E.periodic_vortex_state(x, y, t, self.L, m=self.m)
        return None

    def fixed_dt(self, eng, q0):
        """Apollo's dt formula from q0 (the Maxwell wave speed is c0 = 1);
        `eng` is an Engine on this workload's mesh."""
        if self.linear:
            return eng.suggested_dt(self.cfl, 1.0)
        import numpy as np
        return eng.rhs(0.0, q0, np.zeros_like(q0), cfl=self.cfl)


def get(name):
    if name == "vortex_64_N3":
        return Workload(name, 64, (-5.0, 5.0), 0.1, 0, True, 3, E.EQN_EULER4, E.FLUX_LF, [], 0.5,
                        100, "vortex",
                        "2D Euler isentropic vortex (strength 5, centre (0,0)), [-5,5]^2 periodic, "
                        "64x64x2 jittered 0.1h seed 0, K=8192, N=3, LF flux, CFL 0.5, 100 steps")
    if name == "tm_cavity_256_N5":
        return Workload(name, 256, (0.0, 1.0), 0.1, 1, False, 5, E.EQN_MAXWELL_TM3, E.FLUX_UPWIND,
                        [(E.BC_MAXWELL_TM, [])], 0.5 / 11, 500, "cavity",
                        "2D Maxwell TM (Ez,Bx,By), unit square, PEC walls (Apollo TM boundary "
                        "state), 256x256x2 jittered 0.1h seed 1, K=131072, N=5, upwind flux, "
                        "Ez=sin(3 pi x)sin(4 pi y), H=0, 500 steps")
    if name == "vortex_512_N4":
        return Workload(name, 512, (-10.0, 10.0), 0.15, 2, True, 4, E.EQN_EULER4, E.FLUX_LF, [],
                        0.5, 200, "vortex",
                        "2D Euler isentropic vortex, [-10,10]^2 periodic, 512x512x2 jittered "
                        "0.15h seed 2, K=524288, N=4, LF flux, 200 steps")
    if name == "adjoint_vortex_256_N3":
        return Workload(name, 256, (-5.0, 5.0), 0.1, 3, True, 3, E.EQN_EULER4, E.FLUX_LF, [], 0.5,
                        200, "vortex",
                        "adjoint: 2D Euler isentropic vortex, [-5,5]^2 periodic, 256x256x2 "
                        "jittered 0.1h seed 3, K=131072, N=3, LF flux, J = int_{r<1} rho^2 at "
                        "step 200, checkpoint every 20", ckpt=20)
    if name == "tm_fourier_384_N8":
        return Workload(name, 384, (0.0, 1.0), 0.1, 4, False, 8, E.EQN_MAXWELL_TM3, E.FLUX_UPWIND,
                        [(E.BC_MAXWELL_TM, [])], 0.5 / 17, 300, "fourier",
                        "2D Maxwell TM, unit square, PEC walls, 384x384x2 jittered 0.1h seed 4, "
                        "K=294912, N=8, upwind flux, Ez from 16 Fourier modes with N(0,1)/k^2 "
                        "amplitudes, 300 steps")
    raise KeyError("unknown workload %r (known: %s)" % (name, ", ".join(NAMES)))
