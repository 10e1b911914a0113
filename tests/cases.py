"""Shared test cases: meshes, parameters and states for the engine and the
oracle, plus the maps between the engine's variable sets and Apollo's.

The engine runs EULER4 / MAXWELL_TM3 (the physical 2D variables) as well as
Apollo's own EULER5 / PHMAXWELL8 layouts; the oracle always uses Apollo's 5
(Euler) or 8 (phMaxwell) variables.  EULER4 embeds with rho w = 0, TM3 embeds
as (0, 0, Ez, Bx, By, 0, 0, 0)."""
import os

import numpy as np

from apollo_amd import engine as E
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

EMBED = {E.EQN_EULER5: None, E.EQN_EULER4: [0, 1, 2, 4],
         E.EQN_PHMAXWELL8: None, E.EQN_MAXWELL_TM3: [2, 3, 4]}


def oracle_eqn(eqn):
    return O.ORC_EULER if eqn in (E.EQN_EULER5, E.EQN_EULER4) else O.ORC_PHMAXWELL


def to_oracle(eqn, q):
    """Engine state [K,Np,m] -> oracle state [K,Np,5|8]."""
    idx = EMBED[eqn]
    if idx is None:
        return np.ascontiguousarray(q)
    mo = 5 if oracle_eqn(eqn) == O.ORC_EULER else 8
    out = np.zeros(q.shape[:-1] + (mo,))
    out[..., idx] = q
    return out


def from_oracle(eqn, qo):
    idx = EMBED[eqn]
    return np.ascontiguousarray(qo if idx is None else qo[..., idx])


def vortex_bc(u0=1.0, v0=0.0, x0=0.0, y0=0.0, gamma=1.4, beta=5.0):
    return (E.BC_ISENTROPIC_VORTEX, [u0, v0, x0, y0, gamma, beta])


def oracle_phys(params):
    eqn = oracle_eqn(params.eqn)
    flux = {E.FLUX_LF: O.ORC_FLUX_LF, E.FLUX_UPWIND: O.ORC_FLUX_UPWIND, E.FLUX_HLL: O.ORC_FLUX_HLL,
            E.FLUX_ROE: O.ORC_FLUX_ROE, E.FLUX_WAVE: O.ORC_FLUX_WAVE}[params.flux]
    return O.phys(eqn, flux=flux, gamma=params.gas_gamma, min_dens=params.min_density,
                  min_pres=params.min_pressure, c0=params.c0, chi=params.chi,
                  gph=params.gamma_ph, bcs=params.bcs, limiter=params.limiter)


class Case:
    """A mesh + parameters, with the oracle's view of the same problem."""

    def __init__(self, mesh, params, name=""):
        self.mesh, self.params, self.name = mesh, params, name
        self.ops = O.Ops(params.order)
        self.omesh = O.Mesh(self.ops, mesh.VX, mesh.VY, mesh.EToV, mesh.vert_id, mesh.face_tag)
        self.ophys = oracle_phys(params)

    @property
    def shape(self):
        return (self.mesh.K, self.ops.Np, self.params.m)

    def state(self, kind="smooth", seed=0, amp=1e-2):
        """A valid state at the nodes (engine layout)."""
        x, y = self.omesh.x, self.omesh.y
        rng = np.random.default_rng(seed)
        eqn, m = self.params.eqn, self.params.m
        if eqn in (E.EQN_EULER4, E.EQN_EULER5):
            q = E.vortex_state(x, y, 0.0, m=m)
            q = q * (1.0 + amp * rng.standard_normal(q.shape))
            if eqn == E.EQN_EULER5:
                q[..., 3] = amp * rng.standard_normal(q.shape[:-1])
            return np.ascontiguousarray(q)
        if kind == "cavity":
            return np.ascontiguousarray(E.tm_cavity_mode(x, y, 0.0, m=m))
        return np.ascontiguousarray(rng.standard_normal(self.shape))

    def oracle_rhs(self, q, t=0.0):
        r, ms = O.rhs(self.omesh, self.ophys, t, to_oracle(self.params.eqn, q))
        return from_oracle(self.params.eqn, r), ms


def rounding_floor(case, q, t=0.0, seeds=(1, 2)):
    """Rounding floor of the FP64 RHS at state q, measured on the
    exact-geometry oracle in two ways, of which the larger is returned:
      - its 1-ulp input sensitivity: the largest relative L2 change of the
        RHS when every entry of q is moved by -1, 0 or +1 ulp at random (one
        draw per seed);
      - the relative L2 difference between the oracle and the same oracle
        with every matrix-vector product summed in reverse order.
    Two FP64 evaluations that order their operations differently cannot be
    expected to agree more closely than this.  On fine meshes it is well
    above 1e-12 for smooth states: the volume and surface terms, each of
    order |F| / h, cancel to a much smaller RHS (in the far field of the
    vortex to almost nothing)."""
    eps = np.finfo(np.float64).eps
    with O.exact_geometry():
        ref, _ = case.oracle_rhs(q, t)
        worst = 0.0
        for s in seeds:
            r = np.random.default_rng(s).integers(-1, 2, size=q.shape)
            moved, _ = case.oracle_rhs(np.ascontiguousarray(q * (1.0 + eps * r)), t)
            worst = max(worst, rel_l2(moved, ref))
        with O.reverse_sums():
            rev, _ = case.oracle_rhs(q, t)
    return max(worst, rel_l2(rev, ref))


def fp64_parity(case, got, q, t=0.0, tol=1e-12, floor=None):
    """FP64 parity bar.  The engine evaluates the geometric factors of each
    straight-sided element exactly, Apollo recomputes them at every cubature
    point from Dr_c x (wxcubature2d.cc:305), which adds rounding noise of order
    N^2 |x| / h * eps.  The engine must be within `tol` of the oracle run with
    exact geometric factors, and no farther from the faithful oracle than
    max(tol, 2 x that noise).

    With `floor` (rounding_floor of this state), the exact-geometry bar is
    max(tol, 4 x floor): the engine differs from the oracle by its summation
    order, its folded operators (V V^T, and the cubature folded into nodal
    operators for linear fluxes) and its exact geometric factors, each an
    FP64 rounding of about that size.  Returns (err_exact, err_faithful,
    noise)."""
    ref, _ = case.oracle_rhs(q, t)
    with O.exact_geometry():
        ref_exact, _ = case.oracle_rhs(q, t)
    noise = rel_l2(ref, ref_exact)
    e_exact, e_faith = rel_l2(got, ref_exact), rel_l2(got, ref)
    bar = tol if floor is None else max(tol, 4.0 * floor)
    print("%s: rel L2 vs exact-geometry oracle %.3e, vs faithful oracle %.3e "
          "(oracle q8 noise %.3e%s; bar %.2e)"
          % (case.name, e_exact, e_faith, noise,
             "" if floor is None else ", FP64 rounding floor %.3e" % floor, bar))
    assert e_exact <= bar, (e_exact, bar)
    assert e_faith <= max(bar, 2.0 * noise), (e_faith, noise)
    return e_exact, e_faith, noise


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def frame_sensitivity(case, q0, t1, dt0, cfl, qo, dto, seeds=(1, 2, 3), centred=True):
    """How far the oracle's own FP64 roundings move one output frame of
    Apollo's stepping (O.ssp_advance over [0, t1] from q0), as the largest
    relative L2 change of the increment qo - q0 and the largest relative
    change of a dt, over
      - a random one-ulp relative perturbation of q0 (each seed);
      - the oracle with the elements' exact affine geometric factors
        (Apollo recomputes them at the cubature points, quirk q8; the
        engine uses the affine constants);
      - with `centred`, the oracle on the mesh translated to be centred on
        the origin (Apollo's limiter forms absolute node coordinates and
        subtracts their average, a cancellation of size eps |x| / h; the
        engine's reconstruction works with coordinate offsets and does not
        see it).  Only for problems whose boundary states do not depend on
        position.
    Returns (sensitivity, dt sensitivity, {kind: (increment, dt)})."""
    eqn = case.params.eqn
    qo0 = to_oracle(eqn, q0)
    inc = qo - q0
    parts = {}

    def record(kind, q, dts):
        n = min(len(dts), len(dto))
        parts[kind] = (rel_l2(from_oracle(eqn, q) - q0, inc),
                       float(np.abs(dts[:n] - dto[:n]).max() / dto.max()) if n else 0.0)

    for seed in seeds:
        rng = np.random.default_rng(seed)
        record("seed %d" % seed, *O.ssp_advance(case.omesh, case.ophys, 0.0, t1, dt0, cfl,
                                                qo0 * (1.0 + 2.2e-16 * rng.standard_normal(qo0.shape)),
                                                max_steps=20000))
    with O.exact_geometry():
        record("exact geometry", *O.ssp_advance(case.omesh, case.ophys, 0.0, t1, dt0, cfl, qo0,
                                                max_steps=20000))
    if centred:
        m = case.mesh
        cx, cy = 0.5 * (m.VX.min() + m.VX.max()), 0.5 * (m.VY.min() + m.VY.max())
        om = O.Mesh(case.ops, m.VX - cx, m.VY - cy, m.EToV, m.vert_id, m.face_tag)
        record("centred mesh", *O.ssp_advance(om, case.ophys, 0.0, t1, dt0, cfl, qo0,
                                              max_steps=20000))
    return max(v[0] for v in parts.values()), max(v[1] for v in parts.values()), parts


# --- standard meshes ------------------------------------------------------

def periodic_square(n, L=10.0, jitter=0.1, seed=0):
    return E.structured_mesh(n, -L / 2, L / 2, -L / 2, L / 2, jitter, seed, True)


def walled_unit_square(n, jitter=0.1, seed=1):
    return E.structured_mesh(n, 0.0, 1.0, 0.0, 1.0, jitter, seed, False)


def euler_case(n=16, N=3, m=4, periodic=True, precision=E.PREC_FP64, L=10.0, seed=0,
               flux=E.FLUX_LF):
    eqn = E.EQN_EULER4 if m == 4 else E.EQN_EULER5
    if periodic:
        mesh = periodic_square(n, L=L, seed=seed)
        bcs = []
    else:
        mesh = E.structured_mesh(n, -L / 2, L / 2, -L / 2, L / 2, 0.1, seed, False)
        bcs = [vortex_bc()]
    return Case(mesh, E.Params(N, eqn, flux=flux, precision=precision, bcs=bcs), "euler")


# Apollo's forward-facing-step deck (examples/unstructuredDG/euler/
# forwardFacingStep/forwardFacingStep.pin): Mach 3 inflow of rho = gamma,
# p = 5 over the step mesh grid.msh, boundary list [inflow, bottomWall,
# topWall, noGrad], HLL flux, eulerLimiterHW with min_pressure =
# min_density = 1e-2, cfl = 1/(2(N+1)-1).
FFS_MACH, FFS_P, FFS_GAMMA = 3.0, 5.0, 1.4


def ffs_case(N=1, m=5, flux=E.FLUX_HLL, limiter=True, precision=E.PREC_FP64):
    from apollo_amd.gmsh import read_gmsh
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "ffs_grid.msh"))
    inflow = (E.BC_EULER_INFLOW, [FFS_MACH, 0.0, FFS_GAMMA, FFS_P, FFS_GAMMA])
    bcs = [inflow, (E.BC_EULER_WALL, []), (E.BC_EULER_WALL, []), (E.BC_COPY, [])]
    eqn = E.EQN_EULER5 if m == 5 else E.EQN_EULER4
    lim = (FFS_GAMMA, 1e-2, 1e-2) if limiter else None
    return Case(mesh, E.Params(N, eqn, flux=flux, precision=precision, gas_gamma=FFS_GAMMA,
                               bcs=bcs, limiter=lim), "ffs")


def ffs_state(case, kind="uniform", seed=0):
    """The deck's uniform Mach-3 state, or that state with a smooth
    perturbation of amplitude 10% (still positive)."""
    x, y = case.omesh.x, case.omesh.y
    g, p0 = FFS_GAMMA, FFS_P
    rho = np.full_like(x, g)
    u = np.full_like(x, FFS_MACH * np.sqrt(g * p0 / g))
    v = np.zeros_like(x)
    p = np.full_like(x, p0)
    if kind != "uniform":
        rng = np.random.default_rng(seed)
        for arr, amp in ((rho, 0.1), (u, 0.1), (v, 0.1), (p, 0.1)):
            k = rng.standard_normal(4)
            arr *= 1.0 + amp * np.sin(k[0] * 3 * x + k[1] * 3 * y + k[2]) * np.cos(k[3] * 2 * y)
            if arr is v:
                arr += amp * np.sin(2 * x + 3 * y + k[2])
    q = np.zeros(x.shape + (case.params.m,))
    E_ = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    q[..., 0], q[..., 1], q[..., 2] = rho, rho * u, rho * v
    q[..., -1] = E_
    return np.ascontiguousarray(q)


# Apollo's other Euler shock decks (examples/unstructuredDG/euler/*/*.pin),
# each on the mesh it names except the backward-facing step, whose grid2.msh
# (63280 triangles) is replaced by the deck's own grid.msh (same geometry
# and boundary tags, 15820 triangles).  Boundary list entry i is physical
# tag i + 1.  All run at P_ORDER = 1 with cfl = 1/(2(N+1)-1).
#   name: (mesh fixture, triangles, flux, bcs, limiter (gamma, min_dens,
#          min_pres) or None, initial state)
def _uniform_flow(mach, p):
    def q(x, y, g):
        rho = np.full_like(x, g)
        rhou = rho * mach * np.sqrt(g * p / rho)
        return rho, rhou, np.zeros_like(x), np.full_like(x, p)
    return q


def _bfs_state(x, y, g):
    # backwardFacingStep.pin: rho = gm, rhou = if(x > 0.5, 0, rho M a)
    rho, rhou, rhov, p = _uniform_flow(5.09, 5.0)(x, y, g)
    return rho, np.where(x > 0.5, 0.0, rhou), rhov, p


def _explosion_state(x, y, g):
    # explosion.pin: r = |(x, y) - (1, 1)|, rho = if(r > 0.4, 0.125, 1),
    # p = if(r > 0.4, 0.1, 1), at rest
    r = np.sqrt((x - 1.0) * (x - 1.0) + (y - 1.0) * (y - 1.0))
    z = np.zeros_like(x)
    return np.where(r > 0.4, 0.125, 1.0), z, z, np.where(r > 0.4, 0.1, 1.0)


def _pulse_state(x, y, g):
    # eulerCircularPulse.pin: rho = rho0 (10 + exp(-10 r^2)), p = p0 (10 +
    # exp(-10 r^2)), rho0 = gamma, p0 = 1, at rest
    b = 10.0 + 1.0 * np.exp(-10.0 * (x * x + y * y))
    z = np.zeros_like(x)
    return g * b, z, z, 1.0 * b


_INFLOW = lambda mach, p: (E.BC_EULER_INFLOW, [mach, 0.0, 1.4, p, 1.4])
_WALL, _COPY = (E.BC_EULER_WALL, []), (E.BC_COPY, [])
DECKS = {
    "bfs": ("bfs_grid.msh.gz", 15820, E.FLUX_HLL, [_INFLOW(5.09, 5.0), _WALL, _COPY],
            (1.4, 1e-2, 1e-2), _bfs_state),
    "explosion": ("explosion_gridBox.msh.gz", 16790, E.FLUX_LF,
                  [(E.BC_EULER_DIRICHLET, [0.125, 0.0, 0.0, 0.0, 1.4, 0.1])],
                  (1.4, 1e-2, 1e-2), _explosion_state),
    "pulse": ("pulse_grid2.msh.gz", 3472, E.FLUX_HLL, [_WALL], None, _pulse_state),
    "scramjet": ("scramjet_grid.msh.gz", 10414, E.FLUX_HLL,
                 [_WALL, _WALL, _INFLOW(3.0, 5.0), _COPY], (1.4, 1e-2, 1e-2),
                 _uniform_flow(3.0, 5.0)),
}


def deck_case(name, N=1):
    from apollo_amd.gmsh import read_gmsh
    fixture, _, flux, bcs, lim, _ = DECKS[name]
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", fixture))
    return Case(mesh, E.Params(N, E.EQN_EULER5, flux=flux, gas_gamma=1.4, bcs=bcs,
                               limiter=lim), name)


def deck_state(name, case):
    """The deck's initial state at the nodes, [K][Np][5]."""
    g = 1.4
    rho, rhou, rhov, p = DECKS[name][5](case.omesh.x, case.omesh.y, g)
    q = np.zeros(case.omesh.x.shape + (5,))
    q[..., 0], q[..., 1], q[..., 2] = rho, rhou, rhov
    q[..., 4] = p / (g - 1) + 0.5 * (rhou * rhou + rhov * rhov) / rho
    return np.ascontiguousarray(q)


def maxwell_case(n=12, N=5, m=3, bc=E.BC_MAXWELL_TM, flux=E.FLUX_LF, precision=E.PREC_FP64,
                 seed=1, chi=0.0, gph=0.0):
    eqn = E.EQN_MAXWELL_TM3 if m == 3 else E.EQN_PHMAXWELL8
    mesh = walled_unit_square(n, seed=seed)
    return Case(mesh, E.Params(N, eqn, flux=flux, precision=precision, chi=chi, gamma_ph=gph,
                               bcs=[(bc, [])]), "maxwell")
