"""Apollo's two decks on the hot path, end to end on the GPU through
apl_ssp_advance (TSSSP rks2 under ApSolver's dt control) against the oracle:

  - examples/unstructuredDG/euler/isentropicVortex/isentropicVortex.pin:
    Apollo's 5-variable Euler (gas_gamma 1.4, LF flux, the deck's default),
    N = 1, cfl = 1/(2(N+1)-1), on grid2.msh ([0,10] x [-5,5], 1030
    triangles, every boundary face tag 1 -> `allSides`, the isentropic-vortex
    boundary state u0 = 1, v0 = 0, x0 = 5, y0 = 0, beta = 5), the vortex
    centred at (5, 0) as initial condition; Time = [0, 1], one output frame,
    Dt = 0.01.
  - examples/unstructuredDG/maxwell/transverseMagnetic/transverseMagnetic.pin:
    Apollo's 8-variable phMaxwell (c0 = 1, chi = gamma = 0 by default, LF
    flux), N = 2, cfl = 0.5/(2(N+1)-1), on square3.msh ([-1,1]^2, 3968
    triangles, tag 1 -> the maxwellTransverseMagneticBC state),
    Ez = sin(pi x) sin(pi y); Time = [0, 10], one frame, Dt = 0.01.

The first step's dt is min(Dt, init's suggested dt) (ApSolver::init,
src/solvers/apsolver.cc:273-283; WxNodalDG2dMethod::init,
src/subsolvers/wxnodaldg2dmethod.cc:201-228).  Bar, as for the shock decks
(test_gpu_decks.py): the increment q(T) - q(0) within max(1e-12, 10 x the
oracle's own sensitivity) in relative L2, the same step count, each dt
within max(1e-12, 10 x the oracle's dt sensitivity).  The transverse-magnetic deck runs 14567 steps; the
oracle follows its first 2 time units (2914 steps), and the engine's whole
frame is compared with the analytic cavity mode."""
import os

import numpy as np
import pytest

from apollo_amd import engine as E
from apollo_amd.gmsh import read_gmsh
from oracle import oracle as O

from cases import GOLDEN, Case, frame_sensitivity, from_oracle, rel_l2, to_oracle

pytestmark = pytest.mark.gpu


def _vortex_deck():
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "vortex_grid2.msh.gz"))
    bc = (E.BC_ISENTROPIC_VORTEX, [1.0, 0.0, 5.0, 0.0, 1.4, 5.0])   # u0 v0 x0 y0 gamma beta
    case = Case(mesh, E.Params(1, E.EQN_EULER5, gas_gamma=1.4, bcs=[bc]), "isentropicVortex")
    q0 = np.ascontiguousarray(E.vortex_state(case.omesh.x, case.omesh.y, 0.0, x0=5.0, m=5))
    return case, q0


def _tm_deck():
    mesh = read_gmsh(os.path.join(GOLDEN, "meshes", "tm_square3.msh.gz"))
    case = Case(mesh, E.Params(2, E.EQN_PHMAXWELL8, c0=1.0, bcs=[(E.BC_MAXWELL_TM, [])]),
                "transverseMagnetic")
    x, y = case.omesh.x, case.omesh.y
    q0 = np.zeros(x.shape + (8,))
    q0[..., 2] = np.sin(np.pi * x) * np.sin(np.pi * y)
    return case, np.ascontiguousarray(q0)


def _tm_exact(x, y, t):
    """The (1,1) cavity mode of [-1,1]^2 with Ez = 0 on the walls:
    Ez = sin(pi x) sin(pi y) cos(w t), Bx = -(pi/w) sin(pi x) cos(pi y)
    sin(w t), By = (pi/w) cos(pi x) sin(pi y) sin(w t), w = sqrt(2) pi."""
    w = np.sqrt(2.0) * np.pi
    q = np.zeros(x.shape + (8,))
    q[..., 2] = np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(w * t)
    q[..., 3] = -(np.pi / w) * np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(w * t)
    q[..., 4] = (np.pi / w) * np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(w * t)
    return q


def _against_oracle(case, q0, t_end, dt0, cfl, q, dts):
    """Increment and dt sequence of the engine's frame [0, t_end] against
    the oracle's, with the oracle's own sensitivities as the scale."""
    qo0 = to_oracle(case.params.eqn, q0)
    qo, dto = O.ssp_advance(case.omesh, case.ophys, 0.0, t_end, dt0, cfl, qo0, max_steps=20000)
    qo = from_oracle(case.params.eqn, qo)
    # the vortex boundary state depends on position: no translated mesh
    sens, sens_dt, parts = frame_sensitivity(case, q0, t_end, dt0, cfl, qo, dto, centred=False)
    err = rel_l2(q - q0, qo - q0)
    ddt = np.abs(dts - dto).max() / dto.max() if len(dts) == len(dto) else np.nan
    print("%s: [0, %g], %d steps: increment rel L2 vs oracle %.3e, dt max diff %.3e; oracle "
          "sensitivities (increment, dt): %s"
          % (case.name, t_end, len(dts), err, ddt,
             ", ".join("%s %.2e %.2e" % (k, *v) for k, v in parts.items())), flush=True)
    assert len(dts) == len(dto)
    assert ddt <= max(1e-12, 10.0 * sens_dt)
    assert err <= max(1e-12, 10.0 * sens)
    return qo


def test_isentropic_vortex_deck():
    case, q0 = _vortex_deck()
    assert case.mesh.K == 1030
    cfl = 1.0 / (2 * (1 + 1) - 1)
    dt0 = min(0.01, O.initial_dt(case.omesh, case.ophys, to_oracle(case.params.eqn, q0), cfl))
    eng = E.Engine(case.mesh, case.params)
    q = q0.copy()
    dts = eng.ssp_advance(0.0, 1.0, dt0, cfl, q)
    eng.close()
    qo = _against_oracle(case, q0, 1.0, dt0, cfl, q, dts)
    ex = E.vortex_state(case.omesh.x, case.omesh.y, 1.0, x0=5.0, m=5)
    e_eng, e_or = rel_l2(q, ex), rel_l2(qo, ex)
    print("isentropicVortex at t = 1: rel L2 vs the translated vortex: engine %.4e, oracle %.4e"
          % (e_eng, e_or), flush=True)
    assert abs(e_eng - e_or) <= 1e-6 * e_or
    assert np.isfinite(q).all() and (q[..., 0] > 0).all()


def test_transverse_magnetic_deck():
    case, q0 = _tm_deck()
    assert case.mesh.K == 3968
    cfl = 0.5 / (2 * (2 + 1) - 1)
    dt0 = min(0.01, O.initial_dt(case.omesh, case.ophys, to_oracle(case.params.eqn, q0), cfl))
    eng = E.Engine(case.mesh, case.params)
    q = q0.copy()
    dts = eng.ssp_advance(0.0, 2.0, dt0, cfl, q)
    _against_oracle(case, q0, 2.0, dt0, cfl, q, dts)
    # the rest of the deck's frame: restarting at t = 2 with the last dt is
    # the same stepping, since the dt control keeps dt constant here (the
    # LF speed of phMaxwell is c0 max(chi, gamma, 1))
    assert np.all(dts[:-1] == dts[0])
    q_full = q0.copy()
    dts_full = eng.ssp_advance(0.0, 10.0, dt0, cfl, q_full)
    eng.close()
    assert np.array_equal(dts_full[:len(dts) - 1], dts[:-1])
    x, y = case.omesh.x, case.omesh.y
    err = rel_l2(q_full, _tm_exact(x, y, 10.0))
    e0 = float(np.sum(q0 * q0))
    e1 = float(np.sum(q_full * q_full))
    print("transverseMagnetic: [0, 10], %d steps: rel L2 vs the cavity mode %.3e; "
          "nodal energy %.6e -> %.6e" % (len(dts_full), err, e0, e1), flush=True)
    assert len(dts_full) >= 14000
    assert err <= 1e-2
    assert e1 <= e0 * (1 + 1e-12)
    assert np.all(q_full[..., [0, 1, 5, 6, 7]] == 0.0)
