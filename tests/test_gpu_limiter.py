"""GPU parity of Apollo's Hesthaven-Warburton Euler limiter (apl_limit) and
of the steppers that limit every RHS input, against the oracle.

Bar: FP64, max |engine - oracle| <= 1e-12 max |oracle| for the limiter, and
for the steppers the same dt sequence (1e-12 relative) and the state within
1e-11 relative L2.  Cases: the forward-facing-step deck (Apollo's mesh
grid.msh, inflow / wall / zero-gradient boundaries, HLL flux), a periodic
mesh, Dirichlet walls with both floors active."""
import numpy as np
import pytest

from apollo_amd import engine as E
from oracle import oracle as O

from cases import (Case, euler_case, ffs_case, ffs_state, frame_sensitivity, from_oracle, rel_l2,
                   to_oracle)

pytestmark = pytest.mark.gpu


def _close(got, want, tol=1e-12):
    err = np.abs(got - want).max() / np.abs(want).max()
    print("limiter max rel err %.3e" % err)
    assert err <= tol, err


def _engine_limit(case, q, inplace=False, device=False):
    eng = E.Engine(case.mesh, case.params)
    if device:
        import torch
        qd = torch.from_numpy(q.copy()).cuda()
        out = qd if inplace else torch.zeros_like(qd)
        eng.limit(qd, out)
        res = out.cpu().numpy()
    else:
        out = q.copy() if inplace else np.zeros_like(q)
        eng.limit(out if inplace else q, out)
        res = out
    eng.close()
    return res


@pytest.mark.parametrize("m", [4, 5])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_limiter_ffs_matches_oracle(N, m):
    case = ffs_case(N=N, m=m)
    q = ffs_state(case, "smooth", seed=N)
    want = from_oracle(case.params.eqn, O.limit(case.omesh, case.ophys, to_oracle(case.params.eqn, q)))
    _close(_engine_limit(case, q), want)
    _close(_engine_limit(case, q, inplace=True, device=True), want)


def test_limiter_floors_and_dirichlet_match_oracle():
    case = euler_case(n=6, N=3, m=5, periodic=False)
    bcs = [(E.BC_EULER_DIRICHLET, [0.125, 0.0, 0.0, 0.0, 1.4, 0.1])]
    params = E.Params(3, E.EQN_EULER5, bcs=bcs, limiter=(1.4, 0.3, 0.3))
    c = Case(case.mesh, params, "dirichlet")
    x, y = c.omesh.x, c.omesh.y
    rho = np.where(x < 0.3, 1.0, 0.125) + 0.05 * np.sin(y)
    p = np.where(x * x + y * y < 4.0, 1.0, 0.1)
    u, v, w = 0.2 * np.cos(y), 0.1 * np.sin(x), 0.05 * np.cos(x + y)
    q = np.ascontiguousarray(np.stack([rho, rho * u, rho * v, rho * w,
                                       p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], -1))
    want = O.limit(c.omesh, c.ophys, q)
    _close(_engine_limit(c, q), want)


def test_limiter_periodic_matches_oracle():
    case = euler_case(n=8, N=2, m=4)
    c = Case(case.mesh, E.Params(2, E.EQN_EULER4, limiter=(1.4, 1e-2, 1e-2)), "periodic")
    q = case.state(amp=5e-2)
    want = from_oracle(c.params.eqn, O.limit(c.omesh, c.ophys, to_oracle(c.params.eqn, q)))
    _close(_engine_limit(c, q), want)


def test_limit_needs_a_limiter_handle():
    case = ffs_case(N=1, limiter=False)
    eng = E.Engine(case.mesh, case.params)
    q = ffs_state(case)
    with pytest.raises(E.ApolloError) as ei:
        eng.limit(q, q.copy())
    assert ei.value.status == 1
    eng.close()


@pytest.mark.parametrize("N", [1, 4])
def test_limit_reports_nan_then_starts_clean(N):
    """A NaN in the state reaches the limited state and apl_limit reports
    APL_ERR_NAN (7, as wxnodaldg2dmethod.cc:246 does for the RHS); the next
    call on the same handle starts from cleared flags, succeeds and matches
    the oracle."""
    case = ffs_case(N=N)
    q = ffs_state(case, "smooth", seed=3)
    bad = q.copy()
    bad[5, 0, 0] = np.nan
    eng = E.Engine(case.mesh, case.params)
    with pytest.raises(E.ApolloError) as ei:
        eng.limit(bad, np.zeros_like(q))
    assert ei.value.status == 7
    out = np.zeros_like(q)
    eng.limit(q, out)
    eng.close()
    want = from_oracle(case.params.eqn, O.limit(case.omesh, case.ophys, to_oracle(case.params.eqn, q)))
    _close(out, want)


@pytest.mark.parametrize("nsteps", [4, 13])
@pytest.mark.parametrize("N", [1, 2])
def test_ssp_ffs_deck_matches_oracle(N, nsteps):
    """Apollo's forward-facing-step deck: the uniform Mach-3 start, a frame
    of TSSSP rks2 steps with the limiter on every stage vector.

    This start is ill-conditioned: a one-ulp relative perturbation of the
    initial state moves the oracle's own increment by ~1e-12 after 4 steps
    and by 4e-8 (N = 1) to 5e-6 (N = 2) after 13 (the limiter's gradient
    weights switch between neighbours as the flow at the step develops).
    The bar is therefore stated against the oracle's own sensitivities
    (cases.frame_sensitivity: one-ulp input perturbations, Apollo's
    geometric factors, the translated mesh): the engine's increment is
    within max(1e-12, 10 x sensitivity) of the oracle's, each dt within
    max(1e-12, 10 x the dt sensitivity)."""
    case = ffs_case(N=N)
    q0 = ffs_state(case)
    cfl = 1.0 / (2 * (N + 1) - 1)
    qo0 = to_oracle(case.params.eqn, q0)
    dt0 = O.initial_dt(case.omesh, case.ophys, qo0, cfl)
    t1 = (nsteps - 0.5) * dt0
    eng = E.Engine(case.mesh, case.params)
    q = q0.copy()
    dts = eng.ssp_advance(0.0, t1, dt0, cfl, q)
    eng.close()
    qo, dto = O.ssp_advance(case.omesh, case.ophys, 0.0, t1, dt0, cfl, qo0)
    qo = from_oracle(case.params.eqn, qo)
    sens, sens_dt, parts = frame_sensitivity(case, q0, t1, dt0, cfl, qo, dto)
    err = rel_l2(q - q0, qo - q0)
    print("N=%d: %d steps, increment rel L2 vs oracle %.3e; oracle sensitivities (increment, "
          "dt): %s" % (N, len(dts), err, ", ".join("%s %.2e %.2e" % (k, *v) for k, v in parts.items())),
          flush=True)
    assert len(dts) == len(dto) == nsteps
    assert np.abs(dts - dto).max() <= max(1e-12, 10.0 * sens_dt) * dto.max()
    assert err <= max(1e-12, 10.0 * sens)
    assert rel_l2(q, q0) > 1e-3                              # the steps did act


def test_lserk4_with_limiter_matches_oracle():
    case = ffs_case(N=2, m=4, flux=E.FLUX_LF)
    q0 = ffs_state(case, "smooth", seed=7)
    eng = E.Engine(case.mesh, case.params)
    q = q0.copy()
    dt = 2e-4
    eng.lserk4(0.0, dt, 6, q)
    eng.close()
    qo = from_oracle(case.params.eqn, O.lserk4(case.omesh, case.ophys, 0.0, dt, 6,
                                               to_oracle(case.params.eqn, q0)))
    err = rel_l2(q - q0, qo - q0)
    print("LSERK4 + limiter increment rel L2 %.3e" % err)
    assert err <= 1e-11


def test_solve_writes_apollo_frames(tmp_path):
    """ApSolver::solve on the forward-facing-step deck: frame 0 is the
    initial state, frames 1..nout follow each TSSolve; the last frame holds
    the engine's state."""
    from apollo_amd import vtu
    case = ffs_case(N=1)
    q = ffs_state(case)
    eng = E.Engine(case.mesh, case.params)
    dt0 = eng.initial_dt(q, 1.0 / 3.0)
    run = str(tmp_path / "ffstep")
    steps = eng.solve(6 * dt0, 2, dt0, 1.0 / 3.0, q, run_name=run, mesh=case.mesh)
    eng.close()
    assert len(steps) == 2 and min(steps) >= 3
    pts, cells, arrays = vtu.read_frame(vtu.frame_name(run, 2))
    m = case.params.m
    got = np.stack([arrays[1 + j][1] for j in range(case.ops.Np * m)], -1).reshape(q.shape)
    assert np.array_equal(got, q)
    _, _, a0 = vtu.read_frame(vtu.frame_name(run, 0))
    assert np.allclose(a0[1][1], ffs_state(case)[:, 0, 0])


def test_cpp_host_runs_the_ffs_deck(tmp_path):
    """The C++ host (apollo_amd/host: the Gmsh reader with DMPlex's
    numbering and the Apollo-shaped subsolver) runs the forward-facing-step
    deck: step() limits its input in place and returns the RHS; solve() runs
    two frames.  The results equal the Python binding's bit for bit (same
    library, same sequence of launches) and the oracle to the FP64 bar."""
    import os
    import subprocess
    from cases import GOLDEN
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ffs_deck_demo")
    assert os.path.exists(exe), "build tests/cpp first (make -C tests/cpp)"
    out = tmp_path / "ffs.bin"
    r = subprocess.run([exe, os.path.join(GOLDEN, "meshes", "ffs_grid.msh"), "1", "2", "6.0",
                        str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(str(out), dtype=np.float64)
    dt0, sdt, n = raw[0], raw[1], int(raw[2])
    q0, qs, rhs, qf = (raw[3 + i * n:3 + (i + 1) * n] for i in range(4))
    case = ffs_case(N=1)
    shape = case.shape
    q0, qs, rhs, qf = (a.reshape(shape) for a in (q0, qs, rhs, qf))
    # the deck's state (the C++ expression orders the products differently)
    assert np.abs(q0 - ffs_state(case)).max() <= 1e-14 * np.abs(q0).max()
    eng = E.Engine(case.mesh, case.params)
    cfl = 1.0 / 3.0
    assert dt0 == eng.initial_dt(q0, cfl)
    ql = np.zeros_like(q0)
    eng.limit(q0, ql)
    assert np.array_equal(qs, ql)
    r2 = np.zeros_like(q0)
    assert eng.rhs(0.0, ql, r2, cfl=cfl) == sdt
    assert np.array_equal(rhs, r2)
    qe = q0.copy()
    eng.solve(6.0 * dt0, 2, dt0, cfl, qe)
    eng.close()
    assert np.array_equal(qf, qe)
    # against the oracle: the limiter and the RHS of the limited state
    qo = to_oracle(case.params.eqn, q0)
    _close(qs, from_oracle(case.params.eqn, O.limit(case.omesh, case.ophys, qo)))
    from cases import fp64_parity
    fp64_parity(case, rhs, qs)
